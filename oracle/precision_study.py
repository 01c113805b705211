"""Storage-precision study on the CPU: how the growth-rate error of the
reduced storage forms scales with the stored significand width.

Test infrastructure (it drives the oracle, ``oracle/vlasov_oracle.py``); it is
not part of the product.  It follows ``Stepper.nonlinear_ssp_rk``
(main/timestep.py:141-187) statement by statement and, after each stage value
is formed, replaces the state's deviation from the ring background,
g = f - f0 (f0 = the ring, main/grid.py:244-279), by g rounded to a p-bit
significand with stochastic rounding (unbounded exponent range).  p = 8 is
BF16's significand, p = 11 FP16's, p = 24 FP32's.  Two placements:

* ``all``   -- f1, f2 and f^{n+1} rounded: what the 16-bit GPU forms store;
* ``final`` -- only f^{n+1} rounded, the stage values kept exact: the best any
  scheme can do that stores the state itself at p bits (the stage buffers
  are then wider than 2 bytes per value).

For each run the field energy W(t) is recorded as the reference's loop does
(main/timestep.py:81-139) and the growth rate is the least-squares slope of
log W / 2 over the fit window; the report is the relative error against the
unrounded FP64 run of the same grid.

    python -m oracle.precision_study --bits 8 9 10 11 --seeds 0 1 --out study.json
"""
import argparse
import json
import time

import numpy as np

from .vlasov_oracle import NONLINEAR_SSP_RK, Case, ring_distribution

EIGEN = dict(ic="eigen", k0=0.0886, om=-0.349j, om_pc=10.0, delta_n=1e-3, B=0.1, ring_j=6)


def round_significand(x, bits, rng):
    """x rounded stochastically to a `bits`-bit significand (element-wise)."""
    out = np.zeros_like(x)
    nz = x != 0
    m, e = np.frexp(x[nz])                 # x = m 2^e, 0.5 <= |m| < 1
    scaled = np.ldexp(m, bits)             # |scaled| in [2^(bits-1), 2^bits)
    lo = np.floor(scaled)
    up = rng.random(lo.shape) < (scaled - lo)
    out[nz] = np.ldexp(lo + up, e - bits)
    return out


class RoundedStepper:
    """SSP-RK3 of the oracle with the stage values' deviation from f0 rounded."""

    coefficients = np.array(NONLINEAR_SSP_RK[3])

    def __init__(self, dt, f0, bits, place, seed):
        self.dt, self.f0, self.bits, self.place = dt, f0, bits, place
        self.rng = np.random.default_rng(seed)

    def _round(self, arr, ng):
        if self.bits is None:
            return
        g = arr[ng] - self.f0[ng]
        arr[ng] = self.f0[ng] + round_significand(g, self.bits, self.rng)

    def step(self, case):
        d, e, b, grids, flux = case.dist, case.elliptic, case.basis, case.grids, case.flux
        d.swap_ghost_cells_of_array()
        stages = np.zeros((3,) + d.arr.shape)
        stages[:, :, :, 0] = d.arr[:, :, 0]
        stages[:, :, :, -1] = d.arr[:, :, -1]
        stages[:, :, :, :, :, 0] = d.arr[:, :, :, :, 0]
        stages[:, :, :, :, :, -1] = d.arr[:, :, :, :, -1]
        ng = grids.no_ghost_slice
        density = d.moment_zero()
        e.poisson2(np.mean(density) - density, grids.x, b.b1)
        stages[(0,) + ng] = d.arr[ng] + self.dt * flux.semi_discrete_rhs(d.arr, e, b, grids)[ng]
        if self.place == "all":
            self._round(stages[0], ng)
        for i in range(1, 3):
            stages[i - 1, 0] = stages[i - 1, -2]
            stages[i - 1, -1] = stages[i - 1, 1]
            density = d.moment_zero_of_arr(stages[i - 1])
            e.poisson2(np.mean(density) - density, grids.x, b.b1)
            df_dt = flux.semi_discrete_rhs(stages[i - 1], e, b, grids)[ng]
            c = self.coefficients[i - 1]
            stages[(i,) + ng] = c[0] * d.arr[ng] + c[1] * stages[(i - 1,) + ng] + c[2] * self.dt * df_dt
            if self.place == "all" or i == 2:
                self._round(stages[i], ng)
        d.arr[ng] = stages[-1][ng]


def run(n, nx, nv, dt, t_end, bits, place, seed):
    case = Case(n=n, nx=nx, nu=nv, nv=nv, **EIGEN)
    f0 = ring_distribution(case.grids, case.dist.vt, case.dist.ring_j)
    st = RoundedStepper(dt, f0, bits, place, seed)
    steps = int(round(t_end / dt))
    energy = [case.elliptic.electric_energy(case.grids.x)]
    for _ in range(steps):
        st.step(case)
        energy.append(case.elliptic.electric_energy(case.grids.x))
    return np.arange(steps + 1) * dt, np.asarray(energy), case.dist.arr


def growth_rate(t, w, window):
    sel = (t >= window[0]) & (t <= window[1])
    return 0.5 * np.polyfit(t[sel], np.log(w[sel]), 1)[0]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=6)
    ap.add_argument("--nx", type=int, default=8)
    ap.add_argument("--nv", type=int, default=12)
    ap.add_argument("--dt", type=float, default=0.01)
    ap.add_argument("--t-end", type=float, default=20.0)
    ap.add_argument("--bits", type=int, nargs="+", default=[8, 9, 10, 11])
    ap.add_argument("--places", nargs="+", default=["all", "final"])
    ap.add_argument("--seeds", type=int, nargs="+", default=[0, 1])
    ap.add_argument("--out", default=None)
    a = ap.parse_args()
    window = (a.t_end / 2, a.t_end)
    t0 = time.time()
    t, w64, f64 = run(a.n, a.nx, a.nv, a.dt, a.t_end, None, "all", 0)
    g64 = growth_rate(t, w64, window)
    rep = {"grid": {"n": a.n, "nx": a.nx, "nu": a.nv, "nv": a.nv, "dt": a.dt, "t_end": a.t_end,
                    "window": window, **{k: (str(v) if isinstance(v, complex) else v) for k, v in EIGEN.items()}},
           "gamma_f64": g64, "runs": []}
    print(json.dumps({"gamma_f64": g64, "seconds": time.time() - t0}), flush=True)
    for place in a.places:
        for bits in a.bits:
            for seed in a.seeds:
                t, w, f = run(a.n, a.nx, a.nv, a.dt, a.t_end, bits, place, seed)
                g = growth_rate(t, w, window)
                ng = (slice(1, -1), slice(None)) * 3
                r = {"place": place, "bits": bits, "seed": seed, "gamma": g,
                     "rel_gamma_err": abs(g - g64) / abs(g64),
                     "rel_l2_f": float(np.linalg.norm(f[ng] - f64[ng]) / np.linalg.norm(f64[ng])),
                     "rel_energy_end": abs(w[-1] - w64[-1]) / w64[-1]}
                rep["runs"].append(r)
                print(json.dumps(r), flush=True)
    if a.out:
        with open(a.out, "w") as fh:
            json.dump(rep, fh, indent=1)


if __name__ == "__main__":
    main()
