"""Error report of the reduced-precision storage forms against the FP64 form.

    python -m vlasov1d2v_amd.errors [--workload ring64] [--t-end 20] [--forms f32,bf16,f16]
                                    [--ic eigen --k0 0.0886 --om -0.349j --om-pc 10 --delta-n 1e-3]

Every form runs the same seeded workload to ``t_end`` with the same fixed dt
(steps replayed from a HIP graph).  The FP64 GPU form is the reference: it is
the reference's arithmetic, pinned to the NumPy oracle at 1e-12 relative L2
on f and on the energy series (ring16 over 50 steps, ring64 over 2 steps).
Reported per form:

* rel_l2_f           ||f - f64|| / ||f64|| over the interior at t_end (f is
                     the total distribution, background + stored deviation);
* rel_energy_end     relative error of the field energy at t_end;
* rel_l2_energy      relative L2 error of the whole field-energy series;
* gamma              growth rate fitted to the field energy W(t) ~ exp(2 gamma t)
                     by least squares on log W over the fit window, and its
                     relative error against the FP64 form's gamma;
* gamma_env          the same fit on the envelope of the oscillating field:
                     log of the maximum of W within +-0.25 of each local
                     maximum of the FP64 series, over the whole run.

The energy series is the reference's (main/timestep.py:81-139): one entry
per step, the field of the last stage solve, weights without the dx/2
Jacobian (main/elliptic.py:150-151).  Prints one JSON line.
"""
import argparse
import json
import time

import numpy as np


def run_form(w, storage, t_end, dt, graph_steps, device, mfma=True, **ic):
    import torch
    from . import _lib
    from .grid import STORAGE_CODES
    from .workload import Run
    run = Run.from_workload(w, dt=dt, device=device, storage=storage, **ic)
    if not mfma:
        run.grids.geometry(STORAGE_CODES[storage]).set_tuning(_lib.VL_TUNE_MFMA, 0)
    steps = int(round(t_end / dt))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    st = run.run(steps, graph_steps=graph_steps)
    torch.cuda.synchronize()
    seconds = time.perf_counter() - t0
    f = run.dist.total()[1:-1, :, 1:-1, :, 1:-1, :]
    return np.asarray(st.time_array), np.asarray(st.field_energy), f, seconds, run


def growth_rate(t, energy, window):
    """gamma of W ~ exp(2 gamma t): least squares of log W on [t0, t1]."""
    t0, t1 = window
    sel = (t >= t0) & (t <= t1) & (energy > 0)
    slope, _ = np.polyfit(t[sel], np.log(energy[sel]), 1)
    return 0.5 * slope


def envelope_rate(t, energy, ref_energy, half_width=0.25):
    """gamma of the peak envelope: peaks located on the reference series,
    each form's maximum within +-half_width of every reference peak time."""
    i = np.arange(1, len(ref_energy) - 1)
    pk = i[(ref_energy[i] > ref_energy[i - 1]) & (ref_energy[i] >= ref_energy[i + 1])]
    if len(pk) < 2:
        return float("nan")
    vals = np.array([energy[(t >= t[p] - half_width) & (t <= t[p] + half_width)].max() for p in pk])
    slope, _ = np.polyfit(t[pk], np.log(vals), 1)
    return 0.5 * slope


def main(argv=None):
    import torch
    from .workload import CONFIGS
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="ring64")
    ap.add_argument("--t-end", type=float, default=20.0)
    ap.add_argument("--dt", type=float, default=0.0, help="time step (default: the workload's stable dt)")
    ap.add_argument("--forms", default="f32,bf16,f16")
    ap.add_argument("--fit", default="", help="fit window t0,t1 (default: the second half of the run)")
    ap.add_argument("--graph-steps", type=int, default=50)
    ap.add_argument("--save", default="", help="write the time and energy series of every form to this .npz")
    ap.add_argument("--no-mfma", action="store_true",
                    help="16-bit forms on the vector-ALU stage kernel (element-major) instead of the matrix cores")
    ap.add_argument("--ic", default="modulated", choices=["modulated", "eigen"],
                    help="the workload's modulated ring, or the reference driver's ring + Bernstein eigenfunction")
    ap.add_argument("--k0", type=float, default=0.0, help="wave number (default: the workload's)")
    ap.add_argument("--om", type=complex, default=-0.349j, help="eigenfunction frequency (ic=eigen)")
    ap.add_argument("--om-pc", type=float, default=10.0, help="omega_p / omega_c of the eigenfunction (ic=eigen)")
    ap.add_argument("--delta-n", type=float, default=1e-3, help="eigenfunction amplitude (ic=eigen)")
    args = ap.parse_args(argv)
    w = CONFIGS[args.workload]
    if args.k0 > 0:
        import dataclasses
        w = dataclasses.replace(w, k0=args.k0)
    ic = {} if args.ic == "modulated" else dict(ic="eigen", om=args.om, om_pc=args.om_pc, delta_n=args.delta_n)
    dt = args.dt if args.dt > 0 else w.dt
    window = tuple(float(x) for x in args.fit.split(",")) if args.fit else (0.5 * args.t_end, args.t_end)
    dev = torch.device("cuda:0")
    t, e64, f64, s64, _ = run_form(w, "f64", args.t_end, dt, args.graph_steps, dev, **ic)
    g64 = growth_rate(t, e64, window)
    ge64 = envelope_rate(t, e64, e64)
    nf64 = torch.linalg.vector_norm(f64).item()
    out = {"workload": w.name, "dof": w.dof, "k0": w.k0, "ic": args.ic if not ic else
           {"ic": "eigen", "om": str(args.om), "om_pc": args.om_pc, "delta_n": args.delta_n},
           "dt": dt, "t_end": args.t_end, "steps": len(t) - 1,
           "fit_window": list(window),
           "f64": {"seconds": s64, "gamma": g64, "gamma_env": ge64, "energy_end": float(e64[-1]),
                   "energy_range": [float(e64.min()), float(e64.max())]},
           "forms": {}}
    series = {"t": t, "f64": e64}
    for storage in [s for s in args.forms.split(",") if s]:
        _, e, f, sec, _ = run_form(w, storage, args.t_end, dt, args.graph_steps, dev, mfma=not args.no_mfma, **ic)
        series[storage] = e
        g = growth_rate(t, e, window)
        ge = envelope_rate(t, e, e64)
        out["forms"][storage] = {
            "seconds": sec,
            "rel_l2_f": torch.linalg.vector_norm(f - f64).item() / nf64,
            "rel_energy_end": abs(float(e[-1]) - float(e64[-1])) / abs(float(e64[-1])),
            "rel_l2_energy": float(np.linalg.norm(e - e64) / np.linalg.norm(e64)),
            "gamma": g,
            "rel_err_gamma": abs(g - g64) / abs(g64) if g64 != 0 else None,
            "gamma_env": ge,
            "rel_err_gamma_env": abs(ge - ge64) / abs(ge64) if ge64 != 0 else None,
            "finite": bool(np.all(np.isfinite(e))),
        }
        del f
        torch.cuda.empty_cache()
    if args.save:
        np.savez(args.save, **series)
    print(json.dumps(out), flush=True)
    return out


if __name__ == "__main__":
    main()
