"""Ensembles of independent runs stepped together on one GPU.

Each member is a full run of the path -- its own n, grid, ring index, k0 and
magnetic field, built exactly as a single run is (main/main.py:37-166 via
``Run``) -- and the ensemble steps all of them with one launch per kernel
kind and RK stage (``vl_ensemble_rk3_step``: 13 launches per ensemble step),
captured K steps at a time in a HIP graph.  A run of an ensemble computes
what ``Stepper.nonlinear_ssp_rk`` computes for it alone, to rounding (the
same kernel statements; the compiler may fuse multiply-adds differently).
The runs' potentials (``elliptic.potential``) are not updated.

    python -m vlasov1d2v_amd.ensemble [--runs 48] [--steps 200] [--storage f32] [--check-f64]
                                      [--schedule work,grid,order]

runs the seeded ensemble workload (48 runs; NX in {16, 32, 48, 64}, NU = NV
in [24, 48], n in {6, 7, 8, 9}, ring index j in [1, 8], k0 in [0.05, 0.2],
omega_c/omega_p in [0.05, 0.2]), autotunes the stage schedule, and prints one
JSON line with the aggregate DOF-updates/s and, with --check-f64, every run's
relative error against the same run in FP64.
"""
import argparse
import ctypes
import dataclasses
import gc
import json
import time

import numpy as np

from . import _lib
from .timestep import COURANT_NUMBERS, get_max_speeds
from .workload import Run


@dataclasses.dataclass(frozen=True)
class Member:
    n: int
    nx: int
    nu: int
    nv: int
    ring_j: int
    k0: float
    b_field: float
    seed: int

    @property
    def dof(self):
        return self.nx * self.nu * self.nv * self.n ** 3


def draw_members(count=48, seed=0):
    """Seeded draw of the ensemble workload (one numpy Generator, members in order)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        nx = int(rng.choice([16, 32, 48, 64]))
        nuv = int(rng.integers(24, 49))
        n = int(rng.integers(6, 10))
        j = int(rng.integers(1, 9))
        k0 = float(rng.uniform(0.05, 0.2))
        b = float(rng.uniform(0.05, 0.2))
        out.append(Member(n=n, nx=nx, nu=nuv, nv=nuv, ring_j=j, k0=k0, b_field=b, seed=1000 * seed + i))
    return out


def cfl_dt(run):
    """The reference's adaptive step at the run's current field
    (main/timestep.py:231-242, Courant number of the 3-stage row)."""
    g = run.grids
    speeds = get_max_speeds(g, run.elliptic, run.flux.acceleration_coefficient)
    c = COURANT_NUMBERS[3][g.x.order - 1]
    return c / (speeds[0] / g.x.dx + speeds[1] / g.u.dx + speeds[2] / g.v.dx)


class Ensemble:
    """Members stepped together; states stay resident (element-major) until ``finish()``.

    ``dt``: one fixed step for every run, a sequence of one per run, None
    for 0.85 x each run's CFL step at t = 0 (the margin the single-run
    workloads keep below the rule), or "adaptive" for the reference's CFL
    rule after every step of every run (main/timestep.py:184,231-242),
    evaluated on the device per run (``clocks()`` reads each run's time and dt).
    """

    def __init__(self, members, storage="f32", device="cuda", dt=None, safety=0.85):
        import torch
        if storage not in ("f64", "f32"):
            raise ValueError("ensembles run FP64 or FP32 storage")
        self.members = list(members)
        self.storage = storage
        self.device = torch.device(device)
        self.runs = []
        self.dts = []
        self.adaptive = isinstance(dt, str) and dt == "adaptive"
        for m in self.members:
            run = Run(m.n, m.nx, m.nu, m.nv, k0=m.k0, b_field=m.b_field, ring_j=m.ring_j, seed=m.seed,
                      dt=1.0, device=self.device, storage=storage)
            if self.adaptive:
                step = cfl_dt(run)          # the rule at t = 0 (the clock starts from it)
            elif dt is None:
                step = safety * cfl_dt(run)
            elif np.ndim(dt) == 0:
                step = float(dt)
            else:
                step = float(dt[len(self.runs)])
            run.stepper.dt = run.stepper.fixed_dt = step
            self.runs.append(run)
            self.dts.append(step)
        self.energy0 = torch.stack([r.elliptic.electric_energy(r.grids.x).reshape(()) for r in self.runs])
        lib = _lib.load()
        self._geoms = []
        for run in self.runs:
            run.resident_begin(layout=_lib.VL_LAYOUT_ELEMENT_MAJOR)
            self._geoms.append(run.stepper._res["geom"])
        handles = (ctypes.c_void_p * len(self.runs))(*[g.handle for g in self._geoms])
        h = ctypes.c_void_p()
        _lib.check(lib.vl_ensemble_create(handles, len(self.runs), ctypes.byref(h)), "vl_ensemble_create")
        self.handle = h
        for r, run in enumerate(self.runs):
            f, f1, f2 = run.stepper._res["bufs"]
            ws = self._geoms[r].workspace(self.device)
            _lib.check(lib.vl_ensemble_bind(h, r, _lib.ptr(f), _lib.ptr(f1), _lib.ptr(f2),
                                            _lib.ptr(run.elliptic.electric_field), _lib.ptr(ws),
                                            self.dts[r]), "vl_ensemble_bind")
        if self.adaptive:
            for r, run in enumerate(self.runs):
                g = self._geoms[r]
                gr = run.grids
                g.set_cfl(run.stepper.courant, gr.x.dx, gr.u.dx, gr.v.dx)
                _lib.check(lib.vl_clock_init(g.handle, _lib.ptr(run.elliptic.electric_field), 0.0,
                                             _lib.ptr(g.workspace(self.device)), _lib.stream_ptr()),
                           "vl_clock_init")
                _lib.check(lib.vl_ensemble_set_adaptive(h, r, 1), "vl_ensemble_set_adaptive")
        _lib.check(lib.vl_ensemble_moments(h, _lib.stream_ptr()), "vl_ensemble_moments")
        self.steps_done = 0
        self.tuning = None
        self.schedule_generation = 0     # bumped when the schedule changes (captured graphs go stale)

    @property
    def dof(self):
        return sum(m.dof for m in self.members)

    def schedule(self):
        w, g, o, it = (ctypes.c_int() for _ in range(4))
        _lib.check(_lib.load().vl_ensemble_get_schedule(self.handle, ctypes.byref(w), ctypes.byref(g),
                                                        ctypes.byref(o), ctypes.byref(it)),
                   "vl_ensemble_get_schedule")
        return {"work_per_item": w.value, "grid": g.value, "order": o.value, "items": it.value}

    def set_schedule(self, work_per_item, grid=0, order=0):
        _lib.check(_lib.load().vl_ensemble_set_schedule(self.handle, int(work_per_item), int(grid), int(order)),
                   "vl_ensemble_set_schedule")
        self.schedule_generation += 1

    def autotune(self, reps=2):
        """Time the candidate schedules of the stage kernel and keep the fastest."""
        table = (ctypes.c_float * (4 * 64))()
        rows = ctypes.c_int()
        _lib.check(_lib.load().vl_ensemble_autotune(self.handle, int(reps), table, 64, ctypes.byref(rows),
                                                    _lib.stream_ptr()), "vl_ensemble_autotune")
        self.schedule_generation += 1
        t = np.frombuffer(table, dtype=np.float32)[:4 * rows.value].reshape(-1, 4)
        self.tuning = [{"work_per_item": int(a), "grid": int(b), "order": int(c), "ms": float(d)}
                       for a, b, c, d in t]
        return self.schedule()

    def step(self, energy_out=None):
        """One SSP-RK3 step of every run; energy_out: (runs,) FP64 GPU tensor or None."""
        _lib.check(_lib.load().vl_ensemble_rk3_step(self.handle, _lib.ptr(energy_out), _lib.stream_ptr()),
                   "vl_ensemble_rk3_step")
        self.steps_done += 1

    def clocks(self):
        """(time, dt) of every adaptive run after the steps so far, (runs, 2) on the host."""
        import torch
        out = torch.zeros((len(self.runs), 2), dtype=torch.float64, device=self.device)
        _lib.check(_lib.load().vl_ensemble_clocks(self.handle, _lib.ptr(out), _lib.stream_ptr()),
                   "vl_ensemble_clocks")
        return out.cpu().numpy()

    def check_finite(self):
        """Raise VlasovError naming the runs whose steps met a non-finite
        field (each run's workspace status; one synchronisation per run)."""
        bad = [r for r, g in enumerate(self._geoms) if g.read_status(self.device) & _lib.VL_STATUS_NONFINITE]
        if bad:
            raise _lib.VlasovError(f"ensemble runs {bad}: the field energy or max|acc E| of a stage-3 field solve "
                                   "is not finite (the state has overflowed or holds NaN)")

    def capture(self, steps):
        """A HIP graph of ``steps`` ensemble steps (step once eagerly first)."""
        return EnsembleGraph(self, steps)

    def run(self, steps, graph_steps=0):
        """``steps`` steps; returns the energy series (steps + 1, runs) on the host."""
        import torch
        en = torch.empty((steps + 1, len(self.runs)), dtype=torch.float64, device=self.device)
        en[0] = self.energy0
        k = 0
        graph = None
        while k < steps:
            if graph_steps > 0 and self.steps_done > 0 and steps - k >= graph_steps:
                if graph is None:
                    graph = self.capture(graph_steps)
                graph.replay()
                en[k + 1:k + 1 + graph_steps].copy_(graph.energies)
                k += graph_steps
            else:
                self.step(en[k + 1])
                k += 1
        torch.cuda.synchronize(self.device)
        self.check_finite()
        return en.cpu().numpy()

    def finish(self):
        """Write every run's state back to its distribution (reference layout)."""
        for run in self.runs:
            run.resident_end()

    def close(self):
        if getattr(self, "handle", None):
            _lib.load().vl_ensemble_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class EnsembleGraph:
    """``steps`` ensemble steps captured once; ``replay()`` advances every run
    by ``steps`` steps and leaves the energies in ``energies`` (steps, runs).
    The graph holds the schedule of its capture: ``replay()`` raises once
    ``set_schedule`` or ``autotune`` has changed it (capture again)."""

    def __init__(self, ens, steps):
        import torch
        if ens.steps_done == 0:
            raise RuntimeError("step the ensemble once eagerly before capturing")
        dev = ens.device
        self.steps = int(steps)
        self.ens = ens
        self.schedule_generation = ens.schedule_generation
        self.energies = torch.zeros((self.steps, len(ens.runs)), dtype=torch.float64, device=dev)
        side = torch.cuda.Stream(device=dev)
        side.wait_stream(torch.cuda.current_stream(dev))
        self.graph = _lib.HipGraph()
        gc.collect()
        gc_was_enabled = gc.isenabled()
        gc.disable()
        try:
            with self.graph.capture(side):
                for k in range(self.steps):
                    _lib.check(_lib.load().vl_ensemble_rk3_step(ens.handle, _lib.ptr(self.energies[k]),
                                                                _lib.stream_ptr(side)), "vl_ensemble_rk3_step")
        finally:
            if gc_was_enabled:
                gc.enable()
        torch.cuda.current_stream(dev).wait_stream(side)

    def replay(self):
        if self.ens.schedule_generation != self.schedule_generation:
            raise _lib.VlasovError("the ensemble schedule changed after this graph was captured: capture again")
        self.graph.launch()
        self.ens.steps_done += self.steps


def time_steps(ens, steps, graph_steps):
    """Time ``steps`` ensemble steps replayed from one graph of ``graph_steps``
    steps (after one eager step).  Returns (seconds of the replays, steps
    replayed, energy series (1 + 1 + steps, runs) on the host: the initial
    field, the eager step, the replayed steps)."""
    import torch
    reps = max(1, steps // graph_steps)
    series = torch.empty((2 + reps * graph_steps, len(ens.runs)), dtype=torch.float64, device=ens.device)
    series[0] = ens.energy0
    ens.step(series[1])
    graph = ens.capture(graph_steps)
    torch.cuda.synchronize(ens.device)
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    t0.record()
    for k in range(reps):
        graph.replay()
        series[2 + k * graph_steps:2 + (k + 1) * graph_steps].copy_(graph.energies)
    t1.record()
    torch.cuda.synchronize(ens.device)
    ens.check_finite()
    return t0.elapsed_time(t1) / 1e3, reps * graph_steps, series.cpu().numpy()


def run_workload(runs=48, seed=0, steps=200, graph_steps=20, storage="f32", autotune=True, check_f64=False,
                 device="cuda:0", schedule=None):
    """The seeded ensemble workload: build, autotune (or take ``schedule`` =
    (work_per_item, grid, order)), time ``steps`` steps from graph replays;
    optionally every member's error against FP64."""
    import torch
    members = draw_members(runs, seed)
    dev = torch.device(device)
    t0 = time.perf_counter()
    ens = Ensemble(members, storage=storage, device=dev)
    setup_s = time.perf_counter() - t0
    if schedule is not None:
        ens.set_schedule(*schedule)
    sched = ens.autotune() if autotune and schedule is None else ens.schedule()
    sec, done, series = time_steps(ens, steps, graph_steps)
    out = {"metric": "ensemble RK3 DOF-updates/s", "runs": len(members), "seed": seed, "dof": ens.dof,
           "storage": storage, "steps": done, "seconds": sec,
           "dof_updates_per_s": ens.dof * done / sec, "ms_per_step": 1e3 * sec / done,
           "schedule": sched, "setup_s": setup_s, "launches_per_step": 13,
           "members": [dataclasses.asdict(m) for m in members]}
    if ens.tuning:
        out["autotune"] = ens.tuning
    if check_f64:
        # same members and dt, same step count (one eager step + graph replays)
        total = ens.steps_done
        ens.finish()
        ref = Ensemble(members, storage="f64", device=dev, dt=ens.dts)
        ref_series = ref.run(total, graph_steps=graph_steps)
        ref.finish()
        errs, eerrs = [], []
        for r, (a, b) in enumerate(zip(ens.runs, ref.runs)):
            fa = a.dist.total()[1:-1, :, 1:-1, :, 1:-1, :]
            fb = b.dist.arr[1:-1, :, 1:-1, :, 1:-1, :]
            errs.append(torch.linalg.vector_norm(fa - fb).item() / torch.linalg.vector_norm(fb).item())
            eerrs.append(float(np.linalg.norm(series[:, r] - ref_series[:, r]) / np.linalg.norm(ref_series[:, r])))
        out["rel_l2_f_vs_f64"] = {"max": max(errs), "median": float(np.median(errs)), "per_run": errs}
        out["rel_l2_energy_vs_f64"] = {"max": max(eerrs), "median": float(np.median(eerrs)), "per_run": eerrs}
        ref.close()
    ens.close()
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--runs", type=int, default=48)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--graph-steps", type=int, default=20)
    ap.add_argument("--storage", default="f32")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--schedule", default="", help="work_per_item,grid,order instead of autotuning")
    ap.add_argument("--check-f64", action="store_true",
                    help="also run every member in FP64 and report per-run errors")
    args = ap.parse_args(argv)
    sched = tuple(int(v) for v in args.schedule.split(",")) if args.schedule else None
    out = run_workload(args.runs, args.seed, args.steps, args.graph_steps, args.storage,
                       autotune=not args.no_autotune, check_f64=args.check_f64, schedule=sched)
    print(json.dumps(out), flush=True)
    return out


if __name__ == "__main__":
    main()
