"""RK3 DOF-updates/s of the GPU SSP-RK3 path on one MI355X.

    python bench.py [--steps K] [--warmup W] [--workload ring64] [--no-cpu-baseline] [--no-ensemble]

A step is one full Shu-Osher SSP-RK3 step (3 fused stage kernels + moment
reduction + spectral Poisson solve per stage) over the 116.64M-DOF workload
(n=9 Lobatto nodes per axis, 64 x 50 x 50 elements).  The state is resident
in HBM before the timed region starts.  K steps are captured in one HIP graph;
after one untimed replay, HIP events on the launch stream bracket exactly one
replay (K steps).  Per-stage kernel times come from K eager steps with events
recorded around each stage kernel.  Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0       # MI355X HBM3E, published


def cpu_baseline(w, seconds_target=15.0):
    """Time the NumPy oracle on a bounded sample of the same workload: same n,
    NU, NV, physics; fewer x elements (per-DOF cost does not depend on NX)."""
    import numpy as np
    import oracle
    try:
        from threadpoolctl import threadpool_info
        threads = max([i.get("num_threads", 1) for i in threadpool_info()] or [1])
    except Exception:
        threads = 1
    nx = 2
    case = oracle.Case(n=w.n, nx=nx, nu=w.nu, nv=w.nv, k0=w.k0, B=w.b_field, ring_j=w.ring_j, seed=w.seed)
    st = oracle.Stepper(w.dt)
    dof = nx * w.nu * w.nv * w.n ** 3
    t0 = time.perf_counter()
    steps = 0
    while True:
        st.nonlinear_ssp_rk(case.dist, case.basis, case.elliptic, case.grids, case.flux)
        steps += 1
        if time.perf_counter() - t0 > seconds_target or steps >= 50:
            break
    dt = time.perf_counter() - t0
    out = {"value": dof * steps / dt, "unit": "DOF-updates/s", "cores": int(threads), "kind": "port",
           "sample": f"NumPy FP64 oracle, {steps} SSP-RK3 steps of the same workload with NX={nx} "
                     f"instead of {w.nx} ({dof} DOF), {dt:.1f} s"}
    # the same sample on one thread (BLAS pools limited to 1), a few seconds
    try:
        from threadpoolctl import threadpool_limits
        with threadpool_limits(limits=1):
            t0 = time.perf_counter()
            steps1 = 0
            while True:
                st.nonlinear_ssp_rk(case.dist, case.basis, case.elliptic, case.grids, case.flux)
                steps1 += 1
                if time.perf_counter() - t0 > seconds_target / 3 or steps1 >= 20:
                    break
            dt1 = time.perf_counter() - t0
        out["single_thread"] = {"value": dof * steps1 / dt1, "cores": 1,
                                "sample": f"{steps1} steps of the same sample, {dt1:.1f} s"}
    except Exception:
        pass
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="ring64")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--storage", default="all", choices=["all", "f64", "f32", "bf16", "f16"],
                    help="storage form(s) to time; the headline value is the FP64 form")
    ap.add_argument("--layout", default="elem", choices=["elem", "ref"],
                    help="resident element-major layout (default) or the reference layout")
    ap.add_argument("--xseg", type=int, default=0,
                    help="x elements each stage-kernel workgroup walks (0: the library default)")
    ap.add_argument("--no-mfma", action="store_true",
                    help="run the BF16/FP16 stages on the vector ALUs instead of the matrix cores")
    ap.add_argument("--rs-kernel", type=int, default=0, choices=[0, 1, 2],
                    help="row-split stage kernel of the 16-bit forms at n = 9 (0: the library default)")
    ap.add_argument("--no-ensemble", action="store_true",
                    help="skip the 48-run mixed-shape FP32 ensemble (one persistent launch per RK stage "
                         "for all runs, autotuned schedule), which is timed by default and reported as "
                         "`ensemble`")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a plain 1 GiB device-to-device copy (reference kernel for "
                         "scaling FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        a = torch.ones(1 << 27, dtype=torch.float64, device="cuda:0")   # 1 GiB
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": a.numel() * 8, "copy_bytes_written": b.numel() * 8}))
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); "
              "the multi-GPU x-slab decomposition is a design, not a build", file=sys.stderr)
        sys.exit(2)

    import torch
    from vlasov1d2v_amd import CONFIGS
    from vlasov1d2v_amd.grid import STORAGE_CODES

    w = CONFIGS[args.workload]
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    forms = {}
    storages = ["f64", "f32", "bf16", "f16"] if args.storage == "all" else [args.storage]
    for storage in storages:
        forms[storage] = measure(w, storage, args.steps, args.warmup, args.layout, dev, mfma=not args.no_mfma,
                                 xseg=args.xseg, rs_kernel=args.rs_kernel)
        torch.cuda.empty_cache()
    head = forms["f64"] if "f64" in forms else forms[storages[0]]
    N = w.dof
    s_bytes = head["bytes_per_value"]
    # dominant kernel: the stage-2/3 launches (k_stage MODE 2: read the stage
    # input and f^n, write the stage output) -> 3 N s algorithmic bytes each
    launch_bytes = 3 * N * s_bytes
    launch_ms = 0.5 * (head["stage_ms"][1] + head["stage_ms"][2])
    ceiling = stream_ceiling(N * s_bytes, dev)
    achieved = launch_bytes / (launch_ms / 1e3) / 1e9
    code = STORAGE_CODES[head["storage"]]
    if head["storage"] in ("bf16", "f16") and w.n in (8, 9) and args.layout == "elem":
        kname = f"k_stage_rs<{w.n},2,{code}>"        # the row-split matrix-core kernel
    else:
        kname = f"k_stage<{w.n},2,{code},{1 if args.layout == 'elem' else 0}>"
    traffic = profiled_traffic(kname, head["storage"])
    out = {
        "metric": "RK3 DOF-updates/s at 116.6M DOF, FP64 vs BF16-store; % HBM roofline",
        "value": head["dof_updates_per_s"],
        "unit": "DOF-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": head["storage"],
        "data": "synthetic (seeded ring distribution with density modulation)",
        "config": {"workload": w.name, "dof": w.dof, "n": w.n, "elements": [w.nx, w.nu, w.nv],
                   "dt": w.dt, "storage": head["storage"], "layout": args.layout,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "stream_ceiling": ceiling, "frac_of_stream_ceiling": achieved / ceiling["achieved"],
                     "kernel": kname + " (fused RHS + stage combination + moment partials; stages 2 and 3)",
                     "launch_ms": launch_ms, "algorithmic_bytes_per_launch": launch_bytes,
                     "stage_ms": head["stage_ms"]},
        "forms": {k: {kk: v[kk] for kk in ("dof_updates_per_s", "ms_per_step", "ms_per_step_eager", "stage_ms",
                                           "stage23_hbm_frac", "step_hbm_frac", "finite")}
                  for k, v in forms.items()},
    }
    if "bf16" in forms and "f64" in forms:
        out["bf16_over_f64"] = forms["bf16"]["dof_updates_per_s"] / forms["f64"]["dof_updates_per_s"]
    if not args.no_ensemble and args.workload == "ring64":
        from vlasov1d2v_amd.ensemble import run_workload
        torch.cuda.empty_cache()
        ens = run_workload(runs=48, seed=0, steps=200, graph_steps=20, storage="f32")
        out["ensemble"] = {k: ens[k] for k in ("runs", "dof", "storage", "steps", "dof_updates_per_s",
                                               "ms_per_step", "schedule", "launches_per_step")}
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(w)
    print(json.dumps(out), flush=True)


def stream_ceiling(array_bytes, dev, reps=10):
    """Best rate (GB/s, 3 array_bytes moved per launch) of the streaming probe
    vl_stream_probe (two non-temporal reads, one non-temporal write, 16-byte
    vectors) over a few grid sizes, timed with events on the launch stream."""
    import torch
    from vlasov1d2v_amd import _lib
    n = array_bytes // 16 * 2
    a = torch.ones(n, dtype=torch.float64, device=dev)
    b = torch.ones(n, dtype=torch.float64, device=dev)
    out = torch.empty(n, dtype=torch.float64, device=dev)
    lib = _lib.load()
    rates = {}
    for nt in (1, 0):
        for grid in (4096, 16384, 32768, 65536, 131072):
            def launch():
                _lib.check(lib.vl_stream_probe(_lib.ptr(a), _lib.ptr(b), _lib.ptr(out), n * 8, grid, nt,
                                               _lib.stream_ptr()), "vl_stream_probe")
            launch()
            t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            t0.record()
            for _ in range(reps):
                launch()
            t1.record()
            torch.cuda.synchronize(dev)
            rates[(nt, grid)] = 3 * n * 8 * reps / (t0.elapsed_time(t1) / 1e3) / 1e9
    ok = bool(torch.all(out == 2.0).item())
    del a, b, out
    torch.cuda.empty_cache()
    (nt, grid), best = max(rates.items(), key=lambda kv: kv[1])
    return {"kernel": "k_stream_probe (2 reads + 1 write, 16-byte vectors)", "achieved": best,
            "unit": "GB/s", "nontemporal": bool(nt), "grid": grid, "array_bytes": n * 8, "checked": ok,
            "best_nontemporal": max(v for (t, _), v in rates.items() if t),
            "best_default_policy": max(v for (t, _), v in rates.items() if not t)}


def profiled_traffic(kernel, storage):
    """HBM bytes per launch of `kernel` from the committed FETCH_SIZE /
    WRITE_SIZE profile of this workload (counter readings in KiB per
    dispatch, scaled by the copy-kernel calibration); None if not profiled."""
    path = os.path.join(ROOT, "profiles", "k_stage_ring64_n9_64x50x50.md")
    if not os.path.exists(path):
        return None
    cal, rows = None, {}
    with open(path) as fh:
        for line in fh:
            t = line.split()
            if len(t) == 5 and t[0] == "calibration":
                cal = (float(t[2]), float(t[4]))
            elif len(t) == 5 and t[0] == "traffic":
                rows[(t[1], t[2])] = (float(t[3]), float(t[4]))
    if cal is None or (kernel, storage) not in rows:
        return None
    fetch, write = rows[(kernel, storage)]
    return (fetch * cal[0] + write * cal[1]) * 1024.0


def measure(w, storage, K, W, layout, dev, mfma=True, xseg=0, rs_kernel=0):
    """Time K SSP-RK3 steps (after W warmup steps) of one storage form."""
    import ctypes
    import torch
    from vlasov1d2v_amd import Run, _lib
    from vlasov1d2v_amd.grid import STORAGE_CODES

    run = Run.from_workload(w, device=dev, storage=storage)
    g = run.grids.geometry(STORAGE_CODES[storage])
    g.set_tuning(_lib.VL_TUNE_MFMA, 1 if mfma else 0)
    if xseg > 0:
        g.set_tuning(_lib.VL_TUNE_XSEG, xseg)
    if rs_kernel > 0:
        g.set_tuning(_lib.VL_TUNE_RS_KERNEL, rs_kernel)
    energies = torch.empty(W + K + 1, dtype=torch.float64, device=dev)
    if layout == "elem":
        run.resident_begin()
    for k in range(W):
        run.step(energy_out=energies[k:k + 1])
    torch.cuda.synchronize()
    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(6)] for _ in range(K)]
    for row in evs:
        for e in row:
            e.record()
    torch.cuda.synchronize()
    ev_arrays = [(ctypes.c_void_p * 6)(*[e.cuda_event for e in row]) for row in evs]
    lib = _lib.load()
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    start.record()
    for k in range(K):
        lib.vl_set_stage_events(g.handle, ctypes.cast(ev_arrays[k], ctypes.c_void_p))
        run.step(energy_out=energies[W + k:W + k + 1])
    stop.record()
    torch.cuda.synchronize()
    lib.vl_set_stage_events(g.handle, None)
    ms_eager = start.elapsed_time(stop)
    # the timed measurement: the same K steps captured in one HIP graph,
    # replayed once untimed, then exactly one timed replay
    graph = run.capture(K)
    graph.replay()
    torch.cuda.synchronize()
    start.record()
    graph.replay(check=False)
    stop.record()
    torch.cuda.synchronize()
    ms = start.elapsed_time(stop)
    graph.check()
    energies[W + K:W + K + 1].copy_(graph.energies[-1:])
    if layout == "elem":
        run.resident_end()
    stage_ms = [sum(row[2 * s].elapsed_time(row[2 * s + 1]) for row in evs) / K for s in range(3)]
    bpv = {"f64": 8, "f32": 4, "bf16": 2, "f16": 2}[storage]
    res = {"storage": storage, "bytes_per_value": bpv, "ms_per_step": ms / K, "ms_per_step_eager": ms_eager / K,
           "dof_updates_per_s": w.dof * K / (ms / 1e3), "stage_ms": stage_ms,
           "step_hbm_frac": 8 * w.dof * bpv / (ms / K / 1e3) / 1e9 / HBM_PEAK_GBS,
           "stage23_hbm_frac": 3 * w.dof * bpv / (0.5 * (stage_ms[1] + stage_ms[2]) / 1e3) / 1e9
                               / HBM_PEAK_GBS,
           "finite": bool(torch.isfinite(energies).all().item())}
    del run
    return res


if __name__ == "__main__":
    main()
