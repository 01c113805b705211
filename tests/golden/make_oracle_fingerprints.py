"""Write compact fingerprints of oracle runs at the two benchmark sizes.

    python tests/golden/make_oracle_fingerprints.py [ring16] [ring64]

These are outputs of this project's NumPy oracle (oracle/), not of the
reference: the full state is too large to commit, so each fixture keeps the
field-energy series, and at a few checkpoint steps the L2 norm of the interior
of f, the charge density and f at 4096 seeded random interior positions.  The
GPU tests run the same workloads and compare against them
(tests/test_gpu_workloads.py).  ring64 runs the configured 200 steps at the
stable dt = 0.002 (about 30 s of oracle time per step).
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402

RUNS = {
    # name: (n, nx, nu, nv, steps, dt, checkpoint steps)
    "ring16": (8, 16, 16, 16, 50, 0.008, (2, 10, 50)),
    "ring64": (9, 64, 50, 50, 200, 0.002, (2, 20, 50, 100, 200)),
}
SAMPLE_SEED = 1234
SAMPLES = 4096


def sample_index(shape, seed=SAMPLE_SEED, count=SAMPLES):
    """Seeded flat indices into the interior block of shape `shape` (interior dims)."""
    rng = np.random.default_rng(seed)
    return np.sort(rng.choice(int(np.prod(shape)), size=count, replace=False))


def fingerprint(case, step):
    f = case.dist.arr[1:-1, :, 1:-1, :, 1:-1, :]
    idx = sample_index(f.shape)
    return {
        "step": step,
        "f_interior_l2": float(np.linalg.norm(f)),
        "density": [float(v) for v in case.dist.moment_zero().reshape(-1)],
        "sample_values": [float(v) for v in f.reshape(-1)[idx]],
    }


def main(names):
    for name in names:
        n, nx, nu, nv, steps, dt, checkpoints = RUNS[name]
        t0 = time.time()
        case = oracle.Case(n=n, nx=nx, nu=nu, nv=nv, seed=0)
        marks = []

        def on_step(st, case=case, marks=marks, checkpoints=checkpoints, t0=t0):
            if st.steps_counter in checkpoints:
                marks.append(fingerprint(case, st.steps_counter))
                print(f"{name}: step {st.steps_counter} at {time.time() - t0:.0f} s", flush=True)

        st = oracle.Stepper(dt)
        st.main_loop(steps, case.dist, case.basis, case.elliptic, case.grids, case.flux, on_step=on_step)
        last = marks[-1]
        assert last["step"] == steps
        out = {
            "workload": name, "n": n, "elements": [nx, nu, nv], "steps": steps, "dt": dt, "seed": 0,
            "k0": 0.1, "b_field": 0.1, "ring_j": 6, "amp1": 1e-3, "amps": [1e-6, 1e-6, 1e-6],
            "phases": [float(p) for p in case.phases],
            "field_energy": [float(e) for e in st.field_energy],
            "f_interior_l2": last["f_interior_l2"],
            "sample_seed": SAMPLE_SEED,
            "sample_values": last["sample_values"],
            "checkpoints": marks,
            "oracle_seconds": time.time() - t0,
        }
        path = os.path.join(HERE, f"{name}_oracle.json")
        with open(path + ".tmp", "w") as fh:
            json.dump(out, fh, indent=0)
        os.replace(path + ".tmp", path)
        print(f"{name}: {steps} steps in {out['oracle_seconds']:.1f} s -> {path}", flush=True)


if __name__ == "__main__":
    main(sys.argv[1:] or ["ring16"])
