"""Write tests/golden/ring16_dt002_blowup.json: the oracle's field-energy
series of the 2.1M-DOF workload (n = 8, 16 x 16 x 16) at the configured
dt = 0.02, which is above the stability limit of the upwind DG + SSP-RK3
scheme (about 1.26x growth per step of grid-scale modes): stepped until the
energy is no longer finite or `max_steps` is reached.

    python tests/golden/make_blowup_fixture.py [max_steps]

Output of this project's NumPy oracle (oracle/), not of the reference.
Entries are floats, or null once non-finite.
"""
import json
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402


def main(max_steps=6000):
    t0 = time.time()
    case = oracle.Case(n=8, nx=16, nu=16, nv=16, seed=0)
    st = oracle.Stepper(0.02)
    done = {"stop": None}

    class Stop(Exception):
        pass

    def on_step(s):
        e = s.field_energy[-1]
        if s.steps_counter % 100 == 0:
            print(f"step {s.steps_counter}: energy {e:.3e} ({time.time() - t0:.0f} s)", flush=True)
        if not math.isfinite(e) or not np.isfinite(case.dist.arr).all():
            done["stop"] = s.steps_counter
            raise Stop()

    with np.errstate(all="ignore"):
        try:
            st.main_loop(max_steps, case.dist, case.basis, case.elliptic, case.grids, case.flux, on_step=on_step)
        except Stop:
            pass
    energy = [float(e) if math.isfinite(e) else None for e in st.field_energy]
    out = {"workload": "ring16", "n": 8, "elements": [16, 16, 16], "dt": 0.02, "seed": 0,
           "steps": len(energy) - 1, "first_non_finite_step": done["stop"], "field_energy": energy,
           "oracle_seconds": time.time() - t0}
    path = os.path.join(HERE, "ring16_dt002_blowup.json")
    with open(path, "w") as fh:
        json.dump(out, fh, indent=0)
    print(f"{out['steps']} steps, first non-finite at {done['stop']} -> {path}", flush=True)


if __name__ == "__main__":
    main(int(sys.argv[1]) if len(sys.argv) > 1 else 6000)
