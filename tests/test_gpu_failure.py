"""Run-time failure detection and the configured (unstable) time step.

The 2.1M-DOF workload (n = 8, 16^3 elements) at its configured dt = 0.02 is
above the stability limit of the upwind DG + SSP-RK3 scheme: grid-scale modes
grow from rounding until the state overflows.  The oracle's energy series of
that run (tests/golden/ring16_dt002_blowup.json) is finite through step 45.
The GPU run follows it to 1e-10 while rounding noise is small, overflows at
the same step to within a few steps, and the workspace status word turns the
overflow into a VlasovError instead of a silent NaN series.
"""
import json
import math
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def _fixture():
    with open(os.path.join(HERE, "golden", "ring16_dt002_blowup.json")) as fh:
        return json.load(fh)


def test_configured_dt_blows_up_like_the_oracle(cuda):
    import torch
    from vlasov1d2v_amd import Run, _lib
    fx = _fixture()
    stop = fx["first_non_finite_step"]
    assert stop is not None
    run = Run(8, 16, 16, 16, seed=0, dt=0.02, device=cuda)
    g = run.flux._geom(run.grids, run.elliptic)
    e0 = float(run.elliptic.electric_energy(run.grids.x).item())
    energies, flagged = [e0], None
    run.resident_begin()
    e = torch.empty(1, dtype=torch.float64, device=cuda)
    for k in range(1, stop + 8):
        run.step(energy_out=e)
        energies.append(float(e.item()))
        if g.read_status(cuda) & _lib.VL_STATUS_NONFINITE:
            flagged = k
            break
    run.resident_end()
    ref = fx["field_energy"]
    print(f"oracle non-finite at step {stop}; GPU flagged at {flagged}", flush=True)
    # rounding-level agreement before the unstable modes dominate
    early = np.asarray(ref[:31])
    assert np.max(np.abs(np.asarray(energies[:31]) - early) / early) < 1e-10
    # both blow up at the same point
    assert flagged is not None and abs(flagged - stop) <= 3
    assert all(math.isfinite(x) for x in energies[:stop - 4])
    assert max(x for x in energies[:flagged] if math.isfinite(x)) > 1e3 * e0


def test_time_loop_raises_on_overflow(cuda):
    """Stepper.main_loop with graph replays stops with a VlasovError."""
    from vlasov1d2v_amd import Run, _lib
    run = Run(8, 16, 16, 16, seed=0, dt=0.02, device=cuda)
    with pytest.raises(_lib.VlasovError, match="not finite"):
        run.run(80, graph_steps=10)


def test_nan_in_the_state_is_detected(cuda):
    import torch
    from vlasov1d2v_amd import Run, _lib
    run = Run(5, 3, 4, 4, dt=0.005, device=cuda)
    g = run.flux._geom(run.grids, run.elliptic)
    run.step()
    g.check_finite(cuda, "clean step")
    run.dist.arr[2, 1, 2, 2, 2, 3] = float("nan")
    run.step()
    torch.cuda.synchronize()
    with pytest.raises(_lib.VlasovError, match="not finite"):
        g.check_finite(cuda, "step after NaN injection")
    # the bit is sticky until cleared
    assert g.read_status(cuda, clear=True) & _lib.VL_STATUS_NONFINITE
    assert g.read_status(cuda) == 0
