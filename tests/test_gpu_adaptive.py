"""Adaptive time step (main/timestep.py:88-89,184,231-242) evaluated on the
device: ``Stepper(dt=None).main_loop`` keeps dt and time in the workspace,
so adaptive loops replay HIP graphs.

The device rule forms the same IEEE divisions and sums as the host rule, and
the stage kernels form c2 * dt as the host does, so the device loop must
reproduce the host-adaptive loop bit for bit (f, energies, times), and both
agree with the oracle's adaptive loop at 1e-12.
"""
import numpy as np
import pytest

from helpers import interior, make_pair, rel_l2

pytestmark = pytest.mark.gpu


def _run(device_dt, steps=None, final_time=None, graph_steps=0, n=6, storage="f64"):
    from vlasov1d2v_amd import Run
    run = Run(n, 4, 5, 5, dt=None, storage=storage)
    if final_time is not None:
        run.stepper.final_time = final_time
    st = run.stepper
    st.main_loop(run.dist, run.basis, run.elliptic, run.grids, run.flux, steps=steps, graph_steps=graph_steps,
                 device_dt=device_dt)
    return run, st


@pytest.mark.parametrize("graph_steps", [0, 4])
def test_device_dt_loop_matches_host_dt_loop(cuda, graph_steps):
    import torch
    rh, sh = _run(False, steps=11)
    rd, sd = _run(True, steps=11, graph_steps=graph_steps)
    print(f"graph_steps={graph_steps}: host dt {sh.dt!r} device dt {sd.dt!r}; times {sd.time_array[-3:]}",
          flush=True)
    assert sd.steps_counter == sh.steps_counter == 11
    np.testing.assert_array_equal(sd.time_array, sh.time_array)
    np.testing.assert_array_equal(sd.field_energy, sh.field_energy)
    assert sd.dt == sh.dt and sd.time == sh.time
    assert torch.equal(rd.dist.arr, rh.dist.arr)


def test_device_dt_loop_stops_at_final_time(cuda):
    import torch
    # dt is about 0.045 here: 0.5 takes ~11 steps, so 4-step replays run and
    # the tail is stepped one at a time
    rh, sh = _run(False, final_time=0.5)
    rd, sd = _run(True, final_time=0.5, graph_steps=4)
    print(f"steps {sd.steps_counter} (host {sh.steps_counter}), final time {sd.time!r}", flush=True)
    assert sd.steps_counter == sh.steps_counter and sd.time_array[-2] < 0.5 <= sd.time_array[-1]
    np.testing.assert_array_equal(sd.time_array, sh.time_array)
    np.testing.assert_array_equal(sd.field_energy, sh.field_energy)
    assert torch.equal(rd.dist.arr, rh.dist.arr)


def test_device_dt_loop_matches_oracle(cuda):
    steps = 5
    case, run = make_pair(5, 3, 4, 4, dt=None)
    st_ref = case.run(steps, None)
    st = run.stepper
    st.main_loop(run.dist, run.basis, run.elliptic, run.grids, run.flux, steps=steps, graph_steps=2)
    err_f = rel_l2(interior(run.dist.arr.cpu().numpy()), interior(case.dist.arr))
    err_e = rel_l2(st.field_energy, st_ref.field_energy)
    err_t = float(np.max(np.abs(st.time_array - st_ref.time_array) / np.maximum(st_ref.time_array, 1e-300)))
    print(f"rel-L2 f {err_f:.2e}, energy {err_e:.2e}, time {err_t:.2e}; dt {st.dt!r} vs {st_ref.dt!r}", flush=True)
    assert err_f < 1e-12 and err_e < 1e-12 and err_t < 1e-13


def test_device_dt_reduced_storage_graph_matches_eager(cuda):
    """BF16 row-split storage: adaptive replays reproduce adaptive eager steps."""
    import torch
    ra, sa = _run(True, steps=9, n=8, storage="bf16")
    rb, sb = _run(True, steps=9, graph_steps=4, n=8, storage="bf16")
    np.testing.assert_array_equal(sa.time_array, sb.time_array)
    np.testing.assert_array_equal(sa.field_energy, sb.field_energy)
    assert torch.equal(ra.dist.arr, rb.dist.arr)
