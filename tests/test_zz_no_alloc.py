"""vl_rk3_step_noalloc -- the SSP-RK3 step whose workspace is the library's
own device array -- against vl_rk3_step with a caller workspace on the same
inputs: bit-identical state, field, potential and energy, on two calls in a
row (the second on a smaller grid, so the array holds the first call's
larger layout), and a clear refusal of a grid too large for the array."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _run(cuda, n, nx, nu, nv, dt):
    from vlasov1d2v_amd import Run
    return Run(n, nx, nu, nv, dt=dt, device=cuda)


def _step(run, noalloc, steps):
    import torch
    from vlasov1d2v_amd import _lib
    lib = _lib.load()
    g = run.flux._geom(run.grids, run.elliptic)
    f = run.dist.arr
    f1, f2 = torch.empty_like(f), torch.empty_like(f)
    E = run.elliptic.electric_field
    phi = torch.zeros((run.grids.x.res, run.grids.x.order), dtype=torch.float64, device=f.device)
    energy = torch.zeros(steps, dtype=torch.float64, device=f.device)
    for k in range(steps):
        e_k = energy[k:k + 1]
        if noalloc:
            rc = lib.vl_rk3_step_noalloc(g.handle, _lib.ptr(f), _lib.ptr(f1), _lib.ptr(f2), _lib.ptr(E),
                                         _lib.ptr(phi), float(run.stepper.dt), _lib.ptr(e_k), 0, _lib.stream_ptr())
            _lib.check(rc, "vl_rk3_step_noalloc")
        else:
            ws = g.workspace(f.device)
            rc = lib.vl_rk3_step(g.handle, _lib.ptr(f), _lib.ptr(f1), _lib.ptr(f2), _lib.ptr(E), _lib.ptr(phi),
                                 float(run.stepper.dt), _lib.ptr(e_k), _lib.ptr(ws), 0, _lib.stream_ptr())
            _lib.check(rc, "vl_rk3_step")
    torch.cuda.synchronize()
    return f.clone(), E.clone(), phi, energy


@pytest.mark.parametrize("shape", [(6, 6, 8, 8), (5, 3, 4, 4)])
def test_noalloc_step_matches_the_workspace_step(cuda, shape):
    """Parametrised larger grid first, smaller second: pytest runs them in
    this order in one process, so the second call finds the first call's
    contents in the shared array."""
    import torch
    n, nx, nu, nv = shape
    print(f"noalloc parity n={n} elements={nx}x{nu}x{nv}", flush=True)
    ref = _step(_run(cuda, n, nx, nu, nv, 0.01), False, 3)
    got = _step(_run(cuda, n, nx, nu, nv, 0.01), True, 3)
    for a, b in zip(got, ref):
        assert torch.equal(a, b)
    assert bool(torch.isfinite(ref[0]).all()) and float(ref[3][-1]) > 0


def test_noalloc_refuses_a_workspace_above_the_array(cuda):
    from vlasov1d2v_amd import _lib
    from vlasov1d2v_amd.basis import Basis3D
    from vlasov1d2v_amd.grid import Grid3D
    L = 2 * np.pi / 0.1
    grids = Grid3D(Basis3D(np.array([9, 9, 9])), np.array([-L / 2, -8.5, -8.5]), np.array([L / 2, 8.5, 8.5]),
                   np.array([64, 64, 64]))
    g = grids.geometry()
    assert g.workspace_bytes > (32 << 20)
    lib = _lib.load()
    rc = lib.vl_rk3_step_noalloc(g.handle, None, None, None, None, None, 0.01, None, 0, _lib.stream_ptr())
    assert rc == _lib.VL_E_WORKSPACE
    assert b"scratch array" in lib.vl_last_error_string()
