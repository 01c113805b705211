"""Element-wise entry points over whole states (vl_axpby, vl_init_separable,
vl_deviation_to_storage, vl_total_f64, vl_density_deviation) against NumPy
and torch's own dtype conversions, through the C ABI."""
import numpy as np
import pytest

from vlasov1d2v_amd import _lib

pytestmark = pytest.mark.gpu

SHAPE = (3, 4, 2, 5)          # n, nx, nu, nv


def _geom(storage=_lib.VL_F64):
    from vlasov1d2v_amd import Run
    n, nx, nu, nv = SHAPE
    run = Run(n, nx, nu, nv, k0=0.1, b_field=0.1, ring_j=6, seed=0, dt=0.01)
    return run, run.grids.geometry(storage)


def _interior_mask(shape):
    m = np.zeros(shape, dtype=bool)
    m[1:-1, :, 1:-1, :, 1:-1, :] = True
    return m


def test_axpby_interior_and_full(cuda):
    import torch
    run, g = _geom()
    shape = run.grids.shape
    rng = np.random.default_rng(3)
    y0, x0 = rng.standard_normal(shape), rng.standard_normal(shape)
    y = torch.from_numpy(y0).to(cuda)
    x = torch.from_numpy(x0).to(cuda)
    _lib.axpby(g, y, x, 0.75, -1.25, interior=True)
    want = np.where(_interior_mask(shape), 0.75 * y0 - 1.25 * x0, y0)
    # one fused multiply-add per node: within two ulps of the three-rounding form
    assert np.abs(y.cpu().numpy() - want).max() <= 4.5e-16 * np.abs(want).max()
    assert np.array_equal(y.cpu().numpy()[~_interior_mask(shape)], y0[~_interior_mask(shape)])
    _lib.axpby(g, y, x, 0.0, 1.0, interior=False)          # a copy, ghosts included
    assert np.array_equal(y.cpu().numpy(), x0)


def test_axpby_skips_unread_operands(cuda):
    import torch
    run, g = _geom()
    shape = run.grids.shape
    y = torch.full(shape, float("nan"), dtype=torch.float64, device=cuda)
    x = torch.ones(shape, dtype=torch.float64, device=cuda)
    _lib.axpby(g, y, x, 0.0, 2.0, interior=False)            # a == 0: y not read
    assert torch.all(y == 2.0)
    _lib.axpby(g, y, None, 3.0, 0.0, interior=True)          # b == 0: x not read (NULL)
    yn = y.cpu().numpy()
    assert np.all(yn[_interior_mask(shape)] == 6.0) and np.all(yn[~_interior_mask(shape)] == 2.0)
    with pytest.raises(_lib.VlasovError):
        _lib.axpby(g, y, None, 1.0, 1.0)


@pytest.mark.parametrize("storage", ["f64", "f32", "bf16", "f16"])
def test_init_separable_and_total(cuda, storage):
    import torch
    from vlasov1d2v_amd.grid import STORAGE_CODES, STORAGE_SCALE, storage_dtype
    run, g = _geom(STORAGE_CODES[storage])
    n, nx, nu, nv = SHAPE
    rng = np.random.default_rng(5)
    vel = rng.random(((nu + 2) * n, (nv + 2) * n)).reshape(nu + 2, n, nv + 2, n)
    xp = 1e-3 * rng.standard_normal((nx + 2, n))
    out = torch.empty(run.grids.shape, dtype=storage_dtype(storage), device=cuda)
    run.dist._separable(g, vel, xp, out)
    prod = (torch.from_numpy(xp)[:, :, None, None, None, None] * torch.from_numpy(vel)[None, None]).to(cuda)
    want = (prod * STORAGE_SCALE[storage]).to(storage_dtype(storage))
    assert torch.equal(out.view(torch.int16 if out.element_size() == 2 else out.dtype),
                       want.view(torch.int16 if want.element_size() == 2 else want.dtype))
    if storage == "f64":
        return
    _lib.check(_lib.load().vl_set_background(g.handle, _lib.ptr(np.ascontiguousarray(vel))), "vl_set_background")
    tot = torch.empty(run.grids.shape, dtype=torch.float64, device=cuda)
    _lib.check(_lib.load().vl_total_f64(g.handle, _lib.ptr(out), _lib.ptr(tot), _lib.stream_ptr()), "vl_total_f64")
    ref = vel[None, None] + out.to(torch.float64).cpu().numpy() / STORAGE_SCALE[storage]
    assert np.array_equal(tot.cpu().numpy(), ref)
    # deviation -> storage of the same FP64 deviation gives the same words
    dev = prod.contiguous()
    g2 = torch.empty_like(out)
    _lib.check(_lib.load().vl_deviation_to_storage(g.handle, _lib.ptr(dev), _lib.ptr(g2), _lib.stream_ptr()),
               "vl_deviation_to_storage")
    assert torch.equal(g2.view(torch.int16 if g2.element_size() == 2 else g2.dtype),
                       want.view(torch.int16 if want.element_size() == 2 else want.dtype))


@pytest.mark.parametrize("sign", [1, -1])
def test_density_deviation(cuda, sign):
    import torch
    run, g = _geom()
    rng = np.random.default_rng(11)
    rho = 1.0 + 0.01 * rng.standard_normal((g.nx, g.n))
    out = _lib.density_deviation(g, torch.from_numpy(rho).to(cuda), sign=sign).cpu().numpy()
    want = rho.mean() - rho if sign > 0 else rho - rho.mean()
    assert np.abs(out - want).max() < 4e-16
    assert abs(out.sum()) < 1e-13


def test_native_graph_replays(cuda):
    """vl_graph_*: calls recorded on a side stream replay on the current
    stream; every replay is checked, with eager library work between them."""
    import torch
    run, g = _geom()
    shape = run.grids.shape
    y = torch.ones(shape, dtype=torch.float64, device=cuda)
    x = torch.full(shape, 0.5, dtype=torch.float64, device=cuda)
    side = torch.cuda.Stream(device=cuda)
    side.wait_stream(torch.cuda.current_stream(cuda))
    graph = _lib.HipGraph()
    with graph.capture(side):
        _lib.axpby(g, y, x, 1.0, 1.0, interior=False)      # y += x, twice
        _lib.axpby(g, y, x, 1.0, 1.0, interior=False)
    torch.cuda.current_stream(cuda).wait_stream(side)
    torch.cuda.synchronize()
    assert torch.all(y == 1.0)                              # captured, not run
    want = 1.0
    for rep in range(3):
        graph.launch()
        want += 1.0
        torch.cuda.synchronize()
        assert torch.all(y == want), rep
        _lib.axpby(g, y, None, 2.0, 0.0, interior=False)    # eager work between replays
        want *= 2.0
    torch.cuda.synchronize()
    assert torch.all(y == want)


def test_native_graph_capture_error_ends_capture(cuda):
    """A failing call inside the capture block leaves the stream usable."""
    import torch
    run, g = _geom()
    y = torch.zeros(run.grids.shape, dtype=torch.float64, device=cuda)
    side = torch.cuda.Stream(device=cuda)
    graph = _lib.HipGraph()
    with pytest.raises(_lib.VlasovError):
        with graph.capture(side):
            _lib.axpby(g, y, None, 1.0, 1.0)                 # x NULL with b != 0: refused
    assert graph.handle is None
    with torch.cuda.stream(side):
        _lib.axpby(g, y, None, 0.0, 0.0, interior=False)   # the side stream runs eagerly again
    torch.cuda.synchronize()
    assert torch.all(y == 0.0)


@pytest.mark.parametrize("nt", [1, 0])
@pytest.mark.parametrize("nvals,grid", [(2, 1), (1000, 3), (123456, 64)])
def test_stream_probe(cuda, nvals, grid, nt):
    import torch
    a = torch.arange(nvals, dtype=torch.float64, device=cuda)
    b = torch.full((nvals,), 0.25, dtype=torch.float64, device=cuda)
    out = torch.empty_like(a)
    _lib.check(_lib.load().vl_stream_probe(_lib.ptr(a), _lib.ptr(b), _lib.ptr(out), nvals * 8, grid, nt,
                                           _lib.stream_ptr()), "vl_stream_probe")
    assert torch.equal(out, a + 0.25)
    with pytest.raises(_lib.VlasovError):
        _lib.check(_lib.load().vl_stream_probe(_lib.ptr(a), _lib.ptr(b), _lib.ptr(out), 8, grid, nt,
                                               _lib.stream_ptr()), "vl_stream_probe")
