"""CPU tests of the product's host-side setup code and of the C-ABI library
(symbols only; no device calls)."""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle
from helpers import rel_max

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("n", range(2, 11))
def test_basis_matches_oracle(n):
    from vlasov1d2v_amd.basis import Basis1D
    b = Basis1D(n)
    o = oracle.Basis1D(n)
    assert np.allclose(b.nodes, o.nodes, rtol=0, atol=4e-16)
    assert np.allclose(b.weights, o.weights, rtol=0, atol=4e-16)
    assert np.max(np.abs(b.up - o.up)) < 5e-13 * np.max(np.abs(o.up))
    assert np.max(np.abs(b.xi - o.xi)) < 5e-13 * np.max(np.abs(o.xi))
    assert np.max(np.abs(b.vandermonde_inverse - o.vandermonde_inverse)) < 1e-13
    # both zero the same entries of up (threshold 1e-10, main/basis.py:186)
    assert np.array_equal(b.up == 0, o.up == 0)


@pytest.mark.parametrize("n,nx", [(8, 16), (9, 64), (6, 48)])
def test_spectral_tables_match_oracle(n, nx):
    from vlasov1d2v_amd.basis import Basis1D
    from vlasov1d2v_amd.grid import Grid1D
    L = 2 * np.pi / 0.1
    g = Grid1D(-L / 2, L / 2, nx, Basis1D(n), spectrum=True)
    o = oracle.Grid1D(-L / 2, L / 2, nx, oracle.Basis1D(n), spectrum=True)
    assert g.nyquist_number == o.nyquist_number
    assert np.array_equal(g.wave_numbers, o.wave_numbers)
    assert np.array_equal(g.arr, o.arr)
    assert np.max(np.abs(g.spectral_transform - o.spectral_transform)) < 1e-13
    assert np.max(np.abs(g.grid_phases - o.grid_phases)) < 1e-15


def test_mode_count_float_expression():
    """M = 2*floor(2.5 L // dx) - 1 in float arithmetic (main/grid.py:46): 119, not 120, at nx=48."""
    from vlasov1d2v_amd.basis import Basis1D
    from vlasov1d2v_amd.grid import Grid1D
    for k0 in (0.1, 0.05, 0.2):
        L = 2 * np.pi / k0
        g = Grid1D(-L / 2, L / 2, 48, Basis1D(3), spectrum=True)
        assert g.nyquist_number == 2.5 * g.length // g.dx
    L = 2 * np.pi / 0.1
    assert Grid1D(-L / 2, L / 2, 48, Basis1D(3), spectrum=True).nyquist_number == 119.0
    assert Grid1D(-L / 2, L / 2, 64, Basis1D(3), spectrum=True).nyquist_number == 160.0


def test_ring_ic_matches_oracle():
    from vlasov1d2v_amd.basis import Basis3D
    from vlasov1d2v_amd.grid import Distribution, Grid3D
    L = 2 * np.pi / 0.1
    lows, highs = np.array([-L / 2, -8.5, -8.5]), np.array([L / 2, 8.5, 8.5])
    g = Grid3D(Basis3D([5, 5, 5]), lows, highs, [3, 4, 5])
    d = Distribution(1.0, 6, [3, 4, 5], [5, 5, 5])
    ring = d.ring(g)
    case = oracle.Case(n=5, nx=3, nu=4, nv=5)
    ref = oracle.ring_distribution(case.grids, 1.0, 6)
    assert rel_max(ring[None, None], ref) < 1e-15


def _header_symbols():
    text = open(os.path.join(ROOT, "include", "vlasov1d2v.h")).read()
    return sorted(set(re.findall(r"\b(vl_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_header_symbol():
    from vlasov1d2v_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    syms = _header_symbols()
    assert set(syms) == set(_lib.EXPORTS)
    for s in syms:
        assert hasattr(lib, s), s


def test_library_loads_and_rejects_bad_geometry():
    """Validation paths return negative codes without touching a device."""
    from vlasov1d2v_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    lib = _lib.load()
    d = _lib.VlGeomDesc()
    d.n = 11
    h = ctypes.c_void_p()
    assert lib.vl_geom_create(ctypes.byref(d), ctypes.byref(h)) == -1
    assert b"n=11" in lib.vl_last_error_string()
    d.n, d.nx, d.nu, d.nv = 4, 0, 1, 1
    assert lib.vl_geom_create(ctypes.byref(d), ctypes.byref(h)) == -1
    assert lib.vl_rhs(None, None, None, None, None) == -4
    assert lib.vl_geom_create(None, ctypes.byref(h)) == -2


def test_ulp_distance_helper():
    """The ulp metric of the layout tests: neighbours are 1 apart, across
    zero too, and +0 / -0 coincide."""
    import torch
    from helpers import ulp_distance
    for dt, it in ((torch.float16, torch.int16), (torch.bfloat16, torch.int16), (torch.float32, torch.int32)):
        a = torch.tensor([1.0, 3.5, -2.0, 0.0], dtype=dt)
        up = (a.view(it) + 1).view(dt)             # away from zero for both signs
        assert list(ulp_distance(a, up)) == [1, 1, 1, 1]
        tiny = torch.tensor([0.0], dtype=dt).view(it) + 1
        pos, neg = tiny.view(dt), (-tiny.view(dt))
        assert ulp_distance(pos, neg)[0] == 2
        assert ulp_distance(torch.tensor([0.0], dtype=dt), torch.tensor([-0.0], dtype=dt))[0] == 0


def test_ensemble_members_follow_the_workload_ranges():
    from vlasov1d2v_amd import draw_members
    a = draw_members(48, seed=0)
    assert a == draw_members(48, seed=0) and a != draw_members(48, seed=1)
    assert len(a) == 48 and len({m.seed for m in a}) == 48
    for m in a:
        assert m.nx in (16, 32, 48, 64) and 24 <= m.nu <= 48 and m.nu == m.nv
        assert 6 <= m.n <= 9 and 1 <= m.ring_j <= 8
        assert 0.05 <= m.k0 <= 0.2 and 0.05 <= m.b_field <= 0.2
    # every node count and x resolution occurs in the 48 draws
    assert {m.n for m in a} == {6, 7, 8, 9} and {m.nx for m in a} == {16, 32, 48, 64}


def test_ensemble_entry_points_validate_without_a_device():
    from vlasov1d2v_amd import _lib, Ensemble
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    lib = _lib.load()
    h = ctypes.c_void_p()
    assert lib.vl_ensemble_create(None, 1, ctypes.byref(h)) == -2
    handles = (ctypes.c_void_p * 1)(None)
    assert lib.vl_ensemble_create(handles, 0, ctypes.byref(h)) == -1
    assert lib.vl_ensemble_create(handles, 1, ctypes.byref(h)) == -1      # no geometry: no storage
    assert lib.vl_ensemble_bind(None, 0, None, None, None, None, None, 0.01) == -4
    assert lib.vl_ensemble_rk3_step(None, None, None) == -4
    assert lib.vl_ensemble_set_schedule(None, 1, 0, 0) == -4
    assert lib.vl_ensemble_destroy(None) == 0
    with pytest.raises(ValueError):
        Ensemble([], storage="bf16")


def _geom_desc(n, up):
    """A small valid geometry description around the given volume operator."""
    from vlasov1d2v_amd import _lib
    keep = []

    def arr(a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        keep.append(a)
        return a.ctypes.data
    d = _lib.VlGeomDesc()
    d.n, d.nx, d.nu, d.nv, d.n_modes, d.storage = n, 2, 2, 2, 3, 0
    d.jx = d.ju = d.jv = 1.0
    d.acc, d.b_field, d.poisson_coef, d.x_dx, d.x_length = -1.0, 0.1, -1.0, 1.0, 2.0
    d.up = arr(up)
    d.xi = arr(np.zeros((n, 2)))
    d.weights = arr(np.ones(n))
    d.x_nodes = arr(np.zeros(4 * n))
    d.u_nodes = arr(np.zeros(4 * n))
    d.v_nodes = arr(np.zeros(4 * n))
    d.wave_numbers = arr(np.ones(3))
    d.transform = arr(np.zeros(2 * 2 * n * 3))
    d.phases = arr(np.zeros(2 * 2 * n * 3))
    return d, keep


@pytest.mark.parametrize("n", list(range(2, 11)))
def test_geometry_accepts_the_reference_volume_operator(n):
    """The reference's own up table (scipy polynomials on its literal nodes)
    is centro-antisymmetric only to rounding; the library accepts it (the
    validation passes; without a device the call then fails in HIP with a
    positive code) and rejects a table off by more than 1e-10 of max|up|."""
    from helpers import reference_way_up
    from vlasov1d2v_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    lib = _lib.load()
    up = reference_way_up(n)
    defect = np.max(np.abs(up + up[::-1, ::-1])) / np.max(np.abs(up))
    assert defect <= 1e-11
    h = ctypes.c_void_p()
    d, keep = _geom_desc(n, up)
    rc = lib.vl_geom_create(ctypes.byref(d), ctypes.byref(h))
    assert rc >= 0, lib.vl_last_error_string()
    if rc == 0:
        lib.vl_geom_destroy(h)
    bad = up.copy()
    bad[0, 1] += 1e-6 * np.max(np.abs(up))
    d, keep = _geom_desc(n, bad)
    assert lib.vl_geom_create(ctypes.byref(d), ctypes.byref(h)) == -1
    assert b"centro-antisymmetric" in lib.vl_last_error_string()


def test_reference_volume_operator_is_not_exactly_antisymmetric():
    """The defect that the tolerance above is for: nonzero at n = 9 and 10."""
    from helpers import reference_way_up
    for n in (9, 10):
        up = reference_way_up(n)
        assert np.max(np.abs(up + up[::-1, ::-1])) > 0


def test_status_word_raises_on_a_non_finite_field():
    """Host side of the failure detection: a workspace status word with the
    non-finite bit set (as a stage-3 field solve writes it) raises, a clear
    word does not -- checked on a stand-in geometry whose status is injected."""
    from vlasov1d2v_amd import _lib
    _lib.raise_for_status(0, "step")
    with pytest.raises(_lib.VlasovError, match="not finite"):
        _lib.raise_for_status(_lib.VL_STATUS_NONFINITE, "step")

    class _Stub:
        def __init__(self, word):
            self.word = word

        def read_status(self, device, clear=False):
            return self.word

    _lib.Geometry.check_finite(_Stub(0), "cpu", "step")
    with pytest.raises(_lib.VlasovError, match="graph replay"):
        _lib.Geometry.check_finite(_Stub(1), "cpu", "graph replay of 10 SSP-RK3 steps")
