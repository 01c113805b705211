"""Reduced-precision storage forms (FP32 / BF16 / FP16 storage of the
deviation g = (f - f0) * scale from the ring background, FP32 arithmetic,
FP64 moment / Poisson / energy) against the FP64 GPU path, which is pinned to
the oracle by tests/test_gpu_parity.py.

Tolerances are error budgets of each storage format, not parity bars.
"""
import numpy as np
import pytest

from helpers import rel_l2

pytestmark = pytest.mark.gpu

# storage: (rel-L2 bound on total f, rel-L2 bound on the field-energy series)
BOUNDS = {"f32": (1e-8, 1e-5), "bf16": (1e-4, 5e-3), "f16": (2e-5, 2e-3)}


def _run(storage, n=6, nx=8, nu=10, nv=10, steps=20, dt=0.005, seed=3):
    from vlasov1d2v_amd import Run
    run = Run(n, nx, nu, nv, seed=seed, dt=dt, storage=storage)
    st = run.run(steps)
    f = run.dist.total().cpu().numpy()[1:-1, :, 1:-1, :, 1:-1, :]
    return f, st.field_energy


@pytest.mark.parametrize("storage", ["f32", "bf16", "f16"])
def test_reduced_storage_tracks_fp64(cuda, storage):
    f64, e64 = _run("f64")
    f, e = _run(storage)
    ef, ee = rel_l2(f, f64), rel_l2(e, e64)
    print(f"{storage}: rel-L2 f {ef:.3e}, energy {ee:.3e}", flush=True)
    bf, be = BOUNDS[storage]
    assert np.all(np.isfinite(f)) and np.all(np.isfinite(e))
    assert ef < bf
    assert ee < be


def test_reduced_storage_initial_state_is_exact_background_plus_deviation(cuda):
    import torch
    from vlasov1d2v_amd import Run
    a = Run(5, 4, 6, 6, storage="f64")
    b = Run(5, 4, 6, 6, storage="f32")
    fa = a.dist.arr.cpu().numpy()
    fb = b.dist.total().cpu().numpy()
    assert rel_l2(fb, fa) < 1e-9
    assert b.dist.arr.dtype == torch.float32


@pytest.mark.parametrize("storage", ["bf16", "f16"])
def test_row_split_path_is_as_accurate_as_vector_path(cuda, storage):
    """Same check for the row-split matrix-core kernel (n = 9, the default
    resident layout of the 16-bit forms) against the element-major
    vector-ALU kernel, both measured against the FP64 path."""
    from vlasov1d2v_amd import Run, _lib
    from vlasov1d2v_amd.grid import STORAGE_CODES
    f64, e64 = _run("f64", n=9, nx=6, nu=8, nv=9, steps=12, dt=0.003)
    errs = {}
    for mfma in (0, 1):
        run = Run(9, 6, 8, 9, seed=3, dt=0.003, storage=storage)
        run.grids.geometry(STORAGE_CODES[storage]).set_tuning(_lib.VL_TUNE_MFMA, mfma)
        run.grids.geometry(STORAGE_CODES[storage]).set_tuning(_lib.VL_TUNE_ROUNDING, 0)
        st = run.run(12)
        assert run.stepper._res["layout"] == (_lib.VL_LAYOUT_ROW_SPLIT if mfma else _lib.VL_LAYOUT_ELEMENT_MAJOR)
        f = run.dist.total().cpu().numpy()[1:-1, :, 1:-1, :, 1:-1, :]
        errs[mfma] = (rel_l2(f, f64), rel_l2(st.field_energy, e64))
    print(f"{storage}: vector {errs[0]}, row-split {errs[1]}", flush=True)
    assert errs[1][0] <= 1.1 * errs[0][0]
    assert errs[1][1] <= 1.1 * errs[0][1] + 1e-7


@pytest.mark.parametrize("storage", ["bf16", "f16"])
def test_matrix_core_path_is_as_accurate_as_vector_path(cuda, storage):
    """Against the FP64 path, the matrix-core stage kernel (the default for
    the 16-bit forms in the resident layout) is as accurate as the
    vector-ALU kernel: its operator rounding (hi + lo 16-bit parts) is far
    below the storage rounding."""
    from vlasov1d2v_amd import Run, _lib
    from vlasov1d2v_amd.grid import STORAGE_CODES
    f64, e64 = _run("f64")
    errs = {}
    for mfma in (0, 1):
        run = Run(6, 8, 10, 10, seed=3, dt=0.005, storage=storage)
        run.grids.geometry(STORAGE_CODES[storage]).set_tuning(_lib.VL_TUNE_MFMA, mfma)
        st = run.run(20)
        f = run.dist.total().cpu().numpy()[1:-1, :, 1:-1, :, 1:-1, :]
        errs[mfma] = (rel_l2(f, f64), rel_l2(st.field_energy, e64))
    print(f"{storage}: vector {errs[0]}, matrix {errs[1]}", flush=True)
    assert errs[1][0] <= 1.1 * errs[0][0]
    assert errs[1][1] <= 1.1 * errs[0][1] + 1e-7


@pytest.mark.parametrize("storage", ["bf16", "f16"])
def test_stochastic_rounding_keeps_the_field_evolution(cuda, storage):
    """Over a long run at small dt, stage updates are about one storage ulp of
    the stored deviation; round-to-nearest then loses part of every update
    and damps the field, stochastic rounding (the row-split kernel's
    default) keeps it unbiased.  Bars: the energy-series error against the
    FP64 path is at least 3x smaller with stochastic rounding, and the
    stochastic run is reproducible (same seeds, same result)."""
    import torch
    from vlasov1d2v_amd import Run, _lib
    from vlasov1d2v_amd.grid import STORAGE_CODES
    n, nx, nu, nv, steps, dt = 8, 4, 8, 8, 600, 0.004
    ref = Run(n, nx, nu, nv, seed=5, dt=dt).run(steps).field_energy
    errs, states = {}, {}
    for mode in (0, 1, 1):
        run = Run(n, nx, nu, nv, seed=5, dt=dt, storage=storage)
        run.grids.geometry(STORAGE_CODES[storage]).set_tuning(_lib.VL_TUNE_ROUNDING, mode)
        e = run.run(steps, graph_steps=50).field_energy
        assert run.stepper._res["layout"] == _lib.VL_LAYOUT_ROW_SPLIT
        if mode in states:
            assert torch.equal(states[mode], run.dist.arr), "stochastic rounding must be reproducible"
        states[mode] = run.dist.arr.clone()
        errs[mode] = rel_l2(e, ref)
    print(f"{storage}: energy-series rel-L2 nearest {errs[0]:.3e}, stochastic {errs[1]:.3e}", flush=True)
    assert np.all(np.isfinite(errs[1]))
    assert errs[1] * 3 < errs[0]
