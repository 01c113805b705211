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
