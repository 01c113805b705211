"""The natural-frame row-split stage kernel (k_stage_rs2, n = 9, the
VL_TUNE_RS_KERNEL = 2 path) against the 9 x 9-frame kernel (k_stage_rs) on
the same inputs, and both against the FP64 path.

With round-to-nearest stores both kernels compute the same FP32 stage values
up to the order of the FP32 sums, so the stored 16-bit values agree to one
or two ulps (entries above 1e-3 of the maximum; smaller ones, which cancelled
down from larger intermediates, are covered by the norm bound) and the field
energies to 1e-5.  Against FP64 the new kernel is as accurate as the old.
"""
import numpy as np
import pytest

from helpers import rel_l2, significant_ulp_distance

pytestmark = pytest.mark.gpu


def _run(storage, kernel, steps, rounding, n=9, nx=5, nu=6, nv=7, dt=0.003, seed=3):
    from vlasov1d2v_amd import Run, _lib
    from vlasov1d2v_amd.grid import STORAGE_CODES
    run = Run(n, nx, nu, nv, seed=seed, dt=dt, storage=storage)
    if storage != "f64":
        g = run.grids.geometry(STORAGE_CODES[storage])
        g.set_tuning(_lib.VL_TUNE_RS_KERNEL, kernel)
        g.set_tuning(_lib.VL_TUNE_ROUNDING, rounding)
    st = run.run(steps)
    if storage != "f64":
        assert run.stepper._res["layout"] == _lib.VL_LAYOUT_ROW_SPLIT
    return run, np.asarray(st.field_energy)


@pytest.mark.parametrize("storage", ["bf16", "f16"])
@pytest.mark.parametrize("steps", [1, 4])
def test_natural_frame_kernel_matches_the_frame_kernel(cuda, storage, steps):
    import torch
    old, e_old = _run(storage, 1, steps, 0)
    new, e_new = _run(storage, 2, steps, 0)
    a, b = new.dist.arr, old.dist.arr
    ulp = significant_ulp_distance(a, b)
    print(f"{storage} {steps} steps: ulp max {ulp.max()}, >1 ulp {np.mean(ulp > 1):.2e}, "
          f"energy {rel_l2(e_new, e_old):.2e}", flush=True)
    fa = new.dist.total().cpu().numpy()
    fb = old.dist.total().cpu().numpy()
    if steps == 1:
        # one step: the same stored values to within one ulp
        assert ulp.max() <= 1
        assert rel_l2(e_new, e_old) < 1e-6
    else:
        # later steps: one-ulp flips of the two kernels' roundings propagate
        assert np.mean(ulp > 1) < 1e-3
        assert rel_l2(e_new, e_old) < 1e-4
    assert rel_l2(fa, fb) < 1e-5
    assert bool(torch.isfinite(a.float()).all())


@pytest.mark.parametrize("storage", ["bf16", "f16"])
@pytest.mark.parametrize("rounding", [0, 1])
def test_natural_frame_kernel_is_as_accurate_as_the_frame_kernel(cuda, storage, rounding):
    ref, e64 = _run("f64", 0, 12, 0)
    f64 = ref.dist.arr.cpu().numpy()[1:-1, :, 1:-1, :, 1:-1, :]
    errs = {}
    for kernel in (1, 2):
        run, e = _run(storage, kernel, 12, rounding)
        f = run.dist.total().cpu().numpy()[1:-1, :, 1:-1, :, 1:-1, :]
        errs[kernel] = (rel_l2(f, f64), rel_l2(e, e64))
    print(f"{storage} rounding {rounding}: frame {errs[1]}, natural frame {errs[2]}", flush=True)
    # f: the same error to a few parts in 1e3; the energy series (12 steps of a
    # small grid) is dominated by the rounding noise of each run: storage budgets
    assert errs[2][0] <= 1.1 * errs[1][0] + 1e-8
    assert errs[2][1] <= {"bf16": 1e-3, "f16": 2e-4}[storage]


def test_natural_frame_kernel_stochastic_rounding_is_reproducible(cuda):
    import torch
    a, ea = _run("bf16", 2, 6, 1)
    b, eb = _run("bf16", 2, 6, 1)
    assert torch.equal(a.dist.arr, b.dist.arr)
    assert np.array_equal(ea, eb)
