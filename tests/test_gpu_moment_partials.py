"""Moment partials written by a stage kernel against those vl_moment_partials
recomputes from the stored state.

A step skips its first moment pass when the workspace still holds the
partials the previous step's last stage wrote (VL_STEP_RHO_VALID); graph
replays and the first step of a loop recompute them.  Where the stage kernel
and the moment kernel sum the same stored values in the same order (the
vector-ALU kernels of every storage, reference and element-major layouts, and
the row-split kernel with k_moment_rs), the two must agree bit for bit, so a
run gives the same numbers whichever path it took.  The matrix-core
element-major kernel and the natural-frame row-split kernel sum FP32 row
partials first; theirs agree to FP32 rounding of those partial sums.

The comparison covers the whole workspace: after the step (A), after
vl_moment_partials of an all-zero state (Z: every slot the kernel writes is
0) and after vl_moment_partials of the stepped state (B).  B == A bytewise
and Z != A show that the kernel rewrote every slot the stage kernel left and
that the two agree.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EM, RS, REF = 1, 2, 0

# (storage, n, nx, nu, nv, layout, mfma, row-split kernel, bit-exact)
CASES = [
    ("f64", 9, 5, 6, 7, REF, 0, 1, True),
    ("f64", 9, 5, 6, 7, EM, 0, 1, True),
    ("f64", 4, 3, 3, 17, EM, 0, 1, True),
    ("f32", 9, 5, 6, 7, EM, 0, 1, True),
    ("f32", 7, 2, 5, 6, REF, 0, 1, True),
    ("bf16", 9, 5, 6, 7, EM, 0, 1, True),
    ("f16", 8, 3, 3, 5, EM, 0, 1, True),
    ("bf16", 9, 5, 6, 7, RS, 1, 1, True),
    ("f16", 9, 5, 6, 7, RS, 1, 1, True),
    ("bf16", 8, 6, 2, 7, RS, 1, 1, True),
    ("f16", 8, 3, 3, 5, RS, 1, 1, True),
    ("bf16", 9, 2, 1, 1, RS, 1, 1, True),
    ("bf16", 9, 5, 6, 7, EM, 1, 1, False),
    ("f16", 9, 5, 6, 7, RS, 1, 2, False),
]


def _ids(c):
    lay = {REF: "ref", EM: "em", RS: "rs"}[c[5]]
    return f"{c[0]}-n{c[1]}-{c[2]}x{c[3]}x{c[4]}-{lay}-mx{c[6]}-k{c[7]}"


@pytest.mark.parametrize("case", CASES, ids=_ids)
def test_stage_partials_equal_recomputed_partials(cuda, case):
    import torch
    from vlasov1d2v_amd import Run, _lib
    from vlasov1d2v_amd.grid import STORAGE_CODES
    storage, n, nx, nu, nv, layout, mfma, rs_kernel, exact = case
    run = Run(n, nx, nu, nv, seed=5, dt=0.004, storage=storage)
    g = run.grids.geometry(STORAGE_CODES[storage])
    if storage != "f64":
        g.set_tuning(_lib.VL_TUNE_MFMA, mfma)
        g.set_tuning(_lib.VL_TUNE_RS_KERNEL, rs_kernel)
    if layout == REF:
        f = run.dist.arr
        flags = 0
    else:
        run.resident_begin(layout=layout)
        f = run.stepper._res["bufs"][0]
        flags = _lib.VL_STEP_ROW_SPLIT if layout == RS else _lib.VL_STEP_ELEMENT_MAJOR
    for _ in range(2):
        run.step()
    torch.cuda.synchronize()
    ws = g.workspace(f.device)
    lib = _lib.load()
    after_step = ws.clone()
    zero = torch.zeros_like(f)
    _lib.check(lib.vl_moment_partials(g.handle, _lib.ptr(zero), _lib.ptr(ws), flags, _lib.stream_ptr()),
               "vl_moment_partials")
    of_zero = ws.clone()
    _lib.check(lib.vl_moment_partials(g.handle, _lib.ptr(f), _lib.ptr(ws), flags, _lib.stream_ptr()),
               "vl_moment_partials")
    again = ws.clone()
    torch.cuda.synchronize()
    a = after_step.cpu().numpy()
    b = again.cpu().numpy()
    z = of_zero.cpu().numpy()
    assert np.any(z != a), "the moment kernel wrote no partial"
    a64, b64 = a.view(np.float64), b.view(np.float64)
    differ = a64.view(np.int64) != b64.view(np.int64)
    scale = float(np.max(np.abs(np.nan_to_num(a64[z.view(np.int64) != a64.view(np.int64)]))))
    worst = float(np.max(np.abs(a64[differ] - b64[differ]))) / scale if differ.any() else 0.0
    print(f"{_ids(case)}: {int(differ.sum())} of {int((z.view(np.int64) != a64.view(np.int64)).sum())} "
          f"partial words differ, max {worst:.2e} of the largest", flush=True)
    if exact:
        assert not differ.any()
    else:
        assert worst < 1e-6
