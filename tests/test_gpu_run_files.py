"""Run-file records from the GPU state (vl_export_f32 / vl_import_f32) and
the asynchronous writer: records equal the float32 rounding of f for every
storage form, resumed states equal the record, and records written while
the run keeps stepping are the states at their save calls."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _need_io():
    from vlasov1d2v_amd import data_management as dm
    if not os.path.exists(dm.IO_LIB_PATH):
        pytest.skip("libvlasov1d2v_io.so not built (no HDF5 C library)")
    try:
        dm.load_io()
    except OSError as e:      # the HDF5 C library it links is not installed here
        pytest.skip(f"libvlasov1d2v_io.so does not load: {e}")
    return dm


@pytest.mark.parametrize("storage", ["f64", "f32", "bf16", "f16"])
def test_export_and_import_records(cuda, storage):
    import torch
    from vlasov1d2v_amd import Run, _lib
    from vlasov1d2v_amd.data_management import _state_geometry
    n = 8 if storage in ("bf16", "f16") else 5
    run = Run(n, 3, 4, 4, dt=0.01, storage=storage)
    run.run(2)
    d = run.dist
    g = _state_geometry(d)
    out = torch.empty(d.arr.shape, dtype=torch.float32, device=cuda)
    _lib.check(_lib.load().vl_export_f32(g.handle, _lib.ptr(d.arr), _lib.ptr(out), _lib.stream_ptr()),
               "vl_export_f32")
    ref = d.total().to(torch.float32)     # f in FP64 (f0 + g / scale), rounded once
    print(f"{storage}: export max |diff| {(out - ref).abs().max().item():.3e}", flush=True)
    assert torch.equal(out, ref)
    # import: FP64 takes the float32 values; reduced forms store (in - f0) * scale rounded to nearest
    back = torch.empty_like(d.arr)
    _lib.check(_lib.load().vl_import_f32(g.handle, _lib.ptr(out), _lib.ptr(back), _lib.stream_ptr()),
               "vl_import_f32")
    if storage == "f64":
        exp = out.to(torch.float64)
    else:
        f0 = torch.from_numpy(d.f0).to(cuda)
        # (in - f0) * scale in FP64, rounded to FP32, then to the storage type
        exp = ((out.to(torch.float64) - f0[None, None]) * d.scale).to(torch.float32).to(back.dtype)
    assert torch.equal(back, exp)


def test_async_records_and_resume(cuda, tmp_path):
    import torch
    dm = _need_io()
    from vlasov1d2v_amd import Run
    run = Run(5, 3, 4, 4, dt=0.01)
    rd = dm.RunData(str(tmp_path) + os.sep, "run", run.dist.arr.shape, None, None)
    rho = run.dist.moment_zero()
    run.elliptic.poisson2(rho, run.grids.x, run.basis)
    rd.create_file(run.dist, run.elliptic, rho)
    snaps, times, energies = [run.dist.arr.to(torch.float32).cpu()], [0.0], [0.0]
    for k in range(3):
        st = run.run(2)
        rho = run.dist.moment_zero()
        rd.save_data(run.dist, run.elliptic, rho, st.time, st.field_energy[-1])
        snaps.append(run.dist.arr.to(torch.float32).cpu())
        times.append(st.time)
        energies.append(float(st.field_energy[-1]))
    rd.close()
    t, pdf, pot, den, en = dm.ReadData(str(tmp_path) + os.sep, "run").read_data()
    assert pdf.shape[0] == 4
    for k in range(4):
        assert np.array_equal(pdf[k], snaps[k].numpy()), k
    np.testing.assert_array_equal(t, times)
    np.testing.assert_array_equal(en, energies)
    np.testing.assert_array_equal(pot[-1], run.elliptic.potential.to(torch.float32).cpu().numpy())
    # resume from record 2 into a fresh run: the state is that record
    run2 = Run(5, 3, 4, 4, dt=0.01)
    t2 = dm.resume(run2.dist, str(tmp_path / "run.hdf5"), idx=2)
    assert t2 == times[2]
    assert torch.equal(run2.dist.arr.cpu(), snaps[2].to(torch.float64))
