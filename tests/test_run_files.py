"""Run files (RunData / ReadData, main/data_management.py:5-126) on the host:
the HDF5 library writes and reads them, and h5py -- the reference's own IO
library, run in a separate interpreter that has it -- checks both directions:
files written here read back through h5py with the reference's dataset names,
dtypes, shapes and maxshapes, and files written with the reference's h5py
calls read back here."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

from vlasov1d2v_amd import data_management as dm

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROBE = os.path.join(ROOT, "tests", "h5py_probe.py")
SHAPE = (5, 3, 6, 3, 4, 3)        # (NX+2, n, NU+2, n, NV+2, n)


def _need_io():
    if not os.path.exists(dm.IO_LIB_PATH):
        pytest.skip("libvlasov1d2v_io.so not built (no HDF5 C library)")


def _h5py_python():
    """An interpreter with h5py, run without this environment's paths."""
    for cand in (os.environ.get("VL_H5PY_PYTHON"), "/opt/conda/bin/python3.9"):
        if cand and os.path.exists(cand):
            return cand
    return None


def _probe(*args):
    py = _h5py_python()
    if py is None:
        pytest.skip("no interpreter with h5py")
    env = {"PATH": os.path.dirname(py) + ":/usr/bin:/bin", "HOME": "/tmp"}
    r = subprocess.run([py, "-E", "-s", PROBE, *args], capture_output=True, text=True, env=env, timeout=300)
    if r.returncode != 0 and "No module named 'h5py'" in r.stderr:
        pytest.skip("h5py not importable")
    assert r.returncode == 0, r.stderr[-2000:]
    return r.stdout


class _Field:
    def __init__(self, potential):
        self.potential = potential


def _records(count, seed=3):
    rng = np.random.default_rng(seed)
    pdf = [rng.standard_normal(SHAPE) * 10.0 ** rng.integers(-8, 2, SHAPE) for _ in range(count)]
    pot = [rng.standard_normal((SHAPE[0] - 2, SHAPE[1])) for _ in range(count)]
    den = [1.0 + 1e-3 * rng.standard_normal((SHAPE[0] - 2, SHAPE[1])) for _ in range(count)]
    return pdf, pot, den


def _write(tmp_path, count=3):
    pdf, pot, den = _records(count)
    rd = dm.RunData(str(tmp_path) + os.sep, "run", SHAPE, None, None)
    rd.create_file(pdf[0], _Field(pot[0]), den[0])
    for k in range(1, count):
        rd.save_data(pdf[k], _Field(pot[k]), den[k], 0.1 * k, 1e-3 * k)
    rd.close()
    return pdf, pot, den


def test_write_then_read_back(tmp_path):
    _need_io()
    pdf, pot, den = _write(tmp_path, 4)
    rd = dm.ReadData(str(tmp_path) + os.sep, "run")
    t, p, po, de, en = rd.read_data()
    assert p.dtype == np.float32 and p.shape == (4,) + SHAPE
    np.testing.assert_array_equal(p, np.array(pdf, dtype=np.float32))
    np.testing.assert_array_equal(po, np.array(pot, dtype=np.float32))
    np.testing.assert_array_equal(de, np.array(den, dtype=np.float32))
    # time and energy keep their values (float64); record 0 is (0, 0) as in create_file
    np.testing.assert_array_equal(t, [0.0, 0.1, 0.2, 0.30000000000000004])
    np.testing.assert_array_equal(en, [0.0, 1e-3, 2e-3, 3e-3])
    t2, p2, po2 = rd.read_specific_time(-1)
    assert t2 == t[-1]
    np.testing.assert_array_equal(p2, p[-1])
    np.testing.assert_array_equal(po2, po[-1])
    t3, po3 = rd.read_potential_only([1, 3])
    np.testing.assert_array_equal(t3, t[1:3])
    np.testing.assert_array_equal(po3, po[1:3])
    # the dataset layout of RunData.create_file (main/data_management.py:16-35)
    assert rd.dims("pdf") == ((4,) + SHAPE, (-1,) + SHAPE)
    assert rd.dims("potential") == ((4, SHAPE[0] - 2, SHAPE[1]), (-1, SHAPE[0], SHAPE[1]))
    assert rd.dims("density") == ((4, SHAPE[0] - 2, SHAPE[1]), (-1, SHAPE[0], SHAPE[1]))
    assert rd.dims("time") == ((4,), (-1,)) and rd.dims("energy") == ((4,), (-1,))
    with pytest.raises(IndexError):
        rd.read_specific_time(4)


def test_h5py_reads_files_written_here(tmp_path):
    _need_io()
    pdf, pot, den = _write(tmp_path, 3)
    desc = json.loads(_probe("describe", str(tmp_path / "run.hdf5")))
    assert sorted(desc) == ["density", "energy", "pdf", "potential", "time"]
    assert desc["pdf"]["dtype"] == "float32" and desc["pdf"]["shape"] == [3] + list(SHAPE)
    assert desc["pdf"]["maxshape"] == [-1] + list(SHAPE)
    assert desc["potential"]["dtype"] == "float32" and desc["potential"]["maxshape"] == [-1, SHAPE[0], SHAPE[1]]
    assert desc["density"]["shape"] == [3, SHAPE[0] - 2, SHAPE[1]]
    assert desc["time"]["dtype"] == "float64" and desc["time"]["maxshape"] == [-1]
    ref = np.array(pdf, dtype=np.float32).astype(np.float64)
    assert desc["pdf"]["sum"] == pytest.approx(ref.sum(), rel=1e-12, abs=1e-300)
    assert desc["pdf"]["head"] == ref.ravel()[:4].tolist()
    assert desc["potential"]["head"] == np.array(pot, dtype=np.float32).astype(np.float64).ravel()[:4].tolist()
    assert desc["energy"]["head"] == [0.0, 1e-3, 2e-3]


def test_files_written_by_h5py_read_back_here(tmp_path):
    """A file made with RunData's own h5py calls: float32 records read back
    exactly; its int64 "time" / "energy" (from data=[0]) read as float64
    holding the truncated values that h5py stored."""
    _need_io()
    path, npz = tmp_path / "ref.hdf5", tmp_path / "ref.npz"
    _probe("write", str(path), str(npz), "3", *[str(s) for s in SHAPE])
    exp = np.load(npz)
    rd = dm.ReadData(str(tmp_path) + os.sep, "ref")
    t, p, po, de, en = rd.read_data()
    np.testing.assert_array_equal(p, exp["pdf"])
    np.testing.assert_array_equal(po, exp["potential"])
    np.testing.assert_array_equal(de, exp["density"])
    np.testing.assert_array_equal(t, np.trunc(exp["time"]))
    np.testing.assert_array_equal(en, np.trunc(exp["energy"]))
    assert en.dtype == np.float64
    # appending to it with this library works too
    pdf, pot, den = _records(1, seed=9)
    wr = dm.RunData(str(tmp_path) + os.sep, "ref", SHAPE, None, None)
    wr.save_data(pdf[0], _Field(pot[0]), den[0], 7.5, 0.125)
    wr.close()
    t, p, po, de, en = rd.read_data()
    assert p.shape[0] == 4
    np.testing.assert_array_equal(p[-1], pdf[0].astype(np.float32))
    assert t[-1] == 7.0 and en[-1] == 0.0   # int64 datasets truncate, as in the reference


def test_errors_are_reported(tmp_path):
    _need_io()
    rd = dm.ReadData(str(tmp_path) + os.sep, "missing")
    with pytest.raises(dm._lib.VlasovError, match="cannot open"):
        rd.read_data()
    _write(tmp_path, 1)
    with pytest.raises(dm._lib.VlasovError, match="no dataset"):
        dm.ReadData(str(tmp_path) + os.sep, "run").dims("phi")
    lib = dm.load_io()
    import ctypes
    bad = (ctypes.c_int64 * 6)(5, 3, 0, 3, 4, 3)
    buf = np.zeros(4, dtype=np.float32)
    assert lib.vl_run_data_create(str(tmp_path / "x.hdf5").encode(), bad, buf.ctypes.data, 1, 1,
                                  buf.ctypes.data, buf.ctypes.data) == -1
    assert lib.vl_run_data_create(None, bad, None, 1, 1, None, None) == -2
    with pytest.raises(ValueError):
        dm.RunData(str(tmp_path) + os.sep, "run", SHAPE[:5], None, None)
    w = dm.RunData(str(tmp_path) + os.sep, "run", SHAPE, None, None)
    with pytest.raises(ValueError, match="state shape"):
        w.save_data(np.zeros((2, 2)), _Field(np.zeros((3, 3))), np.zeros((3, 3)), 0.0, 0.0)
    # potential / density records must have the file's (NX, n) shape
    with pytest.raises(ValueError, match="records are"):
        w.save_data(np.zeros(SHAPE), _Field(np.zeros((3, 2))), np.zeros((3, 2)), 0.0, 0.0)
    with pytest.raises(ValueError, match="must be"):
        w.save_data(np.zeros(SHAPE), _Field(np.zeros((3, 3))), np.zeros((3, 2)), 0.0, 0.0)
    w.close()
    # the C side checks the buffer sizes itself and leaves the file unchanged
    path = str(tmp_path / "run.hdf5").encode()
    before = dm.ReadData(str(tmp_path) + os.sep, "run")._records("pdf")
    pdf = np.zeros(SHAPE, dtype=np.float32)
    small = np.zeros(4, dtype=np.float32)
    assert lib.vl_run_data_append(path, pdf.ctypes.data, pdf.size, small.ctypes.data, small.size,
                                  small.ctypes.data, small.size, 0.0, 0.0) == -1
    assert b"buffer 4" in lib.vl_io_last_error_string()
    assert dm.ReadData(str(tmp_path) + os.sep, "run")._records("pdf") == before


def test_read_info_parses_the_reference_lines(tmp_path):
    """ReadData.read_info reads fixed lines of the run-parameter file that the
    reference's Reference.info_file writes (main/reference.py:66-141)."""
    lines = ["#"] * 50
    lines[6] = "Ref density: 1.00E+20 [m^-3]"
    lines[7] = "Ref temperature: 2.50E+01 [eV]"
    lines[8] = "Ref length (Debye): 3.20E-05 [m]"
    lines[34] = "X elements: 64"
    lines[35] = "U elements: 50"
    lines[36] = "V elements: 50"
    lines[37] = "X domain minimum: -3.55E+01"
    lines[38] = "X domain maximum: 3.55E+01"
    lines[39] = "Electron u-velocity minimum -8.50E+00"
    lines[40] = "Electron u-velocity maximum 8.50E+00"
    lines[41] = "Electron v-velocity minimum -8.50E+00"
    lines[42] = "Electron v-velocity maximum 8.50E+00"
    lines[43] = "Time to run: 20.00"
    lines[44] = "Time per data write-out 0.10"
    lines[46] = "Spatial element order: 9"
    lines[47] = "Velocity-u element order: 9"
    lines[48] = "Velocity-v element order: 9"
    lines[49] = "Temporal stepping order: 3"
    (tmp_path / "run_info.txt").write_text("\n".join(lines))
    orders, res, lows, highs, time_info, refs = dm.ReadData(str(tmp_path) + os.sep, "run").read_info()
    assert orders.tolist() == [9, 9, 9] and res.tolist() == [64, 50, 50]
    assert lows.tolist() == [-35.5, -8.5, -8] and highs.tolist() == [35.5, 8.5, 8]   # v bounds via int()
    assert time_info.tolist() == [20.0, 0.1, 3.0]
    assert refs.tolist() == [1e20, 25.0, 3.2e-05]


def test_io_library_exports_every_header_symbol():
    _need_io()
    import re
    text = open(os.path.join(ROOT, "include", "vlasov1d2v_io.h")).read()
    syms = set(re.findall(r"\b(vl_[a-z0-9_]+)\s*\(", text))
    assert syms == set(dm.IO_EXPORTS)
    import ctypes
    lib = ctypes.CDLL(dm.IO_LIB_PATH)
    for s in syms:
        assert hasattr(lib, s), s
    assert shutil.which("gcc") is not None


def test_async_write_keeps_the_values_at_the_call(tmp_path):
    """save_data returns before the file is written; changing the caller's
    arrays afterwards must not change the record."""
    _need_io()
    pdf, pot, den = _records(2, seed=11)
    rd = dm.RunData(str(tmp_path) + os.sep, "run", SHAPE, None, None)
    rd.create_file(pdf[0], _Field(pot[0]), den[0])
    a = pdf[1].astype(np.float32)
    p = pot[1].copy()
    rd.save_data(a, _Field(p), den[1], 1.0, 2.0)
    a[...] = -1.0
    p[...] = -1.0
    rd.close()
    t, got, gp, _, _ = dm.ReadData(str(tmp_path) + os.sep, "run").read_data()
    np.testing.assert_array_equal(got[1], pdf[1].astype(np.float32))
    np.testing.assert_array_equal(gp[1], pot[1].astype(np.float32))
