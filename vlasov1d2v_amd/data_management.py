"""Run files: RunData / ReadData of main/data_management.py:5-126 on the HDF5
C library (``libvlasov1d2v_io.so``), with the state converted on the GPU.

A run file ``folder + filename + '.hdf5'`` holds the datasets "pdf" (the
ghosted state as float32, one record per write-out), "potential" and
"density" ((NX, n) float32 records), "time" and "energy" (one value per
record), with an unlimited record axis, as RunData.create_file lays them out
(main/data_management.py:13-35).  "time" and "energy" are float64 here; the
reference creates them from ``data=[0]``, which makes them int64 and
truncates what it writes later.  Files either way read back with ReadData.

Writes are asynchronous: ``save_data`` converts the state to float32 on the
GPU (``vl_export_f32``, any storage form), copies it to one of two pinned
host buffers and hands the HDF5 append to a writer thread, so stepping goes
on while the previous record is written.  ``flush()`` (or ``close()`` / the
context manager) waits for pending writes and raises their errors.
"""
import concurrent.futures
import ctypes
import os

import numpy as np

from . import _lib

HERE = os.path.dirname(os.path.abspath(__file__))
IO_LIB_PATH = os.path.join(HERE, "libvlasov1d2v_io.so")
VL_E_IO = -7

# every symbol include/vlasov1d2v_io.h declares
IO_EXPORTS = ("vl_io_last_error_string", "vl_run_data_create", "vl_run_data_append", "vl_read_data_dims",
              "vl_read_data")

_io = None


def load_io(path=IO_LIB_PATH):
    """Load the run-file library (raises if it has not been built)."""
    global _io
    if _io is not None:
        return _io
    if not os.path.exists(path):
        raise _lib.VlasovError(f"{path} is missing: build it with `python -m vlasov1d2v_amd.build` "
                               "(needs the HDF5 C library)")
    lib = ctypes.CDLL(path)
    vp, i32, i64, f64, cp = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_double, ctypes.c_char_p
    p64 = ctypes.POINTER(ctypes.c_int64)
    sig = {
        "vl_run_data_create": [cp, p64, vp, i64, i64, vp, vp],
        "vl_run_data_append": [cp, vp, i64, vp, i64, vp, i64, f64, f64],
        "vl_read_data_dims": [cp, cp, ctypes.POINTER(i32), p64, p64],
        "vl_read_data": [cp, cp, i64, i64, i32, vp],
    }
    for name, args in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = i32
    lib.vl_io_last_error_string.restype = cp
    lib.vl_io_last_error_string.argtypes = []
    _io = lib
    return lib


def _check_io(rc, what):
    if rc != 0:
        msg = load_io().vl_io_last_error_string().decode(errors="replace")
        raise _lib.VlasovError(f"{what} failed ({rc}): {msg}")


def _dataset_dims(path, name):
    """(dims, maxdims) of dataset `name` of the run file at `path`."""
    rank = ctypes.c_int()
    d = (ctypes.c_int64 * 8)()
    m = (ctypes.c_int64 * 8)()
    _check_io(load_io().vl_read_data_dims(path.encode(), name.encode(), ctypes.byref(rank), d, m),
              f"vl_read_data_dims({name})")
    return tuple(d[:rank.value]), tuple(m[:rank.value])


def _numel(a):
    return int(a.numel()) if hasattr(a, "numel") else int(a.size)


def _host_f32(a):
    """A float32 host copy of an array (the writer thread may use it after the
    caller has changed the original)."""
    try:
        import torch
        if isinstance(a, torch.Tensor):
            a = a.detach().to("cpu").numpy()
    except ImportError:
        pass
    if hasattr(a, "get"):
        a = a.get()
    return np.array(a, dtype=np.float32, order="C", copy=True)


def _state_geometry(distribution):
    """The HIP geometry of the Distribution's storage form (its background is
    registered by Distribution.to_storage)."""
    from .grid import STORAGE_CODES
    if distribution._grids is None:
        raise RuntimeError("Distribution has no grids: call initialize_quad_weights first")
    return distribution._grids.geometry(STORAGE_CODES[distribution.storage])


class _Staging:
    """Device float32 record and two pinned host buffers for one state shape."""

    def __init__(self, shape, device):
        import torch
        self.dev = torch.empty(shape, dtype=torch.float32, device=device)
        self.host = [torch.empty(shape, dtype=torch.float32, pin_memory=True) for _ in range(2)]
        self.pending = [None, None]      # write future of each host buffer
        self.next = 0


class RunData:
    """Writer of a run file (main/data_management.py:5-52)."""

    def __init__(self, folder, filename, shape, geometry, time, refs=None):
        self.write_filename = folder + filename + '.hdf5'
        self.info_name = folder + filename + '_info.txt'
        self.shape = tuple(int(s) for s in shape)
        if len(self.shape) != 6:
            raise ValueError(f"shape must be the 6-D ghosted state shape, got {self.shape}")
        # the reference always writes the run-parameter file here (main/data_management.py:11)
        if refs is not None and hasattr(refs, "info_file"):
            refs.info_file(self.info_name, geometry, time)
        self._pool = concurrent.futures.ThreadPoolExecutor(max_workers=1)
        self._staging = None
        self._last = None

    # -- state conversion ------------------------------------------------
    def _pdf_record(self, distribution):
        """(host float32 array or pinned tensor, event or None, slot or None)."""
        import torch
        from .grid import Distribution
        if isinstance(distribution, Distribution):
            arr = distribution.arr
            if not arr.is_cuda:
                raise ValueError("the Distribution's state is not on a GPU")
            if tuple(arr.shape) != self.shape:
                raise ValueError(f"state shape {tuple(arr.shape)} != file shape {self.shape}")
            g = _state_geometry(distribution)
            st = self._staging
            if st is None or st.dev.device != arr.device:
                st = self._staging = _Staging(self.shape, arr.device)
            k = st.next
            st.next ^= 1
            if st.pending[k] is not None:          # the buffer's previous write must be done
                st.pending[k].result()
                st.pending[k] = None
            _lib.check(_lib.load().vl_export_f32(g.handle, _lib.ptr(arr), _lib.ptr(st.dev), _lib.stream_ptr()),
                       "vl_export_f32")
            st.host[k].copy_(st.dev, non_blocking=True)
            ev = torch.cuda.Event()
            ev.record()
            return st.host[k], ev, k
        if isinstance(distribution, torch.Tensor):
            if distribution.is_cuda:
                raise ValueError("pass the Distribution (GPU state) or a host array")
        a = _host_f32(distribution)
        if a.shape != self.shape:
            raise ValueError(f"state shape {a.shape} != file shape {self.shape}")
        return a, None, None

    @staticmethod
    def _ptr(a):
        import torch
        return ctypes.c_void_p(a.data_ptr() if isinstance(a, torch.Tensor) else a.ctypes.data)

    # -- reference API ----------------------------------------------------
    def create_file(self, distribution, elliptic, density):
        """Creates the file with the initial record (main/data_management.py:13-35)."""
        self.flush()
        pdf, ev, _ = self._pdf_record(distribution)
        pot = _host_f32(elliptic.potential)
        den = _host_f32(density)
        if pot.ndim != 2 or den.shape != pot.shape:
            raise ValueError(f"potential {pot.shape} and density {den.shape} must be (NX, n)")
        if ev is not None:
            ev.synchronize()
        shape = (ctypes.c_int64 * 6)(*self.shape)
        _check_io(load_io().vl_run_data_create(self.write_filename.encode(), shape, self._ptr(pdf), pot.shape[0],
                                               pot.shape[1], self._ptr(pot), self._ptr(den)),
                  "vl_run_data_create")

    def save_data(self, distribution, elliptic, density, time, field_energy):
        """Appends a record (main/data_management.py:37-52).  Returns once the
        state is staged; the file write runs on the writer thread."""
        pot = _host_f32(elliptic.potential)
        den = _host_f32(density)
        # the record shapes of the file (potential and density share theirs)
        if pot.ndim != 2 or den.shape != pot.shape:
            raise ValueError(f"potential {pot.shape} and density {den.shape} must be (NX, n)")
        want = tuple(_dataset_dims(self.write_filename, "potential")[0][1:])
        if pot.shape != want:
            raise ValueError(f"potential/density records are {want}, got {pot.shape}")
        pdf, ev, slot = self._pdf_record(distribution)
        t = float(time)
        e = float(np.asarray(field_energy.cpu() if hasattr(field_energy, "cpu") else field_energy))
        path = self.write_filename.encode()
        lib = load_io()

        def write():
            if ev is not None:
                ev.synchronize()
            _check_io(lib.vl_run_data_append(path, self._ptr(pdf), _numel(pdf), self._ptr(pot), pot.size,
                                             self._ptr(den), den.size, t, e),
                      "vl_run_data_append")

        fut = self._pool.submit(write)
        if slot is not None:
            self._staging.pending[slot] = fut
        self._last = fut
        return fut

    def flush(self):
        """Waits for every pending write (re-raises a failed one)."""
        st = self._staging
        futs = [self._last] + (list(st.pending) if st is not None else [])
        for f in futs:
            if f is not None:
                f.result()
        self._last = None
        if st is not None:
            st.pending = [None, None]

    def close(self):
        self.flush()
        self._pool.shutdown(wait=True)
        self._staging = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
        return False


class ReadData:
    """Reader of a run file (main/data_management.py:55-126)."""

    def __init__(self, folder, filename):
        self.write_filename = folder + filename + '.hdf5'
        self.info_name = folder + filename + '_info.txt'

    def dims(self, name):
        """(dims, maxdims) of a dataset; maxdims -1 = unlimited."""
        return _dataset_dims(self.write_filename, name)

    def _read(self, name, start, count, as_double):
        dims, _ = self.dims(name)
        out = np.empty((count,) + tuple(dims[1:]), dtype=np.float64 if as_double else np.float32)
        _check_io(load_io().vl_read_data(self.write_filename.encode(), name.encode(), start, count,
                                         1 if as_double else 0, out.ctypes.data_as(ctypes.c_void_p)),
                  f"vl_read_data({name})")
        return out

    def _records(self, name):
        return self.dims(name)[0][0]

    def _index(self, name, idx):
        n = self._records(name)
        i = idx + n if idx < 0 else idx
        if not 0 <= i < n:
            raise IndexError(f"record {idx} of '{name}' out of range ({n} records)")
        return i

    def _range(self, name, idx_range):
        start, stop, _ = slice(idx_range[0], idx_range[1]).indices(self._records(name))
        return start, max(stop - start, 0)

    def read_data(self):
        """(time, pdf, potential, density, energy), whole datasets (main/data_management.py:60-68)."""
        t = self._read("time", 0, self._records("time"), True)
        pdf = self._read("pdf", 0, self._records("pdf"), False)
        pot = self._read("potential", 0, self._records("potential"), False)
        den = self._read("density", 0, self._records("density"), False)
        eng = self._read("energy", 0, self._records("energy"), True)
        return t, pdf, pot, den, eng

    def read_specific_time(self, idx):
        """(time, pdf, potential) of record idx (main/data_management.py:70-76)."""
        t = self._read("time", self._index("time", idx), 1, True)[0]
        pdf = self._read("pdf", self._index("pdf", idx), 1, False)[0]
        pot = self._read("potential", self._index("potential", idx), 1, False)[0]
        return t, pdf, pot

    def read_potential_only(self, idx_range):
        """(time, potential) of records idx_range[0]:idx_range[1] (main/data_management.py:78-84)."""
        s, c = self._range("time", idx_range)
        t = self._read("time", s, c, True)
        s, c = self._range("potential", idx_range)
        pot = self._read("potential", s, c, False)
        return t, pot

    def read_info(self):
        """Run parameters from the ``_info.txt`` file that the reference's
        Reference.info_file writes, parsed from the same lines
        (main/data_management.py:87-126)."""
        with open(self.info_name, 'r') as params:
            p = params.readlines()
        order_x = int(float(p[46].split()[3]))
        res_x = int(float(p[34].split()[2]))
        x_low = float(p[37].split()[3])
        x_high = float(p[38].split()[3])
        order_u = int(float(p[47].split()[3]))
        res_u = int(float(p[35].split()[2]))
        u_low = float(p[39].split()[3])
        u_high = float(p[40].split()[3])
        order_v = int(float(p[48].split()[3]))
        res_v = int(float(p[36].split()[2]))
        # the reference parses the v bounds with int() (main/data_management.py:106-107)
        v_low = int(float(p[41].split()[3]))
        v_high = int(float(p[42].split()[3]))
        final_time = float(p[43].split()[3])
        write_time = float(p[44].split()[4])
        order_time = float(p[49].split()[3])
        n0 = float(p[6].split()[2])
        T0 = float(p[7].split()[2])
        L0 = float(p[8].split()[3])
        orders = np.array([order_x, order_u, order_v])
        resolutions = np.array([res_x, res_u, res_v])
        lows = np.array([x_low, u_low, v_low])
        highs = np.array([x_high, u_high, v_high])
        time_info = np.array([final_time, write_time, order_time])
        refs = np.array([n0, T0, L0])
        return orders, resolutions, lows, highs, time_info, refs


def resume(distribution, reader, idx=-1):
    """Restarts from record idx of a run file: sets the Distribution's GPU
    state from the float32 "pdf" record (``vl_import_f32``: FP64 f, or the
    stored deviation of a reduced form) and returns the record's time.  The
    reference has no resume; a record holds f to float32 precision, so a
    resumed run continues from f rounded to float32."""
    import torch
    if not isinstance(reader, ReadData):
        folder, filename = os.path.split(str(reader))
        reader = ReadData(folder + os.sep if folder else "", filename[:-5] if filename.endswith(".hdf5") else filename)
    t, pdf, _ = reader.read_specific_time(idx)
    arr = distribution.arr
    if tuple(pdf.shape) != tuple(arr.shape):
        raise ValueError(f"record shape {pdf.shape} != state shape {tuple(arr.shape)}")
    src = torch.from_numpy(pdf).to(arr.device)
    g = _state_geometry(distribution)
    _lib.check(_lib.load().vl_import_f32(g.handle, _lib.ptr(src), _lib.ptr(arr), _lib.stream_ptr()),
               "vl_import_f32")
    torch.cuda.current_stream(arr.device).synchronize()
    return float(t)
