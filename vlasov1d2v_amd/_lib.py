"""ctypes binding of ``libvlasov1d2v.so`` (declared in include/vlasov1d2v.h).

The library is loaded from this package directory only.  A missing library is
an error: there is no CPU fallback for any device entry point.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvlasov1d2v.so")

VL_F64, VL_F32, VL_BF16, VL_F16 = 0, 1, 2, 3
VL_OK, VL_E_ARG, VL_E_NULL, VL_E_ALIGN, VL_E_HANDLE, VL_E_WORKSPACE, VL_E_NOTFINITE = 0, -1, -2, -3, -4, -5, -6
VL_STEP_RHO_VALID = 1

# Every symbol the header declares (checked by the CPU test suite).
EXPORTS = (
    "vl_geom_create", "vl_geom_destroy", "vl_workspace_bytes", "vl_state_elems",
    "vl_rhs", "vl_moment0", "vl_poisson", "vl_field_energy", "vl_max_abs_acc_e",
    "vl_swap_ghosts", "vl_rk3_step", "vl_set_background", "vl_set_physics", "vl_set_stage_events",
    "vl_workspace_init", "vl_set_tuning", "vl_get_tuning", "vl_state_elems_layout",
    "vl_convert_layout", "vl_resident_layout", "vl_last_error_string",
    "vl_ensemble_create", "vl_ensemble_destroy", "vl_ensemble_bind", "vl_ensemble_set_schedule",
    "vl_ensemble_get_schedule", "vl_ensemble_moments", "vl_ensemble_rk3_step", "vl_ensemble_autotune",
    "vl_set_cfl", "vl_clock_init", "vl_rk3_step_adaptive", "vl_init_eigen", "vl_set_central_flux",
    "vl_poisson_central", "vl_export_f32", "vl_import_f32", "vl_ensemble_set_adaptive", "vl_ensemble_clocks",
    "vl_moment_partials", "vl_rk3_step_noalloc", "vl_read_status",
    "vl_axpby", "vl_init_separable", "vl_deviation_to_storage", "vl_total_f64", "vl_density_deviation",
    "vl_graph_begin", "vl_graph_end", "vl_graph_launch", "vl_graph_destroy", "vl_stream_probe",
)
VL_TUNE_XSEG = 1
VL_TUNE_MFMA = 2
VL_TUNE_ROW_SPLIT = 3
VL_TUNE_ROUNDING = 4
VL_TUNE_RS_KERNEL = 5
VL_STEP_ELEMENT_MAJOR = 2
VL_STEP_ROW_SPLIT = 4
VL_LAYOUT_REFERENCE, VL_LAYOUT_ELEMENT_MAJOR, VL_LAYOUT_ROW_SPLIT = 0, 1, 2


class VlGeomDesc(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int32), ("nx", ctypes.c_int32), ("nu", ctypes.c_int32),
        ("nv", ctypes.c_int32), ("n_modes", ctypes.c_int32), ("storage", ctypes.c_int32),
        ("jx", ctypes.c_double), ("ju", ctypes.c_double), ("jv", ctypes.c_double),
        ("acc", ctypes.c_double), ("b_field", ctypes.c_double), ("poisson_coef", ctypes.c_double),
        ("x_dx", ctypes.c_double), ("x_length", ctypes.c_double),
        ("up", ctypes.c_void_p), ("xi", ctypes.c_void_p), ("weights", ctypes.c_void_p),
        ("x_nodes", ctypes.c_void_p), ("u_nodes", ctypes.c_void_p), ("v_nodes", ctypes.c_void_p),
        ("wave_numbers", ctypes.c_void_p), ("transform", ctypes.c_void_p), ("phases", ctypes.c_void_p),
    ]


class VlEigenDesc(ctypes.Structure):
    _fields_ = [
        ("ring_j", ctypes.c_int32), ("terms", ctypes.c_int32), ("perturbation", ctypes.c_int32),
        ("pert_only", ctypes.c_int32), ("vt", ctypes.c_double), ("k1", ctypes.c_double),
        ("om_re", ctypes.c_double), ("om_im", ctypes.c_double), ("om_pc", ctypes.c_double),
        ("delta_n", ctypes.c_double),
    ]


class VlasovError(RuntimeError):
    pass


_lib = None


def load(path=LIB_PATH):
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise VlasovError(
            f"{path} is missing: build it with `python -m vlasov1d2v_amd.build` "
            "(there is no CPU fallback)")
    lib = ctypes.CDLL(path)
    vp, i32, f64, sz = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_size_t
    sig = {
        "vl_geom_create": [ctypes.POINTER(VlGeomDesc), ctypes.POINTER(vp)],
        "vl_geom_destroy": [vp],
        "vl_workspace_bytes": [vp, ctypes.POINTER(sz)],
        "vl_state_elems": [vp, ctypes.POINTER(sz)],
        "vl_rhs": [vp, vp, vp, vp, vp],
        "vl_moment0": [vp, vp, vp, vp, vp],
        "vl_poisson": [vp, vp, vp, vp, vp, vp],
        "vl_field_energy": [vp, vp, vp, vp],
        "vl_max_abs_acc_e": [vp, vp, vp, vp],
        "vl_swap_ghosts": [vp, vp, vp],
        "vl_rk3_step": [vp, vp, vp, vp, vp, vp, f64, vp, vp, i32, vp],
        "vl_set_background": [vp, vp],
        "vl_set_physics": [vp, f64, f64, f64],
        "vl_set_stage_events": [vp, vp],
        "vl_workspace_init": [vp, vp, vp],
        "vl_set_tuning": [vp, i32, i32],
        "vl_get_tuning": [vp, i32, ctypes.POINTER(i32)],
        "vl_state_elems_layout": [vp, i32, ctypes.POINTER(sz)],
        "vl_convert_layout": [vp, vp, i32, vp, i32, vp],
        "vl_resident_layout": [vp, ctypes.POINTER(i32)],
        "vl_init_eigen": [vp, ctypes.POINTER(VlEigenDesc), vp, vp],
        "vl_set_central_flux": [vp, vp, vp, vp],
        "vl_poisson_central": [vp, vp, vp, vp, i32, vp],
        "vl_export_f32": [vp, vp, vp, vp],
        "vl_import_f32": [vp, vp, vp, vp],
        "vl_axpby": [vp, vp, vp, f64, f64, i32, vp],
        "vl_init_separable": [vp, vp, vp, vp, vp],
        "vl_deviation_to_storage": [vp, vp, vp, vp],
        "vl_total_f64": [vp, vp, vp, vp],
        "vl_density_deviation": [vp, vp, vp, i32, vp],
        "vl_graph_begin": [vp],
        "vl_graph_end": [vp, ctypes.POINTER(vp)],
        "vl_graph_launch": [vp, vp],
        "vl_graph_destroy": [vp],
        "vl_stream_probe": [vp, vp, vp, sz, i32, i32, vp],
        "vl_set_cfl": [vp, f64, f64, f64, f64],
        "vl_clock_init": [vp, vp, f64, vp, vp],
        "vl_rk3_step_adaptive": [vp, vp, vp, vp, vp, vp, vp, vp, vp, i32, vp],
        "vl_moment_partials": [vp, vp, vp, i32, vp],
        "vl_read_status": [vp, vp, i32, ctypes.POINTER(ctypes.c_uint), vp],
        "vl_rk3_step_noalloc": [vp, vp, vp, vp, vp, vp, f64, vp, i32, vp],
        "vl_ensemble_create": [ctypes.POINTER(vp), i32, ctypes.POINTER(vp)],
        "vl_ensemble_destroy": [vp],
        "vl_ensemble_set_adaptive": [vp, i32, i32],
        "vl_ensemble_clocks": [vp, vp, vp],
        "vl_ensemble_bind": [vp, i32, vp, vp, vp, vp, vp, f64],
        "vl_ensemble_set_schedule": [vp, i32, i32, i32],
        "vl_ensemble_get_schedule": [vp, ctypes.POINTER(i32), ctypes.POINTER(i32), ctypes.POINTER(i32),
                                     ctypes.POINTER(i32)],
        "vl_ensemble_moments": [vp, vp],
        "vl_ensemble_rk3_step": [vp, vp, vp],
        "vl_ensemble_autotune": [vp, i32, ctypes.POINTER(ctypes.c_float), i32, ctypes.POINTER(i32), vp],
    }
    for name, args in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int
    lib.vl_last_error_string.argtypes = []
    lib.vl_last_error_string.restype = ctypes.c_char_p
    _lib = lib
    return lib


# Entry points that write a workspace's moment partials or a state array.
# A stepper passes VL_STEP_RHO_VALID only while none of them ran since its
# own previous step (see ``generation``).
WRITERS = frozenset({
    "vl_moment0", "vl_moment_partials", "vl_rk3_step", "vl_rk3_step_adaptive", "vl_rk3_step_noalloc",
    "vl_convert_layout",
    "vl_import_f32", "vl_init_eigen", "vl_axpby", "vl_init_separable", "vl_deviation_to_storage",
    "vl_ensemble_rk3_step", "vl_ensemble_moments", "vl_ensemble_autotune",
})
_generation = 0


def generation():
    """Count of the calls of WRITERS entry points made so far (through check)."""
    return _generation


def check(rc, what):
    global _generation
    if what.split("(")[0] in WRITERS:
        _generation += 1
    if rc != 0:
        msg = load().vl_last_error_string().decode(errors="replace")
        raise VlasovError(f"{what} failed (code {rc}): {msg}")


VL_STATUS_NONFINITE = 1


def raise_for_status(status, what):
    """Raise VlasovError when a workspace status word (vl_read_status) has
    VL_STATUS_NONFINITE set."""
    if int(status) & VL_STATUS_NONFINITE:
        raise VlasovError(f"{what}: the field energy or max|acc E| of a stage-3 field solve is not finite "
                          "(the state has overflowed -- e.g. a time step above the stability limit -- "
                          "or holds NaN)")


def ptr(t):
    """Device (or host) address of a tensor / array, or None."""
    if t is None:
        return None
    if isinstance(t, np.ndarray):
        return t.ctypes.data
    return t.data_ptr()


def axpby(geom, y, x, a, b, interior=True):
    """y = a y + b x on the (interior) nodes of an FP64 state (vl_axpby)."""
    check(load().vl_axpby(geom.handle, ptr(y), ptr(x), float(a), float(b), 1 if interior else 0, stream_ptr()),
          "vl_axpby")


def density_deviation(geom, rho, sign=1):
    """mean(rho) - rho (sign 1) or rho - mean(rho) (sign -1), a new (NX, n)
    FP64 GPU tensor (vl_density_deviation)."""
    import torch
    if not isinstance(rho, torch.Tensor) or rho.device.type != "cuda" or rho.dtype != torch.float64 \
            or not rho.is_contiguous() or tuple(rho.shape) != (geom.nx, geom.n):
        raise ValueError(f"density_deviation: rho must be a contiguous float64 GPU tensor of shape "
                         f"{(geom.nx, geom.n)}")
    out = torch.empty_like(rho)
    check(load().vl_density_deviation(geom.handle, ptr(rho), ptr(out), int(sign), stream_ptr()),
          "vl_density_deviation")
    return out


class HipGraph:
    """A HIP graph of library calls (vl_graph_*): ``with g.capture(stream):``
    records the calls made on ``stream`` (a non-default torch stream) inside
    the block; ``launch()`` replays them on the current stream."""

    def __init__(self):
        self.handle = None

    def capture(self, stream):
        import contextlib
        import torch

        @contextlib.contextmanager
        def cm():
            if self.handle is not None:
                raise RuntimeError("HipGraph: already captured")
            h = ctypes.c_void_p()
            with torch.cuda.stream(stream):
                check(load().vl_graph_begin(stream.cuda_stream), "vl_graph_begin")
                try:
                    yield self
                except BaseException:
                    # end the capture (the stream must leave capture mode) and drop the graph
                    if load().vl_graph_end(stream.cuda_stream, ctypes.byref(h)) == 0 and h.value:
                        load().vl_graph_destroy(h.value)
                    raise
                check(load().vl_graph_end(stream.cuda_stream, ctypes.byref(h)), "vl_graph_end")
            self.handle = h.value
        return cm()

    def launch(self, stream=None):
        if self.handle is None:
            raise RuntimeError("HipGraph: nothing captured")
        check(load().vl_graph_launch(self.handle, stream_ptr(stream)), "vl_graph_launch")

    def __del__(self):
        h, self.handle = self.handle, None
        if h is not None and _lib is not None:
            _lib.vl_graph_destroy(h)


def stream_ptr(stream=None):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream


class Geometry:
    """Owner of one ``vl_geom_t`` handle (operator tables on the device)."""

    def __init__(self, n, nx, nu, nv, up, xi, weights, x_nodes, u_nodes, v_nodes,
                 jx, ju, jv, x_dx, x_length, wave_numbers, transform, phases,
                 acc=-1.0, b_field=0.0, poisson_coef=-1.0, storage=VL_F64):
        lib = load()
        self._keep = []

        def arr(a, dtype=np.float64):
            a = np.ascontiguousarray(a, dtype=dtype)
            self._keep.append(a)
            return a.ctypes.data

        tr = np.ascontiguousarray(transform, dtype=np.complex128)
        ph = np.ascontiguousarray(phases, dtype=np.complex128)
        d = VlGeomDesc()
        d.n, d.nx, d.nu, d.nv = int(n), int(nx), int(nu), int(nv)
        d.n_modes = int(len(wave_numbers))
        d.storage = int(storage)
        d.jx, d.ju, d.jv = float(jx), float(ju), float(jv)
        d.acc, d.b_field, d.poisson_coef = float(acc), float(b_field), float(poisson_coef)
        d.x_dx, d.x_length = float(x_dx), float(x_length)
        d.up = arr(up)
        d.xi = arr(xi)
        d.weights = arr(weights)
        d.x_nodes = arr(x_nodes)
        d.u_nodes = arr(u_nodes)
        d.v_nodes = arr(v_nodes)
        d.wave_numbers = arr(wave_numbers)
        d.transform = arr(tr.view(np.float64))
        d.phases = arr(ph.view(np.float64))
        h = ctypes.c_void_p()
        check(lib.vl_geom_create(ctypes.byref(d), ctypes.byref(h)), "vl_geom_create")
        self._keep = []
        self.handle = h
        self.n, self.nx, self.nu, self.nv = int(n), int(nx), int(nu), int(nv)
        self.storage = storage
        self.acc, self.b_field, self.poisson_coef = float(acc), float(b_field), float(poisson_coef)
        b = ctypes.c_size_t()
        check(lib.vl_workspace_bytes(h, ctypes.byref(b)), "vl_workspace_bytes")
        self.workspace_bytes = int(b.value)
        e = ctypes.c_size_t()
        check(lib.vl_state_elems(h, ctypes.byref(e)), "vl_state_elems")
        self.state_elems = int(e.value)
        self._ws = None

    def set_physics(self, acc, b_field, poisson_coef):
        acc, b_field, poisson_coef = float(acc), float(b_field), float(poisson_coef)
        if (acc, b_field, poisson_coef) != (self.acc, self.b_field, self.poisson_coef):
            check(load().vl_set_physics(self.handle, acc, b_field, poisson_coef), "vl_set_physics")
            self.acc, self.b_field, self.poisson_coef = acc, b_field, poisson_coef

    def workspace(self, device):
        """Device workspace owned by this geometry (allocated once)."""
        import torch
        if self._ws is None or self._ws.device != torch.device(device):
            self._ws = torch.zeros(max(self.workspace_bytes, 256), dtype=torch.uint8, device=device)
            check(load().vl_workspace_init(self.handle, ptr(self._ws), stream_ptr()), "vl_workspace_init")
        return self._ws

    def read_status(self, device, clear=False):
        """The status word of this geometry's workspace (synchronises)."""
        st = ctypes.c_uint()
        check(load().vl_read_status(self.handle, ptr(self.workspace(device)), 1 if clear else 0, ctypes.byref(st),
                                    stream_ptr()), "vl_read_status")
        return int(st.value)

    def check_finite(self, device, what):
        """Raise VlasovError if a step on this workspace met a non-finite field."""
        raise_for_status(self.read_status(device), what)

    def set_tuning(self, key, value):
        check(load().vl_set_tuning(self.handle, int(key), int(value)), "vl_set_tuning")

    def state_elems_layout(self, layout):
        e = ctypes.c_size_t()
        check(load().vl_state_elems_layout(self.handle, int(layout), ctypes.byref(e)), "vl_state_elems_layout")
        return int(e.value)

    def convert_layout(self, src, src_layout, dst, dst_layout):
        check(load().vl_convert_layout(self.handle, ptr(src), int(src_layout), ptr(dst), int(dst_layout),
                                       stream_ptr()), "vl_convert_layout")

    def resident_layout(self):
        """The layout vl_rk3_step runs fastest in (VL_LAYOUT_*)."""
        v = ctypes.c_int()
        check(load().vl_resident_layout(self.handle, ctypes.byref(v)), "vl_resident_layout")
        return int(v.value)

    def set_cfl(self, courant, dx, du, dv):
        check(load().vl_set_cfl(self.handle, float(courant), float(dx), float(du), float(dv)), "vl_set_cfl")

    def get_tuning(self, key):
        v = ctypes.c_int()
        check(load().vl_get_tuning(self.handle, int(key), ctypes.byref(v)), "vl_get_tuning")
        return int(v.value)

    def close(self):
        if getattr(self, "handle", None):
            load().vl_geom_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
