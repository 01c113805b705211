"""Grids and the distribution state: drop-ins for the reference's ``Grid1D``,
``Grid3D`` and ``Distribution`` (main/grid.py:10-321).

Coordinates, quadrature weights and spectral tables are host FP64 setup data.
The distribution function lives on the GPU as a torch tensor of the
reference's ghosted layout (NX+2, n, NU+2, n, NV+2, n); its moment and ghost
exchange run in the HIP library.
"""
import ctypes
import math

import numpy as np

from . import _lib
from .basis import ring_factor


class Grid1D:
    """One phase-space axis with one ghost element per side (main/grid.py:10-88)."""

    def __init__(self, low, high, res, basis, spectrum=False, fine=False):
        self.low = float(low)
        self.high = float(high)
        self.res = int(res)
        self.res_ghosts = int(res + 2)
        self.order = basis.order
        self.basis = basis
        self.length = self.high - self.low
        self.dx = self.length / self.res
        self.J = 2.0 / self.dx
        self.quad_weights = np.multiply.outer(np.ones(self.res), np.asarray(basis.weights))
        self.arr = np.zeros((self.res_ghosts, self.order))
        self.create_grid(basis.nodes)
        self.midpoints = np.array([(self.arr[i, -1] + self.arr[i, 0]) / 2.0
                                   for i in range(1, self.res_ghosts - 1)])
        self.arr_max = float(np.amax(np.abs(self.arr)))
        self.one_negatives = np.where(self.arr < 0, 1, 0)
        self.one_positives = np.where(self.arr >= 0, 1, 0)
        if fine:
            fine_num = 25
            self.arr_fine = np.array([np.linspace(self.arr[i, 0], self.arr[i, -1], num=fine_num)
                                      for i in range(self.res_ghosts)])
        self.spectrum = bool(spectrum)
        self._geometry = None
        self._owner = None
        if spectrum:
            # mode count from the reference's float floor-division (main/grid.py:46)
            self.nyquist_number = 2.5 * self.length // self.dx
            self.k1 = 2.0 * np.pi / self.length
            self.wave_numbers = self.k1 * np.arange(1 - self.nyquist_number, self.nyquist_number)
            self.d_wave_numbers = self.wave_numbers
            self.grid_phases = np.exp(1j * np.multiply.outer(self.wave_numbers, self.arr[1:-1, :]))
            self.spectral_transform = basis.fourier_transform_array(self.midpoints, self.J,
                                                                    self.wave_numbers)

    @property
    def arr_cp(self):
        return self.arr

    def create_grid(self, nodes):
        """Element coordinates including ghosts (main/grid.py:55-70)."""
        min_gs = self.low - self.dx
        max_gs = self.high
        nodes = (np.array(nodes) + 1) / 2
        xl = np.linspace(min_gs, max_gs, num=self.res_ghosts)
        for i in range(self.res_ghosts):
            self.arr[i, :] = xl[i] + self.dx * nodes

    # Host-side spectral utilities (setup / analysis; the per-step Poisson
    # solve runs on the GPU in Elliptic.poisson2).
    def fourier_basis(self, function):
        """Nodal DFT coefficients of an (res, n) grid function (main/grid.py:78-82)."""
        return np.tensordot(np.asarray(function), self.spectral_transform,
                            axes=([0, 1], [0, 1])) * self.dx / self.length

    def sum_fourier(self, coefficients):
        """Real part of the Fourier re-summation at the nodes (main/grid.py:84-88)."""
        return np.real(np.tensordot(np.asarray(coefficients), self.grid_phases, axes=([0], [0])))

    def geometry(self):
        """HIP geometry handle carrying this axis' spectral tables."""
        if self._owner is not None:
            return self._owner.geometry()
        if self._geometry is None:
            if not self.spectrum:
                raise ValueError("the Poisson solve needs a Grid1D built with spectrum=True")
            b = self.basis
            dummy = np.array([[-1.0] * self.order, [0.0] * self.order, [1.0] * self.order])
            self._geometry = _lib.Geometry(
                self.order, self.res, 1, 1, b.up, b.xi, b.weights, self.arr, dummy, dummy,
                self.J, 1.0, 1.0, self.dx, self.length, self.wave_numbers,
                self.spectral_transform, self.grid_phases)
        return self._geometry


class Grid3D:
    """x (spectral), u and v axes (main/grid.py:91-109)."""

    def __init__(self, basis, lows, highs, resolutions, fine_x=False, fine_all=False):
        fx = bool(fine_x or fine_all)
        self.basis = basis
        self.x = Grid1D(lows[0], highs[0], resolutions[0], basis.b1, spectrum=True, fine=fx)
        self.u = Grid1D(lows[1], highs[1], resolutions[1], basis.b2, fine=fine_all)
        self.v = Grid1D(lows[2], highs[2], resolutions[2], basis.b3, fine=fine_all)
        if not (basis.b1.order == basis.b2.order == basis.b3.order):
            raise ValueError("the HIP path needs the same number of nodes on x, u and v")
        self.x._owner = self
        self._geometries = {}
        self.no_ghost_slice = (slice(1, self.x.res_ghosts - 1), slice(self.x.order),
                               slice(1, self.u.res_ghosts - 1), slice(self.u.order),
                               slice(1, self.v.res_ghosts - 1), slice(self.v.order))
        self.shape = (self.x.res_ghosts, self.x.order, self.u.res_ghosts, self.u.order,
                      self.v.res_ghosts, self.v.order)

    def geometry(self, storage=_lib.VL_F64):
        """The HIP geometry handle of this phase-space grid for one storage
        type of f (created once per storage type)."""
        g = self._geometries.get(storage)
        if g is None:
            b = self.basis.b1
            g = _lib.Geometry(
                b.order, self.x.res, self.u.res, self.v.res, b.up, b.xi, b.weights,
                self.x.arr, self.u.arr, self.v.arr, self.x.J, self.u.J, self.v.J,
                self.x.dx, self.x.length, self.x.wave_numbers, self.x.spectral_transform,
                self.x.grid_phases, storage=storage)
            self._geometries[storage] = g
        return g


STORAGE_CODES = {"f64": _lib.VL_F64, "f32": _lib.VL_F32, "bf16": _lib.VL_BF16, "f16": _lib.VL_F16}
# stored g = (f - f0) * scale for the reduced forms (FP16 keeps g in its normal range)
STORAGE_SCALE = {"f64": 1.0, "f32": 1.0, "bf16": 1.0, "f16": 65536.0}


def storage_dtype(storage):
    import torch
    return {"f64": torch.float64, "f32": torch.float32, "bf16": torch.bfloat16,
            "f16": torch.float16}[storage]


def storage_of(arr):
    import torch
    m = {torch.float64: "f64", torch.float32: "f32", torch.bfloat16: "bf16", torch.float16: "f16"}
    if arr.dtype not in m:
        raise TypeError(f"unsupported state dtype {arr.dtype}")
    return m[arr.dtype]


def _check_state(arr, shape, what, allow_reduced=False):
    import torch
    if not isinstance(arr, torch.Tensor):
        raise TypeError(f"{what}: expected a torch tensor on the GPU")
    if arr.device.type != "cuda":
        raise _lib.VlasovError(f"{what}: the state must be a GPU tensor (no CPU fallback)")
    if tuple(arr.shape) != tuple(shape):
        raise ValueError(f"{what}: shape {tuple(arr.shape)} != {tuple(shape)}")
    if not arr.is_contiguous():
        raise ValueError(f"{what}: tensor must be contiguous")
    if arr.dtype != torch.float64 and not (allow_reduced and storage_of(arr) != "f64"):
        raise TypeError(f"{what}: expected float64")


class Distribution:
    """Phase-space distribution f on the GPU (main/grid.py:137-321)."""

    def __init__(self, vt, ring_j, resolutions, orders, perturbation=True, om=0.9, om_pc=1.0,
                 delta_n=0.01, device="cuda"):
        self.ring_j = ring_j
        self.vt = vt
        self.om = om
        self.om_pc = om_pc
        self.delta_n = delta_n
        self.perturbation = perturbation
        self.arr = None
        self.x_res, self.u_res, self.v_res = (int(r) for r in resolutions)
        self.x_ord, self.u_ord, self.v_ord = (int(o) for o in orders)
        self.quad_weights = None
        self.device = device
        self._grids = None
        self.storage = "f64"
        self.scale = 1.0
        self.f0 = None          # background ring (NU+2, n, NV+2, n) of the reduced forms

    def initialize_quad_weights(self, grids):
        """W[eu,iu,ev,iv] = w_iu w_iv / (Ju Jv) (main/grid.py:297-301); host copy."""
        self.quad_weights = np.multiply.outer(grids.u.quad_weights,
                                              grids.v.quad_weights) / (grids.u.J * grids.v.J)
        self._grids = grids

    def ring(self, grids):
        """Ring distribution on every ghosted node, host FP64 (main/grid.py:252-268)."""
        factor = ring_factor(self.ring_j, self.vt)
        u2 = grids.u.arr ** 2.0
        v2 = grids.v.arr ** 2.0
        rsq = (u2[:, :, None, None] + v2[None, None, :, :]) / (self.vt ** 2.0)
        pol = np.multiply(rsq ** self.ring_j, np.exp(-0.5 * rsq))
        return factor * pol                                   # (NU+2, n, NV+2, n)

    def initialize_gpu(self, grids, modulation=None):
        """Initial state on every ghosted node.

        Without ``modulation`` this is the reference's own initial condition
        (main/grid.py:244-279): the ring plus, when ``perturbation``, the
        Bernstein-mode eigenfunction with this distribution's ``om``,
        ``om_pc`` and ``delta_n`` (main/grid.py:207-242), computed on the
        device (``vl_init_eigen``).  With ``modulation``, an (NX+2, n) host
        array, it is ring(u, v) * modulation(x) (the workloads' condition).
        """
        import torch
        self._grids = grids
        if modulation is None:
            self.arr = self._eigen_state(grids, pert_only=False)
            self._ic = "eigen"
            if self.quad_weights is None:
                self.initialize_quad_weights(grids)
            return
        self.arr = torch.empty(grids.shape, dtype=torch.float64, device=self.device)
        self._separable(grids.geometry(), self.ring(grids), modulation, self.arr)
        if self.quad_weights is None:
            self.initialize_quad_weights(grids)

    def _separable(self, geom, vel, xprofile, out):
        """out = xprofile(x) vel(u, v) in geom's storage form (vl_init_separable);
        vel (NU+2, n, NV+2, n) and xprofile (NX+2, n) host arrays, uploaded as buffers."""
        import torch
        vel_d = torch.from_numpy(np.ascontiguousarray(vel, dtype=np.float64)).to(self.device)
        xp_d = torch.from_numpy(np.ascontiguousarray(xprofile, dtype=np.float64)).to(self.device)
        if vel_d.numel() != (geom.nu + 2) * geom.n * (geom.nv + 2) * geom.n or \
                xp_d.numel() != (geom.nx + 2) * geom.n:
            raise ValueError("separable state: table sizes do not match the geometry")
        _lib.check(_lib.load().vl_init_separable(geom.handle, _lib.ptr(vel_d), _lib.ptr(xp_d), _lib.ptr(out),
                                                 _lib.stream_ptr()), "vl_init_separable")

    def _eigen_state(self, grids, pert_only):
        """ring + eigenfunction (or the eigenfunction alone) as an FP64 GPU tensor."""
        import torch
        g = grids.geometry()
        om = complex(self.om)
        d = _lib.VlEigenDesc(ring_j=int(self.ring_j), terms=20, perturbation=1 if self.perturbation else 0,
                             pert_only=1 if pert_only else 0, vt=float(self.vt), k1=float(grids.x.k1),
                             om_re=om.real, om_im=om.imag, om_pc=float(self.om_pc), delta_n=float(self.delta_n))
        out = torch.empty(grids.shape, dtype=torch.float64, device=self.device)
        _lib.check(_lib.load().vl_init_eigen(g.handle, ctypes.byref(d), _lib.ptr(out), _lib.stream_ptr()),
                   "vl_init_eigen")
        return out

    def initialize_ring_modulated(self, grids, amp1, amps, phases):
        """f = ring(u,v) * (1 + amp1 cos(k1 x) + sum_m amps[m] cos((m+2) k1 x + phases[m]))."""
        x = grids.x.arr
        k1 = grids.x.k1
        mod = 1.0 + amp1 * np.cos(k1 * x)
        pert = amp1 * np.cos(k1 * x)
        for m, (a, ph) in enumerate(zip(amps, phases)):
            mod = mod + a * np.cos((m + 2) * k1 * x + ph)
            pert = pert + a * np.cos((m + 2) * k1 * x + ph)
        self._pert = pert
        self._ic = "modulated"
        self.initialize_gpu(grids, modulation=mod)

    def to_storage(self, grids, storage):
        """Switch to a reduced-precision form: the state becomes
        g = (f - f0(u, v)) * scale in the storage dtype, with f0 the ring and
        f - f0 = ring * (modulation - 1) formed without cancellation.
        The background f0 is registered with the HIP geometry."""
        import torch
        if storage == "f64":
            return
        ic = getattr(self, "_ic", None)
        if ic not in ("modulated", "eigen"):
            raise RuntimeError("to_storage needs initialize_ring_modulated or initialize_gpu first")
        self.storage = storage
        self.scale = STORAGE_SCALE[storage]
        ring_h = self.ring(grids)
        self.f0 = ring_h
        geo = grids.geometry(STORAGE_CODES[storage])
        out = torch.empty(grids.shape, dtype=storage_dtype(storage), device=self.device)
        if ic == "eigen":
            d = self._eigen_state(grids, pert_only=True)   # f - f0: the eigenfunction alone
            _lib.check(_lib.load().vl_deviation_to_storage(geo.handle, _lib.ptr(d), _lib.ptr(out),
                                                           _lib.stream_ptr()), "vl_deviation_to_storage")
            del d
        else:
            self._separable(geo, ring_h, self._pert, out)  # f - f0 = ring (modulation - 1)
        self.arr = out
        _lib.check(_lib.load().vl_set_background(geo.handle, _lib.ptr(np.ascontiguousarray(ring_h))),
                   "vl_set_background")

    def total(self):
        """f in FP64 on the GPU (f itself, or f0 + g / scale for the reduced forms)."""
        import torch
        if self.storage == "f64":
            return self.arr
        geo = self._grids.geometry(STORAGE_CODES[self.storage])
        out = torch.empty(tuple(self.arr.shape), dtype=torch.float64, device=self.arr.device)
        _lib.check(_lib.load().vl_total_f64(geo.handle, _lib.ptr(self.arr.contiguous()), _lib.ptr(out),
                                            _lib.stream_ptr()), "vl_total_f64")
        return out

    def _geom(self):
        if self._grids is None:
            raise RuntimeError("Distribution has no grids: call initialize_quad_weights first")
        return self._grids.geometry()

    def swap_ghost_cells_of_array(self):
        """Periodic x ghosts (main/grid.py:187-189), in place on the GPU."""
        g = self._geom()
        _check_state(self.arr, self._grids.shape, "swap_ghost_cells_of_array")
        _lib.check(_lib.load().vl_swap_ghosts(g.handle, _lib.ptr(self.arr), _lib.stream_ptr()),
                   "vl_swap_ghosts")

    def moment_zero(self):
        """Zeroth velocity moment of the interior (main/grid.py:303-312) -> (NX, n) FP64."""
        return self.moment_zero_of_arr(self.arr)

    def moment_zero_of_arr(self, arr):
        """Zeroth moment of a state array (main/grid.py:319-321)."""
        import torch
        g = self._geom()
        _check_state(arr, self._grids.shape, "moment_zero_of_arr")
        rho = torch.empty((g.nx, g.n), dtype=torch.float64, device=arr.device)
        ws = g.workspace(arr.device)
        _lib.check(_lib.load().vl_moment0(g.handle, _lib.ptr(arr), _lib.ptr(rho), _lib.ptr(ws),
                                          _lib.stream_ptr()), "vl_moment0")
        return rho

    def grid_flatten_gpu(self):
        return self.arr.reshape((self.x_res + 2) * self.x_ord, (self.u_res + 2) * self.u_ord,
                                (self.v_res + 2) * self.v_ord)


def seeded_phases(seed, count=3):
    """Phases phi_m ~ U[0, 2 pi) from numpy's default_rng(seed)."""
    rng = np.random.default_rng(seed)
    return rng.uniform(0.0, 2.0 * math.pi, size=count)
