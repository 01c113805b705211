"""Semi-discrete upwind DG right-hand side on the GPU: drop-in for the
reference's ``DGFlux`` (main/fluxes.py:13-207)."""
import numpy as np

from . import _lib


class DGFlux:
    """``resolutions`` are the ghosted element counts (main/main.py:148)."""

    def __init__(self, resolutions, orders, flux_coefficients):
        self.resolutions = np.asarray(resolutions)
        self.orders = np.asarray(orders)
        self.acceleration_coefficient = float(flux_coefficients)
        self.grid_axis = np.array([0, 2, 4])
        self.sub_element_axis = np.array([1, 3, 5])

    def _geom(self, grids, elliptic, storage=_lib.VL_F64):
        g = grids.geometry(storage)
        exp = tuple(int(r) for r in self.resolutions)
        got = (grids.x.res_ghosts, grids.u.res_ghosts, grids.v.res_ghosts)
        if exp != got:
            raise ValueError(f"DGFlux resolutions {exp} do not match the ghosted grid {got}")
        g.set_physics(self.acceleration_coefficient, elliptic.magnetic_field, elliptic.poisson_coefficient)
        return g

    def semi_discrete_rhs(self, function, elliptic, basis, grids):
        """L(f) = Lx Jx + Lu Ju + Lv Jv (main/fluxes.py:59-105) on the GPU.

        Returns a new GPU tensor of the ghosted shape; its interior holds the
        RHS and its ghost entries are zero (the reference leaves values there
        that every caller discards).  Ghost elements of ``function`` are used
        as given.
        """
        import torch
        g = self._geom(grids, elliptic)
        if not isinstance(function, torch.Tensor) or function.device.type != "cuda":
            raise _lib.VlasovError("semi_discrete_rhs: function must be a GPU tensor (no CPU fallback)")
        if tuple(function.shape) != grids.shape or function.dtype != torch.float64 or not function.is_contiguous():
            raise ValueError("semi_discrete_rhs: function must be a contiguous float64 tensor of the ghosted shape")
        E = elliptic.electric_field
        if E is None or tuple(E.shape) != (grids.x.res_ghosts, grids.x.order):
            raise ValueError("semi_discrete_rhs: elliptic.electric_field is not set")
        out = torch.zeros_like(function)
        _lib.check(_lib.load().vl_rhs(g.handle, _lib.ptr(function), _lib.ptr(E.contiguous()),
                                      _lib.ptr(out), _lib.stream_ptr()), "vl_rhs")
        return out
