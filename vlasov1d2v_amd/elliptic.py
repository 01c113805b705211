"""Spectral Poisson solve and field energy on the GPU: drop-in for the
reference's ``Elliptic`` (main/elliptic.py:9-151, spectral path).

The central-flux DG Poisson operator of the reference
(main/elliptic.py:24-115) is not part of this path.
"""
import numpy as np

from . import _lib


class Elliptic:
    """Holds E (NX+2, n), phi (NX, n) and B; solves with the high-order nodal DFT."""

    def __init__(self, poisson_coefficient):
        self.poisson_coefficient = float(poisson_coefficient)
        self.potential = None
        self.electric_field = None
        self.magnetic_field = None

    def set_magnetic_field(self, magnetic_field):          # main/elliptic.py:147-148
        self.magnetic_field = float(magnetic_field)

    def _geom(self, grid):
        g = grid.geometry()
        acc = g.acc
        b = self.magnetic_field if self.magnetic_field is not None else g.b_field
        g.set_physics(acc, b, self.poisson_coefficient)
        return g

    def poisson2(self, charge_density, grid, basis=None):
        """E = Re sum_m -rho_m/(i k_m) e^{i k_m x}, k=0 removed; periodic ghosts
        (main/elliptic.py:117-145).  ``charge_density`` is an (NX, n) GPU
        tensor (the reference's caller passes mean(rho) - rho)."""
        import torch
        g = self._geom(grid)
        cd = charge_density
        if not isinstance(cd, torch.Tensor) or cd.device.type != "cuda":
            raise _lib.VlasovError("poisson2: charge_density must be a GPU tensor (no CPU fallback)")
        cd = cd.to(torch.float64).contiguous()
        if tuple(cd.shape) != (grid.res, grid.order):
            raise ValueError(f"poisson2: charge_density shape {tuple(cd.shape)} != {(grid.res, grid.order)}")
        dev = cd.device
        if self.electric_field is None or self.electric_field.device != dev:
            self.electric_field = torch.zeros((grid.res_ghosts, grid.order), dtype=torch.float64, device=dev)
        if self.potential is None or self.potential.device != dev:
            self.potential = torch.zeros((grid.res, grid.order), dtype=torch.float64, device=dev)
        ws = g.workspace(dev)
        _lib.check(_lib.load().vl_poisson(g.handle, _lib.ptr(cd), _lib.ptr(self.electric_field),
                                          _lib.ptr(self.potential), _lib.ptr(ws), _lib.stream_ptr()),
                   "vl_poisson")

    def electric_energy(self, grid):
        """0.5 sum_{e,j} w_j E[e,j]^2 over the interior (main/elliptic.py:150-151); 0-d GPU tensor."""
        import torch
        g = grid.geometry()
        out = torch.empty((), dtype=torch.float64, device=self.electric_field.device)
        _lib.check(_lib.load().vl_field_energy(g.handle, _lib.ptr(self.electric_field), _lib.ptr(out),
                                               _lib.stream_ptr()), "vl_field_energy")
        return out

    def max_abs_acc_e(self, grid, acc):
        """max |acc E| over the ghosted field (main/timestep.py:237) as a 0-d GPU tensor."""
        import torch
        g = grid.geometry()
        g.set_physics(acc, g.b_field, g.poisson_coef)
        out = torch.empty((), dtype=torch.float64, device=self.electric_field.device)
        _lib.check(_lib.load().vl_max_abs_acc_e(g.handle, _lib.ptr(self.electric_field), _lib.ptr(out),
                                                _lib.stream_ptr()), "vl_max_abs_acc_e")
        return out
