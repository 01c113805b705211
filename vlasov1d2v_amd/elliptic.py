"""Poisson solves and field energy on the GPU: drop-in for the reference's
``Elliptic`` (main/elliptic.py:9-151).

``poisson2`` is the spectral (high-order nodal DFT) solve of the hot path.
``poisson`` is the reference's stabilized central-flux DG solve: its operator
is assembled and inverted on the host exactly as the reference does
(``build_central_flux_operator`` / ``invert``), then applied on the device.
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

    # -- stabilized central-flux DG Poisson (main/elliptic.py:24-115) ---------
    def build_central_flux_operator(self, grid, basis):
        """Assemble the central-flux DG Laplacian column by column from nodal
        indicators, with the gauge row/column (zero quadrature mean of the
        potential), as main/elliptic.py:24-81 does (host, FP64)."""
        res, order = grid.res, grid.order
        indicator = np.zeros((res, order))
        face_diff0 = np.zeros((res, 2))
        face_diff1 = np.zeros((res, 2))
        num_flux = np.zeros_like(face_diff0)
        grad_num_flux = np.zeros_like(face_diff1)
        op = np.zeros((res, order, res, order))
        self.gradient_operator = np.zeros_like(op)
        for i in range(res):
            for j in range(order):
                indicator[i, j] = 1.0
                face_diff0[:, 0] = indicator[:, 0] - np.roll(indicator[:, -1], 1)
                face_diff0[:, 1] = indicator[:, -1] - np.roll(indicator[:, 0], -1)
                num_flux[:, 0] = 0.5 * face_diff0[:, 0]
                num_flux[:, 1] = -0.5 * face_diff0[:, 1]
                grad = (np.tensordot(basis.der, indicator, axes=([1], [1])) +
                        np.tensordot(basis.np_xi, num_flux, axes=([1], [1]))).T
                face_diff1[:, 0] = grad[:, 0] - np.roll(grad[:, -1], 1)
                face_diff1[:, 1] = grad[:, -1] - np.roll(grad[:, 0], -1)
                grad_num_flux[:, 0] = 0.5 * face_diff1[:, 0]
                grad_num_flux[:, 1] = -0.5 * face_diff1[:, 1]
                op[i, j] = (np.tensordot(basis.stf, grad, axes=([1], [1])) +
                            np.tensordot(basis.face_mass, grad_num_flux + face_diff0, axes=([1], [1]))).T
                self.gradient_operator[i, j] = grad
                indicator[i, j] = 0
        qw = np.asarray(grid.quad_weights, dtype=np.float64)
        op0 = np.hstack([op.reshape(res * order, res * order), qw.reshape(res * order, 1)])
        self.central_flux_operator = np.vstack([op0, np.append(qw.flatten(), 0)])
        self.central_flux_operator[np.abs(self.central_flux_operator) < 1.0e-15] = 0
        self._cf_geom = None

    def invert(self):
        """Dense inverse of the central-flux operator (main/elliptic.py:83-84), host FP64."""
        self.inv_op = np.linalg.inv(self.central_flux_operator)
        self._cf_geom = None

    def poisson(self, charge_density, grid, basis, anti_alias=True):
        """phi and E from the central-flux operator (main/elliptic.py:86-115) on
        the GPU; ``charge_density`` is an (NX, n) GPU tensor."""
        import torch
        if getattr(self, "inv_op", None) is None:
            raise RuntimeError("poisson: call build_central_flux_operator and invert first")
        g = self._geom(grid)
        if getattr(self, "_cf_geom", None) is not g:
            K = grid.res * grid.order
            _lib.check(_lib.load().vl_set_central_flux(
                g.handle, _lib.ptr(np.ascontiguousarray(self.inv_op, dtype=np.float64)),
                _lib.ptr(np.ascontiguousarray(self.gradient_operator.reshape(K, K), dtype=np.float64)),
                _lib.ptr(np.ascontiguousarray(basis.mass, dtype=np.float64))), "vl_set_central_flux")
            self._cf_geom = g
        cd = charge_density
        if not isinstance(cd, torch.Tensor) or cd.device.type != "cuda":
            raise _lib.VlasovError("poisson: charge_density must be a GPU tensor (no CPU fallback)")
        cd = cd.to(torch.float64).contiguous()
        dev = cd.device
        if self.electric_field is None or self.electric_field.device != dev:
            self.electric_field = torch.zeros((grid.res_ghosts, grid.order), dtype=torch.float64, device=dev)
        if self.potential is None or self.potential.device != dev:
            self.potential = torch.zeros((grid.res, grid.order), dtype=torch.float64, device=dev)
        _lib.check(_lib.load().vl_poisson_central(g.handle, _lib.ptr(cd), _lib.ptr(self.electric_field),
                                                  _lib.ptr(self.potential), 1 if anti_alias else 0,
                                                  _lib.stream_ptr()), "vl_poisson_central")

    def max_abs_acc_e(self, grid, acc):
        """max |acc E| over the ghosted field (main/timestep.py:237) as a 0-d GPU tensor."""
        import torch
        g = grid.geometry()
        g.set_physics(acc, g.b_field, g.poisson_coef)
        out = torch.empty((), dtype=torch.float64, device=self.electric_field.device)
        _lib.check(_lib.load().vl_max_abs_acc_e(g.handle, _lib.ptr(self.electric_field), _lib.ptr(out),
                                                _lib.stream_ptr()), "vl_max_abs_acc_e")
        return out
