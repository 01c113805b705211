"""Poisson solves and field energy on the GPU: drop-in for the reference's
``Elliptic`` (main/elliptic.py:9-151).

``poisson2`` is the spectral (high-order nodal DFT) solve of the hot path.
``poisson`` is the reference's stabilized central-flux DG solve: its operator
is assembled (as a product of block operators) and inverted on the host, then
applied on the device.
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
    @staticmethod
    def _face_jumps(res, n):
        """Face-jump matrix F, shape (2 res, res n): row (e, 0) is the jump
        phi_e(left node) - phi_{e-1}(right node) at element e's left face, row
        (e, 1) the jump phi_e(right node) - phi_{e+1}(left node) at its right
        face (periodic element neighbours)."""
        F = np.zeros((res, 2, res, n))
        e = np.arange(res)
        F[e, 0, e, 0] += 1.0
        F[e, 0, (e - 1) % res, n - 1] -= 1.0
        F[e, 1, e, n - 1] += 1.0
        F[e, 1, (e + 1) % res, 0] -= 1.0
        return F.reshape(2 * res, res * n)

    def build_central_flux_operator(self, grid, basis):
        """Assemble the central-flux DG Laplacian of main/elliptic.py:24-81 as a
        product of block operators (host, FP64).

        With F the face jumps and s = (+1/2, -1/2) the central-flux sign per
        face, the DG gradient is G = I (x) D + (I (x) Xi) diag(s) F, and the
        operator is L = (I (x) S^T) G + (I (x) M_f) (diag(s) F G + F): stiffness
        against the gradient plus the face terms of the gradient's central flux
        and of the potential's jump penalty.  The reference stores both
        operators input-major (row = perturbed node, column = response node),
        so ``gradient_operator`` holds G^T as (res, n, res, n), and
        ``central_flux_operator`` is [[L^T, w], [w^T, 0]] with the quadrature
        weights w fixing the gauge (zero mean potential); entries below 1e-15
        in magnitude are set to zero, as the reference does."""
        res, n = grid.res, grid.order
        K = res * n
        D = np.asarray(basis.der, dtype=np.float64)           # (n, n) nodal derivative
        Xi = np.asarray(basis.np_xi, dtype=np.float64)        # (n, 2) face lift
        St = np.asarray(basis.stf, dtype=np.float64)          # (n, n) stiffness (transposed)
        Mf = np.asarray(basis.face_mass, dtype=np.float64)    # (n, 2) face mass
        F = self._face_jumps(res, n)
        sF = np.tile([0.5, -0.5], res)[:, None] * F           # central numerical flux of a field

        def per_element(A, B):
            """(I (x) A) B for A of shape (n, m) and B of shape (res m, K)."""
            m = A.shape[1]
            return np.einsum("am,emk->eak", A, B.reshape(res, m, K)).reshape(K, K)

        G = per_element(Xi, sF)
        e = np.arange(res)
        G.reshape(res, n, res, n)[e, :, e, :] += D            # the element-local derivative blocks
        L = per_element(St, G) + per_element(Mf, sF @ G + F)
        self.gradient_operator = np.ascontiguousarray(G.T).reshape(res, n, res, n)
        w = np.asarray(grid.quad_weights, dtype=np.float64).reshape(K)
        op = np.zeros((K + 1, K + 1))
        op[:K, :K] = L.T
        op[:K, K] = w
        op[K, :K] = w
        op[np.abs(op) < 1.0e-15] = 0.0
        self.central_flux_operator = op
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
