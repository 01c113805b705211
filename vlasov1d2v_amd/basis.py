"""Lobatto nodal basis: drop-in for the reference's ``Basis1D``/``Basis3D``
(main/basis.py:87-289).

Host FP64 setup code (NumPy).  The operators the hot path uses (``up``,
``xi``, ``weights``) are handed to the HIP library through the geometry
descriptor; nothing here runs per time step.
"""
import math

import numpy as np


def legendre_table(x, nmax):
    """P_0..P_{nmax-1} and their derivatives at points x (three-term recurrence).

    Returns (P, dP) of shape (nmax, len(x)).
    """
    x = np.asarray(x, dtype=np.float64)
    P = np.zeros((max(nmax, 2), x.size))
    dP = np.zeros_like(P)
    P[0] = 1.0
    P[1] = x
    dP[1] = 1.0
    for k in range(2, nmax):
        P[k] = ((2 * k - 1) * x * P[k - 1] - (k - 1) * P[k - 2]) / k
        # P'_k = P'_{k-2} + (2k-1) P_{k-1}
        dP[k] = dP[k - 2] + (2 * k - 1) * P[k - 1]
    return P[:nmax], dP[:nmax]


def lobatto(n):
    """Legendre-Gauss-Lobatto nodes and weights on [-1, 1], ascending.

    The reference lists them as literal tables (main/basis.py:9-57); here they
    are computed by Newton iteration on x P_{n-1} - P_{n-2} (the classic
    Lobatto iteration of Hesthaven & Warburton) in long double, then rounded.
    """
    n = int(n)
    if n < 2:
        raise ValueError("Lobatto rules need n >= 2")
    N = n - 1
    x = -np.cos(np.pi * np.arange(n, dtype=np.longdouble) / N)

    def table(x):
        P = np.zeros((n, n), dtype=np.longdouble)
        P[:, 0] = 1
        P[:, 1] = x
        for k in range(2, n):
            P[:, k] = ((2 * k - 1) * x * P[:, k - 1] - (k - 1) * P[:, k - 2]) / k
        return P

    for _ in range(200):
        P = table(x)
        step = (x * P[:, N] - P[:, N - 1]) / (n * P[:, N])
        x = x - step
        if np.max(np.abs(step)) < 1e-30:
            break
    P = table(x)
    w = 2 / (N * n * P[:, N] ** 2)
    x = x.astype(np.float64)
    x[0], x[-1] = -1.0, 1.0
    if n % 2 == 1:
        x[n // 2] = 0.0
    x = 0.5 * (x - x[::-1])
    w = w.astype(np.float64)
    w = 0.5 * (w + w[::-1])
    return x, w


class Basis1D:
    """Nodal DG operators on n Lobatto nodes (main/basis.py:87-228)."""

    def __init__(self, order):
        self.order = int(order)
        if not 2 <= self.order <= 10:
            raise ValueError(f"order (nodes per axis) must be in [2, 10], got {order}")
        n = self.order
        self.nodes, self.weights = lobatto(n)
        self.eigenvalues = np.append(np.array([(2.0 * s + 1) / 2.0 for s in range(n - 1)]),
                                     (n - 1) / 2.0)                      # main/basis.py:126-132
        P, dP = legendre_table(self.nodes, n)
        self._P, self._dP = P, dP
        # V^{-1}[s, j] = w_j lambda_s P_s(x_j)                              main/basis.py:134-137
        self.vandermonde_inverse = self.weights[None, :] * self.eigenvalues[:, None] * P
        # exact mass matrix and inverse (rank-one corrections)            main/basis.py:139-174
        v = self.weights * P[n - 1]
        a = -n * (n - 1) / (2 * (2 * n - 1))
        self.mass = np.diag(self.weights) + a * np.outer(v, v)
        pn = P[n - 1]
        self.invm = np.diag(1.0 / self.weights) + (n / 2) * np.outer(pn, pn)
        self.face_mass = np.eye(n)[:, np.array([0, -1])]
        # weak-form volume operator                                       main/basis.py:176-188
        # up[i, j] = w_j sum_s (2s+1)/2 P_s(x_i) P_s'(x_j), |up| < 1e-10 -> 0
        c = (2.0 * np.arange(n) + 1.0) / 2.0
        up = (P.T * c[None, :]) @ dP * self.weights[None, :]
        up[np.abs(up) < 1.0e-10] = 0
        self.up = up
        self.xi = self.invm[:, np.array([0, -1])]                         # main/basis.py:110
        self.np_up, self.np_xi = self.up, self.xi
        # strong-form derivative / advection operators (off the hot path)  main/basis.py:150-202
        adv = self.weights[:, None] * self.weights[None, :] * (
            (P.T * self.eigenvalues[None, :]) @ dP)
        adv[np.abs(adv) < 1.0e-15] = 0
        self.adv = adv
        self.stf = adv.T
        der = self.weights[None, :] * ((dP.T * self.eigenvalues[None, :]) @ P)
        der[np.abs(der) < 1.0e-15] = 0
        self.der = der

    def get_nodes(self):
        return self.nodes

    def get_weights(self):
        return self.weights

    def fourier_transform_array(self, midpoints, J, wave_numbers):
        """T[e, j, m]: exact Fourier transform of nodal basis function j of element e
        at wave number k_m (main/basis.py:204-228)."""
        from scipy.special import spherical_jn
        k = np.asarray(wave_numbers, dtype=np.float64)
        signs = np.sign(k)
        signs[k == 0] = 1.0
        p_tilde = np.array([(signs ** s) * np.exp(-1j * np.pi / 2.0 * s) *
                            spherical_jn(s, np.abs(k) / J) for s in range(self.order)])
        ell_tilde = self.vandermonde_inverse.T @ p_tilde                  # (n, M)
        phase = np.exp(-1j * np.multiply.outer(np.asarray(midpoints), k))  # (res, M)
        return np.ascontiguousarray(np.transpose(phase[:, :, None] * ell_tilde.T[None], (0, 2, 1)))


class Basis3D:
    """Three 1-D bases for x, u, v (main/basis.py:283-289)."""

    def __init__(self, orders):
        self.orders = np.asarray(orders)
        self.b1 = Basis1D(self.orders[0])
        self.b2 = Basis1D(self.orders[1])
        self.b3 = Basis1D(self.orders[2])


def ring_factor(ring_j, vt):
    """1 / (2^j 2 pi vt^2 j!) (main/grid.py:253)."""
    return 1.0 / ((2.0 ** ring_j) * (2.0 * np.pi) * (vt ** 2.0) * math.factorial(ring_j))
