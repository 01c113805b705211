"""Pins of the oracle's DG right-hand side (main/fluxes.py:59-207).

The brute-force check assembles the 1-D upwind DG weak form
  d/dt f_e = J M^-1 [ S (a f_e) + e_0 Fhat_L - e_{n-1} Fhat_R ]
from Lagrange polynomials and Gauss-Legendre quadrature, with the upwind
numerical flux written out face by face, and applies it along x, u and v
with the reference's speeds a_x = u, a_u = acc (E + B v), a_v = -acc B u.
It shares no code with the oracle's tensordot/roll formulation.
"""
import numpy as np
import pytest
from numpy.polynomial import legendre as npleg

import oracle
from helpers import interior, rel_max
from test_oracle_basis import lagrange_deriv_matrix, lagrange_matrix


def dg_matrices(n):
    x, _ = oracle.lgl(n)
    g, gw = npleg.leggauss(n + 3)
    L = lagrange_matrix(x, g)
    dL = lagrange_deriv_matrix(x, g)
    M = (L * gw) @ L.T
    S = (dL * gw) @ L.T
    Minv = np.linalg.inv(M)
    return Minv @ S, Minv[:, 0], Minv[:, -1]


def line_rhs(f, a, J, ops):
    """Upwind DG along the last two axes (element, node) of f, with speed a
    broadcastable to f; element 0 and -1 are ghosts (inputs only)."""
    D, l0, l1 = ops
    F = f * a
    out = np.zeros_like(f)
    ne = f.shape[-2]
    for e in range(1, ne - 1):
        Fe = F[..., e, :]
        ae = a[..., e, :] if np.ndim(a) == np.ndim(f) else a
        pos = ae >= 0
        # upwind flux at the left and right faces of element e
        FL = np.where(pos[..., 0], F[..., e - 1, -1], F[..., e, 0]) if np.ndim(pos) else None
        FR = np.where(pos[..., -1], F[..., e, -1], F[..., e + 1, 0]) if np.ndim(pos) else None
        vol = np.einsum("ij,...j->...i", D, Fe)
        out[..., e, :] = J * (vol + FL[..., None] * l0 - FR[..., None] * l1)
    return out


def brute_rhs(case, f, E):
    n = case.n
    g = case.grids
    ops = dg_matrices(n)
    acc, B = case.acc, case.elliptic.magnetic_field
    u = g.u.arr            # (nu+2, n)
    v = g.v.arr
    # x direction: move (ex, ix) to the last two axes
    fx = np.moveaxis(f, (0, 1), (4, 5))          # (eu,iu,ev,iv,ex,ix)
    ax = np.broadcast_to(u[:, :, None, None, None, None], fx.shape)
    Lx = np.moveaxis(line_rhs(fx, ax, g.x.J, ops), (4, 5), (0, 1))
    # u direction
    fu = np.moveaxis(f, (2, 3), (4, 5))          # (ex,ix,ev,iv,eu,iu)
    au = acc * (E[:, :, None, None, None, None] + B * v[None, None, :, :, None, None])
    au = np.broadcast_to(au, fu.shape)
    Lu = np.moveaxis(line_rhs(fu, au, g.u.J, ops), (4, 5), (2, 3))
    # v direction
    av = acc * (-B * u)
    av = np.broadcast_to(av[None, None, :, :, None, None], f.shape)
    Lv = line_rhs(f, av, g.v.J, ops)
    return Lx, Lu, Lv


@pytest.mark.parametrize("n,nx,nu,nv,B,acc", [(2, 3, 3, 3, 0.1, -1.0), (3, 2, 3, 4, 0.3, -1.0),
                                              (4, 3, 2, 2, -0.2, 1.0), (5, 2, 3, 2, 0.1, -1.0),
                                              (7, 2, 2, 2, 0.15, -1.0)])
def test_rhs_against_independent_dg_assembly(n, nx, nu, nv, B, acc):
    case = oracle.Case(n=n, nx=nx, nu=nu, nv=nv, B=B, acc=acc)
    rng = np.random.default_rng(n)
    f = rng.standard_normal(case.dist.arr.shape)
    E = rng.standard_normal((nx + 2, n))
    case.elliptic.electric_field = E
    ref = case.rhs(f)
    Lx, Lu, Lv = brute_rhs(case, f, E)
    assert rel_max(interior(ref), interior(Lx + Lu + Lv)) < 1e-11


def test_rhs_of_constant_vanishes():
    """f = const, E = 0: transport and rotation of a constant give zero."""
    case = oracle.Case(n=6, nx=3, nu=4, nv=4)
    f = np.full(case.dist.arr.shape, 0.7)
    case.elliptic.electric_field = np.zeros_like(case.elliptic.electric_field)
    r = case.rhs(f)
    assert np.max(np.abs(interior(r))) < 1e-12


def test_x_flux_conserves_mass_on_periodic_x():
    case = oracle.Case(n=5, nx=4, nu=3, nv=3)
    rng = np.random.default_rng(1)
    f = rng.standard_normal(case.dist.arr.shape)
    f[0] = f[-2]
    f[-1] = f[1]
    g = case.grids
    Lx = case.flux.x_flux(f, case.basis.b1, g.u) * g.x.J
    w = case.basis.b1.weights
    tot = np.einsum("xiujvk,i->ujvk", interior(Lx), w)
    assert np.max(np.abs(tot)) < 1e-12 * np.max(np.abs(Lx))


@pytest.mark.parametrize("n", [4, 6, 8])
def test_x_derivative_converges_spectrally(n):
    """f = g(u,v) sin(k x): L_x -> -u k g cos(k x); error falls fast with n."""
    errs = []
    for nn in (n, n + 2):
        case = oracle.Case(n=nn, nx=4, nu=2, nv=2)
        g = case.grids
        k = g.x.k1
        x = g.x.arr[:, :, None, None, None, None]
        u = g.u.arr[None, None, :, :, None, None]
        G = np.exp(-0.5 * (g.u.arr[None, None, :, :, None, None] ** 2 + g.v.arr[None, None, None, None, :, :] ** 2))
        f = G * np.sin(k * x)
        Lx = case.flux.x_flux(f, case.basis.b1, g.u) * g.x.J
        exact = -u * k * G * np.cos(k * x)
        errs.append(np.max(np.abs(interior(Lx - exact))) / np.max(np.abs(interior(exact))))
    assert errs[1] < errs[0] / 10
    assert errs[0] < 5e-2
