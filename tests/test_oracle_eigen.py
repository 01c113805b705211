"""Pins of the oracle's Bernstein-eigenfunction initial condition
(main/grid.py:207-242)."""
import math

import numpy as np
import pytest
import scipy.special as sp

import oracle


def _series_jn(n, x, terms=40):
    """J_n(x) from its power series (independent of scipy)."""
    s = 0.0
    for k in range(terms):
        s += (-1) ** k * (x / 2) ** (2 * k + n) / (math.factorial(k) * math.factorial(k + n))
    return s


@pytest.mark.parametrize("x", [-3.7, -0.4, 0.0, 1e-3, 0.9, 2.5, 6.0])
def test_scipy_jv_matches_the_power_series(x):
    for n in range(0, 21):
        ref = _series_jn(n, x, terms=60)
        assert sp.jv(n, x) == pytest.approx(ref, rel=1e-12, abs=1e-300 + 1e-17)
        # negative orders: J_{-n} = (-1)^n J_n
        assert sp.jv(-n, x) == pytest.approx((-1) ** n * ref, rel=1e-12, abs=1e-17)


def test_eigenfunction_reduces_to_jacobi_anger_at_zero_frequency():
    """At om -> 0 (om = 1e-20, where n/(n - om) rounds to exactly 1 for n != 0
    and to 0 for n = 0; om = 0 itself divides 0 by 0), so (Jacobi-Anger)
    Upsilon = exp(-i beta sin phi) - J_0(beta) and the velocity mode is
    -i df_dv (1 - exp(i beta sin phi) J_0(beta)) / k1^2.  Small k1 om_pc keeps
    |beta| <= 2.5, where the 20-term truncation is far below 1e-15."""
    case = oracle.Case(n=4, nx=2, nu=3, nv=3, k0=0.1)
    g = case.grids
    vt, j, om_pc, dn = 1.0, 3, 1.0, 0.01
    got = oracle.eigenfunction(g, vt, j, 1e-20, om_pc, dn)
    u = np.tensordot(g.u.arr, np.ones_like(g.v.arr), axes=0)
    v = np.tensordot(np.ones_like(g.u.arr), g.v.arr, axes=0)
    r = np.sqrt(u * u + v * v)
    k1 = g.x.k1
    beta = -k1 * r * om_pc
    assert np.max(np.abs(beta)) < 2.5
    x = 0.5 * r * r
    f0 = x ** j * np.exp(-x) / (2 * np.pi * math.factorial(j))
    df = f0 * (j / (x + 1e-16) - 1.0)
    vel = -1j * df * (1.0 - np.exp(1j * beta * v / np.where(r > 0, r, 1.0)) * sp.j0(beta)) / k1 ** 2
    ref = 2 * dn * np.real(np.exp(1j * k1 * g.x.arr)[:, :, None, None, None, None] * vel[None, None])
    assert np.max(np.abs(got - ref)) <= 1e-13 * np.max(np.abs(ref))


def test_eigenfunction_is_a_single_x_harmonic():
    """The perturbation is Re(exp(i k1 x) V(u, v)) = a cos(k1 x) + b sin(k1 x)
    per velocity node: a and b from two x nodes predict every other node
    (ghost nodes included)."""
    case = oracle.Case(n=3, nx=4, nu=3, nv=3, k0=0.2)
    g = case.grids
    p = oracle.eigenfunction(g, 1.0, 2, 0.5 - 0.2j, 3.0, 1e-3)
    xs = g.x.arr.reshape(-1)
    P = p.reshape(xs.size, -1)
    k1 = g.x.k1
    M = np.array([[np.cos(k1 * xs[1]), np.sin(k1 * xs[1])], [np.cos(k1 * xs[4]), np.sin(k1 * xs[4])]])
    ab = np.linalg.solve(M, P[[1, 4]])
    pred = np.stack([np.cos(k1 * xs), np.sin(k1 * xs)], axis=1) @ ab
    assert np.max(np.abs(pred - P)) <= 1e-12 * np.max(np.abs(P))
    assert np.max(np.abs(P)) > 0
