"""Pins of the oracle's spectral Poisson path (main/basis.py:204-228,
main/grid.py:45-88, main/elliptic.py:117-151) against closed forms and
independently computed Fourier integrals."""
import numpy as np
import pytest
from numpy.polynomial import legendre as npleg

import oracle


def grid(n, nx, k0=0.1):
    L = 2 * np.pi / k0
    return oracle.Grid1D(-L / 2, L / 2, nx, oracle.Basis1D(n), spectrum=True)


@pytest.mark.parametrize("n,nx,p", [(4, 6, 3), (8, 16, 5), (9, 8, 8), (6, 5, 0), (3, 7, 2)])
def test_nodal_dft_of_piecewise_polynomial_is_exact(n, nx, p):
    """f = (x - xbar_e)^p on each element is represented exactly by the nodal
    basis; its Fourier coefficients (1/L) int f e^{-ikx} dx are computed here by
    60-point Gauss-Legendre quadrature per element (no spherical Bessel)."""
    g = grid(n, nx)
    xs = g.arr[1:-1, :]
    f = (xs - g.midpoints[:, None]) ** p
    c = g.fourier_basis(f)
    gx, gw = npleg.leggauss(60)
    exact = np.zeros(len(g.wave_numbers), dtype=complex)
    for e in range(nx):
        x = g.midpoints[e] + gx * g.dx / 2
        fe = (x - g.midpoints[e]) ** p
        exact += (g.dx / 2) * np.exp(-1j * np.outer(g.wave_numbers, x)) @ (gw * fe)
    exact /= g.length
    scale = max(np.max(np.abs(exact)), 1e-300)
    assert np.max(np.abs(c - exact)) < 1e-12 * max(scale, (g.dx / 2) ** p)


def test_sum_fourier_of_band_limited_function():
    g = grid(8, 16)
    k1 = g.k1
    m = np.round(g.wave_numbers / k1).astype(int)
    c = np.zeros(len(m), dtype=complex)
    c[m == 1] = 0.5
    c[m == -1] = 0.5                       # cos(k1 x)
    c[m == 3] = -0.15j
    c[m == -3] = 0.15j                     # 0.3 sin(3 k1 x)
    x = g.arr[1:-1]
    want = np.cos(k1 * x) + 0.3 * np.sin(3 * k1 * x)
    assert np.max(np.abs(g.sum_fourier(c) - want)) < 1e-13


@pytest.mark.parametrize("coef", [-1.0, 1.0, -0.5])
def test_poisson_exact_solution(coef):
    """charge_density = sin(k1 x)/coef -> E = cos(k1 x)/k1, phi = -sin(k1 x)/k1^2
    (the method of main/main.py:276-301)."""
    g = grid(8, 16)
    k1 = g.k1
    x = g.arr[1:-1]
    e = oracle.Elliptic(coef)
    e.poisson2(np.sin(k1 * x) / coef, g)
    assert np.max(np.abs(e.electric_field[1:-1] - np.cos(k1 * x) / k1)) < 1e-9 / k1
    assert np.max(np.abs(e.potential + np.sin(k1 * x) / k1 ** 2)) < 1e-9 / k1 ** 2
    # periodic ghosts
    assert np.array_equal(e.electric_field[0], e.electric_field[-2])
    assert np.array_equal(e.electric_field[-1], e.electric_field[1])
    # energy: 0.5 sum w E^2 = 0.5 * J * int E^2 dx (weights omit dx/2)
    W = e.electric_energy(g)
    assert abs(W - 0.5 * g.J * (g.length / 2) / k1 ** 2) < 1e-9 * W


def test_poisson_drops_the_mean():
    """k = 0 contributes nothing (CuPy's NaN -> 0 at main/elliptic.py:124-128)."""
    g = grid(6, 10)
    e = oracle.Elliptic(-1.0)
    e.poisson2(np.full((10, 6), 3.7), g)
    assert np.max(np.abs(e.electric_field)) < 1e-13
    assert np.all(np.isfinite(e.potential))
    assert np.max(np.abs(e.potential)) < 1e-11


def test_mode_count_is_the_float_expression():
    """nyquist = 2.5*length // dx in float arithmetic (main/grid.py:46)."""
    assert grid(3, 48).nyquist_number == 119.0
    assert grid(3, 64).nyquist_number == 160.0
    assert grid(3, 16).nyquist_number == 40.0
    g = grid(3, 48)
    assert len(g.wave_numbers) == 2 * 119 - 1
    assert 0.0 in g.wave_numbers


@pytest.mark.parametrize("n", [4, 6, 8])
def test_central_flux_poisson_single_mode(n):
    """Central-flux DG Poisson (main/elliptic.py:24-115): charge density
    sin(k x)/pc gives E = cos(k x)/k; the potential has zero quadrature mean;
    without anti-aliasing E is minus the discrete gradient of the potential."""
    case = oracle.Case(n=n, nx=8, nu=2, nv=2, k0=0.1)
    g, b = case.grids.x, case.basis.b1
    ell = oracle.Elliptic(-1.0)
    ell.build_central_flux_operator(g, b)
    ell.invert()
    k = g.k1
    x = g.arr[1:-1]
    ell.poisson(np.sin(k * x) / -1.0, g, b)
    err = np.max(np.abs(ell.electric_field[1:-1] - np.cos(k * x) / k)) / (1 / k)
    # high-order convergence on 8 elements (measured 2.1e-3, 1.4e-5, 1.1e-8)
    assert err < {4: 5e-3, 6: 5e-5, 8: 5e-8}[n]
    assert abs(np.sum(ell.potential * g.quad_weights)) < 1e-12 * np.max(np.abs(ell.potential))
    ell.poisson(np.sin(k * x) / -1.0, g, b, anti_alias=False)
    grad = np.tensordot(ell.gradient_operator, ell.potential, axes=([0, 1], [0, 1]))
    np.testing.assert_allclose(ell.electric_field[1:-1], -g.J * grad, rtol=1e-13, atol=1e-15)
    # periodic ghosts
    np.testing.assert_array_equal(ell.electric_field[0], ell.electric_field[-2])
