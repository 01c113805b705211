"""Pins of the oracle's Lobatto basis (main/basis.py:9-188) against
mathematics computed independently: Gauss-Legendre quadrature of Lagrange
polynomials and barycentric differentiation."""
import numpy as np
import pytest
from numpy.polynomial import legendre as npleg

import oracle

NS = list(range(2, 11))


def lagrange_matrix(nodes, pts):
    """L[i, p] = ell_i(pts[p]) for the Lagrange basis on `nodes` (direct products)."""
    n = len(nodes)
    L = np.ones((n, len(pts)))
    for i in range(n):
        for m in range(n):
            if m != i:
                L[i] *= (pts - nodes[m]) / (nodes[i] - nodes[m])
    return L


def lagrange_deriv_matrix(nodes, pts):
    """dL[i, p] = ell_i'(pts[p]) by the product rule."""
    n = len(nodes)
    dL = np.zeros((n, len(pts)))
    for i in range(n):
        for k in range(n):
            if k == i:
                continue
            term = np.ones(len(pts)) / (nodes[i] - nodes[k])
            for m in range(n):
                if m != i and m != k:
                    term = term * (pts - nodes[m]) / (nodes[i] - nodes[m])
            dL[i] += term
    return dL


@pytest.mark.parametrize("n", NS)
def test_lgl_nodes_are_lobatto_points(n):
    x, w = oracle.lgl(n)
    # interior nodes are the roots of P'_{n-1}
    c = np.zeros(n)
    c[-1] = 1
    dP = npleg.legval(x[1:-1], npleg.legder(c))
    assert np.all(np.abs(dP) < 1e-12 * n * n)
    assert x[0] == -1.0 and x[-1] == 1.0
    assert np.all(np.diff(x) > 0)
    # weights: 2 / (n (n-1) P_{n-1}(x)^2)
    P = npleg.legval(x, c)
    assert np.allclose(w, 2.0 / (n * (n - 1) * P ** 2), rtol=1e-14, atol=0)
    # exactness to degree 2n-3
    for k in range(2 * n - 2):
        exact = 0.0 if k % 2 else 2.0 / (k + 1)
        assert abs(np.sum(w * x ** k) - exact) < 1e-14


@pytest.mark.parametrize("n", NS)
def test_inverse_mass_matrix(n):
    b = oracle.Basis1D(n)
    g, gw = npleg.leggauss(n + 3)
    L = lagrange_matrix(b.nodes, g)
    M = (L * gw) @ L.T                       # exact mass matrix
    assert np.max(np.abs(b.invm @ M - np.eye(n))) < 1e-12


@pytest.mark.parametrize("n", NS)
def test_volume_operator_is_minv_stiffness(n):
    """up = M^-1 S with S_ij = int ell_i' ell_j, both by exact Gauss quadrature."""
    b = oracle.Basis1D(n)
    g, gw = npleg.leggauss(n + 3)
    L = lagrange_matrix(b.nodes, g)
    dL = lagrange_deriv_matrix(b.nodes, g)
    M = (L * gw) @ L.T
    S = (dL * gw) @ L.T
    up = np.linalg.solve(M, S)
    assert np.max(np.abs(b.up - up)) < 1e-10 * max(1.0, np.max(np.abs(up)))
    # face lift columns: M^-1 e_0 and M^-1 e_{n-1}
    e = np.eye(n)
    assert np.allclose(b.xi[:, 0], np.linalg.solve(M, e[:, 0]), rtol=0, atol=1e-11)
    assert np.allclose(b.xi[:, 1], np.linalg.solve(M, e[:, -1]), rtol=0, atol=1e-11)


@pytest.mark.parametrize("n", NS)
def test_free_stream_and_conservation_identities(n):
    b = oracle.Basis1D(n)
    scale = np.max(np.abs(b.up))
    # up . 1 = xi[:,1] - xi[:,0]  (constants are transported exactly)
    assert np.max(np.abs(b.up @ np.ones(n) - (b.xi[:, 1] - b.xi[:, 0]))) < 1e-11 * scale
    # w^T up = 0 (the volume term integrates to zero)
    assert np.max(np.abs(b.weights @ b.up)) < 1e-12 * scale
    # cleaning threshold of main/basis.py:186: no surviving entry below 1e-10
    nz = b.up[b.up != 0]
    assert np.all(np.abs(nz) >= 1e-10)


@pytest.mark.parametrize("n", NS)
def test_vandermonde_inverse(n):
    b = oracle.Basis1D(n)
    c = np.eye(n)
    V = np.array([npleg.legval(b.nodes, c[s]) for s in range(n)]).T   # V[j, s] = P_s(x_j)
    # the reference's V^-1 uses the Lobatto eigenvalue (n-1)/2 in the last row,
    # which makes it the exact inverse of V
    assert np.max(np.abs(b.vandermonde_inverse @ V - np.eye(n))) < 1e-12


@pytest.mark.parametrize("n", NS)
def test_mass_advection_and_derivative_matrices(n):
    """mass = int l_i l_j (Gauss-Legendre quadrature of the Lagrange basis),
    der[i, j] = l_j'(x_i) (product rule), adv = der^T W (main/basis.py:139-163,190-202)."""
    b = oracle.Basis1D(n)
    xg, wg = npleg.leggauss(n + 2)
    L = lagrange_matrix(b.nodes, xg)
    M = (L * wg) @ L.T
    assert np.max(np.abs(b.mass - M)) < 1e-14
    D = lagrange_deriv_matrix(b.nodes, b.nodes).T        # D[i, j] = l_j'(x_i)
    D[np.abs(D) < 1e-15] = 0
    assert np.max(np.abs(b.der - D)) < 1e-12 * max(1.0, np.max(np.abs(D)))
    assert np.max(np.abs(b.adv - (D.T * b.weights[None, :]))) < 1e-12 * max(1.0, np.max(np.abs(D)))
    assert np.array_equal(b.stf, b.adv.T)
    assert np.array_equal(b.face_mass, np.eye(n)[:, [0, -1]])


def _literals():
    import json
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "lgl_literals.json")
    with open(path) as fh:
        return json.load(fh)


@pytest.mark.parametrize("n", [7, 8, 9, 10])
def test_lgl_matches_reference_literals(n):
    """The oracle's Newton-iterated Lobatto nodes and weights equal the decimal
    tables of main/basis.py:9-57 (fixture tests/golden/lgl_literals.json,
    written by make_lgl_fixture.py from the reference's text) to a few ulp."""
    lit = _literals()
    nodes = np.array([float(t) for t in lit["nodes"][str(n)]])
    weights = np.array([float(t) for t in lit["weights"][str(n)]])
    x, w = oracle.lgl(n)
    assert np.max(np.abs(x - nodes)) <= 4.5e-16
    assert np.max(np.abs(w - weights)) <= 2e-16


CLOSED_FORMS = {
    # Lobatto nodes/weights written as closed forms in main/basis.py:9-57
    # (textbook values: Abramowitz & Stegun 25.4.32)
    2: ([-1.0, 1.0], [1.0, 1.0]),
    3: ([-1.0, 0.0, 1.0], [1 / 3, 4 / 3, 1 / 3]),
    4: ([-1.0, -np.sqrt(1 / 5), np.sqrt(1 / 5), 1.0], [1 / 6, 5 / 6, 5 / 6, 1 / 6]),
    5: ([-1.0, -np.sqrt(3 / 7), 0.0, np.sqrt(3 / 7), 1.0], [1 / 10, 49 / 90, 32 / 45, 49 / 90, 1 / 10]),
    6: ([-1.0, -np.sqrt(1 / 3 + 2 * np.sqrt(7) / 21), -np.sqrt(1 / 3 - 2 * np.sqrt(7) / 21),
         np.sqrt(1 / 3 - 2 * np.sqrt(7) / 21), np.sqrt(1 / 3 + 2 * np.sqrt(7) / 21), 1.0],
        [1 / 15, (14 - np.sqrt(7)) / 30, (14 + np.sqrt(7)) / 30, (14 + np.sqrt(7)) / 30,
         (14 - np.sqrt(7)) / 30, 1 / 15]),
}


@pytest.mark.parametrize("n", sorted(CLOSED_FORMS))
def test_lgl_matches_closed_forms(n):
    x, w = oracle.lgl(n)
    assert np.max(np.abs(x - np.array(CLOSED_FORMS[n][0]))) <= 2.3e-16
    assert np.max(np.abs(w - np.array(CLOSED_FORMS[n][1]))) <= 2.3e-16
