"""NumPy FP64 restatement of the reference's DG Vlasov-Poisson RK3 path.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``): imported by tests, by
``__graft_entry__.smoke()`` and by the CPU-baseline leg of ``bench.py``; never
by the product package.

Each class or function mirrors the array-op sequence of the reference
(crewsdw/Vlasov1D2V, CuPy) with NumPy in place of CuPy, citing the
``file:line`` it follows.  Deviations, all forced by running on NumPy or by
the workload definition:

* The k=0 Fourier mode of the Poisson solve is set to 0 explicitly.  CuPy's
  complex divide by 0j yields NaN which ``nan_to_num`` maps to 0; NumPy would
  give inf -> 1.8e308 (main/elliptic.py:124-128).
* ``math.factorial`` replaces the removed ``np.math.factorial``
  (main/grid.py:253).
* The time step is fixed unless ``Stepper.main_loop`` is given a Courant
  number; then it follows the CFL rule of main/timestep.py:88-89,184,231-242
  (``get_max_speeds``/``adapt_dt``).
* Lobatto nodes/weights are computed by the Newton iteration of
  Hesthaven & Warburton (``JacobiGL``-style, von Winckel's ``lglnodes``) in
  80-bit long double and rounded to FP64; they equal the literal tables of
  main/basis.py:9-57 to within a few ulp.
* The workloads' initial condition is the ring distribution of
  main/grid.py:244-268 times a density modulation
  ``1 + a1 cos(k1 x) + sum_m a_m cos(m k1 x + phi_m)`` (the workload's
  definition); the reference's Bernstein eigenfunction (main/grid.py:207-242)
  is ``eigenfunction`` / ``Distribution.initialize_eigen``.

Pinning (tests/test_oracle_*.py):

* ``lgl``              -- roots of (1-x^2)P'_{n-1}, weight closed form, and
                          exactness of the quadrature to degree 2n-3.
* ``Basis1D``          -- M^-1 against a mass matrix built by Gauss-Legendre
                          quadrature of Lagrange polynomials; ``up`` against
                          M^-1 D^T W with D from barycentric differentiation;
                          free-stream and conservation identities.
* ``fourier_transform_array`` / ``Grid1D.fourier_basis`` / ``sum_fourier``
                       -- analytic Fourier coefficients of band-limited and
                          Jacobi-elliptic test functions (the method of
                          main/high_order_dft.py:41-95).
* ``Elliptic.poisson2`` -- exact solution rho = sin(k x) -> E = cos(k x)/k
                          (main/main.py:276-301 method) and k=0 removal.
* ``moment_zero``      -- closed-form ring integral (incomplete gamma tail).
* ``DGFlux.semi_discrete_rhs`` -- an independently assembled dense upwind DG
                          operator (Lagrange-quadrature weak form) applied to
                          unit vectors on tiny grids, plus free-stream,
                          conservation and spectral-convergence checks.
* ``lgl``              -- also the decimal tables of main/basis.py:9-57
                          (tests/golden/lgl_literals.json) and their closed
                          forms, to a few ulp.
* ``Stepper``          -- SSP-RK3 amplification polynomial on a linear
                          problem, Shu-Osher table; on the coupled
                          Vlasov-Poisson system third-order temporal
                          self-convergence (a stepper that reuses the f^n
                          field for every stage measures first order).
* ``electric_energy``  -- closed form for E = A cos(kx).
* ``eigenfunction``    -- the Jacobi-Anger sum exp(-i b sin t) = sum_n J_n(b)
                          exp(-i n t) at om = 0, and independent Bessel values
                          (series) for the scipy ``jv`` it calls.
* ``Basis1D.mass_matrix`` / ``advection_matrix`` / ``derivative_matrix`` --
                          Gauss-Legendre quadrature of Lagrange polynomials,
                          barycentric differentiation, adv = D^T W.
* ``Elliptic.poisson`` (central flux) -- the exact solution of a single
                          Fourier mode, the gauge (zero quadrature mean of the
                          potential) and E = -d(phi)/dx of the discrete potential.
* ``Stepper.ssp_rk_update`` -- the amplification polynomial of the linear
                          SSP-RK3 (1 + z + z^2/2 + z^3/6).
* ``get_max_speeds`` / ``adapt_dt`` -- the rule's value at the 2.1M-DOF
                          set-up (0.00944, computed independently from the
                          grid and field) and the zero-field closed form; the
                          loop order (adapt after the step, time += new dt)
                          follows the reference text.
* ``Stepper.main_loop`` energy series -- entry k >= 1 equals the energy of
                          the field solved from f2 recomputed independently
                          with the pinned flux and Poisson solve, and differs
                          from the energy of a re-solve of f^{n+1}.
The frozen u/v ghost planes follow the reference text (main/timestep.py:
148-151,187); no reference-held vectors exist to pin them further.
"""
import math

import numpy as np
import scipy.special as sp

# Shu-Osher SSP-RK coefficient rows [a_fn, a_prev, a_dtL]
# (main/timestep.py:35-39).
NONLINEAR_SSP_RK = {
    2: [[1 / 2, 1 / 2, 1 / 2]],
    3: [[3 / 4, 1 / 4, 1 / 4],
        [1 / 3, 2 / 3, 2 / 3]],
}

# Cockburn-Shu Courant numbers, RK3 row (main/timestep.py:27).
COURANT_RK3 = [1.256, 0.409, 0.209, 0.130, 0.089, 0.066, 0.051, 0.040, 0.033]


def lgl(n):
    """Legendre-Gauss-Lobatto nodes and weights on [-1, 1] (ascending).

    Values of main/basis.py:9-57 (which lists them as literals); computed here
    by Newton iteration on x P_N - P_{N-1} in long double.
    """
    if n == 1:
        return np.array([0.0]), np.array([2.0])
    N = n - 1
    x = -np.cos(np.pi * np.arange(n, dtype=np.longdouble) / N)
    x = x.astype(np.longdouble)
    for _ in range(100):
        P = np.zeros((n, n), dtype=np.longdouble)
        P[:, 0] = 1
        P[:, 1] = x
        for k in range(2, n):
            P[:, k] = ((2 * k - 1) * x * P[:, k - 1] - (k - 1) * P[:, k - 2]) / k
        dx = (x * P[:, N] - P[:, N - 1]) / (n * P[:, N])
        x = x - dx
        if np.max(np.abs(dx)) < 1e-30:
            break
    P = np.zeros((n, n), dtype=np.longdouble)
    P[:, 0] = 1
    P[:, 1] = x
    for k in range(2, n):
        P[:, k] = ((2 * k - 1) * x * P[:, k - 1] - (k - 1) * P[:, k - 2]) / k
    w = 2 / (N * n * P[:, N] ** 2)
    x = x.astype(np.float64)
    x[0], x[-1] = -1.0, 1.0
    if n % 2 == 1:
        x[n // 2] = 0.0
    # enforce exact antisymmetry as the literal tables have it
    xs = 0.5 * (x - x[::-1])
    w = w.astype(np.float64)
    ws = 0.5 * (w + w[::-1])
    return xs, ws


class Basis1D:
    """main/basis.py:87-228 (Basis1D)."""

    def __init__(self, order):
        self.order = int(order)
        self.nodes, self.weights = lgl(self.order)
        self.eigenvalues = self.set_eigenvalues()
        self.vandermonde_inverse = self.set_vandermonde_inverse()
        self.invm = self.inv_mass_matrix()
        self.up = self.internal_flux()
        self.xi = self.invm[:, np.array([0, -1])]          # main/basis.py:110
        # central-flux Poisson operators (main/basis.py:98-116)
        self.mass = self.mass_matrix()
        self.face_mass = np.eye(self.order)[:, np.array([0, -1])]
        self.adv = self.advection_matrix()
        self.stf = self.adv.T
        self.np_xi = self.xi
        self.der = self.derivative_matrix()

    def mass_matrix(self):                                  # main/basis.py:139-148
        approx_mass = np.diag(self.weights)
        p = sp.legendre(self.order - 1)
        v = np.multiply(self.weights, p(self.nodes))
        a = -self.order * (self.order - 1) / (2 * (2 * self.order - 1))
        return approx_mass + a * np.outer(v, v)

    def advection_matrix(self):                             # main/basis.py:150-163
        n = self.order
        adv = np.zeros((n, n))
        for i in range(n):
            for j in range(n):
                adv[i, j] = self.weights[i] * self.weights[j] * sum(
                    self.eigenvalues[s] * sp.legendre(s)(self.nodes[i]) *
                    sp.legendre(s).deriv()(self.nodes[j]) for s in range(n))
        adv[np.abs(adv) < 1.0e-15] = 0
        return adv

    def derivative_matrix(self):                            # main/basis.py:190-202
        n = self.order
        der = np.zeros((n, n))
        for i in range(n):
            for j in range(n):
                der[i, j] = self.weights[j] * sum(
                    self.eigenvalues[s] * sp.legendre(s).deriv()(self.nodes[i]) *
                    sp.legendre(s)(self.nodes[j]) for s in range(n))
        der[np.abs(der) < 1.0e-15] = 0
        return der

    def set_eigenvalues(self):                              # main/basis.py:126-132
        ev = np.array([(2.0 * s + 1) / 2.0 for s in range(self.order - 1)])
        return np.append(ev, (self.order - 1) / 2.0)

    def set_vandermonde_inverse(self):                      # main/basis.py:134-137
        return np.array([[self.weights[j] * self.eigenvalues[s] * sp.legendre(s)(self.nodes[j])
                          for j in range(self.order)] for s in range(self.order)])

    def inv_mass_matrix(self):                              # main/basis.py:165-174
        approx_inv = np.diag(np.divide(1.0, self.weights))
        v = sp.legendre(self.order - 1)(self.nodes)
        return approx_inv + self.order / 2 * np.outer(v, v)

    def internal_flux(self):                                # main/basis.py:176-188
        n = self.order
        up = np.zeros((n, n))
        for i in range(n):
            for j in range(n):
                up[i, j] = self.weights[j] * sum(
                    (2 * s + 1) / 2 * sp.legendre(s)(self.nodes[i]) *
                    sp.legendre(s).deriv()(self.nodes[j]) for s in range(n))
        up[np.abs(up) < 1.0e-10] = 0
        return up

    def fourier_transform_array(self, midpoints, J, wave_numbers):  # main/basis.py:204-228
        signs = np.sign(wave_numbers)
        signs[np.where(wave_numbers == 0)] = 1.0
        p_tilde = np.array([(signs ** s) * np.exp(-1j * np.pi / 2.0 * s) *
                            sp.spherical_jn(s, np.absolute(wave_numbers) / J)
                            for s in range(self.order)])
        ell_tilde = np.matmul(self.vandermonde_inverse.T, p_tilde)
        phase = np.exp(-1j * np.tensordot(midpoints, wave_numbers, axes=0))
        transform_array = np.multiply(phase[:, :, None], ell_tilde.T)
        return np.transpose(transform_array, (0, 2, 1))


class Basis3D:
    """main/basis.py:283-289."""

    def __init__(self, orders):
        self.orders = orders
        self.b1 = Basis1D(orders[0])
        self.b2 = Basis1D(orders[1])
        self.b3 = Basis1D(orders[2])


class Grid1D:
    """main/grid.py:10-88."""

    def __init__(self, low, high, res, basis, spectrum=False):
        self.low, self.high = low, high
        self.res = int(res)
        self.res_ghosts = int(res + 2)
        self.order = basis.order
        self.length = self.high - self.low
        self.dx = self.length / self.res
        self.J = 2.0 / self.dx
        self.quad_weights = np.tensordot(np.ones(self.res), np.asarray(basis.weights), axes=0)
        self.arr = np.zeros((self.res_ghosts, self.order))
        self.create_grid(basis.nodes)
        self.midpoints = np.array([(self.arr[i, -1] + self.arr[i, 0]) / 2.0
                                   for i in range(1, self.res_ghosts - 1)])
        self.arr_max = np.amax(abs(self.arr))
        self.one_negatives = np.where(self.arr < 0, 1, 0)
        self.one_positives = np.where(self.arr >= 0, 1, 0)
        if spectrum:                                        # main/grid.py:45-53
            self.nyquist_number = 2.5 * self.length // self.dx
            self.k1 = 2.0 * np.pi / self.length
            self.wave_numbers = self.k1 * np.arange(1 - self.nyquist_number, self.nyquist_number)
            self.grid_phases = np.exp(1j * np.tensordot(self.wave_numbers, self.arr[1:-1, :], axes=0))
            self.spectral_transform = basis.fourier_transform_array(self.midpoints, self.J,
                                                                    self.wave_numbers)

    def create_grid(self, nodes):                           # main/grid.py:55-70
        min_gs = self.low - self.dx
        max_gs = self.high
        nodes = (np.array(nodes) + 1) / 2
        xl = np.linspace(min_gs, max_gs, num=self.res_ghosts)
        for i in range(self.res_ghosts):
            self.arr[i, :] = xl[i] + self.dx * nodes

    def fourier_basis(self, function):                      # main/grid.py:78-82
        return np.tensordot(function, self.spectral_transform, axes=([0, 1], [0, 1])) * self.dx / self.length

    def sum_fourier(self, coefficients):                    # main/grid.py:84-88
        return np.real(np.tensordot(coefficients, self.grid_phases, axes=([0], [0])))


class Grid3D:
    """main/grid.py:91-103."""

    def __init__(self, basis, lows, highs, resolutions):
        self.x = Grid1D(lows[0], highs[0], resolutions[0], basis.b1, spectrum=True)
        self.u = Grid1D(lows[1], highs[1], resolutions[1], basis.b2)
        self.v = Grid1D(lows[2], highs[2], resolutions[2], basis.b3)
        self.no_ghost_slice = (slice(1, self.x.res_ghosts - 1), slice(self.x.order),
                               slice(1, self.u.res_ghosts - 1), slice(self.u.order),
                               slice(1, self.v.res_ghosts - 1), slice(self.v.order))


def ring_distribution(grids, vt, ring_j):
    """Ring distribution on every ghosted node (main/grid.py:244-268)."""
    factor = 1.0 / ((2.0 ** ring_j) * (2.0 * np.pi) * (vt ** 2.0) * math.factorial(ring_j))
    ix = np.tensordot(np.ones(grids.x.res_ghosts), np.ones(grids.x.order), axes=0)
    iu = np.tensordot(np.ones(grids.u.res_ghosts), np.ones(grids.u.order), axes=0)
    iv = np.tensordot(np.ones(grids.v.res_ghosts), np.ones(grids.v.order), axes=0)
    rsq = (np.tensordot(grids.u.arr ** 2.0, iv, axes=0) +
           np.tensordot(iu, grids.v.arr ** 2.0, axes=0)) / (vt ** 2.0)
    pol_eig = np.multiply(rsq ** ring_j, np.exp(-0.5 * rsq))
    return factor * np.tensordot(ix, pol_eig, axes=0)


def eigenfunction(grids, vt, ring_j, om, om_pc, delta_n, terms=20):
    """Bernstein-mode perturbation of the ring (main/grid.py:207-242, the
    non-symmetric branch): on every ghosted node
        2 delta_n Re( exp(i k1 x) * (-i) exp(i beta sin phi) * df_dv * Upsilon / k1^2 ),
    Upsilon = sum_{n=-terms..terms} n / (n - om) J_n(beta) exp(-i n phi),
    beta = -k1 r om_pc, phi = atan2(v, u), df_dv = f0 (j / (x + 1e-16) - 1) / vt^2,
    f0 = x^j e^-x / (2 pi vt^2 j!), x = (r / vt)^2 / 2."""
    u = np.tensordot(grids.u.arr, np.ones_like(grids.v.arr), axes=0)
    v = np.tensordot(np.ones_like(grids.u.arr), grids.v.arr, axes=0)
    r = np.sqrt(u ** 2.0 + v ** 2.0)
    phi = np.arctan2(v, u)
    k1 = grids.x.k1
    beta = -k1 * r * om_pc
    x = 0.5 * (r / vt) ** 2.0
    f0 = 1 / (2.0 * np.pi * (vt ** 2.0) * math.factorial(ring_j)) * np.multiply(x ** ring_j, np.exp(-x))
    df_dv = np.multiply(f0, (ring_j / (x + 1.0e-16) - 1.0)) / (vt ** 2.0)
    upsilon = np.array([n / (n - om) * np.multiply(sp.jv(n, beta), np.exp(-1j * n * phi))
                        for n in range(-terms, terms + 1)]).sum(axis=0)
    eig = np.multiply(df_dv, upsilon) / k1 ** 2.0
    vel_mode = -1j * np.multiply(np.exp(1j * np.multiply(beta, np.sin(phi))), eig)
    return 2.0 * delta_n * np.real(np.tensordot(np.exp(1j * k1 * grids.x.arr), vel_mode, axes=0))


def density_modulation(x, k1, amp1, amps, phases):
    """1 + amp1 cos(k1 x) + sum_m amps[m] cos((m+2) k1 x + phases[m])."""
    out = 1.0 + amp1 * np.cos(k1 * x)
    for m, (a, ph) in enumerate(zip(amps, phases)):
        out = out + a * np.cos((m + 2) * k1 * x + ph)
    return out


class Distribution:
    """State container (main/grid.py:137-321)."""

    def __init__(self, vt, ring_j, resolutions, orders):
        self.vt, self.ring_j = vt, ring_j
        self.arr = None
        self.quad_weights = None

    def initialize_quad_weights(self, grids):               # main/grid.py:297-301
        self.quad_weights = np.tensordot(grids.u.quad_weights, grids.v.quad_weights,
                                         axes=0) / (grids.u.J * grids.v.J)

    def initialize(self, grids, amp1=1e-3, amps=(1e-6, 1e-6, 1e-6), phases=(0.0, 0.0, 0.0)):
        ring = ring_distribution(grids, self.vt, self.ring_j)
        mod = density_modulation(grids.x.arr, grids.x.k1, amp1, amps, phases)
        self.arr = ring * mod[:, :, None, None, None, None]

    def initialize_eigen(self, grids, om, om_pc, delta_n, perturbation=True):
        """The reference's own initial condition (main/grid.py:244-279): the
        ring plus, when ``perturbation``, the Bernstein eigenfunction."""
        self.arr = ring_distribution(grids, self.vt, self.ring_j)
        if perturbation:
            self.arr = self.arr + eigenfunction(grids, self.vt, self.ring_j, om, om_pc, delta_n)

    def swap_ghost_cells_of_array(self):                    # main/grid.py:187-189
        self.arr[0, :, :, :, :, :] = np.copy(self.arr[-2, :, :, :, :, :])
        self.arr[-1, :, :, :, :, :] = np.copy(self.arr[1, :, :, :, :, :])

    def moment_zero(self):                                  # main/grid.py:303-312
        return self.moment_zero_of_arr(self.arr)

    def moment_zero_of_arr(self, arr):                      # main/grid.py:319-321
        return np.tensordot(arr[1:-1, :, 1:-1, :, 1:-1, :], self.quad_weights,
                            axes=([2, 3, 4, 5], [0, 1, 2, 3]))


class Elliptic:
    """Spectral Poisson solve and field energy (main/elliptic.py:9-22,117-151)."""

    def __init__(self, poisson_coefficient):
        self.poisson_coefficient = poisson_coefficient
        self.potential = None
        self.electric_field = None
        self.magnetic_field = None

    def poisson2(self, charge_density, grid, basis=None):   # main/elliptic.py:117-145
        rhs = self.poisson_coefficient * charge_density
        rhs_coefficients = grid.fourier_basis(rhs)
        k = grid.wave_numbers
        nz = k != 0
        field_coefficients = np.zeros_like(rhs_coefficients)
        potential_coefficients = np.zeros_like(rhs_coefficients)
        # k = 0: CuPy gives NaN -> nan_to_num -> 0 (main/elliptic.py:124-128)
        field_coefficients[nz] = -np.divide(rhs_coefficients[nz], 1j * k[nz])
        potential_coefficients[nz] = -np.divide(rhs_coefficients[nz], k[nz] ** 2.0)
        self.electric_field = np.zeros_like(grid.arr)
        self.electric_field[1:-1, :] = grid.sum_fourier(field_coefficients)
        self.potential = grid.sum_fourier(potential_coefficients)
        self.electric_field[0, :] = self.electric_field[-2, :]
        self.electric_field[-1, :] = self.electric_field[1, :]

    def build_central_flux_operator(self, grid, basis):     # main/elliptic.py:24-81
        """The stabilized central-flux DG Laplacian, assembled column by column
        from nodal indicators as the reference does, with the gauge row and
        column (zero quadrature mean of the potential)."""
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
                operator = (np.tensordot(basis.stf, grad, axes=([1], [1])) +
                            np.tensordot(basis.face_mass, grad_num_flux + face_diff0, axes=([1], [1]))).T
                op[i, j, :, :] = operator
                self.gradient_operator[i, j, :, :] = grad
                indicator[i, j] = 0
        qw = grid.quad_weights.reshape(res * order, 1)
        op0 = np.hstack([op.reshape(res * order, res * order), qw])
        self.central_flux_operator = np.vstack([op0, np.append(grid.quad_weights.flatten(), 0)])
        self.central_flux_operator[np.abs(self.central_flux_operator) < 1.0e-15] = 0

    def invert(self):                                       # main/elliptic.py:83-84
        self.inv_op = np.linalg.inv(self.central_flux_operator)

    def poisson(self, charge_density, grid, basis, anti_alias=True):   # main/elliptic.py:86-115
        rhs = np.zeros((grid.res * grid.order + 1))
        rhs[:-1] = self.poisson_coefficient * np.tensordot(charge_density, basis.mass, axes=([1], [1])).flatten()
        sol = np.matmul(self.inv_op, rhs)[:-1] / (grid.J ** 2.0)
        self.potential = sol.reshape(grid.res, grid.order)
        if anti_alias:
            self.potential = grid.sum_fourier(grid.fourier_basis(self.potential))
        self.electric_field = np.zeros_like(grid.arr)
        self.electric_field[1:-1, :] = -1.0 * (grid.J * np.tensordot(self.gradient_operator, self.potential,
                                                                     axes=([0, 1], [0, 1])))
        if anti_alias:
            self.electric_field[1:-1, :] = grid.sum_fourier(grid.fourier_basis(self.electric_field[1:-1, :]))
        self.electric_field[0, :] = self.electric_field[-2, :]
        self.electric_field[-1, :] = self.electric_field[1, :]

    def set_magnetic_field(self, magnetic_field):           # main/elliptic.py:147-148
        self.magnetic_field = magnetic_field

    def electric_energy(self, grid):                        # main/elliptic.py:150-151
        return 0.5 * np.tensordot(self.electric_field[1:-1, :] ** 2.0, grid.quad_weights,
                                  axes=([0, 1], [0, 1]))


def basis_product(flux, basis_arr, axis, permutation):      # main/fluxes.py:8-10
    return np.transpose(np.tensordot(flux, basis_arr, axes=([axis], [1])), axes=permutation)


class DGFlux:
    """Semi-discrete upwind DG right-hand side (main/fluxes.py:13-207).

    ``resolutions`` are the ghosted element counts, as main/main.py:148 passes.
    """

    def __init__(self, resolutions, orders, flux_coefficients):
        r, o = resolutions, orders
        self.resolutions, self.orders = r, o
        self.permutations = [(0, 5, 1, 2, 3, 4), (0, 1, 2, 5, 3, 4), (0, 1, 2, 3, 4, 5)]
        s = slice
        self.boundary_slices = [
            [(s(r[0]), 0, s(r[1]), s(o[1]), s(r[2]), s(o[2])),
             (s(r[0]), -1, s(r[1]), s(o[1]), s(r[2]), s(o[2]))],
            [(s(r[0]), s(o[0]), s(r[1]), 0, s(r[2]), s(o[2])),
             (s(r[0]), s(o[0]), s(r[1]), -1, s(r[2]), s(o[2]))],
            [(s(r[0]), s(o[0]), s(r[1]), s(o[1]), s(r[2]), 0),
             (s(r[0]), s(o[0]), s(r[1]), s(o[1]), s(r[2]), -1)]]
        self.speed_slices = [(None, s(r[1]), s(o[1]), None, None),
                             (s(r[0]), s(o[0]), None, s(r[2]), s(o[2])),
                             (None, None, s(r[1]), s(o[1]), None)]
        self.grid_axis = np.array([0, 2, 4])
        self.sub_element_axis = np.array([1, 3, 5])
        self.acceleration_coefficient = flux_coefficients
        self.num_flux_sizes = [(r[0], 2, r[1], o[1], r[2], o[2]),
                               (r[0], o[0], r[1], 2, r[2], o[2]),
                               (r[0], o[0], r[1], o[1], r[2], 2)]

    def semi_discrete_rhs(self, function, elliptic, basis, grids):  # main/fluxes.py:59-105
        return ((self.x_flux(function, basis.b1, grids.u) * grids.x.J) +
                (self.u_flux(function, basis.b2, elliptic, grids.v) * grids.u.J) +
                (self.v_flux(function, basis.b3, elliptic, grids.u) * grids.v.J))

    def x_flux(self, function, basis, grid_u):              # main/fluxes.py:107-115
        dim = 0
        flux = np.multiply(function, grid_u.arr[None, None, :, :, None, None])
        return (basis_product(flux, basis.up, self.sub_element_axis[dim], self.permutations[dim])
                - self.spatial_flux(flux, grid_u, basis, dim))

    def u_flux(self, function, basis, elliptic, grid_v):    # main/fluxes.py:117-138
        dim = 1
        speed = self.acceleration_coefficient * (elliptic.electric_field[:, :, None, None] +
                                                 elliptic.magnetic_field * grid_v.arr[None, None, :, :])
        flux = np.multiply(function, speed[:, :, None, None, :, :])
        return (basis_product(flux, basis.up, self.sub_element_axis[dim], self.permutations[dim])
                - self.velocity_flux(flux, speed, basis, dim))

    def v_flux(self, function, basis, elliptic, grid_u):    # main/fluxes.py:140-153
        dim = 2
        speed = self.acceleration_coefficient * (-elliptic.magnetic_field * grid_u.arr)
        flux = np.multiply(function, speed[None, None, :, :, None, None])
        return (basis_product(flux, basis.up, self.sub_element_axis[dim], self.permutations[dim])
                - self.velocity_flux(flux, speed, basis, dim))

    def _upwind(self, flux, pos, neg, basis, dim):          # main/fluxes.py:158-178 / 183-207
        num_flux = np.zeros(self.num_flux_sizes[dim])
        bl, br = self.boundary_slices[dim]
        ss = self.speed_slices[dim]
        ax = self.grid_axis[dim]
        num_flux[bl] = -1.0 * (np.multiply(np.roll(flux[br], shift=1, axis=ax), pos[ss]) +
                               np.multiply(flux[bl], neg[ss]))
        num_flux[br] = (np.multiply(flux[br], pos[ss]) +
                        np.multiply(np.roll(flux[bl], shift=-1, axis=ax), neg[ss]))
        return basis_product(num_flux, basis.xi, self.sub_element_axis[dim], self.permutations[dim])

    def spatial_flux(self, flux, speed, basis, dim):        # main/fluxes.py:156-178
        return self._upwind(flux, speed.one_positives, speed.one_negatives, basis, dim)

    def velocity_flux(self, flux, speed, basis, dim):       # main/fluxes.py:181-207
        one_negatives = np.where(speed < 0, 1, 0)
        one_positives = np.where(speed >= 0, 1, 0)
        return self._upwind(flux, one_positives, one_negatives, basis, dim)


def get_max_speeds(grids, elliptic, acc):                   # main/timestep.py:235-242
    return np.absolute(np.array([
        grids.u.arr_max,
        np.amax(np.absolute(acc * elliptic.electric_field)) + abs(acc * elliptic.magnetic_field * grids.v.arr_max),
        acc * abs(elliptic.magnetic_field * grids.u.arr_max)]))


# Linear SSP-RK coefficients (main/timestep.py:11-21)
SSP_RK_LINEAR = {
    1: [1],
    2: [1 / 2, 1 / 2],
    3: [1 / 3, 1 / 2, 1 / 6],
    4: [3 / 8, 1 / 3, 1 / 4, 1 / 24],
}

# Courant numbers of the 3-stage row, n = 1..9 (main/timestep.py:24-33)
COURANT_NUMBERS_RK3 = [1.256, 0.409, 0.209, 0.130, 0.089, 0.066, 0.051, 0.040, 0.033]


def adapt_dt(courant, max_speeds, dx, du, dv):              # main/timestep.py:231-232
    return courant / ((max_speeds[0] / dx) + (max_speeds[1] / du) + (max_speeds[2] / dv))


class Stepper:
    """Shu-Osher SSP-RK3 (main/timestep.py:42-187) with a fixed time step."""

    def __init__(self, dt, time_order=3):
        self.time_order = time_order
        self.coefficients = np.array(NONLINEAR_SSP_RK[time_order])
        self.dt = dt
        self.time = 0.0
        self.steps_counter = 0
        self.time_array = np.array([0.0])
        self.field_energy = np.array([])

    def nonlinear_ssp_rk(self, distribution, basis, elliptic, grids, dg_flux):  # main/timestep.py:141-187
        distribution.swap_ghost_cells_of_array()
        t_shape = tuple([self.time_order] + list(distribution.arr.shape))
        arr_stages = np.zeros(t_shape)
        arr_stages[:, :, :, 0, :, :, :] = distribution.arr[:, :, 0, :, :, :]
        arr_stages[:, :, :, -1, :, :, :] = distribution.arr[:, :, -1, :, :, :]
        arr_stages[:, :, :, :, :, 0, :] = distribution.arr[:, :, :, :, 0, :]
        arr_stages[:, :, :, :, :, -1, :] = distribution.arr[:, :, :, :, -1, :]
        density = distribution.moment_zero()
        elliptic.poisson2(np.mean(density) - density, grids.x, basis.b1)
        ng = grids.no_ghost_slice
        no_ghost_idx = tuple([0] + list(ng))
        arr_stages[no_ghost_idx] = distribution.arr[ng] + self.dt * dg_flux.semi_discrete_rhs(
            distribution.arr, elliptic, basis, grids)[ng]
        for i in range(1, self.time_order):
            arr_stages[i - 1, 0, :, :, :, :, :] = arr_stages[i - 1, -2, :, :, :, :, :]
            arr_stages[i - 1, -1, :, :, :, :, :] = arr_stages[i - 1, 1, :, :, :, :, :]
            density = distribution.moment_zero_of_arr(arr_stages[i - 1])
            elliptic.poisson2(np.mean(density) - density, grids.x, basis.b1)
            df_dt = dg_flux.semi_discrete_rhs(arr_stages[i - 1], elliptic, basis, grids)[ng]
            g_i = tuple([i] + list(ng))
            g_i1 = tuple([i - 1] + list(ng))
            c = self.coefficients[i - 1]
            arr_stages[g_i] = c[0] * distribution.arr[ng] + c[1] * arr_stages[g_i1] + c[2] * self.dt * df_dt
        distribution.arr[ng] = arr_stages[-1, 1:-1, :, 1:-1, :, 1:-1, :]

    def ssp_rk_update(self, distribution, basis, elliptic, grids, dg_flux, coefficients=None):
        """The linear SSP-RK update (main/timestep.py:189-229) with the
        central-flux Poisson solve (elliptic.poisson); the reference's
        closing adapt_time_step is left to the caller."""
        c = SSP_RK_LINEAR[self.time_order] if coefficients is None else coefficients
        ng = grids.no_ghost_slice
        distribution.swap_ghost_cells_of_array()
        copy = np.copy(distribution.arr)
        distribution.arr[ng] = c[0] * distribution.arr[ng]

        def field_of(arr):
            arr[0] = np.copy(arr[-2])
            arr[-1] = np.copy(arr[1])
            density = distribution.moment_zero_of_arr(arr)
            elliptic.poisson(np.mean(density) - density, grids.x, basis.b1)
        for i in range(1, len(c)):
            field_of(copy)
            copy[ng] += self.dt * dg_flux.semi_discrete_rhs(copy, elliptic, basis, grids)[ng]
            distribution.arr[ng] += c[i] * copy[ng]
        field_of(copy)
        distribution.arr[ng] += c[-1] * self.dt * dg_flux.semi_discrete_rhs(copy, elliptic, basis, grids)[ng]

    def main_loop(self, steps, distribution, basis, elliptic, grids, dg_flux, courant=None, acc=None,
                  final_time=None, on_step=None):  # main/timestep.py:81-139
        """Fixed dt, or with ``courant`` the reference's adaptive rule: dt from
        the set-up field before the loop (main/timestep.py:88-89), again after
        every step from the field of the last stage solve (:184), and the time
        advanced by that new dt (:96).  ``final_time``: loop while time <
        final_time instead of ``steps`` steps.  ``on_step(self)`` runs after
        every step (the fingerprint writer records checkpoints with it)."""
        self.field_energy = np.append(self.field_energy, elliptic.electric_energy(grids.x))
        adapt = (lambda: adapt_dt(courant, get_max_speeds(grids, elliptic, acc), grids.x.dx, grids.u.dx,
                                  grids.v.dx)) if courant is not None else None
        if adapt:
            self.dt = adapt()
        while (self.time < final_time) if final_time is not None else (self.steps_counter < steps):
            self.nonlinear_ssp_rk(distribution, basis, elliptic, grids, dg_flux)
            if adapt:
                self.dt = adapt()
            self.time += self.dt
            self.steps_counter += 1
            self.time_array = np.append(self.time_array, self.time)
            self.field_energy = np.append(self.field_energy, elliptic.electric_energy(grids.x))
            if on_step is not None:
                on_step(self)


class Case:
    """One seeded run set up in the order of main/main.py:37-166.

    Physics: x in [-L/2, L/2) with L = 2 pi / k0, velocity box [-vmax, vmax]^2,
    ``flux_coefficients = acc``, ``poisson_coefficient = poisson``,
    ``magnetic_field = B``, ``vt = 1``.
    """

    def __init__(self, n, nx, nu, nv, k0=0.1, vmax=8.5, B=0.1, ring_j=6, acc=-1.0,
                 poisson=-1.0, vt=1.0, amp1=1e-3, amps=(1e-6, 1e-6, 1e-6), seed=0, phases=None,
                 ic="modulated", om=0.9, om_pc=1.0, delta_n=0.01):
        self.n = n
        self.basis = Basis3D([n, n, n])
        L = 2.0 * np.pi / k0
        lows = np.array([-L / 2.0, -vmax * vt, -vmax * vt])
        highs = np.array([L / 2.0, vmax * vt, vmax * vt])
        self.grids = Grid3D(self.basis, lows, highs, np.array([nx, nu, nv]))
        if phases is None:
            phases = seeded_phases(seed, len(amps))
        self.phases = np.asarray(phases, dtype=np.float64)
        self.dist = Distribution(vt, ring_j, [nx, nu, nv], [n, n, n])
        self.dist.initialize_quad_weights(self.grids)
        if ic == "eigen":                                    # main/main.py:73-77
            self.dist.initialize_eigen(self.grids, om, om_pc, delta_n)
        else:
            self.dist.initialize(self.grids, amp1, amps, self.phases)
        self.elliptic = Elliptic(poisson)
        self.elliptic.set_magnetic_field(B)
        n0 = self.dist.moment_zero()
        self.elliptic.poisson2(n0 - np.mean(n0), self.grids.x, self.basis.b1)  # main/main.py:131
        self.flux = DGFlux(np.array([nx + 2, nu + 2, nv + 2]), np.array([n, n, n]), acc)
        self.acc = acc

    def rhs(self, arr=None):
        arr = self.dist.arr if arr is None else arr
        return self.flux.semi_discrete_rhs(arr, self.elliptic, self.basis, self.grids)

    def run(self, steps, dt, final_time=None):
        """``steps`` steps of a fixed dt, or with dt=None the adaptive rule
        (Courant number of the 3-stage row for n nodes)."""
        st = Stepper(dt)
        courant = COURANT_NUMBERS_RK3[self.n - 1] if dt is None else None
        st.main_loop(steps, self.dist, self.basis, self.elliptic, self.grids, self.flux, courant=courant,
                     acc=self.acc, final_time=final_time)
        return st


def seeded_phases(seed, count=3):
    """Random phases phi_m ~ U[0, 2 pi) from numpy's default_rng(seed)."""
    rng = np.random.default_rng(seed)
    return rng.uniform(0.0, 2.0 * np.pi, size=count)
