"""Pins of the oracle's SSP-RK3 stepper (main/timestep.py:35-39,141-187) and
of its initial condition and moment (main/grid.py:244-321)."""
import math

import numpy as np
import pytest
from scipy.special import gammaincc

import oracle
from helpers import interior


class _LinearFlux:
    def __init__(self, lam):
        self.lam = lam

    def semi_discrete_rhs(self, function, elliptic, basis, grids):
        return self.lam * function


class _NoField:
    def poisson2(self, *a, **k):
        pass


@pytest.mark.parametrize("z", [-0.5, 0.3, -2.0, 1e-3])
def test_rk3_amplification_polynomial(z):
    """With L f = lam f the step multiplies the interior by 1 + z + z^2/2 + z^3/6."""
    case = oracle.Case(n=3, nx=2, nu=2, nv=2)
    f0 = case.dist.arr.copy()
    dt = 0.1
    st = oracle.Stepper(dt)
    st.nonlinear_ssp_rk(case.dist, case.basis, _NoField(), case.grids, _LinearFlux(z / dt))
    R = 1 + z + z * z / 2 + z ** 3 / 6
    assert np.allclose(interior(case.dist.arr), R * interior(f0), rtol=1e-14, atol=0)


def test_shu_osher_table():
    assert oracle.NONLINEAR_SSP_RK[3] == [[3 / 4, 1 / 4, 1 / 4], [1 / 3, 2 / 3, 2 / 3]]


def test_velocity_ghosts_stay_frozen():
    case = oracle.Case(n=4, nx=3, nu=4, nv=4)
    f0 = case.dist.arr.copy()
    case.run(3, 0.01)
    f = case.dist.arr
    for sl in [(slice(1, -1), slice(None), 0), (slice(1, -1), slice(None), -1)]:
        assert np.array_equal(f[sl], f0[sl])
    assert np.array_equal(f[1:-1, :, :, :, 0], f0[1:-1, :, :, :, 0])
    assert np.array_equal(f[1:-1, :, :, :, -1], f0[1:-1, :, :, :, -1])


@pytest.mark.parametrize("j", [1, 6])
def test_ring_density_integrates_to_one(j):
    """The quadrature moment of the unmodulated ring equals 1 minus the tail
    beyond the box, P(j+1, vmax^2/2)-ish bounds; here checked against the
    incomplete-gamma tail of the disk r < vmax (inside the square box)."""
    case = oracle.Case(n=8, nx=2, nu=12, nv=12, ring_j=j, amp1=0.0, amps=(0.0, 0.0, 0.0))
    rho = case.dist.moment_zero()
    tail_disk = gammaincc(j + 1, 8.5 ** 2 / 2)       # mass outside r = 8.5
    assert np.all(rho <= 1.0 + 1e-10)
    assert np.all(np.abs(rho - 1.0) < max(2 * tail_disk, 1e-9))
    assert np.ptp(rho) < 1e-14                       # x-independent


def test_ring_normalisation_factor():
    # 1 / (2^j 2 pi j!) for vt = 1 (main/grid.py:253)
    case = oracle.Case(n=3, nx=1, nu=1, nv=1, ring_j=4, amp1=0.0, amps=(0.0, 0.0, 0.0))
    u = case.grids.u.arr[1, 1]
    v = case.grids.v.arr[1, 1]
    r2 = u * u + v * v
    want = r2 ** 4 * math.exp(-r2 / 2) / (2 ** 4 * 2 * math.pi * math.factorial(4))
    assert abs(case.dist.arr[1, 0, 1, 1, 1, 1] - want) < 1e-15 * max(want, 1e-300)


def test_energy_series_is_lagged():
    """Entry k (k >= 1) is the energy of the field solved from f2 of step k
    (main/timestep.py:102 after :172)."""
    case = oracle.Case(n=4, nx=3, nu=4, nv=4)
    st = case.run(2, 0.01)
    assert len(st.field_energy) == 3
    assert st.field_energy[-1] == case.elliptic.electric_energy(case.grids.x)


def test_adaptive_dt_known_answer_at_the_reference_size_setup():
    """The CFL rule (main/timestep.py:231-242) on the 2.1M-DOF workload's grid
    (n=8, 16^3 elements, k0=0.1, B=0.1, [-8.5,8.5]^2): max|u| = max|v| =
    8.5 + du over the ghosted nodes (du = 17/16), dx = 2 pi / (0.1 * 16),
    Courant number 0.040 -> 0.00944 for a vanishing field (the rule's value
    quoted for this workload); the set-up field, E ~ 1e-3/k0 from the 1e-3
    density modulation, adds max|acc E| ~ 0.01 to the u speed."""
    case = oracle.Case(n=8, nx=16, nu=16, nv=16)
    sp = oracle.get_max_speeds(case.grids, case.elliptic, case.acc)
    um = 8.5 + 17 / 16
    assert sp[0] == pytest.approx(um, rel=1e-15)
    assert sp[1] - 0.1 * um == pytest.approx(1e-3 / 0.1, rel=0.02)
    dt = oracle.adapt_dt(oracle.COURANT_NUMBERS_RK3[7], sp, case.grids.x.dx, case.grids.u.dx, case.grids.v.dx)
    # zero field: the closed form with speeds (max|u|, |B| max|v|, |B| max|u|)
    dx, du = 2 * math.pi / (0.1 * 16), 17 / 16
    dt0 = 0.040 / (um / dx + 0.1 * um / du + 0.1 * um / du)
    assert round(dt0, 5) == 0.00944
    case.elliptic.electric_field = np.zeros_like(case.elliptic.electric_field)
    sp0 = oracle.get_max_speeds(case.grids, case.elliptic, case.acc)
    assert oracle.adapt_dt(0.040, sp0, case.grids.x.dx, case.grids.u.dx, case.grids.v.dx) == pytest.approx(dt0, rel=1e-14)
    assert dt == pytest.approx(0.040 / (0.040 / dt0 + (sp[1] - 0.1 * um) / du), rel=1e-14) and dt < dt0


def test_adaptive_loop_advances_time_by_the_new_dt():
    """main/timestep.py:184 adapts dt at the end of the step and :96 then adds
    the new dt: time_array[k] = sum of the dts adapted after steps 1..k."""
    case = oracle.Case(n=3, nx=3, nu=3, nv=3)
    st = oracle.Stepper(None)
    dts = []
    from oracle import vlasov_oracle as vo
    orig = vo.adapt_dt

    def spy(*a):
        dts.append(orig(*a))
        return dts[-1]
    vo.adapt_dt = spy
    try:
        st.main_loop(4, case.dist, case.basis, case.elliptic, case.grids, case.flux, courant=0.209, acc=case.acc)
    finally:
        vo.adapt_dt = orig
    assert len(dts) == 5 and st.steps_counter == 4
    np.testing.assert_allclose(st.time_array[1:], np.cumsum(dts[1:]), rtol=1e-15)


@pytest.mark.parametrize("z", [-0.5, 0.3, -2.0, 1e-3])
def test_linear_ssp_rk3_amplification_polynomial(z):
    """main/timestep.py:189-229 with L f = lam f: the interior is multiplied by
    1 + z + z^2/2 + z^3/6 (coefficients 1/3, 1/2, 1/6)."""

    class _Field:
        def poisson(self, *a, **k):
            pass
    case = oracle.Case(n=3, nx=2, nu=2, nv=2)
    f0 = case.dist.arr.copy()
    dt = 0.1
    st = oracle.Stepper(dt)
    st.ssp_rk_update(case.dist, case.basis, _Field(), case.grids, _LinearFlux(z / dt))
    R = 1 + z + z * z / 2 + z ** 3 / 6
    assert np.allclose(interior(case.dist.arr), R * interior(f0), rtol=1e-14, atol=0)
    assert oracle.SSP_RK_LINEAR[3] == [1 / 3, 1 / 2, 1 / 6]


# ---- coupled Vlasov-Poisson stepping: temporal self-convergence -------------

class _StaleFieldStepper(oracle.Stepper):
    """A planted mistake: every stage advances with the field solved from f^n
    instead of the field of its own input (main/timestep.py:153-182 solves
    rho -> E before each of the three right-hand sides)."""

    def nonlinear_ssp_rk(self, distribution, basis, elliptic, grids, dg_flux):
        distribution.swap_ghost_cells_of_array()
        density = distribution.moment_zero()
        elliptic.poisson2(np.mean(density) - density, grids.x, basis.b1)
        ng = grids.no_ghost_slice
        f = distribution.arr
        g = f.copy()
        g[ng] = f[ng] + self.dt * dg_flux.semi_discrete_rhs(f, elliptic, basis, grids)[ng]
        for i in range(1, self.time_order):
            g[0] = g[-2]
            g[-1] = g[1]
            c = self.coefficients[i - 1]
            h = g.copy()
            h[ng] = c[0] * f[ng] + c[1] * g[ng] + c[2] * self.dt * dg_flux.semi_discrete_rhs(
                g, elliptic, basis, grids)[ng]
            g = h
        f[ng] = g[ng]


# A strongly coupled small case: a 10 % density modulation at k0 = 0.5 drives
# a field of O(0.2) that oscillates at about the plasma frequency.
_COUPLED = dict(n=4, nx=4, nu=6, nv=6, k0=0.5, amp1=0.1, amps=(0.0, 0.0, 0.0), B=0.1)


def _final_state(stepper_cls, steps, t_end=1.0):
    case = oracle.Case(**_COUPLED)
    st = stepper_cls(t_end / steps)
    st.main_loop(steps, case.dist, case.basis, case.elliptic, case.grids, case.flux)
    return interior(case.dist.arr).copy()


def _convergence_orders(stepper_cls):
    """log2 error ratios of dt = 1/32, 1/64, 1/128 against dt = 1/512 at t = 1."""
    ref = _final_state(stepper_cls, 512)
    errs = [np.linalg.norm(_final_state(stepper_cls, m) - ref) / np.linalg.norm(ref) for m in (32, 64, 128)]
    return [math.log2(errs[i] / errs[i + 1]) for i in range(2)]


def test_coupled_rk3_is_third_order_in_time():
    """Shu-Osher SSP-RK3 on the coupled Vlasov-Poisson system converges at
    third order in dt only if each stage solves its field from its own input
    (main/timestep.py:153-182); the error ratios per halving are ~8."""
    orders = _convergence_orders(oracle.Stepper)
    assert all(2.8 < p < 3.3 for p in orders), orders


def test_coupled_convergence_detects_stale_stage_field():
    """The same measurement on the planted stale-field stepper gives first
    order, so the test above fails for that mistake."""
    orders = _convergence_orders(_StaleFieldStepper)
    assert all(p < 1.5 for p in orders), orders


# ---- the lagged energy series -----------------------------------------------

def _field_energy_of(case, arr):
    """Energy of the field solved from a stage array (x ghosts refreshed)."""
    a = arr.copy()
    a[0] = a[-2]
    a[-1] = a[1]
    rho = case.dist.moment_zero_of_arr(a)
    ell = oracle.Elliptic(case.elliptic.poisson_coefficient)
    ell.set_magnetic_field(case.elliptic.magnetic_field)
    ell.poisson2(np.mean(rho) - rho, case.grids.x, case.basis.b1)
    return ell, ell.electric_energy(case.grids.x)


def test_energy_series_holds_the_stage3_field_energy():
    """After a step the series appends 0.5 sum w E^2 of the field solved from
    f2, the input of the third stage (main/timestep.py:102 reads the elliptic
    state the stage-3 solve left, :170-172), not of the field of f^{n+1}.
    f1 and f2 are recomputed here with the oracle's flux and Poisson solve
    (pinned on their own by test_oracle_flux / test_oracle_poisson)."""
    case = oracle.Case(**_COUPLED)
    dt = 1.0 / 32
    ng = case.grids.no_ghost_slice
    f = case.dist.arr.copy()
    f[0] = f[-2]
    f[-1] = f[1]
    e0 = case.elliptic.electric_energy(case.grids.x)     # set-up solve (main/main.py:131)
    ell, _ = _field_energy_of(case, f)
    f1 = f.copy()
    f1[ng] = f[ng] + dt * case.flux.semi_discrete_rhs(f, ell, case.basis, case.grids)[ng]
    ell, _ = _field_energy_of(case, f1)
    f1g = f1.copy()
    f1g[0] = f1g[-2]
    f1g[-1] = f1g[1]
    f2 = f.copy()
    f2[ng] = 0.75 * f[ng] + 0.25 * f1g[ng] + 0.25 * dt * case.flux.semi_discrete_rhs(
        f1g, ell, case.basis, case.grids)[ng]
    _, w_f2 = _field_energy_of(case, f2)

    st = oracle.Stepper(dt)
    st.main_loop(1, case.dist, case.basis, case.elliptic, case.grids, case.flux)
    _, w_new = _field_energy_of(case, case.dist.arr)
    assert st.field_energy[0] == e0
    assert abs(st.field_energy[1] - w_f2) <= 1e-13 * w_f2
    # the two candidates differ by far more than the tolerance, so a series
    # built from a re-solve of f^{n+1} fails the check above
    assert abs(w_new - w_f2) > 1e-5 * w_f2
