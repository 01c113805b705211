"""The reference driver's initial condition on the device (ring + Bernstein
eigenfunction, main/grid.py:207-279; vl_init_eigen) against the oracle's
NumPy/SciPy restatement, and runs started from it against the oracle."""
import numpy as np
import pytest

from helpers import interior, rel_l2

pytestmark = pytest.mark.gpu

# (om, om_pc, delta_n, k0, ring_j): the reference driver's values (om = -0.349j,
# om_pc = 10, delta_n = 1e-3, k = 0.886 / om_pc, j = 6; main/main.py:32-34,52,71),
# the class defaults (om = 0.9, om_pc = 1, delta_n = 0.01) and a complex om
PARAMS = [(-0.349j, 10.0, 1e-3, 0.0886, 6), (0.9, 1.0, 0.01, 0.1, 6), (1.2 - 0.3j, 4.0, 1e-3, 0.15, 3)]


@pytest.mark.parametrize("om,om_pc,delta_n,k0,ring_j", PARAMS)
@pytest.mark.parametrize("n", [4, 8])
def test_eigen_initial_condition_matches_oracle(cuda, om, om_pc, delta_n, k0, ring_j, n):
    import oracle
    from vlasov1d2v_amd import Basis3D, Distribution, Grid3D
    nx, nu, nv = 3, 6, 5
    L = 2 * np.pi / k0
    lows, highs = np.array([-L / 2, -8.5, -8.5]), np.array([L / 2, 8.5, 8.5])
    grids = Grid3D(Basis3D([n, n, n]), lows, highs, [nx, nu, nv])
    d = Distribution(1.0, ring_j, [nx, nu, nv], [n, n, n], om=om, om_pc=om_pc, delta_n=delta_n, device=cuda)
    d.initialize_gpu(grids)
    got = d.arr.cpu().numpy()
    case = oracle.Case(n=n, nx=nx, nu=nu, nv=nv, k0=k0, ring_j=ring_j, ic="eigen", om=om, om_pc=om_pc,
                       delta_n=delta_n)
    ref = case.dist.arr
    pert_ref = oracle.eigenfunction(case.grids, 1.0, ring_j, om, om_pc, delta_n)
    pert_got = d._eigen_state(grids, pert_only=True).cpu().numpy()
    e_tot = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
    e_pert = np.max(np.abs(pert_got - pert_ref)) / np.max(np.abs(pert_ref))
    print(f"om={om} n={n}: total {e_tot:.2e}, perturbation {e_pert:.2e} (|pert|max {np.abs(pert_ref).max():.2e})",
          flush=True)
    assert e_tot < 1e-13 and e_pert < 1e-12
    # without perturbation: the ring alone
    d2 = Distribution(1.0, ring_j, [nx, nu, nv], [n, n, n], perturbation=False, device=cuda)
    d2.initialize_gpu(grids)
    ring = oracle.ring_distribution(case.grids, 1.0, ring_j)
    assert np.max(np.abs(d2.arr.cpu().numpy() - ring)) <= 1e-14 * np.max(ring)


def test_run_from_eigen_initial_condition_matches_oracle(cuda):
    """Three SSP-RK3 steps from the reference driver's initial condition."""
    from vlasov1d2v_amd import Run
    import oracle
    n, nx, nu, nv, dt, steps = 5, 3, 4, 4, 0.01, 3
    kw = dict(k0=0.0886, ring_j=6, om=-0.349j, om_pc=10.0, delta_n=1e-3)
    case = oracle.Case(n=n, nx=nx, nu=nu, nv=nv, ic="eigen", B=0.1, **kw)
    st_ref = case.run(steps, dt)
    run = Run(n, nx, nu, nv, dt=dt, ic="eigen", b_field=0.1, **kw)
    st = run.run(steps)
    err_f = rel_l2(interior(run.dist.arr.cpu().numpy()), interior(case.dist.arr))
    err_e = rel_l2(st.field_energy, st_ref.field_energy)
    print(f"rel-L2 f {err_f:.2e}, energy {err_e:.2e}", flush=True)
    assert err_f < 1e-12 and err_e < 1e-12


def test_reduced_storage_from_eigen_initial_condition(cuda):
    """FP32 storage of the deviation from the ring: the stored deviation is
    the eigenfunction itself, so the total state matches FP64 to FP32
    rounding of the perturbation."""
    from vlasov1d2v_amd import Run
    kw = dict(k0=0.0886, ring_j=6, om=-0.349j, om_pc=10.0, delta_n=1e-3)
    a = Run(6, 3, 4, 4, dt=0.01, ic="eigen", **kw)
    b = Run(6, 3, 4, 4, dt=0.01, ic="eigen", storage="f32", **kw)
    fa = a.dist.arr
    fb = b.dist.total()
    err = (fa - fb).abs().max().item() / (fa - fb.new_tensor(b.dist.f0)[None, None]).abs().max().item()
    print(f"FP32 stored deviation vs FP64: {err:.2e}", flush=True)
    assert err < 1e-6
