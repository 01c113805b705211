"""The reference's stabilized central-flux DG Poisson solve
(main/elliptic.py:24-115) and linear SSP-RK update (main/timestep.py:189-229)
on the GPU against the oracle's restatement."""
import numpy as np
import pytest

from helpers import interior, make_pair, rel_l2, rel_max

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,nx", [(4, 6), (6, 8), (8, 5)])
@pytest.mark.parametrize("anti_alias", [True, False])
def test_central_flux_poisson_matches_oracle(cuda, n, nx, anti_alias):
    import torch
    case, run = make_pair(n, nx, 2, 2)
    gx_o, b_o = case.grids.x, case.basis.b1
    gx, b = run.grids.x, run.basis.b1
    rng = np.random.default_rng(7)
    k = gx_o.k1
    x = gx_o.arr[1:-1]
    cd = np.sin(k * x) + 0.3 * np.cos(2 * k * x + 0.4) + 1e-3 * rng.standard_normal(x.shape)
    case.elliptic.build_central_flux_operator(gx_o, b_o)
    case.elliptic.invert()
    case.elliptic.poisson(cd, gx_o, b_o, anti_alias=anti_alias)
    run.elliptic.build_central_flux_operator(gx, b)
    # the two Lobatto bases agree to a few ulp, so the assembled operators do too
    assert rel_max(run.elliptic.central_flux_operator, case.elliptic.central_flux_operator) < 1e-13
    run.elliptic.invert()
    run.elliptic.poisson(torch.from_numpy(cd).to(cuda), gx, b, anti_alias=anti_alias)
    e_E = rel_max(run.elliptic.electric_field.cpu().numpy(), case.elliptic.electric_field)
    e_p = rel_max(run.elliptic.potential.cpu().numpy(), case.elliptic.potential)
    print(f"n={n} nx={nx} anti_alias={anti_alias}: E {e_E:.2e}, phi {e_p:.2e}", flush=True)
    assert e_E < 1e-12 and e_p < 1e-12


def test_linear_ssp_rk_update_matches_oracle(cuda):
    import oracle
    from vlasov1d2v_amd import Stepper
    n, dt, steps = 5, 0.01, 2
    case, run = make_pair(n, 3, 4, 4)
    case.elliptic.build_central_flux_operator(case.grids.x, case.basis.b1)
    case.elliptic.invert()
    run.elliptic.build_central_flux_operator(run.grids.x, run.basis.b1)
    run.elliptic.invert()
    st_o = oracle.Stepper(dt)
    st = Stepper(3, n, 1e30, 1e30, linear=True, dt=dt)
    for _ in range(steps):
        st_o.ssp_rk_update(case.dist, case.basis, case.elliptic, case.grids, case.flux)
        st.ssp_rk_update(run.dist, run.basis, run.grids, run.elliptic, run.flux)
    err_f = rel_l2(interior(run.dist.arr.cpu().numpy()), interior(case.dist.arr))
    err_e = rel_max(run.elliptic.electric_field.cpu().numpy(), case.elliptic.electric_field)
    print(f"rel-L2 f {err_f:.2e}, E {err_e:.2e}", flush=True)
    assert err_f < 1e-12 and err_e < 1e-11
