"""Ensembles of runs with different n, grids and physics stepped together
(one launch per kernel kind and RK stage for all runs).

Every run of an ensemble executes the single-run kernels' statements; the
compiler may fuse multiply-adds differently in the two kernels, so a run
agrees with the single-run resident loop to rounding (FP64: 1e-12 relative
L2 on f and on the energy series; FP32: 1e-5).  FP64 members are also
checked against the NumPy oracle at 1e-12.  Within the ensemble the stage
schedule (x-segment length, persistent grid, item order) changes nothing:
results are bit for bit the same, and graph replays reproduce eager steps
bit for bit.
"""
import numpy as np
import pytest

from helpers import interior, rel_l2

pytestmark = pytest.mark.gpu

# (n, nx, nu, nv, ring_j, k0, b_field, seed): every n of the ensemble
# kernels, partial v-chunks (nv not a multiple of the tile width), nx = 2
MEMBERS = [(6, 4, 5, 6, 6, 0.1, 0.1, 0), (7, 3, 4, 4, 3, 0.15, 0.08, 1), (8, 5, 3, 5, 1, 0.07, 0.19, 2),
           (9, 2, 4, 3, 8, 0.2, 0.05, 3), (6, 3, 3, 9, 5, 0.05, 0.12, 4)]
DT = 0.004


def _members():
    from vlasov1d2v_amd import Member
    return [Member(n=a, nx=b, nu=c, nv=d, ring_j=e, k0=f, b_field=g, seed=h) for a, b, c, d, e, f, g, h in MEMBERS]


def _single(m, storage, steps):
    from vlasov1d2v_amd import Run
    run = Run(m.n, m.nx, m.nu, m.nv, k0=m.k0, b_field=m.b_field, ring_j=m.ring_j, seed=m.seed, dt=DT,
              storage=storage)
    st = run.run(steps)
    return run, np.asarray(st.field_energy)


@pytest.mark.parametrize("storage", ["f64", "f32"])
def test_ensemble_matches_single_runs(cuda, storage):
    import torch
    from vlasov1d2v_amd import Ensemble
    steps = 4
    members = _members()
    ens = Ensemble(members, storage=storage, device=cuda, dt=DT)
    print(f"{storage}: schedule {ens.schedule()}", flush=True)
    energies = ens.run(steps)
    ens.finish()
    torch.cuda.synchronize()
    tol = 1e-12 if storage == "f64" else 1e-5
    for r, m in enumerate(members):
        run, e_ref = _single(m, storage, steps)
        err_e = rel_l2(energies[:, r], e_ref)
        err_f = rel_l2(interior(ens.runs[r].dist.total().cpu().numpy()), interior(run.dist.total().cpu().numpy()))
        err_E = rel_l2(ens.runs[r].elliptic.electric_field.cpu().numpy(), run.elliptic.electric_field.cpu().numpy())
        print(f"{storage}: member {m}: rel-L2 f {err_f:.2e}, energy {err_e:.2e}, E {err_E:.2e}", flush=True)
        assert energies[0, r] == e_ref[0]
        assert err_f < tol and err_e < tol and err_E < tol, (m, err_f, err_e, err_E)


def test_ensemble_f64_members_match_oracle(cuda):
    import oracle
    from vlasov1d2v_amd import Ensemble
    steps = 3
    members = _members()[:3]
    ens = Ensemble(members, storage="f64", device=cuda, dt=DT)
    energies = ens.run(steps)
    ens.finish()
    for r, m in enumerate(members):
        case = oracle.Case(n=m.n, nx=m.nx, nu=m.nu, nv=m.nv, k0=m.k0, B=m.b_field, ring_j=m.ring_j, seed=m.seed)
        st = case.run(steps, DT)
        err_f = rel_l2(interior(ens.runs[r].dist.arr.cpu().numpy()), interior(case.dist.arr))
        err_e = rel_l2(energies[:, r], st.field_energy)
        print(f"member {m}: rel-L2 f {err_f:.2e}, energy {err_e:.2e}", flush=True)
        assert err_f < 1e-12 and err_e < 1e-12, (m, err_f, err_e)


@pytest.mark.parametrize("work,grid,order", [(500, 0, 0), (3000, 7, 1), (100000, 3, 0), (9000, 64, 1)])
def test_ensemble_schedule_does_not_change_results(cuda, work, grid, order):
    import torch
    from vlasov1d2v_amd import Ensemble
    steps = 3
    members = _members()
    a = Ensemble(members, storage="f32", device=cuda, dt=DT)
    b = Ensemble(members, storage="f32", device=cuda, dt=DT)
    b.set_schedule(work, grid, order)
    print(f"schedule {b.schedule()}", flush=True)
    ea = a.run(steps)
    eb = b.run(steps)
    np.testing.assert_array_equal(ea, eb)
    a.finish()
    b.finish()
    torch.cuda.synchronize()
    for ra, rb in zip(a.runs, b.runs):
        assert torch.equal(ra.dist.arr, rb.dist.arr)


def test_ensemble_graph_and_autotune_match_eager(cuda):
    import torch
    from vlasov1d2v_amd import Ensemble
    members = _members()
    eager = Ensemble(members, storage="f32", device=cuda, dt=DT)
    tuned = Ensemble(members, storage="f32", device=cuda, dt=DT)
    sched = tuned.autotune(reps=1)
    print(f"autotuned schedule {sched}; {len(tuned.tuning)} candidates", flush=True)
    assert len(tuned.tuning) == 48 and all(t["ms"] > 0 for t in tuned.tuning)
    e_eager = eager.run(9)
    e_graph = tuned.run(9, graph_steps=4)     # 1 eager + 2 replays of 4
    scratch = torch.ones(1 << 16, device=cuda)
    scratch.mul_(1.0001)
    np.testing.assert_array_equal(e_eager, e_graph)
    # more replays of the same graph, with eager work in between
    g = tuned.capture(2)
    e_eager2 = np.stack([_step_energy(eager) for _ in range(4)])
    rows = []
    for rep in range(2):
        g.replay()
        rows.append(g.energies.cpu().numpy())
        scratch.add_(1.0)
        torch.cuda.synchronize()
    np.testing.assert_array_equal(np.concatenate(rows), e_eager2)
    eager.finish()
    tuned.finish()
    torch.cuda.synchronize()
    for ra, rb in zip(eager.runs, tuned.runs):
        assert torch.equal(ra.dist.arr, rb.dist.arr)


def _step_energy(ens):
    import torch
    e = torch.empty(len(ens.runs), dtype=torch.float64, device=ens.device)
    ens.step(e)
    return e.cpu().numpy()


@pytest.mark.parametrize("graph_steps", [0, 2])
def test_ensemble_adaptive_dt_matches_single_device_dt_loops(cuda, graph_steps):
    """dt="adaptive": every run takes the CFL step of its own field after each
    step (main/timestep.py:184,231-242), as a single run's device-dt loop does."""
    import torch
    from vlasov1d2v_amd import Ensemble, Run
    steps = 5
    members = _members()
    ens = Ensemble(members, storage="f64", device=cuda, dt="adaptive")
    energies = ens.run(steps, graph_steps=graph_steps)
    clocks = ens.clocks()
    ens.finish()
    torch.cuda.synchronize()
    for r, m in enumerate(members):
        run = Run(m.n, m.nx, m.nu, m.nv, k0=m.k0, b_field=m.b_field, ring_j=m.ring_j, seed=m.seed, dt=None)
        st = run.stepper
        st.main_loop(run.dist, run.basis, run.elliptic, run.grids, run.flux, steps=steps, device_dt=True)
        err_e = rel_l2(energies[:, r], np.asarray(st.field_energy))
        err_f = rel_l2(interior(ens.runs[r].dist.total().cpu().numpy()), interior(run.dist.total().cpu().numpy()))
        print(f"member {m}: time {clocks[r, 0]!r} vs {st.time!r}, dt {clocks[r, 1]!r} vs {st.dt!r}, "
              f"rel-L2 f {err_f:.2e}, energy {err_e:.2e}", flush=True)
        assert abs(clocks[r, 0] - st.time) <= 1e-12 * st.time
        assert abs(clocks[r, 1] - st.dt) <= 1e-12 * st.dt
        assert clocks[r, 1] != ens.dts[r]           # the step changed with the field
        assert err_f < 1e-12 and err_e < 1e-12, (m, err_f, err_e)
