"""HIP-graph capture of the fixed-dt SSP-RK3 step loop against eager steps.

The captured sequence is the same kernel sequence as the eager one, so
replays must reproduce eager stepping bit for bit.  Replays alternate with
eager work and synchronisations, and every replay is checked.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("storage,n,nx,nu,nv", [("f64", 6, 4, 6, 5), ("bf16", 9, 3, 4, 5), ("f32", 5, 3, 4, 4)])
def test_graph_replays_match_eager_steps(cuda, storage, n, nx, nu, nv):
    import torch
    from vlasov1d2v_amd import Run
    K = 3
    eager = Run(n, nx, nu, nv, dt=0.004, storage=storage)
    graphed = Run(n, nx, nu, nv, dt=0.004, storage=storage)
    for r in (eager, graphed):
        r.resident_begin()
    e_eager = torch.zeros(1 + 3 * K, dtype=torch.float64, device=cuda)
    for k in range(1 + 3 * K):
        eager.step(energy_out=e_eager[k:k + 1])
    e_graph = torch.zeros(1 + 3 * K, dtype=torch.float64, device=cuda)
    graphed.step(energy_out=e_graph[0:1])
    g = graphed.capture(K)
    scratch = torch.ones(1 << 16, device=cuda)
    for rep in range(3):
        g.replay()
        e_graph[1 + rep * K:1 + (rep + 1) * K].copy_(g.energies)
        # eager work and a synchronisation between replays
        scratch.mul_(1.0001).add_(1.0)
        torch.cuda.synchronize()
        print(f"{storage}: replay {rep} energies {g.energies.cpu().numpy()}", flush=True)
        np.testing.assert_array_equal(e_graph[1 + rep * K:1 + (rep + 1) * K].cpu().numpy(),
                                      e_eager[1 + rep * K:1 + (rep + 1) * K].cpu().numpy())
    for r in (eager, graphed):
        r.resident_end()
    torch.cuda.synchronize()
    assert torch.equal(graphed.dist.arr, eager.dist.arr)
    assert torch.equal(graphed.elliptic.electric_field, eager.elliptic.electric_field)


def test_main_loop_with_graph_matches_eager_loop(cuda):
    from vlasov1d2v_amd import Run
    a = Run(6, 4, 5, 5, dt=0.005)
    b = Run(6, 4, 5, 5, dt=0.005)
    sa = a.run(11)
    sb = b.run(11, graph_steps=4)        # 1 eager + 2 replays of 4 + 2 eager
    np.testing.assert_array_equal(sa.field_energy, sb.field_energy)
    assert sb.steps_counter == 11 and len(sb.field_energy) == 12
    np.testing.assert_array_equal(a.dist.arr.cpu().numpy(), b.dist.arr.cpu().numpy())
