"""The two benchmark workloads on the GPU against committed oracle
fingerprints (tests/golden/*_oracle.json, written by
tests/golden/make_oracle_fingerprints.py from the NumPy oracle).

Tolerance: 1e-12 relative L2 on the field-energy series, on the interior L2
norm of f and on f at 4096 seeded interior positions.
"""
import json
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

pytestmark = pytest.mark.gpu


def _fixture(name):
    path = os.path.join(HERE, "golden", f"{name}_oracle.json")
    if not os.path.exists(path):
        pytest.skip(f"{path} not generated")
    with open(path) as fh:
        return json.load(fh)


@pytest.mark.parametrize("name", ["ring16", "ring64"])
def test_workload_matches_oracle_fingerprint(cuda, name):
    import torch
    from make_oracle_fingerprints import sample_index
    from vlasov1d2v_amd import Run

    fx = _fixture(name)
    nx, nu, nv = fx["elements"]
    print(f"workload {name}: n={fx['n']} elements={fx['elements']} steps={fx['steps']}", flush=True)
    run = Run(fx["n"], nx, nu, nv, k0=fx["k0"], b_field=fx["b_field"], ring_j=fx["ring_j"],
              amp1=fx["amp1"], amps=tuple(fx["amps"]), seed=fx["seed"], dt=fx["dt"], device=cuda)
    assert np.array_equal(run.phases, np.asarray(fx["phases"]))
    st = run.run(fx["steps"])
    torch.cuda.synchronize()
    e_ref = np.asarray(fx["field_energy"])
    assert np.linalg.norm(st.field_energy - e_ref) / np.linalg.norm(e_ref) < 1e-12
    f = run.dist.arr[1:-1, :, 1:-1, :, 1:-1, :]
    l2 = float(torch.linalg.vector_norm(f).item())
    assert abs(l2 - fx["f_interior_l2"]) < 1e-12 * fx["f_interior_l2"]
    idx = torch.from_numpy(sample_index(tuple(f.shape), fx["sample_seed"], len(fx["sample_values"])))
    got = f.reshape(-1)[idx.to(cuda)].cpu().numpy()
    want = np.asarray(fx["sample_values"])
    assert np.linalg.norm(got - want) / np.linalg.norm(want) < 1e-12


@pytest.mark.parametrize("name", ["ring16", "ring64"])
def test_workload_checkpoints_match_oracle(cuda, name):
    """The same run checked at every checkpoint of the fixture (ring64: steps
    2, 20, 50, 100 and the configured 200): f's interior L2 norm, the charge
    density, f at the seeded positions, and the field-energy segment since
    the previous checkpoint, all to 1e-12."""
    import torch
    from make_oracle_fingerprints import sample_index
    from vlasov1d2v_amd import Run

    fx = _fixture(name)
    if "checkpoints" not in fx:
        pytest.skip("fixture without checkpoints")
    nx, nu, nv = fx["elements"]
    run = Run(fx["n"], nx, nu, nv, k0=fx["k0"], b_field=fx["b_field"], ring_j=fx["ring_j"],
              amp1=fx["amp1"], amps=tuple(fx["amps"]), seed=fx["seed"], dt=fx["dt"], device=cuda)
    e_ref = np.asarray(fx["field_energy"])
    done = 0
    for cp in fx["checkpoints"]:
        st = run.run(cp["step"] - done)
        torch.cuda.synchronize()
        seg = e_ref[done:cp["step"] + 1]
        # each loop records its starting energy, then one per step
        e_err = np.linalg.norm(np.asarray(st.field_energy)[-len(seg):] - seg) / np.linalg.norm(seg)
        done = cp["step"]
        f = run.dist.arr[1:-1, :, 1:-1, :, 1:-1, :]
        l2 = float(torch.linalg.vector_norm(f).item())
        rho = run.dist.moment_zero().cpu().numpy().reshape(-1)
        rho_ref = np.asarray(cp["density"])
        idx = torch.from_numpy(sample_index(tuple(f.shape), fx["sample_seed"], len(cp["sample_values"])))
        got = f.reshape(-1)[idx.to(cuda)].cpu().numpy()
        want = np.asarray(cp["sample_values"])
        errs = (e_err, abs(l2 - cp["f_interior_l2"]) / cp["f_interior_l2"],
                np.linalg.norm(rho - rho_ref) / np.linalg.norm(rho_ref),
                np.linalg.norm(got - want) / np.linalg.norm(want))
        print(f"{name} step {done}: energy {errs[0]:.1e} |f| {errs[1]:.1e} rho {errs[2]:.1e} "
              f"samples {errs[3]:.1e}", flush=True)
        assert max(errs) < 1e-12, (done, errs)
