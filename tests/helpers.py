"""Shared builders for the parity tests: one oracle case and one GPU run on
the same seeded inputs."""
import numpy as np

import oracle


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def rel_max(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def interior(arr):
    return arr[1:-1, :, 1:-1, :, 1:-1, :]


def make_pair(n, nx, nu, nv, seed=0, k0=0.1, b_field=0.1, ring_j=6, dt=0.01, **kw):
    """(oracle.Case, vlasov1d2v_amd.Run) built from the same parameters."""
    from vlasov1d2v_amd import Run
    case = oracle.Case(n=n, nx=nx, nu=nu, nv=nv, k0=k0, B=b_field, ring_j=ring_j, seed=seed, **kw)
    run = Run(n, nx, nu, nv, k0=k0, b_field=b_field, ring_j=ring_j, seed=seed, dt=dt, **kw)
    return case, run


def random_state(shape, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(shape)


def ulp_distance(a, b):
    """Distance in units in the last place between two arrays of the same
    IEEE storage type (torch tensors of float32 / bfloat16 / float16):
    the bit patterns are mapped to a monotonic integer order."""
    import torch
    ity = {torch.float32: torch.int32, torch.bfloat16: torch.int16, torch.float16: torch.int16}[a.dtype]
    bits = 31 if ity == torch.int32 else 15

    def key(t):
        i = t.contiguous().view(ity).cpu().numpy().astype(np.int64)
        return np.where(i >= 0, i, -(i & ((1 << bits) - 1)))
    return np.abs(key(a) - key(b))


def significant_ulp_distance(a, b, floor=1e-3):
    """ulp_distance over the entries whose magnitude is at least `floor`
    times the array maximum.  Near zero (values that cancelled down from
    larger intermediates) a rounding flip of an intermediate shows up as many
    ulps of a tiny number; those entries are covered by a norm bound."""
    import torch
    big = torch.maximum(a.float().abs(), b.float().abs())
    mask = (big >= floor * float(a.float().abs().max())).cpu().numpy()
    return ulp_distance(a, b)[mask]
