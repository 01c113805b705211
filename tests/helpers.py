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


def reference_lgl(n):
    """Lobatto nodes and weights as the reference's tables give them
    (main/basis.py:9-57): the decimal literals of tests/golden/lgl_literals.json
    for 7..10 nodes, the closed forms (float64 evaluation) below that."""
    import json
    import os
    if n >= 7:
        path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "lgl_literals.json")
        with open(path) as fh:
            lit = json.load(fh)
        return (np.array([float(t) for t in lit["nodes"][str(n)]]),
                np.array([float(t) for t in lit["weights"][str(n)]]))
    s7 = np.sqrt(7)
    closed = {
        2: ([-1, 1], [1, 1]),
        3: ([-1, 0, 1], [1 / 3, 4 / 3, 1 / 3]),
        4: ([-1, -np.sqrt(1 / 5), np.sqrt(1 / 5), 1], [1 / 6, 5 / 6, 5 / 6, 1 / 6]),
        5: ([-1, -np.sqrt(3 / 7), 0, np.sqrt(3 / 7), 1], [1 / 10, 49 / 90, 32 / 45, 49 / 90, 1 / 10]),
        6: ([-1, -np.sqrt(1 / 3 + 2 * s7 / 21), -np.sqrt(1 / 3 - 2 * s7 / 21), np.sqrt(1 / 3 - 2 * s7 / 21),
             np.sqrt(1 / 3 + 2 * s7 / 21), 1],
            [1 / 15, (14 - s7) / 30, (14 + s7) / 30, (14 + s7) / 30, (14 - s7) / 30, 1 / 15]),
    }
    x, w = closed[n]
    return np.array(x, dtype=np.float64), np.array(w, dtype=np.float64)


def reference_way_up(n):
    """The volume operator evaluated as main/basis.py:176-188 evaluates it:
    up[i, j] = w_j sum_s (2s+1)/2 P_s(x_i) P_s'(x_j) with scipy's
    ``legendre(s)`` polynomial objects (their coefficients come from roots,
    so odd/even terms are not exactly zero) on the reference's node tables,
    then entries below 1e-10 cleared.  Not exactly centro-antisymmetric."""
    import scipy.special as sp
    x, w = reference_lgl(n)
    up = np.zeros((n, n))
    for i in range(n):
        for j in range(n):
            up[i, j] = w[j] * sum((2 * s + 1) / 2 * sp.legendre(s)(x[i]) * sp.legendre(s).deriv()(x[j])
                                  for s in range(n))
    up[np.abs(up) < 1.0e-10] = 0
    return up
