"""Pins the CPU restatement of the MLP (oracle/mlp_ref.py): its analytic
gradient equals central finite differences (fp64, no rounding), and the
bf16-rounded variant the device follows stays close to the exact one."""
import numpy as np

import mlp_ref as M


def small_problem(seed=0, n=12):
    rng = np.random.default_rng(seed)
    p = np.zeros(M.N_PARAMS)
    p[M.OFF_W1:M.OFF_B1] = rng.normal(0, 0.08, M.IN * M.H)
    p[M.OFF_B1:M.OFF_W2] = rng.normal(0.05, 0.05, M.H)
    p[M.OFF_W2:M.OFF_B2] = rng.normal(0, 0.1, M.H * M.H)
    p[M.OFF_B2:M.OFF_W3] = rng.normal(0.05, 0.05, M.H)
    p[M.OFF_W3:M.OFF_B3] = rng.normal(0, 0.1, M.H)
    p[M.OFF_B3] = 3.0
    boards = np.stack([rng.permutation(25) for _ in range(n)])
    h = rng.integers(30, 60, n)
    thr = h + 2 * rng.integers(0, 8, n)
    x = M.features(boards, thr, h)
    t = rng.uniform(0, 20, n)
    return p, x, t


def test_features_layout():
    b = np.arange(25)[None, :]
    x = M.features(b, [50], [40])
    assert x.shape == (1, 640)
    assert x[0, :625].sum() == 25
    for c in range(25):
        assert x[0, c * 25 + c] == 1.0  # tile c at cell c
    assert x[0, 625] == 40 / 32 and x[0, 626] == 10 / 32 and x[0, 627:].sum() == 0


def test_gradient_matches_finite_differences():
    p, x, t = small_problem()
    loss, g = M.grad(p, x, t, rounding=False)
    rng = np.random.default_rng(1)
    idx = np.concatenate([rng.integers(M.OFF_W1, M.OFF_B1, 12), rng.integers(M.OFF_B1, M.OFF_W2, 4),
                          rng.integers(M.OFF_W2, M.OFF_B2, 12), rng.integers(M.OFF_B2, M.OFF_W3, 4),
                          rng.integers(M.OFF_W3, M.OFF_B3, 6), [M.OFF_B3]])
    # only coordinates that affect the loss through active units are informative;
    # every sampled coordinate is checked anyway (inactive ones have zero gradient)
    eps = 1e-6
    for i in idx:
        q = p.copy()
        q[i] += eps
        lp = M.grad(q, x, t, rounding=False)[0]
        q[i] -= 2 * eps
        lm = M.grad(q, x, t, rounding=False)[0]
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-5 * max(1.0, abs(fd)), (i, fd, g[i])


def test_rounded_gradient_close_to_exact():
    p, x, t = small_problem(3, 64)
    _, ge = M.grad(p, x, t, rounding=False)
    _, gr = M.grad(p, x, t, rounding=True)
    cos = ge @ gr / np.linalg.norm(ge) / np.linalg.norm(gr)
    assert cos > 0.999
