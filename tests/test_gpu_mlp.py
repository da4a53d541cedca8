"""The learned load balancer on the GPU: MFMA forward and backward against
the fp64 CPU restatement (oracle/mlp_ref.py), training on logged samples,
the sample log itself, and invariance of every result under every
dispatch policy (the policy changes the schedule only)."""
import sys

import numpy as np
import pytest

import mlp_ref as M
import oracle as O
import spz

pytestmark = pytest.mark.gpu

# forward outputs and gradients: fp32 accumulation on the device vs fp64 on
# the CPU, same bf16 rounding points
FWD_ATOL = 2e-3
GRAD_REL_L2 = 1e-2


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def pdbs():
    say("PDB context for the MLP tests")
    s = spz.Solver(width=5, heuristic=spz.H_PDB666, frontier_target=256, log_samples=400_000)
    s.build_pdb()
    yield s
    s.close()


def rows(s, n, seed):
    boards = np.stack([O.walk(5, 120, seed + i) for i in range(n)])
    h = s.eval_h(boards)
    thr = (h + 2 * (np.arange(n) % 6)).astype(np.int32)
    return boards, thr, h


def test_forward_matches_reference(pdbs):
    pdbs.mlp_init(7)
    p = pdbs.mlp_get_params()
    b, thr, h = rows(pdbs, 300, 11)
    y = pdbs.mlp_forward(b, thr)
    yr, _ = M.forward(p, M.features(b, thr, h))
    assert np.max(np.abs(y - yr)) <= FWD_ATOL * max(1.0, float(np.max(np.abs(yr))))


def test_fused_prediction_matches_forward(pdbs):
    """The learned balancer's fused kernel (layer 1 as one-hot row sums plus a
    rank-1 threshold term, layers 2-3 in one MFMA pass per node) against the
    plain forward pass and the fp64 restatement, at all 16 thresholds."""
    pdbs.mlp_init(5)
    p = pdbs.mlp_get_params()
    b, thr, h = rows(pdbs, 96, 31)
    pred = pdbs.mlp_predict(b, thr)
    assert pred.shape == (96, 16)
    rb = np.repeat(b, 16, axis=0)
    rt = (np.repeat(thr, 16) + 2 * np.tile(np.arange(16), 96)).astype(np.int32)
    fwd = pdbs.mlp_forward(rb, rt).reshape(96, 16)
    yr, _ = M.forward(p, M.features(rb, rt, np.repeat(h, 16)))
    scale = max(1.0, float(np.max(np.abs(yr))))
    assert np.max(np.abs(pred - fwd)) <= FWD_ATOL * scale
    assert np.max(np.abs(pred - yr.reshape(96, 16))) <= FWD_ATOL * scale


def test_fused_prediction_15_puzzle():
    boards = np.stack([O.walk(4, 50, 9_100_000 + i) for i in range(40)])
    with spz.Solver(width=4) as s:
        s.mlp_init(3)
        p = s.mlp_get_params()
        h = s.eval_h(boards)
        thr = (h + 2).astype(np.int32)
        pred = s.mlp_predict(boards, thr)
        rb = np.repeat(boards, 16, axis=0)
        rt = (np.repeat(thr, 16) + 2 * np.tile(np.arange(16), 40)).astype(np.int32)
        yr, _ = M.forward(p, M.features(rb, rt, np.repeat(h, 16)))
    scale = max(1.0, float(np.max(np.abs(yr))))
    assert np.max(np.abs(pred.reshape(-1) - yr)) <= FWD_ATOL * scale


def test_gradient_matches_fp64_reference(pdbs):
    pdbs.mlp_init(9)
    p = pdbs.mlp_get_params()
    b, thr, h = rows(pdbs, 200, 21)
    t = np.random.default_rng(3).uniform(0, 25, len(b)).astype(np.float32)
    loss, g = pdbs.mlp_grad(b, thr, t)
    lref, gref = M.grad(p, M.features(b, thr, h), t)
    assert abs(loss - lref) <= 1e-3 * max(1.0, lref)
    blocks = [(M.OFF_W1, M.OFF_B1), (M.OFF_B1, M.OFF_W2), (M.OFF_W2, M.OFF_B2),
              (M.OFF_B2, M.OFF_W3), (M.OFF_W3, M.OFF_B3), (M.OFF_B3, M.N_PARAMS)]
    for a, z in blocks:
        err = np.linalg.norm(g[a:z] - gref[a:z]) / max(np.linalg.norm(gref[a:z]), 1e-12)
        assert err <= GRAD_REL_L2, (a, z, err)


def test_sample_log_equals_iteration_counts():
    """With no root split the only frontier node is the root, so its logged
    subtree count at threshold L is the count of iteration L."""
    boards = [O.walk(4, 40, 1_000_000 + i) for i in range(6)]
    with spz.Solver(width=4, frontier_target=1, log_samples=1000) as s:
        for b in boards:
            r = s.solve([b])[0]
            sb, st, sl = s.get_samples()
            assert len(sb) == r.n_iters - 1
            for bb, t, ld in zip(sb, st, sl):
                assert (bb == b).all()
                k = r.thresholds.index(int(t))
                assert int(ld) == r.counts[k]


@pytest.fixture(scope="module")
def trained(pdbs):
    """Samples logged by uniform-order solves of 24-puzzle walks, then GPU
    training; returns held-out data for the error check."""
    boards = [O.walk(5, 130, 6_000_000 + i) for i in range(24)]
    say("logging samples")
    pdbs.set_lb(spz.LB_UNIFORM)
    pdbs.solve(boards)
    sb, st, sl = pdbs.get_samples()
    say("samples", len(sb))
    target = np.log2(sl.astype(np.float64) + 1).astype(np.float32)
    rng = np.random.default_rng(5)
    idx = rng.permutation(len(sb))
    cut = int(0.8 * len(idx))
    tr, te = idx[:cut], idx[cut:]
    pdbs.mlp_init(1)
    losses = pdbs.mlp_train(sb[tr], st[tr], target[tr], epochs=12, batch=512, lr=1e-3, seed=2)
    say("epoch losses", list(np.round(losses, 3)))
    return sb[te], st[te], target[te], target[tr], losses


def test_training_reduces_loss_and_beats_uniform(pdbs, trained):
    """The trained predictor's held-out error is below that of the uniform
    predictor (one constant for every node, the training mean)."""
    xb, xt, yt, ytr, losses = trained
    assert losses[-1] < losses[0]
    pred = pdbs.mlp_forward(xb, xt)
    mse = float(np.mean((pred - yt) ** 2))
    mse_uniform = float(np.mean((ytr.mean() - yt) ** 2))
    say("held-out mse", mse, "uniform", mse_uniform)
    assert mse < mse_uniform


def test_results_invariant_under_lb_policy(pdbs, trained, ref_sets, ref_results):
    rows_ = ref_results["pdb_expand_nodes"]["rows"]
    idx = [i for i, r in enumerate(rows_) if r["nodes"] < 30_000_000][:20]
    boards = [ref_sets["yama24_50_hard_new"][i] for i in idx]
    boards += [O.walk(5, 120, 7_000_000 + i) for i in range(8)]
    ref = None
    for pol in (spz.LB_UNIFORM, spz.LB_PREVLOAD, spz.LB_LEARNED):
        pdbs.set_lb(pol)
        say("policy", pol)
        res = [(r.opt_len, r.counts) for r in pdbs.solve(boards)]
        for i, (ln, cnt) in zip(idx, res):
            assert (ln, sum(cnt)) == (rows_[i]["len"], rows_[i]["nodes"])
        if ref is None:
            ref = res
        assert res == ref
    # garbage parameters change nothing either
    pdbs.mlp_set_params(np.random.default_rng(0).normal(0, 5, spz.Solver.mlp_param_count()))
    assert [(r.opt_len, r.counts) for r in pdbs.solve(boards)] == ref
    pdbs.set_lb(spz.LB_UNIFORM)
