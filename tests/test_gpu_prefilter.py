"""Prefilter network on the GPU vs its FP64 CPU restatement (oracle/pbw_mlp_ref.py).

The GPU runs BF16 operands on the matrix cores with FP32 accumulation; the
reference rounds at the same points in FP64.  A BF16 rounding of an
intermediate can land on the other side of a boundary when the FP32 and FP64
accumulations differ in the last bits, so comparisons use a tolerance a wrong
operand layout, index or sign would exceed by orders of magnitude.
"""
import numpy as np
import pytest

import probebwa_amd as P
from oracle import pbw_mlp_ref as R
from probebwa_amd import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    contigs = synth.small_genome(71, n_contigs=1, length=5000)
    g = P.Genome()
    for name, s in contigs:
        g.add_contig(name, bytes(s))
    return P.Context(g)


def rand_rows(rng, n):
    x = P.to_bf16(rng.normal(0.0, 1.0, (n, P.MLP_INPUTS)).astype(np.float32))
    return x, P.from_bf16(x).astype(np.float64)


def test_mlp_forward_matches_fp64_reference(ctx):
    rng = np.random.default_rng(5)
    ctx.mlp_init(seed=3)
    theta = ctx.mlp_params
    th = theta.copy()
    th[R.OFF_B1:R.OFF_W2] = rng.normal(0, 0.1, R.H)  # nonzero biases exercise the bias path
    th[R.OFF_B2:R.OFF_W3] = rng.normal(0, 0.1, R.H)
    th[R.OFF_B3] = 0.25
    ctx.mlp_params = th
    assert np.array_equal(ctx.mlp_params, th)
    xb, x = rand_rows(rng, 1000)  # not a multiple of 32: ragged last tile
    z = ctx.mlp_forward(xb)
    zr, _ = R.forward(th, x)
    err = np.abs(z - zr)
    assert np.max(err) < 2e-2 * max(1.0, np.max(np.abs(zr))), np.max(err)
    assert np.mean(err) < 1e-3 * max(1.0, np.mean(np.abs(zr)))


@pytest.mark.parametrize("n,pos_weight", [(256, 1.0), (77, 4.0)])
def test_mlp_gradient_matches_fp64_reference(ctx, n, pos_weight):
    rng = np.random.default_rng(11 + n)
    ctx.mlp_init(seed=7)
    th = ctx.mlp_params
    th[R.OFF_B1:R.OFF_W2] = rng.normal(0, 0.1, R.H)
    th[R.OFF_B2:R.OFF_W3] = rng.normal(0, 0.1, R.H)
    ctx.mlp_params = th
    xb, x = rand_rows(rng, n)
    y = (rng.random(n) < 0.3).astype(np.float32)
    g, loss = ctx.mlp_gradient(xb, y, pos_weight)
    gr, lr = R.gradient(th, x, y.astype(np.float64), pos_weight)
    assert abs(loss - lr) < 1e-3 * max(1.0, abs(lr))
    for name, a, b in (("W1", R.OFF_W1, R.OFF_B1), ("b1", R.OFF_B1, R.OFF_W2), ("W2", R.OFF_W2, R.OFF_B2),
                       ("b2", R.OFF_B2, R.OFF_W3), ("W3", R.OFF_W3, R.OFF_B3), ("b3", R.OFF_B3, R.NP)):
        d = np.linalg.norm(g[a:b] - gr[a:b])
        ref = np.linalg.norm(gr[a:b])
        assert d <= 2e-2 * ref + 1e-7, (name, d, ref)
    # a mistake that swaps an operand or drops a term fails by far more: the
    # gradient of a wrong layer is nearly orthogonal to the right one
    cos = float(g @ gr / (np.linalg.norm(g) * np.linalg.norm(gr)))
    assert cos > 0.9999, cos


def test_mlp_training_learns_a_linear_rule(ctx):
    rng = np.random.default_rng(21)
    xb, x = rand_rows(rng, 20000)
    y = ((x[:, 0] + 0.5 * x[:, 1] - x[:, 2]) > 0.3).astype(np.float32)
    ctx.mlp_init(seed=5)
    z0 = ctx.mlp_forward(xb[:4000])
    losses = ctx.mlp_train(xb[:16000], y[:16000], epochs=15, batch=512, lr=1e-3, seed=9)
    assert losses[-1] < 0.5 * losses[0], losses
    z = ctx.mlp_forward(xb[16000:])
    acc = np.mean((z > 0) == (y[16000:] > 0.5))
    acc0 = np.mean((z0 > 0) == (y[:4000] > 0.5))
    assert acc > 0.95 and acc > acc0 + 0.2, (acc, acc0)


def test_adam_steps_match_reference(ctx):
    """Two training steps (one batch each) vs Adam applied to the GPU's own gradient.
    The gradient of a step is summed in a shuffled row order, so it can differ from
    pbw_mlp_gradient's in the last FP32 bits; Adam's first steps are about lr*sign(g),
    so only an element whose gradient is ~0 may move by up to 2 lr."""
    rng = np.random.default_rng(31)
    n, lr = 512, 1e-3
    xb, _ = rand_rows(rng, n)
    y = (rng.random(n) < 0.3).astype(np.float32)
    ctx.mlp_init(seed=13)  # resets the Adam moments and step count
    th = ctx.mlp_params.astype(np.float64)
    m = np.zeros(P.MLP_PARAMS)
    v = np.zeros(P.MLP_PARAMS)
    for t in (1, 2):
        g, _ = ctx.mlp_gradient(xb, y, 2.0)
        ctx.mlp_train(xb, y, epochs=1, batch=n, lr=lr, pos_weight=2.0, seed=t)
        th, m, v = R.adam_step(th, m, v, g.astype(np.float64), t, lr=lr)
        got = ctx.mlp_params.astype(np.float64)
        err = np.abs(got - th)
        tol = 1e-6 * np.abs(th) + 1e-3 * lr
        assert np.mean(err > tol) < 1e-3, (t, np.mean(err > tol))
        assert np.all(err <= 2 * t * lr + 1e-6 * np.abs(th)), (t, err.max())
        th = got  # continue from the GPU's parameters


# ---------------------------------------------------------------------------
# prefilter in the mapping path: features from clustering, labels from plain DP
# ---------------------------------------------------------------------------
sam_identical = P.primary_identical


@pytest.fixture(scope="module")
def probe_genome():
    contigs = synth.genome_probe_se_2mbp()
    g = P.Genome()
    for name, s in contigs:
        g.add_contig(name, s.tobytes())
    (train, _) = synth.reads_probe_se(contigs, n=12000, seed=101)
    (held, _) = synth.reads_probe_se(contigs, n=6000, seed=11)
    return g, P.ReadBatch.from_arrays(*train), P.ReadBatch.from_arrays(*held)


def test_prefilter_dataset_rows_and_labels(probe_genome):
    g, train, _ = probe_genome
    ctx = P.Context(g)
    sub = train.subset(0, 2000)
    x, y, rank = ctx.prefilter_dataset(sub, label="loglk", margin=40.0)
    plain = ctx.map_se(sub)
    # one row per aligned candidate; the best candidate of every mapped read is a positive
    assert len(x) == len(y) == len(rank) and len(x) >= int((plain.rec["pos"] > 0).sum())
    assert rank.max() < 20 and set(np.unique(y)) <= {0.0, 1.0}
    assert y.sum() >= (plain.rec["pos"] > 0).sum()
    f = P.from_bf16(x)
    assert np.all(np.isfinite(f))
    # feature 9 is rank / 20 rounded to bf16
    assert np.all(f[:, 9] == P.from_bf16(P.to_bf16(rank.astype(np.float32) / np.float32(20))))
    assert np.all((f[:, 10] == 1) == (rank == 0))  # feature 10 flags the first candidate


def test_prefilter_label_rules(probe_genome):
    g, train, _ = probe_genome
    ctx = P.Context(g)
    sub = train.subset(0, 1500)
    x, y_nats, rank = ctx.prefilter_dataset(sub, label="loglk", margin=40.0)
    x2, y_score, rank2 = ctx.prefilter_dataset(sub, label="score", margin=10.0)
    assert np.array_equal(x, x2) and np.array_equal(rank, rank2)  # same rows, other labels
    _, y_all, _ = ctx.prefilter_dataset(sub, label="score", margin=1e9)
    _, y_tight, _ = ctx.prefilter_dataset(sub, label="score", margin=0.0)
    assert y_all.min() == 1.0
    # rows of a read are contiguous and start at rank 0: every read has a best, so
    # at least one positive at margin 0, and widening the margin only adds positives
    starts = np.nonzero(rank == 0)[0]
    assert np.all(np.add.reduceat(y_tight, starts) >= 1)
    assert np.all(y_score >= y_tight) and np.all(y_all >= y_score)
    assert y_nats.sum() > y_tight.sum()
    with pytest.raises(KeyError):
        ctx.prefilter_dataset(sub, label="depth")


def test_prefilter_path_is_sam_identical_and_skips_dp(probe_genome):
    g, train, held = probe_genome
    ctx = P.Context(g)
    rep = ctx.prefilter_train(train, label="loglk", margin=40.0, recall=0.9999, epochs=20, batch=4096)
    assert rep["loss_last"] < rep["loss_first"] and rep["val_recall"] >= 0.999
    plain = ctx.map_se(held)
    cells_plain = ctx.stats()["dp_cells"]
    ctx.set_prefilter(enabled=1, audit_period=16)
    pre = ctx.map_se(held)
    st = ctx.stats()
    same = sam_identical(plain, pre)
    assert same.mean() >= 0.999, same.mean()
    assert st["pf_jobs_aligned"] < 0.9 * st["pf_jobs"], st
    assert st["dp_cells"] < 0.9 * cells_plain
    assert st["pf_audit_reads"] > 0 and st["pf_audit_mismatches"] == 0
    # posteriors of identical records agree to the FP64 tolerance of the plain path
    pp, pq = plain.rec["posterior"][same], pre.rec["posterior"][same]
    assert np.all(np.abs(pp - pq) <= 1e-12 * np.maximum(np.abs(pp), 1e-300) + 1e-300) or np.mean(pp == pq) > 0.999


def test_trained_network_beats_the_rank_rule(probe_genome):
    """What the network adds: at the same top-k and fallbacks, the rank rule alone
    (threshold above every score) skips more but loses reads; the trained network
    keeps SAM identity with a large skip."""
    g, train, held = probe_genome
    ctx = P.Context(g)
    ctx.prefilter_train(train, label="loglk", margin=40.0, recall=0.9999, epochs=20, batch=4096)
    plain = ctx.map_se(held)
    thr = ctx.prefilter.threshold
    ctx.set_prefilter(enabled=1, audit_period=0)
    net = sam_identical(plain, ctx.map_se(held)).mean()
    net_jobs = ctx.stats()["pf_jobs_aligned"]
    ctx.set_prefilter(enabled=1, threshold=3.0e38, audit_period=0)
    rank_only = sam_identical(plain, ctx.map_se(held)).mean()
    ctx.set_prefilter(enabled=1, threshold=thr)
    assert net >= 0.999
    assert net > rank_only, (net, rank_only)
    assert net_jobs > 0
