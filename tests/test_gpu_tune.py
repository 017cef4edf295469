"""Launch shapes (pbw_tune) and the autotuner change the speed of a mapping call,
never its records: every setting must give the same bytes as the default one,
and the default one matches the oracle."""
import numpy as np
import pytest

import probebwa_amd as P
from oracle import pbw_oracle as O
from probebwa_amd import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup():
    contigs = synth.small_genome(41, n_contigs=3, length=80_000)
    gp, go = P.Genome(), O.Genome()
    for name, s in contigs:
        gp.add_contig(name, bytes(s))
        go.add_contig(name, bytes(s))
    # lengths 30..420: the quad kernel, its length split and the one-job-per-wave
    # kernel (reads over 320 bases) all take part
    flat = synth.sample_reads(contigs, 3000, seed=43, lmin=30, lmax=420)
    return contigs, gp, go, flat


def same_bytes(a, b):
    return (np.array_equal(np.asarray(a.rec).view(np.uint8), np.asarray(b.rec).view(np.uint8))
            and np.array_equal(np.asarray(a.cigar), np.asarray(b.cigar)))


def test_default_shape_matches_oracle(setup):
    contigs, gp, go, flat = setup
    ctx = P.Context(gp)
    d = ctx.upload_reads(P.ReadBatch.from_arrays(*flat))
    rp = ctx.map_se_device(d)
    ro = O.Mapper(go, O.HashTable(go)).map_se(O.ReadBatch.from_arrays(*flat))
    assert P.primary_identical(ro, rp).all()


@pytest.mark.parametrize("shape", [
    dict(split_len=96, long_waves=0),
    dict(split_len=96, long_waves=1),
    dict(split_len=160, long_waves=1, dp_grid=64),
    dict(split_len=256, long_waves=0, dp_grid=8192),
    dict(batch_reads=777, dp_grid=4),
    dict(dp_pack16=1),
    dict(dp_pack16=1, split_len=96, long_waves=0, dp_grid=64),
    dict(seed_pool=16),
])
def test_launch_shapes_give_identical_records(setup, shape):
    _, gp, _, flat = setup
    ctx = P.Context(gp)
    d = ctx.upload_reads(P.ReadBatch.from_arrays(*flat))
    base = ctx.map_se_device(d)
    ctx.tune = shape
    got = ctx.map_se_device(d)
    assert ctx.tune.as_dict()["split_len"] == shape.get("split_len", 0)
    assert same_bytes(base, got), shape


def test_autotune_keeps_records_and_reports_trials(setup):
    _, gp, _, flat = setup
    ctx = P.Context(gp)
    d = ctx.upload_reads(P.ReadBatch.from_arrays(*flat))
    base = ctx.map_se_device(d)
    rep = ctx.autotune(d)
    assert rep["best_ms"] > 0 and len(rep["trials"]) >= 6
    best = {k: v for k, v in rep["best"].items()}
    assert best in [{k: t[k] for k in best} for t in rep["trials"]]
    assert ctx.tune.as_dict() == best
    assert same_bytes(base, ctx.map_se_device(d))
