"""The five benchmark workloads, each end to end on its own generator, HIP path against
the oracle through the C ABI.

- config1_probe_se_2mbp: all 20,000 probe reads (120 bp) on the 2.1 Mbp probe genome;
- config2_uneven_se_100mbp: 20,000 reads (80-150 bp, decaying qualities) on the
  100 Mbp / 10-contig genome with segmental duplications and homopolymers;
- config3_probe_pe_2mbp: 10,000 FR pairs of 2 x 150 (fragment N(350, 50), 10% with one
  mate in a repeat copy), records, SAM text and the learned insert model;
- config4_prefilter_se_2mbp: the learned path (network trained on 500,000 reads of the
  configuration's generator, 20 epochs) on 50,000 held-out reads -- plain records equal
  to the oracle's, learned records >= 99.9% SAM-identical to the oracle's;
- config5_genome_3gbp: 4,096 reads (60-250 bp) and 1,024 pairs on the 3.1 Gbp genome,
  against the oracle's records in tests/golden/genome_3gbp.npz (written by
  tests/golden/make_genome_3gbp.py).

Every record field is compared exactly except the posterior, held to 1e-12 relative.
"""
import hashlib
import os
import sys

import numpy as np
import pytest

import probebwa_amd as P
from oracle import pbw_oracle as O
from probebwa_amd import synth
from test_gpu_parity import compare_records, pair_of_genomes

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
if os.path.join(HERE, "golden") not in sys.path:
    sys.path.insert(0, os.path.join(HERE, "golden"))
ORACLE_THREADS = 16


def oracle_mapper(go):
    return O.Mapper(go, O.HashTable(go, threads=ORACLE_THREADS), threads=ORACLE_THREADS)


def pair_batches(a, b, n):
    ids = ["p%d" % i for i in range(n)]
    o1, o2 = O.ReadBatch.from_arrays(*a, ids=ids), O.ReadBatch.from_arrays(*b, ids=ids)
    p1, p2 = P.ReadBatch.from_arrays(*a, ids=ids), P.ReadBatch.from_arrays(*b, ids=ids)
    return o1, o2, p1, p2


def interleaved(o1, o2):
    """The interleaved r1, r2 batch of map_pe's record order, for both bindings."""
    ob = O.interleave(o1, o2)
    return ob, P.ReadBatch.from_arrays(ob.seq, ob.seq_off, ob.qual, ob.qual_off, ob.ids)


@pytest.mark.timeout(600)
def test_config1_probe_se_2mbp():
    contigs = synth.genome_probe_se_2mbp()
    go, gp = pair_of_genomes(contigs)
    (seq, so, qual, qo), _truth = synth.reads_probe_se(contigs, n=20_000, L=120, seed=11)
    reads = P.ReadBatch.from_arrays(seq, so, qual, qo)
    ro = oracle_mapper(go).map_se(O.ReadBatch.from_arrays(seq, so, qual, qo))
    rp = P.Context(gp).map_se(reads)
    compare_records(ro, rp, reads, go, gp, label="config1")


@pytest.mark.timeout(600)
def test_config2_uneven_se_100mbp():
    contigs = synth.genome_uneven_100mbp()
    go, gp = pair_of_genomes(contigs)
    seq, so, qual, qo = synth.reads_uneven_se(contigs, n=20_000, seed=12)
    reads = P.ReadBatch.from_arrays(seq, so, qual, qo)
    ro = oracle_mapper(go).map_se(O.ReadBatch.from_arrays(seq, so, qual, qo))
    rp = P.Context(gp).map_se(reads)
    compare_records(ro, rp, reads, go, gp, label="config2")


@pytest.mark.timeout(600)
def test_config3_probe_pe_2mbp():
    contigs = synth.genome_probe_se_2mbp()
    go, gp = pair_of_genomes(contigs)
    n = 10_000
    a, b = synth.reads_pairs_probe(contigs, n=n, seed=3, repeat_spans=synth.probe_repeat_spans())
    o1, o2, p1, p2 = pair_batches(a, b, n)
    m = oracle_mapper(go)
    ro = m.map_pe(o1, o2)
    ctx = P.Context(gp)
    rp = ctx.map_pe(p1, p2)
    ob, pb = interleaved(o1, o2)
    compare_records(ro, rp, pb, go, gp, label="config3")
    assert tuple(m.insert_model) == tuple(ctx.insert_model)
    proper = (rp.rec["flag"] & 0x2) != 0
    assert proper.mean() > 0.9 and (rp.rec["tlen"] != 0).mean() > 0.9


@pytest.mark.timeout(900)
def test_config4_prefilter_se_2mbp():
    contigs = synth.genome_probe_se_2mbp()
    go, gp = pair_of_genomes(contigs)
    ctx = P.Context(gp)
    train = P.ReadBatch.from_arrays(*synth.reads_probe_uneven_se(contigs, n=500_000, seed=41))
    rep = ctx.prefilter_train(train, recall=0.9999, epochs=20, batch=4096, lr=1e-3)
    assert rep["loss_last"] < rep["loss_first"]
    seq, so, qual, qo = synth.reads_probe_uneven_se(contigs, n=50_000, seed=42)
    reads = P.ReadBatch.from_arrays(seq, so, qual, qo)
    ro = oracle_mapper(go).map_se(O.ReadBatch.from_arrays(seq, so, qual, qo))
    ctx.set_prefilter(enabled=0)
    plain = ctx.map_se(reads)
    compare_records(ro, plain, reads, go, gp, label="config4 plain")
    ctx.set_prefilter(enabled=1, audit_period=64)
    learned = ctx.map_se(reads)
    st = ctx.stats()
    same = P.primary_identical(ro, learned)
    assert same.mean() >= 0.999, "learned path %.5f SAM-identical to the oracle" % same.mean()
    assert st["pf_audit_reads"] > 0 and st["pf_audit_mismatches"] <= max(1, st["pf_audit_reads"] // 1000)
    assert st["pf_jobs_aligned"] < st["pf_jobs"]


# ---------------------------------------------------------------------------
# genome scale, against the golden records
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def genome_3gbp():
    import make_genome_3gbp as MK  # tests/golden
    gold = np.load(os.path.join(HERE, "golden", "genome_3gbp.npz"))
    contigs = synth.genome_scale_3gbp()
    assert MK.genome_digest(contigs) == str(gold["genome_digest"]), "genome generator changed: rerun the script"
    se, pe = MK.samples(contigs)
    assert MK.reads_digest(*se) == str(gold["se_digest"]) and MK.reads_digest(*pe[0], *pe[1]) == str(gold["pe_digest"])
    gp = P.Genome()
    for name, s in contigs:
        gp.add_contig(name, s.tobytes())
    del contigs
    ctx = P.Context(gp)
    yield gold, gp, ctx, se, pe
    del ctx


def _golden_records(gold, tag):
    return O.Records(gold[tag + "_rec"], gold[tag + "_cigar"])


@pytest.mark.timeout(900)
def test_config5_genome_se_3gbp(genome_3gbp):
    gold, gp, ctx, se, _pe = genome_3gbp
    reads = P.ReadBatch.from_arrays(*se)
    rp = ctx.map_se(reads)
    compare_records(_golden_records(gold, "se"), rp, label="config5 single-end")
    sam = P.format_sam(gp, reads, rp)
    assert hashlib.sha256(sam).hexdigest() == str(gold["se_sam_sha256"])
    assert ctx.stats()["tier2_cluster_reads"] > 0  # the block clustering tier is exercised


@pytest.mark.timeout(900)
def test_config5_genome_pe_3gbp(genome_3gbp):
    gold, gp, ctx, _se, pe = genome_3gbp
    n = len(pe[0][1]) - 1
    o1, o2, p1, p2 = pair_batches(pe[0], pe[1], n)
    rp = ctx.map_pe(p1, p2)
    compare_records(_golden_records(gold, "pe"), rp, label="config5 paired-end")
    _ob, pb = interleaved(o1, o2)
    assert hashlib.sha256(P.format_sam(gp, pb, rp)).hexdigest() == str(gold["pe_sam_sha256"])
    assert tuple(ctx.insert_model) == tuple(float(x) for x in gold["pe_model"][:2]) + (int(gold["pe_model"][2]),)
