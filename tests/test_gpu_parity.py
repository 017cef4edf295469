"""HIP path vs the CPU oracle: bit-exact SAM fields, posteriors within 1e-12.

Every test here runs the HIP kernels through the C ABI (libprobebwa_amd.so)
and checks the result against the oracle on the same seeded inputs.
"""
import numpy as np
import pytest

import probebwa_amd as P
from oracle import pbw_oracle as O
from probebwa_amd import synth
from refcases import ECOLI_READ, ECOLI_WINDOW, Xorshift64, homopolymer_battery, homopolymer_scenario, q40, revcomp

pytestmark = pytest.mark.gpu

POSTERIOR_RTOL = 1e-12  # north-star tolerance for raw posteriors (FP64)


def pair_of_genomes(contigs):
    go, gp = O.Genome(), P.Genome()
    for name, seq in contigs:
        s = bytes(seq)
        go.add_contig(name, s)
        gp.add_contig(name, s)
    return go, gp


def compare_records(ro, rp, reads=None, go=None, gp=None, label=""):
    """Field-by-field comparison; returns number of mismatching records (asserted 0)."""
    assert len(ro) == len(rp)
    bad = []
    for i in range(len(ro)):
        a, b = ro.rec[i], rp.rec[i]
        same = (a["pos"] == b["pos"] and a["contig"] == b["contig"] and a["flag"] == b["flag"]
                and a["mapq"] == b["mapq"] and a["strand"] == b["strand"] and a["seq_rc"] == b["seq_rc"]
                and a["mate_pos"] == b["mate_pos"] and a["mate_contig"] == b["mate_contig"] and a["tlen"] == b["tlen"]
                and ro.cigar_words(i) == rp.cigar_words(i) and a["log_lk"] == b["log_lk"])
        pa, pb = float(a["posterior"]), float(b["posterior"])
        if pa != pb and abs(pa - pb) > POSTERIOR_RTOL * max(abs(pa), abs(pb)):
            same = False
        if not same:
            bad.append(i)
    if bad:
        i = bad[0]
        msg = "%s: %d/%d records differ; first #%d oracle=%s %s gpu=%s %s" % (
            label, len(bad), len(ro), i, ro.rec[i], ro.cigar_str(i), rp.rec[i], rp.cigar_str(i))
        raise AssertionError(msg)
    if reads is not None:
        so = O.format_sam(go, O.ReadBatch.from_arrays(reads.seq, reads.seq_off, reads.qual, reads.qual_off, reads.ids), ro)
        sp = P.format_sam(gp, reads, rp)
        assert so == sp, label + ": SAM text differs"


def both_reads(seq, so, qual, qo):
    return (O.ReadBatch.from_arrays(seq, so, qual, qo), P.ReadBatch.from_arrays(seq, so, qual, qo))


# ---------------------------------------------------------------------------
# reference known answers through the HIP path (tests/mapping.rs)
# ---------------------------------------------------------------------------
def se_hip(seq, reads, quals=None, name="chr1"):
    g = P.Genome().add_contig(name, seq)
    ctx = P.Context(g)
    quals = quals or [q40(r) for r in reads]
    return ctx.map_se(P.ReadBatch(reads, quals))


def test_hip_forward_and_reverse_strand_mapping():  # tests/mapping.rs:67-103
    seq = Xorshift64(42).gen_seq(400)
    r = se_hip(seq, [seq[40:80], revcomp(seq[200:240])])
    assert (r.rec[0]["pos"], r.rec[0]["strand"], r.cigar_str(0), r.rec[0]["mapq"]) == (41, 0, "40M", 80)
    assert (r.rec[1]["pos"], r.rec[1]["strand"], r.cigar_str(1), r.rec[1]["mapq"]) == (201, 1, "40M", 80)


def test_hip_mismatch_deletion_and_unmapped():  # tests/mapping.rs:105-160
    seq = Xorshift64(7).gen_seq(400)
    base = seq[40:80]
    mm = list(base)
    mm[10] = "C" if mm[10] == "A" else "A"
    r = se_hip(seq, ["".join(mm), base[:20] + base[22:], "T" * 40])
    assert (r.rec[0]["pos"], r.cigar_str(0), r.rec[0]["mapq"]) == (41, "40M", 80)
    ops = [(w & 15, w >> 4) for w in r.cigar_words(1)]
    assert r.rec[1]["pos"] == 41 and sum(n for o, n in ops if o == 0) == 38 and sum(n for o, n in ops if o == 2) == 2
    assert r.rec[1]["mapq"] > 0
    assert r.rec[2]["pos"] == 0 and r.rec[2]["mapq"] == 0 and r.cigar_str(2) == "40S" and r.rec[2]["flag"] == 4


def test_hip_large_deletion_and_large_positions():  # tests/mapping.rs:171-193, :240-258
    seq = Xorshift64(555).gen_seq(500)
    w = seq[100:160]
    r = se_hip(seq, [w[:25] + w[50:]])
    assert r.rec[0]["pos"] == 101 and sum(w >> 4 for w in r.cigar_words(0) if w & 15 == 2) == 25
    seq = Xorshift64(2024).gen_seq(80000)
    r = se_hip(seq, [seq[70000:70050]])
    assert r.rec[0]["pos"] == 70001 and r.rec[0]["mapq"] == 99


def test_hip_n_run_round_trip():  # tests/mapping.rs:304-334
    rng = Xorshift64(7)
    left, right = rng.gen_seq(100), rng.gen_seq(100)
    seq = left + "N" * 20 + right
    r = se_hip(seq, [seq[105:135], seq[150:190]], name="chrN")
    assert r.rec[0]["pos"] == 106
    assert r.rec[1]["pos"] == 151 and r.rec[1]["mapq"] == 80


def test_hip_divergent_region_does_not_invent_spurious_indels():  # tests/integration.rs:130-153
    g = P.Genome().add_contig("x", "ACGTACGTACGTACGTACGT")
    ctx = P.Context(g)
    (score, cig, off), = ctx.align_batch([(ECOLI_READ, bytes([30] * 100), ECOLI_WINDOW, len(ECOLI_WINDOW), 513)])
    assert off == 513 and all(w & 15 == 0 for w in cig)
    s, c, o, _ = O.SmithWatermanAligner(len(ECOLI_WINDOW)).align(ECOLI_READ, bytes([30] * 100), ECOLI_WINDOW, 513)
    assert (score, cig, off) == (s, c, o)


def test_hip_homopolymer_battery_matches_oracle():  # examples/tune_homopolymer.rs:87-146
    correct = 0
    for run_len, del_len, n_err, seed, vs in homopolymer_battery():
        genome, read, true_pos = homopolymer_scenario(vs, run_len, del_len, n_err)
        g = P.Genome().add_contig("chr1", genome)
        r = P.Context(g).map_se(P.ReadBatch([read], [bytes([10] * len(read))]))
        dels = sum(w >> 4 for w in r.cigar_words(0) if w & 15 == 2)
        if abs(int(r.rec[0]["pos"]) - true_pos) <= 2 and dels == del_len:
            correct += 1
    assert correct == 285


# ---------------------------------------------------------------------------
# kernel-level parity
# ---------------------------------------------------------------------------
def test_index_buckets_match_oracle():
    contigs = synth.small_genome(5, n_contigs=3, length=30000)
    go, gp = pair_of_genomes(contigs)
    ht = O.HashTable(go)
    ctx = P.Context(gp)
    assert ctx.total_entries == ht.total_entries
    rng = np.random.default_rng(3)
    seqs = [bytes(c[1]) for c in contigs]
    checked = 0
    for _ in range(3000):
        s = seqs[rng.integers(0, len(seqs))]
        p = int(rng.integers(0, len(s) - 15))
        kmer = s[p:p + 15]
        if rng.random() < 0.3:
            kmer = bytes(synth.random_bases(rng, 15))
        if b"N" in kmer:
            continue
        assert list(ctx.lookup_code(P.code_of(kmer))) == ht.lookup(kmer)
        checked += 1
    assert checked > 2000


def _bucket_code(contigs, packed):
    """Code of the 15-mer a packed position (hashtable.rs:157-164) indexes."""
    packed = int(packed)
    cid, pos = (packed >> 40) & ((1 << 23) - 1), packed & ((1 << 40) - 1)
    kmer = np.asarray(contigs[cid][1][pos:pos + 15], np.uint8)
    if packed >> 63:
        kmer = synth.revcomp(kmer)
    return P.code_of(kmer.tobytes())


def test_index_full_table_matches_oracle():
    """The device-built bucket table equals HashTableBuilder::build (hashtable.rs:248-322) as a whole:
    every oracle bucket equals the device bucket of its code, in scan order, and the totals agree,
    so the device holds no other entry.  The genome hits the build's edge cases: contigs shorter
    than, equal to and just over k (no window, one window per strand, one window), lengths at every
    residue mod 5, N runs, and a poly-A run whose 15-mer passes the 300 cap across strands and
    contigs (the cap counts in scan order: contig, forward windows, then reverse windows)."""
    rng = np.random.default_rng(np.random.PCG64(11))
    contigs = []
    for i, ln in enumerate([0, 7, 14, 15, 16, 19, 20, 21, 22, 23, 24, 25, 40, 1234]):
        contigs.append(("t%d" % i, synth.random_bases(rng, ln)))
    big = synth.random_bases(rng, 20000)
    big[3000:3900] = ord("A")   # 900 A's: 178 forward AAA.. windows, 178 reverse TTT.. windows
    big[9000:9040] = ord("N")
    big[12001:12003] = ord("N")
    contigs.append(("big1", big))
    big2 = synth.random_bases(rng, 8003)
    big2[100:900] = ord("T")    # the T-run's reverse windows add AAA.. entries past the cap
    big2[5000] = ord("N")
    contigs.append(("big2", big2))
    contigs += synth.small_genome(5, n_contigs=2, length=15001)
    go, gp = pair_of_genomes(contigs)
    ht = O.HashTable(go)
    ctx = P.Context(gp)
    assert ctx.total_entries == ht.total_entries
    n_keys, n_entries, capped = 0, 0, 0
    for _h, bucket in ht.buckets():
        code = _bucket_code(contigs, bucket[0])
        dev = ctx.lookup_code(code)
        assert np.array_equal(dev, bucket), "bucket of code %d differs" % code
        n_keys += 1
        n_entries += len(bucket)
        capped += len(bucket) == 300
    assert n_entries == ctx.total_entries
    assert capped >= 2 and n_keys > 10000


def random_align_jobs(rng, n, wide=False):
    jobs = []
    for _ in range(n):
        m = int(rng.integers(1, 60 if not wide else 200))
        nref = int(rng.integers(1, 90 if not wide else 600))
        read = bytes(synth.random_bases(rng, m))
        ref = bytearray(synth.random_bases(rng, nref))
        if rng.random() < 0.6 and nref > 5:  # plant the read (with edits) in the ref
            p = int(rng.integers(0, nref))
            src = synth.mutate(rng, np.frombuffer(read, np.uint8), 0.05).tobytes()
            ref[p:p + m] = src[: max(0, min(m, nref - p))]
            ref = ref[:nref]
        if rng.random() < 0.1:
            ref[int(rng.integers(0, len(ref)))] = ord("N")
        if rng.random() < 0.1:
            read = read[: max(0, m - 1)] + b"N"
        if rng.random() < 0.2:  # homopolymers
            p = int(rng.integers(0, len(ref)))
            ref[p:p + 6] = b"AAAAAA"
        qual = bytes(rng.integers(0, 45, len(read) - (1 if rng.random() < 0.1 and len(read) > 1 else 0)).astype(np.uint8))
        band = int(rng.integers(0, 32)) if not wide else int(rng.choice([len(ref), 40, 100]))
        hint = int(rng.integers(-40, 60)) if rng.random() < 0.9 else int(rng.integers(-200, 200))
        jobs.append((read, qual, bytes(ref), band, hint))
    return jobs


@pytest.mark.parametrize("wide", [False, True])
def test_align_batch_matches_oracle(wide):
    rng = np.random.default_rng(17 + wide)
    jobs = random_align_jobs(rng, 3000 if not wide else 600, wide)
    g = P.Genome().add_contig("x", "ACGT" * 10)
    ctx = P.Context(g)
    for gop, gep in ((40, 3), (10, 20)):
        got = ctx.align_batch(jobs, gop, gep)
        for i, (read, qual, ref, band, hint) in enumerate(jobs):
            s, c, o, _ = O.SmithWatermanAligner(band, gop, gep).align(read, qual, ref, hint)
            assert got[i] == (s, c, o), "job %d (m=%d n=%d band=%d hint=%d)" % (i, len(read), len(ref), band, hint)


# ---------------------------------------------------------------------------
# whole single-end pipeline parity
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("seed,n_contigs", [(21, 1), (22, 3), (23, 7)])
def test_map_se_matches_oracle(seed, n_contigs):
    contigs = synth.small_genome(seed, n_contigs=n_contigs, length=40000)
    go, gp = pair_of_genomes(contigs)
    seq, so, qual, qo = synth.sample_reads(contigs, 1500, seed + 100)
    ro_b, rp_b = both_reads(seq, so, qual, qo)
    ro = O.Mapper(go, O.HashTable(go)).map_se(ro_b)
    rp = P.Context(gp).map_se(rp_b)
    compare_records(ro, rp, rp_b, go, gp, "seed %d" % seed)
    assert (rp.rec["pos"] > 0).mean() > 0.8


def test_map_se_small_batches_and_options():
    """Output must not depend on the device batch size; non-default penalties."""
    contigs = synth.small_genome(31, n_contigs=2, length=30000)
    go, gp = pair_of_genomes(contigs)
    seq, so, qual, qo = synth.sample_reads(contigs, 700, 131)
    ro_b, rp_b = both_reads(seq, so, qual, qo)
    opts = dict(substitution_rate=0.01, gap_open_phred=30, gap_extend_phred=6)
    ro = O.Mapper(go, O.HashTable(go), **opts).map_se(ro_b)
    rp = P.Context(gp, batch_reads=97, **opts).map_se(rp_b)
    compare_records(ro, rp, label="batch 97")


def test_map_se_repeat_heavy_reads():
    """Hyper-repetitive genome: buckets at the 300 cap, tier-2 clustering, hit overflow."""
    rng = np.random.default_rng(41)
    unit = synth.random_bases(rng, 37)
    seqs = []
    s = np.concatenate([synth.mutate(rng, unit, 0.03) for _ in range(1500)])
    seqs.append(("rep", s))
    seqs.append(("uniq", synth.random_bases(rng, 20000)))
    go, gp = pair_of_genomes(seqs)
    seq, so, qual, qo = synth.sample_reads(seqs, 200, 141, lmin=60, lmax=180)
    ro_b, rp_b = both_reads(seq, so, qual, qo)
    ro = O.Mapper(go, O.HashTable(go)).map_se(ro_b)
    ctx = P.Context(gp)
    rp = ctx.map_se(rp_b)
    compare_records(ro, rp, label="repeats")
    st = ctx.stats()
    assert st["tier2_cluster_reads"] > 0 or st["seed_overflow_reads"] > 0


def test_map_se_first_pass_capacity_adapts():
    """When a batch's hits outgrow the seed pass's chunk pool, the reads left without a chunk
    are seeded again into exact slots and the later batches take a larger pool; the records
    stay the oracle's in every batch and call."""
    rng = np.random.default_rng(43)
    unit = synth.random_bases(rng, 41)
    seqs = [("rep", np.concatenate([synth.mutate(rng, unit, 0.03) for _ in range(1200)])),
            ("uniq", synth.random_bases(rng, 30000))]
    go, gp = pair_of_genomes(seqs)
    seq, so, qual, qo = synth.sample_reads(seqs, 240, 143, lmin=60, lmax=200)
    ro_b, rp_b = both_reads(seq, so, qual, qo)
    ro = O.Mapper(go, O.HashTable(go)).map_se(ro_b)
    ctx = P.Context(gp, batch_reads=40)
    rp1 = ctx.map_se(rp_b)
    ov1 = ctx.stats()["seed_overflow_reads"]
    rp2 = ctx.map_se(rp_b)
    ov2 = ctx.stats()["seed_overflow_reads"]
    compare_records(ro, rp1, label="first call")
    compare_records(ro, rp2, label="second call")
    assert ov1 > 0 and ov2 <= ov1, (ov1, ov2)


def test_map_se_batch_over_a_million_reads_matches_smaller_batches():
    """1.1 M reads in one batch with a seed pool of 1000 chunks: nearly every read's
    hits find no chunk, so the reads seeded again into exact slots are listed over
    more than 1024 blocks of 1024 reads and the block scan runs in two chunks.  The
    records equal those of 65,536-read batches with the grown pool, byte for byte."""
    rng = np.random.default_rng(7)
    unit = synth.random_bases(rng, 41)
    rep = np.concatenate([synth.mutate(rng, unit, 0.03) for _ in range(400)])
    uniq = synth.random_bases(rng, 3_000_000)
    g = P.Genome()
    g.add_contig("rep", rep.tobytes())
    g.add_contig("uniq", uniq.tobytes())
    n, L, n_rep = 1_100_000, 36, 3000
    src = np.where(np.arange(n) < n_rep, 0, 1)
    rng.shuffle(src)
    starts = np.where(src == 0, rng.integers(0, len(rep) - L, n), rng.integers(0, len(uniq) - L, n))
    idx = starts[:, None] + np.arange(L)[None, :]
    seq = np.where((src == 0)[:, None], rep[np.minimum(idx, len(rep) - 1)], uniq[idx]).astype(np.uint8).ravel()
    qual = rng.integers(20, 41, n * L).astype(np.uint8)
    off = np.arange(0, n * L + 1, L, dtype=np.uint64)
    reads = P.ReadBatch.from_arrays(seq, off, qual, off)
    ctx = P.Context(g)
    ctx.tune = dict(batch_reads=1_200_000, seed_pool=1000)
    big = ctx.map_se(reads)
    assert ctx.stats()["seed_overflow_reads"] > 1_000_000
    ctx2 = P.Context(g)
    ctx2.tune = dict(batch_reads=65536)
    small = ctx2.map_se(reads)
    assert np.array_equal(np.asarray(big.rec).view(np.uint8), np.asarray(small.rec).view(np.uint8))
    assert np.array_equal(np.asarray(big.cigar), np.asarray(small.cigar))


@pytest.mark.parametrize("n_contigs,lmax", [(60, 30), (150, 60)])
def test_map_se_many_contig_groups(n_contigs, lmax):
    """A 400 bp unit in every contig, in alternating orientation: short reads
    from it hit > 56 and > 128 (contig, strand) groups, which takes the table
    and block-tier paths of the emulated FxHashMap iteration order."""
    rng = np.random.default_rng(61 + n_contigs)
    unit = synth.random_bases(rng, 400)
    contigs = []
    for c in range(n_contigs):
        u = synth.mutate(rng, unit, 0.01)
        if c % 2:
            u = synth.revcomp(u)
        flank = lambda: synth.random_bases(rng, 300)
        contigs.append(("c%d" % c, np.concatenate([flank(), u, flank()])))
    go, gp = pair_of_genomes(contigs)
    seq, so, qual, qo = synth.sample_reads(contigs, 400, 161 + n_contigs, lmin=18, lmax=lmax)
    ro_b, rp_b = both_reads(seq, so, qual, qo)
    ro = O.Mapper(go, O.HashTable(go)).map_se(ro_b)
    ctx = P.Context(gp)
    rp = ctx.map_se(rp_b)
    compare_records(ro, rp, label="%d contigs" % n_contigs)


def test_map_se_over_32768_contigs():
    """A genome of 33,000 short contigs: more contigs than an index entry with the flank
    fingerprint inside can name (2^15), so the index keeps the reference's packed positions
    (hashtable.rs:157-164) with the fingerprints in an array beside them.  Records equal the
    oracle's, and bucket lookups return the reference's packing."""
    rng = np.random.default_rng(71)
    contigs = [("u%d" % c, synth.random_bases(rng, int(rng.integers(40, 161)))) for c in range(33000)]
    go, gp = pair_of_genomes(contigs)
    ht = O.HashTable(go)
    ctx = P.Context(gp)
    assert ctx.total_entries == ht.total_entries
    checked = 0
    for _ in range(2000):
        s = bytes(contigs[int(rng.integers(0, len(contigs)))][1])
        p = int(rng.integers(0, len(s) - 15))
        kmer = s[p:p + 15]
        if b"N" in kmer:
            continue
        assert list(ctx.lookup_code(P.code_of(kmer))) == ht.lookup(kmer)
        checked += 1
    assert checked > 1500
    seq, so, qual, qo = synth.sample_reads(contigs, 3000, 171, lmin=30, lmax=120)
    ro_b, rp_b = both_reads(seq, so, qual, qo)
    compare_records(O.Mapper(go, ht).map_se(ro_b), ctx.map_se(rp_b), label="33000 contigs")


def test_map_se_edge_reads():
    """Empty reads, reads shorter than k, all-N reads, quality tracks of other lengths."""
    contigs = synth.small_genome(51, n_contigs=1, length=20000)
    go, gp = pair_of_genomes(contigs)
    g = bytes(contigs[0][1])
    reads = [b"", b"ACGT", g[100:114], g[100:115], b"N" * 50, g[5000:5100], g[:60], g[-60:], g[300:420]]
    quals = [b"", b"\x05\x05", bytes([30] * 14), bytes([30] * 15), bytes([2] * 50), bytes([40] * 90), bytes([20] * 60),
             bytes([35] * 70), bytes([11] * 120)]
    ro = O.Mapper(go, O.HashTable(go)).map_se(O.ReadBatch(reads, quals))
    rp = P.Context(gp).map_se(P.ReadBatch(reads, quals))
    compare_records(ro, rp, label="edge")


# ---------------------------------------------------------------------------
# paired-end (paired.rs, lib.rs:145-185)
# ---------------------------------------------------------------------------
def pe_hip(seq, m1, m2, q1=None, q2=None):
    g = P.Genome().add_contig("chr1", seq)
    ctx = P.Context(g)
    q1 = q1 or [q40(r) for r in m1]
    q2 = q2 or [q40(r) for r in m2]
    return ctx.map_pe(P.ReadBatch(m1, q1), P.ReadBatch(m2, q2))


def test_hip_paired_end_proper_and_improper():  # tests/mapping.rs:195-233
    seq = Xorshift64(99).gen_seq(400)
    r = pe_hip(seq, [seq[40:80]], [revcomp(seq[200:240])])
    a, b = r.rec[0], r.rec[1]
    assert a["flag"] & 0x40 and b["flag"] & 0x80 and a["flag"] & 2 and b["flag"] & 2
    assert (a["strand"], b["strand"], a["tlen"], b["tlen"]) == (0, 1, 200, -200)
    r = pe_hip(seq, [seq[40:80]], [seq[200:240]])
    assert not (r.rec[0]["flag"] & 2) and not (r.rec[1]["flag"] & 2)


def test_hip_mate_rescue_recovers_unseedable_mate():  # tests/mapping.rs:267-302
    seq = Xorshift64(2468).gen_seq(400)
    m2 = list(revcomp(seq[200:240]))
    for p in range(0, len(m2), 6):
        m2[p] = next(b for b in "ACGT" if b != m2[p])
    m2 = "".join(m2)
    r = pe_hip(seq, [seq[40:80]], [m2])
    assert r.rec[0]["pos"] == 41 and r.rec[1]["strand"] == 1 and abs(int(r.rec[1]["pos"]) - 201) <= 5


def test_hip_paired_disambiguates_repeat_copy_via_anchor():  # tests/mapping.rs:421-507
    rng = Xorshift64(2026)
    unique_a, spacer, unit, filler = rng.gen_seq(100), rng.gen_seq(110), rng.gen_seq(100), rng.gen_seq(10000)
    tc = list(unit)
    tc[20] = {"A": "C", "C": "A", "G": "T"}.get(tc[20], "G")
    genome = unique_a + spacer + "".join(tc) + filler + unit
    q2 = [35] * 40
    q2[19] = 10
    r = pe_hip(genome, [unique_a[:40]], [revcomp(unit[:40])], [bytes([35] * 40)], [bytes(q2)])
    assert r.rec[0]["pos"] == 1 and r.rec[1]["pos"] == 211 and r.rec[1]["strand"] == 1
    assert (r.rec[0]["flag"] & 2) and (r.rec[1]["flag"] & 2) and r.rec[1]["mapq"] > 0


@pytest.mark.parametrize("seed,n_pairs", [(61, 120), (62, 700)])
def test_map_pe_matches_oracle(seed, n_pairs):
    contigs = synth.small_genome(seed, n_contigs=2, length=40000, repeat_copies=8)
    go, gp = pair_of_genomes(contigs)
    (a, b) = synth.reads_pairs(contigs, n=n_pairs, L=100, seed=seed + 1,
                               repeat_spans=())
    o1, p1 = both_reads(*a)
    o2, p2 = both_reads(*b)
    m = O.Mapper(go, O.HashTable(go))
    ro = m.map_pe(o1, o2)
    ctx = P.Context(gp)
    rp = ctx.map_pe(p1, p2)
    compare_records(ro, rp, label="pairs seed %d" % seed)
    assert m.insert_model[2] == ctx.insert_model[2]
    assert m.insert_model[0] == ctx.insert_model[0] and m.insert_model[1] == ctx.insert_model[1]


def test_map_pe_many_batches_matches_oracle():
    """4700 pairs: the 500 warm-up pairs, then 2000-pair batches on frozen model
    snapshots (lib.rs:145-185).  With the default read batch the three pair
    batches share one seeding/clustering/DP pass; with 4000 reads per batch each
    gets its own.  Both give the oracle's records and insert-size model."""
    contigs = synth.small_genome(63, n_contigs=2, length=40000, repeat_copies=8)
    go, gp = pair_of_genomes(contigs)
    (a, b) = synth.reads_pairs(contigs, n=4700, L=100, seed=64, repeat_spans=())
    o1, p1 = both_reads(*a)
    o2, p2 = both_reads(*b)
    m = O.Mapper(go, O.HashTable(go))
    ro = m.map_pe(o1, o2)
    for shape in (None, dict(batch_reads=4000)):
        ctx = P.Context(gp)
        if shape:
            ctx.tune = shape
        rp = ctx.map_pe(p1, p2)
        compare_records(ro, rp, label="4700 pairs, tune %s" % shape)
        assert tuple(m.insert_model) == tuple(ctx.insert_model)


# ---------------------------------------------------------------------------
# shapes beyond LDS: long reads, mate windows of thousands of columns
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("frag", [(1200, 200, 400, 2500), (2000, 350, 800, 4000), (3000, 1000, 500, 8000)])
def test_map_pe_wide_insert_windows_match_oracle(frag):
    """Mate windows of 2 (ceil(5 sd) + L) + L columns (align_mate_near,
    paired.rs:450-468): about 2,450 and 3,950 columns (k_dp_wide with 12 and 16
    tiles per job) and about 10,450 (the tiled kernel, staged in device memory).
    The model starts at the library's own (mean, sd), so the windows stay that wide."""
    contigs = synth.small_genome(71, n_contigs=2, length=150000, repeat_copies=8)
    go, gp = pair_of_genomes(contigs)
    a, b = synth.reads_pairs(contigs, n=700, L=150, seed=72, frag=frag)
    o1, p1 = both_reads(*a)
    o2, p2 = both_reads(*b)
    m = O.Mapper(go, O.HashTable(go), threads=16)
    ro = m.map_pe(o1, o2, mean=float(frag[0]), sd=float(frag[1]))
    ctx = P.Context(gp)
    rp = ctx.map_pe(p1, p2, mean=float(frag[0]), sd=float(frag[1]))
    compare_records(ro, rp, label="pairs, fragments N(%d, %d)" % frag[:2])
    assert tuple(m.insert_model) == tuple(ctx.insert_model)
    assert ctx.stats()["wide_jobs"] > 0
    assert ((rp.rec["flag"] & 2) != 0).mean() > 0.8


def test_map_se_long_reads_match_oracle():
    """Reads of 30-6000 bases in one call: the longer than 1,024 are seeded from
    device memory and their banded DP is staged in device memory (the reference
    has no read-length limit)."""
    contigs = synth.small_genome(81, n_contigs=2, length=120000)
    go, gp = pair_of_genomes(contigs)
    seq, so, qual, qo = synth.sample_reads(contigs, 300, 82, lmin=30, lmax=6000)
    assert (np.diff(so) > 1024).sum() > 200
    ro_b, rp_b = both_reads(seq, so, qual, qo)
    ro = O.Mapper(go, O.HashTable(go), threads=16).map_se(ro_b)
    for shape in (None, dict(batch_reads=64)):
        ctx = P.Context(gp)
        if shape:
            ctx.tune = shape
        rp = ctx.map_se(rp_b)
        compare_records(ro, rp, rp_b, go, gp, "long reads, tune %s" % shape)
    assert (rp.rec["pos"] > 0).mean() > 0.8


def test_align_batch_long_reads_and_wide_windows_match_oracle():
    """`align` on reads over 1,024 bases and windows over 8,000 columns, banded and
    whole-window, through the banded, wide and tiled kernels (staging in device
    memory where LDS is too small)."""
    rng = np.random.default_rng(91)
    jobs = []
    for m, nref, band in [(1500, 1560, 30), (3000, 3100, 20), (2000, 9000, None), (300, 12000, None),
                          (1200, 5000, 100), (150, 3900, None), (700, 700, 31), (40, 20000, None)]:
        ref = bytearray(synth.random_bases(rng, nref))
        p = int(rng.integers(0, max(1, nref - m)))
        read = synth.mutate(rng, np.frombuffer(bytes(ref[p:p + m]), np.uint8), 0.03).tobytes()
        if len(read) < m:
            read = read + bytes(synth.random_bases(rng, m - len(read)))
        qual = bytes(rng.integers(2, 42, m).astype(np.uint8))
        b = nref if band is None else band
        hint = p + int(rng.integers(-5, 6)) if band is not None else p
        jobs.append((read, qual, bytes(ref), b, hint))
    g = P.Genome().add_contig("x", "ACGT" * 10)
    ctx = P.Context(g)
    got = ctx.align_batch(jobs)
    for i, (read, qual, ref, band, hint) in enumerate(jobs):
        s, c, o, _ = O.SmithWatermanAligner(band, 40, 3).align(read, qual, ref, hint)
        assert got[i] == (s, c, o), "job %d (m=%d n=%d band=%d hint=%d)" % (i, len(read), len(ref), band, hint)
