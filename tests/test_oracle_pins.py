"""Pins of the CPU oracle beyond the reference's own known answers.

Each check is an independent computation (brute force, textbook recurrence,
closed form) rather than the oracle's formula retyped, chosen so that a
plausible slip in the oracle (a transposed index, a dropped term, a wrong tie
order) fails one of them.
"""
import math
import os
import tempfile

import numpy as np
import pytest

from oracle import pbw_oracle as O
from probebwa_amd import synth
from refcases import Xorshift64, revcomp

NEG = -(1 << 30)  # i32::MIN / 2


# ---------------------------------------------------------------------------
# DP: independent full-matrix Gotoh (smith_waterman.rs:163-345 semantics,
# implemented from the recurrence with explicit per-cell band membership)
# ---------------------------------------------------------------------------
def mis_pen(q):
    """round(-30 * clamp(q/30, 0.2, 1)) in closed form."""
    return -min(max(q, 6), 30)


def brute_align(read, qual, ref, band, hint, go=-90, ge=-1):
    m, n = len(read), len(ref)
    if m == 0 or n == 0:
        return 0, [(4, m)], 0
    run = [0] * n
    for x in range(n):
        a = x
        while a > 0 and ref[a - 1] == ref[x]:
            a -= 1
        b = x
        while b + 1 < n and ref[b + 1] == ref[x]:
            b += 1
        run[x] = b - a + 1

    def gop(j):
        return min(go + (run[j - 1] - 1), -1)

    def in_band(i, j):
        lo = max(i + hint - band, 1)
        hi = max(min(i + hint + band, n), 0)
        return lo <= j <= hi

    rows = [i for i in range(1, m + 1) if max(i + hint - band, 1) <= max(min(i + hint + band, n), 0)]
    M = {0: [0] * (n + 1)}
    I = {0: [NEG] * (n + 1)}
    D = {0: [NEG] * (n + 1)}
    tb = {}
    prev = 0
    for i in rows:
        M[i], I[i], D[i] = [NEG] * (n + 1), [NEG] * (n + 1), [NEG] * (n + 1)
        q = qual[i - 1] if i - 1 < len(qual) else 30
        for j in range(1, n + 1):
            if not in_band(i, j):
                continue
            sc = 20 if read[i - 1] == ref[j - 1] else mis_pen(q)
            cands = [(M[prev][j - 1] + sc, 0), (I[prev][j - 1] + sc, 1), (D[prev][j - 1] + sc, 2)]
            best = max(c[0] for c in cands)
            src = next(s for v, s in cands if v == best)  # first = M > I > D preference
            io, ie = M[prev][j] + gop(j), I[prev][j] + ge
            do, de = M[i][j - 1] + gop(j), D[i][j - 1] + ge
            M[i][j], I[i][j], D[i][j] = best, max(io, ie), max(do, de)
            tb[(i, j)] = (src, 0 if I[i][j] == io else 1, 0 if D[i][j] == do else 1)
        prev = i
    last = prev
    best = None
    for j in range(n + 1):
        for s, v in ((0, M[last][j]), (1, I[last][j])):
            if best is None or v >= best[0]:
                best = (v, j, s)
    score, j, state = best
    ops, i = [], m
    while i > 0:
        if j == 0:
            ops.append(1)
            i -= 1
            continue
        t = tb.get((i, j), (0, 0, 0))
        if state == 0:
            ops.append(0)
            i, j, state = i - 1, j - 1, t[0]
        elif state == 1:
            ops.append(1)
            i, state = i - 1, (0 if t[1] == 0 else 1)
        else:
            ops.append(2)
            j, state = j - 1, (0 if t[2] == 0 else 2)
    ops.reverse()
    cig = []
    for o in ops:
        if cig and cig[-1][0] == o:
            cig[-1][1] += 1
        else:
            cig.append([o, 1])
    return score, [tuple(c) for c in cig], j


def test_mismatch_penalty_closed_form():  # smith_waterman.rs:272-275
    # every byte a quality can hold: the DP kernel uses -clamp(q, 6, 30) directly
    for q in range(0, 256):
        want = -6 if q <= 6 else (-q if q <= 30 else -30)
        assert O.mismatch_penalty(q) == want, q


def test_homopolymer_gap_open_closed_form():  # smith_waterman.rs:288-291
    for run in range(0, 200):
        assert O.homopolymer_gap_open(-90, run) == min(-90 + max(run - 1, 0), -1)


def test_dp_matches_independent_full_matrix_gotoh():
    rng = np.random.default_rng(5)
    alpha = b"ACGTN"
    for trial in range(1500):
        m = int(rng.integers(0, 12))
        n = int(rng.integers(0, 20))
        probs = [0.24, 0.24, 0.24, 0.24, 0.04]
        read = bytes(rng.choice(list(alpha), m, p=probs))
        ref = bytes(rng.choice(list(alpha[:2] if trial % 3 == 0 else alpha), n))
        qual = bytes(rng.integers(0, 45, m if rng.random() < 0.9 else max(0, m - 2)))
        band = int(rng.integers(0, 12))
        hint = int(rng.integers(-15, 25))
        s, c, off, _ = O.SmithWatermanAligner(band).align(read, qual, ref, hint)
        bs, bc, boff = brute_align(read, qual, ref, band, hint)
        assert (s, [(w & 15, w >> 4) for w in c], off) == (bs, bc, boff), (read, qual, ref, band, hint)


def test_dp_perfect_match_closed_form():
    rng = np.random.default_rng(6)
    for _ in range(50):
        m = int(rng.integers(20, 120))  # 61L-210 holds once the band is no longer clipped at j=1 (L >= 20)
        ref = synth.random_bases(rng, m + 40).tobytes()
        read = ref[10:10 + m]
        s, c, off, cells = O.SmithWatermanAligner(30).align(read, bytes([30] * m), ref, 10)
        assert (s, O.cigar_to_str(c), off) == (20 * m, "%dM" % m, 10)
        assert cells == 61 * m - 210  # unclipped window: cell count of the reference loop


# ---------------------------------------------------------------------------
# seeding + clustering: naive re-derivation on tiny genomes
# ---------------------------------------------------------------------------
def naive_table(contigs, k=15, stride=5, cap=300):
    table, counts = {}, {}
    for cid, seq in enumerate(contigs):
        L = len(seq)
        rc = revcomp(seq)
        scans = [(seq[i:i + k], i, 0) for i in range(0, L - k + 1, stride)] + \
                [(rc[i:i + k], L - i - k, 1) for i in range(0, L - k + 1, stride)]
        for kmer, pos, strand in scans:
            if any(ch not in "ACGT" for ch in kmer):
                continue
            if counts.get(kmer, 0) < cap:
                table.setdefault(kmer, []).append((cid, pos, strand))
                counts[kmer] = counts.get(kmer, 0) + 1
    return table


def comp_counts(s):
    return [s.count(ch) for ch in "ACGT"]


def naive_hits(read, qual, contigs, table):
    k, L = 15, len(read)
    rc = revcomp(read)
    hits, rep = [], [0.0] * max(0, L - k + 1)

    def passes(p, cid, pos, strand, nb):
        g = contigs[cid]
        local, ws = (read, p) if strand == 0 else (rc, L - p - k)
        r_after = local[ws + k: ws + k + 8]
        r_before = local[max(0, ws - 4): ws]
        g_after = g[pos + k: min(pos + k + 8, len(g))]
        g_before = g[max(0, pos - 4): pos]
        al, bl = min(len(r_after), len(g_after)), min(len(r_before), len(g_before))
        if al + bl < 12:
            return True
        rf = [a + b for a, b in zip(comp_counts(r_after[:al]), comp_counts(r_before[len(r_before) - bl:]))]
        gf = [a + b for a, b in zip(comp_counts(g_after[:al]), comp_counts(g_before[len(g_before) - bl:]))]
        return sum(abs(a - b) for a, b in zip(rf, gf)) <= (8 if nb else 10)

    for p in range(0, L - k + 1):
        w = read[p:p + k]
        if any(ch not in "ACGT" for ch in w):
            continue
        ex = table.get(w, [])
        hits += [(p, c, pos, s, 0) for c, pos, s in ex if passes(p, c, pos, s, False)]
        if len(ex) >= 64:
            rep[p] = 1.0
            continue
        probed = L <= 34 or (L <= 49 and (p // 5) % 2 == 0) or (L > 49 and p % 15 < 5)
        if not probed:
            continue
        q = qual[p] if p < len(qual) else 30
        for i in range(k):
            for b in "ACGT":
                if b == w[i]:
                    continue
                nbk = w[:i] + b + w[i + 1:]
                nb = table.get(nbk, [])
                if len(nb) >= 64:
                    rep[p] = max(rep[p], 10 ** (-q / 10.0))
                hits += [(p, c, pos, s, 1) for c, pos, s in nb if passes(p, c, pos, s, True)]
    return hits, rep


def naive_clusters(hits, L, order):
    groups = {}
    for p, c, pos, s, _ in hits:
        groups.setdefault((c, s), []).append(pos - p if s == 0 else pos + p)
    out = []
    for key in order(list(groups)):
        d = sorted(groups[key])
        start = 0
        for idx in range(1, len(d) + 1):
            if idx == len(d) or d[idx] - d[idx - 1] > 50:
                g = d[start:idx]
                avg = sum(g) // len(g) if sum(g) >= 0 else -((-sum(g)) // len(g))
                pos = avg if key[1] == 0 else avg - L + 15
                out.append((key[0], max(pos, 0), key[1], len(g)))
                start = idx
    out.sort(key=lambda t: -t[3])  # stable
    return out


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_seed_hits_and_clusters_match_naive_scan(seed):
    rng = np.random.default_rng(seed)
    unit = synth.random_bases(rng, 60).tobytes().decode()
    contigs = []
    for c in range(2):  # one contig: FxHashMap order = forward group then reverse
        s = list(synth.random_bases(rng, 3000).tobytes().decode())
        for _ in range(70):  # repeat copies push some buckets past 64
            p = int(rng.integers(0, 2900))
            s[p:p + 60] = list(synth.mutate(rng, np.frombuffer(unit.encode(), np.uint8), 0.02).tobytes().decode())
        p = int(rng.integers(0, 2900))
        s[p:p + 10] = ["N"] * 10
        contigs.append("".join(s[:3000]))
    contigs = contigs[:1]
    g = O.Genome()
    for i, s in enumerate(contigs):
        g.add_contig("c%d" % i, s)
    m = O.Mapper(g, O.HashTable(g))
    table = naive_table(contigs)
    for t in range(60):
        L = int(rng.integers(15, 130))
        s = contigs[0]
        st = int(rng.integers(0, len(s) - L))
        read = s[st:st + L]
        read = synth.mutate(rng, np.frombuffer(read.encode(), np.uint8), 0.04).tobytes().decode()
        if rng.random() < 0.5:
            read = revcomp(read)
        qual = bytes(rng.integers(2, 42, L))
        hits, rp, cands = m.seed_and_cluster(read, qual)
        nh, nrp = naive_hits(read, qual, contigs, table)
        assert [tuple(int(x) for x in h) for h in hits] == nh
        assert list(rp) == nrp
        want = naive_clusters(nh, L, lambda keys: sorted(keys))
        assert [tuple(int(x) for x in c) for c in cands] == want


# ---------------------------------------------------------------------------
# MAPQ pieces: closed forms
# ---------------------------------------------------------------------------
def test_phred_n_choose_m_matches_lgamma():  # bayesian.rs:72-86
    for n in range(0, 300, 7):
        for m in range(0, n + 1, 3):
            exact = 10 * (math.lgamma(n + 1) - math.lgamma(m + 1) - math.lgamma(n - m + 1)) / math.log(10)
            assert abs(O.phred_n_choose_m(n, m) - exact) <= 0.5 + 1e-9


def test_effective_alphabet_size_closed_form():  # bayesian.rs:463-466
    assert abs(O.effective_alphabet_size(0.5) - 4.0) < 1e-12
    for f in np.linspace(0.01, 0.99, 37):
        want = math.exp(-f * math.log(f / 2) - (1 - f) * math.log((1 - f) / 2))
        assert abs(O.effective_alphabet_size(f) - want) < 1e-12 * want
        assert O.effective_alphabet_size(f) <= 4.0 + 1e-12


def test_missing_locus_closed_form():  # bayesian.rs:405-439 (uniform quality)
    for L, q in ((40, 40), (50, 40), (120, 30), (90, 20)):
        e = O.logadd(q, 30.0)
        p = min(10 ** (-e / 10), 1 - 1e-12)
        win = 1 - (1 - p) ** 15 - 15 * p * (1 - p) ** 14
        want = sum(win ** ((L - ph) // 15) for ph in range(5)) / 5
        got = O.missing_locus_probability(bytes([q] * L))
        assert abs(got - min(want, 0.99)) <= 1e-9 * max(want, 1e-300) + 1e-300
    assert O.missing_locus_probability(bytes([40] * 14)) == 0.5


def test_logadd_table():  # bayesian.rs:47-58
    assert O.logadd(30, 45) == 30 and O.logadd(45, 30) == 30
    assert O.logadd(30, 36) == 29 and O.logadd(30, 33) == 28 and O.logadd(30, 31) == 27
    assert O.logadd(1, 2) == 0


def test_insert_model_mean_is_running_mean():  # insert_size.rs:32-44
    rng = np.random.default_rng(9)
    obs = rng.normal(350, 50, 500)
    mean, sd, n = 250.0, 60.0, 0
    for x in obs:
        mean, sd, n = O.insert_update(mean, sd, n, float(x))
        assert sd >= 1.0
    assert n == 500 and abs(mean - obs.mean()) < 1e-9
    mean, sd, n = O.insert_update(250.0, 60.0, 0, 400.0)  # the first update replaces the mean
    assert (mean, sd, n) == (400.0, 60.0, 1)
    ll = O.insert_log_likelihood(350.0, 50.0, 400.0)
    assert abs(ll - (-0.5 - 0.5 * math.log(2 * math.pi * 2500))) < 1e-12


def test_mapq_80_and_99_closed_forms():
    """SURVEY 8.3 derivations: 40 bp Q40 unique read -> 80; 50 bp in 80 kb -> 99."""
    for L, gsize, want in ((40, 400, 80), (50, 80000, 99)):
        rng = np.random.default_rng(L)
        q = bytes([40] * L)
        seq = synth.random_bases(rng, L).tobytes()
        import ctypes as C
        cig = np.array([(L << 4) | 0], np.uint32)
        mapq, post = C.c_uint8(), C.c_double()
        O.lib().orc_finalize_mapq(0.001, gsize, 1.0, q, L, cig.ctypes.data, 1, seq, L, q, L, seq, L, None, 0,
                                  C.byref(mapq), C.byref(post))
        assert mapq.value == want


# ---------------------------------------------------------------------------
# FASTQ/FASTA/SAM (fastx.rs, sam.rs; tests/integration.rs:66-113, tests/cli.rs)
# ---------------------------------------------------------------------------
def test_fastq_empty_sequence_record_does_not_desync_parser():
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "e.fq")
        open(p, "w").write("@empty\n\n+\n\n@normal\nACGT\n+\nIIII\n")
        r = O.read_reads(p)
        assert r.ids == ["empty", "normal"] and r.seq_of(0) == b"" and r.seq_of(1) == b"ACGT"
        assert list(r.qual_of(1)) == [40, 40, 40, 40]


def test_fasta_multi_record_and_mate_suffix():
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "m.fa")
        open(p, "w").write(">r1/1 x\nACGT\n>r2/2\nTTTT\n>r3\nGGGG\n>r4\nCCCC\n")
        r = O.read_reads(p, fasta=True)
        assert r.ids == ["r1", "r2", "r3", "r4"] and list(r.qual_of(0)) == [40] * 4
        gp = os.path.join(d, "g.fa")
        open(gp, "w").write(">c1\nAAAA\n>c2\nCCCC\n>c3\nGGGG\n>c4\nTTTT\n")
        g = O.Genome().add_fasta(gp)
        assert g.names == ["c1", "c2", "c3", "c4"]


def test_adapter_strip_and_sam_fields():  # tests/cli.rs:113-146
    seq = Xorshift64(101).gen_seq(400)
    adapter = "AGATCGGAAGAGCACACGTC"
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "r.fq")
        read = seq[40:80] + adapter
        open(p, "w").write("@r1\n%s\n+\n%s\n" % (read, "I" * len(read)))
        rb = O.read_reads(p, adapter=adapter)
        g = O.Genome().add_contig("chr1", seq)
        recs = O.Mapper(g, O.HashTable(g)).map_se(rb)
        line = [l for l in O.format_sam(g, rb, recs).decode().split("\n") if l.startswith("r1\t")][0]
        f = line.split("\t")
        assert f[3] == "41" and f[5] == "40M" and len(f[9]) == 40


# ---------------------------------------------------------------------------
# log_entropy (bayesian.rs:306-373): the random-match entropy from the counts
# (n HQ aligned bases, m HQ mismatches, read length consumed, insertion and
# deletion runs) with exact integer binomials.  The reference counts a "run"
# at every change of operation type, a return to M after an indel included
# (bayesian.rs:323-346), so 20M 2I 20M has two non-deletion runs; the counts
# of each hand case below are derived from the code by hand and written out.
# ---------------------------------------------------------------------------
LN10 = math.log(10.0)


def pyround(x):
    return math.trunc(x + 0.5)


def phred_choose_exact(n, m):
    """py_int_round(10 log10 C(n, m)) from the exact binomial; 0 when the reduced m <= 0 (bayesian.rs:72-86)."""
    m = n - m if m > n // 2 else m
    if m <= 0:
        return 0
    return pyround(10.0 * math.log10(math.comb(n, m)))


def entropy_from_counts(n, m, readlen, ins_runs, del_runs, window, gsize):
    gc = sum(window.upper().count(c) for c in "GC")
    total = gc + sum(window.upper().count(c) for c in "AT")
    f = 0.5 if total == 0 else gc / total
    f = min(max(f, 1e-6), 1.0 - 1e-6)
    b = math.exp(-f * math.log(f / 2.0) - (1.0 - f) * math.log((1.0 - f) / 2.0))
    match = pyround(10.0 * (n * math.log(b) - math.log(2.0 * gsize)) / LN10)
    mis = pyround(10.0 * m * math.log(max(b - 1.0, 1e-6)) / LN10)
    dterm = pyround(del_runs * 10.0 * math.log(30.0) / LN10) if del_runs else 0
    return match - mis - phred_choose_exact(n, m) - phred_choose_exact(readlen, 2 * ins_runs + del_runs) - dterm


def words(*ops):
    code = {"M": 0, "I": 1, "D": 2, "S": 4}
    return [(n << 4) | code[o] for n, o in ops]


def _mut(s, at, to):
    s = list(s)
    for i in at:
        s[i] = to if s[i] != to else ("C" if to == "A" else "A")
    return "".join(s)


def test_log_entropy_hand_cases_closed_form():
    gsize = 2_000_000
    w40 = Xorshift64(31).gen_seq(43)
    ref40 = w40[:40]
    q40s = bytes([40] * 40)
    cases = []
    # 40M perfect: n 40, m 0, readlen 40
    cases.append((words((40, "M")), ref40, q40s, ref40, (40, 0, 40, 0, 0)))
    # mismatches: HQ at 3, 10, 17; read N at 20 (not counted); Q5 mismatches at 25, 33 (not HQ: off n and m)
    q = bytearray(q40s)
    q[25] = q[33] = 5
    rd = _mut(ref40, (3, 10, 17, 25, 33), "T")
    rd = rd[:20] + "N" + rd[21:]
    cases.append((words((40, "M")), rd, bytes(q), ref40, (38, 3, 40, 0, 0)))
    # 20M 2I 20M: runs I (entry), M (return): two non-deletion runs; read length 42
    ins = ref40[:20] + "GG" + ref40[20:]
    cases.append((words((20, "M"), (2, "I"), (20, "M")), ins, bytes([40] * 42), ref40, (40, 0, 42, 2, 0)))
    # 20M 3D 20M: D (entry, a deletion run) then M (return): one deletion run, one other run
    dele = w40[:20] + w40[23:43]
    cases.append((words((20, "M"), (3, "D"), (20, "M")), dele, q40s, w40, (40, 0, 40, 1, 1)))
    # 10M 2D 2I 10M: D, I, M changes: one deletion run, two others; read 22, window 22
    rd5 = ref40[:10] + "AC" + ref40[12:22]
    cases.append((words((10, "M"), (2, "D"), (2, "I"), (10, "M")), rd5, bytes([40] * 22), ref40[:22], (20, 0, 22, 2, 1)))
    # 2I 38M: a leading insertion is a change from the initial M, then the return to M
    cases.append((words((2, "I"), (38, "M")), "TT" + ref40[:38], q40s, ref40[:38], (38, 0, 40, 2, 0)))
    # window shorter than the alignment: M columns stop at the window's end (n 30, read consumed 30)
    cases.append((words((40, "M")), ref40, q40s, ref40[:30], (30, 0, 30, 0, 0)))
    # qualities shorter than the read: missing qualities read 0 (not HQ)
    cases.append((words((40, "M")), ref40, bytes([40] * 25), ref40, (25, 0, 40, 0, 0)))
    # GC-rich window (32 of 40 G/C): b < 4 moves the match and mismatch terms
    gcw = "GCGCCGGCAT" * 4
    gcr = _mut(gcw, (5, 15), "A")
    cases.append((words((40, "M")), gcr, q40s, gcw, (40, 2, 40, 0, 0)))
    # soft clips are skipped entirely
    cases.append((words((5, "S"), (35, "M")), ref40[:35], q40s, ref40[:35], (35, 0, 35, 0, 0)))
    seen = set()
    for cig, rd, qual, win, counts in cases:
        got = O.log_entropy(cig, rd, qual, win, gsize)
        assert got == entropy_from_counts(*counts, win, gsize), (cig, counts, got)
        seen.add(got)
    assert len(seen) >= 8  # the cases move different terms


def _entropy_counts_by_columns(cig, rd, qual, win):
    """Counts from the expanded operation string: runs are the type changes of the I/D/M sequence
    (starting from M); M columns count until the first that reaches past the read or the window."""
    ops = "".join({0: "M", 1: "I", 2: "D"}.get(w & 15, "") * (w >> 4) for w in cig)
    prev = ["M"] + list(ops[:-1]) if ops else []
    changes = [o for o, p in zip(ops, prev) if o != p]
    del_runs = changes.count("D")
    r = p = n = m = 0
    stopped = False
    for o in ops:
        if o == "I":
            r += 1
        elif o == "D":
            p += 1
        else:
            if stopped or r >= len(rd) or p >= len(win):
                stopped = True
                continue
            hq = (qual[r] if r < len(qual) else 0) >= 10
            n += hq
            m += hq and rd[r] != win[p] and rd[r] != "N" and win[p] != "N"
            r += 1
            p += 1
    return n, m, r, len(changes) - del_runs, del_runs


def test_log_entropy_random_alignments_match_closed_form():
    rng = np.random.default_rng(17)
    checked = 0
    for _ in range(1500):
        cig, L = [], 0
        for _k in range(int(rng.integers(1, 6))):
            op = int(rng.choice([0, 0, 0, 1, 2, 4]))
            ln = int(rng.integers(1, 30))
            cig.append((ln << 4) | op)
            L += ln if op in (0, 1, 4) else 0
        rd = "".join(rng.choice(list("ACGTN"), L, p=[0.24, 0.24, 0.24, 0.24, 0.04]))
        win = "".join(rng.choice(list("ACGTN"), int(rng.integers(0, L + 20)), p=[0.3, 0.2, 0.2, 0.28, 0.02]))
        qual = bytes(rng.integers(0, 42, int(rng.integers(0, L + 3))).tolist())
        counts = _entropy_counts_by_columns(cig, rd, qual, win)
        want = entropy_from_counts(*counts, win, 5_000_000)
        assert O.log_entropy(cig, rd, qual, win, 5_000_000) == want, (cig, counts)
        checked += 1
    assert checked == 1500


# ---------------------------------------------------------------------------
# missing_locus_probability with a repeat mask (bayesian.rs:405-439): per
# window a probabilistic OR of P(>= 2 errors in 15) and the mask at the
# window's start, product over a phase's windows, mean over 5 phases
# ---------------------------------------------------------------------------
def test_missing_locus_with_repeat_mask_closed_form():
    L = 50
    e = 0.01  # Q20 with a substitution-rate Phred of 30: logadd(20, 30) = 20
    p2 = 1.0 - (1.0 - e) ** 15 - 15 * e * (1.0 - e) ** 14
    mask = np.zeros(L - 14)
    mask[15], mask[31], mask[4], mask[19], mask[34] = 0.5, 1.0, 0.2, 0.9, 0.05
    phases = []
    for ph in range(5):
        prod = 1.0
        for start in range(ph, L - 15 + 1, 15):
            prod *= min(max(1.0 - (1.0 - p2) * (1.0 - mask[start]), 1e-300), 1.0)
        phases.append(prod)
    want = min(sum(phases) / 5.0, 0.99)
    got = O.missing_locus_probability(bytes([20] * L), mask)
    assert abs(got - want) <= 1e-12 * want
    assert abs(got - O.missing_locus_probability(bytes([20] * L))) > 1e-3 * want  # the mask matters
    shifted = np.zeros(L - 14)
    shifted[[16, 32, 5, 20, 35]] = mask[[15, 31, 4, 19, 34]]  # the mask one base later (window end-ish)
    assert abs(O.missing_locus_probability(bytes([20] * L), shifted) - want) > 1e-3 * want


# ---------------------------------------------------------------------------
# shortlist_indices (paired.rs:553-566): the shortest prefix of the stable
# descending order that is 20 long, or at least 3 long with mass >= 0.999
# ---------------------------------------------------------------------------
def test_shortlist_indices_brute_force():
    rng = np.random.default_rng(5)
    for trial in range(3000):
        n = int(rng.integers(0, 41))
        if trial % 3 == 0:  # ties
            post = rng.choice([0.0, 0.1, 0.25, 0.5], n)
        else:
            post = rng.dirichlet(np.full(max(n, 1), float(rng.choice([0.05, 0.5, 5.0]))))[:n]
        order = sorted(range(n), key=lambda i: (-post[i], i))
        take = n
        for t in range(n + 1):
            if t == 20 or (t >= 3 and sum(post[order[k]] for k in range(t)) >= 0.999):
                take = t
                break
        assert O.shortlist_indices(post) == order[:take], (trial, list(post))


# ---------------------------------------------------------------------------
# joint_map's dedup of raw pairs (paired.rs:346-351): first occurrence of each
# (mate-1 locus, mate-2 locus), order kept
# ---------------------------------------------------------------------------
def test_pair_dedup_keeps_first_occurrence():
    rng = np.random.default_rng(9)
    for _ in range(500):
        pool = [(int(rng.integers(0, 3)), int(rng.integers(0, 3)), int(rng.integers(0, 2))) for _ in range(4)]
        raw = [(pool[int(rng.integers(0, 4))], pool[int(rng.integers(0, 4))]) for _ in range(int(rng.integers(0, 12)))]
        first = {}
        for i, pr in enumerate(raw):
            first.setdefault(pr, i)
        assert O.dedup_pairs(raw) == sorted(first.values())


# ---------------------------------------------------------------------------
# mapq_for_winning_pair (paired.rs:398-435): both mates own-best -> each its
# own finalize; exactly one -> the anchor's (MAPQ, posterior) for both;
# neither -> each finalized with the pair's sum of exponentials
# ---------------------------------------------------------------------------
def test_mapq_for_winning_pair_branches():
    genome = Xorshift64(77).gen_seq(3000)
    g = O.Genome().add_contig("c", genome)
    mapper = O.Mapper(g, O.HashTable(g))
    gsize = 3000

    def mate(at, lk):
        s = genome[at:at + 50]
        return dict(seq=s, qual=bytes([30] * 50), cigar=words((50, "M")), ref_window=s, lk=lk, mask=())

    a = mate(100, [-20.0, -45.0, -60.5])
    b = mate(900, [-30.0, -31.0])

    def fin(se, d):
        return O.finalize_mapq(se, d["qual"], d["cigar"], d["seq"], d["qual"], d["ref_window"], gsize)

    def se(lk):
        mx = max(lk)
        return sum(math.exp(x - mx) for x in lk)

    pair_se = 3.0
    want = {
        (True, True): (fin(se(a["lk"]), a), fin(se(b["lk"]), b)),
        (True, False): (fin(se(a["lk"]), a),) * 2,
        (False, True): (fin(se(b["lk"]), b),) * 2,
        (False, False): (fin(pair_se, a), fin(pair_se, b)),
    }
    for (oa, ob), w in want.items():
        assert O.mapq_for_winning_pair(mapper, pair_se, oa, ob, a, b) == w, (oa, ob)
    assert len(set(want.values())) == 4  # the branches give four different outcomes here
