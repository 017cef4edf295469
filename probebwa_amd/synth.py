"""Deterministic synthetic genomes and reads for the benchmark workloads.

Generator: numpy PCG64 seeded per workload (recorded in each function), so a
workload is reproducible byte for byte on any machine.

Workloads (the reference publishes no dataset for its mapping path):
  probe_se_2mbp      2,000,000 bp uniform chromosome + 50 copies of a 2 kb unit
                     at 1% divergence; 20,000 x 120 bp reads, 1% substitutions,
                     0.05% 1-3 bp indel events, qualities U{20..40}.
  uneven_se_100mbp   10 contigs x 10 Mbp: uniform background, 5% segmental
                     duplications (1-10 kb copies at 0.5-3% divergence),
                     homopolymer runs of 5-15 bp every ~2 kb; 2M reads with
                     L ~ U{80..150}, Q(pos) = 38 - 16 pos/(L-1) + N(0,3)
                     clamped to [2,41], substitutions at the Phred rate.
  pairs_2x150        pairs on the probe_se_2mbp genome: fragment ~ N(350,50)
                     clipped to [200,600], FR orientation, 10% of pairs with
                     one mate inside a repeat copy.
"""
import concurrent.futures
import multiprocessing
import os

import numpy as np

BASES = np.frombuffer(b"ACGT", dtype=np.uint8)
COMP = np.zeros(256, np.uint8)
COMP[:] = ord("N")
for a, b in zip(b"ACGT", b"TGCA"):
    COMP[a] = b


def revcomp(a):
    return COMP[a[::-1]]


def random_bases(rng, n):
    return BASES[rng.integers(0, 4, n)]


def mutate(rng, seq, rate):
    """Substitute each base with probability `rate` by a different base."""
    s = seq.copy()
    hit = np.nonzero(rng.random(len(s)) < rate)[0]
    if len(hit):
        codes = np.searchsorted(BASES, s[hit])
        s[hit] = BASES[(codes + rng.integers(1, 4, len(hit))) % 4]
    return s


def genome_probe_se_2mbp(seed=1):
    rng = np.random.default_rng(np.random.PCG64(seed))
    bg = random_bases(rng, 2_000_000)
    unit = random_bases(rng, 2000)
    pieces, last = [], 0
    for p in np.sort(rng.choice(2_000_000, 50, replace=False)):
        pieces.append(bg[last:p])
        pieces.append(mutate(rng, unit, 0.01))
        last = p
    pieces.append(bg[last:])
    return [("chr1", np.concatenate(pieces))]


def probe_repeat_spans(seed=1):
    """[start, end) of the 50 inserted 2 kb copies in genome_probe_se_2mbp(seed): the
    same draws replayed up to the insertion points (copy k lands after k earlier copies)."""
    rng = np.random.default_rng(np.random.PCG64(seed))
    random_bases(rng, 2_000_000)
    random_bases(rng, 2000)
    pos = np.sort(rng.choice(2_000_000, 50, replace=False))
    return [(int(p) + 2000 * k, int(p) + 2000 * k + 2000) for k, p in enumerate(pos)]


def genome_uneven_100mbp(seed=2, n_contigs=10, contig_len=10_000_000):
    rng = np.random.default_rng(np.random.PCG64(seed))
    contigs = []
    for c in range(n_contigs):
        contigs.append(random_bases(rng, contig_len))
    total = n_contigs * contig_len
    # segmental duplications: copy 1-10 kb segments (0.5-3% divergence) until 5% of the genome
    dup = 0
    while dup < total * 0.05:
        ln = int(rng.integers(1000, 10001))
        sc, dc = int(rng.integers(0, n_contigs)), int(rng.integers(0, n_contigs))
        sp, dp = int(rng.integers(0, contig_len - ln)), int(rng.integers(0, contig_len - ln))
        contigs[dc][dp:dp + ln] = mutate(rng, contigs[sc][sp:sp + ln], float(rng.uniform(0.005, 0.03)))
        dup += ln
    # homopolymer runs of 5-15 bp every ~2 kb
    for c in range(n_contigs):
        pos = 0
        while True:
            pos += int(rng.integers(1000, 3001))
            ln = int(rng.integers(5, 16))
            if pos + ln >= contig_len:
                break
            contigs[c][pos:pos + ln] = BASES[rng.integers(0, 4)]
    return [("chr%d" % (c + 1), contigs[c]) for c in range(n_contigs)]


def _mutated_copies(rng, consensus, n, dmin, dmax):
    """n copies of `consensus` with per-copy substitution divergence U[dmin, dmax], half of
    them reverse-complemented.  Returns an (n, len) uint8 array."""
    L = len(consensus)
    out = np.empty((n, L), np.uint8)
    for c0 in range(0, n, 65536):
        c1 = min(n, c0 + 65536)
        m = c1 - c0
        x = np.broadcast_to(consensus, (m, L)).copy()
        d = rng.uniform(dmin, dmax, m)
        err = rng.random((m, L)) < d[:, None]
        codes = np.searchsorted(BASES, x)
        x = np.where(err, BASES[(codes + rng.integers(1, 4, (m, L), dtype=np.uint8)) % 4], x)
        rc = rng.random(m) < 0.5
        x[rc] = COMP[x[rc, ::-1]]
        out[c0:c1] = x
    return out


def _genome_scale_contig(args):
    """Contig `ci` of genome_scale_3gbp: its own generator (child `ci` of the workload seed)
    and the two repeat consensi (drawn from the workload seed itself)."""
    seed, ci, L = args
    rng0 = np.random.default_rng(np.random.PCG64(seed))
    fam_s = random_bases(rng0, 300)
    fam_l = random_bases(rng0, 6000)
    rng = np.random.default_rng(np.random.PCG64(np.random.SeedSequence(seed).spawn(ci + 1)[ci]))
    g = BASES[rng.integers(0, 4, L, dtype=np.uint8)]
    ns = int(0.225 * L / 300)
    nl = int(0.225 * L / 6000)
    lens = np.concatenate([np.full(ns, 300, np.int64), np.full(nl, 6000, np.int64)])
    kind = np.concatenate([np.zeros(ns, np.int8), np.ones(nl, np.int8)])
    order = rng.permutation(len(lens))
    lens, kind = lens[order], kind[order]
    gaps = rng.exponential(1.0, len(lens) + 1)
    gaps = np.floor(gaps / gaps.sum() * (L - lens.sum())).astype(np.int64)
    starts = np.cumsum(gaps[:-1]) + np.concatenate([[0], np.cumsum(lens)[:-1]])
    for k, cons, dlo, dhi in ((0, fam_s, 0.10, 0.20), (1, fam_l, 0.05, 0.15)):
        sel = np.nonzero(kind == k)[0]
        cp = _mutated_copies(rng, cons, len(sel), dlo, dhi)
        idx = starts[sel][:, None] + np.arange(len(cons))[None, :]
        g[idx] = cp
    # segmental duplications: 5% of the contig
    done = 0
    while done < 0.05 * L:
        ln = int(rng.integers(1000, 10001))
        src = int(rng.integers(0, L - ln))
        dst = int(rng.integers(0, L - ln))
        seg = mutate(rng, g[src:src + ln].copy(), float(rng.uniform(0.005, 0.03)))
        g[dst:dst + ln] = seg
        done += ln
    return g


def genome_scale_3gbp(seed=5, n_contigs=24, total=3_100_000_000, workers=None):
    """Genome-scale synthetic reference: `n_contigs` contigs of equal length summing to
    `total` bases; a uniform background with 45% of the bases in interspersed repeats from
    two families (a 300 bp consensus whose copies diverge 10-20%, a 6 kb consensus at
    5-15%; half of the repeat bases each, both strands, in random order with exponential
    background gaps) and 5% in segmental duplications (copies of 1-10 kb segments of the
    same contig at 0.5-3% divergence).  Contig ci draws from child ci of the seed, so the
    contigs are built in parallel processes (spawned: safe after a GPU runtime has started)
    and the genome does not depend on the process count."""
    L = total // n_contigs
    jobs = [(seed, ci, L) for ci in range(n_contigs)]
    workers = workers or max(1, min(n_contigs, 16, os.cpu_count() or 1))
    if workers == 1:
        seqs = [_genome_scale_contig(j) for j in jobs]
    else:
        ctx = multiprocessing.get_context("spawn")
        with concurrent.futures.ProcessPoolExecutor(workers, mp_context=ctx) as ex:
            seqs = list(ex.map(_genome_scale_contig, jobs))
    return [("chr%d" % (ci + 1), g) for ci, g in enumerate(seqs)]


def _flat(seqs, quals):
    so = np.zeros(len(seqs) + 1, np.uint64)
    qo = np.zeros(len(quals) + 1, np.uint64)
    so[1:] = np.cumsum([len(s) for s in seqs])
    qo[1:] = np.cumsum([len(q) for q in quals])
    seq = np.concatenate(seqs + [np.zeros(1, np.uint8)]).astype(np.uint8)
    qual = np.concatenate(quals + [np.zeros(1, np.uint8)]).astype(np.uint8)
    return seq, so, qual, qo


def _indels(rng, s, rate):
    """0.05%-style per-base indel events of length U{1,2,3}, ins/del 50/50."""
    out, i = [], 0
    ev = np.nonzero(rng.random(len(s)) < rate)[0]
    for p in ev:
        out.append(s[i:p])
        ln = int(rng.integers(1, 4))
        if rng.random() < 0.5:
            out.append(random_bases(rng, ln))
            i = p
        else:
            i = min(len(s), p + ln)
    out.append(s[i:])
    return np.concatenate(out) if out else s


def reads_probe_se(contigs, n=20_000, L=120, seed=11):
    """20,000 x 120 bp reads: uniform start, strand 50/50, 1% subs, 0.05% indels, Q ~ U{20..40}."""
    rng = np.random.default_rng(np.random.PCG64(seed))
    g = contigs[0][1]
    seqs, quals, truth = [], [], []
    for _ in range(n):
        start = int(rng.integers(0, len(g) - L - 4))
        s = _indels(rng, g[start:start + L + 3], 0.0005)[:L]
        s = mutate(rng, s, 0.01)
        rev = bool(rng.random() < 0.5)
        if rev:
            s = revcomp(s)
        seqs.append(s)
        quals.append(rng.integers(20, 41, len(s)).astype(np.uint8))
        truth.append((0, start, rev))
    return _flat(seqs, quals), truth


def reads_uneven_se(contigs, n=2_000_000, seed=12, lmin=80, lmax=150, chunk=65536):
    """Uneven single-end reads: L ~ U{80..150}, Q(pos) = 38 - 16 pos/(L-1) + N(0,3)
    rounded and clamped to [2,41], substitutions at the Phred-implied rate,
    strand 50/50 (reverse reads are the reverse complement, quals reversed).
    Vectorised in chunks; returns flat (seq, seq_off, qual, qual_off)."""
    rng = np.random.default_rng(np.random.PCG64(seed))
    genome = np.concatenate([c[1] for c in contigs])
    lens = np.array([len(c[1]) for c in contigs], np.int64)
    starts = np.concatenate([[0], np.cumsum(lens)[:-1]])
    L = rng.integers(lmin, lmax + 1, n)
    cid = rng.choice(len(contigs), n, p=lens / lens.sum())
    st = (rng.random(n) * (lens[cid] - L)).astype(np.int64) + starts[cid]
    rev = rng.random(n) < 0.5
    so = np.zeros(n + 1, np.uint64)
    so[1:] = np.cumsum(L)
    seq = np.zeros(int(so[-1]) + 1, np.uint8)
    qual = np.zeros(int(so[-1]) + 1, np.uint8)
    col = np.arange(lmax)
    for c0 in range(0, n, chunk):
        c1 = min(n, c0 + chunk)
        Lc, sc, rc = L[c0:c1], st[c0:c1], rev[c0:c1]
        valid = col[None, :] < Lc[:, None]
        # forward-strand bases of the template; reverse reads read it backwards
        idx = np.where(rc[:, None], sc[:, None] + Lc[:, None] - 1 - col[None, :], sc[:, None] + col[None, :])
        idx = np.where(valid, idx, 0)
        b = genome[idx]
        b = np.where(rc[:, None], COMP[b], b)
        q = np.rint(38.0 - 16.0 * col[None, :] / (Lc[:, None] - 1) + rng.normal(0.0, 3.0, (c1 - c0, lmax)))
        q = np.clip(q, 2, 41).astype(np.uint8)
        err = rng.random((c1 - c0, lmax)) < np.power(10.0, -q.astype(np.float64) / 10.0)
        codes = np.searchsorted(BASES, b)
        b = np.where(err, BASES[(codes + rng.integers(1, 4, (c1 - c0, lmax))) % 4], b)
        o0 = int(so[c0])
        seq[o0:int(so[c1])] = b[valid]
        qual[o0:int(so[c1])] = q[valid]
    return seq, so, qual, so.copy()


def reads_probe_uneven_se(contigs, n=2_000_000, seed=31, lmin=80, lmax=150, chunk=65536):
    """Probe-style single-end reads of uneven length: L ~ U{lmin..lmax}, start uniform over
    the contigs, strand 50/50, 0.05% per-base indel events of length U{1,2,3} (ins/del 50/50),
    then 1% substitutions, qualities i.i.d. U{20..40}.  Vectorised in chunks (only the reads
    with an indel event take a per-read pass); returns flat (seq, seq_off, qual, qual_off)."""
    rng = np.random.default_rng(np.random.PCG64(seed))
    genome = np.concatenate([c[1] for c in contigs])
    lens = np.array([len(c[1]) for c in contigs], np.int64)
    starts = np.concatenate([[0], np.cumsum(lens)[:-1]])
    pad = 12  # template slack: deletions consume bases past the read's end
    L = rng.integers(lmin, lmax + 1, n)
    cid = rng.choice(len(contigs), n, p=lens / lens.sum())
    W = lmax + pad
    st = (rng.random(n) * (lens[cid] - W)).astype(np.int64) + starts[cid]  # template inside its contig
    rev = rng.random(n) < 0.5
    so = np.zeros(n + 1, np.uint64)
    so[1:] = np.cumsum(L)
    seq = np.zeros(int(so[-1]) + 1, np.uint8)
    qual = np.zeros(int(so[-1]) + 1, np.uint8)
    col = np.arange(W)
    for c0 in range(0, n, chunk):
        c1 = min(n, c0 + chunk)
        m = c1 - c0
        Lc, sc, rc = L[c0:c1], st[c0:c1], rev[c0:c1]
        t = genome[sc[:, None] + col[None, :]]
        ev = (rng.random((m, W)) < 0.0005) & (col[None, :] < Lc[:, None] + 3)
        for r in np.nonzero(ev.any(axis=1))[0]:
            parts, i = [], 0
            for p in np.nonzero(ev[r])[0]:
                if p < i:
                    continue
                parts.append(t[r, i:p])
                ln = int(rng.integers(1, 4))
                if rng.random() < 0.5:
                    parts.append(random_bases(rng, ln))
                    i = p
                else:
                    i = min(W, p + ln)
            parts.append(t[r, i:])
            row = np.concatenate(parts)
            if len(row) < W:
                row = np.concatenate([row, random_bases(rng, W - len(row))])
            t[r] = row[:W]
        b = t[:, :lmax]
        codes = np.searchsorted(BASES, b)
        err = rng.random((m, lmax)) < 0.01
        b = np.where(err, BASES[(codes + rng.integers(1, 4, (m, lmax))) % 4], b)
        # reverse reads are the reverse complement of the first L template bases
        ridx = np.where(rc[:, None], Lc[:, None] - 1 - col[None, :lmax], col[None, :lmax])
        valid = col[None, :lmax] < Lc[:, None]
        b = np.take_along_axis(b, np.where(valid, ridx, 0), axis=1)
        b = np.where(rc[:, None], COMP[b], b)
        q = rng.integers(20, 41, (m, lmax)).astype(np.uint8)
        o0 = int(so[c0])
        seq[o0:int(so[c1])] = b[valid]
        qual[o0:int(so[c1])] = q[valid]
    return seq, so, qual, so.copy()


def _probe_errors(rng, t, L):
    """Rows of template bases (each at least L + 12 long) -> rows of L bases under the
    probe read model: 0.05% per-base indel events of length U{1,2,3} (ins/del 50/50),
    then 1% substitutions."""
    m, W = t.shape
    ev = (rng.random((m, W)) < 0.0005) & (np.arange(W)[None, :] < L + 3)
    t = t.copy()
    for r in np.nonzero(ev.any(axis=1))[0]:
        parts, i = [], 0
        for p in np.nonzero(ev[r])[0]:
            if p < i:
                continue
            parts.append(t[r, i:p])
            ln = int(rng.integers(1, 4))
            if rng.random() < 0.5:
                parts.append(random_bases(rng, ln))
                i = p
            else:
                i = min(W, p + ln)
        parts.append(t[r, i:])
        row = np.concatenate(parts)
        if len(row) < W:
            row = np.concatenate([row, random_bases(rng, W - len(row))])
        t[r] = row[:W]
    b = t[:, :L]
    err = rng.random((m, L)) < 0.01
    codes = np.searchsorted(BASES, b)
    return np.where(err, BASES[(codes + rng.integers(1, 4, (m, L))) % 4], b)


def reads_pairs_probe(contigs, n=1_000_000, L=150, seed=3, repeat_spans=(), chunk=65536):
    """FR pairs on the probe genome: fragment ~ N(350, 50) rounded and clipped to
    [200, 600]; 10% of the pairs placed so that one mate lies inside a repeat copy;
    mates under the probe read model (indels, 1% substitutions, Q ~ U{20..40});
    mate 2 is the reverse complement of the fragment's end; mates swapped for half
    of the pairs.  Vectorised; returns ((seq, so, qual, qo) of mate 1, ... of mate 2)."""
    rng = np.random.default_rng(np.random.PCG64(seed))
    g = contigs[0][1]
    pad = 12
    frag = np.clip(np.rint(rng.normal(350, 50, n)), 200, 600).astype(np.int64)
    start = (rng.random(n) * (len(g) - frag - 1 - pad)).astype(np.int64)
    if repeat_spans:
        anch = np.nonzero(rng.random(n) < 0.1)[0]
        sp = np.asarray(repeat_spans, np.int64)[rng.integers(0, len(repeat_spans), len(anch))]
        off = (rng.random(len(anch)) * np.maximum(1, sp[:, 1] - sp[:, 0] - L)).astype(np.int64)
        start[anch] = np.clip(sp[:, 0] + off - frag[anch] + L, 0, len(g) - frag[anch] - 1 - pad)
    swap = rng.random(n) < 0.5
    so = np.arange(n + 1, dtype=np.uint64) * np.uint64(L)
    out = [np.zeros(n * L + 1, np.uint8) for _ in range(4)]
    col = np.arange(L + pad)
    for c0 in range(0, n, chunk):
        c1 = min(n, c0 + chunk)
        st, fr = start[c0:c1], frag[c0:c1]
        a = _probe_errors(rng, g[st[:, None] + col[None, :]], L)
        # mate 2: reverse complement of the fragment's last L bases; its template runs
        # backwards from the fragment end so deletions consume bases beyond its 5' end
        e = st + fr - 1
        tb = COMP[g[np.maximum(e[:, None] - col[None, :], 0)]]
        b2 = _probe_errors(rng, tb, L)
        sw = swap[c0:c1][:, None]
        m1, m2 = np.where(sw, b2, a), np.where(sw, a, b2)
        q1 = rng.integers(20, 41, (c1 - c0, L)).astype(np.uint8)
        q2 = rng.integers(20, 41, (c1 - c0, L)).astype(np.uint8)
        out[0][c0 * L:c1 * L] = m1.ravel()
        out[1][c0 * L:c1 * L] = q1.ravel()
        out[2][c0 * L:c1 * L] = m2.ravel()
        out[3][c0 * L:c1 * L] = q2.ravel()
    return (out[0], so, out[1], so.copy()), (out[2], so.copy(), out[3], so.copy())


def reads_pairs_genome(contigs, n=1_000_000, L=150, seed=58, chunk=65536):
    """FR pairs of 2 x L over a multi-contig genome (the genome-scale pairs): contig by
    length, fragment ~ N(350, 50) rounded and clipped to [200, 600] with a uniform start;
    mate 1 = the fragment's first L bases, mate 2 = the reverse complement of its last L;
    mates swapped for half of the pairs.  Each mate has the decaying quality model of
    reads_uneven_se (Q(pos) = 38 - 16 pos/(L-1) + N(0,3), clamped to [2,41]) and
    substitutions at the Phred-implied rate.  Returns ((seq, so, qual, qo) of mate 1, ... of mate 2)."""
    rng = np.random.default_rng(np.random.PCG64(seed))
    genome = np.concatenate([c[1] for c in contigs])
    lens = np.array([len(c[1]) for c in contigs], np.int64)
    starts = np.concatenate([[0], np.cumsum(lens)[:-1]])
    frag = np.clip(np.rint(rng.normal(350, 50, n)), 200, 600).astype(np.int64)
    cid = rng.choice(len(contigs), n, p=lens / lens.sum())
    st = (rng.random(n) * (lens[cid] - frag)).astype(np.int64) + starts[cid]
    swap = rng.random(n) < 0.5
    so = np.arange(n + 1, dtype=np.uint64) * np.uint64(L)
    out = [np.zeros(n * L + 1, np.uint8) for _ in range(4)]
    col = np.arange(L)

    def mate(b):
        m = b.shape[0]
        q = np.rint(38.0 - 16.0 * col[None, :] / (L - 1) + rng.normal(0.0, 3.0, (m, L)))
        q = np.clip(q, 2, 41).astype(np.uint8)
        err = rng.random((m, L)) < np.power(10.0, -q.astype(np.float64) / 10.0)
        codes = np.searchsorted(BASES, b)
        return np.where(err, BASES[(codes + rng.integers(1, 4, (m, L))) % 4], b), q

    for c0 in range(0, n, chunk):
        c1 = min(n, c0 + chunk)
        s, fr = st[c0:c1], frag[c0:c1]
        a, qa = mate(genome[s[:, None] + col[None, :]])
        b, qb = mate(COMP[genome[(s + fr - 1)[:, None] - col[None, :]]])
        sw = swap[c0:c1][:, None]
        out[0][c0 * L:c1 * L] = np.where(sw, b, a).ravel()
        out[1][c0 * L:c1 * L] = np.where(sw, qb, qa).ravel()
        out[2][c0 * L:c1 * L] = np.where(sw, a, b).ravel()
        out[3][c0 * L:c1 * L] = np.where(sw, qa, qb).ravel()
    return (out[0], so, out[1], so.copy()), (out[2], so.copy(), out[3], so.copy())


def reads_pairs(contigs, n=10_000, L=150, seed=13, repeat_spans=(), frag=(350, 50, 200, 600)):
    """FR pairs: fragment ~ N(mean, sd) clipped to [lo, hi] (default N(350, 50) in
    [200, 600]); 10% anchored in a repeat span."""
    rng = np.random.default_rng(np.random.PCG64(seed))
    g = contigs[0][1]
    fmean, fsd, flo, fhi = frag
    s1, q1, s2, q2 = [], [], [], []
    for i in range(n):
        flen = int(np.clip(np.rint(rng.normal(fmean, fsd)), flo, fhi))
        if repeat_spans and rng.random() < 0.1:
            a, b = repeat_spans[int(rng.integers(0, len(repeat_spans)))]
            start = int(np.clip(a + rng.integers(0, max(1, b - a - L)) - flen + L, 0, len(g) - flen - 1))
        else:
            start = int(rng.integers(0, len(g) - flen - 1))
        m1 = mutate(rng, g[start:start + L], 0.01)
        m2 = revcomp(mutate(rng, g[start + flen - L:start + flen], 0.01))
        if rng.random() < 0.5:
            m1, m2 = m2, m1
        s1.append(m1)
        s2.append(m2)
        q1.append(rng.integers(20, 41, L).astype(np.uint8))
        q2.append(rng.integers(20, 41, L).astype(np.uint8))
    return _flat(s1, q1), _flat(s2, q2)


def small_genome(seed, n_contigs=2, length=60_000, repeat_copies=6, unit_len=700, n_run=True):
    """Small multi-contig genome with repeats, homopolymers and an N run (parity tests)."""
    rng = np.random.default_rng(np.random.PCG64(seed))
    unit = random_bases(rng, unit_len)
    out = []
    for c in range(n_contigs):
        s = random_bases(rng, length)
        for _ in range(repeat_copies):
            p = int(rng.integers(0, length - unit_len))
            s[p:p + unit_len] = mutate(rng, unit, 0.01)
        for _ in range(length // 3000):
            p = int(rng.integers(0, length - 20))
            s[p:p + int(rng.integers(5, 16))] = BASES[rng.integers(0, 4)]
        if n_run:
            p = int(rng.integers(100, length - 200))
            s[p:p + int(rng.integers(5, 60))] = ord("N")
        out.append(("ctg%d" % c, s))
    return out


def sample_reads(contigs, n, seed, lmin=30, lmax=200, sub=0.02, indel=0.002, n_rate=0.001, qmin=2, qmax=41):
    """Mixed reads for parity: uneven lengths, both strands, subs/indels, Ns, random-sequence reads."""
    rng = np.random.default_rng(np.random.PCG64(seed))
    seqs, quals = [], []
    for i in range(n):
        L = int(rng.integers(lmin, lmax + 1))
        kind = rng.random()
        if kind < 0.05:
            s = random_bases(rng, L)
        else:
            c = contigs[int(rng.integers(0, len(contigs)))][1]
            if kind < 0.1:  # near contig ends
                st = int(rng.integers(0, 40)) if rng.random() < 0.5 else max(0, len(c) - L - int(rng.integers(-30, 20)))
            else:
                st = int(rng.integers(0, max(1, len(c) - L)))
            s = _indels(rng, c[st:st + L + 4].copy(), indel)[:L]
            s = mutate(rng, s, sub)
            nn = np.nonzero(rng.random(len(s)) < n_rate)[0]
            s[nn] = ord("N")
            if rng.random() < 0.5:
                s = revcomp(s)
        q = rng.integers(qmin, qmax + 1, len(s)).astype(np.uint8)
        if rng.random() < 0.02 and len(q) > 3:
            q = q[: len(q) - int(rng.integers(1, 3))]  # quality track shorter than the sequence
        seqs.append(s)
        quals.append(q)
    return _flat(seqs, quals)
