"""ctypes binding of oracle/libpbw_oracle.so (test infrastructure only).

The oracle is a scalar C++ restatement of the reference mapper
(GUIBA-EX/probebwa).  This module mirrors the reference's Rust API names where
one exists (HashTable.lookup, SmithWatermanAligner.align, map_reads ...), so the
parity tests read like the reference's own tests.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpbw_oracle.so")

RECORD_DTYPE = np.dtype(
    [
        ("pos", "<u8"),
        ("mate_pos", "<u8"),
        ("tlen", "<i8"),
        ("log_lk", "<f8"),
        ("posterior", "<f8"),
        ("contig", "<i4"),
        ("mate_contig", "<i4"),
        ("cigar_off", "<u4"),
        ("cigar_n", "<u4"),
        ("flag", "<u2"),
        ("mapq", "u1"),
        ("strand", "u1"),
        ("seq_rc", "u1"),
        ("pad", "u1", (3,)),
    ]
)
assert RECORD_DTYPE.itemsize == 64

CIGAR_CHARS = "MIDNSHP=X"


class _Options(C.Structure):
    _fields_ = [
        ("substitution_rate", C.c_double),
        ("insert_size_mean", C.c_double),
        ("insert_size_sd", C.c_double),
        ("sv_prior_phred", C.c_double),
        ("max_indel_len", C.c_uint32),
        ("gap_open_phred", C.c_uint32),
        ("gap_extend_phred", C.c_uint32),
        ("threads", C.c_uint32),
    ]


class _Stats(C.Structure):
    _fields_ = [
        (n, C.c_uint64)
        for n in ("lookups", "positions", "filtered_hits", "candidates", "dp_jobs", "dp_cells", "wide_jobs", "wide_cells")
    ]


class _Reads(C.Structure):
    _fields_ = [
        ("n", C.c_uint64),
        ("seq", C.c_void_p),
        ("seq_off", C.c_void_p),
        ("qual", C.c_void_p),
        ("qual_off", C.c_void_p),
    ]


class _Records(C.Structure):
    _fields_ = [("n", C.c_uint64), ("rec", C.c_void_p), ("cigar", C.c_void_p), ("n_cigar", C.c_uint64)]


class _Parsed(C.Structure):
    _fields_ = [
        ("n", C.c_uint64),
        ("seq", C.c_void_p),
        ("seq_off", C.c_void_p),
        ("qual", C.c_void_p),
        ("qual_off", C.c_void_p),
        ("ids", C.c_void_p),
        ("id_off", C.c_void_p),
    ]


_lib = None


class _PairMate(C.Structure):
    _fields_ = [("seq", C.c_char_p), ("len", C.c_uint64), ("qual", C.c_char_p), ("qlen", C.c_uint64),
                ("cigar", C.c_void_p), ("n_ops", C.c_uint32), ("ref_window", C.c_char_p), ("rwlen", C.c_uint64),
                ("lk", C.c_void_p), ("nlk", C.c_uint64), ("mask", C.c_void_p), ("mlen", C.c_uint64)]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle library missing: run `make -C oracle` (or __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        vp, u64, i64, u32, i32, dbl = C.c_void_p, C.c_uint64, C.c_int64, C.c_uint32, C.c_int, C.c_double
        sig = {
            "orc_genome_new": (vp, []),
            "orc_genome_free": (None, [vp]),
            "orc_genome_add_contig": (i32, [vp, C.c_char_p, vp, u64]),
            "orc_genome_add_fasta": (i32, [vp, C.c_char_p]),
            "orc_genome_total_length": (u64, [vp]),
            "orc_genome_n_contigs": (i32, [vp]),
            "orc_genome_contig_length": (u64, [vp, i32]),
            "orc_genome_contig_name": (C.c_char_p, [vp, i32]),
            "orc_genome_slice": (u64, [vp, i32, u64, u64, vp]),
            "orc_hash_kmer": (u32, [C.c_char_p, u64]),
            "orc_pack_position": (u64, [u64, u64, i32]),
            "orc_unpack_position": (None, [u64, C.POINTER(u64), C.POINTER(u64), C.POINTER(i32)]),
            "orc_is_acgt_only": (i32, [C.c_char_p, u64]),
            "orc_hash_build": (vp, [vp, i32, i32, i32]),
            "orc_hash_free": (None, [vp]),
            "orc_hash_lookup": (u64, [vp, C.c_char_p, u64, C.POINTER(C.POINTER(u64))]),
            "orc_hash_total_entries": (u64, [vp]),
            "orc_hash_n_keys": (u64, [vp]),
            "orc_hash_key_at": (u64, [vp, u64, C.POINTER(u32), C.POINTER(C.POINTER(u64))]),
            "orc_mix_is_bijective": (i32, []),
            "orc_sw_align": (
                i32,
                [u32, u32, u32, C.c_char_p, u64, C.c_char_p, u64, C.c_char_p, u64, i64, C.POINTER(C.c_int32), vp, u32,
                 C.POINTER(u32), C.POINTER(u64), C.POINTER(u64)],
            ),
            "orc_mismatch_penalty": (i32, [C.c_uint8]),
            "orc_homopolymer_gap_open": (i32, [i32, C.c_uint16]),
            "orc_logadd": (dbl, [dbl, dbl]),
            "orc_phred_n_choose_m": (dbl, [i64, i64]),
            "orc_effective_alphabet_size": (dbl, [dbl]),
            "orc_missing_locus_probability": (dbl, [dbl, C.c_char_p, u64, vp, u64]),
            "orc_finalize_mapq": (
                None,
                [dbl, u64, dbl, C.c_char_p, u64, vp, u32, C.c_char_p, u64, C.c_char_p, u64, C.c_char_p, u64, vp, u64,
                 C.POINTER(C.c_uint8), C.POINTER(dbl)],
            ),
            "orc_log_entropy": (dbl, [dbl, u64, vp, u32, C.c_char_p, u64, C.c_char_p, u64, C.c_char_p, u64]),
            "orc_shortlist_indices": (u64, [vp, u64, vp]),
            "orc_dedup_pairs": (u64, [vp, u64, vp]),
            "orc_mapq_for_winning_pair": (None, [vp, dbl, i32, i32, C.POINTER(_PairMate), C.POINTER(_PairMate),
                                                 vp, vp]),
            "orc_insert_update": (None, [C.POINTER(dbl), C.POINTER(dbl), C.POINTER(u64), dbl]),
            "orc_insert_log_likelihood": (dbl, [dbl, dbl, dbl]),
            "orc_default_options": (None, [C.POINTER(_Options)]),
            "orc_mapper_new": (vp, [vp, vp, C.POINTER(_Options)]),
            "orc_mapper_free": (None, [vp]),
            "orc_seed_and_cluster": (
                i32, [vp, C.c_char_p, u64, C.c_char_p, u64, vp, u64, C.POINTER(u64), vp, vp, u64, C.POINTER(u64)]),
            "orc_map_se": (i32, [vp, C.POINTER(_Reads), C.POINTER(_Records), C.POINTER(_Stats)]),
            "orc_map_pe": (
                i32,
                [vp, C.POINTER(_Reads), C.POINTER(_Reads), C.POINTER(_Records), C.POINTER(dbl), C.POINTER(dbl),
                 C.POINTER(u64), C.POINTER(_Stats)],
            ),
            "orc_records_release": (None, [C.POINTER(_Records)]),
            "orc_set_group_order": (None, [i32]),
            "orc_set_progress": (None, [i32]),
            "orc_set_index_threads": (None, [i32]),
            "orc_read_reads": (i32, [C.c_char_p, i32, C.c_uint8, C.c_char_p, C.POINTER(_Parsed)]),
            "orc_parsed_reads_free": (None, [C.POINTER(_Parsed)]),
            "orc_format_sam": (C.c_void_p, [vp, C.POINTER(_Parsed), C.POINTER(_Records), C.c_char_p, C.c_char_p,
                                            C.POINTER(u64)]),
            "orc_free_text": (None, [C.c_void_p]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _b(x):
    if isinstance(x, str):
        return x.encode()
    return bytes(x)


# ---------------------------------------------------------------------------
# genome.rs / hashtable.rs
# ---------------------------------------------------------------------------
class Genome:
    """In-memory `GenomeIndex` (genome.rs:71-99)."""

    def __init__(self):
        self.h = lib().orc_genome_new()
        self.names = []
        self.lengths = []

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_genome_free(self.h)
            self.h = None

    def add_contig(self, name, seq):
        seq = _b(seq)
        buf = C.create_string_buffer(seq, len(seq))
        lib().orc_genome_add_contig(self.h, _b(name), buf, len(seq))
        self.names.append(name if isinstance(name, str) else name.decode())
        self.lengths.append(len(seq))
        return self

    def add_fasta(self, path):
        if lib().orc_genome_add_fasta(self.h, _b(path)) != 0:
            raise IOError(path)
        n = lib().orc_genome_n_contigs(self.h)
        self.names = [lib().orc_genome_contig_name(self.h, i).decode() for i in range(n)]
        self.lengths = [lib().orc_genome_contig_length(self.h, i) for i in range(n)]
        return self

    @property
    def total_length(self):
        return lib().orc_genome_total_length(self.h)

    def slice(self, cid, start, end):
        out = C.create_string_buffer(max(1, int(end) - int(start)))
        n = lib().orc_genome_slice(self.h, cid, start, end, out)
        return out.raw[:n]


def hash_kmer(kmer):
    kmer = _b(kmer)
    return lib().orc_hash_kmer(kmer, len(kmer))


def pack_position(contig, pos, strand):
    return lib().orc_pack_position(contig, pos, strand)


def unpack_position(packed):
    c, p, s = C.c_uint64(), C.c_uint64(), C.c_int()
    lib().orc_unpack_position(packed, C.byref(c), C.byref(p), C.byref(s))
    return c.value, p.value, s.value


def is_acgt_only(kmer):
    kmer = _b(kmer)
    return bool(lib().orc_is_acgt_only(kmer, len(kmer)))


class HashTable:
    """`HashTableBuilder::new(&genome).with_k(k).with_stride(s).with_max_count(c).build()` (hashtable.rs:248)."""

    def __init__(self, genome, k=15, stride=5, max_count=300, progress=False, threads=1):
        self.genome = genome
        self.k = k
        lib().orc_set_progress(1 if progress else 0)
        lib().orc_set_index_threads(int(threads))
        self.h = lib().orc_hash_build(genome.h, k, stride, max_count)
        if not self.h:
            raise RuntimeError("index larger than 2^32 entries")

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_hash_free(self.h)
            self.h = None

    def lookup(self, kmer):
        kmer = _b(kmer)
        p = C.POINTER(C.c_uint64)()
        n = lib().orc_hash_lookup(self.h, kmer, len(kmer), C.byref(p))
        return [p[i] for i in range(n)]

    @property
    def total_entries(self):
        return lib().orc_hash_total_entries(self.h)

    def buckets(self):
        """Yield (hash, [packed...]) in ascending hash order."""
        n = lib().orc_hash_n_keys(self.h)
        h = C.c_uint32()
        p = C.POINTER(C.c_uint64)()
        for i in range(n):
            cnt = lib().orc_hash_key_at(self.h, i, C.byref(h), C.byref(p))
            yield h.value, np.ctypeslib.as_array(p, shape=(cnt,)).copy() if cnt else np.zeros(0, np.uint64)


def mix_is_bijective():
    return bool(lib().orc_mix_is_bijective())


# ---------------------------------------------------------------------------
# smith_waterman.rs
# ---------------------------------------------------------------------------
def cigar_to_str(words):
    return "".join("%d%s" % (w >> 4, CIGAR_CHARS[w & 15]) for w in words)


def cigar_ops(words):
    """[(op_char, length)] from u32 words (len << 4 | op)."""
    return [(CIGAR_CHARS[w & 15], int(w >> 4)) for w in words]


class SmithWatermanAligner:
    """`SmithWatermanAligner::new(band, gap_open_phred, gap_extend_phred)` (smith_waterman.rs:113)."""

    def __init__(self, band, gap_open_phred=40, gap_extend_phred=3):
        self.band, self.gop, self.gep = band, gap_open_phred, gap_extend_phred

    def align(self, read, qual, ref_seg, ref_offset_hint):
        """Returns (score, cigar words, ref_start_offset, cells) — smith_waterman.rs:163."""
        read, qual, ref_seg = _b(read), _b(qual), _b(ref_seg)
        cap = len(read) + len(ref_seg) + 2
        cig = (C.c_uint32 * cap)()
        score, nops, off, cells = C.c_int32(), C.c_uint32(), C.c_uint64(), C.c_uint64()
        rc = lib().orc_sw_align(self.band, self.gop, self.gep, read, len(read), qual, len(qual), ref_seg, len(ref_seg),
                                int(ref_offset_hint), C.byref(score), cig, cap, C.byref(nops), C.byref(off),
                                C.byref(cells))
        assert rc == 0
        return score.value, [cig[i] for i in range(nops.value)], off.value, cells.value


def mismatch_penalty(q):
    return lib().orc_mismatch_penalty(q)


def homopolymer_gap_open(gap_open, run):
    return lib().orc_homopolymer_gap_open(gap_open, run)


def logadd(a, b):
    return lib().orc_logadd(a, b)


def phred_n_choose_m(n, m):
    return lib().orc_phred_n_choose_m(n, m)


def effective_alphabet_size(gc):
    return lib().orc_effective_alphabet_size(gc)


def missing_locus_probability(qual, repeat_mask=(), sub_rate=0.001):
    qual = _b(qual)
    mask = np.ascontiguousarray(np.asarray(repeat_mask, dtype=np.float64))
    return lib().orc_missing_locus_probability(sub_rate, qual, len(qual), mask.ctypes.data if len(mask) else None,
                                               len(mask))


def _u32(words):
    return np.ascontiguousarray(np.asarray(words, dtype=np.uint32).reshape(-1)) if len(words) else np.zeros(1, np.uint32)


def _f64(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.float64).reshape(-1)) if len(x) else np.zeros(1, np.float64)


def log_entropy(cigar_words, query, qual, ref_window, genome_size, sub_rate=0.001):
    """`MapqCalculator::log_entropy` (bayesian.rs:306-373); CIGAR words are (len << 4) | op."""
    cig = _u32(cigar_words)
    query, qual, ref_window = _b(query), _b(qual), _b(ref_window)
    return lib().orc_log_entropy(sub_rate, genome_size, cig.ctypes.data, len(cigar_words), query, len(query), qual,
                                 len(qual), ref_window, len(ref_window))


def finalize_mapq(sum_exp, read_qual, cigar_words, query, qual, ref_window, genome_size, repeat_mask=(),
                  sub_rate=0.001):
    """`MapqCalculator::finalize_mapq` (bayesian.rs:251-282) -> (mapq, posterior)."""
    cig, mask = _u32(cigar_words), _f64(repeat_mask)
    read_qual, query, qual, ref_window = _b(read_qual), _b(query), _b(qual), _b(ref_window)
    mq, post = C.c_uint8(), C.c_double()
    lib().orc_finalize_mapq(sub_rate, genome_size, sum_exp, read_qual, len(read_qual), cig.ctypes.data,
                            len(cigar_words), query, len(query), qual, len(qual), ref_window, len(ref_window),
                            mask.ctypes.data if len(repeat_mask) else None, len(repeat_mask), C.byref(mq),
                            C.byref(post))
    return mq.value, post.value


def shortlist_indices(post):
    """`shortlist_indices` (paired.rs:553-566)."""
    p = _f64(post)
    out = np.zeros(max(1, len(post)), np.uint64)
    n = lib().orc_shortlist_indices(p.ctypes.data, len(post), out.ctypes.data)
    return [int(x) for x in out[:n]]


def dedup_pairs(loci):
    """The raw-pair dedup of joint_map (paired.rs:346-351): loci = [((contig, pos, strand), (contig, pos, strand))];
    returns the kept indices."""
    arr = np.ascontiguousarray(np.asarray([a + b for a, b in loci], dtype=np.uint64).reshape(-1))
    out = np.zeros(max(1, len(loci)), np.uint64)
    n = lib().orc_dedup_pairs(arr.ctypes.data if len(loci) else None, len(loci), out.ctypes.data)
    return [int(x) for x in out[:n]]


def mapq_for_winning_pair(mapper, pair_se, own1, own2, mate1, mate2):
    """`mapq_for_winning_pair` (paired.rs:398-435) -> ((mapq1, posterior1), (mapq2, posterior2)).  Each mate
    is a dict: seq, qual, cigar (words), ref_window, lk (its candidates' log-likelihoods), mask."""
    keep = []

    def mk(d):
        cig, lk, mask = _u32(d["cigar"]), _f64(d["lk"]), _f64(d.get("mask", ()))
        keep.extend([cig, lk, mask])
        seq, qual, rw = _b(d["seq"]), _b(d["qual"]), _b(d["ref_window"])
        return _PairMate(seq, len(seq), qual, len(qual), cig.ctypes.data, len(d["cigar"]), rw, len(rw),
                         lk.ctypes.data, len(d["lk"]), mask.ctypes.data if len(d.get("mask", ())) else None,
                         len(d.get("mask", ())))
    a, b = mk(mate1), mk(mate2)
    q = np.zeros(2, np.uint8)
    p = np.zeros(2, np.float64)
    lib().orc_mapq_for_winning_pair(mapper.h, pair_se, int(own1), int(own2), C.byref(a), C.byref(b), q.ctypes.data,
                                    p.ctypes.data)
    return (int(q[0]), float(p[0])), (int(q[1]), float(p[1]))


def insert_update(mean, sd, n_obs, observed):
    m, s, n = C.c_double(mean), C.c_double(sd), C.c_uint64(n_obs)
    lib().orc_insert_update(C.byref(m), C.byref(s), C.byref(n), observed)
    return m.value, s.value, n.value


def insert_log_likelihood(mean, sd, x):
    return lib().orc_insert_log_likelihood(mean, sd, x)


# ---------------------------------------------------------------------------
# reads / records
# ---------------------------------------------------------------------------
class ReadBatch:
    """Flat read batch: bases upper-cased, quals offset-subtracted."""

    def __init__(self, seqs, quals, ids=None):
        seqs = [_b(s) for s in seqs]
        quals = [_b(q) for q in quals]
        self.n = len(seqs)
        self.seq = np.frombuffer(b"".join(seqs) + b"\0", dtype=np.uint8).copy()
        self.qual = np.frombuffer(b"".join(quals) + b"\0", dtype=np.uint8).copy()
        self.seq_off = np.zeros(self.n + 1, np.uint64)
        self.qual_off = np.zeros(self.n + 1, np.uint64)
        self.seq_off[1:] = np.cumsum([len(s) for s in seqs]) if seqs else []
        self.qual_off[1:] = np.cumsum([len(q) for q in quals]) if quals else []
        self.ids = list(ids) if ids is not None else ["r%d" % i for i in range(self.n)]

    @classmethod
    def from_arrays(cls, seq, seq_off, qual, qual_off, ids=None):
        rb = cls.__new__(cls)
        rb.n = len(seq_off) - 1
        rb.seq = np.ascontiguousarray(seq, dtype=np.uint8)
        rb.qual = np.ascontiguousarray(qual, dtype=np.uint8)
        rb.seq_off = np.ascontiguousarray(seq_off, dtype=np.uint64)
        rb.qual_off = np.ascontiguousarray(qual_off, dtype=np.uint64)
        rb.ids = ids if ids is not None else ["r%d" % i for i in range(rb.n)]
        return rb

    def subset(self, lo, hi):
        so, qo = self.seq_off, self.qual_off
        return ReadBatch.from_arrays(self.seq[so[lo]:so[hi]], so[lo:hi + 1] - so[lo], self.qual[qo[lo]:qo[hi]],
                                     qo[lo:hi + 1] - qo[lo], self.ids[lo:hi] if self.ids is not None else None)

    def seq_of(self, i):
        return self.seq[self.seq_off[i]:self.seq_off[i + 1]].tobytes()

    def qual_of(self, i):
        return self.qual[self.qual_off[i]:self.qual_off[i + 1]].tobytes()

    def cstruct(self):
        return _Reads(self.n, self.seq.ctypes.data, self.seq_off.ctypes.data, self.qual.ctypes.data,
                      self.qual_off.ctypes.data)


class Records:
    """Records of one map call: numpy structured array + cigar word array."""

    def __init__(self, rec, cigar):
        self.rec = rec
        self.cigar = cigar

    def __len__(self):
        return len(self.rec)

    def cigar_words(self, i):
        r = self.rec[i]
        return [int(x) for x in self.cigar[r["cigar_off"]:r["cigar_off"] + r["cigar_n"]]]

    def cigar_str(self, i):
        return cigar_to_str(self.cigar_words(i))


def _take_records(cr):
    n = cr.n
    rec = np.frombuffer((C.c_char * (n * 64)).from_address(cr.rec), dtype=RECORD_DTYPE).copy() if n else np.zeros(
        0, RECORD_DTYPE)
    cig = np.frombuffer((C.c_char * (cr.n_cigar * 4)).from_address(cr.cigar), dtype=np.uint32).copy() \
        if cr.n_cigar else np.zeros(0, np.uint32)
    lib().orc_records_release(C.byref(cr))
    return Records(rec, cig)


DEFAULT_OPTIONS = dict(substitution_rate=0.001, insert_size_mean=250.0, insert_size_sd=60.0, sv_prior_phred=55.0,
                       max_indel_len=30, gap_open_phred=40, gap_extend_phred=3, threads=1)


class Mapper:
    """`SingleEndMapper` / `PairedEndMapper` over one genome + table (single.rs:52, paired.rs:85)."""

    def __init__(self, genome, table, **options):
        o = dict(DEFAULT_OPTIONS)
        o.update(options)
        self.options = o
        self.genome, self.table = genome, table
        self.h = lib().orc_mapper_new(genome.h, table.h, C.byref(_Options(**o)))
        self.last_stats = None

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_mapper_free(self.h)
            self.h = None

    def seed_and_cluster(self, seq, qual):
        seq, qual = _b(seq), _b(qual)
        cap = 1 << 16
        while True:
            hits = np.zeros(cap * 5, np.uint64)
            cands = np.zeros(cap * 4, np.uint64)
            rp = np.zeros(max(1, len(seq)), np.float64)
            nh, nc = C.c_uint64(), C.c_uint64()
            rc = lib().orc_seed_and_cluster(self.h, seq, len(seq), qual, len(qual), hits.ctypes.data, cap,
                                            C.byref(nh), rp.ctypes.data, cands.ctypes.data, cap, C.byref(nc))
            if rc == 0:
                break
            cap *= 4
        nrp = max(0, len(seq) - 14)
        return hits[:nh.value * 5].reshape(-1, 5), rp[:nrp], cands[:nc.value * 4].reshape(-1, 4)

    def map_se(self, reads, threads=None):
        if threads is not None:
            lib().orc_mapper_free(self.h)
            o = dict(self.options, threads=threads)
            self.h = lib().orc_mapper_new(self.genome.h, self.table.h, C.byref(_Options(**o)))
        out, st = _Records(), _Stats()
        rs = reads.cstruct()
        lib().orc_map_se(self.h, C.byref(rs), C.byref(out), C.byref(st))
        self.last_stats = {f: getattr(st, f) for f, _ in _Stats._fields_}
        return _take_records(out)

    def map_pe(self, r1, r2, mean=None, sd=None, n_obs=0):
        mean = self.options["insert_size_mean"] if mean is None else mean
        sd = self.options["insert_size_sd"] if sd is None else sd
        out, st = _Records(), _Stats()
        m, s, n = C.c_double(mean), C.c_double(sd), C.c_uint64(n_obs)
        a, b = r1.cstruct(), r2.cstruct()
        rc = lib().orc_map_pe(self.h, C.byref(a), C.byref(b), C.byref(out), C.byref(m), C.byref(s), C.byref(n),
                              C.byref(st))
        if rc != 0:
            raise ValueError("paired-end input files have different read counts (%d vs %d)" % (r1.n, r2.n))
        self.last_stats = {f: getattr(st, f) for f, _ in _Stats._fields_}
        self.insert_model = (m.value, s.value, n.value)
        return _take_records(out)


def set_group_order(mode):
    lib().orc_set_group_order(mode)


# ---------------------------------------------------------------------------
# fastx.rs / sam.rs
# ---------------------------------------------------------------------------
def read_reads(path, fasta=False, quality_offset=33, adapter=None):
    """`read_reads_files_with_preprocessing` for one file: (ReadBatch, ids)."""
    p = _Parsed()
    rc = lib().orc_read_reads(_b(path), 1 if fasta else 0, quality_offset, _b(adapter) if adapter else None,
                              C.byref(p))
    if rc != 0:
        raise IOError(path)
    n = p.n

    def arr(ptr, cnt, dt):
        if cnt == 0:
            return np.zeros(0, dt)
        return np.frombuffer((C.c_char * (cnt * np.dtype(dt).itemsize)).from_address(ptr), dtype=dt).copy()

    so = arr(p.seq_off, n + 1, np.uint64)
    qo = arr(p.qual_off, n + 1, np.uint64)
    io = arr(p.id_off, n + 1, np.uint64)
    seq = arr(p.seq, int(so[-1]) + 1, np.uint8)
    qual = arr(p.qual, int(qo[-1]) + 1, np.uint8)
    idb = arr(p.ids, int(io[-1]) + 1, np.uint8).tobytes()
    ids = [idb[int(io[i]):int(io[i + 1]) - 1].decode("utf-8", "replace") for i in range(n)]
    lib().orc_parsed_reads_free(C.byref(p))
    return ReadBatch.from_arrays(seq, so, qual, qo, ids)


def format_sam(genome, reads, records, command_line="probebwa", read_group=None):
    """SAM text (sam.rs:83-107), records in the order given (reads aligned 1:1)."""
    ids = b"".join(_b(i) + b"\0" for i in reads.ids)
    id_off = np.zeros(reads.n + 1, np.uint64)
    id_off[1:] = np.cumsum([len(_b(i)) + 1 for i in reads.ids]) if reads.n else []
    idbuf = np.frombuffer(ids + b"\0", dtype=np.uint8).copy()
    p = _Parsed(reads.n, reads.seq.ctypes.data, reads.seq_off.ctypes.data, reads.qual.ctypes.data,
                reads.qual_off.ctypes.data, idbuf.ctypes.data, id_off.ctypes.data)
    rec = np.ascontiguousarray(records.rec)
    cig = np.ascontiguousarray(records.cigar) if len(records.cigar) else np.zeros(1, np.uint32)
    cr = _Records(len(rec), rec.ctypes.data, cig.ctypes.data, len(records.cigar))
    ln = C.c_uint64()
    t = lib().orc_format_sam(genome.h, C.byref(p), C.byref(cr), _b(command_line),
                             _b(read_group) if read_group else None, C.byref(ln))
    s = C.string_at(t, ln.value)
    lib().orc_free_text(t)
    return s


def interleave(r1, r2):
    """Pairs as one batch r1[0], r2[0], r1[1], ... (the record order of map_pe)."""
    seqs, quals, ids = [], [], []
    for i in range(r1.n):
        for rb in (r1, r2):
            seqs.append(rb.seq_of(i))
            quals.append(rb.qual_of(i))
            ids.append(rb.ids[i])
    return ReadBatch(seqs, quals, ids)
