"""probebwa_amd — MI355X (gfx950) implementation of probebwa's `map` hot path.

Python binding (ctypes) of libprobebwa_amd.so, whose C ABI is declared in
include/probebwa_amd.h.  The names mirror the reference's Rust API
(GUIBA-EX/probebwa): `GenomeIndexBuilder`/`Genome`, `MapperOptions`,
`map_reads`, `SingleEndMapper.map`, `PairedEndMapper.map_pair`,
`SmithWatermanAligner.align`.  Every compute call runs the HIP kernels; there
is no CPU fallback — importing works without a GPU, but creating a context
fails loudly when no HIP device or no built library is present.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libprobebwa_amd.so")

PBW_OK = 0
ERRORS = {-1: "invalid argument", -2: "HIP error", -3: "limit exceeded", -4: "I/O error", -5: "count mismatch"}

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

# Every symbol include/probebwa_amd.h declares (checked by the CPU test-suite).
EXPORTED_SYMBOLS = (
    "pbw_genome_create", "pbw_genome_add_contig", "pbw_genome_add_fasta", "pbw_genome_n_contigs",
    "pbw_genome_contig_name", "pbw_genome_contig_length", "pbw_genome_destroy", "pbw_default_options",
    "pbw_ctx_create", "pbw_ctx_destroy", "pbw_last_error", "pbw_global_error", "pbw_map_se", "pbw_reads_upload",
    "pbw_map_se_device", "pbw_reads_free", "pbw_map_pe",
    "pbw_align_batch", "pbw_align_batch_noalloc", "pbw_index_lookup", "pbw_index_total_entries", "pbw_records_release", "pbw_get_stats",
    "pbw_format_sam", "pbw_free_text", "pbw_default_train_options", "pbw_mlp_init", "pbw_mlp_set_params",
    "pbw_mlp_get_params", "pbw_mlp_forward", "pbw_mlp_gradient", "pbw_mlp_train", "pbw_default_prefilter_options",
    "pbw_prefilter_set", "pbw_prefilter_get", "pbw_prefilter_dataset", "pbw_prefilter_train",
    "pbw_tune_get", "pbw_tune_set", "pbw_autotune",
    "pbw_ctx_create_from_sthash",
    "pbw_ctx_genome",
    "pbw_ctx_get_options",
    "pbw_default_io_options",
    "pbw_fastx_bytes_read",
    "pbw_fastx_close",
    "pbw_fastx_next",
    "pbw_fastx_open",
    "pbw_genome_get_sequence",
    "pbw_genome_load_stidx",
    "pbw_genome_save_stidx",
    "pbw_index_save_sthash",
    "pbw_map_files",
    "pbw_write_bam",
)

MLP_INPUTS = 32
MLP_PARAMS = 20865


class TrainOptions(C.Structure):
    """Prefilter training options (Adam over shuffled mini-batches)."""

    _fields_ = [("epochs", C.c_uint32), ("batch", C.c_uint32), ("lr", C.c_float), ("beta1", C.c_float),
                ("beta2", C.c_float), ("eps", C.c_float), ("pos_weight", C.c_float), ("seed", C.c_uint64)]

    def __init__(self, **kw):
        super().__init__()
        lib().pbw_default_train_options(C.byref(self))
        for k, v in kw.items():
            setattr(self, k, v)


class PrefilterOptions(C.Structure):
    """Candidate prefilter in the mapping path (see include/probebwa_amd.h)."""

    _fields_ = [("enabled", C.c_int32), ("threshold", C.c_float), ("top_k", C.c_uint32),
                ("mapq_fallback", C.c_uint32), ("close_nats", C.c_float), ("audit_period", C.c_uint32)]

    def __init__(self, **kw):
        super().__init__()
        lib().pbw_default_prefilter_options(C.byref(self))
        for k, v in kw.items():
            setattr(self, k, v)


class Tune(C.Structure):
    """Launch shapes of a mapping call (`pbw_tune`): they change its speed, never its records."""

    _fields_ = [("batch_reads", C.c_uint32), ("dp_grid", C.c_uint32), ("split_len", C.c_uint32),
                ("long_waves", C.c_uint32), ("dp_pack16", C.c_uint32), ("seed_pool", C.c_uint32)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


TUNE_MAX_TRIALS = 48


class AutotuneReport(C.Structure):
    _fields_ = [("best", Tune), ("best_ms", C.c_double), ("trials", C.c_uint32),
                ("tried", Tune * TUNE_MAX_TRIALS), ("ms", C.c_double * TUNE_MAX_TRIALS)]

    def as_dict(self):
        return {"best": self.best.as_dict(), "best_ms": self.best_ms,
                "trials": [dict(self.tried[i].as_dict(), ms=self.ms[i]) for i in range(self.trials)]}


# training label rules (PBW_LABEL_LOGLK, PBW_LABEL_SCORE)
LABEL_KINDS = {"loglk": 0, "score": 1}


class PrefilterReport(C.Structure):
    _fields_ = [("rows", C.c_uint64), ("positives", C.c_uint64), ("val_rows", C.c_uint64),
                ("val_positives", C.c_uint64), ("threshold", C.c_float), ("val_recall", C.c_double),
                ("val_skip", C.c_double), ("loss_first", C.c_double), ("loss_last", C.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


def to_bf16(a):
    """float array -> bf16 bit patterns (uint16), round to nearest even."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return u.astype(np.uint16)


def from_bf16(h):
    return (np.asarray(h, dtype=np.uint16).astype(np.uint32) << 16).view(np.float32)


class MapperOptions(C.Structure):
    """`MapperOptions` (src/types.rs:207-249) plus the device batch size."""

    _fields_ = [
        ("substitution_rate", C.c_double),
        ("insert_size_mean", C.c_double),
        ("insert_size_sd", C.c_double),
        ("sv_prior_phred", C.c_double),
        ("max_indel_len", C.c_uint32),
        ("gap_open_phred", C.c_uint32),
        ("gap_extend_phred", C.c_uint32),
        ("batch_reads", C.c_uint32),
    ]

    def __init__(self, **kw):
        super().__init__()
        lib().pbw_default_options(C.byref(self))
        for k, v in kw.items():
            setattr(self, k, v)


class _ReadBatch(C.Structure):
    _fields_ = [("n", C.c_uint64), ("seq", C.c_void_p), ("seq_off", C.c_void_p), ("qual", C.c_void_p),
                ("qual_off", C.c_void_p)]


class _RecordBatch(C.Structure):
    _fields_ = [("n", C.c_uint64), ("rec", C.c_void_p), ("cigar", C.c_void_p), ("n_cigar", C.c_uint64)]


class InsertModel(C.Structure):
    """`InsertSizeDistribution` (src/types.rs:266-270)."""

    _fields_ = [("mean", C.c_double), ("sd", C.c_double), ("n_observed", C.c_uint64)]


class _AlignJob(C.Structure):
    _fields_ = [("read", C.c_void_p), ("qual", C.c_void_p), ("ref", C.c_void_p), ("m", C.c_uint32),
                ("qlen", C.c_uint32), ("n", C.c_uint32), ("band", C.c_uint32), ("hint", C.c_int64)]


class _AlignResult(C.Structure):
    _fields_ = [("score", C.c_int32), ("ref_offset", C.c_uint32), ("cigar_off", C.c_uint32), ("cigar_n", C.c_uint32)]


class Stats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("reads", "lookups", "positions", "filtered_hits", "dp_jobs", "dp_cells",
                                           "wide_jobs", "wide_cells", "seed_overflow_reads", "tier2_cluster_reads",
                                           "generic_dp_jobs", "band_dp_jobs")] + \
               [(n, C.c_double) for n in ("ms_seed", "ms_cluster", "ms_dp", "ms_mapq", "ms_pair", "ms_total",
                                           "ms_k_seed", "ms_k_dp")] + \
               [(n, C.c_uint64) for n in ("n_k_seed", "n_k_dp", "pf_jobs", "pf_jobs_aligned", "pf_fallback_reads",
                                           "pf_audit_reads", "pf_audit_mismatches")] + \
               [("ms_pf", C.c_double), ("ms_k_cluster", C.c_double), ("n_k_cluster", C.c_uint64),
                ("mapq_near_boundary", C.c_uint64), ("ms_k_wide", C.c_double), ("n_k_wide", C.c_uint64)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class PbwError(RuntimeError):
    pass


class IoOptions(C.Structure):
    """pbw_io_options: the map subcommand's read and output flags (src/main.rs:40-111)."""
    _fields_ = [("input_format", C.c_int), ("quality_offset", C.c_uint32), ("adapter", C.c_char_p),
                ("label_prefix", C.c_char_p), ("command_line", C.c_char_p), ("read_group", C.c_char_p),
                ("output_format", C.c_int), ("index", C.c_int), ("raw_qual", C.c_int), ("threads", C.c_uint32),
                ("batch_reads", C.c_uint64)]


class IoReport(C.Structure):
    _fields_ = [("reads", C.c_uint64), ("records", C.c_uint64), ("bytes_in", C.c_uint64),
                ("bytes_out", C.c_uint64), ("ms_parse", C.c_double), ("ms_map", C.c_double),
                ("ms_write", C.c_double), ("ms_total", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


def lib():
    """Loads libprobebwa_amd.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise PbwError("libprobebwa_amd.so is not built: run `make -C probebwa_amd/csrc` "
                           "(or __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        vp, u64, i64, u32, i32 = C.c_void_p, C.c_uint64, C.c_int64, C.c_uint32, C.c_int
        sig = {
            "pbw_genome_create": (i32, [C.POINTER(vp)]),
            "pbw_genome_add_contig": (i32, [vp, C.c_char_p, vp, u64]),
            "pbw_genome_add_fasta": (i32, [vp, C.c_char_p]),
            "pbw_genome_n_contigs": (i32, [vp]),
            "pbw_genome_contig_name": (C.c_char_p, [vp, i32]),
            "pbw_genome_contig_length": (u64, [vp, i32]),
            "pbw_genome_destroy": (None, [vp]),
            "pbw_genome_get_sequence": (u64, [vp, i32, u64, u64, vp]),
            "pbw_genome_load_stidx": (i32, [C.c_char_p, C.POINTER(vp)]),
            "pbw_genome_save_stidx": (i32, [vp, C.c_char_p, C.c_char_p, C.c_char_p]),
            "pbw_ctx_create_from_sthash": (i32, [vp, C.c_char_p, C.POINTER(MapperOptions), i32, C.POINTER(vp)]),
            "pbw_index_save_sthash": (i32, [vp, C.c_char_p]),
            "pbw_ctx_genome": (vp, [vp]),
            "pbw_ctx_get_options": (i32, [vp, C.POINTER(MapperOptions)]),
            "pbw_write_bam": (i32, [vp, C.POINTER(_ReadBatch), C.c_char_p, C.POINTER(_RecordBatch), C.c_char_p,
                                    C.c_char_p, C.c_char_p, i32, i32, u32]),
            "pbw_fastx_open": (i32, [C.c_char_p, i32, u32, C.c_char_p, C.POINTER(vp)]),
            "pbw_fastx_next": (i32, [vp, u64, C.POINTER(_ReadBatch), C.POINTER(C.c_void_p)]),
            "pbw_fastx_bytes_read": (u64, [vp]),
            "pbw_fastx_close": (None, [vp]),
            "pbw_default_io_options": (None, [C.POINTER(IoOptions)]),
            "pbw_map_files": (i32, [vp, C.c_char_p, C.c_char_p, C.c_char_p, C.POINTER(IoOptions),
                                    C.POINTER(IoReport)]),
            "pbw_default_options": (None, [C.POINTER(MapperOptions)]),
            "pbw_ctx_create": (i32, [vp, C.POINTER(MapperOptions), i32, C.POINTER(vp)]),
            "pbw_ctx_destroy": (None, [vp]),
            "pbw_last_error": (C.c_char_p, [vp]),
            "pbw_global_error": (C.c_char_p, []),
            "pbw_map_se": (i32, [vp, C.POINTER(_ReadBatch), C.POINTER(_RecordBatch)]),
            "pbw_reads_upload": (i32, [vp, C.POINTER(_ReadBatch), C.POINTER(vp)]),
            "pbw_map_se_device": (i32, [vp, vp, u64, u64, C.POINTER(_RecordBatch)]),
            "pbw_reads_free": (None, [vp]),
            "pbw_tune_get": (i32, [vp, vp]),
            "pbw_tune_set": (i32, [vp, vp]),
            "pbw_autotune": (i32, [vp, vp, u64, u64, vp]),
            "pbw_map_pe": (i32, [vp, C.POINTER(_ReadBatch), C.POINTER(_ReadBatch), C.POINTER(InsertModel),
                                 C.POINTER(_RecordBatch)]),
            "pbw_align_batch": (i32, [vp, u32, u32, C.POINTER(_AlignJob), u64, C.POINTER(_AlignResult), vp, u64]),
            "pbw_align_batch_noalloc": (i32, [vp, u32, u32, C.POINTER(_AlignJob), u64, C.POINTER(_AlignResult), vp,
                                              u64]),
            "pbw_index_lookup": (i64, [vp, u32, vp, u64]),
            "pbw_index_total_entries": (u64, [vp]),
            "pbw_records_release": (None, [C.POINTER(_RecordBatch)]),
            "pbw_get_stats": (i32, [vp, C.POINTER(Stats)]),
            "pbw_format_sam": (vp, [vp, C.POINTER(_ReadBatch), C.c_char_p, C.POINTER(_RecordBatch), C.c_char_p,
                                    C.c_char_p, C.POINTER(u64)]),
            "pbw_free_text": (None, [vp]),
            "pbw_default_train_options": (None, [vp]),
            "pbw_mlp_init": (i32, [vp, u64]),
            "pbw_mlp_set_params": (i32, [vp, vp, u64]),
            "pbw_mlp_get_params": (i32, [vp, vp, u64]),
            "pbw_mlp_forward": (i32, [vp, vp, u64, vp]),
            "pbw_mlp_gradient": (i32, [vp, vp, vp, u64, C.c_float, vp, C.POINTER(C.c_double)]),
            "pbw_mlp_train": (i32, [vp, vp, vp, u64, vp, vp]),
            "pbw_default_prefilter_options": (None, [vp]),
            "pbw_prefilter_set": (i32, [vp, vp]),
            "pbw_prefilter_get": (i32, [vp, vp]),
            "pbw_prefilter_dataset": (i32, [vp, C.POINTER(_ReadBatch), i32, C.c_float, vp, vp, vp, u64,
                                            C.POINTER(u64)]),
            "pbw_prefilter_train": (i32, [vp, C.POINTER(_ReadBatch), vp, i32, C.c_float, C.c_float, vp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _b(x):
    return x.encode() if isinstance(x, str) else bytes(x)


def _check(rc, ctx=None):
    if rc != PBW_OK:
        msg = (lib().pbw_last_error(ctx) if ctx else lib().pbw_global_error()) or b""
        raise PbwError("%s: %s" % (ERRORS.get(rc, rc), msg.decode(errors="replace")))


class Genome:
    """`GenomeIndexBuilder` + `GenomeIndex` (src/index/genome.rs:71-147), kept in host memory."""

    def __init__(self):
        h = C.c_void_p()
        _check(lib().pbw_genome_create(C.byref(h)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            lib().pbw_genome_destroy(self.h)
            self.h = None

    def add_contig(self, name, seq):
        seq = _b(seq)
        buf = C.create_string_buffer(seq, len(seq))
        _check(lib().pbw_genome_add_contig(self.h, _b(name), buf, len(seq)))
        return self

    def add_fasta(self, path):
        _check(lib().pbw_genome_add_fasta(self.h, _b(path)))
        return self

    @classmethod
    def load_stidx(cls, path):
        """`GenomeIndex::load` (src/index/genome.rs:80-86): a bincode-1 `.stidx` file."""
        g = cls.__new__(cls)
        h = C.c_void_p()
        _check(lib().pbw_genome_load_stidx(_b(path), C.byref(h)))
        g.h = h
        return g

    def save_stidx(self, path, species="", assembly=""):
        """`GenomeIndex::save` (src/index/genome.rs:88-94)."""
        _check(lib().pbw_genome_save_stidx(self.h, _b(species), _b(assembly), _b(path)))

    def get_sequence(self, cid, start, end):
        """`GenomeIndex::get_sequence` (src/index/genome.rs:97-99): bases [start, end), 'N' for non-ACGT."""
        n = max(0, min(end, self.lengths[cid]) - start)
        buf = C.create_string_buffer(max(n, 1))
        got = lib().pbw_genome_get_sequence(self.h, cid, start, end, buf)
        return buf.raw[:got]

    @property
    def names(self):
        return [lib().pbw_genome_contig_name(self.h, i).decode() for i in range(lib().pbw_genome_n_contigs(self.h))]

    @property
    def lengths(self):
        return [lib().pbw_genome_contig_length(self.h, i) for i in range(lib().pbw_genome_n_contigs(self.h))]


class ReadBatch:
    """Flat reads: bases upper-cased, quals offset-subtracted (src/types.rs:12-15, :57-60)."""

    def __init__(self, seqs, quals, ids=None):
        seqs = [_b(s).upper() for s in seqs]
        quals = [_b(q) for q in quals]
        self.n = len(seqs)
        self.seq = np.frombuffer(b"".join(seqs) + b"\0", dtype=np.uint8).copy()
        self.qual = np.frombuffer(b"".join(quals) + b"\0", dtype=np.uint8).copy()
        self.seq_off = np.zeros(self.n + 1, np.uint64)
        self.qual_off = np.zeros(self.n + 1, np.uint64)
        if self.n:
            self.seq_off[1:] = np.cumsum([len(s) for s in seqs])
            self.qual_off[1:] = np.cumsum([len(q) for q in quals])
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

    def cstruct(self):
        return _ReadBatch(self.n, self.seq.ctypes.data, self.seq_off.ctypes.data, self.qual.ctypes.data,
                          self.qual_off.ctypes.data)

    def subset(self, begin, end):
        """Reads [begin, end) as a new batch."""
        so, qo = self.seq_off, self.qual_off
        return ReadBatch.from_arrays(self.seq[so[begin]:so[end]], so[begin:end + 1] - so[begin],
                                     self.qual[qo[begin]:qo[end]], qo[begin:end + 1] - qo[begin],
                                     self.ids[begin:end])

    def seq_of(self, i):
        return self.seq[self.seq_off[i]:self.seq_off[i + 1]].tobytes()

    def qual_of(self, i):
        return self.qual[self.qual_off[i]:self.qual_off[i + 1]].tobytes()


class Records:
    """Records of one map call (numpy structured array + CIGAR words)."""

    def __init__(self, rec, cigar):
        self.rec = rec
        self.cigar = cigar

    def __len__(self):
        return len(self.rec)

    def cigar_words(self, i):
        r = self.rec[i]
        return [int(x) for x in self.cigar[r["cigar_off"]:r["cigar_off"] + r["cigar_n"]]]

    def cigar_str(self, i):
        return "".join("%d%s" % (w >> 4, CIGAR_CHARS[w & 15]) for w in self.cigar_words(i))


class _RecordMemory:
    """Owns the arrays of one pbw_record_batch; frees them when no view is left."""

    def __init__(self, rb):
        self.rb = rb

    def __del__(self):
        rb, self.rb = self.rb, None
        if rb is not None:
            try:
                lib().pbw_records_release(C.byref(rb))
            except Exception:  # interpreter shutdown
                pass


def _view(keeper, addr, nbytes, dtype):
    buf = (C.c_char * nbytes).from_address(addr)
    buf._keeper = keeper  # the numpy view holds buf, buf holds the owner
    return np.frombuffer(buf, dtype=dtype)


def _take(rb):
    """Records of a pbw_record_batch as numpy views of the library's arrays (no copy)."""
    keeper = _RecordMemory(rb)
    rec = _view(keeper, rb.rec, rb.n * 64, RECORD_DTYPE) if rb.n else np.zeros(0, RECORD_DTYPE)
    cig = _view(keeper, rb.cigar, rb.n_cigar * 4, np.uint32) if rb.n_cigar else np.zeros(0, np.uint32)
    return Records(rec, cig)


class Context:
    """Device context: uploads the genome, builds the 15-mer index on it
    (HashTableBuilder, src/index/hashtable.rs:248-322) and owns the HIP stream.
    """

    def __init__(self, genome, options=None, device=0, sthash=None, **kw):
        """With `sthash`, the bucket table is read from that `.sthash` file (HashTable::load,
        src/index/hashtable.rs:66-72) instead of built from the genome."""
        self.genome = genome
        self.options = options if options is not None else MapperOptions(**kw)
        h = C.c_void_p()
        if sthash is None:
            _check(lib().pbw_ctx_create(genome.h, C.byref(self.options), device, C.byref(h)))
        else:
            _check(lib().pbw_ctx_create_from_sthash(genome.h, _b(sthash), C.byref(self.options), device,
                                                    C.byref(h)))
        self.h = h

    def save_sthash(self, path):
        """`HashTable::save` (src/index/hashtable.rs:74-80) of this context's buckets."""
        _check(lib().pbw_index_save_sthash(self.h, _b(path)), self.h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().pbw_ctx_destroy(self.h)
            self.h = None

    def map_se(self, reads):
        """`map_reads` single-end loop (src/lib.rs:84-104); records in input order."""
        out = _RecordBatch()
        rs = reads.cstruct()
        _check(lib().pbw_map_se(self.h, C.byref(rs), C.byref(out)), self.h)
        return _take(out)

    def upload_reads(self, reads):
        """Copies a read batch into device memory once (`pbw_reads_upload`)."""
        return DeviceReads(self, reads)

    @property
    def tune(self):
        t = Tune()
        _check(lib().pbw_tune_get(self.h, C.byref(t)), self.h)
        return t

    @tune.setter
    def tune(self, t):
        if isinstance(t, dict):
            cur = self.tune
            for k, v in t.items():
                setattr(cur, k, v)
            t = cur
        _check(lib().pbw_tune_set(self.h, C.byref(t)), self.h)

    def autotune(self, dreads, begin=0, end=None):
        """Times candidate launch shapes on reads [begin, end) of `dreads` and keeps the fastest."""
        end = dreads.n if end is None else end
        rep = AutotuneReport()
        _check(lib().pbw_autotune(self.h, dreads.h, begin, end, C.byref(rep)), self.h)
        return rep.as_dict()

    def map_se_device(self, dreads, begin=0, end=None):
        """Single-end map of reads [begin, end) already resident in HBM."""
        end = dreads.n if end is None else end
        out = _RecordBatch()
        _check(lib().pbw_map_se_device(self.h, dreads.h, begin, end, C.byref(out)), self.h)
        return _take(out)

    def map_pe(self, r1, r2, mean=None, sd=None, n_observed=0):
        """`map_paired` (src/lib.rs:145-185); records interleaved r1, r2."""
        model = InsertModel(self.options.insert_size_mean if mean is None else mean,
                            self.options.insert_size_sd if sd is None else sd, n_observed)
        out = _RecordBatch()
        a, b = r1.cstruct(), r2.cstruct()
        _check(lib().pbw_map_pe(self.h, C.byref(a), C.byref(b), C.byref(model), C.byref(out)), self.h)
        self.insert_model = (model.mean, model.sd, model.n_observed)
        return _take(out)

    def align_batch(self, jobs, gap_open_phred=40, gap_extend_phred=3):
        """Batched `SmithWatermanAligner::new(band,..).align(read, qual, ref, hint)`.

        jobs: list of (read, qual, ref_seg, band, hint).  Returns a list of
        (score, cigar words, ref_start_offset)."""
        return self._align(lib().pbw_align_batch, jobs, gap_open_phred, gap_extend_phred)

    def align_batch_noalloc(self, jobs, gap_open_phred=40, gap_extend_phred=3):
        """`align_batch` through pbw_align_batch_noalloc: the library's fixed device region
        (64 MiB per GPU) instead of allocations; a batch that needs more is an error."""
        return self._align(lib().pbw_align_batch_noalloc, jobs, gap_open_phred, gap_extend_phred)

    def _align(self, fn, jobs, gap_open_phred, gap_extend_phred):
        n = len(jobs)
        keep = []
        arr = (_AlignJob * max(n, 1))()
        total = 0
        for i, (read, qual, ref, band, hint) in enumerate(jobs):
            bufs = [C.create_string_buffer(_b(x), max(1, len(x))) for x in (read, qual, ref)]
            keep.append(bufs)
            arr[i] = _AlignJob(C.addressof(bufs[0]), C.addressof(bufs[1]), C.addressof(bufs[2]), len(read), len(qual),
                               len(ref), band, hint)
            total += len(read) + len(ref) + 2
        res = (_AlignResult * max(n, 1))()
        cig = np.zeros(max(total, 1), np.uint32)
        _check(fn(self.h, gap_open_phred, gap_extend_phred, arr, n, res, cig.ctypes.data, len(cig)), self.h)
        return [(res[i].score, [int(x) for x in cig[res[i].cigar_off:res[i].cigar_off + res[i].cigar_n]],
                 res[i].ref_offset) for i in range(n)]

    def lookup_code(self, code):
        """`HashTable::lookup` by 30-bit code: the bucket's packed positions."""
        n = lib().pbw_index_lookup(self.h, code, None, 0)
        if n < 0:
            _check(int(n), self.h)
        out = np.zeros(max(n, 1), np.uint64)
        lib().pbw_index_lookup(self.h, code, out.ctypes.data, n)
        return out[:n]

    @property
    def total_entries(self):
        return lib().pbw_index_total_entries(self.h)

    def stats(self):
        s = Stats()
        _check(lib().pbw_get_stats(self.h, C.byref(s)), self.h)
        return s.as_dict()

    # ---- candidate prefilter network ----
    def mlp_init(self, seed=1):
        _check(lib().pbw_mlp_init(self.h, seed), self.h)

    @property
    def mlp_params(self):
        th = np.zeros(MLP_PARAMS, np.float32)
        _check(lib().pbw_mlp_get_params(self.h, th.ctypes.data, MLP_PARAMS), self.h)
        return th

    @mlp_params.setter
    def mlp_params(self, theta):
        th = np.ascontiguousarray(theta, dtype=np.float32)
        _check(lib().pbw_mlp_set_params(self.h, th.ctypes.data, len(th)), self.h)

    def mlp_forward(self, x_bf16):
        x = np.ascontiguousarray(x_bf16, dtype=np.uint16).reshape(-1, MLP_INPUTS)
        z = np.zeros(len(x), np.float32)
        _check(lib().pbw_mlp_forward(self.h, x.ctypes.data, len(x), z.ctypes.data), self.h)
        return z

    def mlp_gradient(self, x_bf16, y, pos_weight=1.0):
        x = np.ascontiguousarray(x_bf16, dtype=np.uint16).reshape(-1, MLP_INPUTS)
        yy = np.ascontiguousarray(y, dtype=np.float32)
        g = np.zeros(MLP_PARAMS, np.float32)
        loss = C.c_double()
        _check(lib().pbw_mlp_gradient(self.h, x.ctypes.data, yy.ctypes.data, len(x), pos_weight, g.ctypes.data,
                                      C.byref(loss)), self.h)
        return g, loss.value

    # ---- prefilter in the mapping path ----
    @property
    def prefilter(self):
        o = PrefilterOptions()
        _check(lib().pbw_prefilter_get(self.h, C.byref(o)), self.h)
        return o

    @prefilter.setter
    def prefilter(self, o):
        _check(lib().pbw_prefilter_set(self.h, C.byref(o)), self.h)

    def set_prefilter(self, **kw):
        o = self.prefilter
        for k, v in kw.items():
            setattr(o, k, v)
        self.prefilter = o
        return o

    def prefilter_dataset(self, reads, label="loglk", margin=40.0):
        """(x_bf16 [n, 32], y [n], rank [n]) training rows from plain mapping of `reads`.
        label "loglk": y = 1 when the candidate's log-likelihood is within `margin` nats of the read's
        best; "score": when its DP score is within `margin` score units."""
        rs = reads.cstruct()
        kind = LABEL_KINDS[label]
        n = C.c_uint64()
        _check(lib().pbw_prefilter_dataset(self.h, C.byref(rs), kind, margin, None, None, None, 0, C.byref(n)),
               self.h)
        x = np.zeros((max(n.value, 1), MLP_INPUTS), np.uint16)
        y = np.zeros(max(n.value, 1), np.float32)
        r = np.zeros(max(n.value, 1), np.uint8)
        _check(lib().pbw_prefilter_dataset(self.h, C.byref(rs), kind, margin, x.ctypes.data, y.ctypes.data,
                                           r.ctypes.data, n.value, C.byref(n)), self.h)
        return x[:n.value], y[:n.value], r[:n.value]

    def prefilter_train(self, reads, label="loglk", margin=40.0, recall=0.9999, **kw):
        """Trains the network on `reads` (plain-DP labels, as in prefilter_dataset) and calibrates
        the threshold; `kw` are TrainOptions fields."""
        rs = reads.cstruct()
        o = TrainOptions(**kw)
        rep = PrefilterReport()
        _check(lib().pbw_prefilter_train(self.h, C.byref(rs), C.byref(o), LABEL_KINDS[label], margin, recall,
                                         C.byref(rep)), self.h)
        return rep.as_dict()

    def mlp_train(self, x_bf16, y, **kw):
        x = np.ascontiguousarray(x_bf16, dtype=np.uint16).reshape(-1, MLP_INPUTS)
        yy = np.ascontiguousarray(y, dtype=np.float32)
        o = TrainOptions(**kw)
        losses = np.zeros(o.epochs, np.float64)
        _check(lib().pbw_mlp_train(self.h, x.ctypes.data, yy.ctypes.data, len(x), C.byref(o), losses.ctypes.data),
               self.h)
        return losses


class DeviceReads:
    """Reads resident in device memory (`pbw_reads_upload` / `pbw_reads_free`)."""

    def __init__(self, ctx, reads):
        self.ctx = ctx  # keeps the context (and its device) alive
        self.n = reads.n
        h = C.c_void_p()
        rs = reads.cstruct()
        _check(lib().pbw_reads_upload(ctx.h, C.byref(rs), C.byref(h)), ctx.h)
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            lib().pbw_reads_free(self.h)
            self.h = None


def primary_identical(a, b):
    """Per record of two record sets for the same reads (same order): True where the SAM
    fields of the primary alignment agree (RNAME, POS, FLAG, MAPQ, strand and every CIGAR run)."""
    ra, rb = a.rec, b.rec
    if len(ra) != len(rb):
        raise ValueError("record sets differ in length")
    same = ((ra["pos"] == rb["pos"]) & (ra["contig"] == rb["contig"]) & (ra["flag"] == rb["flag"])
            & (ra["mapq"] == rb["mapq"]) & (ra["strand"] == rb["strand"]) & (ra["cigar_n"] == rb["cigar_n"]))
    n = np.where(same, ra["cigar_n"], 0).astype(np.int64)
    tot = int(n.sum())
    if tot:
        rid = np.repeat(np.arange(len(ra)), n)
        first = np.cumsum(n) - n
        k = np.arange(tot) - np.repeat(first, n)
        ca, cb = np.asarray(a.cigar), np.asarray(b.cigar)
        diff = ca[ra["cigar_off"].astype(np.int64)[rid] + k] != cb[rb["cigar_off"].astype(np.int64)[rid] + k]
        same[rid[diff]] = False
    return same


def format_sam(genome, reads, records, command_line="probebwa", read_group=None):
    """SAM text (src/io/sam.rs:37-107) for records aligned 1:1 with `reads`."""
    ids = b"".join(_b(i) + b"\0" for i in reads.ids)
    rs = reads.cstruct()
    rec = np.ascontiguousarray(records.rec)
    cig = np.ascontiguousarray(records.cigar) if len(records.cigar) else np.zeros(1, np.uint32)
    rb = _RecordBatch(len(rec), rec.ctypes.data, cig.ctypes.data, len(records.cigar))
    ln = C.c_uint64()
    t = lib().pbw_format_sam(genome.h, C.byref(rs), ids, C.byref(rb), _b(command_line),
                             _b(read_group) if read_group else None, C.byref(ln))
    if not t:
        _check(-1)
    s = C.string_at(t, ln.value)
    lib().pbw_free_text(t)
    return s


def _record_batch(records):
    rec = np.ascontiguousarray(records.rec)
    cig = np.ascontiguousarray(records.cigar) if len(records.cigar) else np.zeros(1, np.uint32)
    return _RecordBatch(len(rec), rec.ctypes.data, cig.ctypes.data, len(records.cigar)), (rec, cig)


def write_bam(genome, reads, records, path, command_line="probebwa", read_group=None, index=False,
              raw_qual=False, threads=0):
    """`write_bam` / `write_bam_indexed` (src/io/bam.rs:19-57) of the records `format_sam` would
    write; with `index` the records are coordinate-sorted and `path`.bai is written."""
    ids = b"".join(_b(i) + b"\0" for i in reads.ids)
    rs = reads.cstruct()
    rb, _keep = _record_batch(records)
    _check(lib().pbw_write_bam(genome.h, C.byref(rs), ids, C.byref(rb), _b(command_line),
                               _b(read_group) if read_group else None, _b(path), int(bool(index)),
                               int(bool(raw_qual)), int(threads)))


def read_fastx(path, fasta=False, quality_offset=33, adapter=None, batch=1 << 20):
    """`read_reads_files_with_preprocessing` for one file (src/io/fastx.rs:51-305):
    a ReadBatch with the record ids, through the library's reader."""
    h = C.c_void_p()
    _check(lib().pbw_fastx_open(_b(path), 1 if fasta else 0, int(quality_offset),
                                _b(adapter) if adapter else None, C.byref(h)))
    seqs, quals, ids = [], [], []
    try:
        while True:
            rb = _ReadBatch()
            idp = C.c_void_p()
            _check(lib().pbw_fastx_next(h, batch, C.byref(rb), C.byref(idp)))
            if rb.n == 0:
                break
            so = np.ctypeslib.as_array(C.cast(rb.seq_off, C.POINTER(C.c_uint64)), (rb.n + 1,))
            qo = np.ctypeslib.as_array(C.cast(rb.qual_off, C.POINTER(C.c_uint64)), (rb.n + 1,))
            sq = C.string_at(rb.seq, int(so[-1])) if so[-1] else b""
            ql = C.string_at(rb.qual, int(qo[-1])) if qo[-1] else b""
            at = idp.value
            for i in range(rb.n):
                seqs.append(sq[so[i]:so[i + 1]])
                quals.append(ql[qo[i]:qo[i + 1]])
                t = C.string_at(at)
                ids.append(t.decode("utf-8", "replace"))
                at += len(t) + 1
    finally:
        lib().pbw_fastx_close(h)
    return ReadBatch(seqs, quals, ids)


def io_options(**kw):
    o = IoOptions()
    lib().pbw_default_io_options(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, _b(v) if isinstance(v, str) else v)
    return o


def map_files(ctx, reads1, reads2=None, out_path=None, **kw):
    """The map subcommand (src/main.rs:130-195): read file(s) in, SAM or BAM out; returns the
    pbw_io_report as a dict."""
    o = io_options(**kw)
    rep = IoReport()
    _check(lib().pbw_map_files(ctx.h, _b(reads1), _b(reads2) if reads2 else None,
                               _b(out_path) if out_path else None, C.byref(o), C.byref(rep)))
    return rep.as_dict()


def code_of(kmer):
    """30-bit code of a 15-mer (the pre-mix() packing of hash_kmer, hashtable.rs:107-113)."""
    c = 0
    for ch in _b(kmer)[:15]:
        c = (c << 2) | {65: 0, 67: 1, 71: 2, 84: 3, 97: 0, 99: 1, 103: 2, 116: 3}.get(ch, 0)
    return c


def map_reads(genome, reads, options=None, device=0):
    """Single-end `map_reads` (src/lib.rs:49-56) over an in-memory genome and read batch."""
    return Context(genome, options, device).map_se(reads)
