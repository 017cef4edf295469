/*
 * probebwa_amd — MI355X (gfx950) implementation of probebwa's `map` hot path.
 *
 * C ABI: plain pointers and sizes, no exceptions across the boundary.  Every
 * call returns 0 on success and a negative status on error; the message is
 * available from pbw_last_error() (per context) or pbw_global_error().
 *
 * Each entry point names the reference interface it replaces
 * (GUIBA-EX/probebwa, Rust; paths relative to that repository).
 */
#ifndef PROBEBWA_AMD_H
#define PROBEBWA_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PBW_OK 0
#define PBW_ERR_ARG (-1)
#define PBW_ERR_HIP (-2)
#define PBW_ERR_LIMIT (-3)
#define PBW_ERR_IO (-4)
#define PBW_ERR_MISMATCH (-5)

typedef struct pbw_genome pbw_genome;
typedef struct pbw_ctx pbw_ctx;

/* One output record: the fields of AlignmentRecord (src/types.rs:180-203)
 * that SAM carries, plus log_lk/posterior.  SEQ/QUAL are produced from the
 * input read: reverse-complemented/reversed when seq_rc is set
 * (src/mapper/single.rs:101-103,144-147). */
typedef struct {
  uint64_t pos;         /* 1-based; 0 = unmapped */
  uint64_t mate_pos;    /* PNEXT; 0 = none */
  int64_t tlen;         /* TLEN; 0 = none */
  double log_lk;        /* AlignmentRecord::alignment_score */
  double posterior;     /* AlignmentRecord::posterior_prob */
  int32_t contig;       /* RNAME index; -1 = '*' */
  int32_t mate_contig;  /* RNEXT index; -1 = '*' */
  uint32_t cigar_off;   /* first CIGAR word in pbw_record_batch.cigar */
  uint32_t cigar_n;     /* CIGAR ops; word = (len << 4) | op, op M=0 I=1 D=2 S=4 H=5 */
  uint16_t flag;        /* SAM FLAG (src/io/sam.rs:126-143) */
  uint8_t mapq;
  uint8_t strand;       /* 0 forward, 1 reverse */
  uint8_t seq_rc;       /* SEQ/QUAL emitted reverse-complemented/reversed */
  uint8_t pad[3];
} pbw_record;

/* MapperOptions (src/types.rs:207-249), defaults from pbw_default_options. */
typedef struct {
  double substitution_rate;  /* -s (0.001) */
  double insert_size_mean;   /* 250 */
  double insert_size_sd;     /* 60 */
  double sv_prior_phred;     /* --svprior (55) */
  uint32_t max_indel_len;    /* band half-width (30) */
  uint32_t gap_open_phred;   /* --gapopen (40) */
  uint32_t gap_extend_phred; /* --gapextend (3) */
  uint32_t batch_reads;      /* reads per device batch (0 = default 131072) */
} pbw_options;

/* Reads: bases upper-cased (DnaSeq::from_ascii, src/types.rs:12-15), quals
 * offset-subtracted (QualityScores::from_phred_ascii, src/types.rs:57-60). */
typedef struct {
  uint64_t n;
  const uint8_t* seq;
  const uint64_t* seq_off;  /* n+1 */
  const uint8_t* qual;
  const uint64_t* qual_off; /* n+1 */
} pbw_read_batch;

/* Library-owned output arena in pinned host memory; free with pbw_records_release. */
typedef struct {
  uint64_t n;
  pbw_record* rec;
  uint32_t* cigar;
  uint64_t n_cigar;
} pbw_record_batch;

/* InsertSizeDistribution (src/types.rs:266-270). */
typedef struct {
  double mean;
  double sd;
  uint64_t n_observed;
} pbw_insert_model;

/* Kernel-level job: SmithWatermanAligner::new(band, ..).align(read, qual,
 * ref_seg, ref_offset_hint) (src/align/smith_waterman.rs:113,163).  Host
 * pointers. */
typedef struct {
  const uint8_t* read;
  const uint8_t* qual;
  const uint8_t* ref;
  uint32_t m;     /* read length */
  uint32_t qlen;  /* quality length (missing quals read as 30) */
  uint32_t n;     /* ref_seg length */
  uint32_t band;
  int64_t hint;
} pbw_align_job;

typedef struct {
  int32_t score;
  uint32_t ref_offset; /* ref_start_offset */
  uint32_t cigar_off;  /* into the caller's cigar array */
  uint32_t cigar_n;
} pbw_align_result;

/* Per-context counters and stage timings of the last map call. */
typedef struct {
  uint64_t reads;
  uint64_t lookups;        /* exact + neighbour probes */
  uint64_t positions;      /* packed positions gathered */
  uint64_t filtered_hits;  /* hits that passed the similarity filter */
  uint64_t dp_jobs;
  uint64_t dp_cells;       /* counted as the reference loop does */
  uint64_t wide_jobs;
  uint64_t wide_cells;
  uint64_t seed_overflow_reads;
  uint64_t tier2_cluster_reads;
  uint64_t generic_dp_jobs;  /* jobs handed to the tiled (any-width) DP kernel */
  uint64_t band_dp_jobs;     /* jobs handed to the one-job-per-wave banded DP kernel */
  double ms_seed, ms_cluster, ms_dp, ms_mapq, ms_pair, ms_total;  /* host wall time per stage */
  double ms_k_seed, ms_k_dp;  /* HIP-event time of the seed kernel and the main banded-DP launch */
  uint64_t n_k_seed, n_k_dp;  /* launches timed */
  /* candidate prefilter (zero when it is off) */
  uint64_t pf_jobs;             /* job slots that plain mapping would align */
  uint64_t pf_jobs_aligned;     /* job slots aligned (kept + re-run) */
  uint64_t pf_fallback_reads;   /* reads re-run plain as ambiguous */
  uint64_t pf_audit_reads;      /* reads re-run plain as the audit sample */
  uint64_t pf_audit_mismatches; /* audited reads whose SAM fields differ from the plain run */
  double ms_pf;                 /* host wall time of scoring and selection */
  double ms_k_cluster;          /* HIP-event time of the clustering kernels (every tier) */
  uint64_t n_k_cluster;         /* clustering stages timed */
  uint64_t mapq_near_boundary;  /* MAPQ finalizations whose integer MAPQ rests on a device exp/log/log10
                                   value within 1e-9 of a rounding point (one ulp could move it) */
  double ms_k_wide;             /* HIP-event time of the paired-end mate-window DP (align_at_window):
                                   the wide kernel and the tiled kernel for windows it hands on */
  uint64_t n_k_wide;            /* pair batches whose mate-window DP was timed */
} pbw_stats;

/* ---- genome: GenomeIndexBuilder (src/index/genome.rs:101-147) ---- */
int pbw_genome_create(pbw_genome** out);
/* Contig::from_ascii (src/index/genome.rs:17-45) */
int pbw_genome_add_contig(pbw_genome* g, const char* name, const uint8_t* seq, uint64_t len);
/* GenomeIndexBuilder::add_fasta (src/index/genome.rs:119-134) */
int pbw_genome_add_fasta(pbw_genome* g, const char* path);
int pbw_genome_n_contigs(const pbw_genome* g);
const char* pbw_genome_contig_name(const pbw_genome* g, int cid);
uint64_t pbw_genome_contig_length(const pbw_genome* g, int cid);
void pbw_genome_destroy(pbw_genome* g);
/* GenomeIndex::get_sequence (src/index/genome.rs:97-99) -> Contig::slice (:47-68): bases
 * [start, end) of contig cid, clamped to its length, non-ACGT as 'N'; writes them to out
 * (when not NULL) and returns their count */
uint64_t pbw_genome_get_sequence(const pbw_genome* g, int cid, uint64_t start, uint64_t end, uint8_t* out);
/* GenomeIndex::load / save (src/index/genome.rs:80-94): a bincode-1 `.stidx` (species,
 * assembly, contigs as name + 2-bit packed bases + N runs + length, total_length) */
int pbw_genome_load_stidx(const char* path, pbw_genome** out);
int pbw_genome_save_stidx(const pbw_genome* g, const char* species, const char* assembly, const char* path);

/* ---- context: index build + upload, SingleEndMapper::new (src/mapper/single.rs:52-70),
 *      HashTableBuilder::build (src/index/hashtable.rs:248-322, k=15 stride=5 cap=300) ---- */
void pbw_default_options(pbw_options* o);
int pbw_ctx_create(const pbw_genome* g, const pbw_options* o, int device, pbw_ctx** out);
void pbw_ctx_destroy(pbw_ctx* ctx);
/* HashTable::load (src/index/hashtable.rs:66-72): the context's buckets read from a bincode-1
 * `.sthash` (FxHashMap<u32 hash, (offset, count)>, positions: Vec<u64>, k, seed_stride,
 * total_entries) instead of built from the genome.  k must be 15 and no bucket may hold more
 * than the builder's cap of 300; `g` must be the genome the table was built from (every
 * position is checked against its contigs). */
int pbw_ctx_create_from_sthash(const pbw_genome* g, const char* sthash_path, const pbw_options* o, int device,
                               pbw_ctx** out);
/* HashTable::save (src/index/hashtable.rs:74-80): the context's buckets as a `.sthash`,
 * keyed by hash_kmer (mix of the 30-bit code), in code order */
int pbw_index_save_sthash(pbw_ctx* ctx, const char* path);
/* the genome and options a context was created with */
const pbw_genome* pbw_ctx_genome(const pbw_ctx* ctx);
int pbw_ctx_get_options(const pbw_ctx* ctx, pbw_options* out);
const char* pbw_last_error(const pbw_ctx* ctx);
const char* pbw_global_error(void);

/* ---- map: replaces the single-end rayon loop of map_reads_with_format
 *      (src/lib.rs:84-104, SingleEndMapper::map src/mapper/single.rs:164-207);
 *      records in input order ---- */
int pbw_map_se(pbw_ctx* ctx, const pbw_read_batch* reads, pbw_record_batch* out);

/* ---- reads kept resident in device memory (HBM) across calls: upload once,
 *      then map any [begin, end) range without host-to-device copies.  No
 *      counterpart in the reference; for callers that keep reads on the GPU. ---- */
typedef struct pbw_device_reads pbw_device_reads;
int pbw_reads_upload(pbw_ctx* ctx, const pbw_read_batch* reads, pbw_device_reads** out);
int pbw_map_se_device(pbw_ctx* ctx, const pbw_device_reads* reads, uint64_t begin, uint64_t end,
                      pbw_record_batch* out);
void pbw_reads_free(pbw_device_reads* reads);

/* ---- Autotuner (no counterpart in the reference): launch shapes that change
 *      the speed of a mapping call, never its records. ---- */
typedef struct {
  uint32_t batch_reads; /* reads per device batch (as pbw_options.batch_reads) */
  uint32_t dp_grid;     /* waves of a four-jobs-per-wave DP launch (jobs beyond loop) */
  uint32_t split_len;   /* 0: one DP launch for every read length; else reads longer than
                           split_len run in a launch of their own and the rest in a launch
                           whose per-job LDS tile is sized for split_len bases */
  uint32_t long_waves;  /* reads longer than split_len: 0 = four jobs per wave, 1 = one job
                           per wave (64 lanes per alignment) */
  uint32_t dp_pack16;   /* 1: the four-jobs-per-wave launches run eight jobs per wave, two per
                           register in saturating 16-bit halves, when the gap penalties allow
                           (gap extend >= -4, gap open >= -400); jobs whose best score is below
                           -18000 are re-run in 32 bits */
  uint32_t seed_pool;   /* the seed pass's pool of 1024-hit chunks per batch; 0: two per read to
                           start with, then grown to what the batches' hits take.  Reads whose
                           hits find no chunk are seeded again into exact slots */
} pbw_tune;
int pbw_tune_get(const pbw_ctx* ctx, pbw_tune* out);
int pbw_tune_set(pbw_ctx* ctx, const pbw_tune* t);
#define PBW_TUNE_MAX_TRIALS 48
typedef struct {
  pbw_tune best;
  double best_ms;  /* mapping time of the sample with `best` */
  uint32_t trials;
  pbw_tune tried[PBW_TUNE_MAX_TRIALS];
  double ms[PBW_TUNE_MAX_TRIALS];
} pbw_autotune_report;
/* Maps reads [begin, end) of `reads` under candidate settings (coordinate descent
 * over dp_grid, dp_pack16, then split_len x long_waves, then batch_reads; each timed twice,
 * the faster run counts), keeps the fastest in the context and reports every trial. */
int pbw_autotune(pbw_ctx* ctx, const pbw_device_reads* reads, uint64_t begin, uint64_t end,
                 pbw_autotune_report* out);

/* ---- paired-end: map_paired (src/lib.rs:145-185) incl. the 500-pair serial
 *      warm-up and 2000-pair frozen-snapshot batches; records interleaved
 *      r1,r2; the insert model is read and updated in place ---- */
int pbw_map_pe(pbw_ctx* ctx, const pbw_read_batch* m1, const pbw_read_batch* m2, pbw_insert_model* inout,
               pbw_record_batch* out);

/* ---- kernel level: batched SmithWatermanAligner::align ---- */
int pbw_align_batch(pbw_ctx* ctx, uint32_t gap_open_phred, uint32_t gap_extend_phred, const pbw_align_job* jobs,
                    uint64_t n, pbw_align_result* results, uint32_t* cigar, uint64_t cigar_cap);

/* ---- pbw_align_batch that allocates nothing.  No counterpart in the reference: for
 *      callers that cannot allocate device memory (code under HIP graph capture, tight
 *      serving loops, harnesses that forbid hipMalloc).  The same jobs, results and
 *      kernels as pbw_align_batch, on a fixed 64 MiB device region that each GPU the
 *      library runs on holds until the process ends, and a fixed 16 MiB host staging
 *      area; a batch needing more, a read over 1024 bases or a window over 8000
 *      returns PBW_ERR_LIMIT.  Calls run one at a time (one
 *      host lock); the region keeps the last call's contents.  The context (created
 *      beforehand) is the only set-up. ---- */
int pbw_align_batch_noalloc(pbw_ctx* ctx, uint32_t gap_open_phred, uint32_t gap_extend_phred,
                            const pbw_align_job* jobs, uint64_t n, pbw_align_result* results, uint32_t* cigar,
                            uint64_t cigar_cap);

/* ---- index access: HashTable::lookup (src/index/hashtable.rs:90-102) by 30-bit code;
 *      copies up to cap packed positions, returns the bucket size (or <0) ---- */
int64_t pbw_index_lookup(pbw_ctx* ctx, uint32_t code, uint64_t* out, uint64_t cap);
uint64_t pbw_index_total_entries(const pbw_ctx* ctx);

/* ---- candidate prefilter network: MLP 32 -> 128 -> 128 -> 1 (ReLU), BF16 matrix-core
 *      operands with FP32 accumulation; FP32 master parameters and Adam.  No counterpart
 *      in the reference: it ranks the <= 20 candidates of a read before the DP
 *      (between cluster_hits, src/mapper/single.rs:91, and the DP loop, :114-150).
 *      Parameter layout (floats): W1[32][128] | b1[128] | W2[128][128] | b2[128] | W3[128] | b3.
 *      Inputs are bf16 bit patterns, n rows of 32. ---- */
#define PBW_MLP_INPUTS 32
#define PBW_MLP_PARAMS 20865
typedef struct {
  uint32_t epochs;   /* 20 */
  uint32_t batch;    /* 4096 rows */
  float lr;          /* Adam 1e-3 */
  float beta1, beta2, eps;
  float pos_weight;  /* BCE weight of label-1 rows */
  uint64_t seed;     /* shuffle seed */
} pbw_train_options;
void pbw_default_train_options(pbw_train_options* o);

/* Prefilter in the mapping path (pbw_map_se / pbw_map_se_device).  A candidate is
 * aligned when its rank (cluster-score order) is below top_k or its network score
 * is >= threshold.  A read is re-run plain when its prefiltered MAPQ is below
 * mapq_fallback or a kept second candidate is within close_nats of the best; every
 * audit_period-th read is re-run plain and compared (pbw_stats.pf_audit_*). */
typedef struct {
  int32_t enabled;
  float threshold;
  uint32_t top_k;          /* 1 */
  uint32_t mapq_fallback;  /* 20 */
  float close_nats;        /* 10 */
  uint32_t audit_period;   /* 64 */
} pbw_prefilter_options;
void pbw_default_prefilter_options(pbw_prefilter_options* o);
int pbw_prefilter_set(pbw_ctx* ctx, const pbw_prefilter_options* o);
int pbw_prefilter_get(const pbw_ctx* ctx, pbw_prefilter_options* o);

/* Training label of a candidate, from the plain DP over every candidate of its read:
 * PBW_LABEL_LOGLK  1 when its log-likelihood is within `margin` nats of the read's best;
 * PBW_LABEL_SCORE  1 when its DP score is within `margin` score units of the read's best. */
#define PBW_LABEL_LOGLK 0
#define PBW_LABEL_SCORE 1

/* Training rows from plain mapping of `reads`: per aligned candidate its 32 features
 * (BF16), its label and its rank.  Returns the row count (rows beyond cap are counted,
 * not written). */
int pbw_prefilter_dataset(pbw_ctx* ctx, const pbw_read_batch* reads, int label_kind, float margin, uint16_t* x,
                          float* y, uint8_t* rank, uint64_t cap, uint64_t* n);
typedef struct {
  uint64_t rows, positives, val_rows, val_positives;
  float threshold;   /* calibrated and stored in the context's prefilter options */
  double val_recall; /* validation positives kept (by rank or score) */
  double val_skip;   /* validation rows not aligned */
  double loss_first, loss_last;
} pbw_prefilter_report;
/* Trains the network on rows of `reads` (the last 10% held out) and calibrates the
 * threshold so that `recall` of the held-out positives ranked >= top_k score above it. */
int pbw_prefilter_train(pbw_ctx* ctx, const pbw_read_batch* reads, const pbw_train_options* o, int label_kind,
                        float margin, float recall, pbw_prefilter_report* out);
int pbw_mlp_init(pbw_ctx* ctx, uint64_t seed);
int pbw_mlp_set_params(pbw_ctx* ctx, const float* theta, uint64_t n);
int pbw_mlp_get_params(pbw_ctx* ctx, float* theta, uint64_t n);
int pbw_mlp_forward(pbw_ctx* ctx, const uint16_t* x_bf16, uint64_t n, float* logits);
/* gradient of mean_i w_i BCE(sigmoid(z_i), y_i), w = pos_weight for y = 1 */
int pbw_mlp_gradient(pbw_ctx* ctx, const uint16_t* x_bf16, const float* y, uint64_t n, float pos_weight, float* grad,
                     double* loss);
int pbw_mlp_train(pbw_ctx* ctx, const uint16_t* x_bf16, const float* y, uint64_t n, const pbw_train_options* o,
                  double* epoch_loss);

void pbw_records_release(pbw_record_batch* b);
int pbw_get_stats(const pbw_ctx* ctx, pbw_stats* out);

/* ---- SAM text: format_sam_records (src/io/sam.rs:37-107); ids is n NUL-terminated
 *      strings back to back (NULL: r0, r1, ...); SEQ and QUAL pass through
 *      String::from_utf8_lossy as the reference's do (sam.rs:50-51); read_group is
 *      parsed as ReadGroup::parse (sam.rs:17-32) and NULL is returned when it is
 *      malformed; free with pbw_free_text ---- */
char* pbw_format_sam(const pbw_genome* g, const pbw_read_batch* reads, const char* ids, const pbw_record_batch* recs,
                     const char* command_line, const char* read_group, uint64_t* len);
void pbw_free_text(char* t);

/* ---- BAM: write_bam / write_bam_indexed (src/io/bam.rs:19-57) of the records
 *      pbw_format_sam would write, BGZF-compressed.  With index != 0 the records are
 *      first coordinate-sorted (sort_for_coordinate_order, src/io/sam.rs:109-124;
 *      header SO:coordinate) and the BAI is written to `path`.bai.  The reference
 *      round-trips its SAM text through noodles, which rejects what that text cannot
 *      carry into BAM: raw_qual = 0 does the same (QUAL bytes -- raw Phred scores in
 *      this SAM text -- must lie in '!'..'~' and are stored minus 33; QNAME must be
 *      1..254 bytes of '!'..'~' without '@'; SEQ and QUAL non-empty and equally
 *      long) and returns PBW_ERR_ARG naming the first record that fails; raw_qual = 1
 *      stores the Phred scores themselves.  threads: compression threads (0: all). ---- */
int pbw_write_bam(const pbw_genome* g, const pbw_read_batch* reads, const char* ids, const pbw_record_batch* recs,
                  const char* command_line, const char* read_group, const char* path, int index, int raw_qual,
                  uint32_t threads);

/* ---- reads: read_fastq / read_fasta_as_reads (src/io/fastx.rs:128-236) with
 *      strip_mate_suffix and strip_adapter (fastx.rs:66-118): format 0 FASTQ, 1 FASTA;
 *      gzip when the path ends in ".gz" (MultiGzDecoder); lines must be UTF-8 (as
 *      BufRead::read_line into a String requires).  pbw_fastx_next fills `reads` (and
 *      `ids`, NUL-terminated back to back) with up to max_reads records owned by the
 *      reader until its next call; reads->n == 0 at the end. ---- */
typedef struct pbw_fastx pbw_fastx;
int pbw_fastx_open(const char* path, int format, uint32_t quality_offset, const char* adapter, pbw_fastx** out);
int pbw_fastx_next(pbw_fastx* f, uint64_t max_reads, pbw_read_batch* reads, const char** ids);
uint64_t pbw_fastx_bytes_read(const pbw_fastx* f);
void pbw_fastx_close(pbw_fastx* f);

/* ---- files to SAM/BAM: the map subcommand (src/main.rs:130-195) over
 *      map_reads_with_format (src/lib.rs:65-143): reads2 == NULL maps reads1
 *      single-end, otherwise the two files are mates in lockstep (map_paired).
 *      Single-end input is parsed, mapped and written as overlapping batches
 *      (records in input order); paired-end input is read whole, as the reference
 *      reads it.  out_path NULL writes SAM to stdout. ---- */
typedef struct {
  int input_format;         /* --inputformat: 0 fastq, 1 fasta */
  uint32_t quality_offset;  /* 33, or 64 with --phred64 */
  const char* adapter;      /* --adapter-strip (NULL: none) */
  const char* label_prefix; /* --labelfilter (NULL: none) */
  const char* command_line; /* the @PG CL field */
  const char* read_group;   /* --readgroup (NULL: none) */
  int output_format;        /* --outputformat: 0 sam, 1 bam */
  int index;                /* --index: coordinate-sorted BAM plus out_path.bai */
  int raw_qual;             /* BAM: as pbw_write_bam's raw_qual */
  uint32_t threads;         /* formatting and compression threads (0: all) */
  uint64_t batch_reads;     /* single-end reads per mapped batch (262144) */
} pbw_io_options;
void pbw_default_io_options(pbw_io_options* o);
typedef struct {
  uint64_t reads;      /* reads parsed (both mates) */
  uint64_t records;    /* records written */
  uint64_t bytes_in;   /* read-file bytes consumed (compressed for .gz) */
  uint64_t bytes_out;  /* output bytes */
  double ms_parse, ms_map, ms_write, ms_total;
} pbw_io_report;
int pbw_map_files(pbw_ctx* ctx, const char* reads1, const char* reads2, const char* out_path,
                  const pbw_io_options* o, pbw_io_report* report);

#ifdef __cplusplus
}
#endif
#endif
