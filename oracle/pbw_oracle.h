/*
 * CPU oracle for the probebwa map path — TEST INFRASTRUCTURE ONLY.
 *
 * This library is a scalar C++ restatement of the reference mapper's
 * algorithm (GUIBA-EX/probebwa, Rust).  It exists so that tests, the smoke
 * check and the CPU-baseline leg of bench.py can check the HIP product path
 * against it.  Nothing in probebwa_amd/ links, loads or calls it.
 *
 * Every function in pbw_oracle.cpp cites the reference file:line it restates.
 * Parity status of each restated function (what pins it) is listed at the top
 * of pbw_oracle.cpp.
 */
#ifndef PBW_ORACLE_H
#define PBW_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Record layout shared (byte for byte) with include/probebwa_amd.h's
 * pbw_record so one test harness can diff both. */
typedef struct {
  uint64_t pos;         /* 1-based; 0 = unmapped */
  uint64_t mate_pos;    /* 0 = none */
  int64_t tlen;         /* 0 = none */
  double log_lk;        /* AlignmentRecord::alignment_score */
  double posterior;     /* AlignmentRecord::posterior_prob */
  int32_t contig;       /* -1 = '*' */
  int32_t mate_contig;  /* -1 = none */
  uint32_t cigar_off;   /* into the batch's cigar word array */
  uint32_t cigar_n;     /* number of ops */
  uint16_t flag;        /* SAM FLAG */
  uint8_t mapq;
  uint8_t strand;       /* 0 forward, 1 reverse */
  uint8_t seq_rc;       /* SEQ/QUAL are reverse-complemented/reversed */
  uint8_t pad[3];
} orc_record;

typedef struct {
  double substitution_rate;  /* -s, default 0.001 */
  double insert_size_mean;   /* default 250 */
  double insert_size_sd;     /* default 60 */
  double sv_prior_phred;     /* --svprior, default 55 */
  uint32_t max_indel_len;    /* default 30 */
  uint32_t gap_open_phred;   /* --gapopen, default 40 */
  uint32_t gap_extend_phred; /* --gapextend, default 3 */
  uint32_t threads;          /* oracle worker threads (SE only) */
} orc_options;

typedef struct {
  uint64_t lookups;        /* exact + neighbour hash probes */
  uint64_t positions;      /* packed positions returned by those probes */
  uint64_t filtered_hits;  /* seed hits that passed the similarity filter */
  uint64_t candidates;     /* clusters produced */
  uint64_t dp_jobs;        /* SmithWatermanAligner::align calls */
  uint64_t dp_cells;       /* cells computed, counted as the reference loop does */
  uint64_t wide_jobs;      /* align_at_window calls (paired-end) */
  uint64_t wide_cells;
} orc_stats;

/* Reads as one flat batch: bases already upper-cased, quals already offset-
 * subtracted (exactly what the reference's Read holds). */
typedef struct {
  uint64_t n;
  const uint8_t* seq;
  const uint64_t* seq_off;  /* n+1 */
  const uint8_t* qual;
  const uint64_t* qual_off; /* n+1 */
} orc_reads;

typedef struct {
  uint64_t n;
  orc_record* rec;
  uint32_t* cigar;
  uint64_t n_cigar;
} orc_records;

/* ---- genome (genome.rs) ---- */
void* orc_genome_new(void);
void orc_genome_free(void* g);
int orc_genome_add_contig(void* g, const char* name, const uint8_t* seq, uint64_t len);
int orc_genome_add_fasta(void* g, const char* path);
uint64_t orc_genome_total_length(const void* g);
int orc_genome_n_contigs(const void* g);
uint64_t orc_genome_contig_length(const void* g, int cid);
const char* orc_genome_contig_name(const void* g, int cid);
/* writes min(end,len)-min(start,len) bytes (or 0) and returns that count */
uint64_t orc_genome_slice(const void* g, int cid, uint64_t start, uint64_t end, uint8_t* out);

/* ---- hash table (hashtable.rs) ---- */
uint32_t orc_hash_kmer(const uint8_t* kmer, uint64_t len);
uint64_t orc_pack_position(uint64_t contig, uint64_t pos, int strand);
void orc_unpack_position(uint64_t packed, uint64_t* contig, uint64_t* pos, int* strand);
int orc_is_acgt_only(const uint8_t* kmer, uint64_t len);
/* threads used by the next orc_hash_build calls (default 1); the table is the same for any count */
void orc_set_index_threads(int n);
void* orc_hash_build(const void* g, int k, int stride, int max_count);
void orc_hash_free(void* h);
/* returns bucket size; *out points into the table's position store */
uint64_t orc_hash_lookup(const void* h, const uint8_t* kmer, uint64_t len, const uint64_t** out);
uint64_t orc_hash_total_entries(const void* h);
uint64_t orc_hash_n_keys(const void* h);
/* i-th key in ascending hash order: hash, bucket size, pointer to bucket */
uint64_t orc_hash_key_at(const void* h, uint64_t i, uint32_t* hash, const uint64_t** bucket);
/* mix() never maps two 30-bit 15-mer codes to one hash (checks all 2^30). */
int orc_mix_is_bijective(void);

/* ---- aligner (smith_waterman.rs) ---- */
int orc_sw_align(uint32_t band, uint32_t gap_open_phred, uint32_t gap_extend_phred,
                 const uint8_t* read, uint64_t m, const uint8_t* qual, uint64_t qlen,
                 const uint8_t* ref, uint64_t n, int64_t hint,
                 int32_t* score, uint32_t* cigar, uint32_t cigar_cap, uint32_t* n_ops,
                 uint64_t* ref_offset, uint64_t* cells);
int orc_mismatch_penalty(uint8_t q);
int orc_homopolymer_gap_open(int gap_open, uint16_t run);

/* ---- mapq (bayesian.rs) building blocks exposed for closed-form tests ---- */
double orc_logadd(double a, double b);
double orc_phred_n_choose_m(int64_t n, int64_t m);
double orc_effective_alphabet_size(double gc);
double orc_missing_locus_probability(double sub_rate, const uint8_t* qual, uint64_t len,
                                     const double* repeat_mask, uint64_t mask_len);
void orc_finalize_mapq(double sub_rate, uint64_t genome_size, double sum_exp,
                       const uint8_t* read_qual, uint64_t read_qlen,
                       const uint32_t* cigar, uint32_t n_ops,
                       const uint8_t* qbases, uint64_t qlen, const uint8_t* qqual, uint64_t qqlen,
                       const uint8_t* ref_window, uint64_t rwlen,
                       const double* repeat_mask, uint64_t mask_len,
                       uint8_t* mapq, double* posterior);
/* bayesian.rs:306-373 log_entropy of one alignment (cigar words (len << 4) | op) */
double orc_log_entropy(double sub_rate, uint64_t gsize, const uint32_t* cigar, uint32_t n_ops, const uint8_t* qb,
                       uint64_t qlen, const uint8_t* qq, uint64_t qqlen, const uint8_t* rw, uint64_t rwlen);
/* paired.rs:553-566 shortlist_indices; out holds up to n (at most 20) indices */
uint64_t orc_shortlist_indices(const double* post, uint64_t n, uint64_t* out);
/* paired.rs:346-351 dedup of raw pairs: loci are n x (contig, position, strand, contig, position, strand);
 * out receives the kept indices in order */
uint64_t orc_dedup_pairs(const uint64_t* loci, uint64_t n, uint64_t* out);
/* paired.rs:398-435 mapq_for_winning_pair: one mate's read, winning alignment (forward orientation),
 * own candidate log-likelihoods and repeat mask */
typedef struct {
  const uint8_t* seq;
  uint64_t len;
  const uint8_t* qual;
  uint64_t qlen;
  const uint32_t* cigar;
  uint32_t n_ops;
  const uint8_t* ref_window;
  uint64_t rwlen;
  const double* lk;
  uint64_t nlk;
  const double* mask;
  uint64_t mlen;
} orc_pair_mate;
void orc_mapq_for_winning_pair(const void* mapper, double pair_se, int bxo, int byo, const orc_pair_mate* a,
                               const orc_pair_mate* b, uint8_t* q, double* p);
void orc_insert_update(double* mean, double* sd, uint64_t* n_obs, double observed);
double orc_insert_log_likelihood(double mean, double sd, double insert_size);

/* ---- mapping (single.rs, paired.rs, lib.rs) ---- */
void orc_default_options(orc_options* o);
void* orc_mapper_new(const void* g, const void* h, const orc_options* o);
void orc_mapper_free(void* m);
/* Seed hits and candidates of one read (seeds.rs + candidates.rs), for kernel-level
 * parity.  hits: 5 u64 per hit (read_pos, contig, ref_pos, strand, is_neighbour).
 * cands: 4 u64 per candidate (contig, position, strand, score). */
int orc_seed_and_cluster(const void* m, const uint8_t* seq, uint64_t L, const uint8_t* qual, uint64_t qlen,
                         uint64_t* hits, uint64_t hits_cap, uint64_t* n_hits,
                         double* repeat_prob, uint64_t* cands, uint64_t cands_cap, uint64_t* n_cands);
int orc_map_se(const void* m, const orc_reads* reads, orc_records* out, orc_stats* stats);
int orc_map_pe(const void* m, const orc_reads* r1, const orc_reads* r2, orc_records* out,
               double* insert_mean, double* insert_sd, uint64_t* n_obs, orc_stats* stats);
void orc_records_release(orc_records* r);
/* Emulated cross-group order of cluster_hits: 0 = FxHashMap/SwissTable emulation
 * (default), 1 = (contig, forward<reverse) lexicographic. */
void orc_set_group_order(int mode);

/* ---- I/O (fastx.rs, sam.rs) ---- */
/* Parses a FASTQ (or FASTA when fasta!=0) file into a read batch plus ids
 * (NUL-separated).  adapter may be NULL.  Returns 0 on success. */
typedef struct {
  uint64_t n;
  uint8_t* seq; uint64_t* seq_off;
  uint8_t* qual; uint64_t* qual_off;
  char* ids; uint64_t* id_off;
} orc_parsed_reads;
int orc_read_reads(const char* path, int fasta, uint8_t quality_offset, const char* adapter,
                   orc_parsed_reads* out);
void orc_parsed_reads_free(orc_parsed_reads* p);
/* SAM text for the records (header + one line per record); caller frees with
 * orc_free_text. */
char* orc_format_sam(const void* g, const orc_parsed_reads* reads, const orc_records* recs,
                     const char* command_line, const char* read_group, uint64_t* len);
void orc_free_text(char* t);

#ifdef __cplusplus
}
#endif
#endif
