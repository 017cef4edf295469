// Shared host/device definitions of the probebwa_amd HIP path.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

namespace pbw {

// ---- constants of the reference algorithm ----
constexpr int K = 15;                    // hashtable.rs:191 (k = 15)
constexpr int STRIDE = 5;                // hashtable.rs:187 DEFAULT_STRIDE
constexpr int MAX_COUNT = 300;           // hashtable.rs:51 DEFAULT_MAX_COUNT
constexpr int REPEAT_SKIP = 64;          // seeds.rs:45
constexpr int MERGE_DISTANCE = 50;       // candidates.rs:27
constexpr int TOP_CANDIDATES = 20;       // single.rs:114
constexpr int LEFT_PAD = 10;             // single.rs:109
constexpr int MATCH_SCORE = 20;          // smith_waterman.rs:81
constexpr int32_t NEG_INF = INT32_MIN / 2;  // smith_waterman.rs:173
constexpr uint64_t CODE_SPACE = 1ull << 30;  // 4^15 15-mer codes
constexpr uint16_t GENOME_FP_OPEN = 0xFFFF;  // no complete genome flanks (counts never reach 15)
constexpr int DIAG_BITS = 41;            // hit key: (group << 41) | (diag + 2^40)
constexpr int64_t DIAG_BIAS = 1ll << 40;

// ---- limits of this implementation ----
constexpr int MAX_READ_LEN = 1024;       // reads up to this are staged in LDS by the seed kernel; longer ones are read from device memory
constexpr int HIT_CAP = 1024;            // least hits per read in the first seed pass (overflow -> exact rerun)
constexpr int HIT_CAP_MAX = 32768;       // most: the first-pass capacity adapts between the two
constexpr int CLUSTER_LDS_CAP = 512;     // hits per read clustered in LDS by one wave (else tier 2)
constexpr int LN_INT_MAX = 4096;         // ln(k) table for phred_n_choose_m
constexpr int GC_LUT_MAX = 1024;         // (gc, total) table for the GC-adjusted alphabet size
constexpr int QUAD_MAX_M = 320;         // longest read the four-jobs-per-wave DP kernel takes
constexpr int QUAD_MAX_N = QUAD_MAX_M + 64;  // longest reference window it takes

struct DevGenome {
  const uint64_t* g2b;    // 2-bit bases, 32 per word, LSB first (A0 C1 G2 T3), contigs back to back
  const uint64_t* nmask;  // 1 bit per base: base was not A/C/G/T (decodes as 'N')
  const uint64_t* coff;   // base offset of each contig in the arrays above (multiple of 64)
  const uint64_t* clen;   // contig lengths
  int32_t ncontigs;
};

// Index entries with their flank fingerprint inside, for genomes of at most 2^15
// contigs of at most 2^32 bases: bit 63 strand, 62..48 contig, 47..32 fingerprint,
// 31..0 position (one 8-byte load per entry in the seed pass; larger genomes keep
// the reference's packing and the fingerprints in an array beside it).
constexpr uint64_t FUSED_MAX_CONTIGS = 1ull << 15, FUSED_MAX_LEN = 1ull << 32;
__host__ __device__ inline uint64_t fused_entry(uint64_t strand, uint64_t contig, uint64_t fp, uint64_t pos) {
  return (strand << 63) | (contig << 48) | (fp << 32) | pos;
}
// a fused entry in the reference's packing (pack_position, hashtable.rs:157-164)
__host__ __device__ inline uint64_t fused_to_packed(uint64_t e) {
  return (e & (1ull << 63)) | (((e >> 48) & 0x7FFFull) << 40) | (e & 0xFFFFFFFFull);
}
struct DevIndex {
  const uint32_t* koff;  // [2^30 + 1] bucket start by 15-mer code; bucket = [koff[c], koff[c+1])
  const uint64_t* kpos;  // entries: fused (kfp null), else packed positions (bit 63 strand, 62..40 contig, 39..0 position)
  const uint64_t* pres;  // [2^24] bit c%64 of word c/64 set when bucket c is non-empty
  // the same bitmap indexed by the code rotated right by 10 and by 20 bits, so that
  // the 1-mismatch neighbours at window positions 5-9 (resp. 0-4) share the
  // original code's 1024-bit line, as positions 10-14 do in `pres`
  const uint64_t* pres_r10;
  const uint64_t* pres_r20;
  const uint16_t* kfp;   // per entry of kpos: genome flank fingerprint (GENOME_FP_OPEN: flanks off the
                         // contig); null when the entries are fused
};

struct DevReads {
  const uint8_t* seq;
  const uint64_t* seq_off;
  const uint8_t* qual;
  const uint64_t* qual_off;
  uint32_t n;
};

// Tables computed on the host with the same libm calls the reference makes,
// so every quality-indexed FP64 term is the identical double on the device.
struct DevLuts {
  const double* lk_match;  // [256] ln(max(1-e_q, 1e-300)), e_q = 10^(-logadd(q,sub)/10)
  const double* lk_mis;    // [256] ln(max(e_q/3, 1e-300))
  const double* omq;       // [256] 1 - min(e_q, 1-1e-12)
  const double* qratio;    // [256] q'/(1-q'), q' = min(e_q, 1-1e-12)
  const double* nrp;       // [256] 10^(-q/10) (neighbour repeat probability)
  const int32_t* mis_pen;  // [256] mismatch_penalty(q)
  const double* ln_int;    // [LN_INT_MAX+1] ln(k)
  const double* gc_lnb;    // [(GC_LUT_MAX+1)(GC_LUT_MAX+2)/2] ln(b(gc/total))
  const double* gc_lnbm1;  // same layout: ln(max(b-1, 1e-6))
  double ln_quarter;       // ln(0.25)
  double ln_2g;            // ln(2 * genome_size)
  double ln_30;            // ln(30)
  double sub_phred;
  double posterior_floor;  // exp(99 * DB)
  double genome_size;
};

// One banded DP job (SmithWatermanAligner::align on one window).
struct DpJob {
  uint64_t ref_gb;    // global base index of ref_seg[0] (genome source), or byte offset (explicit source)
  uint32_t read;      // read index in the batch
  uint32_t n;         // ref_seg length
  int32_t hint;       // ref_offset_hint
  int32_t band;       // band half-width
  uint8_t rc;         // query = reverse_complement(read), quals reversed
  uint8_t explicit_ref;
  uint16_t pad;
  uint32_t cand;      // candidate slot (read * 20 + k) or job id
};

struct DpOut {
  double log_lk;
  uint32_t ref_off;   // ref_start_offset
  int32_t score;
  uint32_t cigar_n;   // number of runs (may exceed the cap: overflow)
  uint32_t span;      // cigar_reference_span
  uint64_t cells;
};

struct Cand {
  uint64_t pos;       // candidate start (clamped >= 0)
  uint32_t contig;
  uint32_t strand_score;  // bit 31 strand, low bits = hit count
};

struct Counters {
  unsigned long long lookups, positions, filtered, cells, overflow;
  unsigned long long mapq_near;  // MAPQ finalizations within 1e-9 of a rounding point (finalize_mapq_from)
};

}  // namespace pbw

// ---- kernel launchers (pbw_kernels.hip) ----
namespace pbw {
struct SeedArgs {
  DevGenome G;
  DevIndex X;
  DevReads R;
  const uint32_t* rlist;   // reads to process (nullptr = 0..n-1)
  uint32_t nlist;
  uint64_t* hoff;          // per list entry: hit buffer offset (written per read in the chunked first pass)
  const uint32_t* hcap;    // per list entry: capacity (nullptr = hit_cap)
  uint32_t hit_cap;        // capacity of every read when hcap is null
  uint64_t* hits;
  uint32_t* nhits;         // per read (indexed by read id)
  uint8_t* rcode;          // per window repeat code, indexed by seq_off[read] + p
  Counters* ctr;
  uint8_t* hflag;          // optional, parallel to hits: bit 7 neighbour probe, bits 0..3 bucket-size class
  uint32_t has_long;       // some read is longer than MAX_READ_LEN (seeded from device memory by a second kernel)
  // first pass in chunks (chunk_tab set): read r's hit i at chunk_tab[r * SEED_MAXCH + i / 1024] * 1024 + i % 1024,
  // chunks taken from sub-pool r % SEED_SUBPOOLS (pool_chunks / SEED_SUBPOOLS each, counter pool_next[16 k]);
  // hoff[r] = chunk 0's first slot; dropped[r] = 1 when a chunk could not be had (re-seeded into an exact slot)
  uint32_t* chunk_tab;
  uint32_t* pool_next;
  uint32_t pool_chunks;
  uint8_t* dropped;
  uint32_t* mark_tab;      // exact-slot passes: marks the read's chunk row SEED_CONTIGUOUS (hits contiguous at hoff)
};
constexpr uint32_t SEED_CH_BITS = 10, SEED_MAXCH = 128;  // 1024-hit chunks, at most 131072 hits a read
constexpr uint32_t SEED_SUBPOOLS = 64;  // chunk counters: one atomic address per 64th of the reads
constexpr uint32_t SEED_CONTIGUOUS = 0xFFFFFFFFu;
// Read r's hit i in the hits array under a chunk table (or contiguous from hoff[r]).
__host__ __device__ inline uint64_t hit_at(const uint32_t* tab, const uint64_t* hoff, uint32_t r, uint32_t i) {
  if (!tab) return hoff[r] + i;
  const uint32_t* row = tab + (uint64_t)r * SEED_MAXCH;
  if (row[0] == SEED_CONTIGUOUS) return hoff[r] + i;
  return ((uint64_t)row[i >> SEED_CH_BITS] << SEED_CH_BITS) + (i & ((1u << SEED_CH_BITS) - 1));
}
void launch_seed(const SeedArgs& a, hipStream_t s);

struct ClusterArgs {
  DevReads R;
  const uint64_t* hits;
  const uint64_t* hoff;    // per read hit offset
  const uint32_t* chunk_tab;  // when set: the seed pass's chunk rows (hit_at); reads over 512 hits may be chunked
  const uint32_t* nhits;
  Cand* cands;             // read * 20
  uint8_t* ncand;
  uint32_t* mid;           // reads past the small wave tier (count at nmid, zero before launch)
  uint32_t* nmid;
  uint32_t* tier2;         // list of reads needing the block-level path
  uint32_t* ntier2;
  uint64_t* scratch;       // tier 2: per-block scratch (keys)
  uint64_t scratch_per_block;
  // prefilter features (optional): per hit flags from the seed pass, per
  // candidate slot (read * 20 + rank) 32 BF16 features
  const uint8_t* hflag;
  const uint8_t* rcode;
  uint16_t* feat;
};
void launch_cluster(const ClusterArgs& a, hipStream_t s);
// Block tier in three launches (k_cg_group, k_cg_sort, k_cg_select) for reads of
// at most CG_GMAX groups whose diagonals each span less than 2^32 - 1: the hits as
// composite keys in `ck` (group rank << 32 | biased diagonal - the group's lowest,
// in a region of the read's hit count), the groups in `cg` (CG_GMAX + 1 per list entry), the
// group count in cg_n (0: the read is appended to fb_list / fb_n for
// k_cluster_tier2, as is a read whose 20th cluster count the selection cannot
// resolve); `tmp` (as large as `ck`) is the merges' second buffer.  The hits
// stay intact.
constexpr uint32_t CG_GMAX = 56;
struct CgGroup {
  unsigned long long lo;  // lowest biased diagonal
  uint32_t key;           // contig * 2 + strand
  uint32_t start;         // segment start within the read's keys ([G]: the hit count)
  uint32_t rbits;         // bits of the highest diagonal offset (hi - lo)
  uint32_t pad;
};
struct CgArgs {
  uint64_t* ck;            // composite keys: each read's region taken from *ck_next by k_cg_group
  uint64_t* tmp;
  unsigned long long* ck_next;
  uint64_t* ck_off;        // per list entry: the read's region in ck / tmp
  CgGroup* cg;
  uint32_t* cg_n;
  uint32_t* fb_list;
  uint32_t* fb_n;
  uint64_t* bm;            // cluster-start bitmaps: each read's (n >> 6) + 1 words taken from *bm_next
  unsigned long long* bm_next;
  uint64_t* bm_off;        // per list entry: the read's first bitmap word
};
void launch_cluster_grid(const ClusterArgs& a, const CgArgs& g, const uint32_t* list, uint32_t n, hipStream_t s);
uint64_t cluster_scratch_words(uint32_t max_hits);  // per-block scratch of the block tier (u64 words)
void launch_cluster_tier2(const ClusterArgs& a, const uint32_t* list, uint32_t n, uint32_t blocks, hipStream_t s);

struct JobArgs {
  DevGenome G;
  DevReads R;
  const Cand* cands;
  const uint8_t* ncand;
  int32_t right_pad;
  int32_t band;
  DpJob* jobs;             // read * 20 + k (valid for k < njobs[read])
  uint8_t* njobs;
};
void launch_build_jobs(const JobArgs& a, hipStream_t s);

struct DpArgs {
  DevGenome G;
  DevReads R;
  DevLuts T;
  const DpJob* jobs;
  const uint32_t* jlist;   // job indices to run (nullptr = 0..njobs-1)
  uint32_t njobs;
  const uint32_t* njobs_dev;  // k_dp_wide / k_dp_tiled: when set, the job count (<= njobs, which sizes the launch)
  const uint8_t* job_valid;  // optional: per read count of valid slots (slot = read*20+k)
  const uint8_t* ref_bytes;  // explicit-ref source
  int32_t gap_open, gap_extend;
  uint32_t cigar_cap;
  uint32_t* cigar;         // job jid at (jid - cigar_base) * cigar_cap
  uint64_t cigar_base;     // first job id the CIGAR pool holds
  DpOut* out;
  uint32_t* flag_list;  // jobs this kernel does not handle, appended in any order (flag_n counts them)
  uint32_t* flag_n;
  Counters* ctr;
  uint32_t max_m, max_n;   // LDS sizing
  // tiled kernel scratch
  int32_t* rows;           // per wave: 3 * (max_n + 1)
  uint64_t* tb;            // per wave: (max_m + 1) * tiles * 4
  uint32_t tiles;          // ceil((max_n + 1) / 64)
  // k_dp_band / k_dp_tiled staging in device memory instead of LDS (for tiles
  // beyond LDS: long reads, windows of many thousand columns): wave w uses
  // gstage + w * gstage_stride; the launch runs at most gstage_waves waves
  uint8_t* gstage;
  uint64_t gstage_stride;
  uint32_t gstage_waves;
  uint32_t wpb;            // waves per block (banded kernel; set by the launcher)
  uint32_t* tbq;           // k_dp_quad traceback scratch: quad_scratch_words() words
  int64_t* tbend;          // k_dp_quad -> k_dp_walk: per job of the launch, its endpoint key (INT64_MIN: none)
  uint32_t quad_grid;      // k_dp_quad grid cap (0: QUAD_MAX_WAVES)
  uint32_t pack16;         // 1: eight jobs per wave in 16-bit halves (k_dp_pair16) where allowed
};
void launch_dp_band(const DpArgs& a, hipStream_t s);
// Bytes per wave of k_dp_band's staging in device memory when its tile for
// (max_m, max_n) does not fit LDS; 0 when it does (gstage unused).
uint64_t dp_band_gstage_bytes(uint32_t max_m, uint32_t max_n);
void launch_zero_u32(uint32_t* p, uint32_t n, hipStream_t s);  // a fill kernel (no memset node)
void launch_dp_tiled(const DpArgs& a, uint32_t nwaves, hipStream_t s);
// Bytes per wave of k_dp_tiled's staging and previous row in device memory
// when its staging for (max_m, max_n) does not fit LDS; 0 when it does.
uint64_t dp_tiled_gstage_bytes(uint32_t max_m, uint32_t max_n);
// Windows whose band covers every row (paired-end mate windows) up to 4096
// columns: k_dp_wide, which flags every other job to (flag_list, flag_n).
uint32_t dp_wide_tiles(uint32_t max_m, uint32_t max_n);  // 256-column tiles per job, 0: too wide
void launch_dp_wide(const DpArgs& a, uint32_t nwaves, hipStream_t s);
// true: the tiled kernel keeps its previous row in LDS (pass rows = nullptr)
bool dp_tiled_rows_in_lds(uint32_t max_m, uint32_t max_n);
// k_dp_quad (returns 1: its jobs then need launch_dp_walk with the same arguments)
// or k_dp_pair16 (2: done); 0: layout too large, nothing launched
int launch_dp_quad(const DpArgs& a, hipStream_t s);
void launch_dp_walk(const DpArgs& a, hipStream_t s);  // k_dp_walk, then k_dp_lk
constexpr uint32_t QUAD_MAX_WAVES = 32768;  // grid cap of k_dp_quad (more blocks than resident waves: dynamic balance)
size_t quad_scratch_words(uint32_t max_m, uint32_t max_n, uint32_t njobs, uint32_t grid_cap);
constexpr uint64_t QUAD_CHUNK_BYTES = 8ull << 30;  // traceback planes of one k_dp_quad launch, at most (two buffers)
// Jobs of `a` (jlist / njobs / job_valid) sorted longest read first into `out`, count
// to total[0]; `hist` holds MAX_READ_LEN + 1 words of scratch.
void launch_job_sort(const DpArgs& a, uint32_t* hist, uint32_t* out, uint32_t* total, hipStream_t s);

struct MapqArgs {
  DevGenome G;
  DevReads R;
  DevLuts T;
  const DpJob* jobs;
  const DpOut* out;
  const uint32_t* cigar;
  uint32_t cigar_cap;
  const uint8_t* njobs;    // per read
  const Cand* cands;
  const uint8_t* rcode;
  uint8_t* rec;            // pbw_record array (64 B each)
  uint32_t* best;          // per read best job slot or UINT32_MAX
  const uint32_t* keep;    // optional per read mask of the job slots to use (prefilter)
  const uint32_t* rlist;   // optional list of reads to finalize (nullptr = all)
  uint32_t nlist;
  Counters* ctr;           // optional: counts near-boundary finalizations
};
void launch_mapq(const MapqArgs& a, hipStream_t s);

// Prefilter selection, re-run verdict and merge (pbw_kernels.hip).
struct PfArgs {
  DevReads R;
  const DpJob* jobs;
  const DpOut* out;
  const uint8_t* njobs;
  const float* logit;        // per candidate slot (read * 20 + rank)
  float threshold;           // keep a candidate whose logit >= threshold
  uint32_t top_k;            // always keep ranks < top_k
  uint32_t* keep;            // per read: mask of kept job slots
  uint32_t* list;            // job slots to align
  uint32_t* nlist;
  unsigned long long* ctr;   // [0] jobs, [1] kept jobs, [2] cells skipped (reference count), [3] re-run jobs
  const uint8_t* rec;        // prefiltered records
  uint32_t mapq_min;         // re-run when the prefiltered MAPQ is below this
  float close_nats;          // .. or a kept second candidate is within this of the best
  uint32_t audit_period;     // re-run every audit_period-th read to measure concordance (0: never)
  uint64_t read_base;        // global index of the batch's first read
  uint8_t* verdict;          // per read: 0 as is, 1 fallback, 2 audit
  uint32_t* rlist;           // re-run reads
  uint32_t* nrlist;
  uint8_t* rec_out;          // records to patch (fallback) or compare (audit)
  const uint8_t* rec2;       // plain records of the re-run reads
  uint32_t* best;
  const uint32_t* best2;
  const uint32_t* cigar;
  uint32_t cigar_cap;
  unsigned long long* audit; // [0] audited reads, [1] audited reads whose SAM fields differ, [2] fallback reads
};
void launch_pf_select(const PfArgs& a, hipStream_t s);
void launch_pf_verdict(const PfArgs& a, hipStream_t s);
void launch_pf_merge(const PfArgs& a, uint32_t n, hipStream_t s);

struct PairArgs {
  DevGenome G;
  DevReads R;              // 2P reads: read 2p = mate 1, read 2p+1 = mate 2 of pair p
  DevLuts T;
  const DpJob* jobs;       // single-end candidate stage (slot = read * 20 + k)
  const DpOut* out;
  const uint32_t* cigar;
  uint32_t cigar_cap;
  const uint8_t* njobs;
  const Cand* cands;
  const uint8_t* rcode;
  DpJob* wjobs;            // cross-alignment jobs, slot = pair * 40 + s
  uint8_t* wvalid;
  uint32_t* wlist;         // slots with a window, appended by the plan (count in *wn)
  uint32_t* wn;
  uint8_t* slcand;         // shortlist candidate of each slot
  const DpOut* wout;
  const uint32_t* wcigar;  // slot s at (s - wbase) * wcap
  uint32_t wcap;
  uint64_t wbase;          // first slot the cross-alignment CIGAR pool holds
  uint32_t* plan;
  double mean, sd, ln_norm, sv_prior_log;  // frozen insert-size snapshot
  uint64_t mean_usize, radius_base;        // mean.max(0) as usize, ceil(5 sd) as usize
  const uint32_t* name_class;              // contig -> first contig with the same name
  uint8_t* rec;            // 2P records
  uint64_t* cig_src;       // per record CIGAR location (bit 63: cross-alignment pool)
  double* obs;             // per pair of the launch (index p - p_base)
  uint8_t* has_obs;
  uint32_t P;              // pairs handled by this launch: p_base .. p_base + P - 1 of the batch
  uint32_t p_base;
  Counters* ctr;           // optional: counts near-boundary finalizations
};
void launch_pair_plan(const PairArgs& a, hipStream_t s);
void launch_pair_finalize(const PairArgs& a, hipStream_t s);
// records of a batch to host memory (device-side CIGAR offsets and runs)
constexpr uint32_t REC_BLOCK = 1024;
void launch_rec_src_se(const uint32_t* best, uint32_t n, uint32_t cap, uint64_t* src, hipStream_t s);
void launch_rec_offsets(uint8_t* rec, const uint64_t* src, uint32_t n, uint32_t base, uint32_t* bsum,
                        uint32_t* total, hipStream_t s);
void launch_rec_cigar(const uint8_t* rec, const uint64_t* src, const uint32_t* cig0, const uint32_t* cig1,
                      const uint64_t* seq_off, uint32_t n, uint32_t base, uint32_t* out, hipStream_t s);
void gather_cigar2(const uint32_t* cig0, const uint32_t* cig1, const uint64_t* src, const uint32_t* off,
                   const uint32_t* cnt, uint32_t n, uint32_t* out, hipStream_t s);

void gather_cigar(const uint32_t* cig, const uint64_t* src, const uint32_t* off, const uint32_t* cnt, uint32_t n,
                  uint32_t* out, hipStream_t s);
void launch_fill_pres(const uint32_t* koff, uint64_t* pres, uint32_t rot, hipStream_t s);
// out[i] = i * stride for i < n
void launch_fill_stride(uint64_t* out, uint32_t n, uint64_t stride, hipStream_t s);
// Reads whose hits passed their first-pass slot (cap1), or with `dropped` the
// reads flagged there, listed in read order with exact slots from `base` on.
// part: seed_overflow_blocks(B) * 12 words of
// scratch; st (11 words): [0] listed reads, [1] most hits of a read, [2..5] reads
// over HIT_CAP << k hits (k = 0..3), [6..7] the listed reads' hits, [8..9] every
// read's hits (64-bit values low word first), [10] reads over HIT_CAP << 4.
uint32_t seed_overflow_blocks(uint32_t B);
void launch_seed_overflow(const uint32_t* nhits, const uint8_t* dropped, uint32_t B, uint32_t cap1, uint64_t base,
                          uint32_t* part, uint32_t* list, uint64_t* off, uint32_t* cap, uint32_t* st, hipStream_t s);
// Hits (and flags when hflag is set) of the listed reads copied contiguously:
// read r's region taken from *next, its offset to off[r].
void launch_hits_gather(const uint64_t* hits, const uint8_t* hflag, const uint64_t* hoff, const uint32_t* tab,
                        const uint32_t* nhits, const uint32_t* list, uint32_t nlist, uint64_t* dst, uint8_t* dflag,
                        uint64_t* off, unsigned long long* next, hipStream_t s);
// out[idx[k]] = val[k] for k < n
void launch_scatter_u64(uint64_t* out, const uint32_t* idx, const uint64_t* val, uint32_t n, hipStream_t s);
// Bucket table on the device (pbw_index.hip): fills koff[0 .. 2^30] and allocates
// *kpos (hipMalloc, owned by the caller) with *n_entries entries: fused ones when
// `fused` (*kfp null), else packed positions with their fingerprints in *kfp.
hipError_t gpu_build_index(const DevGenome& G, const uint64_t* h_clen, uint32_t* koff, uint64_t** kpos,
                           uint16_t** kfp, uint64_t* n_entries, bool fused, hipStream_t s, std::string& err);

// A persisted `.sthash` (HashTable, src/index/hashtable.rs:53-72) read into host
// arrays: per bucket its key (mix(code)), offset and count into `pos`.
struct HostTable {
  std::vector<uint32_t> key, off, cnt;
  std::vector<uint64_t> pos;
  uint64_t k = 0, stride = 0, total_entries = 0;
};
// Reads a bincode-1 `.sthash` file (pbw_io.cpp); false with `err` set on a
// malformed or unreadable file.
bool read_sthash(const char* path, HostTable& t, std::string& err);
bool write_sthash(const char* path, const HostTable& t, std::string& err);
// Sets the message pbw_global_error() returns on this thread (pbw_host.cpp).
void set_global_error(const std::string& e);
// The same bucket table built from a loaded HostTable instead of the genome:
// koff by code (code = the inverse of mix of each key), kpos in bucket order,
// fingerprints from the genome (fused or in kfp); every position is checked
// against the contigs.
hipError_t gpu_index_from_table(const DevGenome& G, const uint64_t* h_clen, const HostTable& t, uint32_t* koff,
                                uint64_t** kpos, uint16_t** kfp, uint64_t* n_entries, bool fused, hipStream_t s,
                                std::string& err);
// hashtable.rs:117-124 mix and its inverse (each step is a bijection of u32)
__host__ __device__ inline uint32_t kmer_mix(uint32_t h) {
  h ^= h >> 16;
  h *= 0x7feb352du;
  h ^= h >> 15;
  h *= 0x846ca68bu;
  h ^= h >> 16;
  return h;
}
__host__ __device__ inline uint32_t kmer_unmix(uint32_t h) {
  constexpr uint32_t INV_B = 0x43021123u;  // 0x846ca68b^-1 mod 2^32
  constexpr uint32_t INV_A = 0x1d69e2a5u;  // 0x7feb352d^-1 mod 2^32
  h ^= h >> 16;
  h *= INV_B;
  h ^= (h >> 15) ^ (h >> 30);
  h *= INV_A;
  h ^= h >> 16;
  return h;
}

__device__ __forceinline__ uint16_t float_to_bf16(float f) {  // round to nearest even
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

// ---- candidate prefilter network (pbw_mlp.hip) ----
constexpr int MLP_IN = 32, MLP_H = 128;
constexpr int MLP_OFF_W1 = 0;                          // W1 [f][k], 32 x 128
constexpr int MLP_OFF_B1 = MLP_OFF_W1 + MLP_IN * MLP_H;
constexpr int MLP_OFF_W2 = MLP_OFF_B1 + MLP_H;         // W2 [k][n], 128 x 128
constexpr int MLP_OFF_B2 = MLP_OFF_W2 + MLP_H * MLP_H;
constexpr int MLP_OFF_W3 = MLP_OFF_B2 + MLP_H;         // W3 [n]
constexpr int MLP_OFF_B3 = MLP_OFF_W3 + MLP_H;
constexpr int MLP_NP = MLP_OFF_B3 + 1;                 // 20865 parameters
struct MlpDev {
  float* theta;   // FP32 master parameters, layout above
  float* m;       // Adam moments
  float* v;
  uint16_t* w1t;  // BF16 W1 as [k][f]
  uint16_t* w2t;  // BF16 W2 as [n][k]
  uint16_t* w2;   // BF16 W2 as [k][n]
  uint16_t* w3;   // BF16 W3 [n]
  float* b1;
  float* b2;
  float* b3;
};
size_t mlp_lds_bytes();
void launch_mlp_forward(const MlpDev& P, const uint16_t* X, uint64_t N, float* z, hipStream_t s);
void launch_mlp_grad(const MlpDev& P, const uint16_t* X, const float* Y, const uint32_t* perm, uint64_t row0,
                     uint64_t nrows, uint64_t N, float pos_weight, float inv_b, float* part, float* loss_part,
                     hipStream_t s);
void launch_mlp_reduce_adam(const MlpDev& P, const float* part, uint32_t ntiles, float* grad_out, int adam, float lr,
                            float b1, float b2, float eps, float bc1, float bc2, hipStream_t s);
void launch_mlp_publish(const MlpDev& P, hipStream_t s);

}  // namespace pbw
