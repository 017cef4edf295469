// Host side of probebwa_amd: genome packing, index build, device context,
// batch driver and the C ABI declared in include/probebwa_amd.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <random>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/probebwa_amd.h"
#include "pbw_internal.h"

using namespace pbw;

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
static thread_local std::string g_error;
void pbw::set_global_error(const std::string& e) { g_error = e; }

struct PbwError {
  int code;
  std::string msg;
};

#define HIP_CHECK(x)                                                                                 \
  do {                                                                                               \
    hipError_t e_ = (x);                                                                             \
    if (e_ != hipSuccess)                                                                            \
      throw PbwError{PBW_ERR_HIP, std::string(#x " failed: ") + hipGetErrorString(e_)};              \
  } while (0)

template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  void ensure(size_t count) {
    if (count <= n && p) return;
    release();
    size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    HIP_CHECK(hipMalloc((void**)&p, bytes));
    n = std::max<size_t>(count, 1);
  }
  void grow_keep(size_t count, hipStream_t s) {  // as ensure, keeping the current contents
    if (count <= n && p) return;
    T* q = nullptr;
    HIP_CHECK(hipMalloc((void**)&q, std::max<size_t>(count, 1) * sizeof(T)));
    if (p && n) {
      HIP_CHECK(hipMemcpyAsync(q, p, n * sizeof(T), hipMemcpyDeviceToDevice, s));
      HIP_CHECK(hipStreamSynchronize(s));
    }
    release();
    p = q;
    n = std::max<size_t>(count, 1);
  }
  void upload(const T* h, size_t count, hipStream_t s) {
    ensure(count);
    if (count) HIP_CHECK(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
  }
};

// ---------------------------------------------------------------------------
// genome (genome.rs:17-45 packing, 2-bit + N mask; contigs aligned to 64 bases)
// ---------------------------------------------------------------------------
struct pbw_genome {
  std::vector<std::string> names;
  std::vector<uint64_t> lengths;
  std::vector<uint64_t> coff;
  std::vector<uint64_t> g2b;    // 32 bases per word
  std::vector<uint64_t> nmask;  // 64 bases per word
  uint64_t total_length = 0;
  uint64_t next_base = 0;       // first free (64-aligned) base slot
};

static void genome_add(pbw_genome* g, const std::string& name, const uint8_t* seq, uint64_t len) {
  uint64_t base = g->next_base;
  uint64_t end = base + len;
  uint64_t words2 = (end + 31) / 32 + 2, wordsn = (end + 63) / 64 + 2;
  if (g->g2b.size() < words2) g->g2b.resize(words2, 0);
  if (g->nmask.size() < wordsn) g->nmask.resize(wordsn, 0);
  for (uint64_t p = 0; p < len; p++) {
    uint64_t gb = base + p;
    uint64_t code = 0;
    bool acgt = true;
    switch (seq[p]) {
      case 'A': case 'a': code = 0; break;
      case 'C': case 'c': code = 1; break;
      case 'G': case 'g': code = 2; break;
      case 'T': case 't': code = 3; break;
      default: acgt = false; break;
    }
    g->g2b[gb >> 5] |= code << ((gb & 31) * 2);
    if (!acgt) g->nmask[gb >> 6] |= 1ull << (gb & 63);
  }
  g->names.push_back(name);
  g->lengths.push_back(len);
  g->coff.push_back(base);
  g->total_length += len;  // genome.rs:119
  g->next_base = (end + 63) / 64 * 64;
}

static inline uint32_t gbase2(const pbw_genome* g, uint64_t gb) { return (uint32_t)(g->g2b[gb >> 5] >> ((gb & 31) * 2)) & 3u; }
static inline bool gisn(const pbw_genome* g, uint64_t gb) { return (g->nmask[gb >> 6] >> (gb & 63)) & 1ull; }

// ---------------------------------------------------------------------------
// host copies of the reference's scalar math (for LUTs)
// ---------------------------------------------------------------------------
static const double H_LN_10 = 2.302585092994045684017991454684364208;
static const double H_DB = -0.1 * H_LN_10;
static double h_logadd(double a, double b) {
  if (a > b) std::swap(a, b);
  if (b >= a + 10.0) return a;
  if (b >= a + 4.0) return std::max(a - 1.0, 0.0);
  if (b >= a + 2.0) return std::max(a - 2.0, 0.0);
  return std::max(a - 3.0, 0.0);
}
static double h_py_round(double x) { return std::trunc(x + 0.5); }
static int32_t h_f64_i32(double x) {
  if (std::isnan(x)) return 0;
  if (x >= 2147483647.0) return INT32_MAX;
  if (x <= -2147483648.0) return INT32_MIN;
  return (int32_t)x;
}

// ---------------------------------------------------------------------------
// Fixed device region of pbw_align_batch_noalloc: every GPU the library runs on
// holds its own copy until the process ends.  Host code never uses the array's
// name as an address: the device address comes from hipGetSymbolAddress, once
// per device, when a context is created on it.
// ---------------------------------------------------------------------------
constexpr size_t NOALLOC_SCRATCH_BYTES = 64ull << 20;
constexpr size_t NOALLOC_STAGE_BYTES = 16ull << 20;
static __device__ __align__(256) uint8_t g_noalloc_scratch[NOALLOC_SCRATCH_BYTES];
static uint8_t* g_noalloc_ptr[256];  // per device
static std::mutex g_noalloc_mu;
alignas(256) static uint8_t g_noalloc_stage[NOALLOC_STAGE_BYTES];  // host staging of those calls

static void noalloc_setup(int device) {  // with `device` current
  if (device < 0 || device >= 256) throw PbwError{PBW_ERR_ARG, "device index out of range"};
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  if (g_noalloc_ptr[device]) return;
  void* ptr = nullptr;
  HIP_CHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(g_noalloc_scratch)));
  g_noalloc_ptr[device] = (uint8_t*)ptr;
}

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
struct pbw_ctx {
  const pbw_genome* genome;
  pbw_options opt;
  int device;
  hipStream_t stream = nullptr;
  hipEvent_t ev[8] = {};  // [0,1] seed, [2,3] DP, [4,5] wave-tier and [6,7] block-tier clustering
  hipStream_t stream2 = nullptr;  // the DP walks of one launch beside the next launch's rows
  hipEvent_t dev[4] = {};         // [0,1] rows of a launch done (by buffer), [2,3] walks done (by buffer)
  hipEvent_t evw[2] = {};         // the paired-end mate-window DP
  std::string error;
  pbw_stats stats{};
  // genome + index
  DevBuf<uint64_t> d_g2b, d_nmask, d_coff, d_clen;
  DevBuf<uint32_t> d_koff;
  DevBuf<uint64_t> d_pres;
  DevBuf<uint64_t> d_kpos;
  DevBuf<uint16_t> d_kfp;
  uint64_t total_entries = 0;
  bool fused = false;  // d_kpos holds fused entries (fingerprint inside, d_kfp empty)
  const HostTable* table_src = nullptr;  // set while a context is built from a loaded `.sthash`
  // LUTs
  DevBuf<double> d_lut;  // packed double tables
  DevBuf<int32_t> d_mis;
  DevLuts luts{};
  int32_t gap_open = -90, gap_extend = -1;
  // batch workspace
  DevBuf<uint8_t> d_seq, d_qual, d_rcode;
  DevBuf<uint64_t> d_seq_off, d_qual_off, d_hoff, d_hits, d_hits_ov, d_scratch, d_ovoff;
  uint32_t pool_chunks = 0;  // the seed pass's 1024-hit chunks, grown batch to batch to what the hits took
  DevBuf<uint32_t> d_chunk_tab, d_pool_next;
  DevBuf<uint8_t> d_dropped;
  DevBuf<uint64_t> d_hits_g, d_hoff_g;  // contiguous copies of reads' hits for k_cluster_tier2
  DevBuf<uint8_t> d_hflag_g;
  DevBuf<uint32_t> d_seed_st;  // the seed pass's overflow counts
  DevBuf<uint32_t> d_nhits, d_list, d_hcap, d_tier2, d_mid, d_ntier2, d_best, d_cigar, d_cigar_out, d_cigar_off;
  DevBuf<Cand> d_cands;
  DevBuf<uint8_t> d_ncand, d_njobs, d_rec, d_refbytes;
  DevBuf<uint32_t> d_flag_a, d_flag_b, d_flag_c, d_flag_n, d_sorted, d_sort_hist, d_tbq, d_flag_w, d_flag_wn;
  DevBuf<int64_t> d_tbend;  // k_dp_quad's endpoint keys for k_dp_walk
  pbw_tune tune{0, 0, 0, 0};  // batch_reads mirrors opt.batch_reads; zeros elsewhere: defaults
  DevBuf<DpJob> d_jobs;
  DevBuf<DpOut> d_out;
  DevBuf<Counters> d_ctr;
  DevBuf<int32_t> d_rows;
  DevBuf<uint8_t> d_gstage;  // DP staging in device memory for tiles beyond LDS
  DevBuf<uint64_t> d_ck, d_cktmp, d_ck_off;  // block-tier composite keys, the merges' second buffer, regions
  DevBuf<uint64_t> d_cbm, d_cbm_off;        // block-tier cluster-start bitmaps and each read's first word
  DevBuf<CgGroup> d_cg;
  DevBuf<uint32_t> d_cg_n, d_fb;
  DevBuf<uint64_t> d_tb;
  // paired-end workspace
  DevBuf<DpJob> d_wjobs;
  DevBuf<DpOut> d_wout;
  DevBuf<uint8_t> d_wvalid, d_slcand, d_pe_io;
  DevBuf<uint32_t> d_plan, d_wcigar, d_name_class, d_wlist;
  DevBuf<Counters> d_ctr_w;  // the cross-alignment DP's counters
  uint8_t* h_pe = nullptr;   // pinned: a pair batch's counts and observations
  size_t h_pe_cap = 0;
  uint8_t* h_up[2] = {nullptr, nullptr};  // pinned: bases and qualities of a pair batch on their way up
  size_t h_up_cap[2] = {0, 0};
  DevBuf<uint64_t> d_cig_src, d_rec_src;
  DevBuf<uint32_t> d_rec_bsum;
  // candidate prefilter network
  DevBuf<float> d_mlp_f;     // theta | m | v | b1 | b2 | b3
  DevBuf<uint16_t> d_mlp_h;  // w1t | w2t | w2 | w3
  DevBuf<float> d_mlp_part, d_mlp_loss, d_mlp_y, d_mlp_z, d_mlp_grad;
  DevBuf<uint16_t> d_mlp_x;
  DevBuf<uint32_t> d_mlp_perm;
  bool mlp_ready = false;
  uint64_t mlp_step = 0;
  // prefilter in the mapping path
  pbw_prefilter_options pf{};
  DevBuf<uint8_t> d_hflag, d_hflag_ov, d_verdict, d_rec2;
  DevBuf<uint16_t> d_feat;
  DevBuf<uint32_t> d_pf_keep, d_pf_list, d_pf_n, d_pf_rlist, d_best2;
  DevBuf<unsigned long long> d_pf_ctr;

  MlpDev mlp() {
    d_mlp_f.ensure(3 * (size_t)MLP_NP + 2 * MLP_H + 1);
    d_mlp_h.ensure((size_t)MLP_IN * MLP_H + 2 * (size_t)MLP_H * MLP_H + MLP_H);
    MlpDev P;
    P.theta = d_mlp_f.p;
    P.m = P.theta + MLP_NP;
    P.v = P.m + MLP_NP;
    P.b1 = P.v + MLP_NP;
    P.b2 = P.b1 + MLP_H;
    P.b3 = P.b2 + MLP_H;
    P.w1t = d_mlp_h.p;
    P.w2t = P.w1t + MLP_IN * MLP_H;
    P.w2 = P.w2t + MLP_H * MLP_H;
    P.w3 = P.w2 + MLP_H * MLP_H;
    return P;
  }

  DevGenome dev_genome() const {
    DevGenome G;
    G.g2b = d_g2b.p;
    G.nmask = d_nmask.p;
    G.coff = d_coff.p;
    G.clen = d_clen.p;
    G.ncontigs = (int32_t)genome->lengths.size();
    return G;
  }
};

static void ctx_build(pbw_ctx* c) {
  const pbw_genome* g = c->genome;
  HIP_CHECK(hipSetDevice(c->device));
  noalloc_setup(c->device);
  HIP_CHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  for (auto& e : c->ev) HIP_CHECK(hipEventCreate(&e));
  HIP_CHECK(hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking));
  for (auto& e : c->dev) HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  for (auto& e : c->evw) HIP_CHECK(hipEventCreate(&e));
  hipStream_t s = c->stream;
  if (g->lengths.size() >= (1u << 22)) throw PbwError{PBW_ERR_LIMIT, "more than 2^22 contigs"};
  for (uint64_t L : g->lengths)
    if (L >= (1ull << 40)) throw PbwError{PBW_ERR_LIMIT, "contig longer than 2^40 bases"};
  std::vector<uint64_t> g2b = g->g2b, nm = g->nmask;
  g2b.resize(std::max<size_t>(g2b.size(), 2) + 2, 0);
  nm.resize(std::max<size_t>(nm.size(), 2) + 2, 0);
  c->d_g2b.upload(g2b.data(), g2b.size(), s);
  c->d_nmask.upload(nm.data(), nm.size(), s);
  c->d_coff.upload(g->coff.data(), g->coff.size(), s);
  c->d_clen.upload(g->lengths.data(), g->lengths.size(), s);
  {  // contig -> first contig with the same name (paired.rs:188 compares names)
    std::vector<uint32_t> nc(g->names.size());
    std::unordered_map<std::string, uint32_t> first;
    first.reserve(g->names.size());
    for (size_t i = 0; i < g->names.size(); i++) nc[i] = first.emplace(g->names[i], (uint32_t)i).first->second;
    c->d_name_class.upload(nc.data(), nc.size(), s);
    HIP_CHECK(hipStreamSynchronize(s));
  }
  // index: HashTableBuilder::build (hashtable.rs:248-322) on the device
  c->d_koff.ensure(CODE_SPACE + 1);
  {
    uint64_t* kpos = nullptr;
    uint16_t* kfp = nullptr;
    uint64_t nent = 0;
    std::string err;
    // fused entries (fingerprint inside) when every contig id and position fits them
    c->fused = g->lengths.size() <= FUSED_MAX_CONTIGS;
    for (uint64_t len : g->lengths) c->fused = c->fused && len <= FUSED_MAX_LEN;
    if (c->table_src) {  // HashTable::load (hashtable.rs:66-72)
      if (gpu_index_from_table(c->dev_genome(), g->lengths.data(), *c->table_src, c->d_koff.p, &kpos, &kfp, &nent,
                               c->fused, s, err) != hipSuccess)
        throw PbwError{PBW_ERR_IO, "table: " + err};
    } else if (gpu_build_index(c->dev_genome(), g->lengths.data(), c->d_koff.p, &kpos, &kfp, &nent, c->fused, s,
                               err) != hipSuccess) {
      throw PbwError{err.find("2^32") != std::string::npos ? PBW_ERR_LIMIT : PBW_ERR_HIP, err};
    }
    c->d_kpos.release();
    c->d_kpos.p = kpos;
    c->d_kpos.n = std::max<uint64_t>(nent, 1);
    c->d_kfp.release();
    c->d_kfp.p = kfp;
    c->d_kfp.n = kfp ? std::max<uint64_t>(nent, 1) : 0;
    c->total_entries = nent;
  }
  c->d_pres.ensure(3 * CODE_SPACE / 64);  // natural, rotated by 10, rotated by 20
  launch_fill_pres(c->d_koff.p, c->d_pres.p, 0, s);
  launch_fill_pres(c->d_koff.p, c->d_pres.p + CODE_SPACE / 64, 10, s);
  launch_fill_pres(c->d_koff.p, c->d_pres.p + 2 * CODE_SPACE / 64, 20, s);
  HIP_CHECK(hipGetLastError());
  // LUTs (bayesian.rs:103-110, :142-179, :405-439; seeds.rs:133-134; smith_waterman.rs:272-275)
  const pbw_options& o = c->opt;
  double sub_phred = h_py_round(std::log(std::max(o.substitution_rate, 1e-30)) / H_DB);
  std::vector<double> lut;
  auto region = [&](size_t n) {
    size_t off = lut.size();
    lut.resize(off + n, 0.0);
    return off;
  };
  size_t o_match = region(256), o_mis = region(256), o_omq = region(256), o_qr = region(256), o_nrp = region(256);
  size_t o_ln = region(LN_INT_MAX + 1);
  size_t ngc = (size_t)(GC_LUT_MAX + 1) * (GC_LUT_MAX + 2) / 2;
  size_t o_lnb = region(ngc), o_lnbm1 = region(ngc);
  std::vector<int32_t> mis(256);
  for (int q = 0; q < 256; q++) {
    double eff = h_logadd((double)q, sub_phred);
    double e = std::pow(10.0, -eff / 10.0);
    double pm = 1.0 - e;
    lut[o_match + q] = std::log(std::max(pm, 1e-300));
    lut[o_mis + q] = std::log(std::max(e / 3.0, 1e-300));
    double qv = std::min(std::pow(10.0, -eff / 10.0), 1.0 - 1e-12);
    lut[o_omq + q] = 1.0 - qv;
    lut[o_qr + q] = qv / (1.0 - qv);
    lut[o_nrp + q] = std::pow(10.0, -(double)q / 10.0);
    double scale = std::min(std::max((double)q / 30.0, 0.2), 1.0);
    mis[q] = h_f64_i32(std::round(-30.0 * scale));
  }
  for (int k = 0; k <= LN_INT_MAX; k++) lut[o_ln + k] = std::log((double)k);
  for (uint32_t t = 0; t <= (uint32_t)GC_LUT_MAX; t++) {
    for (uint32_t gc = 0; gc <= t; gc++) {
      double f = t == 0 ? 0.5 : (double)gc / (double)t;
      double ff = std::min(std::max(f, 1e-6), 1.0 - 1e-6);
      double b = std::exp(-ff * std::log(ff / 2.0) - (1.0 - ff) * std::log((1.0 - ff) / 2.0));
      size_t idx = (size_t)t * (t + 1) / 2 + gc;
      lut[o_lnb + idx] = std::log(b);
      lut[o_lnbm1 + idx] = std::log(std::max(b - 1.0, 1e-6));
    }
  }
  c->d_lut.upload(lut.data(), lut.size(), s);
  c->d_mis.upload(mis.data(), mis.size(), s);
  DevLuts& T = c->luts;
  T.lk_match = c->d_lut.p + o_match;
  T.lk_mis = c->d_lut.p + o_mis;
  T.omq = c->d_lut.p + o_omq;
  T.qratio = c->d_lut.p + o_qr;
  T.nrp = c->d_lut.p + o_nrp;
  T.mis_pen = c->d_mis.p;
  T.ln_int = c->d_lut.p + o_ln;
  T.gc_lnb = c->d_lut.p + o_lnb;
  T.gc_lnbm1 = c->d_lut.p + o_lnbm1;
  T.ln_quarter = std::log(0.25);
  T.genome_size = (double)std::max<uint64_t>(g->total_length, 1);
  T.ln_2g = std::log(2.0 * T.genome_size);
  T.ln_30 = std::log(30.0);
  T.sub_phred = sub_phred;
  T.posterior_floor = std::exp(99.0 * H_DB);
  // gap costs: smith_waterman.rs:113-130
  c->gap_open = std::min(h_f64_i32(std::round(-90.0 * (double)o.gap_open_phred / 40.0)), -1);
  c->gap_extend = std::min(h_f64_i32(std::round(-1.0 * (double)o.gap_extend_phred / 3.0)), -1);
  c->d_ctr.ensure(1);
  HIP_CHECK(hipStreamSynchronize(s));
}

// ---------------------------------------------------------------------------
// single-end batch: seed -> cluster -> jobs -> DP -> MAPQ
// ---------------------------------------------------------------------------
// A pinned host array that an output batch takes over (pbw_records_release
// frees it with hipHostFree).  Device-to-host copies into pinned memory run at
// the link's rate; into fresh pageable pages they run several times slower.
template <typename T>
struct HostArr {
  T* p = nullptr;
  size_t n = 0, cap = 0;
  HostArr() = default;
  HostArr(const HostArr&) = delete;
  HostArr& operator=(const HostArr&) = delete;
  ~HostArr() {
    if (p) (void)hipHostFree(p);
  }
  void reserve(size_t want) {
    if (want <= cap) return;
    T* q = nullptr;
    if (hipHostMalloc((void**)&q, want * sizeof(T), hipHostMallocDefault) != hipSuccess || !q)
      throw PbwError{PBW_ERR_LIMIT, "pinned host allocation of " + std::to_string(want * sizeof(T)) + " bytes failed"};
    if (n) memcpy(q, p, n * sizeof(T));
    if (p) (void)hipHostFree(p);
    p = q;
    cap = want;
  }
  T* grow(size_t add) {  // appends `add` uninitialised elements
    if (n + add > cap) reserve(std::max(n + add, cap * 2));
    T* at = p + n;
    n += add;
    return at;
  }
  void push_back(const T& v) { *grow(1) = v; }
  size_t size() const { return n; }
  bool empty() const { return n == 0; }
  T& operator[](size_t i) { return p[i]; }
  T* release() {
    reserve(1);  // never hand out a null array
    T* q = p;
    p = nullptr;
    n = cap = 0;
    return q;
  }
};

struct HostRecords {
  HostArr<pbw_record> rec;
  HostArr<uint32_t> cigar;
  size_t expect = 0;  // records the call will produce (sizes the first reservation)
};

static double ms_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

static DpArgs dp_args_base(pbw_ctx* c, const DevReads& R) {
  DpArgs a{};
  a.G = c->dev_genome();
  a.R = R;
  a.T = c->luts;
  a.gap_open = c->gap_open;
  a.gap_extend = c->gap_extend;
  a.ctr = c->d_ctr.p;
  return a;
}

// Runs the tiled DP kernel over a job list (wide windows, odd shapes).
// One device word to the host (synchronises the stream).
static uint32_t download_u32(pbw_ctx* c, const uint32_t* d) {
  uint32_t v = 0;
  HIP_CHECK(hipMemcpyAsync(&v, d, 4, hipMemcpyDeviceToHost, c->stream));
  HIP_CHECK(hipStreamSynchronize(c->stream));
  return v;
}

// Device-memory staging of `bytes` per wave for up to `want` waves (within
// 2 GiB) when a DP launch's tile does not fit LDS; bytes = 0 clears it.
static void dp_gstage(pbw_ctx* c, DpArgs& a, uint64_t bytes, uint32_t want) {
  a.gstage = nullptr;
  a.gstage_stride = 0;
  a.gstage_waves = 0;
  if (!bytes) return;
  const uint64_t fit = std::max<uint64_t>(1, (2ull << 30) / bytes);
  const uint32_t w = (uint32_t)std::min<uint64_t>(std::max<uint32_t>(want, 1), fit);
  c->d_gstage.ensure((size_t)w * bytes);
  a.gstage = c->d_gstage.p;
  a.gstage_stride = bytes;
  a.gstage_waves = w;
}

// k_dp_band over a's job list, its staging in device memory when the tile for
// (max_m, max_n) does not fit LDS (long reads).
static void run_dp_band(pbw_ctx* c, DpArgs a) {
  if (!a.njobs) return;
  dp_gstage(c, a, dp_band_gstage_bytes(a.max_m, a.max_n), std::min<uint32_t>(a.njobs, 16384));
  launch_dp_band(a, c->stream);
}

// Scratch of a k_dp_tiled launch of up to `want` waves: the previous row in
// LDS when it fits (windows up to ~2000 columns), else per wave in device
// memory; staging in device memory when even that does not fit LDS (windows
// of many thousand columns, long reads).  Waves are limited so that the
// traceback bits stay within 4 GiB.  Returns the waves to launch.
static uint32_t tiled_scratch(pbw_ctx* c, DpArgs& a, uint32_t want) {
  a.tiles = (a.max_n + 1 + 63) / 64;
  const uint64_t gs = dp_tiled_gstage_bytes(a.max_m, a.max_n);
  const bool lds_rows = !gs && dp_tiled_rows_in_lds(a.max_m, a.max_n);
  const uint64_t tb_words = (uint64_t)(a.max_m + 1) * a.tiles * 4;
  const uint64_t row_words = (gs || lds_rows) ? 0 : 3ull * (a.max_n + 1);
  const uint64_t per = tb_words * 8 + row_words * 4 + gs;
  const uint32_t nw =
      (uint32_t)std::min<uint64_t>(std::max<uint32_t>(want, 1), std::max<uint64_t>(1, (4ull << 30) / per));
  dp_gstage(c, a, gs, nw);
  if (row_words) {
    c->d_rows.ensure((size_t)nw * row_words);
    a.rows = c->d_rows.p;
  } else {
    a.rows = nullptr;
  }
  c->d_tb.ensure((size_t)nw * tb_words);
  a.tb = c->d_tb.p;
  return nw;
}

// Jobs of a device list through the tiled kernel.  It flags only jobs beyond the
// batch's max_m / max_n, which the callers never pass; any flag is an error.
static void run_dp_tiled_dev(pbw_ctx* c, DpArgs a, const uint32_t* dlist, uint32_t n) {
  if (!n) return;
  hipStream_t s = c->stream;
  a.jlist = dlist;
  a.njobs = n;
  a.job_valid = nullptr;
  c->d_flag_n.ensure(4);
  c->d_flag_c.ensure(n);
  a.flag_list = c->d_flag_c.p;
  a.flag_n = c->d_flag_n.p + 2;
  HIP_CHECK(hipMemsetAsync(a.flag_n, 0, 4, s));
  const bool lds_rows = !dp_tiled_gstage_bytes(a.max_m, a.max_n) && dp_tiled_rows_in_lds(a.max_m, a.max_n);
  const uint32_t nwaves = tiled_scratch(c, a, std::min<uint32_t>(n, lds_rows ? 8192 : 2048));
  launch_dp_tiled(a, nwaves, s);
  HIP_CHECK(hipGetLastError());
  if (download_u32(c, a.flag_n)) throw PbwError{PBW_ERR_LIMIT, "alignment window beyond the tiled kernel's shape"};
}
static void run_dp_tiled(pbw_ctx* c, DpArgs a, const std::vector<uint32_t>& list) {
  if (list.empty()) return;
  hipStream_t s = c->stream;
  const uint32_t n = (uint32_t)list.size();
  c->d_list.upload(list.data(), n, s);
  const uint32_t T = dp_wide_tiles(a.max_m, a.max_n);
  if (!T) {
    run_dp_tiled_dev(c, a, c->d_list.p, n);
    return;
  }
  // whole-window rectangles in k_dp_wide; the jobs it flags go to k_dp_tiled
  DpArgs w = a;
  w.jlist = c->d_list.p;
  w.njobs = n;
  w.job_valid = nullptr;
  c->d_flag_w.ensure(n);
  c->d_flag_wn.ensure(1);
  HIP_CHECK(hipMemsetAsync(c->d_flag_wn.p, 0, 4, s));
  w.flag_list = c->d_flag_w.p;
  w.flag_n = c->d_flag_wn.p;
  const uint32_t nwaves = std::min<uint32_t>(n, 8192);
  c->d_tb.ensure((size_t)nwaves * (a.max_m + 1) * T * 16);  // 16 bits per lane, tile and row
  w.tb = c->d_tb.p;
  launch_dp_wide(w, nwaves, s);
  HIP_CHECK(hipGetLastError());
  const uint32_t nflag = download_u32(c, c->d_flag_wn.p);
  if (nflag) run_dp_tiled_dev(c, a, c->d_flag_w.p, nflag);
}

// Device state of one batch between the stages.
struct Batch {
  uint32_t B = 0;
  uint32_t max_m = 0, max_n = 0, cap = 0;
  std::vector<uint64_t> so, qo;  // rebased offsets
  DevReads R{};
};

// Pairs [p0, p1) of two read batches, interleaved (read 2i = mate 1 and read
// 2i + 1 = mate 2 of pair p0 + i): offsets, longest read, and bases then
// qualities packed into the context's pinned staging slot `slot`.  Touches no
// device state, so it may run on another thread while the stream works.
struct PairPack {
  std::vector<uint64_t> so, qo;
  uint32_t B = 0, max_m = 0;
  uint8_t* buf = nullptr;
};
static void pair_pack(pbw_ctx* c, int slot, const pbw_read_batch* m1, const pbw_read_batch* m2, uint64_t p0,
                      uint64_t p1, PairPack& pk) {
  const uint32_t P = (uint32_t)(p1 - p0), B = 2 * P;
  const pbw_read_batch* mm[2] = {m1, m2};
  pk.B = B;
  pk.so.assign(B + 1, 0);
  pk.qo.assign(B + 1, 0);
  pk.max_m = 0;
  for (uint32_t i = 0; i < B; i++) {
    const pbw_read_batch* x = mm[i & 1];
    const uint64_t r = p0 + i / 2;
    const uint64_t sl = x->seq_off[r + 1] - x->seq_off[r], ql = x->qual_off[r + 1] - x->qual_off[r];
    if (sl > 0x7FFFFFFFull || ql > 0x7FFFFFFFull) throw PbwError{PBW_ERR_LIMIT, "read longer than 2^31 bases"};
    pk.so[i + 1] = pk.so[i] + sl;
    pk.qo[i + 1] = pk.qo[i] + ql;
    pk.max_m = std::max<uint32_t>(pk.max_m, (uint32_t)std::max(sl, ql));
  }
  const size_t ns = pk.so[B] + 1, nq = pk.qo[B] + 1;
  if (c->h_up_cap[slot] < ns + nq) {
    const size_t want = std::max(ns + nq, (c->h_up_cap[slot] * 3) / 2);
    if (c->h_up[slot]) (void)hipHostFree(c->h_up[slot]);
    c->h_up[slot] = nullptr;
    c->h_up_cap[slot] = 0;
    if (hipHostMalloc((void**)&c->h_up[slot], want, hipHostMallocDefault) != hipSuccess || !c->h_up[slot])
      throw PbwError{PBW_ERR_LIMIT, "pinned host allocation of " + std::to_string(want) + " bytes failed"};
    c->h_up_cap[slot] = want;
  }
  pk.buf = c->h_up[slot];
  uint8_t* seq = pk.buf;
  uint8_t* qual = pk.buf + ns;
  for (uint32_t i = 0; i < B; i++) {
    const pbw_read_batch* x = mm[i & 1];
    const uint64_t r = p0 + i / 2;
    const uint64_t sl = pk.so[i + 1] - pk.so[i], ql = pk.qo[i + 1] - pk.qo[i];
    if (sl) memcpy(seq + pk.so[i], x->seq + x->seq_off[r], sl);
    if (ql) memcpy(qual + pk.qo[i], x->qual + x->qual_off[r], ql);
  }
  seq[ns - 1] = 0;
  qual[nq - 1] = 0;
}

// A packed pair batch to the device (the staging slot is free again on return).
static void batch_send(pbw_ctx* c, Batch& b, const PairPack& pk) {
  hipStream_t s = c->stream;
  const uint32_t B = pk.B;
  b.B = B;
  b.so = pk.so;
  b.qo = pk.qo;
  b.max_m = pk.max_m;
  const size_t ns = pk.so[B] + 1, nq = pk.qo[B] + 1;
  c->d_seq.upload(pk.buf, ns, s);
  c->d_qual.upload(pk.buf + ns, nq, s);
  c->d_seq_off.upload(b.so.data(), B + 1, s);
  c->d_qual_off.upload(b.qo.data(), B + 1, s);
  b.R = DevReads{c->d_seq.p, c->d_seq_off.p, c->d_qual.p, c->d_qual_off.p, B};
  c->d_rcode.ensure(b.so[B] + 1);
  HIP_CHECK(hipStreamSynchronize(s));
}

// Uploads a contiguous slice [b0, b1) of a read batch.
static void batch_upload_slice(pbw_ctx* c, Batch& b, const pbw_read_batch* in, uint64_t b0, uint64_t b1,
                               bool resident) {
  hipStream_t s = c->stream;
  const uint32_t B = (uint32_t)(b1 - b0);
  b.B = B;
  b.so.resize(B + 1);
  b.qo.resize(B + 1);
  b.max_m = 0;
  for (uint32_t i = 0; i <= B; i++) {
    b.so[i] = in->seq_off[b0 + i] - in->seq_off[b0];
    b.qo[i] = in->qual_off[b0 + i] - in->qual_off[b0];
  }
  for (uint32_t i = 0; i < B; i++) {
    uint64_t L = b.so[i + 1] - b.so[i], ql = b.qo[i + 1] - b.qo[i];
    if (L > 0x7FFFFFFFull || ql > 0x7FFFFFFFull) throw PbwError{PBW_ERR_LIMIT, "read longer than 2^31 bases"};
    b.max_m = std::max<uint32_t>(b.max_m, (uint32_t)std::max(L, ql));
  }
  c->d_seq_off.upload(b.so.data(), B + 1, s);
  c->d_qual_off.upload(b.qo.data(), B + 1, s);
  if (resident) {  // bases and quals already in device memory
    b.R = DevReads{in->seq + in->seq_off[b0], c->d_seq_off.p, in->qual + in->qual_off[b0], c->d_qual_off.p, B};
  } else {
    c->d_seq.upload(in->seq + in->seq_off[b0], b.so[B], s);
    c->d_qual.upload(in->qual + in->qual_off[b0], b.qo[B], s);
    b.R = DevReads{c->d_seq.p, c->d_seq_off.p, c->d_qual.p, c->d_qual_off.p, B};
  }
  c->d_rcode.ensure(b.so[B] + 1);
  HIP_CHECK(hipStreamSynchronize(s));
}

// map_candidates (single.rs:86-162) for every read of the batch: seed,
// cluster, DP jobs, banded DP.  Results stay on the device.
// `pre` (optional): output arrays to reserve on the host while the seed kernel runs.
static void batch_candidates(pbw_ctx* c, Batch& b, bool feat = false, HostRecords* pre = nullptr) {
  hipStream_t s = c->stream;
  const uint32_t B = b.B;
  const DevReads R = b.R;
  auto t0 = std::chrono::steady_clock::now();
  // ---- seed ----
  // hits in 1024-hit chunks taken from a pool as they come (a read's chunk row in
  // d_chunk_tab), then slack for the exact slots of the reads a full pool left out
  const uint32_t pool0 = c->tune.seed_pool ? c->tune.seed_pool : (c->pool_chunks ? c->pool_chunks : 2 * B);
  const uint32_t pool = (pool0 + SEED_SUBPOOLS - 1) / SEED_SUBPOOLS * SEED_SUBPOOLS;  // whole sub-pools
  c->d_hoff.ensure(B);
  const size_t base_cap = (size_t)pool << SEED_CH_BITS, slack = (size_t)B * 64;
  c->d_hits.ensure(base_cap + slack);
  c->d_nhits.ensure(B);
  c->d_chunk_tab.ensure((size_t)B * SEED_MAXCH);
  c->d_dropped.ensure(B);
  c->d_pool_next.ensure(16 * SEED_SUBPOOLS);
  HIP_CHECK(hipMemsetAsync(c->d_pool_next.p, 0, 64 * SEED_SUBPOOLS, s));
  SeedArgs sa{};
  sa.G = c->dev_genome();
  sa.X = DevIndex{c->d_koff.p, c->d_kpos.p, c->d_pres.p, c->d_pres.p + CODE_SPACE / 64,
                  c->d_pres.p + 2 * CODE_SPACE / 64, c->d_kfp.p};
  sa.R = R;
  sa.rlist = nullptr;
  sa.nlist = B;
  sa.hoff = c->d_hoff.p;
  sa.hcap = nullptr;
  sa.hit_cap = 0;
  sa.chunk_tab = c->d_chunk_tab.p;
  sa.pool_next = c->d_pool_next.p;
  sa.pool_chunks = pool;
  sa.dropped = c->d_dropped.p;
  sa.hits = c->d_hits.p;
  sa.nhits = c->d_nhits.p;
  sa.rcode = c->d_rcode.p;
  sa.ctr = c->d_ctr.p;
  sa.has_long = b.max_m > (uint32_t)MAX_READ_LEN;
  if (feat) {
    c->d_hflag.ensure(base_cap + slack);
    sa.hflag = c->d_hflag.p;
  }
  HIP_CHECK(hipEventRecord(c->ev[0], s));
  launch_seed(sa, s);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipEventRecord(c->ev[1], s));
  if (pre) {  // pinned allocations overlap the seed kernel: one record and ~4 CIGAR runs per read
    const size_t wr = std::max(pre->rec.n + B, pre->expect);
    if (wr > pre->rec.cap) pre->rec.reserve(std::max(wr, pre->rec.cap * 2));
    const size_t wc = std::max(pre->cigar.n + (size_t)B * 4, pre->expect * 4);
    if (wc > pre->cigar.cap) pre->cigar.reserve(std::max(wc, pre->cigar.cap * 2));
  }
  // overflow reads listed on the device; their counts to the host
  c->d_list.ensure(B);
  c->d_ovoff.ensure(B);
  c->d_hcap.ensure(B);
  c->d_seed_st.ensure(16 + 12 * (size_t)seed_overflow_blocks(B));
  launch_seed_overflow(c->d_nhits.p, c->d_dropped.p, B, 0, (uint64_t)base_cap, c->d_seed_st.p + 16, c->d_list.p,
                       c->d_ovoff.p, c->d_hcap.p, c->d_seed_st.p, s);
  HIP_CHECK(hipGetLastError());
  uint32_t st[11];
  HIP_CHECK(hipMemcpyAsync(st, c->d_seed_st.p, sizeof(st), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  {
    float ms = 0.f;
    HIP_CHECK(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
    c->stats.ms_k_seed += ms;
    c->stats.n_k_seed += 1;
  }
  const uint32_t n_ov = st[0], max_hits = st[1];
  const uint64_t ov_total = (uint64_t)st[6] | ((uint64_t)st[7] << 32);
  const uint64_t all_hits = (uint64_t)st[8] | ((uint64_t)st[9] << 32);
  {
    // the next batches' pool: the chunks this batch's hits needed (at most one
    // partial chunk per read) and an eighth more, within 64 GiB of hits and flags
    const uint64_t need = (all_hits >> SEED_CH_BITS) + B;
    if (need > pool && !c->tune.seed_pool) {
      const uint64_t want = std::min<uint64_t>(need + need / 8, (64ull << 30) / (9ull << SEED_CH_BITS));
      c->pool_chunks = (uint32_t)std::max<uint64_t>(pool, want);
    }
  }
  const uint64_t* hits_base = c->d_hits.p;
  const uint8_t* hflag_base = feat ? c->d_hflag.p : nullptr;
  if (n_ov) {
    // overflow slots after the fixed ones: in place when the slack holds them,
    // else in a second arena that starts with a copy of the fixed slots
    uint64_t* arena = c->d_hits.p;
    uint8_t* farena = feat ? c->d_hflag.p : nullptr;
    if (ov_total > slack) {
      c->d_hits_ov.ensure(base_cap + ov_total + (c->d_hits_ov.n < base_cap + ov_total ? ov_total / 4 : 0));
      HIP_CHECK(hipMemcpyAsync(c->d_hits_ov.p, c->d_hits.p, base_cap * 8, hipMemcpyDeviceToDevice, s));
      arena = c->d_hits_ov.p;
      if (feat) {
        c->d_hflag_ov.ensure(base_cap + ov_total + (c->d_hflag_ov.n < base_cap + ov_total ? ov_total / 4 : 0));
        HIP_CHECK(hipMemcpyAsync(c->d_hflag_ov.p, c->d_hflag.p, base_cap, hipMemcpyDeviceToDevice, s));
        farena = c->d_hflag_ov.p;
      }
    }
    SeedArgs sb = sa;
    sb.rlist = c->d_list.p;
    sb.nlist = n_ov;
    sb.hoff = c->d_ovoff.p;
    sb.hcap = c->d_hcap.p;
    sb.chunk_tab = nullptr;  // exact slots; the reads' chunk rows marked contiguous
    sb.dropped = nullptr;
    sb.mark_tab = c->d_chunk_tab.p;
    sb.hits = arena;
    sb.ctr = nullptr;  // counted once, in the first pass
    sb.hflag = farena;
    launch_seed(sb, s);
    HIP_CHECK(hipGetLastError());
    // the overflow reads' slots replace their fixed ones in the per-read offsets
    launch_scatter_u64(c->d_hoff.p, c->d_list.p, c->d_ovoff.p, n_ov, s);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipStreamSynchronize(s));
    hits_base = arena;
    hflag_base = farena;
    c->stats.seed_overflow_reads += n_ov;
  }
  c->stats.ms_seed += ms_since(t0);
  t0 = std::chrono::steady_clock::now();
  // ---- cluster ----
  c->d_cands.ensure((size_t)B * TOP_CANDIDATES);
  c->d_ncand.ensure(B);
  c->d_tier2.ensure(B);
  c->d_mid.ensure(B);
  c->d_ntier2.ensure(4);  // [0] block tier, [1] second wave tier, [2] block tier handed to k_cluster_tier2
  HIP_CHECK(hipMemsetAsync(c->d_ntier2.p, 0, 16, s));
  ClusterArgs ca{};
  ca.R = R;
  ca.hits = hits_base;
  ca.hoff = c->d_hoff.p;
  ca.chunk_tab = c->d_chunk_tab.p;
  ca.nhits = c->d_nhits.p;
  ca.cands = c->d_cands.p;
  ca.ncand = c->d_ncand.p;
  ca.tier2 = c->d_tier2.p;
  ca.ntier2 = c->d_ntier2.p;
  ca.mid = c->d_mid.p;
  ca.nmid = c->d_ntier2.p + 1;
  if (feat) {
    c->d_feat.ensure((size_t)B * TOP_CANDIDATES * MLP_IN);
    ca.hflag = hflag_base;
    ca.rcode = c->d_rcode.p;
    ca.feat = c->d_feat.p;
  }
  HIP_CHECK(hipEventRecord(c->ev[4], s));
  launch_cluster(ca, s);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipEventRecord(c->ev[5], s));
  uint32_t nt2 = 0;
  HIP_CHECK(hipMemcpyAsync(&nt2, c->d_ntier2.p, 4, hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  float ms_cl = 0.f;
  HIP_CHECK(hipEventElapsedTime(&ms_cl, c->ev[4], c->ev[5]));
  if (nt2) {
    const uint32_t* t2list = c->d_tier2.p;
    uint32_t nt2b = nt2;  // reads for k_cluster_tier2
    HIP_CHECK(hipEventRecord(c->ev[6], s));
    {
      // three launches (group / sort / select) for the reads they take (features
      // too); the rest, listed on the device, to k_cluster_tier2 below
      if (c->d_ck.n < all_hits) {  // with headroom: the count varies from batch to batch
        c->d_ck.ensure(all_hits + all_hits / 4);
        c->d_cktmp.ensure(all_hits + all_hits / 4);
      }
      c->d_cg.ensure((size_t)nt2 * (CG_GMAX + 1));
      c->d_cg_n.ensure(nt2);
      c->d_fb.ensure(nt2);
      c->d_ck_off.ensure(nt2 + 2);  // [nt2]: the key region counter, [nt2 + 1]: the bitmap's
      HIP_CHECK(hipMemsetAsync(c->d_ck_off.p + nt2, 0, 16, s));
      {  // a read's bitmap: (n >> 6) + 1 words
        const uint64_t bw = (all_hits >> 6) + nt2;
        if (c->d_cbm.n < bw) c->d_cbm.ensure(bw + bw / 4);
      }
      c->d_cbm_off.ensure(nt2);
      CgArgs g{c->d_ck.p, c->d_cktmp.p, reinterpret_cast<unsigned long long*>(c->d_ck_off.p + nt2), c->d_ck_off.p,
               c->d_cg.p, c->d_cg_n.p, c->d_fb.p, c->d_ntier2.p + 2, c->d_cbm.p,
               reinterpret_cast<unsigned long long*>(c->d_ck_off.p + nt2 + 1), c->d_cbm_off.p};
      launch_cluster_grid(ca, g, c->d_tier2.p, nt2, s);
      HIP_CHECK(hipGetLastError());
      nt2b = download_u32(c, c->d_ntier2.p + 2);
      t2list = c->d_fb.p;
    }
    if (nt2b) {
      // the general block tier takes each read's hits contiguously: copies of the
      // listed reads' (chunked) hits first
      ClusterArgs cb = ca;
      if (c->d_hits_g.n < all_hits) {
        c->d_hits_g.ensure(all_hits + all_hits / 4);
        if (feat) c->d_hflag_g.ensure(all_hits + all_hits / 4);
      } else if (feat && c->d_hflag_g.n < all_hits) {
        c->d_hflag_g.ensure(all_hits + all_hits / 4);
      }
      c->d_hoff_g.ensure((size_t)B + 1);  // [B]: the region counter
      HIP_CHECK(hipMemsetAsync(c->d_hoff_g.p + B, 0, 8, s));
      launch_hits_gather(hits_base, hflag_base, c->d_hoff.p, c->d_chunk_tab.p, c->d_nhits.p, t2list, nt2b,
                         c->d_hits_g.p, feat ? c->d_hflag_g.p : nullptr, c->d_hoff_g.p,
                         reinterpret_cast<unsigned long long*>(c->d_hoff_g.p + B), s);
      HIP_CHECK(hipGetLastError());
      cb.hits = c->d_hits_g.p;
      cb.hoff = c->d_hoff_g.p;
      cb.chunk_tab = nullptr;
      if (feat) cb.hflag = c->d_hflag_g.p;
      // one 256-thread block per read, per-block scratch sized for the largest read; up to
      // 4096 blocks within 16 GiB (repeat-rich genomes send most reads here, with 10^4 hits each)
      cb.scratch_per_block = cluster_scratch_words(max_hits);
      const uint64_t fit = std::max<uint64_t>(1, (1ull << 31) / cb.scratch_per_block);
      const uint32_t blocks = (uint32_t)std::min<uint64_t>({(uint64_t)nt2b, 4096, fit});
      c->d_scratch.ensure((size_t)blocks * cb.scratch_per_block);
      cb.scratch = c->d_scratch.p;
      launch_cluster_tier2(cb, t2list, nt2b, blocks, s);
      HIP_CHECK(hipGetLastError());
    }
    HIP_CHECK(hipEventRecord(c->ev[7], s));
    c->stats.tier2_cluster_reads += nt2;
  }
  HIP_CHECK(hipStreamSynchronize(s));
  if (nt2) {
    float ms2 = 0.f;
    HIP_CHECK(hipEventElapsedTime(&ms2, c->ev[6], c->ev[7]));
    ms_cl += ms2;
  }
  c->stats.ms_k_cluster += ms_cl;
  c->stats.n_k_cluster += 1;
  c->stats.ms_cluster += ms_since(t0);
  t0 = std::chrono::steady_clock::now();
  // ---- DP jobs ----
  const int32_t band = (int32_t)c->opt.max_indel_len;
  const int32_t right_pad = (int32_t)std::max<uint32_t>(c->opt.max_indel_len, 10);  // single.rs:110
  c->d_jobs.ensure((size_t)B * TOP_CANDIDATES);
  c->d_njobs.ensure(B);
  JobArgs ja{};
  ja.G = c->dev_genome();
  ja.R = R;
  ja.cands = c->d_cands.p;
  ja.ncand = c->d_ncand.p;
  ja.right_pad = right_pad;
  ja.band = band;
  ja.jobs = c->d_jobs.p;
  ja.njobs = c->d_njobs.p;
  launch_build_jobs(ja, s);
  HIP_CHECK(hipGetLastError());
  b.max_n = b.max_m + (uint32_t)right_pad + LEFT_PAD;
  b.cap = b.max_m + b.max_n + 2;
  c->d_out.ensure((size_t)B * TOP_CANDIDATES);
  c->d_cigar.ensure((size_t)B * TOP_CANDIDATES * b.cap);
}

// Banded DP (single.rs:114-150) of the batch's job slots: every valid slot
// (jlist == nullptr) or the nlist slots of a device list.
static void batch_dp(pbw_ctx* c, Batch& b, const uint32_t* jlist, uint32_t nlist) {
  hipStream_t s = c->stream;
  const uint32_t B = b.B;
  const DevReads R = b.R;
  auto t0 = std::chrono::steady_clock::now();
  const int32_t band = (int32_t)c->opt.max_indel_len;
  const int32_t right_pad = (int32_t)std::max<uint32_t>(c->opt.max_indel_len, 10);  // single.rs:110
  const size_t nslots = (size_t)B * TOP_CANDIDATES;
  // flagged-job lists: quad -> band (a), band -> tiled (b); counters [0], [1]
  c->d_flag_a.ensure(nslots);
  c->d_flag_b.ensure(nslots);
  c->d_flag_n.ensure(4);
  HIP_CHECK(hipMemsetAsync(c->d_flag_n.p, 0, 8, s));
  DpArgs da = dp_args_base(c, R);
  da.jobs = c->d_jobs.p;
  da.jlist = jlist;
  da.njobs = jlist ? nlist : (uint32_t)nslots;
  da.job_valid = jlist ? nullptr : c->d_njobs.p;
  da.cigar_cap = b.cap;
  da.cigar = c->d_cigar.p;
  da.out = c->d_out.p;
  da.max_m = b.max_m;
  da.max_n = b.max_n;
  if (2 * band + 1 <= 63) {
    // four jobs per wave for the regular shapes; the flagged rest goes to the
    // one-job-per-wave banded kernel, and what that flags to the tiled one
    // valid jobs only, longest read first
    c->d_sorted.ensure(nslots);
    c->d_sort_hist.ensure(MAX_READ_LEN + 1);
    launch_job_sort(da, c->d_sort_hist.p, c->d_sorted.p, c->d_flag_n.p + 3, s);
    HIP_CHECK(hipGetLastError());
    const uint32_t nsorted = download_u32(c, c->d_flag_n.p + 3);
    // length split (pbw_tune): sorted jobs [0, n_long) are the reads longer than split_len;
    // after the scatter, hist[k] is the end of key k (key = MAX_READ_LEN - read length)
    const pbw_tune& tn = c->tune;
    const bool split = tn.split_len > 0 && tn.split_len < b.max_m;
    const uint32_t n_long = split ? download_u32(c, c->d_sort_hist.p + (MAX_READ_LEN - tn.split_len - 1)) : 0u;
    auto quad_args = [&](uint32_t lo, uint32_t cnt, uint32_t mm) {
      DpArgs q = da;
      q.jlist = c->d_sorted.p + lo;
      q.njobs = cnt;
      q.job_valid = nullptr;
      q.max_m = std::min<uint32_t>(mm, QUAD_MAX_M);
      q.max_n = std::min<uint32_t>(b.max_n, q.max_m + (uint32_t)right_pad + LEFT_PAD);
      q.flag_list = c->d_flag_a.p;
      q.flag_n = c->d_flag_n.p;
      q.quad_grid = tn.dp_grid;
      q.pack16 = tn.dp_pack16;
      return q;
    };
    DpArgs ba = da;
    ba.flag_list = c->d_flag_b.p;
    ba.flag_n = c->d_flag_n.p + 1;
    auto band_args = [&](uint32_t lo, uint32_t cnt) {
      DpArgs q = ba;
      q.jlist = c->d_sorted.p + lo;
      q.njobs = cnt;
      q.job_valid = nullptr;
      return q;
    };
    std::vector<DpArgs> quads, bands;
    if (!split) {
      quads.push_back(quad_args(0, nsorted, b.max_m));
    } else {
      if (tn.long_waves) bands.push_back(band_args(0, n_long));
      else quads.push_back(quad_args(0, n_long, b.max_m));
      quads.push_back(quad_args(n_long, nsorted - n_long, tn.split_len));
    }
    // every job of a four-jobs-per-wave launch keeps its traceback planes until
    // k_dp_walk reads them: launches of at most QUAD_CHUNK_BYTES of planes, each
    // cut where its reads fall below 3/4 of its longest, whose length sizes the
    // launch's LDS tile (shorter reads, more waves per CU)
    std::vector<DpArgs> chunks;
    if (!quads.empty()) {
      const uint32_t nkeys = (uint32_t)MAX_READ_LEN + 1;
      std::vector<uint32_t> hend(nkeys);  // end of each length key in the sorted list (key = MAX_READ_LEN - length)
      HIP_CHECK(hipMemcpy(hend.data(), c->d_sort_hist.p, nkeys * 4, hipMemcpyDeviceToHost));
      for (const DpArgs& q : quads) {
        const uint32_t base = (uint32_t)(q.jlist - c->d_sorted.p);
        uint32_t lo = 0;
        while (lo < q.njobs) {
          const uint32_t k0 = (uint32_t)(std::upper_bound(hend.begin(), hend.end(), base + lo) - hend.begin());
          const uint32_t len = (uint32_t)MAX_READ_LEN - std::min<uint32_t>(k0, MAX_READ_LEN);
          DpArgs k = q;
          k.max_m = std::max<uint32_t>(1, std::min(q.max_m, len));
          k.max_n = std::min<uint32_t>(q.max_n, k.max_m + (uint32_t)right_pad + LEFT_PAD);
          const uint64_t per = quad_scratch_words(k.max_m, k.max_n, 4, k.quad_grid);  // bytes per job
          uint32_t cj = (uint32_t)std::max<uint64_t>(4, std::min<uint64_t>(q.njobs - lo, QUAD_CHUNK_BYTES / per));
          // keys up to length 3/4 of the longest stay in this launch
          const uint32_t kcut = std::min<uint32_t>(nkeys - 1, (uint32_t)MAX_READ_LEN - (len * 3 + 3) / 4);
          if (len > 32 && hend[kcut] > base + lo) cj = std::min<uint32_t>(cj, hend[kcut] - (base + lo));
          if (cj < q.njobs - lo) cj = std::max<uint32_t>(4, cj & ~3u);
          k.jlist = q.jlist + lo;
          k.njobs = std::min(cj, q.njobs - lo);
          chunks.push_back(k);
          lo += k.njobs;
        }
      }
    }
    // two scratch buffers: launch i's walks (k_dp_walk, k_dp_lk) run on the second
    // stream while launch i + 1's rows run; launch i + 2 reuses the buffer once
    // launch i's walks are done
    size_t tbw = 0, tbe = 0;
    for (const DpArgs& q : chunks)
      if (q.max_m <= QUAD_MAX_M) {
        tbw = std::max(tbw, quad_scratch_words(q.max_m, q.max_n, q.njobs, q.quad_grid));
        tbe = std::max<size_t>(tbe, q.njobs);
      }
    const size_t nbuf = chunks.size() > 1 ? 2 : 1;
    c->d_tbq.ensure(std::max<size_t>(tbw, 1) * nbuf);
    c->d_tbend.ensure(std::max<size_t>(tbe, 1) * nbuf);
    HIP_CHECK(hipEventRecord(c->ev[2], s));
    bool walks = false;
    for (size_t ci = 0; ci < chunks.size(); ci++) {
      DpArgs& q = chunks[ci];
      const size_t bi = ci % nbuf;
      q.tbq = c->d_tbq.p + bi * std::max<size_t>(tbw, 1);
      q.tbend = c->d_tbend.p + bi * std::max<size_t>(tbe, 1);
      if (ci >= nbuf) HIP_CHECK(hipStreamWaitEvent(s, c->dev[2 + bi], 0));  // the buffer's last walks
      // a launch whose tile does not fit LDS goes to the one-job-per-wave kernel
      const int kind = launch_dp_quad(q, s);
      if (kind == 0) bands.push_back(band_args((uint32_t)(q.jlist - c->d_sorted.p), q.njobs));
      if (kind == 1) {
        HIP_CHECK(hipEventRecord(c->dev[bi], s));
        HIP_CHECK(hipStreamWaitEvent(c->stream2, c->dev[bi], 0));
        launch_dp_walk(q, c->stream2);
        HIP_CHECK(hipEventRecord(c->dev[2 + bi], c->stream2));
        walks = true;
      } else if (ci + nbuf < chunks.size()) {
        HIP_CHECK(hipEventRecord(c->dev[2 + bi], s));  // the buffer is free once this launch is done
      }
    }
    if (walks) {  // the main stream continues after every walk
      HIP_CHECK(hipEventRecord(c->dev[0], c->stream2));
      HIP_CHECK(hipStreamWaitEvent(s, c->dev[0], 0));
    }
    for (const DpArgs& q : bands) run_dp_band(c, q);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipEventRecord(c->ev[3], s));
    const uint32_t nflag = download_u32(c, c->d_flag_n.p);
    {
      float ms = 0.f;
      HIP_CHECK(hipEventSynchronize(c->ev[3]));
      HIP_CHECK(hipEventElapsedTime(&ms, c->ev[2], c->ev[3]));
      c->stats.ms_k_dp += ms;
      c->stats.n_k_dp += 1;
    }
    if (split && tn.long_waves) c->stats.band_dp_jobs += n_long;
    if (nflag) {
      c->stats.band_dp_jobs += nflag;
      ba.jlist = c->d_flag_a.p;
      ba.njobs = nflag;
      ba.job_valid = nullptr;
      run_dp_band(c, ba);
      HIP_CHECK(hipGetLastError());
    }
    const uint32_t ntiled = download_u32(c, ba.flag_n);
    run_dp_tiled_dev(c, da, c->d_flag_b.p, ntiled);
    c->stats.generic_dp_jobs += ntiled;
  } else {
    std::vector<uint32_t> list;
    if (jlist) {
      list.resize(nlist);
      if (nlist) HIP_CHECK(hipMemcpyAsync(list.data(), jlist, (size_t)nlist * 4, hipMemcpyDeviceToHost, s));
      HIP_CHECK(hipStreamSynchronize(s));
    } else {
      std::vector<uint8_t> nj(B);
      HIP_CHECK(hipMemcpyAsync(nj.data(), c->d_njobs.p, B, hipMemcpyDeviceToHost, s));
      HIP_CHECK(hipStreamSynchronize(s));
      for (uint32_t r = 0; r < B; r++)
        for (uint32_t k = 0; k < nj[r]; k++) list.push_back(r * TOP_CANDIDATES + k);
    }
    run_dp_tiled(c, da, list);
    c->stats.generic_dp_jobs += list.size();
  }
  HIP_CHECK(hipStreamSynchronize(s));
  c->stats.ms_dp += ms_since(t0);
}

// Downloads records (already in d_rec) and gathers their CIGARs from one or
// two pools (src bit 63 selects the second); unmapped records get {L}S.
// Records [0, nrec) of a batch (in d_rec) and their CIGAR runs to host memory.
// d_src[i]: first word of record i's runs in its pool (bit 63: pool1), ~0 when
// unmapped.  Word counts and offsets are scanned on the device; the host reads
// the batch's word count and copies records and runs straight into the arrays
// the output batch takes over.
static void download_records(pbw_ctx* c, const Batch& b, uint32_t nrec, const uint64_t* d_src, const uint32_t* pool1,
                             HostRecords& hr) {
  hipStream_t s = c->stream;
  if (!nrec) {
    HIP_CHECK(hipStreamSynchronize(s));
    return;
  }
  if (hr.cigar.size() > 0xFFFFFFFFull)
    throw PbwError{PBW_ERR_LIMIT, "more than 2^32 CIGAR words in one call; map fewer reads per call"};
  const uint32_t base = (uint32_t)hr.cigar.size();
  const uint32_t nb = (nrec + REC_BLOCK - 1) / REC_BLOCK;
  c->d_rec_bsum.ensure(nb + 1);
  uint32_t* d_total = c->d_rec_bsum.p + nb;
  uint8_t* d_rec = c->d_rec.p;
  launch_rec_offsets(d_rec, d_src, nrec, base, c->d_rec_bsum.p, d_total, s);
  HIP_CHECK(hipGetLastError());
  pbw_record* hrec = hr.rec.grow(nrec);
  uint32_t total = 0;
  HIP_CHECK(hipMemcpyAsync(hrec, d_rec, (size_t)nrec * sizeof(pbw_record), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipMemcpyAsync(&total, d_total, 4, hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  if ((uint64_t)base + total > 0xFFFFFFFFull)
    throw PbwError{PBW_ERR_LIMIT, "more than 2^32 CIGAR words in one call; map fewer reads per call"};
  c->d_cigar_out.ensure(std::max<uint32_t>(total, 1));
  launch_rec_cigar(d_rec, d_src, c->d_cigar.p, pool1 ? pool1 : c->d_cigar.p, b.R.seq_off, nrec, base,
                   c->d_cigar_out.p, s);
  HIP_CHECK(hipGetLastError());
  uint32_t* hc = hr.cigar.grow(total);
  HIP_CHECK(hipMemcpyAsync(hc, c->d_cigar_out.p, (size_t)total * 4, hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
}

static void collect_counters(pbw_ctx* c, uint32_t reads) {
  Counters ctr{};
  HIP_CHECK(hipMemcpy(&ctr, c->d_ctr.p, sizeof(Counters), hipMemcpyDeviceToHost));
  c->stats.lookups += ctr.lookups;
  c->stats.positions += ctr.positions;
  c->stats.filtered_hits += ctr.filtered;
  c->stats.dp_cells += ctr.cells;
  c->stats.mapq_near_boundary += ctr.mapq_near;
  c->stats.reads += reads;
  HIP_CHECK(hipMemsetAsync(c->d_ctr.p, 0, sizeof(Counters), c->stream));
}

static void mlp_forward_dev(pbw_ctx* c, const uint16_t* dx, uint64_t n);

// MAPQ arguments of a batch (records to d_rec, best slots to d_best).
static MapqArgs mapq_args(pbw_ctx* c, const Batch& b) {
  MapqArgs ma{};
  ma.G = c->dev_genome();
  ma.R = b.R;
  ma.T = c->luts;
  ma.jobs = c->d_jobs.p;
  ma.out = c->d_out.p;
  ma.cigar = c->d_cigar.p;
  ma.cigar_cap = b.cap;
  ma.njobs = c->d_njobs.p;
  ma.cands = c->d_cands.p;
  ma.rcode = c->d_rcode.p;
  ma.rec = c->d_rec.p;
  ma.best = c->d_best.p;
  ma.ctr = c->d_ctr.p;
  return ma;
}

// Candidate prefilter between clustering and the DP (no counterpart in the
// reference): the network scores every candidate, the DP aligns the kept ones,
// MAPQ uses the kept ones, and ambiguous or audited reads are re-run plain.
static void prefilter_dp_mapq(pbw_ctx* c, Batch& b, uint64_t read_base) {
  hipStream_t s = c->stream;
  const uint32_t B = b.B;
  const uint64_t nslots = (uint64_t)B * TOP_CANDIDATES;
  auto tp = std::chrono::steady_clock::now();
  mlp_forward_dev(c, c->d_feat.p, nslots);
  c->d_pf_keep.ensure(B);
  c->d_pf_list.ensure(nslots);
  c->d_pf_n.ensure(2);
  c->d_pf_ctr.ensure(8);
  HIP_CHECK(hipMemsetAsync(c->d_pf_n.p, 0, 8, s));
  HIP_CHECK(hipMemsetAsync(c->d_pf_ctr.p, 0, 8 * sizeof(unsigned long long), s));
  PfArgs pa{};
  pa.R = b.R;
  pa.jobs = c->d_jobs.p;
  pa.out = c->d_out.p;
  pa.njobs = c->d_njobs.p;
  pa.logit = c->d_mlp_z.p;
  pa.threshold = c->pf.threshold;
  pa.top_k = c->pf.top_k;
  pa.keep = c->d_pf_keep.p;
  pa.list = c->d_pf_list.p;
  pa.nlist = c->d_pf_n.p;
  pa.ctr = c->d_pf_ctr.p;
  launch_pf_select(pa, s);
  HIP_CHECK(hipGetLastError());
  const uint32_t nk = download_u32(c, c->d_pf_n.p);
  c->stats.ms_pf += ms_since(tp);
  batch_dp(c, b, c->d_pf_list.p, nk);
  auto t0 = std::chrono::steady_clock::now();
  c->d_rec.ensure((size_t)B * sizeof(pbw_record));
  c->d_best.ensure(B);
  MapqArgs ma = mapq_args(c, b);
  ma.keep = c->d_pf_keep.p;
  launch_mapq(ma, s);
  HIP_CHECK(hipGetLastError());
  // re-run verdicts
  c->d_verdict.ensure(B);
  c->d_pf_rlist.ensure(B);
  HIP_CHECK(hipMemsetAsync(c->d_pf_n.p, 0, 8, s));
  pa.rec = c->d_rec.p;
  pa.mapq_min = c->pf.mapq_fallback;
  pa.close_nats = c->pf.close_nats;
  pa.audit_period = c->pf.audit_period;
  pa.read_base = read_base;
  pa.verdict = c->d_verdict.p;
  pa.rlist = c->d_pf_rlist.p;
  pa.nrlist = c->d_pf_n.p + 1;
  pa.audit = c->d_pf_ctr.p + 4;
  launch_pf_verdict(pa, s);
  HIP_CHECK(hipGetLastError());
  uint32_t cnt[2];
  HIP_CHECK(hipMemcpyAsync(cnt, c->d_pf_n.p, 8, hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  c->stats.ms_mapq += ms_since(t0);
  if (cnt[0]) batch_dp(c, b, c->d_pf_list.p, cnt[0]);
  if (cnt[1]) {
    t0 = std::chrono::steady_clock::now();
    c->d_rec2.ensure((size_t)B * sizeof(pbw_record));
    c->d_best2.ensure(B);
    MapqArgs m2 = mapq_args(c, b);
    m2.rec = c->d_rec2.p;
    m2.best = c->d_best2.p;
    m2.rlist = c->d_pf_rlist.p;
    m2.nlist = cnt[1];
    launch_mapq(m2, s);
    pa.rec_out = c->d_rec.p;
    pa.rec2 = c->d_rec2.p;
    pa.best = c->d_best.p;
    pa.best2 = c->d_best2.p;
    pa.cigar = c->d_cigar.p;
    pa.cigar_cap = b.cap;
    launch_pf_merge(pa, cnt[1], s);
    HIP_CHECK(hipGetLastError());
    c->stats.ms_mapq += ms_since(t0);
  }
  unsigned long long ctr[8];
  HIP_CHECK(hipMemcpyAsync(ctr, c->d_pf_ctr.p, sizeof(ctr), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  c->stats.pf_jobs += ctr[0];
  c->stats.pf_jobs_aligned += ctr[1] + ctr[3];
  c->stats.pf_audit_reads += ctr[4];
  c->stats.pf_audit_mismatches += ctr[5];
  c->stats.pf_fallback_reads += ctr[6];
}

// SingleEndMapper::map (single.rs:164-207) for every read of a batch.
// Device words a batch reserves per read for its candidates' CIGAR slots:
// 20 slots of max_m + max_n + 2 runs, max_n = max_m + right pad + LEFT_PAD.
static uint64_t cigar_words_per_read(const pbw_ctx* c, uint64_t max_m) {
  const uint64_t pad = std::max<uint32_t>(c->opt.max_indel_len, 10) + LEFT_PAD;
  return (uint64_t)TOP_CANDIDATES * (2 * max_m + pad + 2);
}
constexpr uint64_t CIGAR_BUDGET_WORDS = (16ull << 30) / 4;  // per batch

// End of the single-end batch that starts at read b0: at most bs reads before
// `end`, cut where the batch's CIGAR slots would pass 16 GiB (only long reads
// reach it; the batch boundaries never change a record).
static uint64_t se_batch_end(const pbw_ctx* c, const pbw_read_batch* in, uint64_t b0, uint64_t end, uint64_t bs) {
  const uint64_t e = std::min(end, b0 + bs);
  uint64_t mx = 0;
  for (uint64_t r = b0; r < e; r++) {
    mx = std::max({mx, in->seq_off[r + 1] - in->seq_off[r], in->qual_off[r + 1] - in->qual_off[r]});
    if (r > b0 && (r + 1 - b0) * cigar_words_per_read(c, mx) > CIGAR_BUDGET_WORDS) return r;
  }
  return e;
}

static void map_se_batch(pbw_ctx* c, const pbw_read_batch* in, uint64_t b0, uint64_t b1, HostRecords& hr,
                         bool resident) {
  hipStream_t s = c->stream;
  Batch b;
  batch_upload_slice(c, b, in, b0, b1, resident);
  HIP_CHECK(hipMemsetAsync(c->d_ctr.p, 0, sizeof(Counters), s));
  const bool pf = c->pf.enabled && c->mlp_ready;
  batch_candidates(c, b, pf, &hr);
  const uint32_t B = b.B;
  if (pf) {
    prefilter_dp_mapq(c, b, b0);
  } else {
    batch_dp(c, b, nullptr, 0);
    auto t0 = std::chrono::steady_clock::now();
    c->d_rec.ensure((size_t)B * sizeof(pbw_record));
    c->d_best.ensure(B);
    launch_mapq(mapq_args(c, b), s);
    HIP_CHECK(hipGetLastError());
    c->stats.ms_mapq += ms_since(t0);
  }
  auto t0 = std::chrono::steady_clock::now();
  c->d_rec_src.ensure(B);
  launch_rec_src_se(c->d_best.p, B, b.cap, c->d_rec_src.p, s);
  download_records(c, b, B, c->d_rec_src.p, nullptr, hr);
  c->stats.ms_mapq += ms_since(t0);
  collect_counters(c, B);
}

// ---------------------------------------------------------------------------
// paired-end: paired.rs:101-234 per batch with a frozen insert-size snapshot
// ---------------------------------------------------------------------------
struct HostInsert {  // insert_size.rs:23-44
  double mean, sd;
  uint64_t n;
  void update(double obs) {
    n += 1;
    double nn = (double)n;
    double delta = obs - mean;
    mean += delta / nn;
    double delta2 = obs - mean;
    double var = nn > 1.0 ? ((sd * sd) * (nn - 1.0) + delta * delta2) / nn : sd * sd;
    sd = std::max(std::sqrt(var), 1.0);
  }
};

static uint64_t rust_f64_to_usize(double x) {  // `as usize`: saturating, NaN -> 0
  if (!(x > 0.0)) return 0;
  if (x >= 18446744073709551615.0) return UINT64_MAX;
  return (uint64_t)x;
}

// The part of a batch of pairs that does not depend on the insert-size model:
// both mates' reads (interleaved r1, r2) through seeding, clustering and the
// single-end DP.
static void pe_prepare(pbw_ctx* c, const PairPack& pk, Batch& b, HostRecords* pre = nullptr) {
  const uint32_t P = pk.B / 2;
  batch_send(c, b, pk);
  batch_candidates(c, b, false, pre);
  batch_dp(c, b, nullptr, 0);
  const size_t nws = (size_t)P * 40;
  c->d_wjobs.ensure(nws);
  c->d_wvalid.ensure(nws);
  c->d_slcand.ensure(nws);
  c->d_plan.ensure(P);
  c->d_wout.ensure(nws);
  c->d_rec.ensure((size_t)2 * P * sizeof(pbw_record));
  c->d_cig_src.ensure(2 * (size_t)P);
}

// Cross-alignment windows from a device list whose length is on the device
// (*d_n <= cap): k_dp_wide for the whole-window rectangles, k_dp_tiled for the
// jobs it flags (or for all when the windows are wider than k_dp_wide's
// tiles).  No host round trip: jobs beyond the tiled kernel's shape are counted
// in *d_bad for the caller to check.
static void run_dp_wide_dev(pbw_ctx* c, DpArgs a, const uint32_t* d_list, const uint32_t* d_n, uint32_t cap,
                            uint32_t* d_flag_n, uint32_t* d_bad) {
  if (!cap) return;
  hipStream_t s = c->stream;
  a.job_valid = nullptr;
  const uint32_t T = dp_wide_tiles(a.max_m, a.max_n);
  const uint32_t* tl = d_list;
  const uint32_t* tn = d_n;
  uint32_t tw = std::min<uint32_t>(cap, 2048);
  if (T) {
    DpArgs w = a;
    w.jlist = d_list;
    w.njobs = cap;
    w.njobs_dev = d_n;
    c->d_flag_w.ensure(cap);
    w.flag_list = c->d_flag_w.p;
    w.flag_n = d_flag_n;
    const uint32_t nwaves = std::min<uint32_t>(cap, 8192);
    c->d_tb.ensure((size_t)nwaves * (a.max_m + 1) * T * 16);  // 16 bits per lane, tile and row
    w.tb = c->d_tb.p;
    launch_dp_wide(w, nwaves, s);
    HIP_CHECK(hipGetLastError());
    tl = c->d_flag_w.p;
    tn = d_flag_n;
    tw = std::min<uint32_t>(cap, 64);  // the flagged jobs: few or none
  }
  a.jlist = tl;
  a.njobs = cap;
  a.njobs_dev = tn;
  c->d_flag_c.ensure(cap);
  a.flag_list = c->d_flag_c.p;
  a.flag_n = d_bad;
  tw = tiled_scratch(c, a, tw);
  launch_dp_tiled(a, tw, s);
  HIP_CHECK(hipGetLastError());
}

// The cross-alignment CIGAR pool: slots base .. base + nslots - 1 at one stride,
// which only grows while the pool is in use (slots written earlier keep their
// places, since a record's CIGAR source is an absolute offset, and slots are
// filled in increasing order, so a larger stride never overlaps them).
struct WidePool {
  uint64_t base;
  size_t nslots;
  uint32_t stride;
};

// Pairs [q0, q1) of a prepared batch under the model snapshot `snap`: cross-
// alignment plan, wide mate windows, joint posterior and records, queued on the
// stream.  One synchronisation brings back the counts and the observations
// (obs_out / has_out) and, when hr is given, the records and the CIGAR total;
// the CIGAR runs then follow asynchronously and are in hr by the stream's next
// synchronisation.
static void pe_pairs(pbw_ctx* c, const Batch& b, uint32_t q0, uint32_t q1, const HostInsert& snap, WidePool& pool,
                     HostRecords* hr, std::vector<double>& obs_out, std::vector<uint8_t>& has_out) {
  hipStream_t s = c->stream;
  const uint32_t P = q1 - q0, nws = P * 40;
  // the call's device words for the host, one copy up at the end: 4 counts
  // ([0] windows listed, [1] flagged by k_dp_wide, [2] beyond the tiled shape,
  // [3] the records' CIGAR total), then the observations and their flags
  const size_t io_bytes = 16 + (size_t)P * 9;
  c->d_pe_io.ensure(io_bytes);
  uint32_t* d_wn = reinterpret_cast<uint32_t*>(c->d_pe_io.p);
  HIP_CHECK(hipMemsetAsync(d_wn, 0, 4 * sizeof(uint32_t), s));
  c->d_wlist.ensure(nws);
  PairArgs pa{};
  pa.ctr = c->d_ctr.p;
  pa.G = c->dev_genome();
  pa.R = b.R;
  pa.T = c->luts;
  pa.jobs = c->d_jobs.p;
  pa.out = c->d_out.p;
  pa.cigar = c->d_cigar.p;
  pa.cigar_cap = b.cap;
  pa.njobs = c->d_njobs.p;
  pa.cands = c->d_cands.p;
  pa.rcode = c->d_rcode.p;
  pa.wjobs = c->d_wjobs.p;
  pa.wvalid = c->d_wvalid.p;
  pa.wlist = c->d_wlist.p;
  pa.wn = d_wn;
  pa.slcand = c->d_slcand.p;
  pa.plan = c->d_plan.p;
  pa.mean = snap.mean;
  pa.sd = snap.sd;
  pa.ln_norm = std::log(2.0 * M_PI * snap.sd * snap.sd);  // insert_size.rs:27-28
  pa.sv_prior_log = -c->opt.sv_prior_phred / 10.0 * H_LN_10;  // paired.rs:97
  pa.mean_usize = rust_f64_to_usize(std::max(snap.mean, 0.0));
  pa.radius_base = rust_f64_to_usize(std::ceil(5.0 * snap.sd));
  pa.name_class = c->d_name_class.p;
  pa.P = P;
  pa.p_base = q0;
  launch_pair_plan(pa, s);
  HIP_CHECK(hipGetLastError());
  // wide DP (align_at_window, single.rs:220-275): window <= 2 radius + L.  A
  // CIGAR has at most 2 m + 1 runs (every run but deletions takes a read base,
  // and two deletion runs are never adjacent), so slots of 2 max_m + 2 words
  // hold any window's alignment.
  const uint64_t wmax = 2 * (pa.radius_base + b.max_m) + b.max_m;
  if (wmax > 0x7FFFFFFFull) throw PbwError{PBW_ERR_LIMIT, "insert-size search window beyond 2^31 bases"};
  const uint32_t max_n = (uint32_t)wmax;
  const uint32_t wcap = 2 * b.max_m + 2;
  if (wcap > pool.stride) {
    if (pool.stride) c->d_wcigar.grow_keep(pool.nslots * (size_t)wcap, s);
    else c->d_wcigar.ensure(pool.nslots * (size_t)wcap);
    pool.stride = wcap;
  }
  const uint64_t wbase = pool.base;
  c->d_ctr_w.ensure(1);
  DpArgs da = dp_args_base(c, b.R);
  da.jobs = c->d_wjobs.p;
  da.cigar_cap = pool.stride;
  da.cigar = c->d_wcigar.p;
  da.cigar_base = wbase;
  da.out = c->d_wout.p;
  da.max_m = b.max_m;
  da.max_n = max_n;
  da.ctr = c->d_ctr_w.p;
  HIP_CHECK(hipEventRecord(c->evw[0], s));
  run_dp_wide_dev(c, da, c->d_wlist.p, d_wn, nws, d_wn + 1, d_wn + 2);
  HIP_CHECK(hipEventRecord(c->evw[1], s));
  // joint posterior + records
  pa.wout = c->d_wout.p;
  pa.wcigar = c->d_wcigar.p;
  pa.wcap = pool.stride;
  pa.wbase = wbase;
  pa.rec = c->d_rec.p;
  pa.cig_src = c->d_cig_src.p;
  pa.obs = reinterpret_cast<double*>(c->d_pe_io.p + 16);
  pa.has_obs = c->d_pe_io.p + 16 + (size_t)P * 8;
  launch_pair_finalize(pa, s);
  HIP_CHECK(hipGetLastError());
  // records [2 q0, 2 q1): CIGAR offsets on the device, then one synchronisation
  const uint32_t nrec = 2 * P, rec0 = 2 * q0;
  uint32_t base = 0;
  uint8_t* d_rec = c->d_rec.p + (size_t)rec0 * sizeof(pbw_record);
  const uint64_t* d_src = c->d_cig_src.p + rec0;
  uint32_t* d_total = nullptr;
  if (hr) {
    if (hr->cigar.size() > 0xFFFFFFFFull)
      throw PbwError{PBW_ERR_LIMIT, "more than 2^32 CIGAR words in one call; map fewer reads per call"};
    base = (uint32_t)hr->cigar.size();
    const uint32_t nb = (nrec + REC_BLOCK - 1) / REC_BLOCK;
    c->d_rec_bsum.ensure(nb);
    d_total = d_wn + 3;
    launch_rec_offsets(d_rec, d_src, nrec, base, c->d_rec_bsum.p, d_total, s);
    HIP_CHECK(hipGetLastError());
    pbw_record* hrec = hr->rec.grow(nrec);
    HIP_CHECK(hipMemcpyAsync(hrec, d_rec, (size_t)nrec * sizeof(pbw_record), hipMemcpyDeviceToHost, s));
  }
  const size_t stage = io_bytes;
  if (c->h_pe_cap < stage) {
    if (c->h_pe) (void)hipHostFree(c->h_pe);
    c->h_pe = nullptr;
    c->h_pe_cap = 0;
    if (hipHostMalloc((void**)&c->h_pe, stage, hipHostMallocDefault) != hipSuccess || !c->h_pe)
      throw PbwError{PBW_ERR_LIMIT, "pinned host allocation of " + std::to_string(stage) + " bytes failed"};
    c->h_pe_cap = stage;
  }
  uint32_t* h_cnt = reinterpret_cast<uint32_t*>(c->h_pe);
  double* h_obs = reinterpret_cast<double*>(c->h_pe + 16);
  uint8_t* h_has = c->h_pe + 16 + (size_t)P * 8;
  HIP_CHECK(hipMemcpyAsync(c->h_pe, c->d_pe_io.p, io_bytes, hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  {
    float ms = 0.f;
    HIP_CHECK(hipEventElapsedTime(&ms, c->evw[0], c->evw[1]));
    c->stats.ms_k_wide += ms;
    c->stats.n_k_wide += 1;
  }
  if (h_cnt[2]) throw PbwError{PBW_ERR_LIMIT, "alignment window beyond the tiled kernel's shape"};
  c->stats.wide_jobs += h_cnt[0];
  obs_out.assign(h_obs, h_obs + P);
  has_out.assign(h_has, h_has + P);
  if (!hr) return;
  const uint32_t total = h_cnt[3];
  if ((uint64_t)base + total > 0xFFFFFFFFull)
    throw PbwError{PBW_ERR_LIMIT, "more than 2^32 CIGAR words in one call; map fewer reads per call"};
  c->d_cigar_out.ensure(std::max<uint32_t>(total, 1));
  launch_rec_cigar(d_rec, d_src, c->d_cigar.p, c->d_wcigar.p, b.R.seq_off + rec0, nrec, base, c->d_cigar_out.p, s);
  HIP_CHECK(hipGetLastError());
  uint32_t* hc = hr->cigar.grow(total);
  if (total) HIP_CHECK(hipMemcpyAsync(hc, c->d_cigar_out.p, (size_t)total * 4, hipMemcpyDeviceToHost, s));
}

// Cross-alignment DP cells to the statistics (synchronises the stream).
static void collect_wide_cells(pbw_ctx* c) {
  if (!c->d_ctr_w.p) return;
  Counters w{};
  HIP_CHECK(hipMemcpyAsync(&w, c->d_ctr_w.p, sizeof(Counters), hipMemcpyDeviceToHost, c->stream));
  HIP_CHECK(hipStreamSynchronize(c->stream));
  c->stats.wide_cells += w.cells;
  HIP_CHECK(hipMemsetAsync(c->d_ctr_w.p, 0, sizeof(Counters), c->stream));
}

// Packs a pair batch on a worker thread while the caller keeps the stream
// busy; wait() hands over the pack (or rethrows the worker's error).  The
// destructor waits too, so an error on the caller's side never leaves the
// worker running.
struct PackAhead {
  std::thread th;
  std::exception_ptr err;
  void start(pbw_ctx* c, int slot, const pbw_read_batch* m1, const pbw_read_batch* m2, uint64_t p0, uint64_t p1,
             PairPack& pk) {
    err = nullptr;
    th = std::thread([this, c, slot, m1, m2, p0, p1, &pk] {
      try {
        pair_pack(c, slot, m1, m2, p0, p1, pk);
      } catch (...) {
        err = std::current_exception();
      }
    });
  }
  void wait() {
    if (th.joinable()) th.join();
    if (err) std::rethrow_exception(err);
  }
  ~PackAhead() {
    if (th.joinable()) th.join();
  }
};

// lib.rs:145-185: 500 serial warm-up pairs, then 2000-pair batches
static void map_pe_all(pbw_ctx* c, const pbw_read_batch* m1, const pbw_read_batch* m2, pbw_insert_model* model,
                       HostRecords& hr) {
  HostInsert im{model->mean, model->sd, model->n_observed};
  const uint64_t n = m1->n;
  std::vector<double> obs;
  std::vector<uint8_t> has;
  const uint64_t warm = std::min<uint64_t>(n, 500);
  const uint64_t bs = c->opt.batch_reads ? c->opt.batch_reads : 131072;
  // pairs prepared together: whole 2000-pair batches, fewer when long mates
  // would take the CIGAR slots past their budget
  uint64_t mx = 0;
  for (const pbw_read_batch* x : {m1, m2})
    for (uint64_t r = 0; r < n; r++)
      mx = std::max({mx, x->seq_off[r + 1] - x->seq_off[r], x->qual_off[r + 1] - x->qual_off[r]});
  const uint64_t fit = CIGAR_BUDGET_WORDS / (2 * cigar_words_per_read(c, mx) * 2000);
  const uint64_t per_prep = std::max<uint64_t>(1, std::min<uint64_t>(bs / 2 / 2000, fit)) * 2000;
  PairPack packs[2];
  PackAhead ahead;
  int cur = 0;
  if (warm < n) ahead.start(c, 0, m1, m2, warm, std::min<uint64_t>(warm + per_prep, n), packs[0]);
  if (warm) {
    // the warm-up pairs go through seeding, clustering and the single-end DP
    // together (none of it reads the model); then pair by pair, each on the
    // model its predecessors left, through the model-dependent stages
    PairPack wp;
    pair_pack(c, 1, m1, m2, 0, warm, wp);
    Batch b;
    pe_prepare(c, wp, b, &hr);  // the records' pinned space is reserved during the seed kernel
    auto t0 = std::chrono::steady_clock::now();
    WidePool pool{0, (size_t)warm * 40, 0};
    for (uint64_t i = 0; i < warm; i++) {
      HostInsert snap = im;
      pe_pairs(c, b, (uint32_t)i, (uint32_t)i + 1, snap, pool, nullptr, obs, has);
      if (has[0]) im.update(obs[0]);
    }
    download_records(c, b, 2 * (uint32_t)warm, c->d_cig_src.p, c->d_wcigar.p, hr);
    c->stats.ms_pair += ms_since(t0);
    collect_counters(c, 2 * (uint32_t)warm);
    collect_wide_cells(c);
  }
  // the 2000-pair batches: seeding, clustering and the single-end DP for many
  // of them at once (whole batches, so the boundaries stay warm + 2000 k); then
  // batch by batch, each on the model the batches before it left, the
  // model-dependent stages and that batch's records.  The next group of
  // batches is packed for upload meanwhile, on a worker thread.
  for (uint64_t s0 = warm; s0 < n; s0 += per_prep) {
    const uint64_t s1 = std::min<uint64_t>(s0 + per_prep, n);
    ahead.wait();
    Batch b;
    pe_prepare(c, packs[cur], b, s0 == 0 ? &hr : nullptr);
    if (s1 < n) ahead.start(c, cur ^ 1, m1, m2, s1, std::min<uint64_t>(s1 + per_prep, n), packs[cur ^ 1]);
    cur ^= 1;
    for (uint64_t b0 = s0; b0 < s1; b0 += 2000) {
      const uint64_t b1 = std::min<uint64_t>(b0 + 2000, s1);
      const uint32_t q0 = (uint32_t)(b0 - s0), q1 = (uint32_t)(b1 - s0);
      HostInsert snap = im;
      auto t0 = std::chrono::steady_clock::now();
      WidePool pool{(uint64_t)q0 * 40, (size_t)(q1 - q0) * 40, 0};
      pe_pairs(c, b, q0, q1, snap, pool, &hr, obs, has);
      c->stats.ms_pair += ms_since(t0);
      for (uint64_t k = 0; k < b1 - b0; k++)
        if (has[k]) im.update(obs[k]);
    }
    HIP_CHECK(hipStreamSynchronize(c->stream));
    collect_counters(c, 2 * (uint32_t)(s1 - s0));
    collect_wide_cells(c);
  }
  model->mean = im.mean;
  model->sd = im.sd;
  model->n_observed = im.n;
}

static void emit_batch(HostRecords& hr, pbw_record_batch* out) {
  out->n = hr.rec.size();
  out->n_cigar = hr.cigar.size();
  out->rec = hr.rec.release();
  out->cigar = hr.cigar.release();
}

template <typename F>
static int guarded(pbw_ctx* c, F&& f) {
  try {
    f();
    return PBW_OK;
  } catch (const PbwError& e) {
    if (c) c->error = e.msg;
    g_error = e.msg;
    return e.code;
  } catch (const std::exception& e) {
    if (c) c->error = e.what();
    g_error = e.what();
    return PBW_ERR_ARG;
  }
}

// ===========================================================================
// candidate prefilter network: parameter upload, forward, gradient, training
// ===========================================================================
static void mlp_set(pbw_ctx* c, const float* theta) {
  hipStream_t s = c->stream;
  MlpDev P = c->mlp();
  HIP_CHECK(hipMemcpyAsync(P.theta, theta, MLP_NP * sizeof(float), hipMemcpyHostToDevice, s));
  HIP_CHECK(hipMemsetAsync(P.m, 0, 2 * (size_t)MLP_NP * sizeof(float), s));
  launch_mlp_publish(P, s);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipStreamSynchronize(s));
  c->mlp_ready = true;
  c->mlp_step = 0;
}

// He-uniform weights (U(-sqrt(6/fan_in), +)), zero biases; splitmix64 stream.
static void mlp_init_host(uint64_t seed, std::vector<float>& th) {
  th.assign(MLP_NP, 0.f);
  uint64_t x = seed * 0x9E3779B97F4A7C15ull + 1;
  auto uni = [&]() {
    x += 0x9E3779B97F4A7C15ull;
    uint64_t z = x;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
  };
  const double a1 = std::sqrt(6.0 / MLP_IN), a2 = std::sqrt(6.0 / MLP_H);
  for (int p = MLP_OFF_W1; p < MLP_OFF_B1; p++) th[p] = (float)((2 * uni() - 1) * a1);
  for (int p = MLP_OFF_W2; p < MLP_OFF_B2; p++) th[p] = (float)((2 * uni() - 1) * a2);
  for (int p = MLP_OFF_W3; p < MLP_OFF_B3; p++) th[p] = (float)((2 * uni() - 1) * a2);
}

static void mlp_require(pbw_ctx* c) {
  if (!c->mlp_ready) throw PbwError{PBW_ERR_ARG, "prefilter network has no parameters (pbw_mlp_init or pbw_mlp_set_params)"};
}

// Forward over n rows already on the device; logits to d_mlp_z.
static void mlp_forward_dev(pbw_ctx* c, const uint16_t* dx, uint64_t n) {
  c->d_mlp_z.ensure(n);
  launch_mlp_forward(c->mlp(), dx, n, c->d_mlp_z.p, c->stream);
  HIP_CHECK(hipGetLastError());
}

// One Adam epoch over the device dataset (d_mlp_x, d_mlp_y) in a shuffled order; returns the mean loss.
static double mlp_epoch(pbw_ctx* c, uint64_t n, const pbw_train_options& o, uint64_t epoch) {
  hipStream_t s = c->stream;
  std::vector<uint32_t> perm(n);
  for (uint64_t i = 0; i < n; i++) perm[i] = (uint32_t)i;
  std::mt19937_64 rng(o.seed * 1000003ull + epoch);
  std::shuffle(perm.begin(), perm.end(), rng);
  c->d_mlp_perm.upload(perm.data(), n, s);
  const uint64_t B = o.batch;
  const uint64_t steps = (n + B - 1) / B;
  const uint32_t tiles_max = (uint32_t)((B + 31) / 32);
  c->d_mlp_part.ensure((size_t)tiles_max * MLP_NP);
  c->d_mlp_loss.ensure((size_t)steps * tiles_max);
  HIP_CHECK(hipMemsetAsync(c->d_mlp_loss.p, 0, (size_t)steps * tiles_max * sizeof(float), s));
  MlpDev P = c->mlp();
  for (uint64_t st = 0; st < steps; st++) {
    const uint64_t r0 = st * B, nr = std::min<uint64_t>(B, n - r0);
    launch_mlp_grad(P, c->d_mlp_x.p, c->d_mlp_y.p, c->d_mlp_perm.p, r0, nr, n, o.pos_weight, 1.0f / (float)nr,
                    c->d_mlp_part.p, c->d_mlp_loss.p + st * tiles_max, s);
    c->mlp_step++;
    const float bc1 = 1.f - std::pow(o.beta1, (float)c->mlp_step), bc2 = 1.f - std::pow(o.beta2, (float)c->mlp_step);
    launch_mlp_reduce_adam(P, c->d_mlp_part.p, (uint32_t)((nr + 31) / 32), nullptr, 1, o.lr, o.beta1, o.beta2, o.eps,
                           bc1, bc2, s);
  }
  HIP_CHECK(hipGetLastError());
  std::vector<float> lp((size_t)steps * tiles_max);
  HIP_CHECK(hipMemcpyAsync(lp.data(), c->d_mlp_loss.p, lp.size() * sizeof(float), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  double tot = 0.0;
  for (float v : lp) tot += v;
  return tot / (double)n;
}

// Training rows of the prefilter: plain mapping with features on; per aligned
// candidate its features, label (log-likelihood within `margin` nats, or DP
// score within `margin` units, of the read's best) and rank.
struct PfRows {
  std::vector<uint16_t> x;
  std::vector<float> y;
  std::vector<uint8_t> rank;
};
static void prefilter_rows(pbw_ctx* c, const pbw_read_batch* in, int label_kind, float margin, PfRows& out) {
  if (label_kind != PBW_LABEL_LOGLK && label_kind != PBW_LABEL_SCORE)
    throw PbwError{PBW_ERR_ARG, "unknown prefilter label kind"};
  hipStream_t s = c->stream;
  const uint64_t bs = c->opt.batch_reads ? c->opt.batch_reads : 131072;
  for (uint64_t b0 = 0; b0 < in->n; b0 += bs) {
    const uint64_t b1 = std::min<uint64_t>(in->n, b0 + bs);
    Batch b;
    batch_upload_slice(c, b, in, b0, b1, false);
    HIP_CHECK(hipMemsetAsync(c->d_ctr.p, 0, sizeof(Counters), s));
    batch_candidates(c, b, true);
    batch_dp(c, b, nullptr, 0);
    const uint32_t B = b.B;
    const size_t ns = (size_t)B * TOP_CANDIDATES;
    std::vector<uint8_t> nj(B);
    std::vector<DpJob> jobs(ns);
    std::vector<DpOut> dout(ns);
    std::vector<uint16_t> feat(ns * MLP_IN);
    HIP_CHECK(hipMemcpyAsync(nj.data(), c->d_njobs.p, B, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipMemcpyAsync(jobs.data(), c->d_jobs.p, ns * sizeof(DpJob), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipMemcpyAsync(dout.data(), c->d_out.p, ns * sizeof(DpOut), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipMemcpyAsync(feat.data(), c->d_feat.p, ns * MLP_IN * 2, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    for (uint32_t r = 0; r < B; r++) {
      const size_t s0 = (size_t)r * TOP_CANDIDATES;
      if (!nj[r]) continue;
      const bool by_score = label_kind == PBW_LABEL_SCORE;
      auto key = [&](uint32_t k) { return by_score ? (double)dout[s0 + k].score : dout[s0 + k].log_lk; };
      double best = key(0);
      for (uint32_t k = 1; k < nj[r]; k++) best = std::max(best, key(k));
      for (uint32_t k = 0; k < nj[r]; k++) {
        const uint32_t cand = jobs[s0 + k].cand;
        out.x.insert(out.x.end(), feat.begin() + (size_t)cand * MLP_IN, feat.begin() + (size_t)(cand + 1) * MLP_IN);
        out.y.push_back(key(k) >= best - (double)margin ? 1.f : 0.f);
        out.rank.push_back((uint8_t)(cand - s0));
      }
    }
  }
  HIP_CHECK(hipMemsetAsync(c->d_ctr.p, 0, sizeof(Counters), s));
  HIP_CHECK(hipStreamSynchronize(s));
}

static void prefilter_train(pbw_ctx* c, const pbw_read_batch* in, const pbw_train_options& o, int label_kind,
                            float margin, float recall, pbw_prefilter_report& rep) {
  PfRows rows;
  prefilter_rows(c, in, label_kind, margin, rows);
  const uint64_t n = rows.y.size();
  if (n < 20) throw PbwError{PBW_ERR_ARG, "too few candidate rows to train the prefilter"};
  const uint64_t nval = n / 10, ntr = n - nval;
  if (!c->mlp_ready) {
    std::vector<float> th;
    mlp_init_host(o.seed, th);
    mlp_set(c, th.data());
  }
  hipStream_t s = c->stream;
  c->d_mlp_x.upload(rows.x.data(), ntr * MLP_IN, s);
  c->d_mlp_y.upload(rows.y.data(), ntr, s);
  for (uint32_t e = 0; e < o.epochs; e++) {
    const double l = mlp_epoch(c, ntr, o, e);
    if (e == 0) rep.loss_first = l;
    rep.loss_last = l;
  }
  // threshold: the (1 - recall) quantile of held-out positives the rank rule does not keep
  c->d_mlp_x.upload(rows.x.data() + ntr * MLP_IN, nval * MLP_IN, s);
  mlp_forward_dev(c, c->d_mlp_x.p, nval);
  std::vector<float> z(nval);
  if (nval) HIP_CHECK(hipMemcpyAsync(z.data(), c->d_mlp_z.p, nval * 4, hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  std::vector<float> pos;
  uint64_t npos = 0;
  for (uint64_t i = 0; i < nval; i++) {
    const bool p = rows.y[ntr + i] > 0.5f;
    npos += p;
    if (p && rows.rank[ntr + i] >= c->pf.top_k) pos.push_back(z[i]);
  }
  std::sort(pos.begin(), pos.end());
  float thr = 3.0e38f;
  if (!pos.empty()) {
    const size_t idx = (size_t)std::floor((1.0 - (double)recall) * (double)pos.size());
    thr = pos[std::min(idx, pos.size() - 1)];
  }
  uint64_t kept_pos = 0, skipped = 0;
  for (uint64_t i = 0; i < nval; i++) {
    const bool keep = rows.rank[ntr + i] < c->pf.top_k || z[i] >= thr;
    if (rows.y[ntr + i] > 0.5f && keep) kept_pos++;
    if (!keep) skipped++;
  }
  uint64_t tot_pos = 0;
  for (float v : rows.y) tot_pos += v > 0.5f;
  rep.rows = n;
  rep.positives = tot_pos;
  rep.val_rows = nval;
  rep.val_positives = npos;
  rep.threshold = thr;
  rep.val_recall = npos ? (double)kept_pos / (double)npos : 1.0;
  rep.val_skip = nval ? (double)skipped / (double)nval : 0.0;
  c->pf.threshold = thr;
}

// ===========================================================================
// C ABI
// ===========================================================================
// Bump allocation over a fixed region; a request past its end is an error.
struct FixedArena {
  uint8_t* base;
  size_t cap, used;
  const char* what;
  template <typename T>
  T* take(size_t count) {
    const size_t off = (used + 255) & ~(size_t)255;
    const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    if (off > cap || bytes > cap - off)
      throw PbwError{PBW_ERR_LIMIT, std::string("the batch needs more than the ") + what};
    used = off + bytes;
    return reinterpret_cast<T*>(base + off);
  }
  size_t left() const {
    const size_t off = (used + 255) & ~(size_t)255;
    return off < cap ? cap - off : 0;
  }
};

extern "C" {

const char* pbw_global_error(void) { return g_error.c_str(); }
const char* pbw_last_error(const pbw_ctx* c) { return c ? c->error.c_str() : g_error.c_str(); }

int pbw_genome_create(pbw_genome** out) {
  if (!out) return PBW_ERR_ARG;
  *out = new pbw_genome();
  return PBW_OK;
}
int pbw_genome_add_contig(pbw_genome* g, const char* name, const uint8_t* seq, uint64_t len) {
  if (!g || !name || (!seq && len)) return PBW_ERR_ARG;
  genome_add(g, name, seq, len);
  return PBW_OK;
}
int pbw_genome_n_contigs(const pbw_genome* g) { return g ? (int)g->lengths.size() : 0; }
const char* pbw_genome_contig_name(const pbw_genome* g, int cid) { return g->names[cid].c_str(); }
uint64_t pbw_genome_contig_length(const pbw_genome* g, int cid) { return g->lengths[cid]; }
void pbw_genome_destroy(pbw_genome* g) { delete g; }

void pbw_default_options(pbw_options* o) {
  o->substitution_rate = 0.001;
  o->insert_size_mean = 250.0;
  o->insert_size_sd = 60.0;
  o->sv_prior_phred = 55.0;
  o->max_indel_len = 30;
  o->gap_open_phred = 40;
  o->gap_extend_phred = 3;
  o->batch_reads = 0;
}

// Contig::slice (genome.rs:47-68) through GenomeIndex::get_sequence (:97-99):
// bases [start, end) of contig cid clamped to its length, non-ACGT as 'N'.
uint64_t pbw_genome_get_sequence(const pbw_genome* g, int cid, uint64_t start, uint64_t end, uint8_t* out) {
  if (!g || cid < 0 || (size_t)cid >= g->lengths.size()) return 0;
  const uint64_t L = g->lengths[cid];
  end = std::min(end, L);
  if (start >= end) return 0;
  static const char BASES[4] = {'A', 'C', 'G', 'T'};
  const uint64_t base = g->coff[cid];
  if (out)
    for (uint64_t p = start; p < end; p++)
      out[p - start] = gisn(g, base + p) ? 'N' : (uint8_t)BASES[gbase2(g, base + p)];
  return end - start;
}

static int ctx_create(const pbw_genome* g, const pbw_options* o, int device, const HostTable* table, pbw_ctx** out) {
  if (!g || !out) return PBW_ERR_ARG;
  pbw_ctx* c = new pbw_ctx();
  c->genome = g;
  c->table_src = table;
  if (o) c->opt = *o;
  else pbw_default_options(&c->opt);
  c->device = device;
  pbw_default_prefilter_options(&c->pf);
  int rc = guarded(c, [&] { ctx_build(c); });
  if (rc != PBW_OK) {
    g_error = c->error;
    if (c->stream) (void)hipStreamDestroy(c->stream);
    if (c->stream2) (void)hipStreamDestroy(c->stream2);
    delete c;
    *out = nullptr;
    return rc;
  }
  c->table_src = nullptr;
  *out = c;
  return PBW_OK;
}

int pbw_ctx_create(const pbw_genome* g, const pbw_options* o, int device, pbw_ctx** out) {
  return ctx_create(g, o, device, nullptr, out);
}

int pbw_ctx_create_from_sthash(const pbw_genome* g, const char* sthash_path, const pbw_options* o, int device,
                               pbw_ctx** out) {
  if (!g || !sthash_path || !out) return PBW_ERR_ARG;
  *out = nullptr;
  HostTable t;
  std::string err;
  if (!read_sthash(sthash_path, t, err)) {
    g_error = err;
    return PBW_ERR_IO;
  }
  return ctx_create(g, o, device, &t, out);
}

// HashTable::save (hashtable.rs:74-80): the context's buckets in code order, each
// keyed by mix(code) (hash_kmer, hashtable.rs:107-113), positions in bucket order.
int pbw_index_save_sthash(pbw_ctx* c, const char* path) {
  if (!c || !path) return PBW_ERR_ARG;
  return guarded(c, [&] {
    HIP_CHECK(hipSetDevice(c->device));
    HostTable t;
    t.k = K;
    t.stride = STRIDE;
    t.total_entries = c->total_entries;
    t.pos.resize(c->total_entries);
    if (c->total_entries)
      HIP_CHECK(hipMemcpy(t.pos.data(), c->d_kpos.p, c->total_entries * sizeof(uint64_t), hipMemcpyDeviceToHost));
    if (c->fused)
      for (uint64_t& e : t.pos) e = fused_to_packed(e);
    const uint64_t CH = 1ull << 24;
    std::vector<uint32_t> ko(CH + 1);
    for (uint64_t c0 = 0; c0 < CODE_SPACE; c0 += CH) {
      HIP_CHECK(hipMemcpy(ko.data(), c->d_koff.p + c0, (CH + 1) * sizeof(uint32_t), hipMemcpyDeviceToHost));
      for (uint64_t j = 0; j < CH; j++)
        if (ko[j + 1] > ko[j]) {
          t.key.push_back(kmer_mix((uint32_t)(c0 + j)));
          t.off.push_back(ko[j]);
          t.cnt.push_back(ko[j + 1] - ko[j]);
        }
    }
    std::string err;
    if (!write_sthash(path, t, err)) throw PbwError{PBW_ERR_IO, err};
  });
}

void pbw_ctx_destroy(pbw_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->stream2) (void)hipStreamSynchronize(c->stream2);
  for (auto& e : c->ev)
    if (e) (void)hipEventDestroy(e);
  for (auto& e : c->dev)
    if (e) (void)hipEventDestroy(e);
  for (auto& e : c->evw)
    if (e) (void)hipEventDestroy(e);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  if (c->stream2) (void)hipStreamDestroy(c->stream2);
  if (c->h_pe) (void)hipHostFree(c->h_pe);
  for (uint8_t* p : c->h_up)
    if (p) (void)hipHostFree(p);
  delete c;
}

int pbw_map_se(pbw_ctx* c, const pbw_read_batch* reads, pbw_record_batch* out) {
  if (!c || !reads || !out) return PBW_ERR_ARG;
  return guarded(c, [&] {
    HIP_CHECK(hipSetDevice(c->device));
    c->stats = pbw_stats{};
    auto t0 = std::chrono::steady_clock::now();
    HostRecords hr;  // reserved while the first batch's seed kernel runs
    hr.expect = reads->n;
    uint64_t bs = c->opt.batch_reads ? c->opt.batch_reads : 131072;
    for (uint64_t b = 0, e; b < reads->n; b = e) {
      e = se_batch_end(c, reads, b, reads->n, bs);
      map_se_batch(c, reads, b, e, hr, false);
    }
    emit_batch(hr, out);
    c->stats.ms_total = ms_since(t0);
  });
}

struct pbw_device_reads {
  int device;
  uint64_t n;
  DevBuf<uint8_t> seq, qual;
  std::vector<uint64_t> seq_off, qual_off;
};

int pbw_reads_upload(pbw_ctx* c, const pbw_read_batch* reads, pbw_device_reads** out) {
  if (!c || !reads || !out) return PBW_ERR_ARG;
  return guarded(c, [&] {
    HIP_CHECK(hipSetDevice(c->device));
    auto* d = new pbw_device_reads();
    d->device = c->device;
    d->n = reads->n;
    d->seq_off.assign(reads->seq_off, reads->seq_off + reads->n + 1);
    d->qual_off.assign(reads->qual_off, reads->qual_off + reads->n + 1);
    try {
      d->seq.upload(reads->seq + reads->seq_off[0], reads->seq_off[reads->n] - reads->seq_off[0], c->stream);
      d->qual.upload(reads->qual + reads->qual_off[0], reads->qual_off[reads->n] - reads->qual_off[0], c->stream);
      HIP_CHECK(hipStreamSynchronize(c->stream));
    } catch (...) {
      delete d;
      throw;
    }
    for (auto& x : d->seq_off) x -= reads->seq_off[0];
    for (auto& x : d->qual_off) x -= reads->qual_off[0];
    *out = d;
  });
}

void pbw_reads_free(pbw_device_reads* d) {
  if (!d) return;
  (void)hipSetDevice(d->device);
  delete d;
}

int pbw_map_se_device(pbw_ctx* c, const pbw_device_reads* d, uint64_t begin, uint64_t end, pbw_record_batch* out) {
  if (!c || !d || !out || begin > end || end > d->n) return PBW_ERR_ARG;
  return guarded(c, [&] {
    HIP_CHECK(hipSetDevice(c->device));
    c->stats = pbw_stats{};
    auto t0 = std::chrono::steady_clock::now();
    HostRecords hr;  // reserved while the first batch's seed kernel runs
    hr.expect = end - begin;
    pbw_read_batch view{d->n, d->seq.p, d->seq_off.data(), d->qual.p, d->qual_off.data()};
    uint64_t bs = c->opt.batch_reads ? c->opt.batch_reads : 131072;
    for (uint64_t b = begin, e; b < end; b = e) {
      e = se_batch_end(c, &view, b, end, bs);
      map_se_batch(c, &view, b, e, hr, true);
    }
    emit_batch(hr, out);
    c->stats.ms_total = ms_since(t0);
  });
}

int pbw_tune_get(const pbw_ctx* c, pbw_tune* out) {
  if (!c || !out) return PBW_ERR_ARG;
  *out = c->tune;
  out->batch_reads = c->opt.batch_reads ? c->opt.batch_reads : 131072;
  if (!out->dp_grid) out->dp_grid = QUAD_MAX_WAVES;
  return PBW_OK;
}

int pbw_tune_set(pbw_ctx* c, const pbw_tune* t) {
  if (!c || !t || t->long_waves > 1 || t->dp_pack16 > 1 || t->split_len > (uint32_t)MAX_READ_LEN) return PBW_ERR_ARG;
  c->tune = *t;
  if (t->batch_reads) c->opt.batch_reads = t->batch_reads;
  return PBW_OK;
}

// Coordinate descent over the launch shapes; every trial maps the same reads, so
// the only difference between trials is time (the records are identical).
int pbw_autotune(pbw_ctx* c, const pbw_device_reads* d, uint64_t begin, uint64_t end, pbw_autotune_report* out) {
  if (!c || !d || !out || begin >= end || end > d->n) return PBW_ERR_ARG;
  return guarded(c, [&] {
    HIP_CHECK(hipSetDevice(c->device));
    const pbw_stats saved = c->stats;
    pbw_read_batch view{d->n, d->seq.p, d->seq_off.data(), d->qual.p, d->qual_off.data()};
    uint32_t max_len = 0;
    for (uint64_t r = begin; r < end; r++) max_len = std::max<uint32_t>(max_len, (uint32_t)(d->seq_off[r + 1] - d->seq_off[r]));
    pbw_autotune_report rep{};
    auto trial = [&](const pbw_tune& t) {
      c->tune = t;
      c->opt.batch_reads = t.batch_reads;
      double best = 1e30;
      for (int rep_i = 0; rep_i < 2; rep_i++) {
        HostRecords hr;
        hr.rec.reserve(end - begin);
        c->stats = pbw_stats{};
        auto t0 = std::chrono::steady_clock::now();
        for (uint64_t b0 = begin; b0 < end; b0 += t.batch_reads)
          map_se_batch(c, &view, b0, std::min<uint64_t>(end, b0 + t.batch_reads), hr, true);
        best = std::min(best, ms_since(t0));
      }
      if (rep.trials < PBW_TUNE_MAX_TRIALS) {
        rep.tried[rep.trials] = t;
        rep.ms[rep.trials] = best;
        rep.trials++;
      }
      return best;
    };
    pbw_tune cur;
    pbw_tune_get(c, &cur);
    double cur_ms = trial(cur);
    auto consider = [&](pbw_tune t) {
      const double ms = trial(t);
      if (ms < cur_ms) {
        cur = t;
        cur_ms = ms;
      }
    };
    const pbw_tune base0 = cur;
    for (uint32_t g : {8192u, 16384u, 32768u, 65536u})
      if (g != base0.dp_grid) consider(pbw_tune{cur.batch_reads, g, cur.split_len, cur.long_waves, cur.dp_pack16});
    consider(pbw_tune{cur.batch_reads, cur.dp_grid, cur.split_len, cur.long_waves, 1u - cur.dp_pack16});
    const pbw_tune base1 = cur;
    for (uint32_t sl : {96u, 128u, 160u, 192u, 256u})
      if (sl < max_len)
        for (uint32_t lw : {0u, 1u}) consider(pbw_tune{base1.batch_reads, base1.dp_grid, sl, lw, base1.dp_pack16});
    const pbw_tune base2 = cur;
    for (uint32_t bs : {65536u, 131072u, 262144u, 524288u})
      if (bs != base2.batch_reads && bs <= std::max<uint64_t>(end - begin, 65536))
        consider(pbw_tune{bs, base2.dp_grid, base2.split_len, base2.long_waves, base2.dp_pack16});
    c->tune = cur;
    c->opt.batch_reads = cur.batch_reads;
    c->stats = saved;
    rep.best = cur;
    rep.best_ms = cur_ms;
    *out = rep;
  });
}

int pbw_map_pe(pbw_ctx* c, const pbw_read_batch* m1, const pbw_read_batch* m2, pbw_insert_model* model,
               pbw_record_batch* out) {
  if (!c || !m1 || !m2 || !model || !out) return PBW_ERR_ARG;
  if (m1->n != m2->n) {
    c->error = "paired-end input files have different read counts (" + std::to_string(m1->n) + " vs " +
               std::to_string(m2->n) + ")";
    return PBW_ERR_MISMATCH;
  }
  return guarded(c, [&] {
    HIP_CHECK(hipSetDevice(c->device));
    c->stats = pbw_stats{};
    auto t0 = std::chrono::steady_clock::now();
    HostRecords hr;
    hr.expect = 2 * m1->n;
    map_pe_all(c, m1, m2, model, hr);
    emit_batch(hr, out);
    c->stats.ms_total = ms_since(t0);
  });
}

int pbw_align_batch(pbw_ctx* c, uint32_t gop, uint32_t gep, const pbw_align_job* jobs, uint64_t n,
                    pbw_align_result* res, uint32_t* cigar, uint64_t cigar_cap) {
  if (!c || (!jobs && n) || (!res && n)) return PBW_ERR_ARG;
  return guarded(c, [&] {
    HIP_CHECK(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    if (n == 0) return;
    if (n > (1u << 30)) throw PbwError{PBW_ERR_LIMIT, "too many jobs"};
    std::vector<uint8_t> seq, qual, ref;
    std::vector<uint64_t> so(n + 1), qo(n + 1);
    std::vector<DpJob> dj(n);
    uint32_t max_m = 0, max_n = 0;
    for (uint64_t i = 0; i < n; i++) {
      const pbw_align_job& j = jobs[i];
      if (j.m > 0x7FFFFFFFu || j.qlen > 0x7FFFFFFFu || j.n > 0x7FFFFFFFu)
        throw PbwError{PBW_ERR_LIMIT, "job beyond 2^31 bases"};
      so[i] = seq.size();
      qo[i] = qual.size();
      seq.insert(seq.end(), j.read, j.read + j.m);
      qual.insert(qual.end(), j.qual, j.qual + j.qlen);
      DpJob d{};
      d.ref_gb = ref.size();
      d.read = (uint32_t)i;
      d.n = j.n;
      d.hint = (int32_t)std::max<int64_t>(std::min<int64_t>(j.hint, INT32_MAX / 2), INT32_MIN / 2);
      d.band = (int32_t)std::min<uint32_t>(j.band, (uint32_t)(INT32_MAX / 4));
      d.rc = 0;
      d.explicit_ref = 1;
      d.cand = (uint32_t)i;
      dj[i] = d;
      ref.insert(ref.end(), j.ref, j.ref + j.n);
      max_m = std::max(max_m, std::max(j.m, j.qlen));
      max_n = std::max(max_n, j.n);
    }
    so[n] = seq.size();
    qo[n] = qual.size();
    c->d_seq.upload(seq.data(), seq.size(), s);
    c->d_qual.upload(qual.data(), qual.size(), s);
    c->d_seq_off.upload(so.data(), n + 1, s);
    c->d_qual_off.upload(qo.data(), n + 1, s);
    c->d_refbytes.upload(ref.data(), ref.size(), s);
    c->d_jobs.upload(dj.data(), n, s);
    DevReads R{c->d_seq.p, c->d_seq_off.p, c->d_qual.p, c->d_qual_off.p, (uint32_t)n};
    // at most 2 m + 1 runs per alignment (every run but deletions takes a read
    // base, and two deletion runs are never adjacent)
    const uint32_t cap = 2 * max_m + 2;
    c->d_out.ensure(n);
    c->d_cigar.ensure(n * cap);
    c->d_flag_a.ensure(n);
    c->d_flag_n.ensure(4);
    HIP_CHECK(hipMemsetAsync(c->d_flag_n.p, 0, 4, s));
    HIP_CHECK(hipMemsetAsync(c->d_ctr.p, 0, sizeof(Counters), s));
    DpArgs da = dp_args_base(c, R);
    da.gap_open = std::min(h_f64_i32(std::round(-90.0 * (double)gop / 40.0)), -1);
    da.gap_extend = std::min(h_f64_i32(std::round(-1.0 * (double)gep / 3.0)), -1);
    da.jobs = c->d_jobs.p;
    da.ref_bytes = c->d_refbytes.p;
    da.cigar_cap = cap;
    da.cigar = c->d_cigar.p;
    da.out = c->d_out.p;
    da.flag_list = c->d_flag_a.p;
    da.flag_n = c->d_flag_n.p;
    // banded kernel for narrow bands and moderate windows, tiled for the rest
    std::vector<uint32_t> band_list, tiled_list;
    uint32_t bm = 0, bn = 0;
    for (uint64_t i = 0; i < n; i++) {
      if (2ull * jobs[i].band + 1 <= 63 && jobs[i].n <= 2048) {
        band_list.push_back((uint32_t)i);
        bm = std::max(bm, std::max(jobs[i].m, jobs[i].qlen));
        bn = std::max(bn, jobs[i].n);
      } else {
        tiled_list.push_back((uint32_t)i);
      }
    }
    if (!band_list.empty()) {
      DevBuf<uint32_t> dl;
      dl.upload(band_list.data(), band_list.size(), s);
      DpArgs db = da;
      db.jlist = dl.p;
      db.njobs = (uint32_t)band_list.size();
      db.max_m = bm;
      db.max_n = bn;
      run_dp_band(c, db);
      HIP_CHECK(hipGetLastError());
      const uint32_t nf = download_u32(c, c->d_flag_n.p);
      std::vector<uint32_t> fl(nf);
      if (nf) HIP_CHECK(hipMemcpyAsync(fl.data(), c->d_flag_a.p, (size_t)nf * 4, hipMemcpyDeviceToHost, s));
      HIP_CHECK(hipStreamSynchronize(s));
      tiled_list.insert(tiled_list.end(), fl.begin(), fl.end());
    }
    da.max_m = max_m;
    da.max_n = max_n;
    run_dp_tiled(c, da, tiled_list);
    std::vector<DpOut> out(n);
    HIP_CHECK(hipMemcpyAsync(out.data(), c->d_out.p, n * sizeof(DpOut), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    uint64_t w = 0;
    for (uint64_t i = 0; i < n; i++) {
      res[i].score = out[i].score;
      res[i].ref_offset = out[i].ref_off;
      res[i].cigar_off = (uint32_t)w;
      res[i].cigar_n = out[i].cigar_n;
      if (cigar && w + out[i].cigar_n <= cigar_cap && out[i].cigar_n)
        HIP_CHECK(hipMemcpyAsync(cigar + w, c->d_cigar.p + (uint64_t)i * cap, out[i].cigar_n * 4,
                                 hipMemcpyDeviceToHost, s));
      w += out[i].cigar_n;
    }
    HIP_CHECK(hipStreamSynchronize(s));
    if (cigar && w > cigar_cap) throw PbwError{PBW_ERR_LIMIT, "cigar buffer too small"};
  });
}

// pbw_align_batch on the fixed device region and host stage: the same kernels,
// grids sized to the region (the DP kernels loop over their job lists).
int pbw_align_batch_noalloc(pbw_ctx* c, uint32_t gop, uint32_t gep, const pbw_align_job* jobs, uint64_t n,
                            pbw_align_result* res, uint32_t* cigar, uint64_t cigar_cap) {
  if (!c || (!jobs && n) || (!res && n)) return PBW_ERR_ARG;
  return guarded(c, [&] {
    HIP_CHECK(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    if (n == 0) return;
    if (n > (1u << 30)) throw PbwError{PBW_ERR_LIMIT, "too many jobs"};
    uint8_t* dev = (c->device >= 0 && c->device < 256) ? g_noalloc_ptr[c->device] : nullptr;
    if (!dev) throw PbwError{PBW_ERR_HIP, "the no-allocation scratch is not set up on this device"};
    std::lock_guard<std::mutex> lk(g_noalloc_mu);  // one host stage: calls run one at a time
    FixedArena D{dev, NOALLOC_SCRATCH_BYTES, 0, "no-allocation device scratch (64 MiB)"};
    FixedArena H{g_noalloc_stage, NOALLOC_STAGE_BYTES, 0, "no-allocation host stage (16 MiB)"};
    uint64_t sm = 0, sq = 0, sr = 0;
    uint32_t max_m = 0, max_n = 0;
    for (uint64_t i = 0; i < n; i++) {
      const pbw_align_job& j = jobs[i];
      // staging in LDS only (no device-memory staging here): reads up to
      // MAX_READ_LEN bases, windows up to 8000
      if (j.m > (uint32_t)MAX_READ_LEN || j.qlen > (uint32_t)MAX_READ_LEN || j.n > 8000)
        throw PbwError{PBW_ERR_LIMIT, "job beyond the no-allocation path's shape (reads of 1024, windows of 8000 bases)"};
      sm += j.m;
      sq += j.qlen;
      sr += j.n;
      max_m = std::max(max_m, std::max(j.m, j.qlen));
      max_n = std::max(max_n, j.n);
    }
    const uint32_t cap = 2 * max_m + 2;  // at most 2 m + 1 runs per alignment
    // host stage
    uint8_t* hseq = H.take<uint8_t>(sm);
    uint8_t* hqual = H.take<uint8_t>(sq);
    uint8_t* href = H.take<uint8_t>(sr);
    uint64_t* hso = H.take<uint64_t>(n + 1);
    uint64_t* hqo = H.take<uint64_t>(n + 1);
    DpJob* hjob = H.take<DpJob>(n);
    uint32_t* hband = H.take<uint32_t>(n);
    uint32_t* htiled = H.take<uint32_t>(n);
    uint32_t* hcnt = H.take<uint32_t>(4);
    DpOut* hout = H.take<DpOut>(n);
    uint64_t os = 0, oq = 0, orf = 0;
    uint32_t nb = 0, nt = 0, bm = 0, bn = 0;
    for (uint64_t i = 0; i < n; i++) {
      const pbw_align_job& j = jobs[i];
      hso[i] = os;
      hqo[i] = oq;
      if (j.m) memcpy(hseq + os, j.read, j.m);
      if (j.qlen) memcpy(hqual + oq, j.qual, j.qlen);
      if (j.n) memcpy(href + orf, j.ref, j.n);
      DpJob d{};
      d.ref_gb = orf;
      d.read = (uint32_t)i;
      d.n = j.n;
      d.hint = (int32_t)std::max<int64_t>(std::min<int64_t>(j.hint, INT32_MAX / 2), INT32_MIN / 2);
      d.band = (int32_t)std::min<uint32_t>(j.band, (uint32_t)(INT32_MAX / 4));
      d.explicit_ref = 1;
      d.cand = (uint32_t)i;
      hjob[i] = d;
      os += j.m;
      oq += j.qlen;
      orf += j.n;
      if (2ull * j.band + 1 <= 63 && j.n <= 2048) {  // banded kernel, as pbw_align_batch
        hband[nb++] = (uint32_t)i;
        bm = std::max(bm, std::max(j.m, j.qlen));
        bn = std::max(bn, j.n);
      } else {
        htiled[nt++] = (uint32_t)i;
      }
    }
    hso[n] = os;
    hqo[n] = oq;
    // device region
    uint8_t* dseq = D.take<uint8_t>(sm);
    uint8_t* dqual = D.take<uint8_t>(sq);
    uint8_t* dref = D.take<uint8_t>(sr);
    uint64_t* dso = D.take<uint64_t>(n + 1);
    uint64_t* dqo = D.take<uint64_t>(n + 1);
    DpJob* djob = D.take<DpJob>(n);
    DpOut* dout = D.take<DpOut>(n);
    uint32_t* dcig = D.take<uint32_t>((size_t)n * cap);
    uint32_t* dband = D.take<uint32_t>(n);
    uint32_t* dtiled = D.take<uint32_t>(n);
    uint32_t* dflag = D.take<uint32_t>(n);
    uint32_t* dflag2 = D.take<uint32_t>(n);
    uint32_t* dcnt = D.take<uint32_t>(4);
    Counters* dctr = D.take<Counters>(1);
    auto up = [&](void* d, const void* h, size_t bytes) {
      if (bytes) HIP_CHECK(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, s));
    };
    up(dseq, hseq, sm);
    up(dqual, hqual, sq);
    up(dref, href, sr);
    up(dso, hso, (n + 1) * 8);
    up(dqo, hqo, (n + 1) * 8);
    up(djob, hjob, n * sizeof(DpJob));
    up(dband, hband, (size_t)nb * 4);
    launch_zero_u32(dcnt, 4, s);
    launch_zero_u32((uint32_t*)dctr, (uint32_t)(sizeof(Counters) / 4), s);
    HIP_CHECK(hipGetLastError());
    DevReads R{dseq, dso, dqual, dqo, (uint32_t)n};
    DpArgs da = dp_args_base(c, R);
    da.ctr = dctr;
    da.gap_open = std::min(h_f64_i32(std::round(-90.0 * (double)gop / 40.0)), -1);
    da.gap_extend = std::min(h_f64_i32(std::round(-1.0 * (double)gep / 3.0)), -1);
    da.jobs = djob;
    da.ref_bytes = dref;
    da.cigar_cap = cap;
    da.cigar = dcig;
    da.out = dout;
    auto count = [&](const uint32_t* d) {
      HIP_CHECK(hipMemcpyAsync(hcnt, d, 4, hipMemcpyDeviceToHost, s));
      HIP_CHECK(hipStreamSynchronize(s));
      return hcnt[0];
    };
    if (nb) {  // jobs the banded kernel flags join the tiled list
      DpArgs db = da;
      db.jlist = dband;
      db.njobs = nb;
      db.max_m = bm;
      db.max_n = bn;
      db.flag_list = dflag;
      db.flag_n = dcnt;
      launch_dp_band(db, s);
      HIP_CHECK(hipGetLastError());
      const uint32_t nf = count(dcnt);
      if (nf) {
        HIP_CHECK(hipMemcpyAsync(htiled + nt, dflag, (size_t)nf * 4, hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipStreamSynchronize(s));
        nt += nf;
      }
    }
    da.max_m = max_m;
    da.max_n = max_n;
    const size_t mark = D.used;
    auto tiled = [&](const uint32_t* list, uint32_t cnt) {
      DpArgs a = da;
      a.jlist = list;
      a.njobs = cnt;
      a.flag_list = dflag;
      a.flag_n = dcnt + 2;
      a.tiles = (a.max_n + 1 + 63) / 64;
      const bool lds_rows = dp_tiled_rows_in_lds(a.max_m, a.max_n);
      const size_t per = (lds_rows ? 0 : 12ull * (a.max_n + 1) + 256) + 32ull * (a.max_m + 1) * a.tiles + 256;
      const uint32_t nwaves = (uint32_t)std::min<size_t>({(size_t)cnt, lds_rows ? 8192u : 2048u, D.left() / per});
      if (!nwaves) throw PbwError{PBW_ERR_LIMIT, "the batch needs more than the no-allocation device scratch (64 MiB)"};
      a.rows = lds_rows ? nullptr : D.take<int32_t>((size_t)nwaves * 3 * (a.max_n + 1));
      a.tb = D.take<uint64_t>((size_t)nwaves * (a.max_m + 1) * a.tiles * 4);
      launch_dp_tiled(a, nwaves, s);
      HIP_CHECK(hipGetLastError());
      if (count(dcnt + 2)) throw PbwError{PBW_ERR_LIMIT, "alignment window beyond the tiled kernel's shape"};
      D.used = mark;
    };
    if (nt) {
      up(dtiled, htiled, (size_t)nt * 4);
      const uint32_t T = dp_wide_tiles(max_m, max_n);
      if (!T) {
        tiled(dtiled, nt);
      } else {  // whole-window rectangles first; the jobs k_dp_wide flags go to the tiled kernel
        DpArgs w = da;
        w.jlist = dtiled;
        w.njobs = nt;
        w.flag_list = dflag2;
        w.flag_n = dcnt + 1;
        const size_t per = 128ull * (max_m + 1) * T + 256;
        const uint32_t nwaves = (uint32_t)std::min<size_t>({(size_t)nt, (size_t)8192, D.left() / per});
        if (!nwaves)
          throw PbwError{PBW_ERR_LIMIT, "the batch needs more than the no-allocation device scratch (64 MiB)"};
        w.tb = D.take<uint64_t>((size_t)nwaves * (max_m + 1) * T * 16);
        launch_dp_wide(w, nwaves, s);
        HIP_CHECK(hipGetLastError());
        D.used = mark;
        const uint32_t nflag = count(dcnt + 1);
        if (nflag) tiled(dflag2, nflag);
      }
    }
    HIP_CHECK(hipMemcpyAsync(hout, dout, n * sizeof(DpOut), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    uint64_t w = 0;
    for (uint64_t i = 0; i < n; i++) {
      res[i].score = hout[i].score;
      res[i].ref_offset = hout[i].ref_off;
      res[i].cigar_off = (uint32_t)w;
      res[i].cigar_n = hout[i].cigar_n;
      if (cigar && w + hout[i].cigar_n <= cigar_cap && hout[i].cigar_n)
        HIP_CHECK(hipMemcpyAsync(cigar + w, dcig + (uint64_t)i * cap, hout[i].cigar_n * 4, hipMemcpyDeviceToHost, s));
      w += hout[i].cigar_n;
    }
    HIP_CHECK(hipStreamSynchronize(s));
    if (cigar && w > cigar_cap) throw PbwError{PBW_ERR_LIMIT, "cigar buffer too small"};
  });
}

int64_t pbw_index_lookup(pbw_ctx* c, uint32_t code, uint64_t* out, uint64_t cap) {
  if (!c || code >= CODE_SPACE) return PBW_ERR_ARG;
  int64_t result = 0;
  int rc = guarded(c, [&] {
    HIP_CHECK(hipSetDevice(c->device));
    uint32_t off[2];
    HIP_CHECK(hipMemcpy(off, c->d_koff.p + code, 8, hipMemcpyDeviceToHost));
    uint64_t n = off[1] - off[0];
    if (out && n) {
      HIP_CHECK(hipMemcpy(out, c->d_kpos.p + off[0], std::min(n, cap) * 8, hipMemcpyDeviceToHost));
      if (c->fused)
        for (uint64_t k = 0; k < std::min(n, cap); k++) out[k] = fused_to_packed(out[k]);
    }
    result = (int64_t)n;
  });
  return rc == PBW_OK ? result : rc;
}
uint64_t pbw_index_total_entries(const pbw_ctx* c) { return c ? c->total_entries : 0; }
const pbw_genome* pbw_ctx_genome(const pbw_ctx* c) { return c ? c->genome : nullptr; }
int pbw_ctx_get_options(const pbw_ctx* c, pbw_options* out) {
  if (!c || !out) return PBW_ERR_ARG;
  *out = c->opt;
  return PBW_OK;
}

void pbw_records_release(pbw_record_batch* b) {
  if (!b) return;
  if (b->rec) (void)hipHostFree(b->rec);
  if (b->cigar) (void)hipHostFree(b->cigar);
  b->rec = nullptr;
  b->cigar = nullptr;
  b->n = b->n_cigar = 0;
}
int pbw_get_stats(const pbw_ctx* c, pbw_stats* out) {
  if (!c || !out) return PBW_ERR_ARG;
  *out = c->stats;
  return PBW_OK;
}


void pbw_default_train_options(pbw_train_options* o) {
  if (!o) return;
  o->epochs = 20;
  o->batch = 4096;
  o->lr = 1e-3f;
  o->beta1 = 0.9f;
  o->beta2 = 0.999f;
  o->eps = 1e-8f;
  o->pos_weight = 1.0f;
  o->seed = 1;
}

void pbw_default_prefilter_options(pbw_prefilter_options* o) {
  if (!o) return;
  o->enabled = 0;
  o->threshold = 0.f;
  o->top_k = 1;
  o->mapq_fallback = 20;
  o->close_nats = 10.f;
  o->audit_period = 64;
}

int pbw_prefilter_set(pbw_ctx* c, const pbw_prefilter_options* o) {
  if (!c || !o) return PBW_ERR_ARG;
  if (o->enabled && !c->mlp_ready) {
    c->error = "prefilter enabled without network parameters";
    return PBW_ERR_ARG;
  }
  c->pf = *o;
  return PBW_OK;
}

int pbw_prefilter_get(const pbw_ctx* c, pbw_prefilter_options* o) {
  if (!c || !o) return PBW_ERR_ARG;
  *o = c->pf;
  return PBW_OK;
}

int pbw_prefilter_dataset(pbw_ctx* c, const pbw_read_batch* reads, int label_kind, float margin, uint16_t* x,
                          float* y, uint8_t* rank, uint64_t cap, uint64_t* n) {
  if (!c || !reads || !n) return PBW_ERR_ARG;
  return guarded(c, [&] {
    PfRows rows;
    prefilter_rows(c, reads, label_kind, margin, rows);
    const uint64_t m = rows.y.size(), w = std::min<uint64_t>(m, cap);
    if (x && w) std::memcpy(x, rows.x.data(), w * MLP_IN * 2);
    if (y && w) std::memcpy(y, rows.y.data(), w * 4);
    if (rank && w) std::memcpy(rank, rows.rank.data(), w);
    *n = m;
  });
}

int pbw_prefilter_train(pbw_ctx* c, const pbw_read_batch* reads, const pbw_train_options* o, int label_kind,
                        float margin, float recall, pbw_prefilter_report* out) {
  if (!c || !reads || !o || !o->batch) return PBW_ERR_ARG;
  return guarded(c, [&] {
    pbw_prefilter_report rep{};
    prefilter_train(c, reads, *o, label_kind, margin, recall, rep);
    if (out) *out = rep;
  });
}

int pbw_mlp_init(pbw_ctx* c, uint64_t seed) {
  if (!c) return PBW_ERR_ARG;
  return guarded(c, [&] {
    std::vector<float> th;
    mlp_init_host(seed, th);
    mlp_set(c, th.data());
  });
}

int pbw_mlp_set_params(pbw_ctx* c, const float* theta, uint64_t n) {
  if (!c || !theta || n != (uint64_t)MLP_NP) return PBW_ERR_ARG;
  return guarded(c, [&] { mlp_set(c, theta); });
}

int pbw_mlp_get_params(pbw_ctx* c, float* theta, uint64_t n) {
  if (!c || !theta || n != (uint64_t)MLP_NP) return PBW_ERR_ARG;
  return guarded(c, [&] {
    mlp_require(c);
    HIP_CHECK(hipMemcpyAsync(theta, c->mlp().theta, MLP_NP * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
  });
}

int pbw_mlp_forward(pbw_ctx* c, const uint16_t* x, uint64_t n, float* logits) {
  if (!c || (n && (!x || !logits))) return PBW_ERR_ARG;
  return guarded(c, [&] {
    mlp_require(c);
    if (!n) return;
    c->d_mlp_x.upload(x, n * MLP_IN, c->stream);
    mlp_forward_dev(c, c->d_mlp_x.p, n);
    HIP_CHECK(hipMemcpyAsync(logits, c->d_mlp_z.p, n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
  });
}

int pbw_mlp_gradient(pbw_ctx* c, const uint16_t* x, const float* y, uint64_t n, float pos_weight, float* grad,
                     double* loss) {
  if (!c || !n || !x || !y || !grad) return PBW_ERR_ARG;
  return guarded(c, [&] {
    mlp_require(c);
    hipStream_t s = c->stream;
    c->d_mlp_x.upload(x, n * MLP_IN, s);
    c->d_mlp_y.upload(y, n, s);
    const uint32_t tiles = (uint32_t)((n + 31) / 32);
    c->d_mlp_part.ensure((size_t)tiles * MLP_NP);
    c->d_mlp_loss.ensure(tiles);
    c->d_mlp_grad.ensure(MLP_NP);
    MlpDev P = c->mlp();
    launch_mlp_grad(P, c->d_mlp_x.p, c->d_mlp_y.p, nullptr, 0, n, n, pos_weight, 1.0f / (float)n, c->d_mlp_part.p,
                    c->d_mlp_loss.p, s);
    launch_mlp_reduce_adam(P, c->d_mlp_part.p, tiles, c->d_mlp_grad.p, 0, 0.f, 0.f, 0.f, 0.f, 1.f, 1.f, s);
    HIP_CHECK(hipGetLastError());
    std::vector<float> lp(tiles);
    HIP_CHECK(hipMemcpyAsync(grad, c->d_mlp_grad.p, MLP_NP * sizeof(float), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipMemcpyAsync(lp.data(), c->d_mlp_loss.p, tiles * sizeof(float), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    if (loss) {
      double t = 0.0;
      for (float v : lp) t += v;
      *loss = t / (double)n;
    }
  });
}

int pbw_mlp_train(pbw_ctx* c, const uint16_t* x, const float* y, uint64_t n, const pbw_train_options* o,
                  double* epoch_loss) {
  if (!c || !n || !x || !y || !o || !o->batch || n > 0xFFFFFFFFull) return PBW_ERR_ARG;
  return guarded(c, [&] {
    mlp_require(c);
    c->d_mlp_x.upload(x, n * MLP_IN, c->stream);
    c->d_mlp_y.upload(y, n, c->stream);
    for (uint32_t e = 0; e < o->epochs; e++) {
      const double l = mlp_epoch(c, n, *o, e);
      if (epoch_loss) epoch_loss[e] = l;
    }
  });
}

}  // extern "C"
