// HIP kernels (gfx950) of the probebwa map path.
//
//   k_seed         seeds.rs:91-151 + hashtable.rs:90-124 — wavefront per read:
//                  15-mer codes in LDS, exact probes of the dense code-indexed
//                  bucket table, the 45 one-mismatch neighbours' presence from
//                  rotated presence bitmaps, lookups of the present ones, a
//                  queue of buckets whose positions are read in full rounds of
//                  64, similarity filter on the 2-bit genome, hits compacted in
//                  emission order with ballots.
//   k_cluster      candidates.rs:30-78 — wavefront per read: bitonic sort of
//                  (group, diagonal) keys in registers, segmented merge,
//                  FxHashMap group order emulation, top-20 selection; a block
//                  per read (k_cluster_tier2) for reads with many hits.
//   k_build_jobs   single.rs:105-133 — candidate windows -> DP jobs.
//   k_dp_band      smith_waterman.rs:163-255 — wavefront per job, one lane per
//                  band column, row state in registers, the deletion recurrence
//                  as an exact max-plus prefix scan across lanes, traceback
//                  bits in LDS, traceback + log-likelihood on lane 0.
//   k_dp_quad      the same recurrence, four jobs per wavefront (one 16-lane
//                  DPP row each, four band columns per lane): the main DP.
//   k_dp_pair16    k_dp_quad with two jobs per row in saturating 16-bit halves
//                  (a tune option).
//   k_dp_wide      windows whose band covers every row (paired-end mate
//                  windows): four columns per lane, previous row in registers.
//   k_dp_tiled     the same recurrence for any band width (64-column tiles,
//                  row state in LDS or global scratch): the remaining shapes.
//   k_mapq         bayesian.rs:142-439 — FP64 select_best + finalize_mapq per read.
//
// All FP64 code is compiled with -ffp-contract=off: the reference (Rust)
// never fuses a*b+c, so neither may this file.
#include "pbw_internal.h"

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <rocprim/block/block_radix_sort.hpp>

namespace pbw {

// ---------------------------------------------------------------------------
// small helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & 63); }
__device__ __forceinline__ int64_t max64(int64_t a, int64_t b) { return a > b ? a : b; }
__device__ __forceinline__ int64_t min64(int64_t a, int64_t b) { return a < b ? a : b; }
// (score, j, state) endpoint key ordered by score, then j, then state
__device__ __forceinline__ int64_t end_key(int32_t v, uint32_t low) {
  return (int64_t)(((uint64_t)(uint32_t)v << 32) | (uint64_t)low);
}

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Orders a wave's global-memory stores before later loads of the same data by
// other lanes of the wave (a block of one wave: workgroup scope suffices).
__device__ __forceinline__ void mem_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

__device__ __forceinline__ uint64_t lanemask_lt() { return (1ull << lane_id()) - 1ull; }

__device__ __forceinline__ int32_t shfl_up_i32(int32_t v, int d) { return __shfl_up(v, (unsigned)d, 64); }
__device__ __forceinline__ int32_t shfl_down_i32(int32_t v, int d) { return __shfl_down(v, (unsigned)d, 64); }
__device__ __forceinline__ uint32_t shfl_u32(uint32_t v, int src) { return (uint32_t)__shfl((int)v, src, 64); }
__device__ __forceinline__ uint64_t shfl_u64(uint64_t v, int src) {
  uint32_t lo = shfl_u32((uint32_t)v, src), hi = shfl_u32((uint32_t)(v >> 32), src);
  return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ uint64_t shfl_up_u64(uint64_t v, int d) {
  const uint32_t lo = (uint32_t)__shfl_up((int)(uint32_t)v, (unsigned)d, 64);
  const uint32_t hi = (uint32_t)__shfl_up((int)(uint32_t)(v >> 32), (unsigned)d, 64);
  return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ int64_t shfl_xor_i64(int64_t v, int m) {
  int lo = __shfl_xor((int)(uint32_t)v, m, 64), hi = __shfl_xor((int)(uint32_t)((uint64_t)v >> 32), m, 64);
  return (int64_t)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}

// DPP cross-lane moves (GFX9 encodings; all VALU, no LDS traffic).
// wave_shr:1 -> lane i reads lane i-1; wave_shl:1 -> lane i reads lane i+1;
// lanes without a source keep `old`.
__device__ __forceinline__ int32_t dpp_wave_shr1(int32_t old, int32_t x) {
  return __builtin_amdgcn_update_dpp(old, x, 0x138, 0xf, 0xf, false);
}
__device__ __forceinline__ int32_t dpp_wave_shl1(int32_t old, int32_t x) {
  return __builtin_amdgcn_update_dpp(old, x, 0x130, 0xf, 0xf, false);
}
__device__ __forceinline__ int32_t readlane_i32(int32_t x, int lane) { return __builtin_amdgcn_readlane(x, lane); }

// Inclusive prefix sum over the wave: row_shr 1,2,4,8 inside rows of 16,
// then row_bcast 15 / 31 across rows.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x) {
  int32_t v = (int32_t)x;
  v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false);
  return (uint32_t)v;
}
// Inclusive prefix max over the wave (same pattern, identity INT32_MIN).
// The group of NL lanes (NL a power of two, 8..64, aligned in the wave) that
// holds one read: every lane of it active, in the same control flow.
template <int NL>
__device__ __forceinline__ uint32_t group_sum_u32(uint32_t x) {
#pragma unroll
  for (int o = NL / 2; o > 0; o >>= 1) x += (uint32_t)__shfl_xor((int)x, o, 64);
  return x;
}

__device__ __forceinline__ int32_t wave_incl_max(int32_t x) {
  const int32_t I = INT32_MIN;
  x = max(x, __builtin_amdgcn_update_dpp(I, x, 0x111, 0xf, 0xf, false));
  x = max(x, __builtin_amdgcn_update_dpp(I, x, 0x112, 0xf, 0xf, false));
  x = max(x, __builtin_amdgcn_update_dpp(I, x, 0x114, 0xf, 0xf, false));
  x = max(x, __builtin_amdgcn_update_dpp(I, x, 0x118, 0xf, 0xf, false));
  x = max(x, __builtin_amdgcn_update_dpp(I, x, 0x142, 0xa, 0xf, false));
  x = max(x, __builtin_amdgcn_update_dpp(I, x, 0x143, 0xc, 0xf, false));
  return x;
}
__device__ __forceinline__ int32_t wave_suffix_min(int32_t x) {
  int l = lane_id();
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    int32_t t = shfl_down_i32(x, o);
    if (l + o < 64) x = min(x, t);
  }
  return x;
}
__device__ __forceinline__ int64_t wave_max_i64(int64_t x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    int64_t t = shfl_xor_i64(x, o);
    x = x > t ? x : t;
  }
  return x;
}

__device__ __forceinline__ uint32_t base2bits(uint8_t b) {
  switch (b) {
    case 'A': case 'a': return 0;
    case 'C': case 'c': return 1;
    case 'G': case 'g': return 2;
    case 'T': case 't': return 3;
    default: return 0;
  }
}
__device__ __forceinline__ bool is_acgt(uint8_t b) {
  return b == 'A' || b == 'C' || b == 'G' || b == 'T' || b == 'a' || b == 'c' || b == 'g' || b == 't';
}
// types.rs:34-47 reverse_complement base map
__device__ __forceinline__ uint8_t comp_base(uint8_t b) {
  switch (b) {
    case 'A': return 'T';
    case 'T': return 'A';
    case 'C': return 'G';
    case 'G': return 'C';
    default: return 'N';
  }
}
// genome.rs:48-68 Contig::slice, one base: 2-bit code with the N-run overlay
__device__ __forceinline__ uint8_t genome_base(const DevGenome& G, uint64_t gb) {
  uint32_t c = (uint32_t)(G.g2b[gb >> 5] >> ((gb & 31) * 2)) & 3u;
  bool isn = (G.nmask[gb >> 6] >> (gb & 63)) & 1ull;
  return isn ? (uint8_t)'N' : (uint8_t)("ACGT"[c]);
}

// Rotation of a 30-bit 15-mer code (index of the rotated presence bitmaps).
__host__ __device__ __forceinline__ uint32_t rotr30(uint32_t x, uint32_t r) {
  return r ? (((x >> r) | (x << (30 - r))) & 0x3FFFFFFFu) : x;
}

// ---------------------------------------------------------------------------
// seed kernel: seeds.rs:91-151
// ---------------------------------------------------------------------------
constexpr int SEED_WPB = 4;

// Pending-bucket queue: fewer than 64 positions are carried from one lookup
// round to the next, so at most 63 buckets wait when a round appends its 64.
constexpr int SEED_QCAP = 128;

struct SeedLds {
  uint8_t seq[MAX_READ_LEN];
  uint32_t wcode[64], woff[64], wcnt[64], lbase[64];
  uint16_t fpF[64], fpR[64];
  uint64_t wmask[64];  // per window: present neighbours (bit 3i + alt)
  uint8_t wflags[64];  // 1 valid, 2 probed, 4 fullF, 8 fullR, 16 exact-repetitive
  uint8_t nbrep[64];
  // looked-up buckets holding at least one position, first in first out in
  // emission order; q_incl is the inclusive prefix of their sizes
  uint32_t q_incl[SEED_QCAP], q_off[SEED_QCAP], q_fp[SEED_QCAP];  // q_fp: fpF | fpR << 16
  uint32_t q_pos[SEED_QCAP];                                        // window start in the read
  uint8_t q_fl[SEED_QCAP];   // 1 neighbour probe, 4 fullF, 8 fullR
  uint8_t q_hfl[SEED_QCAP];  // prefilter flags of the hits from this bucket
};

// seeds.rs:276-286
__device__ __forceinline__ bool neighbour_probed(uint32_t p, uint32_t L) {
  if (L <= 34) return true;
  if (L <= 49) return ((p / 5) % 2) == 0;
  return (p % 15) < 5;
}

// Base-composition counts (A,C,G,T; others ignored) packed 4 bits each.
__device__ __forceinline__ uint32_t fp_count(uint8_t b, bool complement) {
  int c;
  switch (b) {
    case 'A': c = 0; break;
    case 'C': c = 1; break;
    case 'G': c = 2; break;
    case 'T': c = 3; break;
    default: return 0;
  }
  if (complement) c = 3 - c;
  return 1u << (4 * c);
}

// seeds.rs:162-219 for the only case that can reject: both flanks complete on
// both sides (after_len + before_len == 12); otherwise the hit passes.  The
// genome side is the entry's stored flank fingerprint (the counts of the 4
// bases before and the 8 after the 15-mer, N not counted, written by the index
// build), read with the position itself instead of from the 2-bit genome.
__device__ __forceinline__ bool seed_filter(uint16_t gfp, bool reverse, bool neighbour, uint8_t wflags, uint16_t fpF,
                                            uint16_t fpR) {
  bool rfull = reverse ? (wflags & 8) : (wflags & 4);
  if (!rfull || gfp == GENOME_FP_OPEN) return true;
  uint16_t rf = reverse ? fpR : fpF;
  int dist = 0;
#pragma unroll
  for (int c = 0; c < 4; c++) dist += abs((int)((rf >> (4 * c)) & 15) - (int)((gfp >> (4 * c)) & 15));
  return dist <= (neighbour ? 8 : 10);
}

// Which of the 45 single-substitution neighbours of a 15-mer have a non-empty
// bucket: bit 3i + alt for position i and the alt-th other base, the lookup
// order of seeds.rs:291-309.  The five positions of a group (10..14, 5..9,
// 0..4) differ in bits 0..9 of the presence copy rotated by 0, 10 or 20 bits,
// so a group reads its code's word (bits 0..5) and three words for each of
// the two positions in bits 6..9: 21 words per window, all issued together.
__device__ __forceinline__ uint64_t neighbour_presence(const DevIndex& X, uint32_t code) {
  uint64_t mask = 0;
#pragma unroll 1
  for (int g = 0; g < 3; g++) {  // one group's seven words in flight at a time
    const uint32_t rg = 10u * (uint32_t)g;
    const uint64_t* pb = g == 0 ? X.pres : (g == 1 ? X.pres_r10 : X.pres_r20);
    const uint32_t base = rotr30(code, rg), W = base >> 6, off = base & 63u;
    const uint32_t o3 = (base >> 6) & 3u, o4 = (base >> 8) & 3u;
    const uint64_t w0 = pb[W];
    uint64_t w3[3], w4[3];
#pragma unroll
    for (uint32_t alt = 0; alt < 3; alt++) {
      w3[alt] = pb[(W & ~3u) | (alt < o3 ? alt : alt + 1)];
      w4[alt] = pb[(W & ~12u) | ((alt < o4 ? alt : alt + 1) << 2)];
    }
#pragma unroll
    for (uint32_t k = 0; k < 5; k++) {
      const uint32_t i = 14u - rg / 2u - k, ls = 2u * k, orig = (base >> ls) & 3u;
#pragma unroll
      for (uint32_t alt = 0; alt < 3; alt++) {
        const uint32_t nb = alt < orig ? alt : alt + 1;
        uint64_t bit;
        if (k < 3) bit = (w0 >> ((off & ~(3u << ls)) | (nb << ls))) & 1ull;
        else if (k == 3) bit = (w3[alt] >> off) & 1ull;
        else bit = (w4[alt] >> off) & 1ull;
        mask |= bit << (3u * i + alt);
      }
    }
  }
  return mask;
}

// Position of the k-th (from 0) set bit of m; m has more than k set bits.
__device__ __forceinline__ uint32_t select_set_bit(uint64_t m, uint32_t k) {
  uint32_t pos = 0, c;
  c = (uint32_t)__popcll(m & 0xFFFFFFFFull);
  if (k >= c) { k -= c; m >>= 32; pos += 32; }
  c = (uint32_t)__popcll(m & 0xFFFFull);
  if (k >= c) { k -= c; m >>= 16; pos += 16; }
  c = (uint32_t)__popcll(m & 0xFFull);
  if (k >= c) { k -= c; m >>= 8; pos += 8; }
  c = (uint32_t)__popcll(m & 0xFull);
  if (k >= c) { k -= c; m >>= 4; pos += 4; }
  c = (uint32_t)__popcll(m & 0x3ull);
  if (k >= c) { k -= c; m >>= 2; pos += 2; }
  if (k >= (uint32_t)(m & 1ull)) pos += 1;
  return pos;
}

// Where a read's hits go: a slot of `cap` hits from hbase on (fixed or exact
// slots), or, with a chunk table, 1024-hit chunks taken from a pool as the hits
// come (nch chunks held so far, `last` the id of the newest; `dropped` once a
// chunk could not be had: the hits are then only counted).
struct HitSlots {
  uint64_t hbase;
  uint32_t cap;
  uint32_t nch, last, first;
  bool dropped;
};

// One round of the 64 queued positions from queue position qc on: flank
// filter, candidate key and compaction of the passing hits in emission order
// (seeds.rs:120-147, candidates.rs:37-40).
__device__ __forceinline__ void seed_positions(const SeedArgs& a, const SeedLds& S, uint32_t qn, uint32_t qc,
                                               uint32_t qtot, uint32_t r, HitSlots& hs, uint32_t& hcount,
                                               unsigned long long& c_filt) {
  const uint32_t e = qc + lane_id();
  bool pass = false;
  uint64_t key = 0;
  uint32_t hfl = 0;
  if (e < qtot) {
    int lo = 0, hi = (int)qn - 1;  // first bucket l with q_incl[l] > e
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (S.q_incl[mid] > e) hi = mid;
      else lo = mid + 1;
    }
    const uint32_t excl = lo ? S.q_incl[lo - 1] : 0u;
    const uint64_t entry = (uint64_t)S.q_off[lo] + (e - excl);
    const uint64_t packed = a.X.kpos[entry];
    const bool rev = (packed >> 63) & 1ull;
    uint16_t gfp;
    uint32_t contig;
    uint64_t pos;
    if (a.X.kfp) {
      gfp = a.X.kfp[entry];
      contig = (uint32_t)((packed >> 40) & ((1ull << 23) - 1));
      pos = packed & ((1ull << 40) - 1);
    } else {  // fused entry
      gfp = (uint16_t)(packed >> 32);
      contig = (uint32_t)(packed >> 48) & 0x7FFFu;
      pos = packed & 0xFFFFFFFFull;
    }
    const uint32_t fl = S.q_fl[lo], fp = S.q_fp[lo];
    pass = seed_filter(gfp, rev, (fl & 1u) != 0, (uint8_t)fl, (uint16_t)fp, (uint16_t)(fp >> 16));
    const int64_t rp = (int64_t)S.q_pos[lo];
    const int64_t diag = rev ? (int64_t)pos + rp : (int64_t)pos - rp;
    key = ((uint64_t)(contig * 2u + (rev ? 1u : 0u)) << DIAG_BITS) | (uint64_t)(diag + DIAG_BIAS);
    hfl = S.q_hfl[lo];
  }
  const uint64_t m = __ballot(pass);
  const uint32_t pre = (uint32_t)__popcll(m & lanemask_lt());
  const uint32_t np = (uint32_t)__popcll(m);
  if (a.chunk_tab) {
    if (np && !hs.dropped) {
      // chunks up to the one holding hit hcount + np - 1 (one or two new at most)
      const uint32_t j1 = (hcount + np - 1) >> SEED_CH_BITS, nch0 = hs.nch, last0 = hs.last;
      uint32_t got = 0;
      if (j1 >= nch0) {
        // from the read's sub-pool (SEED_SUBPOOLS counters on lines of their own)
        const uint32_t need = j1 + 1 - nch0, sub = r % SEED_SUBPOOLS, per = a.pool_chunks / SEED_SUBPOOLS;
        if (lane_id() == 0 && j1 < (uint32_t)SEED_MAXCH) got = atomicAdd(a.pool_next + sub * 16, need);
        got = (uint32_t)__builtin_amdgcn_readfirstlane((int)got);
        if (j1 >= (uint32_t)SEED_MAXCH || got + need > per) {
          hs.dropped = true;
        } else {
          got += sub * per;
          if ((uint32_t)lane_id() < need) a.chunk_tab[(uint64_t)r * SEED_MAXCH + nch0 + lane_id()] = got + lane_id();
          if (nch0 == 0) hs.first = got;
          hs.nch = j1 + 1;
          hs.last = got + need - 1;
        }
      }
      if (pass && !hs.dropped) {
        const uint32_t h = hcount + pre, j = h >> SEED_CH_BITS;
        const uint32_t id = j < nch0 ? last0 : got + (j - nch0);
        const uint64_t at = ((uint64_t)id << SEED_CH_BITS) + (h & ((1u << SEED_CH_BITS) - 1));
        a.hits[at] = key;
        if (a.hflag) a.hflag[at] = (uint8_t)hfl;
      }
    }
  } else if (pass && hcount + pre < hs.cap) {
    a.hits[hs.hbase + hcount + pre] = key;
    if (a.hflag) a.hflag[hs.hbase + hcount + pre] = (uint8_t)hfl;
  }
  hcount += np;
  c_filt += pass ? 1 : 0;
}

// LONG = false: reads of at most MAX_READ_LEN bases, staged in LDS; LONG =
// true: the longer ones, read from device memory (the two launches split a
// list between them).
template <bool LONG>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6))) void k_seed(SeedArgs a) {
  __shared__ SeedLds lds_all[SEED_WPB];
  const int wib = threadIdx.x >> 6, lane = lane_id();
  SeedLds& S = lds_all[wib];
  const uint32_t gw = blockIdx.x * SEED_WPB + wib, nwaves = gridDim.x * SEED_WPB;
  unsigned long long c_look = 0, c_pos = 0, c_filt = 0;
  for (uint32_t li = gw; li < a.nlist; li += nwaves) {
    const uint32_t r = a.rlist ? a.rlist[li] : li;
    const uint64_t so = a.R.seq_off[r];
    const uint32_t L = (uint32_t)(a.R.seq_off[r + 1] - so);
    const uint64_t qo = a.R.qual_off[r];
    const uint32_t qlen = (uint32_t)(a.R.qual_off[r + 1] - qo);
    (void)qlen;
    HitSlots hs{};
    if (!a.chunk_tab) {
      hs.hbase = a.hoff[li];
      hs.cap = a.hcap ? a.hcap[li] : a.hit_cap;
    }
    uint32_t hcount = 0;
    uint32_t qn = 0, qtot = 0, qc = 0;  // queued buckets, their positions, positions consumed
    if (LONG != (L > (uint32_t)MAX_READ_LEN)) continue;  // the other launch's read
    if (L < (uint32_t)K) {
      if (lane == 0) {
        a.nhits[r] = 0;
        if (a.chunk_tab) {
          a.hoff[r] = 0;
          a.dropped[r] = 0;
        } else if (a.mark_tab) {
          a.mark_tab[(uint64_t)r * SEED_MAXCH] = SEED_CONTIGUOUS;
        }
      }
      continue;
    }
    const uint8_t* rs = a.R.seq + so;
    auto base_at = [&](uint32_t x) -> uint8_t { return LONG ? rs[x] : S.seq[x]; };
    if (!LONG) {
      for (uint32_t x = lane; x < L; x += 64) S.seq[x] = rs[x];
      wave_sync();
    }
    const uint32_t nwin = L - K + 1;
    for (uint32_t cb = 0; cb < nwin; cb += 64) {
      const uint32_t p = cb + lane;
      const bool act = p < nwin;
      uint32_t code = 0, off = 0, cnt = 0, nlook = 0;
      uint64_t nbm = 0;
      uint8_t fl = 0;
      uint16_t fpF = 0, fpR = 0;
      if (act) {
        bool valid = true;
        for (int t = 0; t < K; t++) {
          uint8_t b = base_at(p + t);
          code = (code << 2) | base2bits(b);
          valid &= is_acgt(b);
        }
        if (valid) {
          fl |= 1;
          off = a.X.koff[code];  // hashtable.rs:90-102 lookup
          cnt = a.X.koff[code + 1] - off;
          if (cnt >= (uint32_t)REPEAT_SKIP) fl |= 16;
          else if (neighbour_probed(p, L)) fl |= 2;
          // the reference looks up the window and, when probed, its 45 neighbours;
          // only the neighbours with a non-empty bucket are enumerated below
          c_look += 1 + ((fl & 2) ? 45u : 0u);
          if (fl & 2) nbm = neighbour_presence(a.X, code);
          nlook = 1 + (uint32_t)__popcll(nbm);
        }
        // read-side fingerprints (seeds.rs:181-211) for complete flanks only
        if (p >= 4 && p + 23 <= L) {
          fl |= 4;
          uint32_t f = 0;
          for (uint32_t t = p + 15; t < p + 23; t++) f += fp_count(base_at(t), false);
          for (uint32_t t = p - 4; t < p; t++) f += fp_count(base_at(t), false);
          fpF = (uint16_t)f;
        }
        if (p >= 8 && p + 19 <= L) {  // rc(read): window start L-p-15
          fl |= 8;
          uint32_t f = 0;
          for (uint32_t t = p - 8; t < p; t++) f += fp_count(base_at(t), true);
          for (uint32_t t = p + 15; t < p + 19; t++) f += fp_count(base_at(t), true);
          fpR = (uint16_t)f;
        }
      }
      S.wcode[lane] = code;
      S.woff[lane] = off;
      S.wcnt[lane] = cnt;
      S.wflags[lane] = fl;
      S.wmask[lane] = nbm;
      S.fpF[lane] = fpF;
      S.fpR[lane] = fpR;
      S.nbrep[lane] = 0;
      uint32_t incl = wave_incl_scan(nlook);
      S.lbase[lane] = incl - nlook;
      const uint32_t T = (uint32_t)readlane_i32((int32_t)incl, 63);
      wave_sync();
      // lookups of this chunk in emission order: window, exact, then the
      // neighbours with a non-empty bucket in (position, base) order
      // (seeds.rs:107-147, :291-309)
      for (uint32_t ub = 0; ub < T; ub += 64) {
        const uint32_t u = ub + lane;
        const bool act2 = u < T;
        uint32_t w = 0, j = 0, loff = 0, lcnt = 0;
        if (act2) {
          int lo = 0, hi = 63;  // largest w with lbase[w] <= u
          while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if (S.lbase[mid] <= u) lo = mid;
            else hi = mid - 1;
          }
          w = (uint32_t)lo;
          j = u - S.lbase[w];
          if (j == 0) {
            loff = S.woff[w];
            lcnt = S.wcnt[w];
          } else {
            const uint32_t wc = S.wcode[w];
            const uint32_t t = select_set_bit(S.wmask[w], j - 1);  // neighbour 3i + alt
            const uint32_t i = t / 3, alt = t % 3;
            const uint32_t sh = 2 * (14 - i);
            const uint32_t orig = (wc >> sh) & 3u;
            const uint32_t nbase = alt < orig ? alt : alt + 1;
            const uint32_t ncode = (wc & ~(3u << sh)) | (nbase << sh);
            loff = a.X.koff[ncode];
            lcnt = a.X.koff[ncode + 1] - loff;
            if (lcnt >= (uint32_t)REPEAT_SKIP) S.nbrep[w] = 1;
          }
          c_pos += lcnt;
        }
        // append the buckets with positions to the queue in emission order
        const uint32_t pincl = wave_incl_scan(lcnt);
        const uint32_t PT = (uint32_t)readlane_i32((int32_t)pincl, 63);
        const uint64_t qm = __ballot(lcnt != 0);
        if (lcnt) {
          const uint32_t qi = qn + (uint32_t)__popcll(qm & lanemask_lt());
          S.q_incl[qi] = qtot + pincl;
          S.q_off[qi] = loff;
          S.q_fp[qi] = (uint32_t)S.fpF[w] | ((uint32_t)S.fpR[w] << 16);
          S.q_pos[qi] = cb + w;
          S.q_fl[qi] = (uint8_t)((S.wflags[w] & 12u) | (j != 0 ? 1u : 0u));
          // prefilter flags: neighbour probe, bit length of the bucket size
          S.q_hfl[qi] = (uint8_t)((j != 0 ? 0x80u : 0u) | (uint32_t)(32 - __builtin_clz(lcnt)));
        }
        qn += (uint32_t)__popcll(qm);
        qtot += PT;
        wave_sync();
        if (qtot - qc >= 64) {
          do {
            seed_positions(a, S, qn, qc, qtot, r, hs, hcount, c_filt);
            qc += 64;
          } while (qtot - qc >= 64);
          // drop the buckets consumed whole and rebase the prefix sums; fewer
          // than 64 positions remain, so fewer than 64 buckets move
          const uint64_t d0 = __ballot(lane < qn && S.q_incl[lane] <= qc);
          const uint64_t d1 = __ballot(lane + 64 < qn && S.q_incl[lane + 64] <= qc);
          const uint32_t l0 = (uint32_t)(__popcll(d0) + __popcll(d1));
          const uint32_t base = l0 ? S.q_incl[l0 - 1] : 0u;
          const bool mv = l0 + lane < qn;
          uint32_t vi = 0, vo = 0, vf = 0;
          uint32_t vp = 0;
          uint8_t vfl = 0, vh = 0;
          if (mv) {
            const uint32_t sidx = l0 + lane;
            vi = S.q_incl[sidx];
            vo = S.q_off[sidx];
            vf = S.q_fp[sidx];
            vp = S.q_pos[sidx];
            vfl = S.q_fl[sidx];
            vh = S.q_hfl[sidx];
          }
          wave_sync();
          if (mv) {
            S.q_incl[lane] = vi - base;
            S.q_off[lane] = vo;
            S.q_fp[lane] = vf;
            S.q_pos[lane] = vp;
            S.q_fl[lane] = vfl;
            S.q_hfl[lane] = vh;
          }
          qn -= l0;
          qtot -= base;
          qc -= base;
          wave_sync();
        }
      }
      if (act) {
        uint8_t rc = 0;
        if (fl & 16) rc = 1;
        else if (S.nbrep[lane]) rc = 2;
        a.rcode[so + p] = rc;
      }
      wave_sync();
    }
    // the positions still queued at the end of the read
    for (; qc < qtot; qc += 64) seed_positions(a, S, qn, qc, qtot, r, hs, hcount, c_filt);
    if (lane == 0) {
      a.nhits[r] = hcount;
      if (a.chunk_tab) {  // chunk 0's first slot: reads of up to 1024 hits are contiguous there
        a.hoff[r] = hs.nch ? ((uint64_t)hs.first << SEED_CH_BITS) : 0ull;
        a.dropped[r] = hs.dropped ? 1 : 0;
      } else if (a.mark_tab) {  // an exact slot: contiguous at hoff
        a.mark_tab[(uint64_t)r * SEED_MAXCH] = SEED_CONTIGUOUS;
      }
    }
  }
  // per-wave counter reduction
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    c_look += (unsigned long long)__shfl_xor((long long)c_look, o, 64);
    c_pos += (unsigned long long)__shfl_xor((long long)c_pos, o, 64);
    c_filt += (unsigned long long)__shfl_xor((long long)c_filt, o, 64);
  }
  if (lane == 0 && a.ctr) {
    atomicAdd(&a.ctr->lookups, c_look);
    atomicAdd(&a.ctr->positions, c_pos);
    atomicAdd(&a.ctr->filtered, c_filt);
  }
}

void launch_seed(const SeedArgs& a, hipStream_t s) {
  if (a.nlist == 0) return;
  uint32_t blocks = (a.nlist + SEED_WPB - 1) / SEED_WPB;
  if (blocks > 16384) blocks = 16384;
  hipLaunchKernelGGL(k_seed<false>, dim3(blocks), dim3(64 * SEED_WPB), 0, s, a);
  if (a.has_long) hipLaunchKernelGGL(k_seed<true>, dim3(blocks), dim3(64 * SEED_WPB), 0, s, a);
}

// ---------------------------------------------------------------------------
// cluster kernel: candidates.rs:30-78
// ---------------------------------------------------------------------------
// FxHasher (rustc-hash 2.x) of the key (contig, strand): add-then-multiply by
// K per word, finish with rotate_left(26).
__device__ __forceinline__ uint64_t fx_hash_group(uint64_t contig, uint64_t disc) {
  const uint64_t KK = 0xf1357aea2e62a9c5ull;
  uint64_t h = 0;
  h = (h + contig) * KK;
  h = (h + disc) * KK;
  return (h << 26) | (h >> 38);
}
__device__ __forceinline__ uint32_t sw_capacity(uint32_t buckets) {
  if (buckets == 0) return 0;
  uint32_t mask = buckets - 1;
  return mask < 8 ? mask : (buckets / 8) * 7;
}
__device__ __forceinline__ uint32_t sw_buckets_for(uint32_t cap) {
  if (cap < 8) return cap < 4 ? 4 : 8;
  uint32_t adj = cap * 8 / 7, p = 1;
  while (p < adj) p <<= 1;
  return p;
}
// SwissTable insert slot (16-wide groups, triangular probing, small-table fix)
__device__ uint32_t sw_find_slot(const int32_t* tab, uint32_t nb, uint64_t h) {
  uint32_t mask = nb - 1, pos = (uint32_t)h & mask, stride = 0;
  for (;;) {
    for (uint32_t b = 0; b < 16; b++) {
      uint32_t idx = pos + b;
      bool empty;
      if (idx < nb) empty = tab[idx] < 0;
      else if (nb < 16) empty = idx < 16 ? true : tab[idx - 16] < 0;
      else empty = tab[idx - nb] < 0;
      if (empty) {
        uint32_t index = (pos + b) & mask;
        if (tab[index] >= 0) {
          for (uint32_t c = 0; c < 16; c++) {
            bool e2 = c < nb ? tab[c] < 0 : true;
            if (e2) return c;
          }
        }
        return index;
      }
    }
    stride += 16;
    pos = (pos + stride) & mask;
  }
}

// ---- emulated FxHashMap iteration order of the (contig, strand) groups ----
// Groups are inserted in first-appearance order (candidates.rs:36-43); the
// map iterates in ascending bucket index.

// General path (any number of groups): one thread walks an explicit table.
// g_sort[t] = group inserted t-th; writes g_rank[group].
__device__ void group_ranks_table(uint32_t G, const uint32_t* g_key, const uint32_t* g_sort, uint32_t* g_rank,
                                  int32_t* table, int32_t* tmp) {
  uint32_t nb = 0, items = 0, growth_left = 0;
  for (uint32_t t = 0; t < G; t++) {
    const uint32_t gi = g_sort[t];
    const uint64_t h = fx_hash_group(g_key[gi] >> 1, g_key[gi] & 1u);
    if (growth_left == 0) {
      const uint32_t nnb = sw_buckets_for(max(items + 1, sw_capacity(nb) + 1));
      uint32_t cnt = 0;
      for (uint32_t b = 0; b < nb; b++)
        if (table[b] >= 0) tmp[cnt++] = table[b];
      for (uint32_t b = 0; b < nnb; b++) table[b] = -1;
      nb = nnb;
      for (uint32_t c = 0; c < cnt; c++) {
        const int32_t gj = tmp[c];
        table[sw_find_slot(table, nb, fx_hash_group(g_key[gj] >> 1, g_key[gj] & 1u))] = gj;
      }
      growth_left = sw_capacity(nb) - items;
    }
    table[sw_find_slot(table, nb, h)] = (int32_t)gi;
    growth_left--;
    items++;
  }
  uint32_t rank = 0;
  for (uint32_t b = 0; b < nb; b++)
    if (table[b] >= 0) g_rank[table[b]] = rank++;
}

// Insert slot of sw_find_slot for a table of at most 64 buckets whose
// occupancy is the bit mask `occ` (wave-uniform).  A probe group's 16 control
// positions pos..pos+15 are a 16-bit window of the occupancy: rotated within
// the nb buckets when nb >= 16 (positions past the end mirror the first
// buckets), and for nb < 16 the mask followed by empty padding up to 16 and
// a copy of the mask.  The first empty position of the window is the slot;
// in a small table it can fall on the copy of an occupied bucket, and the
// slot is then the first empty position of group 0.
__device__ __forceinline__ uint32_t sw_find_slot_mask(uint64_t occ, uint32_t nb, uint32_t h) {
  const uint32_t mask = nb - 1;
  const uint64_t lowm = nb >= 64 ? ~0ull : ((1ull << nb) - 1ull);
  const uint64_t o = occ & lowm;
  const uint32_t g0 = ~(uint32_t)o & 0xFFFFu;  // empty positions of group 0
  uint32_t pos = h & mask, stride = 0;
  for (;;) {
    uint32_t win;
    if (nb >= 16) win = (uint32_t)((pos ? (((o >> pos) | (o << (nb - pos))) & lowm) : o) & 0xFFFFull);
    else win = (uint32_t)(((o | (o << 16)) >> pos) & 0xFFFFull);
    const uint32_t empty = ~win & 0xFFFFu;
    if (empty) {
      const uint32_t index = (pos + (uint32_t)__builtin_ctz(empty)) & mask;
      return ((o >> index) & 1ull) ? (uint32_t)__builtin_ctz(g0) : index;
    }
    stride += 16;
    pos = (pos + stride) & mask;
  }
}

// Register path for G <= 56 groups (final table <= 64 buckets): one wave,
// lane b holds the group stored in bucket b, occupancy is a uniform mask.
// All 64 lanes of the calling wave must be active.
__device__ void group_ranks_wave(uint32_t G, const uint32_t* g_key, const uint32_t* g_sort, uint32_t* g_rank) {
  const int lane = lane_id();
  // lane g: low word of the group's hash; lane t: group inserted t-th
  uint32_t hl = 0, ins = 0;
  if ((uint32_t)lane < G) {
    hl = (uint32_t)fx_hash_group(g_key[lane] >> 1, g_key[lane] & 1u);
    ins = g_sort[lane];
  }
  int32_t tab = -1;
  uint64_t occ = 0;
  uint32_t nb = 0, items = 0, growth_left = 0;
  for (uint32_t t = 0; t < G; t++) {
    const uint32_t gi = (uint32_t)readlane_i32((int32_t)ins, (int)t);
    const uint32_t h = (uint32_t)readlane_i32((int32_t)hl, (int)gi);
    if (growth_left == 0) {
      const uint32_t nnb = sw_buckets_for(max(items + 1, sw_capacity(nb) + 1));
      uint64_t old = occ, nocc = 0;
      int32_t ntab = -1;
      while (old) {  // re-insert in ascending old bucket order
        const uint32_t b = (uint32_t)__builtin_ctzll(old);
        old &= old - 1;
        const int32_t gj = readlane_i32(tab, (int)b);
        const uint32_t s = sw_find_slot_mask(nocc, nnb, (uint32_t)readlane_i32((int32_t)hl, gj));
        nocc |= 1ull << s;
        if ((uint32_t)lane == s) ntab = gj;
      }
      occ = nocc;
      tab = ntab;
      nb = nnb;
      growth_left = sw_capacity(nb) - items;
    }
    const uint32_t s = sw_find_slot_mask(occ, nb, h);
    occ |= 1ull << s;
    if ((uint32_t)lane == s) tab = (int32_t)gi;
    growth_left--;
    items++;
  }
  if ((occ >> lane) & 1ull) g_rank[tab] = (uint32_t)__builtin_popcountll(occ & lanemask_lt());
}

// ---- team primitives: T = 64 (one wave) or a block of T threads ----
template <int T>
__device__ __forceinline__ void team_sync() {
  if constexpr (T == 64) wave_sync();
  else __syncthreads();
}
// Exclusive prefix sum of one value per thread; `total` = team sum.
// sh: T/64 words of LDS (unused for T = 64).
template <int T>
__device__ __forceinline__ uint32_t team_excl_scan(uint32_t v, uint32_t& total, uint32_t* sh) {
  const uint32_t incl = wave_incl_scan(v);
  const uint32_t wtot = (uint32_t)readlane_i32((int32_t)incl, 63);
  if constexpr (T == 64) {
    total = wtot;
    return incl - v;
  } else {
    const uint32_t w = threadIdx.x >> 6;
    if (lane_id() == 0) sh[w] = wtot;
    __syncthreads();
    uint32_t before = 0, tot = 0;
    for (uint32_t k = 0; k < (uint32_t)(T / 64); k++) {
      const uint32_t x = sh[k];
      if (k < w) before += x;
      tot += x;
    }
    __syncthreads();
    total = tot;
    return before + incl - v;
  }
}
// Bitonic sort of p2 (power of two) 64-bit keys; `val` (optional) follows.
template <int T>
__device__ void team_bitonic(uint64_t* key, uint32_t* val, uint32_t p2, uint32_t tid) {
  for (uint32_t k = 2; k <= p2; k <<= 1) {
    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
      for (uint32_t i = tid; i < p2; i += T) {
        const uint32_t ixj = i ^ j;
        if (ixj > i) {
          const bool up = (i & k) == 0;
          const uint64_t x = key[i], y = key[ixj];
          if ((x > y) == up) {
            key[i] = y;
            key[ixj] = x;
            if (val) {
              const uint32_t t = val[i];
              val[i] = val[ixj];
              val[ixj] = t;
            }
          }
        }
      }
      team_sync<T>();
    }
  }
}

// The same sorting network over keys in global memory, staged through LDS
// chunks of CH keys: every step whose partner distance is below the chunk runs
// on a chunk in LDS, so only the few steps with distance >= CH touch global
// memory (about 3 of the 78 steps at 4096 keys).
template <int T, uint32_t CH>
__device__ void team_bitonic_staged(uint64_t* key, uint32_t* val, uint32_t p2, uint32_t tid, uint64_t* sk,
                                    uint32_t* sv) {
  const uint32_t ch = p2 < CH ? p2 : CH;
  auto chunk_pass = [&](uint32_t k_lo, uint32_t k_hi) {  // stages k_lo..k_hi, steps below ch
    for (uint32_t c0 = 0; c0 < p2; c0 += ch) {
      for (uint32_t i = tid; i < ch; i += T) {
        sk[i] = key[c0 + i];
        if (val) sv[i] = val[c0 + i];
      }
      team_sync<T>();
      for (uint32_t k = k_lo; k <= k_hi; k <<= 1) {
        for (uint32_t j = min(k >> 1, ch >> 1); j > 0; j >>= 1) {
          for (uint32_t i = tid; i < ch; i += T) {
            const uint32_t ixj = i ^ j;
            if (ixj > i) {
              const bool up = ((c0 + i) & k) == 0;
              const uint64_t x = sk[i], y = sk[ixj];
              if ((x > y) == up) {
                sk[i] = y;
                sk[ixj] = x;
                if (val) {
                  const uint32_t t = sv[i];
                  sv[i] = sv[ixj];
                  sv[ixj] = t;
                }
              }
            }
          }
          team_sync<T>();
        }
      }
      for (uint32_t i = tid; i < ch; i += T) {
        key[c0 + i] = sk[i];
        if (val) val[c0 + i] = sv[i];
      }
      team_sync<T>();
    }
  };
  chunk_pass(2, ch);
  for (uint32_t k = ch << 1; k <= p2; k <<= 1) {
    for (uint32_t j = k >> 1; j >= ch; j >>= 1) {
      for (uint32_t i = tid; i < p2; i += T) {
        const uint32_t ixj = i ^ j;
        if (ixj > i) {
          const bool up = (i & k) == 0;
          const uint64_t x = key[i], y = key[ixj];
          if ((x > y) == up) {
            key[i] = y;
            key[ixj] = x;
            if (val) {
              const uint32_t t = val[i];
              val[i] = val[ixj];
              val[ixj] = t;
            }
          }
        }
      }
      team_sync<T>();
    }
    chunk_pass(k, k);
  }
}

// Working arrays of one read's clustering (LDS for the wave tier, global
// scratch for the block tier).  Sizes: key/idx p2, cl_* n (+1), g_* gcap,
// table tcap, tmp gcap.
struct ClusterMem {
  uint64_t* key;       // hit keys; later the cluster sort keys
  uint32_t* idx;       // emission index of each hit; later cluster ids
  unsigned long long* cl_sum;  // per cluster: sum of diagonals (two's complement)
  uint32_t* cl_start;  // per cluster: first sorted hit; [C] = n
  uint32_t* cl_gid;    // per cluster: group id
  uint32_t* g_key;     // per group: contig*2 + strand
  uint32_t* g_min;     // per group: first emission index
  uint32_t* g_rank;    // per group: FxHashMap iteration rank
  uint32_t* g_sort;    // groups in insertion (first-appearance) order
  int32_t* table;      // general-path hash table
  int32_t* tmp;
  uint32_t gcap;       // capacity of the g_* arrays and tmp
  uint64_t* stage_key;  // block tier: LDS chunks for the sorts (CLUSTER_STAGE keys)
  uint32_t* stage_val;
  uint32_t* gt_key;     // block tier: LDS table of the read's distinct groups (CLUSTER_GT slots)
  uint32_t* gt_cnt;     //   hits per group, then the group's scatter cursor
  uint32_t* gt_misc;    //   [0] distinct groups, [1] listed groups
  unsigned long long* gt_lo;  //   lowest biased diagonal per group slot, then per group rank
  unsigned long long* gt_hi;  //   highest per slot, then the diagonal range per rank
  void* rsort;                //   LDS of the block radix sort
  uint32_t* gt_emin;          //   first emission index per slot, then per group rank
  uint32_t* g_c0;             // block tier after the segmented sort: first cluster per group ([G] = C)
};
constexpr uint32_t CLUSTER_STAGE = 1024;
constexpr uint32_t CLUSTER_GT = 1024, CLUSTER_GT_MAX = 512;  // table slots, groups handled
// Block-tier runs of whole segments sorted by one block radix sort (256 threads).
constexpr uint32_t RSORT_T = 256, RSORT_BIG = 16, RSORT_SMALL = 4, RSORT_TILE = RSORT_T * RSORT_BIG;
// rocPRIM's block radix sort with its wave-match ranking: 8-bit digits, so a
// run of keys of up to 32 bits takes at most four passes.
template <bool VALS, int IPT>
using RSort = rocprim::block_radix_sort<uint32_t, RSORT_T, IPT, typename std::conditional<VALS, uint32_t, rocprim::empty_type>::type,
                                        1, 1, 8, rocprim::block_radix_rank_algorithm::match>;
union RSortLds {
  typename RSort<true, RSORT_BIG>::storage_type big;
  typename RSort<true, RSORT_SMALL>::storage_type small;
  typename RSort<false, RSORT_BIG>::storage_type big_keys;
  typename RSort<false, RSORT_SMALL>::storage_type small_keys;
  uint64_t raw[2048];  // also the merge windows and the top-20 selection's tables
};

// The same network for p2 <= 64 * E keys held by one wave in registers:
// element e * 64 + lane sits in slot e of lane `lane`; partners at distance
// j < 64 are exchanged across lanes, at j >= 64 between slots of a lane.
// Every compare-exchange takes the same decision as team_bitonic's, ties
// included, so the (key, val) arrangement is the same.
template <int E>
__device__ void wave_bitonic_reg(uint64_t* key, uint32_t* val, uint32_t p2) {
  const uint32_t lane = lane_id();
  uint64_t kr[E];
  uint32_t vr[E];
#pragma unroll
  for (int e = 0; e < E; e++) {
    const uint32_t i = (uint32_t)e * 64u + lane;
    kr[e] = i < p2 ? key[i] : ~0ull;
    vr[e] = (val && i < p2) ? val[i] : 0u;
  }
  for (uint32_t k = 2; k <= p2; k <<= 1) {
    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
      if (j >= 64) {
        const uint32_t js = j >> 6;
#pragma unroll
        for (int e = 0; e < E; e++) {
          const uint32_t ep = (uint32_t)e ^ js;
          if ((uint32_t)e < ep && ep < (uint32_t)E) {
            const bool up = (((uint32_t)e * 64u) & k) == 0;
            const uint64_t x = kr[e], y = kr[ep];
            if ((x > y) == up) {
              kr[e] = y;
              kr[ep] = x;
              const uint32_t t = vr[e];
              vr[e] = vr[ep];
              vr[ep] = t;
            }
          }
        }
      } else {
        const bool lower = (lane & j) == 0;
#pragma unroll
        for (int e = 0; e < E; e++) {
          const uint32_t i = (uint32_t)e * 64u + lane;
          const uint64_t pk = (uint64_t)__shfl_xor((long long)kr[e], (int)j, 64);
          const uint32_t pv = (uint32_t)__shfl_xor((int)vr[e], (int)j, 64);
          const bool up = (i & k) == 0;
          // the pair (x at the lower index, y at the upper) swaps when (x > y) == up
          const bool sw = lower ? ((kr[e] > pk) == up) : ((pk > kr[e]) == up);
          if (sw) {
            kr[e] = pk;
            vr[e] = pv;
          }
        }
      }
    }
  }
#pragma unroll
  for (int e = 0; e < E; e++) {
    const uint32_t i = (uint32_t)e * 64u + lane;
    if (i < p2) {
      key[i] = kr[e];
      if (val) val[i] = vr[e];
    }
  }
  wave_sync();
}

template <int T>
__device__ __forceinline__ void cluster_sort(const ClusterMem& M, uint64_t* key, uint32_t* val, uint32_t p2,
                                             uint32_t tid) {
  if constexpr (T == 64) {
    if (p2 <= 64) wave_bitonic_reg<1>(key, val, p2);
    else if (p2 <= 128) wave_bitonic_reg<2>(key, val, p2);
    else if (p2 <= 256) wave_bitonic_reg<4>(key, val, p2);
    else if (p2 <= 512) wave_bitonic_reg<8>(key, val, p2);
    else team_bitonic<T>(key, val, p2, tid);
  } else {
    team_bitonic_staged<T, CLUSTER_STAGE>(key, val, p2, tid, M.stage_key, M.stage_val);
  }
}

__device__ __forceinline__ int64_t hit_diag(uint64_t k) {
  return (int64_t)(k & ((1ull << DIAG_BITS) - 1)) - DIAG_BIAS;
}

// The prefilter's 32 BF16 features of the candidate at rank `rank` (its cluster
// pc) of a read (pr): the table of the network's inputs, slots 27-31 zero.
struct PfCluster {
  uint32_t cnt;   // hits
  uint32_t nb;    // of them from 1-mismatch neighbours
  uint32_t ms;    // sum of the hits' bucket-size classes
  uint32_t mn;    // least bucket-size class
  uint32_t span;  // diagonal span
  bool rev;
};
struct PfRead {
  uint32_t c0, c1;               // the top and second candidates' hits
  uint32_t ties;                 // clusters with the top count
  uint32_t qsum, q10, q20, qlen; // quality sum, bases below Q10 and Q20, quality length
  uint32_t rep;                  // repeat-skipped windows
  uint32_t C, n, L, G;           // clusters, hits, read length, groups
};
__device__ __forceinline__ void pf_feature_row(uint16_t* out, uint32_t rank, const PfCluster& pc, const PfRead& pr) {
  const float cnt = (float)pc.cnt, c0 = (float)pr.c0, c1 = (float)pr.c1;
  const float nb = (float)pc.nb, mn = (float)pc.mn, ms = (float)pc.ms, span = (float)pc.span;
  const float nwin = (float)(pr.L >= (uint32_t)K ? pr.L - K + 1 : 1);
  const float ql = (float)max(pr.qlen, 1u);
  float f[32];
  f[0] = log2f(1.f + cnt);
  f[1] = cnt / c0;
  f[2] = log2f(1.f + nb);
  f[3] = nb / cnt;
  f[4] = (cnt - nb) / cnt;
  f[5] = log2f(1.f + span);
  f[6] = span / fmaxf(1.f, cnt - 1.f);
  f[7] = ms / cnt / 9.f;
  f[8] = mn / 9.f;
  f[9] = (float)rank / 20.f;
  f[10] = rank == 0 ? 1.f : 0.f;
  f[11] = pc.rev ? 1.f : 0.f;
  f[12] = log2f(1.f + c0);
  f[13] = log2f(1.f + c1);
  f[14] = cnt == c0 ? 1.f : 0.f;
  f[15] = (float)pr.ties / 20.f;
  f[16] = log2f(1.f + (float)pr.C) / 10.f;
  f[17] = log2f(1.f + (float)pr.n) / 10.f;
  f[18] = (float)pr.L / 150.f;
  f[19] = (float)pr.qsum / ql / 40.f;
  f[20] = (float)pr.q10 / ql;
  f[21] = (float)pr.q20 / ql;
  f[22] = cnt / nwin;
  f[23] = (float)pr.rep / nwin;
  f[24] = (float)pr.G / 20.f;
  f[25] = f[0] - f[12];
  f[26] = (c0 - c1) / c0;
  for (int u = 27; u < 32; u++) f[u] = 0.f;
  for (int u = 0; u < 32; u++) out[u] = float_to_bf16(f[u]);
}

// find_candidates (candidates.rs:30-78) of one read by a team of T threads.
// Returns false (uniformly) when the read has more groups than M.gcap; the
// caller then hands it to the block tier.
// Prefilter features of one read (only for the FEAT instances of the cluster kernels).
struct ClusterFeat {
  const uint8_t* hflag;  // this read's hit flags, emission order
  uint8_t* hf;           // working copy of hflag (LDS in the wave tier; hflag itself in the block tier)
  const uint8_t* qual;
  uint32_t qlen;
  const uint8_t* rcode;  // this read's window repeat codes
  uint16_t* out;         // this read's 20 feature rows of 32 BF16
  uint32_t* f1;          // per cluster: neighbour hits
  uint32_t* f2;          // per cluster: sum of bucket classes
  uint32_t* f3;          // per cluster: diagonal span
  uint32_t* f4;          // per cluster: minimum bucket class
  uint32_t* cnt;         // team counters: clusters tied at the top count, quality sum, q < 10, q < 20, repeat windows
};

__device__ __forceinline__ uint32_t gt_slot(uint32_t g) { return (g * 0x9E3779B1u) >> 22; }  // 10 bits

// Merge path: how many of the first d outputs of merging the sorted runs
// A[0, na) and B[0, nb) come from A (A first on ties).
__device__ __forceinline__ uint32_t merge_split(const uint64_t* A, uint32_t na, const uint64_t* B, uint32_t nb,
                                                uint32_t d, uint64_t cm = ~0ull) {
  uint32_t lo = d > nb ? d - nb : 0u, hi = min(d, na);
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if ((A[mid] & cm) <= (B[d - 1 - mid] & cm)) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// The same split found by one wave: a 64-way search that evaluates the merge
// path's predicate (A[i] <= B[d - 1 - i], true below the split) at 64 evenly
// spaced points per round, so a split in a run of m keys takes about log64(m)
// rounds of parallel loads instead of log2(m) dependent ones.  All 64 lanes call
// it; every lane returns the split.
__device__ __forceinline__ uint32_t merge_split_wave(const uint64_t* A, uint32_t na, const uint64_t* B, uint32_t nb,
                                                     uint32_t d, uint64_t cm = ~0ull) {
  uint32_t lo = d > nb ? d - nb : 0u, hi = min(d, na);
  const uint32_t lane = (uint32_t)lane_id();
  while (lo < hi) {
    const uint32_t step = (hi - lo + 63) / 64;
    const uint32_t i = lo + lane * step;
    const bool t = i < hi && (A[i] & cm) <= (B[d - 1 - i] & cm);
    const uint32_t nt = (uint32_t)__popcll(__ballot(t));  // the true points are a prefix
    if (nt == 0) {
      hi = lo;
    } else {
      const uint32_t last = lo + (nt - 1) * step;
      hi = min(hi, lo + nt * step);
      lo = last + 1;
    }
  }
  return lo;
}

// One merge step of a block: the sorted runs src[s, s + na) and
// src[s + na, s + na + nb) into dst[s, s + na + nb) (keys, with their values when
// VALS), MERGE_CHUNK outputs at a time.  The waves first find where every chunk's
// outputs begin along the merge path (one split per wave at a time, kept in
// stage_val); then for each chunk the block loads the two windows of the runs
// that feed it into LDS (the radix sort's storage and stage_key, free while the
// sort merges) with coalesced loads, and each thread merges its MERGE_CHUNK / 256
// outputs from there.
constexpr uint32_t MERGE_CHUNK = 2048;
template <bool VALS>
__device__ __forceinline__ void block_merge(const ClusterMem& M, const uint64_t* sk, const uint32_t* sv,
                                            uint64_t* dk, uint32_t* dv, uint32_t s, uint32_t na, uint32_t nb,
                                            uint32_t tid, uint64_t cm = ~0ull) {
  static_assert(sizeof(RSortLds) >= MERGE_CHUNK * sizeof(uint64_t), "merge window keys fit the sort's LDS");
  static_assert(CLUSTER_STAGE * sizeof(uint64_t) >= MERGE_CHUNK * sizeof(uint32_t), "merge window values fit");
  constexpr uint32_t PER = MERGE_CHUNK / RSORT_T, NW = RSORT_T / 64, BATCH = CLUSTER_STAGE - 1;
  uint64_t* lk = reinterpret_cast<uint64_t*>(M.rsort);
  uint32_t* lv = reinterpret_cast<uint32_t*>(M.stage_key);
  uint32_t* sp = M.stage_val;  // splits of up to BATCH chunks (BATCH + 1 boundaries)
  const uint64_t* A = sk + s;
  const uint64_t* B = sk + s + na;
  const uint32_t total = na + nb;
  for (uint32_t D0 = 0; D0 < total; D0 += BATCH * MERGE_CHUNK) {
    const uint32_t nch = min(BATCH, (total - D0 + MERGE_CHUNK - 1) / MERGE_CHUNK);
    for (uint32_t k = tid >> 6; k <= nch; k += NW)
      sp[k] = merge_split_wave(A, na, B, nb, min(total, D0 + k * MERGE_CHUNK), cm);
    __syncthreads();
    for (uint32_t k = 0; k < nch; k++) {
      const uint32_t D = D0 + k * MERGE_CHUNK, E = min(total, D + MERGE_CHUNK);
      const uint32_t ia = sp[k], ja = D - ia;
      const uint32_t an = sp[k + 1] - ia, bn = (E - sp[k + 1]) - ja;
      uint64_t wk[PER];
      uint32_t wv[PER];
#pragma unroll
      for (uint32_t u = 0; u < PER; u++) {  // all of a thread's window loads in flight at once
        const uint32_t x = tid + u * RSORT_T;
        const uint32_t src = x < an ? s + ia + x : s + na + ja + (x - an);
        wk[u] = x < an + bn ? sk[src] : 0ull;
        if constexpr (VALS) wv[u] = x < an + bn ? sv[src] : 0u;
      }
#pragma unroll
      for (uint32_t u = 0; u < PER; u++) {
        const uint32_t x = tid + u * RSORT_T;
        if (x < an + bn) {
          lk[x] = wk[u];
          if constexpr (VALS) lv[x] = wv[u];
        }
      }
      __syncthreads();
      const uint32_t t0 = min(an + bn, tid * PER), t1 = min(an + bn, t0 + PER);
      uint32_t i = merge_split(lk, an, lk + an, bn, t0, cm), j = t0 - i;
      for (uint32_t d = t0; d < t1; d++) {
        const bool takeA = j >= bn || (i < an && (lk[i] & cm) <= (lk[an + j] & cm));
        const uint32_t x = takeA ? i : an + j;
        dk[s + D + d] = lk[x];
        if constexpr (VALS) dv[s + D + d] = lv[x];
        if (takeA) i++;
        else j++;
      }
      __syncthreads();
    }
  }
}

// Rank of the run segment holding composite key x: the last q in [r, r1) with pre[q] <= x.
__device__ __forceinline__ uint32_t run_segment(const uint32_t* pre, uint32_t r, uint32_t r1, uint32_t x) {
  uint32_t lo = r, hi = r1 - 1;
  while (lo < hi) {
    const uint32_t mid = (lo + hi + 1) >> 1;
    if (pre[mid] <= x) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Sorts hits [a0, a0 + cnt) = whole segments r..r1-1 with one block radix sort
// of 32-bit keys: segment q's biased diagonals [lo_q, lo_q + range_q] map to
// pre_q + (d - lo_q), the segments' ranges laid end to end in group order, so the
// composite key orders (group, diagonal) as the 64-bit hit key does.  The
// emission index rides along as the value.
template <int IPT, bool VALS>
__device__ __forceinline__ void run_radix_sort(const ClusterMem& M, uint32_t a0, uint32_t cnt, uint32_t r, uint32_t r1,
                               int bits, uint32_t tid) {
  using BRS = RSort<VALS, IPT>;
  const uint32_t* pre = M.gt_cnt + CLUSTER_GT_MAX;
  uint32_t k[IPT], v[IPT];
  const uint32_t pad = bits >= 32 ? 0xFFFFFFFFu : ((1u << bits) - 1u);
#pragma unroll
  for (int j = 0; j < IPT; j++) {  // striped loads: a sort takes its input in any order
    const uint32_t pos = (uint32_t)j * RSORT_T + tid;
    k[j] = pad;
    v[j] = 0;
    if (pos < cnt) {
      const uint64_t key = M.key[a0 + pos];
      const uint32_t q = run_segment(M.gt_key, r, r1, a0 + pos);
      k[j] = pre[q] + (uint32_t)((key & ((1ull << DIAG_BITS) - 1)) - M.gt_lo[q]);
      if constexpr (VALS) v[j] = M.idx[a0 + pos];
    }
  }
  auto& tmp = *reinterpret_cast<typename BRS::storage_type*>(M.rsort);
  if constexpr (VALS) BRS().sort_to_striped(k, v, tmp, 0u, (unsigned)bits);
  else BRS().sort_to_striped(k, tmp, 0u, (unsigned)bits);
#pragma unroll
  for (int j = 0; j < IPT; j++) {  // striped: thread tid holds ranks j * RSORT_T + tid (coalesced stores)
    const uint32_t pos = (uint32_t)j * RSORT_T + tid;
    if (pos < cnt) {
      const uint32_t q = run_segment(pre, r, r1, k[j]);
      M.key[a0 + pos] = ((uint64_t)M.gt_cnt[q] << DIAG_BITS) | (M.gt_lo[q] + (k[j] - pre[q]));
      if constexpr (VALS) M.idx[a0 + pos] = v[j];
    }
  }
  __syncthreads();
}

// The block tier's group table: every hit's (contig, strand) group counted in
// the LDS table gt_* (CLUSTER_GT slots; per slot the group key, its hits, its
// lowest and highest biased diagonal and its first emission index); gt_misc[0]
// = distinct groups (may exceed CLUSTER_GT_MAX: the table then holds only some).
// hits(i): the read's hit i.
template <int T, typename HitAt>
__device__ __forceinline__ void group_table_count(HitAt hits, uint32_t n, const ClusterMem& M, uint32_t tid) {
  constexpr uint32_t U = 8;  // hits loaded ahead per thread: the loads of a pass overlap
  for (uint32_t i = tid; i < CLUSTER_GT; i += T) {
    M.gt_key[i] = 0xFFFFFFFFu;
    M.gt_cnt[i] = 0;
    M.gt_lo[i] = ~0ull;
    M.gt_hi[i] = 0ull;
    M.gt_emin[i] = 0xFFFFFFFFu;
  }
  if (tid < 2) M.gt_misc[tid] = 0;
  team_sync<T>();
  for (uint32_t i0 = tid; i0 < n; i0 += U * T) {
    uint64_t kb[U];
#pragma unroll
    for (uint32_t u = 0; u < U; u++) kb[u] = i0 + u * T < n ? hits(i0 + u * T) : 0ull;
#pragma unroll
    for (uint32_t u = 0; u < U; u++) {
      const uint32_t i = i0 + u * T;
      if (i >= n) break;
      const uint64_t k = kb[u];
      const uint32_t g = (uint32_t)(k >> DIAG_BITS);
      const unsigned long long d = k & ((1ull << DIAG_BITS) - 1);
      // the wave's hits mostly share one group (a bucket's entries are in group
      // order): a full wave of one group probes the table once, from lane 0, and
      // updates its slot once (lane 0 holds the wave's first emission index)
      const uint32_t g0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)g);
      const uint64_t act = __ballot(true);
      const bool uni = act == ~0ull && __ballot(g == g0) == act;
      // a slot, once keyed, keeps its key: a plain read finds the group, and only
      // an empty slot takes a compare-and-swap; the minima and maxima below are
      // only ever lowered / raised, so a read that shows no change needed is final
      uint32_t sl = gt_slot(g);
      if (!uni || lane_id() == 0) {
        for (uint32_t probe = 0; probe < CLUSTER_GT; probe++) {
          const uint32_t cur = __atomic_load_n(&M.gt_key[sl], __ATOMIC_RELAXED);
          if (cur == g) break;
          if (cur == 0xFFFFFFFFu) {
            const uint32_t prev = atomicCAS(&M.gt_key[sl], 0xFFFFFFFFu, g);
            if (prev == 0xFFFFFFFFu) atomicAdd(&M.gt_misc[0], 1u);
            if (prev == 0xFFFFFFFFu || prev == g) break;
          }
          sl = (sl + 1) & (CLUSTER_GT - 1);
        }
      }
      if (uni) {
        const uint32_t sl0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)sl);
        unsigned long long lo = d, hi = d;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
          lo = min(lo, (unsigned long long)shfl_xor_i64((int64_t)lo, o));
          hi = max(hi, (unsigned long long)shfl_xor_i64((int64_t)hi, o));
        }
        if (lane_id() == 0) {
          atomicAdd(&M.gt_cnt[sl0], 64u);
          if (lo < __atomic_load_n(&M.gt_lo[sl0], __ATOMIC_RELAXED)) atomicMin(&M.gt_lo[sl0], lo);
          if (hi > __atomic_load_n(&M.gt_hi[sl0], __ATOMIC_RELAXED)) atomicMax(&M.gt_hi[sl0], hi);
          if (i < __atomic_load_n(&M.gt_emin[sl0], __ATOMIC_RELAXED)) atomicMin(&M.gt_emin[sl0], i);
        }
      } else {
        // runs of one group over consecutive lanes (a bucket's entries come in
        // group order; the active lanes are a prefix): a run's first lane adds
        // the run's count and, holding its lowest emission index, the index
        const uint32_t gp = (uint32_t)__shfl_up((int)g, 1, 64);
        const uint32_t ln = (uint32_t)lane_id();
        const bool head = ln == 0 || gp != g;
        const uint64_t ends = __ballot(head) | ~act;  // run heads and the inactive tail
        if (head) {
          const uint64_t after = ends & ~((2ull << ln) - 1ull);  // lanes above this one (none for lane 63)
          const uint32_t nxt = after ? (uint32_t)__builtin_ctzll(after) : 64u;
          atomicAdd(&M.gt_cnt[sl], nxt - ln);
          if (i < __atomic_load_n(&M.gt_emin[sl], __ATOMIC_RELAXED)) atomicMin(&M.gt_emin[sl], i);
        }
        if (d < __atomic_load_n(&M.gt_lo[sl], __ATOMIC_RELAXED)) atomicMin(&M.gt_lo[sl], d);
        if (d > __atomic_load_n(&M.gt_hi[sl], __ATOMIC_RELAXED)) atomicMax(&M.gt_hi[sl], d);
      }
    }
  }
  team_sync<T>();
}

// Block tier: the (group, diagonal) sort as a segmented sort.  The hits are
// counted per group in an LDS table (with each group's lowest and highest
// diagonal and its first emission index), the groups ranked in ascending order,
// the hits scattered into one segment per group (key[], with the emission index
// in idx[] when VALS), and then runs of consecutive whole segments are sorted by
// a block radix sort of 32-bit composite keys (run_radix_sort) -- as many
// segments as fit one 4096-hit tile and a 2^32 span of diagonals.  A segment
// larger than a tile is sorted tile by tile and the sorted tiles merged pairwise
// (block_merge) through the scratch that cl_sum / cl_start provide until the
// clustering needs them; one whose diagonals span 2^32 or more goes through the
// staged bitonic network there.  Ties between equal keys may land in any order,
// as in the full sort: the clustering only sums diagonals (and, with VALS, reads
// per-hit flags).  On return gt_key[r] is the segment start of group rank r
// (gt_key[G] = n), gt_cnt[r] its group key and gt_emin[r] its first emission
// index.  Returns false (nothing written) when the read has more than
// CLUSTER_GT_MAX groups.
template <int T, bool VALS>
__device__ __forceinline__ bool group_segment_sort(const uint64_t* hits, uint32_t n, const ClusterMem& M, uint32_t tid,
                                                   bool& by_rank) {
  static_assert(T == RSORT_T, "the block tier runs 256 threads");
  constexpr uint32_t U = 8;  // hits loaded ahead per thread in the scatter
  group_table_count<T>([&](uint32_t i) { return hits[i]; }, n, M, tid);
  const uint32_t G = M.gt_misc[0];
  if (G > CLUSTER_GT_MAX) return false;
  // Segment order.  Without per-hit values and for at most 56 groups: the
  // FxHashMap iteration order (rank) of the groups, found here from their first
  // emission indices, so the sorted keys run in the order the candidates are
  // enumerated (candidates.rs:45-72).  Otherwise ascending group keys.
  by_rank = !VALS && G <= 56;
  uint32_t g2 = 1;
  while (g2 < G) g2 <<= 1;
  for (uint32_t i = tid; i < CLUSTER_GT; i += T)
    if (M.gt_key[i] != 0xFFFFFFFFu) {
      const uint32_t at = atomicAdd(&M.gt_misc[1], 1u);
      M.stage_key[at] = ((uint64_t)(by_rank ? M.gt_emin[i] : M.gt_key[i]) << 32) | i;
    }
  team_sync<T>();
  for (uint32_t i = G + tid; i < g2; i += T) M.stage_key[i] = ~0ull;
  team_sync<T>();
  team_bitonic<T>(M.stage_key, nullptr, g2, tid);
  if (by_rank) {
    // stage_key[t]: the group inserted t-th (first appearance order); its key in
    // stage_val[t], t itself as the insertion list, its rank to stage_val[128 + t]
    static_assert(T >= 64, "one entry per thread for G <= 56");
    if (tid < G) {
      M.stage_val[tid] = M.gt_key[(uint32_t)M.stage_key[tid]];
      M.stage_val[64 + tid] = tid;
    }
    team_sync<T>();
    if (tid < 64) group_ranks_wave(G, M.stage_val, M.stage_val + 64, M.stage_val + 128);
    team_sync<T>();
    const uint32_t sl = tid < G ? (uint32_t)M.stage_key[tid] : 0u;
    const uint32_t rk = tid < G ? M.stage_val[128 + tid] : 0u;
    team_sync<T>();
    if (tid < G) M.stage_key[rk] = ((uint64_t)M.gt_key[sl] << 32) | sl;  // (group << 32 | slot) by rank
    team_sync<T>();
  }
  // segment starts in group order: the scatter cursor of every slot; the start of
  // group rank r stays in stage_val[r] (stage_val[G] = n)
  uint32_t tot;
  const uint32_t per = (G + T - 1) / T;  // ranks per thread (G <= 512: at most 2)
  uint32_t local = 0;
  for (uint32_t r = tid * per; r < min(G, tid * per + per); r++) local += M.gt_cnt[(uint32_t)M.stage_key[r]];
  uint32_t run = team_excl_scan<T>(local, tot, M.gt_misc + 2 /* scan scratch */);
  team_sync<T>();
  for (uint32_t r = tid * per; r < min(G, tid * per + per); r++) {
    const uint32_t sl = (uint32_t)M.stage_key[r];
    M.stage_val[r] = run;
    run += M.gt_cnt[sl];
    M.gt_cnt[sl] = M.stage_val[r];
  }
  if (tid == 0) M.stage_val[G] = n;
  team_sync<T>();
  for (uint32_t i0 = tid; i0 < n; i0 += U * T) {
    uint64_t kb[U];
#pragma unroll
    for (uint32_t u = 0; u < U; u++) kb[u] = i0 + u * T < n ? hits[i0 + u * T] : 0ull;
#pragma unroll
    for (uint32_t u = 0; u < U; u++) {
      const uint32_t i = i0 + u * T;
      if (i >= n) break;
      const uint64_t k = kb[u];
      const uint32_t g = (uint32_t)(k >> DIAG_BITS);
      const uint32_t g0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)g);
      const uint64_t act = __ballot(true);
      uint32_t pos;
      if (__ballot(g == g0) == act) {
        // one group across the active lanes (a prefix of the wave): one cursor
        // update, the lanes at consecutive positions
        uint32_t sl = gt_slot(g0);
        while (M.gt_key[sl] != g0) sl = (sl + 1) & (CLUSTER_GT - 1);
        uint32_t base = 0;
        if (lane_id() == 0) base = atomicAdd(&M.gt_cnt[sl], (uint32_t)__popcll(act));
        pos = (uint32_t)__builtin_amdgcn_readfirstlane((int)base) + lane_id();
      } else {
        uint32_t sl = gt_slot(g);
        while (M.gt_key[sl] != g) sl = (sl + 1) & (CLUSTER_GT - 1);
        pos = atomicAdd(&M.gt_cnt[sl], 1u);
      }
      M.key[pos] = k;
      if constexpr (VALS) M.idx[pos] = i;
    }
  }
  team_sync<T>();
  // per group rank r (r < G <= 512, two per thread at most): segment start in
  // gt_key[r] (gt_key[G] = n), group key in gt_cnt[r], lowest diagonal in
  // gt_lo[r], the diagonal range in gt_hi[r] and the first emission index in
  // gt_emin[r]; slot data read before the barrier, rank data written after it
  uint64_t rl[2] = {0, 0}, rh[2] = {0, 0};
  uint32_t rg[2] = {0, 0}, re[2] = {0, 0};
  for (uint32_t u = 0; u < 2; u++) {
    const uint32_t r = tid + u * T;
    if (r < G) {
      const uint32_t sl = (uint32_t)M.stage_key[r];
      rg[u] = (uint32_t)(M.stage_key[r] >> 32);
      rl[u] = M.gt_lo[sl];
      rh[u] = M.gt_hi[sl] - M.gt_lo[sl];
      re[u] = M.gt_emin[sl];
    }
  }
  team_sync<T>();
  for (uint32_t u = 0; u < 2; u++) {
    const uint32_t r = tid + u * T;
    if (r < G) {
      M.gt_cnt[r] = rg[u];
      M.gt_lo[r] = rl[u];
      M.gt_hi[r] = rh[u];
      M.gt_emin[r] = re[u];
    }
  }
  for (uint32_t r = tid; r <= G; r += T) M.gt_key[r] = M.stage_val[r];
  team_sync<T>();
  uint32_t* pre = M.gt_cnt + CLUSTER_GT_MAX;  // composite-key offsets of the current run
  uint32_t r = 0;
  while (r < G) {
    // the run r..r1-1: thread 0 lays the segments' ranges end to end
    if (tid == 0) {
      const uint32_t a0 = M.gt_key[r];
      uint64_t span = 0;
      uint32_t r1 = r;
      while (r1 < G && M.gt_key[r1 + 1] - a0 <= RSORT_TILE && span + M.gt_hi[r1] + 1 < (1ull << 32)) {
        pre[r1] = (uint32_t)span;
        span += M.gt_hi[r1] + 1;
        r1++;
      }
      M.gt_misc[0] = r1;
      M.gt_misc[1] = (uint32_t)(span ? 64 - __builtin_clzll(span) : 0);  // key bits
    }
    team_sync<T>();
    const uint32_t r1 = M.gt_misc[0], bits = M.gt_misc[1];
    const uint32_t a0 = M.gt_key[r];
    team_sync<T>();
    if (r1 > r) {
      const uint32_t cnt = M.gt_key[r1] - a0;
      if (cnt > 1) {
        if (cnt <= RSORT_SMALL * RSORT_T) run_radix_sort<RSORT_SMALL, VALS>(M, a0, cnt, r, r1, (int)bits, tid);
        else run_radix_sort<RSORT_BIG, VALS>(M, a0, cnt, r, r1, (int)bits, tid);
      }
      r = r1;
      continue;
    }
    const uint32_t cnt = M.gt_key[r + 1] - a0;
    uint64_t* tk = reinterpret_cast<uint64_t*>(M.cl_sum);
    uint32_t* tv = VALS ? M.cl_start : nullptr;
    if (M.gt_hi[r] + 1 < (1ull << 32)) {
      // segment r alone exceeds a tile: each tile sorted on its own (a one-segment
      // run), then pairs of sorted runs merged, alternating between key/idx and
      // the cl_sum / cl_start scratch (both hold n entries)
      if (tid == 0) {
        pre[r] = 0;
        const uint64_t span = M.gt_hi[r] + 1;
        M.gt_misc[1] = (uint32_t)(64 - __builtin_clzll(span));
      }
      team_sync<T>();
      const int sbits = (int)M.gt_misc[1];
      for (uint32_t t0 = 0; t0 < cnt; t0 += RSORT_TILE)
        run_radix_sort<RSORT_BIG, VALS>(M, a0 + t0, min(RSORT_TILE, cnt - t0), r, r + 1, sbits, tid);
      bool in_key = true;
      for (uint32_t w = RSORT_TILE; w < cnt; w <<= 1) {
        const uint64_t* sk = in_key ? M.key : tk;
        const uint32_t* sv = in_key ? M.idx : tv;
        uint64_t* dk = in_key ? tk : M.key;
        uint32_t* dv = in_key ? tv : M.idx;
        for (uint32_t s0 = 0; s0 < cnt; s0 += 2 * w) {
          const uint32_t na = min(w, cnt - s0), nb = min(w, cnt - s0 - na);
          block_merge<VALS>(M, sk, sv, dk, dv, a0 + s0, na, nb, tid);
        }
        team_sync<T>();
        in_key = !in_key;
      }
      if (!in_key) {
        for (uint32_t i = tid; i < cnt; i += T) {
          M.key[a0 + i] = tk[a0 + i];
          if constexpr (VALS) M.idx[a0 + i] = tv[a0 + i];
        }
        team_sync<T>();
      }
      r = r + 1;
      continue;
    }
    // a segment whose diagonals span 2^32 or more: padded copy in cl_sum / cl_start
    uint32_t p2 = 1;
    while (p2 < cnt) p2 <<= 1;
    for (uint32_t i = tid; i < p2; i += T) {
      tk[i] = i < cnt ? M.key[a0 + i] : ~0ull;
      if constexpr (VALS) tv[i] = i < cnt ? M.idx[a0 + i] : 0u;
    }
    team_sync<T>();
    team_bitonic_staged<T, CLUSTER_STAGE>(tk, tv, p2, tid, M.stage_key, M.stage_val);
    for (uint32_t i = tid; i < cnt; i += T) {
      M.key[a0 + i] = tk[i];
      if constexpr (VALS) M.idx[a0 + i] = tv[i];
    }
    team_sync<T>();
    r = r + 1;
  }
  return true;
}

// Block tier, many clusters: only the nt smallest of the C (distinct) score keys
// in key[0..C) matter.  An MSB-first radix select (one LDS histogram of a key
// byte per pass) finds the nt-th smallest key; the keys up to it are sorted and
// swapped to the front of key[] / idx[], so key[] still holds every cluster's key
// (the prefilter features count ties over all of them).  idx[c] must be c.
template <int T>
__device__ __forceinline__ void top_clusters(const ClusterMem& M, uint32_t C, uint32_t nt, uint32_t tid) {
  uint32_t* hist = M.gt_cnt + CLUSTER_GT_MAX;  // gt_cnt[0, G) may hold the group keys
  uint32_t* sh = M.gt_misc;
  uint64_t prefix = 0, mask = 0;
  uint32_t need = nt;
  for (int byte = 7; byte >= 0; byte--) {
    const int sh8 = 8 * byte;
    for (uint32_t i = tid; i < 256; i += T) hist[i] = 0;
    team_sync<T>();
    for (uint32_t c = tid; c < C; c += T) {
      const uint64_t k = M.key[c];
      if ((k & mask) == prefix) atomicAdd(&hist[(uint32_t)(k >> sh8) & 255u], 1u);
    }
    team_sync<T>();
    if (tid == 0) {
      uint32_t acc = 0, d = 0;
      for (; d < 255; d++) {
        if (acc + hist[d] >= need) break;
        acc += hist[d];
      }
      sh[0] = d;
      sh[1] = need - acc;
    }
    team_sync<T>();
    prefix |= (uint64_t)sh[0] << sh8;
    mask |= 0xFFull << sh8;
    need = sh[1];
    team_sync<T>();
  }
  if (tid == 0) sh[2] = 0;
  team_sync<T>();
  for (uint32_t c = tid; c < C; c += T)
    if (M.key[c] <= prefix) {
      const uint32_t at = atomicAdd(&sh[2], 1u);
      M.stage_key[at] = M.key[c];
      M.stage_val[at] = c;
    }
  team_sync<T>();
  uint32_t p2 = 1;
  while (p2 < nt) p2 <<= 1;
  for (uint32_t i = nt + tid; i < p2; i += T) M.stage_key[i] = ~0ull;
  team_sync<T>();
  team_bitonic<T>(M.stage_key, M.stage_val, p2, tid);
  if (tid == 0) {
    uint32_t pos[TOP_CANDIDATES];
    for (uint32_t j = 0; j < nt; j++) pos[j] = M.stage_val[j];
    for (uint32_t j = 0; j < nt; j++) {
      const uint32_t q = pos[j];
      if (q == j) continue;
      const uint64_t tk = M.key[j];
      M.key[j] = M.key[q];
      M.key[q] = tk;
      const uint32_t tv = M.idx[j];
      M.idx[j] = M.idx[q];
      M.idx[q] = tv;
      for (uint32_t m2 = j + 1; m2 < nt; m2++)
        if (pos[m2] == j) pos[m2] = q;
    }
  }
  team_sync<T>();
}

// Reduction over the team (T = 64: one wave; else the block, through sh[T/64]).
template <int T, typename Op>
__device__ __forceinline__ uint32_t team_reduce(uint32_t v, Op op, uint32_t* sh) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = op(v, (uint32_t)__shfl_xor((int)v, o, 64));
  if constexpr (T == 64) {
    return v;
  } else {
    const uint32_t w = threadIdx.x >> 6;
    if (lane_id() == 0) sh[w] = v;
    __syncthreads();
    uint32_t r = sh[0];
    for (uint32_t k = 1; k < (uint32_t)(T / 64); k++) r = op(r, sh[k]);
    __syncthreads();
    return r;
  }
}

// Block tier, many clusters, 32-bit selection keys key[0..C) (unique, ascending in
// the order of the score sort): the nt smallest by an MSB-first radix select --
// bytes in which every key agrees (kor ^ kand) are taken without a pass, the bin
// of the nt-th key found by a scan over the 256 bins -- then those nt sorted in
// LDS; their cluster ids go to idx[0..nt).
template <int T>
__device__ __forceinline__ void top_clusters32(const ClusterMem& M, const uint32_t* key, uint32_t C, uint32_t nt, uint32_t kor,
                               uint32_t kand, uint32_t tid) {
  static_assert(T == 256, "one histogram bin per thread");
  uint32_t* hist = M.gt_cnt + CLUSTER_GT_MAX;  // gt_cnt[0, G) may hold the group keys
  uint32_t* sh = M.gt_misc;
  uint32_t prefix = 0, mask = 0, need = nt;
  for (int byte = 3; byte >= 0; byte--) {
    const int sh8 = 8 * byte;
    const uint32_t bm = 0xFFu << sh8;
    if (((kor ^ kand) & bm) == 0) {
      prefix |= kand & bm;
      mask |= bm;
      continue;
    }
    hist[tid] = 0;
    team_sync<T>();
    for (uint32_t c = tid; c < C; c += T) {
      const uint32_t k = key[c];
      if ((k & mask) == prefix) atomicAdd(&hist[(k >> sh8) & 255u], 1u);
    }
    team_sync<T>();
    const uint32_t v = hist[tid];
    uint32_t tot;
    const uint32_t ex = team_excl_scan<T>(v, tot, sh + 2);
    if (ex < need && ex + v >= need) {
      sh[0] = tid;
      sh[1] = need - ex;
    }
    team_sync<T>();
    prefix |= sh[0] << sh8;
    mask |= bm;
    need = sh[1];
    team_sync<T>();
  }
  if (tid == 0) sh[0] = 0;
  team_sync<T>();
  for (uint32_t c = tid; c < C; c += T)
    if (key[c] <= prefix) {
      const uint32_t at = atomicAdd(&sh[0], 1u);
      M.stage_key[at] = key[c];
      M.stage_val[at] = c;
    }
  team_sync<T>();
  for (uint32_t j = nt + tid; j < 32; j += T) M.stage_key[j] = ~0ull;
  team_sync<T>();
  team_bitonic<T>(M.stage_key, M.stage_val, 32, tid);
  for (uint32_t j = tid; j < nt; j += T) M.idx[j] = M.stage_val[j];
  team_sync<T>();
}

// Adds 1 to hist[v] for every active lane of the wave: one LDS add per distinct
// value in the wave (cluster counts are mostly 1 or 2, so a per-lane add would
// queue most of the wave on one address).  All 64 lanes must call it.
__device__ __forceinline__ void wave_hist_add(uint32_t* hist, uint32_t v, bool act) {
  uint64_t pend = __ballot(act);
  while (pend) {
    const int l0 = __builtin_ctzll(pend);
    const uint32_t v0 = (uint32_t)readlane_i32((int32_t)v, l0);
    const uint64_t m = __ballot(act && v == v0);
    if (lane_id() == l0) atomicAdd(&hist[v0], (uint32_t)__popcll(m));
    pend &= ~m;
  }
}

// Block tier after the segmented sort (G <= CLUSTER_GT_MAX groups, per-group
// arrays in LDS, g_c0 filled), C > nt clusters: the nt first clusters of the
// score order (count descending, then group rank, then position in the group)
// without a sort of all C.  A histogram of the counts (min(count, 255)) gives the
// count thr of the nt-th cluster: every cluster counted above thr is taken, and
// of those counted exactly thr the first `need` in (group rank, position) order
// -- each group's share found from the per-rank tallies of such clusters, and
// taken by one wave walking the group's clusters in order.  The nt chosen are
// sorted in LDS and their cluster ids written to idx[0..nt).  Returns false
// (uniformly, nothing written) when thr falls in the top bin (counts >= 255);
// the caller then selects by the general path.
template <int T>
__device__ __forceinline__ bool top_by_count(const ClusterMem& M, uint32_t C, uint32_t G, uint32_t nt, uint32_t tid) {
  static_assert(T == 256, "one histogram bin per thread");
  constexpr uint32_t NW = T / 64;
  uint32_t* hist = M.gt_cnt + CLUSTER_GT_MAX;  // gt_cnt[0, G) holds the group keys
  uint32_t* sh = M.gt_misc;
  uint32_t* g_inv = M.g_c0 + CLUSTER_GT_MAX + 1;  // group of each rank
  uint32_t* eqr = g_inv + CLUSTER_GT_MAX;          // clusters counted thr, per group rank
  const uint32_t lane = (uint32_t)lane_id(), wv = tid >> 6;
  hist[tid] = 0;
  for (uint32_t q = tid; q < G; q += T) {
    eqr[q] = 0;
    g_inv[M.g_rank[q]] = q;
  }
  team_sync<T>();
  constexpr uint32_t U = 4;  // windows of clusters loaded ahead per wave
  for (uint32_t cb = tid - lane; cb < C; cb += U * T) {
    uint32_t cnt[U];
#pragma unroll
    for (uint32_t u = 0; u < U; u++) {
      const uint32_t c = cb + u * T + lane;
      cnt[u] = c < C ? M.cl_start[c + 1] - M.cl_start[c] : 0u;
    }
#pragma unroll
    for (uint32_t u = 0; u < U; u++) wave_hist_add(hist, min(cnt[u], 255u), cb + u * T + lane < C);
  }
  team_sync<T>();
  // thread t holds bin 255 - t: the inclusive scan is the number counted >= that bin
  const uint32_t x = hist[255 - tid];
  uint32_t tot;
  const uint32_t ex = team_excl_scan<T>(x, tot, sh + 2);
  if (ex < nt && ex + x >= nt) {
    sh[0] = 255 - tid;  // thr
    sh[1] = ex;         // clusters counted above thr
  }
  team_sync<T>();
  const uint32_t thr = sh[0], ngt = sh[1];
  team_sync<T>();
  if (thr == 255) return false;
  if (tid == 0) sh[0] = 0;  // entries appended
  team_sync<T>();
  for (uint32_t cb = tid - lane; cb < C; cb += U * T) {
    uint32_t cnt[U], gi[U];
#pragma unroll
    for (uint32_t u = 0; u < U; u++) {
      const uint32_t c = cb + u * T + lane;
      cnt[u] = c < C ? M.cl_start[c + 1] - M.cl_start[c] : 0u;
      gi[u] = c < C ? M.cl_gid[c] : 0u;
    }
#pragma unroll
    for (uint32_t u = 0; u < U; u++) {
      const uint32_t c = cb + u * T + lane;
      const bool act = c < C;
      const uint32_t q = act ? M.g_rank[gi[u]] : 0u;
      if (act && cnt[u] > thr) {
        const uint32_t at = atomicAdd(&sh[0], 1u);
        M.stage_key[at] = ((uint64_t)(0xFFFFFFFFu - cnt[u]) << 32) | ((uint64_t)q << 23) | c;
        M.stage_val[at] = c;
      }
      wave_hist_add(eqr, q, act && cnt[u] == thr);
    }
  }
  team_sync<T>();
  // group ranks in order: those whose thr-counted clusters come first take them,
  // until `need` are taken; the groups with a share are listed in stage_val[32..]
  const uint32_t need = nt - ngt;
  if (tid == 0) sh[1] = 0;
  const uint32_t per = (G + T - 1) / T;
  uint32_t local = 0;
  for (uint32_t q = tid * per; q < min(G, tid * per + per); q++) local += eqr[q];
  uint32_t before = team_excl_scan<T>(local, tot, sh + 2);
  for (uint32_t q = tid * per; q < min(G, tid * per + per); q++) {
    const uint32_t e = eqr[q];
    if (e && before < need) {
      const uint32_t at = atomicAdd(&sh[1], 1u);
      M.stage_val[32 + 2 * at] = q;
      M.stage_val[33 + 2 * at] = min(e, need - before);
    }
    before += e;
  }
  team_sync<T>();
  const uint32_t ns = sh[1];
  for (uint32_t si = wv; si < ns; si += NW) {
    const uint32_t q = M.stage_val[32 + 2 * si], quota = M.stage_val[33 + 2 * si];
    const uint32_t g = g_inv[q], cend = M.g_c0[g + 1];
    uint32_t taken = 0;
    for (uint32_t cb = M.g_c0[g]; cb < cend && taken < quota; cb += 64) {
      const uint32_t c = cb + lane;
      const bool eq = c < cend && M.cl_start[c + 1] - M.cl_start[c] == thr;
      const uint64_t m = __ballot(eq);
      if (eq && taken + (uint32_t)__popcll(m & lanemask_lt()) < quota) {
        const uint32_t at = atomicAdd(&sh[0], 1u);
        M.stage_key[at] = ((uint64_t)(0xFFFFFFFFu - thr) << 32) | ((uint64_t)q << 23) | c;
        M.stage_val[at] = c;
      }
      taken += (uint32_t)__popcll(m);
    }
  }
  team_sync<T>();
  for (uint32_t j = nt + tid; j < 32; j += T) M.stage_key[j] = ~0ull;
  team_sync<T>();
  team_bitonic<T>(M.stage_key, M.stage_val, 32, tid);
  for (uint32_t j = tid; j < nt; j += T) M.idx[j] = M.stage_val[j];
  team_sync<T>();
  return true;
}

// Block tier with the segments in group iteration order (group_segment_sort's
// by_rank): the sorted keys M.key[0, n) run in the order find_candidates
// enumerates its clusters (groups in FxHashMap order, diagonals ascending within
// a group, candidates.rs:45-72), so the stable sort by score (:76) keeps, among
// equal counts, the order of the clusters' first keys.  Two walks over the keys
// and no per-cluster arrays:
//  1. every cluster's count: counts 1..CNT_REG tallied in registers per thread,
//     larger ones in an LDS histogram whose last bin holds every count of
//     CNT_BINS - 1 or more; from them T, the count of the nt-th cluster in score
//     order, and gt, the number of clusters counted above T;
//  2. the clusters counted above T, and of those counted exactly T the first
//     nt - gt in key order (a thread's ties offset by the ties of the threads
//     before it), each with its diagonal sum; the nt chosen are then ordered by
//     (count desc, first key) and written as candidates.
// A thread owns the clusters whose first key lies in its chunk and follows each
// to its end, past the chunk when it runs on.  Returns false (uniformly, nothing
// written) when T falls in the saturated bin; the caller's general path then
// handles the read.  Keys: hit keys (CK false), or composite keys (group rank <<
// 32 | biased diagonal - the group's lowest, CK true) with the groups in grp[].
// lds: CNT_BINS + 16 + 2 * TOP_CANDIDATES words and TOP_CANDIDATES doubles.
// Composite keys of the three-launch block tier: bits 0..31 the diagonal offset,
// 32..37 the group rank (together the sort key, CG_KEY_MASK), 40..47 the hit's
// prefilter flags when features are on (carried, never compared).
constexpr uint64_t CG_KEY_MASK = (1ull << 38) - 1;
constexpr uint32_t CG_FLAG_SHIFT = 40;
__device__ __forceinline__ uint32_t cg_rank(uint64_t k) { return (uint32_t)(k >> 32) & 63u; }

constexpr uint32_t CNT_BINS = 1024, CNT_REG = 8;
constexpr size_t TOP20_LDS_BYTES = (CNT_BINS + 16 + 2 * TOP_CANDIDATES) * 4 + TOP_CANDIDATES * 8;
template <int T, bool CK>
__device__ __forceinline__ bool top20_ordered(const uint64_t* key, void* lds, const CgGroup* grp, uint32_t n,
                                              uint32_t L, Cand* out, uint8_t* ncand_out, uint32_t tid) {
  // group of a key, and the part of its diagonal that differs within a group
  auto kgroup = [](uint64_t k) -> uint64_t { return CK ? (uint64_t)cg_rank(k) : (k >> DIAG_BITS); };
  auto kdiag = [](uint64_t k) -> int64_t { return CK ? (int64_t)(uint32_t)k : hit_diag(k); };
  uint32_t* hist = reinterpret_cast<uint32_t*>(lds);  // CNT_BINS counters
  uint32_t* misc = hist + CNT_BINS;  // [0] chosen so far, [1] T, [2] ties wanted, [3] nt, [4] T saturated, [6..] scan
  uint32_t* ocnt = misc + 16;        // chosen clusters: count, first key, diagonal sum
  uint32_t* ostart = ocnt + TOP_CANDIDATES;
  long long* osum = reinterpret_cast<long long*>(ostart + TOP_CANDIDATES);
  static_assert(((CNT_BINS + 16 + 2 * TOP_CANDIDATES) * 4) % 8 == 0, "sums 8-byte aligned");
  for (uint32_t i = tid; i < CNT_BINS; i += T) hist[i] = 0;
  if (tid < 16) misc[tid] = 0;
  team_sync<T>();
  const uint32_t chunk = (n + T - 1) / T;
  const uint32_t c0 = min(n, tid * chunk), c1 = min(n, c0 + chunk);
  // f(count, first key, diagonal sum) for each cluster this thread owns, in order
  auto walk = [&](auto&& f) {
    if (c0 >= c1) return;
    uint64_t kp = c0 > 0 ? key[c0 - 1] : 0ull;
    bool active = false, done = false;
    uint32_t cs = 0, cnt = 0;
    int64_t sum = 0;
    constexpr uint32_t U = 16;  // keys loaded ahead
    for (uint32_t e0 = c0; e0 < n && !done; e0 += U) {
      uint64_t kb[U];
#pragma unroll
      for (uint32_t u = 0; u < U; u++) kb[u] = e0 + u < n ? key[e0 + u] : 0ull;
#pragma unroll
      for (uint32_t u = 0; u < U; u++) {
        const uint32_t e = e0 + u;
        if (done || e >= n) break;
        const uint64_t k = kb[u];
        const bool st = e == 0 || kgroup(k) != kgroup(kp) || kdiag(k) - kdiag(kp) > MERGE_DISTANCE;
        if (st) {
          if (active) f(cnt, cs, sum);
          active = false;
          if (e >= c1) {
            done = true;
            break;
          }
          active = true;
          cs = e;
          cnt = 0;
          sum = 0;
        }
        cnt++;
        sum += kdiag(k);
        kp = k;
      }
    }
    if (active) f(cnt, cs, sum);
  };
  // 1. counts
  uint32_t tally[CNT_REG];
#pragma unroll
  for (uint32_t j = 0; j < CNT_REG; j++) tally[j] = 0;
  walk([&](uint32_t c, uint32_t, int64_t) {
    if (c <= CNT_REG) {
#pragma unroll
      for (uint32_t j = 0; j < CNT_REG; j++) tally[j] += c == j + 1 ? 1u : 0u;
    } else {
      atomicAdd(&hist[min(c, CNT_BINS - 1)], 1u);
    }
  });
#pragma unroll
  for (uint32_t j = 0; j < CNT_REG; j++) {
    const uint32_t v = group_sum_u32<64>(tally[j]);
    if (lane_id() == 0 && v) atomicAdd(&hist[j + 1], v);
  }
  team_sync<T>();
  {
    // bins in descending count order, CNT_BINS / T per thread: the clusters counted
    // above a thread's bins are the scan of the sums of the threads before it
    constexpr uint32_t PB = CNT_BINS / T;
    static_assert(CNT_BINS % T == 0, "whole bins per thread");
    const uint32_t hi = CNT_BINS - 1 - tid * PB;
    uint32_t local = 0;
#pragma unroll
    for (uint32_t j = 0; j < PB; j++) local += hist[hi - j];
    uint32_t C;
    uint32_t above = team_excl_scan<T>(local, C, misc + 6);
    const uint32_t nt = min(C, (uint32_t)TOP_CANDIDATES);
#pragma unroll
    for (uint32_t j = 0; j < PB; j++) {
      const uint32_t h = hist[hi - j];
      if (above < nt && above + h >= nt) {  // the nt-th cluster is counted hi - j
        misc[1] = hi - j;
        misc[2] = nt - above;
        misc[4] = hi - j == CNT_BINS - 1 ? 1u : 0u;
      }
      above += h;
    }
    if (tid == 0) misc[3] = nt;
  }
  team_sync<T>();
  const uint32_t Tc = misc[1], want = misc[2], nt = misc[3];
  if (misc[4]) return false;
  // 2. the chosen clusters
  uint32_t mine = 0;
  if (Tc <= CNT_REG) {
#pragma unroll
    for (uint32_t j = 0; j < CNT_REG; j++) mine += Tc == j + 1 ? tally[j] : 0u;
  } else {
    walk([&](uint32_t c, uint32_t, int64_t) { mine += c == Tc ? 1u : 0u; });
  }
  uint32_t tot;
  uint32_t tb = team_excl_scan<T>(mine, tot, misc + 6);
  walk([&](uint32_t c, uint32_t cs, int64_t sum) {
    bool take = c > Tc;
    if (c == Tc) {
      take = tb < want;
      tb++;
    }
    if (take) {
      const uint32_t at = atomicAdd(&misc[0], 1u);
      if (at < (uint32_t)TOP_CANDIDATES) {
        ocnt[at] = c;
        ostart[at] = cs;
        osum[at] = (long long)sum;
      }
    }
  });
  team_sync<T>();
  if (tid < nt) {
    const uint32_t c = ocnt[tid], cs = ostart[tid];
    uint32_t rk = 0;
    for (uint32_t j = 0; j < nt; j++) rk += (ocnt[j] > c || (ocnt[j] == c && ostart[j] < cs)) ? 1u : 0u;
    uint32_t gk;
    int64_t dsum = (int64_t)osum[tid];
    if constexpr (CK) {  // the diagonals' sum from the offsets' sum: c (lowest - bias) + sum
      const CgGroup& g = grp[cg_rank(key[cs])];
      gk = g.key;
      dsum += (int64_t)c * ((int64_t)g.lo - DIAG_BIAS);
    } else {
      gk = (uint32_t)(key[cs] >> DIAG_BITS);
    }
    const int64_t avg = dsum / (int64_t)c;
    const bool rev = gk & 1u;
    int64_t position = rev ? avg - (int64_t)L + K : avg;
    if (position < 0) position = 0;
    Cand cd;
    cd.pos = (uint64_t)position;
    cd.contig = gk >> 1;
    cd.strand_score = (rev ? 0x80000000u : 0u) | min(c, 0x7FFFFFFFu);
    out[rk] = cd;
  }
  if (tid == 0) *ncand_out = (uint8_t)nt;
  team_sync<T>();
  return true;
}

template <int T, bool FEAT>
__device__ __forceinline__ bool cluster_read(const uint64_t* hits, uint32_t n, uint32_t L, const ClusterMem& M0, Cand* out,
                             uint8_t* ncand_out, uint32_t* sh, const ClusterFeat& F) {
  const uint32_t tid = threadIdx.x % T;
  uint32_t p2 = 1;
  while (p2 < n) p2 <<= 1;
  ClusterMem M = M0;
  // group by (contig, strand), sort each group by diagonal (:36-49): one sort by (group, diagonal)
  bool sorted = false, by_rank = false;
  if constexpr (T != 64) sorted = group_segment_sort<T, FEAT>(hits, n, M, tid, by_rank);
  if (sorted) {
    // per-group arrays in LDS (at most CLUSTER_GT_MAX groups): the group keys
    // and first emission indices by segment as the segmented sort left them, ranks
    // and insertion order in gt_lo, free once the sort is done
    M.g_key = M.gt_cnt;
    M.g_min = M.gt_emin;
    M.g_rank = reinterpret_cast<uint32_t*>(M.gt_lo);
    M.g_sort = M.g_rank + CLUSTER_GT_MAX;
    M.g_c0 = reinterpret_cast<uint32_t*>(M.gt_hi);  // g_c0, then top_by_count's rank tables
    if constexpr (T != 64 && !FEAT) {
      static_assert(sizeof(RSortLds) >= TOP20_LDS_BYTES, "the selection's tables fit the sort's LDS");
      if (by_rank && top20_ordered<T, false>(M.key, M.rsort, nullptr, n, L, out, ncand_out, tid)) return true;
    }
  } else {
    for (uint32_t i = tid; i < p2; i += T) {
      M.key[i] = i < n ? hits[i] : ~0ull;
      M.idx[i] = i;
    }
    team_sync<T>();
    cluster_sort<T>(M, M.key, M.idx, p2, tid);
  }
  if constexpr (FEAT) {
    for (uint32_t i = tid; i < n; i += T)
      if (F.hf != F.hflag) F.hf[i] = F.hflag[i];
  }
  team_sync<T>();
  // group and cluster starts (sliding merge, :51-71), one contiguous chunk per thread
  const uint32_t chunk = (n + T - 1) / T;
  const uint32_t c0 = min(n, tid * chunk), c1 = min(n, c0 + chunk);
  uint32_t ng = 0, nc = 0;
  constexpr uint32_t U = T == 64 ? 1 : 16;  // keys loaded ahead per thread in the block tier
  {
    uint64_t kp = c0 > 0 ? M.key[c0 - 1] : 0;
    for (uint32_t e0 = c0; e0 < c1; e0 += U) {
      uint64_t kb[U];
#pragma unroll
      for (uint32_t u = 0; u < U; u++) kb[u] = e0 + u < c1 ? M.key[e0 + u] : 0ull;
#pragma unroll
      for (uint32_t u = 0; u < U; u++) {
        const uint32_t e = e0 + u;
        if (e >= c1) break;
        const uint64_t k = kb[u];
        const bool gs = e == 0 || (k >> DIAG_BITS) != (kp >> DIAG_BITS);
        ng += gs;
        nc += gs || hit_diag(k) - hit_diag(kp) > MERGE_DISTANCE;
        kp = k;
      }
    }
  }
  uint32_t G, C;
  const uint32_t gbase = team_excl_scan<T>(ng, G, sh);
  const uint32_t cbase = team_excl_scan<T>(nc, C, sh);
  if (G > M.gcap) return false;
  if (!sorted)
    for (uint32_t g = tid; g < G; g += T) M.g_min[g] = 0xFFFFFFFFu;
  if constexpr (FEAT) {
    for (uint32_t c = tid; c < C; c += T) {
      F.f1[c] = 0;
      F.f2[c] = 0;
      F.f4[c] = 15;
    }
  }
  if (tid == 0) {
    M.cl_start[C] = n;
    if (sorted) M.g_c0[G] = C;
  }
  // the only clusters summed atomically are those that can run past a chunk's end:
  // each chunk's last cluster start, zeroed here; every other sum is a plain store
  if (nc) M.cl_sum[cbase + nc - 1] = 0ull;
  team_sync<T>();
  if (c1 > c0) {
    int32_t gid = (int32_t)gbase - 1, cid = (int32_t)cbase - 1;
    uint64_t kp = c0 > 0 ? M.key[c0 - 1] : 0;
    int64_t psum = 0;
    uint32_t pmin = 0xFFFFFFFFu;
    uint32_t fnb = 0, fms = 0, fmn = 15;  // feature sums of the current run of a cluster
    // A cluster that lies wholly inside this thread's chunk is written with plain
    // stores; only the first (begun in an earlier chunk) and the last (it may run
    // on into the next chunk) can have other contributors, and they add atomically
    // to the zeroed sums.
    bool shared = !(c0 == 0 || (M.key[c0] >> DIAG_BITS) != (kp >> DIAG_BITS) ||
                    hit_diag(M.key[c0]) - hit_diag(kp) > MERGE_DISTANCE);
    auto flush = [&](int32_t c, bool atomic) {
      if (atomic) {
        atomicAdd(&M.cl_sum[c], (unsigned long long)psum);
        if constexpr (FEAT) {
          if (fnb) atomicAdd(&F.f1[c], fnb);
          atomicAdd(&F.f2[c], fms);
          atomicMin(&F.f4[c], fmn);
        }
      } else {
        M.cl_sum[c] = (unsigned long long)psum;
        if constexpr (FEAT) {
          F.f1[c] = fnb;
          F.f2[c] = fms;
          F.f4[c] = fmn;
        }
      }
    };
    // emission indices: per-hit flags (FEAT) and, unless the segmented sort
    // found them, each group's first index
    const bool need_idx = FEAT || !sorted;
    for (uint32_t e0 = c0; e0 < c1; e0 += U) {
      uint64_t kb[U];
      uint32_t ib[U];
#pragma unroll
      for (uint32_t u = 0; u < U; u++) {
        kb[u] = e0 + u < c1 ? M.key[e0 + u] : 0ull;
        ib[u] = need_idx && e0 + u < c1 ? M.idx[e0 + u] : 0u;
      }
#pragma unroll
      for (uint32_t u = 0; u < U; u++) {
        const uint32_t e = e0 + u;
        if (e >= c1) break;
        const uint64_t k = kb[u];
        const bool gs = e == 0 || (k >> DIAG_BITS) != (kp >> DIAG_BITS);
        const bool cs = gs || hit_diag(k) - hit_diag(kp) > MERGE_DISTANCE;
        if (gs) {
          if (!sorted && gid >= 0 && pmin != 0xFFFFFFFFu) atomicMin(&M.g_min[gid], pmin);
          pmin = 0xFFFFFFFFu;
          gid++;
          M.g_key[gid] = (uint32_t)(k >> DIAG_BITS);
        }
        if (cs) {
          if (cid >= 0 && e > c0) flush(cid, shared);
          shared = false;
          psum = 0;
          fnb = fms = 0;
          fmn = 15;
          cid++;
          M.cl_start[cid] = e;
          M.cl_gid[cid] = (uint32_t)gid;
          if (sorted && gs) M.g_c0[gid] = (uint32_t)cid;
        }
        psum += hit_diag(k);
        const uint32_t ei = ib[u];
        pmin = min(pmin, ei);
        if constexpr (FEAT) {
          const uint32_t f = F.hf[ei];
          fnb += f >> 7;
          fms += f & 15u;
          fmn = min(fmn, f & 15u);
        }
        kp = k;
      }
    }
    // the last cluster ends at c1 unless the next chunk's first hit continues it
    const bool ends = c1 == n || (M.key[c1] >> DIAG_BITS) != (kp >> DIAG_BITS) ||
                      hit_diag(M.key[c1]) - hit_diag(kp) > MERGE_DISTANCE;
    flush(cid, shared || !ends);
    if (!sorted) atomicMin(&M.g_min[gid], pmin);
  }
  team_sync<T>();
  if constexpr (FEAT) {
    // diagonal span of every cluster, before the keys are overwritten
    for (uint32_t c = tid; c < C; c += T) {
      const uint32_t s0 = M.cl_start[c], e0 = M.cl_start[c + 1];
      F.f3[c] = (uint32_t)min((int64_t)0xFFFFFFF, hit_diag(M.key[e0 - 1]) - hit_diag(M.key[s0]));
    }
    if (tid < 5) F.cnt[tid] = 0;
    team_sync<T>();
    uint32_t qs = 0, q10 = 0, q20 = 0, rep = 0;
    for (uint32_t x = tid; x < F.qlen; x += T) {
      const uint32_t q = F.qual[x];
      qs += q;
      q10 += q < 10;
      q20 += q < 20;
    }
    const uint32_t nwin = L >= (uint32_t)K ? L - K + 1 : 0;
    for (uint32_t x = tid; x < nwin; x += T) rep += F.rcode[x] != 0;
    qs = group_sum_u32<64>(qs);
    q10 = group_sum_u32<64>(q10);
    q20 = group_sum_u32<64>(q20);
    rep = group_sum_u32<64>(rep);
    if (lane_id() == 0) {
      if (qs) atomicAdd(&F.cnt[1], qs);
      if (q10) atomicAdd(&F.cnt[2], q10);
      if (q20) atomicAdd(&F.cnt[3], q20);
      if (rep) atomicAdd(&F.cnt[4], rep);
    }
  }
  team_sync<T>();
  // insertion order of the groups = order of first appearance in the hit list
  if (G <= 1024) {
    for (uint32_t g = tid; g < G; g += T) {
      const uint32_t mg = M.g_min[g];
      uint32_t r = 0;
      for (uint32_t u = 0; u < G; u++) r += M.g_min[u] < mg;
      M.g_sort[r] = g;
    }
    team_sync<T>();
  } else {  // hit keys are no longer needed: sort (first index, group) in key[]
    uint32_t g2 = 1;
    while (g2 < G) g2 <<= 1;
    for (uint32_t g = tid; g < g2; g += T) M.key[g] = g < G ? (((uint64_t)M.g_min[g] << 32) | g) : ~0ull;
    team_sync<T>();
    cluster_sort<T>(M, M.key, nullptr, g2, tid);
    for (uint32_t g = tid; g < G; g += T) M.g_sort[g] = (uint32_t)M.key[g];
    team_sync<T>();
  }
  if (G == 1) {
    if (tid == 0) M.g_rank[0] = 0;
  } else if (G <= 56) {
    if (tid < 64) group_ranks_wave(G, M.g_key, M.g_sort, M.g_rank);
  } else if (tid == 0) {
    group_ranks_table(G, M.g_key, M.g_sort, M.g_rank, M.table, M.tmp);
  }
  team_sync<T>();
  // stable sort by score (:76): key (count desc, group rank, position in the group)
  uint32_t pc = 1;
  while (pc < C) pc <<= 1;
  const uint32_t nt = min(C, (uint32_t)TOP_CANDIDATES);
  bool compact = false;
  if constexpr (T != 64) {
    if (sorted && C > 256 && C < (1u << 23)) compact = top_by_count<T>(M, C, G, nt, tid);
    if (!compact && C > 256) {
      // 32-bit keys (max count - count, group rank, cluster id) when they fit: the
      // cluster id orders a group's clusters as their first sorted hits do
      uint32_t mx = 0;
      for (uint32_t c = tid; c < C; c += T) mx = max(mx, M.cl_start[c + 1] - M.cl_start[c]);
      mx = team_reduce<T>(mx, [](uint32_t a, uint32_t b) { return a > b ? a : b; }, sh);
      const uint32_t cb = 32 - __builtin_clz(mx), gb = G > 1 ? 32 - __builtin_clz(G - 1) : 0,
                     sb = 32 - __builtin_clz(C - 1);
      if (cb + gb + sb <= 32) {
        uint32_t* k32 = reinterpret_cast<uint32_t*>(M.key);
        uint32_t kor = 0, kand = 0xFFFFFFFFu;
        for (uint32_t c = tid; c < C; c += T) {
          const uint32_t k = ((mx - (M.cl_start[c + 1] - M.cl_start[c])) << (gb + sb)) |
                             (M.g_rank[M.cl_gid[c]] << sb) | c;
          k32[c] = k;
          kor |= k;
          kand &= k;
        }
        kor = team_reduce<T>(kor, [](uint32_t a, uint32_t b) { return a | b; }, sh);
        kand = team_reduce<T>(kand, [](uint32_t a, uint32_t b) { return a & b; }, sh);
        top_clusters32<T>(M, k32, C, nt, kor, kand, tid);
        compact = true;
      }
    }
  }
  if (!compact) {
    for (uint32_t c = tid; c < pc; c += T) {
      uint64_t sk = ~0ull;
      if (c < C) {
        const uint32_t s = M.cl_start[c], cnt = M.cl_start[c + 1] - s;
        sk = ((uint64_t)(0x3FFFFFu - min(cnt, 0x3FFFFFu)) << 42) |
             ((uint64_t)min(M.g_rank[M.cl_gid[c]], 0xFFFFFu) << 22) | (uint64_t)s;
      }
      M.key[c] = sk;
      M.idx[c] = c;
    }
    team_sync<T>();
    if constexpr (T != 64) {
      if (C > 256) top_clusters<T>(M, C, nt, tid);
      else cluster_sort<T>(M, M.key, M.idx, pc, tid);
    } else {
      cluster_sort<T>(M, M.key, M.idx, pc, tid);
    }
  }
  if constexpr (FEAT) {
    const uint32_t top = M.cl_start[M.idx[0] + 1] - M.cl_start[M.idx[0]];  // the largest count
    uint32_t ties = 0;
    for (uint32_t c = tid; c < C; c += T) ties += M.cl_start[c + 1] - M.cl_start[c] == top;
    ties = group_sum_u32<64>(ties);
    if (lane_id() == 0 && ties) atomicAdd(&F.cnt[0], ties);
    team_sync<T>();
    if (tid < nt) {
      const uint32_t c = M.idx[tid];
      PfCluster pc;
      pc.cnt = M.cl_start[c + 1] - M.cl_start[c];
      pc.nb = F.f1[c];
      pc.ms = F.f2[c];
      pc.mn = F.f4[c];
      pc.span = F.f3[c];
      pc.rev = (M.g_key[M.cl_gid[c]] & 1u) != 0;
      PfRead pr;
      pr.c0 = M.cl_start[M.idx[0] + 1] - M.cl_start[M.idx[0]];
      pr.c1 = C > 1 ? M.cl_start[M.idx[1] + 1] - M.cl_start[M.idx[1]] : 0u;
      pr.ties = F.cnt[0];
      pr.qsum = F.cnt[1];
      pr.q10 = F.cnt[2];
      pr.q20 = F.cnt[3];
      pr.rep = F.cnt[4];
      pr.qlen = F.qlen;
      pr.C = C;
      pr.n = n;
      pr.L = L;
      pr.G = G;
      pf_feature_row(F.out + tid * 32, tid, pc, pr);
    }
  }
  if (tid < nt) {
    const uint32_t c = M.idx[tid];
    const uint32_t s = M.cl_start[c], cnt = M.cl_start[c + 1] - s;
    const uint32_t gk = M.g_key[M.cl_gid[c]];
    const int64_t avg = (int64_t)M.cl_sum[c] / (int64_t)cnt;
    const bool rev = gk & 1u;
    int64_t position = rev ? avg - (int64_t)L + K : avg;
    if (position < 0) position = 0;
    Cand cd;
    cd.pos = (uint64_t)position;
    cd.contig = gk >> 1;
    cd.strand_score = (rev ? 0x80000000u : 0u) | min(cnt, 0x7FFFFFFFu);
    out[tid] = cd;
  }
  if (tid == 0) *ncand_out = (uint8_t)nt;
  team_sync<T>();
  return true;
}

// Wave tier: one wave per read, all arrays in LDS.
constexpr uint32_t CL_NCAP = CLUSTER_LDS_CAP;  // hits
constexpr uint32_t CL_GCAP = 128;              // groups
constexpr uint32_t CL_SMALL = 128;             // hits of the small wave tier (most reads)
template <uint32_t NC>
struct ClusterLds {
  uint64_t key[NC];
  unsigned long long cl_sum[NC];
  uint32_t idx[NC];
  uint32_t cl_start[NC + 1];
  uint32_t cl_gid[NC];
  uint32_t g_key[CL_GCAP], g_min[CL_GCAP], g_rank[CL_GCAP], g_sort[CL_GCAP];
  int32_t table[2 * CL_GCAP], tmp[CL_GCAP];
};

// Wave tier: one wave per read, every array in LDS for up to NC hits.  The
// reads of `list` (all reads when null; count at list_n[0]) that hold more
// hits, or more groups than CL_GCAP, go to (defer, defer_n) for the next tier.
// Two sizes run in turn: NC = 128 takes most reads at a small LDS footprint
// (more waves per CU), NC = 512 the rest.
template <bool FEAT, uint32_t NC>
__global__ __launch_bounds__(64) void k_cluster(ClusterArgs a, const uint32_t* list, const uint32_t* list_n,
                                                uint32_t* defer, uint32_t* defer_n) {
  __shared__ ClusterLds<NC> S;
  __shared__ uint32_t FX[FEAT ? 4 * NC + 8 + NC / 4 : 1];
  const ClusterMem M{S.key, S.idx, S.cl_sum, S.cl_start, S.cl_gid, S.g_key, S.g_min, S.g_rank, S.g_sort, S.table,
                     S.tmp, CL_GCAP, nullptr, nullptr, nullptr, nullptr, nullptr};
  const uint32_t total = list ? list_n[0] : a.R.n;
  const uint32_t lane = (uint32_t)lane_id();
  // 64 list entries a round: their hit counts read at once, the reads over NC
  // hits deferred with one atomic for the wave, the rest clustered one by one
  for (uint32_t l0 = blockIdx.x * 64; l0 < total; l0 += gridDim.x * 64) {
    const uint32_t li = l0 + lane;
    const uint32_t rl = li < total ? (list ? list[li] : li) : 0u;
    const uint32_t nl = li < total ? a.nhits[rl] : 0u;
    if (li < total && nl == 0) a.ncand[rl] = 0;
    const uint64_t over = __ballot(nl > NC);
    if (over) {
      uint32_t base = 0;
      if (lane == 0) base = atomicAdd(defer_n, (uint32_t)__popcll(over));
      base = (uint32_t)__builtin_amdgcn_readfirstlane((int)base);
      if (nl > NC) defer[base + (uint32_t)__popcll(over & lanemask_lt())] = rl;
    }
    uint64_t work = __ballot(nl > 0 && nl <= NC);
    while (work) {
      const uint32_t src = (uint32_t)__builtin_ctzll(work);
      work &= work - 1;
      const uint32_t r = (uint32_t)__shfl((int)rl, (int)src, 64), n = (uint32_t)__shfl((int)nl, (int)src, 64);
      const uint32_t L = (uint32_t)(a.R.seq_off[r + 1] - a.R.seq_off[r]);
      ClusterFeat F{};
      if (FEAT) {
        const uint64_t so = a.R.seq_off[r], qo = a.R.qual_off[r];
        F.hflag = a.hflag + a.hoff[r];
        F.qual = a.R.qual + qo;
        F.qlen = (uint32_t)(a.R.qual_off[r + 1] - qo);
        F.rcode = a.rcode + so;
        F.out = a.feat + (uint64_t)r * TOP_CANDIDATES * MLP_IN;
        F.f1 = FX;
        F.f2 = FX + NC;
        F.f3 = FX + 2 * NC;
        F.f4 = FX + 3 * NC;
        F.cnt = FX + 4 * NC;
        F.hf = (uint8_t*)(FX + 4 * NC + 8);
      }
      const bool done = cluster_read<64, FEAT>(a.hits + a.hoff[r], n, L, M, a.cands + (uint64_t)r * TOP_CANDIDATES,
                                               a.ncand + r, nullptr, F);
      if (!done && lane == 0) defer[atomicAdd(defer_n, 1u)] = r;
    }
  }
}

// Block tier: reads with more hits (or groups) than the wave tier holds; a
// block of 256 threads per read, arrays in per-block global scratch of
// cluster_scratch_words(max hits) words.
template <bool FEAT>
__global__ __launch_bounds__(256) void k_cluster_tier2(ClusterArgs a, const uint32_t* list, uint32_t nlist) {
  __shared__ uint32_t sh[4];
  __shared__ uint64_t stage_key[CLUSTER_STAGE];
  __shared__ uint32_t stage_val[CLUSTER_STAGE];
  __shared__ uint32_t gt_key[CLUSTER_GT], gt_cnt[CLUSTER_GT], gt_misc[2 + 256 / 64];
  __shared__ unsigned long long gt_lo[CLUSTER_GT], gt_hi[CLUSTER_GT];
  __shared__ uint32_t gt_emin[CLUSTER_GT];
  __shared__ RSortLds rsort;
  for (uint32_t li = blockIdx.x; li < nlist; li += gridDim.x) {
    const uint32_t r = list[li];
    const uint32_t n = a.nhits[r];
    const uint32_t L = (uint32_t)(a.R.seq_off[r + 1] - a.R.seq_off[r]);
    uint32_t p2 = 1;
    while (p2 < n) p2 <<= 1;
    uint64_t* base = a.scratch + (uint64_t)blockIdx.x * a.scratch_per_block;
    ClusterMem M;
    M.key = base;
    M.cl_sum = (unsigned long long*)(base + p2);
    uint32_t* w = (uint32_t*)(base + 2 * (uint64_t)p2);
    M.idx = w;
    M.cl_start = w + p2;
    M.cl_gid = w + 2 * (uint64_t)p2 + 1;
    M.g_key = w + 3 * (uint64_t)p2 + 1;
    M.g_min = M.g_key + p2;
    M.g_rank = M.g_min + p2;
    M.g_sort = M.g_rank + p2;
    M.tmp = (int32_t*)(M.g_sort + p2);
    M.table = M.tmp + p2;  // 4 * p2 buckets
    M.gcap = p2;
    M.stage_key = stage_key;
    M.stage_val = stage_val;
    M.gt_key = gt_key;
    M.gt_cnt = gt_cnt;
    M.gt_misc = gt_misc;
    M.gt_lo = gt_lo;
    M.gt_hi = gt_hi;
    M.rsort = &rsort;
    M.gt_emin = gt_emin;
    ClusterFeat F{};
    if (FEAT) {
      const uint64_t so = a.R.seq_off[r], qo = a.R.qual_off[r];
      uint32_t* fx = (uint32_t*)(M.table + 4 * p2);
      F.hflag = a.hflag + a.hoff[r];
      F.hf = const_cast<uint8_t*>(F.hflag);  // read in place
      F.qual = a.R.qual + qo;
      F.qlen = (uint32_t)(a.R.qual_off[r + 1] - qo);
      F.rcode = a.rcode + so;
      F.out = a.feat + (uint64_t)r * TOP_CANDIDATES * MLP_IN;
      F.f1 = fx;
      F.f2 = fx + p2;
      F.f3 = fx + 2 * p2;
      F.f4 = fx + 3 * p2;
      F.cnt = fx + 4 * p2;
    }
    cluster_read<256, FEAT>(a.hits + a.hoff[r], n, L, M, a.cands + (uint64_t)r * TOP_CANDIDATES, a.ncand + r, sh,
                            F);
  }
}

void launch_cluster(const ClusterArgs& a, hipStream_t s) {
  if (a.R.n == 0) return;
  const uint32_t blocks = (a.R.n + 63) / 64 < 65536 ? (a.R.n + 63) / 64 : 65536;  // 64 reads a block and round
  const uint32_t mid_blocks = blocks < 8192 ? blocks : 8192;  // the count is on the device
  if (a.feat) {
    hipLaunchKernelGGL((k_cluster<true, CL_SMALL>), dim3(blocks), dim3(64), 0, s, a, nullptr, nullptr, a.mid, a.nmid);
    hipLaunchKernelGGL((k_cluster<true, CL_NCAP>), dim3(mid_blocks), dim3(64), 0, s, a, a.mid, a.nmid, a.tier2,
                       a.ntier2);
  } else {
    hipLaunchKernelGGL((k_cluster<false, CL_SMALL>), dim3(blocks), dim3(64), 0, s, a, nullptr, nullptr, a.mid, a.nmid);
    hipLaunchKernelGGL((k_cluster<false, CL_NCAP>), dim3(mid_blocks), dim3(64), 0, s, a, a.mid, a.nmid, a.tier2,
                       a.ntier2);
  }
}
uint64_t cluster_scratch_words(uint32_t max_hits) {
  uint64_t p2 = 1;
  while (p2 < max_hits) p2 <<= 1;
  // key + cl_sum: 2 p2 words; idx, cl_start (+1), cl_gid, g_key, g_min, g_rank, g_sort, tmp
  // (p2 u32 each), table (4 p2 u32) and the feature arrays (4 p2 + 8 u32): 16 p2 + 9 u32
  return 2 * p2 + 8 * p2 + 16;
}
void launch_cluster_tier2(const ClusterArgs& a, const uint32_t* list, uint32_t n, uint32_t blocks, hipStream_t s) {
  if (n == 0) return;
  if (a.feat) hipLaunchKernelGGL(k_cluster_tier2<true>, dim3(blocks), dim3(256), 0, s, a, list, n);
  else hipLaunchKernelGGL(k_cluster_tier2<false>, dim3(blocks), dim3(256), 0, s, a, list, n);
}

// ---------------------------------------------------------------------------
// Block tier in three launches, for reads of at most CG_GMAX groups whose
// diagonals each span less than 2^32 - 1 (at genome scale every read): each phase
// runs at the occupancy its own registers and LDS allow, where one kernel for all
// of them would run every phase at the lowest.
//  k_cg_group:  the group table (group_table_count), the groups ranked in
//               FxHashMap iteration order from their first emission indices,
//               segments laid out by rank, every hit scattered into its
//               segment as a composite key (rank << 32 | biased diagonal - the
//               group's lowest; with features, the hit's flags in bits 40..47);
//  k_cg_sort:   runs of whole segments of up to RSORT_TILE keys sorted by one
//               bucket pass over their offsets (cg_bucket_sort), or, when a
//               bucket holds over 512 keys, by one block radix sort of the
//               composite keys (8-bit digits over the 38 bits that vary); a
//               larger segment tile by tile on its low 32 bits, the tiles then
//               merged pairwise (block_merge) through a second buffer;
//  k_cg_select: the top 20 from a bitmap of cluster starts (top20_bitmap; the
//               key walk top20_ordered for reads of 2^17 hits or more), and
//               with features each candidate's feature row.
// A read that fails the conditions (or whose 20th count lies in the selection's
// saturated bin) is listed for k_cluster_tier2, its hits left intact.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_cg_group(ClusterArgs a, CgArgs g, const uint32_t* list, uint32_t nlist) {
  __shared__ uint32_t gt_key[CLUSTER_GT], gt_cnt[CLUSTER_GT], gt_misc[2 + 256 / 64];
  __shared__ unsigned long long gt_lo[CLUSTER_GT], gt_hi[CLUSTER_GT];
  __shared__ uint32_t gt_emin[CLUSTER_GT];  // first emission index per slot, then the slot's rank
  __shared__ uint64_t sk[64];               // (first emission index << 32 | slot), insertion order
  __shared__ uint32_t sv[192];              // group key and insertion list for the ranks, then ranks
  __shared__ uint32_t rslot[64];            // slot of each rank
  __shared__ uint32_t ctab[SEED_MAXCH];     // the read's hit chunks
  __shared__ uint32_t fail;
  ClusterMem M{};
  M.gt_key = gt_key;
  M.gt_cnt = gt_cnt;
  M.gt_misc = gt_misc;
  M.gt_lo = gt_lo;
  M.gt_hi = gt_hi;
  M.gt_emin = gt_emin;
  const uint32_t tid = threadIdx.x;
  for (uint32_t li = blockIdx.x; li < nlist; li += gridDim.x) {
    const uint32_t r = list[li];
    const uint32_t n = a.nhits[r];
    const uint64_t hb = a.hoff[r];
    const bool chunked = a.chunk_tab && a.chunk_tab[(uint64_t)r * SEED_MAXCH] != SEED_CONTIGUOUS;
    if (chunked)
      for (uint32_t j = tid; j < SEED_MAXCH; j += 256) ctab[j] = a.chunk_tab[(uint64_t)r * SEED_MAXCH + j];
    if (tid < 64) sk[tid] = ~0ull;
    if (tid == 0) fail = 0;
    __syncthreads();
    auto hidx = [&](uint32_t i) -> uint64_t {
      return chunked ? ((uint64_t)ctab[i >> SEED_CH_BITS] << SEED_CH_BITS) + (i & ((1u << SEED_CH_BITS) - 1)) : hb + i;
    };
    auto hit = [&](uint32_t i) -> uint64_t { return a.hits[hidx(i)]; };
    group_table_count<256>(hit, n, M, tid);
    const uint32_t G = gt_misc[0];
    if (G <= CG_GMAX) {
      for (uint32_t i = tid; i < CLUSTER_GT; i += 256)
        if (gt_key[i] != 0xFFFFFFFFu) {
          const uint32_t at = atomicAdd(&gt_misc[1], 1u);
          sk[at] = ((uint64_t)gt_emin[i] << 32) | i;
        }
      __syncthreads();
      if (tid < 64) {  // wave 0: insertion order, ranks, segment starts
        const uint32_t lane = tid;
        wave_bitonic_reg<1>(sk, nullptr, 64);
        const uint32_t slt = (uint32_t)sk[lane];
        sv[lane] = lane < G ? gt_key[slt] : 0u;
        sv[64 + lane] = lane;
        wave_sync();
        group_ranks_wave(G, sv, sv + 64, sv + 128);
        wave_sync();
        if (lane < G) rslot[sv[128 + lane]] = slt;
        wave_sync();
        uint32_t cnt = 0, sl = 0, gkey = 0;
        unsigned long long lo = 0, hi = 0;
        if (lane < G) {
          sl = rslot[lane];
          cnt = gt_cnt[sl];
          lo = gt_lo[sl];
          hi = gt_hi[sl];
          gkey = gt_key[sl];
        }
        const uint32_t start = wave_incl_scan(cnt) - cnt;
        const bool wide = lane < G && hi - lo >= 0xFFFFFFFFull;
        CgGroup* cg = g.cg + (uint64_t)li * (CG_GMAX + 1);
        if (lane < G) {
          cg[lane].lo = lo;
          cg[lane].key = gkey;
          cg[lane].start = start;
          cg[lane].rbits = hi - lo ? 64u - (uint32_t)__builtin_clzll(hi - lo) : 0u;
          gt_cnt[sl] = start;  // the slot's scatter cursor
          gt_emin[sl] = lane;  // the slot's rank
        }
        if (lane == 0) {
          cg[G].lo = 0;
          cg[G].key = 0;
          cg[G].start = n;
          cg[G].rbits = 0;
        }
        if (__ballot(wide) && lane == 0) fail = 1;
      }
    } else if (tid == 0) {
      fail = 1;
    }
    __syncthreads();
    if (fail) {
      if (tid == 0) {
        g.cg_n[li] = 0;
        g.fb_list[atomicAdd(g.fb_n, 1u)] = r;
      }
      __syncthreads();
      continue;
    }
    if (tid == 0) {
      g.cg_n[li] = G;
      sk[0] = atomicAdd(g.ck_next, (unsigned long long)n);  // this read's region in ck and tmp
      g.ck_off[li] = sk[0];
      sk[1] = atomicAdd(g.bm_next, (unsigned long long)((n >> 6) + 1));  // and its start bitmap
      g.bm_off[li] = sk[1];
    }
    __syncthreads();
    for (uint32_t w = tid; w <= (n >> 6); w += 256) g.bm[sk[1] + w] = 0;
    uint64_t* ck = g.ck + sk[0];
    constexpr uint32_t U = 8;
    for (uint32_t i0 = tid; i0 < n; i0 += U * 256) {
      uint64_t kb[U];
      uint32_t fb[U];  // prefilter flags (features on), carried in the keys' bits 40..47
#pragma unroll
      for (uint32_t u = 0; u < U; u++) {
        const bool in = i0 + u * 256 < n;
        const uint64_t x = in ? hidx(i0 + u * 256) : 0ull;
        kb[u] = in ? a.hits[x] : 0ull;
        fb[u] = in && a.feat ? (uint32_t)a.hflag[x] : 0u;
      }
#pragma unroll
      for (uint32_t u = 0; u < U; u++) {
        const uint32_t i = i0 + u * 256;
        if (i >= n) break;
        const uint64_t k = kb[u];
        const uint32_t gg = (uint32_t)(k >> DIAG_BITS);
        const unsigned long long d = k & ((1ull << DIAG_BITS) - 1);
        const uint32_t g0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)gg);
        const uint64_t act = __ballot(true);
        uint32_t pos, sl;
        if (__ballot(gg == g0) == act) {
          // one group across the active lanes (a prefix of the wave): one cursor update
          sl = gt_slot(g0);
          while (gt_key[sl] != g0) sl = (sl + 1) & (CLUSTER_GT - 1);
          uint32_t base = 0;
          if (lane_id() == 0) base = atomicAdd(&gt_cnt[sl], (uint32_t)__popcll(act));
          pos = (uint32_t)__builtin_amdgcn_readfirstlane((int)base) + lane_id();
        } else {
          sl = gt_slot(gg);
          while (gt_key[sl] != gg) sl = (sl + 1) & (CLUSTER_GT - 1);
          pos = atomicAdd(&gt_cnt[sl], 1u);
        }
        ck[pos] = ((uint64_t)fb[u] << CG_FLAG_SHIFT) | ((uint64_t)gt_emin[sl] << 32) | (uint64_t)(uint32_t)(d - gt_lo[sl]);
      }
    }
    __syncthreads();
  }
}

// One block radix sort of cnt <= 256 IPT composite keys in place (the padding
// keys sort last: every real key is below 2^end_bit - 1).
template <int IPT>
using CgSort = rocprim::block_radix_sort<uint64_t, 256, IPT, rocprim::empty_type, 1, 1, 8,
                                         rocprim::block_radix_rank_algorithm::match>;
union CgSortLds {
  typename CgSort<16>::storage_type s16;
  typename CgSort<8>::storage_type s8;
  typename CgSort<4>::storage_type s4;
  typename CgSort<2>::storage_type s2;
  uint64_t raw[MERGE_CHUNK];  // block_merge's windows
};
template <int IPT>
__device__ __forceinline__ void cg_run_sort(uint64_t* key, uint32_t cnt, uint32_t end_bit, CgSortLds& lds,
                                            uint32_t tid) {
  uint64_t k[IPT];
#pragma unroll
  for (int j = 0; j < IPT; j++) {
    const uint32_t pos = (uint32_t)j * 256 + tid;
    k[j] = pos < cnt ? key[pos] : ~0ull;
  }
  CgSort<IPT>().sort_to_striped(k, *reinterpret_cast<typename CgSort<IPT>::storage_type*>(&lds), 0u, end_bit);
#pragma unroll
  for (int j = 0; j < IPT; j++) {  // striped: thread tid holds ranks j * 256 + tid
    const uint32_t pos = (uint32_t)j * 256 + tid;
    if (pos < cnt) key[pos] = k[j];
  }
  __syncthreads();
}

// Bitonic sort of the s <= 64 E keys at key[0, s) by one wave, in registers
// (element e * 64 + lane in slot e of a lane), padded with ~0; keys compare
// under the mask cm.
template <int E>
__device__ __forceinline__ void wave_bitonic_n(uint64_t* key, uint32_t s, uint64_t cm) {
  const uint32_t lane = (uint32_t)lane_id();
  uint32_t p2 = 1;
  while (p2 < s) p2 <<= 1;
  uint64_t kr[E];
#pragma unroll
  for (int e = 0; e < E; e++) {
    const uint32_t i = (uint32_t)e * 64u + lane;
    kr[e] = i < s ? key[i] : ~0ull;
  }
  for (uint32_t k = 2; k <= p2; k <<= 1) {
    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
      if (j >= 64) {
        const uint32_t js = j >> 6;
#pragma unroll
        for (int e = 0; e < E; e++) {
          const uint32_t ep = (uint32_t)e ^ js;
          if ((uint32_t)e < ep && ep < (uint32_t)E) {
            const bool up = (((uint32_t)e * 64u) & k) == 0;
            const uint64_t x = kr[e], y = kr[ep];
            if (((x & cm) > (y & cm)) == up) {
              kr[e] = y;
              kr[ep] = x;
            }
          }
        }
      } else {
        const bool lower = (lane & j) == 0;
#pragma unroll
        for (int e = 0; e < E; e++) {
          const uint32_t i = (uint32_t)e * 64u + lane;
          const uint64_t pk = (uint64_t)__shfl_xor((long long)kr[e], (int)j, 64);
          const bool up = (i & k) == 0;
          const bool sw = lower ? (((kr[e] & cm) > (pk & cm)) == up) : (((pk & cm) > (kr[e] & cm)) == up);
          if (sw) kr[e] = pk;
        }
      }
    }
  }
#pragma unroll
  for (int e = 0; e < E; e++) {
    const uint32_t i = (uint32_t)e * 64u + lane;
    if (i < s) key[i] = kr[e];
  }
}

// The cluster-start bits of the sorted keys at read positions [a0, a0 + cnt), a run
// of whole segments (get(i): the key at a0 + i): bit p set when key p starts a
// group or lies more than MERGE_DISTANCE above key p - 1 (the first key of the
// run starts its group).  One wave per bitmap word; the words the run shares with
// its neighbours are ORed in (the bitmap starts zeroed), the rest stored.
template <class Get>
__device__ __forceinline__ void cg_start_bits(Get get, uint32_t a0, uint32_t cnt, uint64_t* bmg, uint32_t tid) {
  constexpr uint32_t U = 4;  // words per wave and round: their loads issued together
  const uint32_t lane = tid & 63u, wave = tid >> 6;
  const uint32_t w0 = a0 >> 6, w1 = (a0 + cnt - 1) >> 6;
  for (uint32_t wb = w0 + wave * U; wb <= w1; wb += 4 * U) {
    uint64_t k[U], kp[U];
    bool in[U];
#pragma unroll
    for (uint32_t u = 0; u < U; u++) {
      const uint32_t p = ((wb + u) << 6) + lane;
      in[u] = wb + u <= w1 && p >= a0 && p < a0 + cnt;
      const uint32_t i = p - a0;
      k[u] = in[u] ? get(i) : 0ull;
      kp[u] = in[u] && i > 0 ? get(i - 1) : 0ull;
    }
#pragma unroll
    for (uint32_t u = 0; u < U; u++) {
      const uint32_t w = wb + u;
      const uint32_t p = (w << 6) + lane;
      const bool sb = in[u] && (p == a0 || cg_rank(k[u]) != cg_rank(kp[u]) ||
                                (uint32_t)k[u] - (uint32_t)kp[u] > (uint32_t)MERGE_DISTANCE);
      const uint64_t word = __ballot(sb);
      if (lane == 0 && w <= w1) {
        if (w > w0 && w < w1) bmg[w] = word;
        else if (word) atomicOr(reinterpret_cast<unsigned long long*>(bmg + w), (unsigned long long)word);
      }
    }
  }
}

// A run of whole segments (ranks q0..q1-1, cnt <= CG_BTILE keys at key[0, cnt))
// sorted by one counting pass over buckets: segment q gets the largest power of two
// of buckets not above its key count, each a 2^sh_q-wide slice of its offsets, so
// bucket order is key order and a bucket holds one or two keys; every key is counted
// into its bucket, the buckets' ends come from a block scan, every key is placed
// in LDS, and then each bucket of two to CG_BSMALL keys is put in order by one
// lane (a sorting network in registers, bucket_net), each larger one of up to 512
// by a wave (bitonic in registers).  One segment's tile of a larger segment is
// sorted the same way (one_cnt: its key count, q1 = q0 + 1), its output to `out`
// when that is set (without bmg).  With bmg set, the
// run's cluster-start bits (read positions from a0 on) are written from LDS.  Returns false, with
// nothing written, when a bucket holds more than 512 keys (the caller then sorts
// the run by radix).  The keys of one bucket share their rank, so within a bucket
// they compare by their offsets (bits 0..31) alone.
constexpr uint32_t CG_BTILE = 4096, CG_BSMALL = 16, CG_NBMAX = CG_BTILE;
static_assert(RSORT_TILE <= CG_BTILE, "a radix tile fits the bucket sort");
// The c <= N keys at p[0, c) put in order by one lane: loaded together, a bitonic
// network over N registers (positions from c on padded with offset 0xFFFFFFFF,
// above every offset of this tier), the first c stored back.  c == 0: no access.
template <int N>
__device__ __forceinline__ void bucket_net(uint64_t* p, uint32_t c) {
  uint64_t r[N];
#pragma unroll
  for (int i = 0; i < N; i++) r[i] = (uint32_t)i < c ? p[i] : 0xFFFFFFFFull;
#pragma unroll
  for (int k = 2; k <= N; k <<= 1) {
#pragma unroll
    for (int j = k >> 1; j > 0; j >>= 1) {
#pragma unroll
      for (int i = 0; i < N; i++) {
        const int l = i ^ j;
        if (l > i) {
          const bool up = (i & k) == 0;
          const uint64_t x = r[i], y = r[l];
          if (((uint32_t)x > (uint32_t)y) == up) {
            r[i] = y;
            r[l] = x;
          }
        }
      }
    }
  }
#pragma unroll
  for (int i = 0; i < N; i++)
    if ((uint32_t)i < c) p[i] = r[i];
}
struct CgBucketLds {
  uint64_t keys[CG_BTILE];
  uint32_t end[CG_NBMAX];   // bucket counts, then (scan) each bucket's end
  uint32_t base[CG_GMAX + 1], sh[CG_GMAX + 1];
  uint32_t big[CG_BTILE / CG_BSMALL];  // buckets of more than CG_BSMALL keys
  uint32_t wq[4][128];                 // per wave: buckets of 5..CG_BSMALL keys waiting for a network
  uint32_t nbig, fail, scan[8];
};
__device__ __forceinline__ bool cg_bucket_sort(uint64_t* key, uint32_t cnt, uint32_t q0, uint32_t q1, const uint32_t* st,
                                               const uint32_t* rng_bits, CgBucketLds& B, uint32_t tid,
                                               uint32_t one_cnt = 0, uint64_t* bmg = nullptr, uint32_t a0 = 0,
                                               uint64_t* bits = nullptr, uint64_t* out = nullptr) {
  constexpr uint32_t IPT = CG_BTILE / 256;
  // buckets per segment
  if (tid == 0) {
    uint32_t nb = 0;
    for (uint32_t q = q0; q < q1; q++) {
      const uint32_t c = one_cnt ? one_cnt : st[q + 1] - st[q];
      uint32_t lg = 0;  // 2^lg <= c buckets: about one or two keys each
      while ((2u << lg) <= c) lg++;
      B.base[q - q0] = nb;
      B.sh[q - q0] = rng_bits[q] > lg ? rng_bits[q] - lg : 0u;
      nb += 1u << (rng_bits[q] > lg ? lg : rng_bits[q]);
    }
    B.base[q1 - q0] = nb;
    B.nbig = 0;
    B.fail = 0;
  }
  __syncthreads();
  const uint32_t NB = B.base[q1 - q0];
  for (uint32_t b = tid; b < NB; b += 256) B.end[b] = 0;
  __syncthreads();
  uint64_t kk[IPT];
  uint32_t bk[IPT];  // bucket << 12 | the key's rank among its bucket's keys (both < 4096)
#pragma unroll
  for (uint32_t j = 0; j < IPT; j++) {
    const uint32_t pos = j * 256 + tid;
    kk[j] = pos < cnt ? key[pos] : 0ull;
  }
#pragma unroll
  for (uint32_t j = 0; j < IPT; j++) {
    const uint32_t pos = j * 256 + tid;
    if (pos < cnt) {
      const uint32_t ql = cg_rank(kk[j]) - q0;
      const uint32_t b = B.base[ql] + (uint32_t)((uint64_t)(uint32_t)kk[j] >> B.sh[ql]);  // sh <= 32
      bk[j] = (b << 12) | atomicAdd(&B.end[b], 1u);
    }
  }
  __syncthreads();
  {  // exclusive scan of the counts: bucket starts
    const uint32_t per = (NB + 255) / 256, b0 = min(NB, tid * per), b1 = min(NB, b0 + per);
    uint32_t local = 0;
    for (uint32_t b = b0; b < b1; b++) local += B.end[b];
    uint32_t tot;
    uint32_t run = team_excl_scan<256>(local, tot, B.scan);
    for (uint32_t b = b0; b < b1; b++) {
      const uint32_t c = B.end[b];
      if (c > CG_BSMALL) {
        if (c > 512) B.fail = 1;
        else B.big[atomicAdd(&B.nbig, 1u)] = b;
      }
      B.end[b] = run;
      run += c;
    }
  }
  __syncthreads();
  if (B.fail) {  // every thread has read the flag before the caller reuses the LDS
    __syncthreads();
    return false;
  }
  // place: bucket start + the key's rank
#pragma unroll
  for (uint32_t j = 0; j < IPT; j++) {
    const uint32_t pos = j * 256 + tid;
    if (pos < cnt) B.keys[B.end[bk[j] >> 12] + (bk[j] & 4095u)] = kk[j];
  }
  __syncthreads();
  // B.end[b] is bucket b's start; its end is the next bucket's start (cnt for the
  // last).  Buckets of 2..4 keys: a 2- or 4-key network per lane, sized per wave;
  // buckets of 5..CG_BSMALL keys: queued per wave and put in order 64 at a time
  // with every lane busy (an 8- or 16-key network, sized per wave).
  {
    const uint32_t lane = tid & 63u;
    uint32_t* wq = B.wq[tid >> 6];
    uint32_t qn = 0;  // wave-uniform
    auto drain = [&](uint32_t k) {  // the first k <= 64 queued buckets, one per lane
      wave_sync();
      uint32_t s0 = 0, c = 0;
      if (lane < k) {
        const uint32_t b = wq[lane];
        s0 = B.end[b];
        c = (b + 1 < NB ? B.end[b + 1] : cnt) - s0;
      }
      if (__ballot(c > 8)) bucket_net<16>(B.keys + s0, c);
      else bucket_net<8>(B.keys + s0, c);
    };
    for (uint32_t b0 = 0; b0 < NB; b0 += 256) {
      const uint32_t b = b0 + tid;
      uint32_t s0 = 0, c = 0;
      if (b < NB) {
        s0 = B.end[b];
        c = (b + 1 < NB ? B.end[b + 1] : cnt) - s0;
      }
      const bool mid = c >= 5 && c <= CG_BSMALL;
      const uint64_t mm = __ballot(mid);
      if (mid) wq[qn + (uint32_t)__popcll(mm & lanemask_lt())] = b;
      qn += (uint32_t)__popcll(mm);
      const uint32_t cs = c >= 2 && c <= 4 ? c : 0u;
      if (__ballot(cs > 2)) bucket_net<4>(B.keys + s0, cs);
      else if (__ballot(cs != 0)) bucket_net<2>(B.keys + s0, cs);
      if (qn >= 64) {
        drain(64);
        const uint32_t rest = qn - 64, v = lane < rest ? wq[64 + lane] : 0u;
        wave_sync();
        if (lane < rest) wq[lane] = v;
        qn = rest;
      }
    }
    if (qn) drain(qn);
  }
  for (uint32_t i = tid >> 6; i < B.nbig; i += 4) {
    const uint32_t b = B.big[i];
    const uint32_t s0 = B.end[b], s1 = b + 1 < NB ? B.end[b + 1] : cnt;
    const uint32_t c = s1 - s0;
    if (c <= 128) wave_bitonic_n<2>(B.keys + s0, c, CG_KEY_MASK);
    else if (c <= 256) wave_bitonic_n<4>(B.keys + s0, c, CG_KEY_MASK);
    else wave_bitonic_n<8>(B.keys + s0, c, CG_KEY_MASK);
  }
  __syncthreads();
  if (!bmg) {
    uint64_t* dst = out ? out : key;
#pragma unroll
    for (uint32_t j = 0; j < IPT; j++) {
      const uint32_t pos = j * 256 + tid;
      if (pos < cnt) dst[pos] = B.keys[pos];
    }
    __syncthreads();
    return true;
  }
  // write-back with the cluster-start bits: each wave's ballot over 64 run positions
  // ORed into the run's word-aligned LDS bitmap (bits, zero on entry), which then
  // goes to the read's bitmap (the two end words ORed in) and is left zero
  const uint32_t lane = tid & 63u, wb0 = a0 >> 6;
#pragma unroll
  for (uint32_t j = 0; j < IPT; j++) {
    const uint32_t pos = j * 256 + tid;
    bool sb = false;
    if (pos < cnt) {
      const uint64_t k = B.keys[pos];
      key[pos] = k;
      if (pos == 0) {
        sb = true;
      } else {
        const uint64_t kp = B.keys[pos - 1];
        sb = cg_rank(k) != cg_rank(kp) || (uint32_t)k - (uint32_t)kp > (uint32_t)MERGE_DISTANCE;
      }
    }
    const uint64_t m = __ballot(sb);
    if (lane == 0 && m) {
      const uint32_t p = a0 + pos, w = (p >> 6) - wb0, sh = p & 63u;
      atomicOr(reinterpret_cast<unsigned long long*>(bits + w), (unsigned long long)(m << sh));
      if (sh) atomicOr(reinterpret_cast<unsigned long long*>(bits + w + 1), (unsigned long long)(m >> (64u - sh)));
    }
  }
  __syncthreads();
  const uint32_t nwd = ((a0 + cnt - 1) >> 6) - wb0 + 1;
  for (uint32_t w = tid; w < nwd; w += 256) {
    const uint64_t v = bits[w];
    bits[w] = 0;
    if (w > 0 && w + 1 < nwd) bmg[wb0 + w] = v;
    else if (v) atomicOr(reinterpret_cast<unsigned long long*>(bmg + wb0 + w), (unsigned long long)v);
  }
  return true;
}

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_cg_sort(ClusterArgs a, CgArgs g,
                                                                                      const uint32_t* list,
                                                                                      uint32_t nlist) {
  // the bucket sort's tables, or the radix sort's and the merges' (splits in stage_val)
  __shared__ union {
    struct {
      CgSortLds rs;
      uint32_t stage_val[CLUSTER_STAGE];
    } r;
    CgBucketLds bs;
  } u;
  __shared__ uint32_t st[CG_GMAX + 1], rb[CG_GMAX + 1];
  __shared__ uint64_t bits[CG_BTILE / 64 + 2];  // a run's start bits, word-aligned (zero between runs)
  CgSortLds& rs = u.r.rs;
  for (uint32_t w = threadIdx.x; w < CG_BTILE / 64 + 2; w += 256) bits[w] = 0;
  ClusterMem M{};
  M.rsort = &rs;
  M.stage_val = u.r.stage_val;
  const uint32_t tid = threadIdx.x;
  constexpr uint32_t RUN_BITS = 38;  // rank < 64 above 32 bits of diagonal offset
  static_assert(CG_GMAX < 64, "ranks in 6 bits");
  for (uint32_t li = blockIdx.x; li < nlist; li += gridDim.x) {
    const uint32_t G = g.cg_n[li];
    if (G == 0) continue;
    const CgGroup* cg = g.cg + (uint64_t)li * (CG_GMAX + 1);
    if (tid <= G) {
      st[tid] = cg[tid].start;
      rb[tid] = cg[tid].rbits;
    }
    __syncthreads();
    uint64_t* key = g.ck + g.ck_off[li];
    uint64_t* tmp = g.tmp + g.ck_off[li];
    uint64_t* bmg = g.bm + g.bm_off[li];
    uint32_t q = 0;
    while (q < G) {
      const uint32_t a0 = st[q];
      uint32_t q1 = q + 1;
      while (q1 < G && st[q1 + 1] - a0 <= RSORT_TILE) q1++;
      const uint32_t cnt = st[q1] - a0;
      if (cnt <= RSORT_TILE) {
        if (cnt == 1) {
          cg_start_bits([&](uint32_t i) { return key[a0 + i]; }, a0, cnt, bmg, tid);
        } else if (!cg_bucket_sort(key + a0, cnt, q, q1, st, rb, u.bs, tid, 0, bmg, a0, bits)) {
          // a bucket of over 512 keys: the smallest radix-sort tile that holds the run
          if (cnt <= 2 * 256) cg_run_sort<2>(key + a0, cnt, RUN_BITS, rs, tid);
          else if (cnt <= 4 * 256) cg_run_sort<4>(key + a0, cnt, RUN_BITS, rs, tid);
          else if (cnt <= 8 * 256) cg_run_sort<8>(key + a0, cnt, RUN_BITS, rs, tid);
          else cg_run_sort<16>(key + a0, cnt, RUN_BITS, rs, tid);
          cg_start_bits([&](uint32_t i) { return key[a0 + i]; }, a0, cnt, bmg, tid);
        }
      } else {  // one segment over a tile: its keys share the rank bits
        // the tiles go to the buffer that leaves the last merge round's output in key
        uint32_t rounds = 0;
        for (uint32_t w = RSORT_TILE; w < cnt; w <<= 1) rounds++;
        const bool to_tmp = rounds & 1u;
        for (uint32_t t0 = 0; t0 < cnt; t0 += RSORT_TILE) {
          const uint32_t tc = min(RSORT_TILE, cnt - t0);
          if (!cg_bucket_sort(key + a0 + t0, tc, q, q + 1, st, rb, u.bs, tid, tc, nullptr, 0, nullptr,
                              to_tmp ? tmp + a0 + t0 : nullptr)) {
            cg_run_sort<16>(key + a0 + t0, tc, 32u, rs, tid);
            if (to_tmp) {
              for (uint32_t i = tid; i < tc; i += 256) tmp[a0 + t0 + i] = key[a0 + t0 + i];
              __syncthreads();
            }
          }
        }
        bool in_key = !to_tmp;
        for (uint32_t w = RSORT_TILE; w < cnt; w <<= 1) {
          const uint64_t* sk = in_key ? key : tmp;
          uint64_t* dk = in_key ? tmp : key;
          for (uint32_t s0 = 0; s0 < cnt; s0 += 2 * w) {
            const uint32_t na = min(w, cnt - s0), nb = min(w, cnt - s0 - na);
            block_merge<false>(M, sk, nullptr, dk, nullptr, a0 + s0, na, nb, tid, CG_KEY_MASK);
          }
          __syncthreads();
          in_key = !in_key;
        }
        if (!in_key) {
          for (uint32_t i = tid; i < cnt; i += 256) key[a0 + i] = tmp[a0 + i];
          __syncthreads();
        }
        cg_start_bits([&](uint32_t i) { return key[a0 + i]; }, a0, cnt, bmg, tid);
      }
      q = q1;
    }
    __syncthreads();
  }
}

// First cluster start after position q in the start bitmap (the sentinel bit at
// the key count bounds the search).
__device__ __forceinline__ uint32_t bm_next_start(const uint64_t* bm, uint32_t q) {
  uint32_t w = q >> 6;
  const uint32_t o = q & 63u;
  uint64_t b = o == 63u ? 0ull : (bm[w] >> (o + 1)) << (o + 1);
  while (!b) b = bm[++w];
  return (w << 6) + (uint32_t)__builtin_ctzll(b);
}

// top20_ordered's selection over composite keys, bit-parallel: a bitmap of the
// cluster starts (bit i set when key i begins a cluster, bit n set as a
// sentinel), loaded as k_cg_sort wrote it (bmg, cg_start_bits); a cluster of s
// keys starting at p is a set bit p whose next set bit is p + s, so the clusters
// of 1..CNT_REG keys are counted per 64-bit word with shifted masks, and only the
// starts of larger ones are followed to their ends.  The same masks give, once T
// is known, the clusters counted above it and the ties, whose order is the word
// order (a scan of the words' tie counts) and the bit order within a word.  The
// chosen clusters' offset sums are then added up by the waves.  n < 64 BM_WORDS.
// With features (fout set), the chosen clusters' prefilter statistics come from
// the flags in their keys' bits 40..47 and each candidate's feature row is written
// (pr: the read's quality and repeat figures; the rest is filled in here).
constexpr uint32_t BM_WORDS = 2048;
struct Top20Feat {
  uint32_t nb[TOP_CANDIDATES], ms[TOP_CANDIDATES], mn[TOP_CANDIDATES], span[TOP_CANDIDATES], rcnt[TOP_CANDIDATES];
};
template <int T>
__device__ __forceinline__ bool top20_bitmap(const uint64_t* key, uint64_t* bm, void* tabs, const CgGroup* grp,
                                             uint32_t n, uint32_t L, Cand* out, uint8_t* ncand_out, uint32_t tid,
                                             const uint64_t* bmg, uint16_t* fout = nullptr, Top20Feat* tf = nullptr,
                                             PfRead pr = PfRead{}) {
  uint32_t* hist = reinterpret_cast<uint32_t*>(tabs);  // CNT_BINS counters
  uint32_t* misc = hist + CNT_BINS;  // [0] chosen, [1] T, [2] ties wanted, [3] nt, [4] T saturated, [5] ties so far, [6..] scans
  uint32_t* ocnt = misc + 16;        // chosen clusters: count, first key, offset sum
  uint32_t* ostart = ocnt + TOP_CANDIDATES;
  long long* osum = reinterpret_cast<long long*>(ostart + TOP_CANDIDATES);
  const uint32_t lane = (uint32_t)lane_id(), wave = tid >> 6;
  const uint32_t nw = n / 64 + 1;  // words holding bits 0..n
  for (uint32_t i = tid; i < CNT_BINS; i += T) hist[i] = 0;
  if (tid < 16) misc[tid] = 0;
  // A. the start bitmap, and the sentinel bit n
  for (uint32_t w = tid; w < nw; w += T) bm[w] = bmg[w] | (w == (n >> 6) ? 1ull << (n & 63u) : 0ull);
  team_sync<T>();
  const uint32_t nwd = n >> 6;
  const uint64_t sent = 1ull << (n & 63u);  // the sentinel bit, in word nwd
  // starts of clusters of exactly s keys (s <= CNT_REG) in word w, from the starts
  // m whose size is not known yet; m keeps those of more than s keys
  auto size_masks = [&](uint32_t w, uint64_t& m, uint64_t* eq) {
    const uint64_t b = bm[w], nx = w + 1 < nw ? bm[w + 1] : 0ull;
    m = w == nwd ? b & ~sent : b;
#pragma unroll
    for (uint32_t s1 = 1; s1 <= CNT_REG; s1++) {
      const uint64_t sh = (b >> s1) | (nx << (64 - s1));  // bit p: a start at p + s1
      eq[s1 - 1] = m & sh;
      m &= ~sh;
    }
  };
  // B. counts
  uint32_t tally[CNT_REG];
#pragma unroll
  for (uint32_t j = 0; j < CNT_REG; j++) tally[j] = 0;
  for (uint32_t w = tid; w < nw; w += T) {
    uint64_t m, eq[CNT_REG];
    size_masks(w, m, eq);
#pragma unroll
    for (uint32_t j = 0; j < CNT_REG; j++) tally[j] += (uint32_t)__popcll(eq[j]);
    while (m) {
      const uint32_t p = (w << 6) + (uint32_t)__builtin_ctzll(m);
      m &= m - 1;
      atomicAdd(&hist[min(bm_next_start(bm, p) - p, CNT_BINS - 1)], 1u);
    }
  }
#pragma unroll
  for (uint32_t j = 0; j < CNT_REG; j++) {
    const uint32_t v = group_sum_u32<64>(tally[j]);
    if (lane == 0 && v) atomicAdd(&hist[j + 1], v);
  }
  team_sync<T>();
  {
    constexpr uint32_t PB = CNT_BINS / T;
    static_assert(CNT_BINS % T == 0, "whole bins per thread");
    const uint32_t hi = CNT_BINS - 1 - tid * PB;
    uint32_t local = 0;
#pragma unroll
    for (uint32_t j = 0; j < PB; j++) local += hist[hi - j];
    uint32_t C;
    uint32_t above = team_excl_scan<T>(local, C, misc + 6);
    const uint32_t nt = min(C, (uint32_t)TOP_CANDIDATES);
#pragma unroll
    for (uint32_t j = 0; j < PB; j++) {
      const uint32_t h = hist[hi - j];
      if (above < nt && above + h >= nt) {
        misc[1] = hi - j;
        misc[2] = nt - above;
        misc[4] = hi - j == CNT_BINS - 1 ? 1u : 0u;
      }
      above += h;
    }
    if (tid == 0) misc[3] = nt;
  }
  team_sync<T>();
  const uint32_t Tc = misc[1], want = misc[2], nt = misc[3];
  if (misc[4]) return false;
  // C. the chosen clusters, word by word in order (one word per thread per round)
  for (uint32_t r0 = 0; r0 < nw; r0 += T) {
    const uint32_t w = r0 + tid;
    uint64_t eqT = 0, gtT = 0;
    if (w < nw) {
      uint64_t m, eq[CNT_REG];
      size_masks(w, m, eq);
      if (Tc <= CNT_REG) {
        // sizes above T: the starts of sizes T + 1 .. CNT_REG, and the rest in m
#pragma unroll
        for (uint32_t j = 0; j < CNT_REG; j++) {
          if (j + 1 == Tc) eqT = eq[j];
          if (j + 1 > Tc) gtT |= eq[j];
        }
        gtT |= m;
      } else {
        uint64_t mm = m;
        while (mm) {
          const uint32_t b = (uint32_t)__builtin_ctzll(mm);
          mm &= mm - 1;
          const uint32_t p = (w << 6) + b, sz = bm_next_start(bm, p) - p;
          if (sz == Tc) eqT |= 1ull << b;
          else if (sz > Tc) gtT |= 1ull << b;
        }
      }
    }
    uint32_t tot;
    const uint32_t ex = team_excl_scan<T>((uint32_t)__popcll(eqT), tot, misc + 6) + misc[5];
    while (gtT) {
      const uint32_t b = (uint32_t)__builtin_ctzll(gtT);
      gtT &= gtT - 1;
      const uint32_t p = (w << 6) + b;
      const uint32_t at = atomicAdd(&misc[0], 1u);
      if (at < (uint32_t)TOP_CANDIDATES) {
        ostart[at] = p;
        ocnt[at] = bm_next_start(bm, p) - p;
      }
    }
    uint32_t ti = ex;
    while (eqT && ti < want) {
      const uint32_t b = (uint32_t)__builtin_ctzll(eqT);
      eqT &= eqT - 1;
      const uint32_t at = atomicAdd(&misc[0], 1u);
      if (at < (uint32_t)TOP_CANDIDATES) {
        ostart[at] = (w << 6) + b;
        ocnt[at] = Tc;
      }
      ti++;
    }
    team_sync<T>();
    if (tid == 0) misc[5] += tot;
    team_sync<T>();
  }
  // D. offset sums of the chosen clusters, one wave each (with features: their
  // flags' neighbour hits, class sum and least class, and the offsets' span)
  for (uint32_t e = wave; e < nt; e += T / 64) {
    const uint32_t p = ostart[e], c = ocnt[e];
    int64_t sum = 0;
    uint32_t fnb = 0, fms = 0, fmn = 15;
    for (uint32_t x = lane; x < c; x += 64) {
      const uint64_t k = key[p + x];
      sum += (int64_t)(uint32_t)k;
      const uint32_t f = (uint32_t)(k >> CG_FLAG_SHIFT) & 0xFFu;
      fnb += f >> 7;
      fms += f & 15u;
      fmn = min(fmn, f & 15u);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) sum += shfl_xor_i64(sum, o);
    if (fout) {
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        fnb += (uint32_t)__shfl_xor((int)fnb, o, 64);
        fms += (uint32_t)__shfl_xor((int)fms, o, 64);
        fmn = min(fmn, (uint32_t)__shfl_xor((int)fmn, o, 64));
      }
      if (lane == 0) {
        tf->nb[e] = fnb;
        tf->ms[e] = fms;
        tf->mn[e] = fmn;
        tf->span[e] = min((uint32_t)key[p + c - 1] - (uint32_t)key[p], 0xFFFFFFFu);
      }
    }
    if (lane == 0) osum[e] = (long long)sum;
  }
  team_sync<T>();
  // E. the candidates, ordered by (count desc, first key)
  uint32_t rk = 0;
  if (tid < nt) {
    const uint32_t c = ocnt[tid], cs = ostart[tid];
    for (uint32_t j = 0; j < nt; j++) rk += (ocnt[j] > c || (ocnt[j] == c && ostart[j] < cs)) ? 1u : 0u;
    if (fout) tf->rcnt[rk] = c;
    const CgGroup& g = grp[cg_rank(key[cs])];
    const int64_t dsum = (int64_t)osum[tid] + (int64_t)c * ((int64_t)g.lo - DIAG_BIAS);
    const int64_t avg = dsum / (int64_t)c;
    const bool rev = g.key & 1u;
    int64_t position = rev ? avg - (int64_t)L + K : avg;
    if (position < 0) position = 0;
    Cand cd;
    cd.pos = (uint64_t)position;
    cd.contig = g.key >> 1;
    cd.strand_score = (rev ? 0x80000000u : 0u) | min(c, 0x7FFFFFFFu);
    out[rk] = cd;
  }
  if (tid == 0) *ncand_out = (uint8_t)nt;
  if (fout) {
    team_sync<T>();
    if (tid < nt) {
      PfCluster pc;
      pc.cnt = ocnt[tid];
      pc.nb = tf->nb[tid];
      pc.ms = tf->ms[tid];
      pc.mn = tf->mn[tid];
      pc.span = tf->span[tid];
      pc.rev = (grp[cg_rank(key[ostart[tid]])].key & 1u) != 0;
      PfRead q = pr;
      q.c0 = tf->rcnt[0];
      q.c1 = nt > 1 ? tf->rcnt[1] : 0u;
      // clusters counted as the top one: every one above T is chosen; when the top
      // count is T itself, all the ties are (hist holds them: T is below the last bin)
      uint32_t ties = 0;
      if (q.c0 == Tc) {
        ties = hist[Tc];
      } else {
        for (uint32_t j = 0; j < nt; j++) ties += ocnt[j] == q.c0 ? 1u : 0u;
      }
      q.ties = ties;
      uint32_t C = 0;
      for (uint32_t b = 0; b < CNT_BINS; b++) C += hist[b];
      q.C = C;
      q.n = n;
      q.L = L;
      pf_feature_row(fout + rk * 32, rk, pc, q);
    }
  }
  team_sync<T>();
  return true;
}

__global__ __launch_bounds__(256) void k_cg_select(ClusterArgs a, CgArgs g, const uint32_t* list, uint32_t nlist) {
  __shared__ uint64_t lds[(TOP20_LDS_BYTES + 7) / 8];
  __shared__ uint64_t bm[BM_WORDS + 1];
  __shared__ CgGroup gs[CG_GMAX + 1];
  __shared__ Top20Feat tf;
  __shared__ uint32_t qs[4];
  const uint32_t tid = threadIdx.x;
  for (uint32_t li = blockIdx.x; li < nlist; li += gridDim.x) {
    const uint32_t G = g.cg_n[li];
    if (G == 0) continue;
    const uint32_t r = list[li];
    const CgGroup* cg = g.cg + (uint64_t)li * (CG_GMAX + 1);
    if (tid <= G) gs[tid] = cg[tid];
    if (tid < 4) qs[tid] = 0;
    __syncthreads();
    const uint32_t L = (uint32_t)(a.R.seq_off[r + 1] - a.R.seq_off[r]);
    const uint32_t n = a.nhits[r];
    Cand* out = a.cands + (uint64_t)r * TOP_CANDIDATES;
    const uint64_t* key = g.ck + g.ck_off[li];
    PfRead pr{};
    if (a.feat) {  // the read's quality sum, bases below Q10 and Q20, and repeat-skipped windows
      const uint64_t qo = a.R.qual_off[r];
      pr.qlen = (uint32_t)(a.R.qual_off[r + 1] - qo);
      uint32_t q0 = 0, q10 = 0, q20 = 0, rep = 0;
      for (uint32_t x = tid; x < pr.qlen; x += 256) {
        const uint32_t q = a.R.qual[qo + x];
        q0 += q;
        q10 += q < 10;
        q20 += q < 20;
      }
      const uint32_t nwin = L >= (uint32_t)K ? L - K + 1 : 0;
      const uint8_t* rc = a.rcode + a.R.seq_off[r];
      for (uint32_t x = tid; x < nwin; x += 256) rep += rc[x] != 0;
      q0 = group_sum_u32<64>(q0);
      q10 = group_sum_u32<64>(q10);
      q20 = group_sum_u32<64>(q20);
      rep = group_sum_u32<64>(rep);
      if (lane_id() == 0) {
        atomicAdd(&qs[0], q0);
        atomicAdd(&qs[1], q10);
        atomicAdd(&qs[2], q20);
        atomicAdd(&qs[3], rep);
      }
      __syncthreads();
      pr.qsum = qs[0];
      pr.q10 = qs[1];
      pr.q20 = qs[2];
      pr.rep = qs[3];
      pr.G = G;
    }
    bool ok;
    if (n < 64 * BM_WORDS)
      ok = top20_bitmap<256>(key, bm, lds, gs, n, L, out, a.ncand + r, tid, g.bm + g.bm_off[li],
                             a.feat ? a.feat + (uint64_t)r * TOP_CANDIDATES * MLP_IN : nullptr, &tf, pr);
    else  // the walk over the keys (no features: such a read goes to the general tier with them)
      ok = !a.feat && top20_ordered<256, true>(key, lds, gs, n, L, out, a.ncand + r, tid);
    if (!ok) {
      // the 20th count lies in the saturated bin (20 clusters of 1023 hits or
      // more): the general tier takes the read from its intact hits
      if (tid == 0) g.fb_list[atomicAdd(g.fb_n, 1u)] = r;
    }
    __syncthreads();
  }
}

void launch_cluster_grid(const ClusterArgs& a, const CgArgs& g, const uint32_t* list, uint32_t n, hipStream_t s) {
  if (n == 0) return;
  const uint32_t blocks = n < 65536 ? n : 65536;
  hipLaunchKernelGGL(k_cg_group, dim3(blocks), dim3(256), 0, s, a, g, list, n);
  hipLaunchKernelGGL(k_cg_sort, dim3(blocks), dim3(256), 0, s, a, g, list, n);
  hipLaunchKernelGGL(k_cg_select, dim3(blocks), dim3(256), 0, s, a, g, list, n);
}

// ---------------------------------------------------------------------------
// DP job construction: single.rs:105-133
// ---------------------------------------------------------------------------
__global__ void k_build_jobs(JobArgs a) {
  uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= a.R.n) return;
  const uint64_t L = a.R.seq_off[r + 1] - a.R.seq_off[r];
  uint32_t nc = a.ncand[r], k = 0;
  for (uint32_t c = 0; c < nc; c++) {
    Cand cd = a.cands[(uint64_t)r * TOP_CANDIDATES + c];
    uint64_t clen = a.G.clen[cd.contig];
    uint64_t ref_start = cd.pos >= (uint64_t)LEFT_PAD ? cd.pos - LEFT_PAD : 0;
    uint64_t ref_end = cd.pos + L + (uint64_t)a.right_pad;
    if (ref_end > clen) ref_end = clen;
    if (ref_start >= ref_end) continue;  // single.rs:118
    DpJob j;
    j.ref_gb = a.G.coff[cd.contig] + ref_start;
    j.read = r;
    j.n = (uint32_t)(ref_end - ref_start);
    j.hint = LEFT_PAD;
    j.band = a.band;
    j.rc = (cd.strand_score >> 31) & 1u;
    j.explicit_ref = 0;
    j.pad = 0;
    j.cand = r * TOP_CANDIDATES + c;
    a.jobs[(uint64_t)r * TOP_CANDIDATES + k] = j;
    k++;
  }
  a.njobs[r] = (uint8_t)k;
}

void launch_build_jobs(const JobArgs& a, hipStream_t s) {
  if (a.R.n == 0) return;
  hipLaunchKernelGGL(k_build_jobs, dim3((a.R.n + 255) / 256), dim3(256), 0, s, a);
}

// ---------------------------------------------------------------------------
// DP: smith_waterman.rs:163-255
// ---------------------------------------------------------------------------
// Per-job LDS staging shared by both DP kernels.
struct DpStage {
  uint8_t* q;      // query bases [m]
  uint8_t* qq;     // query quals [qlen]
  uint8_t* ref;    // ref_seg [n]
  int16_t* go;     // gap-open at column j (index j-1) [n]
  uint32_t* runs;  // CIGAR runs (reverse order) [m + n + 2]
};

// Loads query (read or its reverse complement with reversed quals,
// single.rs:101-103), ref_seg and the homopolymer gap-open per column
// (smith_waterman.rs:288-291, :348-363).
__device__ void dp_stage_job(const DpArgs& a, const DpJob& job, uint32_t m, uint32_t qlen, uint64_t so, uint64_t qo,
                             DpStage& st) {
  const int lane = lane_id();
  for (uint32_t x = lane; x < m; x += 64) st.q[x] = job.rc ? comp_base(a.R.seq[so + m - 1 - x]) : a.R.seq[so + x];
  for (uint32_t x = lane; x < qlen; x += 64) st.qq[x] = job.rc ? a.R.qual[qo + qlen - 1 - x] : a.R.qual[qo + x];
  const uint32_t n = job.n;
  for (uint32_t x = lane; x < n; x += 64)
    st.ref[x] = job.explicit_ref ? a.ref_bytes[job.ref_gb + x] : genome_base(a.G, job.ref_gb + x);
  mem_sync();
  // homopolymer run of every position: start = prefix max of run starts,
  // end = suffix min of run ends (64-column chunks with carries)
  int32_t carry = 0;
  for (uint32_t cb = 0; cb < n; cb += 64) {
    uint32_t x = cb + lane;
    int32_t s = 0;
    if (x < n) s = (x == 0 || st.ref[x] != st.ref[x - 1]) ? (int32_t)x : 0;
    s = max(wave_incl_max(s), carry);
    if (x < n) st.go[x] = (int16_t)0;  // placeholder; overwritten below
    carry = __shfl(s, 63, 64);
    if (x < n) ((int32_t*)st.runs)[x] = s;  // run start (runs[] is free scratch here)
  }
  mem_sync();
  int32_t ecarry = INT32_MAX;
  uint32_t nchunks = (n + 63) / 64;
  for (int c = (int)nchunks - 1; c >= 0; c--) {
    uint32_t x = (uint32_t)c * 64 + lane;
    int32_t e = INT32_MAX;
    if (x < n) e = (x == n - 1 || st.ref[x] != st.ref[x + 1]) ? (int32_t)x : INT32_MAX;
    e = min(wave_suffix_min(e), ecarry);
    ecarry = __shfl(e, 0, 64);
    if (x < n) {
      int32_t run = e - ((int32_t*)st.runs)[x] + 1;
      if (run > 65535) run = 65535;
      int32_t g = a.gap_open + (run - 1);
      st.go[x] = (int16_t)max(min(g, -1), -32768);
    }
  }
  mem_sync();
}

// Lane 0: traceback (smith_waterman.rs:299-345) from (m, best_j, state) with
// tb(i, j) supplied by the caller; run-length CIGAR in st.runs (reverse order).
template <typename TbFn>
__device__ void dp_traceback(uint32_t m, uint32_t best_j, uint32_t state, TbFn tb, DpStage& st, uint32_t& nruns,
                             uint32_t& ref_off) {
  uint32_t i = m, j = best_j, nr = 0, cur_op = 0xFF, cur_len = 0;
  auto push = [&](uint32_t op) {
    if (op == cur_op) {
      cur_len++;
    } else {
      if (cur_len) st.runs[nr++] = (cur_len << 4) | cur_op;
      cur_op = op;
      cur_len = 1;
    }
  };
  while (i > 0) {
    if (j == 0) {
      push(1);
      i--;
      continue;
    }
    uint32_t bits = tb(i, j);  // bit0|bit1 = tb_m, bit2 = tb_i, bit3 = tb_d
    if (state == 0) {
      push(0);
      i--;
      j--;
      state = bits & 3u;
    } else if (state == 1) {
      push(1);
      i--;
      state = (bits & 4u) ? 1u : 0u;
    } else {
      push(2);
      j--;
      state = (bits & 8u) ? 2u : 0u;
    }
  }
  if (cur_len) st.runs[nr++] = (cur_len << 4) | cur_op;
  nruns = nr;
  ref_off = j;
}

// dp_traceback's walk taken a run at a time by the whole wave: in state M lane l
// looks at the cell l steps up the diagonal, in I l rows up, in D l columns left
// (tb_at(i, j): the cell's bits, as dp_traceback's tb), and a ballot finds where
// the state's run ends -- the first cell whose bits leave the state, or the edge
// of the matrix -- so a run of up to 64 steps costs one round.  Every lane ends
// with the same runs and offset as dp_traceback.
template <typename TbAt>
__device__ void dp_traceback_wave(uint32_t m, uint32_t best_j, uint32_t state, TbAt tb_at, DpStage& st,
                                  uint32_t& nruns, uint32_t& ref_off) {
  const uint32_t lane = (uint32_t)lane_id();
  uint32_t i = m, j = best_j, nr = 0, cur_op = 0xFF, cur_len = 0;
  auto push = [&](uint32_t op, uint32_t len) {
    if (op == cur_op) {
      cur_len += len;
    } else {
      if (cur_len) {
        if (lane == 0) st.runs[nr] = (cur_len << 4) | cur_op;
        nr++;
      }
      cur_op = op;
      cur_len = len;
    }
  };
  while (i > 0) {
    if (j == 0) {  // column 0: single I steps to the first row
      push(1, i);
      i = 0;
      break;
    }
    const uint32_t room = state == 0 ? min(i, j) : state == 1 ? i : j;  // steps before an edge
    bool stop = true;
    uint32_t bits = 0;
    if (lane < room) {
      const uint32_t ii = state == 2 ? i : i - lane, jj = state == 1 ? j : j - lane;
      bits = tb_at(ii, jj);
      stop = state == 0 ? (bits & 3u) != 0 : state == 1 ? (bits & 4u) == 0 : (bits & 8u) == 0;
    }
    const uint64_t sm = __ballot(stop);  // lanes at or past the edge stop
    const uint32_t f = sm ? (uint32_t)__builtin_ctzll(sm) : 64u;
    const uint32_t fb = (uint32_t)__shfl((int)bits, (int)(f & 63u), 64);
    uint32_t len, nstate;
    if (f < room && f < 64u) {  // the run ends at cell f, whose bits give the next state
      len = f + 1;
      nstate = state == 0 ? (fb & 3u) : 0u;
    } else {  // 64 steps, or the run reaches the edge, in its own state
      len = f;
      nstate = state;
    }
    push(state == 0 ? 0u : state == 1 ? 1u : 2u, len);
    if (state != 2) i -= len;
    if (state != 1) j -= len;
    state = nstate;
  }
  if (cur_len) {
    if (lane == 0) st.runs[nr] = (cur_len << 4) | cur_op;
    nr++;
  }
  nruns = nr;
  ref_off = j;
}

// Lane 0: alignment_log_likelihood_bases (bayesian.rs:142-179) over the
// reverse-ordered runs; ref_window = ref_seg[off .. off+span).
__device__ double dp_log_lk(const DevLuts& T, const DpStage& st, uint32_t m, uint32_t qlen, uint32_t nruns,
                            uint32_t off, uint32_t span) {
  double lk = 0.0;
  uint32_t qi = 0, ri = 0;
  for (int t = (int)nruns - 1; t >= 0; t--) {
    uint32_t w = st.runs[t], op = w & 15u, len = w >> 4;
    if (op == 0) {
      for (uint32_t u = 0; u < len; u++) {
        if (qi >= m || ri >= span) break;
        uint32_t q = qi < qlen ? st.qq[qi] : 30u;
        lk += (st.q[qi] == st.ref[off + ri]) ? T.lk_match[q] : T.lk_mis[q];
        qi++;
        ri++;
      }
    } else if (op == 1) {
      for (uint32_t u = 0; u < len; u++) {
        if (qi >= m) break;
        lk += T.ln_quarter;
        qi++;
      }
    } else if (op == 2) {
      ri += len;
    }
  }
  return lk;
}

// dp_log_lk by a whole wave: each round its lanes take 64 consecutive read
// positions, find their run with a cursor that only moves forward and write the
// position's term to terms[lane]; lane 0 adds them in read order (the same
// doubles in the same order as dp_log_lk; every lane returns the sum).
__device__ double dp_log_lk_wave(const DevLuts& T, const DpStage& st, uint32_t m, uint32_t qlen, uint32_t nruns,
                                 uint32_t off, uint32_t span, double* terms) {
  const uint32_t lane = (uint32_t)lane_id();
  int u = (int)nruns - 1;   // runs in read order: st.runs[nruns - 1] first
  uint32_t qs = 0, rs = 0;  // read and reference bases before run u
  double lk = 0.0;
  for (uint32_t base = 0; base < m; base += 64) {
    const uint32_t qi = base + lane;
    double tv = 0.0;
    if (qi < m) {
      uint32_t op = 3u;
      while (u >= 0) {
        const uint32_t w = st.runs[u], len = w >> 4;
        op = w & 15u;
        const uint32_t qc = op <= 1u ? len : 0u;
        if (qi < qs + qc) break;
        qs += qc;
        rs += (op == 0u || op == 2u) ? len : 0u;
        u--;
        op = 3u;
      }
      if (op == 0u) {
        const uint32_t ri = rs + (qi - qs);
        if (ri < span) {
          const uint32_t q = qi < qlen ? st.qq[qi] : 30u;
          tv = (st.q[qi] == st.ref[off + ri]) ? T.lk_match[q] : T.lk_mis[q];
        }
      } else if (op == 1u) {
        tv = T.ln_quarter;
      }
    }
    terms[lane] = tv;
    wave_sync();
    if (lane == 0) {
      const uint32_t c = min(64u, m - base);
      for (uint32_t v = 0; v < c; v++) lk += terms[v];
    }
    wave_sync();
  }
  const long long b = __double_as_longlong(lk);
  const uint32_t lo = (uint32_t)__shfl((int)(uint32_t)b, 0, 64), hi = (uint32_t)__shfl((int)(uint32_t)(b >> 32), 0, 64);
  return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// Copies the CIGAR (forward order) and fills DpOut; lane-parallel copy.
// Hands job `jid` to the next kernel of the chain (one lane calls it).
__device__ __forceinline__ void dp_flag(const DpArgs& a, uint32_t jid) {
  const uint32_t at = atomicAdd(a.flag_n, 1u);
  a.flag_list[at] = jid;
}

__device__ void dp_emit(const DpArgs& a, uint32_t jid, const DpStage& st, uint32_t nruns, uint32_t off, uint32_t span,
                        int32_t score, double lk, uint64_t cells) {
  const int lane = lane_id();
  uint32_t ncopy = nruns < a.cigar_cap ? nruns : a.cigar_cap;
  for (uint32_t t = lane; t < ncopy; t += 64) a.cigar[((uint64_t)jid - a.cigar_base) * a.cigar_cap + t] = st.runs[nruns - 1 - t];
  if (lane == 0) {
    DpOut o;
    o.log_lk = lk;
    o.ref_off = off;
    o.score = score;
    o.cigar_n = nruns;
    o.span = span;
    o.cells = cells;
    a.out[jid] = o;
  }
}

constexpr int DP_WPB = 4;

// Dynamic LDS per wave for the banded kernel.
__host__ __device__ inline uint32_t dp_band_lds_bytes(uint32_t max_m, uint32_t max_n) {
  uint32_t b = 0;
  b += (max_m + 15) & ~15u;            // q
  b += (max_m + 15) & ~15u;            // qq
  b += (max_n + 15) & ~15u;            // ref
  b += ((max_n * 2) + 15) & ~15u;      // go
  b += ((max_m + max_n + 2) * 4 + 15) & ~15u;  // runs (also run-start scratch: needs n ints)
  b += (max_m + 1) * 32;               // tb rows: 4 x u64
  return b;
}

// Band of at most 63 columns (2*band+1 <= 63): lane d holds column
// j = i + hint - band + d of row i; M/I/D of the previous row stay in
// registers (same lane = diagonal predecessor, lane+1 = vertical).
__global__ __launch_bounds__(64 * DP_WPB) void k_dp_band(DpArgs a) {
  extern __shared__ __align__(16) uint8_t dyn[];
  const int wib = threadIdx.x >> 6, lane = lane_id();
  // the wave's staging: its LDS slice, or its slice of device memory when the
  // tile does not fit LDS
  uint8_t* base = a.gstage ? a.gstage + (uint64_t)(blockIdx.x * a.wpb + wib) * a.gstage_stride
                           : dyn + (size_t)wib * dp_band_lds_bytes(a.max_m, a.max_n);
  DpStage st;
  st.q = base;
  st.qq = st.q + ((a.max_m + 15) & ~15u);
  st.ref = st.qq + ((a.max_m + 15) & ~15u);
  st.go = (int16_t*)(st.ref + ((a.max_n + 15) & ~15u));
  st.runs = (uint32_t*)((uint8_t*)st.go + (((a.max_n * 2) + 15) & ~15u));
  uint64_t* tbl = (uint64_t*)((uint8_t*)st.runs + (((a.max_m + a.max_n + 2) * 4 + 15) & ~15u));
  const uint32_t gw = blockIdx.x * a.wpb + wib, nwaves = gridDim.x * a.wpb;
  unsigned long long c_cells = 0;
  for (uint32_t t = gw; t < a.njobs; t += nwaves) {
    const uint32_t jid = a.jlist ? a.jlist[t] : t;
    if (a.job_valid) {  // slot = read * 20 + k, valid while k < njobs[read]
      if ((jid % TOP_CANDIDATES) >= a.job_valid[jid / TOP_CANDIDATES]) continue;
    }
    const DpJob job = a.jobs[jid];
    const uint64_t so = a.R.seq_off[job.read], qo = a.R.qual_off[job.read];
    const uint32_t m = (uint32_t)(a.R.seq_off[job.read + 1] - so);
    const uint32_t qlen = (uint32_t)(a.R.qual_off[job.read + 1] - qo);
    const uint32_t n = job.n;
    const int32_t b = job.band, h = job.hint, hb = h - b;
    if (m == 0 || n == 0) {  // smith_waterman.rs:166-168
      if (lane == 0) {
        st.runs[0] = (m << 4) | 4u;
      }
      wave_sync();
      dp_emit(a, jid, st, 1, 0, 0, 0, 0.0, 0);
      continue;
    }
    if (m > a.max_m || n > a.max_n || 2 * b + 1 > 63 || b < 0) {
      if (lane == 0) dp_flag(a, jid);
      continue;
    }
    dp_stage_job(a, job, m, qlen, so, qo, st);
    for (uint32_t x = lane; x < 4 * (m + 1); x += 64) tbl[x] = 0;
    // computed rows: i + h + b >= 1 and i + h - b <= n
    const int32_t i0 = max(1, 1 - h - b);
    const int32_t i1 = min((int32_t)m, (int32_t)n + b - h);
    int32_t Mp, Ip, Dp;
    {
      int32_t col = (i0 - 1) + hb + lane;  // row 0 in row (i0-1)'s lane coordinates
      Mp = (col >= 0 && col <= (int32_t)n) ? 0 : NEG_INF;
      Ip = NEG_INF;
      Dp = NEG_INF;
    }
    mem_sync();  // the cleared rows before lane 0 fills them
    uint64_t cells = 0;
    const int32_t ge = a.gap_extend;
    for (int32_t i = i0; i <= i1; i++) {
      const int32_t j = i + hb + lane;
      const bool valid = lane <= 2 * b && j >= 1 && j <= (int32_t)n;
      const uint8_t rb = st.q[i - 1];
      const uint32_t qv = (uint32_t)(i - 1) < qlen ? st.qq[i - 1] : 30u;
      const int32_t mis = a.T.mis_pen[qv];
      uint8_t g = 0;
      int32_t go = 0;
      if (valid) {
        g = st.ref[j - 1];
        go = st.go[j - 1];
      }
      const int32_t sc = (rb == g) ? MATCH_SCORE : mis;
      const int32_t fm = Mp + sc, fi = Ip + sc, fd = Dp + sc;
      int32_t M = max(max(fm, fi), fd);
      const int32_t Mup = dpp_wave_shl1(NEG_INF, Mp), Iup = dpp_wave_shl1(NEG_INF, Ip);
      const int32_t io = Mup + go, ie = Iup + ge;
      int32_t I = max(io, ie);
      if (!valid) M = NEG_INF;
      const int32_t Ml = dpp_wave_shr1(NEG_INF, M);
      const int32_t A = Ml + go;
      const int32_t d0 = max(0, 1 - (i + hb));
      const int32_t Sent = NEG_INF - (d0 - 1) * ge;
      int32_t B = valid ? A - lane * ge : INT32_MIN;
      B = wave_incl_max(B);
      int32_t D = lane * ge + max(B, Sent);
      const uint64_t vm = __ballot(valid);
      const uint64_t ne_fm = __ballot(valid && M != fm);
      const uint64_t eq_fi = __ballot(valid && M == fi);
      const uint64_t ti = __ballot(valid && I != io);
      const uint64_t td = __ballot(valid && D != A);
      if (lane == 0) {
        uint64_t* row = tbl + 4 * (uint32_t)i;
        row[0] = ne_fm & eq_fi;
        row[1] = ne_fm & ~eq_fi;
        row[2] = ti;
        row[3] = td;
      }
      cells += (uint64_t)__popcll(vm);
      if (!valid) {
        I = NEG_INF;
        D = NEG_INF;
      }
      Mp = M;
      Ip = I;
      Dp = D;
    }
    wave_sync();
    // endpoint: last maximum over (j asc; M, I) of the last computed row
    uint32_t best_j, best_state;
    int32_t best_score;
    if (i1 < i0) {  // no row computed: row 0 (M = 0 for all j)
      best_j = n;
      best_state = 0;
      best_score = 0;
    } else {
      const int32_t j = i1 + hb + lane;
      const bool valid = lane <= 2 * b && j >= 1 && j <= (int32_t)n;
      int64_t key = INT64_MIN;
      if (valid) {
        int64_t km = end_key(Mp, (uint32_t)(2 * j));
        int64_t ki = end_key(Ip, (uint32_t)(2 * j + 1));
        key = km > ki ? km : ki;
      }
      const int32_t js = max(i1 + hb, 1), je = max(min(i1 + h + b, (int32_t)n), 0);
      const int32_t joob = je < (int32_t)n ? (int32_t)n : js - 1;
      const int64_t koob = end_key(NEG_INF, (uint32_t)(2 * joob + 1));
      key = wave_max_i64(key);
      key = key > koob ? key : koob;
      best_score = (int32_t)(key >> 32);
      best_j = ((uint32_t)key) >> 1;
      best_state = ((uint32_t)key) & 1u;
    }
    uint32_t nruns = 0, off = 0, span = 0;
    double lk = 0.0;
    if (lane == 0) {
      auto tb = [&](uint32_t ii, uint32_t jj) -> uint32_t {
        int32_t d = (int32_t)jj - ((int32_t)ii + hb);
        if (d < 0 || d >= 64) return 0u;
        const uint64_t* row = tbl + 4 * ii;
        return (uint32_t)((row[0] >> d) & 1ull) | ((uint32_t)((row[1] >> d) & 1ull) << 1) |
               ((uint32_t)((row[2] >> d) & 1ull) << 2) | ((uint32_t)((row[3] >> d) & 1ull) << 3);
      };
      dp_traceback(m, best_j, best_state, tb, st, nruns, off);
      for (uint32_t t2 = 0; t2 < nruns; t2++) {
        uint32_t op = st.runs[t2] & 15u;
        if (op == 0 || op == 2) span += st.runs[t2] >> 4;
      }
      lk = dp_log_lk(a.T, st, m, qlen, nruns, off, span);
    }
    nruns = shfl_u32(nruns, 0);
    off = shfl_u32(off, 0);
    span = shfl_u32(span, 0);
    mem_sync();  // lane 0's runs to every lane (LDS or device-memory staging)
    dp_emit(a, jid, st, nruns, off, span, best_score, lk, cells);
    c_cells += cells;
    wave_sync();
  }
  if (lane == 0 && a.ctr) atomicAdd(&a.ctr->cells, c_cells);
}

uint64_t dp_band_gstage_bytes(uint32_t max_m, uint32_t max_n) {
  const uint64_t per_wave = dp_band_lds_bytes(max_m, max_n);
  return per_wave <= 65536 ? 0 : (per_wave + 255) & ~255ull;
}
void launch_dp_band(const DpArgs& a0, hipStream_t s) {
  if (a0.njobs == 0) return;
  DpArgs a = a0;
  uint32_t blocks;
  size_t lds = 0;
  if (a.gstage) {  // staging in device memory: gstage_waves slices of gstage_stride bytes
    a.wpb = a.gstage_waves < (uint32_t)DP_WPB ? (a.gstage_waves ? a.gstage_waves : 1u) : (uint32_t)DP_WPB;
    blocks = (a.njobs + a.wpb - 1) / a.wpb;
    const uint32_t cap = a.gstage_waves / a.wpb;
    if (blocks > cap) blocks = cap ? cap : 1u;
  } else {
    const uint32_t per_wave = dp_band_lds_bytes(a.max_m, a.max_n);
    a.wpb = 65536 / per_wave;
    if (a.wpb > DP_WPB) a.wpb = DP_WPB;
    if (a.wpb < 1) a.wpb = 1;
    blocks = (a.njobs + a.wpb - 1) / a.wpb;
    lds = (size_t)a.wpb * per_wave;
  }
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(k_dp_band, dim3(blocks), dim3(64 * a.wpb), lds, s, a);
}

// ---------------------------------------------------------------------------
// DP, four jobs per wave (k_dp_quad): smith_waterman.rs:163-345 for bands of
// W = 2*band+1 <= 63 columns whose rows are all computed (no skipped row),
// genome-sourced reference windows and reads of at most QUAD_MAX_M bases.
// Each job owns one 16-lane DPP row of the wave; lane L holds the columns
// d' = 4L..4L+3 of a 64-column window that puts the band at [s, s+W),
// s = 63 - W, and a guard column at d' = 63.  Column d' of row i is the
// reference column j = i + o + d', o = hint + band - 62, so the diagonal
// predecessor keeps d', the vertical one is d'+1 and the horizontal one d'-1.
//  - M and I of the previous row stay in registers, 4 per lane.
//  - D is the max-plus prefix scan D[d'] = max_k (M[k-1] + go[k] + (d'-k) ge)
//    over the band, exact in int32: a lane-local prefix over 4 columns, then
//    an exclusive scan over the 16 lanes with row_shr DPP moves.  The row
//    keeps P = D - d'*ge.
//  - Columns left of the band never feed it (the scan excludes them).  The
//    guard column scores -2^28 per row, so from row 1 on it is unreachable,
//    as the reference's NEG_INF fill makes it; row 0 keeps 0 there.  Cells
//    with j <= 0 are reset to NEG_INF in the rows where the band reaches
//    them.  Cells with j > n are computed but never read by a cell with
//    j <= n, never reach the endpoint scan and never lie on the traceback
//    path.  Unreachable values stay within [-2^31, -2^29).
//  - Traceback bits go into four words per lane with v_alignbit (planes:
//    M came from I or D; M came from D; tb_i; tb_d) and to the job's area in
//    global scratch every 8 rows, one 16-byte store per lane ([block][lane]
//    [plane]).  The words of the five lanes around the endpoint's are then
//    staged into LDS for every block with loads issued together, and one lane
//    per job walks them, taking runs of diagonal M steps from one word (global
//    loads only where the path leaves that strip), and pushes CIGAR runs into
//    the consumed top of its traceback area.
// ---------------------------------------------------------------------------
constexpr int32_t QUAD_GUARD = -(1 << 28);
constexpr int64_t QUAD_NO_JOB = INT64_MIN;  // endpoint-key slot of a job k_dp_quad did not run (no key is this low)
constexpr int32_t QUAD_EXCL = -(1 << 29);

struct QuadLayout {  // byte offsets inside one job's LDS area
  uint32_t tb, gc, scw, q, qq, ref, stride, tb_words;
};
__host__ __device__ inline QuadLayout quad_layout(uint32_t max_m, uint32_t max_n) {
  QuadLayout l;
  const uint32_t blocks = (max_m + 7) / 8 + 1;  // + one block for the CIGAR runs
  l.tb_words = 64 * blocks;  // per job, in global scratch (DpArgs::tbq)
  l.tb = 0;
  l.gc = 0;
  l.scw = l.gc + 8 * (max_m + 64);
  l.q = l.scw + 4 * ((max_m > 128 ? max_m : 128) + 4);
  l.qq = l.q + ((max_m + 15) & ~15u);
  l.ref = l.qq + ((max_m + 15) & ~15u);
  l.stride = (l.ref + ((max_n + 15) & ~15u) + 15) & ~15u;  // 8-byte keys and terms in every group's area
  return l;
}

template <int CTRL>
__device__ __forceinline__ int32_t dpp_mov(int32_t old, int32_t x) {
  return __builtin_amdgcn_update_dpp(old, x, CTRL, 0xf, 0xf, false);
}
template <int CTRL>
__device__ __forceinline__ int32_t dpp_mov_z(int32_t x) {  // lanes without a source read 0
  return __builtin_amdgcn_update_dpp(0, x, CTRL, 0xf, 0xf, true);
}
constexpr int DPP_ROW_SHL1 = 0x101;
constexpr int DPP_ROW_SHR1 = 0x111, DPP_ROW_SHR2 = 0x112, DPP_ROW_SHR4 = 0x114, DPP_ROW_SHR8 = 0x118;
constexpr int DPP_ROW_SHL2 = 0x102, DPP_ROW_SHL4 = 0x104, DPP_ROW_SHL8 = 0x108;

__device__ __forceinline__ uint32_t tb_push_bit(uint32_t w, int32_t x) {  // (w << 1) | sign(x)
  return __builtin_amdgcn_alignbit(w, (uint32_t)x, 31);
}
__device__ __forceinline__ int32_t sub_sat(int32_t x, int32_t y) {  // x - y clamped to int32
  return __builtin_elementwise_sub_sat(x, y);
}
__device__ __forceinline__ int32_t bfe_i32(int32_t x, int32_t off, int32_t w) {
  return __builtin_amdgcn_sbfe(x, (uint32_t)off, (uint32_t)w);
}

// Sum over rows 1..m of the reference's computed cells per row,
// min(i+h+b, n) - max(i+h-b, 1) + 1 (all rows computed).
__device__ uint64_t quad_cells(int64_t m, int64_t h, int64_t b, int64_t n) {
  const int64_t c = h + b;
  int64_t t = n - c;
  t = t < 0 ? 0 : (t > m ? m : t);
  const int64_t hi = t * (t + 1) / 2 + c * t + (m - t) * n;
  const int64_t c2 = h - b;
  int64_t u = -c2;
  u = u < 0 ? 0 : (u > m ? m : u);
  const int64_t lo = u + (m * (m + 1) / 2 - u * (u + 1) / 2) + c2 * (m - u);
  return (uint64_t)(hi - lo + m);
}

// Three waves per SIMD (held to 128 VGPRs for four it spills and runs slower).
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) void k_dp_quad(DpArgs a) {
  extern __shared__ __align__(16) uint8_t dyn[];
  const int lane = lane_id();
  const int g = lane >> 4, L = lane & 15;
  const QuadLayout Y = quad_layout(a.max_m, a.max_n);
  uint8_t* jb = dyn + (size_t)g * Y.stride;
  // per window column c at [c - (1 + o)]: gap open, then 6 * (A, C, G, T, N -> 0..4)
  int2* GC = (int2*)(jb + Y.gc);
  int32_t* SCW = (int32_t*)(jb + Y.scw);
  uint8_t* Q = jb + Y.q;
  uint8_t* QQ = jb + Y.qq;
  uint8_t* REF = jb + Y.ref;
  const int32_t ge = a.gap_extend;
  int32_t gev;  // the same value held per lane
  asm volatile("v_mov_b32 %0, %1" : "=v"(gev) : "s"(ge));
  const uint32_t nwaves = gridDim.x;
  const uint32_t nquads = (a.njobs + 3) / 4;
  for (uint32_t wq = blockIdx.x; wq < nquads; wq += nwaves) {
    // ---- this group's job and its shape ----
    const uint32_t t = wq * 4 + (uint32_t)g;
    // the job's traceback planes in global scratch (read by k_dp_walk): LDS keeps
    // only the per-row inputs, so more waves fit per CU
    uint32_t* TB = a.tbq + (uint64_t)t * Y.tb_words;
    bool valid = t < a.njobs;
    uint32_t jid = 0;
    if (valid) {
      jid = a.jlist ? a.jlist[t] : t;
      if (a.job_valid && (jid % TOP_CANDIDATES) >= a.job_valid[jid / TOP_CANDIDATES]) valid = false;
    }
    DpJob job{};
    uint32_t m = 0, qlen = 0, n = 0;
    uint64_t so = 0, qo = 0;
    int32_t h = 0, b = 0;
    if (valid) {
      job = a.jobs[jid];
      so = a.R.seq_off[job.read];
      qo = a.R.qual_off[job.read];
      m = (uint32_t)(a.R.seq_off[job.read + 1] - so);
      qlen = (uint32_t)(a.R.qual_off[job.read + 1] - qo);
      n = job.n;
      h = job.hint;
      b = job.band;
      const bool regular = m >= 1 && n >= 1 && m <= a.max_m && n <= a.max_n && n <= (uint32_t)QUAD_MAX_N && b >= 0 && b <= 31 &&
                           !job.explicit_ref && (int64_t)h + b >= 0 && (int64_t)m + h - b <= (int64_t)n;
      if (!regular) {
        if (L == 0) dp_flag(a, jid);
        valid = false;
      }
    }
    if (!valid) {
      m = 0;
      n = 0;
      h = 0;
      b = 0;
    }
    const int32_t W = 2 * b + 1, s = 63 - W, o = h + b - 62;
    const uint32_t ql = qlen < m ? qlen : m;
    // ---- staging: query (reverse complement for reverse-strand jobs, single.rs:101-103) ----
    // Loads first, at clamped addresses so they are all in flight together, then the stores.
    for (uint32_t x = L; x < m; x += 16) Q[x] = job.rc ? comp_base(a.R.seq[so + m - 1 - x]) : a.R.seq[so + x];
    for (uint32_t x = L; x < ql; x += 16) QQ[x] = job.rc ? a.R.qual[qo + qlen - 1 - x] : a.R.qual[qo + x];
    // reference window: 2-bit words and N-mask words through LDS, then bytes
    uint64_t* gw = (uint64_t*)SCW;
    const uint64_t gb0 = job.ref_gb;
    const uint32_t nw = m ? (uint32_t)(((gb0 & 31) + n + 31) >> 5) : 0;
    const uint32_t nwm = m ? (uint32_t)(((gb0 & 63) + n + 63) >> 6) : 0;
    for (uint32_t x = L; x < nw; x += 16) gw[x] = a.G.g2b[(gb0 >> 5) + x];
    for (uint32_t x = L; x < nwm; x += 16) gw[nw + x] = a.G.nmask[(gb0 >> 6) + x];
    wave_sync();
    for (uint32_t x = L; x < n && m; x += 16) {
      const uint64_t p = (gb0 & 31) + x, pm = (gb0 & 63) + x;
      const uint32_t c2 = (uint32_t)(gw[p >> 5] >> ((p & 31) * 2)) & 3u;
      const bool isn = (gw[nw + (pm >> 6)] >> (pm & 63)) & 1ull;
      REF[x] = isn ? (uint8_t)'N' : (uint8_t)("ACGT"[c2]);
    }
    wave_sync();
    constexpr int NR = QUAD_MAX_N / 16;
    // homopolymer run of every column (smith_waterman.rs:348-363): each lane
    // takes a contiguous segment in registers; run starts carry in by a prefix
    // max and run ends by a suffix min over the 16 lanes
    {
      const uint32_t seg = (n + 15) / 16;  // <= NR
      const uint32_t x0 = min(n, L * seg), cnt = min(n, x0 + seg) - x0;
      uint32_t bb[NR + 1];
      const uint32_t prevb = (x0 > 0 && x0 < n) ? REF[x0 - 1] : 0x100u;  // no byte equals 0x100
#pragma unroll
      for (int u = 0; u <= NR; u++) {
        const uint32_t x = x0 + u;
        const uint32_t v = REF[n ? min(x, n - 1) : 0];
        bb[u] = x < n ? v : 0x101u;
      }
      int32_t lb = -1, fe = INT32_MAX;
#pragma unroll
      for (int u = NR - 1; u >= 0; u--) {
        if ((uint32_t)u < cnt) {
          if (bb[u] != bb[u + 1]) fe = (int32_t)(x0 + u);
        }
      }
#pragma unroll
      for (int u = 0; u < NR; u++) {
        if ((uint32_t)u < cnt) {
          if (bb[u] != (u == 0 ? prevb : bb[u - 1])) lb = (int32_t)(x0 + u);
        }
      }
      int32_t pm = lb;
      pm = max(pm, dpp_mov<DPP_ROW_SHR1>(-1, pm));
      pm = max(pm, dpp_mov<DPP_ROW_SHR2>(-1, pm));
      pm = max(pm, dpp_mov<DPP_ROW_SHR4>(-1, pm));
      pm = max(pm, dpp_mov<DPP_ROW_SHR8>(-1, pm));
      int32_t start = dpp_mov<DPP_ROW_SHR1>(-1, pm);
      int32_t sm = fe;
      sm = min(sm, dpp_mov<DPP_ROW_SHL1>(INT32_MAX, sm));
      sm = min(sm, dpp_mov<DPP_ROW_SHL2>(INT32_MAX, sm));
      sm = min(sm, dpp_mov<DPP_ROW_SHL4>(INT32_MAX, sm));
      sm = min(sm, dpp_mov<DPP_ROW_SHL8>(INT32_MAX, sm));
      int32_t end = dpp_mov<DPP_ROW_SHL1>(INT32_MAX, sm);
      // window columns c = x + 1 at index c - (1 + o); default for c outside [1, n]
      const int32_t c_lo = 1 + o;
      const int32_t nent = (int32_t)m + 63;
      for (int32_t e = L; e < nent && m; e += 16) GC[e] = make_int2(-1, 24);
      int32_t st[NR];
#pragma unroll
      for (int u = 0; u < NR; u++) {
        if (bb[u] != (u == 0 ? prevb : bb[u - 1])) start = (int32_t)(x0 + u);
        st[u] = start;
      }
      wave_sync();
#pragma unroll
      for (int u = NR - 1; u >= 0; u--) {
        const int32_t x = (int32_t)x0 + u;
        if ((uint32_t)u < cnt) {
          if (bb[u] != bb[u + 1]) end = x;
          const int32_t e = x + 1 - c_lo;
          if (e >= 0 && e < nent) {
            int32_t run = end - st[u] + 1;
            if (run > 65535) run = 65535;
            const uint32_t rb = bb[u];
            GC[e] = make_int2(min(a.gap_open + (run - 1), -1),
                              rb == 'A' ? 0 : rb == 'C' ? 6 : rb == 'G' ? 12 : rb == 'T' ? 18 : 24);
          }
        }
      }
    }
    wave_sync();
    // per-row score word: 6-bit signed fields for reference A, C, G, T, N;
    // mismatch_penalty(q) = round(-30 clamp(q/30, 0.2, 1)) = -clamp(q, 6, 30)
    for (uint32_t x = L; x < m; x += 16) {
      const uint8_t rb = Q[x];
      const uint32_t qv = x < qlen ? QQ[x] : 30u;
      const int32_t mis = -(int32_t)min(max(qv, 6u), 30u);
      uint32_t w = 0;
      w |= (uint32_t)((rb == 'A' ? MATCH_SCORE : mis) & 63) << 0;
      w |= (uint32_t)((rb == 'C' ? MATCH_SCORE : mis) & 63) << 6;
      w |= (uint32_t)((rb == 'G' ? MATCH_SCORE : mis) & 63) << 12;
      w |= (uint32_t)((rb == 'T' ? MATCH_SCORE : mis) & 63) << 18;
      w |= (uint32_t)((rb == 'N' ? MATCH_SCORE : mis) & 63) << 24;
      SCW[x] = (int32_t)w;
    }
    wave_sync();
    // ---- row 0 and per-column constants ----
    int32_t Mo[4], Io[4], Po[4], kge[4], cst[4];
    const int32_t cp = (L == 15) ? QUAD_GUARD : 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const int32_t d = 4 * L + k;
      kge[k] = d * ge;
      cst[k] = -d * ge + (d <= s ? QUAD_EXCL : 0);
      const int32_t j0 = o + d;
      Mo[k] = (d >= s && j0 >= 0 && j0 <= (int32_t)n) ? 0 : NEG_INF;
      Io[k] = NEG_INF;
      Po[k] = NEG_INF - kge[k];
    }
    uint32_t w0 = 0, w1 = 0, w2 = 0, w3 = 0;
    uint32_t mmax = m;
    int32_t emax = m ? b - h : 0;  // rows i <= band - hint reach columns j <= 0
    for (int gg = 0; gg < 64; gg += 16) {
      mmax = max(mmax, (uint32_t)readlane_i32((int32_t)m, gg));
      emax = max(emax, readlane_i32(m ? b - h : 0, gg));
    }
    // Inputs of one row: score word and, per column of this lane, the gap
    // open and the base code.  Row i+1's inputs are loaded while row i runs.
    struct RowIn {
      int32_t scw, g0, g1, g2, g3, c0, c1, c2, c3;
    };
    auto load_row = [&](RowIn& x, uint32_t i) {  // in bounds for i <= m + 1
      const int2* gc = GC + (i - 1 + 4 * (uint32_t)L);
      x.scw = SCW[i - 1];
      const int2 v0 = gc[0], v1 = gc[1], v2 = gc[2], v3 = gc[3];
      x.g0 = v0.x;
      x.g1 = v1.x;
      x.g2 = v2.x;
      x.g3 = v3.x;
      x.c0 = v0.y;
      x.c1 = v1.y;
      x.c2 = v2.y;
      x.c3 = v3.y;
    };
    const int32_t cp2 = (L == 15) ? QUAD_GUARD : 0;  // the guard column's vertical neighbour
    // One row of the band.  EDGE: reset cells with j <= 0 (rows i <= band - hint).
    auto row = [&](uint32_t i, const RowIn& R, auto edge_tag) {
      constexpr bool EDGE = decltype(edge_tag)::value;
      // out-of-row lanes read 0: the guard column adds its sentinel, lane 0's
      // left neighbour is excluded from the scan
      const int32_t mu3 = dpp_mov_z<DPP_ROW_SHL1>(Mo[0]);
      const int32_t iu3 = dpp_mov_z<DPP_ROW_SHL1>(Io[0]);
      int32_t M0, M1, M2, M3, I0, I1, I2, I3;
      // M = max(M, I, D of the diagonal predecessor) + score: the predecessor
      // whose value is below the maximum sets its plane bit (w0: M, w1: I; the
      // row flush keeps w0 & w1, "M came from D")
      {
        const int32_t X = max(max(Mo[0], Io[0]), Po[0] + kge[0]);
        M0 = X + bfe_i32(R.scw, R.c0, 6);
        const int32_t io = Mo[1] + R.g0;
        I0 = max(io, Io[1] + gev);
        w0 = tb_push_bit(w0, sub_sat(Mo[0], X));
        w1 = tb_push_bit(w1, sub_sat(Io[0], X));
        w2 = tb_push_bit(w2, sub_sat(io, I0));
      }
      {
        const int32_t X = max(max(Mo[1], Io[1]), Po[1] + kge[1]);
        M1 = X + bfe_i32(R.scw, R.c1, 6);
        const int32_t io = Mo[2] + R.g1;
        I1 = max(io, Io[2] + gev);
        w0 = tb_push_bit(w0, sub_sat(Mo[1], X));
        w1 = tb_push_bit(w1, sub_sat(Io[1], X));
        w2 = tb_push_bit(w2, sub_sat(io, I1));
      }
      {
        const int32_t X = max(max(Mo[2], Io[2]), Po[2] + kge[2]);
        M2 = X + bfe_i32(R.scw, R.c2, 6);
        const int32_t io = Mo[3] + R.g2;
        I2 = max(io, Io[3] + gev);
        w0 = tb_push_bit(w0, sub_sat(Mo[2], X));
        w1 = tb_push_bit(w1, sub_sat(Io[2], X));
        w2 = tb_push_bit(w2, sub_sat(io, I2));
      }
      {
        const int32_t X = max(max(Mo[3], Io[3]), Po[3] + kge[3]);
        M3 = X + bfe_i32(R.scw, R.c3, 6) + cp;
        const int32_t io = mu3 + R.g3 + cp2;
        I3 = max(io, iu3 + gev + cp2);
        w0 = tb_push_bit(w0, sub_sat(Mo[3], X));
        w1 = tb_push_bit(w1, sub_sat(Io[3], X));
        w2 = tb_push_bit(w2, sub_sat(io, I3));
      }
      const int32_t jl = (int32_t)i + o + 4 * L;  // j of column k is jl + k
      if (EDGE) {
        if (jl + 0 <= 0) M0 = I0 = NEG_INF;
        if (jl + 1 <= 0) M1 = I1 = NEG_INF;
        if (jl + 2 <= 0) M2 = I2 = NEG_INF;
        if (jl + 3 <= 0) M3 = I3 = NEG_INF;
      }
      // D: max-plus prefix scan along the row
      const int32_t ml0 = dpp_mov_z<DPP_ROW_SHR1>(M3);
      const int32_t B0 = ml0 + R.g0 + cst[0], B1 = M0 + R.g1 + cst[1], B2 = M1 + R.g2 + cst[2],
                    B3 = M2 + R.g3 + cst[3];
      const int32_t L0 = B0, L1 = max(L0, B1), L2 = max(L1, B2), L3 = max(L2, B3);
      int32_t x = L3;
      x = max(x, dpp_mov<DPP_ROW_SHR1>(INT32_MIN, x));
      x = max(x, dpp_mov<DPP_ROW_SHR2>(INT32_MIN, x));
      x = max(x, dpp_mov<DPP_ROW_SHR4>(INT32_MIN, x));
      x = max(x, dpp_mov<DPP_ROW_SHR8>(INT32_MIN, x));
      const int32_t carry = dpp_mov<DPP_ROW_SHR1>(INT32_MIN, x);
      int32_t P0 = max(L0, carry), P1 = max(L1, carry), P2 = max(L2, carry), P3 = max(L3, carry);
      w3 = tb_push_bit(w3, sub_sat(B0, P0));
      w3 = tb_push_bit(w3, sub_sat(B1, P1));
      w3 = tb_push_bit(w3, sub_sat(B2, P2));
      w3 = tb_push_bit(w3, sub_sat(B3, P3));
      if (EDGE) {
        if (jl + 0 <= 0) P0 = NEG_INF - kge[0];
        if (jl + 1 <= 0) P1 = NEG_INF - kge[1];
        if (jl + 2 <= 0) P2 = NEG_INF - kge[2];
        if (jl + 3 <= 0) P3 = NEG_INF - kge[3];
      }
      Mo[0] = M0;
      Mo[1] = M1;
      Mo[2] = M2;
      Mo[3] = M3;
      Io[0] = I0;
      Io[1] = I1;
      Io[2] = I2;
      Io[3] = I3;
      Po[0] = P0;
      Po[1] = P1;
      Po[2] = P2;
      Po[3] = P3;
      if ((i & 7u) == 0) {
        *reinterpret_cast<uint4*>(TB + ((i - 1) >> 3) * 64 + L * 4) = make_uint4(w0, w0 & w1, w2, w3);
        w0 = w1 = w2 = w3 = 0;
      }
    };
    RowIn RA, RB;
    load_row(RA, 1);
    uint32_t i = 1;
    // rows the band reaches j <= 0 in, then the rest; two rows per trip so the
    // prefetch alternates between RA and RB without copies
    const uint32_t elim = (uint32_t)min((int32_t)mmax, emax);
    for (; i + 1 <= elim; i += 2) {
      load_row(RB, i + 1);
      if (i <= m) row(i, RA, std::true_type{});
      load_row(RA, i + 2);
      if (i + 1 <= m) row(i + 1, RB, std::true_type{});
    }
    if (i <= elim) {
      load_row(RB, i + 1);
      if (i <= m) row(i, RA, std::true_type{});
      RA = RB;
      i++;
    }
    for (; i + 1 <= mmax; i += 2) {
      load_row(RB, i + 1);
      if (i <= m) row(i, RA, std::false_type{});
      load_row(RA, i + 2);
      if (i + 1 <= m) row(i + 1, RB, std::false_type{});
    }
    if (i <= mmax) {
      load_row(RB, i + 1);
      if (i <= m) row(i, RA, std::false_type{});
    }
    if (m & 7u) {
      const uint32_t sh = 4 * (8 - (m & 7u));
      *reinterpret_cast<uint4*>(TB + ((m - 1) >> 3) * 64 + L * 4) =
          make_uint4(w0 << sh, (w0 & w1) << sh, w2 << sh, w3 << sh);
    }
    // ---- endpoint: last maximum over (j asc; M then I) of row m ----
    int64_t key = end_key(NEG_INF, 2 * n + 1);
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const int32_t d = 4 * L + k, j = (int32_t)m + o + d;
      if (m && d >= s && d < s + W && j >= 1 && j <= (int32_t)n) {
        const int64_t km = end_key(Mo[k], (uint32_t)(2 * j)), ki = end_key(Io[k], (uint32_t)(2 * j + 1));
        key = key > km ? key : km;
        key = key > ki ? key : ki;
      }
    }
    wave_sync();
    int64_t* keys = (int64_t*)SCW;
    keys[L] = key;
    wave_sync();
    mem_sync();  // the traceback words stored by every lane
    if (L == 0 && t < a.njobs) {  // the endpoint key for k_dp_walk (QUAD_NO_JOB: not this kernel's job)
      int64_t best = keys[0];
      for (int u = 1; u < 16; u++) best = best > keys[u] ? best : keys[u];
      a.tbend[t] = m ? best : QUAD_NO_JOB;
    }
    wave_sync();
  }
}

// ---------------------------------------------------------------------------
// Traceback and outputs of k_dp_quad's jobs, one lane per job
// (smith_waterman.rs:299-345): every lane of the wave walks a job, where inside
// k_dp_quad one lane of each 16-lane row would.
// The walk takes a whole run per step from at most two traceback words of the
// job's area ([8-row block][lane][plane], written by k_dp_quad) -- diagonal M
// steps (rows r, r-1, .. of a column, into the block below), horizontal D steps
// (columns k, k-1, .., into the lane to the left) or vertical I steps (cells
// (r, k), (r-1, k+1), .., into the lane to the right) -- with the selection
// branch-free; the pending CIGAR run is rewritten in place every step and its
// slot advances when the operation changes, the runs going into the consumed
// top of the area, where k_dp_lk reads them.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_dp_walk(DpArgs a) {
  const uint32_t t = blockIdx.x * 256 + threadIdx.x;
  const QuadLayout Y = quad_layout(a.max_m, a.max_n);
  unsigned long long c_cells = 0;
  const int64_t best = t < a.njobs ? a.tbend[t] : QUAD_NO_JOB;
  if (best != QUAD_NO_JOB) {
    const uint32_t jid = a.jlist ? a.jlist[t] : t;
    const DpJob job = a.jobs[jid];
    const uint32_t m = (uint32_t)(a.R.seq_off[job.read + 1] - a.R.seq_off[job.read]);
    const uint32_t n = job.n;
    const int32_t h = job.hint, b = job.band, o = h + b - 62;
    const uint32_t* TB = a.tbq + (uint64_t)t * Y.tb_words;
    uint32_t* runs_top = a.tbq + (uint64_t)t * Y.tb_words + Y.tb_words;
    uint32_t nruns = 0, off = 0;
    {
    uint32_t i = m, j = ((uint32_t)best) >> 1, st = ((uint32_t)best) & 1u;
    uint32_t cur_op = 0xFFu, cur_len = 0;
    int32_t cb = -1, cl = -1;
    uint32_t v0 = 0, v1 = 0, v2 = 0, v3 = 0, p0 = 0, p1 = 0, n2 = 0, l3 = 0, have = 0;
    while (i > 0) {
      int32_t d = (int32_t)j - (int32_t)i - o;
      d = d < 0 ? 0 : (d > 63 ? 63 : d);  // cells on the path always lie in the band
      const int32_t lq = d >> 2, rb = (int32_t)(i - 1) >> 3;
      // select: reference column 0 forces a single I step with the state kept
      const uint32_t sel = (j == 0) ? 1u : st;
      if (rb != cb || lq != cl) {  // this lane's four planes
        const uint4 tv = *reinterpret_cast<const uint4*>(TB + rb * 64 + lq * 4);
        v0 = tv.x;
        v1 = tv.y;
        v2 = tv.z;
        v3 = tv.w;
        cb = rb;
        cl = lq;
        have = 0;
      }
      // the words only one state reads, loaded once per (block, lane) when that
      // state walks: the two M planes of the block above (M), plane 2 of the lane
      // to the right (I), plane 3 of the lane to the left (D); own lane at the edges
      if (sel == 0u && !(have & 1u)) {
        const uint2 pv = *reinterpret_cast<const uint2*>(TB + (rb > 0 ? rb - 1 : 0) * 64 + lq * 4);
        p0 = pv.x;
        p1 = pv.y;
        have |= 1u;
      } else if (sel == 1u && !(have & 2u)) {
        n2 = lq < 15 ? TB[rb * 64 + lq * 4 + 6] : v2;
        have |= 2u;
      } else if (sel == 2u && !(have & 4u)) {
        l3 = lq > 0 ? TB[rb * 64 + lq * 4 - 1] : v3;
        have |= 4u;
      }
      const uint32_t k = (uint32_t)d & 3u, r = (i - 1) & 7u;
      const uint32_t pos = 31u - (4u * r + k);
      // M: stops after the first row whose M came from I or D
      const uint32_t mav = r + 1u + (rb > 0 ? 8u : 0u);
      const uint64_t xm = ((((uint64_t)p0 << 32) | v0) >> pos) & 0x1111111111111111ull &
                          (mav >= 16u ? ~0ull : ((1ull << (4u * mav)) - 1ull));
      const uint32_t f = xm ? ((uint32_t)__builtin_ctzll(xm) >> 2) : 64u;
      const uint32_t mst = min(min(f + 1u, mav), j);
      const uint64_t x1 = (((uint64_t)p1 << 32) | v1) >> pos;
      const uint32_t mpar = (mst == f + 1u) ? 1u + (uint32_t)((x1 >> ((4u * f) & 63u)) & 1ull) : 0u;
      // D: stops after the first column whose D was opened
      const uint32_t dav = k + 1u + (lq > 0 ? 4u : 0u);
      const uint32_t yd = ((v3 >> pos) & ((2u << k) - 1u)) | (((l3 >> (28u - 4u * r)) & 0xFu) << (k + 1u));
      const uint32_t y = ~yd & ((1u << dav) - 1u);
      const uint32_t fz = y ? (uint32_t)__builtin_ctz(y) : 32u;
      const uint32_t dst = min(min(fz + 1u, dav), j);
      const uint32_t dpar = (dst == fz + 1u) ? 0u : 2u;
      // I: stops after the first cell whose I was opened (stride-3 bits)
      const uint32_t tav = min((lq < 15 ? 7u : 3u) - k, r);
      const uint32_t ma = 0x249249u & ((1u << (3u * (3u - k) + 1u)) - 1u);
      const uint32_t zv = ((v2 >> pos) & ma) | ((r > 0 ? (n2 >> ((pos + 4u) & 31u)) : 0u) & 0x249249u & ~ma);
      const uint32_t zi = ~zv & 0x249249u & ((1u << (3u * tav + 1u)) - 1u);
      const uint32_t fi = zi ? (uint32_t)__builtin_ctz(zi) / 3u : 32u;
      const uint32_t ist = min(fi + 1u, tav + 1u);
      const uint32_t ipar = (ist == fi + 1u) ? 0u : 1u;
      const uint32_t isM = 0u - (uint32_t)(sel == 0u), isD = 0u - (uint32_t)(sel == 2u);
      const uint32_t isI = ~(isM | isD);
      const uint32_t jz = 0u - (uint32_t)(j == 0);
      const uint32_t len = (mst & isM) | (dst & isD) | (((ist & ~jz) | (1u & jz)) & isI);
      const uint32_t nst = (mpar & isM) | (dpar & isD) | (((ipar & ~jz) | (st & jz)) & isI);
      i -= len & ~isD;
      j -= len & ~isI;
      st = nst;
      const uint32_t fresh = 0u - (uint32_t)(sel != cur_op);
      nruns += (fresh & (uint32_t)(cur_len != 0)) ? 1u : 0u;
      cur_len = (len & fresh) | ((cur_len + len) & ~fresh);
      cur_op = sel;
      runs_top[-1 - (int32_t)nruns] = (cur_len << 4) | cur_op;
    }
      nruns += cur_len != 0 ? 1u : 0u;
      off = j;
    }
    const uint32_t* runs = runs_top - nruns;
    uint32_t span = 0;
    for (uint32_t u = 0; u < nruns; u++) {
      const uint32_t w = runs[u], op = w & 15u;
      if (op == 0u || op == 2u) span += w >> 4;
    }
    const uint64_t cells = quad_cells(m, h, b, n);
    c_cells = cells;
    DpOut oo;
    oo.log_lk = 0.0;  // k_dp_lk's
    oo.ref_off = off;
    oo.score = (int32_t)(best >> 32);
    oo.cigar_n = nruns;
    oo.span = span;
    oo.cells = cells;
    a.out[jid] = oo;
    // ---- CIGAR runs to the job's slot, forward order ----
    const uint32_t ncopy = nruns < a.cigar_cap ? nruns : a.cigar_cap;
    uint32_t* dst = a.cigar + ((uint64_t)jid - a.cigar_base) * a.cigar_cap;
    for (uint32_t u = 0; u < ncopy; u++) dst[u] = runs[u];
  }
  for (int sft = 32; sft > 0; sft >>= 1) c_cells += (unsigned long long)__shfl_xor((long long)c_cells, sft, 64);
  if (lane_id() == 0 && a.ctr && c_cells) atomicAdd(&a.ctr->cells, c_cells);
}

// The alignment log-likelihood of k_dp_quad's jobs (single.rs:228-262,
// bayesian.rs:142-179), sixteen lanes per job: each round the lanes take 16
// consecutive read positions (coalesced read, quality and genome loads) and
// find their CIGAR run with a cursor that only moves forward; M positions add
// ln(1-e) or ln(e/3) by base match against the window, I positions ln(1/4),
// and the group's first lane adds the 16 terms in read order, as the
// reference sums them.
__global__ __launch_bounds__(256) void k_dp_lk(DpArgs a) {
  __shared__ double lut_m[256], lut_x[256];  // per-base log-likelihoods of a match / mismatch by quality
  __shared__ double terms[16][16];
  lut_m[threadIdx.x] = a.T.lk_match[threadIdx.x];
  lut_x[threadIdx.x] = a.T.lk_mis[threadIdx.x];
  __syncthreads();
  const uint32_t grp = threadIdx.x >> 4, L = threadIdx.x & 15u;
  const uint32_t t = blockIdx.x * 16 + grp;
  const QuadLayout Y = quad_layout(a.max_m, a.max_n);
  const int64_t best = t < a.njobs ? a.tbend[t] : QUAD_NO_JOB;
  if (best != QUAD_NO_JOB) {
    const uint32_t jid = a.jlist ? a.jlist[t] : t;
    const DpJob job = a.jobs[jid];
    const uint64_t so = a.R.seq_off[job.read], qo = a.R.qual_off[job.read];
    const uint32_t m = (uint32_t)(a.R.seq_off[job.read + 1] - so);
    const uint32_t qlen = (uint32_t)(a.R.qual_off[job.read + 1] - qo);
    const DpOut& oo = a.out[jid];
    const uint32_t nruns = oo.cigar_n, off = oo.ref_off;
    const uint32_t* runs = a.tbq + (uint64_t)t * Y.tb_words + Y.tb_words - nruns;
    uint32_t u = 0, qs = 0, rs = 0;
    double lk = 0.0;
    for (uint32_t base = 0; base < m; base += 16) {
      const uint32_t qi = base + L;
      double tv = 0.0;
      if (qi < m) {
        uint32_t op = 1u;
        while (u < nruns) {  // the run that holds read position qi
          const uint32_t w = runs[u], len = w >> 4;
          op = w & 15u;
          const uint32_t qc = op <= 1u ? len : 0u;
          if (qi < qs + qc) break;
          qs += qc;
          rs += (op == 0u || op == 2u) ? len : 0u;
          u++;
        }
        tv = a.T.ln_quarter;
        if (op == 0u) {
          const uint32_t qb = job.rc ? comp_base(a.R.seq[so + m - 1 - qi]) : a.R.seq[so + qi];
          const uint32_t qv = qi < qlen ? (job.rc ? a.R.qual[qo + qlen - 1 - qi] : a.R.qual[qo + qi]) : 30u;
          const uint64_t p = job.ref_gb + off + rs + (qi - qs);
          const uint32_t c2 = (uint32_t)(a.G.g2b[p >> 5] >> ((p & 31) * 2)) & 3u;
          const uint32_t gb = ((a.G.nmask[p >> 6] >> (p & 63)) & 1ull) ? (uint32_t)'N' : (uint32_t)(uint8_t)("ACGT"[c2]);
          tv = qb == gb ? lut_m[qv] : lut_x[qv];
        }
      }
      terms[grp][L] = tv;
      wave_sync();
      if (L == 0) {
        const uint32_t c = min(16u, m - base);
        double t16[16];
#pragma unroll
        for (uint32_t v = 0; v < 16; v++) t16[v] = terms[grp][v];
#pragma unroll
        for (uint32_t v = 0; v < 16; v++)
          if (v < c) lk += t16[v];
      }
      wave_sync();
    }
    if (L == 0) a.out[jid].log_lk = lk;
  }
}

// ---------------------------------------------------------------------------
// DP, eight jobs per wave in 16-bit halves (k_dp_pair16): the k_dp_quad
// recurrence with two jobs per 16-lane DPP row, job A in the low and job B
// in the high half of every 32-bit register, so one v_pk_* instruction
// advances two cells and one DPP move serves both scans.
//  - Saturating 16-bit arithmetic.  A cell on the optimal path whose score is
//    s has every argmax predecessor >= s - 20 per row, so for a job whose best
//    score is >= PAIR16_MIN_SCORE all those values stay >= -25,000 (reads of
//    at most 320 bases) and are exact, while every value derived from a
//    saturated one stays <= -32,768 + 320 * 20 = -26,368 below them: the best
//    score, its endpoint and every traceback decision on the path equal the
//    32-bit kernel's.  A job whose best is lower is flagged to the
//    one-job-per-wave kernel.  The launcher takes this kernel only when
//    -4 <= gap extend and gap open >= -400, which bound the shifts of the D
//    scan's representation.
//  - Plane bits: the signs of both halves of a saturating difference, pushed
//    into one word per plane (bit 0 and bit 16); every 4 rows (16 pushes) the
//    words are split per job into 16-bit plane fields, 8 bytes per lane and
//    4-row block.  The walk joins two blocks into the 32-bit 8-row words of
//    k_dp_quad and runs k_dp_quad's walker unchanged.
//  - Walk, log-likelihood terms, ordered sum and outputs per job as in
//    k_dp_quad; lane h of a row walks and sums its job h, so eight lanes of
//    the wave do the serial work.
// ---------------------------------------------------------------------------
constexpr int32_t PAIR16_MIN_SCORE = -18000;

// The shapes k_dp_quad runs (the rest is flagged to the one-job-per-wave kernel).
__device__ __forceinline__ bool quad_regular(const DpArgs& a, const DpJob& job, uint32_t m, uint32_t n) {
  const int32_t h = job.hint, b = job.band;
  return m >= 1 && n >= 1 && m <= a.max_m && n <= a.max_n && n <= (uint32_t)QUAD_MAX_N && b >= 0 && b <= 31 &&
         !job.explicit_ref && (int64_t)h + b >= 0 && (int64_t)m + h - b <= (int64_t)n;
}
constexpr uint32_t PK_NEG = 0x80008000u;  // -32768 in both halves

typedef short s16x2_t __attribute__((ext_vector_type(2)));
typedef unsigned short u16x2_t __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t pk_add(uint32_t x, uint32_t y) {
  return __builtin_bit_cast(uint32_t, __builtin_elementwise_add_sat(__builtin_bit_cast(s16x2_t, x), __builtin_bit_cast(s16x2_t, y)));
}
__device__ __forceinline__ uint32_t pk_sub(uint32_t x, uint32_t y) {
  return __builtin_bit_cast(uint32_t, __builtin_elementwise_sub_sat(__builtin_bit_cast(s16x2_t, x), __builtin_bit_cast(s16x2_t, y)));
}
__device__ __forceinline__ uint32_t pk_max(uint32_t x, uint32_t y) {
  return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(__builtin_bit_cast(s16x2_t, x), __builtin_bit_cast(s16x2_t, y)));
}
// (w << 1) with the sign of d's low half at bit 0 and of its high half at bit 16
__device__ __forceinline__ uint32_t pk_push(uint32_t w, uint32_t d) {
  const uint32_t sg = __builtin_bit_cast(uint32_t, __builtin_bit_cast(u16x2_t, d) >> (unsigned short)15);
  return (w << 1) | sg;
}
__device__ __forceinline__ uint32_t pk_make(int32_t lo, int32_t hi) {
  return ((uint32_t)lo & 0xFFFFu) | ((uint32_t)hi << 16);
}
__device__ __forceinline__ int32_t pk_lo(uint32_t v) { return (int32_t)(int16_t)(v & 0xFFFFu); }
__device__ __forceinline__ int32_t pk_hi(uint32_t v) { return (int32_t)(int16_t)(v >> 16); }

struct Pair16Layout {  // byte offsets inside one row's (two jobs') LDS area
  uint32_t gc, scw, q, qq, ref, keys, stride, tb_words;
  uint32_t scw_len, q_len, ref_len;
};
__host__ __device__ inline Pair16Layout pair16_layout(uint32_t max_m, uint32_t max_n) {
  Pair16Layout l;
  const uint32_t blocks = (max_m + 3) / 4 + 2;  // 4-row blocks + two for the CIGAR runs
  l.tb_words = 32 * blocks;                     // per job, in global scratch (DpArgs::tbq)
  l.gc = 0;                                     // (max_m + 64) x {gap open A | B, code A | code B << 8}
  l.scw_len = 4 * ((max_m > 128 ? max_m : 128) + 4);
  l.scw = l.gc + 8 * (max_m + 64);              // score words of job A, then of job B
  l.q_len = (max_m + 15) & ~15u;
  l.q = l.scw + 2 * l.scw_len;                  // Q A, Q B
  l.qq = l.q + 2 * l.q_len;                     // QQ A, QQ B
  l.ref_len = (max_n + 15) & ~15u;
  l.ref = l.qq + 2 * l.q_len;                   // REF A, REF B
  l.keys = (l.ref + 2 * l.ref_len + 7) & ~7u;    // 16 endpoint keys of job A, then of job B
  l.stride = (l.keys + 256 + 15) & ~15u;
  return l;
}

__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) void k_dp_pair16(DpArgs a) {
  extern __shared__ __align__(16) uint8_t dyn[];
  const int lane = lane_id();
  const int g = lane >> 4, L = lane & 15;
  const Pair16Layout Y = pair16_layout(a.max_m, a.max_n);
  uint8_t* jb = dyn + (size_t)g * Y.stride;
  uint2* GC = (uint2*)(jb + Y.gc);
  const int32_t ge = a.gap_extend;
  int32_t gev;  // the same value held per lane
  asm volatile("v_mov_b32 %0, %1" : "=v"(gev) : "s"(ge));
  const uint32_t ge_pk = pk_make(gev, gev);
  unsigned long long c_cells = 0;
  const uint32_t nwaves = gridDim.x;
  const uint32_t noct = (a.njobs + 7) / 8;
  for (uint32_t wq = blockIdx.x; wq < noct; wq += nwaves) {
    // ---- the row's two jobs and their shapes ----
    uint32_t jidv[2], mv[2], qlenv[2], nv[2];
    uint64_t sov[2], qov[2], gbv[2];
    int32_t hv[2], bv[2], ov[2], sv[2], Wv[2];
    bool rcv[2];
#pragma unroll
    for (int hh = 0; hh < 2; hh++) {
      const uint32_t t = wq * 8 + (uint32_t)g * 2 + (uint32_t)hh;
      bool valid = t < a.njobs;
      uint32_t jid = 0;
      if (valid) {
        jid = a.jlist[t];
        if (a.job_valid && (jid % TOP_CANDIDATES) >= a.job_valid[jid / TOP_CANDIDATES]) valid = false;
      }
      DpJob job{};
      uint32_t m = 0, qlen = 0, n = 0;
      uint64_t so = 0, qo = 0;
      if (valid) {
        job = a.jobs[jid];
        so = a.R.seq_off[job.read];
        qo = a.R.qual_off[job.read];
        m = (uint32_t)(a.R.seq_off[job.read + 1] - so);
        qlen = (uint32_t)(a.R.qual_off[job.read + 1] - qo);
        n = job.n;
        if (!quad_regular(a, job, m, n)) {
          if (L == 0) dp_flag(a, jid);
          valid = false;
        }
      }
      if (!valid) {
        m = 0;
        n = 0;
        job.hint = 0;
        job.band = 0;
      }
      jidv[hh] = jid;
      mv[hh] = m;
      qlenv[hh] = qlen;
      nv[hh] = n;
      sov[hh] = so;
      qov[hh] = qo;
      gbv[hh] = job.ref_gb;
      hv[hh] = job.hint;
      bv[hh] = job.band;
      rcv[hh] = job.rc != 0;
      Wv[hh] = 2 * job.band + 1;
      sv[hh] = 63 - Wv[hh];
      ov[hh] = job.hint + job.band - 62;
    }
    // ---- staging, one job after the other with the row's 16 lanes ----
    for (int32_t e = L; e < (int32_t)max(mv[0], mv[1]) + 63; e += 16) GC[e] = make_uint2(0xFFFFFFFFu, 24u | (24u << 8));
    wave_sync();
#pragma unroll
    for (int hh = 0; hh < 2; hh++) {
      const uint32_t m = mv[hh], qlen = qlenv[hh], n = nv[hh];
      const uint64_t so = sov[hh], qo = qov[hh];
      const bool rc = rcv[hh];
      const int32_t o = ov[hh];
      int32_t* SCW = (int32_t*)(jb + Y.scw + hh * Y.scw_len);
      uint8_t* Q = jb + Y.q + hh * Y.q_len;
      uint8_t* QQ = jb + Y.qq + hh * Y.q_len;
      uint8_t* REF = jb + Y.ref + hh * Y.ref_len;
      const uint32_t ql = qlen < m ? qlen : m;
      for (uint32_t x = L; x < m; x += 16) Q[x] = rc ? comp_base(a.R.seq[so + m - 1 - x]) : a.R.seq[so + x];
      for (uint32_t x = L; x < ql; x += 16) QQ[x] = rc ? a.R.qual[qo + qlen - 1 - x] : a.R.qual[qo + x];
      uint64_t* gw = (uint64_t*)SCW;
      const uint64_t gb0 = gbv[hh];
      const uint32_t nw = m ? (uint32_t)(((gb0 & 31) + n + 31) >> 5) : 0;
      const uint32_t nwm = m ? (uint32_t)(((gb0 & 63) + n + 63) >> 6) : 0;
      for (uint32_t x = L; x < nw; x += 16) gw[x] = a.G.g2b[(gb0 >> 5) + x];
      for (uint32_t x = L; x < nwm; x += 16) gw[nw + x] = a.G.nmask[(gb0 >> 6) + x];
      wave_sync();
      for (uint32_t x = L; x < n && m; x += 16) {
        const uint64_t p = (gb0 & 31) + x, pm = (gb0 & 63) + x;
        const uint32_t c2 = (uint32_t)(gw[p >> 5] >> ((p & 31) * 2)) & 3u;
        const bool isn = (gw[nw + (pm >> 6)] >> (pm & 63)) & 1ull;
        REF[x] = isn ? (uint8_t)'N' : (uint8_t)("ACGT"[c2]);
      }
      wave_sync();
      constexpr int NR = QUAD_MAX_N / 16;
      // homopolymer run of every column (smith_waterman.rs:348-363), as in k_dp_quad
      {
        const uint32_t seg = (n + 15) / 16;
        const uint32_t x0 = min(n, L * seg), cnt = min(n, x0 + seg) - x0;
        uint32_t bb[NR + 1];
        const uint32_t prevb = (x0 > 0 && x0 < n) ? REF[x0 - 1] : 0x100u;
#pragma unroll
        for (int u = 0; u <= NR; u++) {
          const uint32_t x = x0 + u;
          const uint32_t v = REF[n ? min(x, n - 1) : 0];
          bb[u] = x < n ? v : 0x101u;
        }
        int32_t lb = -1, fe = INT32_MAX;
#pragma unroll
        for (int u = NR - 1; u >= 0; u--) {
          if ((uint32_t)u < cnt) {
            if (bb[u] != bb[u + 1]) fe = (int32_t)(x0 + u);
          }
        }
#pragma unroll
        for (int u = 0; u < NR; u++) {
          if ((uint32_t)u < cnt) {
            if (bb[u] != (u == 0 ? prevb : bb[u - 1])) lb = (int32_t)(x0 + u);
          }
        }
        int32_t pm = lb;
        pm = max(pm, dpp_mov<DPP_ROW_SHR1>(-1, pm));
        pm = max(pm, dpp_mov<DPP_ROW_SHR2>(-1, pm));
        pm = max(pm, dpp_mov<DPP_ROW_SHR4>(-1, pm));
        pm = max(pm, dpp_mov<DPP_ROW_SHR8>(-1, pm));
        int32_t start = dpp_mov<DPP_ROW_SHR1>(-1, pm);
        int32_t sm = fe;
        sm = min(sm, dpp_mov<DPP_ROW_SHL1>(INT32_MAX, sm));
        sm = min(sm, dpp_mov<DPP_ROW_SHL2>(INT32_MAX, sm));
        sm = min(sm, dpp_mov<DPP_ROW_SHL4>(INT32_MAX, sm));
        sm = min(sm, dpp_mov<DPP_ROW_SHL8>(INT32_MAX, sm));
        int32_t end = dpp_mov<DPP_ROW_SHL1>(INT32_MAX, sm);
        const int32_t c_lo = 1 + o;
        const int32_t nent = (int32_t)m + 63;
        int32_t st[NR];
#pragma unroll
        for (int u = 0; u < NR; u++) {
          if (bb[u] != (u == 0 ? prevb : bb[u - 1])) start = (int32_t)(x0 + u);
          st[u] = start;
        }
        int16_t* gch = (int16_t*)GC;  // gap-open halves: [4e + hh]
        uint8_t* gcc = (uint8_t*)GC;  // code bytes: [8e + 4 + hh]
#pragma unroll
        for (int u = NR - 1; u >= 0; u--) {
          const int32_t x = (int32_t)x0 + u;
          if ((uint32_t)u < cnt) {
            if (bb[u] != bb[u + 1]) end = x;
            const int32_t e = x + 1 - c_lo;
            if (e >= 0 && e < nent) {
              int32_t run = end - st[u] + 1;
              if (run > 65535) run = 65535;
              const uint32_t rb = bb[u];
              gch[4 * e + hh] = (int16_t)min(a.gap_open + (run - 1), -1);
              gcc[8 * e + 4 + hh] = (uint8_t)(rb == 'A' ? 0 : rb == 'C' ? 6 : rb == 'G' ? 12 : rb == 'T' ? 18 : 24);
            }
          }
        }
      }
      wave_sync();
      // per-row score word (6-bit fields for reference A, C, G, T, N)
      for (uint32_t x = L; x < m; x += 16) {
        const uint8_t rb = Q[x];
        const uint32_t qv = x < qlen ? QQ[x] : 30u;
        const int32_t mis = -(int32_t)min(max(qv, 6u), 30u);
        uint32_t w = 0;
        w |= (uint32_t)((rb == 'A' ? MATCH_SCORE : mis) & 63) << 0;
        w |= (uint32_t)((rb == 'C' ? MATCH_SCORE : mis) & 63) << 6;
        w |= (uint32_t)((rb == 'G' ? MATCH_SCORE : mis) & 63) << 12;
        w |= (uint32_t)((rb == 'T' ? MATCH_SCORE : mis) & 63) << 18;
        w |= (uint32_t)((rb == 'N' ? MATCH_SCORE : mis) & 63) << 24;
        SCW[x] = (int32_t)w;
      }
      wave_sync();
    }
    const int32_t* SCWA = (const int32_t*)(jb + Y.scw);
    const int32_t* SCWB = (const int32_t*)(jb + Y.scw + Y.scw_len);
    // ---- row 0 and per-column constants ----
    uint32_t Mo[4], Io[4], Po[4], kge[4], cst[4];
    const uint32_t cp = (L == 15) ? PK_NEG : 0u;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const int32_t d = 4 * L + k;
      kge[k] = pk_make(d * gev, d * gev);
      cst[k] = pk_make(d <= sv[0] ? -32768 : -d * gev, d <= sv[1] ? -32768 : -d * gev);
      const int32_t j0a = ov[0] + d, j0b = ov[1] + d;
      Mo[k] = pk_make((d >= sv[0] && j0a >= 0 && j0a <= (int32_t)nv[0]) ? 0 : -32768,
                      (d >= sv[1] && j0b >= 0 && j0b <= (int32_t)nv[1]) ? 0 : -32768);
      Io[k] = PK_NEG;
      Po[k] = pk_sub(PK_NEG, kge[k]);
    }
    uint32_t w0 = 0, w1 = 0, w2 = 0, w3 = 0;
    const uint32_t mrow = max(mv[0], mv[1]);
    uint32_t mmax = mrow;
    int32_t emax = max(mv[0] ? bv[0] - hv[0] : 0, mv[1] ? bv[1] - hv[1] : 0);
    for (int gg = 0; gg < 64; gg += 16) {
      mmax = max(mmax, (uint32_t)readlane_i32((int32_t)mrow, gg));
      emax = max(emax, readlane_i32(emax, gg));
    }
    uint32_t* TBA = a.tbq + (((size_t)blockIdx.x * 4 + (size_t)g) * 2) * Y.tb_words;
    uint32_t* TBB = TBA + Y.tb_words;
    int64_t* keys = (int64_t*)(jb + Y.keys);
    keys[L] = end_key(NEG_INF, 2 * nv[0] + 1);
    keys[16 + L] = end_key(NEG_INF, 2 * nv[1] + 1);
    // endpoint of job hh at its last row: last maximum over (j asc; M then I)
    auto endpoint = [&](int hh) {
      int64_t key = end_key(NEG_INF, 2 * nv[hh] + 1);
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const int32_t d = 4 * L + k, j = (int32_t)mv[hh] + ov[hh] + d;
        if (d >= sv[hh] && d < sv[hh] + Wv[hh] && j >= 1 && j <= (int32_t)nv[hh]) {
          const int32_t Mk = hh ? pk_hi(Mo[k]) : pk_lo(Mo[k]), Ik = hh ? pk_hi(Io[k]) : pk_lo(Io[k]);
          const int64_t km = end_key(Mk, (uint32_t)(2 * j)), ki = end_key(Ik, (uint32_t)(2 * j + 1));
          key = key > km ? key : km;
          key = key > ki ? key : ki;
        }
      }
      keys[16 * hh + L] = key;
    };
    struct RowIn {
      int32_t sa, sb;
      uint32_t g0, g1, g2, g3, c0, c1, c2, c3;
    };
    auto load_row = [&](RowIn& x, uint32_t i) {  // in bounds for i <= max_m + 1
      const uint2* gc = GC + (i - 1 + 4 * (uint32_t)L);
      x.sa = SCWA[i - 1];
      x.sb = SCWB[i - 1];
      const uint2 v0 = gc[0], v1 = gc[1], v2 = gc[2], v3 = gc[3];
      x.g0 = v0.x;
      x.g1 = v1.x;
      x.g2 = v2.x;
      x.g3 = v3.x;
      x.c0 = v0.y;
      x.c1 = v1.y;
      x.c2 = v2.y;
      x.c3 = v3.y;
    };
    // scores of both jobs' cells of a column: code A in bits 0..4, code B in 8..12
    auto score = [&](const RowIn& R, uint32_t c) {
      return __builtin_amdgcn_perm((uint32_t)bfe_i32(R.sb, (int32_t)(c >> 8), 6), (uint32_t)bfe_i32(R.sa, (int32_t)c, 6),
                                   0x05040100u);
    };
    auto row = [&](uint32_t i, const RowIn& R, auto edge_tag) {
      constexpr bool EDGE = decltype(edge_tag)::value;
      const uint32_t mu3 = (uint32_t)dpp_mov_z<DPP_ROW_SHL1>((int32_t)Mo[0]);
      const uint32_t iu3 = (uint32_t)dpp_mov_z<DPP_ROW_SHL1>((int32_t)Io[0]);
      uint32_t M0, M1, M2, M3, I0, I1, I2, I3;
      {
        const uint32_t X = pk_max(pk_max(Mo[0], Io[0]), pk_add(Po[0], kge[0]));
        M0 = pk_add(X, score(R, R.c0));
        const uint32_t io = pk_add(Mo[1], R.g0);
        I0 = pk_max(io, pk_add(Io[1], ge_pk));
        w0 = pk_push(w0, pk_sub(Mo[0], X));
        w1 = pk_push(w1, pk_sub(Io[0], X));
        w2 = pk_push(w2, pk_sub(io, I0));
      }
      {
        const uint32_t X = pk_max(pk_max(Mo[1], Io[1]), pk_add(Po[1], kge[1]));
        M1 = pk_add(X, score(R, R.c1));
        const uint32_t io = pk_add(Mo[2], R.g1);
        I1 = pk_max(io, pk_add(Io[2], ge_pk));
        w0 = pk_push(w0, pk_sub(Mo[1], X));
        w1 = pk_push(w1, pk_sub(Io[1], X));
        w2 = pk_push(w2, pk_sub(io, I1));
      }
      {
        const uint32_t X = pk_max(pk_max(Mo[2], Io[2]), pk_add(Po[2], kge[2]));
        M2 = pk_add(X, score(R, R.c2));
        const uint32_t io = pk_add(Mo[3], R.g2);
        I2 = pk_max(io, pk_add(Io[3], ge_pk));
        w0 = pk_push(w0, pk_sub(Mo[2], X));
        w1 = pk_push(w1, pk_sub(Io[2], X));
        w2 = pk_push(w2, pk_sub(io, I2));
      }
      {
        const uint32_t X = pk_max(pk_max(Mo[3], Io[3]), pk_add(Po[3], kge[3]));
        M3 = pk_add(pk_add(X, score(R, R.c3)), cp);
        const uint32_t io = pk_add(pk_add(mu3, R.g3), cp);
        I3 = pk_max(io, pk_add(pk_add(iu3, ge_pk), cp));
        w0 = pk_push(w0, pk_sub(Mo[3], X));
        w1 = pk_push(w1, pk_sub(Io[3], X));
        w2 = pk_push(w2, pk_sub(io, I3));
      }
      const int32_t jla = (int32_t)i + ov[0] + 4 * L, jlb = (int32_t)i + ov[1] + 4 * L;
      uint32_t em[4] = {0, 0, 0, 0};
      if (EDGE) {
#pragma unroll
        for (int k = 0; k < 4; k++) em[k] = (jla + k <= 0 ? 0xFFFFu : 0u) | (jlb + k <= 0 ? 0xFFFF0000u : 0u);
        M0 = (M0 & ~em[0]) | (PK_NEG & em[0]);
        I0 = (I0 & ~em[0]) | (PK_NEG & em[0]);
        M1 = (M1 & ~em[1]) | (PK_NEG & em[1]);
        I1 = (I1 & ~em[1]) | (PK_NEG & em[1]);
        M2 = (M2 & ~em[2]) | (PK_NEG & em[2]);
        I2 = (I2 & ~em[2]) | (PK_NEG & em[2]);
        M3 = (M3 & ~em[3]) | (PK_NEG & em[3]);
        I3 = (I3 & ~em[3]) | (PK_NEG & em[3]);
      }
      // D: max-plus prefix scan along the row
      const uint32_t ml0 = (uint32_t)dpp_mov_z<DPP_ROW_SHR1>((int32_t)M3);
      const uint32_t B0 = pk_add(pk_add(ml0, R.g0), cst[0]), B1 = pk_add(pk_add(M0, R.g1), cst[1]),
                     B2 = pk_add(pk_add(M1, R.g2), cst[2]), B3 = pk_add(pk_add(M2, R.g3), cst[3]);
      const uint32_t L0 = B0, L1 = pk_max(L0, B1), L2 = pk_max(L1, B2), L3 = pk_max(L2, B3);
      uint32_t x = L3;
      x = pk_max(x, (uint32_t)dpp_mov<DPP_ROW_SHR1>((int32_t)PK_NEG, (int32_t)x));
      x = pk_max(x, (uint32_t)dpp_mov<DPP_ROW_SHR2>((int32_t)PK_NEG, (int32_t)x));
      x = pk_max(x, (uint32_t)dpp_mov<DPP_ROW_SHR4>((int32_t)PK_NEG, (int32_t)x));
      x = pk_max(x, (uint32_t)dpp_mov<DPP_ROW_SHR8>((int32_t)PK_NEG, (int32_t)x));
      const uint32_t carry = (uint32_t)dpp_mov<DPP_ROW_SHR1>((int32_t)PK_NEG, (int32_t)x);
      uint32_t P0 = pk_max(L0, carry), P1 = pk_max(L1, carry), P2 = pk_max(L2, carry), P3 = pk_max(L3, carry);
      w3 = pk_push(w3, pk_sub(B0, P0));
      w3 = pk_push(w3, pk_sub(B1, P1));
      w3 = pk_push(w3, pk_sub(B2, P2));
      w3 = pk_push(w3, pk_sub(B3, P3));
      if (EDGE) {
        P0 = (P0 & ~em[0]) | (pk_sub(PK_NEG, kge[0]) & em[0]);
        P1 = (P1 & ~em[1]) | (pk_sub(PK_NEG, kge[1]) & em[1]);
        P2 = (P2 & ~em[2]) | (pk_sub(PK_NEG, kge[2]) & em[2]);
        P3 = (P3 & ~em[3]) | (pk_sub(PK_NEG, kge[3]) & em[3]);
      }
      Mo[0] = M0;
      Mo[1] = M1;
      Mo[2] = M2;
      Mo[3] = M3;
      Io[0] = I0;
      Io[1] = I1;
      Io[2] = I2;
      Io[3] = I3;
      Po[0] = P0;
      Po[1] = P1;
      Po[2] = P2;
      Po[3] = P3;
      if (i == mv[0]) endpoint(0);
      if (i == mv[1]) endpoint(1);
      if ((i & 3u) == 0) {  // 16 pushes: one 4-row block per job
        const uint32_t wd = w0 & w1;
        const uint32_t q = (i - 1) >> 2;
        *(uint2*)(TBA + q * 32 + 2 * L) = make_uint2(__builtin_amdgcn_perm(wd, w0, 0x05040100u), __builtin_amdgcn_perm(w3, w2, 0x05040100u));
        *(uint2*)(TBB + q * 32 + 2 * L) = make_uint2(__builtin_amdgcn_perm(wd, w0, 0x07060302u), __builtin_amdgcn_perm(w3, w2, 0x07060302u));
        w0 = w1 = w2 = w3 = 0;
      }
    };
    RowIn RA, RB;
    load_row(RA, 1);
    uint32_t i = 1;
    const uint32_t elim = (uint32_t)min((int32_t)mmax, emax);
    for (; i + 1 <= elim; i += 2) {
      load_row(RB, i + 1);
      if (i <= mrow) row(i, RA, std::true_type{});
      load_row(RA, i + 2);
      if (i + 1 <= mrow) row(i + 1, RB, std::true_type{});
    }
    if (i <= elim) {
      load_row(RB, i + 1);
      if (i <= mrow) row(i, RA, std::true_type{});
      RA = RB;
      i++;
    }
    for (; i + 1 <= mmax; i += 2) {
      load_row(RB, i + 1);
      if (i <= mrow) row(i, RA, std::false_type{});
      load_row(RA, i + 2);
      if (i + 1 <= mrow) row(i + 1, RB, std::false_type{});
    }
    if (i <= mmax) {
      load_row(RB, i + 1);
      if (i <= mrow) row(i, RA, std::false_type{});
    }
    if (mrow & 3u) {  // the row's last, partial block: left-align both halves' fields
      const uint32_t sh = 4 * (4 - (mrow & 3u));
      const u16x2_t shv = {(unsigned short)sh, (unsigned short)sh};
      const uint32_t a0 = __builtin_bit_cast(uint32_t, __builtin_bit_cast(u16x2_t, w0) << shv);
      const uint32_t a1 = __builtin_bit_cast(uint32_t, __builtin_bit_cast(u16x2_t, w0 & w1) << shv);
      const uint32_t a2 = __builtin_bit_cast(uint32_t, __builtin_bit_cast(u16x2_t, w2) << shv);
      const uint32_t a3 = __builtin_bit_cast(uint32_t, __builtin_bit_cast(u16x2_t, w3) << shv);
      const uint32_t q = (mrow - 1) >> 2;
      *(uint2*)(TBA + q * 32 + 2 * L) = make_uint2(__builtin_amdgcn_perm(a1, a0, 0x05040100u), __builtin_amdgcn_perm(a3, a2, 0x05040100u));
      *(uint2*)(TBB + q * 32 + 2 * L) = make_uint2(__builtin_amdgcn_perm(a1, a0, 0x07060302u), __builtin_amdgcn_perm(a3, a2, 0x07060302u));
    }
    wave_sync();
    mem_sync();  // the traceback words stored by every lane
    // ---- traceback by lane hh of the row, for job hh (smith_waterman.rs:299-345) ----
    uint32_t nruns = 0, off = 0;
    int64_t best = 0;
    bool wvalid = false;
    if (L < 2) {
      const int hh = L;
      const uint32_t m = mv[hh];
      if (m) {
        best = keys[16 * hh];
        for (int u = 1; u < 16; u++) best = best > keys[16 * hh + u] ? best : keys[16 * hh + u];
        wvalid = (int32_t)(best >> 32) >= PAIR16_MIN_SCORE;
        if (!wvalid) dp_flag(a, jidv[hh]);
      }
      if (wvalid) {
        const uint32_t* TB = hh ? TBB : TBA;
        uint32_t* runs_top = (hh ? TBB : TBA) + Y.tb_words;
        const int32_t o = ov[hh];
        uint32_t i = m, j = ((uint32_t)best) >> 1, st = ((uint32_t)best) & 1u;
        uint32_t cur_op = 0xFFu, cur_len = 0;
        int32_t cb = -1, cl = -1;
        uint32_t v0 = 0, v1 = 0, v2 = 0, v3 = 0, p0 = 0, p1 = 0, n2 = 0, l3 = 0;
        // the 32-bit 8-row word of plane (field) f of lane ln and 8-row block rb:
        // 4-row block 2rb in the high half, 2rb + 1 in the low half
        while (i > 0) {
          int32_t d = (int32_t)j - (int32_t)i - o;
          d = d < 0 ? 0 : (d > 63 ? 63 : d);
          const int32_t lq = d >> 2, rb = (int32_t)(i - 1) >> 3;
          if (rb != cb || lq != cl) {
            const uint2 h0 = *(const uint2*)(TB + (2 * rb) * 32 + 2 * lq);
            const uint2 h1 = *(const uint2*)(TB + (2 * rb + 1) * 32 + 2 * lq);
            const int32_t pb = rb > 0 ? rb - 1 : 0;
            const uint32_t ph0 = TB[(2 * pb) * 32 + 2 * lq], ph1 = TB[(2 * pb + 1) * 32 + 2 * lq];
            const int32_t ln2 = lq < 15 ? lq + 1 : lq, ll3 = lq > 0 ? lq - 1 : lq;
            const uint32_t n20 = TB[(2 * rb) * 32 + 2 * ln2 + 1], n21 = TB[(2 * rb + 1) * 32 + 2 * ln2 + 1];
            const uint32_t l30 = TB[(2 * rb) * 32 + 2 * ll3 + 1], l31 = TB[(2 * rb + 1) * 32 + 2 * ll3 + 1];
            v0 = __builtin_amdgcn_perm(h0.x, h1.x, 0x05040100u);  // {h0.x lo16, h1.x lo16}: plane 0
            v1 = __builtin_amdgcn_perm(h0.x, h1.x, 0x07060302u);  // plane 1
            v2 = __builtin_amdgcn_perm(h0.y, h1.y, 0x05040100u);  // plane 2
            v3 = __builtin_amdgcn_perm(h0.y, h1.y, 0x07060302u);  // plane 3
            p0 = __builtin_amdgcn_perm(ph0, ph1, 0x05040100u);
            p1 = __builtin_amdgcn_perm(ph0, ph1, 0x07060302u);
            n2 = __builtin_amdgcn_perm(n20, n21, 0x05040100u);
            l3 = __builtin_amdgcn_perm(l30, l31, 0x07060302u);
            cb = rb;
            cl = lq;
          }
          const uint32_t k = (uint32_t)d & 3u, r = (i - 1) & 7u;
          const uint32_t pos = 31u - (4u * r + k);
          const uint32_t mav = r + 1u + (rb > 0 ? 8u : 0u);
          const uint64_t xm = ((((uint64_t)p0 << 32) | v0) >> pos) & 0x1111111111111111ull &
                              (mav >= 16u ? ~0ull : ((1ull << (4u * mav)) - 1ull));
          const uint32_t f = xm ? ((uint32_t)__builtin_ctzll(xm) >> 2) : 64u;
          const uint32_t mst = min(min(f + 1u, mav), j);
          const uint64_t x1 = (((uint64_t)p1 << 32) | v1) >> pos;
          const uint32_t mpar = (mst == f + 1u) ? 1u + (uint32_t)((x1 >> ((4u * f) & 63u)) & 1ull) : 0u;
          const uint32_t dav = k + 1u + (lq > 0 ? 4u : 0u);
          const uint32_t yd = ((v3 >> pos) & ((2u << k) - 1u)) | (((l3 >> (28u - 4u * r)) & 0xFu) << (k + 1u));
          const uint32_t y = ~yd & ((1u << dav) - 1u);
          const uint32_t fz = y ? (uint32_t)__builtin_ctz(y) : 32u;
          const uint32_t dst = min(min(fz + 1u, dav), j);
          const uint32_t dpar = (dst == fz + 1u) ? 0u : 2u;
          const uint32_t tav = min((lq < 15 ? 7u : 3u) - k, r);
          const uint32_t ma = 0x249249u & ((1u << (3u * (3u - k) + 1u)) - 1u);
          const uint32_t zv = ((v2 >> pos) & ma) | ((r > 0 ? (n2 >> ((pos + 4u) & 31u)) : 0u) & 0x249249u & ~ma);
          const uint32_t zi = ~zv & 0x249249u & ((1u << (3u * tav + 1u)) - 1u);
          const uint32_t fi = zi ? (uint32_t)__builtin_ctz(zi) / 3u : 32u;
          const uint32_t ist = min(fi + 1u, tav + 1u);
          const uint32_t ipar = (ist == fi + 1u) ? 0u : 1u;
          const uint32_t sel = (j == 0) ? 1u : st;
          const uint32_t isM = 0u - (uint32_t)(sel == 0u), isD = 0u - (uint32_t)(sel == 2u);
          const uint32_t isI = ~(isM | isD);
          const uint32_t jz = 0u - (uint32_t)(j == 0);
          const uint32_t len = (mst & isM) | (dst & isD) | (((ist & ~jz) | (1u & jz)) & isI);
          const uint32_t nst = (mpar & isM) | (dpar & isD) | (((ipar & ~jz) | (st & jz)) & isI);
          i -= len & ~isD;
          j -= len & ~isI;
          st = nst;
          const uint32_t fresh = 0u - (uint32_t)(sel != cur_op);
          nruns += (fresh & (uint32_t)(cur_len != 0)) ? 1u : 0u;
          cur_len = (len & fresh) | ((cur_len + len) & ~fresh);
          cur_op = sel;
          runs_top[-1 - (int32_t)nruns] = (cur_len << 4) | cur_op;
        }
        nruns += cur_len != 0 ? 1u : 0u;
        off = j;
      }
    }
    mem_sync();  // the CIGAR runs the walkers stored
    // ---- per job: log-likelihood terms (single.rs:228-262), ordered sum, outputs ----
    double* terms0 = (double*)(jb + Y.gc);  // dead per-row inputs: job A's terms, then job B's
#pragma unroll
    for (int hh = 0; hh < 2; hh++) {
      const uint32_t nr = (uint32_t)__shfl((int)nruns, (lane & ~15) + hh, 64);
      const uint32_t of = (uint32_t)__shfl((int)off, (lane & ~15) + hh, 64);
      const bool ok = __shfl((int)wvalid, (lane & ~15) + hh, 64) != 0;
      const uint32_t rd = a.jobs[jidv[hh]].read;
      const uint32_t m = ok ? mv[hh] : 0u, qlen = (uint32_t)(a.R.qual_off[rd + 1] - a.R.qual_off[rd]);
      const uint32_t* runs = (hh ? TBB : TBA) + Y.tb_words - nr;
      const uint8_t* Q = jb + Y.q + hh * Y.q_len;
      const uint8_t* QQ = jb + Y.qq + hh * Y.q_len;
      const uint8_t* REF = jb + Y.ref + hh * Y.ref_len;
      double* terms = terms0 + (hh ? a.max_m : 0u);
      uint32_t u = 0, qs = 0, rs = 0;
      for (uint32_t qi = L; qi < m; qi += 16) {
        uint32_t op = 1u;
        while (u < nr) {
          const uint32_t w = runs[u], len = w >> 4;
          op = w & 15u;
          const uint32_t qc = op <= 1u ? len : 0u;
          if (qi < qs + qc) break;
          qs += qc;
          rs += (op == 0u || op == 2u) ? len : 0u;
          u++;
        }
        double tv = a.T.ln_quarter;
        if (op == 0u) {
          const uint32_t qv = qi < qlen ? QQ[qi] : 30u;
          tv = (Q[qi] == REF[of + rs + (qi - qs)]) ? a.T.lk_match[qv] : a.T.lk_mis[qv];
        }
        terms[qi] = tv;
      }
    }
    wave_sync();  // the terms of all 16 lanes
    if (L < 2 && wvalid) {
      const int hh = L;
      const uint32_t m = mv[hh];
      const double* terms = terms0 + (hh ? a.max_m : 0u);
      double lk = 0.0;
      uint32_t qi = 0;
      for (; qi + 8 <= m; qi += 8) {  // read order, as the reference sums
        double t8[8];
#pragma unroll
        for (int u = 0; u < 8; u++) t8[u] = terms[qi + u];
#pragma unroll
        for (int u = 0; u < 8; u++) lk += t8[u];
      }
      for (; qi < m; qi++) lk += terms[qi];
      const uint32_t* runs = (hh ? TBB : TBA) + Y.tb_words - nruns;
      uint32_t span = 0;
      for (uint32_t u = 0; u < nruns; u++) {
        const uint32_t w = runs[u], op = w & 15u;
        if (op == 0u || op == 2u) span += w >> 4;
      }
      const DpJob jh = a.jobs[jidv[hh]];
      const uint64_t cells = quad_cells(m, jh.hint, jh.band, nv[hh]);
      c_cells += cells;
      DpOut oo;
      oo.log_lk = lk;
      oo.ref_off = off;
      oo.score = (int32_t)(best >> 32);
      oo.cigar_n = nruns;
      oo.span = span;
      oo.cells = cells;
      a.out[jidv[hh]] = oo;
    }
    // ---- CIGAR runs to each job's slot, forward order ----
#pragma unroll
    for (int hh = 0; hh < 2; hh++) {
      const uint32_t nr = (uint32_t)__shfl((int)nruns, (lane & ~15) + hh, 64);
      const bool ok = __shfl((int)wvalid, (lane & ~15) + hh, 64) != 0;
      if (ok) {
        const uint32_t* runs = (hh ? TBB : TBA) + Y.tb_words - nr;
        const uint32_t ncopy = nr < a.cigar_cap ? nr : a.cigar_cap;
        for (uint32_t u = L; u < ncopy; u += 16) a.cigar[((uint64_t)jidv[hh] - a.cigar_base) * a.cigar_cap + u] = runs[u];
      }
    }
    wave_sync();
  }
#pragma unroll
  for (int sft = 32; sft > 0; sft >>= 1) c_cells += (unsigned long long)__shfl_xor((long long)c_cells, sft, 64);
  if (lane == 0 && a.ctr && c_cells) atomicAdd(&a.ctr->cells, c_cells);
}

// Launches k_dp_quad; returns false (nothing launched) when one wave's four
// job areas do not fit in 64 KiB of LDS.
__global__ void k_job_zero(uint32_t* p, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}
void launch_zero_u32(uint32_t* p, uint32_t n, hipStream_t s) {
  if (n) hipLaunchKernelGGL(k_job_zero, dim3((n + 255) / 256), dim3(256), 0, s, p, n);
}

// ---------------------------------------------------------------------------
// Counting sort of the DP jobs by read length, longest first: the four jobs
// that share a wave of k_dp_quad then run the same number of rows, and slots
// without a job (past the read's job count) drop out.  Results do not depend
// on the order (each job writes only its own slot); inside one length the
// order is whatever the atomics give.
// ---------------------------------------------------------------------------
constexpr uint32_t SORT_KEYS = MAX_READ_LEN + 1, SORT_BLOCK = 1024;

__device__ __forceinline__ bool sort_job(const DpArgs& a, uint32_t t, uint32_t& jid, uint32_t& key) {
  if (t >= a.njobs) return false;
  jid = a.jlist ? a.jlist[t] : t;
  if (a.job_valid && (jid % TOP_CANDIDATES) >= a.job_valid[jid / TOP_CANDIDATES]) return false;
  const uint32_t r = a.jobs[jid].read;
  const uint32_t m = (uint32_t)(a.R.seq_off[r + 1] - a.R.seq_off[r]);
  key = (uint32_t)MAX_READ_LEN - min(m, (uint32_t)MAX_READ_LEN);  // 0 = longest
  return true;
}

__global__ __launch_bounds__(SORT_BLOCK) void k_job_hist(DpArgs a, uint32_t* hist) {
  __shared__ uint32_t h[SORT_KEYS];
  for (uint32_t i = threadIdx.x; i < SORT_KEYS; i += SORT_BLOCK) h[i] = 0;
  __syncthreads();
  uint32_t jid, key;
  if (sort_job(a, blockIdx.x * SORT_BLOCK + threadIdx.x, jid, key)) atomicAdd(&h[key], 1u);
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < SORT_KEYS; i += SORT_BLOCK)
    if (h[i]) atomicAdd(&hist[i], h[i]);
}

// Exclusive scan of the key counts in place; the job count to total[0].
__global__ __launch_bounds__(SORT_BLOCK) void k_job_scan(uint32_t* hist, uint32_t* total) {
  __shared__ uint32_t sh[SORT_BLOCK / 64];
  uint32_t carry = 0;
  for (uint32_t k0 = 0; k0 < SORT_KEYS; k0 += SORT_BLOCK) {
    const uint32_t k = k0 + threadIdx.x;
    uint32_t tot;
    const uint32_t ex = team_excl_scan<SORT_BLOCK>(k < SORT_KEYS ? hist[k] : 0u, tot, sh);
    if (k < SORT_KEYS) hist[k] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0) total[0] = carry;
}

// Each block reserves one range per key with one global atomic, then places its
// jobs inside it by their LDS rank.
__global__ __launch_bounds__(SORT_BLOCK) void k_job_scatter(DpArgs a, uint32_t* off, uint32_t* out) {
  __shared__ uint32_t h[SORT_KEYS];
  for (uint32_t i = threadIdx.x; i < SORT_KEYS; i += SORT_BLOCK) h[i] = 0;
  __syncthreads();
  uint32_t jid = 0, key = 0, rank = 0;
  const bool has = sort_job(a, blockIdx.x * SORT_BLOCK + threadIdx.x, jid, key);
  if (has) rank = atomicAdd(&h[key], 1u);
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < SORT_KEYS; i += SORT_BLOCK)
    if (h[i]) h[i] = atomicAdd(&off[i], h[i]);
  __syncthreads();
  if (has) out[h[key] + rank] = jid;
}

void launch_job_sort(const DpArgs& a, uint32_t* hist, uint32_t* out, uint32_t* total, hipStream_t s) {
  const uint32_t blocks = (a.njobs + SORT_BLOCK - 1) / SORT_BLOCK;
  hipLaunchKernelGGL(k_job_zero, dim3((SORT_KEYS + 255) / 256), dim3(256), 0, s, hist, SORT_KEYS);
  if (blocks) hipLaunchKernelGGL(k_job_hist, dim3(blocks), dim3(SORT_BLOCK), 0, s, a, hist);
  hipLaunchKernelGGL(k_job_scan, dim3(1), dim3(SORT_BLOCK), 0, s, hist, total);
  if (blocks) hipLaunchKernelGGL(k_job_scatter, dim3(blocks), dim3(SORT_BLOCK), 0, s, a, hist, out);
}

uint32_t quad_blocks(uint32_t njobs, uint32_t grid_cap) {
  const uint32_t cap = grid_cap ? grid_cap : QUAD_MAX_WAVES;
  const uint32_t q = (njobs + 3) / 4;
  return q < cap ? q : cap;
}
static uint32_t pair16_blocks(uint32_t njobs, uint32_t grid_cap) {
  const uint32_t cap = grid_cap ? grid_cap : QUAD_MAX_WAVES;
  const uint32_t q = (njobs + 7) / 8;
  return q < cap ? q : cap;
}
size_t quad_scratch_words(uint32_t max_m, uint32_t max_n, uint32_t njobs, uint32_t grid_cap) {
  const size_t quad = (size_t)((njobs + 3) & ~3u) * quad_layout(max_m, max_n).tb_words;
  const size_t pair = (size_t)pair16_blocks(njobs, grid_cap) * 8 * pair16_layout(max_m, max_n).tb_words;
  return quad > pair ? quad : pair;
}
int launch_dp_quad(const DpArgs& a, hipStream_t s) {
  if (a.njobs == 0) return 2;
  // eight jobs per wave in 16-bit halves while the gap penalties keep the D
  // scan's column shifts small (see k_dp_pair16)
  if (a.pack16 && a.max_m <= (uint32_t)QUAD_MAX_M && a.gap_extend >= -4 && a.gap_open >= -400 && a.tbq && a.jlist) {
    const Pair16Layout P = pair16_layout(a.max_m, a.max_n);
    const size_t lds = 4 * (size_t)P.stride;
    const size_t need = 8 * ((size_t)((a.max_n + 62) >> 5) + ((a.max_n + 126) >> 6));
    if (lds <= 64 * 1024 && need <= P.scw_len) {
      hipLaunchKernelGGL(k_dp_pair16, dim3(pair16_blocks(a.njobs, a.quad_grid)), dim3(64), lds, s, a);
      return 2;
    }
  }
  const QuadLayout Y = quad_layout(a.max_m, a.max_n);
  const size_t lds = 4 * (size_t)Y.stride;
  if (a.max_m > (uint32_t)QUAD_MAX_M || lds > 64 * 1024 || !a.tbq || !a.tbend) return 0;
  // genome-word staging reuses the score-word area
  const size_t need = 8 * ((size_t)((a.max_n + 62) >> 5) + ((a.max_n + 126) >> 6));
  if (need > Y.q - Y.scw) return 0;
  hipLaunchKernelGGL(k_dp_quad, dim3(quad_blocks(a.njobs, a.quad_grid)), dim3(64), lds, s, a);
  return 1;
}
void launch_dp_walk(const DpArgs& a, hipStream_t s) {
  if (a.njobs == 0) return;
  hipLaunchKernelGGL(k_dp_walk, dim3((a.njobs + 255) / 256), dim3(256), 0, s, a);
  hipLaunchKernelGGL(k_dp_lk, dim3((a.njobs + 15) / 16), dim3(256), 0, s, a);
}

// Any band width: the band of row i, [js, je], is processed in 64-column
// tiles; the previous row lives in per-wave global scratch (M, I, D arrays
// over columns 0..n); traceback bits per (row, tile) in global scratch.
__host__ __device__ inline uint32_t dp_tiled_lds_bytes(uint32_t max_m, uint32_t max_n) {
  uint32_t b = 0;
  b += (max_m + 15) & ~15u;
  b += (max_m + 15) & ~15u;
  b += (max_n + 15) & ~15u;
  b += ((max_n * 2) + 15) & ~15u;
  b += ((max_m + max_n + 2) * 4 + 15) & ~15u;
  return b;
}

__global__ __launch_bounds__(64) void k_dp_tiled(DpArgs a) {
  extern __shared__ __align__(16) uint8_t dyn[];
  const int lane = lane_id();
  // staging in LDS, or in this wave's slice of device memory when it does not fit
  uint8_t* sbase = a.gstage ? a.gstage + (uint64_t)blockIdx.x * a.gstage_stride : dyn;
  DpStage st;
  st.q = sbase;
  st.qq = st.q + ((a.max_m + 15) & ~15u);
  st.ref = st.qq + ((a.max_m + 15) & ~15u);
  st.go = (int16_t*)(st.ref + ((a.max_n + 15) & ~15u));
  st.runs = (uint32_t*)((uint8_t*)st.go + (((a.max_n * 2) + 15) & ~15u));
  // the previous row's M, I, D: after the staging arrays when the launch gave
  // it room there (a.rows null), else in this wave's global scratch
  int32_t* RM = a.rows ? a.rows + (uint64_t)blockIdx.x * 3 * (a.max_n + 1)
                       : (int32_t*)(sbase + dp_tiled_lds_bytes(a.max_m, a.max_n));
  int32_t* RI = RM + (a.max_n + 1);
  int32_t* RD = RI + (a.max_n + 1);
  uint64_t* TB = a.tb + (uint64_t)blockIdx.x * (uint64_t)(a.max_m + 1) * a.tiles * 4;
  unsigned long long c_cells = 0;
  const uint32_t nj = a.njobs_dev ? *a.njobs_dev : a.njobs;
  for (uint32_t t = blockIdx.x; t < nj; t += gridDim.x) {
    const uint32_t jid = a.jlist ? a.jlist[t] : t;
    if (a.job_valid) {
      if ((jid % TOP_CANDIDATES) >= a.job_valid[jid / TOP_CANDIDATES]) continue;
    }
    const DpJob job = a.jobs[jid];
    const uint64_t so = a.R.seq_off[job.read], qo = a.R.qual_off[job.read];
    const uint32_t m = (uint32_t)(a.R.seq_off[job.read + 1] - so);
    const uint32_t qlen = (uint32_t)(a.R.qual_off[job.read + 1] - qo);
    const uint32_t n = job.n;
    const int64_t b = job.band, h = job.hint;
    if (m == 0 || n == 0) {
      if (lane == 0) st.runs[0] = (m << 4) | 4u;
      wave_sync();
      dp_emit(a, jid, st, 1, 0, 0, 0, 0.0, 0);
      continue;
    }
    if (m > a.max_m || n > a.max_n) {
      if (lane == 0) dp_flag(a, jid);  // the host reports PBW_ERR_LIMIT
      continue;
    }
    dp_stage_job(a, job, m, qlen, so, qo, st);
    // row 0: M = 0 on [0, n], I = D = NEG_INF; "previous band" = [0, n]
    for (uint32_t x = lane; x <= n; x += 64) {
      RM[x] = 0;
      RI[x] = NEG_INF;
      RD[x] = NEG_INF;
    }
    __threadfence_block();
    wave_sync();
    int64_t pjs = 0, pje = n;
    int64_t last_row = 0;
    uint64_t cells = 0;
    const int32_t ge = a.gap_extend;
    for (int64_t i = 1; i <= (int64_t)m; i++) {
      const int64_t center = i + h;
      const int64_t js = max64(center - b, (int64_t)1);
      const int64_t je = max64(min64(center + b, (int64_t)n), (int64_t)0);
      if (js > je) continue;  // skipped row: no reset, no swap
      const uint8_t rb = st.q[i - 1];
      const uint32_t qv = (uint64_t)(i - 1) < qlen ? st.qq[i - 1] : 30u;
      const int32_t mis = a.T.mis_pen[qv];
      // carry into the first tile: column js-1 of the previous row (not yet
      // overwritten) and of the current row (outside the band: NEG_INF)
      int32_t cM, cI, cD;
      {
        int64_t c = js - 1;
        bool in = c >= pjs && c <= pje;
        cM = in ? RM[c] : NEG_INF;
        cI = in ? RI[c] : NEG_INF;
        cD = in ? RD[c] : NEG_INF;
      }
      int32_t curMl = NEG_INF, curDl = NEG_INF;  // current row, column tile_start-1
      const uint32_t ntile = (uint32_t)((je - js) / 64 + 1);
      for (uint32_t tt = 0; tt < ntile; tt++) {
        const int64_t j = js + (int64_t)tt * 64 + lane;
        const bool valid = j <= je;
        int32_t uM = NEG_INF, uI = NEG_INF, uD = NEG_INF;
        if (valid && j >= pjs && j <= pje) {
          uM = RM[j];
          uI = RI[j];
          uD = RD[j];
        }
        const int32_t dM = dpp_wave_shr1(cM, uM), dI = dpp_wave_shr1(cI, uI), dD = dpp_wave_shr1(cD, uD);
        cM = readlane_i32(uM, 63);
        cI = readlane_i32(uI, 63);
        cD = readlane_i32(uD, 63);
        uint8_t g = 0;
        int32_t go = 0;
        if (valid) {
          g = st.ref[j - 1];
          go = st.go[j - 1];
        }
        const int32_t sc = (rb == g) ? MATCH_SCORE : mis;
        const int32_t fm = dM + sc, fi = dI + sc, fd = dD + sc;
        int32_t M = max(max(fm, fi), fd);
        const int32_t io = uM + go, ie = uI + ge;
        const int32_t I = max(io, ie);
        if (!valid) M = NEG_INF;
        const int32_t Ml = dpp_wave_shr1(curMl, M);
        const int32_t A = Ml + go;
        const int32_t Sent = curDl + ge;  // D(tile_start - 1) + ge
        int32_t B = valid ? A - lane * ge : INT32_MIN;
        B = wave_incl_max(B);
        const int32_t D = lane * ge + max(B, Sent);
        const uint64_t ne_fm = __ballot(valid && M != fm);
        const uint64_t eq_fi = __ballot(valid && M == fi);
        const uint64_t ti = __ballot(valid && I != io);
        const uint64_t td = __ballot(valid && D != A);
        cells += (uint64_t)__popcll(__ballot(valid));
        if (lane == 0) {
          uint64_t* row = TB + ((uint64_t)i * a.tiles + tt) * 4;
          row[0] = ne_fm & eq_fi;
          row[1] = ne_fm & ~eq_fi;
          row[2] = ti;
          row[3] = td;
        }
        // last valid lane's current-row values feed the next tile
        const int lastl = (int)min64((int64_t)63, je - (js + (int64_t)tt * 64));
        curMl = readlane_i32(M, lastl);
        curDl = readlane_i32(D, lastl);
        if (valid) {
          RM[j] = M;
          RI[j] = I;
          RD[j] = D;
        }
      }
      __threadfence_block();
      wave_sync();
      pjs = js;
      pje = je;
      last_row = i;
    }
    // endpoint over the last computed row (or row 0)
    uint32_t best_j, best_state;
    int32_t best_score;
    {
      int64_t key = INT64_MIN;
      for (int64_t j = pjs + lane; j <= pje; j += 64) {
        int64_t km = end_key(RM[j], (uint32_t)(2 * j));
        int64_t ki = end_key(RI[j], (uint32_t)(2 * j + 1));
        int64_t k2 = km > ki ? km : ki;
        key = key > k2 ? key : k2;
      }
      key = wave_max_i64(key);
      if (last_row > 0) {
        const int64_t joob = pje < (int64_t)n ? (int64_t)n : pjs - 1;
        const int64_t koob = end_key(NEG_INF, (uint32_t)(2 * joob + 1));
        key = key > koob ? key : koob;
      }
      best_score = (int32_t)(key >> 32);
      best_j = ((uint32_t)key) >> 1;
      best_state = ((uint32_t)key) & 1u;
    }
    uint32_t nruns = 0, off = 0, span = 0;
    double lk = 0.0;
    if (lane == 0) {
      const int64_t i0 = max64((int64_t)1, 1 - h - b);
      const int64_t i1 = min64((int64_t)m, (int64_t)n + b - h);
      auto tb = [&](uint32_t ii, uint32_t jj) -> uint32_t {
        if ((int64_t)ii < i0 || (int64_t)ii > i1) return 0u;
        const int64_t js = max64((int64_t)ii + h - b, (int64_t)1);
        const int64_t je = max64(min64((int64_t)ii + h + b, (int64_t)n), (int64_t)0);
        if ((int64_t)jj < js || (int64_t)jj > je) return 0u;
        const uint64_t rel = (uint64_t)((int64_t)jj - js);
        const uint64_t* row = TB + ((uint64_t)ii * a.tiles + rel / 64) * 4;
        const uint32_t d = (uint32_t)(rel % 64);
        return (uint32_t)((row[0] >> d) & 1ull) | ((uint32_t)((row[1] >> d) & 1ull) << 1) |
               ((uint32_t)((row[2] >> d) & 1ull) << 2) | ((uint32_t)((row[3] >> d) & 1ull) << 3);
      };
      dp_traceback(m, best_j, best_state, tb, st, nruns, off);
      for (uint32_t t2 = 0; t2 < nruns; t2++) {
        uint32_t op = st.runs[t2] & 15u;
        if (op == 0 || op == 2) span += st.runs[t2] >> 4;
      }
      lk = dp_log_lk(a.T, st, m, qlen, nruns, off, span);
    }
    nruns = shfl_u32(nruns, 0);
    off = shfl_u32(off, 0);
    span = shfl_u32(span, 0);
    mem_sync();  // lane 0's runs to every lane (LDS or device-memory staging)
    dp_emit(a, jid, st, nruns, off, span, best_score, lk, cells);
    c_cells += cells;
    wave_sync();
  }
  if (lane == 0 && a.ctr) atomicAdd(&a.ctr->cells, c_cells);
}

bool dp_tiled_rows_in_lds(uint32_t max_m, uint32_t max_n) {
  return dp_tiled_lds_bytes(max_m, max_n) + 12ull * (max_n + 1) <= 48 * 1024;
}
uint64_t dp_tiled_gstage_bytes(uint32_t max_m, uint32_t max_n) {
  const uint64_t st = dp_tiled_lds_bytes(max_m, max_n);
  return st <= 64 * 1024 ? 0 : (st + 12ull * (max_n + 1) + 255) & ~255ull;
}
void launch_dp_tiled(const DpArgs& a, uint32_t nwaves, hipStream_t s) {
  if (a.njobs == 0) return;
  uint32_t blocks = nwaves < a.njobs ? nwaves : a.njobs;
  size_t lds = 0;
  if (a.gstage) {  // staging and the previous row in device memory
    if (blocks > a.gstage_waves) blocks = a.gstage_waves;
  } else {
    lds = dp_tiled_lds_bytes(a.max_m, a.max_n) + (a.rows ? 0 : 12ull * (a.max_n + 1));
  }
  hipLaunchKernelGGL(k_dp_tiled, dim3(blocks), dim3(64), lds, s, a);
}

// ---------------------------------------------------------------------------
// Wide windows whose band covers every row entirely (the align_at_window mate
// windows of paired-end: js = 1 and je = n on every row, single.rs:220-258):
// k_dp_tiled's recurrence with four columns per lane and 256-column tiles, one
// wave per tile (a block of T waves per job, n <= 256 T), the previous row's
// M, I and D in registers.
//  - A step is two rows; wave w runs one step behind wave w - 1 (step s: rows
//    2 (s - w) - 1 and 2 (s - w)) and takes the previous tile's last column of
//    each row (M, I, D) from LDS: M and D as the row's left neighbours, and,
//    for the next row, all three as the diagonal.  Four LDS slots per tile
//    (row mod 4) and a block barrier per step keep writer and reader apart.
//  - Up neighbour: the same register; diagonal: the lane's previous column or,
//    for its first column, the left lane's last one by a wave_shr DPP move.
//  - D: lane-local max-plus prefix over four columns, then an inclusive max
//    scan over the wave of (A - position * ge), exactly as k_dp_tiled's 64
//    lanes but one scan per 256 columns.
//  - Traceback bits as k_dp_tiled's (M from I, M from D, I extended, D
//    extended), 16 bits per lane, tile and row in global scratch; wave 0 walks
//    them from LDS-staged row blocks.
// Jobs of another shape, or wider than 256 T columns, are flagged to
// k_dp_tiled.
// ---------------------------------------------------------------------------
// traceback rows staged in LDS at a time: 16, or 8 for blocks of more than four tiles
__host__ __device__ constexpr uint32_t wide_tb_rows(uint32_t T) { return T <= 4 ? 16u : 8u; }
__host__ __device__ inline uint32_t dp_wide_lds_bytes(uint32_t max_m, uint32_t max_n, uint32_t T) {
  return dp_tiled_lds_bytes(max_m, max_n) + wide_tb_rows(T) * T * 128 + 56 * T;
}
template <int T>
__global__ __launch_bounds__(64 * T) __attribute__((amdgpu_waves_per_eu(4))) void k_dp_wide(DpArgs a) {
  extern __shared__ __align__(16) uint8_t dyn[];
  const int lane = lane_id();
  const uint32_t w = threadIdx.x >> 6;  // this wave's tile
  DpStage st;
  st.q = dyn;
  st.qq = st.q + ((a.max_m + 15) & ~15u);
  st.ref = st.qq + ((a.max_m + 15) & ~15u);
  st.go = (int16_t*)(st.ref + ((a.max_n + 15) & ~15u));
  st.runs = (uint32_t*)((uint8_t*)st.go + (((a.max_n * 2) + 15) & ~15u));
  uint16_t* TB = (uint16_t*)a.tb + (uint64_t)blockIdx.x * (uint64_t)(a.max_m + 1) * T * 64;
  constexpr uint32_t WIDE_TB_ROWS = wide_tb_rows(T);
  uint32_t* tbs = (uint32_t*)(dyn + dp_tiled_lds_bytes(a.max_m, a.max_n));  // traceback staging
  int64_t* ekey = (int64_t*)(tbs + WIDE_TB_ROWS * T * 32);                  // per tile: endpoint key
  int32_t* carry = (int32_t*)(ekey + T);                                    // [tile][row % 4][M, I, D]
  const int32_t ge = a.gap_extend;
  unsigned long long c_cells = 0;
  const uint32_t nj = a.njobs_dev ? *a.njobs_dev : a.njobs;
  for (uint32_t t = blockIdx.x; t < nj; t += gridDim.x) {
    const uint32_t jid = a.jlist ? a.jlist[t] : t;
    const DpJob job = a.jobs[jid];
    const uint64_t so = a.R.seq_off[job.read], qo = a.R.qual_off[job.read];
    const uint32_t m = (uint32_t)(a.R.seq_off[job.read + 1] - so);
    const uint32_t qlen = (uint32_t)(a.R.qual_off[job.read + 1] - qo);
    const uint32_t n = job.n;
    const int64_t b = job.band, h = job.hint;
    const bool rect = m >= 1 && n >= 1 && m <= a.max_m && n <= a.max_n && n <= (uint32_t)(256 * T) &&
                      (int64_t)m + h - b <= 1 && 1 + h + b >= (int64_t)n;
    if (!rect) {
      if (threadIdx.x == 0) dp_flag(a, jid);
      continue;
    }
    // staging, then each row's mismatch penalty (st.runs is free until the
    // traceback): no table load on the row loop's path
    int32_t* rmis = (int32_t*)st.runs;
    if (w == 0) {
      dp_stage_job(a, job, m, qlen, so, qo, st);
      for (uint32_t x = (uint32_t)lane; x < m; x += 64) rmis[x] = a.T.mis_pen[x < qlen ? st.qq[x] : 30u];
    }
    __syncthreads();
    // per-column constants: reference byte and gap open of column c = 256 w + 4 lane + k + 1
    uint32_t rbyte[4];
    int32_t gop[4], gk[4];  // gap open, and gap open - position * ge
    int32_t PM[4], PI[4], PD[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const uint32_t c = 256u * w + 4u * lane + k + 1u;
      const bool v = c <= n;
      rbyte[k] = v ? st.ref[c - 1] : 0x100u;
      gop[k] = v ? (int32_t)st.go[c - 1] : -1;
      gk[k] = gop[k] - (int32_t)(4 * lane + k) * ge;
      PM[k] = 0;  // row 0: M = 0 on [0, n], I = D = NEG_INF
      PI[k] = NEG_INF;
      PD[k] = NEG_INF;
    }
    // previous row at column 256 w (row 0 to start with)
    int32_t cM = 0, cI = NEG_INF, cD = NEG_INF;
    // Row i reads the previous row from one register set and writes the other:
    // odd rows read P* and write N*, even rows the reverse, so no copies.
    int32_t NM[4], NI[4], ND[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
      NM[k] = PM[k];
      NI[k] = PI[k];
      ND[k] = PD[k];
    }
    auto row = [&](int32_t i, const int32_t (&PM)[4], const int32_t (&PI)[4], const int32_t (&PD)[4],
                   int32_t (&NM)[4], int32_t (&NI)[4], int32_t (&ND)[4]) {
      if (i >= 1 && i <= (int32_t)m) {
        const uint32_t rb = st.q[i - 1];
        const int32_t mis = rmis[i - 1];
        int32_t dgM, dgI, dgD, lM, lD;  // column 256 w: previous row, current row
        if (w == 0) {  // column 0 is in the band only in row 0
          dgM = i == 1 ? 0 : NEG_INF;
          dgI = dgD = NEG_INF;
          lM = lD = NEG_INF;
        } else {
          const int32_t* cr = carry + ((w - 1) * 4 + ((uint32_t)i & 3u)) * 3;
          dgM = cM;
          dgI = cI;
          dgD = cD;
          cM = cr[0];
          cI = cr[1];
          cD = cr[2];
          lM = cM;
          lD = cD;
        }
        const int32_t dM0 = dpp_wave_shr1(dgM, PM[3]);
        const int32_t dI0 = dpp_wave_shr1(dgI, PI[3]);
        const int32_t dD0 = dpp_wave_shr1(dgD, PD[3]);
        int32_t M[4], I[4];
        uint32_t bits = 0;
        // Columns past n only feed columns past n (every dependency runs
        // rightwards or down), so they are computed like the others and never
        // read: no masks.  With one score added to all three sources, the
        // maximum and the traceback choice (M, then I, then D on ties) follow
        // from the sources alone.
#pragma unroll
        for (int k = 0; k < 4; k++) {
          const int32_t dM = k ? PM[k - 1] : dM0, dI = k ? PI[k - 1] : dI0, dD = k ? PD[k - 1] : dD0;
          const int32_t sc = (rb == rbyte[k]) ? MATCH_SCORE : mis;
          const int32_t xid = max(dI, dD);
          const int32_t Mk = max(dM, xid) + sc;
          const int32_t io = PM[k] + gop[k], ie = PI[k] + ge;
          const int32_t Ik = max(io, ie);
          M[k] = Mk;
          I[k] = Ik;
          const uint32_t tm = dM >= xid ? 0u : (dI >= dD ? 1u : 2u);
          bits |= (tm | (ie > io ? 4u : 0u)) << (4 * k);
        }
        // D: A(c) = M(c - 1) + go(c); D(c) = max(A(c), D(c - 1) + ge), in key
        // space (key = A - position * ge, so key = M(c - 1) + gk(c)): the prefix
        // maximum there, and D differs from A exactly where it differs from key
        const int32_t Ml0 = dpp_wave_shr1(lM, M[3]);
        int32_t key[4];
#pragma unroll
        for (int k = 0; k < 4; k++) key[k] = (k ? M[k - 1] : Ml0) + gk[k];
        const int32_t l1 = max(key[0], key[1]), l2 = max(l1, key[2]), l3 = max(l2, key[3]);
        const int32_t incl = wave_incl_max(l3);
        const int32_t ex = dpp_wave_shr1(INT32_MIN, incl);  // lanes left of this one
        const int32_t base = max(ex, lD + ge);
        const int32_t dk[4] = {max(key[0], base), max(l1, base), max(l2, base), max(l3, base)};
        int32_t D[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
          D[k] = (4 * lane + k) * ge + dk[k];
          bits |= (dk[k] != key[k] ? 8u : 0u) << (4 * k);
        }
        TB[((uint64_t)i * T + w) * 64 + lane] = (uint16_t)bits;
#pragma unroll
        for (int k = 0; k < 4; k++) {
          NM[k] = M[k];
          NI[k] = I[k];
          ND[k] = D[k];
        }
        if (w + 1 < (uint32_t)T && lane == 63) {  // the tile's last column to wave w + 1
          int32_t* cw = carry + (w * 4 + ((uint32_t)i & 3u)) * 3;
          cw[0] = M[3];
          cw[1] = I[3];
          cw[2] = D[3];
        }
      }
    };
    // step s: rows 2 g - 1 and 2 g of group g = s - w (the first from P* into
    // N*, the second back), then the block barrier
    const uint32_t ng = (m + 1) / 2;
    for (uint32_t s = 1; s < ng + T; s++) {
      const int32_t g = (int32_t)s - (int32_t)w;
      row(2 * g - 1, PM, PI, PD, NM, NI, ND);
      row(2 * g, NM, NI, ND, PM, PI, PD);
      __syncthreads();
    }
    if (m & 1u) {  // row m was the first of its group: in N*
#pragma unroll
      for (int k = 0; k < 4; k++) {
        PM[k] = NM[k];
        PI[k] = NI[k];
      }
    }
    // endpoint over row m: last maximum over (j asc; M then I)
    int64_t key = INT64_MIN;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const uint32_t c = 256u * w + 4u * lane + k + 1u;
      if (c <= n) {
        const int64_t km = end_key(PM[k], 2u * c), ki = end_key(PI[k], 2u * c + 1u);
        const int64_t k2 = km > ki ? km : ki;
        key = key > k2 ? key : k2;
      }
    }
    key = wave_max_i64(key);
    if (lane == 0) ekey[w] = key;
    __syncthreads();  // also orders every wave's traceback words before wave 0 reads them
    if (w == 0) {
      key = end_key(NEG_INF, 1u);  // column 0 of the last row lies outside its band
#pragma unroll
      for (int tt = 0; tt < T; tt++) key = key > ekey[tt] ? key : ekey[tt];
      const int32_t best_score = (int32_t)(key >> 32);
      const uint32_t best_j = ((uint32_t)key) >> 1, best_state = ((uint32_t)key) & 1u;
      // The wave walks a run per round (dp_traceback_wave), each lane reading
      // one cell's bits from the planes in global scratch.
      uint32_t nruns = 0, off = 0, span = 0;
      auto tb = [&](uint32_t ii, uint32_t jj) -> uint32_t {
        const uint32_t rel = jj - 1, tt = rel >> 8, ln = (rel >> 2) & 63u, k = rel & 3u;
        const uint16_t hw = TB[((uint64_t)ii * T + tt) * 64 + ln];
        return ((uint32_t)hw >> (4 * k)) & 15u;
      };
      dp_traceback_wave(m, best_j, best_state, tb, st, nruns, off);
      wave_sync();  // the runs lane 0 stored
      for (uint32_t t2 = 0; t2 < nruns; t2++) {
        const uint32_t op = st.runs[t2] & 15u;
        if (op == 0 || op == 2) span += st.runs[t2] >> 4;
      }
      const double lk = dp_log_lk_wave(a.T, st, m, qlen, nruns, off, span, reinterpret_cast<double*>(tbs));
      wave_sync();
      const uint64_t cells = (uint64_t)m * n;
      dp_emit(a, jid, st, nruns, off, span, best_score, lk, cells);
      c_cells += cells;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0 && a.ctr) atomicAdd(&a.ctr->cells, c_cells);
}

// Tiles of k_dp_wide per job for windows of up to max_n columns: the least
// instantiated count that covers them (0: too wide for it, or its staging for
// (max_m, max_n) beyond 64 KiB of LDS).
uint32_t dp_wide_tiles(uint32_t max_m, uint32_t max_n) {
  const uint32_t need = (max_n + 255) / 256;
  static const uint32_t inst[] = {1, 2, 3, 4, 6, 8, 12, 16};
  for (uint32_t T : inst)
    if (T >= need) return dp_wide_lds_bytes(max_m, max_n, T) <= 64 * 1024 ? T : 0u;
  return 0;
}
void launch_dp_wide(const DpArgs& a, uint32_t nwaves, hipStream_t s) {
  if (a.njobs == 0) return;
  const uint32_t blocks = nwaves < a.njobs ? nwaves : a.njobs;
  const uint32_t T = dp_wide_tiles(a.max_m, a.max_n);
  const size_t lds = dp_wide_lds_bytes(a.max_m, a.max_n, T);
  switch (T) {
    case 1: hipLaunchKernelGGL(k_dp_wide<1>, dim3(blocks), dim3(64), lds, s, a); break;
    case 2: hipLaunchKernelGGL(k_dp_wide<2>, dim3(blocks), dim3(128), lds, s, a); break;
    case 3: hipLaunchKernelGGL(k_dp_wide<3>, dim3(blocks), dim3(192), lds, s, a); break;
    case 4: hipLaunchKernelGGL(k_dp_wide<4>, dim3(blocks), dim3(256), lds, s, a); break;
    case 6: hipLaunchKernelGGL(k_dp_wide<6>, dim3(blocks), dim3(384), lds, s, a); break;
    case 8: hipLaunchKernelGGL(k_dp_wide<8>, dim3(blocks), dim3(512), lds, s, a); break;
    case 12: hipLaunchKernelGGL(k_dp_wide<12>, dim3(blocks), dim3(768), lds, s, a); break;
    case 16: hipLaunchKernelGGL(k_dp_wide<16>, dim3(blocks), dim3(1024), lds, s, a); break;
    default: break;
  }
}

// ---------------------------------------------------------------------------
// MAPQ: bayesian.rs:193-439 (FP64; a group of lanes per read)
// ---------------------------------------------------------------------------
constexpr double LN_10 = 2.302585092994045684017991454684364208;
constexpr double DB = -0.1 * LN_10;

__device__ __forceinline__ double logadd(double a, double b) {  // bayesian.rs:47-58
  if (a > b) {
    double t = a;
    a = b;
    b = t;
  }
  if (b >= a + 10.0) return a;
  if (b >= a + 4.0) return fmax(a - 1.0, 0.0);
  if (b >= a + 2.0) return fmax(a - 2.0, 0.0);
  return fmax(a - 3.0, 0.0);
}
__device__ __forceinline__ double py_int_round(double x) { return trunc(x + 0.5); }
__device__ __forceinline__ double ln_int(const DevLuts& T, int64_t k) {
  return k <= LN_INT_MAX ? T.ln_int[k] : log((double)k);
}
// bayesian.rs:72-86
__device__ double phred_n_choose_m(const DevLuts& T, int64_t n, int64_t m) {
  if (m > n / 2) m = n - m;
  if (m <= 0) return 0.0;
  double phred = 0.0;
  while (m > 1) {
    double t = ln_int(T, n) - ln_int(T, m);
    phred += t;
    n -= 1;
    m -= 1;
  }
  phred += ln_int(T, n);
  return py_int_round(10.0 * phred / LN_10);
}

struct QueryView {  // query bases/quals in the aligned orientation
  const uint8_t* seq;
  const uint8_t* qual;
  uint32_t L, qlen;
  bool rc;
  __device__ uint8_t base(uint32_t x) const { return rc ? comp_base(seq[L - 1 - x]) : seq[x]; }
  __device__ uint8_t q(uint32_t x) const { return rc ? qual[qlen - 1 - x] : qual[x]; }
};

// Sequential reads of the 2-bit genome: one word load per 32 bases.
struct GenomeCursor {
  const DevGenome& G;
  uint64_t w = ~0ull, g = 0;
  uint32_t n = 0;  // N bits of the 32 bases of word w
  __device__ explicit GenomeCursor(const DevGenome& g_) : G(g_) {}
  __device__ __forceinline__ uint8_t at(uint64_t p) {  // same value as genome_base(G, p)
    const uint64_t ww = p >> 5;
    if (ww != w) {
      w = ww;
      g = G.g2b[ww];
      n = (uint32_t)(G.nmask[ww >> 1] >> ((ww & 1) * 32));
    }
    const uint32_t i = (uint32_t)(p & 31);
    return ((n >> i) & 1u) ? (uint8_t)'N' : (uint8_t)("ACGT"[(g >> (2 * i)) & 3u]);
  }
};

// gc_content (bayesian.rs:445-456) of genome[gb, gb+span): G/C bases and
// non-N bases, counted 32 at a time (C = 01 and G = 10 are the codes whose
// two bits differ).
// Word w's share of genome[gb, end): G/C bases and non-N bases.
__device__ __forceinline__ void gc_word(const DevGenome& G, uint64_t w, uint64_t gb, uint64_t end, uint32_t& gc,
                                        uint32_t& total) {
  const uint64_t lo = gb > (w << 5) ? gb : (w << 5), hi = end < ((w + 1) << 5) ? end : ((w + 1) << 5);
  const uint32_t b0 = (uint32_t)(lo & 31), cnt = (uint32_t)(hi - lo);
  const uint32_t sel = (cnt == 32 ? 0xFFFFFFFFu : ((1u << cnt) - 1u)) << b0;
  const uint64_t g = G.g2b[w];
  uint64_t x = (g ^ (g >> 1)) & 0x5555555555555555ull;  // bit 2i: the two bits of base i differ
  x = (x | (x >> 1)) & 0x3333333333333333ull;
  x = (x | (x >> 2)) & 0x0F0F0F0F0F0F0F0Full;
  x = (x | (x >> 4)) & 0x00FF00FF00FF00FFull;
  x = (x | (x >> 8)) & 0x0000FFFF0000FFFFull;
  x = (x | (x >> 16)) & 0x00000000FFFFFFFFull;
  const uint32_t nb = (uint32_t)(G.nmask[w >> 1] >> ((w & 1) * 32));
  const uint32_t ok = sel & ~nb;
  gc += (uint32_t)__popc((uint32_t)x & ok);
  total += (uint32_t)__popc(ok);
}
__device__ void gc_count(const DevGenome& G, uint64_t gb, uint32_t span, uint32_t& gc, uint32_t& total) {
  gc = total = 0;
  if (!span) return;
  const uint64_t end = gb + span;
  for (uint64_t w = gb >> 5; w <= (end - 1) >> 5; w++) gc_word(G, w, gb, end, gc, total);
}

// The FP64 part of log_entropy (bayesian.rs:306-373) from its integer counts.
__device__ double entropy_from_counts(const DevLuts& T, int64_t n, int64_t mm, uint32_t readlen, int64_t insertion_runs,
                                      int64_t deletion_runs, uint32_t gc, uint32_t total) {
  double lnb, lnbm1;
  if (total <= (uint32_t)GC_LUT_MAX) {
    uint64_t idx = (uint64_t)total * (total + 1) / 2 + gc;
    lnb = T.gc_lnb[idx];
    lnbm1 = T.gc_lnbm1[idx];
  } else {
    double f = total == 0 ? 0.5 : (double)gc / (double)total;
    f = fmin(fmax(f, 1e-6), 1.0 - 1e-6);
    double bb = exp(-f * log(f / 2.0) - (1.0 - f) * log((1.0 - f) / 2.0));
    lnb = log(bb);
    lnbm1 = log(fmax(bb - 1.0, 1e-6));
  }
  double match_term = py_int_round(10.0 * ((double)n * lnb - T.ln_2g) / LN_10);
  double mismatch_term = py_int_round(10.0 * (double)mm * lnbm1 / LN_10);
  double position_choice = phred_n_choose_m(T, (int64_t)readlen, 2 * insertion_runs + deletion_runs);
  double deletion_length = 0.0;
  if (deletion_runs > 0) deletion_length = py_int_round((double)deletion_runs * 10.0 * T.ln_30 / LN_10);
  return match_term - mismatch_term - phred_n_choose_m(T, n, mm) - position_choice - deletion_length;
}

// bayesian.rs:306-373 log_entropy; ref_window = genome[gb, gb+span)
__device__ double log_entropy(const DevGenome& G, const DevLuts& T, const uint32_t* cig, uint32_t ncig,
                              const QueryView& Q, uint64_t gb, uint32_t span) {
  int64_t n = 0, mm = 0;
  uint32_t readlen = 0, refpos = 0;
  int64_t indel_runs = 0, deletion_runs = 0;
  uint32_t curtype = 0;
  GenomeCursor gcur(G);
  for (uint32_t c = 0; c < ncig; c++) {
    uint32_t op = cig[c] & 15u, len = cig[c] >> 4;
    if (op == 1) {
      for (uint32_t u = 0; u < len; u++) {
        if (curtype != 1) indel_runs++;
        curtype = 1;
        readlen++;
      }
    } else if (op == 2) {
      for (uint32_t u = 0; u < len; u++) {
        if (curtype != 2) {
          deletion_runs++;
          indel_runs++;
        }
        curtype = 2;
        refpos++;
      }
    } else if (op == 0) {
      for (uint32_t u = 0; u < len; u++) {
        if (curtype != 0) indel_runs++;
        curtype = 0;
        if (readlen >= Q.L || refpos >= span) break;
        uint32_t q = readlen < Q.qlen ? Q.q(readlen) : 0u;
        bool hq = q >= 10u;
        if (hq) n++;
        uint8_t aa = Q.base(readlen), rr = gcur.at(gb + refpos);
        if (aa != rr && aa != 'N' && rr != 'N' && hq) mm++;
        readlen++;
        refpos++;
      }
    }
  }
  uint32_t gc, total;
  gc_count(G, gb, span, gc, total);
  return entropy_from_counts(T, n, mm, readlen, indel_runs - deletion_runs, deletion_runs, gc, total);
}

// The log term of one missing-locus window starting at `start`, from its
// quality product z and ratio sum p1 (bayesian.rs:418-433).
__device__ __forceinline__ double missing_window_log(const DevLuts& T, double z, double p1, const uint8_t* qual,
                                                     uint32_t qlen, const uint8_t* rcode, uint32_t nwin,
                                                     uint32_t start) {
  double rep = 0.0;
  if (start < nwin) {
    const uint8_t c = rcode[start];
    if (c == 1) rep = 1.0;
    else if (c == 2) rep = T.nrp[start < qlen ? qual[start] : 30];
  }
  const double p2 = fmin(fmax(1.0 - z * (1.0 + p1), 0.0), 1.0);
  const double pm = fmin(fmax(1.0 - (1.0 - p2) * (1.0 - rep), 1e-300), 1.0);
  return log(pm);
}

// bayesian.rs:405-439 missing_locus_probability; repeat mask from the seed
// pass (0: 0.0, 1: 1.0, 2: 10^(-q/10) at the window start, seeds.rs:125-139)
__device__ double missing_locus(const DevLuts& T, const uint8_t* qual, uint32_t qlen, const uint8_t* rcode,
                                uint32_t nwin) {
  constexpr uint32_t k = K, PH = 5;
  if (qlen < k) return 0.5;
  // The reference loops phase { window { position } }.  One pass over the
  // positions feeds the five phases together: each window's product and sum
  // still run over its positions in order, and each phase still adds its
  // windows' log terms in window order, so every double is the same.
  double z[PH], p1[PH], log_p[PH];
  uint32_t rel[PH];  // position inside the phase's current window
#pragma unroll
  for (uint32_t ph = 0; ph < PH; ph++) {
    z[ph] = 1.0;
    p1[ph] = -0.0;
    log_p[ph] = 0.0;
    rel[ph] = 0;
  }
  for (uint32_t t = 0; t < qlen; t++) {
    const uint32_t q = qual[t];
    const double om = T.omq[q], qr = T.qratio[q];
#pragma unroll
    for (uint32_t ph = 0; ph < PH; ph++) {
      if (t < ph) continue;
      const uint32_t start = t - rel[ph];
      if (start + k <= qlen) {  // windows past the end do not exist in the reference
        z[ph] = z[ph] * om;
        p1[ph] = p1[ph] + qr;
        if (rel[ph] == k - 1) {
          log_p[ph] += missing_window_log(T, z[ph], p1[ph], qual, qlen, rcode, nwin, start);
          z[ph] = 1.0;
          p1[ph] = -0.0;
        }
      }
      rel[ph] = rel[ph] == k - 1 ? 0u : rel[ph] + 1u;
    }
  }
  double offset_means = 0.0;
#pragma unroll
  for (uint32_t ph = 0; ph < PH; ph++) offset_means += exp(log_p[ph]);
  return fmin(offset_means / 5.0, 0.99);
}

// ---- the same two terms over a whole wave (one read, all 64 lanes active) ----
__device__ __forceinline__ double readlane_f64(double v, int l) {
  const uint64_t b = (uint64_t)__double_as_longlong(v);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, l);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), l);
  return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}
__device__ __forceinline__ uint64_t readlane_u64(uint64_t v, int l) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, l);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), l);
  return ((uint64_t)hi << 32) | lo;
}


// log_entropy with the per-base and per-word counting spread over the group:
// integer counts (exact in any order), then the same FP64 tail.  The operation
// scan runs on every lane alike; an M run's loop in log_entropy stops at the
// read's or the window's end, which bounds the bases counted here the same way.
template <int NL>
__device__ double group_log_entropy(const DevGenome& G, const DevLuts& T, const uint32_t* cig, uint32_t ncig,
                                    const QueryView& Q, uint64_t gb, uint32_t span) {
  const uint32_t gl = (uint32_t)lane_id() & (uint32_t)(NL - 1);
  int64_t indel_runs = 0, deletion_runs = 0;
  uint32_t readlen = 0, refpos = 0, curtype = 0, n_l = 0, mm_l = 0;
  for (uint32_t c = 0; c < ncig; c++) {
    const uint32_t op = cig[c] & 15u, len = cig[c] >> 4;
    if (len == 0) continue;
    if (op == 1) {
      if (curtype != 1) indel_runs++;
      curtype = 1;
      readlen += len;
    } else if (op == 2) {
      if (curtype != 2) {
        deletion_runs++;
        indel_runs++;
      }
      curtype = 2;
      refpos += len;
    } else if (op == 0) {
      if (curtype != 0) indel_runs++;
      curtype = 0;
      const uint32_t ra = readlen < Q.L ? Q.L - readlen : 0u, rb = refpos < span ? span - refpos : 0u;
      const uint32_t cnt = min(len, min(ra, rb));
      for (uint32_t u = gl; u < cnt; u += NL) {
        const uint32_t x = readlen + u;
        const uint32_t q = x < Q.qlen ? Q.q(x) : 0u;
        const bool hq = q >= 10u;
        const uint8_t aa = Q.base(x), rr = genome_base(G, gb + refpos + u);
        n_l += hq ? 1u : 0u;
        mm_l += (aa != rr && aa != 'N' && rr != 'N' && hq) ? 1u : 0u;
      }
      readlen += cnt;
      refpos += cnt;
    }
  }
  uint32_t gc_l = 0, tot_l = 0;
  if (span) {
    const uint64_t end = gb + span;
    for (uint64_t w = (gb >> 5) + gl; w <= (end - 1) >> 5; w += NL) gc_word(G, w, gb, end, gc_l, tot_l);
  }
  const int64_t n = group_sum_u32<NL>(n_l), mm = group_sum_u32<NL>(mm_l);
  return entropy_from_counts(T, n, mm, readlen, indel_runs - deletion_runs, deletion_runs, group_sum_u32<NL>(gc_l),
                             group_sum_u32<NL>(tot_l));
}

// missing_locus with one window per lane: window 5 j + ph starts at ph + 15 j
// and multiplies / adds its 15 positions in order, as there; lane ph of the
// group then adds phase ph's window terms in window order, and the five phases
// are summed in order, so every double is the same.
template <int NL>
__device__ double group_missing_locus(const DevLuts& T, const uint8_t* qual, uint32_t qlen, const uint8_t* rcode,
                                      uint32_t nwin) {
  static_assert(NL >= 8 && NL <= 64 && (NL & (NL - 1)) == 0, "group of 8..64 lanes");
  constexpr uint32_t k = K, PH = 5;
  if (qlen < k) return 0.5;
  const uint32_t lane = (uint32_t)lane_id(), gl = lane & (uint32_t)(NL - 1), g0 = lane - gl;
  const uint32_t nids = PH * ((qlen - k) / k + 1);
  double acc = 0.0;  // lanes 0..4 of the group: log_p of phase `gl`
  for (uint32_t base = 0; base < nids; base += NL) {
    const uint32_t id = base + gl, start = id % PH + k * (id / PH);
    const bool valid = id < nids && start + k <= qlen;
    double term = 0.0;
    if (valid) {
      double z = 1.0, p1 = -0.0;
      for (uint32_t i = 0; i < k; i++) {
        const uint32_t q = qual[start + i];
        z = z * T.omq[q];
        p1 = p1 + T.qratio[q];
      }
      term = missing_window_log(T, z, p1, qual, qlen, rcode, nwin, start);
    }
    const uint32_t first = (gl + PH - base % PH) % PH;  // first offset in this pass of phase `gl`
    for (uint32_t o0 = 0; o0 < (uint32_t)NL; o0 += PH) {
      const uint32_t o = first + o0;
      const int src = (int)(g0 + (o < (uint32_t)NL ? o : 0u));
      const double v = __shfl(term, src, 64);
      const int ok = __shfl(valid ? 1 : 0, src, 64);
      if (gl < PH && o < (uint32_t)NL && ok) acc += v;
    }
  }
  double offset_means = 0.0;
  for (uint32_t ph = 0; ph < PH; ph++) offset_means += exp(__shfl(acc, (int)(g0 + ph), 64));
  return fmin(offset_means / 5.0, 0.99);
}

// bayesian.rs:251-282 finalize_mapq, given its two long parts: log_ent =
// log_entropy(..) and missing = missing_locus(..), the latter read only when
// the entropy term is positive (as the reference computes it only then).
// Whether the MAPQ rests on a value that a one-ulp difference between this
// device's exp / log / log10 and the reference's libm could move across a
// rounding point: x within 1e-9 of an integer.
__device__ __forceinline__ bool near_integer(double x) { return fabs(x - rint(x)) < 1e-9; }
// logadd's branch thresholds |a - b| = 10, 4, 2 (bayesian.rs:47-58), where a step of 1 sits.
__device__ __forceinline__ bool near_logadd_step(double a, double b) {
  const double d = fabs(a - b);
  return fabs(d - 10.0) < 1e-9 || fabs(d - 4.0) < 1e-9 || fabs(d - 2.0) < 1e-9;
}

// finalize_mapq (bayesian.rs:251-282).  `near` reports whether the integer MAPQ
// depends on a device-transcendental value within 1e-9 of a rounding point:
// trunc(ln(p_wrong)/DB + 0.5), a logadd threshold against the continuous
// not-found Phred, or round(pp).  (The entropy is exact: its logarithms come
// from host tables.)
__device__ void finalize_mapq_from(const DevLuts& T, double sum_exp, double log_ent, double missing, uint8_t& mapq,
                                   double& posterior, bool& near) {
  double bp = 1.0 / fmax(sum_exp, 1.0);
  double pw = fmax(1.0 - bp, T.posterior_floor);
  const double xw = log(pw) / DB;
  double wrongmapping = py_int_round(xw);
  double entropy = log_ent - 3.0;
  near = false;
  if (entropy <= 0.0) {
    mapq = 0;
    posterior = 0.0;
    return;
  }
  double p_missing = fmax(missing, 1e-30);
  double notfound = -10.0 * log10(p_missing);
  double combined = logadd(wrongmapping, notfound);
  double pp = fmin(fmax(logadd(entropy, combined), 0.0), 99.0);
  near = near_integer(xw + 0.5) || near_logadd_step(wrongmapping, notfound) || near_logadd_step(entropy, combined) ||
         near_integer(pp - 0.5);
  mapq = (uint8_t)round(pp);
  posterior = 1.0 - pow(10.0, -pp / 10.0);
}

struct RecOut {  // pbw_record layout
  uint64_t pos, mate_pos;
  int64_t tlen;
  double log_lk, posterior;
  int32_t contig, mate_contig;
  uint32_t cigar_off, cigar_n;
  uint16_t flag;
  uint8_t mapq, strand, seq_rc, pad[3];
};
static_assert(sizeof(RecOut) == 64, "record layout");

// MAPQ_LANES lanes per read, all in the same control flow: every lane takes
// the selection; the group then computes the log-entropy and missing-locus
// terms together, and its first lane combines them and writes the record.
constexpr int MAPQ_LANES = 8;
__global__ void k_mapq(MapqArgs a) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t gl = t & (uint32_t)(MAPQ_LANES - 1);
  uint32_t r = t / MAPQ_LANES;
  if (a.rlist) {
    if (r >= a.nlist) return;
    r = a.rlist[r];
  } else if (r >= a.R.n) {
    return;
  }
  const uint64_t so = a.R.seq_off[r], qo = a.R.qual_off[r];
  const uint32_t L = (uint32_t)(a.R.seq_off[r + 1] - so), qlen = (uint32_t)(a.R.qual_off[r + 1] - qo);
  RecOut o;
  o.mate_pos = 0;
  o.tlen = 0;
  o.mate_contig = -1;
  o.cigar_off = 0;
  o.pad[0] = o.pad[1] = o.pad[2] = 0;
  const uint32_t nj = a.njobs[r];
  if (nj == 0) {  // single.rs:362-383 unmapped_record
    if (gl) return;
    o.pos = 0;
    o.log_lk = 0.0;
    o.posterior = 0.0;
    o.contig = -1;
    o.cigar_n = 1;
    o.flag = 4;
    o.mapq = 0;
    o.strand = 0;
    o.seq_rc = 0;
    a.best[r] = 0xFFFFFFFFu;
    ((RecOut*)a.rec)[r] = o;
    return;
  }
  const uint64_t s0 = (uint64_t)r * TOP_CANDIDATES;
  // slots in use: all, or the prefilter's kept ones (never empty: the top ranks are kept)
  uint32_t use = nj >= 32 ? 0xFFFFFFFFu : ((1u << nj) - 1u);
  if (a.keep && (a.keep[r] & use)) use &= a.keep[r];
  const uint32_t k0 = (uint32_t)__builtin_ctz(use);
  double mx = a.out[s0 + k0].log_lk;
  uint32_t bi = k0;
  for (uint32_t k = 0; k < nj; k++) {  // select_best: last maximum
    if (!((use >> k) & 1u)) continue;
    double v = a.out[s0 + k].log_lk;
    if (v >= mx) {
      mx = v;
      bi = k;
    }
  }
  double se = -0.0;
  for (uint32_t k = 0; k < nj; k++)
    if ((use >> k) & 1u) se += exp(a.out[s0 + k].log_lk - mx);
  const DpJob job = a.jobs[s0 + bi];
  const DpOut dout = a.out[s0 + bi];
  QueryView Q{a.R.seq + so, a.R.qual + qo, L, qlen, job.rc != 0};
  const uint32_t* cig = a.cigar + (s0 + bi) * a.cigar_cap;
  const double le =
      group_log_entropy<MAPQ_LANES>(a.G, a.T, cig, dout.cigar_n, Q, job.ref_gb + dout.ref_off, dout.span);
  const double miss = le - 3.0 <= 0.0 ? 1.0
                                      : group_missing_locus<MAPQ_LANES>(a.T, a.R.qual + qo, qlen, a.rcode + so,
                                                                        L >= (uint32_t)K ? L - K + 1 : 0);
  if (gl) return;
  uint8_t mapq;
  double post;
  bool near;
  finalize_mapq_from(a.T, se, le, miss, mapq, post, near);
  if (near && a.ctr) atomicAdd(&a.ctr->mapq_near, 1ull);
  const uint32_t contig = a.cands[job.cand].contig;
  o.pos = job.ref_gb - a.G.coff[contig] + dout.ref_off + 1;  // SAM is 1-based (single.rs:192)
  o.log_lk = mx;
  o.posterior = post;
  o.contig = (int32_t)contig;
  o.cigar_n = dout.cigar_n;
  o.flag = job.rc ? 16 : 0;
  o.mapq = mapq;
  o.strand = job.rc;
  o.seq_rc = job.rc;
  a.best[r] = (uint32_t)(s0 + bi);
  ((RecOut*)a.rec)[r] = o;
}

void launch_mapq(const MapqArgs& a, hipStream_t s) {
  const uint32_t n = a.rlist ? a.nlist : a.R.n;
  if (n == 0) return;
  hipLaunchKernelGGL(k_mapq, dim3(((uint64_t)MAPQ_LANES * n + 127) / 128), dim3(128), 0, s, a);
}

// ---------------------------------------------------------------------------
// Candidate prefilter in the mapping path (no counterpart in the reference).
// It sits between cluster_hits (single.rs:91) and the DP loop (:114-150).
// ---------------------------------------------------------------------------
// Per read: keep the job slots whose candidate ranks < top_k or scores
// >= threshold; list them for the DP.
__global__ void k_pf_select(PfArgs a) {
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= a.R.n) return;
  const uint32_t nj = a.njobs[r];
  const uint32_t m = (uint32_t)(a.R.seq_off[r + 1] - a.R.seq_off[r]);
  const uint64_t s0 = (uint64_t)r * TOP_CANDIDATES;
  uint32_t mask = 0, nk = 0;
  unsigned long long skipped = 0;
  for (uint32_t k = 0; k < nj; k++) {
    const DpJob job = a.jobs[s0 + k];
    const uint32_t rank = (uint32_t)(job.cand - s0);
    if (rank < a.top_k || a.logit[job.cand] >= a.threshold) {
      mask |= 1u << k;
      nk++;
    } else if (m >= 1 && job.n >= 1 && (int64_t)job.hint + job.band >= 0 &&
               (int64_t)m + job.hint - job.band <= (int64_t)job.n) {
      skipped += quad_cells(m, job.hint, job.band, job.n);
    }
  }
  a.keep[r] = mask;
  if (nk) {
    uint32_t pos = atomicAdd(a.nlist, nk);
    for (uint32_t k = 0; k < nj; k++)
      if ((mask >> k) & 1u) a.list[pos++] = (uint32_t)(s0 + k);
  }
  if (nj) atomicAdd(&a.ctr[0], (unsigned long long)nj);
  if (nk) atomicAdd(&a.ctr[1], (unsigned long long)nk);
  if (skipped) atomicAdd(&a.ctr[2], skipped);
}

// Per read with skipped slots: re-run it plain when its prefiltered result is
// ambiguous (MAPQ below mapq_min, or a kept second candidate within
// close_nats of the best), or when it falls on the audit sample.
__global__ void k_pf_verdict(PfArgs a) {
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= a.R.n) return;
  const uint32_t nj = a.njobs[r];
  const uint32_t full = nj >= 32 ? 0xFFFFFFFFu : ((1u << nj) - 1u);
  const uint32_t mask = a.keep[r] & full;
  a.verdict[r] = 0;
  if (mask == full) return;
  const uint64_t s0 = (uint64_t)r * TOP_CANDIDATES;
  double mx = -HUGE_VAL, sec = -HUGE_VAL;
  uint32_t bi = 0;
  for (uint32_t k = 0; k < nj; k++) {
    if (!((mask >> k) & 1u)) continue;
    const double v = a.out[s0 + k].log_lk;
    if (v >= mx) {
      if (mx != -HUGE_VAL) sec = mx > sec ? mx : sec;
      mx = v;
      bi = k;
    } else if (v > sec) {
      sec = v;
    }
  }
  (void)bi;
  const uint8_t mapq = ((const RecOut*)a.rec)[r].mapq;
  const bool fallback = mapq < a.mapq_min || sec >= mx - (double)a.close_nats;
  const bool audit = !fallback && a.audit_period && ((a.read_base + r) % a.audit_period) == 0;
  if (!fallback && !audit) return;
  a.verdict[r] = fallback ? 1 : 2;
  const uint32_t p = atomicAdd(a.nrlist, 1u);
  a.rlist[p] = r;
  const uint32_t ns = (uint32_t)__builtin_popcount(full & ~mask);
  uint32_t pos = atomicAdd(a.nlist, ns);
  for (uint32_t k = 0; k < nj; k++)
    if (!((mask >> k) & 1u)) a.list[pos++] = (uint32_t)(s0 + k);
  atomicAdd(&a.ctr[3], (unsigned long long)ns);
  if (fallback) atomicAdd(&a.audit[2], 1ull);
}

// Fallback reads take the plain record; audited reads are compared field by
// field (POS, RNAME, FLAG, MAPQ, strand and the CIGAR words).
__global__ void k_pf_merge(PfArgs a, uint32_t n) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const uint32_t r = a.rlist[t];
  const RecOut* p2 = (const RecOut*)(a.rec2) + r;
  RecOut* p1 = (RecOut*)(a.rec_out) + r;
  if (a.verdict[r] == 1) {
    *p1 = *p2;
    a.best[r] = a.best2[r];
    return;
  }
  const RecOut x = *p1, y = *p2;
  bool same = x.pos == y.pos && x.contig == y.contig && x.flag == y.flag && x.mapq == y.mapq && x.strand == y.strand &&
              x.cigar_n == y.cigar_n;
  const uint32_t b1 = a.best[r], b2 = a.best2[r];
  if (same && b1 != b2 && b1 != 0xFFFFFFFFu && b2 != 0xFFFFFFFFu) {
    const uint32_t* c1 = a.cigar + (uint64_t)b1 * a.cigar_cap;
    const uint32_t* c2 = a.cigar + (uint64_t)b2 * a.cigar_cap;
    for (uint32_t u = 0; u < x.cigar_n && u < a.cigar_cap; u++) same &= c1[u] == c2[u];
  }
  atomicAdd(&a.audit[0], 1ull);
  if (!same) atomicAdd(&a.audit[1], 1ull);
}

void launch_pf_select(const PfArgs& a, hipStream_t s) {
  if (a.R.n == 0) return;
  hipLaunchKernelGGL(k_pf_select, dim3((a.R.n + 127) / 128), dim3(128), 0, s, a);
}
void launch_pf_verdict(const PfArgs& a, hipStream_t s) {
  if (a.R.n == 0) return;
  hipLaunchKernelGGL(k_pf_verdict, dim3((a.R.n + 127) / 128), dim3(128), 0, s, a);
}
void launch_pf_merge(const PfArgs& a, uint32_t n, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_pf_merge, dim3((n + 127) / 128), dim3(128), 0, s, a, n);
}

// Winning CIGARs into one compact array (one wave per record).
__global__ void k_gather_cigar(const uint32_t* cig, const uint64_t* src, const uint32_t* off, const uint32_t* cnt,
                               uint32_t n, uint32_t* out) {
  uint32_t w = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
  if (w >= n) return;
  for (uint32_t t = lane_id(); t < cnt[w]; t += 64) out[off[w] + t] = cig[src[w] + t];
}
void gather_cigar(const uint32_t* cig, const uint64_t* src, const uint32_t* off, const uint32_t* cnt, uint32_t n,
                  uint32_t* out, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_gather_cigar, dim3((n + 3) / 4), dim3(256), 0, s, cig, src, off, cnt, n, out);
}

// ---------------------------------------------------------------------------
// records to host memory: CIGAR word counts, offsets and runs of one batch.
// src[i] is the first CIGAR word of record i in its pool (bit 63 selects the
// second pool), or ~0 for an unmapped read, whose CIGAR is the single run
// SoftClip(L) (single.rs:368).
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t rec_words(const RecOut* rec, const uint64_t* src, uint32_t i) {
  return src[i] == ~0ull ? 1u : rec[i].cigar_n;
}

__global__ __launch_bounds__(REC_BLOCK) void k_rec_block_sums(const RecOut* rec, const uint64_t* src, uint32_t n,
                                                              uint32_t* bsum) {
  __shared__ uint32_t sh[REC_BLOCK / 64];
  const uint32_t i = blockIdx.x * REC_BLOCK + threadIdx.x;
  uint32_t tot;
  team_excl_scan<REC_BLOCK>(i < n ? rec_words(rec, src, i) : 0u, tot, sh);
  if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

// Exclusive scan of the block sums, starting at `base`; the batch's word count to total[0].
__global__ __launch_bounds__(REC_BLOCK) void k_rec_scan_sums(uint32_t* bsum, uint32_t nb, uint32_t base,
                                                             uint32_t* total) {
  __shared__ uint32_t sh[REC_BLOCK / 64];
  uint32_t carry = base;
  for (uint32_t b0 = 0; b0 < nb; b0 += REC_BLOCK) {
    const uint32_t b = b0 + threadIdx.x;
    uint32_t tot;
    const uint32_t ex = team_excl_scan<REC_BLOCK>(b < nb ? bsum[b] : 0u, tot, sh);
    if (b < nb) bsum[b] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0) total[0] = carry - base;
}

__global__ __launch_bounds__(REC_BLOCK) void k_rec_offsets(RecOut* rec, const uint64_t* src, uint32_t n,
                                                           const uint32_t* bsum) {
  __shared__ uint32_t sh[REC_BLOCK / 64];
  const uint32_t i = blockIdx.x * REC_BLOCK + threadIdx.x;
  const uint32_t v = i < n ? rec_words(rec, src, i) : 0u;
  uint32_t tot;
  const uint32_t ex = team_excl_scan<REC_BLOCK>(v, tot, sh);
  if (i < n) {
    rec[i].cigar_off = bsum[blockIdx.x] + ex;
    rec[i].cigar_n = v;
  }
}

// Thread per record: its runs to out[cigar_off - base ...].
__global__ void k_rec_cigar(const RecOut* rec, const uint64_t* src, const uint32_t* cig0, const uint32_t* cig1,
                            const uint64_t* seq_off, uint32_t n, uint32_t base, uint32_t* out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t sv = src[i];
  uint32_t* o = out + (rec[i].cigar_off - base);
  if (sv == ~0ull) {
    o[0] = (uint32_t)((seq_off[i + 1] - seq_off[i]) << 4) | 4u;
    return;
  }
  const uint32_t* pool = (sv >> 63) ? cig1 : cig0;
  const uint64_t at = sv & ~(1ull << 63);
  const uint32_t cnt = rec[i].cigar_n;
  for (uint32_t t = 0; t < cnt; t++) o[t] = pool[at + t];
}

// src of single-end records: slot `best[i]` of the DP output pool (cap words per slot).
__global__ void k_rec_src_se(const uint32_t* best, uint32_t n, uint32_t cap, uint64_t* src) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) src[i] = best[i] == 0xFFFFFFFFu ? ~0ull : (uint64_t)best[i] * cap;
}

void launch_rec_src_se(const uint32_t* best, uint32_t n, uint32_t cap, uint64_t* src, hipStream_t s) {
  if (n) hipLaunchKernelGGL(k_rec_src_se, dim3((n + 255) / 256), dim3(256), 0, s, best, n, cap, src);
}
void launch_rec_offsets(uint8_t* recb, const uint64_t* src, uint32_t n, uint32_t base, uint32_t* bsum,
                        uint32_t* total, hipStream_t s) {
  RecOut* rec = reinterpret_cast<RecOut*>(recb);
  const uint32_t nb = (n + REC_BLOCK - 1) / REC_BLOCK;
  if (nb) hipLaunchKernelGGL(k_rec_block_sums, dim3(nb), dim3(REC_BLOCK), 0, s, rec, src, n, bsum);
  hipLaunchKernelGGL(k_rec_scan_sums, dim3(1), dim3(REC_BLOCK), 0, s, bsum, nb, base, total);
  if (nb) hipLaunchKernelGGL(k_rec_offsets, dim3(nb), dim3(REC_BLOCK), 0, s, rec, src, n, bsum);
}
void launch_rec_cigar(const uint8_t* recb, const uint64_t* src, const uint32_t* cig0, const uint32_t* cig1,
                      const uint64_t* seq_off, uint32_t n, uint32_t base, uint32_t* out, hipStream_t s) {
  const RecOut* rec = reinterpret_cast<const RecOut*>(recb);
  if (n)
    hipLaunchKernelGGL(k_rec_cigar, dim3((n + 255) / 256), dim3(256), 0, s, rec, src, cig0, cig1, seq_off, n, base,
                       out);
}

// Same, from two pools (bit 63 of src selects the second).
__global__ void k_gather_cigar2(const uint32_t* cig0, const uint32_t* cig1, const uint64_t* src, const uint32_t* off,
                                const uint32_t* cnt, uint32_t n, uint32_t* out) {
  uint32_t w = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
  if (w >= n) return;
  const uint64_t sv = src[w];
  const uint32_t* base = (sv >> 63) ? cig1 : cig0;
  const uint64_t o = sv & ~(1ull << 63);
  for (uint32_t t = lane_id(); t < cnt[w]; t += 64) out[off[w] + t] = base[o + t];
}
void gather_cigar2(const uint32_t* cig0, const uint32_t* cig1, const uint64_t* src, const uint32_t* off,
                   const uint32_t* cnt, uint32_t n, uint32_t* out, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_gather_cigar2, dim3((n + 3) / 4), dim3(256), 0, s, cig0, cig1, src, off, cnt, n, out);
}

#include "pbw_pair_kernels.inc"

// ---------------------------------------------------------------------------
// dense bucket table from the sorted key list
// ---------------------------------------------------------------------------
// Presence bitmap of the bucket table, bit i = bucket of code rotl30(i, rot)
// non-empty: one wave per 64 consecutive bits.
__global__ void k_fill_pres(const uint32_t* koff, uint64_t* pres, uint32_t rot) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < CODE_SPACE; i += stride) {
    const uint32_t c = rot ? rotr30((uint32_t)i, 30 - rot) : (uint32_t)i;  // rotl by rot
    const uint64_t m = __ballot(koff[c + 1] > koff[c]);
    if (lane_id() == 0) pres[i >> 6] = m;
  }
}
__global__ void k_fill_stride(uint64_t* out, uint32_t n, uint64_t stride) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (uint64_t)i * stride;
}
void launch_fill_stride(uint64_t* out, uint32_t n, uint64_t stride, hipStream_t s) {
  if (n) hipLaunchKernelGGL(k_fill_stride, dim3((n + 255) / 256), dim3(256), 0, s, out, n, stride);
}
__global__ void k_scatter_u64(uint64_t* out, const uint32_t* idx, const uint64_t* val, uint32_t n) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) out[idx[k]] = val[k];
}
// The seed pass's overflow reads (more than cap1 hits), found on the device in
// three passes without global atomics: per-block counts, one block's scan of
// them, and the list itself in read order (entry k: read, exact hit count,
// and slots base + the hits of the overflow reads before it).  st: [0]
// overflow reads, [1] most hits of any read, [2..5] reads with more than
// HIT_CAP << k hits (k = 0..3), [6..7] total overflow hits, [8..9] every read's
// hits (64 bits), [10] reads with more than HIT_CAP << 4 hits.
constexpr uint32_t OV_BLOCK = 1024, OV_PART = 12;
// Block-wide exclusive sum of x (OV_BLOCK threads); *total gets the sum.
__device__ __forceinline__ uint32_t ov_block_scan(uint32_t x, uint32_t* sh, uint32_t& total) {
  const uint32_t w = threadIdx.x >> 6, l = (uint32_t)lane_id();
  const uint32_t incl = wave_incl_scan(x);
  if (l == 63) sh[w] = incl;
  __syncthreads();
  uint32_t before = 0, tot = 0;
  for (uint32_t v = 0; v < OV_BLOCK / 64; v++) {
    const uint32_t t = sh[v];
    before += v < w ? t : 0u;
    tot += t;
  }
  __syncthreads();
  total = tot;
  return before + incl - x;
}
// pass 1: per block (OV_PART words) [count, max, over0..3, overflow hits lo, hi,
// all hits lo, hi, over4] (hit sums in 64 bits: the reads' low and high 16-bit
// halves summed apart, each sum < 2^26)
__global__ __launch_bounds__(OV_BLOCK) void k_ov_count(const uint32_t* nhits, const uint8_t* dropped, uint32_t B,
                                                      uint32_t cap1, uint32_t* part) {
  __shared__ uint32_t sh[OV_BLOCK / 64];
  __shared__ uint32_t smax[OV_BLOCK / 64];
  const uint32_t i = blockIdx.x * OV_BLOCK + threadIdx.x;
  const uint32_t nh = i < B ? nhits[i] : 0u;
  const bool ov = dropped ? (i < B && dropped[i] != 0) : nh > cap1;
  uint32_t mx = nh;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) mx = max(mx, (uint32_t)__shfl_xor((int)mx, o, 64));
  if (lane_id() == 0) smax[threadIdx.x >> 6] = mx;
  uint32_t c[10] = {ov ? 1u : 0u, ov ? (nh & 0xFFFFu) : 0u, ov ? (nh >> 16) : 0u, 0, 0, 0, 0, nh & 0xFFFFu, nh >> 16,
                    (ov && nh > ((uint32_t)HIT_CAP << 4)) ? 1u : 0u};
#pragma unroll
  for (int j = 0; j < 4; j++) c[3 + j] = (ov && nh > ((uint32_t)HIT_CAP << j)) ? 1u : 0u;
  uint32_t tot[10];
#pragma unroll
  for (int j = 0; j < 10; j++) (void)ov_block_scan(c[j], sh, tot[j]);
  if (threadIdx.x == 0) {
    uint32_t m = 0;
    for (uint32_t v = 0; v < OV_BLOCK / 64; v++) m = max(m, smax[v]);
    uint32_t* p = part + (uint64_t)blockIdx.x * OV_PART;
    const uint64_t hits = (uint64_t)tot[1] + ((uint64_t)tot[2] << 16);
    const uint64_t all = (uint64_t)tot[7] + ((uint64_t)tot[8] << 16);
    p[0] = tot[0];
    p[1] = m;
    p[2] = tot[3];
    p[3] = tot[4];
    p[4] = tot[5];
    p[5] = tot[6];
    p[6] = (uint32_t)hits;
    p[7] = (uint32_t)(hits >> 32);
    p[8] = (uint32_t)all;
    p[9] = (uint32_t)(all >> 32);
    p[10] = tot[9];
  }
}
// pass 2 (one block): block bases (count, hits) in place of the counts, and st
__global__ __launch_bounds__(OV_BLOCK) void k_ov_scan(uint32_t* part, uint32_t nb, uint32_t* st) {
  __shared__ uint32_t sh[OV_BLOCK / 64];
  __shared__ unsigned long long carry_h, carry_all;
  __shared__ uint32_t carry_c, cmax, cover[5];
  if (threadIdx.x == 0) {
    carry_h = carry_all = 0;
    carry_c = cmax = 0;
    for (int j = 0; j < 5; j++) cover[j] = 0;
  }
  __syncthreads();
  for (uint32_t b0 = 0; b0 < nb; b0 += OV_BLOCK) {
    const uint32_t bi = b0 + threadIdx.x;
    uint32_t* p = part + (uint64_t)bi * OV_PART;
    const bool v = bi < nb;
    const uint32_t cnt = v ? p[0] : 0u;
    const uint64_t hits = v ? ((uint64_t)p[6] | ((uint64_t)p[7] << 32)) : 0ull;  // < 2^42
    // the hits in 64 bits: three 16-bit parts scanned apart (each sum < 2^26)
    uint32_t tc, tl, tu, tv;
    const uint32_t ec = ov_block_scan(cnt, sh, tc);
    const uint32_t el = ov_block_scan((uint32_t)hits & 0xFFFFu, sh, tl);
    const uint32_t eu = ov_block_scan((uint32_t)(hits >> 16) & 0xFFFFu, sh, tu);
    const uint32_t ev = ov_block_scan((uint32_t)(hits >> 32), sh, tv);
    const uint64_t eh = (uint64_t)el + ((uint64_t)eu << 16) + ((uint64_t)ev << 32);
    const uint64_t th = (uint64_t)tl + ((uint64_t)tu << 16) + ((uint64_t)tv << 32);
    // every read's hits: only the total
    const uint64_t all = v ? ((uint64_t)p[8] | ((uint64_t)p[9] << 32)) : 0ull;
    uint32_t al, au, av;
    (void)ov_block_scan((uint32_t)all & 0xFFFFu, sh, al);
    (void)ov_block_scan((uint32_t)(all >> 16) & 0xFFFFu, sh, au);
    (void)ov_block_scan((uint32_t)(all >> 32), sh, av);
    const uint64_t ta = (uint64_t)al + ((uint64_t)au << 16) + ((uint64_t)av << 32);
    const uint32_t m = v ? p[1] : 0u;
    uint32_t o[5];
    for (int j = 0; j < 4; j++) o[j] = v ? p[2 + j] : 0u;
    o[4] = v ? p[10] : 0u;
    const unsigned long long hb = carry_h;
    const uint32_t cb = carry_c;
    if (v) {
      p[0] = cb + ec;
      p[6] = (uint32_t)(hb + eh);
      p[7] = (uint32_t)((hb + eh) >> 32);
    }
    uint32_t mm = m;
#pragma unroll
    for (int q = 32; q > 0; q >>= 1) mm = max(mm, (uint32_t)__shfl_xor((int)mm, q, 64));
    uint32_t to[5];
    for (int j = 0; j < 5; j++) (void)ov_block_scan(o[j], sh, to[j]);
    if (lane_id() == 0) atomicMax(&cmax, mm);
    __syncthreads();
    if (threadIdx.x == 0) {
      carry_c = cb + tc;
      carry_h = hb + th;
      carry_all += ta;
      for (int j = 0; j < 5; j++) cover[j] += to[j];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    st[0] = carry_c;
    st[1] = cmax;
    for (int j = 0; j < 4; j++) st[2 + j] = cover[j];
    st[6] = (uint32_t)carry_h;
    st[7] = (uint32_t)(carry_h >> 32);
    st[8] = (uint32_t)carry_all;
    st[9] = (uint32_t)(carry_all >> 32);
    st[10] = cover[4];
  }
}
// pass 3: the list, in read order
__global__ __launch_bounds__(OV_BLOCK) void k_ov_write(const uint32_t* nhits, const uint8_t* dropped, uint32_t B,
                                                      uint32_t cap1, uint64_t base, const uint32_t* part,
                                                      uint32_t* list, uint64_t* off, uint32_t* cap) {
  __shared__ uint32_t sh[OV_BLOCK / 64];
  const uint32_t i = blockIdx.x * OV_BLOCK + threadIdx.x;
  const uint32_t nh = i < B ? nhits[i] : 0u;
  const bool ov = dropped ? (i < B && dropped[i] != 0) : nh > cap1;
  uint32_t t0, t1, t2;
  const uint32_t ec = ov_block_scan(ov ? 1u : 0u, sh, t0);
  // the hits before this read in 64 bits: low and high 16-bit halves apart
  const uint32_t el = ov_block_scan(ov ? (nh & 0xFFFFu) : 0u, sh, t1);
  const uint32_t eu = ov_block_scan(ov ? (nh >> 16) : 0u, sh, t2);
  const uint64_t eh = (uint64_t)el + ((uint64_t)eu << 16);
  if (!ov) return;
  const uint32_t* p = part + (uint64_t)blockIdx.x * OV_PART;
  const uint32_t k = p[0] + ec;
  const uint64_t hb = (uint64_t)p[6] | ((uint64_t)p[7] << 32);
  list[k] = i;
  off[k] = base + hb + eh;
  cap[k] = nh;
}
uint32_t seed_overflow_blocks(uint32_t B) { return (B + OV_BLOCK - 1) / OV_BLOCK; }
void launch_seed_overflow(const uint32_t* nhits, const uint8_t* dropped, uint32_t B, uint32_t cap1, uint64_t base,
                          uint32_t* part, uint32_t* list, uint64_t* off, uint32_t* cap, uint32_t* st, hipStream_t s) {
  const uint32_t nb = seed_overflow_blocks(B);
  if (!nb) return;
  hipLaunchKernelGGL(k_ov_count, dim3(nb), dim3(OV_BLOCK), 0, s, nhits, dropped, B, cap1, part);
  hipLaunchKernelGGL(k_ov_scan, dim3(1), dim3(OV_BLOCK), 0, s, part, nb, st);
  hipLaunchKernelGGL(k_ov_write, dim3(nb), dim3(OV_BLOCK), 0, s, nhits, dropped, B, cap1, base, part, list, off, cap);
}

// The listed reads' hits (and flags), chunked or not, copied contiguously for the
// general block tier: read r's region taken from *next, its offset to off[r].
__global__ __launch_bounds__(256) void k_hits_gather(const uint64_t* hits, const uint8_t* hflag, const uint64_t* hoff,
                                                    const uint32_t* tab, const uint32_t* nhits, const uint32_t* list,
                                                    uint32_t nlist, uint64_t* dst, uint8_t* dflag, uint64_t* off,
                                                    unsigned long long* next) {
  __shared__ unsigned long long base;
  for (uint32_t li = blockIdx.x; li < nlist; li += gridDim.x) {
    const uint32_t r = list[li];
    const uint32_t n = nhits[r];
    if (threadIdx.x == 0) {
      base = atomicAdd(next, (unsigned long long)n);
      off[r] = base;
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < n; i += 256) {
      const uint64_t src = hit_at(tab, hoff, r, i);
      dst[base + i] = hits[src];
      if (hflag) dflag[base + i] = hflag[src];
    }
    __syncthreads();
  }
}
void launch_hits_gather(const uint64_t* hits, const uint8_t* hflag, const uint64_t* hoff, const uint32_t* tab,
                        const uint32_t* nhits, const uint32_t* list, uint32_t nlist, uint64_t* dst, uint8_t* dflag,
                        uint64_t* off, unsigned long long* next, hipStream_t s) {
  if (!nlist) return;
  const uint32_t blocks = nlist < 65536 ? nlist : 65536;
  hipLaunchKernelGGL(k_hits_gather, dim3(blocks), dim3(256), 0, s, hits, hflag, hoff, tab, nhits, list, nlist, dst,
                     dflag, off, next);
}
void launch_scatter_u64(uint64_t* out, const uint32_t* idx, const uint64_t* val, uint32_t n, hipStream_t s) {
  if (n) hipLaunchKernelGGL(k_scatter_u64, dim3((n + 255) / 256), dim3(256), 0, s, out, idx, val, n);
}

void launch_fill_pres(const uint32_t* koff, uint64_t* pres, uint32_t rot, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_pres, dim3(65536), dim3(256), 0, s, koff, pres, rot);
}


}  // namespace pbw
