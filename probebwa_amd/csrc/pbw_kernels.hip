// HIP kernels (gfx950) of the probebwa map path.
//
//   k_seed         seeds.rs:91-151 + hashtable.rs:90-124 — wavefront per read:
//                  15-mer codes in LDS, exact + 1-mismatch-neighbour probes of
//                  the dense code-indexed bucket table, coalesced bucket reads,
//                  similarity filter on the 2-bit genome, hits compacted in
//                  emission order with ballots.
//   k_cluster      candidates.rs:30-78 — wavefront per read: bitonic sort of
//                  (group, diagonal) keys in LDS, segmented merge, FxHashMap
//                  group order emulation, top-20 selection.
//   k_build_jobs   single.rs:105-133 — candidate windows -> DP jobs.
//   k_dp_band      smith_waterman.rs:163-255 — wavefront per job, one lane per
//                  band column, row state in registers, the deletion recurrence
//                  as an exact max-plus prefix scan across lanes, traceback
//                  bits in LDS, traceback + log-likelihood on lane 0.
//   k_dp_tiled     the same recurrence for any band width (64-column tiles,
//                  row state in global scratch): wide windows of align_at_window.
//   k_mapq         bayesian.rs:142-439 — FP64 select_best + finalize_mapq per read.
//
// All FP64 code is compiled with -ffp-contract=off: the reference (Rust)
// never fuses a*b+c, so neither may this file.
#include "pbw_internal.h"

#include <hip/hip_runtime.h>

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

__device__ __forceinline__ uint64_t lanemask_lt() { return (1ull << lane_id()) - 1ull; }

__device__ __forceinline__ int32_t shfl_up_i32(int32_t v, int d) { return __shfl_up(v, (unsigned)d, 64); }
__device__ __forceinline__ int32_t shfl_down_i32(int32_t v, int d) { return __shfl_down(v, (unsigned)d, 64); }
__device__ __forceinline__ uint32_t shfl_u32(uint32_t v, int src) { return (uint32_t)__shfl((int)v, src, 64); }
__device__ __forceinline__ uint64_t shfl_u64(uint64_t v, int src) {
  uint32_t lo = shfl_u32((uint32_t)v, src), hi = shfl_u32((uint32_t)(v >> 32), src);
  return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ int64_t shfl_xor_i64(int64_t v, int m) {
  int lo = __shfl_xor((int)(uint32_t)v, m, 64), hi = __shfl_xor((int)(uint32_t)((uint64_t)v >> 32), m, 64);
  return (int64_t)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x) {
  int l = lane_id();
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t t = (uint32_t)__shfl_up((int)x, (unsigned)o, 64);
    if (l >= o) x += t;
  }
  return x;
}
__device__ __forceinline__ int32_t wave_incl_max(int32_t x) {
  int l = lane_id();
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    int32_t t = shfl_up_i32(x, o);
    if (l >= o) x = max(x, t);
  }
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

// ---------------------------------------------------------------------------
// seed kernel: seeds.rs:91-151
// ---------------------------------------------------------------------------
constexpr int SEED_WPB = 4;

struct SeedLds {
  uint8_t seq[MAX_READ_LEN];
  uint32_t wcode[64], woff[64], wcnt[64], lbase[64];
  uint32_t r_incl[64], r_off[64];
  uint16_t fpF[64], fpR[64];
  uint8_t wflags[64];  // 1 valid, 2 probed, 4 fullF, 8 fullR, 16 exact-repetitive
  uint8_t nbrep[64];
  uint8_t r_w[64], r_nb[64];
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
// both sides (after_len + before_len == 12); otherwise the hit passes.
__device__ __forceinline__ bool seed_filter(const DevGenome& G, uint32_t contig, uint64_t pos, bool reverse,
                                            bool neighbour, uint8_t wflags, uint16_t fpF, uint16_t fpR) {
  bool rfull = reverse ? (wflags & 8) : (wflags & 4);
  if (!rfull) return true;
  uint64_t clen = G.clen[contig];
  if (pos < 4 || pos + 23 > clen) return true;
  uint64_t gb = G.coff[contig] + pos - 4;  // bases gb .. gb+26
  uint64_t w0 = gb >> 5;
  uint32_t s = (uint32_t)(gb & 31) * 2;
  uint64_t lo = G.g2b[w0], hi = G.g2b[w0 + 1];
  uint64_t bits = s ? ((lo >> s) | (hi << (64 - s))) : lo;
  uint64_t n0 = gb >> 6;
  uint32_t ns = (uint32_t)(gb & 63);
  uint64_t nlo = G.nmask[n0], nhi = G.nmask[n0 + 1];
  uint64_t nb = ns ? ((nlo >> ns) | (nhi << (64 - ns))) : nlo;
  int cnt[4] = {0, 0, 0, 0};
#pragma unroll
  for (int t = 0; t < 27; t++) {
    if (t >= 4 && t < 19) continue;
    if ((nb >> t) & 1ull) continue;
    cnt[(bits >> (2 * t)) & 3]++;
  }
  uint16_t rf = reverse ? fpR : fpF;
  int dist = 0;
#pragma unroll
  for (int c = 0; c < 4; c++) dist += abs((int)((rf >> (4 * c)) & 15) - cnt[c]);
  return dist <= (neighbour ? 8 : 10);
}

__global__ __launch_bounds__(256) void k_seed(SeedArgs a) {
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
    const uint64_t hbase = a.hoff[li];
    const uint32_t cap = a.hcap ? a.hcap[li] : (uint32_t)HIT_CAP;
    uint32_t hcount = 0;
    if (L < (uint32_t)K) {
      if (lane == 0) a.nhits[r] = 0;
      continue;
    }
    for (uint32_t x = lane; x < L; x += 64) S.seq[x] = a.R.seq[so + x];
    wave_sync();
    const uint32_t nwin = L - K + 1;
    for (uint32_t cb = 0; cb < nwin; cb += 64) {
      const uint32_t p = cb + lane;
      const bool act = p < nwin;
      uint32_t code = 0, off = 0, cnt = 0, nlook = 0;
      uint8_t fl = 0;
      uint16_t fpF = 0, fpR = 0;
      if (act) {
        bool valid = true;
        for (int t = 0; t < K; t++) {
          uint8_t b = S.seq[p + t];
          code = (code << 2) | base2bits(b);
          valid &= is_acgt(b);
        }
        if (valid) {
          fl |= 1;
          off = a.X.koff[code];  // hashtable.rs:90-102 lookup
          cnt = a.X.koff[code + 1] - off;
          if (cnt >= (uint32_t)REPEAT_SKIP) fl |= 16;
          else if (neighbour_probed(p, L)) fl |= 2;
          nlook = 1 + ((fl & 2) ? 45u : 0u);
        }
        // read-side fingerprints (seeds.rs:181-211) for complete flanks only
        if (p >= 4 && p + 23 <= L) {
          fl |= 4;
          uint32_t f = 0;
          for (uint32_t t = p + 15; t < p + 23; t++) f += fp_count(S.seq[t], false);
          for (uint32_t t = p - 4; t < p; t++) f += fp_count(S.seq[t], false);
          fpF = (uint16_t)f;
        }
        if (p >= 8 && p + 19 <= L) {  // rc(read): window start L-p-15
          fl |= 8;
          uint32_t f = 0;
          for (uint32_t t = p - 8; t < p; t++) f += fp_count(S.seq[t], true);
          for (uint32_t t = p + 15; t < p + 19; t++) f += fp_count(S.seq[t], true);
          fpR = (uint16_t)f;
        }
      }
      S.wcode[lane] = code;
      S.woff[lane] = off;
      S.wcnt[lane] = cnt;
      S.wflags[lane] = fl;
      S.fpF[lane] = fpF;
      S.fpR[lane] = fpR;
      S.nbrep[lane] = 0;
      uint32_t incl = wave_incl_scan(nlook);
      S.lbase[lane] = incl - nlook;
      const uint32_t T = shfl_u32(incl, 63);
      wave_sync();
      // lookups of this chunk in emission order: window, exact, then the 45
      // neighbours in (position, base) order (seeds.rs:107-147, :291-309)
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
            uint32_t wc = S.wcode[w];
            uint32_t i = (j - 1) / 3, alt = (j - 1) % 3;
            uint32_t sh = 2 * (14 - i);
            uint32_t orig = (wc >> sh) & 3u;
            uint32_t nbase = alt < orig ? alt : alt + 1;
            uint32_t ncode = (wc & ~(3u << sh)) | (nbase << sh);
            loff = a.X.koff[ncode];
            lcnt = a.X.koff[ncode + 1] - loff;
            if (lcnt >= (uint32_t)REPEAT_SKIP) S.nbrep[w] = 1;
          }
          c_look++;
          c_pos += lcnt;
        }
        uint32_t pincl = wave_incl_scan(lcnt);
        const uint32_t PT = shfl_u32(pincl, 63);
        S.r_incl[lane] = pincl;
        S.r_off[lane] = loff;
        S.r_w[lane] = (uint8_t)w;
        S.r_nb[lane] = (uint8_t)(j != 0);
        wave_sync();
        for (uint32_t pb = 0; pb < PT; pb += 64) {
          const uint32_t e = pb + lane;
          bool pass = false;
          uint64_t key = 0;
          if (e < PT) {
            int lo = 0, hi = 63;  // first l with r_incl[l] > e
            while (lo < hi) {
              int mid = (lo + hi) >> 1;
              if (S.r_incl[mid] > e) hi = mid;
              else lo = mid + 1;
            }
            const int l = lo;
            const uint32_t excl = l ? S.r_incl[l - 1] : 0u;
            const uint64_t packed = a.X.kpos[(uint64_t)S.r_off[l] + (e - excl)];
            const uint32_t contig = (uint32_t)((packed >> 40) & ((1ull << 23) - 1));
            const uint64_t pos = packed & ((1ull << 40) - 1);
            const bool rev = (packed >> 63) & 1ull;
            const uint32_t ww = S.r_w[l];
            pass = seed_filter(a.G, contig, pos, rev, S.r_nb[l] != 0, S.wflags[ww], S.fpF[ww], S.fpR[ww]);
            const int64_t rp = (int64_t)(cb + ww);
            const int64_t diag = rev ? (int64_t)pos + rp : (int64_t)pos - rp;  // candidates.rs:37-40
            key = ((uint64_t)(contig * 2u + (rev ? 1u : 0u)) << DIAG_BITS) | (uint64_t)(diag + DIAG_BIAS);
          }
          const uint64_t m = __ballot(pass);
          const uint32_t pre = (uint32_t)__popcll(m & lanemask_lt());
          if (pass && hcount + pre < cap) a.hits[hbase + hcount + pre] = key;
          hcount += (uint32_t)__popcll(m);
          c_filt += pass ? 1 : 0;
        }
        wave_sync();
      }
      if (act) {
        uint8_t rc = 0;
        if (fl & 16) rc = 1;
        else if (S.nbrep[lane]) rc = 2;
        a.rcode[so + p] = rc;
      }
      wave_sync();
    }
    if (lane == 0) a.nhits[r] = hcount;
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
  hipLaunchKernelGGL(k_seed, dim3(blocks), dim3(64 * SEED_WPB), 0, s, a);
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

// Scratch used by the (single-lane) group/cluster pass.  All pointers are
// generic: LDS for the wave tier, global memory for the block tier.
struct ClusterWork {
  uint32_t* g_key;    // group key (contig*2 + strand)
  uint32_t* g_min;    // group's first hit index (emission order)
  uint32_t* g_rank;   // iteration rank of the group in the emulated FxHashMap
  uint32_t* g_sort;   // (first hit << 16 | group) sorted
  int32_t* table;     // emulated SwissTable buckets
  int32_t* tmp;       // resize re-insertion list
  uint32_t table_cap; // capacity of table[]
};

// Lane-0 part 1: groups from the sorted keys.
__device__ uint32_t cluster_groups(const uint64_t* key, const uint32_t* idx, uint32_t n, ClusterWork& W) {
  uint32_t G = 0;
  uint32_t prev = 0xFFFFFFFFu;
  for (uint32_t e = 0; e < n; e++) {
    uint32_t gk = (uint32_t)(key[e] >> DIAG_BITS);
    if (gk != prev) {
      W.g_key[G] = gk;
      W.g_min[G] = idx[e];
      G++;
      prev = gk;
    } else if (idx[e] < W.g_min[G - 1]) {
      W.g_min[G - 1] = idx[e];
    }
  }
  return G;
}

// Lane-0 part 2: iteration rank of each group in the emulated FxHashMap
// (keys inserted in first-appearance order; iteration = ascending bucket).
__device__ void cluster_group_ranks(uint32_t G, ClusterWork& W) {
  uint32_t nb = 0, items = 0, growth_left = 0;
  for (uint32_t t = 0; t < G; t++) {
    uint32_t gi = W.g_sort[t] & 0xFFFFu;
    if (G > 0xFFFF) gi = W.g_sort[t];  // block tier stores plain indices
    uint64_t h = fx_hash_group(W.g_key[gi] >> 1, W.g_key[gi] & 1u);
    if (growth_left == 0) {
      uint32_t cap = max(items + 1, sw_capacity(nb) + 1);
      uint32_t nnb = sw_buckets_for(cap);
      uint32_t cnt = 0;
      for (uint32_t b = 0; b < nb; b++)
        if (W.table[b] >= 0) W.tmp[cnt++] = W.table[b];
      for (uint32_t b = 0; b < nnb; b++) W.table[b] = -1;
      nb = nnb;
      for (uint32_t c = 0; c < cnt; c++) {
        int32_t gj = W.tmp[c];
        uint64_t hj = fx_hash_group(W.g_key[gj] >> 1, W.g_key[gj] & 1u);
        W.table[sw_find_slot(W.table, nb, hj)] = gj;
      }
      growth_left = sw_capacity(nb) - items;
    }
    W.table[sw_find_slot(W.table, nb, h)] = (int32_t)gi;
    growth_left--;
    items++;
  }
  uint32_t rank = 0;
  for (uint32_t b = 0; b < nb; b++)
    if (W.table[b] >= 0) W.g_rank[W.table[b]] = rank++;
}

// Lane-0 part 3: clusters (sliding merge, candidates.rs:51-71) and the
// top-20 of the stable sort by score (:76): key (count desc, group rank, order).
__device__ uint32_t cluster_select(const uint64_t* key, uint32_t n, uint32_t L, const ClusterWork& W, Cand* out) {
  uint64_t top_k[TOP_CANDIDATES];
  Cand top_c[TOP_CANDIDATES];
  uint32_t nt = 0;
  uint32_t gi = 0, start = 0;
  int64_t sum = 0;
  uint32_t gk_prev = (uint32_t)(key[0] >> DIAG_BITS);
  int64_t d_prev = (int64_t)(key[0] & ((1ull << DIAG_BITS) - 1)) - DIAG_BIAS;
  sum = d_prev;
  for (uint32_t e = 1; e <= n; e++) {
    bool boundary;
    uint32_t gk = 0;
    int64_t d = 0;
    if (e == n) {
      boundary = true;
    } else {
      gk = (uint32_t)(key[e] >> DIAG_BITS);
      d = (int64_t)(key[e] & ((1ull << DIAG_BITS) - 1)) - DIAG_BIAS;
      boundary = gk != gk_prev || d - d_prev > MERGE_DISTANCE;
    }
    if (boundary) {
      uint32_t count = e - start;
      int64_t avg = sum / (int64_t)count;
      bool rev = gk_prev & 1u;
      int64_t position = rev ? avg - (int64_t)L + K : avg;
      if (position < 0) position = 0;
      uint64_t sk = ((uint64_t)(0x3FFFFFu - min(count, 0x3FFFFFu)) << 42) |
                    ((uint64_t)min(W.g_rank[gi], 0xFFFFFu) << 22) | (uint64_t)min(start, 0x3FFFFFu);
      if (nt < (uint32_t)TOP_CANDIDATES || sk < top_k[nt - 1]) {
        uint32_t pos = nt < (uint32_t)TOP_CANDIDATES ? nt : nt - 1;
        while (pos > 0 && top_k[pos - 1] > sk) {
          top_k[pos] = top_k[pos - 1];
          top_c[pos] = top_c[pos - 1];
          pos--;
        }
        top_k[pos] = sk;
        Cand c;
        c.pos = (uint64_t)position;
        c.contig = gk_prev >> 1;
        c.strand_score = (rev ? 0x80000000u : 0u) | min(count, 0x7FFFFFFFu);
        top_c[pos] = c;
        if (nt < (uint32_t)TOP_CANDIDATES) nt++;
      }
      if (e < n) {
        if (gk != gk_prev) gi++;
        start = e;
        sum = 0;
      }
    }
    if (e < n) {
      sum += d;
      gk_prev = gk;
      d_prev = d;
    }
  }
  for (uint32_t t = 0; t < nt; t++) out[t] = top_c[t];
  return nt;
}

struct ClusterLds {
  uint64_t key[CLUSTER_LDS_CAP];
  uint32_t idx[CLUSTER_LDS_CAP];
  uint32_t g_key[CLUSTER_LDS_CAP], g_min[CLUSTER_LDS_CAP], g_rank[CLUSTER_LDS_CAP], g_sort[CLUSTER_LDS_CAP];
  int32_t table[2048], tmp[CLUSTER_LDS_CAP];
  uint32_t G;
};

__device__ void wave_bitonic_u64(uint64_t* key, uint32_t* idx, uint32_t p2) {
  for (uint32_t k = 2; k <= p2; k <<= 1) {
    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
      for (uint32_t i = lane_id(); i < p2; i += 64) {
        uint32_t ixj = i ^ j;
        if (ixj > i) {
          bool up = (i & k) == 0;
          uint64_t x = key[i], y = key[ixj];
          if ((x > y) == up) {
            key[i] = y;
            key[ixj] = x;
            uint32_t t = idx[i];
            idx[i] = idx[ixj];
            idx[ixj] = t;
          }
        }
      }
      wave_sync();
    }
  }
}
__device__ void wave_bitonic_u32(uint32_t* key, uint32_t p2) {
  for (uint32_t k = 2; k <= p2; k <<= 1) {
    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
      for (uint32_t i = lane_id(); i < p2; i += 64) {
        uint32_t ixj = i ^ j;
        if (ixj > i) {
          bool up = (i & k) == 0;
          uint32_t x = key[i], y = key[ixj];
          if ((x > y) == up) {
            key[i] = y;
            key[ixj] = x;
          }
        }
      }
      wave_sync();
    }
  }
}

__global__ __launch_bounds__(64) void k_cluster(ClusterArgs a) {
  __shared__ ClusterLds S;
  const int lane = lane_id();
  for (uint32_t r = blockIdx.x; r < a.R.n; r += gridDim.x) {
    const uint32_t n = a.nhits[r];
    const uint32_t L = (uint32_t)(a.R.seq_off[r + 1] - a.R.seq_off[r]);
    if (n == 0) {
      if (lane == 0) a.ncand[r] = 0;
      continue;
    }
    if (n > (uint32_t)CLUSTER_LDS_CAP) {
      if (lane == 0) {
        uint32_t slot = atomicAdd(a.ntier2, 1u);
        a.tier2[slot] = r;
      }
      continue;
    }
    uint32_t p2 = 1;
    while (p2 < n) p2 <<= 1;
    const uint64_t* h = a.hits + a.hoff[r];
    for (uint32_t i = lane; i < p2; i += 64) {
      S.key[i] = i < n ? h[i] : ~0ull;
      S.idx[i] = i;
    }
    wave_sync();
    wave_bitonic_u64(S.key, S.idx, p2);
    ClusterWork W{S.g_key, S.g_min, S.g_rank, S.g_sort, S.table, S.tmp, 2048};
    if (lane == 0) S.G = cluster_groups(S.key, S.idx, n, W);
    wave_sync();
    const uint32_t G = S.G;
    uint32_t g2 = 1;
    while (g2 < G) g2 <<= 1;
    for (uint32_t i = lane; i < g2; i += 64) S.g_sort[i] = i < G ? ((S.g_min[i] << 16) | i) : 0xFFFFFFFFu;
    wave_sync();
    wave_bitonic_u32(S.g_sort, g2);
    if (lane == 0) {
      cluster_group_ranks(G, W);
      a.ncand[r] = (uint8_t)cluster_select(S.key, n, L, W, a.cands + (uint64_t)r * TOP_CANDIDATES);
    }
    wave_sync();
  }
}

// Block tier for reads with more hits than one wave sorts in LDS: bitonic
// sort in global scratch by a 256-thread block, then the same lane-0 passes.
__global__ __launch_bounds__(256) void k_cluster_tier2(ClusterArgs a, const uint32_t* list, uint32_t nlist) {
  uint64_t* base = a.scratch + (uint64_t)blockIdx.x * a.scratch_per_block;
  for (uint32_t li = blockIdx.x; li < nlist; li += gridDim.x) {
    const uint32_t r = list[li];
    const uint32_t n = a.nhits[r];
    const uint32_t L = (uint32_t)(a.R.seq_off[r + 1] - a.R.seq_off[r]);
    uint32_t p2 = 1;
    while (p2 < n) p2 <<= 1;
    uint64_t* key = base;
    uint32_t* idx = (uint32_t*)(base + p2);
    uint32_t* g_key = idx + p2;
    uint32_t* g_min = g_key + p2;
    uint32_t* g_rank = g_min + p2;
    uint32_t* g_sort = g_rank + p2;
    int32_t* tmp = (int32_t*)(g_sort + p2);
    int32_t* table = tmp + p2;  // up to 4*p2 buckets
    const uint64_t* h = a.hits + a.hoff[r];
    for (uint32_t i = threadIdx.x; i < p2; i += blockDim.x) {
      key[i] = i < n ? h[i] : ~0ull;
      idx[i] = i;
    }
    __syncthreads();
    for (uint32_t k = 2; k <= p2; k <<= 1) {
      for (uint32_t j = k >> 1; j > 0; j >>= 1) {
        for (uint32_t i = threadIdx.x; i < p2; i += blockDim.x) {
          uint32_t ixj = i ^ j;
          if (ixj > i) {
            bool up = (i & k) == 0;
            uint64_t x = key[i], y = key[ixj];
            if ((x > y) == up) {
              key[i] = y;
              key[ixj] = x;
              uint32_t t = idx[i];
              idx[i] = idx[ixj];
              idx[ixj] = t;
            }
          }
        }
        __syncthreads();
      }
    }
    __shared__ uint32_t Gs;
    ClusterWork W{g_key, g_min, g_rank, g_sort, table, tmp, 4 * p2};
    if (threadIdx.x == 0) Gs = cluster_groups(key, idx, n, W);
    __syncthreads();
    const uint32_t G = Gs;
    uint32_t g2 = 1;
    while (g2 < G) g2 <<= 1;
    // sort groups by first appearance: keys (g_min, index) as u64 in key[] (hits no longer needed
    // after cluster_groups only if we re-read them; keep hits: use g_sort pairs with a stable
    // O(G log G) bitonic over 64-bit packed keys stored in g_rank/g_sort temporarily)
    uint64_t* gk64 = (uint64_t*)table;  // table area is free until cluster_group_ranks
    for (uint32_t i = threadIdx.x; i < g2; i += blockDim.x) gk64[i] = i < G ? (((uint64_t)g_min[i] << 32) | i) : ~0ull;
    __syncthreads();
    for (uint32_t k = 2; k <= g2; k <<= 1) {
      for (uint32_t j = k >> 1; j > 0; j >>= 1) {
        for (uint32_t i = threadIdx.x; i < g2; i += blockDim.x) {
          uint32_t ixj = i ^ j;
          if (ixj > i) {
            bool up = (i & k) == 0;
            uint64_t x = gk64[i], y = gk64[ixj];
            if ((x > y) == up) {
              gk64[i] = y;
              gk64[ixj] = x;
            }
          }
        }
        __syncthreads();
      }
    }
    for (uint32_t i = threadIdx.x; i < G; i += blockDim.x) g_sort[i] = (uint32_t)(gk64[i] & 0xFFFFFFFFu);
    __syncthreads();
    if (threadIdx.x == 0) {
      if (G <= 0xFFFF)
        for (uint32_t i = 0; i < G; i++) g_sort[i] = (g_sort[i] & 0xFFFFu);
      cluster_group_ranks(G, W);
      a.ncand[r] = (uint8_t)cluster_select(key, n, L, W, a.cands + (uint64_t)r * TOP_CANDIDATES);
    }
    __syncthreads();
  }
}

void launch_cluster(const ClusterArgs& a, hipStream_t s) {
  if (a.R.n == 0) return;
  uint32_t blocks = a.R.n < 65536 ? a.R.n : 65536;
  hipLaunchKernelGGL(k_cluster, dim3(blocks), dim3(64), 0, s, a);
}
void launch_cluster_tier2(const ClusterArgs& a, const uint32_t* list, uint32_t n, hipStream_t s) {
  if (n == 0) return;
  uint32_t blocks = n < 64 ? n : 64;
  hipLaunchKernelGGL(k_cluster_tier2, dim3(blocks), dim3(256), 0, s, a, list, n);
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
  wave_sync();
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
  wave_sync();
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
  wave_sync();
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

// Copies the CIGAR (forward order) and fills DpOut; lane-parallel copy.
__device__ void dp_emit(const DpArgs& a, uint32_t jid, const DpStage& st, uint32_t nruns, uint32_t off, uint32_t span,
                        int32_t score, double lk, uint64_t cells) {
  const int lane = lane_id();
  uint32_t ncopy = nruns < a.cigar_cap ? nruns : a.cigar_cap;
  for (uint32_t t = lane; t < ncopy; t += 64) a.cigar[(uint64_t)jid * a.cigar_cap + t] = st.runs[nruns - 1 - t];
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
  uint8_t* base = dyn + (size_t)wib * dp_band_lds_bytes(a.max_m, a.max_n);
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
      if (lane == 0) a.needs_generic[jid] = 1;
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
    wave_sync();
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
      const int32_t Mup = shfl_down_i32(Mp, 1), Iup = shfl_down_i32(Ip, 1);
      const int32_t io = Mup + go, ie = Iup + ge;
      int32_t I = max(io, ie);
      if (!valid) M = NEG_INF;
      int32_t Ml = shfl_up_i32(M, 1);
      if (lane == 0) Ml = NEG_INF;
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
    wave_sync();
    dp_emit(a, jid, st, nruns, off, span, best_score, lk, cells);
    c_cells += cells;
    wave_sync();
  }
  if (lane == 0 && a.ctr) atomicAdd(&a.ctr->cells, c_cells);
}

void launch_dp_band(const DpArgs& a0, hipStream_t s) {
  if (a0.njobs == 0) return;
  DpArgs a = a0;
  uint32_t per_wave = dp_band_lds_bytes(a.max_m, a.max_n);
  a.wpb = 65536 / per_wave;
  if (a.wpb > DP_WPB) a.wpb = DP_WPB;
  if (a.wpb < 1) a.wpb = 1;
  uint32_t blocks = (a.njobs + a.wpb - 1) / a.wpb;
  if (blocks > 65536) blocks = 65536;
  size_t lds = (size_t)a.wpb * per_wave;
  hipLaunchKernelGGL(k_dp_band, dim3(blocks), dim3(64 * a.wpb), lds, s, a);
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
  DpStage st;
  st.q = dyn;
  st.qq = st.q + ((a.max_m + 15) & ~15u);
  st.ref = st.qq + ((a.max_m + 15) & ~15u);
  st.go = (int16_t*)(st.ref + ((a.max_n + 15) & ~15u));
  st.runs = (uint32_t*)((uint8_t*)st.go + (((a.max_n * 2) + 15) & ~15u));
  int32_t* RM = a.rows + (uint64_t)blockIdx.x * 3 * (a.max_n + 1);
  int32_t* RI = RM + (a.max_n + 1);
  int32_t* RD = RI + (a.max_n + 1);
  uint64_t* TB = a.tb + (uint64_t)blockIdx.x * (uint64_t)(a.max_m + 1) * a.tiles * 4;
  unsigned long long c_cells = 0;
  for (uint32_t t = blockIdx.x; t < a.njobs; t += gridDim.x) {
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
      if (lane == 0) a.needs_generic[jid] = 2;  // host reports PBW_ERR_LIMIT
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
        int32_t dM = shfl_up_i32(uM, 1), dI = shfl_up_i32(uI, 1), dD = shfl_up_i32(uD, 1);
        if (lane == 0) {
          dM = cM;
          dI = cI;
          dD = cD;
        }
        cM = __shfl(uM, 63, 64);
        cI = __shfl(uI, 63, 64);
        cD = __shfl(uD, 63, 64);
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
        int32_t Ml = shfl_up_i32(M, 1);
        if (lane == 0) Ml = curMl;
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
        curMl = __shfl(M, lastl, 64);
        curDl = __shfl(D, lastl, 64);
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
    wave_sync();
    dp_emit(a, jid, st, nruns, off, span, best_score, lk, cells);
    c_cells += cells;
    wave_sync();
  }
  if (lane == 0 && a.ctr) atomicAdd(&a.ctr->cells, c_cells);
}

void launch_dp_tiled(const DpArgs& a, uint32_t nwaves, hipStream_t s) {
  if (a.njobs == 0) return;
  uint32_t blocks = nwaves < a.njobs ? nwaves : a.njobs;
  size_t lds = dp_tiled_lds_bytes(a.max_m, a.max_n);
  hipLaunchKernelGGL(k_dp_tiled, dim3(blocks), dim3(64), lds, s, a);
}

// ---------------------------------------------------------------------------
// MAPQ: bayesian.rs:193-439 (thread per read, FP64)
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

// bayesian.rs:306-373 log_entropy; ref_window = genome[gb, gb+span)
__device__ double log_entropy(const DevGenome& G, const DevLuts& T, const uint32_t* cig, uint32_t ncig,
                              const QueryView& Q, uint64_t gb, uint32_t span) {
  int64_t n = 0, mm = 0;
  uint32_t readlen = 0, refpos = 0;
  int64_t indel_runs = 0, deletion_runs = 0;
  uint32_t curtype = 0;
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
        uint8_t aa = Q.base(readlen), rr = genome_base(G, gb + refpos);
        if (aa != rr && aa != 'N' && rr != 'N' && hq) mm++;
        readlen++;
        refpos++;
      }
    }
  }
  int64_t insertion_runs = indel_runs - deletion_runs;
  // gc_content (bayesian.rs:445-456) of the ref window
  uint32_t gc = 0, total = 0;
  for (uint32_t x = 0; x < span; x++) {
    uint8_t bb = genome_base(G, gb + x);
    if (bb == 'G' || bb == 'C') {
      gc++;
      total++;
    } else if (bb == 'A' || bb == 'T') {
      total++;
    }
  }
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

// bayesian.rs:405-439 missing_locus_probability; repeat mask from the seed
// pass (0: 0.0, 1: 1.0, 2: 10^(-q/10) at the window start, seeds.rs:125-139)
__device__ double missing_locus(const DevLuts& T, const uint8_t* qual, uint32_t qlen, const uint8_t* rcode,
                                uint32_t nwin) {
  const uint32_t k = K;
  if (qlen < k) return 0.5;
  double offset_means = 0.0;
  for (uint32_t phase = 0; phase < 5; phase++) {
    double log_p = 0.0;
    for (uint32_t start = phase; start + k <= qlen; start += k) {
      double rep = 0.0;
      if (start < nwin) {
        uint8_t c = rcode[start];
        if (c == 1) rep = 1.0;
        else if (c == 2) rep = T.nrp[start < qlen ? qual[start] : 30];
      }
      double z = 1.0;
      for (uint32_t t = start; t < start + k; t++) z = z * T.omq[qual[t]];
      double p1 = -0.0;
      for (uint32_t t = start; t < start + k; t++) p1 = p1 + T.qratio[qual[t]];
      double p2 = fmin(fmax(1.0 - z * (1.0 + p1), 0.0), 1.0);
      double pm = fmin(fmax(1.0 - (1.0 - p2) * (1.0 - rep), 1e-300), 1.0);
      log_p += log(pm);
    }
    offset_means += exp(log_p);
  }
  return fmin(offset_means / 5.0, 0.99);
}

// bayesian.rs:251-282 finalize_mapq
__device__ void finalize_mapq(const DevGenome& G, const DevLuts& T, double sum_exp, const uint8_t* read_qual,
                              uint32_t read_qlen, const uint8_t* rcode, uint32_t nwin, const uint32_t* cig,
                              uint32_t ncig, const QueryView& Q, uint64_t gb, uint32_t span, uint8_t& mapq,
                              double& posterior) {
  double bp = 1.0 / fmax(sum_exp, 1.0);
  double pw = fmax(1.0 - bp, T.posterior_floor);
  double wrongmapping = py_int_round(log(pw) / DB);
  double entropy = log_entropy(G, T, cig, ncig, Q, gb, span) - 3.0;
  if (entropy <= 0.0) {
    mapq = 0;
    posterior = 0.0;
    return;
  }
  double p_missing = fmax(missing_locus(T, read_qual, read_qlen, rcode, nwin), 1e-30);
  double notfound = -10.0 * log10(p_missing);
  double combined = logadd(wrongmapping, notfound);
  double pp = fmin(fmax(logadd(entropy, combined), 0.0), 99.0);
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

__global__ void k_mapq(MapqArgs a) {
  uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= a.R.n) return;
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
  double mx = a.out[s0].log_lk;
  uint32_t bi = 0;
  for (uint32_t k = 0; k < nj; k++) {  // select_best: last maximum
    double v = a.out[s0 + k].log_lk;
    if (v >= mx) {
      mx = v;
      bi = k;
    }
  }
  double se = -0.0;
  for (uint32_t k = 0; k < nj; k++) se += exp(a.out[s0 + k].log_lk - mx);
  const DpJob job = a.jobs[s0 + bi];
  const DpOut dout = a.out[s0 + bi];
  QueryView Q{a.R.seq + so, a.R.qual + qo, L, qlen, job.rc != 0};
  const uint32_t* cig = a.cigar + (s0 + bi) * a.cigar_cap;
  uint8_t mapq;
  double post;
  finalize_mapq(a.G, a.T, se, a.R.qual + qo, qlen, a.rcode + so, L >= (uint32_t)K ? L - K + 1 : 0, cig,
                dout.cigar_n, Q, job.ref_gb + dout.ref_off, dout.span, mapq, post);
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
  if (a.R.n == 0) return;
  hipLaunchKernelGGL(k_mapq, dim3((a.R.n + 127) / 128), dim3(128), 0, s, a);
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
__global__ void k_fill_koff(uint32_t* koff, const uint32_t* keys, const uint32_t* kstart, uint64_t nkeys,
                            uint32_t total) {
  uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; c <= CODE_SPACE; c += stride) {
    uint64_t lo = 0, hi = nkeys;  // first key >= c
    while (lo < hi) {
      uint64_t mid = (lo + hi) >> 1;
      if ((uint64_t)keys[mid] < c) lo = mid + 1;
      else hi = mid;
    }
    koff[c] = lo < nkeys ? kstart[lo] : total;
  }
}
void launch_fill_koff(uint32_t* koff, const uint32_t* keys, const uint32_t* kstart, uint64_t nkeys, uint32_t total,
                      hipStream_t s) {
  hipLaunchKernelGGL(k_fill_koff, dim3(65536), dim3(256), 0, s, koff, keys, kstart, nkeys, total);
}

}  // namespace pbw
