// Bucket table built on the device: HashTableBuilder::build
// (src/index/hashtable.rs:248-322) with k = 15, stride 5 and the cap of 300
// entries per 15-mer applied in scan order.
//
// Scan order is the reference's: contigs ascending; within a contig the
// forward windows i = 0, 5, 10, ... (hashtable.rs:266-275), then the windows
// i = 0, 5, ... of the reverse-complement string, stored at position
// len - i - 15 on the reverse strand (hashtable.rs:279-291).  Windows holding
// a non-ACGT base are not indexed (is_acgt_only, hashtable.rs:147-149).
//
// Steps, all on the library's stream:
//   1. one thread per scan slot j (global scan index) writes (code, j); an
//      N window writes the sentinel code 2^30 so that it sorts last;
//   2. a stable LSD radix sort of (code, j) by the low 31 bits of the code:
//      within a code, entries stay in scan order (the reference's stable
//      sort_by_key, hashtable.rs:297, over hashes; mix() is a bijection so
//      code order and hash order group identically);
//   3. every sorted entry's run start (a max-scan of run-start marks) and,
//      at each run's end, its capped length min(run, 300) into koff[code]: the
//      reference stops pushing a k-mer's positions once 300 are kept
//      (hashtable.rs:270-273, 286-289), i.e. it keeps the first 300 of the scan;
//   4. an in-place exclusive scan of koff over all 4^15 codes gives the
//      bucket starts (bucket c = [koff[c], koff[c+1]));
//   5. the first 300 entries of each run are decoded from j and placed, each
//      with the genome flank fingerprint the seed filter compares against.
// Temporaries are sized by the number of windows, not by the code space.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <string>
#include <vector>

#include "pbw_internal.h"

namespace pbw {

namespace {

constexpr uint32_t IDX_SENTINEL = 1u << 30;

// Contig of global scan slot j: the last c with cbase[c] <= j.
__device__ __forceinline__ int32_t slot_contig(const uint64_t* cbase, int32_t nctg, uint64_t j) {
  int32_t lo = 0, hi = nctg - 1;
  while (lo < hi) {
    int32_t mid = (lo + hi + 1) >> 1;
    if (cbase[mid] <= j) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Window position and strand of slot j of contig c (length L >= 15).
__device__ __forceinline__ void slot_window(uint64_t t, uint64_t L, uint64_t& pos, uint32_t& strand) {
  const uint64_t nf = (L - K) / STRIDE + 1;  // windows i = 0, 5, ..., <= L - 15 on each strand
  if (t < nf) {
    pos = t * STRIDE;
    strand = 0;
  } else {
    pos = L - (t - nf) * STRIDE - K;  // hashtable.rs:284 (len - i - k)
    strand = 1;
  }
}

__global__ void k_idx_windows(const uint64_t* __restrict__ g2b, const uint64_t* __restrict__ nmask,
                              const uint64_t* __restrict__ coff, const uint64_t* __restrict__ clen,
                              const uint64_t* __restrict__ cbase, int32_t nctg, uint64_t n_total,
                              uint32_t* __restrict__ key, uint32_t* __restrict__ val) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n_total; j += stride) {
    const int32_t c = slot_contig(cbase, nctg, j);
    uint64_t pos;
    uint32_t strand;
    slot_window(j - cbase[c], clen[c], pos, strand);
    const uint64_t gb = coff[c] + pos;
    // 15 bases, LSB first: base k of the window at bits 2k..2k+1
    const uint64_t w = gb >> 5;
    const uint32_t sh = (uint32_t)(gb & 31) * 2;
    uint64_t bits = g2b[w] >> sh;
    if (sh) bits |= g2b[w + 1] << (64 - sh);
    const uint32_t lsb = (uint32_t)bits & 0x3FFFFFFFu;
    const uint64_t nw = gb >> 6;
    const uint32_t nsh = (uint32_t)(gb & 63);
    uint64_t nb = nmask[nw] >> nsh;
    if (nsh > 64 - K) nb |= nmask[nw + 1] << (64 - nsh);
    uint32_t code;
    if (nb & ((1ull << K) - 1)) {
      code = IDX_SENTINEL;
    } else if (strand == 0) {
      // MSB-first packing (hash_kmer, hashtable.rs:107-113): reverse the order of the 2-bit fields
      uint32_t r = __builtin_bitreverse32(lsb) >> 2;           // fields reversed, bits swapped within each
      code = ((r & 0x15555555u) << 1) | ((r >> 1) & 0x15555555u);
    } else {
      // reverse complement, MSB first: its base k is the complement of window base 14 - k,
      // which lands at bits 2(14 - k) -- the LSB-first field of that base, complemented
      code = ~lsb & 0x3FFFFFFFu;
    }
    key[j] = code;
    val[j] = (uint32_t)j;
  }
}

// Run start of every sorted entry, step 1: i at a run's first entry, else 0
// (an inclusive max-scan then carries each run's start along it).
__global__ void k_idx_startmark(const uint32_t* __restrict__ key, uint64_t n, uint32_t* __restrict__ mark) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    mark[i] = (i == 0 || key[i - 1] != key[i]) ? (uint32_t)i : 0u;
}

// Capped run length (min(run, 300)) into koff[code] at each run's last entry.
__global__ void k_idx_counts(const uint32_t* __restrict__ key, uint64_t n, const uint32_t* __restrict__ rstart,
                             uint32_t* __restrict__ koff) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const uint32_t k = key[i];
    if (k == IDX_SENTINEL) continue;
    if (i + 1 == n || key[i + 1] != k) {
      const uint64_t run = i + 1 - rstart[i];
      koff[k] = (uint32_t)(run < (uint64_t)MAX_COUNT ? run : (uint64_t)MAX_COUNT);
    }
  }
}

// Genome side of the seed similarity filter (seeds.rs:162-219) for a window at
// `pos` of contig c: the A/C/G/T counts (4 bits each, A lowest) of the 4 bases
// before and the 8 after the 15-mer, N not counted; GENOME_FP_OPEN when those
// flanks run past the contig (the filter then passes every hit).
__device__ __forceinline__ uint16_t genome_flank_fp(const uint64_t* g2b, const uint64_t* nmask, uint64_t coff,
                                                    uint64_t clen, uint64_t pos) {
  if (pos < 4 || pos + 23 > clen) return GENOME_FP_OPEN;
  const uint64_t gb = coff + pos - 4;  // bases gb .. gb + 26
  const uint64_t w0 = gb >> 5;
  const uint32_t s = (uint32_t)(gb & 31) * 2;
  const uint64_t bits = s ? ((g2b[w0] >> s) | (g2b[w0 + 1] << (64 - s))) : g2b[w0];
  const uint64_t n0 = gb >> 6;
  const uint32_t ns = (uint32_t)(gb & 63);
  const uint64_t nb = ns ? ((nmask[n0] >> ns) | (nmask[n0 + 1] << (64 - ns))) : nmask[n0];
  uint32_t fp = 0;
#pragma unroll
  for (int t = 0; t < 27; t++) {
    if (t >= 4 && t < 19) continue;
    if ((nb >> t) & 1ull) continue;
    fp += 1u << (4 * ((bits >> (2 * t)) & 3));
  }
  return (uint16_t)fp;
}

// The first MAX_COUNT entries of every run, decoded to packed positions
// (pack_position, hashtable.rs:157-164: bit 63 strand, 62..40 contig, 39..0 position)
// with each entry's genome flank fingerprint beside it, or (kfp null) to fused
// entries that hold the fingerprint.
__global__ void k_idx_place(const uint32_t* __restrict__ key, const uint32_t* __restrict__ val, uint64_t n,
                            const uint32_t* __restrict__ rstart, const uint32_t* __restrict__ koff,
                            const uint64_t* __restrict__ g2b, const uint64_t* __restrict__ nmask,
                            const uint64_t* __restrict__ coff, const uint64_t* __restrict__ clen,
                            const uint64_t* __restrict__ cbase, int32_t nctg, uint64_t* __restrict__ kpos,
                            uint16_t* __restrict__ kfp) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const uint32_t k = key[i];
    if (k == IDX_SENTINEL) continue;
    const uint64_t r = i - rstart[i];
    if (r >= (uint64_t)MAX_COUNT) continue;
    const uint64_t j = val[i];
    const int32_t c = slot_contig(cbase, nctg, j);
    uint64_t pos;
    uint32_t strand;
    slot_window(j - cbase[c], clen[c], pos, strand);
    const uint64_t at = (uint64_t)koff[k] + r;
    const uint16_t fp = genome_flank_fp(g2b, nmask, coff[c], clen[c], pos);
    if (kfp) {
      kpos[at] = ((uint64_t)strand << 63) | ((uint64_t)c << 40) | pos;
      kfp[at] = fp;
    } else {
      kpos[at] = fused_entry(strand, (uint64_t)c, fp, pos);
    }
  }
}

// In-place exclusive sum of x[0, n) in three passes: tile sums, one block's
// scan of them, and each tile's scan from its base.  A tile is read and written
// only by its own block, so the scan is safe in place.
constexpr uint32_t SCAN_T = 256, SCAN_E = 16, SCAN_TILE = SCAN_T * SCAN_E;

__global__ __launch_bounds__(SCAN_T) void k_scan_tiles(const uint32_t* __restrict__ x, uint64_t n,
                                                       uint32_t* __restrict__ part) {
  using BR = hipcub::BlockReduce<uint32_t, SCAN_T>;
  __shared__ typename BR::TempStorage tmp;
  const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE;
  uint32_t s = 0;
  for (uint32_t e = 0; e < SCAN_E; e++) {
    const uint64_t i = base + (uint64_t)e * SCAN_T + threadIdx.x;
    if (i < n) s += x[i];
  }
  s = BR(tmp).Sum(s);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ __launch_bounds__(1024) void k_scan_parts(uint32_t* part, uint32_t np) {
  using BS = hipcub::BlockScan<uint32_t, 1024>;
  __shared__ typename BS::TempStorage tmp;
  __shared__ uint32_t carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (uint32_t b0 = 0; b0 < np; b0 += 1024) {
    const uint32_t i = b0 + threadIdx.x;
    const uint32_t v = i < np ? part[i] : 0u;
    uint32_t ex, tot;
    BS(tmp).ExclusiveSum(v, ex, tot);
    if (i < np) part[i] = carry + ex;
    __syncthreads();
    if (threadIdx.x == 0) carry += tot;
    __syncthreads();
  }
}

__global__ __launch_bounds__(SCAN_T) void k_scan_apply(uint32_t* __restrict__ x, uint64_t n,
                                                       const uint32_t* __restrict__ part) {
  using BL = hipcub::BlockLoad<uint32_t, SCAN_T, SCAN_E, hipcub::BLOCK_LOAD_TRANSPOSE>;
  using BSt = hipcub::BlockStore<uint32_t, SCAN_T, SCAN_E, hipcub::BLOCK_STORE_TRANSPOSE>;
  using BS = hipcub::BlockScan<uint32_t, SCAN_T>;
  __shared__ union {
    typename BL::TempStorage load;
    typename BSt::TempStorage store;
    typename BS::TempStorage scan;
  } tmp;
  const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE;
  const uint32_t valid = (uint32_t)min<uint64_t>(SCAN_TILE, n - base);
  uint32_t v[SCAN_E];
  BL(tmp.load).Load(x + base, v, valid, 0u);
  __syncthreads();
  BS(tmp.scan).ExclusiveSum(v, v);
  __syncthreads();
  const uint32_t b = part[blockIdx.x];
  for (uint32_t e = 0; e < SCAN_E; e++) v[e] += b;
  BSt(tmp.store).Store(x + base, v, valid);
}

__global__ __launch_bounds__(SCAN_T) void k_max_tiles(const uint32_t* __restrict__ x, uint64_t n,
                                                      uint32_t* __restrict__ part) {
  using BR = hipcub::BlockReduce<uint32_t, SCAN_T>;
  __shared__ typename BR::TempStorage tmp;
  const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE;
  uint32_t s = 0;
  for (uint32_t e = 0; e < SCAN_E; e++) {
    const uint64_t i = base + (uint64_t)e * SCAN_T + threadIdx.x;
    if (i < n) s = max(s, x[i]);
  }
  s = BR(tmp).Reduce(s, hipcub::Max());
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// Inclusive max-scan of the tile maxima, shifted by one: part[b] = max of tiles < b.
__global__ __launch_bounds__(1024) void k_max_parts(uint32_t* part, uint32_t np) {
  using BS = hipcub::BlockScan<uint32_t, 1024>;
  __shared__ typename BS::TempStorage tmp;
  __shared__ uint32_t carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (uint32_t b0 = 0; b0 < np; b0 += 1024) {
    const uint32_t i = b0 + threadIdx.x;
    const uint32_t v = i < np ? part[i] : 0u;
    uint32_t ex, tot;
    BS(tmp).ExclusiveScan(v, ex, 0u, hipcub::Max(), tot);
    if (i < np) part[i] = max(carry, ex);
    __syncthreads();
    if (threadIdx.x == 0) carry = max(carry, tot);
    __syncthreads();
  }
}

__global__ __launch_bounds__(SCAN_T) void k_max_apply(uint32_t* __restrict__ x, uint64_t n,
                                                      const uint32_t* __restrict__ part) {
  using BL = hipcub::BlockLoad<uint32_t, SCAN_T, SCAN_E, hipcub::BLOCK_LOAD_TRANSPOSE>;
  using BSt = hipcub::BlockStore<uint32_t, SCAN_T, SCAN_E, hipcub::BLOCK_STORE_TRANSPOSE>;
  using BS = hipcub::BlockScan<uint32_t, SCAN_T>;
  __shared__ union {
    typename BL::TempStorage load;
    typename BSt::TempStorage store;
    typename BS::TempStorage scan;
  } tmp;
  const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE;
  const uint32_t valid = (uint32_t)min<uint64_t>(SCAN_TILE, n - base);
  uint32_t v[SCAN_E];
  BL(tmp.load).Load(x + base, v, valid, 0u);
  __syncthreads();
  BS(tmp.scan).InclusiveScan(v, v, hipcub::Max());
  __syncthreads();
  const uint32_t b = part[blockIdx.x];
  for (uint32_t e = 0; e < SCAN_E; e++) v[e] = max(v[e], b);
  BSt(tmp.store).Store(x + base, v, valid);
}

// Buckets of a loaded table: the capped count of each key's code into koff
// (counts first, scanned into starts afterwards).  bad[0] counts keys that are
// not the mix of a 15-mer code and bucket ranges outside the position store.
__global__ void k_tab_counts(const uint32_t* __restrict__ key, const uint32_t* __restrict__ off,
                             const uint32_t* __restrict__ cnt, uint64_t n, uint64_t npos, uint32_t* __restrict__ koff,
                             uint32_t* __restrict__ bad) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const uint32_t code = kmer_unmix(key[i]);
    if (code >= CODE_SPACE || (uint64_t)off[i] + cnt[i] > npos) {
      atomicAdd(bad, 1u);
      continue;
    }
    // a second bucket of the same code (a malformed map) would overrun the first
    if (atomicExch(&koff[code], cnt[i] | 0x80000000u) != 0u) atomicAdd(bad, 1u);
  }
}

__global__ void k_tab_strip(uint32_t* __restrict__ koff, uint64_t n) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) koff[i] &= 0x7FFFFFFFu;
}

// Each bucket's positions copied to its place in code order, with the genome
// flank fingerprint of every entry; bad[1] counts positions whose contig or
// window does not exist in the genome (nothing is written for them).
__global__ void k_tab_place(const uint32_t* __restrict__ key, const uint32_t* __restrict__ off,
                            const uint32_t* __restrict__ cnt, uint64_t n, const uint64_t* __restrict__ src,
                            const uint32_t* __restrict__ koff, const uint64_t* __restrict__ g2b,
                            const uint64_t* __restrict__ nmask, const uint64_t* __restrict__ coff,
                            const uint64_t* __restrict__ clen, int32_t nctg, uint64_t* __restrict__ kpos,
                            uint16_t* __restrict__ kfp, uint32_t* __restrict__ bad) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const uint32_t code = kmer_unmix(key[i]);
    if (code >= CODE_SPACE) continue;
    const uint64_t at = koff[code];
    for (uint32_t j = 0; j < cnt[i]; j++) {
      const uint64_t p = src[(uint64_t)off[i] + j];
      const uint64_t c = (p >> 40) & ((1ull << 23) - 1), pos = p & ((1ull << 40) - 1);
      if (c >= (uint64_t)nctg || pos + K > clen[c]) {
        atomicAdd(bad + 1, 1u);
        kpos[at + j] = 0;
        if (kfp) kfp[at + j] = GENOME_FP_OPEN;
        continue;
      }
      const uint16_t fp = genome_flank_fp(g2b, nmask, coff[c], clen[c], pos);
      if (kfp) {
        kpos[at + j] = p;
        kfp[at + j] = fp;
      } else {
        kpos[at + j] = fused_entry(p >> 63, c, fp, pos);
      }
    }
  }
}

struct Tmp {
  void* p = nullptr;
  ~Tmp() {
    if (p) (void)hipFree(p);
  }
};

}  // namespace

#define IDX_CHECK(x)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (x);                                                               \
    if (e_ != hipSuccess) {                                                            \
      err = std::string(#x " failed: ") + hipGetErrorString(e_);                       \
      return e_;                                                                       \
    }                                                                                  \
  } while (0)

hipError_t gpu_build_index(const DevGenome& G, const uint64_t* h_clen, uint32_t* koff, uint64_t** kpos_out,
                           uint16_t** kfp_out, uint64_t* n_entries, bool fused, hipStream_t s, std::string& err) {
  *kpos_out = nullptr;
  *kfp_out = nullptr;
  *n_entries = 0;
  const int32_t nctg = G.ncontigs;
  std::vector<uint64_t> cbase(nctg + 1, 0);
  for (int32_t c = 0; c < nctg; c++) {
    const uint64_t L = h_clen[c];
    const uint64_t per = L >= (uint64_t)K ? 2 * ((L - K) / STRIDE + 1) : 0;
    cbase[c + 1] = cbase[c] + per;
  }
  const uint64_t n = cbase[nctg];
  if (n >= (1ull << 32)) {
    err = "more than 2^32 indexed windows";
    return hipErrorInvalidValue;
  }
  // koff first holds the capped run length of every code (0 for absent codes)
  IDX_CHECK(hipMemsetAsync(koff, 0, (CODE_SPACE + 1) * sizeof(uint32_t), s));
  const uint64_t nk = CODE_SPACE + 1;
  const uint32_t np_k = (uint32_t)((nk + SCAN_TILE - 1) / SCAN_TILE);
  const uint32_t np_n = (uint32_t)((n + SCAN_TILE - 1) / SCAN_TILE);
  Tmp t_cbase, t_k0, t_k1, t_v0, t_v1, t_sort, t_part;
  IDX_CHECK(hipMalloc(&t_part.p, (size_t)std::max(np_k, np_n) * sizeof(uint32_t)));
  uint32_t* part = (uint32_t*)t_part.p;
  uint64_t* kpos = nullptr;
  uint16_t* kfp = nullptr;
  uint32_t total = 0;
  if (n) {
    IDX_CHECK(hipMalloc(&t_cbase.p, (nctg + 1) * sizeof(uint64_t)));
    IDX_CHECK(hipMemcpyAsync(t_cbase.p, cbase.data(), (nctg + 1) * sizeof(uint64_t), hipMemcpyHostToDevice, s));
    IDX_CHECK(hipMalloc(&t_k0.p, n * sizeof(uint32_t)));
    IDX_CHECK(hipMalloc(&t_k1.p, n * sizeof(uint32_t)));
    IDX_CHECK(hipMalloc(&t_v0.p, n * sizeof(uint32_t)));
    IDX_CHECK(hipMalloc(&t_v1.p, n * sizeof(uint32_t)));
    const uint64_t* dcb = (const uint64_t*)t_cbase.p;
    uint32_t *k0 = (uint32_t*)t_k0.p, *k1 = (uint32_t*)t_k1.p, *v0 = (uint32_t*)t_v0.p, *v1 = (uint32_t*)t_v1.p;
    const dim3 grid(16384), block(256);
    hipLaunchKernelGGL(k_idx_windows, grid, block, 0, s, G.g2b, G.nmask, G.coff, G.clen, dcb, nctg, n, k0, v0);
    IDX_CHECK(hipGetLastError());
    size_t sort_bytes = 0;
    IDX_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, k0, k1, v0, v1, n, 0, 31, s));
    IDX_CHECK(hipMalloc(&t_sort.p, sort_bytes));
    IDX_CHECK(hipcub::DeviceRadixSort::SortPairs(t_sort.p, sort_bytes, k0, k1, v0, v1, n, 0, 31, s));
    // run start of every sorted entry (in k0, free after the sort): marks, then a max-scan in place
    uint32_t* rstart = k0;
    hipLaunchKernelGGL(k_idx_startmark, grid, block, 0, s, k1, n, rstart);
    hipLaunchKernelGGL(k_max_tiles, dim3(np_n), dim3(SCAN_T), 0, s, rstart, n, part);
    hipLaunchKernelGGL(k_max_parts, dim3(1), dim3(1024), 0, s, part, np_n);
    hipLaunchKernelGGL(k_max_apply, dim3(np_n), dim3(SCAN_T), 0, s, rstart, n, part);
    hipLaunchKernelGGL(k_idx_counts, grid, block, 0, s, k1, n, rstart, koff);
    IDX_CHECK(hipGetLastError());
    // koff: exclusive sum of the capped lengths over every code, in place
    hipLaunchKernelGGL(k_scan_tiles, dim3(np_k), dim3(SCAN_T), 0, s, koff, nk, part);
    hipLaunchKernelGGL(k_scan_parts, dim3(1), dim3(1024), 0, s, part, np_k);
    hipLaunchKernelGGL(k_scan_apply, dim3(np_k), dim3(SCAN_T), 0, s, koff, nk, part);
    IDX_CHECK(hipGetLastError());
    IDX_CHECK(hipMemcpyAsync(&total, koff + CODE_SPACE, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    IDX_CHECK(hipStreamSynchronize(s));
    IDX_CHECK(hipMalloc((void**)&kpos, std::max<uint64_t>(total, 1) * sizeof(uint64_t)));
    if (!fused) IDX_CHECK(hipMalloc((void**)&kfp, std::max<uint64_t>(total, 1) * sizeof(uint16_t)));
    hipLaunchKernelGGL(k_idx_place, grid, block, 0, s, k1, v1, n, rstart, koff, G.g2b, G.nmask, G.coff, G.clen, dcb,
                       nctg, kpos, kfp);
  } else {
    IDX_CHECK(hipMalloc((void**)&kpos, sizeof(uint64_t)));
    if (!fused) IDX_CHECK(hipMalloc((void**)&kfp, sizeof(uint16_t)));
  }
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    (void)hipFree(kpos);
    (void)hipFree(kfp);
    err = std::string("index build failed: ") + hipGetErrorString(e);
    return e;
  }
  *kpos_out = kpos;
  *kfp_out = kfp;
  *n_entries = total;
  return hipSuccess;
}

hipError_t gpu_index_from_table(const DevGenome& G, const uint64_t* h_clen, const HostTable& t, uint32_t* koff,
                                uint64_t** kpos_out, uint16_t** kfp_out, uint64_t* n_entries, bool fused,
                                hipStream_t s, std::string& err) {
  (void)h_clen;
  *kpos_out = nullptr;
  *kfp_out = nullptr;
  *n_entries = 0;
  const uint64_t n = t.key.size(), npos = t.pos.size();
  if (t.off.size() != n || t.cnt.size() != n) {
    err = "malformed table";
    return hipErrorInvalidValue;
  }
  uint64_t total = 0;
  for (uint64_t i = 0; i < n; i++) {
    if (t.cnt[i] > (uint32_t)MAX_COUNT) {
      err = "a table bucket holds more than the 300 entries of the reference's cap";
      return hipErrorInvalidValue;
    }
    total += t.cnt[i];
  }
  if (total >= (1ull << 32)) {
    err = "more than 2^32 - 1 indexed positions";
    return hipErrorInvalidValue;
  }
  IDX_CHECK(hipMemsetAsync(koff, 0, (CODE_SPACE + 1) * sizeof(uint32_t), s));
  const uint64_t nk = CODE_SPACE + 1;
  const uint32_t np_k = (uint32_t)((nk + SCAN_TILE - 1) / SCAN_TILE);
  Tmp t_part, t_key, t_off, t_cnt, t_src, t_bad;
  IDX_CHECK(hipMalloc(&t_part.p, (size_t)np_k * sizeof(uint32_t)));
  IDX_CHECK(hipMalloc(&t_bad.p, 2 * sizeof(uint32_t)));
  IDX_CHECK(hipMemsetAsync(t_bad.p, 0, 2 * sizeof(uint32_t), s));
  uint32_t* part = (uint32_t*)t_part.p;
  uint32_t* bad = (uint32_t*)t_bad.p;
  const dim3 grid(16384), block(256);
  uint32_t *dkey = nullptr, *doff = nullptr, *dcnt = nullptr;
  uint64_t* dsrc = nullptr;
  if (n) {
    IDX_CHECK(hipMalloc(&t_key.p, n * sizeof(uint32_t)));
    IDX_CHECK(hipMalloc(&t_off.p, n * sizeof(uint32_t)));
    IDX_CHECK(hipMalloc(&t_cnt.p, n * sizeof(uint32_t)));
    IDX_CHECK(hipMalloc(&t_src.p, std::max<uint64_t>(npos, 1) * sizeof(uint64_t)));
    dkey = (uint32_t*)t_key.p;
    doff = (uint32_t*)t_off.p;
    dcnt = (uint32_t*)t_cnt.p;
    dsrc = (uint64_t*)t_src.p;
    IDX_CHECK(hipMemcpyAsync(dkey, t.key.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice, s));
    IDX_CHECK(hipMemcpyAsync(doff, t.off.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice, s));
    IDX_CHECK(hipMemcpyAsync(dcnt, t.cnt.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice, s));
    if (npos) IDX_CHECK(hipMemcpyAsync(dsrc, t.pos.data(), npos * sizeof(uint64_t), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_tab_counts, grid, block, 0, s, dkey, doff, dcnt, n, npos, koff, bad);
    hipLaunchKernelGGL(k_tab_strip, grid, block, 0, s, koff, nk);
    IDX_CHECK(hipGetLastError());
  }
  hipLaunchKernelGGL(k_scan_tiles, dim3(np_k), dim3(SCAN_T), 0, s, koff, nk, part);
  hipLaunchKernelGGL(k_scan_parts, dim3(1), dim3(1024), 0, s, part, np_k);
  hipLaunchKernelGGL(k_scan_apply, dim3(np_k), dim3(SCAN_T), 0, s, koff, nk, part);
  IDX_CHECK(hipGetLastError());
  uint32_t h_bad[2] = {0, 0}, placed = 0;
  IDX_CHECK(hipMemcpyAsync(h_bad, bad, 2 * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  IDX_CHECK(hipMemcpyAsync(&placed, koff + CODE_SPACE, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  IDX_CHECK(hipStreamSynchronize(s));
  if (h_bad[0]) {
    err = std::to_string(h_bad[0]) +
          " table buckets whose key is no 15-mer's hash, repeats another's, or points past the positions";
    return hipErrorInvalidValue;
  }
  uint64_t* kpos = nullptr;
  uint16_t* kfp = nullptr;
  IDX_CHECK(hipMalloc((void**)&kpos, std::max<uint64_t>(placed, 1) * sizeof(uint64_t)));
  if (!fused) IDX_CHECK(hipMalloc((void**)&kfp, std::max<uint64_t>(placed, 1) * sizeof(uint16_t)));
  if (n) {
    hipLaunchKernelGGL(k_tab_place, grid, block, 0, s, dkey, doff, dcnt, n, dsrc, koff, G.g2b, G.nmask, G.coff,
                       G.clen, G.ncontigs, kpos, kfp, bad);
  }
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(h_bad, bad, 2 * sizeof(uint32_t), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess || h_bad[1]) {
    (void)hipFree(kpos);
    (void)hipFree(kfp);
    err = e != hipSuccess ? std::string("index from table failed: ") + hipGetErrorString(e)
                          : std::to_string(h_bad[1]) + " table positions outside the genome's contigs";
    return e != hipSuccess ? e : hipErrorInvalidValue;
  }
  *kpos_out = kpos;
  *kfp_out = kfp;
  *n_entries = placed;
  return hipSuccess;
}

}  // namespace pbw
