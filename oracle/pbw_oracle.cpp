// CPU oracle for the probebwa map path — TEST INFRASTRUCTURE ONLY.
//
// A scalar C++ restatement of the reference mapper (GUIBA-EX/probebwa, Rust)
// used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as
// the checker.  The HIP product library never links or calls it.
//
// Build: oracle/Makefile (g++ -O2 -ffp-contract=off, glibc libm).  Rust never
// contracts a*b+c into an FMA, so neither may this file.
//
// Parity pins (what checks each restated function; see tests/test_oracle_*.py):
//   genome.rs from_ascii/slice ........ round-trip + N-run tests, reference N-run test
//   hashtable.rs hash/pack/build ....... reference known answers (integration.rs:31-64,
//                                        hashtable.rs:336-437), bijectivity brute force,
//                                        naive-scan bucket invariants
//   seeds.rs ........................... brute-force naive seed scan on tiny genomes
//   candidates.rs ...................... brute-force clustering; cross-group order for
//                                        >1 contig is an EMULATION of FxHashMap
//                                        iteration (parity unpinned, see fx_group_order)
//   smith_waterman.rs .................. independent full-matrix Gotoh on tiny inputs,
//                                        E. coli window (integration.rs:131-153),
//                                        285-case battery (recall 284/285)
//   bayesian.rs / insert_size.rs ....... closed forms (MAPQ 80 / 99), lgamma bound,
//                                        Welford closed form
//   single.rs / paired.rs / lib.rs ..... end-to-end known answers of tests/mapping.rs
//   fastx.rs / sam.rs .................. reference parser regressions + field checks
//   glibc exp/log/pow last ulp ......... parity unpinned vs Rust's lowering
#include "pbw_oracle.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

namespace {

// ---------------------------------------------------------------------------
// types.rs
// ---------------------------------------------------------------------------
enum Strand : uint8_t { Forward = 0, Reverse = 1 };

enum CigarKind : uint8_t { OP_M = 0, OP_I = 1, OP_D = 2, OP_S = 4, OP_H = 5 };
struct CigarOp {
  CigarKind kind;
  uint32_t n;
};

// types.rs:34-47 reverse_complement: A<->T, C<->G, everything else -> N.
static std::vector<uint8_t> reverse_complement(const uint8_t* s, size_t n) {
  std::vector<uint8_t> rc(n);
  for (size_t i = 0; i < n; i++) {
    uint8_t b = s[n - 1 - i];
    uint8_t c;
    switch (b) {
      case 'A': c = 'T'; break;
      case 'T': c = 'A'; break;
      case 'C': c = 'G'; break;
      case 'G': c = 'C'; break;
      default: c = 'N'; break;
    }
    rc[i] = c;
  }
  return rc;
}

struct Read {
  const uint8_t* seq;
  size_t len;
  const uint8_t* qual;
  size_t qlen;
};

// ---------------------------------------------------------------------------
// genome.rs
// ---------------------------------------------------------------------------
static const uint8_t BITS2BASE[4] = {'A', 'C', 'G', 'T'};

struct Contig {
  std::string name;
  std::vector<uint8_t> packed;
  std::vector<std::pair<uint32_t, uint32_t>> n_runs;
  uint64_t length = 0;
};

struct Genome {
  std::vector<Contig> contigs;
  uint64_t total_length = 0;
};

// genome.rs:17-45 Contig::from_ascii
static Contig contig_from_ascii(const std::string& name, const uint8_t* seq, uint64_t len) {
  Contig c;
  c.name = name;
  c.length = len;
  c.packed.assign((len + 3) / 4, 0);
  bool in_run = false;
  uint64_t run_start = 0;
  for (uint64_t pos = 0; pos < len; pos++) {
    uint8_t bits = 0;
    bool acgt = true;
    switch (seq[pos]) {
      case 'A': case 'a': bits = 0; break;
      case 'C': case 'c': bits = 1; break;
      case 'G': case 'g': bits = 2; break;
      case 'T': case 't': bits = 3; break;
      default: bits = 0; acgt = false; break;
    }
    c.packed[pos / 4] |= (uint8_t)(bits << ((pos % 4) * 2));
    if (acgt) {
      if (in_run) {
        c.n_runs.push_back({(uint32_t)run_start, (uint32_t)pos});
        in_run = false;
      }
    } else if (!in_run) {
      in_run = true;
      run_start = pos;
    }
  }
  if (in_run) c.n_runs.push_back({(uint32_t)run_start, (uint32_t)len});
  return c;
}

// genome.rs:48-68 Contig::slice (clamped; N overlay scans every run)
static void contig_slice(const Contig& c, uint64_t start, uint64_t end, std::vector<uint8_t>& out) {
  start = std::min(start, c.length);
  end = std::min(end, c.length);
  out.clear();
  if (start >= end) return;
  out.resize(end - start);
  for (uint64_t pos = start; pos < end; pos++) {
    uint8_t byte = c.packed[pos / 4];
    out[pos - start] = BITS2BASE[(byte >> ((pos % 4) * 2)) & 3];
  }
  for (auto& r : c.n_runs) {
    uint64_t os = std::max<uint64_t>(r.first, start);
    uint64_t oe = std::min<uint64_t>(r.second, end);
    for (uint64_t p = os; p < oe; p++) out[p - start] = 'N';
  }
}

// ---------------------------------------------------------------------------
// hashtable.rs
// ---------------------------------------------------------------------------
static const uint64_t POSITION_BITS = 40;
static const uint64_t POSITION_MASK = (1ull << 40) - 1;
static const uint64_t CONTIG_MASK = (1ull << 23) - 1;

// hashtable.rs:126-134 base2bits
static inline uint32_t base2bits(uint8_t b) {
  switch (b) {
    case 'A': case 'a': return 0;
    case 'C': case 'c': return 1;
    case 'G': case 'g': return 2;
    case 'T': case 't': return 3;
    default: return 0;
  }
}
// hashtable.rs:117-124 mix
static inline uint32_t mix(uint32_t h) {
  h ^= h >> 16;
  h *= 0x7feb352du;
  h ^= h >> 15;
  h *= 0x846ca68bu;
  h ^= h >> 16;
  return h;
}
// hashtable.rs:107-113 hash_kmer (only the first 15 bases are packed)
static inline uint32_t hash_kmer(const uint8_t* kmer, size_t len) {
  uint32_t raw = 0;
  for (size_t i = 0; i < len && i < 15; i++) raw = (raw << 2) | base2bits(kmer[i]);
  return mix(raw);
}
// hashtable.rs:147-149 is_acgt_only
static inline bool is_acgt_only(const uint8_t* kmer, size_t len) {
  for (size_t i = 0; i < len; i++) {
    switch (kmer[i]) {
      case 'A': case 'a': case 'C': case 'c': case 'G': case 'g': case 'T': case 't': break;
      default: return false;
    }
  }
  return true;
}
// hashtable.rs:157-164 pack_position
static inline uint64_t pack_position(uint64_t cid, uint64_t pos, Strand s) {
  uint64_t sb = (s == Forward) ? 0 : 1;
  return (sb << 63) | ((cid & CONTIG_MASK) << POSITION_BITS) | (pos & POSITION_MASK);
}
// hashtable.rs:166-171 unpack_position
static inline void unpack_position(uint64_t p, uint64_t& cid, uint64_t& pos, Strand& s) {
  s = ((p >> 63) & 1) == 0 ? Forward : Reverse;
  cid = (p >> POSITION_BITS) & CONTIG_MASK;
  pos = p & POSITION_MASK;
}

// Open-addressing u32 -> (u32,u32) map standing in for FxHashMap<u32,(u32,u32)>
// (hashtable.rs:62).  Only lookup results matter on the path, not its layout;
// a slot holds the key and both values (b = 0 marks an empty slot: every
// stored count is at least 1), so a probe reads one line.
struct U32Map {
  struct Slot {
    uint32_t key, a, b;
  };
  std::vector<Slot> s;
  uint64_t mask = 0;
  void init(uint64_t expected) {
    uint64_t cap = 16;
    while (cap < expected + expected / 2 + 16) cap <<= 1;
    s.assign(cap, Slot{0, 0, 0});
    mask = cap - 1;
  }
  static inline uint64_t slot_of(uint32_t k) { return (uint64_t)k * 0x9E3779B97F4A7C15ull >> 20; }
  // slot of a key not yet present (b is set by the caller)
  uint64_t insert_new(uint32_t k) {
    uint64_t i = slot_of(k) & mask;
    while (s[i].b) i = (i + 1) & mask;
    s[i].key = k;
    return i;
  }
  bool find(uint32_t k, uint64_t& idx) const {
    if (s.empty()) return false;
    uint64_t i = slot_of(k) & mask;
    while (s[i].b) {
      if (s[i].key == k) {
        idx = i;
        return true;
      }
      i = (i + 1) & mask;
    }
    return false;
  }
};

// The build's per-k-mer counter (`seen`, hashtable.rs:253): u32 -> u32 count.
struct CountMap {
  struct Slot {
    uint32_t key, count;  // count 0: empty
  };
  std::vector<Slot> s;
  uint64_t mask = 0;
  void init(uint64_t expected) {
    uint64_t cap = 16;
    while (cap < expected + expected / 2 + 16) cap <<= 1;
    s.assign(cap, Slot{0, 0});
    mask = cap - 1;
  }
  // the count of k, inserting k (count 0) if absent
  uint32_t& at(uint32_t k) {
    uint64_t i = U32Map::slot_of(k) & mask;
    while (s[i].count) {
      if (s[i].key == k) return s[i].count;
      i = (i + 1) & mask;
    }
    s[i].key = k;
    return s[i].count;  // 0 until the caller counts the window
  }
};

// The table is held as 2^pbits sub-maps, one per range of hash values (the top
// pbits bits); `positions` holds every bucket in ascending hash order, so the
// whole is one FxHashMap<u32,(u32,u32)> over one positions vector.
struct HashTable {
  std::vector<U32Map> tables;       // hash -> (offset, count), by hash range
  uint32_t pbits = 0;
  std::vector<uint64_t> positions;  // flat store, buckets in ascending hash order
  std::vector<uint32_t> sorted_keys;
  std::vector<uint32_t> key_off;
  size_t k = 15, stride = 5;
  uint64_t total_entries = 0;

  // hashtable.rs:90-102 lookup/get
  inline const uint64_t* lookup(const uint8_t* kmer, size_t len, uint64_t& count) const {
    uint32_t h = hash_kmer(kmer, len);
    const U32Map& t = tables[pbits ? (h >> (32 - pbits)) : 0];
    uint64_t idx;
    if (t.find(h, idx)) {
      count = t.s[idx].b;
      return positions.data() + t.s[idx].a;
    }
    count = 0;
    return nullptr;
  }
};

// Stable LSD radix sort of (hash, packed) pairs by hash: same result as the
// reference's stable `sort_by_key(|&(h, _)| h)` (hashtable.rs:297).
static void stable_sort_by_hash(std::vector<uint32_t>& h, std::vector<uint64_t>& p) {
  size_t n = h.size();
  std::vector<uint32_t> h2(n);
  std::vector<uint64_t> p2(n);
  for (int pass = 0; pass < 2; pass++) {
    int shift = pass * 16;
    std::vector<uint64_t> cnt(65537, 0);
    for (size_t i = 0; i < n; i++) cnt[((h[i] >> shift) & 0xFFFF) + 1]++;
    for (int i = 0; i < 65536; i++) cnt[i + 1] += cnt[i];
    for (size_t i = 0; i < n; i++) {
      uint64_t d = cnt[(h[i] >> shift) & 0xFFFF]++;
      h2[d] = h[i];
      p2[d] = p[i];
    }
    h.swap(h2);
    p.swap(p2);
  }
}

// Progress lines of long index builds (genome-scale references), on stderr.
static int g_progress = 0;
static unsigned g_index_threads = 1;

// hashtable.rs:248-322 HashTableBuilder::build.
//
// The reference scans contigs in order, forward windows i = 0, stride, ... then
// the windows of the reverse complement, keeping a k-mer's first max_count
// windows in that scan order, then stably sorts (hash, packed) by hash and
// records one (offset, count) per hash run.  Here the scan is cut into units
// that run on `threads` threads, each unit sorting its windows into 2^pbits
// lists by hash range (the top pbits bits); range p then concatenates its lists
// in unit order -- the scan order -- and applies the cap, the stable sort and
// the run table on its own.  A k-mer's windows all fall into one range and stay
// in scan order, so every bucket, and the ascending-hash concatenation of the
// ranges, is the reference's.
static HashTable* hash_build(const Genome& g, size_t k, size_t stride, size_t max_count, unsigned threads) {
  threads = std::max(1u, threads);
  uint64_t windows = 0;
  for (auto& c : g.contigs) windows += 2 * (c.length / std::max<size_t>(stride, 1) + 1);
  // at least one range per thread, and ranges of at most ~2^25 windows (fewer large
  // transients alive at once); the table does not depend on the count
  uint32_t pbits = 0;
  while (((1u << pbits) < threads || (windows >> pbits) > (1ull << 25)) && pbits < 8) pbits++;
  const uint32_t P = 1u << pbits;
  auto range_of = [&](uint32_t h) { return pbits ? (h >> (32 - pbits)) : 0u; };
  const size_t UNIT = 1u << 22;  // windows per unit
  // unit lists: per unit, per range, (hash, packed) in scan order
  struct Unit {
    std::vector<std::vector<uint32_t>> h;
    std::vector<std::vector<uint64_t>> p;
  };
  std::vector<Unit> units;
  std::vector<uint8_t> seq, rc;
  stride = std::max<size_t>(stride, 1);
  for (size_t cid = 0; cid < g.contigs.size(); cid++) {
    const Contig& c = g.contigs[cid];
    if (g_progress) fprintf(stderr, "[oracle] index: scanning contig %zu/%zu\n", cid + 1, g.contigs.size());
    contig_slice(c, 0, c.length, seq);
    if (seq.size() < k) continue;
    rc = reverse_complement(seq.data(), seq.size());  // hashtable.rs:279
    const size_t nw = (seq.size() - k) / stride + 1;  // windows i = 0, stride, ... <= len - k, per strand
    // units of this contig: forward windows [w0, w1), then reverse windows [w0, w1)
    std::vector<std::pair<int, std::pair<size_t, size_t>>> todo;
    for (int strand = 0; strand < 2; strand++)
      for (size_t w0 = 0; w0 < nw; w0 += UNIT) todo.push_back({strand, {w0, std::min(nw, w0 + UNIT)}});
    const size_t u0 = units.size();
    units.resize(u0 + todo.size());
    std::atomic<size_t> next{0};
    auto work = [&]() {
      for (size_t t; (t = next.fetch_add(1)) < todo.size();) {
        Unit& U = units[u0 + t];
        U.h.assign(P, {});
        U.p.assign(P, {});
        const bool fwd = todo[t].first == 0;
        const uint8_t* src = fwd ? seq.data() : rc.data();
        for (size_t w = todo[t].second.first; w < todo[t].second.second; w++) {
          const size_t i = w * stride;
          const uint8_t* km = src + i;
          if (!is_acgt_only(km, k)) continue;  // hashtable.rs:268, 282
          const uint32_t h = hash_kmer(km, k);
          const uint32_t r = range_of(h);
          U.h[r].push_back(h);
          // hashtable.rs:272 (forward), :284-288 (reverse: position len - i - k)
          U.p[r].push_back(fwd ? pack_position(cid, i, Forward) : pack_position(cid, seq.size() - i - k, Reverse));
        }
      }
    };
    std::vector<std::thread> th;
    for (unsigned t = 1; t < std::min<size_t>(threads, todo.size()); t++) th.emplace_back(work);
    work();
    for (auto& t : th) t.join();
  }
  if (g_progress) fprintf(stderr, "[oracle] index: capping, sorting and tabling %u hash ranges\n", P);
  HashTable* ht = new HashTable();
  ht->k = k;
  ht->stride = stride;
  ht->pbits = pbits;
  ht->tables.resize(P);
  std::vector<std::vector<uint32_t>> rh(P), rkeys(P), roff(P);
  std::vector<std::vector<uint64_t>> rp(P);
  std::atomic<uint32_t> next{0};
  auto range_work = [&]() {
    for (uint32_t r; (r = next.fetch_add(1)) < P;) {
      std::vector<uint32_t>& ph = rh[r];
      std::vector<uint64_t>& pp = rp[r];
      uint64_t n = 0;
      for (auto& U : units) n += U.h[r].size();
      CountMap seen;
      seen.init(std::max<uint64_t>(n, 16));
      ph.reserve(n);
      pp.reserve(n);
      for (auto& U : units) {  // scan order
        for (size_t e = 0; e < U.h[r].size(); e++) {
          uint32_t& cnt = seen.at(U.h[r][e]);
          if (cnt < max_count) {  // hashtable.rs:270-273, 286-289
            ph.push_back(U.h[r][e]);
            pp.push_back(U.p[r][e]);
            cnt++;
          }
        }
        std::vector<uint32_t>().swap(U.h[r]);
        std::vector<uint64_t>().swap(U.p[r]);
      }
      seen = CountMap();
      stable_sort_by_hash(ph, pp);
      // hashtable.rs:303-313: one (offset, count) per hash run; offsets within the range
      uint64_t nkeys = 0;
      for (size_t i = 0; i < ph.size(); i++)
        if (i == 0 || ph[i] != ph[i - 1]) nkeys++;
      U32Map& tab = ht->tables[r];
      tab.init(nkeys);
      size_t i = 0;
      while (i < ph.size()) {
        uint32_t h = ph[i];
        size_t start = i;
        while (i < ph.size() && ph[i] == h) i++;
        uint64_t si = tab.insert_new(h);
        tab.s[si].a = (uint32_t)start;
        tab.s[si].b = (uint32_t)(i - start);
        rkeys[r].push_back(h);
        roff[r].push_back((uint32_t)start);
      }
      std::vector<uint32_t>().swap(ph);
    }
  };
  {
    std::vector<std::thread> th;
    for (unsigned t = 1; t < std::min<unsigned>(threads, P); t++) th.emplace_back(range_work);
    range_work();
    for (auto& t : th) t.join();
  }
  units.clear();
  // ranges in ascending hash order: shift each range's offsets by the entries before it
  uint64_t total = 0;
  for (uint32_t r = 0; r < P; r++) total += rp[r].size();
  if (total >= (1ull << 32)) {
    delete ht;
    return nullptr;
  }
  ht->positions.resize(total);
  uint64_t base = 0;
  for (uint32_t r = 0; r < P; r++) {
    std::copy(rp[r].begin(), rp[r].end(), ht->positions.begin() + base);
    for (auto& sl : ht->tables[r].s)
      if (sl.b) sl.a += (uint32_t)base;
    for (size_t q = 0; q < rkeys[r].size(); q++) {
      ht->sorted_keys.push_back(rkeys[r][q]);
      ht->key_off.push_back((uint32_t)(roff[r][q] + base));
    }
    base += rp[r].size();
    std::vector<uint64_t>().swap(rp[r]);
  }
  ht->key_off.push_back((uint32_t)total);
  ht->total_entries = total;
  return ht;
}

// ---------------------------------------------------------------------------
// align/smith_waterman.rs, align/cigar.rs
// ---------------------------------------------------------------------------
static const int32_t MATCH_SCORE = 20;                // smith_waterman.rs:81
static const int32_t NEG_INF = INT32_MIN / 2;         // smith_waterman.rs:173

// Rust `f64 as i32` saturates; values here are always in range.
static inline int32_t f64_to_i32(double x) {
  if (std::isnan(x)) return 0;
  if (x >= 2147483647.0) return INT32_MAX;
  if (x <= -2147483648.0) return INT32_MIN;
  return (int32_t)x;
}

struct SwAligner {
  int32_t gap_open, gap_extend;
  int64_t band_width;
  double homopolymer_slope = 1.0;  // smith_waterman.rs:97
  double mismatch_floor = 0.2;     // smith_waterman.rs:98

  // smith_waterman.rs:113-130 new
  SwAligner(size_t band, uint32_t go_phred, uint32_t ge_phred) {
    int32_t go = f64_to_i32(std::round(-90.0 * (double)go_phred / 40.0));
    int32_t ge = f64_to_i32(std::round(-1.0 * (double)ge_phred / 3.0));
    gap_open = std::min(go, -1);
    gap_extend = std::min(ge, -1);
    band_width = (int64_t)band;
  }
  // smith_waterman.rs:272-275 mismatch_penalty
  int32_t mismatch_penalty(uint8_t q) const {
    double scale = std::min(std::max((double)q / 30.0, mismatch_floor), 1.0);
    return f64_to_i32(std::round(-30.0 * scale));
  }
  // smith_waterman.rs:288-291 homopolymer_gap_open
  int32_t homopolymer_gap_open(uint16_t run) const {
    uint16_t r1 = run >= 1 ? (uint16_t)(run - 1) : 0;
    int32_t discount = f64_to_i32(std::round((double)r1 * homopolymer_slope));
    return std::min(gap_open + discount, -1);
  }

  // smith_waterman.rs:163-255 align (+ traceback :299-345, collapse :365-381)
  int32_t align(const uint8_t* read, size_t m, const uint8_t* qual, size_t qlen, const uint8_t* ref,
                size_t n, int64_t hint, std::vector<CigarOp>& cigar, uint64_t& ref_offset,
                uint64_t* cells) const {
    cigar.clear();
    if (m == 0 || n == 0) {
      cigar.push_back({OP_S, (uint32_t)m});
      ref_offset = 0;
      return 0;
    }
    // smith_waterman.rs:348-363 precompute_homopolymer_runs
    std::vector<uint16_t> run(n, 1);
    {
      size_t start = 0;
      for (size_t idx = 1; idx <= n; idx++) {
        if (idx == n || ref[idx] != ref[start]) {
          size_t len = std::min<size_t>(idx - start, 65535);
          for (size_t r = start; r < idx; r++) run[r] = (uint16_t)len;
          start = idx;
        }
      }
    }
    size_t width = n + 1;
    std::vector<int32_t> dm(width, 0), di(width, NEG_INF), dd(width, NEG_INF);
    std::vector<int32_t> dmn(width, NEG_INF), din(width, NEG_INF), ddn(width, NEG_INF);
    std::vector<uint8_t> tbm((m + 1) * width, 0), tbi((m + 1) * width, 0), tbd((m + 1) * width, 0);
    for (size_t i = 1; i <= m; i++) {
      int64_t center = (int64_t)i + hint;
      int64_t js = std::max<int64_t>(center - band_width, 1);
      int64_t je = std::max<int64_t>(std::min<int64_t>(center + band_width, (int64_t)n), 0);
      if (js > je) continue;  // skipped row: no reset, no swap
      std::fill(dmn.begin(), dmn.end(), NEG_INF);
      std::fill(din.begin(), din.end(), NEG_INF);
      std::fill(ddn.begin(), ddn.end(), NEG_INF);
      uint8_t q = (i - 1 < qlen) ? qual[i - 1] : 30;
      size_t rb = i * width;
      for (int64_t jj = js; jj <= je; jj++) {
        size_t j = (size_t)jj;
        int32_t sc = (read[i - 1] == ref[j - 1]) ? MATCH_SCORE : mismatch_penalty(q);
        int32_t fm = dm[j - 1] + sc, fi = di[j - 1] + sc, fd = dd[j - 1] + sc;
        int32_t ms = std::max(std::max(fm, fi), fd);
        int32_t goh = homopolymer_gap_open(run[j - 1]);
        int32_t io = dm[j] + goh, ie = di[j] + gap_extend;
        int32_t is = std::max(io, ie);
        int32_t dO = dmn[j - 1] + goh, de = ddn[j - 1] + gap_extend;
        int32_t ds = std::max(dO, de);
        dmn[j] = ms;
        din[j] = is;
        ddn[j] = ds;
        tbm[rb + j] = (ms == fm) ? 0 : ((ms == fi) ? 1 : 2);
        tbi[rb + j] = (is == io) ? 0 : 1;
        tbd[rb + j] = (ds == dO) ? 0 : 1;
      }
      if (cells) *cells += (uint64_t)(je - js + 1);
      dm.swap(dmn);
      di.swap(din);
      dd.swap(ddn);
    }
    // endpoint: last maximum over (j asc; M then I) — Iterator::max_by_key
    size_t best_j = n;
    int32_t best = 0;
    uint8_t best_state = 0;
    bool first = true;
    for (size_t j = 0; j <= n; j++) {
      for (uint8_t s = 0; s < 2; s++) {
        int32_t v = s == 0 ? dm[j] : di[j];
        if (first || v >= best) {
          best = v;
          best_j = j;
          best_state = s;
          first = false;
        }
      }
    }
    // traceback
    std::vector<CigarOp> ops;
    size_t i = m, j = best_j;
    uint8_t state = best_state;
    while (i > 0) {
      if (j == 0) {
        ops.push_back({OP_I, 1});
        i -= 1;
        continue;
      }
      if (state == 0) {
        uint8_t parent = tbm[i * width + j];
        ops.push_back({OP_M, 1});
        i -= 1;
        j -= 1;
        state = parent;
      } else if (state == 1) {
        bool opened = tbi[i * width + j] == 0;
        ops.push_back({OP_I, 1});
        i -= 1;
        state = opened ? 0 : 1;
      } else {
        bool opened = tbd[i * width + j] == 0;
        ops.push_back({OP_D, 1});
        j -= 1;
        state = opened ? 0 : 2;
      }
    }
    std::reverse(ops.begin(), ops.end());
    for (auto& op : ops) {
      if (!cigar.empty() && cigar.back().kind == op.kind &&
          (op.kind == OP_M || op.kind == OP_I || op.kind == OP_D || op.kind == OP_S)) {
        cigar.back().n += op.n;
      } else {
        cigar.push_back(op);
      }
    }
    ref_offset = j;
    return best;
  }
};

// cigar.rs:23-28 cigar_reference_span
static inline uint64_t cigar_reference_span(const std::vector<CigarOp>& c) {
  uint64_t s = 0;
  for (auto& op : c)
    if (op.kind == OP_M || op.kind == OP_D) s += op.n;
  return s;
}

// ---------------------------------------------------------------------------
// mapq/bayesian.rs
// ---------------------------------------------------------------------------
static const double LN_10 = 2.302585092994045684017991454684364208;
static const double DB = -0.1 * LN_10;  // bayesian.rs:36

// bayesian.rs:39-41
static inline double posterior_floor() { return std::exp(99.0 * DB); }

// bayesian.rs:47-58 logadd
static inline double logadd(double a, double b) {
  if (a > b) std::swap(a, b);
  if (b >= a + 10.0) return a;
  if (b >= a + 4.0) return std::max(a - 1.0, 0.0);
  if (b >= a + 2.0) return std::max(a - 2.0, 0.0);
  return std::max(a - 3.0, 0.0);
}
// bayesian.rs:63-65 py_int_round
static inline double py_int_round(double x) { return std::trunc(x + 0.5); }

// bayesian.rs:72-86 phred_n_choose_m
static double phred_n_choose_m(int64_t n, int64_t m) {
  if (m > n / 2) m = n - m;
  if (m <= 0) return 0.0;
  double phred = 0.0;
  while (m > 1) {
    double t = std::log((double)n) - std::log((double)m);
    phred += t;
    n -= 1;
    m -= 1;
  }
  phred += std::log((double)n);
  return py_int_round(10.0 * phred / LN_10);
}

// bayesian.rs:445-456 gc_content
static double gc_content(const uint8_t* s, size_t n) {
  uint64_t gc = 0, total = 0;
  for (size_t i = 0; i < n; i++) {
    uint8_t b = s[i];
    if (b >= 'a' && b <= 'z') b = (uint8_t)(b - 32);
    if (b == 'G' || b == 'C') {
      gc++;
      total++;
    } else if (b == 'A' || b == 'T') {
      total++;
    }
  }
  return total == 0 ? 0.5 : (double)gc / (double)total;
}
// bayesian.rs:463-466 effective_alphabet_size
static double effective_alphabet_size(double gc) {
  double f = std::min(std::max(gc, 1e-6), 1.0 - 1e-6);
  return std::exp(-f * std::log(f / 2.0) - (1.0 - f) * std::log((1.0 - f) / 2.0));
}

struct Mapq {
  double sub_phred;
  size_t seed_k;
  double genome_size;

  // bayesian.rs:103-110 new
  Mapq(double sub_rate, size_t k, uint64_t gsize) {
    sub_phred = py_int_round(std::log(std::max(sub_rate, 1e-30)) / DB);
    seed_k = k;
    genome_size = (double)std::max<uint64_t>(gsize, 1);
  }

  // bayesian.rs:142-179 alignment_log_likelihood_bases
  double log_lk(const uint8_t* bases, size_t blen, const uint8_t* qual, size_t qlen, const uint8_t* ref,
                size_t rlen, const std::vector<CigarOp>& cigar) const {
    double lk = 0.0;
    size_t qi = 0, ri = 0;
    for (auto& op : cigar) {
      switch (op.kind) {
        case OP_M:
          for (uint32_t t = 0; t < op.n; t++) {
            if (qi >= blen || ri >= rlen) break;
            double raw_q = (double)(qi < qlen ? qual[qi] : 30);
            double eff_q = logadd(raw_q, sub_phred);
            double e = std::pow(10.0, -eff_q / 10.0);
            double p_match = 1.0 - e;
            double p = bases[qi] == ref[ri] ? p_match : e / 3.0;
            lk += std::log(std::max(p, 1e-300));
            qi++;
            ri++;
          }
          break;
        case OP_I:
          for (uint32_t t = 0; t < op.n; t++) {
            if (qi >= blen) break;
            lk += std::log(0.25);
            qi++;
          }
          break;
        case OP_D: ri += op.n; break;
        case OP_S: qi += op.n; break;
        default: break;
      }
    }
    return lk;
  }

  // bayesian.rs:193-201 select_best: last maximum wins; sum in index order
  void select_best(const std::vector<double>& lks, size_t& best_idx, double& sum_exp) const {
    double mx = lks[0];
    size_t bi = 0;
    for (size_t i = 0; i < lks.size(); i++)
      if (lks[i] >= mx) {
        mx = lks[i];
        bi = i;
      }
    double s = -0.0;
    for (double v : lks) s += std::exp(v - mx);
    best_idx = bi;
    sum_exp = s;
  }
  // bayesian.rs:208-216 posterior_fractions
  std::vector<double> posterior_fractions(const std::vector<double>& lks) const {
    std::vector<double> out;
    if (lks.empty()) return out;
    double mx = -INFINITY;
    for (double v : lks) mx = std::fmax(mx, v);
    std::vector<double> ex;
    for (double v : lks) ex.push_back(std::exp(v - mx));
    double s = -0.0;
    for (double e : ex) s += e;
    for (double e : ex) out.push_back(e / s);
    return out;
  }
  // bayesian.rs:227-230
  double wrong_mapping_probability(double sum_exp) const {
    double bp = 1.0 / std::max(sum_exp, 1.0);
    return std::max(1.0 - bp, posterior_floor());
  }

  // bayesian.rs:306-373 log_entropy
  double log_entropy(const std::vector<CigarOp>& cigar, const uint8_t* qb, size_t qblen, const uint8_t* qq,
                     size_t qqlen, const uint8_t* rw, size_t rwlen) const {
    const uint8_t MIN_Q = 10;
    const double MAX_INDEL_LEN = 30.0;
    int64_t n = 0, mm = 0;
    size_t readlen = 0, refpos = 0;
    int64_t indel_runs = 0, deletion_runs = 0;
    uint8_t curtype = 0;
    for (auto& op : cigar) {
      if (op.kind == OP_I) {
        for (uint32_t t = 0; t < op.n; t++) {
          if (curtype != 1) indel_runs++;
          curtype = 1;
          readlen++;
        }
      } else if (op.kind == OP_D) {
        for (uint32_t t = 0; t < op.n; t++) {
          if (curtype != 2) {
            deletion_runs++;
            indel_runs++;
          }
          curtype = 2;
          refpos++;
        }
      } else if (op.kind == OP_M) {
        for (uint32_t t = 0; t < op.n; t++) {
          if (curtype != 0) indel_runs++;
          curtype = 0;
          if (readlen >= qblen || refpos >= rwlen) break;
          uint8_t q = readlen < qqlen ? qq[readlen] : 0;
          bool hq = q >= MIN_Q;
          if (hq) n++;
          uint8_t a = qb[readlen], r = rw[refpos];
          if (a != r && a != 'N' && r != 'N' && hq) mm++;
          readlen++;
          refpos++;
        }
      }
    }
    int64_t insertion_runs = indel_runs - deletion_runs;
    double gc = gc_content(rw, rwlen);
    double b = effective_alphabet_size(gc);
    double two_g = 2.0 * genome_size;
    double match_term = py_int_round(10.0 * ((double)n * std::log(b) - std::log(two_g)) / LN_10);
    double mismatch_term = py_int_round(10.0 * (double)mm * std::log(std::max(b - 1.0, 1e-6)) / LN_10);
    double position_choice = phred_n_choose_m((int64_t)readlen, 2 * insertion_runs + deletion_runs);
    double deletion_length = 0.0;
    if (deletion_runs > 0)
      deletion_length = py_int_round((double)deletion_runs * 10.0 * std::log(MAX_INDEL_LEN) / LN_10);
    return match_term - mismatch_term - phred_n_choose_m(n, mm) - position_choice - deletion_length;
  }

  // bayesian.rs:405-439 missing_locus_probability
  double missing_locus_probability(const uint8_t* qual, size_t len, const double* mask, size_t mlen) const {
    size_t k = seed_k;
    if (k == 0 || len < k) return 0.5;
    std::vector<double> q(len);
    for (size_t i = 0; i < len; i++) {
      double eff_q = logadd((double)qual[i], sub_phred);
      q[i] = std::min(std::pow(10.0, -eff_q / 10.0), 1.0 - 1e-12);
    }
    double offset_means = 0.0;
    size_t phases = std::min<size_t>(k, 5);
    for (size_t phase = 0; phase < phases; phase++) {
      double log_p = 0.0;
      size_t start = phase;
      while (start + k <= len) {
        double rep = start < mlen ? mask[start] : 0.0;
        double z = 1.0;
        for (size_t t = start; t < start + k; t++) z = z * (1.0 - q[t]);
        double p1 = -0.0;
        for (size_t t = start; t < start + k; t++) p1 = p1 + q[t] / (1.0 - q[t]);
        double p2 = std::min(std::max(1.0 - z * (1.0 + p1), 0.0), 1.0);
        double pm = std::min(std::max(1.0 - (1.0 - p2) * (1.0 - rep), 1e-300), 1.0);
        log_p += std::log(pm);
        start += k;
      }
      offset_means += std::exp(log_p);
    }
    return std::min(offset_means / (double)phases, 0.99);
  }

  // bayesian.rs:251-282 finalize_mapq
  void finalize_mapq(double sum_exp, const Read& read, const std::vector<CigarOp>& cigar, const uint8_t* qb,
                     size_t qblen, const uint8_t* qq, size_t qqlen, const uint8_t* rw, size_t rwlen,
                     const double* mask, size_t mlen, uint8_t& mapq, double& posterior) const {
    double pw = wrong_mapping_probability(sum_exp);
    double wrongmapping = py_int_round(std::log(pw) / DB);
    double entropy = log_entropy(cigar, qb, qblen, qq, qqlen, rw, rwlen) - 3.0;
    if (entropy <= 0.0) {
      mapq = 0;
      posterior = 0.0;
      return;
    }
    double p_missing = std::max(missing_locus_probability(read.qual, read.qlen, mask, mlen), 1e-30);
    double notfound = -10.0 * std::log10(p_missing);
    double combined = logadd(wrongmapping, notfound);
    double pp = std::min(std::max(logadd(entropy, combined), 0.0), 99.0);
    double r = std::round(pp);
    mapq = (uint8_t)(r < 0 ? 0 : (r > 255 ? 255 : r));
    double post = std::pow(10.0, -pp / 10.0);
    posterior = 1.0 - post;
  }
};

// mapq/insert_size.rs:23-44
struct InsertModel {
  double mean, sd;
  uint64_t n_observed = 0;
  double log_likelihood(double x) const {
    if (sd <= 0.0) return -1e300;
    double diff = x - mean;
    return -0.5 * (diff * diff) / (sd * sd) - 0.5 * std::log(2.0 * M_PI * sd * sd);
  }
  void update(double obs) {
    n_observed += 1;
    double n = (double)n_observed;
    double delta = obs - mean;
    mean += delta / n;
    double delta2 = obs - mean;
    double var = n > 1.0 ? ((sd * sd) * (n - 1.0) + delta * delta2) / n : sd * sd;
    sd = std::max(std::sqrt(var), 1.0);
  }
};

// ---------------------------------------------------------------------------
// mapper/seeds.rs
// ---------------------------------------------------------------------------
struct SeedHit {
  size_t read_pos;
  size_t contig;
  size_t ref_pos;
  Strand strand;
  bool neighbour;
};

struct Stats {
  uint64_t lookups = 0, positions = 0, filtered = 0, candidates = 0, dp_jobs = 0, dp_cells = 0,
           wide_jobs = 0, wide_cells = 0;
  void add(const Stats& o) {
    lookups += o.lookups;
    positions += o.positions;
    filtered += o.filtered;
    candidates += o.candidates;
    dp_jobs += o.dp_jobs;
    dp_cells += o.dp_cells;
    wide_jobs += o.wide_jobs;
    wide_cells += o.wide_cells;
  }
};

// seeds.rs:224-232 clip_flanks
static inline void clip_flanks(size_t seq_len, size_t ws, size_t k, size_t aw, size_t bw, size_t& as,
                               size_t& ae, size_t& bs, size_t& be) {
  as = std::min(ws + k, seq_len);
  ae = std::min(as + aw, seq_len);
  be = std::min(ws, seq_len);
  bs = be >= bw ? be - bw : 0;
}

struct Fingerprint {
  uint32_t a = 0, c = 0, g = 0, t = 0;
};
// seeds.rs:241-253 compute_fingerprint (only A/C/G/T counted)
static inline void fp_add(Fingerprint& fp, const uint8_t* s, size_t n) {
  for (size_t i = 0; i < n; i++) {
    uint8_t b = s[i];
    if (b >= 'a' && b <= 'z') b = (uint8_t)(b - 32);
    if (b == 'A') fp.a++;
    else if (b == 'C') fp.c++;
    else if (b == 'G') fp.g++;
    else if (b == 'T') fp.t++;
  }
}
static inline uint32_t absdiff(uint32_t x, uint32_t y) { return x > y ? x - y : y - x; }

// seeds.rs:162-219 passes_similarity_filter
static bool passes_similarity_filter(const Read& read, const std::vector<uint8_t>& rc, size_t read_pos, size_t k,
                                     const Genome& g, size_t cid, size_t ref_pos, Strand strand,
                                     bool is_neighbour, std::vector<uint8_t>& scratch) {
  const Contig& c = g.contigs[cid];
  const uint8_t* rl;
  size_t ws;
  if (strand == Forward) {
    rl = read.seq;
    ws = read_pos;
  } else {
    rl = rc.data();
    ws = read.len - read_pos - k;
  }
  size_t ras, rae, rbs, rbe;
  clip_flanks(read.len, ws, k, 8, 4, ras, rae, rbs, rbe);
  size_t region_start = ref_pos >= 4 ? ref_pos - 4 : 0;
  size_t region_end = std::min<uint64_t>(ref_pos + k + 8, c.length);
  if (region_start >= region_end) return true;
  contig_slice(c, region_start, region_end, scratch);
  size_t gws = ref_pos - region_start;
  size_t gas, gae, gbs, gbe;
  clip_flanks(scratch.size(), gws, k, 8, 4, gas, gae, gbs, gbe);
  size_t after_len = std::min(rae - ras, gae - gas);
  size_t before_len = std::min(rbe - rbs, gbe - gbs);
  if (after_len + before_len < 12) return true;
  Fingerprint rf, gf;
  fp_add(rf, rl + ras, after_len);
  fp_add(rf, rl + rbe - before_len, before_len);
  fp_add(gf, scratch.data() + gas, after_len);
  fp_add(gf, scratch.data() + gbe - before_len, before_len);
  uint32_t dist = absdiff(rf.a, gf.a) + absdiff(rf.c, gf.c) + absdiff(rf.g, gf.g) + absdiff(rf.t, gf.t);
  uint32_t thr = 10 - (is_neighbour ? 2 : 0);
  return dist <= thr;
}

// seeds.rs:276-286 neighbour_probed_position
static inline bool neighbour_probed_position(size_t p, size_t L) {
  if (L <= 34) return true;
  if (L <= 49) return (p / 5) % 2 == 0;
  return (p % 15) < 5;
}

// seeds.rs:91-151 find_seed_hits
static void find_seed_hits(const Read& read, const HashTable& ht, const Genome& g, std::vector<SeedHit>& hits,
                           std::vector<double>& repeat_prob, Stats& st) {
  size_t k = ht.k;
  hits.clear();
  repeat_prob.clear();
  size_t L = read.len;
  if (L < k || k > 32) return;
  repeat_prob.assign(L - k + 1, 0.0);
  std::vector<uint8_t> rc = reverse_complement(read.seq, L);
  std::vector<uint8_t> scratch;
  static const uint8_t BASES[4] = {'A', 'C', 'G', 'T'};
  for (size_t p = 0; p + k <= L; p++) {
    const uint8_t* w = read.seq + p;
    if (!is_acgt_only(w, k)) continue;
    uint64_t cnt;
    const uint64_t* ex = ht.lookup(w, k, cnt);
    st.lookups++;
    st.positions += cnt;
    for (uint64_t e = 0; e < cnt; e++) {
      uint64_t cid, rpos;
      Strand s;
      unpack_position(ex[e], cid, rpos, s);
      if (passes_similarity_filter(read, rc, p, k, g, cid, rpos, s, false, scratch))
        hits.push_back({p, cid, rpos, s, false});
    }
    bool rep = cnt >= 64;
    if (rep) repeat_prob[p] = 1.0;
    if (neighbour_probed_position(p, L) && !rep) {
      uint8_t q = p < read.qlen ? read.qual[p] : 30;
      double nrp = std::pow(10.0, -(double)q / 10.0);
      uint8_t buf[32];
      memcpy(buf, w, k);
      // seeds.rs:291-309 probe_1mm_neighbours: i in 0..k, base in ACGT != orig
      for (size_t i = 0; i < k; i++) {
        uint8_t orig = buf[i];
        for (int bi = 0; bi < 4; bi++) {
          if (BASES[bi] == orig) continue;
          buf[i] = BASES[bi];
          uint64_t nc;
          const uint64_t* nh = ht.lookup(buf, k, nc);
          st.lookups++;
          st.positions += nc;
          if (nc >= 64) repeat_prob[p] = std::fmax(repeat_prob[p], nrp);
          for (uint64_t e = 0; e < nc; e++) {
            uint64_t cid, rpos;
            Strand s;
            unpack_position(nh[e], cid, rpos, s);
            if (passes_similarity_filter(read, rc, p, k, g, cid, rpos, s, true, scratch))
              hits.push_back({p, cid, rpos, s, true});
          }
        }
        buf[i] = orig;
      }
    }
  }
  st.filtered += hits.size();
}

// ---------------------------------------------------------------------------
// mapper/candidates.rs
// ---------------------------------------------------------------------------
struct Candidate {
  size_t contig;
  size_t position;
  Strand strand;
  double score;
};

static int g_group_order_mode = 0;

// Emulated iteration order of FxHashMap<(usize, Strand), _> (candidates.rs:34)
// given its keys in first-insertion order.  Declared emulation of rustc-hash
// 2.x FxHasher (K = 0xf1357aea2e62a9c5, add-then-multiply, finish rotl 26) and
// std's SwissTable (16-wide groups, triangular probing, growth 4->8->16->32...,
// resize re-inserts old buckets in ascending index order, iteration in
// ascending bucket index).  Parity unpinned: no reference test fixes the
// cross-group order, and the unpinned rustc fixes hashbrown's version.
static uint64_t fx_hash_group(uint64_t cid, uint64_t disc) {
  const uint64_t K = 0xf1357aea2e62a9c5ull;
  uint64_t h = 0;
  h = (h + cid) * K;
  h = (h + disc) * K;
  return (h << 26) | (h >> 38);
}
static size_t sw_capacity(size_t buckets) {
  if (buckets == 0) return 0;
  size_t mask = buckets - 1;
  return mask < 8 ? mask : (buckets / 8) * 7;
}
static size_t sw_buckets_for(size_t cap) {
  if (cap < 8) return cap < 4 ? 4 : 8;
  size_t adj = cap * 8 / 7;
  size_t p = 1;
  while (p < adj) p <<= 1;
  return p;
}
static size_t sw_find_slot(const std::vector<int>& sk, uint64_t h) {
  size_t nb = sk.size(), mask = nb - 1;
  size_t pos = (size_t)h & mask, stride = 0;
  auto ext_empty = [&](size_t idx) -> bool {
    if (idx < nb) return sk[idx] < 0;
    if (nb < 16) return idx < 16 ? true : sk[idx - 16] < 0;
    return sk[idx - nb] < 0;
  };
  for (;;) {
    for (size_t b = 0; b < 16; b++) {
      if (ext_empty(pos + b)) {
        size_t index = (pos + b) & mask;
        if (sk[index] >= 0) {  // fix_insert_slot (tables smaller than a group)
          for (size_t c = 0; c < 16; c++)
            if (ext_empty(c)) return c;
        }
        return index;
      }
    }
    stride += 16;
    pos = (pos + stride) & mask;
  }
}
static std::vector<size_t> fx_group_order(const std::vector<std::pair<size_t, Strand>>& keys) {
  std::vector<size_t> order;
  if (g_group_order_mode == 1) {
    for (size_t i = 0; i < keys.size(); i++) order.push_back(i);
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) {
      if (keys[a].first != keys[b].first) return keys[a].first < keys[b].first;
      return keys[a].second < keys[b].second;
    });
    return order;
  }
  std::vector<int> sk;  // bucket -> key index or -1
  std::vector<uint64_t> hs(keys.size());
  size_t items = 0, growth_left = 0;
  for (size_t ki = 0; ki < keys.size(); ki++) {
    hs[ki] = fx_hash_group(keys[ki].first, keys[ki].second == Forward ? 0 : 1);
    if (growth_left == 0) {  // reserve(1) -> reserve_rehash -> resize
      size_t cap = std::max(items + 1, sw_capacity(sk.size()) + 1);
      std::vector<int> nsk(sw_buckets_for(cap), -1);
      for (size_t b = 0; b < sk.size(); b++)
        if (sk[b] >= 0) nsk[sw_find_slot(nsk, hs[sk[b]])] = sk[b];
      sk.swap(nsk);
      growth_left = sw_capacity(sk.size()) - items;
    }
    sk[sw_find_slot(sk, hs[ki])] = (int)ki;
    growth_left--;
    items++;
  }
  for (size_t b = 0; b < sk.size(); b++)
    if (sk[b] >= 0) order.push_back((size_t)sk[b]);
  return order;
}

// candidates.rs:30-78 cluster_hits
static std::vector<Candidate> cluster_hits(const std::vector<SeedHit>& hits, size_t read_len, size_t k) {
  std::vector<std::pair<size_t, Strand>> keys;
  std::unordered_map<uint64_t, size_t> index;
  std::vector<std::vector<int64_t>> diags;
  for (auto& h : hits) {
    int64_t d = h.strand == Forward ? (int64_t)h.ref_pos - (int64_t)h.read_pos
                                    : (int64_t)h.ref_pos + (int64_t)h.read_pos;
    uint64_t key = (uint64_t)h.contig * 2 + (h.strand == Forward ? 0 : 1);
    auto it = index.find(key);
    size_t gi;
    if (it == index.end()) {
      gi = keys.size();
      index.emplace(key, gi);
      keys.push_back({h.contig, h.strand});
      diags.emplace_back();
    } else {
      gi = it->second;
    }
    diags[gi].push_back(d);
  }
  std::vector<Candidate> cands;
  for (size_t gi : fx_group_order(keys)) {
    std::vector<int64_t>& dh = diags[gi];
    std::stable_sort(dh.begin(), dh.end());
    size_t cs = 0;
    for (size_t idx = 1; idx <= dh.size(); idx++) {
      bool boundary = idx == dh.size() || dh[idx] - dh[idx - 1] > 50;
      if (!boundary) continue;
      int64_t sum = 0;
      for (size_t t = cs; t < idx; t++) sum += dh[t];
      int64_t avg = sum / (int64_t)(idx - cs);
      int64_t position = keys[gi].second == Forward ? avg : avg - (int64_t)read_len + (int64_t)k;
      cands.push_back({keys[gi].first, (size_t)std::max<int64_t>(position, 0), keys[gi].second,
                       (double)(idx - cs)});
      cs = idx;
    }
  }
  std::stable_sort(cands.begin(), cands.end(),
                   [](const Candidate& a, const Candidate& b) { return a.score > b.score; });
  return cands;
}

// ---------------------------------------------------------------------------
// mapper/single.rs
// ---------------------------------------------------------------------------
struct CandAln {
  size_t contig;
  size_t position;  // 0-based
  Strand strand;
  std::vector<CigarOp> cigar;
  double log_lk;
  std::vector<uint8_t> ref_window;
  bool rc;  // out_seq/out_qual = rc(read)/reversed quals
};

struct Options {
  double substitution_rate = 0.001, insert_size_mean = 250.0, insert_size_sd = 60.0, sv_prior_phred = 55.0;
  size_t max_indel_len = 30;
  uint32_t gap_open_phred = 40, gap_extend_phred = 3, threads = 1;
};

struct OutRec {
  int32_t contig = -1;
  uint64_t pos = 0;
  uint8_t mapq = 0;
  std::vector<CigarOp> cigar;
  Strand strand = Forward;
  bool seq_rc = false;
  bool is_paired = false, is_proper = false;
  int32_t mate_contig = -1;
  uint64_t mate_pos = 0;
  bool has_mate_pos = false;
  int64_t tlen = 0;
  bool has_tlen = false;
  double log_lk = 0.0, posterior = 0.0;
  int read_number = 0;
  int mate_strand = -1;
  bool mate_unmapped = false;
};

struct Mapper {
  const Genome* g;
  const HashTable* ht;
  Options opt;
  SwAligner sw;
  Mapq mq;
  double sv_prior_log;

  Mapper(const Genome* g_, const HashTable* h_, const Options& o)
      : g(g_), ht(h_), opt(o), sw(o.max_indel_len, o.gap_open_phred, o.gap_extend_phred),
        mq(o.substitution_rate, h_->k, g_->total_length) {
    sv_prior_log = -opt.sv_prior_phred / 10.0 * LN_10;  // paired.rs:97
  }

  static void query_of(const Read& r, bool rc, std::vector<uint8_t>& qs, std::vector<uint8_t>& qq) {
    if (!rc) {
      qs.assign(r.seq, r.seq + r.len);
      qq.assign(r.qual, r.qual + r.qlen);
    } else {
      qs = reverse_complement(r.seq, r.len);
      qq.assign(r.qual, r.qual + r.qlen);
      std::reverse(qq.begin(), qq.end());
    }
  }

  // single.rs:86-162 map_candidates
  void map_candidates(const Read& r, std::vector<CandAln>& out, std::vector<double>& mask, Stats& st) const {
    out.clear();
    std::vector<SeedHit> hits;
    find_seed_hits(r, *ht, *g, hits, mask, st);
    std::vector<Candidate> cands = cluster_hits(hits, r.len, ht->k);
    st.candidates += cands.size();
    std::vector<uint8_t> fs, fq, rs, rq;
    query_of(r, false, fs, fq);
    query_of(r, true, rs, rq);
    size_t left_pad = 10, right_pad = std::max<size_t>(opt.max_indel_len, 10);
    std::vector<uint8_t> seg;
    size_t ncand = std::min<size_t>(cands.size(), 20);
    for (size_t ci = 0; ci < ncand; ci++) {
      const Candidate& cd = cands[ci];
      const Contig& c = g->contigs[cd.contig];
      size_t ref_start = cd.position >= left_pad ? cd.position - left_pad : 0;
      size_t ref_end = std::min<uint64_t>(cd.position + r.len + right_pad, c.length);
      if (ref_start >= ref_end) continue;
      contig_slice(c, ref_start, ref_end, seg);
      bool rc = cd.strand == Reverse;
      const std::vector<uint8_t>& qs = rc ? rs : fs;
      const std::vector<uint8_t>& qq = rc ? rq : fq;
      CandAln a;
      uint64_t off = 0;
      st.dp_jobs++;
      sw.align(qs.data(), qs.size(), qq.data(), qq.size(), seg.data(), seg.size(), (int64_t)left_pad, a.cigar, off,
               &st.dp_cells);
      a.contig = cd.contig;
      a.position = ref_start + off;
      a.strand = cd.strand;
      uint64_t span = cigar_reference_span(a.cigar);
      contig_slice(c, a.position, std::min<uint64_t>(a.position + span, c.length), a.ref_window);
      a.log_lk = mq.log_lk(qs.data(), qs.size(), qq.data(), qq.size(), a.ref_window.data(), a.ref_window.size(),
                           a.cigar);
      a.rc = rc;
      out.push_back(std::move(a));
    }
  }

  void finalize(double sum_exp, const Read& r, const CandAln& a, const std::vector<double>& mask, uint8_t& mapq,
                double& post) const {
    std::vector<uint8_t> qs, qq;
    query_of(r, a.rc, qs, qq);
    mq.finalize_mapq(sum_exp, r, a.cigar, qs.data(), qs.size(), qq.data(), qq.size(), a.ref_window.data(),
                     a.ref_window.size(), mask.data(), mask.size(), mapq, post);
  }

  // single.rs:338-359 to_record
  OutRec to_record(const CandAln& a, uint8_t mapq, double post) const {
    OutRec o;
    o.contig = (int32_t)a.contig;
    o.pos = a.position + 1;
    o.mapq = mapq;
    o.cigar = a.cigar;
    o.strand = a.strand;
    o.seq_rc = a.rc;
    o.log_lk = a.log_lk;
    o.posterior = post;
    return o;
  }
  // single.rs:362-383 unmapped_record
  static OutRec unmapped(const Read& r) {
    OutRec o;
    o.cigar.push_back({OP_S, (uint32_t)r.len});
    return o;
  }

  // single.rs:164-207 map
  OutRec map_se(const Read& r, Stats& st) const {
    std::vector<CandAln> al;
    std::vector<double> mask;
    map_candidates(r, al, mask, st);
    if (al.empty()) return unmapped(r);
    std::vector<double> lks;
    for (auto& a : al) lks.push_back(a.log_lk);
    size_t bi;
    double se;
    mq.select_best(lks, bi, se);
    uint8_t mapq;
    double post;
    finalize(se, r, al[bi], mask, mapq, post);
    return to_record(al[bi], mapq, post);
  }

  // single.rs:220-275 align_at_window
  bool align_at_window(const Read& r, size_t cid, size_t expected, size_t radius, Strand strand, CandAln& a,
                       Stats& st) const {
    const Contig& c = g->contigs[cid];
    size_t ref_start = expected >= radius ? expected - radius : 0;
    size_t ref_end = std::min<uint64_t>(expected + r.len + radius, c.length);
    if (ref_start >= ref_end) return false;
    std::vector<uint8_t> seg;
    contig_slice(c, ref_start, ref_end, seg);
    std::vector<uint8_t> qs, qq;
    bool rc = strand == Reverse;
    query_of(r, rc, qs, qq);
    SwAligner wa(seg.size(), opt.gap_open_phred, opt.gap_extend_phred);
    int64_t hint = (int64_t)(expected >= ref_start ? expected - ref_start : 0);
    uint64_t off = 0;
    st.wide_jobs++;
    wa.align(qs.data(), qs.size(), qq.data(), qq.size(), seg.data(), seg.size(), hint, a.cigar, off, &st.wide_cells);
    a.contig = cid;
    a.position = ref_start + off;
    a.strand = strand;
    uint64_t span = cigar_reference_span(a.cigar);
    contig_slice(c, a.position, std::min<uint64_t>(a.position + span, c.length), a.ref_window);
    a.log_lk = mq.log_lk(qs.data(), qs.size(), qq.data(), qq.size(), a.ref_window.data(), a.ref_window.size(),
                         a.cigar);
    a.rc = rc;
    return true;
  }

  // -------------------------------------------------------------------------
  // mapper/paired.rs
  // -------------------------------------------------------------------------
  static bool same_locus(const CandAln& a, const CandAln& b) {  // paired.rs:517-519
    return a.contig == b.contig && a.position == b.position && a.strand == b.strand;
  }
  static int64_t implied_insert_size(const CandAln& a, const CandAln& b) {  // paired.rs:534-538
    const CandAln& up = a.position <= b.position ? a : b;
    const CandAln& dn = a.position <= b.position ? b : a;
    uint64_t dend = dn.position + cigar_reference_span(dn.cigar);
    return (int64_t)dend - (int64_t)up.position;
  }
  static bool is_fr_pair(const CandAln& a, const CandAln& b) {  // paired.rs:544-547
    const CandAln& up = a.position <= b.position ? a : b;
    const CandAln& dn = a.position <= b.position ? b : a;
    return up.strand == Forward && dn.strand != Forward;
  }
  static int argmax_index(const std::vector<double>& v) {  // paired.rs:511-515 (last max)
    if (v.empty()) return -1;
    size_t bi = 0;
    double mx = v[0];
    for (size_t i = 0; i < v.size(); i++)
      if (v[i] >= mx) {
        mx = v[i];
        bi = i;
      }
    return (int)bi;
  }
  static std::vector<size_t> shortlist_indices(const std::vector<double>& post) {  // paired.rs:553-566
    std::vector<size_t> order(post.size());
    for (size_t i = 0; i < post.size(); i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return post[b] < post[a]; });
    double cum = 0.0;
    size_t take = 0;
    for (size_t i : order) {
      if (take >= 20) break;
      if (take >= 3 && cum >= 0.999) break;
      cum += post[i];
      take++;
    }
    order.resize(take);
    return order;
  }
  // paired.rs:450-468 align_mate_near
  bool align_mate_near(const Read& mate, const CandAln& anchor, const InsertModel& im, CandAln& out,
                       Stats& st) const {
    double mf = std::max(im.mean, 0.0);
    size_t mean = mf >= 18446744073709551615.0 ? SIZE_MAX : (size_t)mf;
    size_t mlen = mate.len;
    double rad = std::ceil(5.0 * im.sd);
    size_t radius = (rad <= 0 ? 0 : (size_t)rad) + mlen;
    size_t expected;
    Strand ms;
    if (anchor.strand == Forward) {
      expected = anchor.position + (mean >= mlen ? mean - mlen : 0);
      ms = Reverse;
    } else {
      size_t aend = anchor.position + cigar_reference_span(anchor.cigar);
      expected = aend >= mean ? aend - mean : 0;
      ms = Forward;
    }
    return align_at_window(mate, anchor.contig, expected, radius, ms, out, st);
  }
  double pair_distance_log_lk(const CandAln& a, const CandAln& b, const InsertModel& im) const {  // :486-496
    if (a.contig != b.contig || !is_fr_pair(a, b)) return sv_prior_log;
    double gap = (double)implied_insert_size(a, b);
    return std::max(im.log_likelihood(gap), sv_prior_log);
  }
  // paired.rs:499-509 independent_record
  OutRec independent_record(const Read& r, const std::vector<CandAln>& c, const std::vector<double>& mask) const {
    if (c.empty()) return unmapped(r);
    std::vector<double> lk;
    for (auto& a : c) lk.push_back(a.log_lk);
    size_t bi;
    double se;
    mq.select_best(lk, bi, se);
    uint8_t mapq;
    double post;
    finalize(se, r, c[bi], mask, mapq, post);
    return to_record(c[bi], mapq, post);
  }
  // paired.rs:346-351: the raw pairs kept, first occurrence of each pair of loci
  static std::vector<size_t> dedup_pairs(const std::vector<std::pair<CandAln, CandAln>>& raw) {
    std::vector<size_t> kept;
    for (size_t i = 0; i < raw.size(); i++) {
      bool dup = false;
      for (size_t k : kept)
        if (same_locus(raw[k].first, raw[i].first) && same_locus(raw[k].second, raw[i].second)) {
          dup = true;
          break;
        }
      if (!dup) kept.push_back(i);
    }
    return kept;
  }
  // paired.rs:398-435 mapq_for_winning_pair
  void mapq_for_winning_pair(const Read& r1, const Read& r2, const CandAln& bx, const CandAln& by, double pair_se,
                             const std::vector<double>& lk1, const std::vector<double>& lk2,
                             const std::vector<double>& m1, const std::vector<double>& m2, bool bxo, bool byo,
                             uint8_t& q1, double& p1, uint8_t& q2, double& p2) const {
    size_t bi;
    if (bxo && byo) {
      double s1, s2;
      mq.select_best(lk1, bi, s1);
      mq.select_best(lk2, bi, s2);
      finalize(s1, r1, bx, m1, q1, p1);
      finalize(s2, r2, by, m2, q2, p2);
    } else if (bxo != byo) {
      double s;
      uint8_t q;
      double p;
      if (bxo) {
        mq.select_best(lk1, bi, s);
        finalize(s, r1, bx, m1, q, p);
      } else {
        mq.select_best(lk2, bi, s);
        finalize(s, r2, by, m2, q, p);
      }
      q1 = q2 = q;
      p1 = p2 = p;
    } else {
      finalize(pair_se, r1, bx, m1, q1, p1);
      finalize(pair_se, r2, by, m2, q2, p2);
    }
  }
  // paired.rs:239-371 joint_map
  void joint_map(const Read& r1, const Read& r2, const std::vector<CandAln>& c1, const std::vector<double>& m1,
                 const std::vector<CandAln>& c2, const std::vector<double>& m2, const InsertModel& im, OutRec& o1,
                 OutRec& o2, bool& confident, Stats& st) const {
    std::vector<double> lk1, lk2;
    for (auto& a : c1) lk1.push_back(a.log_lk);
    for (auto& a : c2) lk2.push_back(a.log_lk);
    int b1 = argmax_index(lk1), b2 = argmax_index(lk2);
    if (b1 >= 0 && b2 >= 0) {  // fast path :266-291
      size_t bi;
      double s1, s2;
      mq.select_best(lk1, bi, s1);
      mq.select_best(lk2, bi, s2);
      double w1 = mq.wrong_mapping_probability(s1), w2 = mq.wrong_mapping_probability(s2);
      if (w1 < 0.01 && w2 < 0.01) {
        double mi1 = mq.missing_locus_probability(r1.qual, r1.qlen, m1.data(), m1.size());
        double mi2 = mq.missing_locus_probability(r2.qual, r2.qlen, m2.data(), m2.size());
        if (mi1 < 0.01 && mi2 < 0.01) {
          const CandAln& bx = c1[b1];
          const CandAln& by = c2[b2];
          double gap = (double)implied_insert_size(bx, by);
          bool close = bx.contig == by.contig && is_fr_pair(bx, by) && std::fabs(gap - im.mean) <= 4.0 * im.sd;
          if (close) {
            uint8_t q1, q2;
            double p1, p2;
            mapq_for_winning_pair(r1, r2, bx, by, 1.0, lk1, lk2, m1, m2, true, true, q1, p1, q2, p2);
            o1 = to_record(bx, q1, p1);
            o2 = to_record(by, q2, p2);
            confident = true;
            return;
          }
        }
      }
    }
    std::vector<std::pair<CandAln, CandAln>> raw;
    CandAln tmp;
    if (!c1.empty() && !c2.empty()) {
      raw.push_back({c1[b1], c2[b2]});
      std::vector<double> post1 = mq.posterior_fractions(lk1);
      for (size_t i : shortlist_indices(post1))
        if (align_mate_near(r2, c1[i], im, tmp, st)) raw.push_back({c1[i], tmp});
      std::vector<double> post2 = mq.posterior_fractions(lk2);
      for (size_t j : shortlist_indices(post2))
        if (align_mate_near(r1, c2[j], im, tmp, st)) raw.push_back({tmp, c2[j]});
    } else if (c1.empty()) {
      if (align_mate_near(r1, c2[b2], im, tmp, st)) raw.push_back({tmp, c2[b2]});
    } else {
      if (align_mate_near(r2, c1[b1], im, tmp, st)) raw.push_back({c1[b1], tmp});
    }
    if (raw.empty()) {
      o1 = independent_record(r1, c1, m1);
      o2 = independent_record(r2, c2, m2);
      confident = true;
      return;
    }
    std::vector<std::pair<CandAln, CandAln>> pairs;  // dedup :346-351
    for (size_t i : dedup_pairs(raw)) pairs.push_back(raw[i]);
    std::vector<double> jl;
    for (auto& pr : pairs) jl.push_back(pr.first.log_lk + pr.second.log_lk + pair_distance_log_lk(pr.first, pr.second, im));
    size_t bi;
    double se;
    mq.select_best(jl, bi, se);
    const CandAln& bx = pairs[bi].first;
    const CandAln& by = pairs[bi].second;
    bool bxo = b1 >= 0 && same_locus(bx, c1[b1]);
    bool byo = b2 >= 0 && same_locus(by, c2[b2]);
    uint8_t q1, q2;
    double p1, p2;
    mapq_for_winning_pair(r1, r2, bx, by, se, lk1, lk2, m1, m2, bxo, byo, q1, p1, q2, p2);
    o1 = to_record(bx, q1, p1);
    o2 = to_record(by, q2, p2);
    confident = bxo && byo;
  }
  static bool is_fr_orientation(const OutRec& r1, const OutRec& r2) {  // paired.rs:572-575
    const OutRec& up = r1.pos <= r2.pos ? r1 : r2;
    const OutRec& dn = r1.pos <= r2.pos ? r2 : r1;
    return up.strand == Forward && dn.strand != Forward;
  }
  static uint64_t span_of(const OutRec& r) { return cigar_reference_span(r.cigar); }
  // paired.rs:154-234 map_pair_readonly
  void map_pair_readonly(const Read& r1, const Read& r2, const InsertModel& im, OutRec& o1, OutRec& o2, bool& has_obs,
                         double& obs, Stats& st) const {
    std::vector<CandAln> c1, c2;
    std::vector<double> m1, m2;
    map_candidates(r1, c1, m1, st);
    map_candidates(r2, c2, m2, st);
    bool conf = false;
    if (c1.empty() && c2.empty()) {
      o1 = unmapped(r1);
      o2 = unmapped(r2);
    } else {
      joint_map(r1, r2, c1, m1, c2, m2, im, o1, o2, conf, st);
    }
    o1.is_paired = o2.is_paired = true;
    o1.read_number = 1;
    o2.read_number = 2;
    bool m1d = o1.pos != 0, m2d = o2.pos != 0;
    o1.mate_unmapped = !m2d;
    o2.mate_unmapped = !m1d;
    o1.mate_strand = m2d ? (int)o2.strand : -1;
    o2.mate_strand = m1d ? (int)o1.strand : -1;
    has_obs = false;
    if (m1d && m2d && g->contigs[o1.contig].name == g->contigs[o2.contig].name) {
      const OutRec& up = o1.pos <= o2.pos ? o1 : o2;
      const OutRec& dn = o1.pos <= o2.pos ? o2 : o1;
      uint64_t dend = dn.pos + span_of(dn) - 1;
      int64_t ins = std::max<int64_t>((int64_t)dend - (int64_t)up.pos + 1, 0);
      bool within = std::fabs((double)ins - im.mean) < 5.0 * im.sd;
      bool proper = within && is_fr_orientation(o1, o2);
      o1.is_proper = o2.is_proper = proper;
      o1.mate_contig = o2.contig;
      o2.mate_contig = o1.contig;
      o1.mate_pos = o2.pos;
      o2.mate_pos = o1.pos;
      o1.has_mate_pos = o2.has_mate_pos = true;
      o1.tlen = o1.pos <= o2.pos ? ins : -ins;
      o2.tlen = o2.pos <= o1.pos ? ins : -ins;
      o1.has_tlen = o2.has_tlen = true;
      if (proper && conf) {
        has_obs = true;
        obs = (double)ins;
      }
    }
  }
};

// sam.rs:126-143 sam_flag
static uint16_t sam_flag(const OutRec& r) {
  uint16_t f = 0;
  bool unmapped = r.pos == 0;
  if (r.is_paired) f |= 0x01;
  if (r.is_proper) f |= 0x02;
  if (unmapped) f |= 0x04;
  if (r.mate_unmapped) f |= 0x08;
  if (!unmapped && r.strand != Forward) f |= 0x10;
  if (r.mate_strand == 1) f |= 0x20;
  if (r.read_number == 1) f |= 0x40;
  if (r.read_number == 2) f |= 0x80;
  return f;
}

static void emit_records(const std::vector<OutRec>& recs, orc_records* out) {
  out->n = recs.size();
  out->rec = (orc_record*)calloc(std::max<size_t>(recs.size(), 1), sizeof(orc_record));
  uint64_t nc = 0;
  for (auto& r : recs) nc += r.cigar.size();
  out->cigar = (uint32_t*)calloc(std::max<uint64_t>(nc, 1), sizeof(uint32_t));
  out->n_cigar = nc;
  uint64_t co = 0;
  for (size_t i = 0; i < recs.size(); i++) {
    const OutRec& r = recs[i];
    orc_record& o = out->rec[i];
    o.pos = r.pos;
    o.mate_pos = r.has_mate_pos ? r.mate_pos : 0;
    o.tlen = r.has_tlen ? r.tlen : 0;
    o.log_lk = r.log_lk;
    o.posterior = r.posterior;
    o.contig = r.pos == 0 ? -1 : r.contig;
    o.mate_contig = r.mate_contig;
    o.cigar_off = (uint32_t)co;
    o.cigar_n = (uint32_t)r.cigar.size();
    for (auto& c : r.cigar) out->cigar[co++] = (c.n << 4) | (uint32_t)c.kind;
    o.flag = sam_flag(r);
    o.mapq = r.mapq;
    o.strand = r.strand == Forward ? 0 : 1;
    o.seq_rc = r.seq_rc ? 1 : 0;
  }
}

static Read read_at(const orc_reads* rs, uint64_t i) {
  Read r;
  r.seq = rs->seq + rs->seq_off[i];
  r.len = rs->seq_off[i + 1] - rs->seq_off[i];
  r.qual = rs->qual + rs->qual_off[i];
  r.qlen = rs->qual_off[i + 1] - rs->qual_off[i];
  return r;
}

// ---------------------------------------------------------------------------
// io/fastx.rs
// ---------------------------------------------------------------------------
// Rust's str view of a line: UTF-8 (BufRead::read_line into a String rejects
// anything else), char::is_whitespace for trim / split_whitespace.
static int utf8_len_at(const uint8_t* p, size_t n, size_t i) {  // > 0 valid length, else -(invalid prefix)
  uint8_t b0 = p[i];
  if (b0 < 0x80) return 1;
  int need;
  uint8_t lo = 0x80, hi = 0xBF;
  if (b0 >= 0xC2 && b0 <= 0xDF) need = 1;
  else if (b0 >= 0xE0 && b0 <= 0xEF) { need = 2; if (b0 == 0xE0) lo = 0xA0; if (b0 == 0xED) hi = 0x9F; }
  else if (b0 >= 0xF0 && b0 <= 0xF4) { need = 3; if (b0 == 0xF0) lo = 0x90; if (b0 == 0xF4) hi = 0x8F; }
  else return -1;
  for (int k = 1; k <= need; k++) {
    if (i + k >= n) return -k;
    uint8_t b = p[i + k];
    if (b < (k == 1 ? lo : 0x80) || b > (k == 1 ? hi : 0xBF)) return -k;
  }
  return need + 1;
}
static uint32_t utf8_code(const uint8_t* p, int len) {
  if (len == 1) return p[0];
  uint32_t c = p[0] & (0x7F >> len);
  for (int k = 1; k < len; k++) c = (c << 6) | (p[k] & 0x3F);
  return c;
}
static bool unicode_ws(uint32_t c) {
  return (c >= 9 && c <= 13) || c == 32 || c == 0x85 || c == 0xA0 || c == 0x1680 || (c >= 0x2000 && c <= 0x200A) ||
         c == 0x2028 || c == 0x2029 || c == 0x202F || c == 0x205F || c == 0x3000;
}
static std::vector<std::pair<size_t, int>> utf8_chars(const std::string& s) {  // (offset, length); s valid
  std::vector<std::pair<size_t, int>> v;
  const uint8_t* p = (const uint8_t*)s.data();
  for (size_t i = 0; i < s.size();) {
    int l = utf8_len_at(p, s.size(), i);
    v.push_back({i, l});
    i += l;
  }
  return v;
}
static bool char_ws(const std::string& s, std::pair<size_t, int> c) {
  return unicode_ws(utf8_code((const uint8_t*)s.data() + c.first, c.second));
}
static std::string trim(const std::string& s) {  // str::trim
  auto cs = utf8_chars(s);
  size_t a = 0, b = cs.size();
  while (a < b && char_ws(s, cs[a])) a++;
  while (b > a && char_ws(s, cs[b - 1])) b--;
  if (a == b) return std::string();
  return s.substr(cs[a].first, cs[b - 1].first + cs[b - 1].second - cs[a].first);
}
static std::string first_token(const std::string& s) {  // split_whitespace().next().unwrap_or("")
  auto cs = utf8_chars(s);
  size_t a = 0;
  while (a < cs.size() && char_ws(s, cs[a])) a++;
  size_t b = a;
  while (b < cs.size() && !char_ws(s, cs[b])) b++;
  if (a == b) return std::string();
  return s.substr(cs[a].first, cs[b - 1].first + cs[b - 1].second - cs[a].first);
}
static bool valid_utf8(const std::string& s) {
  const uint8_t* p = (const uint8_t*)s.data();
  for (size_t i = 0; i < s.size();) {
    int l = utf8_len_at(p, s.size(), i);
    if (l < 0) return false;
    i += l;
  }
  return true;
}
// String::from_utf8_lossy: every maximal invalid subpart becomes U+FFFD
static void append_lossy(std::string& out, const uint8_t* p, size_t n) {
  for (size_t i = 0; i < n;) {
    int l = utf8_len_at(p, n, i);
    if (l > 0) {
      out.append((const char*)p + i, l);
      i += l;
    } else {
      out += "\xEF\xBF\xBD";
      i += -l;
    }
  }
}
struct LineReader {
  FILE* f;
  bool bad = false;  // a line was not UTF-8: the reference's read_line fails (io::ErrorKind::InvalidData)
  bool read_line(std::string& line) {  // keeps the trailing newline like BufRead::read_line
    line.clear();
    if (bad) return false;
    int c;
    bool any = false;
    while ((c = fgetc(f)) != EOF) {
      any = true;
      line.push_back((char)c);
      if (c == '\n') break;
    }
    if (any && !valid_utf8(line)) {
      bad = true;
      return false;
    }
    return any;
  }
};
struct PReads {
  std::vector<std::string> ids;
  std::vector<std::vector<uint8_t>> seqs, quals;
};
// fastx.rs:255-305 FastqParser::next
static bool parse_fastq(FILE* f, uint8_t offset, PReads& out) {
  LineReader lr{f};
  std::string buf;
  for (;;) {
    if (!lr.read_line(buf)) break;
    size_t a = 0;
    while (a < buf.size() && buf[a] == '@') a++;
    std::string id = first_token(buf.substr(a));
    std::vector<uint8_t> seq, qual;
    for (;;) {
      if (!lr.read_line(buf)) break;
      std::string t = trim(buf);
      if (!t.empty() && t[0] == '+') break;
      seq.insert(seq.end(), t.begin(), t.end());
    }
    for (;;) {
      if (!lr.read_line(buf)) break;
      std::string t = trim(buf);
      qual.insert(qual.end(), t.begin(), t.end());
      if (qual.size() >= seq.size()) break;
    }
    for (auto& b : seq)  // types.rs:12-15 DnaSeq::from_ascii
      if (b >= 'a' && b <= 'z') b = (uint8_t)(b - 32);
    for (auto& q : qual) q = q >= offset ? (uint8_t)(q - offset) : 0;  // types.rs:57-60
    out.ids.push_back(id);
    out.seqs.push_back(seq);
    out.quals.push_back(qual);
  }
  return !lr.bad;
}
// fastx.rs:194-236 FastaAsReadsParser::next
static bool parse_fasta_reads(FILE* f, PReads& out) {
  LineReader lr{f};
  std::string line, pending;
  bool has_pending = false;
  for (;;) {
    std::string header;
    if (has_pending) {
      header = pending;
      has_pending = false;
    } else {
      bool found = false;
      while (lr.read_line(line))
        if (!line.empty() && line[0] == '>') {
          found = true;
          break;
        }
      if (!found) break;
      header = line;
    }
    size_t a = 0;
    while (a < header.size() && header[a] == '>') a++;
    std::string id = first_token(header.substr(a));
    std::vector<uint8_t> seq;
    while (lr.read_line(line)) {
      std::string t = trim(line);
      if (!t.empty() && t[0] == '>') {
        pending = line;
        has_pending = true;
        break;
      }
      seq.insert(seq.end(), t.begin(), t.end());
    }
    for (auto& b : seq)
      if (b >= 'a' && b <= 'z') b = (uint8_t)(b - 32);
    out.ids.push_back(id);
    out.quals.push_back(std::vector<uint8_t>(seq.size(), 40));  // fastx.rs:157 FASTA_DEFAULT_QUAL
    out.seqs.push_back(seq);
  }
  return !lr.bad;
}
static inline uint8_t ascii_lower(uint8_t c) { return (c >= 'A' && c <= 'Z') ? (uint8_t)(c + 32) : c; }
// fastx.rs:97-117 strip_adapter
static void strip_adapter(std::vector<uint8_t>& seq, std::vector<uint8_t>& qual, const std::string& adapter) {
  size_t L = seq.size();
  size_t max_overlap = std::min(L, adapter.size());
  if (max_overlap < 6) return;
  for (size_t ov = max_overlap; ov >= 6; ov--) {
    size_t mm = 0;
    for (size_t i = 0; i < ov; i++)
      if (ascii_lower(seq[L - ov + i]) != ascii_lower((uint8_t)adapter[i])) mm++;
    if (mm <= ov / 10) {
      size_t keep = L - ov;
      seq.resize(keep);
      if (qual.size() > keep) qual.resize(keep);
      return;
    }
  }
}
// fastx.rs:77-81 strip_mate_suffix
static void strip_mate_suffix(std::string& id) {
  if (id.size() >= 2 && id[id.size() - 2] == '/' && (id.back() == '1' || id.back() == '2')) id.resize(id.size() - 2);
}

}  // namespace

// ===========================================================================
// C API
// ===========================================================================
extern "C" {

void* orc_genome_new(void) { return new Genome(); }
void orc_genome_free(void* g) { delete (Genome*)g; }
int orc_genome_add_contig(void* gp, const char* name, const uint8_t* seq, uint64_t len) {
  Genome* g = (Genome*)gp;
  g->contigs.push_back(contig_from_ascii(name, seq, len));
  g->total_length += len;  // genome.rs:119 (sum of contig lengths)
  return 0;
}
// genome.rs:101-147 GenomeIndexBuilder::add_fasta + FastaReader (:150-213)
int orc_genome_add_fasta(void* gp, const char* path) {
  FILE* f = fopen(path, "rb");
  if (!f) return -1;
  LineReader lr{f};
  std::string line, pending;
  bool has_pending = false;
  for (;;) {
    std::string header;
    if (has_pending) {
      header = pending;
      has_pending = false;
    } else {
      bool found = false;
      while (lr.read_line(line))
        if (!line.empty() && line[0] == '>') {
          found = true;
          break;
        }
      if (!found) break;
      header = line;
    }
    size_t a = 0;
    while (a < header.size() && header[a] == '>') a++;
    std::string id = first_token(header.substr(a));
    std::string name = first_token(id);
    std::vector<uint8_t> seq;
    while (lr.read_line(line)) {
      std::string t = trim(line);
      if (!t.empty() && t[0] == '>') {
        pending = line;
        has_pending = true;
        break;
      }
      seq.insert(seq.end(), t.begin(), t.end());
    }
    orc_genome_add_contig(gp, name.c_str(), seq.data(), seq.size());
  }
  fclose(f);
  return 0;
}
uint64_t orc_genome_total_length(const void* g) { return ((const Genome*)g)->total_length; }
int orc_genome_n_contigs(const void* g) { return (int)((const Genome*)g)->contigs.size(); }
uint64_t orc_genome_contig_length(const void* g, int cid) { return ((const Genome*)g)->contigs[cid].length; }
const char* orc_genome_contig_name(const void* g, int cid) { return ((const Genome*)g)->contigs[cid].name.c_str(); }
uint64_t orc_genome_slice(const void* g, int cid, uint64_t start, uint64_t end, uint8_t* out) {
  std::vector<uint8_t> v;
  contig_slice(((const Genome*)g)->contigs[cid], start, end, v);
  if (!v.empty()) memcpy(out, v.data(), v.size());
  return v.size();
}

uint32_t orc_hash_kmer(const uint8_t* kmer, uint64_t len) { return hash_kmer(kmer, len); }
uint64_t orc_pack_position(uint64_t c, uint64_t p, int s) { return pack_position(c, p, s ? Reverse : Forward); }
void orc_unpack_position(uint64_t packed, uint64_t* c, uint64_t* p, int* s) {
  Strand st;
  unpack_position(packed, *c, *p, st);
  *s = st == Forward ? 0 : 1;
}
int orc_is_acgt_only(const uint8_t* k, uint64_t len) { return is_acgt_only(k, len) ? 1 : 0; }
void orc_set_progress(int on) { g_progress = on; }
void orc_set_index_threads(int n) { g_index_threads = n > 0 ? (unsigned)n : 1u; }
void* orc_hash_build(const void* g, int k, int stride, int max_count) {
  return hash_build(*(const Genome*)g, (size_t)k, (size_t)stride, (size_t)max_count, g_index_threads);
}
void orc_hash_free(void* h) { delete (HashTable*)h; }
uint64_t orc_hash_lookup(const void* h, const uint8_t* kmer, uint64_t len, const uint64_t** out) {
  uint64_t c;
  *out = ((const HashTable*)h)->lookup(kmer, len, c);
  return c;
}
uint64_t orc_hash_total_entries(const void* h) { return ((const HashTable*)h)->total_entries; }
uint64_t orc_hash_n_keys(const void* h) { return ((const HashTable*)h)->sorted_keys.size(); }
uint64_t orc_hash_key_at(const void* hp, uint64_t i, uint32_t* hash, const uint64_t** bucket) {
  const HashTable* h = (const HashTable*)hp;
  *hash = h->sorted_keys[i];
  *bucket = h->positions.data() + h->key_off[i];
  return h->key_off[i + 1] - h->key_off[i];
}
int orc_mix_is_bijective(void) {
  // Injective over every 30-bit 15-mer code (the whole key space of the table).
  std::vector<uint64_t> seen((1ull << 32) / 64, 0);
  for (uint32_t x = 0; x < (1u << 30); x++) {
    uint32_t y = mix(x);
    uint64_t& w = seen[y >> 6];
    uint64_t bit = 1ull << (y & 63);
    if (w & bit) return 0;
    w |= bit;
  }
  return 1;
}

int orc_sw_align(uint32_t band, uint32_t gop, uint32_t gep, const uint8_t* read, uint64_t m, const uint8_t* qual,
                 uint64_t qlen, const uint8_t* ref, uint64_t n, int64_t hint, int32_t* score, uint32_t* cigar,
                 uint32_t cap, uint32_t* n_ops, uint64_t* ref_offset, uint64_t* cells) {
  SwAligner sw(band, gop, gep);
  std::vector<CigarOp> c;
  uint64_t off = 0, cl = 0;
  int32_t s = sw.align(read, m, qual, qlen, ref, n, hint, c, off, &cl);
  *score = s;
  *ref_offset = off;
  *n_ops = (uint32_t)c.size();
  if (cells) *cells = cl;
  if (c.size() > cap) return -1;
  for (size_t i = 0; i < c.size(); i++) cigar[i] = (c[i].n << 4) | (uint32_t)c[i].kind;
  return 0;
}
int orc_mismatch_penalty(uint8_t q) { return SwAligner(30, 40, 3).mismatch_penalty(q); }
int orc_homopolymer_gap_open(int gap_open, uint16_t run) {
  SwAligner s(30, 40, 3);
  s.gap_open = gap_open;
  return s.homopolymer_gap_open(run);
}

double orc_logadd(double a, double b) { return logadd(a, b); }
double orc_phred_n_choose_m(int64_t n, int64_t m) { return phred_n_choose_m(n, m); }
double orc_effective_alphabet_size(double gc) { return effective_alphabet_size(gc); }
double orc_missing_locus_probability(double sub_rate, const uint8_t* q, uint64_t len, const double* mask,
                                     uint64_t mlen) {
  Mapq mq(sub_rate, 15, 1);
  return mq.missing_locus_probability(q, len, mask, mlen);
}
static std::vector<CigarOp> decode_cigar(const uint32_t* w, uint32_t n) {
  std::vector<CigarOp> c;
  for (uint32_t i = 0; i < n; i++) c.push_back({(CigarKind)(w[i] & 15), w[i] >> 4});
  return c;
}
void orc_finalize_mapq(double sub_rate, uint64_t gsize, double sum_exp, const uint8_t* rq, uint64_t rql,
                       const uint32_t* cigar, uint32_t n_ops, const uint8_t* qb, uint64_t qlen, const uint8_t* qq,
                       uint64_t qqlen, const uint8_t* rw, uint64_t rwlen, const double* mask, uint64_t mlen,
                       uint8_t* mapq, double* posterior) {
  Mapq mq(sub_rate, 15, gsize);
  Read r{qb, qlen, rq, rql};
  mq.finalize_mapq(sum_exp, r, decode_cigar(cigar, n_ops), qb, qlen, qq, qqlen, rw, rwlen, mask, mlen, *mapq,
                   *posterior);
}
double orc_log_entropy(double sub_rate, uint64_t gsize, const uint32_t* cigar, uint32_t n_ops, const uint8_t* qb,
                       uint64_t qlen, const uint8_t* qq, uint64_t qqlen, const uint8_t* rw, uint64_t rwlen) {
  Mapq mq(sub_rate, 15, gsize);
  return mq.log_entropy(decode_cigar(cigar, n_ops), qb, qlen, qq, qqlen, rw, rwlen);
}
uint64_t orc_shortlist_indices(const double* post, uint64_t n, uint64_t* out) {
  std::vector<size_t> idx = Mapper::shortlist_indices(std::vector<double>(post, post + n));
  for (size_t i = 0; i < idx.size(); i++) out[i] = idx[i];
  return idx.size();
}
uint64_t orc_dedup_pairs(const uint64_t* loci, uint64_t n, uint64_t* out) {
  std::vector<std::pair<CandAln, CandAln>> raw(n);
  for (uint64_t i = 0; i < n; i++) {
    const uint64_t* l = loci + 6 * i;  // contig, position, strand of mate 1, then of mate 2
    raw[i].first.contig = l[0];
    raw[i].first.position = l[1];
    raw[i].first.strand = l[2] ? Reverse : Forward;
    raw[i].second.contig = l[3];
    raw[i].second.position = l[4];
    raw[i].second.strand = l[5] ? Reverse : Forward;
  }
  std::vector<size_t> kept = Mapper::dedup_pairs(raw);
  for (size_t i = 0; i < kept.size(); i++) out[i] = kept[i];
  return kept.size();
}
void orc_mapq_for_winning_pair(const void* mp, double pair_se, int bxo, int byo, const orc_pair_mate* a,
                               const orc_pair_mate* b, uint8_t* q, double* p) {
  const Mapper* m = (const Mapper*)mp;
  const orc_pair_mate* in[2] = {a, b};
  Read r[2];
  CandAln c[2];
  std::vector<double> lk[2], mask[2];
  for (int k = 0; k < 2; k++) {
    r[k] = Read{in[k]->seq, in[k]->len, in[k]->qual, in[k]->qlen};
    c[k].cigar = decode_cigar(in[k]->cigar, in[k]->n_ops);
    c[k].ref_window.assign(in[k]->ref_window, in[k]->ref_window + in[k]->rwlen);
    c[k].rc = false;
    lk[k].assign(in[k]->lk, in[k]->lk + in[k]->nlk);
    mask[k].assign(in[k]->mask, in[k]->mask + in[k]->mlen);
  }
  m->mapq_for_winning_pair(r[0], r[1], c[0], c[1], pair_se, lk[0], lk[1], mask[0], mask[1], bxo != 0, byo != 0, q[0],
                           p[0], q[1], p[1]);
}
void orc_insert_update(double* mean, double* sd, uint64_t* n_obs, double observed) {
  InsertModel im{*mean, *sd, *n_obs};
  im.update(observed);
  *mean = im.mean;
  *sd = im.sd;
  *n_obs = im.n_observed;
}
double orc_insert_log_likelihood(double mean, double sd, double x) {
  InsertModel im{mean, sd, 0};
  return im.log_likelihood(x);
}

void orc_default_options(orc_options* o) {
  o->substitution_rate = 0.001;
  o->insert_size_mean = 250.0;
  o->insert_size_sd = 60.0;
  o->sv_prior_phred = 55.0;
  o->max_indel_len = 30;
  o->gap_open_phred = 40;
  o->gap_extend_phred = 3;
  o->threads = 1;
}
void* orc_mapper_new(const void* g, const void* h, const orc_options* o) {
  Options op;
  op.substitution_rate = o->substitution_rate;
  op.insert_size_mean = o->insert_size_mean;
  op.insert_size_sd = o->insert_size_sd;
  op.sv_prior_phred = o->sv_prior_phred;
  op.max_indel_len = o->max_indel_len;
  op.gap_open_phred = o->gap_open_phred;
  op.gap_extend_phred = o->gap_extend_phred;
  op.threads = o->threads;
  return new Mapper((const Genome*)g, (const HashTable*)h, op);
}
void orc_mapper_free(void* m) { delete (Mapper*)m; }
void orc_set_group_order(int mode) { g_group_order_mode = mode; }

int orc_seed_and_cluster(const void* mp, const uint8_t* seq, uint64_t L, const uint8_t* qual, uint64_t qlen,
                         uint64_t* hits, uint64_t hcap, uint64_t* n_hits, double* repeat_prob, uint64_t* cands,
                         uint64_t ccap, uint64_t* n_cands) {
  const Mapper* m = (const Mapper*)mp;
  Read r{seq, L, qual, qlen};
  std::vector<SeedHit> h;
  std::vector<double> rp;
  Stats st;
  find_seed_hits(r, *m->ht, *m->g, h, rp, st);
  std::vector<Candidate> c = cluster_hits(h, L, m->ht->k);
  *n_hits = h.size();
  *n_cands = c.size();
  for (size_t i = 0; i < rp.size(); i++) repeat_prob[i] = rp[i];
  for (size_t i = 0; i < h.size() && i < hcap; i++) {
    hits[5 * i] = h[i].read_pos;
    hits[5 * i + 1] = h[i].contig;
    hits[5 * i + 2] = h[i].ref_pos;
    hits[5 * i + 3] = h[i].strand;
    hits[5 * i + 4] = h[i].neighbour;
  }
  for (size_t i = 0; i < c.size() && i < ccap; i++) {
    cands[4 * i] = c[i].contig;
    cands[4 * i + 1] = c[i].position;
    cands[4 * i + 2] = c[i].strand;
    cands[4 * i + 3] = (uint64_t)c[i].score;
  }
  return (h.size() > hcap || c.size() > ccap) ? -1 : 0;
}

// lib.rs:84-104 (single-end loop; per-read dynamic scheduling in chunks of 64)
int orc_map_se(const void* mp, const orc_reads* reads, orc_records* out, orc_stats* stats) {
  const Mapper* m = (const Mapper*)mp;
  std::vector<OutRec> recs(reads->n);
  unsigned nt = std::max<uint32_t>(m->opt.threads, 1);
  std::atomic<uint64_t> next(0);
  std::vector<Stats> sts(nt);
  auto worker = [&](unsigned t) {
    for (;;) {
      uint64_t b = next.fetch_add(64);
      if (b >= reads->n) break;
      uint64_t e = std::min<uint64_t>(b + 64, reads->n);
      for (uint64_t i = b; i < e; i++) recs[i] = m->map_se(read_at(reads, i), sts[t]);
    }
  };
  if (nt == 1) {
    worker(0);
  } else {
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; t++) th.emplace_back(worker, t);
    for (auto& x : th) x.join();
  }
  Stats tot;
  for (auto& s : sts) tot.add(s);
  if (stats) {
    stats->lookups = tot.lookups;
    stats->positions = tot.positions;
    stats->filtered_hits = tot.filtered;
    stats->candidates = tot.candidates;
    stats->dp_jobs = tot.dp_jobs;
    stats->dp_cells = tot.dp_cells;
    stats->wide_jobs = tot.wide_jobs;
    stats->wide_cells = tot.wide_cells;
  }
  emit_records(recs, out);
  return 0;
}

// lib.rs:145-185 map_paired: 500 serial warm-up pairs, then 2000-pair batches
// against a frozen insert-size snapshot with in-order updates afterwards.
int orc_map_pe(const void* mp, const orc_reads* r1, const orc_reads* r2, orc_records* out, double* mean, double* sd,
               uint64_t* n_obs, orc_stats* stats) {
  const Mapper* m = (const Mapper*)mp;
  if (r1->n != r2->n) return -2;  // lib.rs:157-161
  uint64_t n = r1->n;
  InsertModel im{*mean, *sd, *n_obs};
  std::vector<OutRec> recs(2 * n);
  Stats st;
  uint64_t warm = std::min<uint64_t>(n, 500);
  for (uint64_t i = 0; i < warm; i++) {
    InsertModel snap = im;
    bool ho;
    double ob;
    m->map_pair_readonly(read_at(r1, i), read_at(r2, i), snap, recs[2 * i], recs[2 * i + 1], ho, ob, st);
    if (ho) im.update(ob);
  }
  unsigned nt = std::max<uint32_t>(m->opt.threads, 1);
  for (uint64_t b = warm; b < n; b += 2000) {
    uint64_t e = std::min<uint64_t>(b + 2000, n);
    InsertModel snap = im;
    std::vector<uint8_t> has(e - b, 0);
    std::vector<double> obs(e - b, 0.0);
    std::vector<Stats> sts(nt);
    std::atomic<uint64_t> next(b);
    auto worker = [&](unsigned t) {
      for (;;) {
        uint64_t c = next.fetch_add(16);
        if (c >= e) break;
        uint64_t ce = std::min<uint64_t>(c + 16, e);
        for (uint64_t i = c; i < ce; i++) {
          bool ho;
          double ob;
          m->map_pair_readonly(read_at(r1, i), read_at(r2, i), snap, recs[2 * i], recs[2 * i + 1], ho, ob, sts[t]);
          has[i - b] = ho;
          obs[i - b] = ob;
        }
      }
    };
    if (nt == 1) {
      worker(0);
    } else {
      std::vector<std::thread> th;
      for (unsigned t = 0; t < nt; t++) th.emplace_back(worker, t);
      for (auto& x : th) x.join();
    }
    for (auto& s : sts) st.add(s);
    for (uint64_t i = 0; i < e - b; i++)
      if (has[i]) im.update(obs[i]);
  }
  *mean = im.mean;
  *sd = im.sd;
  *n_obs = im.n_observed;
  if (stats) {
    stats->lookups = st.lookups;
    stats->positions = st.positions;
    stats->filtered_hits = st.filtered;
    stats->candidates = st.candidates;
    stats->dp_jobs = st.dp_jobs;
    stats->dp_cells = st.dp_cells;
    stats->wide_jobs = st.wide_jobs;
    stats->wide_cells = st.wide_cells;
  }
  emit_records(recs, out);
  return 0;
}
void orc_records_release(orc_records* r) {
  free(r->rec);
  free(r->cigar);
  r->rec = nullptr;
  r->cigar = nullptr;
  r->n = r->n_cigar = 0;
}

// fastx.rs:51-71 read_reads_files_with_preprocessing (one file)
int orc_read_reads(const char* path, int fasta, uint8_t offset, const char* adapter, orc_parsed_reads* out) {
  FILE* f = fopen(path, "rb");
  if (!f) return -1;
  PReads pr;
  const bool ok = fasta ? parse_fasta_reads(f, pr) : parse_fastq(f, offset, pr);
  fclose(f);
  if (!ok) return -2;
  for (size_t i = 0; i < pr.ids.size(); i++) {
    strip_mate_suffix(pr.ids[i]);
    if (adapter) strip_adapter(pr.seqs[i], pr.quals[i], adapter);
  }
  uint64_t n = pr.ids.size();
  out->n = n;
  uint64_t ts = 0, tq = 0, ti = 0;
  for (uint64_t i = 0; i < n; i++) {
    ts += pr.seqs[i].size();
    tq += pr.quals[i].size();
    ti += pr.ids[i].size() + 1;
  }
  out->seq = (uint8_t*)malloc(ts + 1);
  out->qual = (uint8_t*)malloc(tq + 1);
  out->ids = (char*)malloc(ti + 1);
  out->seq_off = (uint64_t*)malloc((n + 1) * 8);
  out->qual_off = (uint64_t*)malloc((n + 1) * 8);
  out->id_off = (uint64_t*)malloc((n + 1) * 8);
  uint64_t a = 0, b = 0, c = 0;
  for (uint64_t i = 0; i < n; i++) {
    out->seq_off[i] = a;
    out->qual_off[i] = b;
    out->id_off[i] = c;
    if (!pr.seqs[i].empty()) memcpy(out->seq + a, pr.seqs[i].data(), pr.seqs[i].size());
    if (!pr.quals[i].empty()) memcpy(out->qual + b, pr.quals[i].data(), pr.quals[i].size());
    memcpy(out->ids + c, pr.ids[i].c_str(), pr.ids[i].size() + 1);
    a += pr.seqs[i].size();
    b += pr.quals[i].size();
    c += pr.ids[i].size() + 1;
  }
  out->seq_off[n] = a;
  out->qual_off[n] = b;
  out->id_off[n] = c;
  return 0;
}
void orc_parsed_reads_free(orc_parsed_reads* p) {
  free(p->seq);
  free(p->qual);
  free(p->ids);
  free(p->seq_off);
  free(p->qual_off);
  free(p->id_off);
  memset(p, 0, sizeof(*p));
}

// sam.rs:37-58 format_sam, :83-107 header; QUAL is raw Phred bytes (sam.rs:52).
char* orc_format_sam(const void* gp, const orc_parsed_reads* rd, const orc_records* recs, const char* cmd,
                     const char* rg, uint64_t* len) {
  const Genome* g = (const Genome*)gp;
  std::string out = "@HD\tVN:1.6\tSO:unsorted\n";
  for (auto& c : g->contigs) out += "@SQ\tSN:" + c.name + "\tLN:" + std::to_string(c.length) + "\n";
  out += std::string("@PG\tID:probebwa\tPN:probebwa\tVN:0.1.0\tCL:") + (cmd ? cmd : "") + "\n";
  std::string rg_id;
  if (rg) {  // sam.rs:11-32 ReadGroup::parse
    std::string spec(rg), hdr = "@RG";
    size_t s = 0;
    while (true) {
      size_t e = spec.find(',', s);
      std::string pair = spec.substr(s, e == std::string::npos ? std::string::npos : e - s);
      size_t colon = pair.find(':');
      std::string tag = pair.substr(0, colon), val = colon == std::string::npos ? "" : pair.substr(colon + 1);
      if (tag == "ID") rg_id = val;
      hdr += "\t" + tag + ":" + val;
      if (e == std::string::npos) break;
      s = e + 1;
    }
    out += hdr + "\n";
  }
  const char* ops = "MIDNSHP=X";
  for (uint64_t i = 0; i < recs->n; i++) {
    const orc_record& r = recs->rec[i];
    out += std::string(rd->ids + rd->id_off[i]);
    out += "\t" + std::to_string(r.flag);
    out += "\t" + (r.contig < 0 ? std::string("*") : g->contigs[r.contig].name);
    out += "\t" + std::to_string(r.pos);
    out += "\t" + std::to_string(r.mapq);
    out += "\t";
    if (r.cigar_n == 0) out += "*";
    for (uint32_t c = 0; c < r.cigar_n; c++) {
      uint32_t w = recs->cigar[r.cigar_off + c];
      out += std::to_string(w >> 4);
      out += ops[w & 15];
    }
    out += "\t" + (r.mate_contig < 0 ? std::string("*") : g->contigs[r.mate_contig].name);
    out += "\t" + std::to_string(r.mate_pos);
    out += "\t" + std::to_string(r.tlen);
    out += "\t";
    const uint8_t* s = rd->seq + rd->seq_off[i];
    uint64_t sl = rd->seq_off[i + 1] - rd->seq_off[i];
    const uint8_t* q = rd->qual + rd->qual_off[i];
    uint64_t ql = rd->qual_off[i + 1] - rd->qual_off[i];
    // sam.rs:50-51: SEQ and QUAL through String::from_utf8_lossy
    std::vector<uint8_t> sv = r.seq_rc ? reverse_complement(s, sl) : std::vector<uint8_t>(s, s + sl);
    std::vector<uint8_t> qv(q, q + ql);
    if (r.seq_rc) std::reverse(qv.begin(), qv.end());
    append_lossy(out, sv.data(), sv.size());
    out += "\t";
    append_lossy(out, qv.data(), qv.size());
    if (rg) out += "\tRG:Z:" + rg_id;
    out += "\n";
  }
  char* t = (char*)malloc(out.size() + 1);
  memcpy(t, out.data(), out.size());
  t[out.size()] = 0;
  *len = out.size();
  return t;
}
void orc_free_text(char* t) { free(t); }

}  // extern "C"
