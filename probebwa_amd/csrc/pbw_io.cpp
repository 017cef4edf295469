// Host I/O of probebwa_amd: the reference's persisted index formats (bincode-1
// `.stidx` / `.sthash`), read files (FASTQ / FASTA, plain or gzip), and the
// output side (SAM text, BGZF-compressed BAM with its BAI index).
//
// Formats follow the reference's serde types and bincode 1.3's default
// options (Cargo.toml: bincode = "1.3"): little-endian fixed-width integers,
// usize as u64, every String / Vec / map prefixed by its u64 length, struct
// fields in declaration order, a map's entries as (key, value) pairs.
#include <zlib.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/probebwa_amd.h"
#include "pbw_internal.h"

namespace pbw {

namespace {

// ---------------------------------------------------------------------------
// bincode-1 primitives
// ---------------------------------------------------------------------------
struct BinReader {
  FILE* f = nullptr;
  uint64_t left = 0;  // bytes not yet read: a length prefix may not claim more
  std::string err;
  explicit BinReader(const char* path) {
    f = fopen(path, "rb");
    if (!f) {
      err = std::string("cannot open ") + path;
      return;
    }
    fseek(f, 0, SEEK_END);
    long sz = ftell(f);
    fseek(f, 0, SEEK_SET);
    left = sz > 0 ? (uint64_t)sz : 0;
  }
  ~BinReader() {
    if (f) fclose(f);
  }
  bool bytes(void* p, uint64_t n) {
    if (!err.empty()) return false;
    if (n > left || fread(p, 1, n, f) != n) {
      err = "truncated file";
      return false;
    }
    left -= n;
    return true;
  }
  bool u32(uint32_t& v) {
    uint8_t b[4];
    if (!bytes(b, 4)) return false;
    v = (uint32_t)b[0] | ((uint32_t)b[1] << 8) | ((uint32_t)b[2] << 16) | ((uint32_t)b[3] << 24);
    return true;
  }
  bool u64(uint64_t& v) {
    uint8_t b[8];
    if (!bytes(b, 8)) return false;
    v = 0;
    for (int i = 7; i >= 0; i--) v = (v << 8) | b[i];
    return true;
  }
  // a length prefix of `n` elements of `elem` bytes each, checked against the file size
  bool len(uint64_t& n, uint64_t elem) {
    if (!u64(n)) return false;
    if (elem && n > left / elem) {
      err = "length prefix past the end of the file";
      return false;
    }
    return true;
  }
  bool str(std::string& s) {
    uint64_t n;
    if (!len(n, 1)) return false;
    s.resize(n);
    return n == 0 || bytes(&s[0], n);
  }
  bool at_end() const { return left == 0; }
};

struct BinWriter {
  FILE* f = nullptr;
  std::string err;
  explicit BinWriter(const char* path) {
    f = fopen(path, "wb");
    if (!f) err = std::string("cannot create ") + path;
  }
  bool close() {
    if (f && fclose(f) != 0 && err.empty()) err = "write failed";
    f = nullptr;
    return err.empty();
  }
  ~BinWriter() {
    if (f) fclose(f);
  }
  void bytes(const void* p, uint64_t n) {
    if (err.empty() && n && fwrite(p, 1, n, f) != n) err = "write failed";
  }
  void u32(uint32_t v) {
    uint8_t b[4] = {(uint8_t)v, (uint8_t)(v >> 8), (uint8_t)(v >> 16), (uint8_t)(v >> 24)};
    bytes(b, 4);
  }
  void u64(uint64_t v) {
    uint8_t b[8];
    for (int i = 0; i < 8; i++) b[i] = (uint8_t)(v >> (8 * i));
    bytes(b, 8);
  }
  void str(const std::string& s) {
    u64(s.size());
    bytes(s.data(), s.size());
  }
};

}  // namespace

// HashTable (hashtable.rs:53-64): table: FxHashMap<u32, (u32, u32)>, positions:
// Vec<u64>, k: usize, seed_stride: usize, total_entries: usize.
bool read_sthash(const char* path, HostTable& t, std::string& err) {
  BinReader r(path);
  uint64_t nb = 0, np = 0;
  if (r.len(nb, 12)) {
    t.key.resize(nb);
    t.off.resize(nb);
    t.cnt.resize(nb);
    for (uint64_t i = 0; i < nb && r.err.empty(); i++) {
      r.u32(t.key[i]);
      r.u32(t.off[i]);
      r.u32(t.cnt[i]);
    }
  }
  if (r.err.empty() && r.len(np, 8)) {
    t.pos.resize(np);
    std::vector<uint8_t> buf;
    const uint64_t CH = 1u << 20;
    for (uint64_t i = 0; i < np && r.err.empty(); i += CH) {
      const uint64_t m = std::min(CH, np - i);
      buf.resize(m * 8);
      if (!r.bytes(buf.data(), m * 8)) break;
      for (uint64_t j = 0; j < m; j++) {
        uint64_t v = 0;
        for (int b = 7; b >= 0; b--) v = (v << 8) | buf[j * 8 + b];
        t.pos[i + j] = v;
      }
    }
  }
  if (r.err.empty()) {
    r.u64(t.k);
    r.u64(t.stride);
    r.u64(t.total_entries);
  }
  if (r.err.empty() && !r.at_end()) r.err = "trailing bytes after the table";
  if (r.err.empty() && t.k != (uint64_t)K) r.err = "k = " + std::to_string(t.k) + " (this path indexes 15-mers)";
  if (!r.err.empty()) {
    err = std::string(path) + ": " + r.err;
    return false;
  }
  return true;
}

bool write_sthash(const char* path, const HostTable& t, std::string& err) {
  BinWriter w(path);
  w.u64(t.key.size());
  for (size_t i = 0; i < t.key.size(); i++) {
    w.u32(t.key[i]);
    w.u32(t.off[i]);
    w.u32(t.cnt[i]);
  }
  w.u64(t.pos.size());
  std::vector<uint8_t> buf;
  const size_t CH = 1u << 20;
  for (size_t i = 0; i < t.pos.size(); i += CH) {
    const size_t m = std::min(CH, t.pos.size() - i);
    buf.resize(m * 8);
    for (size_t j = 0; j < m; j++)
      for (int b = 0; b < 8; b++) buf[j * 8 + b] = (uint8_t)(t.pos[i + j] >> (8 * b));
    w.bytes(buf.data(), buf.size());
  }
  w.u64(t.k);
  w.u64(t.stride);
  w.u64(t.total_entries);
  if (!w.close()) {
    err = std::string(path) + ": " + w.err;
    return false;
  }
  return true;
}

}  // namespace pbw

using namespace pbw;

extern "C" {

// GenomeIndex (genome.rs:70-76 / types.rs:107-121): species, assembly, contigs
// (name, packed 2-bit bases four per byte with base p at bits 2 (p % 4), N runs
// as half-open (start, end) pairs, length), total_length.  Contig::slice
// (genome.rs:47-68) decodes: packed bits, then 'N' over every run.
int pbw_genome_load_stidx(const char* path, pbw_genome** out) {
  if (!path || !out) return PBW_ERR_ARG;
  *out = nullptr;
  BinReader r(path);
  std::string species, assembly;
  uint64_t nctg = 0;
  r.str(species);
  r.str(assembly);
  r.len(nctg, 40);  // a contig takes at least 40 bytes (four length prefixes + length)
  if (!r.err.empty()) {
    set_global_error(std::string(path) + ": " + r.err);
    return PBW_ERR_IO;
  }
  pbw_genome* g = nullptr;
  if (pbw_genome_create(&g) != PBW_OK) return PBW_ERR_ARG;
  static const char BASES[4] = {'A', 'C', 'G', 'T'};
  std::vector<uint8_t> packed, ascii;
  uint64_t total = 0;
  for (uint64_t c = 0; c < nctg && r.err.empty(); c++) {
    std::string name;
    uint64_t np = 0, nr = 0, length = 0;
    if (!r.str(name) || !r.len(np, 1)) break;
    packed.resize(np);
    if (np && !r.bytes(packed.data(), np)) break;
    if (!r.len(nr, 8)) break;
    std::vector<std::pair<uint32_t, uint32_t>> runs(nr);
    for (uint64_t i = 0; i < nr && r.err.empty(); i++) {
      r.u32(runs[i].first);
      r.u32(runs[i].second);
    }
    if (!r.u64(length)) break;
    if (np < (length + 3) / 4) {
      r.err = "contig " + name + ": fewer packed bytes than its length needs";
      break;
    }
    ascii.resize(length);
    for (uint64_t p = 0; p < length; p++) ascii[p] = (uint8_t)BASES[(packed[p >> 2] >> (2 * (p & 3))) & 3];
    for (auto& run : runs)
      for (uint64_t p = run.first; p < std::min<uint64_t>(run.second, length); p++) ascii[p] = 'N';
    if (pbw_genome_add_contig(g, name.c_str(), ascii.data(), length) != PBW_OK) {
      r.err = "contig " + name + " rejected";
      break;
    }
    total += length;
  }
  uint64_t total_length = 0;
  if (r.err.empty()) r.u64(total_length);
  if (r.err.empty() && !r.at_end()) r.err = "trailing bytes after the genome";
  if (r.err.empty() && total_length != total) r.err = "total_length differs from the sum of the contig lengths";
  if (!r.err.empty()) {
    set_global_error(std::string(path) + ": " + r.err);
    pbw_genome_destroy(g);
    return PBW_ERR_IO;
  }
  *out = g;
  return PBW_OK;
}

// GenomeIndex::save (genome.rs:88-94) of `g`, packed as Contig::from_ascii
// (genome.rs:17-45) packs: non-ACGT bases as 0 bits plus their runs.
int pbw_genome_save_stidx(const pbw_genome* g, const char* species, const char* assembly, const char* path) {
  if (!g || !path) return PBW_ERR_ARG;
  BinWriter w(path);
  w.str(species ? species : "");
  w.str(assembly ? assembly : "");
  const int n = pbw_genome_n_contigs(g);
  w.u64((uint64_t)n);
  uint64_t total = 0;
  std::vector<uint8_t> seq, packed;
  for (int c = 0; c < n; c++) {
    const uint64_t L = pbw_genome_contig_length(g, c);
    seq.resize(L);
    pbw_genome_get_sequence(g, c, 0, L, seq.data());
    packed.assign((L + 3) / 4, 0);
    std::vector<std::pair<uint32_t, uint32_t>> runs;
    uint64_t run_start = UINT64_MAX;
    for (uint64_t p = 0; p < L; p++) {
      uint8_t bits = 0;
      bool acgt = true;
      switch (seq[p]) {
        case 'A': bits = 0; break;
        case 'C': bits = 1; break;
        case 'G': bits = 2; break;
        case 'T': bits = 3; break;
        default: acgt = false;
      }
      packed[p >> 2] |= (uint8_t)(bits << (2 * (p & 3)));
      if (acgt) {
        if (run_start != UINT64_MAX) runs.push_back({(uint32_t)run_start, (uint32_t)p});
        run_start = UINT64_MAX;
      } else if (run_start == UINT64_MAX) {
        run_start = p;
      }
    }
    if (run_start != UINT64_MAX) runs.push_back({(uint32_t)run_start, (uint32_t)L});
    w.str(pbw_genome_contig_name(g, c));
    w.u64(packed.size());
    w.bytes(packed.data(), packed.size());
    w.u64(runs.size());
    for (auto& r : runs) {
      w.u32(r.first);
      w.u32(r.second);
    }
    w.u64(L);
    total += L;
  }
  w.u64(total);
  if (!w.close()) {
    set_global_error(std::string(path) + ": " + w.err);
    return PBW_ERR_IO;
  }
  return PBW_OK;
}

}  // extern "C"

// ===========================================================================
// text: UTF-8 as Rust's String sees it
// ===========================================================================
namespace pbw {
namespace {

// Length of the valid UTF-8 sequence at s[i] (> 0), or minus the length of the
// maximal invalid subpart there (the unit String::from_utf8_lossy replaces).
int utf8_at(const uint8_t* s, size_t n, size_t i) {
  const uint8_t b0 = s[i];
  if (b0 < 0x80) return 1;
  int need;
  uint8_t lo = 0x80, hi = 0xBF;
  if (b0 >= 0xC2 && b0 <= 0xDF) {
    need = 1;
  } else if (b0 >= 0xE0 && b0 <= 0xEF) {
    need = 2;
    if (b0 == 0xE0) lo = 0xA0;
    if (b0 == 0xED) hi = 0x9F;
  } else if (b0 >= 0xF0 && b0 <= 0xF4) {
    need = 3;
    if (b0 == 0xF0) lo = 0x90;
    if (b0 == 0xF4) hi = 0x8F;
  } else {
    return -1;
  }
  for (int k = 1; k <= need; k++) {
    if (i + k >= n) return -k;
    const uint8_t b = s[i + k];
    const uint8_t l = k == 1 ? lo : 0x80, h = k == 1 ? hi : 0xBF;
    if (b < l || b > h) return -k;
  }
  return need + 1;
}

uint32_t utf8_cp(const uint8_t* s, int len) {
  if (len == 1) return s[0];
  uint32_t cp = s[0] & (0x7F >> len);
  for (int k = 1; k < len; k++) cp = (cp << 6) | (s[k] & 0x3F);
  return cp;
}

// char::is_whitespace (the Unicode White_Space property)
bool rust_is_ws(uint32_t c) {
  return (c >= 0x09 && c <= 0x0D) || c == 0x20 || c == 0x85 || c == 0xA0 || c == 0x1680 ||
         (c >= 0x2000 && c <= 0x200A) || c == 0x2028 || c == 0x2029 || c == 0x202F || c == 0x205F || c == 0x3000;
}

bool valid_utf8(const std::string& s) {
  const uint8_t* p = (const uint8_t*)s.data();
  for (size_t i = 0; i < s.size();) {
    const int l = utf8_at(p, s.size(), i);
    if (l < 0) return false;
    i += l;
  }
  return true;
}

// str::trim of valid UTF-8: [a, b) without leading / trailing whitespace
void trim_bounds(const std::string& s, size_t& a, size_t& b) {
  const uint8_t* p = (const uint8_t*)s.data();
  a = 0;
  b = s.size();
  while (a < b) {
    const int l = utf8_at(p, s.size(), a);
    if (!rust_is_ws(utf8_cp(p + a, l))) break;
    a += l;
  }
  while (b > a) {  // step back over one character
    size_t c = b - 1;
    while (c > a && (p[c] & 0xC0) == 0x80) c--;
    const int l = utf8_at(p, s.size(), c);
    if (!rust_is_ws(utf8_cp(p + c, l))) break;
    b = c;
  }
}
std::string rust_trim(const std::string& s) {
  size_t a, b;
  trim_bounds(s, a, b);
  return s.substr(a, b - a);
}
// s[from..].split_whitespace().next().unwrap_or("")
std::string first_token(const std::string& s, size_t from) {
  const uint8_t* p = (const uint8_t*)s.data();
  size_t a = from;
  while (a < s.size()) {
    const int l = utf8_at(p, s.size(), a);
    if (!rust_is_ws(utf8_cp(p + a, l))) break;
    a += l;
  }
  size_t b = a;
  while (b < s.size()) {
    const int l = utf8_at(p, s.size(), b);
    if (rust_is_ws(utf8_cp(p + b, l))) break;
    b += l;
  }
  return s.substr(a, b - a);
}

// String::from_utf8_lossy appended to out
void append_lossy(std::string& out, const uint8_t* s, size_t n) {
  for (size_t i = 0; i < n;) {
    const int l = utf8_at(s, n, i);
    if (l > 0) {
      out.append((const char*)s + i, l);
      i += l;
    } else {
      out.append("\xEF\xBF\xBD");
      i += -l;
    }
  }
}

// ===========================================================================
// line source: BufRead::read_line over a plain or gzip (MultiGzDecoder) file
// ===========================================================================
struct LineSource {
  FILE* f = nullptr;
  gzFile gz = nullptr;
  std::vector<char> buf;
  size_t pos = 0, end = 0;
  bool eof = false;
  uint64_t bytes = 0;  // file bytes consumed (compressed for gzip)
  std::string err;
  bool open(const char* path) {
    const size_t n = strlen(path);
    const bool is_gz = n > 3 && strcmp(path + n - 3, ".gz") == 0;  // extension "gz" (fastx.rs:128-133)
    if (is_gz) {
      gz = gzopen(path, "rb");
      if (gz) gzbuffer(gz, 1 << 20);
    } else {
      f = fopen(path, "rb");
    }
    if (!f && !gz) {
      err = std::string("cannot open ") + path;
      return false;
    }
    buf.resize(1 << 20);
    return true;
  }
  ~LineSource() {
    if (f) fclose(f);
    if (gz) gzclose(gz);
  }
  bool fill() {
    if (eof) return false;
    int got;
    if (gz) {
      got = gzread(gz, buf.data(), (unsigned)buf.size());
      if (got < 0) {
        int e;
        err = std::string("gzip: ") + gzerror(gz, &e);
        eof = true;
        return false;
      }
    } else {
      got = (int)fread(buf.data(), 1, buf.size(), f);
      bytes += (uint64_t)got;
    }
    pos = 0;
    end = (size_t)got;
    if (got == 0) eof = true;
    return got > 0;
  }
  // 1: a line (with its '\n' when there is one); 0: end of file; -1: error
  int read_line(std::string& line) {
    line.clear();
    bool any = false;
    for (;;) {
      if (pos == end && !fill()) break;
      const char* s = buf.data() + pos;
      const char* nl = (const char*)memchr(s, '\n', end - pos);
      const size_t take = nl ? (size_t)(nl - s) + 1 : end - pos;
      line.append(s, take);
      pos += take;
      any = true;
      if (nl) break;
    }
    if (!err.empty()) return -1;
    if (!any) return 0;
    if (!valid_utf8(line)) {  // read_line appends to a String (io::ErrorKind::InvalidData)
      err = "stream did not contain valid UTF-8";
      return -1;
    }
    return 1;
  }
};

}  // namespace
}  // namespace pbw

// ===========================================================================
// reads: FASTQ / FASTA (fastx.rs:51-305)
// ===========================================================================
struct pbw_fastx {
  pbw::LineSource src;
  int format = 0;  // 0 FASTQ, 1 FASTA
  uint8_t qoff = 33;
  bool has_adapter = false;
  std::string adapter;
  std::string pending;  // FASTA: the header that ended the previous record
  bool has_pending = false;
  bool done = false;
  // the current batch (reader-owned)
  std::vector<uint8_t> seq, qual;
  std::vector<uint64_t> so, qo;
  std::string ids;
  std::string line;
};

namespace pbw {
namespace {

// fastx.rs:77-81
void strip_mate_suffix(std::string& id) {
  if (id.size() >= 2 && id[id.size() - 2] == '/' && (id.back() == '1' || id.back() == '2')) id.resize(id.size() - 2);
}
inline uint8_t ascii_lower(uint8_t c) { return (c >= 'A' && c <= 'Z') ? (uint8_t)(c + 32) : c; }

// One record into the reader's batch; false at the end of the stream (or on
// an error, left in src.err).  fastx.rs:255-305 (FASTQ), :194-236 (FASTA).
bool next_record(pbw_fastx* f) {
  std::string& line = f->line;
  std::string id;
  std::vector<uint8_t> seq, qual;
  if (f->format == 0) {
    const int r = f->src.read_line(line);
    if (r <= 0) return false;
    size_t a = 0;
    while (a < line.size() && line[a] == '@') a++;  // trim_start_matches('@')
    id = first_token(line, a);
    for (;;) {
      const int q = f->src.read_line(line);
      if (q < 0) return false;
      if (q == 0) break;
      const std::string t = rust_trim(line);
      if (!t.empty() && t[0] == '+') break;
      seq.insert(seq.end(), t.begin(), t.end());
    }
    for (;;) {  // at least one quality line, even for an empty sequence
      const int q = f->src.read_line(line);
      if (q < 0) return false;
      if (q == 0) break;
      const std::string t = rust_trim(line);
      qual.insert(qual.end(), t.begin(), t.end());
      if (qual.size() >= seq.size()) break;
    }
  } else {
    std::string header;
    if (f->has_pending) {
      header = f->pending;
      f->has_pending = false;
    } else {
      for (;;) {
        const int r = f->src.read_line(line);
        if (r <= 0) return false;
        if (!line.empty() && line[0] == '>') break;
      }
      header = line;
    }
    size_t a = 0;
    while (a < header.size() && header[a] == '>') a++;
    id = first_token(header, a);
    for (;;) {
      const int q = f->src.read_line(line);
      if (q < 0) return false;
      if (q == 0) break;
      const std::string t = rust_trim(line);
      if (!t.empty() && t[0] == '>') {
        f->pending = line;
        f->has_pending = true;
        break;
      }
      seq.insert(seq.end(), t.begin(), t.end());
    }
    qual.assign(seq.size(), 40);  // fastx.rs:157 FASTA_DEFAULT_QUAL (already a score)
  }
  for (auto& b : seq)  // DnaSeq::from_ascii (types.rs:12-15)
    if (b >= 'a' && b <= 'z') b = (uint8_t)(b - 32);
  if (f->format == 0)
    for (auto& q : qual) q = q >= f->qoff ? (uint8_t)(q - f->qoff) : 0;  // from_phred_ascii (types.rs:57-60)
  strip_mate_suffix(id);
  if (f->has_adapter) {  // fastx.rs:97-117 strip_adapter
    const size_t L = seq.size(), max_ov = std::min(L, f->adapter.size());
    if (max_ov >= 6) {
      for (size_t ov = max_ov; ov >= 6; ov--) {
        size_t mm = 0;
        for (size_t i = 0; i < ov; i++) mm += ascii_lower(seq[L - ov + i]) != ascii_lower((uint8_t)f->adapter[i]);
        if (mm <= ov / 10) {
          seq.resize(L - ov);
          if (qual.size() > L - ov) qual.resize(L - ov);
          break;
        }
      }
    }
  }
  f->seq.insert(f->seq.end(), seq.begin(), seq.end());
  f->qual.insert(f->qual.end(), qual.begin(), qual.end());
  f->so.push_back(f->seq.size());
  f->qo.push_back(f->qual.size());
  f->ids.append(id);
  f->ids.push_back('\0');
  return true;
}

}  // namespace
}  // namespace pbw

extern "C" {

int pbw_fastx_open(const char* path, int format, uint32_t quality_offset, const char* adapter, pbw_fastx** out) {
  if (!path || !out || (format != 0 && format != 1) || quality_offset > 255) return PBW_ERR_ARG;
  *out = nullptr;
  pbw_fastx* f = new pbw_fastx();
  if (!f->src.open(path)) {
    set_global_error(f->src.err);
    delete f;
    return PBW_ERR_IO;
  }
  f->format = format;
  f->qoff = (uint8_t)quality_offset;
  if (adapter) {
    f->has_adapter = true;
    f->adapter = adapter;
  }
  *out = f;
  return PBW_OK;
}

int pbw_fastx_next(pbw_fastx* f, uint64_t max_reads, pbw_read_batch* reads, const char** ids) {
  if (!f || !reads) return PBW_ERR_ARG;
  f->seq.clear();
  f->qual.clear();
  f->so.assign(1, 0);
  f->qo.assign(1, 0);
  f->ids.clear();
  uint64_t n = 0;
  while (!f->done && n < max_reads) {
    if (!pbw::next_record(f)) {
      f->done = true;
      break;
    }
    n++;
  }
  if (!f->src.err.empty()) {
    set_global_error(f->src.err);
    return PBW_ERR_IO;
  }
  reads->n = n;
  reads->seq = f->seq.data();
  reads->seq_off = f->so.data();
  reads->qual = f->qual.data();
  reads->qual_off = f->qo.data();
  if (ids) *ids = f->ids.c_str();
  return PBW_OK;
}

uint64_t pbw_fastx_bytes_read(const pbw_fastx* f) { return f ? f->src.bytes : 0; }

void pbw_fastx_close(pbw_fastx* f) { delete f; }

// GenomeIndexBuilder::add_fasta (genome.rs:116-134) with its FastaReader
// (:150-213): plain or gzip by the ".gz" extension; the name is the header's
// first whitespace-separated token.
int pbw_genome_add_fasta(pbw_genome* g, const char* path) {
  if (!g || !path) return PBW_ERR_ARG;
  pbw::LineSource src;
  if (!src.open(path)) {
    set_global_error(src.err);
    return PBW_ERR_IO;
  }
  std::string line, pending;
  bool has_pending = false;
  for (;;) {
    std::string header;
    if (has_pending) {
      header = pending;
      has_pending = false;
    } else {
      int r;
      while ((r = src.read_line(line)) > 0)
        if (!line.empty() && line[0] == '>') break;
      if (r < 0) break;
      if (r == 0) break;
      header = line;
    }
    size_t a = 0;
    while (a < header.size() && header[a] == '>') a++;
    const std::string name = pbw::first_token(header, a);
    std::vector<uint8_t> seq;
    int r;
    while ((r = src.read_line(line)) > 0) {
      const std::string t = pbw::rust_trim(line);
      if (!t.empty() && t[0] == '>') {
        pending = line;
        has_pending = true;
        break;
      }
      seq.insert(seq.end(), t.begin(), t.end());
    }
    if (r < 0) break;
    const int rc = pbw_genome_add_contig(g, name.c_str(), seq.data(), seq.size());
    if (rc != PBW_OK) return rc;
  }
  if (!src.err.empty()) {
    set_global_error(std::string(path) + ": " + src.err);
    return PBW_ERR_IO;
  }
  return PBW_OK;
}

}  // extern "C"

// ===========================================================================
// output: SAM text (sam.rs:37-143), BAM (bam.rs:19-57) and BAI
// ===========================================================================
#include <atomic>
#include <condition_variable>
#include <deque>
#include <functional>
#include <mutex>
#include <thread>
#include <chrono>
#include <map>
#include <memory>

namespace pbw {
namespace {

struct OutGenome {
  std::vector<std::string> names;
  std::vector<uint64_t> lengths;
  std::vector<uint32_t> sort_rank;  // sam.rs:111-123: a name maps to its last contig index
};
OutGenome out_genome(const pbw_genome* g) {
  OutGenome G;
  const int n = pbw_genome_n_contigs(g);
  for (int c = 0; c < n; c++) {
    G.names.push_back(pbw_genome_contig_name(g, c));
    G.lengths.push_back(pbw_genome_contig_length(g, c));
  }
  G.sort_rank.resize(n);
  for (int c = 0; c < n; c++) {
    uint32_t last = (uint32_t)c;
    for (int d = c + 1; d < n; d++)
      if (G.names[d] == G.names[c]) last = (uint32_t)d;
    G.sort_rank[c] = last;
  }
  return G;
}

// ReadGroup::parse (sam.rs:17-32)
bool parse_read_group(const char* spec, std::string& id, std::string& line, std::string& err) {
  std::string s(spec), hdr = "@RG";
  bool has_id = false;
  size_t st = 0;
  for (;;) {
    const size_t e = s.find(',', st);
    const std::string pair = s.substr(st, e == std::string::npos ? std::string::npos : e - st);
    const size_t colon = pair.find(':');
    if (colon == std::string::npos) {
      err = "--readgroup: expected TAG:VALUE, got '" + pair + "'";
      return false;
    }
    const std::string tag = pair.substr(0, colon), val = pair.substr(colon + 1);
    if (tag == "ID") {
      id = val;
      has_id = true;
    }
    hdr += "\t" + tag + ":" + val;
    if (e == std::string::npos) break;
    st = e + 1;
  }
  if (!has_id) {
    err = "--readgroup must set at least an ID tag";
    return false;
  }
  line = hdr;
  return true;
}

// format_sam_records_sorted's header (sam.rs:83-107)
std::string sam_header(const OutGenome& G, const char* cmd, const std::string* rg_line, bool sorted) {
  std::string out = std::string("@HD\tVN:1.6\tSO:") + (sorted ? "coordinate" : "unsorted") + "\n";
  for (size_t i = 0; i < G.names.size(); i++)
    out += "@SQ\tSN:" + G.names[i] + "\tLN:" + std::to_string(G.lengths[i]) + "\n";
  out += std::string("@PG\tID:probebwa\tPN:probebwa\tVN:0.1.0\tCL:") + (cmd ? cmd : "") + "\n";
  if (rg_line) out += *rg_line + "\n";
  return out;
}

// The output side of one read: SEQ / QUAL as the record carries them
// (reverse-complemented / reversed when seq_rc; single.rs:101-103, 144-147).
struct ReadView {
  const uint8_t* seq;
  uint64_t sl;
  const uint8_t* qual;
  uint64_t ql;
  bool rc;
  uint8_t base(uint64_t t) const {
    if (!rc) return seq[t];
    const uint8_t b = seq[sl - 1 - t];
    return b == 'A' ? 'T' : b == 'T' ? 'A' : b == 'C' ? 'G' : b == 'G' ? 'C' : 'N';
  }
  uint8_t q(uint64_t t) const { return rc ? qual[ql - 1 - t] : qual[t]; }
};
ReadView read_view(const pbw_read_batch* rd, uint64_t i, const pbw_record& r) {
  ReadView v;
  v.seq = rd->seq + rd->seq_off[i];
  v.sl = rd->seq_off[i + 1] - rd->seq_off[i];
  v.qual = rd->qual + rd->qual_off[i];
  v.ql = rd->qual_off[i + 1] - rd->qual_off[i];
  v.rc = r.seq_rc != 0;
  return v;
}

// format_sam (sam.rs:37-58): SEQ and QUAL through String::from_utf8_lossy
void sam_record(std::string& out, const OutGenome& G, const pbw_read_batch* rd, uint64_t i, const char* id,
                const pbw_record& r, const uint32_t* cigar, const std::string* rg_id) {
  static const char OPS[] = "MIDNSHP=X";
  out += id;
  out += '\t';
  out += std::to_string(r.flag);
  out += '\t';
  out += r.contig < 0 ? std::string("*") : G.names[r.contig];
  out += '\t';
  out += std::to_string(r.pos);
  out += '\t';
  out += std::to_string(r.mapq);
  out += '\t';
  if (r.cigar_n == 0) out += '*';
  for (uint32_t k = 0; k < r.cigar_n; k++) {
    const uint32_t w = cigar[r.cigar_off + k];
    out += std::to_string(w >> 4);
    out += OPS[w & 15];
  }
  out += '\t';
  out += r.mate_contig < 0 ? std::string("*") : G.names[r.mate_contig];
  out += '\t';
  out += std::to_string(r.mate_pos);
  out += '\t';
  out += std::to_string(r.tlen);
  out += '\t';
  const ReadView v = read_view(rd, i, r);
  std::string tmp;
  tmp.resize(std::max(v.sl, v.ql));
  for (uint64_t t = 0; t < v.sl; t++) tmp[t] = (char)v.base(t);
  append_lossy(out, (const uint8_t*)tmp.data(), v.sl);
  out += '\t';
  for (uint64_t t = 0; t < v.ql; t++) tmp[t] = (char)v.q(t);
  append_lossy(out, (const uint8_t*)tmp.data(), v.ql);
  if (rg_id) {
    out += "\tRG:Z:";
    out += *rg_id;
  }
  out += '\n';
}

// ---- BAM ----
// reg2bin of the SAM specification (section 5.3): [beg, end) 0-based
int reg2bin(int64_t beg, int64_t end) {
  --end;
  if (beg >> 14 == end >> 14) return (int)(((1 << 15) - 1) / 7 + (beg >> 14));
  if (beg >> 17 == end >> 17) return (int)(((1 << 12) - 1) / 7 + (beg >> 17));
  if (beg >> 20 == end >> 20) return (int)(((1 << 9) - 1) / 7 + (beg >> 20));
  if (beg >> 23 == end >> 23) return (int)(((1 << 6) - 1) / 7 + (beg >> 23));
  if (beg >> 26 == end >> 26) return (int)(((1 << 3) - 1) / 7 + (beg >> 26));
  return 0;
}

inline void put32(std::string& o, uint32_t v) { o.append((const char*)&v, 4); }  // little-endian host
inline void put16(std::string& o, uint16_t v) { o.append((const char*)&v, 2); }

// Placement of an encoded record, for the index.
struct BamPlace {
  int32_t ref;
  int64_t beg, end;  // 0-based [beg, end) on ref (ref >= 0)
  bool mapped;
};

// The BAM encoding of the record that noodles would parse from the SAM text
// (bam.rs:19-32): QNAME 1..254 bytes of '!'..'~' without '@', QUAL bytes
// '!'..'~' stored minus 33 (raw_qual: the scores themselves), QUAL as long as
// SEQ.  Returns false with `err` where noodles-sam 0.86 / noodles-bam 0.91
// reject the record.
bool bam_record(std::string& out, const OutGenome& G, const pbw_read_batch* rd, uint64_t i, const char* id,
                const pbw_record& r, const uint32_t* cigar, const std::string* rg_id, bool raw_qual, BamPlace& pl,
                std::string& err) {
  static uint8_t CODE[256];
  static std::once_flag once;
  std::call_once(once, [] {
    const char* alpha = "=ACMGRSVTWYHKDBN";
    for (int c = 0; c < 256; c++) CODE[c] = 15;
    for (int k = 0; k < 16; k++) CODE[(uint8_t)alpha[k]] = (uint8_t)k;
  });
  const size_t nl = strlen(id);
  if (nl == 0 || nl > 254) {
    err = "read name '" + std::string(id) + "' is not a valid QNAME (1 to 254 characters)";
    return false;
  }
  for (size_t k = 0; k < nl; k++) {
    const uint8_t c = (uint8_t)id[k];
    if (c < '!' || c > '~' || c == '@') {
      err = "read name '" + std::string(id) + "' is not a valid QNAME";
      return false;
    }
  }
  const ReadView v = read_view(rd, i, r);
  if (v.sl == 0 || v.ql == 0) {
    err = "read '" + std::string(id) + "' has an empty SEQ or QUAL field";
    return false;
  }
  if (v.ql != v.sl) {
    err = "read '" + std::string(id) + "': QUAL length differs from SEQ length";
    return false;
  }
  if (r.cigar_n > 65535) {
    err = "read '" + std::string(id) + "': more than 65535 CIGAR operations";
    return false;
  }
  int64_t span = 0;
  for (uint32_t k = 0; k < r.cigar_n; k++) {
    const uint32_t w = cigar[r.cigar_off + k], op = w & 15;
    if (op == 0 || op == 2 || op == 3 || op == 7 || op == 8) span += w >> 4;
  }
  const int32_t ref = r.contig < 0 ? -1 : r.contig;
  const int32_t pos = r.pos ? (int32_t)(r.pos - 1) : -1;
  pl.ref = ref;
  pl.beg = pos;
  pl.end = pos + std::max<int64_t>(span, 1);
  pl.mapped = !(r.flag & 4);
  const uint16_t bin = (uint16_t)(pos < 0 ? 4680 : reg2bin(pos, pl.end));
  const size_t at = out.size();
  put32(out, 0);  // block_size, set below
  put32(out, (uint32_t)ref);
  put32(out, (uint32_t)pos);
  out.push_back((char)(nl + 1));
  out.push_back((char)r.mapq);
  put16(out, bin);
  put16(out, (uint16_t)r.cigar_n);
  put16(out, r.flag);
  put32(out, (uint32_t)v.sl);
  put32(out, (uint32_t)(r.mate_contig < 0 ? -1 : r.mate_contig));
  put32(out, (uint32_t)(r.mate_pos ? (int32_t)(r.mate_pos - 1) : -1));
  put32(out, (uint32_t)(int32_t)r.tlen);
  out.append(id, nl + 1);
  for (uint32_t k = 0; k < r.cigar_n; k++) put32(out, cigar[r.cigar_off + k]);
  for (uint64_t t = 0; t < v.sl; t += 2) {
    const uint8_t hi = CODE[v.base(t)], lo = t + 1 < v.sl ? CODE[v.base(t + 1)] : 0;
    out.push_back((char)((hi << 4) | lo));
  }
  for (uint64_t t = 0; t < v.ql; t++) {
    const uint8_t q = v.q(t);
    if (raw_qual) {
      out.push_back((char)q);
    } else {
      if (q < '!' || q > '~') {
        err = "read '" + std::string(id) + "': QUAL byte " + std::to_string(q) +
              " outside '!'..'~' (the SAM text carries raw Phred values; noodles-sam rejects it)";
        return false;
      }
      out.push_back((char)(q - 33));
    }
  }
  if (rg_id) {
    out += "RGZ";
    out += *rg_id;
    out.push_back('\0');
  }
  const uint32_t bs = (uint32_t)(out.size() - at - 4);
  memcpy(&out[at], &bs, 4);
  return true;
}

std::string bam_header(const OutGenome& G, const std::string& text) {
  std::string o = "BAM\1";
  put32(o, (uint32_t)text.size());
  o += text;
  put32(o, (uint32_t)G.names.size());
  for (size_t c = 0; c < G.names.size(); c++) {
    put32(o, (uint32_t)G.names[c].size() + 1);
    o.append(G.names[c].c_str(), G.names[c].size() + 1);
    put32(o, (uint32_t)G.lengths[c]);
  }
  return o;
}

// ---- BGZF (SAM specification, section 4.1) ----
constexpr size_t BGZF_BLOCK = 0xFF00;  // uncompressed bytes per block
const uint8_t BGZF_EOF[28] = {0x1f, 0x8b, 8, 4, 0, 0, 0, 0, 0, 0xff, 6, 0, 0x42, 0x43, 2, 0,
                              0x1b, 0, 3, 0, 0, 0, 0, 0, 0, 0, 0, 0};

bool bgzf_block(const uint8_t* src, size_t n, int level, std::string& out) {
  z_stream zs{};
  if (deflateInit2(&zs, level, Z_DEFLATED, -15, 8, Z_DEFAULT_STRATEGY) != Z_OK) return false;
  std::vector<uint8_t> cd(deflateBound(&zs, n) + 16);
  zs.next_in = const_cast<uint8_t*>(src);
  zs.avail_in = (uInt)n;
  zs.next_out = cd.data();
  zs.avail_out = (uInt)cd.size();
  const int rc = deflate(&zs, Z_FINISH);
  const size_t clen = zs.total_out;
  deflateEnd(&zs);
  if (rc != Z_STREAM_END || clen + 26 > 65536) return false;
  const uint32_t crc = (uint32_t)crc32(0, src, (uInt)n);
  const uint16_t bsize = (uint16_t)(clen + 25);
  const uint8_t hdr[18] = {0x1f, 0x8b, 8, 4, 0, 0, 0, 0, 0, 0xff, 6, 0, 0x42, 0x43, 2, 0,
                           (uint8_t)bsize, (uint8_t)(bsize >> 8)};
  out.assign((const char*)hdr, 18);
  out.append((const char*)cd.data(), clen);
  put32(out, crc);
  put32(out, (uint32_t)n);
  return true;
}

void parallel_for(unsigned threads, size_t n, const std::function<void(size_t)>& fn) {
  threads = (unsigned)std::max<size_t>(1, std::min<size_t>(threads, n));
  if (threads <= 1) {
    for (size_t i = 0; i < n; i++) fn(i);
    return;
  }
  std::atomic<size_t> next{0};
  std::vector<std::thread> pool;
  for (unsigned t = 0; t < threads; t++)
    pool.emplace_back([&] {
      for (size_t i; (i = next.fetch_add(1)) < n;) fn(i);
    });
  for (auto& th : pool) th.join();
}

// A BGZF file written block by block; blocks are compressed in parallel.
struct BgzfWriter {
  FILE* f = nullptr;
  std::string pending;     // uncompressed bytes not yet in a block
  uint64_t coff = 0;       // compressed bytes written
  uint64_t data_end = 0;   // compressed offset of the end-of-file block
  uint64_t uoff = 0;       // uncompressed bytes flushed into blocks
  std::vector<uint64_t> block_coff;  // compressed offset of every block written
  int level = 6;
  unsigned threads = 1;
  std::string err;
  // compresses every full block of `pending` (all of it when `all`)
  void flush(bool all) {
    const size_t nb = all ? (pending.size() + BGZF_BLOCK - 1) / BGZF_BLOCK : pending.size() / BGZF_BLOCK;
    if (!nb || !err.empty()) return;
    std::vector<std::string> blk(nb);
    std::atomic<int> bad{0};
    parallel_for(threads, nb, [&](size_t b) {
      const size_t a = b * BGZF_BLOCK, n = std::min(BGZF_BLOCK, pending.size() - a);
      if (!bgzf_block((const uint8_t*)pending.data() + a, n, level, blk[b])) bad = 1;
    });
    if (bad) {
      err = "BGZF compression failed";
      return;
    }
    for (size_t b = 0; b < nb; b++) {
      block_coff.push_back(coff);
      if (fwrite(blk[b].data(), 1, blk[b].size(), f) != blk[b].size()) err = "write failed";
      coff += blk[b].size();
    }
    const size_t used = std::min(pending.size(), nb * BGZF_BLOCK);
    uoff += used;
    pending.erase(0, used);
  }
  void write(const std::string& s) {
    pending += s;
    if (pending.size() >= 64 * BGZF_BLOCK) flush(false);
  }
  bool finish() {
    flush(true);
    data_end = coff;
    if (err.empty() && fwrite(BGZF_EOF, 1, 28, f) != 28) err = "write failed";
    coff += 28;
    return err.empty();
  }
  // virtual offset of uncompressed stream offset u (u < uoff after the flush that covered it)
  uint64_t voff(uint64_t u) const {
    const uint64_t b = u / BGZF_BLOCK;
    const uint64_t c = b < block_coff.size() ? block_coff[b] : data_end;
    return (c << 16) | (b < block_coff.size() ? u % BGZF_BLOCK : 0);
  }
};

// BAI (SAM specification, section 5.2): bins with chunks of virtual offsets,
// the 16 kb linear index, the metadata pseudo-bin 37450 (reference span,
// mapped and unmapped counts) and the count of unplaced unmapped records.
struct BaiBuilder {
  struct Ref {
    std::map<uint32_t, std::vector<std::pair<uint64_t, uint64_t>>> bins;
    std::vector<uint64_t> lin;
    uint64_t beg = UINT64_MAX, end = 0, n_mapped = 0, n_unmapped = 0;
    bool any = false;
  };
  std::vector<Ref> refs;
  uint64_t n_no_coor = 0;
  explicit BaiBuilder(size_t n) : refs(n) {}
  // a record at [v0, v1) of the stream (records arrive in coordinate order)
  void add(const BamPlace& p, int bin, uint64_t v0, uint64_t v1) {
    if (p.ref < 0) {
      n_no_coor++;
      return;
    }
    Ref& R = refs[p.ref];
    R.any = true;
    R.beg = std::min(R.beg, v0);
    R.end = std::max(R.end, v1);
    (p.mapped ? R.n_mapped : R.n_unmapped)++;
    auto& ch = R.bins[(uint32_t)bin];
    if (!ch.empty() && ch.back().second == v0) ch.back().second = v1;  // contiguous: one chunk
    else ch.push_back({v0, v1});
    const uint64_t w0 = (uint64_t)p.beg >> 14, w1 = (uint64_t)(p.end - 1) >> 14;
    if (R.lin.size() <= w1) R.lin.resize(w1 + 1, 0);
    for (uint64_t w = w0; w <= w1; w++)
      if (!R.lin[w]) R.lin[w] = v0;
  }
  std::string bytes() {
    std::string o = "BAI\1";
    put32(o, (uint32_t)refs.size());
    for (Ref& R : refs) {
      put32(o, (uint32_t)(R.bins.size() + (R.any ? 1 : 0)));
      for (auto& b : R.bins) {
        put32(o, b.first);
        put32(o, (uint32_t)b.second.size());
        for (auto& c : b.second) {
          o.append((const char*)&c.first, 8);
          o.append((const char*)&c.second, 8);
        }
      }
      if (R.any) {
        put32(o, 37450);
        put32(o, 2);
        o.append((const char*)&R.beg, 8);
        o.append((const char*)&R.end, 8);
        o.append((const char*)&R.n_mapped, 8);
        o.append((const char*)&R.n_unmapped, 8);
      }
      for (size_t w = 1; w < R.lin.size(); w++)
        if (!R.lin[w]) R.lin[w] = R.lin[w - 1];
      put32(o, (uint32_t)R.lin.size());
      for (uint64_t x : R.lin) o.append((const char*)&x, 8);
    }
    o.append((const char*)&n_no_coor, 8);
    return o;
  }
};

}  // namespace
}  // namespace pbw

// ===========================================================================
// the map subcommand's file path (main.rs:130-195, lib.rs:65-143)
// ===========================================================================
namespace pbw {
namespace {

double ms_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// One mapped batch: reads (owned copies), their records and the records kept.
struct MappedBatch {
  std::vector<uint8_t> seq, qual;
  std::vector<uint64_t> so, qo;
  std::string ids;
  std::vector<const char*> idp;
  pbw_record_batch rec{};
  std::vector<uint32_t> keep;
  MappedBatch() = default;
  MappedBatch(const MappedBatch&) = delete;
  ~MappedBatch() {
    if (rec.rec) pbw_records_release(&rec);
  }
  pbw_read_batch view() const {
    pbw_read_batch b;
    b.n = so.empty() ? 0 : so.size() - 1;
    b.seq = seq.data();
    b.seq_off = so.data();
    b.qual = qual.data();
    b.qual_off = qo.data();
    return b;
  }
  void index_ids() {
    idp.clear();
    for (size_t p = 0; p < ids.size(); p += strlen(ids.c_str() + p) + 1) idp.push_back(ids.c_str() + p);
  }
};

template <typename T>
struct Queue {  // bounded, closed by pushing nullptr
  std::mutex mu;
  std::condition_variable cv;
  std::deque<T*> q;
  size_t cap = 2;
  void push(T* x) {
    std::unique_lock<std::mutex> l(mu);
    cv.wait(l, [&] { return q.size() < cap || x == nullptr; });
    q.push_back(x);
    cv.notify_all();
  }
  T* pop() {
    std::unique_lock<std::mutex> l(mu);
    cv.wait(l, [&] { return !q.empty(); });
    T* x = q.front();
    q.pop_front();
    cv.notify_all();
    return x;
  }
};

// The output stream: SAM text to a FILE, or BAM through BGZF (+ its index).
struct Output {
  const OutGenome* G = nullptr;
  int format = 0;
  bool raw_qual = false;
  unsigned threads = 1;
  const std::string* rg_id = nullptr;
  FILE* f = nullptr;
  bool own_file = false;
  BgzfWriter bgzf;
  uint64_t bytes = 0, records = 0;
  std::string err;
  // placement and [begin, end) stream offsets of every record (index builds)
  bool track = false;
  std::vector<BamPlace> place;
  std::vector<int> bins;
  std::vector<std::pair<uint64_t, uint64_t>> uoffs;

  void header(const std::string& text) {
    if (format == 0) {
      if (fwrite(text.data(), 1, text.size(), f) != text.size()) err = "write failed";
      bytes += text.size();
    } else {
      bgzf.write(bam_header(*G, text));
    }
  }
  // records `sel` of batch `b`, formatted in parallel chunks and written in order
  void records_of(const std::vector<std::pair<const MappedBatch*, uint32_t>>& sel) {
    const size_t CH = 2048, nch = (sel.size() + CH - 1) / CH;
    std::vector<std::string> txt(nch);
    std::vector<std::vector<BamPlace>> pl(nch);
    std::vector<std::vector<uint32_t>> ends(nch);
    std::vector<std::string> errs(nch);
    parallel_for(threads, nch, [&](size_t c) {
      for (size_t k = c * CH; k < std::min(sel.size(), (c + 1) * CH); k++) {
        const MappedBatch* b = sel[k].first;
        const uint32_t i = sel[k].second;
        const pbw_read_batch rb = b->view();
        if (format == 0) {
          sam_record(txt[c], *G, &rb, i, b->idp[i], b->rec.rec[i], b->rec.cigar, rg_id);
        } else {
          BamPlace p;
          if (!bam_record(txt[c], *G, &rb, i, b->idp[i], b->rec.rec[i], b->rec.cigar, rg_id, raw_qual, p,
                          errs[c]))
            return;
          if (track) {
            pl[c].push_back(p);
            ends[c].push_back((uint32_t)txt[c].size());
          }
        }
      }
    });
    for (size_t c = 0; c < nch; c++)
      if (!errs[c].empty()) {
        err = errs[c];
        return;
      }
    for (size_t c = 0; c < nch; c++) {
      if (format == 0) {
        if (fwrite(txt[c].data(), 1, txt[c].size(), f) != txt[c].size()) err = "write failed";
        bytes += txt[c].size();
      } else {
        const uint64_t u0 = bgzf.uoff + bgzf.pending.size();
        if (track) {
          uint32_t prev = 0;
          for (size_t k = 0; k < pl[c].size(); k++) {
            place.push_back(pl[c][k]);
            const BamPlace& p = pl[c][k];
            bins.push_back(p.beg < 0 ? 4680 : reg2bin(p.beg, p.end));
            uoffs.push_back({u0 + prev, u0 + ends[c][k]});
            prev = ends[c][k];
          }
        }
        bgzf.write(txt[c]);
      }
    }
    records += sel.size();
  }
  bool finish(const char* path) {
    if (format == 1) {
      if (!bgzf.finish() && err.empty()) err = bgzf.err;
      bytes = bgzf.coff;
    }
    if (own_file && f && fclose(f) != 0 && err.empty()) err = "write failed";
    f = nullptr;
    if (err.empty() && track) {
      BaiBuilder bai(G->names.size());
      for (size_t k = 0; k < place.size(); k++)
        bai.add(place[k], bins[k], bgzf.voff(uoffs[k].first), bgzf.voff(uoffs[k].second));
      const std::string idx = bai.bytes();
      const std::string p = std::string(path) + ".bai";
      FILE* fi = fopen(p.c_str(), "wb");
      if (!fi || fwrite(idx.data(), 1, idx.size(), fi) != idx.size()) err = "cannot write " + p;
      if (fi) fclose(fi);
    }
    return err.empty();
  }
};

// sort_for_coordinate_order (sam.rs:109-124): stable by (contig rank, position),
// unmapped records (position 0) last
void coordinate_sort(std::vector<std::pair<const MappedBatch*, uint32_t>>& sel, const OutGenome& G) {
  auto key = [&](const std::pair<const MappedBatch*, uint32_t>& x) {
    const pbw_record& r = x.first->rec.rec[x.second];
    if (r.pos == 0) return std::make_pair(UINT64_MAX, UINT64_MAX);
    return std::make_pair((uint64_t)G.sort_rank[r.contig], r.pos);
  };
  std::stable_sort(sel.begin(), sel.end(), [&](const auto& a, const auto& b) { return key(a) < key(b); });
}

bool keep_record(const char* id, const char* prefix) {
  return !prefix || strncmp(id, prefix, strlen(prefix)) == 0;
}

// every record of a read file, through the reader
bool read_all(const char* path, const pbw_io_options& o, MappedBatch& b, uint64_t& bytes, std::string& err) {
  pbw_fastx* fx = nullptr;
  if (pbw_fastx_open(path, o.input_format, o.quality_offset, o.adapter, &fx) != PBW_OK) {
    err = pbw_global_error();
    return false;
  }
  b.so.assign(1, 0);
  b.qo.assign(1, 0);
  for (;;) {
    pbw_read_batch rb;
    const char* ids = nullptr;
    if (pbw_fastx_next(fx, 1u << 20, &rb, &ids) != PBW_OK) {
      err = pbw_global_error();
      pbw_fastx_close(fx);
      return false;
    }
    if (rb.n == 0) break;
    const uint64_t s0 = b.seq.size(), q0 = b.qual.size();
    b.seq.insert(b.seq.end(), rb.seq, rb.seq + rb.seq_off[rb.n]);
    b.qual.insert(b.qual.end(), rb.qual, rb.qual + rb.qual_off[rb.n]);
    for (uint64_t i = 1; i <= rb.n; i++) {
      b.so.push_back(s0 + rb.seq_off[i]);
      b.qo.push_back(q0 + rb.qual_off[i]);
    }
    const char* p = ids;
    for (uint64_t i = 0; i < rb.n; i++) {
      const size_t l = strlen(p);
      b.ids.append(p, l + 1);
      p += l + 1;
    }
  }
  bytes = pbw_fastx_bytes_read(fx);
  pbw_fastx_close(fx);
  return true;
}

}  // namespace
}  // namespace pbw

extern "C" {

// format_sam_records (sam.rs:60-107) of one batch: records[i] belongs to reads[i];
// ids is n NUL-terminated strings back to back (NULL: r0, r1, ...).
char* pbw_format_sam(const pbw_genome* g, const pbw_read_batch* rd, const char* ids, const pbw_record_batch* recs,
                     const char* cmd, const char* rg, uint64_t* len) {
  if (!g || !rd || !recs || recs->n > rd->n) return nullptr;
  const OutGenome G = out_genome(g);
  std::string rg_id, rg_line, err;
  if (rg && !parse_read_group(rg, rg_id, rg_line, err)) {
    set_global_error(err);
    return nullptr;
  }
  std::string out = sam_header(G, cmd, rg ? &rg_line : nullptr, false);
  const char* idp = ids;
  std::string gen;
  for (uint64_t i = 0; i < recs->n; i++) {
    const char* id = idp;
    if (idp) {
      idp += strlen(idp) + 1;
    } else {
      gen = "r" + std::to_string(i);
      id = gen.c_str();
    }
    sam_record(out, G, rd, i, id, recs->rec[i], recs->cigar, rg ? &rg_id : nullptr);
  }
  char* t = (char*)malloc(out.size() + 1);
  memcpy(t, out.data(), out.size());
  t[out.size()] = 0;
  if (len) *len = out.size();
  return t;
}
void pbw_free_text(char* t) { free(t); }

void pbw_default_io_options(pbw_io_options* o) {
  if (!o) return;
  memset(o, 0, sizeof(*o));
  o->input_format = 0;
  o->quality_offset = 33;
  o->threads = 0;
  o->batch_reads = 262144;
}

// write_bam / write_bam_indexed (bam.rs:19-57) of one batch's records.
int pbw_write_bam(const pbw_genome* g, const pbw_read_batch* rd, const char* ids, const pbw_record_batch* recs,
                  const char* cmd, const char* rg, const char* path, int index, int raw_qual, uint32_t threads) {
  if (!g || !rd || !recs || !path || recs->n > rd->n) return PBW_ERR_ARG;
  const OutGenome G = out_genome(g);
  std::string rg_id, rg_line, err;
  if (rg && !parse_read_group(rg, rg_id, rg_line, err)) {
    set_global_error(err);
    return PBW_ERR_ARG;
  }
  MappedBatch b;  // a view of the caller's batch (records borrowed, not released)
  const uint64_t n = recs->n;
  b.so.assign(rd->seq_off, rd->seq_off + n + 1);
  b.qo.assign(rd->qual_off, rd->qual_off + n + 1);
  b.seq.assign(rd->seq, rd->seq + rd->seq_off[n]);
  b.qual.assign(rd->qual, rd->qual + rd->qual_off[n]);
  for (uint64_t i = 0; i < n; i++) {
    if (ids) {
      const size_t l = strlen(ids);
      b.ids.append(ids, l + 1);
      ids += l + 1;
    } else {
      b.ids += "r" + std::to_string(i);
      b.ids.push_back('\0');
    }
  }
  b.index_ids();
  b.rec = *recs;
  std::vector<std::pair<const MappedBatch*, uint32_t>> sel;
  for (uint64_t i = 0; i < n; i++) sel.push_back({&b, (uint32_t)i});
  if (index) coordinate_sort(sel, G);
  Output out;
  out.G = &G;
  out.format = 1;
  out.raw_qual = raw_qual != 0;
  out.threads = threads ? threads : std::max(1u, std::thread::hardware_concurrency());
  out.rg_id = rg ? &rg_id : nullptr;
  out.track = index != 0;
  out.bgzf.threads = out.threads;
  out.f = out.bgzf.f = fopen(path, "wb");
  out.own_file = true;
  int rc = PBW_OK;
  if (!out.f) {
    set_global_error(std::string("cannot create ") + path);
    rc = PBW_ERR_IO;
  } else {
    out.header(sam_header(G, cmd, rg ? &rg_line : nullptr, index != 0));
    out.records_of(sel);
    if (!out.finish(path)) {
      set_global_error(out.err);
      rc = out.err.find("write") != std::string::npos || out.err.find("cannot") != std::string::npos ? PBW_ERR_IO
                                                                                                      : PBW_ERR_ARG;
    }
  }
  b.rec = pbw_record_batch{};  // borrowed
  return rc;
}

int pbw_map_files(pbw_ctx* ctx, const char* reads1, const char* reads2, const char* out_path,
                  const pbw_io_options* opt, pbw_io_report* rep) {
  if (!ctx || !reads1) return PBW_ERR_ARG;
  pbw_io_options o;
  if (opt) o = *opt;
  else pbw_default_io_options(&o);
  if (o.output_format != 0 && o.output_format != 1) return PBW_ERR_ARG;
  if (o.index && o.output_format != 1) {
    set_global_error("--index requires --outputformat=bam");
    return PBW_ERR_ARG;
  }
  if (o.output_format == 1 && !out_path) {
    set_global_error("--outputformat=bam requires --output FILE (binary output shouldn't go to a terminal)");
    return PBW_ERR_ARG;
  }
  const auto t0 = std::chrono::steady_clock::now();
  pbw_io_report R{};
  const OutGenome G = out_genome(pbw_ctx_genome(ctx));
  std::string rg_id, rg_line, err;
  if (o.read_group && !parse_read_group(o.read_group, rg_id, rg_line, err)) {
    set_global_error(err);
    return PBW_ERR_ARG;
  }
  Output out;
  out.G = &G;
  out.format = o.output_format;
  out.raw_qual = o.raw_qual != 0;
  out.threads = o.threads ? o.threads : std::max(1u, std::thread::hardware_concurrency());
  out.rg_id = o.read_group ? &rg_id : nullptr;
  out.track = o.index != 0;
  out.bgzf.threads = out.threads;
  if (out_path) {
    out.f = fopen(out_path, "wb");
    out.own_file = true;
    if (!out.f) {
      set_global_error(std::string("cannot create ") + out_path);
      return PBW_ERR_IO;
    }
  } else {
    out.f = stdout;
  }
  out.bgzf.f = out.f;
  const std::string header = sam_header(G, o.command_line, o.read_group ? &rg_line : nullptr, o.index != 0);
  std::vector<std::unique_ptr<MappedBatch>> kept;  // batches whose records are still to be written
  std::vector<std::pair<const MappedBatch*, uint32_t>> all;
  int rc = PBW_OK;
  auto fail = [&](int code, const std::string& m) {
    if (rc == PBW_OK) {
      rc = code;
      err = m;
    }
  };
  if (reads2) {
    // map_paired (lib.rs:145-185): both files read whole, counts checked, then the pairs
    MappedBatch m1, m2;
    std::string e1, e2;
    uint64_t b1 = 0, b2 = 0;
    const auto tp = std::chrono::steady_clock::now();
    std::thread th([&] {
      if (!read_all(reads2, o, m2, b2, e2)) m2.so.clear();
    });
    const bool ok1 = read_all(reads1, o, m1, b1, e1);
    th.join();
    R.ms_parse = ms_since(tp);
    R.bytes_in = b1 + b2;
    if (!ok1 || !e2.empty()) {
      set_global_error(!ok1 ? e1 : e2);
      return PBW_ERR_IO;
    }
    const pbw_read_batch r1 = m1.view(), r2 = m2.view();
    R.reads = r1.n + r2.n;
    if (r1.n != r2.n) {
      set_global_error("paired-end input files have different read counts (" + std::to_string(r1.n) + " vs " +
                       std::to_string(r2.n) + ")");
      return PBW_ERR_MISMATCH;
    }
    auto pb = std::make_unique<MappedBatch>();  // r1, r2 interleaved: record j belongs to read j
    pb->so.assign(1, 0);
    pb->qo.assign(1, 0);
    m1.index_ids();
    m2.index_ids();
    for (uint64_t i = 0; i < r1.n; i++)
      for (const MappedBatch* m : {&m1, &m2}) {
        pb->seq.insert(pb->seq.end(), m->seq.begin() + m->so[i], m->seq.begin() + m->so[i + 1]);
        pb->qual.insert(pb->qual.end(), m->qual.begin() + m->qo[i], m->qual.begin() + m->qo[i + 1]);
        pb->so.push_back(pb->seq.size());
        pb->qo.push_back(pb->qual.size());
        pb->ids.append(m->idp[i], strlen(m->idp[i]) + 1);
      }
    pb->index_ids();
    pbw_options co;
    pbw_ctx_get_options(ctx, &co);
    pbw_insert_model model{co.insert_size_mean, co.insert_size_sd, 0};
    const auto tm = std::chrono::steady_clock::now();
    const int mrc = pbw_map_pe(ctx, &r1, &r2, &model, &pb->rec);
    R.ms_map = ms_since(tm);
    if (mrc != PBW_OK) {
      set_global_error(pbw_last_error(ctx));
      if (out.own_file) fclose(out.f);
      return mrc;
    }
    for (uint32_t j = 0; j < pb->rec.n; j++)
      if (keep_record(pb->idp[j], o.label_prefix)) all.push_back({pb.get(), j});
    kept.push_back(std::move(pb));
    const auto tw = std::chrono::steady_clock::now();
    if (o.index) coordinate_sort(all, G);
    out.header(header);
    out.records_of(all);
    R.ms_write = ms_since(tw);
  } else {
    // single-end: reads parsed, mapped and written as overlapping batches
    pbw_fastx* fx = nullptr;
    if (pbw_fastx_open(reads1, o.input_format, o.quality_offset, o.adapter, &fx) != PBW_OK) {
      if (out.own_file) fclose(out.f);
      return PBW_ERR_IO;
    }
    const uint64_t bs = o.batch_reads ? o.batch_reads : 262144;
    Queue<MappedBatch> parsed, mapped;
    std::string perr;
    std::thread parser([&] {
      const auto tp = std::chrono::steady_clock::now();
      for (;;) {
        pbw_read_batch rb;
        const char* ids = nullptr;
        if (pbw_fastx_next(fx, bs, &rb, &ids) != PBW_OK) {
          perr = pbw_global_error();
          break;
        }
        if (rb.n == 0) break;
        auto* b = new MappedBatch();
        b->seq.assign(rb.seq, rb.seq + rb.seq_off[rb.n]);
        b->qual.assign(rb.qual, rb.qual + rb.qual_off[rb.n]);
        b->so.assign(rb.seq_off, rb.seq_off + rb.n + 1);
        b->qo.assign(rb.qual_off, rb.qual_off + rb.n + 1);
        const char* p = ids;
        for (uint64_t i = 0; i < rb.n; i++) p += strlen(p) + 1;
        b->ids.assign(ids, p - ids);
        b->index_ids();
        parsed.push(b);
      }
      R.ms_parse = ms_since(tp);
      parsed.push(nullptr);
    });
    std::thread writer;
    if (!o.index) {
      writer = std::thread([&] {
        out.header(header);
        for (MappedBatch* b; (b = mapped.pop());) {
          const auto tw = std::chrono::steady_clock::now();
          std::vector<std::pair<const MappedBatch*, uint32_t>> sel;
          for (uint32_t i : b->keep) sel.push_back({b, i});
          if (out.err.empty()) out.records_of(sel);
          R.ms_write += ms_since(tw);
          delete b;
        }
      });
    }
    for (MappedBatch* b; (b = parsed.pop());) {
      const pbw_read_batch rb = b->view();
      R.reads += rb.n;
      if (rc == PBW_OK) {
        const auto tm = std::chrono::steady_clock::now();
        const int mrc = pbw_map_se(ctx, &rb, &b->rec);
        R.ms_map += ms_since(tm);
        if (mrc != PBW_OK) fail(mrc, pbw_last_error(ctx));
      }
      if (rc != PBW_OK) {
        delete b;
        continue;  // drain the parser
      }
      for (uint32_t i = 0; i < b->rec.n; i++)
        if (keep_record(b->idp[i], o.label_prefix)) b->keep.push_back(i);
      if (o.index) {
        for (uint32_t i : b->keep) all.push_back({b, i});
        kept.emplace_back(b);
      } else {
        mapped.push(b);
      }
    }
    parser.join();
    R.bytes_in = pbw_fastx_bytes_read(fx);
    pbw_fastx_close(fx);
    if (!o.index) {
      mapped.push(nullptr);
      writer.join();
    }
    if (!perr.empty()) fail(PBW_ERR_IO, perr);
    if (o.index && rc == PBW_OK) {
      const auto tw = std::chrono::steady_clock::now();
      coordinate_sort(all, G);
      out.header(header);
      out.records_of(all);
      R.ms_write = ms_since(tw);
    }
  }
  if (!out.err.empty()) fail(PBW_ERR_ARG, out.err);
  const std::string oerr = out.err;
  if (!out.finish(out_path ? out_path : "") && rc == PBW_OK) fail(PBW_ERR_IO, out.err);
  if (rc != PBW_OK) {
    set_global_error(err);
    return rc;
  }
  R.records = out.records;
  R.bytes_out = out.bytes;
  R.ms_total = ms_since(t0);
  if (rep) *rep = R;
  return PBW_OK;
}

}  // extern "C"
