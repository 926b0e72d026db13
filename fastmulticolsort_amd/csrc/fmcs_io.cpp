// fmcs_io.cpp -- column file I/O (host code): the reference's text column
// files (src/generator.h:43-227), parsed by several threads, and the binary
// column file of include/fmcs_io.h.
#include "../../include/fmcs_io.h"

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cerrno>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

namespace {

typedef uint16_t u16;
typedef uint32_t u32;
typedef uint64_t u64;

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

// Bank of a text column (ReadFromFile, src/generator.h:77-97).
u32 text_bank(u32 w) { return w <= 16 ? 2u : (w <= 32 ? 4u : 8u); }

u64 bank_max(u32 bank) { return bank >= 8 ? ~0ull : ((1ull << (8 * bank)) - 1); }

u32 pick_threads(u32 want, u64 work_bytes) {
  u32 hw = std::thread::hardware_concurrency();
  if (hw == 0) hw = 1;
  u32 t = want ? want : hw;
  t = std::min<u32>(t, 64);
  // below ~1 MiB a thread costs more than it saves
  const u64 cap = std::max<u64>(1, work_bytes >> 20);
  return (u32)std::max<u64>(1, std::min<u64>(t, cap));
}

template <class F>
void parallel_for(u32 nt, F f) {
  if (nt <= 1) {
    f(0u);
    return;
  }
  std::vector<std::thread> th;
  th.reserve(nt);
  for (u32 t = 0; t < nt; ++t) th.emplace_back(f, t);
  for (auto& x : th) x.join();
}

// Read-only map of a whole file (an empty file maps to nothing).
struct Mapped {
  int fd = -1;
  const char* p = nullptr;
  size_t n = 0;
  ~Mapped() {
    if (p) munmap((void*)p, n);
    if (fd >= 0) close(fd);
  }
  int open(const char* path) {
    fd = ::open(path, O_RDONLY);
    if (fd < 0) return fail(FMCS_E_IO, std::string("cannot open ") + path + ": " + std::strerror(errno));
    struct stat st;
    if (fstat(fd, &st) != 0) return fail(FMCS_E_IO, std::string("cannot stat ") + path);
    n = (size_t)st.st_size;
    if (n) {
      void* q = mmap(nullptr, n, PROT_READ, MAP_PRIVATE, fd, 0);
      if (q == MAP_FAILED) return fail(FMCS_E_IO, std::string("cannot map ") + path);
      p = (const char*)q;
      madvise(q, n, MADV_SEQUENTIAL);
    }
    return FMCS_OK;
  }
};

// White space as the C locale's isspace (what `istream >> v` skips).
inline bool is_ws(char c) { return c == ' ' || c == '\n' || c == '\t' || c == '\r' || c == '\v' || c == '\f'; }

// Chunk t of nt of [0, n): tokens that start inside it belong to it.  A
// chunk whose start falls inside a token skips that token's tail (it belongs
// to the previous chunk).
void chunk_bounds(const char* p, size_t n, u32 t, u32 nt, size_t* s, size_t* e) {
  size_t a = n * t / nt, b = n * (t + 1) / nt;
  if (a > 0 && a < n && !is_ws(p[a - 1]))
    while (a < n && !is_ws(p[a])) ++a;
  *s = a;
  *e = b;
}

u64 count_tokens(const char* p, size_t n, size_t s, size_t e) {
  u64 c = 0;
  size_t i = s;
  while (i < e) {
    while (i < e && is_ws(p[i])) ++i;
    if (i >= e) break;
    ++c;
    while (i < n && !is_ws(p[i])) ++i;  // a token that starts here may end past e
  }
  return c;
}

// Parses one unsigned decimal token at p[i..); returns false if it is not
// one or does not fit 64 bits.
inline bool parse_u64(const char* p, size_t n, size_t* i, u64* v) {
  u64 x = 0;
  size_t j = *i;
  const size_t j0 = j;
  for (; j < n && !is_ws(p[j]); ++j) {
    const unsigned d = (unsigned)(unsigned char)p[j] - '0';
    if (d > 9) return false;
    if (x > (~0ull - d) / 10) return false;
    x = x * 10 + d;
  }
  *i = j;
  *v = x;
  return j > j0;
}

struct ParseErr {
  size_t at = ~(size_t)0;
  int code = FMCS_OK;
  std::string msg;
};

// Distinct values of a column: values are split by hash into partitions,
// one per thread, each counted in its own open-addressing table (grown at
// half load).
u64 mix64(u64 z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

template <class T>
u64 distinct_count(const T* v, u64 n, u32 nt) {
  std::vector<u64> part(nt, 0);
  parallel_for(nt, [&](u32 t) {
    std::vector<u64> tab(1024);
    std::vector<uint8_t> used(1024, 0);
    u64 mask = 1023, cnt = 0;
    for (u64 i = 0; i < n; ++i) {
      const u64 x = (u64)v[i];
      const u64 h = mix64(x);
      if (nt > 1 && (u32)((h >> 40) % nt) != t) continue;
      u64 k = h & mask;
      while (used[k] && tab[k] != x) k = (k + 1) & mask;
      if (used[k]) continue;
      used[k] = 1;
      tab[k] = x;
      if (++cnt * 2 > mask + 1) {  // grow and reinsert
        std::vector<u64> nt2((mask + 1) * 2);
        std::vector<uint8_t> nu((mask + 1) * 2, 0);
        const u64 m2 = (mask + 1) * 2 - 1;
        for (u64 j = 0; j <= mask; ++j)
          if (used[j]) {
            u64 q = mix64(tab[j]) & m2;
            while (nu[q]) q = (q + 1) & m2;
            nu[q] = 1;
            nt2[q] = tab[j];
          }
        tab.swap(nt2);
        used.swap(nu);
        mask = m2;
      }
    }
    part[t] = cnt;
  });
  u64 s = 0;
  for (u64 c : part) s += c;
  return s;
}

u64 column_distinct(const void* col, u32 bank, u64 n, u32 nt) {
  switch (bank) {
    case 1: return distinct_count((const uint8_t*)col, n, nt);
    case 2: return distinct_count((const u16*)col, n, nt);
    case 4: return distinct_count((const u32*)col, n, nt);
    default: return distinct_count((const u64*)col, n, nt);
  }
}

inline u64 load_bank(const void* col, u32 bank, u64 i) {
  switch (bank) {
    case 1: return ((const uint8_t*)col)[i];
    case 2: return ((const u16*)col)[i];
    case 4: return ((const u32*)col)[i];
    default: return ((const u64*)col)[i];
  }
}

inline void store_bank(void* col, u32 bank, u64 i, u64 v) {
  switch (bank) {
    case 2: ((u16*)col)[i] = (u16)v; break;
    case 4: ((u32*)col)[i] = (u32)v; break;
    default: ((u64*)col)[i] = v; break;
  }
}

// ---- binary column file --------------------------------------------------

const char MAGIC[8] = {'F', 'M', 'C', 'S', 'C', 'O', 'L', 'S'};
constexpr u32 VERSION = 1;
constexpr u64 ALIGN = 4096;
constexpr u64 GOLDEN = 0x9E3779B97F4A7C15ull;

#pragma pack(push, 1)
struct FileHead {
  char magic[8];
  u32 version;
  u32 ncols;
  u64 nrows;
  u64 header_bytes;
};
struct FileCol {
  u32 width;
  u32 bank_bytes;
  u64 cardinality;
  u64 offset;
  u64 bytes;
  u64 digest;
  u32 flags;
  u32 reserved;
  char name[16];
};
#pragma pack(pop)
static_assert(sizeof(FileHead) == 32, "header layout");
static_assert(sizeof(FileCol) == 64, "column record layout");

u64 round_up(u64 x, u64 a) { return (x + a - 1) / a * a; }

// Digest of bytes [0, nbytes) of a column, as 8-byte words starting at word
// index w0 (the last word zero-padded).
u64 digest_words(const uint8_t* p, u64 nbytes, u64 w0) {
  u64 s = 0;
  const u64 nw = nbytes / 8;
  for (u64 i = 0; i < nw; ++i) {
    u64 w;
    std::memcpy(&w, p + 8 * i, 8);
    s += mix64(w + (w0 + i) * GOLDEN);
  }
  if (nbytes % 8) {
    u64 w = 0;
    std::memcpy(&w, p + 8 * nw, nbytes % 8);
    s += mix64(w + (w0 + nw) * GOLDEN);
  }
  return s;
}

u64 digest_parallel(const uint8_t* p, u64 nbytes) {
  const u32 nt = pick_threads(0, nbytes);
  std::vector<u64> part(nt, 0);
  const u64 nw = (nbytes + 7) / 8;
  parallel_for(nt, [&](u32 t) {
    const u64 a = nw * t / nt, b = nw * (t + 1) / nt;
    const u64 ba = a * 8, bb = std::min<u64>(b * 8, nbytes);
    if (bb > ba) part[t] = digest_words(p + ba, bb - ba, a);
  });
  u64 s = 0;
  for (u64 x : part) s += x;
  return s;
}

int pread_all(int fd, void* buf, u64 n, u64 off) {
  uint8_t* b = (uint8_t*)buf;
  while (n) {
    const ssize_t r = ::pread(fd, b, (size_t)std::min<u64>(n, 1ull << 30), (off_t)off);
    if (r < 0 && errno == EINTR) continue;
    if (r <= 0) return fail(FMCS_E_IO, r == 0 ? "unexpected end of file" : std::strerror(errno));
    b += r;
    n -= (u64)r;
    off += (u64)r;
  }
  return FMCS_OK;
}

int pwrite_all(int fd, const void* buf, u64 n, u64 off) {
  const uint8_t* b = (const uint8_t*)buf;
  while (n) {
    const ssize_t r = ::pwrite(fd, b, (size_t)std::min<u64>(n, 1ull << 30), (off_t)off);
    if (r < 0 && errno == EINTR) continue;
    if (r <= 0) return fail(FMCS_E_IO, std::string("write failed: ") + std::strerror(errno));
    b += r;
    n -= (u64)r;
    off += (u64)r;
  }
  return FMCS_OK;
}

struct Fd {
  int fd = -1;
  ~Fd() {
    if (fd >= 0) close(fd);
  }
};

// Reads and checks the header of an open column file.
int read_header(int fd, const char* path, FileHead* h, std::vector<FileCol>* cols) {
  struct stat st;
  if (fstat(fd, &st) != 0) return fail(FMCS_E_IO, std::string("cannot stat ") + path);
  const u64 fsize = (u64)st.st_size;
  if (fsize < sizeof(FileHead)) return fail(FMCS_E_FORMAT, std::string(path) + ": too short for a column file");
  int rc = pread_all(fd, h, sizeof(FileHead), 0);
  if (rc) return rc;
  if (std::memcmp(h->magic, MAGIC, 8) != 0) return fail(FMCS_E_FORMAT, std::string(path) + ": not a column file");
  if (h->version != VERSION) return fail(FMCS_E_FORMAT, std::string(path) + ": unknown version");
  if (h->ncols == 0 || h->ncols > FMCS_COLFILE_MAX_COLS)
    return fail(FMCS_E_FORMAT, std::string(path) + ": bad column count");
  if (h->header_bytes < sizeof(FileHead) + (u64)h->ncols * sizeof(FileCol) || h->header_bytes > fsize)
    return fail(FMCS_E_FORMAT, std::string(path) + ": bad header size");
  cols->resize(h->ncols);
  rc = pread_all(fd, cols->data(), (u64)h->ncols * sizeof(FileCol), sizeof(FileHead));
  if (rc) return rc;
  for (u32 c = 0; c < h->ncols; ++c) {
    const FileCol& fc = (*cols)[c];
    const bool bank_ok = fc.bank_bytes == 1 || fc.bank_bytes == 2 || fc.bank_bytes == 4 || fc.bank_bytes == 8;
    if (!bank_ok || fc.width == 0 || fc.width > 8 * fc.bank_bytes || fc.bytes != h->nrows * fc.bank_bytes ||
        fc.offset < h->header_bytes || fc.offset + fc.bytes > fsize || fc.offset + fc.bytes < fc.offset)
      return fail(FMCS_E_FORMAT, std::string(path) + ": bad record of column " + std::to_string(c));
  }
  return FMCS_OK;
}

}  // namespace

extern "C" {

const char* fmcs_io_last_error(void) { return g_err.c_str(); }

int fmcs_text_scan(const char* path, int tpch, uint64_t* nrows, uint32_t* ncols, uint64_t* nvalues) {
  if (!path) return fail(FMCS_E_ARG, "null path");
  Mapped m;
  int rc = m.open(path);
  if (rc) return rc;
  const u32 nt = pick_threads(0, m.n);
  std::vector<u64> cnt(nt, 0);
  parallel_for(nt, [&](u32 t) {
    size_t s, e;
    chunk_bounds(m.p, m.n, t, nt, &s, &e);
    cnt[t] = count_tokens(m.p, m.n, s, e);
  });
  u64 total = 0;
  for (u64 c : cnt) total += c;
  u64 hr = 0, hc = 0;
  if (tpch) {
    size_t i = 0;
    while (i < m.n && is_ws(m.p[i])) ++i;
    if (!parse_u64(m.p, m.n, &i, &hr)) return fail(FMCS_E_FORMAT, std::string(path) + ": bad row count in header");
    while (i < m.n && is_ws(m.p[i])) ++i;
    if (!parse_u64(m.p, m.n, &i, &hc) || hc > 0xffffffffull)
      return fail(FMCS_E_FORMAT, std::string(path) + ": bad column count in header");
    total -= 2;
  }
  if (nrows) *nrows = hr;
  if (ncols) *ncols = (u32)hc;
  if (nvalues) *nvalues = total;
  return FMCS_OK;
}

int fmcs_text_read(const char* path, int tpch, uint32_t ncols, const uint32_t* widths, uint64_t nrows,
                   void* const* cols, uint64_t* cardinality, uint32_t nthreads) {
  if (!path || !widths || !cols || ncols == 0) return fail(FMCS_E_ARG, "null argument or no columns");
  for (u32 c = 0; c < ncols; ++c) {
    if (widths[c] == 0 || widths[c] > 64) return fail(FMCS_E_ARG, "column width outside 1..64");
    if (nrows && !cols[c]) return fail(FMCS_E_ARG, "null column buffer");
  }
  Mapped m;
  int rc = m.open(path);
  if (rc) return rc;
  const u32 nt = pick_threads(nthreads, m.n);
  // pass 1: tokens per chunk
  std::vector<u64> cnt(nt, 0), base(nt + 1, 0);
  std::vector<size_t> cs(nt), ce(nt);
  parallel_for(nt, [&](u32 t) {
    chunk_bounds(m.p, m.n, t, nt, &cs[t], &ce[t]);
    cnt[t] = count_tokens(m.p, m.n, cs[t], ce[t]);
  });
  for (u32 t = 0; t < nt; ++t) base[t + 1] = base[t] + cnt[t];
  const u64 skip = tpch ? 2 : 0;
  const u64 need = (u64)ncols * nrows;
  if (base[nt] < skip + need)
    return fail(FMCS_E_FORMAT, std::string(path) + ": " + std::to_string(base[nt] - std::min(base[nt], skip)) +
                                   " values, " + std::to_string(need) + " needed");
  if (tpch) {
    u64 hr = 0;
    size_t i = 0;
    while (i < m.n && is_ws(m.p[i])) ++i;
    if (!parse_u64(m.p, m.n, &i, &hr)) return fail(FMCS_E_FORMAT, std::string(path) + ": bad row count in header");
    if (hr != nrows)
      return fail(FMCS_E_ARG, "nrows " + std::to_string(nrows) + " differs from the file's " + std::to_string(hr));
  }
  std::vector<u32> bank(ncols);
  for (u32 c = 0; c < ncols; ++c) bank[c] = text_bank(widths[c]);
  // pass 2: parse; token g (after the header) is row g % nrows of column g / nrows
  std::vector<ParseErr> err(nt);
  parallel_for(nt, [&](u32 t) {
    u64 g = base[t];
    if (g >= skip + need) return;
    const char* p = m.p;
    size_t i = cs[t];
    const size_t e = ce[t];
    // position of the first value token of this chunk
    u64 col = 0, row = 0;
    if (g >= skip) {
      col = (g - skip) / std::max<u64>(nrows, 1);
      row = (g - skip) % std::max<u64>(nrows, 1);
    }
    while (i < e && g < skip + need) {
      while (i < e && is_ws(p[i])) ++i;
      if (i >= e) break;
      const size_t at = i;
      u64 v;
      if (!parse_u64(p, m.n, &i, &v)) {
        err[t].at = at;
        err[t].code = FMCS_E_FORMAT;
        err[t].msg = "not an unsigned decimal value";
        return;
      }
      if (g >= skip) {
        if (v > bank_max(bank[col])) {
          err[t].at = at;
          err[t].code = FMCS_E_FORMAT;
          err[t].msg = "value " + std::to_string(v) + " does not fit column " + std::to_string(col) + "'s " +
                       std::to_string(8 * bank[col]) + "-bit bank";
          return;
        }
        store_bank(cols[col], bank[col], row, v);
        if (++row == nrows) {
          row = 0;
          ++col;
        }
      }
      ++g;
    }
  });
  const ParseErr* first = nullptr;
  for (const auto& x : err)
    if (x.code && (!first || x.at < first->at)) first = &x;
  if (first) {
    // line number of the offending token
    u64 line = 1;
    for (size_t i = 0; i < first->at; ++i) line += m.p[i] == '\n';
    return fail(first->code, std::string(path) + ":" + std::to_string(line) + ": " + first->msg);
  }
  if (cardinality) {
    // every partition thread scans the whole column: a few suffice
    const u32 ntc = std::min<u32>(16, pick_threads(nthreads, nrows * 8));
    for (u32 c = 0; c < ncols; ++c) cardinality[c] = column_distinct(cols[c], bank[c], nrows, ntc);
  }
  return FMCS_OK;
}

int fmcs_text_write(const char* path, int tpch, uint32_t ncols, const uint32_t* widths, uint64_t nrows,
                    const void* const* cols) {
  if (!path || !widths || !cols || ncols == 0) return fail(FMCS_E_ARG, "null argument or no columns");
  for (u32 c = 0; c < ncols; ++c) {
    if (widths[c] == 0 || widths[c] > 64) return fail(FMCS_E_ARG, "column width outside 1..64");
    if (nrows && !cols[c]) return fail(FMCS_E_ARG, "null column buffer");
  }
  FILE* f = std::fopen(path, "wb");
  if (!f) return fail(FMCS_E_IO, std::string("cannot create ") + path + ": " + std::strerror(errno));
  std::vector<char> buf(1 << 20);
  size_t used = 0;
  bool ok = true;
  auto flush = [&]() {
    if (used && std::fwrite(buf.data(), 1, used, f) != used) ok = false;
    used = 0;
  };
  auto put = [&](u64 v, char end) {
    if (used + 24 > buf.size()) flush();
    char tmp[24];
    int k = 0;
    do {
      tmp[k++] = (char)('0' + v % 10);
      v /= 10;
    } while (v);
    while (k) buf[used++] = tmp[--k];
    buf[used++] = end;
  };
  if (tpch) {
    put(nrows, ' ');
    put(ncols, '\n');
  }
  for (u32 c = 0; c < ncols && ok; ++c) {
    const u32 b = text_bank(widths[c]);
    for (u64 r = 0; r < nrows; ++r) put(load_bank(cols[c], b, r), '\n');
  }
  flush();
  if (std::fclose(f) != 0) ok = false;
  if (!ok) return fail(FMCS_E_IO, std::string("write to ") + path + " failed");
  return FMCS_OK;
}

int fmcs_colfile_write(const char* path, uint64_t nrows, uint32_t ncols, const fmcs_colfile_column* meta,
                       const void* const* cols) {
  if (!path || !meta || !cols || ncols == 0 || ncols > FMCS_COLFILE_MAX_COLS)
    return fail(FMCS_E_ARG, "null argument or column count outside 1..64");
  FileHead h;
  std::memset(&h, 0, sizeof(h));
  std::memcpy(h.magic, MAGIC, 8);
  h.version = VERSION;
  h.ncols = ncols;
  h.nrows = nrows;
  h.header_bytes = round_up(sizeof(FileHead) + (u64)ncols * sizeof(FileCol), ALIGN);
  std::vector<FileCol> fc(ncols);
  u64 off = h.header_bytes;
  for (u32 c = 0; c < ncols; ++c) {
    const fmcs_colfile_column& mc = meta[c];
    const u32 bank = mc.bank_bytes ? mc.bank_bytes : text_bank(mc.width);
    if (!(bank == 1 || bank == 2 || bank == 4 || bank == 8) || mc.width == 0 || mc.width > 8 * bank)
      return fail(FMCS_E_ARG, "column " + std::to_string(c) + ": width and bank do not agree");
    if (nrows && !cols[c]) return fail(FMCS_E_ARG, "null column buffer");
    if (::strnlen(mc.name, sizeof(mc.name)) >= sizeof(mc.name))
      return fail(FMCS_E_ARG, "column " + std::to_string(c) + ": name not NUL-terminated within 16 bytes");
    std::memset(&fc[c], 0, sizeof(FileCol));
    fc[c].width = mc.width;
    fc[c].bank_bytes = bank;
    fc[c].cardinality = mc.cardinality;
    fc[c].flags = mc.flags;
    std::memcpy(fc[c].name, mc.name, sizeof(mc.name));
    fc[c].offset = off;
    fc[c].bytes = nrows * bank;
    off = round_up(off + fc[c].bytes, ALIGN);
    // values wider than the width, and the digest
    const u64 lim = mc.width >= 64 ? ~0ull : ((1ull << mc.width) - 1);
    if (lim != bank_max(bank)) {
      const u32 nt = pick_threads(0, fc[c].bytes);
      std::vector<uint8_t> bad(nt, 0);
      parallel_for(nt, [&](u32 t) {
        const u64 a = nrows * t / nt, b = nrows * (t + 1) / nt;
        for (u64 r = a; r < b; ++r)
          if (load_bank(cols[c], bank, r) > lim) {
            bad[t] = 1;
            break;
          }
      });
      for (uint8_t x : bad)
        if (x) return fail(FMCS_E_ARG, "column " + std::to_string(c) + " holds values wider than its width");
    }
    fc[c].digest = digest_parallel((const uint8_t*)cols[c], fc[c].bytes);
  }
  Fd f;
  f.fd = ::open(path, O_CREAT | O_TRUNC | O_WRONLY, 0644);
  if (f.fd < 0) return fail(FMCS_E_IO, std::string("cannot create ") + path + ": " + std::strerror(errno));
  std::vector<uint8_t> head(h.header_bytes, 0);
  std::memcpy(head.data(), &h, sizeof(h));
  std::memcpy(head.data() + sizeof(h), fc.data(), (size_t)ncols * sizeof(FileCol));
  int rc = pwrite_all(f.fd, head.data(), head.size(), 0);
  for (u32 c = 0; c < ncols && !rc; ++c) rc = pwrite_all(f.fd, cols[c], fc[c].bytes, fc[c].offset);
  if (!rc && ::ftruncate(f.fd, (off_t)off) != 0) rc = fail(FMCS_E_IO, std::string("cannot size ") + path);
  return rc;
}

int fmcs_colfile_info(const char* path, uint64_t* nrows, uint32_t* ncols, fmcs_colfile_column* meta, uint32_t cap) {
  if (!path) return fail(FMCS_E_ARG, "null path");
  Fd f;
  f.fd = ::open(path, O_RDONLY);
  if (f.fd < 0) return fail(FMCS_E_IO, std::string("cannot open ") + path + ": " + std::strerror(errno));
  FileHead h;
  std::vector<FileCol> fc;
  int rc = read_header(f.fd, path, &h, &fc);
  if (rc) return rc;
  if (nrows) *nrows = h.nrows;
  if (ncols) *ncols = h.ncols;
  if (meta)
    for (u32 c = 0; c < std::min(cap, h.ncols); ++c) {
      std::memset(&meta[c], 0, sizeof(fmcs_colfile_column));
      meta[c].width = fc[c].width;
      meta[c].bank_bytes = fc[c].bank_bytes;
      meta[c].cardinality = fc[c].cardinality;
      meta[c].flags = fc[c].flags;
      std::memcpy(meta[c].name, fc[c].name, sizeof(meta[c].name));
      meta[c].name[sizeof(meta[c].name) - 1] = 0;
    }
  return FMCS_OK;
}

int fmcs_colfile_read(const char* path, uint32_t col, uint64_t row0, uint64_t nrows, void* out) {
  if (!path || (nrows && !out)) return fail(FMCS_E_ARG, "null argument");
  Fd f;
  f.fd = ::open(path, O_RDONLY);
  if (f.fd < 0) return fail(FMCS_E_IO, std::string("cannot open ") + path + ": " + std::strerror(errno));
  FileHead h;
  std::vector<FileCol> fc;
  int rc = read_header(f.fd, path, &h, &fc);
  if (rc) return rc;
  if (col >= h.ncols) return fail(FMCS_E_ARG, "column index beyond the file's columns");
  if (row0 > h.nrows || nrows > h.nrows - row0) return fail(FMCS_E_ARG, "row range beyond the file's rows");
  const u64 b = fc[col].bank_bytes;
  return pread_all(f.fd, out, nrows * b, fc[col].offset + row0 * b);
}

int fmcs_colfile_verify(const char* path) {
  if (!path) return fail(FMCS_E_ARG, "null path");
  Mapped m;
  int rc = m.open(path);
  if (rc) return rc;
  FileHead h;
  std::vector<FileCol> fc;
  rc = read_header(m.fd, path, &h, &fc);
  if (rc) return rc;
  for (u32 c = 0; c < h.ncols; ++c) {
    const u64 d = digest_parallel((const uint8_t*)m.p + fc[c].offset, fc[c].bytes);
    if (d != fc[c].digest) return fail(FMCS_E_FORMAT, std::string(path) + ": column " + std::to_string(c) +
                                                          " does not match its digest");
  }
  return FMCS_OK;
}

}  // extern "C"
