// MatrixMarket coordinate files <-> host CSR, in time linear in the file size (host code of the bclx C ABI).
//
// Reading follows the reference's mmread + CSRMatrix::read_MatrixMarket
// (bcl/containers/sequential/matrix_io.hpp:221-357, bcl/containers/sequential/CSRMatrix.hpp:369-411):
//   - the banner must be "%%MatrixMarket matrix coordinate <field> <symmetry>"; field "pattern" gives every entry
//     the value 1, any other field is read as one number per entry (the first number of the line); symmetry is
//     "general" or "symmetric" (an off-diagonal entry (i, j) of a symmetric file also stands for (j, i)); anything
//     else is rejected;
//   - comment lines start with '%'; the size line holds m, n and the stored entry count;
//   - indices are 1-based unless one_indexed is 0; an index outside the matrix, or more entry lines than the size
//     line announces, is an error;
//   - an entry whose (i, j) was already seen is dropped: the first occurrence in file order wins, with the mirror
//     of a symmetric entry inserted right after it (coo_matrix::insert, matrix_io.hpp:166-177);
//   - the result is sorted by (i, j).
// The reference does this with a linear scan per inserted entry (quadratic in the entry count) and one
// istringstream per line. Here the file is memory-mapped and its entry lines are cut into chunks parsed by
// parallel threads, the rows are formed by a stable counting sort, and each row is sorted by column and
// de-duplicated on its own: the same result in O(file size + nnz log(row length)).
// Differences, all on files the reference mishandles: blank lines are skipped (the reference parses them as entries
// with undefined indices), negative indices are rejected (a signed index_type lets the reference accept them), and
// a value that does not parse is an error.
//
// Writing follows CSRMatrix::write_MatrixMarket (CSRMatrix.hpp:532-548) byte for byte: the banner
// "%%MatrixMarket matrix coordinate integer general", the size line, then "i j v" per entry with 1-based indices
// and the value printed by "%f".
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <cerrno>
#include <climits>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../include/bclx.h"

namespace bclx {
void set_error(const std::string& msg);
}

namespace {

struct MmError {
  bclx_status status;
  std::string msg;
};

template <typename F>
bclx_status mm_guarded(F&& f) {
  try {
    f();
    return BCLX_OK;
  } catch (const MmError& e) {
    bclx::set_error(e.msg);
    return e.status;
  } catch (const std::bad_alloc&) {
    bclx::set_error("host allocation failed");
    return BCLX_ERR_ALLOC;
  } catch (const std::exception& e) {
    bclx::set_error(e.what());
    return BCLX_ERR_INVALID;
  }
}

[[noreturn]] void fail(bclx_status s, const std::string& msg) { throw MmError{s, msg}; }

int worker_count() {
  unsigned hw = std::thread::hardware_concurrency();
  if (const char* e = getenv("OMP_NUM_THREADS")) {
    const long v = strtol(e, nullptr, 10);
    if (v > 0) hw = (unsigned)v;
  }
  return (int)std::min<unsigned>(std::max<unsigned>(hw, 1u), 64u);
}

template <typename F>
void parallel_for(int nt, F&& f) {
  if (nt <= 1) {
    f(0);
    return;
  }
  std::vector<std::thread> th;
  for (int t = 1; t < nt; t++) th.emplace_back([&f, t] { f(t); });
  f(0);
  for (auto& x : th) x.join();
}

// Read-only mapping of a whole file.
struct Mapped {
  int fd = -1;
  const char* p = nullptr;
  size_t len = 0;
  explicit Mapped(const char* path) {
    fd = open(path, O_RDONLY);
    if (fd < 0) fail(BCLX_ERR_INVALID, std::string("mmread: cannot open ") + path);
    struct stat st;
    if (fstat(fd, &st) != 0) fail(BCLX_ERR_INVALID, std::string("mmread: cannot stat ") + path);
    len = (size_t)st.st_size;
    if (len > 0) {
      void* a = mmap(nullptr, len, PROT_READ, MAP_PRIVATE, fd, 0);
      if (a == MAP_FAILED) fail(BCLX_ERR_INVALID, std::string("mmread: cannot map ") + path);
      p = static_cast<const char*>(a);
    }
  }
  ~Mapped() {
    if (p) munmap(const_cast<char*>(p), len);
    if (fd >= 0) close(fd);
  }
};

inline bool is_space(char c) { return c == ' ' || c == '\t' || c == '\r' || c == '\v' || c == '\f'; }

// Next line [b, e) starting at pos; returns the position after its newline.
inline size_t next_line(const char* s, size_t len, size_t pos, size_t* b, size_t* e) {
  const void* nl = pos < len ? memchr(s + pos, '\n', len - pos) : nullptr;
  const size_t end = nl ? (size_t)(static_cast<const char*>(nl) - s) : len;
  *b = pos;
  *e = end;
  return nl ? end + 1 : len;
}

std::vector<std::string> tokens(const char* b, const char* e) {
  std::vector<std::string> t;
  while (b < e) {
    while (b < e && is_space(*b)) b++;
    const char* s = b;
    while (b < e && !is_space(*b)) b++;
    if (b > s) t.emplace_back(s, b);
  }
  return t;
}

struct Header {
  bool pattern = false, symmetric = false;
  int64_t m = 0, n = 0, nnz = 0;
  size_t body = 0;  // offset of the first entry line
};

// Parses a signed integer token at s (bounded by e); advances s. Returns false when there is none.
inline bool parse_i64(const char*& s, const char* e, int64_t* out) {
  while (s < e && is_space(*s)) s++;
  bool neg = false;
  if (s < e && (*s == '-' || *s == '+')) neg = *s++ == '-';
  if (s >= e || *s < '0' || *s > '9') return false;
  uint64_t v = 0;
  while (s < e && *s >= '0' && *s <= '9') {
    v = v * 10 + (uint64_t)(*s - '0');
    if (v > (uint64_t)INT64_MAX) return false;
    s++;
  }
  if (s < e && !is_space(*s)) return false;
  *out = neg ? -(int64_t)v : (int64_t)v;
  return true;
}

// One number token parsed as T the way an istream >> T does (strtof / strtod, round to nearest).
template <typename T>
inline bool parse_val(const char*& s, const char* e, T* out) {
  while (s < e && is_space(*s)) s++;
  if (s >= e) return false;
  char buf[128];
  const char* t = s;
  while (t < e && !is_space(*t)) t++;
  const size_t n = (size_t)(t - s);
  if (n == 0 || n >= sizeof(buf)) return false;
  memcpy(buf, s, n);
  buf[n] = '\0';
  char* end = nullptr;
  errno = 0;
  T v;
  if (sizeof(T) == 4) v = (T)strtof(buf, &end);
  else v = (T)strtod(buf, &end);
  if (end != buf + n) return false;
  *out = v;
  s = t;
  return true;
}

Header read_header(const Mapped& f, const char* path) {
  Header h;
  const std::string what = std::string(path) + " could not be parsed as a Matrix Market file.";
  size_t b, e;
  size_t pos = next_line(f.p, f.len, 0, &b, &e);
  auto t = tokens(f.p + b, f.p + e);
  if (t.size() < 1 || t[0] != "%%MatrixMarket") fail(BCLX_ERR_INVALID, what);
  if (t.size() < 2 || t[1] != "matrix") fail(BCLX_ERR_INVALID, what);
  if (t.size() < 3 || t[2] != "coordinate") fail(BCLX_ERR_INVALID, what);
  h.pattern = t.size() > 3 && t[3] == "pattern";
  const std::string sym = t.size() > 4 ? t[4] : "";
  if (sym == "general") h.symmetric = false;
  else if (sym == "symmetric") h.symmetric = true;
  else fail(BCLX_ERR_INVALID, std::string(path) + " has an unsupported matrix type");
  for (;;) {  // comments and blank lines, then the size line
    if (pos >= f.len) fail(BCLX_ERR_INVALID, what + " (no size line)");
    pos = next_line(f.p, f.len, pos, &b, &e);
    const char* s = f.p + b;
    while (s < f.p + e && is_space(*s)) s++;
    if (s == f.p + e || *s == '%') continue;
    const char* q = f.p + b;
    if (!parse_i64(q, f.p + e, &h.m) || !parse_i64(q, f.p + e, &h.n) || !parse_i64(q, f.p + e, &h.nnz) ||
        h.m < 0 || h.n < 0 || h.nnz < 0)
      fail(BCLX_ERR_INVALID, what + " (bad size line)");
    break;
  }
  h.body = pos;
  return h;
}

struct Coo {
  std::vector<int64_t> i, j;  // entries in file order, mirrors right after their entry
};

template <typename T>
struct Chunk {
  std::vector<int64_t> i, j;
  std::vector<T> v;
  int64_t lines = 0;         // entry lines in this chunk
  int64_t first_bad = -1;    // entry line (within the chunk) of the first out-of-bounds or malformed entry
  bool bad_parse = false;
};

template <typename T>
void parse_chunk(const Mapped& f, size_t b0, size_t b1, const Header& h, bool one_indexed, Chunk<T>& c) {
  size_t pos = b0;
  const int64_t base = one_indexed ? 1 : 0;
  while (pos < b1) {
    size_t lb, le;
    pos = next_line(f.p, f.len, pos, &lb, &le);
    const char* s = f.p + lb;
    const char* e = f.p + le;
    const char* q = s;
    while (q < e && is_space(*q)) q++;
    if (q == e) continue;  // blank line
    const int64_t line = c.lines++;
    if (c.first_bad >= 0) continue;
    int64_t i, j;
    T v = T(1);
    if (!parse_i64(s, e, &i) || !parse_i64(s, e, &j) || (!h.pattern && !parse_val<T>(s, e, &v))) {
      c.first_bad = line;
      c.bad_parse = true;
      continue;
    }
    i -= base;
    j -= base;
    if (i < 0 || j < 0 || i >= h.m || j >= h.n) {
      c.first_bad = line;
      continue;
    }
    c.i.push_back(i);
    c.j.push_back(j);
    c.v.push_back(v);
    if (h.symmetric && i != j) {
      c.i.push_back(j);
      c.j.push_back(i);
      c.v.push_back(v);
    }
  }
}

template <typename T, typename RP>
void read_csr(const char* path, bool one_indexed, int32_t ptr_width, bclx_csr* out) {
  Mapped f(path);
  if (f.len == 0) fail(BCLX_ERR_INVALID, std::string(path) + " could not be parsed as a Matrix Market file.");
  const Header h = read_header(f, path);
  if (h.n > (int64_t)INT32_MAX) fail(BCLX_ERR_OVERFLOW, "read_MatrixMarket: more than 2^31 - 1 columns");
  // chunks of whole lines
  const int nt = std::max(1, std::min<int>(worker_count(), (int)((f.len - h.body) / (1 << 16)) + 1));
  std::vector<size_t> cut(nt + 1);
  cut[0] = h.body;
  cut[nt] = f.len;
  for (int t = 1; t < nt; t++) {
    size_t p = h.body + (f.len - h.body) * (size_t)t / (size_t)nt;
    p = std::max(p, cut[t - 1]);
    const void* nl = p < f.len ? memchr(f.p + p, '\n', f.len - p) : nullptr;
    cut[t] = nl ? (size_t)(static_cast<const char*>(nl) - f.p) + 1 : f.len;
  }
  std::vector<Chunk<T>> ch(nt);
  parallel_for(nt, [&](int t) { parse_chunk<T>(f, cut[t], cut[t + 1], h, one_indexed, ch[t]); });
  // errors in file order: a bad entry line, or an entry line past the announced count
  int64_t before = 0;
  for (int t = 0; t < nt; t++) {
    if (ch[t].first_bad >= 0) {
      const int64_t line = before + ch[t].first_bad;
      if (line >= h.nnz)
        fail(BCLX_ERR_INVALID,
             "read_MatrixMarket: error reading Matrix Market file, file has more nonzeros than reported.");
      fail(BCLX_ERR_INVALID, ch[t].bad_parse ? "read_MatrixMarket: malformed entry line " + std::to_string(line + 1)
                                             : std::string("read_MatrixMarket: file has nonzero out of bounds."));
    }
    before += ch[t].lines;
  }
  if (before > h.nnz)
    fail(BCLX_ERR_INVALID, "read_MatrixMarket: error reading Matrix Market file, file has more nonzeros than reported.");
  // stable counting sort of the entries by row (file order kept within a row)
  const int64_t m = h.m;
  std::vector<std::vector<int64_t>> hist(nt, std::vector<int64_t>(m + 1, 0));
  parallel_for(nt, [&](int t) {
    for (int64_t r : ch[t].i) hist[t][r]++;
  });
  std::vector<int64_t> rstart(m + 1, 0);
  for (int64_t r = 0; r < m; r++) {
    int64_t s = 0;
    for (int t = 0; t < nt; t++) s += hist[t][r];
    rstart[r + 1] = rstart[r] + s;
  }
  const int64_t total = rstart[m];
  // per chunk, the first slot of each row: the row's start plus its entries in earlier chunks (threads split rows)
  parallel_for(nt, [&](int t) {
    const int64_t r0 = m * t / nt, r1 = m * (t + 1) / nt;
    for (int64_t r = r0; r < r1; r++) {
      int64_t s = rstart[r];
      for (int u = 0; u < nt; u++) {
        const int64_t c = hist[u][r];
        hist[u][r] = s;
        s += c;
      }
    }
  });
  std::vector<int32_t> col(std::max<int64_t>(total, 1));
  std::vector<T> val(std::max<int64_t>(total, 1));
  parallel_for(nt, [&](int t) {
    auto& c = ch[t];
    for (size_t q = 0; q < c.i.size(); q++) {
      const int64_t d = hist[t][c.i[q]]++;
      col[d] = (int32_t)c.j[q];
      val[d] = c.v[q];
    }
    std::vector<int64_t>().swap(c.i);
    std::vector<int64_t>().swap(c.j);
    std::vector<T>().swap(c.v);
  });
  hist.clear();
  // each row: stable sort by column, keep the first of equal columns
  std::vector<int64_t> keep(m + 1, 0);
  std::atomic<int64_t> next_row{0};
  parallel_for(nt, [&](int) {
    std::vector<std::pair<int32_t, int64_t>> tmp;
    std::vector<T> tv;
    for (;;) {
      const int64_t r0 = next_row.fetch_add(4096);
      if (r0 >= m) break;
      const int64_t r1 = std::min<int64_t>(m, r0 + 4096);
      for (int64_t r = r0; r < r1; r++) {
        const int64_t a = rstart[r], b = rstart[r + 1];
        if (b - a <= 1) {
          keep[r + 1] = b - a;
          continue;
        }
        bool sorted_unique = true;
        for (int64_t q = a + 1; q < b && sorted_unique; q++) sorted_unique = col[q - 1] < col[q];
        if (sorted_unique) {
          keep[r + 1] = b - a;
          continue;
        }
        tmp.clear();
        for (int64_t q = a; q < b; q++) tmp.emplace_back(col[q], q);
        std::stable_sort(tmp.begin(), tmp.end(),
                         [](const auto& x, const auto& y) { return x.first < y.first; });
        tv.clear();
        for (auto& x : tmp) tv.push_back(val[x.second]);
        int64_t w = a;
        for (size_t q = 0; q < tmp.size(); q++) {
          if (q > 0 && tmp[q].first == tmp[q - 1].first) continue;
          col[w] = tmp[q].first;
          val[w] = tv[q];
          w++;
        }
        keep[r + 1] = w - a;
      }
    }
  });
  for (int64_t r = 0; r < m; r++) keep[r + 1] += keep[r];
  const int64_t nnz = keep[m];
  if (ptr_width == 4 && nnz > (int64_t)INT32_MAX)
    fail(BCLX_ERR_OVERFLOW, "read_MatrixMarket: " + std::to_string(nnz) + " entries do not fit a 4-byte row_ptr");
  RP* rp = static_cast<RP*>(malloc(sizeof(RP) * (size_t)(m + 1)));
  int32_t* oc = static_cast<int32_t*>(malloc(sizeof(int32_t) * (size_t)std::max<int64_t>(nnz, 1)));
  T* ov = static_cast<T*>(malloc(sizeof(T) * (size_t)std::max<int64_t>(nnz, 1)));
  if (!rp || !oc || !ov) {
    free(rp);
    free(oc);
    free(ov);
    throw std::bad_alloc();
  }
  std::atomic<int64_t> next{0};
  parallel_for(nt, [&](int) {
    for (;;) {
      const int64_t r0 = next.fetch_add(16384);
      if (r0 >= m) break;
      const int64_t r1 = std::min<int64_t>(m, r0 + 16384);
      for (int64_t r = r0; r < r1; r++) {
        rp[r] = (RP)keep[r];
        const int64_t cnt = keep[r + 1] - keep[r];
        if (cnt > 0) {
          memcpy(oc + keep[r], col.data() + rstart[r], sizeof(int32_t) * (size_t)cnt);
          memcpy(ov + keep[r], val.data() + rstart[r], sizeof(T) * (size_t)cnt);
        }
      }
    }
  });
  rp[m] = (RP)nnz;
  out->m = h.m;
  out->n = h.n;
  out->nnz = nnz;
  out->ptr_width = ptr_width;
  out->dtype = sizeof(T) == 4 ? BCLX_F32 : BCLX_F64;
  out->row_ptr = rp;
  out->col_idx = oc;
  out->vals = ov;
}

template <typename T, typename RP>
void write_csr(const char* path, const bclx_csr* a) {
  const RP* rp = static_cast<const RP*>(a->row_ptr);
  const T* v = static_cast<const T*>(a->vals);
  const int64_t m = a->m;
  // rows formatted by parallel threads into their own buffers, written in row order
  const int nt = std::max(1, std::min<int>(worker_count(), (int)(a->nnz / 65536) + 1));
  std::vector<std::string> part(nt);
  parallel_for(nt, [&](int t) {
    const int64_t r0 = m * t / nt, r1 = m * (t + 1) / nt;
    std::string& s = part[t];
    char buf[512];
    for (int64_t i = r0; i < r1; i++)
      for (int64_t q = (int64_t)rp[i]; q < (int64_t)rp[i + 1]; q++) {
        const int k = snprintf(buf, sizeof(buf), "%d %d %f\n", (int)(i + 1), (int)(a->col_idx[q] + 1), (double)v[q]);
        s.append(buf, (size_t)k);
      }
  });
  FILE* f = fopen(path, "w");
  if (!f) fail(BCLX_ERR_INVALID, std::string("write_MatrixMarket: cannot open ") + path);
  bool ok = fprintf(f, "%%%%MatrixMarket matrix coordinate integer general\n") > 0 &&
            fprintf(f, "%lu %lu %lu\n", (unsigned long)a->m, (unsigned long)a->n, (unsigned long)a->nnz) > 0;
  for (auto& s : part) ok = ok && fwrite(s.data(), 1, s.size(), f) == s.size();
  ok = (fclose(f) == 0) && ok;
  if (!ok) fail(BCLX_ERR_INVALID, std::string("write_MatrixMarket: short write to ") + path);
}

void check_host_csr(const bclx_csr* a) {
  if (!a || a->m < 0 || a->n < 0 || a->nnz < 0) fail(BCLX_ERR_INVALID, "bad matrix descriptor");
  if (a->ptr_width != 4 && a->ptr_width != 8) fail(BCLX_ERR_INVALID, "ptr_width must be 4 or 8");
  if (a->dtype != BCLX_F32 && a->dtype != BCLX_F64) fail(BCLX_ERR_INVALID, "dtype must be BCLX_F32 or BCLX_F64");
  if (!a->row_ptr || (a->nnz > 0 && (!a->col_idx || !a->vals))) fail(BCLX_ERR_INVALID, "null array");
}

}  // namespace

extern "C" {

bclx_status bclx_mm_info(const char* path, int32_t one_indexed, int64_t* m, int64_t* n, int64_t* nnz) {
  (void)one_indexed;  // the size line does not depend on the index base
  return mm_guarded([&] {
    if (!path || !m || !n || !nnz) fail(BCLX_ERR_INVALID, "null argument");
    Mapped f(path);
    if (f.len == 0) fail(BCLX_ERR_INVALID, std::string(path) + " could not be parsed as a Matrix Market file.");
    const Header h = read_header(f, path);
    *m = h.m;
    *n = h.n;
    *nnz = h.nnz;
  });
}

bclx_status bclx_mm_read(const char* path, int32_t one_indexed, int32_t dtype, int32_t ptr_width, bclx_csr* out) {
  return mm_guarded([&] {
    if (!path || !out) fail(BCLX_ERR_INVALID, "null argument");
    if (ptr_width != 4 && ptr_width != 8) fail(BCLX_ERR_INVALID, "ptr_width must be 4 or 8");
    const bool one = one_indexed != 0;
    if (dtype == BCLX_F32 && ptr_width == 4) read_csr<float, int32_t>(path, one, 4, out);
    else if (dtype == BCLX_F32) read_csr<float, int64_t>(path, one, 8, out);
    else if (dtype == BCLX_F64 && ptr_width == 4) read_csr<double, int32_t>(path, one, 4, out);
    else if (dtype == BCLX_F64) read_csr<double, int64_t>(path, one, 8, out);
    else fail(BCLX_ERR_INVALID, "dtype must be BCLX_F32 or BCLX_F64");
  });
}

bclx_status bclx_mm_write(const char* path, const bclx_csr* a) {
  return mm_guarded([&] {
    if (!path) fail(BCLX_ERR_INVALID, "null path");
    check_host_csr(a);
    if (a->dtype == BCLX_F32 && a->ptr_width == 4) write_csr<float, int32_t>(path, a);
    else if (a->dtype == BCLX_F32) write_csr<float, int64_t>(path, a);
    else if (a->ptr_width == 4) write_csr<double, int32_t>(path, a);
    else write_csr<double, int64_t>(path, a);
  });
}

void bclx_host_free(void* p) { free(p); }

}  // extern "C"
