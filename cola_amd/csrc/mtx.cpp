// MatrixMarket coordinate files <-> global CSR graph, and locality reorders
// of an existing graph.  No GPU calls in this file.
//
// Reference behaviour followed (paths relative to the CoLa tree):
//   prep_torch_par.py:14-20      load_mtx: scipy.io.mmread of <graph>.mtx
//   prep_torch_par.py:44-47,62-65  tocsr()/tocsc(): duplicates summed
//   rabbit_order_write.py:46-77  reorder, then scipy.io.mmwrite
//
// The reader maps the file, cuts the entry section into one piece per host
// thread at line boundaries and parses the pieces in parallel with
// std::from_chars; each piece files its entries into one bucket per
// contiguous row range.  Thread t then gathers bucket t of every piece in
// piece order (so the entries of a row keep file order and duplicate sums are
// deterministic), counting-sorts them by row and sorts each row by column.
#include <omp.h>

#include <algorithm>
#include <cctype>
#include <charconv>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>
#include <vector>

#include "cola_internal.hpp"

using cola::fail;
using cola::Graph;

namespace {

enum Symmetry { GENERAL = 0, SYMMETRIC = 1, SKEW = 2 };

struct Header {
  bool pattern = false, integer = false;
  Symmetry sym = GENERAL;
};

bool ieq(const std::string &a, const char *b) {
  if (a.size() != std::strlen(b)) return false;
  for (size_t i = 0; i < a.size(); ++i)
    if (std::tolower((unsigned char)a[i]) != b[i]) return false;
  return true;
}

// Read-only private mapping of a whole regular file.
struct Mapped {
  const char *data = nullptr;
  size_t size = 0;
  ~Mapped() {
    if (data) munmap(const_cast<char *>(data), size);
  }
};

cola_status map_file(const char *path, Mapped *m) {
  const int fd = open(path, O_RDONLY);
  if (fd < 0) return fail(COLA_ERR_IO, std::string("from_mtx: cannot open ") + path);
  struct stat st;
  if (fstat(fd, &st) != 0 || !S_ISREG(st.st_mode)) {
    close(fd);
    return fail(COLA_ERR_IO, std::string("from_mtx: not a regular file: ") + path);
  }
  if (st.st_size == 0) {
    close(fd);
    return fail(COLA_ERR_INVALID, std::string("from_mtx: empty file ") + path);
  }
  void *p = mmap(nullptr, (size_t)st.st_size, PROT_READ, MAP_PRIVATE, fd, 0);
  close(fd);
  if (p == MAP_FAILED) return fail(COLA_ERR_IO, std::string("from_mtx: cannot map ") + path);
  m->data = static_cast<const char *>(p);
  m->size = (size_t)st.st_size;
  return COLA_OK;
}

// End of the line starting at p: its '\n', or e for an unterminated last line.
inline const char *line_end(const char *p, const char *e) {
  const void *q = std::memchr(p, '\n', (size_t)(e - p));
  return q ? static_cast<const char *>(q) : e;
}

// Whitespace-separated tokens of [p, e).
std::vector<std::string> tokens(const char *p, const char *e) {
  std::vector<std::string> out;
  while (p < e) {
    while (p < e && (*p == ' ' || *p == '\t' || *p == '\r')) ++p;
    const char *s = p;
    while (p < e && *p != ' ' && *p != '\t' && *p != '\r') ++p;
    if (p > s) out.emplace_back(s, p);
  }
  return out;
}

cola_status parse_header(const std::vector<std::string> &t, Header *h) {
  if (t.size() != 5 || !ieq(t[0], "%%matrixmarket"))
    return fail(COLA_ERR_INVALID, "from_mtx: first line is not a %%MatrixMarket banner with 4 keywords");
  if (!ieq(t[1], "matrix")) return fail(COLA_ERR_INVALID, "from_mtx: object '" + t[1] + "' is not 'matrix'");
  if (ieq(t[2], "array")) return fail(COLA_ERR_UNSUPPORTED, "from_mtx: dense 'array' format is not a graph");
  if (!ieq(t[2], "coordinate")) return fail(COLA_ERR_INVALID, "from_mtx: unknown format '" + t[2] + "'");
  if (ieq(t[3], "complex")) return fail(COLA_ERR_UNSUPPORTED, "from_mtx: complex values are not supported");
  if (ieq(t[3], "pattern")) h->pattern = true;
  else if (ieq(t[3], "integer")) h->integer = true;
  else if (!ieq(t[3], "real") && !ieq(t[3], "double"))
    return fail(COLA_ERR_INVALID, "from_mtx: unknown field '" + t[3] + "'");
  if (ieq(t[4], "general")) h->sym = GENERAL;
  else if (ieq(t[4], "symmetric")) h->sym = SYMMETRIC;
  else if (ieq(t[4], "skew-symmetric")) h->sym = SKEW;
  else if (ieq(t[4], "hermitian")) return fail(COLA_ERR_UNSUPPORTED, "from_mtx: hermitian matrices are complex");
  else return fail(COLA_ERR_INVALID, "from_mtx: unknown symmetry '" + t[4] + "'");
  if (h->pattern && h->sym == SKEW) return fail(COLA_ERR_INVALID, "from_mtx: pattern skew-symmetric is undefined");
  return COLA_OK;
}

inline const char *skip_blank(const char *p, const char *e) {
  while (p < e && (*p == ' ' || *p == '\t' || *p == '\r')) ++p;
  return p;
}

inline bool parse_i64(const char *&p, const char *e, int64_t *v) {
  p = skip_blank(p, e);
  if (p < e && *p == '+') ++p;
  auto r = std::from_chars(p, e, *v);
  if (r.ec != std::errc() || r.ptr == p) return false;
  p = r.ptr;
  return p == e || *p == ' ' || *p == '\t' || *p == '\r';
}

inline bool parse_f64(const char *&p, const char *e, double *v) {
  p = skip_blank(p, e);
  if (p < e && *p == '+') ++p;
  auto r = std::from_chars(p, e, *v);
  if (r.ec != std::errc() || r.ptr == p) return false;
  p = r.ptr;
  return p == e || *p == ' ' || *p == '\t' || *p == '\r';
}

struct Entry {
  int32_t r, c;
  double v;
};

// Row range of bucket b of T: rows r with r*T/n == b, i.e. [ceil(b*n/T), ...).
inline int64_t bucket_begin(int64_t b, int64_t T, int64_t n) { return (b * n + T - 1) / T; }

struct Piece {
  std::vector<std::vector<Entry>> bucket;  // one per row range
  int64_t entries = 0;                     // lines parsed (before mirroring)
  std::string err;
  cola_status st = COLA_OK;
};

// Parses the entry lines of [b, e).
void parse_piece(const char *b, const char *e, const Header &h, int64_t n, int T, Piece *out) {
  out->bucket.assign(T, {});
  const int64_t guess = std::count(b, e, '\n') / T + 16;
  for (auto &bk : out->bucket) bk.reserve((size_t)(h.sym != GENERAL ? 2 * guess : guess));
  const char *p = b;
  while (p < e) {
    const char *eol = line_end(p, e);
    const char *q = skip_blank(p, eol);
    if (q == eol || *q == '%') {  // blank line or comment
      p = eol + 1;
      continue;
    }
    int64_t i, j;
    double v = 1.0;
    bool ok = parse_i64(q, eol, &i) && parse_i64(q, eol, &j);
    if (ok && !h.pattern) {
      if (h.integer) {
        int64_t iv = 0;
        ok = parse_i64(q, eol, &iv);
        v = (double)iv;
      } else {
        ok = parse_f64(q, eol, &v);
      }
    }
    if (ok) ok = skip_blank(q, eol) == eol;
    if (!ok) {
      out->st = COLA_ERR_INVALID;
      out->err = "from_mtx: malformed entry line '" + std::string(p, std::min<size_t>(eol - p, 80)) + "'";
      return;
    }
    if (i < 1 || i > n || j < 1 || j > n) {
      out->st = COLA_ERR_INVALID;
      out->err = "from_mtx: entry (" + std::to_string(i) + ", " + std::to_string(j) +
                 ") outside the 1-based " + std::to_string(n) + " x " + std::to_string(n) + " matrix";
      return;
    }
    --i;
    --j;
    out->bucket[i * T / n].push_back({(int32_t)i, (int32_t)j, v});
    if (h.sym != GENERAL && i != j)
      out->bucket[j * T / n].push_back({(int32_t)j, (int32_t)i, h.sym == SKEW ? -v : v});
    ++out->entries;
    p = eol + 1;
  }
}

}  // namespace

extern "C" {

cola_status cola_graph_from_mtx(const char *path, int values, cola_graph *out) {
  if (!out) return fail(COLA_ERR_INVALID, "from_mtx: out is NULL");
  *out = nullptr;
  if (!path) return fail(COLA_ERR_INVALID, "from_mtx: path is NULL");
  if (values < COLA_MTX_VALUES_FILE || values > COLA_MTX_VALUES_GCN)
    return fail(COLA_ERR_INVALID, "from_mtx: values must be FILE, ONE or GCN");
  Mapped file;
  cola_status st = map_file(path, &file);
  if (st) return st;
  const char *p = file.data, *end = file.data + file.size;
  // Banner.
  const char *eol = line_end(p, end);
  Header h;
  if ((st = parse_header(tokens(p, eol), &h))) return st;
  p = eol + 1;
  // Comments, then the size line.
  int64_t rows = -1, cols = -1, nnz = -1;
  while (p < end) {
    eol = line_end(p, end);
    const char *q = skip_blank(p, eol);
    if (q == eol || *q == '%') {
      p = eol + 1;
      continue;
    }
    if (!parse_i64(q, eol, &rows) || !parse_i64(q, eol, &cols) || !parse_i64(q, eol, &nnz) ||
        skip_blank(q, eol) != eol || rows < 0 || cols < 0 || nnz < 0)
      return fail(COLA_ERR_INVALID, "from_mtx: size line must hold three non-negative integers");
    p = eol + 1;
    break;
  }
  if (nnz < 0) return fail(COLA_ERR_IO, "from_mtx: file ends before the size line");
  if (rows != cols)
    return fail(COLA_ERR_UNSUPPORTED, "from_mtx: matrix is " + std::to_string(rows) + " x " +
                                          std::to_string(cols) + "; a graph adjacency must be square");
  if (rows > INT32_MAX) return fail(COLA_ERR_INVALID, "from_mtx: more than 2^31-1 nodes");
  const int64_t n = rows;
  p = std::min(p, end);
  // Pieces cut at line boundaries, parsed in parallel; T pieces and T row buckets.
  const int T = (int)std::max<int64_t>(1, std::min<int64_t>({omp_get_max_threads(), ((end - p) >> 16) + 1,
                                                             std::max<int64_t>(n, 1)}));
  std::vector<const char *> cut(T + 1);
  cut[0] = p;
  cut[T] = end;
  for (int t = 1; t < T; ++t) {
    const char *c = std::max(p + (end - p) * t / T, cut[t - 1]);
    cut[t] = c < end ? std::min(line_end(c, end) + 1, end) : end;
  }
  std::vector<Piece> pieces(T);
#pragma omp parallel for schedule(static, 1) num_threads(T)
  for (int t = 0; t < T; ++t) {
    try {
      parse_piece(cut[t], cut[t + 1], h, n, T, &pieces[t]);
    } catch (const std::bad_alloc &) {
      pieces[t].st = COLA_ERR_NOMEM;
      pieces[t].err = "from_mtx: out of host memory";
    }
  }
  int64_t lines = 0;
  for (const Piece &pc : pieces) {
    if (pc.st) return fail(pc.st, pc.err);
    lines += pc.entries;
  }
  if (lines != nnz)
    return fail(COLA_ERR_INVALID, "from_mtx: size line declares " + std::to_string(nnz) + " entries, file holds " +
                                      std::to_string(lines));
  auto *gh = new (std::nothrow) cola_graph_s;
  if (!gh) return fail(COLA_ERR_NOMEM, "from_mtx: out of host memory");
  Graph &g = gh->g;
  g.n = n;
  try {
    // Bucket t = rows [rb[t], rb[t+1]), entries at [eb[t], eb[t+1]).
    std::vector<int64_t> rb(T + 1), eb(T + 1, 0);
    for (int t = 0; t <= T; ++t) rb[t] = bucket_begin(t, T, n);
    for (int t = 0; t < T; ++t) {
      eb[t + 1] = eb[t];
      for (const Piece &pc : pieces) eb[t + 1] += (int64_t)pc.bucket[t].size();
    }
    std::vector<int64_t> start(n, 0);  // first entry of each row in cs / vs
    std::vector<int64_t> uniq(n, 0);
    std::vector<int32_t> cs(eb[T]);
    std::vector<double> vs(eb[T]);
#pragma omp parallel for schedule(static, 1) num_threads(T)
    for (int t = 0; t < T; ++t) {
      const int64_t r0 = rb[t], r1 = rb[t + 1];
      // Stable counting sort of bucket t by row: file order within a row.
      // off[i - r0] = first entry of row i; only thread t touches rows
      // [r0, r1) of start and uniq.
      std::vector<int64_t> off(r1 - r0 + 1, 0);
      for (const Piece &pc : pieces)
        for (const Entry &x : pc.bucket[t]) ++off[x.r - r0 + 1];
      off[0] = eb[t];
      for (int64_t k = 0; k < r1 - r0; ++k) off[k + 1] += off[k];
      std::vector<int64_t> pos(off.begin(), off.end() - 1);
      for (const Piece &pc : pieces)
        for (const Entry &x : pc.bucket[t]) {
          const int64_t w = pos[x.r - r0]++;
          cs[w] = x.c;
          vs[w] = x.v;
        }
      // Per row: stable sort by column, duplicates summed in FP64.
      std::vector<std::pair<int32_t, double>> row;
      for (int64_t i = r0; i < r1; ++i) {
        const int64_t b = off[i - r0], e = off[i - r0 + 1];
        start[i] = b;
        row.clear();
        for (int64_t k = b; k < e; ++k) row.emplace_back(cs[k], vs[k]);
        std::stable_sort(row.begin(), row.end(),
                         [](const auto &x, const auto &y) { return x.first < y.first; });
        int64_t w = b;
        for (size_t k = 0; k < row.size(); ++k) {
          if (k && row[k].first == row[k - 1].first) {
            vs[w - 1] += row[k].second;
          } else {
            cs[w] = row[k].first;
            vs[w] = row[k].second;
            ++w;
          }
        }
        uniq[i] = w - b;
      }
    }
    std::vector<Piece>().swap(pieces);
    g.row_ptr.assign(n + 1, 0);
    for (int64_t i = 0; i < n; ++i) g.row_ptr[i + 1] = g.row_ptr[i] + uniq[i];
    if (g.row_ptr[n] > INT32_MAX) {
      delete gh;
      return fail(COLA_ERR_INVALID, "from_mtx: nnz overflows int32");
    }
    g.col_idx.resize(g.row_ptr[n]);
    g.val.resize(g.row_ptr[n]);
#pragma omp parallel for schedule(dynamic, 4096)
    for (int64_t i = 0; i < n; ++i)
      for (int64_t k = 0; k < uniq[i]; ++k) {
        g.col_idx[g.row_ptr[i] + k] = cs[start[i] + k];
        g.val[g.row_ptr[i] + k] = (float)vs[start[i] + k];
      }
    if (values != COLA_MTX_VALUES_FILE) cola::assign_values(&g, values == COLA_MTX_VALUES_GCN);
  } catch (const std::bad_alloc &) {
    delete gh;
    return fail(COLA_ERR_NOMEM, "from_mtx: out of host memory");
  }
  *out = gh;
  return COLA_OK;
}

cola_status cola_graph_write_mtx(cola_graph g, const char *path, int with_values) {
  if (!g || !path) return fail(COLA_ERR_INVALID, "write_mtx: NULL graph or path");
  const Graph &G = g->g;
  const int64_t n = G.n;
  const int T = std::max(1, std::min<int>(omp_get_max_threads(), (int)(n >> 12) + 1));
  std::vector<std::string> text(T);
#pragma omp parallel for schedule(static, 1) num_threads(T)
  for (int t = 0; t < T; ++t) {
    const int64_t r0 = n * t / T, r1 = n * (t + 1) / T;
    std::string &s = text[t];
    s.reserve((size_t)(G.row_ptr[r1] - G.row_ptr[r0]) * (with_values ? 30 : 16));
    char line[128];  // a line needs at most 11 + 1 + 11 + 1 + 16 + 1 bytes
    for (int64_t i = r0; i < r1; ++i)
      for (int64_t k = G.row_ptr[i]; k < G.row_ptr[i + 1]; ++k) {
        char *q = line, *e = line + 96;
        q = std::to_chars(q, e, i + 1).ptr;
        *q++ = ' ';
        q = std::to_chars(q, e, (int64_t)G.col_idx[k] + 1).ptr;
        if (with_values) {
          *q++ = ' ';
          q = std::to_chars(q, e, G.val[k], std::chars_format::general, 9).ptr;
        }
        *q++ = '\n';
        s.append(line, q);
      }
  }
  std::FILE *f = std::fopen(path, "wb");
  if (!f) return fail(COLA_ERR_IO, std::string("write_mtx: cannot create ") + path);
  const std::string head = std::string("%%MatrixMarket matrix coordinate ") + (with_values ? "real" : "pattern") +
                           " general\n" + std::to_string(n) + " " + std::to_string(n) + " " +
                           std::to_string(G.col_idx.size()) + "\n";
  bool ok = std::fwrite(head.data(), 1, head.size(), f) == head.size();
  for (int t = 0; ok && t < T; ++t) ok = std::fwrite(text[t].data(), 1, text[t].size(), f) == text[t].size();
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) return fail(COLA_ERR_IO, std::string("write_mtx: short write to ") + path);
  return COLA_OK;
}

cola_status cola_graph_reorder(cola_graph g, int method, int32_t *new2old, cola_graph *out) {
  if (!out) return fail(COLA_ERR_INVALID, "graph_reorder: out is NULL");
  *out = nullptr;
  if (!g) return fail(COLA_ERR_INVALID, "graph_reorder: NULL graph");
  const Graph &G = g->g;
  const int64_t n = G.n;
  std::vector<int32_t> order;
  switch (method) {
    case COLA_ORDER_CUTHILL_MCKEE:
      order = cola::cuthill_mckee(G);
      break;
    case COLA_ORDER_REVERSE_CM:
      order = cola::cuthill_mckee(G);
      std::reverse(order.begin(), order.end());
      break;
    case COLA_ORDER_COMMUNITY:
      order = cola::community_order(G);
      break;
    case COLA_ORDER_DEGREE: {
      order.resize(n);
      for (int64_t i = 0; i < n; ++i) order[i] = (int32_t)i;
      std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) {
        return G.row_ptr[x + 1] - G.row_ptr[x] > G.row_ptr[y + 1] - G.row_ptr[y];
      });
      break;
    }
    default:
      return fail(COLA_ERR_INVALID, "graph_reorder: unknown method " + std::to_string(method));
  }
  {
    std::vector<uint8_t> hit(n, 0);
    bool ok = (int64_t)order.size() == n;
    for (size_t i = 0; ok && i < order.size(); ++i) {
      const int32_t v = order[i];
      ok = v >= 0 && v < n && !hit[v];
      if (ok) hit[v] = 1;
    }
    if (!ok) return fail(COLA_ERR_INVALID, "graph_reorder: the order is not a permutation");
  }
  auto *h = new (std::nothrow) cola_graph_s;
  if (!h) return fail(COLA_ERR_NOMEM, "graph_reorder: out of host memory");
  cola::relabel(G, order, &h->g);
  if (new2old) std::copy(order.begin(), order.end(), new2old);
  *out = h;
  return COLA_OK;
}

}  // extern "C"
