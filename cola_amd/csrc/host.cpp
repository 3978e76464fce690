// Host side of the single-partition SpMM engine: synthetic graph generation,
// the on-disk partition layout (write + load), the reference's nnz_split and
// the degree-binned plan builder.  No GPU calls in this file.
//
// Reference behaviour followed (paths relative to the CoLa tree):
//   prep_torch_par.py:24-101      row_partition / serialize (file layout)
//   src/distSpMM_ppl_batch.cu:49-99  partition load + value override
//   include/dist_spmm_ppl_batch.cuh:7-27   bin2vec
//   include/dist_spmm_ppl_batch.cuh:29-56  nnz_split
#include <omp.h>

#include <algorithm>
#include <atomic>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <numeric>
#include <sys/stat.h>

#include "cola_internal.hpp"

namespace cola {

static thread_local std::string g_last_error;

cola_status fail(cola_status code, const std::string &msg) {
  g_last_error = msg;
  return code;
}

// ---------------------------------------------------------------------------
// Pseudo-random numbers: SplitMix64 (Steele, Lea & Flood, 2014) seeds one
// xoshiro256** stream (Blackman & Vigna, 2018) per fixed-size chunk of edges,
// so the edge set does not depend on the number of host threads.
// ---------------------------------------------------------------------------
static inline uint64_t splitmix64(uint64_t &x) {
  uint64_t z = (x += 0x9e3779b97f4a7c15ULL);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

struct Xoshiro256ss {
  uint64_t s[4];
  explicit Xoshiro256ss(uint64_t seed) {
    uint64_t x = seed;
    for (auto &w : s) w = splitmix64(x);
  }
  static inline uint64_t rotl(uint64_t v, int k) {
    return (v << k) | (v >> (64 - k));
  }
  inline uint64_t next() {
    const uint64_t result = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return result;
  }
};

static constexpr int64_t kEdgeChunk = 1 << 16;

// R-MAT edge e of chunk `chunk`: one 32-bit uniform per level picks the
// quadrant; level 0 sets the most significant bit.
template <typename F>
static void rmat_chunk(int log_n, int64_t chunk, int64_t m, uint64_t seed,
                       uint32_t ta, uint32_t tab, uint32_t tabc, F &&emit) {
  uint64_t sx = seed ^ (0xd1b54a32d192ed03ULL * (uint64_t)(chunk + 1));
  Xoshiro256ss rng(splitmix64(sx));
  const int64_t e0 = chunk * kEdgeChunk;
  const int64_t e1 = std::min(m, e0 + kEdgeChunk);
  for (int64_t e = e0; e < e1; ++e) {
    uint32_t u = 0, v = 0;
    uint64_t bits = 0;
    for (int lvl = 0; lvl < log_n; ++lvl) {
      if ((lvl & 1) == 0) bits = rng.next();
      const uint32_t r = (lvl & 1) ? (uint32_t)(bits >> 32) : (uint32_t)bits;
      uint32_t ub, vb;
      if (r < ta) { ub = 0; vb = 0; }
      else if (r < tab) { ub = 0; vb = 1; }
      else if (r < tabc) { ub = 1; vb = 0; }
      else { ub = 1; vb = 1; }
      u = (u << 1) | ub;
      v = (v << 1) | vb;
    }
    emit(u, v);
  }
}

// Sort + dedupe every row of a CSR given by counts/positions; returns the
// compacted graph structure.
static void finish_rows(int64_t n, std::vector<int64_t> &start,
                        std::vector<int32_t> &cols, Graph *g) {
  std::vector<int64_t> uniq(n, 0);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t i = 0; i < n; ++i) {
    auto b = cols.begin() + start[i], e = cols.begin() + start[i + 1];
    std::sort(b, e);
    uniq[i] = std::unique(b, e) - b;
  }
  g->n = n;
  g->row_ptr.assign(n + 1, 0);
  for (int64_t i = 0; i < n; ++i) g->row_ptr[i + 1] = g->row_ptr[i] + uniq[i];
  g->col_idx.resize(g->row_ptr[n]);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t i = 0; i < n; ++i)
    std::copy(cols.begin() + start[i], cols.begin() + start[i] + uniq[i],
              g->col_idx.begin() + g->row_ptr[i]);
}

void assign_values(Graph *g, bool gcn) {
  const int64_t n = g->n;
  g->val.resize(g->col_idx.size());
  if (!gcn) {
    std::fill(g->val.begin(), g->val.end(), 1.0f);
    return;
  }
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t i = 0; i < n; ++i) {
    const double di = (double)(g->row_ptr[i + 1] - g->row_ptr[i]);
    for (int64_t k = g->row_ptr[i]; k < g->row_ptr[i + 1]; ++k) {
      const int32_t j = g->col_idx[k];
      const double dj = (double)(g->row_ptr[j + 1] - g->row_ptr[j]);
      g->val[k] = (float)(1.0 / std::sqrt(di * dj));
    }
  }
}

// Cuthill-McKee order: components in order of their minimum-degree vertex,
// breadth-first over out-edges, neighbours visited by increasing (degree,
// id); vertices never reached (isolated ones) last.  Returns new->old, a
// permutation of [0, n) for any graph, symmetric or not.
std::vector<int32_t> cuthill_mckee(const Graph &g) {
  const int64_t n = g.n;
  std::vector<int32_t> deg(n);
  for (int64_t i = 0; i < n; ++i) deg[i] = (int32_t)(g.row_ptr[i + 1] - g.row_ptr[i]);
  std::vector<int32_t> by_deg(n);
  std::iota(by_deg.begin(), by_deg.end(), 0);
  std::stable_sort(by_deg.begin(), by_deg.end(),
                   [&](int32_t x, int32_t y) { return deg[x] < deg[y]; });
  std::vector<int32_t> order;
  order.reserve(n);
  std::vector<uint8_t> seen(n, 0);
  std::vector<int32_t> nb;
  for (int32_t s : by_deg) {
    if (deg[s] == 0 || seen[s]) continue;
    size_t head = order.size();
    order.push_back(s);
    seen[s] = 1;
    while (head < order.size()) {
      const int32_t v = order[head++];
      nb.clear();
      for (int64_t k = g.row_ptr[v]; k < g.row_ptr[v + 1]; ++k) {
        const int32_t w = g.col_idx[k];
        if (!seen[w]) {
          seen[w] = 1;
          nb.push_back(w);
        }
      }
      std::sort(nb.begin(), nb.end(), [&](int32_t x, int32_t y) {
        return deg[x] != deg[y] ? deg[x] < deg[y] : x < y;
      });
      order.insert(order.end(), nb.begin(), nb.end());
    }
  }
  // Vertices never reached: isolated ones, and (in a directed graph) those
  // without out-edges whose in-neighbours were not visited.
  for (int64_t i = 0; i < n; ++i)
    if (!seen[i]) order.push_back((int32_t)i);
  return order;
}

void relabel(const Graph &in, const std::vector<int32_t> &new2old, Graph *out) {
  const int64_t n = in.n;
  const bool has_val = !in.val.empty();
  std::vector<int32_t> old2new(n);
  for (int64_t i = 0; i < n; ++i) old2new[new2old[i]] = (int32_t)i;
  out->n = n;
  out->row_ptr.assign(n + 1, 0);
  for (int64_t i = 0; i < n; ++i) {
    const int32_t o = new2old[i];
    out->row_ptr[i + 1] = out->row_ptr[i] + (in.row_ptr[o + 1] - in.row_ptr[o]);
  }
  out->col_idx.resize(in.col_idx.size());
  out->val.resize(has_val ? in.val.size() : 0);
#pragma omp parallel
  {
    std::vector<std::pair<int32_t, float>> row;
#pragma omp for schedule(dynamic, 1024)
    for (int64_t i = 0; i < n; ++i) {
      const int32_t o = new2old[i];
      const int64_t w = out->row_ptr[i];
      if (!has_val) {
        for (int64_t k = in.row_ptr[o]; k < in.row_ptr[o + 1]; ++k)
          out->col_idx[w + k - in.row_ptr[o]] = old2new[in.col_idx[k]];
        std::sort(out->col_idx.begin() + w, out->col_idx.begin() + out->row_ptr[i + 1]);
        continue;
      }
      row.clear();
      for (int64_t k = in.row_ptr[o]; k < in.row_ptr[o + 1]; ++k)
        row.emplace_back(old2new[in.col_idx[k]], in.val[k]);
      std::sort(row.begin(), row.end(),
                [](const auto &x, const auto &y) { return x.first < y.first; });
      for (size_t k = 0; k < row.size(); ++k) {
        out->col_idx[w + k] = row[k].first;
        out->val[w + k] = row[k].second;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Partitioning (prep_torch_par.py:24-79).
// ---------------------------------------------------------------------------
static cola_status row_partition(const Graph &g, int nGPU, int rank,
                                 Partition *p) {
  if (nGPU < 1 || rank < 0 || rank >= nGPU)
    return fail(COLA_ERR_INVALID, "row_partition: rank out of range");
  const int64_t n = g.n;
  const int64_t P = (n + nGPU - 1) / nGPU;
  if (P > INT32_MAX) return fail(COLA_ERR_INVALID, "row_partition: nodesPerPE overflows int32");
  const int64_t lb = (int64_t)rank * P;
  const int64_t ub = std::min(n, lb + P);
  const int64_t M = std::max<int64_t>(0, ub - lb);
  const int64_t clo = (int64_t)rank * P, chi = (int64_t)(rank + 1) * P;
  p->num_nodes = n;
  p->nGPU = nGPU;
  p->rank = rank;
  p->nodesPerPE = (int32_t)P;
  p->num_rows = (int32_t)M;

  // Local CSR with empty rows collapsed (np.unique of indptr / rows).
  std::vector<int64_t> lcnt(M, 0), rcnt_row(M, 0);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t i = 0; i < M; ++i) {
    int64_t c = 0;
    for (int64_t k = g.row_ptr[lb + i]; k < g.row_ptr[lb + i + 1]; ++k) {
      const int64_t col = g.col_idx[k];
      c += (col >= clo && col < chi);
    }
    lcnt[i] = c;
    rcnt_row[i] = (g.row_ptr[lb + i + 1] - g.row_ptr[lb + i]) - c;
  }
  int64_t nnz_l = 0, L = 0;
  for (int64_t i = 0; i < M; ++i) {
    nnz_l += lcnt[i];
    L += lcnt[i] > 0;
  }
  if (nnz_l > INT32_MAX) return fail(COLA_ERR_INVALID, "row_partition: local nnz overflows int32");
  p->csr_rowPtr_l.assign(1, 0);
  p->csr_rowPtr_l.reserve(L + 1);
  p->csr_rowIdx_l.clear();
  p->csr_rowIdx_l.reserve(L);
  std::vector<int64_t> lpos(M, 0);
  {
    int64_t acc = 0;
    for (int64_t i = 0; i < M; ++i) {
      lpos[i] = acc;
      if (lcnt[i] > 0) {
        acc += lcnt[i];
        p->csr_rowIdx_l.push_back((int32_t)i);
        p->csr_rowPtr_l.push_back((int32_t)acc);
      }
    }
  }
  p->csr_colIdx_l.resize(nnz_l);
  p->csr_val_l.resize(nnz_l);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t i = 0; i < M; ++i) {
    int64_t w = lpos[i];
    for (int64_t k = g.row_ptr[lb + i]; k < g.row_ptr[lb + i + 1]; ++k) {
      const int64_t col = g.col_idx[k];
      if (col >= clo && col < chi) {
        p->csr_colIdx_l[w] = (int32_t)(col % P);
        p->csr_val_l[w] = g.val[k];
        ++w;
      }
    }
  }

  // Remote CSC over the (M x n) panel, columns compressed to the unique
  // remote ids; rows ascending within each column (scipy canonical order).
  std::vector<int32_t> ccnt(n, 0);
  for (int64_t i = 0; i < M; ++i)
    for (int64_t k = g.row_ptr[lb + i]; k < g.row_ptr[lb + i + 1]; ++k) {
      const int64_t col = g.col_idx[k];
      if (!(col >= clo && col < chi)) ccnt[col]++;
    }
  int64_t nnz_r = 0, U = 0;
  for (int64_t c = 0; c < n; ++c) {
    nnz_r += ccnt[c];
    U += ccnt[c] > 0;
  }
  if (nnz_r > INT32_MAX) return fail(COLA_ERR_INVALID, "row_partition: remote nnz overflows int32");
  p->csc_colIdx_r.clear();
  p->csc_colIdx_r.reserve(U);
  p->csc_colPtr_r.assign(1, 0);
  p->csc_colPtr_r.reserve(U + 1);
  std::vector<int64_t> cpos(n, 0);
  {
    int64_t acc = 0;
    for (int64_t c = 0; c < n; ++c) {
      cpos[c] = acc;
      if (ccnt[c] > 0) {
        acc += ccnt[c];
        p->csc_colIdx_r.push_back((int32_t)c);
        p->csc_colPtr_r.push_back((int32_t)acc);
      }
    }
  }
  p->csc_rowIdx_r.resize(nnz_r);
  p->csc_val_r.resize(nnz_r);
  for (int64_t i = 0; i < M; ++i)
    for (int64_t k = g.row_ptr[lb + i]; k < g.row_ptr[lb + i + 1]; ++k) {
      const int64_t col = g.col_idx[k];
      if (!(col >= clo && col < chi)) {
        const int64_t w = cpos[col]++;
        p->csc_rowIdx_r[w] = (int32_t)i;
        p->csc_val_r[w] = g.val[k];
      }
    }
  return COLA_OK;
}

cola_status validate_partition(const Partition &p) {
  const int64_t P = p.nodesPerPE, M = p.num_rows, n = p.num_nodes;
  auto bad = [](const char *what) {
    return fail(COLA_ERR_INVALID, std::string("partition: ") + what);
  };
  if (p.nGPU < 1 || p.rank < 0 || p.rank >= p.nGPU) return bad("rank out of range");
  if (P != (n + p.nGPU - 1) / p.nGPU) return bad("nodesPerPE != ceil(n/nGPU)");
  if (M != std::max<int64_t>(0, std::min(n, (int64_t)(p.rank + 1) * P) - (int64_t)p.rank * P))
    return bad("row count does not match rank");
  const auto &rp = p.csr_rowPtr_l, &ri = p.csr_rowIdx_l;
  if (rp.empty() || rp[0] != 0) return bad("csr_rowPtr_l must start at 0");
  if (ri.size() + 1 != rp.size()) return bad("len(csr_rowIdx_l) != len(csr_rowPtr_l)-1");
  for (size_t i = 1; i < rp.size(); ++i)
    if (rp[i] <= rp[i - 1]) return bad("csr_rowPtr_l not strictly increasing");
  if ((size_t)rp.back() != p.csr_colIdx_l.size()) return bad("csr_rowPtr_l end != nnz_l");
  if (p.csr_val_l.size() != p.csr_colIdx_l.size()) return bad("csr_val_l size");
  for (size_t i = 0; i < ri.size(); ++i) {
    if (ri[i] < 0 || ri[i] >= M) return bad("csr_rowIdx_l out of [0,M)");
    if (i && ri[i] <= ri[i - 1]) return bad("csr_rowIdx_l not strictly increasing");
  }
  // Local column c is graph node rank*P + c: below P, and below n on the
  // last rank, whose slice of B has only M rows.
  const int64_t lcols = std::min<int64_t>(P, n - (int64_t)p.rank * P);
  for (int32_t c : p.csr_colIdx_l)
    if (c < 0 || c >= lcols) return bad("csr_colIdx_l out of [0, min(P, n - rank*P))");
  const auto &cp = p.csc_colPtr_r, &ci = p.csc_colIdx_r;
  if (cp.empty() || cp[0] != 0) return bad("csc_colPtr_r must start at 0");
  if (ci.size() + 1 != cp.size()) return bad("len(csc_colIdx_r) != len(csc_colPtr_r)-1");
  for (size_t i = 1; i < cp.size(); ++i)
    if (cp[i] <= cp[i - 1]) return bad("csc_colPtr_r not strictly increasing");
  if ((size_t)cp.back() != p.csc_rowIdx_r.size()) return bad("csc_colPtr_r end != nnz_r");
  if (p.csc_val_r.size() != p.csc_rowIdx_r.size()) return bad("csc_val_r size");
  const int64_t clo = (int64_t)p.rank * P, chi = (int64_t)(p.rank + 1) * P;
  for (size_t i = 0; i < ci.size(); ++i) {
    if (ci[i] < 0 || ci[i] >= n) return bad("csc_colIdx_r out of [0,n)");
    if (ci[i] >= clo && ci[i] < chi) return bad("csc_colIdx_r inside the owned column range");
    if (i && ci[i] <= ci[i - 1]) return bad("csc_colIdx_r not strictly increasing");
  }
  for (int32_t r : p.csc_rowIdx_r)
    if (r < 0 || r >= M) return bad("csc_rowIdx_r out of [0,M)");
  return COLA_OK;
}

// ---------------------------------------------------------------------------
// Files (raw little-endian, no header).
// ---------------------------------------------------------------------------
template <typename T>
static cola_status bin2vec(const std::string &path, std::vector<T> *out,
                           bool *missing = nullptr) {
  std::ifstream f(path, std::ios::binary);
  if (!f.is_open()) {
    if (missing) {
      *missing = true;
      return COLA_OK;
    }
    return fail(COLA_ERR_IO, "cannot open " + path);
  }
  if (missing) *missing = false;
  f.seekg(0, std::ios::end);
  const int64_t size = (int64_t)f.tellg();
  f.seekg(0, std::ios::beg);
  if (size < 0) return fail(COLA_ERR_IO, "cannot size " + path);
  if (size % (int64_t)sizeof(T) != 0)
    return fail(COLA_ERR_IO, path + ": size " + std::to_string(size) +
                                 " is not a multiple of " + std::to_string(sizeof(T)));
  out->resize(size / sizeof(T));
  if (size > 0 && !f.read(reinterpret_cast<char *>(out->data()), size))
    return fail(COLA_ERR_IO, "short read on " + path);
  return COLA_OK;
}

template <typename T>
static cola_status vec2bin(const std::string &path, const T *data, size_t count) {
  FILE *f = std::fopen(path.c_str(), "wb");
  if (!f) return fail(COLA_ERR_IO, "cannot create " + path + ": " + std::strerror(errno));
  const size_t w = count ? std::fwrite(data, sizeof(T), count, f) : 0;
  const int rc = std::fclose(f);
  if (w != count || rc != 0) return fail(COLA_ERR_IO, "short write on " + path);
  return COLA_OK;
}

static cola_status mkdirs(const std::string &path) {
  std::string cur;
  size_t i = 0;
  while (i <= path.size()) {
    size_t j = path.find('/', i);
    if (j == std::string::npos) j = path.size();
    cur = path.substr(0, j);
    if (!cur.empty()) {
      if (mkdir(cur.c_str(), 0755) != 0 && errno != EEXIST)
        return fail(COLA_ERR_IO, "cannot create directory " + cur);
    }
    i = j + 1;
  }
  return COLA_OK;
}

static std::string part_name(const std::string &dir, const char *stem, int nGPU,
                             int rank) {
  return dir + "/" + stem + "_" + std::to_string(nGPU) + "_" + std::to_string(rank) + ".bin";
}

// ---------------------------------------------------------------------------
// Plan builder.
// ---------------------------------------------------------------------------
cola_status build_plan(const Partition &p, const cola_plan_opts &o,
                       HostPlan *out) {
  if (o.hot_cols > 0 && o.col_panels > 1)
    return fail(COLA_ERR_INVALID, "plan: hot_cols and col_panels are exclusive");
  const int32_t hub = o.hub_threshold > 0 ? o.hub_threshold : 1024;
  const int32_t chunk = o.chunk_nnz > 0 ? o.chunk_nnz : 512;
  const int32_t kdeg = o.kahan_degree > 0 ? o.kahan_degree : 1024;
  const int64_t M = p.num_rows;
  out->num_rows = (int32_t)M;
  out->nodesPerPE = p.nodesPerPE;
  out->U_r = (int32_t)p.csc_colIdx_r.size();
  out->max_dim = o.max_dim > 0 ? o.max_dim : 256;
  out->remote_cols = p.csc_colIdx_r;

  // Per-row entry counts: local (zip rows) + remote (transposed CSC).
  std::vector<int64_t> cnt(M + 1, 0);
  for (size_t z = 0; z < p.csr_rowIdx_l.size(); ++z)
    cnt[p.csr_rowIdx_l[z] + 1] += p.csr_rowPtr_l[z + 1] - p.csr_rowPtr_l[z];
  for (int32_t r : p.csc_rowIdx_r) cnt[r + 1]++;
  for (int64_t i = 0; i < M; ++i) cnt[i + 1] += cnt[i];
  const int64_t nnz = cnt[M];
  if (nnz > INT32_MAX) return fail(COLA_ERR_INVALID, "plan: partition nnz overflows int32");
  out->nnz = nnz;
  out->cols.resize(nnz);
  out->vals.resize(nnz);
  std::vector<int64_t> cur(cnt.begin(), cnt.end() - 1);
  for (size_t z = 0; z < p.csr_rowIdx_l.size(); ++z) {
    const int32_t r = p.csr_rowIdx_l[z];
    for (int32_t k = p.csr_rowPtr_l[z]; k < p.csr_rowPtr_l[z + 1]; ++k) {
      out->cols[cur[r]] = p.csr_colIdx_l[k];
      out->vals[cur[r]++] = p.csr_val_l[k];
    }
  }
  for (size_t u = 0; u + 1 < p.csc_colPtr_r.size(); ++u)
    for (int32_t k = p.csc_colPtr_r[u]; k < p.csc_colPtr_r[u + 1]; ++k) {
      const int32_t r = p.csc_rowIdx_r[k];
      out->cols[cur[r]] = ~(int32_t)u;
      out->vals[cur[r]++] = p.csc_val_r[k];
    }

  // Local entries of each row come first: lcnt[i] of them.
  std::vector<int64_t> lcnt(M, 0);
  for (size_t z = 0; z < p.csr_rowIdx_l.size(); ++z)
    lcnt[p.csr_rowIdx_l[z]] += p.csr_rowPtr_l[z + 1] - p.csr_rowPtr_l[z];

  // Distinct local columns (for the algorithmic byte count).
  {
    std::vector<uint8_t> used(p.nodesPerPE, 0);
    for (int32_t c : p.csr_colIdx_l) used[c] = 1;
    int64_t ul = 0;
    for (uint8_t b : used) ul += b;
    out->U_l = (int32_t)ul;
  }

  // Hot / cold split (hot_cols = H > 0): the H most referenced B rows (key c
  // for local column c, P + u for packed row u; ties by key) are hot.  Each
  // row's entries are reordered stably to hot-local, hot-remote, cold-local,
  // cold-remote; hcnt / hloc count the hot and hot-local entries per row.
  const int64_t H = o.hot_cols > 0 ? o.hot_cols : 0;
  std::vector<int64_t> hcnt, hloc;
  out->hot_nnz = nnz;
  if (H > 0) {
    const int64_t Pn = p.nodesPerPE, Un = out->U_r;
    std::vector<int64_t> refs(Pn + Un, 0);
    for (int32_t c : out->cols) ++refs[c >= 0 ? (int64_t)c : Pn + (int64_t)~c];
    std::vector<int32_t> keys(Pn + Un);
    std::iota(keys.begin(), keys.end(), 0);
    const int64_t h = std::min<int64_t>(H, Pn + Un);
    std::nth_element(keys.begin(), keys.begin() + (h - 1), keys.end(), [&](int32_t x, int32_t y) {
      return refs[x] != refs[y] ? refs[x] > refs[y] : x < y;
    });
    std::vector<uint8_t> hot(Pn + Un, 0);
    for (int64_t k = 0; k < h; ++k) hot[keys[k]] = 1;
    hcnt.assign(M, 0);
    hloc.assign(M, 0);
#pragma omp parallel
    {
      std::vector<std::pair<int32_t, float>> row, cold;
#pragma omp for schedule(dynamic, 1024)
      for (int64_t i = 0; i < M; ++i) {
        row.clear();
        cold.clear();
        for (int64_t k = cnt[i]; k < cnt[i + 1]; ++k) {
          const int32_t c = out->cols[k];
          const bool is_hot = hot[c >= 0 ? (int64_t)c : Pn + (int64_t)~c] != 0;
          (is_hot ? row : cold).emplace_back(c, out->vals[k]);
          if (is_hot && c >= 0) ++hloc[i];
        }
        hcnt[i] = (int64_t)row.size();
        row.insert(row.end(), cold.begin(), cold.end());
        for (size_t k = 0; k < row.size(); ++k) {
          out->cols[cnt[i] + k] = row[k].first;
          out->vals[cnt[i] + k] = row[k].second;
        }
      }
    }
    out->hot_nnz = std::accumulate(hcnt.begin(), hcnt.end(), (int64_t)0);
  }

  // Work items in row order; hub rows become chunk items.  flags: bit 0 =
  // compensated sums allowed, bits 1.. = number of local entries in the item
  // (the local / remote boundary of the item's range).
  const int32_t sdeg = o.short_degree > 0 ? o.short_degree : 32;
  out->items_short.clear();
  out->items_long.clear();
  out->items_short_r.clear();
  out->items_long_r.clear();
  out->items_short2.clear();
  out->items_long2.clear();
  out->empty_rows.clear();
  out->hub_rows.clear();
  out->hub_chunk_ptr.assign(1, 0);
  int32_t nchunks = 0;
  std::vector<WorkItem> unused;
  auto add = [&](std::vector<WorkItem> &all, std::vector<WorkItem> &rem, int32_t outr, int64_t b,
                 int64_t e, int64_t mid, int32_t kflag) {
    const int64_t nloc = std::max<int64_t>(0, std::min(e, mid) - b);
    const WorkItem w{outr, (int32_t)b, (int32_t)e, (int32_t)((nloc << 1) | kflag)};
    all.push_back(w);
    if (mid < e) rem.push_back(w);
  };
  // Item order: row order, or rows by decreasing entry count (rows that
  // share hub columns run together).
  std::vector<int32_t> order(M);
  std::iota(order.begin(), order.end(), 0);
  if (o.row_order == 1)
    std::stable_sort(order.begin(), order.end(),
                     [&](int32_t x, int32_t y) { return cnt[x + 1] - cnt[x] > cnt[y + 1] - cnt[y]; });
  for (int64_t oi = 0; oi < M; ++oi) {
    const int64_t i = order[oi];
    const int64_t b = cnt[i], e = cnt[i + 1], d = e - b, mid = b + lcnt[i];
    if (d == 0) {
      out->empty_rows.push_back((int32_t)i);
      continue;
    }
    const int32_t kflag = d > kdeg ? 1 : 0;
    if (H > 0) {
      // Hot part [b, hb) in the first pass, cold part [hb, e) in the second;
      // each part keeps its local entries first (hmid / cmid bound them).
      // The first pass also gets its remote-entry items, so the
      // local-then-remote phases can run it; the second pass runs whole.
      const int64_t hb = b + hcnt[i], hmid = b + hloc[i], cmid = hb + (lcnt[i] - hloc[i]);
      if (d > sdeg && d > hub) {
        for (int64_t s = b; s < hb; s += chunk) {
          add(out->items_long, out->items_long_r, -(nchunks + 1), s, std::min<int64_t>(hb, s + chunk), hmid, kflag);
          ++nchunks;
        }
        for (int64_t s = hb; s < e; s += chunk) {
          add(out->items_long2, unused, -(nchunks + 1), s, std::min<int64_t>(e, s + chunk), cmid, kflag);
          ++nchunks;
        }
        out->hub_rows.push_back((int32_t)i);
        out->hub_chunk_ptr.push_back(nchunks);
      } else {
        // Every row gets a first-pass item (possibly empty: it writes the
        // row so the second pass can add onto it).
        if (hb - b <= sdeg) add(out->items_short, out->items_short_r, (int32_t)i, b, hb, hmid, kflag);
        else add(out->items_long, out->items_long_r, (int32_t)i, b, hb, hmid, kflag);
        if (e > hb) add(e - hb <= sdeg ? out->items_short2 : out->items_long2, unused, (int32_t)i, hb, e, cmid, kflag);
      }
      continue;
    }
    if (d <= sdeg) {
      add(out->items_short, out->items_short_r, (int32_t)i, b, e, mid, kflag);
    } else if (d <= hub) {
      add(out->items_long, out->items_long_r, (int32_t)i, b, e, mid, kflag);
    } else {
      for (int64_t s = b; s < e; s += chunk) {
        add(out->items_long, out->items_long_r, -(nchunks + 1), s, std::min<int64_t>(e, s + chunk), mid,
            kflag);
        ++nchunks;
      }
      out->hub_rows.push_back((int32_t)i);
      out->hub_chunk_ptr.push_back(nchunks);
    }
  }
  // Contiguous item ranges per XCD, balanced by entries + a per-item cost.
  auto split = [](const std::vector<WorkItem> &items, std::vector<int32_t> *xb) {
    const int kXcd = 8;
    const int64_t kItemCost = 8;
    int64_t total = 0;
    for (const auto &it : items) total += (it.end - it.begin) + kItemCost;
    xb->assign(kXcd + 1, (int32_t)items.size());
    (*xb)[0] = 0;
    int64_t acc = 0;
    int x = 1;
    for (size_t k = 0; k < items.size() && x < kXcd; ++k) {
      while (x < kXcd && acc >= (total * x) / kXcd) (*xb)[x++] = (int32_t)k;
      acc += (items[k].end - items[k].begin) + kItemCost;
    }
  };
  split(out->items_short, &out->xcd_short);
  split(out->items_long, &out->xcd_long);
  split(out->items_short_r, &out->xcd_short_r);
  split(out->items_long_r, &out->xcd_long_r);
  split(out->items_short2, &out->xcd_short2);
  split(out->items_long2, &out->xcd_long2);

  // Column panels.
  const int32_t K = o.col_panels > 1 ? o.col_panels : 0;
  out->panels = K;
  out->panel_items.clear();
  out->panel_begin.assign(1, 0);
  out->panel_xcd.clear();
  out->panel_hub_ptr.assign(1, 0);
  out->panel_chunks = 0;
  if (K > 1) {
    const int64_t P = std::max<int32_t>(1, p.nodesPerPE), U = std::max<int32_t>(1, out->U_r);
    auto panel_of = [&](int32_t c) {
      return c >= 0 ? (int)((int64_t)c * K / P) : K + (int)((int64_t)(~c) * K / U);
    };
    // Each row becomes one item per panel it touches, so its entries must be
    // grouped by panel: two items of one row in the same panel launch would
    // both read-add-write that row of C at once.  Rows are stably sorted by
    // panel (local panels come before remote ones, so local entries stay
    // first; the order within a panel is kept).  Rows from the reference's
    // writer are already sorted by column and are left as they are.
#pragma omp parallel
    {
      std::vector<std::pair<int32_t, float>> row;
#pragma omp for schedule(dynamic, 1024)
      for (int64_t i = 0; i < M; ++i) {
        bool sorted = true;
        for (int64_t k = cnt[i] + 1; k < cnt[i + 1] && sorted; ++k)
          sorted = panel_of(out->cols[k - 1]) <= panel_of(out->cols[k]);
        if (sorted) continue;
        row.clear();
        for (int64_t k = cnt[i]; k < cnt[i + 1]; ++k) row.emplace_back(out->cols[k], out->vals[k]);
        std::stable_sort(row.begin(), row.end(), [&](const std::pair<int32_t, float> &x,
                                                     const std::pair<int32_t, float> &y) {
          return panel_of(x.first) < panel_of(y.first);
        });
        for (size_t k = 0; k < row.size(); ++k) {
          out->cols[cnt[i] + k] = row[k].first;
          out->vals[cnt[i] + k] = row[k].second;
        }
      }
    }
    std::vector<std::vector<WorkItem>> per(2 * K);
    int32_t pch = 0;
    // Same rows, order and hub test as the items above, so panel_hub_ptr
    // lines up with hub_rows (the hub reduction pairs them).
    for (int64_t oi = 0; oi < M; ++oi) {
      const int64_t i = order[oi];
      const int64_t b = cnt[i], e = cnt[i + 1], d = e - b;
      if (d == 0) continue;
      const bool is_hub = d > sdeg && d > hub;
      const int32_t kflag = d > kdeg ? 1 : 0;
      int64_t s = b;
      while (s < e) {
        const int key = panel_of(out->cols[s]);
        int64_t t = s + 1;
        while (t < e && panel_of(out->cols[t]) == key) ++t;
        if (!is_hub) {
          per[key].push_back({(int32_t)i, (int32_t)s, (int32_t)t, kflag});
        } else {
          for (int64_t q = s; q < t; q += chunk)
            per[key].push_back({-(pch++ + 1), (int32_t)q, (int32_t)std::min<int64_t>(t, q + chunk), kflag});
        }
        s = t;
      }
      if (is_hub) out->panel_hub_ptr.push_back(pch);
    }
    out->panel_chunks = pch;
    for (int k = 0; k < 2 * K; ++k) {
      std::vector<int32_t> xb;
      split(per[k], &xb);
      out->panel_xcd.insert(out->panel_xcd.end(), xb.begin(), xb.end());
      out->panel_items.insert(out->panel_items.end(), per[k].begin(), per[k].end());
      out->panel_begin.push_back((int64_t)out->panel_items.size());
    }
    if (out->panel_items.size() > (size_t)INT32_MAX) return fail(COLA_ERR_INVALID, "plan: too many panel items");
  }

  // Tiny-row bundles: short rows of L <= 4 entries (not flagged for
  // compensated sums) in item order, grouped by L, floor(8/L) rows a bundle.
  // A row of one entry would otherwise occupy a lane group for a whole batch
  // of 8 row loads, 7 of them padding.  The entries of a bundle are copied
  // back to back after the partition's own; the copies keep each row's entry
  // order, so a row's sum is formed in the same order as in the short bin.
  out->items_tiny.clear();
  out->items_short_nt.clear();
  out->tiny_rows.clear();
  if (H == 0) {
    constexpr int kTinyMax = 4, kBatch = 8;
    std::vector<WorkItem> byL[kTinyMax + 1];
    for (const WorkItem &w : out->items_short) {
      const int64_t d = (int64_t)w.end - w.begin;
      if (d >= 1 && d <= kTinyMax && (w.flags & 1) == 0) byL[d].push_back(w);
      else out->items_short_nt.push_back(w);
    }
    for (int L = kTinyMax; L >= 1; --L) {
      const int R = kBatch / L;
      for (size_t i = 0; i < byL[L].size(); i += R) {
        const size_t n = std::min<size_t>(R, byL[L].size() - i);
        const int64_t b = (int64_t)out->cols.size();
        if (b + (int64_t)n * L > INT32_MAX || out->tiny_rows.size() + n > (size_t)INT32_MAX)
          return fail(COLA_ERR_INVALID, "plan: tiny-row bundles overflow int32 offsets");
        const int32_t roff = (int32_t)out->tiny_rows.size();
        for (size_t k = 0; k < n; ++k) {
          const WorkItem &w = byL[L][i + k];
          out->tiny_rows.push_back(w.out);
          for (int32_t q = w.begin; q < w.end; ++q) {
            const int32_t c = out->cols[q];
            const float v = out->vals[q];
            out->cols.push_back(c);
            out->vals.push_back(v);
          }
        }
        out->items_tiny.push_back({roff, (int32_t)b, (int32_t)(b + (int64_t)n * L), L});
      }
    }
    if (out->items_tiny.empty()) out->items_short_nt.clear();
  }
  split(out->items_tiny, &out->xcd_tiny);
  split(out->items_short_nt, &out->xcd_short_nt);
  return COLA_OK;
}

}  // namespace cola

using namespace cola;

// ---------------------------------------------------------------------------
// C ABI (host part).
// ---------------------------------------------------------------------------
extern "C" {

const char *cola_last_error(void) { return g_last_error.c_str(); }
const char *cola_version(void) { return "cola-mi355x 0.1"; }

cola_status cola_rmat_graph(int log_n, double avg_degree, double a, double b,
                            double c, uint64_t seed, int reorder,
                            int gcn_values, cola_graph *out) {
  if (!out) return fail(COLA_ERR_INVALID, "rmat: out is NULL");
  *out = nullptr;
  if (log_n < 1 || log_n > 30) return fail(COLA_ERR_INVALID, "rmat: log_n must be in [1,30]");
  if (!(avg_degree > 0) || a < 0 || b < 0 || c < 0 || a + b + c > 1.0)
    return fail(COLA_ERR_INVALID, "rmat: bad degree or quadrant probabilities");
  if (reorder < 0 || reorder > 2) return fail(COLA_ERR_INVALID, "rmat: reorder must be 0, 1 or 2");
  const int64_t n = (int64_t)1 << log_n;
  const int64_t m = (int64_t)std::llround((double)n * avg_degree / 2.0);
  const double s32 = 4294967296.0;
  auto th = [&](double p) -> uint32_t {
    const double v = p * s32;
    return v >= s32 ? 0xffffffffu : (uint32_t)v;
  };
  const uint32_t ta = th(a), tab = th(a + b), tabc = th(a + b + c);
  const int64_t nchunks = (m + kEdgeChunk - 1) / kEdgeChunk;

  // Pass 1: degree counts of both directions (self-loops dropped).
  std::vector<std::atomic<int64_t>> deg(n);
  for (auto &d : deg) d.store(0, std::memory_order_relaxed);
#pragma omp parallel for schedule(dynamic, 4)
  for (int64_t ch = 0; ch < nchunks; ++ch)
    rmat_chunk(log_n, ch, m, seed, ta, tab, tabc, [&](uint32_t u, uint32_t v) {
      if (u == v) return;
      deg[u].fetch_add(1, std::memory_order_relaxed);
      deg[v].fetch_add(1, std::memory_order_relaxed);
    });
  std::vector<int64_t> start(n + 1, 0);
  for (int64_t i = 0; i < n; ++i) start[i + 1] = start[i] + deg[i].load(std::memory_order_relaxed);
  std::vector<int32_t> cols(start[n]);
  for (int64_t i = 0; i < n; ++i) deg[i].store(start[i], std::memory_order_relaxed);
  // Pass 2: regenerate the same edges and place them.
#pragma omp parallel for schedule(dynamic, 4)
  for (int64_t ch = 0; ch < nchunks; ++ch)
    rmat_chunk(log_n, ch, m, seed, ta, tab, tabc, [&](uint32_t u, uint32_t v) {
      if (u == v) return;
      cols[deg[u].fetch_add(1, std::memory_order_relaxed)] = (int32_t)v;
      cols[deg[v].fetch_add(1, std::memory_order_relaxed)] = (int32_t)u;
    });
  auto *h = new (std::nothrow) cola_graph_s;
  if (!h) return fail(COLA_ERR_NOMEM, "rmat: out of host memory");
  finish_rows(n, start, cols, &h->g);
  std::vector<int32_t>().swap(cols);
  if (h->g.col_idx.size() > (size_t)INT32_MAX) {
    delete h;
    return fail(COLA_ERR_INVALID, "rmat: nnz overflows int32");
  }
  if (reorder != 0) {
    Graph re;
    relabel(h->g, reorder == 1 ? cuthill_mckee(h->g) : community_order(h->g), &re);
    h->g = std::move(re);
  }
  assign_values(&h->g, gcn_values != 0);
  *out = h;
  return COLA_OK;
}

cola_status cola_graph_from_csr(int64_t n, const int64_t *row_ptr,
                                const int32_t *col_idx, const float *val,
                                cola_graph *out) {
  if (!out) return fail(COLA_ERR_INVALID, "graph_from_csr: out is NULL");
  *out = nullptr;
  if (n < 0 || n > INT32_MAX || !row_ptr) return fail(COLA_ERR_INVALID, "graph_from_csr: bad n or row_ptr");
  if (row_ptr[0] != 0) return fail(COLA_ERR_INVALID, "graph_from_csr: row_ptr[0] != 0");
  for (int64_t i = 0; i < n; ++i)
    if (row_ptr[i + 1] < row_ptr[i]) return fail(COLA_ERR_INVALID, "graph_from_csr: row_ptr decreasing");
  const int64_t nnz = row_ptr[n];
  if (nnz > 0 && !col_idx) return fail(COLA_ERR_INVALID, "graph_from_csr: col_idx is NULL");
  for (int64_t k = 0; k < nnz; ++k)
    if (col_idx[k] < 0 || col_idx[k] >= n) return fail(COLA_ERR_INVALID, "graph_from_csr: column out of range");
  auto *h = new (std::nothrow) cola_graph_s;
  if (!h) return fail(COLA_ERR_NOMEM, "graph_from_csr: out of host memory");
  Graph &g = h->g;
  g.n = n;
  g.row_ptr.assign(n + 1, 0);
  std::vector<std::pair<int32_t, float>> row;
  std::vector<int32_t> cols;
  std::vector<float> vals;
  cols.reserve(nnz);
  vals.reserve(nnz);
  for (int64_t i = 0; i < n; ++i) {
    row.clear();
    for (int64_t k = row_ptr[i]; k < row_ptr[i + 1]; ++k)
      row.emplace_back(col_idx[k], val ? val[k] : 1.0f);
    std::stable_sort(row.begin(), row.end(),
                     [](const auto &x, const auto &y) { return x.first < y.first; });
    for (size_t k = 0; k < row.size(); ++k) {
      if (k && row[k].first == row[k - 1].first)
        vals.back() += row[k].second;  // duplicates summed (scipy tocsr)
      else {
        cols.push_back(row[k].first);
        vals.push_back(row[k].second);
      }
    }
    g.row_ptr[i + 1] = (int64_t)cols.size();
  }
  g.col_idx = std::move(cols);
  g.val = std::move(vals);
  *out = h;
  return COLA_OK;
}

cola_status cola_graph_info(cola_graph g, int64_t *n, int64_t *nnz,
                            int64_t *max_degree) {
  if (!g) return fail(COLA_ERR_INVALID, "graph_info: NULL graph");
  if (n) *n = g->g.n;
  if (nnz) *nnz = (int64_t)g->g.col_idx.size();
  if (max_degree) {
    int64_t md = 0;
    for (int64_t i = 0; i < g->g.n; ++i)
      md = std::max(md, g->g.row_ptr[i + 1] - g->g.row_ptr[i]);
    *max_degree = md;
  }
  return COLA_OK;
}

cola_status cola_graph_arrays(cola_graph g, const int64_t **row_ptr,
                              const int32_t **col_idx, const float **val) {
  if (!g) return fail(COLA_ERR_INVALID, "graph_arrays: NULL graph");
  if (row_ptr) *row_ptr = g->g.row_ptr.data();
  if (col_idx) *col_idx = g->g.col_idx.data();
  if (val) *val = g->g.val.data();
  return COLA_OK;
}

cola_status cola_graph_destroy(cola_graph g) {
  delete g;
  return COLA_OK;
}

cola_status cola_row_partition(cola_graph g, int nGPU, int rank,
                               cola_partition *out) {
  if (!g || !out) return fail(COLA_ERR_INVALID, "row_partition: NULL argument");
  *out = nullptr;
  auto *h = new (std::nothrow) cola_partition_s;
  if (!h) return fail(COLA_ERR_NOMEM, "row_partition: out of host memory");
  cola_status st = row_partition(g->g, nGPU, rank, &h->p);
  if (st != COLA_OK) {
    delete h;
    return st;
  }
  *out = h;
  return COLA_OK;
}

cola_status cola_partition_from_arrays(int64_t num_nodes, int nGPU, int rank,
                                       const cola_partition_arrays *a,
                                       int64_t L, int64_t nnz_l, int64_t U_r,
                                       int64_t nnz_r, cola_partition *out) {
  if (!a || !out) return fail(COLA_ERR_INVALID, "partition_from_arrays: NULL argument");
  *out = nullptr;
  if (num_nodes < 0 || num_nodes > INT32_MAX || nGPU < 1 || L < 0 || nnz_l < 0 ||
      U_r < 0 || nnz_r < 0 || nnz_l > INT32_MAX || nnz_r > INT32_MAX)
    return fail(COLA_ERR_INVALID, "partition_from_arrays: bad sizes");
  if (!a->csr_rowPtr_l || !a->csc_colPtr_r || (L && !a->csr_rowIdx_l) ||
      (nnz_l && !a->csr_colIdx_l) || (U_r && !a->csc_colIdx_r) ||
      (nnz_r && !a->csc_rowIdx_r))
    return fail(COLA_ERR_INVALID, "partition_from_arrays: NULL array");
  auto *h = new (std::nothrow) cola_partition_s;
  if (!h) return fail(COLA_ERR_NOMEM, "partition_from_arrays: out of host memory");
  Partition &p = h->p;
  p.num_nodes = num_nodes;
  p.nGPU = nGPU;
  p.rank = rank;
  const int64_t P = (num_nodes + nGPU - 1) / nGPU;
  p.nodesPerPE = (int32_t)P;
  p.num_rows = (int32_t)std::max<int64_t>(0, std::min(num_nodes, (int64_t)(rank + 1) * P) - (int64_t)rank * P);
  p.csr_rowPtr_l.assign(a->csr_rowPtr_l, a->csr_rowPtr_l + L + 1);
  p.csr_colIdx_l.assign(a->csr_colIdx_l, a->csr_colIdx_l + nnz_l);
  p.csr_rowIdx_l.assign(a->csr_rowIdx_l, a->csr_rowIdx_l + L);
  if (a->csr_val_l) p.csr_val_l.assign(a->csr_val_l, a->csr_val_l + nnz_l);
  else p.csr_val_l.assign(nnz_l, 1.0f);
  p.csc_colPtr_r.assign(a->csc_colPtr_r, a->csc_colPtr_r + U_r + 1);
  p.csc_rowIdx_r.assign(a->csc_rowIdx_r, a->csc_rowIdx_r + nnz_r);
  p.csc_colIdx_r.assign(a->csc_colIdx_r, a->csc_colIdx_r + U_r);
  if (a->csc_val_r) p.csc_val_r.assign(a->csc_val_r, a->csc_val_r + nnz_r);
  else p.csc_val_r.assign(nnz_r, 1.0f);
  cola_status st = validate_partition(p);
  if (st != COLA_OK) {
    delete h;
    return st;
  }
  *out = h;
  return COLA_OK;
}

cola_status cola_partition_get_info(cola_partition h, cola_partition_info *info) {
  if (!h || !info) return fail(COLA_ERR_INVALID, "partition_info: NULL argument");
  const Partition &p = h->p;
  info->num_nodes = p.num_nodes;
  info->nGPU = p.nGPU;
  info->rank = p.rank;
  info->nodesPerPE = p.nodesPerPE;
  info->num_rows = p.num_rows;
  info->L = (int32_t)p.csr_rowIdx_l.size();
  info->U_r = (int32_t)p.csc_colIdx_r.size();
  info->nnz_l = (int64_t)p.csr_colIdx_l.size();
  info->nnz_r = (int64_t)p.csc_rowIdx_r.size();
  std::vector<uint8_t> used(p.nodesPerPE, 0);
  for (int32_t c : p.csr_colIdx_l) used[c] = 1;
  int64_t ul = 0;
  for (uint8_t b : used) ul += b;
  info->U_l = (int32_t)ul;
  std::vector<int32_t> deg(p.num_rows, 0);
  for (size_t z = 0; z < p.csr_rowIdx_l.size(); ++z)
    deg[p.csr_rowIdx_l[z]] += p.csr_rowPtr_l[z + 1] - p.csr_rowPtr_l[z];
  for (int32_t r : p.csc_rowIdx_r) deg[r]++;
  info->max_row_nnz = deg.empty() ? 0 : *std::max_element(deg.begin(), deg.end());
  return COLA_OK;
}

cola_status cola_partition_get_arrays(cola_partition h, cola_partition_arrays *a) {
  if (!h || !a) return fail(COLA_ERR_INVALID, "partition_arrays: NULL argument");
  const Partition &p = h->p;
  a->csr_rowPtr_l = p.csr_rowPtr_l.data();
  a->csr_colIdx_l = p.csr_colIdx_l.data();
  a->csr_rowIdx_l = p.csr_rowIdx_l.data();
  a->csr_val_l = p.csr_val_l.data();
  a->csc_colPtr_r = p.csc_colPtr_r.data();
  a->csc_rowIdx_r = p.csc_rowIdx_r.data();
  a->csc_colIdx_r = p.csc_colIdx_r.data();
  a->csc_val_r = p.csc_val_r.data();
  return COLA_OK;
}

cola_status cola_partition_destroy(cola_partition p) {
  delete p;
  return COLA_OK;
}

cola_status cola_serialize(cola_graph g, const char *bin_path,
                           const char *graph_name, int nGPU, int write_values) {
  if (!g || !bin_path || !graph_name) return fail(COLA_ERR_INVALID, "serialize: NULL argument");
  if (nGPU < 1) return fail(COLA_ERR_INVALID, "serialize: nGPU < 1");
  const Graph &G = g->g;
  if (G.col_idx.size() > (size_t)INT32_MAX) return fail(COLA_ERR_INVALID, "serialize: nnz overflows int32");
  const std::string gdir = std::string(bin_path) + "/" + graph_name;
  const std::string dir = gdir + "/GPU" + std::to_string(nGPU);
  cola_status st = mkdirs(dir);
  if (st) return st;
  std::vector<int32_t> rp32(G.row_ptr.begin(), G.row_ptr.end());
  if ((st = vec2bin(gdir + "/global_rowPtr.bin", rp32.data(), rp32.size()))) return st;
  if ((st = vec2bin(gdir + "/global_colIdx.bin", G.col_idx.data(), G.col_idx.size()))) return st;
  for (int r = 0; r < nGPU; ++r) {
    Partition p;
    if ((st = row_partition(G, nGPU, r, &p))) return st;
    struct {
      const char *stem;
      const std::vector<int32_t> *v;
    } files[] = {{"csr_rowPtr_l", &p.csr_rowPtr_l}, {"csr_colIdx_l", &p.csr_colIdx_l},
                 {"csr_rowIdx_l", &p.csr_rowIdx_l}, {"csc_rowIdx_r", &p.csc_rowIdx_r},
                 {"csc_colPtr_r", &p.csc_colPtr_r}, {"csc_colIdx_r", &p.csc_colIdx_r}};
    for (auto &f : files)
      if ((st = vec2bin(part_name(dir, f.stem, nGPU, r), f.v->data(), f.v->size()))) return st;
    if (write_values) {
      if ((st = vec2bin(part_name(dir, "csr_val_l", nGPU, r), p.csr_val_l.data(), p.csr_val_l.size())))
        return st;
      if ((st = vec2bin(part_name(dir, "csc_val_r", nGPU, r), p.csc_val_r.data(), p.csc_val_r.size())))
        return st;
    }
  }
  return COLA_OK;
}

cola_status cola_serialize_ppl(cola_graph g, const char *bin_path, const char *graph_name, int nGPU,
                               int group_size) {
  if (!g || !bin_path || !graph_name) return fail(COLA_ERR_INVALID, "serialize_ppl: NULL argument");
  if (nGPU < 2 || group_size < 1 || group_size >= nGPU)
    return fail(COLA_ERR_INVALID, "serialize_ppl: need 1 <= group_size < nGPU");
  const Graph &G = g->g;
  const int64_t n = G.n;
  const int stages = nGPU / group_size;
  const int64_t P = (n + nGPU - 1) / nGPU;
  const std::string dir = std::string(bin_path) + "/" + graph_name + "/ppl" + std::to_string(stages);
  cola_status st = mkdirs(dir);
  if (st) return st;
  // Stage s covers rows [s*gs*P, min((s+1)*gs*P, n)); its file holds the
  // sorted unique columns of those rows.
  std::vector<uint8_t> seen(n, 0);
  std::vector<int32_t> cols;
  for (int s = 0; s < stages; ++s) {
    const int64_t lb = std::min<int64_t>(n, (int64_t)s * group_size * P);
    const int64_t ub = std::min<int64_t>(n, (int64_t)(s + 1) * group_size * P);
    std::fill(seen.begin(), seen.end(), 0);
    for (int64_t k = G.row_ptr[lb]; k < G.row_ptr[ub]; ++k) seen[G.col_idx[k]] = 1;
    cols.clear();
    for (int64_t c = 0; c < n; ++c)
      if (seen[c]) cols.push_back((int32_t)c);
    const std::string path = dir + "/colIdx_" + std::to_string(nGPU) + "_" + std::to_string(stages) + "_" +
                             std::to_string(s) + ".bin";
    if ((st = vec2bin(path, cols.data(), cols.size()))) return st;
  }
  return COLA_OK;
}

cola_status cola_partition_load(const char *bin_path, const char *graph_name,
                                int nGPU, int rank, cola_partition *out) {
  if (!bin_path || !graph_name || !out) return fail(COLA_ERR_INVALID, "partition_load: NULL argument");
  *out = nullptr;
  if (nGPU < 1 || rank < 0 || rank >= nGPU) return fail(COLA_ERR_INVALID, "partition_load: rank out of range");
  const std::string gdir = std::string(bin_path) + "/" + graph_name;
  const std::string dir = gdir + "/GPU" + std::to_string(nGPU);
  std::vector<int32_t> grp;
  cola_status st = bin2vec(gdir + "/global_rowPtr.bin", &grp);
  if (st) return st;
  if (grp.empty()) return fail(COLA_ERR_IO, "global_rowPtr.bin is empty");
  auto *h = new (std::nothrow) cola_partition_s;
  if (!h) return fail(COLA_ERR_NOMEM, "partition_load: out of host memory");
  Partition &p = h->p;
  p.num_nodes = (int64_t)grp.size() - 1;
  p.nGPU = nGPU;
  p.rank = rank;
  const int64_t P = (p.num_nodes + nGPU - 1) / nGPU;
  p.nodesPerPE = (int32_t)P;
  p.num_rows = (int32_t)std::max<int64_t>(0, std::min(p.num_nodes, (int64_t)(rank + 1) * P) - (int64_t)rank * P);
  bool miss_l = false, miss_r = false;
  if ((st = bin2vec(part_name(dir, "csr_rowPtr_l", nGPU, rank), &p.csr_rowPtr_l)) ||
      (st = bin2vec(part_name(dir, "csr_colIdx_l", nGPU, rank), &p.csr_colIdx_l)) ||
      (st = bin2vec(part_name(dir, "csr_rowIdx_l", nGPU, rank), &p.csr_rowIdx_l)) ||
      (st = bin2vec(part_name(dir, "csc_colPtr_r", nGPU, rank), &p.csc_colPtr_r)) ||
      (st = bin2vec(part_name(dir, "csc_rowIdx_r", nGPU, rank), &p.csc_rowIdx_r)) ||
      (st = bin2vec(part_name(dir, "csc_colIdx_r", nGPU, rank), &p.csc_colIdx_r)) ||
      (st = bin2vec(part_name(dir, "csr_val_l", nGPU, rank), &p.csr_val_l, &miss_l)) ||
      (st = bin2vec(part_name(dir, "csc_val_r", nGPU, rank), &p.csc_val_r, &miss_r))) {
    delete h;
    return st;
  }
  // Absent value files: every value is 1.0 (distSpMM_ppl_batch.cu:83,99).
  if (miss_l) p.csr_val_l.assign(p.csr_colIdx_l.size(), 1.0f);
  if (miss_r) p.csc_val_r.assign(p.csc_rowIdx_r.size(), 1.0f);
  if ((st = validate_partition(p))) {
    delete h;
    return st;
  }
  *out = h;
  return COLA_OK;
}

cola_status cola_nnz_split(const int32_t *rowIdx, int64_t nrows,
                           const int32_t *rowPtr, int64_t nnz, int32_t *target,
                           int32_t *ptr, int64_t *n_virtual, int32_t *gran_out) {
  if (!n_virtual) return fail(COLA_ERR_INVALID, "nnz_split: n_virtual is NULL");
  if (nrows < 0 || nnz < 0) return fail(COLA_ERR_INVALID, "nnz_split: negative size");
  // rowIdx.size()==0: nothing is produced (cuh:33-34).
  if (nrows == 0) {
    *n_virtual = 0;
    if (gran_out) *gran_out = 0;
    return COLA_OK;
  }
  if (!rowPtr) return fail(COLA_ERR_INVALID, "nnz_split: rowPtr is NULL");
  const int64_t avg = nnz / nrows;  // integer average (cuh:34)
  const int64_t gran = (avg + 32 - 1) / 32 * 32;
  if (gran <= 0) return fail(COLA_ERR_INVALID, "nnz_split: fewer entries than rows (gran would be 0)");
  if (gran_out) *gran_out = (int32_t)gran;
  int64_t nv = 0;
  for (int64_t i = 0; i < nrows; ++i) {
    const int64_t lb = rowPtr[i], hb = rowPtr[i + 1];
    if (hb > lb) nv += (hb - lb + gran - 1) / gran;
  }
  if (target == nullptr && ptr == nullptr) {
    *n_virtual = nv;
    return COLA_OK;
  }
  if (!target || !ptr || !rowIdx) return fail(COLA_ERR_INVALID, "nnz_split: output arrays are NULL");
  if (*n_virtual < nv) return fail(COLA_ERR_INVALID, "nnz_split: output capacity too small");
  int64_t w = 0;
  ptr[0] = 0;
  for (int64_t i = 0; i < nrows; ++i) {
    for (int64_t lb = rowPtr[i], hb = rowPtr[i + 1]; lb < hb; lb += gran) {
      target[w] = rowIdx[i];
      ptr[++w] = (int32_t)(lb + gran > hb ? hb : lb + gran);
    }
  }
  *n_virtual = nv;
  return COLA_OK;
}

}  // extern "C"
