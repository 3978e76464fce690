// Host-side code under AddressSanitizer + UBSan (built by
// tests/test_host_sanitize.py with g++, no GPU): R-MAT graphs of every
// small size and order, all four relabelings, row partitions for several
// partition counts, plan building over a spread of plan options, nnz_split,
// the on-disk writers and loader (partition and stage files), MatrixMarket
// write/read round trips, and rejection of malformed MatrixMarket text and of
// partition files that are truncated or hold out-of-range indices.  Exits
// non-zero on the first unexpected status; the sanitizers abort on any
// memory or undefined-behaviour error.
#include <unistd.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../cola_amd/csrc/cola_internal.hpp"

#define CK(x)                                                              \
  do {                                                                     \
    cola_status s_ = (x);                                                  \
    if (s_) {                                                              \
      std::printf("FAIL %s: %d %s\n", #x, (int)s_, cola_last_error());     \
      std::exit(1);                                                        \
    }                                                                      \
  } while (0)

static int g_plans = 0;

static void build_plans(const cola::Partition &p) {
  // hub_threshold, chunk_nnz, kahan_degree, max_dim, short_degree, col_panels, row_order, hot_cols
  const int32_t sets[][8] = {{0, 0, 0, 0, 0, 0, 1, 0}, {0, 0, 0, 0, 0, 0, 0, 0}, {4, 2, 2, 0, 3, 0, 1, 0},
                             {8, 4, 0, 16, 1, 2, 1, 0}, {0, 0, 0, 0, 0, 4, 0, 0}, {0, 0, 0, 0, 0, 0, 1, 16},
                             {3, 1, 1, 0, 2, 0, 0, 5}, {2, 1, 1, 0, 0, 3, 1, 0}};
  for (const auto &o : sets) {
    cola_plan_opts po;
    static_assert(sizeof po == sizeof o, "cola_plan_opts layout");
    std::memcpy(&po, o, sizeof po);
    cola::HostPlan hp;
    CK(cola::build_plan(p, po, &hp));
    ++g_plans;
  }
}

static void split_both_sides(cola_partition p) {
  cola_partition_info info;
  cola_partition_arrays a;
  CK(cola_partition_get_info(p, &info));
  CK(cola_partition_get_arrays(p, &a));
  int64_t nv = 0;
  int32_t gran = 0;
  CK(cola_nnz_split(a.csr_rowIdx_l, info.L, a.csr_rowPtr_l, info.nnz_l, nullptr, nullptr, &nv, &gran));
  std::vector<int32_t> t(nv + 1), pt(nv + 1);
  CK(cola_nnz_split(a.csr_rowIdx_l, info.L, a.csr_rowPtr_l, info.nnz_l, t.data(), pt.data(), &nv, &gran));
  CK(cola_nnz_split(a.csc_colIdx_r, info.U_r, a.csc_colPtr_r, info.nnz_r, nullptr, nullptr, &nv, &gran));
  std::vector<int32_t> t2(nv + 1), p2(nv + 1);
  CK(cola_nnz_split(a.csc_colIdx_r, info.U_r, a.csc_colPtr_r, info.nnz_r, t2.data(), p2.data(), &nv, &gran));
  cola_partition q;
  CK(cola_partition_from_arrays(info.num_nodes, info.nGPU, info.rank, &a, info.L, info.nnz_l, info.U_r,
                                info.nnz_r, &q));
  cola_partition_destroy(q);
}

static void write_text(const std::string &path, const char *text) {
  FILE *f = std::fopen(path.c_str(), "wb");
  if (!f) {
    std::printf("FAIL cannot create %s\n", path.c_str());
    std::exit(1);
  }
  std::fwrite(text, 1, std::strlen(text), f);
  std::fclose(f);
}

int main(int argc, char **argv) {
  if (argc < 2) {
    std::printf("usage: driver <scratch dir>\n");
    return 2;
  }
  const std::string dir = argv[1];
  int parts = 0;
  for (int log_n = 1; log_n <= 10; ++log_n)
    for (int reorder = 0; reorder <= 2; ++reorder)
      for (double deg : {1.0, 5.0, 24.0}) {
        cola_graph g;
        CK(cola_rmat_graph(log_n, deg, 0.57, 0.19, 0.19, 7 + log_n, reorder, 1, &g));
        int64_t n, nnz, md;
        CK(cola_graph_info(g, &n, &nnz, &md));
        for (int m = COLA_ORDER_CUTHILL_MCKEE; m <= COLA_ORDER_COMMUNITY; ++m) {
          std::vector<int32_t> n2o(n);
          cola_graph r;
          CK(cola_graph_reorder(g, m, n2o.data(), &r));
          cola_graph_destroy(r);
        }
        for (int nGPU : {1, 2, 3, 8, 9}) {
          for (int rank = 0; rank < nGPU; ++rank) {
            cola_partition p;
            CK(cola_row_partition(g, nGPU, rank, &p));
            build_plans(p->p);
            split_both_sides(p);
            cola_partition_destroy(p);
            ++parts;
          }
          if (log_n < 3 || reorder != 2) continue;
          const std::string name = "g" + std::to_string(log_n) + "_" + std::to_string((int)deg);
          CK(cola_serialize(g, dir.c_str(), name.c_str(), nGPU, 1));
          for (int gs = 1; gs < nGPU; ++gs) CK(cola_serialize_ppl(g, dir.c_str(), name.c_str(), nGPU, gs));
          for (int rank = 0; rank < nGPU; ++rank) {
            cola_partition p;
            CK(cola_partition_load(dir.c_str(), name.c_str(), nGPU, rank, &p));
            build_plans(p->p);
            cola_partition_destroy(p);
          }
          const std::string mtx = dir + "/" + name + ".mtx";
          for (int wv = 0; wv <= 1; ++wv) {
            CK(cola_graph_write_mtx(g, mtx.c_str(), wv));
            for (int vp = COLA_MTX_VALUES_FILE; vp <= COLA_MTX_VALUES_GCN; ++vp) {
              cola_graph h;
              CK(cola_graph_from_mtx(mtx.c_str(), vp, &h));
              int64_t n2, nnz2, md2;
              CK(cola_graph_info(h, &n2, &nnz2, &md2));
              if (n2 != n || nnz2 != nnz) {
                std::printf("FAIL mtx round trip %s: %lld/%lld vs %lld/%lld\n", mtx.c_str(), (long long)n2,
                            (long long)nnz2, (long long)n, (long long)nnz);
                return 1;
              }
              cola_graph_destroy(h);
            }
          }
        }
        cola_graph_destroy(g);
      }

  // MatrixMarket text: which must be read and which rejected.
  struct Case {
    const char *text;
    bool ok;
  };
  const Case mtx_cases[] = {
      {"%%MatrixMarket matrix coordinate real general\n3 3 2\n1 1 1.0\n", false},
      {"%%MatrixMarket matrix coordinate real general\n3 3 1\n4 1 1.0\n", false},
      {"%%MatrixMarket matrix coordinate pattern symmetric\n3 3 2\n1 2\n3 3\n", true},
      {"%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n", false},
      {"%%MatrixMarket matrix coordinate real general\n", false},
      {"", false},
      {"%%MatrixMarket matrix coordinate integer general\n% c\n\n2 2 1\n2 1 -7\n", true},
      {"%%MatrixMarket matrix coordinate real general\n2 2 1\n1 x 1\n", false},
      {"%%MatrixMarket matrix coordinate real general\n2 2 1\n1 2\n", false},
      {"%%MatrixMarket matrix coordinate complex general\n2 2 1\n1 2 1 1\n", false},
      {"%%MatrixMarket matrix coordinate real general\n2 3 1\n1 2 1\n", false},
      {"%%MatrixMarket matrix coordinate real general\n2 2 -1\n", false},
      {"%%MatrixMarket matrix coordinate real general\n0 0 0\n", true},
  };
  int k = 0;
  for (const Case &c : mtx_cases) {
    const std::string path = dir + "/case" + std::to_string(k++) + ".mtx";
    write_text(path, c.text);
    cola_graph h;
    const cola_status s = cola_graph_from_mtx(path.c_str(), COLA_MTX_VALUES_FILE, &h);
    if ((s == COLA_OK) != c.ok) {
      std::printf("FAIL mtx case %d: status %d (%s)\n", k - 1, (int)s, s ? cola_last_error() : "read");
      return 1;
    }
    if (s == COLA_OK) cola_graph_destroy(h);
  }

  // Partition files of rank 1 of 2, each integer file damaged three ways
  // (cut short, an index of 2^30 in the middle, -5 first): every one is
  // rejected; damaged values load (any float is a value).
  cola_graph g;
  CK(cola_rmat_graph(8, 8, 0.57, 0.19, 0.19, 3, 0, 1, &g));
  const char *stems[] = {"csr_rowPtr_l", "csr_colIdx_l", "csr_rowIdx_l", "csc_colPtr_r",
                         "csc_rowIdx_r", "csc_colIdx_r", "csr_val_l",    "csc_val_r"};
  for (const char *stem : stems)
    for (int mode = 0; mode < 3; ++mode) {
      CK(cola_serialize(g, dir.c_str(), "damaged", 2, 1));
      const std::string path = dir + "/damaged/GPU2/" + stem + "_2_1.bin";
      FILE *f = std::fopen(path.c_str(), "r+b");
      if (!f) {
        std::printf("FAIL no %s\n", path.c_str());
        return 1;
      }
      std::fseek(f, 0, SEEK_END);
      const long size = std::ftell(f);
      if (mode == 0) {
        std::fclose(f);
        if (truncate(path.c_str(), size > 4 ? size - 4 : 0) != 0) return 1;
      } else {
        const int32_t v = mode == 1 ? (1 << 30) : -5;
        std::fseek(f, mode == 1 ? (size / 2) & ~3L : 0, SEEK_SET);
        std::fwrite(&v, 4, 1, f);
        std::fclose(f);
      }
      const bool values = std::strstr(stem, "_val_") != nullptr;
      cola_partition p;
      const cola_status s = cola_partition_load(dir.c_str(), "damaged", 2, 1, &p);
      const bool want_ok = values && mode != 0;
      if ((s == COLA_OK) != want_ok) {
        std::printf("FAIL damaged %s mode %d: status %d (%s)\n", stem, mode, (int)s,
                    s ? cola_last_error() : "loaded");
        return 1;
      }
      if (s == COLA_OK) {
        build_plans(p->p);
        cola_partition_destroy(p);
      }
    }
  cola_graph_destroy(g);
  std::printf("ok: %d partitions, %d plans\n", parts, g_plans);
  return 0;
}
