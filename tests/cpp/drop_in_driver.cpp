// Host driver written the way the reference's src/distSpMM_ppl_batch.cu:17-275
// uses include/sparse_data.h, compiled against this engine's headers: load
// one rank's partition files, nnz_split both halves, build d_csr_zip /
// d_csc_zip from std::vectors, fill B with ones (fill_certain, utils.cuh:31-35),
// run the single-stage wrapper, and check C[i,:] = deg(i) with the
// reference's 1e-2 relative test (utils.cuh:71-100).
//
// With a sixth argument N, it then times N more calls after 10 warm-up
// calls with HIP events and prints the average, as the reference's harness
// does (run_distSpMM_ppl_batch, dist_spmm_ppl_batch.cuh:920-988; C keeps
// accumulating, as there).
//
// usage: drop_in_driver <bin_path> <graph> <num_GPUs> <rank> <dim> [N]
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "sparse_data.h"

static std::vector<int> split(const int32_t *rowIdx, int64_t nrows, const int32_t *rowPtr, int64_t nnz,
                              std::vector<int> *ptr) {
  int64_t nv = 0;
  int32_t gran = 0;
  cola_throw_if(cola_nnz_split(rowIdx, nrows, rowPtr, nnz, nullptr, nullptr, &nv, &gran));
  std::vector<int> target(nv > 0 ? nv : 0);
  ptr->assign(nv + 1, 0);
  if (nrows == 0) {
    ptr->clear();
    return target;
  }
  cola_throw_if(cola_nnz_split(rowIdx, nrows, rowPtr, nnz, target.data(), ptr->data(), &nv, &gran));
  return target;
}

int main(int argc, char **argv) {
  if (argc < 6) {
    fprintf(stderr, "usage: %s <bin_path> <graph> <num_GPUs> <rank> <dim>\n", argv[0]);
    return 2;
  }
  const char *bin_path = argv[1], *graph = argv[2];
  const int num_GPUs = atoi(argv[3]), rank = atoi(argv[4]), dim = atoi(argv[5]);
  try {
    cola_partition part = nullptr;
    cola_throw_if(cola_partition_load(bin_path, graph, num_GPUs, rank, &part));
    cola_partition_info info;
    cola_partition_arrays a;
    cola_throw_if(cola_partition_get_info(part, &info));
    cola_throw_if(cola_partition_get_arrays(part, &a));
    const int P = info.nodesPerPE, M = info.num_rows;

    std::vector<int> rowPtr_l, colPtr_r;
    std::vector<int> rowIdx_l = split(a.csr_rowIdx_l, info.L, a.csr_rowPtr_l, info.nnz_l, &rowPtr_l);
    std::vector<int> colIdx_r = split(a.csc_colIdx_r, info.U_r, a.csc_colPtr_r, info.nnz_r, &colPtr_r);
    std::vector<int> colIdx_l(a.csr_colIdx_l, a.csr_colIdx_l + info.nnz_l);
    std::vector<int> rowIdx_r(a.csc_rowIdx_r, a.csc_rowIdx_r + info.nnz_r);
    std::vector<float> val_l(info.nnz_l, 1.0f), val_r(info.nnz_r, 1.0f);
    if (rowPtr_l.empty()) rowPtr_l.push_back(0);
    if (colPtr_r.empty()) colPtr_r.push_back(0);
    d_csr_zip *csr_zip_l = new d_csr_zip(rowPtr_l, colIdx_l, rowIdx_l, val_l);
    d_csc_zip *csc_zip_r = new d_csc_zip(colPtr_r, rowIdx_r, colIdx_r, val_r);

    const size_t rowsB = (size_t)P * num_GPUs;
    std::vector<float> h_in(rowsB * dim, 1.0f);
    float *d_in = nullptr, *d_out = nullptr, *d_packed = nullptr;
    const size_t packed_rows = csc_zip_r->num_unique() > 0 ? csc_zip_r->num_unique() : 1;
    if (hipMalloc((void **)&d_in, h_in.size() * 4) != hipSuccess ||
        hipMalloc((void **)&d_out, (size_t)P * dim * 4) != hipSuccess ||
        hipMalloc((void **)&d_packed, packed_rows * dim * 4) != hipSuccess)
      throw cola_error(COLA_ERR_HIP, "hipMalloc");
    if (hipMemcpy(d_in, h_in.data(), h_in.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(d_out, 0, (size_t)P * dim * 4) != hipSuccess)
      throw cola_error(COLA_ERR_HIP, "hipMemcpy");
    hipStream_t stream;
    if (hipStreamCreate(&stream) != hipSuccess) throw cola_error(COLA_ERR_HIP, "hipStreamCreate");
    cola_throw_if(cola_dist_spmm_ppl_batch(d_out, d_in, d_packed, csr_zip_l->get(), csc_zip_r->get(), dim, P,
                                           rank, stream));
    if (hipStreamSynchronize(stream) != hipSuccess) throw cola_error(COLA_ERR_HIP, "sync");
    std::vector<float> h_out((size_t)P * dim);
    if (hipMemcpy(h_out.data(), d_out, h_out.size() * 4, hipMemcpyDeviceToHost) != hipSuccess)
      throw cola_error(COLA_ERR_HIP, "D2H");
    // expected: C[i,:] = deg(i), from the partition's own arrays
    std::vector<double> deg(M, 0.0);
    for (int64_t z = 0; z < info.L; ++z) deg[a.csr_rowIdx_l[z]] += a.csr_rowPtr_l[z + 1] - a.csr_rowPtr_l[z];
    for (int64_t k = 0; k < info.nnz_r; ++k) deg[a.csc_rowIdx_r[k]] += 1.0;
    int64_t bad = 0;
    for (int i = 0; i < M; ++i)
      for (int j = 0; j < dim; ++j)
        if (std::fabs(h_out[(size_t)i * dim + j] - deg[i]) > 1e-2 * std::fabs(deg[i])) ++bad;
    if (bad) {
      printf("PEid-%d failed: %lld wrong elements\n", rank, (long long)bad);
      return 1;
    }
    printf("PEid-%d passed\n", rank);
    const int runs = argc > 6 ? atoi(argv[6]) : 0;
    if (runs > 0) {
      for (int r = 0; r < 10; ++r)
        cola_throw_if(cola_dist_spmm_ppl_batch(d_out, d_in, d_packed, csr_zip_l->get(), csc_zip_r->get(), dim, P,
                                               rank, stream));
      hipEvent_t e0, e1;
      if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
        throw cola_error(COLA_ERR_HIP, "hipEventCreate");
      (void)hipEventRecord(e0, stream);
      for (int r = 0; r < runs; ++r)
        cola_throw_if(cola_dist_spmm_ppl_batch(d_out, d_in, d_packed, csr_zip_l->get(), csc_zip_r->get(), dim, P,
                                               rank, stream));
      (void)hipEventRecord(e1, stream);
      if (hipEventSynchronize(e1) != hipSuccess) throw cola_error(COLA_ERR_HIP, "hipEventSynchronize");
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, e0, e1);
      printf("PEid-%d average time over %d runs: %.4f ms\n", rank, runs, ms / runs);
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
    }
    delete csr_zip_l;
    delete csc_zip_r;
    (void)hipFree(d_in);
    (void)hipFree(d_out);
    (void)hipFree(d_packed);
    (void)hipStreamDestroy(stream);
    cola_partition_destroy(part);
  } catch (const cola_error &e) {
    fprintf(stderr, "cola error %d: %s\n", e.status, e.what());
    return 1;
  }
  return 0;
}
