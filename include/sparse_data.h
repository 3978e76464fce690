// sparse_data.h — C++ device sparse structures of the single-partition SpMM
// engine, with the member names and vector-taking constructors of the
// reference's include/sparse_data.h:6-77 (d_csr, d_csc_zip, d_csr_zip), so
// host code written against those structs compiles unchanged on MI355X.
//
// Differences from the reference structs:
//   - memory comes from hipMalloc (through the C ABI for the zip forms);
//   - failures throw cola_error (carrying the cola_status and
//     cola_last_error() text) instead of printing and calling exit();
//   - the structs are move-only (the reference's are copyable and would free
//     the same device pointers twice);
//   - d_csc_zip also exposes d_csc_packIdx (packed-buffer row of every
//     virtual column) and d_csc_uniq (the unique remote column ids), which
//     cola_pack_remote / cola_ppl0_batch_spmm take.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "cola_spmm.h"

struct cola_error : std::runtime_error {
  cola_status status;
  cola_error(cola_status s, const std::string &what) : std::runtime_error(what), status(s) {}
};

inline void cola_throw_if(cola_status s) {
  if (s != COLA_OK) throw cola_error(s, cola_last_error());
}

// sparse_data.h:6-27 d_csr: plain device CSR (row pointer, columns, values).
struct d_csr {
  int *d_csr_rowPtr = nullptr;
  int *d_csr_colIdx = nullptr;
  float *d_csr_val = nullptr;

  d_csr(std::vector<int> rowPtr, std::vector<int> colIdx, std::vector<float> val) {
    try {
      d_csr_rowPtr = upload(rowPtr);
      d_csr_colIdx = upload(colIdx);
      d_csr_val = upload(val);
    } catch (...) {
      release();
      throw;
    }
  }
  d_csr(const d_csr &) = delete;
  d_csr &operator=(const d_csr &) = delete;
  d_csr(d_csr &&o) noexcept { swap(o); }
  d_csr &operator=(d_csr &&o) noexcept {
    swap(o);
    return *this;
  }
  ~d_csr() { release(); }

 private:
  template <typename T>
  static T *upload(const std::vector<T> &v) {
    T *d = nullptr;
    if (v.empty()) return d;
    if (hipMalloc((void **)&d, v.size() * sizeof(T)) != hipSuccess)
      throw cola_error(COLA_ERR_HIP, "d_csr: hipMalloc failed");
    if (hipMemcpy(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess) {
      (void)hipFree(d);
      throw cola_error(COLA_ERR_HIP, "d_csr: hipMemcpy failed");
    }
    return d;
  }
  void release() {
    (void)hipFree(d_csr_rowPtr);
    (void)hipFree(d_csr_colIdx);
    (void)hipFree(d_csr_val);
    d_csr_rowPtr = d_csr_colIdx = nullptr;
    d_csr_val = nullptr;
  }
  void swap(d_csr &o) {
    std::swap(d_csr_rowPtr, o.d_csr_rowPtr);
    std::swap(d_csr_colIdx, o.d_csr_colIdx);
    std::swap(d_csr_val, o.d_csr_val);
  }
};

namespace cola_detail {
// Shared body of the two zip structs: one cola_zip handle.
struct zip_base {
  cola_zip handle = nullptr;
  int32_t *ptr = nullptr, *idx = nullptr, *zip = nullptr, *pack = nullptr;
  float *val = nullptr;
  int64_t nptr = 0, nnz = 0, n_unique = 0;

  zip_base(const std::vector<int> &p, const std::vector<int> &i, const std::vector<int> &z,
           const std::vector<float> &v) {
    // cola_zip_create reads size(p)-1 ids and size(i) values.
    if (z.size() != (p.empty() ? 0 : p.size() - 1) || (!v.empty() && v.size() != i.size()))
      throw cola_error(COLA_ERR_INVALID, "zip: need one id per pointer interval and one value per index");
    cola_throw_if(cola_zip_create(p.data(), (int64_t)p.size(), i.data(), (int64_t)i.size(), z.data(),
                                  v.empty() ? nullptr : v.data(), &handle));
    cola_throw_if(cola_zip_device_arrays(handle, &ptr, &idx, &zip, &val, &pack, &nptr, &nnz, &n_unique));
  }
  zip_base(const zip_base &) = delete;
  zip_base &operator=(const zip_base &) = delete;
  zip_base(zip_base &&o) noexcept { swap(o); }
  zip_base &operator=(zip_base &&o) noexcept {
    swap(o);
    return *this;
  }
  ~zip_base() {
    if (handle) cola_zip_destroy(handle);
  }
  void swap(zip_base &o) {
    std::swap(handle, o.handle);
    std::swap(ptr, o.ptr);
    std::swap(idx, o.idx);
    std::swap(zip, o.zip);
    std::swap(pack, o.pack);
    std::swap(val, o.val);
    std::swap(nptr, o.nptr);
    std::swap(nnz, o.nnz);
    std::swap(n_unique, o.n_unique);
  }
};
}  // namespace cola_detail

// sparse_data.h:28-52 d_csc_zip: remote CSC over the (split) unique remote
// columns; colIdx_zip holds global column ids.
struct d_csc_zip : private cola_detail::zip_base {
  int *d_csc_colPtr = nullptr;
  int *d_csc_rowIdx = nullptr;
  int *d_csc_colIdx_zip = nullptr;
  float *d_csc_val = nullptr;
  int *d_csc_packIdx = nullptr;

  d_csc_zip(std::vector<int> colPtr, std::vector<int> rowIdx, std::vector<int> colIdx_zip,
            std::vector<float> val)
      : zip_base(colPtr, rowIdx, colIdx_zip, val) {
    d_csc_colPtr = ptr;
    d_csc_rowIdx = idx;
    d_csc_colIdx_zip = zip;
    d_csc_val = val_ptr();
    d_csc_packIdx = pack;
  }
  cola_zip get() const { return handle; }
  int64_t num_columns() const { return nptr > 0 ? nptr - 1 : 0; }
  int64_t num_unique() const { return n_unique; }

 private:
  float *val_ptr() const { return zip_base::val; }
};

// sparse_data.h:53-77 d_csr_zip: local CSR over non-empty (split) rows;
// rowIdx_zip maps each virtual row to its output row.
struct d_csr_zip : private cola_detail::zip_base {
  int *d_csr_rowPtr = nullptr;
  int *d_csr_colIdx = nullptr;
  int *d_csr_rowIdx_zip = nullptr;
  float *d_csr_val = nullptr;

  d_csr_zip(std::vector<int> rowPtr, std::vector<int> colIdx, std::vector<int> rowIdx_zip,
            std::vector<float> val)
      : zip_base(rowPtr, colIdx, rowIdx_zip, val) {
    d_csr_rowPtr = ptr;
    d_csr_colIdx = idx;
    d_csr_rowIdx_zip = zip;
    d_csr_val = zip_base::val;
  }
  cola_zip get() const { return handle; }
  int64_t num_rows() const { return nptr > 0 ? nptr - 1 : 0; }
};
