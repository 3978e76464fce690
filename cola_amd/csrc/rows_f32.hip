// Row-kernel instantiations for B stored as f32 (EPL = 4 elements per
// 16-byte lane vector).  One translation unit per storage type so the
// instantiations compile in parallel.
#include "rows.hpp"

namespace cola_dev {
bool launch_rows_f32(const Shape &sh, const RowArgs &a, int blocks, int threads, hipStream_t s,
                      bool unroll16) {
  return launch_rows_impl<COLA_F32, 4>(sh, a, blocks, threads, s, unroll16);
}
bool launch_tiny_f32(const Shape &sh, const RowArgs &a, const int32_t *rowtab, int blocks, hipStream_t s) {
  return launch_tiny_impl<COLA_F32, 4>(sh, a, rowtab, blocks, s);
}
}  // namespace cola_dev
