// Dirichlet rows on the device: f_r = x_r - g_i and row r of the CRS matrix replaced by
// the identity row, for each listed LID r = lids[i]
// (disc-fe/src/lof/Panzer_TpetraLinearObjContainer.hpp:228-237, 300-318).
// Shared by the single-mesh pass and the batch pass.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace txa {

// G lanes per row (64/G rows per wave) and kDirU passes per wave whose loads (LID, row range,
// the row's first G columns) are all issued before any store, so a wave has 64/G * kDirU rows
// in flight.  G: the longest row rounded up to a power of two, 8 to 64.
constexpr int kDirU = 8;

inline int dirichlet_lanes(int max_row) { return max_row <= 8 ? 8 : max_row <= 16 ? 16 : max_row <= 32 ? 32 : 64; }

// Rows i0 + u * 64/G + lane/G (u < kDirU, i < n) of one wave.  A LID outside [0, nrows)
// touches no row and sets *err.
template <int G>
__device__ __forceinline__ void dirichlet_wave(int64_t i0, int64_t n, int64_t nrows, const int32_t *__restrict__ lids,
                                               const double *__restrict__ g, const double *__restrict__ x,
                                               const int64_t *__restrict__ rowptr, const int32_t *__restrict__ col,
                                               double *__restrict__ f, double *__restrict__ vals, int *err) {
  constexpr int RPW = 64 / G;
  const int lane = threadIdx.x & 63, gl = lane % G;
  int64_t idx[kDirU], b[kDirU], e[kDirU];
  int32_t r[kDirU];
  bool ok[kDirU];
#pragma unroll
  for (int u = 0; u < kDirU; ++u) {
    idx[u] = i0 + u * RPW + lane / G;
    r[u] = idx[u] < n ? lids[idx[u]] : -1;
  }
#pragma unroll
  for (int u = 0; u < kDirU; ++u) {
    ok[u] = idx[u] < n && r[u] >= 0 && r[u] < nrows;
    if (idx[u] < n && !ok[u] && gl == 0 && err) atomicOr(err, 1);
    b[u] = ok[u] ? rowptr[r[u]] : 0;
    e[u] = ok[u] ? rowptr[r[u] + 1] : 0;
  }
  if (f)
#pragma unroll
    for (int u = 0; u < kDirU; ++u)
      if (ok[u] && gl == 0) f[r[u]] = x[r[u]] - (g ? g[idx[u]] : 0.0);
  if (!vals) return;
  int32_t c0[kDirU];
#pragma unroll
  for (int u = 0; u < kDirU; ++u) c0[u] = b[u] + gl < e[u] ? col[b[u] + gl] : -1;
#pragma unroll
  for (int u = 0; u < kDirU; ++u) {
    if (b[u] + gl < e[u]) vals[b[u] + gl] = c0[u] == r[u] ? 1.0 : 0.0;
    for (int64_t k = b[u] + gl + G; k < e[u]; k += G) vals[k] = col[k] == r[u] ? 1.0 : 0.0;
  }
}

}  // namespace txa
