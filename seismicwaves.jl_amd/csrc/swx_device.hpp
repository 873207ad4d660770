// Device helpers shared by the engines.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "swx_internal.hpp"

namespace swx {

// Store of a value the next launch reads back from HBM (a new time level,
// replay history, a gradient field): non-temporal, so the stream of new
// levels does not evict the halo lines neighbouring tiles reuse from L2.
template <typename T>
__device__ __forceinline__ void st_nt(T *p, T v) {
  __builtin_nontemporal_store(v, p);
}

// Index of the unique point group at linear cell `cell` (grid row `row`) of
// shot s, or -1.  rowptr/cells are a PointTables row index (sorted cells).
__device__ __forceinline__ int find_point(const int *rowptr, const int64_t *cells, int64_t nrows, int s,
                                          int64_t row, int64_t cell) {
  SWX_DASSERT(row >= 0 && row < nrows && s >= 0);
  const int *rp = rowptr + (int64_t)s * (nrows + 1) + row;
  int lo = rp[0], hi = rp[1];
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    int64_t v = cells[mid];
    if (v == cell) return mid;
    if (v < cell) lo = mid + 1;
    else hi = mid;
  }
  return -1;
}

// Packed C-PML coefficients of one axis: a (2h+2) | a_h (2h) | b (2h+2) |
// b_h (2h).  One pointer per axis (the four vectors are found by offset), so
// a kernel keeps few uniform registers for them.
template <typename T>
struct AxisCoef {
  const T *p;
  int h;
  __device__ __forceinline__ T a(int k) const { SWX_DASSERT(k >= 0 && k < 2 * h + 2); return p[k]; }
  __device__ __forceinline__ T ah(int k) const { SWX_DASSERT(k >= 0 && k < 2 * h); return p[2 * h + 2 + k]; }
  __device__ __forceinline__ T b(int k) const { SWX_DASSERT(k >= 0 && k < 2 * h + 2); return p[4 * h + 2 + k]; }
  __device__ __forceinline__ T bh(int k) const { SWX_DASSERT(k >= 0 && k < 2 * h); return p[6 * h + 4 + k]; }
};
template <typename T>
__device__ __forceinline__ AxisCoef<T> axis_coef(const T *base, int h, int axis) {
  return AxisCoef<T>{base + axis * (8 * h + 4), h};
}

// Sum over shots (in shot order) times `scale`, or a scaled copy per shot.
template <typename T>
__global__ void reduce_shots(T *out, const T *grad, int64_t N, int ns, T scale, int per_shot) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < N; q += (int64_t)gridDim.x * blockDim.x) {
    if (per_shot) {
      for (int s = 0; s < ns; s++) out[(int64_t)s * N + q] = grad[(int64_t)s * N + q] * scale;
    } else {
      T acc = 0;
      for (int s = 0; s < ns; s++) acc += grad[(int64_t)s * N + q] * scale;
      out[q] = acc;
    }
  }
}

// Copies n elements per shot (shot slabs of n) from src to dst for the
// shots being replayed (nt_s > act_min).
template <typename T>
__global__ void copy_active_shots(T *dst, const T *src, int64_t n, const int *nt_shot, int act_min) {
  const int s = blockIdx.y;
  if (nt_shot[s] <= act_min) return;
  const int64_t o = (int64_t)s * n;
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x)
    dst[o + q] = src[o + q];
}

// Packs the per-shot C-PML coefficient vectors of the shot descriptors.
template <typename T>
std::vector<T> pack_cpml(int nshots, const swx_shot_desc *shots, int ndim, int h) {
  const int stride = ndim * (8 * h + 4);
  std::vector<T> hc((size_t)nshots * stride);
  const int la = 2 * h + 2, lh = 2 * h;
  for (int s = 0; s < nshots; s++) {
    T *dst = hc.data() + (size_t)s * stride;
    for (int ax = 0; ax < ndim; ax++) {
      const int len[4] = {la, lh, la, lh};
      T *p = dst + ax * (8 * h + 4);
      for (int k = 0; k < 4; k++) {
        const T *src = static_cast<const T *>(shots[s].cpml[ax][k]);
        if (len[k] && !src) throw Error(SWX_EINVAL, "missing C-PML coefficient array");
        if (len[k]) std::memcpy(p, src, sizeof(T) * len[k]);
        p += len[k];
      }
    }
  }
  return hc;
}

}  // namespace swx
