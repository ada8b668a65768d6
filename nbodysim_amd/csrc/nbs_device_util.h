// nbs_device_util.h -- wave64 helpers shared by the kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace nbs {

__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }

// Number of set bits of `mask` in lanes below this one.
__device__ __forceinline__ int mbcnt(unsigned long long mask) {
  return __builtin_amdgcn_mbcnt_hi((unsigned)(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0));
}

__device__ __forceinline__ unsigned long long ballot(bool p) {
  return __ballot(p);
}

template <class T>
__device__ __forceinline__ T wave_inclusive_scan(T v) {
  const int lane = lane_id();
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    T o = __shfl_up(v, d, 64);
    if (lane >= d) v += o;
  }
  return v;
}

template <class T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
  return v;
}

__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v = fmin(v, __shfl_xor(v, d, 64));
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v = fmax(v, __shfl_xor(v, d, 64));
  return v;
}

__device__ __forceinline__ int wave_min_int(int v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v = min(v, __shfl_xor(v, d, 64));
  return v;
}
__device__ __forceinline__ int wave_max_int(int v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v = max(v, __shfl_xor(v, d, 64));
  return v;
}

// Order-preserving map of doubles onto unsigned 64-bit integers, so
// min/max can use integer atomics.
__device__ __host__ __forceinline__ unsigned long long ord_enc(double v) {
  unsigned long long u = __builtin_bit_cast(unsigned long long, v);
  return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}
__device__ __host__ __forceinline__ double ord_dec(unsigned long long u) {
  u = (u >> 63) ? (u & 0x7fffffffffffffffull) : ~u;
  return __builtin_bit_cast(double, u);
}

// Bounding-box partial of one wave: part[k * nwaves + w] for k = min x,
// y, z, max x, y, z over the wave's valid lanes (every lane participates).
__device__ __forceinline__ void wave_bbox_partial(bool v, double x, double y,
                                                  double z, int w, int nwaves,
                                                  double* part) {
  const double mnx = wave_min(v ? x : INFINITY);
  const double mny = wave_min(v ? y : INFINITY);
  const double mnz = wave_min(v ? z : INFINITY);
  const double mxx = wave_max(v ? x : -INFINITY);
  const double mxy = wave_max(v ? y : -INFINITY);
  const double mxz = wave_max(v ? z : -INFINITY);
  if (lane_id() == 0) {
    part[w] = mnx;
    part[nwaves + w] = mny;
    part[2 * nwaves + w] = mnz;
    part[3 * nwaves + w] = mxx;
    part[4 * nwaves + w] = mxy;
    part[5 * nwaves + w] = mxz;
  }
}

// System owning walk wave w / particle slot i (batched mode): the last
// s with start[s] <= w (binary search over nsys+1 offsets).
__device__ __forceinline__ int sys_of_wave(int w, const int* __restrict__ ws,
                                           int nsys) {
  int lo = 0, hi = nsys - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (ws[mid] <= w)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

__device__ __forceinline__ int sys_of_slot(int i, const int* __restrict__ st,
                                           int nsys) {
  int lo = 0, hi = nsys - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (st[mid] <= i)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

}  // namespace nbs
