// scan.hip -- device-wide exclusive prefix sum of int32: one block for up
// to 16K items, else three phases (tile sums, scan of tile sums, tile
// scans).  Used for the node offsets of the tree emission and the radix
// sort's digit offsets.
#include "nbs_device_util.h"
#include "nbs_internal.h"

namespace nbs {
namespace {

constexpr int kScanBlock = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanBlock * kScanItems;

__device__ int block_exclusive_scan(int v, int* lds_wave, int* total) {
  const int lane = lane_id();
  const int w = threadIdx.x >> 6;
  int inc = wave_inclusive_scan(v);
  if (lane == 63) lds_wave[w] = inc;
  __syncthreads();
  int wbase = 0, tot = 0;
  for (int k = 0; k < kScanBlock / 64; k++) {
    int c = lds_wave[k];
    if (k < w) wbase += c;
    tot += c;
  }
  __syncthreads();
  *total = tot;
  return wbase + inc - v;
}

__global__ void __launch_bounds__(kScanBlock)
    tile_sum_kernel(const int* __restrict__ in, int n, int* __restrict__ sums) {
  __shared__ int lds[kScanBlock / 64];
  const int base = blockIdx.x * kScanTile;
  int v = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; k++) {
    int i = base + k * kScanBlock + threadIdx.x;
    if (i < n) v += in[i];
  }
  v = wave_sum(v);
  if (lane_id() == 0) lds[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    int t = 0;
    for (int k = 0; k < kScanBlock / 64; k++) t += lds[k];
    sums[blockIdx.x] = t;
  }
}

// One block scans the tile sums in place (exclusive) in chunks.
__global__ void __launch_bounds__(1024)
    scan_sums_kernel(int* __restrict__ sums, int ntiles) {
  __shared__ int lds[1024 / 64];
  __shared__ int carry_s;
  if (threadIdx.x == 0) carry_s = 0;
  __syncthreads();
  for (int base = 0; base < ntiles; base += 1024) {
    int i = base + threadIdx.x;
    int v = i < ntiles ? sums[i] : 0;
    const int lane = lane_id();
    const int w = threadIdx.x >> 6;
    int inc = wave_inclusive_scan(v);
    if (lane == 63) lds[w] = inc;
    __syncthreads();
    int wbase = 0, tot = 0;
    for (int k = 0; k < 16; k++) {
      int c = lds[k];
      if (k < w) wbase += c;
      tot += c;
    }
    int carry = carry_s;
    if (i < ntiles) sums[i] = carry + wbase + inc - v;
    __syncthreads();
    if (threadIdx.x == 0) carry_s = carry + tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) sums[ntiles] = carry_s;
}

__global__ void __launch_bounds__(kScanBlock)
    tile_scan_kernel(const int* __restrict__ in, int* __restrict__ out, int n,
                     const int* __restrict__ sums) {
  __shared__ int lds[kScanBlock / 64];
  const int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
  int v[kScanItems];
  int s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; k++) {
    int i = base + k;
    v[k] = i < n ? in[i] : 0;
    s += v[k];
  }
  int total;
  int run = block_exclusive_scan(s, lds, &total) + sums[blockIdx.x];
#pragma unroll
  for (int k = 0; k < kScanItems; k++) {
    int i = base + k;
    if (i < n) out[i] = run;
    run += v[k];
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) out[n] = sums[gridDim.x];
}

// Short inputs: one block of 1024 threads scans chunks of 1024 x 16
// contiguous items with a running carry, one launch instead of three.
constexpr int kOneBlock = 1024;
constexpr int kOneItems = 16;
constexpr int kOneChunk = kOneBlock * kOneItems;
// A single block is latency-bound past one chunk (~17 us per chunk of
// 16K items measured, against ~16 us for the three-kernel scan of 64K).
constexpr int kOneBlockMax = kOneChunk;

__global__ void __launch_bounds__(kOneBlock)
    one_block_scan_kernel(const int* __restrict__ in, int* __restrict__ out, int n) {
  __shared__ int lds[kOneBlock / 64];
  __shared__ int carry_s;
  const int lane = lane_id();
  const int w = threadIdx.x >> 6;
  if (threadIdx.x == 0) carry_s = 0;
  __syncthreads();
  for (int base = 0; base < n; base += kOneChunk) {
    const int i0 = base + threadIdx.x * kOneItems;
    int v[kOneItems];
    int sum = 0;
#pragma unroll
    for (int k = 0; k < kOneItems; k++) {
      v[k] = i0 + k < n ? in[i0 + k] : 0;
      sum += v[k];
    }
    const int inc = wave_inclusive_scan(sum);
    if (lane == 63) lds[w] = inc;
    __syncthreads();
    int wbase = 0, tot = 0;
    for (int k = 0; k < kOneBlock / 64; k++) {
      const int c = lds[k];
      if (k < w) wbase += c;
      tot += c;
    }
    int run = carry_s + wbase + inc - sum;
#pragma unroll
    for (int k = 0; k < kOneItems; k++) {
      if (i0 + k < n) out[i0 + k] = run;
      run += v[k];
    }
    __syncthreads();
    if (threadIdx.x == 0) carry_s += tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) out[n] = carry_s;
}

}  // namespace

size_t scan_tmp_ints(size_t n) { return (n + kScanTile - 1) / kScanTile + 2; }

void exclusive_scan(const int* in, int* out, int n, int* tmp, hipStream_t s) {
  if (n <= kOneBlockMax) {
    one_block_scan_kernel<<<1, kOneBlock, 0, s>>>(in, out, n);
    return;
  }
  int ntiles = (n + kScanTile - 1) / kScanTile;
  if (ntiles == 0) ntiles = 1;
  tile_sum_kernel<<<ntiles, kScanBlock, 0, s>>>(in, n, tmp);
  scan_sums_kernel<<<1, 1024, 0, s>>>(tmp, ntiles);
  tile_scan_kernel<<<ntiles, kScanBlock, 0, s>>>(in, out, n, tmp);
}

}  // namespace nbs
