// radix_sort.hip -- stable LSD radix sort of (u64 key, u32 value) pairs,
// 8-bit digits.  Per pass, three kernels: an upsweep writes each tile's
// digit histogram (digit-major); a digit scan (one workgroup per digit)
// turns a digit's tile counts into tile offsets within the digit and the
// digit's total; a downsweep scans the 256 digit totals itself, ranks the
// tile's keys stably with wave64 ballot matching, sorts the tile in LDS and
// writes digit runs out contiguously.
#include "nbs_device_util.h"
#include "nbs_internal.h"

namespace nbs {
namespace {

constexpr int kRsBlock = 256;
constexpr int kRsItems = 16;
constexpr int kRsTile = kRsBlock * kRsItems;
constexpr int kRadix = 256;
constexpr int kRsWaves = kRsBlock / 64;

__global__ void __launch_bounds__(kRsBlock)
    upsweep_kernel(const unsigned long long* __restrict__ keys, int n,
                   int shift, int ntiles, int* __restrict__ hist) {
  __shared__ int h[kRadix];
  h[threadIdx.x] = 0;
  __syncthreads();
  const int base = blockIdx.x * kRsTile;
#pragma unroll 4
  for (int k = 0; k < kRsItems; k++) {
    int i = base + k * kRsBlock + threadIdx.x;
    if (i < n) atomicAdd(&h[(keys[i] >> shift) & 255], 1);
  }
  __syncthreads();
  hist[threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

// Workgroup d: exclusive scan of hist[d * ntiles + t] over the tiles t
// (in place) and total[d] = the digit's count.
__global__ void __launch_bounds__(kRsBlock)
    digit_scan_kernel(int* __restrict__ hist, int ntiles, int* __restrict__ total) {
  __shared__ int wsum[kRsWaves];
  __shared__ int carry_s;
  const int t = threadIdx.x, lane = lane_id(), w = t >> 6;
  int* h = hist + (size_t)blockIdx.x * ntiles;
  if (t == 0) carry_s = 0;
  __syncthreads();
  for (int base = 0; base < ntiles; base += kRsBlock) {
    const int i = base + t;
    const int v = i < ntiles ? h[i] : 0;
    const int inc = wave_inclusive_scan(v);
    if (lane == 63) wsum[w] = inc;
    __syncthreads();
    int wb = 0, tot = 0;
    for (int k = 0; k < kRsWaves; k++) {
      if (k < w) wb += wsum[k];
      tot += wsum[k];
    }
    const int carry = carry_s;
    if (i < ntiles) h[i] = carry + wb + inc - v;
    __syncthreads();
    if (t == 0) carry_s = carry + tot;
    __syncthreads();
  }
  if (t == 0) total[blockIdx.x] = carry_s;
}

// Downsweep: 1024 threads, 4 keys each, rank a 4096-key tile in 4 rounds
// (each: per-wave ballot matching, per-digit prefix over the 16 waves, a
// stable scatter into LDS), then write the tile's digit runs out.
constexpr int kDsBlock = 1024;
constexpr int kDsItems = kRsTile / kDsBlock;
constexpr int kDsWaves = kDsBlock / 64;

__global__ void __launch_bounds__(kDsBlock)
    downsweep_kernel(const unsigned long long* __restrict__ kin,
                     const unsigned* __restrict__ vin,
                     unsigned long long* __restrict__ kout,
                     unsigned* __restrict__ vout, int n, int shift,
                     int ntiles, const int* __restrict__ toff,
                     const int* __restrict__ total) {
  __shared__ unsigned long long sk[kRsTile];
  __shared__ unsigned sv[kRsTile];
  __shared__ int tile_start[kRadix];
  __shared__ int gbase[kRadix];
  __shared__ int run[kRadix];
  __shared__ int round_total[kRadix];
  __shared__ int wcnt[kDsWaves][kRadix];
  __shared__ int wsum[kRadix / 64];
  __shared__ int gsum[kRadix / 64];

  const int t = threadIdx.x;
  const int lane = lane_id();
  const int w = t >> 6;
  const int base = blockIdx.x * kRsTile;
  const int tile_n = min(kRsTile, n - base);

  unsigned long long key[kDsItems];
  unsigned val[kDsItems];
  int dig[kDsItems];

  if (t < kRadix) run[t] = 0;
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kDsItems; k++) {
    const int j = k * kDsBlock + t;
    if (j < tile_n) {
      key[k] = kin[base + j];
      val[k] = vin[base + j];
      dig[k] = (int)((key[k] >> shift) & 255);
      atomicAdd(&run[dig[k]], 1);
    } else {
      key[k] = 0;
      val[k] = 0;
      dig[k] = kRadix;
    }
  }
  __syncthreads();
  // Tile-local exclusive scan of the digit counts, and the digit's global
  // start: exclusive scan of the digit totals plus this tile's offset
  // within the digit (one digit per thread of the first four waves).
  int c = 0, g = 0, inc = 0, ginc = 0;
  if (t < kRadix) {
    c = run[t];
    g = total[t];
    inc = wave_inclusive_scan(c);
    ginc = wave_inclusive_scan(g);
    if (lane == 63) {
      wsum[w] = inc;
      gsum[w] = ginc;
    }
  }
  __syncthreads();
  if (t < kRadix) {
    int wb = 0, gb = 0;
    for (int k = 0; k < w; k++) {
      wb += wsum[k];
      gb += gsum[k];
    }
    tile_start[t] = wb + inc - c;
    gbase[t] = gb + ginc - g + toff[t * ntiles + blockIdx.x];
    run[t] = 0;
  }
  __syncthreads();

  for (int k = 0; k < kDsItems; k++) {
    for (int q = lane; q < kRadix; q += 64) wcnt[w][q] = 0;
    const int d = dig[k];
    const bool valid = d < kRadix;
    unsigned long long peers = ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; b++) {
      const bool bit = (d >> b) & 1;
      const unsigned long long bb = ballot(bit);
      peers &= bit ? bb : ~bb;
    }
    const int rank = mbcnt(peers);
    if (valid && rank == 0) wcnt[w][d] = __popcll(peers);
    __syncthreads();
    if (t < kRadix) {  // exclusive prefix over the waves, per digit
      int acc = 0;
      for (int q = 0; q < kDsWaves; q++) {
        const int c = wcnt[q][t];
        wcnt[q][t] = acc;
        acc += c;
      }
      round_total[t] = acc;
    }
    __syncthreads();
    if (valid) {
      const int pos = tile_start[d] + run[d] + wcnt[w][d] + rank;
      sk[pos] = key[k];
      sv[pos] = val[k];
    }
    __syncthreads();
    if (t < kRadix) run[t] += round_total[t];
  }
  __syncthreads();

  for (int j = t; j < tile_n; j += kDsBlock) {
    const unsigned long long kk = sk[j];
    const int d = (int)((kk >> shift) & 255);
    const int o = gbase[d] + (j - tile_start[d]);
    if (o >= 0 && o < n) {
      kout[o] = kk;
      vout[o] = sv[j];
    }
  }
}

}  // namespace

size_t radix_tmp_ints(size_t n) {
  size_t ntiles = (n + kRsTile - 1) / kRsTile;
  if (ntiles == 0) ntiles = 1;
  return kRadix * ntiles + kRadix;
}

int radix_sort_pairs(unsigned long long* k0, unsigned* v0,
                     unsigned long long* k1, unsigned* v1, int n, int nbits,
                     int* tmp, hipStream_t s) {
  if (n <= 1 || nbits <= 0) return 0;
  int ntiles = (n + kRsTile - 1) / kRsTile;
  int* hist = tmp;                   // [kRadix][ntiles], scanned in place
  int* total = tmp + kRadix * ntiles;  // [kRadix]
  unsigned long long* ka = k0;
  unsigned* va = v0;
  unsigned long long* kb = k1;
  unsigned* vb = v1;
  int which = 0;
  for (int shift = 0; shift < nbits; shift += 8) {
    upsweep_kernel<<<ntiles, kRsBlock, 0, s>>>(ka, n, shift, ntiles, hist);
    digit_scan_kernel<<<kRadix, kRsBlock, 0, s>>>(hist, ntiles, total);
    downsweep_kernel<<<ntiles, kDsBlock, 0, s>>>(ka, va, kb, vb, n, shift,
                                                 ntiles, hist, total);
    unsigned long long* tk = ka;
    ka = kb;
    kb = tk;
    unsigned* tv = va;
    va = vb;
    vb = tv;
    which ^= 1;
  }
  return which;
}

}  // namespace nbs
