// image.hip -- the frame image of write_image (src/main.cpp:556-678) on the
// device: every particle adds `brightness` to the pixel it projects to, the
// sums become 8-bit grey or a two-segment heat map, optionally normalised
// to the brightest pixel and optionally square-rooted first.
//
// The reference accumulates temp[x][y] += brightness particle by particle.
// All addends are equal, so a pixel's value depends only on how many
// particles it holds: here a splat counts them (32-bit atomics), and one
// thread per pixel forms the value by the same sequence of additions.  The
// projections, the scale law and the bounds keep the reference's
// expressions and quirks:
//   * front (x, z), side (y, z), top (x, y), each offset by img/2 in
//     integer division; iso ((sqrt(3)/2 (x - y) + w) / 2,
//     (-(x + y + 2z)/2 + h) / 2), with the scale doubled;
//   * x += (x - w/2)(scale - 1), the same for y with h/2;
//   * a particle is dropped when x or y is negative or above w - 1 (y too
//     is bounded by the width), then y is clamped to h - 1 and, in iso,
//     flipped;
//   * the pixel is (int) x, (int) y.
// A particle whose projection is not a number is dropped (the reference's
// (int) cast of it is undefined).  Compiled with -ffp-contract=off.
#include <hip/hip_runtime.h>

#include "nbs_device_util.h"
#include "nbs_internal.h"

namespace nbs {
namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ double clampd(double a, double x, double b) {
  if (x < a) return a;  // src/main.cpp:177-182
  if (x > b) return b;
  return x;
}

__global__ void splat_kernel(const double* __restrict__ px, const double* __restrict__ py,
                             const double* __restrict__ pz, int n, ImageParams p,
                             unsigned* __restrict__ counts) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double x0 = px[i], y0 = py[i], z0 = pz[i];
  const double hw = (double)(p.width / 2), hh = (double)(p.height / 2);
  double x, y, scale = p.scale;
  if (p.projection == 0) {
    x = x0 + hw;
    y = z0 + hh;
  } else if (p.projection == 1) {
    x = y0 + hw;
    y = z0 + hh;
  } else if (p.projection == 2) {
    x = x0 + hw;
    y = y0 + hh;
  } else {
    scale = scale * 2.0;
    x = ((1.7320508075688772 / 2.0) * (x0 - y0) + (double)p.width) / 2.0;
    y = ((-0.5) * (x0 + y0 + 2 * z0) + (double)p.height) / 2.0;
  }
  x += (x - hw) * (scale - 1);
  y += (y - hh) * (scale - 1);
  const double wmax = (double)(p.width - 1), hmax = (double)(p.height - 1);
  if (x < 0 || y < 0 || x > wmax || y > wmax) return;
  if (x != x || y != y) return;
  x = clampd(0, x, wmax);
  y = clampd(0, y, hmax);
  if (p.projection == 3) y = hmax - y;
  atomicAdd(counts + (size_t)(int)x * p.height + (int)y, 1u);
}

// temp = brightness added `count` times; the block's largest value goes to
// partial[block] (one contended atomic per block on a single word cost
// 95 us at 8100 blocks; max_kernel reduces the partials instead).
// With brightness = M 2^E (M odd) every partial sum j M 2^E is exact while
// count M < 2^53, and the sum is then the product count * brightness; the
// additions are repeated only otherwise (a brightness such as 0.1).
__global__ void value_kernel(const unsigned* __restrict__ counts, size_t npix, ImageParams p,
                             double* __restrict__ temp, double* __restrict__ partial) {
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  double t = 0.0;
  if (k < npix) {
    const unsigned c = counts[k];
    const unsigned long long bits = (unsigned long long)__double_as_longlong(p.brightness);
    const unsigned ex = (unsigned)(bits >> 52) & 0x7ffu;
    unsigned long long odd = 0;
    if (ex != 0 && ex != 0x7ffu) {  // a normal number
      const unsigned long long mant = (bits & ((1ull << 52) - 1)) | (1ull << 52);
      odd = mant >> __builtin_ctzll(mant);
    }
    // c * odd < 2^53, tested without forming the product (it can pass
    // 2^64 for a 53-bit odd mantissa such as 0.1's)
    if (odd != 0 && (c == 0 || odd <= ((1ull << 53) - 1) / c)) {
      t = p.brightness * (double)c;
    } else {
      for (unsigned j = 0; j < c; j++) t += p.brightness;
    }
    temp[k] = t;
  }
  __shared__ double red[kBlock / kWave];
  const double m = wave_max(k < npix ? t : 0.0);
  if (lane_id() == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double b = red[0];
    for (int w = 1; w < kBlock / kWave; w++) b = b > red[w] ? b : red[w];
    partial[blockIdx.x] = b;
  }
}

// The reference's running maximum: from 0, over every block's partial.
constexpr int kMaxThreads = 1024;
__global__ void __launch_bounds__(kMaxThreads)
    max_kernel(const double* __restrict__ partial, unsigned nblocks,
               unsigned long long* __restrict__ vmax) {
  double m = 0.0;
  for (unsigned b = threadIdx.x; b < nblocks; b += kMaxThreads)
    m = partial[b] > m ? partial[b] : m;
  m = wave_max(m);
  __shared__ double red[kMaxThreads / kWave];
  if (lane_id() == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kMaxThreads / kWave; w++) m = red[w] > m ? red[w] : m;
    m = red[0] > m ? red[0] : m;
    *vmax = ord_enc(m);
  }
}

// Pixel colours (src/main.cpp:616-670), written as PNG rows: row r holds
// the reference's y = h - 1 - r (FreeImage stores the bottom line first).
__global__ void color_kernel(const double* __restrict__ temp, ImageParams p,
                             const unsigned long long* __restrict__ vmax,
                             unsigned char* __restrict__ rgb) {
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (size_t)p.width * p.height) return;
  const unsigned xi = (unsigned)(k / p.height), yi = (unsigned)(k % p.height);
  double t = temp[k];
  double mx = ord_dec(*vmax);
  if (p.nonlinear) {
    mx = __dsqrt_rn(mx);
    t = __dsqrt_rn(t);
  }
  // A value that is not a number (adaptive scaling of an empty image, 0 /
  // 0) is black: the reference's (int) cast gives INT_MIN on x86, which its
  // byte stores and heat-map arithmetic turn into 0, 0, 0.
  unsigned char r = 0, g = 0, b = 0;
  const double v = p.color == 0
                       ? (p.adaptive ? clampd(0, t * 255.0 / mx, 255) : clampd(0, t, 255))
                       : (p.adaptive ? clampd(0, t * 511.0 / mx, 511) : clampd(0, t, 511));
  if (v != v) {
  } else if (p.color == 0) {
    r = g = b = (unsigned char)(int)v;
  } else {
    int iv = (int)v;
    if (iv < 256) {
      b = (unsigned char)(int)clampd(0, 255 - iv, 255);
      g = (unsigned char)(int)clampd(0, iv, 255);
      r = 0;
    } else {
      iv -= 256;
      b = 0;
      g = (unsigned char)(int)clampd(0, 255 - iv, 255);
      r = (unsigned char)(int)clampd(0, iv, 255);
    }
  }
  const size_t o = ((size_t)(p.height - 1 - yi) * p.width + xi) * 3;
  rgb[o] = r;
  rgb[o + 1] = g;
  rgb[o + 2] = b;
}

}  // namespace

size_t image_partials(size_t npix) { return (npix + kBlock - 1) / kBlock; }

void launch_image(const State& st, int n, const ImageParams& p, unsigned* counts, double* temp,
                  double* partial, unsigned long long* vmax, unsigned char* rgb,
                  hipStream_t s) {
  const size_t npix = (size_t)p.width * p.height;
  launch_fill_int((int*)counts, 0, npix, s);
  if (n > 0) splat_kernel<<<(n + kBlock - 1) / kBlock, kBlock, 0, s>>>(st.x, st.y, st.z, n, p,
                                                                      counts);
  const unsigned blocks = (unsigned)image_partials(npix);
  value_kernel<<<blocks, kBlock, 0, s>>>(counts, npix, p, temp, partial);
  max_kernel<<<1, kMaxThreads, 0, s>>>(partial, blocks, vmax);
  color_kernel<<<blocks, kBlock, 0, s>>>(temp, p, vmax, rgb);
}

}  // namespace nbs
