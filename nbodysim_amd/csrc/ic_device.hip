// ic_device.hip -- seeded initial conditions generated on the device, one
// thread per particle, draw for draw as generate_particle
// (src/main.cpp:184-275) with std::default_random_engine (libstdc++'s
// minstd_rand0: x' = 16807 x mod 2^31 - 1) and uniform_real_distribution.
//
// With LINEAR mass and velocity distributions every particle of a system
// consumes a fixed number D of engine draws (two per double: libstdc++'s
// generate_canonical<double, 53> needs ceil(53 / 30) = 2 words of
// 30 bits), so particle i of a system starts at draw i * D, reached by
// skip-ahead: x_{iD} = 16807^{iD} x_0 mod (2^31 - 1).  Per particle
// (FP64, the reference's -DDOUBLE build):
//   mass                          2 draws
//   cube:   position z, y, x      6 draws (g++ evaluates the three vector
//           velocity z, y, x      6 draws  constructor arguments right to
//           speed                 2 draws  left, src/main.cpp:222)
//   sphere: radius fraction, azimuth, cos(polar)       6 draws
//   shell:  [radius factor, shell_jitter != 0] azimuth, cos(polar)
// The arithmetic is the host generator's with each operation rounded
// separately (this file is compiled with -ffp-contract=off): mass and the
// cube's positions and velocities are bit-identical to it.  Spheres and shells also call cbrt,
// acos, sin and cos, where the device math library and glibc may differ
// in the last bit.
#include <hip/hip_runtime.h>

#include "nbs_internal.h"

namespace nbs {
namespace {

constexpr unsigned long long kM = 2147483647ull;  // 2^31 - 1
constexpr unsigned long long kA = 16807ull;

// a * b mod 2^31 - 1 for a, b < 2^31 (fold the 62-bit product twice).
__device__ __forceinline__ unsigned long long mulmod(unsigned long long a,
                                                     unsigned long long b) {
  unsigned long long p = a * b;
  p = (p & kM) + (p >> 31);
  p = (p & kM) + (p >> 31);
  return p >= kM ? p - kM : p;
}

__device__ __forceinline__ unsigned long long powmod(unsigned long long a,
                                                     unsigned long long e) {
  unsigned long long r = 1;
  while (e) {
    if (e & 1) r = mulmod(r, a);
    a = mulmod(a, a);
    e >>= 1;
  }
  return r;
}

struct Engine {
  unsigned long long x;
  __device__ unsigned long long next() {
    x = mulmod(kA, x);
    return x;
  }
  // generate_canonical<double, 53>: two words of range r = 2^31 - 2,
  // sum = w1 + w2 * r, divided by r^2 (rounded to double), clamped below 1.
  __device__ double canonical() {
    const double r = 2147483646.0;
    const double w1 = (double)(next() - 1);
    const double w2 = (double)(next() - 1);
    const double sum = __dadd_rn(w1, __dmul_rn(w2, r));
    const double v = __ddiv_rn(sum, __dmul_rn(r, r));
    return v >= 1.0 ? 0x1.fffffffffffffp-1 : v;
  }
  // uniform_real_distribution<double>(a, b): canonical * (b - a) + a
  __device__ double uniform(double a, double b) {
    return __dadd_rn(__dmul_rn(canonical(), __dsub_rn(b, a)), a);
  }
};

__global__ void generate_kernel(IcParams p, int nsys, const long long* __restrict__ off,
                                const IcSystem* __restrict__ sys, State st) {
  const long long n = off[nsys];
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int lo = 0, hi = nsys - 1;  // last system with off[s] <= i
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= i) lo = mid; else hi = mid - 1;
  }
  const IcSystem q = sys[lo];
  const unsigned long long k = (unsigned long long)(i - off[lo]) * (unsigned)q.draws;
  Engine g{mulmod(powmod(kA, k), q.x0)};
  const double m = g.uniform(p.min_mass, p.max_mass);
  double px, py, pz, vx, vy, vz;
  if (q.gen_type == kIcCube) {
    const double h = __dmul_rn(p.size, 0.5);
    pz = g.uniform(-h, h);
    py = g.uniform(-h, h);
    px = g.uniform(-h, h);
    px = __dmul_rn(px, p.scale[0]);
    py = __dmul_rn(py, p.scale[1]);
    pz = __dmul_rn(pz, p.scale[2]);
    vz = g.uniform(-1.0, 1.0);
    vy = g.uniform(-1.0, 1.0);
    vx = g.uniform(-1.0, 1.0);
    const double vm = __dsqrt_rn(__dadd_rn(__dadd_rn(__dmul_rn(vx, vx), __dmul_rn(vy, vy)),
                                           __dmul_rn(vz, vz)));
    vx = __ddiv_rn(vx, vm);
    vy = __ddiv_rn(vy, vm);
    vz = __ddiv_rn(vz, vm);
    const double speed = g.uniform(p.min_vel, p.max_vel);
    vx = __dmul_rn(vx, speed);
    vy = __dmul_rn(vy, speed);
    vz = __dmul_rn(vz, speed);
  } else {
    double radius;
    if (q.gen_type == kIcSphere) {
      radius = __dmul_rn(cbrt(g.uniform(0.0, 1.0)), p.r_sphere);
    } else {
      radius = p.r_sphere;
      if (p.shell_jitter != 0)
        radius = __dmul_rn(radius,
                           __dadd_rn(1.0, __dmul_rn(p.shell_jitter, g.uniform(-1.0, 1.0))));
    }
    const double az = g.uniform(0.0, 6.28318530718);
    const double polar = acos(g.uniform(-1.0, 1.0));
    const double rs = __dmul_rn(radius, sin(polar));
    px = __dmul_rn(__dmul_rn(rs, cos(az)), p.scale[0]);
    py = __dmul_rn(__dmul_rn(rs, sin(az)), p.scale[1]);
    pz = __dmul_rn(__dmul_rn(radius, cos(polar)), p.scale[2]);
    // cross(rotation, point), src/vector.cpp:162-165
    vx = __dsub_rn(__dmul_rn(p.rot[1], pz), __dmul_rn(p.rot[2], py));
    vy = __dsub_rn(__dmul_rn(p.rot[2], px), __dmul_rn(p.rot[0], pz));
    vz = __dsub_rn(__dmul_rn(p.rot[0], py), __dmul_rn(p.rot[1], px));
  }
  st.x[i] = px;
  st.y[i] = py;
  st.z[i] = pz;
  st.vx[i] = vx;
  st.vy[i] = vy;
  st.vz[i] = vz;
  st.m[i] = m;
  st.id[i] = (unsigned)i;
}

}  // namespace

void launch_generate(const IcParams& p, int nsys, const long long* off, const IcSystem* sys,
                     long long n, State st, hipStream_t s) {
  if (n > 0)
    generate_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(p, nsys, off, sys, st);
}

}  // namespace nbs
