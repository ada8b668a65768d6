// integrate.hip -- adaptive step length and the separate kick-drift update
// (the fixed-dt path fuses the update into the walk kernel).
//
// Adaptive dt: documented by the reference README (README.md:54-60:
// min_adaptive_dt, max_pos_change, max_vel_change) but not implemented in
// its code; the only hook is particle::update(dv, dp) returning |a| and |v|
// (src/particle.cpp:76-80).  This build defines it per system as
//   dt_s = max(min(dt, max_vel_change / max|a|, max_pos_change / max|v|),
//              min_adaptive_dt)
// with |a| from the current forces and |v| before the kick.
#include <cmath>

#include "nbs_device_util.h"
#include "nbs_internal.h"

namespace nbs {
namespace {

constexpr int kBlock = 256;

// Per-wave maxima of |a| and |v| (plain stores of wave partials; the
// per-system reduction follows in adaptive_dt_kernel).
__global__ void __launch_bounds__(kBlock)
    adaptive_max_kernel(const double* __restrict__ vx,
                        const double* __restrict__ vy,
                        const double* __restrict__ vz,
                        const double* __restrict__ ax,
                        const double* __restrict__ ay,
                        const double* __restrict__ az,
                        const int* __restrict__ start,
                        const int* __restrict__ wstart, int nsys, int nwaves,
                        double* __restrict__ amax, double* __restrict__ vmax,
                        const int* __restrict__ flags) {
  // |a| and |v| as magnitude() forms them (src/vector.cpp:104-107): each
  // square and sum rounded, summed left to right, so the step length can
  // be restated exactly on the host
#pragma clang fp contract(off)
  if (*flags) return;
  const int w = (blockIdx.x * kBlock + threadIdx.x) >> 6;
  if (w >= nwaves) return;
  const int s = nsys > 1 ? sys_of_wave(w, wstart, nsys) : 0;
  const int i = start[s] + (w - wstart[s]) * 64 + lane_id();
  const bool v = i < start[s + 1];
  double a2 = 0, v2 = 0;
  if (v) {
    a2 = sqrt(ax[i] * ax[i] + ay[i] * ay[i] + az[i] * az[i]);
    v2 = sqrt(vx[i] * vx[i] + vy[i] * vy[i] + vz[i] * vz[i]);
  }
  a2 = wave_max(a2);
  v2 = wave_max(v2);
  if (lane_id() == 0) {
    amax[w] = a2;
    vmax[w] = v2;
  }
}

// One workgroup per system: the system's maxima over its wave partials,
// then its step length.
constexpr int kDtBlock = 1024;
__global__ void __launch_bounds__(kDtBlock)
    adaptive_dt_kernel(const double* __restrict__ amax, const double* __restrict__ vmax,
                       const int* __restrict__ wstart, int nsys, double dt, double min_dt,
                       double max_dp, double max_dv, double* __restrict__ sys_dt,
                       double* __restrict__ sys_time, const int* __restrict__ flags) {
  const int s = blockIdx.x;
  if (s >= nsys || *flags) return;
  __shared__ double red[2][kDtBlock / 64];
  double am = 0, vm = 0;
#pragma unroll 4
  for (int w = wstart[s] + threadIdx.x; w < wstart[s + 1]; w += kDtBlock) {
    am = fmax(am, amax[w]);
    vm = fmax(vm, vmax[w]);
  }
  am = wave_max(am);
  vm = wave_max(vm);
  if (lane_id() == 0) {
    red[0][threadIdx.x >> 6] = am;
    red[1][threadIdx.x >> 6] = vm;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  for (int q = 1; q < kDtBlock / 64; q++) {
    am = fmax(am, red[0][q]);
    vm = fmax(vm, red[1][q]);
  }
  am = fmax(am, red[0][0]);
  vm = fmax(vm, red[1][0]);
  double h = dt;
  if (am > 0) h = fmin(h, max_dv / am);
  if (vm > 0) h = fmin(h, max_dp / vm);
  h = fmax(h, min_dt);
  sys_dt[s] = h;
  sys_time[s] += h;
}

__global__ void __launch_bounds__(kBlock)
    update_kernel(State st, const double* __restrict__ ax,
                  const double* __restrict__ ay, const double* __restrict__ az,
                  const int* __restrict__ start,
                  const int* __restrict__ wstart, int nsys, int nwaves,
                  const double* __restrict__ sys_dt, double* __restrict__ bbox,
                  const int* __restrict__ flags) {
#pragma clang fp contract(off)
  if (*flags) return;
  const int w = (blockIdx.x * kBlock + threadIdx.x) >> 6;
  if (w >= nwaves) return;
  const int s = nsys > 1 ? sys_of_wave(w, wstart, nsys) : 0;
  const int i = start[s] + (w - wstart[s]) * 64 + lane_id();
  const bool v = i < start[s + 1];
  const double dt = sys_dt[s];
  double x = 0, y = 0, z = 0;
  if (v) {
    // src/particle.cpp:66-74
    double vx = st.vx[i], vy = st.vy[i], vz = st.vz[i];
    const double tx = ax[i] * dt, ty = ay[i] * dt, tz = az[i] * dt;
    vx = vx + tx;
    vy = vy + ty;
    vz = vz + tz;
    const double px = vx * dt, py = vy * dt, pz = vz * dt;
    x = st.x[i] + px;
    y = st.y[i] + py;
    z = st.z[i] + pz;
    st.vx[i] = vx;
    st.vy[i] = vy;
    st.vz[i] = vz;
    st.x[i] = x;
    st.y[i] = y;
    st.z[i] = z;
  }
  wave_bbox_partial(v, x, y, z, w, nwaves, bbox);
}

}  // namespace

void launch_adaptive_dt(const State& st, const double* ax, const double* ay,
                        const double* az, int n, Systems& sys,
                        const StepParams& p, const int* flags, hipStream_t s) {
  (void)n;
  const int nw = sys.n_waves;
  if (nw > 0)
    adaptive_max_kernel<<<(nw * 64 + kBlock - 1) / kBlock, kBlock, 0, s>>>(
        st.vx, st.vy, st.vz, ax, ay, az, sys.start, sys.wave_start, sys.nsys,
        nw, sys.amax, sys.vmax, flags);
  adaptive_dt_kernel<<<sys.nsys, kDtBlock, 0, s>>>(
      sys.amax, sys.vmax, sys.wave_start, sys.nsys, p.dt, p.min_adaptive_dt,
      p.max_pos_change, p.max_vel_change, sys.dt, sys.time, flags);
}

void launch_update(State& st, const double* ax, const double* ay,
                   const double* az, int n, Systems& sys, const int* flags,
                   hipStream_t s) {
  (void)n;
  const int nw = sys.n_waves;
  if (nw > 0)
    update_kernel<<<(nw * 64 + kBlock - 1) / kBlock, kBlock, 0, s>>>(
        st, ax, ay, az, sys.start, sys.wave_start, sys.nsys, nw, sys.dt,
        sys.bbox, flags);
}

}  // namespace nbs
