// adjoint.hip -- reverse mode of the step with the walk's cell decisions
// (tree topology and per-particle accepted cells) held fixed.
//
// Forward step t (fixed dt):  a = F(x, m);  v' = v + dt a;  x' = x + dt v'
// with F_i = sum over accepted cells n of G M_n K(c_n - x_i),
// K(r) = r f(|r|), f(d) = d^-3 (damped: 1 / (d (d^2 + e^-d))), and for a
// cell c_n = sum_j m_j x_j / M_n, M_n = sum_j m_j over its particles.
//
// Reverse step, given adjoints gx, gv of (x', v'):
//   gv <- gv + dt gx;  u = dt gv                   (u = adjoint of a)
//   per accepted (i, n):  w = G M_n J(r) u_i,  J = dK/dr (symmetric)
//       gx_i -= w;   gc_n += w;   gM_n += G K(r) . u_i
//   per particle j, over its ancestors n (leaf included):
//       gx_j += m_j sum_n gc_n / M_n
//       gm_j += x_j . sum_n gc_n / M_n + sum_n (gM_n - gc_n . c_n / M_n)
//   gx <- gx + (those);  gm accumulates over steps.
// The walk below visits cells exactly as the forward walk does (same MAC in
// the same precision), so the decisions are the forward's.
#include <cmath>

#include "nbs_device_util.h"
#include "nbs_internal.h"

namespace nbs {
namespace {

constexpr int kAdjBlock = 256;

typedef __attribute__((address_space(4))) const WalkNode* ConstNodes;

__device__ __forceinline__ double sq_sum_ref_a(double dx, double dy,
                                               double dz) {
#pragma clang fp contract(off)
  return dx * dx + dy * dy + dz * dz;
}

__device__ __forceinline__ bool mac_open_f64_a(double d2, double side,
                                               double theta, double theta2) {
#pragma clang fp contract(off)
  const double s2 = side * side;
  const double t = s2 - theta2 * d2;
  const double tol = 1e-9 * s2;
  if (t > tol) return true;
  if (t < -tol) return false;
  return side / __builtin_sqrt(d2) > theta;
}

// Totals of four values over the wave: lane l ends with the total of value
// l >> 4 (lanes 0-15 a, 16-31 b, 32-47 c, 48-63 d).  Each exchange halves
// the values a lane carries (two at xor 32, one at xor 16), so the four
// sums take seven shuffles instead of twenty-four.
__device__ __forceinline__ double wave_sum4(double a, double b, double c, double d) {
  const int lane = lane_id();
  const bool h32 = (lane & 32) != 0, h16 = (lane & 16) != 0;
  double k0 = h32 ? c : a, k1 = h32 ? d : b;
  const double s0 = h32 ? a : c, s1 = h32 ? b : d;
  k0 += __shfl_xor(s0, 32, 64);
  k1 += __shfl_xor(s1, 32, 64);
  double k = h16 ? k1 : k0;
  k += __shfl_xor(h16 ? k0 : k1, 16, 64);
#pragma unroll
  for (int q = 8; q >= 1; q >>= 1) k += __shfl_xor(k, q, 64);
  return k;
}

// 1 / sqrt(d2) in FP64: the hardware estimate and one Newton-Raphson step
// (relative error ~1e-15).
__device__ __forceinline__ double rsqrt_f64_a(double d2) {
  const double r = __builtin_amdgcn_rsq(d2);
  const double h = d2 * r;
  const double e = fma(-h, r, 1.0);
  return fma(0.5 * r, e, r);
}

// Where a piece of the reverse walk starts (split > 1): the walk state at
// node b of the whole traversal, from a descent from the root with this
// walk's opening tests (lanes that accept an ancestor wait at its skip,
// the rest take b), as the forward walk's pieces start.  With no lane
// taking b the piece resumes at the stack top, or has nothing to do.
template <bool MAC32>
__device__ __forceinline__ void adj_enter(ConstNodes nodes, int root, int b, int e, double x,
                                          double y, double z, double theta, double theta2,
                                          unsigned long long& A, int& n, int& top,
                                          int& st_pos, int& st_lo, int& st_hi) {
  const int lane = lane_id();
  int X = root;
  while (X < b) {
    const int leaf = nodes[X].leaf;
    const int skip = nodes[X].skip;
    bool open;
    if (leaf == -2) {
      open = true;
    } else {
      const double dx = nodes[X].cx - x, dy = nodes[X].cy - y, dz = nodes[X].cz - z;
      if (MAC32) {
        const float fdx = (float)dx, fdy = (float)dy, fdz = (float)dz;
        open = fmaf(fdx, fdx, fmaf(fdy, fdy, fdz * fdz)) < nodes[X].r2open32;
      } else {
        open = mac_open_f64_a(sq_sum_ref_a(dx, dy, dz), nodes[X].side, theta, theta2);
      }
    }
    const bool in = (A >> lane) & 1ull;
    const unsigned long long O = ballot(open && in);
    const unsigned long long C = A & ~O;
    if (C) {
      if (top > 0 && __builtin_amdgcn_readlane(st_pos, top - 1) == skip) {
        const int slot = top - 1;
        st_lo = lane == slot ? (st_lo | (int)(unsigned)C) : st_lo;
        st_hi = lane == slot ? (st_hi | (int)(unsigned)(C >> 32)) : st_hi;
      } else {
        const int slot = top;
        st_pos = lane == slot ? skip : st_pos;
        st_lo = lane == slot ? (int)(unsigned)C : st_lo;
        st_hi = lane == slot ? (int)(unsigned)(C >> 32) : st_hi;
        if (top < kWave - 1) top++;
      }
    }
    A = O;
    int c = X + 1;  // the child of X whose subtree holds b
    for (;;) {
      const int sk = max(nodes[c].skip, c + 1);
      if (sk > b) break;
      c = sk;
    }
    X = c;
  }
  n = b;
  if (!A) {
    const int p = top > 0 ? __builtin_amdgcn_readlane(st_pos, top - 1) : e;
    if (p < e) {
      n = p;
      const unsigned lo = (unsigned)__builtin_amdgcn_readlane(st_lo, top - 1);
      const unsigned hi = (unsigned)__builtin_amdgcn_readlane(st_hi, top - 1);
      A = ((unsigned long long)hi << 32) | lo;
      top--;
    } else {
      n = e;
    }
  }
}

// Eight wavefronts per SIMD (at most 64 VGPRs and 80 SGPRs each on gfx950)
// where that needs no spills: without damping (54-56 VGPRs; 74-76 before
// the bound); the damped variants keep their registers.  SPLIT: wavefront
// q walks piece q % split of traversal q / split.
template <bool MAC32, bool DAMP, bool SPLIT>
__global__ void __launch_bounds__(kAdjBlock) __attribute__((amdgpu_waves_per_eu(DAMP ? 1 : 8, 8)))
adj_walk_kernel(AdjArgs a) {
  if (*a.flags) return;
  const int lane = lane_id();
  const int q = __builtin_amdgcn_readfirstlane((int)((blockIdx.x * kAdjBlock + threadIdx.x) >> 6));
  const int wave = SPLIT ? q / a.split : q;
  const int piece = SPLIT ? q - wave * a.split : 0;
  if (wave >= a.n_waves) return;
  const int s = a.nsys > 1 ? __builtin_amdgcn_readfirstlane(
                                 sys_of_wave(wave, a.wave_start, a.nsys))
                           : 0;
  const int pbeg = a.sys_start[s] + (wave - a.wave_start[s]) * kWave;
  const int pend = a.sys_start[s + 1];
  const int i = pbeg + lane;
  const bool valid = i < pend;
  const double x = valid ? a.x[i] : 0.0;
  const double y = valid ? a.y[i] : 0.0;
  const double z = valid ? a.z[i] : 0.0;
  const double ux = valid ? a.ux[i] : 0.0;
  const double uy = valid ? a.uy[i] : 0.0;
  const double uz = valid ? a.uz[i] : 0.0;

  ConstNodes nodes = (ConstNodes)a.nodes;
  unsigned long long A = ballot(valid);
  int n = a.node_start[s];
  int nend = a.node_start[s + 1];
  const double theta = a.theta, theta2 = theta * theta;
  int st_pos = 0, st_lo = 0, st_hi = 0, top = 0;
  double gx = 0, gy = 0, gz = 0;
  if constexpr (SPLIT) {
    const long long len = nend - n;
    const int b = n + (int)(len * piece / a.split);
    const int e = n + (int)(len * (piece + 1) / a.split);
    adj_enter<MAC32>(nodes, n, b, e, x, y, z, theta, theta2, A, n, top, st_pos, st_lo, st_hi);
    nend = e;
  }

  while (n < nend) {
    const int leaf = nodes[n].leaf;
    const int skip = nodes[n].skip;
    const double dx = nodes[n].cx - x;
    const double dy = nodes[n].cy - y;
    const double dz = nodes[n].cz - z;
    const bool in = (A >> lane) & 1ull;
    const double d2 = sq_sum_ref_a(dx, dy, dz);
    bool inter;
    int next;
    if (leaf >= 0) {
      inter = in && leaf != i;
      next = n + 1;
    } else {
      bool open;
      if (leaf == -2) {
        open = true;
      } else if (MAC32) {
        const float fdx = (float)dx, fdy = (float)dy, fdz = (float)dz;
        const float fd2 = fmaf(fdx, fdx, fmaf(fdy, fdy, fdz * fdz));
        open = fd2 < nodes[n].r2open32;
      } else {
        open = mac_open_f64_a(d2, nodes[n].side, theta, theta2);
      }
      open = open && in;
      const unsigned long long O = ballot(open);
      inter = in && !open;
      if (O) {
        const unsigned long long C = A & ~O;
        if (C) {
          if (top > 0 && __builtin_amdgcn_readlane(st_pos, top - 1) == skip) {
            const int slot = top - 1;
            st_lo = lane == slot ? (st_lo | (int)(unsigned)C) : st_lo;
            st_hi = lane == slot ? (st_hi | (int)(unsigned)(C >> 32)) : st_hi;
          } else {
            const int slot = top;
            st_pos = lane == slot ? skip : st_pos;
            st_lo = lane == slot ? (int)(unsigned)C : st_lo;
            st_hi = lane == slot ? (int)(unsigned)(C >> 32) : st_hi;
            if (top < kWave - 1) top++;
          }
        }
        A = O;
        next = n + 1;
      } else {
        next = skip;
      }
    }
    const unsigned long long I = ballot(inter);
    if (I) {
      double wx = 0, wy = 0, wz = 0, wm = 0;
      if (inter) {
        const double rd = rsqrt_f64_a(d2);
        const double d = d2 * rd;
        double f, fp_over_d;  // f(d), f'(d) / d
        if (DAMP) {
          const double e = exp(-d);
          const double q = d2 + e;
          f = rd / q;
          fp_over_d = -(3.0 * d2 + e * (1.0 - d)) * rd * rd * rd / (q * q);
        } else {
          f = rd * rd * rd;
          fp_over_d = -3.0 * f * rd * rd;
        }
        const double GM = nodes[n].gm;
        const double ru = dx * ux + dy * uy + dz * uz;
        // w = G M (f u + (f'/d) r (r.u))
        wx = GM * (f * ux + fp_over_d * dx * ru);
        wy = GM * (f * uy + fp_over_d * dy * ru);
        wz = GM * (f * uz + fp_over_d * dz * ru);
        wm = kG * f * ru;  // G K(r) . u
        gx -= wx;
        gy -= wy;
        gz -= wz;
      }
      // the wave's totals: lanes 0, 16, 32, 48 add gc_n (x, y, z) and gM_n,
      // four words of one 32-byte record
      const double t = wave_sum4(wx, wy, wz, wm);
      if ((lane & 15) == 0) atomicAdd(a.gnode + 4 * (size_t)n + (lane >> 4), t);
    }
    n = max(next, n + 1);
    if (top > 0 && __builtin_amdgcn_readlane(st_pos, top - 1) == n) {
      const unsigned lo = (unsigned)__builtin_amdgcn_readlane(st_lo, top - 1);
      const unsigned hi = (unsigned)__builtin_amdgcn_readlane(st_hi, top - 1);
      A |= ((unsigned long long)hi << 32) | lo;
      top--;
    }
  }
  if (valid) {
    if constexpr (SPLIT) {
      const size_t np = (size_t)a.n;
      a.part[((size_t)piece * 3 + 0) * np + i] = gx;
      a.part[((size_t)piece * 3 + 1) * np + i] = gy;
      a.part[((size_t)piece * 3 + 2) * np + i] = gz;
    } else {
      a.gx[i] = gx;
      a.gy[i] = gy;
      a.gz[i] = gz;
    }
  }
}

// The pieces' position adjoints added in piece order.
__global__ void __launch_bounds__(kAdjBlock) adj_combine_kernel(AdjArgs a) {
  const int i = blockIdx.x * kAdjBlock + threadIdx.x;
  if (i >= a.n || *a.flags) return;
  const size_t np = (size_t)a.n;
  double gx = 0, gy = 0, gz = 0;
  for (int q = 0; q < a.split; q++) {
    gx = gx + a.part[((size_t)q * 3 + 0) * np + i];
    gy = gy + a.part[((size_t)q * 3 + 1) * np + i];
    gz = gz + a.part[((size_t)q * 3 + 2) * np + i];
  }
  a.gx[i] = gx;
  a.gy[i] = gy;
  a.gz[i] = gz;
}

// Leaf-to-root sums of the cell adjoints, per particle (slot order).
__global__ void __launch_bounds__(kAdjBlock)
    adj_propagate_kernel(const int* __restrict__ offsets, int n,
                         const WalkNode* __restrict__ nodes,
                         const int* __restrict__ parent,
                         const double* __restrict__ gnode,
                         const double* __restrict__ x,
                         const double* __restrict__ y,
                         const double* __restrict__ z,
                         const double* __restrict__ m, double* __restrict__ gx,
                         double* __restrict__ gy, double* __restrict__ gz,
                         double* __restrict__ gm, const int* __restrict__ flags) {
  const int i = blockIdx.x * kAdjBlock + threadIdx.x;
  if (i >= n || *flags) return;
  int k = offsets[i + 1] - 1;
  double Ax = 0, Ay = 0, Az = 0, B = 0;
  for (int level = 0; k >= 0 && level <= kLevels + 2; level++) {
    const double M = nodes[k].m;
    if (M != 0) {
      const double inv = 1.0 / M;
      const double cx = gnode[4 * (size_t)k], cy = gnode[4 * (size_t)k + 1],
                   cz = gnode[4 * (size_t)k + 2];
      Ax += cx * inv;
      Ay += cy * inv;
      Az += cz * inv;
      B += gnode[4 * (size_t)k + 3] - (cx * nodes[k].cx + cy * nodes[k].cy + cz * nodes[k].cz) * inv;
    }
    k = parent[k];
  }
  const double mi = m[i];
  gx[i] += mi * Ax;
  gy[i] += mi * Ay;
  gz[i] += mi * Az;
  gm[i] = x[i] * Ax + y[i] * Ay + z[i] * Az + B;
}

__global__ void adj_kick_kernel(double* __restrict__ gv,
                                const double* __restrict__ gx,
                                double* __restrict__ u, double dt, int n) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= 3 * n) return;
  const double g = gv[k] + dt * gx[k];
  gv[k] = g;
  u[k] = dt * g;
}

__global__ void gather_comp_kernel(const double* __restrict__ src,
                                   double* __restrict__ dst,
                                   const unsigned* __restrict__ id, int n,
                                   int ncomp) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const unsigned j = id[s];
  for (int c = 0; c < ncomp; c++) dst[(size_t)c * n + s] = src[(size_t)c * n + j];
}

__global__ void scatter_comp_kernel(const double* __restrict__ src,
                                    double* __restrict__ dst,
                                    const unsigned* __restrict__ id, int n,
                                    int ncomp, int add) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const unsigned j = id[s];
  for (int c = 0; c < ncomp; c++) {
    const double v = src[(size_t)c * n + s];
    double* d = dst + (size_t)c * n + j;
    *d = add ? *d + v : v;
  }
}

}  // namespace

void launch_adj_kick(double* gv, const double* gx, double* u, double dt,
                     int n, hipStream_t s) {
  if (n > 0) adj_kick_kernel<<<(3 * n + 255) / 256, 256, 0, s>>>(gv, gx, u, dt, n);
}

void launch_gather_comp(const double* src, double* dst, const unsigned* id,
                        int n, int ncomp, hipStream_t s) {
  if (n > 0)
    gather_comp_kernel<<<(n + 255) / 256, 256, 0, s>>>(src, dst, id, n, ncomp);
}

void launch_scatter_comp(const double* src, double* dst, const unsigned* id,
                         int n, int ncomp, int add, hipStream_t s) {
  if (n > 0)
    scatter_comp_kernel<<<(n + 255) / 256, 256, 0, s>>>(src, dst, id, n, ncomp,
                                                        add);
}

void launch_adj_walk(const AdjArgs& a, int mac32, int damping, hipStream_t s) {
  if (a.n_waves <= 0) return;
  if (a.split > 1) {
    const int blocks = (int)(((long long)a.n_waves * a.split * kWave + kAdjBlock - 1) / kAdjBlock);
    if (mac32) {
      if (damping)
        adj_walk_kernel<true, true, true><<<blocks, kAdjBlock, 0, s>>>(a);
      else
        adj_walk_kernel<true, false, true><<<blocks, kAdjBlock, 0, s>>>(a);
    } else {
      if (damping)
        adj_walk_kernel<false, true, true><<<blocks, kAdjBlock, 0, s>>>(a);
      else
        adj_walk_kernel<false, false, true><<<blocks, kAdjBlock, 0, s>>>(a);
    }
    adj_combine_kernel<<<(a.n + kAdjBlock - 1) / kAdjBlock, kAdjBlock, 0, s>>>(a);
    return;
  }
  const int blocks = (a.n_waves * kWave + kAdjBlock - 1) / kAdjBlock;
  if (mac32) {
    if (damping)
      adj_walk_kernel<true, true, false><<<blocks, kAdjBlock, 0, s>>>(a);
    else
      adj_walk_kernel<true, false, false><<<blocks, kAdjBlock, 0, s>>>(a);
  } else {
    if (damping)
      adj_walk_kernel<false, true, false><<<blocks, kAdjBlock, 0, s>>>(a);
    else
      adj_walk_kernel<false, false, false><<<blocks, kAdjBlock, 0, s>>>(a);
  }
}

void launch_adj_propagate(const AdjArgs& a, const int* offsets,
                          const int* parent, const double* m, double* gm,
                          int n, hipStream_t s) {
  if (n <= 0) return;
  adj_propagate_kernel<<<(n + kAdjBlock - 1) / kAdjBlock, kAdjBlock, 0, s>>>(
      offsets, n, a.nodes, parent, a.gnode, a.x, a.y,
      a.z, m, a.gx, a.gy, a.gz, gm, a.flags);
}

}  // namespace nbs
