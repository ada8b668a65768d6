// walk.hip -- theta-MAC Barnes-Hut walk, one wavefront per 64 particles
// that are consecutive in octree (Morton) order.
//
// Semantics are the reference's per-particle walk (src/main.cpp:335-421):
// for each particle, a cell is opened iff side / |com - x| > theta and it
// is not a leaf; a leaf is never opened and interacts unless it is the
// particle itself; every accepted cell adds G m (com - x) / |com - x|^3
// (src/main.cpp:306-333).  The set of accepted cells of every particle is
// exactly the reference's; only the summation order differs (depth-first
// here, breadth-first in the reference).
//
// The 64 lanes share one traversal of the pre-order node array.  The wave
// visits a node only if some lane needs it; lanes that open it descend
// (next entry = first child), lanes that accept it wait at the node's
// `skip`.  Waiting lanes sit on a small stack of (position, lane mask)
// entries kept in registers across lanes (lane k holds entry k) --
// positions on it are skips of ancestors, so it never holds more entries
// than the tree has levels.  Because the node index is wave-uniform, node
// records are read with scalar loads through the scalar cache, and the
// traversal bookkeeping is scalar.
//
// Precision modes:
//   FP64: the MAC decision is bit-exact with the reference's
//     side / sqrt(dx*dx + dy*dy + dz*dz) > theta (correctly rounded sqrt and
//     division, no contraction); a squared test with a 1e-9 relative margin
//     settles all but the near-ties without the sqrt and division.  Forces:
//     1/r via v_rsq_f64 and two Newton steps.
//   FP32: com - x is formed in FP64 (full rate on CDNA4) and rounded to
//     FP32, which keeps near-field differences exact to FP32 precision
//     (node-relative coordinates); d^2, the MAC (d^2 < (side/theta)^2) and
//     the force (v_rsq_f32) are FP32, accumulated in FP32.
#include <cmath>
#include <type_traits>

#include "nbs_device_util.h"
#include "nbs_internal.h"

namespace nbs {
namespace {


typedef __attribute__((address_space(4))) const WalkNode* ConstNodes;

// The fields of a WalkNode the walk reads, loaded with scalar loads (the
// node index is wave-uniform); unused fields are dropped by the compiler.
struct NodeRegs {
  double cx, cy, cz, side, gm;
  float gm32, r2open32;
  int skip, leaf;
};

__device__ __forceinline__ NodeRegs load_node(ConstNodes nodes, int k) {
  NodeRegs r;
  r.cx = nodes[k].cx;
  r.cy = nodes[k].cy;
  r.cz = nodes[k].cz;
  r.side = nodes[k].side;
  r.gm = nodes[k].gm;
  r.gm32 = nodes[k].gm32;
  r.r2open32 = nodes[k].r2open32;
  r.skip = nodes[k].skip;
  r.leaf = nodes[k].leaf;
  return r;
}

__device__ __forceinline__ double sq_sum_ref(double dx, double dy, double dz) {
#pragma clang fp contract(off)
  return dx * dx + dy * dy + dz * dz;  // src/vector.cpp:146 rounding
}

__device__ __forceinline__ bool mac_open_f64(double d2, double side,
                                             double theta, double theta2) {
#pragma clang fp contract(off)
  const double s2 = side * side;
  const double t = s2 - theta2 * d2;
  const double tol = 1e-9 * s2;
  if (t > tol) return true;
  if (t < -tol) return false;
  return side / __builtin_sqrt(d2) > theta;
}

__device__ __forceinline__ double rsqrt_f64(double d2) {
  double r = __builtin_amdgcn_rsq(d2);
  double h = d2 * r;
  double e = fma(-h, r, 1.0);
  r = fma(0.5 * r, e, r);
  h = d2 * r;
  e = fma(-h, r, 1.0);
  r = fma(0.5 * r, e, r);
  return r;
}

// Kick-drift update of src/particle.cpp:66-74: t = a dt; v += t;
// t = v dt; x += t, each operation rounded.
__device__ __forceinline__ void kick_drift(double& x, double& v, double a,
                                           double dt) {
#pragma clang fp contract(off)
  const double t1 = a * dt;
  v = v + t1;
  const double t2 = v * dt;
  x = x + t2;
}

// FROZEN (FP64 only): the MAC is evaluated on a second tree (mac_nodes,
// same topology) against second positions (mac_x/y/z) -- a recorded base
// trajectory -- while forces use this step's tree and positions.  That
// makes the step a smooth function of the perturbed state, for finite
// differences of the adjoint.
// Workgroup -> tile order.  Workgroups are handed to the 8 XCDs round-robin
// by workgroup id, each XCD with its own L2.  With `groups` > 1 the tiles
// (runs of Morton-consecutive waves) are cut into `groups` contiguous
// regions and workgroup b works in region b % groups, so XCD k walks the
// regions k, k+8, ...: neighbouring waves, which open the same cells, share
// an L2.  groups = 1 keeps the plain order.  A bijection on [0, nblocks).
__device__ __forceinline__ int tile_of_block(int b, int nblocks, int groups) {
  if (groups <= 1) return b;
  const int v = b % groups, j = b / groups;
  const int q = nblocks / groups, r = nblocks % groups;
  return v * q + min(v, r) + j;
}

template <bool F32, bool DAMP, bool FROZEN, int BLOCK>
__global__ void __launch_bounds__(BLOCK) walk_kernel(WalkArgs a) {
  if (*a.flags) return;
  const int lane = lane_id();
  const int tile = tile_of_block(blockIdx.x, gridDim.x, a.groups);
  const int wave = __builtin_amdgcn_readfirstlane(
      (int)((tile * BLOCK + threadIdx.x) >> 6));
  if (wave >= a.n_waves) return;
  const int s = a.nsys > 1 ? __builtin_amdgcn_readfirstlane(
                                 sys_of_wave(wave, a.wave_start, a.nsys))
                           : 0;
  const int pbeg = a.sys_start[s] + (wave - a.wave_start[s]) * kWave;
  const int pend = a.sys_start[s + 1];
  const int i = pbeg + lane;
  const bool valid = i < pend;
  const double mx = FROZEN && valid ? a.mac_x[i] : 0.0;
  const double my = FROZEN && valid ? a.mac_y[i] : 0.0;
  const double mz = FROZEN && valid ? a.mac_z[i] : 0.0;
  ConstNodes mac_nodes = (ConstNodes)a.mac_nodes;
  double x = valid ? a.x[i] : 0.0;
  double y = valid ? a.y[i] : 0.0;
  double z = valid ? a.z[i] : 0.0;

  ConstNodes nodes = (ConstNodes)a.nodes;
  unsigned long long A = ballot(valid);
  int n = a.node_start[s];
  const int nend = a.node_start[s + 1];
  const double theta = a.theta;
  const double theta2 = theta * theta;

  int st_pos = 0;
  int st_lo = 0, st_hi = 0;
  int top = 0;
  int top_pos = -1;  // st_pos of the top entry, kept in an SGPR

  typedef typename std::conditional<F32, float, double>::type acc_t;
  acc_t ax = 0, ay = 0, az = 0;
  int cnt = 0;

  // One scalar load of the node record per visit.  (Loading both possible
  // successors a visit ahead measured slower: the doubled scalar-memory
  // traffic costs more than the latency it hides.)
  while (n < nend) {
    const NodeRegs cur = load_node(nodes, n);
    const int leaf = cur.leaf;
    const int skip = cur.skip;
    const double dx = cur.cx - x;
    const double dy = cur.cy - y;
    const double dz = cur.cz - z;
    const bool in = (A >> lane) & 1ull;
    float fdx = 0, fdy = 0, fdz = 0, fd2 = 0;
    double d2 = 0;
    if (F32) {
      fdx = (float)dx;
      fdy = (float)dy;
      fdz = (float)dz;
      fd2 = fmaf(fdx, fdx, fmaf(fdy, fdy, fdz * fdz));
    } else {
      d2 = sq_sum_ref(dx, dy, dz);
    }
    bool inter;
    int next;
    if (leaf >= 0) {
      inter = in && leaf != i;
      next = n + 1;
    } else {
      bool open;
      if (leaf == -2)
        open = true;
      else if (F32)
        open = fd2 < cur.r2open32;
      else
        open = mac_open_f64(
            FROZEN ? sq_sum_ref(mac_nodes[n].cx - mx, mac_nodes[n].cy - my,
                                mac_nodes[n].cz - mz)
                   : d2,
            cur.side, theta, theta2);
      open = open && in;
      const unsigned long long O = ballot(open);
      inter = in && !open;
      if (O) {
        const unsigned long long C = A & ~O;
        if (C) {
          // lane `slot` of the st_* registers holds stack entry `slot`
          if (top > 0 && top_pos == skip) {
            const int slot = top - 1;
            st_lo = lane == slot ? (st_lo | (int)(unsigned)C) : st_lo;
            st_hi = lane == slot ? (st_hi | (int)(unsigned)(C >> 32)) : st_hi;
          } else {
            const int slot = top;
            st_pos = lane == slot ? skip : st_pos;
            st_lo = lane == slot ? (int)(unsigned)C : st_lo;
            st_hi = lane == slot ? (int)(unsigned)(C >> 32) : st_hi;
            top_pos = skip;
            if (top < kWave - 1)
              top++;
            else if (lane == 0)
              atomicOr(a.flags_out, kFlagStack);
          }
        }
        A = O;
        next = n + 1;
      } else {
        next = skip;
      }
    }
    if (inter) {
      if (F32) {
        const float r = __builtin_amdgcn_rsqf(fd2);
        float sc;
        if (DAMP) {
          const float d = fd2 * r;
          sc = cur.gm32 * r / (fd2 + __expf(-d));
        } else {
          sc = cur.gm32 * (r * r * r);
        }
        ax = fmaf(sc, fdx, (float)ax);
        ay = fmaf(sc, fdy, (float)ay);
        az = fmaf(sc, fdz, (float)az);
      } else {
        const double r = rsqrt_f64(d2);
        double sc;
        if (DAMP) {
          const double d = d2 * r;
          sc = cur.gm * r / (d2 + exp(-d));
        } else {
          sc = cur.gm * (r * r * r);
        }
        ax = fma(sc, dx, (double)ax);
        ay = fma(sc, dy, (double)ay);
        az = fma(sc, dz, (double)az);
      }
      cnt++;
    }
    n = max(next, n + 1);  // always advance, even on a corrupt skip
    if (top > 0 && top_pos == n) {
      const unsigned lo = (unsigned)__builtin_amdgcn_readlane(st_lo, top - 1);
      const unsigned hi = (unsigned)__builtin_amdgcn_readlane(st_hi, top - 1);
      A |= ((unsigned long long)hi << 32) | lo;
      top--;
      top_pos = top > 0 ? __builtin_amdgcn_readlane(st_pos, top - 1) : -1;
    }
  }

  const double fax = (double)ax, fay = (double)ay, faz = (double)az;
  if (valid) {
    a.ax[i] = fax;
    a.ay[i] = fay;
    a.az[i] = faz;
    a.interactions[i] = cnt;
  }
  if (a.fused_update) {
    const double dt = a.dt;
    if (valid) {
      double vx = a.vx[i], vy = a.vy[i], vz = a.vz[i];
      kick_drift(x, vx, fax, dt);
      kick_drift(y, vy, fay, dt);
      kick_drift(z, vz, faz, dt);
      a.x[i] = x;
      a.y[i] = y;
      a.z[i] = z;
      a.vx[i] = vx;
      a.vy[i] = vy;
      a.vz[i] = vz;
    }
    wave_bbox_partial(valid, x, y, z, wave, a.n_waves, a.bbox);
    if (lane == 0 && wave == a.wave_start[s]) {
      a.sys_dt[s] = dt;
      a.sys_time[s] += dt;
    }
  }
  // per-wave running interaction count (no contended atomics)
  const int wtot = wave_sum(valid ? cnt : 0);
  if (lane == 0) a.inter_total[wave] += (unsigned long long)wtot;
  if (wave == 0 && lane == 0) atomicAdd(a.steps_done, 1);
}

// ------------------------------------------------------------ direct sum --
constexpr int kDirectChunks = 64;

__device__ __forceinline__ void kahan_add(double& s, double& c, double v) {
  const double y = v - c;
  const double t = s + y;
  c = (t - s) - y;
  s = t;
}

__global__ void __launch_bounds__(64)
    direct_kernel(const double* __restrict__ x, const double* __restrict__ y,
                  const double* __restrict__ z, const double* __restrict__ m,
                  int n, const int* __restrict__ samples, int ms, int damping,
                  double* __restrict__ part) {
#pragma clang fp contract(off)
  const int k = blockIdx.x * 64 + threadIdx.x;
  const int c = blockIdx.y;
  const bool valid = k < ms;
  const int si = valid ? samples[k] : 0;
  const double xi = x[si], yi = y[si], zi = z[si];
  const int per = (n + kDirectChunks - 1) / kDirectChunks;
  const int jb = c * per;
  const int je = min(n, jb + per);
  double sx = 0, sy = 0, sz = 0, cx = 0, cy = 0, cz = 0;
  for (int j = jb; j < je; j++) {
    const double dx = x[j] - xi, dy = y[j] - yi, dz = z[j] - zi;
    if (j != si) {
      const double d2 = dx * dx + dy * dy + dz * dz;
      const double d = __builtin_sqrt(d2);
      double sc;
      if (damping)
        sc = kG * m[j] / ((d2 + exp(-d)) * d);
      else
        sc = kG * m[j] / (d2 * d);
      kahan_add(sx, cx, sc * dx);
      kahan_add(sy, cy, sc * dy);
      kahan_add(sz, cz, sc * dz);
    }
  }
  if (valid) {
    double* p = part + ((size_t)c * ms + k) * 3;
    p[0] = sx;
    p[1] = sy;
    p[2] = sz;
  }
}

__global__ void direct_reduce_kernel(const double* __restrict__ part, int ms,
                                     double* __restrict__ out) {
#pragma clang fp contract(off)
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= ms) return;
  for (int d = 0; d < 3; d++) {
    double s = 0, c = 0;
    for (int q = 0; q < kDirectChunks; q++)
      kahan_add(s, c, part[((size_t)q * ms + k) * 3 + d]);
    out[3 * k + d] = s;
  }
}

__global__ void inverse_perm_kernel(const unsigned* __restrict__ id, int n,
                                    int* __restrict__ slot) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) slot[id[i]] = i;
}

}  // namespace

template <int BLOCK>
void launch_walk_block(const WalkArgs& a, int f32, int damping, hipStream_t s) {
  const int blocks = (a.n_waves * kWave + BLOCK - 1) / BLOCK;
  if (f32) {
    if (damping)
      walk_kernel<true, true, false, BLOCK><<<blocks, BLOCK, 0, s>>>(a);
    else
      walk_kernel<true, false, false, BLOCK><<<blocks, BLOCK, 0, s>>>(a);
  } else if (a.mac_nodes) {
    if (damping)
      walk_kernel<false, true, true, BLOCK><<<blocks, BLOCK, 0, s>>>(a);
    else
      walk_kernel<false, false, true, BLOCK><<<blocks, BLOCK, 0, s>>>(a);
  } else {
    if (damping)
      walk_kernel<false, true, false, BLOCK><<<blocks, BLOCK, 0, s>>>(a);
    else
      walk_kernel<false, false, false, BLOCK><<<blocks, BLOCK, 0, s>>>(a);
  }
}

void launch_walk(const WalkArgs& a, int f32, int damping, int block,
                 hipStream_t s) {
  if (a.n_waves <= 0) return;
  switch (block) {
    case 64:
      launch_walk_block<64>(a, f32, damping, s);
      break;
    case 128:
      launch_walk_block<128>(a, f32, damping, s);
      break;
    case 512:
      launch_walk_block<512>(a, f32, damping, s);
      break;
    case 1024:
      launch_walk_block<1024>(a, f32, damping, s);
      break;
    default:
      launch_walk_block<256>(a, f32, damping, s);
  }
}

size_t direct_tmp_doubles(size_t m) { return (size_t)kDirectChunks * m * 3; }

void launch_direct(const State& st, int n, const int* sample_slots, int m,
                   int damping, double* tmp, double* acc_out, hipStream_t s) {
  if (m <= 0) return;
  dim3 grid((m + 63) / 64, kDirectChunks);
  direct_kernel<<<grid, 64, 0, s>>>(st.x, st.y, st.z, st.m, n, sample_slots, m,
                                    damping, tmp);
  direct_reduce_kernel<<<(m + 255) / 256, 256, 0, s>>>(tmp, m, acc_out);
}

void launch_inverse_perm(const unsigned* id, int n, int* slot_of_id,
                         hipStream_t s) {
  if (n > 0)
    inverse_perm_kernel<<<(n + 255) / 256, 256, 0, s>>>(id, n, slot_of_id);
}

}  // namespace nbs
