// collide.hip -- collision merging on the device (src/main.cpp:384-396,
// :423-436, :438-550), after a walk that recorded its pairs.
//
// The reference's walk records every pair (curr, leaf particle) closer than
// collision_range into a hash set; after the walk it merges the pairs one
// by one, redirecting merged particles through an update table, so each
// connected group of the pair graph becomes one particle.  Each pairwise
// merge is collide(a, b): mass a.m + b.m, position and velocity
// weighted_average(a, b, a.m, b.m) = (a * a.m + b * b.m) / (a.m + b.m)
// (src/vector.cpp:129-132), acceleration zero for the frame's update.  The
// merged particle replaces the group in the set (update_collision,
// src/main.cpp:500-550) before update_all moves every particle.
//
// Here: the groups are found by a lock-free union-find over the recorded
// pairs, listed by (group root, caller id) and sorted, and one thread per
// group folds collide over the members in increasing caller id -- the
// reference's order is that of its hash set, which is not reproducible;
// for groups of two the result does not depend on it (both products and
// the sum commute).  The merged particle takes the slot and caller
// position of its lowest-id member.  After the kick-drift update, a stable
// compaction removes the absorbed members: survivors keep their relative
// order in slot order and in caller order, and each system's start moves
// down by the members its predecessors lost.
//
// Compiled with -ffp-contract=off: the weighted averages are rounded
// operation by operation, as in the reference.
#include <hip/hip_runtime.h>

#include "nbs_device_util.h"
#include "nbs_internal.h"

namespace nbs {
namespace {

constexpr int kBlock = 256;

int grid_for(int n) {
  const int b = (n + kBlock - 1) / kBlock;
  return b < 1 ? 1 : (b > 2048 ? 2048 : b);
}

__device__ __forceinline__ int load_parent(const int* parent, int x) {
  return __hip_atomic_load(parent + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Parents only ever point to smaller slots, so the chain ends.
__device__ int find_root(const int* parent, int x) {
  int p = load_parent(parent, x);
  while (p != x) {
    x = p;
    p = load_parent(parent, x);
  }
  return x;
}

// Every endpoint starts as its own group.  More detections than the pair
// buffer holds: the step is flagged and every later kernel of it does
// nothing (the host grows the buffer and redoes the step).  Also sets each
// system's step length for the update.
__global__ void init_kernel(const int2* __restrict__ pairs, const int* __restrict__ count,
                            int cap, int* __restrict__ parent, double* __restrict__ sys_dt,
                            int nsys, double dt, int* __restrict__ flags) {
  if (blockIdx.x == 0)
    for (int k = threadIdx.x; k < nsys; k += blockDim.x) sys_dt[k] = dt;
  if (*flags) return;
  const int c = count[0];
  if (c > cap) {
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicOr(flags, kFlagPairs);
    return;
  }
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < c; k += gridDim.x * blockDim.x) {
    const int2 q = pairs[k];
    parent[q.x] = q.x;
    parent[q.y] = q.y;
  }
}

// Union: the larger root is hooked under the smaller one by compare-and-
// swap.  A failed swap means another hook of that root succeeded; the
// roots are looked up again.  Each group ends with its smallest slot as
// root.
__global__ void union_kernel(const int2* __restrict__ pairs, const int* __restrict__ count,
                             int* parent, const int* __restrict__ flags) {
  if (*flags) return;
  const int c = count[0];
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < c; k += gridDim.x * blockDim.x) {
    const int2 q = pairs[k];
    int a = find_root(parent, q.x), b = find_root(parent, q.y);
    while (a != b) {
      const int lo = min(a, b), hi = max(a, b);
      if (atomicCAS(parent + hi, hi, lo) == hi) break;
      a = find_root(parent, lo);
      b = find_root(parent, hi);
    }
  }
}

// Lists each endpoint once (the first pair to mark it): key = group root
// in the high word, caller id in the low word; value = slot.
__global__ void members_kernel(const int2* __restrict__ pairs, int* count,
                               const int* parent, int* __restrict__ mark,
                               const unsigned* __restrict__ id,
                               unsigned long long* __restrict__ key, unsigned* __restrict__ val,
                               const int* __restrict__ flags) {
  if (*flags) return;
  const int c = count[0];
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < c; k += gridDim.x * blockDim.x) {
    const int2 q = pairs[k];
#pragma unroll
    for (int t = 0; t < 2; t++) {
      const int e = t ? q.y : q.x;
      if (atomicExch(mark + e, 1) == 0) {
        const int r = find_root(parent, e);
        const int slot = atomicAdd(count + 1, 1);
        key[slot] = ((unsigned long long)(unsigned)r << 32) | id[e];
        val[slot] = (unsigned)e;
      }
    }
  }
}

// One thread per group (its first entry in sorted order): collide folded
// over the members in increasing caller id.
__global__ void fold_kernel(const unsigned long long* __restrict__ key,
                            const unsigned* __restrict__ val, const int* __restrict__ count,
                            State st, double* __restrict__ ax, double* __restrict__ ay,
                            double* __restrict__ az, int* __restrict__ alive,
                            const int* __restrict__ flags) {
#pragma clang fp contract(off)
  if (*flags) return;
  const int mc = count[1];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= mc) return;
  const unsigned g = (unsigned)(key[i] >> 32);
  if (i > 0 && (unsigned)(key[i - 1] >> 32) == g) return;
  const int s0 = (int)val[i];
  double m = st.m[s0];
  double x = st.x[s0], y = st.y[s0], z = st.z[s0];
  double vx = st.vx[s0], vy = st.vy[s0], vz = st.vz[s0];
  for (int j = i + 1; j < mc && (unsigned)(key[j] >> 32) == g; j++) {
    const int s = (int)val[j];
    const double mb = st.m[s];
    const double w = m + mb;
    x = (x * m + st.x[s] * mb) / w;
    y = (y * m + st.y[s] * mb) / w;
    z = (z * m + st.z[s] * mb) / w;
    vx = (vx * m + st.vx[s] * mb) / w;
    vy = (vy * m + st.vy[s] * mb) / w;
    vz = (vz * m + st.vz[s] * mb) / w;
    m = w;
    alive[s] = 0;
  }
  st.x[s0] = x;
  st.y[s0] = y;
  st.z[s0] = z;
  st.vx[s0] = vx;
  st.vy[s0] = vy;
  st.vz[s0] = vz;
  st.m[s0] = m;
  ax[s0] = 0.0;
  ay[s0] = 0.0;
  az[s0] = 0.0;
}

__global__ void id_alive_kernel(const int* __restrict__ alive, const unsigned* __restrict__ id,
                                int n, int* __restrict__ id_alive) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < n) id_alive[id[s]] = alive[s];
}

__global__ void compact_kernel(State src, State dst, const int* __restrict__ alive,
                               const int* __restrict__ newpos, const int* __restrict__ newid,
                               const double* __restrict__ ax, const double* __restrict__ ay,
                               const double* __restrict__ az, const int* __restrict__ inter,
                               double* __restrict__ cax, double* __restrict__ cay,
                               double* __restrict__ caz, int* __restrict__ cinter, int n,
                               const int* __restrict__ flags) {
  if (*flags) return;
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n || !alive[s]) return;
  const int d = newpos[s];
  dst.x[d] = src.x[s];
  dst.y[d] = src.y[s];
  dst.z[d] = src.z[s];
  dst.vx[d] = src.vx[s];
  dst.vy[d] = src.vy[s];
  dst.vz[d] = src.vz[s];
  dst.m[d] = src.m[s];
  dst.id[d] = (unsigned)newid[src.id[s]];
  cax[d] = ax[s];
  cay[d] = ay[s];
  caz[d] = az[s];
  cinter[d] = inter[s];
}

// info[0] = particles left, info[1 + k] = new start of system k (k <=
// nsys); the system starts and simulated times are updated in place.
__global__ void finish_kernel(const int* __restrict__ newpos, int* __restrict__ start, int nsys,
                              const double* __restrict__ sys_dt, double* __restrict__ sys_time,
                              int advance, int* __restrict__ info, const int* __restrict__ flags) {
  if (*flags) return;
  for (int k = threadIdx.x; k <= nsys; k += blockDim.x) {
    const int ns = newpos[start[k]];
    start[k] = ns;
    info[1 + k] = ns;
    if (k == nsys) info[0] = ns;
    if (k < nsys && advance) sys_time[k] += sys_dt[k];
  }
}

// ------------------------------------------------------- minimum size --
// apply_minimum_size (src/main.cpp:1001-1057): after the build, a cell of
// side < min_size takes the particles of its leaf children -- merged by
// collide() in child order 0..7 when there are several -- and, post-order,
// that repeats up to the largest cells below min_size.  So every maximal
// cell of side < min_size ends as one particle: the left fold of collide
// over its children's results, each child's result the same fold over its
// own children (a leaf's result is its particle).  One thread per such
// cell (an internal node with side < min_size whose parent, of side 2
// side, is not, or a system root) folds its pre-order subtree with a stack
// of partial results, one entry per open cell.  The merged particle takes
// the slot of the member with the lowest caller id; the others are marked
// absorbed.  A depth-42 bucket's particles count as its children in slot
// order (the reference would keep splitting).
struct Part {
  double m, x, y, z, vx, vy, vz;
};

__device__ __forceinline__ void collide_into(Part& a, const Part& b) {
#pragma clang fp contract(off)
  const double w = a.m + b.m;  // src/main.cpp:423-430
  a.x = (a.x * a.m + b.x * b.m) / w;
  a.y = (a.y * a.m + b.y * b.m) / w;
  a.z = (a.z * a.m + b.z * b.m) / w;
  a.vx = (a.vx * a.m + b.vx * b.m) / w;
  a.vy = (a.vy * a.m + b.vy * b.m) / w;
  a.vz = (a.vz * a.m + b.vz * b.m) / w;
  a.m = w;
}

constexpr int kMinSizeStack = 48;  // open cells below a group root (depth <= 42)

__global__ void min_size_kernel(const WalkNode* __restrict__ nodes,
                                const int* __restrict__ node_start, int nsys, double min_size,
                                State st, int* __restrict__ alive, int* __restrict__ flags) {
  if (*flags) return;
  const int total = node_start[nsys];
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= total) return;
  const WalkNode nd = nodes[g];
  if (nd.leaf >= 0 || !(nd.side < min_size)) return;
  if (2.0 * nd.side < min_size) {
    // the parent is below min_size too: only a system root (no parent)
    // starts a group here
    int lo = 0, hi = nsys;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (node_start[mid] < g) lo = mid + 1; else hi = mid;
    }
    if (lo >= nsys || node_start[lo] != g) return;
  }
  Part acc[kMinSizeStack];
  int end[kMinSizeStack];
  bool has[kMinSizeStack];
  int top = 0;
  end[0] = nd.skip;
  has[0] = false;
  unsigned best_id = 0xffffffffu;
  int best = -1;
  for (int k = g + 1; k < nd.skip; k++) {
    while (top > 0 && end[top] <= k) {  // close finished cells into their parents
      if (has[top]) {
        if (has[top - 1]) collide_into(acc[top - 1], acc[top]); else acc[top - 1] = acc[top];
        has[top - 1] = true;
      }
      top--;
    }
    const WalkNode c = nodes[k];
    if (c.leaf >= 0) {
      const int sl = c.leaf;
      const Part q = {st.m[sl], st.x[sl], st.y[sl], st.z[sl], st.vx[sl], st.vy[sl], st.vz[sl]};
      if (has[top]) collide_into(acc[top], q); else acc[top] = q;
      has[top] = true;
      alive[sl] = 0;
      if (st.id[sl] < best_id) {
        best_id = st.id[sl];
        best = sl;
      }
    } else if (top + 1 < kMinSizeStack) {
      top++;
      end[top] = c.skip;
      has[top] = false;
    } else {
      atomicOr(flags, kFlagStack);
      return;
    }
  }
  while (top > 0) {
    if (has[top]) {
      if (has[top - 1]) collide_into(acc[top - 1], acc[top]); else acc[top - 1] = acc[top];
      has[top - 1] = true;
    }
    top--;
  }
  if (best < 0) return;
  const Part r = acc[0];
  st.m[best] = r.m;
  st.x[best] = r.x;
  st.y[best] = r.y;
  st.z[best] = r.z;
  st.vx[best] = r.vx;
  st.vy[best] = r.vy;
  st.vz[best] = r.vz;
  alive[best] = 1;
}

}  // namespace

size_t collide_sort_tmp_ints(int pair_cap) { return radix_tmp_ints(2 * (size_t)pair_cap); }

void launch_collide(CollideBuffers& cb, State& cur, State& nxt, double* ax, double* ay,
                    double* az, const int* inter, int n, Systems& sys, double dt, int* flags,
                    hipStream_t s) {
  const int L = 2 * cb.pair_cap;
  launch_fill_int(cb.mark, 0, n, s);
  launch_fill_int(cb.alive, 1, n, s);
  launch_fill_u64(cb.key0, ~0ull, L, s);
  init_kernel<<<grid_for(cb.pair_cap), kBlock, 0, s>>>(cb.pairs, cb.count, cb.pair_cap,
                                                       cb.parent, sys.dt, sys.nsys, dt, flags);
  union_kernel<<<grid_for(cb.pair_cap), kBlock, 0, s>>>(cb.pairs, cb.count, cb.parent, flags);
  members_kernel<<<grid_for(cb.pair_cap), kBlock, 0, s>>>(cb.pairs, cb.count, cb.parent, cb.mark,
                                                          cur.id, cb.key0, cb.val0, flags);
  const int which = radix_sort_pairs(cb.key0, cb.val0, cb.key1, cb.val1, L, 64, cb.sort_tmp, s);
  const unsigned long long* key = which ? cb.key1 : cb.key0;
  const unsigned* val = which ? cb.val1 : cb.val0;
  fold_kernel<<<(L + kBlock - 1) / kBlock, kBlock, 0, s>>>(key, val, cb.count, cur, ax, ay, az,
                                                          cb.alive, flags);
  launch_update(cur, ax, ay, az, n, sys, flags, s);
  launch_compact(cb, cur, nxt, ax, ay, az, inter, n, sys, 1, flags, s);
}

void launch_compact(CollideBuffers& cb, State& cur, State& nxt, const double* ax,
                    const double* ay, const double* az, const int* inter, int n, Systems& sys,
                    int advance_time, int* flags, hipStream_t s) {
  exclusive_scan(cb.alive, cb.newpos, n, cb.scan_tmp, s);
  id_alive_kernel<<<(n + kBlock - 1) / kBlock, kBlock, 0, s>>>(cb.alive, cur.id, n, cb.id_alive);
  exclusive_scan(cb.id_alive, cb.newid, n, cb.scan_tmp, s);
  compact_kernel<<<(n + kBlock - 1) / kBlock, kBlock, 0, s>>>(
      cur, nxt, cb.alive, cb.newpos, cb.newid, ax, ay, az, inter, cb.cax, cb.cay, cb.caz,
      cb.cinter, n, flags);
  finish_kernel<<<1, kBlock, 0, s>>>(cb.newpos, sys.start, sys.nsys, sys.dt, sys.time,
                                     advance_time, cb.info, flags);
}

void launch_min_size(CollideBuffers& cb, const WalkNode* nodes, int node_cap, double min_size,
                     State& cur, State& nxt, const double* ax, const double* ay,
                     const double* az, const int* inter, int n, Systems& sys, int* flags,
                     hipStream_t s) {
  launch_fill_int(cb.alive, 1, n, s);
  min_size_kernel<<<(node_cap + kBlock - 1) / kBlock, kBlock, 0, s>>>(
      nodes, sys.node_start, sys.nsys, min_size, cur, cb.alive, flags);
  launch_compact(cb, cur, nxt, ax, ay, az, inter, n, sys, 0, flags, s);
}

}  // namespace nbs
