// tree_threaded.hip -- the octree the reference builds when threads > 1
// (gen_root_threaded / gen_root_thread, src/main.cpp:817-886): the root
// allocates all eight children first (octree::allocate_child,
// src/octree.cpp:54-85), then child i receives every particle its
// octree::inside() accepts (src/octree.cpp:320-336).  inside() is
// inclusive, so a particle within rounding of the root's centre planes
// has a leaf in every child whose closed cell holds it (up to eight), one
// outside every child has none, and an octant nobody occupies is a
// childless cell of mass 0 whose COM is (0, 0, 0) (calc_com, :259-262).
//
// The build runs the serial build's kernels over tree *entries* instead of
// particles: entry (particle p, child o) for every o in p's inside() mask,
// keyed by o at level 0 and by child o's rounded centre chain below it
// (add_particle inside child o), plus one placeholder entry per empty
// octant.  After the sort:
//   * the particles are permuted into the order of their first entries
//     (then the particles outside every child), so walk waves stay
//     Morton-consecutive and every particle keeps one slot;
//   * the emission runs over the entries; each leaf is pointed at its
//     particle's slot (two leaves of a duplicated particle point at the
//     same slot, so its own walk skips both and every other particle
//     interacts with both, as in the reference), and each placeholder leaf
//     becomes the empty child cell.
// The host reads the entry count back once per build (so these steps are
// not graph-captured).
#include <cmath>

#include "nbs_device_util.h"
#include "nbs_internal.h"

namespace nbs {
namespace {

constexpr int kBlock = 256;

inline int blocks_for(size_t n, int b) { return (int)((n + b - 1) / b); }

// Child o of the root: centre = root centre + (+-1 * half) / 2 per axis,
// side = half = root side / 2, as allocate_child forms them.
__device__ __forceinline__ void child_cell(const double* root, int o, double* c, double* half) {
#pragma clang fp contract(off)
  const double h = root[3] / 2;
  const double q = h / 2;
  c[0] = root[0] + ((o & 1) ? q : -q);
  c[1] = root[1] + ((o & 2) ? q : -q);
  c[2] = root[2] + ((o & 4) ? q : -q);
  *half = h;
}

// inside() mask of every particle over the eight children, its popcount,
// whether it is outside every child, and each block's OR of the masks (the
// occupied octants).
__global__ void __launch_bounds__(kBlock)
    member_kernel(const double* __restrict__ x, const double* __restrict__ y,
                  const double* __restrict__ z, int n, const double* __restrict__ root,
                  int* __restrict__ mask, int* __restrict__ cnt, int* __restrict__ outside,
                  int* __restrict__ occ_part) {
#pragma clang fp contract(off)
  __shared__ int occ;
  if (threadIdx.x == 0) occ = 0;
  __syncthreads();
  const int i = blockIdx.x * kBlock + threadIdx.x;
  int mk = 0;
  if (i < n) {
    const double p[3] = {x[i], y[i], z[i]};
    for (int o = 0; o < 8; o++) {
      double c[3], side;
      child_cell(root, o, c, &side);
      const double hi = side / 2.0, lo = (-1.0 * side) / 2.0;
      bool in = true;
      for (int k = 0; k < 3; k++) {
        const double d = p[k] - c[k];
        if (d > hi || d < lo) in = false;
      }
      if (in) mk |= 1 << o;
    }
    mask[i] = mk;
    cnt[i] = __popc(mk);
    outside[i] = mk == 0;
  }
  int wo = mk;
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) wo |= __shfl_xor(wo, d, 64);
  if (lane_id() == 0 && wo) atomicOr(&occ, wo);
  __syncthreads();
  if (threadIdx.x == 0) occ_part[blockIdx.x] = occ;
}

// Keys of the entries: entry eoff[i] + r is particle i in the r-th child
// of its mask (child order), keyed by that child at level 0 and by the
// child's centre chain below (keys_kernel's arithmetic); entries
// e_real + k are the placeholders of the empty octants empty[k].
__global__ void __launch_bounds__(kBlock)
    entry_keys_kernel(const double* __restrict__ x, const double* __restrict__ y,
                      const double* __restrict__ z, const unsigned* __restrict__ id, int n,
                      const double* __restrict__ root, const int* __restrict__ mask,
                      const int* __restrict__ eoff, int e_real, int4 empty_lo,
                      int4 empty_hi, int n_empty, int prim_levels,
                      unsigned long long* __restrict__ prim, unsigned* __restrict__ val,
                      unsigned long long* __restrict__ khi, unsigned long long* __restrict__ klo,
                      int* __restrict__ epart, unsigned* __restrict__ eid) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * kBlock + threadIdx.x;
  const int pshift = 63 - 3 * prim_levels;
  if (i < n_empty) {
    const int o = i < 4 ? (i == 0 ? empty_lo.x : i == 1 ? empty_lo.y : i == 2 ? empty_lo.z
                                                                              : empty_lo.w)
                        : (i == 4 ? empty_hi.x : i == 5 ? empty_hi.y : i == 6 ? empty_hi.z
                                                                              : empty_hi.w);
    const int e = e_real + i;
    const unsigned long long hi = (unsigned long long)o << 60;
    khi[e] = hi;
    klo[e] = 0;
    prim[e] = hi >> pshift;
    val[e] = (unsigned)e;
    epart[e] = -1;
    eid[e] = 0xffffffffu;
  }
  if (i >= n) return;
  const int mk = mask[i];
  if (!mk) return;
  const double px = x[i], py = y[i], pz = z[i];
  const unsigned pid = id[i];
  int e = eoff[i];
  for (int o = 0; o < 8; o++) {
    if (!((mk >> o) & 1)) continue;
    double c[3], sd;
    child_cell(root, o, c, &sd);
    double cx = c[0], cy = c[1], cz = c[2];
    unsigned long long hi = (unsigned long long)o << 60, lo = 0;
    for (int d = 1; d < kLevels; d++) {
      const bool bx = px > cx, by = py > cy, bz = pz > cz;
      const unsigned q3 = (unsigned)bx | ((unsigned)by << 1) | ((unsigned)bz << 2);
      const double half = sd * 0.5;
      const double q = half * 0.5;
      cx = cx + (bx ? q : -q);
      cy = cy + (by ? q : -q);
      cz = cz + (bz ? q : -q);
      sd = half;
      if (d < 21)
        hi |= (unsigned long long)q3 << (60 - 3 * d);
      else
        lo |= (unsigned long long)q3 << (60 - 3 * (d - 21));
    }
    khi[e] = hi;
    klo[e] = lo;
    prim[e] = hi >> pshift;
    val[e] = (unsigned)e;
    epart[e] = i;
    eid[e] = pid;
    e++;
  }
}

// Sorted entry j: its particle, that particle's position and mass (the
// leaf data of the emission; placeholders: mass 0 at the origin) and
// whether it is the particle's first entry.
__global__ void __launch_bounds__(kBlock)
    entry_gather_kernel(const unsigned* __restrict__ perm, const int* __restrict__ epart,
                        const int* __restrict__ eoff, State cur, int ne,
                        double* __restrict__ ex, double* __restrict__ ey,
                        double* __restrict__ ez, double* __restrict__ em,
                        int* __restrict__ sp, int* __restrict__ first) {
  const int j = blockIdx.x * kBlock + threadIdx.x;
  if (j >= ne) return;
  const unsigned e = perm[j];
  const int p = e < (unsigned)ne ? epart[e] : -1;
  sp[j] = p;
  if (p >= 0) {
    ex[j] = cur.x[p];
    ey[j] = cur.y[p];
    ez[j] = cur.z[p];
    em[j] = cur.m[p];
    first[j] = (int)e == eoff[p] ? 1 : 0;
  } else {
    ex[j] = ey[j] = ez[j] = 0.0;
    em[j] = 0.0;
    first[j] = 0;
  }
}

// New slots: the particles in the order of their first entries
// (pslot = rank among first entries), then those with no entry in caller
// slot order (outside[i] = 1 for them; orank its exclusive scan).  gperm is
// the gather permutation of the particle state.
__global__ void __launch_bounds__(kBlock)
    slots_first_kernel(const int* __restrict__ sp, const int* __restrict__ first,
                       const int* __restrict__ pslot, int ne, int* __restrict__ new_slot,
                       unsigned* __restrict__ gperm) {
  const int j = blockIdx.x * kBlock + threadIdx.x;
  if (j >= ne || !first[j]) return;
  const int s = pslot[j];
  new_slot[sp[j]] = s;
  gperm[s] = (unsigned)sp[j];
}

__global__ void __launch_bounds__(kBlock)
    slots_outside_kernel(const int* __restrict__ cnt, const int* __restrict__ orank,
                         const int* __restrict__ pslot, int ne, int n,
                         int* __restrict__ new_slot, unsigned* __restrict__ gperm) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n || cnt[i] != 0) return;
  const int s = pslot[ne] + orank[i];
  new_slot[i] = s;
  gperm[s] = (unsigned)i;
}

// After the emission over the entries: each leaf takes its particle's new
// slot; a placeholder's leaf becomes the empty child cell (mass 0, COM at
// the origin, no children, opened where d^2 < (side / theta)^2 like any
// cell).  The emission counted every entry as a leaf: the placeholders are
// cells.
__global__ void __launch_bounds__(kBlock)
    entry_patch_kernel(const int* __restrict__ offsets, const int* __restrict__ sp,
                       const int* __restrict__ new_slot, int ne, double theta,
                       int n_empty, WalkNode* __restrict__ nodes,
                       WalkNode32* __restrict__ walk32, int* __restrict__ stats,
                       const int* __restrict__ flags) {
  const int j = blockIdx.x * kBlock + threadIdx.x;
  if (j >= ne || *flags) return;
  const int leafn = offsets[j + 1] - 1;
  const int p = sp[j];
  if (p >= 0) {
    nodes[leafn].leaf = new_slot[p];
    if (walk32) walk32[leafn].link = kLink32Leaf | (unsigned)new_slot[p];
  } else {
    WalkNode w = nodes[leafn];
    w.cx = w.cy = w.cz = 0.0;
    w.m = w.gm = 0.0;
    w.gm32 = 0.0f;
    const double ro = w.side / theta;
    w.r2open32 = (float)(ro * ro);
    w.leaf = -1;
    nodes[leafn] = w;
    if (walk32) {  // a root child: depth 1, no children (skip = next)
      WalkNode32 v;
      v.cx = v.cy = v.cz = 0.0;
      v.link = link32_cell(1, w.skip);
      v.gm32 = 0.0f;
      walk32[leafn] = v;
    }
  }
  if (j == 0) stats[1] += n_empty;
}

}  // namespace

void launch_member(const State& st, int n, const double* root, int* mask, int* cnt,
                   int* outside, int* occ_part, hipStream_t s) {
  if (n > 0)
    member_kernel<<<blocks_for(n, kBlock), kBlock, 0, s>>>(st.x, st.y, st.z, n, root, mask,
                                                           cnt, outside, occ_part);
}

int member_blocks(int n) { return blocks_for(n, kBlock); }

void launch_entry_keys(const State& st, int n, const double* root, const int* mask,
                       const int* eoff, int e_real, const int* empty, int n_empty,
                       int prim_levels, BuildBuffers& b, int* epart, unsigned* eid,
                       hipStream_t s) {
  const int4 lo = make_int4(empty[0], empty[1], empty[2], empty[3]);
  const int4 hi = make_int4(empty[4], empty[5], empty[6], empty[7]);
  const int threads = n > n_empty ? n : n_empty;
  if (threads > 0)
    entry_keys_kernel<<<blocks_for(threads, kBlock), kBlock, 0, s>>>(
        st.x, st.y, st.z, st.id, n, root, mask, eoff, e_real, lo, hi, n_empty, prim_levels,
        b.prim0, b.val0, b.khi_u, b.klo_u, epart, eid);
}

void launch_entry_gather(const unsigned* perm, const int* epart, const int* eoff,
                         const State& cur, int ne, double* ex, double* ey, double* ez,
                         double* em, int* sp, int* first, hipStream_t s) {
  if (ne > 0)
    entry_gather_kernel<<<blocks_for(ne, kBlock), kBlock, 0, s>>>(perm, epart, eoff, cur, ne,
                                                                  ex, ey, ez, em, sp, first);
}

void launch_entry_slots(const int* sp, const int* first, const int* pslot, const int* cnt,
                        const int* orank, int ne, int n, int* new_slot, unsigned* gperm,
                        hipStream_t s) {
  if (ne > 0)
    slots_first_kernel<<<blocks_for(ne, kBlock), kBlock, 0, s>>>(sp, first, pslot, ne,
                                                                 new_slot, gperm);
  if (n > 0)
    slots_outside_kernel<<<blocks_for(n, kBlock), kBlock, 0, s>>>(cnt, orank, pslot, ne, n,
                                                                  new_slot, gperm);
}

void launch_entry_patch(const int* offsets, const int* sp, const int* new_slot, int ne,
                        double theta, int n_empty, WalkNode* nodes, WalkNode32* walk32,
                        int* stats, const int* flags, hipStream_t s) {
  if (ne > 0)
    entry_patch_kernel<<<blocks_for(ne, kBlock), kBlock, 0, s>>>(
        offsets, sp, new_slot, ne, theta, n_empty, nodes, walk32, stats, flags);
}

}  // namespace nbs
