// tree_build.hip -- Morton-keyed octree build on the GPU that reproduces
// the reference's pointer octree node for node.
//
// Reference tree (src/octree.cpp:54-196): one particle per leaf; a cell
// splits exactly when it holds two or more particles; child index
// (x>cx) + 2(y>cy) + 4(z>cz); child centre = centre + (+-1 * half) / 2
// with half = side / 2, every addition rounded in the datatype.  That tree
// is canonical (independent of insertion order), so it is fully described
// by each particle's path of child indices computed with the same rounded
// centre chain (keys_kernel), sorted (radix_sort.hip), plus the lengths of
// the common prefixes of sorted neighbours:
//   * l_i = levels shared by keys i and i+1 (-1 across a system boundary);
//   * internal cells starting at slot i have depths (l_{i-1}, l_i];
//   * particle i is a leaf at depth 1 + max(l_{i-1}, l_i).
// Cells are emitted in DFS pre-order (children 0..7), the order of the
// reference's recursive calc_mass/calc_com, so a cell's subtree is the run
// up to its `skip`.  Mass and COM are then reduced bottom-up, each cell
// summing its children in child order with the reference's operation
// sequence (src/octree.cpp:203-221, :252-278), so they match bit for bit.
// The reduction runs one kernel per tree level from the deepest up
// (internal cells are bucketed by depth during emission), so a parent reads
// its children after a kernel boundary and no device-wide fences are
// needed.
#include <cmath>

#include "nbs_device_util.h"
#include "nbs_internal.h"

namespace nbs {
namespace {

constexpr int kBlock = 256;

// Common prefix of two 42-level keys, in levels (0..42).
__device__ __forceinline__ int cpl(unsigned long long ahi,
                                   unsigned long long alo,
                                   unsigned long long bhi,
                                   unsigned long long blo) {
  unsigned long long x = ahi ^ bhi;
  if (x) return (__clzll(x) - 1) / 3;
  x = alo ^ blo;
  if (x) return 21 + (__clzll(x) - 1) / 3;
  return kLevels;
}

__global__ void fill_int_kernel(int* p, int v, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void fill_u64_kernel(unsigned long long* p, unsigned long long v,
                                size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Bounding-box partials: every wave writes the min/max of its particles to
// part[k * nwaves + w] (k = min x, y, z, max x, y, z); root_kernel reduces a
// system's waves.  No atomics, so no contention on a few addresses.
__global__ void __launch_bounds__(kBlock)
    bbox_kernel(const double* __restrict__ x, const double* __restrict__ y,
                const double* __restrict__ z, const int* __restrict__ start,
                const int* __restrict__ wstart, int nsys, int nwaves,
                double* __restrict__ part) {
  const int w = (blockIdx.x * kBlock + threadIdx.x) >> 6;
  if (w >= nwaves) return;
  const int s = nsys > 1 ? sys_of_wave(w, wstart, nsys) : 0;
  const int i = start[s] + (w - wstart[s]) * 64 + lane_id();
  const bool v = i < start[s + 1];
  const double px = v ? x[i] : 0, py = v ? y[i] : 0, pz = v ? z[i] : 0;
  wave_bbox_partial(v, px, py, pz, w, nwaves, part);
}

// Root cell per system: src/main.cpp:908-920 in FP64 (midpoint of the
// bbox; side = largest extent + 2.0, ties falling through to z).  One block
// per system reduces the system's wave partials (1024 threads, four
// partials in flight per thread: the reduction is latency-bound).
constexpr int kRootBlock = 1024;
__global__ void __launch_bounds__(kRootBlock)
    root_kernel(const double* __restrict__ part, int nwaves,
                const int* __restrict__ wstart, double* __restrict__ root,
                int nsys, int mode, double cx, double cy, double cz,
                double side, const int* __restrict__ flags) {
#pragma clang fp contract(off)
  const int s = blockIdx.x;
  if (s >= nsys || *flags) return;
  __shared__ double red[6][kRootBlock / 64];
  double b[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
  if (mode != 1) {
#pragma unroll 4
    for (int w = wstart[s] + threadIdx.x; w < wstart[s + 1]; w += kRootBlock)
#pragma unroll
      for (int k = 0; k < 6; k++) {
        const double v = part[(size_t)k * nwaves + w];
        b[k] = k < 3 ? fmin(b[k], v) : fmax(b[k], v);
      }
    for (int k = 0; k < 6; k++) {
      b[k] = k < 3 ? wave_min(b[k]) : wave_max(b[k]);
      if (lane_id() == 0) red[k][threadIdx.x >> 6] = b[k];
    }
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  if (mode == 1) {
    root[4 * s] = cx;
    root[4 * s + 1] = cy;
    root[4 * s + 2] = cz;
    root[4 * s + 3] = side;
    return;
  }
  for (int k = 0; k < 6; k++) {
    double r = red[k][0];
    for (int q = 1; q < kRootBlock / 64; q++) r = k < 3 ? fmin(r, red[k][q]) : fmax(r, red[k][q]);
    b[k] = r;
  }
  const double min_x = b[0], min_y = b[1], min_z = b[2];
  const double max_x = b[3], max_y = b[4], max_z = b[5];
  root[4 * s] = (min_x + max_x) / 2.0;
  root[4 * s + 1] = (min_y + max_y) / 2.0;
  root[4 * s + 2] = (min_z + max_z) / 2.0;
  const double ex = max_x - min_x, ey = max_y - min_y, ez = max_z - min_z;
  double sz;
  if (ex > ey && ex > ez)
    sz = ex + 2.0;
  else if (ey > ex && ey > ez)
    sz = ey + 2.0;
  else
    sz = ez + 2.0;
  root[4 * s + 3] = sz;
}

// 42-level keys by the reference's rounded centre chain.  The primary sort
// key is the key's top levels behind the system index (batched mode).
__global__ void __launch_bounds__(kBlock)
    keys_kernel(const double* __restrict__ x, const double* __restrict__ y,
                const double* __restrict__ z, const int* __restrict__ start,
                const int* __restrict__ wstart, int nsys, int nwaves,
                const double* __restrict__ root, int sys_shift,
                int prim_levels, unsigned long long* __restrict__ prim,
                unsigned* __restrict__ val,
                unsigned long long* __restrict__ khi,
                unsigned long long* __restrict__ klo) {
#pragma clang fp contract(off)
  const int w = (blockIdx.x * kBlock + threadIdx.x) >> 6;
  if (w >= nwaves) return;
  const int s = nsys > 1 ? sys_of_wave(w, wstart, nsys) : 0;
  const int i = start[s] + (w - wstart[s]) * 64 + lane_id();
  if (i >= start[s + 1]) return;
  const double px = x[i], py = y[i], pz = z[i];
  double cx = root[4 * s], cy = root[4 * s + 1], cz = root[4 * s + 2];
  double sd = root[4 * s + 3];
  unsigned long long hi = 0, lo = 0;
#pragma unroll
  for (int d = 0; d < kLevels; d++) {
    const bool bx = px > cx, by = py > cy, bz = pz > cz;
    const unsigned o = (unsigned)bx | ((unsigned)by << 1) | ((unsigned)bz << 2);
    const double half = sd * 0.5;  // side / 2 (exact)
    const double q = half * 0.5;   // (+-1 * half) / 2 (exact)
    cx = cx + (bx ? q : -q);
    cy = cy + (by ? q : -q);
    cz = cz + (bz ? q : -q);
    sd = half;
    if (d < 21)
      hi |= (unsigned long long)o << (60 - 3 * d);
    else
      lo |= (unsigned long long)o << (60 - 3 * (d - 21));
  }
  khi[i] = hi;
  klo[i] = lo;
  // primary sort key: system, then the first prim_levels levels (runs of
  // equal primary keys are finished by fixup_kernel on the full key)
  prim[i] = ((unsigned long long)s << sys_shift) | (hi >> (63 - 3 * prim_levels));
  val[i] = (unsigned)i;
}

__global__ void gather_keys_kernel(const unsigned* __restrict__ perm, int n,
                                   const unsigned long long* __restrict__ hu,
                                   const unsigned long long* __restrict__ lu,
                                   unsigned long long* __restrict__ h,
                                   unsigned long long* __restrict__ l) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const unsigned p = perm[j];
  if (p >= (unsigned)n) return;
  h[j] = hu[p];
  l[j] = lu[p];
}

// Runs of equal primary keys (particles closer than the primary key's
// resolution, common only while a cold system collapses) are sorted by
// (full key, caller index).  One thread per run start: a run of up to
// kShortRun entries is sorted by that thread in registers; a longer one is
// listed, runs of up to kWaveRun entries in one list (short_run_kernel, one
// wavefront each) and longer ones in another (long_run_kernel, one
// workgroup each).
// (A cold 4M-particle shell falling through its centre goes from ~10^6
// runs of 2-8 entries through ~10^5 of 9-64 to ~10^4 of 65-1024 and a
// thousand longer ones.)  Layout of `runs`: [0] entries of the short list,
// [1] long_run_kernel's queue counter over the long list, [2] entries of
// the long list, [3] unused; the short list's (start, length) pairs from
// [4], the long list's from [big].  Appends are aggregated per workgroup
// (append_runs).
constexpr int kShortRun = 8;
constexpr int kWaveRun = 256;

// Appends go through the workgroup: per-wave counts in LDS, one
// device-scope atomic per workgroup and list, offsets back through LDS.
// (One atomic per wavefront on the two shared counters cost ~0.6 ms per
// build in a 4M-particle collapse, where nearly every wavefront holds a
// run: atomics on one word serialise across the XCDs.)  Every thread of
// the workgroup calls it.
constexpr int kFixBlock = 1024;
constexpr int kFixWaves = kFixBlock / kWave;

__device__ __forceinline__ void append_runs(bool to_short, bool to_long, int beg, int len,
                                            int* runs, int big) {
  __shared__ int sh[2][kFixWaves + 1];
  const unsigned long long m0 = ballot(to_short), m1 = ballot(to_long);
  const int w = threadIdx.x >> 6;
  if (lane_id() == 0) {
    sh[0][w] = __popcll(m0);
    sh[1][w] = __popcll(m1);
  }
  __syncthreads();
  if (threadIdx.x < 2) {  // list q: wave offsets, then the workgroup's base
    const int q = threadIdx.x;
    int t = 0;
    for (int k = 0; k < kFixWaves; k++) {
      const int c = sh[q][k];
      sh[q][k] = t;
      t += c;
    }
    sh[q][kFixWaves] = t ? atomicAdd(&runs[2 * q], t) : 0;
  }
  __syncthreads();
  if (to_short) {
    const int r = sh[0][kFixWaves] + sh[0][w] + mbcnt(m0);
    runs[4 + 2 * r] = beg;
    runs[4 + 2 * r + 1] = len;
  }
  if (to_long) {
    const int r = sh[1][kFixWaves] + sh[1][w] + mbcnt(m1);
    runs[big + 2 * r] = beg;
    runs[big + 2 * r + 1] = len;
  }
}

// (h, l, q) < (h', l', q') lexicographically: key_hi, key_lo, caller index.
__device__ __forceinline__ bool run_less(unsigned long long h, unsigned long long l, unsigned q,
                                         unsigned long long h2, unsigned long long l2,
                                         unsigned q2) {
  return h < h2 || (h == h2 && (l < l2 || (l == l2 && q < q2)));
}

// Compare-exchange of slots a < b of a run held in registers.
__device__ __forceinline__ void run_cx(unsigned long long* H, unsigned long long* L, unsigned* Q,
                                       unsigned* P, int a, int b) {
  if (run_less(H[b], L[b], Q[b], H[a], L[a], Q[a])) {
    const unsigned long long th = H[a], tl = L[a];
    const unsigned tq = Q[a], tp = P[a];
    H[a] = H[b]; L[a] = L[b]; Q[a] = Q[b]; P[a] = P[b];
    H[b] = th; L[b] = tl; Q[b] = tq; P[b] = tp;
  }
}

__global__ void __launch_bounds__(kFixBlock)
    fixup_kernel(const unsigned long long* __restrict__ prim, unsigned* __restrict__ perm,
                 unsigned long long* __restrict__ khi, unsigned long long* __restrict__ klo,
                 const unsigned* __restrict__ id, int n, int* __restrict__ runs, int big) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  // every thread reaches the appends below (workgroup barriers)
  int len = 0;
  if (j + 1 < n) {
    const unsigned long long pk = prim[j];
    if (pk == prim[j + 1] && !(j > 0 && prim[j - 1] == pk)) {
      // last index of the run: galloping, then bisection (prim is sorted)
      int good = j + 1, bad = n, step = 1;
      for (;;) {
        const int probe = good + step;
        if (probe >= n) break;
        if (prim[probe] == pk) {
          good = probe;
          step <<= 1;
        } else {
          bad = probe;
          break;
        }
      }
      while (bad - good > 1) {
        const int mid = good + ((bad - good) >> 1);
        if (prim[mid] == pk)
          good = mid;
        else
          bad = mid;
      }
      len = good - j + 1;
    }
  }
  append_runs(len > kShortRun && len <= kWaveRun, len > kWaveRun, j, len, runs, big);
  if (len < 2 || len > kShortRun) return;
  // The run in registers (its loads issued together, not one dependent
  // chain per insertion), eight slots, those past the run holding a key
  // above every real one (key_hi never has bit 63 set), sorted by a fixed
  // 19-comparator network of depth 6 and written back.  The key includes
  // the caller index, so the order is total and equals any other sort's.
  static_assert(kShortRun == 8, "the network sorts eight slots");
  unsigned long long H[kShortRun], L[kShortRun];
  unsigned Q[kShortRun], P[kShortRun];
#pragma unroll
  for (int k = 0; k < kShortRun; k++) {
    const bool in = k < len;
    H[k] = in ? khi[j + k] : ~0ull;
    L[k] = in ? klo[j + k] : ~0ull;
    P[k] = in ? perm[j + k] : 0u;
  }
#pragma unroll
  for (int k = 0; k < kShortRun; k++) Q[k] = k < len ? id[P[k]] : ~0u;
  constexpr int kNet[19][2] = {{0, 2}, {1, 3}, {4, 6}, {5, 7}, {0, 4}, {1, 5}, {2, 6},
                               {3, 7}, {0, 1}, {2, 3}, {4, 5}, {6, 7}, {2, 4}, {3, 5},
                               {1, 4}, {3, 6}, {1, 2}, {3, 4}, {5, 6}};
#pragma unroll
  for (int c = 0; c < 19; c++) run_cx(H, L, Q, P, kNet[c][0], kNet[c][1]);
#pragma unroll
  for (int k = 0; k < kShortRun; k++) {
    if (k < len) {
      khi[j + k] = H[k];
      klo[j + k] = L[k];
      perm[j + k] = P[k];
    }
  }
}

// One wavefront sorts a run of up to 64 * E entries by (key_hi, key_lo,
// caller index): a bitonic network over elements i = lane * E + e (E
// consecutive elements per lane), exchanging with shuffles where the
// partner sits in another lane (j >= E) and within the lane's registers
// otherwise.  The triple is unique, so the order is total;
// padding entries (all ones) sort behind the run.
template <int E>
__device__ __forceinline__ void wave_sort_run(int len, unsigned long long* __restrict__ h0,
                                              unsigned long long* __restrict__ l0,
                                              unsigned* __restrict__ p0,
                                              const unsigned* __restrict__ id) {
  const int lane = lane_id();
  unsigned long long h[E], l[E];
  unsigned q[E], p[E];
#pragma unroll
  for (int e = 0; e < E; e++) {
    const int i = lane * E + e;
    const bool v = i < len;
    p[e] = v ? p0[i] : 0u;
    h[e] = v ? h0[i] : ~0ull;
    l[e] = v ? l0[i] : ~0ull;
    q[e] = v ? id[p[e]] : ~0u;
  }
#pragma unroll
  for (int k = 2; k <= 64 * E; k <<= 1) {
#pragma unroll
    for (int j = k >> 1; j > 0; j >>= 1) {
      if (j >= E) {
#pragma unroll
        for (int e = 0; e < E; e++) {
          const int i = lane * E + e;
          const unsigned long long oh = __shfl_xor(h[e], j / E, 64);
          const unsigned long long ol = __shfl_xor(l[e], j / E, 64);
          const unsigned oq = __shfl_xor(q[e], j / E, 64);
          const unsigned op = __shfl_xor(p[e], j / E, 64);
          const bool want_min = ((i & j) == 0) == ((i & k) == 0);
          if (want_min == run_less(oh, ol, oq, h[e], l[e], q[e])) {
            h[e] = oh;
            l[e] = ol;
            q[e] = oq;
            p[e] = op;
          }
        }
      } else {
#pragma unroll
        for (int e = 0; e < E; e++) {
          if (e & j) continue;
          const int f = e | j;
          const bool asc = ((lane * E + e) & k) == 0;
          if (asc == run_less(h[f], l[f], q[f], h[e], l[e], q[e])) {
            const unsigned long long th = h[e], tl = l[e];
            const unsigned tq = q[e], tp = p[e];
            h[e] = h[f];
            l[e] = l[f];
            q[e] = q[f];
            p[e] = p[f];
            h[f] = th;
            l[f] = tl;
            q[f] = tq;
            p[f] = tp;
          }
        }
      }
    }
  }
#pragma unroll
  for (int e = 0; e < E; e++) {
    const int i = lane * E + e;
    if (i < len) {
      h0[i] = h[e];
      l0[i] = l[e];
      p0[i] = p[e];
    }
  }
}

// The short list's runs sorted by (key_hi, key_lo, caller index):
// wavefront g of the grid takes runs g, g + G, ... (G wavefronts in the
// grid) and sorts each in registers (wave_sort_run: no barrier, no
// atomic).  A kernel of its own: the unrolled register sort takes ~128
// VGPRs, which would cost the workgroup sorts below their occupancy.
constexpr int kShortBlock = 256;
constexpr int kShortGrid = 1024;
__global__ void __launch_bounds__(kShortBlock)
    short_run_kernel(const int* __restrict__ runs, unsigned long long* __restrict__ khi,
                     unsigned long long* __restrict__ klo, unsigned* __restrict__ perm,
                     const unsigned* __restrict__ id) {
  const int nshort = runs[0];
  const int waves = gridDim.x * (kShortBlock / 64);
  for (int r = (blockIdx.x * kShortBlock + threadIdx.x) / 64; r < nshort; r += waves) {
    const int beg = runs[4 + 2 * r], len = runs[5 + 2 * r];
    if (len <= 64)
      wave_sort_run<1>(len, khi + beg, klo + beg, perm + beg, id);
    else if (len <= 128)
      wave_sort_run<2>(len, khi + beg, klo + beg, perm + beg, id);
    else
      wave_sort_run<4>(len, khi + beg, klo + beg, perm + beg, id);
  }
}

// The long list's runs sorted by (key_hi, key_lo, caller index), one run
// per workgroup from a queue (runs[1]).  Up to kRunBitonic entries: a
// bitonic sort in LDS (the triple is unique, so stability does not arise).  Longer runs: stable LSD radix sort, 8-bit
// digits, least significant first (caller index bytes 0-3, key_lo bytes
// 0-7, key_hi bytes 0-7), skipping digits that are equal over the whole
// run; entries ping-pong between the run's range of (khi, klo, perm) and
// the same range of three free arrays.  Each chunk of 1024 entries is
// ranked stably with ballots (equal digits within a wave) and per-wave
// digit counts.
constexpr int kRunBlock = 1024;
constexpr int kRunWaves = kRunBlock / 64;
constexpr int kRunGrid = 1024;
constexpr int kRunBitonic = 1024;

__global__ void __launch_bounds__(kRunBlock)
    long_run_kernel(int* __restrict__ runs, unsigned long long* __restrict__ khi,
                    unsigned long long* __restrict__ klo, unsigned* __restrict__ perm,
                    unsigned long long* __restrict__ shi, unsigned long long* __restrict__ slo,
                    unsigned* __restrict__ sperm, const unsigned* __restrict__ id, int big) {
  __shared__ int hist[256];
  __shared__ int wcnt[kRunWaves][256];
  __shared__ int total[256];
  __shared__ int run_sh;
  __shared__ int skip_sh;
  __shared__ unsigned long long bh[kRunBitonic], bl[kRunBitonic];
  __shared__ unsigned bid[kRunBitonic], bp[kRunBitonic];
  const int t = threadIdx.x, lane = lane_id(), w = t >> 6;
  const int nlong = runs[2];
  // no long run (the usual case): leave without touching the shared queue
  // counter (1024 workgroups' atomics on one word cost ~10 us)
  if (nlong == 0) return;
  const int* lst = runs + big;
  for (;;) {
    if (t == 0) run_sh = atomicAdd(&runs[1], 1);
    __syncthreads();
    const int r = run_sh;
    __syncthreads();
    if (r >= nlong) return;
    const int beg = lst[2 * r], len = lst[2 * r + 1];
    unsigned long long *h0 = khi + beg, *l0 = klo + beg, *h1 = shi + beg, *l1 = slo + beg;
    unsigned *p0 = perm + beg, *p1 = sperm + beg;
    if (len <= kRunBitonic) {
      int P = 2;
      while (P < len) P <<= 1;
      for (int e = t; e < P; e += kRunBlock) {
        const bool v = e < len;
        const unsigned pe = v ? p0[e] : 0u;
        bh[e] = v ? h0[e] : ~0ull;
        bl[e] = v ? l0[e] : ~0ull;
        bid[e] = v ? id[pe] : ~0u;
        bp[e] = pe;
      }
      __syncthreads();
      for (int k = 2; k <= P; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
          for (int i = t; i < P; i += kRunBlock) {
            const int q = i ^ j;
            if (q > i) {
              const bool gt = bh[i] > bh[q] ||
                              (bh[i] == bh[q] && (bl[i] > bl[q] || (bl[i] == bl[q] && bid[i] > bid[q])));
              if (gt == ((i & k) == 0)) {  // ascending where bit k of i is clear
                const unsigned long long th = bh[i], tl = bl[i];
                const unsigned ti = bid[i], tp = bp[i];
                bh[i] = bh[q];
                bl[i] = bl[q];
                bid[i] = bid[q];
                bp[i] = bp[q];
                bh[q] = th;
                bl[q] = tl;
                bid[q] = ti;
                bp[q] = tp;
              }
            }
          }
          __syncthreads();
        }
      }
      for (int e = t; e < len; e += kRunBlock) {
        h0[e] = bh[e];
        l0[e] = bl[e];
        p0[e] = bp[e];
      }
      __syncthreads();
      continue;
    }
    for (int pass = 0; pass < 20; pass++) {
      // digit of entry e of the current source arrays
      auto digit = [&](int e) -> int {
        if (pass < 4) return (id[p0[e]] >> (8 * pass)) & 255;
        if (pass < 12) return (int)((l0[e] >> (8 * (pass - 4))) & 255);
        return (int)((h0[e] >> (8 * (pass - 12))) & 255);
      };
      if (t < 256) hist[t] = 0;
      if (t == 0) skip_sh = 0;
      __syncthreads();
      for (int e = t; e < len; e += kRunBlock) atomicAdd(&hist[digit(e)], 1);
      __syncthreads();
      if (t < 256 && hist[t] == len) skip_sh = 1;  // one digit value: order unchanged
      __syncthreads();
      // Every wave reads the flag before any wave can leave for the next
      // pass, whose first statements clear hist and skip_sh: without this
      // barrier a wave that reads late sees the next pass's 0 and falls
      // into this pass's scatter, out of step with the other waves'
      // barriers.
      const int skip = skip_sh;
      __syncthreads();
      if (skip) continue;
      if (w == 0) {  // exclusive scan of the 256 bins, 4 per lane
        int v[4], s4 = 0;
#pragma unroll
        for (int k = 0; k < 4; k++) {
          v[k] = hist[lane * 4 + k];
          s4 += v[k];
        }
        int run = wave_inclusive_scan(s4) - s4;
#pragma unroll
        for (int k = 0; k < 4; k++) {
          hist[lane * 4 + k] = run;
          run += v[k];
        }
      }
      __syncthreads();
      for (int c0 = 0; c0 < len; c0 += kRunBlock) {
        const int e = c0 + t;
        const bool v = e < len;
        const int d = v ? digit(e) : 0;
        unsigned long long same = ballot(v);
#pragma unroll
        for (int bit = 0; bit < 8; bit++) {
          const unsigned long long m = ballot(v && ((d >> bit) & 1));
          same &= ((d >> bit) & 1) ? m : ~m;
        }
        const int rank = mbcnt(same);
        for (int k = lane; k < 256; k += 64) wcnt[w][k] = 0;
        __syncthreads();
        if (v && rank == 0) wcnt[w][d] = __popcll(same);
        __syncthreads();
        if (t < 256) {  // prefix over the waves, per digit
          int acc = 0;
          for (int k = 0; k < kRunWaves; k++) {
            const int c = wcnt[k][t];
            wcnt[k][t] = acc;
            acc += c;
          }
          total[t] = acc;
        }
        __syncthreads();
        if (v) {
          const int dst = hist[d] + wcnt[w][d] + rank;
          h1[dst] = h0[e];
          l1[dst] = l0[e];
          p1[dst] = p0[e];
        }
        __syncthreads();
        if (t < 256) hist[t] += total[t];
        __syncthreads();
      }
      unsigned long long* th = h0; h0 = h1; h1 = th;
      unsigned long long* tl = l0; l0 = l1; l1 = tl;
      unsigned* tp = p0; p0 = p1; p1 = tp;
      __syncthreads();
    }
    if (h0 != khi + beg) {  // result in the free arrays: copy back
      for (int e = t; e < len; e += kRunBlock) {
        khi[beg + e] = h0[e];
        klo[beg + e] = l0[e];
        perm[beg + e] = p0[e];
      }
    }
    __syncthreads();
  }
}

__global__ void permute_kernel(State src, State dst,
                               const unsigned* __restrict__ perm, int n) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const unsigned p = perm[j];
  if (p >= (unsigned)n) return;
  dst.x[j] = src.x[p];
  dst.y[j] = src.y[p];
  dst.z[j] = src.z[p];
  dst.vx[j] = src.vx[p];
  dst.vy[j] = src.vy[p];
  dst.vz[j] = src.vz[p];
  dst.m[j] = src.m[p];
  dst.id[j] = src.id[p];
}


struct Neigh {
  int s, a, e, lp, ln;
};

__device__ __forceinline__ Neigh neighbours(
    int i, const int* __restrict__ start, int nsys,
    const unsigned long long* __restrict__ khi,
    const unsigned long long* __restrict__ klo) {
  Neigh r;
  r.s = nsys > 1 ? sys_of_slot(i, start, nsys) : 0;
  r.a = start[r.s];
  r.e = start[r.s + 1];
  const unsigned long long hi = khi[i], lo = klo[i];
  r.lp = i > r.a ? cpl(khi[i - 1], klo[i - 1], hi, lo) : -1;
  r.ln = i + 1 < r.e ? cpl(hi, lo, khi[i + 1], klo[i + 1]) : -1;
  return r;
}

// Cell openings and closings.  With l_i the levels keys i and i+1 share,
// the internal cells that start at slot i have depths (l_{i-1}, l_i] and
// those that end at slot i depths (l_i, l_{i-1}].  Cells of one depth are
// disjoint runs of slots, so in slot order they open and close
// alternately: the k-th cell of depth d to open ends at the k-th closing
// of depth d.  Every opening and closing gets its rank among the events of
// its depth from per-block histograms (depth-major, scanned) and, inside a
// block, from per-wave ballots; a cell's last slot is then one lookup, not
// a search of the keys.  The opening ranks also order the depth buckets
// the per-level reduction walks.
constexpr int kEmitWaves = kBlock / 64;
typedef unsigned long long DepthMasks[kEmitWaves][kDepthBins];

// m[w][d] = lanes of wave w with an event at depth d (lo < d <= hi).  Only
// the depths some lane of the wave has, (min lo, max hi], take a ballot
// (a wave of Morton-consecutive slots spans ~10-20 of the 43); the others
// are cleared one per lane first (same wave, so the writes stay in order).
__device__ __forceinline__ void depth_ballots(int lo, int hi, DepthMasks& m) {
  const int w = threadIdx.x >> 6, lane = lane_id();
  for (int d = lane; d < kDepthBins; d += kWave) m[w][d] = 0;
  const int d0 = max(wave_min_int(lo) + 1, 0);
  const int d1 = min(wave_max_int(hi), kDepthBins - 1);
  for (int d = d0; d <= d1; d++) {
    const unsigned long long b = ballot(lo < d && d <= hi);
    if (lane == 0) m[w][d] = b;
  }
}

// Events at depth d in this block before this thread (after depth_ballots
// and a barrier).
__device__ __forceinline__ int block_rank(int d, const DepthMasks& m) {
  const int w = threadIdx.x >> 6;
  int r = mbcnt(m[w][d]);
  for (int q = 0; q < w; q++) r += __popcll(m[q][d]);
  return r;
}

// Entries (internal cells + the leaf) that start at slot i, and this
// block's histograms of openings (dhist[d * nb + blk]) and closings
// (dhist[(kDepthBins + d) * nb + blk]) per depth.
__global__ void __launch_bounds__(kBlock)
    emit_count_kernel(const int* __restrict__ start, int nsys, int n,
                      const unsigned long long* __restrict__ khi,
                      const unsigned long long* __restrict__ klo,
                      int* __restrict__ counts, int* __restrict__ dhist) {
  __shared__ DepthMasks mo, mc;
  const int i = blockIdx.x * kBlock + threadIdx.x;
  int lp = 0, ln = 0;
  if (i < n) {
    const Neigh g = neighbours(i, start, nsys, khi, klo);
    lp = g.lp;
    ln = g.ln;
    counts[i] = max(0, ln - lp) + 1;
  }
  depth_ballots(lp, ln, mo);
  depth_ballots(ln, lp, mc);
  __syncthreads();
  for (int d = threadIdx.x; d < kDepthBins; d += kBlock) {
    int o = 0, c = 0;
    for (int q = 0; q < kEmitWaves; q++) {
      o += __popcll(mo[q][d]);
      c += __popcll(mc[q][d]);
    }
    dhist[d * gridDim.x + blockIdx.x] = o;
    dhist[(kDepthBins + d) * gridDim.x + blockIdx.x] = c;
  }
}

// close_at[rank of a cell's opening] = the cell's last slot, written by the
// slot that closes it.  doff is dhist exclusive-scanned; closings of depth
// d are ranked from the total number of openings of shallower depths,
// doff[(kDepthBins + d) * nb] - doff[kDepthBins * nb], which is where the
// openings of depth d start.  Same grid as emit_count_kernel.
__global__ void __launch_bounds__(kBlock)
    close_kernel(const int* __restrict__ start, int nsys, int n,
                 const unsigned long long* __restrict__ khi,
                 const unsigned long long* __restrict__ klo,
                 const int* __restrict__ doff, int* __restrict__ close_at) {
  __shared__ DepthMasks mc;
  const int i = blockIdx.x * kBlock + threadIdx.x;
  int lp = 0, ln = 0;
  if (i < n) {
    const Neigh g = neighbours(i, start, nsys, khi, klo);
    lp = g.lp;
    ln = g.ln;
  }
  depth_ballots(ln, lp, mc);
  __syncthreads();
  const int nb = gridDim.x;
  const int total = doff[kDepthBins * nb];
  for (int d = ln + 1; d <= lp; d++)
    close_at[doff[(kDepthBins + d) * nb + blockIdx.x] - total + block_rank(d, mc)] = i;
}

__global__ void sys_nodes_kernel(const int* __restrict__ start,
                                 const int* __restrict__ offsets, int nsys,
                                 int* __restrict__ node_start) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s <= nsys) node_start[s] = offsets[start[s]];
}

// Writes the cells and the leaf that start at slot i (pre-order), files
// each internal cell into its depth bucket, and fills the leaf's mass/COM
// from its particle.  Same grid as emit_count_kernel.  W32: the FP32
// walk's 32-byte records too (aux.walk32).
template <bool W32>
__global__ void __launch_bounds__(kBlock)
    emit_kernel(const int* __restrict__ start, int nsys, int n,
                const unsigned long long* __restrict__ khi,
                const unsigned long long* __restrict__ klo,
                const int* __restrict__ offsets,
                const int* __restrict__ doff,
                const int* __restrict__ close_at,
                const double* __restrict__ root, double theta, int cap,
                const double* __restrict__ x, const double* __restrict__ y,
                const double* __restrict__ z, const double* __restrict__ m,
                WalkNode* __restrict__ nodes, TreeAux aux,
                int* __restrict__ bucket, int* __restrict__ flags,
                int* __restrict__ stats, int* __restrict__ blkstats) {
  __shared__ DepthMasks mo;
  const int i = blockIdx.x * kBlock + threadIdx.x;
  Neigh g;
  g.s = g.a = g.e = 0;
  g.lp = g.ln = 0;
  if (i < n) g = neighbours(i, start, nsys, khi, klo);
  depth_ballots(g.lp, g.ln, mo);
  __syncthreads();
  const int nb = gridDim.x;
  const int total = offsets[n];
  const bool fits = total <= cap;
  if (!fits && i == 0) atomicOr(flags, kFlagNodeOverflow);
  int my_depth = 0, my_bucket = 0;
  if (i < n && fits) {
    const double s0 = root[4 * g.s + 3];
    const int base = offsets[i];
    const int nint = max(0, g.ln - g.lp);
    for (int k = 0; k < nint; k++) {
      const int d = g.lp + 1 + k;
      const int node = base + k;
      const int r = doff[d * nb + blockIdx.x] + block_rank(d, mo);
      // (clamped into this system's slots: a wrong rank could not fault)
      const int b = min(max(close_at[r], i), g.e - 1);
      WalkNode w;
      w.cx = w.cy = w.cz = 0;
      w.m = w.gm = 0;
      w.gm32 = 0;
      w.side = ldexp(s0, -d);
      const bool bkt = d >= kLevels;
      const double ro = w.side / theta;
      w.r2open32 = bkt ? INFINITY : (float)(ro * ro);
      w.skip = offsets[b + 1];
      w.leaf = bkt ? -2 : -1;
      nodes[node] = w;
      if constexpr (W32) aux.walk32[node].link = link32_cell(d, w.skip);
      if (d == 0 && aux.parent) aux.parent[node] = -1;
      bucket[r] = node;
      if (bkt) my_bucket++;
    }
    const int leafn = base + nint;
    const int ld = 1 + max(g.lp, g.ln);
    const double mi = m[i];
    WalkNode w;
    w.cx = x[i];
    w.cy = y[i];
    w.cz = z[i];
    w.m = mi;
    w.gm = kG * mi;
    w.gm32 = (float)(kG * mi);
    w.side = ldexp(s0, -ld);
    w.r2open32 = -1.0f;
    w.skip = leafn + 1;
    w.leaf = i;
    nodes[leafn] = w;
    if constexpr (W32) {
      WalkNode32 v;
      v.cx = w.cx;
      v.cy = w.cy;
      v.cz = w.cz;
      v.link = kLink32Leaf | (unsigned)i;
      v.gm32 = w.gm32;
      aux.walk32[leafn] = v;
    }
    if (ld == 0 && aux.parent) aux.parent[leafn] = -1;
    my_depth = ld;
  }
  // Deepest leaf and number of buckets of this block: plain stores of
  // per-block partials (the host reduces them when asked), no atomics on a
  // shared word from every wave.
  __shared__ int wmd[kEmitWaves], wnb[kEmitWaves];
  int md = my_depth;
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) md = max(md, __shfl_xor(md, d, 64));
  const int nbk = wave_sum(my_bucket);
  if (lane_id() == 0) {
    wmd[threadIdx.x >> 6] = md;
    wnb[threadIdx.x >> 6] = nbk;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int bmd = 0, bnb = 0;
    for (int q = 0; q < kEmitWaves; q++) {
      bmd = max(bmd, wmd[q]);
      bnb += wnb[q];
    }
    blkstats[2 * blockIdx.x] = bmd;
    blkstats[2 * blockIdx.x + 1] = bnb;
  }
  if (i == 0) {
    stats[0] = total;
    stats[1] = total - n;
  }
}

// Leaves take their particle's position and mass (frozen replays refresh
// the leaves of an existing topology with new particle data).
__global__ void __launch_bounds__(kBlock)
    leaf_refresh_kernel(const int* __restrict__ offsets, int n,
                        const double* __restrict__ x,
                        const double* __restrict__ y,
                        const double* __restrict__ z,
                        const double* __restrict__ m,
                        WalkNode* __restrict__ nodes,
                        WalkNode32* __restrict__ walk32,
                        const int* __restrict__ flags) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n || *flags) return;
  const int leafn = offsets[i + 1] - 1;
  const double mi = m[i];
  nodes[leafn].cx = x[i];
  nodes[leafn].cy = y[i];
  nodes[leafn].cz = z[i];
  nodes[leafn].m = mi;
  nodes[leafn].gm = kG * mi;
  nodes[leafn].gm32 = (float)(kG * mi);
  if (walk32) {
    walk32[leafn].cx = x[i];
    walk32[leafn].cy = y[i];
    walk32[leafn].cz = z[i];
    walk32[leafn].gm32 = (float)(kG * mi);
  }
}

// One internal cell p sums its children (already final) in child order
// 0..7 with the reference's operation sequence, src/octree.cpp:203-221
// (mass) and :252-278 (COM), and links them to itself.  One pass along the
// sibling chain (child c's record holds the next sibling, skip) keeps four
// running sums: the mass sum and the three COM sums are independent
// sequences, each added in child order exactly as the reference's two
// loops add them; the COM sums are used only when the mass is not zero.
// The parent links are written only when parent is not null (adjoint).
// (Running sums instead of the children's values: 30 VGPRs instead of 106,
// eight wavefronts per SIMD instead of four for this latency-bound chain.)
template <bool W32>
__device__ __forceinline__ void reduce_cell(int p, WalkNode* __restrict__ nodes,
                                            int* __restrict__ parent,
                                            WalkNode32* __restrict__ walk32) {
#pragma clang fp contract(off)
  const int end = nodes[p].skip;
  double M = 0, tx = 0, ty = 0, tz = 0;
  for (int c = p + 1; c < end;) {
    const double cm = nodes[c].m;
    const double px = nodes[c].cx * cm;
    const double py = nodes[c].cy * cm;
    const double pz = nodes[c].cz * cm;
    const int next = nodes[c].skip;
    if (parent) parent[c] = p;
    M = M + cm;
    tx = tx + px;
    ty = ty + py;
    tz = tz + pz;
    c = max(next, c + 1);
  }
  double cxv = 0, cyv = 0, czv = 0;
  if (M != 0) {  // else COM (0,0,0), src/octree.cpp:259-262
    cxv = tx / M;
    cyv = ty / M;
    czv = tz / M;
  }
  nodes[p].cx = cxv;
  nodes[p].cy = cyv;
  nodes[p].cz = czv;
  nodes[p].m = M;
  nodes[p].gm = kG * M;
  nodes[p].gm32 = (float)(kG * M);
  if constexpr (W32) {
    walk32[p].cx = cxv;
    walk32[p].cy = cyv;
    walk32[p].cz = czv;
    walk32[p].gm32 = (float)(kG * M);
  }
}

// One tree level: every internal cell of depth d.  Launched from the
// deepest level up; kernel boundaries order the levels, so no fences are
// needed.  W32: the FP32 walk's 32-byte records are written too (a
// variant of its own, so that the pass without them costs what it did).
template <bool W32>
__global__ void __launch_bounds__(kBlock)
    level_kernel(int d, int nblocks, const int* __restrict__ doff,
                 const int* __restrict__ bucket, WalkNode* __restrict__ nodes,
                 int* __restrict__ parent, WalkNode32* __restrict__ walk32,
                 const int* __restrict__ flags) {
  if (*flags) return;
  const int b0 = doff[d * nblocks], b1 = doff[(d + 1) * nblocks];
  for (int q = b0 + blockIdx.x * kBlock + threadIdx.x; q < b1; q += gridDim.x * kBlock)
    reduce_cell<W32>(bucket[q], nodes, parent, walk32);
}

// Levels d1 down to d0 in one workgroup, a barrier between levels (which
// makes each level's results visible to the workgroup): the top levels,
// which hold few cells, in one launch instead of one per level.
constexpr int kFewBlock = 1024;
template <bool W32>
__global__ void __launch_bounds__(kFewBlock)
    few_levels_kernel(int d1, int d0, int nblocks, const int* __restrict__ doff,
                      const int* __restrict__ bucket, WalkNode* __restrict__ nodes,
                      int* __restrict__ parent, WalkNode32* __restrict__ walk32,
                      const int* __restrict__ flags) {
  if (*flags) return;
  for (int d = d1; d >= d0; d--) {
    const int b0 = doff[d * nblocks], b1 = doff[(d + 1) * nblocks];
    for (int q = b0 + threadIdx.x; q < b1; q += kFewBlock)
      reduce_cell<W32>(bucket[q], nodes, parent, walk32);
    __syncthreads();
  }
}

// Grid of every level launch (grid-stride over the level's cells).  1024
// blocks shortened the widest levels (35 -> 27 us at N = 2^20) but made
// each of the ~30 near-empty levels cost 4.4 us instead of ~1.8 us.
constexpr int kLevelBlocks = 256;

inline int blocks_for(size_t n, int b) { return (int)((n + b - 1) / b); }

}  // namespace

void launch_fill_int(int* p, int v, size_t n, hipStream_t s) {
  if (n) fill_int_kernel<<<blocks_for(n, kBlock), kBlock, 0, s>>>(p, v, n);
}
void launch_fill_u64(unsigned long long* p, unsigned long long v, size_t n,
                     hipStream_t s) {
  if (n) fill_u64_kernel<<<blocks_for(n, kBlock), kBlock, 0, s>>>(p, v, n);
}

void launch_bbox(const State& st, Systems& sys, hipStream_t s) {
  const int nw = sys.n_waves;
  if (nw <= 0) return;
  bbox_kernel<<<blocks_for((size_t)nw * 64, kBlock), kBlock, 0, s>>>(
      st.x, st.y, st.z, sys.start, sys.wave_start, sys.nsys, nw, sys.bbox);
}

void launch_root(Systems& sys, const StepParams& p, const int* flags,
                 hipStream_t s) {
  root_kernel<<<sys.nsys, kRootBlock, 0, s>>>(
      sys.bbox, sys.n_waves, sys.wave_start, sys.root, sys.nsys, p.root_mode,
      p.root_center[0], p.root_center[1], p.root_center[2], p.root_size, flags);
}

int max_prim_levels(int nsys) {
  int sys_bits = 0;
  while ((1 << sys_bits) < nsys) sys_bits++;
  const int levels = (64 - sys_bits) / 3;
  return levels > 21 ? 21 : levels;
}

int launch_keys(const State& st, Systems& sys, BuildBuffers& b, int prim_levels,
                hipStream_t s) {
  const int nw = sys.n_waves;
  int sys_bits = 0;
  while ((1 << sys_bits) < sys.nsys) sys_bits++;
  const int levels = prim_levels < 1 ? 1 : (prim_levels > max_prim_levels(sys.nsys)
                                                ? max_prim_levels(sys.nsys)
                                                : prim_levels);
  if (nw > 0)
    keys_kernel<<<blocks_for((size_t)nw * 64, kBlock), kBlock, 0, s>>>(
        st.x, st.y, st.z, sys.start, sys.wave_start, sys.nsys, nw, sys.root,
        3 * levels, levels, b.prim0, b.val0, b.khi_u, b.klo_u);
  return sys_bits + 3 * levels;
}

void launch_fixup(BuildBuffers& b, const unsigned long long* prim,
                  unsigned* perm, unsigned* free_perm, const State& st, int n,
                  hipStream_t s) {
  if (n <= 0) return;
  gather_keys_kernel<<<blocks_for(n, kBlock), kBlock, 0, s>>>(
      perm, n, b.khi_u, b.klo_u, b.khi, b.klo);
  launch_fill_int(b.runs, 0, 4, s);
  fixup_kernel<<<blocks_for(n, kFixBlock), kFixBlock, 0, s>>>(prim, perm, b.khi,
                                                        b.klo, st.id, n,
                                                        b.runs, b.runs_big);
  short_run_kernel<<<kShortGrid, kShortBlock, 0, s>>>(b.runs, b.khi, b.klo, perm, st.id);
  // khi_u / klo_u (already gathered) and the sort's other value array are
  // free until the next build: scratch for the long runs
  long_run_kernel<<<kRunGrid, kRunBlock, 0, s>>>(b.runs, b.khi, b.klo, perm, b.khi_u,
                                                 b.klo_u, free_perm, st.id, b.runs_big);
}

// Runs listed for long_run_kernel have more than kShortRun entries: at
// most cap / (kShortRun + 1) in the short list, cap / (kWaveRun + 1) in
// the long one.
int run_list_big(size_t cap) { return (int)(4 + 2 * (cap / (kShortRun + 1) + 1)); }
size_t run_list_ints(size_t cap) {
  return (size_t)run_list_big(cap) + 2 * (cap / (kWaveRun + 1) + 1);
}

void launch_permute(const State& src, State& dst, const unsigned* perm, int n,
                    hipStream_t s) {
  if (n) permute_kernel<<<blocks_for(n, kBlock), kBlock, 0, s>>>(src, dst, perm, n);
}

int emit_blocks(size_t n) { return blocks_for(n, kBlock); }
size_t depth_hist_ints(size_t n) { return 2 * (size_t)kDepthBins * blocks_for(n, kBlock) + 1; }

void launch_emit(BuildBuffers& b, Systems& sys, const StepParams& p,
                 const State& st, WalkNode* nodes, TreeAux& aux, int n,
                 int cap, hipStream_t s) {
  const int nb = blocks_for(n, kBlock);
  emit_count_kernel<<<nb, kBlock, 0, s>>>(sys.start, sys.nsys, n, b.khi, b.klo,
                                          b.counts, b.dhist);
  exclusive_scan(b.counts, b.offsets, n, b.scan_tmp, s);
  exclusive_scan(b.dhist, b.doff, 2 * kDepthBins * nb, b.scan_tmp, s);
  close_kernel<<<nb, kBlock, 0, s>>>(sys.start, sys.nsys, n, b.khi, b.klo, b.doff,
                                     b.close_at);
  sys_nodes_kernel<<<blocks_for(sys.nsys + 1, 64), 64, 0, s>>>(
      sys.start, b.offsets, sys.nsys, sys.node_start);
  if (aux.walk32)
    emit_kernel<true><<<nb, kBlock, 0, s>>>(
        sys.start, sys.nsys, n, b.khi, b.klo, b.offsets, b.doff, b.close_at, sys.root,
        p.theta, cap, st.x, st.y, st.z, st.m, nodes, aux, b.bucket, b.flags,
        b.stats, b.blkstats);
  else
    emit_kernel<false><<<nb, kBlock, 0, s>>>(
        sys.start, sys.nsys, n, b.khi, b.klo, b.offsets, b.doff, b.close_at, sys.root,
        p.theta, cap, st.x, st.y, st.z, st.m, nodes, aux, b.bucket, b.flags,
        b.stats, b.blkstats);
}

void launch_bottom_up(BuildBuffers& b, WalkNode* nodes, TreeAux& aux, int n,
                      int nsys, hipStream_t s) {
  const int nb = blocks_for(n, kBlock);
  // One launch per level, except the top levels of all systems while they
  // hold at most kFewBlock cells (nsys * 8^d), which one workgroup reduces
  // in one launch.  (The deep levels keep a launch each although they are
  // usually empty: in a collapse -- a cold shell passing through its
  // centre -- nearly the whole tree lies below 2^-24 of the root's side,
  // and one workgroup would then reduce millions of cells alone.)
  int top = -1;
  for (long long c = nsys; c <= kFewBlock; c *= 8) top++;
  const bool w32 = aux.walk32 != nullptr;
  for (int d = kLevels; d > top; d--) {
    if (w32)
      level_kernel<true><<<kLevelBlocks, kBlock, 0, s>>>(d, nb, b.doff, b.bucket, nodes,
                                                         aux.parent, aux.walk32, b.flags);
    else
      level_kernel<false><<<kLevelBlocks, kBlock, 0, s>>>(d, nb, b.doff, b.bucket, nodes,
                                                          aux.parent, nullptr, b.flags);
  }
  if (top >= 0) {
    if (w32)
      few_levels_kernel<true><<<1, kFewBlock, 0, s>>>(top, 0, nb, b.doff, b.bucket, nodes,
                                                      aux.parent, aux.walk32, b.flags);
    else
      few_levels_kernel<false><<<1, kFewBlock, 0, s>>>(top, 0, nb, b.doff, b.bucket, nodes,
                                                       aux.parent, nullptr, b.flags);
  }
}

void launch_leaf_refresh(BuildBuffers& b, const State& st, WalkNode* nodes,
                         WalkNode32* walk32, int n, hipStream_t s) {
  if (n > 0)
    leaf_refresh_kernel<<<blocks_for(n, kBlock), kBlock, 0, s>>>(
        b.offsets, n, st.x, st.y, st.z, st.m, nodes, walk32, b.flags);
}

}  // namespace nbs
