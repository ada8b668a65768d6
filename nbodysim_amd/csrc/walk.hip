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
//     division, no contraction); a margin test (the generic loop: squared,
//     1e-9 relative; the hand-written loop: f32(d^2) against the record's
//     f32 radius widened by 8 ulps) settles all but the near-ties without
//     the sqrt and division.  Forces: G m / |d|^3 from the v_rsq_f64
//     estimate with a first-order correction of its cube (inv_cube_gm);
//     the damped law takes 1/r with one Newton step (rsqrt_f64).  Either
//     way the estimate's ~2^-26 relative error enters squared, ~1e-15.
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
typedef __attribute__((address_space(4))) const WalkNode32* ConstNodes32;

// The fields of a WalkNode the walk reads, loaded as 64-bit words with
// scalar loads (the node index is wave-uniform); the compiler merges the
// adjacent words of a visit into two wide loads.  (Loading the record as
// one vector value and extracting its elements miscompiled with this
// toolchain: extractions from the 512-bit register tuple read the wrong
// registers.  Separate scalar loads merged by the compiler do not.)
struct NodeRegs {
  double cx, cy, cz, side, gm;
  float gm32, r2open32;
  int skip, leaf;
};

typedef __attribute__((address_space(4))) const unsigned long long* ConstQwords;
typedef __attribute__((address_space(4))) const unsigned* ConstDwords;

// 64-bit words of the record: 0-2 cx cy cz, 3 (skip, leaf), 4 gm, 5 side,
// 6 (gm32, r2open32).  skip and leaf are read as 32-bit words so their
// tests stay on the scalar unit.
__device__ __forceinline__ NodeRegs load_node(ConstNodes nodes, int k) {
  static_assert(offsetof(WalkNode, skip) == 24 && offsetof(WalkNode, leaf) == 28 &&
                    offsetof(WalkNode, gm) == 32 && offsetof(WalkNode, side) == 40 &&
                    offsetof(WalkNode, gm32) == 48 && offsetof(WalkNode, r2open32) == 52,
                "load_node follows the WalkNode layout");
  const ConstQwords q = (ConstQwords)(nodes + k);
  const ConstDwords d = (ConstDwords)(nodes + k);
  NodeRegs r;
  r.cx = __builtin_bit_cast(double, q[0]);
  r.cy = __builtin_bit_cast(double, q[1]);
  r.cz = __builtin_bit_cast(double, q[2]);
  r.skip = (int)d[6];
  r.leaf = (int)d[7];
  r.gm = __builtin_bit_cast(double, q[4]);
  r.side = __builtin_bit_cast(double, q[5]);
  r.gm32 = __builtin_bit_cast(float, d[12]);
  r.r2open32 = __builtin_bit_cast(float, d[13]);
  return r;
}

__device__ __forceinline__ double sq_sum_ref(double dx, double dy, double dz) {
#pragma clang fp contract(off)
  return dx * dx + dy * dy + dz * dz;  // src/vector.cpp:146 rounding
}

// |d|^2 of the FP64 force: two fused multiply-adds (the opening decision
// keeps the reference's rounding, sq_sum_ref; the force needs no bit
// pattern of the reference's, only its value to well within 1e-12).
__device__ __forceinline__ double sq_sum_fused(double dx, double dy, double dz) {
#pragma clang fp contract(off)
  return fma(dx, dx, fma(dy, dy, dz * dz));
}

// Called wave-uniformly (every lane of the wave, divergence-free).  The
// exact test runs only when some lane is within the margin; the empty asm
// keeps the compiler from hoisting the sqrt and division (cheap-looking
// in IR, ~30 FP64 instructions in ISA) out of that rare branch.
__device__ __forceinline__ bool mac_open_f64(double d2, double side,
                                             double theta, double theta2) {
#pragma clang fp contract(off)
  const double s2 = side * side;
  const double t = s2 - theta2 * d2;
  const double tol = 1e-9 * s2;
  bool open = t > tol;
  const bool tie = !(t > tol) && !(t < -tol);
  if (ballot(tie)) {
    __asm__ volatile("");
    if (tie) open = side / __builtin_sqrt(d2) > theta;
  }
  return open;
}

// G m / d^3 from d^2 in FP64: with the hardware estimate r = R (1 + eps)
// of R = 1 / sqrt(d^2) and e = 1 - d^2 r^2 = -2 eps - eps^2,
// r^3 (1 + 1.5 e) = R^3 (1 + O(7.5 eps^2)): relative error ~2e-15 for the
// estimate's ~2^-26 (the same order as a Newton step on r, cubed), in
// seven operations.
__device__ __forceinline__ double inv_cube_gm(double d2, double gm) {
#pragma clang fp contract(off)
  const double r = __builtin_amdgcn_rsq(d2);
  const double r2 = r * r;
  const double e = fma(-d2, r2, 1.0);
  const double g = gm * (r2 * r);
  return fma(g, fma(e, 0.5, e), g);
}

// 1/sqrt(d2) in FP64: the hardware estimate refined by one Newton-Raphson
// step (relative error ~1e-15 after it; the estimate's ~2^-26 squares).
__device__ __forceinline__ double rsqrt_f64(double d2) {
  double r = __builtin_amdgcn_rsq(d2);
  const double h = d2 * r;
  const double e = fma(-h, r, 1.0);
  return fma(0.5 * r, e, r);
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

// One accepted body-node interaction of one lane: a += G m d / |d|^3
// (src/main.cpp:306-333), d = com - x formed in FP64 (and, for FP32 forces,
// rounded to FP32 with its square already formed by the caller).
template <bool F32, bool DAMP, typename AccT>
__device__ __forceinline__ void interact(const NodeRegs& nd, double dx, double dy,
                                         double dz, float fdx, float fdy, float fdz,
                                         float fd2, AccT& ax, AccT& ay, AccT& az,
                                         int& cnt) {
  if (F32) {
    const float r = __builtin_amdgcn_rsqf(fd2);
    float sc;
    if (DAMP) {
      const float d = fd2 * r;
      sc = nd.gm32 * r / (fd2 + __expf(-d));
    } else {
      sc = nd.gm32 * (r * r * r);
    }
    ax = fmaf(sc, fdx, (float)ax);
    ay = fmaf(sc, fdy, (float)ay);
    az = fmaf(sc, fdz, (float)az);
  } else {
    const double d2 = sq_sum_fused(dx, dy, dz);
    double sc;
    if (DAMP) {
      const double r = rsqrt_f64(d2);
      const double d = d2 * r;
      sc = nd.gm * r / (d2 + exp(-d));
    } else {
      sc = inv_cube_gm(d2, nd.gm);
    }
    ax = fma(sc, dx, (double)ax);
    ay = fma(sc, dy, (double)ay);
    az = fma(sc, dz, (double)az);
  }
  cnt++;
}

// The FP32 walk loop (one particle per lane) written out for gfx950, with
// the plain force (walk_loop_f32) or the damped one (walk_loop_f32_damped,
// which differs only in the force macro).  Same traversal, lane decisions and arithmetic, in the same
// order, as the generic loop in walk_kernel; what differs is the
// instruction stream.  Per visit: 10 VALU instructions for d = com - x
// (FP64, rounded to FP32), d^2 and the opening test, 8 for the force on
// the interacting lanes, and scalar code for everything else; the common
// case (a cell every lane accepts) takes 9 scalar ALU instructions (11 with
// the wave's interaction count) and 3 branches on the 64-byte records
// (WalkNode), whose buckets need no test of their own, 5 more and a branch
// on the 32-byte ones (WalkNode32), which decode the skip link and the
// opening radius from a packed link word but put twice as many records in
// a cache line and in the L2.  The walk waits on each next record (the scalar cache misses most
// of them), so the next record is requested as early as it is known: for
// a cell every taking lane accepts, after the opening test, before the
// force arithmetic; for an opened cell (next = this + 1) right after the
// opening test, before the stack push; for a leaf (next = this + 1) as
// soon as the position differences have read the current record and its
// particle slot is taken.
//
// Fixed SGPRs (clobbered): the 64-byte record in s[48:63] (cx s[48:49],
// cy s[50:51], cz s[52:53], skip s54, leaf s55, gm32 s60, r2open32 s61),
// or the 32-byte one in s[48:55] (cx s[48:49], cy s[50:51], cz s[52:53],
// link s54, gm32 s55) with, from a cell's link, its skip s56, its
// (side/theta)^2 s57 and its depth s58;
// s[46:47] A (lanes taking the node), s[44:45] I (lanes interacting),
// s[42:43] O (lanes opening), s[40:41] saved exec, s[38:39] a popped
// stack mask.  The stack of waiting lanes is entry `slot` in lane `slot`
// of sp (position), slo/shi (mask); lane 63 of sp stays -1, the position
// an empty stack reports.  M0 carries the lane index of stack writes (a
// VALU instruction reads one SGPR at most); the block restores it.  Every
// scalar load is waited for inside the block.
//
// The record after the last node of a system is requested (and not used)
// when the walk ends: each record array holds one beyond its capacity.
// Record offsets are 32-bit: callers keep (capacity + 1) * 64 below 2^32,
// and the 32-byte records exist only below 2^25 nodes (their skip links).

// Interaction counting on the lanes of MASK (the interacting lanes), one
// of three forms chosen per launch: a per-lane count (one VALU instruction
// per interacting visit), the wave's total on the scalar unit (the mask's
// population count, no VALU), or none.  Each form sits between v_rsq and
// the first instruction that reads its result, and each is at least one
// instruction: on gfx950 a VALU instruction may read a transcendental
// result only after one wait state, which the compiler inserts for its own
// code but not inside this asm (without it the multiply reads the old
// register).
#define NBS_CNT_LANE(MASK) "v_add_u32 %[cnt], 1, %[cnt]\n\t"
#define NBS_CNT_WAVE(MASK)                                            \
  "s_bcnt1_i32_b64 %[tc], " MASK "\n\t"                               \
  "s_add_u32 %[tot], %[tot], %[tc]\n\t"
#define NBS_CNT_NONE(MASK) "s_nop 0\n\t"

// The force on the lanes of MASK and the loop test (the next record has
// been requested); NBS_WALK_TAIL requests it first.
#define NBS_WALK_FORCE_PLAIN(MASK, CNT)                               \
  "s_and_saveexec_b64 s[40:41], " MASK "\n\t"                          \
  "v_rsq_f32 %[tv], %[d2]\n\t"                                        \
  CNT(MASK)                                                           \
  "v_mul_f32 %[d2], %[tv], %[tv]\n\t"                                 \
  "v_mul_f32 %[d2], %[d2], %[tv]\n\t"                                 \
  "v_mul_f32 %[d2], %[gm], %[d2]\n\t"                                 \
  "v_fmac_f32 %[ax], %[d2], %[fx]\n\t"                                \
  "v_fmac_f32 %[ay], %[d2], %[fy]\n\t"                                \
  "v_fmac_f32 %[az], %[d2], %[fz]\n\t"                                \
  "s_mov_b64 exec, s[40:41]\n\t"                                      \
  "s_cmp_lt_u32 %[n], %[lim]\n\t"                                     \
  "s_cbranch_scc1 L_loop%=\n\t"                                       \
  "s_branch L_slow%=\n"

// The two record formats (FMT W64: WalkNode, W32: WalkNode32): the load of
// record n, the leaf test, where G*m, the skip link and (side/theta)^2
// are, the decoding a 32-byte cell needs (its skip s56, depth s58 and
// (side/theta)^2 s57 = the root's value minus 2 * depth in the exponent;
// rmask 0 keeps an infinite root value), the bucket test and a leaf's
// particle slot minus pbeg into t.
#define NBS_REC_LOAD_W64                                              \
  "s_lshl_b32 %[t], %[n], 6\n\t"                                      \
  "s_load_dwordx16 s[48:63], %[base], %[t]\n\t"
#define NBS_REC_LOAD_W32                                              \
  "s_lshl_b32 %[t], %[n], 5\n\t"                                      \
  "s_load_dwordx8 s[48:55], %[base], %[t]\n\t"
// The same, the offset formed in sk (t holds a leaf's slot meanwhile).
#define NBS_REC_LOADSK_W64                                            \
  "s_lshl_b32 %[sk], %[n], 6\n\t"                                     \
  "s_load_dwordx16 s[48:63], %[base], %[sk]\n\t"
#define NBS_REC_LOADSK_W32                                            \
  "s_lshl_b32 %[sk], %[n], 5\n\t"                                     \
  "s_load_dwordx8 s[48:55], %[base], %[sk]\n\t"
#define NBS_REC_LEAF_W64 "s_cmp_gt_i32 s55, -1\n\t"
#define NBS_REC_LEAF_W32 "s_cmp_lt_i32 s54, 0\n\t"
#define NBS_REC_GM_W64 "s60"
#define NBS_REC_GM_W32 "s55"
#define NBS_REC_SKIP_W64 "s54"
#define NBS_REC_SKIP_W32 "s56"
#define NBS_REC_R2_W64 "s61"
#define NBS_REC_R2_W32 "s57"
#define NBS_REC_DECODE1_W64 ""
#define NBS_REC_DECODE1_W32                                           \
  "s_bfe_u32 s58, s54, 0x60019\n\t"                                   \
  "s_and_b32 s56, s54, 0x1ffffff\n\t"
#define NBS_REC_DECODE2_W64 ""
#define NBS_REC_DECODE2_W32                                           \
  "s_lshl_b32 s57, s58, 24\n\t"                                       \
  "s_and_b32 s57, s57, %[rmask]\n\t"                                  \
  "s_sub_u32 s57, %[rbase], s57\n\t"
// A depth-42 bucket is opened by every lane that takes it.  The 64-byte
// record of a bucket carries an infinite opening radius, so the opening
// test itself opens it for every lane with a finite d^2 and no bucket test
// is needed; the 32-byte record derives its radius from the depth and tests
// the depth.
#define NBS_REC_BUCKET_W64 ""
#define NBS_REC_BUCKET_W32                                            \
  "s_cmp_eq_u32 s58, 42\n\t"                                         \
  "s_cbranch_scc1 L_bucket%=\n\t"
#define NBS_REC_SLOT_W64 "s_sub_u32 %[t], s55, %[pbeg]\n\t"
#define NBS_REC_SLOT_W32                                              \
  "s_and_b32 %[t], s54, 0x7fffffff\n\t"                               \
  "s_sub_u32 %[t], %[t], %[pbeg]\n\t"

#define NBS_WALK_TAIL(FORCE, MASK, CNT, FMT)                          \
  NBS_REC_LOAD_##FMT                                                  \
  FORCE(MASK, CNT)

// The damped force (src/main.cpp:318: |a| = G m / (r^2 + exp(-r))), in the
// generic loop's operations: d = r^2 * (1/r) = r, e = exp2(-d log2 e) (the
// hardware exponential, __expf), q = (G m * (1/r)) / (r^2 + e) correctly
// rounded by the scale / reciprocal / Newton / fixup division sequence.
// s[38:39] receives the first scale's flag (unused); vcc the second's,
// which v_div_fmas reads six instructions later.
#define NBS_WALK_FORCE_DAMPED(MASK, CNT)                              \
  "s_and_saveexec_b64 s[40:41], " MASK "\n\t"                          \
  "v_rsq_f32 %[tv], %[d2]\n\t"                                        \
  CNT(MASK)                                                           \
  "v_mul_f32 %[te], %[d2], %[tv]\n\t"                                 \
  "v_mul_f32 %[te], 0xbfb8aa3b, %[te]\n\t"                            \
  "v_exp_f32 %[te], %[te]\n\t"                                        \
  "v_mul_f32 %[tn], %[gm], %[tv]\n\t"                                 \
  "v_add_f32 %[td], %[d2], %[te]\n\t"                                 \
  "v_div_scale_f32 %[q0], s[38:39], %[td], %[td], %[tn]\n\t"          \
  "v_rcp_f32 %[q1], %[q0]\n\t"                                        \
  "v_div_scale_f32 %[q2], vcc, %[tn], %[td], %[tn]\n\t"               \
  "v_fma_f32 %[q3], -%[q0], %[q1], 1.0\n\t"                           \
  "v_fmac_f32 %[q1], %[q3], %[q1]\n\t"                                \
  "v_mul_f32 %[q3], %[q2], %[q1]\n\t"                                 \
  "v_fma_f32 %[q4], -%[q0], %[q3], %[q2]\n\t"                         \
  "v_fmac_f32 %[q3], %[q4], %[q1]\n\t"                                \
  "v_fma_f32 %[q0], -%[q0], %[q3], %[q2]\n\t"                         \
  "v_div_fmas_f32 %[q0], %[q0], %[q1], %[q3]\n\t"                     \
  "v_div_fixup_f32 %[q0], %[q0], %[td], %[tn]\n\t"                    \
  "v_fmac_f32 %[ax], %[q0], %[fx]\n\t"                                \
  "v_fmac_f32 %[ay], %[q0], %[fy]\n\t"                                \
  "v_fmac_f32 %[az], %[q0], %[fz]\n\t"                                \
  "s_mov_b64 exec, s[40:41]\n\t"                                      \
  "s_cmp_lt_u32 %[n], %[lim]\n\t"                                     \
  "s_cbranch_scc1 L_loop%=\n\t"                                       \
  "s_branch L_slow%=\n"

#define NBS_WALK_F32_LOOP(FORCE, CNT, FMT) \
      "s_mov_b32 %[m0save], m0\n\t" \
      "s_mov_b64 s[46:47], %[A0]\n\t" \
      "s_min_u32 %[lim], %[tpos], %[nend]\n\t" \
      "s_cmp_lt_i32 %[n], %[nend]\n\t" \
      "s_cbranch_scc0 L_done%=\n\t" \
      NBS_REC_LOAD_##FMT \
      "L_loop%=:\n\t" \
      "s_waitcnt lgkmcnt(0)\n\t" \
      NBS_REC_LEAF_##FMT \
      "s_cbranch_scc1 L_leaf%=\n\t" \
      "v_add_f64 %[dx], s[48:49], -%[x]\n\t" \
      "v_add_f64 %[dy], s[50:51], -%[y]\n\t" \
      "v_add_f64 %[dz], s[52:53], -%[z]\n\t" \
      "s_mov_b32 %[gm], " NBS_REC_GM_##FMT "\n\t" \
      NBS_REC_DECODE1_##FMT \
      "v_cvt_f32_f64 %[fx], %[dx]\n\t" \
      "v_cvt_f32_f64 %[fy], %[dy]\n\t" \
      "v_cvt_f32_f64 %[fz], %[dz]\n\t" \
      NBS_REC_DECODE2_##FMT \
      "v_mul_f32 %[d2], %[fz], %[fz]\n\t" \
      "v_fmac_f32 %[d2], %[fy], %[fy]\n\t" \
      "v_fmac_f32 %[d2], %[fx], %[fx]\n\t" \
      NBS_REC_BUCKET_##FMT \
      /* cell: open where d^2 < (side/theta)^2 */ \
      "v_cmp_gt_f32 vcc, " NBS_REC_R2_##FMT ", %[d2]\n\t" \
      "s_and_b64 s[42:43], vcc, s[46:47]\n\t" \
      "s_cbranch_scc1 L_opened%=\n\t" \
      /* every lane that takes it accepts it: all interact, jump the */ \
      /* subtree (never backwards, even on a corrupt skip) */ \
      "s_add_i32 %[n], %[n], 1\n\t" \
      "s_max_i32 %[n], %[n], " NBS_REC_SKIP_##FMT "\n\t" \
      NBS_WALK_TAIL(FORCE, "s[46:47]", CNT, FMT) \
      "L_opened%=:\n\t" \
      /* I = A & ~O interact and wait at skip; O descend (A = O), so the */ \
      /* next record is n + 1: requested at once, the skip link kept in */ \
      /* sk for the stack; with no lane in I the force is skipped */ \
      "s_add_i32 %[n], %[n], 1\n\t" \
      "s_mov_b32 %[sk], " NBS_REC_SKIP_##FMT "\n\t" \
      NBS_REC_LOAD_##FMT \
      "s_andn2_b64 s[44:45], s[46:47], s[42:43]\n\t" \
      "s_mov_b64 s[46:47], s[42:43]\n\t" \
      "s_cbranch_scc0 L_next%=\n\t" \
      "s_cmp_eq_u32 %[tpos], %[sk]\n\t" \
      "s_cbranch_scc1 L_merge%=\n\t" \
      "s_cmp_gt_u32 %[top], 62\n\t" \
      "s_cbranch_scc1 L_ovf%=\n\t" \
      "s_mov_b32 m0, %[top]\n\t" \
      "v_writelane_b32 %[sp], %[sk], m0\n\t" \
      "v_writelane_b32 %[slo], s44, m0\n\t" \
      "v_writelane_b32 %[shi], s45, m0\n\t" \
      "s_add_u32 %[top], %[top], 1\n\t" \
      "s_mov_b32 %[tpos], %[sk]\n\t" \
      "s_min_u32 %[lim], %[tpos], %[nend]\n\t" \
      "s_branch L_int%=\n" \
      "L_merge%=:\n\t" \
      "s_sub_u32 m0, %[top], 1\n\t" \
      "v_readlane_b32 s38, %[slo], m0\n\t" \
      "v_readlane_b32 s39, %[shi], m0\n\t" \
      "s_or_b64 s[38:39], s[38:39], s[44:45]\n\t" \
      "v_writelane_b32 %[slo], s38, m0\n\t" \
      "v_writelane_b32 %[shi], s39, m0\n\t" \
      "s_branch L_int%=\n" \
      "L_ovf%=:\n\t" \
      "s_mov_b32 %[ovf], 1\n\t" \
      "s_branch L_int%=\n" \
      "L_bucket%=:\n\t" \
      /* every lane opens a bucket; nobody interacts */ \
      "s_add_i32 %[n], %[n], 1\n" \
      "L_noint%=:\n\t" \
      NBS_REC_LOAD_##FMT \
      "L_next%=:\n\t" \
      "s_cmp_lt_u32 %[n], %[lim]\n\t" \
      "s_cbranch_scc1 L_loop%=\n\t" \
      "s_branch L_slow%=\n" \
      "L_int%=:\n\t" \
      FORCE("s[44:45]", CNT) \
      "L_leaf%=:\n\t" \
      /* leaf: the next entry is known at once, so its record is requested */ \
      /* as soon as the three differences have read this one's position */ \
      /* and its slot is taken (the offset in sk); every lane that takes */ \
      /* the leaf interacts, except its own particle's (bit slot - pbeg */ \
      /* when that is below 64) */ \
      "v_add_f64 %[dx], s[48:49], -%[x]\n\t" \
      "v_add_f64 %[dy], s[50:51], -%[y]\n\t" \
      "v_add_f64 %[dz], s[52:53], -%[z]\n\t" \
      "s_mov_b32 %[gm], " NBS_REC_GM_##FMT "\n\t" \
      NBS_REC_SLOT_##FMT \
      "s_add_i32 %[n], %[n], 1\n\t" \
      NBS_REC_LOADSK_##FMT \
      "s_lshl_b64 s[38:39], 1, %[t]\n\t" \
      "s_cmp_lt_u32 %[t], 64\n\t" \
      "s_cselect_b64 s[38:39], s[38:39], 0\n\t" \
      "s_andn2_b64 s[44:45], s[46:47], s[38:39]\n\t" \
      /* only its own particle's lane takes the leaf: no force */ \
      "s_cmp_eq_u64 s[44:45], 0\n\t" \
      "s_cbranch_scc1 L_next%=\n\t" \
      "v_cvt_f32_f64 %[fx], %[dx]\n\t" \
      "v_cvt_f32_f64 %[fy], %[dy]\n\t" \
      "v_cvt_f32_f64 %[fz], %[dz]\n\t" \
      "v_mul_f32 %[d2], %[fz], %[fz]\n\t" \
      "v_fmac_f32 %[d2], %[fy], %[fy]\n\t" \
      "v_fmac_f32 %[d2], %[fx], %[fx]\n\t" \
      FORCE("s[44:45]", CNT) \
      "L_slow%=:\n\t" \
      /* n reached lim = min(stack top, nend): the lanes waiting at n */ \
      /* rejoin (the stack top is never passed), or the walk is done */ \
      "s_cmp_eq_u32 %[n], %[tpos]\n\t" \
      "s_cbranch_scc0 L_end%=\n\t" \
      "s_sub_u32 %[top], %[top], 1\n\t" \
      "v_readlane_b32 s38, %[slo], %[top]\n\t" \
      "v_readlane_b32 s39, %[shi], %[top]\n\t" \
      "s_add_u32 %[t], %[top], 63\n\t" \
      "s_and_b32 %[t], %[t], 63\n\t" \
      "v_readlane_b32 %[tpos], %[sp], %[t]\n\t" \
      "s_or_b64 s[46:47], s[46:47], s[38:39]\n\t" \
      "s_min_u32 %[lim], %[tpos], %[nend]\n\t" \
      "s_cmp_lt_u32 %[n], %[lim]\n\t" \
      "s_cbranch_scc1 L_loop%=\n" \
      "L_end%=:\n\t" \
      "s_waitcnt lgkmcnt(0)\n" \
      "L_done%=:\n\t" \
      "s_mov_b32 m0, %[m0save]"

// Operands of the FP32 loops: the shared ones, then those of each counting
// form (the lane count in a VGPR, the wave total and a temporary in SGPRs).
#define NBS_F32_OUTS                                                                      \
  [n] "+s"(n), [ovf] "+s"(overflow), [ax] "+v"(ax), [ay] "+v"(ay), [az] "+v"(az),         \
      [top] "+s"(top), [tpos] "+s"(tpos), [lim] "=&s"(lim), [t] "=&s"(t), [gm] "=&s"(gm),   \
      [m0save] "=&s"(m0save), [sp] "+v"(sp), [slo] "+v"(slo), [shi] "+v"(shi),             \
      [dx] "=&v"(dx), [dy] "=&v"(dy), [dz] "=&v"(dz), [fx] "=&v"(fx), [fy] "=&v"(fy),       \
      [fz] "=&v"(fz), [d2] "=&v"(d2), [tv] "=&v"(tv), [sk] "=&s"(sk)
#define NBS_F32_DAMPED_OUTS                                                               \
  NBS_F32_OUTS, [te] "=&v"(te), [tn] "=&v"(tn), [td] "=&v"(td), [q0] "=&v"(q0),          \
      [q1] "=&v"(q1), [q2] "=&v"(q2), [q3] "=&v"(q3), [q4] "=&v"(q4)
#define NBS_F32_INS_W64                                                                   \
  [base] "s"(nodes), [nend] "s"(nend), [pbeg] "s"(pbeg), [A0] "s"(A), [x] "v"(x),          \
      [y] "v"(y), [z] "v"(z)
#define NBS_F32_INS_W32                                                                   \
  NBS_F32_INS_W64, [rbase] "s"(__builtin_amdgcn_readfirstlane(rbase)),                     \
      [rmask] "s"(__builtin_amdgcn_readfirstlane(rmask))
#define NBS_F32_CLOBBERS_W32                                                              \
  "s38", "s39", "s40", "s41", "s42", "s43", "s44", "s45", "s46", "s47", "s48", "s49",     \
      "s50", "s51", "s52", "s53", "s54", "s55", "s56", "s57", "s58", "vcc", "scc", "memory"
#define NBS_WALK_CLOBBERS                                                                 \
  "s38", "s39", "s40", "s41", "s42", "s43", "s44", "s45", "s46", "s47", "s48", "s49",     \
      "s50", "s51", "s52", "s53", "s54", "s55", "s56", "s57", "s58", "s59", "s60", "s61", \
      "s62", "s63", "vcc", "scc", "memory"
#define NBS_F32_CLOBBERS_W64 NBS_WALK_CLOBBERS

// One loop: force law FORCE, counting form CM (kCountLanes adds to cnt on
// the interacting lanes, kCountWave adds the interacting lanes' number to
// tot, kCountNone counts nothing), record format FMT.
#define NBS_F32_ASM(FORCE, OUTS, FMT)                                                     \
  if constexpr (CM == kCountLanes)                                                        \
    __asm__ volatile(NBS_WALK_F32_LOOP(FORCE, NBS_CNT_LANE, FMT)                           \
                     : OUTS, [cnt] "+v"(cnt) : NBS_F32_INS_##FMT : NBS_F32_CLOBBERS_##FMT);  \
  else if constexpr (CM == kCountWave)                                                    \
    __asm__ volatile(NBS_WALK_F32_LOOP(FORCE, NBS_CNT_WAVE, FMT)                           \
                     : OUTS, [tot] "+s"(tot), [tc] "=&s"(tc)                              \
                     : NBS_F32_INS_##FMT : NBS_F32_CLOBBERS_##FMT);                       \
  else                                                                                    \
    __asm__ volatile(NBS_WALK_F32_LOOP(FORCE, NBS_CNT_NONE, FMT)                           \
                     : OUTS : NBS_F32_INS_##FMT : NBS_F32_CLOBBERS_##FMT);

// W32: `nodes` is the 32-byte record array (rbase, rmask: the opening radii's
// root value and mask); otherwise the 64-byte one.  The walk starts at node
// n with the lanes of A and the stack (top entries, the top's position
// tpos, entry k in lane k of sp, slo, shi; lane 63 of sp -1) a piece's
// entry left (empty for a whole traversal) and ends at nend.
template <int CM, bool W32, class Nodes>
__device__ __forceinline__ void walk_loop_f32(Nodes nodes, int n, int nend, int pbeg,
                                              unsigned long long A, int top, int tpos, int sp,
                                              int slo, int shi, double x, double y,
                                              double z, float& ax, float& ay, float& az,
                                              int& cnt, int& tot, int& overflow, unsigned rbase,
                                              unsigned rmask) {
  int lim, t, gm, m0save, tc, sk;
  double dx, dy, dz;
  float fx, fy, fz, d2, tv;
  if constexpr (W32) {
    NBS_F32_ASM(NBS_WALK_FORCE_PLAIN, NBS_F32_OUTS, W32)
  } else {
    NBS_F32_ASM(NBS_WALK_FORCE_PLAIN, NBS_F32_OUTS, W64)
  }
}

template <int CM, bool W32, class Nodes>
__device__ __forceinline__ void walk_loop_f32_damped(Nodes nodes, int n, int nend, int pbeg,
                                                     unsigned long long A, int top, int tpos,
                                                     int sp, int slo, int shi, double x,
                                                     double y, double z, float& ax, float& ay,
                                                     float& az, int& cnt, int& tot,
                                                     int& overflow, unsigned rbase,
                                                     unsigned rmask) {
  int lim, t, gm, m0save, tc, sk;
  double dx, dy, dz;
  float fx, fy, fz, d2, tv, te, tn, td, q0, q1, q2, q3, q4;
  if constexpr (W32) {
    NBS_F32_ASM(NBS_WALK_FORCE_DAMPED, NBS_F32_DAMPED_OUTS, W32)
  } else {
    NBS_F32_ASM(NBS_WALK_FORCE_DAMPED, NBS_F32_DAMPED_OUTS, W64)
  }
}
#undef NBS_F32_ASM
#undef NBS_F32_OUTS
#undef NBS_F32_DAMPED_OUTS
#undef NBS_F32_INS_W64
#undef NBS_F32_INS_W32
#undef NBS_F32_CLOBBERS_W32
#undef NBS_F32_CLOBBERS_W64
#undef NBS_WALK_F32_LOOP
#undef NBS_WALK_TAIL
#undef NBS_REC_LOADSK_W64
#undef NBS_REC_LOADSK_W32
#undef NBS_WALK_FORCE_DAMPED
#undef NBS_WALK_FORCE_PLAIN

// The FP64 walk loop (no damping, one particle per lane) written out for
// gfx950, the FP64 counterpart of walk_loop_f32 with the same registers,
// traversal and stack.  Per visit: 3 FP64 subtracts, d^2 fused (3), and
// for a cell the margin test below (a conversion and two compares); for
// the interacting lanes G m / |d|^3 from v_rsq_f64 with a first-order
// correction, and the force (10), the operations and their order those of
// the generic loop (interact<false, false>).
//
// The margin test settles every lane whose d^2 is clearly on one side of
// the opening boundary r2 = (side / theta)^2, with no FP64 arithmetic
// beyond the conversion: d^2 rounded to f32 is compared with the record's
// r2 in f32 lowered and raised by 8 units in the last place (integer
// arithmetic on its bits, on the scalar unit).  Both roundings to f32 are
// within half a unit of the values, so a lane below the lowered bound is
// inside the exact boundary and one above the raised bound outside it, in
// the normal and the subnormal range alike; a bound that wraps to a NaN
// pattern (an r2 of 0 or infinity) decides nothing on its side.  The force
// takes G m / |d|^3 as (G m r^3)(1 + 1.5 e) from the hardware estimate
// r = v_rsq_f64(d^2) and e = 1 - d^2 r^2 (inv_cube_gm).  When a
// lane that takes a cell is between the bounds, or its d^2 is a NaN, the
// loop returns 1 before touching the cell (n, A and the stack describe the
// walk at that cell); the caller then visits it with the reference's exact
// test (walk_visit_f64) and calls the loop again.  So the decisions are
// the reference's whatever the margin.  It returns 0 when the walk is
// done.  The record's FP64 mass-times-G is copied to `gm` before the next
// record is requested.
#define NBS_WALK64_FORCE(MASK, CNT)                                   \
  "s_and_saveexec_b64 s[40:41], " MASK "\n\t"                          \
  "v_rsq_f64 %[tv], %[d2]\n\t"                                        \
  CNT(MASK)                                                           \
  "v_mul_f64 %[s2], %[tv], %[tv]\n\t"                                 \
  "v_fma_f64 %[tmp], -%[d2], %[s2], 1.0\n\t"                          \
  "v_mul_f64 %[s2], %[s2], %[tv]\n\t"                                 \
  "v_mul_f64 %[s2], %[gm], %[s2]\n\t"                                 \
  "v_fma_f64 %[tmp], %[tmp], 0.5, %[tmp]\n\t"                         \
  "v_fma_f64 %[tmp], %[s2], %[tmp], %[s2]\n\t"                        \
  "v_fma_f64 %[ax], %[tmp], %[dx], %[ax]\n\t"                         \
  "v_fma_f64 %[ay], %[tmp], %[dy], %[ay]\n\t"                         \
  "v_fma_f64 %[az], %[tmp], %[dz], %[az]\n\t"                         \
  "s_mov_b64 exec, s[40:41]\n\t"                                      \
  "s_cmp_lt_u32 %[n], %[lim]\n\t"                                     \
  "s_cbranch_scc1 L_loop%=\n\t"                                       \
  "s_branch L_slow%=\n"

#define NBS_WALK64_TAIL(MASK, CNT)                                    \
  "s_lshl_b32 %[t], %[n], 6\n\t"                                      \
  "s_load_dwordx16 s[48:63], %[base], %[t]\n\t"                       \
  NBS_WALK64_FORCE(MASK, CNT)

// d^2 = fma(dx, dx, fma(dy, dy, dz * dz)) (sq_sum_fused)
#define NBS_WALK64_D2                                                 \
  "v_mul_f64 %[d2], %[dz], %[dz]\n\t"                                 \
  "v_fma_f64 %[d2], %[dy], %[dy], %[d2]\n\t"                          \
  "v_fma_f64 %[d2], %[dx], %[dx], %[d2]\n\t"

#define NBS_WALK64_LOOP(CNT) \
      "s_mov_b32 %[m0save], m0\n\t" \
      "s_mov_b32 %[tie], 0\n\t" \
      "s_mov_b64 s[46:47], %[A]\n\t" \
      "s_min_u32 %[lim], %[tpos], %[nend]\n\t" \
      "s_cmp_lt_i32 %[n], %[nend]\n\t" \
      "s_cbranch_scc0 L_done%=\n\t" \
      "s_lshl_b32 %[t], %[n], 6\n\t" \
      "s_load_dwordx16 s[48:63], %[base], %[t]\n" \
      "L_loop%=:\n\t" \
      "s_waitcnt lgkmcnt(0)\n\t" \
      "s_cmp_gt_i32 s55, -1\n\t" \
      "s_cbranch_scc1 L_leaf%=\n\t" \
      "v_add_f64 %[dx], s[48:49], -%[x]\n\t" \
      "v_add_f64 %[dy], s[50:51], -%[y]\n\t" \
      "v_add_f64 %[dz], s[52:53], -%[z]\n\t" \
      "s_mov_b64 %[gm], s[56:57]\n\t" \
      NBS_WALK64_D2 \
      /* cell: open where f32(d^2) < r2 - 8 ulp, accept where it is */ \
      /* > r2 + 8 ulp (s60, s62: the record's unused gm32 and mass); a */ \
      /* taking lane in neither (or NaN) needs the exact test.  A bucket */ \
      /* (r2 infinite: its lower bound the largest float, its upper one */ \
      /* a NaN) is opened here by every lane with a finite d^2 */ \
      "v_cvt_f32_f64 %[f2], %[d2]\n\t" \
      "s_sub_u32 s60, s61, 8\n\t" \
      "s_add_u32 s62, s61, 8\n\t" \
      "v_cmp_gt_f32 vcc, s60, %[f2]\n\t" \
      "s_and_b64 s[42:43], vcc, s[46:47]\n\t" \
      "v_cmp_lt_f32 s[38:39], s62, %[f2]\n\t" \
      "s_or_b64 s[38:39], s[38:39], vcc\n\t" \
      "s_andn2_b64 s[38:39], s[46:47], s[38:39]\n\t" \
      "s_cbranch_scc1 L_tie%=\n\t" \
      "s_cmp_lg_u64 s[42:43], 0\n\t" \
      "s_cbranch_scc1 L_opened%=\n\t" \
      "s_add_i32 %[n], %[n], 1\n\t" \
      "s_max_i32 %[n], %[n], s54\n\t" \
      NBS_WALK64_TAIL("s[46:47]", CNT) \
      "L_opened%=:\n\t" \
      "s_add_i32 %[n], %[n], 1\n\t" \
      "s_andn2_b64 s[44:45], s[46:47], s[42:43]\n\t" \
      "s_mov_b64 s[46:47], s[42:43]\n\t" \
      "s_cbranch_scc0 L_noint%=\n\t" \
      "s_cmp_eq_u32 %[tpos], s54\n\t" \
      "s_cbranch_scc1 L_merge%=\n\t" \
      "s_cmp_gt_u32 %[top], 62\n\t" \
      "s_cbranch_scc1 L_ovf%=\n\t" \
      "s_mov_b32 m0, %[top]\n\t" \
      "v_writelane_b32 %[sp], s54, m0\n\t" \
      "v_writelane_b32 %[slo], s44, m0\n\t" \
      "v_writelane_b32 %[shi], s45, m0\n\t" \
      "s_add_u32 %[top], %[top], 1\n\t" \
      "s_mov_b32 %[tpos], s54\n\t" \
      "s_min_u32 %[lim], %[tpos], %[nend]\n\t" \
      "s_branch L_int%=\n" \
      "L_merge%=:\n\t" \
      "s_sub_u32 m0, %[top], 1\n\t" \
      "v_readlane_b32 s38, %[slo], m0\n\t" \
      "v_readlane_b32 s39, %[shi], m0\n\t" \
      "s_or_b64 s[38:39], s[38:39], s[44:45]\n\t" \
      "v_writelane_b32 %[slo], s38, m0\n\t" \
      "v_writelane_b32 %[shi], s39, m0\n\t" \
      "s_branch L_int%=\n" \
      "L_ovf%=:\n\t" \
      "s_mov_b32 %[ovf], 1\n\t" \
      "s_branch L_int%=\n" \
      /* no lane interacts: the next record, no force */ \
      "L_noint%=:\n\t" \
      "s_lshl_b32 %[t], %[n], 6\n\t" \
      "s_load_dwordx16 s[48:63], %[base], %[t]\n" \
      "L_next%=:\n\t" \
      "s_cmp_lt_u32 %[n], %[lim]\n\t" \
      "s_cbranch_scc1 L_loop%=\n\t" \
      "s_branch L_slow%=\n" \
      "L_int%=:\n\t" \
      NBS_WALK64_TAIL("s[44:45]", CNT) \
      "L_leaf%=:\n\t" \
      "v_add_f64 %[dx], s[48:49], -%[x]\n\t" \
      "v_add_f64 %[dy], s[50:51], -%[y]\n\t" \
      "v_add_f64 %[dz], s[52:53], -%[z]\n\t" \
      "s_mov_b64 %[gm], s[56:57]\n\t" \
      "s_sub_u32 %[t], s55, %[pbeg]\n\t" \
      "s_lshl_b64 s[38:39], 1, %[t]\n\t" \
      "s_cmp_lt_u32 %[t], 64\n\t" \
      "s_cselect_b64 s[38:39], s[38:39], 0\n\t" \
      "s_andn2_b64 s[44:45], s[46:47], s[38:39]\n\t" \
      "s_add_i32 %[n], %[n], 1\n\t" \
      "s_lshl_b32 %[t], %[n], 6\n\t" \
      "s_load_dwordx16 s[48:63], %[base], %[t]\n\t" \
      "s_cmp_eq_u64 s[44:45], 0\n\t" \
      "s_cbranch_scc1 L_next%=\n\t" \
      NBS_WALK64_D2 \
      NBS_WALK64_FORCE("s[44:45]", CNT) \
      "L_tie%=:\n\t" \
      "s_mov_b32 %[tie], 1\n\t" \
      "s_branch L_end%=\n" \
      "L_slow%=:\n\t" \
      "s_cmp_eq_u32 %[n], %[tpos]\n\t" \
      "s_cbranch_scc0 L_end%=\n\t" \
      "s_sub_u32 %[top], %[top], 1\n\t" \
      "v_readlane_b32 s38, %[slo], %[top]\n\t" \
      "v_readlane_b32 s39, %[shi], %[top]\n\t" \
      "s_add_u32 %[t], %[top], 63\n\t" \
      "s_and_b32 %[t], %[t], 63\n\t" \
      "v_readlane_b32 %[tpos], %[sp], %[t]\n\t" \
      "s_or_b64 s[46:47], s[46:47], s[38:39]\n\t" \
      "s_min_u32 %[lim], %[tpos], %[nend]\n\t" \
      "s_cmp_lt_u32 %[n], %[lim]\n\t" \
      "s_cbranch_scc1 L_loop%=\n" \
      "L_end%=:\n\t" \
      "s_waitcnt lgkmcnt(0)\n" \
      "L_done%=:\n\t" \
      "s_mov_b64 %[A], s[46:47]\n\t" \
      "s_mov_b32 m0, %[m0save]"

#define NBS_F64_OUTS                                                                      \
  [n] "+s"(n), [ovf] "+s"(overflow), [A] "+s"(A), [top] "+s"(top), [tpos] "+s"(tpos),     \
      [ax] "+v"(ax), [ay] "+v"(ay), [az] "+v"(az), [sp] "+v"(sp), [slo] "+v"(slo),        \
      [shi] "+v"(shi), [lim] "=&s"(lim), [t] "=&s"(t), [tie] "=&s"(tie), [gm] "=&s"(gm),  \
      [m0save] "=&s"(m0save), [dx] "=&v"(dx), [dy] "=&v"(dy), [dz] "=&v"(dz),              \
      [d2] "=&v"(d2), [tmp] "=&v"(tmp), [tv] "=&v"(tv), [s2] "=&v"(s2), [f2] "=&v"(f2)
#define NBS_F64_INS                                                                       \
  [base] "s"(nodes), [nend] "s"(nend), [pbeg] "s"(pbeg), [x] "v"(x), [y] "v"(y),          \
      [z] "v"(z)

template <int CM>
__device__ __forceinline__ int walk_loop_f64(ConstNodes nodes, int& n, int nend, int pbeg,
                                             unsigned long long& A, int& top, int& tpos,
                                             int& sp, int& slo, int& shi, double x, double y,
                                             double z, double& ax, double& ay, double& az, int& cnt,
                                             int& tot, int& overflow) {
  int lim, t, m0save, tie, tc;
  double gm, dx, dy, dz, d2, tmp, tv, s2;
  float f2;
  if constexpr (CM == kCountLanes)
    __asm__ volatile(NBS_WALK64_LOOP(NBS_CNT_LANE)
                     : NBS_F64_OUTS, [cnt] "+v"(cnt) : NBS_F64_INS : NBS_WALK_CLOBBERS);
  else if constexpr (CM == kCountWave)
    __asm__ volatile(NBS_WALK64_LOOP(NBS_CNT_WAVE)
                     : NBS_F64_OUTS, [tot] "+s"(tot), [tc] "=&s"(tc)
                     : NBS_F64_INS : NBS_WALK_CLOBBERS);
  else
    __asm__ volatile(NBS_WALK64_LOOP(NBS_CNT_NONE)
                     : NBS_F64_OUTS : NBS_F64_INS : NBS_WALK_CLOBBERS);
  return tie;
}
#undef NBS_F64_OUTS
#undef NBS_F64_INS
#undef NBS_WALK64_LOOP

#undef NBS_WALK64_D2
#undef NBS_WALK64_TAIL
#undef NBS_WALK64_FORCE
#undef NBS_CNT_LANE
#undef NBS_CNT_WAVE
#undef NBS_CNT_NONE
#undef NBS_WALK_CLOBBERS

// One visit of the FP64 walk (no damping, one particle per lane) at node
// n with the exact opening test: the generic loop's body for one node,
// used by the hand-written loop for the rare cell where a lane is within
// rounding of the opening boundary.
__device__ __forceinline__ void walk_visit_f64(ConstNodes nodes, int& n, int i,
                                               unsigned long long& A, int& top, int& top_pos,
                                               int& st_pos, int& st_lo, int& st_hi,
                                               int& overflow, double x, double y, double z,
                                               double theta, double theta2, double& ax,
                                               double& ay, double& az, int& cnt) {
  const int lane = lane_id();
  const NodeRegs cur = load_node(nodes, n);
  const double dx = cur.cx - x, dy = cur.cy - y, dz = cur.cz - z;
  const double d2 = sq_sum_ref(dx, dy, dz);
  const bool in = __builtin_amdgcn_inverse_ballot_w64(A);
  bool inter;
  int next;
  if (cur.leaf >= 0) {
    inter = in && cur.leaf != i;
    next = n + 1;
  } else {
    bool open = cur.leaf == -2 ? true : mac_open_f64(d2, cur.side, theta, theta2);
    open = open && in;
    const unsigned long long O = ballot(open);
    inter = in && !open;
    if (O) {
      const unsigned long long C = A & ~O;
      if (C) {
        if (top > 0 && top_pos == cur.skip) {
          const int slot = top - 1;
          st_lo = lane == slot ? (st_lo | (int)(unsigned)C) : st_lo;
          st_hi = lane == slot ? (st_hi | (int)(unsigned)(C >> 32)) : st_hi;
        } else if (top < kWave - 1) {
          const int slot = top;
          st_pos = lane == slot ? cur.skip : st_pos;
          st_lo = lane == slot ? (int)(unsigned)C : st_lo;
          st_hi = lane == slot ? (int)(unsigned)(C >> 32) : st_hi;
          top_pos = cur.skip;
          top++;
        } else {
          overflow = 1;
        }
      }
      A = O;
      next = n + 1;
    } else {
      next = cur.skip;
    }
  }
  if (inter) interact<false, false>(cur, dx, dy, dz, 0.f, 0.f, 0.f, 0.f, ax, ay, az, cnt);
  n = max(next, n + 1);
  if (top > 0 && top_pos == n) {
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane(st_lo, top - 1);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane(st_hi, top - 1);
    A |= ((unsigned long long)hi << 32) | lo;
    top--;
    top_pos = top > 0 ? __builtin_amdgcn_readlane(st_pos, top - 1) : -1;
  }
}

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

// Where a piece of a wavefront's traversal starts.  With split > 1 the
// system's node range is cut into `split` equal ranges of pre-order index
// and each wavefront walks each range as a separate piece (more wavefronts
// for a system too small to fill the GPU).  The walk state at the first
// node b of a range is what the whole traversal has there: descending from
// the system's root to b, each cell on the path is opened or accepted by
// the lanes that take it with the walk's own test, lanes that accept one
// wait at its skip (the stack, merged as the walk merges), and the rest
// take b.  Each lane's decisions, and so its interactions and their order
// within the range, are the whole traversal's; a piece only adds the
// range's contribution.  When no lane takes b the piece resumes at the
// stack top, or has nothing to do (n = e).
template <bool F32, bool FROZEN, int P>
__device__ __forceinline__ void walk_enter(ConstNodes nodes, ConstNodes mac_nodes, int root,
                                           int b, int e, const double* x, const double* y,
                                           const double* z, const double* mx,
                                           const double* my, const double* mz, double theta,
                                           double theta2, unsigned long long* A, int& n,
                                           int& top, int& top_pos, int& st_pos, int* st_lo,
                                           int* st_hi, bool& overflow) {
  const int lane = lane_id();
  int X = root;
  while (X < b) {
    const NodeRegs cur = load_node(nodes, X);
    unsigned long long O[P], C[P], any_wait = 0;
#pragma unroll
    for (int p = 0; p < P; p++) {
      bool open;
      if (cur.leaf == -2) {
        open = true;
      } else if (F32) {
        const float fdx = (float)(cur.cx - x[p]), fdy = (float)(cur.cy - y[p]),
                    fdz = (float)(cur.cz - z[p]);
        open = fmaf(fdx, fdx, fmaf(fdy, fdy, fdz * fdz)) < cur.r2open32;
      } else {
        const double d2 = FROZEN ? sq_sum_ref(mac_nodes[X].cx - mx[p], mac_nodes[X].cy - my[p],
                                              mac_nodes[X].cz - mz[p])
                                 : sq_sum_ref(cur.cx - x[p], cur.cy - y[p], cur.cz - z[p]);
        open = mac_open_f64(d2, cur.side, theta, theta2);
      }
      O[p] = ballot(open && __builtin_amdgcn_inverse_ballot_w64(A[p]));
      C[p] = A[p] & ~O[p];
      any_wait |= C[p];
    }
    if (any_wait) {
      if (top > 0 && top_pos == cur.skip) {
        const int slot = top - 1;
#pragma unroll
        for (int p = 0; p < P; p++) {
          st_lo[p] = lane == slot ? (st_lo[p] | (int)(unsigned)C[p]) : st_lo[p];
          st_hi[p] = lane == slot ? (st_hi[p] | (int)(unsigned)(C[p] >> 32)) : st_hi[p];
        }
      } else if (top < kWave - 1) {
        const int slot = top;
        st_pos = lane == slot ? cur.skip : st_pos;
#pragma unroll
        for (int p = 0; p < P; p++) {
          st_lo[p] = lane == slot ? (int)(unsigned)C[p] : st_lo[p];
          st_hi[p] = lane == slot ? (int)(unsigned)(C[p] >> 32) : st_hi[p];
        }
        top_pos = cur.skip;
        top++;
      } else {
        overflow = true;
      }
    }
#pragma unroll
    for (int p = 0; p < P; p++) A[p] = O[p];
    // the child of X whose subtree holds b (siblings are chained by skip)
    int c = X + 1;
    for (;;) {
      const int sk = max(load_node(nodes, c).skip, c + 1);
      if (sk > b) break;
      c = sk;
    }
    X = c;
  }
  n = b;
  unsigned long long any = 0;
#pragma unroll
  for (int p = 0; p < P; p++) any |= A[p];
  if (!any) {
    if (top > 0 && top_pos < e) {
      n = top_pos;
#pragma unroll
      for (int p = 0; p < P; p++) {
        const unsigned lo = (unsigned)__builtin_amdgcn_readlane(st_lo[p], top - 1);
        const unsigned hi = (unsigned)__builtin_amdgcn_readlane(st_hi[p], top - 1);
        A[p] = ((unsigned long long)hi << 32) | lo;
      }
      top--;
      top_pos = top > 0 ? __builtin_amdgcn_readlane(st_pos, top - 1) : -1;
    } else {
      n = e;
    }
  }
}

// One wavefront walks P * 64 particles that are consecutive in octree
// order (lane l holds particles pbeg + l, pbeg + 64 + l, ...), each set
// with its own lane mask; a node is visited if any particle of any set
// needs it.  P = 2 halves the node loads and traversal bookkeeping per
// particle at the price of a larger union of paths.
//
// FROZEN (FP64 only): the MAC is evaluated on a second tree (mac_nodes,
// same topology) against second positions (mac_x/y/z) -- a recorded base
// trajectory -- while forces use this step's tree and positions.  That
// makes the step a smooth function of the perturbed state, for finite
// differences of the adjoint.
//
// Eight wavefronts per SIMD: gfx950 holds 640 SGPRs per SIMD, so eight
// wavefronts need at most 80 each.  (The FP64 walks with damping or two
// particles per lane would spill to scratch within that bound: they keep
// their registers.)  With the scalar loop's fixed registers
// at s38-s63 and the compiler held to eight wavefronts (it moves the few
// values it cannot keep into VGPR lanes outside the loop) the walk uses 78;
// at the 95 it used before, only seven fitted and the FP32 walk took 10 %
// longer.
template <bool F32, bool DAMP, bool FROZEN, int BLOCK, int P, bool W32 = false,
          bool SPLIT = false>
__global__ void __launch_bounds__(BLOCK)
__attribute__((amdgpu_waves_per_eu(!F32 && (DAMP || P == 2) ? 1 : 8, 8)))
walk_kernel(WalkArgs a) {
  if (*a.flags) return;
  const int lane = lane_id();
  // split > 1: wavefront q walks piece q % split of traversal q / split
  // (the pieces of one traversal on consecutive workgroups, which the
  // hardware hands to different XCDs: each XCD walks its own node range)
  const int split = SPLIT ? a.split : 1;
  int tile = 0, hw, piece = 0;
  if constexpr (SPLIT) {
    const int q = (int)((blockIdx.x * BLOCK + threadIdx.x) >> 6);
    hw = __builtin_amdgcn_readfirstlane(q / split);
    piece = __builtin_amdgcn_readfirstlane(q - (q / split) * split);
  } else {
    tile = a.order ? a.order[blockIdx.x] : tile_of_block(blockIdx.x, gridDim.x, a.groups);
    hw = __builtin_amdgcn_readfirstlane((int)((tile * BLOCK + threadIdx.x) >> 6));
  }
  const int* hw_start = P == 1 ? a.wave_start : a.wave2_start;
  if (hw >= (P == 1 ? a.n_waves : a.n_waves2)) return;
  const int s = a.nsys > 1 ? __builtin_amdgcn_readfirstlane(
                                 sys_of_wave(hw, hw_start, a.nsys))
                           : 0;
  const int k = hw - hw_start[s];          // wavefront within the system
  const int w0 = a.wave_start[s] + k * P;  // its first 64-particle wave
  const int wend = a.wave_start[s + 1];
  const int pbeg = a.sys_start[s] + k * (kWave * P);
  const int pend = a.sys_start[s + 1];

  int i[P];
  bool valid[P];
  double x[P], y[P], z[P], mx[P], my[P], mz[P];
  // Lane masks, wave-uniform (SGPRs): A[p] = lanes of set p that take the
  // current node.
  unsigned long long A[P];
#pragma unroll
  for (int p = 0; p < P; p++) {
    i[p] = pbeg + p * kWave + lane;
    valid[p] = i[p] < pend;
    x[p] = valid[p] ? a.x[i[p]] : 0.0;
    y[p] = valid[p] ? a.y[i[p]] : 0.0;
    z[p] = valid[p] ? a.z[i[p]] : 0.0;
    mx[p] = FROZEN && valid[p] ? a.mac_x[i[p]] : 0.0;
    my[p] = FROZEN && valid[p] ? a.mac_y[i[p]] : 0.0;
    mz[p] = FROZEN && valid[p] ? a.mac_z[i[p]] : 0.0;
    A[p] = ballot(valid[p]);
  }
  ConstNodes nodes = (ConstNodes)a.nodes;
  ConstNodes mac_nodes = (ConstNodes)a.mac_nodes;
  int n = a.node_start[s];
  int nend = a.node_start[s + 1];
  const double theta = a.theta;
  const double theta2 = theta * theta;

  bool overflow = false;
  // the stack of waiting lanes: entry `slot` lives in lane `slot` of
  // st_pos, st_lo, st_hi (lane 63 of st_pos stays -1); top_pos is the top
  // entry's position, kept in an SGPR
  int top = 0, top_pos = -1, st_pos = -1;
  int st_lo[P], st_hi[P];
#pragma unroll
  for (int p = 0; p < P; p++) st_lo[p] = st_hi[p] = 0;
  if constexpr (SPLIT) {
    const long long len = nend - n;
    const int b = n + (int)(len * piece / split);
    const int e = n + (int)(len * (piece + 1) / split);
    walk_enter<F32, FROZEN, P>(nodes, mac_nodes, n, b, e, x, y, z, mx, my, mz, theta, theta2,
                               A, n, top, top_pos, st_pos, st_lo, st_hi, overflow);
    nend = e;
  }
  if constexpr (SPLIT) {
    top = __builtin_amdgcn_readfirstlane(top);
    top_pos = __builtin_amdgcn_readfirstlane(top_pos);
  }

  typedef typename std::conditional<F32, float, double>::type acc_t;
  acc_t ax[P], ay[P], az[P];
  int cnt[P];
#pragma unroll
  for (int p = 0; p < P; p++) {
    ax[p] = ay[p] = az[p] = 0;
    cnt[p] = 0;
  }

  unsigned visits = 0;
  // interactions the hand-written loops counted on the scalar unit
  // (kCountWave); cnt holds the per-lane counts of every other path
  int wave_count = 0;
  // The FP32 hand-written loop reads the 32-byte records in the W32
  // variant (launched when the context keeps them), else the 64-byte
  // ones.  A 32-byte cell carries its depth
  // instead of (side / theta)^2 in f32: that is the root's value
  // r0 = (float)((side_0 / theta)^2) times 4^-depth exactly (dividing a
  // power-of-two multiple of side_0 by theta and squaring it scale
  // exactly, in FP64 and in the rounding to f32) while every value down to
  // depth 41 is a normal float: r0's exponent field at least 85.  So the
  // loop subtracts 2 * depth from r0's exponent.  theta = 0 makes r0
  // infinite, as every cell's value is: rmask then keeps r0.  Any other
  // root (an overflowing r0, or one too small) takes the generic loop.
  // (the hand-written loops do not count visits: per-tile costs come from
  // the generic loop)
  unsigned rbase = 0, rmask = 0;
  bool hand32 = F32 && !FROZEN && P == 1 && a.asm_loop && !a.tile_cost;
  if constexpr (W32) {
    if (hand32) {
      const double ro = a.root[4 * s + 3] / a.theta;
      rbase = __float_as_uint((float)(ro * ro));
      const unsigned ex = (rbase >> 23) & 255u;
      if (a.theta == 0.0 && rbase == 0x7f800000u) {
        rmask = 0u;
      } else if (ex >= 85u && ex < 255u) {
        rmask = ~0u;
      } else {
        hand32 = false;
      }
      rbase = __builtin_amdgcn_readfirstlane(rbase);
      rmask = __builtin_amdgcn_readfirstlane(rmask);
    }
  }
  if (hand32) {
    int ovf = overflow ? 1 : 0;
    float fax = 0.f, fay = 0.f, faz = 0.f;
    ConstNodes32 nodes32 = (ConstNodes32)a.nodes32;
#define NBS_F32_WALK(CM, W, ND)                                                             \
  if constexpr (DAMP)                                                                       \
    walk_loop_f32_damped<CM, W>(ND, n, nend, pbeg, A[0], top, top_pos, st_pos, st_lo[0],    \
                                st_hi[0], x[0], y[0], z[0], fax, fay, faz, cnt[0],          \
                                wave_count, ovf, rbase, rmask);                             \
  else                                                                                      \
    walk_loop_f32<CM, W>(ND, n, nend, pbeg, A[0], top, top_pos, st_pos, st_lo[0], st_hi[0], \
                         x[0], y[0], z[0], fax, fay, faz, cnt[0], wave_count, ovf, rbase,   \
                         rmask);
#define NBS_F32_WALK_FMT(CM)                                                                \
  if constexpr (W32) {                                                                      \
    NBS_F32_WALK(CM, true, nodes32)                                                         \
  } else {                                                                                  \
    NBS_F32_WALK(CM, false, nodes)                                                          \
  }
    if (a.count_mode == kCountLanes) {
      NBS_F32_WALK_FMT(kCountLanes)
    } else if (a.count_mode == kCountWave) {
      NBS_F32_WALK_FMT(kCountWave)
    } else {
      NBS_F32_WALK_FMT(kCountNone)
    }
#undef NBS_F32_WALK_FMT
#undef NBS_F32_WALK
    ax[0] = fax;
    ay[0] = fay;
    az[0] = faz;
    overflow = ovf != 0;
  } else if (!F32 && !DAMP && !FROZEN && P == 1 && a.asm_loop && !a.tile_cost) {
    int ovf = overflow ? 1 : 0, slo = st_lo[0], shi = st_hi[0];
    double fax = 0, fay = 0, faz = 0;
    unsigned long long Acur = A[0];
#define NBS_F64_WALK(CM)                                                                    \
  while (walk_loop_f64<CM>(nodes, n, nend, pbeg, Acur, top, top_pos, st_pos, slo, shi,      \
                           x[0], y[0], z[0], fax, fay, faz, cnt[0], wave_count,             \
                           ovf))                                                            \
    walk_visit_f64(nodes, n, i[0], Acur, top, top_pos, st_pos, slo, shi, ovf, x[0], y[0],     \
                   z[0], theta, theta2, fax, fay, faz, cnt[0]);
    if (a.count_mode == kCountLanes) {
      NBS_F64_WALK(kCountLanes)
    } else if (a.count_mode == kCountWave) {
      NBS_F64_WALK(kCountWave)
    } else {
      NBS_F64_WALK(kCountNone)
    }
#undef NBS_F64_WALK
    ax[0] = fax;
    ay[0] = fay;
    az[0] = faz;
    overflow = ovf != 0;
  } else {
    // One scalar load of the node record per visit.  (Loading both possible
    // successors a visit ahead measured slower: the doubled scalar-memory
    // traffic costs more than the latency it hides.)
    while (n < nend) {
      visits++;
      const NodeRegs cur = load_node(nodes, n);
      const int leaf = cur.leaf;
      const int skip = cur.skip;
      double dx[P], dy[P], dz[P], d2[P];
      float fdx[P], fdy[P], fdz[P], fd2[P];
      bool in[P], inter[P];
#pragma unroll
      for (int p = 0; p < P; p++) {
        dx[p] = cur.cx - x[p];
        dy[p] = cur.cy - y[p];
        dz[p] = cur.cz - z[p];
        in[p] = __builtin_amdgcn_inverse_ballot_w64(A[p]);  // the mask itself
        if (F32) {
          fdx[p] = (float)dx[p];
          fdy[p] = (float)dy[p];
          fdz[p] = (float)dz[p];
          fd2[p] = fmaf(fdx[p], fdx[p], fmaf(fdy[p], fdy[p], fdz[p] * fdz[p]));
          d2[p] = 0;
        } else {
          fdx[p] = fdy[p] = fdz[p] = fd2[p] = 0;
          d2[p] = sq_sum_ref(dx[p], dy[p], dz[p]);
        }
      }
      int next;
      if (leaf >= 0) {
#pragma unroll
        for (int p = 0; p < P; p++) inter[p] = in[p] && leaf != i[p];
        next = n + 1;
      } else {
        unsigned long long O[P], any_open = 0;
#pragma unroll
        for (int p = 0; p < P; p++) {
          bool open;
          if (leaf == -2)
            open = true;
          else if (F32)
            open = fd2[p] < cur.r2open32;
          else
            open = mac_open_f64(
                FROZEN ? sq_sum_ref(mac_nodes[n].cx - mx[p], mac_nodes[n].cy - my[p],
                                    mac_nodes[n].cz - mz[p])
                       : d2[p],
                cur.side, theta, theta2);
          open = open && in[p];
          O[p] = ballot(open);
          inter[p] = in[p] && !open;
          any_open |= O[p];
        }
        if (any_open) {
          unsigned long long C[P], any_wait = 0;
#pragma unroll
          for (int p = 0; p < P; p++) {
            C[p] = A[p] & ~O[p];
            any_wait |= C[p];
          }
          if (any_wait) {
            if (top > 0 && top_pos == skip) {
              const int slot = top - 1;
#pragma unroll
              for (int p = 0; p < P; p++) {
                st_lo[p] = lane == slot ? (st_lo[p] | (int)(unsigned)C[p]) : st_lo[p];
                st_hi[p] = lane == slot ? (st_hi[p] | (int)(unsigned)(C[p] >> 32)) : st_hi[p];
              }
            } else {
              const int slot = top;
              st_pos = lane == slot ? skip : st_pos;
#pragma unroll
              for (int p = 0; p < P; p++) {
                st_lo[p] = lane == slot ? (int)(unsigned)C[p] : st_lo[p];
                st_hi[p] = lane == slot ? (int)(unsigned)(C[p] >> 32) : st_hi[p];
              }
              top_pos = skip;
              if (top < kWave - 1)
                top++;
              else
                overflow = true;
            }
          }
#pragma unroll
          for (int p = 0; p < P; p++) A[p] = O[p];
          next = n + 1;
        } else {
          next = skip;
        }
      }
#pragma unroll
      for (int p = 0; p < P; p++)
        if (inter[p])
          interact<F32, DAMP>(cur, dx[p], dy[p], dz[p], fdx[p], fdy[p], fdz[p], fd2[p],
                              ax[p], ay[p], az[p], cnt[p]);
      n = max(next, n + 1);  // always advance, even on a corrupt skip
      if (top > 0 && top_pos == n) {
#pragma unroll
        for (int p = 0; p < P; p++) {
          const unsigned lo = (unsigned)__builtin_amdgcn_readlane(st_lo[p], top - 1);
          const unsigned hi = (unsigned)__builtin_amdgcn_readlane(st_hi[p], top - 1);
          A[p] |= ((unsigned long long)hi << 32) | lo;
        }
        top--;
        top_pos = top > 0 ? __builtin_amdgcn_readlane(st_pos, top - 1) : -1;
      }
    }
  }
  if (overflow && lane == 0) atomicOr(a.flags_out, kFlagStack);

  if constexpr (SPLIT) {
    // a piece: its partial sums and counts, which walk_combine_kernel adds
    // in piece order and finishes (update, bbox partials, step counters)
    acc_t* part = (acc_t*)a.part;
    const size_t np = (size_t)a.n;
    int wt = wave_count;
#pragma unroll
    for (int p = 0; p < P; p++) {
      const int ip = i[p];
      if (valid[p]) {
        part[((size_t)piece * 3 + 0) * np + ip] = ax[p];
        part[((size_t)piece * 3 + 1) * np + ip] = ay[p];
        part[((size_t)piece * 3 + 2) * np + ip] = az[p];
        if (a.count_mode == kCountLanes) a.cpart[(size_t)piece * np + ip] = cnt[p];
      }
      if (a.count_mode == kCountWave) wt += wave_sum(valid[p] ? cnt[p] : 0);
    }
    if (a.count_mode == kCountWave && lane == 0 && wt)
      atomicAdd(&a.inter_total[w0], (unsigned long long)wt);
    return;
  }

  int wtot = 0;
#pragma unroll
  for (int p = 0; p < P; p++) {
    const double fax = (double)ax[p], fay = (double)ay[p], faz = (double)az[p];
    const int ip = i[p];
    if (valid[p]) {
      a.ax[ip] = fax;
      a.ay[ip] = fay;
      a.az[ip] = faz;
      if (a.count_mode == kCountLanes) a.interactions[ip] = cnt[p];
    }
    if (a.fused_update) {
      if (valid[p]) {
        const double dt = a.dt;
        double vx = a.vx[ip], vy = a.vy[ip], vz = a.vz[ip];
        kick_drift(x[p], vx, fax, dt);
        kick_drift(y[p], vy, fay, dt);
        kick_drift(z[p], vz, faz, dt);
        a.x[ip] = x[p];
        a.y[ip] = y[p];
        a.z[ip] = z[p];
        a.vx[ip] = vx;
        a.vy[ip] = vy;
        a.vz[ip] = vz;
      }
      // the next step's bbox partial of 64-particle wave w0 + p (if this
      // system has it)
      if (w0 + p < wend)
        wave_bbox_partial(valid[p], x[p], y[p], z[p], w0 + p, a.n_waves, a.bbox);
    }
    if (a.count_mode != kCountNone) wtot += wave_sum(valid[p] ? cnt[p] : 0);
  }
  wtot += wave_count;
  if (a.fused_update && lane == 0 && k == 0) {
    a.sys_dt[s] = a.dt;
    a.sys_time[s] += a.dt;
  }
  // per-wave running interaction count (no contended atomics)
  if (lane == 0 && a.count_mode != kCountNone) a.inter_total[w0] += (unsigned long long)wtot;
  if (a.tile_cost && lane == 0) atomicAdd(&a.tile_cost[tile], visits);
  if (hw == 0 && lane == 0) atomicAdd(a.steps_done, 1);
}

// The sums of a split walk: per particle the pieces' partial sums added in
// piece order (FP32 sums in FP32), then what the whole walk does after its
// loop: accelerations and counts, the fused kick-drift and the next bbox
// partials, the system's time, the per-wave interaction totals and the
// completed-walk counter.  One thread per particle, 64-particle waves as in
// the walk.
template <bool F32>
__global__ void __launch_bounds__(256) walk_combine_kernel(WalkArgs a) {
  if (*a.flags) return;
  typedef typename std::conditional<F32, float, double>::type acc_t;
  const int lane = lane_id();
  const int hw = __builtin_amdgcn_readfirstlane((int)((blockIdx.x * 256 + threadIdx.x) >> 6));
  if (hw >= a.n_waves) return;
  const int s = a.nsys > 1 ? __builtin_amdgcn_readfirstlane(sys_of_wave(hw, a.wave_start, a.nsys))
                           : 0;
  const int k = hw - a.wave_start[s];
  const int i = a.sys_start[s] + k * kWave + lane;
  const bool valid = i < a.sys_start[s + 1];
  const acc_t* part = (const acc_t*)a.part;
  const size_t np = (size_t)a.n;
  acc_t sx = 0, sy = 0, sz = 0;
  int cnt = 0;
  if (valid) {
    for (int q = 0; q < a.split; q++) {
      sx = sx + part[((size_t)q * 3 + 0) * np + i];
      sy = sy + part[((size_t)q * 3 + 1) * np + i];
      sz = sz + part[((size_t)q * 3 + 2) * np + i];
      if (a.count_mode == kCountLanes) cnt += a.cpart[(size_t)q * np + i];
    }
  }
  const double fax = (double)sx, fay = (double)sy, faz = (double)sz;
  double x = valid ? a.x[i] : 0.0, y = valid ? a.y[i] : 0.0, z = valid ? a.z[i] : 0.0;
  if (valid) {
    a.ax[i] = fax;
    a.ay[i] = fay;
    a.az[i] = faz;
    if (a.count_mode == kCountLanes) a.interactions[i] = cnt;
  }
  if (a.fused_update) {
    if (valid) {
      const double dt = a.dt;
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
    wave_bbox_partial(valid, x, y, z, hw, a.n_waves, a.bbox);
    if (lane == 0 && k == 0) {
      a.sys_dt[s] = a.dt;
      a.sys_time[s] += a.dt;
    }
  }
  if (a.count_mode == kCountLanes) {
    const int wt = wave_sum(valid ? cnt : 0);
    if (lane == 0) a.inter_total[hw] += (unsigned long long)wt;
  }
  if (hw == 0 && lane == 0) atomicAdd(a.steps_done, 1);
}

// ------------------------------------------------- collision detection --
// The pairs the reference's walk records for collision merging
// (src/main.cpp:386-396): particle i and a leaf j that i's walk reaches
// (every cell above j opened by the MAC) with distance(j, i) < range.  A
// pruned walk finds exactly those: it opens a cell when the MAC opens it --
// the reference's FP64 test, whatever the precision of the forces -- and
// the cell's box comes within range of the lane's particle.  The box is
// the octree cell of side s that holds the centre of mass: per axis
// [r0 + k s, r0 + (k + 1) s] with k = floor((com - r0) / s) from the root's
// low corner r0.  Rounding: com - r0 and its product with 1 / s are off by
// a few ulps of the root side S, under 1e-15 S / s cells, and the
// reference's chain of rounded centres moves the boundaries by less than
// 1e-13 S.  So the box is taken one cell wider on a side where the centre
// of mass lies within 1e-6 + 1e-15 S / s cells of a boundary, and widened
// by 1e-6 s + 1e-13 S.  A lane follows about one cell per level.  A fixed
// root can leave particles outside it (and so outside their cells'
// boxes): with a fixed root only the MAC prunes.  Lanes that do not open a
// cell wait at its skip, as in the walk.  A pair is recorded as (min slot,
// max slot), usually once from each end; one atomic per wave.
__device__ __forceinline__ double box_gap(double x, double com, double r0, double s,
                                          double inv_s, double tol, double margin) {
  const double u = (com - r0) * inv_s;
  const double lo = r0 + floor(u - tol) * s - margin;
  const double hi = r0 + (floor(u + tol) + 1.0) * s + margin;
  return x < lo ? lo - x : (x > hi ? x - hi : 0.0);
}

__global__ void __launch_bounds__(256) detect_kernel(DetectArgs a) {
  if (*a.flags) return;
  const int lane = lane_id();
  const int hw = __builtin_amdgcn_readfirstlane((int)((blockIdx.x * 256 + threadIdx.x) >> 6));
  if (hw >= a.n_waves) return;
  const int s = a.nsys > 1 ? __builtin_amdgcn_readfirstlane(sys_of_wave(hw, a.wave_start, a.nsys))
                           : 0;
  const int pbeg = a.sys_start[s] + (hw - a.wave_start[s]) * kWave;
  const int i = pbeg + lane;
  const bool valid = i < a.sys_start[s + 1];
  const double x = valid ? a.x[i] : 0.0, y = valid ? a.y[i] : 0.0, z = valid ? a.z[i] : 0.0;
  unsigned long long A = ballot(valid);
  ConstNodes nodes = (ConstNodes)a.nodes;
  int n = a.node_start[s];
  const int nend = a.node_start[s + 1];
  const double theta = a.theta, theta2 = theta * theta, range = a.range;
  const double rside = a.root[4 * s + 3];
  const double rx = a.root[4 * s] - 0.5 * rside, ry = a.root[4 * s + 1] - 0.5 * rside,
               rz = a.root[4 * s + 2] - 0.5 * rside;
  int st_pos = 0, st_lo = 0, st_hi = 0, top = 0, top_pos = -1;
  bool overflow = false;
  while (n < nend) {
    const NodeRegs cur = load_node(nodes, n);
    const double dx = cur.cx - x, dy = cur.cy - y, dz = cur.cz - z;
    const double d2 = sq_sum_ref(dx, dy, dz);
    const bool in = __builtin_amdgcn_inverse_ballot_w64(A);
    int next;
    if (cur.leaf >= 0) {
      // the correctly rounded root only where d^2 is near range^2 or below
      bool hit = in && cur.leaf != i && d2 < range * range * (1.0 + 1e-9);
      if (hit) hit = __dsqrt_rn(d2) < range;
      const unsigned long long h = ballot(hit);
      if (h) {
        const int first = __builtin_ctzll(h);
        int base = 0;
        if (lane == first) base = atomicAdd(a.pair_count, __popcll(h));
        base = __builtin_amdgcn_readlane(base, first);
        const int slot = base + mbcnt(h);
        if (hit && slot < a.pair_cap)
          a.pairs[slot] = make_int2(min(i, cur.leaf), max(i, cur.leaf));
      }
      next = n + 1;
    } else {
      const double inv_s = 1.0 / cur.side;
      const double tol = 1e-6 + 1e-15 * (rside * inv_s);
      const double margin = 1e-6 * cur.side + 1e-13 * rside;
      const double gx = box_gap(x, cur.cx, rx, cur.side, inv_s, tol, margin);
      const double gy = box_gap(y, cur.cy, ry, cur.side, inv_s, tol, margin);
      const double gz = box_gap(z, cur.cz, rz, cur.side, inv_s, tol, margin);
      const bool near =
          !a.boxes || gx * gx + gy * gy + gz * gz < range * range * (1.0 + 1e-9);
      const bool open =
          (cur.leaf == -2 || mac_open_f64(d2, cur.side, theta, theta2)) && near && in;
      const unsigned long long O = ballot(open);
      if (O) {
        const unsigned long long C = A & ~O;
        if (C) {
          if (top > 0 && top_pos == cur.skip) {
            const int slot = top - 1;
            st_lo = lane == slot ? (st_lo | (int)(unsigned)C) : st_lo;
            st_hi = lane == slot ? (st_hi | (int)(unsigned)(C >> 32)) : st_hi;
          } else if (top < kWave - 1) {
            const int slot = top;
            st_pos = lane == slot ? cur.skip : st_pos;
            st_lo = lane == slot ? (int)(unsigned)C : st_lo;
            st_hi = lane == slot ? (int)(unsigned)(C >> 32) : st_hi;
            top_pos = cur.skip;
            top++;
          } else {
            overflow = true;
          }
        }
        A = O;
        next = n + 1;
      } else {
        next = cur.skip;
      }
    }
    n = max(next, n + 1);
    if (top > 0 && top_pos == n) {
      const unsigned lo = (unsigned)__builtin_amdgcn_readlane(st_lo, top - 1);
      const unsigned hi = (unsigned)__builtin_amdgcn_readlane(st_hi, top - 1);
      A |= ((unsigned long long)hi << 32) | lo;
      top--;
      top_pos = top > 0 ? __builtin_amdgcn_readlane(st_pos, top - 1) : -1;
    }
  }
  if (overflow && lane == 0) atomicOr(a.flags_out, kFlagStack);
}

// ------------------------------------------------------- wave ordering --
// Longest-first order of the walk's workgroups.  The walk adds each
// wavefront's node visits to tile_cost[tile] (a tile is the run of
// wavefronts one workgroup takes); the next walk dispatches the tiles of
// every XCD region (see tile_of_block) in decreasing order of that cost, so
// the expensive tiles start first and cheap ones fill the end of the
// launch.  One workgroup per region: a stable counting sort on 32 cost
// buckets (linear in cost up to the region's maximum; equal buckets keep
// the plain order, so with no cost recorded the order is tile_of_block's).
// The costs are cleared for the next walk to accumulate.
constexpr int kOrderBuckets = 32;
constexpr int kOrderThreads = 1024;
constexpr int kOrderWaves = kOrderThreads / kWave;

__global__ void __launch_bounds__(kOrderThreads) tile_order_kernel(
    unsigned* __restrict__ cost, int n_tiles, int groups, int* __restrict__ order) {
  const int G = groups <= 1 ? 1 : groups;
  const int r = blockIdx.x;
  const int q = n_tiles / G, rem = n_tiles % G;
  const int beg = r * q + min(r, rem);
  const int cnt = q + (r < rem ? 1 : 0);
  const int lane = lane_id(), w = threadIdx.x >> 6;
  __shared__ unsigned wmax[kOrderWaves];
  __shared__ int base[kOrderBuckets];
  __shared__ int wcount[kOrderWaves][kOrderBuckets];
  __shared__ int chunk_total[kOrderBuckets];

  // maximum cost of the region
  unsigned mx = 0;
  for (int k = threadIdx.x; k < cnt; k += kOrderThreads) mx = max(mx, cost[beg + k]);
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) mx = max(mx, (unsigned)__shfl_xor((int)mx, d, 64));
  if (lane == 0) wmax[w] = mx;
  if (threadIdx.x < kOrderBuckets) base[threadIdx.x] = 0;
  __syncthreads();
  mx = 0;
  for (int k = 0; k < kOrderWaves; k++) mx = max(mx, wmax[k]);
  const unsigned long long den = (unsigned long long)mx + 1;
  auto bucket = [&](unsigned c) {  // 0 = most expensive
    return kOrderBuckets - 1 - (int)(((unsigned long long)c * kOrderBuckets) / den);
  };

  // histogram, exclusive scan into base
  for (int k = threadIdx.x; k < cnt; k += kOrderThreads) atomicAdd(&base[bucket(cost[beg + k])], 1);
  __syncthreads();
  if (w == 0) {
    const int h = lane < kOrderBuckets ? base[lane] : 0;
    const int incl = wave_inclusive_scan(h);
    if (lane < kOrderBuckets) base[lane] = incl - h;
  }
  __syncthreads();

  // stable scatter, one chunk of kOrderThreads tiles at a time
  for (int k0 = 0; k0 < cnt; k0 += kOrderThreads) {
    const int k = k0 + threadIdx.x;
    const bool v = k < cnt;
    const int b = v ? bucket(cost[beg + k]) : 0;
    unsigned long long same = ballot(v);
#pragma unroll
    for (int bit = 0; bit < 5; bit++) {
      const unsigned long long m = ballot(v && ((b >> bit) & 1));
      same &= ((b >> bit) & 1) ? m : ~m;
    }
    const int rank = mbcnt(same);
    for (int j = lane; j < kOrderBuckets; j += kWave) wcount[w][j] = 0;
    __syncthreads();
    if (v && rank == 0) wcount[w][b] = __popcll(same);
    __syncthreads();
    if (threadIdx.x < kOrderBuckets) {  // prefix over the waves, per bucket
      int acc = 0;
      for (int j = 0; j < kOrderWaves; j++) {
        const int c = wcount[j][threadIdx.x];
        wcount[j][threadIdx.x] = acc;
        acc += c;
      }
      chunk_total[threadIdx.x] = acc;
    }
    __syncthreads();
    if (v) order[(base[b] + wcount[w][b] + rank) * G + r] = beg + k;
    __syncthreads();
    if (threadIdx.x < kOrderBuckets) base[threadIdx.x] += chunk_total[threadIdx.x];
    __syncthreads();
  }
  for (int k = threadIdx.x; k < cnt; k += kOrderThreads) cost[beg + k] = 0;
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

template <int BLOCK, int P>
void launch_walk_shape(const WalkArgs& a, int f32, int damping, hipStream_t s) {
  const int nhw = P == 1 ? a.n_waves : a.n_waves2;
  const int blocks = (nhw * kWave + BLOCK - 1) / BLOCK;
  if (f32 && P == 1 && a.nodes32) {
    if (damping)
      walk_kernel<true, true, false, BLOCK, 1, true><<<blocks, BLOCK, 0, s>>>(a);
    else
      walk_kernel<true, false, false, BLOCK, 1, true><<<blocks, BLOCK, 0, s>>>(a);
  } else if (f32) {
    if (damping)
      walk_kernel<true, true, false, BLOCK, P><<<blocks, BLOCK, 0, s>>>(a);
    else
      walk_kernel<true, false, false, BLOCK, P><<<blocks, BLOCK, 0, s>>>(a);
  } else if (a.mac_nodes) {
    if (damping)
      walk_kernel<false, true, true, BLOCK, 1><<<blocks, BLOCK, 0, s>>>(a);
    else
      walk_kernel<false, false, true, BLOCK, 1><<<blocks, BLOCK, 0, s>>>(a);
  } else {
    if (damping)
      walk_kernel<false, true, false, BLOCK, P><<<blocks, BLOCK, 0, s>>>(a);
    else
      walk_kernel<false, false, false, BLOCK, P><<<blocks, BLOCK, 0, s>>>(a);
  }
}

template <int P>
void launch_walk_pack(const WalkArgs& a, int f32, int damping, int block,
                      hipStream_t s) {
  switch (block) {
    case 64:
      launch_walk_shape<64, P>(a, f32, damping, s);
      break;
    case 128:
      launch_walk_shape<128, P>(a, f32, damping, s);
      break;
    case 512:
      launch_walk_shape<512, P>(a, f32, damping, s);
      break;
    case 1024:
      launch_walk_shape<1024, P>(a, f32, damping, s);
      break;
    default:
      launch_walk_shape<256, P>(a, f32, damping, s);
  }
}

void launch_walk(const WalkArgs& a, int f32, int damping, int block, int pack,
                 hipStream_t s) {
  if (a.n_waves <= 0) return;
  if (a.split > 1) {
    // pieces: 64-thread workgroups, one particle per lane (the pieces of
    // one traversal on consecutive workgroups, spread over the XCDs)
    const int blocks = (int)(((long long)a.n_waves * a.split * kWave + 63) / 64);
    if (f32 && a.nodes32) {
      if (damping)
        walk_kernel<true, true, false, 64, 1, true, true><<<blocks, 64, 0, s>>>(a);
      else
        walk_kernel<true, false, false, 64, 1, true, true><<<blocks, 64, 0, s>>>(a);
    } else if (f32) {
      if (damping)
        walk_kernel<true, true, false, 64, 1, false, true><<<blocks, 64, 0, s>>>(a);
      else
        walk_kernel<true, false, false, 64, 1, false, true><<<blocks, 64, 0, s>>>(a);
    } else if (a.mac_nodes) {
      if (damping)
        walk_kernel<false, true, true, 64, 1, false, true><<<blocks, 64, 0, s>>>(a);
      else
        walk_kernel<false, false, true, 64, 1, false, true><<<blocks, 64, 0, s>>>(a);
    } else {
      if (damping)
        walk_kernel<false, true, false, 64, 1, false, true><<<blocks, 64, 0, s>>>(a);
      else
        walk_kernel<false, false, false, 64, 1, false, true><<<blocks, 64, 0, s>>>(a);
    }
  } else if (pack == 2 && !a.mac_nodes) {
    // the frozen-decision walk has a one-particle-per-lane form only
    launch_walk_pack<2>(a, f32, damping, block, s);
  } else {
    launch_walk_pack<1>(a, f32, damping, block, s);
  }
  if (a.split > 1) {
    const int blocks = (int)(((long long)a.n_waves * kWave + 255) / 256);
    if (f32)
      walk_combine_kernel<true><<<blocks, 256, 0, s>>>(a);
    else
      walk_combine_kernel<false><<<blocks, 256, 0, s>>>(a);
  }
}

void launch_tile_order(unsigned* tile_cost, int n_tiles, int groups, int* order,
                       hipStream_t s) {
  if (n_tiles <= 0) return;
  const int G = groups <= 1 ? 1 : groups;
  tile_order_kernel<<<G, kOrderThreads, 0, s>>>(tile_cost, n_tiles, G, order);
}

void launch_detect(const DetectArgs& a, hipStream_t s) {
  if (a.n_waves > 0) detect_kernel<<<(a.n_waves + 3) / 4, 256, 0, s>>>(a);
}

int walk_tiles(int n_waves, int n_waves2, int block, int pack) {
  const int nhw = pack == 1 ? n_waves : n_waves2;
  return (nhw * kWave + block - 1) / block;
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
