// api.hip -- the C ABI of include/nbs.h: context, device buffers and the
// step pipeline
//   [bbox] -> root -> keys -> radix sort -> fix-up -> permute ->
//   emit -> link -> bottom-up mass/COM -> walk (+ fused kick-drift) ->
//   [adaptive dt -> update]
// all enqueued on the context's stream with no host synchronisation inside
// a step.  Errors are returned, never exit()ed (the reference's
// cuda_assert exits, src/cuda_code.cu:29-37).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <initializer_list>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/nbs.h"
#include "../../include/nbs_host.h"
#include "nbs_device_util.h"
#include "nbs_internal.h"

using namespace nbs;

namespace {

constexpr int kPhaseBuild = 0, kPhaseWalk = 1, kPhaseUpdate = 2;

__global__ void deinterleave_kernel(const double* __restrict__ pos,
                                    const double* __restrict__ vel,
                                    const double* __restrict__ mass, int n,
                                    State st, int has_vel) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  st.x[i] = pos[3 * i];
  st.y[i] = pos[3 * i + 1];
  st.z[i] = pos[3 * i + 2];
  st.vx[i] = has_vel ? vel[3 * i] : 0.0;
  st.vy[i] = has_vel ? vel[3 * i + 1] : 0.0;
  st.vz[i] = has_vel ? vel[3 * i + 2] : 0.0;
  st.m[i] = mass[i];
  st.id[i] = (unsigned)i;
}

// SoA in slot order -> AoS in caller order.
__global__ void interleave_kernel(const double* __restrict__ a,
                                  const double* __restrict__ b,
                                  const double* __restrict__ c,
                                  const unsigned* __restrict__ id, int n,
                                  double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const unsigned k = id[i];
  out[3 * (size_t)k] = a[i];
  out[3 * (size_t)k + 1] = b[i];
  out[3 * (size_t)k + 2] = c[i];
}

// slot order -> caller order, one value per particle
__global__ void scatter_kernel(const double* __restrict__ a, const unsigned* __restrict__ id,
                               int n, double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[id[i]] = a[i];
}

// [n][3] <-> [3][n] (caller order both ways): the adjoint's seeds and
// gradients between the caller's device arrays and the engine's layout.
__global__ void aos_soa_kernel(const double* __restrict__ in, double* __restrict__ out,
                               int n, int to_soa) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int k = 0; k < 3; k++) {
    if (to_soa)
      out[(size_t)k * n + i] = in[3 * (size_t)i + k];
    else
      out[3 * (size_t)i + k] = in[(size_t)k * n + i];
  }
}

// slot order -> caller order, int widened to int64
__global__ void scatter_count_kernel(const int* __restrict__ a, const unsigned* __restrict__ id,
                                     int n, int64_t* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[id[i]] = a[i];
}

__global__ void gather_slots_kernel(const long long* __restrict__ ids, int m,
                                    const int* __restrict__ slot_of_id,
                                    int* __restrict__ slots) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < m) slots[k] = slot_of_id[ids[k]];
}

// State from caller-order [3][n] arrays (vel may be null), identity order.
__global__ void load_soa_kernel(const double* __restrict__ pos,
                                const double* __restrict__ vel,
                                const double* __restrict__ mass, int n,
                                State st) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  st.x[i] = pos[i];
  st.y[i] = pos[n + i];
  st.z[i] = pos[2 * (size_t)n + i];
  st.vx[i] = vel ? vel[i] : 0.0;
  st.vy[i] = vel ? vel[n + i] : 0.0;
  st.vz[i] = vel ? vel[2 * (size_t)n + i] : 0.0;
  st.m[i] = mass[i];
  st.id[i] = (unsigned)i;
}

// slot-order triple -> caller-order [3][n]
__global__ void scatter3_kernel(const double* __restrict__ a,
                                const double* __restrict__ b,
                                const double* __restrict__ c,
                                const unsigned* __restrict__ id, int n,
                                double* __restrict__ dst) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const unsigned j = id[s];
  dst[j] = a[s];
  dst[n + j] = b[s];
  dst[2 * (size_t)n + j] = c[s];
}

// caller-order [3][n] -> slot-order triple
__global__ void gather3_kernel(const double* __restrict__ src,
                               const unsigned* __restrict__ id, int n,
                               double* __restrict__ a, double* __restrict__ b,
                               double* __restrict__ c) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const unsigned j = id[s];
  a[s] = src[j];
  b[s] = src[n + j];
  c[s] = src[2 * (size_t)n + j];
}

template <class T>
T* dalloc(std::vector<void*>& owner, size_t count, hipError_t* err) {
  void* p = nullptr;
  size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
  bytes = (bytes + 255) & ~size_t(255);
  hipError_t e = hipMalloc(&p, bytes);
  if (e != hipSuccess) {
    *err = e;
    return nullptr;
  }
  owner.push_back(p);
  return (T*)p;
}

void free_all(std::vector<void*>& v) {
  for (void* p : v) (void)hipFree(p);
  v.clear();
}

// A fixed block that buffers are carved from in order, 256-byte aligned,
// instead of allocated (the no-allocation entry point's scratch).
struct Arena {
  char* base = nullptr;
  size_t size = 0, used = 0;
};

template <class T>
T* carve(Arena& a, size_t count, hipError_t* err) {
  size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
  bytes = (bytes + 255) & ~size_t(255);
  if (a.used + bytes > a.size) {
    *err = hipErrorOutOfMemory;
    return nullptr;
  }
  T* p = (T*)(a.base + a.used);
  a.used += bytes;
  return p;
}

}  // namespace

struct nbs_ctx {
  nbs_config cfg;
  StepParams p;
  std::string err;
  int device = 0;
  hipStream_t stream = nullptr;
  // When arena.base is set, every buffer of the step is carved from it
  // (nothing is allocated or freed): the no-allocation entry point.
  Arena arena;

  int n = 0, cap_n = 0;
  int node_cap = 0;
  int build_cap = 0;  // keys the build buffers hold
  std::vector<void*> part_bufs, build_bufs, node_bufs, sys_bufs, misc_bufs, split_bufs;

  State st[2];
  int cur = 0;
  double *ax = nullptr, *ay = nullptr, *az = nullptr;
  int* inter = nullptr;
  double* stage = nullptr;
  double* stage_v = nullptr;
  double* stage_m = nullptr;
  int* slot_of_id = nullptr;
  BuildBuffers bb{};
  WalkNode* nodes = nullptr;
  WalkNode32* nodes32 = nullptr;  // the FP32 walk's records, or null
  TreeAux aux{};
  Systems sys{};
  std::vector<int> h_start, h_wave_start, h_wave2_start;
  int* flags = nullptr;
  int* steps_done = nullptr;
  int* stats = nullptr;
  unsigned long long* inter_total = nullptr;

  // direct-sum scratch
  int direct_cap = 0;
  long long* d_ids = nullptr;
  int* d_slots = nullptr;
  double* d_direct_tmp = nullptr;
  double* d_direct_out = nullptr;

  bool has_state = false, bbox_valid = false, tree_valid = false;
  // builds record each node's parent (the adjoint's propagation needs them)
  bool want_parent = false;
  bool acc_valid = false;
  int nsys_req = 1;
  std::vector<int64_t> sys_off_req;

  bool prof = false;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev[4];
  double prof_ms[4] = {0, 0, 0, 0};

  // launch shape of the walk (nbs_autotune / nbs_set_walk_shape)
  int walk_block = 256;
  int walk_groups = 1;
  int walk_pack = 1;
  // key levels in the radix-sorted primary key (3 bits each).  All 21 that
  // fit: with fewer, the fixup's per-run insertion sort met runs of
  // thousands of equal keys while a cold sphere collapses (seconds per step).
  int prim_levels = 21;
  // longest-first workgroup order: per-tile visit counts of the last walk
  // (recorded under cost_block / cost_pack) and the order derived from them
  int walk_order = 0;
  // 1: the compiler-generated walk loop even where the hand-written one applies
  int walk_generic = 0;
  // pieces per walk traversal: requested (0: by wavefront count) and in use,
  // with the pieces' partial sums ([split][3][n] doubles, [split][n] ints)
  int walk_split_req = 0;
  int walk_split = 1;
  double* walk_part = nullptr;
  int* walk_cpart = nullptr;
  // records of the FP32 hand-written loop: 0 by particle count
  // (kWalk32MinN), 32 or 64 bytes (nbs_set_walk_code 3 / 2)
  int walk_records = 0;
  // interaction counting of the walks (nbs_set_counting): per particle by
  // default; lane_counts: `inter` holds the last walk's per-particle counts
  int count_mode = kCountLanes;
  bool lane_counts = false;
  unsigned* tile_cost = nullptr;
  int* tile_order = nullptr;
  int cost_block = -1, cost_pack = -1, cost_tiles = -1;

  // Two captured steps (the state ping-pongs between its two buffers, so
  // a pair of steps returns to the same buffer); valid while the
  // configuration, state size and node capacity do not change.
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  int graph_cur = -1;

  // Adjoint: checkpoints (caller order, [3][n] SoA) every `adj_every`
  // steps, the recomputed positions of one segment, optionally every base
  // position (frozen replays), and the adjoint accumulators.
  std::vector<void*> adj_bufs;
  int adj_steps = 0, adj_every = 0, adj_n = 0;
  bool adj_record = false;
  double adj_dt = 0;
  std::vector<double*> ck_x, ck_v;
  double *adj_m = nullptr, *seg_x = nullptr, *base_x = nullptr;
  double *fin_x = nullptr, *fin_v = nullptr;
  double *g_x = nullptr, *g_v = nullptr, *g_m = nullptr, *g_u = nullptr;
  // g_x, g_v, g_m at the start of the segment being reversed
  double *b_x = nullptr, *b_v = nullptr, *b_m = nullptr;
  double *s_u = nullptr, *s_gx = nullptr, *s_gm = nullptr;
  double* node_g = nullptr;
  int node_g_cap = 0;
  WalkNode* mac_nodes = nullptr;
  int mac_nodes_cap = 0;
  double* mac_x = nullptr;
  double *p_x = nullptr, *p_v = nullptr, *p_m = nullptr;

  // Snapshots for frame dumps (nbs_snapshot_begin / _end): pos, vel, acc
  // gathered into caller order [n][3] on the device, then copied on a
  // second stream into one of two pinned host buffers while the following
  // steps run on `stream`.
  hipStream_t copy_stream = nullptr;
  hipEvent_t snap_ready = nullptr, snap_done = nullptr;
  double* snap_dev = nullptr;                  // 9 * snap_cap doubles
  double* snap_host[2] = {nullptr, nullptr};   // pinned, 9 * snap_cap each
  size_t snap_cap = 0;
  int snap_next = 0;     // host buffer the next snapshot fills
  int snap_pending = -1; // host buffer being filled, -1 when none
  int snap_n = 0;

  // Collision merging (nbs_set_collisions): buffers sized for col_cap_n
  // particles and cb.pair_cap pairs; running totals of detections and
  // absorbed particles.
  bool collide = false;
  double collide_range = 0;
  double min_size = 0;  // nbs_set_min_node_size (0: off)
  int64_t min_removed = 0;
  CollideBuffers cb{};
  std::vector<void*> col_bufs, col_pair_bufs;
  int col_cap_n = 0;
  int col_pair_want = 0;
  int col_info_cap = 0;
  int64_t col_pairs = 0, col_removed = 0;

  // The tree of the reference's threads > 1 path (cfg.threaded_tree):
  // per-particle inside() masks and new slots (thr_cap_n particles),
  // per-entry data (thr_cap_e entries), the last build's entry count, and
  // an error of the build's host part (reported after the step).
  std::vector<void*> thr_bufs, thr_ebufs;
  int thr_cap_n = 0, thr_cap_e = 0, n_entries = 0;
  int *t_mask = nullptr, *t_cnt = nullptr, *t_outside = nullptr, *t_eoff = nullptr;
  int *t_orank = nullptr, *t_new_slot = nullptr, *t_occ = nullptr;
  unsigned *t_gperm = nullptr, *t_eid = nullptr;
  int *t_epart = nullptr, *t_sp = nullptr, *t_first = nullptr, *t_pslot = nullptr;
  int* t_estart = nullptr;
  double *t_ex = nullptr, *t_ey = nullptr, *t_ez = nullptr, *t_em = nullptr;
  int thr_err = 0;
  std::string thr_msg;

  // Frame images (nbs_render_image): per-pixel counts and values, the
  // running maximum and the RGB rows, for img_pixels pixels.
  std::vector<void*> img_bufs;
  size_t img_pixels = 0;
  unsigned* img_counts = nullptr;
  double* img_temp = nullptr;
  double* img_partial = nullptr;
  unsigned long long* img_max = nullptr;
  unsigned char* img_rgb = nullptr;
};

namespace {

int fail(nbs_ctx* c, int code, const std::string& msg) {
  c->err = msg;
  return code;
}

int hip_fail(nbs_ctx* c, hipError_t e, const char* what) {
  return fail(c, e == hipErrorOutOfMemory ? NBS_ERR_NOMEM : NBS_ERR_HIP,
              std::string(what) + ": " + hipGetErrorString(e));
}

// A device buffer of the context: carved from its arena when it has one,
// else allocated (and owned by `owner`, freed by free_all).
template <class T>
T* cbuf(nbs_ctx* c, std::vector<void*>& owner, size_t count, hipError_t* err) {
  return c->arena.base ? carve<T>(c->arena, count, err) : dalloc<T>(owner, count, err);
}

#define HIPCHK(c, call)                                   \
  do {                                                    \
    hipError_t e_ = (call);                               \
    if (e_ != hipSuccess) return hip_fail((c), e_, #call); \
  } while (0)

StepParams params_of(const nbs_config& c) {
  StepParams p{};
  p.theta = c.theta;
  p.dt = c.dt;
  p.damping = c.damping;
  p.root_mode = c.root_mode;
  for (int k = 0; k < 3; k++) p.root_center[k] = c.root_center[k];
  p.root_size = c.root_size;
  p.adaptive = c.adaptive;
  p.min_adaptive_dt = c.min_adaptive_dt;
  p.max_pos_change = c.max_pos_change;
  p.max_vel_change = c.max_vel_change;
  return p;
}

int check_config(nbs_ctx* c, const nbs_config* cfg) {
  if (!(cfg->theta >= 0)) return fail(c, NBS_ERR_ARG, "theta must be >= 0");
  if (!(cfg->dt >= 0)) return fail(c, NBS_ERR_ARG, "dt must be >= 0");
  if (cfg->force_precision != NBS_FORCE_F64 &&
      cfg->force_precision != NBS_FORCE_F32)
    return fail(c, NBS_ERR_ARG, "force_precision must be NBS_FORCE_F64/F32");
  if (cfg->root_mode != NBS_ROOT_BBOX && cfg->root_mode != NBS_ROOT_FIXED)
    return fail(c, NBS_ERR_ARG, "root_mode must be NBS_ROOT_BBOX/FIXED");
  if (cfg->root_mode == NBS_ROOT_FIXED && !(cfg->root_size > 0))
    return fail(c, NBS_ERR_ARG, "root_size must be > 0 for NBS_ROOT_FIXED");
  if (cfg->threaded_tree != 0 && cfg->threaded_tree != 1)
    return fail(c, NBS_ERR_ARG, "threaded_tree must be 0 or 1");
  if (cfg->threaded_tree && cfg->root_mode != NBS_ROOT_BBOX)
    return fail(c, NBS_ERR_ARG, "the threaded tree has the bbox + 2 root (NBS_ROOT_BBOX)");
  return NBS_OK;
}

// What the threaded-root tree does not support (one system; no
// minimum-size merge, adjoint or frozen replay).
int check_threaded(nbs_ctx* c, const char* what) {
  if (!c->cfg.threaded_tree) return NBS_OK;
  return fail(c, NBS_ERR_STATE, std::string(what) + " is not supported with the threaded tree");
}

void drop_graph(nbs_ctx* c);
int alloc_build(nbs_ctx* c, int cap);

int alloc_particles(nbs_ctx* c, int n) {
  if (n <= c->cap_n && c->cap_n > 0) return NBS_OK;
  free_all(c->part_bufs);
  const int cap = std::max(n, 64);
  hipError_t e = hipSuccess;
  auto& o = c->part_bufs;
  for (int b = 0; b < 2; b++) {
    State& s = c->st[b];
    s.x = cbuf<double>(c, o, cap, &e);
    s.y = cbuf<double>(c, o, cap, &e);
    s.z = cbuf<double>(c, o, cap, &e);
    s.vx = cbuf<double>(c, o, cap, &e);
    s.vy = cbuf<double>(c, o, cap, &e);
    s.vz = cbuf<double>(c, o, cap, &e);
    s.m = cbuf<double>(c, o, cap, &e);
    s.id = cbuf<unsigned>(c, o, cap, &e);
  }
  c->ax = cbuf<double>(c, o, cap, &e);
  c->ay = cbuf<double>(c, o, cap, &e);
  c->az = cbuf<double>(c, o, cap, &e);
  c->inter = cbuf<int>(c, o, cap, &e);
  c->stage = cbuf<double>(c, o, 3 * (size_t)cap, &e);
  c->stage_v = cbuf<double>(c, o, 3 * (size_t)cap, &e);
  c->stage_m = cbuf<double>(c, o, cap, &e);
  c->slot_of_id = cbuf<int>(c, o, cap, &e);
  if (e != hipSuccess) {
    free_all(c->part_bufs);
    c->cap_n = 0;
    return hip_fail(c, e, "hipMalloc(particle buffers)");
  }
  c->cap_n = cap;
  return alloc_build(c, cap);
}

// The build's buffers (keys, sort, emission) for `cap` keys: particles, or
// tree entries with the threaded root.
int alloc_build(nbs_ctx* c, int cap) {
  free_all(c->build_bufs);
  c->build_cap = 0;
  hipError_t e = hipSuccess;
  auto& o = c->build_bufs;
  BuildBuffers& b = c->bb;
  b.prim0 = cbuf<unsigned long long>(c, o, cap, &e);
  b.prim1 = cbuf<unsigned long long>(c, o, cap, &e);
  b.val0 = cbuf<unsigned>(c, o, cap, &e);
  b.val1 = cbuf<unsigned>(c, o, cap, &e);
  b.khi_u = cbuf<unsigned long long>(c, o, cap, &e);
  b.klo_u = cbuf<unsigned long long>(c, o, cap, &e);
  b.khi = cbuf<unsigned long long>(c, o, cap, &e);
  b.klo = cbuf<unsigned long long>(c, o, cap, &e);
  b.counts = cbuf<int>(c, o, cap, &e);
  b.close_at = cbuf<int>(c, o, cap, &e);
  b.blkstats = cbuf<int>(c, o, 2 * (size_t)emit_blocks(cap), &e);
  b.offsets = cbuf<int>(c, o, (size_t)cap + 1, &e);
  b.sort_tmp = cbuf<int>(c, o, radix_tmp_ints(cap), &e);
  b.runs = cbuf<int>(c, o, run_list_ints(cap), &e);
  b.runs_big = run_list_big(cap);
  b.dhist = cbuf<int>(c, o, depth_hist_ints(cap), &e);
  b.doff = cbuf<int>(c, o, depth_hist_ints(cap), &e);
  b.scan_tmp = cbuf<int>(
      c, o, scan_tmp_ints(std::max((size_t)cap + 1, depth_hist_ints(cap))), &e);
  if (e != hipSuccess) {
    free_all(c->build_bufs);
    return hip_fail(c, e, "hipMalloc(build buffers)");
  }
  c->build_cap = cap;
  c->cb.scan_tmp = b.scan_tmp;  // the collision kernels borrow it
  return NBS_OK;
}

// Particle count from which the FP32 walk reads 32-byte records.  The
// 2^20-particle sphere's 1.6M 64-byte records (100 MB) fit in MI355X's
// 256 MB of last-level cache and its walk gains nothing from the smaller
// ones (which also cost the build 32 B per node); the 2^22-particle shell's
// 6.2M (400 MB) do not, and its walk takes 4-8 % less time with them.
constexpr int kWalk32MinN = 1 << 21;

// Whether node arrays of capacity `cap` come with the 32-byte copy: FP32
// forces only, while its 25-bit skip links can address every node, not in a
// context carved from the fixed scratch array (its budget holds the full
// records only), and as the walk code asks.
bool want_walk32(const nbs_ctx* c, int cap) {
  if (c->cfg.force_precision != NBS_FORCE_F32 || c->arena.base ||
      (size_t)cap + 1 >= (size_t)kLink32MaxNodes || c->walk_generic)
    return false;
  return c->walk_records == 32 || (c->walk_records == 0 && c->n >= kWalk32MinN);
}

int alloc_nodes(nbs_ctx* c, int cap) {
  drop_graph(c);
  free_all(c->node_bufs);
  hipError_t e = hipSuccess;
  auto& o = c->node_bufs;
  // one record beyond the capacity: the hand-written FP32 walk requests the
  // record after a system's last node (unused) without a bounds test
  c->nodes = cbuf<WalkNode>(c, o, (size_t)cap + 1, &e);
  // the FP32 walk's 32-byte copy of the records
  c->nodes32 = nullptr;
  if (want_walk32(c, cap)) c->nodes32 = cbuf<WalkNode32>(c, o, (size_t)cap + 1, &e);
  c->aux.walk32 = c->nodes32;
  c->aux.parent = cbuf<int>(c, o, cap, &e);
  c->bb.bucket = cbuf<int>(c, o, cap, &e);
  if (e != hipSuccess) {
    free_all(c->node_bufs);
    c->node_cap = 0;
    return hip_fail(c, e, "hipMalloc(node buffers)");
  }
  c->node_cap = cap;
  return NBS_OK;
}

// Pieces per walk traversal: a system of few wavefronts (a small N) leaves
// most of the GPU's wavefront slots (1024 SIMDs x 8) empty and each
// wavefront latency-bound, so each traversal is walked as up to 8 pieces,
// the fewest that give about kWalkWaveSlots wavefronts.  Fixed by the
// wavefront count, so a given state always takes the same pieces.
constexpr long long kWalkWaveSlots = 8192;
constexpr int kWalkMaxPieces = 32;
int walk_split_for(const nbs_ctx* c) {
  if (c->walk_split_req > 0) return c->walk_split_req;
  int k = 1;
  while (k < 8 && (long long)c->sys.n_waves * k * 2 <= kWalkWaveSlots) k *= 2;
  return k;
}

// The pieces' partial-sum buffers for the current systems and request.
int alloc_walk_split(nbs_ctx* c) {
  drop_graph(c);
  free_all(c->split_bufs);
  c->walk_part = nullptr;
  c->walk_cpart = nullptr;
  c->walk_split = walk_split_for(c);
  if (c->walk_split > 1 && c->n > 0) {
    hipError_t e = hipSuccess;
    c->walk_part = cbuf<double>(c, c->split_bufs, (size_t)c->walk_split * 3 * c->n, &e);
    c->walk_cpart = cbuf<int>(c, c->split_bufs, (size_t)c->walk_split * c->n, &e);
    if (e != hipSuccess) {
      free_all(c->split_bufs);
      c->walk_part = nullptr;
      c->walk_cpart = nullptr;
      c->walk_split = 1;
      return hip_fail(c, e, "hipMalloc(walk pieces)");
    }
  }
  return NBS_OK;
}

// Per-system arrays from caller offsets (one system when offsets empty).
int setup_systems(nbs_ctx* c) {
  const int nsys = c->nsys_req;
  drop_graph(c);
  free_all(c->sys_bufs);
  hipError_t e = hipSuccess;
  auto& o = c->sys_bufs;
  Systems& s = c->sys;
  s.nsys = nsys;
  s.start = cbuf<int>(c, o, nsys + 1, &e);
  s.wave_start = cbuf<int>(c, o, nsys + 1, &e);
  s.wave2_start = cbuf<int>(c, o, nsys + 1, &e);
  s.node_start = cbuf<int>(c, o, nsys + 1, &e);
  s.root = cbuf<double>(c, o, 4 * (size_t)nsys, &e);
  s.dt = cbuf<double>(c, o, nsys, &e);
  s.time = cbuf<double>(c, o, nsys, &e);
  if (e != hipSuccess) return hip_fail(c, e, "hipMalloc(system buffers)");
  c->h_start.assign(nsys + 1, 0);
  c->h_wave_start.assign(nsys + 1, 0);
  c->h_wave2_start.assign(nsys + 1, 0);
  for (int k = 0; k <= nsys; k++)
    c->h_start[k] = (int)(c->sys_off_req.empty() ? (k == 0 ? 0 : c->n)
                                                 : c->sys_off_req[k]);
  for (int k = 0; k < nsys; k++) {
    const int cnt = c->h_start[k + 1] - c->h_start[k];
    c->h_wave_start[k + 1] = c->h_wave_start[k] + (cnt + kWave - 1) / kWave;
    c->h_wave2_start[k + 1] =
        c->h_wave2_start[k] + (cnt + 2 * kWave - 1) / (2 * kWave);
  }
  s.n_waves = c->h_wave_start[nsys];
  s.n_waves2 = c->h_wave2_start[nsys];
  c->prim_levels = max_prim_levels(nsys);
  s.bbox = cbuf<double>(c, o, 6 * (size_t)std::max(s.n_waves, 1), &e);
  s.amax = cbuf<double>(c, o, std::max(s.n_waves, 1), &e);
  s.vmax = cbuf<double>(c, o, std::max(s.n_waves, 1), &e);
  c->inter_total = cbuf<unsigned long long>(c, o, std::max(s.n_waves, 1), &e);
  c->tile_cost = cbuf<unsigned>(c, o, std::max(s.n_waves, 1), &e);
  c->tile_order = cbuf<int>(c, o, std::max(s.n_waves, 1), &e);
  c->cost_tiles = -1;
  if (e != hipSuccess) return hip_fail(c, e, "hipMalloc(system buffers)");
  launch_fill_u64(c->inter_total, 0, std::max(s.n_waves, 1), c->stream);
  {
    const int r = alloc_walk_split(c);
    if (r) return r;
  }
  HIPCHK(c, hipMemcpyAsync(s.start, c->h_start.data(), (nsys + 1) * sizeof(int),
                           hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(s.wave_start, c->h_wave_start.data(),
                           (nsys + 1) * sizeof(int), hipMemcpyHostToDevice,
                           c->stream));
  HIPCHK(c, hipMemcpyAsync(s.wave2_start, c->h_wave2_start.data(),
                           (nsys + 1) * sizeof(int), hipMemcpyHostToDevice,
                           c->stream));
  std::vector<double> zeros(nsys, 0.0), dts(nsys, c->cfg.dt);
  HIPCHK(c, hipMemcpyAsync(s.time, zeros.data(), nsys * sizeof(double),
                           hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(s.dt, dts.data(), nsys * sizeof(double),
                           hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->bbox_valid = false;
  c->tree_valid = false;
  return NBS_OK;
}

void prof_begin(nbs_ctx* c, int phase, hipEvent_t* ev) {
  if (!c->prof) return;
  (void)phase;
  hipEventCreate(ev);
  hipEventRecord(*ev, c->stream);
}
void prof_end(nbs_ctx* c, int phase, hipEvent_t ev0) {
  if (!c->prof) return;
  hipEvent_t ev1;
  hipEventCreate(&ev1);
  hipEventRecord(ev1, c->stream);
  c->ev[phase].push_back({ev0, ev1});
}

int ensure_nodes(nbs_ctx* c) {
  const int want = std::max(c->cfg.node_factor > 0 ? c->cfg.node_factor : 4, 2) * c->n + 4096;
  if (c->node_cap >= want && (c->nodes32 != nullptr) == want_walk32(c, c->node_cap))
    return NBS_OK;
  return alloc_nodes(c, std::max(want, c->node_cap));
}

// Buffers of the threaded-root build for the current particle count and
// `ne` entries (entries grow with the particles on the root's centre
// planes; the build buffers then hold entries, not particles).
int ensure_threaded(nbs_ctx* c, int ne) {
  hipError_t e = hipSuccess;
  const int n = std::max(c->n, 1);
  if (c->thr_cap_n < n) {  // per particle (filled before the entry count is known)
    free_all(c->thr_bufs);
    c->thr_cap_n = 0;
    const int cn = std::max(n, c->cap_n);
    auto& o = c->thr_bufs;
    c->t_mask = dalloc<int>(o, cn, &e);
    c->t_cnt = dalloc<int>(o, cn, &e);
    c->t_outside = dalloc<int>(o, cn, &e);
    c->t_eoff = dalloc<int>(o, (size_t)cn + 1, &e);
    c->t_orank = dalloc<int>(o, (size_t)cn + 1, &e);
    c->t_new_slot = dalloc<int>(o, cn, &e);
    c->t_gperm = dalloc<unsigned>(o, cn, &e);
    c->t_occ = dalloc<int>(o, member_blocks(cn), &e);
    c->t_estart = dalloc<int>(o, 2, &e);
    if (e != hipSuccess) {
      free_all(c->thr_bufs);
      return hip_fail(c, e, "hipMalloc(threaded-root buffers)");
    }
    c->thr_cap_n = cn;
  }
  if (c->thr_cap_e < ne) {  // per entry
    free_all(c->thr_ebufs);
    c->thr_cap_e = 0;
    const int ce = std::max(ne + ne / 4, c->thr_cap_n + 64);
    auto& o = c->thr_ebufs;
    c->t_eid = dalloc<unsigned>(o, ce, &e);
    c->t_epart = dalloc<int>(o, ce, &e);
    c->t_sp = dalloc<int>(o, ce, &e);
    c->t_first = dalloc<int>(o, ce, &e);
    c->t_pslot = dalloc<int>(o, (size_t)ce + 1, &e);
    c->t_ex = dalloc<double>(o, ce, &e);
    c->t_ey = dalloc<double>(o, ce, &e);
    c->t_ez = dalloc<double>(o, ce, &e);
    c->t_em = dalloc<double>(o, ce, &e);
    if (e != hipSuccess) {
      free_all(c->thr_ebufs);
      return hip_fail(c, e, "hipMalloc(threaded-root entry buffers)");
    }
    c->thr_cap_e = ce;
  }
  if (c->build_cap < ne) return alloc_build(c, c->thr_cap_e);
  return NBS_OK;
}

// The tree gen_root_threaded builds (threads > 1, one system): eight
// eager root children, each filled with the particles its inclusive
// inside() accepts, built over (particle, child) entries (see
// tree_threaded.hip).  The entry count is read back from the device
// before the keys are formed.  A failure of the host part is kept in
// thr_err and reported by the caller after the step; the step's kernels
// then do nothing (the flags word is set).
void enqueue_build_threaded(nbs_ctx* c, bool keep_root) {
  hipStream_t s = c->stream;
  State& cur = c->st[c->cur];
  State& nxt = c->st[1 - c->cur];
  const int n = c->n;
  auto bail = [&](int code, const std::string& msg) {
    if (!c->thr_err) {
      c->thr_err = code;
      c->thr_msg = msg;
    }
    launch_fill_int(c->flags, kFlagStack, 1, s);
  };
  if (!keep_root) {
    if (!c->bbox_valid) {
      launch_bbox(cur, c->sys, s);
      c->bbox_valid = true;
    }
    launch_root(c->sys, c->p, c->flags, s);
  }
  int r = ensure_threaded(c, 0);
  if (r) return bail(r, c->err);
  launch_member(cur, n, c->sys.root, c->t_mask, c->t_cnt, c->t_outside, c->t_occ, s);
  exclusive_scan(c->t_cnt, c->t_eoff, n, c->bb.scan_tmp, s);
  int e_real = 0;
  std::vector<int> occ(member_blocks(n));
  if (hipMemcpyAsync(&e_real, c->t_eoff + n, sizeof(int), hipMemcpyDeviceToHost, s) !=
          hipSuccess ||
      hipMemcpyAsync(occ.data(), c->t_occ, occ.size() * sizeof(int), hipMemcpyDeviceToHost,
                     s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return bail(NBS_ERR_HIP, "threaded root: entry count read-back failed");
  int mask = 0;
  for (int v : occ) mask |= v;
  int empty[8] = {0, 0, 0, 0, 0, 0, 0, 0}, n_empty = 0;
  for (int o = 0; o < 8; o++)
    if (!((mask >> o) & 1)) empty[n_empty++] = o;
  const int ne = e_real + n_empty;
  r = ensure_threaded(c, ne);
  if (r) return bail(r, c->err);
  BuildBuffers& b = c->bb;
  const int levels = std::min(c->prim_levels, max_prim_levels(1));
  launch_entry_keys(cur, n, c->sys.root, c->t_mask, c->t_eoff, e_real, empty, n_empty,
                    levels, b, c->t_epart, c->t_eid, s);
  const int which = radix_sort_pairs(b.prim0, b.val0, b.prim1, b.val1, ne, 3 * levels,
                                     b.sort_tmp, s);
  unsigned long long* prim = which ? b.prim1 : b.prim0;
  unsigned* perm = which ? b.val1 : b.val0;
  State by_entry = cur;
  by_entry.id = c->t_eid;  // runs of equal keys: caller index of each entry
  launch_fixup(b, prim, perm, which ? b.val0 : b.val1, by_entry, ne, s);
  launch_entry_gather(perm, c->t_epart, c->t_eoff, cur, ne, c->t_ex, c->t_ey, c->t_ez,
                      c->t_em, c->t_sp, c->t_first, s);
  exclusive_scan(c->t_first, c->t_pslot, ne, b.scan_tmp, s);
  exclusive_scan(c->t_outside, c->t_orank, n, b.scan_tmp, s);
  launch_entry_slots(c->t_sp, c->t_first, c->t_pslot, c->t_cnt, c->t_orank, ne, n,
                     c->t_new_slot, c->t_gperm, s);
  launch_permute(cur, nxt, c->t_gperm, n, s);
  c->cur ^= 1;
  launch_fill_int(c->t_estart, 0, 1, s);
  launch_fill_int(c->t_estart + 1, ne, 1, s);
  Systems es = c->sys;
  es.nsys = 1;
  es.start = c->t_estart;
  State leaves{};
  leaves.x = c->t_ex;
  leaves.y = c->t_ey;
  leaves.z = c->t_ez;
  leaves.m = c->t_em;
  TreeAux aux = c->aux;
  if (!c->want_parent) aux.parent = nullptr;
  launch_emit(b, es, c->p, leaves, c->nodes, aux, ne, c->node_cap, s);
  launch_entry_patch(b.offsets, c->t_sp, c->t_new_slot, ne, c->p.theta, n_empty, c->nodes,
                     c->nodes32, b.stats, c->flags, s);
  launch_bottom_up(b, c->nodes, aux, ne, 1, s);
  c->n_entries = ne;
}

// Enqueues one tree build of the current state.
void enqueue_build(nbs_ctx* c, bool keep_root = false) {
  if (c->cfg.threaded_tree) {
    hipEvent_t e0 = nullptr;
    prof_begin(c, kPhaseBuild, &e0);
    enqueue_build_threaded(c, keep_root);
    prof_end(c, kPhaseBuild, e0);
    return;
  }
  c->n_entries = c->n;
  hipStream_t s = c->stream;
  hipEvent_t e0 = nullptr;
  prof_begin(c, kPhaseBuild, &e0);
  State& cur = c->st[c->cur];
  State& nxt = c->st[1 - c->cur];
  if (!keep_root) {  // keep_root: the roots of the last build (a rebuild)
    if (!c->bbox_valid) {
      launch_bbox(cur, c->sys, s);
      c->bbox_valid = true;  // positions do not change during a build
    }
    launch_root(c->sys, c->p, c->flags, s);
  }
  const int nbits = launch_keys(cur, c->sys, c->bb, c->prim_levels, s);
  const int which = radix_sort_pairs(c->bb.prim0, c->bb.val0, c->bb.prim1,
                                     c->bb.val1, c->n, nbits, c->bb.sort_tmp, s);
  unsigned long long* prim = which ? c->bb.prim1 : c->bb.prim0;
  unsigned* perm = which ? c->bb.val1 : c->bb.val0;
  launch_fixup(c->bb, prim, perm, which ? c->bb.val0 : c->bb.val1, cur, c->n, s);
  launch_permute(cur, nxt, perm, c->n, s);
  c->cur ^= 1;
  TreeAux aux = c->aux;
  if (!c->want_parent) aux.parent = nullptr;
  launch_emit(c->bb, c->sys, c->p, c->st[c->cur], c->nodes, aux, c->n, c->node_cap, s);
  launch_bottom_up(c->bb, c->nodes, aux, c->n, c->sys.nsys, s);
  prof_end(c, kPhaseBuild, e0);
}

// Tiles of the current walk shape; clears the recorded tile costs when they
// were recorded under another shape (another tiling or traversal).
int sync_tile_costs(nbs_ctx* c) {
  const int tiles = walk_tiles(c->sys.n_waves, c->sys.n_waves2, c->walk_block, c->walk_pack);
  if (tiles != c->cost_tiles || c->walk_block != c->cost_block || c->walk_pack != c->cost_pack) {
    launch_fill_int((int*)c->tile_cost, 0, std::max(tiles, 1), c->stream);
    c->cost_tiles = tiles;
    c->cost_block = c->walk_block;
    c->cost_pack = c->walk_pack;
  }
  return tiles;
}

void enqueue_walk(nbs_ctx* c, int fused) {
  hipEvent_t e0 = nullptr;
  prof_begin(c, kPhaseWalk, &e0);
  State& st = c->st[c->cur];
  WalkArgs a{};
  a.nodes = c->nodes;
  a.nodes32 = c->nodes32;
  a.root = c->sys.root;
  a.node_start = c->sys.node_start;
  a.sys_start = c->sys.start;
  a.wave_start = c->sys.wave_start;
  a.wave2_start = c->sys.wave2_start;
  a.nsys = c->sys.nsys;
  a.n_waves = c->sys.n_waves;
  a.n_waves2 = c->sys.n_waves2;
  a.x = st.x;
  a.y = st.y;
  a.z = st.z;
  a.vx = st.vx;
  a.vy = st.vy;
  a.vz = st.vz;
  a.ax = c->ax;
  a.ay = c->ay;
  a.az = c->az;
  a.interactions = c->inter;
  a.flags = c->flags;
  a.flags_out = c->flags;
  a.theta = c->p.theta;
  a.dt = c->p.dt;
  a.bbox = c->sys.bbox;
  a.sys_dt = c->sys.dt;
  a.sys_time = c->sys.time;
  a.steps_done = c->steps_done;
  a.inter_total = c->inter_total;
  a.fused_update = fused;
  a.groups = c->walk_groups;
  a.count_mode = c->count_mode;
  c->lane_counts = c->count_mode == kCountLanes;
  // the hand-written loop addresses node records with 32-bit byte offsets
  a.asm_loop = !c->walk_generic && ((size_t)c->node_cap + 1) * sizeof(WalkNode) < (1ull << 32);
  if (c->walk_split > 1) {
    // the pieces take their own order (the traversal's pieces on
    // consecutive workgroups); no tile costs
    a.split = c->walk_split;
    a.part = c->walk_part;
    a.cpart = c->walk_cpart;
    a.n = c->n;
  } else if (c->walk_order) {
    const int tiles = sync_tile_costs(c);
    launch_tile_order(c->tile_cost, tiles, c->walk_groups, c->tile_order, c->stream);
    a.order = c->tile_order;
    a.tile_cost = c->tile_cost;
  }
  launch_walk(a, c->cfg.force_precision == NBS_FORCE_F32, c->p.damping,
              c->walk_block, c->walk_pack, c->stream);
  if (fused) c->bbox_valid = true;
  prof_end(c, kPhaseWalk, e0);
}

void enqueue_adaptive_update(nbs_ctx* c) {
  hipEvent_t e0 = nullptr;
  prof_begin(c, kPhaseUpdate, &e0);
  State& st = c->st[c->cur];
  launch_adaptive_dt(st, c->ax, c->ay, c->az, c->n, c->sys, c->p, c->flags,
                     c->stream);
  launch_update(st, c->ax, c->ay, c->az, c->n, c->sys, c->flags, c->stream);
  c->bbox_valid = true;
  prof_end(c, kPhaseUpdate, e0);
}

// The walk and update of one step (after its build).
void enqueue_step_tail(nbs_ctx* c) {
  if (c->p.adaptive) {
    enqueue_walk(c, 0);
    enqueue_adaptive_update(c);
  } else {
    enqueue_walk(c, 1);
  }
}

void drop_graph(nbs_ctx* c) {
  if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  c->graph_exec = nullptr;
  c->graph = nullptr;
  c->graph_cur = -1;
}

int read_flags(nbs_ctx* c, int* flags, int* done) {
  int h[2] = {0, 0};
  HIPCHK(c, hipMemcpyAsync(&h[0], c->flags, sizeof(int), hipMemcpyDeviceToHost,
                           c->stream));
  HIPCHK(c, hipMemcpyAsync(&h[1], c->steps_done, sizeof(int),
                           hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  *flags = h[0];
  *done = h[1];
  return NBS_OK;
}

void reset_flags(nbs_ctx* c) {
  launch_fill_int(c->flags, 0, 1, c->stream);
  launch_fill_int(c->steps_done, 0, 1, c->stream);
}

// Device flags that larger buffers cannot cure.  kFlagStack: a walk's
// stack of waiting lanes (64 entries, one per lane; a tree is at most 43
// levels deep) or the minimum-size fold's stack ran past its end, which
// only a corrupt tree can cause.  Node and pair overflows return NBS_OK
// here: the caller grows that buffer and redoes the step.
int fatal_flags(nbs_ctx* c, int flags) {
  if (c->thr_err) {  // the threaded-root build's host part failed
    const int e = c->thr_err;
    c->thr_err = 0;
    return fail(c, e, c->thr_msg);
  }
  if (flags & kFlagStack)
    return fail(c, NBS_ERR_OVERFLOW,
                "tree walk stack overflow (a tree deeper than its 43 levels: corrupt build)");
  return NBS_OK;
}

// ------------------------------------------------------------ collide --
// Buffers of collision merging: per particle (cap_n), per pair (the
// wanted capacity) and the read-back info (nsys + 2 ints).
int ensure_collide(nbs_ctx* c) {
  hipError_t e = hipSuccess;
  CollideBuffers& b = c->cb;
  if (c->col_cap_n < c->cap_n) {
    free_all(c->col_bufs);
    c->col_info_cap = 0;
    const int cap = c->cap_n;
    auto& o = c->col_bufs;
    b.parent = dalloc<int>(o, cap, &e);
    b.mark = dalloc<int>(o, cap, &e);
    b.alive = dalloc<int>(o, cap, &e);
    b.id_alive = dalloc<int>(o, cap, &e);
    b.newpos = dalloc<int>(o, (size_t)cap + 1, &e);
    b.newid = dalloc<int>(o, (size_t)cap + 1, &e);
    b.cax = dalloc<double>(o, cap, &e);
    b.cay = dalloc<double>(o, cap, &e);
    b.caz = dalloc<double>(o, cap, &e);
    b.cinter = dalloc<int>(o, cap, &e);
    b.count = dalloc<int>(o, 2, &e);
    if (e != hipSuccess) {
      free_all(c->col_bufs);
      c->col_cap_n = 0;
      return hip_fail(c, e, "hipMalloc(collision buffers)");
    }
    c->col_cap_n = cap;
  }
  if (c->col_info_cap < c->sys.nsys + 2) {
    const int m = c->sys.nsys + 2;
    b.info = dalloc<int>(c->col_bufs, m, &e);
    if (e != hipSuccess) return hip_fail(c, e, "hipMalloc(collision info)");
    c->col_info_cap = m;
  }
  const int want = std::max(c->col_pair_want, std::max(4096, c->n / 16));
  if (b.pair_cap < want) {
    free_all(c->col_pair_bufs);
    b.pair_cap = 0;
    auto& o = c->col_pair_bufs;
    const size_t L = 2 * (size_t)want;
    b.pairs = dalloc<int2>(o, want, &e);
    b.key0 = dalloc<unsigned long long>(o, L, &e);
    b.key1 = dalloc<unsigned long long>(o, L, &e);
    b.val0 = dalloc<unsigned>(o, L, &e);
    b.val1 = dalloc<unsigned>(o, L, &e);
    b.sort_tmp = dalloc<int>(o, collide_sort_tmp_ints(want), &e);
    if (e != hipSuccess) {
      free_all(c->col_pair_bufs);
      return hip_fail(c, e, "hipMalloc(collision pair buffers)");
    }
    b.pair_cap = want;
  }
  b.scan_tmp = c->bb.scan_tmp;
  return NBS_OK;
}

// Host copies of the system layout after a compaction moved the starts
// (the device starts were moved by the compaction itself).
int shrink_systems(nbs_ctx* c, const std::vector<int>& info) {
  const int nsys = c->sys.nsys;
  for (int k = 0; k <= nsys; k++) c->h_start[k] = info[1 + k];
  for (int k = 0; k < nsys; k++) {
    const int cnt = c->h_start[k + 1] - c->h_start[k];
    c->h_wave_start[k + 1] = c->h_wave_start[k] + (cnt + kWave - 1) / kWave;
    c->h_wave2_start[k + 1] = c->h_wave2_start[k] + (cnt + 2 * kWave - 1) / (2 * kWave);
  }
  c->sys.n_waves = c->h_wave_start[nsys];
  c->sys.n_waves2 = c->h_wave2_start[nsys];
  if (!c->sys_off_req.empty())
    for (int k = 0; k <= nsys; k++) c->sys_off_req[k] = c->h_start[k];
  c->n = c->h_start[nsys];
  HIPCHK(c, hipMemcpyAsync(c->sys.wave_start, c->h_wave_start.data(), (nsys + 1) * sizeof(int),
                           hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->sys.wave2_start, c->h_wave2_start.data(),
                           (nsys + 1) * sizeof(int), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return NBS_OK;
}

// Steps with collision merging, one at a time (src/main.cpp:1420-1492 with
// config.collide): build, walk (forces only), the pruned walk that records
// the pairs, merge, update, compact.
// The particle count is read back after each step; a step whose pair or
// node buffers overflowed changed nothing but the slot order, and is
// redone with larger buffers.
int run_collide(nbs_ctx* c, int nsteps) {
  if (!c->has_state) return fail(c, NBS_ERR_STATE, "no particle state set");
  if (c->min_size > 0 || c->sys.nsys > 1) {
    const int r = check_threaded(c, c->min_size > 0 ? "min_node_size" : "batched systems");
    if (r) return r;
  }
  if (c->p.adaptive)
    return fail(c, NBS_ERR_STATE, "merging particles needs adaptive dt off");
  for (int k = 0; k < nsteps && c->n > 0; k++) {
    // once this step's minimum-size merge is done, a redo (pair buffer
    // grown) keeps it and its root
    bool min_done = false;
    for (int attempt = 0;; attempt++) {
      if (attempt == 8) return fail(c, NBS_ERR_OVERFLOW, "collision step overflow after retries");
      int r = ensure_nodes(c);
      if (r) return r;
      r = ensure_collide(c);
      if (r) return r;
      reset_flags(c);
      launch_fill_int(c->cb.count, 0, 2, c->stream);
      enqueue_build(c, min_done);
      if (c->min_size > 0 && !min_done) {
        // apply_minimum_size on this build, then the tree of the merged
        // particles on the same root (src/main.cpp:1421-1434)
        const int n0 = c->n;
        launch_min_size(c->cb, c->nodes, c->node_cap, c->min_size, c->st[c->cur],
                        c->st[1 - c->cur], c->ax, c->ay, c->az, c->inter, n0, c->sys, c->flags,
                        c->stream);
        int flags = 0, done = 0;
        r = read_flags(c, &flags, &done);
        if (r) return r;
        c->bbox_valid = false;
        c->tree_valid = false;
        if (flags) {
          r = fatal_flags(c, flags);
          if (r) return r;
          r = alloc_nodes(c, c->node_cap * 2);
          if (r) return r;
          continue;
        }
        std::vector<int> info(c->sys.nsys + 2, 0);
        HIPCHK(c, hipMemcpy(info.data(), c->cb.info, info.size() * sizeof(int),
                            hipMemcpyDeviceToHost));
        c->cur ^= 1;
        min_done = true;
        c->min_removed += n0 - info[0];
        if (info[0] != n0) {
          drop_graph(c);
          r = shrink_systems(c, info);
          if (r) return r;
        }
        enqueue_build(c, true);
        if (!c->collide) {
          enqueue_walk(c, 1);
          r = read_flags(c, &flags, &done);
          if (r) return r;
          if (flags)
            return fail(c, NBS_ERR_OVERFLOW, "tree overflow after the minimum-size merge");
          c->bbox_valid = true;
          c->acc_valid = true;
          break;
        }
      }
      const int n = c->n, nsys = c->sys.nsys;
      enqueue_walk(c, 0);
      hipEvent_t e0 = nullptr;
      prof_begin(c, kPhaseUpdate, &e0);
      State& cur = c->st[c->cur];
      State& nxt = c->st[1 - c->cur];
      DetectArgs d{};
      d.nodes = c->nodes;
      d.node_start = c->sys.node_start;
      d.sys_start = c->sys.start;
      d.wave_start = c->sys.wave_start;
      d.nsys = c->sys.nsys;
      d.n_waves = c->sys.n_waves;
      d.x = cur.x;
      d.y = cur.y;
      d.z = cur.z;
      d.root = c->sys.root;
      d.boxes = c->p.root_mode == NBS_ROOT_BBOX;
      d.theta = c->p.theta;
      d.range = c->collide_range;
      d.pairs = c->cb.pairs;
      d.pair_count = c->cb.count;
      d.pair_cap = c->cb.pair_cap;
      d.flags = c->flags;
      d.flags_out = c->flags;
      launch_detect(d, c->stream);
      launch_collide(c->cb, cur, nxt, c->ax, c->ay, c->az, c->inter, n, c->sys, c->p.dt,
                     c->flags, c->stream);
      prof_end(c, kPhaseUpdate, e0);
      HIPCHK(c, hipGetLastError());
      int flags = 0, done = 0;
      r = read_flags(c, &flags, &done);
      if (r) return r;
      int cnt[2] = {0, 0};
      std::vector<int> info(nsys + 2, 0);
      HIPCHK(c, hipMemcpy(cnt, c->cb.count, sizeof(cnt), hipMemcpyDeviceToHost));
      c->bbox_valid = false;
      c->tree_valid = false;
      if (flags) {
        r = fatal_flags(c, flags);
        if (r) return r;
        if (flags & kFlagPairs)
          c->col_pair_want = std::max(2 * c->cb.pair_cap, cnt[0] + cnt[0] / 4);
        if (flags & kFlagNodeOverflow) {
          r = alloc_nodes(c, c->node_cap * 2);
          if (r) return r;
        }
        continue;
      }
      HIPCHK(c, hipMemcpy(info.data(), c->cb.info, info.size() * sizeof(int),
                          hipMemcpyDeviceToHost));
      const size_t m = (size_t)n;
      HIPCHK(c, hipMemcpyAsync(c->ax, c->cb.cax, m * sizeof(double), hipMemcpyDeviceToDevice,
                               c->stream));
      HIPCHK(c, hipMemcpyAsync(c->ay, c->cb.cay, m * sizeof(double), hipMemcpyDeviceToDevice,
                               c->stream));
      HIPCHK(c, hipMemcpyAsync(c->az, c->cb.caz, m * sizeof(double), hipMemcpyDeviceToDevice,
                               c->stream));
      HIPCHK(c, hipMemcpyAsync(c->inter, c->cb.cinter, m * sizeof(int), hipMemcpyDeviceToDevice,
                               c->stream));
      c->cur ^= 1;
      c->col_pairs += cnt[0];
      c->col_removed += n - info[0];
      if (info[0] != n) {
        drop_graph(c);
        r = shrink_systems(c, info);
        if (r) return r;
      }
      c->acc_valid = true;
      break;
    }
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return NBS_OK;
}

// Runs `nsteps` steps (or one force evaluation when nsteps == 0), growing
// the node arrays and redoing the remaining steps on overflow.
int run(nbs_ctx* c, int nsteps) {
  if (!c->has_state) return fail(c, NBS_ERR_STATE, "no particle state set");
  if (c->n == 0) return NBS_OK;
  if (c->sys.nsys > 1) {
    const int r = check_threaded(c, "batched systems");
    if (r) return r;
  }
  int r = ensure_nodes(c);
  if (r) return r;
  const bool forces_only = nsteps == 0;
  int remaining = forces_only ? 1 : nsteps;
  for (int attempt = 0; attempt < 6; attempt++) {
    reset_flags(c);
    for (int k = 0; k < remaining; k++) {
      enqueue_build(c);
      if (forces_only)
        enqueue_walk(c, 0);
      else
        enqueue_step_tail(c);
    }
    int flags = 0, done = 0;
    r = read_flags(c, &flags, &done);
    if (r) return r;
    if (!flags) {
      c->tree_valid = true;
      c->acc_valid = true;
      return NBS_OK;
    }
    c->bbox_valid = false;
    c->tree_valid = false;
    r = fatal_flags(c, flags);
    if (r) return r;
    remaining -= done;
    r = alloc_nodes(c, c->node_cap * 2);
    if (r) return r;
  }
  return fail(c, NBS_ERR_OVERFLOW, "node capacity overflow after retries");
}

// Enqueues the two-step graph replays for an even number of steps.  The
// first call captures the graph (steps enqueued while capturing are not
// executed by the capture itself).
int run_graph_pairs(nbs_ctx* c, int pairs) {
  if (!c->graph_exec || c->graph_cur != c->cur) {
    drop_graph(c);
    const bool prof = c->prof;
    c->prof = false;  // no event records inside a captured graph
    const int cur0 = c->cur;
    const bool bbox0 = c->bbox_valid;
    if (c->walk_order) sync_tile_costs(c);  // no clearing inside the graph
    HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    for (int k = 0; k < 2; k++) {
      enqueue_build(c);
      enqueue_step_tail(c);
    }
    hipError_t e = hipStreamEndCapture(c->stream, &c->graph);
    c->prof = prof;
    c->cur = cur0;
    c->bbox_valid = bbox0;
    if (e != hipSuccess) return hip_fail(c, e, "hipStreamEndCapture");
    e = hipGraphInstantiate(&c->graph_exec, c->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      drop_graph(c);
      return hip_fail(c, e, "hipGraphInstantiate");
    }
    c->graph_cur = cur0;
  }
  for (int k = 0; k < pairs; k++) HIPCHK(c, hipGraphLaunch(c->graph_exec, c->stream));
  c->bbox_valid = true;
  return NBS_OK;
}

}  // namespace

namespace {

void free_snapshots(nbs_ctx* c) {
  if (c->copy_stream) (void)hipStreamSynchronize(c->copy_stream);
  if (c->snap_dev) (void)hipFree(c->snap_dev);
  for (double*& h : c->snap_host) {
    if (h) (void)hipHostFree(h);
    h = nullptr;
  }
  c->snap_dev = nullptr;
  c->snap_cap = 0;
  c->snap_pending = -1;
}

int set_state_from(nbs_ctx* c, size_t n, const double* pos, const double* vel,
                   const double* mass, hipMemcpyKind kind) {
  if (!c) return NBS_ERR_ARG;
  if (n > 0 && (!pos || !mass)) return fail(c, NBS_ERR_ARG, "pos and mass are required");
  if (n > (size_t)0x3fffffff) return fail(c, NBS_ERR_ARG, "n too large");
  HIPCHK(c, hipSetDevice(c->device));
  drop_graph(c);
  int r = alloc_particles(c, (int)n);
  if (r) return r;
  c->n = (int)n;
  c->cur = 0;
  if (n > 0) {
    HIPCHK(c, hipMemcpyAsync(c->stage, pos, 3 * n * sizeof(double),
                             kind, c->stream));
    if (vel)
      HIPCHK(c, hipMemcpyAsync(c->stage_v, vel, 3 * n * sizeof(double),
                               kind, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->stage_m, mass, n * sizeof(double),
                             kind, c->stream));
    deinterleave_kernel<<<(int)((n + 255) / 256), 256, 0, c->stream>>>(
        c->stage, c->stage_v, c->stage_m, (int)n, c->st[0], vel ? 1 : 0);
    launch_fill_int(c->inter, 0, n, c->stream);
  }
  c->nsys_req = 1;
  c->sys_off_req.clear();
  r = setup_systems(c);
  if (r) return r;
  c->has_state = true;
  c->acc_valid = false;
  return NBS_OK;
}


// The state in caller order, [n][3] each, into host arrays or (device)
// into device arrays of this library's runtime: interleaved straight into
// the caller's array on the device, through the staging buffers to the
// host.  Returns after the copies have completed.
int get_state_into(nbs_ctx* c, double* pos, double* vel, double* acc, bool device) {
  if (!c) return NBS_ERR_ARG;
  if (!c->has_state) return fail(c, NBS_ERR_STATE, "no particle state set");
  HIPCHK(c, hipSetDevice(c->device));
  const int n = c->n;
  if (n == 0) return NBS_OK;
  State& st = c->st[c->cur];
  const int blocks = (n + 255) / 256;
  const size_t bytes = 3 * (size_t)n * sizeof(double);
  if (pos) {
    interleave_kernel<<<blocks, 256, 0, c->stream>>>(st.x, st.y, st.z, st.id, n,
                                                     device ? pos : c->stage);
    if (!device)
      HIPCHK(c, hipMemcpyAsync(pos, c->stage, bytes, hipMemcpyDeviceToHost, c->stream));
  }
  if (vel) {
    interleave_kernel<<<blocks, 256, 0, c->stream>>>(st.vx, st.vy, st.vz, st.id, n,
                                                     device ? vel : c->stage_v);
    if (!device)
      HIPCHK(c, hipMemcpyAsync(vel, c->stage_v, bytes, hipMemcpyDeviceToHost, c->stream));
  }
  if (acc) {
    if (!c->acc_valid) {
      if (device)
        launch_fill_u64((unsigned long long*)acc, 0, 3 * (size_t)n, c->stream);
      else
        std::memset(acc, 0, bytes);
    } else {
      interleave_kernel<<<blocks, 256, 0, c->stream>>>(c->ax, c->ay, c->az, st.id, n,
                                                       device ? acc : c->stage);
      if (!device)
        HIPCHK(c, hipMemcpyAsync(acc, c->stage, bytes, hipMemcpyDeviceToHost, c->stream));
    }
  }
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return NBS_OK;
}

// Every non-null array is device memory of the HIP runtime this library is
// bound to (a pointer of another runtime in the same process -- e.g. a
// PyTorch tensor when PyTorch loaded its own HIP after this library -- or a
// host pointer is refused before any copy).
int check_device_arrays(nbs_ctx* c, std::initializer_list<const void*> ptrs) {
  for (const void* p : ptrs) {
    if (!p) continue;
    hipPointerAttribute_t at{};
    const hipError_t e = hipPointerGetAttributes(&at, p);
    if (e != hipSuccess || (at.type != hipMemoryTypeDevice && at.type != hipMemoryTypeManaged)) {
      (void)hipGetLastError();
      return fail(c, NBS_ERR_ARG,
                  "not device memory of this library's HIP runtime (a host pointer, or "
                  "memory of another HIP runtime loaded in the same process)");
    }
  }
  return NBS_OK;
}

}  // namespace

extern "C" {

void nbs_config_default(nbs_config* cfg) {
  std::memset(cfg, 0, sizeof(*cfg));
  cfg->theta = 0.5;
  cfg->dt = 0.0333;
  cfg->damping = 0;
  cfg->force_precision = NBS_FORCE_F64;
  cfg->root_mode = NBS_ROOT_BBOX;
  cfg->adaptive = 0;
  cfg->root_size = 512;
  cfg->min_adaptive_dt = 0.00333;
  cfg->max_pos_change = 3;
  cfg->max_vel_change = 3;
  cfg->device = 0;
  cfg->node_factor = 4;
}

int nbs_create(const nbs_config* cfg, nbs_ctx** out) {
  if (!cfg || !out) return NBS_ERR_ARG;
  *out = nullptr;
  nbs_ctx* c = new nbs_ctx();
  int r = check_config(c, cfg);
  if (r) {
    delete c;
    return r;
  }
  c->cfg = *cfg;
  c->p = params_of(*cfg);
  c->device = cfg->device;
  hipError_t e = hipSetDevice(c->device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete c;
    return NBS_ERR_HIP;
  }
  e = hipSuccess;
  c->flags = dalloc<int>(c->misc_bufs, 1, &e);
  c->steps_done = dalloc<int>(c->misc_bufs, 1, &e);
  c->bb.flags = c->flags;
  c->bb.stats = dalloc<int>(c->misc_bufs, 2, &e);
  c->stats = c->bb.stats;
  if (e != hipSuccess) {
    free_all(c->misc_bufs);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return NBS_ERR_NOMEM;
  }
  launch_fill_int(c->flags, 0, 1, c->stream);
  launch_fill_int(c->stats, 0, 2, c->stream);
  *out = c;
  return NBS_OK;
}

void nbs_destroy(nbs_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  drop_graph(c);
  free_all(c->adj_bufs);
  for (auto& v : c->ev)
    for (auto& p : v) {
      (void)hipEventDestroy(p.first);
      (void)hipEventDestroy(p.second);
    }
  free_all(c->part_bufs);
  free_all(c->build_bufs);
  free_all(c->thr_bufs);
  free_all(c->thr_ebufs);
  free_all(c->node_bufs);
  free_all(c->sys_bufs);
  free_all(c->split_bufs);
  free_all(c->misc_bufs);
  free_all(c->col_bufs);
  free_all(c->col_pair_bufs);
  free_all(c->img_bufs);
  free_snapshots(c);
  if (c->copy_stream) (void)hipStreamDestroy(c->copy_stream);
  if (c->snap_ready) (void)hipEventDestroy(c->snap_ready);
  if (c->snap_done) (void)hipEventDestroy(c->snap_done);
  (void)hipStreamDestroy(c->stream);
  delete c;
}

const char* nbs_last_error(const nbs_ctx* c) {
  return c ? c->err.c_str() : "null context";
}

int nbs_configure(nbs_ctx* c, const nbs_config* cfg) {
  if (!c || !cfg) return NBS_ERR_ARG;
  int r = check_config(c, cfg);
  if (r) return r;
  const int old_node_factor = c->cfg.node_factor;
  drop_graph(c);
  c->cfg = *cfg;
  c->p = params_of(*cfg);
  if (cfg->node_factor != old_node_factor) c->node_cap = 0;
  c->tree_valid = false;
  return NBS_OK;
}

int nbs_set_state(nbs_ctx* c, size_t n, const double* pos, const double* vel,
                  const double* mass) {
  return set_state_from(c, n, pos, vel, mass, hipMemcpyHostToDevice);
}

int nbs_set_state_device(nbs_ctx* c, size_t n, const double* pos, const double* vel,
                         const double* mass) {
  if (!c) return NBS_ERR_ARG;
  if (n > 0) {
    int r = check_device_arrays(c, {pos, vel, mass});
    if (r) return r;
  }
  int r = set_state_from(c, n, pos, vel, mass, hipMemcpyDeviceToDevice);
  if (r) return r;
  // the caller may reuse its arrays as soon as this returns
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return NBS_OK;
}

int nbs_set_systems(nbs_ctx* c, int nsys, const int64_t* off) {
  if (!c || nsys < 1 || !off) return NBS_ERR_ARG;
  if (!c->has_state) return fail(c, NBS_ERR_STATE, "set state before systems");
  if (off[0] != 0 || off[nsys] != c->n)
    return fail(c, NBS_ERR_ARG, "offsets must run from 0 to n");
  for (int k = 0; k < nsys; k++)
    if (off[k + 1] <= off[k])
      return fail(c, NBS_ERR_ARG, "offsets must increase (every system holds a particle)");
  // Particles are in caller order only until the first build.
  if (c->cur != 0 || c->tree_valid)
    return fail(c, NBS_ERR_STATE, "set systems right after set_state");
  c->nsys_req = nsys;
  c->sys_off_req.assign(off, off + nsys + 1);
  return setup_systems(c);
}

int nbs_generate_state(nbs_ctx* c, const nbs_settings* s, int nsys, const int64_t* off,
                       const uint64_t* seeds, const uint32_t* gen_types) {
  if (!c) return NBS_ERR_ARG;
  if (!s || nsys < 1 || !off || !seeds)
    return fail(c, NBS_ERR_ARG, "settings, nsys >= 1, offsets and seeds are required");
  if (off[0] != 0) return fail(c, NBS_ERR_ARG, "offsets must start at 0");
  for (int k = 0; k < nsys; k++)
    if (off[k + 1] <= off[k] && !(nsys == 1 && off[1] == 0))
      return fail(c, NBS_ERR_ARG, "offsets must increase (every system holds a particle)");
  const int64_t n = off[nsys];
  if (n > 0x3fffffff) return fail(c, NBS_ERR_ARG, "n too large");
  if (s->mass_dist != NBS_DIST_LINEAR)
    return fail(c, NBS_ERR_ARG,
                "device generation needs mass_dist linear (exp and normal draw a "
                "variable number of words; use nbs_generate)");
  std::vector<IcSystem> sys(nsys);
  for (int k = 0; k < nsys; k++) {
    const uint32_t t = gen_types ? gen_types[k] : s->gen_type;
    if (t > NBS_GEN_SHELL) return fail(c, NBS_ERR_ARG, "gen_type out of range");
    if (t == NBS_GEN_CUBE && s->vel_dist != NBS_DIST_LINEAR)
      return fail(c, NBS_ERR_ARG, "device generation of cubes needs vel_dist linear");
    sys[k].gen_type = (int)t;
    // two engine words per double draw: mass, then the gen type's draws
    sys[k].draws = 2 + (t == NBS_GEN_CUBE ? 14 : t == NBS_GEN_SPHERE ? 6
                                            : (s->shell_jitter != 0 ? 6 : 4));
    // minstd_rand0::seed: seed mod 2^31 - 1, with 0 replaced by 1
    const uint64_t x0 = seeds[k] % 2147483647ull;
    sys[k].x0 = (unsigned)(x0 ? x0 : 1);
    sys[k].pad = 0;
  }
  IcParams p;
  p.min_mass = s->min_mass;
  p.max_mass = s->max_mass;
  p.size = s->size;
  p.min_vel = s->min_vel;
  p.max_vel = s->max_vel;
  p.r_sphere = s->r_sphere;
  p.shell_jitter = s->shell_jitter;
  p.scale[0] = s->scale_x;
  p.scale[1] = s->scale_y;
  p.scale[2] = s->scale_z;
  {  // the rotation vector as generate_particle's caller forms it
    const double r0 = s->rotation_vector[0], r1 = s->rotation_vector[1],
                 r2 = s->rotation_vector[2];
    const double rm = std::sqrt(r0 * r0 + r1 * r1 + r2 * r2);
    for (int k = 0; k < 3; k++) p.rot[k] = s->rotation_vector[k] / rm * s->rotation_magnitude;
  }
  HIPCHK(c, hipSetDevice(c->device));
  drop_graph(c);
  int r = alloc_particles(c, (int)n);
  if (r) return r;
  c->n = (int)n;
  c->cur = 0;
  if (n > 0) {
    std::vector<long long> h_off(off, off + nsys + 1);
    const size_t off_bytes = (nsys + 1) * sizeof(long long);
    void* tmp = nullptr;
    HIPCHK(c, hipMalloc(&tmp, off_bytes + nsys * sizeof(IcSystem)));
    long long* d_off = (long long*)tmp;
    IcSystem* d_sys = (IcSystem*)((char*)tmp + off_bytes);
    hipError_t e = hipMemcpyAsync(d_off, h_off.data(), off_bytes, hipMemcpyHostToDevice,
                                  c->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(d_sys, sys.data(), nsys * sizeof(IcSystem), hipMemcpyHostToDevice,
                         c->stream);
    if (e == hipSuccess) {
      launch_generate(p, nsys, d_off, d_sys, n, c->st[0], c->stream);
      launch_fill_int(c->inter, 0, n, c->stream);
      e = hipGetLastError();
    }
    const hipError_t e2 = hipStreamSynchronize(c->stream);
    (void)hipFree(tmp);
    if (e == hipSuccess) e = e2;
    if (e != hipSuccess) return hip_fail(c, e, "device generation");
  }
  c->nsys_req = nsys;
  if (nsys > 1)
    c->sys_off_req.assign(off, off + nsys + 1);
  else
    c->sys_off_req.clear();
  r = setup_systems(c);
  if (r) return r;
  c->has_state = true;
  c->acc_valid = false;
  return NBS_OK;
}

int nbs_get_state(nbs_ctx* c, double* pos, double* vel, double* acc) {
  return get_state_into(c, pos, vel, acc, false);
}

int nbs_get_state_device(nbs_ctx* c, double* pos, double* vel, double* acc) {
  if (!c) return NBS_ERR_ARG;
  if (c->has_state && c->n > 0) {
    int r = check_device_arrays(c, {pos, vel, acc});
    if (r) return r;
  }
  return get_state_into(c, pos, vel, acc, true);
}

int nbs_get_mass(nbs_ctx* c, double* mass) {
  if (!c || !mass) return NBS_ERR_ARG;
  if (!c->has_state) return fail(c, NBS_ERR_STATE, "no particle state set");
  HIPCHK(c, hipSetDevice(c->device));
  const int n = c->n;
  if (n == 0) return NBS_OK;
  const State& st = c->st[c->cur];
  scatter_kernel<<<(n + 255) / 256, 256, 0, c->stream>>>(st.m, st.id, n, c->stage_m);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(mass, c->stage_m, (size_t)n * sizeof(double),
                           hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return NBS_OK;
}

int nbs_snapshot_begin(nbs_ctx* c) {
  if (!c) return NBS_ERR_ARG;
  if (!c->has_state) return fail(c, NBS_ERR_STATE, "no particle state set");
  if (c->snap_pending >= 0)
    return fail(c, NBS_ERR_STATE, "end the pending snapshot first");
  HIPCHK(c, hipSetDevice(c->device));
  const size_t n = (size_t)c->n;
  if (!c->copy_stream)
    HIPCHK(c, hipStreamCreateWithFlags(&c->copy_stream, hipStreamNonBlocking));
  if (!c->snap_ready) HIPCHK(c, hipEventCreateWithFlags(&c->snap_ready, hipEventDisableTiming));
  if (!c->snap_done) HIPCHK(c, hipEventCreateWithFlags(&c->snap_done, hipEventDisableTiming));
  if (n > c->snap_cap) {
    free_snapshots(c);
    hipError_t e = hipMalloc(&c->snap_dev, 9 * n * sizeof(double));
    for (double*& h : c->snap_host)
      if (e == hipSuccess) e = hipHostMalloc(&h, 9 * n * sizeof(double), hipHostMallocDefault);
    if (e != hipSuccess) {
      free_snapshots(c);
      return hip_fail(c, e, "snapshot buffers");
    }
    c->snap_cap = n;
  }
  const int b = c->snap_next;
  if (n > 0) {
    const State& st = c->st[c->cur];
    const int blocks = (int)((n + 255) / 256);
    double* d = c->snap_dev;
    interleave_kernel<<<blocks, 256, 0, c->stream>>>(st.x, st.y, st.z, st.id, (int)n, d);
    interleave_kernel<<<blocks, 256, 0, c->stream>>>(st.vx, st.vy, st.vz, st.id, (int)n,
                                                     d + 3 * n);
    if (c->acc_valid)
      interleave_kernel<<<blocks, 256, 0, c->stream>>>(c->ax, c->ay, c->az, st.id, (int)n,
                                                       d + 6 * n);
    else
      launch_fill_u64((unsigned long long*)(d + 6 * n), 0, 3 * n, c->stream);
    HIPCHK(c, hipGetLastError());
  }
  HIPCHK(c, hipEventRecord(c->snap_ready, c->stream));
  HIPCHK(c, hipStreamWaitEvent(c->copy_stream, c->snap_ready, 0));
  if (n > 0)
    HIPCHK(c, hipMemcpyAsync(c->snap_host[b], c->snap_dev, 9 * n * sizeof(double),
                             hipMemcpyDeviceToHost, c->copy_stream));
  HIPCHK(c, hipEventRecord(c->snap_done, c->copy_stream));
  c->snap_pending = b;
  c->snap_n = (int)n;
  c->snap_next = b ^ 1;
  return NBS_OK;
}

int nbs_snapshot_end(nbs_ctx* c, const double** pos, const double** vel,
                     const double** acc, size_t* n) {
  if (!c) return NBS_ERR_ARG;
  if (c->snap_pending < 0) return fail(c, NBS_ERR_STATE, "no snapshot pending");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipEventSynchronize(c->snap_done));
  const double* h = c->snap_host[c->snap_pending];
  const size_t m = (size_t)c->snap_n;
  c->snap_pending = -1;
  if (pos) *pos = h;
  if (vel) *vel = h ? h + 3 * m : nullptr;
  if (acc) *acc = h ? h + 6 * m : nullptr;
  if (n) *n = m;
  return NBS_OK;
}

int nbs_set_collisions(nbs_ctx* c, int enabled, double range) {
  if (!c) return NBS_ERR_ARG;
  if (enabled && !(range >= 0)) return fail(c, NBS_ERR_ARG, "collision range must be >= 0");
  c->collide = enabled != 0;
  c->collide_range = enabled ? range : 0.0;
  c->col_pairs = 0;
  c->col_removed = 0;
  drop_graph(c);
  return NBS_OK;
}

int nbs_set_min_node_size(nbs_ctx* c, double min_size, int64_t* merged) {
  if (!c) return NBS_ERR_ARG;
  if (merged) *merged = c->min_removed;
  if (!(min_size >= 0)) return fail(c, NBS_ERR_ARG, "min_node_size must be >= 0");
  if (min_size != c->min_size) c->min_removed = 0;
  c->min_size = min_size;
  drop_graph(c);
  return NBS_OK;
}

int nbs_collision_stats(nbs_ctx* c, int64_t* pairs, int64_t* removed) {
  if (!c) return NBS_ERR_ARG;
  if (pairs) *pairs = c->col_pairs;
  if (removed) *removed = c->col_removed;
  return NBS_OK;
}

int nbs_render_image(nbs_ctx* c, const nbs_image_params* ip, uint8_t* rgb) {
  if (!c || !ip || !rgb) return NBS_ERR_ARG;
  if (!c->has_state) return fail(c, NBS_ERR_STATE, "no particle state set");
  if (ip->width < 1 || ip->height < 1 || ip->projection > 3 || ip->color > 1)
    return fail(c, NBS_ERR_ARG, "image: width, height >= 1, projection 0-3, color 0-1");
  const size_t npix = (size_t)ip->width * ip->height;
  if (npix > ((size_t)1 << 31)) return fail(c, NBS_ERR_ARG, "image too large");
  HIPCHK(c, hipSetDevice(c->device));
  if (npix > c->img_pixels) {
    free_all(c->img_bufs);
    c->img_pixels = 0;
    hipError_t e = hipSuccess;
    c->img_counts = dalloc<unsigned>(c->img_bufs, npix, &e);
    c->img_temp = dalloc<double>(c->img_bufs, npix, &e);
    c->img_partial = dalloc<double>(c->img_bufs, image_partials(npix), &e);
    c->img_max = dalloc<unsigned long long>(c->img_bufs, 1, &e);
    c->img_rgb = dalloc<unsigned char>(c->img_bufs, 3 * npix, &e);
    if (e != hipSuccess) {
      free_all(c->img_bufs);
      return hip_fail(c, e, "hipMalloc(image buffers)");
    }
    c->img_pixels = npix;
  }
  ImageParams p;
  p.width = ip->width;
  p.height = ip->height;
  p.projection = ip->projection;
  p.color = ip->color;
  p.adaptive = ip->adaptive != 0;
  p.nonlinear = ip->nonlinear != 0;
  p.scale = ip->scale;
  p.brightness = ip->brightness;
  launch_image(c->st[c->cur], c->n, p, c->img_counts, c->img_temp, c->img_partial, c->img_max,
               c->img_rgb, c->stream);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(rgb, c->img_rgb, 3 * npix, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return NBS_OK;
}

int nbs_particle_count(nbs_ctx* c, int64_t* n, int64_t* offsets, int nsys_cap) {
  if (!c) return NBS_ERR_ARG;
  if (n) *n = c->has_state ? c->n : 0;
  if (offsets) {
    if (!c->has_state) return fail(c, NBS_ERR_STATE, "no particle state set");
    const int nsys = c->sys.nsys;
    if (nsys_cap < nsys + 1) return fail(c, NBS_ERR_ARG, "offsets need nsys + 1 entries");
    for (int k = 0; k <= nsys; k++) offsets[k] = c->h_start[k];
  }
  return NBS_OK;
}

int nbs_forces(nbs_ctx* c, double* acc) {
  if (!c) return NBS_ERR_ARG;
  HIPCHK(c, hipSetDevice(c->device));
  int r = run(c, 0);
  if (r) return r;
  if (acc) return nbs_get_state(c, nullptr, nullptr, acc);
  return NBS_OK;
}

}  // extern "C"

// ------------------------------------------------- no-allocation forces --
// The device memory of nbs_forces_noalloc: every buffer of one force
// evaluation (state, keys, sort, nodes, per-wave partials, flags) is carved
// from this array, NBS_NOALLOC_MAX_N particles at node_factor 4 fit (about
// 580 B per particle plus fixed parts).  Each device holds its own copy
// until the process ends; all calls on one device share it.
namespace {
constexpr size_t kScratchBytes = 56ull << 20;
constexpr int kMaxDevices = 64;
}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices];  // the array's device address, per device

// The address of `scratch` on the current device, looked up on the first
// call there (the array's name in host code is only a host-side stand-in).
hipError_t scratch_on_device(int dev, void** out) {
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!ptr) return hipErrorInvalidSymbol;
    g_scratch_ptr[dev] = ptr;
  }
  *out = g_scratch_ptr[dev];
  return hipSuccess;
}

int noalloc_fail(char* msg, size_t msg_len, int code, const std::string& text) {
  if (msg && msg_len) snprintf(msg, msg_len, "%s", text.c_str());
  return code;
}
}  // namespace

extern "C" {

int nbs_forces_noalloc(const nbs_config* cfg, size_t n, const double* pos,
                       const double* mass, double* acc, int64_t* counts, void* stream,
                       char* msg, size_t msg_len) {
  if (msg && msg_len) msg[0] = 0;
  if (!cfg) return noalloc_fail(msg, msg_len, NBS_ERR_ARG, "cfg is required");
  if (n > (size_t)NBS_NOALLOC_MAX_N)
    return noalloc_fail(msg, msg_len, NBS_ERR_ARG,
                        "n = " + std::to_string(n) + " exceeds the no-allocation maximum of " +
                            std::to_string(NBS_NOALLOC_MAX_N) + " particles");
  if (n == 0) return NBS_OK;
  if (!pos || !mass || !acc)
    return noalloc_fail(msg, msg_len, NBS_ERR_ARG, "pos, mass and acc are required");
  nbs_ctx c;
  int r = check_config(&c, cfg);
  if (r) return noalloc_fail(msg, msg_len, r, c.err);
  if (cfg->threaded_tree)
    return noalloc_fail(msg, msg_len, NBS_ERR_ARG, "the threaded tree allocates per build");
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  void* base = nullptr;
  if (e == hipSuccess) e = scratch_on_device(dev, &base);
  if (e != hipSuccess)
    return noalloc_fail(msg, msg_len, NBS_ERR_HIP,
                        std::string("scratch address: ") + hipGetErrorString(e));
  c.cfg = *cfg;
  c.cfg.device = dev;
  c.p = params_of(c.cfg);
  c.device = dev;
  c.stream = (hipStream_t)stream;
  c.arena.base = (char*)base;
  c.arena.size = kScratchBytes;
  // the flags and counters a previous call on this device left behind are
  // cleared below (reset_flags, the stats fill, setup_systems' counters)
  e = hipSuccess;
  c.flags = carve<int>(c.arena, 1, &e);
  c.steps_done = carve<int>(c.arena, 1, &e);
  c.stats = carve<int>(c.arena, 2, &e);
  c.bb.flags = c.flags;
  c.bb.stats = c.stats;
  if (e != hipSuccess) return noalloc_fail(msg, msg_len, NBS_ERR_NOMEM, "scratch exhausted");
  launch_fill_int(c.stats, 0, 2, c.stream);
  r = alloc_particles(&c, (int)n);
  if (r == NBS_OK) {
    c.n = (int)n;
    c.cur = 0;
    deinterleave_kernel<<<(int)((n + 255) / 256), 256, 0, c.stream>>>(
        pos, nullptr, mass, (int)n, c.st[0], 0);
    launch_fill_int(c.inter, 0, n, c.stream);
    c.nsys_req = 1;
    r = setup_systems(&c);
  }
  if (r == NBS_OK) r = ensure_nodes(&c);
  if (r != NBS_OK) {
    if (r == NBS_ERR_NOMEM)
      return noalloc_fail(msg, msg_len, NBS_ERR_ARG,
                          "the no-allocation scratch (" + std::to_string(kScratchBytes) +
                              " bytes) cannot hold " + std::to_string(n) +
                              " particles at node_factor " +
                              std::to_string(cfg->node_factor > 0 ? cfg->node_factor : 4));
    return noalloc_fail(msg, msg_len, r, c.err);
  }
  c.has_state = true;
  reset_flags(&c);
  enqueue_build(&c);
  enqueue_walk(&c, 0);
  int flags = 0, done = 0;
  r = read_flags(&c, &flags, &done);
  if (r) return noalloc_fail(msg, msg_len, r, c.err);
  if (flags) {
    r = fatal_flags(&c, flags);
    if (r) return noalloc_fail(msg, msg_len, r, c.err);
    return noalloc_fail(msg, msg_len, NBS_ERR_OVERFLOW,
                        "tree needs more than the " + std::to_string(c.node_cap) +
                            " nodes the no-allocation scratch holds");
  }
  const State& st = c.st[c.cur];
  interleave_kernel<<<(int)((n + 255) / 256), 256, 0, c.stream>>>(c.ax, c.ay, c.az, st.id,
                                                                   (int)n, acc);
  if (counts)
    scatter_count_kernel<<<(int)((n + 255) / 256), 256, 0, c.stream>>>(c.inter, st.id,
                                                                       (int)n, counts);
  e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(c.stream);
  if (e != hipSuccess)
    return noalloc_fail(msg, msg_len, NBS_ERR_HIP,
                        std::string("nbs_forces_noalloc: ") + hipGetErrorString(e));
  return NBS_OK;
}

int nbs_step(nbs_ctx* c, int nsteps) {
  if (!c) return NBS_ERR_ARG;
  if (nsteps < 0) return fail(c, NBS_ERR_ARG, "nsteps must be >= 0");
  if (nsteps == 0) return NBS_OK;
  HIPCHK(c, hipSetDevice(c->device));
  if (c->collide || c->min_size > 0) return run_collide(c, nsteps);
  return run(c, nsteps);
}

int nbs_step_graph(nbs_ctx* c, int nsteps) {
  if (!c) return NBS_ERR_ARG;
  if (nsteps < 0) return fail(c, NBS_ERR_ARG, "nsteps must be >= 0");
  if (!c->has_state) return fail(c, NBS_ERR_STATE, "no particle state set");
  if (c->collide || c->min_size > 0)
    return fail(c, NBS_ERR_STATE, "merging particles runs eager steps (nbs_step)");
  if (nsteps == 0 || c->n == 0) return NBS_OK;
  HIPCHK(c, hipSetDevice(c->device));
  // the threaded-root build reads its entry count back: eager steps
  if (c->cfg.threaded_tree) return run(c, nsteps);
  int r = ensure_nodes(c);
  if (r) return r;
  reset_flags(c);
  int left = nsteps;
  // The graph assumes the bbox accumulators were filled by the previous
  // step's update; a first eager step provides that.
  if (!c->bbox_valid) {
    enqueue_build(c);
    if (c->p.adaptive) {
      enqueue_walk(c, 0);
      enqueue_adaptive_update(c);
    } else {
      enqueue_walk(c, 1);
    }
    left--;
  }
  r = run_graph_pairs(c, left / 2);
  if (r) return r;
  if (left % 2) {
    enqueue_build(c);
    if (c->p.adaptive) {
      enqueue_walk(c, 0);
      enqueue_adaptive_update(c);
    } else {
      enqueue_walk(c, 1);
    }
  }
  int flags = 0, done = 0;
  r = read_flags(c, &flags, &done);
  if (r) return r;
  if (!flags) {
    c->tree_valid = true;
    c->acc_valid = true;
    return NBS_OK;
  }
  // overflow: the failing step and all after it did nothing
  drop_graph(c);
  c->bbox_valid = false;
  c->tree_valid = false;
  r = fatal_flags(c, flags);
  if (r) return r;
  r = alloc_nodes(c, c->node_cap * 2);
  if (r) return r;
  return run(c, nsteps - done);
}

int nbs_set_walk_shape(nbs_ctx* c, int block, int groups, int pack) {
  if (!c) return NBS_ERR_ARG;
  bool okb = false, okg = false, okp = false;
  for (int b : kWalkBlocks) okb = okb || b == block;
  for (int g : kWalkGroups) okg = okg || g == groups;
  for (int q : kWalkPacks) okp = okp || q == pack;
  if (!okb) return fail(c, NBS_ERR_ARG, "walk block must be 64, 128, 256, 512 or 1024");
  if (!okg) return fail(c, NBS_ERR_ARG, "walk groups must be 1, 8, 16, 32 or 64");
  if (!okp) return fail(c, NBS_ERR_ARG, "walk pack must be 1 or 2");
  drop_graph(c);
  c->walk_block = block;
  c->walk_groups = groups;
  c->walk_pack = pack;
  return NBS_OK;
}

int nbs_set_walk_split(nbs_ctx* c, int pieces) {
  if (!c) return NBS_ERR_ARG;
  if (pieces < 0 || pieces > kWalkMaxPieces || (pieces & (pieces - 1)))
    return fail(c, NBS_ERR_ARG, "walk pieces must be 0 (by wavefront count) or a power of two "
                                "up to " + std::to_string(kWalkMaxPieces));
  c->walk_split_req = pieces;
  if (!c->has_state) return NBS_OK;
  return alloc_walk_split(c);
}

int nbs_walk_split(nbs_ctx* c) { return c ? c->walk_split : 0; }

int nbs_set_walk_code(nbs_ctx* c, int code) {
  if (!c) return NBS_ERR_ARG;
  drop_graph(c);
  if (code < 0 || code > 3)
    return fail(c, NBS_ERR_ARG, "walk code must be 0 (hand-written), 1 (generic), "
                                "2 (hand-written, 64-byte records) or 3 (32-byte records)");
  c->walk_generic = code == 1;
  c->walk_records = code == 2 ? 64 : (code == 3 ? 32 : 0);
  return NBS_OK;
}

int nbs_set_counting(nbs_ctx* c, int mode) {
  if (!c) return NBS_ERR_ARG;
  if (mode < NBS_COUNT_NONE || mode > NBS_COUNT_PARTICLES)
    return fail(c, NBS_ERR_ARG, "counting mode must be 0 (none), 1 (total) or 2 (per particle)");
  drop_graph(c);
  c->count_mode = mode == NBS_COUNT_PARTICLES ? kCountLanes
                                              : (mode == NBS_COUNT_TOTAL ? kCountWave : kCountNone);
  return NBS_OK;
}

int nbs_set_walk_order(nbs_ctx* c, int on) {
  if (!c) return NBS_ERR_ARG;
  drop_graph(c);
  c->walk_order = on ? 1 : 0;
  c->cost_tiles = -1;
  return NBS_OK;
}

int nbs_autotune(nbs_ctx* c, int reps, int* best_block, int* best_groups,
                 int* best_pack, double* ms) {
  if (!c) return NBS_ERR_ARG;
  if (!c->has_state) return fail(c, NBS_ERR_STATE, "no particle state set");
  if (c->n == 0) return NBS_OK;
  HIPCHK(c, hipSetDevice(c->device));
  reps = std::max(reps, 1);
  int r = ensure_nodes(c);
  if (r) return r;
  drop_graph(c);
  HIPCHK(c, hipStreamSynchronize(c->stream));

  // Save what the timing steps change: the current state buffer, the
  // accelerations and interaction counts of the last walk (slot order of
  // that buffer), simulated time and step length, the running total.
  const size_t n = c->n;
  const int nsys = c->sys.nsys, nw = std::max(c->sys.n_waves, 1);
  std::vector<void*> tmp;
  hipError_t e = hipSuccess;
  double* sv = dalloc<double>(tmp, 10 * n, &e);
  unsigned* sid = dalloc<unsigned>(tmp, n, &e);
  int* sint = dalloc<int>(tmp, n, &e);
  if (e != hipSuccess) {
    free_all(tmp);
    return hip_fail(c, e, "hipMalloc(autotune save)");
  }
  const int cur0 = c->cur;
  const bool acc0 = c->acc_valid;
  State& st0 = c->st[cur0];
  double* const live[10] = {st0.x, st0.y, st0.z, st0.vx, st0.vy, st0.vz,
                            st0.m, c->ax, c->ay, c->az};
  std::vector<double> h_time(nsys), h_dt(nsys);
  std::vector<unsigned long long> h_total(nw);
  auto copy_state = [&](bool save) -> hipError_t {
    hipError_t err = hipSuccess;
    for (int k = 0; k < 10 && err == hipSuccess; k++)
      err = save ? hipMemcpyAsync(sv + k * n, live[k], n * 8, hipMemcpyDeviceToDevice, c->stream)
                 : hipMemcpyAsync(live[k], sv + k * n, n * 8, hipMemcpyDeviceToDevice, c->stream);
    if (err == hipSuccess)
      err = hipMemcpyAsync(save ? (void*)sid : (void*)st0.id, save ? (void*)st0.id : (void*)sid,
                           n * sizeof(unsigned), hipMemcpyDeviceToDevice, c->stream);
    if (err == hipSuccess)
      err = hipMemcpyAsync(save ? (void*)sint : (void*)c->inter, save ? (void*)c->inter : (void*)sint,
                           n * sizeof(int), hipMemcpyDeviceToDevice, c->stream);
    c->cur = cur0;
    c->bbox_valid = false;
    return err;
  };
  e = copy_state(true);
  if (e == hipSuccess)
    e = hipMemcpyAsync(h_time.data(), c->sys.time, nsys * 8, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(h_dt.data(), c->sys.dt, nsys * 8, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(h_total.data(), c->inter_total, nw * 8, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) {
    free_all(tmp);
    return hip_fail(c, e, "autotune save");
  }

  // Each candidate: restore, one untimed step (first use of the shape), then
  // `reps` whole steps timed with events on the context's stream.
  const bool prof0 = c->prof;
  c->prof = false;
  const int block0 = c->walk_block, groups0 = c->walk_groups;
  const int pack0 = c->walk_pack;
  double best = 1e300;
  int bb = block0, bg = groups0, bp = pack0;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  const int nshapes = kNumWalkPacks * kNumWalkBlocks * kNumWalkGroups;
  for (int sh = 0; sh < nshapes && e == hipSuccess && r == NBS_OK; sh++) {
    {
      const int j = sh % kNumWalkGroups;
      const int i = (sh / kNumWalkGroups) % kNumWalkBlocks;
      const int q = sh / (kNumWalkGroups * kNumWalkBlocks);
      c->walk_block = kWalkBlocks[i];
      c->walk_groups = kWalkGroups[j];
      c->walk_pack = kWalkPacks[q];
      e = copy_state(false);
      if (e != hipSuccess) break;
      r = run(c, 1);  // grows the node arrays if needed
      if (r) break;
      e = copy_state(false);
      if (e != hipSuccess) break;
      reset_flags(c);
      enqueue_build(c);  // untimed: first use of this shape
      enqueue_step_tail(c);
      e = copy_state(false);
      if (e != hipSuccess) break;
      e = hipEventRecord(e0, c->stream);
      for (int q = 0; q < reps && e == hipSuccess; q++) {
        enqueue_build(c);
        enqueue_step_tail(c);
      }
      if (e == hipSuccess) e = hipEventRecord(e1, c->stream);
      if (e == hipSuccess) e = hipEventSynchronize(e1);
      float t = 0;
      if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
      int flags = 0, done = 0;
      if (e == hipSuccess) r = read_flags(c, &flags, &done);
      if (e != hipSuccess || r) break;
      const double per = flags ? 1e300 : t / reps;  // overflow: not timed
      if (ms) ms[sh] = flags ? -1.0 : per;
      if (per < best) {
        best = per;
        bb = kWalkBlocks[i];
        bg = kWalkGroups[j];
        bp = kWalkPacks[q];
      }
    }
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  c->prof = prof0;

  // Restore (also after a failure above).
  hipError_t e2 = copy_state(false);
  if (e2 == hipSuccess)
    e2 = hipMemcpyAsync(c->sys.time, h_time.data(), nsys * 8, hipMemcpyHostToDevice, c->stream);
  if (e2 == hipSuccess)
    e2 = hipMemcpyAsync(c->sys.dt, h_dt.data(), nsys * 8, hipMemcpyHostToDevice, c->stream);
  if (e2 == hipSuccess)
    e2 = hipMemcpyAsync(c->inter_total, h_total.data(), nw * 8, hipMemcpyHostToDevice, c->stream);
  if (e2 == hipSuccess) e2 = hipStreamSynchronize(c->stream);
  reset_flags(c);
  free_all(tmp);
  c->tree_valid = false;  // the node array holds a tuning step's tree
  c->acc_valid = acc0;
  drop_graph(c);
  if (r || e != hipSuccess || e2 != hipSuccess) {
    c->walk_block = block0;
    c->walk_groups = groups0;
    c->walk_pack = pack0;
    return r ? r : hip_fail(c, e != hipSuccess ? e : e2, "autotune");
  }
  c->walk_block = bb;
  c->walk_groups = bg;
  c->walk_pack = bp;
  if (best_block) *best_block = bb;
  if (best_groups) *best_groups = bg;
  if (best_pack) *best_pack = bp;
  return NBS_OK;
}

int nbs_synchronize(nbs_ctx* c) {
  if (!c) return NBS_ERR_ARG;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return NBS_OK;
}

int nbs_direct(nbs_ctx* c, size_t m, const int64_t* ids, double* acc) {
  if (!c) return NBS_ERR_ARG;
  if (!c->has_state) return fail(c, NBS_ERR_STATE, "no particle state set");
  if (m == 0) return NBS_OK;
  if (!ids || !acc) return fail(c, NBS_ERR_ARG, "ids and acc are required");
  for (size_t k = 0; k < m; k++)
    if (ids[k] < 0 || ids[k] >= c->n) return fail(c, NBS_ERR_ARG, "id out of range");
  HIPCHK(c, hipSetDevice(c->device));
  if ((int)m > c->direct_cap) {
    for (void* p : {(void*)c->d_ids, (void*)c->d_slots, (void*)c->d_direct_tmp,
                    (void*)c->d_direct_out})
      if (p) {
        (void)hipFree(p);
        c->misc_bufs.erase(std::remove(c->misc_bufs.begin(), c->misc_bufs.end(), p),
                           c->misc_bufs.end());
      }
    hipError_t e = hipSuccess;
    c->d_ids = dalloc<long long>(c->misc_bufs, m, &e);
    c->d_slots = dalloc<int>(c->misc_bufs, m, &e);
    c->d_direct_tmp = dalloc<double>(c->misc_bufs, direct_tmp_doubles(m), &e);
    c->d_direct_out = dalloc<double>(c->misc_bufs, 3 * m, &e);
    if (e != hipSuccess) {
      c->direct_cap = 0;
      return hip_fail(c, e, "hipMalloc(direct)");
    }
    c->direct_cap = (int)m;
  }
  State& st = c->st[c->cur];
  launch_inverse_perm(st.id, c->n, c->slot_of_id, c->stream);
  HIPCHK(c, hipMemcpyAsync(c->d_ids, ids, m * sizeof(long long), hipMemcpyHostToDevice,
                           c->stream));
  gather_slots_kernel<<<(int)((m + 255) / 256), 256, 0, c->stream>>>(
      c->d_ids, (int)m, c->slot_of_id, c->d_slots);
  launch_direct(st, c->n, c->d_slots, (int)m, c->p.damping, c->d_direct_tmp,
                c->d_direct_out, c->stream);
  HIPCHK(c, hipMemcpyAsync(acc, c->d_direct_out, 3 * m * sizeof(double),
                           hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return NBS_OK;
}

int nbs_interactions(nbs_ctx* c, int64_t* total, int64_t* per_particle) {
  if (!c) return NBS_ERR_ARG;
  if (!c->acc_valid) return fail(c, NBS_ERR_STATE, "no walk has run");
  if (!c->lane_counts)
    return fail(c, NBS_ERR_STATE,
                "the last walk did not count per particle (nbs_set_counting(ctx, 2))");
  const int n = c->n;
  std::vector<int> h(n);
  std::vector<unsigned> id(n);
  HIPCHK(c, hipMemcpyAsync(h.data(), c->inter, n * sizeof(int), hipMemcpyDeviceToHost,
                           c->stream));
  HIPCHK(c, hipMemcpyAsync(id.data(), c->st[c->cur].id, n * sizeof(unsigned),
                           hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  int64_t t = 0;
  for (int i = 0; i < n; i++) {
    t += h[i];
    if (per_particle) per_particle[id[i]] = h[i];
  }
  if (total) *total = t;
  return NBS_OK;
}

int nbs_interactions_total(nbs_ctx* c, int64_t* total, int reset) {
  if (!c || !total) return NBS_ERR_ARG;
  if (!c->has_state) return fail(c, NBS_ERR_STATE, "no particle state set");
  const int nw = std::max(c->sys.n_waves, 1);
  std::vector<unsigned long long> h(nw);
  HIPCHK(c, hipMemcpyAsync(h.data(), c->inter_total, nw * sizeof(h[0]),
                           hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  unsigned long long t = 0;
  for (unsigned long long v : h) t += v;
  *total = (int64_t)t;
  if (reset) launch_fill_u64(c->inter_total, 0, nw, c->stream);
  return NBS_OK;
}

int nbs_time(nbs_ctx* c, double* t, int cap) {
  if (!c || !t) return NBS_ERR_ARG;
  if (!c->has_state) return fail(c, NBS_ERR_STATE, "no particle state set");
  const int k = std::min(cap, c->sys.nsys);
  HIPCHK(c, hipMemcpyAsync(t, c->sys.time, k * sizeof(double), hipMemcpyDeviceToHost,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return NBS_OK;
}

int nbs_tree_info_get(nbs_ctx* c, nbs_tree_info* info) {
  if (!c || !info) return NBS_ERR_ARG;
  if (!c->tree_valid) return fail(c, NBS_ERR_STATE, "no tree built");
  int s[2];
  const int nb = emit_blocks(c->n_entries);
  std::vector<int> bs(2 * (size_t)nb);
  HIPCHK(c, hipMemcpyAsync(s, c->stats, sizeof(s), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(bs.data(), c->bb.blkstats, bs.size() * sizeof(int),
                           hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  int max_depth = 0, buckets = 0;
  for (int k = 0; k < nb; k++) {
    max_depth = std::max(max_depth, bs[2 * k]);
    buckets += bs[2 * k + 1];
  }
  info->n_particles = c->n;
  info->n_nodes = s[0];
  info->n_internal = s[1];
  info->max_depth = max_depth;
  info->n_buckets = buckets;
  info->node_capacity = c->node_cap;
  return NBS_OK;
}

int nbs_tree_dump(nbs_ctx* c, int32_t* depth, int32_t* leaf, int64_t* first,
                  int64_t* count, double* side, double* mass, double* com,
                  uint64_t* key_hi, uint64_t* key_lo, int32_t* order) {
  nbs_tree_info info;
  int r = nbs_tree_info_get(c, &info);
  if (r) return r;
  const int T = (int)info.n_nodes, n = c->n, ne = c->n_entries;
  std::vector<WalkNode> nd(T);
  std::vector<int> dep(T), lpos(T + 1), offs(ne + 1);
  std::vector<unsigned> id(n);
  std::vector<unsigned long long> khi(ne), klo(ne);
  hipStream_t s = c->stream;
  HIPCHK(c, hipMemcpyAsync(nd.data(), c->nodes, T * sizeof(WalkNode), hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipMemcpyAsync(id.data(), c->st[c->cur].id, n * sizeof(unsigned),
                           hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipMemcpyAsync(khi.data(), c->bb.khi, ne * sizeof(unsigned long long),
                           hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipMemcpyAsync(klo.data(), c->bb.klo, ne * sizeof(unsigned long long),
                           hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipMemcpyAsync(offs.data(), c->bb.offsets, (ne + 1) * sizeof(int),
                           hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipStreamSynchronize(s));
  // leaves before each node in pre-order (a leaf's position in the DFS leaf
  // order; with the threaded root a particle can own several leaves),
  // depth = the cells enclosing a node in pre-order
  lpos[0] = 0;
  for (int k = 0; k < T; k++) lpos[k + 1] = lpos[k] + (nd[k].leaf >= 0 ? 1 : 0);
  {
    std::vector<int> ends;
    for (int k = 0; k < T; k++) {
      while (!ends.empty() && ends.back() <= k) ends.pop_back();
      dep[k] = (int)ends.size();
      if (nd[k].leaf < 0) ends.push_back(std::min(std::max(nd[k].skip, k + 1), T));
    }
  }
  for (int k = 0; k < T; k++) {
    const int d = dep[k];
    const bool is_leaf = nd[k].leaf >= 0;
    const int e = is_leaf ? k + 1 : std::min(std::max(nd[k].skip, k + 1), T);
    if (depth) depth[k] = d;
    if (leaf) leaf[k] = is_leaf ? (int32_t)id[nd[k].leaf] : nd[k].leaf;
    if (first) first[k] = lpos[k];
    if (count) count[k] = lpos[e] - lpos[k];
    if (side) side[k] = nd[k].side;
    if (mass) mass[k] = nd[k].m;
    if (com) {
      com[3 * k] = nd[k].cx;
      com[3 * k + 1] = nd[k].cy;
      com[3 * k + 2] = nd[k].cz;
    }
    if (key_hi || key_lo) {
      // the sorted key whose nodes include this one: its path's prefix
      const int j = (int)(std::upper_bound(offs.begin(), offs.end(), k) - offs.begin()) - 1;
      const int dd = std::min(d, kLevels);
      const unsigned long long h = khi[std::max(0, std::min(j, ne - 1))];
      const unsigned long long l = klo[std::max(0, std::min(j, ne - 1))];
      unsigned long long mh, ml;
      if (dd <= 21) {
        mh = dd == 0 ? 0 : (h & (~0ull << (63 - 3 * dd)));
        ml = 0;
      } else {
        mh = h;
        ml = l & (~0ull << (63 - 3 * (dd - 21)));
      }
      if (key_hi) key_hi[k] = mh;
      if (key_lo) key_lo[k] = ml;
    }
    if (order && is_leaf) order[lpos[k]] = (int32_t)id[nd[k].leaf];
  }
  return NBS_OK;
}

int nbs_keys(nbs_ctx* c, uint64_t* key_hi, uint64_t* key_lo) {
  if (!c) return NBS_ERR_ARG;
  {
    const int r = check_threaded(c, "per-particle keys (nbs_keys)");
    if (r) return r;
  }
  if (!c->tree_valid) return fail(c, NBS_ERR_STATE, "no tree built");
  const int n = c->n;
  std::vector<unsigned> id(n);
  std::vector<unsigned long long> khi(n), klo(n);
  HIPCHK(c, hipMemcpyAsync(id.data(), c->st[c->cur].id, n * sizeof(unsigned),
                           hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(khi.data(), c->bb.khi, n * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(klo.data(), c->bb.klo, n * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int i = 0; i < n; i++) {
    if (key_hi) key_hi[id[i]] = khi[i];
    if (key_lo) key_lo[id[i]] = klo[i];
  }
  return NBS_OK;
}

int nbs_roots(nbs_ctx* c, double* center, double* side, int cap) {
  if (!c) return NBS_ERR_ARG;
  if (!c->tree_valid) return fail(c, NBS_ERR_STATE, "no tree built");
  const int k = std::min(cap, c->sys.nsys);
  std::vector<double> r(4 * (size_t)c->sys.nsys);
  HIPCHK(c, hipMemcpyAsync(r.data(), c->sys.root, r.size() * sizeof(double),
                           hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int s = 0; s < k; s++) {
    if (center)
      for (int d = 0; d < 3; d++) center[3 * s + d] = r[4 * s + d];
    if (side) side[s] = r[4 * s + 3];
  }
  return NBS_OK;
}

int nbs_profile_enable(nbs_ctx* c, int on) {
  if (!c) return NBS_ERR_ARG;
  c->prof = on != 0;
  return NBS_OK;
}

int nbs_profile_read(nbs_ctx* c, double* ms4, int reset) {
  if (!c || !ms4) return NBS_ERR_ARG;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int ph = 0; ph < 3; ph++) {
    for (auto& p : c->ev[ph]) {
      float ms = 0;
      if (hipEventElapsedTime(&ms, p.first, p.second) == hipSuccess) c->prof_ms[ph] += ms;
      (void)hipEventDestroy(p.first);
      (void)hipEventDestroy(p.second);
      if (ph == kPhaseWalk) c->prof_ms[3] += 1;
    }
    c->ev[ph].clear();
  }
  for (int k = 0; k < 4; k++) ms4[k] = c->prof_ms[k];
  if (reset)
    for (int k = 0; k < 4; k++) c->prof_ms[k] = 0;
  return NBS_OK;
}

void* nbs_stream(nbs_ctx* c) { return c ? (void*)c->stream : nullptr; }

}  // extern "C"

// ---------------------------------------------------------------- adjoint --
namespace {

void adj_free(nbs_ctx* c) {
  free_all(c->adj_bufs);
  c->ck_x.clear();
  c->ck_v.clear();
  c->adj_steps = c->adj_every = c->adj_n = 0;
  c->node_g = nullptr;
  c->node_g_cap = 0;
  c->mac_nodes = nullptr;
  c->mac_nodes_cap = 0;
  c->mac_x = nullptr;
  c->base_x = nullptr;
  c->adj_record = false;
}

// Current state (slot order) -> caller-order [3][n] positions / velocities.
void save_state(nbs_ctx* c, double* x3, double* v3) {
  State& st = c->st[c->cur];
  const int b = (c->n + 255) / 256;
  if (x3) scatter3_kernel<<<b, 256, 0, c->stream>>>(st.x, st.y, st.z, st.id, c->n, x3);
  if (v3) scatter3_kernel<<<b, 256, 0, c->stream>>>(st.vx, st.vy, st.vz, st.id, c->n, v3);
}

// Caller-order arrays -> state buffer 0 in caller order (the next build
// sorts it); invalidates the bbox and the tree.
void load_state(nbs_ctx* c, const double* x3, const double* v3, const double* m) {
  load_soa_kernel<<<(c->n + 255) / 256, 256, 0, c->stream>>>(x3, v3, m, c->n, c->st[0]);
  c->cur = 0;
  c->bbox_valid = false;
  c->tree_valid = false;
}

int ensure_node_adjoints(nbs_ctx* c) {
  if (c->node_g && c->node_g_cap >= c->node_cap) return NBS_OK;
  hipError_t e = hipSuccess;
  c->node_g = dalloc<double>(c->adj_bufs, 4 * (size_t)c->node_cap, &e);
  if (e != hipSuccess) return hip_fail(c, e, "hipMalloc(node adjoints)");
  c->node_g_cap = c->node_cap;
  return NBS_OK;
}

// One build of the current state with an overflow check.
int build_checked(nbs_ctx* c) {
  for (int attempt = 0; attempt < 6; attempt++) {
    reset_flags(c);
    enqueue_build(c);
    int flags = 0, done = 0;
    int r = read_flags(c, &flags, &done);
    if (r) return r;
    if (!flags) return NBS_OK;
    r = fatal_flags(c, flags);
    if (r) return r;
    // state was permuted but not changed: rebuild with room to spare
    c->bbox_valid = false;
    r = alloc_nodes(c, c->node_cap * 2);
    if (r) return r;
  }
  return fail(c, NBS_ERR_OVERFLOW, "node capacity overflow after retries");
}

// Reverse step at positions x3 (caller order): g_x += dF/dx^T u,
// g_m += dF/dm^T u, with u = dt * (g_v + dt g_x) and g_v updated.  checked:
// the build's flags are read (and the node arrays grown) before the walk;
// otherwise the caller reads them after a stretch of steps.
int adj_reverse_step(nbs_ctx* c, const double* x3, bool checked) {
  const int n = c->n;
  hipStream_t s = c->stream;
  launch_adj_kick(c->g_v, c->g_x, c->g_u, c->adj_dt, n, s);
  load_state(c, x3, nullptr, c->adj_m);
  int r = checked ? build_checked(c) : NBS_OK;
  if (!checked) enqueue_build(c);
  if (r) return r;
  r = ensure_node_adjoints(c);
  if (r) return r;
  State& st = c->st[c->cur];
  gather3_kernel<<<(n + 255) / 256, 256, 0, s>>>(c->g_u, st.id, n, c->s_u, c->s_u + n,
                                                 c->s_u + 2 * (size_t)n);
  launch_fill_u64((unsigned long long*)c->node_g, 0, 4 * (size_t)c->node_cap, s);
  AdjArgs a{};
  a.nodes = c->nodes;
  a.node_start = c->sys.node_start;
  a.sys_start = c->sys.start;
  a.wave_start = c->sys.wave_start;
  a.nsys = c->sys.nsys;
  a.n_waves = c->sys.n_waves;
  a.x = st.x;
  a.y = st.y;
  a.z = st.z;
  a.ux = c->s_u;
  a.uy = c->s_u + n;
  a.uz = c->s_u + 2 * (size_t)n;
  a.gx = c->s_gx;
  a.gy = c->s_gx + n;
  a.gz = c->s_gx + 2 * (size_t)n;
  a.gnode = c->node_g;
  a.flags = c->flags;
  a.theta = c->p.theta;
  if (c->walk_split > 1) {
    a.split = c->walk_split;
    a.part = c->walk_part;
    a.n = n;
  }
  launch_adj_walk(a, c->cfg.force_precision == NBS_FORCE_F32, c->p.damping, s);
  launch_adj_propagate(a, c->bb.offsets, c->aux.parent, st.m, c->s_gm, n, s);
  launch_scatter_comp(c->s_gx, c->g_x, st.id, n, 3, 1, s);
  launch_scatter_comp(c->s_gm, c->g_m, st.id, n, 1, 1, s);
  return NBS_OK;
}

void to_soa(const double* aos, std::vector<double>& soa, size_t n) {
  soa.resize(3 * n);
  for (size_t i = 0; i < n; i++)
    for (int k = 0; k < 3; k++) soa[k * n + i] = aos[3 * i + k];
}
void to_aos(const std::vector<double>& soa, double* aos, size_t n) {
  for (size_t i = 0; i < n; i++)
    for (int k = 0; k < 3; k++) aos[3 * i + k] = soa[k * n + i];
}

// Steps t0 .. t1 - 1 of an adjoint pass from the current state, save(t)
// enqueued before step t: enqueued together with one read of the flags at
// the end instead of one per step.  A node overflow at step t leaves the
// state as it was before step t (every later kernel of the stretch is a
// no-op; the builds only permute it), so the node arrays grow and the
// stretch resumes at t, saves included.
template <class Save>
int adj_run_steps(nbs_ctx* c, int t0, int t1, Save save) {
  int r = ensure_nodes(c);
  if (r) return r;
  for (int attempt = 0; attempt < 8; attempt++) {
    if (t0 >= t1) return NBS_OK;
    reset_flags(c);
    for (int t = t0; t < t1; t++) {
      save(t);
      enqueue_build(c);
      enqueue_step_tail(c);
    }
    int flags = 0, done = 0;
    r = read_flags(c, &flags, &done);
    if (r) return r;
    if (!flags) {
      c->tree_valid = true;
      c->acc_valid = true;
      return NBS_OK;
    }
    c->bbox_valid = false;
    c->tree_valid = false;
    r = fatal_flags(c, flags);
    if (r) return r;
    t0 += done;
    r = alloc_nodes(c, c->node_cap * 2);
    if (r) return r;
  }
  return fail(c, NBS_ERR_OVERFLOW, "node capacity overflow after retries");
}

// dL/dx_0, dL/dv_0, dL/dm from dL/dx_T, dL/dv_T: host [n][3] / [n] arrays,
// or (device) device arrays of this library's runtime.
int adjoint_backward(nbs_ctx* c, const double* gx_T, const double* gv_T, double* gx0,
                     double* gv0, double* gm, bool device) {
  if (!c) return NBS_ERR_ARG;
  {
    const int r = check_threaded(c, "the adjoint");
    if (r) return r;
  }
  if (!c->adj_steps || c->adj_n != c->n)
    return fail(c, NBS_ERR_STATE, "run nbs_adjoint_forward first");
  HIPCHK(c, hipSetDevice(c->device));
  const size_t n = c->n;
  hipStream_t s = c->stream;
  const int nb = (int)((n + 255) / 256);
  if (device) {
    for (int k = 0; k < 2; k++) {
      const double* g = k ? gv_T : gx_T;
      double* dst = k ? c->g_v : c->g_x;
      if (g)
        aos_soa_kernel<<<nb, 256, 0, s>>>(g, dst, (int)n, 1);
      else
        launch_fill_u64((unsigned long long*)dst, 0, 3 * n, s);
    }
  } else {
    std::vector<double> h(3 * n, 0.0);
    if (gx_T) to_soa(gx_T, h, n); else std::fill(h.begin(), h.end(), 0.0);
    HIPCHK(c, hipMemcpyAsync(c->g_x, h.data(), 3 * n * 8, hipMemcpyHostToDevice, s));
    HIPCHK(c, hipStreamSynchronize(s));
    if (gv_T) to_soa(gv_T, h, n); else std::fill(h.begin(), h.end(), 0.0);
    HIPCHK(c, hipMemcpyAsync(c->g_v, h.data(), 3 * n * 8, hipMemcpyHostToDevice, s));
    HIPCHK(c, hipStreamSynchronize(s));
  }
  launch_fill_u64((unsigned long long*)c->g_m, 0, n, s);
  const int T = c->adj_steps, K = c->adj_every;
  const int nck = (T + K - 1) / K;
  for (int seg = nck - 1; seg >= 0; seg--) {
    const int t0 = seg * K, t1 = std::min(T, t0 + K);
    // recompute the segment's positions from its checkpoint (bit-exact:
    // the build sorts canonically, so every step repeats the forward)
    load_state(c, c->ck_x[seg], c->ck_v[seg], c->adj_m);
    int r = adj_run_steps(c, t0, t1 - 1, [&](int t) {
      save_state(c, c->seg_x + 3 * n * (t - t0), nullptr);
    });
    if (r) return r;
    save_state(c, c->seg_x + 3 * n * (t1 - 1 - t0), nullptr);
    // the reverse steps, their flags read once: the trees are the
    // forward's, which the node arrays already hold, so an overflow is not
    // expected; if one comes, the segment is redone from saved adjoints
    // with a checked build per step
    HIPCHK(c, hipMemcpyAsync(c->b_x, c->g_x, 3 * n * 8, hipMemcpyDeviceToDevice, s));
    HIPCHK(c, hipMemcpyAsync(c->b_v, c->g_v, 3 * n * 8, hipMemcpyDeviceToDevice, s));
    HIPCHK(c, hipMemcpyAsync(c->b_m, c->g_m, n * 8, hipMemcpyDeviceToDevice, s));
    reset_flags(c);
    for (int t = t1 - 1; t >= t0; t--) {
      r = adj_reverse_step(c, c->seg_x + 3 * n * (t - t0), false);
      if (r) return r;
    }
    int flags = 0, done = 0;
    r = read_flags(c, &flags, &done);
    if (r) return r;
    if (flags) {
      r = fatal_flags(c, flags);
      if (r) return r;
      HIPCHK(c, hipMemcpyAsync(c->g_x, c->b_x, 3 * n * 8, hipMemcpyDeviceToDevice, s));
      HIPCHK(c, hipMemcpyAsync(c->g_v, c->b_v, 3 * n * 8, hipMemcpyDeviceToDevice, s));
      HIPCHK(c, hipMemcpyAsync(c->g_m, c->b_m, n * 8, hipMemcpyDeviceToDevice, s));
      for (int t = t1 - 1; t >= t0; t--) {
        r = adj_reverse_step(c, c->seg_x + 3 * n * (t - t0), true);
        if (r) return r;
      }
    }
  }
  if (device) {
    if (gx0) aos_soa_kernel<<<nb, 256, 0, s>>>(c->g_x, gx0, (int)n, 0);
    if (gv0) aos_soa_kernel<<<nb, 256, 0, s>>>(c->g_v, gv0, (int)n, 0);
    if (gm) HIPCHK(c, hipMemcpyAsync(gm, c->g_m, n * 8, hipMemcpyDeviceToDevice, s));
    HIPCHK(c, hipGetLastError());
    // leave the context at the forward's final state
    load_state(c, c->fin_x, c->fin_v, c->adj_m);
    HIPCHK(c, hipStreamSynchronize(s));
    return NBS_OK;
  }
  std::vector<double> hx(3 * n), hv(3 * n), hm(n);
  HIPCHK(c, hipMemcpyAsync(hx.data(), c->g_x, 3 * n * 8, hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipMemcpyAsync(hv.data(), c->g_v, 3 * n * 8, hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipMemcpyAsync(hm.data(), c->g_m, n * 8, hipMemcpyDeviceToHost, s));
  // leave the context at the forward's final state
  load_state(c, c->fin_x, c->fin_v, c->adj_m);
  HIPCHK(c, hipStreamSynchronize(s));
  if (gx0) to_aos(hx, gx0, n);
  if (gv0) to_aos(hv, gv0, n);
  if (gm) std::memcpy(gm, hm.data(), n * 8);
  return NBS_OK;
}

}  // namespace

extern "C" {

int nbs_adjoint_forward(nbs_ctx* c, int nsteps, int every, int record_all) {
  if (c) c->want_parent = true;  // the backward pass propagates along parents
  if (!c) return NBS_ERR_ARG;
  {
    const int r = check_threaded(c, "the adjoint");
    if (r) return r;
  }
  if (!c->has_state) return fail(c, NBS_ERR_STATE, "no particle state set");
  if (nsteps < 1 || every < 1) return fail(c, NBS_ERR_ARG, "nsteps and every must be >= 1");
  if (c->p.adaptive) return fail(c, NBS_ERR_ARG, "the adjoint needs a fixed dt");
  if (c->sys.nsys != 1) return fail(c, NBS_ERR_ARG, "the adjoint runs one system");
  HIPCHK(c, hipSetDevice(c->device));
  adj_free(c);
  const int n = c->n;
  const int nck = (nsteps + every - 1) / every;
  hipError_t e = hipSuccess;
  auto& o = c->adj_bufs;
  for (int k = 0; k < nck; k++) {
    c->ck_x.push_back(dalloc<double>(o, 3 * (size_t)n, &e));
    c->ck_v.push_back(dalloc<double>(o, 3 * (size_t)n, &e));
  }
  c->adj_m = dalloc<double>(o, n, &e);
  c->seg_x = dalloc<double>(o, 3 * (size_t)n * every, &e);
  c->fin_x = dalloc<double>(o, 3 * (size_t)n, &e);
  c->fin_v = dalloc<double>(o, 3 * (size_t)n, &e);
  c->g_x = dalloc<double>(o, 3 * (size_t)n, &e);
  c->g_v = dalloc<double>(o, 3 * (size_t)n, &e);
  c->g_u = dalloc<double>(o, 3 * (size_t)n, &e);
  c->g_m = dalloc<double>(o, n, &e);
  c->b_x = dalloc<double>(o, 3 * (size_t)n, &e);
  c->b_v = dalloc<double>(o, 3 * (size_t)n, &e);
  c->b_m = dalloc<double>(o, n, &e);
  c->s_u = dalloc<double>(o, 3 * (size_t)n, &e);
  c->s_gx = dalloc<double>(o, 3 * (size_t)n, &e);
  c->s_gm = dalloc<double>(o, n, &e);
  c->p_x = dalloc<double>(o, 3 * (size_t)n, &e);
  c->p_v = dalloc<double>(o, 3 * (size_t)n, &e);
  c->p_m = dalloc<double>(o, n, &e);
  if (record_all) c->base_x = dalloc<double>(o, 3 * (size_t)n * nsteps, &e);
  if (e != hipSuccess) {
    adj_free(c);
    return hip_fail(c, e, "hipMalloc(adjoint buffers)");
  }
  {
    State& st = c->st[c->cur];
    launch_scatter_comp(st.m, c->adj_m, st.id, n, 1, 0, c->stream);
  }
  {
    const int r = adj_run_steps(c, 0, nsteps, [&](int t) {
      if (t % every == 0) save_state(c, c->ck_x[t / every], c->ck_v[t / every]);
      if (record_all) save_state(c, c->base_x + 3 * (size_t)n * t, nullptr);
    });
    if (r) return r;
  }
  save_state(c, c->fin_x, c->fin_v);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->adj_steps = nsteps;
  c->adj_every = every;
  c->adj_n = n;
  c->adj_record = record_all != 0;
  c->adj_dt = c->p.dt;
  return NBS_OK;
}

int nbs_adjoint_backward(nbs_ctx* c, const double* gx_T, const double* gv_T,
                         double* gx0, double* gv0, double* gm) {
  return adjoint_backward(c, gx_T, gv_T, gx0, gv0, gm, false);
}

int nbs_adjoint_backward_device(nbs_ctx* c, const double* gx_T, const double* gv_T,
                                double* gx0, double* gv0, double* gm) {
  if (!c) return NBS_ERR_ARG;
  if (c->n > 0) {
    int r = check_device_arrays(c, {gx_T, gv_T, gx0, gv0, gm});
    if (r) return r;
  }
  return adjoint_backward(c, gx_T, gv_T, gx0, gv0, gm, true);
}

int nbs_frozen_forward(nbs_ctx* c, const double* pos0, const double* vel0,
                       const double* mass, double* posT, double* velT) {
  if (!c || !pos0 || !mass) return NBS_ERR_ARG;
  {
    const int r = check_threaded(c, "the adjoint");
    if (r) return r;
  }
  if (!c->adj_steps || !c->adj_record || c->adj_n != c->n)
    return fail(c, NBS_ERR_STATE, "run nbs_adjoint_forward with record_all first");
  if (c->cfg.force_precision != NBS_FORCE_F64)
    return fail(c, NBS_ERR_ARG, "frozen replays are FP64 only");
  HIPCHK(c, hipSetDevice(c->device));
  const size_t n = c->n;
  hipStream_t s = c->stream;
  hipError_t e = hipSuccess;
  if (!c->mac_x) c->mac_x = dalloc<double>(c->adj_bufs, 3 * n, &e);
  if (e != hipSuccess) return hip_fail(c, e, "hipMalloc(frozen)");
  std::vector<double> h;
  to_soa(pos0, h, n);
  HIPCHK(c, hipMemcpyAsync(c->p_x, h.data(), 3 * n * 8, hipMemcpyHostToDevice, s));
  HIPCHK(c, hipStreamSynchronize(s));
  if (vel0) to_soa(vel0, h, n); else h.assign(3 * n, 0.0);
  HIPCHK(c, hipMemcpyAsync(c->p_v, h.data(), 3 * n * 8, hipMemcpyHostToDevice, s));
  HIPCHK(c, hipMemcpyAsync(c->p_m, mass, n * 8, hipMemcpyHostToDevice, s));
  HIPCHK(c, hipStreamSynchronize(s));
  for (int t = 0; t < c->adj_steps; t++) {
    load_state(c, c->base_x + 3 * n * t, nullptr, c->adj_m);
    int r = build_checked(c);
    if (r) return r;
    if (!c->mac_nodes || c->mac_nodes_cap < c->node_cap) {
      c->mac_nodes = dalloc<WalkNode>(c->adj_bufs, c->node_cap, &e);
      if (e != hipSuccess) return hip_fail(c, e, "hipMalloc(frozen nodes)");
      c->mac_nodes_cap = c->node_cap;
    }
    State& st = c->st[c->cur];
    const int b = (int)((n + 255) / 256);
    HIPCHK(c, hipMemcpyAsync(c->mac_nodes, c->nodes, (size_t)c->node_cap * sizeof(WalkNode),
                             hipMemcpyDeviceToDevice, s));
    HIPCHK(c, hipMemcpyAsync(c->mac_x, st.x, n * 8, hipMemcpyDeviceToDevice, s));
    HIPCHK(c, hipMemcpyAsync(c->mac_x + n, st.y, n * 8, hipMemcpyDeviceToDevice, s));
    HIPCHK(c, hipMemcpyAsync(c->mac_x + 2 * n, st.z, n * 8, hipMemcpyDeviceToDevice, s));
    // perturbed state into slot order, COM of the same topology
    gather3_kernel<<<b, 256, 0, s>>>(c->p_x, st.id, (int)n, st.x, st.y, st.z);
    gather3_kernel<<<b, 256, 0, s>>>(c->p_v, st.id, (int)n, st.vx, st.vy, st.vz);
    launch_gather_comp(c->p_m, st.m, st.id, (int)n, 1, s);
    launch_leaf_refresh(c->bb, st, c->nodes, c->nodes32, (int)n, s);
    launch_bottom_up(c->bb, c->nodes, c->aux, (int)n, c->sys.nsys, s);
    WalkArgs a{};
    a.nodes = c->nodes;
    a.node_start = c->sys.node_start;
    a.sys_start = c->sys.start;
    a.wave_start = c->sys.wave_start;
    a.wave2_start = c->sys.wave2_start;
    a.nsys = c->sys.nsys;
    a.n_waves = c->sys.n_waves;
    a.n_waves2 = c->sys.n_waves2;
    a.x = st.x;
    a.y = st.y;
    a.z = st.z;
    a.vx = st.vx;
    a.vy = st.vy;
    a.vz = st.vz;
    a.ax = c->ax;
    a.ay = c->ay;
    a.az = c->az;
    a.interactions = c->inter;
    a.flags = c->flags;
    a.flags_out = c->flags;
    a.theta = c->p.theta;
    a.dt = c->adj_dt;
    a.bbox = c->sys.bbox;
    a.sys_dt = c->sys.dt;
    a.sys_time = c->sys.time;
    a.steps_done = c->steps_done;
    a.inter_total = c->inter_total;
    a.fused_update = 1;
    a.count_mode = kCountLanes;
    a.mac_nodes = c->mac_nodes;
    a.mac_x = c->mac_x;
    a.mac_y = c->mac_x + n;
    a.mac_z = c->mac_x + 2 * n;
    if (c->walk_split > 1) {  // the pieces of the recorded forward's walks
      a.split = c->walk_split;
      a.part = c->walk_part;
      a.cpart = c->walk_cpart;
      a.n = (int)n;
    }
    launch_walk(a, 0, c->p.damping, c->walk_block, c->walk_pack, s);
    scatter3_kernel<<<b, 256, 0, s>>>(st.x, st.y, st.z, st.id, (int)n, c->p_x);
    scatter3_kernel<<<b, 256, 0, s>>>(st.vx, st.vy, st.vz, st.id, (int)n, c->p_v);
  }
  std::vector<double> hx(3 * n), hv(3 * n);
  HIPCHK(c, hipMemcpyAsync(hx.data(), c->p_x, 3 * n * 8, hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipMemcpyAsync(hv.data(), c->p_v, 3 * n * 8, hipMemcpyDeviceToHost, s));
  load_state(c, c->fin_x, c->fin_v, c->adj_m);
  HIPCHK(c, hipStreamSynchronize(s));
  if (posT) to_aos(hx, posT, n);
  if (velT) to_aos(hv, velT, n);
  return NBS_OK;
}

}  // extern "C"
