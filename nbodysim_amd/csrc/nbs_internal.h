// nbs_internal.h -- device data layout and kernel launchers of libnbs.
//
// Data layout in HBM (per context, FP64 state):
//   particles, structure-of-arrays, in octree (Morton / DFS-leaf) order of
//   the last build, double-buffered so a build can permute into the other
//   half:  x y z vx vy vz m (f64), id (u32 caller index), ax ay az (f64).
//   nodes: one 64-byte WalkNode per tree entry (internal cells and leaves)
//   in DFS pre-order with children 0..7 -- exactly the order of the
//   reference octree's recursive traversal (src/octree.cpp:203-313) -- so a
//   subtree is a contiguous run and `skip` jumps over it.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace nbs {

constexpr double kG = 6.67384e-11;  // src/main.cpp:322
constexpr int kLevels = 42;         // key depth: 21 levels per u64 half
constexpr int kDepthBins = kLevels + 1;  // internal-cell depths 0..42
constexpr int kWave = 64;

// One tree entry, one 64-byte line.  The FP64 walk reads it with one
// wave-uniform scalar load.
struct alignas(64) WalkNode {
  double cx, cy, cz;  // centre of mass (src/octree.cpp:252-278)
  int skip;           // pre-order index after this subtree
  int leaf;           // >= 0 leaf (particle slot); -1 internal; -2 bucket
  double gm;          // G * m
  double side;        // cell side: root side / 2^depth
  float gm32;         // G * m rounded to f32
  float r2open32;     // (side / theta)^2 in f32: open if d^2 < this
  double m;           // mass (src/octree.cpp:203-221)
};
static_assert(sizeof(WalkNode) == 64, "WalkNode must be one 64 B line");

// What the FP32 walk reads of a tree entry, 32 bytes: half a line, so twice
// the entries per cache line and per L2.  The centre of mass and G * m are
// the WalkNode's, bit for bit.  `link`: a leaf has bit 31 set and its
// particle slot in bits 30..0; a cell has its depth in bits 30..25 (a
// depth-42 bucket is opened by every lane) and its skip in bits 24..0, so
// the copy exists only while the node capacity stays below 2^25.  The
// opening radius is not stored: (side / theta)^2 in f32 is the root's
// value times 4^-depth, exactly, as long as every such value is a normal
// float, which the walk checks per system.
struct alignas(32) WalkNode32 {
  double cx, cy, cz;
  unsigned link;
  float gm32;
};
static_assert(sizeof(WalkNode32) == 32, "WalkNode32 must be half a 64 B line");
constexpr unsigned kLink32Leaf = 0x80000000u;
constexpr int kLink32DepthShift = 25;
constexpr int kLink32MaxNodes = 1 << 25;  // capacity + 1 must stay below
__host__ __device__ inline unsigned link32_cell(int depth, int skip) {
  return ((unsigned)depth << kLink32DepthShift) | (unsigned)skip;
}

// Per-build tree bookkeeping beside the WalkNode array: each node's parent
// (-1 for a root), written only when the adjoint needs it (null otherwise).
// A node's depth, first particle and particle count follow from the
// pre-order array and its skip links (nbs_tree_dump derives them).
struct TreeAux {
  int* parent;  // [cap] or null
  WalkNode32* walk32;  // [cap + 1]: the FP32 walk's copy, or null
};

// Device-side per-system data (batched mode; one system otherwise).
struct Systems {
  int nsys;
  int n_waves;       // host copy of wave_start[nsys]
  int* start;        // [nsys+1] particle slots
  int* wave_start;   // [nsys+1] walk waves (64 particles each, per system)
  int n_waves2;      // host copy of wave2_start[nsys]
  int* wave2_start;  // [nsys+1] wavefronts of 128 particles (2 per lane)
  int* node_start;   // [nsys+1] pre-order node offsets (written by emit)
  double* bbox;      // [6][n_waves] per-wave min xyz, max xyz
  double* root;      // [nsys*4] centre xyz, side
  double* dt;        // [nsys] step length used by the last update
  double* time;      // [nsys] simulated time
  double* amax;      // [n_waves] adaptive: per-wave max |a|
  double* vmax;      // [n_waves] adaptive: per-wave max |v|
};

struct State {
  double *x, *y, *z, *vx, *vy, *vz, *m;
  unsigned* id;
};

// Flags word bits (device int).
enum : int { kFlagNodeOverflow = 1, kFlagStack = 4, kFlagPairs = 8 };

struct StepParams {
  double theta;
  double dt;
  int damping;
  int root_mode;
  double root_center[3];
  double root_size;
  int adaptive;
  double min_adaptive_dt, max_pos_change, max_vel_change;
};

// ---------------------------------------------------------------- scan --
// Exclusive prefix sum of n ints into out (n+1 entries; out[n] = total).
// tmp needs scan_tmp_ints(n) ints.
size_t scan_tmp_ints(size_t n);
void exclusive_scan(const int* in, int* out, int n, int* tmp,
                    hipStream_t s);

// ----------------------------------------------------------- radix sort --
// Stable LSD sort of (u64 key, u32 value) pairs over bits [0, nbits).
// Ping-pongs between (k0,v0) and (k1,v1); returns 0 if the result ends in
// (k0, v0), 1 if in (k1, v1).  tmp needs radix_tmp_ints(n) ints.
size_t radix_tmp_ints(size_t n);
int radix_sort_pairs(unsigned long long* k0, unsigned* v0,
                     unsigned long long* k1, unsigned* v1, int n, int nbits,
                     int* tmp, hipStream_t s);

// ----------------------------------------------------------- tree build --
struct BuildBuffers {
  unsigned long long *prim0, *prim1;  // sort keys
  unsigned *val0, *val1;              // sort values (source slot)
  unsigned long long *khi_u, *klo_u;  // keys in unsorted (current) order
  unsigned long long *khi, *klo;      // keys in sorted order
  int* counts;                        // [n]
  int* offsets;                       // [n+1] node offsets per slot
  int* dhist;                         // [depth_hist_ints(n)] per-block depths
  int* doff;                          // same, exclusive-scanned
  int* bucket;                        // [node cap] internal cells by depth
  int* close_at;                      // [n] last slot of each cell, by opening rank
  int* sort_tmp;
  int* scan_tmp;
  int* runs;                          // runs of equal primary keys: run_list_ints(cap)
  int runs_big;                       // index in runs of the list of runs over 256 entries
  int* flags;                         // device flags word
  int* stats;                         // [2]: nodes, internal cells
  int* blkstats;                      // [2 * emit_blocks(n)]: per emit block deepest leaf, buckets
};

void launch_bbox(const State& st, Systems& sys, hipStream_t s);
void launch_root(Systems& sys, const StepParams& p, const int* flags,
                 hipStream_t s);
// Keys of the particles in `st` (current order) into prim0/val0 and
// khi_u/klo_u; the primary (sort) key holds the system and the first
// prim_levels key levels.  Returns the number of significant primary-key
// bits.
int launch_keys(const State& st, Systems& sys, BuildBuffers& b, int prim_levels,
                hipStream_t s);
// Deepest primary key for nsys systems (system bits + 3 bits per level
// within 64 bits, at most 21 levels).
int max_prim_levels(int nsys);
// Gathers the full keys into sorted order and sorts runs of equal primary
// keys by (full key, caller index).  perm is updated in place; free_perm
// (the sort's other value array) is scratch.
// Ints of the run lists for up to cap keys (runs of 9-256 and of more than
// 256 equal primary keys); the second list starts at run_list_big(cap).
size_t run_list_ints(size_t cap);
int run_list_big(size_t cap);
void launch_fixup(BuildBuffers& b, const unsigned long long* prim,
                  unsigned* perm, unsigned* free_perm, const State& st, int n,
                  hipStream_t s);
void launch_permute(const State& src, State& dst, const unsigned* perm,
                    int n, hipStream_t s);
// Emits the pre-order node array (capacity `cap`), leaves filled from `st`.
size_t depth_hist_ints(size_t n);
// Blocks of the emission kernels (per-block tree statistics).
int emit_blocks(size_t n);
void launch_emit(BuildBuffers& b, Systems& sys, const StepParams& p,
                 const State& st, WalkNode* nodes, TreeAux& aux, int n,
                 int cap, hipStream_t s);
// Mass/COM of every internal cell, level by level from the deepest up
// (nsys systems: the top levels hold nsys * 8^d cells at most).
void launch_bottom_up(BuildBuffers& b, WalkNode* nodes, TreeAux& aux, int n,
                      int nsys, hipStream_t s);
// Leaves of the current topology take new particle data.
void launch_leaf_refresh(BuildBuffers& b, const State& st, WalkNode* nodes,
                         WalkNode32* walk32,
                         int n, hipStream_t s);

// -------------------------------------------------- threaded-root tree --
// The tree of the reference's threads > 1 path (tree_threaded.hip): the
// build over (particle, child) entries.
int member_blocks(int n);
void launch_member(const State& st, int n, const double* root, int* mask, int* cnt,
                   int* outside, int* occ_part, hipStream_t s);
void launch_entry_keys(const State& st, int n, const double* root, const int* mask,
                       const int* eoff, int e_real, const int* empty, int n_empty,
                       int prim_levels, BuildBuffers& b, int* epart, unsigned* eid,
                       hipStream_t s);
void launch_entry_gather(const unsigned* perm, const int* epart, const int* eoff,
                         const State& cur, int ne, double* ex, double* ey, double* ez,
                         double* em, int* sp, int* first, hipStream_t s);
void launch_entry_slots(const int* sp, const int* first, const int* pslot, const int* cnt,
                        const int* orank, int ne, int n, int* new_slot, unsigned* gperm,
                        hipStream_t s);
void launch_entry_patch(const int* offsets, const int* sp, const int* new_slot, int ne,
                        double theta, int n_empty, WalkNode* nodes, WalkNode32* walk32,
                        int* stats, const int* flags, hipStream_t s);

// ----------------------------------------------------------------- walk --
enum CountMode { kCountNone = 0, kCountWave = 1, kCountLanes = 2 };

struct WalkArgs {
  const WalkNode* nodes;
  const int* node_start;
  const int* sys_start;
  const int* wave_start;
  const int* wave2_start;  // per-system wavefront offsets, 2 particles/lane
  int nsys;
  int n_waves;
  int n_waves2;
  double *x, *y, *z, *vx, *vy, *vz;
  double *ax, *ay, *az;
  int* interactions;
  const int* flags;
  int* flags_out;
  double theta;
  double dt;
  double* bbox;  // next step's per-wave bbox partials (fused update)
  double* sys_dt;
  double* sys_time;
  int* steps_done;  // incremented once per completed walk
  unsigned long long* inter_total;  // running sum of interactions
  int fused_update;
  int groups;  // workgroup -> tile interleave over the XCDs (1 = plain)
  int asm_loop;  // FP32 walk: the hand-written loop where it applies
  // interaction counting: kCountLanes per particle (interactions[] and the
  // per-wave totals), kCountWave the per-wave totals only, kCountNone none
  int count_mode;
  const int* order;     // workgroup -> tile (longest first), or null
  unsigned* tile_cost;  // per tile: node visits of this walk added, or null
  // frozen decisions (FP64 only; null otherwise): base tree + positions
  const WalkNode* mac_nodes;
  const double *mac_x, *mac_y, *mac_z;
  // the FP32 hand-written loop's 32-byte records (null: 64-byte records)
  // and the roots ([4 s + 3] = side of system s) their opening radii follow
  const WalkNode32* nodes32;
  const double* root;
  // split > 1: each traversal walked as `split` pieces over equal ranges of
  // its system's nodes; part [split][3][n] partial accelerations (f32 or
  // f64, the force precision's) and cpart [split][n] partial counts, added
  // by the combine kernel; n particles in all systems
  int split;
  void* part;
  int* cpart;
  int n;
};

// Collision pairs: every (particle i, leaf j) that i's walk reaches -- the
// FP64 MAC -- with distance < range appends (min(i, j), max(i, j));
// pair_count counts every detection, including those beyond pair_cap.
struct DetectArgs {
  const WalkNode* nodes;
  const int* node_start;
  const int* sys_start;
  const int* wave_start;
  int nsys;
  int n_waves;
  const double *x, *y, *z;
  const double* root;  // per system: centre xyz, side
  int boxes;           // 1: every particle inside its root (bbox root)
  double theta, range;
  int2* pairs;
  int* pair_count;
  int pair_cap;
  const int* flags;
  int* flags_out;
};
void launch_detect(const DetectArgs& a, hipStream_t s);

// Launch shape of the walk: block = workgroup size (64, 128, 256, 512 or
// 1024 threads); a.groups = contiguous tile regions interleaved over the
// XCDs; pack = particles per lane (1, or 2 sharing one traversal).
void launch_walk(const WalkArgs& a, int f32, int damping, int block, int pack,
                 hipStream_t s);
// Workgroups (tiles) of a walk launch of this shape.
int walk_tiles(int n_waves, int n_waves2, int block, int pack);
// Longest-first tile order from the previous walk's tile costs (cleared).
void launch_tile_order(unsigned* tile_cost, int n_tiles, int groups, int* order,
                       hipStream_t s);
constexpr int kWalkBlocks[] = {64, 128, 256, 512, 1024};
constexpr int kWalkGroups[] = {1, 8, 16, 32, 64};
constexpr int kWalkPacks[] = {1, 2};
constexpr int kNumWalkBlocks = sizeof(kWalkBlocks) / sizeof(int);
constexpr int kNumWalkGroups = sizeof(kWalkGroups) / sizeof(int);
constexpr int kNumWalkPacks = sizeof(kWalkPacks) / sizeof(int);

// FP64 direct sum on sample slots; partial buffer of
// direct_tmp_doubles(m) doubles.
size_t direct_tmp_doubles(size_t m);
void launch_direct(const State& st, int n, const int* sample_slots, int m,
                   int damping, double* tmp, double* acc_out, hipStream_t s);
void launch_inverse_perm(const unsigned* id, int n, int* slot_of_id,
                         hipStream_t s);

// -------------------------------------------------------------- adjoint --
struct AdjArgs {
  const WalkNode* nodes;
  const int* node_start;
  const int* sys_start;
  const int* wave_start;
  int nsys;
  int n_waves;
  const double *x, *y, *z;       // slot order
  const double *ux, *uy, *uz;    // adjoint of the acceleration, slot order
  double *gx, *gy, *gz;          // out: position adjoint, slot order
  double* gnode;  // per node [4]: adjoint of the COM (x, y, z), of the mass
  const int* flags;
  double theta;
  // split > 1: each traversal as `split` pieces over equal node ranges (as
  // the forward walk), partial position adjoints part [split][3][n] added
  // in piece order into gx, gy, gz
  int split;
  double* part;
  int n;
};

void launch_adj_walk(const AdjArgs& a, int mac32, int damping, hipStream_t s);
void launch_adj_propagate(const AdjArgs& a, const int* offsets,
                          const int* parent, const double* m, double* gm,
                          int n, hipStream_t s);
// gv += dt gx; u = dt gv (caller order, 3 components of n each)
void launch_adj_kick(double* gv, const double* gx, double* u, double dt,
                     int n, hipStream_t s);
// dst[c][s] = src[c][id[s]] for c < ncomp (caller -> slot order)
void launch_gather_comp(const double* src, double* dst, const unsigned* id,
                        int n, int ncomp, hipStream_t s);
// dst[c][id[s]] (+)= src[c][s] (slot -> caller order)
void launch_scatter_comp(const double* src, double* dst, const unsigned* id,
                         int n, int ncomp, int add, hipStream_t s);

// ------------------------------------------------------------ integrate --
void launch_adaptive_dt(const State& st, const double* ax, const double* ay,
                        const double* az, int n, Systems& sys,
                        const StepParams& p, const int* flags, hipStream_t s);
void launch_update(State& st, const double* ax, const double* ay,
                   const double* az, int n, Systems& sys, const int* flags,
                   hipStream_t s);

// ------------------------------------------------------------- collide --
// Collision merging after a walk that recorded pairs (WalkArgs::pairs).
// Buffers: parent, mark, alive, id_alive [cap_n]; newpos, newid [cap_n+1];
// the compacted acceleration / interaction arrays [cap_n]; sort keys and
// values [2][2 * pair_cap]; info = {members, total, new starts[nsys+1]}.
struct CollideBuffers {
  int2* pairs;
  int* count;  // [2]: pairs detected, members listed
  int pair_cap;
  int *parent, *mark, *alive, *id_alive, *newpos, *newid;
  double *cax, *cay, *caz;
  int* cinter;
  unsigned long long *key0, *key1;
  unsigned *val0, *val1;
  int* sort_tmp;
  int* scan_tmp;
  int* info;
};
size_t collide_sort_tmp_ints(int pair_cap);
// Merges the groups of the recorded pairs, updates (kick-drift; merged
// particles with zero acceleration) and compacts `cur` into `nxt`,
// accelerations and interaction counts into cb.cax.. / cb.cinter.
void launch_collide(CollideBuffers& cb, State& cur, State& nxt, double* ax, double* ay,
                    double* az, const int* inter, int n, Systems& sys, double dt,
                    int* flags, hipStream_t s);

// --------------------------------------------------------------- image --
// write_image's settings (src/main.cpp:556): projection 0 front, 1 side,
// 2 top, 3 iso; color 0 bw, 1 heat.
struct ImageParams {
  unsigned width, height, projection, color;
  int adaptive, nonlinear;
  double scale, brightness;
};
// counts [w*h] u32, temp [w*h] doubles, partial [image_partials(w*h)]
// doubles, vmax one word; rgb [h][w][3] bytes, PNG row order (top row
// first).
size_t image_partials(size_t npix);
void launch_image(const State& st, int n, const ImageParams& p, unsigned* counts, double* temp,
                  double* partial, unsigned long long* vmax, unsigned char* rgb,
                  hipStream_t s);

// Stable compaction of `cur` into `nxt` by cb.alive (accelerations and
// interaction counts into cb.cax.. / cb.cinter), the system starts moved;
// advance_time adds each system's dt to its time.
void launch_compact(CollideBuffers& cb, State& cur, State& nxt, const double* ax,
                    const double* ay, const double* az, const int* inter, int n, Systems& sys,
                    int advance_time, int* flags, hipStream_t s);
// apply_minimum_size on the built tree: every largest cell of side <
// min_size becomes one particle, then `cur` is compacted into `nxt`.
void launch_min_size(CollideBuffers& cb, const WalkNode* nodes, int node_cap, double min_size,
                     State& cur, State& nxt, const double* ax, const double* ay,
                     const double* az, const int* inter, int n, Systems& sys, int* flags,
                     hipStream_t s);

// Small fills (no hipMemsetAsync: these run inside captured graphs too).
void launch_fill_int(int* p, int v, size_t n, hipStream_t s);
void launch_fill_u64(unsigned long long* p, unsigned long long v, size_t n,
                     hipStream_t s);

// Device initial conditions (ic_device.hip): settings shared by all
// systems, and per system its gen type, seeded engine state and draws per
// particle.
enum { kIcCube = 0, kIcSphere = 1, kIcShell = 2 };
struct IcParams {
  double min_mass, max_mass, size, min_vel, max_vel, r_sphere, shell_jitter;
  double scale[3];
  double rot[3];  // rotation_vector normalised, times rotation_magnitude
};
struct IcSystem {
  int gen_type;
  int draws;     // engine draws per particle
  unsigned x0;   // engine state after seeding: seed mod 2^31 - 1, 0 -> 1
  int pad;
};
// Particles [off[s], off[s + 1]) of system s into st (caller order, id = i).
void launch_generate(const IcParams& p, int nsys, const long long* off, const IcSystem* sys,
                     long long n, State st, hipStream_t s);

}  // namespace nbs
