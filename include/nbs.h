/* nbs.h -- C ABI of the MI355X Barnes-Hut step engine (libnbs.so).
 *
 * Drop-in for nbodysim's per-frame step.  The reference crosses the
 * host/device boundary once per frame, at
 *     void barnes_hut_cuda(std::unordered_set<particle*>*, octree*,
 *                          bool damping, datatype theta);
 *     (reference src/cuda_helper.h:8, called from src/main.cpp:1449-1452)
 * with the octree built on the host (src/main.cpp:887-927,
 * src/octree.cpp:134-313) and the update done on the host
 * (src/main.cpp:929-973).  Here the whole frame -- root, octree, mass/COM,
 * walk, update -- runs on the GPU, so the boundary moves to step
 * granularity: the caller hands over particle state once and asks for
 * forces (nbs_forces, barnes_hut_cuda's role) or whole frames (nbs_step).
 *
 * Conventions: every function returns NBS_OK (0) or an NBS_ERR_* code and
 * never exits the process (the reference's cuda_assert calls exit,
 * src/cuda_code.cu:29-37); nbs_last_error() gives the message.  Host
 * arrays are particle-major [n][3] doubles (x, y, z), in the caller's
 * particle order.  One context = one device, one stream, one host thread;
 * calls on one context must not overlap.
 */
#ifndef NBS_H_
#define NBS_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct nbs_ctx nbs_ctx;

enum {
  NBS_OK = 0,
  NBS_ERR_ARG = 1,      /* bad argument */
  NBS_ERR_HIP = 2,      /* HIP runtime failure */
  NBS_ERR_STATE = 3,    /* call out of order (e.g. no state set) */
  NBS_ERR_OVERFLOW = 4, /* tree or sort scratch exceeded its capacity */
  NBS_ERR_NOMEM = 5,    /* device allocation failed */
  NBS_ERR_IO = 6        /* file I/O (dump / resume) */
};

/* Force arithmetic.  State (positions, velocities), the octree keys and the
 * centre-of-mass reduction are FP64 in both modes. */
enum { NBS_FORCE_F64 = 0, NBS_FORCE_F32 = 1 };

/* Root cell.  BBOX is the reference's gen_root (bounding-box midpoint,
 * largest extent + 2, src/main.cpp:887-921); FIXED uses root_center and
 * root_size. */
enum { NBS_ROOT_BBOX = 0, NBS_ROOT_FIXED = 1 };

typedef struct nbs_config {
  double theta;          /* settings.cfg `theta` (src/main.cpp:1287) */
  double dt;             /* settings.cfg `dt` (src/main.cpp:1288) */
  int damping;           /* settings.cfg `damping` (src/main.cpp:1193-1198) */
  int force_precision;   /* NBS_FORCE_F64 | NBS_FORCE_F32 */
  int root_mode;         /* NBS_ROOT_BBOX | NBS_ROOT_FIXED */
  int adaptive;          /* adaptive dt (documented in the reference README,
                            README.md:54-60; not implemented there) */
  double root_center[3]; /* NBS_ROOT_FIXED only */
  double root_size;      /* NBS_ROOT_FIXED only */
  double min_adaptive_dt;
  double max_pos_change;
  double max_vel_change;
  int device;            /* HIP device ordinal */
  int node_factor;       /* node capacity = node_factor * n + 4096 (0: 4) */
  int threaded_tree;     /* 1: the tree of the reference's threads > 1 path
                            (gen_root_threaded, src/main.cpp:817-886): all
                            eight root children exist (an empty one is a
                            cell of mass 0 at the origin) and a particle on
                            the root's centre planes has a leaf in every
                            child whose closed cell holds it (inside(),
                            src/octree.cpp:320-336).  One system, bbox
                            root; no min_node_size, adjoint or graph
                            capture (nbs_step_graph runs eager steps). */
  int reserved[5];
} nbs_config;

/* Tree summary of the last build. */
typedef struct nbs_tree_info {
  int64_t n_particles;
  int64_t n_nodes;       /* internal nodes + leaves, all systems */
  int64_t n_internal;
  int32_t max_depth;
  int32_t n_buckets;     /* depth-42 cells holding >= 2 particles (the
                            reference would split deeper); 0 = exact */
  int64_t node_capacity;
} nbs_tree_info;

/* Defaults of src/main.cpp:1312-1357 that the step uses (theta 0.5,
 * dt 0.0333, no damping) plus FP64 forces and the bbox root. */
void nbs_config_default(nbs_config* cfg);

int nbs_create(const nbs_config* cfg, nbs_ctx** out);
void nbs_destroy(nbs_ctx* ctx);
const char* nbs_last_error(const nbs_ctx* ctx);

/* Replaces the configuration (theta, dt, precision, ...) of a context. */
int nbs_configure(nbs_ctx* ctx, const nbs_config* cfg);

/* Uploads n particles (host arrays; vel may be NULL = zero).  Replaces the
 * reference's particle_set contents (src/data_structures.h:9). */
int nbs_set_state(nbs_ctx* ctx, size_t n, const double* pos,
                  const double* vel, const double* mass);

/* Batched mode: splits the uploaded particles into nsys independent
 * systems, system s = caller indices [offsets[s], offsets[s+1]).
 * offsets has nsys + 1 entries, offsets[0] = 0, offsets[nsys] = n, and
 * increases strictly (every system holds at least one particle;
 * NBS_ERR_ARG otherwise). */
int nbs_set_systems(nbs_ctx* ctx, int nsys, const int64_t* offsets);

/* Downloads state in the caller's order; any pointer may be NULL.  acc is
 * the acceleration of the last force evaluation (what the reference keeps
 * in particle::acceleration, src/particle.h:35). */
int nbs_get_state(nbs_ctx* ctx, double* pos, double* vel, double* acc);

/* Device-memory forms of nbs_set_state and nbs_get_state (no counterpart in
 * the reference; for callers whose particles already live on the GPU, such
 * as PyTorch tensors): the same layouts and meaning, every array in device
 * memory of the HIP runtime this library is bound to (checked with
 * hipPointerGetAttributes: a host pointer, or memory of a second HIP
 * runtime loaded in the same process, is NBS_ERR_ARG before any copy).
 * The arrays must be ready when the call is made (their producers'
 * streams synchronized); both return after their copies have completed,
 * with no transfer over PCIe. */
int nbs_set_state_device(nbs_ctx* ctx, size_t n, const double* pos,
                         const double* vel, const double* mass);
int nbs_get_state_device(nbs_ctx* ctx, double* pos, double* vel, double* acc);

/* Downloads the masses in the caller's order (the mass field of the
 * reference's particle record, src/particle.h:36). */
int nbs_get_mass(nbs_ctx* ctx, double* mass);

/* Asynchronous snapshot for frame dumps: the reference copies every
 * particle into its dump thread's set before writing
 * (src/main.cpp:1462-1489, dump_threaded :768-788); here
 * nbs_snapshot_begin enqueues, behind the work already submitted, a gather
 * of pos, vel and acc into caller order on the device and a copy into a
 * pinned host buffer on a second stream, and returns at once, so the copy
 * overlaps the steps submitted next.  nbs_snapshot_end waits for that copy
 * and returns [n][3] pos, vel, acc (acc zero before the first force
 * evaluation) and n.  The context owns two pinned buffers and alternates
 * between them: the arrays stay valid until the second nbs_snapshot_begin
 * after the one they came from (or until a snapshot of more particles than
 * any before, or nbs_destroy).  One snapshot may be pending at a time.  No
 * counterpart in the reference's GPU boundary. */
int nbs_snapshot_begin(nbs_ctx* ctx);
int nbs_snapshot_end(nbs_ctx* ctx, const double** pos, const double** vel,
                     const double** acc, size_t* n);

/* Accelerations for the current positions: barnes_hut_cuda's role
 * (src/cuda_helper.cpp:218-246) including the tree build it was handed.
 * acc (host, may be NULL to keep the result on the device). */
int nbs_forces(nbs_ctx* ctx, double* acc);

/* nbs_forces for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  One force evaluation of
 * the n particles in the caller's DEVICE arrays pos ([n][3] FP64) and mass
 * ([n]) with the configuration cfg (device: the current one): the same
 * root, tree, walk and results as nbs_set_state + nbs_forces on a context.
 * acc (device, [n][3], caller order) receives the accelerations and counts
 * (device, [n] int64, may be NULL) each particle's interactions.  It calls
 * no hipMalloc / hipFree: every buffer comes from one fixed library-owned
 * device array per GPU, sized for at most NBS_NOALLOC_MAX_N particles at
 * node_factor <= 4 (56 MiB, held until the process ends).  Larger n
 * (or a tree that outgrows the fixed node capacity) is rejected with
 * NBS_ERR_ARG / NBS_ERR_OVERFLOW and a message in msg (may be NULL).  Work
 * is enqueued on `stream` (hipStream_t, NULL = default stream) and the call
 * returns after it has finished.  Every call on one GPU shares that GPU's
 * array: calls that overlap on one GPU (other threads, other streams)
 * race and must be serialised by the caller. */
#define NBS_NOALLOC_MAX_N 65536
int nbs_forces_noalloc(const nbs_config* cfg, size_t n, const double* pos,
                       const double* mass, double* acc, int64_t* counts, void* stream,
                       char* msg, size_t msg_len);

/* nsteps frames of the reference loop body (src/main.cpp:1422-1480):
 * root, octree, mass, COM, walk, kick-drift update. */
int nbs_step(nbs_ctx* ctx, int nsteps);

/* nbs_step through a captured HIP graph: two steps are captured once
 * (kernels only, no memset or copy nodes) and replayed, which removes the
 * per-kernel launch cost of the ~35 kernels of a step.  Results are
 * identical to nbs_step.  The graph is rebuilt after any change of state
 * size, systems, configuration or node capacity. */
int nbs_step_graph(nbs_ctx* ctx, int nsteps);

/* Collision merging (config.collide / collision_range, src/main.cpp:1442-1483;
 * the reference's CUDA path ignores it, :1445-1448).  While enabled, each
 * nbs_step step records the pairs (particle, leaf its walk reaches under
 * the FP64 MAC) closer than `range` (src/main.cpp:386-396), merges every connected
 * group of pairs into one particle -- collide() folded over the members in
 * increasing caller index: mass sum, mass-weighted position and velocity
 * (src/main.cpp:423-430, src/vector.cpp:129-132) -- with zero acceleration
 * for the step's update, then updates and removes the absorbed particles.
 * Survivors keep their caller order; the merged particle takes its
 * lowest-index member's place.  The count shrinks (nbs_particle_count).
 * Eager steps only: nbs_step_graph returns NBS_ERR_STATE while enabled;
 * adaptive dt must be off.  nbs_forces never merges. */
int nbs_set_collisions(nbs_ctx* ctx, int enabled, double range);
/* apply_minimum_size (src/main.cpp:1001-1057; the reference runs it when
 * min_node_size > 0 and verbose is set, :1425): while min_size > 0, each
 * nbs_step step builds the tree, merges the particles of every largest
 * cell of side < min_size into one -- collide() folded along the tree, in
 * child order 0..7 at each level -- and walks the tree of the merged
 * particles on the same root.  The merged particle takes its lowest-index
 * member's place; the count shrinks.  Eager steps only, adaptive dt off.
 * merged (may be NULL) receives the particles absorbed so far (before this
 * call; the count restarts when the size changes). */
int nbs_set_min_node_size(nbs_ctx* ctx, double min_size, int64_t* merged);
/* Pairs detected and particles absorbed since collisions were enabled. */
int nbs_collision_stats(nbs_ctx* ctx, int64_t* pairs, int64_t* removed);
/* Current particle count and, if offsets is not NULL, the nsys + 1 system
 * offsets (nsys_cap >= nsys + 1). */
int nbs_particle_count(nbs_ctx* ctx, int64_t* n, int64_t* offsets, int nsys_cap);

/* The frame image of write_image (src/main.cpp:556-678) for the current
 * positions, rendered on the device: projection 0 front, 1 side, 2 top,
 * 3 iso; color 0 bw, 1 heat; adaptive / nonlinear are the settings'
 * adaptive_brightness / nonlinear_brightness.  rgb receives height rows of
 * width RGB triples, top row first (the PNG order; the reference's pixel
 * (x, y) is row height - 1 - y).  width and height must be at least 1. */
typedef struct nbs_image_params {
  uint32_t width, height, projection, color;
  int32_t adaptive, nonlinear;
  double scale, brightness;
} nbs_image_params;
int nbs_render_image(nbs_ctx* ctx, const nbs_image_params* p, uint8_t* rgb);

/* Launch-shape tuning of the walk (no counterpart in the reference, whose
 * kernel shape was fixed, src/cuda_helper.cpp:183-188).  The shape is the
 * workgroup size (block: 64, 128, 256, 512, 1024 threads), the number of
 * contiguous particle regions interleaved over the 8 XCDs (groups: 1 =
 * plain order, 8, 16, 32, 64) and the particles per lane (pack: 1, or 2
 * sharing one traversal); it changes the speed only, never a result bit.
 * nbs_autotune times `reps` whole steps from the current state for each
 * of the 50 shapes, keeps the
 * fastest for this context, reports it (best_*; may be NULL) and the times
 * in ms per step (ms[(pack_index * 5 + block_index) * 5 + groups_index],
 * may be NULL; -1 = not timed), and restores the state, accelerations, time
 * and interaction counters as they were (the tree must be rebuilt:
 * nbs_tree_dump etc. need a new step or nbs_forces).  nbs_set_walk_shape
 * sets the shape directly. */
int nbs_autotune(nbs_ctx* ctx, int reps, int* best_block, int* best_groups,
                 int* best_pack, double* ms);
int nbs_set_walk_shape(nbs_ctx* ctx, int block, int groups, int pack);

/* Longest-first dispatch (default off; no counterpart in the reference):
 * each walk records the node visits of every workgroup's tile and the next
 * walk dispatches the tiles of each XCD region most expensive first, so a
 * launch does not end with a few long workgroups running alone.  Changes
 * the speed only, never a result bit.  On the uniform-sphere bench it
 * shortens a repeated walk of one tree by ~4% but slows whole steps by
 * ~5%: the cost order scatters Morton-neighbour tiles in time, and a
 * freshly built tree then misses more in L2. */
int nbs_set_walk_order(nbs_ctx* ctx, int on);

/* Which instruction stream runs the FP32 walk (no counterpart in the
 * reference; for tests and diagnostics): 0 (default) = the hand-written
 * gfx950 loop where it applies (FP32 forces, one particle per lane),
 * reading 32-byte node records from 2^21 particles on and the 64-byte
 * ones below; 1 = the compiler-generated loop everywhere; 2 / 3 = the
 * hand-written loop on 64-byte / 32-byte records at any size (32-byte
 * records need fewer than 2^25 nodes).  All make the same decisions and
 * the same arithmetic in the same order: changes the speed only, never a
 * result bit. */
int nbs_set_walk_code(nbs_ctx* ctx, int code);

/* Pieces per walk traversal (no counterpart in the reference): 0 (default)
 * = by the wavefront count, so that a system too small to fill the GPU's
 * wavefront slots (fewer than ~4096 64-particle wavefronts) walks each
 * traversal as 2, 4 or 8 pieces over equal ranges of its nodes, whose
 * partial sums are added in piece order; a power of two up to 32 = that
 * many (more pieces than the GPU has room for only add work).  Every
 * particle's accepted cells and interaction count are those of the whole
 * traversal; the forces differ from it only by the summation order, and a
 * given state and setting always give the same bits.  nbs_walk_split
 * returns the pieces in use. */
int nbs_set_walk_split(nbs_ctx* ctx, int pieces);
int nbs_walk_split(nbs_ctx* ctx);

/* Blocks until the context's stream has drained. */
int nbs_synchronize(nbs_ctx* ctx);

/* FP64 direct sum (no tree) at the current positions for the m particles
 * whose caller indices are in ids; acc is host [m][3].  The accuracy
 * reference for the tree walk. */
int nbs_direct(nbs_ctx* ctx, size_t m, const int64_t* ids, double* acc);

/* Interaction counting of the walks (no counterpart in the reference, whose
 * step counts nothing): NBS_COUNT_PARTICLES (default) keeps every
 * particle's count of the last walk (nbs_interactions) and the running
 * total (nbs_interactions_total); NBS_COUNT_TOTAL only the running total,
 * counted per wavefront on the scalar unit; NBS_COUNT_NONE nothing.  The
 * per-particle count is one vector instruction per interacting node visit
 * of the walk, so the frame loop steps with NBS_COUNT_TOTAL.  Changes no
 * force, position or velocity bit. */
enum { NBS_COUNT_NONE = 0, NBS_COUNT_TOTAL = 1, NBS_COUNT_PARTICLES = 2 };
int nbs_set_counting(nbs_ctx* ctx, int mode);

/* Interaction count (accepted nodes per particle) of the last walk, summed
 * over all particles; per_particle (host, caller order) may be NULL.
 * NBS_ERR_STATE when the last walk ran without NBS_COUNT_PARTICLES. */
int nbs_interactions(nbs_ctx* ctx, int64_t* total, int64_t* per_particle);

/* Interactions summed over every walk since the last reset (reset != 0
 * clears the counter after reading); walks under NBS_COUNT_NONE add
 * nothing. */
int nbs_interactions_total(nbs_ctx* ctx, int64_t* total, int reset);

/* Simulation time advanced so far, summed over steps (differs from
 * steps * dt only with adaptive dt); per system when batched. */
int nbs_time(nbs_ctx* ctx, double* t, int nsys_cap);

int nbs_tree_info_get(nbs_ctx* ctx, nbs_tree_info* info);

/* Test hook: the last tree in pre-order, children 0..7, every system after
 * the other.  Arrays have n_nodes entries (order: one per leaf,
 * n_nodes - n_internal; n_particles except with the threaded tree):
 *   depth, leaf (caller index of a leaf's particle, -1 internal,
 *   -2 bucket), first / count (position range in `order`), side, mass,
 *   com [n_nodes][3], key_hi / key_lo (node path, 3 bits per level,
 *   level 0 in bits 62..60 of key_hi), order (caller indices in DFS leaf
 *   order).  Any pointer may be NULL. */
int nbs_tree_dump(nbs_ctx* ctx, int32_t* depth, int32_t* leaf,
                  int64_t* first, int64_t* count, double* side,
                  double* mass, double* com, uint64_t* key_hi,
                  uint64_t* key_lo, int32_t* order);

/* Per-particle 42-level keys (caller order) of the last build. */
int nbs_keys(nbs_ctx* ctx, uint64_t* key_hi, uint64_t* key_lo);

/* Root (centre, side) per system of the last build; center [nsys][3]. */
int nbs_roots(nbs_ctx* ctx, double* center, double* side, int nsys_cap);

/* Reverse mode (no counterpart in the reference).  nbs_adjoint_forward
 * runs nsteps frames from the current state (fixed dt, one system),
 * keeping (x, v) checkpoints every `every` steps and, with record_all, the
 * positions of every step.  nbs_adjoint_backward takes dL/dx_T and
 * dL/dv_T (host [n][3], NULL = 0) and returns dL/dx_0, dL/dv_0 (host
 * [n][3]) and dL/dm (host [n]): each segment is recomputed from its
 * checkpoint and reversed with the walk's cell decisions of the forward.
 * The context is left at the forward's final state.  The gradient is that
 * of the step with the tree topology and accepted cells held fixed (the
 * decisions are step functions of the positions). */
int nbs_adjoint_forward(nbs_ctx* ctx, int nsteps, int every, int record_all);
int nbs_adjoint_backward(nbs_ctx* ctx, const double* gx_T, const double* gv_T,
                         double* gx0, double* gv0, double* gm);
/* The same with every array in device memory of this library's HIP runtime
 * (checked as for nbs_set_state_device); returns after the gradients are
 * written. */
int nbs_adjoint_backward_device(nbs_ctx* ctx, const double* gx_T, const double* gv_T,
                                double* gx0, double* gv0, double* gm);

/* Replays the recorded forward from perturbed initial data with the tree
 * topology and cell decisions of the recorded trajectory (FP64 only): the
 * smooth map the adjoint differentiates, for finite-difference checks.
 * Outputs host [n][3]; the context is left at the forward's final state. */
int nbs_frozen_forward(nbs_ctx* ctx, const double* pos0, const double* vel0,
                       const double* mass, double* posT, double* velT);

/* Phase timing on the context's stream with HIP events (ms, summed since
 * the last reset): [0] tree build (bbox..mass/COM), [1] walk (incl. fused
 * update), [2] separate update/adaptive-dt kernels, [3] walk launches. */
int nbs_profile_enable(nbs_ctx* ctx, int on);
int nbs_profile_read(nbs_ctx* ctx, double* ms4, int reset);

/* Raw device stream (hipStream_t) of the context, for interop. */
void* nbs_stream(nbs_ctx* ctx);

#ifdef __cplusplus
}
#endif

#endif /* NBS_H_ */
