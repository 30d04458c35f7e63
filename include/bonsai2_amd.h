/* bonsai2_amd.h -- C ABI of the MI355X Barnes-Hut treecode (libbonsai2_amd.so).
 *
 * The reference (treecode/Bonsai2) has no C ABI: its boundary is the C++
 * class template Treecode<real_t> (code/Treecode.h:123-223), instantiated for
 * float only (code/TreecodeInstances.h:3).  Each entry point below replaces
 * one member of that class; include/Treecode.h rebuilds the class on top of
 * these functions so that reference callers compile unchanged.
 *
 * Conventions
 *   - Every call returns 0 on success or a negative BT_ERR_* code; the text
 *     of the last error is available from bt_last_error().
 *   - All state lives in the context (no process-global device symbols, unlike
 *     code/buildTree.cu:23-36 and code/computeForces.cu:465-473).  One context
 *     is used by one host thread at a time.
 *   - Particle arrays are 4 floats per particle: pos = (x, y, z, mass),
 *     vel = (vx, vy, vz, mass) as the reference driver packs them
 *     (code/treecode.cu:73-86).  The caller owns host buffers; the context
 *     owns device buffers.
 *   - Orders: BT_ORDER_INPUT = the order given to bt_set_particles;
 *     BT_ORDER_TREE = octree order after bt_build_tree (reference d_ptclPos);
 *     BT_ORDER_PH = Peano-Hilbert order after bt_make_groups (reference
 *     d_ptclPos_tmp / d_ptclAcc).
 *   - Work is queued on the context's stream; calls that return host data
 *     synchronise that stream.
 */
#ifndef BONSAI2_AMD_H
#define BONSAI2_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bt_ctx bt_ctx;

enum {
  BT_OK = 0,
  BT_ERR_ARG = -1,            /* bad argument (NLEAF, NCRIT, sizes, order) */
  BT_ERR_HIP = -2,            /* HIP runtime error */
  BT_ERR_STATE = -3,          /* call order violated (e.g. forces before groups) */
  BT_ERR_CAPACITY = -4,       /* cell or group capacity exceeded */
  BT_ERR_STACK = -5,          /* walk stack overflow (reference: :302-303, :565) */
  BT_ERR_NODEVICE = -6,       /* no HIP device */
  BT_ERR_IO = -7              /* tipsy file error */
};

enum { BT_ORDER_INPUT = 0, BT_ORDER_TREE = 1, BT_ORDER_PH = 2 };

/* Build flags returned by bt_get_counts()[7]. */
enum { BT_FLAG_DEPTH_CAP = 1 };

/* ---- lifetime: Treecode::Treecode / ~Treecode / alloc ---------------- */

/* Treecode(eps, theta, ncrit), code/Treecode.h:164-173. */
int bt_create(bt_ctx **ctx, float eps, float theta, int ncrit);
/* ~Treecode (RAII cuda_mem frees, code/cudamem.h:12). */
int bt_destroy(bt_ctx *ctx);
/* Treecode::alloc(n), code/Treecode.h:175-198.  nsys >= 1 independent
 * systems ("forest"); offsets[0..nsys] partition [0, n) (null when nsys=1).
 * A later call with a larger n or different systems reallocates. */
int bt_alloc(bt_ctx *ctx, int n, int nsys, const int *offsets);
/* Queue all work on this stream (default: the context's own stream).  A step
 * (bt_step, bt_step_graph) also runs part of its work on a second stream the
 * context owns, forked from and joined back into this one (bt_set_overlap), so
 * work queued on this stream after a step sees all of it. */
int bt_set_stream(bt_ctx *ctx, void *hip_stream);
/* eps/theta are public members in the reference (code/Treecode.h:133). */
int bt_set_params(bt_ctx *ctx, float eps, float theta);
const char *bt_last_error(const bt_ctx *ctx);

/* ---- particles: Treecode::ptcl_h2d / ptcl_d2h --------------------------- */

/* ptcl_h2d(), code/Treecode.h:207-214: copies n particles (input order). */
int bt_set_particles(bt_ctx *ctx, const float *pos4, const float *vel4);
/* Same, from device pointers (no host staging): float4 arrays, or double4
 * for a Treecode<double> context (bt_create_f64). */
int bt_set_particles_device(bt_ctx *ctx, const void *d_pos4, const void *d_vel4);
/* ptcl_d2h(), code/Treecode.h:200-204, in the requested order. */
int bt_get_particles(bt_ctx *ctx, float *pos4, float *vel4, int order);

/* ---- the gravity step -------------------------------------------------- */

/* buildTree(nLeaf), code/buildTree.cu:845-975.  nleaf in {16,24,32,48,64}
 * (:849).  Leaves the particles in tree order. */
int bt_build_tree(bt_ctx *ctx, int nleaf);
/* computeMultipoles(), code/computeMultipoles.cu:194-234. */
int bt_compute_multipoles(bt_ctx *ctx);
/* makeGroups(levelSplit, nCrit), code/makeGroups.cu:236-341.
 * 1 <= level_split <= 7, 1 <= ncrit <= 64. */
int bt_make_groups(bt_ctx *ctx, int level_split, int ncrit);
/* computeForces(INTCOUNT), code/computeForces.cu:622-727.  When stats is not
 * null it receives {direct_sum/N, direct_max, approx_sum/N, approx_max}
 * (:701-715); grav_potential (:711) is returned by bt_get_potential(). */
int bt_compute_forces(bt_ctx *ctx, int intcount, double stats[4]);
/* moveParticles(), code/moveParticles.cu:3-7 (a stub in the reference):
 * the first half of a kick-drift-kick leapfrog step, v += a*dt/2 with the
 * current accelerations, then x += v*dt.  The step's second half kick is
 * owed to the next force evaluation: the next bt_compute_forces (or the
 * force evaluation bt_compute_energies makes) applies v += a_new*dt/2, so
 * the reference's loop buildTree .. computeForces, moveParticles integrates
 * a KDK leapfrog. */
int bt_move_particles(bt_ctx *ctx, float dt);
/* Half kick only (closes a kick-drift-kick step after new forces). */
int bt_kick(bt_ctx *ctx, float dt_half);
/* computeEnergies(), code/computeEnergies.cu:3-7 (a stub in the reference):
 * FP64 reductions K = 1/2 sum m v^2 and W = 1/2 sum m (phi + m/eps), per
 * system (arrays of nsys) when ekin/epot are not null.  When the forces are
 * not current (after bt_move_particles, or none yet), it first evaluates them
 * (buildTree, computeMultipoles, makeGroups, computeForces with the last
 * NLEAF, levelSplit and NCRIT), which also applies a pending half kick. */
int bt_compute_energies(bt_ctx *ctx, double *ekin, double *epot);
/* One full kick-drift-kick step: forces are (re)computed as needed.  If the
 * cell arrays overflow, they are grown and the step's second half is run
 * again.  On BT_ERR_STACK the particles hold the step's first half kick and
 * drift (the closing kick is not applied with invalid forces). */
int bt_step(bt_ctx *ctx, float dt, int nleaf, int level_split, int ncrit);

/* ---- results ----------------------------------------------------------- */

/* d_ptclAcc (code/Treecode.h:144): (ax, ay, az, phi) per particle. */
int bt_get_acc(bt_ctx *ctx, float *acc4, int order);
/* Interaction statistics of the last walk, as bt_compute_forces returns them:
 * {direct_sum/N, direct_max, approx_sum/N, approx_max} (code/computeForces.cu
 * :701-715), plus the raw integer totals {direct_sum, approx_sum} in totals2
 * (may be null). */
int bt_get_walk_stats(bt_ctx *ctx, double stats[4], double totals2[2]);
/* grav_potential (code/Treecode.h:162): 1/2 sum m phi incl. self term. */
int bt_get_potential(bt_ctx *ctx, double *pot);
/* Counters: {n, ncells, nnodes, nleaves, nlevels, ngroups, nsys, flags}
 * (reference private members nCells/nNodes/nLeaves/nLevels/nGroups,
 * code/Treecode.h:135). */
int bt_get_counts(bt_ctx *ctx, int counts[8]);
/* Copies of the device tree in the reference layouts:
 * cells = CellData uint4 (code/Treecode.h:45-89), level_range = int2[32]
 * (d_level_begIdx), size/mono/quad0 = float4, quad1 = float2 (:158-160).
 * Any pointer may be null.  Arrays must hold ncells entries. */
int bt_get_tree(bt_ctx *ctx, uint32_t *cells4, int *level_range64, float *size4,
                float *mono4, float *quad0_4, float *quad1_2);
/* (l, s) per cell: bbox size and bbox-centre to com distance
 * (code/computeMultipoles.cu:171-175). */
int bt_get_cell_ls(bt_ctx *ctx, float *ls2);
/* Group list (code/Treecode.h:29-43): (pbeg, np) in PH order. */
int bt_get_groups(bt_ctx *ctx, int *groups2);
/* Per-group interaction counts (n_approx, n_direct) of the last walk. */
int bt_get_group_counts(bt_ctx *ctx, int *counts2);
/* Sorted PH keys and permutation (d_key/d_value, code/makeGroups.cu:270-280):
 * perm[k] = tree index of PH index k.  Either pointer may be null. */
int bt_get_ph(bt_ctx *ctx, uint64_t *keys, int *perm);
/* Original (input) index of each particle in the given order. */
int bt_get_ids(bt_ctx *ctx, int *ids, int order);
/* Root box per system {cx, cy, cz, hsize} (d_domain, code/buildTree.cu:161). */
int bt_get_boxes(bt_ctx *ctx, float *boxes4);

/* ---- checks and measurement ------------------------------------------- */

/* FP64 direct sum (the error oracle of the walk) for targets given as PH
 * indices; sources = all particles of the target's system.  out = ns x
 * (ax, ay, az, phi) doubles, self term included in phi. */
int bt_direct_fp64(bt_ctx *ctx, int ns, const int *ph_idx, double *out4);
/* bt_direct_fp64 for callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops, harnesses that forbid hipMalloc).  No
 * counterpart in the reference.  Targets d_ph_idx (ns ints) and results
 * d_out4 (ns x 4 doubles) are device memory the caller owns; the call is
 * enqueued on the context stream and does not synchronise.  The partial sums
 * live in a fixed 16 MiB device array per GPU (held until the process ends),
 * so ns <= 4096 (else BT_ERR_ARG) and calls on one GPU that overlap in time,
 * from any context or stream, race on that array.  An index outside [0, n)
 * gives a NaN row.  No hipMalloc, hipFree or other allocating call. */
int bt_direct_fp64_noalloc(bt_ctx *ctx, int ns, const int *d_ph_idx, double *d_out4);
/* Milliseconds of the phases of the last step, measured with HIP events on
 * the context stream: {bbox+keys, sort, tree levels, multipoles, ph keys,
 * ph sort, groups, walk, integrate, energies}.  In a step run with
 * bt_set_overlap(ctx, 1) the last three of the first seven (on a second
 * stream) overlap the levels and multipoles, and each is timed with the
 * other branch beside it. */
int bt_get_timings(bt_ctx *ctx, float ms[10]);
/* Walk options: kahan = compensated accumulators (default 1);
 * stack_ints = per-wave traversal stack capacity (default 16384, at least 64;
 *   a walk that needs more returns BT_ERR_STACK);
 * blocks_per_cu = persistent grid size (default 0 = automatic). */
int bt_set_walk_options(bt_ctx *ctx, int kahan, int stack_ints, int blocks_per_cu);
/* How the walk feeds interaction data to the lanes: 0 (default) = one
 * i-particle per lane, the wave stages each batch of 64 cells / particles in
 * LDS and sweeps it with broadcast reads; 1 = one i-particle per lane, each
 * batch entry read by scalar loads straight from the immutable cell /
 * particle arrays and used from SGPRs; 2 = each lane carries two of the
 * group's i-particles and each half-wave sweeps half of the LDS batch (every
 * LDS read serves two interactions, at the cost of registers and occupancy).
 * All evaluate the same interaction lists; forces agree to FP32 rounding. */
int bt_set_walk_variant(bt_ctx *ctx, int variant);
/* Treecode<float> walks (no reference counterpart): entries of each wave's
 * traversal stack kept in LDS, 512..1024 (default 1024); deeper entries spill
 * to the global stack of bt_set_walk_options and come back in LIFO order, so
 * the traversal and the forces do not depend on it.  A walk whose stack
 * needs more than these entries plus stack_ints returns BT_ERR_STACK. */
int bt_set_walk_lds_stack(bt_ctx *ctx, int entries);
/* Steps of bt_step and bt_step_graph (no reference counterpart: the
 * reference's phases run in sequence on one stream, code/treecode.cu:137-165):
 * on = 1 (default) runs makeGroups' work (PH keys, their sort, the groups) on
 * a second stream beside the level build and the multipoles, both after the
 * geometric-key sort, and the walk after both; on = 0 runs the phases in
 * sequence, so that bt_get_timings gives each phase alone.  The results are
 * bitwise the same.  bt_build_tree, bt_compute_multipoles and bt_make_groups
 * always run on the context stream. */
int bt_set_overlap(bt_ctx *ctx, int on);
/* Groups of the walk (no reference use; code/buildTree.cu:954-958 collects the
 * leaf list d_leafList and leaves it unused): mode 0 (default) = makeGroups'
 * PH-key groups (code/makeGroups.cu); mode 1 = one group per leaf cell, in
 * tree order; mode 2 = one group per tree cell of at most ncrit particles
 * whose parent holds more (and per leaf of more than ncrit), in tree order.
 * In modes 1 and 2 the "PH order" of bt_get_acc and the other per-group
 * arrays is the tree order (perm = identity), level_split is not used, and a
 * leaf of more than 64 particles (a leaf capped at the deepest level) returns
 * BT_ERR_CAPACITY.  Takes effect at the next makeGroups. */
int bt_set_group_mode(bt_ctx *ctx, int mode);
/* Deepest traversal stack of any group in the walks of the last call that
 * walked (entries, LDS and global parts together; all replays of a
 * bt_step_graph): the stack_ints a walk of this tree needs. */
int bt_get_walk_stack_max(bt_ctx *ctx, int *entries);
/* code/Treecode.h:14-16 (#define QUADRUPOLE): 1 (default) = monopole +
 * quadrupole P2C (code/computeForces.cu:83-128, 64 flops); 0 = monopole-only
 * P2C, the cell evaluated as a point mass at its centre of mass
 * (code/computeForces.cu:198-225, 20 flops; code/treecode.cu:149-153). */
int bt_set_quadrupole(bt_ctx *ctx, int on);

/* ---- captured step and autotuner (no reference counterpart) ---------------
 * bt_step_graph: the step of bt_step (kick, drift, buildTree,
 * computeMultipoles, makeGroups, computeForces, kick) captured once into a
 * HIP graph and replayed nsteps times with no host work between kernels.  The
 * graph is re-captured when any value it holds changes (dt, nleaf,
 * level_split, ncrit, eps, theta, walk options, the opening mode, budget and
 * band, the check's settings), and dropped when arrays are reallocated.
 * Needs forces (or one bt_step) first.  With the learned angle the per-step
 * FP64 check and its fixed-angle fallback are part of the graph (the device
 * decides).  Errors of every replay are kept (a sticky device word that only
 * the host clears); after one, later replays leave the particles alone.  A
 * cell-capacity overflow grows the cells, finishes the failed step and runs
 * the remaining steps eagerly; a walk-stack overflow returns BT_ERR_STACK
 * with the particles as after the failed step's first half kick and drift. */
int bt_step_graph(bt_ctx *ctx, float dt, int nleaf, int level_split, int ncrit, int nsteps);
/* Picks NLEAF, NCRIT (3 x 3 candidates, timed zero-dt steps), the persistent
 * walk grid (4, 8, 12, 16 blocks per CU) and the traversal stack depth
 * (4x the deepest stack seen) for the current particles, and leaves them set.
 * table (may be NULL) gets 6 doubles per timed candidate: nleaf, ncrit,
 * blocks_per_cu, step ms, walk ms, deepest stack; *ntable their count (<= 13).
 * best = {nleaf, ncrit, blocks_per_cu, stack entries}.  flags: reserved, 0. */
int bt_autotune(bt_ctx *ctx, int level_split, int flags, double *table, int *ntable,
                int best[4]);

/* ---- learned opening angle (no reference counterpart) --------------------
 * The reference opens every cell with one angle theta (code/computeForces.cu
 * :30-54).  With opening mode 1 a network picks the angle of each
 * (group, candidate cell) pair from 8 geometric features:
 *   theta = clamp(exp2(y + log2(budget) / 3), theta_lo, theta_hi),
 * and the cell is opened iff ds^2 < (l / theta + s)^2 (or it holds < 3
 * particles).  The network is 8 -> 32 (SiLU) -> 32 (SiLU) -> 1 with FP16
 * parameters and FP32 accumulation; its 1377 parameters are, in order,
 * W1[32][8], b1[32], W2[32][32], b2[32], w3[32], b3.  It is trained on the
 * GPU on pairs the method generates: target = log2(theta_need) -
 * log2(e) / 3, where theta_need = l / (ds - s) is the pair's acceptance
 * boundary and e the pair's largest relative P2C error against the FP64
 * direct force of the cell's particles; accepting a pair iff its predicted
 * e <= budget is then the same as the angle rule above. */
int bt_mlp_init(bt_ctx *ctx, uint64_t seed, float b3);  /* Xavier-uniform weights, b3 */
int bt_mlp_set_params(bt_ctx *ctx, const float *params);  /* 1377 floats, stored FP16 */
int bt_mlp_get_params(bt_ctx *ctx, float *params);        /* the FP32 master copy */
/* y[k] = network(x[8k .. 8k+7]) by the walk's matrix-core routine. */
int bt_mlp_forward(bt_ctx *ctx, int n, const float *x, float *y);
/* mean squared error of the network on (x, t) and its gradient (FP64 out). */
int bt_mlp_grad(bt_ctx *ctx, int n, const float *x, const float *t, double *grad,
                double *loss);
/* Adam on (x, t): steps minibatches of batch samples, cosine-decayed lr;
 * loss_curve (may be NULL) gets the minibatch loss of every step. */
int bt_mlp_train(bt_ctx *ctx, int n, const float *x, const float *t, int steps, int batch,
                 float lr, uint64_t seed, float *loss_curve);
/* Training pairs of the groups gsel[0..nsel) (indices into bt_get_groups)
 * from the current tree: every cell the walk visits at theta_lo whose
 * theta_need lies in [theta_lo, theta_hi] and that holds >= 3 particles.
 * Writes up to max_pairs rows of feat[8], target, err, theta_need and
 * pair_gc[2] = (group, cell) (any output may be NULL); *npairs = pairs
 * found (may exceed max_pairs). */
int bt_learn_collect(bt_ctx *ctx, int nsel, const int *gsel, float theta_lo, float theta_hi,
                     int max_pairs, float *feat, float *target, float *err, float *theta_need,
                     int *pair_gc, int *npairs);
/* mode 0: fixed theta (default); mode 1: the network's per-pair angle. */
int bt_set_opening(bt_ctx *ctx, int mode, float budget, float theta_lo, float theta_hi);
/* After each learned-angle walk, nsamples (<= 1024; 0 = off) seeded particles
 * are re-checked against the FP64 direct sum; if their p99 relative force
 * error exceeds p99_max the forces are recomputed with the fixed angle.  The
 * decision is taken on the device (no host synchronisation).  The samples of
 * check k are the particles at positions mix64(seed ^ (k << 32) ^ j) mod n,
 * j < nsamples, of the particle arrays when the check is drawn (k counted
 * from this call): in bt_step and bt_step_graph the drifted particles before
 * the step's build, whose FP64 sums then run on a third stream beside the
 * build (bt_set_overlap(ctx, 0): before it, on the context stream); in
 * bt_compute_forces the tree order, after the walk. */
int bt_set_opening_check(bt_ctx *ctx, int nsamples, float p99_max, uint64_t seed);
/* {p50, p99, max of the last check, last call fell back (0/1), fallbacks,
 * checks} */
int bt_get_opening_check(bt_ctx *ctx, double out[6]);
/* The PH indices of the last check's samples (min(nsamples, n) ints), in draw
 * order; no reference counterpart (the check itself has none). */
int bt_get_check_samples(bt_ctx *ctx, int *ph_idx);

/* ---- Treecode<double> ----------------------------------------------------
 * The reference declares Particle4<double> (code/Particle4.h:106-125) but
 * instantiates Treecode for float only (code/TreecodeInstances.h:3).  A
 * context made by bt_create_f64 is Treecode<double>: particles, root box,
 * geometric and PH keys, moments, walk and integration in FP64 (double4
 * particles (x, y, z, m), double2 / double4 cell data).  The root box keeps
 * its extremes in FP64 (the reference's double build would pass them through
 * its float2 block reduction, code/buildTree.cu:102-105).  Entry points
 * without a real-typed argument (build, multipoles, groups, forces, energies,
 * counts, groups, keys, statistics, bt_direct_fp64, bt_step_graph, bt_step,
 * bt_move_particles, the device buffers) serve both kinds; the float data
 * entry points (bt_set_particles, bt_get_particles, bt_get_acc, the moments of
 * bt_get_tree, bt_get_cell_ls, bt_get_boxes) return BT_ERR_STATE on a double
 * context, and the _f64 ones on a float context.  The walk evaluates one
 * i-particle per lane (bt_set_walk_variant has no effect); the learned
 * opening angle and its training are float only. */
int bt_create_f64(bt_ctx **ctx, double eps, double theta, int ncrit);
int bt_set_params_f64(bt_ctx *ctx, double eps, double theta);
int bt_set_particles_f64(bt_ctx *ctx, const double *pos4, const double *vel4);
int bt_get_particles_f64(bt_ctx *ctx, double *pos4, double *vel4, int order);
int bt_get_acc_f64(bt_ctx *ctx, double *acc4, int order);
/* size4 = (com, cellOp^2), mono4 = (com, M), quad0_4, quad1_2 in FP64. */
int bt_get_tree_f64(bt_ctx *ctx, double *size4, double *mono4, double *quad0_4, double *quad1_2);
int bt_get_boxes_f64(bt_ctx *ctx, double *boxes4);
int bt_move_particles_f64(bt_ctx *ctx, double dt);
int bt_step_f64(bt_ctx *ctx, double dt, int nleaf, int level_split, int ncrit);

/* ---- memory of the reference's public members ---------------------------- */

/* Pinned host memory, the reference's host_mem<T> (code/cudamem.h:64-96,
 * cudaMallocHost): hipHostMalloc / hipHostFree.  Needs no context. */
int bt_host_alloc(void **ptr, size_t bytes);
int bt_host_free(void *ptr);

/* The device arrays behind the reference's public cuda_mem members
 * (code/Treecode.h:142-160): current address and element count.  Particle
 * arrays change address at every buildTree (the reference swaps its pointers
 * too, code/buildTree.cu:929), so a pointer is valid until the next build.
 * In a Treecode<double> context the float4 / float2 arrays hold double4 /
 * double2. */
enum {
  BT_BUF_PTCL_POS = 0,     /* d_ptclPos: float4 (x,y,z,m), tree order after buildTree */
  BT_BUF_PTCL_VEL = 1,     /* d_ptclVel: float4 (vx,vy,vz,m), same order */
  BT_BUF_PTCL_POS_TMP = 2, /* d_ptclPos_tmp: float4, PH order after makeGroups */
  BT_BUF_PTCL_ACC = 3,     /* d_ptclAcc: float4 (ax,ay,az,phi), PH order */
  BT_BUF_DOMAIN = 4,       /* d_domain: float4 {cx, cy, cz, hsize} per system */
  BT_BUF_LEVEL_BEGIDX = 5, /* d_level_begIdx: int2[32] */
  BT_BUF_CELL_DATA = 6,    /* d_cellDataList: CellData uint4 */
  BT_BUF_CELL_SIZE = 7,    /* d_cellSize: float4 (com, cellOp^2) */
  BT_BUF_CELL_MONOPOLE = 8,/* d_cellMonopole: float4 (com, M) */
  BT_BUF_CELL_QUAD0 = 9,   /* d_cellQuad0: float4 (xx, yy, zz, xy) */
  BT_BUF_CELL_QUAD1 = 10,  /* d_cellQuad1: float2 (xz, yz) */
  BT_BUF_GROUP_LIST = 11,  /* d_groupList: int2 (pbeg, np) */
  BT_BUF_PH_KEY = 12,      /* sorted PH keys, uint64 (makeGroups' d_key) */
  BT_BUF_PH_PERM = 13      /* PH index -> tree index, int (makeGroups' d_value) */
};
int bt_get_device_buffer(bt_ctx *ctx, int which, void **ptr, int64_t *count);
/* Copies the first count elements of one of those arrays to host memory
 * (cuda_mem<T>::d2h, code/cudamem.h:40-46); synchronises the stream. */
int bt_copy_device_buffer(bt_ctx *ctx, int which, void *host, int64_t count);

/* ---- initial conditions and I/O (host only, no device needed) ---------- */

/* Plummer sphere (Aarseth-Henon-Wielen, code/plummer.h:37-117 algorithm)
 * with a counter-based RNG: radius truncated at rmax (structural units),
 * Henon units, m = 1/n, centre of mass removed in FP64.  Fills pos4/vel4. */
int bt_ic_plummer(int n, uint64_t seed, double rmax, float *pos4, float *vel4);
/* nclust Plummer subclusters: centres U[-box,box]^3, masses log-uniform in
 * [mmin, 1] normalised to 1, scale radius a0 * m^(1/3), particle count
 * proportional to mass (equal particle masses 1/n). */
int bt_ic_clustered(int n, int nclust, uint64_t seed, double box, double mmin, double a0,
                    float *pos4, float *vel4);
/* Plummer halo (mass mh) + exponential disk (mass 1-mh, scale length rd,
 * sech^2 scale height z0) on circular orbits. */
int bt_ic_halo_disk(int n, uint64_t seed, double mh, double rd, double z0, float *pos4,
                    float *vel4);
/* Tipsy snapshot reader (code/read_tipsy.h:84-168, tipsydefs.h:22-51).
 * With pos4 == null only *n is returned. */
int bt_read_tipsy(const char *path, int *n, float *pos4, float *vel4, int cap);
/* Tipsy snapshot writer (dark particles; eps written as given). */
int bt_write_tipsy(const char *path, int n, const float *pos4, const float *vel4, double time,
                   float eps);

#ifdef __cplusplus
}
#endif
#endif /* BONSAI2_AMD_H */
