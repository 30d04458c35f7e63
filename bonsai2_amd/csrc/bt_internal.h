// bt_internal.h -- context state and kernel launchers of libbonsai2_amd.so.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/bonsai2_amd.h"

namespace bt {

constexpr int kSortErr = 8;   // DevCounters::err bit: a sort pass's look-back wait timed out
constexpr int kLevels = 32;   // level table slots (reference int2[32], code/Treecode.h:171)
#ifndef BT_LEVEL_BLOCKS
#define BT_LEVEL_BLOCKS 2048
#endif
constexpr int kLevelBlocks = BT_LEVEL_BLOCKS;  // fixed grid of the per-level tree kernels
constexpr int kFp64Levels = 4;     // cells at levels <= this combine moments in FP64

struct MlpDev;   // matrix-core operands of the opening-angle network (bt_learn.h)

// Device buffers of one training-pair collection (bt_learn_collect).
struct LearnBuffers {
  int nsel = 0;
  int *gsel = nullptr;
  float theta_lo = 0.25f, theta_hi = 1.0f;
  int *count = nullptr;
  int max_pairs = 0;
  int2 *pairs = nullptr;
  double *tneed = nullptr;
  double *amag = nullptr;
  float *feat = nullptr, *target = nullptr, *err = nullptr, *tneed_f = nullptr;
  float *sink = nullptr;
};

// Device-resident counters of one context (one allocation).
// Blocks of the walk's statistics reduction; d_wstats holds the 4 results,
// then 4 words of partials per block.
constexpr int kWalkStatBlocks = 256;

// FP32 walks: stack entries per wave kept in LDS (bt_set_walk_lds_stack).
constexpr int kLdsStack = 1024;
#ifndef BT_LDS_STACK_MIN
#define BT_LDS_STACK_MIN 512   // the children one walk step can push (8 * 64)
#endif

struct DevCounters {
  int lvl_start[kLevels + 2];
  int lvl_count[kLevels + 2];
  int nnodes, nleaves, nlevels, ncells;
  int ngroups;
  int flags;   // BT_FLAG_*
  int err;     // 1 = cell capacity, 2 = group capacity, 4 = walk stack, 8 = sort
  int queue;   // walk group queue head
  int stack_max;  // deepest walk stack of the last walk (entries)
  int xq[8];      // walk group queues, one per XCD range
  // per-step check of the learned angle, decided on the device
  int check_calls;  // checks drawn so far (seeds the sample indices)
  int fallback;     // 1: the last check failed and the fixed-angle walk ran
  int checks, fallbacks;
  // Sticky state, cleared only by the host (launch_clear_sticky) at the start
  // of each API call that steps or walks: every error bit any phase raised
  // since (err is cleared by each phase for itself), the deepest walk stack of
  // every walk since, and the closing half kicks done since.  After an error
  // the kick and drift kernels do nothing, so a failed step leaves the
  // particles as after its first half kick and drift.
  int err_sticky;
  int stack_hw;
  int steps_done;
};

// The values captured graphs bake into their kernel arguments: a graph is
// replayed only while all of them are unchanged (compared bytewise).
struct GraphKey {
  double dt, eps, theta;
  int nleaf, level_split, ncrit;
  int opening, kahan, quadrupole, walk_variant, stack_ints, blocks_per_cu, lds_stack, group_mode;
  int overlap;
  float learn_shift, learn_theta_lo, learn_theta_hi;
  int check_ns;
  float check_p99;
  unsigned long long check_seed;
};

struct Ctx {
  int dev = 0;
  // Treecode<double>: particle and cell arrays hold double4 / double2 (twice
  // the float4 / float2 element count is allocated) and the step computes in
  // FP64; otherwise Treecode<float>
  bool fp64 = false;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  double eps = 0.01, theta = 0.75;   // a float context holds float values
  int ncrit = 64, nleaf = 16, level_split = 1;

  // sizes
  int n = 0, cap_n = 0, nsys = 1, max_sys_n = 0;
  std::vector<int> h_off;
  int *d_off = nullptr;

  // particles: "current" arrays hold input order until the first build, then
  // tree order.  *_2 are the reorder targets.
  float4 *pos = nullptr, *vel = nullptr, *pos2 = nullptr, *vel2 = nullptr;
  int *id = nullptr, *id2 = nullptr;

  // sort buffers
  uint64_t *keyB = nullptr;   // sort scratch
  int *valA = nullptr, *valB = nullptr;
  int *sort_hist = nullptr;   // 256 x blocks
  int sort_hist_cap = 0;
  // one-sweep passes: per-pass digit counts of every pass (from one histogram
  // launch), tile status words (two buffers, pass parity) and tile counters
  int *sort_hpart = nullptr;        // kSortHistBlocks x passes x 256 partial counts
  int *sort_gcount = nullptr;       // passes x 256 global digit counts
  unsigned int *sort_status = nullptr;  // 2 x tiles x 256
  int *sort_ctr = nullptr;          // 2 tile counters
  int sort_tiles_cap = 0;
  bool sort_onesweep = false;
  uint64_t *geo_keys = nullptr;   // sorted geometric keys (tree order)

  // scans
  int *scan_blk = nullptr;
  int scan_blk_cap = 0;

  // boxes
  float4 *box = nullptr;       // per system {cx, cy, cz, hsize}
  float4 *bbox_part = nullptr; // partial min/max
  int bbox_blocks = 1;

  // cells
  int cap_cells = 0;
  uint4 *cells = nullptr;
  float4 *csize = nullptr, *cmono = nullptr, *cq0 = nullptr;
  float2 *cq1 = nullptr, *cls = nullptr;
  float4 *cscr = nullptr;      // 6 float4 per cell (bottom-up moments)
  float4 *cwalk = nullptr;     // 3 float4 per cell: walk record (com,M) (3Q) (3Qxz,3Qyz,trQ/2,1.5trQ)
  int *child_bnd = nullptr;    // 8 ints per parent of the current level
  int *lvl_blk = nullptr;      // block sums / offsets of the level scan
  int2 *level_beg = nullptr;   // int2[32] (reference d_level_begIdx)
  int2 *root_cells = nullptr;  // per system (first level-1 cell, count)

  // PH order
  float4 *pos_ph = nullptr;
  int *perm = nullptr;         // PH index -> tree index
  int *id_ph = nullptr;        // PH index -> input index
  uint64_t *ph_keys = nullptr; // sorted PH keys
  float4 *acc = nullptr;       // (ax, ay, az, phi) in PH order

  // groups
  int cap_groups = 0;
  int2 *groups = nullptr;
  int *group_sys = nullptr;
  int *gtmp0 = nullptr, *gtmp1 = nullptr;  // N-sized scan scratch
  int2 *gcounts = nullptr;     // per group (n_approx, n_direct)
  double *gpot = nullptr;      // per group sum m*phi

  // walk
  int *stack = nullptr;
  size_t stack_alloc = 0;
  int stack_ints = 16384;
  int lds_stack = kLdsStack;   // FP32 walks: stack entries per wave kept in LDS
  int group_mode = 0;   // 0: PH-key groups (makeGroups); 1: leaves; 2: tree cells <= ncrit
  int walk_blocks = 0;         // persistent grid (blocks of 256)
  int blocks_per_cu = 0;
  int kahan = 1;
  int walk_variant = 0;        // 0: LDS j-data, 1: scalar loads, 2: two i-particles per lane
  int quadrupole = 1;          // 0: monopole-only P2C (the reference without QUADRUPOLE)
  int num_cus = 256;
  unsigned long long *d_wstats = nullptr;  // direct_sum, direct_max, approx_sum, approx_max
  // host copy of d_wstats fetched with the device counters; valid until the
  // next walk is queued
  unsigned long long h_wstats[4] = {};
  bool h_wstats_valid = false;
  double *d_wpot = nullptr;

  // learned opening angle
  float2 *sys_scale = nullptr;   // per system (root half size, total mass)
  float *mlp_params = nullptr;   // 1377 FP32 master parameters
  float *mlp_m = nullptr, *mlp_v = nullptr;  // Adam moments
  MlpDev *mlp_dev = nullptr;     // FP16 per-lane operands for inference
  bool have_mlp = false;
  int opening = 0;               // 0 fixed theta, 1 learned per-pair theta
  float learn_shift = 0.f, learn_theta_lo = 0.25f, learn_theta_hi = 1.0f;
  int check_ns = 0;              // FP64 samples re-checked after each learned walk
  float check_p99 = 1e-3f;       // above this p99 the step falls back to fixed theta
  unsigned long long check_seed = 1;
  // 4 * ns ints: the draws' indices into the particle arrays as they were when
  // drawn, the draws' input ids sorted, the draw slot of each sorted id, and
  // each draw's PH index (filled once the step's groups exist)
  int *d_check_idx = nullptr;
  double *d_check = nullptr;     // direct results (4 * ns) + {p50, p99, max}
  // A step sums the check's FP64 forces on stream3 beside the build, from the
  // particles as the drift left them (the tree-order gather writes the other
  // particle buffer).  check_state: 0 no sum queued for the coming check, 1
  // queued on the context stream, 2 queued on stream3 (joins the context
  // stream at chk_join_ev).  check_slices: source slices of the sum (fewer
  // slices leave the build kernels room on the CUs)
  hipStream_t stream3 = nullptr;
  hipEvent_t chk_fork_ev = nullptr, chk_join_ev = nullptr;
  int check_state = 0;
  int check_slices = 768;

  // energies / direct sum scratch
  double *d_energy = nullptr;  // 2 * nsys
  double *d_direct = nullptr;
  int direct_cap = 0;

  DevCounters *dc = nullptr;
  DevCounters hc;

  // moveParticles' half kick still owed: the next computeForces closes the
  // kick-drift-kick step with it (0 = none)
  double pending_kick = 0.0;
  int check_cap = 0;             // samples d_check / d_check_idx are sized for

  // state
  bool have_particles = false, have_tree = false, have_mp = false, have_groups = false,
       have_acc = false;

  // captured step (bt_step_graph).  The tree-order gather writes the particles
  // into the other of two buffers, so there is one graph per buffer holding
  // the particles at the start of the step: [0] for graph_pos0, [1] for the other.
  hipGraph_t graph[2] = {nullptr, nullptr};
  hipGraphExec_t graph_exec[2] = {nullptr, nullptr};
  float4 *graph_pos0 = nullptr;
  GraphKey graph_key{};   // every by-value launch parameter the graphs hold
  int tree_version = 0;

  // timing
  hipEvent_t ev[12] = {};
  // A step runs the groups (PH keys, PH sort, groups) on stream2 beside the
  // level build and multipoles on stream: both need only the tree-order
  // particles and sorted keys (overlap = false runs them in sequence on
  // stream).  forked: the last step's phase events follow that layout
  // (multipoles end at ev[10], groups at ev[11]).
  hipStream_t stream2 = nullptr;
  hipEvent_t fork_ev = nullptr, join_ev = nullptr;
  bool overlap = true;
  bool forked = false;
  float ms[10] = {};
  bool timing = true;

  std::string err;
};

// ---- launchers (each queues work on c.stream) --------------------------
void launch_bbox(Ctx &c);
void launch_geo_keys(Ctx &c);
void launch_gather_tree(Ctx &c);
void launch_levels(Ctx &c);
void launch_multipoles(Ctx &c);

void radix_sort_pairs(Ctx &c, uint64_t *keys, int *vals, uint64_t *keys_tmp, int *vals_tmp,
                      int n, int key_bits, bool by_system_after);
void scan_exclusive_sum(Ctx &c, const int *in, int *out, int n, int *total_dev);
void scan_inclusive_max(Ctx &c, const int *in, int *out, int n);
void fill_int(Ctx &c, int *p, int value, int n);

void launch_ph_keys(Ctx &c);
void launch_gather_ph(Ctx &c);
void launch_groups(Ctx &c);
// groups = tree cells (mode 1: the leaves; mode 2: cells of <= ncrit
// particles), in tree order; the group order is the tree order
void launch_leaf_groups(Ctx &c);

// run_if: device flag; the walk (reset, traversal, statistics) does nothing
// unless it is nonzero (nullptr = always run)
void launch_walk(Ctx &c, const int *run_if = nullptr);

void launch_mlp_prepare(Ctx &c);
void launch_mlp_forward(Ctx &c, int n, const float *x, float *y);
void launch_sys_scale(Ctx &c);
void launch_learn_pairs(Ctx &c, LearnBuffers &L);
void launch_learn_errors(Ctx &c, LearnBuffers &L, int npairs);
void launch_learn_amag(Ctx &c, int ns, const int *d_idx, const double *d_direct, double *amag);
int mlp_grad_blocks();
void launch_mlp_grad(Ctx &c, const float *params, int n, const float *feat, const float *target,
                     int batch, unsigned long long seed, int step, float *partial, double *grad);
void launch_mlp_adam(Ctx &c, float *p, float *m, float *v, const double *grad, float lr, int step);
// targets of the next check, drawn on the device from check_seed and the
// device's check count: indices into the current particle arrays (c.pos,
// c.id), and their input ids sorted for launch_check_map
void launch_check_idx(Ctx &c, int ns);
// each target's PH index (after the step's groups), from c.id_ph
void launch_check_map(Ctx &c, int ns);
// p50/p99/max of the relative force errors into out[0..2]; sets the device's
// fallback flag (p99 > check_p99) and counts checks and fallbacks
void launch_learn_check(Ctx &c, int ns, const int *d_idx, const double *d_direct, double *out);

// closing: the second half kick of a step (counted in steps_done)
void launch_kick(Ctx &c, double dt_half, bool closing = false);
// the opening half kick (dth) and the drift (dt) of a step, in one pass
void launch_kick_drift(Ctx &c, double dth, double dt);
void launch_clear_sticky(Ctx &c);
void launch_energies(Ctx &c);
int direct_slices(int ns);   // d_out of launch_direct: ns * (slices + 1) double4
void launch_direct(Ctx &c, int ns, const int *d_idx, double *d_out, bool check = false);
// src: the particles targets and sources are read from (null: c.pos_ph)
void launch_direct_parts(Ctx &c, int ns, const int *d_idx, double *d_out, double *d_part,
                         int slices, bool check = false, const void *src = nullptr);

// host-side helpers
int grid_for(int n, int block, int cap);

}  // namespace bt
