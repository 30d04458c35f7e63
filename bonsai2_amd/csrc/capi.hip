// capi.hip -- the C ABI of libbonsai2_amd.so (include/bonsai2_amd.h).
//
// Host orchestration of the gravity step.  Each exported function replaces
// one member of the reference's Treecode<real_t> (code/Treecode.h:123-223);
// unlike the reference (process-global __device__ symbols, exit() on error,
// code/cutil.h:772-780) all state is per context and errors are returned.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "bt_internal.h"
#include "bt_learn.h"

using namespace bt;

struct bt_ctx {
  Ctx c;
};

namespace {

#define BT_CK(call)                                                                  \
  do {                                                                               \
    const hipError_t e_ = (call);                                                    \
    if (e_ != hipSuccess) {                                                          \
      c.err = std::string(#call) + ": " + hipGetErrorString(e_);                     \
      return BT_ERR_HIP;                                                             \
    }                                                                                \
  } while (0)

#define BT_LAUNCH_CK()                                                               \
  do {                                                                               \
    const hipError_t e_ = hipGetLastError();                                         \
    if (e_ != hipSuccess) {                                                          \
      c.err = std::string("kernel launch: ") + hipGetErrorString(e_);                \
      return BT_ERR_HIP;                                                             \
    }                                                                                \
    if (!c.err.empty()) return BT_ERR_HIP;                                           \
  } while (0)

int fail(Ctx &c, int code, const char *msg) {
  c.err = msg;
  return code;
}

template <typename T>
int dev_alloc(Ctx &c, T **p, size_t count) {
  if (*p) {
    (void)hipFree(*p);
    *p = nullptr;
  }
  if (count == 0) return 0;
  const hipError_t e = hipMalloc((void **)p, sizeof(T) * count);
  if (e != hipSuccess) {
    *p = nullptr;
    c.err = std::string("hipMalloc failed: ") + hipGetErrorString(e);
    return BT_ERR_HIP;
  }
  return 0;
}

template <typename T>
void dev_free(T **p) {
  if (*p) (void)hipFree(*p);
  *p = nullptr;
}

void free_particle_arrays(Ctx &c) {
  dev_free(&c.pos); dev_free(&c.vel); dev_free(&c.pos2); dev_free(&c.vel2);
  dev_free(&c.id); dev_free(&c.id2);
  dev_free(&c.keyB); dev_free(&c.valA); dev_free(&c.valB); dev_free(&c.geo_keys);
  dev_free(&c.pos_ph); dev_free(&c.perm); dev_free(&c.id_ph); dev_free(&c.ph_keys);
  dev_free(&c.acc); dev_free(&c.groups); dev_free(&c.group_sys); dev_free(&c.gtmp0);
  dev_free(&c.gtmp1); dev_free(&c.gcounts); dev_free(&c.gpot);
  dev_free(&c.box); dev_free(&c.bbox_part); dev_free(&c.root_cells); dev_free(&c.d_off);
  dev_free(&c.d_energy); dev_free(&c.sys_scale);
}

void free_cell_arrays(Ctx &c) {
  dev_free(&c.cells); dev_free(&c.csize); dev_free(&c.cmono); dev_free(&c.cq0);
  dev_free(&c.cq1); dev_free(&c.cls); dev_free(&c.cscr); dev_free(&c.child_bnd);
  dev_free(&c.cwalk);
}

// float4 / float2 elements per value of the context's real type (2 for
// Treecode<double>, whose arrays hold double4 / double2)
size_t rw(const Ctx &c) { return c.fp64 ? 2 : 1; }

int alloc_cells(Ctx &c, int cap) {
  int rc = 0;
  c.cap_cells = cap;
  const size_t w = rw(c), n = (size_t)cap;
  rc |= dev_alloc(c, &c.cells, n);
  rc |= dev_alloc(c, &c.csize, n * w);
  rc |= dev_alloc(c, &c.cmono, n * w);
  rc |= dev_alloc(c, &c.cq0, n * w);
  rc |= dev_alloc(c, &c.cq1, n * w);
  rc |= dev_alloc(c, &c.cls, n * w);
  rc |= dev_alloc(c, &c.cscr, n * 6 * w);
  rc |= dev_alloc(c, &c.cwalk, n * 3 * w);
  rc |= dev_alloc(c, &c.child_bnd, (size_t)cap * 8 + 8 * (size_t)c.nsys);
  return rc ? BT_ERR_HIP : 0;
}

// Waits for all of the context's work, including work a step forked onto its
// second and third streams (joined into the context stream unless the call
// that queued it failed), before buffers are freed or replaced.
void sync_streams(Ctx &c) {
  if (c.stream2) (void)hipStreamSynchronize(c.stream2);
  if (c.stream3) (void)hipStreamSynchronize(c.stream3);
  (void)hipStreamSynchronize(c.stream);
}

void release_graph(Ctx &c) {
  for (int k = 0; k < 2; k++) {
    if (c.graph_exec[k]) (void)hipGraphExecDestroy(c.graph_exec[k]);
    if (c.graph[k]) (void)hipGraphDestroy(c.graph[k]);
    c.graph_exec[k] = nullptr;
    c.graph[k] = nullptr;
  }
  c.graph_pos0 = nullptr;
}

// The host-side effect of the tree-order gather: the particle arrays and
// their reorder targets trade places.
void swap_particle_buffers(Ctx &c) {
  std::swap(c.pos, c.pos2);
  std::swap(c.vel, c.vel2);
  std::swap(c.id, c.id2);
}

int ensure_stack(Ctx &c) {
  const int bpc = c.blocks_per_cu > 0 ? c.blocks_per_cu : 8;
  c.walk_blocks = std::max(1, c.num_cus * bpc);
  const size_t need = (size_t)c.walk_blocks * 4 * (size_t)c.stack_ints;
  if (need > c.stack_alloc) {
    release_graph(c);
    if (dev_alloc(c, &c.stack, need)) return BT_ERR_HIP;
    c.stack_alloc = need;
  }
  return 0;
}

void record(Ctx &c, int k) {
  if (c.timing) (void)hipEventRecord(c.ev[k], c.stream);
}

// Device-to-host copy ordered after the work queued on the context stream
// (the stream is non-blocking: a plain hipMemcpy would not wait for it).
hipError_t d2h(Ctx &c, void *dst, const void *src, size_t bytes) {
  const hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c.stream);
  return e != hipSuccess ? e : hipStreamSynchronize(c.stream);
}

int read_counters(Ctx &c) {
  BT_CK(hipMemcpyAsync(&c.hc, c.dc, sizeof(DevCounters), hipMemcpyDeviceToHost, c.stream));
  if (c.d_wstats)   // the last walk's statistics ride along (one synchronisation)
    BT_CK(hipMemcpyAsync(c.h_wstats, c.d_wstats, sizeof(c.h_wstats), hipMemcpyDeviceToHost,
                         c.stream));
  BT_CK(hipStreamSynchronize(c.stream));
  c.h_wstats_valid = c.d_wstats != nullptr;
  return 0;
}

// ---- phases (asynchronous; no host synchronisation) ----------------------

// Box, geometric keys, their sort and the tree-order gather.
void phase_sort(Ctx &c) {
  record(c, 0);
  launch_bbox(c);
  launch_geo_keys(c);
  record(c, 1);
  radix_sort_pairs(c, c.geo_keys, c.valA, c.keyB, c.valB, c.n, 63, true);
  record(c, 2);
  launch_gather_tree(c);
}

int phase_build(Ctx &c) {
  c.forked = false;
  phase_sort(c);
  launch_levels(c);
  record(c, 3);
  BT_LAUNCH_CK();
  return 0;
}

// end_ev: the event recorded when the phase's work is queued.
int phase_multipoles(Ctx &c, int end_ev = 4) {
  launch_multipoles(c);
  if (!c.fp64) launch_sys_scale(c);   // the learned angle's scales (Treecode<float> only)
  record(c, end_ev);
  BT_LAUNCH_CK();
  return 0;
}

int phase_groups(Ctx &c, int end_ev = 7) {
  if (c.group_mode != 0) {
    record(c, 5);
    record(c, 6);
    launch_leaf_groups(c);
    record(c, end_ev);
    BT_LAUNCH_CK();
    return 0;
  }
  launch_ph_keys(c);
  record(c, 5);
  radix_sort_pairs(c, c.ph_keys, c.perm, c.keyB, c.valB, c.n, 63, true);
  record(c, 6);
  launch_gather_ph(c);
  launch_groups(c);
  record(c, end_ev);
  BT_LAUNCH_CK();
  return 0;
}

// The tree and the groups of a step, the groups on stream2 beside the level
// build and the multipoles: after the tree-order gather, the groups read
// only the particles, the root box and their own scratch (PH keys, the sort
// buffers, the group arrays), the levels and multipoles only the sorted
// geometric keys, the particles and the cell arrays, and the device counters
// they share are updated with atomics.  The walk waits for both.
int phase_tree_and_groups(Ctx &c) {
  phase_sort(c);
  c.forked = true;
  BT_CK(hipEventRecord(c.fork_ev, c.stream));
  BT_CK(hipStreamWaitEvent(c.stream2, c.fork_ev, 0));
  std::swap(c.stream, c.stream2);   // the launches below go to stream2
  record(c, 4);
  const int rc_groups = phase_groups(c, 11);
  const hipError_t e_join = hipEventRecord(c.join_ev, c.stream);
  std::swap(c.stream, c.stream2);
  launch_levels(c);
  record(c, 3);
  const int rc_mp = phase_multipoles(c, 10);
  // join before any return, so that a capture never ends with stream2 forked
  if (e_join == hipSuccess) BT_CK(hipStreamWaitEvent(c.stream, c.join_ev, 0));
  record(c, 7);
  if (e_join != hipSuccess) {
    c.err = std::string("hipEventRecord: ") + hipGetErrorString(e_join);
    return BT_ERR_HIP;
  }
  return rc_groups ? rc_groups : rc_mp;
}

int phase_forces(Ctx &c) {
  if (ensure_stack(c)) return BT_ERR_HIP;
  launch_walk(c);
  record(c, 8);
  BT_LAUNCH_CK();
  return 0;
}

const char *const kSortErrMsg = "radix sort: a pass's look-back wait timed out";
const char *const kGroupCapacityMsg =
    "group capacity exceeded (leaf groups: a leaf holds more than 64 particles)";

static uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// {p50, p99, max} of the last check, after the direct-sum buffer of check_ns
// samples (every ns <= 1024 uses the same number of source slices).
double *check_result(Ctx &c) {
  return c.d_check + (size_t)c.check_ns * (direct_slices(c.check_ns) + 1) * 4;
}

// After a learned-angle walk: the relative force errors of check_ns seeded
// samples against the FP64 direct sum, and the decision on the device: if
// their p99 exceeds check_p99 the fixed-angle walk (the fallback) recomputes
// the forces, otherwise its launches return at once.  No host
// synchronisation, so a captured step holds the check too.
bool check_active(const Ctx &c) { return c.opening == 1 && c.have_mlp && c.check_ns > 0; }

// Makes the context stream wait for a check sum queued on stream3.
int join_check(Ctx &c) {
  if (c.check_state == 2) {
    c.check_state = 1;
    BT_CK(hipStreamWaitEvent(c.stream, c.chk_join_ev, 0));
  }
  return 0;
}

// The learned walk's check, first part: draw the targets among the particle
// arrays as they are now and sum their FP64 forces.  The step queues it
// before its build, with fork = true on stream3 beside the build (the
// tree-order gather writes the other particle buffer, so the arrays read
// here stay unchanged until the next step's gather, which follows the join).
int queue_check_direct(Ctx &c, bool fork) {
  if (join_check(c)) return BT_ERR_HIP;   // one left forked by a call that failed
  c.check_state = 0;
  if (!check_active(c)) return 0;
  const int ns = std::min(c.check_ns, c.n);
  const int slices = std::max(1, std::min(c.check_slices, direct_slices(ns)));
  if (fork) {
    BT_CK(hipEventRecord(c.chk_fork_ev, c.stream));
    BT_CK(hipStreamWaitEvent(c.stream3, c.chk_fork_ev, 0));
    std::swap(c.stream, c.stream3);   // the launches below go to stream3
  }
  launch_check_idx(c, ns);
  launch_direct_parts(c, ns, c.d_check_idx, c.d_check, c.d_check + (size_t)4 * ns, slices, true,
                      c.pos);
  if (fork) {
    const hipError_t e = hipEventRecord(c.chk_join_ev, c.stream);
    std::swap(c.stream, c.stream3);
    if (e != hipSuccess) {
      (void)hipStreamSynchronize(c.stream3);   // nothing left running unjoined
      c.err = std::string("hipEventRecord: ") + hipGetErrorString(e);
      return BT_ERR_HIP;
    }
  }
  c.check_state = fork ? 2 : 1;
  return 0;
}

// The check, second part (after the walk): each target's PH index, the
// relative errors of the walk's forces, the decision, and the fixed-angle
// walk, which runs only when the check failed.
int launch_learned_check(Ctx &c) {
  if (!check_active(c)) return 0;
  if (c.check_state == 0 && queue_check_direct(c, false)) return BT_ERR_HIP;
  const int ns = std::min(c.check_ns, c.n);
  launch_check_map(c, ns);
  if (join_check(c)) return BT_ERR_HIP;
  c.check_state = 0;
  launch_learn_check(c, ns, c.d_check_idx + (size_t)3 * ns, c.d_check, check_result(c));
  c.opening = 0;
  launch_walk(c, &c.dc->fallback);
  c.opening = 1;
  return 0;
}

int check_build_errors(Ctx &c) {
  if (read_counters(c)) return BT_ERR_HIP;
  if (c.hc.err & kSortErr) return fail(c, BT_ERR_HIP, kSortErrMsg);
  if (c.hc.err & 1) return fail(c, BT_ERR_CAPACITY, "cell capacity exceeded");
  return 0;
}

// Grows the cell arrays to their largest size (3n + 64 per system + 1024:
// every level adds at most n cells, and a cell holds at least one particle
// unless it is a node, whose children hold more than NLEAF).  Captured graphs
// hold the old arrays and are dropped.
int regrow_cells(Ctx &c) {
  const long long cap = 3LL * c.n + 64LL * c.nsys + 1024;
  if (cap > 0x1FFFFFFF || cap <= c.cap_cells)
    return fail(c, BT_ERR_CAPACITY, "cell capacity exceeded");
  sync_streams(c);
  release_graph(c);
  free_cell_arrays(c);
  if (alloc_cells(c, (int)cap)) return BT_ERR_HIP;
  return 0;
}

// Build, growing the cell arrays once if they were too small.
int build_checked(Ctx &c) {
  if (phase_build(c)) return BT_ERR_HIP;
  if (read_counters(c)) return BT_ERR_HIP;
  if (c.hc.err & kSortErr) return fail(c, BT_ERR_HIP, kSortErrMsg);
  if (c.hc.err & 1) {
    const int rc = regrow_cells(c);
    if (rc) return rc;
    launch_levels(c);
    BT_LAUNCH_CK();
    if (check_build_errors(c)) return BT_ERR_CAPACITY;
  }
  return 0;
}

// The second half of a kick-drift-kick step, after the drift: the forces at
// the new positions and the closing half kick.  Queued without host
// synchronisation.
int step_second_half(Ctx &c, double dt) {
  if (queue_check_direct(c, c.overlap)) return BT_ERR_HIP;
  if (c.overlap) {
    if (phase_tree_and_groups(c)) return BT_ERR_HIP;
  } else if (phase_build(c) || phase_multipoles(c) || phase_groups(c)) {
    return BT_ERR_HIP;
  }
  if (phase_forces(c) || launch_learned_check(c)) return BT_ERR_HIP;
  launch_kick(c, 0.5 * dt, true);
  record(c, 9);
  BT_LAUNCH_CK();
  return 0;
}

// After the queued second half of a step: on a cell-capacity failure (the
// walk and the closing kick were skipped, the particles hold the step's first
// half kick and drift) grow the cells once and run the second half again.
int finish_step(Ctx &c, double dt) {
  if (read_counters(c)) return BT_ERR_HIP;
  if (c.hc.err_sticky == 1) {
    int rc = regrow_cells(c);
    if (rc) return rc;
    launch_clear_sticky(c);
    rc = step_second_half(c, dt);
    if (rc) return rc;
    if (read_counters(c)) return BT_ERR_HIP;
  }
  if (c.hc.err_sticky & kSortErr) return fail(c, BT_ERR_HIP, kSortErrMsg);
  if (c.hc.err_sticky & 1) return fail(c, BT_ERR_CAPACITY, "cell capacity exceeded");
  if (c.hc.err_sticky & 2) return fail(c, BT_ERR_CAPACITY, kGroupCapacityMsg);
  if (c.hc.err_sticky & 4) return fail(c, BT_ERR_STACK, "walk stack overflow");
  return 0;
}

int finish_timings(Ctx &c, int last) {
  if (!c.timing) return 0;
  for (int k = 0; k < 10; k++) c.ms[k] = 0.f;
  // phases: 0 bbox+keys, 1 sort, 2 levels, 3 multipoles, 4 ph keys, 5 ph sort, 6 groups,
  // 7 walk, 8 integrate, 9 energies
  // phase k runs from event from[k] to event to[k]; after a forked step the
  // multipoles end at ev[10] and the groups (on stream2) at ev[11]
  for (int k = 0; k < last && k < 10; k++) {
    int to = k + 1;
    if (c.forked && k == 3) to = 10;
    if (c.forked && k == 6) to = 11;
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, c.ev[k], c.ev[to]) == hipSuccess) c.ms[k] = ms;
  }
  return 0;
}

bool valid_nleaf(int nleaf) {
  return nleaf == 16 || nleaf == 24 || nleaf == 32 || nleaf == 48 || nleaf == 64;
}

}  // namespace

extern "C" {

}  // extern "C"

namespace {

// ---- the device scratch of bt_direct_fp64_noalloc ------------------------
// The partial sums of that entry point live in this fixed array, one copy per
// GPU, held until the process ends.  Host code never passes the array's name
// to a kernel (on the host it is only a stand-in for the symbol): its device
// address is looked up once per device with hipGetSymbolAddress, at the first
// context created on that device, and kept in scratch_dev[].
constexpr size_t kScratchBytes = size_t(16) << 20;
constexpr size_t kScratchD4 = kScratchBytes / sizeof(double4);
constexpr int kNoAllocMaxTargets = 4096;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) double4 scratch[kScratchD4];
std::mutex scratch_mu;
double4 *scratch_dev[kMaxDevices];

// With device `dev` current: its copy of the array, looked up on first use.
int scratch_lookup(int dev, double4 **out) {
  *out = nullptr;
  if (dev < 0 || dev >= kMaxDevices) return BT_ERR_ARG;
  std::lock_guard<std::mutex> lk(scratch_mu);
  if (!scratch_dev[dev]) {
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return BT_ERR_HIP;
    scratch_dev[dev] = static_cast<double4 *>(p);
  }
  *out = scratch_dev[dev];
  return 0;
}

int create_ctx(bt_ctx **out, double eps, double theta, int ncrit, bool fp64) {
  if (!out) return BT_ERR_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return BT_ERR_NODEVICE;
  bt_ctx *h = new bt_ctx();
  Ctx &c = h->c;
  c.fp64 = fp64;
  c.eps = eps;
  c.theta = theta;
  c.ncrit = ncrit;
  if (hipGetDevice(&c.dev) != hipSuccess) c.dev = 0;
  (void)hipSetDevice(c.dev);
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c.dev) == hipSuccess &&
      cus > 0)
    c.num_cus = cus;
  if (hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking) != hipSuccess) {
    delete h;
    return BT_ERR_HIP;
  }
  c.own_stream = true;
  if (hipStreamCreateWithFlags(&c.stream2, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&c.stream3, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&c.fork_ev, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c.join_ev, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c.chk_fork_ev, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c.chk_join_ev, hipEventDisableTiming) != hipSuccess) {
    bt_destroy(h);
    return BT_ERR_HIP;
  }
  // BT_SORT_ONESWEEP=1 selects the one-sweep radix passes (one launch per pass,
  // decoupled look-back) instead of the three-launch passes (histogram, scan,
  // scatter): same result, measured slower on MI355X, kept for comparison
  if (const char *e = std::getenv("BT_SORT_ONESWEEP")) c.sort_onesweep = std::atoi(e) != 0;
  // BT_CHECK_SLICES: source slices of the learned walk's per-step FP64 check
  // (1..2048; a tuning knob, the results differ only in FP64 rounding)
  if (const char *e = std::getenv("BT_CHECK_SLICES"))
    c.check_slices = std::max(1, std::min(2048, std::atoi(e)));
  for (int k = 0; k < 12; k++) (void)hipEventCreate(&c.ev[k]);
  if (dev_alloc(c, &c.dc, 1) || dev_alloc(c, &c.level_beg, kLevels) ||
      dev_alloc(c, &c.d_wstats, 4 + 4 * (size_t)kWalkStatBlocks) || dev_alloc(c, &c.d_wpot, 1) ||
      dev_alloc(c, &c.lvl_blk, kLevelBlocks) || dev_alloc(c, &c.mlp_params, kMlpParams) ||
      dev_alloc(c, &c.mlp_m, kMlpParams) || dev_alloc(c, &c.mlp_v, kMlpParams) ||
      dev_alloc(c, &c.mlp_dev, 1)) {
    delete h;
    return BT_ERR_HIP;
  }
  double4 *scr = nullptr;
  if (scratch_lookup(c.dev, &scr)) {
    bt_destroy(h);
    return BT_ERR_HIP;
  }
  std::memset(&c.hc, 0, sizeof(c.hc));
  (void)hipMemcpy(c.dc, &c.hc, sizeof(DevCounters), hipMemcpyHostToDevice);
  std::vector<int2> z(kLevels, make_int2(0, 0));
  (void)hipMemcpy(c.level_beg, z.data(), sizeof(int2) * kLevels, hipMemcpyHostToDevice);
  *out = h;
  return 0;
}

// The real type of a data entry point must match the context's.
template <typename T>
int check_real(Ctx &c) {
  if (c.fp64 == (sizeof(T) == 8)) return 0;
  return fail(c, BT_ERR_STATE, c.fp64 ? "Treecode<double> context: use the _f64 entry point"
                                      : "Treecode<float> context: use the float entry point");
}

}  // namespace

extern "C" {

int bt_create(bt_ctx **out, float eps, float theta, int ncrit) {
  return create_ctx(out, eps, theta, ncrit, false);
}

int bt_create_f64(bt_ctx **out, double eps, double theta, int ncrit) {
  return create_ctx(out, eps, theta, ncrit, true);
}

int bt_destroy(bt_ctx *h) {
  if (!h) return 0;
  Ctx &c = h->c;
  sync_streams(c);
  release_graph(c);
  free_particle_arrays(c);
  free_cell_arrays(c);
  dev_free(&c.stack);
  dev_free(&c.dc);
  dev_free(&c.level_beg);
  dev_free(&c.d_wstats);
  dev_free(&c.d_wpot);
  dev_free(&c.sort_hist);
  dev_free(&c.sort_hpart); dev_free(&c.sort_gcount); dev_free(&c.sort_status);
  dev_free(&c.sort_ctr);
  dev_free(&c.scan_blk);
  dev_free(&c.lvl_blk);
  dev_free(&c.d_direct);
  dev_free(&c.mlp_params); dev_free(&c.mlp_m); dev_free(&c.mlp_v); dev_free(&c.mlp_dev);
  dev_free(&c.d_check); dev_free(&c.d_check_idx);
  for (int k = 0; k < 12; k++)
    if (c.ev[k]) (void)hipEventDestroy(c.ev[k]);
  if (c.own_stream && c.stream) (void)hipStreamDestroy(c.stream);
  if (c.stream2) (void)hipStreamDestroy(c.stream2);
  if (c.stream3) (void)hipStreamDestroy(c.stream3);
  if (c.fork_ev) (void)hipEventDestroy(c.fork_ev);
  if (c.join_ev) (void)hipEventDestroy(c.join_ev);
  if (c.chk_fork_ev) (void)hipEventDestroy(c.chk_fork_ev);
  if (c.chk_join_ev) (void)hipEventDestroy(c.chk_join_ev);
  delete h;
  return 0;
}

const char *bt_last_error(const bt_ctx *h) { return h ? h->c.err.c_str() : "null context"; }

int bt_set_stream(bt_ctx *h, void *s) {
  if (!h) return BT_ERR_ARG;
  Ctx &c = h->c;
  sync_streams(c);
  if (c.own_stream && c.stream) (void)hipStreamDestroy(c.stream);
  c.own_stream = false;
  c.stream = (hipStream_t)s;
  return 0;
}

int bt_set_params(bt_ctx *h, float eps, float theta) { return bt_set_params_f64(h, eps, theta); }

int bt_set_params_f64(bt_ctx *h, double eps, double theta) {
  if (!h) return BT_ERR_ARG;
  Ctx &c = h->c;
  if (!c.fp64) {   // a float context holds float values
    eps = (float)eps;
    theta = (float)theta;
  }
  c.eps = eps;
  if (c.theta != theta) c.have_mp = false;  // theta is baked into cellOp^2
  c.theta = theta;
  return 0;
}

int bt_set_walk_options(bt_ctx *h, int kahan, int stack_ints, int blocks_per_cu) {
  if (!h) return BT_ERR_ARG;
  Ctx &c = h->c;
  // the stack must hold the level-1 cells a walk starts from (<= 8); every
  // push beyond its capacity is caught (BT_ERR_STACK)
  if (stack_ints < 64 || blocks_per_cu < 0 || blocks_per_cu > 64)
    return fail(c, BT_ERR_ARG, "bad walk options");
  c.kahan = kahan ? 1 : 0;
  c.stack_ints = stack_ints;
  c.blocks_per_cu = blocks_per_cu;
  return 0;
}

int bt_set_walk_lds_stack(bt_ctx *h, int entries) {
  if (!h) return BT_ERR_ARG;
  if (entries < BT_LDS_STACK_MIN || entries > kLdsStack)
    return fail(h->c, BT_ERR_ARG, "LDS stack entries out of range");
  h->c.lds_stack = entries;
  return 0;
}

int bt_set_group_mode(bt_ctx *h, int mode) {
  if (!h) return BT_ERR_ARG;
  if (mode < 0 || mode > 2) return fail(h->c, BT_ERR_ARG, "group mode must be 0, 1 or 2");
  if (h->c.group_mode != mode) h->c.have_groups = h->c.have_acc = false;
  h->c.group_mode = mode;
  return 0;
}

int bt_get_walk_stack_max(bt_ctx *h, int *entries) {
  if (!h || !entries) return BT_ERR_ARG;
  Ctx &c = h->c;
  if (c.dc && read_counters(c)) return BT_ERR_HIP;
  *entries = c.hc.stack_hw;
  return 0;
}

int bt_set_walk_variant(bt_ctx *h, int variant) {
  if (!h) return BT_ERR_ARG;
  if (variant < 0 || variant > 2) return fail(h->c, BT_ERR_ARG, "walk variant must be 0, 1 or 2");
  h->c.walk_variant = variant;
  return 0;
}

int bt_set_overlap(bt_ctx *h, int on) {
  if (!h) return BT_ERR_ARG;
  if (on != 0 && on != 1) return fail(h->c, BT_ERR_ARG, "overlap must be 0 or 1");
  h->c.overlap = on != 0;
  return 0;
}

int bt_set_quadrupole(bt_ctx *h, int on) {
  if (!h) return BT_ERR_ARG;
  Ctx &c = h->c;
  if (on != 0 && on != 1) return fail(c, BT_ERR_ARG, "quadrupole must be 0 or 1");
  if (!on && c.opening == 1) return fail(c, BT_ERR_STATE, "the learned angle needs quadrupoles");
  c.quadrupole = on;
  return 0;
}

int bt_alloc(bt_ctx *h, int n, int nsys, const int *offsets) {
  if (!h) return BT_ERR_ARG;
  Ctx &c = h->c;
  c.err.clear();
  if (n <= 0 || nsys <= 0 || nsys > n) return fail(c, BT_ERR_ARG, "bad particle/system count");
  std::vector<int> off(nsys + 1);
  if (offsets) {
    for (int k = 0; k <= nsys; k++) off[k] = offsets[k];
    if (off[0] != 0 || off[nsys] != n) return fail(c, BT_ERR_ARG, "offsets must span [0, n)");
    for (int k = 0; k < nsys; k++)
      if (off[k + 1] <= off[k]) return fail(c, BT_ERR_ARG, "empty or unordered system");
  } else {
    if (nsys != 1) return fail(c, BT_ERR_ARG, "offsets required for nsys > 1");
    off[0] = 0;
    off[1] = n;
  }
  sync_streams(c);
  release_graph(c);
  const bool regrow = n > c.cap_n || nsys != c.nsys || !c.pos;
  c.n = n;
  c.nsys = nsys;
  c.h_off = off;
  c.max_sys_n = 0;
  for (int k = 0; k < nsys; k++) c.max_sys_n = std::max(c.max_sys_n, off[k + 1] - off[k]);
  if (regrow) {
    free_particle_arrays(c);
    free_cell_arrays(c);
    c.cap_n = n;
    int rc = 0;
    const size_t w = rw(c), nw = (size_t)n * w;
    rc |= dev_alloc(c, &c.pos, nw); rc |= dev_alloc(c, &c.vel, nw);
    rc |= dev_alloc(c, &c.pos2, nw); rc |= dev_alloc(c, &c.vel2, nw);
    rc |= dev_alloc(c, &c.id, n); rc |= dev_alloc(c, &c.id2, n);
    rc |= dev_alloc(c, &c.keyB, n); rc |= dev_alloc(c, &c.valA, n);
    rc |= dev_alloc(c, &c.valB, n); rc |= dev_alloc(c, &c.geo_keys, n);
    rc |= dev_alloc(c, &c.pos_ph, nw); rc |= dev_alloc(c, &c.perm, n);
    rc |= dev_alloc(c, &c.id_ph, n); rc |= dev_alloc(c, &c.ph_keys, n);
    rc |= dev_alloc(c, &c.acc, ((size_t)n + 64) * w);  // + spare slots for idle walk lanes
    c.cap_groups = n;
    rc |= dev_alloc(c, &c.groups, n); rc |= dev_alloc(c, &c.group_sys, n);
    rc |= dev_alloc(c, &c.gtmp0, n); rc |= dev_alloc(c, &c.gtmp1, n);
    rc |= dev_alloc(c, &c.gcounts, n); rc |= dev_alloc(c, &c.gpot, n);
    rc |= dev_alloc(c, &c.box, (size_t)nsys * w); rc |= dev_alloc(c, &c.root_cells, nsys);
    rc |= dev_alloc(c, &c.sys_scale, nsys);
    rc |= dev_alloc(c, &c.d_off, nsys + 1);
    rc |= dev_alloc(c, &c.d_energy, 2 * (size_t)nsys * 65);
    if (rc) return BT_ERR_HIP;
    if (alloc_cells(c, n + 1024 * nsys + 64)) return BT_ERR_HIP;
  }
  c.bbox_blocks = std::max(1, std::min(1024, (c.max_sys_n + 16383) / 16384));
  if (dev_alloc(c, &c.bbox_part, 2 * (size_t)nsys * c.bbox_blocks * rw(c))) return BT_ERR_HIP;
  BT_CK(hipMemcpy(c.d_off, off.data(), sizeof(int) * (nsys + 1), hipMemcpyHostToDevice));
  c.have_particles = c.have_tree = c.have_mp = c.have_groups = c.have_acc = false;
  return 0;
}

}  // extern "C"

namespace {

// Copies n particles in (input order), from host or device memory; bytes =
// n values of the context's 4-vector type.
int put_particles(Ctx &c, const void *pos4, const void *vel4, hipMemcpyKind kind) {
  if (c.n <= 0) return fail(c, BT_ERR_STATE, "bt_alloc first");
  const size_t bytes = sizeof(float4) * rw(c) * (size_t)c.n;
  BT_CK(hipMemcpyAsync(c.pos, pos4, bytes, kind, c.stream));
  BT_CK(hipMemcpyAsync(c.vel, vel4, bytes, kind, c.stream));
  std::vector<int> ids(c.n);
  for (int i = 0; i < c.n; i++) ids[i] = i;
  BT_CK(hipMemcpyAsync(c.id, ids.data(), sizeof(int) * c.n, hipMemcpyHostToDevice, c.stream));
  BT_CK(hipStreamSynchronize(c.stream));
  c.have_particles = true;
  c.pending_kick = 0.0;
  c.have_tree = c.have_mp = c.have_groups = c.have_acc = false;
  return 0;
}

}  // namespace

extern "C" {

int bt_set_particles(bt_ctx *h, const float *pos4, const float *vel4) {
  if (!h || !pos4 || !vel4) return BT_ERR_ARG;
  if (check_real<float>(h->c)) return BT_ERR_STATE;
  return put_particles(h->c, pos4, vel4, hipMemcpyHostToDevice);
}

int bt_set_particles_f64(bt_ctx *h, const double *pos4, const double *vel4) {
  if (!h || !pos4 || !vel4) return BT_ERR_ARG;
  if (check_real<double>(h->c)) return BT_ERR_STATE;
  return put_particles(h->c, pos4, vel4, hipMemcpyHostToDevice);
}

int bt_set_particles_device(bt_ctx *h, const void *dpos, const void *dvel) {
  if (!h || !dpos || !dvel) return BT_ERR_ARG;
  return put_particles(h->c, dpos, dvel, hipMemcpyDeviceToDevice);
}

int bt_build_tree(bt_ctx *h, int nleaf) {
  if (!h) return BT_ERR_ARG;
  Ctx &c = h->c;
  c.err.clear();
  if (!valid_nleaf(nleaf)) return fail(c, BT_ERR_ARG, "nleaf must be 16, 24, 32, 48 or 64");
  if (!c.have_particles) return fail(c, BT_ERR_STATE, "no particles");
  c.nleaf = nleaf;
  c.have_tree = c.have_mp = c.have_groups = c.have_acc = false;
  const int rc = build_checked(c);
  if (rc) return rc;
  finish_timings(c, 3);
  c.have_tree = true;
  c.tree_version++;
  return 0;
}

int bt_compute_multipoles(bt_ctx *h) {
  if (!h) return BT_ERR_ARG;
  Ctx &c = h->c;
  c.err.clear();
  if (!c.have_tree) return fail(c, BT_ERR_STATE, "buildTree first");
  record(c, 3);
  if (phase_multipoles(c)) return BT_ERR_HIP;
  BT_CK(hipStreamSynchronize(c.stream));
  c.have_mp = true;
  return 0;
}

int bt_make_groups(bt_ctx *h, int level_split, int ncrit) {
  if (!h) return BT_ERR_ARG;
  Ctx &c = h->c;
  c.err.clear();
  if (ncrit < 1 || ncrit > 64) return fail(c, BT_ERR_ARG, "ncrit must be in [1, 64]");
  if (level_split > 21) return fail(c, BT_ERR_ARG, "level_split must be <= 21");
  if (!c.have_tree) return fail(c, BT_ERR_STATE, "buildTree first (makeGroups reads its box)");
  c.level_split = std::max(1, level_split);
  c.ncrit = ncrit;
  record(c, 4);
  if (phase_groups(c)) return BT_ERR_HIP;
  if (read_counters(c)) return BT_ERR_HIP;
  if (c.hc.err & kSortErr) return fail(c, BT_ERR_HIP, kSortErrMsg);
  if (c.hc.err & 2) return fail(c, BT_ERR_CAPACITY, kGroupCapacityMsg);
  c.have_groups = true;
  c.have_acc = false;
  return 0;
}

int bt_compute_forces(bt_ctx *h, int intcount, double stats[4]) {
  if (!h) return BT_ERR_ARG;
  Ctx &c = h->c;
  c.err.clear();
  if (!c.have_mp || !c.have_groups)
    return fail(c, BT_ERR_STATE, "computeMultipoles and makeGroups first");
  launch_clear_sticky(c);
  record(c, 7);
  // the check's sum runs after this walk, from the tree-order particles
  if (join_check(c)) return BT_ERR_HIP;
  c.check_state = 0;
  if (phase_forces(c) || launch_learned_check(c)) return BT_ERR_HIP;
  // a moveParticles before this evaluation owes its step's closing half kick
  if (c.pending_kick != 0.0) launch_kick(c, c.pending_kick, true);
  BT_LAUNCH_CK();
  if (read_counters(c)) return BT_ERR_HIP;
  if (c.hc.err_sticky & 4) return fail(c, BT_ERR_STACK, "walk stack overflow");
  c.pending_kick = 0.0;
  c.have_acc = true;
  if (intcount && stats) {
    const unsigned long long *w = c.h_wstats;   // fetched by read_counters above
    stats[0] = (double)w[0] / c.n;
    stats[1] = (double)w[1];
    stats[2] = (double)w[2] / c.n;
    stats[3] = (double)w[3];
  }
  finish_timings(c, 8);
  return 0;
}

int bt_kick(bt_ctx *h, float dth) {
  if (!h) return BT_ERR_ARG;
  Ctx &c = h->c;
  if (!c.have_acc) return fail(c, BT_ERR_STATE, "computeForces first");
  launch_clear_sticky(c);
  launch_kick(c, dth);
  BT_LAUNCH_CK();
  BT_CK(hipStreamSynchronize(c.stream));
  return 0;
}

int bt_move_particles(bt_ctx *h, float dt) { return bt_move_particles_f64(h, dt); }

int bt_move_particles_f64(bt_ctx *h, double dt) {
  if (!h) return BT_ERR_ARG;
  Ctx &c = h->c;
  if (!c.have_acc) return fail(c, BT_ERR_STATE, "computeForces first");
  launch_clear_sticky(c);
  launch_kick_drift(c, 0.5 * dt, dt);
  BT_LAUNCH_CK();
  BT_CK(hipStreamSynchronize(c.stream));
  // positions moved: the tree, moments, groups and forces are stale; the
  // next force evaluation applies the step's second half kick
  c.pending_kick = 0.5 * dt;
  c.have_tree = c.have_mp = c.have_groups = c.have_acc = false;
  return 0;
}

int bt_compute_energies(bt_ctx *h, double *ekin, double *epot) {
  if (!h) return BT_ERR_ARG;
  Ctx &c = h->c;
  if (!c.have_particles) return fail(c, BT_ERR_STATE, "no particles");
  if (!c.have_acc) {
    // phi of the current positions (after moveParticles: the new forces also
    // close the step's kick, so K and W are taken at the same time)
    int rc = bt_build_tree(h, c.nleaf);
    if (!rc) rc = bt_compute_multipoles(h);
    if (!rc) rc = bt_make_groups(h, c.level_split, c.ncrit);
    if (!rc) rc = bt_compute_forces(h, 0, nullptr);
    if (rc) return rc;
  }
  launch_energies(c);
  BT_LAUNCH_CK();
  std::vector<double> e(2 * (size_t)c.nsys);
  BT_CK(hipMemcpyAsync(e.data(), c.d_energy, sizeof(double) * 2 * c.nsys,
                       hipMemcpyDeviceToHost, c.stream));
  BT_CK(hipStreamSynchronize(c.stream));
  for (int r = 0; r < c.nsys; r++) {
    if (ekin) ekin[r] = e[2 * r];
    if (epot) epot[r] = e[2 * r + 1];
  }
  return 0;
}

// One kick-drift-kick step (forces first if none are current), queued without
// host synchronisation until the end.
int bt_step(bt_ctx *h, float dt, int nleaf, int level_split, int ncrit) {
  return bt_step_f64(h, dt, nleaf, level_split, ncrit);
}

int bt_step_f64(bt_ctx *h, double dt, int nleaf, int level_split, int ncrit) {
  if (!h) return BT_ERR_ARG;
  Ctx &c = h->c;
  c.err.clear();
  if (!valid_nleaf(nleaf)) return fail(c, BT_ERR_ARG, "nleaf must be 16, 24, 32, 48 or 64");
  if (ncrit < 1 || ncrit > 64) return fail(c, BT_ERR_ARG, "ncrit must be in [1, 64]");
  if (!c.have_particles) return fail(c, BT_ERR_STATE, "no particles");
  c.nleaf = nleaf;
  c.ncrit = ncrit;
  c.level_split = std::max(1, level_split);
  if (!c.have_acc) {
    int rc = bt_build_tree(h, nleaf);
    if (!rc) rc = bt_compute_multipoles(h);
    if (!rc) rc = bt_make_groups(h, level_split, ncrit);
    if (!rc) rc = bt_compute_forces(h, 1, nullptr);
    if (rc) return rc;
  }
  launch_clear_sticky(c);
  launch_kick_drift(c, 0.5 * dt, dt);
  c.have_tree = c.have_mp = c.have_groups = c.have_acc = false;
  int rc = step_second_half(c, dt);
  if (!rc) rc = finish_step(c, dt);
  if (rc) return rc;
  c.have_tree = c.have_mp = c.have_groups = c.have_acc = true;
  finish_timings(c, 9);
  return 0;
}

// ---- results --------------------------------------------------------------

}  // extern "C"

namespace {

// acc (PH order on the device) in the requested order; T = the context's real type
template <typename T>
int get_acc_t(Ctx &c, T *acc4, int order) {
  typedef typename RealVec<T>::v4 V;
  if (check_real<T>(c)) return BT_ERR_STATE;
  if (!c.have_acc) return fail(c, BT_ERR_STATE, "no forces");
  if (order != BT_ORDER_PH && order != BT_ORDER_TREE && order != BT_ORDER_INPUT)
    return fail(c, BT_ERR_ARG, "bad order");
  std::vector<V> a(c.n);
  BT_CK(hipMemcpyAsync(a.data(), c.acc, sizeof(V) * c.n, hipMemcpyDeviceToHost, c.stream));
  V *o = reinterpret_cast<V *>(acc4);
  if (order == BT_ORDER_PH) {
    BT_CK(hipStreamSynchronize(c.stream));
    std::memcpy(o, a.data(), sizeof(V) * c.n);
    return 0;
  }
  std::vector<int> map(c.n);
  BT_CK(hipMemcpyAsync(map.data(), order == BT_ORDER_TREE ? c.perm : c.id_ph,
                       sizeof(int) * c.n, hipMemcpyDeviceToHost, c.stream));
  BT_CK(hipStreamSynchronize(c.stream));
  for (int k = 0; k < c.n; k++) o[map[k]] = a[k];
  return 0;
}

template <typename T>
int get_particles_t(Ctx &c, T *pos4, T *vel4, int order) {
  typedef typename RealVec<T>::v4 V;
  if (check_real<T>(c)) return BT_ERR_STATE;
  if (!c.have_particles) return fail(c, BT_ERR_STATE, "no particles");
  if (order != BT_ORDER_PH && order != BT_ORDER_TREE && order != BT_ORDER_INPUT)
    return fail(c, BT_ERR_ARG, "bad order");
  if (order == BT_ORDER_PH && !c.have_groups)
    return fail(c, BT_ERR_STATE, "no PH order before makeGroups");
  std::vector<V> p(c.n), v(c.n);
  std::vector<int> id(c.n), perm;
  BT_CK(hipMemcpyAsync(p.data(), c.pos, sizeof(V) * c.n, hipMemcpyDeviceToHost, c.stream));
  BT_CK(hipMemcpyAsync(v.data(), c.vel, sizeof(V) * c.n, hipMemcpyDeviceToHost, c.stream));
  BT_CK(hipMemcpyAsync(id.data(), c.id, sizeof(int) * c.n, hipMemcpyDeviceToHost, c.stream));
  if (order == BT_ORDER_PH) {
    perm.resize(c.n);
    BT_CK(hipMemcpyAsync(perm.data(), c.perm, sizeof(int) * c.n, hipMemcpyDeviceToHost,
                         c.stream));
  }
  BT_CK(hipStreamSynchronize(c.stream));
  V *po = reinterpret_cast<V *>(pos4);
  V *vo = reinterpret_cast<V *>(vel4);
  for (int k = 0; k < c.n; k++) {
    // tree order: the arrays as they are; input order: through the ids;
    // PH order: slot k holds tree particle perm[k]
    const int src = order == BT_ORDER_PH ? perm[k] : k;
    const int dst = order == BT_ORDER_INPUT ? id[k] : k;
    if (po) po[dst] = p[src];
    if (vo) vo[dst] = v[src];
  }
  return 0;
}

}  // namespace

extern "C" {

int bt_get_acc(bt_ctx *h, float *acc4, int order) {
  if (!h || !acc4) return BT_ERR_ARG;
  return get_acc_t(h->c, acc4, order);
}

int bt_get_acc_f64(bt_ctx *h, double *acc4, int order) {
  if (!h || !acc4) return BT_ERR_ARG;
  return get_acc_t(h->c, acc4, order);
}

int bt_get_particles(bt_ctx *h, float *pos4, float *vel4, int order) {
  if (!h) return BT_ERR_ARG;
  return get_particles_t(h->c, pos4, vel4, order);
}

int bt_get_particles_f64(bt_ctx *h, double *pos4, double *vel4, int order) {
  if (!h) return BT_ERR_ARG;
  return get_particles_t(h->c, pos4, vel4, order);
}

int bt_get_walk_stats(bt_ctx *h, double stats[4], double totals2[2]) {
  if (!h || !stats) return BT_ERR_ARG;
  Ctx &c = h->c;
  if (!c.have_acc) return fail(c, BT_ERR_STATE, "no forces");
  if (!c.h_wstats_valid && read_counters(c)) return BT_ERR_HIP;
  const unsigned long long *w = c.h_wstats;
  stats[0] = (double)w[0] / c.n;
  stats[1] = (double)w[1];
  stats[2] = (double)w[2] / c.n;
  stats[3] = (double)w[3];
  if (totals2) {
    totals2[0] = (double)w[0];
    totals2[1] = (double)w[2];
  }
  return 0;
}

int bt_get_potential(bt_ctx *h, double *pot) {
  if (!h || !pot) return BT_ERR_ARG;
  Ctx &c = h->c;
  if (!c.have_acc) return fail(c, BT_ERR_STATE, "no forces");
  // on the context stream: a plain hipMemcpy would not wait for its work
  BT_CK(hipMemcpyAsync(pot, c.d_wpot, sizeof(double), hipMemcpyDeviceToHost, c.stream));
  BT_CK(hipStreamSynchronize(c.stream));
  return 0;
}

int bt_get_counts(bt_ctx *h, int counts[8]) {
  if (!h || !counts) return BT_ERR_ARG;
  Ctx &c = h->c;
  if (c.dc && read_counters(c)) return BT_ERR_HIP;
  counts[0] = c.n;
  counts[1] = c.hc.ncells;
  counts[2] = c.hc.nnodes;
  counts[3] = c.hc.nleaves;
  counts[4] = c.hc.nlevels;
  counts[5] = c.hc.ngroups;
  counts[6] = c.nsys;
  counts[7] = c.hc.flags;
  return 0;
}

int bt_get_tree(bt_ctx *h, uint32_t *cells4, int *level_range64, float *size4, float *mono4,
                float *q0_4, float *q1_2) {
  if (!h) return BT_ERR_ARG;
  Ctx &c = h->c;
  if (!c.have_tree) return fail(c, BT_ERR_STATE, "no tree");
  if ((size4 || mono4 || q0_4 || q1_2) && check_real<float>(c)) return BT_ERR_STATE;
  if (read_counters(c)) return BT_ERR_HIP;
  const size_t nc = (size_t)c.hc.ncells;
  if (cells4) BT_CK(d2h(c, cells4, c.cells, sizeof(uint4) * nc));
  if (level_range64)
    BT_CK(d2h(c, level_range64, c.level_beg, sizeof(int2) * kLevels));
  if ((size4 || mono4 || q0_4 || q1_2) && !c.have_mp)
    return fail(c, BT_ERR_STATE, "no multipoles");
  if (size4) BT_CK(d2h(c, size4, c.csize, sizeof(float4) * nc));
  if (mono4) BT_CK(d2h(c, mono4, c.cmono, sizeof(float4) * nc));
  if (q0_4) BT_CK(d2h(c, q0_4, c.cq0, sizeof(float4) * nc));
  if (q1_2) BT_CK(d2h(c, q1_2, c.cq1, sizeof(float2) * nc));
  return 0;
}

int bt_get_tree_f64(bt_ctx *h, double *size4, double *mono4, double *q0_4, double *q1_2) {
  if (!h) return BT_ERR_ARG;
  Ctx &c = h->c;
  if (check_real<double>(c)) return BT_ERR_STATE;
  if (!c.have_mp) return fail(c, BT_ERR_STATE, "no multipoles");
  if (read_counters(c)) return BT_ERR_HIP;
  const size_t nc = (size_t)c.hc.ncells;
  if (size4) BT_CK(d2h(c, size4, c.csize, sizeof(double4) * nc));
  if (mono4) BT_CK(d2h(c, mono4, c.cmono, sizeof(double4) * nc));
  if (q0_4) BT_CK(d2h(c, q0_4, c.cq0, sizeof(double4) * nc));
  if (q1_2) BT_CK(d2h(c, q1_2, c.cq1, sizeof(double2) * nc));
  return 0;
}

int bt_get_cell_ls(bt_ctx *h, float *ls2) {
  if (!h || !ls2) return BT_ERR_ARG;
  Ctx &c = h->c;
  if (check_real<float>(c)) return BT_ERR_STATE;
  if (!c.have_mp) return fail(c, BT_ERR_STATE, "no multipoles");
  if (read_counters(c)) return BT_ERR_HIP;
  BT_CK(d2h(c, ls2, c.cls, sizeof(float2) * (size_t)c.hc.ncells));
  return 0;
}

int bt_get_groups(bt_ctx *h, int *groups2) {
  if (!h || !groups2) return BT_ERR_ARG;
  Ctx &c = h->c;
  if (!c.have_groups) return fail(c, BT_ERR_STATE, "no groups");
  if (read_counters(c)) return BT_ERR_HIP;
  BT_CK(d2h(c, groups2, c.groups, sizeof(int2) * (size_t)c.hc.ngroups));
  return 0;
}

int bt_get_group_counts(bt_ctx *h, int *counts2) {
  if (!h || !counts2) return BT_ERR_ARG;
  Ctx &c = h->c;
  if (!c.have_acc) return fail(c, BT_ERR_STATE, "no forces");
  if (read_counters(c)) return BT_ERR_HIP;
  BT_CK(d2h(c, counts2, c.gcounts, sizeof(int2) * (size_t)c.hc.ngroups));
  return 0;
}

int bt_get_ph(bt_ctx *h, uint64_t *keys, int *perm) {
  if (!h) return BT_ERR_ARG;
  Ctx &c = h->c;
  if (!c.have_groups) return fail(c, BT_ERR_STATE, "no groups");
  if (keys) BT_CK(d2h(c, keys, c.ph_keys, sizeof(uint64_t) * c.n));
  if (perm) BT_CK(d2h(c, perm, c.perm, sizeof(int) * c.n));
  return 0;
}

int bt_get_ids(bt_ctx *h, int *ids, int order) {
  if (!h || !ids) return BT_ERR_ARG;
  Ctx &c = h->c;
  if (order == BT_ORDER_TREE) {
    if (!c.have_tree) return fail(c, BT_ERR_STATE, "no tree");
    BT_CK(d2h(c, ids, c.id, sizeof(int) * c.n));
  } else if (order == BT_ORDER_PH) {
    if (!c.have_groups) return fail(c, BT_ERR_STATE, "no groups");
    BT_CK(d2h(c, ids, c.id_ph, sizeof(int) * c.n));
  } else {
    for (int i = 0; i < c.n; i++) ids[i] = i;
  }
  return 0;
}

int bt_get_boxes(bt_ctx *h, float *boxes4) {
  if (!h || !boxes4) return BT_ERR_ARG;
  Ctx &c = h->c;
  if (check_real<float>(c)) return BT_ERR_STATE;
  if (!c.have_tree) return fail(c, BT_ERR_STATE, "no tree");
  BT_CK(d2h(c, boxes4, c.box, sizeof(float4) * c.nsys));
  return 0;
}

int bt_get_boxes_f64(bt_ctx *h, double *boxes4) {
  if (!h || !boxes4) return BT_ERR_ARG;
  Ctx &c = h->c;
  if (check_real<double>(c)) return BT_ERR_STATE;
  if (!c.have_tree) return fail(c, BT_ERR_STATE, "no tree");
  BT_CK(d2h(c, boxes4, c.box, sizeof(double4) * c.nsys));
  return 0;
}

int bt_direct_fp64(bt_ctx *h, int ns, const int *ph_idx, double *out4) {
  if (!h || ns <= 0 || !ph_idx || !out4) return BT_ERR_ARG;
  Ctx &c = h->c;
  if (!c.have_groups) return fail(c, BT_ERR_STATE, "makeGroups first (targets are PH indices)");
  for (int k = 0; k < ns; k++)
    if (ph_idx[k] < 0 || ph_idx[k] >= c.n) return fail(c, BT_ERR_ARG, "target index out of range");
  const size_t need = (size_t)ns * (direct_slices(ns) + 1) * 4 + (size_t)ns;  // + the indices
  if ((int)need > c.direct_cap || !c.d_direct) {
    if (dev_alloc(c, &c.d_direct, need)) return BT_ERR_HIP;
    c.direct_cap = (int)need;
  }
  int *d_idx = reinterpret_cast<int *>(c.d_direct + (size_t)ns * (direct_slices(ns) + 1) * 4);
  BT_CK(hipMemcpyAsync(d_idx, ph_idx, sizeof(int) * ns, hipMemcpyHostToDevice, c.stream));
  launch_direct(c, ns, d_idx, c.d_direct);
  BT_LAUNCH_CK();
  BT_CK(hipMemcpyAsync(out4, c.d_direct, sizeof(double) * 4 * ns, hipMemcpyDeviceToHost,
                       c.stream));
  BT_CK(hipStreamSynchronize(c.stream));
  return 0;
}

int bt_direct_fp64_noalloc(bt_ctx *h, int ns, const int *d_ph_idx, double *d_out4) {
  if (!h) return BT_ERR_ARG;
  Ctx &c = h->c;
  if (!d_ph_idx || !d_out4) return fail(c, BT_ERR_ARG, "null device pointer");
  if (ns <= 0 || ns > kNoAllocMaxTargets)
    return fail(c, BT_ERR_ARG, "bt_direct_fp64_noalloc: ns must be in [1, 4096]");
  if (!c.have_groups) return fail(c, BT_ERR_STATE, "makeGroups first (targets are PH indices)");
  double4 *scr = nullptr;
  if (scratch_lookup(c.dev, &scr)) return fail(c, BT_ERR_HIP, "scratch symbol lookup failed");
  // as many source slices as bt_direct_fp64 uses, up to what the scratch holds
  const int slices = std::min(direct_slices(ns), (int)(kScratchD4 / (size_t)ns));
  launch_direct_parts(c, ns, d_ph_idx, d_out4, reinterpret_cast<double *>(scr), slices);
  BT_LAUNCH_CK();
  return 0;
}

int bt_get_timings(bt_ctx *h, float ms[10]) {
  if (!h || !ms) return BT_ERR_ARG;
  for (int k = 0; k < 10; k++) ms[k] = h->c.ms[k];
  return 0;
}


// ---- memory of the reference's public members ------------------------------

int bt_host_alloc(void **ptr, size_t bytes) {
  if (!ptr) return BT_ERR_ARG;
  *ptr = nullptr;
  if (bytes == 0) return 0;
  return hipHostMalloc(ptr, bytes, hipHostMallocDefault) == hipSuccess ? 0 : BT_ERR_HIP;
}

int bt_host_free(void *ptr) {
  if (!ptr) return 0;
  return hipHostFree(ptr) == hipSuccess ? 0 : BT_ERR_HIP;
}

namespace {

// Address, element count and element size of one of the BT_BUF_* arrays.
int device_buffer(Ctx &c, int which, void **ptr, int64_t *count, size_t *elem) {
  if (c.dc && read_counters(c)) return BT_ERR_HIP;
  const int64_t n = c.n, nc = c.hc.ncells, ng = c.hc.ngroups;
  void *p = nullptr;
  int64_t k = 0;
  size_t e = 0;
  const size_t r4 = sizeof(float4) * rw(c), r2 = sizeof(float2) * rw(c);
  switch (which) {
    case BT_BUF_PTCL_POS: p = c.pos; k = n; e = r4; break;
    case BT_BUF_PTCL_VEL: p = c.vel; k = n; e = r4; break;
    case BT_BUF_PTCL_POS_TMP: p = c.pos_ph; k = n; e = r4; break;
    case BT_BUF_PTCL_ACC: p = c.acc; k = n; e = r4; break;
    case BT_BUF_DOMAIN: p = c.box; k = c.nsys; e = r4; break;
    case BT_BUF_LEVEL_BEGIDX: p = c.level_beg; k = kLevels; e = sizeof(int2); break;
    case BT_BUF_CELL_DATA: p = c.cells; k = nc; e = sizeof(uint4); break;
    case BT_BUF_CELL_SIZE: p = c.csize; k = nc; e = r4; break;
    case BT_BUF_CELL_MONOPOLE: p = c.cmono; k = nc; e = r4; break;
    case BT_BUF_CELL_QUAD0: p = c.cq0; k = nc; e = r4; break;
    case BT_BUF_CELL_QUAD1: p = c.cq1; k = nc; e = r2; break;
    case BT_BUF_GROUP_LIST: p = c.groups; k = ng; e = sizeof(int2); break;
    case BT_BUF_PH_KEY: p = c.ph_keys; k = n; e = sizeof(uint64_t); break;
    case BT_BUF_PH_PERM: p = c.perm; k = n; e = sizeof(int); break;
    default: return fail(c, BT_ERR_ARG, "unknown device buffer");
  }
  if (ptr) *ptr = p;
  if (count) *count = p ? k : 0;
  if (elem) *elem = e;
  return 0;
}

}  // namespace

int bt_get_device_buffer(bt_ctx *h, int which, void **ptr, int64_t *count) {
  if (!h) return BT_ERR_ARG;
  return device_buffer(h->c, which, ptr, count, nullptr);
}

int bt_copy_device_buffer(bt_ctx *h, int which, void *host, int64_t count) {
  if (!h || !host || count < 0) return BT_ERR_ARG;
  Ctx &c = h->c;
  void *p = nullptr;
  int64_t k = 0;
  size_t e = 0;
  const int rc = device_buffer(c, which, &p, &k, &e);
  if (rc) return rc;
  if (count > k) return fail(c, BT_ERR_ARG, "count exceeds the buffer");
  if (count == 0) return 0;
  BT_CK(hipMemcpyAsync(host, p, e * (size_t)count, hipMemcpyDeviceToHost, c.stream));
  BT_CK(hipStreamSynchronize(c.stream));
  return 0;
}

// ---- learned opening angle ------------------------------------------------

int bt_mlp_set_params(bt_ctx *h, const float *params) {
  if (!h || !params) return BT_ERR_ARG;
  Ctx &c = h->c;
  for (int i = 0; i < kMlpParams; i++)
    if (!std::isfinite(params[i])) return fail(c, BT_ERR_ARG, "non-finite network parameter");
  BT_CK(hipMemcpyAsync(c.mlp_params, params, sizeof(float) * kMlpParams, hipMemcpyHostToDevice,
                       c.stream));
  launch_mlp_prepare(c);
  BT_LAUNCH_CK();
  BT_CK(hipStreamSynchronize(c.stream));
  c.have_mlp = true;
  return 0;
}

int bt_mlp_get_params(bt_ctx *h, float *params) {
  if (!h || !params) return BT_ERR_ARG;
  Ctx &c = h->c;
  if (!c.have_mlp) return fail(c, BT_ERR_STATE, "no network parameters");
  BT_CK(hipMemcpyAsync(params, c.mlp_params, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost,
                       c.stream));
  BT_CK(hipStreamSynchronize(c.stream));
  return 0;
}

int bt_mlp_init(bt_ctx *h, uint64_t seed, float b3) {
  if (!h) return BT_ERR_ARG;
  std::vector<float> p(kMlpParams, 0.0f);
  uint64_t st = seed;
  auto uni = [&st]() {  // uniform in [-1, 1)
    st = mix64(st);
    return (float)((double)(st >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0);
  };
  const float g1 = std::sqrt(6.0f / (kMlpIn + kMlpH)), g2 = std::sqrt(6.0f / (2 * kMlpH)),
              g3 = std::sqrt(6.0f / (kMlpH + 1));
  for (int i = 0; i < kMlpIn * kMlpH; i++) p[kOffW1 + i] = g1 * uni();
  for (int i = 0; i < kMlpH * kMlpH; i++) p[kOffW2 + i] = g2 * uni();
  for (int i = 0; i < kMlpH; i++) p[kOffW3 + i] = g3 * uni();
  p[kOffB3] = b3;
  return bt_mlp_set_params(h, p.data());
}

int bt_mlp_forward(bt_ctx *h, int n, const float *x, float *y) {
  if (!h || n <= 0 || !x || !y) return BT_ERR_ARG;
  Ctx &c = h->c;
  if (!c.have_mlp) return fail(c, BT_ERR_STATE, "no network parameters");
  float *dx = nullptr, *dy = nullptr;
  if (dev_alloc(c, &dx, (size_t)n * 8) || dev_alloc(c, &dy, (size_t)n)) {
    dev_free(&dx);
    return BT_ERR_HIP;
  }
  int rc = 0;
  if (hipMemcpyAsync(dx, x, sizeof(float) * 8 * (size_t)n, hipMemcpyHostToDevice, c.stream) !=
      hipSuccess)
    rc = BT_ERR_HIP;
  if (!rc) {
    launch_mlp_forward(c, n, dx, dy);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(y, dy, sizeof(float) * (size_t)n, hipMemcpyDeviceToHost, c.stream) !=
            hipSuccess ||
        hipStreamSynchronize(c.stream) != hipSuccess)
      rc = BT_ERR_HIP;
  }
  dev_free(&dx);
  dev_free(&dy);
  return rc ? fail(c, rc, "network forward failed") : 0;
}

namespace {

struct TrainBufs {
  float *x = nullptr, *t = nullptr, *partial = nullptr;
  double *grad = nullptr, *losses = nullptr;
  void release() {
    dev_free(&x); dev_free(&t); dev_free(&partial); dev_free(&grad); dev_free(&losses);
  }
};

int upload_samples(Ctx &c, TrainBufs &b, int n, const float *x, const float *t, int nloss) {
  if (dev_alloc(c, &b.x, (size_t)n * 8) || dev_alloc(c, &b.t, (size_t)n) ||
      dev_alloc(c, &b.partial, (size_t)mlp_grad_blocks() * (kMlpParams + 1)) ||
      dev_alloc(c, &b.grad, (size_t)kMlpParams + 1) ||
      dev_alloc(c, &b.losses, (size_t)std::max(nloss, 1)))
    return BT_ERR_HIP;
  BT_CK(hipMemcpyAsync(b.x, x, sizeof(float) * 8 * (size_t)n, hipMemcpyHostToDevice, c.stream));
  BT_CK(hipMemcpyAsync(b.t, t, sizeof(float) * (size_t)n, hipMemcpyHostToDevice, c.stream));
  return 0;
}

}  // namespace

int bt_mlp_grad(bt_ctx *h, int n, const float *x, const float *t, double *grad, double *loss) {
  if (!h || n <= 0 || !x || !t || !grad) return BT_ERR_ARG;
  Ctx &c = h->c;
  if (!c.have_mlp) return fail(c, BT_ERR_STATE, "no network parameters");
  TrainBufs b;
  int rc = upload_samples(c, b, n, x, t, 1);
  if (!rc) {
    launch_mlp_grad(c, c.mlp_params, n, b.x, b.t, n, 0ull, 0, b.partial, b.grad);
    std::vector<double> g(kMlpParams + 1);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(g.data(), b.grad, sizeof(double) * (kMlpParams + 1), hipMemcpyDeviceToHost,
                       c.stream) != hipSuccess ||
        hipStreamSynchronize(c.stream) != hipSuccess) {
      rc = BT_ERR_HIP;
    } else {
      std::memcpy(grad, g.data(), sizeof(double) * kMlpParams);
      if (loss) *loss = g[kMlpParams];
    }
  }
  b.release();
  return rc ? fail(c, rc, "network gradient failed") : 0;
}

int bt_mlp_train(bt_ctx *h, int n, const float *x, const float *t, int steps, int batch, float lr,
                 uint64_t seed, float *loss_curve) {
  if (!h || n <= 0 || !x || !t || steps <= 0 || batch <= 0 || !(lr > 0.f)) return BT_ERR_ARG;
  Ctx &c = h->c;
  if (!c.have_mlp) return fail(c, BT_ERR_STATE, "initialise or set the network first");
  batch = std::min(batch, n);
  TrainBufs b;
  int rc = upload_samples(c, b, n, x, t, steps);
  if (!rc) {
    if (hipMemsetAsync(c.mlp_m, 0, sizeof(float) * kMlpParams, c.stream) != hipSuccess ||
        hipMemsetAsync(c.mlp_v, 0, sizeof(float) * kMlpParams, c.stream) != hipSuccess)
      rc = BT_ERR_HIP;
    for (int k = 0; k < steps && !rc; k++) {
      // cosine decay to lr / 20
      const float lr_k =
          lr * (0.05f + 0.95f * 0.5f * (1.0f + std::cos(3.14159265f * (float)k / (float)steps)));
      launch_mlp_grad(c, c.mlp_params, n, b.x, b.t, batch, seed, k, b.partial, b.grad);
      launch_mlp_adam(c, c.mlp_params, c.mlp_m, c.mlp_v, b.grad, lr_k, k + 1);
      if (hipMemcpyAsync(b.losses + k, b.grad + kMlpParams, sizeof(double),
                         hipMemcpyDeviceToDevice, c.stream) != hipSuccess)
        rc = BT_ERR_HIP;
    }
    if (!rc) {
      launch_mlp_prepare(c);
      std::vector<double> l(steps);
      if (hipGetLastError() != hipSuccess ||
          hipMemcpyAsync(l.data(), b.losses, sizeof(double) * steps, hipMemcpyDeviceToHost,
                         c.stream) != hipSuccess ||
          hipStreamSynchronize(c.stream) != hipSuccess)
        rc = BT_ERR_HIP;
      else if (loss_curve)
        for (int k = 0; k < steps; k++) loss_curve[k] = (float)l[k];
    }
  }
  b.release();
  return rc ? fail(c, rc, "network training failed") : 0;
}

int bt_learn_collect(bt_ctx *h, int nsel, const int *gsel, float theta_lo, float theta_hi,
                     int max_pairs, float *feat, float *target, float *err, float *theta_need,
                     int *pair_gc, int *npairs) {
  if (!h || nsel <= 0 || !gsel || !npairs || max_pairs <= 0) return BT_ERR_ARG;
  Ctx &c = h->c;
  c.err.clear();
  if (!(theta_lo > 0.f) || !(theta_hi > theta_lo)) return fail(c, BT_ERR_ARG, "bad theta band");
  if (c.fp64) return fail(c, BT_ERR_STATE, "training pairs come from Treecode<float> contexts");
  if (!c.have_mp || !c.have_groups) return fail(c, BT_ERR_STATE, "computeMultipoles and makeGroups first");
  if (read_counters(c)) return BT_ERR_HIP;
  const int ng = c.hc.ngroups;
  std::vector<int2> grp(ng);
  BT_CK(d2h(c, grp.data(), c.groups, sizeof(int2) * ng));
  std::vector<int> pidx;
  for (int k = 0; k < nsel; k++) {
    if (gsel[k] < 0 || gsel[k] >= ng) return fail(c, BT_ERR_ARG, "group index out of range");
    for (int i = 0; i < grp[gsel[k]].y; i++) pidx.push_back(grp[gsel[k]].x + i);
  }
  if (ensure_stack(c)) return BT_ERR_HIP;
  LearnBuffers L;
  L.nsel = nsel;
  L.theta_lo = theta_lo;
  L.theta_hi = theta_hi;
  L.max_pairs = max_pairs;
  const int ns = (int)pidx.size();
  double *dd = nullptr;
  int *didx = nullptr;
  int rc = 0;
  rc |= dev_alloc(c, &L.gsel, nsel);
  rc |= dev_alloc(c, &L.count, 1);
  rc |= dev_alloc(c, &L.pairs, max_pairs);
  rc |= dev_alloc(c, &L.tneed, max_pairs);
  rc |= dev_alloc(c, &L.amag, (size_t)c.n + 64);
  rc |= dev_alloc(c, &L.feat, (size_t)max_pairs * 8);
  rc |= dev_alloc(c, &L.target, max_pairs);
  rc |= dev_alloc(c, &L.err, max_pairs);
  rc |= dev_alloc(c, &L.tneed_f, max_pairs);
  rc |= dev_alloc(c, &L.sink, (size_t)c.walk_blocks * 4 * 64);
  rc |= dev_alloc(c, &dd, (size_t)ns * (direct_slices(ns) + 1) * 4);
  rc |= dev_alloc(c, &didx, ns);
  int found = 0;
  if (!rc) {
    hipError_t e = hipMemcpyAsync(L.gsel, gsel, sizeof(int) * nsel, hipMemcpyHostToDevice, c.stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(didx, pidx.data(), sizeof(int) * ns, hipMemcpyHostToDevice, c.stream);
    if (e == hipSuccess) e = hipMemsetAsync(L.count, 0, sizeof(int), c.stream);
    if (e == hipSuccess) e = hipMemsetAsync(L.amag, 0, sizeof(double) * ((size_t)c.n + 64), c.stream);
    if (e == hipSuccess) {
      launch_direct(c, ns, didx, dd);
      launch_learn_amag(c, ns, didx, dd, L.amag);
      launch_learn_pairs(c, L);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(&found, L.count, sizeof(int), hipMemcpyDeviceToHost, c.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c.stream);
    const int np = std::min(found, max_pairs);
    if (e == hipSuccess && np > 0) {
      launch_learn_errors(c, L, np);
      e = hipGetLastError();
      std::vector<int2> pg(np);
      if (e == hipSuccess && feat)
        e = hipMemcpyAsync(feat, L.feat, sizeof(float) * 8 * np, hipMemcpyDeviceToHost, c.stream);
      if (e == hipSuccess && target)
        e = hipMemcpyAsync(target, L.target, sizeof(float) * np, hipMemcpyDeviceToHost, c.stream);
      if (e == hipSuccess && err)
        e = hipMemcpyAsync(err, L.err, sizeof(float) * np, hipMemcpyDeviceToHost, c.stream);
      if (e == hipSuccess && theta_need)
        e = hipMemcpyAsync(theta_need, L.tneed_f, sizeof(float) * np, hipMemcpyDeviceToHost,
                           c.stream);
      if (e == hipSuccess && pair_gc)
        e = hipMemcpyAsync(pair_gc, L.pairs, sizeof(int2) * np, hipMemcpyDeviceToHost, c.stream);
      if (e == hipSuccess) e = hipStreamSynchronize(c.stream);
    }
    if (e != hipSuccess) {
      c.err = std::string("learn_collect: ") + hipGetErrorString(e);
      rc = BT_ERR_HIP;
    }
  }
  dev_free(&L.gsel); dev_free(&L.count); dev_free(&L.pairs); dev_free(&L.tneed);
  dev_free(&L.amag); dev_free(&L.feat); dev_free(&L.target); dev_free(&L.err);
  dev_free(&L.tneed_f); dev_free(&L.sink); dev_free(&dd); dev_free(&didx);
  if (rc) return rc;
  if (read_counters(c)) return BT_ERR_HIP;
  if (c.hc.err & 4) return fail(c, BT_ERR_STACK, "traversal stack overflow");
  *npairs = found;
  return 0;
}

int bt_set_opening(bt_ctx *h, int mode, float budget, float theta_lo, float theta_hi) {
  if (!h) return BT_ERR_ARG;
  Ctx &c = h->c;
  if (mode != 0 && mode != 1) return fail(c, BT_ERR_ARG, "opening mode must be 0 or 1");
  if (mode == 1) {
    if (c.fp64) return fail(c, BT_ERR_STATE, "the learned angle runs on Treecode<float> only");
    if (!c.have_mlp) return fail(c, BT_ERR_STATE, "no network parameters");
    if (!c.quadrupole) return fail(c, BT_ERR_STATE, "the learned angle needs quadrupoles");
    if (!(budget > 0.f) || !(theta_lo > 0.f) || !(theta_hi > theta_lo))
      return fail(c, BT_ERR_ARG, "bad budget or theta band");
    c.learn_shift = std::log2(budget) / 3.0f;
    c.learn_theta_lo = theta_lo;
    c.learn_theta_hi = theta_hi;
  }
  c.opening = mode;
  return 0;
}

int bt_set_opening_check(bt_ctx *h, int nsamples, float p99_max, uint64_t seed) {
  if (!h) return BT_ERR_ARG;
  Ctx &c = h->c;
  if (nsamples < 0 || nsamples > 1024 || !(p99_max > 0.f))
    return fail(c, BT_ERR_ARG, "samples must be in [0, 1024], p99_max > 0");
  // no check sum of an earlier call may still run on the check's stream
  sync_streams(c);
  c.check_state = 0;
  if (nsamples > c.check_cap) {
    // captured steps hold the old buffers: drop them before the buffers go
    release_graph(c);
    c.check_cap = 0;
    if (dev_alloc(c, &c.d_check, (size_t)nsamples * (direct_slices(nsamples) + 1) * 4 + 4) ||
        dev_alloc(c, &c.d_check_idx, (size_t)nsamples * 4))
      return BT_ERR_HIP;
    c.check_cap = nsamples;
  }
  c.check_ns = nsamples;
  c.check_p99 = p99_max;
  c.check_seed = seed;
  // restart the device's check count (the sample draws) and statistics
  const int zero[4] = {0, 0, 0, 0};
  BT_CK(hipMemcpyAsync(&c.dc->check_calls, zero, sizeof(zero), hipMemcpyHostToDevice, c.stream));
  if (nsamples > 0) {
    const double z3[3] = {0.0, 0.0, 0.0};
    BT_CK(hipMemcpyAsync(check_result(c), z3, sizeof(z3), hipMemcpyHostToDevice, c.stream));
  }
  BT_CK(hipStreamSynchronize(c.stream));
  return 0;
}

int bt_get_check_samples(bt_ctx *h, int *ph_idx) {
  if (!h || !ph_idx) return BT_ERR_ARG;
  Ctx &c = h->c;
  if (c.check_ns <= 0 || !c.d_check_idx) return fail(c, BT_ERR_STATE, "no opening check set");
  const int ns = std::min(c.check_ns, c.n);
  BT_CK(d2h(c, ph_idx, c.d_check_idx + (size_t)3 * ns, sizeof(int) * ns));
  return 0;
}

int bt_get_opening_check(bt_ctx *h, double out[6]) {
  if (!h || !out) return BT_ERR_ARG;
  Ctx &c = h->c;
  for (int k = 0; k < 6; k++) out[k] = 0.0;
  if (c.check_ns > 0 && c.d_check) {
    BT_CK(hipMemcpyAsync(out, check_result(c), sizeof(double) * 3, hipMemcpyDeviceToHost,
                         c.stream));
  }
  if (read_counters(c)) return BT_ERR_HIP;
  out[3] = c.hc.fallback;
  out[4] = (double)c.hc.fallbacks;
  out[5] = (double)c.hc.checks;
  return 0;
}


// ---- captured step and autotuner -------------------------------------------

int bt_step_graph(bt_ctx *h, float dt, int nleaf, int level_split, int ncrit, int nsteps) {
  if (!h || nsteps <= 0) return BT_ERR_ARG;
  Ctx &c = h->c;
  c.err.clear();
  if (!valid_nleaf(nleaf)) return fail(c, BT_ERR_ARG, "nleaf must be 16, 24, 32, 48 or 64");
  if (ncrit < 1 || ncrit > 64) return fail(c, BT_ERR_ARG, "ncrit must be in [1, 64]");
  if (!c.have_acc) return fail(c, BT_ERR_STATE, "compute forces (or one bt_step) first");
  GraphKey key;
  std::memset(&key, 0, sizeof(key));   // padding too: keys are compared bytewise
  key.dt = dt;
  key.eps = c.eps;
  key.theta = c.theta;
  key.nleaf = nleaf;
  key.level_split = std::max(1, level_split);
  key.ncrit = ncrit;
  key.opening = c.opening;
  key.kahan = c.kahan;
  key.quadrupole = c.quadrupole;
  key.walk_variant = c.walk_variant;
  key.stack_ints = c.stack_ints;
  key.blocks_per_cu = c.blocks_per_cu;
  key.lds_stack = c.lds_stack;
  key.group_mode = c.group_mode;
  key.overlap = c.overlap;
  if (c.opening == 1) {
    key.learn_shift = c.learn_shift;
    key.learn_theta_lo = c.learn_theta_lo;
    key.learn_theta_hi = c.learn_theta_hi;
    key.check_ns = c.check_ns;
    key.check_p99 = c.check_p99;
    key.check_seed = c.check_seed;
  }
  if (c.graph_exec[0] && std::memcmp(&key, &c.graph_key, sizeof(key)) != 0) release_graph(c);
  if (!c.graph_exec[0]) {
    if (ensure_stack(c)) return BT_ERR_HIP;
    c.nleaf = nleaf;
    c.ncrit = ncrit;
    c.level_split = key.level_split;
    if (join_check(c)) return BT_ERR_HIP;   // nothing forked may precede the capture
    c.check_state = 0;
    BT_CK(hipStreamSynchronize(c.stream));
    const bool timing = c.timing;
    c.timing = false;
    float4 *const pos0 = c.pos;
    int rc = 0;
    for (int g = 0; g < 2 && !rc; g++) {
      // graph g starts from the particles in buffer g (0: pos0, 1: the other)
      if (g == 1) swap_particle_buffers(c);
      if (hipStreamBeginCapture(c.stream, hipStreamCaptureModeRelaxed) != hipSuccess) {
        if (g == 1) swap_particle_buffers(c);
        rc = BT_ERR_HIP;
        break;
      }
      launch_kick_drift(c, 0.5 * (double)dt, dt);
      rc = step_second_half(c, dt);   // the build swaps the buffers on the host side
      if (hipStreamEndCapture(c.stream, &c.graph[g]) != hipSuccess) rc = BT_ERR_HIP;
      c.check_state = 0;   // captured work is not queued: nothing of it to join
      if (!rc && hipGraphInstantiate(&c.graph_exec[g], c.graph[g], nullptr, nullptr, 0) !=
                     hipSuccess)
        rc = BT_ERR_HIP;
      swap_particle_buffers(c);  // undo the capture's host-side swap
      if (g == 1) swap_particle_buffers(c);
    }
    c.timing = timing;
    if (rc || c.pos != pos0) {
      release_graph(c);
      return fail(c, BT_ERR_HIP, "step capture failed");
    }
    c.graph_pos0 = pos0;
    std::memcpy(&c.graph_key, &key, sizeof(key));
  }
  // errors of every replay stay on record (err_sticky); after one, the later
  // replays leave the particles alone (their kicks and drifts do nothing)
  launch_clear_sticky(c);
  for (int k = 0; k < nsteps; k++) {
    BT_CK(hipGraphLaunch(c.graph_exec[c.pos == c.graph_pos0 ? 0 : 1], c.stream));
    swap_particle_buffers(c);
  }
  if (read_counters(c)) return BT_ERR_HIP;
  c.have_tree = c.have_mp = c.have_groups = c.have_acc = false;
  if (c.hc.err_sticky == 1) {
    // cell capacity: steps_done steps completed, the next one stopped after
    // its drift; grow the cells, finish that step, run the rest eagerly
    const int done = c.hc.steps_done;
    int rc = regrow_cells(c);
    if (!rc) {
      launch_clear_sticky(c);
      rc = step_second_half(c, dt);
    }
    if (!rc) rc = finish_step(c, dt);
    if (rc) return rc;
    c.have_tree = c.have_mp = c.have_groups = c.have_acc = true;
    for (int k = done + 1; k < nsteps; k++) {
      rc = bt_step(h, dt, nleaf, level_split, ncrit);
      if (rc) return rc;
    }
    return 0;
  }
  if (c.hc.err_sticky & 3) {
    release_graph(c);
    return fail(c, BT_ERR_CAPACITY, kGroupCapacityMsg);
  }
  if (c.hc.err_sticky & 4) {
    c.err = "walk stack overflow in replay " + std::to_string(c.hc.steps_done + 1) + " of " +
            std::to_string(nsteps) + " (the particles hold that step's first half kick and drift)";
    return BT_ERR_STACK;
  }
  c.have_tree = c.have_mp = c.have_groups = c.have_acc = true;
  return 0;
}

// Times one step with dt = 0 (build, multipoles, groups, walk; particles do
// not move) on the GPU's own clock: returns the step's and the walk's ms.
static int timed_evaluation(bt_ctx *h, int nleaf, int level_split, int ncrit, float *total_ms,
                            float *walk_ms) {
  Ctx &c = h->c;
  const int rc = bt_step(h, 0.0f, nleaf, level_split, ncrit);
  if (rc) return rc;
  float t = 0.f;
  for (int k = 0; k < 9; k++) t += c.ms[k];
  *total_ms = t;
  *walk_ms = c.ms[7];
  return 0;
}

int bt_autotune(bt_ctx *h, int level_split, int flags, double *table, int *ntable, int best[4]) {
  if (!h || !best) return BT_ERR_ARG;
  Ctx &c = h->c;
  c.err.clear();
  if (!c.have_particles) return fail(c, BT_ERR_STATE, "no particles");
  (void)flags;
  const bool timing = c.timing;
  c.timing = true;
  static const int kLeaf[] = {16, 32, 64};
  static const int kCrit[] = {32, 48, 64};
  static const int kBpc[] = {4, 8, 12, 16};
  int nt = 0;
  float best_t = 1e30f;
  int bl = 16, bc = 64, bb = c.blocks_per_cu > 0 ? c.blocks_per_cu : 8;
  auto row = [&](int nleaf, int ncrit, int bpc, float tot, float walk) {
    if (table) {
      double *r = table + 6 * nt;
      r[0] = nleaf; r[1] = ncrit; r[2] = bpc; r[3] = tot; r[4] = walk; r[5] = c.hc.stack_hw;
    }
    nt++;
  };
  // NLEAF x NCRIT on the whole force evaluation (second run of each is timed)
  for (int nleaf : kLeaf)
    for (int ncrit : kCrit) {
      float tot = 0.f, walk = 0.f;
      int rc = timed_evaluation(h, nleaf, level_split, ncrit, &tot, &walk);
      if (!rc) rc = timed_evaluation(h, nleaf, level_split, ncrit, &tot, &walk);
      if (rc) { c.timing = timing; return rc; }
      row(nleaf, ncrit, bb, tot, walk);
      if (tot < best_t) { best_t = tot; bl = nleaf; bc = ncrit; }
    }
  // persistent-grid occupancy for the chosen NLEAF, NCRIT
  float best_w = 1e30f;
  int deepest = 0, rc = 0;
  for (int bpc : kBpc) {
    c.blocks_per_cu = bpc;
    float tot = 0.f, w = 0.f;
    rc = timed_evaluation(h, bl, level_split, bc, &tot, &w);
    if (!rc) rc = timed_evaluation(h, bl, level_split, bc, &tot, &w);
    if (rc) { c.timing = timing; return rc; }
    deepest = std::max(deepest, c.hc.stack_hw);
    row(bl, bc, bpc, tot, w);
    if (w < best_w) { best_w = w; bb = bpc; }
  }
  c.blocks_per_cu = bb;
  // traversal stack: 4x the deepest stack seen, at least 1024 entries
  const int stack = std::max(1024, ((4 * deepest + 255) / 256) * 256);
  c.stack_ints = std::min(stack, 1 << 20);
  if (ensure_stack(c)) { c.timing = timing; return BT_ERR_HIP; }
  float tot = 0.f, w = 0.f;
  rc = timed_evaluation(h, bl, level_split, bc, &tot, &w);  // forces of the chosen set-up
  c.timing = timing;
  if (rc) return rc;
  best[0] = bl;
  best[1] = bc;
  best[2] = bb;
  best[3] = c.stack_ints;
  if (ntable) *ntable = nt;
  return 0;
}

}  // extern "C"
