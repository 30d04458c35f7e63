// capi.hip -- the C ABI of libbonsai2_amd.so (include/bonsai2_amd.h).
//
// Host orchestration of the gravity step.  Each exported function replaces
// one member of the reference's Treecode<real_t> (code/Treecode.h:123-223);
// unlike the reference (process-global __device__ symbols, exit() on error,
// code/cutil.h:772-780) all state is per context and errors are returned.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <vector>

#include "bt_internal.h"

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
  dev_free(&c.d_energy);
}

void free_cell_arrays(Ctx &c) {
  dev_free(&c.cells); dev_free(&c.csize); dev_free(&c.cmono); dev_free(&c.cq0);
  dev_free(&c.cq1); dev_free(&c.cls); dev_free(&c.cscr); dev_free(&c.child_bnd);
  dev_free(&c.cwalk);
}

int alloc_cells(Ctx &c, int cap) {
  int rc = 0;
  c.cap_cells = cap;
  rc |= dev_alloc(c, &c.cells, cap);
  rc |= dev_alloc(c, &c.csize, cap);
  rc |= dev_alloc(c, &c.cmono, cap);
  rc |= dev_alloc(c, &c.cq0, cap);
  rc |= dev_alloc(c, &c.cq1, cap);
  rc |= dev_alloc(c, &c.cls, cap);
  rc |= dev_alloc(c, &c.cscr, (size_t)cap * 6);
  rc |= dev_alloc(c, &c.cwalk, (size_t)cap * 3);
  rc |= dev_alloc(c, &c.child_bnd, (size_t)cap * 8 + 8 * (size_t)c.nsys);
  return rc ? BT_ERR_HIP : 0;
}

int ensure_stack(Ctx &c) {
  const int bpc = c.blocks_per_cu > 0 ? c.blocks_per_cu : 8;
  c.walk_blocks = std::max(1, c.num_cus * bpc);
  const size_t need = (size_t)c.walk_blocks * 4 * (size_t)c.stack_ints;
  if (need > c.stack_alloc) {
    if (dev_alloc(c, &c.stack, need)) return BT_ERR_HIP;
    c.stack_alloc = need;
  }
  return 0;
}

void record(Ctx &c, int k) {
  if (c.timing) (void)hipEventRecord(c.ev[k], c.stream);
}

int read_counters(Ctx &c) {
  BT_CK(hipMemcpyAsync(&c.hc, c.dc, sizeof(DevCounters), hipMemcpyDeviceToHost, c.stream));
  BT_CK(hipStreamSynchronize(c.stream));
  return 0;
}

// ---- phases (asynchronous; no host synchronisation) ----------------------

int phase_build(Ctx &c) {
  record(c, 0);
  launch_bbox(c);
  launch_geo_keys(c);
  record(c, 1);
  radix_sort_pairs(c, c.geo_keys, c.valA, c.keyB, c.valB, c.n, 63, true);
  record(c, 2);
  launch_gather_tree(c);
  launch_levels(c);
  record(c, 3);
  BT_LAUNCH_CK();
  return 0;
}

int phase_multipoles(Ctx &c) {
  launch_multipoles(c);
  record(c, 4);
  BT_LAUNCH_CK();
  return 0;
}

int phase_groups(Ctx &c) {
  launch_ph_keys(c);
  record(c, 5);
  radix_sort_pairs(c, c.ph_keys, c.perm, c.keyB, c.valB, c.n, 63, true);
  record(c, 6);
  launch_gather_ph(c);
  launch_groups(c);
  record(c, 7);
  BT_LAUNCH_CK();
  return 0;
}

int phase_forces(Ctx &c) {
  if (ensure_stack(c)) return BT_ERR_HIP;
  launch_walk(c);
  record(c, 8);
  BT_LAUNCH_CK();
  return 0;
}

int check_build_errors(Ctx &c) {
  if (read_counters(c)) return BT_ERR_HIP;
  if (c.hc.err & 1) return fail(c, BT_ERR_CAPACITY, "cell capacity exceeded");
  return 0;
}

// Build, growing the cell arrays once if they were too small.
int build_checked(Ctx &c) {
  if (phase_build(c)) return BT_ERR_HIP;
  if (read_counters(c)) return BT_ERR_HIP;
  if (c.hc.err & 1) {
    const long long cap = 3LL * c.n + 64LL * c.nsys + 1024;
    if (cap > 0x1FFFFFFF) return fail(c, BT_ERR_CAPACITY, "cell capacity exceeded");
    free_cell_arrays(c);
    if (alloc_cells(c, (int)cap)) return BT_ERR_HIP;
    launch_levels(c);
    BT_LAUNCH_CK();
    if (check_build_errors(c)) return BT_ERR_CAPACITY;
  }
  return 0;
}

int finish_timings(Ctx &c, int last) {
  if (!c.timing) return 0;
  for (int k = 0; k < 10; k++) c.ms[k] = 0.f;
  // phases: 0 bbox+keys, 1 sort, 2 levels, 3 multipoles, 4 ph keys, 5 ph sort, 6 groups,
  // 7 walk, 8 integrate, 9 energies
  for (int k = 0; k < last && k < 10; k++) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, c.ev[k], c.ev[k + 1]) == hipSuccess) c.ms[k] = ms;
  }
  return 0;
}

bool valid_nleaf(int nleaf) {
  return nleaf == 16 || nleaf == 24 || nleaf == 32 || nleaf == 48 || nleaf == 64;
}

}  // namespace

extern "C" {

int bt_create(bt_ctx **out, float eps, float theta, int ncrit) {
  if (!out) return BT_ERR_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return BT_ERR_NODEVICE;
  bt_ctx *h = new bt_ctx();
  Ctx &c = h->c;
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
  for (int k = 0; k < 12; k++) (void)hipEventCreate(&c.ev[k]);
  if (dev_alloc(c, &c.dc, 1) || dev_alloc(c, &c.level_beg, kLevels) ||
      dev_alloc(c, &c.d_wstats, 4) || dev_alloc(c, &c.d_wpot, 1) ||
      dev_alloc(c, &c.lvl_blk, kLevelBlocks)) {
    delete h;
    return BT_ERR_HIP;
  }
  std::memset(&c.hc, 0, sizeof(c.hc));
  (void)hipMemcpy(c.dc, &c.hc, sizeof(DevCounters), hipMemcpyHostToDevice);
  std::vector<int2> z(kLevels, make_int2(0, 0));
  (void)hipMemcpy(c.level_beg, z.data(), sizeof(int2) * kLevels, hipMemcpyHostToDevice);
  *out = h;
  return 0;
}

int bt_destroy(bt_ctx *h) {
  if (!h) return 0;
  Ctx &c = h->c;
  (void)hipStreamSynchronize(c.stream);
  free_particle_arrays(c);
  free_cell_arrays(c);
  dev_free(&c.stack);
  dev_free(&c.dc);
  dev_free(&c.level_beg);
  dev_free(&c.d_wstats);
  dev_free(&c.d_wpot);
  dev_free(&c.lvl_blk);
  dev_free(&c.sort_hist);
  dev_free(&c.scan_blk);
  dev_free(&c.d_direct);
  for (int k = 0; k < 12; k++)
    if (c.ev[k]) (void)hipEventDestroy(c.ev[k]);
  if (c.own_stream && c.stream) (void)hipStreamDestroy(c.stream);
  delete h;
  return 0;
}

const char *bt_last_error(const bt_ctx *h) { return h ? h->c.err.c_str() : "null context"; }

int bt_set_stream(bt_ctx *h, void *s) {
  if (!h) return BT_ERR_ARG;
  Ctx &c = h->c;
  (void)hipStreamSynchronize(c.stream);
  if (c.own_stream && c.stream) (void)hipStreamDestroy(c.stream);
  c.own_stream = false;
  c.stream = (hipStream_t)s;
  return 0;
}

int bt_set_params(bt_ctx *h, float eps, float theta) {
  if (!h) return BT_ERR_ARG;
  h->c.eps = eps;
  if (h->c.theta != theta) h->c.have_mp = false;  // theta is baked into cellOp^2
  h->c.theta = theta;
  return 0;
}

int bt_set_walk_options(bt_ctx *h, int kahan, int stack_ints, int blocks_per_cu) {
  if (!h) return BT_ERR_ARG;
  Ctx &c = h->c;
  if (stack_ints < 256 || blocks_per_cu < 0 || blocks_per_cu > 64)
    return fail(c, BT_ERR_ARG, "bad walk options");
  c.kahan = kahan ? 1 : 0;
  c.stack_ints = stack_ints;
  c.blocks_per_cu = blocks_per_cu;
  return 0;
}

int bt_set_walk_variant(bt_ctx *h, int variant) {
  if (!h) return BT_ERR_ARG;
  if (variant < 0 || variant > 1) return fail(h->c, BT_ERR_ARG, "walk variant must be 0 or 1");
  h->c.walk_variant = variant;
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
  (void)hipStreamSynchronize(c.stream);
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
    rc |= dev_alloc(c, &c.pos, n); rc |= dev_alloc(c, &c.vel, n);
    rc |= dev_alloc(c, &c.pos2, n); rc |= dev_alloc(c, &c.vel2, n);
    rc |= dev_alloc(c, &c.id, n); rc |= dev_alloc(c, &c.id2, n);
    rc |= dev_alloc(c, &c.keyB, n); rc |= dev_alloc(c, &c.valA, n);
    rc |= dev_alloc(c, &c.valB, n); rc |= dev_alloc(c, &c.geo_keys, n);
    rc |= dev_alloc(c, &c.pos_ph, n); rc |= dev_alloc(c, &c.perm, n);
    rc |= dev_alloc(c, &c.id_ph, n); rc |= dev_alloc(c, &c.ph_keys, n);
    rc |= dev_alloc(c, &c.acc, (size_t)n + 64);  // + spare slots for idle walk lanes
    c.cap_groups = n;
    rc |= dev_alloc(c, &c.groups, n); rc |= dev_alloc(c, &c.group_sys, n);
    rc |= dev_alloc(c, &c.gtmp0, n); rc |= dev_alloc(c, &c.gtmp1, n);
    rc |= dev_alloc(c, &c.gcounts, n); rc |= dev_alloc(c, &c.gpot, n);
    rc |= dev_alloc(c, &c.box, nsys); rc |= dev_alloc(c, &c.root_cells, nsys);
    rc |= dev_alloc(c, &c.d_off, nsys + 1);
    rc |= dev_alloc(c, &c.d_energy, 2 * (size_t)nsys * 65);
    if (rc) return BT_ERR_HIP;
    if (alloc_cells(c, n + 1024 * nsys + 64)) return BT_ERR_HIP;
  }
  c.bbox_blocks = std::max(1, std::min(1024, (c.max_sys_n + 16383) / 16384));
  if (dev_alloc(c, &c.bbox_part, 2 * (size_t)nsys * c.bbox_blocks)) return BT_ERR_HIP;
  BT_CK(hipMemcpy(c.d_off, off.data(), sizeof(int) * (nsys + 1), hipMemcpyHostToDevice));
  c.have_particles = c.have_tree = c.have_mp = c.have_groups = c.have_acc = false;
  return 0;
}

int bt_set_particles(bt_ctx *h, const float *pos4, const float *vel4) {
  if (!h || !pos4 || !vel4) return BT_ERR_ARG;
  Ctx &c = h->c;
  if (c.n <= 0) return fail(c, BT_ERR_STATE, "bt_alloc first");
  BT_CK(hipMemcpyAsync(c.pos, pos4, sizeof(float4) * c.n, hipMemcpyHostToDevice, c.stream));
  BT_CK(hipMemcpyAsync(c.vel, vel4, sizeof(float4) * c.n, hipMemcpyHostToDevice, c.stream));
  std::vector<int> ids(c.n);
  for (int i = 0; i < c.n; i++) ids[i] = i;
  BT_CK(hipMemcpyAsync(c.id, ids.data(), sizeof(int) * c.n, hipMemcpyHostToDevice, c.stream));
  BT_CK(hipStreamSynchronize(c.stream));
  c.have_particles = true;
  c.have_tree = c.have_mp = c.have_groups = c.have_acc = false;
  return 0;
}

int bt_set_particles_device(bt_ctx *h, const void *dpos, const void *dvel) {
  if (!h || !dpos || !dvel) return BT_ERR_ARG;
  Ctx &c = h->c;
  if (c.n <= 0) return fail(c, BT_ERR_STATE, "bt_alloc first");
  BT_CK(hipMemcpyAsync(c.pos, dpos, sizeof(float4) * c.n, hipMemcpyDeviceToDevice, c.stream));
  BT_CK(hipMemcpyAsync(c.vel, dvel, sizeof(float4) * c.n, hipMemcpyDeviceToDevice, c.stream));
  std::vector<int> ids(c.n);
  for (int i = 0; i < c.n; i++) ids[i] = i;
  BT_CK(hipMemcpyAsync(c.id, ids.data(), sizeof(int) * c.n, hipMemcpyHostToDevice, c.stream));
  BT_CK(hipStreamSynchronize(c.stream));
  c.have_particles = true;
  c.have_tree = c.have_mp = c.have_groups = c.have_acc = false;
  return 0;
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
  if (c.hc.err & 2) return fail(c, BT_ERR_CAPACITY, "group capacity exceeded");
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
  record(c, 7);
  if (phase_forces(c)) return BT_ERR_HIP;
  if (read_counters(c)) return BT_ERR_HIP;
  if (c.hc.err & 4) return fail(c, BT_ERR_STACK, "walk stack overflow");
  c.have_acc = true;
  if (intcount && stats) {
    unsigned long long w[4];
    BT_CK(hipMemcpy(w, c.d_wstats, sizeof(w), hipMemcpyDeviceToHost));
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
  launch_kick(c, dth);
  BT_LAUNCH_CK();
  BT_CK(hipStreamSynchronize(c.stream));
  return 0;
}

int bt_move_particles(bt_ctx *h, float dt) {
  if (!h) return BT_ERR_ARG;
  Ctx &c = h->c;
  if (!c.have_acc) return fail(c, BT_ERR_STATE, "computeForces first");
  launch_kick(c, 0.5f * dt);
  launch_drift(c, dt);
  BT_LAUNCH_CK();
  BT_CK(hipStreamSynchronize(c.stream));
  // positions moved: the tree, moments, groups and forces are stale
  c.have_tree = c.have_mp = c.have_groups = c.have_acc = false;
  return 0;
}

int bt_compute_energies(bt_ctx *h, double *ekin, double *epot) {
  if (!h) return BT_ERR_ARG;
  Ctx &c = h->c;
  if (!c.have_acc) return fail(c, BT_ERR_STATE, "computeForces first (phi is needed)");
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
  launch_kick(c, 0.5f * dt);
  launch_drift(c, dt);
  if (phase_build(c) || phase_multipoles(c) || phase_groups(c) || phase_forces(c))
    return BT_ERR_HIP;
  launch_kick(c, 0.5f * dt);
  record(c, 9);
  BT_LAUNCH_CK();
  if (read_counters(c)) return BT_ERR_HIP;
  c.have_tree = c.have_mp = c.have_groups = c.have_acc = false;
  if (c.hc.err & 1) return fail(c, BT_ERR_CAPACITY, "cell capacity exceeded");
  if (c.hc.err & 2) return fail(c, BT_ERR_CAPACITY, "group capacity exceeded");
  if (c.hc.err & 4) return fail(c, BT_ERR_STACK, "walk stack overflow");
  c.have_tree = c.have_mp = c.have_groups = c.have_acc = true;
  finish_timings(c, 9);
  return 0;
}

// ---- results --------------------------------------------------------------

int bt_get_acc(bt_ctx *h, float *acc4, int order) {
  if (!h || !acc4) return BT_ERR_ARG;
  Ctx &c = h->c;
  if (!c.have_acc) return fail(c, BT_ERR_STATE, "no forces");
  std::vector<float4> a(c.n);
  BT_CK(hipMemcpyAsync(a.data(), c.acc, sizeof(float4) * c.n, hipMemcpyDeviceToHost, c.stream));
  if (order == BT_ORDER_PH) {
    BT_CK(hipStreamSynchronize(c.stream));
    std::memcpy(acc4, a.data(), sizeof(float4) * c.n);
    return 0;
  }
  std::vector<int> map(c.n);
  if (order == BT_ORDER_TREE)
    BT_CK(hipMemcpyAsync(map.data(), c.perm, sizeof(int) * c.n, hipMemcpyDeviceToHost, c.stream));
  else if (order == BT_ORDER_INPUT)
    BT_CK(hipMemcpyAsync(map.data(), c.id_ph, sizeof(int) * c.n, hipMemcpyDeviceToHost, c.stream));
  else
    return fail(c, BT_ERR_ARG, "bad order");
  BT_CK(hipStreamSynchronize(c.stream));
  float4 *o = reinterpret_cast<float4 *>(acc4);
  for (int k = 0; k < c.n; k++) o[map[k]] = a[k];
  return 0;
}

int bt_get_particles(bt_ctx *h, float *pos4, float *vel4, int order) {
  if (!h) return BT_ERR_ARG;
  Ctx &c = h->c;
  if (!c.have_particles) return fail(c, BT_ERR_STATE, "no particles");
  std::vector<float4> p(c.n), v(c.n);
  std::vector<int> id(c.n);
  BT_CK(hipMemcpyAsync(p.data(), c.pos, sizeof(float4) * c.n, hipMemcpyDeviceToHost, c.stream));
  BT_CK(hipMemcpyAsync(v.data(), c.vel, sizeof(float4) * c.n, hipMemcpyDeviceToHost, c.stream));
  BT_CK(hipMemcpyAsync(id.data(), c.id, sizeof(int) * c.n, hipMemcpyDeviceToHost, c.stream));
  std::vector<int> perm;
  if (order == BT_ORDER_PH) {
    if (!c.have_groups) return fail(c, BT_ERR_STATE, "no PH order before makeGroups");
    perm.resize(c.n);
    BT_CK(hipMemcpyAsync(perm.data(), c.perm, sizeof(int) * c.n, hipMemcpyDeviceToHost,
                         c.stream));
  }
  BT_CK(hipStreamSynchronize(c.stream));
  float4 *po = reinterpret_cast<float4 *>(pos4);
  float4 *vo = reinterpret_cast<float4 *>(vel4);
  for (int i = 0; i < c.n; i++) {
    int dst = i;
    if (order == BT_ORDER_INPUT) dst = id[i];
    if (order == BT_ORDER_PH) continue;
    if (po) po[dst] = p[i];
    if (vo) vo[dst] = v[i];
  }
  if (order == BT_ORDER_PH)
    for (int k = 0; k < c.n; k++) {
      if (po) po[k] = p[perm[k]];
      if (vo) vo[k] = v[perm[k]];
    }
  if (order != BT_ORDER_INPUT && order != BT_ORDER_TREE && order != BT_ORDER_PH)
    return fail(c, BT_ERR_ARG, "bad order");
  return 0;
}

int bt_get_walk_stats(bt_ctx *h, double stats[4], double totals2[2]) {
  if (!h || !stats) return BT_ERR_ARG;
  Ctx &c = h->c;
  if (!c.have_acc) return fail(c, BT_ERR_STATE, "no forces");
  unsigned long long w[4];
  BT_CK(hipMemcpy(w, c.d_wstats, sizeof(w), hipMemcpyDeviceToHost));
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
  BT_CK(hipMemcpy(pot, c.d_wpot, sizeof(double), hipMemcpyDeviceToHost));
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
  if (read_counters(c)) return BT_ERR_HIP;
  const size_t nc = (size_t)c.hc.ncells;
  if (cells4) BT_CK(hipMemcpy(cells4, c.cells, sizeof(uint4) * nc, hipMemcpyDeviceToHost));
  if (level_range64)
    BT_CK(hipMemcpy(level_range64, c.level_beg, sizeof(int2) * kLevels, hipMemcpyDeviceToHost));
  if ((size4 || mono4 || q0_4 || q1_2) && !c.have_mp)
    return fail(c, BT_ERR_STATE, "no multipoles");
  if (size4) BT_CK(hipMemcpy(size4, c.csize, sizeof(float4) * nc, hipMemcpyDeviceToHost));
  if (mono4) BT_CK(hipMemcpy(mono4, c.cmono, sizeof(float4) * nc, hipMemcpyDeviceToHost));
  if (q0_4) BT_CK(hipMemcpy(q0_4, c.cq0, sizeof(float4) * nc, hipMemcpyDeviceToHost));
  if (q1_2) BT_CK(hipMemcpy(q1_2, c.cq1, sizeof(float2) * nc, hipMemcpyDeviceToHost));
  return 0;
}

int bt_get_cell_ls(bt_ctx *h, float *ls2) {
  if (!h || !ls2) return BT_ERR_ARG;
  Ctx &c = h->c;
  if (!c.have_mp) return fail(c, BT_ERR_STATE, "no multipoles");
  if (read_counters(c)) return BT_ERR_HIP;
  BT_CK(hipMemcpy(ls2, c.cls, sizeof(float2) * (size_t)c.hc.ncells, hipMemcpyDeviceToHost));
  return 0;
}

int bt_get_groups(bt_ctx *h, int *groups2) {
  if (!h || !groups2) return BT_ERR_ARG;
  Ctx &c = h->c;
  if (!c.have_groups) return fail(c, BT_ERR_STATE, "no groups");
  if (read_counters(c)) return BT_ERR_HIP;
  BT_CK(hipMemcpy(groups2, c.groups, sizeof(int2) * (size_t)c.hc.ngroups,
                  hipMemcpyDeviceToHost));
  return 0;
}

int bt_get_group_counts(bt_ctx *h, int *counts2) {
  if (!h || !counts2) return BT_ERR_ARG;
  Ctx &c = h->c;
  if (!c.have_acc) return fail(c, BT_ERR_STATE, "no forces");
  if (read_counters(c)) return BT_ERR_HIP;
  BT_CK(hipMemcpy(counts2, c.gcounts, sizeof(int2) * (size_t)c.hc.ngroups,
                  hipMemcpyDeviceToHost));
  return 0;
}

int bt_get_ph(bt_ctx *h, uint64_t *keys, int *perm) {
  if (!h) return BT_ERR_ARG;
  Ctx &c = h->c;
  if (!c.have_groups) return fail(c, BT_ERR_STATE, "no groups");
  if (keys) BT_CK(hipMemcpy(keys, c.ph_keys, sizeof(uint64_t) * c.n, hipMemcpyDeviceToHost));
  if (perm) BT_CK(hipMemcpy(perm, c.perm, sizeof(int) * c.n, hipMemcpyDeviceToHost));
  return 0;
}

int bt_get_ids(bt_ctx *h, int *ids, int order) {
  if (!h || !ids) return BT_ERR_ARG;
  Ctx &c = h->c;
  if (order == BT_ORDER_TREE) {
    if (!c.have_tree) return fail(c, BT_ERR_STATE, "no tree");
    BT_CK(hipMemcpy(ids, c.id, sizeof(int) * c.n, hipMemcpyDeviceToHost));
  } else if (order == BT_ORDER_PH) {
    if (!c.have_groups) return fail(c, BT_ERR_STATE, "no groups");
    BT_CK(hipMemcpy(ids, c.id_ph, sizeof(int) * c.n, hipMemcpyDeviceToHost));
  } else {
    for (int i = 0; i < c.n; i++) ids[i] = i;
  }
  return 0;
}

int bt_get_boxes(bt_ctx *h, float *boxes4) {
  if (!h || !boxes4) return BT_ERR_ARG;
  Ctx &c = h->c;
  if (!c.have_tree) return fail(c, BT_ERR_STATE, "no tree");
  BT_CK(hipMemcpy(boxes4, c.box, sizeof(float4) * c.nsys, hipMemcpyDeviceToHost));
  return 0;
}

int bt_direct_fp64(bt_ctx *h, int ns, const int *ph_idx, double *out4) {
  if (!h || ns <= 0 || !ph_idx || !out4) return BT_ERR_ARG;
  Ctx &c = h->c;
  if (!c.have_groups) return fail(c, BT_ERR_STATE, "makeGroups first (targets are PH indices)");
  for (int k = 0; k < ns; k++)
    if (ph_idx[k] < 0 || ph_idx[k] >= c.n) return fail(c, BT_ERR_ARG, "target index out of range");
  const size_t need = (size_t)ns * 65 * 4 + (size_t)ns;  // doubles (+ room for the indices)
  if ((int)need > c.direct_cap || !c.d_direct) {
    if (dev_alloc(c, &c.d_direct, need)) return BT_ERR_HIP;
    c.direct_cap = (int)need;
  }
  int *d_idx = reinterpret_cast<int *>(c.d_direct + (size_t)ns * 65 * 4);
  BT_CK(hipMemcpyAsync(d_idx, ph_idx, sizeof(int) * ns, hipMemcpyHostToDevice, c.stream));
  launch_direct(c, ns, d_idx, c.d_direct);
  BT_LAUNCH_CK();
  BT_CK(hipMemcpyAsync(out4, c.d_direct, sizeof(double) * 4 * ns, hipMemcpyDeviceToHost,
                       c.stream));
  BT_CK(hipStreamSynchronize(c.stream));
  return 0;
}

int bt_get_timings(bt_ctx *h, float ms[10]) {
  if (!h || !ms) return BT_ERR_ARG;
  for (int k = 0; k < 10; k++) ms[k] = h->c.ms[k];
  return 0;
}

}  // extern "C"
