// k_groups.hip -- makeGroups on CDNA4 (reference code/makeGroups.cu:236-341).
//
// PH keys of the tree-ordered particles (code/makeGroups.cu:108-135, the
// Hilbert state machine of :19-106), a stable radix sort carrying the tree
// index (:270-277), the gather into PH order (:280), then groups: every run
// of equal keys masked to the top level_split octal digits (:137-192) is cut
// into chunks of ncrit (:196-223).  The reference finds run bounds with two
// inclusive_scan_by_key calls and takes group slots from an atomic counter;
// here a max-scan gives each particle its run start and a sum-scan over the
// "first particle of a group" flags gives deterministic group slots in PH
// order.
#include "bt_device.h"
#include "bt_internal.h"

namespace bt {

namespace {

constexpr int kBlock = 256;

// Output digit per Hilbert index (x<<2 | y<<1 | z): {0,1,7,6,3,2,4,5}, packed
// 3 bits per entry.
constexpr uint32_t kHilbertDigits = (0u << 0) | (1u << 3) | (7u << 6) | (6u << 9) | (3u << 12) |
                                    (2u << 15) | (4u << 18) | (5u << 21);

__device__ __forceinline__ uint64_t hilbert_key21(int x, int y, int z) {
  uint64_t key = 0;
#pragma unroll
  for (int lev = 20; lev >= 0; lev--) {
    const int bit = 1 << lev;
    const int idx = ((x & bit) ? 4 : 0) | ((y & bit) ? 2 : 0) | ((z & bit) ? 1 : 0);
    int nx = x, ny = y, nz = z;
    if (idx == 0) {               // swap y and z
      ny = z; nz = y;
    } else if (idx == 1 || idx == 5) {  // swap x and y
      nx = y; ny = x;
    } else if (idx == 4 || idx == 6) {  // complement x and z
      nx = ~x; nz = ~z;
    } else if (idx == 3 || idx == 7) {  // x <- ~y, y <- ~x
      nx = ~y; ny = ~x;
    } else {                      // idx 2: y <- ~z, z <- ~y
      ny = ~z; nz = ~y;
    }
    x = nx; y = ny; z = nz;
    key = (key << 3) + ((kHilbertDigits >> (3 * idx)) & 7u);
  }
  return key;
}

template <typename R>
__global__ __launch_bounds__(kBlock) void k_ph_keys(int n, int nsys, const int *__restrict__ off,
                                                    const V4<R> *__restrict__ box,
                                                    const V4<R> *__restrict__ pos,
                                                    uint64_t *__restrict__ keys,
                                                    int *__restrict__ vals) {
  BT_BUILD_PRIO();
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    const int r = nsys > 1 ? system_of(i, nsys, off) : 0;
    const V4<R> b = box[r];
    const V4<R> p = pos[i];
    int xc, yc, zc;
    {
#pragma clang fp contract(off)
      // code/makeGroups.cu:123-131, evaluated left to right in real_t
      const R inv = R(0.5) / b.w;
      const R bx = b.x - b.w, by = b.y - b.w, bz = b.z - b.w;
      const R scale = R(2097152);  // 1 << 21
      xc = (int)((p.x - bx) * inv * scale);
      yc = (int)((p.y - by) * inv * scale);
      zc = (int)((p.z - bz) * inv * scale);
    }
    keys[i] = hilbert_key21(xc, yc, zc);
    vals[i] = i;
  }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_gather_ph(int n, const int *__restrict__ perm,
                                                      const T *__restrict__ pos,
                                                      const int *__restrict__ id,
                                                      T *__restrict__ pos_ph,
                                                      int *__restrict__ id_ph) {
  BT_BUILD_PRIO();
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    const int j = perm[i];
    pos_ph[i] = pos[j];
    id_ph[i] = id[j];
  }
}

// v[i] = i at the start of a run of equal masked keys (or of a system), else 0.
__global__ __launch_bounds__(kBlock) void k_run_starts(int n, int nsys, const int *__restrict__ off,
                                                       const uint64_t *__restrict__ keys,
                                                       uint64_t mask, int *__restrict__ v) {
  BT_BUILD_PRIO();
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    bool start = i == 0;
    if (!start) {
      start = (keys[i] & mask) != (keys[i - 1] & mask);
      if (!start && nsys > 1) start = system_of(i, nsys, off) != system_of(i - 1, nsys, off);
    }
    v[i] = start ? i : 0;
  }
}

__global__ __launch_bounds__(kBlock) void k_group_flags(int n, int ncrit,
                                                        const int *__restrict__ run_start,
                                                        int *__restrict__ flag) {
  BT_BUILD_PRIO();
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
    flag[i] = ((i - run_start[i]) % ncrit) == 0 ? 1 : 0;
}

__global__ __launch_bounds__(kBlock) void k_group_emit(int n, int nsys, const int *__restrict__ off,
                                                       const int *__restrict__ flag,
                                                       const int *__restrict__ slot,
                                                       int cap, int2 *__restrict__ groups,
                                                       int *__restrict__ group_sys,
                                                       DevCounters *dc) {
  BT_BUILD_PRIO();
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    if (!flag[i]) continue;
    const int g = slot[i];
    if (g >= cap) {
      { atomicOr(&dc->err, 2); atomicOr(&dc->err_sticky, 2); }
      continue;
    }
    groups[g] = make_int2(i, 0);
    group_sys[g] = nsys > 1 ? system_of(i, nsys, off) : 0;
  }
}

// Groups tile [0, n) in order, so a group ends where the next begins.
__global__ __launch_bounds__(kBlock) void k_group_sizes(int n, int cap, int2 *__restrict__ groups,
                                                        DevCounters *dc) {
  BT_BUILD_PRIO();
  const int ng = min(dc->ngroups, cap);
  for (int g = blockIdx.x * kBlock + threadIdx.x; g < ng; g += gridDim.x * kBlock) {
    const int next = g + 1 < ng ? groups[g + 1].x : n;
    groups[g].y = next - groups[g].x;
  }
}

}  // namespace

void launch_ph_keys(Ctx &c) {
  const dim3 g(grid_for(c.n, kBlock, 8192));
  if (c.fp64)
    hipLaunchKernelGGL(k_ph_keys<double>, g, dim3(kBlock), 0, c.stream, c.n, c.nsys, c.d_off,
                       as4<double>(c.box), as4<double>(c.pos), c.ph_keys, c.perm);
  else
    hipLaunchKernelGGL(k_ph_keys<float>, g, dim3(kBlock), 0, c.stream, c.n, c.nsys, c.d_off, c.box,
                       c.pos, c.ph_keys, c.perm);
}

void launch_gather_ph(Ctx &c) {
  const dim3 g(grid_for(c.n, kBlock, 8192));
  if (c.fp64)
    hipLaunchKernelGGL(k_gather_ph<double4>, g, dim3(kBlock), 0, c.stream, c.n, c.perm,
                       as4<double>(c.pos), c.id, as4<double>(c.pos_ph), c.id_ph);
  else
    hipLaunchKernelGGL(k_gather_ph<float4>, g, dim3(kBlock), 0, c.stream, c.n, c.perm, c.pos,
                       c.id, c.pos_ph, c.id_ph);
}

// A new set of groups starts without the group error bit.
__global__ void k_groups_init(DevCounters *dc) {
  if (threadIdx.x == 0) atomicAnd(&dc->err, ~2);   // the level build may run beside this
}

// Tree groups: flag the first particle of every group cell.  ncrit = 0: the
// leaves (they tile the particles in tree order).  ncrit > 0: the cells of at
// most ncrit particles whose parent holds more (level-1 cells have no parent
// cell), and the leaves of more than ncrit particles; every particle lies in
// exactly one of them.  A group of more than 64 particles (only a leaf capped
// at the deepest level can be one) cannot be one wave's group: err 2.
__global__ __launch_bounds__(kBlock) void k_tree_group_starts(int cap_cells, int ncrit,
                                                              const uint4 *__restrict__ cells,
                                                              int *__restrict__ flag,
                                                              DevCounters *dc) {
  const int nc = min(dc->ncells, cap_cells);
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < nc; i += gridDim.x * kBlock) {
    const uint4 c = cells[i];
    const int np = (int)c.w - (int)c.z;
    bool group;
    if (ncrit == 0) {
      group = cell_leaf(c);
    } else {
      const int level = (int)(c.x >> 27);
      bool parent_splits = true;
      if (level > 1) {
        const uint4 p = cells[c.x & 0x07FFFFFFu];
        parent_splits = (int)p.w - (int)p.z > ncrit;
      }
      group = (cell_leaf(c) || np <= ncrit) && parent_splits;
    }
    if (!group) continue;
    flag[c.z] = 1;
    if (np > kWave) { atomicOr(&dc->err, 2); atomicOr(&dc->err_sticky, 2); }
  }
}

// The group order of leaf groups is the tree order itself.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_identity_order(int n, const T *__restrict__ pos,
                                                           const int *__restrict__ id,
                                                           int *__restrict__ perm,
                                                           T *__restrict__ pos_g,
                                                           int *__restrict__ id_g) {
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    perm[i] = i;
    pos_g[i] = pos[i];
    id_g[i] = id[i];
  }
}

void launch_leaf_groups(Ctx &c) {
  const int g = grid_for(c.n, kBlock, 8192);
  hipLaunchKernelGGL(k_groups_init, dim3(1), dim3(64), 0, c.stream, c.dc);
  if (c.fp64)
    hipLaunchKernelGGL(k_identity_order<double4>, dim3(g), dim3(kBlock), 0, c.stream, c.n,
                       as4<double>(c.pos), c.id, c.perm, as4<double>(c.pos_ph), c.id_ph);
  else
    hipLaunchKernelGGL(k_identity_order<float4>, dim3(g), dim3(kBlock), 0, c.stream, c.n, c.pos,
                       c.id, c.perm, c.pos_ph, c.id_ph);
  fill_int(c, c.gtmp0, 0, c.n);
  hipLaunchKernelGGL(k_tree_group_starts, dim3(grid_for(c.cap_cells, kBlock, 8192)), dim3(kBlock),
                     0, c.stream, c.cap_cells, c.group_mode == 2 ? c.ncrit : 0, c.cells, c.gtmp0,
                     c.dc);
  scan_exclusive_sum(c, c.gtmp0, c.gtmp1, c.n, &c.dc->ngroups);
  hipLaunchKernelGGL(k_group_emit, dim3(g), dim3(kBlock), 0, c.stream, c.n, c.nsys, c.d_off,
                     c.gtmp0, c.gtmp1, c.cap_groups, c.groups, c.group_sys, c.dc);
  hipLaunchKernelGGL(k_group_sizes, dim3(grid_for(c.cap_groups, kBlock, 4096)), dim3(kBlock), 0,
                     c.stream, c.n, c.cap_groups, c.groups, c.dc);
}

void launch_groups(Ctx &c) {
  const int ls = c.level_split < 1 ? 1 : c.level_split;
  uint64_t mask = 0;  // top ls octal digits of the 63-bit key (code/makeGroups.cu:259-266)
  for (int i = 0; i < 21; i++) {
    mask <<= 3;
    if (i < ls) mask |= 7;
  }
  const int g = grid_for(c.n, kBlock, 8192);
  hipLaunchKernelGGL(k_groups_init, dim3(1), dim3(64), 0, c.stream, c.dc);
  hipLaunchKernelGGL(k_run_starts, dim3(g), dim3(kBlock), 0, c.stream, c.n, c.nsys, c.d_off,
                     c.ph_keys, mask, c.gtmp0);
  scan_inclusive_max(c, c.gtmp0, c.gtmp1, c.n);
  hipLaunchKernelGGL(k_group_flags, dim3(g), dim3(kBlock), 0, c.stream, c.n, c.ncrit, c.gtmp1,
                     c.gtmp0);
  scan_exclusive_sum(c, c.gtmp0, c.gtmp1, c.n, &c.dc->ngroups);
  hipLaunchKernelGGL(k_group_emit, dim3(g), dim3(kBlock), 0, c.stream, c.n, c.nsys, c.d_off,
                     c.gtmp0, c.gtmp1, c.cap_groups, c.groups, c.group_sys, c.dc);
  hipLaunchKernelGGL(k_group_sizes, dim3(grid_for(c.cap_groups, kBlock, 4096)), dim3(kBlock), 0,
                     c.stream, c.n, c.cap_groups, c.groups, c.dc);
}

}  // namespace bt
