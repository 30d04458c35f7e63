// k_tree.hip -- buildTree and computeMultipoles on CDNA4.
//
// The reference builds the octree by recursive dynamic parallelism
// (code/buildTree.cu:170-584: every split cell's last block launches the
// children's grid from the device).  HIP has no device-side launch, so the
// build here is level-synchronous and key-based:
//
//   1. root box per system (code/buildTree.cu:80-166), bit-exact float ops;
//   2. per particle, its geometric key: the 21 octant digits it takes down the
//      same float ChildBox/Octant recursion the reference evaluates per cell
//      (code/Particle4.h:129-148, code/buildTree.cu:233-247);
//   3. a stable radix sort of the keys = the reference's scatter into child
//      ranges in octant order (:444-471), within-leaf order canonical;
//   4. cells, level by level: the children of a level-(L-1) cell are the runs
//      of equal digit L inside its particle range; count > NLEAF -> node,
//      0 < count <= NLEAF -> leaf (:394-395); children of one parent are
//      [nodes in octant order][leaves in octant order] (:393-400, :549) and
//      levels are contiguous, parents in order (the level sort, :932-951).
//      Each level is three fixed-size launches (count, scan, emit) whose work
//      sizes are read on the device, so the whole build can be graph-captured.
//
// Multipoles (code/computeMultipoles.cu:81-190) are computed bottom-up, one
// launch per level from the deepest: a leaf sums its particles, a node
// combines its children with the parallel-axis theorem.  Moments are carried
// relative to a reference point inside the cell's particle bbox (a leaf's
// first particle; a node's first child's reference), so every offset is at
// most the particle extent l and the FP32 sums keep their digits for small
// cells far from the origin and for sparse cells; levels <= kFp64Levels
// combine in FP64.  Outputs keep the reference layout (:180-188).
#include "bt_device.h"
#include "bt_internal.h"

namespace bt {

namespace {

constexpr int kBlock = 256;

// ---- root box -----------------------------------------------------------

template <typename R>
struct Ext {
  R nx, ny, nz, xx, xy, xz;   // min and max per axis
};

template <typename R>
__device__ __forceinline__ void ext_add(Ext<R> &e, R x, R y, R z) {
  e.nx = rmin(e.nx, x); e.ny = rmin(e.ny, y); e.nz = rmin(e.nz, z);
  e.xx = rmax(e.xx, x); e.xy = rmax(e.xy, y); e.xz = rmax(e.xz, z);
}
template <typename R>
__device__ __forceinline__ Ext<R> ext_empty() {
  return Ext<R>{R(1e10), R(1e10), R(1e10), R(-1e10), R(-1e10), R(-1e10)};
}
// Union of two extents, either of which may be empty (min > max): minima with
// minima, maxima with maxima.
template <typename R>
__device__ __forceinline__ void ext_merge(Ext<R> &e, R nx, R ny, R nz, R xx, R xy, R xz) {
  e.nx = rmin(e.nx, nx); e.ny = rmin(e.ny, ny); e.nz = rmin(e.nz, nz);
  e.xx = rmax(e.xx, xx); e.xy = rmax(e.xy, xy); e.xz = rmax(e.xz, xz);
}

// block-wide min/max; the block's result in thread 0
template <typename R>
__device__ __forceinline__ void block_minmax(Ext<R> &e, Ext<R> *sh) {
  e.nx = wave_min_r(e.nx); e.ny = wave_min_r(e.ny); e.nz = wave_min_r(e.nz);
  e.xx = wave_max_r(e.xx); e.xy = wave_max_r(e.xy); e.xz = wave_max_r(e.xz);
  const int w = threadIdx.x >> 6;
  if (lane_id() == 0) sh[w] = e;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < (int)(blockDim.x >> 6); k++)
      ext_merge(e, sh[k].nx, sh[k].ny, sh[k].nz, sh[k].xx, sh[k].xy, sh[k].xz);
  }
}

template <typename R>
__global__ __launch_bounds__(kBlock) void k_bbox_partial(const V4<R> *__restrict__ pos,
                                                         const int *__restrict__ off, int nb,
                                                         V4<R> *__restrict__ part) {
  BT_BUILD_PRIO();
  __shared__ Ext<R> sh[kBlock / kWave];
  const int r = blockIdx.y;
  const int b = off[r], e = off[r + 1];
  Ext<R> x = ext_empty<R>();
  for (int i = b + blockIdx.x * kBlock + threadIdx.x; i < e; i += nb * kBlock) {
    const V4<R> p = pos[i];
    ext_add(x, p.x, p.y, p.z);
  }
  block_minmax(x, sh);
  if (threadIdx.x == 0) {
    part[2 * (r * nb + blockIdx.x)] = make_v4<R>(x.nx, x.ny, x.nz, R(0));
    part[2 * (r * nb + blockIdx.x) + 1] = make_v4<R>(x.xx, x.xy, x.xz, R(0));
  }
}

// Box rules of code/buildTree.cu:145-161, same operations in real_t.  (The
// reference's double instantiation would round the extremes to float in its
// float2 block reduction, code/buildTree.cu:102-105; they stay double here.)
template <typename R>
__device__ V4<R> box_from_extent(const Ext<R> &e) {
#pragma clang fp contract(off)
  const R cx = (e.xx + e.nx) * R(0.5), cy = (e.xy + e.ny) * R(0.5), cz = (e.xz + e.nz) * R(0.5);
  const R hx = (e.xx - e.nx) * R(0.5), hy = (e.xy - e.ny) * R(0.5), hz = (e.xz - e.nz) * R(0.5);
  const R h = rmax(hz, rmax(hy, hx));
  R hs = R(1);
  if (h > R(0)) {
    while (hs > h) hs *= R(0.5);
    while (hs < h) hs *= R(2);
  }
  const R hq = hs / R(1048576);
  const long long qx = (long long)(cx / hq), qy = (long long)(cy / hq), qz = (long long)(cz / hq);
  return make_v4<R>(hq * (R)qx, hq * (R)qy, hq * (R)qz, hs);
}

template <typename R>
__global__ __launch_bounds__(kBlock) void k_bbox_final(const V4<R> *__restrict__ part, int nb,
                                                       V4<R> *__restrict__ box) {
  __shared__ Ext<R> sh[kBlock / kWave];
  const int r = blockIdx.x;
  Ext<R> x = ext_empty<R>();
  for (int k = threadIdx.x; k < nb; k += kBlock) {
    const V4<R> a = part[2 * (r * nb + k)], b = part[2 * (r * nb + k) + 1];
    ext_merge(x, a.x, a.y, a.z, b.x, b.y, b.z);   // a block may have had no particles
  }
  block_minmax(x, sh);
  if (threadIdx.x == 0) box[r] = box_from_extent(x);
}

// ---- geometric keys -----------------------------------------------------

// One ChildBox step (code/Particle4.h:138-148): the centre moves by hsize/2.
template <typename R>
__device__ __forceinline__ void child_step(R &cx, R &cy, R &cz, R &h, int o) {
#pragma clang fp contract(off)
  const R s = R(0.5) * h;
  cx = (o & 1) ? cx + s : cx - s;
  cy = (o & 2) ? cy + s : cy - s;
  cz = (o & 4) ? cz + s : cz - s;
  h = s;
}

template <typename R>
__global__ __launch_bounds__(kBlock) void k_geo_keys(int n, int nsys, const int *__restrict__ off,
                                                     const V4<R> *__restrict__ box,
                                                     const V4<R> *__restrict__ pos,
                                                     uint64_t *__restrict__ keys,
                                                     int *__restrict__ vals) {
  BT_BUILD_PRIO();
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    const int r = nsys > 1 ? system_of(i, nsys, off) : 0;
    const V4<R> b = box[r];
    const V4<R> p = pos[i];
    R cx = b.x, cy = b.y, cz = b.z, h = b.w;
    uint64_t k = 0;
#pragma unroll
    for (int l = 0; l < kMaxLevel; l++) {
      // Octant (code/Particle4.h:129-136): bit set iff centre <= pos
      const int o = (int)(cx <= p.x) | ((int)(cy <= p.y) << 1) | ((int)(cz <= p.z) << 2);
      k = (k << 3) | (uint64_t)o;
      child_step(cx, cy, cz, h, o);
    }
    keys[i] = k;
    vals[i] = i;
  }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_gather3(int n, const int *__restrict__ perm,
                                                    const T *__restrict__ a,
                                                    const T *__restrict__ b,
                                                    const int *__restrict__ c,
                                                    T *__restrict__ a2,
                                                    T *__restrict__ b2,
                                                    int *__restrict__ c2) {
  BT_BUILD_PRIO();
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    const int j = perm[i];
    a2[i] = a[j];
    b2[i] = b[j];
    c2[i] = c[j];
  }
}

// ---- level-synchronous cell emission ------------------------------------
//
// Parents of level L are the cells of level L-1 (the roots for L = 1).  Eight
// lanes serve one parent; lane o finds the end of child octant o (first index
// in the parent's range whose digit L exceeds o) by binary search on the
// sorted keys.  Parents are split into kLevelBlocks contiguous chunks so the
// block sums give each parent's first child slot in parent order.

struct LevelArgs {
  int L, nsys, nleaf, cap_cells, n;
  const int *off;
  const uint64_t *keys;
  uint4 *cells;
  int *bnd;          // 8 ints per parent: end of each child octant
  int *blk;          // kLevelBlocks block sums -> offsets
  int2 *root_cells;
  DevCounters *dc;
};

__device__ __forceinline__ int level_parent_count(const LevelArgs &a) {
  return a.L == 1 ? a.nsys : a.dc->lvl_count[a.L - 1];
}

__device__ __forceinline__ void parent_range(const LevelArgs &a, int p, int &b, int &e,
                                             bool &splits) {
  if (a.L == 1) {
    b = a.off[p];
    e = a.off[p + 1];
    splits = e > b;  // the root always splits (code/buildTree.cu:702)
  } else {
    const uint4 c = a.cells[a.dc->lvl_start[a.L - 1] + p];
    b = (int)c.z;
    e = (int)c.w;
    splits = !cell_leaf(c);
  }
}

// Per-lane child boundaries of one parent; returns the child count of octant o.
__device__ __forceinline__ int child_count(const LevelArgs &a, int p, bool active, int o,
                                           int *end_out) {
  int b = 0, e = 0;
  bool splits = false;
  if (active) parent_range(a, p, b, e, splits);
  int end = b;
  if (active && splits) {
    const int shift = 63 - 3 * a.L;
    // first index in [b, e) with digit > o
    int lo = b, hi = e;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      const int d = (int)((a.keys[mid] >> shift) & 7ull);
      if (d <= o) lo = mid + 1; else hi = mid;
    }
    end = lo;
  }
  const int beg_lane = __shfl_up(end, 1, 8);
  const int beg = (o == 0) ? b : beg_lane;
  *end_out = end;
  return (active && splits) ? end - beg : 0;
}

__global__ __launch_bounds__(kBlock) void k_level_count(LevelArgs a) {
  BT_BUILD_PRIO();
  __shared__ int shw[4];
  const int np = level_parent_count(a);
  const int chunk = (np + kLevelBlocks - 1) / kLevelBlocks;
  const int p0 = blockIdx.x * chunk, p1 = min(np, p0 + chunk);
  const int o = threadIdx.x & 7;
  int sum = 0;
  for (int pb = p0; pb < p1; pb += kBlock / 8) {
    const int p = pb + (threadIdx.x >> 3);
    const bool active = p < p1;
    int end;
    const int cnt = child_count(a, p, active, o, &end);
    if (active) a.bnd[8 * p + o] = end;
    sum += cnt > 0 ? 1 : 0;
  }
  sum = wave_sum(sum);
  if (lane_id() == 0) shw[threadIdx.x >> 6] = sum;
  __syncthreads();
  if (threadIdx.x == 0) a.blk[blockIdx.x] = shw[0] + shw[1] + shw[2] + shw[3];
}

__global__ __launch_bounds__(kBlock) void k_level_scan(LevelArgs a) {
  BT_BUILD_PRIO();
  constexpr int kItems = (kLevelBlocks + kBlock - 1) / kBlock;
  __shared__ int shw[kBlock / kWave];
  __shared__ int s_run;
  // thread t scans block sums [t * kItems, (t + 1) * kItems)
  int v[kItems];
  int own = 0;
#pragma unroll
  for (int k = 0; k < kItems; k++) {
    const int i = threadIdx.x * kItems + k;
    v[k] = i < kLevelBlocks ? a.blk[i] : 0;
    own += v[k];
  }
  int incl = wave_incl_scan(own);
  if (lane_id() == kWave - 1) shw[threadIdx.x >> 6] = incl;
  __syncthreads();
  int pre = 0;
  for (int w = 0; w < (int)(threadIdx.x >> 6); w++) pre += shw[w];
  if (threadIdx.x == kBlock - 1) s_run = pre + incl;
  int off = pre + incl - own;
  __syncthreads();
  if (threadIdx.x == 0) {
    const int run = s_run;
    DevCounters *dc = a.dc;
    const int start = a.L == 1 ? 0 : dc->lvl_start[a.L - 1] + dc->lvl_count[a.L - 1];
    int total = run;
    if (start + total > a.cap_cells) {
      atomicOr(&dc->err, 1);
      atomicOr(&dc->err_sticky, 1);
      total = 0;
    }
    dc->lvl_start[a.L] = start;
    dc->lvl_count[a.L] = total;
    if (total > 0) dc->nlevels = a.L;
  }
#pragma unroll
  for (int k = 0; k < kItems; k++) {
    const int i = threadIdx.x * kItems + k;
    if (i < kLevelBlocks) a.blk[i] = off;
    off += v[k];
  }
}

__global__ __launch_bounds__(kBlock) void k_level_emit(LevelArgs a) {
  BT_BUILD_PRIO();
  __shared__ int shw[4];
  __shared__ int sh_nodes, sh_leaves;
  DevCounters *dc = a.dc;
  const int total = dc->lvl_count[a.L];
  if (total == 0) return;  // nothing at this level (or capacity exceeded)
  if (threadIdx.x == 0) { sh_nodes = 0; sh_leaves = 0; }
  const int np = level_parent_count(a);
  const int chunk = (np + kLevelBlocks - 1) / kLevelBlocks;
  const int p0 = blockIdx.x * chunk, p1 = min(np, p0 + chunk);
  const int o = threadIdx.x & 7;
  const int gsub = (threadIdx.x & 63) >> 3;  // 8-lane group within the wave
  const int start = dc->lvl_start[a.L];
  const int pstart = a.L == 1 ? 0 : dc->lvl_start[a.L - 1];
  int run = a.blk[blockIdx.x];
  int nodes = 0, leaves = 0, capped = 0;
  __syncthreads();
  for (int pb = p0; pb < p1; pb += kBlock / 8) {
    const int p = pb + (threadIdx.x >> 3);
    const bool active = p < p1;
    int b = 0, e = 0;
    bool splits = false;
    if (active) parent_range(a, p, b, e, splits);
    int end = b, beg = b;
    if (active && splits) {
      end = a.bnd[8 * p + o];
      beg = (o == 0) ? b : a.bnd[8 * p + o - 1];
    }
    const int cnt = (active && splits) ? end - beg : 0;
    const bool is_child = cnt > 0;
    const bool is_node = is_child && cnt > a.nleaf && a.L < kMaxLevel;
    // 8-bit masks of this parent's children / node children
    const uint64_t mc = ballot64(is_child), mn = ballot64(is_node);
    const int sh = gsub * 8;
    const uint32_t gc = (uint32_t)((mc >> sh) & 0xFFull), gn = (uint32_t)((mn >> sh) & 0xFFull);
    const int nchild = __popc(gc), nnode = __popc(gn);
    // exclusive scan of nchild over the parents of this block, in order
    const int per_parent = (o == 0) ? nchild : 0;
    const int incl = wave_incl_scan(per_parent);
    if (lane_id() == 63) shw[threadIdx.x >> 6] = incl;
    __syncthreads();
    int wpre = 0;
    for (int k = 0; k < (threadIdx.x >> 6); k++) wpre += shw[k];
    const int tot = shw[0] + shw[1] + shw[2] + shw[3];
    __syncthreads();
    // value at the group's lane 0 (its exclusive prefix)
    const int excl_lane0 = __shfl(incl - per_parent, (threadIdx.x & 63) & ~7, kWave);
    const int first = start + run + wpre + excl_lane0;
    run += tot;
    if (is_child) {
      const uint32_t below = (1u << o) - 1u;
      const int rank = is_node ? __popc(gn & below) : nnode + __popc((gc & ~gn) & below);
      const int ci = first + rank;
      const uint32_t parent_idx = a.L == 1 ? 0u : (uint32_t)(pstart + p);
      uint4 cell;
      cell.x = parent_idx | ((uint32_t)a.L << 27);
      cell.y = is_node ? 0u : 0xFFFFFFFFu;  // node's first/nchild set one level down
      cell.z = (uint32_t)beg;
      cell.w = (uint32_t)end;
      a.cells[ci] = cell;
      if (is_node) nodes++; else leaves++;
      if (cnt > a.nleaf && !is_node) capped = 1;
    }
    if (active && splits && o == 0) {
      if (a.L == 1) {
        a.root_cells[p] = make_int2(first, nchild);
      } else {
        a.cells[pstart + p].y = nchild > 0 ? cell_pack_y(first, nchild) : 0xFFFFFFFFu;
      }
    }
  }
  nodes = wave_sum(nodes);
  leaves = wave_sum(leaves);
  if (lane_id() == 0) {
    atomicAdd(&sh_nodes, nodes);
    atomicAdd(&sh_leaves, leaves);
  }
  if (capped) atomicOr(&dc->flags, (int)BT_FLAG_DEPTH_CAP);
  __syncthreads();
  if (threadIdx.x == 0) {
    if (sh_nodes) atomicAdd(&dc->nnodes, sh_nodes);
    if (sh_leaves) atomicAdd(&dc->nleaves, sh_leaves);
  }
}

__global__ void k_levels_init(DevCounters *dc) {
  if (threadIdx.x < kLevels + 2) {
    dc->lvl_start[threadIdx.x] = 0;
    dc->lvl_count[threadIdx.x] = 0;
  }
  if (threadIdx.x == 0) {
    dc->nnodes = dc->nleaves = dc->nlevels = dc->ncells = 0;
    dc->flags = 0;
    atomicAnd(&dc->err, ~1);   // other bits belong to phases that may run beside this one
  }
}

__global__ void k_levels_finish(DevCounters *dc, int2 *level_beg) {
  const int L = threadIdx.x;
  if (L < kLevels) {
    const int cnt = (L >= 1 && L <= kMaxLevel) ? dc->lvl_count[L] : 0;
    level_beg[L] = cnt > 0 ? make_int2(dc->lvl_start[L], dc->lvl_start[L] + cnt - 1)
                           : make_int2(0, 0);
  }
  if (L == 0) dc->ncells = dc->nnodes + dc->nleaves;
}

// ---- bottom-up multipoles -------------------------------------------------

// T: the real type of particles and cell data (float or double).
template <typename T>
struct MpArgs {
  int L;
  T inv_theta;
  const V4<T> *pos;       // tree order
  const uint4 *cells;
  V4<T> *scr;             // 6 per cell: (ref,M) (com-ref,-) (Qxx,Qyy,Qzz,Qxy) (Qxz,Qyz,-,-) bmin bmax
  V4<T> *size, *mono, *q0;
  V2<T> *q1, *ls;
  V4<T> *wrec;            // 3 per cell, the walk's P2C record
  const DevCounters *dc;
};

template <typename R>
struct Moments {
  R M, sx, sy, sz, xx, yy, zz, xy, xz, yz;
};

template <typename T>
__device__ void write_cell(const MpArgs<T> &a, int c, T rx, T ry, T rz, double M, double dx,
                           double dy, double dz, double qxx, double qyy, double qzz, double qxy,
                           double qxz, double qyz, const Ext<T> &e) {
#pragma clang fp contract(off)
  const T comx = (T)((double)rx + dx), comy = (T)((double)ry + dy), comz = (T)((double)rz + dz);
  const T bcx = (e.xx + e.nx) * T(0.5), bcy = (e.xy + e.ny) * T(0.5), bcz = (e.xz + e.nz) * T(0.5);
  const T hx = (e.xx - e.nx) * T(0.5), hy = (e.xy - e.ny) * T(0.5), hz = (e.xz - e.nz) * T(0.5);
  const T ex = bcx - comx, ey = bcy - comy, ez = bcz - comz;
  T s;
  if constexpr (sizeof(T) == 4) s = sqrtf(ex * ex + ey * ey + ez * ez);
  else s = sqrt(ex * ex + ey * ey + ez * ez);
  const T l = rmax(T(2) * rmax(hx, rmax(hy, hz)), T(1.0e-6));
  const T op = l * a.inv_theta + s;
  a.size[c] = make_v4<T>(comx, comy, comz, op * op);
  a.mono[c] = make_v4<T>(comx, comy, comz, (T)M);
  a.q0[c] = make_v4<T>((T)qxx, (T)qyy, (T)qzz, (T)qxy);
  a.q1[c] = make_v2<T>((T)qxz, (T)qyz);
  a.ls[c] = make_v2<T>(l, s);
  // walk record: quadrupole pre-multiplied by 3, trace terms precomputed
  const T fqxx = (T)qxx, fqyy = (T)qyy, fqzz = (T)qzz;
  const T tr = fqxx + fqyy + fqzz;
  a.wrec[3 * (size_t)c] = make_v4<T>(comx, comy, comz, (T)M);
  a.wrec[3 * (size_t)c + 1] = make_v4<T>(T(3) * fqxx, T(3) * fqyy, T(3) * fqzz, T(3) * (T)qxy);
  a.wrec[3 * (size_t)c + 2] =
      make_v4<T>(T(3) * (T)qxz, T(3) * (T)qyz, T(0.5) * tr, T(1.5) * tr);
  V4<T> *s6 = a.scr + 6 * (size_t)c;
  s6[0] = make_v4<T>(rx, ry, rz, (T)M);
  s6[1] = make_v4<T>((T)dx, (T)dy, (T)dz, T(0));
  s6[2] = make_v4<T>((T)qxx, (T)qyy, (T)qzz, (T)qxy);
  s6[3] = make_v4<T>((T)qxz, (T)qyz, T(0), T(0));
  s6[4] = make_v4<T>(e.nx, e.ny, e.nz, T(0));
  s6[5] = make_v4<T>(e.xx, e.xy, e.xz, T(0));
}

template <typename T, typename R>
__device__ void finish_moments(const MpArgs<T> &a, int c, T rx, T ry, T rz, const Moments<R> &m,
                               const Ext<T> &e) {
  if (!(m.M > R(0))) {
    write_cell<T>(a, c, rx, ry, rz, 0.0, 0, 0, 0, 0, 0, 0, 0, 0, 0, e);
    return;
  }
  const R im = R(1) / m.M;
  const R dx = m.sx * im, dy = m.sy * im, dz = m.sz * im;
  write_cell<T>(a, c, rx, ry, rz, (double)m.M, (double)dx, (double)dy, (double)dz,
                (double)(m.xx * im - dx * dx), (double)(m.yy * im - dy * dy),
                (double)(m.zz * im - dz * dz), (double)(m.xy * im - dx * dy),
                (double)(m.xz * im - dx * dz), (double)(m.yz * im - dy * dz), e);
}

template <typename T, typename R>
__device__ void node_moments(const MpArgs<T> &a, int c, const uint4 cd) {
  Moments<R> m{};
  Ext<T> e = ext_empty<T>();
  const int f = cell_first(cd), nc = cell_nchild(cd);
  const V4<T> r0 = a.scr[6 * (size_t)f];  // this node's reference = first child's
  for (int k = 0; k < nc; k++) {
    const V4<T> *s6 = a.scr + 6 * (size_t)(f + k);
    const V4<T> g = s6[0], d = s6[1], qa = s6[2], qb = s6[3], bmn = s6[4], bmx = s6[5];
    // child com relative to this node's reference: both references lie in the
    // node's bbox, so their difference is at most l and (near-)exact in T
    const R ox = R(g.x - r0.x) + R(d.x), oy = R(g.y - r0.y) + R(d.y),
            oz = R(g.z - r0.z) + R(d.z);
    const R Mc = R(g.w);
    m.M += Mc;
    m.sx += Mc * ox; m.sy += Mc * oy; m.sz += Mc * oz;
    m.xx += Mc * (R(qa.x) + ox * ox); m.yy += Mc * (R(qa.y) + oy * oy);
    m.zz += Mc * (R(qa.z) + oz * oz); m.xy += Mc * (R(qa.w) + ox * oy);
    m.xz += Mc * (R(qb.x) + ox * oz); m.yz += Mc * (R(qb.y) + oy * oz);
    ext_merge(e, bmn.x, bmn.y, bmn.z, bmx.x, bmx.y, bmx.z);
  }
  finish_moments<T, R>(a, c, r0.x, r0.y, r0.z, m, e);
}

// Treecode<float>: leaves sum in FP32, nodes combine in FP32 below level
// kFp64Levels and in FP64 at and above it.  Treecode<double>: FP64 throughout.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_multipole_level(MpArgs<T> a) {
  BT_BUILD_PRIO();
  const int cnt = a.dc->lvl_count[a.L];
  const int c0 = a.dc->lvl_start[a.L];
  for (int t = blockIdx.x * kBlock + threadIdx.x; t < cnt; t += gridDim.x * kBlock) {
    const int c = c0 + t;
    const uint4 cd = a.cells[c];
    const int pb = (int)cd.z, pe = (int)cd.w;
    if (cell_leaf(cd)) {
      Moments<T> m{};
      Ext<T> e = ext_empty<T>();
      const V4<T> r0 = a.pos[pb];  // reference point: the first particle
      auto add = [&](const V4<T> p) {
        ext_add(e, p.x, p.y, p.z);
        const T dx = p.x - r0.x, dy = p.y - r0.y, dz = p.z - r0.z, w = p.w;
        m.M += w;
        m.sx += w * dx; m.sy += w * dy; m.sz += w * dz;
        m.xx += w * dx * dx; m.yy += w * dy * dy; m.zz += w * dz * dz;
        m.xy += w * dx * dy; m.xz += w * dx * dz; m.yz += w * dy * dz;
      };
      // particles in order, four loads issued together
      int i = pb;
      for (; i + 4 <= pe; i += 4) {
        const V4<T> p0 = a.pos[i], p1 = a.pos[i + 1], p2 = a.pos[i + 2], p3 = a.pos[i + 3];
        add(p0);
        add(p1);
        add(p2);
        add(p3);
      }
      for (; i < pe; i++) add(a.pos[i]);
      finish_moments<T, T>(a, c, r0.x, r0.y, r0.z, m, e);
    } else if (sizeof(T) == 8 || a.L <= kFp64Levels) {
      node_moments<T, double>(a, c, cd);
    } else {
      node_moments<T, float>(a, c, cd);
    }
  }
}

}  // namespace

void launch_bbox(Ctx &c) {
  const int nb = c.bbox_blocks;
  if (c.fp64) {
    hipLaunchKernelGGL(k_bbox_partial<double>, dim3(nb, c.nsys), dim3(kBlock), 0, c.stream,
                       as4<double>(c.pos), c.d_off, nb, as4<double>(c.bbox_part));
    hipLaunchKernelGGL(k_bbox_final<double>, dim3(c.nsys), dim3(kBlock), 0, c.stream,
                       as4<double>(c.bbox_part), nb, as4<double>(c.box));
  } else {
    hipLaunchKernelGGL(k_bbox_partial<float>, dim3(nb, c.nsys), dim3(kBlock), 0, c.stream, c.pos,
                       c.d_off, nb, c.bbox_part);
    hipLaunchKernelGGL(k_bbox_final<float>, dim3(c.nsys), dim3(kBlock), 0, c.stream, c.bbox_part,
                       nb, c.box);
  }
}

void launch_geo_keys(Ctx &c) {
  const dim3 g(grid_for(c.n, kBlock, 8192));
  if (c.fp64)
    hipLaunchKernelGGL(k_geo_keys<double>, g, dim3(kBlock), 0, c.stream, c.n, c.nsys, c.d_off,
                       as4<double>(c.box), as4<double>(c.pos), c.geo_keys, c.valA);
  else
    hipLaunchKernelGGL(k_geo_keys<float>, g, dim3(kBlock), 0, c.stream, c.n, c.nsys, c.d_off,
                       c.box, c.pos, c.geo_keys, c.valA);
}

void launch_gather_tree(Ctx &c) {
  const dim3 g(grid_for(c.n, kBlock, 8192));
  if (c.fp64)
    hipLaunchKernelGGL(k_gather3<double4>, g, dim3(kBlock), 0, c.stream, c.n, c.valA,
                       as4<double>(c.pos), as4<double>(c.vel), c.id, as4<double>(c.pos2),
                       as4<double>(c.vel2), c.id2);
  else
    hipLaunchKernelGGL(k_gather3<float4>, g, dim3(kBlock), 0, c.stream, c.n, c.valA, c.pos,
                       c.vel, c.id, c.pos2, c.vel2, c.id2);
  std::swap(c.pos, c.pos2);
  std::swap(c.vel, c.vel2);
  std::swap(c.id, c.id2);
}

void launch_levels(Ctx &c) {
  hipLaunchKernelGGL(k_levels_init, dim3(1), dim3(64), 0, c.stream, c.dc);
  LevelArgs a;
  a.nsys = c.nsys;
  a.nleaf = c.nleaf;
  a.cap_cells = c.cap_cells;
  a.n = c.n;
  a.off = c.d_off;
  a.keys = c.geo_keys;
  a.cells = c.cells;
  a.bnd = c.child_bnd;
  a.blk = c.lvl_blk;
  a.root_cells = c.root_cells;
  a.dc = c.dc;
  for (int L = 1; L <= kMaxLevel; L++) {
    a.L = L;
    hipLaunchKernelGGL(k_level_count, dim3(kLevelBlocks), dim3(kBlock), 0, c.stream, a);
    hipLaunchKernelGGL(k_level_scan, dim3(1), dim3(kBlock), 0, c.stream, a);
    hipLaunchKernelGGL(k_level_emit, dim3(kLevelBlocks), dim3(kBlock), 0, c.stream, a);
  }
  hipLaunchKernelGGL(k_levels_finish, dim3(1), dim3(64), 0, c.stream, c.dc, c.level_beg);
}

namespace {

template <typename T>
void multipoles_t(Ctx &c) {
  MpArgs<T> a;
  a.inv_theta = (T)(1.0 / (double)c.theta);
  a.pos = as4<T>(c.pos);
  a.cells = c.cells;
  a.scr = as4<T>(c.cscr);
  a.size = as4<T>(c.csize);
  a.mono = as4<T>(c.cmono);
  a.q0 = as4<T>(c.cq0);
  a.q1 = as2<T>(c.cq1);
  a.ls = as2<T>(c.cls);
  a.wrec = as4<T>(c.cwalk);
  a.dc = c.dc;
  const int grid = grid_for(c.cap_cells, kBlock, 2048);
  for (int L = kMaxLevel; L >= 1; L--) {
    a.L = L;
    hipLaunchKernelGGL(k_multipole_level<T>, dim3(grid), dim3(kBlock), 0, c.stream, a);
  }
}

}  // namespace

void launch_multipoles(Ctx &c) {
  if (c.fp64) multipoles_t<double>(c);
  else multipoles_t<float>(c);
}

}  // namespace bt
