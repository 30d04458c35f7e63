// oracle.cpp -- CPU restatement of the Bonsai2 gravity step, used ONLY as the
// checker by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.
// Nothing in the product (bonsai2_amd/, include/) links, loads or calls it.
//
// Each function cites the reference file:line whose behaviour it restates.
// Paths are relative to the reference repository root.
//
// Parity pins (what checks each function against something independent):
//   or_bounding_box      brute-force min/max + closed-form box rules (tests)
//   or_build_tree        treebuild/readme:35-39 known-answer counts on the
//                        reference Plummer sampler; invariants of
//                        SWtest/treebuild.cu:1477; order == stable sort of
//                        or_geo_keys (independent route)
//   or_multipoles        FP64 brute force + parallel-axis theorem (tests)
//   or_ph_keys           exhaustive small-NBITS Hilbert bijection/adjacency,
//                        hand-derived golden keys
//   or_make_groups       closed-form chopping of sorted keys (tests)
//   or_walk              theta->0 equals direct sum; two-body and quadrupole
//                        closed forms; FP64 direct sum error bound; its lists
//                        (or_walk_lists) against an FP64 statement of the
//                        opening rule: accepted cells are far enough by the
//                        clamp distance, hold >= 3 particles, have an opened
//                        parent, and with the opened leaves tile [0, N)
//   or_direct_fp64       long-double brute force on tiny inputs
//   or_plummer_ref       bit-exact against the reference's own plummer.h
//                        compiled into oracle/_ref (when that builds)
//   or_energies          closed forms (two-body)
//   or_f16_round         numpy float16 conversion (round to nearest even)
//   or_pair_features     closed forms on hand-built pairs (tests)
//   or_mlp_forward/grad  finite differences of the forward pass in FP64;
//                        a hand-computed 8-32-32-1 case with one-hot weights
//   or_walk (learned)    a network that outputs a constant log2(theta)
//                        reproduces the fixed-theta walk exactly
//   or_learn_pairs       brute-force pair error on a 2-cell toy; theta_need
//                        is the exact acceptance boundary of the walk's test
// Every function has at least one such check; none is "parity unpinned".
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <type_traits>
#include <vector>
#include <omp.h>

namespace {

struct F4 { float x, y, z, w; };
struct F2 { float x, y; };
struct U4 { uint32_t x, y, z, w; };
struct I2 { int32_t x, y; };
struct D4 { double x, y, z, w; };
struct D2 { double x, y; };

// 4- and 2-vectors of the real type (Treecode<float> / Treecode<double>)
template <typename R> struct V;
template <> struct V<float> { typedef F4 v4; typedef F2 v2; };
template <> struct V<double> { typedef D4 v4; typedef D2 v2; };

constexpr int kMaxLevel = 21;   // depth of the geometric octant key (63 bits)

// ---------------------------------------------------------------------------
// Octant / ChildBox: code/Particle4.h:129-148.  Bit b of the octant is set
// iff centre <= pos on that axis (x -> 1, y -> 2, z -> 4).  The child centre
// moves by +-hsize/2 per axis; all values are exact dyadic floats.
// ---------------------------------------------------------------------------
template <typename R>
inline int octant_of(R cx, R cy, R cz, R px, R py, R pz) {
  return (int)(cx <= px) | ((int)(cy <= py) << 1) | ((int)(cz <= pz) << 2);
}

template <typename R>
struct BoxT { R cx, cy, cz, h; };
typedef BoxT<float> BoxF;

template <typename R>
inline BoxT<R> child_box(const BoxT<R> &b, int oct) {
  const R s = R(0.5) * b.h;
  BoxT<R> c;
  c.cx = b.cx + s * ((oct & 1) ? R(1) : R(-1));
  c.cy = b.cy + s * ((oct & 2) ? R(1) : R(-1));
  c.cz = b.cz + s * ((oct & 4) ? R(1) : R(-1));
  c.h = s;
  return c;
}

// CellData packing: code/Treecode.h:45-89.
inline U4 pack_cell(int level, uint32_t parent, uint32_t pbeg, uint32_t pend,
                    bool leaf, uint32_t first, uint32_t nchild) {
  U4 c;
  c.x = parent | ((uint32_t)level << 27);
  c.y = leaf ? 0xFFFFFFFFu : (first | ((nchild - 1u) << 29));
  c.z = pbeg;
  c.w = pend;
  return c;
}
inline bool cell_is_leaf(const U4 &c) { return c.y == 0xFFFFFFFFu; }
inline int cell_first(const U4 &c) { return (int)(c.y & 0x1FFFFFFFu); }
inline int cell_n(const U4 &c) { return (int)(c.y >> 29) + 1; }

}  // namespace

extern "C" {

int or_version() { return 3; }

// ---------------------------------------------------------------------------
// Root box: code/buildTree.cu:80-166 (computeBoundingBox).  Min/max of the
// particle coordinates (initial +-1e10), centre and half extent in float,
// hsize = smallest power of two >= max half extent (loop from 1.0,
// :147-150), centre quantised toward zero to hsize/2^20 (:152-159).
// A zero extent (the reference's loop would underflow to 0 and divide by
// it) is given hsize = 1.
// out box = {cx, cy, cz, hsize}; bmin/bmax = raw extremes.
// Treecode<double> (R = double) keeps the extremes in double; the
// reference's double instantiation would pass them through its float2 block
// reduction (code/buildTree.cu:102-105), rounding them to float.
// ---------------------------------------------------------------------------
}  // extern "C"
template <typename R>
static void bounding_box_t(int n, const typename V<R>::v4 *pos, R *box, R *bmin, R *bmax) {
  R mn[3] = {R(1e10), R(1e10), R(1e10)}, mx[3] = {R(-1e10), R(-1e10), R(-1e10)};
  for (int i = 0; i < n; i++) {
    mn[0] = std::fmin(mn[0], pos[i].x); mx[0] = std::fmax(mx[0], pos[i].x);
    mn[1] = std::fmin(mn[1], pos[i].y); mx[1] = std::fmax(mx[1], pos[i].y);
    mn[2] = std::fmin(mn[2], pos[i].z); mx[2] = std::fmax(mx[2], pos[i].z);
  }
  R c[3], hv[3];
  for (int k = 0; k < 3; k++) {
    c[k] = (mx[k] + mn[k]) * R(0.5);
    hv[k] = (mx[k] - mn[k]) * R(0.5);
  }
  const R h = std::fmax(hv[2], std::fmax(hv[1], hv[0]));
  R hsize = R(1);
  if (h > R(0)) {
    while (hsize > h) hsize *= R(0.5);
    while (hsize < h) hsize *= R(2);
  }
  const R hquant = hsize / (R)(1 << 20);
  for (int k = 0; k < 3; k++) {
    const long long q = (long long)(c[k] / hquant);
    box[k] = hquant * (R)q;
  }
  box[3] = hsize;
  for (int k = 0; k < 3; k++) { bmin[k] = mn[k]; bmax[k] = mx[k]; }
}
extern "C" {

void or_bounding_box(int n, const F4 *pos, float *box, float *bmin, float *bmax) {
  bounding_box_t<float>(n, pos, box, bmin, bmax);
}
void or_bounding_box_d(int n, const D4 *pos, double *box, double *bmin, double *bmax) {
  bounding_box_t<double>(n, pos, box, bmin, bmax);
}

// ---------------------------------------------------------------------------
// Geometric octant key: the sequence of Octant() digits a particle takes on
// the way down the ChildBox recursion of code/buildTree.cu:233-247, 21
// levels, most significant digit first (bits 62..60 = level 1).
// ---------------------------------------------------------------------------
}  // extern "C"
template <typename R>
static void geo_keys_t(int n, const typename V<R>::v4 *pos, const R *box, uint64_t *keys) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; i++) {
    BoxT<R> b{box[0], box[1], box[2], box[3]};
    uint64_t k = 0;
    for (int l = 0; l < kMaxLevel; l++) {
      const int o = octant_of<R>(b.cx, b.cy, b.cz, pos[i].x, pos[i].y, pos[i].z);
      k = (k << 3) | (uint64_t)o;
      b = child_box(b, o);
    }
    keys[i] = k;
  }
}
extern "C" {

void or_geo_keys(int n, const F4 *pos, const float *box, uint64_t *keys) {
  geo_keys_t<float>(n, pos, box, keys);
}
void or_geo_keys_d(int n, const D4 *pos, const double *box, uint64_t *keys) {
  geo_keys_t<double>(n, pos, box, keys);
}

// ---------------------------------------------------------------------------
// Octree build: code/buildTree.cu:170-584 (buildOctant), :624-722
// (buildOctree), level sort and level table :725-804, :932-951.
//
// Restated as a recursive STABLE partition of particle indices by Octant()
// against each cell's ChildBox centre, in octant order 0..7 (:444-471).
// The reference's within-leaf order is nondeterministic (block races on
// atomicAdd, :262); the canonical order used here keeps partitioning below
// the leaves down to level 21, i.e. leaf particles are ordered by their
// deeper octant digits and then by input index.
//
// Cells: a child with count > nleaf is a node, 0 < count <= nleaf a leaf,
// empty children make no cell (:394-395); the root (level 0) always splits
// and is not stored, so level-1 cells start at index 0 (:424).  Cells are
// numbered level by level (stable sort by level, :940-944); within a level,
// children of earlier parents come first and each parent's children are
// [nodes in octant order][leaves in octant order] (:393-400, :446, :549).
// A cell at level 21 that still holds > nleaf particles becomes an
// (oversized) leaf and the depth-cap flag is raised; the reference has no
// depth guard at all (5-bit level field, code/Treecode.h:50-51).
//
// Forest: nroot roots given by offsets[0..nroot] over the input, each with
// its own box (4 floats).  Root r's particles stay in [off[r], off[r+1]).
//
// Outputs: order[n] (tree position -> input index), cells[ncell_max] U4,
// level_range[2*32] = (first, last inclusive) per level (code/buildTree.cu
// :748-780), root_cells[2*nroot] = (first level-1 cell, count),
// stats[6] = ncells, nnodes, nleaves, nlevels, depth_cap_hit, nbodies_leaf.
// Returns 0, or -1 if ncell_max is too small.
// ---------------------------------------------------------------------------
// ocb: octant scratch of n bytes; a call uses only [b, e), disjoint from
// every concurrent call's range and released before its children run.
}  // extern "C"
template <typename R>
static void stable_partition_rec(std::vector<int> &idx, std::vector<int> &tmp,
                                 std::vector<unsigned char> &ocb, const typename V<R>::v4 *pos,
                                 int b, int e, BoxT<R> box, int level) {
  if (e - b <= 1 || level >= kMaxLevel) return;
  int cnt[8] = {0};
  unsigned char *oc = ocb.data() + b;
  for (int i = b; i < e; i++) {
    const typename V<R>::v4 &p = pos[idx[i]];
    const int o = octant_of<R>(box.cx, box.cy, box.cz, p.x, p.y, p.z);
    oc[i - b] = (unsigned char)o;
    cnt[o]++;
  }
  int off[9];
  off[0] = b;
  for (int o = 0; o < 8; o++) off[o + 1] = off[o] + cnt[o];
  int cur[8];
  for (int o = 0; o < 8; o++) cur[o] = off[o];
  for (int i = b; i < e; i++) tmp[cur[oc[i - b]]++] = idx[i];
  for (int i = b; i < e; i++) idx[i] = tmp[i];
  const bool par = (e - b) > 65536 && level < 3;
  for (int o = 0; o < 8; o++) {
    if (cnt[o] == 0) continue;
    const BoxT<R> cb = child_box(box, o);
    const int cb0 = off[o], ce0 = off[o + 1];
    if (par) {
#pragma omp task shared(idx, tmp, ocb)
      stable_partition_rec<R>(idx, tmp, ocb, pos, cb0, ce0, cb, level + 1);
    } else {
      stable_partition_rec<R>(idx, tmp, ocb, pos, cb0, ce0, cb, level + 1);
    }
  }
  if (par) {
#pragma omp taskwait
  }
}

template <typename R>
static int build_tree_t(int n, const typename V<R>::v4 *pos, int nroot, const int *offsets,
                        const R *boxes, int nleaf, int ncell_max,
                        int *order, U4 *cells, int *level_range, int *root_cells,
                        int *stats) {
  std::vector<int> idx(n), tmp(n);
  std::vector<unsigned char> ocb(n);
  for (int i = 0; i < n; i++) idx[i] = i;
#pragma omp parallel
#pragma omp single
  for (int r = 0; r < nroot; r++) {
    const BoxT<R> b{boxes[4 * r], boxes[4 * r + 1], boxes[4 * r + 2], boxes[4 * r + 3]};
    stable_partition_rec<R>(idx, tmp, ocb, pos, offsets[r], offsets[r + 1], b, 0);
  }
  for (int i = 0; i < n; i++) order[i] = idx[i];

  struct Par { int b, e; BoxT<R> box; int cell; int level; };
  std::vector<Par> parents, next;
  for (int r = 0; r < nroot; r++) {
    const BoxT<R> b{boxes[4 * r], boxes[4 * r + 1], boxes[4 * r + 2], boxes[4 * r + 3]};
    parents.push_back({offsets[r], offsets[r + 1], b, -1 - r, 0});
  }
  for (int l = 0; l < 64; l++) { level_range[l] = 0; }
  int ncells = 0, nnodes = 0, nleaves = 0, nlevels = 0, capped = 0;
  long long nb_leaf = 0;
  for (int level = 1; !parents.empty(); level++) {
    next.clear();
    const int lvl_first = ncells;
    for (const Par &p : parents) {
      int cnt[8] = {0};
      int prev = -1;
      for (int i = p.b; i < p.e; i++) {
        const typename V<R>::v4 &q = pos[order[i]];
        const int o = octant_of<R>(p.box.cx, p.box.cy, p.box.cz, q.x, q.y, q.z);
        if (o < prev) return -2;  // partition order broken (cannot happen)
        prev = o;
        cnt[o]++;
      }
      int off[9];
      off[0] = p.b;
      for (int o = 0; o < 8; o++) off[o + 1] = off[o] + cnt[o];
      int nchild = 0;
      for (int o = 0; o < 8; o++) nchild += cnt[o] > 0;
      const int first = ncells;
      if (ncells + nchild > ncell_max) return -1;
      if (p.cell >= 0) {
        const U4 pc = cells[p.cell];
        cells[p.cell] = pack_cell(level - 1, pc.x & 0x07FFFFFFu, pc.z, pc.w, false,
                                  (uint32_t)first, (uint32_t)nchild);
      } else {
        root_cells[2 * (-1 - p.cell)] = first;
        root_cells[2 * (-1 - p.cell) + 1] = nchild;
      }
      const uint32_t parent_idx = p.cell >= 0 ? (uint32_t)p.cell : 0u;
      for (int pass = 0; pass < 2; pass++) {
        for (int o = 0; o < 8; o++) {
          const int c = cnt[o];
          if (c == 0) continue;
          const bool split = c > nleaf && level < kMaxLevel;
          if ((pass == 0) != split) continue;
          if (c > nleaf && level >= kMaxLevel) capped = 1;
          const int ci = ncells++;
          // node records get first/nchild when their own children are made
          cells[ci] = pack_cell(level, parent_idx, (uint32_t)off[o], (uint32_t)off[o + 1],
                                !split, 0u, 1u);
          if (split) {
            nnodes++;
            next.push_back({off[o], off[o + 1], child_box(p.box, o), ci, level});
          } else {
            nleaves++;
            nb_leaf += c;
          }
        }
      }
    }
    if (ncells > lvl_first) {
      level_range[2 * level] = lvl_first;
      level_range[2 * level + 1] = ncells - 1;
      nlevels = level;
    }
    parents.swap(next);
  }
  stats[0] = ncells; stats[1] = nnodes; stats[2] = nleaves; stats[3] = nlevels;
  stats[4] = capped; stats[5] = (int)nb_leaf;
  return 0;
}
extern "C" {

int or_build_tree(int n, const F4 *pos, int nroot, const int *offsets, const float *boxes,
                  int nleaf, int ncell_max, int *order, U4 *cells, int *level_range,
                  int *root_cells, int *stats) {
  return build_tree_t<float>(n, pos, nroot, offsets, boxes, nleaf, ncell_max, order, cells,
                             level_range, root_cells, stats);
}
int or_build_tree_d(int n, const D4 *pos, int nroot, const int *offsets, const double *boxes,
                    int nleaf, int ncell_max, int *order, U4 *cells, int *level_range,
                    int *root_cells, int *stats) {
  return build_tree_t<double>(n, pos, nroot, offsets, boxes, nleaf, ncell_max, order, cells,
                              level_range, root_cells, stats);
}

// ---------------------------------------------------------------------------
// Cell multipoles: code/computeMultipoles.cu:81-190 (computeCellMultipoles),
// addMonopole/addQuadrupole :12-68, addBoxSize code/cuda_primitives.h:56-83.
// Per cell over its particles [pbeg, pend) in tree order: particle bbox,
// mass M, centre of mass, second moments about the centre of mass divided
// by M (Q_ab = sum m x_a x_b / M - com_a com_b), l = max(2*max half
// extent, 1e-6), s = |bbox centre - com|, cellOp = l*inv_theta + s,
// outputs (:180-188): size = (com, cellOp^2), monopole = (com, M),
// quad0 = (xx, yy, zz, xy), quad1 = (xz, yz).
//
// mode 0: the moments are summed in FP64 with the two-pass form
//         sum m (x-com)(x-com)^T / M (the same quantity, well conditioned).
// mode 1: the reference's own arithmetic: FP32 products m*x, m*x*x, a
//         32-lane xor-butterfly FP32 sum per chunk of 32 particles, chunk
//         sums accumulated in FP64, then Q = S2/M - com*com in FP64
//         (:113-114, :127-158).  Kept to document that this form loses
//         most digits of Q for small cells far from the origin.
// The bbox/size part is FP32 and identical in both modes.
// inv_theta is passed as the reference passes it: float(1.0/double(theta)).
// ls_out (may be null) receives (l, s) per cell.
// ---------------------------------------------------------------------------
static float butterfly32(float v[32]) {
  float a[32];
  for (int l = 0; l < 32; l++) a[l] = v[l];
  for (int i = 4; i >= 0; i--) {
    float b[32];
    for (int l = 0; l < 32; l++) b[l] = a[l] + a[l ^ (1 << i)];
    for (int l = 0; l < 32; l++) a[l] = b[l];
  }
  return a[0];
}

void or_multipoles(int ncells, const U4 *cells, const F4 *pos, float inv_theta, int mode,
                   F4 *size, F4 *mono, F4 *quad0, F2 *quad1, F2 *ls_out) {
#pragma omp parallel for schedule(dynamic, 64)
  for (int c = 0; c < ncells; c++) {
    const int b = (int)cells[c].z, e = (int)cells[c].w;
    float mn[3] = {1e10f, 1e10f, 1e10f}, mx[3] = {-1e10f, -1e10f, -1e10f};
    for (int i = b; i < e; i++) {
      mn[0] = std::fmin(mn[0], pos[i].x); mx[0] = std::fmax(mx[0], pos[i].x);
      mn[1] = std::fmin(mn[1], pos[i].y); mx[1] = std::fmax(mx[1], pos[i].y);
      mn[2] = std::fmin(mn[2], pos[i].z); mx[2] = std::fmax(mx[2], pos[i].z);
    }
    double M = 0, cx = 0, cy = 0, cz = 0;
    double qxx = 0, qyy = 0, qzz = 0, qxy = 0, qxz = 0, qyz = 0;
    if (mode == 0) {
      for (int i = b; i < e; i++) {
        const double m = pos[i].w;
        M += m; cx += m * pos[i].x; cy += m * pos[i].y; cz += m * pos[i].z;
      }
      const double im = 1.0 / M;
      cx *= im; cy *= im; cz *= im;
      for (int i = b; i < e; i++) {
        const double m = pos[i].w;
        const double dx = pos[i].x - cx, dy = pos[i].y - cy, dz = pos[i].z - cz;
        qxx += m * dx * dx; qyy += m * dy * dy; qzz += m * dz * dz;
        qxy += m * dx * dy; qxz += m * dx * dz; qyz += m * dy * dz;
      }
      qxx *= im; qyy *= im; qzz *= im; qxy *= im; qxz *= im; qyz *= im;
    } else {
      for (int i0 = b; i0 < e; i0 += 32) {
        float v[10][32];
        for (int l = 0; l < 32; l++) {
          const F4 p = pos[std::min(i0 + l, e - 1)];
          const float m = (i0 + l < e) ? p.w : 0.0f;
          v[0][l] = m * p.x; v[1][l] = m * p.y; v[2][l] = m * p.z; v[3][l] = m;
          v[4][l] = m * p.x * p.x; v[5][l] = m * p.y * p.y; v[6][l] = m * p.z * p.z;
          v[7][l] = m * p.x * p.y; v[8][l] = m * p.x * p.z; v[9][l] = m * p.y * p.z;
        }
        cx += butterfly32(v[0]); cy += butterfly32(v[1]); cz += butterfly32(v[2]);
        M += butterfly32(v[3]);
        qxx += butterfly32(v[4]); qyy += butterfly32(v[5]); qzz += butterfly32(v[6]);
        qxy += butterfly32(v[7]); qxz += butterfly32(v[8]); qyz += butterfly32(v[9]);
      }
      const double im = 1.0 / M;
      cx *= im; cy *= im; cz *= im;
      qxx = qxx * im - cx * cx; qyy = qyy * im - cy * cy; qzz = qzz * im - cz * cz;
      qxy = qxy * im - cx * cy; qxz = qxz * im - cx * cz; qyz = qyz * im - cy * cz;
    }
    const float comx = (float)cx, comy = (float)cy, comz = (float)cz;
    const float bcx = (mx[0] + mn[0]) * 0.5f, bcy = (mx[1] + mn[1]) * 0.5f,
                bcz = (mx[2] + mn[2]) * 0.5f;
    const float hx = (mx[0] - mn[0]) * 0.5f, hy = (mx[1] - mn[1]) * 0.5f,
                hz = (mx[2] - mn[2]) * 0.5f;
    const float dx = bcx - comx, dy = bcy - comy, dz = bcz - comz;
    const float s = std::sqrt(dx * dx + dy * dy + dz * dz);
    const float l = std::max(2.0f * std::max(hx, std::max(hy, hz)), 1.0e-6f);
    const float op = l * inv_theta + s;
    size[c] = F4{comx, comy, comz, op * op};
    mono[c] = F4{comx, comy, comz, (float)M};
    quad0[c] = F4{(float)qxx, (float)qyy, (float)qzz, (float)qxy};
    quad1[c] = F2{(float)qxz, (float)qyz};
    if (ls_out) ls_out[c] = F2{l, s};
  }
}

// Treecode<double>: the same cell quantities from FP64 particles, all in
// FP64 (two-pass moments as mode 0), outputs double.
void or_multipoles_d(int ncells, const U4 *cells, const D4 *pos, double inv_theta, D4 *size,
                     D4 *mono, D4 *quad0, D2 *quad1, D2 *ls_out) {
#pragma omp parallel for schedule(dynamic, 64)
  for (int c = 0; c < ncells; c++) {
    const int b = (int)cells[c].z, e = (int)cells[c].w;
    double mn[3] = {1e10, 1e10, 1e10}, mx[3] = {-1e10, -1e10, -1e10};
    double M = 0, cx = 0, cy = 0, cz = 0;
    for (int i = b; i < e; i++) {
      mn[0] = std::fmin(mn[0], pos[i].x); mx[0] = std::fmax(mx[0], pos[i].x);
      mn[1] = std::fmin(mn[1], pos[i].y); mx[1] = std::fmax(mx[1], pos[i].y);
      mn[2] = std::fmin(mn[2], pos[i].z); mx[2] = std::fmax(mx[2], pos[i].z);
      const double m = pos[i].w;
      M += m; cx += m * pos[i].x; cy += m * pos[i].y; cz += m * pos[i].z;
    }
    const double im = 1.0 / M;
    cx *= im; cy *= im; cz *= im;
    double qxx = 0, qyy = 0, qzz = 0, qxy = 0, qxz = 0, qyz = 0;
    for (int i = b; i < e; i++) {
      const double m = pos[i].w;
      const double dx = pos[i].x - cx, dy = pos[i].y - cy, dz = pos[i].z - cz;
      qxx += m * dx * dx; qyy += m * dy * dy; qzz += m * dz * dz;
      qxy += m * dx * dy; qxz += m * dx * dz; qyz += m * dy * dz;
    }
    qxx *= im; qyy *= im; qzz *= im; qxy *= im; qxz *= im; qyz *= im;
    const double ex = (mx[0] + mn[0]) * 0.5 - cx, ey = (mx[1] + mn[1]) * 0.5 - cy,
                 ez = (mx[2] + mn[2]) * 0.5 - cz;
    const double s = std::sqrt(ex * ex + ey * ey + ez * ez);
    const double l = std::max(2.0 * std::max((mx[0] - mn[0]) * 0.5,
                                             std::max((mx[1] - mn[1]) * 0.5, (mx[2] - mn[2]) * 0.5)),
                              1.0e-6);
    const double op = l * inv_theta + s;
    size[c] = D4{cx, cy, cz, op * op};
    mono[c] = D4{cx, cy, cz, M};
    quad0[c] = D4{qxx, qyy, qzz, qxy};
    quad1[c] = D2{qxz, qyz};
    if (ls_out) ls_out[c] = D2{l, s};
  }
}

// ---------------------------------------------------------------------------
// Peano-Hilbert key: code/makeGroups.cu:19-106 (get_key<21>) and the
// quantisation of computeKeys :108-135.  Coordinates are quantised in
// float to 21-bit integers in the root box; then, MSB first, each 3-bit
// digit (x<<2 | y<<1 | z) selects the Hilbert state change (coordinate
// swaps and bitwise complements) and the output digit from the table
// {0,1,7,6,3,2,4,5}.  box = {cx, cy, cz, hsize}.
// ---------------------------------------------------------------------------
static uint64_t hilbert_key(int x, int y, int z, int nbits) {
  static const int kDigit[8] = {0, 1, 7, 6, 3, 2, 4, 5};
  uint64_t key = 0;
  for (int lev = nbits - 1; lev >= 0; lev--) {
    const int bit = 1 << lev;
    const int bx = (x & bit) ? 1 : 0, by = (y & bit) ? 1 : 0, bz = (z & bit) ? 1 : 0;
    const int idx = (bx << 2) | (by << 1) | bz;
    int t;
    switch (idx) {
      case 0: t = z; z = y; y = t; break;                       // swap y,z
      case 1: case 5: t = x; x = y; y = t; break;               // swap x,y
      case 4: case 6: x = ~x; z = ~z; break;                    // flip x,z
      case 3: case 7: t = ~x; x = ~y; y = t; break;             // x<-~y, y<-~x
      default: t = ~z; z = ~y; y = t; break;                    // idx 2: z<-~y, y<-~z
    }
    key = (key << 3) + (uint64_t)kDigit[idx];
  }
  return key;
}

uint64_t or_hilbert_key(int x, int y, int z, int nbits) { return hilbert_key(x, y, z, nbits); }

void or_ph_keys(int n, const F4 *pos, const float *box, uint64_t *keys) {
  const float inv = 0.5f / box[3];
  const float bx = box[0] - box[3], by = box[1] - box[3], bz = box[2] - box[3];
  const float scale = (float)(1 << 21);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; i++) {
    const int xc = (int)((pos[i].x - bx) * inv * scale);
    const int yc = (int)((pos[i].y - by) * inv * scale);
    const int zc = (int)((pos[i].z - bz) * inv * scale);
    keys[i] = hilbert_key(xc, yc, zc, 21);
  }
}

void or_ph_keys_d(int n, const D4 *pos, const double *box, uint64_t *keys) {
  const double inv = 0.5 / box[3];
  const double bx = box[0] - box[3], by = box[1] - box[3], bz = box[2] - box[3];
  const double scale = (double)(1 << 21);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; i++) {
    const int xc = (int)((pos[i].x - bx) * inv * scale);
    const int yc = (int)((pos[i].y - by) * inv * scale);
    const int zc = (int)((pos[i].z - bz) * inv * scale);
    keys[i] = hilbert_key(xc, yc, zc, 21);
  }
}

// Stable sort of (key, index): the contract of thrust::sort_by_key at
// code/makeGroups.cu:270-277 (nondecreasing keys, ties keep input order).
void or_sort_keys(int n, const uint64_t *keys, int *perm) {
  for (int i = 0; i < n; i++) perm[i] = i;
  std::stable_sort(perm, perm + n, [&](int a, int b) { return keys[a] < keys[b]; });
}

// ---------------------------------------------------------------------------
// Groups: code/makeGroups.cu:137-223 (mask_keys, make_groups) and
// :259-309.  The sorted keys are masked to their top level_split octal
// digits (mask built at :259-266); each run [s,e) of equal masked keys is
// chopped into groups [s + k*ncrit, min(s + (k+1)*ncrit, e)).  The
// reference's group slots come from atomicAdd (:219); here groups are
// listed in order of pbeg.  seg_id (may be null) splits runs further
// (different systems of a forest never share a group).
// Returns the group count; groups[] = (pbeg, np).
// ---------------------------------------------------------------------------
int or_make_groups(int n, const uint64_t *sorted_keys, const int *seg_id,
                   int level_split, int ncrit, I2 *groups) {
  level_split = std::max(1, level_split);
  uint64_t mask = 0;
  for (int i = 0; i < 21; i++) {
    mask <<= 3;
    if (i < level_split) mask |= 7;
  }
  int ng = 0, s = 0;
  while (s < n) {
    int e = s + 1;
    while (e < n && (sorted_keys[e] & mask) == (sorted_keys[s] & mask) &&
           (!seg_id || seg_id[e] == seg_id[s]))
      e++;
    for (int b = s; b < e; b += ncrit) groups[ng++] = I2{b, std::min(ncrit, e - b)};
    s = e;
  }
  return ng;
}

// ---------------------------------------------------------------------------
// Tree walk: code/computeForces.cu:30-54 (opening test), :58-78 (P2P),
// :83-128 (P2C, monopole + quadrupole), :229-463 (treewalk_warp),
// :475-619 (treewalk), :622-727 (host).
//
// For each group: the i-particles are pos_ph[pbeg, pbeg+np) (PH order);
// the group box is their min/max (:539-555).  Breadth-first from the
// group's level-1 cells (:259-263, :511-512): a cell is opened iff
// ds^2 < cellOp^2 (improved Barnes-Hut distance from the group box) or it
// holds fewer than 3 particles (:284-285).  Opened nodes push their
// children; unopened cells (nodes or leaves) go to the approximate (P2C)
// list; opened leaves put every particle (tree order) on the direct (P2P)
// list.  Lists are evaluated in list order, approximate list first.
// acc_out[k] = (ax, ay, az, phi) at PH index k; phi includes the self
// term -m_i/eps (P2P with dr = 0).
// Stats (:567-609): per valid i-particle add its group's list lengths;
// stats_out = {direct_sum, direct_max, approx_sum, approx_max} (integers),
// pot_out = 0.5 * sum m_i phi_i.
// group_cells[2*g] = (first level-1 cell, count) for group g's tree
// (null: use level_range[2], [3]).
// acc_type 0 = FP32 accumulation (as the reference), 1 = FP64.
// group_inv_theta (may be null): per-group 1/theta_g; the cell's opening
// radius is then recomputed as (l * inv_theta_g + s)^2 from cell_ls[c] =
// (l, s), the same float operations as code/computeMultipoles.cu:174-177
// (used by the learned-opening-angle path; null = the stored cellOp^2).
// counts_out (may be null): per-group (n_approx, n_direct).
// ---------------------------------------------------------------------------
}  // extern "C"

// ---------------------------------------------------------------------------
// Learned opening angle (no reference counterpart).  A per-(group, cell)
// network y = MLP(x) picks the opening angle of each candidate cell:
// theta = exp2(y + shift), clamped to [theta_lo, theta_hi]; the cell is then
// opened iff ds^2 < (l / theta + s)^2, the improved Barnes-Hut test of
// code/computeForces.cu:30-54 with a per-pair angle (or if it holds < 3
// particles, :284-285).
//
// Features (8, all in [0, 1]; IEEE +,-,*,/,sqrt here, the hardware reciprocal
// and square root (about 1 ulp) on the GPU): with o = com - group centre,
// d = |o| (>= 1e-30):
//   x0 = rg / (rg + d)         rg = |group half-size vector|   group radius
//   x1 = np / 64                                              group count
//   x2 = l / (l + d)           l = cell bbox size              cell size
//   x3 = d / (d + R)           R = root box half size          distance
//   x4 = min(M / M_sys, 1)                                     mass fraction
//   x5 = |Q|_F / (|Q|_F + M l^2 + 1e-30)                       quadrupole ratio
//   x6 = |o_z| / d, x7 = |o_k| / d (k = longest group axis)    direction cosines
// Network: 8 -> 32 (SiLU) -> 32 (SiLU) -> 1, weights and biases stored in
// FP16, FP32 accumulation; the input and the first hidden layer are rounded
// to FP16 (they feed FP16 matrix products), the second hidden layer is used
// in FP32 by the output dot product.  Parameter order (1377 floats):
//   W1[32][8], b1[32], W2[32][32], b2[32], w3[32], b3.
// ---------------------------------------------------------------------------
constexpr int kMlpIn = 8, kMlpH = 32;
constexpr int kOffW1 = 0, kOffB1 = 256, kOffW2 = 288, kOffB2 = 1312, kOffW3 = 1344,
              kOffB3 = 1376, kMlpParams = 1377;

inline uint16_t f32_to_f16_bits(float f) {
  uint32_t x;
  std::memcpy(&x, &f, 4);
  const uint32_t sign = (x >> 16) & 0x8000u;
  uint32_t mant = x & 0x7FFFFFu;
  const int exp = (int)((x >> 23) & 0xFF);
  if (exp == 0xFF) return (uint16_t)(sign | 0x7C00u | (mant ? 0x200u : 0u));
  const int e = exp - 127 + 15;
  if (e >= 31) return (uint16_t)(sign | 0x7C00u);
  if (e <= 0) {
    if (e < -10) return (uint16_t)sign;
    mant |= 0x800000u;
    const int shift = 14 - e;
    uint32_t hm = mant >> shift;
    const uint32_t rem = mant & ((1u << shift) - 1u), half = 1u << (shift - 1);
    if (rem > half || (rem == half && (hm & 1u))) hm++;
    return (uint16_t)(sign | hm);
  }
  uint32_t h = sign | ((uint32_t)e << 10) | (mant >> 13);
  const uint32_t rem = mant & 0x1FFFu;
  if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h++;
  return (uint16_t)h;
}

inline float f16_bits_to_f32(uint16_t h) {
  const uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
  const int e = (h >> 10) & 0x1F;
  uint32_t m = h & 0x3FFu, x;
  if (e == 0) {
    if (m == 0) {
      x = sign;
    } else {  // subnormal: normalise
      int ee = -1;
      do { m <<= 1; ee++; } while (!(m & 0x400u));
      x = sign | ((uint32_t)(127 - 15 - ee) << 23) | ((m & 0x3FFu) << 13);
    }
  } else if (e == 31) {
    x = sign | 0x7F800000u | (m << 13);
  } else {
    x = sign | ((uint32_t)(e - 15 + 127) << 23) | (m << 13);
  }
  float f;
  std::memcpy(&f, &x, 4);
  return f;
}

inline float f16r(float x) { return f16_bits_to_f32(f32_to_f16_bits(x)); }

struct PairIn {
  float gcx, gcy, gcz, gsx, gsy, gsz;
  int np;
  float cx, cy, cz, M, l;
  float qxx, qyy, qzz, qxy, qxz, qyz;
  float R, Msys;
};

inline void pair_features(const PairIn &p, float x[8]) {
  const float ox = p.cx - p.gcx, oy = p.cy - p.gcy, oz = p.cz - p.gcz;
  const float d = std::fmax(std::sqrt(ox * ox + oy * oy + oz * oz), 1e-30f);
  const float rg = std::sqrt(p.gsx * p.gsx + p.gsy * p.gsy + p.gsz * p.gsz);
  x[0] = rg / (rg + d);
  x[1] = (float)p.np * 0.015625f;
  x[2] = p.l / (p.l + d);
  x[3] = d / (d + p.R);
  x[4] = std::fmin(p.M / p.Msys, 1.0f);
  const float qn = std::sqrt(p.qxx * p.qxx + p.qyy * p.qyy + p.qzz * p.qzz +
                             2.0f * (p.qxy * p.qxy + p.qxz * p.qxz + p.qyz * p.qyz));
  x[5] = qn / (qn + p.M * p.l * p.l + 1e-30f);
  x[6] = std::fabs(oz) / d;
  const float ok = (p.gsx >= p.gsy && p.gsx >= p.gsz) ? ox : (p.gsy >= p.gsz ? oy : oz);
  x[7] = std::fabs(ok) / d;
}

inline double silu_d(double z) { return z / (1.0 + std::exp(-z)); }

// Forward pass on FP16-rounded parameters pr[] (see above).  Optional caches
// for the backward pass: xh[8], z1[32], a1h[32], z2[32], a2[32].
inline double mlp_forward_one(const float *pr, const float *x, float *xh_c, double *z1_c,
                              float *a1h_c, double *z2_c, float *a2_c, bool round16) {
  float xh[kMlpIn], a1h[kMlpH], a2[kMlpH];
  for (int k = 0; k < kMlpIn; k++) xh[k] = round16 ? f16r(x[k]) : x[k];
  for (int h = 0; h < kMlpH; h++) {
    double z = pr[kOffB1 + h];
    for (int k = 0; k < kMlpIn; k++) z += (double)pr[kOffW1 + h * kMlpIn + k] * xh[k];
    const float a = (float)silu_d((double)(float)z);
    a1h[h] = round16 ? f16r(a) : a;
    if (z1_c) z1_c[h] = (double)(float)z;
  }
  for (int h = 0; h < kMlpH; h++) {
    double z = pr[kOffB2 + h];
    for (int k = 0; k < kMlpH; k++) z += (double)pr[kOffW2 + h * kMlpH + k] * a1h[k];
    a2[h] = (float)silu_d((double)(float)z);
    if (z2_c) z2_c[h] = (double)(float)z;
  }
  double y = pr[kOffB3];
  for (int h = 0; h < kMlpH; h++) y += (double)pr[kOffW3 + h] * a2[h];
  if (xh_c) for (int k = 0; k < kMlpIn; k++) xh_c[k] = xh[k];
  if (a1h_c) for (int h = 0; h < kMlpH; h++) a1h_c[h] = a1h[h];
  if (a2_c) for (int h = 0; h < kMlpH; h++) a2_c[h] = a2[h];
  return y;
}

struct LearnCfg {
  const float *pr;     // FP16-rounded parameters
  float shift;         // log2(theta) offset (error budget)
  float theta_lo, theta_hi;
  const F2 *group_scale;  // per group (R, M_sys)
  long long *visits;   // optional: {cells tested, cells whose decision needed the network}
};

// 1 / theta for one pair, clamped to [1/theta_hi, 1/theta_lo]
inline float learned_inv_theta(const LearnCfg &lc, const float x[8]) {
  const double y = mlp_forward_one(lc.pr, x, nullptr, nullptr, nullptr, nullptr, nullptr, true);
  const float it = std::exp2(-((float)y + lc.shift));
  return std::fmin(std::fmax(it, 1.0f / lc.theta_hi), 1.0f / lc.theta_lo);
}

template <typename A, typename R = float>
static void p2p(A acc[4], R px, R py, R pz, const typename V<R>::v4 &j, R eps2) {
  const R dx = j.x - px, dy = j.y - py, dz = j.z - pz;
  const R r2 = dx * dx + dy * dy + dz * dz + eps2;
  const R rinv = R(1) / std::sqrt(r2);
  const R rinv2 = rinv * rinv;
  const R mrinv = j.w * rinv;
  const R mrinv3 = mrinv * rinv2;
  acc[3] -= mrinv;
  acc[0] += mrinv3 * dx;
  acc[1] += mrinv3 * dy;
  acc[2] += mrinv3 * dz;
}

template <typename A, typename R = float>
static void p2c(A acc[4], R px, R py, R pz, const typename V<R>::v4 &m,
                const typename V<R>::v4 &q0, const typename V<R>::v2 &q1, R eps2) {
  const R dx = px - m.x, dy = py - m.y, dz = pz - m.z;
  const R r2 = dx * dx + dy * dy + dz * dz + eps2;
  const R rinv = R(1) / std::sqrt(r2);
  const R rinv2 = rinv * rinv;
  const R mrinv = m.w * rinv;
  const R mrinv3 = rinv2 * mrinv;
  const R mrinv5 = rinv2 * mrinv3;
  const R mrinv7 = rinv2 * mrinv5;
  const R D0 = mrinv, D1 = -mrinv3, D2 = mrinv5 * R(3), D3 = mrinv7 * R(-15);
  const R q11 = q0.x, q22 = q0.y, q33 = q0.z, q12 = q0.w, q13 = q1.x, q23 = q1.y;
  const R q = q11 + q22 + q33;
  const R qRx = q11 * dx + q12 * dy + q13 * dz;
  const R qRy = q12 * dx + q22 * dy + q23 * dz;
  const R qRz = q13 * dx + q23 * dy + q33 * dz;
  const R qRR = qRx * dx + qRy * dy + qRz * dz;
  acc[3] -= D0 + R(0.5) * (D1 * q + D2 * qRR);
  const R C = D1 + R(0.5) * (D2 * q + D3 * qRR);
  acc[0] += C * dx + D2 * qRx;
  acc[1] += C * dy + D2 * qRy;
  acc[2] += C * dz + D2 * qRz;
}

template <typename R = float>
static inline bool open_test(const typename V<R>::v4 &cs, R gcx, R gcy, R gcz, R gsx, R gsy,
                             R gsz, const F2 *ls, float inv_theta) {
  R dx = std::fabs(gcx - cs.x) - gsx;
  R dy = std::fabs(gcy - cs.y) - gsy;
  R dz = std::fabs(gcz - cs.z) - gsz;
  dx += std::fabs(dx); dx *= R(0.5);
  dy += std::fabs(dy); dy *= R(0.5);
  dz += std::fabs(dz); dz *= R(0.5);
  const R ds2 = dx * dx + dy * dy + dz * dz;
  if (!ls) return ds2 < std::fabs(cs.w);
  const R op = ls->x * inv_theta + ls->y;
  return ds2 < op * op;
}

// R: the real type of positions and cell data (float: Treecode<float>,
// double: Treecode<double>, fixed angle only); A: the accumulator type.
template <typename A, typename R = float>
static void walk_impl(int ngroups, const I2 *groups, const int *group_cells,
                      const int *level_range, const typename V<R>::v4 *pos_ph,
                      const typename V<R>::v4 *pos_tree, const U4 *cells,
                      const typename V<R>::v4 *csize, const typename V<R>::v4 *mono,
                      const typename V<R>::v4 *quad0, const typename V<R>::v2 *quad1, R eps2,
                      const F2 *cell_ls, const float *group_inv_theta,
                      typename V<R>::v4 *acc_out,
                      long long *stats_out, double *pot_out, int *counts_out,
                      const LearnCfg *lc = nullptr, bool monopole_only = false) {
  long long dsum = 0, asum = 0, dmax = 0, amax = 0;
  double pot = 0;
#pragma omp parallel reduction(+ : dsum, asum, pot) reduction(max : dmax, amax)
  {
    std::vector<int> cur, nxt, alist, dlist;
#pragma omp for schedule(dynamic, 16)
    for (int g = 0; g < ngroups; g++) {
      const int pbeg = groups[g].x, np = groups[g].y;
      R mn[3] = {pos_ph[pbeg].x, pos_ph[pbeg].y, pos_ph[pbeg].z};
      R mx[3] = {mn[0], mn[1], mn[2]};
      for (int i = pbeg; i < pbeg + np; i++) {
        mn[0] = std::min(mn[0], pos_ph[i].x); mx[0] = std::max(mx[0], pos_ph[i].x);
        mn[1] = std::min(mn[1], pos_ph[i].y); mx[1] = std::max(mx[1], pos_ph[i].y);
        mn[2] = std::min(mn[2], pos_ph[i].z); mx[2] = std::max(mx[2], pos_ph[i].z);
      }
      const R gcx = R(0.5) * (mx[0] + mn[0]), gcy = R(0.5) * (mx[1] + mn[1]),
              gcz = R(0.5) * (mx[2] + mn[2]);
      const R gsx = R(0.5) * (mx[0] - mn[0]), gsy = R(0.5) * (mx[1] - mn[1]),
              gsz = R(0.5) * (mx[2] - mn[2]);
      const float it = group_inv_theta ? group_inv_theta[g] : 0.0f;
      const F2 *lsp = group_inv_theta ? cell_ls : nullptr;
      int c0, nc;
      if (group_cells) { c0 = group_cells[2 * g]; nc = group_cells[2 * g + 1]; }
      else { c0 = level_range[2]; nc = level_range[3] - level_range[2] + 1; }
      cur.clear(); alist.clear(); dlist.clear();
      for (int c = c0; c < c0 + nc; c++) cur.push_back(c);
      while (!cur.empty()) {
        nxt.clear();
        for (int c : cur) {
          const U4 cd = cells[c];
          bool split;
          if constexpr (!std::is_same<R, float>::value) {
            split = open_test<R>(csize[c], gcx, gcy, gcz, gsx, gsy, gsz, nullptr, 0.f) ||
                    ((int)cd.w - (int)cd.z < 3);
          } else if (lc) {
            const F4 &m = mono[c];
            const PairIn pin{gcx, gcy, gcz, gsx, gsy, gsz, np, m.x, m.y, m.z, m.w,
                             cell_ls[c].x, quad0[c].x, quad0[c].y, quad0[c].z, quad0[c].w,
                             quad1[c].x, quad1[c].y, lc->group_scale[g].x,
                             lc->group_scale[g].y};
            float xf[8];
            pair_features(pin, xf);
            split = open_test(csize[c], gcx, gcy, gcz, gsx, gsy, gsz, cell_ls + c,
                              learned_inv_theta(*lc, xf)) ||
                    ((int)cd.w - (int)cd.z < 3);
            if (lc->visits) {
              const bool open_hi = open_test(csize[c], gcx, gcy, gcz, gsx, gsy, gsz, cell_ls + c,
                                             1.0f / lc->theta_hi);
              const bool open_lo = open_test(csize[c], gcx, gcy, gcz, gsx, gsy, gsz, cell_ls + c,
                                             1.0f / lc->theta_lo);
#pragma omp atomic
              lc->visits[0]++;
              if (open_lo && !open_hi) {
#pragma omp atomic
                lc->visits[1]++;
              }
            }
          } else {
            split = open_test(csize[c], gcx, gcy, gcz, gsx, gsy, gsz,
                              lsp ? lsp + c : nullptr, it) ||
                    ((int)cd.w - (int)cd.z < 3);
          }
          const bool leaf = cell_is_leaf(cd);
          if (!split) {
            alist.push_back(c);
          } else if (!leaf) {
            const int fc = cell_first(cd), n = cell_n(cd);
            for (int k = 0; k < n; k++) nxt.push_back(fc + k);
          } else {
            for (int p = (int)cd.z; p < (int)cd.w; p++) dlist.push_back(p);
          }
        }
        cur.swap(nxt);
      }
      const long long na = (long long)alist.size(), nd = (long long)dlist.size();
      if (counts_out) { counts_out[2 * g] = (int)na; counts_out[2 * g + 1] = (int)nd; }
      for (int i = pbeg; i < pbeg + np; i++) {
        A acc[4] = {0, 0, 0, 0};
        const R px = pos_ph[i].x, py = pos_ph[i].y, pz = pos_ph[i].z;
        if (monopole_only)  // code/computeForces.cu:198-225: the cell as a point mass
          for (int c : alist) p2p<A, R>(acc, px, py, pz, mono[c], eps2);
        else
          for (int c : alist) p2c<A, R>(acc, px, py, pz, mono[c], quad0[c], quad1[c], eps2);
        for (int p : dlist) p2p<A, R>(acc, px, py, pz, pos_tree[p], eps2);
        acc_out[i] = typename V<R>::v4{(R)acc[0], (R)acc[1], (R)acc[2], (R)acc[3]};
        pot += 0.5 * (double)pos_ph[i].w * (double)(R)acc[3];
      }
      dsum += nd * np; asum += na * np;
      dmax = std::max(dmax, nd); amax = std::max(amax, na);
    }
  }
  stats_out[0] = dsum; stats_out[1] = dmax; stats_out[2] = asum; stats_out[3] = amax;
  *pot_out = pot;
}

extern "C" {

void or_walk(int ngroups, const I2 *groups, const int *group_cells, const int *level_range,
             const F4 *pos_ph, const F4 *pos_tree, const U4 *cells, const F4 *csize,
             const F4 *mono, const F4 *quad0, const F2 *quad1, float eps2, int acc_type,
             const F2 *cell_ls, const float *group_inv_theta, F4 *acc_out,
             long long *stats_out, double *pot_out, int *counts_out, int monopole_only) {
  if (acc_type == 0)
    walk_impl<float>(ngroups, groups, group_cells, level_range, pos_ph, pos_tree, cells, csize,
                     mono, quad0, quad1, eps2, cell_ls, group_inv_theta, acc_out, stats_out,
                     pot_out, counts_out, nullptr, monopole_only != 0);
  else
    walk_impl<double>(ngroups, groups, group_cells, level_range, pos_ph, pos_tree, cells,
                      csize, mono, quad0, quad1, eps2, cell_ls, group_inv_theta, acc_out,
                      stats_out, pot_out, counts_out, nullptr, monopole_only != 0);
}

// Treecode<double>: the same walk with FP64 positions, cell data and sums
// (fixed opening angle).
void or_walk_d(int ngroups, const I2 *groups, const int *level_range, const D4 *pos_ph,
               const D4 *pos_tree, const U4 *cells, const D4 *csize, const D4 *mono,
               const D4 *quad0, const D2 *quad1, double eps2, D4 *acc_out, long long *stats_out,
               double *pot_out, int *counts_out, int monopole_only) {
  walk_impl<double, double>(ngroups, groups, nullptr, level_range, pos_ph, pos_tree, cells,
                            csize, mono, quad0, quad1, eps2, nullptr, nullptr, acc_out,
                            stats_out, pot_out, counts_out, nullptr, monopole_only != 0);
}

// ---------------------------------------------------------------------------
// FP64 direct sum (no reference counterpart; the error oracle).  For each
// target t in idx[0..ns): a = sum_j m_j (x_j - x_t) / (r^2+eps2)^{3/2},
// phi = -sum_j m_j / sqrt(r^2+eps2), over ALL j including j == t (so phi
// carries the same self term as the walk).  Sources pos[0..n).
// ---------------------------------------------------------------------------
void or_direct_fp64(int n, const F4 *pos, int ns, const int *idx, double eps2, D4 *out) {
#pragma omp parallel for schedule(dynamic, 4)
  for (int s = 0; s < ns; s++) {
    const F4 pt = pos[idx[s]];
    double ax = 0, ay = 0, az = 0, ph = 0;
    for (int j = 0; j < n; j++) {
      const double dx = (double)pos[j].x - pt.x, dy = (double)pos[j].y - pt.y,
                   dz = (double)pos[j].z - pt.z;
      const double r2 = dx * dx + dy * dy + dz * dz + eps2;
      const double ri = 1.0 / std::sqrt(r2);
      const double mr = pos[j].w * ri;
      const double mr3 = mr * ri * ri;
      ax += mr3 * dx; ay += mr3 * dy; az += mr3 * dz; ph -= mr;
    }
    out[s] = D4{ax, ay, az, ph};
  }
}

// The same for FP64 positions (Treecode<double>).
void or_direct_fp64_d(int n, const D4 *pos, int ns, const int *idx, double eps2, D4 *out) {
#pragma omp parallel for schedule(dynamic, 4)
  for (int s = 0; s < ns; s++) {
    const D4 t = pos[idx[s]];
    double ax = 0, ay = 0, az = 0, ph = 0;
    for (int j = 0; j < n; j++) {
      const double dx = pos[j].x - t.x, dy = pos[j].y - t.y, dz = pos[j].z - t.z;
      const double r2 = dx * dx + dy * dy + dz * dz + eps2;
      const double ri = 1.0 / std::sqrt(r2);
      const double mr3 = pos[j].w * ri * ri * ri;
      ax += mr3 * dx; ay += mr3 * dy; az += mr3 * dz; ph -= pos[j].w * ri;
    }
    out[s] = D4{ax, ay, az, ph};
  }
}

// Long-double brute force for pinning or_direct_fp64 on tiny inputs.
void or_direct_ld(int n, const F4 *pos, int ns, const int *idx, double eps2, D4 *out) {
  for (int s = 0; s < ns; s++) {
    const F4 pt = pos[idx[s]];
    long double ax = 0, ay = 0, az = 0, ph = 0;
    for (int j = 0; j < n; j++) {
      const long double dx = (long double)pos[j].x - pt.x, dy = (long double)pos[j].y - pt.y,
                        dz = (long double)pos[j].z - pt.z;
      const long double r2 = dx * dx + dy * dy + dz * dz + (long double)eps2;
      const long double ri = 1.0L / std::sqrt(r2);
      ax += pos[j].w * ri * ri * ri * dx; ay += pos[j].w * ri * ri * ri * dy;
      az += pos[j].w * ri * ri * ri * dz; ph -= pos[j].w * ri;
    }
    out[s] = D4{(double)ax, (double)ay, (double)az, (double)ph};
  }
}

// ---------------------------------------------------------------------------
// Energies (the reference's computeEnergies is a stub, code/computeEnergies
// .cu:3-7).  FP64: K = 1/2 sum m v^2, W = 1/2 sum m (phi + m/eps) (the walk's
// phi includes the self term -m/eps, code/computeForces.cu:64-72).
// ---------------------------------------------------------------------------
void or_energies(int n, const F4 *pos, const F4 *vel, const F4 *acc, double eps,
                 double *ekin, double *epot) {
  double k = 0, w = 0;
#pragma omp parallel for reduction(+ : k, w)
  for (int i = 0; i < n; i++) {
    const double m = pos[i].w;
    k += 0.5 * m * ((double)vel[i].x * vel[i].x + (double)vel[i].y * vel[i].y +
                    (double)vel[i].z * vel[i].z);
    w += 0.5 * m * ((double)acc[i].w + (eps > 0 ? m / eps : 0.0));
  }
  *ekin = k;
  *epot = w;
}

// ---------------------------------------------------------------------------
// Reference Plummer sampler: code/plummer.h:37-117 with glibc srand/rand
// (my_rand = rand()/(1+RAND_MAX), :135-137).  Radius from the cumulative
// mass, isotropic direction, Aarseth-Henon-Wielen velocity rejection,
// Henon scaling 3pi/16, m = 1/n, centre-of-mass position and velocity
// removed in FP64.  rmax is the truncation radius in structural units
// (the reference uses 100).  Output in double: pos/vel [3n], mass [n].
// ---------------------------------------------------------------------------
void or_plummer_ref(long n, unsigned seed, double rmax, double *mass, double *pos,
                    double *vel) {
  srand(seed);
  auto urand = []() { return rand() / (1. + RAND_MAX); };
  long i = 0;
  while (i < n) {
    const double X1 = urand(), X2 = urand(), X3 = urand();
    const double R = 1.0 / std::sqrt(std::pow(X1, -2.0 / 3.0) - 1.0);
    if (!(R < rmax)) continue;
    const double Z = (1.0 - 2.0 * X2) * R;
    const double X = std::sqrt(R * R - Z * Z) * std::cos(2.0 * M_PI * X3);
    const double Y = std::sqrt(R * R - Z * Z) * std::sin(2.0 * M_PI * X3);
    const double Ve = std::sqrt(2.0) * std::pow(1.0 + R * R, -0.25);
    double X4 = 0.0, X5 = 0.0;
    while (0.1 * X5 >= X4 * X4 * std::pow(1.0 - X4 * X4, 3.5)) { X4 = urand(); X5 = urand(); }
    const double V = Ve * X4;
    const double X6 = urand(), X7 = urand();
    const double Vz = (1.0 - 2.0 * X6) * V;
    const double Vx = std::sqrt(V * V - Vz * Vz) * std::cos(2.0 * M_PI * X7);
    const double Vy = std::sqrt(V * V - Vz * Vz) * std::sin(2.0 * M_PI * X7);
    const double conv = 3.0 * M_PI / 16.0;
    mass[i] = 1.0 / n;
    pos[3 * i] = X * conv; pos[3 * i + 1] = Y * conv; pos[3 * i + 2] = Z * conv;
    vel[3 * i] = Vx / std::sqrt(conv); vel[3 * i + 1] = Vy / std::sqrt(conv);
    vel[3 * i + 2] = Vz / std::sqrt(conv);
    i++;
  }
  double mc = 0, xc[3] = {0, 0, 0}, vc[3] = {0, 0, 0};
  for (long k = 0; k < n; k++) {
    mc += mass[k];
    for (int d = 0; d < 3; d++) { xc[d] += mass[k] * pos[3 * k + d]; vc[d] += mass[k] * vel[3 * k + d]; }
  }
  for (int d = 0; d < 3; d++) { xc[d] /= mc; vc[d] /= mc; }
  for (long k = 0; k < n; k++)
    for (int d = 0; d < 3; d++) { pos[3 * k + d] -= xc[d]; vel[3 * k + d] -= vc[d]; }
}


// ---------------------------------------------------------------------------
// Learned opening angle: primitives exported for the tests (see the block
// above walk_impl for the definitions).
// ---------------------------------------------------------------------------
void or_f16_round(int n, const float *in, float *out) {
  for (int i = 0; i < n; i++) out[i] = f16r(in[i]);
}

// pin[20*k ..]: gc.xyz, gs.xyz, np, com.xyz, M, l, Qxx, Qyy, Qzz, Qxy, Qxz, Qyz, R, M_sys
void or_pair_features(int n, const float *pin, float *feat) {
  for (int k = 0; k < n; k++) {
    const float *q = pin + 20 * k;
    const PairIn p{q[0], q[1], q[2], q[3], q[4], q[5], (int)q[6], q[7], q[8], q[9], q[10],
                   q[11], q[12], q[13], q[14], q[15], q[16], q[17], q[18], q[19]};
    pair_features(p, feat + 8 * k);
  }
}

// y[k] = MLP(x[8k..]) with FP16-rounded parameters (round16 = 1) or raw
// FP32 parameters and no activation rounding (round16 = 0).
void or_mlp_forward(int n, const float *params, const float *x, int round16, float *y) {
  std::vector<float> pr(params, params + kMlpParams);
  if (round16) for (auto &v : pr) v = f16r(v);
  for (int k = 0; k < n; k++)
    y[k] = (float)mlp_forward_one(pr.data(), x + 8 * k, nullptr, nullptr, nullptr, nullptr,
                                  nullptr, round16 != 0);
}

// Mean squared error L = 1/n sum (y - t)^2 and dL/dparams (FP64), the
// rounding to FP16 treated as the identity in the backward pass.
void or_mlp_grad(int n, const float *params, const float *x, const float *t, int round16,
                 double *grad, double *loss) {
  std::vector<float> pr(params, params + kMlpParams);
  if (round16) for (auto &v : pr) v = f16r(v);
  for (int i = 0; i < kMlpParams; i++) grad[i] = 0.0;
  double L = 0.0;
  for (int k = 0; k < n; k++) {
    float xh[kMlpIn], a1h[kMlpH], a2[kMlpH];
    double z1[kMlpH], z2[kMlpH];
    const double y = mlp_forward_one(pr.data(), x + 8 * k, xh, z1, a1h, z2, a2, round16 != 0);
    const double r = y - (double)t[k];
    L += r * r;
    const double g = 2.0 * r / n;
    grad[kOffB3] += g;
    double dz2[kMlpH];
    for (int h = 0; h < kMlpH; h++) {
      grad[kOffW3 + h] += g * a2[h];
      const double sg = 1.0 / (1.0 + std::exp(-z2[h]));
      dz2[h] = g * pr[kOffW3 + h] * sg * (1.0 + z2[h] * (1.0 - sg));
      grad[kOffB2 + h] += dz2[h];
      for (int j = 0; j < kMlpH; j++) grad[kOffW2 + h * kMlpH + j] += dz2[h] * a1h[j];
    }
    for (int j = 0; j < kMlpH; j++) {
      double da = 0.0;
      for (int h = 0; h < kMlpH; h++) da += dz2[h] * pr[kOffW2 + h * kMlpH + j];
      const double sg = 1.0 / (1.0 + std::exp(-z1[j]));
      const double dz1 = da * sg * (1.0 + z1[j] * (1.0 - sg));
      grad[kOffB1 + j] += dz1;
      for (int f = 0; f < kMlpIn; f++) grad[kOffW1 + j * kMlpIn + f] += dz1 * xh[f];
    }
  }
  *loss = L / n;
}

// The interaction lists of one group (code/computeForces.cu:229-463, the
// same traversal and opening rule as walk_impl): the cells it evaluates by
// P2C (alist, breadth-first) and the leaves whose particles it evaluates by
// P2P (leaves), with their counts in *na / *nl.  The lists are checked
// against an FP64 statement of the opening rule in the tests.  mutate plants
// a known mistake, so that the tests can show their check catches it:
//   0 none; 1 no "< 3 particles" rule (:284-285); 2 the group is its centre
//   point, not its box (no box half-sizes in ds, :30-54); 3 no s term in the
//   opening radius (op = l / theta, needs cell_ls and inv_theta).
void or_walk_lists(int g, const I2 *groups, const int *level_range, const F4 *pos_ph,
                   const U4 *cells, const F4 *csize, const F2 *cell_ls, float inv_theta,
                   int mutate, int *alist, int *na, int *leaves, int *nl) {
  const int pbeg = groups[g].x, np = groups[g].y;
  float mn[3] = {pos_ph[pbeg].x, pos_ph[pbeg].y, pos_ph[pbeg].z};
  float mx[3] = {mn[0], mn[1], mn[2]};
  for (int i = pbeg; i < pbeg + np; i++) {
    mn[0] = std::min(mn[0], pos_ph[i].x); mx[0] = std::max(mx[0], pos_ph[i].x);
    mn[1] = std::min(mn[1], pos_ph[i].y); mx[1] = std::max(mx[1], pos_ph[i].y);
    mn[2] = std::min(mn[2], pos_ph[i].z); mx[2] = std::max(mx[2], pos_ph[i].z);
  }
  const float gcx = 0.5f * (mx[0] + mn[0]), gcy = 0.5f * (mx[1] + mn[1]),
              gcz = 0.5f * (mx[2] + mn[2]);
  float gsx = 0.5f * (mx[0] - mn[0]), gsy = 0.5f * (mx[1] - mn[1]), gsz = 0.5f * (mx[2] - mn[2]);
  if (mutate == 2) gsx = gsy = gsz = 0.f;
  std::vector<int> cur, nxt;
  for (int c = level_range[2]; c <= level_range[3]; c++) cur.push_back(c);
  int ka = 0, kl = 0;
  while (!cur.empty()) {
    nxt.clear();
    for (int c : cur) {
      const U4 cd = cells[c];
      F4 cs = csize[c];
      if (mutate == 3) {
        const float op = cell_ls[c].x * inv_theta;
        cs.w = op * op;
      }
      bool split = open_test(cs, gcx, gcy, gcz, gsx, gsy, gsz, nullptr, 0.f);
      if (mutate != 1) split = split || ((int)cd.w - (int)cd.z < 3);
      if (!split) {
        alist[ka++] = c;
      } else if (!cell_is_leaf(cd)) {
        for (int k = 0; k < cell_n(cd); k++) nxt.push_back(cell_first(cd) + k);
      } else {
        leaves[kl++] = c;
      }
    }
    cur.swap(nxt);
  }
  *na = ka;
  *nl = kl;
}

// The walk with the learned per-pair opening angle (FP32 accumulation).
// params: raw FP32 (rounded to FP16 here); group_scale[g] = (R, M_sys).
void or_walk_learned(int ngroups, const I2 *groups, const int *group_cells,
                     const int *level_range, const F4 *pos_ph, const F4 *pos_tree,
                     const U4 *cells, const F4 *csize, const F4 *mono, const F4 *quad0,
                     const F2 *quad1, const F2 *cell_ls, float eps2, const float *params,
                     float shift, float theta_lo, float theta_hi, const F2 *group_scale,
                     F4 *acc_out, long long *stats_out, double *pot_out, int *counts_out,
                     long long *visits) {
  std::vector<float> pr(params, params + kMlpParams);
  for (auto &v : pr) v = f16r(v);
  if (visits) visits[0] = visits[1] = 0;
  const LearnCfg lc{pr.data(), shift, theta_lo, theta_hi, group_scale, visits};
  walk_impl<float>(ngroups, groups, group_cells, level_range, pos_ph, pos_tree, cells, csize,
                   mono, quad0, quad1, eps2, cell_ls, nullptr, acc_out, stats_out, pot_out,
                   counts_out, &lc);
}

// Training pairs for the groups gsel[0..nsel): every cell the walk visits at
// theta_lo (a superset of what any angle in [theta_lo, theta_hi] visits)
// whose acceptance boundary theta_need = l / (ds - s) lies in
// [theta_lo, theta_hi] and that holds >= 3 particles.  For each: the
// features, theta_need, the pair error
//   e = max_i |a_P2C(i) - a_direct(i)| / amag[i]     (FP64, softened)
// over the group's particles i (amag = |a_i| of the FP64 direct sum, PH
// order), and the target log2(theta_need) - log2(max(e, 1e-12)) / 3.
// Groups in gsel order, cells breadth-first.  Returns the pair count (only
// the first max_pairs are written).
int or_learn_pairs(int nsel, const int *gsel, const I2 *groups, const int *group_cells,
                   const int *level_range, const F4 *pos_ph, const F4 *pos_tree,
                   const U4 *cells, const F4 *csize, const F4 *mono, const F4 *quad0,
                   const F2 *quad1, const F2 *cell_ls, const F2 *group_scale, float theta_lo,
                   float theta_hi, float eps2, const double *amag, int max_pairs, int *pair_gc,
                   float *feat, float *theta_need, float *target, float *err) {
  int np_out = 0;
  const float it_lo = 1.0f / theta_lo;
  std::vector<int> cur, nxt;
  for (int si = 0; si < nsel; si++) {
    const int g = gsel[si];
    const int pbeg = groups[g].x, np = groups[g].y;
    float mn[3] = {pos_ph[pbeg].x, pos_ph[pbeg].y, pos_ph[pbeg].z};
    float mx[3] = {mn[0], mn[1], mn[2]};
    for (int i = pbeg; i < pbeg + np; i++) {
      mn[0] = std::min(mn[0], pos_ph[i].x); mx[0] = std::max(mx[0], pos_ph[i].x);
      mn[1] = std::min(mn[1], pos_ph[i].y); mx[1] = std::max(mx[1], pos_ph[i].y);
      mn[2] = std::min(mn[2], pos_ph[i].z); mx[2] = std::max(mx[2], pos_ph[i].z);
    }
    const float gcx = 0.5f * (mx[0] + mn[0]), gcy = 0.5f * (mx[1] + mn[1]),
                gcz = 0.5f * (mx[2] + mn[2]);
    const float gsx = 0.5f * (mx[0] - mn[0]), gsy = 0.5f * (mx[1] - mn[1]),
                gsz = 0.5f * (mx[2] - mn[2]);
    int c0, nc;
    if (group_cells) { c0 = group_cells[2 * g]; nc = group_cells[2 * g + 1]; }
    else { c0 = level_range[2]; nc = level_range[3] - level_range[2] + 1; }
    cur.clear();
    for (int c = c0; c < c0 + nc; c++) cur.push_back(c);
    while (!cur.empty()) {
      nxt.clear();
      for (int c : cur) {
        const U4 cd = cells[c];
        const int nc_p = (int)cd.w - (int)cd.z;
        const F4 &cs = csize[c];
        float dx = std::fabs(gcx - cs.x) - gsx, dy = std::fabs(gcy - cs.y) - gsy,
              dz = std::fabs(gcz - cs.z) - gsz;
        dx += std::fabs(dx); dx *= 0.5f;
        dy += std::fabs(dy); dy *= 0.5f;
        dz += std::fabs(dz); dz *= 0.5f;
        const float ds2 = dx * dx + dy * dy + dz * dz;
        const float l = cell_ls[c].x, sc = cell_ls[c].y;
        const double ds = std::sqrt((double)ds2);
        if (nc_p >= 3 && ds > (double)sc) {
          const double tn = (double)l / (ds - (double)sc);
          if (tn >= (double)theta_lo && tn <= (double)theta_hi) {
            if (np_out < max_pairs) {
              const F4 &m = mono[c];
              const PairIn pin{gcx, gcy, gcz, gsx, gsy, gsz, np, m.x, m.y, m.z, m.w, l,
                               quad0[c].x, quad0[c].y, quad0[c].z, quad0[c].w, quad1[c].x,
                               quad1[c].y, group_scale[g].x, group_scale[g].y};
              pair_features(pin, feat + 8 * np_out);
              double e = 0.0;
              for (int i = pbeg; i < pbeg + np; i++) {
                const double px = pos_ph[i].x, py = pos_ph[i].y, pz = pos_ph[i].z;
                double ax = 0, ay = 0, az = 0;
                for (int j = (int)cd.z; j < (int)cd.w; j++) {
                  const double rx = pos_tree[j].x - px, ry = pos_tree[j].y - py,
                               rz = pos_tree[j].z - pz;
                  const double r2 = rx * rx + ry * ry + rz * rz + (double)eps2;
                  const double ri = 1.0 / std::sqrt(r2);
                  const double w = pos_tree[j].w * ri * ri * ri;
                  ax += w * rx; ay += w * ry; az += w * rz;
                }
                // quadrupole P2C in FP64 (same terms as the walk's p2c)
                const double qx_ = px - m.x, qy_ = py - m.y, qz_ = pz - m.z;
                const double r2 = qx_ * qx_ + qy_ * qy_ + qz_ * qz_ + (double)eps2;
                const double ri = 1.0 / std::sqrt(r2), ri2 = ri * ri;
                const double m1 = m.w * ri, m3 = m1 * ri2, m5 = m3 * ri2, m7 = m5 * ri2;
                const double Qxx = quad0[c].x, Qyy = quad0[c].y, Qzz = quad0[c].z,
                             Qxy = quad0[c].w, Qxz = quad1[c].x, Qyz = quad1[c].y;
                const double tq = Qxx + Qyy + Qzz;
                const double Qx = Qxx * qx_ + Qxy * qy_ + Qxz * qz_;
                const double Qy = Qxy * qx_ + Qyy * qy_ + Qyz * qz_;
                const double Qz = Qxz * qx_ + Qyz * qy_ + Qzz * qz_;
                const double qrr = Qx * qx_ + Qy * qy_ + Qz * qz_;
                const double C = -m3 + 1.5 * tq * m5 - 7.5 * qrr * m7;
                const double bx = C * qx_ + 3.0 * m5 * Qx, by = C * qy_ + 3.0 * m5 * Qy,
                             bz = C * qz_ + 3.0 * m5 * Qz;
                const double ex = bx - ax, ey = by - ay, ez = bz - az;
                const double ei = std::sqrt(ex * ex + ey * ey + ez * ez) / amag[i];
                e = std::max(e, ei);
              }
              pair_gc[2 * np_out] = g;
              pair_gc[2 * np_out + 1] = c;
              theta_need[np_out] = (float)tn;
              err[np_out] = (float)e;
              target[np_out] = (float)(std::log2(tn) - std::log2(std::max(e, 1e-12)) / 3.0);
            }
            np_out++;
          }
        }
        const bool split = ds2 < (l * it_lo + sc) * (l * it_lo + sc) || nc_p < 3;
        if (split && !cell_is_leaf(cd)) {
          const int fc = cell_first(cd), n = cell_n(cd);
          for (int k = 0; k < n; k++) nxt.push_back(fc + k);
        }
      }
      cur.swap(nxt);
    }
  }
  return np_out;
}

}  // extern "C"
