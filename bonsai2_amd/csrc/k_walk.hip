// k_walk.hip -- computeForces on CDNA4 (reference code/computeForces.cu).
//
// One wave64 per group of <= 64 i-particles (one per lane), pulled from a
// device-wide atomic queue by a persistent grid (the reference's
// retired_groupCount loop, :514-522).  Per group:
//   * the group box is the min/max of its particles (:539-555);
//   * the walk starts at the group's level-1 cells (:259-263) and keeps a
//     per-wave LIFO of cells in global memory (the learned-angle walk keeps
//     its top 1024 entries in LDS); each step pops up to 64 cells,
//     one per lane, and applies the reference's rules: open iff
//     ds^2 < cellOp^2 (:30-54) or the cell holds < 3 particles (:284-285);
//     opened nodes push their children, unopened cells join the P2C list,
//     opened leaves put their particles on the P2P list (:293-430);
//   * lists are kept in LDS, compacted with ballots and prefix sums; every 64
//     entries the wave evaluates the batch.  Two ways to feed the j-data:
//     LDS (default) -- the wave stages the batch (relative to the group
//     centre) in LDS and sweeps it with broadcast ds_read_b128; scalar --
//     the batch index is moved to an SGPR with readlane and the 48-byte cell
//     record (or particle) is fetched by s_load through the scalar cache.
//     Both run at the same rate on MI355X: the FP32 VALU issue rate bounds
//     the walk, not LDS or the scalar cache;
//   * each batch is summed in FP32 and folded into Kahan-compensated
//     accumulators.
// The interaction set equals the reference's (the order of traversal does
// not change which cells are opened); FP32 sums run in a different order.
// P2P is code/computeForces.cu:58-78 (20 flops), P2C :83-128 (64 flops),
// rearranged algebraically on a per-cell record with the quadrupole
// pre-scaled by 3 and the trace terms precomputed (35 VALU instructions).
#include "bt_device.h"
#include "bt_internal.h"
#include "bt_interact.h"
#include "bt_learn.h"

namespace bt {

namespace {

constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr int kMaxWalkSteps = 1 << 20;

// Unroll of the full-batch j-loops and the occupancy target of the walk.
#ifndef BT_UNROLL_LDS
#define BT_UNROLL_LDS 8
#endif
#ifndef BT_UNROLL_S
#define BT_UNROLL_S 8
#endif
#ifndef BT_WALK_WAVES_PER_EU
#define BT_WALK_WAVES_PER_EU 1
#endif
#ifndef BT_LEARN_WAVES_PER_EU
#define BT_LEARN_WAVES_PER_EU 4
#endif
#ifndef BT_XCD_QUEUE
#define BT_XCD_QUEUE 1
#endif
constexpr int kXcd = 8;   // MI355X: 8 XCDs, workgroups dispatched round-robin over them
// Entries of a learned walk's stack held in LDS per wave (kLdsStack = 1024,
// 4 KiB).  The learned walk runs at 4 waves per SIMD, so 4 blocks share a CU's
// 160 KiB of LDS; the fixed-angle walk runs at up to 8 and keeps its whole
// stack in global memory.  A deeper stack spills its oldest entries to the
// wave's global stack.  The part used (WalkArgs::lstk_cap) is at least the
// 8 * 64 children one step can push, so a step's pushes always fit after a
// spill.

// Next group for this wave.  With per-XCD queues the groups (PH order) are cut
// into kXcd contiguous ranges; a wave drains the range of its own XCD first,
// so neighbouring groups share that XCD's L2, then helps the other ranges.
// Every lane joins the atomics (lane 0 adds 1) and all branches are uniform.
__device__ __forceinline__ int next_group(DevCounters *dc, int ngroups, int home, int &xr) {
  const int lane = lane_id();
  if (!BT_XCD_QUEUE) return wave_uniform(atomicAdd(&dc->queue, lane == 0 ? 1 : 0));
  while (xr < kXcd) {
    const int x = (home + xr) & (kXcd - 1);
    const int lo = (int)((long long)ngroups * x / kXcd);
    const int hi = (int)((long long)ngroups * (x + 1) / kXcd);
    const int k = wave_uniform(atomicAdd(&dc->xq[x], lane == 0 ? 1 : 0));
    if (lo + k < hi) return lo + k;
    xr++;
  }
  return ngroups;
}

struct WalkArgs {
  int n, nsys;
  float eps2;
  const int2 *groups;
  const int *group_sys;
  const int2 *root_cells;
  const float4 *pos_ph;
  const float4 *pos_tree;
  const uint4 *cells;
  const float4 *csize;
  const float4 *cmono;
  const float4 *cq0;
  const float2 *cq1;
  const float4 *cwalk;   // 3 per cell: (com, M) (3Qxx,3Qyy,3Qzz,3Qxy) (3Qxz,3Qyz,trQ/2,1.5trQ)
  // learned opening angle
  const float2 *cls;     // (l, s) per cell
  const float2 *sys_scale;
  const MlpDev *mlp;
  float shift, it_min, it_max;
  float4 *acc;
  int2 *gcounts;
  double *gpot;
  int *stack;
  int stack_ints;
  int lstk_cap;          // LDS part of a learned walk's stack, entries per wave
  const int *run_if;     // nullptr, or a device flag: walk only if it is nonzero
  DevCounters *dc;
};

struct WaveShared {
  float4 jA[kWave];   // P2C: (com - gc, M)         P2P: (x - gc, m)
  float4 jB[kWave];   // P2C: (3Qxx, 3Qyy, 3Qzz, 3Qxy)
  float4 jC[kWave];   // P2C: (3Qxz, 3Qyz, trQ/2, 1.5 trQ)
  int alist[2 * kWave];
  int dlist[2 * kWave];
  int lpre[kWave];
  int lbeg[kWave];
};

// Improved Barnes-Hut opening test, code/computeForces.cu:30-54 (same float ops).
__device__ __forceinline__ bool open_cell(const float4 cs, float gcx, float gcy, float gcz,
                                          float gsx, float gsy, float gsz) {
#pragma clang fp contract(off)
  float dx = fabsf(gcx - cs.x) - gsx;
  float dy = fabsf(gcy - cs.y) - gsy;
  float dz = fabsf(gcz - cs.z) - gsz;
  dx += fabsf(dx); dx *= 0.5f;
  dy += fabsf(dy); dy *= 0.5f;
  dz += fabsf(dz); dz *= 0.5f;
  const float ds2 = dx * dx + dy * dy + dz * dz;
  return ds2 < fabsf(cs.w);
}

// ds^2 of the opening test above.
__device__ __forceinline__ float box_ds2(const float4 cs, float gcx, float gcy, float gcz,
                                         float gsx, float gsy, float gsz) {
#pragma clang fp contract(off)
  float dx = fabsf(gcx - cs.x) - gsx;
  float dy = fabsf(gcy - cs.y) - gsy;
  float dz = fabsf(gcz - cs.z) - gsz;
  dx += fabsf(dx); dx *= 0.5f;
  dy += fabsf(dy); dy *= 0.5f;
  dz += fabsf(dz); dz *= 0.5f;
  return dx * dx + dy * dy + dz * dz;
}

template <bool KAHAN>
__device__ __forceinline__ void fold(Acc4 &s, Acc4 &c, const Acc4 &b) {
  if (KAHAN) {
    float y, t;
    y = b.x - c.x; t = s.x + y; c.x = (t - s.x) - y; s.x = t;
    y = b.y - c.y; t = s.y + y; c.y = (t - s.y) - y; s.y = t;
    y = b.z - c.z; t = s.z + y; c.z = (t - s.z) - y; s.z = t;
    y = b.w - c.w; t = s.w + y; c.w = (t - s.w) - y; s.w = t;
  } else {
    s.x += b.x; s.y += b.y; s.z += b.z; s.w += b.w;
  }
}

template <bool KAHAN>
__device__ __forceinline__ void eval_p2c(const WalkArgs &a, WaveShared &S, int cnt, float gcx,
                                         float gcy, float gcz, float xi, float yi, float zi,
                                         Acc4 &s, Acc4 &c) {
  const int lane = lane_id();
  if (lane < cnt) {
    const int ci = S.alist[lane];
    const float4 r0 = a.cwalk[3 * ci];
    S.jA[lane] = make_float4(r0.x - gcx, r0.y - gcy, r0.z - gcz, r0.w);
    S.jB[lane] = a.cwalk[3 * ci + 1];
    S.jC[lane] = a.cwalk[3 * ci + 2];
  }
  wave_lds_sync();
  Acc4 b = {0.f, 0.f, 0.f, 0.f};
  if (cnt == kWave) {
#pragma unroll BT_UNROLL_LDS
    for (int j = 0; j < kWave; j++) p2c_rec(b, xi, yi, zi, S.jA[j], S.jB[j], S.jC[j], a.eps2);
  } else {
#pragma unroll 2
    for (int j = 0; j < cnt; j++) p2c_rec(b, xi, yi, zi, S.jA[j], S.jB[j], S.jC[j], a.eps2);
  }
  fold<KAHAN>(s, c, b);
  wave_lds_sync();
}

template <bool KAHAN>
__device__ __forceinline__ void eval_p2p(const WalkArgs &a, WaveShared &S, int cnt, float gcx,
                                         float gcy, float gcz, float xi, float yi, float zi,
                                         Acc4 &s, Acc4 &c) {
  const int lane = lane_id();
  if (lane < cnt) {
    const float4 p = a.pos_tree[S.dlist[lane]];
    S.jA[lane] = make_float4(p.x - gcx, p.y - gcy, p.z - gcz, p.w);
  }
  wave_lds_sync();
  Acc4 b = {0.f, 0.f, 0.f, 0.f};
  if (cnt == kWave) {
#pragma unroll BT_UNROLL_LDS
    for (int j = 0; j < kWave; j++) p2p(b, xi, yi, zi, S.jA[j], a.eps2);
  } else {
#pragma unroll 2
    for (int j = 0; j < cnt; j++) p2p(b, xi, yi, zi, S.jA[j], a.eps2);
  }
  fold<KAHAN>(s, c, b);
  wave_lds_sync();
}

// Monopole-only P2C (the reference built without QUADRUPOLE,
// code/computeForces.cu:198-225): the cell's mass at its centre of mass,
// evaluated like a particle (20 flops).
template <bool KAHAN>
__device__ __forceinline__ void eval_p2c_mono(const WalkArgs &a, WaveShared &S, int cnt,
                                              float gcx, float gcy, float gcz, float xi,
                                              float yi, float zi, Acc4 &s, Acc4 &c) {
  const int lane = lane_id();
  if (lane < cnt) {
    const float4 r0 = a.cwalk[3 * S.alist[lane]];
    S.jA[lane] = make_float4(r0.x - gcx, r0.y - gcy, r0.z - gcz, r0.w);
  }
  wave_lds_sync();
  Acc4 b = {0.f, 0.f, 0.f, 0.f};
  if (cnt == kWave) {
#pragma unroll 8
    for (int j = 0; j < kWave; j++) p2p(b, xi, yi, zi, S.jA[j], a.eps2);
  } else {
#pragma unroll 2
    for (int j = 0; j < cnt; j++) p2p(b, xi, yi, zi, S.jA[j], a.eps2);
  }
  fold<KAHAN>(s, c, b);
  wave_lds_sync();
}

// Cell records and particles do not change during the walk: reading them
// through the constant address space lets the compiler fetch a wave-uniform
// record with s_load (scalar cache) instead of a vector load.
typedef float f4v __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(4))) const f4v cf4v;

__device__ __forceinline__ float4 ld_uniform(const cf4v *p, int i) {
  const f4v v = p[i];
  return make_float4(v[0], v[1], v[2], v[3]);
}

template <bool KAHAN>
__device__ __forceinline__ void eval_p2c_s(const WalkArgs &a, WaveShared &S, int cnt, float xi,
                                           float yi, float zi, Acc4 &s, Acc4 &c) {
  const cf4v *rec = (const cf4v *)a.cwalk;
  const int my = S.alist[lane_id()];
  Acc4 b = {0.f, 0.f, 0.f, 0.f};
  if (cnt == kWave) {
#pragma unroll BT_UNROLL_S
    for (int j = 0; j < kWave; j++) {
      const int r = 3 * __builtin_amdgcn_readlane(my, j);
      p2c_rec(b, xi, yi, zi, ld_uniform(rec, r), ld_uniform(rec, r + 1), ld_uniform(rec, r + 2),
              a.eps2);
    }
  } else {
#pragma unroll 2
    for (int j = 0; j < cnt; j++) {
      const int r = 3 * __builtin_amdgcn_readlane(my, j);
      p2c_rec(b, xi, yi, zi, ld_uniform(rec, r), ld_uniform(rec, r + 1), ld_uniform(rec, r + 2),
              a.eps2);
    }
  }
  fold<KAHAN>(s, c, b);
}

template <bool KAHAN>
__device__ __forceinline__ void eval_p2p_s(const WalkArgs &a, WaveShared &S, int cnt, float xi,
                                           float yi, float zi, Acc4 &s, Acc4 &c) {
  const cf4v *pj = (const cf4v *)a.pos_tree;
  const int my = S.dlist[lane_id()];
  Acc4 b = {0.f, 0.f, 0.f, 0.f};
  if (cnt == kWave) {
#pragma unroll BT_UNROLL_S
    for (int j = 0; j < kWave; j++) p2p(b, xi, yi, zi, ld_uniform(pj, __builtin_amdgcn_readlane(my, j)), a.eps2);
  } else {
#pragma unroll 2
    for (int j = 0; j < cnt; j++) p2p(b, xi, yi, zi, ld_uniform(pj, __builtin_amdgcn_readlane(my, j)), a.eps2);
  }
  fold<KAHAN>(s, c, b);
}

// One batch of cnt P2C (list in S.alist) or P2P (S.dlist) interactions.  The
// LDS path takes i relative to the group centre, the scalar path absolute i.
template <bool KAHAN, bool SJ, bool MONO>
__device__ __forceinline__ void batch_p2c(const WalkArgs &a, WaveShared &S, int cnt, float gcx,
                                          float gcy, float gcz, float xi, float yi, float zi,
                                          Acc4 &s, Acc4 &c) {
  if (MONO) eval_p2c_mono<KAHAN>(a, S, cnt, gcx, gcy, gcz, xi, yi, zi, s, c);
  else if (SJ) eval_p2c_s<KAHAN>(a, S, cnt, xi, yi, zi, s, c);
  else eval_p2c<KAHAN>(a, S, cnt, gcx, gcy, gcz, xi, yi, zi, s, c);
}

template <bool KAHAN, bool SJ>
__device__ __forceinline__ void batch_p2p(const WalkArgs &a, WaveShared &S, int cnt, float gcx,
                                          float gcy, float gcz, float xi, float yi, float zi,
                                          Acc4 &s, Acc4 &c) {
  if (SJ) eval_p2p_s<KAHAN>(a, S, cnt, xi, yi, zi, s, c);
  else eval_p2p<KAHAN>(a, S, cnt, gcx, gcy, gcz, xi, yi, zi, s, c);
}

// LEARN: each popped cell gets its own opening angle from the network
// (bt_learn.h), evaluated by the wave on the matrix cores for up to 64 cells.
// MONO: monopole-only P2C (the reference's build without QUADRUPOLE).
template <bool KAHAN, bool SJ, bool LEARN, bool MONO = false>
__global__ __launch_bounds__(kBlock)
__attribute__((amdgpu_waves_per_eu(LEARN ? BT_LEARN_WAVES_PER_EU : BT_WALK_WAVES_PER_EU)))
void k_walk(WalkArgs a) {
  __shared__ WaveShared sh[kWavesPerBlock];
  __shared__ float s_mlp[LEARN ? 64 : 1];
  __shared__ int s_stack[LEARN ? kWavesPerBlock * kLdsStack : 1];
  const int lane = lane_id();
  const int wv = threadIdx.x >> 6;
  WaveShared &S = sh[wv];
  // LSTK: the top of the stack (up to kLdsStack entries) lives in LDS, the
  // older entries below it in global memory (stk[0, gtop)); otherwise the
  // whole stack is in global memory (stk[0, top))
  constexpr bool LSTK = LEARN;
  int *stk = a.stack + (size_t)(blockIdx.x * kWavesPerBlock + wv) * (size_t)a.stack_ints;
  int *lstk = s_stack + (LSTK ? wv * kLdsStack : 0);
  half8_t ma1, ma2_0, ma2_1;
  float mb3 = 0.f;
  if (LEARN) {
    if (threadIdx.x < 32) s_mlp[threadIdx.x] = a.mlp->c2p[threadIdx.x >> 4][threadIdx.x & 15];
    else if (threadIdx.x < 64)
      s_mlp[threadIdx.x] = a.mlp->w3p[(threadIdx.x - 32) >> 4][threadIdx.x & 15];
    ma1 = a.mlp->a1[lane];
    ma2_0 = a.mlp->a2[0][lane];
    ma2_1 = a.mlp->a2[1][lane];
    mb3 = a.mlp->b3;
    __syncthreads();
  }
  if (a.run_if && *a.run_if == 0) return;   // a fallback walk that is not needed
  const int ngroups = a.dc->ngroups;
  // a failed build (cell or group capacity) leaves placeholder links: do not walk it
  if (a.dc->err & 3) return;

  // The persistent loop has no lane-dependent branches at its ends: two
  // "if (lane == 0)" blocks around the back edge (queue fetch, stats store) are
  // merged by the compiler into a region that only lane 0 re-enters, leaving
  // the other lanes to loop on a stale group index.  Every lane therefore
  // joins the queue atomic (lane 0 adds 1, the others 0) and the per-group
  // stores below are branch-free.
  const int home = (int)(blockIdx.x % kXcd);
  int xr = 0;
  for (;;) {
    const int g = next_group(a.dc, ngroups, home, xr);
    if (g >= ngroups) break;
    const int2 grp = a.groups[g];
    const int pbeg = grp.x, np = max(grp.y, 1);
    const int sys = a.nsys > 1 ? a.group_sys[g] : 0;
    const bool valid = lane < np;
    const float4 pi = a.pos_ph[pbeg + min(lane, np - 1)];

    // group box (code/computeForces.cu:539-555)
    const float mnx = wave_min(pi.x), mny = wave_min(pi.y), mnz = wave_min(pi.z);
    const float mxx = wave_max(pi.x), mxy = wave_max(pi.y), mxz = wave_max(pi.z);
    float gcx, gcy, gcz, gsx, gsy, gsz;
    {
#pragma clang fp contract(off)
      gcx = 0.5f * (mxx + mnx); gcy = 0.5f * (mxy + mny); gcz = 0.5f * (mxz + mnz);
      gsx = 0.5f * (mxx - mnx); gsy = 0.5f * (mxy - mny); gsz = 0.5f * (mxz - mnz);
    }
    GroupGeom geom;
    if (LEARN) geom = make_group_geom(gcx, gcy, gcz, gsx, gsy, gsz, np, a.sys_scale[sys]);
    // i-particle: relative to the group centre (LDS path) or absolute (scalar path)
    const float xi = SJ ? pi.x : pi.x - gcx, yi = SJ ? pi.y : pi.y - gcy,
                zi = SJ ? pi.z : pi.z - gcz;

    Acc4 s = {0.f, 0.f, 0.f, 0.f}, cmp = {0.f, 0.f, 0.f, 0.f};
    const int2 rc = a.root_cells[sys];
    if (lane < rc.y) {
      if (LSTK) lstk[lane] = rc.x + lane;
      else stk[lane] = rc.x + lane;
    }
    int top = rc.y;
    int gtop = 0;                // LSTK: entries spilled to global memory
    int na = 0, nd = 0;          // pending list lengths
    int n_app = 0, n_dir = 0;    // interactions evaluated (stats)
    int deepest = top;           // stack high-water mark of this group
    bool overflow = false;
    int steps = 0;
    __builtin_amdgcn_s_waitcnt(0);

    while (top > 0 || gtop > 0) {
      // every cell is visited at most once per group; a longer walk means a
      // corrupt tree, so stop rather than loop
      if (++steps > kMaxWalkSteps) {
        overflow = true;
        break;
      }
      if (LSTK && top < kWave && gtop > 0) {
        // top the LDS part up with the newest spilled entries (slid in below
        // the ones left), so that the step pops the same cells as it would
        // from one stack
        const int r = min(gtop, a.lstk_cap / 2);
        const int v = lane < top ? lstk[lane] : 0;
        wave_lds_sync();
        if (lane < top) lstk[r + lane] = v;
        for (int k = lane; k < r; k += kWave) lstk[k] = stk[gtop - r + k];
        gtop -= r;
        top += r;
        wave_lds_sync();
      }
      const int take = min(kWave, top);
      top -= take;
      const bool use = lane < take;
      int ci = 0;
      uint4 cd = make_uint4(0, 0, 0, 0);
      float4 cs = make_float4(0.f, 0.f, 0.f, 0.f);
      if (use) {
        ci = LSTK ? lstk[top + lane] : stk[top + lane];
        cd = a.cells[ci];
        cs = a.csize[ci];
      }
      bool split = false, leaf = false;
      if (LEARN) {
        float2 lsv = make_float2(0.f, 0.f);
        if (use) lsv = a.cls[ci];
        // A cell opened even at theta_hi or accepted even at theta_lo is
        // decided without the network: (l/theta + s)^2 is monotone in 1/theta,
        // so every angle of the band gives it the same answer.  Only a step
        // with an undecided cell loads the moments and computes features, and
        // only a 32-lane block holding one evaluates the network.
        const float ds2 = box_ds2(cs, gcx, gcy, gcz, gsx, gsy, gsz);
        float op_lo, op_hi;
        {
#pragma clang fp contract(off)
          op_lo = lsv.x * a.it_max + lsv.y;
          op_hi = lsv.x * a.it_min + lsv.y;
          op_lo = op_lo * op_lo;
          op_hi = op_hi * op_hi;
        }
        const bool undecided = use && ds2 < op_lo && ds2 >= op_hi;
        const uint64_t um = ballot64(undecided);
        float it = a.it_min;
        if (um != 0) {
          // every lane evaluates the network (decided and idle lanes on a zero cell)
          float4 m = make_float4(0.f, 0.f, 0.f, 0.f), q0 = m;
          float2 q1 = make_float2(0.f, 0.f);
          if (undecided) {
            m = a.cmono[ci];
            q0 = a.cq0[ci];
            q1 = a.cq1[ci];
          }
          float x[8];
          pair_features(geom, m, lsv.x, q0, q1, x);
          const float y = mlp_wave(ma1, ma2_0, ma2_1, s_mlp, mb3, pack_features(x),
                                   (um & 0xFFFFFFFFull) != 0, (um >> 32) != 0);
          it = learned_inv_theta(y, a.shift, a.it_min, a.it_max);
        }
        float op;
        {
#pragma clang fp contract(off)
          op = lsv.x * it + lsv.y;
          op = op * op;
        }
        if (use) {
          // the opening test of open_cell with this pair's radius (same ds^2)
          split = ds2 < op || ((int)cd.w - (int)cd.z < 3);
          leaf = cell_leaf(cd);
        }
      } else if (use) {
        split = open_cell(cs, gcx, gcy, gcz, gsx, gsy, gsz) || ((int)cd.w - (int)cd.z < 3);
        leaf = cell_leaf(cd);
      }
      const bool push = use && split && !leaf;
      const bool approx = use && !split;
      const bool direct = use && split && leaf;

      // children of opened nodes go on the stack
      const int nch = push ? cell_nchild(cd) : 0;
      const int chi = wave_incl_scan(nch);
      const int ctot = __shfl(chi, kWave - 1, kWave);
      if (LSTK && top + ctot > a.lstk_cap) {
        // spill the oldest LDS entries to the global stack, leaving a quarter
        // of the LDS part free after this step's pushes where possible
        const int sp = min(top, top + ctot - a.lstk_cap + a.lstk_cap / 4);
        if (top - sp + ctot > a.lstk_cap || gtop + sp > a.stack_ints) {
          overflow = true;
          break;
        }
        for (int k = lane; k < sp; k += kWave) stk[gtop + k] = lstk[k];
        wave_lds_sync();
        // move the rest down; each chunk is read in full before it is written
        for (int k0 = 0; k0 < top - sp; k0 += kWave) {
          const int k = k0 + lane;
          const int v = k < top - sp ? lstk[sp + k] : 0;
          wave_lds_sync();
          if (k < top - sp) lstk[k] = v;
          wave_lds_sync();
        }
        gtop += sp;
        top -= sp;
      } else if (!LSTK && top + ctot > a.stack_ints) {
        overflow = true;
        break;
      }
      if (push) {
        const int f = cell_first(cd);
        const int base = top + chi - nch;
        for (int k = 0; k < nch; k++) {
          if (LSTK) lstk[base + k] = f + k;
          else stk[base + k] = f + k;
        }
      }
      top += ctot;
      deepest = max(deepest, top + gtop);

      // approximate list (P2C)
      const uint64_t am = ballot64(approx);
      if (approx) S.alist[na + __popcll(am & lanemask_lt())] = ci;
      na += __popcll(am);

      // direct list (P2P): expand the opened leaves' particle ranges
      const uint64_t dm = ballot64(direct);
      const int ndl = __popcll(dm);
      if (ndl > 0) {
        const int lsz = direct ? (int)cd.w - (int)cd.z : 0;
        const int li = wave_incl_scan(lsz);
        const int ltot = __shfl(li, kWave - 1, kWave);
        if (direct) {
          const int r = __popcll(dm & lanemask_lt());
          S.lpre[r] = li - lsz;
          S.lbeg[r] = (int)cd.z;
        }
        wave_lds_sync();
        for (int k0 = 0; k0 < ltot; k0 += kWave) {
          const int k = k0 + lane;
          if (k < ltot) {
            int lo = 0, hi = ndl - 1;  // last leaf with lpre <= k
            while (lo < hi) {
              const int mid = (lo + hi + 1) >> 1;
              if (S.lpre[mid] <= k) lo = mid; else hi = mid - 1;
            }
            S.dlist[nd + lane] = S.lbeg[lo] + (k - S.lpre[lo]);
          }
          nd += min(kWave, ltot - k0);
          wave_lds_sync();
          if (nd >= kWave) {
            batch_p2p<KAHAN, SJ>(a, S, kWave, gcx, gcy, gcz, xi, yi, zi, s, cmp);
            n_dir += kWave;
            nd -= kWave;
            if (lane < nd) S.dlist[lane] = S.dlist[kWave + lane];
            wave_lds_sync();
          }
        }
      }
      wave_lds_sync();
      if (na >= kWave) {
        batch_p2c<KAHAN, SJ, MONO>(a, S, kWave, gcx, gcy, gcz, xi, yi, zi, s, cmp);
        n_app += kWave;
        na -= kWave;
        if (lane < na) S.alist[lane] = S.alist[kWave + lane];
        wave_lds_sync();
      }
      // stack pushes of this step (stores by some lanes) must have completed
      // before the next step's pops (loads by other lanes)
      __builtin_amdgcn_s_waitcnt(0);
    }
    if (overflow) {
      atomicOr(&a.dc->err, 4);  // every lane, same bit
      atomicOr(&a.dc->err_sticky, 4);
      na = 0;
      nd = 0;
    }
    if (na > 0) {
      batch_p2c<KAHAN, SJ, MONO>(a, S, na, gcx, gcy, gcz, xi, yi, zi, s, cmp);
      n_app += na;
    }
    if (nd > 0) {
      batch_p2p<KAHAN, SJ>(a, S, nd, gcx, gcy, gcz, xi, yi, zi, s, cmp);
      n_dir += nd;
    }
    // lanes past the end of the group write to the kWave spare slots after acc[n)
    a.acc[valid ? pbeg + lane : a.n + lane] = make_float4(s.x, s.y, s.z, s.w);
    const double gp = wave_sum(valid ? (double)pi.w * (double)s.w : 0.0);
    a.gcounts[g] = overflow ? make_int2(-1, -1) : make_int2(n_app, n_dir);
    a.gpot[g] = gp;
    atomicMax(&a.dc->stack_max, lane == 0 ? deepest : 0);
    atomicMax(&a.dc->stack_hw, lane == 0 ? deepest : 0);
  }
}

// Interaction statistics (code/computeForces.cu:567-609, :701-715): per valid
// i-particle its group's counts; {direct_sum, direct_max, approx_sum,
// approx_max} and 1/2 sum m phi.  kWalkStatBlocks blocks reduce contiguous
// group ranges, one block combines their partials: a fixed order, so the
// FP64 potential sum is the same on every run.
struct WStat {
  unsigned long long ds, as;
  unsigned int dm, am;
  double p;
};

// 256 threads; the block's total in thread 0
__device__ __forceinline__ WStat block_stats(WStat v, WStat *sh) {
  v.ds = wave_sum(v.ds);
  v.as = wave_sum(v.as);
  v.p = wave_sum(v.p);
  for (int m = 32; m >= 1; m >>= 1) {
    v.dm = max(v.dm, (unsigned int)__shfl_xor((int)v.dm, m, kWave));
    v.am = max(v.am, (unsigned int)__shfl_xor((int)v.am, m, kWave));
  }
  const int w = threadIdx.x >> 6;
  if (lane_id() == 0) sh[w] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < kWavesPerBlock; k++) {
      v.ds += sh[k].ds; v.as += sh[k].as; v.p += sh[k].p;
      v.dm = max(v.dm, sh[k].dm); v.am = max(v.am, sh[k].am);
    }
  }
  return v;
}

__global__ __launch_bounds__(kBlock) void k_walk_stats_part(const int2 *__restrict__ groups,
                                                           const int2 *__restrict__ gcounts,
                                                           const double *__restrict__ gpot,
                                                           const DevCounters *dc, WStat *part,
                                                           const int *run_if) {
  if (run_if && *run_if == 0) return;
  __shared__ WStat sh[kWavesPerBlock];
  const int ng = dc->ngroups;
  const int chunk = (ng + gridDim.x - 1) / gridDim.x;
  const int g0 = blockIdx.x * chunk, g1 = min(ng, g0 + chunk);
  WStat v = {0ull, 0ull, 0u, 0u, 0.0};
  for (int g = g0 + threadIdx.x; g < g1; g += kBlock) {
    const int2 c = gcounts[g];
    const int np = groups[g].y;
    v.ds += (unsigned long long)c.y * (unsigned long long)np;
    v.as += (unsigned long long)c.x * (unsigned long long)np;
    v.dm = max(v.dm, (unsigned int)c.y);
    v.am = max(v.am, (unsigned int)c.x);
    v.p += gpot[g];
  }
  v = block_stats(v, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = v;
}

__global__ __launch_bounds__(kBlock) void k_walk_stats_final(const WStat *__restrict__ part,
                                                            unsigned long long *out, double *pot,
                                                            const int *run_if) {
  if (run_if && *run_if == 0) return;
  __shared__ WStat sh[kWavesPerBlock];
  WStat v = {0ull, 0ull, 0u, 0u, 0.0};
  for (int k = threadIdx.x; k < kWalkStatBlocks; k += kBlock) {
    const WStat q = part[k];
    v.ds += q.ds; v.as += q.as; v.p += q.p;
    v.dm = max(v.dm, q.dm); v.am = max(v.am, q.am);
  }
  v = block_stats(v, sh);
  if (threadIdx.x == 0) {
    out[0] = v.ds; out[1] = v.dm; out[2] = v.as; out[3] = v.am;
    *pot = 0.5 * v.p;
  }
}

__global__ void k_walk_reset(DevCounters *dc, const int *run_if) {
  if (threadIdx.x == 0 && !(run_if && *run_if == 0)) {
    dc->queue = 0;
    dc->stack_max = 0;
    for (int x = 0; x < kXcd; x++) dc->xq[x] = 0;
    dc->err &= ~4;
  }
}

}  // namespace

void launch_walk(Ctx &c, const int *run_if) {
  WalkArgs a;
  a.n = c.n;
  a.nsys = c.nsys;
  a.eps2 = c.eps * c.eps;
  a.groups = c.groups;
  a.group_sys = c.group_sys;
  a.root_cells = c.root_cells;
  a.pos_ph = c.pos_ph;
  a.pos_tree = c.pos;
  a.cells = c.cells;
  a.csize = c.csize;
  a.cmono = c.cmono;
  a.cq0 = c.cq0;
  a.cq1 = c.cq1;
  a.cwalk = c.cwalk;
  a.cls = c.cls;
  a.sys_scale = c.sys_scale;
  a.mlp = c.mlp_dev;
  a.shift = c.learn_shift;
  a.it_min = 1.0f / c.learn_theta_hi;
  a.it_max = 1.0f / c.learn_theta_lo;
  a.acc = c.acc;
  a.gcounts = c.gcounts;
  a.gpot = c.gpot;
  a.stack = c.stack;
  a.stack_ints = c.stack_ints;
  a.lstk_cap = c.lds_stack;
  a.run_if = run_if;
  a.dc = c.dc;
  hipLaunchKernelGGL(k_walk_reset, dim3(1), dim3(64), 0, c.stream, c.dc, run_if);
  void (*kern)(WalkArgs);
  if (!c.quadrupole)
    kern = c.kahan ? k_walk<true, false, false, true> : k_walk<false, false, false, true>;
  else if (c.opening == 1 && c.have_mlp && c.walk_variant == 1)
    kern = c.kahan ? k_walk<true, true, true> : k_walk<false, true, true>;
  else if (c.opening == 1 && c.have_mlp)
    kern = c.kahan ? k_walk<true, false, true> : k_walk<false, false, true>;
  else if (c.walk_variant == 1)
    kern = c.kahan ? k_walk<true, true, false> : k_walk<false, true, false>;
  else
    kern = c.kahan ? k_walk<true, false, false> : k_walk<false, false, false>;
  hipLaunchKernelGGL(kern, dim3(c.walk_blocks), dim3(kBlock), 0, c.stream, a);
  WStat *part = reinterpret_cast<WStat *>(c.d_wstats + 4);
  hipLaunchKernelGGL(k_walk_stats_part, dim3(kWalkStatBlocks), dim3(kBlock), 0, c.stream, c.groups,
                     c.gcounts, c.gpot, c.dc, part, run_if);
  hipLaunchKernelGGL(k_walk_stats_final, dim3(1), dim3(kBlock), 0, c.stream, part, c.d_wstats,
                     c.d_wpot, run_if);
}

}  // namespace bt
