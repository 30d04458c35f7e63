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
#ifndef BT_UNROLL_LDS2
#define BT_UNROLL_LDS2 4
#endif
#ifndef BT_WALK_WAVES_PER_EU
#define BT_WALK_WAVES_PER_EU 1
#endif
#ifndef BT_LEARN_WAVES_PER_EU
#define BT_LEARN_WAVES_PER_EU 4
#endif
#ifndef BT_LEARN2_WAVES_PER_EU
#define BT_LEARN2_WAVES_PER_EU 3   // two i-particles per lane: 4 would spill VGPRs
#endif
#ifndef BT_XCD_QUEUE
#define BT_XCD_QUEUE 1
#endif
constexpr int kXcd = 8;   // MI355X: 8 XCDs, workgroups dispatched round-robin over them
// Entries of an FP32 walk's stack held in LDS per wave (kLdsStack = 1024,
// 4 KiB).  The FP32 walks run at 4 waves per SIMD (their registers allow no
// more), so 4 blocks of 34 KiB share a CU's 160 KiB of LDS; the FP64 walk
// keeps its whole stack in global memory.  A deeper stack spills its oldest
// entries to the wave's global stack.  The part used (WalkArgs::lstk_cap) is
// at least the 8 * 64 children one step can push, so a step's pushes always
// fit after a spill.

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

// R: the real type of particles, cell data and arithmetic (float for
// Treecode<float>, double for Treecode<double>).
template <typename R>
struct WalkArgs {
  int n, nsys;
  R eps2;
  const int2 *groups;
  const int *group_sys;
  const int2 *root_cells;
  const V4<R> *pos_ph;
  const V4<R> *pos_tree;
  const uint4 *cells;
  const V4<R> *csize;
  const V4<R> *cmono;
  const V4<R> *cq0;
  const V2<R> *cq1;
  const V4<R> *cwalk;    // 3 per cell: (com, M) (3Qxx,3Qyy,3Qzz,3Qxy) (3Qxz,3Qyz,trQ/2,1.5trQ)
  // learned opening angle (Treecode<float> only)
  const float2 *cls;     // (l, s) per cell
  const float2 *sys_scale;
  const MlpDev *mlp;
  float shift, it_min, it_max;
  V4<R> *acc;
  int2 *gcounts;
  double *gpot;
  int *stack;
  int stack_ints;
  int lstk_cap;          // LDS part of a learned walk's stack, entries per wave
  const int *run_if;     // nullptr, or a device flag: walk only if it is nonzero
  DevCounters *dc;
};

// The i-particles one lane carries.  IPL = 1: particle `lane` of the group.
// IPL = 2: particles 2q and 2q + 1 (q = lane & 31), carried by both halves of
// the wave; each batch of j-entries is split between the half-waves (lanes
// 0-31 sweep its first half, lanes 32-63 its second), so every broadcast LDS
// read serves two interactions, and the halves' sums are added once per group.
template <typename R, int IPL>
struct IParts {
  R x[IPL], y[IPL], z[IPL];
  Acc4T<R> s[IPL], c[IPL];   // Kahan sums and compensations
};

// LDS slots of one batch: 64 entries; with IPL = 2 the second half-wave's
// entries start one slot later (its broadcast address differs from the first
// half's by 132 dwords, not a multiple of the bank count) and one zero-mass
// slot follows an odd count.
constexpr int kJSlots = kWave + 2;

template <typename R>
struct WaveShared {
  V4<R> jA[kJSlots];    // P2C: (com - gc, M)         P2P: (x - gc, m)
  V4<R> jB[kJSlots];    // P2C: (3Qxx, 3Qyy, 3Qzz, 3Qxy)
  V4<R> jC[kJSlots];    // P2C: (3Qxz, 3Qyz, trQ/2, 1.5 trQ)
  int alist[2 * kWave];
  int dlist[2 * kWave];
  int lpre[kWave];
  int lbeg[kWave];
};

__device__ __forceinline__ float rabs(float v) { return __builtin_fabsf(v); }
__device__ __forceinline__ double rabs(double v) { return __builtin_fabs(v); }

// ds^2 of the improved Barnes-Hut opening test, code/computeForces.cu:30-54
// (same operations in real_t).
template <typename R>
__device__ __forceinline__ R box_ds2(const V4<R> cs, R gcx, R gcy, R gcz, R gsx, R gsy, R gsz) {
#pragma clang fp contract(off)
  R dx = rabs(gcx - cs.x) - gsx;
  R dy = rabs(gcy - cs.y) - gsy;
  R dz = rabs(gcz - cs.z) - gsz;
  dx += rabs(dx); dx *= R(0.5);
  dy += rabs(dy); dy *= R(0.5);
  dz += rabs(dz); dz *= R(0.5);
  return dx * dx + dy * dy + dz * dz;
}

// The test itself: open iff ds^2 < cellOp^2.
template <typename R>
__device__ __forceinline__ bool open_cell(const V4<R> cs, R gcx, R gcy, R gcz, R gsx, R gsy,
                                          R gsz) {
  return box_ds2<R>(cs, gcx, gcy, gcz, gsx, gsy, gsz) < rabs(cs.w);
}

template <bool KAHAN, typename R>
__device__ __forceinline__ void fold(Acc4T<R> &s, Acc4T<R> &c, const Acc4T<R> &b) {
  if (KAHAN) {
    R y, t;
    y = b.x - c.x; t = s.x + y; c.x = (t - s.x) - y; s.x = t;
    y = b.y - c.y; t = s.y + y; c.y = (t - s.y) - y; s.y = t;
    y = b.z - c.z; t = s.z + y; c.z = (t - s.z) - y; s.z = t;
    y = b.w - c.w; t = s.w + y; c.w = (t - s.w) - y; s.w = t;
  } else {
    s.x += b.x; s.y += b.y; s.z += b.z; s.w += b.w;
  }
}

// Entries per half-wave of a batch of cnt, and the LDS slot of entry k.
template <int IPL>
__device__ __forceinline__ int half_count(int cnt) { return IPL == 2 ? (cnt + 1) >> 1 : cnt; }
template <int IPL>
__device__ __forceinline__ int jslot(int k, int hc) { return IPL == 2 ? k + (k >= hc ? 1 : 0) : k; }

// A zero-mass entry far from the group: every term it adds is exactly zero.
template <typename R>
__device__ __forceinline__ void stage_pad(WaveShared<R> &S, int slot) {
  S.jA[slot] = make_v4<R>(R(1.0e18), R(0), R(0), R(0));
  S.jB[slot] = make_v4<R>(R(0), R(0), R(0), R(0));
  S.jC[slot] = make_v4<R>(R(0), R(0), R(0), R(0));
}

// Sweeps a staged batch: P2C (QUAD) or particle-like (P2P and monopole-only
// P2C) interactions of the lane's i-particles with its half of the entries.
template <bool KAHAN, bool QUAD, typename R, int IPL>
__device__ __forceinline__ void sweep(const WaveShared<R> &S, int cnt, R eps2,
                                      IParts<R, IPL> &ip) {
  const int lane = lane_id();
  Acc4T<R> b[IPL];
#pragma unroll
  for (int k = 0; k < IPL; k++) b[k] = Acc4T<R>{R(0), R(0), R(0), R(0)};
  // One entry's j-data, and its interactions with the lane's i-particles.
  struct J {
    V4<R> A, B, C;
  };
  auto load = [&](int slot) {
    J j;
    j.A = S.jA[slot];
    if (QUAD) {
      j.B = S.jB[slot];
      j.C = S.jC[slot];
    }
    return j;
  };
  auto eval = [&](const J &j) {
    if (QUAD) {
#pragma unroll
      for (int k = 0; k < IPL; k++) p2c_rec(b[k], ip.x[k], ip.y[k], ip.z[k], j.A, j.B, j.C, eps2);
    } else {
#pragma unroll
      for (int k = 0; k < IPL; k++) p2p(b[k], ip.x[k], ip.y[k], ip.z[k], j.A, eps2);
    }
  };
  // Entry j + 1 is read from LDS while entry j is evaluated, so the reads'
  // latency is hidden behind one interaction's arithmetic.  The read past the
  // last entry stays inside the kJSlots slots and its value is not used.
  if (IPL == 2) {
    const int hc = half_count<2>(cnt);
    const int base = lane >= 32 ? hc + 1 : 0;
    J cur = load(base);
    if (cnt == kWave) {
#pragma unroll BT_UNROLL_LDS2
      for (int j = 0; j < kWave / 2; j++) {
        const J nxt = load(base + j + 1);
        __builtin_amdgcn_sched_barrier(0);   // keep the reads ahead of the arithmetic
        eval(cur);
        cur = nxt;
      }
    } else {
#pragma unroll 2
      for (int j = 0; j < hc; j++) {
        const J nxt = load(base + j + 1);
        __builtin_amdgcn_sched_barrier(0);   // keep the reads ahead of the arithmetic
        eval(cur);
        cur = nxt;
      }
    }
  } else {
    J cur = load(0);
    if (cnt == kWave) {
#pragma unroll BT_UNROLL_LDS
      for (int j = 0; j < kWave; j++) {
        const J nxt = load(j + 1);
        __builtin_amdgcn_sched_barrier(0);   // keep the reads ahead of the arithmetic
        eval(cur);
        cur = nxt;
      }
    } else {
#pragma unroll 2
      for (int j = 0; j < cnt; j++) {
        const J nxt = load(j + 1);
        __builtin_amdgcn_sched_barrier(0);   // keep the reads ahead of the arithmetic
        eval(cur);
        cur = nxt;
      }
    }
  }
#pragma unroll
  for (int k = 0; k < IPL; k++) fold<KAHAN, R>(ip.s[k], ip.c[k], b[k]);
}

template <bool KAHAN, typename R, int IPL>
__device__ __forceinline__ void eval_p2c(const WalkArgs<R> &a, WaveShared<R> &S, int cnt, R gcx,
                                         R gcy, R gcz, IParts<R, IPL> &ip) {
  const int lane = lane_id();
  const int hc = half_count<IPL>(cnt);
  if (lane < cnt) {
    const int ci = S.alist[lane];
    const V4<R> r0 = a.cwalk[3 * ci];
    const int d = jslot<IPL>(lane, hc);
    S.jA[d] = make_v4<R>(r0.x - gcx, r0.y - gcy, r0.z - gcz, r0.w);
    S.jB[d] = a.cwalk[3 * ci + 1];
    S.jC[d] = a.cwalk[3 * ci + 2];
  } else if (IPL == 2 && lane == cnt) {
    stage_pad(S, cnt + 1);
  }
  wave_lds_sync();
  sweep<KAHAN, true, R, IPL>(S, cnt, a.eps2, ip);
  wave_lds_sync();
}

template <bool KAHAN, typename R, int IPL>
__device__ __forceinline__ void eval_p2p(const WalkArgs<R> &a, WaveShared<R> &S, int cnt, R gcx,
                                         R gcy, R gcz, IParts<R, IPL> &ip) {
  const int lane = lane_id();
  const int hc = half_count<IPL>(cnt);
  if (lane < cnt) {
    const V4<R> p = a.pos_tree[S.dlist[lane]];
    S.jA[jslot<IPL>(lane, hc)] = make_v4<R>(p.x - gcx, p.y - gcy, p.z - gcz, p.w);
  } else if (IPL == 2 && lane == cnt) {
    stage_pad(S, cnt + 1);
  }
  wave_lds_sync();
  sweep<KAHAN, false, R, IPL>(S, cnt, a.eps2, ip);
  wave_lds_sync();
}

// Monopole-only P2C (the reference built without QUADRUPOLE,
// code/computeForces.cu:198-225): the cell's mass at its centre of mass,
// evaluated like a particle (20 flops).
template <bool KAHAN, typename R, int IPL>
__device__ __forceinline__ void eval_p2c_mono(const WalkArgs<R> &a, WaveShared<R> &S, int cnt,
                                              R gcx, R gcy, R gcz, IParts<R, IPL> &ip) {
  const int lane = lane_id();
  const int hc = half_count<IPL>(cnt);
  if (lane < cnt) {
    const V4<R> r0 = a.cwalk[3 * S.alist[lane]];
    S.jA[jslot<IPL>(lane, hc)] = make_v4<R>(r0.x - gcx, r0.y - gcy, r0.z - gcz, r0.w);
  } else if (IPL == 2 && lane == cnt) {
    stage_pad(S, cnt + 1);
  }
  wave_lds_sync();
  sweep<KAHAN, false, R, IPL>(S, cnt, a.eps2, ip);
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

// Scalar-load j-data (one i-particle per lane, absolute coordinates).
template <bool KAHAN>
__device__ __forceinline__ void eval_p2c_s(const WalkArgs<float> &a, WaveShared<float> &S,
                                           int cnt, IParts<float, 1> &ip) {
  const cf4v *rec = (const cf4v *)a.cwalk;
  const int my = S.alist[lane_id()];
  Acc4 b = {0.f, 0.f, 0.f, 0.f};
  if (cnt == kWave) {
#pragma unroll BT_UNROLL_S
    for (int j = 0; j < kWave; j++) {
      const int r = 3 * __builtin_amdgcn_readlane(my, j);
      p2c_rec(b, ip.x[0], ip.y[0], ip.z[0], ld_uniform(rec, r), ld_uniform(rec, r + 1),
              ld_uniform(rec, r + 2), a.eps2);
    }
  } else {
#pragma unroll 2
    for (int j = 0; j < cnt; j++) {
      const int r = 3 * __builtin_amdgcn_readlane(my, j);
      p2c_rec(b, ip.x[0], ip.y[0], ip.z[0], ld_uniform(rec, r), ld_uniform(rec, r + 1),
              ld_uniform(rec, r + 2), a.eps2);
    }
  }
  fold<KAHAN>(ip.s[0], ip.c[0], b);
}

template <bool KAHAN>
__device__ __forceinline__ void eval_p2p_s(const WalkArgs<float> &a, WaveShared<float> &S,
                                           int cnt, IParts<float, 1> &ip) {
  const cf4v *pj = (const cf4v *)a.pos_tree;
  const int my = S.dlist[lane_id()];
  Acc4 b = {0.f, 0.f, 0.f, 0.f};
  if (cnt == kWave) {
#pragma unroll BT_UNROLL_S
    for (int j = 0; j < kWave; j++)
      p2p(b, ip.x[0], ip.y[0], ip.z[0], ld_uniform(pj, __builtin_amdgcn_readlane(my, j)), a.eps2);
  } else {
#pragma unroll 2
    for (int j = 0; j < cnt; j++)
      p2p(b, ip.x[0], ip.y[0], ip.z[0], ld_uniform(pj, __builtin_amdgcn_readlane(my, j)), a.eps2);
  }
  fold<KAHAN>(ip.s[0], ip.c[0], b);
}

// One batch of cnt P2C (list in S.alist) or P2P (S.dlist) interactions.  The
// LDS path takes i relative to the group centre, the scalar path absolute i.
template <bool KAHAN, bool SJ, bool MONO, typename R, int IPL>
__device__ __forceinline__ void batch_p2c(const WalkArgs<R> &a, WaveShared<R> &S, int cnt, R gcx,
                                          R gcy, R gcz, IParts<R, IPL> &ip) {
  if constexpr (MONO) eval_p2c_mono<KAHAN, R, IPL>(a, S, cnt, gcx, gcy, gcz, ip);
  else if constexpr (SJ) eval_p2c_s<KAHAN>(a, S, cnt, ip);
  else eval_p2c<KAHAN, R, IPL>(a, S, cnt, gcx, gcy, gcz, ip);
}

template <bool KAHAN, bool SJ, typename R, int IPL>
__device__ __forceinline__ void batch_p2p(const WalkArgs<R> &a, WaveShared<R> &S, int cnt, R gcx,
                                          R gcy, R gcz, IParts<R, IPL> &ip) {
  if constexpr (SJ) eval_p2p_s<KAHAN>(a, S, cnt, ip);
  else eval_p2p<KAHAN, R, IPL>(a, S, cnt, gcx, gcy, gcz, ip);
}

// LEARN: each popped cell gets its own opening angle from the network
// (bt_learn.h), evaluated by the wave on the matrix cores for up to 64 cells.
// MONO: monopole-only P2C (the reference's build without QUADRUPOLE).
template <typename R, bool KAHAN, bool SJ, bool LEARN, bool MONO, int IPL>
__global__ __launch_bounds__(kBlock)
__attribute__((amdgpu_waves_per_eu(LEARN ? (IPL == 2 ? BT_LEARN2_WAVES_PER_EU : BT_LEARN_WAVES_PER_EU)
                                         : BT_WALK_WAVES_PER_EU)))
void k_walk(WalkArgs<R> a) {
  static_assert(!(LEARN || SJ) || sizeof(R) == 4, "learned angle and scalar loads: FP32 only");
  __shared__ WaveShared<R> sh[kWavesPerBlock];
  __shared__ float s_mlp[LEARN ? 64 : 1];
  constexpr bool LSTK = sizeof(R) == 4;
  __shared__ int s_stack[LSTK ? kWavesPerBlock * kLdsStack : 1];
  const int lane = lane_id();
  const int wv = threadIdx.x >> 6;
  WaveShared<R> &S = sh[wv];
  // LSTK: the top of the stack (up to kLdsStack entries) lives in LDS, the
  // older entries below it in global memory (stk[0, gtop)); otherwise the
  // whole stack is in global memory (stk[0, top))
  int *stk = a.stack + (size_t)(blockIdx.x * kWavesPerBlock + wv) * (size_t)a.stack_ints;
  int *lstk = s_stack + (LSTK ? wv * kLdsStack : 0);
  half8_t ma1, ma2_0, ma2_1;
  float mb3 = 0.f;
  if constexpr (LEARN) {
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
  if (a.dc->err & (3 | kSortErr)) return;

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
    // the lane's i-particles (clamped to the last one, :534)
    V4<R> pi[IPL];
#pragma unroll
    for (int k = 0; k < IPL; k++)
      pi[k] = a.pos_ph[pbeg + min(IPL == 1 ? lane : 2 * (lane & 31) + k, np - 1)];

    // group box (code/computeForces.cu:539-555)
    R lmnx = pi[0].x, lmny = pi[0].y, lmnz = pi[0].z;
    R lmxx = pi[0].x, lmxy = pi[0].y, lmxz = pi[0].z;
#pragma unroll
    for (int k = 1; k < IPL; k++) {
      lmnx = rmin(lmnx, pi[k].x); lmny = rmin(lmny, pi[k].y); lmnz = rmin(lmnz, pi[k].z);
      lmxx = rmax(lmxx, pi[k].x); lmxy = rmax(lmxy, pi[k].y); lmxz = rmax(lmxz, pi[k].z);
    }
    const R mnx = wave_min_r(lmnx), mny = wave_min_r(lmny), mnz = wave_min_r(lmnz);
    const R mxx = wave_max_r(lmxx), mxy = wave_max_r(lmxy), mxz = wave_max_r(lmxz);
    R gcx, gcy, gcz, gsx, gsy, gsz;
    {
#pragma clang fp contract(off)
      gcx = R(0.5) * (mxx + mnx); gcy = R(0.5) * (mxy + mny); gcz = R(0.5) * (mxz + mnz);
      gsx = R(0.5) * (mxx - mnx); gsy = R(0.5) * (mxy - mny); gsz = R(0.5) * (mxz - mnz);
    }
    GroupGeom geom;
    if constexpr (LEARN) geom = make_group_geom(gcx, gcy, gcz, gsx, gsy, gsz, np, a.sys_scale[sys]);
    // i-particles: relative to the group centre (LDS path) or absolute (scalar path)
    IParts<R, IPL> ip;
#pragma unroll
    for (int k = 0; k < IPL; k++) {
      ip.x[k] = SJ ? pi[k].x : pi[k].x - gcx;
      ip.y[k] = SJ ? pi[k].y : pi[k].y - gcy;
      ip.z[k] = SJ ? pi[k].z : pi[k].z - gcz;
      ip.s[k] = Acc4T<R>{R(0), R(0), R(0), R(0)};
      ip.c[k] = Acc4T<R>{R(0), R(0), R(0), R(0)};
    }

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
      V4<R> cs = make_v4<R>(R(0), R(0), R(0), R(0));
      if (use) {
        ci = LSTK ? lstk[top + lane] : stk[top + lane];
        cd = a.cells[ci];
        cs = a.csize[ci];
      }
      bool split = false, leaf = false;
      if constexpr (LEARN) {
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
      } else {
        if (use) {
          split = open_cell<R>(cs, gcx, gcy, gcz, gsx, gsy, gsz) || ((int)cd.w - (int)cd.z < 3);
          leaf = cell_leaf(cd);
        }
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
            batch_p2p<KAHAN, SJ, R, IPL>(a, S, kWave, gcx, gcy, gcz, ip);
            n_dir += kWave;
            nd -= kWave;
            if (lane < nd) S.dlist[lane] = S.dlist[kWave + lane];
            wave_lds_sync();
          }
        }
      }
      wave_lds_sync();
      if (na >= kWave) {
        batch_p2c<KAHAN, SJ, MONO, R, IPL>(a, S, kWave, gcx, gcy, gcz, ip);
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
      batch_p2c<KAHAN, SJ, MONO, R, IPL>(a, S, na, gcx, gcy, gcz, ip);
      n_app += na;
    }
    if (nd > 0) {
      batch_p2p<KAHAN, SJ, R, IPL>(a, S, nd, gcx, gcy, gcz, ip);
      n_dir += nd;
    }
    // this lane stores particle `idx`; lanes past the end of the group write
    // to the kWave spare slots after acc[n)
    Acc4T<R> res;
    R mi;
    int idx;
    if (IPL == 1) {
      res = ip.s[0];
      mi = pi[0].w;
      idx = lane;
    } else {
      // each half-wave summed half of every batch: add the other half's sums
      // (compensated values s - c) to get the particles' totals
      const int h = lane >> 5;
      Acc4T<R> t[IPL];
#pragma unroll
      for (int k = 0; k < IPL; k++) {
        Acc4T<R> v = ip.s[k];
        if (KAHAN) {
          v.x -= ip.c[k].x; v.y -= ip.c[k].y; v.z -= ip.c[k].z; v.w -= ip.c[k].w;
        }
        t[k].x = v.x + __shfl_xor(v.x, 32, kWave);
        t[k].y = v.y + __shfl_xor(v.y, 32, kWave);
        t[k].z = v.z + __shfl_xor(v.z, 32, kWave);
        t[k].w = v.w + __shfl_xor(v.w, 32, kWave);
      }
      res = h ? t[IPL - 1] : t[0];
      mi = h ? pi[IPL - 1].w : pi[0].w;
      idx = 2 * (lane & 31) + h;
    }
    const bool valid = idx < np;
    a.acc[valid ? pbeg + idx : a.n + lane] = make_v4<R>(res.x, res.y, res.z, res.w);
    const double gp = wave_sum(valid ? (double)mi * (double)res.w : 0.0);
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
    atomicAnd(&dc->err, ~4);
  }
}

}  // namespace

namespace {

template <typename R>
WalkArgs<R> walk_args(Ctx &c, const int *run_if) {
  WalkArgs<R> a;
  a.n = c.n;
  a.nsys = c.nsys;
  a.eps2 = (R)c.eps * (R)c.eps;
  a.groups = c.groups;
  a.group_sys = c.group_sys;
  a.root_cells = c.root_cells;
  a.pos_ph = as4<R>(c.pos_ph);
  a.pos_tree = as4<R>(c.pos);
  a.cells = c.cells;
  a.csize = as4<R>(c.csize);
  a.cmono = as4<R>(c.cmono);
  a.cq0 = as4<R>(c.cq0);
  a.cq1 = as2<R>(c.cq1);
  a.cwalk = as4<R>(c.cwalk);
  a.cls = c.cls;
  a.sys_scale = c.sys_scale;
  a.mlp = c.mlp_dev;
  a.shift = c.learn_shift;
  a.it_min = 1.0f / c.learn_theta_hi;
  a.it_max = 1.0f / c.learn_theta_lo;
  a.acc = as4<R>(c.acc);
  a.gcounts = c.gcounts;
  a.gpot = c.gpot;
  a.stack = c.stack;
  a.stack_ints = c.stack_ints;
  a.lstk_cap = c.lds_stack;
  a.run_if = run_if;
  a.dc = c.dc;
  return a;
}

}  // namespace

void launch_walk(Ctx &c, const int *run_if) {
  c.h_wstats_valid = false;   // d_wstats changes with this walk
  hipLaunchKernelGGL(k_walk_reset, dim3(1), dim3(64), 0, c.stream, c.dc, run_if);
  if (c.fp64) {
    // Treecode<double>: one i-particle per lane, LDS j-data, FP64 sums
    const WalkArgs<double> a = walk_args<double>(c, run_if);
    void (*kern)(WalkArgs<double>) = c.quadrupole ? k_walk<double, false, false, false, false, 1>
                                                  : k_walk<double, false, false, false, true, 1>;
    hipLaunchKernelGGL(kern, dim3(c.walk_blocks), dim3(kBlock), 0, c.stream, a);
  } else {
    // variant 0: LDS j-data, one i-particle per lane; 1: scalar-load j-data,
    // one i-particle per lane; 2: two i-particles per lane, batches split
    // between the half-waves (LDS j-data)
    const WalkArgs<float> a = walk_args<float>(c, run_if);
    void (*kern)(WalkArgs<float>);
    const bool learn = c.opening == 1 && c.have_mlp;
    const int v = c.walk_variant;
#define BT_PICK(K, LEARN, MONO)                                                     \
  (v == 1 ? k_walk<float, K, true, LEARN, MONO, 1>                                  \
          : v == 2 ? k_walk<float, K, false, LEARN, MONO, 2> : k_walk<float, K, false, LEARN, MONO, 1>)
    if (!c.quadrupole)
      kern = c.kahan ? (v == 2 ? k_walk<float, true, false, false, true, 2>
                               : k_walk<float, true, false, false, true, 1>)
                     : (v == 2 ? k_walk<float, false, false, false, true, 2>
                               : k_walk<float, false, false, false, true, 1>);
    else if (learn)
      kern = c.kahan ? BT_PICK(true, true, false) : BT_PICK(false, true, false);
    else
      kern = c.kahan ? BT_PICK(true, false, false) : BT_PICK(false, false, false);
#undef BT_PICK
    hipLaunchKernelGGL(kern, dim3(c.walk_blocks), dim3(kBlock), 0, c.stream, a);
  }
  WStat *part = reinterpret_cast<WStat *>(c.d_wstats + 4);
  hipLaunchKernelGGL(k_walk_stats_part, dim3(kWalkStatBlocks), dim3(kBlock), 0, c.stream, c.groups,
                     c.gcounts, c.gpot, c.dc, part, run_if);
  hipLaunchKernelGGL(k_walk_stats_final, dim3(1), dim3(kBlock), 0, c.stream, part, c.d_wstats,
                     c.d_wpot, run_if);
}

}  // namespace bt
