// k_learn.hip -- the learned opening angle: network preparation, batched
// inference, training-pair generation, the pair-error measurement, the
// gradient / Adam kernels that train the network on the GPU, and the
// per-step FP64 check of a learned walk.  No reference counterpart (the
// reference's walk uses one fixed angle, code/computeForces.cu:30-54).
#include "bt_device.h"
#include "bt_internal.h"
#include "bt_learn.h"

namespace bt {

namespace {

constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr int kMaxSteps = 1 << 20;
constexpr int kGradBlocks = 128;   // partial-gradient rows (fixed: deterministic sums)
constexpr int kRow = 139;          // per-sample LDS row of the gradient kernel (odd stride)
// row layout: xh[8] a1h[32] a2[32] dz1[32] dz2[32] g3 one
constexpr int kRowXh = 0, kRowA1 = 8, kRowA2 = 40, kRowDz1 = 72, kRowDz2 = 104, kRowG3 = 136,
              kRowOne = 137;

__device__ __forceinline__ double wave_max_d(double v) {
  for (int m = 32; m >= 1; m >>= 1) v = fmax(v, __shfl_xor(v, m, kWave));
  return v;
}

__device__ __forceinline__ float f16r(float x) { return (float)(_Float16)x; }

__device__ __forceinline__ float silu_acc(float z) { return z / (1.0f + expf(-z)); }
__device__ __forceinline__ float dsilu_acc(float z) {
  const float sg = 1.0f / (1.0f + expf(-z));
  return sg * (1.0f + z * (1.0f - sg));
}

// ---- network preparation ---------------------------------------------------

// Builds the per-lane operands (see MlpDev) from FP32 parameters, rounding
// every parameter to FP16.  One wave.
__global__ void k_mlp_prepare(const float *__restrict__ p, MlpDev *out) {
  const int l = threadIdx.x;
  if (l >= kWave) return;
  const int i = l & 31, hh = l >> 5;
  half8_t a1;
  for (int t = 0; t < 8; t++) {
    float v;
    if (hh == 0) v = p[kOffW1 + i * kMlpIn + t];
    else v = (t == 0) ? p[kOffB1 + i] : 0.0f;
    a1[t] = (_Float16)v;
  }
  out->a1[l] = a1;
  for (int s = 0; s < 2; s++) {
    half8_t a2;
    for (int t = 0; t < 8; t++) a2[t] = (_Float16)p[kOffW2 + i * kMlpH + mlp_unit(s, t, hh)];
    out->a2[s][l] = a2;
  }
  if (l < 32) {
    const int h2 = l >> 4, r = l & 15;
    const int unit = 8 * (r / 4) + 4 * h2 + (r % 4);
    out->c2p[h2][r] = f16r(p[kOffB2 + unit]);
    out->w3p[h2][r] = f16r(p[kOffW3 + unit]);
  }
  if (l == 0) out->b3 = f16r(p[kOffB3]);
}

// Batched inference y[k] = MLP(x[8k..]) through the same matrix-core routine
// as the walk.  Grid-stride over batches of 64 samples per wave.
__global__ __launch_bounds__(kBlock) void k_mlp_forward(const MlpDev *__restrict__ M, int n,
                                                        const float *__restrict__ x,
                                                        float *__restrict__ y) {
  __shared__ float s_mlp[64];
  if (threadIdx.x < 32) s_mlp[threadIdx.x] = M->c2p[threadIdx.x >> 4][threadIdx.x & 15];
  else if (threadIdx.x < 64) s_mlp[threadIdx.x] = M->w3p[(threadIdx.x - 32) >> 4][threadIdx.x & 15];
  __syncthreads();
  const int lane = lane_id();
  const half8_t a1 = M->a1[lane], a2_0 = M->a2[0][lane], a2_1 = M->a2[1][lane];
  const float b3 = M->b3;
  const int nw = gridDim.x * kWavesPerBlock;
  for (int base = (blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6)) * kWave; base < n;
       base += nw * kWave) {
    const int k = base + lane;
    float f[8];
    for (int t = 0; t < 8; t++) f[t] = k < n ? x[8 * (size_t)k + t] : 0.0f;
    const float yk = mlp_wave(a1, a2_0, a2_1, s_mlp, b3, pack_features(f), true, true);
    if (k < n) y[k] = yk;
  }
}

// ---- system scales -----------------------------------------------------------

// (R, M_sys) per system: the root box half size and the total mass of its
// level-1 cells.
__global__ void k_sys_scale(int nsys, const float4 *__restrict__ box,
                            const int2 *__restrict__ root_cells,
                            const float4 *__restrict__ cmono, float2 *__restrict__ out) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nsys) return;
  const int2 rc = root_cells[s];
  double m = 0.0;
  for (int k = 0; k < rc.y; k++) m += (double)cmono[rc.x + k].w;
  out[s] = make_float2(box[s].w, m > 0.0 ? (float)m : 1.0f);
}

// ---- training pairs ------------------------------------------------------------

struct PairArgs {
  int nsel;
  const int *gsel;
  const int2 *groups;
  const int *group_sys;
  int nsys;
  const int2 *root_cells;
  const float4 *pos_ph;
  const float4 *pos_tree;
  const uint4 *cells;
  const float4 *csize;
  const float4 *cmono;
  const float4 *cq0;
  const float2 *cq1;
  const float2 *cls;
  const float2 *sys_scale;
  float theta_lo, theta_hi, it_lo, eps2;
  int *stack;
  int stack_ints;
  int *count;        // pairs found (may exceed max_pairs)
  int max_pairs;
  int2 *pairs;       // (group, cell)
  double *tneed;     // theta_need
  const double *amag;
  float *feat, *target, *err, *tneed_f;
  float *sink;       // 64 floats per wave: stores of idle lanes
  DevCounters *dc;
};

__device__ __forceinline__ void group_box(const PairArgs &a, int g, int &pbeg, int &np,
                                          float4 &pi, float &gcx, float &gcy, float &gcz,
                                          float &gsx, float &gsy, float &gsz) {
  const int lane = lane_id();
  const int2 grp = a.groups[g];
  pbeg = grp.x;
  np = max(grp.y, 1);
  pi = a.pos_ph[pbeg + min(lane, np - 1)];
  const float mnx = wave_min(pi.x), mny = wave_min(pi.y), mnz = wave_min(pi.z);
  const float mxx = wave_max(pi.x), mxy = wave_max(pi.y), mxz = wave_max(pi.z);
  {
#pragma clang fp contract(off)
    gcx = 0.5f * (mxx + mnx); gcy = 0.5f * (mxy + mny); gcz = 0.5f * (mxz + mnz);
    gsx = 0.5f * (mxx - mnx); gsy = 0.5f * (mxy - mny); gsz = 0.5f * (mxz - mnz);
  }
}

// Box distance squared from the group box to the cell's centre of mass, the
// float operations of the walk's opening test.
__device__ __forceinline__ float box_ds2(float4 cs, float gcx, float gcy, float gcz, float gsx,
                                         float gsy, float gsz) {
#pragma clang fp contract(off)
  float dx = fabsf(gcx - cs.x) - gsx;
  float dy = fabsf(gcy - cs.y) - gsy;
  float dz = fabsf(gcz - cs.z) - gsz;
  dx += fabsf(dx); dx *= 0.5f;
  dy += fabsf(dy); dy *= 0.5f;
  dz += fabsf(dz); dz *= 0.5f;
  return dx * dx + dy * dy + dz * dz;
}

// One wave per selected group: the walk's traversal at theta_lo; every
// visited cell with >= 3 particles, ds > s and theta_need = l/(ds - s) in
// [theta_lo, theta_hi] becomes a pair.
__global__ __launch_bounds__(kBlock) void k_learn_pairs(PairArgs a) {
  const int lane = lane_id();
  const int wv = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
  const int nw = gridDim.x * kWavesPerBlock;
  int *stk = a.stack + (size_t)wv * (size_t)a.stack_ints;
  for (int si = wv; si < a.nsel; si += nw) {
    const int g = a.gsel[si];
    int pbeg, np;
    float4 pi;
    float gcx, gcy, gcz, gsx, gsy, gsz;
    group_box(a, g, pbeg, np, pi, gcx, gcy, gcz, gsx, gsy, gsz);
    const int sys = a.nsys > 1 ? a.group_sys[g] : 0;
    const int2 rc = a.root_cells[sys];
    if (lane < rc.y) stk[lane] = rc.x + lane;
    int top = rc.y;
    int steps = 0;
    __builtin_amdgcn_s_waitcnt(0);
    while (top > 0) {
      // each cell is visited at most once per group: a longer traversal means
      // a corrupt tree
      if (++steps > kMaxSteps) {
        atomicOr(&a.dc->err, 4);
        break;
      }
      const int take = min(kWave, top);
      top -= take;
      const bool use = lane < take;
      int ci = 0;
      uint4 cd = make_uint4(0, 0, 0, 0);
      float4 cs = make_float4(0.f, 0.f, 0.f, 0.f);
      float2 ls = make_float2(0.f, 0.f);
      if (use) {
        ci = stk[top + lane];
        cd = a.cells[ci];
        cs = a.csize[ci];
        ls = a.cls[ci];
      }
      const float ds2 = box_ds2(cs, gcx, gcy, gcz, gsx, gsy, gsz);
      const int ncp = (int)cd.w - (int)cd.z;
      const double ds = sqrt((double)ds2);
      const double tn = (ds > (double)ls.y) ? (double)ls.x / (ds - (double)ls.y) : 1e30;
      const bool inband = use && ncp >= 3 && ds > (double)ls.y && tn >= (double)a.theta_lo &&
                          tn <= (double)a.theta_hi;
      const uint64_t bm = ballot64(inband);
      const int nb = __popcll(bm);
      const int base = wave_uniform(atomicAdd(a.count, lane == 0 ? nb : 0));
      if (inband) {
        const int k = base + __popcll(bm & lanemask_lt());
        if (k < a.max_pairs) {
          a.pairs[k] = make_int2(g, ci);
          a.tneed[k] = tn;
        }
      }
      float op;
      {
#pragma clang fp contract(off)
        op = ls.x * a.it_lo + ls.y;
      }
      const bool split = use && (ds2 < op * op || ncp < 3);
      const bool push = split && !cell_leaf(cd);
      const int nch = push ? cell_nchild(cd) : 0;
      const int chi = wave_incl_scan(nch);
      const int ctot = __shfl(chi, kWave - 1, kWave);
      if (top + ctot > a.stack_ints) {
        atomicOr(&a.dc->err, 4);
        break;
      }
      if (push) {
        const int f = cell_first(cd);
        const int b = top + chi - nch;
        for (int k = 0; k < nch; k++) stk[b + k] = f + k;
      }
      top += ctot;
      __builtin_amdgcn_s_waitcnt(0);
    }
  }
}

// One wave per pair: features, FP64 direct force of the cell's particles on
// the group's particles against the FP64 quadrupole P2C, the pair error
// e = max_i |a_P2C - a_direct| / amag[i] and the target
// log2(theta_need) - log2(max(e, 1e-12)) / 3.
__global__ __launch_bounds__(kBlock) void k_learn_errors(PairArgs a, int npairs) {
  const int lane = lane_id();
  const int wv = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
  const int nw = gridDim.x * kWavesPerBlock;
  for (int p = wv; p < npairs; p += nw) {
    const int2 gc = a.pairs[p];
    const int g = gc.x, c = gc.y;
    int pbeg, np;
    float4 pi;
    float gcx, gcy, gcz, gsx, gsy, gsz;
    group_box(a, g, pbeg, np, pi, gcx, gcy, gcz, gsx, gsy, gsz);
    const bool valid = lane < np;
    const int sys = a.nsys > 1 ? a.group_sys[g] : 0;
    const GroupGeom geom = make_group_geom(gcx, gcy, gcz, gsx, gsy, gsz, np, a.sys_scale[sys]);
    const float4 m = a.cmono[c];
    const float4 q0 = a.cq0[c];
    const float2 q1 = a.cq1[c];
    float x[8];
    pair_features(geom, m, a.cls[c].x, q0, q1, x);
    const uint4 cd = a.cells[c];
    const double px = pi.x, py = pi.y, pz = pi.z;
    double ax = 0.0, ay = 0.0, az = 0.0;
    for (int j = (int)cd.z; j < (int)cd.w; j++) {
      const float4 pj = a.pos_tree[j];
      const double rx = (double)pj.x - px, ry = (double)pj.y - py, rz = (double)pj.z - pz;
      const double r2 = rx * rx + ry * ry + rz * rz + (double)a.eps2;
      const double ri = 1.0 / sqrt(r2);
      const double w = (double)pj.w * ri * ri * ri;
      ax += w * rx; ay += w * ry; az += w * rz;
    }
    const double dx = px - m.x, dy = py - m.y, dz = pz - m.z;
    const double r2 = dx * dx + dy * dy + dz * dz + (double)a.eps2;
    const double ri = 1.0 / sqrt(r2), ri2 = ri * ri;
    const double m1 = (double)m.w * ri, m3 = m1 * ri2, m5 = m3 * ri2, m7 = m5 * ri2;
    const double Qxx = q0.x, Qyy = q0.y, Qzz = q0.z, Qxy = q0.w, Qxz = q1.x, Qyz = q1.y;
    const double tq = Qxx + Qyy + Qzz;
    const double Qx = Qxx * dx + Qxy * dy + Qxz * dz;
    const double Qy = Qxy * dx + Qyy * dy + Qyz * dz;
    const double Qz = Qxz * dx + Qyz * dy + Qzz * dz;
    const double qrr = Qx * dx + Qy * dy + Qz * dz;
    const double C = -m3 + 1.5 * tq * m5 - 7.5 * qrr * m7;
    const double ex = C * dx + 3.0 * m5 * Qx - ax, ey = C * dy + 3.0 * m5 * Qy - ay,
                 ez = C * dz + 3.0 * m5 * Qz - az;
    const double ei = valid ? sqrt(ex * ex + ey * ey + ez * ez) / a.amag[pbeg + lane] : 0.0;
    const double e = wave_max_d(ei);
    const double tn = a.tneed[p];
    const float tgt = (float)(log2(tn) - log2(fmax(e, 1e-12)) / 3.0);
    // branch-free stores: lanes 0..7 features, 8 target, 9 error, 10 theta_need
    float v = 0.f;
    for (int k = 0; k < 8; k++) v = lane == k ? x[k] : v;
    v = lane == 8 ? tgt : (lane == 9 ? (float)e : (lane == 10 ? (float)tn : v));
    float *dst = lane < 8   ? a.feat + 8 * (size_t)p + lane
                 : lane == 8 ? a.target + p
                 : lane == 9 ? a.err + p
                 : lane == 10 ? a.tneed_f + p
                              : a.sink + (size_t)wv * kWave + lane;
    *dst = v;
  }
}

// amag[idx[k]] = |a| of the FP64 direct sum of target k.
__global__ void k_learn_amag(int ns, const int *__restrict__ idx, const double4 *__restrict__ d,
                             double *__restrict__ amag) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= ns) return;
  const double4 v = d[k];
  amag[idx[k]] = sqrt(v.x * v.x + v.y * v.y + v.z * v.z);
}

// ---- training ------------------------------------------------------------------

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Partial gradient of L = 1/B sum (y - t)^2 over a minibatch of B samples:
// sample k of the batch is k itself when B == n, else a hashed draw from
// [0, n).  Block b writes row b of partial[kGradBlocks][kMlpParams + 1]
// (the last entry is the block's sum of squared residuals).  The forward
// pass rounds parameters, inputs and the first hidden layer to FP16 as the
// inference routine does; the backward pass treats the rounding as identity.
__global__ __launch_bounds__(64) void k_mlp_grad(const float *__restrict__ params, int n,
                                                 const float *__restrict__ feat,
                                                 const float *__restrict__ target, int batch,
                                                 unsigned long long seed, int step,
                                                 float *__restrict__ partial) {
  __shared__ float pr[kMlpParams];
  __shared__ float row[64 * kRow];
  const int t = threadIdx.x;
  for (int i = t; i < kMlpParams; i += 64) pr[i] = f16r(params[i]);
  // owned gradient entries: i = t + 64 q
  int ou[22], ov[22];
  float acc[22];
  int nown = 0;
  for (int i = t; i < kMlpParams; i += 64, nown++) {
    int u, v;
    if (i < kOffB1) { u = kRowDz1 + i / kMlpIn; v = kRowXh + i % kMlpIn; }
    else if (i < kOffW2) { u = kRowDz1 + (i - kOffB1); v = kRowOne; }
    else if (i < kOffB2) { u = kRowDz2 + (i - kOffW2) / kMlpH; v = kRowA1 + (i - kOffW2) % kMlpH; }
    else if (i < kOffW3) { u = kRowDz2 + (i - kOffB2); v = kRowOne; }
    else if (i < kOffB3) { u = kRowG3; v = kRowA2 + (i - kOffW3); }
    else { u = kRowG3; v = kRowOne; }
    ou[nown] = u;
    ov[nown] = v;
    acc[nown] = 0.f;
  }
  __syncthreads();
  float sq = 0.f;
  const float inv_b = 1.0f / (float)batch;
  for (int k0 = blockIdx.x * 64; k0 < batch; k0 += gridDim.x * 64) {
    const int k = k0 + t;
    float *r = row + t * kRow;
    const bool live = k < batch;
    int s = 0;
    if (live) {
      s = batch == n ? k
                     : (int)(mix64(seed ^ ((unsigned long long)step << 32) ^ (unsigned long long)k) %
                             (unsigned long long)n);
    }
    float xh[kMlpIn];
    for (int f = 0; f < kMlpIn; f++) {
      xh[f] = live ? f16r(feat[8 * (size_t)s + f]) : 0.f;
      r[kRowXh + f] = xh[f];
    }
    float z1[kMlpH];
    for (int h = 0; h < kMlpH; h++) {
      float z = pr[kOffB1 + h];
      for (int f = 0; f < kMlpIn; f++) z = fmaf(pr[kOffW1 + h * kMlpIn + f], xh[f], z);
      z1[h] = z;
      r[kRowA1 + h] = f16r(silu_acc(z));
    }
    float y = pr[kOffB3];
    for (int h = 0; h < kMlpH; h++) {
      float z = pr[kOffB2 + h];
      for (int j = 0; j < kMlpH; j++) z = fmaf(pr[kOffW2 + h * kMlpH + j], r[kRowA1 + j], z);
      const float a2 = silu_acc(z);
      r[kRowA2 + h] = a2;
      r[kRowDz2 + h] = dsilu_acc(z);  // holds silu'(z2) until g3 is known
      y = fmaf(pr[kOffW3 + h], a2, y);
    }
    const float res = live ? y - target[s] : 0.f;
    sq += res * res;
    const float g3 = 2.0f * res * inv_b;
    r[kRowG3] = g3;
    r[kRowOne] = 1.0f;
    for (int h = 0; h < kMlpH; h++) r[kRowDz2 + h] *= g3 * pr[kOffW3 + h];
    for (int j = 0; j < kMlpH; j++) {
      float da = 0.f;
      for (int h = 0; h < kMlpH; h++) da = fmaf(r[kRowDz2 + h], pr[kOffW2 + h * kMlpH + j], da);
      r[kRowDz1 + j] = da * dsilu_acc(z1[j]);
    }
    __syncthreads();
    for (int q = 0; q < nown; q++) {
      float sacc = 0.f;
      for (int s2 = 0; s2 < 64; s2++) sacc = fmaf(row[s2 * kRow + ou[q]], row[s2 * kRow + ov[q]], sacc);
      acc[q] += sacc;
    }
    __syncthreads();
  }
  float *out = partial + (size_t)blockIdx.x * (kMlpParams + 1);
  for (int q = 0; q < nown; q++) out[t + 64 * q] = acc[q];
  const float sqs = wave_sum(sq);
  if (t == 0) out[kMlpParams] = sqs;
}

// grad[i] = sum over blocks (fixed order) of the partial rows, in FP64.
__global__ void k_mlp_grad_reduce(const float *__restrict__ partial, int nblocks,
                                  double *__restrict__ grad, int batch) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i > kMlpParams) return;
  double s = 0.0;
  for (int b = 0; b < nblocks; b++) s += (double)partial[(size_t)b * (kMlpParams + 1) + i];
  grad[i] = i == kMlpParams ? s / batch : s;   // last entry: the loss
}

__global__ void k_mlp_adam(float *__restrict__ p, float *__restrict__ m, float *__restrict__ v,
                           const double *__restrict__ grad, float lr, float b1, float b2,
                           float eps, float bc1, float bc2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= kMlpParams) return;
  const float g = (float)grad[i];
  const float mi = b1 * m[i] + (1.0f - b1) * g;
  const float vi = b2 * v[i] + (1.0f - b2) * g * g;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

// ---- per-step check of the learned walk -------------------------------------------

// Targets of a check (one block of 1024 threads, ns <= 1024): draw k is
// idx[k] = mix64(seed ^ (calls << 32) ^ k) mod n, an index into the particle
// arrays as they are when the draw runs (calls = the checks drawn before).
// Their input ids, sorted with the draw slot as tie-break, go to
// idx[ns + j] and the slots to idx[2 ns + j], for k_check_map.
__global__ __launch_bounds__(1024) void k_check_idx(int ns, int n, unsigned long long seed,
                                                    DevCounters *dc, const int *__restrict__ id,
                                                    int *__restrict__ idx) {
  __shared__ int s_id[1024];
  __shared__ int s_slot[1024];
  const unsigned long long calls = (unsigned long long)(unsigned)dc->check_calls;
  const int t = threadIdx.x;
  int key = 0x7FFFFFFF;
  if (t < ns) {
    const int k =
        (int)(mix64(seed ^ (calls << 32) ^ (unsigned long long)t) % (unsigned long long)n);
    idx[t] = k;
    key = id[k];
  }
  s_id[t] = key;
  s_slot[t] = t;
  __syncthreads();
  // bitonic sort of (id, slot) pairs, ascending; the slots make keys distinct
  for (int k = 2; k <= 1024; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      const int p = t ^ j;
      if (p > t) {
        const int ai = s_id[t], bi = s_id[p], as = s_slot[t], bs = s_slot[p];
        const bool gt = ai > bi || (ai == bi && as > bs);
        if (gt == ((t & k) == 0)) {
          s_id[t] = bi; s_id[p] = ai;
          s_slot[t] = bs; s_slot[p] = as;
        }
      }
      __syncthreads();
    }
  }
  if (t < ns) {
    idx[ns + t] = s_id[t];
    idx[2 * ns + t] = s_slot[t];
  }
  if (t == 0) dc->check_calls = (int)(calls + 1);
}

// PH index of each check target: particle p (PH order) looks its input id up
// among the sorted target ids (binary search in LDS) and writes p into
// idx[3 ns + slot] for every draw of that id.  ids are a permutation of
// [0, n), so every slot is written exactly once.
__global__ __launch_bounds__(kBlock) void k_check_map(int n, int ns, const int *__restrict__ id_ph,
                                                      int *__restrict__ idx) {
  __shared__ int s_id[1024];
  for (int k = threadIdx.x; k < ns; k += kBlock) s_id[k] = idx[ns + k];
  __syncthreads();
  for (int p = blockIdx.x * kBlock + threadIdx.x; p < n; p += gridDim.x * kBlock) {
    const int i = id_ph[p];
    int lo = 0, hi = ns;
    while (lo < hi) {
      const int m = (lo + hi) >> 1;
      if (s_id[m] < i) lo = m + 1;
      else hi = m;
    }
    for (int q = lo; q < ns && s_id[q] == i; q++) idx[3 * ns + idx[2 * ns + q]] = p;
  }
}

// Relative force errors of the walk against the FP64 direct sum on ns
// samples (d_idx: their PH indices), then (one block) a sort: out = {p50,
// p99, max}.
__global__ __launch_bounds__(1024) void k_learn_check(int ns, const int *__restrict__ idx,
                                                      const double4 *__restrict__ d,
                                                      const float4 *__restrict__ acc,
                                                      double p99_max, DevCounters *dc,
                                                      double *__restrict__ out) {
  __shared__ float e[1024];
  const int t = threadIdx.x;
  float v = 3.4e38f;
  if (t < ns) {
    const double4 r = d[t];
    const float4 w = acc[idx[t]];
    const double ex = (double)w.x - r.x, ey = (double)w.y - r.y, ez = (double)w.z - r.z;
    v = (float)(sqrt(ex * ex + ey * ey + ez * ez) / sqrt(r.x * r.x + r.y * r.y + r.z * r.z));
  }
  e[t] = v;
  __syncthreads();
  // bitonic sort of 1024 floats (ascending)
  for (int k = 2; k <= 1024; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      const int p = t ^ j;
      if (p > t) {
        const float a = e[t], b = e[p];
        const bool up = (t & k) == 0;
        if ((a > b) == up) { e[t] = b; e[p] = a; }
      }
      __syncthreads();
    }
  }
  if (t == 0 && ns > 0) {
    const double p99 = e[(int)(0.99 * (ns - 1))];
    out[0] = e[(ns - 1) / 2];
    out[1] = p99;
    out[2] = e[ns - 1];
    // above the threshold the fixed-angle walk recomputes the forces
    const int fb = p99 > p99_max ? 1 : 0;
    dc->fallback = fb;
    dc->checks += 1;
    dc->fallbacks += fb;
  }
}

}  // namespace

// ---- launchers -------------------------------------------------------------------

void launch_mlp_prepare(Ctx &c) {
  hipLaunchKernelGGL(k_mlp_prepare, dim3(1), dim3(64), 0, c.stream, c.mlp_params, c.mlp_dev);
}

void launch_mlp_forward(Ctx &c, int n, const float *x, float *y) {
  const int waves = (n + kWave - 1) / kWave;
  const int blocks = std::max(1, std::min((waves + kWavesPerBlock - 1) / kWavesPerBlock, 4096));
  hipLaunchKernelGGL(k_mlp_forward, dim3(blocks), dim3(kBlock), 0, c.stream, c.mlp_dev, n, x, y);
}

void launch_sys_scale(Ctx &c) {
  hipLaunchKernelGGL(k_sys_scale, dim3((c.nsys + 63) / 64), dim3(64), 0, c.stream, c.nsys, c.box,
                     c.root_cells, c.cmono, c.sys_scale);
}

static PairArgs pair_args(Ctx &c, LearnBuffers &L) {
  PairArgs a;
  a.nsel = L.nsel;
  a.gsel = L.gsel;
  a.groups = c.groups;
  a.group_sys = c.group_sys;
  a.nsys = c.nsys;
  a.root_cells = c.root_cells;
  a.pos_ph = c.pos_ph;
  a.pos_tree = c.pos;
  a.cells = c.cells;
  a.csize = c.csize;
  a.cmono = c.cmono;
  a.cq0 = c.cq0;
  a.cq1 = c.cq1;
  a.cls = c.cls;
  a.sys_scale = c.sys_scale;
  a.theta_lo = L.theta_lo;
  a.theta_hi = L.theta_hi;
  a.it_lo = 1.0f / L.theta_lo;
  a.eps2 = c.eps * c.eps;
  a.stack = c.stack;
  a.stack_ints = c.stack_ints;
  a.count = L.count;
  a.max_pairs = L.max_pairs;
  a.pairs = L.pairs;
  a.tneed = L.tneed;
  a.amag = L.amag;
  a.feat = L.feat;
  a.target = L.target;
  a.err = L.err;
  a.tneed_f = L.tneed_f;
  a.sink = L.sink;
  a.dc = c.dc;
  return a;
}


void launch_learn_pairs(Ctx &c, LearnBuffers &L) {
  const PairArgs a = pair_args(c, L);
  hipLaunchKernelGGL(k_learn_pairs, dim3(c.walk_blocks), dim3(kBlock), 0, c.stream, a);
}

void launch_learn_errors(Ctx &c, LearnBuffers &L, int npairs) {
  const PairArgs a = pair_args(c, L);
  hipLaunchKernelGGL(k_learn_errors, dim3(c.walk_blocks), dim3(kBlock), 0, c.stream, a, npairs);
}

void launch_learn_amag(Ctx &c, int ns, const int *d_idx, const double *d_direct, double *amag) {
  hipLaunchKernelGGL(k_learn_amag, dim3((ns + 255) / 256), dim3(256), 0, c.stream, ns, d_idx,
                     reinterpret_cast<const double4 *>(d_direct), amag);
}

int mlp_grad_blocks() { return kGradBlocks; }

void launch_mlp_grad(Ctx &c, const float *params, int n, const float *feat, const float *target,
                     int batch, unsigned long long seed, int step, float *partial,
                     double *grad) {
  hipLaunchKernelGGL(k_mlp_grad, dim3(kGradBlocks), dim3(64), 0, c.stream, params, n, feat, target,
                     batch, seed, step, partial);
  hipLaunchKernelGGL(k_mlp_grad_reduce, dim3((kMlpParams + 1 + 255) / 256), dim3(256), 0,
                     c.stream, partial, kGradBlocks, grad, batch);
}

void launch_mlp_adam(Ctx &c, float *p, float *m, float *v, const double *grad, float lr, int step) {
  const float b1 = 0.9f, b2 = 0.999f;
  const float bc1 = 1.0f - powf(b1, (float)step), bc2 = 1.0f - powf(b2, (float)step);
  hipLaunchKernelGGL(k_mlp_adam, dim3((kMlpParams + 255) / 256), dim3(256), 0, c.stream, p, m, v,
                     grad, lr, b1, b2, 1e-8f, bc1, bc2);
}

void launch_check_idx(Ctx &c, int ns) {
  hipLaunchKernelGGL(k_check_idx, dim3(1), dim3(1024), 0, c.stream, ns, c.n,
                     (unsigned long long)c.check_seed, c.dc, c.id, c.d_check_idx);
}

void launch_check_map(Ctx &c, int ns) {
  hipLaunchKernelGGL(k_check_map, dim3(grid_for(c.n, kBlock, 2048)), dim3(kBlock), 0, c.stream,
                     c.n, ns, c.id_ph, c.d_check_idx);
}

void launch_learn_check(Ctx &c, int ns, const int *d_idx, const double *d_direct, double *out) {
  hipLaunchKernelGGL(k_learn_check, dim3(1), dim3(1024), 0, c.stream, ns, d_idx,
                     reinterpret_cast<const double4 *>(d_direct), c.acc, (double)c.check_p99,
                     c.dc, out);
}

}  // namespace bt
