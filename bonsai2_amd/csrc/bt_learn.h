// bt_learn.h -- the learned opening angle: pair features and the 8-32-32-1
// network evaluated by a wave64 with FP16 matrix-core products.
//
// No reference counterpart: the reference opens cells with one fixed angle
// (code/computeForces.cu:30-54).  Here a network picks the angle of each
// (group, candidate cell) pair: theta = exp2(y + shift) clamped to
// [theta_lo, theta_hi], and the cell is opened iff ds^2 < (l/theta + s)^2
// (the same improved Barnes-Hut test with a per-pair angle).
//
// Features (8, in [0, 1]):
//   o = com - group centre, d = max(|o|, 1e-30)
//   x0 = rg / (rg + d)        rg = |group half-size vector|
//   x1 = np / 64
//   x2 = l / (l + d)          l = cell bbox size
//   x3 = d / (d + R)          R = root box half size of the system
//   x4 = min(M / M_sys, 1)
//   x5 = |Q|_F / (|Q|_F + M l^2 + 1e-30)
//   x6 = |o_z| / d, x7 = |o_k| / d, k = the group's longest axis
// Network parameters (1377 floats, stored FP16):
//   W1[32][8], b1[32], W2[32][32], b2[32], w3[32], b3.
// Inputs and the first hidden layer are FP16 operands of
// v_mfma_f32_32x32x16_f16; accumulation is FP32; the second hidden layer
// enters the output dot product in FP32.
#pragma once
#include "bt_device.h"

namespace bt {

constexpr int kMlpIn = 8, kMlpH = 32;
constexpr int kMlpParams = 1377;
constexpr int kOffW1 = 0, kOffB1 = 256, kOffW2 = 288, kOffB2 = 1312, kOffW3 = 1344,
              kOffB3 = 1376;

typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef float float16_t __attribute__((ext_vector_type(16)));

// Per-lane matrix-core operands of the network, built from the parameters by
// k_mlp_prepare.  Layout of the 32x32x16 FP16 product D = A B + C used here:
// lane l, half hh = l / 32, holds A[i = l % 32][k = 8 hh + t] and
// B[k = 8 hh + t][j = l % 32] for t = 0..7, and D[i(r, l)][j = l % 32] in
// register r = 0..15 with i(r, l) = 8 (r / 4) + 4 hh + r % 4.  A holds the
// weights (rows = hidden units), B the activations (columns = cells).
struct MlpDev {
  half8_t a1[64];      // layer 1: hh 0 -> W1[i][0..7]; hh 1 -> (b1[i], 0, ..., 0)
  half8_t a2[2][64];   // layer 2, K steps s = 0, 1 (see mlp_unit)
  float c2p[2][16];    // b2[i(r, hh)]: accumulator start of layer 2
  float w3p[2][16];    // w3[i(r, hh)]
  float b3;
  float pad[3];
};

// Hidden unit carried in slot t of the layer-2 B operand at K step s by a
// lane of half hh: the lane's own layer-1 output register r = 8 s + t.
__host__ __device__ inline int mlp_unit(int s, int t, int hh) {
  const int r = 8 * s + t;
  return 8 * (r / 4) + 4 * hh + (r % 4);
}

// Geometry of one group, shared by all its candidate cells.
struct GroupGeom {
  float gcx, gcy, gcz, gsx, gsy, gsz;
  float rg;        // |group half-size vector|
  float np64;      // np / 64
  int axis;        // longest half size: 0 x, 1 y, 2 z (ties -> lower axis)
  float R, inv_msys;   // system scales
};

__device__ inline GroupGeom make_group_geom(float gcx, float gcy, float gcz, float gsx,
                                            float gsy, float gsz, int np, float2 scale) {
#pragma clang fp contract(off)
  GroupGeom g;
  g.gcx = gcx; g.gcy = gcy; g.gcz = gcz;
  g.gsx = gsx; g.gsy = gsy; g.gsz = gsz;
  g.rg = sqrtf(gsx * gsx + gsy * gsy + gsz * gsz);
  g.np64 = (float)np * 0.015625f;
  g.axis = (gsx >= gsy && gsx >= gsz) ? 0 : (gsy >= gsz ? 1 : 2);
  g.R = scale.x;
  g.inv_msys = 1.0f / scale.y;
  return g;
}

// The hardware reciprocal and square root (about 1 ulp) replace the IEEE
// divide and square-root sequences: the features feed an FP16 network, and
// this keeps their cost at a few instructions per cell.
__device__ inline void pair_features(const GroupGeom &g, float4 m, float l, float4 q0, float2 q1,
                                     float x[8]) {
#pragma clang fp contract(off)
  const float ox = m.x - g.gcx, oy = m.y - g.gcy, oz = m.z - g.gcz;
  const float d = fmaxf(__builtin_amdgcn_sqrtf(ox * ox + oy * oy + oz * oz), 1e-30f);
  const float id = __builtin_amdgcn_rcpf(d);
  x[0] = g.rg * __builtin_amdgcn_rcpf(g.rg + d);
  x[1] = g.np64;
  x[2] = l * __builtin_amdgcn_rcpf(l + d);
  x[3] = d * __builtin_amdgcn_rcpf(d + g.R);
  x[4] = fminf(m.w * g.inv_msys, 1.0f);
  const float qn = __builtin_amdgcn_sqrtf(q0.x * q0.x + q0.y * q0.y + q0.z * q0.z +
                                          2.0f * (q0.w * q0.w + q1.x * q1.x + q1.y * q1.y));
  x[5] = qn * __builtin_amdgcn_rcpf(qn + m.w * l * l + 1e-30f);
  x[6] = fabsf(oz) * id;
  const float ok = g.axis == 0 ? ox : (g.axis == 1 ? oy : oz);
  x[7] = fabsf(ok) * id;
}

__device__ __forceinline__ float silu(float z) {
  return z * __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(-1.4426950408889634f * z));
}

__device__ __forceinline__ half8_t pack_features(const float x[8]) {
  half8_t h;
  h[0] = (_Float16)x[0]; h[1] = (_Float16)x[1]; h[2] = (_Float16)x[2]; h[3] = (_Float16)x[3];
  h[4] = (_Float16)x[4]; h[5] = (_Float16)x[5]; h[6] = (_Float16)x[6]; h[7] = (_Float16)x[7];
  return h;
}

// The layer-1 B operand of half 1: the constant input 1 that carries b1.
__device__ __forceinline__ half8_t bias_column() {
  half8_t h;
  h[0] = (_Float16)1.0f;
  h[1] = h[2] = h[3] = h[4] = h[5] = h[6] = h[7] = (_Float16)0.0f;
  return h;
}

// Network output for the 32 cells j = lane % 32 of one block.  All 64 lanes
// must be active.  xb: for lanes of half 0 the features of cell j, for half 1
// bias_column().  c2p / w3p point at this lane's half (16 floats each).
// Returns y for cell j in both lanes j and j + 32.
__device__ __forceinline__ float mlp_block(half8_t a1, half8_t a2_0, half8_t a2_1,
                                          const float *c2p, const float *w3p, float b3,
                                          half8_t xb) {
  float16_t d1 = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  d1 = __builtin_amdgcn_mfma_f32_32x32x16_f16(a1, xb, d1, 0, 0, 0);
  half8_t h0, h1;
#pragma unroll
  for (int r = 0; r < 8; r++) {
    h0[r] = (_Float16)silu(d1[r]);
    h1[r] = (_Float16)silu(d1[r + 8]);
  }
  float16_t d2;
#pragma unroll
  for (int r = 0; r < 16; r++) d2[r] = c2p[r];
  d2 = __builtin_amdgcn_mfma_f32_32x32x16_f16(a2_0, h0, d2, 0, 0, 0);
  d2 = __builtin_amdgcn_mfma_f32_32x32x16_f16(a2_1, h1, d2, 0, 0, 0);
  float y = 0.f;
#pragma unroll
  for (int r = 0; r < 16; r++) y = fmaf(w3p[r], silu(d2[r]), y);
  y += __shfl_xor(y, 32, kWave);
  return y + b3;
}

// Network output for this lane's own feature vector (all 64 lanes active);
// the wave evaluates the block of lanes 0-31 if need_lo and the block of
// lanes 32-63 if need_hi (wave-uniform flags); skipped lanes get 0.
__device__ __forceinline__ float mlp_wave(const half8_t a1, const half8_t a2_0,
                                          const half8_t a2_1, const float *s_mlp, float b3,
                                          const half8_t xown, bool need_lo, bool need_hi) {
  const int lane = lane_id();
  const int hh = lane >> 5;
  const float *c2p = s_mlp + hh * 16;
  const float *w3p = s_mlp + 32 + hh * 16;
  const half8_t bias = bias_column();
  float y0 = 0.f;
  if (need_lo) y0 = mlp_block(a1, a2_0, a2_1, c2p, w3p, b3, hh == 0 ? xown : bias);
  float y1 = 0.f;
  if (need_hi) {
    // lanes of half 0 take the features of lane + 32
    typedef int int4v __attribute__((ext_vector_type(4)));
    int4v v = __builtin_bit_cast(int4v, xown);
    int4v w;
    w[0] = __shfl_xor(v[0], 32, kWave);
    w[1] = __shfl_xor(v[1], 32, kWave);
    w[2] = __shfl_xor(v[2], 32, kWave);
    w[3] = __shfl_xor(v[3], 32, kWave);
    const half8_t xhi = __builtin_bit_cast(half8_t, w);
    y1 = mlp_block(a1, a2_0, a2_1, c2p, w3p, b3, hh == 0 ? xhi : bias);
  }
  return hh == 0 ? y0 : y1;
}

// 1/theta of a pair from the network output, clamped to [1/theta_hi, 1/theta_lo].
__device__ __forceinline__ float learned_inv_theta(float y, float shift, float it_min,
                                                   float it_max) {
  return fminf(fmaxf(exp2f(-(y + shift)), it_min), it_max);
}

}  // namespace bt
