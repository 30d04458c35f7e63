// bt_interact.h -- the walk's two interaction kernels on one i-particle
// (shared by the walk and the issue-rate probe).
#pragma once
#include "bt_device.h"

namespace bt {

struct Acc4 {
  float x, y, z, w;
};

// Monopole + quadrupole of one cell (code/computeForces.cu:83-128):
//   phi -= M/r - (trQ/2) M/r^3 + 1.5 (dr.Q.dr) M/r^5
//   a   += (-M/r^3 + 1.5 trQ M/r^5 - 7.5 (dr.Q.dr) M/r^7) dr + 3 M/r^5 Q.dr
// evaluated from the cell's walk record, whose quadrupole is pre-scaled by 3
// (B = 3Qxx, 3Qyy, 3Qzz, 3Qxy; C = 3Qxz, 3Qyz, trQ/2, 1.5 trQ): with
// dr = x_i - com and u = dr.3Q.dr / r^2 the terms become
//   phi -= M/r - (trQ/2) M/r^3 + (u/2) M/r^3
//   a   += (M/r^5 (1.5 trQ - 2.5 u) - M/r^3) dr + M/r^5 (3Q).dr
__device__ __forceinline__ void p2c_rec(Acc4 &b, float xi, float yi, float zi, const float4 A,
                                        const float4 B, const float4 C, float eps2) {
  const float dx = xi - A.x, dy = yi - A.y, dz = zi - A.z;
  const float r2 = fmaf(dz, dz, fmaf(dy, dy, fmaf(dx, dx, eps2)));
  const float rinv = __builtin_amdgcn_rsqf(r2);
  const float rinv2 = rinv * rinv;
  const float m1 = A.w * rinv;
  const float m3 = rinv2 * m1;
  const float m5 = rinv2 * m3;
  const float qx = fmaf(B.x, dx, fmaf(B.w, dy, C.x * dz));
  const float qy = fmaf(B.w, dx, fmaf(B.y, dy, C.y * dz));
  const float qz = fmaf(C.x, dx, fmaf(C.y, dy, B.z * dz));
  const float u = fmaf(qx, dx, fmaf(qy, dy, qz * dz)) * rinv2;
  b.w -= m1;
  b.w = fmaf(-m3, fmaf(0.5f, u, -C.z), b.w);
  const float cc = fmaf(m5, fmaf(-2.5f, u, C.w), -m3);
  b.x = fmaf(cc, dx, fmaf(m5, qx, b.x));
  b.y = fmaf(cc, dy, fmaf(m5, qy, b.y));
  b.z = fmaf(cc, dz, fmaf(m5, qz, b.z));
}

// Particle-particle, code/computeForces.cu:58-78.
__device__ __forceinline__ void p2p(Acc4 &b, float xi, float yi, float zi, const float4 A,
                                    float eps2) {
  const float dx = A.x - xi, dy = A.y - yi, dz = A.z - zi;
  const float r2 = fmaf(dz, dz, fmaf(dy, dy, fmaf(dx, dx, eps2)));
  const float rinv = __builtin_amdgcn_rsqf(r2);
  const float mr1 = A.w * rinv;
  const float mr3 = mr1 * rinv * rinv;
  b.w -= mr1;
  b.x = fmaf(mr3, dx, b.x);
  b.y = fmaf(mr3, dy, b.y);
  b.z = fmaf(mr3, dz, b.z);
}

}  // namespace bt
