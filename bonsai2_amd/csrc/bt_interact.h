// bt_interact.h -- the walk's two interaction kernels on one i-particle, for
// Treecode<float> (FP32, hardware 1/sqrt) and Treecode<double> (FP64).
#pragma once
#include "bt_device.h"

namespace bt {

template <typename R>
struct Acc4T {
  R x, y, z, w;
};
typedef Acc4T<float> Acc4;

// 1/sqrt(x): FP32, the hardware estimate v_rsq_f32 (about 1 ulp); FP64, the
// hardware estimate v_rsq_f64 (relative error about 2^-23) refined by NEWTON
// Newton steps (each squares the relative error: ~3e-14 after one, below
// FP64 rounding after two).
template <int NEWTON = 2>
__device__ __forceinline__ double rsqrt_newton(double x) {
  double y = __builtin_amdgcn_rsq(x);
#pragma unroll
  for (int k = 0; k < NEWTON; k++) y = y * (1.5 - 0.5 * x * y * y);
  return y;
}
__device__ __forceinline__ float rsq(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ double rsq(double x) { return rsqrt_newton<2>(x); }
__device__ __forceinline__ float rfma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double rfma(double a, double b, double c) { return __builtin_fma(a, b, c); }

// Monopole + quadrupole of one cell (code/computeForces.cu:83-128):
//   phi -= M/r - (trQ/2) M/r^3 + 1.5 (dr.Q.dr) M/r^5
//   a   += (-M/r^3 + 1.5 trQ M/r^5 - 7.5 (dr.Q.dr) M/r^7) dr + 3 M/r^5 Q.dr
// evaluated from the cell's walk record, whose quadrupole is pre-scaled by 3
// (B = 3Qxx, 3Qyy, 3Qzz, 3Qxy; C = 3Qxz, 3Qyz, trQ/2, 1.5 trQ): with
// dr = x_i - com and u = dr.3Q.dr / r^2 the terms become
//   phi -= M/r - (trQ/2) M/r^3 + (u/2) M/r^3
//   a   += (M/r^5 (1.5 trQ - 2.5 u) - M/r^3) dr + M/r^5 (3Q).dr
template <typename R>
__device__ __forceinline__ void p2c_rec(Acc4T<R> &b, R xi, R yi, R zi, const V4<R> A,
                                        const V4<R> B, const V4<R> C, R eps2) {
  const R dx = xi - A.x, dy = yi - A.y, dz = zi - A.z;
  const R r2 = rfma(dz, dz, rfma(dy, dy, rfma(dx, dx, eps2)));
  const R rinv = rsq(r2);
  const R rinv2 = rinv * rinv;
  const R m1 = A.w * rinv;
  const R m3 = rinv2 * m1;
  const R m5 = rinv2 * m3;
  const R qx = rfma(B.x, dx, rfma(B.w, dy, C.x * dz));
  const R qy = rfma(B.w, dx, rfma(B.y, dy, C.y * dz));
  const R qz = rfma(C.x, dx, rfma(C.y, dy, B.z * dz));
  const R u = rfma(qx, dx, rfma(qy, dy, qz * dz)) * rinv2;
  b.w -= m1;
  b.w = rfma(-m3, rfma(R(0.5), u, -C.z), b.w);
  const R cc = rfma(m5, rfma(R(-2.5), u, C.w), -m3);
  b.x = rfma(cc, dx, rfma(m5, qx, b.x));
  b.y = rfma(cc, dy, rfma(m5, qy, b.y));
  b.z = rfma(cc, dz, rfma(m5, qz, b.z));
}

// Particle-particle, code/computeForces.cu:58-78.
template <typename R>
__device__ __forceinline__ void p2p(Acc4T<R> &b, R xi, R yi, R zi, const V4<R> A, R eps2) {
  const R dx = A.x - xi, dy = A.y - yi, dz = A.z - zi;
  const R r2 = rfma(dz, dz, rfma(dy, dy, rfma(dx, dx, eps2)));
  const R rinv = rsq(r2);
  const R mr1 = A.w * rinv;
  const R mr3 = mr1 * rinv * rinv;
  b.w -= mr1;
  b.x = rfma(mr3, dx, b.x);
  b.y = rfma(mr3, dy, b.y);
  b.z = rfma(mr3, dz, b.z);
}

}  // namespace bt
