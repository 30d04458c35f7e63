// bt_device.h -- wave64 helpers shared by the CDNA4 kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bt {

constexpr int kWave = 64;
constexpr int kMaxLevel = 21;  // octant digits in a 63-bit geometric key

// The kernels of the build (box, keys, sorts, gathers, levels, multipoles,
// groups) raise their waves' issue priority above the default 0: during a
// step the learned walk's FP64 check sum runs beside them on another stream,
// and on a SIMD they share, the build's latency-bound waves then issue ahead
// of its arithmetic.
#ifndef BT_BUILD_WAVE_PRIO
#define BT_BUILD_WAVE_PRIO 1
#endif
#define BT_BUILD_PRIO() __builtin_amdgcn_s_setprio(BT_BUILD_WAVE_PRIO)

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & (kWave - 1)); }

__device__ __forceinline__ uint64_t ballot64(bool p) { return (uint64_t)__ballot(p); }

__device__ __forceinline__ uint64_t lanemask_lt() {
  return (1ull << lane_id()) - 1ull;
}

// Orders LDS accesses of different lanes of one wave: all lanes' prior LDS
// writes are visible to all lanes' later LDS reads (a wave executes its LDS
// instructions in order; this stops the compiler from moving them).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <typename T>
__device__ __forceinline__ T wave_bcast(T v, int src) { return __shfl(v, src, kWave); }

__device__ __forceinline__ int wave_uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

__device__ __forceinline__ float wave_min(float v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v = fminf(v, __shfl_xor(v, m, kWave));
  return v;
}
__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v = fmaxf(v, __shfl_xor(v, m, kWave));
  return v;
}
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, kWave);
  return v;
}
__device__ __forceinline__ int wave_max_int(int v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v = max(v, __shfl_xor(v, m, kWave));
  return v;
}
// Inclusive prefix sum over the 64 lanes.
__device__ __forceinline__ int wave_incl_scan(int v) {
  const int l = lane_id();
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    const int t = __shfl_up(v, d, kWave);
    if (l >= d) v += t;
  }
  return v;
}

// Vector types of the real type: Treecode<float> (float4 particles and cell
// data) and Treecode<double> (double4).  A context allocates its particle and
// cell arrays with the element size of its real type; kernels view them
// through as4 / as2.
template <typename R> struct RealVec;
template <> struct RealVec<float> { typedef float4 v4; typedef float2 v2; };
template <> struct RealVec<double> { typedef double4 v4; typedef double2 v2; };
template <typename R> using V4 = typename RealVec<R>::v4;
template <typename R> using V2 = typename RealVec<R>::v2;

template <typename R>
__host__ __device__ __forceinline__ V4<R> make_v4(R x, R y, R z, R w) {
  V4<R> v;
  v.x = x; v.y = y; v.z = z; v.w = w;
  return v;
}
template <typename R>
__host__ __device__ __forceinline__ V2<R> make_v2(R x, R y) {
  V2<R> v;
  v.x = x; v.y = y;
  return v;
}
template <typename R, typename P>
__host__ __device__ __forceinline__ V4<R> *as4(P *p) { return reinterpret_cast<V4<R> *>(p); }
template <typename R, typename P>
__host__ __device__ __forceinline__ V2<R> *as2(P *p) { return reinterpret_cast<V2<R> *>(p); }

// min / max of the real type, and over the wave
template <typename R>
__device__ __forceinline__ R rmin(R a, R b) { return fmin(a, b); }
template <typename R>
__device__ __forceinline__ R rmax(R a, R b) { return fmax(a, b); }
template <>
__device__ __forceinline__ float rmin(float a, float b) { return fminf(a, b); }
template <>
__device__ __forceinline__ float rmax(float a, float b) { return fmaxf(a, b); }
template <typename R>
__device__ __forceinline__ R wave_min_r(R v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v = rmin(v, __shfl_xor(v, m, kWave));
  return v;
}
template <typename R>
__device__ __forceinline__ R wave_max_r(R v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v = rmax(v, __shfl_xor(v, m, kWave));
  return v;
}

// System (root) of a particle index: offsets[0..nsys] partition [0, n).
__device__ __forceinline__ int system_of(int i, int nsys, const int *__restrict__ off) {
  int lo = 0, hi = nsys - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= i) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// CellData bit layout (reference code/Treecode.h:45-89):
//   x = parent | level << 27; y = first | (nchild-1) << 29, or 0xFFFFFFFF for
//   a leaf; z = pbeg; w = pend (half-open, tree order).
__device__ __forceinline__ bool cell_leaf(const uint4 &c) { return c.y == 0xFFFFFFFFu; }
__device__ __forceinline__ int cell_first(const uint4 &c) { return (int)(c.y & 0x1FFFFFFFu); }
__device__ __forceinline__ int cell_nchild(const uint4 &c) { return (int)(c.y >> 29) + 1; }
__device__ __forceinline__ uint32_t cell_pack_y(int first, int nchild) {
  return (uint32_t)first | ((uint32_t)(nchild - 1) << 29);
}

}  // namespace bt
