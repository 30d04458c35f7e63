// probe_issue.hip -- issue cost of the FP32 VALU operations the walk is made
// of, and of the walk's own P2C / P2P batch loops, on this GPU.
//
//   bonsai2_amd/probe_issue        (built by bonsai2_amd/build.py)
//
// Part 1: one operation kind per kernel, 8 independent chains per lane, 4 and
// 8 waves per SIMD; prints cycles (at the 2.4 GHz peak clock) per wave64
// instruction per SIMD.  Part 2: the walk's P2C (p2c_rec, 35 VALU) and P2P
// (p2p, 14 VALU) on a 64-entry batch staged in LDS and read by broadcast, as
// the walk does, with one i-particle per lane (the walk's shape) and with two
// (each LDS read serving two interactions); prints cycles per
// wave-interaction and TFLOP/s in the reference's count (64 / 20 flops).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

#include "bt_interact.h"

using bt::Acc4;

namespace {

constexpr int kIters = 4096;
constexpr int kReps = 64;
constexpr double kClock = 2.4e9;

#define CK(x)                                                                       \
  do {                                                                              \
    hipError_t e_ = (x);                                                            \
    if (e_ != hipSuccess) {                                                         \
      std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));                  \
      std::exit(1);                                                                 \
    }                                                                               \
  } while (0)

enum Op { kFmac, kFma3, kMul, kAdd, kExp, kRsq };

template <int OP>
__global__ __launch_bounds__(256) void k_op(float *out, float s) {
  float a[8];
  const float b = s * (float)threadIdx.x, c = s + (float)threadIdx.x;
#pragma unroll
  for (int i = 0; i < 8; i++) a[i] = s * (float)(i + 1);
  for (int it = 0; it < kIters; it++) {
#pragma unroll
    for (int i = 0; i < 8; i++) {
      if (OP == kFmac) a[i] = fmaf(b, c, a[i]);                 // v_fmac_f32 (VOP2)
      if (OP == kFma3) a[i] = fmaf(a[i], b, c);                 // v_fma_f32 (VOP3)
      if (OP == kMul) a[i] = a[i] * b;                          // v_mul_f32
      if (OP == kAdd) a[i] = a[i] + b;                          // v_add_f32
      if (OP == kExp) a[i] = __builtin_amdgcn_exp2f(a[i]);      // v_exp_f32
      if (OP == kRsq) a[i] = __builtin_amdgcn_rsqf(a[i]);       // v_rsq_f32
    }
  }
  float r = 0.f;
#pragma unroll
  for (int i = 0; i < 8; i++) r += a[i];
  out[blockIdx.x * 256 + threadIdx.x] = r;
}

template <int NI, bool P2C>
__global__ __launch_bounds__(256) void k_batch(const float4 *__restrict__ rec, float4 *out,
                                               float eps2) {
  __shared__ float4 jA[4][64], jB[4][64], jC[4][64];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  jA[w][l] = rec[3 * l];
  jB[w][l] = rec[3 * l + 1];
  jC[w][l] = rec[3 * l + 2];
  __syncthreads();
  float xi[NI], yi[NI], zi[NI];
  Acc4 b[NI];
#pragma unroll
  for (int i = 0; i < NI; i++) {
    xi[i] = 0.01f * (float)(threadIdx.x + 7 * i);
    yi[i] = 0.02f * (float)i;
    zi[i] = 0.003f * (float)threadIdx.x;
    b[i] = {0.f, 0.f, 0.f, 0.f};
  }
  for (int r = 0; r < kReps; r++) {
#pragma unroll 8
    for (int j = 0; j < 64; j++) {
      const float4 A = jA[w][j];
      if (P2C) {
        const float4 B = jB[w][j], C = jC[w][j];
#pragma unroll
        for (int i = 0; i < NI; i++) bt::p2c_rec(b[i], xi[i], yi[i], zi[i], A, B, C, eps2);
      } else {
#pragma unroll
        for (int i = 0; i < NI; i++) bt::p2p(b[i], xi[i], yi[i], zi[i], A, eps2);
      }
    }
  }
  float4 s = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
  for (int i = 0; i < NI; i++) {
    s.x += b[i].x; s.y += b[i].y; s.z += b[i].z; s.w += b[i].w;
  }
  out[blockIdx.x * 256 + threadIdx.x] = s;
}

template <typename K, typename... Args>
float time_ms(K kern, int blocks, Args... args) {
  hipEvent_t e0, e1;
  CK(hipEventCreate(&e0));
  CK(hipEventCreate(&e1));
  hipLaunchKernelGGL(kern, dim3(blocks), dim3(256), 0, 0, args...);
  CK(hipGetLastError());
  CK(hipEventRecord(e0));
  for (int r = 0; r < 5; r++) hipLaunchKernelGGL(kern, dim3(blocks), dim3(256), 0, 0, args...);
  CK(hipEventRecord(e1));
  CK(hipEventSynchronize(e1));
  float ms = 0.f;
  CK(hipEventElapsedTime(&ms, e0, e1));
  CK(hipEventDestroy(e0));
  CK(hipEventDestroy(e1));
  return ms / 5.f;
}

}  // namespace

int main() {
  hipDeviceProp_t prop;
  CK(hipGetDeviceProperties(&prop, 0));
  const int cus = prop.multiProcessorCount;
  const double simds = 4.0 * cus;
  std::printf("device %s, %d CUs; cycles at %.1f GHz\n", prop.gcnArchName, cus, kClock / 1e9);
  const int max_blocks = 2 * cus * 4;   // 8 waves per SIMD with 256-thread blocks
  float *dout;
  float4 *rec, *dout4;
  CK(hipMalloc(&dout, sizeof(float) * 256 * max_blocks));
  CK(hipMalloc(&dout4, sizeof(float4) * 256 * max_blocks));
  CK(hipMalloc(&rec, sizeof(float4) * 192));
  float4 h[192];
  for (int k = 0; k < 64; k++) {
    h[3 * k] = make_float4(0.5f + 0.01f * k, 0.3f, -0.2f + 0.001f * k, 1e-3f);
    h[3 * k + 1] = make_float4(3e-4f, 2e-4f, -5e-4f, 1e-4f);
    h[3 * k + 2] = make_float4(-1e-4f, 2e-5f, 0.f, 0.f);
  }
  CK(hipMemcpy(rec, h, sizeof(h), hipMemcpyHostToDevice));

  std::printf("\n%-22s %14s %14s\n", "operation", "4 waves/SIMD", "8 waves/SIMD");
  const char *names[] = {"v_fmac_f32 (VOP2)", "v_fma_f32 (VOP3)", "v_mul_f32", "v_add_f32",
                         "v_exp_f32", "v_rsq_f32"};
  using OpK = void (*)(float *, float);
  const OpK ops[] = {k_op<kFmac>, k_op<kFma3>, k_op<kMul>, k_op<kAdd>, k_op<kExp>, k_op<kRsq>};
  for (int o = 0; o < 6; o++) {
    std::printf("%-22s", names[o]);
    for (int wps = 4; wps <= 8; wps += 4) {
      const int blocks = cus * wps;   // 4 waves per block
      const float ms = time_ms(ops[o], blocks, dout, 0.999f);
      const double per_simd = (double)blocks * 4.0 * 8.0 * kIters / simds;
      std::printf(" %10.2f cyc", ms * 1e-3 * kClock / per_simd);
    }
    std::printf("\n");
  }

  std::printf("\n%-28s %12s %22s %12s\n", "batch loop", "waves/SIMD", "cycles/wave-interaction",
              "TFLOP/s");
  struct Loop {
    const char *name;
    void (*k)(const float4 *, float4 *, float);
    int ni, flops;
  };
  const Loop loops[] = {{"P2C, 1 i-particle per lane", k_batch<1, true>, 1, 64},
                        {"P2C, 2 i-particles per lane", k_batch<2, true>, 2, 64},
                        {"P2P, 1 i-particle per lane", k_batch<1, false>, 1, 20},
                        {"P2P, 2 i-particles per lane", k_batch<2, false>, 2, 20}};
  for (const Loop &lp : loops) {
    for (int wps = 4; wps <= 8; wps += 4) {
      const int blocks = cus * wps;
      const float ms = time_ms(lp.k, blocks, (const float4 *)rec, dout4, 1e-4f);
      const double inter = (double)blocks * 4.0 * kReps * 64 * lp.ni;   // wave-interactions
      std::printf("%-28s %12d %22.1f %12.1f\n", lp.name, wps,
                  ms * 1e-3 * kClock / (inter / simds), inter * 64.0 * lp.flops / (ms * 1e-3) / 1e12);
    }
  }
  CK(hipFree(dout));
  CK(hipFree(dout4));
  CK(hipFree(rec));
  return 0;
}
