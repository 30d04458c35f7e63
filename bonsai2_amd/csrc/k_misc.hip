// k_misc.hip -- moveParticles, computeEnergies and the FP64 direct sum.
//
// The reference's moveParticles and computeEnergies are stubs that only
// print (code/moveParticles.cu:3-7, code/computeEnergies.cu:3-7).  Here:
//   * kick:  v += a * dt/2, with a in PH order mapped to the tree-ordered
//            velocities through perm (PH index -> tree index);
//   * drift: x += v * dt (tree order);
//   * energies: FP64 per-system sums K = 1/2 sum m v^2 and
//     W = 1/2 sum m (phi + m/eps) -- the walk's phi includes the self term
//     -m/eps of the P2P kernel with dr = 0 (code/computeForces.cu:64-72);
//     partial sums are reduced in a fixed order (deterministic);
//   * FP64 direct sum: the error oracle for the walk, sources staged through
//     LDS, 1/sqrt by the v_rsq_f64 estimate refined with Newton steps in FP64.
#include <cfloat>

#include <algorithm>

#include "bt_device.h"
#include "bt_internal.h"

namespace bt {

namespace {

constexpr int kBlock = 256;
constexpr int kEnergyBlocks = 64;
constexpr int kDirectSlices = 64;
constexpr int kDirectTargets = kBlock;         // targets per block of the direct sum

// Both integrator kernels do nothing once an error is on record (err_sticky):
// a step whose build or walk failed does not kick with invalid forces.
template <typename R>
__global__ __launch_bounds__(kBlock) void k_kick(int n, const int *__restrict__ perm,
                                                 const V4<R> *__restrict__ acc,
                                                 V4<R> *__restrict__ vel, R dth,
                                                 DevCounters *dc, int closing) {
  if (dc->err_sticky) return;
  if (closing && blockIdx.x == 0 && threadIdx.x == 0) dc->steps_done++;
  for (int k = blockIdx.x * kBlock + threadIdx.x; k < n; k += gridDim.x * kBlock) {
    const int t = perm[k];
    const V4<R> a = acc[k];
    V4<R> v = vel[t];
    v.x += a.x * dth;
    v.y += a.y * dth;
    v.z += a.z * dth;
    vel[t] = v;
  }
}

// The opening half of a kick-drift-kick step in one pass: v += a dth for
// the particle at tree index perm[k], then x += v dt for the same particle;
// the arithmetic of a kick pass followed by a drift pass (so the same
// results), with v kept in registers instead of written and read back.
template <typename R>
__global__ __launch_bounds__(kBlock) void k_kick_drift(int n, const int *__restrict__ perm,
                                                       const V4<R> *__restrict__ acc,
                                                       V4<R> *__restrict__ vel,
                                                       V4<R> *__restrict__ pos, R dth, R dt,
                                                       const DevCounters *dc) {
  if (dc->err_sticky) return;
  for (int k = blockIdx.x * kBlock + threadIdx.x; k < n; k += gridDim.x * kBlock) {
    const int t = perm[k];
    const V4<R> a = acc[k];
    V4<R> v = vel[t];
    V4<R> p = pos[t];
    v.x += a.x * dth;
    v.y += a.y * dth;
    v.z += a.z * dth;
    vel[t] = v;
    p.x += v.x * dt;
    p.y += v.y * dt;
    p.z += v.z * dt;
    pos[t] = p;
  }
}

__global__ void k_clear_sticky(DevCounters *dc) {
  if (threadIdx.x == 0) {
    dc->err_sticky = 0;
    dc->stack_hw = 0;
    dc->steps_done = 0;
  }
}

__device__ __forceinline__ double block_sum_d(double v, double *sh) {
  v = wave_sum(v);
  const int w = threadIdx.x >> 6;
  if (lane_id() == 0) sh[w] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0)
    for (int k = 0; k < (int)(blockDim.x >> 6); k++) t += sh[k];
  __syncthreads();
  return t;
}

// grid (kEnergyBlocks, nsys): partial[(r*B + b)*2 + {0,1}] = (K, W) partial sums
template <typename R>
__global__ __launch_bounds__(kBlock) void k_energy_partial(
    const int *__restrict__ off, const V4<R> *__restrict__ pos, const V4<R> *__restrict__ vel,
    const V4<R> *__restrict__ pos_ph, const V4<R> *__restrict__ acc, double inv_eps,
    double *__restrict__ part) {
  __shared__ double sh[4];
  const int r = blockIdx.y;
  const int b = off[r], e = off[r + 1];
  double k = 0.0, w = 0.0;
  for (int i = b + blockIdx.x * kBlock + threadIdx.x; i < e; i += gridDim.x * kBlock) {
    const V4<R> p = pos[i], v = vel[i];
    const double m = p.w;
    k += 0.5 * m * ((double)v.x * v.x + (double)v.y * v.y + (double)v.z * v.z);
    const double mp = pos_ph[i].w;
    w += 0.5 * mp * ((double)acc[i].w + mp * inv_eps);
  }
  k = block_sum_d(k, sh);
  w = block_sum_d(w, sh);
  if (threadIdx.x == 0) {
    part[2 * (r * gridDim.x + blockIdx.x)] = k;
    part[2 * (r * gridDim.x + blockIdx.x) + 1] = w;
  }
}

__global__ void k_energy_final(int nsys, int nb, const double *__restrict__ part,
                               double *__restrict__ out) {
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < nsys; r += gridDim.x * blockDim.x) {
    double k = 0.0, w = 0.0;
    for (int b = 0; b < nb; b++) {
      k += part[2 * (r * nb + b)];
      w += part[2 * (r * nb + b) + 1];
    }
    out[2 * r] = k;
    out[2 * r + 1] = w;
  }
}

// 1/sqrt in FP64: the hardware FP64 estimate v_rsq_f64 (relative error
// about 2^-23) refined by NEWTON Newton steps (each squares the relative
// error: ~3e-14 after one, below FP64 rounding after two).  CLAMP (no
// softening): r2 is clamped to FLT_MIN so that the self pair stays finite.
template <int NEWTON, bool CLAMP>
__device__ __forceinline__ double rsqrt_fp64(double r2) {
  const double x = CLAMP ? fmax(r2, (double)FLT_MIN) : r2;
  double y = __builtin_amdgcn_rsq(x);
#pragma unroll
  for (int k = 0; k < NEWTON; k++) y = y * (1.5 - 0.5 * x * y * y);
  return y;
}

// grid (ceil(ns/256), slices): each block sums a slice of the source index
// range for 256 targets, the slice staged in LDS as FP64.  FOREST: sources
// outside a target's system are skipped.  The kernel is bound by FP64 VALU
// issue (18 FP64 instructions per pair), not by the LDS broadcasts:
// two targets per thread, halving the LDS reads, measured no faster.
// SOFT: eps > 0, so r2 > 0 for every pair and 1/sqrt needs no clamp.
template <int NEWTON, bool FOREST, bool SOFT, typename R>
__global__ __launch_bounds__(kBlock) void k_direct(int ns, const int *__restrict__ idx, int n,
                                                   int nsys, const int *__restrict__ off,
                                                   const V4<R> *__restrict__ pos, double eps2,
                                                   double4 *__restrict__ part) {
  __shared__ double4 tile[kBlock];
  const int t = blockIdx.x * kBlock + threadIdx.x;
  const bool active = t < ns;
  // a target index outside [0, n) reads particle 0 and yields a NaN row
  const int k_in = active ? idx[t] : 0;
  const bool ok = k_in >= 0 && k_in < n;
  const int k = ok ? k_in : 0;
  const int r = FOREST ? system_of(k, nsys, off) : 0;
  const int lo = FOREST ? off[r] : 0, hi = FOREST ? off[r + 1] : n;
  const V4<R> pt = pos[k];
  const double px = pt.x, py = pt.y, pz = pt.z;
  const int chunk = (n + gridDim.y - 1) / gridDim.y;
  const int s0 = blockIdx.y * chunk, s1 = min(n, s0 + chunk);
  double ax = 0.0, ay = 0.0, az = 0.0, ph = 0.0;
  for (int j0 = s0; j0 < s1; j0 += kBlock) {
    const int j = j0 + threadIdx.x;
    const V4<R> pj = j < s1 ? pos[j] : make_v4<R>(R(0), R(0), R(0), R(0));
    tile[threadIdx.x] = make_double4(pj.x, pj.y, pj.z, pj.w);
    __syncthreads();
    const int cnt = min(kBlock, s1 - j0);
    const int q0 = FOREST ? max(0, lo - j0) : 0;
    const int q1 = FOREST ? min(cnt, hi - j0) : cnt;
#pragma unroll 4
    for (int q = q0; q < q1; q++) {
      const double4 p = tile[q];
      const double dx = p.x - px, dy = p.y - py, dz = p.z - pz;
      const double r2 = fma(dz, dz, fma(dy, dy, fma(dx, dx, eps2)));
      const double ri = rsqrt_fp64<NEWTON, !SOFT>(r2);
      const double mr = p.w * ri;
      const double mr3 = mr * (ri * ri);
      ax += mr3 * dx; ay += mr3 * dy; az += mr3 * dz; ph -= mr;
    }
    __syncthreads();
  }
  if (active) {
    const double nan = __builtin_nan("");
    part[(size_t)blockIdx.y * ns + t] =
        ok ? make_double4(ax, ay, az, ph) : make_double4(nan, nan, nan, nan);
  }
}

// One wave per target: lane l sums slices l, l + 64, ... in order, then the
// wave adds its lanes in a fixed tree order (deterministic).
__global__ __launch_bounds__(kBlock) void k_direct_reduce(int ns, int nslices,
                                                          const double4 *__restrict__ part,
                                                          double4 *__restrict__ out) {
  const int t = blockIdx.x * (kBlock / kWave) + (threadIdx.x >> 6);
  if (t >= ns) return;   // wave-uniform
  double4 s = make_double4(0, 0, 0, 0);
  for (int q = lane_id(); q < nslices; q += kWave) {
    const double4 v = part[(size_t)q * ns + t];
    s.x += v.x; s.y += v.y; s.z += v.z; s.w += v.w;
  }
  s.x = wave_sum(s.x);
  s.y = wave_sum(s.y);
  s.z = wave_sum(s.z);
  s.w = wave_sum(s.w);
  if (lane_id() == 0) out[t] = s;
}

}  // namespace

void launch_kick(Ctx &c, double dth, bool closing) {
  const dim3 g(grid_for(c.n, kBlock, 8192));
  if (c.fp64)
    hipLaunchKernelGGL(k_kick<double>, g, dim3(kBlock), 0, c.stream, c.n, c.perm,
                       as4<double>(c.acc), as4<double>(c.vel), dth, c.dc, closing ? 1 : 0);
  else
    hipLaunchKernelGGL(k_kick<float>, g, dim3(kBlock), 0, c.stream, c.n, c.perm, c.acc, c.vel,
                       (float)dth, c.dc, closing ? 1 : 0);
}

void launch_kick_drift(Ctx &c, double dth, double dt) {
  const dim3 g(grid_for(c.n, kBlock, 8192));
  if (c.fp64)
    hipLaunchKernelGGL(k_kick_drift<double>, g, dim3(kBlock), 0, c.stream, c.n, c.perm,
                       as4<double>(c.acc), as4<double>(c.vel), as4<double>(c.pos), dth, dt, c.dc);
  else
    hipLaunchKernelGGL(k_kick_drift<float>, g, dim3(kBlock), 0, c.stream, c.n, c.perm, c.acc,
                       c.vel, c.pos, (float)dth, (float)dt, c.dc);
}

void launch_clear_sticky(Ctx &c) {
  hipLaunchKernelGGL(k_clear_sticky, dim3(1), dim3(64), 0, c.stream, c.dc);
}

void launch_energies(Ctx &c) {
  const double inv_eps = c.eps > 0.0 ? 1.0 / c.eps : 0.0;
  double *part = c.d_energy + 2 * c.nsys;
  if (c.fp64)
    hipLaunchKernelGGL(k_energy_partial<double>, dim3(kEnergyBlocks, c.nsys), dim3(kBlock), 0,
                       c.stream, c.d_off, as4<double>(c.pos), as4<double>(c.vel),
                       as4<double>(c.pos_ph), as4<double>(c.acc), inv_eps, part);
  else
    hipLaunchKernelGGL(k_energy_partial<float>, dim3(kEnergyBlocks, c.nsys), dim3(kBlock), 0,
                       c.stream, c.d_off, c.pos, c.vel, c.pos_ph, c.acc, inv_eps, part);
  hipLaunchKernelGGL(k_energy_final, dim3(grid_for(c.nsys, 64, 1024)), dim3(64), 0, c.stream,
                     c.nsys, kEnergyBlocks, part, c.d_energy);
}

// Slices of the source range: enough blocks to fill the chip for any number
// of targets (256 targets per block).
int direct_slices(int ns) {
  const int nb = (ns + kDirectTargets - 1) / kDirectTargets;
  return std::min(2048, std::max(kDirectSlices, (8192 + nb - 1) / nb));
}

// d_out must hold ns * (direct_slices(ns) + 1) double4: the result, then the partials.
void launch_direct(Ctx &c, int ns, const int *d_idx, double *d_out, bool check) {
  launch_direct_parts(c, ns, d_idx, d_out, d_out + (size_t)4 * ns, direct_slices(ns), check);
}

// d_out: ns double4; d_part: ns * slices double4, every entry written before
// it is read (nothing needs clearing).  The sums of slices are added in slice
// order, so equal slices give equal results.
void launch_direct_parts(Ctx &c, int ns, const int *d_idx, double *d_out, double *d_part,
                         int slices, bool check, const void *src) {
  if (!src) src = c.pos_ph;
  double4 *part = reinterpret_cast<double4 *>(d_part);
  // the softening the step uses: float(eps) squared in FP64 (Treecode<float>),
  // the FP64 eps squared (Treecode<double>)
  const double eps2 = c.eps * c.eps;
  // the per-step check of a learned walk needs relative errors of ~1e-6:
  // one Newton step (~3e-14) is ample; the measurement path keeps two
  const dim3 grid((ns + kDirectTargets - 1) / kDirectTargets, slices);
  const bool soft = eps2 > 0.0, forest = c.nsys > 1;
  if (c.fp64) {
    void (*kern)(int, const int *, int, int, const int *, const double4 *, double, double4 *);
    if (soft) kern = forest ? k_direct<2, true, true, double> : k_direct<2, false, true, double>;
    else kern = forest ? k_direct<2, true, false, double> : k_direct<2, false, false, double>;
    hipLaunchKernelGGL(kern, grid, dim3(kBlock), 0, c.stream, ns, d_idx, c.n, c.nsys, c.d_off,
                       static_cast<const double4 *>(src), eps2, part);
  } else {
    void (*kern)(int, const int *, int, int, const int *, const float4 *, double, double4 *);
    if (soft) {
      if (forest) kern = check ? k_direct<1, true, true, float> : k_direct<2, true, true, float>;
      else kern = check ? k_direct<1, false, true, float> : k_direct<2, false, true, float>;
    } else {
      if (forest) kern = check ? k_direct<1, true, false, float> : k_direct<2, true, false, float>;
      else kern = check ? k_direct<1, false, false, float> : k_direct<2, false, false, float>;
    }
    hipLaunchKernelGGL(kern, grid, dim3(kBlock), 0, c.stream, ns, d_idx, c.n, c.nsys, c.d_off,
                       static_cast<const float4 *>(src), eps2, part);
  }
  const int tpb = kBlock / kWave;   // targets per reduce block
  hipLaunchKernelGGL(k_direct_reduce, dim3((ns + tpb - 1) / tpb), dim3(kBlock), 0, c.stream, ns,
                     slices, part, reinterpret_cast<double4 *>(d_out));
}

}  // namespace bt
