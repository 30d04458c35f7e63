// k_sort.hip -- stable LSD radix sort of (u64 key, int value) pairs and
// device-wide scans, written for wave64.
//
// Replaces thrust::sort_by_key / thrust::inclusive_scan_by_key of the
// reference (code/makeGroups.cu:270-295, code/buildTree.cu:940-944), whose
// contract is kept: nondecreasing keys, ties in input order.
//
// One pass = 8-bit digit: (1) per-tile histograms in LDS, written digit-major;
// (2) exclusive scan of the digit-major histogram = global scatter offsets;
// (3) stable ranks: each wave ranks its contiguous quarter of the tile with 8
// ballots per 64 pairs (lanes with equal digits form a peer mask) against its
// own digit counters in LDS, keeping the pairs in registers; the waves' counts
// are then chained per digit, each pair is placed at its sorted position
// within the tile in LDS, and the tile is written out in that order, so
// consecutive threads store consecutive addresses of one digit's run (whole
// cache lines instead of one scattered store per pair).
// A tile is 256 threads x 16 rounds = 4096 pairs; HBM traffic per pass is
// ~8 B (histogram read) + 12 B read + 12 B written per pair.
#include <climits>

#include "bt_device.h"
#include "bt_internal.h"

namespace bt {

namespace {

constexpr int kSortThreads = 256;
constexpr int kSortRounds = 16;
constexpr int kSortTile = kSortThreads * kSortRounds;
constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;

struct DigitSrc {
  int shift;        // bit offset of the digit
  int nsys;
  const int *off;
};

// BYSYS: the digit is taken from system_of(value) instead of the key (a
// compile-time choice, so the key passes carry no branch and their loads can
// all be issued before the first is used).
template <bool BYSYS>
__device__ __forceinline__ int digit_of(uint64_t key, int val, const DigitSrc &s) {
  if (BYSYS) return (system_of(val, s.nsys, s.off) >> s.shift) & 0xFF;
  return (int)((key >> s.shift) & 0xFFull);
}

// ---- scans ---------------------------------------------------------------

struct OpSum {
  static __device__ __forceinline__ int id() { return 0; }
  static __device__ __forceinline__ int op(int a, int b) { return a + b; }
};
struct OpMax {
  static __device__ __forceinline__ int id() { return INT_MIN; }
  static __device__ __forceinline__ int op(int a, int b) { return a > b ? a : b; }
};

// Inclusive scan across a 256-thread block; returns inclusive value, *total = block total.
template <class Op>
__device__ __forceinline__ int block_incl_scan(int v, int *sh_wave, int *total) {
  const int l = lane_id(), w = threadIdx.x >> 6;
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    const int t = __shfl_up(v, d, kWave);
    if (l >= d) v = Op::op(t, v);
  }
  if (l == kWave - 1) sh_wave[w] = v;
  __syncthreads();
  int pre = Op::id();
  for (int k = 0; k < w; k++) pre = Op::op(pre, sh_wave[k]);
  int tot = Op::id();
  for (int k = 0; k < (int)(blockDim.x >> 6); k++) tot = Op::op(tot, sh_wave[k]);
  *total = tot;
  __syncthreads();
  return Op::op(pre, v);
}

template <bool BYSYS>
__global__ __launch_bounds__(kSortThreads) void k_radix_hist(const uint64_t *__restrict__ keys,
                                                             const int *__restrict__ vals,
                                                             int n, DigitSrc src, int nblocks,
                                                             int *__restrict__ ghist) {
  __shared__ int hist[256];
  hist[threadIdx.x] = 0;
  const int base = blockIdx.x * kSortTile;
  // all of the tile's loads first, then the LDS counting
  int dg[kSortRounds];
#pragma unroll
  for (int r = 0; r < kSortRounds; r++) {
    const int i = base + r * kSortThreads + threadIdx.x;
    const int ic = min(i, n - 1);   // unconditional loads: no branch between them
    const int d = digit_of<BYSYS>(BYSYS ? 0ull : keys[ic], BYSYS ? vals[ic] : 0, src);
    dg[r] = i < n ? d : -1;
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kSortRounds; r++)
    if (dg[r] >= 0) atomicAdd(&hist[dg[r]], 1);
  __syncthreads();
  ghist[threadIdx.x * nblocks + blockIdx.x] = hist[threadIdx.x];
}

template <bool BYSYS>
__global__ __launch_bounds__(kSortThreads) void k_radix_scatter(
    const uint64_t *__restrict__ keys, const int *__restrict__ vals, int n, DigitSrc src,
    int nblocks, const int *__restrict__ goff, uint64_t *__restrict__ keys_out,
    int *__restrict__ vals_out) {
  __shared__ uint64_t skeys[kSortTile];   // the tile in sorted order
  __shared__ int svals[kSortTile];
  __shared__ int tstart[256];     // first tile-local position of each digit
  __shared__ int gbase[256];      // global position of this tile's first pair of each digit
  __shared__ int wcnt[4][256];    // per-wave digit counts -> per-wave digit bases
  __shared__ int sh_wave[4];
  const int tid = threadIdx.x;
  const int w = tid >> 6;
  const int lane = lane_id();
  const int hist_n = 256 * nblocks;
  {
    // this tile's count of digit tid from the scanned histogram (digit-major:
    // the next entry is the same digit's next tile, or the next digit's first)
    const int idx = tid * nblocks + blockIdx.x;
    const int g0 = goff[idx];
    const int g1 = idx + 1 < hist_n ? goff[idx + 1] : n;
    int tot;
    const int incl = block_incl_scan<OpSum>(g1 - g0, sh_wave, &tot);
    tstart[tid] = incl - (g1 - g0);
    gbase[tid] = g0;
  }
#pragma unroll
  for (int k = 0; k < 4; k++) wcnt[k][tid] = 0;
  __syncthreads();
  // Each wave ranks its own contiguous quarter of the tile (16 rounds of 64
  // pairs) against its private digit counters; pairs stay in registers.
  const int wbase = blockIdx.x * kSortTile + w * (kSortRounds * kWave);
  uint64_t kk[kSortRounds];
  int vv[kSortRounds], dd[kSortRounds], rr[kSortRounds];
  // the wave's 16 rounds of pairs are all loaded before the ranking starts
#pragma unroll
  for (int r = 0; r < kSortRounds; r++) {
    const int i = wbase + r * kWave + lane;
    kk[r] = i < n ? keys[i] : 0ull;
    vv[r] = i < n ? vals[i] : 0;
  }
#pragma unroll
  for (int r = 0; r < kSortRounds; r++) {
    const int i = wbase + r * kWave + lane;
    const bool valid = i < n;
    const uint64_t key = kk[r];
    const int val = vv[r];
    const int d = valid ? digit_of<BYSYS>(key, val, src) : 0;
    uint64_t peers = ballot64(valid);
#pragma unroll
    for (int b = 0; b < 8; b++) {
      const bool bit = (d >> b) & 1;
      const uint64_t m = ballot64(bit);
      peers &= bit ? m : ~m;
    }
    const int rank = __popcll(peers & lanemask_lt());
    const int before = wcnt[w][d];
    wave_lds_sync();
    if (valid && rank == 0) wcnt[w][d] = before + __popcll(peers);
    wave_lds_sync();
    dd[r] = valid ? d : -1;
    rr[r] = before + rank;
  }
  __syncthreads();
  {
    // digit tid: the waves' pairs follow each other in tile order
    int acc = tstart[tid];
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const int c = wcnt[k][tid];
      wcnt[k][tid] = acc;
      acc += c;
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kSortRounds; r++) {
    if (dd[r] >= 0) {
      const int pos = wcnt[w][dd[r]] + rr[r];
      skeys[pos] = kk[r];
      svals[pos] = vv[r];
    }
  }
  __syncthreads();
  const int base = blockIdx.x * kSortTile;
  const int tile_n = min(kSortTile, n - base);
  for (int j = tid; j < tile_n; j += kSortThreads) {
    const uint64_t key = skeys[j];
    const int val = svals[j];
    const int d = digit_of<BYSYS>(key, val, src);
    const int dst = gbase[d] + (j - tstart[d]);
    keys_out[dst] = key;
    vals_out[dst] = val;
  }
}


template <class Op>
__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(const int *__restrict__ in, int n,
                                                              int *__restrict__ blk) {
  __shared__ int shw[4];
  const int base = blockIdx.x * kScanTile;
  int v = Op::id();
#pragma unroll
  for (int k = 0; k < kScanItems; k++) {
    const int i = base + k * kScanThreads + threadIdx.x;
    if (i < n) v = Op::op(v, in[i]);
  }
  int total;
  block_incl_scan<Op>(v, shw, &total);
  if (threadIdx.x == 0) blk[blockIdx.x] = total;
}

// Exclusive scan of the block totals in place (one block, any count).
template <class Op>
__global__ __launch_bounds__(kScanThreads) void k_scan_blocks(int *__restrict__ blk, int nb,
                                                              int *__restrict__ total_dev) {
  __shared__ int shw[4];
  int carry = Op::id();
  for (int b0 = 0; b0 < nb; b0 += kScanThreads) {
    const int i = b0 + threadIdx.x;
    const int v = i < nb ? blk[i] : Op::id();
    int total;
    const int inc = block_incl_scan<Op>(v, shw, &total);
    // exclusive = inclusive of the previous lane
    int exc = __shfl_up(inc, 1, kWave);
    const int w = threadIdx.x >> 6;
    if (lane_id() == 0) {
      exc = Op::id();
      for (int k = 0; k < w; k++) exc = Op::op(exc, shw[k]);
    }
    __syncthreads();
    if (i < nb) blk[i] = Op::op(carry, exc);
    carry = Op::op(carry, total);
    __syncthreads();
  }
  if (threadIdx.x == 0 && total_dev) *total_dev = carry;
}

template <class Op, bool kInclusive>
__global__ __launch_bounds__(kScanThreads) void k_scan_down(const int *__restrict__ in, int n,
                                                            const int *__restrict__ blk,
                                                            int *__restrict__ out) {
  __shared__ int sh[kScanTile];
  __shared__ int shw[4];
  const int base = blockIdx.x * kScanTile;
#pragma unroll
  for (int k = 0; k < kScanItems; k++) {
    const int i = base + k * kScanThreads + threadIdx.x;
    sh[k * kScanThreads + threadIdx.x] = i < n ? in[i] : Op::id();
  }
  __syncthreads();
  // thread t owns items [t*8, t*8+8) of the tile
  int vals[kScanItems];
  int run = Op::id();
#pragma unroll
  for (int k = 0; k < kScanItems; k++) {
    vals[k] = sh[threadIdx.x * kScanItems + k];
    run = Op::op(run, vals[k]);
  }
  int total;
  const int inc = block_incl_scan<Op>(run, shw, &total);
  int exc = __shfl_up(inc, 1, kWave);
  if (lane_id() == 0) {
    exc = Op::id();
    const int w = threadIdx.x >> 6;
    for (int k = 0; k < w; k++) exc = Op::op(exc, shw[k]);
  }
  int acc = Op::op(blk[blockIdx.x], exc);
#pragma unroll
  for (int k = 0; k < kScanItems; k++) {
    const int nv = Op::op(acc, vals[k]);
    sh[threadIdx.x * kScanItems + k] = kInclusive ? nv : acc;
    acc = nv;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kScanItems; k++) {
    const int i = base + k * kScanThreads + threadIdx.x;
    if (i < n) out[i] = sh[k * kScanThreads + threadIdx.x];
  }
}

__global__ void k_fill_int(int *p, int value, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    p[i] = value;
}

template <class Op, bool kInclusive>
void scan_impl(Ctx &c, const int *in, int *out, int n, int *total_dev) {
  if (n <= 0) {
    if (total_dev) fill_int(c, total_dev, 0, 1);
    return;
  }
  const int nb = (n + kScanTile - 1) / kScanTile;
  if (nb > c.scan_blk_cap) {
    if (c.scan_blk) (void)hipFree(c.scan_blk);
    c.scan_blk_cap = nb + 1024;
    if (hipMalloc(&c.scan_blk, sizeof(int) * c.scan_blk_cap) != hipSuccess) {
      c.scan_blk = nullptr;
      c.scan_blk_cap = 0;
      c.err = "scan scratch allocation failed";
      return;
    }
  }
  hipLaunchKernelGGL((k_scan_reduce<Op>), dim3(nb), dim3(kScanThreads), 0, c.stream, in, n,
                     c.scan_blk);
  hipLaunchKernelGGL((k_scan_blocks<Op>), dim3(1), dim3(kScanThreads), 0, c.stream, c.scan_blk,
                     nb, total_dev);
  hipLaunchKernelGGL((k_scan_down<Op, kInclusive>), dim3(nb), dim3(kScanThreads), 0, c.stream,
                     in, n, c.scan_blk, out);
}

}  // namespace

int grid_for(int n, int block, int cap) {
  int g = (n + block - 1) / block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return g;
}

void fill_int(Ctx &c, int *p, int value, int n) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill_int, dim3(grid_for(n, 256, 4096)), dim3(256), 0, c.stream, p, value,
                     n);
}

void scan_exclusive_sum(Ctx &c, const int *in, int *out, int n, int *total_dev) {
  scan_impl<OpSum, false>(c, in, out, n, total_dev);
}

void scan_inclusive_max(Ctx &c, const int *in, int *out, int n) {
  scan_impl<OpMax, true>(c, in, out, n, nullptr);
}

// Sorts (keys, vals)[0..n) stably by the low key_bits bits of the key; when
// by_system_after is set, one or two more stable passes order by the system
// of each value (systems of a forest stay contiguous).  The result is left in
// keys/vals; keys_tmp/vals_tmp are scratch of the same size.
void radix_sort_pairs(Ctx &c, uint64_t *keys, int *vals, uint64_t *keys_tmp, int *vals_tmp,
                      int n, int key_bits, bool by_system_after) {
  if (n <= 1) return;
  const int nblocks = (n + kSortTile - 1) / kSortTile;
  const int hist_n = 256 * nblocks;
  if (hist_n > c.sort_hist_cap) {
    if (c.sort_hist) (void)hipFree(c.sort_hist);
    c.sort_hist_cap = hist_n;
    if (hipMalloc(&c.sort_hist, sizeof(int) * hist_n) != hipSuccess) {
      c.sort_hist = nullptr;
      c.sort_hist_cap = 0;
      c.err = "sort histogram allocation failed";
      return;
    }
  }
  struct Pass { int shift, by_sys; };
  std::vector<Pass> passes;
  for (int s = 0; s < key_bits; s += 8) passes.push_back({s, 0});
  if (by_system_after && c.nsys > 1) {
    passes.push_back({0, 1});
    if (c.nsys > 256) passes.push_back({8, 1});
    if (c.nsys > 65536) passes.push_back({16, 1});
  }
  uint64_t *ks = keys, *kd = keys_tmp;
  int *vs = vals, *vd = vals_tmp;
  for (const Pass &p : passes) {
    DigitSrc src{p.shift, c.nsys, c.d_off};
    auto hist = p.by_sys ? k_radix_hist<true> : k_radix_hist<false>;
    auto scatter = p.by_sys ? k_radix_scatter<true> : k_radix_scatter<false>;
    hipLaunchKernelGGL(hist, dim3(nblocks), dim3(kSortThreads), 0, c.stream, ks, vs, n, src,
                       nblocks, c.sort_hist);
    scan_exclusive_sum(c, c.sort_hist, c.sort_hist, hist_n, nullptr);
    hipLaunchKernelGGL(scatter, dim3(nblocks), dim3(kSortThreads), 0, c.stream, ks, vs, n, src,
                       nblocks, c.sort_hist, kd, vd);
    std::swap(ks, kd);
    std::swap(vs, vd);
  }
  if (ks != keys) {
    (void)hipMemcpyAsync(keys, ks, sizeof(uint64_t) * n, hipMemcpyDeviceToDevice, c.stream);
    (void)hipMemcpyAsync(vals, vs, sizeof(int) * n, hipMemcpyDeviceToDevice, c.stream);
  }
}

}  // namespace bt
