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
// The three-launch passes run the same 4096-pair tile with 512 threads (8
// waves of 8 rounds): twice the waves per CU for the same LDS.
constexpr int kCThreads = 512;
constexpr int kCRounds = kSortTile / kCThreads;
constexpr int kCWaves = kCThreads / kWave;
static_assert(kCThreads * kCRounds == kSortTile, "tile of the three-launch passes");
constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;
// up to this many tiles a scan takes two launches (the down-sweep reduces the
// tile totals before it, at most 1024 ints per block); beyond, three
constexpr int kScanFusedTiles = 1024;

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
__global__ __launch_bounds__(kCThreads) void k_radix_hist(const uint64_t *__restrict__ keys,
                                                          const int *__restrict__ vals,
                                                          int n, DigitSrc src, int nblocks,
                                                          int *__restrict__ ghist) {
  BT_BUILD_PRIO();
  __shared__ int hist[256];
  const int tid = threadIdx.x;
  if (tid < 256) hist[tid] = 0;
  const int base = blockIdx.x * kSortTile;
  // all of the tile's loads first, then the LDS counting
  int dg[kCRounds];
#pragma unroll
  for (int r = 0; r < kCRounds; r++) {
    const int i = base + r * kCThreads + tid;
    const int ic = min(i, n - 1);   // unconditional loads: no branch between them
    const int d = digit_of<BYSYS>(BYSYS ? 0ull : keys[ic], BYSYS ? vals[ic] : 0, src);
    dg[r] = i < n ? d : -1;
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kCRounds; r++)
    if (dg[r] >= 0) atomicAdd(&hist[dg[r]], 1);
  __syncthreads();
  if (tid < 256) ghist[tid * nblocks + blockIdx.x] = hist[tid];
}

template <bool BYSYS>
__global__ __launch_bounds__(kCThreads) void k_radix_scatter(
    const uint64_t *__restrict__ keys, const int *__restrict__ vals, int n, DigitSrc src,
    int nblocks, const int *__restrict__ goff, uint64_t *__restrict__ keys_out,
    int *__restrict__ vals_out) {
  BT_BUILD_PRIO();
  // LDS kept under a third of a CU's 160 KiB, so three blocks share a CU:
  // 16-bit per-wave digit counters (a tile holds 4096 pairs) and, per digit,
  // only (global base - tile-local start)
  __shared__ uint64_t skeys[kSortTile];   // the tile in sorted order
  __shared__ int svals[kSortTile];
  __shared__ int gdelta[256];     // global position of digit d's pairs minus their tile start
  __shared__ unsigned short wcnt[kCWaves][256];   // per-wave digit counts -> digit bases
  __shared__ int sh_wave[kCWaves];
  const int tid = threadIdx.x;
  const int w = tid >> 6;
  const int lane = lane_id();
  const int hist_n = 256 * nblocks;
  int tstart = 0;   // thread d < 256: first tile-local position of digit d
  {
    // thread d < 256: this tile's count of digit d from the scanned histogram
    // (digit-major: the next entry is the same digit's next tile, or the next
    // digit's first); every thread joins the block scan
    int g0 = 0, cnt = 0;
    if (tid < 256) {
      const int idx = tid * nblocks + blockIdx.x;
      g0 = goff[idx];
      const int g1 = idx + 1 < hist_n ? goff[idx + 1] : n;
      cnt = g1 - g0;
    }
    int tot;
    const int incl = block_incl_scan<OpSum>(cnt, sh_wave, &tot);
    if (tid < 256) {
      tstart = incl - cnt;
      gdelta[tid] = g0 - tstart;
    }
  }
  for (int k = tid; k < kCWaves * 256; k += kCThreads) (&wcnt[0][0])[k] = 0;
  __syncthreads();
  // Each wave ranks its own contiguous eighth of the tile (8 rounds of 64
  // pairs) against its private digit counters; pairs stay in registers.
  const int wbase = blockIdx.x * kSortTile + w * (kCRounds * kWave);
  uint64_t kk[kCRounds];
  int vv[kCRounds], dd[kCRounds], rr[kCRounds];
  // the wave's rounds of pairs are all loaded before the ranking starts
#pragma unroll
  for (int r = 0; r < kCRounds; r++) {
    const int i = wbase + r * kWave + lane;
    kk[r] = i < n ? keys[i] : 0ull;
    vv[r] = i < n ? vals[i] : 0;
  }
#pragma unroll
  for (int r = 0; r < kCRounds; r++) {
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
    if (valid && rank == 0) wcnt[w][d] = (unsigned short)(before + __popcll(peers));
    wave_lds_sync();
    dd[r] = valid ? d : -1;
    rr[r] = before + rank;
  }
  __syncthreads();
  if (tid < 256) {
    // digit tid: the waves' pairs follow each other in tile order
    int acc = tstart;
#pragma unroll
    for (int k = 0; k < kCWaves; k++) {
      const int c = wcnt[k][tid];
      wcnt[k][tid] = (unsigned short)acc;
      acc += c;
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kCRounds; r++) {
    if (dd[r] >= 0) {
      const int pos = wcnt[w][dd[r]] + rr[r];
      skeys[pos] = kk[r];
      svals[pos] = vv[r];
    }
  }
  __syncthreads();
  const int base = blockIdx.x * kSortTile;
  const int tile_n = min(kSortTile, n - base);
  for (int j = tid; j < tile_n; j += kCThreads) {
    const uint64_t key = skeys[j];
    const int val = svals[j];
    const int d = digit_of<BYSYS>(key, val, src);
    const int dst = gdelta[d] + j;
    keys_out[dst] = key;
    vals_out[dst] = val;
  }
}

// ---- one-sweep passes ------------------------------------------------------
// The digit counts of every pass come from one histogram launch over the
// input (the digits of a key do not change while it moves between passes).
// Each pass is then one launch: a block takes the next tile from a counter,
// ranks its pairs as above, publishes its per-digit counts and finds the
// counts of all earlier tiles by looking back through their status words
// (decoupled look-back), then writes its pairs to their sorted positions.
// Per pass: 12 B read + 12 B written per pair.
//
// A status word is (flag << 30) | count, flag 1 = this tile's count only,
// 2 = the inclusive count of tiles 0..t.  A block waits only for tiles that
// were handed out before its own, whose blocks are already running and never
// wait for later tiles, so the wait ends; it is bounded anyway and reports
// kSortErr instead of hanging.  Two status buffers and two tile counters
// alternate between passes: a pass clears the other buffer's words of its
// tile and resets the other counter for the next pass.

constexpr int kMaxSortPasses = 11;
// One-sweep tiles: 256 threads x 8 rounds = 2048 pairs (30 KiB of LDS, so
// five blocks share a CU while others wait in their look-back).
constexpr int kOsRounds = 8;
constexpr int kOsTile = kSortThreads * kOsRounds;
constexpr int kSortHistBlocks = 1024;
constexpr unsigned kStAgg = 1u << 30, kStIncl = 2u << 30, kStCount = (1u << 30) - 1u;
constexpr int kLookbackSpins = 1 << 22;

struct PassSet {
  int np;
  int nsys;
  const int *off;
  int shift[kMaxSortPasses];
  int bysys[kMaxSortPasses];
};

// Adds 1 for every valid lane at hist[d].  A wave whose valid lanes all hold
// one digit (the high digits of nearly sorted keys) adds its count with one
// atomic; otherwise each lane adds its own.
__device__ __forceinline__ void wave_count_digit(int *hist, int d, bool valid) {
  const uint64_t vm = ballot64(valid);
  if (vm == 0) return;   // wave-uniform
  const int first = __ffsll((unsigned long long)vm) - 1;
  const int d0 = __shfl(d, first, kWave);
  const uint64_t same = ballot64(valid && d == d0);
  if (same == vm) {
    if (lane_id() == first) atomicAdd(&hist[d0], __popcll(vm));
  } else if (valid) {
    atomicAdd(&hist[d], 1);
  }
}

// grid kSortHistBlocks: block b counts the digits of every pass over a
// contiguous chunk; part[b][p][256].
__global__ __launch_bounds__(kSortThreads) void k_radix_hist_all(const uint64_t *__restrict__ keys,
                                                                 const int *__restrict__ vals,
                                                                 int n, PassSet ps,
                                                                 int *__restrict__ part,
                                                                 DevCounters *dc) {
  __shared__ int hist[kMaxSortPasses * 256];
  // the sort error bit reports this sort only (err_sticky keeps it for the step)
  if (blockIdx.x == 0 && threadIdx.x == 0) atomicAnd(&dc->err, ~kSortErr);
  for (int k = threadIdx.x; k < ps.np * 256; k += kSortThreads) hist[k] = 0;
  __syncthreads();
  const int chunk = (n + gridDim.x - 1) / gridDim.x;
  const int i0 = blockIdx.x * chunk, i1 = min(n, i0 + chunk);
  for (int j0 = i0; j0 < i1; j0 += kSortThreads) {
    const int i = j0 + threadIdx.x;
    const bool valid = i < i1;
    const uint64_t key = valid ? keys[i] : 0ull;
    const int val = valid ? vals[i] : 0;
    for (int p = 0; p < ps.np; p++) {
      const DigitSrc src{ps.shift[p], ps.nsys, ps.off};
      const int d = ps.bysys[p] ? digit_of<true>(key, val, src) : digit_of<false>(key, val, src);
      wave_count_digit(hist + p * 256, d, valid);
    }
  }
  __syncthreads();
  for (int k = threadIdx.x; k < ps.np * 256; k += kSortThreads)
    part[(size_t)blockIdx.x * ps.np * 256 + k] = hist[k];
}

// grid (passes, 8): block (p, g) sums digits [32g, 32g + 32) of pass p over
// the histogram blocks, 8 segments of blocks per digit in parallel.
__global__ __launch_bounds__(256) void k_radix_hist_sum(const int *__restrict__ part, int np,
                                                        int nparts, int *__restrict__ gcount) {
  __shared__ int sh[8][32];
  const int p = blockIdx.x;
  const int d = blockIdx.y * 32 + (threadIdx.x & 31);
  const int seg = threadIdx.x >> 5;
  const int per = (nparts + 7) / 8;
  const int b0 = seg * per, b1 = min(nparts, b0 + per);
  int s = 0;
#pragma unroll 8
  for (int b = b0; b < b1; b++) s += part[(size_t)b * np * 256 + p * 256 + d];
  sh[seg][threadIdx.x & 31] = s;
  __syncthreads();
  if (threadIdx.x < 32) {
    int t = 0;
    for (int k = 0; k < 8; k++) t += sh[k][threadIdx.x];
    gcount[p * 256 + d] = t;
  }
}

template <bool BYSYS>
__global__ __launch_bounds__(kSortThreads) void k_radix_onesweep(
    const uint64_t *__restrict__ keys, const int *__restrict__ vals, int n, DigitSrc src,
    const int *__restrict__ gcount, unsigned int *status, unsigned int *status_next, int *ctr,
    int *ctr_next, int ntiles, DevCounters *dc, uint64_t *__restrict__ keys_out,
    int *__restrict__ vals_out) {
  __shared__ uint64_t skeys[kOsTile];   // the tile in sorted order
  __shared__ int svals[kOsTile];
  __shared__ int tstart[256];     // first tile-local position of each digit
  __shared__ int gbase[256];      // global position of this tile's first pair of each digit
  __shared__ int wcnt[4][256];    // per-wave digit counts -> per-wave digit bases
  __shared__ int sh_wave[4];
  __shared__ int s_tile;
  const int tid = threadIdx.x;
  const int w = tid >> 6;
  const int lane = lane_id();
  if (tid == 0) s_tile = atomicAdd(ctr, 1);
#pragma unroll
  for (int k = 0; k < 4; k++) wcnt[k][tid] = 0;
  __syncthreads();
  const int tile = s_tile;
  if (tile >= ntiles) {   // a counter not reset: never index past the status words
    if (tid == 0) {
      atomicOr(&dc->err, kSortErr);
      atomicOr(&dc->err_sticky, kSortErr);
    }
    return;
  }
  if (tile == 0 && tid == 0) *ctr_next = 0;
  status_next[(size_t)tile * 256 + tid] = 0u;
  // Each wave ranks its own contiguous quarter of the tile (16 rounds of 64
  // pairs) against its private digit counters; pairs stay in registers.
  const int wbase = tile * kOsTile + w * (kOsRounds * kWave);
  uint64_t kk[kOsRounds];
  int vv[kOsRounds], dd[kOsRounds], rr[kOsRounds];
#pragma unroll
  for (int r = 0; r < kOsRounds; r++) {
    const int i = wbase + r * kWave + lane;
    kk[r] = i < n ? keys[i] : 0ull;
    vv[r] = i < n ? vals[i] : 0;
  }
#pragma unroll
  for (int r = 0; r < kOsRounds; r++) {
    const int i = wbase + r * kWave + lane;
    const bool valid = i < n;
    const int d = valid ? digit_of<BYSYS>(kk[r], vv[r], src) : 0;
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
  // digit tid: this tile's count, published before anything else is done
  const int cnt = wcnt[0][tid] + wcnt[1][tid] + wcnt[2][tid] + wcnt[3][tid];
  unsigned int *mine = status + (size_t)tile * 256 + tid;
  __hip_atomic_store(mine, (tile == 0 ? kStIncl : kStAgg) | (unsigned)cnt, __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
  // tile-local start of each digit, and the global start of each digit
  int tot;
  const int incl = block_incl_scan<OpSum>(cnt, sh_wave, &tot);
  tstart[tid] = incl - cnt;
  const int g = gcount[tid];
  const int gincl = block_incl_scan<OpSum>(g, sh_wave, &tot);
  // look back over the earlier tiles until one holds an inclusive count
  int excl = 0;
  if (tile > 0) {
    int j = tile - 1;
    int spins = 0;
    for (;;) {
      const unsigned v = __hip_atomic_load(status + (size_t)j * 256 + tid, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
      if ((v & ~kStCount) == 0u) {
        if (++spins > kLookbackSpins) {
          atomicOr(&dc->err, kSortErr);
          atomicOr(&dc->err_sticky, kSortErr);
          break;
        }
        __builtin_amdgcn_s_sleep(1);
        continue;
      }
      excl += (int)(v & kStCount);
      if ((v & ~kStCount) == kStIncl || j == 0) break;
      j--;
    }
    __hip_atomic_store(mine, kStIncl | (unsigned)(excl + cnt), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  }
  gbase[tid] = gincl - g + excl;
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
  for (int r = 0; r < kOsRounds; r++) {
    if (dd[r] >= 0) {
      const int pos = wcnt[w][dd[r]] + rr[r];
      skeys[pos] = kk[r];
      svals[pos] = vv[r];
    }
  }
  __syncthreads();
  const int tile_n = min(kOsTile, n - tile * kOsTile);
  for (int j = tid; j < tile_n; j += kSortThreads) {
    const uint64_t key = skeys[j];
    const int val = svals[j];
    const int d = digit_of<BYSYS>(key, val, src);
    const int dst = gbase[d] + (j - tstart[d]);
    keys_out[dst] = key;
    vals_out[dst] = val;
  }
}

__global__ void k_sort_clear(unsigned int *status, int nwords, int *ctr) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nwords; i += gridDim.x * blockDim.x)
    status[i] = 0u;
  if (blockIdx.x == 0 && threadIdx.x == 0) *ctr = 0;
}


template <class Op>
__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(const int *__restrict__ in, int n,
                                                              int *__restrict__ blk) {
  BT_BUILD_PRIO();
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
  BT_BUILD_PRIO();
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

// kPrefix: blk holds the tile totals and each block reduces those of the
// tiles before it itself (for few tiles: one launch instead of a separate
// scan of the totals); the last block writes the grand total to total_dev.
// Otherwise blk holds the tiles' exclusive prefixes (k_scan_blocks).
template <class Op, bool kInclusive, bool kPrefix>
__global__ __launch_bounds__(kScanThreads) void k_scan_down(const int *__restrict__ in, int n,
                                                            const int *__restrict__ blk,
                                                            int *__restrict__ out,
                                                            int *__restrict__ total_dev) {
  BT_BUILD_PRIO();
  __shared__ int sh[kScanTile];
  __shared__ int shw[4];
  const int base = blockIdx.x * kScanTile;
  int prefix;
  if constexpr (kPrefix) {
    int v = Op::id();
    for (int k = threadIdx.x; k < (int)blockIdx.x; k += kScanThreads) v = Op::op(v, blk[k]);
    (void)block_incl_scan<Op>(v, shw, &prefix);   // prefix = the reduction over the block
  } else {
    prefix = blk[blockIdx.x];
  }
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
  int acc = Op::op(prefix, exc);
#pragma unroll
  for (int k = 0; k < kScanItems; k++) {
    const int nv = Op::op(acc, vals[k]);
    sh[threadIdx.x * kScanItems + k] = kInclusive ? nv : acc;
    acc = nv;
  }
  if (kPrefix && total_dev && threadIdx.x == 0 && blockIdx.x == gridDim.x - 1)
    *total_dev = Op::op(prefix, total);
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
  if (nb <= kScanFusedTiles) {
    hipLaunchKernelGGL((k_scan_down<Op, kInclusive, true>), dim3(nb), dim3(kScanThreads), 0,
                       c.stream, in, n, c.scan_blk, out, total_dev);
    return;
  }
  hipLaunchKernelGGL((k_scan_blocks<Op>), dim3(1), dim3(kScanThreads), 0, c.stream, c.scan_blk,
                     nb, total_dev);
  hipLaunchKernelGGL((k_scan_down<Op, kInclusive, false>), dim3(nb), dim3(kScanThreads), 0,
                     c.stream, in, n, c.scan_blk, out, nullptr);
}

template <typename T>
void dev_free_raw(T *&p) {
  if (p) (void)hipFree(p);
  p = nullptr;
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
  if (c.sort_onesweep && (int)passes.size() <= kMaxSortPasses) {
    const int np = (int)passes.size();
    const int ntiles = (n + kOsTile - 1) / kOsTile;
    if (ntiles > c.sort_tiles_cap || !c.sort_hpart) {
      dev_free_raw(c.sort_hpart);
      dev_free_raw(c.sort_gcount);
      dev_free_raw(c.sort_status);
      dev_free_raw(c.sort_ctr);
      c.sort_tiles_cap = 0;
      if (hipMalloc(&c.sort_hpart, sizeof(int) * kSortHistBlocks * kMaxSortPasses * 256) ||
          hipMalloc(&c.sort_gcount, sizeof(int) * kMaxSortPasses * 256) ||
          hipMalloc(&c.sort_status, sizeof(unsigned int) * 2 * 256 * (size_t)ntiles) ||
          hipMalloc(&c.sort_ctr, sizeof(int) * 2)) {
        c.err = "sort scratch allocation failed";
        return;
      }
      c.sort_tiles_cap = ntiles;
      hipLaunchKernelGGL(k_sort_clear, dim3(1024), dim3(256), 0, c.stream, c.sort_status,
                         2 * 256 * ntiles, c.sort_ctr);
      hipLaunchKernelGGL(k_sort_clear, dim3(1), dim3(64), 0, c.stream, c.sort_status, 0,
                         c.sort_ctr + 1);
    }
    PassSet ps;
    ps.np = np;
    ps.nsys = c.nsys;
    ps.off = c.d_off;
    for (int k = 0; k < np; k++) {
      ps.shift[k] = passes[k].shift;
      ps.bysys[k] = passes[k].by_sys;
    }
    hipLaunchKernelGGL(k_radix_hist_all, dim3(kSortHistBlocks), dim3(kSortThreads), 0, c.stream,
                       keys, vals, n, ps, c.sort_hpart, c.dc);
    hipLaunchKernelGGL(k_radix_hist_sum, dim3(np, 8), dim3(256), 0, c.stream, c.sort_hpart, np,
                       kSortHistBlocks, c.sort_gcount);
    // pass k uses status buffer k & 1; every sort starts on a clean buffer 0
    const size_t words = (size_t)256 * c.sort_tiles_cap;
    for (int k = 0; k < np; k++) {
      const int cur = k & 1, nxt = cur ^ 1;
      DigitSrc src{passes[k].shift, c.nsys, c.d_off};
      auto kern = passes[k].by_sys ? k_radix_onesweep<true> : k_radix_onesweep<false>;
      hipLaunchKernelGGL(kern, dim3(ntiles), dim3(kSortThreads), 0, c.stream, ks, vs, n, src,
                         c.sort_gcount + 256 * k, c.sort_status + cur * words,
                         c.sort_status + nxt * words, c.sort_ctr + cur, c.sort_ctr + nxt,
                         ntiles, c.dc, kd, vd);
      std::swap(ks, kd);
      std::swap(vs, vd);
    }
    if (np & 1)   // the last pass used buffer 0: clear it for the next sort
      hipLaunchKernelGGL(k_sort_clear, dim3(grid_for(ntiles, 1, 1024)), dim3(256), 0, c.stream,
                         c.sort_status, 256 * ntiles, c.sort_ctr);
  } else {
  for (const Pass &p : passes) {
    DigitSrc src{p.shift, c.nsys, c.d_off};
    auto hist = p.by_sys ? k_radix_hist<true> : k_radix_hist<false>;
    auto scatter = p.by_sys ? k_radix_scatter<true> : k_radix_scatter<false>;
    hipLaunchKernelGGL(hist, dim3(nblocks), dim3(kCThreads), 0, c.stream, ks, vs, n, src,
                       nblocks, c.sort_hist);
    scan_exclusive_sum(c, c.sort_hist, c.sort_hist, hist_n, nullptr);
    hipLaunchKernelGGL(scatter, dim3(nblocks), dim3(kCThreads), 0, c.stream, ks, vs, n, src,
                       nblocks, c.sort_hist, kd, vd);
    std::swap(ks, kd);
    std::swap(vs, vd);
  }
  }
  if (ks != keys) {
    (void)hipMemcpyAsync(keys, ks, sizeof(uint64_t) * n, hipMemcpyDeviceToDevice, c.stream);
    (void)hipMemcpyAsync(vals, vs, sizeof(int) * n, hipMemcpyDeviceToDevice, c.stream);
  }
}

}  // namespace bt
