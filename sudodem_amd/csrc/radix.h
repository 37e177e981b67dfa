// radix.h -- device-wide scans and a stable LSD radix sort for CDNA (wave64).
//
// The broadphase keeps, per axis, the order of the 2N bound endpoints from the previous
// collider run.  The reference's insertion sort of that order (InsertionSortCollider.cpp:29-54)
// yields the stable sort of the old order by the new coordinates, which a stable LSD radix
// sort computes directly.  Tiles are 256 threads x 8 keys; inside a tile, keys are ranked
// stably per 8-bit digit with 64-lane ballots (one match mask per wave) and a per-wave
// digit-count prefix in LDS, so no global atomics decide the order.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sdmrs {

constexpr int kThreads = 256;
constexpr int kItems = 8;
constexpr int kTile = kThreads * kItems;
constexpr int kBins = 256;
constexpr int kWaves = kThreads / 64;

struct OpAdd {
  static __device__ __forceinline__ uint32_t id() { return 0u; }
  static __device__ __forceinline__ uint32_t op(uint32_t a, uint32_t b) { return a + b; }
};
struct OpMin {
  static __device__ __forceinline__ uint32_t id() { return 0xFFFFFFFFu; }
  static __device__ __forceinline__ uint32_t op(uint32_t a, uint32_t b) { return a < b ? a : b; }
};

// Block-wide inclusive scan of one value per thread (256 threads).
template <class Op>
__device__ __forceinline__ uint32_t block_inclusive(uint32_t v, uint32_t* sh) {
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int off = 1; off < kThreads; off <<= 1) {
    uint32_t o = t >= off ? sh[t - off] : Op::id();
    __syncthreads();
    v = Op::op(o, v);
    sh[t] = v;
    __syncthreads();
  }
  return v;
}

// Scan of one tile.  Element j of the logical sequence is in[rev ? n-1-j : j].  Writes the
// exclusive (or inclusive) scan to out at the same physical index and the tile total to sums.
template <class Op, bool kInclusive, bool kRev>
__global__ __launch_bounds__(kThreads) void k_scan_tile(const uint32_t* in, uint32_t* out, int64_t n,
                                                        uint32_t* sums) {
  __shared__ uint32_t sh[kThreads];
  const int64_t base = (int64_t)blockIdx.x * kTile + (int64_t)threadIdx.x * kItems;
  uint32_t v[kItems];
  uint32_t acc = Op::id();
#pragma unroll
  for (int i = 0; i < kItems; i++) {
    int64_t j = base + i;
    uint32_t x = Op::id();
    if (j < n) x = in[kRev ? (n - 1 - j) : j];
    v[i] = x;
    acc = Op::op(acc, x);
  }
  uint32_t incl = block_inclusive<Op>(acc, sh);
  uint32_t run = threadIdx.x ? sh[threadIdx.x - 1] : Op::id();
  __syncthreads();
#pragma unroll
  for (int i = 0; i < kItems; i++) {
    int64_t j = base + i;
    uint32_t nxt = Op::op(run, v[i]);
    if (j < n) out[kRev ? (n - 1 - j) : j] = kInclusive ? nxt : run;
    run = nxt;
  }
  if (threadIdx.x == kThreads - 1 && sums) sums[blockIdx.x] = incl;
}

// Combine tile results with the scanned tile sums (exclusive scan of sums in `pre`).
template <class Op, bool kRev>
__global__ __launch_bounds__(kThreads) void k_scan_fix(uint32_t* out, int64_t n, const uint32_t* pre) {
  const int64_t j = (int64_t)blockIdx.x * kTile + threadIdx.x;
  const uint32_t p = pre[blockIdx.x];
#pragma unroll
  for (int i = 0; i < kItems; i++) {
    int64_t jj = j + (int64_t)i * kThreads;
    if (jj < n) {
      int64_t k = kRev ? (n - 1 - jj) : jj;
      out[k] = Op::op(p, out[k]);
    }
  }
}

// Scratch size (in uint32) needed by scan() for n elements.
inline int64_t scan_scratch(int64_t n) {
  int64_t s = 0;
  while (n > kTile) {
    n = (n + kTile - 1) / kTile;
    s += 2 * n + 2;
  }
  return s + 2;
}

// Device-wide scan; `scratch` holds scan_scratch(n) words.
template <class Op, bool kInclusive, bool kRev>
void scan(const uint32_t* in, uint32_t* out, int64_t n, uint32_t* scratch, hipStream_t st) {
  if (n <= 0) return;
  int64_t nb = (n + kTile - 1) / kTile;
  if (nb == 1) {
    hipLaunchKernelGGL((k_scan_tile<Op, kInclusive, kRev>), dim3(1), dim3(kThreads), 0, st, in, out, n,
                       (uint32_t*)nullptr);
    return;
  }
  uint32_t* sums = scratch;
  uint32_t* pre = scratch + nb + 1;
  hipLaunchKernelGGL((k_scan_tile<Op, kInclusive, kRev>), dim3(nb), dim3(kThreads), 0, st, in, out, n, sums);
  // exclusive scan of the tile sums, in tile order (the tile order already follows kRev)
  scan<Op, false, false>(sums, pre, nb, scratch + 2 * nb + 2, st);
  hipLaunchKernelGGL((k_scan_fix<Op, kRev>), dim3(nb), dim3(kThreads), 0, st, out, n, pre);
}

template <typename K>
__global__ __launch_bounds__(kThreads) void k_rs_hist(const K* keys, int64_t n, int shift, uint32_t* hist,
                                                      int nblocks) {
  __shared__ uint32_t h[kBins];
  h[threadIdx.x] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * kTile;
#pragma unroll
  for (int r = 0; r < kItems; r++) {
    int64_t i = base + (int64_t)r * kThreads + threadIdx.x;
    if (i < n) atomicAdd(&h[(uint32_t)(keys[i] >> shift) & (kBins - 1)], 1u);
  }
  __syncthreads();
  hist[(int64_t)threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

template <typename K, typename V>
__global__ __launch_bounds__(kThreads) void k_rs_scatter(const K* kin, const V* vin, K* kout, V* vout, int64_t n,
                                                         int shift, const uint32_t* off, int nblocks) {
  __shared__ uint32_t base_off[kBins];
  __shared__ uint32_t wcnt[kWaves][kBins];
  const int t = threadIdx.x;
  const int lane = t & 63;
  const int wave = t >> 6;
  base_off[t] = off[(int64_t)t * nblocks + blockIdx.x];
  const uint64_t below_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  for (int r = 0; r < kItems; r++) {
    const int64_t i = (int64_t)blockIdx.x * kTile + (int64_t)r * kThreads + t;
    const bool valid = i < n;
    K k = valid ? kin[i] : K(0);
    const uint32_t d = valid ? ((uint32_t)(k >> shift) & (kBins - 1)) : 0u;
#pragma unroll
    for (int w = 0; w < kWaves; w++) wcnt[w][t] = 0;
    __syncthreads();
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; b++) {
      const bool bit = (d >> b) & 1u;
      const uint64_t m = __ballot(bit);
      peers &= bit ? m : ~m;
    }
    const uint64_t below = peers & below_mask;
    const uint32_t rank = (uint32_t)__popcll(below);
    if (valid && below == 0) wcnt[wave][d] = (uint32_t)__popcll(peers);
    __syncthreads();
    uint32_t c[kWaves];
    uint32_t tot = 0;
#pragma unroll
    for (int w = 0; w < kWaves; w++) {
      c[w] = wcnt[w][t];
    }
    __syncthreads();
#pragma unroll
    for (int w = 0; w < kWaves; w++) {
      wcnt[w][t] = tot;
      tot += c[w];
    }
    __syncthreads();
    if (valid) {
      const uint32_t pos = base_off[d] + wcnt[wave][d] + rank;
      kout[pos] = k;
      vout[pos] = vin[i];
    }
    __syncthreads();
    base_off[t] += tot;
    __syncthreads();
  }
}

// Scratch words needed by sort_pairs for n elements.
inline int64_t sort_scratch(int64_t n) {
  int64_t nb = (n + kTile - 1) / kTile;
  int64_t h = nb * kBins;
  return 2 * h + scan_scratch(h) + 16;
}

// Stable LSD sort of (key, value) by the low `bits` bits of the key.  The result is left in
// (k0, v0); (k1, v1) are buffers of the same size.  Returns nothing; stream ordered.
template <typename K, typename V>
void sort_pairs(K* k0, V* v0, K* k1, V* v1, int64_t n, int bits, uint32_t* scratch, hipStream_t st) {
  if (n <= 1) return;
  const int nb = (int)((n + kTile - 1) / kTile);
  const int64_t h = (int64_t)nb * kBins;
  uint32_t* hist = scratch;
  uint32_t* offs = scratch + h;
  uint32_t* sc = scratch + 2 * h;
  int passes = (bits + 7) / 8;
  K* ka = k0; V* va = v0; K* kb = k1; V* vb = v1;
  for (int p = 0; p < passes; p++) {
    const int shift = 8 * p;
    hipLaunchKernelGGL((k_rs_hist<K>), dim3(nb), dim3(kThreads), 0, st, ka, n, shift, hist, nb);
    scan<OpAdd, false, false>(hist, offs, h, sc, st);
    hipLaunchKernelGGL((k_rs_scatter<K, V>), dim3(nb), dim3(kThreads), 0, st, ka, va, kb, vb, n, shift, offs, nb);
    K* tk = ka; ka = kb; kb = tk;
    V* tv = va; va = vb; vb = tv;
  }
  if (ka != k0) {
    (void)hipMemcpyAsync(k0, ka, n * sizeof(K), hipMemcpyDeviceToDevice, st);
    (void)hipMemcpyAsync(v0, va, n * sizeof(V), hipMemcpyDeviceToDevice, st);
  }
}

}  // namespace sdmrs
