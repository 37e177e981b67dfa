// ring_dev.h -- device kernels of the periodic collider's ring walk (ring.h): the run's
// coordinates, the sequential phases (one lane per axis), and the parallel phases (window
// checks, sort keys, crossings, the permutation).
#pragma once
#include "periodic.h"
#include "ring.h"

// error bits per axis
enum {
  RING_ERR_WIN1 = 1,       // the first parallel phase's conditions failed
  RING_ERR_WIN2 = 2,       // the second's
  RING_ERR_RUN = 64,       // an insertion ran once round the ring (non-finite coordinates)
  RING_ERR_NONFINITE = 128 // a bound coordinate is not finite
};

struct RingSet {
  int32_t* h[3];
  int32_t* per[3];
  double* x[3];
  int64_t n2;
  double L[3];
  int64_t* lo;    // [3] loIdx, read and written by the sequential phases
  uint32_t* err;  // [3] error bits
};

// Positions of one phase per axis.
struct RingSpan {
  int64_t s[3], e[3];
};

// One parallel phase per axis: the elements of [i0, i1) are inserted; the window starts at ws
// (the sort covers [ws, i1)); r0 (device) is where the sorted part of the prefix begins.
struct RingWin {
  int64_t i0[3], i1[3], ws[3];
  int64_t* r0;      // [3] device
  int64_t* r0base;  // [3] device
  uint64_t* minKey; // [3] device: coord_key of the smallest inserted coordinate
  int bit;          // RING_ERR_WIN1 / RING_ERR_WIN2
};

// The run's coordinate of every ring element (InsertionSortCollider.cpp:289-311):
// coord = (bound, or the position of a body without bounds) - L * period.
__global__ void k_ring_coords(BodyArrays B, RingSet R) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= 3 * R.n2) return;
  const int a = (int)(g / R.n2);
  const int64_t k = g - a * R.n2;
  const int32_t h = R.h[a][k];
  const int id = h >> 1;
  const int64_t N = B.n;
  const double b = B.hasBound[id] ? ((h & 1) ? B.bmax[a * N + id] : B.bmin[a * N + id]) : B.pos[a * N + id];
  const double x = b - R.L[a] * (double)R.per[a][k];
  R.x[a][k] = x;
  if (!isfinite(x)) atomicOr(&R.err[a], (uint32_t)RING_ERR_NONFINITE);
}

// handleBoundInversionPeri (:466-475) up to the container lookup: the pair is a candidate
// when spatialOverlapPeri holds; its order key is (axis, position inserted, rank of the
// inversion within that insertion).
struct RingEmit {
  BodyArrays B;
  PeriCell pc;
  Cands C;
  unsigned long long* err;
  uint64_t axisKey;
  __device__ void operator()(int64_t i, uint32_t step, int32_t a, int32_t b) {
    int32_t per[3];
    if (!spatial_overlap_peri(B, pc, a, b, per, err)) return;
    push_cand(C, a, b, axisKey | ((uint64_t)i << 31) | (uint64_t)step);
  }
};

// A sequential phase: block a walks [s[a], e[a]) of axis a on one lane.
__global__ __launch_bounds__(64) void k_ring_walk(BodyArrays B, PeriCell pc, RingSet R, RingSpan sp, Cands C,
                                                  unsigned long long* perr) {
  const int a = blockIdx.x;
  if (threadIdx.x != 0 || a >= 3 || sp.e[a] <= sp.s[a]) return;
  RingAxis ax{R.h[a], R.per[a], R.x[a], R.n2, R.L[a]};
  int64_t lo = R.lo[a];
  RingEmit em{B, pc, C, perr, (uint64_t)a << 62};
  if (ring_walk(ax, sp.s[a], sp.e[a], lo, em) != RING_OK) atomicOr(&R.err[a], (uint32_t)RING_ERR_RUN);
  R.lo[a] = lo;
}

// Window set-up: the window must stay clear of loIdx (lo + 2 when loIdx precedes the
// inserted elements).
__global__ void k_ring_win_init(RingSet R, RingWin w) {
  const int a = threadIdx.x;
  if (a >= 3) return;
  int64_t base = w.ws[a];
  const int64_t lo = R.lo[a];
  if (lo < w.i0[a] && lo + 2 > base) base = lo + 2;
  w.r0base[a] = base;
  w.r0[a] = base;
  w.minKey[a] = ~0ull;
}

// The last descent of the prefix window (the sorted part starts there) and the smallest
// inserted coordinate, all axes in one grid over their windows laid end to end.
__global__ void k_ring_win_scan(RingSet R, RingWin w) {
  int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int a = 0;
  while (a < 3) {
    const int64_t len = w.i1[a] > w.i0[a] ? w.i1[a] - w.ws[a] : 0;
    if (g < len) break;
    g -= len;
    a++;
  }
  if (a >= 3) return;
  const int64_t k = w.ws[a] + g;
  const double* X = R.x[a];
  if (k < w.i0[a]) {
    if (k > w.r0base[a] && X[k - 1] > X[k]) atomicMax((unsigned long long*)&w.r0[a], (unsigned long long)k);
  } else {
    atomicMin((unsigned long long*)&w.minKey[a], (unsigned long long)coord_key(X[k]));
  }
}

// The phase's conditions: a window inside [1, i0], no inserted element below the element
// before the sorted part (then none moves past it), loIdx outside the window.
__global__ void k_ring_win_check(RingSet R, RingWin w) {
  const int a = threadIdx.x;
  if (a >= 3 || w.i1[a] <= w.i0[a]) return;
  const int64_t r0 = w.r0[a];
  const int64_t lo = R.lo[a];
  bool ok = r0 >= 1 && w.r0base[a] <= w.i0[a];
  if (ok) ok = !(coord_key(R.x[a][r0 - 1]) > w.minKey[a]);
  if (ok) ok = !(lo >= r0 - 1 && lo <= w.i1[a]);
  if (!ok) atomicOr(&R.err[a], (uint32_t)w.bit);
}

// Sort keys of one axis's window [ws, i1): the part before the sorted prefix keeps its order
// (its position as key, below every coordinate key), the rest sorts by coordinate.
__global__ void k_ring_keys(RingSet R, int a, int64_t ws, int64_t m, const int64_t* r0, uint64_t* keys, uint32_t* vals) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= m) return;
  const int64_t pos = ws + t;
  keys[t] = pos < r0[a] ? (uint64_t)pos : coord_key(R.x[a][pos]);
  vals[t] = (uint32_t)t;
}

// Crossings of one window: the lower bound now at p, inserted from window offset q >= ins,
// passed the upper bounds now after it with an offset below q, nearest (highest p) first.
__global__ void k_ring_cross(BodyArrays B, PeriCell pc, RingSet R, int a, int64_t ws, int64_t ins, int64_t m,
                             const uint32_t* vals, const uint32_t* smin, Cands C, unsigned long long* perr) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= m - 1) return;
  const uint32_t q = vals[p];
  if ((int64_t)q < ins) return;
  const int32_t* H = R.h[a];
  const int32_t h = H[ws + q];
  if (h & 1) return;
  if (smin[p + 1] >= q) return;
  const int id = h >> 1;
  const uint64_t base = ((uint64_t)a << 62) | ((uint64_t)(ws + q) << 31);
  for (int64_t pp = p + 1; pp < m && smin[pp] < q; pp++) {
    if (vals[pp] >= q) continue;
    const int32_t h2 = H[ws + vals[pp]];
    if (!(h2 & 1)) continue;
    const int id2 = h2 >> 1;
    if (id2 == id) continue;
    int32_t per[3];
    if (!spatial_overlap_peri(B, pc, id, id2, per, perr)) continue;
    push_cand(C, id, id2, base | (uint64_t)(0x7FFFFFFFu - (uint32_t)pp));
  }
}

// The window in its sorted order: gather into the temporaries, then copy back.
__global__ void k_ring_gather(RingSet R, int a, int64_t ws, int64_t m, const uint32_t* vals, int32_t* th, int32_t* tp,
                              double* tx) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= m) return;
  const int64_t src = ws + vals[t];
  th[t] = R.h[a][src];
  tp[t] = R.per[a][src];
  tx[t] = R.x[a][src];
}

// The first run's rings: the bounds sorted by their wrapped coordinate (cellWrap into
// [0, L), :302-304) with the periods of the wrap; loIdx = 0.
__global__ void k_ring_init(BodyArrays B, PeriCell pc, int a, const uint32_t* sorted, int64_t n2, int32_t* h,
                            int32_t* per) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n2) return;
  const int32_t v = (int32_t)sorted[k];
  const int id = v >> 1;
  const int64_t N = B.n;
  const double b = B.hasBound[id] ? ((v & 1) ? B.bmax[a * N + id] : B.bmin[a * N + id]) : B.pos[a * N + id];
  int p;
  cell_wrap(b, 0.0, pc.L[a], p);
  h[k] = v;
  per[k] = p;
}
