// ring_dev.h -- device kernels of the periodic collider's ring walk (ring.h): the run's
// coordinates, the sequential phases (one wave per axis), and the parallel phases (window
// checks, sort keys, crossings, the permutation).
#pragma once
#include "periodic.h"
#include "ring.h"

// error bits per axis
enum {
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
  int64_t* r0;         // [3] device
  int64_t* r0base;     // [3] device
  uint64_t* minKey;    // [3] device: coord_key of the smallest inserted coordinate
  uint64_t* bottom;    // [3] device: the key no inserted element may be below (~0: none checked)
  int64_t* ext;        // [3] device: 0, or the position the sequential phase before must reach
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

// Lane l's value of a wave-uniform loop (readlane: the value lands in a scalar register).
__device__ __forceinline__ double ring_rl(double v, int l) {
  const long long b = __double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b & 0xffffffffll), l);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)((unsigned long long)b >> 32), l);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
__device__ __forceinline__ int32_t ring_rl(int32_t v, int l) { return __builtin_amdgcn_readlane(v, l); }

// ring_walk (ring.h) on one wave of 64 lanes, for vectors of at least 256 elements: the same
// comparisons in the same order, with the memory traffic spread over the lanes.  The outer
// walk reads 64 positions at a time; an insertion reads the 64 positions below the current
// one at a time (the reference reads each element once, before it is overwritten, so the
// values read ahead are the ones it compares), takes its decisions on the wave-uniform
// values lane by lane, then writes the moved elements and reports the inversions from their
// lanes.  The element before position i is tracked in a register: it is either the one the
// walk left there or the one the last insertion shifted into it.
template <class Emit>
__device__ int ring_walk_wave(const RingAxis& R, int64_t s, int64_t e, int64_t& loRef, Emit& emit) {
  const int lane = threadIdx.x & 63;
  const int64_t n2 = R.n2;
  const double L = R.L;
  int32_t* H = R.h;
  int32_t* P = R.per;
  double* X = R.coord;
  int64_t lo = loRef;
  double prevX = X[s == 0 ? n2 - 1 : s - 1];
  for (int64_t c0 = s; c0 < e; c0 += 64) {
    __threadfence_block();
    const int64_t myPos = c0 + lane;
    double cx = 0.0;
    int32_t ch = 0, cp = 0;
    if (myPos < e) {
      cx = X[myPos];
      ch = H[myPos];
      cp = P[myPos];
    }
    const int cnt = (int)((e - c0) < 64 ? (e - c0) : 64);
    const double below = __shfl_up(cx, 1);  // lane t: the value read at position c0 + t - 1
    int t = 0;
    while (t < cnt) {
      // positions t .. cnt-1 that need no work: at loIdx neither, and in order with the
      // element before (the one the walk has left at t - 1 for lane t, the read value after)
      {
        const double prev = lane == t ? prevX : below;
        const bool work = lane >= t && lane < cnt && (myPos == lo || !(prev <= cx));
        const uint64_t wm = __ballot(work);
        const int t1 = wm ? __ffsll((unsigned long long)wm) - 1 : cnt;
        if (t1 > t) prevX = ring_rl(cx, t1 - 1);
        t = t1;
        if (t >= cnt) break;
      }
      const int64_t i = c0 + t;
      double vC = ring_rl(cx, t);
      const int32_t vH = ring_rl(ch, t);
      int32_t vP = ring_rl(cp, t);
      t++;
      if (i == lo && vC < 0) {
        vP -= 1;
        vC += L;
        if (lane == 0) {
          P[i] = vP;
          X[i] = vC;
        }
        lo = lo + 1 == n2 ? 0 : lo + 1;
      }
      const double iCmp = vC + (i == lo ? L : 0.0);
      if (prevX <= iCmp) {
        prevX = vC;
        continue;
      }
      // insertion of the element of position i (at least one move: the test above is the
      // loop's first comparison)
      const bool isMin = !(vH & 1);
      const int32_t vId = vH >> 1;
      uint32_t step = 0;
      int64_t j = i == 0 ? n2 - 1 : i - 1;
      int64_t moves = 0;
      bool firstChunk = true;
      while (true) {
        __threadfence_block();  // the wave's earlier writes are visible to its loads
        int64_t pos = j - lane;
        if (pos < 0) pos += n2;
        const double x = X[pos];
        const int32_t h = H[pos];
        const int32_t p = P[pos];
        int m = 64;
        uint64_t adj = 0, evm = 0;
        // loIdx among the positions j-63 .. j+1 of this chunk's moves?
        int64_t dl = (j + 1 == n2 ? 0 : j + 1) - lo;
        if (dl < 0) dl += n2;
        if (dl > 64) {
          // no ring-head bookkeeping here: the comparisons are independent
          const uint64_t stop = __ballot(!(x > vC));
          m = stop ? __ffsll((unsigned long long)stop) - 1 : 64;
          const uint64_t below_m = m == 64 ? ~0ull : ((1ull << m) - 1ull);
          evm = __ballot(isMin && (h & 1) && (h >> 1) != vId) & below_m;
          moves += m;
        } else {
          for (int l = 0; l < 64; l++) {
            const int64_t jj = j - l < 0 ? j - l + n2 : j - l;
            const int64_t j1 = jj + 1 == n2 ? 0 : jj + 1;
            if (!(ring_rl(x, l) > vC + (j1 == lo ? L : 0.0))) {
              m = l;
              break;
            }
            if (jj == lo && vC < 0) {
              vP -= 1;
              vC += L;
              lo = lo + 1 == n2 ? 0 : lo + 1;
            } else if (j1 == lo) {
              adj |= 1ull << l;
              lo = lo == 0 ? n2 - 1 : lo - 1;
            }
            if (isMin) {
              const int32_t hl = ring_rl(h, l);
              if ((hl & 1) && (hl >> 1) != vId) evm |= 1ull << l;
            }
            moves++;
          }
        }
        // an insertion never moves this far in a ring of finite coordinates; the bound also
        // keeps its writes off the positions read ahead by the outer walk
        if (moves >= n2 - 128) return RING_RUNAWAY;
        if (lane < m) {  // the moved elements, one position up
          const int64_t d = pos + 1 == n2 ? 0 : pos + 1;
          const bool a = (adj >> lane) & 1ull;
          X[d] = a ? x - L : x;
          H[d] = h;
          P[d] = a ? p + 1 : p;
        }
        if (firstChunk) prevX = (adj & 1ull) ? ring_rl(x, 0) - L : ring_rl(x, 0);  // now at position i
        firstChunk = false;
        if ((evm >> lane) & 1ull) emit(i, step + (uint32_t)__popcll(evm & ((1ull << lane) - 1ull)), vId, h >> 1);
        step += (uint32_t)__popcll(evm);
        if (m < 64) {
          int64_t k = j - m + 1;
          if (k < 0) k += n2;
          if (lane == 0) {
            X[k] = vC;
            H[k] = vH;
            P[k] = vP;
          }
          break;
        }
        j -= 64;
        if (j < 0) j += n2;
      }
    }
  }
  __threadfence_block();
  loRef = lo;
  return RING_OK;
}

// A sequential phase: block a walks [s[a], e[a]) of axis a, on one wave (one lane for a
// vector shorter than 256 elements).
__global__ __launch_bounds__(64) void k_ring_walk(BodyArrays B, PeriCell pc, RingSet R, RingSpan sp, Cands C,
                                                  unsigned long long* perr) {
  const int a = blockIdx.x;
  if (a >= 3 || sp.e[a] <= sp.s[a]) return;
  RingAxis ax{R.h[a], R.per[a], R.x[a], R.n2, R.L[a]};
  int64_t lo = R.lo[a];
  RingEmit em{B, pc, C, perr, (uint64_t)a << 62};
  int r;
  if (R.n2 >= 256) {
    r = ring_walk_wave(ax, sp.s[a], sp.e[a], lo, em);
  } else {
    if (threadIdx.x != 0) return;
    r = ring_walk(ax, sp.s[a], sp.e[a], lo, em);
  }
  if (threadIdx.x == 0) {
    if (r != RING_OK) atomicOr(&R.err[a], (uint32_t)RING_ERR_RUN);
    R.lo[a] = lo;
  }
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
  w.bottom[a] = 0;
  w.ext[a] = 0;
}

// The last descent of the prefix window (the sorted part starts there) and the smallest
// inserted coordinate; blockIdx.y is the axis, one atomic per block for the minimum.
__global__ __launch_bounds__(256) void k_ring_win_scan(RingSet R, RingWin w) {
  __shared__ unsigned long long sMin[4];
  const int a = blockIdx.y;
  const int64_t len = w.i1[a] > w.i0[a] ? w.i1[a] - w.ws[a] : 0;
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long key = ~0ull;
  if (g < len) {
    const int64_t k = w.ws[a] + g;
    const double* X = R.x[a];
    if (k < w.i0[a]) {
      if (k > w.r0base[a] && X[k - 1] > X[k]) atomicMax((unsigned long long*)&w.r0[a], (unsigned long long)k);
    } else {
      key = (unsigned long long)coord_key(X[k]);
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long o = __shfl_xor(key, off);
    key = o < key ? o : key;
  }
  if ((threadIdx.x & 63) == 0) sMin[threadIdx.x >> 6] = key;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long m = sMin[0];
    for (int q = 1; q < (int)(blockDim.x >> 6); q++) m = sMin[q] < m ? sMin[q] : m;
    if (m != ~0ull) atomicMin((unsigned long long*)&w.minKey[a], m);
  }
}

// The phase's conditions: a window inside [1, i0]; no inserted element below both the
// bottom of the sorted part (then it stops inside that part) and the element before it (then
// it stops on top of that element), so none moves past r0; loIdx outside the window.  When
// they fail, ext asks the sequential phase before to go on: past loIdx, or (k_ring_win_viol)
// past the last element below the bottom.
__global__ void k_ring_win_check(RingSet R, RingWin w) {
  const int a = threadIdx.x;
  if (a >= 3 || w.i1[a] <= w.i0[a]) return;
  const int64_t r0 = w.r0[a];
  const int64_t lo = R.lo[a];
  const double* X = R.x[a];
  if (w.r0base[a] > w.i0[a] || (lo >= r0 - 1 && lo <= w.i1[a])) {
    const int64_t t = lo + 64;
    w.ext[a] = t < w.i1[a] ? t : w.i1[a];
    return;
  }
  uint64_t bottom = coord_key(X[r0 - 1]);
  if (r0 < w.i0[a]) {
    const uint64_t k0 = coord_key(X[r0]);
    bottom = k0 < bottom ? k0 : bottom;
  }
  if (w.minKey[a] < bottom) w.bottom[a] = bottom;
}

// The last inserted element below the bottom: the sequential phase takes it over.
__global__ void k_ring_win_viol(RingSet R, RingWin w) {
  const int a = blockIdx.y;
  const uint64_t bottom = w.bottom[a];
  if (bottom == 0) return;
  const int64_t k = w.i0[a] + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= w.i1[a]) return;
  if (coord_key(R.x[a][k]) < bottom) atomicMax((unsigned long long*)&w.ext[a], (unsigned long long)(k + 1));
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
// passed the upper bounds now after it with an offset below q; the key counts down from the
// nearest, which the reference passes last.  hn: the handles in new order; smin: suffix
// minimum of the old offsets (non-decreasing), so the passed elements lie in p+1 .. the last
// position where smin < q.  The overlap test is spatialOverlapPeri's with the moving body's
// bounds held in registers.
struct RingCross {
  const uint32_t* vals;
  const uint32_t* smin;
  const int32_t* hn;
  int64_t ws, ins, m;
  int a;
  uint32_t* longList;  // lower bounds that moved far: one wave each (k_ring_cross_long)
  unsigned long long* longCount;
  int64_t longCap;
  const float4* box;   // screen boxes of the bodies (k_screen_boxes): two float4 each
};

// One crossing candidate: the inserted lower bound of `id` (box m1, M1; screen box ca, ha) and
// the element at sorted position pp.  A pair the single-precision screen puts apart is one the
// exact test rejects without an error (both boxes narrow), so it is dropped on 32 bytes of
// the other body instead of six scattered doubles.
__device__ __forceinline__ void ring_cross_one(const BodyArrays& B, const PeriCell& pc, const RingCross& X,
                                               const double m1[3], const double M1[3], float4 ca, float4 ha, int id,
                                               uint32_t q, uint64_t base, int64_t pp, const Cands& C,
                                               unsigned long long* perr) {
  if (X.vals[pp] >= q) return;
  const int32_t h2 = X.hn[pp];
  if (!(h2 & 1)) return;
  const int id2 = h2 >> 1;
  if (id2 == id) return;
  if (ca.w == 0.0f) {
    const float4 cb = X.box[2 * id2];
    if (cb.w == 0.0f && screen_apart(pc, ca, ha, cb, X.box[2 * id2 + 1])) return;
  }
  const int64_t N = B.n;
  bool ok = true;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    if (!ok) break;
    int per;
    const int r = axis_overlap_peri(m1[k], M1[k], B.bmin[k * N + id2], B.bmax[k * N + id2], pc.L[k], per);
    if (r < 0) atomicOr(perr, 1ull);
    ok = r == 1;
  }
  if (ok) push_cand(C, id, id2, base | (uint64_t)(0x7FFFFFFFu - (uint32_t)pp));
}

#define SDM_RING_LONG 192  // a lower bound that moved back further is scanned by a whole wave

__global__ void k_ring_cross(BodyArrays B, PeriCell pc, RingCross X, Cands C, unsigned long long* perr) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= X.m - 1) return;
  const uint32_t q = X.vals[p];
  if ((int64_t)q < X.ins) return;
  const int32_t h = X.hn[p];
  if (h & 1) return;
  if (X.smin[p + 1] >= q) return;
  if ((int64_t)q - p > SDM_RING_LONG) {
    const unsigned long long slot = atomicAdd(X.longCount, 1ull);
    if ((int64_t)slot < X.longCap) {
      X.longList[slot] = (uint32_t)p;
      return;
    }
  }
  const int id = h >> 1;
  const int64_t N = B.n;
  double m1[3], M1[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    m1[k] = B.bmin[k * N + id];
    M1[k] = B.bmax[k * N + id];
  }
  const float4 ca = X.box[2 * id], ha = X.box[2 * id + 1];
  const uint64_t base = ((uint64_t)X.a << 62) | ((uint64_t)(X.ws + q) << 31);
  for (int64_t pp = p + 1; pp < X.m && X.smin[pp] < q; pp++)
    ring_cross_one(B, pc, X, m1, M1, ca, ha, id, q, base, pp, C, perr);
}

// The far-moved lower bounds, one wave each: lane l takes the positions p+1+l, p+65+l, ...
__global__ __launch_bounds__(256) void k_ring_cross_long(BodyArrays B, PeriCell pc, RingCross X, Cands C,
                                                         unsigned long long* perr) {
  const int lane = threadIdx.x & 63;
  const int64_t nItems = (int64_t)(*X.longCount < (unsigned long long)X.longCap ? *X.longCount : X.longCap);
  const int64_t N = B.n;
  for (int64_t it = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); it < nItems;
       it += (int64_t)gridDim.x * (blockDim.x >> 6)) {
    const int64_t p = X.longList[it];
    const uint32_t q = X.vals[p];
    const int id = X.hn[p] >> 1;
    double m1[3], M1[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
      m1[k] = B.bmin[k * N + id];
      M1[k] = B.bmax[k * N + id];
    }
    const float4 ca = X.box[2 * id], ha = X.box[2 * id + 1];
    const uint64_t base = ((uint64_t)X.a << 62) | ((uint64_t)(X.ws + q) << 31);
    for (int64_t c0 = p + 1; c0 < X.m && X.smin[c0] < q; c0 += 64) {
      const int64_t pp = c0 + lane;
      if (pp < X.m && X.smin[pp] < q) ring_cross_one(B, pc, X, m1, M1, ca, ha, id, q, base, pp, C, perr);
    }
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
