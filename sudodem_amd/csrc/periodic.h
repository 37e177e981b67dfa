// periodic.h -- broadphase helpers of a periodic cell.
//
// In a periodic scene InsertionSortCollider keeps each axis as a ring of bounds with a
// period per bound and insertion-sorts the ring (insertionSortPeri,
// pkg/common/InsertionSortCollider.cpp:416-463; ring.h and ring_dev.h walk it on the
// device); a pair is created when the boxes overlap in the sense of spatialOverlapPeri
// (:491-543), which also fixes the pair's cellDist.  This file holds that test, cellWrap,
// the erase flags, and the first run's sweep (:336-397), found here without the sweep: a
// cell list over the periodic cell (each body listed in every cell its box touches, sorted
// by cell), where every pair sharing cells is tested once, in the first cell of the cyclic
// intersection of the two cell ranges, with spatialOverlapPeri's own arithmetic.  Body A
// scans, along the sort axis, the lower bounds that follow its own lower bound in the wrapped
// ring until its upper bound; the pair is (A, B) and is inserted in the order (ring rank of
// A's lower bound, distance to B's lower bound), both from the ranks of the sorted wrapped
// bounds.
#pragma once
#include "collider.h"

struct PeriCell {
  double L[3];   // cell size
  double h[3];   // grid cell size (L / n)
  int32_t n[3];  // grid cells per axis
  double eps;    // absolute widening of every box before it is listed (covers rounding)
  // single-precision screen of box pairs (k_peri_pairs): cell size, its inverse and the
  // margin that covers the screen's rounding
  float Lf[3], iLf[3], mf[3];
};

// Boxes wider than this fraction of the cell on some axis are never screened: only the
// exact test (which reports a box spanning into the next period) decides their pairs.
#define SDM_PERI_WIDE 0.45

// cellWrap / cellWrapRel (InsertionSortCollider.cpp:404-414), same operations.
__device__ __forceinline__ double cell_wrap(double x, double x0, double x1, int& period) {
  const double xNorm = (x - x0) / (x1 - x0);
  period = (int)floor(xNorm);
  return x0 + (xNorm - (double)period) * (x1 - x0);
}
__device__ __forceinline__ double cell_wrap_rel(double x, double x0, double x1) {
  const double xNorm = (x - x0) / (x1 - x0);
  return (xNorm - floor(xNorm)) * (x1 - x0);
}

// One axis of spatialOverlapPeri.  Returns 1 overlap, 0 apart, -1 a box that spans into
// the next period (the reference throws "Body larger than half of the cell size").
__device__ __forceinline__ int axis_overlap_peri(double m1, double M1, double m2, double M2, double dim, int& per) {
  const double wMn = (cell_wrap_rel(m1, m2, m2 + dim) < cell_wrap_rel(m2, m1, m1 + dim)) ? m2 : m1;
  int pmn1, pmx1, pmn2, pmx2;
  const double mn1 = cell_wrap(m1, wMn, wMn + dim, pmn1), mx1 = cell_wrap(M1, wMn, wMn + dim, pmx1);
  const double mn2 = cell_wrap(m2, wMn, wMn + dim, pmn2), mx2 = cell_wrap(M2, wMn, wMn + dim, pmx2);
  if (pmn1 != pmx1 || pmn2 != pmx2) return -1;
  per = pmn1 - pmn2;
  return (mn1 <= mx2 && mx1 >= mn2) ? 1 : 0;
}

// spatialOverlapPeri (:491-543): axes in order, the first apart axis ends the test.
__device__ __forceinline__ bool spatial_overlap_peri(const BodyArrays& B, const PeriCell& pc, int a, int b,
                                                     int32_t per[3], unsigned long long* err) {
  const int64_t N = B.n;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    int p = 0;
    const int r = axis_overlap_peri(B.bmin[k * N + a], B.bmax[k * N + a], B.bmin[k * N + b], B.bmax[k * N + b],
                                    pc.L[k], p);
    if (r < 0) {
      atomicOr(err, 1ull);
      return false;
    }
    per[k] = p;
    if (r == 0) return false;
  }
  return true;
}

// conditionalyEraseNonReal with shouldBeErased = !spatialOverlapPeri (InsertionSortCollider.hpp:179-182).
__global__ void k_erase_flags_peri(BodyArrays B, PairArrays P, int64_t np, PeriCell pc, uint32_t* flag,
                                   unsigned long long* err) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  const bool real = P.r[p].hasGeom[0] && P.r[p].hasPhys[0];
  int32_t per[3];
  flag[p] = (!real && !spatial_overlap_peri(B, pc, P.r[p].id1[0], P.r[p].id2[0], per, err)) ? 1u : 0u;
}

// Box widths: sum and maximum (for the grid size); a body without finite bounds (a wall)
// cannot live in a periodic cell and sets err bit 2.
__global__ void k_width_stats(BodyArrays B, double* sum, unsigned long long* maxBits, unsigned long long* err) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t N = B.n;
  double w = 0;
  if (i < N) {
    for (int k = 0; k < 3; k++) {
      const double d = B.bmax[k * N + i] - B.bmin[k * N + i];
      if (!isfinite(d)) {
        atomicOr(err, 2ull);
        w = 0;
        break;
      }
      w = fmax(w, d);
    }
  }
  __shared__ double ps[16];
  __shared__ double pm[16];
  double s = w, m = w;
  for (int off = 32; off > 0; off >>= 1) {
    s += __shfl_down(s, off);
    m = fmax(m, __shfl_down(m, off));
  }
  if ((threadIdx.x & 63) == 0) {
    ps[threadIdx.x >> 6] = s;
    pm[threadIdx.x >> 6] = m;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double ts = 0, tm = 0;
    for (int k = 0; k < (int)(blockDim.x >> 6); k++) {
      ts += ps[k];
      tm = fmax(tm, pm[k]);
    }
    atomicAdd(sum, ts);
    atomicMax(maxBits, (unsigned long long)__double_as_longlong(tm));  // non-negative: bit order = value order
  }
}

// Single-precision box of the pair screen: wrapped centre and half widths.
__device__ __forceinline__ void screen_box(double m, double M, double L, float& cen, float& hw, bool& wide) {
  int per;
  cen = (float)cell_wrap(0.5 * (m + M), 0.0, L, per);
  hw = (float)(0.5 * (M - m));
  wide = wide || !(M - m < SDM_PERI_WIDE * L);
}

// The screen of a pair of narrow boxes (centre c, half widths h; .w of the centre 1 for a
// wide box): apart on some axis by the minimum-image distance of the centres, beyond a margin
// far above the screen's rounding, so no pair the exact test accepts is dropped.  Wide boxes
// are never screened: only the exact test (which reports a box spanning into the next period)
// decides their pairs.
__device__ __forceinline__ bool screen_apart(const PeriCell& pc, float4 ca, float4 ha, float4 cb, float4 hb) {
  if (ca.w != 0.0f || cb.w != 0.0f) return false;
  const float d[3] = {ca.x - cb.x, ca.y - cb.y, ca.z - cb.z};
  const float h[3] = {ha.x + hb.x, ha.y + hb.y, ha.z + hb.z};
  bool apart = false;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const float dk = d[k] - pc.Lf[k] * rintf(d[k] * pc.iLf[k]);
    apart = apart || fabsf(dk) > h[k] + pc.mf[k];
  }
  return apart;
}

// The screen boxes of every body alone (a later collider run: the ring walk's crossings).
__global__ void k_screen_boxes(BodyArrays B, PeriCell pc, float4* box) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t N = B.n;
  if (i >= N) return;
  float cen[3], hw[3];
  bool wide = false;
#pragma unroll
  for (int k = 0; k < 3; k++) screen_box(B.bmin[k * N + i], B.bmax[k * N + i], pc.L[k], cen[k], hw[k], wide);
  box[2 * i] = make_float4(cen[0], cen[1], cen[2], wide ? 1.0f : 0.0f);
  box[2 * i + 1] = make_float4(hw[0], hw[1], hw[2], 0.0f);
}

// Cell range of every body: start cell s[k] and count c[k] per axis (c = n: the whole
// axis), and the number of cells the box touches.  Also the boxes of the pair screen,
// box[2i], box[2i+1] = centre and half widths; box[2i].w is 1 for a wide box.
__global__ void k_peri_ranges(BodyArrays B, PeriCell pc, int32_t* rng, uint32_t* cnt, float4* box) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t N = B.n;
  if (i >= N) return;
  uint32_t total = 1;
  float cen[3], hw[3];
  bool wide = false;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const int n = pc.n[k];
    screen_box(B.bmin[k * N + i], B.bmax[k * N + i], pc.L[k], cen[k], hw[k], wide);
    const double lo = B.bmin[k * N + i] - pc.eps;
    const double w = (B.bmax[k * N + i] + pc.eps) - lo;
    int per;
    double lw = cell_wrap(lo, 0.0, pc.L[k], per);
    int s = (int)floor(lw / pc.h[k]);
    s = s < 0 ? 0 : (s >= n ? n - 1 : s);
    int e = (int)floor((lw + w) / pc.h[k]);
    int c = e - s + 1;
    if (c < 1) c = 1;
    if (c >= n) {
      c = n;
      s = 0;
    }
    rng[6 * i + k] = s;
    rng[6 * i + 3 + k] = c;
    total *= (uint32_t)c;
  }
  cnt[i] = total;
  box[2 * i] = make_float4(cen[0], cen[1], cen[2], wide ? 1.0f : 0.0f);
  box[2 * i + 1] = make_float4(hw[0], hw[1], hw[2], 0.0f);
}

// Entries (cell, body) of every body at its offset.
__global__ void k_peri_entries(int64_t N, PeriCell pc, const int32_t* rng, const uint32_t* offs, uint32_t* key,
                               uint32_t* val) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int s0 = rng[6 * i], s1 = rng[6 * i + 1], s2 = rng[6 * i + 2];
  const int c0 = rng[6 * i + 3], c1 = rng[6 * i + 4], c2 = rng[6 * i + 5];
  uint32_t o = offs[i];
  for (int z = 0; z < c2; z++) {
    const int cz = (s2 + z) % pc.n[2];
    for (int y = 0; y < c1; y++) {
      const int cy = (s1 + y) % pc.n[1];
      for (int x = 0; x < c0; x++) {
        const int cx = (s0 + x) % pc.n[0];
        key[o] = (uint32_t)(cx + pc.n[0] * (cy + pc.n[1] * cz));
        val[o] = (uint32_t)i;
        o++;
      }
    }
  }
}

// Start and end of the run of equal cells in the sorted entries, per cell.
__global__ void k_cell_ends(const uint32_t* key, int64_t m, uint32_t* runStart, uint32_t* runEnd) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= m) return;
  if (e == 0 || key[e - 1] != key[e]) runStart[key[e]] = (uint32_t)e;
  if (e == m - 1 || key[e + 1] != key[e]) runEnd[key[e]] = (uint32_t)(e + 1);
}

// x in the cyclic range [s, s + c) of 0..n-1 (x, s in [0, n)).
__device__ __forceinline__ bool cyc_in(int x, int s, int c, int n) {
  int d = x - s;
  if (d < 0) d += n;
  return d < c;
}

// Smallest cell index of the intersection of two cyclic ranges (-1 if empty).
__device__ __forceinline__ int first_shared(int sa, int ca, int sb, int cb, int n) {
  int best = n;
  if (cyc_in(0, sa, ca, n) && cyc_in(0, sb, cb, n)) best = 0;
  if (cyc_in(sa, sb, cb, n) && sa < best) best = sa;
  if (cyc_in(sb, sa, ca, n) && sb < best) best = sb;
  return best == n ? -1 : best;
}

struct PeriOrder {
  const uint32_t* rankMin;  // ring rank of each body's lower / upper bound on the sort axis
  const uint32_t* rankMax;
  uint32_t n2;
};

// A pair that survived the screen: the sweep's orientation and order key, then
// spatialOverlapPeri's exact test with (id1, id2) as the reference evaluates it, then the
// container lookup (hash of its unordered pairs); a new overlapping pair becomes a candidate.
__device__ __forceinline__ void peri_candidate(const BodyArrays& B, const PeriCell& pc, const PeriOrder& po, int a, int b,
                                               const uint64_t* tab, uint64_t mask, const Cands& C,
                                               unsigned long long* err) {
  const uint32_t n2 = po.n2;
  const uint32_t dAB = (po.rankMin[b] + n2 - po.rankMin[a]) % n2, dAA = (po.rankMax[a] + n2 - po.rankMin[a]) % n2;
  const uint32_t dBA = (po.rankMin[a] + n2 - po.rankMin[b]) % n2, dBB = (po.rankMax[b] + n2 - po.rankMin[b]) % n2;
  const bool aScans = dAB < dAA, bScans = dBA < dBB;
  if (!aScans && !bScans) return;  // never met in the sweep
  const uint64_t kA = ((uint64_t)po.rankMin[a] << 32) | dAB, kB = ((uint64_t)po.rankMin[b] << 32) | dBA;
  int id1, id2;
  uint64_t okey;
  if (aScans && (!bScans || kA < kB)) {
    id1 = a; id2 = b; okey = kA;
  } else {
    id1 = b; id2 = a; okey = kB;
  }
  int32_t per[3];
  if (!spatial_overlap_peri(B, pc, id1, id2, per, err)) return;
  if (hash_contains(tab, mask, upair(id1, id2))) return;
  const unsigned long long slot = atomicAdd(C.count, 1ull);
  if ((int64_t)slot < C.cap) {
    C.a[slot] = id1;
    C.b[slot] = id2;
    C.key[slot] = okey;
  }
}

// Entries of a cell staged in LDS (the rest of a longer cell is read from global memory).
#define SDM_PERI_TILE 128

// One block per grid cell: every pair of the cell's entries whose first shared cell is this
// one, found overlapping and not in the container, becomes a candidate.  Each unordered pair
// is met in exactly one cell, so the candidates are distinct.  The container is the one
// before this run's erase step, which removes only pairs that do not overlap now, so the
// candidates are the same as after it.
//
// The k entries are paired round-robin: entry i meets the entries i+1 .. i+(k-1)/2 after
// it, cyclically, and for even k the one at distance k/2 when i < k/2, so every unordered
// pair is met once and all lanes loop equally long.  A pair of narrow boxes is screened in
// single precision (minimum-image distance of the centres against the half widths plus a
// margin far above the screen's rounding, so no pair the exact test accepts is dropped).
__global__ __launch_bounds__(64) void k_peri_cell_pairs(BodyArrays B, PeriCell pc, PeriOrder po, const int32_t* rng,
                                                        const float4* box, const uint32_t* val,
                                                        const uint32_t* runStart, const uint32_t* runEnd,
                                                        const uint64_t* tab, uint64_t mask, Cands C,
                                                        unsigned long long* err) {
  __shared__ int32_t sId[SDM_PERI_TILE];
  __shared__ uint32_t sRg[3][SDM_PERI_TILE];
  __shared__ float4 sBox[2][SDM_PERI_TILE];
  const uint32_t cell = blockIdx.x;
  const uint32_t s0 = runStart[cell];
  const int kc = (int)(runEnd[cell] - s0);
  if (kc < 2) return;  // the whole block
  const int X[3] = {(int)(cell % (uint32_t)pc.n[0]), (int)((cell / (uint32_t)pc.n[0]) % (uint32_t)pc.n[1]),
                    (int)(cell / ((uint32_t)pc.n[0] * (uint32_t)pc.n[1]))};
  const int nl = kc < SDM_PERI_TILE ? kc : SDM_PERI_TILE;
  for (int j = threadIdx.x; j < nl; j += blockDim.x) {
    const int a = (int)val[s0 + j];
    sId[j] = a;
#pragma unroll
    for (int k = 0; k < 3; k++) sRg[k][j] = (uint32_t)rng[6 * a + k] | ((uint32_t)rng[6 * a + 3 + k] << 16);
#pragma unroll
    for (int q = 0; q < 2; q++) sBox[q][j] = box[2 * a + q];
  }
  __syncthreads();
  auto ent_id = [&](int j) -> int { return j < SDM_PERI_TILE ? sId[j] : (int)val[s0 + j]; };
  auto ent_rg = [&](int j, int k) -> uint32_t {
    if (j < SDM_PERI_TILE) return sRg[k][j];
    const int a = (int)val[s0 + j];
    return (uint32_t)rng[6 * a + k] | ((uint32_t)rng[6 * a + 3 + k] << 16);
  };
  auto ent_box = [&](int j, int q) -> float4 { return j < SDM_PERI_TILE ? sBox[q][j] : box[2 * (int)val[s0 + j] + q]; };
  for (int i = threadIdx.x; i < kc; i += blockDim.x) {
    const int a = ent_id(i);
    uint32_t ra[3];
#pragma unroll
    for (int k = 0; k < 3; k++) ra[k] = ent_rg(i, k);
    const float4 ca = ent_box(i, 0), ha = ent_box(i, 1);
    const int iters = (kc - 1) / 2 + ((kc % 2 == 0 && i < kc / 2) ? 1 : 0);
    for (int t = 1; t <= iters; t++) {
      int j = i + t;
      if (j >= kc) j -= kc;
      bool first = true;
#pragma unroll
      for (int k = 0; k < 3; k++) {
        if (first) {
          const uint32_t rb = ent_rg(j, k);
          first = first_shared((int)(ra[k] & 0xffffu), (int)(ra[k] >> 16), (int)(rb & 0xffffu), (int)(rb >> 16),
                               pc.n[k]) == X[k];
        }
      }
      if (!first) continue;
      const int b = ent_id(j);
      const float4 cb = ent_box(j, 0);
      if (ca.w == 0.0f && cb.w == 0.0f && screen_apart(pc, ca, ha, cb, ent_box(j, 1))) continue;
      peri_candidate(B, pc, po, a, b, tab, mask, C, err);
    }
  }
}

// Sort keys of the wrapped bounds on the sort axis (first run: coord = cellWrap(bound, 0, L)).
__global__ void k_peri_bound_keys(BodyArrays B, PeriCell pc, int axis, int64_t n2, uint64_t* keys, uint32_t* vals) {
  const int64_t h = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (h >= n2) return;
  const int64_t id = h >> 1;
  const int64_t N = B.n;
  const double x = (h & 1) ? B.bmax[axis * N + id] : B.bmin[axis * N + id];
  int per;
  keys[h] = coord_key(cell_wrap(x, 0.0, pc.L[axis], per));
  vals[h] = (uint32_t)h;
}

__global__ void k_peri_ranks(const uint32_t* sortedVals, int64_t n2, uint32_t* rankMin, uint32_t* rankMax) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n2) return;
  const uint32_t h = sortedVals[q];
  if (h & 1)
    rankMax[h >> 1] = (uint32_t)q;
  else
    rankMin[h >> 1] = (uint32_t)q;
}

// cellDist of the pairs appended at [start, start + count): the periods spatialOverlapPeri
// returns for (id1, id2) (handleBoundInversionPeri, :466-475).
__global__ void k_new_celldist(PairArrays P, int64_t start, int64_t count, BodyArrays B, PeriCell pc,
                               unsigned long long* err) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= count) return;
  PairRec& R = P.r[start + k];
  int32_t per[3] = {0, 0, 0};
  spatial_overlap_peri(B, pc, R.id1[0], R.id2[0], per, err);
  R.cellDist[0] = per[0];
  R.cellDist[1] = per[1];
  R.cellDist[2] = per[2];
}

// ---------------------------------------------------------------------------------------
// Initial sort-and-sweep of an aperiodic scene (InsertionSortCollider.cpp:336-397) without
// the sweep: the sweep creates (A, B) when B's lower bound lies strictly between A's lower
// and upper bounds in the sorted order of the sort axis and the boxes overlap
// (spatialOverlap), and inserts the pairs in the order (rank of A's lower bound, rank of B's
// lower bound).  Every overlapping pair of bounded bodies is found here instead, the bodies
// with finite boxes through a cell list (the pair met once, in its first shared cell) and
// the walls (boxes infinite on two axes) against every body; each gets the sweep's
// orientation and order key from the ranks, and pairs the sweep would not create (boxes
// that only touch, with the tie broken the other way by the sort) are dropped.
// ---------------------------------------------------------------------------------------

// The sweep's decision for an overlapping pair (a, b): scanner s, scanned t, order key.
__device__ __forceinline__ bool sweep_pair(const BodyArrays& B, const uint32_t* rankMin, const uint32_t* rankMax, int a,
                                           int b, int& s, int& t, uint64_t& key) {
  s = rankMin[a] < rankMin[b] ? a : b;
  t = s == a ? b : a;
  if (!(rankMin[t] < rankMax[s])) return false;
  if (!spatial_overlap(B, s, t)) return false;
  key = ((uint64_t)rankMin[s] << 32) | (uint64_t)rankMin[t];
  return true;
}

__device__ __forceinline__ void push_cand(const Cands& C, int s, int t, uint64_t key) {
  const unsigned long long slot = atomicAdd(C.count, 1ull);
  if ((int64_t)slot < C.cap) {
    C.a[slot] = s;
    C.b[slot] = t;
    C.key[slot] = key;
  }
}

// Union box and width statistics of the bounded bodies with finite boxes (no walls), per
// block: lo[3], hi[3], width sum, width max, count.
__global__ __launch_bounds__(256) void k_ic_stats(BodyArrays B, double* partial) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t N = B.n;
  const double inf = __builtin_huge_val();
  double v[9] = {inf, inf, inf, -inf, -inf, -inf, 0.0, 0.0, 0.0};
  if (i < N && B.kind[i] == SDM_KIND_SQ && B.hasBound[i]) {
    double w = 0;
    for (int k = 0; k < 3; k++) {
      const double lo = B.bmin[k * N + i], hi = B.bmax[k * N + i];
      v[k] = lo;
      v[3 + k] = hi;
      w = fmax(w, hi - lo);
    }
    v[6] = w;
    v[7] = w;
    v[8] = 1.0;
  }
  __shared__ double red[9][256];
  for (int k = 0; k < 9; k++) red[k][threadIdx.x] = v[k];
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      for (int k = 0; k < 3; k++) red[k][threadIdx.x] = fmin(red[k][threadIdx.x], red[k][threadIdx.x + s]);
      for (int k = 3; k < 6; k++) red[k][threadIdx.x] = fmax(red[k][threadIdx.x], red[k][threadIdx.x + s]);
      red[6][threadIdx.x] += red[6][threadIdx.x + s];
      red[7][threadIdx.x] = fmax(red[7][threadIdx.x], red[7][threadIdx.x + s]);
      red[8][threadIdx.x] += red[8][threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x < 9) partial[(int64_t)blockIdx.x * 9 + threadIdx.x] = red[threadIdx.x][0];
}

__global__ void k_ic_stats_final(const double* partial, int64_t nb, double* out) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  const double inf = __builtin_huge_val();
  double v[9] = {inf, inf, inf, -inf, -inf, -inf, 0.0, 0.0, 0.0};
  for (int64_t b = 0; b < nb; b++) {
    const double* p = partial + b * 9;
    for (int k = 0; k < 3; k++) v[k] = fmin(v[k], p[k]);
    for (int k = 3; k < 6; k++) v[k] = fmax(v[k], p[k]);
    v[6] += p[6];
    v[7] = fmax(v[7], p[7]);
    v[8] += p[8];
  }
  for (int k = 0; k < 9; k++) out[k] = v[k];
}

// Cell range of every body with a finite box (start cell and count per axis, clamped to the
// grid), 0 cells for the others.
__global__ void k_ic_ranges(BodyArrays B, PeriCell pc, double lo0, double lo1, double lo2, int32_t* rng,
                            uint32_t* cnt) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t N = B.n;
  if (i >= N) return;
  if (!(B.kind[i] == SDM_KIND_SQ && B.hasBound[i])) {
    for (int k = 0; k < 6; k++) rng[6 * i + k] = 0;  // no cells: k_peri_entries writes nothing
    cnt[i] = 0;
    return;
  }
  const double lo[3] = {lo0, lo1, lo2};
  uint32_t total = 1;
  for (int k = 0; k < 3; k++) {
    const int n = pc.n[k];
    int s = (int)floor((B.bmin[k * N + i] - lo[k]) / pc.h[k]);
    int e = (int)floor((B.bmax[k * N + i] - lo[k]) / pc.h[k]);
    s = s < 0 ? 0 : (s >= n ? n - 1 : s);
    e = e < 0 ? 0 : (e >= n ? n - 1 : e);
    if (e < s) e = s;
    rng[6 * i + k] = s;
    rng[6 * i + 3 + k] = e - s + 1;
    total *= (uint32_t)(e - s + 1);
  }
  cnt[i] = total;
}

// One block per grid cell: every pair of the cell's entries whose first shared cell is this
// one (the larger start cell on each axis) becomes a candidate if the sweep would create it.
__global__ __launch_bounds__(64) void k_ic_cell_pairs(BodyArrays B, PeriCell pc, const int32_t* rng,
                                                      const uint32_t* val, const uint32_t* runStart,
                                                      const uint32_t* runEnd, const uint32_t* rankMin,
                                                      const uint32_t* rankMax, Cands C) {
  __shared__ int32_t sId[SDM_PERI_TILE];
  __shared__ uint32_t sRg[3][SDM_PERI_TILE];
  const uint32_t cell = blockIdx.x;
  const uint32_t s0 = runStart[cell];
  const int kc = (int)(runEnd[cell] - s0);
  if (kc < 2) return;  // the whole block
  const int X[3] = {(int)(cell % (uint32_t)pc.n[0]), (int)((cell / (uint32_t)pc.n[0]) % (uint32_t)pc.n[1]),
                    (int)(cell / ((uint32_t)pc.n[0] * (uint32_t)pc.n[1]))};
  const int nl = kc < SDM_PERI_TILE ? kc : SDM_PERI_TILE;
  for (int j = threadIdx.x; j < nl; j += blockDim.x) {
    const int a = (int)val[s0 + j];
    sId[j] = a;
#pragma unroll
    for (int k = 0; k < 3; k++) sRg[k][j] = (uint32_t)rng[6 * a + k];
  }
  __syncthreads();
  auto ent_id = [&](int j) -> int { return j < SDM_PERI_TILE ? sId[j] : (int)val[s0 + j]; };
  auto ent_s = [&](int j, int k) -> int { return j < SDM_PERI_TILE ? (int)sRg[k][j] : rng[6 * (int)val[s0 + j] + k]; };
  for (int i = threadIdx.x; i < kc; i += blockDim.x) {
    const int a = ent_id(i);
    int sa[3];
#pragma unroll
    for (int k = 0; k < 3; k++) sa[k] = ent_s(i, k);
    const int iters = (kc - 1) / 2 + ((kc % 2 == 0 && i < kc / 2) ? 1 : 0);
    for (int t = 1; t <= iters; t++) {
      int j = i + t;
      if (j >= kc) j -= kc;
      bool first = true;
#pragma unroll
      for (int k = 0; k < 3; k++) {
        if (first) {
          const int sb = ent_s(j, k);
          first = (sa[k] > sb ? sa[k] : sb) == X[k];
        }
      }
      if (!first) continue;
      const int b = ent_id(j);
      int s, u;
      uint64_t key;
      if (sweep_pair(B, rankMin, rankMax, a, b, s, u, key)) push_cand(C, s, u, key);
    }
  }
}

// Walls (listed in `walls`) against every bounded body; a pair of walls once.
__global__ __launch_bounds__(256) void k_ic_walls(BodyArrays B, const int32_t* walls, int nw, const uint32_t* rankMin,
                                                  const uint32_t* rankMax, Cands C) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B.n || !B.hasBound[b]) return;
  const bool bWall = B.kind[b] == SDM_KIND_WALL;
  for (int k = 0; k < nw; k++) {
    const int w = walls[k];
    if (w == (int)b || !B.hasBound[w]) continue;
    if (bWall && (int)b < w) continue;
    int s, u;
    uint64_t key;
    if (sweep_pair(B, rankMin, rankMax, w, (int)b, s, u, key)) push_cand(C, s, u, key);
  }
}
