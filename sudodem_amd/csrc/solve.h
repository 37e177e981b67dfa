// solve.h -- persistent, one-evaluation-per-iteration narrowphase solvers.
//
// The two common-normal solvers of the reference take a data-dependent number of distance
// evaluations per pair (Nelder-Mead: 1 to ~60, Levenberg-Marquardt: 1 to ~200).  Run one
// pair per thread to completion and a wavefront waits for its slowest lane.  Here each
// solver is a per-lane state machine whose loop body is exactly one distance evaluation;
// a lane that finishes its pair takes the next one from its wave's batch, so all 64 lanes
// of a wave execute the same evaluation code in lockstep whatever stage their pair is in.
// The sequence of evaluations and decisions of every pair is that of the reference:
//   neldermead_simplex + amotry (pkg/dem/Superquadrics_Ig2.cpp:1681-1875)
//   lmder2 + JacfunctionApp     (Superquadrics_Ig2.cpp:1032-1354)
//
// Mixed precision, Nelder-Mead: the first vertex of a pair is evaluated in double and
// every other vertex as a float difference from it (delta.h), whose rounding bound scales
// with the support displacement and not with the millimetre-size support point.  The
// touching test of an evaluation is taken in float only when |c.(p2 - p1)| exceeds that
// bound; otherwise, and where the difference form leaves its range (a direction component
// coming within 1/64 of its anchor value of zero), the point is evaluated in double and
// becomes the new anchor.
// The simplex vertices, sums and moves stay in double; comparisons and the convergence test
// use the float values, so a comparison closer than the rounding bound can take the other
// branch than the double solver.  The final double evaluations (k_nm_refine) give the depth
// at the best vertex and the points of the last evaluation.
// Mixed precision, Levenberg-Marquardt: double with the device's own functions (a float
// difference form was slower on MI355X, see k_solve_lm).  Both solvers start with a certified
// float screen of the first evaluation: a pair that is not touching there is complete after
// it, in either precision form.
#pragma once
#include "delta.h"
#include "state.h"

struct SolveIO {  // indexed by work slot
  int32_t* pair;   // interaction index
  int32_t* tag;    // LM restart marker added to the reported stop code (100, 200)
  int32_t* id1;
  int32_t* id2;
  double* a0;      // in: start parameters; out: best vertex (NM) / final parameters (LM)
  double* a1;
  double* del;     // NM: initial simplex size
  double* f;       // NM: fmin
  double* la0;     // NM: last evaluated point
  double* la1;
  double* p1;      // [3][cap] points of the last evaluation
  double* p2;
  double* ctr;     // LM: midpoint of the last evaluation
  double* ct;      // LM: contact direction set at the top of the last iteration (in/out)
  int32_t* nev;    // distance evaluations (accumulated over solver passes)
  int32_t* stop;   // LM stop code / NM nfunc
  int32_t* touch;  // 1 = touching
  double* sh;      // LM: [3][cap] periodic shift of body 2 (shift2); nullptr outside a periodic cell
  float* anc;      // mixed NM: [kAnc][cap] the double anchor of the first vertex, as floats
  int32_t* fb;     // mixed NM: 1 = no anchor at the first vertex; handed-over stage and indices
  int64_t cap;
};

// Rows of SolveIO::anc: per body (offset 0 and 10) the reciprocals of the local direction (3),
// sin^2 and cos^2 of phi0 and phi1 (4), the local support point (3); then cos / sin of the
// anchor angles (20-23), the gap vector p2 - p1 as high (24-26) and low (27-29) float parts,
// the contact direction (30-32).
constexpr int kAnc = 33;

struct SolveArgs {
  const double* geo;  // packed body records (state.h): pos, rot_mat2global, semi-axes, exponents, r_max
  int64_t N;
  SolveIO S;
  const int32_t* list;  // work slots to process (nullptr: slots 0..nwork-1)
  int64_t nwork;        // number of work items, or their upper bound when `count` is set
  const unsigned long long* count;  // device-side number of work items (nullptr: nwork)
  unsigned long long* next;   // work counter (64-entry batches), zero at launch
  unsigned long long* evals;  // distance evaluations done by this launch (added)
  unsigned long long* dbl = nullptr;  // mixed NM: evaluations done in double (added)
  int32_t* hoList = nullptr;          // mixed NM: slots handed to the resume launch
  unsigned long long* hoCount = nullptr;  // their number (zero before the main launch)
  bool lmRestarts = false;  // LM (two lanes per pair): the Geom functor's restarts inline
};

// Work items of a launch: the device-side count (written by an earlier kernel of the same
// step, so the host never waits for it) capped by the launch's upper bound.
__device__ __forceinline__ int64_t work_n(const SolveArgs& A) {
  if (!A.count) return A.nwork;
  const int64_t c = (int64_t)*A.count;
  return c < A.nwork ? c : A.nwork;
}

// Add per-thread values a, b into c[0], c[1] with one atomic per block: a device-scope
// atomic on one address from every wavefront serialises across the chip.  Every thread of
// the block must call it.
__device__ __forceinline__ void block_add2(unsigned long long* c, unsigned long long a, unsigned long long b) {
  __shared__ unsigned long long part[2][16];
  for (int off = 32; off > 0; off >>= 1) {
    a += __shfl_down(a, off);
    b += __shfl_down(b, off);
  }
  const int wv = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    part[0][wv] = a;
    part[1][wv] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long sa = 0, sb = 0;
    for (int k = 0; k < (int)((blockDim.x + 63) >> 6); k++) {
      sa += part[0][k];
      sb += part[1][k];
    }
    if (sa) atomicAdd(&c[0], sa);
    if (sb) atomicAdd(&c[1], sb);
  }
  __syncthreads();
}
__device__ __forceinline__ void add_evals(unsigned long long* c, unsigned long long v) {
  __shared__ unsigned long long part[16];
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long s = 0;
    for (int k = 0; k < (int)((blockDim.x + 63) >> 6); k++) s += part[k];
    if (s) atomicAdd(c, s);
  }
  __syncthreads();
}

__device__ __forceinline__ void st3s(double* a, int64_t i, int64_t s, V3<double> v) {
  a[i] = v.x;
  a[s + i] = v.y;
  a[2 * s + i] = v.z;
}
__device__ __forceinline__ V3<double> ld3s(const double* a, int64_t i, int64_t s) {
  return mk3(a[i], a[s + i], a[2 * s + i]);
}

// One packed body record (the part the narrowphase uses: up to the exponents).
struct GeoRec {
  V3<double> pos;
  M3<double> G;
  Shape<double> s;
  double rmax;
};
__device__ __forceinline__ GeoRec load_rec(const double* geo, int id) {
  const double2* r = reinterpret_cast<const double2*>(geo + (int64_t)id * GEO_STRIDE);
  double v[GEO_VEL];
#pragma unroll
  for (int k = 0; k < GEO_VEL / 2; k++) {
    const double2 t = r[k];
    v[2 * k] = t.x;
    v[2 * k + 1] = t.y;
  }
  GeoRec g;
  g.pos = mk3(v[GEO_POS], v[GEO_POS + 1], v[GEO_POS + 2]);
#pragma unroll
  for (int k = 0; k < 9; k++) g.G.m[k] = v[GEO_G + k];
  g.s.rx = v[GEO_R]; g.s.ry = v[GEO_R + 1]; g.s.rz = v[GEO_R + 2];
  g.s.eps1 = v[GEO_E]; g.s.eps2 = v[GEO_E + 1];
  g.rmax = v[GEO_RMAX];
  return g;
}

template <typename T>
__device__ __forceinline__ Shape<T> cast_shape(const Shape<double>& s) {
  Shape<T> r;
  r.rx = T(s.rx); r.ry = T(s.ry); r.rz = T(s.rz); r.eps1 = T(s.eps1); r.eps2 = T(s.eps2);
  return r;
}

template <typename T>
__device__ __forceinline__ void pair_geo(const GeoRec& r1, const GeoRec& r2, bool relative, PairGeo<T>& g) {
  g.s1 = cast_shape<T>(r1.s);
  g.s2 = cast_shape<T>(r2.s);
  g.G1 = cast9<T>(r1.G);
  g.G2 = cast9<T>(r2.G);
  const V3<double> x1 = r1.pos;
  const V3<double> x2 = r2.pos;
  if (relative) {
    g.pos1 = mk3(T(0), T(0), T(0));
    g.pos2 = cast3<T>(x2 - x1);
  } else {
    g.pos1 = cast3<T>(x1);
    g.pos2 = cast3<T>(x2);
  }
}

// Wave-level work fetch: the idle lanes take consecutive queue entries with one atomic per wave.
__device__ __forceinline__ int64_t fetch_work(bool want, unsigned long long* counter) {
  const uint64_t m = __ballot(want);
  if (m == 0) return -1;
  const int lane = threadIdx.x & 63;
  const int leader = __ffsll((unsigned long long)m) - 1;
  unsigned long long base = 0;
  if (lane == leader) base = atomicAdd(counter, (unsigned long long)__popcll(m));
  base = __shfl(base, leader);
  const uint64_t below = lane ? (m & (~0ull >> (64 - lane))) : 0ull;
  return want ? (int64_t)(base + __popcll(below)) : -1;
}

// Work distribution of the persistent solvers: a wavefront claims 64 consecutive list
// entries with one atomic, and its lanes take them one at a time as they go idle.  One
// shared-counter atomic per 64 pairs (instead of one per wavefront iteration) keeps the
// counter off the critical path, while idle lanes still refill at once.
struct WaveBatch {
  int64_t base = 0, end = 0;
  bool exhausted = false;
};
__device__ __forceinline__ int64_t take_work(bool want, WaveBatch& b, unsigned long long* counter, int64_t n) {
  const int lane = threadIdx.x & 63;
  uint64_t m = __ballot(want);
  int64_t got = -1;
  while (m) {
    if (b.base >= b.end) {
      if (b.exhausted) break;
      long long base = 0;
      if (lane == 0) base = (long long)atomicAdd(counter, 64ull);
      base = __shfl(base, 0);
      if (base >= n) {
        b.exhausted = true;
        break;
      }
      b.base = base;
      b.end = base + 64 < n ? base + 64 : n;
    }
    const int64_t avail = b.end - b.base;
    const uint64_t below = lane ? (m & (~0ull >> (64 - lane))) : 0ull;
    const int rank = __popcll(below);
    const bool wanting = (m >> lane) & 1ull;
    if (wanting && rank < avail) got = b.base + rank;
    const int64_t total = __popcll(m);
    b.base += total < avail ? total : avail;
    m = __ballot(wanting && rank >= avail);
  }
  return got;
}

// ---------------------------------------------------------------------------------------
// Nelder-Mead (neldermead_simplex, ftol = 1e-3, NMAX = 5000; nfunc starts at 0)
// ---------------------------------------------------------------------------------------
enum { NM_V = 0, NM_REFL = 1, NM_EXP = 2, NM_CON = 3, NM_SHR = 4 };

// Rounding bound of a difference evaluation (delta_eval): |f - f_exact| <= kDeltaErr 2^-24
// (ds + 4 dk + dl) with ds the support displacements, dk the conditioning term and dl the
// bound of cancelling logarithms; measured on MI355X against double evaluations at the same
// angles over 4.2M random pairs (exponents U[0.6,1.4], steps 1e-7 to 0.5 rad), the largest
// |f - f_exact| / (2^-24 (ds + 4 dk + dl)) was 0.23; with the contact direction forced within
// 1e-7..0.1 of a principal plane of one body, 0.23 for steps above 1e-5 rad, and up to 16 below,
// where the double evaluation itself (cos of an angle within 1e-11 of pi/2) is off by more than
// the difference.  kDeltaErr = 1 keeps a factor 4 over the largest measured ratio.  The
// touching operand cd adds 4 2^-24 f, and the anchor's own double rounding 2^-48 (|q1| + |q2|).
constexpr float kDeltaErr = 1.0f;
// Lanes per wave that the mixed form's resume launch uses.
constexpr int kResumeLanes = 4;
// Mixed form: the main launch re-anchors inline (true) or hands every pair whose evaluation is
// due in double over to the resume launch (false), which keeps the double support map out of
// the main launch and its registers; blocks of 256 per CU of the main launch.  Measured on the
// 200k bed (MI355X): handing over, the main launch takes 0.55 ms at two or three waves per SIMD
// but the resume launch 0.8 ms -- the handed-over solves are chains of evaluations that run
// after the main launch, their length set by the pairs that re-anchor most often -- against
// 0.90 ms for the inline form at two waves per SIMD (1.2 ms at three, where it spills).
#ifndef SDM_NM_INLINE
#define SDM_NM_INLINE 1
#endif
#ifndef SDM_NM_BLOCKS
#define SDM_NM_BLOCKS 2
#endif
constexpr bool kInlineAnchor = SDM_NM_INLINE != 0;
constexpr int kMixedMainBlocks = SDM_NM_BLOCKS;

__device__ __forceinline__ void store_anchor(const SolveIO& S, int64_t w, const DeltaPair& P) {
  const int64_t c = S.cap;
  float* a = S.anc;
  const DeltaBody* b[2] = {&P.b1, &P.b2};
#pragma unroll
  for (int k = 0; k < 2; k++) {
    const int o = 10 * k;
    a[(o + 0) * c + w] = b[k]->in[0]; a[(o + 1) * c + w] = b[k]->in[1]; a[(o + 2) * c + w] = b[k]->in[2];
    a[(o + 3) * c + w] = b[k]->sq0; a[(o + 4) * c + w] = b[k]->sq1;
    a[(o + 5) * c + w] = b[k]->cq0; a[(o + 6) * c + w] = b[k]->cq1;
    a[(o + 7) * c + w] = b[k]->q[0]; a[(o + 8) * c + w] = b[k]->q[1]; a[(o + 9) * c + w] = b[k]->q[2];
  }
  a[20 * c + w] = P.ca0; a[21 * c + w] = P.sa0; a[22 * c + w] = P.ca1; a[23 * c + w] = P.sa1;
  a[24 * c + w] = P.Fh[0]; a[25 * c + w] = P.Fh[1]; a[26 * c + w] = P.Fh[2];
  a[27 * c + w] = P.Fl[0]; a[28 * c + w] = P.Fl[1]; a[29 * c + w] = P.Fl[2];
  a[30 * c + w] = P.cu[0]; a[31 * c + w] = P.cu[1]; a[32 * c + w] = P.cu[2];
}

// The anchor of a queued pair from its stored rows and the two body records.
__device__ __forceinline__ void load_anchor(const SolveIO& S, int64_t w, const GeoRec& r1, const GeoRec& r2,
                                            DeltaPair& P) {
  const int64_t c = S.cap;
  const float* a = S.anc;
  DeltaBody* b[2] = {&P.b1, &P.b2};
  const GeoRec* r[2] = {&r1, &r2};
#pragma unroll
  for (int k = 0; k < 2; k++) {
    const int o = 10 * k;
    b[k]->in[0] = a[(o + 0) * c + w]; b[k]->in[1] = a[(o + 1) * c + w]; b[k]->in[2] = a[(o + 2) * c + w];
    b[k]->sq0 = a[(o + 3) * c + w]; b[k]->sq1 = a[(o + 4) * c + w];
    b[k]->cq0 = a[(o + 5) * c + w]; b[k]->cq1 = a[(o + 6) * c + w];
    b[k]->q[0] = a[(o + 7) * c + w]; b[k]->q[1] = a[(o + 8) * c + w]; b[k]->q[2] = a[(o + 9) * c + w];
    const Shape<double>& s = r[k]->s;
    b[k]->e1 = (float)(1.0 / (2.0 - s.eps1)); b[k]->e2 = (float)(1.0 / (2.0 - s.eps2));
    b[k]->eps1 = (float)s.eps1; b[k]->eps2 = (float)s.eps2; b[k]->t1 = (float)(2.0 - s.eps1);
    b[k]->G = cast9<float>(r[k]->G);
  }
  P.ca0 = a[20 * c + w]; P.sa0 = a[21 * c + w]; P.ca1 = a[22 * c + w]; P.sa1 = a[23 * c + w];
  P.Fh[0] = a[24 * c + w]; P.Fh[1] = a[25 * c + w]; P.Fh[2] = a[26 * c + w];
  P.Fl[0] = a[27 * c + w]; P.Fl[1] = a[28 * c + w]; P.Fl[2] = a[29 * c + w];
  P.cu[0] = a[30 * c + w]; P.cu[1] = a[31 * c + w]; P.cu[2] = a[32 * c + w];
  P.gC = cast9<float>(crotation_mat(r1.pos, r2.pos));
}

// Mixed form, an evaluation the difference form does not take: a new double anchor at
// (b0, b1), whose value and touching test are double.  `anchored` is false where the relative
// form does not apply at that point either (a direction component exactly zero); P is then
// not an anchor and the value not usable.
__device__ __forceinline__ double nm_reanchor(const SolveArgs& A, int64_t w, double b0, double b1, DeltaPair& P,
                                              bool& anchored, bool& touching) {
  const GeoRec r1 = load_rec(A.geo, A.S.id1[w]), r2 = load_rec(A.geo, A.S.id2[w]);
  PairGeo<double> g;
  pair_geo<double>(r1, r2, true, g);
  V3<double> p1, p2;
  double cd = 0.0;
  bool ok = true, t = touching;
  DeltaPair Q;
  const double f = delta_anchor(g, crotation_mat(r1.pos, r2.pos), b0, b1, Q, p1, p2, t, cd, ok);
  // the old anchor is dead from here on (the new one or none): no register keeps it
  P = Q;
  anchored = ok && isfinite(f);
  if (anchored) touching = t;
  return f;
}

// The same, and where no anchor applies, the double distance itself.
__device__ __forceinline__ double nm_mixed_point(const SolveArgs& A, int64_t w, double b0, double b1, DeltaPair& P,
                                                 bool& anchored, bool& touching) {
  const double f = nm_reanchor(A, w, b0, b1, P, anchored, touching);
  if (anchored) return f;
  const GeoRec r1 = load_rec(A.geo, A.S.id1[w]), r2 = load_rec(A.geo, A.S.id2[w]);
  NMGeo<double> n;
  pair_geo<double>(r1, r2, true, n.g);
  n.gC = crotation_mat(r1.pos, r2.pos);
  V3<double> p1, p2;
  int nev = 0;
  return nm_eval<double, true>(n, b0, b1, p1, p2, touching, nev);
}

// T = double: the reference arithmetic (correctly rounded functions), every evaluation in
// double.  T = float: the mixed form -- each evaluation is a float difference from the pair's
// double anchor (delta.h), taken only when it is in the difference form's range and its
// touching test is certified (|c.(p2 - p1)| above the evaluation's rounding bound); any other
// evaluation is done in double at that point, which becomes the new anchor (nm_mixed_point).
// The simplex, its sums and moves stay in double; the vertex comparisons and the convergence
// test are taken on the float values.
//
// The mixed form runs in two launches.  The main one (kResume false) re-anchors inline; a
// pair whose due point admits no anchor (a direction component exactly zero) is handed over
// with its state (the vertices and their values, the sums, the point, the stage and vertex
// indices, the counts), and the lane takes the next pair.  The resume launch (kResume true)
// takes the handed-over pairs -- normally none -- and finishes each solve with the general
// double evaluation inline; its register demand would cut the residency of every wave of
// the main launch.
//
// Residency: two waves per SIMD for the main launches (measured on MI355X: the FP64 form's
// step 17 % faster than at the one wave per SIMD that the unconstrained allocation gives; the
// mixed main launch 19 % faster than at three, where it spills), one for the resume launch,
// whose few pairs want the shortest iteration (no spills).
template <typename T, bool kResume = false>
__global__ __launch_bounds__(256, kResume ? 1 : (sizeof(T) == 4 ? kMixedMainBlocks : 2)) void k_solve_nm(SolveArgs A) {
  constexpr bool kMixed = sizeof(T) == 4;
  const double ftol = 1e-3;
  const double TINY = 1.0e-10;
  const int NMAX = 5000;
  const double U = kDeltaErr * 5.9604644775390625e-08;  // kDeltaErr 2^-24
  NMGeo<double> n;
  DeltaPair P;
  bool anchored = false, due = false;  // mixed form: the pair's anchor is set; a double evaluation is due
  double v[3][2], psum[2], pt[2], anc0 = 0, anc1 = 0;
  double y[3], ysave = 0.0;
  int ihi = 0, ilo = 0, inhi = 0, state = NM_V, vi = 0, nfunc = 0, nev = 0;
  bool touching = true;
  int64_t w = -1;
  V3<double> q1 = mk3(0.0, 0.0, 0.0), q2 = q1;
  bool done = false;
  unsigned long long tot = 0, dbl = 0;
  WaveBatch batch;
  const int64_t nwork = work_n(A);
  const int64_t S = A.S.cap;
  for (;;) {
    const bool idle = (w < 0) && !done;
    int64_t got;
    if constexpr (kResume) {
      // few pairs, each a long serial chain: kResumeLanes lanes per wave, one pair each, so
      // that a wave seldom goes through the double code for one of its lanes
      got = fetch_work(idle && (threadIdx.x & 63) < kResumeLanes, A.next);
      if (got >= nwork) got = -1;
    } else {
      got = take_work(idle, batch, A.next, nwork);
    }
    if (idle) {
      if (got < 0) {
        done = true;
      } else {
        w = A.list ? A.list[got] : got;
        touching = true;
        nev = 0;
        if constexpr (kResume) {
          // the state the main launch handed over, and the double evaluation it left due
          const int32_t code = A.S.fb[w];
          state = code & 7; vi = (code >> 3) & 3; ilo = (code >> 5) & 3; ihi = (code >> 7) & 3; inhi = (code >> 9) & 3;
          v[0][0] = A.S.p1[w]; v[0][1] = A.S.p1[S + w]; v[1][0] = A.S.p1[2 * S + w];
          v[1][1] = A.S.p2[w]; v[2][0] = A.S.p2[S + w]; v[2][1] = A.S.p2[2 * S + w];
          psum[0] = A.S.ctr[w]; psum[1] = A.S.ctr[S + w];
          pt[0] = A.S.la0[w]; pt[1] = A.S.la1[w];
          y[0] = A.S.a0[w]; y[1] = A.S.a1[w]; y[2] = A.S.del[w]; ysave = A.S.f[w];
          nfunc = A.S.stop[w];
          anchored = false;
          due = true;
        } else {
          const GeoRec r1 = load_rec(A.geo, A.S.id1[w]), r2 = load_rec(A.geo, A.S.id2[w]);
          if constexpr (kMixed) {
            anchored = A.S.fb[w] == 0;
            if (anchored) load_anchor(A.S, w, r1, r2, P);
          } else {
            pair_geo<double>(r1, r2, false, n.g);
            n.gC = crotation_mat(r1.pos, r2.pos);
          }
          const double a0 = A.S.a0[w], a1 = A.S.a1[w], del = A.S.del[w];
          anc0 = a0;
          anc1 = a1;
          v[0][0] = a0; v[0][1] = a1;
          v[1][0] = a0 + del; v[1][1] = a1;
          v[2][0] = a0; v[2][1] = a1 + del;
          state = NM_V;
          // the first vertex was evaluated by k_nm_first (touching there): continue at the second
          y[0] = A.S.f[w];
          pt[0] = v[1][0]; pt[1] = v[1][1];
          vi = 1;
          nfunc = 1;
        }
      }
    }
    if (__ballot(!done) == 0) break;
    if (done || w < 0) continue;
    // ---- the one distance evaluation of this iteration ----
    double f = 0.0;
    if constexpr (kMixed) {
      if (!due) {
        nev++;
        bool ok = anchored, t = touching;
        if (anchored) {
          float cd = 0.0f, ds = 0.0f, dk = 0.0f, dl = 0.0f;
          f = delta_eval(P, (float)(pt[0] - anc0), (float)(pt[1] - anc1), t, cd, ds, dk, dl, ok);
          const float qa = fabsf(P.b1.q[0]) + fabsf(P.b1.q[1]) + fabsf(P.b1.q[2]) + fabsf(P.b2.q[0]) +
                           fabsf(P.b2.q[1]) + fabsf(P.b2.q[2]);
          const double ef =
              U * ((double)ds + 4.0 * (double)dk + (double)dl) + 4e-16 * f + 3.552713678800501e-15 * (double)qa;
          ok = ok && fabs((double)cd) > ef + 4.0 * 5.9604644775390625e-08 * f;
        }
        if (ok)
          touching = t;
        else
          due = true;
      }
      if (due) {
        if constexpr (kResume) {
          dbl++;
          f = nm_mixed_point(A, w, pt[0], pt[1], P, anchored, touching);
          anc0 = pt[0];
          anc1 = pt[1];
        } else if constexpr (kInlineAnchor) {
          dbl++;
          f = nm_reanchor(A, w, pt[0], pt[1], P, anchored, touching);
          due = !anchored;
          anc0 = pt[0];
          anc1 = pt[1];
        }
      }
      if constexpr (!kResume) {
        if (due) {
          // hand over: the state before this iteration's evaluation, which is due in double
          A.S.fb[w] = state | (vi << 3) | (ilo << 5) | (ihi << 7) | (inhi << 9);
          A.S.p1[w] = v[0][0]; A.S.p1[S + w] = v[0][1]; A.S.p1[2 * S + w] = v[1][0];
          A.S.p2[w] = v[1][1]; A.S.p2[S + w] = v[2][0]; A.S.p2[2 * S + w] = v[2][1];
          A.S.ctr[w] = psum[0]; A.S.ctr[S + w] = psum[1];
          A.S.la0[w] = pt[0]; A.S.la1[w] = pt[1];
          A.S.a0[w] = y[0]; A.S.a1[w] = y[1]; A.S.del[w] = y[2]; A.S.f[w] = ysave;
          A.S.stop[w] = nfunc;
          A.S.nev[w] += nev;
          tot += nev;
          A.hoList[atomicAdd(A.hoCount, 1ull)] = (int32_t)w;
          due = false;
          w = -1;
          continue;
        }
      }
      due = false;
    } else {
      f = nm_eval<double>(n, pt[0], pt[1], q1, q2, touching, nev);
    }
    nfunc++;
    bool finished = false;
    double fmin = 0.0;
    bool decide = false;
    if (state == NM_V) {
      y[vi] = f;
      if (!touching) {
        fmin = y[0];
        finished = true;
      } else if (vi < 2) {
        vi++;
        pt[0] = v[vi][0];
        pt[1] = v[vi][1];
      } else {
        psum[0] = v[0][0] + v[1][0] + v[2][0];
        psum[1] = v[0][1] + v[1][1] + v[2][1];
        decide = true;
      }
    } else if (state == NM_SHR) {
      y[vi] = f;
      if (!touching) {
        fmin = y[0];
        finished = true;
      } else {
        int nx = vi + 1;
        if (nx == ilo) nx++;
        if (nx < 3) {
          vi = nx;
          psum[0] = 0.5 * (v[vi][0] + v[ilo][0]);
          psum[1] = 0.5 * (v[vi][1] + v[ilo][1]);
          v[vi][0] = psum[0];
          v[vi][1] = psum[1];
          pt[0] = psum[0];
          pt[1] = psum[1];
        } else {
          psum[0] = v[0][0] + v[1][0] + v[2][0];
          psum[1] = v[0][1] + v[1][1] + v[2][1];
          decide = true;
        }
      }
    } else {
      // amotry: accept the trial point if it beats the highest vertex
      const double ytry = f;
      if (ytry < y[ihi]) {
        y[ihi] = ytry;
        psum[0] += pt[0] - v[ihi][0];
        psum[1] += pt[1] - v[ihi][1];
        v[ihi][0] = pt[0];
        v[ihi][1] = pt[1];
      }
      if (!touching) {
        fmin = y[0];
        finished = true;
      } else if (state == NM_REFL) {
        if (ytry <= y[ilo]) {
          const double fac = 2.0, fac1 = (1.0 - fac) / 2.0, fac2 = fac1 - fac;
          pt[0] = psum[0] * fac1 - v[ihi][0] * fac2;
          pt[1] = psum[1] * fac1 - v[ihi][1] * fac2;
          state = NM_EXP;
        } else if (ytry >= y[inhi]) {
          ysave = y[ihi];
          const double fac = 0.5, fac1 = (1.0 - fac) / 2.0, fac2 = fac1 - fac;
          pt[0] = psum[0] * fac1 - v[ihi][0] * fac2;
          pt[1] = psum[1] * fac1 - v[ihi][1] * fac2;
          state = NM_CON;
        } else {
          decide = true;
        }
      } else if (state == NM_EXP) {
        decide = true;
      } else {  // NM_CON
        if (ytry >= ysave) {
          vi = (ilo == 0) ? 1 : 0;
          psum[0] = 0.5 * (v[vi][0] + v[ilo][0]);
          psum[1] = 0.5 * (v[vi][1] + v[ilo][1]);
          v[vi][0] = psum[0];
          v[vi][1] = psum[1];
          pt[0] = psum[0];
          pt[1] = psum[1];
          state = NM_SHR;
        } else {
          decide = true;
        }
      }
    }
    if (decide) {
      // test of doneness (:1779-1806) and the reflection of the next iteration (:1814)
      ilo = 0;
      if (y[0] > y[1]) { ihi = 0; inhi = 1; } else { ihi = 1; inhi = 0; }
      for (int i = 0; i < 3; i++) {
        if (y[i] <= y[ilo]) ilo = i;
        if (y[i] > y[ihi]) {
          inhi = ihi;
          ihi = i;
        } else if (y[i] > y[inhi] && i != ihi)
          inhi = i;
      }
      const double yh = y[ihi], yl = y[ilo];
      const double rtol = 2.0 * fabs(yh - yl) / (fabs(yh) + fabs(yl) + TINY);
      if (rtol < ftol) {
        const double t = y[0]; y[0] = y[ilo]; y[ilo] = t;
        const double t0 = v[ilo][0], t1 = v[ilo][1];
        v[ilo][0] = v[0][0]; v[ilo][1] = v[0][1];
        v[0][0] = t0; v[0][1] = t1;
        fmin = y[0];
        finished = true;
      } else if (nfunc >= NMAX) {
        fmin = y[0];
        finished = true;
      } else {
        const double fac = -1.0, fac1 = (1.0 - fac) / 2.0, fac2 = fac1 - fac;
        pt[0] = psum[0] * fac1 - v[ihi][0] * fac2;
        pt[1] = psum[1] * fac1 - v[ihi][1] * fac2;
        state = NM_REFL;
      }
    }
    if (finished) {
      A.S.a0[w] = v[0][0];
      A.S.a1[w] = v[0][1];
      A.S.f[w] = fmin;
      A.S.la0[w] = pt[0];
      A.S.la1[w] = pt[1];
      if (!kMixed) {  // the mixed form's points come from the double refinement
        st3s(A.S.p1, w, S, q1);
        st3s(A.S.p2, w, S, q2);
      }
      A.S.stop[w] = nfunc;
      A.S.touch[w] = touching ? 1 : 0;
      A.S.nev[w] += nev;
      tot += nev;
      w = -1;
    }
  }
  add_evals(A.evals, tot);
  if constexpr (kMixed) {
    if (A.dbl) add_evals(A.dbl, dbl);
  }
}

// Expected evaluation count of a queued pair, used to start the long solves first.
__device__ __forceinline__ uint32_t nm_cost_key(int prevEval, bool prevContact, double del) {
  const int e = del > 0.01 ? 80 : (prevContact ? prevEval : 30);
  return 255u - (uint32_t)(e < 0 ? 0 : (e > 255 ? 255 : e));
}

// The first distance evaluation of every queued pair (the first simplex vertex,
// neldermead_simplex :1731-1737), one pair per lane.  Pairs that are not touching there are
// complete after it (the caller stores their outcome); the others are listed for the solver
// with their first value in S.f.  Returns whether the pair is listed.
// T = double: the reference arithmetic.  T = float (mixed form): the double anchor of the
// difference evaluations, with the device's double functions (S.fb = 1: the difference form
// does not apply at the first vertex, whose distance is then evaluated in double without an
// anchor).  Only touching pairs store their anchor.
// Certified "not touching" at the first vertex from a float evaluation.  The reference stores
// nothing of a pair that is not touching at its first vertex but that outcome (depth 0, new
// shear, the start angles), so a float evaluation whose touching operand cd = c.(p2 - p1)
// exceeds its error bound decides it exactly.  cd = c.x21 - h1(c) - h2(-c) with h the support
// functions: its gradient in c is bounded by |x21| + |q1| + |q2|, so a direction rounded to
// float (angles of magnitude |a| carry 2^-24 (|a| + 1) each) moves it by a few 2^-24
// (|x21| + r1 + r2) (|a0| + |a1| + 4); the float support points (a few ulp of r each, fpow and
// the rotation), the centre difference and the product add a few 2^-24 (|x21| + r1 + r2).
// The bound takes 256 2^-24 (|x21|_1 + 2 (r1 + r2)) (|a0| + |a1| + 4), over 10x these.  A NaN
// anywhere fails the test.  minTrig > 0: no decision where a cos / sin factor of a support-map
// angle is below it (the Levenberg-Marquardt workloads' exponents up to 1.9 make the
// reference's own double support map depart from the exact one there: cos of an atan2 angle
// within ~1e-16 of pi/2, raised to 1/(2 - eps)).
__device__ __forceinline__ bool nm_first_apart(const GeoRec& r1, const GeoRec& r2, const M3<double>& gC, double a0,
                                               double a1, float minTrig = 0.0f) {
  NMGeo<float> n;
  pair_geo<float>(r1, r2, true, n.g);
  n.gC = cast9<float>(gC);
  nm_prepare(n);
  float s0, c0, s1, c1;
  sincosf((float)a0, &s0, &c0);
  sincosf((float)a1, &s1, &c1);
  const V3<float> c = mul(n.gC, mk3(c0 * c1, s0 * c1, s1));
  float m1 = 1.0f, m2 = 1.0f;
  const V3<float> p1 = support_f(n.f1, n.g.G1, c, minTrig > 0.0f ? &m1 : nullptr);
  const V3<float> p2 = support_f(n.f2, n.g.G2, mk3(-c.x, -c.y, -c.z), minTrig > 0.0f ? &m2 : nullptr) + n.g.pos2;
  if (!(fminf(m1, m2) >= minTrig)) return false;
  const float cd = dot(c, p2 - p1);
  const V3<double> x21 = r2.pos - r1.pos;
  const double scale = (fabs(x21.x) + fabs(x21.y) + fabs(x21.z) + 2.0 * (r1.rmax + r2.rmax)) *
                       (fabs(a0) + fabs(a1) + 4.0);
  return (double)cd > 256.0 * 5.9604644775390625e-08 * scale;
}

template <typename T>
__device__ __forceinline__ bool nm_first_eval(const SolveArgs& A, int64_t w) {
  const GeoRec r1 = load_rec(A.geo, A.S.id1[w]), r2 = load_rec(A.geo, A.S.id2[w]);
  bool touching = true;
  if (nm_first_apart(r1, r2, crotation_mat(r1.pos, r2.pos), A.S.a0[w], A.S.a1[w])) {
    A.S.f[w] = 0.0;  // not read for a pair that is not touching
    if constexpr (sizeof(T) == 4) A.S.fb[w] = 0;
    A.S.nev[w] = 1;
    A.S.stop[w] = 1;
    A.S.touch[w] = 0;
    return false;
  }
  if constexpr (sizeof(T) == 4) {
    PairGeo<double> g;
    pair_geo<double>(r1, r2, true, g);
    const M3<double> gC = crotation_mat(r1.pos, r2.pos);
    const double a0 = A.S.a0[w], a1 = A.S.a1[w];
    DeltaPair P;
    V3<double> p1, p2;
    double cd = 0.0;
    bool ok = true;
    double f = delta_anchor(g, gC, a0, a1, P, p1, p2, touching, cd, ok);
    ok = ok && isfinite(f);
    if (!ok) {
      NMGeo<double> n;
      n.g = g;
      n.gC = gC;
      int nev = 0;
      touching = true;
      f = nm_eval<double, true>(n, a0, a1, p1, p2, touching, nev);
    } else if (touching) {
      store_anchor(A.S, w, P);
    }
    A.S.f[w] = f;
    A.S.fb[w] = ok ? 0 : 1;
    A.S.nev[w] = 1;
    A.S.stop[w] = 1;
    A.S.touch[w] = touching ? 1 : 0;
    return touching;
  } else {
    NMGeo<double> n;
    pair_geo<double>(r1, r2, false, n.g);
    n.gC = crotation_mat(r1.pos, r2.pos);
    V3<double> q1, q2;
    int nev = 0;
    const double f = nm_eval<double>(n, A.S.a0[w], A.S.a1[w], q1, q2, touching, nev);
    A.S.f[w] = f;
    A.S.nev[w] = 1;
    A.S.stop[w] = 1;
    A.S.touch[w] = touching ? 1 : 0;
    return touching;
  }
}

// Mixed form, final refinement in double: the depth is the double distance at the best
// vertex and the points (hence the normal) those of the last evaluated point, as the
// reference takes them (Superquadrics_Ig2.cpp:2439-2454).  One work item per listed pair.
// The evaluation uses the device's double functions (double accuracy, not the FP64 form's
// correctly rounded last bits).
// A double evaluation at (b0, b1) for the refinement: the anchor's formula (delta_anchor:
// powers as exp2(b log2 a), the angles' cos / sin as ratios, no atan2 or sin / cos of them),
// the general evaluation where that formula does not apply (a direction component zero).
__device__ __forceinline__ double nm_double_point(const NMGeo<double>& n, double b0, double b1, V3<double>& p1,
                                                  V3<double>& p2, int& nev) {
  DeltaPair P;
  bool touching = true, ok = true;
  double cd = 0.0;
  nev++;
  const double f = delta_anchor(n.g, n.gC, b0, b1, P, p1, p2, touching, cd, ok);
  if (ok && isfinite(f)) return f;
  nev--;
  return nm_eval<double, true>(n, b0, b1, p1, p2, touching, nev);
}

__global__ __launch_bounds__(256) void k_nm_refine(SolveArgs A) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= work_n(A)) return;
  const int64_t w = A.list ? A.list[k] : k;
  if (!A.S.touch[w]) return;
  NMGeo<double> n;
  const GeoRec r1 = load_rec(A.geo, A.S.id1[w]), r2 = load_rec(A.geo, A.S.id2[w]);
  pair_geo<double>(r1, r2, false, n.g);
  n.gC = crotation_mat(r1.pos, r2.pos);
  V3<double> q1, q2;
  int nev = 0;
  const double a0 = A.S.a0[w], a1 = A.S.a1[w], l0 = A.S.la0[w], l1 = A.S.la1[w];
  const double f = nm_double_point(n, a0, a1, q1, q2, nev);
  // the last point is often the best vertex: its points are those just computed
  if (__double_as_longlong(l0) != __double_as_longlong(a0) || __double_as_longlong(l1) != __double_as_longlong(a1))
    nm_double_point(n, l0, l1, q1, q2, nev);
  const int64_t S = A.S.cap;
  A.S.f[w] = f;
  st3s(A.S.p1, w, S, q1);
  st3s(A.S.p2, w, S, q2);
  A.S.nev[w] += nev;
}

// ---------------------------------------------------------------------------------------
// Levenberg-Marquardt (lmder2 with opts {1, 1e-15, 1e-3, 1e-20}, itmax 50)
// ---------------------------------------------------------------------------------------
enum { LM_INIT = 0, LM_J1 = 2, LM_J2 = 3, LM_TRIAL = 4 };  // f(p) of the Jacobian (1) is never evaluated: see k_solve_lm

// Smallest cos / sin factor of a support-map angle at which the certified float screen of a
// pair's first Levenberg-Marquardt evaluation decides (k_lm_first).  The reference takes cos and
// sin of the rounded atan2 angles (Normal2Phi, getSurface), which carry an absolute error of
// ~1e-16 that the float support map does not have (it forms the factors as ratios): near a
// principal plane -- an exponent close to 2 makes |rx nx|^(1/(2 - eps)) vanish against
// |ry ny|^(1/(2 - eps)) -- the reference's support point parts from the exact one by up to tens
// of percent.  Above 1e-6 that error is below 1e-10 relative in a factor, ~2e-9 after the
// largest exponent 20.
constexpr double kLmMinTrig = 1e-6;

// kFast: the device's double functions (the mixed form); otherwise correctly rounded (FP64).
// Two adjacent lanes take each pair, each holding one body: both run the pair's
// state machine (the same arithmetic on the same values, so they stay in step) and each
// evaluates its own body's support point, which they exchange -- p2 is body 2's support along
// -c (Disfunction's mulT(G2, c) negated is mulT(G2, -c) exactly).  An evaluation then takes
// half the instructions per lane, so a wave's iteration half the time: the launch's length is
// set by its longest chains of dependent evaluations (~250 for the solves that reach itmax),
// which start when the queue has drained.  Measured on MI355X (1M-body periodic cell, FP64):
// main launch 12.0 -> 9.7 ms/step against one lane per pair; restarts 7.5 -> 5.5 ms/step.
// kWaves: resident waves per SIMD the register allocation is held to.
// A.lmRestarts: the two restarts of Ig2_..._SuperquadricsGeom::go
// (Superquadrics_Ig2.cpp:2013-2052) run in the lane pair as soon as its solve ends -- kind 1
// (NaN parameters or stop 7) once after the first solve, kind 2 (a deep contact, 10 |p2 - p1|
// > r_max of body 1, whose points are not inside the other body) after the first solve or the
// kind-1 restart's, each a fresh solve from the line of centres + 1e-5 with the stop code
// tagged 100 / 200 -- so their long chains start during the launch instead of in launches
// after it (1M periodic cell, FP64: main launch + restarts 15.1 -> 13.7 ms/step).
template <bool kFast, int kWaves = 1>
__global__ __launch_bounds__(256, kWaves) void k_solve_lm(SolveArgs A) {
  const bool odd = (threadIdx.x & 1) != 0;  // body 2's lane of the pair
  int phase = 0;          // A.lmRestarts: 0 the first solve, 1 after a kind-1 restart, 2 after a kind-2
  double myRmax = 0.0;    // A.lmRestarts: this lane's body's r_max
  const double tau = 1.0, eps1 = 1e-15, eps2 = 1e-3, eps3 = 1e-20;
  const int itmax = 50;
  PairGeo<double> g;  // this lane's body in s1, G1, pos1
  double p0 = 0, p1v = 0, Dp0 = 0, Dp1 = 0, mu = 0, p_fL2 = 0, A00 = 0, A01 = 0, A11 = 0, dA0 = 0, dA1 = 0;
  double jTf0 = 0, jTf1 = 0, p_L2 = 0, q0 = 0, qq1 = 0;
  V3<double> p_f = mk3(0.0, 0.0, 0.0), f0 = p_f, f1 = p_f, ct = p_f, lp1 = p_f, lp2 = p_f, lc = p_f;
  int nu = 2, stop = 0, k = 0, state = LM_INIT, nev = 0;
  bool stop_flag = false, touching = true;
  int64_t w = -1;
  bool done = false;
  unsigned long long tot = 0;
  WaveBatch batch;
  const int64_t nwork = work_n(A);
  for (;;) {
    const bool idle = (w < 0) && !done;
    int64_t got = take_work(idle && !odd, batch, A.next, nwork);
    got = __shfl((long long)got, (int)(threadIdx.x & 62));
    if (idle) {
      if (got < 0) {
        done = true;
      } else {
        w = A.list ? A.list[got] : got;
        // this lane's body in g.s1, g.G1, g.pos1; position2 = pos + shift2 (:1886)
        GeoRec r = load_rec(A.geo, odd ? A.S.id2[w] : A.S.id1[w]);
        if (odd && A.S.sh) r.pos = r.pos + ld3s(A.S.sh, w, A.S.cap);
        g.s1 = r.s;
        g.G1 = r.G;
        g.pos1 = r.pos;
        myRmax = r.rmax;
        phase = 0;
        p0 = A.S.a0[w];
        p1v = A.S.a1[w];
        ct = ld3s(A.S.ct, w, A.S.cap);
        Dp0 = 1e-7; Dp1 = 1e-7;
        mu = 0; nu = 2; stop = 0; k = 0; nev = 0;
        stop_flag = false;
        touching = true;
        state = LM_INIT;
      }
    }
    if (__ballot(!done) == 0) break;
    if (done) continue;
    // ---- the one distance evaluation of this iteration ----
    double e0 = p0, e1 = p1v;
    if (state == LM_J1) e0 = p0 + Dp0;
    if (state == LM_J2) e1 = p1v + Dp1;
    if (state == LM_TRIAL) { e0 = q0; e1 = qq1; }
    V3<double> fe;
    {  // Disfunction (lm_eval) split over the lane pair
      nev++;
      double sa0, ca0, sa1, ca1;
      msincos<kFast>(e0, sa0, ca0);
      msincos<kFast>(e1, sa1, ca1);
      const V3<double> contact = mk3(ca0 * ca1, sa0 * ca1, sa1);
      const V3<double> dir = odd ? mk3(-contact.x, -contact.y, -contact.z) : contact;
      const V3<double> q = support<double, kFast>(g.s1, g.G1, dir) + g.pos1;
      const V3<double> o = mk3(__shfl_xor(q.x, 1), __shfl_xor(q.y, 1), __shfl_xor(q.z, 1));
      lp1 = odd ? o : q;
      lp2 = odd ? q : o;
      lc = mk3(0.5 * (lp1.x + lp2.x), 0.5 * (lp1.y + lp2.y), 0.5 * (lp1.z + lp2.z));
      fe = lp2 - lp1;
      if (dot(contact, fe) > 0.0) stop_flag = true;
    }
    bool finished = false;
    bool outer = false;  // go to the outer-iteration checks
    bool solve = false;  // go to the damped normal-equation solve
    if (state == LM_INIT) {
      p_f = fe;
      p_fL2 = norm(p_f);
      if (!isfinite(p_fL2)) {
        stop = 7;
        finished = true;
      } else {
        outer = true;
      }
    } else if (state == LM_J1) {
      f1 = fe;
      state = LM_J2;
    } else if (state == LM_J2) {
      const V3<double> f2 = fe;
      const double J00 = (f1.x - f0.x) / Dp0, J10 = (f1.y - f0.y) / Dp0, J20 = (f1.z - f0.z) / Dp0;
      const double J01 = (f2.x - f0.x) / Dp1, J11 = (f2.y - f0.y) / Dp1, J21 = (f2.z - f0.z) / Dp1;
      A00 = J00 * J00 + (J10 * J10 + J20 * J20);
      A01 = J00 * J01 + (J10 * J11 + J20 * J21);
      A11 = J01 * J01 + (J11 * J11 + J21 * J21);
      const V3<double> mf = mk3(-1.0 * p_f.x, -1.0 * p_f.y, -1.0 * p_f.z);
      jTf0 = J00 * mf.x + (J10 * mf.y + J20 * mf.z);
      jTf1 = J01 * mf.x + (J11 * mf.y + J21 * mf.z);
      const double jTf_inf = (fabs(jTf0) > fabs(jTf1) ? fabs(jTf0) : fabs(jTf1));
      p_L2 = sqrt(p0 * p0 + p1v * p1v);
      if (jTf_inf <= eps1) {
        stop = 1;
        finished = true;
      } else {
        if (k == 0) mu = tau * (A00 > A11 ? A00 : A11);
        dA0 = A00;
        dA1 = A11;
        solve = true;
      }
    } else {  // LM_TRIAL
      const V3<double> pDp_f = fe;
      const double pDp_fL2 = norm(pDp_f);
      if (!isfinite(pDp_fL2)) {
        stop = 7;
        k++;
        finished = true;
      } else {
        const double dL = Dp0 * (mu * Dp0 + jTf0) + Dp1 * (mu * Dp1 + jTf1);
        const double dF = p_fL2 - pDp_fL2;
        double tmp = dF / dL;
        if (tmp > 0.0) {
          tmp = 1.0 - mpow<kFast>(2 * tmp - 1., 3.);
          mu = mu * ((tmp >= 0.3333333333333334) ? tmp : 0.3333333333333334);
          nu = 2;
          p0 = q0;
          p1v = qq1;
          p_f = pDp_f;
          p_fL2 = pDp_fL2;
          k++;
          outer = true;
        } else {
          mu *= nu;
          const int nu2 = int(unsigned(nu) << 1);
          if (nu2 <= nu) {
            stop = 5;
            k++;
            finished = true;
          } else {
            nu = nu2;
            A00 = dA0;
            A11 = dA1;
            solve = true;
          }
        }
      }
    }
    // Damped normal equations without further evaluations; loops until it needs an evaluation.
    if (solve) {
      A00 += mu;
      A11 += mu;
      const double n_a = A00 * A11 - A01 * A01;
      Dp0 = (jTf0 * A11 - jTf1 * A01) / n_a;
      Dp1 = (jTf1 * A00 - jTf0 * A01) / n_a;
      const double Dp_L2 = sqrt(Dp0 * Dp0 + Dp1 * Dp1);
      if (Dp_L2 < 1e-5) {
        stop = 2;
        k++;
        finished = true;
      } else {
        q0 = p0 + Dp0;
        qq1 = p1v + Dp1;
        if (Dp_L2 >= (p_L2 + eps2) / (1E-12 * 1E-12)) {
          stop = 4;
          k++;
          finished = true;
        } else {
          state = LM_TRIAL;
        }
      }
    }
    if (outer) {
      if (k >= itmax) {
        finished = true;
      } else if (p_fL2 <= eps3) {
        stop = 6;
        finished = true;
      } else if (stop_flag) {
        stop = 9;
        touching = false;
        finished = true;
      } else {
        double sp0, cp0, sp1, cp1;
        msincos<kFast>(p0, sp0, cp0);
        msincos<kFast>(p1v, sp1, cp1);
        ct = mk3(cp0 * cp1, sp0 * cp1, sp1);
        const double angle = dot(p_f, ct) / (norm(p_f) * norm(ct));
        if (fabs(angle) > 0.95) {
          stop = 8;
          finished = true;
        } else {
          // JacfunctionApp's f(p) (Superquadrics_Ig2.cpp:1037) re-evaluates the point evaluated
          // last -- the start point or the accepted trial, whose result is p_f, lp1, lp2 and
          // lc -- at the same parameters: the same values, so it is counted and taken from there
          f0 = p_f;
          nev++;
          state = LM_J1;
        }
      }
    }
    if (finished && A.lmRestarts && touching && phase < 2) {
      if (k >= itmax) stop = 3;
      // kind 1: NaN parameters or stop 7 (checked once, after the first solve); kind 2: a deep
      // contact whose points are not inside the other body (after the first solve or the
      // kind-1 restart's)
      // (every lane shuffle here runs in both lanes of the pair: a lane reads its partner's
      // register, which must have been written by the same instruction)
      const double rmax1 = __shfl(myRmax, (int)(threadIdx.x & 62));
      int kind = 0;
      if (phase == 0 && (isnan(p0) || isnan(p1v) || stop == 7)) {
        kind = 1;
      } else if (10.0 * norm(lp2 - lp1) > rmax1) {
        const bool in = is_inside(g.s1, g.G1, (odd ? lp1 : lp2) - g.pos1);  // p2 in body 1, p1 in body 2
        const bool inOther = __shfl_xor((int)in, 1) != 0;
        if (!in || !inOther) kind = 2;
      }
      if (kind) {
        const V3<double> o = mk3(__shfl_xor(g.pos1.x, 1), __shfl_xor(g.pos1.y, 1), __shfl_xor(g.pos1.z, 1));
        const V3<double> d0 = (odd ? g.pos1 : o) - (odd ? o : g.pos1);  // (position2 + shift2) - position1
        p0 = xatan2(d0.y, d0.x) + 1e-5;
        p1v = xatan(d0.z / sqrt(d0.x * d0.x + d0.y * d0.y)) + 1e-5;
        if (!odd) A.S.tag[w] = kind == 1 ? 100 : 200;
        phase = kind;
        Dp0 = 1e-7; Dp1 = 1e-7;
        mu = 0; nu = 2; stop = 0; k = 0;
        stop_flag = false;
        state = LM_INIT;
        finished = false;  // the evaluation count runs on: the pair's total is the passes' sum
      }
    }
    if (finished) {
      if (k >= itmax) stop = 3;
      if (!odd) {
        const int64_t S = A.S.cap;
        A.S.a0[w] = p0;
        A.S.a1[w] = p1v;
        st3s(A.S.p1, w, S, lp1);
        st3s(A.S.p2, w, S, lp2);
        st3s(A.S.ctr, w, S, lc);
        st3s(A.S.ct, w, S, ct);
        A.S.nev[w] += nev;
        A.S.stop[w] = stop;
        A.S.touch[w] = touching ? 1 : 0;
        tot += nev;
      }
      w = -1;
    }
  }
  add_evals(A.evals, tot);
  if (A.dbl) add_evals(A.dbl, tot);  // the mixed form: every evaluation in double
}
