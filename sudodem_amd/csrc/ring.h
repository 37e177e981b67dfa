// ring.h -- insertionSortPeri over the bound rings of a periodic cell, as a chain of phases.
//
// InsertionSortCollider keeps, per axis, a vector of bounds (id, isMin, period, coord =
// bound - L * period) sorted as a ring: the element at loIdx is the ring's first, the
// comparison of an element with its predecessor adds L when the element sits at loIdx, and a
// bound that moves across loIdx changes its period by one and its coordinate by L
// (pkg/common/InsertionSortCollider.cpp:416-463).  Every later collider run walks the vector
// once from index 0 (not from loIdx) and insertion-sorts it in place; a lower bound that moves
// backwards over an upper bound calls handleBoundInversionPeri (:466-475).  The walk from
// index 0 makes the first elements' insertions continue backwards over the vector's end,
// through elements the walk has not reached yet; an inversion such a move leaves between the
// vector's last and first elements waits for the next run (the reference creates those pairs
// one run late).
//
// Device form: the walk over positions 0 .. n2-1 is cut into five consecutive phases, run in
// that order, so each phase sees exactly the vector the reference walk has when it gets there:
//   [0, P)   sequential: the reference loop itself (one thread per axis), wrap included;
//   [P, A)   parallel:   insertions into a sorted prefix = a stable sort;
//   [A, B)   sequential: the positions around loIdx;
//   [B, Q)   parallel;
//   [Q, n2)  sequential: the end of the vector, which the first phase may have reshuffled.
// A parallel phase inserts the elements of [I0, I1) into the prefix before I0.  Insertion
// sort inserts each element behind the last element before it that is not greater, so if
// the window [W0, I0) of the prefix is sorted, no inserted element is smaller than both the
// element at W0 and the one at W0 - 1 (exact comparisons; it then stops inside the window or
// on top of W0 - 1), and loIdx lies outside [W0 - 1, I1], then the phase is the stable sort of
// [W0, I1) by coordinate (the ring's +L and period moves happen only at loIdx), and the
// inversions are its crossings: an inserted lower bound passes the upper bounds that precede
// it in position and follow it after the sort, the nearest first.  The device checks these
// conditions when it runs the phase; a run that fails them is undone and repeated with
// longer sequential phases.
#pragma once
#include <stdint.h>

#ifdef __HIPCC__
#define SDM_RING_HD __host__ __device__
#else
#define SDM_RING_HD
#endif

struct RingAxis {
  int32_t* h;     // id << 1 | 1 for an upper bound (the collider.bounds encoding)
  int32_t* per;   // period of the bound
  double* coord;  // bound - L * period, then the walk's own updates
  int64_t n2;
  double L;
};

enum { RING_OK = 0, RING_RUNAWAY = 1 };

// The reference walk over positions [s, e) of one axis (InsertionSortCollider.cpp:424-462),
// with loIdx in `lo` (read and updated).  emit(i, step, id1, id2) receives each inversion of
// a lower bound over an upper bound: i the position being inserted, step the number of its
// inversions before this one.
template <class Emit>
SDM_RING_HD int ring_walk(const RingAxis& R, int64_t s, int64_t e, int64_t& lo, Emit& emit) {
  const int64_t n2 = R.n2;
  const double L = R.L;
  int32_t* H = R.h;
  int32_t* P = R.per;
  double* X = R.coord;
  for (int64_t i = s; i < e; i++) {
    if (i == lo && X[i] < 0) {
      P[i] -= 1;
      X[i] += L;
      lo = lo + 1 == n2 ? 0 : lo + 1;
    }
    const int64_t i_1 = i == 0 ? n2 - 1 : i - 1;
    const double iCmp = X[i] + (i == lo ? L : 0.0);
    if (X[i_1] <= iCmp) continue;
    int64_t j = i_1;
    double vC = X[i];
    const int32_t vH = H[i];
    int32_t vP = P[i];
    const bool isMin = !(vH & 1);
    uint32_t step = 0;
    int64_t moves = 0;
    while (true) {
      const int64_t j1 = j + 1 == n2 ? 0 : j + 1;
      if (!(X[j] > vC + (j1 == lo ? L : 0.0))) break;
      X[j1] = X[j];
      H[j1] = H[j];
      P[j1] = P[j];
      if (j == lo && vC < 0) {
        vP -= 1;
        vC += L;
        lo = lo + 1 == n2 ? 0 : lo + 1;
      } else if (j1 == lo) {
        P[j1] += 1;
        X[j1] -= L;
        lo = lo == 0 ? n2 - 1 : lo - 1;
      }
      if (isMin && (H[j] & 1)) {
        const int32_t other = H[j1] >> 1;
        if ((vH >> 1) != other) emit(i, step++, vH >> 1, other);
      }
      j = j == 0 ? n2 - 1 : j - 1;
      if (++moves >= n2) return RING_RUNAWAY;  // a ring whose coordinates are not finite
    }
    const int64_t k = j + 1 == n2 ? 0 : j + 1;
    X[k] = vC;
    H[k] = vH;
    P[k] = vP;
  }
  return RING_OK;
}
