// Host build of the periodic collider's ring walk (sudodem_amd/csrc/ring.h) for the CPU tests:
// the reference walk over a whole vector, and the device's five-phase plan (sequential
// phases with ring_walk, parallel phases as a stable sort plus crossings, the device's
// checks and extensions) written for the host.  Test infrastructure only.
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../sudodem_amd/csrc/ring.h"

namespace {

struct Ev {
  uint64_t key;
  int32_t a, b;
};

struct Emit {
  std::vector<Ev>* out;
  void operator()(int64_t i, uint32_t step, int32_t a, int32_t b) {
    out->push_back(Ev{((uint64_t)i << 31) | step, a, b});
  }
};

// Order-preserving key of a double (-0 and +0 together), as the device sorts by.
uint64_t key_of(double x) {
  x = x + 0.0;
  uint64_t u;
  std::memcpy(&u, &x, 8);
  return (u & 0x8000000000000000ull) ? ~u : (u | 0x8000000000000000ull);
}

void put(const std::vector<Ev>& E, int64_t* ev, int64_t cap) {
  for (size_t k = 0; k < E.size() && (int64_t)k < cap; k++) {
    ev[3 * k] = (int64_t)E[k].key;
    ev[3 * k + 1] = E[k].a;
    ev[3 * k + 2] = E[k].b;
  }
}

}  // namespace

extern "C" {

// The reference walk over [0, n2): loIdx in *lo (updated); inversions (key, id1, id2) in ev.
// Returns the number of inversions, or -1 when an insertion ran round the ring.
int64_t ring_whole(int32_t* h, int32_t* per, double* x, int64_t n2, double L, int64_t* lo, int64_t* ev, int64_t cap) {
  std::vector<Ev> E;
  Emit em{&E};
  if (ring_walk(RingAxis{h, per, x, n2, L}, 0, n2, *lo, em) != RING_OK) return -1;
  put(E, ev, cap);
  return (int64_t)E.size();
}

// The five phases with start / end length K0, half width KL around loIdx, prefix window W.
// stats[0]: sequential extensions, stats[1]: parallel phases run as sorts.
int64_t ring_phases(int32_t* h, int32_t* per, double* x, int64_t n2, double L, int64_t* lo, int64_t K0, int64_t KL,
                    int64_t W, int64_t* ev, int64_t cap, int64_t* stats) {
  std::vector<Ev> E;
  Emit em{&E};
  stats[0] = stats[1] = 0;
  RingAxis R{h, per, x, n2, L};
  const int64_t lo0 = *lo;
  const int64_t P = std::min(K0, n2), Q = std::max(n2 - K0, P);
  const int64_t A = std::min(std::max(lo0 - KL, P), Q), B = std::min(std::max(lo0 + KL, A), Q);
  int64_t l = lo0;
  bool bad = false;
  auto seq = [&](int64_t s, int64_t e) {
    if (e > s && ring_walk(R, s, e, l, em) != RING_OK) bad = true;
  };
  auto par = [&](int64_t cur, int64_t end) {
    int64_t r0 = 0, ws = 0;
    while (cur < end) {  // the check, and extensions of the sequential phase before
      ws = std::max<int64_t>(1, cur - W);
      int64_t base = ws;
      if (l < cur && l + 2 > base) base = l + 2;
      r0 = base;
      for (int64_t k = base + 1; k < cur; k++)
        if (x[k - 1] > x[k]) r0 = k;
      uint64_t mn = ~0ull;
      for (int64_t k = cur; k < end; k++) mn = std::min(mn, key_of(x[k]));
      int64_t ext = 0;
      if (base > cur || (l >= r0 - 1 && l <= end)) {
        ext = std::min(end, l + 64);
      } else {
        uint64_t bottom = key_of(x[r0 - 1]);
        if (r0 < cur) bottom = std::min(bottom, key_of(x[r0]));
        if (mn < bottom)
          for (int64_t k = cur; k < end; k++)
            if (key_of(x[k]) < bottom) ext = k + 1;
      }
      if (ext == 0) break;
      const int64_t ne = std::min(end, std::max(ext, cur + 64));
      seq(cur, ne);
      cur = ne;
      stats[0]++;
    }
    if (cur >= end) return;
    stats[1]++;
    const int64_t m = end - ws, ins = cur - ws;
    std::vector<uint64_t> key(m);
    std::vector<uint32_t> idx(m);
    for (int64_t t = 0; t < m; t++) {
      key[t] = ws + t < r0 ? (uint64_t)(ws + t) : key_of(x[ws + t]);
      idx[t] = (uint32_t)t;
    }
    std::stable_sort(idx.begin(), idx.end(), [&](uint32_t a, uint32_t b) { return key[a] < key[b]; });
    for (int64_t p = 0; p < m; p++) {
      const int64_t q = idx[p];
      if (q < ins) continue;
      const int32_t hh = h[ws + q];
      if (hh & 1) continue;
      for (int64_t pp = p + 1; pp < m; pp++) {
        if ((int64_t)idx[pp] >= q) continue;
        const int32_t h2 = h[ws + idx[pp]];
        if (!(h2 & 1) || (h2 >> 1) == (hh >> 1)) continue;
        E.push_back(Ev{((uint64_t)(ws + q) << 31) | (uint64_t)(0x7FFFFFFFu - (uint32_t)pp), hh >> 1, h2 >> 1});
      }
    }
    std::vector<int32_t> nh(m), np(m);
    std::vector<double> nx(m);
    for (int64_t t = 0; t < m; t++) {
      nh[t] = h[ws + idx[t]];
      np[t] = per[ws + idx[t]];
      nx[t] = x[ws + idx[t]];
    }
    std::copy(nh.begin(), nh.end(), h + ws);
    std::copy(np.begin(), np.end(), per + ws);
    std::copy(nx.begin(), nx.end(), x + ws);
  };
  seq(0, P);
  par(P, A);
  seq(A, B);
  par(B, Q);
  seq(Q, n2);
  if (bad) return -1;
  std::sort(E.begin(), E.end(), [](const Ev& a, const Ev& b) { return a.key < b.key; });
  *lo = l;
  put(E, ev, cap);
  return (int64_t)E.size();
}
}
