// collider.h -- kernels of the sort-based sweep-and-prune broadphase.
//
// InsertionSortCollider (pkg/common/InsertionSortCollider.cpp:191-400) keeps, for each
// axis, the endpoints of all bounds in the order of the previous run and insertion-sorts
// them; a new interaction is created when a lower bound moves left past an upper bound
// (:46-47), oriented (owner of the moving lower bound, other body).  Here the new order is
// the stable radix sort of the old order by the new coordinates (identical to the
// insertion sort's result), and the inversions are recovered from the old and new ranks:
// a lower bound at new rank p and old rank q crossed exactly the upper bounds that sit
// after it in the new order but had an old rank below q.  The suffix minimum of the old
// ranks over new positions bounds each scan.  The reference inserts these pairs in the
// order (axis, old rank of the moving bound, descending new rank of the crossed bound);
// the candidates carry that order as a 64-bit key and are appended in key order.
#pragma once
#include "pairs.h"

// Order-preserving map of a double to an unsigned key; -0.0 and +0.0 map together (the
// reference compares coordinates with <, under which they are equal).
__device__ __forceinline__ uint64_t coord_key(double x) {
  x = x + 0.0;
  uint64_t u = (uint64_t)__double_as_longlong(x);
  return (u & 0x8000000000000000ull) ? ~u : (u | 0x8000000000000000ull);
}

__device__ __forceinline__ double bound_coord(const BodyArrays& B, int32_t h, int axis) {
  const int id = h >> 1;
  const int64_t N = B.n;
  if (!B.hasBound[id]) return B.pos[axis * N + id];
  return (h & 1) ? B.bmax[axis * N + id] : B.bmin[axis * N + id];
}

// spatialOverlap (InsertionSortCollider.hpp:166-171): closed intervals on all three axes.
__device__ __forceinline__ bool spatial_overlap(const BodyArrays& B, int a, int b) {
  const int64_t N = B.n;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    if (!(B.bmin[k * N + a] <= B.bmax[k * N + b])) return false;
    if (!(B.bmax[k * N + a] >= B.bmin[k * N + b])) return false;
  }
  return true;
}

__global__ void k_bb_keys(BodyArrays B, const int32_t* order, int64_t n2, int axis, uint64_t* keys, uint32_t* vals) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n2) return;
  keys[q] = coord_key(bound_coord(B, order[q], axis));
  vals[q] = (uint32_t)q;
}

// Specimen of the handle now at position p (several specimens in one context): a stable
// sort by it after the coordinate sort keeps every specimen's bounds contiguous, so the
// sweep and the crossings never leave a specimen.
__global__ void k_spec_keys(const int32_t* specimen, const int32_t* oldOrder, const uint32_t* qs, int64_t n2,
                            uint32_t* keys) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n2) return;
  keys[p] = (uint32_t)specimen[oldOrder[qs[p]] >> 1];
}

__global__ void k_bb_apply(const int32_t* oldOrder, const uint32_t* qs, int64_t n2, int32_t* newOrder) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n2) return;
  newOrder[p] = oldOrder[qs[p]];
}

struct Cands {
  int32_t* a;      // owner of the moving lower bound (id1)
  int32_t* b;      // the other body (id2)
  uint64_t* key;   // insertion order key
  unsigned long long* count;
  int64_t cap;
};

// Inversions of one axis.  order/qs: new order and the old rank of each new position;
// smin: suffix minimum of qs.
__global__ void k_crossings(BodyArrays B, const int32_t* order, const uint32_t* qs, const uint32_t* smin, int64_t n2,
                            int axis, Cands C) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n2 - 1) return;
  const int32_t h = order[p];
  if (h & 1) return;  // only a moving lower bound creates an interaction
  const uint32_t q = qs[p];
  if (smin[p + 1] >= q) return;
  const int id = h >> 1;
  if (!B.hasBound[id]) return;
  for (int64_t pp = p + 1; pp < n2 && smin[pp] < q; pp++) {
    if (qs[pp] >= q) continue;
    const int32_t h2 = order[pp];
    if (!(h2 & 1)) continue;
    const int id2 = h2 >> 1;
    if (id2 == id || !B.hasBound[id2]) continue;
    if (!spatial_overlap(B, id, id2)) continue;
    const unsigned long long slot = atomicAdd(C.count, 1ull);
    if ((int64_t)slot < C.cap) {
      C.a[slot] = id;
      C.b[slot] = id2;
      C.key[slot] = ((uint64_t)axis << 62) | ((uint64_t)q << 31) | (uint64_t)(0x7FFFFFFFu - (uint32_t)pp);
    }
  }
}

// Initial sweep on the sort axis (InsertionSortCollider.cpp:346-383): for each lower bound,
// the lower bounds of other bodies before its own upper bound.  Pass 1 counts, pass 2 emits
// in (i, j) order at the scanned offsets.
template <bool kEmit>
__global__ void k_init_sweep(BodyArrays B, const int32_t* V, int64_t n2, uint32_t* counts, const uint32_t* offs,
                             Cands C) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n2) return;
  const int32_t h = V[i];
  uint32_t c = 0;
  if (!(h & 1) && B.hasBound[h >> 1]) {
    const int iid = h >> 1;
    uint32_t o = kEmit ? offs[i] : 0;
    for (int64_t j = i + 1; j < n2 && (V[j] >> 1) != iid; j++) {
      const int32_t hj = V[j];
      if (hj & 1) continue;
      const int jid = hj >> 1;
      if (!B.hasBound[jid]) continue;
      if (!spatial_overlap(B, iid, jid)) continue;
      if (kEmit) {
        C.a[o + c] = iid;
        C.b[o + c] = jid;
        C.key[o + c] = (uint64_t)(o + c);
      }
      c++;
    }
  }
  if (!kEmit) counts[i] = c;
}

// ---- open-addressing hash sets keyed by the unordered pair ----
__device__ __forceinline__ uint64_t upair(int a, int b) {
  if (a > b) { int t = a; a = b; b = t; }
  return ((uint64_t)(uint32_t)a << 32) | (uint32_t)b;
}
__device__ __forceinline__ uint64_t hmix(uint64_t k) {
  k ^= k >> 33;
  k *= 0xff51afd7ed558ccdull;
  k ^= k >> 33;
  k *= 0xc4ceb9fe1a85ec53ull;
  k ^= k >> 33;
  return k;
}
#define SDM_HEMPTY 0xFFFFFFFFFFFFFFFFull

__global__ void k_fill_u64(uint64_t* a, int64_t n, uint64_t v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

// ---- device-side counts: the step's sizes stay on the device, no host round trip ----
// Total of an exclusive scan: dst = out[n-1] + in[n-1].
__global__ void k_scan_total(const uint32_t* in, const uint32_t* out, int64_t n, unsigned long long* dst) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *dst = n > 0 ? (unsigned long long)out[n - 1] + in[n - 1] : 0ull;
}
// Sort padding past a device-side count: the largest key, so that a stable sort of all n
// entries leaves the counted ones first, in their order.
__global__ void k_pad_keys(uint64_t* keys, uint32_t* vals, const unsigned long long* count, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && i >= (int64_t)*count) {
    keys[i] = ~0ull;
    vals[i] = 0;
  }
}
__global__ void k_copy_u32(uint32_t* dst, const uint32_t* src, const unsigned long long* count, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && i < (int64_t)*count) dst[i] = src[i];
}
// Clear counters with a kernel (a memset node misbehaves under graph replay).
__global__ void k_zero_u64(unsigned long long* a, int n) {
  if (blockIdx.x == 0 && (int)threadIdx.x < n) a[threadIdx.x] = 0ull;
}
__global__ void k_zero_u32(uint32_t* a, int n) {
  if (blockIdx.x == 0 && (int)threadIdx.x < n) a[threadIdx.x] = 0u;
}
__global__ void k_set_u64x2(unsigned long long* a, unsigned long long x, unsigned long long y) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    a[0] = x;
    a[1] = y;
  }
}

__global__ void k_hash_insert_pairs(const int32_t* cid, int64_t np, uint64_t* tab, uint64_t mask) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  const int2 ab = reinterpret_cast<const int2*>(cid)[p];
  const uint64_t key = upair(ab.x, ab.y);
  uint64_t h = hmix(key) & mask;
  while (true) {
    const unsigned long long prev = atomicCAS((unsigned long long*)&tab[h], SDM_HEMPTY, (unsigned long long)key);
    if (prev == SDM_HEMPTY || prev == key) return;
    h = (h + 1) & mask;
  }
}

__device__ __forceinline__ bool hash_contains(const uint64_t* tab, uint64_t mask, uint64_t key) {
  uint64_t h = hmix(key) & mask;
  while (true) {
    const uint64_t v = tab[h];
    if (v == key) return true;
    if (v == SDM_HEMPTY) return false;
    h = (h + 1) & mask;
  }
}

// Drop candidates already present; keep the smallest order key per new pair.
__global__ void k_cand_dedup(Cands C, int64_t nc, const uint64_t* found, uint64_t fmask, uint64_t* dkeys,
                             uint64_t* dvals, uint64_t dmask, uint32_t* alive) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  const uint64_t key = upair(C.a[c], C.b[c]);
  if (hash_contains(found, fmask, key)) {
    alive[c] = 0;
    return;
  }
  alive[c] = 1;
  uint64_t h = hmix(key) & dmask;
  while (true) {
    const unsigned long long prev = atomicCAS((unsigned long long*)&dkeys[h], SDM_HEMPTY, (unsigned long long)key);
    if (prev == SDM_HEMPTY || prev == key) break;
    h = (h + 1) & dmask;
  }
  atomicMin((unsigned long long*)&dvals[h], (unsigned long long)C.key[c]);
}

__global__ void k_cand_keep(Cands C, int64_t nc, const uint64_t* dkeys, const uint64_t* dvals, uint64_t dmask,
                            uint32_t* alive) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc || !alive[c]) return;
  const uint64_t key = upair(C.a[c], C.b[c]);
  uint64_t h = hmix(key) & dmask;
  while (dkeys[h] != key) h = (h + 1) & dmask;
  alive[c] = (dvals[h] == C.key[c]) ? 1u : 0u;
}

// Stable compaction of the surviving candidates: their order keys and candidate indices.
__global__ void k_cand_compact(const Cands C, int64_t nc, const uint32_t* alive, const uint32_t* pos, uint64_t* okey,
                               uint32_t* oidx) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc || !alive[c]) return;
  okey[pos[c]] = C.key[c];
  oidx[pos[c]] = (uint32_t)c;
}

// ---- InteractionContainer::conditionalyEraseNonReal (core/InteractionContainer.hpp:104-112)
// The reference walks the container and replaces each erased entry by the current last one.
// The result: the kept entries below K = (number kept) stay; the holes below K receive the
// kept entries at positions >= K, the last one filling the first hole.
__global__ void k_erase_flags(BodyArrays B, PairArrays P, int64_t np, uint32_t* flag) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  const bool real = P.r[p].hasGeom[0] && P.r[p].hasPhys[0];
  flag[p] = (!real && !spatial_overlap(B, P.r[p].id1[0], P.r[p].id2[0])) ? 1u : 0u;
}

__global__ void k_hole_pos(const uint32_t* flag, const uint32_t* E, int64_t K, int32_t* holePos) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= K || !flag[i]) return;
  holePos[E[i]] = (int32_t)i;
}

// For kept j >= K: rank from the end t = #kept in (j, np) = (np-1-j) - #erased in (j, np).
__global__ void k_tail_moves(const uint32_t* flag, const uint32_t* E, int64_t K, int64_t np, uint32_t Etot,
                             const int32_t* holePos, int32_t* msrc, int32_t* mdst) {
  const int64_t j = K + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= np || flag[j]) return;
  const uint32_t erasedAfter = Etot - (E[j] + flag[j]);
  const int64_t t = (np - 1 - j) - (int64_t)erasedAfter;
  msrc[t] = (int32_t)j;
  mdst[t] = holePos[t];
}

// Move whole interaction records: record msrc[k] to slot mdst[k] (the swap-with-last
// compaction of InteractionContainer::erase).  One 16-byte word per lane, 40 lanes per record.
__global__ void k_move_pairs(PairRec* R, int32_t* cid, const int32_t* msrc, const int32_t* mdst, int64_t nm) {
  constexpr int W = (int)(sizeof(PairRec) / 16);
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t k = t / W;
  const int w = (int)(t % W);
  if (k >= nm) return;
  const int64_t s = msrc[k], d = mdst[k];
  reinterpret_cast<uint4*>(R + d)[w] = reinterpret_cast<const uint4*>(R + s)[w];
  if (w == 0) reinterpret_cast<int2*>(cid)[d] = reinterpret_cast<const int2*>(cid)[s];
}

// (id1, id2) of every interaction into the packed endpoint array.
__global__ void k_pair_ids(const PairRec* R, int64_t np, int32_t* cid) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  reinterpret_cast<int2*>(cid)[p] = make_int2(R[p].id1[0], R[p].id2[0]);
}

// List entry e (sorted by body, interaction order within a body) -> its interaction: the
// entry holds p << 2 | side << 1 | active, and cpos remembers where the two entries of each
// interaction are, so the law can mark a contributing contact in both lists.  The lists are
// rebuilt only before the first interaction pass of a step, which withdraws every active
// interaction's entries before any are read, so they start with the bit clear.
__global__ void k_csr_finalize(const uint32_t* sorted, int64_t m, int32_t* cpos, int32_t* csr) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= m) return;
  const uint32_t v = sorted[e];
  cpos[v] = (int32_t)e;
  csr[e] = csr_entry(v >> 1, (int)(v & 1u), false);
}

// List offsets from the sorted body keys: off[b] = first entry whose body is >= b (a
// binary search per body; off[N] = m).
__global__ void k_csr_offsets(const uint32_t* key, int64_t m, int64_t N, int32_t* off) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b > N) return;
  int64_t lo = 0, hi = m;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if ((int64_t)key[mid] < b)
      lo = mid + 1;
    else
      hi = mid;
  }
  off[b] = (int32_t)lo;
}

// One attribute of every interaction between the records and a row-major buffer
// (`words` 32-bit words per interaction, at byte offset `off` in the record).
__global__ void k_gather_field(const PairRec* R, int64_t n, int off, int words, uint32_t* out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * words) return;
  const int64_t i = t / words;
  const int w = (int)(t % words);
  out[t] = reinterpret_cast<const uint32_t*>(reinterpret_cast<const char*>(R + i) + off)[w];
}
__global__ void k_scatter_field(PairRec* R, int64_t n, int off, int words, const uint32_t* in) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * words) return;
  const int64_t i = t / words;
  const int w = (int)(t % words);
  reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(R + i) + off)[w] = in[t];
}

// Fresh Interaction(id1,id2) (core/Interaction.cpp:15): no geom, no phys, geomExists.
__global__ void k_new_pairs(PairArrays P, int64_t start, int64_t count, const Cands C, const uint32_t* idx,
                            int64_t iter) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= count) return;
  const int64_t p = start + k;
  const uint32_t c = idx ? idx[k] : (uint32_t)k;
  PairRec& R = P.r[p];
#define SDM_ZERO(T, m, kk, name)              \
  for (int cc = 0; cc < kk; cc++) R.m[cc] = (T)0;
  SDM_PAIR_FIELDS_NOIDS(SDM_ZERO)
#undef SDM_ZERO
  R.id1[0] = C.a[c];
  R.id2[0] = C.b[c];
  reinterpret_cast<int2*>(P.cid)[p] = make_int2(C.a[c], C.b[c]);
  P.r[p].geomExists[0] = 1;
  P.r[p].isShearNew[0] = 1;
  P.r[p].iterMadeReal[0] = -1;
  P.r[p].tanPhi[0] = __builtin_nan("");
}

// Interaction endpoints for the per-body lists: entry 2p -> (id1, 2p), 2p+1 -> (id2, 2p+1).
__global__ void k_csr_entries(const int32_t* cid, int64_t np, uint32_t* key, uint32_t* val) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  const int2 ab = reinterpret_cast<const int2*>(cid)[p];
  const int a = ab.x, b = ab.y;
  key[2 * p] = (uint32_t)a;
  val[2 * p] = (uint32_t)(2 * p);
  key[2 * p + 1] = (uint32_t)b;
  val[2 * p + 1] = (uint32_t)(2 * p + 1);
}
