// sdm.hip -- libsdm_hip.so: the superellipsoid DEM step on one MI355X (gfx950).
//
// Host side of the library: context management, the engine loop of one
// Scene::moveToNextTimeStep (core/Scene.cpp:70-138) with the canonical engine list
// [ForceResetter, InsertionSortCollider, InteractionLoop, NewtonIntegrator], and the C ABI
// declared in include/sdm.h.  All state lives in HBM; the host keeps only the scalar
// collider/integrator flags the reference keeps in its engine objects.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/sdm.h"
#include "bodies.h"
#include "frozen.h"
#include "observe.h"
#include "periodic.h"
#include "radix.h"
#include "ring_dev.h"
#include "servo.h"

#define SDM_VERSION "sdm-hip 1.0 (gfx950)"

namespace {

std::string g_err;

// kernels timed one by one when timing is on: the integrator, the ordered force sums, the
// bound update of a collider run, the interaction-loop prep
// Kernels timed under sdm_set_timing (events around each launch or group of launches).
enum {
  KT_NEWTON = 0,     // k_newton
  KT_FORCE_SUM = 1,  // k_force_sum
  KT_BOUNDS = 2,     // k_bounds (collider runs)
  KT_PREP = 3,       // k_pairs_prep
  KT_QUEUE = 4,      // the solver queue: flag scan + k_queue_fill
  KT_MISC = 5,       // k_pairs_misc (wall and sphere pairs)
  KT_FIRST = 6,      // k_nm_first
  KT_LIST = 7,       // Nelder-Mead list: flag scan, fill, cost keys, cost sort
  KT_SOLVE = 8,      // main solver launch (k_solve_nm / k_solve_lm)
  KT_RESUME = 9,     // mixed NM resume launch
  KT_REFINE = 10,    // k_nm_refine
  KT_FINISH = 11,    // k_pairs_finish
  KT_RING_WALK = 12, // periodic ring walk: coordinates and the sequential phases
  KT_RING_WIN = 13,  // periodic ring walk: the parallel phases (checks, sorts, crossings)
  KT_N = 14
};

struct HipError {
  std::string msg;
};

#define HIPCHK(x)                                                                              \
  do {                                                                                         \
    hipError_t e_ = (x);                                                                       \
    if (e_ != hipSuccess) throw HipError{std::string(#x) + ": " + hipGetErrorString(e_)};      \
  } while (0)

template <typename T>
T* dalloc(int64_t n) {
  void* p = nullptr;
  if (n <= 0) n = 1;
  HIPCHK(hipMalloc(&p, (size_t)n * sizeof(T)));
  return (T*)p;
}
template <typename T>
void dfree(T*& p) {
  if (p) (void)hipFree((void*)p);
  p = nullptr;
}

inline unsigned nblk(int64_t n, int bs = 256) { return (unsigned)((n + bs - 1) / bs); }

#define KT_START(c, k)                                                      \
  do {                                                                      \
    if ((c)->timing) HIPCHK(hipEventRecord((c)->kev[2 * (k)], (c)->st));     \
  } while (0)
#define KT_STOP(c, k)                                                       \
  do {                                                                      \
    if ((c)->timing) {                                                      \
      HIPCHK(hipEventRecord((c)->kev[2 * (k) + 1], (c)->st));                \
      (c)->kran[k] = true;                                                  \
    }                                                                       \
  } while (0)

}  // namespace

struct sdm_ctx {
  sdm_params prm;
  int dev = 0;
  hipStream_t st = nullptr;
  BodyArrays B;
  PairArrays P;
  int64_t np = 0;
  MaterialDev* dmats = nullptr;
  std::vector<sdm_material> mats;
  // collider persistent state (InsertionSortCollider members)
  int32_t* bb[3] = {nullptr, nullptr, nullptr};
  int32_t* bbTmp = nullptr;
  bool bbValid = false;
  int32_t strideActive = 0;
  double fastestBodyMaxDist = -1;
  int32_t numReinit = 0;
  // NewtonIntegrator::maxVelocitySq, Scene::iter/time
  double maxVelocitySq = std::numeric_limits<double>::quiet_NaN();
  int64_t iter = 0;
  double time = 0;
  // per-body interaction lists
  int32_t* csrOff = nullptr;
  int32_t* csrPair = nullptr;
  int64_t csrCap = 0;
  bool csrValid = false;
  bool cidValid = false;  // P.cid holds (id1, id2) of every interaction
  // scratch
  uint64_t *k0 = nullptr, *k1 = nullptr;
  uint32_t *v0 = nullptr, *v1 = nullptr, *u0 = nullptr, *u1 = nullptr;
  int64_t kcap = 0;
  uint32_t* rsScratch = nullptr;
  int64_t rsCap = 0;
  uint32_t* qs[3] = {nullptr, nullptr, nullptr};
  uint32_t* smin = nullptr;
  uint64_t* htab = nullptr;
  uint64_t* dkeys = nullptr;
  uint64_t* dvals = nullptr;
  int64_t hcap = 0;
  Cands C;
  int32_t *msrc = nullptr, *mdst = nullptr, *holePos = nullptr;
  int64_t mcap = 0;
  // device counters: [0] contact evals, [1] dist evals, [2] candidates, [3] solver queue,
  // [4] solver evaluations, [5] other evaluations, [6] misc list, [7] touching list
  unsigned long long* dcnt = nullptr;
  int* dMaxDisp = nullptr;
  // wall stress servo (servo.h): on/off, its device state, the check's block partials
  bool servoOn = false;
  bool servoDone = false;  // the last step completed the consolidation (stopAtIter = iter + 1)
  ServoDev* dServo = nullptr;
  double* servoPart = nullptr;
  int64_t servoPartCap = 0;
  // narrowphase work queue (indexed by queue slot)
  SolveIO S;
  int32_t* retryList = nullptr;
  int32_t* miscList = nullptr;
  uint32_t* qflag = nullptr;  // per interaction / per slot flags and their exclusive scan
  uint32_t* qpos = nullptr;
  int numCU = 0;
  int residentNM[3] = {0, 0, 0};  // resident blocks of k_solve_nm<float>, <double>, <float, true> (0: not queried)
  int residentLM = 0;  // resident blocks of k_solve_lm
  bool matChecked = false;
  bool solverRan = false;
  bool geoDirty = true;
  int32_t nSpec = 1;  // specimens (body.specimen), contiguous ranges of body ids
  bool pairsWritten = true;  // host wrote interaction data since the last step
  bool skipValid = false;    // P.gap holds this container's settled gaps (no collider run or host write since)
  // Captured steady-state step (hipGraph): rebuilt when any buffer it uses is reallocated or a
  // parameter it bakes in changes (graphGen moves on).
  bool useGraph = true;
  sdm_tuning tune = {0, 1, 2, 1};
  hipGraphExec_t graphExec = nullptr;
  uint64_t graphGen = 1, graphBuiltGen = 0;
  int64_t graphs = 0;
  uint32_t* stage = nullptr;  // attribute staging buffer (32-bit words)
  int64_t stageCap = 0;
  // periodic cell broadphase (periodic.h)
  int32_t* cellRng = nullptr;                     // per body: start cell and count per axis
  uint32_t *rankMin = nullptr, *rankMax = nullptr;
  uint32_t *peCnt = nullptr, *peOff = nullptr;    // cells per body and their exclusive scan
  uint32_t* runEnd = nullptr;                     // [runCap] end, then [runCap] start of each cell's entries
  int64_t runCap = 0;
  float4* pbox = nullptr;                          // [2N] single-precision boxes of the pair screen
  // the rings of a periodic cell (ring.h): bb[a] holds the handles, rper[a] the periods,
  // loIdx[a] the ring head (host copy; the run keeps it in rdev while it walks)
  int32_t* rper[3] = {nullptr, nullptr, nullptr};
  int64_t loIdx[3] = {0, 0, 0};
  double* rx[3] = {nullptr, nullptr, nullptr};  // the run's coordinates
  int32_t *rbh[3] = {nullptr, nullptr, nullptr}, *rbp[3] = {nullptr, nullptr, nullptr};  // the rings before the run
  int32_t *rth = nullptr, *rtp = nullptr;  // window permutation temporaries
  double* rtx = nullptr;
  int64_t* rdev = nullptr;  // [0,3) loIdx, [3,6) error bits (u32 in 8-byte slots), window state
  hipEvent_t rev[6] = {};    // the ring walk's phase boundaries (timing)
  // per-phase timing
  bool timing = false;
  hipEvent_t ev[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  double tms[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  // per-kernel timing of the bandwidth-bound kernels (KT_*): start/stop events, whether the
  // kernel ran in the current step, accumulated milliseconds and launches
  hipEvent_t kev[2 * KT_N] = {};
  bool kran[KT_N] = {};
  double kms[KT_N + 4] = {};  // + the mixed form's touching-list, double-evaluation and hand-over counts,
                              // the ring walk's repeated runs
  double kcount[KT_N] = {};
  sdm_counters cnt;
  std::string err;
};

namespace {

int fail(sdm_ctx* c, const std::string& m) {
  if (c) c->err = m;
  g_err = m;
  return -1;
}

void alloc_bodies(sdm_ctx* c, int32_t n) {
  c->B.n = n;
#define SDM_ALLOC(T, m, k, name) c->B.m = dalloc<T>((int64_t)n * k);
  SDM_BODY_FIELDS(SDM_ALLOC)
#undef SDM_ALLOC
  for (int a = 0; a < 3; a++) c->bb[a] = dalloc<int32_t>(2 * (int64_t)n);
  for (int a = 0; a < 3; a++) {
    c->rper[a] = dalloc<int32_t>(std::max<int64_t>(1, 2 * (int64_t)n));
    HIPCHK(hipMemset(c->rper[a], 0, std::max<int64_t>(1, 2 * (int64_t)n) * sizeof(int32_t)));
  }
  c->bbTmp = dalloc<int32_t>(2 * (int64_t)n);
  for (int a = 0; a < 3; a++) c->qs[a] = dalloc<uint32_t>(2 * (int64_t)n);
  c->smin = dalloc<uint32_t>(2 * (int64_t)n + 1);
  c->csrOff = dalloc<int32_t>((int64_t)n + 1);
  c->B.geo = dalloc<double>((int64_t)n * GEO_STRIDE);
  c->B.lid = dalloc<int32_t>(std::max<int64_t>(1, n));
  c->B.disp = dalloc<float>(std::max<int64_t>(1, n));
}

void free_pairs(PairArrays& P) {
  dfree(P.r);
  dfree(P.cid);
  dfree(P.cpos);
  dfree(P.gap);
}

// body.specimen: ids 0,0,..,1,1,.. (non-decreasing); each body's index within its specimen.
void set_specimens(sdm_ctx* c, const int32_t* spec) {
  const int64_t N = c->B.n;
  std::vector<int32_t> lid(std::max<int64_t>(1, N));
  int32_t cur = 0;
  int64_t start = 0;
  for (int64_t i = 0; i < N; i++) {
    if (spec[i] < cur || spec[i] < 0) throw HipError{"body.specimen must be non-decreasing and >= 0"};
    if (spec[i] != cur) {
      cur = spec[i];
      start = i;
    }
    lid[i] = (int32_t)(i - start);
  }
  const int32_t nSpec = N > 0 ? spec[N - 1] + 1 : 1;
  if (N > 0) HIPCHK(hipMemcpy(c->B.lid, lid.data(), N * sizeof(int32_t), hipMemcpyHostToDevice));
  if (nSpec != c->nSpec) c->bbValid = false;  // the bounds are re-sorted specimen by specimen
  c->nSpec = nSpec;
}

uint32_t* staging(sdm_ctx* c, int64_t words) {
  if (words > c->stageCap) {
    HIPCHK(hipStreamSynchronize(c->st));
    dfree(c->stage);
    c->stageCap = words + words / 4 + 1024;
    c->stage = dalloc<uint32_t>(c->stageCap);
  }
  return c->stage;
}

// Grow the interaction arrays to hold at least `need` entries, keeping [0, np).
void ensure_pair_cap(sdm_ctx* c, int64_t need) {
  if (need <= c->P.cap && c->P.r) return;
  int64_t cap = c->P.cap > 0 ? c->P.cap : 1024;
  while (cap < need) cap = cap * 3 / 2 + 1024;
  PairArrays Q;
  Q.cap = cap;
  Q.r = dalloc<PairRec>(cap);
  Q.cid = dalloc<int32_t>(2 * cap);
  Q.cpos = dalloc<int32_t>(2 * cap);
  Q.gap = dalloc<float>(cap);
  c->skipValid = false;
  c->graphGen++;
  if (c->P.r && c->np > 0) {
    HIPCHK(hipMemcpyAsync(Q.r, c->P.r, c->np * sizeof(PairRec), hipMemcpyDeviceToDevice, c->st));
    HIPCHK(hipMemcpyAsync(Q.cid, c->P.cid, c->np * 2 * sizeof(int32_t), hipMemcpyDeviceToDevice, c->st));
  }
  HIPCHK(hipStreamSynchronize(c->st));
  free_pairs(c->P);
  c->P = Q;
  c->csrValid = false;
}

void ensure_scratch(sdm_ctx* c, int64_t n) {
  if (n <= c->kcap) return;
  HIPCHK(hipStreamSynchronize(c->st));
  int64_t cap = n + n / 2 + 4096;
  c->graphGen++;
  dfree(c->k0); dfree(c->k1); dfree(c->v0); dfree(c->v1); dfree(c->u0); dfree(c->u1);
  c->k0 = dalloc<uint64_t>(cap);
  c->k1 = dalloc<uint64_t>(cap);
  c->v0 = dalloc<uint32_t>(cap);
  c->v1 = dalloc<uint32_t>(cap);
  c->u0 = dalloc<uint32_t>(cap + 1);
  c->u1 = dalloc<uint32_t>(cap + 1);
  c->kcap = cap;
  int64_t rs = std::max(sdmrs::sort_scratch(cap), sdmrs::scan_scratch(cap + 1)) + 64;
  if (rs > c->rsCap) {
    dfree(c->rsScratch);
    c->rsScratch = dalloc<uint32_t>(rs);
    c->rsCap = rs;
  }
}

void ensure_cands(sdm_ctx* c, int64_t n) {
  if (n <= c->C.cap) return;
  HIPCHK(hipStreamSynchronize(c->st));
  int64_t cap = n + n / 2 + 4096;
  dfree(c->C.a); dfree(c->C.b); dfree(c->C.key);
  c->C.a = dalloc<int32_t>(cap);
  c->C.b = dalloc<int32_t>(cap);
  c->C.key = dalloc<uint64_t>(cap);
  c->C.cap = cap;
}

void ensure_hash(sdm_ctx* c, int64_t n) {
  int64_t need = 1024;
  while (need < 2 * n + 16) need <<= 1;
  if (need <= c->hcap) return;
  HIPCHK(hipStreamSynchronize(c->st));
  dfree(c->htab); dfree(c->dkeys); dfree(c->dvals);
  c->htab = dalloc<uint64_t>(need);
  c->dkeys = dalloc<uint64_t>(need);
  c->dvals = dalloc<uint64_t>(need);
  c->hcap = need;
}

void ensure_moves(sdm_ctx* c, int64_t n) {
  if (n <= c->mcap) return;
  HIPCHK(hipStreamSynchronize(c->st));
  int64_t cap = n + n / 2 + 1024;
  dfree(c->msrc); dfree(c->mdst); dfree(c->holePos);
  c->msrc = dalloc<int32_t>(cap);
  c->mdst = dalloc<int32_t>(cap);
  c->holePos = dalloc<int32_t>(cap);
  c->mcap = cap;
}

void free_queue(SolveIO& S) {
  dfree(S.pair); dfree(S.tag); dfree(S.id1); dfree(S.id2); dfree(S.a0); dfree(S.a1); dfree(S.del);
  dfree(S.f); dfree(S.la0); dfree(S.la1); dfree(S.p1); dfree(S.p2); dfree(S.ctr); dfree(S.ct);
  dfree(S.nev); dfree(S.stop); dfree(S.touch); dfree(S.sh); dfree(S.anc); dfree(S.fb);
}

void ensure_queue(sdm_ctx* c, int64_t n) {
  if (n <= c->S.cap && c->S.pair) return;
  HIPCHK(hipStreamSynchronize(c->st));
  c->graphGen++;
  free_queue(c->S);
  dfree(c->retryList);
  dfree(c->miscList);
  dfree(c->qflag);
  dfree(c->qpos);
  const int64_t cap = n + n / 2 + 1024;
  SolveIO& S = c->S;
  S.cap = cap;
  S.pair = dalloc<int32_t>(cap); S.tag = dalloc<int32_t>(cap); S.id1 = dalloc<int32_t>(cap);
  S.id2 = dalloc<int32_t>(cap); S.a0 = dalloc<double>(cap); S.a1 = dalloc<double>(cap);
  S.del = dalloc<double>(cap); S.f = dalloc<double>(cap); S.la0 = dalloc<double>(cap);
  S.la1 = dalloc<double>(cap); S.p1 = dalloc<double>(3 * cap); S.p2 = dalloc<double>(3 * cap);
  S.ctr = dalloc<double>(3 * cap); S.ct = dalloc<double>(3 * cap); S.nev = dalloc<int32_t>(cap);
  S.stop = dalloc<int32_t>(cap); S.touch = dalloc<int32_t>(cap);
  S.sh = c->prm.periodic ? dalloc<double>(3 * cap) : nullptr;
  const bool mixed = c->prm.precision == SDM_MIXED;
  S.anc = mixed ? dalloc<float>(kAnc * cap) : nullptr;
  S.fb = mixed ? dalloc<int32_t>(cap) : nullptr;
  c->retryList = dalloc<int32_t>(cap);
  c->miscList = dalloc<int32_t>(cap);
  c->qflag = dalloc<uint32_t>(cap + 1);
  c->qpos = dalloc<uint32_t>(cap + 1);
}

uint32_t read_u32(sdm_ctx* c, const uint32_t* p) {
  uint32_t v = 0;
  HIPCHK(hipMemcpyAsync(&v, p, 4, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return v;
}

// Exclusive scan of u0[0..n) into u1[0..n], with u1[n] = total; returns the total.
uint32_t exclusive_total(sdm_ctx* c, const uint32_t* in, uint32_t* out, int64_t n) {
  if (n == 0) return 0;
  sdmrs::scan<sdmrs::OpAdd, false, false>(in, out, n, c->rsScratch, c->st);
  uint32_t a = 0, b = 0;
  HIPCHK(hipMemcpyAsync(&a, out + n - 1, 4, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipMemcpyAsync(&b, in + n - 1, 4, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return a + b;
}

// P.cid from the records, after the host wrote interactions.
void refresh_cid(sdm_ctx* c) {
  if (c->cidValid) return;
  if (c->np > 0)
    hipLaunchKernelGGL(k_pair_ids, dim3(nblk(c->np)), dim3(256), 0, c->st, c->P.r, c->np, c->P.cid);
  c->cidValid = true;
}

// Per-body interaction lists, each in interaction order (stable sort by body id).
void build_csr(sdm_ctx* c) {
  refresh_cid(c);
  const int64_t N = c->B.n;
  const int64_t m = 2 * c->np;
  if (m > c->csrCap) {
    HIPCHK(hipStreamSynchronize(c->st));
    dfree(c->csrPair);
    c->graphGen++;
    c->csrCap = m + m / 2 + 1024;
    c->csrPair = dalloc<int32_t>(c->csrCap);
  }
  ensure_scratch(c, std::max<int64_t>(m, N + 1));
  uint32_t* key = c->u1;
  if (c->np > 0) {
    hipLaunchKernelGGL(k_csr_entries, dim3(nblk(c->np)), dim3(256), 0, c->st, c->P.cid, c->np, key, c->v0);
    int bits = 1;
    while ((1ll << bits) < N) bits++;
    // keys are 32-bit: reuse the 64-bit sort buffers through a 32-bit view
    sdmrs::sort_pairs<uint32_t, uint32_t>(key, c->v0, (uint32_t*)c->k1, c->v1, m, bits, c->rsScratch, c->st);
    hipLaunchKernelGGL(k_csr_finalize, dim3(nblk(m)), dim3(256), 0, c->st, c->v0, m, c->P.cpos, c->csrPair);
  }
  hipLaunchKernelGGL(k_csr_offsets, dim3(nblk(N + 1)), dim3(256), 0, c->st, key, m, N, c->csrOff);
  c->csrValid = true;
}

unsigned long long read_counter(sdm_ctx* c, int k) {
  unsigned long long v = 0;
  HIPCHK(hipMemcpyAsync(&v, c->dcnt + k, sizeof(v), hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return v;
}

void zero_counter(sdm_ctx* c, int k) { hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, c->st, c->dcnt + k, 1); }

// Device counter slots: [0] contact evals, [1] distance evals, [2] collider candidates,
// [3] solver batch counter, [4] solver evaluations, [5] other evaluations, [6] wall/sphere
// list, [7]-[9] periodic collider, [10] solver queue, [11] touching list, [12] LM restarts,
// [13] the container size and [14] Scene::iter of the step, [15] the mixed form's solver
// evaluations done in double, [16] its pairs handed to the resume launch.
constexpr int kCounters = 24;
enum { DC_NEXT = 3, DC_MISC = 6, DC_NQ = 10, DC_NT = 11, DC_NP = 13, DC_NDBL = 15, DC_NHO = 16 };

// Exclusive scan of in[0..n) into out[0..n) with the total left on the device in dcnt[slot].
void scan_to_counter(sdm_ctx* c, const uint32_t* in, uint32_t* out, int64_t n, int slot) {
  if (n > 0) sdmrs::scan<sdmrs::OpAdd, false, false>(in, out, n, c->rsScratch, c->st);
  hipLaunchKernelGGL(k_scan_total, dim3(1), dim3(64), 0, c->st, in, out, n, c->dcnt + slot);
}

// Empty the found-set hash and insert the interactions of the container.
void fill_found(sdm_ctx* c) {
  const uint64_t mask = (uint64_t)c->hcap - 1;
  hipLaunchKernelGGL(k_fill_u64, dim3(nblk(c->hcap)), dim3(256), 0, c->st, c->htab, c->hcap, SDM_HEMPTY);
  hipLaunchKernelGGL(k_fill_u64, dim3(nblk(c->hcap)), dim3(256), 0, c->st, c->dkeys, c->hcap, SDM_HEMPTY);
  hipLaunchKernelGGL(k_fill_u64, dim3(nblk(c->hcap)), dim3(256), 0, c->st, c->dvals, c->hcap, SDM_HEMPTY);
  if (c->np > 0)
    hipLaunchKernelGGL(k_hash_insert_pairs, dim3(nblk(c->np)), dim3(256), 0, c->st, c->P.cid, c->np, c->htab, mask);
}

// The new candidates C[0, nc): drop the ones present, keep one per pair (smallest key),
// append them in key order as fresh interactions.  Returns the number appended.
int64_t append_candidates(sdm_ctx* c, int64_t nc) {
  ensure_scratch(c, nc + 1);
  const uint64_t mask = (uint64_t)c->hcap - 1;
  uint32_t* alive = c->u0;
  uint32_t* pos = c->u1;
  hipLaunchKernelGGL(k_cand_dedup, dim3(nblk(nc)), dim3(256), 0, c->st, c->C, nc, c->htab, mask, c->dkeys, c->dvals,
                     mask, alive);
  hipLaunchKernelGGL(k_cand_keep, dim3(nblk(nc)), dim3(256), 0, c->st, c->C, nc, c->dkeys, c->dvals, mask, alive);
  const int64_t ns = exclusive_total(c, alive, pos, nc);
  if (ns > 0) {
    hipLaunchKernelGGL(k_cand_compact, dim3(nblk(nc)), dim3(256), 0, c->st, c->C, nc, alive, pos, c->k0, c->v0);
    sdmrs::sort_pairs<uint64_t, uint32_t>(c->k0, c->v0, c->k1, c->v1, ns, 64, c->rsScratch, c->st);
    ensure_pair_cap(c, c->np + ns);
    hipLaunchKernelGGL(k_new_pairs, dim3(nblk(ns)), dim3(256), 0, c->st, c->P, c->np, ns, c->C, c->v0, c->iter);
    c->np += ns;
    c->cnt.pairs_inserted += ns;
  }
  return ns;
}

// Erase step of the collider (conditionalyEraseNonReal), given the per-interaction flags
// in u0: the swap-with-last compaction of InteractionContainer::erase.
void erase_flagged(sdm_ctx* c) {
  uint32_t* flag = c->u0;
  uint32_t* E = c->u1;
  const uint32_t Etot = exclusive_total(c, flag, E, c->np);
  if (Etot == 0) return;
  const int64_t K = c->np - Etot;
  hipLaunchKernelGGL(k_hole_pos, dim3(nblk(K)), dim3(256), 0, c->st, flag, E, K, c->holePos);
  const int64_t tail = c->np - K;
  hipLaunchKernelGGL(k_tail_moves, dim3(nblk(tail)), dim3(256), 0, c->st, flag, E, K, c->np, Etot, c->holePos,
                     c->msrc, c->mdst);
  // moves = holes below K = erased entries below K = E[K] (= kept entries at >= K)
  const int64_t nmove = read_u32(c, E + K);
  if (nmove > 0)
    hipLaunchKernelGGL(k_move_pairs, dim3(nblk(nmove * (int64_t)(sizeof(PairRec) / 16))), dim3(256), 0, c->st,
                       c->P.r, c->P.cid, c->msrc, c->mdst, nmove);
  c->np = K;
  c->cnt.pairs_erased += Etot;
}

// The ring walk of a later collider run (insertionSortPeri on the three axes, ring.h): the
// candidates (overlapping pairs of the inversions, in the reference's order) land in C; the
// rings and loIdx are updated.  Per axis the walk runs as five phases in pass order:
// sequential [0, P), parallel up to the zone around loIdx, sequential through that zone,
// parallel up to Q, sequential [Q, n2).  Before a parallel phase runs, the device checks its
// conditions; when they fail, the sequential phase before it goes on over the positions the
// check names, and the check is repeated.  Returns the number of candidates.
constexpr int64_t kRingStart = 128;   // sequential phase at the vector's start and at its end
constexpr int64_t kRingLoHalf = 64;   // half width of the sequential zone around loIdx
constexpr int64_t kRingWindow = 1024; // prefix window of a parallel phase

int64_t ring_run(sdm_ctx* c, const PeriCell& pc) {
  const int64_t N = c->B.n;
  const int64_t n2 = 2 * N;
  if (!c->rx[0]) {
    for (int a = 0; a < 3; a++) {
      c->rx[a] = dalloc<double>(n2);
      c->rbh[a] = dalloc<int32_t>(n2);
      c->rbp[a] = dalloc<int32_t>(n2);
    }
    c->rth = dalloc<int32_t>(n2);
    c->rtp = dalloc<int32_t>(n2);
    c->rtx = dalloc<double>(n2);
    c->rdev = dalloc<int64_t>(64);
  }
  // the single-precision screen boxes of this run's bounds, for the crossing scans
  if (!c->pbox) c->pbox = dalloc<float4>(2 * N);
  hipLaunchKernelGGL(k_screen_boxes, dim3(nblk(N)), dim3(256), 0, c->st, c->B, pc, c->pbox);
  RingSet R;
  for (int a = 0; a < 3; a++) {
    R.h[a] = c->bb[a];
    R.per[a] = c->rper[a];
    R.x[a] = c->rx[a];
    R.L[a] = pc.L[a];
  }
  R.n2 = n2;
  R.lo = c->rdev;
  R.err = (uint32_t*)(c->rdev + 3);  // three u32 in rdev[3..4]
  RingWin w;
  w.r0 = c->rdev + 8;
  w.r0base = c->rdev + 12;
  w.minKey = (uint64_t*)(c->rdev + 16);
  w.bottom = (uint64_t*)(c->rdev + 20);
  w.ext = c->rdev + 24;
  ensure_cands(c, std::max<int64_t>(c->C.cap, n2));
  ensure_scratch(c, n2 + 1);
  unsigned long long* perr = c->dcnt + 7;
  for (int tries = 0;; tries++) {
    // the rings before the run (kept for a repeat with more room for candidates)
    for (int a = 0; a < 3; a++) {
      const bool restore = tries > 0;
      HIPCHK(hipMemcpyAsync(restore ? c->bb[a] : c->rbh[a], restore ? c->rbh[a] : c->bb[a], n2 * sizeof(int32_t),
                            hipMemcpyDeviceToDevice, c->st));
      HIPCHK(hipMemcpyAsync(restore ? c->rper[a] : c->rbp[a], restore ? c->rbp[a] : c->rper[a],
                            n2 * sizeof(int32_t), hipMemcpyDeviceToDevice, c->st));
    }
    int64_t hdr[5] = {c->loIdx[0], c->loIdx[1], c->loIdx[2], 0, 0};
    HIPCHK(hipMemcpyAsync(c->rdev, hdr, sizeof(hdr), hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemsetAsync(c->dcnt + 2, 0, sizeof(unsigned long long), c->st));
    c->C.count = c->dcnt + 2;
    // phase boundaries per axis: 0 <= P <= A <= B <= Q <= n2
    int64_t bnd[3][4];
    for (int a = 0; a < 3; a++) {
      const int64_t lo = c->loIdx[a];
      const int64_t P = std::min(kRingStart, n2);
      const int64_t Q = std::max(n2 - kRingStart, P);
      const int64_t A = std::min(std::max(lo - kRingLoHalf, P), Q);
      const int64_t Bq = std::min(std::max(lo + kRingLoHalf, A), Q);
      bnd[a][0] = P; bnd[a][1] = A; bnd[a][2] = Bq; bnd[a][3] = Q;
    }
    if (c->timing) HIPCHK(hipEventRecord(c->rev[0], c->st));
    hipLaunchKernelGGL(k_ring_coords, dim3(nblk(3 * n2)), dim3(256), 0, c->st, c->B, R);
    RingSpan sp;
    for (int a = 0; a < 3; a++) { sp.s[a] = 0; sp.e[a] = bnd[a][0]; }
    hipLaunchKernelGGL(k_ring_walk, dim3(3), dim3(64), 0, c->st, c->B, pc, R, sp, c->C, perr);
    if (c->timing) HIPCHK(hipEventRecord(c->rev[1], c->st));
    for (int wi = 0; wi < 2; wi++) {
      // the parallel phase [cur, end) after the sequential one that ends at cur
      int64_t cur[3], end[3];
      for (int a = 0; a < 3; a++) {
        cur[a] = bnd[a][2 * wi];
        end[a] = bnd[a][2 * wi + 1];
      }
      for (int it = 0;; it++) {
        int64_t tot = 0;
        for (int a = 0; a < 3; a++) {
          w.i0[a] = cur[a];
          w.i1[a] = end[a];
          w.ws[a] = std::max<int64_t>(1, cur[a] - kRingWindow);
          tot = std::max(tot, end[a] > cur[a] ? end[a] - w.ws[a] : 0);
        }
        if (tot == 0) break;
        hipLaunchKernelGGL(k_ring_win_init, dim3(1), dim3(64), 0, c->st, R, w);
        hipLaunchKernelGGL(k_ring_win_scan, dim3(nblk(tot), 3), dim3(256), 0, c->st, R, w);
        hipLaunchKernelGGL(k_ring_win_check, dim3(1), dim3(64), 0, c->st, R, w);
        hipLaunchKernelGGL(k_ring_win_viol, dim3(nblk(tot), 3), dim3(256), 0, c->st, R, w);
        int64_t ext[3];
        HIPCHK(hipMemcpyAsync(ext, w.ext, sizeof(ext), hipMemcpyDeviceToHost, c->st));
        HIPCHK(hipStreamSynchronize(c->st));
        bool any = false;
        RingSpan xs;
        for (int a = 0; a < 3; a++) {
          xs.s[a] = xs.e[a] = 0;
          if (ext[a] > 0 && end[a] > cur[a]) {
            xs.s[a] = cur[a];
            xs.e[a] = std::min(end[a], std::max(ext[a], cur[a] + 64));
            cur[a] = xs.e[a];
            any = true;
          }
        }
        if (!any) break;
        if (it > 4 * n2 / 64 + 8) throw HipError{"periodic collider: the ring walk's phases did not settle"};
        hipLaunchKernelGGL(k_ring_walk, dim3(3), dim3(64), 0, c->st, c->B, pc, R, xs, c->C, perr);
        c->kms[KT_N + 3] += c->timing ? 1 : 0;
      }
      for (int a = 0; a < 3; a++) {
        if (end[a] <= cur[a]) continue;
        const int64_t ws = std::max<int64_t>(1, cur[a] - kRingWindow), m = end[a] - ws;
        hipLaunchKernelGGL(k_ring_keys, dim3(nblk(m)), dim3(256), 0, c->st, R, a, ws, m, w.r0, c->k0, c->v0);
        sdmrs::sort_pairs<uint64_t, uint32_t>(c->k0, c->v0, c->k1, c->v1, m, 64, c->rsScratch, c->st);
        sdmrs::scan<sdmrs::OpMin, true, true>(c->v0, c->u0, m, c->rsScratch, c->st);
        hipLaunchKernelGGL(k_ring_gather, dim3(nblk(m)), dim3(256), 0, c->st, R, a, ws, m, c->v0, c->rth, c->rtp,
                           c->rtx);
        RingCross X;
        X.vals = c->v0;
        X.smin = c->u0;
        X.hn = c->rth;
        X.ws = ws;
        X.ins = cur[a] - ws;
        X.m = m;
        X.a = a;
        X.longList = c->u1;
        X.longCount = (unsigned long long*)(c->rdev + 30);
        X.longCap = m;
        X.box = c->pbox;
        HIPCHK(hipMemsetAsync(X.longCount, 0, sizeof(unsigned long long), c->st));
        hipLaunchKernelGGL(k_ring_cross, dim3(nblk(m)), dim3(256), 0, c->st, c->B, pc, X, c->C, perr);
        hipLaunchKernelGGL(k_ring_cross_long, dim3(std::max(1, c->numCU) * 2), dim3(256), 0, c->st, c->B, pc, X, c->C,
                           perr);
        HIPCHK(hipMemcpyAsync(c->bb[a] + ws, c->rth, m * sizeof(int32_t), hipMemcpyDeviceToDevice, c->st));
        HIPCHK(hipMemcpyAsync(c->rper[a] + ws, c->rtp, m * sizeof(int32_t), hipMemcpyDeviceToDevice, c->st));
        HIPCHK(hipMemcpyAsync(c->rx[a] + ws, c->rtx, m * sizeof(double), hipMemcpyDeviceToDevice, c->st));
      }
      if (c->timing) HIPCHK(hipEventRecord(c->rev[2 + 2 * wi], c->st));
      for (int a = 0; a < 3; a++) {
        sp.s[a] = wi == 0 ? bnd[a][1] : bnd[a][3];
        sp.e[a] = wi == 0 ? bnd[a][2] : n2;
      }
      hipLaunchKernelGGL(k_ring_walk, dim3(3), dim3(64), 0, c->st, c->B, pc, R, sp, c->C, perr);
      if (c->timing) HIPCHK(hipEventRecord(c->rev[3 + 2 * wi], c->st));
    }
    int64_t back[5];
    unsigned long long cnt = 0;
    HIPCHK(hipMemcpyAsync(back, c->rdev, sizeof(back), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipMemcpyAsync(&cnt, c->dcnt + 2, sizeof(cnt), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    uint32_t errs[3];
    std::memcpy(errs, back + 3, sizeof(errs));
    if (c->timing) {
      float t[5];
      for (int k = 0; k < 5; k++) HIPCHK(hipEventElapsedTime(&t[k], c->rev[k], c->rev[k + 1]));
      c->kms[KT_RING_WALK] += t[0] + t[2] + t[4];
      c->kms[KT_RING_WIN] += t[1] + t[3];
      c->kcount[KT_RING_WALK] += 3;
      c->kcount[KT_RING_WIN] += 2;
    }
    for (int a = 0; a < 3; a++) {
      if (errs[a] & RING_ERR_NONFINITE) throw HipError{"periodic collider: a bound coordinate is not finite"};
      if (errs[a] & RING_ERR_RUN) throw HipError{"periodic collider: the ring walk did not end"};
    }
    if ((int64_t)cnt > c->C.cap) {
      if (tries > 2) throw HipError{"periodic collider: candidate buffer"};
      ensure_cands(c, (int64_t)cnt);
      continue;
    }
    for (int a = 0; a < 3; a++) c->loIdx[a] = back[a];
    return (int64_t)cnt;
  }
}

// InsertionSortCollider::action in a periodic cell (InsertionSortCollider.cpp:191-400 with
// its periodic branches).  First run: the rings sorted by the wrapped coordinates, the
// sweep's pairs from the cell list of periodic.h.  Later runs: the erase step by
// spatialOverlapPeri, then insertionSortPeri on the three rings (ring_run), whose inversions
// create the new pairs in the reference's order.
void collider_run_peri(sdm_ctx* c) {
  c->cnt.collider_runs++;
  c->skipValid = false;
  refresh_cid(c);
  const int64_t N = c->B.n;
  const int64_t n2 = 2 * N;
  const bool doInitSort = !c->bbValid;
  KT_START(c, KT_BOUNDS);
  hipLaunchKernelGGL(k_bounds, dim3(nblk(N)), dim3(256), 0, c->st, c->B, c->iter, c->prm.verlet_dist,
                     c->prm.min_sweep_dist_factor, c->prm.target_interv);
  KT_STOP(c, KT_BOUNDS);
  if (!c->strideActive && c->prm.verlet_dist > 0 && c->maxVelocitySq >= 0) c->strideActive = 1;
  ensure_scratch(c, std::max<int64_t>({n2 + 1, c->np + 1, 2 * c->np + 1}));
  ensure_moves(c, c->np + 1);

  // box widths (grid size of the first run) and the wall check: dcnt[7] error bits, [8]
  // width sum, [9] max
  unsigned long long* err = c->dcnt + 7;
  HIPCHK(hipMemsetAsync(c->dcnt + 7, 0, 3 * sizeof(unsigned long long), c->st));
  hipLaunchKernelGGL(k_width_stats, dim3(nblk(N)), dim3(256), 0, c->st, c->B, (double*)(c->dcnt + 8), c->dcnt + 9,
                     err);
  unsigned long long st[3] = {0, 0, 0};
  HIPCHK(hipMemcpyAsync(st, c->dcnt + 7, sizeof(st), hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  if (st[0] & 2) throw HipError{"walls (unbounded boxes) are not supported in a periodic cell"};
  double wsum, wmax;
  std::memcpy(&wsum, &st[1], 8);
  std::memcpy(&wmax, &st[2], 8);
  PeriCell pc;
  const double wmean = N > 0 ? wsum / (double)N : 0.0;
  const double ht = std::max(2.0 * wmean, 0.5 * wmax);
  int64_t ncell = 1;
  for (int k = 0; k < 3; k++) {
    pc.L[k] = c->prm.cell[k];
    int64_t n = ht > 0 ? (int64_t)std::floor(pc.L[k] / ht) : 1;
    pc.n[k] = (int32_t)std::min<int64_t>(std::max<int64_t>(n, 1), 1024);
  }
  while ((int64_t)pc.n[0] * pc.n[1] * pc.n[2] > (1ll << 22)) {
    int k = (pc.n[0] >= pc.n[1] && pc.n[0] >= pc.n[2]) ? 0 : (pc.n[1] >= pc.n[2] ? 1 : 2);
    pc.n[k] = std::max(1, pc.n[k] / 2);
  }
  for (int k = 0; k < 3; k++) {
    pc.h[k] = pc.L[k] / pc.n[k];
    pc.Lf[k] = (float)pc.L[k];
    pc.iLf[k] = (float)(1.0 / pc.L[k]);
    pc.mf[k] = (float)(1e-5 * pc.L[k]);  // the screen's rounding stays below 1e-6 L
  }
  ncell = (int64_t)pc.n[0] * pc.n[1] * pc.n[2];
  pc.eps = 1e-9 * std::max(pc.L[0], std::max(pc.L[1], pc.L[2]));

  // ---- conditionalyEraseNonReal ----
  if (c->np > 0) {
    hipLaunchKernelGGL(k_erase_flags_peri, dim3(nblk(c->np)), dim3(256), 0, c->st, c->B, c->P, c->np, pc, c->u0, err);
    erase_flagged(c);
  }
  if ((read_counter(c, 7) & 1) != 0) throw HipError{"body larger than half of the cell size encountered"};

  int64_t nc = 0;
  if (!doInitSort) {
    // ---- insertionSortPeri on the three rings ----
    nc = ring_run(c, pc);
  } else {
    // ---- first run: the rings sorted by the wrapped coordinates, loIdx = 0 ----
    for (int a = 0; a < 3; a++) {
      hipLaunchKernelGGL(k_peri_bound_keys, dim3(nblk(n2)), dim3(256), 0, c->st, c->B, pc, a, n2, c->k0, c->v0);
      sdmrs::sort_pairs<uint64_t, uint32_t>(c->k0, c->v0, c->k1, c->v1, n2, 64, c->rsScratch, c->st);
      hipLaunchKernelGGL(k_ring_init, dim3(nblk(n2)), dim3(256), 0, c->st, c->B, pc, a, c->v0, n2, c->bb[a],
                         c->rper[a]);
      c->loIdx[a] = 0;
    }
    c->bbValid = true;
    c->numReinit++;
    if (!c->rankMin) {
      c->rankMin = dalloc<uint32_t>(N);
      c->rankMax = dalloc<uint32_t>(N);
      c->cellRng = dalloc<int32_t>(6 * N);
      c->peCnt = dalloc<uint32_t>(N + 1);
      c->peOff = dalloc<uint32_t>(N + 1);
    }
    if (!c->pbox) c->pbox = dalloc<float4>(2 * N);
    PeriOrder po;
    po.rankMin = c->rankMin;
    po.rankMax = c->rankMax;
    po.n2 = (uint32_t)n2;
    hipLaunchKernelGGL(k_peri_ranks, dim3(nblk(n2)), dim3(256), 0, c->st, (const uint32_t*)c->bb[c->prm.sort_axis],
                       n2, c->rankMin, c->rankMax);

    // ---- the sweep's pairs from a cell list ----
    hipLaunchKernelGGL(k_peri_ranges, dim3(nblk(N)), dim3(256), 0, c->st, c->B, pc, c->cellRng, c->peCnt, c->pbox);
    const int64_t m = exclusive_total(c, c->peCnt, c->peOff, N);
    ensure_scratch(c, m + 1);
    uint32_t* ckey = (uint32_t*)c->k0;
    uint32_t* cval = c->v0;
    hipLaunchKernelGGL(k_peri_entries, dim3(nblk(N)), dim3(256), 0, c->st, N, pc, c->cellRng, c->peOff, ckey, cval);
    int bits = 1;
    while ((1ll << bits) < ncell) bits++;
    sdmrs::sort_pairs<uint32_t, uint32_t>(ckey, cval, (uint32_t*)c->k1, c->v1, m, bits, c->rsScratch, c->st);
    if (ncell > c->runCap) {
      HIPCHK(hipStreamSynchronize(c->st));
      dfree(c->runEnd);
      c->runCap = ncell;
      c->runEnd = dalloc<uint32_t>(2 * ncell);
    }
    uint32_t* runStart = c->runEnd + c->runCap;
    HIPCHK(hipMemsetAsync(c->runEnd, 0, ncell * sizeof(uint32_t), c->st));
    HIPCHK(hipMemsetAsync(runStart, 0, ncell * sizeof(uint32_t), c->st));
    hipLaunchKernelGGL(k_cell_ends, dim3(nblk(m)), dim3(256), 0, c->st, ckey, m, runStart, c->runEnd);
    ensure_hash(c, c->np + 1024);
    const uint64_t mask = (uint64_t)c->hcap - 1;
    hipLaunchKernelGGL(k_fill_u64, dim3(nblk(c->hcap)), dim3(256), 0, c->st, c->htab, c->hcap, SDM_HEMPTY);
    if (c->np > 0)
      hipLaunchKernelGGL(k_hash_insert_pairs, dim3(nblk(c->np)), dim3(256), 0, c->st, c->P.cid, c->np, c->htab, mask);
    for (int pass = 0; pass < 2; pass++) {
      HIPCHK(hipMemsetAsync(c->dcnt + 2, 0, sizeof(unsigned long long), c->st));
      c->C.count = c->dcnt + 2;
      hipLaunchKernelGGL(k_peri_cell_pairs, dim3((unsigned)ncell), dim3(64), 0, c->st, c->B, pc, po, c->cellRng,
                         c->pbox, cval, runStart, c->runEnd, c->htab, mask, c->C, err);
      nc = (int64_t)read_counter(c, 2);
      if (nc <= c->C.cap) break;
      ensure_cands(c, nc);  // rerun with room for all
    }
  }
  if ((read_counter(c, 7) & 1) != 0) throw HipError{"body larger than half of the cell size encountered"};

  // ---- append the new pairs in order-key order ----
  if (nc > 0) {
    const int64_t start = c->np;
    ensure_hash(c, c->np + nc);
    fill_found(c);
    const int64_t ns = append_candidates(c, nc);
    if (ns > 0)
      hipLaunchKernelGGL(k_new_celldist, dim3(nblk(ns)), dim3(256), 0, c->st, c->P, start, ns, c->B, pc, err);
  }
  build_csr(c);
}

// The first run's pairs of a one-specimen aperiodic scene without the sweep (periodic.h,
// "Initial sort-and-sweep"): ranks of the sorted bounds of the sort axis, a cell list of the
// finite boxes, the walls against every body.  Candidates in C; returns their number.
int64_t init_pairs_cells(sdm_ctx* c, const int32_t* V) {
  const int64_t N = c->B.n;
  const int64_t n2 = 2 * N;
  if (!c->rankMin) {
    c->rankMin = dalloc<uint32_t>(N);
    c->rankMax = dalloc<uint32_t>(N);
  }
  if (!c->peCnt) {
    c->peCnt = dalloc<uint32_t>(N + 1);
    c->peOff = dalloc<uint32_t>(N + 1);
  }
  if (!c->cellRng) c->cellRng = dalloc<int32_t>(6 * N);
  hipLaunchKernelGGL(k_peri_ranks, dim3(nblk(n2)), dim3(256), 0, c->st, (const uint32_t*)V, n2, c->rankMin, c->rankMax);
  // union box of the finite boxes and their widths
  const int64_t nb = (int64_t)nblk(std::max<int64_t>(1, N));
  double* part = dalloc<double>(9 * nb + 9);
  hipLaunchKernelGGL(k_ic_stats, dim3((unsigned)nb), dim3(256), 0, c->st, c->B, part);
  hipLaunchKernelGGL(k_ic_stats_final, dim3(1), dim3(64), 0, c->st, part, nb, part + 9 * nb);
  double st[9];
  HIPCHK(hipMemcpyAsync(st, part + 9 * nb, sizeof(st), hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  dfree(part);
  // walls (infinite boxes): listed on the host, once
  std::vector<int32_t> kind(N > 0 ? N : 1);
  if (N > 0) HIPCHK(hipMemcpy(kind.data(), c->B.kind, N * sizeof(int32_t), hipMemcpyDeviceToHost));
  std::vector<int32_t> walls;
  for (int64_t i = 0; i < N; i++)
    if (kind[i] == SDM_KIND_WALL) walls.push_back((int32_t)i);
  int32_t* dWalls = nullptr;
  if (!walls.empty()) {
    dWalls = dalloc<int32_t>((int64_t)walls.size());
    HIPCHK(hipMemcpy(dWalls, walls.data(), walls.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  }
  // grid over the union box
  PeriCell pc;
  std::memset(&pc, 0, sizeof(pc));
  const bool finite = st[8] > 0;
  int64_t ncell = 1;
  if (finite) {
    const double wmean = st[6] / st[8], wmax = st[7];
    const double ht = std::max(2.0 * wmean, 0.5 * wmax);
    for (int k = 0; k < 3; k++) {
      const double ext = st[3 + k] - st[k];
      int64_t n = (ht > 0 && ext > 0) ? (int64_t)std::floor(ext / ht) : 1;
      pc.n[k] = (int32_t)std::min<int64_t>(std::max<int64_t>(n, 1), 1024);
    }
    while ((int64_t)pc.n[0] * pc.n[1] * pc.n[2] > (1ll << 22)) {
      int k = (pc.n[0] >= pc.n[1] && pc.n[0] >= pc.n[2]) ? 0 : (pc.n[1] >= pc.n[2] ? 1 : 2);
      pc.n[k] = std::max(1, pc.n[k] / 2);
    }
    for (int k = 0; k < 3; k++) {
      const double ext = st[3 + k] - st[k];
      pc.h[k] = ext > 0 ? ext / pc.n[k] : 1.0;
      pc.L[k] = ext;
    }
    ncell = (int64_t)pc.n[0] * pc.n[1] * pc.n[2];
  }
  int64_t m = 0;
  uint32_t* ckey = nullptr;
  uint32_t* cval = nullptr;
  uint32_t* runStart = nullptr;
  if (finite) {
    hipLaunchKernelGGL(k_ic_ranges, dim3(nblk(N)), dim3(256), 0, c->st, c->B, pc, st[0], st[1], st[2], c->cellRng,
                       c->peCnt);
    m = exclusive_total(c, c->peCnt, c->peOff, N);
    ensure_scratch(c, m + 1);
    ckey = (uint32_t*)c->k0;
    cval = c->v0;
    hipLaunchKernelGGL(k_peri_entries, dim3(nblk(N)), dim3(256), 0, c->st, N, pc, c->cellRng, c->peOff, ckey, cval);
    int bits = 1;
    while ((1ll << bits) < ncell) bits++;
    sdmrs::sort_pairs<uint32_t, uint32_t>(ckey, cval, (uint32_t*)c->k1, c->v1, m, bits, c->rsScratch, c->st);
    if (ncell > c->runCap) {
      HIPCHK(hipStreamSynchronize(c->st));
      dfree(c->runEnd);
      c->runCap = ncell;
      c->runEnd = dalloc<uint32_t>(2 * ncell);
    }
    runStart = c->runEnd + c->runCap;
    HIPCHK(hipMemsetAsync(c->runEnd, 0, ncell * sizeof(uint32_t), c->st));
    HIPCHK(hipMemsetAsync(runStart, 0, ncell * sizeof(uint32_t), c->st));
    if (m > 0) hipLaunchKernelGGL(k_cell_ends, dim3(nblk(m)), dim3(256), 0, c->st, ckey, m, runStart, c->runEnd);
  }
  int64_t nc = 0;
  for (int pass = 0; pass < 2; pass++) {
    HIPCHK(hipMemsetAsync(c->dcnt + 2, 0, sizeof(unsigned long long), c->st));
    c->C.count = c->dcnt + 2;
    if (m > 0)
      hipLaunchKernelGGL(k_ic_cell_pairs, dim3((unsigned)ncell), dim3(64), 0, c->st, c->B, pc, c->cellRng, cval,
                         runStart, c->runEnd, c->rankMin, c->rankMax, c->C);
    if (!walls.empty())
      hipLaunchKernelGGL(k_ic_walls, dim3(nblk(N)), dim3(256), 0, c->st, c->B, dWalls, (int)walls.size(), c->rankMin,
                         c->rankMax, c->C);
    nc = (int64_t)read_counter(c, 2);
    if (nc <= c->C.cap) break;
    ensure_cands(c, nc);  // rerun with room for all
  }
  HIPCHK(hipStreamSynchronize(c->st));
  dfree(dWalls);
  return nc;
}

// InsertionSortCollider::action (InsertionSortCollider.cpp:191-400), non-periodic.
void collider_run(sdm_ctx* c) {
  if (c->prm.periodic) {
    collider_run_peri(c);
    return;
  }
  c->cnt.collider_runs++;
  c->skipValid = false;
  refresh_cid(c);
  const int64_t N = c->B.n;
  const int64_t n2 = 2 * N;
  bool doInitSort = false;
  if (!c->bbValid) {
    // empty bound arrays: fill in body order (min, max) and do the initial sort
    std::vector<int32_t> h(n2);
    for (int64_t i = 0; i < n2; i++) h[i] = (int32_t)i;
    for (int a = 0; a < 3; a++) HIPCHK(hipMemcpy(c->bb[a], h.data(), n2 * sizeof(int32_t), hipMemcpyHostToDevice));
    c->bbValid = true;
    doInitSort = true;
  }
  // BoundDispatcher::action with sweepDist = verletDist
  KT_START(c, KT_BOUNDS);
  hipLaunchKernelGGL(k_bounds, dim3(nblk(N)), dim3(256), 0, c->st, c->B, c->iter, c->prm.verlet_dist,
                     c->prm.min_sweep_dist_factor, c->prm.target_interv);
  KT_STOP(c, KT_BOUNDS);
  if (!c->strideActive && c->prm.verlet_dist > 0 && c->maxVelocitySq >= 0) c->strideActive = 1;
  ensure_scratch(c, std::max<int64_t>({n2 + 1, c->np + 1, 2 * c->np + 1}));
  ensure_moves(c, c->np + 1);

  // ---- conditionalyEraseNonReal ----
  if (c->np > 0) {
    hipLaunchKernelGGL(k_erase_flags, dim3(nblk(c->np)), dim3(256), 0, c->st, c->B, c->P, c->np, c->u0);
    erase_flagged(c);
  }

  // ---- sort the bounds of each axis ----
  for (int a = 0; a < 3; a++) {
    hipLaunchKernelGGL(k_bb_keys, dim3(nblk(n2)), dim3(256), 0, c->st, c->B, c->bb[a], n2, a, c->k0, c->v0);
    sdmrs::sort_pairs<uint64_t, uint32_t>(c->k0, c->v0, c->k1, c->v1, n2, 64, c->rsScratch, c->st);
    if (c->nSpec > 1) {
      uint32_t* sk = (uint32_t*)c->k0;
      hipLaunchKernelGGL(k_spec_keys, dim3(nblk(n2)), dim3(256), 0, c->st, c->B.specimen, c->bb[a], c->v0, n2, sk);
      int bits = 1;
      while ((1ll << bits) < c->nSpec) bits++;
      sdmrs::sort_pairs<uint32_t, uint32_t>(sk, c->v0, (uint32_t*)c->k1, c->v1, n2, bits, c->rsScratch, c->st);
    }
    hipLaunchKernelGGL(k_bb_apply, dim3(nblk(n2)), dim3(256), 0, c->st, c->bb[a], c->v0, n2, c->bbTmp);
    HIPCHK(hipMemcpyAsync(c->qs[a], c->v0, n2 * sizeof(uint32_t), hipMemcpyDeviceToDevice, c->st));
    std::swap(c->bb[a], c->bbTmp);
  }

  // ---- candidates ----
  int64_t nc = 0;
  if (!doInitSort) {
    for (int pass = 0; pass < 2; pass++) {
      HIPCHK(hipMemsetAsync(c->dcnt + 2, 0, sizeof(unsigned long long), c->st));
      for (int a = 0; a < 3; a++) {
        sdmrs::scan<sdmrs::OpMin, true, true>(c->qs[a], c->smin, n2, c->rsScratch, c->st);
        hipLaunchKernelGGL(k_crossings, dim3(nblk(n2)), dim3(256), 0, c->st, c->B, c->bb[a], c->qs[a], c->smin, n2, a,
                           c->C);
      }
      unsigned long long cnt = 0;
      HIPCHK(hipMemcpyAsync(&cnt, c->dcnt + 2, sizeof(cnt), hipMemcpyDeviceToHost, c->st));
      HIPCHK(hipStreamSynchronize(c->st));
      nc = (int64_t)cnt;
      if (nc <= c->C.cap) break;
      ensure_cands(c, nc);  // rerun with room for all
    }
  } else {
    c->numReinit++;
    const int32_t* V = c->bb[c->prm.sort_axis];
    if (c->nSpec == 1) {
      nc = init_pairs_cells(c, V);
    } else {
      // several specimens (each a few thousand bodies, spatially superposed): the sweep itself
      uint32_t* counts = c->u0;
      uint32_t* offs = c->u1;
      hipLaunchKernelGGL(k_init_sweep<false>, dim3(nblk(n2)), dim3(256), 0, c->st, c->B, V, n2, counts,
                         (const uint32_t*)nullptr, c->C);
      nc = exclusive_total(c, counts, offs, n2);
      ensure_cands(c, nc);
      hipLaunchKernelGGL(k_init_sweep<true>, dim3(nblk(n2)), dim3(256), 0, c->st, c->B, V, n2, (uint32_t*)nullptr,
                         offs, c->C);
    }
  }

  if (nc > 0) {
    ensure_hash(c, c->np + nc);
    fill_found(c);
    append_candidates(c, nc);
  }
  build_csr(c);
}


// Grid of a persistent solver: the blocks that are resident at once (occupancy x CUs), or
// fewer when the list is short.  Each lane then works through several pairs, so the lanes
// a short solve frees are refilled from the list instead of idling until the wave ends.
template <typename K>
unsigned resident_grid(sdm_ctx* c, K kernel, int& cache, int64_t nwork) {
  if (cache == 0) {
    int nb = 0;
    HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, 256, 0));
    cache = std::max(1, nb) * std::max(1, c->numCU);
  }
  if (c->tune.solver_blocks_per_cu > 0)
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>((nwork + 255) / 256,
                                                            (int64_t)c->tune.solver_blocks_per_cu * std::max(1, c->numCU)));
  const int64_t b = (nwork + 255) / 256;
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(b, cache));
}

// Nelder-Mead pass of the Geom2 functor over the queued pairs: the first evaluation of every
// pair (dense, one pair per lane), then the persistent solver over the touching ones, longest
// expected solves first, the mixed form's double refinement, and the finish kernel.
template <typename T>
void nm_pass(sdm_ctx* c, const StepArgs& A, SolveArgs SA) {
  // every size below is known only on the device; launches cover the upper bound np
  const int64_t nmax = A.np;
  ensure_scratch(c, nmax + 1);
  SA.evals = c->dcnt + 5;
  KT_START(c, KT_FIRST);
  hipLaunchKernelGGL(k_nm_first<T>, dim3(nblk(nmax)), dim3(256), 0, c->st, A, SA, c->qflag, c->k1);
  KT_STOP(c, KT_FIRST);
  KT_START(c, KT_LIST);
  scan_to_counter(c, c->qflag, c->qpos, nmax, DC_NT);
  const unsigned long long* nt = c->dcnt + DC_NT;
  hipLaunchKernelGGL(k_list_fill, dim3(nblk(nmax)), dim3(256), 0, c->st, c->qflag, c->qpos, c->k1, nmax, c->k0, c->v0);
  // the list in queue order (increasing interaction index) for the finish kernel, whose
  // record reads and writes then follow memory order; with sort_by_cost, a copy sorted by
  // expected cost for the solver (the solver reads only the first nt entries: no padding
  // without the sort)
  const uint32_t* byPair = c->v0;
  if (c->tune.sort_by_cost) {
    hipLaunchKernelGGL(k_pad_keys, dim3(nblk(nmax)), dim3(256), 0, c->st, c->k0, c->v0, nt, nmax);
    hipLaunchKernelGGL(k_copy_u32, dim3(nblk(nmax)), dim3(256), 0, c->st, c->u0, c->v0, nt, nmax);
    byPair = c->u0;
    sdmrs::sort_pairs<uint64_t, uint32_t>(c->k0, c->v0, c->k1, c->v1, nmax, 8, c->rsScratch, c->st);
  }
  KT_STOP(c, KT_LIST);
  const uint32_t* list = c->v0;
  SolveArgs S2 = SA;
  S2.list = reinterpret_cast<const int32_t*>(list);
  S2.nwork = nmax;
  S2.count = nt;
  S2.evals = c->dcnt + 4;
  S2.next = c->dcnt + DC_NEXT;
  S2.dbl = c->dcnt + DC_NDBL;
  S2.hoList = c->retryList;
  S2.hoCount = c->dcnt + DC_NHO;
  zero_counter(c, DC_NEXT);
  zero_counter(c, DC_NDBL);
  zero_counter(c, DC_NHO);
  const int ti = sizeof(T) == 4 ? 0 : 1;
  const unsigned grid = resident_grid(c, k_solve_nm<T>, c->residentNM[ti], nmax);
  if (c->timing) HIPCHK(hipEventRecord(c->ev[4], c->st));
  KT_START(c, KT_SOLVE);
  hipLaunchKernelGGL(k_solve_nm<T>, dim3(grid), dim3(256), 0, c->st, S2);
  KT_STOP(c, KT_SOLVE);
  if (c->timing) HIPCHK(hipEventRecord(c->ev[5], c->st));
  c->solverRan = true;
  if (sizeof(T) == 4) {
    // the pairs the main launch handed over, with their double evaluations
    SolveArgs R = S2;
    R.list = c->retryList;
    R.count = c->dcnt + DC_NHO;
    R.hoList = nullptr;
    R.hoCount = nullptr;
    zero_counter(c, DC_NEXT);
    KT_START(c, KT_RESUME);
    hipLaunchKernelGGL((k_solve_nm<float, true>), dim3(resident_grid(c, k_solve_nm<float, true>, c->residentNM[2], nmax)),
                       dim3(256), 0, c->st, R);
    KT_STOP(c, KT_RESUME);
    SolveArgs RF = S2;
    RF.list = reinterpret_cast<const int32_t*>(byPair);
    RF.evals = c->dcnt + 5;
    KT_START(c, KT_REFINE);
    hipLaunchKernelGGL(k_nm_refine, dim3(nblk(nmax)), dim3(256), 0, c->st, RF);
    KT_STOP(c, KT_REFINE);
  }
  // the mixed form takes the curvature with the device's double functions
  KT_START(c, KT_FINISH);
  hipLaunchKernelGGL((k_pairs_finish<1, sizeof(T) == 4>), dim3(nblk(nmax)), dim3(256), 0, c->st, A, c->S, byPair, nmax,
                     nt);
  KT_STOP(c, KT_FINISH);
}

// InteractionLoop::action: prep (every interaction), wall / sphere pairs, solvers (queued
// pairs), finish.  No host synchronisation: the queue sizes stay on the device.
template <int LAW>
void interaction_loop(sdm_ctx* c, const StepArgs& A) {
  zero_counter(c, DC_NQ);
  if (A.np <= 0) return;
  ensure_queue(c, A.np);
  ensure_scratch(c, A.np + 1);
  zero_counter(c, DC_MISC);
  KT_START(c, KT_PREP);
  hipLaunchKernelGGL((k_pairs_prep<LAW>), dim3(nblk(A.np)), dim3(256), 0, c->st, A, c->qflag, c->miscList,
                     c->dcnt + DC_MISC);
  KT_STOP(c, KT_PREP);
  // the solver queue: scan of the flags (no shared counter), then the slots in interaction order
  KT_START(c, KT_QUEUE);
  scan_to_counter(c, c->qflag, c->qpos, A.np, DC_NQ);
  const unsigned long long* nq = c->dcnt + DC_NQ;
  hipLaunchKernelGGL((k_queue_fill<LAW>), dim3(nblk(A.np)), dim3(256), 0, c->st, A, c->S, c->qflag, c->qpos, A.np);
  KT_STOP(c, KT_QUEUE);
  KT_START(c, KT_MISC);
  hipLaunchKernelGGL((k_pairs_misc<LAW>), dim3(std::max(1, c->numCU) * std::max(1, c->tune.misc_blocks_per_cu)),
                     dim3(256), 0, c->st, A, c->miscList,
                     c->dcnt + DC_MISC);
  KT_STOP(c, KT_MISC);
  SolveArgs SA;
  SA.geo = c->B.geo;
  SA.N = c->B.n;
  SA.S = c->S;
  SA.list = nullptr;
  SA.nwork = A.np;
  SA.count = nq;
  SA.next = c->dcnt + DC_NEXT;
  SA.evals = c->dcnt + 4;
  if (LAW == 0) {
    zero_counter(c, DC_NEXT);
    // The FP64 form solves in double with the correctly rounded functions, the mixed form in
    // double with the device's own; on two lanes per pair (each lane one body's support point)
    // at two waves per SIMD, the Geom functor's two restarts inline
    const bool mixedLM = c->prm.precision == SDM_MIXED;
    const auto lmKernel = mixedLM ? k_solve_lm<true, 2> : k_solve_lm<false, 2>;
    if (mixedLM) {
      SA.dbl = c->dcnt + DC_NDBL;
      zero_counter(c, DC_NDBL);
    }
    // the certified float screen of every pair's first evaluation, then the solver over the
    // pairs it leaves (in queue order)
    SolveArgs SF = SA;
    SF.evals = c->dcnt + 5;  // counted with the other evaluations outside the solver launch
    KT_START(c, KT_FIRST);
    hipLaunchKernelGGL(k_lm_first, dim3(nblk(A.np)), dim3(256), 0, c->st, SF, c->qflag);
    KT_STOP(c, KT_FIRST);
    KT_START(c, KT_LIST);
    scan_to_counter(c, c->qflag, c->qpos, A.np, DC_NT);
    hipLaunchKernelGGL(k_list_compact, dim3(nblk(A.np)), dim3(256), 0, c->st, c->qflag, c->qpos, A.np, c->v0);
    KT_STOP(c, KT_LIST);
    SolveArgs SL = SA;
    SL.list = reinterpret_cast<const int32_t*>(c->v0);
    SL.count = c->dcnt + DC_NT;
    SL.lmRestarts = true;
    const unsigned grid = resident_grid(c, lmKernel, c->residentLM, 2 * A.np);
    if (c->timing) HIPCHK(hipEventRecord(c->ev[4], c->st));
    KT_START(c, KT_SOLVE);
    hipLaunchKernelGGL(lmKernel, dim3(grid), dim3(256), 0, c->st, SL);
    KT_STOP(c, KT_SOLVE);
    if (c->timing) HIPCHK(hipEventRecord(c->ev[5], c->st));
    c->solverRan = true;
    KT_START(c, KT_FINISH);
    hipLaunchKernelGGL(k_pairs_finish<0>, dim3(nblk(A.np)), dim3(256), 0, c->st, A, c->S, nullptr, A.np, nq);
    KT_STOP(c, KT_FINISH);
  } else if (c->prm.precision == SDM_MIXED) {
    nm_pass<float>(c, A, SA);
  } else {
    nm_pass<double>(c, A, SA);
  }
}

bool collider_is_activated(sdm_ctx* c) {
  if (!c->strideActive) return true;
  if (c->fastestBodyMaxDist < 0) {
    c->fastestBodyMaxDist = 0;
    return true;
  }
  c->fastestBodyMaxDist = c->maxVelocitySq;
  if (c->fastestBodyMaxDist >= 1 || c->fastestBodyMaxDist == 0) return true;
  if (!c->bbValid) return true;
  return false;
}

// The part of a step after the collider: InteractionLoop, then the ordered force sums and
// NewtonIntegrator.  Kernel launches only (no allocation, no synchronisation): it can be
// captured into a graph.
void step_body(sdm_ctx* c, const StepArgs& A) {
  if (c->prm.law == SDM_LAW_LINEAR)
    interaction_loop<0>(c, A);
  else
    interaction_loop<1>(c, A);
  if (c->timing) HIPCHK(hipEventRecord(c->ev[2], c->st));
  NewtonArgs NA;
  NA.B = c->B;
  NA.P = c->P;
  NA.csrOff = c->csrOff;
  NA.csrPair = c->csrPair;
  NA.dt = c->prm.dt;
  NA.damping = c->prm.damping;
  NA.g[0] = c->prm.gravity[0];
  NA.g[1] = c->prm.gravity[1];
  NA.g[2] = c->prm.gravity[2];
  NA.exactAsph = c->prm.exact_aspherical_rot;
  NA.maxDisp = c->dMaxDisp;
  NA.quiet = c->servoOn ? &c->dServo->quiet : nullptr;
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, c->st, (uint32_t*)c->dMaxDisp, 1);
  KT_START(c, KT_FORCE_SUM);
  hipLaunchKernelGGL(k_force_sum, dim3((unsigned)((c->B.n + 15) / 16)), dim3(256), 0, c->st, c->B, c->P, c->csrOff,
                     c->csrPair);
  KT_STOP(c, KT_FORCE_SUM);
  if (c->servoOn) {
    // CompressionEngine between InteractionLoop and NewtonIntegrator; every decision on the device
    ServoDev* S = c->dServo;
    const unsigned long long* iterDev = (const unsigned long long*)(c->dcnt + DC_NP + 1);
    hipLaunchKernelGGL(k_servo_begin, dim3(1), dim3(64), 0, c->st, S, c->B, iterDev);
    hipLaunchKernelGGL(k_servo_gain, dim3(6), dim3(64), 0, c->st, S, c->P, c->csrOff, c->csrPair);
    hipLaunchKernelGGL(k_servo_step, dim3(1), dim3(64), 0, c->st, S, c->B, c->prm.dt, iterDev);
    hipLaunchKernelGGL(k_servo_gain, dim3(6), dim3(64), 0, c->st, S, c->P, c->csrOff, c->csrPair);
    const int64_t nb = (int64_t)nblk(std::max<int64_t>({(int64_t)1, A.np, (int64_t)c->B.n}));
    hipLaunchKernelGGL(k_servo_check_partial, dim3((unsigned)nb), dim3(256), 0, c->st, S, c->P,
                       (const unsigned long long*)(c->dcnt + DC_NP), c->B, c->prm.gravity[0], c->prm.gravity[1],
                       c->prm.gravity[2], c->servoPart);
    hipLaunchKernelGGL(k_servo_check_final, dim3(1), dim3(64), 0, c->st, S, c->B, c->servoPart, nb, iterDev);
  }
  KT_START(c, KT_NEWTON);
  hipLaunchKernelGGL(k_newton, dim3(nblk(c->B.n)), dim3(256), 0, c->st, NA);
  KT_STOP(c, KT_NEWTON);
  if (c->servoOn) hipLaunchKernelGGL(k_servo_end, dim3(1), dim3(64), 0, c->st, c->dServo);
}

void one_step(sdm_ctx* c) {
  c->solverRan = false;
  if (c->geoDirty) {
    hipLaunchKernelGGL(k_geo_pack, dim3(nblk(c->B.n)), dim3(256), 0, c->st, c->B);
    c->geoDirty = false;
  }
  if (c->timing) HIPCHK(hipEventRecord(c->ev[0], c->st));
  if (collider_is_activated(c)) collider_run(c);
  if (!c->csrValid) build_csr(c);
  if (c->timing) HIPCHK(hipEventRecord(c->ev[1], c->st));
  StepArgs A;
  A.B = c->B;
  A.P = c->P;
  A.mats = c->dmats;
  A.np = c->np;
  A.iter = c->iter;
  A.dt = c->prm.dt;
  A.counters = c->dcnt;
  A.csr = c->csrPair;
  A.fullWrite = c->pairsWritten ? 1 : 0;
  A.periodic = c->prm.periodic ? 1 : 0;
  A.skip = (c->skipValid && !c->pairsWritten) ? 1 : 0;
  A.cell[0] = c->prm.cell[0];
  A.cell[1] = c->prm.cell[1];
  A.cell[2] = c->prm.cell[2];
  c->pairsWritten = false;
  A.npDev = c->dcnt + DC_NP;
  A.iterDev = c->dcnt + DC_NP + 1;
  hipLaunchKernelGGL(k_set_u64x2, dim3(1), dim3(64), 0, c->st, c->dcnt + DC_NP, (unsigned long long)c->np,
                     (unsigned long long)c->iter);
  if (c->servoOn) {  // block partials of the servo's check, for launches up to the capacity
    const int64_t need = 11 * (int64_t)nblk(std::max<int64_t>({(int64_t)1, c->P.cap, (int64_t)c->B.n}));
    if (need > c->servoPartCap) {
      HIPCHK(hipStreamSynchronize(c->st));
      dfree(c->servoPart);
      c->servoPart = dalloc<double>(need);
      c->servoPartCap = need;
      c->graphGen++;
    }
  }
  // Steady state (no collider run, no host write since the last step, no timing): replay the
  // captured step; its launches cover the container's capacity and read np and iter above.
  const bool graphable = c->useGraph && !c->timing && A.fullWrite == 0 && A.skip == 1 && c->np > 0;
  if (graphable) {
    A.np = c->P.cap;
    ensure_queue(c, A.np);
    ensure_scratch(c, A.np + 1);
    resident_grid(c, c->prm.precision == SDM_MIXED ? k_solve_lm<true, 2> : k_solve_lm<false, 2>,
                  c->residentLM, 1);
    resident_grid(c, k_solve_nm<float>, c->residentNM[0], 1);
    resident_grid(c, k_solve_nm<double>, c->residentNM[1], 1);
    resident_grid(c, k_solve_nm<float, true>, c->residentNM[2], 1);
    if (!c->graphExec || c->graphBuiltGen != c->graphGen) {
      if (c->graphExec) {
        HIPCHK(hipGraphExecDestroy(c->graphExec));
        c->graphExec = nullptr;
      }
      hipGraph_t g = nullptr;
      HIPCHK(hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
      step_body(c, A);
      HIPCHK(hipStreamEndCapture(c->st, &g));
      HIPCHK(hipGraphInstantiate(&c->graphExec, g, nullptr, nullptr, 0));
      HIPCHK(hipGraphDestroy(g));
      c->graphBuiltGen = c->graphGen;
      c->graphs++;
    }
    HIPCHK(hipGraphLaunch(c->graphExec, c->st));
    c->cnt.graph_steps++;
  } else {
    step_body(c, A);
  }
  if (c->timing) HIPCHK(hipEventRecord(c->ev[3], c->st));
  int code = 0, servoDone = 0;
  unsigned long long nqueued = 0;
  HIPCHK(hipMemcpyAsync(&code, c->dMaxDisp, sizeof(int), hipMemcpyDeviceToHost, c->st));
  if (c->servoOn)
    HIPCHK(hipMemcpyAsync(&servoDone, &c->dServo->completedNow, sizeof(int), hipMemcpyDeviceToHost, c->st));
  unsigned long long nlisted[3] = {0, 0, 0};  // touching list, evaluations done in double, handed over
  if (c->timing) {
    HIPCHK(hipMemcpyAsync(&nqueued, c->dcnt + DC_NQ, sizeof(nqueued), hipMemcpyDeviceToHost, c->st));
    if (c->solverRan && c->prm.precision == SDM_MIXED) {
      HIPCHK(hipMemcpyAsync(&nlisted[1], c->dcnt + DC_NDBL, sizeof(unsigned long long), hipMemcpyDeviceToHost, c->st));
      if (c->prm.law == SDM_LAW_HERTZ) {
        HIPCHK(hipMemcpyAsync(&nlisted[0], c->dcnt + DC_NT, sizeof(unsigned long long), hipMemcpyDeviceToHost, c->st));
        HIPCHK(hipMemcpyAsync(&nlisted[2], c->dcnt + DC_NHO, sizeof(unsigned long long), hipMemcpyDeviceToHost, c->st));
      }
    }
  }
  HIPCHK(hipStreamSynchronize(c->st));
  if (c->timing) {
    c->tms[7] += (double)nqueued;
    c->kms[KT_N] += (double)nlisted[0];
    c->kms[KT_N + 1] += (double)nlisted[1];
    c->kms[KT_N + 2] += (double)nlisted[2];
  }
  if (c->timing) {
    float a = 0, b = 0, d = 0;
    HIPCHK(hipEventElapsedTime(&a, c->ev[0], c->ev[1]));
    HIPCHK(hipEventElapsedTime(&b, c->ev[1], c->ev[2]));
    HIPCHK(hipEventElapsedTime(&d, c->ev[2], c->ev[3]));
    c->tms[0] += a;
    c->tms[1] += b;
    c->tms[2] += d;
    c->tms[3] += 1;
    if (c->solverRan) {
      float e = 0;
      HIPCHK(hipEventElapsedTime(&e, c->ev[4], c->ev[5]));
      c->tms[4] += e;
    }
    for (int k = 0; k < KT_N; k++) {
      if (!c->kran[k]) continue;
      float e = 0;
      HIPCHK(hipEventElapsedTime(&e, c->kev[2 * k], c->kev[2 * k + 1]));
      c->kms[k] += e;
      c->kcount[k] += 1;
      c->kran[k] = false;
    }
  }
  c->maxVelocitySq = code == 2 ? 2.0 : (code == 1 ? 0.5 : 0.0);
  c->servoDone = servoDone != 0;
  c->skipValid = true;  // this step wrote every interaction's gap and every body's displacement
  c->iter++;
  c->time += c->prm.dt;
  c->cnt.steps++;
}

// ---- attribute table ----
enum Scope { S_BODY, S_PAIR, S_SCALAR, S_BOUNDS };
struct FieldInfo {
  const char* name;
  Scope scope;
  int dtype;
  int comps;
  int esize;
  size_t off;  // offset of the pointer member in BodyArrays, of the field in PairRec, or scalar id
};
template <typename T>
constexpr int dtype_of() {
  return sizeof(T) == 4 ? SDM_I32 : (std::is_same<T, double>::value ? SDM_F64 : SDM_I64);
}
const FieldInfo kFields[] = {
#define SDM_BF(T, m, k, name) {name, S_BODY, dtype_of<T>(), k, (int)sizeof(T), offsetof(BodyArrays, m)},
    SDM_BODY_FIELDS(SDM_BF)
#undef SDM_BF
#define SDM_PF(T, m, k, name) {name, S_PAIR, dtype_of<T>(), k, (int)sizeof(T), offsetof(PairRec, m)},
        SDM_PAIR_FIELDS(SDM_PF)
#undef SDM_PF
            {"scene.iter", S_SCALAR, SDM_I64, 1, 8, 0},
    {"scene.time", S_SCALAR, SDM_F64, 1, 8, 1},
    {"collider.strideActive", S_SCALAR, SDM_I32, 1, 4, 2},
    {"collider.fastestBodyMaxDist", S_SCALAR, SDM_F64, 1, 8, 3},
    {"newton.maxVelocitySq", S_SCALAR, SDM_F64, 1, 8, 4},
    {"collider.numReinit", S_SCALAR, SDM_I32, 1, 4, 5},
    {"collider.bounds", S_BOUNDS, SDM_I32, 1, 4, 0},
    {"collider.periods", S_BOUNDS, SDM_I32, 1, 4, 1},
    {"collider.loIdx", S_SCALAR, SDM_I64, 3, 8, 9},
    {"scene.dt", S_SCALAR, SDM_F64, 1, 8, 6},
    {"newton.damping", S_SCALAR, SDM_F64, 1, 8, 7},
    {"newton.gravity", S_SCALAR, SDM_F64, 3, 8, 8},
};

const FieldInfo* find_field(const char* name) {
  if (!name) return nullptr;
  for (const FieldInfo& f : kFields)
    if (std::strcmp(f.name, name) == 0) return &f;
  return nullptr;
}

int64_t field_count(sdm_ctx* c, const FieldInfo* f) {
  switch (f->scope) {
    case S_BODY: return (int64_t)c->B.n * f->comps;
    case S_PAIR: return c->np * f->comps;
    case S_SCALAR: return f->comps;
    case S_BOUNDS: return c->bbValid ? 6 * (int64_t)c->B.n : 0;
  }
  return 0;
}

// Fresh interactions for [from, to) after the host resized the container.
void init_fresh_pairs(sdm_ctx* c, int64_t from, int64_t to) {
  if (to <= from) return;
  const int64_t n = to - from;
  ensure_cands(c, n);
  // ids are overwritten by the caller's next "intr.id*" writes; use -1 placeholders
  std::vector<int32_t> ones(n, -1);
  HIPCHK(hipMemcpy(c->C.a, ones.data(), n * 4, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->C.b, ones.data(), n * 4, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_new_pairs, dim3(nblk(n)), dim3(256), 0, c->st, c->P, from, n, c->C, (const uint32_t*)nullptr,
                     c->iter);
  HIPCHK(hipStreamSynchronize(c->st));
}

}  // namespace

// ---------------------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------------------
extern "C" {

const char* sdm_version(void) { return SDM_VERSION; }

sdm_ctx* sdm_create(const sdm_params* params, int32_t n_bodies, int32_t n_materials) {
  if (!params || n_bodies < 0 || n_materials < 0) {
    g_err = "sdm_create: invalid arguments";
    return nullptr;
  }
  if (!(params->verlet_dist >= 0)) {
    g_err = "verletDist must be set explicitly (>= 0): the automatic value needs spheres";
    return nullptr;
  }
  if (params->law != SDM_LAW_LINEAR && params->law != SDM_LAW_HERTZ) {
    g_err = "sdm_create: unknown law";
    return nullptr;
  }
  if (params->precision != SDM_FP64 && params->precision != SDM_MIXED) {
    g_err = "sdm_create: unknown precision";
    return nullptr;
  }
  if (params->sort_axis < 0 || params->sort_axis > 2) {
    g_err = "sdm_create: sortAxis must be 0, 1 or 2";
    return nullptr;
  }
  if (params->periodic && !(params->cell[0] > 0 && params->cell[1] > 0 && params->cell[2] > 0 &&
                            std::isfinite(params->cell[0] + params->cell[1] + params->cell[2]))) {
    g_err = "sdm_create: a periodic cell needs three positive finite sizes";
    return nullptr;
  }
  sdm_ctx* c = new sdm_ctx();
  try {
    c->prm = *params;
    // the cost-ordered Nelder-Mead list pays off for the FP64 solver's long evaluations
    // (measured at 200k: 2.71 against 3.70 ms per solve), not for the mixed form's (0.91 against
    // 0.92 ms, and the sort itself costs 0.06 ms)
    c->tune.sort_by_cost = params->precision == SDM_MIXED ? 0 : 1;
    HIPCHK(hipGetDevice(&c->dev));
    HIPCHK(hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking));
    std::memset(&c->P, 0, sizeof(c->P));
    std::memset(&c->C, 0, sizeof(c->C));
    std::memset(&c->S, 0, sizeof(c->S));
    alloc_bodies(c, n_bodies);
    // defaults of a fresh Body/State/Bound: identity orientation, damped, NaN bound data
    const int64_t N = n_bodies;
    std::vector<double> nan(N * 3, std::numeric_limits<double>::quiet_NaN());
    std::vector<int32_t> zi(N, 0), onei(N, 1);
#define SDM_ZB(T, m, k, name) HIPCHK(hipMemset(c->B.m, 0, N * k * sizeof(T) + (N == 0 ? 1 : 0)));
    SDM_BODY_FIELDS(SDM_ZB)
#undef SDM_ZB
    if (N > 0) {
      std::vector<double> one(N, 1.0);
      HIPCHK(hipMemcpy(c->B.ori, one.data(), N * 8, hipMemcpyHostToDevice));
      HIPCHK(hipMemcpy(c->B.damped, onei.data(), N * 4, hipMemcpyHostToDevice));
      std::vector<int32_t> iota(N);
      for (int64_t i = 0; i < N; i++) iota[i] = (int32_t)i;
      HIPCHK(hipMemcpy(c->B.lid, iota.data(), N * 4, hipMemcpyHostToDevice));
      HIPCHK(hipMemcpy(c->B.refPos, nan.data(), N * 24, hipMemcpyHostToDevice));
      HIPCHK(hipMemcpy(c->B.bmin, nan.data(), N * 24, hipMemcpyHostToDevice));
      HIPCHK(hipMemcpy(c->B.bmax, nan.data(), N * 24, hipMemcpyHostToDevice));
      for (int k : {0, 4, 8}) {
        HIPCHK(hipMemcpy(c->B.rotL + k * N, one.data(), N * 8, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(c->B.rotG + k * N, one.data(), N * 8, hipMemcpyHostToDevice));
      }
    }
    c->mats.resize(n_materials);
    c->dmats = dalloc<MaterialDev>(std::max(1, n_materials));
    c->dcnt = dalloc<unsigned long long>(kCounters);
    HIPCHK(hipMemset(c->dcnt, 0, kCounters * sizeof(unsigned long long)));
    HIPCHK(hipDeviceGetAttribute(&c->numCU, hipDeviceAttributeMultiprocessorCount, c->dev));
    c->dMaxDisp = dalloc<int>(1);
    ensure_pair_cap(c, 1024);
    ensure_cands(c, 4096);
    c->C.count = c->dcnt + 2;
    std::memset(&c->cnt, 0, sizeof(c->cnt));
  } catch (HipError& e) {
    g_err = e.msg;
    delete c;
    return nullptr;
  }
  return c;
}

void sdm_destroy(sdm_ctx* c) {
  if (!c) return;
  (void)hipStreamSynchronize(c->st);
#define SDM_FB(T, m, k, name) dfree(c->B.m);
  SDM_BODY_FIELDS(SDM_FB)
#undef SDM_FB
  dfree(c->B.geo);
  dfree(c->B.lid);
  dfree(c->B.disp);
  if (c->graphExec) (void)hipGraphExecDestroy(c->graphExec);
  free_pairs(c->P);
  free_queue(c->S);
  dfree(c->retryList);
  dfree(c->miscList); dfree(c->qflag); dfree(c->qpos);
  for (int a = 0; a < 3; a++) {
    dfree(c->bb[a]);
    dfree(c->qs[a]);
    dfree(c->rper[a]); dfree(c->rx[a]); dfree(c->rbh[a]); dfree(c->rbp[a]);
  }
  dfree(c->rth); dfree(c->rtp); dfree(c->rtx); dfree(c->rdev);
  dfree(c->bbTmp); dfree(c->smin); dfree(c->csrOff); dfree(c->csrPair);
  dfree(c->k0); dfree(c->k1); dfree(c->v0); dfree(c->v1); dfree(c->u0); dfree(c->u1); dfree(c->rsScratch);
  dfree(c->htab); dfree(c->dkeys); dfree(c->dvals);
  dfree(c->C.a); dfree(c->C.b); dfree(c->C.key);
  dfree(c->msrc); dfree(c->mdst); dfree(c->holePos); dfree(c->stage);
  dfree(c->cellRng); dfree(c->rankMin); dfree(c->rankMax);
  dfree(c->peCnt); dfree(c->peOff); dfree(c->runEnd); dfree(c->pbox);
  dfree(c->dmats); dfree(c->dcnt); dfree(c->dMaxDisp); dfree(c->dServo); dfree(c->servoPart);
  for (int k = 0; k < 6; k++)
    if (c->ev[k]) (void)hipEventDestroy(c->ev[k]);
  for (int k = 0; k < 2 * KT_N; k++)
    if (c->kev[k]) (void)hipEventDestroy(c->kev[k]);
  for (int k = 0; k < 6; k++)
    if (c->rev[k]) (void)hipEventDestroy(c->rev[k]);
  if (c->st) (void)hipStreamDestroy(c->st);
  delete c;
}

const char* sdm_last_error(const sdm_ctx* c) { return c ? c->err.c_str() : g_err.c_str(); }

int sdm_set_materials(sdm_ctx* c, const sdm_material* mats, int32_t n) {
  if (c) {
    c->skipValid = false;
    c->graphGen++;
  }
  if (!c || n < 0 || (n > 0 && !mats)) return fail(c, "sdm_set_materials: invalid arguments");
  try {
    c->matChecked = false;
    c->mats.assign(mats, mats + n);
    dfree(c->dmats);
    c->dmats = dalloc<MaterialDev>(std::max(1, n));
    std::vector<MaterialDev> md(std::max(1, n));
    for (int i = 0; i < n; i++)
      md[i] = MaterialDev{mats[i].density, mats[i].Kn, mats[i].Ks, mats[i].frictionAngle,
                          mats[i].betan,   mats[i].betas, mats[i].G, mats[i].nu, std::tan(mats[i].frictionAngle)};
    HIPCHK(hipMemcpy(c->dmats, md.data(), md.size() * sizeof(MaterialDev), hipMemcpyHostToDevice));
  } catch (HipError& e) {
    return fail(c, e.msg);
  }
  return 0;
}

// Superquadrics::Initial (pkg/dem/Superquadrics.cpp:47-93) with Mathr::beta/gammln
// (lib/base/Math.hpp:193-212): host-side set-up of one body.
static double sdm_gammln(const double xx) {
  static const double cof[14] = {57.1562356658629235,     -59.5979603554754912,    14.1360979747417471,
                                 -0.491913816097620199,   .339946499848118887e-4,  .465236289270485756e-4,
                                 -.983744753048795646e-4, .158088703224912494e-3,  -.210264441724104883e-3,
                                 .217439618115212643e-3,  -.164318106536763890e-3, .844182239838527433e-4,
                                 -.261908384015814087e-4, .368991826595316234e-5};
  double x = xx, y = xx;
  double tmp = x + 5.24218750000000000;
  tmp = (x + 0.5) * std::log(tmp) - tmp;
  double ser = 0.999999999999997092;
  for (int j = 0; j < 14; j++) ser += cof[j] / ++y;
  return tmp + std::log(2.5066282746310005 * ser / x);
}
static double sdm_beta(double z, double w) { return std::exp(sdm_gammln(z) + sdm_gammln(w) - sdm_gammln(z + w)); }

int sdm_superquadric_props(double rx, double ry, double rz, double eps1, double eps2, double* r_max, double* volume,
                           double inertia[3]) {
  if (!r_max || !volume || !inertia) return fail(nullptr, "sdm_superquadric_props: null output");
  if (!(rx > 0 && ry > 0 && rz > 0 && eps1 > 0 && eps2 > 0 && eps1 < 2 && eps2 < 2))
    return fail(nullptr, "sdm_superquadric_props: semi-axes must be > 0 and exponents in (0, 2)");
  double r = (rx > ry) ? rx : ry;
  r = r > rz ? r : rz;
  if ((eps1 < 1.0) || (eps2 < 1.0)) r *= 1.4142135623730951;  // pow(2.0, 0.5)
  *r_max = r;
  const double a = rx * ry * rz * eps1 * eps2;
  const double beta1 = sdm_beta(1.5 * eps1, 0.5 * eps1) * sdm_beta(0.5 * eps2, 2. * eps2 + 1);
  const double beta2 = sdm_beta(0.5 * eps1, 0.5 * eps1 + 1) * sdm_beta(1.5 * eps2, eps2 + 1);
  *volume = 2. * a * sdm_beta(eps1 * 0.5, eps1 * 0.5) * sdm_beta(eps2 * 0.5 + 1, eps2);
  inertia[0] = 0.5 * a * ((ry * ry) * beta1 + 4. * (rz * rz) * beta2);
  inertia[1] = 0.5 * a * ((rx * rx) * beta1 + 4. * (rz * rz) * beta2);
  inertia[2] = 0.5 * a * ((rx * rx) + (ry * ry)) * beta1;
  return 0;
}

int sdm_superquadric_props_n(int64_t n, const double* rxyz, const double* eps, double* r_max, double* volume,
                             double* inertia) {
  if (n < 0 || (n > 0 && (!rxyz || !eps || !r_max || !volume || !inertia)))
    return fail(nullptr, "sdm_superquadric_props_n: bad arguments");
  for (int64_t i = 0; i < n; i++) {
    const int rc = sdm_superquadric_props(rxyz[3 * i], rxyz[3 * i + 1], rxyz[3 * i + 2], eps[2 * i], eps[2 * i + 1],
                                          &r_max[i], &volume[i], &inertia[3 * i]);
    if (rc) return rc;
  }
  return 0;
}

const char* sdm_field_list(void) {
  static std::string s;
  if (s.empty())
    for (const FieldInfo& f : kFields) {
      s += f.name;
      s += "\n";
    }
  return s.c_str();
}

int sdm_field_info(sdm_ctx* c, const char* name, int32_t* dtype, int64_t* count) {
  const FieldInfo* f = find_field(name);
  if (!c || !f) return fail(c, std::string("unknown attribute ") + (name ? name : "(null)"));
  if (dtype) *dtype = f->dtype;
  if (count) *count = field_count(c, f);
  return 0;
}

int sdm_get(sdm_ctx* c, const char* name, void* dst, int64_t count) {
  const FieldInfo* f = find_field(name);
  if (!c || !f || (!dst && count > 0)) return fail(c, std::string("sdm_get: bad attribute ") + (name ? name : "(null)"));
  const int64_t n = field_count(c, f);
  if (count != n) return fail(c, std::string("sdm_get: count mismatch for ") + name);
  try {
    HIPCHK(hipStreamSynchronize(c->st));
    if (f->scope == S_PAIR) {
      // one row per interaction: gathered out of the records on the device, then one copy
      const int words = f->comps * f->esize / 4;
      if (c->np > 0) {
        uint32_t* buf = staging(c, c->np * words);
        hipLaunchKernelGGL(k_gather_field, dim3(nblk(c->np * words)), dim3(256), 0, c->st, c->P.r, c->np, (int)f->off,
                           words, buf);
        HIPCHK(hipMemcpyAsync(dst, buf, (size_t)c->np * words * 4, hipMemcpyDeviceToHost, c->st));
        HIPCHK(hipStreamSynchronize(c->st));
      }
    } else if (f->scope == S_BODY) {
      const int64_t rows = c->B.n;
      const int64_t stride = c->B.n;
      const char* dptr = *(char* const*)((const char*)&c->B + f->off);
      std::vector<char> tmp((size_t)rows * f->esize);
      char* out = (char*)dst;
      for (int k = 0; k < f->comps; k++) {
        if (rows > 0)
          HIPCHK(hipMemcpy(tmp.data(), dptr + (size_t)k * stride * f->esize, rows * f->esize, hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < rows; i++)
          std::memcpy(out + ((size_t)i * f->comps + k) * f->esize, tmp.data() + (size_t)i * f->esize, f->esize);
      }
    } else if (f->scope == S_SCALAR) {
      switch (f->off) {
        case 0: *(int64_t*)dst = c->iter; break;
        case 1: *(double*)dst = c->time; break;
        case 2: *(int32_t*)dst = c->strideActive; break;
        case 3: *(double*)dst = c->fastestBodyMaxDist; break;
        case 4: *(double*)dst = c->maxVelocitySq; break;
        case 5: *(int32_t*)dst = c->numReinit; break;
        case 6: *(double*)dst = c->prm.dt; break;
        case 7: *(double*)dst = c->prm.damping; break;
        case 8: std::memcpy(dst, c->prm.gravity, 3 * sizeof(double)); break;
        case 9: std::memcpy(dst, c->loIdx, 3 * sizeof(int64_t)); break;
      }
    } else {
      const int64_t n2 = 2 * (int64_t)c->B.n;
      for (int a = 0; a < 3 && c->bbValid; a++)
        HIPCHK(hipMemcpy((int32_t*)dst + a * n2, f->off == 0 ? c->bb[a] : c->rper[a], n2 * 4, hipMemcpyDeviceToHost));
    }
  } catch (HipError& e) {
    return fail(c, e.msg);
  }
  return 0;
}

int sdm_set(sdm_ctx* c, const char* name, const void* src, int64_t count) {
  const FieldInfo* f = find_field(name);
  if (!c || !f || (!src && count > 0)) return fail(c, std::string("sdm_set: bad attribute ") + (name ? name : "(null)"));
  if (count < 0) return fail(c, std::string("sdm_set: negative count for ") + name);
  c->skipValid = false;
  c->graphGen++;
  try {
    HIPCHK(hipStreamSynchronize(c->st));
    if (f->scope == S_PAIR && std::strcmp(name, "intr.id1") == 0) {
      const int64_t old = c->np;
      ensure_pair_cap(c, count);
      c->np = count;
      init_fresh_pairs(c, old, count);
      c->csrValid = false;
      c->cidValid = false;
    }
    if (f->scope == S_BOUNDS) {
      // collider.bounds: the handles of the three bound vectors, with every period reset to 0
      // as the reference's fresh Bounds have; collider.periods: the periods of those handles
      const int64_t n2 = 2 * (int64_t)c->B.n;
      if (count != 3 * n2) return fail(c, std::string(name) + ": count must be 6*N");
      for (int a = 0; a < 3; a++) {
        HIPCHK(hipMemcpy(f->off == 0 ? c->bb[a] : c->rper[a], (const int32_t*)src + a * n2, n2 * 4,
                         hipMemcpyHostToDevice));
        if (f->off == 0 && n2 > 0) HIPCHK(hipMemset(c->rper[a], 0, n2 * 4));
      }
      if (f->off == 0) c->bbValid = true;
      return 0;
    }
    const int64_t n = field_count(c, f);
    if (count != n) return fail(c, std::string("sdm_set: count mismatch for ") + name);
    if (f->scope == S_PAIR) {
      const int words = f->comps * f->esize / 4;
      if (c->np > 0) {
        uint32_t* buf = staging(c, c->np * words);
        HIPCHK(hipMemcpyAsync(buf, src, (size_t)c->np * words * 4, hipMemcpyHostToDevice, c->st));
        hipLaunchKernelGGL(k_scatter_field, dim3(nblk(c->np * words)), dim3(256), 0, c->st, c->P.r, c->np, (int)f->off,
                           words, buf);
        HIPCHK(hipStreamSynchronize(c->st));
      }
      c->csrValid = false;
      c->cidValid = false;
      c->pairsWritten = true;
    } else if (f->scope == S_BODY) {
      const int64_t rows = c->B.n;
      const int64_t stride = c->B.n;
      char* dptr = *(char* const*)((const char*)&c->B + f->off);
      std::vector<char> tmp((size_t)rows * f->esize);
      const char* in = (const char*)src;
      for (int k = 0; k < f->comps; k++) {
        for (int64_t i = 0; i < rows; i++)
          std::memcpy(tmp.data() + (size_t)i * f->esize, in + ((size_t)i * f->comps + k) * f->esize, f->esize);
        if (rows > 0)
          HIPCHK(hipMemcpy(dptr + (size_t)k * stride * f->esize, tmp.data(), rows * f->esize, hipMemcpyHostToDevice));
      }
      c->matChecked = false;
      c->geoDirty = true;
      if (std::strcmp(name, "body.specimen") == 0) set_specimens(c, (const int32_t*)src);
    } else {
      switch (f->off) {
        case 0: c->iter = *(const int64_t*)src; break;
        case 1: c->time = *(const double*)src; break;
        case 2: c->strideActive = *(const int32_t*)src; break;
        case 3: c->fastestBodyMaxDist = *(const double*)src; break;
        case 4: c->maxVelocitySq = *(const double*)src; break;
        case 5: c->numReinit = *(const int32_t*)src; break;
        case 6: c->prm.dt = *(const double*)src; break;
        case 7: c->prm.damping = *(const double*)src; break;
        case 8: std::memcpy(c->prm.gravity, src, 3 * sizeof(double)); break;
        case 9:
          for (int a = 0; a < 3; a++) {
            const int64_t v = ((const int64_t*)src)[a];
            if (v < 0 || (v > 0 && v >= 2 * (int64_t)c->B.n)) return fail(c, "collider.loIdx: outside the bound vector");
            c->loIdx[a] = v;
          }
          break;
      }
    }
  } catch (HipError& e) {
    return fail(c, e.msg);
  }
  return 0;
}

int sdm_step(sdm_ctx* c, int32_t nsteps) {
  if (!c) return fail(c, "null context");
  if (nsteps < 0) return fail(c, "sdm_step: nsteps < 0");
  try {
    if (!c->matChecked && c->B.n > 0) {
      std::vector<int32_t> mat(c->B.n);
      HIPCHK(hipMemcpy(mat.data(), c->B.material, c->B.n * 4, hipMemcpyDeviceToHost));
      for (int32_t m : mat)
        if (m < 0 || m >= (int)c->mats.size()) return fail(c, "body material index out of range");
      c->matChecked = true;
    }
    if (c->prm.periodic && c->nSpec > 1) return fail(c, "several specimens in one periodic cell are not supported");
    if (c->servoOn) {
      if (c->prm.periodic) return fail(c, "the wall servo needs the six walls of a cubic specimen, not a periodic cell");
      if (c->nSpec > 1) return fail(c, "the wall servo steps one specimen");
      if (c->B.n < 6) return fail(c, "the wall servo needs bodies 0-5 to be the walls");
      int32_t kind[6];
      HIPCHK(hipMemcpy(kind, c->B.kind, sizeof(kind), hipMemcpyDeviceToHost));
      for (int i = 0; i < 6; i++)
        if (kind[i] != SDM_KIND_WALL) return fail(c, "the wall servo needs bodies 0-5 to be the walls");
    }
    for (int s = 0; s < nsteps; s++) {
      one_step(c);
      if (c->servoOn && c->servoDone) break;  // stopAtIter = iter + 1
    }
  } catch (HipError& e) {
    return fail(c, e.msg);
  }
  return 0;
}

int sdm_sync(sdm_ctx* c) {
  if (!c) return -1;
  hipError_t e = hipStreamSynchronize(c->st);
  if (e != hipSuccess) return fail(c, hipGetErrorString(e));
  return 0;
}

int sdm_get_counters(sdm_ctx* c, sdm_counters* out) {
  if (!c || !out) return fail(c, "sdm_get_counters: bad arguments");
  try {
    unsigned long long d[2];
    HIPCHK(hipStreamSynchronize(c->st));
    HIPCHK(hipMemcpy(d, c->dcnt, sizeof(d), hipMemcpyDeviceToHost));
    *out = c->cnt;
    out->contact_evals = (int64_t)d[0];
    out->dist_evals = (int64_t)d[1];
  } catch (HipError& e) {
    return fail(c, e.msg);
  }
  return 0;
}

int sdm_set_timing(sdm_ctx* c, int32_t on) {
  if (!c) return -1;
  try {
    for (int k = 0; k < 6; k++)
      if (!c->ev[k]) HIPCHK(hipEventCreate(&c->ev[k]));
    for (int k = 0; k < 2 * KT_N; k++)
      if (!c->kev[k]) HIPCHK(hipEventCreate(&c->kev[k]));
    for (int k = 0; k < 6; k++)
      if (!c->rev[k]) HIPCHK(hipEventCreate(&c->rev[k]));
    HIPCHK(hipMemsetAsync(c->dcnt + 4, 0, 2 * sizeof(unsigned long long), c->st));
    HIPCHK(hipStreamSynchronize(c->st));
  } catch (HipError& e) {
    return fail(c, e.msg);
  }
  c->timing = on != 0;
  for (int k = 0; k < 8; k++) c->tms[k] = 0;
  for (int k = 0; k < KT_N; k++) {
    c->kms[k] = 0;
    c->kcount[k] = 0;
    c->kran[k] = false;
  }
  c->kms[KT_N] = c->kms[KT_N + 1] = c->kms[KT_N + 2] = c->kms[KT_N + 3] = 0;
  return 0;
}

int sdm_set_graph(sdm_ctx* c, int32_t on) {
  if (!c) return fail(c, "null context");
  c->useGraph = on != 0;
  c->tune.graph = on != 0 ? 1 : 0;
  return 0;
}

int sdm_set_tuning(sdm_ctx* c, const sdm_tuning* t) {
  if (!c || !t) return fail(c, "sdm_set_tuning: bad arguments");
  if (t->solver_blocks_per_cu < 0 || t->solver_blocks_per_cu > 64 || t->misc_blocks_per_cu < 1 ||
      t->misc_blocks_per_cu > 64)
    return fail(c, "sdm_set_tuning: blocks per CU out of range");
  c->tune = *t;
  c->useGraph = t->graph != 0;
  c->graphGen++;
  return 0;
}

int sdm_mean_coordination(sdm_ctx* c, double* cn) {
  if (!c || !cn) return fail(c, "sdm_mean_coordination: bad arguments");
  try {
    if (c->geoDirty) {  // body kinds written by the host since the last step
      hipLaunchKernelGGL(k_geo_pack, dim3(nblk(c->B.n)), dim3(256), 0, c->st, c->B);
      c->geoDirty = false;
    }
    unsigned long long* buf = dalloc<unsigned long long>(2);
    HIPCHK(hipMemsetAsync(buf, 0, 2 * sizeof(unsigned long long), c->st));
    const int64_t n = std::max<int64_t>(c->np, c->B.n);
    hipLaunchKernelGGL(k_cn_count, dim3(nblk(std::max<int64_t>(1, n))), dim3(256), 0, c->st, c->P, c->np, c->B,
                       (int)SDM_KIND_WALL, buf);
    unsigned long long h[2] = {0, 0};
    HIPCHK(hipMemcpyAsync(h, buf, sizeof(h), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    dfree(buf);
    if (h[1] == 0) return fail(c, "sdm_mean_coordination: no particles (every body is a wall)");
    *cn = 2.0 * (double)h[0] / (double)h[1];
  } catch (HipError& e) {
    return fail(c, e.msg);
  }
  return 0;
}

int sdm_stress(sdm_ctx* c, double volume, double sigma[9]) {
  if (!c || !sigma) return fail(c, "sdm_stress: bad arguments");
  try {
    if (volume == 0) volume = c->prm.periodic ? c->prm.cell[0] * c->prm.cell[1] * c->prm.cell[2] : 1.0;
    const int64_t nb = (int64_t)nblk(std::max<int64_t>(1, c->np));
    double* part = dalloc<double>(9 * nb + 9);
    hipLaunchKernelGGL(k_stress_partial, dim3((unsigned)nb), dim3(256), 0, c->st, c->P, c->np, c->B,
                       c->prm.periodic ? 1 : 0, c->prm.cell[0], c->prm.cell[1], c->prm.cell[2], part);
    hipLaunchKernelGGL(k_stress_final, dim3(1), dim3(64), 0, c->st, part, nb, volume, part + 9 * nb);
    HIPCHK(hipMemcpyAsync(sigma, part + 9 * nb, 9 * sizeof(double), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    dfree(part);
  } catch (HipError& e) {
    return fail(c, e.msg);
  }
  return 0;
}

// Scratch of sdm_stress_noalloc: the per-block partial tensors and the result, in a
// file-scope device array instead of an allocation.  Each GPU holds its own copy (for the
// life of the process) at its own address; host code never passes the array's name (a host
// stand-in), only the address hipGetSymbolAddress returns on that device.
constexpr int64_t kStressScratchBlocks = 131072;  // interactions: up to 131072 * 256 = 33,554,432
static __device__ __align__(256) double scratch[9 * kStressScratchBlocks + 9];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch[kMaxDevices] = {};

// The device address of `scratch` on device `dev`, looked up once per device with it current.
double* stress_scratch(int dev) {
  if (dev < 0 || dev >= kMaxDevices) throw HipError{"sdm_stress_noalloc: device index out of range"};
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[dev]) {
    int cur = 0;
    HIPCHK(hipGetDevice(&cur));
    if (cur != dev) HIPCHK(hipSetDevice(dev));
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != dev) HIPCHK(hipSetDevice(cur));
    if (e != hipSuccess || !p)
      throw HipError{std::string("sdm_stress_noalloc: hipGetSymbolAddress(scratch): ") + hipGetErrorString(e)};
    g_scratch[dev] = (double*)p;
  }
  return g_scratch[dev];
}
}  // namespace

int sdm_stress_noalloc(sdm_ctx* c, double volume, double sigma[9]) {
  if (!c || !sigma) return fail(c, "sdm_stress_noalloc: bad arguments");
  try {
    const int64_t nb = (int64_t)nblk(std::max<int64_t>(1, c->np));
    if (nb > kStressScratchBlocks)
      return fail(c, "sdm_stress_noalloc: " + std::to_string(c->np) + " interactions exceed the scratch limit of " +
                         std::to_string(kStressScratchBlocks * 256) + "; use sdm_stress");
    double* part = stress_scratch(c->dev);
    if (volume == 0) volume = c->prm.periodic ? c->prm.cell[0] * c->prm.cell[1] * c->prm.cell[2] : 1.0;
    // every block writes its partial and the final kernel its result: nothing to clear
    hipLaunchKernelGGL(k_stress_partial, dim3((unsigned)nb), dim3(256), 0, c->st, c->P, c->np, c->B,
                       c->prm.periodic ? 1 : 0, c->prm.cell[0], c->prm.cell[1], c->prm.cell[2], part);
    hipLaunchKernelGGL(k_stress_final, dim3(1), dim3(64), 0, c->st, part, nb, volume,
                       part + 9 * kStressScratchBlocks);
    HIPCHK(hipMemcpyAsync(sigma, part + 9 * kStressScratchBlocks, 9 * sizeof(double), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
  } catch (HipError& e) {
    return fail(c, e.msg);
  }
  return 0;
}

int sdm_body_forces(sdm_ctx* c, const int32_t* ids, int32_t n, double* forces) {
  if (!c || n < 0 || (n > 0 && (!ids || !forces))) return fail(c, "sdm_body_forces: bad arguments");
  if (n == 0) return 0;
  try {
    for (int32_t i = 0; i < n; i++)
      if (ids[i] < 0 || ids[i] >= c->B.n) return fail(c, "sdm_body_forces: body id out of range");
    int32_t* dids = dalloc<int32_t>(n);
    double* dout = dalloc<double>(3 * (int64_t)n);
    HIPCHK(hipMemcpyAsync(dids, ids, n * sizeof(int32_t), hipMemcpyHostToDevice, c->st));
    hipLaunchKernelGGL(k_gather_forces, dim3(nblk(n)), dim3(256), 0, c->st, c->B, dids, n, dout);
    HIPCHK(hipMemcpyAsync(forces, dout, 3 * (size_t)n * sizeof(double), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    dfree(dids);
    dfree(dout);
  } catch (HipError& e) {
    return fail(c, e.msg);
  }
  return 0;
}

int sdm_set_servo(sdm_ctx* c, const sdm_servo* sv) {
  if (!c) return fail(c, "null context");
  try {
    HIPCHK(hipStreamSynchronize(c->st));
    c->graphGen++;
    if (!sv) {
      c->servoOn = false;
      return 0;
    }
    if (sv->iterate_num_max < 1 || sv->solve_num_max < 1 || sv->echo_interval < 1)
      return fail(c, "sdm_set_servo: intervals must be >= 1");
    if (sv->ramp_interval < 0 || sv->ramp_chunks < 1 || sv->shear_mode < 0)
      return fail(c, "sdm_set_servo: ramp_interval >= 0, ramp_chunks >= 1 and shear_mode >= 0 are needed");
    ServoDev h;
    std::memset(&h, 0, sizeof(h));
    if (sv->keep_state) {  // the running engine's state, new attributes
      if (!c->dServo) return fail(c, "sdm_set_servo: keep_state needs a servo that was set before");
      HIPCHK(hipMemcpy(&h, c->dServo, sizeof(h), hipMemcpyDeviceToHost));
    } else {
      h.first = 1;
      h.completedIter = -1;
      h.unbf = 1.0;  // CompressionEngine::UnbalancedForce default
      if (sv->continue_flag) {
        h.width0 = sv->size0[0];
        h.depth0 = sv->size0[1];
        h.height0 = sv->size0[2];
      }
    }
    for (int k = 0; k < 3; k++) h.goal[k] = sv->goal[k];
    h.max_vel = sv->max_vel;
    h.gain_alpha = sv->gain_alpha;
    h.f_threshold = sv->f_threshold;
    h.fmin_threshold = sv->fmin_threshold;
    h.unbf_tol = sv->unbf_tol;
    h.iterate_num_max = sv->iterate_num_max;
    h.solve_num_max = sv->solve_num_max;
    for (int k = 0; k < 4; k++) h.activated[k] = sv->activated[k];
    h.z_servo = sv->z_servo;
    h.hydro = sv->hydro_strain;
    h.hydroRate = sv->hydro_strain_rate;
    h.shearMode = sv->shear_mode;
    h.rampInterval = sv->ramp_interval;
    h.rampChunks = sv->ramp_chunks;
    h.echoInterval = sv->echo_interval;
    h.targetStrain = sv->target_strain;
    h.cont = sv->continue_flag;
    if (!c->dServo) c->dServo = dalloc<ServoDev>(1);
    HIPCHK(hipMemcpy(c->dServo, &h, sizeof(h), hipMemcpyHostToDevice));
    c->servoOn = true;
    c->servoDone = false;
  } catch (HipError& e) {
    return fail(c, e.msg);
  }
  return 0;
}

int sdm_get_servo(sdm_ctx* c, sdm_servo_state* o) {
  if (!c || !o) return fail(c, "sdm_get_servo: bad arguments");
  if (!c->dServo) return fail(c, "sdm_get_servo: the servo was never set");
  try {
    ServoDev h;
    HIPCHK(hipStreamSynchronize(c->st));
    HIPCHK(hipMemcpy(&h, c->dServo, sizeof(h), hipMemcpyDeviceToHost));
    for (int k = 0; k < 6; k++) {
      o->wall_stress[k] = h.ws[k];
      o->gain[k] = h.gain[k];
      o->stiffness[k] = h.stiff[k];
    }
    o->stress[0] = h.wsxx;
    o->stress[1] = h.wsyy;
    o->stress[2] = h.wszz;
    o->size[0] = h.width;
    o->size[1] = h.depth;
    o->size[2] = h.height;
    o->mean_stress = h.mean;
    o->deviator_stress = h.dev;
    o->unbalanced_force = h.unbf;
    o->iterate_num = h.iterate_num;
    o->solve_num = h.solve_num;
    o->completed_iter = h.completedIter;
    o->size0[0] = h.width0;
    o->size0[1] = h.depth0;
    o->size0[2] = h.height0;
    o->ramp_num = h.rampNum;
    o->pad_ = 0;
  } catch (HipError& e) {
    return fail(c, e.msg);
  }
  return 0;
}

int sdm_get_tuning(sdm_ctx* c, sdm_tuning* t) {
  if (!c || !t) return fail(c, "sdm_get_tuning: bad arguments");
  *t = c->tune;
  return 0;
}

int sdm_get_timers(sdm_ctx* c, double out[8]) {
  if (!c || !out) return -1;
  unsigned long long ev[2] = {0, 0};
  if (hipMemcpy(ev, c->dcnt + 4, sizeof(ev), hipMemcpyDeviceToHost) != hipSuccess) return fail(c, "timers");
  for (int k = 0; k < 5; k++) out[k] = c->tms[k];
  out[5] = (double)ev[0];
  out[6] = (double)ev[1];
  out[7] = c->tms[7];
  return 0;
}

int sdm_get_kernel_timers(sdm_ctx* c, double out[32]) {
  if (!c || !out) return -1;
  for (int k = 0; k < KT_N; k++) {
    out[2 * k] = c->kms[k];
    out[2 * k + 1] = c->kcount[k];
  }
  for (int k = 0; k < 4; k++) out[2 * KT_N + k] = c->kms[KT_N + k];
  return 0;
}

int sdm_reset_counters(sdm_ctx* c) {
  if (!c) return -1;
  std::memset(&c->cnt, 0, sizeof(c->cnt));
  if (hipMemset(c->dcnt, 0, 2 * sizeof(unsigned long long)) != hipSuccess) return fail(c, "reset counters");
  return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------------------
// Frozen-contact rollout and its adjoint (frozen.h)
// ---------------------------------------------------------------------------------------
namespace {

int64_t fz_state(const sdm_frozen_problem* p) { return 9 * (int64_t)p->n_bodies + 3 * (int64_t)p->n_contacts; }
int64_t fz_nck(const sdm_frozen_problem* p) { return (p->steps + p->checkpoint_every - 1) / p->checkpoint_every; }

struct FzState {  // views into one state slot [x | v | w | fs]
  double *x, *v, *w, *fs;
};
FzState fz_slot(const sdm_frozen_problem* p, double* base) {
  const int64_t n3 = 3 * (int64_t)p->n_bodies;
  return FzState{base, base + n3, base + 2 * n3, base + 3 * n3};
}

// work layout: [two forward slots][segment slots (checkpoint_every + 1)][adjoint slot]
//              [fc 3m][lF 3n][lT 3n][contrib 18m][acc 8]
struct FzWork {
  double *fwd[2], *seg, *adj, *fc, *lF, *lT, *contrib, *acc;
};
FzWork fz_work(const sdm_frozen_problem* p, double* w) {
  const int64_t S = fz_state(p), n = p->n_bodies, m = p->n_contacts;
  FzWork W;
  W.fwd[0] = w;
  W.fwd[1] = w + S;
  W.seg = w + 2 * S;
  W.adj = W.seg + (p->checkpoint_every + 1) * S;
  W.fc = W.adj + S;
  W.lF = W.fc + 3 * m;
  W.lT = W.lF + 3 * n;
  W.contrib = W.lT + 3 * n;
  W.acc = W.contrib + 18 * m;
  return W;
}

int fz_check(const sdm_frozen_problem* p) {
  if (!p || p->n_bodies <= 0 || p->n_contacts < 0 || p->steps <= 0 || p->checkpoint_every <= 0 ||
      p->loss_steps <= 0 || p->loss_steps > p->steps)
    return fail(nullptr, "sdm_frozen: invalid problem sizes");
  return 0;
}

__global__ void k_fz_init(sdm_frozen_problem P, const double* v0, const double* w0, double* x, double* v, double* w,
                          double* fs) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < P.n_bodies) {
    for (int k = 0; k < 3; k++) {
      x[3 * i + k] = P.pos0[3 * i + k];
      v[3 * i + k] = P.dynamic[i] ? v0[3 * i + k] : P.wall_vel[3 * i + k];
      w[3 * i + k] = P.dynamic[i] ? w0[3 * i + k] : 0.0;
    }
  }
  if (i < P.n_contacts)
    for (int k = 0; k < 3; k++) fs[3 * i + k] = P.fs0[3 * i + k];
}

// One forward step from slot `a` into slot `b`; the loss is added into *loss.
void fz_step(const sdm_frozen_problem& P, const double* par, FzState a, FzState b, double* fc, int t, double* loss) {
  if (P.n_contacts > 0)
    hipLaunchKernelGGL(fz::k_contacts, dim3(nblk(P.n_contacts)), dim3(256), 0, 0, P, par, a.x, a.v, a.w, a.fs, b.fs,
                       fc, t, loss);
  hipLaunchKernelGGL(fz::k_newton, dim3(nblk(P.n_bodies)), dim3(256), 0, 0, P, a.x, a.v, a.w, fc, b.x, b.v, b.w);
}

}  // namespace

int64_t sdm_frozen_checkpoint_size(const sdm_frozen_problem* p) {
  if (fz_check(p)) return -1;
  return fz_nck(p) * fz_state(p);
}

int64_t sdm_frozen_work_size(const sdm_frozen_problem* p) {
  if (fz_check(p)) return -1;
  return (p->checkpoint_every + 4) * fz_state(p) + 3 * (int64_t)p->n_contacts + 6 * (int64_t)p->n_bodies +
         18 * (int64_t)p->n_contacts + 8;
}

int sdm_frozen_forward(const sdm_frozen_problem* p, const double* params, const double* v0, const double* w0,
                       double* checkpoints, double* work, double* loss) {
  if (fz_check(p) || !params || !v0 || !w0 || !checkpoints || !work || !loss)
    return fail(nullptr, "sdm_frozen_forward: invalid arguments");
  try {
    const sdm_frozen_problem& P = *p;
    FzWork W = fz_work(p, work);
    const int64_t S = fz_state(p);
    HIPCHK(hipMemsetAsync(W.acc, 0, 8 * sizeof(double), 0));
    FzState cur = fz_slot(p, W.fwd[0]), nxt = fz_slot(p, W.fwd[1]);
    const int64_t nmax = std::max<int64_t>(P.n_bodies, P.n_contacts);
    hipLaunchKernelGGL(k_fz_init, dim3(nblk(nmax)), dim3(256), 0, 0, P, v0, w0, cur.x, cur.v, cur.w, cur.fs);
    for (int t = 0; t < P.steps; t++) {
      if (t % P.checkpoint_every == 0)
        HIPCHK(hipMemcpyAsync(checkpoints + (t / P.checkpoint_every) * S, cur.x, S * sizeof(double),
                              hipMemcpyDeviceToDevice, 0));
      fz_step(P, params, cur, nxt, W.fc, t, W.acc);
      std::swap(cur, nxt);
    }
    HIPCHK(hipMemcpy(loss, W.acc, sizeof(double), hipMemcpyDeviceToHost));
  } catch (HipError& e) {
    return fail(nullptr, e.msg);
  }
  return 0;
}

int sdm_frozen_backward(const sdm_frozen_problem* p, const double* params, const double* checkpoints, double* work,
                        double* grad_params, double* grad_v0) {
  if (fz_check(p) || !params || !checkpoints || !work || !grad_params || !grad_v0)
    return fail(nullptr, "sdm_frozen_backward: invalid arguments");
  try {
    const sdm_frozen_problem& P = *p;
    FzWork W = fz_work(p, work);
    const int64_t S = fz_state(p);
    HIPCHK(hipMemsetAsync(W.adj, 0, S * sizeof(double), 0));
    HIPCHK(hipMemsetAsync(W.acc, 0, 8 * sizeof(double), 0));
    FzState L = fz_slot(p, W.adj);
    double* dummyLoss = W.acc + 4;
    double* gpar = W.acc + 1;
    const int ck = P.checkpoint_every;
    for (int64_t seg = fz_nck(p) - 1; seg >= 0; seg--) {
      const int t0 = (int)(seg * ck), t1 = std::min(t0 + ck, P.steps);
      // recompute the segment's states from its checkpoint
      HIPCHK(hipMemcpyAsync(W.seg, checkpoints + seg * S, S * sizeof(double), hipMemcpyDeviceToDevice, 0));
      for (int t = t0; t < t1; t++)
        fz_step(P, params, fz_slot(p, W.seg + (t - t0) * S), fz_slot(p, W.seg + (t - t0 + 1) * S), W.fc, t, dummyLoss);
      for (int t = t1 - 1; t >= t0; t--) {
        const FzState s = fz_slot(p, W.seg + (t - t0) * S);
        const FzState s1 = fz_slot(p, W.seg + (t - t0 + 1) * S);
        // the step's contact forces (the shear output goes to a slot that is not read again)
        if (P.n_contacts > 0)
          hipLaunchKernelGGL(fz::k_contacts, dim3(nblk(P.n_contacts)), dim3(256), 0, 0, P, params, s.x, s.v, s.w,
                             s.fs, s1.fs, W.fc, t, dummyLoss);
        hipLaunchKernelGGL(fz::k_back_newton, dim3(nblk(P.n_bodies)), dim3(256), 0, 0, P, s.x, s.v, s.w, W.fc, L.x,
                           L.v, L.w, W.lF, W.lT);
        if (P.n_contacts > 0) {
          hipLaunchKernelGGL(fz::k_back_contacts, dim3(nblk(P.n_contacts)), dim3(256), 0, 0, P, params, s.x, s.v,
                             s.w, s.fs, W.lF, W.lT, L.fs, W.contrib, t, gpar);
          hipLaunchKernelGGL(fz::k_back_gather, dim3(nblk(P.n_bodies)), dim3(256), 0, 0, P, W.contrib, L.x, L.v, L.w);
        }
      }
    }
    HIPCHK(hipMemcpyAsync(grad_v0, L.v, 3 * (size_t)P.n_bodies * sizeof(double), hipMemcpyDeviceToDevice, 0));
    HIPCHK(hipMemcpy(grad_params, gpar, 3 * sizeof(double), hipMemcpyDeviceToHost));
  } catch (HipError& e) {
    return fail(nullptr, e.msg);
  }
  return 0;
}
