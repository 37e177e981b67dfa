// pairs.h -- the interaction loop: Ig2 functor, Ip2 functor and contact law per interaction.
//
// Device restatement of InteractionLoop::action (pkg/common/InteractionLoop.cpp:23-149) with
//   Ig2_Superquadrics_Superquadrics_SuperquadricsGeom::go   (Superquadrics_Ig2.cpp:1880-2317)
//   Ig2_Superquadrics_Superquadrics_SuperquadricsGeom2::go  (Superquadrics_Ig2.cpp:2322-2517)
//   Ig2_Wall_Superquadrics_SuperquadricsGeom(2)::go         (Superquadrics_Ig2.cpp:2522-2703)
//   SuperquadricsGeom::precompute / rotate                 (Superquadrics.cpp:1030-1080)
//   Ip2_SuperquadricsMat(2)_* ::go                         (Superquadrics.cpp:1086-1160)
//   SuperquadricsLaw::go / SuperquadricsLaw2::go           (Superquadrics.cpp:1181-1557)
// split in three kernels so that none carries the register footprint of all of them:
//   k_pairs_prep   every interaction: functor dispatch, wall / sphere / bounding-sphere
//                  cases completely; pairs that need an iterative solve are queued;
//   (solve.h)      persistent solvers over the queue;
//   k_pairs_finish queued pairs: solver outputs -> geometry, curvature and Hertz terms,
//                  precompute, Ip2, law.
// Each contact's force and torques are stored on the interaction; the per-body sums are
// formed by the integrator kernel in interaction order, which is the order of the
// reference's serial addForce/addTorque calls (core/ForceContainer.hpp:218-319).
#pragma once
#include "solve.h"
#include "state.h"

enum { SDM_KIND_SQ = 0, SDM_KIND_WALL = 1 };

struct StepArgs {
  BodyArrays B;
  PairArrays P;
  const MaterialDev* mats;
  int64_t np;
  int64_t iter;
  double dt;
  unsigned long long* counters;  // [0] contact evals, [1] distance evaluations
  int fullWrite;                 // store every value (first step after host writes)
  int32_t* csr;                  // per-body interaction lists (entries carry the active bit)
  int periodic;                  // Scene::isPeriodic
  double cell[3];                // cell size (axis-aligned cell: hSize = diag(cell))
  int skip;                      // settled apart pairs may be skipped (P.gap, B.disp are current)
  // The container size and Scene::iter on the device (set before each step), so that a
  // captured step stays valid while they change; np above is then the launch bound.
  const unsigned long long* npDev;
  const unsigned long long* iterDev;
};

__device__ __forceinline__ int64_t step_np(const StepArgs& A) {
  const int64_t n = (int64_t)*A.npDev;
  return n < A.np ? n : A.np;
}

// shift2 = hSize * cellDist (InteractionLoop.cpp:100-105): zero outside a periodic cell.
__device__ __forceinline__ V3<double> pair_shift(const StepArgs& A, int64_t p) {
  if (!A.periodic) return mk3(0.0, 0.0, 0.0);
  const PairRec& R = A.P.r[p];
  return mk3(A.cell[0] * (double)R.cellDist[0], A.cell[1] * (double)R.cellDist[1], A.cell[2] * (double)R.cellDist[2]);
}

__device__ __forceinline__ int32_t csr_entry(int64_t p, int side, bool active) {
  return (int32_t)((p << 2) | (side << 1) | (active ? 1 : 0));
}

// Store v unless `lazy` and the stored value is bitwise equal.
__device__ __forceinline__ void put(int32_t& d, int32_t v, bool lazy) {
  if (!lazy || d != v) d = v;
}
__device__ __forceinline__ void put(double& d, double v, bool lazy) {
  if (!lazy || __double_as_longlong(d) != __double_as_longlong(v)) d = v;
}

// Three consecutive values of a record field.
__device__ __forceinline__ V3<double> ldv3(const double* a) { return mk3(a[0], a[1], a[2]); }
__device__ __forceinline__ void stv3(double* a, V3<double> v) {
  a[0] = v.x;
  a[1] = v.y;
  a[2] = v.z;
}
__device__ __forceinline__ V3<double> ld3(const double* a, int64_t i, int64_t s) {
  return mk3(a[i], a[s + i], a[2 * s + i]);
}
__device__ __forceinline__ void st3(double* a, int64_t i, int64_t s, V3<double> v) {
  a[i] = v.x;
  a[s + i] = v.y;
  a[2 * s + i] = v.z;
}
__device__ __forceinline__ M3<double> ld9(const double* a, int64_t i, int64_t s) {
  M3<double> m;
#pragma unroll
  for (int k = 0; k < 9; k++) m.m[k] = a[k * s + i];
  return m;
}

// Geometry of one contact as the functors leave it; fields follow SuperquadricsGeom(2).
struct GeomState {
  double depth, a0, a1;
  V3<double> normal, cpt, shearInc, relVn, relVs, twist, orth, point1, point2;
  bool isShearNew;
  double c10, c11, c20, c21, curvAngle, alpha, K, E, curvSum;
  bool gSphere;
  V3<double> shift2;  // periodic shift of body 2 (precompute, SuperquadricsLaw's box test)
};

struct BodyView {
  V3<double> pos, vel, angVel, bmin, bmax;
  double mass;
  int blocked, material, kind, isSphere, wallAxis, wallSense;
  bool low6;  // one of the first six bodies of its specimen (the reference's ids 0-5)
  M3<double> G;  // rot_mat2global
  Shape<double> s;
  double rmax;
};

__device__ __forceinline__ void load_bounds(const double* geo, int id, V3<double>& mn, V3<double>& mx) {
  const double* g = geo + (int64_t)id * GEO_STRIDE + GEO_BMIN;
  const double2* r = reinterpret_cast<const double2*>(g);
  const double2 a = r[0], b = r[1], c = r[2];
  mn = mk3(a.x, a.y, b.x);
  mx = mk3(b.y, c.x, c.y);
}

__device__ __forceinline__ BodyView load_body(const BodyArrays& B, int id) {
  BodyView b;
  const double2* r = reinterpret_cast<const double2*>(B.geo + (int64_t)id * GEO_STRIDE);
  double v[GEO_STRIDE];
#pragma unroll
  for (int k = 0; k < GEO_STRIDE / 2; k++) {
    const double2 t = r[k];
    v[2 * k] = t.x;
    v[2 * k + 1] = t.y;
  }
  b.pos = mk3(v[GEO_POS], v[GEO_POS + 1], v[GEO_POS + 2]);
  b.s.rx = v[GEO_R]; b.s.ry = v[GEO_R + 1]; b.s.rz = v[GEO_R + 2];
  b.s.eps1 = v[GEO_E]; b.s.eps2 = v[GEO_E + 1];
  b.rmax = v[GEO_RMAX];
#pragma unroll
  for (int k = 0; k < 9; k++) b.G.m[k] = v[GEO_G + k];
  b.bmin = mk3(v[GEO_BMIN], v[GEO_BMIN + 1], v[GEO_BMIN + 2]);
  b.bmax = mk3(v[GEO_BMAX], v[GEO_BMAX + 1], v[GEO_BMAX + 2]);
  b.vel = mk3(v[GEO_VEL], v[GEO_VEL + 1], v[GEO_VEL + 2]);
  b.angVel = mk3(v[GEO_ANGVEL], v[GEO_ANGVEL + 1], v[GEO_ANGVEL + 2]);
  b.mass = v[GEO_MASS];
  const uint64_t fl = (uint64_t)__double_as_longlong(v[GEO_FLAGS]);
  b.material = (int)(uint32_t)fl;
  b.blocked = (int)((fl >> 32) & 0xFF);
  b.kind = (int)((fl >> 40) & 3);
  b.isSphere = (int)((fl >> 42) & 1);
  b.wallAxis = (int)((fl >> 43) & 3);
  b.wallSense = (int)((fl >> 45) & 3) - 1;
  b.low6 = ((fl >> 47) & 1) != 0;
  return b;
}

// Previous geometry of an interaction (a fresh SuperquadricsGeom(2) is all zero, isShearNew).
__device__ __forceinline__ void load_geom_prev(const PairArrays& P, int64_t p, bool hadGeom, GeomState& g) {
  const int64_t S = P.cap;
  g.normal = mk3(0.0, 0.0, 0.0);
  g.cpt = g.normal; g.shearInc = g.normal; g.relVn = g.normal; g.relVs = g.normal;
  g.twist = g.normal; g.orth = g.normal; g.point1 = g.normal; g.point2 = g.normal;
  g.c10 = g.c11 = g.c20 = g.c21 = 0; g.curvAngle = g.alpha = g.K = g.E = g.curvSum = 0;
  if (!hadGeom) {
    g.depth = 0; g.a0 = 0; g.a1 = 0;
    g.isShearNew = true;
    g.gSphere = false;
  } else {
    g.depth = P.r[p].depth[0];
    g.a0 = P.r[p].ang[0];
    g.a1 = P.r[p].ang[1];
    g.normal = ldv3(P.r[p].normal);
    g.gSphere = P.r[p].gSphere[0] != 0;
    g.isShearNew = g.depth <= 0;  // Superquadrics_Ig2.cpp:1931, 2344, 2537
  }
}

// SuperquadricsGeom::precompute (Superquadrics.cpp:1030-1050); g.normal is the previous normal.
__device__ __forceinline__ void precompute(GeomState& g, const BodyView& b1, const BodyView& b2, double dt,
                                           V3<double> currentNormal, bool isNew) {
  if (!isNew) {
    g.orth = cross(g.normal, currentNormal);
    double angle = dt * 0.5 * dot(currentNormal, b1.angVel + b2.angVel);
    g.twist = angle * currentNormal;
  } else {
    g.twist = mk3(0.0, 0.0, 0.0);
    g.orth = g.twist;
  }
  g.normal = currentNormal;
  V3<double> c1x = g.cpt - b1.pos;
  V3<double> c2x = (g.cpt - b2.pos) - g.shift2;
  V3<double> rv = (b2.vel + cross(b2.angVel, c2x)) - (b1.vel + cross(b1.angVel, c1x));
  double vn = dot(g.normal, rv);
  g.relVn = vn * g.normal;
  g.relVs = rv - g.relVn;
  g.shearInc = g.relVs * dt;
}

// Outcome of a functor: EARLY = only depth/isShearNew changed; EARLY_ANG also the angles.
enum { GEO_EARLY = 0, GEO_EARLY_ANG = 1, GEO_FULL = 2 };

struct PhysState {
  double kn, ks, tanPhi, betan, betas, nuab, Gab, mu1, mu2;
  V3<double> fn, fs;
};

// SuperquadricsGeom::rotate (Superquadrics.cpp:1053-1080).
__device__ __forceinline__ V3<double> rotate_shear(const GeomState& g, V3<double> fs) {
  fs = fs - cross(fs, g.orth);
  fs = fs - cross(fs, g.twist);
  fs = fs - dot(g.normal, fs) * g.normal;
  return fs;
}

// tan(min(phi1, phi2)) of Ip2 (Superquadrics.cpp:1101-1102, 1140-1141): the tangent of the material whose
// angle std::min picks (the first when equal), as the host's C library evaluates it.
__device__ __forceinline__ double tan_min_friction(const MaterialDev& m1, const MaterialDev& m2) {
  return m2.frictionAngle < m1.frictionAngle ? m2.tanF : m1.tanF;
}

// Ip2 (first time only), the law, and the write-back of one interaction.
template <int LAW>
__device__ bool phys_law_store(const StepArgs& A, int64_t p, const BodyView& b1, const BodyView& b2, int id1,
                               GeomState& g, int outcome, bool hadGeom, bool hadPhys, int nev, int stopCode) {
  const PairArrays& P = A.P;
  const int64_t S = P.cap;
  // ---------------- Ip2 functor ----------------
  PhysState ph;
  if (hadPhys && outcome != GEO_FULL) {
    // no contact this step: the law reads nothing of the phys but the erase test
    ph.fs = mk3(0.0, 0.0, 0.0);
    ph.kn = ph.ks = 0;
  } else if (hadPhys) {
    ph.kn = P.r[p].kn[0]; ph.ks = P.r[p].ks[0]; ph.tanPhi = P.r[p].tanPhi[0];
    ph.betan = P.r[p].betan[0]; ph.betas = P.r[p].betas[0];
    ph.nuab = P.r[p].nuab[0]; ph.Gab = P.r[p].Gab[0]; ph.mu1 = P.r[p].mu1[0]; ph.mu2 = P.r[p].mu2[0];
    ph.fs = ldv3(P.r[p].fs);
  } else {
    const MaterialDev m1 = A.mats[b1.material];
    const MaterialDev m2 = A.mats[b2.material];
    ph.kn = 0; ph.ks = 0; ph.tanPhi = 0; ph.nuab = 0; ph.Gab = 0; ph.mu1 = 0; ph.mu2 = 0;
    ph.fs = mk3(0.0, 0.0, 0.0);
    // "id1 > 5" of Ip2 (Superquadrics.cpp:1100, 1134) and Law2 (:1394): id1 is not one of the
    // first six bodies -- of its specimen when several specimens share the context
    if (LAW == 0) {
      if (!b1.low6) {
        ph.tanPhi = tan_min_friction(m1, m2);
        ph.kn = 2.0 * m1.Kn * m2.Kn / (m1.Kn + m2.Kn);
        ph.ks = 2.0 * m1.Ks * m2.Ks / (m1.Ks + m2.Ks);
      } else {
        ph.tanPhi = m1.tanF;
        ph.kn = m1.Kn;
        ph.ks = m1.Ks;
      }
    } else {
      if (!b1.low6) {
        ph.nuab = 0.5 * (m1.nu + m2.nu);
        ph.Gab = 1.0 / ((1 - m1.nu) / m1.G + (1 - m2.nu) / m2.G);
        ph.mu1 = 1. / m1.G + 1. / m2.G;
        ph.mu2 = m1.nu / m1.G + m2.nu / m2.G;
        ph.tanPhi = tan_min_friction(m1, m2);
      } else {
        ph.tanPhi = m1.tanF;
        ph.kn = m1.Kn;
        ph.ks = m1.Ks;
      }
    }
    ph.betan = fmax(m1.betan, m2.betan);
    ph.betas = fmax(m1.betas, m2.betas);
  }

  // ---------------- law functor ----------------
  bool keep = true;
  bool contributed = false;
  V3<double> F = mk3(0.0, 0.0, 0.0);
  // SuperquadricsLaw shifts B's box by shift2 in a periodic cell (Superquadrics.cpp:1196-1201);
  // SuperquadricsLaw2 does not (:1381)
  V3<double> bmn2 = b2.bmin, bmx2 = b2.bmax;
  if (LAW == 0 && A.periodic) {
    bmn2 = bmn2 + g.shift2;
    bmx2 = bmx2 + g.shift2;
  }
  if (b1.bmin.x >= bmx2.x || bmn2.x >= b1.bmax.x || b1.bmin.y >= bmx2.y || bmn2.y >= b1.bmax.y ||
      b1.bmin.z >= bmx2.z || bmn2.z >= b1.bmax.z) {
    keep = false;  // requestErase: the interaction is reset below
  } else if (!(g.depth > 1E-18)) {
    ph.fn = mk3(0.0, 0.0, 0.0);
    ph.fs = ph.fn;
  } else {
    if (LAW == 1 && !b1.low6) {
      const double depth = g.depth, cs = g.curvSum;
      if (g.gSphere) {
        double tmp = sqrt(depth / cs);
        ph.kn = 8. / 3. * ph.Gab * tmp;
        ph.ks = 8. / (2.0 * ph.mu1 - ph.mu2) * tmp;
      } else {
        // The reference forms these two in long double through M_PIl (Superquadrics.cpp:1413,
        // 1417): each product and the quotient round to a 64-bit significand (emulated on
        // double-doubles), the assignment to double.
        double a = sqrt(depth * g.E / cs / g.K);
        crm::DD t = crm::ext_mul_d(crm::ext_pi(), 4. / 3.);
        t = crm::ext_mul_d(crm::ext_mul_d(crm::ext_mul_d(t, a), ph.Gab), g.alpha);
        ph.kn = crm::ext_to_double(crm::ext_div_d(t, g.K));
        double e2 = 1 - 1 / (g.alpha * g.alpha);
        double ksx = ph.mu1 * g.K - ph.mu2 * (g.K - g.E) / e2;
        double ksy = ph.mu1 * g.K + ph.mu2 * ((1 - e2) * g.K - g.E) / e2;
        const double sinv = 1. / ksx + 1. / ksy;
        ph.ks = crm::ext_to_double(crm::ext_mul_d(crm::ext_mul_d(crm::ext_mul_d(crm::ext_pi(), g.alpha), a), sinv));
      }
    }
    V3<double> fn = mk3(g.normal.x * g.depth * ph.kn, g.normal.y * g.depth * ph.kn, g.normal.z * g.depth * ph.kn);
    V3<double> fs1 = mk3(0.0, 0.0, 0.0);
    const bool useDamping = (ph.betan > 1e-5 || ph.betas > 1e-5);
    V3<double> shearViscous = mk3(0.0, 0.0, 0.0);
    double csd = 0;
    if (useDamping) {
      const bool Adyn = b1.blocked != 63, Bdyn = b2.blocked != 63;
      double mbar = (!Adyn && Bdyn) ? b2.mass : ((!Bdyn && Adyn) ? b1.mass : (b1.mass * b2.mass / (b1.mass + b2.mass)));
      double Cn_crit = 2. * sqrt(mbar * ph.kn);
      double Cs_crit = 2. * sqrt(mbar * ph.ks);
      double cn = Cn_crit * ph.betan;
      csd = Cs_crit * ph.betas;
      V3<double> nv = cn * g.relVn;
      if (dot(fn - nv, g.normal) < 0.0) {
        fn = mk3(0.0, 0.0, 0.0);
      } else {
        fn = fn - nv;
      }
    }
    if (!g.isShearNew) fs1 = rotate_shear(g, ph.fs);
    fs1 = fs1 - ph.ks * g.shearInc;
    const double maxFs = sqn(fn) * (ph.tanPhi * ph.tanPhi);
    if (sqn(fs1) > maxFs) {
      double ratio = sqrt(maxFs) / norm(fs1);
      fs1 = fs1 * ratio;
    } else if (useDamping) {
      shearViscous = csd * g.relVs;
      if (LAW == 1) fs1 = fs1 - shearViscous;
    }
    if (isnan(fs1.x) || isnan(fn.x)) {
      fn = mk3(0.0, 0.0, 0.0);
      fs1 = fn;
    }
    if (LAW == 1)
      F = mk3(-fn.x - fs1.x, -fn.y - fs1.y, -fn.z - fs1.z);
    else
      F = mk3(-fn.x - fs1.x + shearViscous.x, -fn.y - fs1.y + shearViscous.y, -fn.z - fs1.z + shearViscous.z);
    ph.fn = fn;
    ph.fs = fs1;
    contributed = true;
  }

  // ---------------- write back ----------------
  // Outside the first step after the host wrote interaction data (A.fullWrite), a value is
  // only stored when it differs from the stored one: an interaction that stays apart leaves
  // its record untouched.
  PairRec& R = P.r[p];
  const bool lazy = !A.fullWrite;
  const double prevDepth = R.depth[0];
  put(R.nEval[0], nev, lazy);
  put(R.stop[0], stopCode, lazy);
  put(R.active[0], contributed ? 1 : 0, lazy);
  if (!keep) {
    // Interaction::reset (core/Interaction.cpp:25-29): drop geom and phys; store the defaults
    // a fresh geom/phys starts from.
    P.r[p].hasGeom[0] = 0;
    P.r[p].hasPhys[0] = 0;
    P.r[p].geomExists[0] = 1;
    P.r[p].iterMadeReal[0] = -1;
    const V3<double> z = mk3(0.0, 0.0, 0.0);
    P.r[p].depth[0] = 0; P.r[p].ang[0] = 0; P.r[p].ang[1] = 0;
    stv3(P.r[p].cpt, z); stv3(P.r[p].shearInc, z); stv3(P.r[p].relVn, z); stv3(P.r[p].relVs, z);
    stv3(P.r[p].normal, z); stv3(P.r[p].twist, z); stv3(P.r[p].orth, z); stv3(P.r[p].point1, z);
    stv3(P.r[p].point2, z);
    P.r[p].isShearNew[0] = 1;
    P.r[p].curv1[0] = 0; P.r[p].curv1[1] = 0; P.r[p].curv2[0] = 0; P.r[p].curv2[1] = 0;
    P.r[p].curvAngle[0] = 0; P.r[p].alpha[0] = 0; P.r[p].hK[0] = 0; P.r[p].hE[0] = 0; P.r[p].curvSum[0] = 0; P.r[p].gSphere[0] = 0;
    P.r[p].kn[0] = 0; P.r[p].ks[0] = 0; P.r[p].tanPhi[0] = __builtin_nan(""); P.r[p].betan[0] = 0; P.r[p].betas[0] = 0;
    stv3(P.r[p].fn, z); stv3(P.r[p].fs, z);
    P.r[p].nuab[0] = 0; P.r[p].Gab[0] = 0; P.r[p].mu1[0] = 0; P.r[p].mu2[0] = 0;
    return false;
  }
  if (!(hadGeom && hadPhys)) P.r[p].iterMadeReal[0] = (int64_t)*A.iterDev;
  put(R.hasGeom[0], 1, lazy);
  put(R.hasPhys[0], 1, lazy);
  put(R.depth[0], g.depth, lazy);
  put(R.isShearNew[0], g.isShearNew ? 1 : 0, lazy);
  if (outcome != GEO_EARLY) {
    put(R.ang[0], g.a0, lazy);
    put(R.ang[1], g.a1, lazy);
  }
  if (outcome == GEO_FULL) {
    stv3(P.r[p].normal, g.normal); stv3(P.r[p].cpt, g.cpt); stv3(P.r[p].shearInc, g.shearInc);
    stv3(P.r[p].relVn, g.relVn); stv3(P.r[p].relVs, g.relVs); stv3(P.r[p].twist, g.twist);
    stv3(P.r[p].orth, g.orth); stv3(P.r[p].point1, g.point1); stv3(P.r[p].point2, g.point2);
    if (LAW == 1 && b1.kind != SDM_KIND_WALL) {
      P.r[p].curvSum[0] = g.curvSum;
      P.r[p].gSphere[0] = g.gSphere ? 1 : 0;
      if (!g.gSphere) {
        P.r[p].curv1[0] = g.c10; P.r[p].curv1[1] = g.c11; P.r[p].curv2[0] = g.c20; P.r[p].curv2[1] = g.c21;
        P.r[p].curvAngle[0] = g.curvAngle; P.r[p].alpha[0] = g.alpha; P.r[p].hK[0] = g.K; P.r[p].hE[0] = g.E;
      }
    }
  }
  if (!hadPhys) {
    P.r[p].tanPhi[0] = ph.tanPhi; P.r[p].betan[0] = ph.betan; P.r[p].betas[0] = ph.betas;
    P.r[p].nuab[0] = ph.nuab; P.r[p].Gab[0] = ph.Gab; P.r[p].mu1[0] = ph.mu1; P.r[p].mu2[0] = ph.mu2;
  }
  if (!hadPhys || contributed) {  // kn/ks change only in Ip2 and in Law2's contact branch
    P.r[p].kn[0] = ph.kn;
    P.r[p].ks[0] = ph.ks;
  }
  // A kept interaction without force stores zero forces (:1217-1220); after such a step
  // (stored depth <= 1e-18, phys present) they are zero already.
  if (contributed || !lazy || !hadPhys || prevDepth > 1E-18) {
    stv3(R.fn, ph.fn);
    stv3(R.fs, ph.fs);
  }
  if (contributed) {
    A.csr[P.cpos[2 * p]] = csr_entry(p, 0, true);
    A.csr[P.cpos[2 * p + 1]] = csr_entry(p, 1, true);
    stv3(P.r[p].cF, F);
    stv3(P.r[p].cT1, -cross(b1.pos - g.cpt, F));
    stv3(P.r[p].cT2, cross(b2.pos - g.cpt, F));
  }
  return true;
}

// Add the block's counters into the global ones (one atomic per block).
__device__ __forceinline__ void add_counters(unsigned long long* c, unsigned long long ce, unsigned long long de) {
  block_add2(c, ce, de);
}

// Outcome of a superquadric pair that needs no further evaluation: depth 0 and a new
// shear state; EARLY leaves the stored angles, EARLY_ANG stores (a0, a1).  Only the bounds
// (the law's erase test) and, for a first contact, the materials are read.
template <int LAW>
__device__ __forceinline__ bool early_store(const StepArgs& A, int64_t p, int id1, int id2, bool hadGeom,
                                            bool hadPhys, double a0, double a1, int outcome, int nev, int stopCode) {
  const BodyArrays& B = A.B;
  BodyView b1, b2;
  load_bounds(B.geo, id1, b1.bmin, b1.bmax);
  load_bounds(B.geo, id2, b2.bmin, b2.bmax);
  b1.kind = b2.kind = SDM_KIND_SQ;
  b1.mass = b2.mass = 0;
  b1.blocked = b2.blocked = 0;
  b1.material = hadPhys ? 0 : B.material[id1];
  b2.material = hadPhys ? 0 : B.material[id2];
  b1.low6 = b2.low6 = false;
  if (!hadPhys) {
    const uint64_t f1 = (uint64_t)__double_as_longlong(B.geo[(int64_t)id1 * GEO_STRIDE + GEO_FLAGS]);
    b1.low6 = ((f1 >> 47) & 1) != 0;
  }
  GeomState g;
  g.depth = 0;
  g.isShearNew = true;
  g.a0 = a0;
  g.a1 = a1;
  g.gSphere = false;
  g.shift2 = pair_shift(A, p);
  return phys_law_store<LAW>(A, p, b1, b2, id1, g, outcome, hadGeom, hadPhys, nev, stopCode);
}

// Every interaction: functor dispatch; superquadric pairs that the bounding spheres or the
// box bound settle are completed here, the others queued for the solver; wall and sphere
// pairs are listed for k_pairs_misc.
template <int LAW>
__global__ __launch_bounds__(256) void k_pairs_prep(StepArgs A, uint32_t* qflag, int32_t* mlist,
                                                    unsigned long long* mcount) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long ce = 0;
  bool enqueue = false, misc = false;
  int id1 = 0, id2 = 0;
  const PairArrays& P = A.P;
  const BodyArrays& B = A.B;
  // A settled apart pair: its last evaluation ended at the bounding-sphere test with depth 0
  // and kept boxes, so its record holds exactly what that outcome stores, and P.gap[p] is
  // its separation |x1 - x2| - (r_max1 + r_max2) minus both bodies' displacements since the
  // last collider run at that step.  While gap exceeds the displacements now, the spheres
  // are still apart (triangle inequality), the boxes are unchanged (they move only at
  // collider runs), and the functors would store nothing: the interaction is skipped
  // without reading its record or its bodies.
  bool skipped = false;
  const int64_t np = step_np(A);
  if (A.skip && p < np) {
    const float gp = P.gap[p];
    if (gp > 0.0f) {
      const float da = B.disp[P.cid[2 * p]], db = B.disp[P.cid[2 * p + 1]];
      skipped = (double)gp - ((double)da + (double)db) > 1e-13;
    }
  }
  float gapv = __int_as_float(0x7fc00000);
  if (skipped) {
    ce++;  // the reference's functors ran on it
  } else if (p < np) {
    if (P.r[p].active[0]) {  // ForceResetter: last step's contribution is withdrawn
      P.r[p].active[0] = 0;
      A.csr[P.cpos[2 * p]] = csr_entry(p, 0, false);
      A.csr[P.cpos[2 * p + 1]] = csr_entry(p, 1, false);
    }
    if (P.r[p].geomExists[0]) {
      id1 = P.r[p].id1[0];
      id2 = P.r[p].id2[0];
      bool skip = false;
      if (!P.r[p].functorSet[0]) {
        // dispatch order (Wall,Superquadrics) / (Superquadrics,Superquadrics); swapOrder otherwise
        const int k1 = B.kind[id1], k2 = B.kind[id2];
        if (k1 == SDM_KIND_WALL && k2 == SDM_KIND_WALL) {
          P.r[p].geomExists[0] = 0;
          skip = true;
        } else {
          P.r[p].functorSet[0] = 1;
          if (k1 == SDM_KIND_SQ && k2 == SDM_KIND_WALL) {
            int t = id1; id1 = id2; id2 = t;
            P.r[p].id1[0] = id1;
            P.r[p].id2[0] = id2;
            P.cid[2 * p] = id1;
            P.cid[2 * p + 1] = id2;
            // the list entry of the new id1 is the one built for the old id2: the law marks
            // entry cpos[2p] as side 0 (force +F) and cpos[2p + 1] as side 1 (-F)
            const int32_t e = P.cpos[2 * p];
            P.cpos[2 * p] = P.cpos[2 * p + 1];
            P.cpos[2 * p + 1] = e;
          }
        }
      }
      if (!skip) {
        const bool wallPair = B.kind[id1] == SDM_KIND_WALL;
        const bool spherePair = !wallPair && B.isSphere[id1] && B.isSphere[id2];
        if (wallPair || spherePair) {
          misc = true;
        } else {
          // Superquadric pair: bounding spheres, then the separation bound of the solver's
          // first evaluation.  Only the packed records are read on this path.
          const bool hadGeom = P.r[p].hasGeom[0] != 0;
          const bool hadPhys = P.r[p].hasPhys[0] != 0;
          const bool isNew = !hadGeom;
          ce++;
          const GeoRec r1 = load_rec(B.geo, id1), r2 = load_rec(B.geo, id2);
          // Geom (LM) works on position2 = pos + shift2 (Superquadrics_Ig2.cpp:1886), Geom2
          // (NM) on the unshifted position (:2328)
          const V3<double> x1 = r1.pos, x2 = LAW == 0 ? r2.pos + pair_shift(A, p) : r2.pos;
          const double rmax1 = r1.rmax, rmax2 = r2.rmax;
          double para0, para1;
          if (LAW == 0 && isNew) {
            const V3<double> d0 = x2 - x1;
            para0 = xatan2(d0.y, d0.x);
            para1 = xatan(d0.z / sqrt(d0.x * d0.x + d0.y * d0.y));
            para0 += 1e-5;
            para1 += 1e-5;
          } else if (isNew) {
            para0 = 0;
            para1 = 0;
          } else {
            para0 = P.r[p].ang[0];
            para1 = P.r[p].ang[1];
          }
          int outcome = -1, stopCode = 0;
          const double dist = norm(x1 - x2);
          if (dist > rmax1 + rmax2) {
            outcome = GEO_EARLY;  // Superquadrics_Ig2.cpp:1975-1979, 2384-2388
          } else {
            // The first distance evaluation of either solver is at `para` along the contact
            // direction c; it reports "not touching" when c.(p2 - p1) > 0, and
            //   c.(p2 - p1) = c.(x2 - x1) - h1(c) - h2(-c)
            // with h the support function of each body.  Every point of a superellipsoid lies
            // in its box [-rx,rx]x[-ry,ry]x[-rz,rz], so h(n) <= rx|nx| + ry|ny| + rz|nz| in body
            // axes.  When the box bound already proves c.(p2 - p1) > 0 (with a margin far above
            // the rounding of the evaluation), the solver would stop after that evaluation with
            // the parameters unchanged: LM at stop 9 (lmder2 :1145-1149), NM after one function
            // call (neldermead_simplex :1731-1737).  The outcome is then written directly.
            // (c from the device's sin/cos: a last-bit difference moves the gap by ~1e-16 of the
            // scale, far inside the margin, and the solver route gives the same outcome)
            double sp0, cp0, sp1, cp1;
            sincos(para0, &sp0, &cp0);
            sincos(para1, &sp1, &cp1);
            V3<double> c;
            if (LAW == 0) {
              c = mk3(cp0 * cp1, sp0 * cp1, sp1);
            } else {
              const M3<double> gC = crotation_mat(r1.pos, r2.pos);
              c = mul(gC, mk3(cp0 * cp1, sp0 * cp1, sp1));
            }
            const V3<double> n1 = mulT(r1.G, c), n2 = mulT(r2.G, c);
            const double hb1 = r1.s.rx * fabs(n1.x) + r1.s.ry * fabs(n1.y) + r1.s.rz * fabs(n1.z);
            const double hb2 = r2.s.rx * fabs(n2.x) + r2.s.ry * fabs(n2.y) + r2.s.rz * fabs(n2.z);
            const V3<double> dx = x2 - x1;
            const double gap = dot(c, dx) - hb1 - hb2;
            if (gap > 1e-12 * (1.0 + norm(x1) + norm(x2) + hb1 + hb2)) {
              outcome = GEO_EARLY_ANG;
              stopCode = (LAW == 0) ? 9 : 1;
            }
          }
          if (outcome >= 0) {
            // nEval: the reference's count (the one evaluation the box bound stands in for)
            const bool keep = early_store<LAW>(A, p, id1, id2, hadGeom, hadPhys, para0, para1, outcome,
                                               outcome == GEO_EARLY_ANG ? 1 : 0, stopCode);
            if (outcome == GEO_EARLY && keep)
              gapv = __double2float_rd((dist - (rmax1 + rmax2)) - ((double)B.disp[id1] + (double)B.disp[id2]));
          } else {
            enqueue = true;
          }
        }
      }
    }
  }
  if (p < A.np) {
    qflag[p] = enqueue ? 1u : 0u;
    if (!skipped) P.gap[p] = gapv;
  }
  add_counters(A.counters, ce, 0);
  const int64_t ms = fetch_work(misc, mcount);  // wall / sphere pairs are few: rare atomics
  if (misc) mlist[ms] = (int32_t)p;
}

// The solver queue in interaction order: slot qpos[p] for every flagged interaction p, with
// the start parameters of the Ig2 functor (new LM pair: the line of centres plus 1e-5,
// Superquadrics_Ig2.cpp:1919-1924; new NM pair: (0, 0) in its CRotationMat frame, :2355;
// otherwise the stored contactAngle) and the NM simplex size (0.1 new, 0.001 otherwise).
template <int LAW>
__global__ __launch_bounds__(256) void k_queue_fill(StepArgs A, SolveIO S, const uint32_t* qflag,
                                                    const uint32_t* qpos, int64_t np) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np || p >= step_np(A) || !qflag[p]) return;
  const int64_t slot = qpos[p];
  const PairRec& R = A.P.r[p];
  const int id1 = R.id1[0], id2 = R.id2[0];
  const bool isNew = !R.hasGeom[0];
  double a0, a1;
  const V3<double> sh = pair_shift(A, p);
  if (LAW == 0 && S.sh) st3s(S.sh, slot, S.cap, sh);
  if (LAW == 0 && isNew) {
    const double* g1 = A.B.geo + (int64_t)id1 * GEO_STRIDE;
    const double* g2 = A.B.geo + (int64_t)id2 * GEO_STRIDE;
    const V3<double> d0 = (mk3(g2[GEO_POS], g2[GEO_POS + 1], g2[GEO_POS + 2]) + sh) -
                          mk3(g1[GEO_POS], g1[GEO_POS + 1], g1[GEO_POS + 2]);
    a0 = xatan2(d0.y, d0.x) + 1e-5;
    a1 = xatan(d0.z / sqrt(d0.x * d0.x + d0.y * d0.y)) + 1e-5;
  } else if (isNew) {
    a0 = 0;
    a1 = 0;
  } else {
    a0 = R.ang[0];
    a1 = R.ang[1];
  }
  S.pair[slot] = (int32_t)p;
  S.id1[slot] = id1;
  S.id2[slot] = id2;
  S.a0[slot] = a0;
  S.a1[slot] = a1;
  S.del[slot] = isNew ? 0.1 : 0.001;
  S.nev[slot] = 0;
  S.tag[slot] = 0;
  S.touch[slot] = 1;
  st3s(S.ct, slot, S.cap, mk3(0.0, 0.0, 0.0));
}

// Wall-superquadric pairs (Ig2_Wall_Superquadrics_SuperquadricsGeom(2)::go,
// Superquadrics_Ig2.cpp:2522-2703) and sphere-sphere pairs (the sphere branches of both Ig2
// superquadric functors, :1950-1970, :2331-2350), listed by k_pairs_prep.
template <int LAW>
__global__ __launch_bounds__(256) void k_pairs_misc(StepArgs A, const int32_t* mlist,
                                                    const unsigned long long* count) {
  const int64_t nm = (int64_t)*count;
  unsigned long long ce = 0, de = 0;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < nm; k += (int64_t)gridDim.x * blockDim.x) {
    const PairArrays& P = A.P;
    const BodyArrays& B = A.B;
    const int64_t N = B.n;
    const int64_t p = mlist[k];
    const int id1 = P.r[p].id1[0], id2 = P.r[p].id2[0];
    const bool hadGeom = P.r[p].hasGeom[0] != 0;
    const bool hadPhys = P.r[p].hasPhys[0] != 0;
    const bool isNew = !hadGeom;
        const BodyView b1 = load_body(B, id1);
    const BodyView b2 = load_body(B, id2);
    GeomState g;
    load_geom_prev(P, p, hadGeom, g);
    g.shift2 = pair_shift(A, p);
    // position2 of the superquadric functors: shifted in Geom (LM), not in Geom2 (NM)
    const V3<double> x2 = LAW == 0 ? b2.pos + g.shift2 : b2.pos;
    int outcome = GEO_FULL, nev = 0;
    V3<double> cn = mk3(0.0, 0.0, 0.0);
    if (b1.kind == SDM_KIND_WALL) {
      // Ig2_Wall_Superquadrics_SuperquadricsGeom(2)::go
      const int PA = b1.wallAxis;
      V3<double> wn = mk3(0.0, 0.0, 0.0);
      const double sense = (double)b1.wallSense;
      if (PA == 0) wn.x = sense; else if (PA == 1) wn.y = sense; else wn.z = sense;
      const double wallc = PA == 0 ? b1.pos.x : (PA == 1 ? b1.pos.y : b1.pos.z);
      if (b2.isSphere) {
        V3<double> q1 = b2.pos;
        if (PA == 0) q1.x = wallc; else if (PA == 1) q1.y = wallc; else q1.z = wallc;
        const V3<double> point2 = b2.pos - wn * b2.rmax;
        const double depth = b2.rmax - norm(b2.pos - q1);
        if (depth < 0.0) {
          g.depth = 0; g.isShearNew = true; outcome = GEO_EARLY;
        } else {
          g.point1 = q1; g.point2 = point2;
          g.cpt = mk3(0.5 * (point2.x + q1.x), 0.5 * (point2.y + q1.y), 0.5 * (point2.z + q1.z));
          g.depth = depth;
        }
      } else {
        const M3<double>& G2 = b2.G;
        const V3<double> point2 = support(b2.s, G2, -wn) + b2.pos;
        nev = 1;
        V3<double> q1 = point2;
        if (PA == 0) q1.x = wallc; else if (PA == 1) q1.y = wallc; else q1.z = wallc;
        const V3<double> d = point2 - q1;
        const double nPA = PA == 0 ? wn.x : (PA == 1 ? wn.y : wn.z);
        const double dPA = PA == 0 ? d.x : (PA == 1 ? d.y : d.z);
        if (nPA * dPA >= 0.0) {
          g.depth = 0; g.isShearNew = true; outcome = GEO_EARLY;
        } else {
          g.point1 = q1; g.point2 = point2;
          g.cpt = mk3(0.5 * (point2.x + q1.x), 0.5 * (point2.y + q1.y), 0.5 * (point2.z + q1.z));
          g.depth = norm(d);
        }
      }
      cn = wn;
    } else {
      ce++;
      double para0, para1;
      if (LAW == 0 && isNew) {
        const V3<double> d0 = x2 - b1.pos;
        para0 = xatan2(d0.y, d0.x);
        para1 = xatan(d0.z / sqrt(d0.x * d0.x + d0.y * d0.y));
        para0 += 1e-5;
        para1 += 1e-5;
      } else {
        para0 = g.a0;
        para1 = g.a1;
      }
      if (b1.isSphere && b2.isSphere) {
        V3<double> dist = x2 - b1.pos;
        const double r_sum = b1.rmax + b2.rmax;
        const double depth = r_sum - norm(dist);
        if (depth < 1E-18) {
          g.depth = 0; g.isShearNew = true; outcome = GEO_EARLY;
        } else {
          dist = normalized(dist);
          const V3<double> q1 = b1.pos + b1.rmax * dist;
          const V3<double> q2 = x2 - b2.rmax * dist;
          cn = normalized(q1 - q2);
          g.point1 = q1; g.point2 = q2;
          g.cpt = mk3(0.5 * (q1.x + q2.x), 0.5 * (q1.y + q2.y), 0.5 * (q1.z + q2.z));
          g.depth = depth;
          if (LAW == 1) {
            g.curvSum = r_sum / (b1.rmax * b2.rmax);
            g.gSphere = true;
            g.a0 = 0;  // Geom2 keeps its local para(0,0) on this branch (:2335)
            g.a1 = 0;
          } else {
            g.a0 = para0;
            g.a1 = para1;
          }
        }
      }
    }
    if (outcome == GEO_FULL) precompute(g, b1, b2, A.dt, cn, g.isShearNew);
    de += (b1.kind == SDM_KIND_WALL) ? 0 : nev;
    phys_law_store<LAW>(A, p, b1, b2, id1, g, outcome, hadGeom, hadPhys, nev, 0);
  }
  add_counters(A.counters, ce, de);
}

// The first Nelder-Mead evaluation of every queued pair (T: evaluation type).  A pair not
// touching there is complete: its outcome is stored here (angles unchanged, one function
// call).  Touching pairs are listed with a sort key that puts the expected long solves first.
template <typename T>
__global__ __launch_bounds__(256) void k_nm_first(StepArgs A, SolveArgs SA, uint32_t* tflag, uint64_t* tkey) {
  const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nw = work_n(SA);
  bool go = false;
  uint32_t key = 0;
  if (w < nw) {
    const PairArrays& P = A.P;
    const int64_t p = SA.S.pair[w];
    const bool hadGeom = P.r[p].hasGeom[0] != 0;
    if (nm_first_eval<T>(SA, w)) {
      go = true;
      key = nm_cost_key(P.r[p].nEval[0], hadGeom && P.r[p].depth[0] > 0, SA.S.del[w]);
    } else {
      early_store<1>(A, p, SA.S.id1[w], SA.S.id2[w], hadGeom, P.r[p].hasPhys[0] != 0, SA.S.a0[w], SA.S.a1[w],
                     GEO_EARLY_ANG, 1, 1);
    }
  }
  if (w < SA.nwork) {  // flags up to the upper bound: zero past the count, for the scan
    tflag[w] = go ? 1u : 0u;
    tkey[w] = key;
  }
  add_evals(SA.evals, w < nw ? 1ull : 0ull);
  // listed pairs are counted when they finish; the others are complete here
  add_counters(A.counters, 0, (w < nw && !go) ? 1ull : 0ull);
}

// The first distance evaluation of every queued pair of the Levenberg-Marquardt pass, as a
// certified float screen (nm_first_apart, with the global contact direction: no contact
// frame).  A pair whose float touching operand c.(p2 - p1) exceeds its rounding bound is not
// touching at its start angles: lmder2 sets its sticky stop flag at that evaluation and its
// first outer check ends it with stop 9 (Superquadrics_Ig2.cpp:679-683, 1127-1147), angles and
// contact direction unchanged, after one evaluation.  That outcome is stored here; the points
// of a pair that is not touching are never read.  Flags: 1 for the pairs left to the solver.
__global__ __launch_bounds__(256) void k_lm_first(SolveArgs A, uint32_t* tflag) {
  const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nw = work_n(A);
  bool apart = false;
  if (w < nw) {
    GeoRec r2 = load_rec(A.geo, A.S.id2[w]);
    if (A.S.sh) r2.pos = r2.pos + ld3s(A.S.sh, w, A.S.cap);
    const GeoRec r1 = load_rec(A.geo, A.S.id1[w]);
    const M3<double> I = {1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0};
    apart = nm_first_apart(r1, r2, I, A.S.a0[w], A.S.a1[w], (float)kLmMinTrig);
    if (apart) {
      A.S.nev[w] += 1;
      A.S.stop[w] = 9;
      A.S.touch[w] = 0;
    }
  }
  if (w < A.nwork) tflag[w] = (w < nw && !apart) ? 1u : 0u;  // zero past the count, for the scan
  add_evals(A.evals, apart ? 1ull : 0ull);
}

// Compaction of the flagged slots, in slot order.
__global__ __launch_bounds__(256) void k_list_compact(const uint32_t* flag, const uint32_t* pos, int64_t n,
                                                      uint32_t* listOut) {
  const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= n || !flag[w]) return;
  listOut[pos[w]] = (uint32_t)w;
}

// Compaction of the touching slots (in slot order) with their sort keys.
__global__ __launch_bounds__(256) void k_list_fill(const uint32_t* flag, const uint32_t* pos, const uint64_t* key,
                                                   int64_t n, uint64_t* keysOut, uint32_t* listOut) {
  const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= n || !flag[w]) return;
  keysOut[pos[w]] = key[w];
  listOut[pos[w]] = (uint32_t)w;
}

// Curvature directions and the Hertz terms of a Geom2 contact (Superquadrics_Ig2.cpp:2459-2503)
// in double, with the correctly rounded functions (the FP64 form) or the device's (kFast, the
// mixed form).  A float form of these was off by up to 0.1 rad in the angle between the
// curvature directions (an acos near +-1 of nearly parallel directions, principal directions
// of nearly equal curvatures) for ~0.5 % of contacts on MI355X, which moved their Hertz
// stiffness by up to 50 %.
template <bool kFast>
__device__ __forceinline__ void contact_curvature(const BodyView& b1, const BodyView& b2, double a0, double a1,
                                                  V3<double> cn, double out[9]) {
  using T = double;
  const M3<T> gC = crotation_mat(b1.pos, b2.pos);
  const M3<T>& G1 = b1.G;
  const M3<T>& G2 = b2.G;
  const Shape<T>& s1 = b1.s;
  const Shape<T>& s2 = b2.s;
  T st0, ct0, st1, ct1;
  msincos<kFast>(a0, st0, ct0);
  msincos<kFast>(a1, st1, ct1);
  const V3<T> contact = mul(gC, mk3(ct0 * st1, st0 * st1, ct1));
  const V3<T> c1 = mulT(G1, contact), c2 = mulT(G2, contact);
  V3<T> dmax1 = mk3(T(0.0), T(0.0), T(0.0)), dmax2 = dmax1, dmin1, dmin2;
  bool f1, f2;
  T ph0, ph1, k10, k11, k20, k21;
  normal2phi<T, kFast>(s1, c1, ph0, ph1);
  sq_curvatures<kFast>(s1, G1, ph0, ph1, k10, k11, dmax1, dmin1, f1);
  normal2phi<T, kFast>(s2, mk3(T(-1.0) * c2.x, T(-1.0) * c2.y, T(-1.0) * c2.z), ph0, ph1);
  sq_curvatures<kFast>(s2, G2, ph0, ph1, k20, k21, dmax2, dmin2, f2);
  const V3<T> cd1 = (f1 && f2) ? dmax1 : dmin1;
  const V3<T> cd2 = (f1 && f2) ? dmax2 : dmin2;
  const V3<T> dir1 = mk3(cd1.x * cn.x, cd1.y * cn.y, cd1.z * cn.z) - cd1;
  const V3<T> dir2 = mk3(cd2.x * cn.x, cd2.y * cn.y, cd2.z * cn.z) - cd2;
  T angle = macos<kFast>(dot(dir1, dir2) / (norm(dir1) * norm(dir2)));
  if (isnan(angle)) angle = T(0.0);
  T cs, al, K, E;
  hertz_mindlin<kFast>(k10, k11, k20, k21, angle, cs, al, K, E);
  out[0] = k10; out[1] = k11; out[2] = k20; out[3] = k21; out[4] = angle;
  out[5] = cs; out[6] = al; out[7] = K; out[8] = E;
}

// Queued pairs after the solvers: geometry outputs, curvature / Hertz terms, precompute, law.
// kFast: the curvature with the device's double functions (the mixed form).
template <int LAW, bool kFast = false>
__global__ __launch_bounds__(256) void k_pairs_finish(StepArgs A, SolveIO S, const uint32_t* list, int64_t nmax,
                                                      const unsigned long long* count) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nq = count ? ((int64_t)*count < nmax ? (int64_t)*count : nmax) : nmax;
  unsigned long long de = 0;
  if (k < nq) {
    const int64_t w = list ? (int64_t)list[k] : k;
    const PairArrays& P = A.P;
    const BodyArrays& B = A.B;
    const int64_t N = B.n;
    const int64_t p = S.pair[w];
    const int id1 = S.id1[w], id2 = S.id2[w];
    const bool hadGeom = P.r[p].hasGeom[0] != 0;
    const bool hadPhys = P.r[p].hasPhys[0] != 0;
    const BodyView b1 = load_body(B, id1);
    const BodyView b2 = load_body(B, id2);
    GeomState g;
    load_geom_prev(P, p, hadGeom, g);
    g.shift2 = pair_shift(A, p);
    const int nev = S.nev[w];
    de = nev;
    g.a0 = S.a0[w];
    g.a1 = S.a1[w];
    int outcome = GEO_FULL;
    V3<double> cn = mk3(0.0, 0.0, 0.0);
    const int stopCode = S.stop[w] + S.tag[w];
    if (!S.touch[w]) {
      g.depth = 0;
      g.isShearNew = true;
      outcome = GEO_EARLY_ANG;
    } else if (LAW == 0) {
      const V3<double> p1 = ld3s(S.p1, w, S.cap), p2 = ld3s(S.p2, w, S.cap), ct = ld3s(S.ct, w, S.cap);
      g.point1 = p1;
      g.point2 = p2;
      g.cpt = ld3s(S.ctr, w, S.cap);
      g.depth = dot(p1 - p2, ct);
      cn = normalized(ct);
    } else {
      const V3<double> p1 = ld3s(S.p1, w, S.cap), p2 = ld3s(S.p2, w, S.cap);
      cn = normalized(p1 - p2);
      g.point1 = p1;
      g.point2 = p2;
      g.cpt = mk3(0.5 * (p1.x + p2.x), 0.5 * (p1.y + p2.y), 0.5 * (p1.z + p2.z));
      g.depth = S.f[w];
      // curvature direction convention of :2463-2465
      double cv[9];
      contact_curvature<kFast>(b1, b2, g.a0, g.a1, cn, cv);
      g.c10 = cv[0]; g.c11 = cv[1]; g.c20 = cv[2]; g.c21 = cv[3];
      g.curvAngle = cv[4];
      g.curvSum = cv[5]; g.alpha = cv[6]; g.K = cv[7]; g.E = cv[8];
    }
    if (outcome == GEO_FULL) precompute(g, b1, b2, A.dt, cn, g.isShearNew);
    phys_law_store<LAW>(A, p, b1, b2, id1, g, outcome, hadGeom, hadPhys, nev, stopCode);
  }
  add_counters(A.counters, 0, de);
}
