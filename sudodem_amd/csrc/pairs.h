// pairs.h -- one interaction per thread: Ig2 functor, Ip2 functor and contact law.
//
// Device restatement of InteractionLoop::action (pkg/common/InteractionLoop.cpp:23-149)
// for one interaction, with
//   Ig2_Superquadrics_Superquadrics_SuperquadricsGeom::go   (Superquadrics_Ig2.cpp:1880-2317)
//   Ig2_Superquadrics_Superquadrics_SuperquadricsGeom2::go  (Superquadrics_Ig2.cpp:2322-2517)
//   Ig2_Wall_Superquadrics_SuperquadricsGeom(2)::go         (Superquadrics_Ig2.cpp:2522-2703)
//   SuperquadricsGeom::precompute / rotate                 (Superquadrics.cpp:1030-1080)
//   Ip2_SuperquadricsMat(2)_* ::go                         (Superquadrics.cpp:1086-1160)
//   SuperquadricsLaw::go / SuperquadricsLaw2::go           (Superquadrics.cpp:1181-1557)
// The force/torque of each contact is written to the interaction; the per-body sums are
// formed later in interaction order by the integrator kernel, which reproduces the serial
// ForceContainer summation order (core/ForceContainer.hpp:218-319) without atomics.
#pragma once
#include "narrow.h"
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
};

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
};

struct BodyView {
  V3<double> pos, vel, angVel, bmin, bmax;
  double mass;
  int blocked, material, kind, isSphere, wallAxis, wallSense;
  Shape<double> s;
  double rmax;
};

__device__ __forceinline__ BodyView load_body(const BodyArrays& B, int id) {
  const int64_t N = B.n;
  BodyView b;
  b.pos = ld3(B.pos, id, N);
  b.vel = ld3(B.vel, id, N);
  b.angVel = ld3(B.angVel, id, N);
  b.bmin = ld3(B.bmin, id, N);
  b.bmax = ld3(B.bmax, id, N);
  b.mass = B.mass[id];
  b.blocked = B.blocked[id];
  b.material = B.material[id];
  b.kind = B.kind[id];
  b.isSphere = B.isSphere[id];
  b.wallAxis = B.wallAxis[id];
  b.wallSense = B.wallSense[id];
  b.s.rx = B.rxyz[id];
  b.s.ry = B.rxyz[N + id];
  b.s.rz = B.rxyz[2 * N + id];
  b.s.eps1 = B.eps[id];
  b.s.eps2 = B.eps[N + id];
  b.rmax = B.rmax[id];
  return b;
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
  V3<double> c2x = (g.cpt - b2.pos) - mk3(0.0, 0.0, 0.0);
  V3<double> rv = (b2.vel + cross(b2.angVel, c2x)) - (b1.vel + cross(b1.angVel, c1x));
  double vn = dot(g.normal, rv);
  g.relVn = vn * g.normal;
  g.relVs = rv - g.relVn;
  g.shearInc = g.relVs * dt;
}

// Outcome of a functor: EARLY = only depth/isShearNew (and maybe the angles) changed.
enum { GEO_EARLY = 0, GEO_EARLY_ANG = 1, GEO_FULL = 2 };

// Frame of the mixed form: everything relative to body 1's centre, in float.
template <typename T>
__device__ __forceinline__ PairGeo<T> make_geo(const BodyView& b1, const BodyView& b2, const M3<double>& L1,
                                               const M3<double>& G1, const M3<double>& L2, const M3<double>& G2,
                                               bool relative) {
  PairGeo<T> g;
  g.s1.rx = T(b1.s.rx); g.s1.ry = T(b1.s.ry); g.s1.rz = T(b1.s.rz); g.s1.eps1 = T(b1.s.eps1); g.s1.eps2 = T(b1.s.eps2);
  g.s2.rx = T(b2.s.rx); g.s2.ry = T(b2.s.ry); g.s2.rz = T(b2.s.rz); g.s2.eps1 = T(b2.s.eps1); g.s2.eps2 = T(b2.s.eps2);
  g.L1 = cast9<T>(L1); g.G1 = cast9<T>(G1); g.L2 = cast9<T>(L2); g.G2 = cast9<T>(G2);
  if (relative) {
    g.pos1 = mk3(T(0), T(0), T(0));
    g.pos2 = cast3<T>(b2.pos - b1.pos);
  } else {
    g.pos1 = cast3<T>(b1.pos);
    g.pos2 = cast3<T>(b2.pos);
  }
  return g;
}

// Ig2_Superquadrics_Superquadrics_SuperquadricsGeom::go, non-sphere branch after the
// bounding-sphere test: LM solve with the reference's two restarts.
template <int PREC>
__device__ int lm_geometry(GeomState& g, const BodyView& b1, const BodyView& b2, const M3<double>& L1,
                           const M3<double>& G1, const M3<double>& L2, const M3<double>& G2, double& para0,
                           double& para1, V3<double>& contactNormal, int& nev, int& stopOut) {
  const V3<double> position1 = b1.pos, position2 = b2.pos;
  const PairGeo<double> gd = make_geo<double>(b1, b2, L1, G1, L2, G2, false);
  LMResult<double> R;
  R.contact = mk3(0.0, 0.0, 0.0);
  if (PREC == 1) {
    // Mixed form: iterate in float on the body-1 frame, then finish in double from the
    // float solution (the double solver stops after its first angle test when converged).
    const PairGeo<float> gf = make_geo<float>(b1, b2, L1, G1, L2, G2, true);
    LMResult<float> Rf;
    Rf.contact = mk3(0.f, 0.f, 0.f);
    float f0 = float(para0), f1 = float(para1);
    lmder2<float>(gf, f0, f1, 50, Rf, nev);
    if (Rf.touching && isfinite(f0) && isfinite(f1) && Rf.stop != 7) {
      para0 = f0;
      para1 = f1;
    }
  }
  lmder2<double>(gd, para0, para1, 50, R, nev);
  int stop = R.stop;
  if (isnan(para0) || isnan(para1) || 7 == stop) {
    if (R.touching) {
      V3<double> d0 = position2 - position1;
      para0 = atan2(d0.y, d0.x);
      para1 = atan(d0.z / sqrt(d0.x * d0.x + d0.y * d0.y));
      para0 += 1e-5;
      para1 += 1e-5;
      lmder2<double>(gd, para0, para1, 50, R, nev);
      stop = R.stop + 100;
    }
  }
  V3<double> p1 = R.p1, p2 = R.p2;
  double pDepth = dot(p1 - p2, R.contact);
  if (R.touching) {
    if (10.0 * norm(p2 - p1) > b1.rmax) {
      if (!(is_inside(b1.s, L1, p2 - position1)) || !(is_inside(b2.s, L2, p1 - position2))) {
        V3<double> d0 = position2 - position1;
        para0 = atan2(d0.y, d0.x);
        para1 = atan(d0.z / sqrt(d0.x * d0.x + d0.y * d0.y));
        para0 += 1e-5;
        para1 += 1e-5;
        lmder2<double>(gd, para0, para1, 50, R, nev);
        stop = R.stop + 200;
        p1 = R.p1;
        p2 = R.p2;
        pDepth = dot(p1 - p2, R.contact);
      }
    }
  }
  stopOut = stop;
  if (!R.touching) return 0;
  g.point1 = p1;
  g.point2 = p2;
  g.cpt = R.ctr;
  g.depth = pDepth;
  contactNormal = normalized(R.contact);
  return 1;
}

// Ig2_Superquadrics_Superquadrics_SuperquadricsGeom2::go, non-sphere branch after the
// bounding-sphere test: NM solve, curvatures and Hertz geometry.
template <int PREC>
__device__ int nm_geometry(GeomState& g, const BodyView& b1, const BodyView& b2, const M3<double>& L1,
                           const M3<double>& G1, const M3<double>& L2, const M3<double>& G2, double dels,
                           double& para0, double& para1, V3<double>& contactNormal, int& nev, int& nfuncOut) {
  NMGeo<double> nd;
  nd.g = make_geo<double>(b1, b2, L1, G1, L2, G2, false);
  nd.gC = crotation_mat(b1.pos, b2.pos);
  NMResult<double> R;
  if (PREC == 1) {
    NMGeo<float> nf;
    nf.g = make_geo<float>(b1, b2, L1, G1, L2, G2, true);
    nf.gC = cast9<float>(nd.gC);
    NMResult<float> Rf;
    neldermead<float>(nf, float(para0), float(para1), float(dels), Rf, nev);
    if (!Rf.touching) {
      para0 = Rf.a0;
      para1 = Rf.a1;
      nfuncOut = Rf.nfunc;
      return 0;
    }
    // Finish in double: restart the simplex at the float optimum with the warm-start size.
    neldermead<double>(nd, double(Rf.a0), double(Rf.a1), 0.001, R, nev);
    R.nfunc += Rf.nfunc;
  } else {
    neldermead<double>(nd, para0, para1, dels, R, nev);
  }
  para0 = R.a0;
  para1 = R.a1;
  nfuncOut = R.nfunc;
  if (!R.touching) return 0;
  const V3<double> p1 = R.p1, p2 = R.p2;
  contactNormal = normalized(p1 - p2);
  g.point1 = p1;
  g.point2 = p2;
  g.cpt = mk3(0.5 * (p1.x + p2.x), 0.5 * (p1.y + p2.y), 0.5 * (p1.z + p2.z));
  g.depth = R.fmin;
  // curvature direction convention of :2463-2465
  V3<double> contact = mul(nd.gC, mk3(cos(para0) * sin(para1), sin(para0) * sin(para1), cos(para1)));
  V3<double> c1 = mul(L1, contact), c2 = mul(L2, contact);
  V3<double> dmax1 = mk3(0.0, 0.0, 0.0), dmax2 = dmax1, dmin1, dmin2;
  bool f1, f2;
  double ph0, ph1, k10, k11, k20, k21;
  normal2phi(b1.s, c1, ph0, ph1);
  sq_curvatures(b1.s, G1, ph0, ph1, k10, k11, dmax1, dmin1, f1);
  normal2phi(b2.s, mk3(-1.0 * c2.x, -1.0 * c2.y, -1.0 * c2.z), ph0, ph1);
  sq_curvatures(b2.s, G2, ph0, ph1, k20, k21, dmax2, dmin2, f2);
  V3<double> cd1, cd2;
  if (f1 && f2) {
    cd1 = dmax1;
    cd2 = dmax2;
  } else {
    cd1 = dmin1;
    cd2 = dmin2;
  }
  V3<double> dir1 = mk3(cd1.x * contactNormal.x, cd1.y * contactNormal.y, cd1.z * contactNormal.z) - cd1;
  V3<double> dir2 = mk3(cd2.x * contactNormal.x, cd2.y * contactNormal.y, cd2.z * contactNormal.z) - cd2;
  double angle = acos(dot(dir1, dir2) / (norm(dir1) * norm(dir2)));
  if (isnan(angle)) angle = 0.0;
  g.c10 = k10; g.c11 = k11; g.c20 = k20; g.c21 = k21;
  g.curvAngle = angle;
  hertz_mindlin(k10, k11, k20, k21, angle, g.curvSum, g.alpha, g.K, g.E);
  return 1;
}

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

template <int LAW, int PREC>
__device__ void process_pair(const StepArgs& A, int64_t p, unsigned long long& ce, unsigned long long& de) {
  const PairArrays& P = A.P;
  const BodyArrays& B = A.B;
  const int64_t S = P.cap;
  const int64_t N = B.n;
  P.active[p] = 0;
  if (!P.geomExists[p]) return;
  int id1 = P.id1[p], id2 = P.id2[p];
  if (!P.functorSet[p]) {
    // dispatch order (Wall,Superquadrics) / (Superquadrics,Superquadrics); swapOrder otherwise
    const int k1 = B.kind[id1], k2 = B.kind[id2];
    if (k1 == SDM_KIND_WALL && k2 == SDM_KIND_WALL) {
      P.geomExists[p] = 0;
      return;
    }
    P.functorSet[p] = 1;
    if (k1 == SDM_KIND_SQ && k2 == SDM_KIND_WALL) {
      int t = id1; id1 = id2; id2 = t;
      P.id1[p] = id1;
      P.id2[p] = id2;
    }
  }
  const bool hadGeom = P.hasGeom[p] != 0;
  const bool hadPhys = P.hasPhys[p] != 0;
  const BodyView b1 = load_body(B, id1);
  const BodyView b2 = load_body(B, id2);

  // ---------------- geometry functor ----------------
  GeomState g;
  int outcome;
  int nev = 0, stopCode = 0;
  const bool isNew = !hadGeom;
  if (isNew) {
    // fresh SuperquadricsGeom(2): all zero, isShearNew = true
    g.depth = 0; g.a0 = 0; g.a1 = 0;
    g.normal = mk3(0.0, 0.0, 0.0);
    g.cpt = g.normal; g.shearInc = g.normal; g.relVn = g.normal; g.relVs = g.normal;
    g.twist = g.normal; g.orth = g.normal; g.point1 = g.normal; g.point2 = g.normal;
    g.isShearNew = true;
    g.c10 = g.c11 = g.c20 = g.c21 = 0; g.curvAngle = g.alpha = g.K = g.E = g.curvSum = 0;
    g.gSphere = false;
  } else {
    g.depth = P.depth[p];
    g.a0 = P.ang[p];
    g.a1 = P.ang[S + p];
    g.normal = ld3(P.normal, p, S);
    g.gSphere = P.gSphere[p] != 0;
    g.isShearNew = g.depth <= 0;
  }
  V3<double> contactNormal = mk3(0.0, 0.0, 0.0);
  if (b1.kind == SDM_KIND_WALL) {
    // Ig2_Wall_Superquadrics_SuperquadricsGeom(2)::go
    const int PA = b1.wallAxis;
    V3<double> wn = mk3(0.0, 0.0, 0.0);
    double sense = (double)b1.wallSense;
    if (PA == 0) wn.x = sense; else if (PA == 1) wn.y = sense; else wn.z = sense;
    const double wallc = PA == 0 ? b1.pos.x : (PA == 1 ? b1.pos.y : b1.pos.z);
    outcome = GEO_FULL;
    if (b2.isSphere) {
      V3<double> q1 = b2.pos;
      if (PA == 0) q1.x = wallc; else if (PA == 1) q1.y = wallc; else q1.z = wallc;
      V3<double> point2 = b2.pos - wn * b2.rmax;
      double depth = b2.rmax - norm(b2.pos - q1);
      if (depth < 0.0) {
        g.depth = 0; g.isShearNew = true; outcome = GEO_EARLY;
      } else {
        g.point1 = q1; g.point2 = point2;
        g.cpt = mk3(0.5 * (point2.x + q1.x), 0.5 * (point2.y + q1.y), 0.5 * (point2.z + q1.z));
        g.depth = depth;
      }
    } else {
      const M3<double> L2 = ld9(B.rotL, id2, N), G2 = ld9(B.rotG, id2, N);
      V3<double> point2 = support(b2.s, L2, G2, -wn) + b2.pos;
      nev = 1;
      V3<double> q1 = point2;
      if (PA == 0) q1.x = wallc; else if (PA == 1) q1.y = wallc; else q1.z = wallc;
      V3<double> d = point2 - q1;
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
    contactNormal = wn;
  } else {
    ce++;
    double para0, para1;
    if (LAW == 0) {
      if (isNew) {
        V3<double> d0 = b2.pos - b1.pos;
        para0 = atan2(d0.y, d0.x);
        para1 = atan(d0.z / sqrt(d0.x * d0.x + d0.y * d0.y));
        para0 += 1e-5;
        para1 += 1e-5;
      } else {
        para0 = g.a0;
        para1 = g.a1;
      }
    } else {
      para0 = g.a0;
      para1 = g.a1;
    }
    if (b1.isSphere && b2.isSphere) {
      V3<double> dist = b2.pos - b1.pos;
      double r_sum = b1.rmax + b2.rmax;
      double depth = r_sum - norm(dist);
      if (depth < 1E-18) {
        g.depth = 0; g.isShearNew = true; outcome = GEO_EARLY;
      } else {
        dist = normalized(dist);
        V3<double> q1 = b1.pos + b1.rmax * dist;
        V3<double> q2 = b2.pos - b2.rmax * dist;
        contactNormal = normalized(q1 - q2);
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
        outcome = GEO_FULL;
      }
    } else {
      double r_sum = b1.rmax + b2.rmax;
      if (norm(b1.pos - b2.pos) > r_sum) {
        g.depth = 0; g.isShearNew = true; outcome = GEO_EARLY;
      } else {
        const M3<double> L1 = ld9(B.rotL, id1, N), G1 = ld9(B.rotG, id1, N);
        const M3<double> L2 = ld9(B.rotL, id2, N), G2 = ld9(B.rotG, id2, N);
        int ok;
        if (LAW == 0)
          ok = lm_geometry<PREC>(g, b1, b2, L1, G1, L2, G2, para0, para1, contactNormal, nev, stopCode);
        else
          ok = nm_geometry<PREC>(g, b1, b2, L1, G1, L2, G2, isNew ? 0.1 : 0.001, para0, para1, contactNormal, nev,
                                 stopCode);
        g.a0 = para0;
        g.a1 = para1;
        if (!ok) {
          g.depth = 0; g.isShearNew = true; outcome = GEO_EARLY_ANG;
        } else {
          outcome = GEO_FULL;
        }
      }
    }
    de += nev;
  }
  if (outcome == GEO_FULL) precompute(g, b1, b2, A.dt, contactNormal, g.isShearNew);

  // ---------------- Ip2 functor ----------------
  PhysState ph;
  if (hadPhys) {
    ph.kn = P.kn[p]; ph.ks = P.ks[p]; ph.tanPhi = P.tanPhi[p];
    ph.betan = P.betan[p]; ph.betas = P.betas[p];
    ph.nuab = P.nuab[p]; ph.Gab = P.Gab[p]; ph.mu1 = P.mu1[p]; ph.mu2 = P.mu2[p];
    ph.fs = ld3(P.fs, p, S);
  } else {
    const MaterialDev m1 = A.mats[b1.material];
    const MaterialDev m2 = A.mats[b2.material];
    ph.kn = 0; ph.ks = 0; ph.tanPhi = 0; ph.nuab = 0; ph.Gab = 0; ph.mu1 = 0; ph.mu2 = 0;
    ph.fs = mk3(0.0, 0.0, 0.0);
    if (LAW == 0) {
      if (id1 > 5) {
        ph.tanPhi = tan(fmin(m1.frictionAngle, m2.frictionAngle));
        ph.kn = 2.0 * m1.Kn * m2.Kn / (m1.Kn + m2.Kn);
        ph.ks = 2.0 * m1.Ks * m2.Ks / (m1.Ks + m2.Ks);
      } else {
        ph.tanPhi = tan(m1.frictionAngle);
        ph.kn = m1.Kn;
        ph.ks = m1.Ks;
      }
    } else {
      if (id1 > 5) {
        ph.nuab = 0.5 * (m1.nu + m2.nu);
        ph.Gab = 1.0 / ((1 - m1.nu) / m1.G + (1 - m2.nu) / m2.G);
        ph.mu1 = 1. / m1.G + 1. / m2.G;
        ph.mu2 = m1.nu / m1.G + m2.nu / m2.G;
        ph.tanPhi = tan(fmin(m1.frictionAngle, m2.frictionAngle));
      } else {
        ph.tanPhi = tan(m1.frictionAngle);
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
  if (b1.bmin.x >= b2.bmax.x || b2.bmin.x >= b1.bmax.x || b1.bmin.y >= b2.bmax.y || b2.bmin.y >= b1.bmax.y ||
      b1.bmin.z >= b2.bmax.z || b2.bmin.z >= b1.bmax.z) {
    keep = false;  // requestErase: the interaction is reset below
  } else if (!(g.depth > 1E-18)) {
    ph.fn = mk3(0.0, 0.0, 0.0);
    ph.fs = ph.fn;
  } else {
    if (LAW == 1 && id1 > 5) {
      const double depth = g.depth, cs = g.curvSum;
      if (g.gSphere) {
        double tmp = sqrt(depth / cs);
        ph.kn = 8. / 3. * ph.Gab * tmp;
        ph.ks = 8. / (2.0 * ph.mu1 - ph.mu2) * tmp;
      } else {
        // The reference forms these two in long double through M_PIl (:1413, :1417); the
        // device has no extended type, so they are evaluated in double here.
        const double PI = 3.141592653589793238462643383279502884;
        double a = sqrt(depth * g.E / cs / g.K);
        ph.kn = 4. / 3. * PI * a * ph.Gab * g.alpha / g.K;
        double e2 = 1 - 1 / (g.alpha * g.alpha);
        double ksx = ph.mu1 * g.K - ph.mu2 * (g.K - g.E) / e2;
        double ksy = ph.mu1 * g.K + ph.mu2 * ((1 - e2) * g.K - g.E) / e2;
        ph.ks = PI * g.alpha * a * (1. / ksx + 1. / ksy);
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
  P.nEval[p] = nev;
  P.stop[p] = stopCode;
  if (!keep) {
    // Interaction::reset (core/Interaction.cpp:25-29): geom and phys are dropped; store the
    // defaults a fresh geom/phys will start from.
    P.hasGeom[p] = 0;
    P.hasPhys[p] = 0;
    P.geomExists[p] = 1;
    P.iterMadeReal[p] = -1;
    const V3<double> z = mk3(0.0, 0.0, 0.0);
    P.depth[p] = 0; P.ang[p] = 0; P.ang[S + p] = 0;
    st3(P.cpt, p, S, z); st3(P.shearInc, p, S, z); st3(P.relVn, p, S, z); st3(P.relVs, p, S, z);
    st3(P.normal, p, S, z); st3(P.twist, p, S, z); st3(P.orth, p, S, z); st3(P.point1, p, S, z);
    st3(P.point2, p, S, z);
    P.isShearNew[p] = 1;
    P.curv1[p] = 0; P.curv1[S + p] = 0; P.curv2[p] = 0; P.curv2[S + p] = 0;
    P.curvAngle[p] = 0; P.alpha[p] = 0; P.hK[p] = 0; P.hE[p] = 0; P.curvSum[p] = 0; P.gSphere[p] = 0;
    P.kn[p] = 0; P.ks[p] = 0; P.tanPhi[p] = __builtin_nan(""); P.betan[p] = 0; P.betas[p] = 0;
    st3(P.fn, p, S, z); st3(P.fs, p, S, z);
    P.nuab[p] = 0; P.Gab[p] = 0; P.mu1[p] = 0; P.mu2[p] = 0;
    return;
  }
  if (!(hadGeom && hadPhys)) P.iterMadeReal[p] = A.iter;
  P.hasGeom[p] = 1;
  P.hasPhys[p] = 1;
  P.depth[p] = g.depth;
  P.isShearNew[p] = g.isShearNew ? 1 : 0;
  if (outcome != GEO_EARLY) {
    P.ang[p] = g.a0;
    P.ang[S + p] = g.a1;
  }
  if (outcome == GEO_FULL) {
    st3(P.normal, p, S, g.normal); st3(P.cpt, p, S, g.cpt); st3(P.shearInc, p, S, g.shearInc);
    st3(P.relVn, p, S, g.relVn); st3(P.relVs, p, S, g.relVs); st3(P.twist, p, S, g.twist);
    st3(P.orth, p, S, g.orth); st3(P.point1, p, S, g.point1); st3(P.point2, p, S, g.point2);
    if (LAW == 1 && b1.kind != SDM_KIND_WALL) {
      P.curvSum[p] = g.curvSum;
      P.gSphere[p] = g.gSphere ? 1 : 0;
      if (!g.gSphere) {
        P.curv1[p] = g.c10; P.curv1[S + p] = g.c11; P.curv2[p] = g.c20; P.curv2[S + p] = g.c21;
        P.curvAngle[p] = g.curvAngle; P.alpha[p] = g.alpha; P.hK[p] = g.K; P.hE[p] = g.E;
      }
    }
  }
  if (!hadPhys) {
    P.tanPhi[p] = ph.tanPhi; P.betan[p] = ph.betan; P.betas[p] = ph.betas;
    P.nuab[p] = ph.nuab; P.Gab[p] = ph.Gab; P.mu1[p] = ph.mu1; P.mu2[p] = ph.mu2;
  }
  P.kn[p] = ph.kn;
  P.ks[p] = ph.ks;
  st3(P.fn, p, S, ph.fn);
  st3(P.fs, p, S, ph.fs);
  if (contributed) {
    P.active[p] = 1;
    st3(P.cF, p, S, F);
    st3(P.cT1, p, S, -cross(b1.pos - g.cpt, F));
    st3(P.cT2, p, S, cross(b2.pos - g.cpt, F));
  }
}

template <int LAW, int PREC>
__global__ __launch_bounds__(256) void k_interactions(StepArgs A) {
  unsigned long long ce = 0, de = 0;
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < A.np; p += (int64_t)gridDim.x * blockDim.x)
    process_pair<LAW, PREC>(A, p, ce, de);
  // one atomic per wavefront
  for (int off = 32; off > 0; off >>= 1) {
    ce += __shfl_down(ce, off);
    de += __shfl_down(de, off);
  }
  if ((threadIdx.x & 63) == 0) {
    atomicAdd(&A.counters[0], ce);
    atomicAdd(&A.counters[1], de);
  }
}
