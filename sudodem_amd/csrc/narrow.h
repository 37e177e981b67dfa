// narrow.h -- superellipsoid support map and the two common-normal solvers.
//
// Device restatement of the SudoDEM3D narrowphase:
//   support map     Superquadrics::Normal2Phi / getSurface / isInside (pkg/dem/Superquadrics.cpp:103-120, 269-307)
//   curvature       Superquadrics::Curvatures (Superquadrics.cpp:131-255)
//   LM solver       Disfunction, JacfunctionApp, lmder2 (pkg/dem/Superquadrics_Ig2.cpp:621-687, 1032-1354)
//   NM solver       CRotationMat, Disfun, amotry, neldermead_simplex (Superquadrics_Ig2.cpp:295-319, 1624-1875)
//   Hertz geometry  SuperquadricsGeom2::HertzMindlin (Superquadrics.cpp:974-1026)
// The solvers are templated on the evaluation type T: T=double is the reference
// arithmetic; T=float evaluates the support map on frames centred at body 1.
#pragma once
#include "dev_math.h"

template <typename T>
struct Shape {
  T rx, ry, rz, eps1, eps2;
};

// Math of the support map in double: correctly rounded (kFast = false; the FP64 form, whose
// values must round as the reference's glibc rounds them) or the device library's (kFast; the
// mixed form's final double evaluation, which needs double accuracy, not the same last bits).
// Float keeps the device's float functions either way.
template <bool kFast>
SDM_HD double mpow(double a, double b) {
  if constexpr (kFast) return pow(a, b);
  else return crm::cr_pow(a, b);
}
template <bool kFast>
SDM_HD float mpow(float a, float b) { return powf(a, b); }
template <bool kFast>
SDM_HD double matan2(double a, double b) {
  if constexpr (kFast) return atan2(a, b);
  else return crm::cr_atan2(a, b);
}
template <bool kFast>
SDM_HD float matan2(float a, float b) { return atan2f(a, b); }
template <bool kFast>
SDM_HD double mcos(double a) {
  if constexpr (kFast) return cos(a);
  else return crm::cr_cos(a);
}
template <bool kFast>
SDM_HD float mcos(float a) { return cosf(a); }
// sin and cos of one angle (the same values as the separate functions)
template <bool kFast>
SDM_HD void msincos(double a, double& sn, double& cs) {
  if constexpr (kFast) {
    sn = sin(a);
    cs = cos(a);
  } else {
    crm::cr_sincos(a, &sn, &cs);
  }
}
template <bool kFast>
SDM_HD void msincos(float a, float& sn, float& cs) {
  sn = sinf(a);
  cs = cosf(a);
}
template <bool kFast>
SDM_HD double msin(double a) {
  if constexpr (kFast) return sin(a);
  else return crm::cr_sin(a);
}
template <bool kFast>
SDM_HD double mlog(double a) {
  if constexpr (kFast) return log(a);
  else return xlog(a);
}
template <bool kFast>
SDM_HD double mlog10(double a) {
  if constexpr (kFast) return log10(a);
  else return xlog10(a);
}
template <bool kFast>
SDM_HD double macos(double a) {
  if constexpr (kFast) return acos(a);
  else return xacos(a);
}

// Normal2Phi, active branches (Superquadrics.cpp:284, 288).
template <typename T, bool kFast = false>
SDM_HD void normal2phi(const Shape<T>& s, V3<T> n, T& ph0, T& ph1) {
  const T e1 = T(1) / (T(2) - s.eps1);
  const T e2 = T(1) / (T(2) - s.eps2);
  T px1, px2;  // |rx nx|^e1 and |rx nx|^e2
  if constexpr (sizeof(T) == 8 && !kFast) {
    crm::cr_pow_pair(xfabs(s.rx * n.x), e1, e2, &px1, &px2);  // one logarithm for both
  } else {
    px1 = mpow<kFast>(xfabs(s.rx * n.x), e1);
    px2 = mpow<kFast>(xfabs(s.rx * n.x), e2);
  }
  ph0 = matan2<kFast>(sgn(n.y) * mpow<kFast>((xfabs(s.ry * n.y)), e1), sgn(n.x) * px1);
  ph1 = matan2<kFast>(sgn(n.z) * mpow<kFast>(xfabs(s.rz * n.z * mpow<kFast>(xfabs(mcos<kFast>(ph0)), T(2) - s.eps1)), e2),
                      px2);
}

// getSurface (Superquadrics.cpp:110-120): local coordinates of the point with parameters phi.
template <typename T, bool kFast = false>
SDM_HD V3<T> get_surface(const Shape<T>& s, T ph0, T ph1) {
  T c0, s0, c1, s1;
  msincos<kFast>(ph0, s0, c0);
  msincos<kFast>(ph1, s1, c1);
  const T pc1 = mpow<kFast>(xfabs(c1), s.eps2);
  return mk3(sgn(c0) * s.rx * mpow<kFast>(xfabs(c0), s.eps1) * pc1,
             sgn(s0) * s.ry * mpow<kFast>(xfabs(s0), s.eps1) * pc1, sgn(s1) * s.rz * mpow<kFast>(xfabs(s1), s.eps2));
}

// rot_mat2local * v.  The body caches rot_mat2local and rot_mat2global from the same
// normalised quaternion (Superquadrics.hpp:86-91, NewtonIntegrator.cpp:426-427), and
// toRotationMatrix of the conjugate is bitwise the transpose, so only rot_mat2global is read.
template <typename T>
SDM_HD V3<T> mulT(const M3<T>& G, V3<T> v) {
  return mk3(G.m[0] * v.x + (G.m[3] * v.y + G.m[6] * v.z), G.m[1] * v.x + (G.m[4] * v.y + G.m[7] * v.z),
             G.m[2] * v.x + (G.m[5] * v.y + G.m[8] * v.z));
}

// Support point of a body along the global direction dir (global frame, relative to its centre).
template <typename T, bool kFast = false>
SDM_HD V3<T> support(const Shape<T>& s, const M3<T>& G, V3<T> dir) {
  T ph0, ph1;
  normal2phi<T, kFast>(s, mulT(G, dir), ph0, ph1);
  return mul(G, get_surface<T, kFast>(s, ph0, ph1));
}

// ---------------------------------------------------------------------------------------
// Float support map of the mixed form.  Same formula as normal2phi + get_surface, evaluated
// with two identities and a short pow, since the float path only steers the solver (its
// result is re-evaluated in double):
//  * cos(atan2(a, b)) = b / |(a, b)| and sin(atan2(a, b)) = a / |(a, b)| (atan2(0, 0) = 0):
//    the four sin/cos of getSurface and the cos of Normal2Phi follow from the two atan2
//    arguments with one reciprocal square root each;
//  * x^y for x >= 0, y > 0 as 2^(y log2 x): hardware log2, one Newton step on 2^L = x for
//    the logarithm, y L as an exact product (fma residual), and hardware exp2 with a first
//    order correction for the low part -- a few ulp, against ~150 instructions for powf.
//  * sgn(x) * p for p >= 0 as copysign(p, x) (equal but for the sign of a zero result).
//  * the exponents 1/(2 - eps) and 2 - eps of a body are computed once per pair (SupF).
// The evaluation has no branches: special cases are selects after the arithmetic.
// ---------------------------------------------------------------------------------------
SDM_HD float fpow(float x, float y) {
  const float L0 = __builtin_amdgcn_logf(x);               // log2 x, hardware
  const float r = fmaf(x, __builtin_amdgcn_exp2f(-L0), -1.0f);  // x 2^-L0 - 1, |r| ~ 1e-7
  const float Llo = r * 1.44269504088896341f;              // r / ln 2: Newton correction of L0
  const float p = y * L0;
  const float plo = fmaf(y, L0, -p) + y * Llo;             // y L = p + plo
  const float z = __builtin_amdgcn_exp2f(p);
  const float v = fmaf(z, plo * 0.693147180559945309f, z);  // 2^p (1 + plo ln 2)
  return x > 0.0f ? v : 0.0f;                              // 0 for x <= 0 (and NaN)
}

// (cos, sin) of atan2(a, b) without the angle: (b, a) divided by the larger magnitude,
// then normalised.
SDM_HD void unit_angle(float a, float b, float& c, float& s) {
  const float m = fmaxf(fabsf(a), fabsf(b));
  const bool ok = m > 0.0f;
  const float im = 1.0f / (ok ? m : 1.0f);
  const float an = a * im, bn = b * im;
  const float ih = __builtin_amdgcn_rsqf(an * an + bn * bn);  // argument in [1, 2]: no rescaling
  c = ok ? bn * ih : 1.0f;
  s = ok ? an * ih : 0.0f;
}

// A body's shape with the exponents of the float support map.
struct SupF {
  float rx, ry, rz, eps1, eps2;
  float e1, e2, t1;  // 1 / (2 - eps1), 1 / (2 - eps2), 2 - eps1
};
SDM_HD SupF sup_f(const Shape<float>& s) {
  SupF f;
  f.rx = s.rx; f.ry = s.ry; f.rz = s.rz; f.eps1 = s.eps1; f.eps2 = s.eps2;
  f.e1 = 1.0f / (2.0f - s.eps1);
  f.e2 = 1.0f / (2.0f - s.eps2);
  f.t1 = 2.0f - s.eps1;
  return f;
}

// minTrig: the smallest of the four cos / sin factors of the support-map angles is written
// there when given.
SDM_HD V3<float> support_f(const SupF& s, const M3<float>& G, V3<float> dir, float* minTrig = nullptr) {
  const V3<float> n = mulT(G, dir);
  const float ax = fabsf(s.rx * n.x);
  float c0, s0, c1, s1;
  unit_angle(copysignf(fpow(fabsf(s.ry * n.y), s.e1), n.y), copysignf(fpow(ax, s.e1), n.x), c0, s0);
  unit_angle(copysignf(fpow(fabsf(s.rz * n.z * fpow(fabsf(c0), s.t1)), s.e2), n.z), fpow(ax, s.e2), c1, s1);
  if (minTrig) *minTrig = fminf(fminf(fabsf(c0), fabsf(s0)), fminf(fabsf(c1), fabsf(s1)));
  const float pc1 = fpow(fabsf(c1), s.eps2);
  const V3<float> q = mk3(copysignf(s.rx, c0) * fpow(fabsf(c0), s.eps1) * pc1,
                          copysignf(s.ry, s0) * fpow(fabsf(s0), s.eps1) * pc1, copysignf(s.rz, s1) * fpow(fabsf(s1), s.eps2));
  return mul(G, q);
}

// isInside (Superquadrics.cpp:103-108); pg relative to the centre, global frame.
SDM_HD bool is_inside(const Shape<double>& s, const M3<double>& G, V3<double> pg) {
  V3<double> p = mulT(G, pg);
  return 1 > xpow(xpow(fabs(p.x / s.rx), 2.0 / s.eps1) + xpow(fabs(p.y / s.ry), 2.0 / s.eps1), s.eps1 / s.eps2) +
                 xpow(fabs(p.z / s.rz), 2.0 / s.eps2);
}

// The pair geometry a solver needs, in evaluation type T.  pos1/pos2 are the centres in the
// frame the solver works in (global for T=double, body-1-centred for the mixed form).
template <typename T>
struct PairGeo {
  Shape<T> s1, s2;
  M3<T> G1, G2;
  V3<T> pos1, pos2;
};

// Disfunction (Superquadrics_Ig2.cpp:621-687).  stop_flag is sticky: set, never cleared.
template <typename T, bool kFast = false>
SDM_HD void lm_eval(const PairGeo<T>& g, T a0, T a1, V3<T>& p1, V3<T>& p2, V3<T>& center, bool& stop_flag,
                    int& nev) {
  nev++;
  T sa0, ca0, sa1, ca1;
  msincos<kFast>(a0, sa0, ca0);
  msincos<kFast>(a1, sa1, ca1);
  const V3<T> contact = mk3(ca0 * ca1, sa0 * ca1, sa1);
  p1 = support<T, kFast>(g.s1, g.G1, contact) + g.pos1;
  {
    T ph0, ph1;
    V3<T> c2 = mulT(g.G2, contact);
    normal2phi<T, kFast>(g.s2, mk3(T(-1) * c2.x, T(-1) * c2.y, T(-1) * c2.z), ph0, ph1);
    p2 = mul(g.G2, get_surface<T, kFast>(g.s2, ph0, ph1)) + g.pos2;
  }
  V3<T> d = p2 - p1;
  center = mk3(T(0.5) * (p1.x + p2.x), T(0.5) * (p1.y + p2.y), T(0.5) * (p1.z + p2.z));
  if (dot(contact, d) > T(0)) stop_flag = true;
}

// CRotationMat (Superquadrics_Ig2.cpp:295-319).
SDM_HD M3<double> crotation_mat(V3<double> position1, V3<double> position2) {
  V3<double> e1 = position2 - position1;
  V3<double> x1 = mk3(1.0, 0.0, 0.0);
  double cos_phi = dot(x1, e1) / (norm(x1) * norm(e1));
  double cos_halfphi = sqrt(0.5 * (1 + cos_phi));
  double sin_halfphi = sqrt(0.5 * (1 - cos_phi));
  Q4 q;
  if (fabs(cos_phi) > 1.0 - 1e-10) {
    q.w = cos_halfphi; q.x = 0; q.y = 0; q.z = sin_halfphi;
  } else {
    V3<double> n1 = normalized(cross(x1, e1));
    n1 = n1 * sin_halfphi;
    q.w = cos_halfphi; q.x = n1.x; q.y = n1.y; q.z = n1.z;
  }
  return quat_rot(q);
}

template <typename T>
struct NMGeo {
  PairGeo<T> g;
  M3<T> gC;  // CRotationMat frame
  SupF f1, f2;  // float form: the shapes with their exponents (nm_prepare)
};

// Per-pair set-up of the float support map.
template <typename T>
SDM_HD void nm_prepare(NMGeo<T>& n) {
  if constexpr (sizeof(T) == 4) {
    n.f1 = sup_f(n.g.s1);
    n.f2 = sup_f(n.g.s2);
  }
}

// Disfun (Superquadrics_Ig2.cpp:1624-1678): distance |p2-p1| and the contact test.
template <typename T, bool kFast = false>
SDM_HD T nm_eval(const NMGeo<T>& n, T a0, T a1, V3<T>& p1, V3<T>& p2, bool& touching, int& nev) {
  nev++;
  if constexpr (sizeof(T) == 4) {
    float s0, c0, s1, c1;  // one argument reduction per angle
    sincosf(a0, &s0, &c0);
    sincosf(a1, &s1, &c1);
    const V3<T> contact = mul(n.gC, mk3(c0 * c1, s0 * c1, s1));
    p1 = support_f(n.f1, n.g.G1, contact) + n.g.pos1;
    p2 = support_f(n.f2, n.g.G2, mk3(-contact.x, -contact.y, -contact.z)) + n.g.pos2;
    V3<T> d = p2 - p1;
    if (dot(contact, d) > T(0)) touching = false;
    return norm(d);
  } else {
    T sa0, ca0, sa1, ca1;
    msincos<kFast>(a0, sa0, ca0);
    msincos<kFast>(a1, sa1, ca1);
    const V3<T> contact = mul(n.gC, mk3(ca0 * ca1, sa0 * ca1, sa1));
    p1 = support<T, kFast>(n.g.s1, n.g.G1, contact) + n.g.pos1;
    T ph0, ph1;
    V3<T> c2 = mulT(n.g.G2, contact);
    normal2phi<T, kFast>(n.g.s2, mk3(T(-1) * c2.x, T(-1) * c2.y, T(-1) * c2.z), ph0, ph1);
    p2 = mul(n.g.G2, get_surface<T, kFast>(n.g.s2, ph0, ph1)) + n.g.pos2;
    V3<T> d = p2 - p1;
    if (dot(contact, d) > T(0)) touching = false;
    return norm(d);
  }
}

// Curvatures (Superquadrics.cpp:131-255).  Keeps the reference's expressions: K ignores the
// M term, and R11/R22 use the coefficient choices of :172-178.  kFast: the device's double
// functions instead of the correctly rounded ones.
template <bool kFast = false>
SDM_HD void sq_curvatures(const Shape<double>& s, const M3<double>& Rg, double ph0, double ph1, double& k1o, double& k2o,
                          V3<double>& dir_max, V3<double>& dir_min, bool& max_flag) {
  using T = double;
  max_flag = true;
  if (xfabs(ph0) < T(1E-20)) ph0 = T(1E-20);
  if (xfabs(ph1) < T(1E-20)) ph1 = T(1E-20);
  const T rx = s.rx, ry = s.ry, rz = s.rz, eps1 = s.eps1, eps2 = s.eps2;
  T s1 = msin<kFast>(ph0), s2 = msin<kFast>(ph1), c1 = mcos<kFast>(ph0), c2 = mcos<kFast>(ph1);
  T s_s1 = sgn(s1), s_s2 = sgn(s2), s_c1 = sgn(c1), s_c2 = sgn(c2);
  T abs_s1 = xfabs(s1), abs_s2 = xfabs(s2), abs_c1 = xfabs(c1), abs_c2 = xfabs(c2);
  T s10 = s_s1 * mpow<kFast>(abs_s1, eps1);
  T s20 = s_s2 * mpow<kFast>(abs_s2, eps2);
  T c10 = s_c1 * mpow<kFast>(abs_c1, eps1);
  T c20 = s_c2 * mpow<kFast>(abs_c2, eps2);
  T s12 = s_s1 * mpow<kFast>(abs_s1, eps1 - T(2.0));
  T s22 = s_s2 * mpow<kFast>(abs_s2, eps2 - T(2.0));
  T c12 = s_c1 * mpow<kFast>(abs_c1, eps1 - T(2.0));
  T c22 = s_c2 * mpow<kFast>(abs_c2, eps2 - T(2.0));
  V3<T> R1 = mk3(-rx * eps1 * c12 * c1 * c20 * s1, ry * eps1 * c1 * c20 * s12 * s1, T(0.0));
  V3<T> R2 = mk3(-rx * eps2 * c10 * c22 * c2 * s2, -ry * eps2 * c22 * c2 * s10 * s2, rz * eps2 * c2 * s22 * s2);
  V3<T> R11 = mk3(rx * eps1 * c12 * s20 * s1 * s1 * (eps1 - T(1.0)) - rx * eps1 * c10 * c20,
                  ry * eps1 * c1 * c1 * c20 * s12 * (eps1 - T(1.0)) - ry * eps1 * c20 * s10, T(0.0));
  V3<T> R22 = mk3(rx * eps2 * c10 * c22 * s2 * s2 * (eps2 - T(1.0)) - rx * eps2 * c10 * c20,
                  ry * eps2 * s10 * s1 * s1 * (eps2 - T(1.0)) - ry * eps2 * c20 * s10,
                  rz * eps2 * c2 * c2 * s22 * (eps2 - T(1.0)) - rz * eps2 * s20);
  V3<T> RR = cross(R1, R2);
  T nr = norm(RR);
  V3<T> n = mk3(RR.x / nr, RR.y / nr, RR.z / nr);
  T L = dot(R11, n);
  T N = dot(R22, n);
  T E = dot(R1, R1);
  T F = dot(R1, R2);
  T G = dot(R2, R2);
  T K = L * N / (E * G - F * F);
  T H = T(0.5) * (E * N + G * L) / (E * G - F * F);
  T tmp0 = xsqrt(xfabs(H * H - K));
  T k1 = H + tmp0;
  T k2 = H - tmp0;
  const T a1 = N - k1 * G, a2 = N - k2 * G;
  V3<T> dir1 = mk3(R1.x * a1 + R2.x * k1 * F, R1.y * a1 + R2.y * k1 * F, R1.z * a1 + R2.z * k1 * F);
  if (norm(dir1) == T(0)) {
    max_flag = false;
  } else {
    dir1 = normalized(dir1);
    dir_max = mul(Rg, dir1);
  }
  V3<T> dir2 = mk3(R1.x * a2 + R2.x * k2 * F, R1.y * a2 + R2.y * k2 * F, R1.z * a2 + R2.z * k2 * F);
  dir2 = normalized(dir2);
  dir_min = mul(Rg, dir2);
  k1o = k1;
  k2o = k2;
}

// SuperquadricsGeom2::HertzMindlin (Superquadrics.cpp:974-1026).
template <bool kFast = false>
SDM_HD void hertz_mindlin(double c10, double c11, double c20, double c21, double theta, double& curvSum, double& alpha,
                          double& Ko, double& Eo) {
  using T = double;
  T rho11 = xfabs(c10), rho12 = xfabs(c11), rho21 = xfabs(c20), rho22 = xfabs(c21);
  if (rho11 * rho12 < T(1E-18)) rho11 = rho12 = T(0.0);
  if (rho21 * rho22 < T(1E-18)) rho21 = rho22 = T(0.0);
  T sum_rho = rho11 + rho12 + rho21 + rho22;
  T a1 = rho11 - rho12, a2 = rho21 - rho22;
  T F_rho = -xsqrt(a1 * a1 + a2 * a2 + T(2) * (xfabs(rho11 - rho12)) * xfabs(rho21 - rho22) * mcos<kFast>(T(2) * theta));
  T A = (sum_rho + F_rho) / T(4.);
  T B = (sum_rho - F_rho) / T(4.);
  curvSum = A + B;
  T lg = mlog10<kFast>(B / A);
  T x_squ = lg * lg;
  T gamma1 = T(2) * (T(1) + x_squ * (T(0.40227436) + x_squ * (T(3.7491752e-2) + x_squ * (T(7.4855761e-4) +
                                                                                       x_squ * T(2.1667028e-6)))));
  T gamma2 = T(3) * (T(1) + x_squ * (T(0.42678878) + x_squ * (T(4.2605401e-2) + x_squ * (T(9.0786922e-4) +
                                                                                       x_squ * T(2.7868927e-6)))));
  alpha = mpow<kFast>((B / A), gamma1 / gamma2);
  T m1 = T(1) / (alpha * alpha);
  Ko = (T(1.3862944) + m1 * (T(0.1119723) + T(0.0725296) * m1)) -
       mlog<kFast>(m1) * (T(0.5) + m1 * (T(0.1213478) + T(0.0288729) * m1));
  Eo = (T(1) + m1 * (T(0.4630151) + T(0.1077812) * m1)) - mlog<kFast>(m1) * (T(0.2452727) + T(0.0412496) * m1) * m1;
}
