// delta.h -- the distance function of the Nelder-Mead solver as a float difference from a
// double anchor (mixed precision form).
//
// Disfun (pkg/dem/Superquadrics_Ig2.cpp:1624-1678) is |p2 - p1|, two support points of
// millimetre-size bodies whose difference is a micrometre-size gap.  Evaluated in float, each
// support point carries rounding of a few ulp of its millimetre coordinates, ~1e-10 m, which
// is ~1e-4 of the gap: Nelder-Mead comparisons inside that band go either way.  Here one
// evaluation per pair -- the anchor, at the solve's first vertex -- is done in double, and
// every other evaluation computes only its difference from the anchor in float, in relative
// (logarithmic) form, so that its rounding scales with the difference (a few ulp of a
// micrometre-size displacement) and not with the support point.
//
// For a body with local direction n (normalised direction in body axes) the support point of
// Normal2Phi + getSurface (Superquadrics.cpp:110-120, 269-307) is, with e1 = 1/(2 - eps1),
// e2 = 1/(2 - eps2):
//   t  = tan phi0 = sgn(ny) |ry ny|^e1 / (sgn(nx) |rx nx|^e1)
//   t' = tan phi1 = sgn(nz) |rz nz |cos phi0|^(2-eps1)|^e2 / |rx nx|^e2
//   q  = (sgn(c0) rx |c0|^eps1 |c1|^eps2, sgn(s0) ry |s0|^eps1 |c1|^eps2, sgn(s1) rz |s1|^eps2)
// with c = cos, s = sin of phi0, phi1.  When no component of n changes sign, every factor
// changes by a relative amount:
// Every factor changes by a ratio: with lx = ln|1 + dnx/nx| (the same for y, z),
//   u0 = e1 (ly - lx) = d ln|t|,  t'^2 / t^2 = e^(2 u0)
//   d ln|c0| = -1/2 ln(1 + sin^2 phi0 (e^(2 u0) - 1)),  d ln|s0| = u0 + d ln|c0|
//   u1 = e2 (lz + (2 - eps1) d ln|c0| - lx),  d ln c1 = -1/2 ln(1 + sin^2 phi1 (e^(2 u1) - 1))
//   d ln|s1| = u1 + d ln c1,  d ln|qx| = eps1 d ln|c0| + eps2 d ln c1,  ...
//   dq = q (e^(d ln|q|) - 1), or -q (e^(d ln|q|) + 1) where the component's sign changes
// (sgn qx = sgn nx, sgn qy = sgn ny, sgn qz = sgn nz), evaluated with log1p / expm1 built on
// the hardware log2 / exp2: the rounding of each term scales with the change.  The
// difference of the contact direction itself comes from half-angle forms of the angle steps.
// A component of n that comes within kNear of its anchor value of zero (|1 + dn/n| < kNear),
// or a factor outside the exponentials' range, is not evaluated this way: the evaluation
// reports it, and the solver evaluates that point in double.
#pragma once
#include "narrow.h"

// log(1 + x) for x > -1 and exp(x) - 1 in float, a few ulp of the result: the rounding of
// 1 + x and of exp(x) is undone by the ratio (Goldberg's and Kahan's forms).  The ratios use
// the hardware reciprocal (1 ulp) instead of a correctly rounded division.
SDM_HD float dlog1p(float x) {
#pragma clang fp contract(fast)
  const float u = 1.0f + x;
  const float l = __builtin_amdgcn_logf(u) * 0.693147180559945309f;  // ln u
  return u == 1.0f ? x : l * (x * __builtin_amdgcn_rcpf(u - 1.0f));
}
SDM_HD float dexpm1(float x) {
#pragma clang fp contract(fast)
  const float u = __builtin_amdgcn_exp2f(x * 1.44269504088896341f);
  const float l = __builtin_amdgcn_logf(u) * 0.693147180559945309f;
  return (u == 1.0f || l == 0.0f) ? x : (u - 1.0f) * (x * __builtin_amdgcn_rcpf(l));
}
// log|1 + x|: log1p above -1; below, the log of -(1 + x) (a sign change, not near -1 here)
SDM_HD float dlogabs1p(float x) {
  return x > -1.0f ? dlog1p(x) : __builtin_amdgcn_logf(-1.0f - x) * 0.693147180559945309f;
}

// The anchor of one body: the support point and its factors at the anchor direction.
struct DeltaBody {
  float in[3];   // reciprocals of the local direction's components at the anchor
  float sq0, sq1;  // sin^2 phi0, sin^2 phi1
  float cq0, cq1;  // cos^2 phi0, cos^2 phi1 (each to its own relative accuracy)
  float q[3];    // local support point
  float e1, e2, eps1, eps2, t1;  // exponents: 1/(2 - eps1), 1/(2 - eps2), eps1, eps2, 2 - eps1
  M3<float> G;   // rot_mat2global
};

// a^b for a >= 0 as exp2(b log2 a) with the device's double log2 / exp2: a relative error of
// a few 1e-15 (|b log2 a| < 60 here) against the ~1e-16 of pow, far below the float rounding
// of the differences taken from the anchor, and a fraction of pow's registers, which the
// solver's inline re-anchoring has to leave to its other waves.
SDM_HD double apow(double a, double b) { return exp2(b * log2(a)); }

// Double support point along local direction n with the anchor factors; returns false when a
// factor is degenerate (a zero direction component or a zero divisor), where the relative form
// does not apply.  kFast: the device's double functions.
SDM_HD bool delta_anchor_local(const Shape<double>& s, V3<double> n, DeltaBody& b, V3<double>& q) {
  const double e1 = 1.0 / (2.0 - s.eps1), e2 = 1.0 / (2.0 - s.eps2);
  if (n.x == 0.0 || n.y == 0.0 || n.z == 0.0) return false;
  const double ax = fabs(s.rx * n.x);
  const double a = sgn(n.y) * apow(fabs(s.ry * n.y), e1);
  const double bb = sgn(n.x) * apow(ax, e1);
  const double t = a / bb;
  const double h0 = sqrt(a * a + bb * bb);
  const double c0 = bb / h0, s0 = a / h0;
  const double ap = sgn(n.z) * apow(fabs(s.rz * n.z * apow(fabs(c0), 2.0 - s.eps1)), e2);
  const double bp = apow(ax, e2);
  const double tp = ap / bp;
  const double h1 = sqrt(ap * ap + bp * bp);
  const double c1 = bp / h1, s1 = ap / h1;
  const double pc1 = apow(c1, s.eps2);
  q = mk3(sgn(c0) * s.rx * apow(fabs(c0), s.eps1) * pc1, sgn(s0) * s.ry * apow(fabs(s0), s.eps1) * pc1,
          sgn(s1) * s.rz * apow(fabs(s1), s.eps2));
  b.in[0] = (float)(1.0 / n.x); b.in[1] = (float)(1.0 / n.y); b.in[2] = (float)(1.0 / n.z);
  b.sq0 = (float)(s0 * s0);
  b.sq1 = (float)(s1 * s1);
  b.cq0 = (float)(c0 * c0);
  b.cq1 = (float)(c1 * c1);
  b.q[0] = (float)q.x; b.q[1] = (float)q.y; b.q[2] = (float)q.z;
  b.e1 = (float)e1; b.e2 = (float)e2; b.eps1 = (float)s.eps1; b.eps2 = (float)s.eps2; b.t1 = (float)(2.0 - s.eps1);
  const bool ok = h0 > 0.0 && h1 > 0.0 && bb != 0.0 && bp != 0.0 && c0 != 0.0 && s0 != 0.0 && c1 != 0.0 && s1 != 0.0;
  // exponents 1/(2 - eps) above 2 (eps > 1.5, towards the octahedral limit) amplify the float
  // rounding of the logarithms beyond the stated bound: such bodies are evaluated in double
  return ok && fabs(t) < 1e30 && fabs(tp) < 1e30 && s.eps1 <= 1.5 && s.eps2 <= 1.5;
}

// Smallest magnitude of a new direction component relative to its anchor value that the
// relative form takes (|1 + dn/n|): below it the logarithm of the ratio leaves float range.
constexpr float kNear = 9.094947017729282e-13f;  // 2^-40
// Largest magnitude of the logarithmic factors (e^(2 u) stays far from float overflow).
constexpr float kLogMax = 40.0f;

// ln(a + b e^(2u)) for a, b >= 0 with a + b = 1: the log1p of w = b (e^(2u) - 1) while w stays
// above -1/2, accurate to its own relative rounding however small the change; below, the
// logarithm of the sum of two positive terms, which has no cancellation and whose magnitude
// is then above ln 2.
SDM_HD float lnmix(float a, float b, float u) {
#pragma clang fp contract(fast)
  const float w = b * dexpm1(2.0f * u);
  float r = dlog1p(w);
  if (w < -0.5f)  // rare: a branch, not a select, so that waves without such a lane skip it
    r = __builtin_amdgcn_logf(a + b * __builtin_amdgcn_exp2f(2.0f * u * 1.44269504088896341f)) *
        0.693147180559945309f;
  return r;
}

// Change of the global support point of a body when its local direction changes by dn.
// With l_k = ln|1 + dn_k/n_k| the changes of the logarithms of the factors are
//   u0 = e1 (ly - lx)                      tan phi0 ratio
//   lc0 = -1/2 ln(cq0 + sq0 e^(2 u0))      cos phi0 ratio
//   ls0 = -1/2 ln(sq0 + cq0 e^(-2 u0))     sin phi0 ratio
//   u1 = e2 (lz + t1 lc0 - lx) = e2 (lz + t1 ls0 - ly)   (t1 e1 = 1)
// and the same for phi1; each is taken in the form without cancellation for small changes:
// of lc0 and ls0 the one of the larger of cos^2 / sin^2 through lnmix, the other from ls0 - lc0
// = u0; u1 through the form whose middle term (t1 lc0 or t1 ls0) is the smaller.  A change
// of any size is then carried with rounding that scales with it, except where two large
// logarithms cancel in u0, u1 or the exponents g (the direction leaving a principal plane on
// two axes at once); `lgk` bounds that absolute rounding, (|qx| + |qy| + |qz|) (|lx| + |ly| +
// |lz| + |u0| + |u1|).  Sets `ok` false when a component comes within kNear of zero relative
// to its anchor value or a factor leaves kLogMax.  `cond` bounds the conditioning of the ratio
// form, (|qx| + |qy| + |qz|) sum_k 1/min(|n_k|, |n_k + dn_k|): the rounding of a direction
// change (~ulp |dn| in each component) enters through dn_k / n_k and every logarithmic factor
// passes it on with a coefficient of at most one.
SDM_HD V3<float> delta_support(const DeltaBody& b, const float dn[3], bool& ok, float& cond, float& lgk) {
#pragma clang fp contract(fast)
  const float rx = dn[0] * b.in[0], ry = dn[1] * b.in[1], rz = dn[2] * b.in[2];
  const float ax = fabsf(1.0f + rx), ay = fabsf(1.0f + ry), az = fabsf(1.0f + rz);
  const float lx = dlogabs1p(rx), ly = dlogabs1p(ry), lz = dlogabs1p(rz);
  // of the two logarithms of an angle, the one on the side of the larger of cos^2 / sin^2
  // through lnmix, the other as its difference with u (ls - lc = u exactly), which then does
  // not cancel for small changes
  const float u0 = b.e1 * (ly - lx);
  const bool c0big = b.cq0 >= 0.5f;
  const float m0 = -0.5f * lnmix(c0big ? b.cq0 : b.sq0, c0big ? b.sq0 : b.cq0, c0big ? u0 : -u0);
  const float lc0 = c0big ? m0 : m0 - u0, ls0 = c0big ? m0 + u0 : m0;
  const float u1 = b.e2 * (lz + (fabsf(lc0) <= fabsf(ls0) ? b.t1 * lc0 - lx : b.t1 * ls0 - ly));
  const bool c1big = b.cq1 >= 0.5f;
  const float m1 = -0.5f * lnmix(c1big ? b.cq1 : b.sq1, c1big ? b.sq1 : b.cq1, c1big ? u1 : -u1);
  const float lc1 = c1big ? m1 : m1 - u1, ls1 = c1big ? m1 + u1 : m1;
  const float gx = b.eps1 * lc0 + b.eps2 * lc1, gy = b.eps1 * ls0 + b.eps2 * lc1, gz = b.eps2 * ls1;
  ok = ok && ax >= kNear && ay >= kNear && az >= kNear && fabsf(u0) <= kLogMax && fabsf(u1) <= kLogMax &&
       fabsf(gx) <= kLogMax && fabsf(gy) <= kLogMax && fabsf(gz) <= kLogMax;
  const float ex = dexpm1(gx), ey = dexpm1(gy), ez = dexpm1(gz);
  const V3<float> dq = mk3(b.q[0] * (rx < -1.0f ? -(ex + 2.0f) : ex), b.q[1] * (ry < -1.0f ? -(ey + 2.0f) : ey),
                           b.q[2] * (rz < -1.0f ? -(ez + 2.0f) : ez));
  const float qn = fabsf(b.q[0]) + fabsf(b.q[1]) + fabsf(b.q[2]);
  cond = qn * (fabsf(b.in[0]) * __builtin_amdgcn_rcpf(fminf(ax, 1.0f)) +
               fabsf(b.in[1]) * __builtin_amdgcn_rcpf(fminf(ay, 1.0f)) +
               fabsf(b.in[2]) * __builtin_amdgcn_rcpf(fminf(az, 1.0f)));
  lgk = qn * (fabsf(lx) + fabsf(ly) + fabsf(lz) + fabsf(u0) + fabsf(u1));
  return mul(b.G, dq);
}

// A pair's anchor for the solver: both bodies, the contact direction's angles and the gap
// vector at the anchor (global frame, CRotationMat frame gC of Geom2).
struct DeltaPair {
  DeltaBody b1, b2;
  M3<float> gC;
  float ca0, sa0, ca1, sa1;  // cos / sin of the anchor angles
  float Fh[3], Fl[3];        // p2 - p1 at the anchor as a float pair (hi + lo: double accuracy)
  float cu[3];               // contact direction (global) at the anchor
};

// The anchor in double at angles (a0, a1): fills `P` and returns |p2 - p1| in double, the
// points p1, p2, the touching test (Disfun :1670-1674) and its operand cd = c.(p2 - p1).  `ok`
// is false when the relative form does not apply to this pair.
SDM_HD double delta_anchor(const PairGeo<double>& g, const M3<double>& gC, double a0, double a1, DeltaPair& P,
                           V3<double>& p1, V3<double>& p2, bool& touching, double& cd, bool& ok) {
  const double sa0 = sin(a0), ca0 = cos(a0), sa1 = sin(a1), ca1 = cos(a1);
  const V3<double> c = mul(gC, mk3(ca0 * ca1, sa0 * ca1, sa1));
  V3<double> q1, q2;
  ok = delta_anchor_local(g.s1, mulT(g.G1, c), P.b1, q1);
  ok = delta_anchor_local(g.s2, mulT(g.G2, mk3(-c.x, -c.y, -c.z)), P.b2, q2) && ok;
  p1 = mul(g.G1, q1) + g.pos1;
  p2 = mul(g.G2, q2) + g.pos2;
  const V3<double> d = p2 - p1;
  cd = dot(c, d);
  if (cd > 0.0) touching = false;
  P.b1.G = cast9<float>(g.G1);
  P.b2.G = cast9<float>(g.G2);
  P.gC = cast9<float>(gC);
  P.ca0 = (float)ca0; P.sa0 = (float)sa0; P.ca1 = (float)ca1; P.sa1 = (float)sa1;
  P.Fh[0] = (float)d.x; P.Fh[1] = (float)d.y; P.Fh[2] = (float)d.z;
  P.Fl[0] = (float)(d.x - (double)P.Fh[0]);
  P.Fl[1] = (float)(d.y - (double)P.Fh[1]);
  P.Fl[2] = (float)(d.z - (double)P.Fh[2]);
  P.cu[0] = (float)c.x; P.cu[1] = (float)c.y; P.cu[2] = (float)c.z;
  return norm(d);
}

// One evaluation at the angle steps (d0, d1) from the anchor: |p2 - p1|, the touching test
// and its operand cd, and the size `ds` of the support displacements, which bounds the
// rounding of f (the float difference is added to the anchor's gap vector in double) and,
// with ulp(f), of cd.  `ok` is false when the relative form left its range; the value is
// then not usable.
SDM_HD double delta_eval(const DeltaPair& P, float d0, float d1, bool& touching, float& cd, float& ds, float& dk,
                         float& dl, bool& ok) {
#pragma clang fp contract(fast)
  // cos / sin of (a + d) - cos / sin a by half-angle forms: sin d = 2 sin(d/2) cos(d/2),
  // cos d - 1 = -2 sin^2(d/2)
  float h0, g0, h1, g1;
  sincosf(0.5f * d0, &h0, &g0);
  sincosf(0.5f * d1, &h1, &g1);
  const float sh0 = 2.0f * h0 * g0, sh1 = 2.0f * h1 * g1;
  const float vc0 = -2.0f * h0 * h0, vc1 = -2.0f * h1 * h1;  // cos d - 1
  const float dca0 = P.ca0 * vc0 - P.sa0 * sh0, dsa0 = P.sa0 * vc0 + P.ca0 * sh0;
  const float dca1 = P.ca1 * vc1 - P.sa1 * sh1, dsa1 = P.sa1 * vc1 + P.ca1 * sh1;
  const float ca1 = P.ca1 + dca1;
  // u = (cos a0 cos a1, sin a0 cos a1, sin a1): du = d(x) y + x* d(y)
  const V3<float> du = mk3(dca0 * ca1 + P.ca0 * dca1, dsa0 * ca1 + P.sa0 * dca1, dsa1);
  const V3<float> dc = mul(P.gC, du);
  const V3<float> dn1 = mulT(P.b1.G, dc);
  const V3<float> dn2 = mulT(P.b2.G, dc);
  const float n1[3] = {dn1.x, dn1.y, dn1.z};
  const float n2[3] = {-dn2.x, -dn2.y, -dn2.z};
  float k1 = 0.0f, k2 = 0.0f, l1 = 0.0f, l2 = 0.0f;
  const V3<float> ds1 = delta_support(P.b1, n1, ok, k1, l1);
  const V3<float> ds2 = delta_support(P.b2, n2, ok, k2, l2);
  const V3<double> d = mk3(((double)P.Fh[0] + (double)P.Fl[0]) + (double)(ds2.x - ds1.x),
                           ((double)P.Fh[1] + (double)P.Fl[1]) + (double)(ds2.y - ds1.y),
                           ((double)P.Fh[2] + (double)P.Fl[2]) + (double)(ds2.z - ds1.z));
  const V3<float> c = mk3(P.cu[0] + dc.x, P.cu[1] + dc.y, P.cu[2] + dc.z);
  cd = dot(c, cast3<float>(d));
  if (cd > 0.0f) touching = false;
  ds = norm(ds1) + norm(ds2);
  dk = norm(dc) * (k1 + k2);
  dl = l1 + l2;
  return norm(d);
}
