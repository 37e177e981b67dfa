// crmath.h -- correctly rounded double pow, log, log10, atan2, atan, acos, sin, cos for the FP64 path.
//
// The reference runs on glibc, whose pow, log, atan2, atan, acos, sin and cos return the
// correctly rounded result for all but ~0.1 % of arguments (measured against 120-bit mpmath);
// its log10 is not correctly rounded and is reproduced by its own formula (cr_log10).  The
// device library's versions are within one ulp but differ from the correctly rounded result
// for 2-27 % of arguments, and the narrowphase solvers turn such last-bit differences into
// different iteration paths.  These versions evaluate in double-double arithmetic to a
// relative error near 2^-70 and round once, so they return the correctly rounded result
// except for arguments whose exact result lies within ~2^-18 ulp of a rounding boundary.
// Methods: Tang's table-driven logarithm and exponential (log m = -log c_i + log1p(m c_i - 1),
// exp t = 2^(N/128) exp(s)), table-driven arctangent (atan q = atan(i/64) + atan((q - i/64) /
// (1 + q i/64))), Cody-Waite reduction by pi/2 with sin/cos of j/64 tables.  Special
// arguments (zeros, infinities, NaNs, non-positive bases, results near overflow or
// underflow, angles beyond 2^20) go to the device library.  The tables are generated by
// gen_crmath_tables.py.
#pragma once
#include <math.h>

#include "crmath_tables.h"

// Function qualifier: host and device, inlined (a host-only build defines it as `inline`).
#ifndef CRM_FN
#define CRM_FN __host__ __device__ __forceinline__
#endif

namespace crm {

struct DD {
  double hi, lo;
};

CRM_FN DD two_sum(double a, double b) {
  const double s = a + b;
  const double bb = s - a;
  const double e = (a - (s - bb)) + (b - bb);
  return {s, e};
}
CRM_FN DD fast_two_sum(double a, double b) {  // |a| >= |b| or a == 0
  const double s = a + b;
  return {s, b - (s - a)};
}
CRM_FN DD two_prod(double a, double b) {
  const double p = a * b;
  return {p, fma(a, b, -p)};
}
CRM_FN DD dd_add(DD a, DD b) {
  DD s = two_sum(a.hi, b.hi);
  const DD t = two_sum(a.lo, b.lo);
  s.lo += t.hi;
  s = fast_two_sum(s.hi, s.lo);
  s.lo += t.lo;
  return fast_two_sum(s.hi, s.lo);
}
CRM_FN DD dd_add_d(DD a, double b) {
  DD s = two_sum(a.hi, b);
  s.lo += a.lo;
  return fast_two_sum(s.hi, s.lo);
}
CRM_FN DD dd_mul(DD a, DD b) {
  DD p = two_prod(a.hi, b.hi);
  p.lo += a.hi * b.lo + a.lo * b.hi;
  return fast_two_sum(p.hi, p.lo);
}
CRM_FN DD dd_mul_d(DD a, double b) {
  DD p = two_prod(a.hi, b);
  p.lo += a.lo * b;
  return fast_two_sum(p.hi, p.lo);
}
CRM_FN DD dd_div(DD a, DD b) {
  const double q1 = a.hi / b.hi;
  const DD r = dd_add(a, dd_mul_d(b, -q1));
  const double q2 = r.hi / b.hi;
  return fast_two_sum(q1, q2);
}
CRM_FN DD dd_neg(DD a) { return {-a.hi, -a.lo}; }

// ---- pow for x > 0 ----------------------------------------------------------------------
// ln x = k ln2 - ln c_i + log1p(r),  r = m c_i - 1 (exact as a double-double), |r| < 2^-7.9
CRM_FN DD log_dd(double x) {
  int e;
  const double f = frexp(x, &e);  // x = f 2^e, f in [0.5, 1)
  const double m = 2.0 * f;       // [1, 2)
  const int k = e - 1;
  const int i = (int)((m - 1.0) * 128.0);
  const CrT4 L = kCrLog[i];  // c_i, -ln c_i (hi, lo)
  const double c = L.a;
  const DD p = two_prod(m, c);
  const DD r = two_sum(p.hi - 1.0, p.lo);  // p.hi - 1 is exact (p.hi within 1 % of 1)
  // log1p(r) = r - r^2/2 + r^3 (1/3 - r/4 + r^2/5 - ... + r^8/11)
  const double rh = r.hi;
  double q = 1.0 / 11;
  q = q * rh - 1.0 / 10;
  q = q * rh + 1.0 / 9;
  q = q * rh - 1.0 / 8;
  q = q * rh + 1.0 / 7;
  q = q * rh - 1.0 / 6;
  q = q * rh + 1.0 / 5;
  q = q * rh - 1.0 / 4;
  q = q * rh + 1.0 / 3;
  const double r3q = rh * rh * rh * q;
  DD sq = two_prod(rh, rh);
  sq.lo += 2.0 * rh * r.lo;
  // l1p = r - r^2/2 + r^3 q with |r| > r^2/2 > |r^3 q|: exact head, tails summed
  DD l1p = fast_two_sum(rh, -0.5 * sq.hi);
  l1p.lo += (r.lo - 0.5 * sq.lo) + r3q;
  // k ln2 (ln2 hi on 40 bits: k ln2_hi is exact) - ln c_i + l1p; |k ln2| > |ln c_i| when k != 0
  const DD head = fast_two_sum((double)k * kCrLn2Hi, L.b);
  DD t = two_sum(head.hi, l1p.hi);
  t.lo += ((head.lo + l1p.lo) + (double)k * kCrLn2Lo) + L.c;
  return fast_two_sum(t.hi, t.lo);
}

// exp(t) for a double-double t with |t.hi| < 708
CRM_FN DD exp_dd(DD t) {
  const double nd = rint(t.hi * kCrInvLn2x128);
  const int n = (int)nd;
  const int j = n & 127;
  const int m = (n - j) / 128;
  // s = t - n ln2/128, with ln2/128 in three parts (n A and n B exact)
  const double s0 = t.hi - nd * kCrLn2o128A;  // exact
  DD s = two_sum(s0, -nd * kCrLn2o128B);
  s.lo += t.lo - nd * kCrLn2o128C;
  s = fast_two_sum(s.hi, s.lo);
  // exp(s) = 1 + s + s^2/2 + s^3 (1/6 + s/24 + s^2/120 + s^3/720 + s^4/5040), |s| < 2^-8.4
  const double sh = s.hi;
  double q = 1.0 / 5040;
  q = q * sh + 1.0 / 720;
  q = q * sh + 1.0 / 120;
  q = q * sh + 1.0 / 24;
  q = q * sh + 1.0 / 6;
  const double s3q = sh * sh * sh * q;
  DD sq = two_prod(sh, sh);
  sq.lo += 2.0 * sh * s.lo;
  const DD a1 = fast_two_sum(1.0, sh);
  const DD a2 = fast_two_sum(a1.hi, 0.5 * sq.hi);
  const DD e = fast_two_sum(a2.hi, ((a1.lo + a2.lo) + (s.lo + 0.5 * sq.lo)) + s3q);
  const CrT2 T = kCrExp[j];
  const DD tj = {T.hi, T.lo};
  DD r = dd_mul(tj, e);
  r.hi = ldexp(r.hi, m);
  r.lo = ldexp(r.lo, m);
  return r;
}

CRM_FN double cr_pow(double x, double y) {
  if (!(x > 0.0) || !(x < INFINITY) || !(fabs(y) < INFINITY) || x < 2.2250738585072014e-308) return pow(x, y);
  if (x == 1.0 || y == 0.0) return 1.0;
  if (y == 1.0) return x;
  const DD l = log_dd(x);
  const DD t = dd_mul_d(l, y);
  if (!(fabs(t.hi) < 700.0)) return pow(x, y);  // near overflow / underflow
  const DD r = exp_dd(t);
  return r.hi + r.lo;
}

// pow(x, y1) and pow(x, y2) with one logarithm (the same values as two cr_pow calls)
CRM_FN void cr_pow_pair(double x, double y1, double y2, double* o1, double* o2) {
  if (!(x > 0.0) || !(x < INFINITY) || !(fabs(y1) < INFINITY) || !(fabs(y2) < INFINITY) ||
      x < 2.2250738585072014e-308 || x == 1.0) {
    *o1 = cr_pow(x, y1);
    *o2 = cr_pow(x, y2);
    return;
  }
  const DD l = log_dd(x);
  const double y[2] = {y1, y2};
  double r[2];
  for (int k = 0; k < 2; k++) {
    if (y[k] == 0.0) {
      r[k] = 1.0;
    } else if (y[k] == 1.0) {
      r[k] = x;
    } else {
      const DD t = dd_mul_d(l, y[k]);
      if (!(fabs(t.hi) < 700.0)) {
        r[k] = pow(x, y[k]);
      } else {
        const DD e = exp_dd(t);
        r[k] = e.hi + e.lo;
      }
    }
  }
  *o1 = r[0];
  *o2 = r[1];
}

// ---- atan2, atan ------------------------------------------------------------------------
// atan(q) for q = qh + ql in [0, 1]: atan(i/64) + atan(u), u = (q - i/64) / (1 + q i/64)
CRM_FN DD atan01_dd(DD q) {
  const int i = (int)rint(q.hi * 64.0);
  const double c = (double)i * (1.0 / 64);
  const DD num = two_sum(q.hi - c, q.lo);  // q.hi - c is exact (Sterbenz, or c = 0)
  DD den = two_prod(q.hi, c);
  den.lo += q.lo * c;
  den = dd_add_d(den, 1.0);
  const DD u = dd_div(num, den);
  // atan(u) = u + u^3 (-1/3 + u^2/5 - u^4/7 + u^6/9 - u^8/11)
  const double uh = u.hi, u2 = uh * uh;
  double p = -1.0 / 11;
  p = p * u2 + 1.0 / 9;
  p = p * u2 - 1.0 / 7;
  p = p * u2 + 1.0 / 5;
  p = p * u2 - 1.0 / 3;
  const DD au = dd_add_d(u, uh * u2 * p);
  const CrT2 A = kCrAtan[i];
  const DD ai = {A.hi, A.lo};
  return dd_add(ai, au);
}

CRM_FN double cr_atan2(double y, double x) {
  const double ay = fabs(y), ax = fabs(x);
  if (!(ay > 1e-300) || !(ax > 1e-300) || !(ay < 1e300) || !(ax < 1e300)) return atan2(y, x);
  const bool swap = ay > ax;
  const double num = swap ? ax : ay, den = swap ? ay : ax;
  const double qh = num / den;
  if (qh < 1e-290) return atan2(y, x);
  const double ql = fma(-qh, den, num) / den;
  DD r = atan01_dd({qh, ql});
  if (swap) r = dd_add(DD{kCrPio2Hi, kCrPio2Lo}, dd_neg(r));
  if (x < 0.0) r = dd_add(DD{kCrPiHi, kCrPiLo}, dd_neg(r));
  const double v = r.hi + r.lo;
  return y < 0.0 ? -v : v;
}

CRM_FN double cr_atan(double x) {
  if (!(fabs(x) < 1e300) || fabs(x) < 1e-300) return atan(x);
  return cr_atan2(x, 1.0);
}

// ---- log, log10 -------------------------------------------------------------------------
// ln x as a double-double with a small relative error: near 1 (|x - 1| < 2^-7) from the
// series of t = x - 1 (exact there), elsewhere log_dd (absolute error ~2^-100, |ln x| > 2^-8)
CRM_FN DD log_rel_dd(double x) {
  const double t = x - 1.0;
  if (fabs(t) < 0.0078125) {
    // ln(1+t) = t - t^2/2 + t^3/3 + t^4 (-1/4 + t/5 - t^2/6 + ... - t^11/15)
    double q = -1.0 / 15;
    q = q * t + 1.0 / 14;
    q = q * t - 1.0 / 13;
    q = q * t + 1.0 / 12;
    q = q * t - 1.0 / 11;
    q = q * t + 1.0 / 10;
    q = q * t - 1.0 / 9;
    q = q * t + 1.0 / 8;
    q = q * t - 1.0 / 7;
    q = q * t + 1.0 / 6;
    q = q * t - 1.0 / 5;
    q = q * t + 1.0 / 4;
    const double t2h = t * t;
    const double t4q = -(t2h * t2h) * q;
    const DD t2 = two_prod(t, t);
    const DD t3 = dd_mul_d(t2, t);
    const DD third = {0.3333333333333333, 1.850371707708594e-17};  // 1/3 as a double-double
    DD r = fast_two_sum(t, -0.5 * t2.hi);
    r.lo += -0.5 * t2.lo;
    r = dd_add(r, dd_mul(t3, third));
    return dd_add_d(r, t4q);
  }
  return log_dd(x);
}

CRM_FN double cr_log(double x) {
  if (!(x > 2.2250738585072014e-308) || !(x < INFINITY)) return log(x);
  if (x == 1.0) return 0.0;
  const DD r = log_rel_dd(x);
  return r.hi + r.lo;
}

// log10 as glibc computes it (it is not correctly rounded there: 15 % of arguments differ
// from the nearest double): the freely published fdlibm formula, x = m 2^k with m in [1, 2)
// (in [0.5, 1) for k < 0, y = k + 1 then), log10 x = (y log10(2)_lo + log(m) / ln 10) +
// y log10(2)_hi, on the correctly rounded log.
CRM_FN double cr_log10(double x) {
  if (!(x > 2.2250738585072014e-308) || !(x < INFINITY)) return log10(x);
  const double ivln10 = 4.34294481903251816668e-01;
  const double log10_2hi = 3.01029995663611771306e-01;
  const double log10_2lo = 3.69423907715893078616e-13;
  int e;
  const double f = frexp(x, &e);  // x = f 2^e, f in [0.5, 1)
  const int k = e - 1;
  const int i = k < 0 ? 1 : 0;
  const double m = ldexp(f, 1 - i);
  const double y = (double)(k + i);
  const double z = y * log10_2lo + ivln10 * cr_log(m);
  return z + y * log10_2hi;
}

// ---- acos --------------------------------------------------------------------------------
// acos x = atan2(sqrt(1 - x^2), x) with 1 - x^2 = (1 - x)(1 + x) and its square root as
// double-doubles
CRM_FN double cr_acos(double x) {
  if (!(fabs(x) < 1.0)) return acos(x);
  if (x == 0.0) return kCrPio2Hi;
  const DD a = two_sum(1.0, -x), b = two_sum(1.0, x);
  const DD p = dd_mul(a, b);
  const double sh = sqrt(p.hi);
  const DD sq = two_prod(sh, sh);
  const double sl = ((p.hi - sq.hi) - sq.lo + p.lo) / (2.0 * sh);
  const DD s = fast_two_sum(sh, sl);
  const double ax = fabs(x);
  // atan of the ratio of the smaller to the larger of (s, |x|), as in cr_atan2
  const bool swap = s.hi > ax;
  const DD num = swap ? DD{ax, 0.0} : s, den = swap ? s : DD{ax, 0.0};
  DD r = atan01_dd(dd_div(num, den));
  if (swap) r = dd_add(DD{kCrPio2Hi, kCrPio2Lo}, dd_neg(r));
  if (x < 0.0) r = dd_add(DD{kCrPiHi, kCrPiLo}, dd_neg(r));
  return r.hi + r.lo;
}

// ---- sin, cos ---------------------------------------------------------------------------
// a = k pi/2 + r, |r| <= pi/4 (+ rounding), r as a double-double; sin r and cos r from the
// j/64 tables and the series of s = r - j/64, |s| <= 1/128.
CRM_FN void sincos_dd(double a, DD& sr, DD& cr, int& quad) {
  const double kd = rint(a * kCrTwoOverPi);
  quad = ((int)kd) & 3;
  const double r0 = a - kd * kCrPio2A;  // exact: k A is exact and the difference is small
  DD r = two_sum(r0, -kd * kCrPio2B);
  r = dd_add_d(r, -kd * kCrPio2C);
  const double rj = rint(r.hi * 64.0);
  const int j = (int)rj;
  const int aj = j < 0 ? -j : j;
  const DD s = two_sum(r.hi - rj * (1.0 / 64), r.lo);  // r.hi - j/64 is exact
  const double sh = s.hi, s2 = sh * sh;
  // sin s = s + s^3 (-1/6 + s^2/120 - s^4/5040 + s^6/362880)
  double ps = 1.0 / 362880;
  ps = ps * s2 - 1.0 / 5040;
  ps = ps * s2 + 1.0 / 120;
  ps = ps * s2 - 1.0 / 6;
  const DD sins = dd_add_d(s, sh * s2 * ps);
  // cos s = 1 - s^2/2 + s^4 (1/24 - s^2/720 + s^4/40320 - s^6/3628800)
  double pc = -1.0 / 3628800;
  pc = pc * s2 + 1.0 / 40320;
  pc = pc * s2 - 1.0 / 720;
  pc = pc * s2 + 1.0 / 24;
  DD sq = two_prod(sh, sh);
  sq.lo += 2.0 * sh * s.lo;
  sq.hi *= -0.5;
  sq.lo *= -0.5;
  DD coss = fast_two_sum(1.0, sq.hi);
  coss.lo += sq.lo;
  coss = dd_add_d(coss, s2 * s2 * pc);
  const double sg = j < 0 ? -1.0 : 1.0;
  const CrT4 SC = kCrSinCos[aj];  // sin (hi, lo), cos (hi, lo) of |j|/64
  const DD Sj = {sg * SC.a, sg * SC.b}, Cj = {SC.c, SC.d};
  sr = dd_add(dd_mul(Sj, coss), dd_mul(Cj, sins));
  cr = dd_add(dd_mul(Cj, coss), dd_neg(dd_mul(Sj, sins)));
}

CRM_FN void cr_sincos(double a, double* sp, double* cp) {
  if (!(fabs(a) < 1048576.0) || fabs(a) < 1e-300) {
    *sp = sin(a);
    *cp = cos(a);
    return;
  }
  DD s, c;
  int q;
  sincos_dd(a, s, c, q);
  const DD vs = (q == 0) ? s : (q == 1) ? c : (q == 2) ? dd_neg(s) : dd_neg(c);
  const DD vc = (q == 0) ? c : (q == 1) ? dd_neg(s) : (q == 2) ? dd_neg(c) : s;
  *sp = vs.hi + vs.lo;
  *cp = vc.hi + vc.lo;
}

CRM_FN double cr_sin(double a) {
  if (!(fabs(a) < 1048576.0) || fabs(a) < 1e-300) return sin(a);
  DD s, c;
  int q;
  sincos_dd(a, s, c, q);
  const DD v = (q == 0) ? s : (q == 1) ? c : (q == 2) ? dd_neg(s) : dd_neg(c);
  return v.hi + v.lo;
}

CRM_FN double cr_cos(double a) {
  if (!(fabs(a) < 1048576.0) || fabs(a) < 1e-300) return cos(a);
  DD s, c;
  int q;
  sincos_dd(a, s, c, q);
  const DD v = (q == 0) ? c : (q == 1) ? dd_neg(s) : (q == 2) ? dd_neg(c) : s;
  return v.hi + v.lo;
}

// ---- x87 extended precision ----------------------------------------------------------------
// Two expressions of the reference's Hertz law are long double through M_PIl (Superquadrics.cpp:
// 1413, 1417): on x86-64 every operation rounds to a 64-bit significand (nearest-even) and the
// assignment rounds that to double.  An extended value is held as hi + lo with a 64-bit
// significand; products and quotients are formed to ~2^-104 and rounded to 64 bits.
CRM_FN DD ext_round(DD v) {
  if (v.lo == 0.0 || v.hi == 0.0) return v;
  int e;
  const double f = frexp(v.hi, &e);  // ulp of a 64-bit significand at |hi|: 2^(e - 64)
  int eu = e - 64;
  if (fabs(f) == 0.5 && ((v.lo < 0.0) != (v.hi < 0.0))) eu -= 1;  // |v| below a power of two
  const double u = ldexp(1.0, eu);
  return DD{v.hi, rint(v.lo / u) * u};
}
CRM_FN DD ext_mul_d(DD x, double d) {
  return ext_round(dd_add(two_prod(x.hi, d), two_prod(x.lo, d)));
}
CRM_FN DD ext_div_d(DD x, double d) {
  const double q1 = x.hi / d;
  const DD p = two_prod(q1, d);
  DD r = two_sum(x.hi, -p.hi);
  r.lo += x.lo - p.lo;
  const double q2 = (r.hi + r.lo) / d;
  return ext_round(fast_two_sum(q1, q2));
}
CRM_FN double ext_to_double(DD x) { return x.hi + x.lo; }
// M_PIl: pi on a 64-bit significand
CRM_FN DD ext_pi() { return DD{3.141592653589793116, 1.22514845490862e-16}; }

}  // namespace crm
