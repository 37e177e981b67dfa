"""Generate crmath_tables.h: the constant tables of the correctly rounded double functions in
crmath.h, each value a double-double (hi + lo, |lo| <= ulp(hi) / 2) computed with mpmath at
200 bits.  Run from the repository root: python sudodem_amd/csrc/gen_crmath_tables.py"""
import os

import mpmath

mpmath.mp.prec = 200


def dd(v):
    hi = float(v)
    lo = float(v - mpmath.mpf(hi))
    return hi, lo


def short(v, bits):
    """v rounded to `bits` significant bits (exact in a double)."""
    m, e = mpmath.frexp(mpmath.mpf(v))
    return float(mpmath.ldexp(mpmath.nint(mpmath.ldexp(m, bits)), e - bits))


def rec(name, kind, rows):
    """An array of records (CrT2 / CrT4), one table line per entry: one wide load per lookup."""
    out = ["static constexpr %s %s[%d] = {" % (kind, name, len(rows))]
    for r in rows:
        out.append("    {" + ", ".join(repr(float(x)) for x in r) + "},")
    out.append("};")
    return "\n".join(out)


def main():
    parts = ["// Generated by gen_crmath_tables.py (mpmath, 200 bits); do not edit.", "#pragma once", "",
             "struct alignas(16) CrT2 {", "  double hi, lo;", "};",
             "struct alignas(32) CrT4 {", "  double a, b, c, d;", "};", ""]
    # log: m in [1 + i/128, 1 + (i+1)/128) -> c_i ~ 1/(1 + (i+0.5)/128) on 12 bits, -ln(c_i)
    c, lh, ll = [], [], []
    for i in range(128):
        ci = short(1 / (1 + (mpmath.mpf(i) + 0.5) / 128), 12)
        h, l = dd(-mpmath.log(mpmath.mpf(ci)))
        c.append(ci); lh.append(h); ll.append(l)
    parts += [rec("kCrLog", "CrT4", [(a, b, d, 0.0) for a, b, d in zip(c, lh, ll)])]
    # exp: 2^(j/128)
    eh, el = zip(*[dd(mpmath.power(2, mpmath.mpf(j) / 128)) for j in range(128)])
    parts += [rec("kCrExp", "CrT2", list(zip(eh, el)))]
    # atan(i/64), i = 0..64
    ah, al = zip(*[dd(mpmath.atan(mpmath.mpf(i) / 64)) for i in range(65)])
    parts += [rec("kCrAtan", "CrT2", list(zip(ah, al)))]
    # sin(j/64), cos(j/64), j = 0..51
    sh, sl = zip(*[dd(mpmath.sin(mpmath.mpf(j) / 64)) for j in range(52)])
    ch, cl = zip(*[dd(mpmath.cos(mpmath.mpf(j) / 64)) for j in range(52)])
    parts += [rec("kCrSinCos", "CrT4", list(zip(sh, sl, ch, cl)))]
    # constants: ln 2 (hi on 40 bits), ln2/128 in three parts (36, 36, 53 bits), pi/2 in three
    # parts (33, 33, 53 bits), pi/2 and pi as double-doubles
    ln2 = mpmath.log(2)
    ln2_hi = short(ln2, 40)
    ln2_lo = float(ln2 - mpmath.mpf(ln2_hi))
    q = ln2 / 128
    q1 = short(q, 36); q2 = short(q - mpmath.mpf(q1), 36); q3 = float(q - mpmath.mpf(q1) - mpmath.mpf(q2))
    p = mpmath.pi / 2
    p1 = short(p, 33); p2 = short(p - mpmath.mpf(p1), 33); p3 = float(p - mpmath.mpf(p1) - mpmath.mpf(p2))
    pio2 = dd(mpmath.pi / 2); pi = dd(mpmath.pi)
    consts = dict(kCrLn2Hi=ln2_hi, kCrLn2Lo=ln2_lo, kCrInvLn2x128=float(128 / ln2), kCrLn2o128A=q1, kCrLn2o128B=q2,
                  kCrLn2o128C=q3, kCrPio2A=p1, kCrPio2B=p2, kCrPio2C=p3, kCrTwoOverPi=float(2 / mpmath.pi),
                  kCrPio2Hi=pio2[0], kCrPio2Lo=pio2[1], kCrPiHi=pi[0], kCrPiLo=pi[1])
    for k, v in consts.items():
        parts.append("static constexpr double %s = %r;" % (k, float(v)))
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "crmath_tables.h")
    with open(path, "w") as f:
        f.write("\n".join(parts) + "\n")


if __name__ == "__main__":
    main()
