#!/usr/bin/env python
"""Table of the float64 erf used by the HIP kernels (erf_tab in xpg_erf.h), computed here in 80-digit decimal
arithmetic and written to xpg_erf_table.h:

  nodes x_i = i / 256, i = 0 .. 1536 (x <= 6; float64 erf(x) rounds to 1 beyond 5.93),
  per node f0 = erf(x_i) and f1 = erf'(x_i) = (2 / sqrt(pi)) exp(-x_i^2), each correctly rounded to float64,
  and (ERFC_TAB, with f1 again so that erfc_tab reads one 16-byte node) c0 = erfc(x_i) = 1 - erf(x_i), correctly
  rounded: erfc_tab expands about it, which keeps its relative precision in the tail (erfc(6) = 2.2e-17), where
  1 - erf_tab would have none.

erf near a node is its Taylor expansion to degree 5, erf(x_i + d) = f0 + f1 d (1 + c2 d + c3 d^2 + c4 d^3 + c5 d^4),
with the higher derivatives erf^(k) = f1 (-1)^(k-1) H_(k-1)(x) (Hermite polynomials), so that
c2 = -x, c3 = (2 x^2 - 1) / 3, c4 = x (3 - 2 x^2) / 6, c5 = (4 x^4 - 12 x^2 + 3) / 30.  With |d| <= 1/512 the
truncation is below |erf^(6)| d^6 / 720 < 4e-18.

Usage: python gen_erf_table.py > xpg_erf_table.h   (prints the worst error of the evaluation, emulated with exact
rational arithmetic and the same roundings as the device code, against 80-digit erf on stderr)
"""
from decimal import Decimal as D, getcontext
from fractions import Fraction as Fr
import math
import random
import sys

getcontext().prec = 80
INV_H = 256
N = 6 * INV_H + 1


def pi_dec():
    getcontext().prec += 4
    lasts, t, s, n, na, d, da = 0, D(3), 3, 1, 0, 0, 24
    while s != lasts:
        lasts = s
        n, na = n + na, na + 8
        d, da = d + da, da + 32
        t = (t * n) / d
        s += t
    getcontext().prec -= 4
    return +s


TWO_OVER_SQRTPI = 2 / pi_dec().sqrt()


def erf_dec(t):
    """Maclaurin series of erf; at 80 digits the cancellation for t <= 6.5 leaves more than 60 correct digits."""
    t = D(t)
    z = t * t
    term, s, n = t, t, 0
    while True:
        n += 1
        term = -term * z / n
        add = term / (2 * n + 1)
        s += add
        if abs(add) < D(10) ** -75:
            break
    return TWO_OVER_SQRTPI * s


def table():
    out = []
    for i in range(N):
        x = D(i) / INV_H
        out.append((float(erf_dec(x)), float(TWO_OVER_SQRTPI * (-(x * x)).exp())))
    return out


def erfc_table():
    """erfc at the nodes, correctly rounded (1 - the 80-digit erf keeps > 40 digits up to x = 6)."""
    return [float(1 - erf_dec(D(i) / INV_H)) for i in range(N)]


def rnd(q):
    """Round an exact rational to the nearest float64 (ties to even)."""
    return float(q)


def fma(a, b, c):
    return rnd(Fr(a) * Fr(b) + Fr(c))


def mul(a, b):
    return rnd(Fr(a) * Fr(b))


def add(a, b):
    return rnd(Fr(a) + Fr(b))


def erf_model(tab, x):
    """erf_tab of xpg_erf.h, operation for operation (every float64 operation rounded once)."""
    a = abs(x)
    if a != a or a > 6.0:                    # NaN propagates; beyond the last node the device selects +-1
        return x if a != a else math.copysign(1.0, x)
    v = a * INV_H
    i = int(round(v)) if v < N - 1 else N - 1    # rint (ties to even, as Python's round); NaN, inf -> last
    x0 = i / INV_H
    d = a - x0                               # exact: a and x0 within a factor of two, or x0 = 0
    f0, f1 = tab[i]
    x2 = mul(x0, x0)
    c2 = -x0
    c3 = mul(add(mul(2.0, x2), -1.0), 1.0 / 3.0)
    c4 = mul(mul(x0, add(3.0, -mul(2.0, x2))), 1.0 / 6.0)
    c5 = mul(add(mul(x2, add(mul(4.0, x2), -12.0)), 3.0), 1.0 / 30.0)
    q = fma(c5, d, c4)
    q = fma(q, d, c3)
    q = fma(q, d, c2)
    q = fma(q, d, 1.0)
    r = fma(mul(f1, d), q, f0)
    return math.copysign(r, x)


def _q(x0, d):
    x2 = mul(x0, x0)
    c2 = -x0
    c3 = mul(add(mul(2.0, x2), -1.0), 1.0 / 3.0)
    c4 = mul(mul(x0, add(3.0, -mul(2.0, x2))), 1.0 / 6.0)
    c5 = mul(add(mul(x2, add(mul(4.0, x2), -12.0)), 3.0), 1.0 / 30.0)
    q = fma(c5, d, c4)
    q = fma(q, d, c3)
    q = fma(q, d, c2)
    return fma(q, d, 1.0)


def erfc_model(tab, ctab, x):
    """erfc_tab of xpg_erf.h for 0 <= x <= 6, operation for operation: erfc(x0) - erf'(x0) d q(d)."""
    v = x * INV_H
    i = int(round(v)) if v < N - 1 else N - 1
    x0 = i / INV_H
    d = x - x0
    return fma(-mul(tab[i][1], d), _q(x0, d), ctab[i])


def main():
    tab = table()
    ctab = erfc_table()
    random.seed(11)
    pts = [random.uniform(0.0, 6.5) for _ in range(3000)] + [random.uniform(0.0, 0.01) for _ in range(300)]
    pts += [i / INV_H + s / 512 for i in range(0, N - 1, 37) for s in (-1, 1) if i / INV_H + s / 512 >= 0]
    worst_abs = worst_ulp = 0.0
    for x in pts:
        ref = erf_dec(x)
        got = erf_model(tab, x)
        err = abs(D(got) - ref)
        worst_abs = max(worst_abs, float(err))
        worst_ulp = max(worst_ulp, float(err) / math.ulp(float(ref)) if float(ref) else 0.0)
    print(f"erf_tab model: worst |error| {worst_abs:.3e}, {worst_ulp:.3f} ulp at {len(pts)} points in [0, 6.5]",
          file=sys.stderr)
    worst_rel = 0.0
    for x in pts:
        if x <= 6.0:
            ref = 1 - erf_dec(x)
            worst_rel = max(worst_rel, float(abs(D(erfc_model(tab, ctab, x)) - ref) / ref))
    print(f"erfc_tab model: worst relative error {worst_rel:.3e} in [0, 6]", file=sys.stderr)
    w = sys.stdout.write
    w("// generated by gen_erf_table.py: erf, erf' and erfc at x = i / 256, i = 0 .. 1536, correctly rounded to float64;\n")
    w(f"// worst |error| of erf_tab (xpg_erf.h) {worst_abs:.2e} ({worst_ulp:.2f} ulp) at {len(pts)} points in [0, 6.5];\n")
    w(f"// worst relative error of erfc_tab {worst_rel:.2e} in [0, 6]\n")
    w("#pragma once\n\nnamespace xpg {\n\n")
    w("struct alignas(16) ErfNode {\n  double f0, f1;\n};\n")
    w(f"constexpr int ERF_TAB_INV_H = {INV_H}, ERF_TAB_N = {N};\n")
    w(f"constexpr ErfNode ERF_TAB[{N}] = {{\n")
    for i in range(0, N, 2):
        w("    " + " ".join(f"{{{f0!r}, {f1!r}}}," for f0, f1 in tab[i:i + 2]) + "\n")
    w("};\n")
    w("struct alignas(16) ErfcNode {\n  double c0, f1;\n};\n")
    w(f"constexpr ErfcNode ERFC_TAB[{N}] = {{\n")
    for i in range(0, N, 2):
        w("    " + " ".join(f"{{{c!r}, {f1!r}}}," for c, (_, f1) in zip(ctab[i:i + 2], tab[i:i + 2])) + "\n")
    w("};\n\n}  // namespace xpg\n")


if __name__ == "__main__":
    main()
