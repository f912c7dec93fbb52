#!/usr/bin/env python
"""Coefficients of the branch-free float64 erf of xpg_adapt.hip (erf_bf), derived here from the definitions in
80-digit decimal arithmetic:

  |x| <= 0.75: erf(x) = x sum_n a_n (x^2)^n,  a_n = (2/sqrt(pi)) (-1)^n / (n! (2n + 1))     (Maclaurin series)
  |x| >  0.75: erf(x) = sign(x) (1 - exp(-x^2) erfcx(|x|)),
               erfcx(t) = p(u),  u = affine map of y = (t - K)/(t + K) onto [-1, 1] for t in [0.75, T_MAX],
               p the Chebyshev interpolant of erfcx at N nodes, evaluated by Clenshaw's recurrence.

erfc is evaluated from erf's Maclaurin series at 80 digits (all terms summed exactly enough for t <= 7).
Prints C++ initialisers and the worst errors of a float64 Horner evaluation against 80-digit values.
"""
from decimal import Decimal as D, getcontext
import math
import random

getcontext().prec = 80
K, T_SWITCH, T_MAX, N, N_SMALL = 2.5, 0.75, 6.75, 20, 16


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


PI = pi_dec()
TWO_OVER_SQRTPI = 2 / PI.sqrt()


def erf_dec(t):
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


def erfcx_dec(t):
    t = D(t)
    return (1 - erf_dec(t)) * (t * t).exp()


def cos_dec(x):
    x = D(x)
    term, s, n = D(1), D(1), 0
    while True:
        n += 2
        term = -term * x * x / (n * (n - 1))
        s += term
        if abs(term) < D(10) ** -78:
            return s


def cheb_coeffs():
    """Chebyshev interpolation at N nodes, everything in decimal: c_j = (2/N) sum_i g(t(u_i)) T_j(u_i)."""
    ya, yb = (D(T_SWITCH) - D(K)) / (D(T_SWITCH) + D(K)), (D(T_MAX) - D(K)) / (D(T_MAX) + D(K))
    nodes = [cos_dec(PI * (D(i) + D("0.5")) / N) for i in range(N)]
    vals = []
    for u in nodes:
        y = ya + (u + 1) * (yb - ya) / 2
        t = D(K) * (1 + y) / (1 - y)
        vals.append(erfcx_dec(t))
    c = []
    for j in range(N):
        s = sum(v * cos_dec(j * PI * (D(i) + D("0.5")) / N) for i, v in enumerate(vals))
        c.append(s * 2 / N)
    c[0] /= 2
    return c, float(ya), float(yb)


def main():
    c, ya, yb = cheb_coeffs()
    cf = [float(v) for v in c]
    a = [float(TWO_OVER_SQRTPI * (-1) ** n / (math.factorial(n) * (2 * n + 1))) for n in range(N_SMALL)]
    u1, u0 = 2.0 / (yb - ya), -(ya + yb) / (yb - ya)

    def erf_model(x):
        ax = abs(x)
        z = x * x
        if ax <= T_SWITCH:
            s = a[-1]
            for v in a[-2::-1]:
                s = s * z + v
            return x * s
        if ax > T_MAX:
            return math.copysign(1.0, x)
        u = (ax - K) / (ax + K) * u1 + u0
        b1 = b2 = 0.0
        for v in cf[:0:-1]:
            b1, b2 = 2 * u * b1 - b2 + v, b1
        g = u * b1 - b2 + cf[0]
        return math.copysign(1.0 - math.exp(-z) * g, x)

    random.seed(5)
    worst_abs = worst_ulp = 0.0
    for _ in range(1500):
        x = random.uniform(0.0, 7.0) if random.random() < 0.8 else random.uniform(0.0, 1.0)
        ref = erf_dec(x) if x < T_MAX else D(1)
        got = erf_model(x)
        err = abs(D(got) - ref)
        worst_abs = max(worst_abs, float(err))
        worst_ulp = max(worst_ulp, float(err) / math.ulp(float(ref)) if float(ref) else 0.0)
    print(f"// generated by gen_erf_coeffs.py: K = {K}, switch {T_SWITCH}, saturation {T_MAX}, {N} Chebyshev terms, "
          f"{N_SMALL} series terms;")
    print(f"// worst |error| of a float64 model {worst_abs:.2e} ({worst_ulp:.2f} ulp) at 1500 points in [0, 7]")
    print("constexpr double ERF_K = %r, ERF_SWITCH = %r, ERF_TMAX = %r, ERF_U1 = %r, ERF_U0 = %r;"
          % (K, T_SWITCH, T_MAX, u1, u0))
    print("constexpr double ERF_SERIES[%d] = {%s};" % (len(a), ", ".join(repr(v) for v in a)))
    print("constexpr double ERFCX_CHEB[%d] = {%s};" % (len(cf), ", ".join(repr(v) for v in cf)))


if __name__ == "__main__":
    main()
