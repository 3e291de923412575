"""Generate log_coeffs.h: constants of the FP64 natural logarithm used by
the elliptic-integral kernels (spt_device.h log_pos).

Method (the atanh reduction of Cody and Waite, "Software Manual for the
Elementary Functions", 1980, ch. 5): x = 2^e m with m in [sqrt(1/2), sqrt(2)), f = m - 1,
s = f / (2 + f), so that ln m = 2 atanh(s) = 2 s + s R(s^2) with
R(z) = sum_k 2 z^k / (2k + 1), and 2 s = f - s f. Then

    ln x = e ln2_hi + (f - (hf - (s (hf + R) + e ln2_lo))),  hf = f^2 / 2,

where only the small terms carry the rounding of s. R(z) = z T(z) with T a
degree-6 polynomial fitted here (relative least squares at 50 digits, Lawson
reweighting toward minimax) over z in [0, (3 - 2 sqrt 2)^2]. ln2_hi keeps
42 significant bits so e * ln2_hi is exact for every double exponent.

A second, table-driven form (spt_device.h log_tab) serves the elliptic
integrals of the field gather, which need ln p only to an absolute
2^-53 max(1, |ln p|): K(1-p) = P(p) - ln p Q(p) and E are O(1) near p = 1.
x = 2^e m with m in [1/2, 1) (no folding), the top 8 fraction bits of m
pick c_i, the centre of one of 256 intervals of width 2^-9, and

    ln x = e ln2 + ln c_i + log1p(r),  r = m / c_i - 1 = fma(m, 1/c_i, -1),

with the table holding the double nearest 1/c_i and ln c_i := -ln(that
double) rounded, so the identity is exact up to those roundings. |r| <=
2^-9, so log1p(r) = r + r^2 (-1/2 + r/3 - r^2/4 + r^3/5) leaves a
truncation error below r^6 / 6 <= 2^-56.6 (Taylor coefficients suffice).

Run once at development time (mpmath needed); the header it writes is
committed. The script reports the worst error of both FP64 evaluations
(exactly rounded fma emulated with fractions) on random x in (0, 1].
"""
import os
import sys
from fractions import Fraction

import mpmath as mp
import numpy as np

mp.mp.dps = 50
DEG = 6


def t_target(z):
    """T(z) = (2 atanh(s) / s - 2) / z, z = s^2."""
    if z == 0:
        return mp.mpf(2) / 3
    s = mp.sqrt(z)
    return (2 * mp.atanh(s) / s - 2) / z


def fit():
    smax = (mp.sqrt(2) - 1) / (mp.sqrt(2) + 1)
    zmax = smax**2
    n = 200
    xs = [zmax * (1 - mp.cos(mp.pi * (k + mp.mpf(1) / 2) / n)) / 2 for k in range(n)]
    fx = [t_target(x) for x in xs]
    w = [mp.mpf(1)] * n
    for _ in range(6):
        A = mp.matrix(n, DEG + 1)
        b = mp.matrix(n, 1)
        for i, x in enumerate(xs):
            for j in range(DEG + 1):
                A[i, j] = x**j * w[i] / fx[i]
            b[i] = w[i]
        sol = mp.qr_solve(A, b)[0]
        c = [sol[j] for j in range(DEG + 1)]
        errs = [abs(mp.polyval(c[::-1], x) / f - 1) for x, f in zip(xs, fx)]
        top = max(errs)
        w = [wi * mp.sqrt(e / top + mp.mpf(10) ** -6) for wi, e in zip(w, errs)]
    return [float(v) for v in c]


def ln2_split():
    l2 = mp.log(2)
    hi = float(mp.floor(mp.ldexp(l2, 41)) / mp.mpf(2) ** 41)
    lo = float(l2 - mp.mpf(hi))
    return hi, lo


def fma(a, b, c):
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def log_pos(x, C, hi, lo):
    """The device formula in exactly rounded FP64 steps; the hardware
    reciprocal estimate is modelled by a float32-rounded 1/d."""
    m, e = np.frexp(x)
    m, e = float(m), int(e)
    if m < 0.70710678118654752440:
        m, e = m + m, e - 1
    f = m - 1.0
    d = 2.0 + f
    r = float(np.float32(1.0 / d))
    r = fma(r, fma(-d, r, 1.0), r)
    s = f * r
    s = fma(r, fma(-d, s, f), s)
    z = s * s
    t = C[-1]
    for c in reversed(C[:-1]):
        t = fma(t, z, c)
    R = z * t
    hf = 0.5 * (f * f)
    de = float(e)
    return fma(de, hi, -((hf - fma(s, hf + R, de * lo)) - f))


def check(C, hi, lo, n=3000):
    rng = np.random.default_rng(0)
    xs = np.concatenate([rng.uniform(0, 1, n // 3), 10 ** rng.uniform(-300, 0, n // 3),
                         1 - 10 ** rng.uniform(-16, 0, n // 3), [5e-324, 2.0**-1022, 0.5, 1.0]])
    worst = 0.0
    for x in xs:
        if x <= 0:
            continue
        ref = mp.log(mp.mpf(float(x)))
        v = log_pos(float(x), C, hi, lo)
        if ref == 0:
            assert v == 0.0
            continue
        worst = max(worst, float(abs((mp.mpf(v) - ref) / ref)))
    return worst / 2.0**-52


TAB_N = 256
TAB_Q = [-0.5, 1.0 / 3.0, -0.25, 0.2]


def table():
    """(1/c_i, ln c_i) for the centres c_i of [1/2 + i/512, 1/2 + (i+1)/512)."""
    inv, logc = [], []
    for i in range(TAB_N):
        c = mp.mpf(0.5) + (i + mp.mpf(0.5)) / (2 * TAB_N)
        iv = float(1 / c)
        inv.append(iv)
        logc.append(float(-mp.log(mp.mpf(iv))))
    return inv, logc


def log_tab(x, inv, logc, ln2):
    """spt_device.h log_tab in exactly rounded FP64 steps."""
    m, e = np.frexp(x)
    m, e = float(m), int(e)
    i = (int(np.float64(m).view(np.uint64)) >> 44) & (TAB_N - 1)
    r = fma(m, inv[i], -1.0)
    r2 = r * r
    q = TAB_Q[-1]
    for c in reversed(TAB_Q[:-1]):
        q = fma(q, r, c)
    return fma(float(e), ln2, logc[i] + fma(r2, q, r))


def check_tab(inv, logc, ln2, n=3000):
    """Worst |error| / max(1, |ln x|) in units of 2^-52, and the largest |r|."""
    rmax = max(max(abs((0.5 + i / (2 * TAB_N)) * inv[i] - 1), abs((0.5 + (i + 1) / (2 * TAB_N)) * inv[i] - 1))
               for i in range(TAB_N))
    rng = np.random.default_rng(1)
    xs = np.concatenate([rng.uniform(0, 1, n // 3), 10 ** rng.uniform(-300, 0, n // 3),
                         1 - 10 ** rng.uniform(-16, 0, n // 3), [2.0**-1022, 0.5, 1.0]])
    worst = 0.0
    for x in xs:
        ref = mp.log(mp.mpf(float(x)))
        v = log_tab(float(x), inv, logc, ln2)
        worst = max(worst, float(abs(mp.mpf(v) - ref) / max(1, abs(ref))))
    return worst / 2.0**-52, rmax


def main():
    C = fit()
    hi, lo = ln2_split()
    ulp = check(C, hi, lo)
    inv, logc = table()
    ln2 = float(mp.log(2))
    ulp_t, rmax = check_tab(inv, logc, ln2)
    assert rmax <= 2.0**-9 and log_tab(1.0, inv, logc, ln2) == 0.0
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "log_coeffs.h")
    with open(out, "w") as f:
        f.write("// Generated by gen_log_coeffs.py -- do not edit by hand.\n")
        f.write("// ln(1 + f) = 2 s + s R(s^2), R(z) = z T(z), T of degree 6; ln 2 = LN2_HI + LN2_LO.\n")
        f.write(f"// Worst FP64 error on 3000 random x in (0, 1]: {ulp:.2f} (units of 2^-52 rel).\n")
        f.write("#pragma once\n")
        f.write(f"#define SPT_LOG_T {{ {', '.join(repr(v) for v in C)} }}\n")
        f.write(f"#define SPT_LOG_LN2_HI {hi!r}\n")
        f.write(f"#define SPT_LOG_LN2_LO {lo!r}\n")
        f.write(f"// Table form: {TAB_N} intervals, |r| <= {rmax:.4g}; worst |error| / max(1, |ln x|) on 3000\n")
        f.write(f"// random x in (0, 1]: {ulp_t:.2f} x 2^-52.\n")
        f.write(f"#define SPT_LOGT_N {TAB_N}\n")
        f.write(f"#define SPT_LOGT_INVC {{ {', '.join(repr(v) for v in inv)} }}\n")
        f.write(f"#define SPT_LOGT_LOGC {{ {', '.join(repr(v) for v in logc)} }}\n")
        f.write(f"#define SPT_LOGT_Q {{ {', '.join(repr(v) for v in TAB_Q)} }}\n")
        f.write(f"#define SPT_LOG_LN2 {ln2!r}\n")
    print(f"wrote {out}: worst {ulp:.2f} x 2^-52, table form {ulp_t:.2f} x 2^-52 (abs / max(1, |ln x|))",
          file=sys.stderr)


if __name__ == "__main__":
    main()
