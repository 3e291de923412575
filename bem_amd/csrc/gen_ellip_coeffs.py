"""Generate ellip_coeffs.h: FP64 coefficients for the complete elliptic
integrals of the complementary parameter, K(1-p) and E(1-p), p in [0, 1].

Form (from the logarithmic expansions DLMF 19.12.1 and 19.12.2, with
k'^2 = p):

    K(1-p) = PK(p) - ln(p) * QK(p)
    E(1-p) = PE(p) - ln(p) * p * QE(p)

PK, QK, PE, QE are degree-10 polynomials for FP64. A second, degree-6 pair
PK32/QK32 serves the FP32 (mixed-precision) field path. They are fitted jointly by
weighted (relative-error) linear least squares at 50 significant digits,
with two Lawson reweighting passes toward the minimax solution. The
target values come from the DLMF 19.12 series themselves for p < 1/4
(mpmath's own ellipk(1 - p) loses digits there because 1 - p is rounded
at working precision) and from mpmath ellipk/ellipe otherwise.

Run once at development time (mpmath needed); the header it writes is
committed. The script also reports the worst FP64 Horner error it sees on
3000 random p values, in units of 2^-52 relative.
"""
import os
import sys

import mpmath as mp
import numpy as np

mp.mp.dps = 50
DEG = 10
HALF = mp.mpf(1) / 2


def k_of_complement(p):
    """K(1-p) (DLMF 19.12.1 series for small p, else mpmath)."""
    if p < 0.25:
        s = mp.mpf(0)
        for m in range(140):
            a = (mp.rf(HALF, m) / mp.factorial(m)) ** 2
            s += a * p**m * (-mp.log(p) / 2 + mp.digamma(1 + m) - mp.digamma(HALF + m))
        return s
    return mp.ellipk(1 - p)


def e_of_complement(p):
    """E(1-p) (DLMF 19.12.2 series for small p, else mpmath)."""
    if p < 0.25:
        s = mp.mpf(0)
        for m in range(140):
            c = mp.rf(HALF, m) * mp.rf(3 * HALF, m) / (mp.rf(2, m) * mp.factorial(m))
            s += c * p ** (m + 1) * (
                -mp.log(p) / 2 + mp.digamma(1 + m) - mp.digamma(HALF + m) - mp.mpf(1) / ((2 * m + 1) * (2 * m + 2))
            )
        return 1 + s / 2
    return mp.ellipe(1 - p)


def fit_points():
    n = 300
    xs = [(1 - mp.cos(mp.pi * (k + HALF) / n)) / 2 for k in range(n)]
    xs += [mp.mpf(10) ** (-e) for e in np.linspace(3, 30, 40)]
    return xs


def joint_fit(fun, q_shift, xs, weights, deg=None):
    DEG = deg if deg is not None else globals()["DEG"]
    m, n = len(xs), 2 * (DEG + 1)
    A = mp.matrix(m, n)
    b = mp.matrix(m, 1)
    for i, x in enumerate(xs):
        f = fun(x)
        w = weights[i] / f
        lx = mp.log(x)
        for j in range(DEG + 1):
            A[i, j] = x**j * w
            A[i, DEG + 1 + j] = -lx * x ** (j + q_shift) * w
        b[i] = f * w
    sol = mp.qr_solve(A, b)[0]
    return [sol[j] for j in range(DEG + 1)], [sol[DEG + 1 + j] for j in range(DEG + 1)]


def fit(fun, q_shift, deg=None):
    xs = fit_points()
    fx = [fun(x) for x in xs]
    weights = [mp.mpf(1)] * len(xs)
    for _ in range(3):
        P, Q = joint_fit(fun, q_shift, xs, weights, deg)
        errs = []
        for x, f in zip(xs, fx):
            v = mp.polyval(P[::-1], x) - mp.log(x) * x**q_shift * mp.polyval(Q[::-1], x)
            errs.append(abs(v / f - 1))
        top = max(errs)
        weights = [w * mp.sqrt(e / top + mp.mpf(10) ** -6) for w, e in zip(weights, errs)]
    return [float(c) for c in P], [float(c) for c in Q]


def horner(c, x):
    s = 0.0
    for a in reversed(c):
        s = s * x + a
    return s


def check(P, Q, q_shift, fun):
    rng = np.random.default_rng(0)
    ps = np.concatenate([10 ** rng.uniform(-30, 0, 1500), rng.uniform(0, 1, 1500), [1.0, 0.5]])
    worst = 0.0
    for p in ps:
        v = horner(P, p) - np.log(p) * (p if q_shift else 1.0) * horner(Q, p)
        t = fun(mp.mpf(p))
        worst = max(worst, float(abs(v / t - 1)))
    return worst / 2.0**-52


def check32(P, Q):
    """Worst K and E (Legendre-derivative form) error of the FP32 pair,
    evaluated in float32, in units of 2^-23 relative."""
    P32, Q32 = np.float32(P), np.float32(Q)

    def hd(c, x):
        s, d = np.float32(c[-1]), np.float32(0)
        for a in reversed(c[:-1]):
            d = np.float32(d * x + s)
            s = np.float32(s * x + a)
        return s, d

    rng = np.random.default_rng(0)
    ps = np.concatenate([10 ** rng.uniform(-30, 0, 800), rng.uniform(0, 1, 800), [1.0]]).astype(np.float32)
    wk = we = 0.0
    for p in ps:
        Pv, dP = hd(P32, p)
        Qv, dQ = hd(Q32, p)
        lp = np.float32(np.log(p))
        K = np.float32(Pv - lp * Qv)
        pkp = np.float32(p * dP - Qv - (p * lp) * dQ)
        E = np.float32(p * K - np.float32(2) * (np.float32(1) - p) * pkp)
        wk = max(wk, float(abs(K / k_of_complement(mp.mpf(float(p))) - 1)))
        we = max(we, float(abs(E / e_of_complement(mp.mpf(float(p))) - 1)))
    return wk / 2.0**-23, we / 2.0**-23


def main():
    PK32, QK32 = fit(k_of_complement, 0, deg=6)
    ulp32 = check32(PK32, QK32)
    PK, QK = fit(k_of_complement, 0)
    PE, QE = fit(e_of_complement, 1)
    ulp_k = check(PK, QK, 0, k_of_complement)
    ulp_e = check(PE, QE, 1, e_of_complement)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ellip_coeffs.h")
    with open(out, "w") as f:
        f.write("// Generated by gen_ellip_coeffs.py -- do not edit by hand.\n")
        f.write("// K(1-p) = PK(p) - ln(p) QK(p);  E(1-p) = PE(p) - ln(p) p QE(p);  p in [0, 1].\n")
        f.write(f"// Worst FP64 Horner error on 3000 random p: K {ulp_k:.2f}, E {ulp_e:.2f} (units of 2^-52 rel).\n")
        f.write("#pragma once\n")
        for name, c in (("PK", PK), ("QK", QK), ("PE", PE), ("QE", QE)):
            body = ", ".join(repr(x) for x in c)
            f.write(f"#define SPT_ELLIP_{name} {{ {body} }}\n")
        f.write(f"// FP32 pair (degree 6), float32 evaluation: K {ulp32[0]:.2f}, E via dK/dp {ulp32[1]:.2f} "
                "(units of 2^-23 rel).\n")
        for name, c in (("PK32", PK32), ("QK32", QK32)):
            body = ", ".join(repr(float(np.float32(x))) + "f" for x in c)
            f.write(f"#define SPT_ELLIP_{name} {{ {body} }}\n")
    print(f"wrote {out}: K {ulp_k:.2f} ulp, E {ulp_e:.2f} ulp", file=sys.stderr)


if __name__ == "__main__":
    main()
