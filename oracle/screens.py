"""Oracle screen-crossing detector (reference ``screens.py:43-168``).

Events are returned in the reference's discovery order within one call:
(screen, interval, group row, particle)."""
import numpy as np

BISECTION_ITERATIONS = 60   # reference screens.py:15


def hermite_basis(s):
    s2 = s * s
    s3 = s2 * s
    return 2 * s3 - 3 * s2 + 1, s3 - 2 * s2 + s, -2 * s3 + 3 * s2, s3 - s2


def hermite_basis_d(s):
    s2 = s * s
    return 6 * s2 - 6 * s, 3 * s2 - 4 * s + 1, -6 * s2 + 6 * s, 3 * s2 - 2 * s


def hermite_z(p0, v0, p1, v1, dt, s):
    h00, h10, h01, h11 = hermite_basis(s)
    return h00 * p0 + h10 * dt * v0 + h01 * p1 + h11 * dt * v1


def hermite(p0, v0, p1, v1, dt, s):
    h00, h10, h01, h11 = hermite_basis(s)
    return h00[:, None] * p0 + (h10 * dt)[:, None] * v0 + h01[:, None] * p1 + (h11 * dt)[:, None] * v1


def hermite_d(p0, v0, p1, v1, dt, s):
    d00, d10, d01, d11 = hermite_basis_d(s)
    return (d00 / dt)[:, None] * p0 + d10[:, None] * v0 + (d01 / dt)[:, None] * p1 + d11[:, None] * v1


def bisect(p0z, v0z, p1z, v1z, dt, target, lo, hi):
    lo = lo.copy()
    hi = hi.copy()
    flo = hermite_z(p0z, v0z, p1z, v1z, dt, lo) - target
    for _ in range(BISECTION_ITERATIONS):
        mid = 0.5 * (lo + hi)
        fm = hermite_z(p0z, v0z, p1z, v1z, dt, mid) - target
        same = np.sign(fm) == np.sign(flo)
        lo = np.where(same, mid, lo)
        flo = np.where(same, fm, flo)
        hi = np.where(same, hi, mid)
    return 0.5 * (lo + hi)


def find_crossings(screens, t_old, pos_old, vel_old, t_new, pos_new, vel_new):
    """Cubic Hermite z(s) over each accepted step, split at its real
    extrema, sign-change test on the half-open (t_old, t_new] with moving
    particles only, 60-step bisection (reference ``screens.py:43-115``).

    Returns a list of dicts (screen, interval, row, particle, t, pos, vel)."""
    k, n, _ = pos_old.shape
    dt = np.broadcast_to((t_new - t_old)[:, None], (k, n))
    t0 = np.broadcast_to(t_old[:, None], (k, n))
    z0, z1 = pos_old[..., 2], pos_new[..., 2]
    u0, u1 = dt * vel_old[..., 2], dt * vel_new[..., 2]
    A = 3 * (2 * (z0 - z1) + u0 + u1)
    B = 2 * (3 * (z1 - z0) - 2 * u0 - u1)
    C = u0
    sc = np.maximum(np.maximum(np.abs(A), np.abs(B)), np.abs(C))
    sc = np.where(sc > 0.0, sc, 1.0)
    A, B, C = A / sc, B / sc, C / sc
    disc = B * B - 4 * A * C
    q = -0.5 * (B + np.where(B >= 0.0, 1.0, -1.0) * np.sqrt(np.maximum(disc, 0.0)))
    with np.errstate(divide="ignore", invalid="ignore"):
        r0 = q / np.where(A != 0.0, A, 1.0)
        r1 = C / np.where(q != 0.0, q, 1.0)
        lin = -C / np.where(B != 0.0, B, 1.0)
    r0 = np.where(A == 0.0, lin, r0)
    ok0 = (disc >= 0.0) & ((A != 0.0) | (B != 0.0)) & (r0 > 0.0) & (r0 < 1.0)
    ok1 = (disc >= 0.0) & (A != 0.0) & (q != 0.0) & (r1 > 0.0) & (r1 < 1.0)
    cuts = np.sort(np.stack([np.zeros_like(r0), np.where(ok0, r0, 1.0), np.where(ok1, r1, 1.0),
                             np.ones_like(r0)], axis=-1), axis=-1)
    moving = np.any(pos_old != pos_new, axis=-1) | np.any(vel_old != vel_new, axis=-1)
    events = []
    for si, zs in enumerate(screens):
        for iv in range(3):
            left, right = cuts[..., iv], cuts[..., iv + 1]
            f0 = hermite_z(z0, vel_old[..., 2], z1, vel_new[..., 2], dt, left) - zs
            f1 = hermite_z(z0, vel_old[..., 2], z1, vel_new[..., 2], dt, right) - zs
            hit = moving & (right > left) & (dt > 0.0) & (f0 != 0.0) & (np.sign(f0) != np.sign(f1))
            gi, pi = np.nonzero(hit)
            if not gi.size:
                continue
            p0, v0 = pos_old[gi, pi], vel_old[gi, pi]
            p1, v1 = pos_new[gi, pi], vel_new[gi, pi]
            dtc, t0c = dt[gi, pi], t0[gi, pi]
            ss = bisect(p0[:, 2], v0[:, 2], p1[:, 2], v1[:, 2], dtc, zs, left[gi, pi], right[gi, pi])
            P = hermite(p0, v0, p1, v1, dtc, ss)
            V = hermite_d(p0, v0, p1, v1, dtc, ss)
            for e in range(gi.size):
                events.append(dict(screen=si, interval=iv, row=int(gi[e]), particle=int(pi[e]),
                                   t=float(t0c[e] + ss[e] * dtc[e]), pos=P[e].copy(), vel=V[e].copy()))
    return events
