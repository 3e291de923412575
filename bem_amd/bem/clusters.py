"""Axis-centred exterior expansions of runs of consecutive BEM segments
(no counterpart in the reference, which sums every segment at every point,
bem/axisymmetric.py:560-605): a level between the global exterior
multipole zone (bem_amd.bem.multipole) and the per-segment far rule.

Every source ring of an axisymmetric surface is centred on the axis, so a
run of segments has an exact exterior expansion about any axis point
(0, z_c): outside the sphere of radius r_c about it that holds all of the
run's quadrature nodes,

    phi = (1/4 pi) sum_l m_l (r_c / r)^l P_l(cos theta) / r,
    m_l = sum_q (area * sigma)_q (r_q / r_c)^l P_l(cos theta_q),

with r, theta measured from (0, z_c) (the addition theorem for a ring).
A point at distance r >= kappa r_c from a cluster's centre takes its
field from the expansion, truncated where the tail bound
sum_{l > N} (l + 1) q^l with q = r_c / r falls below tol; closer clusters
keep the far rule segment by segment. The nodes are the fixed rule's
(8 sub-panels x Gauss-4), so the expansion converges to exactly what the
plain method sums.
"""
import ctypes

import numpy as np

from .multipole import legendre, quadrature_nodes


class Clusters:
    """Runs of segments seg_begin[c] .. seg_begin[c + 1] - 1 with centre
    z_c, radius r_c and moments (n_clusters, n_terms); `geometry` is the
    BEM geometry whose charge they expand (None for a bare profile)."""

    def __init__(self, seg_begin, zc, rc, moments, kappa, tol, geometry=None):
        self.seg_begin = np.ascontiguousarray(seg_begin, dtype=np.int32)
        self.zc = np.ascontiguousarray(zc, dtype=np.float64)
        self.rc = np.ascontiguousarray(rc, dtype=np.float64)
        self.moments = np.ascontiguousarray(moments, dtype=np.float64)
        self.kappa = float(kappa)
        self.tol = float(tol)
        self.geometry = geometry

    def install(self, ctx):
        """spt_bem_clusters on the context (used by far-rule gathers)."""
        P = ctypes.c_void_p
        ctx.check(ctx.lib.spt_bem_clusters(ctx.h, self.seg_begin.ctypes.data_as(P), self.n_clusters,
                                           self.zc.ctypes.data_as(P), self.rc.ctypes.data_as(P),
                                           self.moments.ctypes.data_as(P), self.n_terms, self.kappa, self.tol),
                  "spt_bem_clusters")

    @staticmethod
    def uninstall(ctx):
        ctx.lib.spt_bem_clusters(ctx.h, None, 0, None, None, None, 0, 0.0, 0.0)

    @property
    def n_clusters(self):
        return len(self.zc)

    @property
    def n_terms(self):
        return self.moments.shape[1]


def _regions(prof):
    """Runs of consecutive segments that are all straight-continued (the
    same direction as their predecessor, to 1e-9 rad) or all turning:
    (start, end) pairs covering the profile."""
    t = np.diff(prof, axis=0)
    ang = np.arctan2(t[:, 1], t[:, 0])
    turn = np.abs(np.angle(np.exp(1j * np.diff(ang)))) > 1e-9  # segment j + 1 turns from segment j
    kind = np.concatenate([[turn[0] if len(turn) else False], turn])
    out, a = [], 0
    for j in range(1, len(kind)):
        if kind[j] != kind[a]:
            out.append((a, j))
            a = j
    out.append((a, len(kind)))
    return out


def boundaries(profile, seg_per_cluster, split_regions=True):
    """Cluster boundaries: runs of seg_per_cluster segments (a short last run
    joins the one before), within each straight or turning region of the
    profile when split_regions (so the hemispherical cap of a post is a
    cluster of its own, whose enclosing sphere is the cap's)."""
    prof = np.asarray(profile, dtype=np.float64)
    n_seg = len(prof) - 1
    regions = _regions(prof) if split_regions else [(0, n_seg)]
    bounds = [0]
    for a, b in regions:
        cuts = list(range(a, b, seg_per_cluster)) + [b]
        if len(cuts) > 2 and cuts[-1] - cuts[-2] < seg_per_cluster // 2:
            cuts.pop(-2)  # no runt cluster at the end of a region
        bounds += cuts[1:]
    return bounds


def enclosing_centre(r, z):
    """The axis point (0, z_c) whose sphere through the farthest of the
    points (r, z) is smallest: minimises the convex max_q r_q^2 + (z_q -
    z_c)^2 by ternary search."""
    lo, hi = float(z.min()), float(z.max())
    for _ in range(200):
        m1, m2 = lo + (hi - lo) / 3.0, hi - (hi - lo) / 3.0
        if np.max(r * r + (z - m1) ** 2) < np.max(r * r + (z - m2) ** 2):
            hi = m2
        else:
            lo = m1
    return 0.5 * (lo + hi)


def build(profile, sigma, n_subdiv=8, seg_per_cluster=25, n_terms=128, kappa=1.6, tol=1e-11, split_regions=True):
    """Clusters of the fixed-rule node table (see boundaries()), each about
    the axis point of its smallest enclosing sphere."""
    prof = np.asarray(profile, dtype=np.float64)
    n_seg = len(prof) - 1
    if n_seg < 1 or seg_per_cluster < 1:
        raise ValueError("need at least one segment per cluster")
    if not kappa > 1.0:
        raise ValueError("kappa must exceed 1 (the expansion converges outside r_c)")
    rho, z, w = quadrature_nodes(prof, sigma, n_subdiv)
    nq = 4 * n_subdiv
    bounds = boundaries(prof, seg_per_cluster, split_regions)
    zc, rc, mom = [], [], []
    for a, b in zip(bounds[:-1], bounds[1:]):
        sl = slice(a * nq, b * nq)
        r_, z_, w_ = rho[sl], z[sl], w[sl]
        c = enclosing_centre(r_, z_)
        d = np.hypot(r_, z_ - c)
        R = float(d.max()) * (1.0 + 1e-12)
        cos = np.divide(z_ - c, d, out=np.ones_like(d), where=d > 0)
        P = legendre(cos, n_terms)
        m = np.zeros(n_terms)
        x = w_.copy()
        for l in range(n_terms):  # x = w (d / R)^l
            m[l] = np.sum(x * P[l])
            x = x * (d / R)
        zc.append(c)
        rc.append(R)
        mom.append(m)
    return Clusters(bounds, zc, rc, np.array(mom), kappa, tol)


def fit(geometry, seg_per_cluster=25, n_terms=128, kappa=1.6, tol=1e-11, split_regions=True):
    """Clusters of a BEM geometry's uploaded table (its profile, sigma and
    n_subdiv)."""
    prof, sig, nsub = geometry.bem_table()
    cl = build(prof, sig, nsub, seg_per_cluster, n_terms, kappa, tol, split_regions)
    cl.geometry = geometry
    return cl


def terms_needed(q, tol, n_terms):
    """Smallest N with sum_{l >= N} (l + 1) q^l <= tol (tail of the
    expansion's field relative to its scale), at most n_terms."""
    q = np.asarray(q, dtype=np.float64)
    n = np.zeros(q.shape, dtype=np.int64)
    tail = 1.0 / (1.0 - q) ** 2  # sum_{l >= 0} (l + 1) q^l
    term = np.ones_like(q)
    for l in range(n_terms):
        mask = tail > tol
        n = np.where(mask, l + 1, n)
        tail = tail - term
        term = term * q * (l + 2) / (l + 1)
    return np.minimum(n, n_terms)


def field(cl, c, rho, z):
    """(E_rho, E_z) of cluster c's expansion at meridian points (host,
    NumPy), with the point-dependent truncation."""
    dz = z - cl.zc[c]
    r = np.hypot(rho, dz)
    q = cl.rc[c] / r
    N = terms_needed(q, cl.tol, cl.n_terms)
    cs, sn = dz / r, rho / r
    n = cl.n_terms
    P = legendre(cs, n + 1)
    dP = np.zeros_like(P)
    if n > 0:
        dP[1] = 1.0
    for l in range(1, n):
        dP[l + 1] = dP[l - 1] + (2 * l + 1) * P[l]
    l = np.arange(n)[:, None]
    keep = l < N[None, :]
    qp = q[None, :] ** l
    m = cl.moments[c][:, None]
    er = np.sum(np.where(keep, m * (l + 1) * qp * P[:n], 0.0), 0) / (4 * np.pi * r * r)
    et = np.sum(np.where(keep, m * qp * dP[:n], 0.0), 0) * sn / (4 * np.pi * r * r)
    return er * sn + et * cs, er * cs - et * sn
