"""Exterior multipole expansion of an axisymmetric BEM surface charge (no
counterpart in the reference, which sums every segment at every point,
bem/axisymmetric.py:560-605).

Outside the sphere r = r_ref about the origin that holds every quadrature
node, the perturbation potential of the node rings is

    phi = (1/4 pi) sum_l m_l (r_ref / r)^l P_l(cos theta) / r,
    m_l = sum_q (area * sigma)_q (r_q / r_ref)^l P_l(z_q / r_q),

(a ring of charge Q at (r_q, theta_q) contributes Q r_q^l P_l(cos theta_q)
/ r^(l+1) P_l(cos theta) by the addition theorem, 1/4 pi being the
normalisation of the unit-charge ring kernel). The library evaluates it
(spt_bem_multipole) at points r >= r_min; fit() chooses r_min from a
validation against the device's segment sums so the field error stays
within a stated tolerance of |E|.
"""
import contextlib

import numpy as np

from .. import _lib

_GX4 = np.array([-0.8611363115940526, -0.33998104358485626, 0.33998104358485626, 0.8611363115940526])
_GW4 = np.array([0.3478548451374537, 0.6521451548625462, 0.6521451548625462, 0.3478548451374537])


def quadrature_nodes(profile, sigma, n_subdiv):
    """(rho_q, z_q, area_q * sigma_q) of the fixed rule, node for node as the
    device table (reference bem/axisymmetric.py:305-326)."""
    prof = np.asarray(profile, dtype=np.float64)
    sig = np.asarray(sigma, dtype=np.float64)
    p0, p1 = prof[:-1], prof[1:]
    L = np.linalg.norm(p1 - p0, axis=1)
    sub = np.arange(n_subdiv)
    t = (sub[:, None] / n_subdiv + 0.5 / n_subdiv * (_GX4[None, :] + 1.0)).ravel()  # (n_sub * 4,)
    w = np.tile(_GW4, n_subdiv)
    rho = p0[:, :1] + t[None] * (p1 - p0)[:, :1]
    z = p0[:, 1:] + t[None] * (p1 - p0)[:, 1:]
    s = sig[:-1, None] + t[None] * (sig[1:] - sig[:-1])[:, None]
    area = 2.0 * np.pi * rho * ((L[:, None] / n_subdiv) * 0.5 * w[None])
    return rho.ravel(), z.ravel(), (area * s).ravel()


def legendre(c, n):
    """P_0 .. P_{n-1} at c, shape (n,) + c.shape."""
    c = np.asarray(c, dtype=np.float64)
    P = np.zeros((n,) + c.shape)
    P[0] = 1.0
    if n > 1:
        P[1] = c
    for l in range(1, n - 1):
        P[l + 1] = ((2 * l + 1) * c * P[l] - l * P[l - 1]) / (l + 1)
    return P


def moments(profile, sigma, n_subdiv, n_terms):
    """(m_l for l < n_terms, r_ref) of the fixed-rule node rings."""
    rho, z, w = quadrature_nodes(profile, sigma, n_subdiv)
    r = np.hypot(rho, z)
    r_ref = float(r.max()) * (1.0 + 1e-12)
    P = legendre(z / r, n_terms)
    m = np.zeros(n_terms)
    x = w.copy()
    for l in range(n_terms):  # x = w (r / r_ref)^l, accumulated in order
        m[l] = np.sum(x * P[l])
        x = x * (r / r_ref)
    return m, r_ref


def field(m, r_ref, rho, z):
    """(E_rho, E_z) of the expansion at meridian points (host, NumPy)."""
    r = np.hypot(rho, z)
    c, s = z / r, rho / r
    n = len(m)
    P = legendre(c, n + 1)
    dP = np.zeros_like(P)
    if n > 0:
        dP[1] = 1.0
    for l in range(1, n):
        dP[l + 1] = dP[l - 1] + (2 * l + 1) * P[l]
    l = np.arange(n).reshape((n,) + (1,) * r.ndim)
    q = (r_ref / r)[None] ** l
    mm = m.reshape((n,) + (1,) * r.ndim)
    er = (mm * (l + 1) * q * P[:n]).sum(0) / (4.0 * np.pi * r * r)
    et = (mm * q * dP[:n]).sum(0) * s / (4.0 * np.pi * r * r)
    return er * s + et * c, er * c - et * s


class Multipole:
    """A validated expansion for one BEM geometry: moments, r_ref, r_min
    (inf: nowhere within tolerance, never installed) and the report."""

    def __init__(self, geometry, m, r_ref, r_min, report=None):
        self.geometry = geometry
        self.m = np.ascontiguousarray(m, dtype=np.float64)
        self.r_ref = float(r_ref)
        self.r_min = float(r_min)
        self.report = report or {}

    @property
    def inert(self):
        return not np.isfinite(self.r_min)

    def install(self, ctx, r_min=None):
        r_min = self.r_min if r_min is None else r_min
        ctx.check(ctx.lib.spt_bem_multipole(ctx.h, self.m.ctypes.data_as(_lib.ctypes.c_void_p), len(self.m),
                                            self.r_ref, float(r_min)), "spt_bem_multipole")

    @contextlib.contextmanager
    def active(self, r_min=None):
        """Within the block, the geometry's field gathers use the expansion
        at r >= r_min (default: the validated radius)."""
        ctx = self.geometry._ctx()
        self.install(ctx, r_min)
        try:
            yield
        finally:
            ctx.lib.spt_bem_multipole(ctx.h, None, 0, 0.0, 0.0)


def _pert_field(geometry, rho, z):
    import torch

    ctx = geometry._ctx()
    pts = torch.as_tensor(np.stack([rho, np.zeros_like(rho), z], -1), dtype=torch.float64,
                          device=ctx.torch_device).contiguous()
    E = torch.empty_like(pts)
    ctx.bind_stream()
    ctx.check(ctx.lib.spt_bem_pert_field(ctx.h, _lib.ptr(pts), pts.shape[0], _lib.ptr(E), _lib.PREC_FP64),
              "spt_bem_pert_field")
    ctx.sync()
    return E.cpu().numpy()


def validation_points(geometry, r_ref, n, r_max, seed):
    """Meridian points with r in [r_ref, r_max] (log-uniform), cos theta
    uniform in [0, 1] (a tenth within 0.02 rad of the axis), above the
    cathode plane and outside the conductor."""
    rng = np.random.default_rng(seed)
    th = np.arccos(rng.uniform(0.0, 1.0, n))
    th[: n // 10] = rng.uniform(0.0, 0.02, n // 10)
    r = r_ref * np.exp(rng.uniform(0.0, np.log(r_max / r_ref), n))
    rho, z = r * np.sin(th), r * np.cos(th)
    H = float(getattr(geometry, "H", 0.0))
    d = np.where(z > H, np.hypot(rho, z - H), rho) - geometry.R
    keep = (z > 0.0) & (d >= 0.0)
    return rho[keep], z[keep], r[keep]


def fit(geometry, n_terms=128, tol=1e-8, n_val=1 << 16, r_max=None, seed=13):
    """Moments of the geometry's BEM charge and the smallest r_min (from
    r_ref (1.01 .. 4)) beyond which the device expansion's field is within
    tol of |E| from the device segment sums on fresh points."""
    if geometry.bem_table() is None or geometry.spt_geometry().kind != _lib.GEOM_BEM_POST:
        raise ValueError("the exterior multipole zone is fitted for protrusions on the cathode (the post and the "
                         "hemispherical tip BEM geometries): its validation points are chosen with their exterior "
                         "test, and a well's profile runs out along the plane, leaving no exterior zone")
    prof, sig, nsub = geometry.bem_table()
    m, r_ref = moments(prof, sig, nsub, n_terms)
    r_max = 10.0 * r_ref if r_max is None else r_max
    rho, z, r = validation_points(geometry, r_ref, n_val, r_max, seed)
    ref = _pert_field(geometry, rho, z)
    mp = Multipole(geometry, m, r_ref, np.inf)
    with mp.active(r_min=r_ref * (1.0 + 1e-9)):
        got = _pert_field(geometry, rho, z)
    Et = np.hypot(ref[:, 0], ref[:, 2] + geometry.E_gun)
    err = np.hypot(got[:, 0] - ref[:, 0], got[:, 2] - ref[:, 2]) / Et
    cands = r_ref * np.geomspace(1.01, 4.0, 60)
    cands = cands[[bool(np.any(r >= c)) for c in cands]]
    worst = np.array([err[r >= c].max() for c in cands])
    r_min = np.inf
    for i in range(len(cands)):
        if np.all(worst[i:] <= tol):
            r_min = float(cands[i])
            break
    beyond = r >= r_min
    rep = dict(n_terms=n_terms, tol=tol, r_ref=r_ref, r_min=r_min, r_min_over_r_ref=r_min / r_ref, n_val=int(len(r)),
               field_err_max_beyond=float(err[beyond].max()) if beyond.any() else None,
               field_err_median_beyond=float(np.median(err[beyond])) if beyond.any() else None,
               candidates=(cands / r_ref).tolist(), worst_beyond=worst.tolist())
    mp.r_min = r_min
    mp.report = rep
    return mp
