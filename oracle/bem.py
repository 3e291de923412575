"""Oracle axisymmetric BEM (m = 0 ring kernels), FP64 NumPy/SciPy.

Restates reference ``bem/_elliptic.py``, ``bem/axisymmetric.py`` and the
evaluate/blend parts of ``bem/fields.py`` / ``bem/geometry.py``.
"""
import numpy as np
from scipy import special

_GX4, _GW4 = np.polynomial.legendre.leggauss(4)   # reference axisymmetric.py:98
_GX8, _GW8 = np.polynomial.legendre.leggauss(8)   # reference axisymmetric.py:94


def ellip_ke_complement(p):
    """K(1-p), E(1-p) (reference ``bem/_elliptic.py:28-39``)."""
    return special.ellipkm1(p), special.ellipe(1.0 - p)


def ring_potential(rho, z, a):
    """Unit-total-charge ring potential, G = 1/(4 pi r) (reference
    ``bem/axisymmetric.py:221-229``)."""
    D = (rho + a) ** 2 + z**2
    p = ((rho - a) ** 2 + z**2) / D
    K, _ = ellip_ke_complement(p)
    return K / (2.0 * np.pi**2 * np.sqrt(D))


def ring_field(rho, z, a):
    """(E_rho, E_z) of a unit-charge ring, with the near-axis Taylor branch
    (reference ``bem/axisymmetric.py:248-265``)."""
    D = (rho + a) ** 2 + z**2
    d2 = (rho - a) ** 2 + z**2
    with np.errstate(divide="ignore", invalid="ignore"):
        K, E = ellip_ke_complement(d2 / D)
        axis = rho**2 < 1e-10 * (a * a + z * z)
        rs = np.where(axis, 1.0, rho)
        er = (K - (a * a - rho * rho + z * z) * E / d2) / (4 * np.pi**2 * rs * np.sqrt(D))
        ez = z * E / (2 * np.pi**2 * np.sqrt(D) * d2)
        ad2 = a * a + z * z
        er_ax = rho * (2 * z * z - a * a) / (8 * np.pi * ad2**2.5)
        ez_ax = z / (4 * np.pi * ad2**1.5) + 3 * rho * rho * z * (3 * a * a - 2 * z * z) / (16 * np.pi * ad2**3.5)
    return np.where(axis, er_ax, er), np.where(axis, ez_ax, ez)


def n_for_length(arc_length, max_length, min_points=2):
    """Reference ``bem/mesh.py:144-150``."""
    if max_length <= 0.0:
        raise ValueError("max_length must be positive")
    return max(min_points, max(1, int(np.ceil(arc_length / max_length))) + 1)


def sphere_profile(radius, n_theta):
    """Reference ``bem/mesh.py:65-73,134-136``."""
    th = np.linspace(0.0, np.pi, n_theta)
    return np.column_stack([radius * np.sin(th), radius * np.cos(th)])


def fixed_quadrature(profile, sigma, n_subdiv):
    """Per segment, per sub-panel, Gauss-4 nodes (rho, z, area, sigma)
    (reference ``bem/axisymmetric.py:305-326``)."""
    out = [[], [], [], []]
    for j in range(len(profile) - 1):
        p0, p1 = profile[j], profile[j + 1]
        s0, s1 = sigma[j], sigma[j + 1]
        L = np.linalg.norm(p1 - p0)
        for k in range(n_subdiv):
            ta, tb = k / n_subdiv, (k + 1) / n_subdiv
            for xi, wi in zip(_GX4, _GW4):
                t = ta + 0.5 * (tb - ta) * (xi + 1.0)
                r = p0[0] + t * (p1[0] - p0[0])
                out[0].append(r)
                out[1].append(p0[1] + t * (p1[1] - p0[1]))
                out[2].append(2.0 * np.pi * r * ((L / n_subdiv) * 0.5 * wi))
                out[3].append(s0 + t * (s1 - s0))
    return tuple(np.array(v) for v in out)


def segment_geometry(profile):
    """(p0, p1, tangent, unit outward normal) per segment (reference
    ``bem/axisymmetric.py:329-338``)."""
    p0, p1 = profile[:-1], profile[1:]
    tan = p1 - p0
    nrm = np.stack([-tan[:, 1], tan[:, 0]], axis=-1)
    nrm = nrm / np.linalg.norm(nrm, axis=-1, keepdims=True)
    return p0, p1, tan, nrm


def closest_point(rho, z, p0, tan, nrm):
    """Argmin over segments of the clamped-projection distance; the first
    (lowest-index) segment wins ties (reference
    ``bem/axisymmetric.py:341-389``)."""
    M = rho.shape[0]
    bd2 = np.full(M, np.inf)
    bseg = np.zeros(M, dtype=np.int64)
    bt = np.zeros(M)
    bn = np.zeros((M, 2))
    for j in range(len(p0)):
        ab = tan[j]
        abab = float(ab[0] * ab[0] + ab[1] * ab[1])
        t = np.clip(((rho - p0[j, 0]) * ab[0] + (z - p0[j, 1]) * ab[1]) / abab, 0.0, 1.0)
        d2 = (rho - (p0[j, 0] + t * ab[0])) ** 2 + (z - (p0[j, 1] + t * ab[1])) ** 2
        better = d2 < bd2
        bd2 = np.where(better, d2, bd2)
        bseg = np.where(better, j, bseg)
        bt = np.where(better, t, bt)
        bn = np.where(better[:, None], nrm[j][None, :], bn)
    return bd2, bseg, bt, bn


def evaluate_field(profile, sigma, points, n_subdiv=8, quadrature=None, return_stats=False):
    """Perturbation field -grad S[sigma] at points (..., 3): cached far
    Gauss quadrature per segment, replaced for segments closer than their
    own length by a sinh-substituted Gauss-8 rule in intervals of width
    <= 0.5; exactly-on-surface points displaced by 1e-4 L along the normal
    (reference ``bem/axisymmetric.py:501-605``)."""
    profile = np.asarray(profile, dtype=float)
    sigma = np.asarray(sigma, dtype=float)
    if quadrature is None:
        quadrature = fixed_quadrature(profile, sigma, n_subdiv)
    rq, zq, aq, sq = quadrature
    p0, p1, tan, nrm = segment_geometry(profile)
    L = np.linalg.norm(tan, axis=-1)
    unit = tan / L[:, None]
    points = np.asarray(points, dtype=float)
    shape = points.shape
    flat = points.reshape(-1, 3)
    if not flat.shape[0]:
        return np.empty_like(points)
    rho = np.hypot(flat[:, 0], flat[:, 1])
    z = flat[:, 2].copy()
    phi = np.arctan2(flat[:, 1], flat[:, 0])

    d2, seg, _, normals = closest_point(rho, z, p0, tan, nrm)
    off = np.where(d2 < (1e-9 * L[seg]) ** 2, 1e-4 * L[seg], 0.0)
    rho_e = rho + off * normals[:, 0]
    rho_e = np.where(rho == 0.0, 0.0, rho_e)
    z = z + off * normals[:, 1]

    n_seg = len(p0)
    er, ez = ring_field(rho_e[:, None], z[:, None] - zq[None, :], rq[None, :])
    w = aq * sq
    er_seg = (er * w[None, :]).reshape(-1, n_seg, len(rq) // n_seg).sum(axis=-1)
    ez_seg = (ez * w[None, :]).reshape(-1, n_seg, len(rq) // n_seg).sum(axis=-1)

    delta = np.stack([rho_e, z], axis=-1)[:, None, :] - p0[None, :, :]
    along = np.sum(delta * unit[None, :, :], axis=-1)
    proj = np.clip(along, 0.0, L[None, :])
    dist = np.linalg.norm(delta - proj[..., None] * unit[None, :, :], axis=-1)
    pi_, si_ = np.nonzero(dist < L[None, :])
    n_near_nodes = 0
    if pi_.size:
        a = p0[si_]
        t = unit[si_]
        Ls = L[si_]
        s0 = proj[pi_, si_]
        d = np.maximum(dist[pi_, si_], 1e-15 * Ls)
        lo = np.arcsinh(-s0 / d)
        hi = np.arcsinh((Ls - s0) / d)
        count = np.maximum(1, np.ceil((hi - lo) / 0.5).astype(np.int64))
        n_near_nodes = int(count.sum()) * 8
        nint = int(count.max())
        v = lo[:, None, None] + (np.arange(nint)[None, :, None] + 0.5 * (_GX8[None, None, :] + 1.0)) * (
            (hi - lo) / count)[:, None, None]
        valid = np.arange(nint)[None, :, None] < count[:, None, None]
        v = np.where(valid, v, 0.0)
        s = s0[:, None, None] + d[:, None, None] * np.sinh(v)
        src = a[:, None, None, :] + s[..., None] * t[:, None, None, :]
        dens = sigma[si_, None, None] + (s / Ls[:, None, None]) * (sigma[si_ + 1] - sigma[si_])[:, None, None]
        wn = dens * (2 * np.pi * src[..., 0]) * d[:, None, None] * np.cosh(v) * (
            0.5 * (hi - lo) / count)[:, None, None] * _GW8[None, None, :]
        erc, ezc = ring_field(rho_e[pi_, None, None], z[pi_, None, None] - src[..., 1], src[..., 0])
        er_seg[pi_, si_] = np.sum(np.where(valid, wn * erc, 0.0), axis=(1, 2))
        ez_seg[pi_, si_] = np.sum(np.where(valid, wn * ezc, 0.0), axis=(1, 2))
    e_rho = er_seg.sum(axis=-1)
    e_z = ez_seg.sum(axis=-1)
    ex = np.where(rho == 0.0, 0.0, e_rho * np.cos(phi))
    ey = np.where(rho == 0.0, 0.0, e_rho * np.sin(phi))
    out = np.stack([ex, ey, e_z], axis=-1).reshape(shape)
    if return_stats:
        return out, dict(n_near_pairs=int(pi_.size), n_near_nodes=n_near_nodes)
    return out


def assemble(profile, refine_ratio=1.0, max_depth=20):
    """Collocation matrix of the P1 hat basis (reference
    ``bem/axisymmetric.py:275-302,411-441``): entry (i, j) is the potential
    at node i of the unit hat on node j, each segment integrated by
    adaptive bisection (split while length > refine_ratio * distance to
    its midpoint, depth <= max_depth) with Gauss-4 leaves.

    Vectorised breadth-first over all (node, segment) pairs; the leaf sum
    order differs from the reference's recursive left+right tree only in
    rounding."""
    profile = np.asarray(profile, dtype=float)
    n = len(profile)
    ii, jj = np.meshgrid(np.arange(n), np.arange(n - 1), indexing="ij")
    ii, jj = ii.ravel(), jj.ravel()
    fr, fz = profile[ii, 0], profile[ii, 1]
    a0 = profile[jj].copy()
    a1 = profile[jj + 1].copy()
    ua = np.zeros(ii.size)   # hat "lo" (1 at segment start) at a0: 1 - u
    ub = np.ones(ii.size)
    lo_acc = np.zeros(ii.size)
    hi_acc = np.zeros(ii.size)
    owner = np.arange(ii.size)
    depth = 0
    while owner.size:
        mid = 0.5 * (a0 + a1)
        length = np.linalg.norm(a1 - a0, axis=-1)
        dist = np.hypot(fr[owner] - mid[:, 0], fz[owner] - mid[:, 1])
        split = (depth < max_depth) & (length > refine_ratio * dist)
        leaf = ~split
        if leaf.any():
            o = owner[leaf]
            b0, b1, u0, u1, ln = a0[leaf], a1[leaf], ua[leaf], ub[leaf], length[leaf]
            tot_lo = np.zeros(o.size)
            tot_hi = np.zeros(o.size)
            for xi, wi in zip(_GX4, _GW4):
                t = 0.5 * (xi + 1.0)
                rs = b0[:, 0] + t * (b1[:, 0] - b0[:, 0])
                zs = b0[:, 1] + t * (b1[:, 1] - b0[:, 1])
                hi_s = u0 + t * (u1 - u0)
                lo_s = (1.0 - u0) + t * ((1.0 - u1) - (1.0 - u0))
                g = 2.0 * np.pi * rs * ring_potential(fr[o], fz[o] - zs, rs) * (ln * 0.5 * wi)
                tot_lo += lo_s * g
                tot_hi += hi_s * g
            np.add.at(lo_acc, o, tot_lo)
            np.add.at(hi_acc, o, tot_hi)
        if split.any():
            pm = mid[split]
            um = 0.5 * (ua[split] + ub[split])
            a0 = np.concatenate([a0[split], pm])
            a1 = np.concatenate([pm, a1[split]])
            ua = np.concatenate([ua[split], um])
            ub = np.concatenate([um, ub[split]])
            owner = np.concatenate([owner[split], owner[split]])
        else:
            owner = owner[:0]
        depth += 1
    lo_m = lo_acc.reshape(n, n - 1)
    hi_m = hi_acc.reshape(n, n - 1)
    A = np.zeros((n, n))
    A[:, :-1] += lo_m
    A[:, 1:] += hi_m
    return A


def potential(profile, sigma, points, refine_ratio=1.0, max_depth=20):
    """Perturbation potential at `points` (..., 3) (reference
    ``bem/axisymmetric.py:445-462``): per segment, adaptive bisection
    (split while length > refine_ratio * distance to the midpoint, depth <=
    max_depth) with Gauss-4 leaves and sigma linear along the segment
    (reference ``bem/axisymmetric.py:275-302``).

    Vectorised breadth-first over all (point, segment) pairs, like
    ``assemble``; the leaf sum order differs from the reference's recursion
    only in rounding."""
    profile = np.asarray(profile, dtype=float)
    sigma = np.asarray(sigma, dtype=float)
    pts = np.asarray(points, dtype=float)
    shape = pts.shape[:-1]
    flat = pts.reshape(-1, 3)
    rho_f = np.hypot(flat[:, 0], flat[:, 1])
    z_f = flat[:, 2]
    n = len(profile)
    ii, jj = np.meshgrid(np.arange(flat.shape[0]), np.arange(n - 1), indexing="ij")
    ii, jj = ii.ravel(), jj.ravel()
    fr, fz = rho_f[ii], z_f[ii]
    a0 = profile[jj].copy()
    a1 = profile[jj + 1].copy()
    s0 = sigma[jj].copy()
    s1 = sigma[jj + 1].copy()
    acc = np.zeros(ii.size)
    owner = np.arange(ii.size)
    depth = 0
    while owner.size:
        mid = 0.5 * (a0 + a1)
        length = np.linalg.norm(a1 - a0, axis=-1)
        dist = np.hypot(fr[owner] - mid[:, 0], fz[owner] - mid[:, 1])
        split = (depth < max_depth) & (length > refine_ratio * dist)
        leaf = ~split
        if leaf.any():
            o = owner[leaf]
            b0, b1, t0, t1, ln = a0[leaf], a1[leaf], s0[leaf], s1[leaf], length[leaf]
            tot = np.zeros(o.size)
            for xi, wi in zip(_GX4, _GW4):
                t = 0.5 * (xi + 1.0)
                rs = b0[:, 0] + t * (b1[:, 0] - b0[:, 0])
                zs = b0[:, 1] + t * (b1[:, 1] - b0[:, 1])
                ss = t0 + t * (t1 - t0)
                tot += ss * 2.0 * np.pi * rs * ring_potential(fr[o], fz[o] - zs, rs) * (ln * 0.5 * wi)
            np.add.at(acc, o, tot)
        if split.any():
            pm = mid[split]
            sm = 0.5 * (s0[split] + s1[split])
            a0 = np.concatenate([a0[split], pm])
            a1 = np.concatenate([pm, a1[split]])
            s0, s1 = np.concatenate([s0[split], sm]), np.concatenate([sm, s1[split]])
            owner = np.concatenate([owner[split], owner[split]])
        else:
            owner = owner[:0]
        depth += 1
    return acc.reshape(flat.shape[0], n - 1).sum(axis=1).reshape(shape)


def solve(profile, dirichlet, refine_ratio=1.0, max_depth=20):
    """sigma with A sigma = g at every node (reference
    ``bem/axisymmetric.py:411-443``); ``dirichlet(rho, z)`` vectorised."""
    profile = np.asarray(profile, dtype=float)
    A = assemble(profile, refine_ratio, max_depth)
    b = dirichlet(profile[:, 0], profile[:, 1])
    return np.linalg.solve(A, b)


def tip_bem_field(pos, Ez, R, profile, sigma, n_subdiv=8, z_handoff=None):
    """HemisphericalTipBEMGeometry.field: uniform + perturbation, zero
    inside (r < R(1-1e-6) or z < 0), blended to the flat-cathode field
    between z_handoff/2 and z_handoff (reference ``bem/fields.py:186-211``,
    ``bem/geometry.py:319-324``)."""
    z_handoff = 20.0 * R if z_handoff is None else z_handoff
    r = np.linalg.norm(pos, axis=-1)
    valid = (r >= R * (1.0 - 1e-6)) & (pos[..., 2] >= 0.0)
    near = np.zeros_like(pos)
    near[..., 2] = Ez
    near = near + evaluate_field(profile, sigma, pos, n_subdiv)
    near = np.where(valid[..., None], near, 0.0)
    from .fields import gun_field, blend
    return blend(near, gun_field(pos, Ez), pos[..., 2], 0.5 * z_handoff, z_handoff)


def post_axis_distance(pos, H):
    """Distance from a point to the axis segment {(0, 0, s): |s| <= H}:
    the capsule (post of half-length H plus hemispherical caps) interior is
    where this is below R. H = 0 is the sphere about the origin."""
    zc = np.clip(pos[..., 2], -H, H)
    return np.sqrt(pos[..., 0] ** 2 + pos[..., 1] ** 2 + (pos[..., 2] - zc) ** 2)


def post_bem_field(pos, Ez, R, H, profile, sigma, z_handoff, n_subdiv=8):
    """Field of a hemisphere-capped post of radius R and height H on the
    grounded plane (solved on its mirror-doubled capsule, the same mirror
    argument as reference ``bem/fields.py:22-33``), with the reference's
    mask/blend conventions applied to the capsule distance."""
    r = post_axis_distance(pos, H)
    valid = (r >= R * (1.0 - 1e-6)) & (pos[..., 2] >= 0.0)
    near = np.zeros_like(pos)
    near[..., 2] = Ez
    near = near + evaluate_field(profile, sigma, pos, n_subdiv)
    near = np.where(valid[..., None], near, 0.0)
    from .fields import gun_field, blend
    return blend(near, gun_field(pos, Ez), pos[..., 2], 0.5 * z_handoff, z_handoff)


def well_contains(rho, z, R, H, bottom_fillet=0.0, rim_fillet=0.0):
    """Inside (or on) the real cylindrical-well conductor (reference
    ``bem/fields.py:89-135``): below the plane outside the well radius,
    below the bottom inside it, the fillet regions decided by the distance
    to the fillet centre (biased 1e-6 R towards inside)."""
    eps = 1e-6 * R
    outside = rho >= R * (1.0 - 1e-6)
    m = np.where(outside, z <= 0.0, z <= -H * (1.0 + 1e-6))
    if rim_fillet > 0.0:
        reg = (rho >= R - eps) & (rho <= R + rim_fillet + eps) & (z >= -rim_fillet - eps) & (z <= eps)
        m = np.where(reg, np.hypot(rho - (R + rim_fillet), z + rim_fillet) <= rim_fillet, m)
    if bottom_fillet > 0.0:
        reg = (rho <= R + eps) & (rho >= R - bottom_fillet - eps) & (z <= -H + bottom_fillet + eps) & (z >= -H - eps)
        m = np.where(reg, np.hypot(rho - (R - bottom_fillet), z - (-H + bottom_fillet)) >= bottom_fillet, m)
    return m


def well_bem_field(pos, Ez, R, H, profile, sigma, z_handoff, bottom_fillet=0.0, rim_fillet=0.0, n_subdiv=8):
    """CylindricalWellBEMGeometry.field (reference ``bem/fields.py:324-352``,
    ``bem/geometry.py:779-784``): uniform + perturbation, zero inside the
    conductor, blended to the flat-cathode field over [z_handoff/2,
    z_handoff]."""
    rho = np.hypot(pos[..., 0], pos[..., 1])
    valid = ~well_contains(rho, pos[..., 2], R, H, bottom_fillet, rim_fillet)
    near = np.zeros_like(pos)
    near[..., 2] = Ez
    near = near + evaluate_field(profile, sigma, pos, n_subdiv)
    near = np.where(valid[..., None], near, 0.0)
    from .fields import gun_field, blend
    return blend(near, gun_field(pos, Ez), pos[..., 2], 0.5 * z_handoff, z_handoff)
