"""Oracle image-charge BEM (azimuthal-mode single-layer solve with a local
mirror charge), FP64 NumPy. Restates reference ``bem/image_charge.py``:
mirror points and weights (333-376), per-mode operator assembly by
adaptive bisection (379-409, 469-493), the per-group joint solve and the
force at the particles (816-1056, 1286-1395). Test infrastructure only.

Pinned by the reference's own values (tests/golden/bem_well_image.npz:
operators, image forces of particle batches) and by mathematics: the
induced force does not depend on the mirror-charge blend once enough
modes are kept (reference module docstring, "Exactness"), and a grounded
sphere reproduces Kelvin's image force.
"""
import numpy as np

from .bem import closest_point, segment_geometry
from .toroidal import point_charge_potential_modes, ring_field_modes, toroidal_Q, chi_of

_GX4, _GW4 = np.polynomial.legendre.leggauss(4)
EPS0 = 8.8541878188e-12


def ring_potential_modes(rho, z, a, za, n_max):
    q, _ = toroidal_Q(chi_of(rho, z, a, za), n_max)
    return np.sqrt(a / rho)[None, ...] * q / (2.0 * np.pi)


def rho_floor(profile):
    return 1e-6 * max(float(np.max(profile[:, 0])), 1e-300)


def assemble_mode_operators(profile, n_max, refine_ratio=1.0, max_depth=20):
    """A[m, i, j]: mode-m potential at node i of the unit hat on node j;
    collocation at rho >= 1e-6 max(rho) (reference ``image_charge.py:469-493``).
    Vectorised breadth-first over (node, segment) pairs like
    ``oracle.bem.assemble``; leaf sums differ from the reference's recursion
    only in rounding."""
    profile = np.asarray(profile, dtype=float)
    n = len(profile)
    fl = rho_floor(profile)
    ii, jj = np.meshgrid(np.arange(n), np.arange(n - 1), indexing="ij")
    ii, jj = ii.ravel(), jj.ravel()
    fr = np.maximum(profile[ii, 0], fl)
    fz = profile[ii, 1]
    a0, a1 = profile[jj].copy(), profile[jj + 1].copy()
    ua, ub = np.zeros(ii.size), np.ones(ii.size)
    lo_acc = np.zeros((n_max + 1, ii.size))
    hi_acc = np.zeros((n_max + 1, ii.size))
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
            tl = np.zeros((n_max + 1, o.size))
            th = np.zeros((n_max + 1, o.size))
            for xi, wi in zip(_GX4, _GW4):
                t = 0.5 * (xi + 1.0)
                rs = b0[:, 0] + t * (b1[:, 0] - b0[:, 0])
                zs = b0[:, 1] + t * (b1[:, 1] - b0[:, 1])
                hs = u0 + t * (u1 - u0)
                ls = (1.0 - u0) + t * ((1.0 - u1) - (1.0 - u0))
                Phi = ring_potential_modes(fr[o], fz[o], rs, zs, n_max) * (ln * 0.5 * wi)
                tl += ls * Phi
                th += hs * Phi
            np.add.at(lo_acc, (slice(None), o), tl)
            np.add.at(hi_acc, (slice(None), o), th)
        if split.any():
            pm = mid[split]
            um = 0.5 * (ua[split] + ub[split])
            a0, a1 = np.concatenate([a0[split], pm]), np.concatenate([pm, a1[split]])
            ua, ub = np.concatenate([ua[split], um]), np.concatenate([um, ub[split]])
            owner = np.concatenate([owner[split], owner[split]])
        else:
            owner = owner[:0]
        depth += 1
    A = np.zeros((n_max + 1, n, n))
    A[:, :, :-1] += lo_acc.reshape(n_max + 1, n, n - 1)
    A[:, :, 1:] += hi_acc.reshape(n_max + 1, n, n - 1)
    return A


def image_weight(d, d_lo, d_hi):
    """1 at d <= d_lo, 0 at d >= d_hi, cubic smoothstep between."""
    t = np.clip((d_hi - d) / (d_hi - d_lo), 0.0, 1.0)
    return t * t * (3.0 - 2.0 * t)


def mirror_points(rho, z, closest, tangent_unit, normal_unit):
    """Reflection through the tangent line at the closest point; the side
    is -1 when the image crosses the axis (reference ``image_charge.py:364-376``)."""
    delta = np.stack([rho, z], axis=-1) - closest
    d_t = np.sum(delta * tangent_unit, axis=-1)
    d_n = np.sum(delta * normal_unit, axis=-1)
    img = closest + d_t[:, None] * tangent_unit - d_n[:, None] * normal_unit
    side = np.where(img[:, 0] < 0.0, -1.0, 1.0)
    return np.abs(img[:, 0]), img[:, 1], side


def quadrature_geometry(profile, n_subdiv):
    """Per segment, sub-panel, Gauss-4 node: rho, z, length weight,
    segment index, local t (reference ``image_charge.py:412-451``)."""
    rq, zq, wq, sq, tq = [], [], [], [], []
    for j in range(len(profile) - 1):
        p0, p1 = profile[j], profile[j + 1]
        L = np.linalg.norm(p1 - p0)
        for k in range(n_subdiv):
            ta, tb = k / n_subdiv, (k + 1) / n_subdiv
            for xi, wi in zip(_GX4, _GW4):
                t = ta + 0.5 * (tb - ta) * (xi + 1.0)
                rq.append(p0[0] + t * (p1[0] - p0[0]))
                zq.append(p0[1] + t * (p1[1] - p0[1]))
                wq.append((L / n_subdiv) * 0.5 * wi)
                sq.append(j)
                tq.append(t)
    return np.array(rq), np.array(zq), np.array(wq), np.array(sq), np.array(tq)


def image_force(profile, A_inv, n_max, position, charge, active, d_lo, d_hi, n_subdiv=4, mirror_plane_z=None):
    """Induced (image) force on every active particle, per group (G, E, 3),
    from one joint solve per group (reference ``image_charge.py:816-1056``).

    mirror_plane_z: the profile is the image surface mirror-doubled through
    that plane (reference ``bem/geometry.py:346-427``): a phantom source -q
    at the reflection of every particle joins the group's solve, local
    mirror charges are kept only where |z - mirror_plane_z| exceeds the
    distance to the surface (``image_charge.py:620-681``), and the forces
    are those on the real particles (the phantoms' direct pull, the plane
    image, is the caller's to add)."""
    E_eval = position.shape[1]
    if mirror_plane_z is not None:
        ph = position.copy()
        ph[..., 2] = 2.0 * mirror_plane_z - position[..., 2]
        position = np.concatenate([position, ph], axis=1)
        charge = np.concatenate([charge, -charge], axis=1)
        active = np.concatenate([active, active], axis=1)
    G, E, _ = position.shape
    fl = rho_floor(profile)
    q = np.where(active, charge, 0.0)
    x, y, z = position[..., 0], position[..., 1], position[..., 2]
    rho = np.maximum(np.hypot(x, y), fl)
    phi = np.arctan2(y, x)
    p0, p1, tan, nrm = segment_geometry(profile)
    bd2, bseg, bt, bn = closest_point(rho.reshape(-1), z.reshape(-1), p0, tan, nrm)
    closest = p0[bseg] + bt[:, None] * tan[bseg]
    tu = tan[bseg] / np.linalg.norm(tan[bseg], axis=-1, keepdims=True)
    dist = np.sqrt(bd2)
    w = image_weight(dist, d_lo, d_hi) if d_hi > d_lo else np.zeros_like(dist)
    if mirror_plane_z is not None:
        w = np.where(np.abs(z.reshape(-1) - mirror_plane_z) > dist, w, 0.0)
    ri, zi, side = mirror_points(rho.reshape(-1), z.reshape(-1), closest, tu, bn)
    phi_i = np.where(side > 0, phi.reshape(-1), phi.reshape(-1) + np.pi)
    q_i = -q.reshape(-1) * w
    ri, zi, phi_i, q_i = (a.reshape(G, E) for a in (ri, zi, phi_i, q_i))
    prho = np.maximum(profile[:, 0], fl)
    pz = profile[:, 1]
    P = len(profile)
    m = np.arange(n_max + 1, dtype=float)

    def rhs_part(sr, sz, sp, sq):
        g = point_charge_potential_modes(prho[None, None, :], pz[None, None, :], sr[:, :, None], sz[:, :, None],
                                         n_max)  # (M, G, E, P)
        c = np.cos(m[:, None, None] * sp[None])
        s = np.sin(m[:, None, None] * sp[None])
        wgt = sq[None, :, :, None] * g
        return np.sum(wgt * c[..., None], axis=2), np.sum(wgt * s[..., None], axis=2)

    grc, grs = rhs_part(rho, z, phi, q)
    gic, gis = rhs_part(ri, zi, phi_i, q_i)
    rhs = np.stack([-(grc + gic), -(grs + gis)], axis=-1)  # (M, G, P, 2)
    rhs = np.moveaxis(rhs, 1, 2).reshape(n_max + 1, P, G * 2)
    sig = np.matmul(A_inv[: n_max + 1], rhs)
    sig = np.moveaxis(sig.reshape(n_max + 1, P, G, 2), 1, 2)
    sC, sS = sig[..., 0], sig[..., 1]
    # field of the residual density at the particles evaluated (the real ones)
    rho, z, phi, q = rho[:, :E_eval], z[:, :E_eval], phi[:, :E_eval], q[:, :E_eval]
    rq, zq, wq, sq, tq = quadrature_geometry(profile, n_subdiv)
    dc = (sC[..., sq] + tq * (sC[..., sq + 1] - sC[..., sq])) * wq  # (M, G, Q)
    ds = (sS[..., sq] + tq * (sS[..., sq + 1] - sS[..., sq])) * wq
    Phi, dPr, dPz = ring_field_modes(rho[:, :, None], z[:, :, None], rq[None, None, :], zq[None, None, :], n_max)
    cm = np.cos(m[:, None, None] * phi[None])
    sm = np.sin(m[:, None, None] * phi[None])
    dens = dc[:, :, None, :] * cm[..., None] + ds[:, :, None, :] * sm[..., None]
    ang = dc[:, :, None, :] * sm[..., None] - ds[:, :, None, :] * cm[..., None]
    Er = np.sum(np.sum(-dens * dPr, axis=-1), axis=0)
    Ez = np.sum(np.sum(-dens * dPz, axis=-1), axis=0)
    Ep = np.sum(np.sum(ang * Phi, axis=-1) * m[:, None, None] / rho[None], axis=0)
    Eres = np.stack([Er * np.cos(phi) - Ep * np.sin(phi), Er * np.sin(phi) + Ep * np.cos(phi), Ez], axis=-1)
    # direct field of every mirror charge of the group
    qxyz = np.stack([rho * np.cos(phi), rho * np.sin(phi), z], axis=-1)
    mxyz = np.stack([ri * np.cos(phi_i), ri * np.sin(phi_i), zi], axis=-1)
    diff = qxyz[:, :, None, :] - mxyz[:, None, :, :]
    r = np.linalg.norm(diff, axis=-1)
    rs = np.where(r > 0.0, r, 1.0)
    coef = np.where(r > 0.0, q_i[:, None, :] / (4.0 * np.pi * rs**3), 0.0)
    Emir = np.einsum("gnm,gnmc->gnc", coef, diff)
    F = q[..., None] * (Eres + Emir) / EPS0
    return np.where(active[:, :E_eval, None], F, 0.0)
