"""Generating profiles (rho, z) for the axisymmetric BEM solves.

``sphere_profile``, ``n_for_length`` and the cylindrical-well profile keep
the reference's definitions (``bem/mesh.py:65-170,317-409``): the BEM
density is solved on these exact nodes, so the node placement is part of
the result. ``post_profile`` is new: the reference has no post geometry."""
import numpy as np


def n_for_length(arc_length, max_length, min_points=2):
    """Points needed so every one of the (n-1) segments along an arc is no
    longer than max_length (reference ``bem/mesh.py:144-150``)."""
    if max_length <= 0.0:
        raise ValueError(f"max_length must be positive (got {max_length!r})")
    return max(min_points, max(1, int(np.ceil(arc_length / max_length))) + 1)


def sphere_profile(radius, n_theta):
    """Full sphere, pole to pole, uniform in polar angle."""
    th = np.linspace(0.0, np.pi, n_theta)
    return np.column_stack([radius * np.sin(th), radius * np.cos(th)])


def post_profile(R, H, n_segments):
    """Mirror-doubled hemisphere-capped post: a capsule of radius R whose
    cylindrical part spans z in [-H, H], from the top pole (0, H + R) down
    to the bottom pole (0, -H - R).

    Solving the closed capsule with odd Dirichlet data Ez*z gives a
    potential that is odd about z = 0, i.e. zero on the cathode plane: the
    same mirror argument the reference uses for the bare hemisphere
    (``bem/fields.py:22-33``), valid for any protrusion.

    Segments are shared between the two quarter-circle caps and the wall in
    proportion to arc length, so the total is exactly ``n_segments``."""
    if R <= 0.0 or H < 0.0:
        raise ValueError("need R > 0 and H >= 0")
    cap = 0.5 * np.pi * R
    total = 2.0 * cap + 2.0 * H
    n_cap = max(2, int(round(n_segments * cap / total)))
    n_wall = n_segments - 2 * n_cap
    if H > 0.0 and n_wall < 1:
        raise ValueError("too few segments for the wall")
    if H == 0.0:
        n_cap = n_segments // 2
        n_wall = 0
        if 2 * n_cap != n_segments:
            raise ValueError("a sphere (H = 0) needs an even segment count")
    th = np.linspace(0.0, 0.5 * np.pi, n_cap + 1)
    top = np.column_stack([R * np.sin(th), H + R * np.cos(th)])
    parts = [top]
    if n_wall:
        zw = np.linspace(H, -H, n_wall + 1)[1:]
        parts.append(np.column_stack([np.full_like(zw, R), zw]))
    th2 = np.linspace(0.5 * np.pi, np.pi, n_cap + 1)[1:]
    parts.append(np.column_stack([R * np.sin(th2), -H + R * np.cos(th2)]))
    prof = np.concatenate(parts, axis=0)
    prof[-1, 0] = 0.0
    prof[0, 0] = 0.0
    return prof


MIN_FILLET_SEGMENTS = 5  # every 90-degree fillet gets at least this many segments


def region_max_length(max_length, region_max_length):
    """A per-region cap can only tighten the global one."""
    return max_length if region_max_length is None else min(max_length, region_max_length)


def flat_annulus(r_inner, r_outer, z, n):
    r = np.linspace(r_inner, r_outer, n)
    return np.column_stack([r, np.full(n, float(z))])


def graded_annulus(r_inner, r_outer, z, first_step, growth=1.15):
    """Nodes from r_inner to r_outer, the spacing starting at first_step and
    growing by `growth` per step; a last gap shorter than half the previous
    step is absorbed by moving the last interior node onto r_outer."""
    r = [float(r_inner)]
    step = float(first_step)
    while r[-1] + step < r_outer:
        r.append(r[-1] + step)
        step *= growth
    if len(r) > 1 and r_outer - r[-1] < 0.5 * (r[-1] - r[-2]):
        r[-1] = float(r_outer)
    else:
        r.append(float(r_outer))
    r = np.asarray(r)
    return np.column_stack([r, np.full(len(r), float(z))])


def arc(center_r, radius, phi0, phi1, n):
    """(center_r + radius cos phi, radius sin phi) for phi from phi0 to phi1."""
    ph = np.linspace(phi0, phi1, n)
    return np.column_stack([center_r + radius * np.cos(ph), radius * np.sin(ph)])


def wall(rho, z0, z1, n):
    z = np.linspace(z0, z1, n)
    return np.column_stack([np.full(n, float(rho)), z])


def cylindrical_well_profile(R, H, plane_radius, max_length, bottom_fillet_radius=None, rim_fillet_radius=None,
                             bottom_max_length=None, bottom_fillet_max_length=None, wall_max_length=None,
                             rim_fillet_max_length=None, plane_max_length=None, plane_growth=1.15):
    """The real surface of a cylindrical well of radius R and depth H cut
    into the z = 0 cathode, from the axis outwards: bottom disk (z = -H),
    optional bottom fillet, wall (rho = R), optional rim fillet, then the
    plane graded out to plane_radius (reference ``bem/mesh.py:317-409``)."""
    rb = 0.0 if bottom_fillet_radius is None else float(bottom_fillet_radius)
    rr = 0.0 if rim_fillet_radius is None else float(rim_fillet_radius)
    if rb < 0.0:
        raise ValueError(f"bottom_fillet_radius must be >= 0 (got {bottom_fillet_radius!r})")
    if rr < 0.0:
        raise ValueError(f"rim_fillet_radius must be >= 0 (got {rim_fillet_radius!r})")
    if rb + rr >= H:
        raise ValueError(f"bottom_fillet_radius + rim_fillet_radius must be < H (got {rb!r} + {rr!r} >= {H!r})")
    parts = []
    n = n_for_length(R - rb, region_max_length(max_length, bottom_max_length))
    parts.append(flat_annulus(0.0, R - rb, -H, n))
    if rb > 0.0:
        n = n_for_length(0.5 * np.pi * rb, region_max_length(max_length, bottom_fillet_max_length),
                         min_points=MIN_FILLET_SEGMENTS + 1)
        a = arc(R - rb, rb, -0.5 * np.pi, 0.0, n)[1:]
        a[:, 1] += -H + rb
        parts.append(a)
    n = n_for_length(H - rr - rb, region_max_length(max_length, wall_max_length))
    parts.append(wall(R, -H + rb, -rr, n)[1:])
    if rr > 0.0:
        n = n_for_length(0.5 * np.pi * rr, region_max_length(max_length, rim_fillet_max_length),
                         min_points=MIN_FILLET_SEGMENTS + 1)
        a = arc(R + rr, rr, np.pi, 0.5 * np.pi, n)[1:]
        a[:, 1] += -rr
        parts.append(a)
    parts.append(graded_annulus(R + rr, plane_radius, 0.0, region_max_length(max_length, plane_max_length),
                                growth=plane_growth)[1:])
    return np.concatenate(parts, axis=0)


def cylindrical_well_image_profile(R, H, z0, plane_radius, max_length, bottom_fillet_radius=None,
                                   rim_fillet_radius=None, bottom_max_length=None, bottom_fillet_max_length=None,
                                   wall_max_length=None, rim_fillet_max_length=None, plane_max_length=None,
                                   plane_growth=1.15):
    """The recessed image surface of the well: the real surface moved z0
    into the conductor (reference ``bem/mesh.py:412-529``). The bottom disk
    sits at z = -H - z0, the wall at rho = R + z0, the plane at z = -z0;
    the bottom (reflex) corner is always rounded with radius
    bottom_fillet + z0 about the real fillet's centre, the rim (convex)
    corner keeps a rounding of radius rim_fillet - z0 only when that is
    positive (else the sharp mitre at (R + z0, -z0))."""
    if not (0.0 < z0 < min(R, H)):
        raise ValueError(f"z0 must satisfy 0 < z0 < min(R, H) (got z0={z0!r}, R={R!r}, H={H!r})")
    rb = 0.0 if bottom_fillet_radius is None else float(bottom_fillet_radius)
    rr = 0.0 if rim_fillet_radius is None else float(rim_fillet_radius)
    if rb < 0.0:
        raise ValueError(f"bottom_fillet_radius must be >= 0 (got {bottom_fillet_radius!r})")
    if rr < 0.0:
        raise ValueError(f"rim_fillet_radius must be >= 0 (got {rim_fillet_radius!r})")
    top = max(rr, z0)
    if rb + top >= H:
        raise ValueError("bottom_fillet_radius + max(rim_fillet_radius, z0) must be < H "
                         f"(got {rb!r} + {top!r} >= {H!r})")
    img_bottom = rb + z0
    img_rim = rr - z0
    parts = []
    n = n_for_length(R - rb, region_max_length(max_length, bottom_max_length))
    parts.append(flat_annulus(0.0, R - rb, -H - z0, n))
    n = n_for_length(0.5 * np.pi * img_bottom, region_max_length(max_length, bottom_fillet_max_length),
                     min_points=MIN_FILLET_SEGMENTS + 1)
    a = arc(R - rb, img_bottom, -0.5 * np.pi, 0.0, n)[1:]
    a[:, 1] += -H + rb
    parts.append(a)
    n = n_for_length(-top - (-H + rb), region_max_length(max_length, wall_max_length))
    parts.append(wall(R + z0, -H + rb, -top, n)[1:])
    if img_rim > 0.0:
        n = n_for_length(0.5 * np.pi * img_rim, region_max_length(max_length, rim_fillet_max_length),
                         min_points=MIN_FILLET_SEGMENTS + 1)
        a = arc(R + rr, img_rim, np.pi, 0.5 * np.pi, n)[1:]
        a[:, 1] += -rr
        parts.append(a)
        r0 = R + rr
    else:
        r0 = R + z0
    parts.append(graded_annulus(r0, plane_radius, -z0, region_max_length(max_length, plane_max_length),
                                growth=plane_growth)[1:])
    return np.concatenate(parts, axis=0)


def sphere_cap(radius, n_theta):
    """Sphere cap from the pole (polar angle 0) to the equator, uniform in
    polar angle."""
    th = np.linspace(0.0, 0.5 * np.pi, n_theta)
    return np.column_stack([radius * np.sin(th), radius * np.cos(th)])


def _recessed_tip(R, z0, max_length, cap_max_length, fillet_max_length):
    """Cap of radius R - z0 (pole first) and the quarter-torus fillet of tube
    radius z0 about rho = R that rounds the recessed ridge down to the
    point (R, -z0), where its tangent is horizontal."""
    if not (0.0 < z0 < R):
        raise ValueError(f"z0 must satisfy 0 < z0 < R (got z0={z0!r}, R={R!r})")
    cap = sphere_cap(R - z0, n_for_length(0.5 * np.pi * (R - z0), region_max_length(max_length, cap_max_length)))
    n = n_for_length(0.5 * np.pi * z0, region_max_length(max_length, fillet_max_length),
                     min_points=MIN_FILLET_SEGMENTS + 1)
    return cap, arc(R, z0, np.pi, 1.5 * np.pi, n)[1:]


def hemisphere_tip_image_profile(R, z0, plane_radius, max_length, cap_max_length=None, fillet_max_length=None,
                                 plane_max_length=None, plane_growth=1.15):
    """The recessed image surface of a hemisphere of radius R on a plane,
    truncated at plane_radius (reference ``bem/mesh.py:200-240``): the cap
    of radius R - z0, the ridge fillet of radius z0 (tangent to the cap and
    to the plane), then the graded plane at z = -z0 out to plane_radius."""
    cap, fillet = _recessed_tip(R, z0, max_length, cap_max_length, fillet_max_length)
    plane = graded_annulus(R, plane_radius, -z0, region_max_length(max_length, plane_max_length),
                           growth=plane_growth)[1:]
    return np.concatenate([cap, fillet, plane], axis=0)


def hemisphere_tip_image_doubled_profile(R, z0, max_length, cap_max_length=None, fillet_max_length=None):
    """The same recessed cap and fillet, closed by their reflection through
    the image plane z = -z0 instead of a truncated plane (reference
    ``bem/mesh.py:243-314``): pole to mirrored pole, the corner (R, -z0)
    shared by both halves. Used with a phantom source -q at the reflection
    of every real charge."""
    cap, fillet = _recessed_tip(R, z0, max_length, cap_max_length, fillet_max_length)
    upper = np.concatenate([cap, fillet], axis=0)
    lower = upper[-2::-1].copy()
    lower[:, 1] = -2.0 * z0 - lower[:, 1]
    return np.concatenate([upper, lower], axis=0)
