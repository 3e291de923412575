"""Generating profiles (rho, z) for the axisymmetric BEM solves.

``sphere_profile`` and ``n_for_length`` keep the reference's definitions
(``bem/mesh.py:65-73,134-150``). ``post_profile`` is new: the reference has
no post geometry."""
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
