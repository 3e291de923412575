"""Oracle pair forces (reference ``forces.py:16-191``).

Arrays are (G, E, 3) positions, (G, E) charges and active masks."""
import numpy as np

from .constants import COULOMB_CONSTANT


def pairwise(pos_i, pos_j, q_i, q_j, active, plummer_radius):
    """sum_j k q_i q_j (r_i - r_j) / (|r_i - r_j|^2 + a^2)^1.5 over
    active-active pairs (reference ``forces.py:16-29``)."""
    diff = pos_i[:, :, None, :] - pos_j[:, None, :, :]
    r2 = np.sum(diff * diff, axis=-1) + plummer_radius**2
    r3 = r2 * np.sqrt(r2)
    both = active[:, :, None] & active[:, None, :]
    c = COULOMB_CONSTANT * q_i[:, :, None] * q_j[:, None, :] / r3
    c = np.where(both, c, 0.0)
    return np.einsum("gij,gijc->gic", c, diff)


def coulomb(pos, q, active, plummer_radius):
    """Reference ``forces.py:32-57``."""
    return pairwise(pos, pos, q, q, active, plummer_radius)


def flat_image(pos, q, active, z0, plummer_radius):
    """Images at z' = -2 z0 - z with charge -q, self included (reference
    ``forces.py:60-98``)."""
    img = pos.copy()
    img[..., 2] = -2.0 * z0 - pos[..., 2]
    return pairwise(pos, img, q, -q, active, plummer_radius)


def tip_image(pos, q, active, sphere_radius, plummer_radius, plane_z0):
    """Three images per member: sphere image, plane image, plane image of
    the sphere image (reference ``forces.py:101-191``)."""
    a = sphere_radius
    d = np.sqrt(np.sum(pos * pos, axis=-1))
    d = np.where(d > 0.0, d, a)
    s_pos = pos * ((a * a) / (d * d))[..., None]
    s_q = -q * (a / d)
    p_pos = pos.copy()
    p_pos[..., 2] = -2.0 * plane_z0 - pos[..., 2]
    ps_pos = s_pos.copy()
    ps_pos[..., 2] = -2.0 * plane_z0 - s_pos[..., 2]
    f1 = pairwise(pos, s_pos, q, s_q, active, plummer_radius)
    f2 = pairwise(pos, p_pos, q, -q, active, plummer_radius)
    f3 = pairwise(pos, ps_pos, q, -s_q, active, plummer_radius)
    return f1 + f2 + f3
