"""Oracle analytic fields (reference ``fields.py:43-160``)."""
import numpy as np


def gun_field(pos, Ez):
    """Uniform Ez for z >= 0, zero below (reference ``fields.py:43-59``)."""
    E = np.zeros_like(pos)
    E[..., 2] = np.where(pos[..., 2] >= 0.0, Ez, 0.0)
    return E


def tip_field(pos, Ez, R):
    """Grounded sphere in a uniform field, exterior of the upper hemisphere
    (reference ``fields.py:119-160``)."""
    x, y, z = pos[..., 0], pos[..., 1], pos[..., 2]
    r = np.sqrt(x * x + y * y + z * z)
    valid = (r >= R * (1.0 - 1e-6)) & (z >= 0.0)
    rs = np.where(valid, r, R)
    r3 = rs * rs * rs
    r5 = r3 * rs * rs
    R3 = R * R * R
    E = np.stack([3.0 * Ez * R3 * x * z / r5, 3.0 * Ez * R3 * y * z / r5,
                  Ez * (1.0 - R3 / r3 + 3.0 * R3 * z * z / r5)], axis=-1)
    return np.where(valid[..., None], E, 0.0)


def smoothstep_weight(d, d_lo, d_hi):
    """1 at d <= d_lo, 0 at d >= d_hi, cubic in between (reference
    ``bem/image_charge.py:328-342``)."""
    t = np.clip((d_hi - d) / (d_hi - d_lo), 0.0, 1.0)
    return t * t * (3.0 - 2.0 * t)


def blend(near, far, z, z_lo, z_hi):
    """Reference ``bem/geometry.py:44-55``."""
    w = smoothstep_weight(z, z_lo, z_hi)
    return w[..., None] * near + (1.0 - w[..., None]) * far
