"""Emission ordering and seeded emission samplers.

``group_indices`` keeps the reference's contract (``emission.py:6-38``):
sort by id (stable), chunk into equal groups, ValueError on a non-divisible
count or duplicate ids. ``group_offsets`` is the uneven-group extension: a
CSR offset array over the id-sorted particles.

The samplers have no counterpart in the reference (its notebooks use the
external distgen tool); each draws from ``numpy.random.default_rng(seed)``
in a fixed order so the same seed gives the same particles everywhere.
"""
import numpy as np

from .constants import ELECTRON_MASS_EV


def group_indices(ids, n_emit):
    ids = np.asarray(ids)
    n = len(ids)
    if n_emit < 1:
        raise ValueError(f"n_emit must be a positive integer, got {n_emit}")
    if n % n_emit != 0:
        raise ValueError(f"Number of particles ({n}) is not evenly divisible by n_emit ({n_emit}).")
    if len(np.unique(ids)) != n:
        raise ValueError("particle_group.id must contain unique values.")
    return np.argsort(ids, kind="stable").reshape(n // n_emit, n_emit)


def group_offsets(ids, group_sizes):
    """(order, offsets): particles sorted by id, then consecutive runs of
    ``group_sizes`` form the groups. ``offsets`` has n_groups + 1 entries."""
    ids = np.asarray(ids)
    sizes = np.asarray(group_sizes, dtype=np.int64)
    if sizes.ndim != 1 or np.any(sizes < 1):
        raise ValueError("group sizes must be positive integers")
    if int(sizes.sum()) != len(ids):
        raise ValueError(f"group sizes sum to {int(sizes.sum())}, not the particle count {len(ids)}")
    if len(np.unique(ids)) != len(ids):
        raise ValueError("particle_group.id must contain unique values.")
    offsets = np.zeros(len(sizes) + 1, dtype=np.int64)
    np.cumsum(sizes, out=offsets[1:])
    return np.argsort(ids, kind="stable"), offsets


def _momentum_ev_c(ke_ev):
    """Non-relativistic |p| [eV/c] for kinetic energy [eV] (as the
    reference's v = p/m model assumes)."""
    return np.sqrt(2.0 * ke_ev * ELECTRON_MASS_EV)


def _cos_weighted(rng, n):
    """Lambertian directions about +z: theta = asin(sqrt(u)), phi ~ U."""
    u = rng.random(n)
    v = rng.random(n)
    th = np.arcsin(np.sqrt(u))
    ph = 2.0 * np.pi * v
    return np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)], axis=-1)


def flat_gaussian_source(n, sigma_r=1e-6, mean_ke_ev=0.1, seed=0):
    """Flat-cathode source: x, y ~ N(0, sigma_r), z = 0, t = 0,
    KE ~ Exp(mean), cos-weighted direction. Returns a dict of arrays in
    openPMD units (m, eV/c, s) with ids 1..n."""
    rng = np.random.default_rng(seed)
    x = rng.normal(0.0, sigma_r, n)
    y = rng.normal(0.0, sigma_r, n)
    ke = rng.exponential(mean_ke_ev, n)
    d = _cos_weighted(rng, n)
    p = _momentum_ev_c(ke)[:, None] * d
    return dict(x=x, y=y, z=np.zeros(n), px=p[:, 0], py=p[:, 1], pz=p[:, 2], t=np.zeros(n),
                id=np.arange(1, n + 1))


def cap_source(n, R, center_z, theta_max_deg=60.0, mean_ke_ev=0.2, seed=0, t_spread=0.0):
    """Emission from a hemispherical cap of radius R centred at (0, 0,
    center_z): polar angle ~ U[0, theta_max], azimuth ~ U, KE ~ Exp(mean),
    cos-weighted direction about the local outward normal. Birth times
    ~ U[0, t_spread] (all zero by default)."""
    rng = np.random.default_rng(seed)
    th = np.deg2rad(theta_max_deg) * rng.random(n)
    ph = 2.0 * np.pi * rng.random(n)
    nrm = np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)], axis=-1)
    pos = R * nrm
    pos[:, 2] += center_z
    ke = rng.exponential(mean_ke_ev, n)
    loc = _cos_weighted(rng, n)
    # rotate local (+z = normal) frame onto the normal
    t1 = np.stack([np.cos(th) * np.cos(ph), np.cos(th) * np.sin(ph), -np.sin(th)], axis=-1)
    t2 = np.stack([-np.sin(ph), np.cos(ph), np.zeros(n)], axis=-1)
    d = loc[:, :1] * t1 + loc[:, 1:2] * t2 + loc[:, 2:] * nrm
    p = _momentum_ev_c(ke)[:, None] * d
    t = t_spread * rng.random(n) if t_spread > 0 else np.zeros(n)
    return dict(x=pos[:, 0], y=pos[:, 1], z=pos[:, 2], px=p[:, 0], py=p[:, 1], pz=p[:, 2], t=t,
                id=np.arange(1, n + 1))


def well_bottom_source(group_sizes, R, H, mean_ke_ev=0.1, t_offset_max=10e-12, radius_fraction=0.9, seed=0):
    """Emission from the bottom disk of a cylindrical well (z = -H),
    uniform in area within radius_fraction * R, KE ~ Exp(mean),
    cos-weighted direction about +z; every group gets its own emission-time
    offset ~ U[0, t_offset_max] (all its members are born then). Returns the
    source dict with ids 1..n in group order."""
    sizes = np.asarray(group_sizes, dtype=np.int64)
    n = int(sizes.sum())
    rng = np.random.default_rng(seed)
    r = radius_fraction * R * np.sqrt(rng.random(n))
    ph = 2.0 * np.pi * rng.random(n)
    ke = rng.exponential(mean_ke_ev, n)
    d = _cos_weighted(rng, n)
    p = _momentum_ev_c(ke)[:, None] * d
    t = np.repeat(t_offset_max * rng.random(len(sizes)), sizes)
    return dict(x=r * np.cos(ph), y=r * np.sin(ph), z=np.full(n, -float(H)), px=p[:, 0], py=p[:, 1], pz=p[:, 2],
                t=t, id=np.arange(1, n + 1))
