"""Oracle tracker wiring (reference ``tracker.py:214-386``,
``geometry.py:97-322``, ``emission.py:6-38``).

Geometries are plain dicts:

* ``{"kind": "flat", "E_gun", "z0" (or None), "kill_z_below" (or None)}``
* ``{"kind": "tip", "E_gun", "R", "z0", "kill_z_below"}`` (analytic)
* ``{"kind": "well_bem", "E_gun", "R", "H", "profile", "sigma", "z_handoff",
  "bottom_fillet", "rim_fillet"}`` (cylindrical well, reference bem/geometry.py:510-812)
* ``{"kind": "tip_bem" | "post_bem", "E_gun", "R", "H" (post only), "profile",
  "sigma", "n_subdiv", "z_handoff", "kill_z_below"}`` (image force off, the
  reference BEM default ``z0=None``); the tip with ``"z0"``, ``"image_profile"``,
  ``"image_A_inv"``, ``"n_max"``, ``"d_lo"``, ``"d_hi"`` and
  ``"image_mirror_symmetric"`` takes the image-charge BEM force
"""
import numpy as np

from . import bem, fields, forces, rk45, screens
from .constants import ELECTRON_MASS_EV, ELEMENTARY_CHARGE, SPEED_OF_LIGHT, ev_c_to_si_momentum, ev_to_kg


def group_indices(ids, n_emit):
    """Reference ``emission.py:6-38``."""
    n = len(ids)
    if n_emit < 1:
        raise ValueError("n_emit must be a positive integer")
    if n % n_emit:
        raise ValueError("not evenly divisible")
    if len(np.unique(ids)) != n:
        raise ValueError("ids must be unique")
    return np.argsort(np.asarray(ids), kind="stable").reshape(n // n_emit, n_emit)


def field(geom, pos):
    k = geom["kind"]
    if k == "flat":
        return fields.gun_field(pos, geom["E_gun"])
    if k == "tip":
        return fields.tip_field(pos, geom["E_gun"], geom["R"])
    if k == "tip_bem":
        return bem.tip_bem_field(pos, geom["E_gun"], geom["R"], geom["profile"], geom["sigma"],
                                 geom.get("n_subdiv", 8), geom.get("z_handoff"))
    if k == "post_bem":
        return bem.post_bem_field(pos, geom["E_gun"], geom["R"], geom["H"], geom["profile"], geom["sigma"],
                                  geom["z_handoff"], geom.get("n_subdiv", 8))
    if k == "well_bem":
        return bem.well_bem_field(pos, geom["E_gun"], geom["R"], geom["H"], geom["profile"], geom["sigma"],
                                  geom["z_handoff"], geom.get("bottom_fillet", 0.0), geom.get("rim_fillet", 0.0),
                                  geom.get("n_subdiv", 8))
    raise ValueError(k)


def image_force(geom, pos, q, active, a):
    k = geom["kind"]
    z0 = geom.get("z0")
    if k in ("well_bem", "tip_bem") and z0 is not None:
        # reference bem/geometry.py:326-344,786-801: image BEM near, plane image far, blended; the tip's
        # mirror-doubled image surface adds the phantom charges' direct pull, the plane image (346-427)
        from . import image

        mpz = -z0 if geom.get("image_mirror_symmetric") else None
        near = image.image_force(geom["image_profile"], geom["image_A_inv"], geom["n_max"], pos, q, active,
                                 geom["d_lo"], geom["d_hi"], mirror_plane_z=mpz)
        far = forces.flat_image(pos, q, active, z0, a)
        if mpz is not None:
            near = near + far
        from .fields import blend
        return blend(near, far, pos[..., 2], 0.5 * geom["z_handoff"], geom["z_handoff"])
    if z0 is None or k in ("tip_bem", "post_bem", "well_bem"):
        return np.zeros_like(pos)
    if k == "flat":
        return forces.flat_image(pos, q, active, z0, a)
    if k == "tip":
        return forces.tip_image(pos, q, active, geom["R"] - z0, a, z0)
    raise ValueError(k)


def kill_mask(geom, pos):
    k = geom["kind"]
    kz = geom.get("kill_z_below", 0.0)
    if k == "flat":
        m = np.zeros(pos.shape[:-1], dtype=bool) if kz is None else pos[..., 2] <= kz
    elif k == "well_bem":  # reference bem/geometry.py:803-812: the conductor only
        m = bem.well_contains(np.hypot(pos[..., 0], pos[..., 1]), pos[..., 2], geom["R"], geom["H"],
                              geom.get("bottom_fillet", 0.0), geom.get("rim_fillet", 0.0))
    else:
        R = geom["R"]
        rk = R * (1.0 - 1e-6)
        if k == "post_bem":
            inside = bem.post_axis_distance(pos, geom["H"]) ** 2 <= rk * rk
        else:
            inside = np.sum(pos * pos, axis=-1) <= rk * rk
        m = inside if kz is None else inside | (pos[..., 2] <= kz)
    zmax = geom.get("z_max")
    if zmax is not None:
        m = m | (pos[..., 2] >= zmax)
    return m


def make_accel(geom, charge, mass, q_over_m, a):
    """Reference ``geometry.py:265-301``."""
    def accel(pos, vel, active, gidx):
        c = charge[gidx]
        m = mass[gidx]
        af = q_over_m * field(geom, pos)
        ac = 0.0 if pos.shape[1] == 1 else forces.coulomb(pos, c, active, a) / m[..., None]
        ai = image_force(geom, pos, c, active, a) / m[..., None]
        return np.where(active[..., None], af + ac + ai, 0.0)
    return accel


def setup(x, y, z, px, py, pz, t, ids, n_emit, weight=None):
    """Grouped SI state (reference ``tracker.py:214-288``) for electrons."""
    mass = ev_to_kg(ELECTRON_MASS_EV)
    idx = group_indices(ids, n_emit)
    w = np.full(len(ids), ELEMENTARY_CHARGE)
    st = dict(
        pos=np.stack([x[idx], y[idx], z[idx]], axis=-1).astype(float),
        vel=np.stack([ev_c_to_si_momentum(px)[idx] / mass, ev_c_to_si_momentum(py)[idx] / mass,
                      ev_c_to_si_momentum(pz)[idx] / mass], axis=-1),
        t_birth=np.asarray(t, dtype=float)[idx],
        charge=(-1.0 * w)[idx],
        mass_p=(mass * (w / ELEMENTARY_CHARGE))[idx],
        ids=np.asarray(ids)[idx],
        q_over_m=-ELEMENTARY_CHARGE / mass,
        mass=mass,
        order=idx,
    )
    return st


def default_t_max(st, E_gun, screen_z, t_out=None):
    """Reference ``tracker.py:290-307``: covers the output times too."""
    a0 = abs(E_gun) * abs(st["q_over_m"])
    zs = np.array(screen_z, dtype=float)
    zp = zs[zs > 0]
    tt = np.sqrt(2 * zp.max() / a0) if len(zp) and a0 > 0 else 1e-12
    t_max = float(st["t_birth"].max()) + 3.0 * tt
    if t_out is not None and len(t_out):
        t_max = max(t_max, float(np.max(t_out)))
    return t_max


def run(st, geom, screen_z, t_max, rtol=1e-7, atol=1e-12, plummer_radius=1e-12, z_max=None,
        max_steps=10_000_000, count_stages=True, t_out=None, raise_at_cap=True):
    """In-process tracker run (reference ``tracker.py:340-386``).

    Returns dict with final pos/vel/t_death, per-group attempts/accepted,
    events (list of dicts with batch, screen, interval, group, particle,
    t, pos, vel) in the reference's recording order, and the number of
    particle-stages evaluated (7 x group size per attempt)."""
    g = dict(geom)
    if z_max is not None:
        g["z_max"] = z_max
    accel = make_accel(g, st["charge"], st["mass_p"], st["q_over_m"], plummer_radius)
    stages = [0]

    def counted(pos, vel, active, gidx):
        stages[0] += pos.shape[0] * pos.shape[1]
        return accel(pos, vel, active, gidx)

    events = []

    def on_step(it, gi, t0, p0, v0, t1, p1, v1):
        for e in screens.find_crossings(screen_z, t0, p0, v0, t1, p1, v1):
            e["batch"] = it
            e["group"] = int(gi[e.pop("row")])
            events.append(e)

    outs = None if t_out is None else np.sort(np.asarray(t_out, dtype=float))
    traj = None if outs is None else [[] for _ in outs]

    def on_output(k, gi, t, pos, vel):
        # reference trajectories.py:44-74: the group's particles born by then
        for r, g_ in enumerate(gi):
            for p in range(pos.shape[1]):
                if st["t_birth"][g_, p] <= t[r]:
                    traj[k].append(dict(group=int(g_), particle=p, t=float(t[r]), pos=pos[r, p].copy(),
                                        vel=vel[r, p].copy()))

    out = rk45.integrate(counted if count_stages else accel, st["pos"], st["vel"], st["t_birth"], t_max,
                         rtol, atol, on_step=on_step, kill=lambda p: kill_mask(g, p), max_steps=max_steps,
                         output_times=outs, on_output=on_output if outs is not None else None,
                         raise_at_cap=raise_at_cap)
    out["trajectories"] = traj
    # screens are recorded per screen; within one screen the reference's
    # order is (batch, interval, group, particle)
    out["events"] = events
    out["stages"] = stages[0]
    return out
