"""Cathode shapes behind the reference's Geometry interface
(reference geometry.py:37-322), evaluated on the MI355X.

Each shape lowers itself to an ``spt_geometry`` parameter struct (kind,
scalars, image model) and, for BEM shapes, a profile + sigma table that the
library uploads once. ``field``, ``image_force`` and ``kill_mask`` accept
NumPy arrays or torch tensors of shape (..., 3) and return the same kind;
the arithmetic runs in the HIP library.
"""
import contextlib
from abc import ABC, abstractmethod

import numpy as np

from . import _lib
from .fields import GunField, HemisphericalTipField, _DeviceField, _to_device, _from_device

DEFAULT_Z0 = 3.0e-9  # m, reference geometry.py:34


class Geometry(ABC):
    """Interface a cathode shape implements (reference geometry.py:37-94)."""

    @abstractmethod
    def field(self, position):
        """External electric field [V/m] at `position`, shape (..., 3)."""

    @abstractmethod
    def image_force(self, position, charge, active, plummer_radius):
        """Image-charge force [N], shape (n_groups, n_emit, 3), within groups."""

    @abstractmethod
    def kill_mask(self, position):
        """Boolean array, shape position.shape[:-1]: inside the conductor?"""

    @abstractmethod
    def worker_args(self):
        """Plain picklable tuple; first element names the class."""

    @staticmethod
    def from_worker_args(args, xp=None):
        kind = args[0]
        if kind not in _WORKER_REGISTRY:
            from .bem import geometry as _bem_geometry  # noqa: F401  (registers BEM shapes)
        return _WORKER_REGISTRY[kind]._from_worker_args(args[1:])

    # -- lowering to the C ABI -------------------------------------------
    def spt_geometry(self, z_max=None):
        """The ``spt_geometry`` struct for this shape (include/spt.h)."""
        raise NotImplementedError

    def bem_table(self):
        """(profile (n, 2), sigma (n,), n_subdiv) for BEM shapes, else None."""
        return None

    def _ctx(self):
        ctx = _lib.Context.get(getattr(self, "device", 0))
        tab = self.bem_table()
        if tab is not None:
            key = (_lib.token(self), tab[0].shape[0], tab[2])
            if ctx.bem_key != key:
                prof, sig, nsub = tab
                prof_c, pp = _lib.host_ptr(np.asarray(prof, dtype=np.float64))
                sig_c, sp = _lib.host_ptr(np.asarray(sig, dtype=np.float64))
                ctx.check(ctx.lib.spt_bem_upload(ctx.h, pp, prof_c.shape[0], sp, int(nsub)), "spt_bem_upload")
                ctx.bem_key = key
                ctx.tiles_key = None  # the upload drops any tiles
        return ctx

    @contextlib.contextmanager
    def far_rule(self, k4=2.0, k2=16.0, near_width=None, clusters=None, tiles=None):
        """Within the block, every field gather on this geometry's context
        sums a segment at distance D >= L over the whole segment by Gauss-8
        below k4 L, Gauss-4 below k2 L and Gauss-2 beyond (spt_bem_far_rule),
        instead of the reference's fixed sub-panel table; near_width widens
        the near rule's intervals (spt_bem_near_width) and clusters (a
        bem_amd.bem.clusters.Clusters of this geometry) replaces runs of
        distant segments by their expansions (spt_bem_clusters) and tiles (a
        bem_amd.bem.tiles.Tiles of this geometry) replace the segments
        outside a point's tile window by the tile's table (spt_bem_tiles).
        Restored on exit."""
        ctx = self._ctx()
        ctx.check(ctx.lib.spt_bem_far_rule(ctx.h, float(k4), float(k2)), "spt_bem_far_rule")
        try:
            if near_width is not None:
                ctx.check(ctx.lib.spt_bem_near_width(ctx.h, float(near_width)), "spt_bem_near_width")
            if clusters is not None:
                clusters.install(ctx)
            if tiles is not None:
                tiles.install(ctx)
            yield
        finally:
            ctx.lib.spt_bem_far_rule(ctx.h, 0.0, 0.0)
            ctx.lib.spt_bem_near_width(ctx.h, 0.0)
            ctx.lib.spt_bem_clusters(ctx.h, None, 0, None, None, None, 0, 0.0, 0.0)
            ctx.lib.spt_bem_tiles_enable(ctx.h, 0)

    def _device_field(self, position, precision="fp64"):
        return _DeviceField(self.spt_geometry(), self._ctx(), _precision_code(precision)).evaluate(position)

    def _device_kill(self, position, z_max=None):
        ctx = self._ctx()
        g = self.spt_geometry(z_max)
        pos, meta = _to_device(position, ctx)
        import torch

        out = torch.empty(pos.shape[0], dtype=torch.uint8, device=ctx.torch_device)
        ctx.bind_stream()
        ctx.check(ctx.lib.spt_kill_mask(ctx.h, g, _lib.ptr(pos), pos.shape[0], _lib.ptr(out)), "spt_kill_mask")
        ctx.sync()
        res = out.bool()
        return _from_device(res, meta, trailing=False)

    def _device_image(self, position, charge, active, plummer_radius, precision="fp64"):
        return _pair_forces(self, position, charge, active, plummer_radius, want="image", precision=precision)


def _precision_code(precision, trace=False):
    codes = {"fp64": _lib.PREC_FP64, "mixed": _lib.PREC_MIXED}
    if trace:
        codes["pairs32"] = _lib.PREC_PAIR32
    if precision not in codes:
        raise ValueError(f"precision must be one of {sorted(codes)}, got {precision!r}")
    return codes[precision]


def _pair_forces(geom, position, charge, active, plummer_radius, want, precision="fp64"):
    """Coulomb or image force over (G, E, 3) arrays via spt_pair_forces."""
    import torch

    ctx = geom._ctx()
    g = geom.spt_geometry()
    is_np = not hasattr(position, "detach")
    P = torch.as_tensor(np.asarray(position) if is_np else position, dtype=torch.float64, device=ctx.torch_device)
    G, E, _ = P.shape
    q = torch.as_tensor(np.asarray(charge) if is_np else charge, dtype=torch.float64, device=ctx.torch_device)
    a = torch.as_tensor(np.asarray(active) if is_np else active, device=ctx.torch_device).to(torch.uint8)
    P = P.reshape(-1, 3).contiguous()
    q = q.reshape(-1).contiguous()
    a = a.reshape(-1).contiguous()
    off = torch.arange(0, G * E + 1, E, dtype=torch.int64, device=ctx.torch_device)
    out = torch.zeros_like(P)
    ctx.bind_stream()
    fc = _lib.ptr(out) if want == "coulomb" else None
    fi = _lib.ptr(out) if want == "image" else None
    ctx.check(ctx.lib.spt_pair_forces(ctx.h, g, _lib.ptr(P), _lib.ptr(q), _lib.ptr(a), _lib.ptr(off), G,
                                      float(plummer_radius), fc, fi, _precision_code(precision)), "spt_pair_forces")
    ctx.sync()
    out = out.reshape(G, E, 3)
    return out.cpu().numpy() if is_np else out


def coulomb_force(position, charge, active, plummer_radius, geometry=None, precision="fp64"):
    """In-group Plummer-softened Coulomb force (reference forces.py:32-57);
    precision "mixed" sums in FP32 with Kahan compensation."""
    geom = geometry if geometry is not None else FlatCathode(0.0, z0=None)
    return _pair_forces(geom, position, charge, active, plummer_radius, want="coulomb", precision=precision)


class FlatCathode(Geometry):
    """The z = 0 plane (reference geometry.py:97-163)."""

    def __init__(self, E_gun, z0=DEFAULT_Z0, kill_z_below=0.0, device=0):
        self.E_gun = float(E_gun)
        self.z0 = z0
        self.kill_z_below = kill_z_below
        self.device = device
        self._field = GunField(E_gun, device=device)

    @property
    def field_solver(self):
        return self._field

    def spt_geometry(self, z_max=None):
        g = _lib.Geometry_t()
        g.kind = _lib.GEOM_FLAT
        g.image = _lib.IMAGE_NONE if self.z0 is None else _lib.IMAGE_PLANE
        g.E_gun = self.E_gun
        g.z0 = 0.0 if self.z0 is None else float(self.z0)
        g.has_kill_z = 0 if self.kill_z_below is None else 1
        g.kill_z_below = 0.0 if self.kill_z_below is None else float(self.kill_z_below)
        g.has_z_max = 0 if z_max is None else 1
        g.z_max = 0.0 if z_max is None else float(z_max)
        return g

    def field(self, position):
        return self._device_field(position)

    def image_force(self, position, charge, active, plummer_radius):
        if self.z0 is None:
            return np.zeros_like(position) if not hasattr(position, "detach") else position.new_zeros(position.shape)
        return self._device_image(position, charge, active, plummer_radius)

    def kill_mask(self, position):
        return self._device_kill(position)

    def worker_args(self):
        return ("flat", self.E_gun, self.z0, self.kill_z_below)

    @classmethod
    def _from_worker_args(cls, args):
        E_gun, z0, kill_z_below = args
        return cls(E_gun, z0=z0, kill_z_below=kill_z_below)


class HemisphericalTip(Geometry):
    """Flat cathode with a grounded hemispherical tip of radius R at the
    origin, closed-form field and 3-image force (reference
    geometry.py:166-256)."""

    def __init__(self, E_gun, R, z0=DEFAULT_Z0, kill_z_below=0.0, device=0):
        if z0 is not None and R <= z0:
            raise ValueError(f"R ({R!r}) must be greater than z0 ({z0!r}) -- a tip no bigger "
                             "than the image-charge smoothing offset isn't a case this "
                             "approximation is meant to cover.")
        self.E_gun = float(E_gun)
        self.R = float(R)
        self.z0 = z0
        self.kill_z_below = kill_z_below
        self.device = device
        self._field = HemisphericalTipField(E_gun, R, device=device)

    @property
    def field_solver(self):
        return self._field

    def spt_geometry(self, z_max=None):
        g = _lib.Geometry_t()
        g.kind = _lib.GEOM_TIP
        g.image = _lib.IMAGE_NONE if self.z0 is None else _lib.IMAGE_TIP3
        g.E_gun = self.E_gun
        g.R = self.R
        g.z0 = 0.0 if self.z0 is None else float(self.z0)
        g.has_kill_z = 0 if self.kill_z_below is None else 1
        g.kill_z_below = 0.0 if self.kill_z_below is None else float(self.kill_z_below)
        g.has_z_max = 0 if z_max is None else 1
        g.z_max = 0.0 if z_max is None else float(z_max)
        return g

    def field(self, position):
        return self._device_field(position)

    def image_force(self, position, charge, active, plummer_radius):
        if self.z0 is None:
            return np.zeros_like(position) if not hasattr(position, "detach") else position.new_zeros(position.shape)
        return self._device_image(position, charge, active, plummer_radius)

    def kill_mask(self, position):
        return self._device_kill(position)

    def worker_args(self):
        return ("hemispherical_tip", self.E_gun, self.R, self.z0, self.kill_z_below)

    @classmethod
    def _from_worker_args(cls, args):
        E_gun, R, z0, kill_z_below = args
        return cls(E_gun, R, z0=z0, kill_z_below=kill_z_below)


_WORKER_REGISTRY = {
    "flat": FlatCathode,
    "hemispherical_tip": HemisphericalTip,
}


def lower(geometry):
    """This package's geometry for a reference Geometry object, or for its
    ``worker_args()`` tuple: the tuple's first element names the class and
    the rest are its constructor arguments, field for field (reference
    geometry.py:61-94; the reference tracker rebuilds every geometry this
    way, tracker.py:199-202). A geometry of this package is returned as is.
    Raises ValueError for a kind this backend does not implement."""
    if isinstance(geometry, Geometry):
        return geometry
    args = geometry.worker_args() if hasattr(geometry, "worker_args") else geometry
    if not isinstance(args, (tuple, list)) or not args or not isinstance(args[0], str):
        raise ValueError("expected a Geometry or its worker_args() tuple (a class name first)")
    args = tuple(args)
    if args[0] not in _WORKER_REGISTRY:
        from .bem import geometry as _bem_geometry  # noqa: F401  (registers BEM shapes)
    if args[0] not in _WORKER_REGISTRY:
        raise ValueError(f"geometry kind {args[0]!r} has no MI355X implementation; "
                         f"known: {sorted(_WORKER_REGISTRY)}")
    return Geometry.from_worker_args(args)


def make_kill_fn(geometry, z_max=None):
    """geometry.kill_mask, optionally OR z >= z_max (reference
    geometry.py:304-322), evaluated on the device."""
    if z_max is None:
        return geometry.kill_mask
    return lambda pos: geometry._device_kill(pos, z_max=z_max)
