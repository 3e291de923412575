"""BEM-backed geometries (reference bem/geometry.py:58-893).

The static field is the uniform asymptotic field plus the perturbation of
an axisymmetric BEM solve (a mirror-doubled closed profile for
protrusions, the real truncated-plane surface for the well), zero inside
the conductor, blended with a cubic smoothstep to the plain flat-cathode
field between z_handoff/2 and z_handoff (reference bem/geometry.py:44-55,
319-324). The collocation solve runs on the device at construction.

With ``z0`` set, the well and the hemispherical tip take the image-charge
BEM force (reference bem/image_charge.py, bem/geometry.py:326-427,786-801)
on their recessed image surfaces; the post has no image surface in the
reference and raises ValueError for ``z0``.
"""
import numpy as np

from .. import _lib
from .. import geometry as geometry_module
from .mesh import (cylindrical_well_image_profile, cylindrical_well_profile, hemisphere_tip_image_doubled_profile,
                   hemisphere_tip_image_profile, n_for_length, post_profile, sphere_profile)
from .solver import AxisymmetricBEMSolution


class _AxisymmetricBEMGeometry(geometry_module.Geometry):
    H = 0.0
    z0 = None
    image_profile = None  # recessed image surface (rho, z) when z0 is set
    image_mode = _lib.IMAGE_BEM
    image_n_max = 16
    image_d_lo = 0.1
    image_d_hi = 0.5

    def _setup(self, E_gun, R, profile, n_subdiv, sigma, z_handoff, kill_z_below, device):
        self.E_gun = float(E_gun)
        self.R = float(R)
        self.n_subdiv = int(n_subdiv)
        self.kill_z_below = kill_z_below
        self.device = device
        self.z_handoff = float(z_handoff)
        if sigma is None:
            self.solution = AxisymmetricBEMSolution.solve(profile, self.E_gun, device=device)
        else:
            self.solution = AxisymmetricBEMSolution(profile, sigma, device=device)
        self._flat_fallback = geometry_module.FlatCathode(E_gun, z0=None, kill_z_below=kill_z_below, device=device)

    @property
    def real_profile(self):
        return self.solution.profile

    def bem_table(self):
        return self.solution.profile, self.solution.sigma, self.n_subdiv

    def _ctx(self):
        ctx = super()._ctx()
        if self.image_profile is not None:
            key = ("image", _lib.token(self), self.image_profile.shape[0], self.image_n_max)
            if getattr(ctx, "image_key", None) != key:
                prof, pp = _lib.host_ptr(np.asarray(self.image_profile, dtype=np.float64))
                ctx.check(ctx.lib.spt_image_upload(ctx.h, pp, prof.shape[0], self.image_n_max, 4,
                                                   self.image_d_lo * self.R, self.image_d_hi * self.R, None),
                          "spt_image_upload")
                ctx.image_key = key
        return ctx

    def image_operators(self):
        """The per-mode operators (n_max + 1, P, P) as assembled on the device."""
        ctx = super()._ctx()
        prof, pp = _lib.host_ptr(np.asarray(self.image_profile, dtype=np.float64))
        P = prof.shape[0]
        A = np.zeros((self.image_n_max + 1, P, P))
        _, ap = _lib.host_ptr(A)
        ctx.check(ctx.lib.spt_image_upload(ctx.h, pp, P, self.image_n_max, 4, self.image_d_lo * self.R,
                                           self.image_d_hi * self.R, ap), "spt_image_upload")
        ctx.image_key = ("image", _lib.token(self), P, self.image_n_max)
        return A

    def spt_geometry(self, z_max=None):
        g = _lib.Geometry_t()
        g.kind = _lib.GEOM_BEM_POST
        g.image = _lib.IMAGE_NONE
        if self.z0 is not None and self.image_profile is not None:
            g.image = self.image_mode
            g.z0 = self.z0
        g.E_gun = self.E_gun
        g.R = self.R
        g.H = float(self.H)
        g.z_handoff = self.z_handoff
        g.has_kill_z = 0 if self.kill_z_below is None else 1
        g.kill_z_below = 0.0 if self.kill_z_below is None else float(self.kill_z_below)
        g.has_z_max = 0 if z_max is None else 1
        g.z_max = 0.0 if z_max is None else float(z_max)
        return g

    def field(self, position):
        return self._device_field(position)

    def image_force(self, position, charge, active, plummer_radius):
        if self.z0 is None:
            # the reference returns the flat fallback's image force, zero for z0 = None
            return np.zeros_like(position) if not hasattr(position, "detach") else position.new_zeros(position.shape)
        import torch

        ctx = self._ctx()
        g = self.spt_geometry()
        is_np = not hasattr(position, "detach")
        P = torch.as_tensor(np.asarray(position) if is_np else position, dtype=torch.float64, device=ctx.torch_device)
        G, E, _ = P.shape
        q = torch.as_tensor(np.asarray(charge) if is_np else charge, dtype=torch.float64, device=ctx.torch_device)
        a = torch.as_tensor(np.asarray(active) if is_np else active, device=ctx.torch_device).to(torch.uint8)
        P, q, a = P.reshape(-1, 3).contiguous(), q.reshape(-1).contiguous(), a.reshape(-1).contiguous()
        off = torch.arange(0, G * E + 1, E, dtype=torch.int64, device=ctx.torch_device)
        out = torch.zeros_like(P)
        ctx.bind_stream()
        ctx.check(ctx.lib.spt_image_force(ctx.h, g, _lib.ptr(P), _lib.ptr(q), _lib.ptr(a), _lib.ptr(off), G,
                                          float(plummer_radius), _lib.ptr(out)), "spt_image_force")
        ctx.sync()
        out = out.reshape(G, E, 3)
        return out.cpu().numpy() if is_np else out

    def kill_mask(self, position):
        return self._device_kill(position)


def _no_image(z0):
    if z0 is not None:
        raise ValueError("the post has no image-charge surface (z0 must be None); the hemispherical tip and the "
                         "well take z0")


class HemisphericalTipBEMGeometry(_AxisymmetricBEMGeometry):
    """Grounded hemispherical tip of radius R on a flat cathode, BEM field
    on the full sphere (reference bem/geometry.py:58-507, bem/fields.py:139-211).

    With ``z0`` set (0 < z0 < R), the image force is the image-charge BEM on
    the recessed image surface: the cap of radius R - z0 with its ridge
    fillet and either the plane at z = -z0 out to 5 R
    (``hemisphere_tip_image_profile``) or, with ``image_mirror_symmetric``,
    the same cap and fillet closed by their reflection through z = -z0
    (``hemisphere_tip_image_doubled_profile``, a phantom -q per particle);
    segments <= (R - z0) pi / 58 by default, ``image_n_max`` = 16 modes,
    mirror-charge blend between ``image_d_lo`` R and ``image_d_hi`` R,
    blended with the plane image at z0 over [z_handoff/2, z_handoff]
    (reference bem/geometry.py:190-344,346-427). Traces with it run in
    lockstep mode.

    ``sigma`` may be passed to reuse an existing solve of the same profile."""

    def __init__(self, E_gun, R, field_max_length=None, n_subdiv=8, z0=None, image_max_length=None,
                 image_cap_max_length=None, image_fillet_max_length=None, image_plane_max_length=None,
                 image_n_max=16, image_d_lo=0.1, image_d_hi=0.5, image_mirror_symmetric=False, z_handoff=None,
                 kill_z_below=0.0, sigma=None, device=0):
        self.field_max_length = field_max_length
        R = float(R)
        self.z0 = None if z0 is None else float(z0)
        self.image_mirror_symmetric = bool(image_mirror_symmetric)
        self.image_max_length = (((R - self.z0) * np.pi / 2.0 / 29.0 if self.z0 is not None else None)
                                 if image_max_length is None else float(image_max_length))
        self.image_cap_max_length = image_cap_max_length
        self.image_fillet_max_length = image_fillet_max_length
        self.image_plane_max_length = image_plane_max_length
        self.image_n_max = int(image_n_max)
        self.image_d_lo = float(image_d_lo)
        self.image_d_hi = float(image_d_hi)
        max_length = R * np.pi / 39.0 if field_max_length is None else field_max_length
        profile = sphere_profile(R, n_for_length(R * np.pi, max_length))
        zh = 20.0 * R if z_handoff is None else z_handoff
        self._setup(E_gun, R, profile, n_subdiv, sigma, zh, kill_z_below, device)
        if self.z0 is not None:
            if self.image_mirror_symmetric:
                self.image_profile = hemisphere_tip_image_doubled_profile(
                    R, self.z0, self.image_max_length, cap_max_length=image_cap_max_length,
                    fillet_max_length=image_fillet_max_length)
                self.image_mode = _lib.IMAGE_BEM_MIRROR
            else:
                self.image_profile = hemisphere_tip_image_profile(
                    R, self.z0, 5.0 * R, self.image_max_length, cap_max_length=image_cap_max_length,
                    fillet_max_length=image_fillet_max_length, plane_max_length=image_plane_max_length)
                self.image_mode = _lib.IMAGE_BEM
            self._flat_fallback = geometry_module.FlatCathode(E_gun, z0=self.z0, kill_z_below=kill_z_below,
                                                              device=device)

    def worker_args(self):
        return ("hemispherical_tip_bem", self.E_gun, self.R, self.field_max_length, self.n_subdiv, self.z0,
                self.image_max_length, self.image_cap_max_length, self.image_fillet_max_length,
                self.image_plane_max_length, self.image_n_max, self.image_d_lo, self.image_d_hi,
                self.image_mirror_symmetric, self.z_handoff, self.kill_z_below)

    @classmethod
    def _from_worker_args(cls, args):
        (E_gun, R, fml, n_subdiv, z0, iml, icml, ifml, ipml, inm, idl, idh, ims, z_handoff, kill_z_below) = args
        return cls(E_gun, R, field_max_length=fml, n_subdiv=n_subdiv, z0=z0, image_max_length=iml,
                   image_cap_max_length=icml, image_fillet_max_length=ifml, image_plane_max_length=ipml,
                   image_n_max=inm, image_d_lo=idl, image_d_hi=idh, image_mirror_symmetric=ims,
                   z_handoff=z_handoff, kill_z_below=kill_z_below)


class PostTipBEMGeometry(_AxisymmetricBEMGeometry):
    """Hemisphere-capped cylindrical post on a grounded flat cathode: post
    radius R, cap centre at height H (apex at H + R). Solved on the
    mirror-doubled capsule with ``n_segments`` panels (``post_profile``).
    No counterpart in the reference; it reduces to
    HemisphericalTipBEMGeometry's conventions (mask r >= R(1 - 1e-6),
    z >= 0; kill inside or at z <= kill_z_below) with r measured from the
    post axis segment.

    ``z_handoff`` defaults to 20 (R + H), the feature-height analogue of
    the reference's 20 R."""

    def __init__(self, E_gun, R, H, n_segments=400, n_subdiv=8, z0=None, z_handoff=None, kill_z_below=0.0,
                 sigma=None, device=0):
        _no_image(z0)
        self.z0 = None
        self.H = float(H)
        self.n_segments = int(n_segments)
        profile = post_profile(float(R), self.H, self.n_segments)
        zh = 20.0 * (float(R) + self.H) if z_handoff is None else z_handoff
        self._setup(E_gun, R, profile, n_subdiv, sigma, zh, kill_z_below, device)

    def worker_args(self):
        return ("post_tip_bem", self.E_gun, self.R, self.H, self.n_segments, self.n_subdiv, None, self.z_handoff,
                self.kill_z_below)

    @classmethod
    def _from_worker_args(cls, args):
        E_gun, R, H, n_segments, n_subdiv, z0, z_handoff, kill_z_below = args
        return cls(E_gun, R, H, n_segments=n_segments, n_subdiv=n_subdiv, z0=z0, z_handoff=z_handoff,
                   kill_z_below=kill_z_below)


geometry_module._WORKER_REGISTRY["hemispherical_tip_bem"] = HemisphericalTipBEMGeometry
geometry_module._WORKER_REGISTRY["post_tip_bem"] = PostTipBEMGeometry


class CylindricalWellBEMGeometry(_AxisymmetricBEMGeometry):
    """Grounded cylindrical well of radius R and depth H cut into a flat
    cathode (reference bem/geometry.py:510-812, bem/fields.py:240-352): BEM
    on the real surface (bottom disk, wall, graded plane out to
    ``field_plane_radius``, optional rounded corners), zero field inside
    the conductor, the same flat-cathode blend; particles are killed only
    when they re-enter the conductor. Default ``z_handoff`` = 20 max(R, H),
    ``field_plane_radius`` = z_handoff, ``field_max_length`` = min(R, H)/19
    (the reference's defaults).

    With ``z0`` set, the image force is the image-charge BEM on the
    recessed image surface (``cylindrical_well_image_profile``, plane out to
    ``image_plane_radius`` = 5 max(R, H), segments <= min(R, H)/14,
    ``image_n_max`` = 16 azimuthal modes, mirror-charge blend between
    ``image_d_lo`` R and ``image_d_hi`` R), blended with the plane image at
    z0 over [z_handoff/2, z_handoff]; the corners then default to rounded
    (bottom min(z0/2, (H - z0)/5), rim z0 plus that), as in the reference.
    Traces with it run in lockstep mode."""

    def __init__(self, E_gun, R, H, field_plane_radius=None, field_max_length=None, n_subdiv=8,
                 bottom_fillet_radius=None, rim_fillet_radius=None, z0=None, image_plane_radius=None,
                 image_max_length=None, image_bottom_max_length=None, image_fillet_max_length=None,
                 image_wall_max_length=None, image_rim_fillet_max_length=None, image_plane_max_length=None,
                 image_n_max=16, image_d_lo=0.1, image_d_hi=0.5, z_handoff=None, sigma=None, device=0):
        self.z0 = None if z0 is None else float(z0)
        self.H = float(H)
        R = float(R)
        zh = 20.0 * max(R, self.H) if z_handoff is None else float(z_handoff)
        self.field_plane_radius = zh if field_plane_radius is None else float(field_plane_radius)
        self.field_max_length = field_max_length
        if self.z0 is not None:
            avail = self.H - self.z0
            r = min(0.5 * self.z0, 0.2 * avail)
            bottom_fillet_radius = r if bottom_fillet_radius is None else float(bottom_fillet_radius)
            rim_fillet_radius = self.z0 + r if rim_fillet_radius is None else float(rim_fillet_radius)
        self.bottom_fillet_radius = bottom_fillet_radius
        self.rim_fillet_radius = rim_fillet_radius
        max_length = min(R, self.H) / 19.0 if field_max_length is None else float(field_max_length)
        profile = cylindrical_well_profile(R, self.H, self.field_plane_radius, max_length,
                                           bottom_fillet_radius=bottom_fillet_radius,
                                           rim_fillet_radius=rim_fillet_radius)
        self._setup(E_gun, R, profile, n_subdiv, sigma, zh, None, device)
        self.image_plane_radius = 5.0 * max(R, self.H) if image_plane_radius is None else float(image_plane_radius)
        self.image_max_length = min(R, self.H) / 14.0 if image_max_length is None else float(image_max_length)
        self.image_bottom_max_length = image_bottom_max_length
        self.image_fillet_max_length = image_fillet_max_length
        self.image_wall_max_length = image_wall_max_length
        self.image_rim_fillet_max_length = image_rim_fillet_max_length
        self.image_plane_max_length = image_plane_max_length
        self.image_n_max = int(image_n_max)
        self.image_d_lo = float(image_d_lo)
        self.image_d_hi = float(image_d_hi)
        self.image_profile = None
        if self.z0 is not None:
            self.image_profile = cylindrical_well_image_profile(
                R, self.H, self.z0, self.image_plane_radius, self.image_max_length,
                bottom_fillet_radius=bottom_fillet_radius, rim_fillet_radius=rim_fillet_radius,
                bottom_max_length=image_bottom_max_length, bottom_fillet_max_length=image_fillet_max_length,
                wall_max_length=image_wall_max_length, rim_fillet_max_length=image_rim_fillet_max_length,
                plane_max_length=image_plane_max_length)
            self._flat_fallback = geometry_module.FlatCathode(E_gun, z0=self.z0, kill_z_below=None, device=device)

    def spt_geometry(self, z_max=None):
        g = super().spt_geometry(z_max)
        g.kind = _lib.GEOM_BEM_WELL
        g.bottom_fillet = float(self.bottom_fillet_radius or 0.0)
        g.rim_fillet = float(self.rim_fillet_radius or 0.0)
        return g

    def worker_args(self):
        """The reference's 21-element tuple (reference bem/geometry.py:814-836),
        field for field."""
        return ("cylindrical_well_bem", self.E_gun, self.R, self.H, self.field_plane_radius, self.field_max_length,
                self.n_subdiv, self.bottom_fillet_radius, self.rim_fillet_radius, self.z0, self.image_plane_radius,
                self.image_max_length, self.image_bottom_max_length, self.image_fillet_max_length,
                self.image_wall_max_length, self.image_rim_fillet_max_length, self.image_plane_max_length,
                self.image_n_max, self.image_d_lo, self.image_d_hi, self.z_handoff)

    @classmethod
    def _from_worker_args(cls, args):
        (E_gun, R, H, fpr, fml, n_subdiv, bf, rf, z0, ipr, iml, ibml, ifml, iwml, irml, ipml, inm, idl, idh,
         z_handoff) = args
        return cls(E_gun, R, H, field_plane_radius=fpr, field_max_length=fml, n_subdiv=n_subdiv,
                   bottom_fillet_radius=bf, rim_fillet_radius=rf, z0=z0, image_plane_radius=ipr,
                   image_max_length=iml, image_bottom_max_length=ibml, image_fillet_max_length=ifml,
                   image_wall_max_length=iwml, image_rim_fillet_max_length=irml, image_plane_max_length=ipml,
                   image_n_max=inm, image_d_lo=idl, image_d_hi=idh, z_handoff=z_handoff)


geometry_module._WORKER_REGISTRY["cylindrical_well_bem"] = CylindricalWellBEMGeometry
