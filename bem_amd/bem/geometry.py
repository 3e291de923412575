"""BEM-backed geometries (reference bem/geometry.py:58-893).

The static field is the uniform asymptotic field plus the perturbation of
an axisymmetric BEM solve (a mirror-doubled closed profile for
protrusions, the real truncated-plane surface for the well), zero inside
the conductor, blended with a cubic smoothstep to the plain flat-cathode
field between z_handoff/2 and z_handoff (reference bem/geometry.py:44-55,
319-324). The collocation solve runs on the device at construction.

Image-charge BEM (``z0`` not None, reference bem/image_charge.py) is not
built on this backend yet: asking for it raises ValueError rather than
silently dropping the force.
"""
import numpy as np

from .. import _lib
from .. import geometry as geometry_module
from .mesh import cylindrical_well_profile, n_for_length, post_profile, sphere_profile
from .solver import AxisymmetricBEMSolution


class _AxisymmetricBEMGeometry(geometry_module.Geometry):
    H = 0.0

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

    def spt_geometry(self, z_max=None):
        g = _lib.Geometry_t()
        g.kind = _lib.GEOM_BEM_POST
        g.image = _lib.IMAGE_NONE
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
        # z0 is None on this backend: the reference returns the flat
        # fallback's image force, which is zero for z0 = None.
        return np.zeros_like(position) if not hasattr(position, "detach") else position.new_zeros(position.shape)

    def kill_mask(self, position):
        return self._device_kill(position)


def _no_image(z0):
    if z0 is not None:
        raise ValueError("image-charge BEM (z0 not None) is not implemented on the mi355x backend yet")


class HemisphericalTipBEMGeometry(_AxisymmetricBEMGeometry):
    """Grounded hemispherical tip of radius R on a flat cathode, BEM field
    on the full sphere (reference bem/geometry.py:58-507, bem/fields.py:139-211).

    ``sigma`` may be passed to reuse an existing solve of the same profile."""

    def __init__(self, E_gun, R, field_max_length=None, n_subdiv=8, z0=None, z_handoff=None, kill_z_below=0.0,
                 sigma=None, device=0):
        _no_image(z0)
        self.field_max_length = field_max_length
        self.z0 = None
        R = float(R)
        max_length = R * np.pi / 39.0 if field_max_length is None else field_max_length
        profile = sphere_profile(R, n_for_length(R * np.pi, max_length))
        zh = 20.0 * R if z_handoff is None else z_handoff
        self._setup(E_gun, R, profile, n_subdiv, sigma, zh, kill_z_below, device)

    def worker_args(self):
        return ("hemispherical_tip_bem", self.E_gun, self.R, self.field_max_length, self.n_subdiv, None,
                self.z_handoff, self.kill_z_below)

    @classmethod
    def _from_worker_args(cls, args):
        E_gun, R, fml, n_subdiv, z0, z_handoff, kill_z_below = args
        return cls(E_gun, R, field_max_length=fml, n_subdiv=n_subdiv, z0=z0, z_handoff=z_handoff,
                   kill_z_below=kill_z_below)


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
    (the reference's defaults)."""

    def __init__(self, E_gun, R, H, field_plane_radius=None, field_max_length=None, n_subdiv=8,
                 bottom_fillet_radius=None, rim_fillet_radius=None, z0=None, z_handoff=None, sigma=None, device=0):
        _no_image(z0)
        self.z0 = None
        self.H = float(H)
        R = float(R)
        zh = 20.0 * max(R, self.H) if z_handoff is None else float(z_handoff)
        self.field_plane_radius = zh if field_plane_radius is None else float(field_plane_radius)
        self.field_max_length = field_max_length
        self.bottom_fillet_radius = bottom_fillet_radius
        self.rim_fillet_radius = rim_fillet_radius
        max_length = min(R, self.H) / 19.0 if field_max_length is None else float(field_max_length)
        profile = cylindrical_well_profile(R, self.H, self.field_plane_radius, max_length,
                                           bottom_fillet_radius=bottom_fillet_radius,
                                           rim_fillet_radius=rim_fillet_radius)
        self._setup(E_gun, R, profile, n_subdiv, sigma, zh, None, device)

    def spt_geometry(self, z_max=None):
        g = super().spt_geometry(z_max)
        g.kind = _lib.GEOM_BEM_WELL
        g.bottom_fillet = float(self.bottom_fillet_radius or 0.0)
        g.rim_fillet = float(self.rim_fillet_radius or 0.0)
        return g

    def worker_args(self):
        return ("cylindrical_well_bem", self.E_gun, self.R, self.H, self.field_plane_radius, self.field_max_length,
                self.n_subdiv, self.bottom_fillet_radius, self.rim_fillet_radius, None, self.z_handoff)

    @classmethod
    def _from_worker_args(cls, args):
        E_gun, R, H, fpr, fml, n_subdiv, bf, rf, z0, z_handoff = args
        return cls(E_gun, R, H, field_plane_radius=fpr, field_max_length=fml, n_subdiv=n_subdiv,
                   bottom_fillet_radius=bf, rim_fillet_radius=rf, z0=z0, z_handoff=z_handoff)


geometry_module._WORKER_REGISTRY["cylindrical_well_bem"] = CylindricalWellBEMGeometry
