"""MI355X-native backend for the specific_particle_tracer tracing path.

Public names mirror the reference package's (reference __init__.py:1-12);
BEM geometries live in ``bem_amd.bem.geometry``."""
from .fields import GunField, HemisphericalTipField
from .geometry import FlatCathode, Geometry, HemisphericalTip
from .tracker import SpecificParticleTracer

__all__ = [
    "SpecificParticleTracer",
    "GunField",
    "HemisphericalTipField",
    "Geometry",
    "FlatCathode",
    "HemisphericalTip",
]
