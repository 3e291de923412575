"""ParticleGroup: openPMD-beamphysics' container when it is installed,
otherwise a minimal data-only stand-in with the attributes the tracker
reads and writes (reference tracker.py:215-284,388-417)."""
import numpy as np

from .constants import ELECTRON_MASS_EV, ELEMENTARY_CHARGE

try:  # pragma: no cover - exercised only where openPMD-beamphysics exists
    from beamphysics import ParticleGroup  # noqa: F401
except ImportError:  # pragma: no cover
    try:
        from pmd_beamphysics import ParticleGroup  # noqa: F401
    except ImportError:

        class ParticleGroup:
            """Arrays x, y, z [m], px, py, pz [eV/c], t [s], weight [C],
            status, id, plus species (electrons only)."""

            _KEYS = ("x", "y", "z", "px", "py", "pz", "t", "weight", "status", "id")

            def __init__(self, data):
                self.species = data.get("species", "electron")
                if self.species != "electron":
                    raise ValueError("only electrons are supported by this stand-in")
                for k in self._KEYS:
                    setattr(self, k, np.asarray(data[k]))

            @property
            def mass(self):
                return ELECTRON_MASS_EV

            @property
            def species_charge(self):
                return -ELEMENTARY_CHARGE

            def __len__(self):
                return len(self.x)
