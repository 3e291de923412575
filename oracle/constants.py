"""CODATA 2022 constants as scipy.constants 1.15 gives them (reference
``constants.py:11-14``); the electron rest energy the reference's openPMD
ParticleGroup supplies (``tests/test_tracker.py:6``). Oracle-side copy."""
import math

ELEMENTARY_CHARGE = 1.602176634e-19
SPEED_OF_LIGHT = 299792458.0
VACUUM_PERMITTIVITY = 8.8541878188e-12
COULOMB_CONSTANT = 1.0 / (4.0 * math.pi * VACUUM_PERMITTIVITY)
ELECTRON_MASS_EV = 510998.95069


def ev_c_to_si_momentum(p):
    return p * ELEMENTARY_CHARGE / SPEED_OF_LIGHT


def si_momentum_to_ev_c(p):
    return p * SPEED_OF_LIGHT / ELEMENTARY_CHARGE


def ev_to_kg(m):
    return m * ELEMENTARY_CHARGE / SPEED_OF_LIGHT**2
