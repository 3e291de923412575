"""Physical constants and unit conversions (reference ``constants.py:11-29``).

CODATA 2022 values, as scipy.constants 1.15 reports them, hard-coded so the
package has no scipy dependency. The tracker works in SI throughout."""
import math

ELEMENTARY_CHARGE = 1.602176634e-19          # C
SPEED_OF_LIGHT = 299792458.0                 # m/s
VACUUM_PERMITTIVITY = 8.8541878188e-12       # F/m
COULOMB_CONSTANT = 1.0 / (4.0 * math.pi * VACUUM_PERMITTIVITY)
ELECTRON_MASS_EV = 510998.95069              # electron rest energy [eV]


def ev_c_to_si_momentum(p_ev_c):
    """[eV/c] -> [kg m/s]."""
    return p_ev_c * ELEMENTARY_CHARGE / SPEED_OF_LIGHT


def si_momentum_to_ev_c(p_si):
    """[kg m/s] -> [eV/c]."""
    return p_si * SPEED_OF_LIGHT / ELEMENTARY_CHARGE


def ev_to_kg(mass_ev):
    """Rest energy [eV] -> rest mass [kg]."""
    return mass_ev * ELEMENTARY_CHARGE / SPEED_OF_LIGHT**2
