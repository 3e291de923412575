"""CPU pins of the oracle toroidal harmonics and ring-mode kernels
(reference bem/toroidal.py, bem/ring_modes.py) by mathematics."""
import numpy as np
import pytest

from oracle import bem as ob
from oracle import toroidal as ot


def test_Q_against_mpmath_legendre_second_kind():
    mpmath = pytest.importorskip("mpmath")
    mpmath.mp.dps = 30
    chis = np.array([1.0 + 1e-9, 1.0 + 1e-5, 1.001, 1.02, 1.3, 2.0, 7.5, 60.0])
    n_max = 24
    q, dq = ot.toroidal_Q(chis, n_max)
    worst = 0.0
    for i, c in enumerate(chis):
        for n in range(n_max + 1):
            ref = float(mpmath.re(mpmath.legenq(n - 0.5, 0, mpmath.mpf(float(c)), type=3)))
            worst = max(worst, abs(q[n, i] - ref) / abs(ref))
    assert worst < 1e-12, worst


def test_dQ_is_the_derivative():
    chis = np.array([1.0005, 1.02, 1.3, 2.0, 7.5])
    n_max = 12
    _, dq = ot.toroidal_Q(chis, n_max)
    h = 1e-7 * (chis - 1.0)
    qp, _ = ot.toroidal_Q(chis + h, n_max)
    qm, _ = ot.toroidal_Q(chis - h, n_max)
    fd = (qp - qm) / (2 * h)
    assert np.max(np.abs(fd - dq) / np.abs(dq)) < 1e-6


def test_mode_zero_is_the_axisymmetric_ring_potential():
    rng = np.random.default_rng(0)
    rho, z, a, za = rng.uniform(0.1, 2, 50), rng.uniform(-1, 1, 50), rng.uniform(0.1, 2, 50), rng.uniform(-1, 1, 50)
    Phi, _, _ = ot.ring_field_modes(rho, z, a, za, 3)
    ref = 2 * np.pi * a * ob.ring_potential(rho, z - za, a)
    assert np.max(np.abs(Phi[0] - ref) / np.abs(ref)) < 1e-13


def test_ring_mode_field_is_minus_gradient():
    rho, z, a, za = 0.7, 0.3, 1.1, -0.2
    Phi, dPr, dPz = ot.ring_field_modes(rho, z, a, za, 8)
    h = 1e-6
    fr = (ot.ring_field_modes(rho + h, z, a, za, 8)[0] - ot.ring_field_modes(rho - h, z, a, za, 8)[0]) / (2 * h)
    fz = (ot.ring_field_modes(rho, z + h, a, za, 8)[0] - ot.ring_field_modes(rho, z - h, a, za, 8)[0]) / (2 * h)
    assert np.max(np.abs(fr - dPr)) < 1e-7 * np.abs(dPr).max()
    assert np.max(np.abs(fz - dPz)) < 1e-7 * np.abs(dPz).max()


def test_point_charge_mode_series_sums_to_the_point_potential():
    """sum_m g_m cos(m phi) -> 1 / (4 pi |r - r0|) for a unit point charge."""
    rho, z, rho0, z0 = 1.0, 0.4, 1.3, -0.1
    g = ot.point_charge_potential_modes(rho, z, rho0, z0, 60)
    for phi in (0.0, 0.7, 2.0, np.pi):
        s = np.sum(g * np.cos(np.arange(61) * phi))
        d = np.sqrt(rho**2 + rho0**2 - 2 * rho * rho0 * np.cos(phi) + (z - z0) ** 2)
        assert abs(s - 1.0 / (4 * np.pi * d)) < 1e-12 / d, (phi, s, 1 / (4 * np.pi * d))


def test_near_coincidence_is_regularised():
    q, dq = ot.toroidal_Q(np.array([1.0, 0.999999]), 4)
    assert np.all(np.isfinite(q)) and np.all(np.isfinite(dq))
