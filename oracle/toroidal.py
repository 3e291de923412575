"""Oracle toroidal harmonics and azimuthal-mode ring kernels, FP64 NumPy.

Restates reference ``bem/toroidal.py:70-219`` (Q_{n-1/2}(chi) and
dQ/dchi by elliptic seeds plus a hybrid recurrence) and
``bem/ring_modes.py:60-158`` (mode-m ring potential/field and point-charge
potential coefficients). Test infrastructure only.

Pinned by mathematics (tests/test_oracle_toroidal.py): Q against mpmath's
Legendre function of the second kind at half-integer degree, dQ/dchi
against central differences, the m = 0 ring mode against the m = 0 ring
potential, and the cosine series of the point-charge coefficients
against the exact point-charge potential.
"""
import numpy as np

from .bem import ellip_ke_complement


def elliptic_seeds(chi):
    """Q_{-1/2}, Q_{1/2} and dQ_{-1/2}/dchi from K, E at the complementary
    parameter (chi-1)/(chi+1) (reference ``bem/toroidal.py:174-182``)."""
    p = (chi - 1.0) / (chi + 1.0)
    K, E = ellip_ke_complement(p)
    root = np.sqrt(2.0 / (chi + 1.0))
    q0 = root * K
    q1 = chi * q0 - np.sqrt(2.0 * (chi + 1.0)) * E
    return q0, q1, -root * E / (2.0 * (chi - 1.0))


def toroidal_Q(chi, n_max, pad_min=50, pad_const=40.0):
    """Q_{n-1/2}(chi), dQ/dchi for n = 0..n_max, shape (n_max+1,) +
    chi.shape (reference ``bem/toroidal.py:188-219``): upward recurrence
    where M eta <= 1 (eta = arccosh chi, M = max(n_max, 1)), otherwise the
    downward ratio recurrence started max(pad_min, ceil(pad_const/eta))
    above M, per element; derivatives by the Legendre identity with
    (chi-1)(chi+1)."""
    if not isinstance(n_max, (int, np.integer)) or n_max < 0:
        raise ValueError("n_max must be a nonnegative integer")
    chi = np.maximum(np.asarray(chi, dtype=float), np.nextafter(1.0, 2.0))
    shape = chi.shape
    c = chi.reshape(-1)
    M = max(n_max, 1)
    q0, q1, dq0 = elliptic_seeds(c)
    out = np.empty((M + 1, c.size))
    dq = np.empty((M + 1, c.size))
    eta = np.arccosh(c)
    den = (c - 1.0) * (c + 1.0)
    near = M * eta <= 1.0
    # upward recurrence, all near elements at once
    out[0] = q0
    iu = np.nonzero(near)[0]
    if iu.size:
        cu = c[iu]
        out[1, iu] = q1[iu]
        for n in range(1, M):
            out[n + 1, iu] = (2.0 * n * cu * out[n, iu] - (n - 0.5) * out[n - 1, iu]) / (n + 0.5)
    # downward ratio recurrence, each element from its own start M + pad (masked)
    idn = np.nonzero(~near)[0]
    if idn.size:
        cd = c[idn]
        pad = np.maximum(pad_min, np.ceil(pad_const / eta[idn]).astype(np.int64))
        start = M + pad
        r_next = np.zeros(idn.size)
        r = np.empty((M + 1, idn.size))
        for n in range(int(start.max()), 0, -1):
            on = n <= start
            rn = (n - 0.5) / (2.0 * n * cd - (n + 0.5) * r_next)
            r_next = np.where(on, rn, r_next)
            if n <= M:
                r[n] = rn
        v = q0[idn]
        for n in range(1, M + 1):
            v = v * r[n]
            out[n, idn] = v
    dq[0] = dq0
    for n in range(1, M + 1):
        dq[n] = (n - 0.5) * (c * out[n] - out[n - 1]) / den
    return out[:n_max + 1].reshape((n_max + 1,) + shape), dq[:n_max + 1].reshape((n_max + 1,) + shape)


def chi_of(rho, z, a, za):
    """1 + ((rho-a)^2 + (z-za)^2) / (2 rho a) (reference ``ring_modes.py:60-67``)."""
    return 1.0 + ((rho - a) ** 2 + (z - za) ** 2) / (2.0 * rho * a)


def ring_field_modes(rho, z, a, za, n_max):
    """(Phi_m, dPhi_m/drho, dPhi_m/dz), m = 0..n_max, of a source ring (a,
    za) with density cos(m phi') (reference ``ring_modes.py:76-143``):
    Phi_m = sqrt(a/rho) Q_{m-1/2}(chi) / (2 pi)."""
    rho, z, a, za = np.broadcast_arrays(*(np.asarray(v, dtype=float) for v in (rho, z, a, za)))
    q, dq = toroidal_Q(chi_of(rho, z, a, za), n_max)
    dchi_drho = (rho**2 - a**2 - (z - za) ** 2) / (2.0 * a * rho**2)
    dchi_dz = (z - za) / (a * rho)
    s = np.sqrt(a / rho)
    Phi = s * q / (2.0 * np.pi)
    dPr = (-s / (2.0 * rho) * q + s * dq * dchi_drho) / (2.0 * np.pi)
    dPz = (s * dq * dchi_dz) / (2.0 * np.pi)
    return Phi, dPr, dPz


def point_charge_potential_modes(rho, z, rho0, z0, n_max):
    """g_m = C_m Q_{m-1/2}(chi) / sqrt(rho rho0), C_0 = 1/(4 pi^2), C_m =
    1/(2 pi^2) (reference ``ring_modes.py:146-158``)."""
    rho, z, rho0, z0 = np.broadcast_arrays(*(np.asarray(v, dtype=float) for v in (rho, z, rho0, z0)))
    q, _ = toroidal_Q(chi_of(rho, z, rho0, z0), n_max)
    c = np.full(n_max + 1, 1.0 / (2.0 * np.pi**2))
    c[0] = 1.0 / (4.0 * np.pi**2)
    return c.reshape((n_max + 1,) + (1,) * rho.ndim) * q / np.sqrt(rho * rho0)[None, ...]
