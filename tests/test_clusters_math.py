"""Axis-centred cluster expansions (bem_amd.bem.clusters) on the CPU: each
cluster's expansion against the direct sum of its fixed-rule node rings
through the oracle's ring field (reference bem/axisymmetric.py:232-265),
on the 400-panel post with the reference's sigma. No GPU."""
import numpy as np
import pytest

from conftest import golden
from bem_amd.bem import clusters as bc
from bem_amd.bem.multipole import quadrature_nodes
from oracle import bem as ob


@pytest.fixture(scope="module")
def post():
    d = golden("bem_post400.npz")
    return d["profile"], d["sigma"], float(d["R"]), float(d["H"])


def _points(R, H, n, seed):
    """Around the cap (1 nm .. 10 um off it) and beside the post wall."""
    rng = np.random.default_rng(seed)
    th = 0.5 * np.pi * rng.random(n)
    d = 10 ** rng.uniform(-9, -5, n)
    rho, z = (R + d) * np.sin(th), H + (R + d) * np.cos(th)
    m = n // 4
    rho[:m] = R + 10 ** rng.uniform(-9, -5.5, m)
    z[:m] = rng.uniform(0.5e-6, 10e-6, m)
    return rho, z


def _direct(cl, c, prof, sig, rho, z):
    rq, zq, wq = quadrature_nodes(prof, sig, 8)
    sl = slice(cl.seg_begin[c] * 32, cl.seg_begin[c + 1] * 32)
    er, ez = ob.ring_field(rho[:, None], z[:, None] - zq[None, sl], rq[None, sl])
    return (er * wq[None, sl]).sum(1), (ez * wq[None, sl]).sum(1)


def test_clusters_cover_the_profile(post):
    prof, sig, R, H = post
    cl = bc.build(prof, sig)
    n_seg = len(prof) - 1
    assert cl.seg_begin[0] == 0 and cl.seg_begin[-1] == n_seg
    sizes = np.diff(cl.seg_begin)
    assert np.all(sizes >= 12) and cl.n_clusters == 16
    rq, zq, _ = quadrature_nodes(prof, sig, 8)
    for c in range(cl.n_clusters):  # every node of a cluster inside its sphere
        sl = slice(cl.seg_begin[c] * 32, cl.seg_begin[c + 1] * 32)
        assert np.all(np.hypot(rq[sl], zq[sl] - cl.zc[c]) <= cl.rc[c])


def test_expansions_equal_the_direct_ring_sums(post):
    """Beyond kappa r_c every cluster's truncated expansion equals the sum of
    its fixed-rule node rings to 1e-12 of |E| (measured 1.5e-14)."""
    prof, sig, R, H = post
    cl = bc.build(prof, sig)
    rho, z = _points(R, H, 1500, 3)
    ref = ob.evaluate_field(prof, sig, np.stack([rho, 0 * rho, z], -1))
    Et = np.hypot(ref[:, 0], ref[:, 2] - 1e8)
    n_far = 0
    for c in range(cl.n_clusters):
        far = np.hypot(rho, z - cl.zc[c]) >= cl.kappa * cl.rc[c]
        if not far.any():
            continue
        n_far += int(far.sum())
        dr, dz = _direct(cl, c, prof, sig, rho[far], z[far])
        gr, gz = bc.field(cl, c, rho[far], z[far])
        assert np.max(np.hypot(gr - dr, gz - dz) / Et[far]) <= 1e-12, c
    assert n_far > 10 * 1500  # most (point, cluster) pairs are far


def test_a_displaced_centre_is_caught(post):
    """A planted mistake (moments about one centre, evaluated about another)
    shows up far above the tolerance."""
    prof, sig, R, H = post
    cl = bc.build(prof, sig)
    c = 5
    rho, z = _points(R, H, 400, 4)
    far = np.hypot(rho, z - cl.zc[c]) >= cl.kappa * cl.rc[c] + 0.1 * cl.rc[c]
    dr, dz = _direct(cl, c, prof, sig, rho[far], z[far])
    cl.zc = cl.zc.copy()
    cl.zc[c] += 0.05 * cl.rc[c]
    gr, gz = bc.field(cl, c, rho[far], z[far])
    assert np.max(np.hypot(gr - dr, gz - dz) / np.hypot(dr, dz)) > 1e-3


def test_truncation_rule_bounds_the_tail():
    """terms_needed(q) is the smallest N with sum_{l >= N} (l + 1) q^l <= tol,
    against a brute-force sum."""
    for q in (0.1, 0.3, 0.5, 0.625, 0.7):
        N = int(bc.terms_needed(np.array([q]), 1e-11, 256)[0])
        tail = lambda n: sum((l + 1) * q ** l for l in range(n, 4000))  # noqa: E731
        assert tail(N) <= 1e-11 < tail(N - 1), q


def test_legendre_bounds_behind_the_tail_rules():
    """The truncation rules of the cluster expansions and of the multipole
    zone bound both field components by sum (l + 1) q^l |m_l|: they need
    |P_l| <= 1 and |sin t P_l'(cos t)| <= l + 1 on [-1, 1] (l <= 256, on a
    fine grid; the largest ratio is about 0.58)."""
    x = np.linspace(-1.0, 1.0, 100001)
    P = np.zeros((258, x.size))
    P[0], P[1] = 1.0, x
    for l in range(1, 257):
        P[l + 1] = ((2 * l + 1) * x * P[l] - l * P[l - 1]) / (l + 1)
    dP = np.zeros_like(P)
    dP[1] = 1.0
    for l in range(1, 257):
        dP[l + 1] = dP[l - 1] + (2 * l + 1) * P[l]
    s = np.sqrt(1.0 - x * x)
    assert np.abs(P).max() <= 1.0 + 1e-12
    ratio = max(np.abs(s * dP[l]).max() / (l + 1) for l in range(258))
    assert ratio <= 1.0, ratio


def test_multipole_truncation_is_far_below_the_budget():
    """multipole_pert stops where sum_{k > l} (k + 1) q^k <= 1e-12, so each
    dropped component is below 1e-12 max|m_l| / 4 pi r^2; for the 400-panel
    post's charge at its validated radius (1.083 r_ref) that is below
    1e-12 of E_gun (1e8 V/m), six orders under the 1e-6 field budget."""
    from bem_amd.bem import multipole as mpl

    d = golden("bem_post400.npz")
    m, r_ref = mpl.moments(d["profile"], d["sigma"], 8, 128)
    bound = 1e-12 * np.abs(m).max() / (4 * np.pi * (1.083 * r_ref) ** 2)
    assert bound <= 1e-12 * 1e8, bound
