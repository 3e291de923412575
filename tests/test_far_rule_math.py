"""The distance-adaptive far rule against the reference's fixed rule, on the
CPU with the oracle's ring kernel (the mathematics the GPU far rule relies
on; the GPU gather itself is checked in test_gpu_far_rule.py).

On the 400-panel post (reference-solved sigma): Gauss-8 over a whole
segment at distance L <= D < 2 L, Gauss-4 at 2 L <= D < 16 L and Gauss-2
beyond change the field by < 1e-7 of |E| from the 8 x Gauss-4 sub-panel
rule, while the midpoint rule at 16 L does not (> 1e-5): the choice of
rule matters and the test would fail for a dropped Gauss node or a wrong
weight.
"""
import numpy as np

from conftest import golden
from oracle import bem as ob

G4X = np.array([-0.8611363115940526, -0.33998104358485626, 0.33998104358485626, 0.8611363115940526])
G4W = np.array([0.3478548451374537, 0.6521451548625462, 0.6521451548625462, 0.3478548451374537])


def _segment_nodes(prof, sig, xs, ws):
    """Gauss nodes over whole segments: (rho, z, area * sigma), (n_seg, k)."""
    p0, p1 = prof[:-1], prof[1:]
    L = np.linalg.norm(p1 - p0, axis=1)
    t = 0.5 * (xs + 1.0)
    r = p0[:, :1] + t[None] * (p1 - p0)[:, :1]
    z = p0[:, 1:] + t[None] * (p1 - p0)[:, 1:]
    s = sig[:-1, None] + t[None] * (sig[1:] - sig[:-1])[:, None]
    return r, z, 2.0 * np.pi * r * (L[:, None] * 0.5 * ws[None]) * s


def _field_parts(nodes, rho, z):
    r, zq, w = nodes
    er, ez = ob.ring_field(rho[:, None, None], z[:, None, None] - zq[None], r[None])
    return (er * w[None]).sum(-1), (ez * w[None]).sum(-1)


def _seg_dist(prof, rho, z):
    p0, t = prof[:-1], prof[1:] - prof[:-1]
    tt = np.clip(((rho[:, None] - p0[None, :, 0]) * t[None, :, 0] + (z[:, None] - p0[None, :, 1]) * t[None, :, 1])
                 / (t ** 2).sum(1)[None], 0.0, 1.0)
    return np.hypot(rho[:, None] - (p0[None, :, 0] + tt * t[None, :, 0]),
                    z[:, None] - (p0[None, :, 1] + tt * t[None, :, 1]))


def test_far_rule_against_fixed_rule_on_the_post():
    d = golden("bem_post400.npz")
    prof, sig, Ez, R, H = d["profile"], d["sigma"], float(d["Ez"]), float(d["R"]), float(d["H"])
    n_seg = len(prof) - 1
    L = np.linalg.norm(np.diff(prof, axis=0), axis=1)
    r, zq, a, s = ob.fixed_quadrature(prof, sig, 8)
    fine = (r.reshape(n_seg, -1), zq.reshape(n_seg, -1), (a * s).reshape(n_seg, -1))
    g8 = _segment_nodes(prof, sig, *np.polynomial.legendre.leggauss(8))
    g4 = _segment_nodes(prof, sig, G4X, G4W)
    g2 = _segment_nodes(prof, sig, np.array([-1.0, 1.0]) / np.sqrt(3.0), np.array([1.0, 1.0]))
    g1 = _segment_nodes(prof, sig, np.array([0.0]), np.array([2.0]))
    rng = np.random.default_rng(5)
    n = 240
    th = rng.uniform(0.0, np.deg2rad(100.0), n)
    dd = np.exp(rng.uniform(np.log(0.1e-9), np.log(50e-6), n))
    rho, z = (R + dd) * np.sin(th), H + (R + dd) * np.cos(th)
    D = _seg_dist(prof, rho, z)
    fr, fz = _field_parts(fine, rho, z)
    far = D >= L[None]
    Et = np.hypot(np.where(far, fr, 0).sum(1), np.where(far, fz, 0).sum(1) + Ez)
    m8 = far & (D < 2 * L[None])
    m4 = (D >= 2 * L[None]) & (D < 16 * L[None])
    m2 = D >= 16 * L[None]
    c8r, c8z = _field_parts(g8, rho, z)
    c4r, c4z = _field_parts(g4, rho, z)
    c2r, c2z = _field_parts(g2, rho, z)
    c1r, c1z = _field_parts(g1, rho, z)
    er = sum(np.where(m, cr - fr, 0).sum(1) for m, cr in ((m8, c8r), (m4, c4r), (m2, c2r)))
    ez = sum(np.where(m, cz - fz, 0).sum(1) for m, cz in ((m8, c8z), (m4, c4z), (m2, c2z)))
    err = np.hypot(er, ez) / Et
    assert err.max() < 1e-7
    assert np.count_nonzero(m8) > 0  # every band is exercised
    assert np.count_nonzero(m2) > 0.5 * m2.size  # the coarse rules do most of the work
    mid = np.hypot(np.where(m2, c1r - fr, 0).sum(1), np.where(m2, c1z - fz, 0).sum(1)) / Et
    assert mid.max() > 1e-5


def test_gauss4_whole_segment_converges_with_distance():
    """One straight segment, sigma linear: Gauss-4 over the whole segment
    against the 8 x Gauss-4 rule falls by more than 4^6 when the distance
    quadruples (Gauss-n error ~ (4 D / L)^(-2n))."""
    prof = np.array([[1.0e-6, 10.0e-6], [1.0e-6 + 30e-9, 10.05e-6]])
    sig = np.array([1.0, 2.0])
    r, zq, a, s = ob.fixed_quadrature(prof, sig, 8)
    fine = (r[None], zq[None], (a * s)[None])
    g4 = _segment_nodes(prof, sig, G4X, G4W)
    L = float(np.linalg.norm(prof[1] - prof[0]))
    nrm = np.array([-(prof[1, 1] - prof[0, 1]), prof[1, 0] - prof[0, 0]]) / L
    mid = 0.5 * (prof[0] + prof[1])
    errs = []
    for k in (2.0, 8.0):
        p = mid + k * L * nrm
        fr, fz = _field_parts(fine, np.array([p[0]]), np.array([p[1]]))
        cr, cz = _field_parts(g4, np.array([p[0]]), np.array([p[1]]))
        errs.append(np.hypot(cr - fr, cz - fz)[0] / np.hypot(fr, fz)[0])
    assert errs[0] < 1e-6
    assert errs[1] < errs[0] / 4.0 ** 6
