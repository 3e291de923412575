"""Exterior multipole expansion of the BEM charge (bem_amd.bem.multipole),
on the CPU against the oracle's ring kernel.

* One ring: the expansion's field equals the exact elliptic-integral ring
  field (oracle.bem.ring_field) off the axis and the closed-form on-axis
  field Q (z - b) / (4 pi ((z - b)^2 + a^2)^(3/2)) to 1e-12.
* The quadrature nodes equal the oracle's fixed rule node for node.
* On the 400-panel post (reference-solved sigma), 128 terms reproduce the
  oracle's far-rule sums at r >= 1.1 r_ref to 1e-9 of |E|, while 32 terms
  do not (> 1e-6): truncation is visible to the test.
"""
import numpy as np

from bem_amd.bem import multipole as mpl
from conftest import golden
from oracle import bem as ob


def test_single_ring_against_closed_forms():
    a, b, Q = 0.7e-6, 0.4e-6, 2.5e-3
    r0 = np.hypot(a, b)
    P = mpl.legendre(np.array([b / r0]), 120)[:, 0]
    m = Q * P  # one node at r_ref = r0: (r_q / r_ref)^l = 1
    z = np.array([1.5, 2.0, 3.0]) * r0
    er, ez = mpl.field(m, r0, np.zeros_like(z), z)
    assert np.all(er == 0.0)
    exact = Q * (z - b) / (4 * np.pi * ((z - b) ** 2 + a * a) ** 1.5)
    assert np.max(np.abs(ez - exact) / np.abs(exact)) < 1e-12
    th = np.array([0.3, 0.9, 1.4])
    rho, zz = 2.2 * r0 * np.sin(th), 2.2 * r0 * np.cos(th)
    er, ez = mpl.field(m, r0, rho, zz)
    fr, fz = ob.ring_field(rho, zz - b, a)
    assert np.max(np.hypot(er - Q * fr, ez - Q * fz) / np.hypot(Q * fr, Q * fz)) < 1e-12


def test_nodes_match_oracle_fixed_rule():
    d = golden("bem_post400.npz")
    rho, z, w = mpl.quadrature_nodes(d["profile"], d["sigma"], 8)
    r, zq, a, s = ob.fixed_quadrature(d["profile"], d["sigma"], 8)
    assert np.allclose(rho, r, rtol=1e-15, atol=0) and np.allclose(z, zq, rtol=1e-15, atol=1e-21)
    assert np.allclose(w, a * s, rtol=1e-13, atol=0)


def test_post_expansion_against_oracle_far_sums():
    d = golden("bem_post400.npz")
    prof, sig, Ez = d["profile"], d["sigma"], float(d["Ez"])
    rq, zq, w = mpl.quadrature_nodes(prof, sig, 8)
    m, r_ref = mpl.moments(prof, sig, 8, 128)
    assert abs(m[0]) < 1e-12 * np.abs(m).max()  # mirror-doubled: no net charge
    rng = np.random.default_rng(2)
    n = 200
    th = np.arccos(rng.uniform(0.0, 1.0, n))
    th[:20] = rng.uniform(0.0, 0.02, 20)
    r = r_ref * np.exp(rng.uniform(np.log(1.1), np.log(8.0), n))
    rho, z = r * np.sin(th), r * np.cos(th)
    er, ez = ob.ring_field(rho[:, None], z[:, None] - zq[None], rq[None])
    fr, fz = (er * w).sum(1), (ez * w).sum(1)
    Et = np.hypot(fr, fz + Ez)
    gr, gz = mpl.field(m, r_ref, rho, z)
    assert np.max(np.hypot(gr - fr, gz - fz) / Et) < 1e-9
    gr, gz = mpl.field(m[:32], r_ref, rho, z)
    assert np.max(np.hypot(gr - fr, gz - fz) / Et) > 1e-6
