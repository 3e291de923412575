"""CPU checks of the oracle parts behind the surrogate: the BEM potential
(reference bem/axisymmetric.py:275-302,445-462) and the FP64 MLP with its
analytic gradient. No GPU."""
import numpy as np
import pytest

from conftest import golden
from oracle import bem as ob
from oracle import mlp as om


def test_potential_reproduces_the_reference_dirichlet_data():
    """The reference solved A sigma = E_z z at the nodes (golden sigma from
    its own solve); the potential of that sigma at each node must give the
    Dirichlet value back."""
    for name in ("bem_sphere.npz", "bem_post80.npz"):
        d = golden(name)
        prof, sig, Ez = d["profile"], d["sigma"], float(d["Ez"])
        nodes = np.stack([prof[:, 0], np.zeros(len(prof)), prof[:, 1]], -1)
        phi = ob.potential(prof, sig, nodes)
        b = Ez * prof[:, 1]
        scale = np.abs(b).max()
        assert np.abs(phi - b).max() / scale < 1e-10, name


def test_potential_far_field_is_the_total_charge():
    """Far away the potential is Q / (4 pi r) with Q the integral of the
    linear sigma over the polygonal surface of revolution (closed form per
    segment); the symmetric polygon has no dipole."""
    prof = ob.sphere_profile(1e-6, 41)
    rng = np.random.default_rng(0)
    sig = 1.0 + 0.3 * np.cos(np.linspace(0, np.pi, 41)) ** 2
    p0, p1 = prof[:-1], prof[1:]
    L = np.linalg.norm(p1 - p0, axis=1)
    s0, s1 = sig[:-1], sig[1:]
    r0, r1 = p0[:, 0], p1[:, 0]
    # integral over t in [0, 1] of (s0 + t ds)(r0 + t dr) dt, times 2 pi L
    Q = np.sum(2 * np.pi * L * (s0 * r0 + 0.5 * (s0 * (r1 - r0) + r0 * (s1 - s0)) + (s1 - s0) * (r1 - r0) / 3.0))
    th = rng.uniform(0, np.pi, 5)
    rr = 1e-2
    pts = np.stack([rr * np.sin(th), np.zeros(5), rr * np.cos(th)], -1)
    phi = ob.potential(prof, sig, pts)
    # quadrupole remainder ~ (1e-6 / 1e-2)^2
    assert np.abs(phi * 4 * np.pi * rr / Q - 1).max() < 1e-7


def test_potential_converges_to_the_sphere_shell_values():
    """Uniform sigma on ever finer spheres: outside sigma R^2 / r, inside
    sigma R, error falling ~4x per doubling (second-order chords)."""
    R = 1e-6
    pts = np.array([[0, 0, 3e-6], [2e-6, 0, 1e-6], [0, 0, 0.3e-6]])
    r = np.linalg.norm(pts, axis=1)
    exact = np.where(r > R, 2.0 * R * R / r, 2.0 * R)
    errs = []
    for n in (51, 101, 201):
        phi = ob.potential(ob.sphere_profile(R, n), np.full(n, 2.0), pts)
        errs.append(np.abs(phi / exact - 1).max())
    assert errs[-1] < 5e-5
    assert 3.0 < errs[0] / errs[1] < 5.0 and 3.0 < errs[1] / errs[2] < 5.0


def _cfg(rng):
    return dict(fourier=rng.normal(0, 1.5, (32, 2)), center_z=10e-6, r_unit=1e-6, u_offset=(2.0, 0.0),
                u_scale=(0.5, 1.0), out_scale=(3.0, 2e8, 5e8), env_r0=1e-6, env_pow=2.5)


def test_mlp_features_closed_form():
    cfg = _cfg(np.random.default_rng(0))
    cfg["fourier"] = np.zeros((32, 2))
    cfg["fourier"][0] = (1.0, 0.0)
    cfg["fourier"][1] = (0.0, 0.5)
    u0, u1, env = om.coords(cfg, np.array([0.0]), np.array([10e-6 + np.e ** 2.5 * 1e-6]))
    assert abs(u0[0] - 0.25) < 1e-12 and abs(u1[0] - 1.0) < 1e-12
    X = om.features(cfg, u0, u1)
    assert abs(X[0, 0] - 1.0) < 1e-12          # sin(2 pi * 0.25)
    assert abs(X[0, 1] - 0.0) < 1e-12          # sin(2 pi * 0.5)
    assert abs(X[0, 32 + 1] + 1.0) < 1e-12     # cos(pi)
    assert abs(X[0, 32 + 2] - 1.0) < 1e-12     # cos(0)
    r = np.e ** 2.5 * 1e-6
    assert abs(env[0] - (1 + r / 1e-6) ** -2.5) < 1e-15


def test_mlp_gradient_matches_finite_differences():
    """The analytic backward pass against central differences of the loss,
    at entries of every parameter block."""
    rng = np.random.default_rng(1)
    off = om.layout()
    p = rng.normal(0, 0.2, off["n"])
    p[off["Wo"] + 3 * 128: off["bo"]] = 0.0
    X0 = rng.normal(0, 1, (8, 64))
    T = rng.normal(0, 1, (8, 3))
    w = (0.5, 1.0, 2.0)
    loss, g = om.loss_and_grad(p, X0, T, w)
    picks = []
    for key in ("W1", "b1", "W2", "b2", "W3", "b3", "W4", "b4", "Wo", "bo"):
        picks += list(off[key] + rng.integers(0, 3 if key == "bo" else 128, 3))
    for i in picks:
        h = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (om.loss_and_grad(pp, X0, T, w)[0] - om.loss_and_grad(pm, X0, T, w)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])), (i, fd, g[i])


def test_mlp_padded_output_rows_get_no_gradient():
    rng = np.random.default_rng(2)
    off = om.layout()
    p = rng.normal(0, 0.2, off["n"])
    _, g = om.loss_and_grad(p, rng.normal(0, 1, (4, 64)), rng.normal(0, 1, (4, 3)), (1, 1, 1))
    assert np.all(g[off["Wo"] + 3 * 128: off["bo"]] == 0) and np.all(g[off["bo"] + 3: off["n"]] == 0)


def test_initial_params_layout_matches_oracle():
    pytest.importorskip("torch")
    from bem_amd import surrogate

    p = surrogate.initial_params(0)
    off = om.layout()
    assert p.size == off["n"]
    Ws, bs = om.unpack(p)
    assert all(np.all(b == 0) for b in bs)
    assert np.all(Ws[-1][3:] == 0) and np.any(Ws[-1][:3] != 0)


def test_surrogate_requires_lockstep_and_its_geometry():
    from bem_amd import FlatCathode, SpecificParticleTracer
    from bem_amd.surrogate import FieldSurrogate
    from bem_amd.particles import ParticleGroup

    geo = FlatCathode(-1e7)
    sur = FieldSurrogate(geo, {}, None)
    n = 4
    pg = ParticleGroup(dict(x=np.zeros(n), y=np.zeros(n), z=np.zeros(n), px=np.zeros(n), py=np.zeros(n),
                            pz=np.ones(n), t=np.zeros(n), id=np.arange(1, n + 1), weight=np.full(n, 1.602176634e-19),
                            status=np.ones(n), species="electron"))
    with pytest.raises(ValueError, match="lockstep"):
        SpecificParticleTracer(pg, 4, geo, [1e-6], mode="persistent", surrogate=sur)
    with pytest.raises(ValueError, match="different geometry"):
        SpecificParticleTracer(pg, 4, FlatCathode(-1e7), [1e-6], surrogate=sur)
