"""The oracle pinned: against the reference's own golden vectors, closed
forms, mpmath and brute force (CPU only)."""
import numpy as np
import pytest

from conftest import golden
from oracle import bem as ob
from oracle import fields as of
from oracle import forces as ofo
from oracle import screens as osc
from oracle import tracker as otr
from oracle.constants import COULOMB_CONSTANT, ELEMENTARY_CHARGE, ELECTRON_MASS_EV, ev_to_kg


# ---------------------------------------------------------------- BEM field
@pytest.mark.parametrize("name", ["bem_sphere.npz", "bem_post80.npz"])
def test_bem_field_matches_reference_golden(name):
    d = golden(name)
    f = ob.evaluate_field(d["profile"], d["sigma"], d["points"])
    np.testing.assert_allclose(f, d["pert"], rtol=1e-13, atol=1e-13 * np.abs(d["pert"]).max())


def test_tip_bem_geometry_field_matches_reference_golden():
    d = golden("bem_sphere.npz")
    g = ob.tip_bem_field(d["gpoints"], float(d["Ez"]), float(d["R"]), d["profile"], d["sigma"],
                         z_handoff=float(d["z_handoff"]))
    np.testing.assert_allclose(g, d["gfield"], rtol=1e-13, atol=1e-13 * np.abs(d["gfield"]).max())


@pytest.mark.parametrize("name", ["bem_sphere.npz", "bem_post80.npz", "bem_post400.npz"])
def test_bem_solve_matches_reference_sigma(name):
    d = golden(name)
    s = ob.solve(d["profile"], lambda r, z: float(d["Ez"]) * z)
    np.testing.assert_allclose(s, d["sigma"], rtol=1e-11, atol=1e-12 * np.abs(d["sigma"]).max())


@pytest.mark.parametrize("tag", ["sharp", "fillet"])
def test_well_profile_field_and_kill_match_reference_golden(tag):
    from bem_amd.bem.mesh import cylindrical_well_profile

    d = golden("bem_well.npz")
    R, H, Ez = float(d["R"]), float(d["H"]), float(d["Ez"])
    bf, rf = float(d[f"{tag}_bottom_fillet"]), float(d[f"{tag}_rim_fillet"])
    prof = cylindrical_well_profile(R, H, float(d[f"{tag}_plane_radius"]), float(d[f"{tag}_max_length"]),
                                    bottom_fillet_radius=bf or None, rim_fillet_radius=rf or None)
    np.testing.assert_array_equal(prof, d[f"{tag}_profile"])
    prof2 = cylindrical_well_profile(R, H, 0.4e-6, 5e-9, bottom_fillet_radius=bf or None,
                                     rim_fillet_radius=rf or None, bottom_max_length=3e-9, wall_max_length=2e-9,
                                     plane_max_length=4e-9, bottom_fillet_max_length=1e-9,
                                     rim_fillet_max_length=1.5e-9, plane_growth=1.2)
    np.testing.assert_array_equal(prof2, d[f"{tag}_profile_regions"])
    pts = d[f"{tag}_points"]
    np.testing.assert_array_equal(ob.evaluate_field(d[f"{tag}_profile"], d[f"{tag}_sigma"], pts), d[f"{tag}_pert"])
    g = ob.well_bem_field(pts, Ez, R, H, d[f"{tag}_profile"], d[f"{tag}_sigma"], float(d[f"{tag}_z_handoff"]), bf, rf)
    np.testing.assert_array_equal(g, d[f"{tag}_gfield"])
    k = ob.well_contains(np.hypot(pts[:, 0], pts[:, 1]), pts[:, 2], R, H, bf, rf)
    np.testing.assert_array_equal(k, d[f"{tag}_kill"])
    sig = ob.solve(d[f"{tag}_profile"], lambda r, z: Ez * z)
    assert np.abs(sig - d[f"{tag}_sigma"]).max() / np.abs(d[f"{tag}_sigma"]).max() < 1e-13


def test_grounded_sphere_potential_and_field():
    """Solved sphere reproduces the closed form E0 a^3 z / r^3 perturbation
    (reference tests/test_bem_axisymmetric.py:118-134,144-221)."""
    R, E0 = 1.0, -1.0
    prof = ob.sphere_profile(R, 40)
    sig = ob.solve(prof, lambda r, z: E0 * z)
    rng = np.random.default_rng(3)
    r = 1.5 + rng.random(20)
    th = np.arccos(rng.uniform(0, 1, 20))
    pts = np.stack([r * np.sin(th), np.zeros(20), r * np.cos(th)], -1)
    E = ob.evaluate_field(prof, sig, pts)
    x, z = pts[:, 0], pts[:, 2]
    rr = np.sqrt(x * x + z * z)
    # perturbation of a grounded sphere in Ez=E0: -grad(E0 R^3 z / r^3)... field of dipole
    Ex = 3 * E0 * R**3 * x * z / rr**5
    Ez = E0 * (-R**3 / rr**3 + 3 * R**3 * z * z / rr**5)
    np.testing.assert_allclose(E[:, 0], Ex, atol=2e-3)
    np.testing.assert_allclose(E[:, 2], Ez, atol=2e-3)


def test_flat_disk_on_axis():
    """Uniform disk: E_z = sigma/2 (1 - h / sqrt(h^2 + a^2)) on the axis
    (reference tests/test_bem_axisymmetric.py:75-88)."""
    a = 1.0
    prof = np.column_stack([np.linspace(0, a, 41), np.zeros(41)])
    sig = np.ones(41)
    h = np.array([0.3, 0.7, 1.5])
    pts = np.stack([np.zeros(3), np.zeros(3), h], -1)
    E = ob.evaluate_field(prof, sig, pts)
    np.testing.assert_allclose(E[:, 2], 0.5 * (1 - h / np.sqrt(h * h + a * a)), rtol=1e-7)


# --------------------------------------------------------------- K, E, ring
def test_ellip_against_mpmath():
    mp = pytest.importorskip("mpmath")
    mp.mp.dps = 40
    for p in [1e-12, 1e-6, 1e-3, 0.1, 0.5, 0.9, 1.0]:
        K, E = ob.ellip_ke_complement(np.array(p))
        assert abs(K / float(mp.ellipk(1 - mp.mpf(p))) - 1) < 4e-16 * 4
        assert abs(E / float(mp.ellipe(1 - mp.mpf(p))) - 1) < 4e-16 * 4


def test_ring_potential_against_brute_quadrature():
    from scipy import integrate

    for rho, z, a in [(0.3, 0.2, 1.0), (1.5, -0.4, 0.7), (0.05, 2.0, 0.3)]:
        f = lambda ph: 1.0 / (4 * np.pi * np.sqrt(rho**2 + a**2 - 2 * rho * a * np.cos(ph) + z**2)) / (2 * np.pi)
        val, _ = integrate.quad(f, 0, 2 * np.pi, epsabs=1e-14, epsrel=1e-13)
        assert abs(ob.ring_potential(rho, z, a) / val - 1) < 1e-10


def test_ring_field_is_minus_gradient_of_potential():
    for rho, z, a in [(0.3, 0.2, 1.0), (1.5, -0.4, 0.7), (1e-7, 0.6, 0.9)]:
        h = 1e-6
        er, ez = ob.ring_field(np.array(rho), np.array(z), np.array(a))
        fd_z = -(ob.ring_potential(rho, z + h, a) - ob.ring_potential(rho, z - h, a)) / (2 * h)
        assert abs(ez - fd_z) < 1e-7 * max(1.0, abs(fd_z))
        if rho > 1e-3:
            fd_r = -(ob.ring_potential(rho + h, z, a) - ob.ring_potential(rho - h, z, a)) / (2 * h)
            assert abs(er - fd_r) < 1e-7 * max(1.0, abs(fd_r))


# ------------------------------------------------------------- pair forces
def test_plummer_two_body():
    """Reference tests/test_coulomb.py:7-21."""
    pos = np.array([[[0.0, 0.0, 0.0], [1e-9, 0.0, 0.0]]])
    q = np.array([[1e-19, 1e-19]])
    act = np.array([[True, True]])
    F = ofo.coulomb(pos, q, act, 1e-12)
    r, a = 1e-9, 1e-12
    mag = COULOMB_CONSTANT * 1e-19 * 1e-19 * r / (r * r + a * a) ** 1.5
    np.testing.assert_allclose(F[0, :, 0], [-mag, mag], rtol=1e-14)
    np.testing.assert_array_equal(F[0, :, 1:], 0.0)
    act2 = np.array([[True, False]])
    np.testing.assert_array_equal(ofo.coulomb(pos, q, act2, 1e-12), 0.0)


def test_flat_self_image():
    """Single charge at height z: k q^2 / (2(z + z0))^2 toward the plane
    (reference tests/test_image_charge.py:8-24)."""
    z, z0, q = 5e-9, 3e-9, -ELEMENTARY_CHARGE
    F = ofo.flat_image(np.array([[[0, 0, z]]]), np.array([[q]]), np.array([[True]]), z0, 1e-12)
    d = 2 * (z + z0)
    expected = -COULOMB_CONSTANT * q * q * d / (d * d + 1e-24) ** 1.5
    np.testing.assert_allclose(F[0, 0, 2], expected, rtol=1e-14)


def test_tip_image_four_charge_reference():
    """Sphere + plane images of one charge against the explicit 4-charge
    sum (reference tests/test_hemisphere.py:66-105)."""
    a, z0, q = 40e-9, 3e-9, -ELEMENTARY_CHARGE
    r0 = np.array([10e-9, -5e-9, 55e-9])
    d = np.linalg.norm(r0)
    imgs = [(r0 * a * a / d / d, -q * a / d), (r0 * [1, 1, -1] - [0, 0, 2 * z0], -q),
            ((r0 * a * a / d / d) * [1, 1, -1] - [0, 0, 2 * z0], q * a / d)]
    F = np.zeros(3)
    for p, qq in imgs:
        dv = r0 - p
        F += COULOMB_CONSTANT * q * qq * dv / np.linalg.norm(dv) ** 3
    got = ofo.tip_image(r0[None, None], np.array([[q]]), np.array([[True]]), a, 0.0, z0)[0, 0]
    np.testing.assert_allclose(got, F, rtol=1e-12)


def test_analytic_tip_enhancement():
    """3 E0 at the pole (reference tests/test_hemisphere.py:40-52)."""
    E = of.tip_field(np.array([[0.0, 0.0, 1e-6]]), -1e8, 1e-6)
    np.testing.assert_allclose(E[0, 2], -3e8, rtol=1e-9)


# ------------------------------------------------------------------ screens
def test_cubic_three_crossings():
    """z = (t-.2)(t-.5)(t-.8) over one step (reference tests/test_screens.py:17-25)."""
    ev = osc.find_crossings([0.0], np.array([0.0]), np.array([[[0, 0, -.08]]]), np.array([[[0, 0, .66]]]),
                            np.array([1.0]), np.array([[[0, 0, .08]]]), np.array([[[0, 0, .66]]]))
    np.testing.assert_allclose(sorted(e["t"] for e in ev), [.2, .5, .8], atol=1e-13)


def test_step_boundary_recorded_once_and_frozen_ignored():
    """Reference tests/test_screens.py:42-47."""
    def step(t0, t1, z0, z1, v0, v1):
        return osc.find_crossings([1.0], np.array([t0]), np.array([[[0, 0, z0]]]), np.array([[[0, 0, v0]]]),
                                  np.array([t1]), np.array([[[0, 0, z1]]]), np.array([[[0, 0, v1]]]))
    evs = step(0, 1, 0, 1, 1, 1) + step(1, 2, 1, 2, 1, 1) + step(2, 3, 1, 1, 1, 1)
    assert [e["t"] for e in evs] == [1.0]


# ---------------------------------------------------------------- integrator
def _geom_for(name):
    if name.startswith("c1"):
        return dict(kind="flat", E_gun=-1e7 if name == "c1" else -1e8, z0=3e-9, kill_z_below=0.0)
    if name == "tipbem":
        R = 1e-6
        prof = ob.sphere_profile(R, ob.n_for_length(R * np.pi, R * np.pi / 39))
        return dict(kind="tip_bem", E_gun=-1e8, R=R, profile=prof, sigma=ob.solve(prof, lambda r, z: -1e8 * z),
                    kill_z_below=0.0)
    if name == "well":
        w = golden("bem_well.npz")
        return dict(kind="well_bem", E_gun=float(w["Ez"]), R=float(w["R"]), H=float(w["H"]),
                    profile=w["sharp_profile"], sigma=w["sharp_sigma"], z_handoff=float(w["sharp_z_handoff"]))
    if name == "tipimg":
        t = golden("bem_tip_image.npz")
        return dict(kind="tip_bem", E_gun=float(t["Ez"]), R=float(t["R"]), profile=t["real_profile"],
                    sigma=t["sigma"], kill_z_below=0.0, z_handoff=float(t["z_handoff"]), z0=float(t["z0"]),
                    image_profile=t["doubled_image_profile"], image_A_inv=np.linalg.inv(t["doubled_A"]),
                    n_max=int(t["n_max"]), d_lo=float(t["d_lo"]), d_hi=float(t["d_hi"]), image_mirror_symmetric=True)
    if name == "wellimg":
        w = golden("bem_well_image.npz")
        return dict(kind="well_bem", E_gun=float(w["Ez"]), R=float(w["R"]), H=float(w["H"]),
                    profile=w["real_profile"], sigma=w["sigma"], z_handoff=float(w["z_handoff"]),
                    bottom_fillet=float(w["bottom_fillet"]), rim_fillet=float(w["rim_fillet"]), z0=float(w["z0"]),
                    image_profile=w["image_profile"], image_A_inv=np.linalg.inv(w["A"]), n_max=int(w["n_max"]),
                    d_lo=float(w["d_lo"]), d_hi=float(w["d_hi"]))
    p = golden("bem_post80.npz")
    R, H = float(p["R"]), float(p["H"])
    return dict(kind="post_bem", E_gun=-1e8, R=R, H=H, profile=p["profile"], sigma=p["sigma"],
                z_handoff=20 * (R + H), kill_z_below=0.0)


def run_oracle(name):
    d = golden(f"run_{name}.npz")
    st = otr.setup(d["in_x"], d["in_y"], d["in_z"], d["in_px"], d["in_py"], d["in_pz"], d["in_t"], d["in_id"],
                   int(d["n_emit"]))
    out = otr.run(st, _geom_for(name), list(d["screens"]), float(d["t_max"]), rtol=float(d["rtol"]),
                  atol=float(d["atol"]))
    return d, st, out


def test_oracle_output_times_match_reference_run():
    d = golden("run_c1out.npz")
    st = otr.setup(d["in_x"], d["in_y"], d["in_z"], d["in_px"], d["in_py"], d["in_pz"], d["in_t"], d["in_id"],
                   int(d["n_emit"]))
    out = otr.run(st, _geom_for("c1small"), list(d["screens"]), float(d["t_max"]), rtol=float(d["rtol"]),
                  atol=float(d["atol"]), t_out=d["t_out"])
    np.testing.assert_array_equal(out["attempts"], d["attempts"])
    for k in range(len(d["t_out"])):
        rec = out["trajectories"][k]
        ids = np.array([st["ids"][r["group"], r["particle"]] for r in rec], dtype=int)
        np.testing.assert_array_equal(ids, d[f"o{k}_id"])
        if len(rec):
            np.testing.assert_allclose([r["t"] for r in rec], d[f"o{k}_t"], rtol=1e-15)
            np.testing.assert_allclose([r["pos"][2] for r in rec], d[f"o{k}_z"], rtol=1e-9, atol=1e-15)


@pytest.mark.parametrize("name", ["c1small", "tipbem", "post80", "well", "wellimg", "tipimg"])
def test_oracle_integrator_matches_reference_run(name):
    d, st, out = run_oracle(name)
    np.testing.assert_array_equal(out["attempts"], d["attempts"])
    np.testing.assert_array_equal(out["accepted"], d["accepted"])
    np.testing.assert_allclose(out["pos"], d["pos"], rtol=1e-9, atol=1e-9 * np.abs(d["pos"]).max())
    np.testing.assert_allclose(out["vel"], d["vel"], rtol=1e-9, atol=1e-9 * np.abs(d["vel"]).max())
    for si in range(len(d["screens"])):
        ev = [e for e in out["events"] if e["screen"] == si]
        ids = np.array([st["ids"][e["group"], e["particle"]] for e in ev])
        np.testing.assert_array_equal(ids, d[f"s{si}_id"])
        np.testing.assert_allclose([e["t"] for e in ev], d[f"s{si}_t"], rtol=1e-12)


def test_ballistic_closed_form():
    """z = a t^2 / 2 at the screen (reference tests/test_tracker.py:10-37)."""
    E = -1e8
    m = ev_to_kg(ELECTRON_MASS_EV)
    a = ELEMENTARY_CHARGE * abs(E) / m
    zs = 1e-6
    ta = np.sqrt(2 * zs / a)
    st = otr.setup(np.zeros(1), np.zeros(1), np.zeros(1), np.zeros(1), np.zeros(1), np.zeros(1), np.zeros(1),
                   np.array([1]), 1)
    out = otr.run(st, dict(kind="flat", E_gun=E, z0=None, kill_z_below=0.0), [zs], ta * 1.5)
    assert len(out["events"]) == 1
    np.testing.assert_allclose(out["events"][0]["t"], ta, rtol=1e-6)
    np.testing.assert_allclose(out["events"][0]["pos"][2], zs, rtol=1e-9)
