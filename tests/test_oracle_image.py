"""CPU checks of the image-charge BEM oracle (reference
bem/image_charge.py) and the well's image profile generator: against the
reference's own golden values, and by the subtraction's mathematics (the
converged force does not depend on the mirror-charge blend, up to the
surface discretisation)."""
import numpy as np

from conftest import golden
from oracle import image as oi
from oracle import tracker as otr


def test_image_and_real_profiles_match_reference():
    from bem_amd.bem.mesh import cylindrical_well_image_profile, cylindrical_well_profile

    d = golden("bem_well_image.npz")
    R, H, z0 = float(d["R"]), float(d["H"]), float(d["z0"])
    bf, rf = float(d["bottom_fillet"]), float(d["rim_fillet"])
    p = cylindrical_well_image_profile(R, H, z0, float(d["image_plane_radius"]), float(d["image_max_length"]),
                                       bottom_fillet_radius=bf, rim_fillet_radius=rf)
    np.testing.assert_array_equal(p, d["image_profile"])
    q = cylindrical_well_profile(R, H, float(d["z_handoff"]), 3e-9, bottom_fillet_radius=bf, rim_fillet_radius=rf)
    np.testing.assert_array_equal(q, d["real_profile"])


def test_operators_and_force_match_reference():
    d = golden("bem_well_image.npz")
    A = oi.assemble_mode_operators(d["image_profile"], int(d["n_max"]))
    assert np.max(np.abs(A - d["A"]) / np.abs(d["A"]).max(axis=(1, 2), keepdims=True)) < 1e-13
    Ai = np.linalg.inv(d["A"])
    F = oi.image_force(d["image_profile"], Ai, int(d["n_max"]), d["pos"], d["charge"], d["active"],
                       float(d["d_lo"]), float(d["d_hi"]))
    assert np.abs(F - d["near"]).max() / np.abs(d["near"]).max() < 1e-13
    geom = dict(kind="well_bem", z0=float(d["z0"]), image_profile=d["image_profile"], image_A_inv=Ai,
                n_max=int(d["n_max"]), d_lo=float(d["d_lo"]), d_hi=float(d["d_hi"]), z_handoff=float(d["z_handoff"]))
    Fb = otr.image_force(geom, d["pos"], d["charge"], d["active"], 1e-12)
    assert np.abs(Fb - d["force"]).max() / np.abs(d["force"]).max() < 1e-13


def test_converged_force_does_not_depend_on_the_mirror_blend():
    d = golden("bem_well_image.npz")
    prof = d["image_profile"]
    R, H = float(d["R"]), float(d["H"])
    Ai = np.linalg.inv(oi.assemble_mode_operators(prof, 24))
    pos = np.array([[[0.3 * R, 0.1 * R, -H + 8e-9], [0.85 * R * np.cos(1.0), 0.85 * R * np.sin(1.0), -0.5 * H]]])
    q = np.full((1, 2), -1.602176634e-19)
    act = np.ones((1, 2), dtype=bool)
    f16 = oi.image_force(prof, Ai, 16, pos, q, act, 0.1 * R, 0.5 * R)
    f24 = oi.image_force(prof, Ai, 24, pos, q, act, 0.1 * R, 0.5 * R)
    bare = oi.image_force(prof, Ai, 24, pos, q, act, 0.0, 0.0)
    # the particle 8 nm above the bottom: modes converged, blend-independent up to the discretisation
    assert np.abs(f24[0, 0] - f16[0, 0]).max() < 1e-6 * np.abs(f24[0, 0]).max()
    assert np.abs(f24[0, 0] - bare[0, 0]).max() < 1e-3 * np.abs(f24[0, 0]).max()
    # the force pulls each electron towards its nearest conductor: down to the bottom, out to the wall
    assert f24[0, 0, 2] < 0
    assert f24[0, 1, 0] * pos[0, 1, 0] + f24[0, 1, 1] * pos[0, 1, 1] > 0


def test_tip_image_profiles_match_reference():
    from bem_amd.bem.mesh import hemisphere_tip_image_doubled_profile, hemisphere_tip_image_profile

    d = golden("bem_tip_image.npz")
    R, z0 = float(d["R"]), float(d["z0"])
    p = hemisphere_tip_image_profile(R, z0, 5.0 * R, float(d["plane_image_max_length"]))
    np.testing.assert_array_equal(p, d["plane_image_profile"])
    q = hemisphere_tip_image_doubled_profile(R, z0, float(d["doubled_image_max_length"]))
    np.testing.assert_array_equal(q, d["doubled_image_profile"])


def _tip_geom(d, tag):
    Ai = np.linalg.inv(d[f"{tag}_A"])
    return dict(kind="tip_bem", z0=float(d["z0"]), image_profile=d[f"{tag}_image_profile"], image_A_inv=Ai,
                n_max=int(d["n_max"]), d_lo=float(d["d_lo"]), d_hi=float(d["d_hi"]), z_handoff=float(d["z_handoff"]),
                image_mirror_symmetric=tag == "doubled")


def test_tip_operators_and_forces_match_reference():
    """Both image surfaces of the reference's HemisphericalTipBEMGeometry
    (truncated plane; mirror-doubled with phantom charges): operators, the
    image-BEM force (plus the phantoms' plane image when doubled) and the
    blended geometry force, particles over the cap, at the belt, over the
    plane and on the axis."""
    d = golden("bem_tip_image.npz")
    z0 = float(d["z0"])
    for tag in ("plane", "doubled"):
        A = oi.assemble_mode_operators(d[f"{tag}_image_profile"], int(d["n_max"]))
        assert np.max(np.abs(A - d[f"{tag}_A"]) / np.abs(d[f"{tag}_A"]).max(axis=(1, 2), keepdims=True)) < 1e-13
        g = _tip_geom(d, tag)
        near = oi.image_force(g["image_profile"], g["image_A_inv"], g["n_max"], d["pos"], d["charge"], d["active"],
                              g["d_lo"], g["d_hi"], mirror_plane_z=-z0 if tag == "doubled" else None)
        if tag == "doubled":
            from oracle import forces
            near = near + forces.flat_image(d["pos"], d["charge"], d["active"], z0, 1e-12)
        assert np.abs(near - d[f"{tag}_near"]).max() / np.abs(d[f"{tag}_near"]).max() < 1e-13, tag
        F = otr.image_force(g, d["pos"], d["charge"], d["active"], 1e-12)
        assert np.abs(F - d[f"{tag}_force"]).max() / np.abs(d[f"{tag}_force"]).max() < 1e-13, tag


def test_tip_image_force_approaches_the_sphere_on_plane_images():
    """On the axis 5 nm above the pole the recessed surface is a sphere of
    radius R - z0 on the plane z = -z0 up to the ridge fillet, so both image
    surfaces must give the classical three-image force (sphere, plane, plane
    image of the sphere image) to the fillet's effect (~7e-4). Dropping the
    phantoms' plane image from the doubled form is 2 % off."""
    from bem_amd.bem.mesh import hemisphere_tip_image_doubled_profile, hemisphere_tip_image_profile
    from oracle import forces

    R, z0 = 50e-9, 3e-9
    pos = np.array([[[1e-12, 0.0, R + 5e-9]]])
    q = np.full((1, 1), -1.602176634e-19)
    act = np.ones((1, 1), dtype=bool)
    ana = forces.tip_image(pos, q, act, R - z0, 1e-12, z0)[0, 0, 2]
    ml = (R - z0) * np.pi / 2 / 29
    for prof, mpz in ((hemisphere_tip_image_profile(R, z0, 5 * R, ml), None),
                      (hemisphere_tip_image_doubled_profile(R, z0, ml), -z0)):
        Ai = np.linalg.inv(oi.assemble_mode_operators(prof, 16))
        f = oi.image_force(prof, Ai, 16, pos, q, act, 0.1 * R, 0.5 * R, mirror_plane_z=mpz)[0, 0, 2]
        if mpz is not None:
            plane = forces.flat_image(pos, q, act, z0, 1e-12)[0, 0, 2]
            assert abs((f - ana) / ana) > 1e-2  # the phantoms' plane image is not optional
            f = f + plane
        assert abs((f - ana) / ana) < 2e-3
