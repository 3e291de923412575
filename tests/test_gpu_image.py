"""Image-charge BEM on the well and on the hemispherical tip (reference
bem/image_charge.py, bem/geometry.py:326-427,786-801; the tip both on its
truncated-plane image surface and on the mirror-doubled one with phantom
charges) on the GPU: per-mode operators vs the
reference's (1e-12 of each mode's largest entry; the adaptive rule and
leaf order are the reference's, K/E and Q differ by ulps), the blended
image force of a particle batch vs the reference (1e-10 of the largest
force: operators inverted by rocSOLVER rather than NumPy, quadrature sums
in another order), and a lockstep trace with image charges vs the
reference run (identical attempt/accept counts, state 1e-9)."""
import sys

import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from bem_amd import _lib

    _lib.load()
    return torch


def _say(*a):
    print(*a, flush=True, file=sys.stderr)


def _geo():
    from bem_amd.bem.geometry import CylindricalWellBEMGeometry

    d = golden("bem_well_image.npz")
    geo = CylindricalWellBEMGeometry(float(d["Ez"]), float(d["R"]), float(d["H"]), z0=float(d["z0"]),
                                     field_max_length=3e-9, image_max_length=float(d["image_max_length"]),
                                     image_n_max=int(d["n_max"]), sigma=d["sigma"])
    return d, geo


def test_image_profile_and_operators_match_reference(gpu):
    _say("image operators")
    d, geo = _geo()
    np.testing.assert_array_equal(geo.image_profile, d["image_profile"])
    np.testing.assert_array_equal(geo.solution.profile, d["real_profile"])
    A = geo.image_operators()
    ref = d["A"]
    err = np.max(np.abs(A - ref) / np.abs(ref).max(axis=(1, 2), keepdims=True))
    _say(f"image operators: max rel {err:.2e}")
    assert err < 1e-12


def test_image_force_matches_reference(gpu):
    _say("image force")
    d, geo = _geo()
    F = geo.image_force(d["pos"], d["charge"], d["active"], 1e-12)
    ref = d["force"]
    err = np.abs(F - ref).max() / np.abs(ref).max()
    _say(f"image force: max rel {err:.2e}")
    assert err < 1e-10
    assert np.all(F[~d["active"]] == 0.0)


@pytest.mark.parametrize("mode", ["lockstep", "persistent"])
def test_image_trace_matches_reference_run(gpu, mode):
    """Both trace kernels: lockstep (per-stage kernels over all groups) and
    the persistent kernel (each workgroup forms its group's right-hand
    sides, solves and forces at every stage)."""
    from bem_amd import SpecificParticleTracer
    from bem_amd.particles import ParticleGroup

    _say(f"image trace [{mode}]")
    d, geo = _geo()
    r = golden("run_wellimg.npz")
    src = {k[3:]: r[k] for k in r.files if k.startswith("in_")}
    n = len(src["x"])
    pg = ParticleGroup(dict(src, weight=np.full(n, 1.602176634e-19), status=np.ones(n), species="electron"))
    tr = SpecificParticleTracer(pg, int(r["n_emit"]), geo, list(r["screens"]), rtol=float(r["rtol"]),
                                atol=float(r["atol"]), t_max=float(r["t_max"]), mode=mode)
    screens, _ = tr.run()
    out = tr.last_run
    G = tr.n_groups
    _say(f"image trace [{mode}]: attempts {out['attempts'].tolist()} vs {r['attempts'].tolist()}")
    np.testing.assert_array_equal(out["attempts"], r["attempts"])
    np.testing.assert_array_equal(out["accepted"], r["accepted"])
    pos = out["pos"].reshape(G, -1, 3)
    vel = out["vel"].reshape(G, -1, 3)
    np.testing.assert_allclose(pos, r["pos"], rtol=1e-9, atol=1e-9 * np.abs(r["pos"]).max())
    np.testing.assert_allclose(vel, r["vel"], rtol=1e-9, atol=1e-9 * np.abs(r["vel"]).max())
    for si, s in enumerate(screens):
        np.testing.assert_array_equal(s.id, r[f"s{si}_id"])
        np.testing.assert_allclose(s.t, r[f"s{si}_t"], rtol=1e-11)


def _tip_geo(tag):
    from bem_amd.bem.geometry import HemisphericalTipBEMGeometry

    d = golden("bem_tip_image.npz")
    geo = HemisphericalTipBEMGeometry(float(d["Ez"]), float(d["R"]), z0=float(d["z0"]), image_n_max=int(d["n_max"]),
                                      image_mirror_symmetric=tag == "doubled", sigma=d["sigma"])
    return d, geo


@pytest.mark.parametrize("tag", ["plane", "doubled"])
def test_tip_image_operators_and_force_match_reference(gpu, tag):
    _say(f"tip image [{tag}]")
    d, geo = _tip_geo(tag)
    np.testing.assert_array_equal(geo.image_profile, d[f"{tag}_image_profile"])
    np.testing.assert_array_equal(geo.solution.profile, d["real_profile"])
    A = geo.image_operators()
    ref = d[f"{tag}_A"]
    err = np.max(np.abs(A - ref) / np.abs(ref).max(axis=(1, 2), keepdims=True))
    F = geo.image_force(d["pos"], d["charge"], d["active"], 1e-12)
    fref = d[f"{tag}_force"]
    ferr = np.abs(F - fref).max() / np.abs(fref).max()
    _say(f"tip image [{tag}]: operators max rel {err:.2e}, force max rel {ferr:.2e}")
    assert err < 1e-12
    assert ferr < 1e-10
    assert np.all(F[~d["active"]] == 0.0)


@pytest.mark.parametrize("mode", ["lockstep", "persistent"])
def test_tip_mirror_image_trace_matches_reference_run(gpu, mode):
    from bem_amd import SpecificParticleTracer
    from bem_amd.particles import ParticleGroup

    _say(f"tip mirror-image trace [{mode}]")
    d, geo = _tip_geo("doubled")
    r = golden("run_tipimg.npz")
    src = {k[3:]: r[k] for k in r.files if k.startswith("in_")}
    n = len(src["x"])
    pg = ParticleGroup(dict(src, weight=np.full(n, 1.602176634e-19), status=np.ones(n), species="electron"))
    tr = SpecificParticleTracer(pg, int(r["n_emit"]), geo, list(r["screens"]), rtol=float(r["rtol"]),
                                atol=float(r["atol"]), t_max=float(r["t_max"]), mode=mode)
    screens, _ = tr.run()
    out = tr.last_run
    G = tr.n_groups
    _say(f"tip mirror-image trace: attempts {out['attempts'].tolist()} vs {r['attempts'].tolist()}")
    np.testing.assert_array_equal(out["attempts"], r["attempts"])
    np.testing.assert_array_equal(out["accepted"], r["accepted"])
    pos = out["pos"].reshape(G, -1, 3)
    vel = out["vel"].reshape(G, -1, 3)
    np.testing.assert_allclose(pos, r["pos"], rtol=1e-9, atol=1e-9 * np.abs(r["pos"]).max())
    np.testing.assert_allclose(vel, r["vel"], rtol=1e-9, atol=1e-9 * np.abs(r["vel"]).max())
    for si, s in enumerate(screens):
        np.testing.assert_array_equal(s.id, r[f"s{si}_id"])
        np.testing.assert_allclose(s.t, r[f"s{si}_t"], rtol=1e-11)
