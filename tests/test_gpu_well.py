"""Cylindrical well (reference bem/geometry.py:510-812) on the GPU: device
BEM solve vs the reference's sigma (1e-10), blended geometry field vs the
reference (1e-11 of the largest component per point), kill mask exact,
and uneven groups with per-group emission offsets from the well bottom
(the C3 shape, small) equal per-group oracle runs: identical attempt and
accept counts, final state 1e-9."""
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


def _geo(tag, solve=True):
    from bem_amd.bem.geometry import CylindricalWellBEMGeometry

    d = golden("bem_well.npz")
    bf, rf = float(d[f"{tag}_bottom_fillet"]), float(d[f"{tag}_rim_fillet"])
    return d, CylindricalWellBEMGeometry(float(d["Ez"]), float(d["R"]), float(d["H"]),
                                         field_max_length=float(d[f"{tag}_max_length"]),
                                         bottom_fillet_radius=bf or None, rim_fillet_radius=rf or None,
                                         sigma=None if solve else d[f"{tag}_sigma"])


@pytest.mark.parametrize("tag", ["sharp", "fillet"])
def test_well_solve_field_and_kill_match_reference(gpu, tag):
    _say("well", tag)
    d, geo = _geo(tag)
    np.testing.assert_array_equal(geo.solution.profile, d[f"{tag}_profile"])
    ref = d[f"{tag}_sigma"]
    err = np.abs(geo.solution.sigma - ref).max() / np.abs(ref).max()
    _say(f"well {tag}: sigma rel {err:.2e}")
    assert err < 1e-10
    d, geo = _geo(tag, solve=False)
    pts = d[f"{tag}_points"]
    g = geo.field(pts)
    ref = d[f"{tag}_gfield"]
    scale = np.abs(ref).max(axis=-1, keepdims=True)
    scale = np.where(scale > 0, scale, 1.0)
    ferr = np.max(np.abs(g - ref) / scale)
    _say(f"well {tag}: field rel {ferr:.2e} over {len(pts)} points")
    assert ferr < 1e-11
    np.testing.assert_array_equal(geo.kill_mask(pts), d[f"{tag}_kill"])


@pytest.mark.parametrize("mode", ["lockstep", "persistent"])
def test_well_uneven_groups_with_offsets_match_per_group_oracle(gpu, mode):
    from bem_amd import SpecificParticleTracer
    from bem_amd.emission import well_bottom_source
    from bem_amd.particles import ParticleGroup
    from oracle import tracker as otr

    _say("well uneven", mode)
    d, geo = _geo("sharp", solve=False)
    R, H = float(d["R"]), float(d["H"])
    sizes = [3, 1, 6, 2]
    src = well_bottom_source(sizes, R, H, t_offset_max=10e-15, seed=12)
    n = sum(sizes)
    pg = ParticleGroup(dict(src, weight=np.full(n, 1.602176634e-19), status=np.ones(n), species="electron"))
    screens = [20e-9, 100e-9]
    t_max = 6e-14
    tr = SpecificParticleTracer(pg, None, geo, screens, rtol=1e-8, t_max=t_max, group_sizes=sizes, mode=mode)
    tr.run()
    out = tr.last_run
    geom = dict(kind="well_bem", E_gun=float(d["Ez"]), R=R, H=H, profile=d["sharp_profile"],
                sigma=d["sharp_sigma"], z_handoff=float(d["sharp_z_handoff"]))
    start = 0
    for g, s in enumerate(sizes):
        sl = slice(start, start + s)
        st = otr.setup(*(src[k][sl] for k in ("x", "y", "z", "px", "py", "pz", "t", "id")), s)
        o = otr.run(st, geom, screens, t_max, rtol=1e-8)
        assert out["attempts"][g] == o["attempts"][0], (g, out["attempts"][g], o["attempts"][0])
        assert out["accepted"][g] == o["accepted"][0]
        np.testing.assert_allclose(out["pos"][sl], o["pos"][0], rtol=1e-9, atol=1e-9 * np.abs(o["pos"][0]).max())
        start += s
    _say(f"well uneven [{mode}]: attempts {out['attempts'].tolist()}")
