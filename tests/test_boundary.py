"""The drop-in boundary on the host (no GPU): reference worker_args()
tuples (captured from the reference by tests/golden/make_golden.py) lowered
onto this package's geometries, the object-cache keys, and the mapping of
device group status to the reference's exceptions."""
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN, golden


@pytest.fixture
def no_device_solve(monkeypatch):
    """Geometry construction without a GPU: the collocation solve (device)
    is replaced by zeros; profiles and parameters are what is checked."""
    from bem_amd.bem import solver

    monkeypatch.setattr(solver, "solve_sigma", lambda profile, *a, **k: np.zeros(len(profile)))


def _tuples():
    with open(os.path.join(GOLDEN, "worker_args.json")) as f:
        return json.load(f)


@pytest.mark.parametrize("name", ["flat", "flat_noimage", "tip", "tip_bem", "tip_bem_mirror", "well_bem",
                                  "well_bem_noimage"])
def test_reference_worker_args_lower_to_the_same_geometry(name, no_device_solve):
    """Every reference tuple rebuilds this package's class of the same name,
    whose own worker_args() gives the tuple back field for field (reference
    geometry.py:61-94, bem/geometry.py:437-507,814-870), and whose profiles
    are the reference's node for node."""
    from bem_amd.geometry import lower

    ref = tuple(_tuples()[name])
    geo = lower(ref)
    assert type(geo).__name__ == {"flat": "FlatCathode", "hemispherical_tip": "HemisphericalTip",
                                  "hemispherical_tip_bem": "HemisphericalTipBEMGeometry",
                                  "cylindrical_well_bem": "CylindricalWellBEMGeometry"}[ref[0]]
    got = geo.worker_args()
    assert len(got) == len(ref)
    for a, b in zip(got, ref):
        assert (a is None and b is None) or a == b, (got, ref)
    prof = golden("worker_args.npz")
    if f"{name}_real_profile" in prof:
        assert np.array_equal(geo.real_profile, prof[f"{name}_real_profile"])
    if f"{name}_image_profile" in prof:
        assert np.array_equal(geo.image_profile, prof[f"{name}_image_profile"])
    elif hasattr(geo, "image_profile"):
        assert geo.image_profile is None


def test_lower_accepts_objects_and_rejects_unknown_kinds(no_device_solve):
    from bem_amd import FlatCathode
    from bem_amd.geometry import lower

    class ReferenceLike:  # anything with the reference's worker_args()
        def worker_args(self):
            return ("hemispherical_tip", -1e8, 1e-6, 3e-9, 0.0)

    g = lower(ReferenceLike())
    assert g.worker_args() == ("hemispherical_tip", -1e8, 1e-6, 3e-9, 0.0)
    own = FlatCathode(-1e7)
    assert lower(own) is own
    with pytest.raises(ValueError):
        lower(("golden_post", 1.0))
    with pytest.raises(ValueError):
        lower((1.0, 2.0))


def test_tracker_takes_a_reference_geometry(no_device_solve):
    """SpecificParticleTracer rebuilds a foreign geometry from its tuple, as
    the reference's tracker does (tracker.py:199-202)."""
    from bem_amd import SpecificParticleTracer
    from bem_amd.emission import flat_gaussian_source
    from bem_amd.particles import ParticleGroup

    class ReferenceFlat:
        def worker_args(self):
            return ("flat", -1e7, 3e-9, 0.0)

    src = flat_gaussian_source(8, seed=1)
    pg = ParticleGroup(dict(src, weight=np.full(8, 1.602176634e-19), status=np.ones(8), species="electron"))
    tr = SpecificParticleTracer(pg, 4, ReferenceFlat(), [1e-6])
    assert type(tr.geometry).__name__ == "FlatCathode" and tr.geometry.E_gun == -1e7


def test_cache_tokens_are_never_reused():
    """Objects freed and reallocated in a row often get the same id(); the
    library context's table keys must not (a new geometry of the same panel
    count would otherwise silently reuse the previous one's sigma)."""
    from bem_amd import _lib

    class Obj:
        pass

    ids, toks = [], []
    for _ in range(50):
        o = Obj()
        ids.append(id(o))
        toks.append(_lib.token(o))
        del o
    assert len(set(toks)) == len(toks)
    assert len(set(ids)) < len(ids)  # CPython did reuse addresses: the case the tokens guard against
    o = Obj()
    assert _lib.token(o) == _lib.token(o)


@pytest.mark.parametrize("status,match", [(1, "within"), (2, "collapsed")])
def test_group_status_maps_to_runtime_error(status, match, monkeypatch):
    """A group that hits max_steps (reference batched_rk45.py:320-321) or
    stalls makes run() raise RuntimeError."""
    from bem_amd import FlatCathode, SpecificParticleTracer
    from bem_amd.emission import flat_gaussian_source
    from bem_amd.particles import ParticleGroup

    src = flat_gaussian_source(8, seed=1)
    pg = ParticleGroup(dict(src, weight=np.full(8, 1.602176634e-19), status=np.ones(8), species="electron"))
    tr = SpecificParticleTracer(pg, 4, FlatCathode(-1e7), [1e-6])
    st = np.zeros(2, dtype=np.int32)
    st[1] = status
    monkeypatch.setattr(tr, "trace", lambda: dict(status=st))
    with pytest.raises(RuntimeError, match=match):
        tr.run()
