"""spt_trace_noalloc (the persistent trace on the library's static device
scratch, for callers that cannot allocate) through the C ABI against
spt_trace on the same inputs: identical counters, states, death times and
screen events, on two calls in a row, the second with fewer particles; and
an input beyond the static scratch rejected with ValueError."""
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


def _tracer(n, groups_of, seed):
    from bem_amd import SpecificParticleTracer
    from bem_amd.bem.geometry import PostTipBEMGeometry
    from bem_amd.emission import cap_source
    from bem_amd.particles import ParticleGroup

    d = golden("bem_post80.npz")
    R, H = float(d["R"]), float(d["H"])
    geo = PostTipBEMGeometry(-1e8, R, H, n_segments=80, sigma=d["sigma"])
    src = cap_source(n, R, H, seed=seed)
    pg = ParticleGroup(dict(src, weight=np.full(n, 1.602176634e-19), status=np.ones(n), species="electron"))
    return SpecificParticleTracer(pg, groups_of, geo, [12e-6, 20e-6], rtol=1e-8, t_max=2e-12, mode="persistent")


def _same(a, b):
    for k in ("attempts", "accepted", "active_stages", "status", "pos", "vel", "t_death", "margin", "t_end"):
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)
    ka = np.lexsort((a["events"]["particle"], a["events"]["group"], a["events"]["screen"]))
    kb = np.lexsort((b["events"]["particle"], b["events"]["group"], b["events"]["screen"]))
    assert len(ka) == len(kb) > 0
    for f in ("group", "particle", "screen", "interval", "batch", "t", "pos", "vel"):
        np.testing.assert_array_equal(a["events"][f][ka], b["events"][f][kb], err_msg=f)


def test_noalloc_trace_equals_spt_trace_twice(gpu):
    for n, g, seed in ((1024, 16, 3), (96, 8, 4)):  # the second call smaller than the first
        tr = _tracer(n, g, seed)
        ref = tr.trace()
        got = tr.trace(no_alloc=True)
        _same(ref, got)
        assert np.all(got["status"] == 0)


def test_noalloc_rejects_what_does_not_fit(gpu):
    from bem_amd import FlatCathode, SpecificParticleTracer
    from bem_amd.emission import flat_gaussian_source
    from bem_amd.particles import ParticleGroup

    n = 155_360  # 54 doubles each: just over the 64 MiB scratch
    src = flat_gaussian_source(n, seed=1)
    pg = ParticleGroup(dict(src, weight=np.full(n, 1.602176634e-19), status=np.ones(n), species="electron"))
    tr = SpecificParticleTracer(pg, 16, FlatCathode(-1e8), [1e-6], t_max=1e-13, mode="persistent")
    with pytest.raises(ValueError, match="static scratch"):
        tr.trace(no_alloc=True)
    lock = SpecificParticleTracer(pg, 16, FlatCathode(-1e8), [1e-6], t_max=1e-13, mode="lockstep")
    with pytest.raises(ValueError, match="persistent"):
        lock.trace(no_alloc=True)
