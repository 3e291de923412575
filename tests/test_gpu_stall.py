"""On-device stall detection: a group whose step is rejected down below
1e-14 |t| stops with SPT_STATUS_STALLED in both trace kernels, and run()
raises the reference's RuntimeError (reference batched_rk45.py:320-321,
whose only exit is max_steps). The stall is forced from the host by an
unreachable tolerance (rtol 1e-30, atol 0): every attempt is rejected and
h shrinks by the minimum factor 0.2, so the check fires after a dozen
attempts. max_steps is kept small, so the kernel would end through the
step cap even if the check did not fire."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from bem_amd import _lib

    _lib.load()
    return torch


def _tracer(mode, rtol, atol, groups=3):
    from bem_amd import FlatCathode, SpecificParticleTracer
    from bem_amd.emission import flat_gaussian_source
    from bem_amd.particles import ParticleGroup

    n = 4 * groups
    src = flat_gaussian_source(n, seed=2)
    src["t"] = np.full(n, 1e-12)
    pg = ParticleGroup(dict(src, weight=np.full(n, 1.602176634e-19), status=np.ones(n), species="electron"))
    return SpecificParticleTracer(pg, 4, FlatCathode(-1e8, z0=3e-9), [1e-6], rtol=rtol, atol=atol, t_max=2e-12,
                                  mode=mode, max_steps=200)


@pytest.mark.parametrize("mode", ["persistent", "lockstep"])
def test_collapsed_step_stops_the_group(gpu, mode):
    from bem_amd import _lib

    tr = _tracer(mode, 1e-30, 0.0)
    out = tr.trace()
    assert np.all(out["status"] == _lib.STATUS_STALLED), out["status"]
    # h0 = 1e-6 (t_max - t) = 1e-18 s; 0.2^k h0 <= 1e-14 * 1e-12 s first at k = 12
    assert np.all(out["attempts"] == 12), out["attempts"]
    assert np.all(out["accepted"] == 0)
    np.testing.assert_array_equal(out["pos"], tr.pos0)  # nothing was accepted
    with pytest.raises(RuntimeError, match="collapsed"):
        tr.run()


@pytest.mark.parametrize("mode", ["persistent", "lockstep"])
def test_ordinary_runs_do_not_stall(gpu, mode):
    from bem_amd import _lib

    tr = _tracer(mode, 1e-8, 1e-12)
    tr.max_steps = 10_000_000
    out = tr.trace()
    assert np.all(out["status"] == _lib.STATUS_DONE)
