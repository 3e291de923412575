"""Distance-adaptive far rule (spt_bem_far_rule) on the GPU.

The far rule is a field-error budget for the fast trace, not part of the
plain method: a segment at distance D >= L from the point is summed over
the whole segment by Gauss-8 below k4 L, Gauss-4 below k2 L and Gauss-2
beyond, instead of the reference's 8 x Gauss-4 sub-panel table. Checks:

* off by default and after the context ends: the gather is bit-identical to
  the fixed rule;
* on the 400-panel post, (k4, k2) = (2, 16): field error <= 1e-7 of |E| at
  points from 1e-3 panel lengths to 110 um off the surface (the CPU
  restatement with the oracle's ring kernel gives <= 3e-8);
* a lockstep trace with it: every electron crosses the screen, crossing
  times within 1e-6 relative of the fixed-rule trace (measured 2.9e-8),
  and it is faster;
* bad (k4, k2) are rejected.
"""
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


def _post400():
    from bem_amd.bem.geometry import PostTipBEMGeometry

    d = golden("bem_post400.npz")
    return PostTipBEMGeometry(-1e8, float(d["R"]), float(d["H"]), n_segments=400, sigma=d["sigma"])


def test_far_rule_off_is_the_fixed_rule(gpu):
    from bem_amd.surrogate import bem_labels, cap_samples

    geo = _post400()
    rng = np.random.default_rng(3)
    rho, z, _ = cap_samples(geo.R, geo.H, 2048, 1e-9, 50e-6, rng, np.deg2rad(100.0))
    a = bem_labels(geo, rho, z).cpu().numpy()
    with geo.far_rule(2.0, 16.0):
        b = bem_labels(geo, rho, z).cpu().numpy()
    c = bem_labels(geo, rho, z).cpu().numpy()
    assert np.array_equal(a, c)
    assert not np.array_equal(a[:, 1:], b[:, 1:])


def test_far_rule_field_error(gpu):
    from bem_amd.surrogate import far_rule_error

    geo = _post400()
    rep = far_rule_error(geo, 2.0, 16.0, n=1 << 14)
    _say(f"far rule (2, 16): field err max {rep['field_err_max']:.2e} (at d {rep['d_at_max']:.2e} m), "
         f"p99 {rep['field_err_p99']:.2e}, median {rep['field_err_median']:.2e} over {rep['n']} points")
    assert rep["field_err_max"] <= 1e-7


def test_far_rule_rejects_bad_parameters(gpu):
    from bem_amd import _lib

    geo = _post400()
    ctx = geo._ctx()
    assert ctx.lib.spt_bem_far_rule(ctx.h, 0.5, 16.0) == _lib.SPT_EINVAL
    assert ctx.lib.spt_bem_far_rule(ctx.h, 4.0, 2.0) == _lib.SPT_EINVAL
    assert ctx.lib.spt_bem_far_rule(ctx.h, 0.0, 0.0) == _lib.SPT_OK


def test_far_rule_trace(gpu):
    """4096 electrons (256 groups of 16, rtol 1e-8, FP32 pairs) traced in
    lockstep with the fixed rule and with the far rule."""
    from bem_amd import SpecificParticleTracer
    from bem_amd.emission import cap_source
    from bem_amd.particles import ParticleGroup

    geo = _post400()
    n = 4096
    src = cap_source(n, geo.R, geo.H, seed=2)
    pg = ParticleGroup(dict(src, weight=np.full(n, 1.602176634e-19), status=np.ones(n), species="electron"))
    runs = {}
    for name, fr in (("fixed", None), ("far", (2.0, 16.0))):
        tr = SpecificParticleTracer(pg, 16, geo, [90e-6], rtol=1e-8, z_max=100e-6, mode="lockstep",
                                    precision="pairs32", far_rule=fr)
        tr.trace()  # warm-up (graph capture, first launches)
        out = tr.trace()
        runs[name] = out
        _say(f"trace [{name}]: {out['device_ms']:.1f} ms, attempts mean {out['attempts'].mean():.2f}")
    a, b = runs["fixed"], runs["far"]
    ea, eb = a["events"], b["events"]
    assert len(ea) == len(eb) == n
    ka = np.lexsort((ea["particle"], ea["group"]))
    kb = np.lexsort((eb["particle"], eb["group"]))
    assert np.array_equal(ea["group"][ka], eb["group"][kb]) and np.array_equal(ea["particle"][ka], eb["particle"][kb])
    dt = np.abs(ea["t"][ka] - eb["t"][kb]) / ea["t"][ka]
    gain = a["device_ms"] / b["device_ms"]
    _say(f"far rule gain {gain:.2f}x; screen-time rel diff max {dt.max():.2e} median {np.median(dt):.2e}; "
         f"attempts differ in {int(np.sum(a['attempts'] != b['attempts']))} of {len(a['attempts'])} groups")
    assert dt.max() < 1e-6
    # the speed of the far rule is bench.py's claim (262k electrons); at 4096
    # electrons the trace is mostly launch latency, and a shared box's timing
    # is no correctness criterion, so the gain is only printed here
