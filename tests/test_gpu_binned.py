"""The binned gather of far-rule lockstep traces (spt_set_lockstep_binning),
on the 400-panel post with the reference's sigma.

Binning hands each gather block 64 points of one distance band and
neighbouring closest segments, in slot order made by atomics; that is sound
only because a far-rule gather's result for a point does not depend on the
other points of its block (each lane picks its own rule). So:

* the far-rule gather (far rule, near width 2, cluster expansions) of a
  permutation of the points is the permutation of the results, bit for
  bit, in the wave-major and the slice-major layouts;
* a binned trace of 16,384 cap electrons under the field budget is
  bit-reproducible, and equal bit for bit to the trace whose every gather
  takes the append path (slots in atomic order, mode 2);
* against the unbinned trace (particle order, whose partial blocks sum with
  another slice count) it keeps every crossing, times within 1e-7
  relative, at the default launch shape and at 64, 128 and 256 threads
  (the few-records form's slice reductions are wavefront shuffles there,
  LDS trees at 192)."""
import sys

import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True, file=sys.stderr)


@pytest.fixture(scope="module")
def post():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from bem_amd import _lib
    from bem_amd.bem.geometry import PostTipBEMGeometry

    _lib.load()
    d = golden("bem_post400.npz")
    return PostTipBEMGeometry(-1e8, float(d["R"]), float(d["H"]), n_segments=400, sigma=d["sigma"])


@pytest.mark.parametrize("shape", [(192, 64), (256, 8)])
def test_far_gather_is_order_independent(post, shape):
    from bem_amd.bem import clusters as bcl
    from bem_amd.bem.multipole import _pert_field

    cl = bcl.fit(post, seg_per_cluster=50, kappa=1.2)
    rng = np.random.default_rng(11)
    n = 64 * 96  # whole blocks at either shape, so every block has the same slice count
    th = np.deg2rad(60.0) * rng.random(n)
    d = post.R + 10 ** rng.uniform(-8.5, -5.5, n)
    rho, z = d * np.sin(th), post.H + d * np.cos(th)
    perm = rng.permutation(n)
    ctx = post._ctx()
    ctx.check(ctx.lib.spt_set_field_shape(ctx.h, *shape), "spt_set_field_shape")
    try:
        with post.far_rule(2.0, 16.0, near_width=2.0, clusters=cl):
            a = _pert_field(post, rho, z)
            b = _pert_field(post, rho[perm], z[perm])
    finally:
        ctx.lib.spt_set_field_shape(ctx.h, 0, 0)
    np.testing.assert_array_equal(b, a[perm])


@pytest.fixture(scope="module")
def workload(post):
    from bem_amd.bem import budget
    from bem_amd.emission import cap_source
    from bem_amd.particles import ParticleGroup

    n = 16384
    src = cap_source(n, post.R, post.H, theta_max_deg=60.0, mean_ke_ev=0.2, seed=0)
    pg = ParticleGroup(dict(src, weight=np.full(n, 1.602176634e-19), status=np.ones(n), species="electron"))
    return pg, budget.fit(post, tol=1e-6, n_val=1 << 12)


def _trace(post, workload, binned):
    from bem_amd import SpecificParticleTracer

    pg, bud = workload
    tr = SpecificParticleTracer(pg, 16, post, [90e-6], rtol=1e-8, atol=1e-12, z_max=100e-6, mode="lockstep",
                                precision="pairs32", binned=binned, **bud.tracer_kwargs())
    return tr.trace()


def _same_crossings(a, b, what):
    ka = np.argsort(a["events"]["group"].astype(np.int64) * 16 + a["events"]["particle"])
    kb = np.argsort(b["events"]["group"].astype(np.int64) * 16 + b["events"]["particle"])
    np.testing.assert_array_equal(a["events"]["group"][ka], b["events"]["group"][kb])
    np.testing.assert_array_equal(a["events"]["particle"][ka], b["events"]["particle"][kb])
    dt = np.abs(a["events"]["t"][ka] - b["events"]["t"][kb]) / b["events"]["t"][kb]
    _say(f"{what}: crossing time rel diff max {dt.max():.2e}, attempts differ in "
         f"{int(np.sum(a['attempts'] != b['attempts']))} of {len(a['attempts'])} groups")
    assert dt.max() <= 1e-7


@pytest.fixture(scope="module")
def traces(post, workload):
    runs = {}
    for name, binned in (("binned", 1), ("binned_again", 1), ("append", 2), ("unbinned", 0)):
        runs[name] = o = _trace(post, workload, binned)
        _say(f"16,384 e- [{name}]: {o['device_ms']:.1f} ms, attempts mean {o['attempts'].mean():.2f}, "
             f"{len(o['events'])} events")
    return runs


def test_binned_trace_is_reproducible_and_slot_order_free(traces):
    a = traces["binned"]
    assert np.all(a["status"] == 0) and len(a["events"]) == 16384
    for other in ("binned_again", "append"):
        b = traces[other]
        for k in ("pos", "vel", "t_death", "attempts", "accepted", "t_end"):
            np.testing.assert_array_equal(a[k], b[k], err_msg=f"{other}: {k}")
        np.testing.assert_array_equal(np.sort(a["events"], order=["group", "particle"]),
                                      np.sort(b["events"], order=["group", "particle"]), err_msg=other)


def test_binned_trace_matches_the_unbinned_one(traces):
    _same_crossings(traces["binned"], traces["unbinned"], "binned vs unbinned")


@pytest.mark.parametrize("threads", [64, 128, 256])
def test_binned_trace_at_each_shuffle_shape(post, workload, traces, threads):
    """At T = 64, 128 and 256 the few-records form (T / 16 = 4, 8, 16 slices
    per point) reduces over a point's slices by wavefront shuffles; the
    binned trace there keeps the unbinned trace's crossings."""
    ctx = post._ctx()
    ctx.check(ctx.lib.spt_set_field_shape(ctx.h, threads, 64), "spt_set_field_shape")
    try:
        a = _trace(post, workload, 1)
    finally:
        ctx.lib.spt_set_field_shape(ctx.h, 0, 0)
    assert np.all(a["status"] == 0) and len(a["events"]) == 16384
    _same_crossings(a, traces["unbinned"], f"binned at ({threads}, 64) vs unbinned")
