"""The fast form's gather refinements on the GPU (spt_bem_clusters,
spt_bem_near_width), on the 400-panel post with the reference's sigma:

* where every cluster is far, the far-rule gather is the sum of the device
  cluster expansions alone, equal to the host formula
  (bem_amd.bem.clusters.field, itself pinned to direct ring sums on the
  CPU) to 1e-12 of |E|;
* the near rule on intervals of width 2 against width 0.5 (the
  reference's) at 1e-5 .. 1 panel lengths from the surface: within 2e-8 of
  |E| (CPU measurement against width 0.25: 6.8e-9);
* the near rule's pair form (far-rule gathers with 2 or 4 slices per
  point: one exp per interval and per node pair) against its node-by-node
  form (1 or 16 slices per point) on and near the cap, at widths 0.5 and
  2: within 1e-11 of |E|. The two node-by-node shapes place every node
  identically and agree to 6e-15; the pair form's products move nodes by
  a few ulp, and a node coordinate (about 1 um) moved by one ulp changes a
  node-point separation of 1e-13 m (1e-5 panel lengths) by ~1e-9 of
  itself, so the forms differ by up to 1.9e-12 (measured). A misplaced or
  dropped node or a wrong weight shows at 1e-8 or more;
* the whole budget (far rule, near width, clusters) against the plain
  fixed rule within 1e-7 of |E| (measured 3.0e-8), and off again after its
  block: bit-identical to a fixed-rule gather."""
import sys

import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu


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


def _say(*a):
    print(*a, flush=True, file=sys.stderr)


def _pert(geo, rho, z):
    from bem_amd.bem.multipole import _pert_field

    return _pert_field(geo, rho, z)


def test_device_expansions_equal_the_host_formula(post):
    from bem_amd.bem import clusters as bc

    cl = bc.fit(post)
    rng = np.random.default_rng(5)
    rho = rng.uniform(2.6e-6, 15e-6, 4000)
    z = rng.uniform(0.5e-6, 25e-6, 4000)
    for c in range(cl.n_clusters):
        assert np.all(np.hypot(rho, z - cl.zc[c]) >= cl.kappa * cl.rc[c])
    want_r, want_z = np.zeros_like(rho), np.zeros_like(rho)
    for c in range(cl.n_clusters):
        gr, gz = bc.field(cl, c, rho, z)
        want_r += gr
        want_z += gz
    with post.far_rule(2.0, 16.0, clusters=cl):
        got = _pert(post, rho, z)
    Et = np.hypot(want_r, want_z - 1e8)
    err = np.hypot(got[:, 0] - want_r, got[:, 2] - want_z) / Et
    assert err.max() <= 1e-12, err.max()


def test_wide_near_rule_within_budget(post):
    prof = post.solution.profile
    seg = np.diff(prof, axis=0)
    L = np.linalg.norm(seg, axis=1)
    nrm = np.stack([-seg[:, 1], seg[:, 0]], 1) / L[:, None]
    rng = np.random.default_rng(6)
    j = rng.choice(np.nonzero(prof[:-1, 1] > 8e-6)[0], 6000)
    u = rng.random(6000)
    d = L[j] * 10 ** rng.uniform(-5, 0, 6000)
    rz = prof[j] + u[:, None] * seg[j] + d[:, None] * nrm[j]
    with post.far_rule(2.0, 16.0):
        ref = _pert(post, rz[:, 0], rz[:, 1])
    with post.far_rule(2.0, 16.0, near_width=2.0):
        got = _pert(post, rz[:, 0], rz[:, 1])
    Et = np.hypot(ref[:, 0], ref[:, 2] - 1e8)
    err = np.hypot(got[:, 0] - ref[:, 0], got[:, 2] - ref[:, 2]) / Et
    assert err.max() <= 2e-8, err.max()


@pytest.mark.parametrize("width", [0.5, 2.0])
def test_pair_form_equals_the_node_by_node_form(post, width):
    prof = post.solution.profile
    seg = np.diff(prof, axis=0)
    L = np.linalg.norm(seg, axis=1)
    nrm = np.stack([-seg[:, 1], seg[:, 0]], 1) / L[:, None]
    rng = np.random.default_rng(8)
    n = 64 * 64
    j = rng.choice(np.nonzero(prof[:-1, 1] > 8e-6)[0], n)
    u = rng.random(n)
    d = L[j] * 10 ** rng.uniform(-6, 0, n)
    d[: n // 4] = 0.0  # on the panels: the on-surface offset, the longest sinh spans
    rz = prof[j] + u[:, None] * seg[j] + d[:, None] * nrm[j]
    ctx = post._ctx()
    out = {}
    try:
        for shape in [(64, 64), (128, 64), (256, 64), (256, 16)]:
            ctx.check(ctx.lib.spt_set_field_shape(ctx.h, *shape), "spt_set_field_shape")
            with post.far_rule(2.0, 16.0, near_width=width):
                out[shape] = _pert(post, rz[:, 0], rz[:, 1])
    finally:
        ctx.lib.spt_set_field_shape(ctx.h, 0, 0)
    ref = out[(256, 16)]  # 16 slices per point: node by node
    Et = np.hypot(ref[:, 0], ref[:, 2] - 1e8)
    dec = np.where(d > 0, np.floor(np.log10(np.maximum(d / L[j], 1e-300))), -9)
    for shape in [(64, 64), (128, 64), (256, 64)]:  # node by node, pairs, pairs
        got = out[shape]
        err = np.hypot(got[:, 0] - ref[:, 0], got[:, 2] - ref[:, 2]) / Et
        by = {int(k): float(err[dec == k].max()) for k in np.unique(dec)}
        _say(f"width {width} {shape}: max difference by log10(d / L) (-9: on the panel) {by}")
        assert err.max() <= 1e-11, (shape, err.max())


def test_budget_against_the_fixed_rule_and_reset(post):
    from bem_amd.bem import budget

    bud = budget.fit(post, tol=1e-6)
    assert bud.report["gather"]["field_err_max"] <= 1e-7
    assert not bud.multipole.inert
    rng = np.random.default_rng(7)
    rho, z = rng.uniform(0, 3e-6, 500), rng.uniform(11.2e-6, 14e-6, 500)
    before = _pert(post, rho, z)
    with post.far_rule(*bud.far_rule, near_width=bud.near_width, clusters=bud.clusters):
        inside = _pert(post, rho, z)
    after = _pert(post, rho, z)
    np.testing.assert_array_equal(before, after)
    assert not np.array_equal(before, inside)
