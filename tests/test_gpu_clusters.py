"""The fast form's gather refinements on the GPU (spt_bem_clusters,
spt_bem_near_width), on the 400-panel post with the reference's sigma:

* where every cluster is far, the far-rule gather is the sum of the device
  cluster expansions alone, equal to the host formula
  (bem_amd.bem.clusters.field, itself pinned to direct ring sums on the
  CPU) to 1e-12 of |E|;
* the near rule on intervals of width 2 against width 0.5 (the
  reference's) at 1e-5 .. 1 panel lengths from the surface: within 2e-8 of
  |E| (CPU measurement against width 0.25: 6.8e-9);
* the whole budget (far rule, near width, clusters) against the plain
  fixed rule within 1e-7 of |E| (measured 3.0e-8), and off again after its
  block: bit-identical to a fixed-rule gather."""
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
