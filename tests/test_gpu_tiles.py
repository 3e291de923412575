"""Chebyshev tiles on the GPU (spt_bem_remainder, spt_bem_tiles), on the
400-panel post with the reference's sigma:

* spt_bem_remainder equals the oracle's fixed-rule sum (reference
  bem/axisymmetric.py:305-326,560-564) over the segments outside random
  windows, to 1e-12 of the field;
* every fitted table matches the direct sum inside its cell within 1e-8 of
  |E_gun|, and the device's table-only points reproduce the host's
  Chebyshev evaluation of the same tables (1e-13);
* spt_bem_tiles refuses tables whose windowed tiles do not come first
  or whose tree places a tile in two cells;
* the tiled far-rule gather stays within 3e-8 of |E| of the plain fixed
  rule from 1e-3 panel lengths to 110 um off the cap, a permutation of
  points in windowed tiles permutes the results bit for bit (the binned
  gather's premise; the points all stay in the gather, so every block
  keeps its slice count -- the standalone gather takes table-only and
  multipole-zone points out of a block, and in the slice-major layout a
  block's slice count follows the points it keeps), and turning the tiles
  off restores the gather exactly.

Traces with tiles (bench.py's value form) are checked against the
reference runs and the plain trace in test_gpu_c5.py and test_gpu_binned.py
(their field budgets include the tiles)."""
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


@pytest.fixture(scope="module")
def tiles(post):
    from bem_amd.bem import tiles as bt

    tl = bt.fit(post, size=11.9e-6, r_drop=11.9e-6)
    _say(f"tiles: {tl.n_tiles}, table error max {tl.report['table_err_max']:.2e}, windows "
         f"{tl.report['window_counts']}")
    return tl


def test_remainder_matches_oracle(post):
    from bem_amd.bem import tiles as bt
    from bem_amd.bem.multipole import quadrature_nodes
    from oracle import bem as ob

    prof, sig, _ = post.bem_table()
    rq, zq, wq = quadrature_nodes(prof, sig, 8)
    seg = np.repeat(np.arange(400), 32)
    rng = np.random.default_rng(1)
    n = 64
    th = 0.5 * np.pi * rng.random(n)
    d = 10 ** rng.uniform(-7, -5.5, n)
    rho, z = (post.R + d) * np.sin(th), post.H + (post.R + d) * np.cos(th)
    lo = rng.integers(0, 390, n)
    win = np.stack([lo, lo + rng.integers(0, 10, n)], -1).astype(np.int32)
    win[:4] = 0  # empty windows: the whole sum
    got = bt.remainder(post, rho, z, win)
    for i in range(n):
        keep = (seg < win[i, 0]) | (seg >= win[i, 1])
        er, ez = ob.ring_field(rho[i], z[i] - zq[keep], rq[keep])
        ref = np.array([er @ wq[keep], ez @ wq[keep]])
        assert np.max(np.abs(got[i] - ref)) <= 1e-12 * np.max(np.abs(ref)), i


def test_tables_and_device_evaluation(post, tiles):
    from bem_amd.bem.multipole import _pert_field

    assert tiles.report["table_err_max"] <= 1e-8
    assert tiles.report["window_counts"][0] > 0  # table-only tiles exist
    # points in table-only tiles: the device's value is the host's table
    rng = np.random.default_rng(2)
    pts = []
    empty = np.nonzero(tiles.tiles[:, 4] <= tiles.tiles[:, 3])[0]
    for t in rng.choice(empty, 200, replace=False):
        x0, z0, s = tiles.tiles[t, :3]
        pts.append((x0 + s * rng.random(), z0 + s * rng.random()))
    rho, z = np.array(pts).T
    host = tiles.table(rho, z)
    with post.far_rule(2.0, 16.0, near_width=2.0, tiles=tiles):
        dev = _pert_field(post, rho, z)
    np.testing.assert_allclose(dev[:, 0], host[:, 0], rtol=1e-13, atol=1e-13 * np.abs(host).max())
    np.testing.assert_allclose(dev[:, 2], host[:, 1], rtol=1e-13, atol=1e-13 * np.abs(host).max())


def test_tiled_gather_against_the_fixed_rule(post, tiles):
    from bem_amd.surrogate import far_rule_error

    err = far_rule_error(post, 2.0, 16.0, n=1 << 15, near_width=2.0, tiles=tiles)
    _say(f"tiled gather vs fixed rule: max {err['field_err_max']:.2e} (at {err['d_at_max']:.2e} m), "
         f"median {err['field_err_median']:.2e}")
    assert err["field_err_max"] <= 3e-8


@pytest.mark.parametrize("shape", [(192, 64), (256, 64), (256, 8)])
def test_tiled_gather_is_order_independent(post, tiles, shape):
    from bem_amd.bem.multipole import _pert_field

    rng = np.random.default_rng(11)
    n = 64 * 96
    th = np.deg2rad(70.0) * rng.random(n)
    d = 10 ** rng.uniform(-10, np.log10(5e-8), n)  # within a panel length: windowed tiles
    rho, z = (post.R + d) * np.sin(th), post.H + (post.R + d) * np.cos(th)
    # a fifth exactly on the cap's panels (the reference's on-surface offset)
    prof = post.bem_table()[0]
    m = n // 5
    j = rng.integers(0, 27, m)
    t = rng.random(m)
    rho[:m] = prof[j, 0] + t * (prof[j + 1, 0] - prof[j, 0])
    z[:m] = prof[j, 1] + t * (prof[j + 1, 1] - prof[j, 1])
    t, _, _, _ = tiles.find(rho, z)
    assert np.all(t >= 0) and np.all(tiles.tiles[t, 4] > tiles.tiles[t, 3])
    perm = rng.permutation(n)
    ctx = post._ctx()
    ctx.check(ctx.lib.spt_set_field_shape(ctx.h, *shape), "spt_set_field_shape")
    try:
        with post.far_rule(2.0, 16.0, near_width=2.0, tiles=tiles):
            a = _pert_field(post, rho, z)
            b = _pert_field(post, rho[perm], z[perm])
    finally:
        ctx.lib.spt_set_field_shape(ctx.h, 0, 0)
    np.testing.assert_array_equal(b, a[perm])


def test_upload_rejects_bad_tables(post, tiles):
    """spt_bem_tiles checks what the device walk and the cell lookup trust:
    windowed tiles first, every tile in exactly one cell."""
    import ctypes

    from bem_amd import _lib

    ctx = post._ctx()
    Pp = ctypes.c_void_p

    def upload(tree, win):
        tree = np.ascontiguousarray(tree, dtype=np.int32)
        win = np.ascontiguousarray(win, dtype=np.int32)
        return ctx.lib.spt_bem_tiles(ctx.h, tree.ctypes.data_as(Pp), tree.shape[0], win.ctypes.data_as(Pp),
                                     tiles.coef.ctypes.data_as(Pp), tiles.n_tiles, tiles.x0, tiles.z0, tiles.size)

    tree, win = tiles.tree.copy(), tiles.win.copy()
    empty = int(np.argmax(win[:, 1] <= win[:, 0]))
    try:
        bad = win.copy()
        bad[[0, empty]] = bad[[empty, 0]]  # an empty window ahead of a windowed tile
        assert upload(tree, bad) == _lib.SPT_EINVAL
        twice = tree.copy()
        leaves = np.nonzero(twice <= -2)[0]
        twice[leaves[1]] = twice[leaves[0]]  # one tile in two cells, another in none
        assert upload(twice, win) == _lib.SPT_EINVAL
        assert upload(tree, win) == _lib.SPT_OK
    finally:
        ctx.tiles_key = None  # the next install uploads again
        ctx.lib.spt_bem_tiles_enable(ctx.h, 0)


def test_tiles_off_restores_the_gather(post, tiles):
    from bem_amd.bem.multipole import _pert_field

    rng = np.random.default_rng(7)
    rho, z = rng.uniform(0, 3e-6, 500), rng.uniform(11.05e-6, 11.8e-6, 500)
    with post.far_rule(2.0, 16.0, near_width=2.0):
        before = _pert_field(post, rho, z)
    with post.far_rule(2.0, 16.0, near_width=2.0, tiles=tiles):
        inside = _pert_field(post, rho, z)
    with post.far_rule(2.0, 16.0, near_width=2.0):
        after = _pert_field(post, rho, z)
    np.testing.assert_array_equal(before, after)
    assert not np.array_equal(before, inside)
