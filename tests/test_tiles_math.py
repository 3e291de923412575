"""Chebyshev tiles (bem_amd.bem.tiles) on the CPU: the rectangle-segment
distance against brute force, the quadtree's window guarantees on the
400-panel post (every segment the reference's near rule or on-surface
offset would use at a point is in its tile's window; everything else is at
least alpha times the cell side away), the tree walk against the cells'
boxes, and the Chebyshev fit against polynomials it must reproduce and the
ring field of distant sources through the oracle (reference
bem/axisymmetric.py:232-265). No GPU."""
import numpy as np
import pytest

from conftest import golden
from bem_amd.bem import tiles as bt
from bem_amd.bem.multipole import quadrature_nodes
from oracle import bem as ob


@pytest.fixture(scope="module")
def post():
    d = golden("bem_post400.npz")
    return d["profile"], d["sigma"], float(d["R"]), float(d["H"])


@pytest.fixture(scope="module")
def tree(post):
    prof = post[0]
    size = 11.9e-6
    tr, tiles = bt.build_tree(prof, 0.0, 0.0, size, alpha=1.0, w_max=5, r_drop=size)
    return tr, tiles, size


def _seg_dist(px, pz, prof):
    return bt._point_seg(px, pz, prof[:-1], prof[1:])


def test_rect_seg_distance_matches_brute_force():
    rng = np.random.default_rng(1)
    for _ in range(200):
        x0, z0 = rng.uniform(-1, 1, 2)
        w, h = rng.uniform(0.05, 1, 2)
        p0 = rng.uniform(-2, 2, (1, 2))
        p1 = p0 + rng.uniform(-1.5, 1.5, (1, 2))
        got = bt.rect_seg_distance(x0, z0, x0 + w, z0 + h, p0, p1)[0]
        # brute force: dense points on the segment against the filled rectangle
        t = np.linspace(0, 1, 4001)[:, None]
        q = p0 + t * (p1 - p0)
        dx = np.maximum(np.maximum(x0 - q[:, 0], q[:, 0] - (x0 + w)), 0)
        dz = np.maximum(np.maximum(z0 - q[:, 1], q[:, 1] - (z0 + h)), 0)
        brute = np.hypot(dx, dz).min()
        assert got <= brute + 1e-12
        assert got >= brute - 1e-3 * np.hypot(*(p1 - p0)[0])  # sampling resolution
        assert (got == 0.0) == (brute == 0.0) or brute < 1e-3


def test_tree_walk_lands_in_the_cell(tree, post):
    tr, tiles, size = tree
    T = bt.Tiles(0.0, 0.0, size, tr, tiles, np.zeros((len(tiles), 2, bt.P, bt.P)), 1.0, 5)
    rng = np.random.default_rng(2)
    rho, z = rng.uniform(0, size, 4000), rng.uniform(0, size, 4000)
    t, cx, cz, h = T.find(rho, z)
    m = t >= 0
    assert m.mean() > 0.5
    c = tiles[t[m]]
    assert np.all(rho[m] >= c[:, 0]) and np.all(rho[m] < c[:, 0] + c[:, 2])
    assert np.all(z[m] >= c[:, 1]) and np.all(z[m] < c[:, 1] + c[:, 2])
    np.testing.assert_array_equal(cx[m], c[:, 0] + 0.5 * c[:, 2])
    np.testing.assert_array_equal(h[m], 0.5 * c[:, 2])
    # windowed tiles are numbered first
    wsize = tiles[:, 4] - tiles[:, 3]
    first_empty = np.argmax(wsize == 0)
    assert np.all(wsize[:first_empty] > 0) and np.all(wsize[first_empty:] == 0)


def test_windows_hold_every_near_segment(tree, post):
    """At points in a tile (near the cap and the wall, and on the surface):
    every segment closer than its own length (the near rule) and every
    segment at the point's closest distance (the on-surface offset's normal)
    is in the window; every segment outside it is at least alpha s away."""
    prof = post[0]
    R, H = post[2], post[3]
    tr, tiles, size = tree
    T = bt.Tiles(0.0, 0.0, size, tr, tiles, np.zeros((len(tiles), 2, bt.P, bt.P)), 1.0, 5)
    L = np.hypot(*np.diff(prof, axis=0).T)
    rng = np.random.default_rng(3)
    n = 3000
    th = 0.5 * np.pi * rng.random(n)
    d = np.concatenate([np.zeros(n // 3), 10 ** rng.uniform(-10, -5.5, n - n // 3)])
    rho, z = (R + d) * np.sin(th), H + (R + d) * np.cos(th)
    # and points on the wall
    rho[:200], z[:200] = R + d[n // 3:n // 3 + 200], rng.uniform(0.2e-6, 10e-6, 200)
    t, _, _, _ = T.find(rho, z)
    assert (t >= 0).mean() > 0.95
    checked = 0
    for i in np.nonzero(t >= 0)[0]:
        lo, hi = int(tiles[t[i], 3]), int(tiles[t[i], 4])
        s = tiles[t[i], 2]
        dist = _seg_dist(rho[i], z[i], prof)
        inside = np.zeros(len(L), dtype=bool)
        inside[lo:hi] = True
        assert np.all(inside[dist < L])
        j = np.argmin(dist)
        if dist[j] < max(s, L[j]):  # then the window's own search finds the same closest segment
            assert inside[j]
        assert np.all(dist[~inside] >= np.maximum(1.0 * s, L[~inside]) * (1 - 1e-9))
        checked += 1
    assert checked > 2500


def test_chebyshev_fit_reproduces_polynomials():
    rng = np.random.default_rng(4)
    c = rng.normal(size=(bt.P, bt.P))
    x = bt.cheb_nodes()
    vals = bt.cheb_eval(c, x[:, None] + 0 * x[None, :], 0 * x[:, None] + x[None, :])
    np.testing.assert_allclose(bt.fit_coefficients(vals), c, atol=1e-12)
    # a product of monomials of degree < P in each variable, at random points
    u, v = rng.uniform(-1, 1, 500), rng.uniform(-1, 1, 500)
    f = lambda a, b: a ** 11 * b ** 3 - 2 * a ** 4 * b ** 10 + a * b + 0.5  # noqa: E731
    cf = bt.fit_coefficients(f(x[:, None], x[None, :]))
    np.testing.assert_allclose(bt.cheb_eval(cf, u, v), f(u, v), atol=1e-13)
    # a transposed table (rho and z swapped) fails
    assert np.max(np.abs(bt.cheb_eval(cf.T, u, v) - f(u, v))) > 1e-2


def test_table_of_distant_rings_converges(tree, post):
    """The ring field of the fixed-rule nodes outside a surface cell's window
    (the sum a tile tabulates), through the oracle, fitted at the cell's
    Chebyshev points and checked inside the cell: within 1e-8 of |E_gun|
    (1e8 V/m)."""
    prof, sig = post[0], post[1]
    tr, tiles, size = tree
    rq, zq, wq = quadrature_nodes(prof, sig, 8)
    seg = np.repeat(np.arange(len(prof) - 1), 32)
    rng = np.random.default_rng(5)
    windowed = np.nonzero(tiles[:, 4] > tiles[:, 3])[0]
    for t in list(rng.choice(windowed, 3, replace=False)) + [len(tiles) - 1]:
        x0, z0, s, lo, hi = tiles[t]
        keep = (seg < lo) | (seg >= hi)
        x = bt.cheb_nodes()
        h = 0.5 * s

        def field(r, zz):
            er, ez = ob.ring_field(r[:, None], zz[:, None] - zq[None, keep], rq[None, keep])
            return er @ wq[keep], ez @ wq[keep]

        R_, Z_ = np.meshgrid(x0 + h + h * x, z0 + h + h * x, indexing="ij")
        er, ez = field(R_.ravel(), Z_.ravel())
        c = bt.fit_coefficients(np.stack([er.reshape(bt.P, bt.P), ez.reshape(bt.P, bt.P)]))
        u, v = rng.uniform(-1, 1, 200), rng.uniform(-1, 1, 200)
        dr, dz = field(x0 + h + h * u, z0 + h + h * v)
        err = np.hypot(bt.cheb_eval(c[0], u, v) - dr, bt.cheb_eval(c[1], u, v) - dz).max() / 1e8
        assert err <= 1e-8, (t, err)


def test_split_divides_the_listed_cells(tree, post):
    """A leaf named in split= becomes four children covering its cell
    (the refinement tiles.fit applies to cells whose table missed its
    tolerance); every other leaf stays as it was."""
    prof = post[0]
    tr, tiles, size = tree
    wins = tiles[:, 4] - tiles[:, 3]
    pick = int(np.nonzero(wins == 0)[0][len(np.nonzero(wins == 0)[0]) // 2])  # a table-only cell
    cell = tuple(float(v) for v in tiles[pick, :3])
    tr2, tiles2 = bt.build_tree(prof, 0.0, 0.0, size, alpha=1.0, w_max=5, r_drop=size, split=frozenset([cell]))
    before = {tuple(float(v) for v in t[:3]) for t in tiles}
    after = {tuple(float(v) for v in t[:3]) for t in tiles2}
    assert cell in before and cell not in after
    kids = after - before
    x0, z0, s = cell
    assert len(kids) == 4
    assert {(k[0], k[1]) for k in kids} == {(x0, z0), (x0 + s / 2, z0), (x0, z0 + s / 2), (x0 + s / 2, z0 + s / 2)}
    assert all(k[2] == s / 2 for k in kids)
    assert before - {cell} == after - kids
