"""Chebyshev tiles of the BEM perturbation field (no counterpart in the
reference, whose plain method sums every segment at every point,
bem/axisymmetric.py:501-605): the level of the field budget that replaces
the distant segments of a far-rule gather by a table.

A quadtree covers a square of the meridian plane (rho, z). A cell of side s
becomes a tile once its window -- the hull [jlo, jhi) of the segments lying
within max(alpha s, L_j) of the cell -- holds at most w_max segments. A
point in the tile sums its window's segments directly (the far-rule
gather's near and far rules) and takes every other segment's field from the
tile's table: 12 x 12 Chebyshev coefficients of the fixed-rule sum over
those segments (8 sub-panels x Gauss-4 nodes, the reference's rule for a
segment at least a panel length away), fitted at the cell's Chebyshev
points from spt_bem_remainder. Those segments lie at least alpha s from the
cell, so their field is analytic over a neighbourhood of it and the tensor
Chebyshev series converges geometrically (about 1e-10 of |E| at alpha = 1).

The window holds every segment within a panel length of the cell, so the
closest segment of any point that needs the reference's on-surface offset,
and every segment its near rule takes, are summed directly as before.
Cells wholly inside a closed conductor profile or wholly beyond r_drop from
the origin get no tile (their points take the gather without tiles)."""
import ctypes

import numpy as np

from .. import _lib

P = 12  # Chebyshev terms per direction (kTileP in the library)
MAX_TILES = 6144
MAX_DEPTH = 24


def _point_seg(px, pz, a, b):
    ab = b - a
    t = np.clip(((px - a[:, 0]) * ab[:, 0] + (pz - a[:, 1]) * ab[:, 1]) / np.einsum("ij,ij->i", ab, ab), 0.0, 1.0)
    return np.hypot(px - a[:, 0] - t * ab[:, 0], pz - a[:, 1] - t * ab[:, 1])


def rect_seg_distance(x0, z0, x1, z1, p0, p1):
    """Distance from the rectangle [x0, x1] x [z0, z1] to every segment
    p0[j] -> p1[j]: zero where they meet, else the smallest
    corner-to-segment or endpoint-to-rectangle distance (two disjoint convex
    sets are closest at a vertex of one of them)."""
    dc = np.min([_point_seg(x, z, p0, p1) for x, z in ((x0, z0), (x1, z0), (x0, z1), (x1, z1))], axis=0)

    def to_rect(q):
        dx = np.maximum(np.maximum(x0 - q[:, 0], q[:, 0] - x1), 0.0)
        dz = np.maximum(np.maximum(z0 - q[:, 1], q[:, 1] - z1), 0.0)
        return np.hypot(dx, dz)

    de = np.minimum(to_rect(p0), to_rect(p1))
    # Liang-Barsky: does the segment meet the rectangle?
    dx = p1[:, 0] - p0[:, 0]
    dz = p1[:, 1] - p0[:, 1]
    t0 = np.zeros(len(p0))
    t1 = np.ones(len(p0))
    ok = np.ones(len(p0), dtype=bool)
    with np.errstate(divide="ignore", invalid="ignore"):
        for pp, qq in ((-dx, p0[:, 0] - x0), (dx, x1 - p0[:, 0]), (-dz, p0[:, 1] - z0), (dz, z1 - p0[:, 1])):
            par = pp == 0.0
            ok &= ~(par & (qq < 0.0))
            r = qq / pp
            t0 = np.where(~par & (pp < 0.0), np.maximum(t0, r), t0)
            t1 = np.where(~par & (pp > 0.0), np.minimum(t1, r), t1)
    meet = ok & (t0 <= t1)
    return np.where(meet, 0.0, np.minimum(dc, de))


def _inside(prof, px, pz):
    """Crossing-number test against a profile closed along the axis."""
    a, b = prof[:-1], prof[1:]
    cross = (a[:, 1] > pz) != (b[:, 1] > pz)
    with np.errstate(divide="ignore", invalid="ignore"):
        xi = a[:, 0] + (pz - a[:, 1]) * (b[:, 0] - a[:, 0]) / (b[:, 1] - a[:, 1])
    return bool(np.count_nonzero(cross & (px < xi)) % 2)


def build_tree(prof, x0, z0, size, alpha=1.0, w_max=5, s_min=None, r_drop=np.inf, split=frozenset()):
    """Quadtree over [x0, x0 + size) x [z0, z0 + size), children visited in
    the library's order (low/high rho, then low/high z) so tiles are
    numbered along a Z curve. Returns (tree int32, tiles (n, 5): cell x0,
    z0, side, jlo, jhi). A cell whose window still exceeds w_max at side
    s_min (default an eighth of the shortest segment) gets no tile; cells
    listed in split (x0, z0, side tuples, as this walk forms them) are
    divided even when their window is small enough (a table that missed
    its tolerance)."""
    prof = np.asarray(prof, dtype=np.float64)
    p0, p1 = prof[:-1], prof[1:]
    L = np.hypot(*(p1 - p0).T)
    s_min = L.min() / 8.0 if s_min is None else s_min
    closed = prof[0, 0] == 0.0 and prof[-1, 0] == 0.0
    tree = [0]
    tiles = []

    def visit(k, cx0, cz0, s, depth):
        nx = min(max(0.0, cx0), cx0 + s)
        nz = min(max(0.0, cz0), cz0 + s)
        if np.hypot(nx, nz) >= r_drop:
            tree[k] = -1
            return
        dist = rect_seg_distance(cx0, cz0, cx0 + s, cz0 + s, p0, p1)
        # margin: a point the device places in this cell may sit a rounding
        # error outside it
        win = np.nonzero(dist < np.maximum(alpha * s, L) + 1e-9 * s)[0]
        lo, hi = (int(win.min()), int(win.max()) + 1) if win.size else (0, 0)
        if hi - lo <= w_max and (cx0, cz0, s) not in split:
            if closed and dist.min() > 0.0 and _inside(prof, cx0 + 0.5 * s, cz0 + 0.5 * s):
                tree[k] = -1  # wholly inside the conductor
                return
            tree[k] = -2 - len(tiles)
            tiles.append((cx0, cz0, s, lo, hi))
            return
        h = s * 0.5
        if depth >= MAX_DEPTH or h < s_min:
            tree[k] = -1
            return
        base = len(tree)
        tree[k] = base
        tree.extend([0, 0, 0, 0])
        for qz in (0, 1):
            for qx in (0, 1):
                visit(base + qx + 2 * qz, cx0 + h if qx else cx0, cz0 + h if qz else cz0, h, depth + 1)

    visit(0, float(x0), float(z0), float(size), 0)
    if len(tiles) > MAX_TILES:
        raise ValueError(f"{len(tiles)} tiles: more than the library's {MAX_TILES}; raise w_max or shrink the box")
    tiles = np.asarray(tiles, dtype=np.float64).reshape(-1, 5)
    tree = np.asarray(tree, dtype=np.int32)
    # tiles with a window first, then the table-only ones (each group in Z
    # order): the binned gather's bins follow tile numbers, so table-only
    # points (evaluated one per thread) come after the windowed ones
    order = np.concatenate([np.nonzero(tiles[:, 4] > tiles[:, 3])[0], np.nonzero(tiles[:, 4] <= tiles[:, 3])[0]])
    new_id = np.empty(len(order), dtype=np.int64)
    new_id[order] = np.arange(len(order))
    leaf = tree <= -2
    tree[leaf] = -2 - new_id[-2 - tree[leaf]]
    return tree, tiles[order]


def cheb_nodes():
    """First-kind Chebyshev points cos(pi (i + 1/2) / P), i < P."""
    return np.cos(np.pi * (np.arange(P) + 0.5) / P)


def fit_coefficients(values):
    """values[..., i, j] at (x_i, y_j) (cheb_nodes) -> c[..., a, b] with
    f = sum_ab c_ab T_a(x) T_b(y) exact at the P x P points."""
    T = np.cos(np.outer(np.arange(P), np.arccos(cheb_nodes())))  # T[a, i]
    c = (2.0 / P) ** 2 * np.einsum("ai,...ij,bj->...ab", T, values, T)
    c[..., 0, :] *= 0.5
    c[..., :, 0] *= 0.5
    return c


def cheb_eval(coef, u, v):
    """sum_ab coef[..., a, b] T_a(u) T_b(v), u, v in [-1, 1] (by the
    three-term recurrence, as the library evaluates it)."""
    u = np.asarray(u, dtype=np.float64)
    v = np.asarray(v, dtype=np.float64)
    Tu = [np.ones_like(u), u]
    Tv = [np.ones_like(v), v]
    for _ in range(2, P):
        Tu.append(2.0 * u * Tu[-1] - Tu[-2])
        Tv.append(2.0 * v * Tv[-1] - Tv[-2])
    Tu = np.stack(Tu, -1)
    Tv = np.stack(Tv, -1)
    return np.einsum("...ab,...a,...b->...", coef, Tu, Tv)


def remainder(geometry, rho, z, win):
    """Device fixed-rule field (E_rho, E_z) at meridian points from every
    segment outside win[i] = [jlo, jhi) (spt_bem_remainder)."""
    import torch

    ctx = geometry._ctx()
    pts = torch.as_tensor(np.stack([rho, z], -1), dtype=torch.float64, device=ctx.torch_device).contiguous()
    w = torch.as_tensor(np.asarray(win, dtype=np.int32).reshape(-1, 2), device=ctx.torch_device).contiguous()
    out = torch.empty_like(pts)
    ctx.bind_stream()
    ctx.check(ctx.lib.spt_bem_remainder(ctx.h, _lib.ptr(pts), _lib.ptr(w), pts.shape[0], _lib.ptr(out)),
              "spt_bem_remainder")
    ctx.sync()
    return out.cpu().numpy()


class Tiles:
    """A tile quadtree with its tables for one BEM geometry's charge."""

    def __init__(self, x0, z0, size, tree, tiles, coef, alpha, w_max, report=None, geometry=None):
        self.x0, self.z0, self.size = float(x0), float(z0), float(size)
        self.tree = np.ascontiguousarray(tree, dtype=np.int32)
        self.tiles = np.asarray(tiles, dtype=np.float64)
        self.win = np.ascontiguousarray(self.tiles[:, 3:5], dtype=np.int32)
        self.coef = np.ascontiguousarray(coef, dtype=np.float64)  # (n, 2, P, P)
        self.alpha = float(alpha)
        self.w_max = int(w_max)
        self.report = report or {}
        self.geometry = geometry

    @property
    def n_tiles(self):
        return self.tiles.shape[0]

    def install(self, ctx):
        """Upload to the context (once per charge upload), or re-enable."""
        key = (_lib.token(self), ctx.bem_key)
        if getattr(ctx, "tiles_key", None) == key:
            ctx.check(ctx.lib.spt_bem_tiles_enable(ctx.h, 1), "spt_bem_tiles_enable")
            return
        Pp = ctypes.c_void_p
        ctx.check(ctx.lib.spt_bem_tiles(ctx.h, self.tree.ctypes.data_as(Pp), self.tree.shape[0],
                                        self.win.ctypes.data_as(Pp), self.coef.ctypes.data_as(Pp), self.n_tiles,
                                        self.x0, self.z0, self.size), "spt_bem_tiles")
        ctx.tiles_key = key

    @staticmethod
    def uninstall(ctx):
        ctx.lib.spt_bem_tiles_enable(ctx.h, 0)

    def find(self, rho, z):
        """Tile index of each point (-1 none) and the cell's centre and
        half-size, walking the tree with the library's arithmetic."""
        rho = np.atleast_1d(np.asarray(rho, dtype=np.float64))
        z = np.atleast_1d(np.asarray(z, dtype=np.float64))
        out = np.full(rho.shape, -1, dtype=np.int64)
        cx = np.zeros(rho.shape)
        cz = np.zeros(rho.shape)
        h = np.ones(rho.shape)
        for i in range(rho.size):
            x0, z0, s = self.x0, self.z0, self.size
            r, zz = rho.flat[i], z.flat[i]
            if not (x0 <= r < x0 + s and z0 <= zz < z0 + s):
                continue
            k = 0
            for _ in range(MAX_DEPTH + 1):
                v = int(self.tree[k])
                if v < 0:
                    h.flat[i] = 0.5 * s
                    cx.flat[i] = x0 + h.flat[i]
                    cz.flat[i] = z0 + h.flat[i]
                    out.flat[i] = -2 - v if v <= -2 else -1
                    break
                s *= 0.5
                qx = 1 if r >= x0 + s else 0
                qz = 1 if zz >= z0 + s else 0
                if qx:
                    x0 += s
                if qz:
                    z0 += s
                k = v + qx + 2 * qz
        return out, cx, cz, h

    def table(self, rho, z):
        """The tables' (E_rho, E_z) at points inside tiles (NaN elsewhere)."""
        t, cx, cz, h = self.find(rho, z)
        out = np.full(t.shape + (2,), np.nan)
        m = t >= 0
        u = (np.atleast_1d(rho)[m] - cx[m]) / h[m]
        v = (np.atleast_1d(z)[m] - cz[m]) / h[m]
        for c in range(2):
            out[m, c] = cheb_eval(self.coef[t[m], c], u, v)
        return out


def fit(geometry, x0=0.0, z0=0.0, size=None, alpha=1.0, w_max=5, r_drop=None, n_check=8, tol=1e-8, seed=3,
        refine=16):
    """Tiles for a BEM geometry's charge over [x0, x0 + size) x [z0, z0 +
    size) (size defaults to 1.05 times the profile's largest coordinate),
    tables from the device (spt_bem_remainder at every tile's Chebyshev
    points), each table checked at n_check random points of its cell
    against the direct sum. A cell whose table misses tol relative to
    |E_gun| is split and its children tabulated again, up to refine rounds
    (large cells away from the surface, where the whole charge is near
    enough to spoil a 12 x 12 table); raises ValueError if a table still
    misses."""
    prof, sigma, nsub = geometry.bem_table()
    prof = np.asarray(prof, dtype=np.float64)
    if size is None:
        size = 1.05 * float(np.max(np.abs(prof)))
    r_drop = np.inf if r_drop is None else float(r_drop)
    split = set()
    for rounds in range(refine + 1):
        out, worst = _fit_tables(geometry, prof, x0, z0, size, alpha, w_max, r_drop, n_check, seed, split)
        bad = np.nonzero(~(worst <= tol))[0]
        if bad.size == 0 or rounds == refine:
            break
        split |= {tuple(float(v) for v in out.tiles[t, :3]) for t in bad}
    out.report["refine_rounds"] = rounds
    out.report["cells_split_for_tables"] = len(split)
    if not worst.max() <= tol:
        raise ValueError(f"tile tables miss {tol:g} of |E_gun|: worst {worst.max():.3e} in tile {int(worst.argmax())}")
    return out


def _fit_tables(geometry, prof, x0, z0, size, alpha, w_max, r_drop, n_check, seed, split):
    """One tree and its tables; returns the Tiles and each table's worst
    error at its check points (relative to |E_gun|)."""
    tree, tiles = build_tree(prof, x0, z0, size, alpha=alpha, w_max=w_max, r_drop=r_drop, split=frozenset(split))
    n = tiles.shape[0]
    if n == 0:
        raise ValueError("no tiles: the box holds no cell outside the conductor")
    xs = cheb_nodes()
    h = 0.5 * tiles[:, 2]
    cx = tiles[:, 0] + h
    cz = tiles[:, 1] + h
    # points [tile, i (rho), j (z)]
    rho = cx[:, None, None] + h[:, None, None] * xs[None, :, None] + 0.0 * xs[None, None, :]
    zz = cz[:, None, None] + 0.0 * xs[None, :, None] + h[:, None, None] * xs[None, None, :]
    win = np.repeat(tiles[:, 3:5].astype(np.int32), P * P, axis=0)
    vals = remainder(geometry, rho.ravel(), zz.ravel(), win).reshape(n, P, P, 2)
    coef = np.ascontiguousarray(fit_coefficients(np.moveaxis(vals, -1, 1)))  # (n, 2, P, P)
    out = Tiles(x0, z0, size, tree, tiles, coef, alpha, w_max, geometry=geometry)
    # check every table at random points of its cell
    rng = np.random.default_rng(seed)
    u = rng.uniform(-1.0, 1.0, (n, n_check))
    v = rng.uniform(-1.0, 1.0, (n, n_check))
    pr = (cx[:, None] + h[:, None] * u).ravel()
    pz = (cz[:, None] + h[:, None] * v).ravel()
    direct = remainder(geometry, pr, pz, np.repeat(tiles[:, 3:5].astype(np.int32), n_check, axis=0))
    tab = np.stack([cheb_eval(np.repeat(coef[:, c], n_check, axis=0), u.ravel(), v.ravel()) for c in range(2)], -1)
    err = np.hypot(*(tab - direct).T) / abs(geometry.E_gun)
    worst = err.reshape(n, n_check).max(axis=1)
    counts = np.bincount((tiles[:, 4] - tiles[:, 3]).astype(np.int64), minlength=w_max + 1)
    out.report = dict(n_tiles=int(n), n_tree=int(tree.shape[0]), alpha=float(alpha), w_max=int(w_max),
                      box=[float(x0), float(z0), float(size)], r_drop=None if np.isinf(r_drop) else r_drop,
                      window_counts=counts.tolist(), table_err_max=float(worst.max()),
                      table_err_median=float(np.median(worst)), table_bytes=int(coef.nbytes),
                      smallest_cell=float(tiles[:, 2].min()), largest_cell=float(tiles[:, 2].max()))
    return out, worst
