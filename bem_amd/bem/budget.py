"""The field-error budget of the fast form (no counterpart in the
reference, whose plain method sums every segment with the fixed rule at
every point, bem/axisymmetric.py:501-605).

A FieldBudget bundles the exact-in-the-limit approximations of the BEM
perturbation that the fast traces use, each converging to the plain
method's fixed-rule sum:

* the distance-adaptive far rule (k4, k2): Gauss-8 / Gauss-4 / Gauss-2 per
  segment instead of 8 sub-panels x Gauss-4 (spt_bem_far_rule);
* a wider near rule: sinh-substituted Gauss-8 on intervals of width
  near_width instead of 0.5 (spt_bem_near_width);
* cluster expansions: axis-centred exterior expansions of runs of
  segments beyond kappa times their radius (spt_bem_clusters,
  bem_amd.bem.clusters);
* tiles: per quadtree cell a Chebyshev table of the fixed-rule field of
  every segment outside the cell's window (spt_bem_tiles,
  bem_amd.bem.tiles); points in a tile sum only the window directly;
* the exterior multipole zone of the whole charge (spt_bem_multipole,
  bem_amd.bem.multipole).

fit() validates the first four together against the device's fixed rule
on fresh points around the cap (field error relative to |E|) and the
multipole zone on its own, and refuses a budget that misses the
tolerance."""
from . import clusters as _clusters
from . import multipole as _multipole
from . import tiles as _tiles


class FieldBudget:
    def __init__(self, geometry, far_rule, near_width, clusters, multipole, report, tiles=None):
        self.geometry = geometry
        self.far_rule = far_rule
        self.near_width = near_width
        self.clusters = clusters
        self.multipole = multipole
        self.tiles = tiles
        self.report = report

    def tracer_kwargs(self):
        """Keyword arguments of SpecificParticleTracer for this budget."""
        return dict(far_rule=self.far_rule, near_width=self.near_width, clusters=self.clusters, tiles=self.tiles,
                    multipole=None if self.multipole is None or self.multipole.inert else self.multipole)


def fit(geometry, tol=1e-6, far_rule=(2.0, 16.0), near_width=2.5, seg_per_cluster=50, kappa=1.2,
        multipole_terms=128, multipole_tol=None, n_val=1 << 16, use_clusters=True, use_multipole=True,
        cluster_terms=128, use_tiles=True, tile_alpha=0.25, tile_w_max=4, zone_factor=2.5):
    """Build and validate a budget for a protrusion BEM geometry (the post or
    the hemispherical tip). Raises ValueError when the validated far-rule
    gather (far rule + near width + clusters + tiles) misses tol.

    The defaults come from sweeps on the 400-panel post's 262,144-electron
    binned trace (far-rule thresholds, near width, cluster size, kappa and
    expansion length against the validated error and the trace time):
    clusters of up to 50 segments within each straight or turning run of
    the profile (the cap is one), each about its smallest enclosing sphere,
    expanded beyond 1.2 r_c with up to 128 terms; near width 2 (2.5 with
    the tiles, below); far rule (2, 16): field error 1.8e-8 of |E|, 86 ms
    per trace without tiles (35
    segments beyond 1.35 r_c about their z-midpoints: 92 ms; kappa 1.1 ..
    1.15 or 70 .. 100 segments: 88 .. 90 ms; near width 2.5: 6.6e-8; far
    rule (2, 10): 1.7e-7). Tiles cover the quarter plane rho, z >= 0 up to
    the multipole zone's radius (the profile's extent without it), windows
    of at most tile_w_max segments within tile_alpha cell sides. The zone
    starts at zone_factor r_ref or at its validated radius, whichever is
    larger: just beyond r_ref the expansion needs all its terms (q =
    r_ref / r near 1), where one table per point is cheaper. Tile defaults
    from a sweep of the same trace (ms per trace): zone 1.083 r_ref (the
    validated radius), w_max 5, alpha 1: 53.5; zone 2.5 r_ref: 50.2; then
    w_max 4: 49.5, alpha 0.5: 47.9, alpha 0.25: 47.4, alpha 0.15: 46.9
    (gather error 1.9e-8 against 7.1e-9); w_max 3 leaves cells without a
    tile (84 ms); the box size moves the quadtree's cell edges against the
    cap (zone 2.0 / 4.0 / 8.0 r_ref: 53.5 .. 53.9, 3.0 / 6.0: 51.1 .. 51.3,
    2.5 / 5.0: 49.4 .. 50.2 at w_max 5, alpha 1). With these tiles, near
    width 2.5 against 2: 44.7 against 46.7 ms at a validated gather error
    of 6.6e-8 against 7.1e-9 (width 3: 44.0 ms at 4.8e-7, half the
    metric's tolerance; not taken)."""
    from ..surrogate import far_rule_error

    cl = (_clusters.fit(geometry, seg_per_cluster=seg_per_cluster, n_terms=cluster_terms, kappa=kappa)
          if use_clusters else None)
    mp = None
    if use_multipole:
        mp = _multipole.fit(geometry, n_terms=multipole_terms, tol=min(1e-8, tol) if multipole_tol is None
                            else multipole_tol)
    tl = None
    if use_tiles:
        live_mp = mp is not None and not mp.inert
        if live_mp:
            # the expansion stays valid beyond its validated radius
            mp.report["r_min_validated"] = mp.r_min
            mp.r_min = max(mp.r_min, zone_factor * mp.r_ref)
            mp.report["r_min"] = mp.r_min
            mp.report["r_min_over_r_ref"] = mp.r_min / mp.r_ref
        tl = _tiles.fit(geometry, size=mp.r_min if live_mp else None, r_drop=mp.r_min if live_mp else None,
                        alpha=tile_alpha, w_max=tile_w_max)
    err = far_rule_error(geometry, *far_rule, n=n_val, near_width=near_width, clusters=cl, tiles=tl)
    if not err["field_err_max"] <= tol:
        raise ValueError(f"the far-rule gather misses the field tolerance {tol:g}: {err['field_err_max']:.3e}")
    rep = dict(tol=tol, far_rule=list(far_rule), near_width=near_width, gather=err,
               clusters=None if cl is None else dict(n=cl.n_clusters, kappa=cl.kappa, tol=cl.tol,
                                                     n_terms=cl.n_terms, seg_per_cluster=seg_per_cluster),
               tiles=None if tl is None else tl.report,
               multipole=None if mp is None else {k: v for k, v in mp.report.items()
                                                 if k not in ("candidates", "worst_beyond")})
    return FieldBudget(geometry, tuple(far_rule), near_width, cl, mp, rep, tiles=tl)
