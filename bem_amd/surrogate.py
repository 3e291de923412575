"""Learned field surrogate for axisymmetric BEM tip geometries (no
counterpart in the reference, which evaluates the BEM field directly; the
plain field gather is the reference the surrogate's error is measured against).

A Fourier-feature MLP maps (rho, z) to the BEM perturbation (phi, E_rho,
E_z). Inputs are polar coordinates about the cap centre (log r, cos
theta), outputs are scaled by a fitted power-law envelope so every target
is O(1). Training samples lie above the cap centre plane, their distance
from the surface log-uniform; labels come from the device field gather
(spt_bem_pert_field) and potential (spt_bem_potential). Training runs on
the device: BF16 MFMA products with FP32 accumulation and FP32 master
weights under Adam, first with one BF16 term per operand, then with the
three-term BF16 split so the trained weights match the ~FP32 inference.

After training, the surrogate is validated against the BEM on fresh points,
and the fallback band (distance from the surface inside which the tracer
keeps the BEM) is the smallest candidate >= `min_band_panels` panel
lengths beyond which the worst field error, relative to the magnitude of
the full external field, stays within `tol`. The tracer uses the
surrogate only inside the sampled cone (polar angle <= theta_max about
the cap centre) and box, outside that band, blending it in between one
and two band widths.
"""
import ctypes
import math

import numpy as np

from . import _lib

LOSS_WEIGHTS = (0.25, 1.0, 1.0)  # phi, E_rho, E_z


def _nparams():
    lib = _lib.load()
    n = ctypes.c_int64(0)
    if lib.spt_mlp_nparams(ctypes.byref(n)) != _lib.SPT_OK:
        raise RuntimeError("spt_mlp_nparams failed")
    return n.value


def cap_samples(R, H, n, d_min, d_max, rng, theta_max=0.5 * np.pi):
    """Points (rho, z) above the cap centre (0, H): polar angle U[0,
    theta_max] about the centre, distance from the cap surface log-uniform
    in [d_min, d_max]; returns (rho, z, d)."""
    th = rng.uniform(0.0, theta_max, n)
    d = np.exp(rng.uniform(np.log(d_min), np.log(d_max), n))
    rr = R + d
    return rr * np.sin(th), H + rr * np.cos(th), d


def _layout():
    F, IN, W, D, OUTP = _lib.MLP_FOURIER, 2 * _lib.MLP_FOURIER, _lib.MLP_WIDTH, _lib.MLP_DEPTH, 32
    off = {"W1": 0, "b1": W * IN}
    o = W * IN + W
    for l in range(2, D + 1):
        off[f"W{l}"] = o
        off[f"b{l}"] = o + W * W
        o += W * W + W
    off["Wo"] = o
    off["bo"] = o + OUTP * W
    off["n"] = o + OUTP * W + OUTP
    return off, IN, W, D


def initial_params(seed=0):
    """FP32 parameters: weights U(-sqrt(6/fan_in), +) scaled for SiLU,
    output layer U(-sqrt(1/fan_in), +) on its 3 live rows, biases 0."""
    rng = np.random.default_rng(seed)
    off, IN, W, D = _layout()
    p = np.zeros(off["n"], dtype=np.float32)
    lim = math.sqrt(6.0 / IN)
    p[off["W1"]:off["W1"] + W * IN] = rng.uniform(-lim, lim, W * IN)
    for l in range(2, D + 1):
        lim = math.sqrt(6.0 / W)
        p[off[f"W{l}"]:off[f"W{l}"] + W * W] = rng.uniform(-lim, lim, W * W)
    lim = math.sqrt(1.0 / W)
    p[off["Wo"]:off["Wo"] + 3 * W] = rng.uniform(-lim, lim, 3 * W)
    assert off["n"] == _nparams()
    return p


class FieldSurrogate:
    """A trained surrogate for one BEM geometry: its spt_mlp_cfg (as a dict
    of plain values), FP32 parameters (a device tensor) and the validation's
    report."""

    def __init__(self, geometry, cfg, params, report=None):
        self.geometry = geometry
        self.cfg = dict(cfg)
        self.params = params
        self.report = report or {}

    @property
    def inert(self):
        """True when validation found no band where the tolerance holds: the
        network is then never used, and a trace skips its eligibility pass."""
        return not np.isfinite(self.cfg["band"])

    # -- the C struct ------------------------------------------------------
    def cfg_struct(self, band=None):
        c = _lib.MlpCfg_t()
        for i, v in enumerate(np.asarray(self.cfg["fourier"], dtype=np.float32).ravel()):
            c.fourier[i] = float(v)
        c.center_z = self.cfg["center_z"]
        c.r_unit = self.cfg["r_unit"]
        for i in range(2):
            c.u_offset[i] = self.cfg["u_offset"][i]
            c.u_scale[i] = self.cfg["u_scale"][i]
        for i in range(3):
            c.out_scale[i] = self.cfg["out_scale"][i]
        c.env_r0 = self.cfg["env_r0"]
        c.env_pow = self.cfg["env_pow"]
        c.rho_max = self.cfg["rho_max"]
        c.z_min = self.cfg["z_min"]
        c.z_max = self.cfg["z_max"]
        c.cos_min = self.cfg.get("cos_min", -1.0)
        c.band = self.cfg["band"] if band is None else band
        return c

    def install(self, ctx):
        """Make this the surrogate spt_trace uses on `ctx` (copied)."""
        c = self.cfg_struct()
        ctx.check(ctx.lib.spt_surrogate_set(ctx.h, ctypes.byref(c), _lib.ptr(self.params)), "spt_surrogate_set")

    # -- evaluation --------------------------------------------------------
    def predict(self, pos, precision="fp64"):
        """(n, 3) (phi, E_rho, E_z) at device or host points (n, 3); precision
        "fp64" (FP64 MFMA, as the FP64 trace uses it) or "mixed" (three-term
        BF16 split, ~FP32)."""
        from .geometry import _precision_code

        import torch

        ctx = self.geometry._ctx()
        is_np = not hasattr(pos, "detach")
        P = torch.as_tensor(np.asarray(pos) if is_np else pos, dtype=torch.float64,
                            device=ctx.torch_device).reshape(-1, 3).contiguous()
        out = torch.empty_like(P)
        c = self.cfg_struct()
        ctx.bind_stream()
        ctx.check(ctx.lib.spt_mlp_infer(ctx.h, ctypes.byref(c), _lib.ptr(self.params), _lib.ptr(P), P.shape[0],
                                        _precision_code(precision), _lib.ptr(out)), "spt_mlp_infer")
        ctx.sync()
        return out.cpu().numpy() if is_np else out


def bem_labels(geometry, rho, z):
    """(n, 3) BEM perturbation (phi, E_rho, E_z) at meridian points, from
    the device field gather and potential."""
    import torch

    ctx = geometry._ctx()
    pts = torch.as_tensor(np.stack([rho, np.zeros_like(rho), z], -1), dtype=torch.float64,
                          device=ctx.torch_device).contiguous()
    E = torch.empty_like(pts)
    phi = torch.empty(pts.shape[0], dtype=torch.float64, device=ctx.torch_device)
    ctx.bind_stream()
    ctx.check(ctx.lib.spt_bem_pert_field(ctx.h, _lib.ptr(pts), pts.shape[0], _lib.ptr(E), _lib.PREC_FP64),
              "spt_bem_pert_field")
    ctx.check(ctx.lib.spt_bem_potential(ctx.h, _lib.ptr(pts), pts.shape[0], _lib.ptr(phi)), "spt_bem_potential")
    ctx.sync()
    return torch.stack([phi, E[:, 0], E[:, 2]], -1)


def make_cfg(geometry, rho, z, y, d_min, d_max, sigma_B=1.5, seed=0, env_r0=None, theta_max=0.5 * np.pi):
    """Input normalisation, output envelope and scales from the samples."""
    R, H = geometry.R, geometry.H
    rng = np.random.default_rng(seed + 1)
    r = np.hypot(rho, z - H)
    lo, hi = math.log((R + d_min) / R), math.log((R + d_max + 1e-12) / R)
    env_r0 = R if env_r0 is None else env_r0
    x = np.log1p(r / env_r0)
    mag = np.hypot(y[:, 1], y[:, 2])
    ok = mag > 0
    slope = np.polyfit(x[ok], np.log(mag[ok]), 1)[0]
    env_pow = float(np.clip(-slope, 0.0, 6.0))
    env = (1.0 + r / env_r0) ** (-env_pow)
    scale = np.sqrt(np.mean((y / env[:, None]) ** 2, axis=0))
    return dict(fourier=rng.normal(0.0, sigma_B, (32, 2)).astype(np.float32), center_z=H, r_unit=R,
                u_offset=(0.5 * (lo + hi), 0.0), u_scale=(2.0 / (hi - lo), 1.0), out_scale=tuple(float(s) for s in scale),
                env_r0=float(env_r0), env_pow=env_pow, rho_max=float(np.max(rho)), z_min=float(H),
                z_max=float(H + R + d_max), cos_min=float(np.cos(theta_max)), band=float("inf"))


def _cosine_lr(i, n, lr0, lr1):
    return lr1 + 0.5 * (lr0 - lr1) * (1.0 + math.cos(math.pi * min(i, n) / n))


def train(geometry, n_samples=1 << 20, steps=20000, batch=4096, split_fraction=0.25, lr=2e-3, lr_final=1e-5,
          d_max=None, tol=1e-6, min_band_panels=3.0, n_val=1 << 16, seed=0, chunk=500, log=None,
          theta_max_deg=80.0, sigma_B=1.5):
    """Generate samples, train on the device and validate against the BEM.
    Returns a FieldSurrogate whose report holds losses, validation errors
    and the chosen band."""
    import time

    import torch

    R, H = geometry.R, geometry.H
    prof = geometry.solution.profile
    panel = float(np.max(np.linalg.norm(np.diff(prof, axis=0), axis=1)))
    d_min = min_band_panels * panel
    d_max = d_max if d_max is not None else 110e-6
    rng = np.random.default_rng(seed)
    ctx = geometry._ctx()
    dev = ctx.torch_device
    t0 = time.perf_counter()
    theta_max = np.deg2rad(theta_max_deg)
    rho, z, d = cap_samples(R, H, n_samples, d_min, d_max, rng, theta_max)
    y = bem_labels(geometry, rho, z)
    t_labels = time.perf_counter() - t0
    cfg = make_cfg(geometry, rho, z, y.cpu().numpy(), d_min, d_max, seed=seed, theta_max=theta_max, sigma_B=sigma_B)
    sur = FieldSurrogate(geometry, cfg, torch.as_tensor(initial_params(seed), device=dev).contiguous())
    x = torch.as_tensor(np.stack([rho, z], -1), dtype=torch.float64, device=dev).contiguous()
    m = torch.zeros_like(sur.params)
    v = torch.zeros_like(sur.params)
    c = sur.cfg_struct()
    w = (ctypes.c_float * 3)(*LOSS_WEIGHTS)
    losses = []
    n_split = int(round(split_fraction * steps))
    t1 = time.perf_counter()
    done = 0
    ctx.bind_stream()
    while done < steps:
        k = min(chunk, steps - done)
        use_split = done >= steps - n_split
        lr_k = _cosine_lr(done, steps, lr, lr_final)
        buf = np.zeros(k, dtype=np.float32)
        ctx.check(ctx.lib.spt_mlp_train(ctx.h, ctypes.byref(c), _lib.ptr(sur.params), _lib.ptr(m), _lib.ptr(v), done,
                                        _lib.ptr(x), _lib.ptr(y), n_samples, batch, k, float(lr_k),
                                        (seed + 1) * 0x9E3779B1, w, 1 if use_split else 0,
                                        buf.ctypes.data_as(ctypes.c_void_p)), "spt_mlp_train")
        losses.append(buf)
        done += k
        if log:
            log(f"surrogate train: step {done}/{steps} lr {lr_k:.2e} {'split3' if use_split else 'bf16'} "
                f"loss {float(np.mean(buf[-50:])):.3e}")
    torch.cuda.synchronize(dev)
    t_train = time.perf_counter() - t1
    losses = np.concatenate(losses)
    rep = validate(sur, n_val, d_min, d_max, tol, panel, seed + 7, theta_max)
    rep.update(t_labels_s=t_labels, t_train_s=t_train, steps=steps, batch=batch, n_samples=n_samples,
               final_loss=float(np.mean(losses[-200:])), panel_length=panel)
    sur.cfg["band"] = rep["band"]
    sur.report = rep
    sur.losses = losses
    return sur


def validate(sur, n_val, d_min, d_max, tol, panel, seed, theta_max=0.5 * np.pi):
    """Surrogate vs BEM on fresh points: errors relative to the full
    external field magnitude |E_gun e_z + E_pert|, worst case per distance
    band, and the smallest band (>= d_min) whose worst error is <= tol."""
    geometry = sur.geometry
    rng = np.random.default_rng(seed)
    rho, z, d = cap_samples(geometry.R, geometry.H, n_val, d_min, d_max, rng, theta_max)
    ref = bem_labels(geometry, rho, z).cpu().numpy()
    got = sur.predict(np.stack([rho, np.zeros_like(rho), z], -1))
    Et = np.hypot(ref[:, 1], ref[:, 2] + geometry.E_gun)
    errE = np.hypot(got[:, 1] - ref[:, 1], got[:, 2] - ref[:, 2]) / Et
    phi_scale = np.abs(geometry.E_gun) * np.maximum(np.hypot(rho, z), geometry.R)
    errP = np.abs(got[:, 0] - ref[:, 0]) / phi_scale
    cands = d_min * np.geomspace(1.0, d_max / d_min, 60)
    # only candidates with validation points beyond them can be chosen
    cands = cands[np.array([np.any(d >= b) for b in cands])]
    worst = np.array([errE[d >= b].max() for b in cands])
    band = band_for(cands, worst, tol)
    edges = np.geomspace(d_min, d_max, 9)
    by_d = []
    for a, b in zip(edges[:-1], edges[1:]):
        sel = (d >= a) & (d < b)
        if sel.any():
            by_d.append(dict(d_lo=float(a), d_hi=float(b), max=float(errE[sel].max()), median=float(np.median(errE[sel]))))
    return dict(tol=tol, band=band, band_panels=band / panel if np.isfinite(band) else None, error_by_distance=by_d,
                field_err_max=float(errE.max()), field_err_median=float(np.median(errE)),
                field_err_p99=float(np.quantile(errE, 0.99)), phi_err_max=float(errP.max()),
                phi_err_median=float(np.median(errP)), band_candidates=cands.tolist(), worst_beyond=worst.tolist(),
                n_val=n_val)


def far_rule_error(geometry, k4=2.0, k2=16.0, n=1 << 16, d_min=None, d_max=110e-6, seed=11,
                   theta_max=np.deg2rad(100.0), near_width=None, clusters=None, tiles=None):
    """Field error of the distance-adaptive far rule against the fixed
    rule on fresh points around the cap and the post's side (distance from
    the cap sphere log-uniform in [d_min, d_max], d_min defaulting to 1e-3
    panel lengths; points inside the conductor, where the tracer uses no
    field, are dropped), relative to the full external field magnitude as
    in validate(). near_width, clusters and tiles add the wider near rule,
    the cluster expansions and the tiles (geometry.far_rule). Returns max, p99 and median, and
    the distance from the conductor surface of the worst point."""
    prof = geometry.solution.profile
    panel = float(np.max(np.linalg.norm(np.diff(prof, axis=0), axis=1)))
    d_min = 1e-3 * panel if d_min is None else d_min
    rng = np.random.default_rng(seed)
    rho, z, _ = cap_samples(geometry.R, geometry.H, n, d_min, d_max, rng, theta_max)
    # distance from the capsule's axis segment z <= H (the post, or the
    # hemisphere for H = 0), minus R: distance from the conductor surface
    d = np.where(z > geometry.H, np.hypot(rho, z - geometry.H), rho) - geometry.R
    keep = (z > 0.0) & (d >= 0.0)
    rho, z, d = rho[keep], z[keep], d[keep]
    ref = bem_labels(geometry, rho, z).cpu().numpy()
    with geometry.far_rule(k4, k2, near_width=near_width, clusters=clusters, tiles=tiles):
        got = bem_labels(geometry, rho, z).cpu().numpy()
    Et = np.hypot(ref[:, 1], ref[:, 2] + geometry.E_gun)
    err = np.hypot(got[:, 1] - ref[:, 1], got[:, 2] - ref[:, 2]) / Et
    return dict(k4=k4, k2=k2, near_width=near_width, clusters=None if clusters is None else clusters.n_clusters,
                tiles=None if tiles is None else tiles.n_tiles,
                n=int(rho.shape[0]), d_min=d_min, d_max=d_max, field_err_max=float(err.max()),
                field_err_p99=float(np.quantile(err, 0.99)), field_err_median=float(np.median(err)),
                d_at_max=float(d[np.argmax(err)]))


def band_for(cands, worst, tol):
    """Smallest candidate band beyond which every larger candidate's worst
    error is within tol (inf if none)."""
    worst = np.asarray(worst)
    for i in range(len(cands)):
        if np.all(worst[i:] <= tol):
            return float(cands[i])
    return float("inf")


def with_tolerance(sur, tol):
    """The same network with the band re-chosen for another tolerance from
    the validation record."""
    rep = dict(sur.report)
    band = band_for(rep["band_candidates"], rep["worst_beyond"], tol)
    rep.update(tol=tol, band=band, band_panels=band / rep["panel_length"] if np.isfinite(band) else None)
    out = FieldSurrogate(sur.geometry, dict(sur.cfg, band=band), sur.params, rep)
    out.losses = getattr(sur, "losses", None)
    return out
