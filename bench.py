"""Benchmark: particle-RK-stages/s of the MI355X tracing path, surrogate +
batched against plain, at a surrogate field error <= 1e-6.

Workload (one "step" = one full trace, every group from birth to t_max):
hemisphere-capped post, R = 1 um on a 10 um post, 400-panel mirror-doubled
axisymmetric BEM (12,800 field quadrature nodes), uniform -1e8 V/m (anode
+10 kV at 100 um; particles absorbed there), 262,144 electrons emitted from
the cap in 16,384 groups of 16 (polar angle U[0, 60 deg], KE ~ Exp(0.2 eV),
cos-weighted, numpy default_rng(0)), RK45 rtol 1e-8, atol 1e-12, screen at
90 um; FP64 field and step control, FP32 + Kahan pair forces; lockstep
schedule (every running group one attempt per graph-replayed iteration).

The surrogate + batched form spends the 1e-6 field budget on three
approximations of the BEM field, each validated against the device's
fixed-rule BEM on fresh points before the timed region: the exterior
multipole expansion of the BEM charge beyond its validated radius (1e-8),
the distance-adaptive far rule (Gauss-8/4/2 per segment) in every remaining
gather, and the learned surrogate (32 Fourier features, 4 x 128 SiLU,
outputs phi, E_rho, E_z; trained on the device on 2^20 BEM samples) beyond
the distance band where its worst field error relative to |E| is <= 1e-6
(none at 1e-6: inert). The plain method (the reference's rule everywhere)
is timed on the same workload; so are the same fast trace without the
network and the network's bands for looser tolerances (1e-4, 1e-3).

value = 7 x sum_g(size_g x attempts_g) / wall time (all particle slots of
running groups, the reference's accel_fn shape). The roofline entry is for
the dominant kernel of the value trace, the far-rule BEM field gather:
algorithmic FP64 flops = points x ring nodes the far rule sums (counted on
the host for the same points) x 120 flops per node, over its average
launch time from HIP events on the launch stream, on 262,144 points
10 nm .. 10 um off the cap; roofline_plain_gather gives the same for the
fixed rule (12,800 nodes per point). cpu_baseline: the oracle
(NumPy FP64 restatement of the reference) on a time-bounded sample of the
same workload, one forked process per host core (up to 16), run before the
GPU is touched.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-RK-stages/sec, 1 GPU, surrogate+batched vs plain at field err ≤1e-6"
PEAK_FP64_TFLOPS = 78.6
C_RING = 120          # FP64 flops per ring-kernel node (K, E, log, kernel algebra, accumulation)
C_TERM = 16    # FP64 flops per term of an axis-centred cluster expansion or of the multipole zone
C_TABLE = 2 * 2 * 12 * 12  # FP64 flops per tile table: two 12 x 12 Chebyshev sums, one FMA per coefficient
# HBM bytes per 262,144-point field-gather launch from the rocprofv3 counter runs (FETCH_SIZE and WRITE_SIZE at
# the timed launch shape, scaled by the same counters' reading for a copy of known size)
TRAFFIC_BYTES = 19.8e6
TRAFFIC_SOURCE = ("FETCH_SIZE x 2.000 + WRITE_SIZE x 1.000 (scales from a 256 MiB copy), fixed rule at (256, 64), "
                  "profiles/bem_field_gather_post400_262k.md")
FAST_TRAFFIC_SHAPE = (128, 64)
FAST_TRAFFIC_BYTES = 14.0e6
FAST_TRAFFIC_SOURCE = ("FETCH_SIZE x 2.000 + WRITE_SIZE x 1.000 (scales from a 256 MiB copy), fast-form gather with "
                       "tiles at (128, 64) on the points in bin order, "
                       "profiles/bem_field_gather_fast_form_post400_262k.md")


def workload(n_el=262144, group=16, n_segments=400, seed=0, rtol=1e-8):
    from bem_amd.emission import cap_source

    R, H = 1e-6, 10e-6
    src = cap_source(n_el, R, H, theta_max_deg=60.0, mean_ke_ev=0.2, seed=seed)
    return dict(R=R, H=H, E_gun=-1e8, n_segments=n_segments, src=src, group=group, screens=[90e-6], z_max=100e-6,
                rtol=rtol, atol=1e-12, n_el=n_el)


def make_tracer(wl, geo, surrogate=None, precision="pairs32", mode="lockstep", far_rule=None, multipole=None,
                near_width=None, clusters=None, tiles=None):
    from bem_amd import SpecificParticleTracer
    from bem_amd.particles import ParticleGroup

    src = wl["src"]
    n = len(src["x"])
    pg = ParticleGroup(dict(src, weight=np.full(n, 1.602176634e-19), status=np.ones(n), species="electron"))
    return SpecificParticleTracer(pg, wl["group"], geo, wl["screens"], rtol=wl["rtol"], atol=wl["atol"],
                                  z_max=wl["z_max"], mode=mode, precision=precision, surrogate=surrogate,
                                  far_rule=far_rule, multipole=multipole, near_width=near_width, clusters=clusters,
                                  tiles=tiles)


def timed_traces(tr, steps, warmup):
    """K traces with the inputs resident in device memory (uploaded by the
    tracer once) and the results left there (trace_device), timed between
    two device synchronisations; the results come back to the host after
    the timed region."""
    import torch

    tr._device_inputs(tr.geometry._ctx())
    for _ in range(warmup):
        tr.trace_device()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    outs = [tr.trace_device() for _ in range(steps)]
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return [tr.to_host(o) for o in outs], wall


def host_io_trace(tr):
    """One trace through the host boundary (host arrays in, host arrays
    out: uploads and downloads over PCIe included), after one warm-up."""
    import torch

    tr.trace()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    o = tr.trace()
    torch.cuda.synchronize()
    return o, time.perf_counter() - t0


def screen_time_diff(a, b):
    ea, eb = a["events"], b["events"]
    if len(ea) != len(eb):
        return None
    ka = np.lexsort((ea["particle"], ea["group"]))
    kb = np.lexsort((eb["particle"], eb["group"]))
    dt = np.abs(ea["t"][ka] - eb["t"][kb]) / ea["t"][ka]
    return dict(max=float(dt.max()), median=float(np.median(dt)))


def gather_points(wl, n_pts=262144):
    """n_pts points displaced 10 nm .. 10 um off the cap (log-uniform)."""
    from bem_amd.emission import cap_source

    src = cap_source(n_pts, wl["R"], wl["H"], theta_max_deg=60.0, mean_ke_ev=0.2, seed=1)
    R, H = wl["R"], wl["H"]
    p = np.stack([src["x"], src["y"], src["z"] - H], -1)
    rc = np.linalg.norm(p, axis=1, keepdims=True)
    d = 10 ** np.random.default_rng(1).uniform(-8, -5, (n_pts, 1))
    pts = p / rc * (R + d)
    pts[:, 2] += H
    return pts


def point_classes(geo, pts, budget):
    """How the fast form's gather takes each point, as the library decides:
    the multipole zone (r >= r_min), a tile (its index; table-only tiles
    have an empty window) or neither (-1: the far-rule gather with cluster
    expansions). Returns (in_mp bool, tile int)."""
    rho, z = np.hypot(pts[:, 0], pts[:, 1]), pts[:, 2]
    mp = budget.multipole
    in_mp = (np.hypot(rho, z) >= mp.r_min) if mp is not None and not mp.inert else np.zeros(len(pts), dtype=bool)
    tile = np.full(len(pts), -1, dtype=np.int64)
    if budget.tiles is not None:
        tile, _, _, _ = budget.tiles.find(rho, z)
    tile = np.where(in_mp, -1, tile)
    return in_mp, tile


def gather_work_per_point(geo, pts, budget, sample=8192):
    """Mean algorithmic work per point of the fast form's gather over the
    first `sample` points, counted on the host from the rules the device
    applies to each point: in the multipole zone its expansion terms (the
    truncation rule of spt_bem_multipole); in a tile the table (two 12 x 12
    Chebyshev sums) plus its window's segments; elsewhere every segment
    outside far clusters plus the far clusters' expansion terms. A summed
    segment costs its far-rule ring nodes (8 at L <= D < k4 L, 4 below k2 L,
    2 beyond) or, closer than its length, its sinh-rule nodes (8 per
    interval of width near_width). Returns per-point means: ring nodes, near
    nodes, expansion terms (clusters and multipole), tables."""
    from bem_amd.bem import clusters as bcl

    k4, k2 = budget.far_rule
    prof = np.asarray(geo.solution.profile)
    p0, t = prof[:-1], np.diff(prof, axis=0)
    L = np.linalg.norm(t, axis=1)
    u = t / L[:, None]
    q = pts[:sample]
    in_mp, tile = point_classes(geo, q, budget)
    rho, z = np.hypot(q[:, 0], q[:, 1]), q[:, 2]
    dr, dz = rho[:, None] - p0[None, :, 0], z[:, None] - p0[None, :, 1]
    s0 = np.clip(dr * u[None, :, 0] + dz * u[None, :, 1], 0.0, L[None])
    D = np.hypot(dr - s0 * u[None, :, 0], dz - s0 * u[None, :, 1])
    Lr = L[None]
    summed = np.ones_like(D, dtype=bool)
    terms = np.zeros(len(q))
    cl = budget.clusters
    if cl is not None:
        fallback = (tile < 0) & ~in_mp
        for c in range(cl.n_clusters):
            r = np.hypot(rho, z - cl.zc[c])
            far = fallback & (r >= cl.kappa * cl.rc[c])
            summed[:, cl.seg_begin[c]:cl.seg_begin[c + 1]] &= ~far[:, None]
            qq = np.where(far, cl.rc[c] / np.maximum(r, 1e-300), 0.0)
            terms += np.where(far, bcl.terms_needed(qq, cl.tol, cl.n_terms), 0)
    if budget.tiles is not None:
        win = budget.tiles.tiles[np.maximum(tile, 0), 3:5].astype(np.int64)
        j = np.arange(len(L))[None, :]
        in_win = (j >= win[:, :1]) & (j < win[:, 1:])
        summed = np.where((tile >= 0)[:, None], in_win, summed)
    summed &= ~in_mp[:, None]
    mp = budget.multipole
    if mp is not None and not mp.inert:
        qq = np.where(in_mp, mp.r_ref / np.maximum(np.hypot(rho, z), 1e-300), 0.0)
        terms += np.where(in_mp, bcl.terms_needed(qq, 1e-12, len(mp.m)), 0)
    ring = np.where(D < Lr, 0, np.where(D < k4 * Lr, 8, np.where(D < k2 * Lr, 4, 2)))
    d = np.maximum(D, 1e-15 * Lr)
    span = np.arcsinh((Lr - s0) / d) - np.arcsinh(-s0 / d)
    near = np.where(D < Lr, 8 * np.maximum(1, np.ceil(span / budget.near_width)), 0)
    return (float(np.where(summed, ring, 0).sum(1).mean()), float(near.sum(1).mean()),
            float(terms.mean()), float((tile >= 0).mean()))


def bin_order(geo, pts, budget=None):
    """The order in which the binned lockstep gather takes points, by the
    library's bin key: points outside tiles and the multipole zone by
    distance band (floor log2 of the distance to the closest segment over
    its length, from -12 to 7) then closest segment (first index on ties),
    then tiles by index (windowed tiles first), then the multipole zone by
    r; ties keep the point order."""
    prof = np.asarray(geo.solution.profile)
    p0, t = prof[:-1], np.diff(prof, axis=0)
    L = np.linalg.norm(t, axis=1)
    u = t / L[:, None]
    rho, z = np.hypot(pts[:, 0], pts[:, 1]), pts[:, 2]
    key = np.empty(len(pts), dtype=np.int64)
    for a in range(0, len(pts), 4096):
        dr, dz = rho[a:a + 4096, None] - p0[None, :, 0], z[a:a + 4096, None] - p0[None, :, 1]
        s0 = np.clip(dr * u[None, :, 0] + dz * u[None, :, 1], 0.0, L[None])
        D = np.hypot(dr - s0 * u[None, :, 0], dz - s0 * u[None, :, 1])
        j = D.argmin(1)
        d = D[np.arange(len(j)), j] / L[j]
        band = np.clip(np.floor(np.log2(np.maximum(d, 1e-300))) + 12, 0, 19).astype(np.int64)
        key[a:a + 4096] = band * len(L) + j
    if budget is not None:
        in_mp, tile = point_classes(geo, pts, budget)
        base = 20 * len(L)
        key = np.where(tile >= 0, base + tile, key)
        n_t = 0 if budget.tiles is None else budget.tiles.n_tiles
        key = np.where(in_mp, base + n_t + np.floor(16 * np.log2(np.maximum(np.hypot(rho, z), 1e-300))).astype(
            np.int64) + 10000, key)
    return np.argsort(key, kind="stable")


def field_gather_timing(wl, geo, n_pts=262144, reps=10, shape=None, budget=None, binned=False):
    """Average launch duration of the BEM field gather over gather_points(),
    timed with HIP events on the stream the kernel is launched on (the
    library is bound to torch's current stream); with a FieldBudget, the
    fast form's gather (far rule, near width, clusters, tiles and the
    multipole zone: table-only and multipole-zone points one per thread,
    the rest 64 per block as the binned gather takes them); binned: the
    points in bin order, as the binned lockstep gather takes them."""
    import contextlib

    pts = gather_points(wl, n_pts)
    if binned:
        pts = np.ascontiguousarray(pts[bin_order(geo, pts, budget)])
    ctx = geo._ctx()

    @contextlib.contextmanager
    def fast_form():
        mp = budget.multipole
        with geo.far_rule(*budget.far_rule, near_width=budget.near_width, clusters=budget.clusters,
                          tiles=budget.tiles):
            if mp is not None and not mp.inert:
                with mp.active():
                    yield
            else:
                yield

    with (fast_form() if budget else contextlib.nullcontext()):
        return _gather_timing(ctx, pts, n_pts, reps, shape)


def _gather_timing(ctx, pts, n_pts, reps, shape):
    import torch

    from bem_amd import _lib

    pt = torch.as_tensor(pts, device=ctx.torch_device).contiguous()
    out = torch.empty_like(pt)
    ctx.bind_stream()
    shape = shape or (0, 0)
    ctx.check(ctx.lib.spt_set_field_shape(ctx.h, int(shape[0]), int(shape[1])), "spt_set_field_shape")
    for _ in range(2):
        ctx.check(ctx.lib.spt_bem_pert_field(ctx.h, _lib.ptr(pt), n_pts, _lib.ptr(out), _lib.PREC_FP64), "field")
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        ctx.check(ctx.lib.spt_bem_pert_field(ctx.h, _lib.ptr(pt), n_pts, _lib.ptr(out), _lib.PREC_FP64), "field")
    e1.record()
    torch.cuda.synchronize()
    ctx.lib.spt_set_field_shape(ctx.h, 0, 0)
    return e0.elapsed_time(e1) / reps, n_pts


_CPU = {}  # oracle state shared with the forked baseline workers


def _cpu_group_rate(gi):
    """Oracle trace of group gi of the workload for _CPU["seconds"] of work:
    (particle-stages, elapsed s)."""
    from threadpoolctl import threadpool_limits

    from oracle import rk45
    from oracle import tracker as otr

    threadpool_limits(1)  # one core per process: no BLAS threads on top
    wl, geom_base, seconds = _CPU["wl"], _CPU["geom"], _CPU["seconds"]
    src, g = wl["src"], wl["group"]
    sl = slice(gi * g, (gi + 1) * g)
    st = otr.setup(*(src[k][sl] for k in ("x", "y", "z", "px", "py", "pz", "t", "id")), g)
    accel = otr.make_accel(geom_base, st["charge"], st["mass_p"], st["q_over_m"], 1e-12)
    count = [0]
    t0 = [None]

    class Done(Exception):
        pass

    def timed(pos, vel, active, gidx):
        if t0[0] is None:
            t0[0] = time.perf_counter()
        if time.perf_counter() - t0[0] > seconds:
            raise Done
        count[0] += pos.shape[0] * pos.shape[1]
        return accel(pos, vel, active, gidx)

    t_max = otr.default_t_max(st, wl["E_gun"], wl["screens"])
    try:
        rk45.integrate(timed, st["pos"], st["vel"], st["t_birth"], t_max, wl["rtol"], wl["atol"],
                       kill=lambda p: otr.kill_mask(geom_base, p))
    except Done:
        pass
    return count[0], time.perf_counter() - t0[0]


def host_cpu():
    """CPU model, sockets, cores per socket and threads per core (lscpu)."""
    import subprocess

    info = {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in ("Model name", "Socket(s)", "Core(s) per socket", "Thread(s) per core", "CPU(s)"):
                info[k.strip()] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return info


def cpu_baseline(wl, seconds=15.0, n_groups=64, workers=None):
    """The oracle on the first n_groups groups of the workload, spread over
    `workers` forked processes (the reference's own process-level group
    partitioning; default: every host core this process may use, at most
    16, one GPU's share of a shared host), each group traced for seconds * workers /
    n_groups of work, so the whole sample takes ~`seconds`. Rate = all
    particle-stages / wall time of the pool. Must run before anything
    touches the GPU (the workers are forked)."""
    import multiprocessing as mp

    from oracle import bem as ob
    from bem_amd.bem.mesh import post_profile

    R, H = wl["R"], wl["H"]
    prof = post_profile(R, H, wl["n_segments"])
    sigma = ob.solve(prof, lambda r, z: wl["E_gun"] * z)
    avail = len(os.sched_getaffinity(0))
    if workers is None:
        workers = max(1, min(16, avail))
    _CPU.update(wl=wl, seconds=seconds * workers / n_groups,
                geom=dict(kind="post_bem", E_gun=wl["E_gun"], R=R, H=H, profile=prof, sigma=sigma,
                          z_handoff=20 * (R + H), kill_z_below=0.0, z_max=wl["z_max"]))
    t0 = time.perf_counter()
    with mp.get_context("fork").Pool(workers) as pool:
        parts = pool.map(_cpu_group_rate, range(n_groups), chunksize=1)
    wall = time.perf_counter() - t0
    count = sum(c for c, _ in parts)
    cpu = host_cpu()
    return dict(value=count / wall, unit="particle-RK-stages/s", cores=workers, kind="port",
                cpu_model=cpu.get("Model name"), sockets=cpu.get("Socket(s)"),
                cores_per_socket=cpu.get("Core(s) per socket"), threads_per_core=cpu.get("Thread(s) per core"),
                host_cpus_online=cpu.get("CPU(s)"), cpus_available_to_process=avail,
                cap=f"{workers} worker processes: every CPU this process may use, capped at 16 (one GPU's "
                    "share of a shared host)",
                sample=f"oracle (NumPy FP64, one BLAS thread per process) on groups 0..{n_groups - 1} "
                       f"({wl['group']} electrons each) of the workload, each traced from birth for "
                       f"{_CPU['seconds']:.2f} s of work (the first, near-surface attempts); {wall:.1f} s wall, "
                       f"{count} particle-stages")


def surrogate_cost(wl, geo, sur, bud, n=262144):
    """Per-point device time in the network's own domain (3 panel lengths to
    1 um off the cap): the fast form's gather against the network's FP64
    inference (HIP events around 5 launches each, after one warm-up)."""
    import torch

    from bem_amd import surrogate

    ctx = geo._ctx()
    panel = float(np.max(np.linalg.norm(np.diff(geo.solution.profile, axis=0), axis=1)))
    rho, z, _ = surrogate.cap_samples(geo.R, geo.H, n, 3 * panel, 1e-6, np.random.default_rng(3), np.deg2rad(80))
    pts = np.stack([rho, np.zeros_like(rho), z], -1)
    with geo.far_rule(*bud.far_rule, near_width=bud.near_width, clusters=bud.clusters):
        fast_ms, _ = _gather_timing(ctx, pts, n, 5, (192, 64))
    P = torch.as_tensor(pts, device=ctx.torch_device).contiguous()
    sur.predict(P)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(5):
        sur.predict(P)
    e1.record()
    torch.cuda.synchronize()
    net_ms = e0.elapsed_time(e1) / 5
    return dict(points=n, domain="3 panel lengths .. 1 um off the cap, polar angle <= 80 deg",
                fast_gather_ns_per_point=fast_ms * 1e6 / n, network_fp64_ns_per_point=net_ms * 1e6 / n)


def c2_plain_rate(steps=3):
    """Plain FP64 method on 1024 electrons (64 groups of 16, rtol 1e-9), the
    configuration the parity tests pin against the oracle."""
    from bem_amd.bem.geometry import PostTipBEMGeometry

    wl = workload(n_el=1024, rtol=1e-9)
    geo = PostTipBEMGeometry(wl["E_gun"], wl["R"], wl["H"], n_segments=wl["n_segments"])
    tr = make_tracer(wl, geo, precision="fp64", mode="lockstep")
    outs, wall = timed_traces(tr, steps, 1)
    return dict(value=sum(o["particle_stages"] for o in outs) / wall, unit="particle-RK-stages/s",
                workload="1024 e- in 64 groups of 16, rtol 1e-9, FP64 field and pairs, lockstep",
                attempts_mean=float(outs[-1]["attempts"].mean()), ms_per_trace=wall / steps * 1e3)


def c3_sample_rate(attempts=8):
    """C3's shape on a capped number of attempts: cylindrical well R 50 nm,
    H 30 nm on a 2000-panel real profile, image-charge BEM (z0 3 nm, 16
    modes), 4096 groups of U{1..256} electrons (530,287) born at per-group
    offsets U[0, 10 ps], rtol 1e-8, lockstep, FP64; every group makes
    `attempts` attempts (a full run is ~100x longer)."""
    import torch

    from bem_amd import SpecificParticleTracer
    from bem_amd.bem.geometry import CylindricalWellBEMGeometry
    from bem_amd.emission import well_bottom_source
    from bem_amd.particles import ParticleGroup

    R, H = 50e-9, 30e-9
    geo = CylindricalWellBEMGeometry(-1e8, R, H, z0=3e-9, field_max_length=0.04223e-9)
    sizes = np.random.default_rng(0).integers(1, 257, 4096)
    src = well_bottom_source(sizes, R, H, t_offset_max=10e-12, seed=0)
    n = len(src["x"])
    pg = ParticleGroup(dict(src, weight=np.full(n, 1.602176634e-19), status=np.ones(n), species="electron"))
    tr = SpecificParticleTracer(pg, 1, geo, [20e-9, 100e-9], rtol=1e-8, atol=1e-12, group_sizes=sizes, z_max=1e-6,
                                max_steps=attempts)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    out = tr.trace()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return dict(value=out["particle_stages"] / wall, unit="particle-RK-stages/s",
                workload=f"well R 50 nm H 30 nm, {len(geo.solution.profile) - 1} panels, image BEM z0 3 nm on "
                         f"{len(geo.image_profile)} nodes x 17 modes, 4096 groups U{{1..256}} ({n} e-), rtol 1e-8, "
                         f"first {attempts} attempts of every group",
                wall_s=wall)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--electrons", type=int, default=262144)
    ap.add_argument("--tol", type=float, default=1e-6, help="field-error tolerance of the headline value")
    ap.add_argument("--train-steps", type=int, default=60000)
    ap.add_argument("--no-sweep", action="store_true", help="skip the looser-tolerance runs")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-c3", action="store_true", help="skip the C3-shaped sample (well with image charges)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--field-only", action="store_true",
                    help="only the field-gather timing and a 256 MiB device copy (for hardware-counter runs)")
    ap.add_argument("--plain-gather", action="store_true",
                    help="with --field-only: the fixed-rule gather instead of the fast form's")
    ap.add_argument("--field-shape", default="256,64",
                    help="threads,points-per-block of the field gather for --field-only and --value-only")
    ap.add_argument("--value-only", action="store_true",
                    help="only the headline trace (the field-error budget; the network is inert at 1e-6) after "
                         "one warm-up trace, for profiler runs")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the tracing path is benchmarked per GPU (groups shard across GPUs "
              "with no exchange during integration); --gpus above 1 is not supported.", file=sys.stderr)
        sys.exit(2)
    from bem_amd import surrogate
    from bem_amd.bem import budget as bbudget
    from bem_amd.bem.geometry import PostTipBEMGeometry

    def say(*a):
        print(*a, file=sys.stderr, flush=True)

    wl = workload(n_el=args.electrons)
    shape_arg = tuple(int(v) for v in args.field_shape.split(","))
    cpu = None
    if not args.no_cpu_baseline and not args.field_only and not args.value_only:
        say(f"bench: CPU baseline (oracle, 64 groups, ~{args.cpu_seconds:.0f} s)")
        cpu = cpu_baseline(wl, args.cpu_seconds)  # before the GPU is touched: the workers are forked
        say(f"bench: CPU baseline {cpu['value']:.0f} particle-stages/s on {cpu['cores']} cores")
    t0 = time.perf_counter()
    geo = PostTipBEMGeometry(wl["E_gun"], wl["R"], wl["H"], n_segments=wl["n_segments"])
    t_solve = time.perf_counter() - t0
    if args.field_only:
        import torch

        bud = None if args.plain_gather else bbudget.fit(geo, tol=args.tol, n_val=1 << 12)
        field_ms, n_pts = field_gather_timing(wl, geo, reps=3, shape=shape_arg, budget=bud, binned=bud is not None)
        src = torch.empty(64 << 20, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        work = None
        if bud is not None:
            ring, near, terms, tiled = gather_work_per_point(geo, gather_points(wl), bud)
            work = dict(ring_nodes=ring, near_nodes=near, expansion_terms=terms, tables=tiled,
                        flops=n_pts * ((ring + near) * C_RING + terms * C_TERM + tiled * C_TABLE))
        print(json.dumps(dict(field_ms=field_ms, points=n_pts, field_shape=shape_arg,
                              gather="fixed rule" if bud is None else "fast form", work_per_point=work,
                              copy_bytes_read=src.numel() * 4, copy_bytes_written=dst.numel() * 4)))
        return
    if args.value_only:
        bud = bbudget.fit(geo, tol=args.tol)
        vt = make_tracer(wl, geo, **bud.tracer_kwargs())
        vt.field_shape = shape_arg
        o, w = timed_traces(vt, args.steps, args.warmup)
        print(json.dumps(dict(value=sum(x["particle_stages"] for x in o) / w, wall_s=w, steps=args.steps,
                              field_shape=shape_arg, budget=bud.report)))
        return
    say(f"bench: BEM solve {t_solve:.2f} s; training the surrogate ({args.train_steps} steps)")
    sur = surrogate.train(geo, n_samples=1 << 20, steps=args.train_steps, batch=8192, tol=args.tol, chunk=10000)
    rep = sur.report
    say(f"bench: trained in {rep['t_train_s']:.1f} s; band at tol {args.tol:g}: {rep['band']:.3e} m")
    from bem_amd import autotune

    plain_tr = make_tracer(wl, geo)
    tuning = autotune.tune(plain_tr, attempts=8)
    say(f"bench: autotuned plain field shape {tuning['best']}")
    bud = bbudget.fit(geo, tol=args.tol)
    say(f"bench: field budget validated: gather err max {bud.report['gather']['field_err_max']:.2e}, multipole "
        f"r_min {bud.multipole.r_min:.3e} m")
    tr = make_tracer(wl, geo, surrogate=sur, **bud.tracer_kwargs())
    tuning_value = autotune.tune(tr, attempts=None)  # whole traces: the point mix changes along the trace
    say(f"bench: autotuned field shape of the fast trace {tuning_value['best']}")
    outs, wall = timed_traces(tr, args.steps, args.warmup)
    say(f"bench: {args.steps} fast traces in {wall:.2f} s")
    stages = sum(o["particle_stages"] for o in outs)
    act = sum(int(o["active_stages"].sum()) for o in outs)
    plain_outs, plain_wall = timed_traces(plain_tr, 1, 0)
    plain_value = plain_outs[0]["particle_stages"] / plain_wall
    say(f"bench: plain trace {plain_wall:.2f} s")
    # the same fast trace through the host boundary (ParticleGroup arrays in,
    # NumPy results out): the PCIe-inclusive rate, reported beside value
    oh, wh = host_io_trace(tr)
    host_io = dict(value=oh["particle_stages"] / wh, wall_s=wh)
    say(f"bench: fast trace with host arrays in and out {wh:.3f} s")
    # the value counts only if the fast trace reproduces the plain one: every
    # electron's screen crossing, at times within the tolerance
    n_el = wl["n_el"]
    st_diff = screen_time_diff(plain_outs[0], outs[-1])
    value_check = dict(events=int(len(outs[-1]["events"])), plain_events=int(len(plain_outs[0]["events"])),
                       screen_t_rel_diff=st_diff, bound=args.tol,
                       groups_done=bool(np.all(outs[-1]["status"] == 0)))
    value_check["pass"] = bool(st_diff is not None and value_check["events"] == value_check["plain_events"] == n_el
                               and st_diff["max"] <= args.tol and value_check["groups_done"])
    if not value_check["pass"]:
        say(f"bench: WARNING the fast trace does not reproduce the plain one: {value_check}")
    t3 = make_tracer(wl, geo, **bud.tracer_kwargs())  # the same fast trace without the network
    t3.field_shape = tr.field_shape
    o3, w3 = timed_traces(t3, 1, 0)
    net_off = dict(value=o3[0]["particle_stages"] / w3, wall_s=w3, attempts_mean=float(o3[0]["attempts"].mean()),
                   screen_t_rel_diff_vs_plain=screen_time_diff(plain_outs[0], o3[0]))
    # a trace that runs the network gathers in particle order (the binned
    # gather is for traces without it): the network's gain is measured
    # against the fast form gathered the same way
    t3u = make_tracer(wl, geo, **bud.tracer_kwargs())
    t3u.field_shape = (tr.field_shape[0], 64) if tr.field_shape else None
    t3u.binned = False
    o3u, w3u = timed_traces(t3u, 1, 0)
    net_off["unbinned"] = dict(value=o3u[0]["particle_stages"] / w3u, wall_s=w3u,
                               screen_t_rel_diff_vs_binned=screen_time_diff(o3[0], o3u[0]))
    say(f"bench: fast trace without the network {w3:.3f} s (particle-order gather {w3u:.3f} s)")
    # what the network gains at the tolerances where validation gives it a
    # band: over the plain method (network + the reference's fixed rule vs the
    # plain trace) and over the fast form without it
    sweep = []
    if not args.no_sweep:
        for tol in (1e-2, 1e-3, 1e-4, 3e-5, 1e-5):
            s = surrogate.with_tolerance(sur, tol)
            if not np.isfinite(s.cfg["band"]):
                sweep.append(dict(tol=tol, band_m=None))
                continue
            t2 = make_tracer(wl, geo, surrogate=s, **bud.tracer_kwargs())
            t2.field_shape = t3u.field_shape
            o, w = timed_traces(t2, 1, 0)
            t4 = make_tracer(wl, geo, surrogate=s)
            t4.field_shape = plain_tr.field_shape
            o4, w4 = timed_traces(t4, 1, 0)
            sweep.append(dict(tol=tol, band_m=s.cfg["band"],
                              with_plain_gather=dict(wall_s=w4, network_gain_wall_vs_plain=plain_wall / w4,
                                                     screen_t_rel_diff_vs_plain=screen_time_diff(plain_outs[0],
                                                                                                 o4[0])),
                              with_fast_form=dict(wall_s=w, value=o[0]["particle_stages"] / w,
                                                  network_gain_wall=w3u / w, network_gain_wall_vs_binned=w3 / w,
                                                  screen_t_rel_diff_vs_plain=screen_time_diff(plain_outs[0], o[0]))))
            say(f"bench: network at tol {tol:g}: band {s.cfg['band']:.2e} m, with the plain gather {w4:.2f} s "
                f"(plain {plain_wall:.2f} s), with the fast form {w:.3f} s (without the network {w3u:.3f} s in "
                f"particle order, {w3:.3f} s binned)")
        cost = surrogate_cost(wl, geo, sur, bud)
        say(f"bench: per point in the network's domain: fast gather {cost['fast_gather_ns_per_point']:.2f} ns, "
            f"network {cost['network_fp64_ns_per_point']:.2f} ns")
    # dominant kernel of the value trace: the fast form's field gather, at the
    # binned gather's launch shape (the tuned thread count, 64 points per
    # block), timed with HIP events on its stream
    pts = gather_points(wl)
    n_pts = len(pts)
    fast_shape = ((tr.field_shape or (256, 64))[0], 64)
    fast_ms, _ = field_gather_timing(wl, geo, shape=fast_shape, budget=bud, binned=True)
    ring, near, terms, tiled = gather_work_per_point(geo, pts, bud)
    fast_flops = n_pts * ((ring + near) * C_RING + terms * C_TERM + tiled * C_TABLE)
    fa = fast_flops / (fast_ms / 1e3) / 1e12
    fast_roof = {"bound": "fp64_valu", "achieved": fa, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                 "frac": fa / PEAK_FP64_TFLOPS, "traffic": FAST_TRAFFIC_BYTES,
                 "traffic_source": FAST_TRAFFIC_SOURCE,
                 "kernel": "bem_pert_kernel with the field budget (far rule "
                           f"{bud.far_rule}, near width {bud.near_width}, {bud.clusters.n_clusters} cluster "
                           f"expansions, {bud.tiles.n_tiles} tiles, the multipole zone) on the points in bin "
                           "order: the device code of the value trace's ls_bin_field_kernel (64 points per "
                           "block, wave-major, per-lane rules; table-only and multipole-zone points one per "
                           "thread)",
                 "kernel_ms_mean": fast_ms, "points_per_launch": n_pts, "field_shape": list(fast_shape),
                 "traffic_field_shape": list(FAST_TRAFFIC_SHAPE),
                 "ring_nodes_per_point": ring, "near_nodes_per_point": near, "expansion_terms_per_point": terms,
                 "tables_per_point": tiled, "flops_per_launch": fast_flops,
                 "note": "FP64 VALU-bound (MI355X FP64 vector peak = FP64 MFMA peak = 78.6 TFLOP/s); algorithmic "
                         "flops = (far-rule ring nodes of the segments each point sums + near-rule nodes) x 120 "
                         "+ expansion terms x 16 + tile tables x 576, counted on the host for the same points"}
    plain_shape = tuple(plain_tr.field_shape or (0, 0))
    field_ms, _ = field_gather_timing(wl, geo, shape=plain_shape)
    n_quad = wl["n_segments"] * 8 * 4
    plain_achieved = n_pts * n_quad * C_RING / (field_ms / 1e3) / 1e12
    plain_roof = {"bound": "fp64_valu", "achieved": plain_achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                  "frac": plain_achieved / PEAK_FP64_TFLOPS, "traffic": TRAFFIC_BYTES,
                  "traffic_source": TRAFFIC_SOURCE,
                  "kernel": "bem_pert_kernel, fixed rule (the plain trace's field gather)",
                  "kernel_ms_mean": field_ms, "points_per_launch": n_pts, "field_shape": list(plain_shape),
                  "flops_per_launch": n_pts * n_quad * C_RING}
    value = stages / wall
    res = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-RK-stages/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded cap emission, numpy default_rng(0)); surrogate trained on BEM samples",
        "config": {"workload": f"post-tip 1 um on 10 um post, 400-panel BEM, {wl['n_el']} e- in "
                               f"{wl['n_el'] // wl['group']} groups of 16, rtol 1e-8, screen 90 um, anode kill 100 um",
                   "form": f"surrogate + batched lockstep at field err <= {args.tol:g}: exterior multipole zone, "
                           "learned network beyond its band (inert if none), BEM gathers with the far rule, a "
                           "wider near rule and cluster expansions, all validated against the plain fixed rule; "
                           "FP64 field and step control, FP32 + Kahan pair forces",
                   "n_quad": n_quad, "field_shape": list(fast_shape), "plain_field_shape": list(plain_shape)},
        "value_check": value_check,
        "value_host_io": host_io,
        "plain_value": plain_value,
        "gain_vs_plain": value / plain_value,
        "budget": bud.report,
        "without_network": net_off,
        "surrogate": {"tol": args.tol, "band_m": rep["band"] if np.isfinite(rep["band"]) else None,
                      "band_panels": rep["band_panels"], "inert": bool(sur.inert),
                      "field_err_max": rep["field_err_max"], "field_err_median": rep["field_err_median"],
                      "field_err_p99": rep["field_err_p99"], "phi_err_max": rep["phi_err_max"],
                      "train_s": rep["t_train_s"], "labels_s": rep["t_labels_s"], "train_steps": rep["steps"],
                      "batch": rep["batch"], "n_samples": rep["n_samples"]},
        "tolerance_sweep": sweep,
        "network_cost": cost if not args.no_sweep else None,
        "network_error_by_distance": sur.report.get("error_by_distance"),
        "autotune": {"best": tuning["best"], "table": tuning["table"], "sample_attempts": tuning["attempts"]},
        "autotune_value": {"best": tuning_value["best"], "table": tuning_value["table"], "sample": "whole traces"},
        "attempts_per_group_mean": float(np.mean(outs[-1]["attempts"])),
        "plain_attempts_per_group_mean": float(np.mean(plain_outs[0]["attempts"])),
        "active_particle_stages_per_s": act / wall,
        "roofline": fast_roof,
        "roofline_plain_gather": plain_roof,
        "c2_plain_fp64": c2_plain_rate(),
    }
    if not args.no_c3:
        say("bench: C3-shaped sample (well, image-charge BEM, 4096 uneven groups, 8 attempts)")
        res["c3_sample"] = c3_sample_rate()
    if cpu is not None:
        res["cpu_baseline"] = cpu
    print(json.dumps(res))


if __name__ == "__main__":
    main()
