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
TRAFFIC_BYTES = 19.8e6  # HBM bytes per 262,144-point field-gather launch at (256, 64), from the rocprofv3 counter runs
FAR_TRAFFIC_BYTES = 13.9e6  # the same with the far rule (2, 16)


def workload(n_el=262144, group=16, n_segments=400, seed=0, rtol=1e-8):
    from bem_amd.emission import cap_source

    R, H = 1e-6, 10e-6
    src = cap_source(n_el, R, H, theta_max_deg=60.0, mean_ke_ev=0.2, seed=seed)
    return dict(R=R, H=H, E_gun=-1e8, n_segments=n_segments, src=src, group=group, screens=[90e-6], z_max=100e-6,
                rtol=rtol, atol=1e-12, n_el=n_el)


def make_tracer(wl, geo, surrogate=None, precision="pairs32", mode="lockstep", far_rule=None, multipole=None):
    from bem_amd import SpecificParticleTracer
    from bem_amd.particles import ParticleGroup

    src = wl["src"]
    n = len(src["x"])
    pg = ParticleGroup(dict(src, weight=np.full(n, 1.602176634e-19), status=np.ones(n), species="electron"))
    return SpecificParticleTracer(pg, wl["group"], geo, wl["screens"], rtol=wl["rtol"], atol=wl["atol"],
                                  z_max=wl["z_max"], mode=mode, precision=precision, surrogate=surrogate,
                                  far_rule=far_rule, multipole=multipole)


def timed_traces(tr, steps, warmup):
    import torch

    for _ in range(warmup):
        tr.trace()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    outs = [tr.trace() for _ in range(steps)]
    torch.cuda.synchronize()
    return outs, time.perf_counter() - t0


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


def far_rule_nodes_per_point(geo, pts, k4, k2, sample=8192):
    """Mean ring-kernel nodes the far rule sums per point (8 per segment at
    L <= D < k4 L, 4 at k4 L <= D < k2 L, 2 beyond; near segments, D < L,
    take the sinh rule and are not counted), over the first `sample`
    points."""
    prof = np.asarray(geo.solution.profile)
    p0, t = prof[:-1], np.diff(prof, axis=0)
    L = np.linalg.norm(t, axis=1)
    q = pts[:sample]
    rho, z = np.hypot(q[:, 0], q[:, 1]), q[:, 2]
    tt = np.clip(((rho[:, None] - p0[None, :, 0]) * t[None, :, 0] + (z[:, None] - p0[None, :, 1]) * t[None, :, 1])
                 / (t ** 2).sum(1)[None], 0.0, 1.0)
    D = np.hypot(rho[:, None] - (p0[None, :, 0] + tt * t[None, :, 0]), z[:, None] - (p0[None, :, 1] + tt * t[None, :, 1]))
    Lr = L[None]
    nodes = np.where(D < Lr, 0, np.where(D < k4 * Lr, 8, np.where(D < k2 * Lr, 4, 2)))
    return float(nodes.sum(1).mean())


def field_gather_timing(wl, geo, n_pts=262144, reps=10, shape=None, far_rule=None):
    """Average launch duration of the BEM field gather over gather_points(),
    timed with HIP events on the stream the kernel is launched on (the
    library is bound to torch's current stream); far_rule = (k4, k2) times
    the distance-adaptive far rule instead of the fixed one."""
    import contextlib

    pts = gather_points(wl, n_pts)
    ctx = geo._ctx()
    with (geo.far_rule(*far_rule) if far_rule else contextlib.nullcontext()):
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


def cpu_baseline(wl, seconds=15.0, workers=None):
    """The oracle on the first `workers` groups of the workload, one forked
    process per group and host core (the reference's own process-level
    group partitioning), each for `seconds` of work. Must run before
    anything touches the GPU (the workers are forked)."""
    import multiprocessing as mp

    from oracle import bem as ob
    from bem_amd.bem.mesh import post_profile

    R, H = wl["R"], wl["H"]
    prof = post_profile(R, H, wl["n_segments"])
    sigma = ob.solve(prof, lambda r, z: wl["E_gun"] * z)
    _CPU.update(wl=wl, seconds=seconds,
                geom=dict(kind="post_bem", E_gun=wl["E_gun"], R=R, H=H, profile=prof, sigma=sigma,
                          z_handoff=20 * (R + H), kill_z_below=0.0, z_max=wl["z_max"]))
    if workers is None:
        workers = max(1, min(16, len(os.sched_getaffinity(0))))
    with mp.get_context("fork").Pool(workers) as pool:
        parts = pool.map(_cpu_group_rate, range(workers))
    count = sum(c for c, _ in parts)
    el = max(e for _, e in parts)
    return dict(value=count / el, unit="particle-RK-stages/s", cores=workers, kind="port",
                sample=f"oracle (NumPy FP64) on groups 0..{workers - 1} ({wl['group']} electrons each) of the "
                       f"workload, one process per group and core, {el:.1f} s each ({count} particle-stages)")


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
    ap.add_argument("--tol", type=float, default=1e-6, help="surrogate field-error tolerance for the headline value")
    ap.add_argument("--train-steps", type=int, default=60000)
    ap.add_argument("--no-sweep", action="store_true", help="skip the looser-tolerance runs")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-c3", action="store_true", help="skip the C3-shaped sample (well with image charges)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--field-only", action="store_true",
                    help="only the field-gather timing and a 256 MiB device copy (for hardware-counter runs)")
    ap.add_argument("--field-shape", default="0,0",
                    help="threads,points-per-block of the field gather for --field-only (0,0 = automatic)")
    ap.add_argument("--value-only", action="store_true",
                    help="only the headline trace (far rule and multipole zone, no network: the network is inert "
                         "at 1e-6) after one warm-up trace, for profiler runs")
    ap.add_argument("--no-multipole", action="store_true",
                    help="no exterior multipole zone in the surrogate + batched traces")
    ap.add_argument("--far-rule", default="2,16",
                    help="k4,k2 of the distance-adaptive far rule in the surrogate + batched traces "
                         "(and in --field-only); 0 = the fixed rule")
    args = ap.parse_args()
    far_rule = tuple(float(v) for v in args.far_rule.split(","))
    far_rule = None if far_rule[0] == 0.0 else far_rule
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the tracing path is benchmarked per GPU (groups shard across GPUs "
              "with no exchange during integration); --gpus above 1 is not supported.", file=sys.stderr)
        sys.exit(2)
    from bem_amd import surrogate
    from bem_amd.bem.geometry import PostTipBEMGeometry

    def say(*a):
        print(*a, file=sys.stderr, flush=True)

    wl = workload(n_el=args.electrons)
    cpu = None
    if not args.no_cpu_baseline and not args.field_only and not args.value_only:
        say(f"bench: CPU baseline (oracle, {args.cpu_seconds:.0f} s per process)")
        cpu = cpu_baseline(wl, args.cpu_seconds)  # before the GPU is touched: the workers are forked
        say(f"bench: CPU baseline {cpu['value']:.0f} particle-stages/s on {cpu['cores']} cores")
    t0 = time.perf_counter()
    geo = PostTipBEMGeometry(wl["E_gun"], wl["R"], wl["H"], n_segments=wl["n_segments"])
    t_solve = time.perf_counter() - t0
    if args.field_only:
        import torch

        shape = tuple(int(v) for v in args.field_shape.split(","))
        field_ms, n_pts = field_gather_timing(wl, geo, reps=3, shape=shape, far_rule=far_rule)
        src = torch.empty(64 << 20, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps(dict(field_ms=field_ms, points=n_pts, field_shape=shape, far_rule=far_rule,
                              copy_bytes_read=src.numel() * 4,
                              copy_bytes_written=dst.numel() * 4)))
        return
    if args.value_only:
        from bem_amd.bem import multipole as mpl

        mp = None if args.no_multipole else mpl.fit(geo, n_terms=128, tol=min(1e-8, args.tol))
        vt = make_tracer(wl, geo, far_rule=far_rule, multipole=mp)
        vt.field_shape = (256, 64)
        o, w = timed_traces(vt, args.steps, args.warmup)
        print(json.dumps(dict(value=sum(x["particle_stages"] for x in o) / w, wall_s=w, steps=args.steps,
                              far_rule=far_rule, multipole_r_min=None if mp is None else mp.r_min)))
        return
    say(f"bench: BEM solve {t_solve:.2f} s; training the surrogate ({args.train_steps} steps)")
    sur = surrogate.train(geo, n_samples=1 << 20, steps=args.train_steps, batch=8192, tol=args.tol, chunk=10000)
    rep = sur.report
    say(f"bench: trained in {rep['t_train_s']:.1f} s; band at tol {args.tol:g}: {rep['band']:.3e} m")
    from bem_amd import autotune

    plain_tr = make_tracer(wl, geo)
    tuning = autotune.tune(plain_tr, attempts=4)
    say(f"bench: autotuned field shape {tuning['best']}")
    far_err = None
    if far_rule:
        far_err = surrogate.far_rule_error(geo, *far_rule, n=1 << 16)
        say(f"bench: far rule {far_rule}: field err max {far_err['field_err_max']:.2e}")
        if far_err["field_err_max"] > args.tol:
            say("bench: the far rule misses the tolerance; the surrogate + batched traces use the fixed rule")
            far_rule = None
    mp = None
    if not args.no_multipole:
        from bem_amd.bem import multipole as mpl

        mp = mpl.fit(geo, n_terms=128, tol=min(1e-8, args.tol))
        say(f"bench: multipole r_min {mp.r_min:.3e} m ({mp.report['r_min_over_r_ref']:.3f} r_ref)")
        if mp.inert:
            mp = None
    tr = make_tracer(wl, geo, surrogate=sur, far_rule=far_rule, multipole=mp)
    tuning_value = autotune.tune(tr, attempts=4) if (far_rule or mp) else None  # its own launch shape
    if tuning_value is None:
        tr.field_shape = plain_tr.field_shape
    else:
        say(f"bench: autotuned field shape of the fast trace {tuning_value['best']}")
    outs, wall = timed_traces(tr, args.steps, args.warmup)
    say(f"bench: {args.steps} surrogate traces in {wall:.2f} s")
    stages = sum(o["particle_stages"] for o in outs)
    act = sum(int(o["active_stages"].sum()) for o in outs)
    plain_outs, plain_wall = timed_traces(plain_tr, 1, 0)
    plain_value = plain_outs[0]["particle_stages"] / plain_wall
    say(f"bench: plain trace {plain_wall:.2f} s")
    net_off = None  # the same batched trace without the network: what the network itself gains
    if far_rule or mp:
        t3 = make_tracer(wl, geo, far_rule=far_rule, multipole=mp)
        t3.field_shape = tr.field_shape
        o3, w3 = timed_traces(t3, 1, 0)
        net_off = dict(value=o3[0]["particle_stages"] / w3, wall_s=w3,
                       attempts_mean=float(o3[0]["attempts"].mean()),
                       screen_t_rel_diff_vs_plain=screen_time_diff(plain_outs[0], o3[0]))
        say(f"bench: far-rule trace without the network {w3:.2f} s")
    base_wall = net_off["wall_s"] if net_off else plain_wall
    sweep = []
    if not args.no_sweep:
        for tol in (1e-4, 1e-3):
            s = surrogate.with_tolerance(sur, tol)
            if not np.isfinite(s.cfg["band"]):
                sweep.append(dict(tol=tol, band_m=None))
                continue
            t2 = make_tracer(wl, geo, surrogate=s, far_rule=far_rule, multipole=mp)
            t2.field_shape = tr.field_shape
            o, w = timed_traces(t2, 1, 0)
            sweep.append(dict(tol=tol, band_m=s.cfg["band"], value=o[0]["particle_stages"] / w,
                              wall_s=w, gain_wall_vs_plain=plain_wall / w, network_gain_wall=base_wall / w,
                              attempts_mean=float(o[0]["attempts"].mean()),
                              screen_t_rel_diff_vs_plain=screen_time_diff(plain_outs[0], o[0])))
            say(f"bench: tol {tol:g} trace {w:.2f} s")
    field_ms, n_pts = field_gather_timing(wl, geo, shape=plain_tr.field_shape)
    n_quad = wl["n_segments"] * 8 * 4
    plain_achieved = n_pts * n_quad * C_RING / (field_ms / 1e3) / 1e12
    far_gather = None
    if far_rule:
        far_ms, _ = field_gather_timing(wl, geo, shape=tr.field_shape, far_rule=far_rule)
        npp = far_rule_nodes_per_point(geo, gather_points(wl, n_pts), *far_rule)
        far_gather = dict(kernel_ms_mean=far_ms, nodes_per_point_mean=npp, points_per_launch=n_pts,
                          flops_per_launch=n_pts * npp * C_RING)
    value = stages / wall
    plain_roof = {"bound": "mfma", "achieved": plain_achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                  "frac": plain_achieved / PEAK_FP64_TFLOPS, "traffic": TRAFFIC_BYTES,
                  "traffic_source": "FETCH_SIZE x 2.000 + WRITE_SIZE x 1.000 (scales from a 256 MiB copy), "
                                    "profiles/bem_field_gather_post400_262k.md",
                  "kernel": "bem_pert_kernel, fixed rule (BEM field gather; the same device code as the plain "
                            "trace's ls_field_kernel)",
                  "kernel_ms_mean": field_ms, "points_per_launch": n_pts, "flops_per_launch": n_pts * n_quad * C_RING,
                  "note": "FP64 VALU-bound; MI355X FP64 vector peak = FP64 MFMA peak = 78.6 TFLOP/s"}
    far_roof = None
    if far_gather:
        fa = far_gather["flops_per_launch"] / (far_gather["kernel_ms_mean"] / 1e3) / 1e12
        far_roof = {"bound": "mfma", "achieved": fa, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                    "frac": fa / PEAK_FP64_TFLOPS, "traffic": FAR_TRAFFIC_BYTES,
                    "traffic_source": "FETCH_SIZE x 2.000 + WRITE_SIZE x 1.000 (scales from a 256 MiB copy), "
                                      "profiles/bem_field_gather_far_rule_post400_262k.md",
                    "kernel": f"bem_pert_kernel, far rule (k4, k2) = {far_rule} (the same device code as the "
                              "value trace's ls_field_kernel)",
                    "kernel_ms_mean": far_gather["kernel_ms_mean"], "points_per_launch": n_pts,
                    "field_shape": list(tr.field_shape or (0, 0)), "traffic_field_shape": [256, 64],
                    "nodes_per_point_mean": far_gather["nodes_per_point_mean"],
                    "flops_per_launch": far_gather["flops_per_launch"],
                    "note": "algorithmic flops = ring-kernel nodes the far rule sums (counted on the host for the "
                            "same points) x 120; near-rule nodes not counted"}
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
                   "form": f"surrogate + batched lockstep at field err <= {args.tol:g}: validated exterior "
                           "multipole expansion of the BEM charge beyond its radius, learned network beyond its "
                           f"band (inert if none), BEM gathers with the far rule {far_rule} (fixed rule if None) "
                           "elsewhere; FP64 field and step control, FP32 + Kahan pair forces",
                   "n_quad": n_quad, "field_shape": list(tr.field_shape or (0, 0)), "plain_field_shape": list(plain_tr.field_shape or (0, 0))},
        "plain_value": plain_value,
        "gain_vs_plain": value / plain_value,
        "far_rule": None if not far_rule else dict(far_err, screen_t_rel_diff_vs_plain=screen_time_diff(
            plain_outs[0], outs[-1])),
        "multipole": None if mp is None else {k: v for k, v in mp.report.items()
                                              if k not in ("candidates", "worst_beyond")},
        "without_network": net_off,
        "surrogate": {"tol": args.tol, "band_m": rep["band"] if np.isfinite(rep["band"]) else None,
                      "band_panels": rep["band_panels"],
                      "inert": bool(sur.inert),
                      "field_err_max": rep["field_err_max"], "field_err_median": rep["field_err_median"],
                      "field_err_p99": rep["field_err_p99"], "phi_err_max": rep["phi_err_max"],
                      "train_s": rep["t_train_s"], "labels_s": rep["t_labels_s"], "train_steps": rep["steps"],
                      "batch": rep["batch"], "n_samples": rep["n_samples"],
                      "screen_t_rel_diff_vs_plain": screen_time_diff(plain_outs[0], outs[-1])},
        "tolerance_sweep": sweep,
        "autotune": {"best": tuning["best"], "table": tuning["table"], "sample_attempts": tuning["attempts"]},
        "autotune_value": None if tuning_value is None else {"best": tuning_value["best"],
                                                             "table": tuning_value["table"]},
        "attempts_per_group_mean": float(np.mean(outs[-1]["attempts"])),
        "plain_attempts_per_group_mean": float(np.mean(plain_outs[0]["attempts"])),
        "active_particle_stages_per_s": act / wall,
        "roofline": far_roof if far_gather else plain_roof,
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
