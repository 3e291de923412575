"""Benchmark: particle-RK-stages/s of the MI355X tracing path.

Workload (one "step" = one full trace, every group from birth to t_max):
hemisphere-capped post, R = 1 um on a 10 um post, 400-panel mirror-doubled
axisymmetric BEM (12,800 field quadrature nodes), uniform -1e8 V/m (anode
+10 kV at 100 um; particles absorbed there), 1024 electrons emitted from the
cap in 64 groups of 16 (polar angle U[0, 60 deg], KE ~ Exp(0.2 eV),
cos-weighted), FP64, RK45 rtol 1e-9, atol 1e-12, screen at 90 um, seed 0.

value = 7 x sum_g(size_g x attempts_g) / time (all particle slots of
running groups, the reference's accel_fn shape). The roofline entry is for
the dominant kernel (the persistent trace kernel, one launch per step):
algorithmic FP64 flops = active particle-stages x 12,800 nodes x 120 flops
per node + pair flops, over its device time from HIP events on the launch
stream. cpu_baseline: the oracle (NumPy FP64 restatement of the reference)
on a time-bounded sample of the same workload, one process.
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
W_PAIR_PER_PARTNER = 20


def workload(n_groups=64, group=16, n_segments=400, seed=0):
    from bem_amd.emission import cap_source

    R, H = 1e-6, 10e-6
    src = cap_source(n_groups * group, R, H, theta_max_deg=60.0, mean_ke_ev=0.2, seed=seed)
    return dict(R=R, H=H, E_gun=-1e8, n_segments=n_segments, src=src, group=group, screens=[90e-6], z_max=100e-6,
                rtol=1e-9, atol=1e-12)


def gpu_bench(wl, steps, warmup, block_threads, mode="auto", precision="fp64"):
    import torch

    from bem_amd import SpecificParticleTracer
    from bem_amd.bem.geometry import PostTipBEMGeometry
    from bem_amd.particles import ParticleGroup

    src = wl["src"]
    n = len(src["x"])
    pg = ParticleGroup(dict(src, weight=np.full(n, 1.602176634e-19), status=np.ones(n), species="electron"))
    geo = PostTipBEMGeometry(wl["E_gun"], wl["R"], wl["H"], n_segments=wl["n_segments"])
    tr = SpecificParticleTracer(pg, wl["group"], geo, wl["screens"], rtol=wl["rtol"], atol=wl["atol"],
                                z_max=wl["z_max"], block_threads=block_threads, mode=mode, precision=precision)
    for _ in range(warmup):
        tr.trace()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    outs = [tr.trace() for _ in range(steps)]
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return tr, outs, wall


def field_gather_timing(wl, reps=20):
    """Average duration of the BEM field gather (the trace's dominant
    kernel; the standalone entry point runs the same device code and launch
    shape) over the workload's particles displaced 10 nm .. 10 um off the
    cap, timed with HIP events on the launch stream."""
    import torch

    from bem_amd import _lib
    from bem_amd.bem.geometry import PostTipBEMGeometry

    src = wl["src"]
    n = len(src["x"])
    R, H = wl["R"], wl["H"]
    p = np.stack([src["x"], src["y"], src["z"] - H], -1)
    rc = np.linalg.norm(p, axis=1, keepdims=True)
    d = 10 ** np.random.default_rng(1).uniform(-8, -5, (n, 1))
    pts = p / rc * (R + d)
    pts[:, 2] += H
    geo = PostTipBEMGeometry(wl["E_gun"], R, H, n_segments=wl["n_segments"])
    sol = geo.solution
    sol.field(pts[:4])  # upload
    ctx = _lib.Context.get(0)
    pt = torch.as_tensor(pts, device="cuda").contiguous()
    out = torch.empty_like(pt)
    ctx.bind_stream()
    for _ in range(2):
        ctx.check(ctx.lib.spt_bem_pert_field(ctx.h, _lib.ptr(pt), n, _lib.ptr(out), 0), "field")
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        ctx.check(ctx.lib.spt_bem_pert_field(ctx.h, _lib.ptr(pt), n, _lib.ptr(out), 0), "field")
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps, n


def cpu_baseline(wl, seconds=15.0):
    """Oracle on the first group of the workload until `seconds` of work."""
    from oracle import bem as ob
    from oracle import tracker as otr

    src = wl["src"]
    g = wl["group"]
    R, H = wl["R"], wl["H"]
    from bem_amd.bem.mesh import post_profile

    prof = post_profile(R, H, wl["n_segments"])
    sigma = ob.solve(prof, lambda r, z: wl["E_gun"] * z)
    st = otr.setup(*(src[k][:g] for k in ("x", "y", "z", "px", "py", "pz", "t", "id")), g)
    geom = dict(kind="post_bem", E_gun=wl["E_gun"], R=R, H=H, profile=prof, sigma=sigma, z_handoff=20 * (R + H),
                kill_z_below=0.0, z_max=wl["z_max"])
    accel = otr.make_accel(geom, st["charge"], st["mass_p"], st["q_over_m"], 1e-12)
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

    from oracle import rk45

    t_max = otr.default_t_max(st, wl["E_gun"], wl["screens"])
    try:
        rk45.integrate(timed, st["pos"], st["vel"], st["t_birth"], t_max, wl["rtol"], wl["atol"],
                       kill=lambda p: otr.kill_mask(geom, p))
    except Done:
        pass
    el = time.perf_counter() - t0[0]
    return dict(value=count[0] / el, unit="particle-RK-stages/s", cores=1, kind="port",
                sample=f"oracle (NumPy FP64) on group 0 ({g} electrons) of the workload for {el:.1f} s "
                       f"({count[0]} particle-stages), 1 process")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--block-threads", type=int, default=0)
    ap.add_argument("--mode", default="auto", choices=["auto", "persistent", "lockstep"])
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the tracing path is benchmarked per GPU (groups shard across GPUs "
              "with no exchange during integration); --gpus above 1 is not supported.", file=sys.stderr)
        sys.exit(2)
    wl = workload()
    tr, outs, wall = gpu_bench(wl, args.steps, args.warmup, args.block_threads, args.mode, args.precision)
    stages = sum(o["particle_stages"] for o in outs)
    act = sum(int(o["active_stages"].sum()) for o in outs)
    dev_ms = np.array([o["device_ms"] for o in outs])
    n_quad = wl["n_segments"] * 8 * 4
    flops = act * (n_quad * C_RING + W_PAIR_PER_PARTNER * wl["group"])
    trace_tflops = flops / (dev_ms.sum() / 1e3) / 1e12
    field_ms, n_pts = field_gather_timing(wl)
    achieved = n_pts * n_quad * C_RING / (field_ms / 1e3) / 1e12
    res = {
        "metric": METRIC,
        "value": stages / wall,
        "unit": "particle-RK-stages/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f64/f32-kahan",
        "data": "synthetic (seeded cap emission, numpy default_rng(0))",
        "config": {"workload": "post-tip 1 um on 10 um post, 400-panel BEM, 1024 e- in 64 groups of 16, fp64 plain, "
                               "rtol 1e-9, screen 90 um, anode kill 100 um",
                   "form": f"{'plain FP64' if args.precision == 'fp64' else 'mixed FP32-Kahan'} BEM field, "
                           f"{args.mode} schedule",
                   "n_quad": n_quad, "block_threads": args.block_threads or 256},
        "active_particle_stages_per_s": act / wall,
        "attempts_per_group_mean": float(np.mean(outs[-1]["attempts"])),
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64_TFLOPS, "traffic": None,
                     "kernel": "bem_pert_kernel (BEM field gather; same device code and launch shape as the "
                               "trace's ls_field_kernel)",
                     "kernel_ms_mean": field_ms, "points_per_launch": n_pts,
                     "flops_per_launch": n_pts * n_quad * C_RING,
                     "whole_trace_tflops": trace_tflops, "trace_ms_mean": float(dev_ms.mean()),
                     "note": "FP64 VALU-bound; MI355X FP64 vector peak = FP64 MFMA peak = 78.6 TFLOP/s"},
    }
    if not args.no_cpu_baseline:
        res["cpu_baseline"] = cpu_baseline(wl, args.cpu_seconds)
    print(json.dumps(res))


if __name__ == "__main__":
    main()
