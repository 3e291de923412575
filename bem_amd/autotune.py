"""Autotuner for the trace's launch configuration (no counterpart in the
reference, which has one NumPy/CuPy code path).

For a configured SpecificParticleTracer it times traces capped at
`attempts` RK45 attempts per group, or whole traces (every candidate does
the same work: the same groups, the same attempts) and keeps the fastest
of:

* mode "persistent" with each workgroup size that holds the largest group
  (64, 128, 256 threads), one group per workgroup from a device queue;
* mode "lockstep" with the BEM field gather at 192 or 256 threads per
  block and 1/4x .. 4x the automatic points per block (the binned gather of
  far-rule traces always takes 64 points per block: only its thread count
  is tuned, 64 .. 256, i.e. one to four slices per point).

The winner is written back to the tracer (mode, block_threads,
field_shape) and returned with the whole timing table. Traces with the
field surrogate (lockstep only) skip the persistent candidates, and so do
far-rule traces, whose binned gather is lockstep's.
"""
import numpy as np

from . import _lib


def _auto_points_per_block(n, n_cu):
    """The library's automatic points per block (field_points_per_block)."""
    if n >= 64 * 12 * n_cu:
        return 64
    want = n // (48 * n_cu)
    pb = 1
    while pb * 2 <= want and pb < 32:
        pb *= 2
    return pb


def candidates(tracer, n_cu):
    sizes = np.diff(tracer.offsets)
    biggest = int(sizes.max())
    n = int(sizes.sum())
    out = []
    # a surrogate with no validated band is never installed (the trace runs without it)
    live_surrogate = tracer.surrogate is not None and not tracer.surrogate.inert
    binned = bool(getattr(tracer, "binned", 0)) and tracer.far_rule is not None and not live_surrogate \
        and tracer.precision != "mixed"
    image = getattr(tracer.geometry, "image_profile", None) is not None
    if not live_surrogate and not binned:
        for T in (64, 128, 256):
            if biggest > T or (tracer.precision == "mixed" and T != 256) or (image and T != 256):
                continue
            if image and tracer.precision == "mixed":
                continue
            out.append(dict(mode="persistent", block_threads=T, field_shape=None))
    pb0 = _auto_points_per_block(n, n_cu)
    pbs = [64] if binned else sorted({max(1, pb0 // 4), max(1, pb0 // 2), pb0, min(64, pb0 * 2), min(64, pb0 * 4)})
    for T in ((64, 128, 192, 256) if binned else (192, 256)):
        for pb in pbs:
            out.append(dict(mode="lockstep", block_threads=0, field_shape=(T, pb)))
    return out


def tune(tracer, attempts=6, log=None):
    """Time every candidate on `attempts` attempts per group (None: whole
    traces, for workloads whose point mix changes along the trace); apply
    and return the best (dict with "best" and "table")."""
    import ctypes

    ctx = tracer.geometry._ctx()
    n_cu = ctypes.c_int(0)
    ctx.check(ctx.lib.spt_device_cus(ctx.h, ctypes.byref(n_cu)), "spt_device_cus")
    saved = (tracer.mode, tracer.block_threads, getattr(tracer, "field_shape", None), tracer.max_steps)
    table = []
    try:
        if attempts is not None:
            tracer.max_steps = int(attempts)
        for cand in candidates(tracer, n_cu.value):
            tracer.mode = cand["mode"]
            tracer.block_threads = cand["block_threads"]
            tracer.field_shape = cand["field_shape"]
            tracer.trace()  # warm-up (graph instantiation, first-touch)
            ms = min(tracer.trace()["device_ms"] for _ in range(2))
            table.append(dict(cand, ms=ms))
            if log:
                log(f"autotune: {cand} -> {ms:.2f} ms")
    finally:
        tracer.mode, tracer.block_threads, tracer.field_shape, tracer.max_steps = saved
    best = min(table, key=lambda r: r["ms"])
    tracer.mode = best["mode"]
    tracer.block_threads = best["block_threads"]
    tracer.field_shape = best["field_shape"]
    tracer.tuning = dict(best=best, table=table, attempts=attempts)
    return tracer.tuning
