"""Surrogate kernels on the GPU against the FP64 oracle, and what the
trained surrogate gains in a trace.

* BEM potential (spt_bem_potential) vs oracle.bem.potential: 1e-11
  relative to the largest |phi| (same adaptive rule; sums differ in order).
* Gradient (spt_mlp_grad) vs oracle.mlp.loss_and_grad on the same batch:
  three-term BF16 split <= 1e-5 of the largest gradient entry per block
  (FP32 level); single BF16 terms <= 3e-2 in norm (BF16 has 8 bits).
* One Adam step through spt_mlp_train vs oracle.mlp.adam: 1e-6 relative.
* Inference (spt_mlp_infer) vs oracle.mlp.predict, relative to the
  output's largest magnitude: FP64 MFMA <= 1e-12; three-term BF16 split
  <= 1e-5.
* Training drives the loss down; a surrogate trained on the post tip is
  validated against the BEM and, used in a lockstep trace beyond its band,
  keeps the trajectories within the validated error of the plain ones (the
  device-time gain is printed; bench.py measures it).
"""
import ctypes
import sys

import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from bem_amd import _lib

    _lib.load()
    return torch


def _say(*a):
    print(*a, flush=True, file=sys.stderr)


def _cfg_dict(rng):
    return dict(fourier=rng.normal(0, 1.5, (32, 2)).astype(np.float32), center_z=10e-6, r_unit=1e-6,
                u_offset=(2.0, 0.0), u_scale=(0.5, 1.0), out_scale=(3.0, 2e8, 5e8), env_r0=1e-6, env_pow=2.5,
                rho_max=1e-4, z_min=10e-6, z_max=1.2e-4, band=0.0)


def _cfg_struct(d):
    from bem_amd.surrogate import FieldSurrogate

    return FieldSurrogate(None, d, None).cfg_struct()


def _samples(rng, n, cfgd):
    """Sample rows whose scaled targets are O(1), like real training data."""
    from oracle import mlp as om

    rho = rng.uniform(0, 20e-6, n)
    z = 10e-6 + rng.uniform(1.2e-6, 30e-6, n)
    x = np.stack([rho, z], -1)
    _, _, env = om.coords(cfgd, rho, z)
    y = rng.normal(0, 1, (n, 3)) * np.asarray(cfgd["out_scale"])[None, :] * env[:, None]
    return x, y


def _ctx():
    from bem_amd import _lib

    ctx = _lib.Context.get(0)
    ctx.bind_stream()
    return ctx


def test_potential_vs_oracle(gpu):
    from bem_amd.bem.solver import AxisymmetricBEMSolution
    from oracle import bem as ob

    d = golden("bem_post80.npz")
    sol = AxisymmetricBEMSolution(d["profile"], d["sigma"])
    rng = np.random.default_rng(3)
    th = rng.uniform(0, np.pi, 40)
    dist = 10 ** rng.uniform(-9, -5, 40)
    R, H = float(d["R"]), float(d["H"])
    pts = np.stack([(R + dist) * np.sin(th), np.zeros(40), H + (R + dist) * np.cos(th)], -1)
    nodes = np.stack([d["profile"][:, 0], np.zeros(81), d["profile"][:, 1]], -1)
    pts = np.concatenate([pts, nodes[::7], d["points"][:20]])
    got = sol.potential(pts)
    ref = ob.potential(d["profile"], d["sigma"], pts)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    _say(f"potential vs oracle: max rel {err:.2e} over {len(pts)} points")
    assert err < 1e-11


@pytest.mark.parametrize("split", [1, 0])
def test_mlp_gradient_vs_oracle(gpu, split):
    from bem_amd import surrogate
    from oracle import mlp as om

    torch = gpu
    rng = np.random.default_rng(5)
    cfgd = _cfg_dict(rng)
    c = _cfg_struct(cfgd)
    B = 96
    x, y = _samples(rng, B, cfgd)
    p = surrogate.initial_params(1)
    off = om.layout()
    p[off["b1"]:off["b1"] + 128] = rng.normal(0, 0.1, 128)  # nonzero biases exercise the epilogues
    ctx = _ctx()
    dev = ctx.torch_device
    P = torch.as_tensor(p, device=dev)
    X = torch.as_tensor(x, device=dev).contiguous()
    Y = torch.as_tensor(y, device=dev).contiguous()
    g = torch.empty_like(P)
    loss = ctypes.c_float(0)
    w = (ctypes.c_float * 3)(0.25, 1.0, 1.0)
    from bem_amd import _lib

    ctx.check(ctx.lib.spt_mlp_grad(ctx.h, ctypes.byref(c), _lib.ptr(P), _lib.ptr(X), _lib.ptr(Y), B, w, split,
                                   _lib.ptr(g), ctypes.byref(loss)), "spt_mlp_grad")
    X0, T = om.batch(cfgd, x, y)
    ref_loss, ref_g = om.loss_and_grad(p.astype(np.float64), X0, T, (0.25, 1.0, 1.0))
    got = g.cpu().numpy().astype(np.float64)
    lerr = abs(loss.value - ref_loss) / ref_loss
    nerr = np.linalg.norm(got - ref_g) / np.linalg.norm(ref_g)
    blocks = {}
    for k in ("W1", "b1", "W2", "b2", "W3", "b3", "W4", "b4", "Wo", "bo"):
        a = off[k]
        b = a + ({"W1": 128 * 64, "Wo": 32 * 128, "bo": 32}.get(k, 128 * 128 if k.startswith("W") else 128))
        blocks[k] = np.abs(got[a:b] - ref_g[a:b]).max() / max(np.abs(ref_g[a:b]).max(), 1e-30)
    _say(f"mlp grad split={split}: loss rel {lerr:.2e}, grad norm rel {nerr:.2e}, "
         + ", ".join(f"{k} {v:.1e}" for k, v in blocks.items()))
    if split:
        assert lerr < 1e-5 and nerr < 1e-5
        assert max(blocks.values()) < 1e-5
    else:
        assert lerr < 3e-2 and nerr < 3e-2


def test_adam_step_vs_oracle(gpu):
    from bem_amd import _lib, surrogate
    from oracle import mlp as om

    torch = gpu
    rng = np.random.default_rng(6)
    cfgd = _cfg_dict(rng)
    c = _cfg_struct(cfgd)
    B = 64
    x, y = _samples(rng, B, cfgd)
    p = surrogate.initial_params(2)
    ctx = _ctx()
    dev = ctx.torch_device
    X = torch.as_tensor(x, device=dev).contiguous()
    Y = torch.as_tensor(y, device=dev).contiguous()
    w = (ctypes.c_float * 3)(1.0, 1.0, 1.0)
    P = torch.as_tensor(p, device=dev)
    g = torch.empty_like(P)
    loss = ctypes.c_float(0)
    ctx.check(ctx.lib.spt_mlp_grad(ctx.h, ctypes.byref(c), _lib.ptr(P), _lib.ptr(X), _lib.ptr(Y), B, w, 1,
                                   _lib.ptr(g), ctypes.byref(loss)), "spt_mlp_grad")
    gh = g.cpu().numpy().astype(np.float64)
    m = torch.zeros_like(P)
    v = torch.zeros_like(P)
    m[:] = 0.01
    v[:] = 1e-4
    buf = np.zeros(1, dtype=np.float32)
    lr = 1e-3
    # seed 0: the batch is rows 0 .. B-1, the same rows spt_mlp_grad used
    ctx.check(ctx.lib.spt_mlp_train(ctx.h, ctypes.byref(c), _lib.ptr(P), _lib.ptr(m), _lib.ptr(v), 4, _lib.ptr(X),
                                    _lib.ptr(Y), B, B, 1, lr, 0, w, 1, buf.ctypes.data_as(ctypes.c_void_p)),
              "spt_mlp_train")
    pr, mr, vr = p.astype(np.float64), np.full(p.size, 0.01), np.full(p.size, 1e-4)
    # the kernel's hyperparameters are FP32: compare with the same rounded values
    om.adam(pr, gh, mr, vr, 5, lr, b1=float(np.float32(0.9)), b2=float(np.float32(0.999)))
    dp = np.abs(P.cpu().numpy() - pr).max() / lr
    dm = np.abs(m.cpu().numpy() - mr).max() / np.abs(mr).max()
    dv = np.abs(v.cpu().numpy() - vr).max() / np.abs(vr).max()
    _say(f"adam step: |dp|/lr {dp:.2e}, m rel {dm:.2e}, v rel {dv:.2e}, loss {buf[0]:.4e} vs {loss.value:.4e}")
    assert dp < 1e-3 and dm < 1e-6 and dv < 1e-6
    assert abs(buf[0] - loss.value) <= 1e-6 * abs(loss.value)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_mlp_inference_vs_oracle(gpu, precision):
    from bem_amd import _lib, surrogate
    from oracle import mlp as om

    torch = gpu
    rng = np.random.default_rng(7)
    cfgd = _cfg_dict(rng)
    c = _cfg_struct(cfgd)
    p = surrogate.initial_params(3)
    off = om.layout()
    p[off["bo"]:off["bo"] + 3] = (0.3, -0.2, 0.1)
    n = 1000  # not a multiple of the 32 points per wavefront
    th = rng.uniform(0, 2 * np.pi, n)
    rho = rng.uniform(0, 20e-6, n)
    pos = np.stack([rho * np.cos(th), rho * np.sin(th), 10e-6 + rng.uniform(1.2e-6, 30e-6, n)], -1)
    pos[:3, :2] = 0.0  # on the axis
    ctx = _ctx()
    P = torch.as_tensor(p, device=ctx.torch_device)
    X = torch.as_tensor(pos, device=ctx.torch_device).contiguous()
    out = torch.empty_like(X)
    ctx.check(ctx.lib.spt_mlp_infer(ctx.h, ctypes.byref(c), _lib.ptr(P), _lib.ptr(X), n,
                                    _lib.PREC_FP64 if precision == "fp64" else _lib.PREC_MIXED, _lib.ptr(out)),
              "spt_mlp_infer")
    ctx.sync()
    got = out.cpu().numpy()
    ref = om.predict(cfgd, p.astype(np.float64), pos)
    err = np.abs(got - ref).max(axis=0) / np.abs(ref).max(axis=0)
    _say(f"mlp inference [{precision}] vs FP64 oracle: rel err (phi, E_rho, E_z) = {err}")
    assert err.max() < (1e-12 if precision == "fp64" else 1e-5)


def test_training_reduces_the_loss(gpu):
    from bem_amd import _lib, surrogate

    torch = gpu
    rng = np.random.default_rng(8)
    cfgd = _cfg_dict(rng)
    cfgd["out_scale"] = (1.0, 1.0, 1.0)
    cfgd["env_pow"] = 0.0
    c = _cfg_struct(cfgd)
    n = 1 << 14
    rho = rng.uniform(0, 20e-6, n)
    z = 10e-6 + rng.uniform(1.2e-6, 30e-6, n)
    r = np.hypot(rho, z - 10e-6)
    y = np.stack([np.log(r / 1e-6), (z - 10e-6) / r, np.sin(rho / 4e-6)], -1)
    ctx = _ctx()
    dev = ctx.torch_device
    P = torch.as_tensor(surrogate.initial_params(4), device=dev)
    m = torch.zeros_like(P)
    v = torch.zeros_like(P)
    X = torch.as_tensor(np.stack([rho, z], -1), device=dev).contiguous()
    Y = torch.as_tensor(y, device=dev).contiguous()
    w = (ctypes.c_float * 3)(1.0, 1.0, 1.0)
    steps = 600
    buf = np.zeros(steps, dtype=np.float32)
    ctx.check(ctx.lib.spt_mlp_train(ctx.h, ctypes.byref(c), _lib.ptr(P), _lib.ptr(m), _lib.ptr(v), 0, _lib.ptr(X),
                                    _lib.ptr(Y), n, 1024, steps, 2e-3, 12345, w, 0,
                                    buf.ctypes.data_as(ctypes.c_void_p)), "spt_mlp_train")
    first, last = float(buf[:20].mean()), float(buf[-20:].mean())
    _say(f"training: loss {first:.3e} -> {last:.3e} over {steps} steps")
    assert np.all(np.isfinite(buf)) and last < 0.02 * first


def test_surrogate_gain_in_a_trace(gpu):
    """Train a small surrogate on the post tip, validate it against the BEM, and
    trace the same 16384 electrons (1024 groups of 16, rtol 1e-8, FP32
    pairs) in lockstep with and without it. The band is the validation's for a
    loose 5e-2 field tolerance (this short training reaches ~1e-2 near the
    surface), so the test shows the mechanism: the network serves the field
    beyond its band and screen-crossing times agree with the plain run to
    1e-3 relative. Its gain over the plain method is printed; the measured
    gains at the validated tolerances of the fully trained network, over the
    plain method and over the fast form, are bench.py's (tolerance_sweep)."""
    from bem_amd import SpecificParticleTracer, surrogate
    from bem_amd.bem.geometry import PostTipBEMGeometry
    from bem_amd.emission import cap_source
    from bem_amd.particles import ParticleGroup

    d = golden("bem_post400.npz")
    R, H = float(d["R"]), float(d["H"])
    geo = PostTipBEMGeometry(-1e8, R, H, n_segments=400, sigma=d["sigma"])
    sur = surrogate.train(geo, n_samples=1 << 16, steps=3000, batch=2048, tol=5e-2, n_val=1 << 14, chunk=1000,
                          d_max=110e-6, log=_say)
    rep = sur.report
    _say(f"validation: band {rep['band']:.3e} m ({rep['band_panels']} panels), field err max {rep['field_err_max']:.2e} "
         f"median {rep['field_err_median']:.2e}, phi err max {rep['phi_err_max']:.2e}; labels {rep['t_labels_s']:.2f}"
         f" s, training {rep['t_train_s']:.2f} s")
    assert np.isfinite(rep["band"])
    n = 16384
    src = cap_source(n, R, H, seed=1)
    pg = ParticleGroup(dict(src, weight=np.full(n, 1.602176634e-19), status=np.ones(n), species="electron"))
    runs = {}
    for name, s in (("plain", None), ("surrogate", sur)):
        tr = SpecificParticleTracer(pg, 16, geo, [90e-6], rtol=1e-8, z_max=100e-6, mode="lockstep",
                                    precision="pairs32", surrogate=s)
        out = tr.trace()
        runs[name] = out
        _say(f"trace [{name}]: {out['device_ms']:.1f} ms, attempts mean {out['attempts'].mean():.1f}, "
             f"{out['particle_stages'] / out['device_ms'] * 1e3:.3e} particle-stages/s")
    a, b = runs["plain"], runs["surrogate"]
    gain = a["device_ms"] / b["device_ms"]
    ea, eb = a["events"], b["events"]
    assert len(ea) == len(eb) == n
    ka = np.lexsort((ea["particle"], ea["group"]))
    kb = np.lexsort((eb["particle"], eb["group"]))
    dt = np.abs(ea["t"][ka] - eb["t"][kb]) / ea["t"][ka]
    _say(f"surrogate gain {gain:.2f}x in device time; screen-crossing time rel diff max {dt.max():.2e} "
         f"median {np.median(dt):.2e}")
    assert dt.max() < 1e-3


def test_surrogate_on_the_hemispherical_tip(gpu):
    """The same surrogate on the reference's own BEM tip geometry (sphere of
    radius R on the plane, cap centre at the origin): trained briefly,
    validated against the BEM (finite band at a loose 5e-2 tolerance), and a
    lockstep trace of 4096 electrons with it crosses the screen at the
    plain run's times to 1e-3 relative."""
    from bem_amd import SpecificParticleTracer, surrogate
    from bem_amd.bem.geometry import HemisphericalTipBEMGeometry
    from bem_amd.emission import cap_source
    from bem_amd.particles import ParticleGroup

    R = 1e-6
    geo = HemisphericalTipBEMGeometry(-1e8, R)
    sur = surrogate.train(geo, n_samples=1 << 16, steps=3000, batch=2048, tol=5e-2, n_val=1 << 14, chunk=1000,
                          d_max=40e-6)
    rep = sur.report
    _say(f"tip surrogate: band {rep['band']:.3e} m, field err max {rep['field_err_max']:.2e} "
         f"median {rep['field_err_median']:.2e}")
    assert np.isfinite(rep["band"])
    n = 4096
    src = cap_source(n, R, 0.0, seed=2)
    pg = ParticleGroup(dict(src, weight=np.full(n, 1.602176634e-19), status=np.ones(n), species="electron"))
    runs = {}
    for name, s in (("plain", None), ("surrogate", sur)):
        tr = SpecificParticleTracer(pg, 16, geo, [30e-6], rtol=1e-8, z_max=40e-6, mode="lockstep", surrogate=s)
        runs[name] = tr.trace()
    a, b = runs["plain"], runs["surrogate"]
    ea, eb = a["events"], b["events"]
    assert len(ea) == len(eb) == n
    ka = np.lexsort((ea["particle"], ea["group"]))
    kb = np.lexsort((eb["particle"], eb["group"]))
    dt = np.abs(ea["t"][ka] - eb["t"][kb]) / ea["t"][ka]
    _say(f"tip surrogate trace: {a['device_ms']:.1f} ms plain vs {b['device_ms']:.1f} ms surrogate; "
         f"screen-time rel diff max {dt.max():.2e}")
    assert dt.max() < 1e-3
