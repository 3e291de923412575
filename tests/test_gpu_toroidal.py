"""Toroidal harmonics and ring-mode kernels on the GPU against the oracle
(reference bem/toroidal.py, bem/ring_modes.py): relative error per value
<= 1e-12 (device K/E are own fits within ~2 ulp of SciPy's; the
recurrences are well conditioned in the branch each element takes)."""
import sys

import numpy as np
import pytest

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


def _rel(got, ref):
    return np.max(np.abs(got - ref) / np.maximum(np.abs(ref), 1e-300))


@pytest.mark.parametrize("n_max", [0, 1, 16, 64])
def test_toroidal_q_vs_oracle(gpu, n_max):
    from bem_amd import _lib
    from oracle import toroidal as ot

    torch = gpu
    rng = np.random.default_rng(n_max)
    chi = np.concatenate([[1.0, 1.0 + 1e-15, 1.0 + 1e-9], 1.0 + 10 ** rng.uniform(-8, 2, 200)])
    ctx = _lib.Context.get(0)
    ctx.bind_stream()
    c = torch.as_tensor(chi, device=ctx.torch_device)
    q = torch.empty((n_max + 1, len(chi)), dtype=torch.float64, device=ctx.torch_device)
    dq = torch.empty_like(q)
    ctx.check(ctx.lib.spt_toroidal_q(ctx.h, _lib.ptr(c), len(chi), n_max, _lib.ptr(q), _lib.ptr(dq)), "toroidal")
    ctx.sync()
    rq, rdq = ot.toroidal_Q(chi, n_max)
    eq, edq = _rel(q.cpu().numpy(), rq), _rel(dq.cpu().numpy(), rdq)
    _say(f"toroidal n_max {n_max}: Q rel {eq:.2e}, dQ rel {edq:.2e}")
    assert eq < 1e-12 and edq < 1e-12


def test_ring_modes_vs_oracle(gpu):
    from bem_amd import _lib
    from oracle import toroidal as ot

    torch = gpu
    rng = np.random.default_rng(5)
    n, n_max = 300, 16
    rho, a = rng.uniform(0.05, 2, n), rng.uniform(0.05, 2, n)
    z, za = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
    rho[:10] = a[:10] * (1 + 1e-6)  # near coincidence
    z[:10] = za[:10]
    ctx = _lib.Context.get(0)
    ctx.bind_stream()
    dev = [torch.as_tensor(v, device=ctx.torch_device) for v in (rho, z, a, za)]
    outs = [torch.empty((n_max + 1, n), dtype=torch.float64, device=ctx.torch_device) for _ in range(4)]
    ctx.check(ctx.lib.spt_ring_modes(ctx.h, *(_lib.ptr(t) for t in dev), n, n_max, *(_lib.ptr(t) for t in outs)),
              "ring modes")
    ctx.sync()
    Phi, dPr, dPz = ot.ring_field_modes(rho, z, a, za, n_max)
    g = ot.point_charge_potential_modes(rho, z, a, za, n_max)
    errs = [_rel(o.cpu().numpy(), r) for o, r in zip(outs, (Phi, dPr, dPz, g))]
    _say("ring modes rel errors (Phi, dPhi/drho, dPhi/dz, g):", ", ".join(f"{e:.2e}" for e in errs))
    assert errs[0] < 1e-12 and errs[3] < 1e-12
    # derivatives: relative to each mode's largest value (cancellation where they cross zero)
    for o, r in ((outs[1], dPr), (outs[2], dPz)):
        got = o.cpu().numpy()
        assert np.max(np.abs(got - r) / np.abs(r).max(axis=1, keepdims=True)) < 1e-12
