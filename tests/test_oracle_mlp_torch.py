"""The oracle surrogate network (oracle/mlp.py) pinned against an
independent implementation: torch.nn layers, torch autograd and
torch.optim.Adam in float64 on the CPU. The device training and inference
kernels are checked against the oracle, so these tests are what ties the
whole chain to a library definition of the same network. No GPU."""
import numpy as np
import pytest

from oracle import mlp as om

torch = pytest.importorskip("torch")


def _params(seed=0, scale=0.15):
    rng = np.random.default_rng(seed)
    p = rng.normal(0.0, scale, om.layout()["n"])
    return p


def _torch_net(p):
    """torch.nn.Sequential of the flat layout: Linear(64, 128) + SiLU, three
    Linear(128, 128) + SiLU, Linear(128, 32) (torch's y = x W^T + b with W
    stored (out, in), as the layout stores it)."""
    Ws, bs = om.unpack(p)
    layers = []
    for l, (W, b) in enumerate(zip(Ws, bs)):
        lin = torch.nn.Linear(W.shape[1], W.shape[0], dtype=torch.float64)
        with torch.no_grad():
            lin.weight.copy_(torch.as_tensor(W))
            lin.bias.copy_(torch.as_tensor(b))
        layers.append(lin)
        if l < om.DEPTH:
            layers.append(torch.nn.SiLU())
    return torch.nn.Sequential(*layers)


def _cfg(seed=1):
    rng = np.random.default_rng(seed)
    return dict(fourier=rng.normal(0.0, 2.0, 2 * om.F), center_z=10e-6, r_unit=1e-6, u_offset=[0.3, 0.4],
                u_scale=[0.25, 1.1], out_scale=[2.0, 3e7, 5e7], env_r0=2e-6, env_pow=1.5)


def test_forward_matches_torch_nn():
    p = _params()
    X = np.random.default_rng(2).uniform(-1.0, 1.0, (257, om.IN))
    Y, _ = om.forward(p, X)
    with torch.no_grad():
        Yt = _torch_net(p)(torch.as_tensor(X)).numpy()
    assert Y.shape == Yt.shape == (257, om.OUTP)
    assert np.max(np.abs(Y - Yt)) <= 1e-13 * np.max(np.abs(Yt))


def test_loss_gradient_matches_torch_autograd():
    """Loss = mean over rows of sum_c w_c (Y_c - T_c)^2 over the first three
    outputs; gradient of every parameter block against autograd."""
    p = _params(3)
    rng = np.random.default_rng(4)
    X = rng.uniform(-1.0, 1.0, (96, om.IN))
    T = rng.normal(0.0, 1.0, (96, 3))
    w = np.array([0.5, 1.0, 2.0])
    loss, g = om.loss_and_grad(p, X, T, w)
    net = _torch_net(p)
    Y = net(torch.as_tensor(X))
    lt = (torch.as_tensor(w) * (Y[:, :3] - torch.as_tensor(T)) ** 2).sum() / X.shape[0]
    lt.backward()
    gt = []
    for m in net:
        if isinstance(m, torch.nn.Linear):
            gt.append(np.concatenate([m.weight.grad.numpy().ravel(), m.bias.grad.numpy()]))
    gt = np.concatenate(gt)
    assert abs(loss - lt.item()) <= 1e-14 * abs(lt.item())
    assert g.shape == gt.shape
    # per block, relative to the block's largest entry
    off = om.layout()
    names = ["W1", "b1"] + [f"{k}{l}" for l in range(2, om.DEPTH + 1) for k in ("W", "b")] + ["Wo", "bo", "n"]
    for a, b in zip(names[:-1], names[1:]):
        ga, gb = g[off[a]:off[b]], gt[off[a]:off[b]]
        assert np.max(np.abs(ga - gb)) <= 1e-12 * max(np.max(np.abs(gb)), 1e-300), a
    # outputs beyond the third carry no loss: their output rows get zero gradient
    Wo = g[off["Wo"]:off["bo"]].reshape(om.OUTP, om.W)
    assert np.all(Wo[3:] == 0.0) and np.all(g[off["bo"] + 3:off["n"]] == 0.0)


def test_adam_matches_torch_optim():
    """Five steps of the oracle's Adam against torch.optim.Adam (lr, betas
    (0.9, 0.999), eps 1e-8) fed the same gradients."""
    rng = np.random.default_rng(5)
    n = 1000
    p = rng.normal(0.0, 1.0, n)
    m = np.zeros(n)
    v = np.zeros(n)
    pt = torch.nn.Parameter(torch.as_tensor(p.copy()))
    opt = torch.optim.Adam([pt], lr=3e-3, betas=(0.9, 0.999), eps=1e-8)
    for t in range(1, 6):
        g = rng.normal(0.0, 10.0 ** rng.uniform(-6, 1), n)
        om.adam(p, g, m, v, t, 3e-3)
        pt.grad = torch.as_tensor(g.copy())
        opt.step()
        assert np.max(np.abs(p - pt.detach().numpy())) <= 1e-12 * np.max(np.abs(p)), t


def test_predict_matches_a_torch_pipeline():
    """Features, network and output scaling end to end: torch's sin/cos of
    2 pi (u B^T), the torch network, times out_scale and the envelope."""
    cfg = _cfg()
    p = _params(6)
    rng = np.random.default_rng(7)
    rho = rng.uniform(0.0, 5e-6, 300)
    z = 10e-6 + rng.uniform(1.2e-6, 8e-6, 300)
    pos = np.stack([rho * np.cos(1.0), rho * np.sin(1.0), z], -1)
    got = om.predict(cfg, p, pos)
    r = np.hypot(rho, z - cfg["center_z"])
    u = np.stack([(np.log(r / cfg["r_unit"]) - cfg["u_offset"][0]) * cfg["u_scale"][0],
                  ((z - cfg["center_z"]) / r - cfg["u_offset"][1]) * cfg["u_scale"][1]], -1)
    B = torch.as_tensor(np.asarray(cfg["fourier"]).reshape(om.F, 2))
    ang = 2.0 * np.pi * torch.as_tensor(u) @ B.T
    X = torch.cat([torch.sin(ang), torch.cos(ang)], dim=1)
    with torch.no_grad():
        Y = _torch_net(p)(X).numpy()[:, :3]
    env = (1.0 + r / cfg["env_r0"]) ** (-cfg["env_pow"])
    want = Y * np.asarray(cfg["out_scale"])[None, :] * env[:, None]
    assert np.max(np.abs(got - want) / np.max(np.abs(want), axis=0)) <= 1e-12


def test_a_planted_mistake_is_caught():
    """The comparison is sharp enough to see a transposed hidden weight."""
    p = _params(8)
    off = om.layout()
    q = p.copy()
    W2 = q[off["W2"]:off["b2"]].reshape(om.W, om.W)
    q[off["W2"]:off["b2"]] = W2.T.ravel()
    X = np.random.default_rng(9).uniform(-1.0, 1.0, (64, om.IN))
    Y, _ = om.forward(q, X)
    with torch.no_grad():
        Yt = _torch_net(p)(torch.as_tensor(X)).numpy()
    assert np.max(np.abs(Y - Yt)) > 1e-3 * np.max(np.abs(Yt))
