"""Oracle Fourier-feature MLP field surrogate, FP64 NumPy: the CPU
reference that the device training and inference kernels are checked
against. The surrogate has no counterpart in the reference (which
evaluates the BEM field directly); this module restates the network and
its gradient as defined by the library's C ABI (include/spt.h,
spt_mlp_cfg and the flat parameter layout).

Test infrastructure only: imported by tests and by bench.py's CPU
baseline, never by the product path.

Pinned by mathematics rather than by the reference: the analytic backward
pass is checked against central finite differences of the forward pass,
and SiLU/feature values against closed forms (tests/test_oracle_mlp.py).
"""
import numpy as np

F, IN, W, DEPTH, OUTP = 32, 64, 128, 4, 32


def layout():
    """Offsets of the flat FP32 parameter vector: W1 (W x IN), b1, then per
    hidden layer W (W x W), b, then Wout (OUTP x W), bout."""
    off = {}
    o = 0
    off["W1"] = o
    o += W * IN
    off["b1"] = o
    o += W
    for l in range(2, DEPTH + 1):
        off[f"W{l}"] = o
        o += W * W
        off[f"b{l}"] = o
        o += W
    off["Wo"] = o
    o += OUTP * W
    off["bo"] = o
    o += OUTP
    off["n"] = o
    return off


def unpack(p):
    off = layout()
    p = np.asarray(p, dtype=np.float64)
    Ws = [p[off["W1"]:off["W1"] + W * IN].reshape(W, IN)]
    bs = [p[off["b1"]:off["b1"] + W]]
    for l in range(2, DEPTH + 1):
        Ws.append(p[off[f"W{l}"]:off[f"W{l}"] + W * W].reshape(W, W))
        bs.append(p[off[f"b{l}"]:off[f"b{l}"] + W])
    Ws.append(p[off["Wo"]:off["Wo"] + OUTP * W].reshape(OUTP, W))
    bs.append(p[off["bo"]:off["bo"] + OUTP])
    return Ws, bs


def pack(Ws, bs):
    return np.concatenate([np.concatenate([w.ravel(), b]) for w, b in zip(Ws, bs)])


def coords(cfg, rho, z):
    """Normalised polar inputs (u0, u1) about (0, center_z) and the output
    envelope (1 + r / env_r0)^-env_pow."""
    dz = z - cfg["center_z"]
    r = np.maximum(np.hypot(rho, dz), 1e-30)
    u0 = (np.log(r / cfg["r_unit"]) - cfg["u_offset"][0]) * cfg["u_scale"][0]
    u1 = (dz / r - cfg["u_offset"][1]) * cfg["u_scale"][1]
    env = (1.0 + r / cfg["env_r0"]) ** (-cfg["env_pow"])
    return u0, u1, env


def features(cfg, u0, u1):
    """(n, 64): sin then cos of 2 pi (B u) for the 32 frequency rows of B."""
    B = np.asarray(cfg["fourier"], dtype=np.float64).reshape(F, 2)
    t = 2.0 * np.pi * (np.outer(u0, B[:, 0]) + np.outer(u1, B[:, 1]))
    return np.concatenate([np.sin(t), np.cos(t)], axis=1)


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(p, X0):
    """Network outputs (n, OUTP) and the cache for backward."""
    Ws, bs = unpack(p)
    H = X0
    Zs, Hs = [], [X0]
    for l in range(DEPTH):
        Z = H @ Ws[l].T + bs[l]
        H = silu(Z)
        Zs.append(Z)
        Hs.append(H)
    Y = H @ Ws[DEPTH].T + bs[DEPTH]
    return Y, (Zs, Hs)


def loss_and_grad(p, X0, T, w):
    """Loss = mean over rows of sum_c w_c (Y_c - T_c)^2 (c < 3) and its
    gradient in the flat layout."""
    Ws, bs = unpack(p)
    Y, (Zs, Hs) = forward(p, X0)
    B = X0.shape[0]
    w = np.asarray(w, dtype=np.float64)
    d = Y[:, :3] - T
    loss = float(np.sum(w * d * d) / B)
    dY = np.zeros_like(Y)
    dY[:, :3] = 2.0 * w * d / B
    gW = [None] * (DEPTH + 1)
    gb = [None] * (DEPTH + 1)
    gW[DEPTH] = dY.T @ Hs[DEPTH]
    gb[DEPTH] = dY.sum(axis=0)
    dH = dY @ Ws[DEPTH]
    for l in range(DEPTH - 1, -1, -1):
        dZ = dH * dsilu(Zs[l])
        gW[l] = dZ.T @ Hs[l]
        gb[l] = dZ.sum(axis=0)
        dH = dZ @ Ws[l]
    return loss, pack(gW, gb)


def batch(cfg, x, y):
    """Features and scaled targets of sample rows x (n, 2: rho, z), y (n, 3:
    phi, E_rho, E_z in physical units)."""
    u0, u1, env = coords(cfg, x[:, 0], x[:, 1])
    X0 = features(cfg, u0, u1)
    T = y / (np.asarray(cfg["out_scale"], dtype=np.float64)[None, :] * env[:, None])
    return X0, T


def predict(cfg, p, pos):
    """(phi, E_rho, E_z) in physical units at points (n, 3)."""
    pos = np.asarray(pos, dtype=np.float64)
    rho = np.hypot(pos[:, 0], pos[:, 1])
    u0, u1, env = coords(cfg, rho, pos[:, 2])
    Y, _ = forward(p, features(cfg, u0, u1))
    return Y[:, :3] * np.asarray(cfg["out_scale"], dtype=np.float64)[None, :] * env[:, None]


def adam(p, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (Kingma & Ba 2015) with bias correction, in place."""
    m[:] = b1 * m + (1 - b1) * g
    v[:] = b2 * v + (1 - b2) * g * g
    p -= lr * (m / (1 - b1**t)) / (np.sqrt(v / (1 - b2**t)) + eps)
