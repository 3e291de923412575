"""Oracle per-group adaptive Dormand-Prince 5(4) integrator (reference
``batched_rk45.py:30-401``), with per-group attempt/accept counters and the
error-norm margin added for parity tests; output times (``t_out``) are
restated too (reference ``batched_rk45.py:181-209,243-249,295-312``)."""
import numpy as np

from .constants import SPEED_OF_LIGHT

# Dormand & Prince (1980) 5(4) tableau (reference batched_rk45.py:30-50).
A = [
    [],
    [1 / 5],
    [3 / 40, 9 / 40],
    [44 / 45, -56 / 15, 32 / 9],
    [19372 / 6561, -25360 / 2187, 64448 / 6561, -212 / 729],
    [9017 / 3168, -355 / 33, 46732 / 5247, 49 / 176, -5103 / 18656],
]
B = [35 / 384, 0.0, 500 / 1113, 125 / 192, -2187 / 6784, 11 / 84]
E = [71 / 57600, 0.0, -71 / 16695, 71 / 1920, -17253 / 339200, 22 / 525, -1 / 40]


def attempt(accel, pos, vel, t, h, t_birth, t_death, gidx, rtol, atol, safety, fmin, fmax):
    """One attempt per running group (reference ``batched_rk45.py:326-401``):
    six stages plus an endpoint evaluation (no FSAL), RMS error over the
    6*E position/velocity components with velocity measured against c*atol,
    step factor 0.9*err^(-1/5) clipped to [0.2, 10]."""
    active = (t[:, None] >= t_birth) & (t[:, None] < t_death)
    kp, kv = [], []
    for i in range(6):
        ps, vs = pos, vel
        for j, a in enumerate(A[i]):
            ps = ps + (h * a)[:, None, None] * kp[j]
            vs = vs + (h * a)[:, None, None] * kv[j]
        kv.append(accel(ps, vs, active, gidx))
        kp.append(np.where(active[..., None], vs, 0.0))
    pos_new = pos + h[:, None, None] * sum(B[i] * kp[i] for i in range(6))
    vel_new = vel + h[:, None, None] * sum(B[i] * kv[i] for i in range(6))
    kv.append(accel(pos_new, vel_new, active, gidx))
    kp.append(np.where(active[..., None], vel_new, 0.0))
    err_p = h[:, None, None] * sum(E[i] * kp[i] for i in range(7))
    err_v = h[:, None, None] * sum(E[i] * kv[i] for i in range(7))
    sp = atol + rtol * np.maximum(np.abs(pos), np.abs(pos_new))
    sv = SPEED_OF_LIGHT * atol + rtol * np.maximum(np.abs(vel), np.abs(vel_new))
    ncomp = 2 * pos.shape[1] * 3
    sq = np.sum((err_p / sp) ** 2, axis=(1, 2)) + np.sum((err_v / sv) ** 2, axis=(1, 2))
    err = np.sqrt(sq / ncomp)
    err = np.where(err == 0.0, np.finfo(float).tiny, err)
    fac = np.clip(safety * err ** (-1.0 / 5), fmin, fmax)
    return pos_new, vel_new, t + h, err <= 1.0, h * fac, err


def integrate(accel, pos0, vel0, t_birth, t_max, rtol, atol, on_step=None, kill=None,
              safety=0.9, fmin=0.2, fmax=10.0, init_frac=1e-6, max_steps=10_000_000, output_times=None,
              on_output=None, raise_at_cap=True):
    """Reference ``batched_rk45.py:55-323``; output times (shared by all
    groups) clip each step's target like birth checkpoints and fire
    ``on_output(k, groups, t, pos, vel)`` when a group lands on the k-th,
    also at a group's start time (reference 181-209, 243-249, 295-312).

    Returns dict(pos, vel, t_death, attempts, accepted, margin, t) where
    attempts/accepted are per group and margin is the smallest |err - 1|
    seen by each group (how close any accept/reject decision came to
    flipping) and t each group's time after its last attempt. With raise_at_cap False, a run that reaches max_steps
    returns its state there instead of raising (for comparisons of capped
    traces)."""
    G, n, _ = pos0.shape
    pos = pos0.copy()
    vel = vel0.copy()
    t_death = np.full((G, n), np.inf)
    cps = np.sort(t_birth, axis=1)
    t = cps[:, 0].copy()
    h = (t_max - t) * init_frac
    have_out = output_times is not None
    if have_out:
        outs = np.sort(np.asarray(output_times, dtype=float))
        n_out = outs.shape[0]
        if on_output is not None:
            last = np.sum(outs[None, :] <= t[:, None], axis=1) - 1
            at_start = (last >= 0) & (outs[np.maximum(last, 0)] == t)
            if np.any(at_start):
                ai0 = np.nonzero(at_start)[0]
                for p in np.unique(last[ai0]):
                    sel = ai0[last[ai0] == p]
                    on_output(int(p), sel, t[sel], pos[sel], vel[sel])
    running = np.ones(G, dtype=bool)
    attempts = np.zeros(G, dtype=np.int64)
    accepted_n = np.zeros(G, dtype=np.int64)
    margin = np.full(G, np.inf)
    idx = np.nonzero(running)[0]
    for it in range(max_steps):
        if idx.shape[0] == 0:
            break
        n_reached = np.sum(cps[idx] <= t[idx][:, None], axis=1)
        cp_idx = np.minimum(n_reached, n - 1)
        exhausted = n_reached >= n
        nxt = np.where(exhausted, t_max, cps[idx, cp_idx])
        target = np.minimum(nxt, t_max)
        if have_out:
            optr = np.sum(outs[None, :] <= t[idx][:, None], axis=1)
            oexh = optr >= n_out
            nout = np.where(oexh, t_max, outs[np.minimum(optr, n_out - 1)])
            target = np.minimum(target, nout)
        h_r = np.minimum(h[idx], target - t[idx])
        pn, vn, tn, acc, hn, err = attempt(accel, pos[idx], vel[idx], t[idx], h_r, t_birth[idx],
                                          t_death[idx], idx, rtol, atol, safety, fmin, fmax)
        attempts[idx] += 1
        margin[idx] = np.minimum(margin[idx], np.abs(err - 1.0))
        h[idx] = hn
        if np.any(acc):
            ai = idx[acc]
            accepted_n[ai] += 1
            t_old, p_old, v_old = t[ai], pos[ai], vel[ai]
            was_active = (t_old[:, None] >= t_birth[ai]) & (t_old[:, None] < t_death[ai])
            pos[ai] = pn[acc]
            vel[ai] = vn[acc]
            t[ai] = tn[acc]
            tol = 1e-9 * max(t_max, 1e-30)
            reset = (~exhausted[acc]) & (t[ai] >= nxt[acc] - tol)
            if np.any(reset):
                h[ai] = np.where(reset, (t_max - t[ai]) * init_frac, h[ai])
            if kill is not None:
                dead = was_active & kill(pos[ai])
                if np.any(dead):
                    t_death[ai] = np.where(dead, t[ai][:, None], t_death[ai])
            if have_out and on_output is not None:
                hit = (~oexh[acc]) & (t[ai] >= nout[acc] - tol)
                if np.any(hit):
                    hl = np.nonzero(hit)[0]
                    pk = optr[acc][hl]
                    for p in np.unique(pk):
                        sel = ai[hl[pk == p]]
                        on_output(int(p), sel, t[sel], pos[sel], vel[sel])
            if on_step is not None:
                on_step(it, ai, t_old, p_old, v_old, t[ai], pos[ai], vel[ai])
            running[ai] = t[ai] < t_max * (1 - 1e-14)
        idx = np.nonzero(running)[0]
    else:
        if raise_at_cap:
            raise RuntimeError(f"batched RK45 did not converge within {max_steps} steps")
    return dict(pos=pos, vel=vel, t_death=t_death, attempts=attempts, accepted=accepted_n, margin=margin, t=t)
