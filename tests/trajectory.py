"""Comparing two integrations of the same trajectories at a common time.

The reference's first step is h0 = 1e-6 (t_max - t0) (batched_rk45.py:
177-178). At that size the embedded error estimate h sum_i E_i k_i is a
near-total cancellation of seven almost equal stage accelerations, so it
carries the field's last-bit rounding: two correct FP64 implementations
that sum the BEM field in different orders get error norms that differ by
~1e-4 relative in the first attempts and hence step sizes that differ by
~1e-5 relative. Accept/reject decisions (margins far from 1) and attempt
counts are unaffected, the trajectories agree to the integration
tolerance, but a particle killed at the end of the step that crosses
z_max (reference batched_rk45.py:290-293), or a run stopped after a fixed
number of attempts, is frozen at an end-of-step time that moves with the
step sizes. So the final states are compared at the device's time: the
oracle's state is carried to it by its own velocity and acceleration,
x + v dt + a dt^2 / 2 and v + a dt, with |dt| / t bounded."""
import numpy as np


def carried(pos, vel, acc, dt):
    """(pos, vel) advanced by dt (per particle) with constant acceleration."""
    dt = np.asarray(dt)[..., None]
    return pos + vel * dt + 0.5 * acc * dt * dt, vel + acc * dt


def assert_same_state(pos, vel, ref_pos, ref_vel, ref_acc, dt, t_ref, rtol=1e-9, dt_rtol=1e-3):
    """Device (pos, vel) against the reference state carried by dt."""
    dt = np.asarray(dt, dtype=float)
    t_ref = np.broadcast_to(np.asarray(t_ref, dtype=float), dt.shape)
    nz = t_ref > 0
    assert np.all(np.abs(dt[nz]) <= dt_rtol * t_ref[nz]), np.max(np.abs(dt[nz]) / t_ref[nz])
    p, v = carried(ref_pos, ref_vel, ref_acc, dt)
    np.testing.assert_allclose(pos, p, rtol=rtol, atol=rtol * np.abs(ref_pos).max())
    np.testing.assert_allclose(vel, v, rtol=rtol, atol=rtol * np.abs(ref_vel).max())
