"""Axisymmetric BEM solve and perturbation-field evaluation on the device
(reference bem/axisymmetric.py:392-605)."""
import numpy as np

from .. import _lib


def solve_sigma(profile, E_gun, refine_ratio=1.0, max_depth=20, device=0):
    """Nodal density sigma with the solved potential equal to E_gun * z at
    every profile node (reference AxisymmetricBEMSolution.solve,
    bem/axisymmetric.py:411-443): device assembly + FP64 LU."""
    ctx = _lib.Context.get(device)
    prof, pp = _lib.host_ptr(np.asarray(profile, dtype=np.float64))
    if prof.ndim != 2 or prof.shape[1] != 2:
        raise ValueError("profile must be (n, 2)")
    sigma = np.empty(prof.shape[0])
    _, sp = _lib.host_ptr(sigma)
    ctx.check(ctx.lib.spt_bem_solve(ctx.h, pp, prof.shape[0], float(E_gun), float(refine_ratio), int(max_depth), sp),
              "spt_bem_solve")
    return sigma


def assemble(profile, refine_ratio=1.0, max_depth=20, device=0):
    """The collocation matrix (row-major), for validation."""
    ctx = _lib.Context.get(device)
    prof, pp = _lib.host_ptr(np.asarray(profile, dtype=np.float64))
    A = np.empty((prof.shape[0], prof.shape[0]))
    _, ap = _lib.host_ptr(A)
    ctx.check(ctx.lib.spt_bem_assemble(ctx.h, pp, prof.shape[0], float(refine_ratio), int(max_depth), ap),
              "spt_bem_assemble")
    return A


class AxisymmetricBEMSolution:
    """Profile + nodal sigma; ``field(points)`` is the reference's
    evaluate_axisymmetric_field on the device."""

    def __init__(self, profile, sigma, device=0):
        self.profile = np.asarray(profile, dtype=np.float64)
        self.sigma = np.asarray(sigma, dtype=np.float64)
        self.device = device

    @classmethod
    def solve(cls, profile, E_gun, refine_ratio=1.0, max_depth=20, device=0):
        return cls(profile, solve_sigma(profile, E_gun, refine_ratio, max_depth, device), device)

    def _upload(self, n_subdiv):
        ctx = _lib.Context.get(self.device)
        key = ("sol", _lib.token(self), self.profile.shape[0], n_subdiv)
        if ctx.bem_key != key:
            _, pp = _lib.host_ptr(self.profile)
            _, sp = _lib.host_ptr(self.sigma)
            ctx.check(ctx.lib.spt_bem_upload(ctx.h, pp, self.profile.shape[0], sp, int(n_subdiv)), "spt_bem_upload")
            ctx.bem_key = key
        return ctx

    def potential(self, points):
        """Perturbation potential at `points` (..., 3), shape points.shape[:-1]
        (reference bem/axisymmetric.py:445-462: adaptive per-segment
        quadrature), on the device."""
        import torch
        from ..fields import _to_device, _from_device

        ctx = self._upload(8)
        pos, meta = _to_device(points, ctx)
        out = torch.empty(pos.shape[0], dtype=torch.float64, device=ctx.torch_device)
        ctx.bind_stream()
        ctx.check(ctx.lib.spt_bem_potential(ctx.h, _lib.ptr(pos), pos.shape[0], _lib.ptr(out)), "spt_bem_potential")
        ctx.sync()
        return _from_device(out, meta, trailing=False)

    def field(self, points, n_subdiv=8, precision="fp64"):
        import torch
        from ..fields import _to_device, _from_device

        ctx = self._upload(n_subdiv)
        pos, meta = _to_device(points, ctx)
        out = torch.empty_like(pos)
        ctx.bind_stream()
        from ..geometry import _precision_code

        ctx.check(ctx.lib.spt_bem_pert_field(ctx.h, _lib.ptr(pos), pos.shape[0], _lib.ptr(out),
                                             _precision_code(precision)), "spt_bem_pert_field")
        ctx.sync()
        return _from_device(out, meta)
