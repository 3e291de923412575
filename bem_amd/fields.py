"""Analytic external fields (reference fields.py:16-187), evaluated by the
HIP library. ``evaluate`` accepts a NumPy array or torch tensor of shape
(..., 3) and returns the same kind."""
import numpy as np

from . import _lib


def _to_device(position, ctx):
    """(flat (n, 3) float64 contiguous device tensor, meta to undo)."""
    import torch

    is_np = not hasattr(position, "detach")
    t = torch.as_tensor(np.asarray(position, dtype=np.float64) if is_np else position,
                        dtype=torch.float64, device=ctx.torch_device)
    shape = tuple(t.shape)
    if len(shape) == 0 or shape[-1] != 3:
        raise ValueError("position must have shape (..., 3)")
    return t.reshape(-1, 3).contiguous(), (is_np, shape, None if is_np else position.device)


def _from_device(out, meta, trailing=True):
    is_np, shape, dev = meta
    out = out.reshape(shape if trailing else shape[:-1])
    if is_np:
        return out.cpu().numpy()
    return out.to(dev)


class _DeviceField:
    def __init__(self, geom_struct, ctx, precision=0):
        self.g = geom_struct
        self.ctx = ctx
        self.precision = precision

    def evaluate(self, position):
        import torch

        ctx = self.ctx
        pos, meta = _to_device(position, ctx)
        out = torch.empty_like(pos)
        ctx.bind_stream()
        ctx.check(ctx.lib.spt_geometry_field(ctx.h, self.g, _lib.ptr(pos), pos.shape[0], _lib.ptr(out),
                                             self.precision), "spt_geometry_field")
        ctx.sync()
        return _from_device(out, meta)


class GunField:
    """Uniform Ez for z >= 0, zero below (reference fields.py:16-79)."""

    def __init__(self, Ez, device=0):
        self.Ez = float(Ez)
        self.device = device

    def _struct(self):
        g = _lib.Geometry_t()
        g.kind = _lib.GEOM_FLAT
        g.E_gun = self.Ez
        return g

    def evaluate(self, position):
        return _DeviceField(self._struct(), _lib.Context.get(self.device)).evaluate(position)


class HemisphericalTipField:
    """Grounded sphere of radius R in a uniform field, restricted to the
    exterior of the upper hemisphere (reference fields.py:82-160)."""

    def __init__(self, Ez, R, device=0):
        self.Ez = float(Ez)
        self.R = float(R)
        self.device = device

    def _struct(self):
        g = _lib.Geometry_t()
        g.kind = _lib.GEOM_TIP
        g.E_gun = self.Ez
        g.R = self.R
        return g

    def evaluate(self, position):
        return _DeviceField(self._struct(), _lib.Context.get(self.device)).evaluate(position)
