"""ctypes binding of libspt_hip.so (declarations in include/spt.h).

Device buffers are torch tensors on the context's device; only their
data_ptr() crosses the boundary. Errors come back as negative codes and
are raised here as the reference's exception types: ValueError for bad
input, RuntimeError otherwise (no convergence, HIP failures).

There is no CPU fallback: if the library or a GPU is missing, every entry
point raises.
"""
import ctypes
import itertools
import os
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libspt_hip.so")

SPT_OK, SPT_EINVAL, SPT_ENOMEM, SPT_EHIP, SPT_ENOBEM, SPT_ESTEPCAP = 0, -1, -2, -3, -4, -5
SPT_MAX_SCREENS = 16
GEOM_FLAT, GEOM_TIP, GEOM_BEM_POST, GEOM_BEM_WELL = 0, 1, 2, 3
IMAGE_NONE, IMAGE_PLANE, IMAGE_TIP3, IMAGE_BEM, IMAGE_BEM_MIRROR = 0, 1, 2, 3, 4
STATUS_DONE, STATUS_STEPCAP, STATUS_STALLED = 0, 1, 2
MODE_PERSISTENT, MODE_LOCKSTEP = 0, 1
PREC_FP64, PREC_MIXED, PREC_PAIR32 = 0, 1, 2

EXPORTS = (
    "spt_create", "spt_destroy", "spt_last_error", "spt_set_stream", "spt_sync", "spt_device_cus",
    "spt_bem_upload", "spt_bem_assemble", "spt_bem_solve", "spt_bem_pert_field", "spt_geometry_field",
    "spt_pair_forces", "spt_kill_mask", "spt_trace", "spt_bem_potential", "spt_mlp_nparams", "spt_mlp_grad",
    "spt_mlp_train", "spt_mlp_infer", "spt_surrogate_set", "spt_toroidal_q", "spt_ring_modes",
    "spt_image_upload", "spt_image_force", "spt_set_field_shape", "spt_bem_far_rule",
    "spt_bem_multipole", "spt_trace_noalloc", "spt_bem_clusters", "spt_bem_near_width",
    "spt_set_lockstep_binning", "spt_bem_tiles", "spt_bem_tiles_enable", "spt_bem_remainder",
)
MLP_FOURIER, MLP_WIDTH, MLP_DEPTH = 32, 128, 4


class Geometry_t(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32), ("image", ctypes.c_int32),
        ("E_gun", ctypes.c_double), ("R", ctypes.c_double), ("H", ctypes.c_double),
        ("z0", ctypes.c_double), ("z_handoff", ctypes.c_double),
        ("has_kill_z", ctypes.c_int32), ("has_z_max", ctypes.c_int32),
        ("kill_z_below", ctypes.c_double), ("z_max", ctypes.c_double),
        ("bottom_fillet", ctypes.c_double), ("rim_fillet", ctypes.c_double),
    ]


class Event_t(ctypes.Structure):
    _fields_ = [
        ("group", ctypes.c_int32), ("particle", ctypes.c_int32), ("screen", ctypes.c_int32),
        ("interval", ctypes.c_int32), ("batch", ctypes.c_int64), ("t", ctypes.c_double),
        ("pos", ctypes.c_double * 3), ("vel", ctypes.c_double * 3),
    ]


EVENT_DTYPE = np.dtype([
    ("group", np.int32), ("particle", np.int32), ("screen", np.int32), ("interval", np.int32),
    ("batch", np.int64), ("t", np.float64), ("pos", np.float64, 3), ("vel", np.float64, 3),
])
assert EVENT_DTYPE.itemsize == ctypes.sizeof(Event_t) == 80

_p = ctypes.c_void_p


class MlpCfg_t(ctypes.Structure):
    _fields_ = [
        ("fourier", ctypes.c_float * (2 * 32)),
        ("center_z", ctypes.c_double), ("r_unit", ctypes.c_double),
        ("u_offset", ctypes.c_double * 2), ("u_scale", ctypes.c_double * 2),
        ("out_scale", ctypes.c_double * 3),
        ("env_r0", ctypes.c_double), ("env_pow", ctypes.c_double),
        ("rho_max", ctypes.c_double), ("z_min", ctypes.c_double), ("z_max", ctypes.c_double),
        ("cos_min", ctypes.c_double), ("band", ctypes.c_double),
    ]


class TraceArgs_t(ctypes.Structure):
    _fields_ = [
        ("geom", Geometry_t),
        ("n_groups", ctypes.c_int32), ("block_threads", ctypes.c_int32), ("n_particles", ctypes.c_int64),
        ("group_offsets", _p), ("group_order", _p),
        ("pos", _p), ("vel", _p), ("t_birth", _p), ("charge", _p), ("mass", _p), ("t_death", _p),
        ("q_over_m", ctypes.c_double), ("plummer_radius", ctypes.c_double), ("t_max", ctypes.c_double),
        ("rtol", ctypes.c_double), ("atol", ctypes.c_double),
        ("safety", ctypes.c_double), ("min_factor", ctypes.c_double), ("max_factor", ctypes.c_double),
        ("initial_step_fraction", ctypes.c_double),
        ("max_steps", ctypes.c_int64),
        ("n_screens", ctypes.c_int32), ("mode", ctypes.c_int32),
        ("precision", ctypes.c_int32), ("use_surrogate", ctypes.c_int32),
        ("screens", ctypes.c_double * SPT_MAX_SCREENS),
        ("attempts", _p), ("accepted", _p), ("active_stages", _p), ("margin", _p), ("status", _p),
        ("events", _p), ("event_capacity", ctypes.c_int64), ("event_count", _p),
        ("t_out", _p), ("n_out", ctypes.c_int32), ("pad1", ctypes.c_int32),
        ("snapshots", _p), ("snapshot_capacity", ctypes.c_int64), ("snapshot_count", _p),
        ("t_end", _p), ("field_order", _p),
    ]


_lib = None
_lock = threading.Lock()
_tokens = itertools.count(1)


def token(obj):
    """A process-unique number for obj, never reused (unlike id(), which
    CPython hands to the next object at the same address once obj is freed):
    the key under which a context remembers which object's tables it holds."""
    t = obj.__dict__.get("_spt_token")
    if t is None:
        t = obj.__dict__["_spt_token"] = next(_tokens)
    return t


def load():
    """Load (once) and return the ctypes library handle; raises if absent."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run __graft_entry__.build() (hipcc, gfx950)")
        # torch first: its HIP runtime (soname libamdhip64.so.7, like the
        # system one the library links) is then the one the library binds
        # to, so both work on the same device memory and streams; loaded the
        # other way round the two runtimes clash and the library sees no
        # device
        import torch  # noqa: F401

        lib = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER
        sig = {
            "spt_create": (ctypes.c_int, [ctypes.c_int, P(_p)]),
            "spt_destroy": (None, [_p]),
            "spt_last_error": (ctypes.c_char_p, [_p]),
            "spt_set_stream": (ctypes.c_int, [_p, _p]),
            "spt_sync": (ctypes.c_int, [_p]),
            "spt_device_cus": (ctypes.c_int, [_p, P(ctypes.c_int)]),
            "spt_bem_upload": (ctypes.c_int, [_p, _p, ctypes.c_int32, _p, ctypes.c_int32]),
            "spt_bem_far_rule": (ctypes.c_int, [_p, ctypes.c_double, ctypes.c_double]),
            "spt_bem_multipole": (ctypes.c_int, [_p, _p, ctypes.c_int32, ctypes.c_double, ctypes.c_double]),
            "spt_bem_clusters": (ctypes.c_int, [_p, _p, ctypes.c_int32, _p, _p, _p, ctypes.c_int32, ctypes.c_double,
                                                ctypes.c_double]),
            "spt_bem_near_width": (ctypes.c_int, [_p, ctypes.c_double]),
            "spt_bem_tiles": (ctypes.c_int, [_p, _p, ctypes.c_int32, _p, _p, ctypes.c_int32, ctypes.c_double,
                                             ctypes.c_double, ctypes.c_double]),
            "spt_bem_tiles_enable": (ctypes.c_int, [_p, ctypes.c_int32]),
            "spt_bem_remainder": (ctypes.c_int, [_p, _p, _p, ctypes.c_int64, _p]),
            "spt_bem_assemble": (ctypes.c_int, [_p, _p, ctypes.c_int32, ctypes.c_double, ctypes.c_int32, _p]),
            "spt_bem_solve": (ctypes.c_int, [_p, _p, ctypes.c_int32, ctypes.c_double, ctypes.c_double,
                                             ctypes.c_int32, _p]),
            "spt_bem_pert_field": (ctypes.c_int, [_p, _p, ctypes.c_int64, _p, ctypes.c_int32]),
            "spt_geometry_field": (ctypes.c_int, [_p, P(Geometry_t), _p, ctypes.c_int64, _p, ctypes.c_int32]),
            "spt_pair_forces": (ctypes.c_int, [_p, P(Geometry_t), _p, _p, _p, _p, ctypes.c_int32,
                                               ctypes.c_double, _p, _p, ctypes.c_int32]),
            "spt_kill_mask": (ctypes.c_int, [_p, P(Geometry_t), _p, ctypes.c_int64, _p]),
            "spt_trace": (ctypes.c_int, [_p, P(TraceArgs_t)]),
            "spt_trace_noalloc": (ctypes.c_int, [_p, P(TraceArgs_t)]),
            "spt_bem_potential": (ctypes.c_int, [_p, _p, ctypes.c_int64, _p]),
            "spt_mlp_nparams": (ctypes.c_int, [P(ctypes.c_int64)]),
            "spt_mlp_grad": (ctypes.c_int, [_p, P(MlpCfg_t), _p, _p, _p, ctypes.c_int32, _p, ctypes.c_int32, _p,
                                            _p]),
            "spt_mlp_train": (ctypes.c_int, [_p, P(MlpCfg_t), _p, _p, _p, ctypes.c_int64, _p, _p, ctypes.c_int64,
                                             ctypes.c_int32, ctypes.c_int32, ctypes.c_float, ctypes.c_uint64, _p,
                                             ctypes.c_int32, _p]),
            "spt_mlp_infer": (ctypes.c_int, [_p, P(MlpCfg_t), _p, _p, ctypes.c_int64, ctypes.c_int32, _p]),
            "spt_surrogate_set": (ctypes.c_int, [_p, P(MlpCfg_t), _p]),
            "spt_toroidal_q": (ctypes.c_int, [_p, _p, ctypes.c_int64, ctypes.c_int32, _p, _p]),
            "spt_set_field_shape": (ctypes.c_int, [_p, ctypes.c_int32, ctypes.c_int32]),
            "spt_set_lockstep_binning": (ctypes.c_int, [_p, ctypes.c_int32]),
            "spt_image_upload": (ctypes.c_int, [_p, _p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                                ctypes.c_double, ctypes.c_double, _p]),
            "spt_image_force": (ctypes.c_int, [_p, P(Geometry_t), _p, _p, _p, _p, ctypes.c_int32, ctypes.c_double,
                                               _p]),
            "spt_ring_modes": (ctypes.c_int, [_p, _p, _p, _p, _p, ctypes.c_int64, ctypes.c_int32, _p, _p, _p, _p]),
        }
        for name, (res, args) in sig.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _lib = lib
        return lib


class Context:
    """One library context on one device (reference: the per-device cache
    of backend.array_cache_key, backend.py:49-51)."""

    _by_device = {}

    def __init__(self, device=0):
        import torch

        if not torch.cuda.is_available():
            raise RuntimeError("backend='mi355x' needs a visible AMD GPU (torch.cuda.is_available() is False)")
        self.lib = load()
        self.device = int(device)
        self.torch_device = torch.device("cuda", self.device)
        h = _p()
        rc = self.lib.spt_create(self.device, ctypes.byref(h))
        if rc != SPT_OK:
            raise RuntimeError(f"spt_create failed ({rc})")
        self.h = h
        self.bem_key = None
        self.tiles_key = None  # (tiles token, bem_key) of the uploaded tiles

    @classmethod
    def get(cls, device=0):
        ctx = cls._by_device.get(device)
        if ctx is None:
            ctx = cls._by_device[device] = Context(device)
        return ctx

    def check(self, rc, what):
        if rc == SPT_OK:
            return
        msg = self.lib.spt_last_error(self.h).decode(errors="replace")
        if rc in (SPT_EINVAL, SPT_ENOBEM):
            raise ValueError(f"{what}: {msg}")
        raise RuntimeError(f"{what} failed ({rc}): {msg}")

    def bind_stream(self):
        import torch

        s = torch.cuda.current_stream(self.torch_device)
        self.check(self.lib.spt_set_stream(self.h, _p(s.cuda_stream)), "spt_set_stream")

    def sync(self):
        self.check(self.lib.spt_sync(self.h), "spt_sync")

    def __del__(self):
        try:
            if getattr(self, "h", None):
                self.lib.spt_destroy(self.h)
        except Exception:
            pass


def ptr(t):
    """data_ptr of a contiguous tensor (or NULL for None)."""
    if t is None:
        return None
    if not t.is_contiguous():
        raise ValueError("device buffers must be contiguous")
    return _p(t.data_ptr())


def host_ptr(a):
    a = np.ascontiguousarray(a)
    return a, a.ctypes.data_as(_p)
