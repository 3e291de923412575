"""Backend selection (reference backend.py:18-51).

This package implements one backend, 'mi355x': the HIP library on an AMD
Instinct GPU, FP64. The reference's 'cpu' (NumPy) and 'gpu' (CuPy) names
stay with the reference; asking this package for them is a ValueError, as
any unknown name is there."""
import numpy as np

_BACKENDS = {"mi355x": "hip"}


def resolve_backend(name, device=0):
    """Return (context, dtype) for a backend name."""
    if name not in _BACKENDS:
        raise ValueError(f"Unknown backend {name!r}; choose from {sorted(_BACKENDS)}")
    from ._lib import Context

    return Context.get(device), np.dtype("float64")


def to_numpy(array):
    """Bring a torch tensor (any device) or array-like back to NumPy."""
    if hasattr(array, "detach"):
        return array.detach().cpu().numpy()
    return np.asarray(array)
