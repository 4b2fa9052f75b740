"""Pitched-row helpers (drop-in for ``kernelkit/data.py`` of the reference).

A pitched array pads its last axis so every row starts on a 32-byte boundary
(``data.py:5-27``).  Works on numpy arrays and PyTorch tensors; for tensors
the pitched buffer is the view's ``_base`` (or the tensor itself).
"""
from __future__ import annotations

import numpy as np

_PITCH_BYTES = 32  # row alignment of pitched buffers, as in the reference


def _is_tensor(x) -> bool:
    try:
        import torch
    except ImportError:  # pragma: no cover
        return False
    return isinstance(x, torch.Tensor)


def _itemsize(array) -> int:
    return array.element_size() if _is_tensor(array) else array.dtype.itemsize


def pitched_shape(array) -> tuple:
    """Shape with the last axis rounded up to a 32-byte multiple."""
    item = _itemsize(array)
    nbytes = -(-array.shape[-1] * item // _PITCH_BYTES) * _PITCH_BYTES
    if nbytes % item:
        raise ValueError("item size does not divide the pitch")
    return (*array.shape[:-1], nbytes // item)


def _base(array):
    if _is_tensor(array):
        return array._base if array._base is not None else array
    return array.base if array.base is not None else array


def ispitched(array) -> bool:
    """True if the array (or its base) has a pitched shape."""
    base = _base(array)
    return tuple(pitched_shape(base)) == tuple(base.shape)


def aspitched(array, xp=None):
    """Copy into a zero-padded pitched buffer; return the unpadded view.

    ``xp`` selects the output kind: ``None`` (same as the input), ``np`` /
    ``"numpy"`` for host memory, or ``"torch"``/``"cuda"`` for device
    memory.
    """
    want_dev = _want_device(array, xp)
    if ispitched(array) and _is_tensor(array) == want_dev:
        return array
    if want_dev:
        import torch
        from kernelkit.kernel import as_device_array
        src = as_device_array(array)
        buf = torch.zeros(pitched_shape(src), dtype=src.dtype,
                          device=src.device)
        buf[..., :src.shape[-1]] = src
        return buf[..., :src.shape[-1]]
    src = array.detach().cpu().numpy() if _is_tensor(array) \
        else np.asarray(array)
    buf = np.zeros(pitched_shape(src), dtype=src.dtype)
    buf[..., :src.shape[-1]] = src
    return buf[..., :src.shape[-1]]


def pitched_like(array, xp=None, fill=None):
    """New pitched buffer shaped like ``array``; returns the unpadded view,
    optionally filled with ``fill`` (``data.py:82-115``)."""
    if not ispitched(array):
        raise ValueError("Input array needs to be pitched.")
    want_dev = _want_device(array, xp)
    shape = pitched_shape(array)
    if want_dev:
        import torch
        from kernelkit.kernel import default_device
        dtype = array.dtype if _is_tensor(array) else torch.float32
        base = torch.empty(shape, dtype=dtype, device=default_device())
    else:
        dtype = array.dtype if not _is_tensor(array) else np.float32
        base = np.empty(shape, dtype=dtype)
    view = base[..., :array.shape[-1]]
    if fill is not None:
        view[...] = fill
    return view


def _want_device(array, xp) -> bool:
    if xp is None:
        return _is_tensor(array)
    if xp is np or xp == "numpy":
        return False
    return True
