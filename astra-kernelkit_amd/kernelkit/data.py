"""Pitched-row helpers (drop-in for ``kernelkit/data.py`` of the reference).

A pitched array pads its last axis so every row starts on a 32-byte boundary
(``data.py:5-27``).  Works on numpy arrays and device arrays (torch tensors,
:class:`~kernelkit.DeviceArray`); the pitched buffer is the view's base
(``base`` for numpy and device arrays, ``_base`` for tensors) or the array
itself.
"""
from __future__ import annotations

import numpy as np

from kernelkit import device as _dev

_PITCH_BYTES = 32  # row alignment of pitched buffers, as in the reference


def _on_device(x) -> bool:
    return not isinstance(x, np.ndarray) and _dev._is_device(x)


def _itemsize(array) -> int:
    if _on_device(array):
        return _dev.asarray(array).element_size()
    return np.dtype(array.dtype).itemsize


def pitched_shape(array) -> tuple:
    """Shape with the last axis rounded up to a 32-byte multiple."""
    item = _itemsize(array)
    nbytes = -(-array.shape[-1] * item // _PITCH_BYTES) * _PITCH_BYTES
    if nbytes % item:
        raise ValueError("item size does not divide the pitch")
    return (*array.shape[:-1], nbytes // item)


def _base(array):
    if _dev.is_torch(array):
        return array._base if array._base is not None else array
    base = getattr(array, "base", None)
    return base if base is not None else array


def ispitched(array) -> bool:
    """True if the array (or its base) has a pitched shape."""
    base = _base(array)
    return tuple(pitched_shape(base)) == tuple(base.shape)


def aspitched(array, xp=None):
    """Copy into a zero-padded pitched buffer; return the unpadded view.

    ``xp`` selects the output kind: ``None`` (same as the input), ``np`` /
    ``"numpy"`` for host memory, or ``"cuda"`` (or ``"torch"``) for device
    memory (of the input's kind for device inputs; a torch tensor for host
    inputs when PyTorch is loaded, else a :class:`~kernelkit.DeviceArray`).
    """
    want_dev = _want_device(array, xp)
    if ispitched(array) and _on_device(array) == want_dev:
        return array
    if want_dev:
        from kernelkit.kernel import as_device_array
        src = as_device_array(array)
        buf = _dev.zeros(pitched_shape(src), like=src)
        view = buf[..., :src.shape[-1]]
        view.copy_(src)
        return view
    src = _dev.to_host(array) if _on_device(array) else np.asarray(array)
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
        from kernelkit.kernel import default_device
        default_device()
        like = array if _on_device(array) else None
        torch = _dev._torch_module()
        if like is None and torch is not None:
            base = torch.empty(shape, dtype=torch.float32,
                               device=default_device())
        else:
            base = _dev.empty(shape, like=like)
    else:
        dtype = array.dtype if not _on_device(array) else np.float32
        base = np.empty(shape, dtype=dtype)
    view = base[..., :array.shape[-1]]
    if fill is not None:
        if _on_device(view):
            view.fill_(fill)
        else:
            view[...] = fill
    return view


def _want_device(array, xp) -> bool:
    if xp is None:
        return _on_device(array)
    if xp is np or xp == "numpy":
        return False
    return True
