"""Device arrays with or without PyTorch.

The reference computes on CuPy arrays (``kernelkit/kernel.py:47-119``,
``projector.py:266-283``).  The gfx950 kernels here take raw device
pointers and element strides, so any device array serves:

* PyTorch-ROCm tensors (``torch.Tensor`` on the HIP device);
* any object exposing ``__cuda_array_interface__`` (CuPy, Numba, ...), used
  in place through a non-owning :class:`DeviceArray` view;
* :class:`DeviceArray`, this package's own buffer type: device memory from
  the library's ``kk_malloc`` (no PyTorch involved).

Arrays the package allocates for a call (outputs, textures, workspaces)
are of the kind of the call's input: torch tensors for torch inputs (so
they come from torch's stream-ordered allocator and work under
``torch.cuda.graph``), :class:`DeviceArray` otherwise.  Host inputs become
torch tensors when PyTorch is already loaded in the process, else
:class:`DeviceArray`; this module never imports torch itself.
"""
from __future__ import annotations

import ctypes
import sys

import numpy as np

from kernelkit import _native

_ELEM = {np.dtype(np.float32): 4, np.dtype(np.int64): 8}


def _torch_module():
    """The torch module if PyTorch is loaded in this process, else None."""
    return sys.modules.get("torch")


def is_torch(x) -> bool:
    torch = _torch_module()
    return torch is not None and isinstance(x, torch.Tensor)


def dtype_name(x) -> str:
    """'float32', 'int64', ... for numpy, torch and device arrays."""
    return str(x.dtype).replace("torch.", "")


class Device:
    """A HIP device, as ``torch.device('cuda', index)`` (``.type``,
    ``.index``)."""

    type = "cuda"

    def __init__(self, index: int):
        self.index = int(index)

    def __eq__(self, other):
        return getattr(other, "type", None) == "cuda" and \
            getattr(other, "index", None) == self.index

    def __hash__(self):
        return hash(("cuda", self.index))

    def __repr__(self):
        return f"Device(cuda:{self.index})"


def current_device() -> int:
    """The current HIP device index of the calling thread."""
    d = ctypes.c_int32(0)
    _native.check(_native.lib().kk_get_device(ctypes.byref(d)))
    return int(d.value)


def device_count() -> int:
    n = ctypes.c_int32(0)
    status = _native.lib().kk_device_count(ctypes.byref(n))
    return int(n.value) if status == _native.KK_OK else 0


class _Buffer:
    """One ``kk_malloc`` allocation, freed with the last array using it."""

    def __init__(self, nbytes: int):
        p = ctypes.c_void_p()
        _native.check(_native.lib().kk_malloc(ctypes.byref(p), int(nbytes)))
        self.ptr = int(p.value or 0)
        self.nbytes = int(nbytes)

    def __del__(self):
        ptr = getattr(self, "ptr", 0)
        if ptr and _native._lib is not None:
            try:
                _native._lib.kk_free(ctypes.c_void_p(ptr))
            except Exception:  # interpreter shutdown
                pass
            self.ptr = 0


def _c_strides(shape):
    st, acc = [], 1
    for s in reversed(shape):
        st.append(acc)
        acc *= max(int(s), 1)
    return tuple(reversed(st))


def _as3(shape, strides):
    """Shape and strides padded (on the left) to three axes."""
    if len(shape) > 3:
        raise ValueError("device copies handle up to three axes")
    pad = 3 - len(shape)
    return (1,) * pad + tuple(shape), (0,) * pad + tuple(strides)


class DeviceArray:
    """A strided device array of float32 or int64 elements.

    Owns a ``kk_malloc`` buffer, or views one (slices, reshapes, the
    ``__cuda_array_interface__`` of another library's array).  The methods
    are the subset of a torch tensor's that this package uses, so kernels
    and projectors take either.  Copies and fills run on the null stream
    (or ``stream=``); :meth:`get` synchronises it.
    """

    is_cuda = True

    def __init__(self, shape, dtype=np.float32, *, ptr=None, strides=None,
                 owner=None, device=None, base=None):
        self.shape = tuple(int(s) for s in shape)
        self.dtype = np.dtype(dtype)
        if self.dtype not in _ELEM:
            raise NotImplementedError(f"dtype {self.dtype} is not supported")
        self._strides = (_c_strides(self.shape) if strides is None
                         else tuple(int(s) for s in strides))
        if len(self._strides) != len(self.shape):
            raise ValueError("shape and strides differ in length")
        if ptr is None:
            cur = current_device()
            device = cur if device is None else int(device)
            n = int(np.prod(self.shape)) if self.shape else 1
            if device != cur:
                _native.check(_native.lib().kk_set_device(device))
            try:
                owner = _Buffer(n * _ELEM[self.dtype])
            finally:
                if device != cur:
                    _native.check(_native.lib().kk_set_device(cur))
            ptr = owner.ptr
        # a view of foreign memory learns its device when asked
        self._device = None if device is None else Device(device)
        self._ptr = int(ptr)
        self._owner = owner      # keeps the memory alive
        self.base = base         # the DeviceArray this one views, or None

    @property
    def device(self) -> Device:
        if self._device is None:
            self._device = Device(current_device())
        return self._device

    def _dev_index(self):
        return None if self._device is None else self._device.index

    # -- the tensor-like subset -------------------------------------------
    @property
    def ndim(self) -> int:
        return len(self.shape)

    def dim(self) -> int:
        return len(self.shape)

    def numel(self) -> int:
        return int(np.prod(self.shape)) if self.shape else 1

    @property
    def size(self) -> int:
        return self.numel()

    def element_size(self) -> int:
        return _ELEM[self.dtype]

    def stride(self, k=None):
        return self._strides if k is None else self._strides[k]

    def data_ptr(self) -> int:
        return self._ptr

    def is_contiguous(self) -> bool:
        return all(a == b or s == 1 for a, b, s in
                   zip(self._strides, _c_strides(self.shape), self.shape))

    @property
    def __cuda_array_interface__(self) -> dict:
        item = _ELEM[self.dtype]
        return {"shape": self.shape,
                "typestr": self.dtype.str,
                "data": (self._ptr, False),
                "strides": (None if self.is_contiguous() else
                            tuple(s * item for s in self._strides)),
                "version": 3}

    def __len__(self):
        return self.shape[0]

    def __repr__(self):
        return (f"DeviceArray(shape={self.shape}, dtype={self.dtype}, "
                f"strides={self._strides}, device={self._dev_index()})")

    def __getitem__(self, idx):
        """Basic indexing (integers, slices with positive steps,
        ``...``): a view."""
        if not isinstance(idx, tuple):
            idx = (idx,)
        if any(i is Ellipsis for i in idx):
            k = idx.index(Ellipsis)
            fill = self.ndim - (len(idx) - 1)
            idx = idx[:k] + (slice(None),) * fill + idx[k + 1:]
        if len(idx) > self.ndim:
            raise IndexError("too many indices")
        idx = idx + (slice(None),) * (self.ndim - len(idx))
        off, shape, strides = 0, [], []
        for i, n, st in zip(idx, self.shape, self._strides):
            if isinstance(i, slice):
                a, b, step = i.indices(n)
                if step <= 0:
                    raise IndexError("only positive slice steps")
                cnt = max(0, (b - a + step - 1) // step)
                off += a * st
                shape.append(cnt)
                strides.append(st * step)
            else:
                i = int(i)
                if i < 0:
                    i += n
                if not 0 <= i < n:
                    raise IndexError("index out of range")
                off += i * st
        return DeviceArray(shape, self.dtype,
                           ptr=self._ptr + off * _ELEM[self.dtype],
                           strides=strides, owner=self._owner,
                           device=self._dev_index(),
                           base=self.base if self.base is not None else self)

    def view(self, *shape):
        """Reshape of a contiguous array (one -1 allowed)."""
        if len(shape) == 1 and isinstance(shape[0], (tuple, list)):
            shape = tuple(shape[0])
        if not self.is_contiguous():
            raise ValueError("view of a non-contiguous array")
        shape = list(int(s) for s in shape)
        if shape.count(-1) == 1:
            known = int(np.prod([s for s in shape if s != -1]))
            shape[shape.index(-1)] = self.numel() // max(known, 1)
        if int(np.prod(shape)) != self.numel():
            raise ValueError(f"cannot view {self.shape} as {tuple(shape)}")
        return DeviceArray(shape, self.dtype, ptr=self._ptr,
                           owner=self._owner, device=self._dev_index(),
                           base=self.base if self.base is not None else self)

    reshape = view

    def expand(self, *shape):
        """Broadcast view (stride 0 along the broadcast axes)."""
        if len(shape) == 1 and isinstance(shape[0], (tuple, list)):
            shape = tuple(shape[0])
        shape = tuple(int(s) for s in shape)
        pad = len(shape) - self.ndim
        if pad < 0:
            raise ValueError("cannot broadcast to fewer axes")
        src = (1,) * pad + self.shape
        sst = (0,) * pad + self._strides
        strides = []
        for n, m, st in zip(shape, src, sst):
            if m == n:
                strides.append(st)
            elif m == 1:
                strides.append(0)
            else:
                raise ValueError(f"cannot broadcast {self.shape} to {shape}")
        return DeviceArray(shape, self.dtype, ptr=self._ptr, strides=strides,
                           owner=self._owner, device=self._dev_index(),
                           base=self.base if self.base is not None else self)

    def copy_(self, src, stream=None):
        """Copy ``src`` (a host array, or a device array of the same
        shape) into this array."""
        if isinstance(src, np.ndarray) or not _is_device(src):
            host = np.ascontiguousarray(np.broadcast_to(
                np.asarray(src, dtype=self.dtype), self.shape))
            if self.is_contiguous():
                _native.check(_native.lib().kk_memcpy(
                    ctypes.c_void_p(self._ptr),
                    host.ctypes.data_as(ctypes.c_void_p), host.nbytes,
                    _native.KK_COPY_H2D, _native.stream_ptr(stream)))
                _native.check(_native.lib().kk_stream_synchronize(
                    _native.stream_ptr(stream)))
                return self
            tmp = DeviceArray(self.shape, self.dtype)
            tmp.copy_(host, stream)
            return self.copy_(tmp, stream)
        s = asarray(src)
        if tuple(s.shape) != self.shape:
            raise ValueError(f"copy of {tuple(s.shape)} into {self.shape}")
        if dtype_name(s) != self.dtype.name:
            raise ValueError(f"copy of {dtype_name(s)} into {self.dtype}")
        shape3, d3 = _as3(self.shape, self._strides)
        _, s3 = _as3(tuple(s.shape), tuple(s.stride()))
        _native.check(_native.lib().kk_copy_strided(
            ctypes.c_void_p(s.data_ptr()), _native.i64(s3),
            ctypes.c_void_p(self._ptr), _native.i64(d3), _native.i64(shape3),
            _ELEM[self.dtype], _native.stream_ptr(stream)))
        return self

    def fill_(self, value, stream=None):
        if self.dtype != np.float32:
            return self.copy_(np.full(self.shape, value, self.dtype), stream)
        if self.is_contiguous():
            _native.check(_native.lib().kk_fill(
                ctypes.c_void_p(self._ptr), self.numel(), float(value),
                _native.stream_ptr(stream)))
            return self
        tmp = DeviceArray(self.shape, self.dtype)
        tmp.fill_(value, stream)
        return self.copy_(tmp, stream)

    def zero_(self, stream=None):
        return self.fill_(0.0, stream)

    def clone(self, stream=None):
        out = DeviceArray(self.shape, self.dtype, device=self.device.index)
        return out.copy_(self, stream)

    def contiguous(self):
        return self if self.is_contiguous() else self.clone()

    def get(self, stream=None) -> np.ndarray:
        """Copy to a new host array (synchronises the stream)."""
        src = self.contiguous()
        out = np.empty(self.shape, dtype=self.dtype)
        _native.check(_native.lib().kk_memcpy(
            out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(src._ptr),
            out.nbytes, _native.KK_COPY_D2H, _native.stream_ptr(stream)))
        _native.check(_native.lib().kk_stream_synchronize(
            _native.stream_ptr(stream)))
        return out

    numpy = get


def _is_device(x) -> bool:
    if isinstance(x, DeviceArray):
        return True
    if is_torch(x):
        return bool(x.is_cuda)
    return hasattr(x, "__cuda_array_interface__")


def asarray(x):
    """A device array for ``x`` without copying: torch tensors and
    :class:`DeviceArray` as they are, other ``__cuda_array_interface__``
    arrays as a :class:`DeviceArray` view (the caller keeps the original
    alive)."""
    if isinstance(x, DeviceArray) or is_torch(x):
        return x
    cai = getattr(x, "__cuda_array_interface__", None)
    if cai is None:
        raise TypeError(f"not a device array: {type(x).__name__}")
    dtype = np.dtype(cai["typestr"])
    if dtype not in _ELEM:
        raise NotImplementedError(f"dtype {dtype} is not supported")
    shape = tuple(int(s) for s in cai["shape"])
    item = dtype.itemsize
    strides = cai.get("strides")
    strides = (None if strides is None
               else tuple(int(s) // item for s in strides))
    return DeviceArray(shape, dtype, ptr=int(cai["data"][0]),
                       strides=strides, owner=x)


def from_host(a, dtype=np.float32) -> DeviceArray:
    """A new contiguous :class:`DeviceArray` holding a copy of host data."""
    a = np.ascontiguousarray(a, dtype=dtype)
    out = DeviceArray(a.shape, a.dtype)
    return out.copy_(a)


def empty(shape, like=None, dtype="float32"):
    """An uninitialised device array: a torch tensor when ``like`` is one
    (same device), else a :class:`DeviceArray`."""
    shape = tuple(int(s) for s in shape)
    if is_torch(like):
        torch = _torch_module()
        return torch.empty(shape, dtype=getattr(torch, dtype),
                           device=like.device)
    dev = asarray(like).device.index if _is_device(like) else None
    return DeviceArray(shape, np.dtype(dtype), device=dev)


def zeros(shape, like=None, dtype="float32"):
    out = empty(shape, like, dtype)
    out.zero_()
    return out


def full(shape, value, like=None, dtype="float32"):
    out = empty(shape, like, dtype)
    out.fill_(value)
    return out


def stack(arrays):
    """A new array holding ``arrays`` (device arrays of one shape) along a
    new first axis."""
    arrays = [asarray(a) for a in arrays]
    if is_torch(arrays[0]):
        return _torch_module().stack(arrays)
    out = DeviceArray((len(arrays),) + tuple(arrays[0].shape),
                      arrays[0].dtype)
    for k, a in enumerate(arrays):
        out[k].copy_(a)
    return out


def broadcast_to(x, shape):
    if is_torch(x):
        return _torch_module().broadcast_to(x, tuple(shape))
    return asarray(x).expand(tuple(shape))


def to_host(x) -> np.ndarray:
    """A host copy (numpy) of a device array."""
    if is_torch(x):
        return x.detach().cpu().numpy()
    return asarray(x).get()
