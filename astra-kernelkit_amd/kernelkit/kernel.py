"""Kernel runtime: base class, exceptions and padded textures.

Mirrors ``kernelkit/kernel.py`` of the reference.  There, a kernel is a
Jinja2-rendered CUDA source compiled by NVRTC (``:188-226``), geometry is
copied into ``__constant__`` memory (``:122-147``) and inputs are sampled
through CUDA texture objects with border addressing (``:22-32``, ``:47-119``).

Here kernels are ahead-of-time gfx950 code in ``libkernelkit_hip.so``:
``compile()`` loads the library and creates a plan that owns the per-angle
device tables.  The HIP kernels sample gather layouts built from the arrays
(:class:`VolumeTexture`, :class:`ProjectionTexture`: zero-halo float2 pairs,
so one 16-byte load fetches the four bilinear taps) with explicit float32
interpolation and no per-tap bounds tests; :class:`Texture` is the plain
snapshot that ``copy_to_texture`` returns.
"""
from __future__ import annotations

from abc import ABC, abstractmethod
from dataclasses import dataclass

import numpy as np

from kernelkit import _native
from kernelkit import device as _dev


class KernelNotCompiledException(Exception):
    """Raised when a kernel is called before it has been compiled."""


class KernelMemoryUnpreparedException(Exception):
    """Raised when a kernel is called before memory has been prepared."""


def _no_device():
    return _native.NativeLibraryError(
        "no HIP device available: the cone-beam operators run on the GPU "
        "only")


def default_device():
    """The HIP device for new arrays: torch's current device when PyTorch
    is loaded in the process, else the library's current device."""
    torch = _dev._torch_module()
    if torch is not None:
        if not torch.cuda.is_available():
            raise _no_device()
        return torch.device("cuda", torch.cuda.current_device())
    if _dev.device_count() == 0:
        raise _no_device()
    return _dev.Device(_dev.current_device())


def is_device_array(x) -> bool:
    """A torch tensor on the HIP device, a :class:`~kernelkit.DeviceArray`
    or any array exposing ``__cuda_array_interface__``."""
    return _dev._is_device(x)


def _is_f32(x) -> bool:
    return _dev.dtype_name(x) == "float32"


def as_device_array(x, dtype=None, copy: bool = False):
    """Float32 device array for ``x``.

    Device arrays are used in place (``__cuda_array_interface__`` arrays
    through a :class:`~kernelkit.DeviceArray` view; a torch tensor of
    another dtype is converted, as torch does).  Host data is copied: into
    a torch tensor when PyTorch is loaded in the process, else into a
    :class:`~kernelkit.DeviceArray`.
    """
    if _dev.is_torch(x):
        torch = _dev._torch_module()
        dtype = dtype or torch.float32
        if x.is_cuda and x.dtype == dtype and not copy:
            return x
        return x.to(device=default_device() if not x.is_cuda else x.device,
                    dtype=dtype, copy=copy)
    if _dev._is_device(x):
        a = _dev.asarray(x)
        if not _is_f32(a):
            raise ValueError(
                f"Array must be of type float32, got {_dev.dtype_name(a)}.")
        return a.clone() if copy else a
    arr = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    torch = _dev._torch_module()
    if torch is not None and torch.cuda.is_available():
        return torch.from_numpy(arr).to(default_device())
    if _dev.device_count() == 0:
        raise _no_device()
    return _dev.from_host(arr)


def _unsqueeze0(a):
    if _dev.is_torch(a):
        return a.unsqueeze(0)
    return _dev.DeviceArray((1,) + tuple(a.shape), a.dtype,
                            ptr=a.data_ptr(), strides=(0,) + a.stride(),
                            owner=a, device=a.device.index)


def _alloc(n: int, like=None, device=None):
    """``n`` floats of device memory of the kind of ``like`` (or of the
    torch ``device`` given)."""
    if like is None and device is not None and not isinstance(
            device, _dev.Device):
        torch = _dev._torch_module()
        return torch.empty(int(n), dtype=torch.float32, device=device)
    return _dev.empty((int(n),), like=like)


class Texture:
    """Snapshot of an array, the analogue of a CUDA array behind a texture
    object (``kernel.py:47-119``): a contiguous float32 device copy in the
    array's own axis order.  Kernels convert it to their gather layout
    (:class:`VolumeTexture` / :class:`ProjectionTexture`) on use.
    """

    def __init__(self, array, layered: bool = False):
        dev = as_device_array(array)
        if not _is_f32(dev):
            raise ValueError(
                f"Array must be of type float32, got {dev.dtype} instead.")
        if dev.dim() == 2:
            dev = _unsqueeze0(dev)
        self.array = dev.contiguous().clone()
        self.shape = tuple(self.array.shape)
        self.layered = bool(layered)
        self.version = 0

    def copy_from(self, array) -> None:
        """Refresh the snapshot with new data of the same shape."""
        dev = as_device_array(array)
        if dev.dim() == 2:
            dev = _unsqueeze0(dev)
        if tuple(dev.shape) != self.shape:
            raise ValueError(f"array shape {tuple(dev.shape)} does not "
                             f"match texture shape {self.shape}")
        self.array.copy_(dev)
        self.version += 1


class VolumeTexture:
    """Forward-projection layout of a volume: one slice-major frame of
    (value, neighbour) float2 pairs with a zero halo per ray axis in
    ``mask`` (bit a = world axis a); see ``include/kernelkit_hip.h``."""

    def __init__(self, world_shape, mask: int, device=None, like=None):
        self.world_shape = tuple(int(s) for s in world_shape)
        self.mask = int(mask)
        n = _native.lib().kk_fp_texture_elems(_native.i32(self.world_shape),
                                              self.mask)
        if n < 0:
            raise ValueError(f"bad volume shape {self.world_shape} or frame "
                             f"mask {mask}")
        self.data = _alloc(n, like, device)

    @property
    def ptr(self):
        return self.data.data_ptr()

    def covers(self, mask: int) -> bool:
        return (int(mask) & ~self.mask) == 0

    def update(self, volume, volume_axes=(0, 1, 2), stream=None) -> None:
        """Copy ``volume`` (array axis k = world axis ``volume_axes[k]``)."""
        volume = as_device_array(volume)
        if not _is_f32(volume):
            raise ValueError(
                f"Volume must be of type float32, got {volume.dtype}.")
        wshape = [0, 0, 0]
        wstride = [0, 0, 0]
        for k, ax in enumerate(volume_axes):
            wshape[ax] = volume.shape[k]
            wstride[ax] = volume.stride(k)
        if tuple(wshape) != self.world_shape:
            raise ValueError(f"volume of world shape {tuple(wshape)} does "
                             f"not match texture shape {self.world_shape}")
        _native.check(_native.lib().kk_volume_to_fp_texture(
            _native._vp(volume.data_ptr()), _native.i32(self.world_shape),
            _native.i64(wstride), self.mask, _native._vp(self.ptr),
            _native.stream_ptr(stream)))


class ProjectionTexture:
    """Backprojection layout of projections: ``[angle][rows+3][cols+3]`` of
    (value, next-row) float2 pairs with a zero halo; see
    ``include/kernelkit_hip.h``."""

    def __init__(self, n_angles: int, rows: int, cols: int, device=None,
                 like=None):
        self.shape = (int(n_angles), int(rows), int(cols))
        n = _native.lib().kk_bp_texture_elems(*self.shape)
        if n < 0:
            raise ValueError(f"bad projection shape {self.shape}")
        self.data = _alloc(n, like, device)

    @property
    def ptr(self):
        return self.data.data_ptr()

    def zero(self, stream=None) -> None:
        _native.check(_native.lib().kk_fill(
            _native._vp(self.ptr), self.data.numel(), 0.0,
            _native.stream_ptr(stream)))

    def update(self, projections, projection_axes=(0, 1, 2),
               stream=None) -> None:
        """Copy projections (array axis k = logical axis
        ``projection_axes[k]`` of (angle, row, col))."""
        projections = as_device_array(projections)
        if not _is_f32(projections):
            raise ValueError("Projections must be of type float32, got "
                             f"{projections.dtype}.")
        lshape = [0, 0, 0]
        lstride = [0, 0, 0]
        for k, ax in enumerate(projection_axes):
            lshape[ax] = projections.shape[k]
            lstride[ax] = projections.stride(k)
        if tuple(lshape) != self.shape:
            raise ValueError(f"projections of shape {tuple(lshape)} do not "
                             f"match texture shape {self.shape}")
        _native.check(_native.lib().kk_projections_to_bp_texture(
            _native._vp(projections.data_ptr()), *self.shape,
            _native.i64(lstride), _native._vp(self.ptr),
            _native.stream_ptr(stream)))

    def values(self):
        """The stored projections as an (angle, row, col) strided view."""
        n, rows, cols = self.shape
        h = _native.HALO
        lo = _native.HALO_LO
        if _dev.is_torch(self.data):
            full = self.data.view(n, rows + h, cols + h, 2)
            return full[:, lo:lo + rows, lo:lo + cols, 0]
        st = (2 * (rows + h) * (cols + h), 2 * (cols + h), 2)
        off = 2 * (lo * (cols + h) + lo)
        return _dev.DeviceArray((n, rows, cols), np.float32,
                                ptr=self.data.data_ptr() + 4 * off,
                                strides=st, owner=self.data,
                                device=self.data.device.index)


def copy_to_texture(array, texture_type: str = "array",
                    layered: bool = False) -> Texture:
    """Texture snapshot of a float32 array (``kernel.py:47-119``).

    ``texture_type`` 'array' expects a 3-D array, 'array2d'/'pitch2d' a 2-D
    one (stored as a single layer); pitched or strided views are accepted.
    """
    if not (isinstance(array, np.ndarray) or is_device_array(array)):
        raise TypeError(
            f"Expected a NumPy array or a device array, got {type(array)} "
            "instead.")
    if _dev.dtype_name(array) != "float32":
        raise ValueError(
            f"Array must be of type float32, got {array.dtype} instead.")
    kind = texture_type.lower()
    if kind == "array":
        if array.ndim != 3:
            raise ValueError(
                f"Type '{texture_type}' is not compatible with array of ndim "
                f"{array.ndim}.")
    elif kind in ("array2d", "pitch2d"):
        if array.ndim != 2:
            raise ValueError(
                f"Type '{texture_type}' is not compatible with array of ndim "
                f"{array.ndim}.")
        if layered:
            raise ValueError("layered textures must be 3-D")
    else:
        raise ValueError(f"Type `{texture_type}` not understood.")
    return Texture(array, layered=layered)


@dataclass(frozen=True, order=True)
class cuda_float4:
    """Four-component record in (x, y, z, w) order (``kernel.py:150-163``)."""

    x: float
    y: float
    z: float
    w: float

    def to_list(self):
        return [self.x, self.y, self.z, self.w]

    def __str__(self):
        return str(self.to_list())


class BaseKernel(ABC):
    """Abstract base of device kernels (``kernel.py:166-239``).

    Kernels work in place and return nothing; a call launches device work
    asynchronously on the current stream and allocates nothing, so it can be
    captured in a HIP graph (``docs/user_guide/kernels.rst:16-21``).
    ``compile()`` loads the gfx950 library and creates the kernel's plan.
    """

    @abstractmethod
    def __init__(self):
        self._plan = None

    @abstractmethod
    def __call__(self, *args, **kwargs) -> None:
        pass

    def _compile(self, template_kwargs: dict | None = None):
        _native.lib()
        self._compiled_kwargs = dict(template_kwargs or {})
        self._plan = _native.Plan()
        return self._plan

    @property
    def _compiled_template_kwargs(self) -> dict:
        try:
            return self._compiled_kwargs
        except AttributeError:
            raise KernelNotCompiledException(
                "Cannot query the compiled template arguments of kernel "
                f"{self.__class__.__name__} before compilation.")

    @property
    def is_compiled(self) -> bool:
        return getattr(self, "_plan", None) is not None
