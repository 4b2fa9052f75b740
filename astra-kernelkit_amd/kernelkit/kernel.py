"""Kernel runtime: base class, exceptions and padded textures.

Mirrors ``kernelkit/kernel.py`` of the reference.  There, a kernel is a
Jinja2-rendered CUDA source compiled by NVRTC (``:188-226``), geometry is
copied into ``__constant__`` memory (``:122-147``) and inputs are sampled
through CUDA texture objects with border addressing (``:22-32``, ``:47-119``).

Here kernels are ahead-of-time gfx950 code in ``libkernelkit_hip.so``:
``compile()`` loads the library and creates a plan that owns the per-angle
device tables; a "texture" is a zero-halo padded copy of the array in its own
axis order (:class:`Texture`), which the HIP kernels sample with explicit
float32 interpolation and no per-tap bounds tests.
"""
from __future__ import annotations

from abc import ABC, abstractmethod
from dataclasses import dataclass

import numpy as np

from kernelkit import _native


class KernelNotCompiledException(Exception):
    """Raised when a kernel is called before it has been compiled."""


class KernelMemoryUnpreparedException(Exception):
    """Raised when a kernel is called before memory has been prepared."""


def _torch():
    import torch
    return torch


def default_device():
    """The HIP device used for new arrays (torch's current device)."""
    torch = _torch()
    if not torch.cuda.is_available():
        raise _native.NativeLibraryError(
            "no HIP device available: the cone-beam operators run on the GPU "
            "only")
    return torch.device("cuda", torch.cuda.current_device())


def is_device_array(x) -> bool:
    torch = _torch()
    return isinstance(x, torch.Tensor) and x.is_cuda


def as_device_array(x, dtype=None, copy: bool = False):
    """Device float32 tensor from a tensor or array-like (copies host data)."""
    torch = _torch()
    dtype = dtype or torch.float32
    if isinstance(x, torch.Tensor):
        if x.is_cuda and x.dtype == dtype and not copy:
            return x
        return x.to(device=default_device() if not x.is_cuda else x.device,
                    dtype=dtype, copy=copy)
    arr = np.asarray(x, dtype=np.float32)
    return torch.from_numpy(np.ascontiguousarray(arr)).to(default_device())


class Texture:
    """Zero-halo padded float32 copy of a 3-D device array.

    The copy keeps the array's axis order (like the CUDA array behind a
    texture object) and pads every axis with ``HALO_LO`` zeros before and
    ``HALO_HI`` after, so kernels can clamp coordinates into the halo instead
    of testing bounds.  ``shape`` is the unpadded array shape.
    """

    def __init__(self, shape, device=None, layered: bool = False):
        torch = _torch()
        self.shape = tuple(int(s) for s in shape)
        if len(self.shape) != 3 or min(self.shape) <= 0:
            raise ValueError(f"texture shape must be 3-D, got {self.shape}")
        self.layered = bool(layered)
        n = _native.lib().kk_texture_elems(_native.i64(self.shape))
        if n < 0:
            raise ValueError(f"bad texture shape {self.shape}")
        self.data = torch.empty(int(n), dtype=torch.float32,
                                device=device or default_device())

    @property
    def padded_shape(self):
        return tuple(s + _native.HALO for s in self.shape)

    @property
    def ptr(self):
        return self.data.data_ptr()

    def interior(self):
        """View of the unpadded region (shape ``self.shape``)."""
        lo = _native.HALO_LO
        full = self.data.view(self.padded_shape)
        return full[lo:lo + self.shape[0], lo:lo + self.shape[1],
                    lo:lo + self.shape[2]]

    def copy_from(self, array, stream=None) -> None:
        """Asynchronously copy ``array`` (same shape, any strides) in."""
        torch = _torch()
        if not (isinstance(array, torch.Tensor) and array.is_cuda):
            array = as_device_array(array)
        if array.dtype != torch.float32:
            raise ValueError(
                f"Array must be of type float32, got {array.dtype} instead.")
        if array.dim() == 2:
            array = array.unsqueeze(0)
        if tuple(array.shape) != self.shape:
            raise ValueError(f"array shape {tuple(array.shape)} does not "
                             f"match texture shape {self.shape}")
        _native.check(_native.lib().kk_array_to_texture(
            _native._vp(array.data_ptr()), _native.i64(array.shape),
            _native.i64(array.stride()), _native._vp(self.ptr),
            _native.stream_ptr(stream)))
        self._source = array  # the caller's array stays referenced

    def copy_to(self, array, stream=None) -> None:
        """Copy the interior back into ``array`` (same shape)."""
        array.copy_(self.interior().reshape(array.shape))


def copy_to_texture(array, texture_type: str = "array",
                    layered: bool = False) -> Texture:
    """Padded texture copy of a float32 array (``kernel.py:47-119``).

    ``texture_type`` 'array' expects a 3-D array, 'array2d'/'pitch2d' a 2-D
    one (stored as a single layer); pitched or strided views are accepted
    as they are (strides are honoured by the copy).
    """
    torch = _torch()
    if not isinstance(array, (np.ndarray, torch.Tensor)):
        raise TypeError(
            f"Expected a NumPy array or a device tensor, got {type(array)} "
            "instead.")
    dtype_ok = (array.dtype == np.float32 if isinstance(array, np.ndarray)
                else array.dtype == torch.float32)
    if not dtype_ok:
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
    dev = as_device_array(array)
    shape = tuple(dev.shape) if dev.dim() == 3 else (1, *dev.shape)
    tex = Texture(shape, device=dev.device, layered=layered)
    tex.copy_from(dev)
    return tex


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
