"""ctypes binding of the gfx950 cone-beam library (``include/kernelkit_hip.h``).

This is the thin C-ABI layer that replaces the reference's NVRTC/CuPy kernel
runtime (``kernelkit/kernel.py:188-226``).  The library is built ahead of
time by ``build_native.py``; if it is missing, every device operation raises
:class:`NativeLibraryError` -- there is no CPU fallback.

Every entry point takes raw device pointers, element strides and a raw
``hipStream_t``; the arrays may be torch tensors, ``__cuda_array_interface__``
arrays or :class:`kernelkit.DeviceArray` (memory from the library's own
``kk_malloc``), and streams torch streams, objects with a ``cuda_stream`` /
``ptr`` attribute, or plain integers.
"""
from __future__ import annotations

import contextlib
import ctypes
import os
import sys
import threading
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "_lib" / "libkernelkit_hip.so"

KK_OK = 0
KK_ERR_INVALID = -1
KK_ERR_HIP = -2
KK_ERR_NOT_SET = -3
KK_ERR_SHAPE = -4
KK_ERR_UNSUPPORTED = -5
KK_COPY_H2D = 1
KK_COPY_D2H = 2
KK_COPY_D2D = 3
HALO_LO = 1
HALO_HI = 2
HALO = HALO_LO + HALO_HI


class NativeLibraryError(RuntimeError):
    """The HIP library is not built or cannot be loaded."""


class KernelError(RuntimeError):
    """A HIP runtime failure reported by the library."""


_lib = None

_c_i32p = ctypes.POINTER(ctypes.c_int32)
_c_i64p = ctypes.POINTER(ctypes.c_int64)
_c_f64p = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p

# name -> (restype, argtypes)
_SIGNATURES = {
    "kk_last_error": (ctypes.c_char_p, []),
    "kk_version": (ctypes.c_char_p, []),
    "kk_plan_create": (ctypes.c_int, [ctypes.POINTER(_vp)]),
    "kk_plan_destroy": (ctypes.c_int, [_vp]),
    "kk_set_params_fp": (ctypes.c_int, [_vp, _c_f64p, ctypes.c_int32, _c_i32p,
                                        ctypes.c_int32, ctypes.c_int32,
                                        _c_i32p, _c_f64p]),
    "kk_fp_frame_mask": (ctypes.c_int32, [_vp]),
    "kk_set_params_bp": (ctypes.c_int, [_vp, _c_f64p, ctypes.c_int32]),
    "kk_fp_texture_elems": (ctypes.c_int64, [_c_i32p, ctypes.c_int32]),
    "kk_bp_texture_elems": (ctypes.c_int64, [ctypes.c_int32, ctypes.c_int32,
                                             ctypes.c_int32]),
    "kk_volume_to_fp_texture": (ctypes.c_int, [_vp, _c_i32p, _c_i64p,
                                               ctypes.c_int32, _vp, _vp]),
    "kk_projections_to_bp_texture": (ctypes.c_int, [_vp, ctypes.c_int32,
                                                    ctypes.c_int32,
                                                    ctypes.c_int32, _c_i64p,
                                                    _vp, _vp]),
    "kk_fill": (ctypes.c_int, [_vp, ctypes.c_int64, ctypes.c_float, _vp]),
    "kk_device_count": (ctypes.c_int, [_c_i32p]),
    "kk_get_device": (ctypes.c_int, [_c_i32p]),
    "kk_set_device": (ctypes.c_int, [ctypes.c_int32]),
    "kk_malloc": (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.c_int64]),
    "kk_free": (ctypes.c_int, [_vp]),
    "kk_memcpy": (ctypes.c_int, [_vp, _vp, ctypes.c_int64, ctypes.c_int32,
                                 _vp]),
    "kk_stream_synchronize": (ctypes.c_int, [_vp]),
    "kk_copy_strided": (ctypes.c_int, [_vp, _c_i64p, _vp, _c_i64p, _c_i64p,
                                       ctypes.c_int32, _vp]),
    "kk_cone_fp": (ctypes.c_int, [_vp, _vp, _vp, _c_i64p, ctypes.c_int32,
                                  ctypes.c_int32, ctypes.c_int32, _vp]),
    "kk_cone_fp_ptrs": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int32,
                                       _c_i64p, ctypes.c_int32,
                                       ctypes.c_int32, ctypes.c_int32, _vp]),
    "kk_batch_create":(ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p)]),
    "kk_batch_destroy": (ctypes.c_int, [_vp]),
    "kk_batch_size": (ctypes.c_int32, [_vp, ctypes.c_int32]),
    "kk_batch_add_fp": (ctypes.c_int, [
        _vp, _vp, _vp, _vp, _c_i64p, ctypes.c_int32, ctypes.c_int32,
        ctypes.c_int32]),
    "kk_batch_add_bp": (ctypes.c_int, [
        _vp, _vp, _vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _vp,
        _c_i32p, _c_i64p, ctypes.c_float, ctypes.c_int32]),
    "kk_batch_add_sirt_fp": (ctypes.c_int, [
        _vp, _vp, _vp, _vp, _c_i64p, _vp, _c_i64p, _vp, ctypes.c_int32,
        ctypes.c_int32, ctypes.c_int32]),
    "kk_batch_add_sirt_bp": (ctypes.c_int, [
        _vp, _vp, _vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _vp,
        _c_i32p, _c_i64p, _vp, _c_i64p, _vp, _c_i64p, ctypes.c_float,
        ctypes.c_float, ctypes.c_int32, ctypes.c_int32, _vp, ctypes.c_int32,
        ctypes.c_float]),
    "kk_batch_configure": (ctypes.c_int, [_vp, ctypes.c_int32,
                                          ctypes.c_int32]),
    "kk_batch_fp": (ctypes.c_int, [_vp, ctypes.c_int32, _vp]),
    "kk_batch_bp": (ctypes.c_int, [_vp, ctypes.c_int32, _vp]),
    "kk_batch_add_residual": (ctypes.c_int, [
        _vp, _vp, _c_i64p, _vp, _c_i64p, _vp, _c_i64p, ctypes.c_int32,
        ctypes.c_int32, ctypes.c_int32, _vp]),
    "kk_batch_add_volume_texture": (ctypes.c_int, [
        _vp, _vp, _c_i32p, _c_i64p, ctypes.c_int32, _vp]),
    "kk_batch_residual": (ctypes.c_int, [_vp, _vp]),
    "kk_device_info": (ctypes.c_int, [_c_i32p, ctypes.c_char_p,
                                      ctypes.c_int32]),
    "kk_stream_create": (ctypes.c_int, [ctypes.POINTER(_vp)]),
    "kk_stream_destroy": (ctypes.c_int, [_vp]),
    "kk_stream_wait_event": (ctypes.c_int, [_vp, _vp]),
    "kk_event_create": (ctypes.c_int, [ctypes.POINTER(_vp)]),
    "kk_event_destroy": (ctypes.c_int, [_vp]),
    "kk_event_record": (ctypes.c_int, [_vp, _vp]),
    "kk_event_synchronize": (ctypes.c_int, [_vp]),
    "kk_event_elapsed_ms": (ctypes.c_int, [_vp, _vp,
                                           ctypes.POINTER(ctypes.c_float)]),
    "kk_graph_begin_capture": (ctypes.c_int, [_vp]),
    "kk_graph_end_capture": (ctypes.c_int, [_vp, ctypes.POINTER(_vp)]),
    "kk_graph_launch": (ctypes.c_int, [_vp, _vp]),
    "kk_graph_destroy": (ctypes.c_int, [_vp]),
    "kk_batch_volume_textures": (ctypes.c_int, [_vp, _vp]),
    "kk_probe_l2_read": (ctypes.c_int, [_vp, ctypes.c_int64, ctypes.c_int32,
                                        ctypes.c_int32, _vp, _vp]),
    "kk_probe_gather": (ctypes.c_int, [
        _vp, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
        ctypes.c_int32, ctypes.c_int32, _vp, _vp]),
    "kk_cone_fp_transpose": (ctypes.c_int, [
        _vp, _vp, _c_i64p, ctypes.c_int32, ctypes.c_int32, _vp, _c_i32p,
        _c_i64p, ctypes.c_int32, _vp, _vp]),
    "kk_cone_fp_transpose_static": (ctypes.c_int, [
        _vp, _vp, _c_i64p, ctypes.c_int32, ctypes.c_int32, _vp, _c_i32p,
        _c_i64p, ctypes.c_int32, _vp]),
    "kk_cone_fp_transpose_static_capacity": (ctypes.c_int64, []),
    "kk_cone_fp_transpose_work_elems": (ctypes.c_int64, [
        _vp, ctypes.c_int32, ctypes.c_int32]),
    "kk_cone_bp":(ctypes.c_int, [_vp, _vp, ctypes.c_int32, ctypes.c_int32,
                                  ctypes.c_int32, _vp, _c_i32p, _c_i64p,
                                  ctypes.c_float, ctypes.c_int32, _vp]),
    "kk_sirt_fp_residual": (ctypes.c_int, [_vp, _vp, _vp, _c_i64p, _vp,
                                           _c_i64p, _vp, ctypes.c_int32,
                                           ctypes.c_int32, ctypes.c_int32,
                                           _vp]),
    "kk_residual_to_bp_texture": (ctypes.c_int, [_vp, _c_i64p, _vp, _c_i64p,
                                                 _vp, _c_i64p, ctypes.c_int32,
                                                 ctypes.c_int32,
                                                 ctypes.c_int32, _vp, _vp]),
    "kk_sirt_bp_update": (ctypes.c_int, [_vp, _vp, ctypes.c_int32,
                                         ctypes.c_int32, ctypes.c_int32, _vp,
                                         _c_i32p, _c_i64p, _vp, _c_i64p, _vp,
                                         _c_i64p, ctypes.c_float,
                                         ctypes.c_float, ctypes.c_int32,
                                         ctypes.c_int32, _vp, ctypes.c_int32,
                                         ctypes.c_float, _vp]),
    "kk_recip_or_zero": (ctypes.c_int, [_vp, ctypes.c_int64, _vp]),
    "kk_fp_table": (ctypes.c_int, [_c_f64p, ctypes.c_int32, _c_i32p,
                                   ctypes.c_int32, ctypes.c_int32, _c_i32p,
                                   _c_f64p, _vp, _c_i32p]),
    "kk_bp_table": (ctypes.c_int, [_c_f64p, ctypes.c_int32, _vp]),
    "kk_filter_create": (ctypes.c_int, [ctypes.POINTER(_vp)]),
    "kk_filter_destroy": (ctypes.c_int, [_vp]),
    "kk_filter_set_response": (ctypes.c_int, [_vp, _c_f64p, ctypes.c_int32]),
    "kk_filter_set_preweight": (ctypes.c_int, [_vp, _c_f64p,
                                               ctypes.c_int32]),
    "kk_preweight": (ctypes.c_int, [_vp, _vp, ctypes.c_int32, ctypes.c_int32,
                                    ctypes.c_int32, _c_i64p, _vp]),
    "kk_filter_rows": (ctypes.c_int, [_vp, _vp, _c_i64p, _vp, _c_i64p,
                                      ctypes.c_int32, ctypes.c_int32,
                                      ctypes.c_int32, ctypes.c_int32, _vp]),
    "kk_filter_rows_to_bp_texture": (ctypes.c_int, [_vp, _vp, _c_i64p, _vp,
                                                    ctypes.c_int32,
                                                    ctypes.c_int32,
                                                    ctypes.c_int32,
                                                    ctypes.c_int32, _vp]),
}


def exported_symbols() -> list:
    """C symbols this binding requires from the library."""
    return sorted(_SIGNATURES)


def lib():
    """Load (once) and return the library; raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    path = Path(os.environ.get("KK_LIBRARY", LIB_PATH))
    if not path.exists():
        raise NativeLibraryError(
            f"HIP library not found at {path}; build it with "
            "`python astra-kernelkit_amd/build_native.py` (hipcc, gfx950).")
    try:
        handle = ctypes.CDLL(str(path))
    except OSError as exc:
        raise NativeLibraryError(f"cannot load {path}: {exc}") from exc
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(handle, name)
        fn.restype = res
        fn.argtypes = args
    _lib = handle
    st = path.stat()
    global _build_id
    _build_id = (f"{handle.kk_version().decode()} "
                 f"{st.st_size}:{st.st_mtime_ns}")
    return _lib


_build_id = None


def build_id() -> str:
    """Version string plus size and mtime of the loaded library file (changes
    with every rebuild)."""
    lib()
    return _build_id


def last_error() -> str:
    return lib().kk_last_error().decode(errors="replace")


def check(status: int, exc_map: dict | None = None) -> None:
    """Raise the Python exception matching a library status code."""
    if status == KK_OK:
        return
    msg = last_error()
    if exc_map and status in exc_map:
        raise exc_map[status](msg)
    if status in (KK_ERR_INVALID, KK_ERR_SHAPE):
        raise ValueError(msg)
    if status == KK_ERR_UNSUPPORTED:
        raise NotImplementedError(msg)
    if status == KK_ERR_NOT_SET:
        from kernelkit.kernel import KernelMemoryUnpreparedException
        raise KernelMemoryUnpreparedException(msg)
    raise KernelError(msg)


def i32(values) -> ctypes.Array:
    vals = [int(v) for v in values]
    return (ctypes.c_int32 * len(vals))(*vals)


def i64(values) -> ctypes.Array:
    vals = [int(v) for v in values]
    return (ctypes.c_int64 * len(vals))(*vals)


def f64(values) -> ctypes.Array:
    vals = [float(v) for v in values]
    return (ctypes.c_double * len(vals))(*vals)


def f64_ptr(arr: np.ndarray):
    """Pointer to a C-contiguous float64 numpy array (caller keeps it alive)."""
    assert arr.dtype == np.float64 and arr.flags["C_CONTIGUOUS"]
    return arr.ctypes.data_as(_c_f64p)


def i32_ptr(arr: np.ndarray):
    assert arr.dtype == np.int32 and arr.flags["C_CONTIGUOUS"]
    return arr.ctypes.data_as(_c_i32p)


_local = threading.local()


@contextlib.contextmanager
def use_stream(stream):
    """Make ``stream`` the default of :func:`stream_ptr` in this thread
    (graph capture records the library's launches on its own stream)."""
    prev = getattr(_local, "stream", None)
    _local.stream = stream_ptr(stream)
    try:
        yield
    finally:
        _local.stream = prev


def stream_ptr(stream=None) -> ctypes.c_void_p:
    """Raw hipStream_t of ``stream``: a torch stream, an object carrying
    the handle as ``cuda_stream`` or ``ptr`` (CuPy's), or an integer.  By
    default the stream set by :func:`use_stream`, else torch's current
    stream when PyTorch is loaded and has set up its HIP runtime, else the
    null stream (torch is never imported here)."""
    if isinstance(stream, ctypes.c_void_p):
        return stream
    if stream is None:
        override = getattr(_local, "stream", None)
        if override is not None:
            return override
        torch = sys.modules.get("torch")
        if torch is not None and torch.cuda.is_initialized():
            stream = torch.cuda.current_stream()
        else:
            return ctypes.c_void_p(None)
    if isinstance(stream, int):
        return ctypes.c_void_p(stream or None)
    for attr in ("cuda_stream", "ptr"):
        handle = getattr(stream, attr, None)
        if handle is not None:
            return ctypes.c_void_p(int(handle) or None)
    raise TypeError(f"not a stream: {type(stream).__name__}")


class Plan:
    """Owner of a library plan (per-angle device tables)."""

    def __init__(self):
        ptr = _vp()
        check(lib().kk_plan_create(ctypes.byref(ptr)))
        self._ptr = ptr

    @property
    def ptr(self) -> ctypes.c_void_p:
        return self._ptr

    def __del__(self):
        ptr = getattr(self, "_ptr", None)
        if ptr is not None and ptr.value and _lib is not None:
            try:
                _lib.kk_plan_destroy(ptr)
            except Exception:  # interpreter shutdown
                pass
            self._ptr = None


class Batch:
    """Owner of a library batch (problem descriptors for one launch per
    operator)."""

    def __init__(self):
        ptr = _vp()
        check(lib().kk_batch_create(ctypes.byref(ptr)))
        self._ptr = ptr

    @property
    def ptr(self) -> ctypes.c_void_p:
        return self._ptr

    def __del__(self):
        ptr = getattr(self, "_ptr", None)
        if ptr is not None and ptr.value and _lib is not None:
            try:
                _lib.kk_batch_destroy(ptr)
            except Exception:  # interpreter shutdown
                pass
            self._ptr = None


class Filter:
    """Owner of a library filter (twiddles, response, pre-weight tables)."""

    def __init__(self):
        ptr = _vp()
        check(lib().kk_filter_create(ctypes.byref(ptr)))
        self._ptr = ptr
        self.length = 0
        self.name = None

    @property
    def ptr(self) -> ctypes.c_void_p:
        return self._ptr

    def __del__(self):
        ptr = getattr(self, "_ptr", None)
        if ptr is not None and ptr.value and _lib is not None:
            try:
                _lib.kk_filter_destroy(ptr)
            except Exception:  # interpreter shutdown
                pass
            self._ptr = None


def device_info() -> tuple:
    """``(compute units, device name)`` of the current device (the
    architecture name where the runtime reports no marketing name)."""
    cus = ctypes.c_int32(0)
    name = ctypes.create_string_buffer(256)
    check(lib().kk_device_info(ctypes.byref(cus), name, len(name)))
    return int(cus.value), name.value.decode(errors="replace")


class Stream:
    """Owner of a library stream (non-blocking, default priority)."""

    def __init__(self):
        ptr = _vp()
        check(lib().kk_stream_create(ctypes.byref(ptr)))
        self._ptr = ptr

    @property
    def ptr(self) -> ctypes.c_void_p:
        return self._ptr

    def wait(self, event: "Event") -> None:
        check(lib().kk_stream_wait_event(self._ptr, event.ptr))

    def synchronize(self) -> None:
        check(lib().kk_stream_synchronize(self._ptr))

    def __del__(self):
        ptr = getattr(self, "_ptr", None)
        if ptr is not None and ptr.value and _lib is not None:
            try:
                _lib.kk_stream_destroy(ptr)
            except Exception:  # interpreter shutdown
                pass
            self._ptr = None


class Event:
    """Owner of a timing event."""

    def __init__(self):
        ptr = _vp()
        check(lib().kk_event_create(ctypes.byref(ptr)))
        self._ptr = ptr

    @property
    def ptr(self) -> ctypes.c_void_p:
        return self._ptr

    def record(self, stream=None) -> None:
        check(lib().kk_event_record(self._ptr, stream_ptr(stream)))

    def synchronize(self) -> None:
        check(lib().kk_event_synchronize(self._ptr))

    def elapsed_ms(self, end: "Event") -> float:
        ms = ctypes.c_float(0.0)
        check(lib().kk_event_elapsed_ms(self._ptr, end.ptr,
                                        ctypes.byref(ms)))
        return float(ms.value)

    def __del__(self):
        ptr = getattr(self, "_ptr", None)
        if ptr is not None and ptr.value and _lib is not None:
            try:
                _lib.kk_event_destroy(ptr)
            except Exception:  # interpreter shutdown
                pass
            self._ptr = None


class Graph:
    """An executable HIP graph of library launches.

    ``Graph.capture(fn)`` runs ``fn`` with :func:`stream_ptr` defaulting to
    a fresh capturing stream, so every launch it makes through this binding
    is recorded instead of run; ``fn`` must not allocate device memory,
    synchronise or launch on an explicit stream.  ``replay`` launches the
    recorded work on ``stream`` (default: :func:`stream_ptr`)."""

    def __init__(self, exec_ptr):
        self._ptr = exec_ptr

    @classmethod
    def capture(cls, fn) -> "Graph":
        s = Stream()
        check(lib().kk_graph_begin_capture(s.ptr))
        try:
            with use_stream(s.ptr):
                fn()
        except BaseException:
            dead = _vp()
            if lib().kk_graph_end_capture(s.ptr, ctypes.byref(dead)) == KK_OK:
                lib().kk_graph_destroy(dead)
            raise
        exec_ptr = _vp()
        check(lib().kk_graph_end_capture(s.ptr, ctypes.byref(exec_ptr)))
        return cls(exec_ptr)

    @property
    def ptr(self) -> ctypes.c_void_p:
        return self._ptr

    def replay(self, stream=None) -> None:
        check(lib().kk_graph_launch(self._ptr, stream_ptr(stream)))

    def __del__(self):
        ptr = getattr(self, "_ptr", None)
        if ptr is not None and ptr.value and _lib is not None:
            try:
                _lib.kk_graph_destroy(ptr)
            except Exception:  # interpreter shutdown
                pass
            self._ptr = None


# numpy views of the device records (csrc/kk_internal.h)
FP_ANGLE_DTYPE = np.dtype([
    ("qa", "<f8"), ("qb", "<f8"), ("qc", "<f8"),
    ("ua", "<f8"), ("ub", "<f8"), ("uc", "<f8"),
    ("va", "<f8"), ("vb", "<f8"), ("vc", "<f8"),
    ("sig", "<f8"), ("taub", "<f8"), ("tauc", "<f8"),
    ("sc1", "<f4"), ("sc2", "<f4"), ("oscale", "<f4"), ("_p0", "<f4"),
    ("na", "<i4"), ("nb", "<i4"), ("nc", "<i4"), ("axis", "<i4"),
    ("tex_off", "<u8"), ("sab", "<u4"), ("sbb", "<u4"),
    ("colc", "<i4"), ("_p1", "<i4", (3,))])
BP_ANGLE_DTYPE = np.dtype([
    ("nux", "<f8"), ("nuy", "<f8"), ("nuz", "<f8"), ("nuw", "<f8"),
    ("nvx", "<f8"), ("nvy", "<f8"), ("nvz", "<f8"), ("nvw", "<f8"),
    ("dx", "<f8"), ("dy", "<f8"), ("dz", "<f8"), ("dw", "<f8"),
    ("fast", "<i4"), ("_p0", "<i4", (3,))])
assert FP_ANGLE_DTYPE.itemsize == 160 and BP_ANGLE_DTYPE.itemsize == 112


def fp_table(params12: np.ndarray, ray_axes, rows, cols, shape,
             voxel_size):
    """Host copy of the forward-projection device table and its frame
    mask."""
    params12 = np.ascontiguousarray(params12, dtype=np.float64)
    n = params12.shape[1]
    out = np.zeros(n, dtype=FP_ANGLE_DTYPE)
    axes = (None if ray_axes is None
            else np.ascontiguousarray(ray_axes, dtype=np.int32))
    mask = ctypes.c_int32(0)
    check(lib().kk_fp_table(f64_ptr(params12), n,
                            None if axes is None else i32_ptr(axes), rows,
                            cols, i32(shape), f64(voxel_size),
                            out.ctypes.data_as(_vp), ctypes.byref(mask)))
    return out, int(mask.value)


def bp_table(params: np.ndarray) -> np.ndarray:
    """Host copy of the backprojection device table."""
    params = np.ascontiguousarray(np.asarray(params, dtype=np.float64).ravel())
    n = params.size // 12
    out = np.zeros(n, dtype=BP_ANGLE_DTYPE)
    check(lib().kk_bp_table(f64_ptr(params), n, out.ctypes.data_as(_vp)))
    return out
