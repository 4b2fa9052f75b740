"""FDK pre-processing on the GPU: cosine pre-weight and row ramp filter.

Drop-in for ``kernelkit/processing.py`` of the reference: ``_filter``
(:44-93), ``filter`` (:96-126) and ``preweight`` (:149-188).  The frequency
response and the per-angle weights are computed in float64 on the host; the
work runs in the gfx950 kernels ``kk_preweight`` and ``kk_filter_rows`` (an
LDS FFT over two rows at a time), one launch for all angles.

Filters: 'ramp'/'ramlak', 'cosine' (as the reference) and 'hann', 'hamming',
'shepp-logan' (scikit-image ``iradon`` windows, which the reference's filter
follows, ``processing.py:49-50``).
"""
from __future__ import annotations

from collections import OrderedDict
from functools import lru_cache

import numpy as np

from kernelkit import _native
from kernelkit.kernel import as_device_array, is_device_array


def padded_length(cols: int) -> int:
    """FFT length ``max(64, 2**ceil(log2(2 cols)))`` (``processing.py:111``)."""
    return max(64, int(2 ** int(np.ceil(np.log2(2 * cols)))))


@lru_cache
def _filter(num: int, filter_name: str = "ramlak", use_cupy: bool = True):
    """Frequency response of length ``num`` (float64, host).

    Built in the spatial domain (Kak & Slaney): ``g[0] = 1/4``,
    ``g[n] = -1/(pi n)^2`` for odd circular distance ``n``, zero otherwise;
    ``H = 4 Re(DFT(g))``, then an optional window.  ``use_cupy`` is accepted
    for compatibility and ignored.
    """
    if num % 2 != 0:
        raise ValueError("Filter must be even.")
    half = num // 2
    dist = np.minimum(np.arange(num), num - np.arange(num))
    g = np.zeros(num)
    g[0] = 0.25
    odd = dist % 2 == 1
    g[odd] = -1.0 / (np.pi * dist[odd]) ** 2
    H = 4.0 * np.fft.fft(g).real
    name = filter_name.lower()
    if name in ("ramp", "ramlak"):
        pass
    elif name == "cosine":
        H = H * np.fft.fftshift(np.sin(np.linspace(0.0, np.pi, num,
                                                   endpoint=False)))
    elif name == "hann":
        H = H * np.fft.fftshift(np.hanning(num))
    elif name == "hamming":
        H = H * np.fft.fftshift(np.hamming(num))
    elif name == "shepp-logan":
        w = np.pi * np.fft.fftfreq(num)[1:]
        H = H.copy()
        H[1:] *= np.sin(w) / w
    else:
        raise ValueError(f"Filter with `filter_name` {filter_name} not known "
                         "or not yet implemented.")
    assert half > 0
    return H


_FILTER_CACHE: dict = {}


def _device_filter(num: int, name: str) -> _native.Filter:
    key = (num, name.lower())
    f = _FILTER_CACHE.get(key)
    if f is None:
        H = np.ascontiguousarray(_filter(num, name), dtype=np.float64)
        f = _native.Filter()
        _native.check(_native.lib().kk_filter_set_response(
            f.ptr, _native.f64_ptr(H), num))
        f.length, f.name = num, name.lower()
        _FILTER_CACHE[key] = f
    return f


def preweight_params(geoms, detector_piercings=None) -> np.ndarray:
    """Per-angle pre-weight geometry (float64, ``N x 10``): ``d_min - s``,
    ``u * pw``, ``v * ph`` and the central ray length ``|p - s|`` where ``p``
    is the detector centre or the given piercing point."""
    geoms = list(geoms)
    n = len(geoms)

    def rows3(attr):
        return np.array([getattr(g, attr) for g in geoms],
                        dtype=np.float64).reshape(n, 3)

    def per_angle(get):
        return np.array([get(g.detector) for g in geoms],
                        dtype=np.float64).reshape(n, 1)

    s, d = rows3("source_position"), rows3("detector_position")
    U = rows3("u") * per_angle(lambda det: det.pixel_width)
    V = rows3("v") * per_angle(lambda det: det.pixel_height)
    dmin = (d - V * (per_angle(lambda det: det.rows) / 2)
            - U * (per_angle(lambda det: det.cols) / 2))
    p = d if detector_piercings is None else np.asarray(
        detector_piercings, dtype=np.float64).reshape(n, 3)
    out = np.empty((n, 10))
    out[:, 0:3] = dmin - s
    out[:, 3:6] = U
    out[:, 6:9] = V
    ps = p - s
    out[:, 9] = np.sqrt(ps[:, 0] * ps[:, 0] + ps[:, 1] * ps[:, 1]
                        + ps[:, 2] * ps[:, 2])
    return out


def _as_stack(projections):
    """(tensor, writeback) for a device tensor or list of 2-D tensors."""
    if isinstance(projections, (list, tuple)):
        from kernelkit import device
        stack = device.stack([as_device_array(p) for p in projections])

        def writeback():
            for k, p in enumerate(projections):
                p.copy_(stack[k])
        return stack, writeback
    if not is_device_array(projections):
        raise TypeError("projections must be a device array or a list of "
                        "device arrays")
    return as_device_array(projections), None


def preweight(projections, geoms, detector_piercings=None,
              verbose: bool = False) -> None:
    """In-place FDK cosine weight ``p *= |p_c - s| / |corner(r, c) - s|``
    (``processing.py:149-188``; pixel corners, not centres)."""
    stack, writeback = _as_stack(projections)
    geoms = list(geoms)
    n, rows, cols = stack.shape
    f = _native.Filter()
    P = np.ascontiguousarray(preweight_params(geoms, detector_piercings))
    _native.check(_native.lib().kk_filter_set_preweight(
        f.ptr, _native.f64_ptr(P), n))
    _native.check(_native.lib().kk_preweight(
        f.ptr, _native._vp(stack.data_ptr()), n, rows, cols,
        _native.i64(stack.stride()), _native.stream_ptr()))
    if writeback:
        writeback()


def filter(projections, verbose: bool = False,
           filter: str = "ramlak") -> None:  # noqa: A001 -- reference name
    """In-place row filter (``processing.py:96-126``)."""
    stack, writeback = _as_stack(projections)
    n, rows, cols = stack.shape
    f = _device_filter(padded_length(cols), filter)
    st = _native.i64(stack.stride())
    _native.check(_native.lib().kk_filter_rows(
        f.ptr, _native._vp(stack.data_ptr()), st,
        _native._vp(stack.data_ptr()), st, n, rows, cols, 0,
        _native.stream_ptr()))
    if writeback:
        writeback()


_TABLES: "OrderedDict[tuple, _native.Filter]" = OrderedDict()
_TABLES_KEPT = 8


def _filter_tables(num: int, filter: str, P) -> "_native.Filter":
    """Device tables (twiddles, response, pre-weights ``P`` or none) of a
    fused filter call, kept for the last few (device, length, filter,
    pre-weight) combinations.  A table set is written once, when made, and
    never changed: a later call with the same key launches on it while
    earlier launches may still read it, with no upload, allocation or free
    (whose device-wide wait would otherwise end every call)."""
    from kernelkit import device
    key = (device.current_device(), num, str(filter).lower(),
           None if P is None else (P.shape, P.tobytes()))
    f = _TABLES.get(key)
    if f is not None:
        _TABLES.move_to_end(key)
        return f
    f = _native.Filter()
    H = np.ascontiguousarray(_filter(num, filter), dtype=np.float64)
    _native.check(_native.lib().kk_filter_set_response(
        f.ptr, _native.f64_ptr(H), num))
    if P is not None:
        _native.check(_native.lib().kk_filter_set_preweight(
            f.ptr, _native.f64_ptr(P), P.shape[0]))
    _TABLES[key] = f
    while len(_TABLES) > _TABLES_KEPT:
        _TABLES.popitem(last=False)
    return f


def filter_into(projections, texture, geoms=None, filter: str = "ramlak",
                detector_piercings=None) -> None:
    """Fused FDK pre-processing into a backprojection texture:
    ``texture <- filter(preweight(projections))``.

    ``projections`` is an ``(angle, row, col)`` device tensor (views
    allowed); ``texture`` a :class:`~kernelkit.kernel.ProjectionTexture` of
    the same shape whose halo is zero.  The pre-weight is applied on load
    when ``geoms`` is given.
    """
    n, rows, cols = projections.shape
    if tuple(texture.shape) != (n, rows, cols):
        raise ValueError(f"texture of shape {texture.shape} does not match "
                         f"projections of shape {(n, rows, cols)}")
    num = padded_length(cols)
    P = None if geoms is None else np.ascontiguousarray(
        preweight_params(geoms, detector_piercings))
    if P is not None and P.shape[0] != n:
        raise ValueError(f"{P.shape[0]} geometries for {n} projections")
    f = _filter_tables(num, filter, P)
    _native.check(_native.lib().kk_filter_rows_to_bp_texture(
        f.ptr, _native._vp(projections.data_ptr()),
        _native.i64(projections.stride()), _native._vp(texture.ptr), n, rows,
        cols, 1 if geoms is not None else 0, _native.stream_ptr()))
