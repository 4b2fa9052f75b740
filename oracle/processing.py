"""FP64 FDK pre-weighting and row filtering for the oracle (test
infrastructure only).

* ``filter_response`` restates ``kernelkit/processing.py:44-93``
  (Kak-Slaney spatial Ram-Lak kernel, 4 * Re(FFT), optional windows);
  ``'hann'``, ``'hamming'`` and ``'shepp-logan'`` are additions that follow
  the scikit-image ``iradon`` convention the reference mirrors
  (``processing.py:49-50``).
* ``filter_rows`` restates ``processing.py:96-126``.
* ``preweight`` restates ``processing.py:149-188`` (pixel corners, world
  units).
"""
from __future__ import annotations

import numpy as np


def padded_length(cols: int) -> int:
    """FFT length: ``max(64, 2**ceil(log2(2*cols)))`` (``processing.py:111``)."""
    return max(64, int(2 ** int(np.ceil(np.log2(2 * cols)))))


def filter_response(num: int, filter_name: str = "ramlak") -> np.ndarray:
    """Frequency response ``H`` of length ``num`` (even), float64.

    ``g[0] = 1/4``, ``g[n] = -1/(pi n)^2`` for odd ``n`` (``n`` is the
    circular distance), else 0; ``H = 4 Re(DFT(g))``
    (``processing.py:76-82``), then a window.
    """
    if num % 2 != 0:
        raise ValueError("Filter must be even.")
    half = num // 2
    n = np.concatenate((np.arange(0, half + 1), np.arange(half - 1, 0, -1)))
    g = np.zeros(num)
    g[0] = 0.25
    odd = (n % 2) == 1
    g[odd] = -1.0 / (np.pi * n[odd]) ** 2
    H = 4.0 * np.real(np.fft.fft(g))
    name = filter_name.lower()
    if name in ("ramp", "ramlak"):
        pass
    elif name == "cosine":
        H = H * np.fft.fftshift(np.sin(np.linspace(0, np.pi, num,
                                                   endpoint=False)))
    elif name == "hann":
        H = H * np.fft.fftshift(np.hanning(num))
    elif name == "hamming":
        H = H * np.fft.fftshift(np.hamming(num))
    elif name == "shepp-logan":
        omega = np.pi * np.fft.fftfreq(num)[1:]
        H = H.copy()
        H[1:] *= np.sin(omega) / omega
    else:
        raise ValueError(f"Filter with `filter_name` {filter_name} not known "
                         "or not yet implemented.")
    return H


def filter_rows(proj: np.ndarray, filter_name: str = "ramlak") -> np.ndarray:
    """Filter every detector row: zero-pad to ``padded_length(cols)``, FFT,
    multiply by ``H``, inverse FFT, real part, truncate (float64)."""
    proj = np.asarray(proj, dtype=np.float64)
    cols = proj.shape[-1]
    N = padded_length(cols)
    H = filter_response(N, filter_name)
    F = np.fft.fft(proj, n=N, axis=-1) * H
    return np.real(np.fft.ifft(F, axis=-1))[..., :cols]


def preweight(proj: np.ndarray, pgs) -> np.ndarray:
    """FDK cosine weight: ``p *= |d - s| / |corner(r, c) - s|`` where
    ``corner(r, c) = d_min + c*u*pw + r*v*ph`` (no half-pixel offset), in
    world units (``processing.py:170-188``)."""
    proj = np.array(proj, dtype=np.float64, copy=True)
    pgs = list(pgs)
    rows, cols = int(pgs[0].detector.rows), int(pgs[0].detector.cols)
    rr = np.arange(rows, dtype=np.float64)[:, None, None]
    cc = np.arange(cols, dtype=np.float64)[None, :, None]
    for k, g in enumerate(pgs):
        s = np.asarray(g.source_position, np.float64)
        d = np.asarray(g.detector_position, np.float64)
        u = np.asarray(g.u, np.float64) * float(g.detector.pixel_width)
        v = np.asarray(g.v, np.float64) * float(g.detector.pixel_height)
        dmin = d - v * rows / 2 - u * cols / 2
        pix = dmin + cc * u + rr * v
        proj[k] *= np.linalg.norm(d - s) / np.linalg.norm(pix - s, axis=2)
    return proj
