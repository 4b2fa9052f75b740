"""FP64 cone-beam projectors for the oracle (test infrastructure only).

* ``cone_fp``  -- ray-driven Joseph forward projection, restating
  ``kernelkit/cuda/cone_fp.cu:141-204`` with the launch logic of
  ``kernelkit/kernels/cone_fp.py:195-262`` (per-angle ray axis, all slices,
  ``+=`` into the projections);
* ``cone_bp``  -- voxel-driven backprojection, restating
  ``kernelkit/cuda/cone_bp.cu:84-158`` with ``kernels/cone_bp.py:292-302``
  (output scaled by the voxel volume);
* ``cone_fp_transpose`` -- the exact transpose of ``cone_fp`` (no reference
  counterpart; the reference backward is the unmatched BP,
  ``kernelkit/torch_support.py:29-33``).

Interpolation follows the texture contract of ``kernelkit/kernel.py:22-32``:
linear filtering, texel centres at ``i + 1/2``, zero outside the array.
Coordinates below are array-index coordinates (texture coordinate minus 1/2),
so ``x = i`` reads element ``i`` exactly.
"""
from __future__ import annotations

import numpy as np

from oracle.geometry import (bp_params, fp_major_axes, fp_output_scale,
                             normalized_frames, voxel_volume)

# ray frame (a, b, c) per major axis: ``cuda/cone_fp.cu:59-61``
_FRAMES = {0: (0, 1, 2), 1: (1, 0, 2), 2: (2, 1, 0)}


def bilinear_border_zero(plane: np.ndarray, xi: np.ndarray, xj: np.ndarray):
    """Bilinear sample of ``plane[i, j]`` at array coordinates (xi, xj),
    with zero outside the array (``kernel.py:22-32``: border address mode)."""
    ni, nj = plane.shape
    i0 = np.floor(xi)
    j0 = np.floor(xj)
    fi = xi - i0
    fj = xj - j0
    i0 = i0.astype(np.int64)
    j0 = j0.astype(np.int64)
    out = np.zeros(np.broadcast(xi, xj).shape)
    for di, wi in ((0, 1.0 - fi), (1, fi)):
        ii = i0 + di
        vi = (ii >= 0) & (ii < ni)
        for dj, wj in ((0, 1.0 - fj), (1, fj)):
            jj = j0 + dj
            ok = vi & (jj >= 0) & (jj < nj)
            vals = plane[np.clip(ii, 0, ni - 1), np.clip(jj, 0, nj - 1)]
            out += np.where(ok, wi * wj * vals, 0.0)
    return out


def _ray_params(fr: dict, k: int, axis: int, shape):
    """Per-ray slope/intercept in the (a, b, c) frame of angle ``k``
    (``cuda/cone_fp.cu:150-178``): returns ``a1, a2`` (slopes of b and c per
    unit step along a) and the array coordinates ``xb0, xc0`` of the ray at
    the centre plane of slice 0, for a ``(rows, cols)`` grid of pixels."""
    a, b, c = _FRAMES[axis]
    rows, cols = fr["rows"], fr["cols"]
    s = fr["s"][k]
    d0, U, V = fr["d0"][k], fr["U"][k], fr["V"][k]
    rr = np.arange(rows, dtype=np.float64)[:, None]
    cc = np.arange(cols, dtype=np.float64)[None, :]
    # pixel centre (cone_fp.cu:150-159)
    D = [d0[t] + (cc + 0.5) * U[t] + (rr + 0.5) * V[t] for t in range(3)]
    sdd = s[a] - D[a]
    a1 = (s[b] - D[b]) / sdd
    a2 = (s[c] - D[c]) / sdd
    b1 = s[b] - a1 * s[a]
    b2 = s[c] - a2 * s[a]
    na, nb, nc = shape[a], shape[b], shape[c]
    w0 = 0.0 - 0.5 * na + 0.5          # centre of slice 0 along a
    xb0 = a1 * w0 + b1 + 0.5 * nb - 0.5  # texture coord minus 1/2
    xc0 = a2 * w0 + b2 + 0.5 * nc - 0.5
    return a1, a2, xb0, xc0


def cone_fp(vol: np.ndarray, pgs, vg, slice_chunk: int = 64,
            axes=None) -> np.ndarray:
    """Joseph forward projection of ``vol[x, y, z]``; returns
    ``proj[angle, row, col]`` (float64).

    For every angle the ray axis ``a`` is fixed (``fp_major_axes``); each ray
    samples the volume at the centre of every slice ``i`` along ``a``
    (bilinear in the other two axes) and the sum is multiplied by
    ``outScale * sqrt(1 + a1^2 scale1 + a2^2 scale2)``
    (``cuda/cone_fp.cu:181-204``).

    ``axes`` (optional) overrides the per-angle ray axes.  The reference
    takes them from the detector centre (``kernels/cone_fp.py:264-285``), so
    a part of a detector (a window cut out to check a restriction) must be
    given the full detector's axes to sample the same way.
    """
    vol = np.asarray(vol, dtype=np.float64)
    shape = vol.shape
    if tuple(shape) != tuple(int(n) for n in vg.shape):
        raise ValueError("volume shape does not match the volume geometry")
    fr = normalized_frames(pgs, vg)
    axes = fp_major_axes(fr) if axes is None else np.asarray(axes)
    scales = fp_output_scale(vg.voxel_size)
    n = len(axes)
    out = np.zeros((n, fr["rows"], fr["cols"]))
    for k in range(n):
        axis = int(axes[k])
        a, b, c = _FRAMES[axis]
        va = np.transpose(vol, (a, b, c))      # va[i, j, l]: frame order
        a1, a2, xb0, xc0 = _ray_params(fr, k, axis, shape)
        acc = np.zeros_like(a1)
        na = shape[a]
        for i0 in range(0, na, slice_chunk):
            idx = np.arange(i0, min(na, i0 + slice_chunk), dtype=np.float64)
            xb = xb0[None] + a1[None] * idx[:, None, None]
            xc = xc0[None] + a2[None] * idx[:, None, None]
            for t, i in enumerate(range(i0, min(na, i0 + slice_chunk))):
                acc += bilinear_border_zero(va[i], xb[t], xc[t])
        sc1, sc2, osc = scales[axis]
        out[k] = acc * np.sqrt(a1 * a1 * sc1 + a2 * a2 * sc2 + 1.0) * osc
    return out


def cone_fp_transpose(proj: np.ndarray, pgs, vg, axes=None) -> np.ndarray:
    """Exact transpose of :func:`cone_fp`: ``vol[x, y, z]`` (float64).

    Every (ray, slice) sample of the forward operator contributes
    ``weight * tap_weight * proj[angle, row, col]`` to each of its in-bounds
    bilinear taps, with ``weight = outScale * sqrt(1 + a1^2 s1 + a2^2 s2)``.
    ``axes`` overrides the per-angle ray axes as in :func:`cone_fp`.
    """
    proj = np.asarray(proj, dtype=np.float64)
    shape = tuple(int(n) for n in vg.shape)
    fr = normalized_frames(pgs, vg)
    axes = fp_major_axes(fr) if axes is None else np.asarray(axes)
    scales = fp_output_scale(vg.voxel_size)
    out = np.zeros(shape)
    for k in range(len(axes)):
        axis = int(axes[k])
        a, b, c = _FRAMES[axis]
        na, nb, nc = shape[a], shape[b], shape[c]
        a1, a2, xb0, xc0 = _ray_params(fr, k, axis, shape)
        sc1, sc2, osc = scales[axis]
        wray = (np.sqrt(a1 * a1 * sc1 + a2 * a2 * sc2 + 1.0) * osc
                * proj[k]).ravel()
        acc = np.zeros(na * nb * nc)
        for i in range(na):
            xb = (xb0 + a1 * i).ravel()
            xc = (xc0 + a2 * i).ravel()
            j0 = np.floor(xb)
            l0 = np.floor(xc)
            fj = xb - j0
            fl = xc - l0
            j0 = j0.astype(np.int64)
            l0 = l0.astype(np.int64)
            for dj, wj in ((0, 1.0 - fj), (1, fj)):
                jj = j0 + dj
                for dl, wl in ((0, 1.0 - fl), (1, fl)):
                    ll = l0 + dl
                    ok = (jj >= 0) & (jj < nb) & (ll >= 0) & (ll < nc)
                    flat = (i * nb + jj[ok]) * nc + ll[ok]
                    acc += np.bincount(flat, weights=(wj * wl * wray)[ok],
                                       minlength=na * nb * nc)
        va = acc.reshape(na, nb, nc)
        inv = np.argsort((a, b, c))
        out += np.transpose(va, inv)
    return out


def cone_bp(proj: np.ndarray, pgs, vg, with_adjoint_scaling: bool = True,
            voxel_chunk: int = 1 << 20) -> np.ndarray:
    """Voxel-driven backprojection of ``proj[angle, row, col]``; returns
    ``vol[x, y, z]`` (float64).

    For voxel centre X (``cuda/cone_bp.cu:84-91``) and every angle:
    ``r = 1/den(X)``, ``u = numU(X) r``, ``v = numV(X) r`` (texture coords),
    ``acc += r^2 * bilinear(proj, v, u)`` (``:103-144``); finally the sum is
    multiplied by the voxel volume (``kernels/cone_bp.py:294``).
    """
    proj = np.asarray(proj, dtype=np.float64)
    shape = tuple(int(n) for n in vg.shape)
    fr = normalized_frames(pgs, vg)
    P = bp_params(fr, vg.voxel_size, with_adjoint_scaling)
    grids = [np.arange(n, dtype=np.float64) - 0.5 * n + 0.5 for n in shape]
    X, Y, Z = np.meshgrid(*grids, indexing="ij")
    X, Y, Z = X.ravel(), Y.ravel(), Z.ravel()
    acc = np.zeros(X.size)
    for lo in range(0, X.size, voxel_chunk):
        x = X[lo:lo + voxel_chunk]
        y = Y[lo:lo + voxel_chunk]
        z = Z[lo:lo + voxel_chunk]
        part = np.zeros(x.size)
        for k in range(len(P)):
            nu, nv, de = P[k]
            num_u = nu[3] + x * nu[0] + y * nu[1] + z * nu[2]
            num_v = nv[3] + x * nv[0] + y * nv[1] + z * nv[2]
            den = de[3] + x * de[0] + y * de[1] + z * de[2]
            r = 1.0 / den
            u = num_u * r
            v = num_v * r
            part += r * r * bilinear_border_zero(proj[k], v - 0.5, u - 0.5)
        acc[lo:lo + voxel_chunk] = part
    return acc.reshape(shape) * voxel_volume(vg)
