"""Exact restrictions of a cone-beam problem (test infrastructure).

Every detector pixel of the forward projection is an independent ray sum,
and every voxel of the backprojection (and of the exact transpose) an
independent sum over angles.  So a GPU result on the full problem can be
checked against the oracle on a part of it:

* forward projection: a band of detector rows, as a detector of its own
  (same pixel grid);
* backprojection: a block of voxels as a volume of its own, with each
  angle's detector cut to the window the block projects into (plus a
  margin for the bilinear taps);
* exact transpose: the same, with the block widened by the support of the
  bilinear taps (a sample touches voxels within one voxel of its
  coordinates).

A detector window keeps the world position of every pixel it holds, so the
per-pixel arithmetic (rays, FDK pre-weight) is unchanged.  Two things a
window does move are carried over from the full detector: the ray axis of
each angle (chosen from the detector centre, ``kernels/cone_fp.py:264-285``:
a window near a corner can flip it on angles near 45 degrees), passed to
the oracle's forward projection and transpose as ``axes``, and the FDK
pre-weight's source to detector-centre distance (:func:`fdk`).
"""
from __future__ import annotations

import numpy as np

import kernelkit as kk
from kernelkit.geom.proj import Detector, ProjectionGeometry


def ray_axes(pgs, vg):
    """Per-angle ray axes of the full detectors (for the oracle's ``axes``
    on windows cut from them)."""
    import oracle
    return oracle.fp_major_axes(oracle.normalized_frames(pgs, vg))


def sub_detector(g, r0: int, c0: int, nr: int, nc: int):
    """Rows ``[r0, r0 + nr)`` and columns ``[c0, c0 + nc)`` of detector
    ``g`` as a detector geometry of their own."""
    rows, cols = int(g.detector.rows), int(g.detector.cols)
    ph, pw = float(g.detector.pixel_height), float(g.detector.pixel_width)
    v = np.asarray(g.v, np.float64)
    u = np.asarray(g.u, np.float64)
    shift = ((r0 + nr / 2.0 - rows / 2.0) * ph * v
             + (c0 + nc / 2.0 - cols / 2.0) * pw * u)
    return ProjectionGeometry(
        np.asarray(g.source_position, np.float64),
        np.asarray(g.detector_position, np.float64) + shift,
        Detector(nr, nc, ph, pw), u=g.u, v=g.v)


def row_band(g, r0: int, nr: int):
    """Rows ``[r0, r0 + nr)`` of detector ``g`` (all columns)."""
    return sub_detector(g, r0, 0, nr, int(g.detector.cols))


def sub_volume(vg, lo, n):
    """Voxels ``[lo, lo + n)`` (per world axis) of ``vg`` as a volume
    geometry of their own (no rotation)."""
    if any(abs(float(r)) > 0 for r in getattr(vg, "rotation", (0, 0, 0))):
        raise ValueError("rotated volumes are not restricted here")
    vs = np.asarray(vg.voxel_size, np.float64)
    emin = np.asarray(vg.extent_min, np.float64) + np.asarray(lo) * vs
    n = tuple(int(k) for k in n)
    return kk.VolumeGeometry(n, tuple(vs), tuple(emin),
                             tuple(emin + np.asarray(n) * vs))


def _project(g, X):
    """Detector index coordinates (row, col) of world points X (k x 3);
    pixel (r, c) has its centre at index coordinates (r, c)."""
    s = np.asarray(g.source_position, np.float64)
    d = np.asarray(g.detector_position, np.float64)
    u = np.asarray(g.u, np.float64)
    v = np.asarray(g.v, np.float64)
    n = np.cross(u, v)
    t = np.dot(d - s, n) / ((X - s) @ n)
    if np.any(t <= 0):
        raise ValueError("block behind the source")
    P = s + t[:, None] * (X - s) - d
    rows, cols = int(g.detector.rows), int(g.detector.cols)
    col = (P @ u) / (float(g.detector.pixel_width) * np.dot(u, u)) \
        + cols / 2.0 - 0.5
    row = (P @ v) / (float(g.detector.pixel_height) * np.dot(v, v)) \
        + rows / 2.0 - 0.5
    return row, col


def window(pgs, vg, lo, n, grow: float = 0.0, margin: int = 2):
    """Common window size ``(nr, nc)`` and per-angle origins ``[(r0, c0)]``
    covering, on every angle's detector, the projection of voxels
    ``[lo, lo + n)`` widened by ``grow`` voxels on every side, plus
    ``margin`` pixels, clipped to the detector."""
    vs = np.asarray(vg.voxel_size, np.float64)
    emin = np.asarray(vg.extent_min, np.float64)
    a = emin + (np.asarray(lo, np.float64) - grow) * vs
    b = emin + (np.asarray(lo, np.float64) + np.asarray(n) + grow) * vs
    corners = np.array([[x, y, z] for x in (a[0], b[0]) for y in (a[1], b[1])
                        for z in (a[2], b[2])])
    spans = []
    for g in pgs:
        row, col = _project(g, corners)
        spans.append((np.floor(row.min()) - margin,
                      np.ceil(row.max()) + margin,
                      np.floor(col.min()) - margin,
                      np.ceil(col.max()) + margin))
    spans = np.array(spans)
    rows, cols = int(pgs[0].detector.rows), int(pgs[0].detector.cols)
    nr = int(min(rows, (spans[:, 1] - spans[:, 0]).max() + 1))
    nc = int(min(cols, (spans[:, 3] - spans[:, 2]).max() + 1))
    origins = [(int(np.clip(s[0], 0, rows - nr)),
                int(np.clip(s[2], 0, cols - nc))) for s in spans]
    return nr, nc, origins


def fdk(part, pgs, geoms, sub, filter="ramlak"):
    """``oracle.fdk`` of the full-width row windows ``part`` (detectors
    ``geoms``, cut from ``pgs``) into the block ``sub``.  The pre-weight's
    numerator is the source to detector-centre distance
    (``processing.py:170-188``), which a window moves, so each angle's
    pre-weight is rescaled to the full detector's centre; the row filter
    and the backprojection are unchanged by a row window."""
    import oracle
    ratio = np.array([
        np.linalg.norm(np.subtract(g.detector_position, g.source_position))
        / np.linalg.norm(np.subtract(w.detector_position, w.source_position))
        for g, w in zip(pgs, geoms)])
    pw = oracle.preweight(part, geoms) * ratio[:, None, None]
    return oracle.cone_bp(oracle.filter_rows(pw, filter), geoms, sub)


def windowed(pgs, proj, nr, nc, origins):
    """Per-angle detector windows as geometries, and the matching parts of
    ``proj[angle, row, col]``."""
    geoms = [sub_detector(g, r0, c0, nr, nc)
             for g, (r0, c0) in zip(pgs, origins)]
    part = np.stack([proj[k, r0:r0 + nr, c0:c0 + nc]
                     for k, (r0, c0) in enumerate(origins)])
    return geoms, part
