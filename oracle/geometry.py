"""FP64 geometry for the oracle: the reference's normalisation and the FP/BP
kernel parameters, restated with numpy (test infrastructure only).

References (paths relative to the reference repository):
* ``kernelkit/geom/proj.py:12-29``    roll/pitch/yaw -> matrix ('sxyz')
* ``kernelkit/geom/proj.py:152-164``, ``:317-325``  detector corner
* ``kernelkit/geom/proj.py:480-504``  scale_ (per-axis voxel size)
* ``kernelkit/geom/proj.py:537-556``  rotate_ (row vector @ R)
* ``kernelkit/geom/__init__.py:305-330``  normalize_
* ``kernelkit/kernels/cone_fp.py:264-300``, ``:302-339``  FP axes/scale/params
* ``kernelkit/kernels/cone_bp.py:342-421``  BP affine parameters
"""
from __future__ import annotations

import numpy as np


def rotation_sxyz(roll: float, pitch: float, yaw: float) -> np.ndarray:
    """``Rz(yaw) @ Ry(pitch) @ Rx(roll)``: extrinsic x, then y, then z
    (``geom/proj.py:9``, transforms3d 'sxyz')."""
    def rx(a):
        c, s = np.cos(a), np.sin(a)
        return np.array([[1, 0, 0], [0, c, -s], [0, s, c]], dtype=np.float64)

    def ry(a):
        c, s = np.cos(a), np.sin(a)
        return np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]], dtype=np.float64)

    def rz(a):
        c, s = np.cos(a), np.sin(a)
        return np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]], dtype=np.float64)

    return rz(yaw) @ ry(pitch) @ rx(roll)


def normalized_frames(pgs, vg) -> dict:
    """Per-angle geometry in the volume's voxel frame (``normalize_``).

    Returns float64 arrays: ``s`` source, ``dc`` detector centre, ``d0``
    corner of pixel (0, 0), ``U`` column step and ``V`` row step (all
    ``(N, 3)``), plus int ``rows``/``cols`` of angle 0 (the reference uses
    angle 0's detector for every angle, ``kernels/cone_fp.py:199-200``).
    Coordinates are volume-centred and measured in voxels, so the volume is
    ``[-n/2, n/2]`` per axis.
    """
    pgs = list(pgs)
    s = np.array([np.asarray(g.source_position, np.float64) for g in pgs])
    dc = np.array([np.asarray(g.detector_position, np.float64) for g in pgs])
    u = np.array([np.asarray(g.u, np.float64) for g in pgs])
    v = np.array([np.asarray(g.v, np.float64) for g in pgs])
    pw = np.array([float(g.detector.pixel_width) for g in pgs])
    ph = np.array([float(g.detector.pixel_height) for g in pgs])
    rows = int(pgs[0].detector.rows)
    cols = int(pgs[0].detector.cols)

    lo = np.asarray(vg.extent_min, np.float64)
    hi = np.asarray(vg.extent_max, np.float64)
    vs = np.asarray(vg.voxel_size, np.float64)
    centre = -(lo + hi) / 2
    # shift_ then scale_ (geom/__init__.py:328-329)
    s = (s + centre) / vs
    dc = (dc + centre) / vs
    U = (u * pw[:, None]) / vs
    V = (v * ph[:, None]) / vs
    # rotate_: row vectors times R (geom/proj.py:552-556)
    R = rotation_sxyz(*[float(r) for r in vg.rotation])
    s, dc, U, V = s @ R, dc @ R, U @ R, V @ R
    d0 = dc - V * (rows / 2) - U * (cols / 2)
    return dict(s=s, dc=dc, d0=d0, U=U, V=V, rows=rows, cols=cols)


def fp_major_axes(frames: dict) -> np.ndarray:
    """Per-angle ray axis: ``argmax_k |s_k - dc_k|`` in the voxel frame,
    detector centre, one choice per angle (``kernels/cone_fp.py:264-285``)."""
    return np.argmax(np.abs(frames["s"] - frames["dc"]), axis=1)


def fp_output_scale(voxel_size) -> list:
    """``(scale1, scale2, outScale)`` per ray axis
    (``kernels/cone_fp.py:287-300``), including its anisotropic table
    verbatim (x-axis uses ``v0**2``; z pairs ``(v0/v2)**2`` with slope 1)."""
    v = [float(x) for x in voxel_size]
    eps = 1e-4
    if abs(v[0] / v[1] - 1) > eps or abs(v[0] / v[2] - 1) > eps:
        return [[(v[1] / v[0]) ** 2, (v[2] / v[0]) ** 2, v[0] ** 2],
                [(v[0] / v[1]) ** 2, (v[2] / v[1]) ** 2, v[0] * v[1]],
                [(v[0] / v[2]) ** 2, (v[1] / v[2]) ** 2, v[0] * v[2]]]
    return [[1.0, 1.0, v[0]]] * 3


def bp_params(frames: dict, voxel_size, with_adjoint_scaling: bool = True):
    """Per-angle affine forms of the voxel-driven BP, shape ``(N, 3, 4)``.

    Row 0 is the column numerator, row 1 the row numerator, row 2 the
    denominator; each row is ``(x, y, z, w)`` so that for a voxel centre X
    (voxel frame) ``form(X) = w + X . xyz``.  Then ``u = numU/den`` and
    ``v = numV/den`` are the detector coordinates of the ray source->X in
    pixel steps from the corner (texel-centre convention), and ``1/den**2``
    is the adjoint weight.  Restates ``kernels/cone_bp.py:350-418``:

    * ``numU = scale * ( s.(V x d0) + X.(V x (s - d0)) )``
    * ``numV = scale * ( -s.(U x d0) + X.((s - d0) x U) )``
    * ``den  = scale * ( s.(U x V) - X.(U x V) )``
    * ``scale = sqrt(|(U x V) * (vy vz, vx vz, vx vy)|) / det[U; V; d0 - s]``
      (without adjoint scaling ``scale = 1``).
    """
    s, d0, U, V = frames["s"], frames["d0"], frames["U"], frames["V"]
    vs = [float(x) for x in voxel_size]
    cr = np.cross(U, V) * np.array([vs[1] * vs[2], vs[0] * vs[2],
                                    vs[0] * vs[1]])
    if with_adjoint_scaling:
        scale = np.sqrt(np.linalg.norm(cr, axis=1))
        scale = scale / np.linalg.det(np.stack([U, V, d0 - s], axis=1))
    else:
        scale = np.ones(len(s))
    smd = s - d0
    out = np.empty((len(s), 3, 4))
    out[:, 0, :3] = scale[:, None] * np.cross(V, smd)
    out[:, 0, 3] = scale * np.einsum("ij,ij->i", s, np.cross(V, d0))
    out[:, 1, :3] = scale[:, None] * np.cross(smd, U)
    out[:, 1, 3] = -scale * np.einsum("ij,ij->i", s, np.cross(U, d0))
    uxv = np.cross(U, V)
    out[:, 2, :3] = -scale[:, None] * uxv
    out[:, 2, 3] = scale * np.einsum("ij,ij->i", s, uxv)
    return out


def voxel_volume(vg) -> float:
    """``prod((extent_max - extent_min) / shape)`` (``geom/vol.py:79-89``)."""
    lo = np.asarray(vg.extent_min, np.float64)
    hi = np.asarray(vg.extent_max, np.float64)
    n = np.asarray(vg.shape, np.float64)
    return float(np.prod((hi - lo) / n))
