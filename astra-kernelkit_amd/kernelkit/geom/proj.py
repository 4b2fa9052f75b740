"""Projection geometry: one source, one flat detector, per angle.

Host-side geometry mirrors ``kernelkit/geom/proj.py`` of the reference
(``Detector`` :32-74, ``Beam`` :77-90, ``ProjectionGeometry`` :102-240,
``ProjectionGeometrySequence`` :243-410, shift/scale/rotate :413-556).

Differences that matter for the device path:

* every vector is stored in float64.  The reference stores float32 in the
  constructor (``geom/proj.py:145-149``); here the per-angle kernel tables are
  derived in float64 and cast to float32 only after the volume-centre shift,
  so the host keeps full precision until then;
* ``ProjectionGeometrySequence`` is numpy-only (no CuPy array module).

Conventions (``docs/user_guide/geometries.rst:38-59``): vectors are (x, y, z),
right-handed, z is the rotation axis; roll/pitch/yaw are extrinsic rotations
about x, y, z.  Pixel (0, 0) sits at the ``-u, -v`` corner of the detector
(``detector_extent_min``).
"""
from __future__ import annotations

import copy
from dataclasses import dataclass
from enum import Enum
from typing import Sequence

import numpy as np

ANGLES_CONVENTION = "sxyz"  # static x, then static y, then static z


def angles2mat(r: float, p: float, y: float) -> np.ndarray:
    """Rotation matrix for extrinsic roll (x), pitch (y), yaw (z) angles.

    Follows the static-frame 'sxyz' convention used by the reference
    (``geom/proj.py:9-29``): the matrix is ``Rz(y) @ Ry(p) @ Rx(r)``, so a
    vector is first rotated about x, then about the fixed y axis, then about
    the fixed z axis.
    """
    cr, sr = np.cos(r), np.sin(r)
    cp_, sp = np.cos(p), np.sin(p)
    cy, sy = np.cos(y), np.sin(y)
    rx = np.array([[1.0, 0.0, 0.0], [0.0, cr, -sr], [0.0, sr, cr]])
    ry = np.array([[cp_, 0.0, sp], [0.0, 1.0, 0.0], [-sp, 0.0, cp_]])
    rz = np.array([[cy, -sy, 0.0], [sy, cy, 0.0], [0.0, 0.0, 1.0]])
    return rz @ ry @ rx


@dataclass
class Detector:
    """Flat detector: ``rows`` x ``cols`` pixels of ``pixel_height`` x
    ``pixel_width`` (``geom/proj.py:32-74``)."""

    __slots__ = "rows", "cols", "pixel_height", "pixel_width"

    rows: int
    cols: int
    pixel_height: float
    pixel_width: float

    @property
    def shape(self):
        return self.rows, self.cols

    @property
    def width(self):
        return self.pixel_width * self.cols

    @property
    def height(self):
        return self.pixel_height * self.rows

    @property
    def pixel_volume(self):
        return self.pixel_width * self.pixel_height


class Beam(Enum):
    """Beam type (``geom/proj.py:77-90``)."""

    CONE = "cone"
    PARALLEL = "parallel"

    @classmethod
    def from_string(cls, value: str):
        value = value.lower()
        for beam in cls:
            if beam.value == value:
                return beam
        raise ValueError(f"Invalid beam type: '{value}'")


def process_beam_type(arg) -> Beam:
    if isinstance(arg, Beam):
        return arg
    if isinstance(arg, str):
        return Beam.from_string(arg)
    raise TypeError("Argument must be of type `str` or `Beam`.")


def _unit_vector(value, name: str) -> np.ndarray:
    arr = np.array(value, dtype=np.float64)
    if np.allclose(arr, 0.0):
        raise ValueError(f"`{name}` must not be zero.")
    if not np.allclose(np.linalg.norm(arr), 1.0):
        raise ValueError(f"`{name}` must be a unit vector. Consider normalizing"
                         f" `{name}` using `{name} /= np.linalg.norm({name})`.")
    return arr


class ProjectionGeometry:
    """Single cone-beam projection: source, detector centre, detector frame.

    Mirrors ``geom/proj.py:102-240``.  ``u`` points along detector columns
    (horizontal), ``v`` along detector rows (vertical); both must be unit
    vectors, as in the reference setters (``:171-216``).
    """

    xp = np

    def __init__(self, source_position, detector_position, detector: Detector,
                 u: Sequence = (0.0, 1.0, 0.0), v: Sequence = (0.0, 0.0, 1.0),
                 beam: Beam | str = Beam.CONE, array_module=None):
        if array_module is not None and array_module is not np:
            raise NotImplementedError("Only numpy host geometry is supported.")
        self.source_position = np.array(source_position, dtype=np.float64)
        self.detector_position = np.array(detector_position, dtype=np.float64)
        self.detector = detector
        self.u = u
        self.v = v
        self.beam = process_beam_type(beam)

    @property
    def u(self):
        return self._u

    @u.setter
    def u(self, value):
        self._u = _unit_vector(value, "u")

    @property
    def v(self):
        return self._v

    @v.setter
    def v(self, value):
        self._v = _unit_vector(value, "v")

    @property
    def detector_extent_min(self):
        """Corner of pixel (0, 0): ``d - v*h/2 - u*w/2`` (``:152-164``)."""
        return (self.detector_position
                - self.v * self.detector.height / 2
                - self.u * self.detector.width / 2)

    def __deepcopy__(self, memodict={}):
        return self.__class__(
            source_position=self.source_position.copy(),
            detector_position=self.detector_position.copy(),
            u=self.u.copy(), v=self.v.copy(),
            detector=Detector(rows=self.detector.rows, cols=self.detector.cols,
                              pixel_width=self.detector.pixel_width,
                              pixel_height=self.detector.pixel_height),
            beam=self.beam)

    def __str__(self):
        return (f"Source {self.source_position} "
                f"Detector {self.detector_position} "
                f"Horizontal vector: {self.u} "
                f"Vertical vector: {self.v} ")


@dataclass
class ProjectionGeometrySequence:
    """Structure-of-arrays view of a list of geometries (``:243-410``).

    Arrays are float64 ``(N, 3)`` vectors, int32 detector sizes and float64
    pixel sizes.
    """

    xp = np

    @dataclass
    class DetectorSequence:
        xp = np
        rows: np.ndarray
        cols: np.ndarray
        pixel_width: np.ndarray
        pixel_height: np.ndarray

        @property
        def height(self):
            return self.pixel_height * self.rows

        @property
        def width(self):
            return self.pixel_width * self.cols

        @property
        def pixel_volume(self):
            return self.pixel_width * self.pixel_height

    source_position: np.ndarray
    detector_position: np.ndarray
    u: np.ndarray
    v: np.ndarray
    detector: DetectorSequence
    beam: Beam

    def __len__(self):
        return len(self.source_position)

    @property
    def detector_extent_min(self):
        """Per-angle corner of pixel (0, 0), shape ``(N, 3)``."""
        return (self.detector_position
                - self.v * (self.detector.height[..., None] / 2)
                - self.u * (self.detector.width[..., None] / 2))

    @classmethod
    def fromList(cls, geometries, xp=None):
        if isinstance(geometries, ProjectionGeometrySequence):
            return copy.deepcopy(geometries)
        geometries = list(geometries)
        if len(geometries) == 0:
            raise ValueError("At least one projection geometry is required.")
        if not all(g.beam == geometries[0].beam for g in geometries):
            raise NotImplementedError(
                "All geometries in the list must have the same beam type.")

        def _arr(values, dtype):
            return np.ascontiguousarray(np.array(values, dtype=dtype))

        det = cls.DetectorSequence(
            rows=_arr([g.detector.rows for g in geometries], np.int32),
            cols=_arr([g.detector.cols for g in geometries], np.int32),
            pixel_width=_arr([g.detector.pixel_width for g in geometries],
                             np.float64),
            pixel_height=_arr([g.detector.pixel_height for g in geometries],
                              np.float64))
        return cls(
            source_position=_arr([g.source_position for g in geometries],
                                 np.float64),
            detector_position=_arr([g.detector_position for g in geometries],
                                   np.float64),
            u=_arr([g.u for g in geometries], np.float64),
            v=_arr([g.v for g in geometries], np.float64),
            detector=det, beam=geometries[0].beam)

    def take(self, indices):
        det = self.DetectorSequence(
            rows=np.atleast_1d(self.detector.rows[indices]),
            cols=np.atleast_1d(self.detector.cols[indices]),
            pixel_width=np.atleast_1d(self.detector.pixel_width[indices]),
            pixel_height=np.atleast_1d(self.detector.pixel_height[indices]))
        return ProjectionGeometrySequence(
            source_position=np.atleast_2d(self.source_position[indices]),
            detector_position=np.atleast_2d(self.detector_position[indices]),
            u=np.atleast_2d(self.u[indices]), v=np.atleast_2d(self.v[indices]),
            detector=det, beam=self.beam)

    def __getitem__(self, item):
        return self.take(item)

    def __iter__(self):
        for i in range(len(self)):
            yield self.geometry(i)

    def geometry(self, i: int) -> ProjectionGeometry:
        """Single ``ProjectionGeometry`` for angle ``i``."""
        return ProjectionGeometry(
            self.source_position[i], self.detector_position[i],
            Detector(int(self.detector.rows[i]), int(self.detector.cols[i]),
                     float(self.detector.pixel_height[i]),
                     float(self.detector.pixel_width[i])),
            u=self.u[i] / np.linalg.norm(self.u[i]),
            v=self.v[i] / np.linalg.norm(self.v[i]), beam=self.beam)

    def __deepcopy__(self, memodict={}):
        det = self.DetectorSequence(
            rows=self.detector.rows.copy(), cols=self.detector.cols.copy(),
            pixel_width=self.detector.pixel_width.copy(),
            pixel_height=self.detector.pixel_height.copy())
        return ProjectionGeometrySequence(
            source_position=self.source_position.copy(),
            detector_position=self.detector_position.copy(),
            u=self.u.copy(), v=self.v.copy(), detector=det, beam=self.beam)


def shift(geom, vector):
    """New geometry translated by ``vector`` (``:413-429``)."""
    geom = copy.deepcopy(geom)
    shift_(geom, vector)
    return geom


def shift_(geom, vector) -> None:
    """In-place translation by ``vector`` (``:432-442``)."""
    geom.source_position = geom.source_position + vector
    geom.detector_position = geom.detector_position + vector


def scale(geom, factor):
    """New geometry with every length divided by ``factor`` (``:445-477``)."""
    geom = copy.deepcopy(geom)
    scale_(geom, factor)
    return geom


def scale_(geom, factor) -> None:
    """In-place division of all lengths by ``factor`` (``:480-504``).

    ``factor`` may be a scalar or a per-axis 3-vector: the pixel step vectors
    ``u*pw`` and ``v*ph`` are divided component-wise and re-split into a
    unit direction and a length, so anisotropic factors tilt the detector
    frame exactly as the reference does.
    """
    factor = np.asarray(factor, dtype=np.float64)
    single = isinstance(geom, ProjectionGeometry)
    pw = np.asarray(geom.detector.pixel_width, dtype=np.float64)
    ph = np.asarray(geom.detector.pixel_height, dtype=np.float64)
    u_step = geom.u * (pw[..., None] if not single else pw) / factor
    v_step = geom.v * (ph[..., None] if not single else ph) / factor
    new_pw = np.linalg.norm(u_step, axis=-1)
    new_ph = np.linalg.norm(v_step, axis=-1)
    if single:
        geom.detector.pixel_width = float(new_pw)
        geom.detector.pixel_height = float(new_ph)
        geom.u = u_step / new_pw
        geom.v = v_step / new_ph
    else:
        geom.detector.pixel_width = new_pw
        geom.detector.pixel_height = new_ph
        geom.u = u_step / new_pw[..., None]
        geom.v = v_step / new_ph[..., None]
    geom.source_position = geom.source_position / factor
    geom.detector_position = geom.detector_position / factor


def rotate(geom, roll: float = 0.0, pitch: float = 0.0, yaw: float = 0.0):
    """New geometry actively rotated by ``R = angles2mat(roll, pitch, yaw)``
    (``:507-534``): every vector ``w`` becomes ``R @ w``."""
    ngeom = copy.deepcopy(geom)
    R = angles2mat(roll, pitch, yaw)
    if isinstance(ngeom, ProjectionGeometry):
        ngeom.source_position = R @ geom.source_position
        ngeom.detector_position = R @ geom.detector_position
        ngeom.u = R @ geom.u
        ngeom.v = R @ geom.v
    else:
        ngeom.source_position = geom.source_position @ R.T
        ngeom.detector_position = geom.detector_position @ R.T
        ngeom.u = geom.u @ R.T
        ngeom.v = geom.v @ R.T
    return ngeom


def rotate_(geom, roll: float = 0.0, pitch: float = 0.0,
            yaw: float = 0.0) -> None:
    """In-place rotation by the *inverse* of ``angles2mat(roll, pitch, yaw)``.

    The reference right-multiplies row vectors, ``w @ R`` (``:537-556``),
    which is ``R.T @ w``.  ``normalize_`` uses this to express the geometry
    in the frame of a rotated volume.
    """
    R = angles2mat(roll, pitch, yaw)
    geom.source_position = geom.source_position @ R
    geom.detector_position = geom.detector_position @ R
    geom.u = geom.u @ R
    geom.v = geom.v @ R
