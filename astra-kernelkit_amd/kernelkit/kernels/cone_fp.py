"""Ray-driven cone-beam forward projection kernel (Joseph's method).

Drop-in for ``kernelkit/kernels/cone_fp.py`` of the reference
(``RayDrivenConeFP``, :18-363): same constructor, ``compile``,
``geoms2params``, ``set_params`` and ``__call__`` contract, backed by the
gfx950 kernel ``kk_cone_fp`` (one launch for all angles and slices).
"""
from __future__ import annotations

import copy

import numpy as np

from kernelkit import _native
from kernelkit import device as _dev
from kernelkit.geom import normalize_
from kernelkit.geom.proj import ProjectionGeometrySequence
from kernelkit.geom.vol import VolumeGeometry
from kernelkit.kernel import (BaseKernel, KernelMemoryUnpreparedException,
                              KernelNotCompiledException, Texture,
                              VolumeTexture, as_device_array,
                              is_device_array)


class FPParams(tuple):
    """The 12 per-angle arrays of :meth:`RayDrivenConeFP.geoms2params`
    (float64), plus ``ray_axes``: the per-angle major axis taken from the
    normalised detector centre."""

    ray_axes: np.ndarray | None = None


def detector_shape(projection_geometry) -> tuple:
    """(rows, cols) of the first angle's detector (the reference uses angle
    0 for every angle, ``kernels/cone_fp.py:199-200``)."""
    det = projection_geometry[0].detector
    return (int(np.asarray(det.rows).ravel()[0]),
            int(np.asarray(det.cols).ravel()[0]))


def _logical_strides(tensor, axes) -> list:
    """(angle, row, col) element strides of an array whose axis k holds
    logical axis ``axes[k]``."""
    st = [0, 0, 0]
    for k, ax in enumerate(axes):
        st[ax] = tensor.stride(k)
    return st


class RayDrivenConeFP(BaseKernel):
    """Ray-driven cone-beam forward projection kernel."""

    FLOAT_DTYPE = np.float32
    SUPPORTED_DTYPES = [np.float32]

    pixels_per_block = (1, 32, 32)
    slices_per_block = 16

    def __init__(self, pixels_per_block: tuple = None,
                 slices_per_block: int = None,
                 volume_axes: tuple = (0, 1, 2),
                 projection_axes: tuple = (0, 1, 2)):
        """Cone-beam forward projection kernel.

        ``volume_axes[k]`` is the world axis (x=0, y=1, z=2) of array axis
        ``k`` of the volume; ``projection_axes[k]`` is the logical axis
        (angle=0, row=1, col=2) of array axis ``k`` of the projections.
        ``pixels_per_block`` and ``slices_per_block`` are accepted for API
        compatibility; the gfx950 kernel processes all slices of a ray in
        one pass and picks its own tile shape.
        """
        super().__init__()
        if pixels_per_block is not None:
            self.pixels_per_block = tuple(pixels_per_block)
        if slices_per_block is not None:
            self.slices_per_block = int(slices_per_block)
        for name, ax in (("volume_axes", volume_axes),
                         ("projection_axes", projection_axes)):
            if sorted(ax) != [0, 1, 2]:
                raise ValueError(f"`{name}` must be a permutation of "
                                 f"(0, 1, 2), got {ax}.")
        self._vol_axs = tuple(volume_axes)
        self._proj_axs = tuple(projection_axes)
        self._params_are_set = False
        self._params = None
        self._uploaded_key = None

    @property
    def volume_axes(self) -> tuple:
        return self._vol_axs

    @property
    def projection_axes(self) -> tuple:
        return self._proj_axs

    def compile(self, max_projs: int | None = None):
        """Load the gfx950 library and create this kernel's plan.

        ``max_projs`` (optional) caps the number of angles like the
        reference's constant-memory size (``kernels/cone_fp.py:88-100``);
        by default there is no cap (tables live in device memory).
        """
        return self._compile({"max_projs": max_projs,
                              "volume_axes": self._vol_axs,
                              "projection_axes": self._proj_axs})

    @staticmethod
    def _geom_axes(projection_geometry: ProjectionGeometrySequence):
        """Per-angle ray axis ``argmax |source - detector|``
        (``kernels/cone_fp.py:264-285``)."""
        return np.argmax(np.abs(projection_geometry.source_position
                                - projection_geometry.detector_position),
                         axis=1).astype(np.int32)

    @staticmethod
    def _output_scale(voxel_size) -> list:
        """``[scale1, scale2, outScale]`` per ray axis
        (``kernels/cone_fp.py:287-300``, including its anisotropic table)."""
        v = [float(x) for x in voxel_size]
        if abs(v[0] / v[1] - 1) > 1e-4 or abs(v[0] / v[2] - 1) > 1e-4:
            return [[(v[1] / v[0]) ** 2, (v[2] / v[0]) ** 2, v[0] ** 2],
                    [(v[0] / v[1]) ** 2, (v[2] / v[1]) ** 2, v[0] * v[1]],
                    [(v[0] / v[2]) ** 2, (v[1] / v[2]) ** 2, v[0] * v[2]]]
        return [[1.0, 1.0, v[0]]] * 3

    @staticmethod
    def geoms2params(projection_geometry, volume_geometry: VolumeGeometry):
        """Per-angle kernel parameters in the volume's voxel frame.

        Returns 12 float64 arrays ``(src.x, src.y, src.z, corner.x, corner.y,
        corner.z, U.x, U.y, U.z, V.x, V.y, V.z)`` as in
        ``kernels/cone_fp.py:302-339``, where ``U``/``V`` are the column/row
        pixel steps.  Unlike the reference the input geometry is not
        normalised in place, and everything stays float64 (the library
        rounds to float32 only after shifting to the pixel-centre/source
        frame).
        """
        if not isinstance(projection_geometry, ProjectionGeometrySequence):
            projection_geometry = ProjectionGeometrySequence.fromList(
                projection_geometry)
        seq = copy.deepcopy(projection_geometry)
        normalize_(seq, volume_geometry)
        src = seq.source_position
        corner = seq.detector_extent_min
        U = seq.u * seq.detector.pixel_width[:, None]
        V = seq.v * seq.detector.pixel_height[:, None]
        out = FPParams(tuple(np.ascontiguousarray(a[:, k])
                             for a in (src, corner, U, V) for k in range(3)))
        out.ray_axes = RayDrivenConeFP._geom_axes(seq)
        return out

    def set_params(self, params, projection_geometry=None,
                   volume_geometry=None):
        """Store the per-angle parameters (``kernels/cone_fp.py:341-363``).

        The device table also depends on the volume shape and detector size;
        it is uploaded here when ``projection_geometry`` and
        ``volume_geometry`` are given, else at the first call.
        """
        if not self.is_compiled:
            raise KernelNotCompiledException(
                f"Compile {self.__class__.__name__} before setting its "
                "parameters.")
        if len(params) != 12:
            raise ValueError("Expected 12 parameter arrays.")
        arr = np.ascontiguousarray(
            np.stack([np.asarray(p, dtype=np.float64).ravel()
                      for p in params]))
        n = arr.shape[1]
        max_projs = self._compiled_template_kwargs.get("max_projs")
        if max_projs is not None and n > max_projs:
            raise ValueError(
                f"Number of projections, {n}, exceeds the maximum of the "
                f"compiled kernel, namely {max_projs}. Please recompile the "
                "kernel with a higher `max_projs`.")
        axes = getattr(params, "ray_axes", None)
        self._params = arr
        self._ray_axes = (None if axes is None
                          else np.ascontiguousarray(axes, dtype=np.int32))
        self._params_are_set = True
        self._uploaded_key = None
        if projection_geometry is not None and volume_geometry is not None:
            rows, cols = detector_shape(projection_geometry)
            self._upload(rows, cols, volume_geometry)

    def _upload(self, rows: int, cols: int, vg: VolumeGeometry) -> None:
        key = (rows, cols, tuple(vg.shape), tuple(vg.voxel_size))
        if key == self._uploaded_key:
            return
        n = self._params.shape[1]
        axes_ptr = (None if self._ray_axes is None
                    else _native.i32_ptr(self._ray_axes))
        _native.check(_native.lib().kk_set_params_fp(
            self._plan.ptr, _native.f64_ptr(self._params), n, axes_ptr,
            rows, cols, _native.i32(vg.shape), _native.f64(vg.voxel_size)))
        self._uploaded_key = key

    @property
    def frame_mask(self) -> int:
        """Frames of the volume texture the uploaded parameters use (bit a:
        ray axis a), 0 before the first upload."""
        if not self.is_compiled or self._uploaded_key is None:
            return 0
        return int(_native.lib().kk_fp_frame_mask(self._plan.ptr))

    def volume_texture(self, volume, volume_geometry: VolumeGeometry,
                       texture: VolumeTexture | None = None) -> VolumeTexture:
        """Build (or refresh) the forward-projection texture of ``volume``
        (in ``volume_axes`` order) for the uploaded parameters; before the
        first upload the frames of the x and y ray axes are built."""
        mask = self.frame_mask or 0b011
        shape = tuple(int(s) for s in volume_geometry.shape)
        volume = as_device_array(volume)
        if (texture is None or texture.world_shape != shape
                or not texture.covers(mask)):
            texture = VolumeTexture(shape, mask, like=volume)
        texture.update(volume, self._vol_axs)
        return texture

    def __call__(self, volume_texture, projection_geometry,
                 volume_geometry: VolumeGeometry, projections,
                 projections_pitch: tuple | None = None, *,
                 additive: bool = True, stream=None):
        """Forward-project the volume texture into ``projections``.

        ``projections`` is a device tensor in ``projection_axes`` order (row
        pitch allowed through its strides) or a list of 2-D tensors, one per
        index of the first axis.  Like the reference kernel the result is
        accumulated (``+=``, ``cuda/cone_fp.cu:201-204``) unless
        ``additive=False``.  ``projections_pitch`` is accepted for API
        compatibility; strides carry the pitch.
        """
        if volume_texture is None:
            raise ValueError("`volume_texture` must be given.")
        if not self.is_compiled:
            raise KernelNotCompiledException(
                "Please compile the kernel with `compile()`.")
        if not self._params_are_set:
            raise KernelMemoryUnpreparedException(
                "Please set the projection geometry with `set_params()`.")
        if not isinstance(volume_texture, (Texture, VolumeTexture)):
            raise TypeError("`volume_texture` must be a kernelkit texture.")
        n = self._params.shape[1]
        if len(projection_geometry) != n:
            raise ValueError(
                f"{len(projection_geometry)} projection geometries but "
                f"parameters for {n} angles.")
        rows, cols = detector_shape(projection_geometry)
        world = tuple(int(s) for s in volume_geometry.shape)
        if isinstance(volume_texture, Texture):
            expect_vol = tuple(world[i] for i in self._vol_axs)
            if tuple(volume_texture.shape) != expect_vol:
                raise ValueError(
                    f"Volume texture shape {volume_texture.shape} does not "
                    f"match volume geometry {world} with axes "
                    f"{self._vol_axs}.")
        elif volume_texture.world_shape != world:
            raise ValueError(
                f"Volume texture shape {volume_texture.world_shape} does not"
                f" match volume geometry {world}.")
        self._upload(rows, cols, volume_geometry)
        if isinstance(volume_texture, Texture):
            # a plain snapshot: convert it to this kernel's gather layout
            volume_texture = self._from_snapshot(volume_texture,
                                                 volume_geometry)
        elif not volume_texture.covers(self.frame_mask):
            raise ValueError(
                "The volume texture lacks frames for ray axes of this "
                f"geometry (has {volume_texture.mask:#b}, needs "
                f"{self.frame_mask:#b}); rebuild it after setting the "
                "geometry.")
        logical = (n, rows, cols)
        expect = tuple(logical[i] for i in self._proj_axs)
        if isinstance(projections, (list, tuple)):
            table, strides = self._pointer_table(projections, expect)
            self._launch_ptrs(volume_texture, table, strides, rows, cols,
                              additive, stream)
            return
        if not is_device_array(projections):
            raise TypeError("Projection data must be a device array or a "
                            f"list of them, got {type(projections)}.")
        projections = _dev.asarray(projections)
        if _dev.dtype_name(projections) == "int64":
            # the reference's form: a device array of entry base addresses
            # along the first projection axis, with the pitched logical
            # dimensions (kernels/cone_fp.py:173, :233-243)
            if projections.dim() != 1 or projections.numel() != expect[0]:
                raise ValueError(
                    f"A pointer array needs {expect[0]} entries (the first "
                    f"axis of `projection_axes` {self._proj_axs}), got shape "
                    f"{tuple(projections.shape)}.")
            pitch = logical if projections_pitch is None else tuple(
                int(p) for p in projections_pitch)
            strides = [0, 0, 0]
            strides[self._proj_axs[1]] = pitch[self._proj_axs[2]]
            strides[self._proj_axs[2]] = 1
            self._launch_ptrs(volume_texture, projections.contiguous(),
                              strides, rows, cols, additive, stream)
            return
        if _dev.dtype_name(projections) != "float32":
            raise NotImplementedError(
                "Currently there is only support for dtype float32, but got "
                f"{projections.dtype}.")
        if tuple(projections.shape) != expect:
            raise ValueError(
                f"Projections shape {tuple(projections.shape)} does not match "
                f"{expect} (angles, rows, cols permuted by "
                f"{self._proj_axs}).")
        self._launch(volume_texture, projections, rows, cols, additive,
                     stream)

    def _from_snapshot(self, snap: Texture, vg) -> VolumeTexture:
        key = (id(snap), snap.version, self.frame_mask)
        cached = getattr(self, "_snap_cache", None)
        if cached is not None and cached[0] == key:
            return cached[1]
        tex = self.volume_texture(snap.array, vg)
        self._snap_cache = (key, tex)
        return tex

    def _pointer_table(self, projections, expect):
        """Device table of the entries' base addresses (one per index of
        the first projection axis, ``projector.py:301-314``) and the
        (angle, row, col) element strides inside an entry.  The table is
        cached per list of addresses, so repeated calls (and graph
        capture) allocate nothing."""
        if len(projections) != expect[0]:
            raise ValueError(
                "The length of projection array must match the number of "
                f"entries along the first axis of `projection_axes` "
                f"{self._proj_axs}, {expect[0]}, but got "
                f"{len(projections)}.")
        for p in projections:
            if not is_device_array(p):
                raise TypeError("Projection data must be device arrays, "
                                f"but got {type(p)}.")
        projections = [_dev.asarray(p) for p in projections]
        for p in projections:
            if _dev.dtype_name(p) != "float32":
                raise NotImplementedError(
                    "Currently there is only support for dtype float32, "
                    f"but got {p.dtype}.")
            if tuple(p.shape) != expect[1:]:
                raise ValueError(
                    f"Projection entry of shape {tuple(p.shape)}, expected "
                    f"{expect[1:]} (axes {self._proj_axs}).")
            if tuple(p.stride()) != tuple(projections[0].stride()):
                raise ValueError("All projection entries need the same "
                                 "strides (row pitch).")
        ptrs = tuple(int(p.data_ptr()) for p in projections)
        if len(set(ptrs)) != len(ptrs):
            raise ValueError(
                "All projection files need to reside in own memory.")
        cached = getattr(self, "_ptr_cache", None)
        if cached is None or cached[0] != ptrs:
            host = np.array(ptrs, dtype=np.int64)
            if _dev.is_torch(projections[0]):
                torch = _dev._torch_module()
                table = torch.from_numpy(host).to(projections[0].device)
            else:
                table = _dev.from_host(host, dtype=np.int64)
            self._ptr_cache = (ptrs, table)
        strides = [0, 0, 0]
        strides[self._proj_axs[1]] = projections[0].stride(0)
        strides[self._proj_axs[2]] = projections[0].stride(1)
        return self._ptr_cache[1], strides

    def _launch_ptrs(self, tex: VolumeTexture, table, strides, rows, cols,
                     additive, stream):
        _native.check(_native.lib().kk_cone_fp_ptrs(
            self._plan.ptr, _native._vp(tex.ptr),
            _native._vp(table.data_ptr()), self._proj_axs[0],
            _native.i64(strides), rows, cols, 1 if additive else 0,
            _native.stream_ptr(stream)))

    def _launch(self, tex: VolumeTexture, proj, rows, cols, additive,
                stream):
        strides = _logical_strides(proj, self._proj_axs)
        _native.check(_native.lib().kk_cone_fp(
            self._plan.ptr, _native._vp(tex.ptr), _native._vp(proj.data_ptr()),
            _native.i64(strides), rows, cols, 1 if additive else 0,
            _native.stream_ptr(stream)))

    def transpose(self, projections, projection_geometry,
                  volume_geometry: VolumeGeometry, volume, *,
                  additive: bool = False, stream=None,
                  allocate: bool = True):
        """Exact transpose of this forward projection (matched adjoint):
        ``volume (+)= A^T projections``.

        No counterpart in the reference (its gradients use the voxel-driven
        backprojector, ``torch_support.py:29-33``).  ``projections`` is a
        device tensor in ``projection_axes`` order, ``volume`` one in
        ``volume_axes`` order.  Every (ray, slice) sample of :meth:`__call__`
        adds its weighted value to its bilinear taps, with weights and
        coordinates computed by the forward kernel's own functions (float64
        coordinates, bit-identical weights), so ``<A x, y> = <x, A^T y>`` up
        to rounding.  ``allocate=False`` uses the library's static
        workspace instead of allocating one (for callers that cannot
        allocate, e.g. under graph capture; limited to
        ``kk_cone_fp_transpose_static_capacity()`` floats of
        ``n_angles * rows * cols``, and calls that overlap on one device
        race on it).
        """
        if not self.is_compiled:
            raise KernelNotCompiledException(
                "Please compile the kernel with `compile()`.")
        if not self._params_are_set:
            raise KernelMemoryUnpreparedException(
                "Please set the projection geometry with `set_params()`.")
        for name, t in (("projections", projections), ("volume", volume)):
            if not is_device_array(t):
                raise TypeError(f"`{name}` must be a device array.")
            if _dev.dtype_name(t) != "float32":
                raise NotImplementedError(
                    "Currently there is only support for dtype float32, but "
                    f"got {t.dtype}.")
        projections, volume = _dev.asarray(projections), _dev.asarray(volume)
        n = self._params.shape[1]
        rows, cols = detector_shape(projection_geometry)
        logical = (n, rows, cols)
        expect = tuple(logical[i] for i in self._proj_axs)
        if tuple(projections.shape) != expect:
            raise ValueError(
                f"Projections shape {tuple(projections.shape)} does not match "
                f"{expect} (angles, rows, cols permuted by "
                f"{self._proj_axs}).")
        world = tuple(int(s) for s in volume_geometry.shape)
        expect_vol = tuple(world[i] for i in self._vol_axs)
        if tuple(volume.shape) != expect_vol:
            raise ValueError(
                f"Volume shape {tuple(volume.shape)} does not match volume "
                f"geometry {world} with axes {self._vol_axs}.")
        self._upload(rows, cols, volume_geometry)
        pst = _logical_strides(projections, self._proj_axs)
        vst = _logical_strides(volume, self._vol_axs)
        if not allocate:
            _native.check(_native.lib().kk_cone_fp_transpose_static(
                self._plan.ptr, _native._vp(projections.data_ptr()),
                _native.i64(pst), rows, cols, _native._vp(volume.data_ptr()),
                _native.i32(world), _native.i64(vst), 1 if additive else 0,
                _native.stream_ptr(stream)))
            return
        # the library's scratch: weighted projections, column-major
        n_work = int(_native.lib().kk_cone_fp_transpose_work_elems(
            self._plan.ptr, rows, cols))
        if n_work < 0:
            raise ValueError(f"bad detector size {rows} x {cols}")
        work = _dev.empty((n_work,), like=projections)
        if stream is not None and _dev.is_torch(work):
            work.record_stream(stream)
        _native.check(_native.lib().kk_cone_fp_transpose(
            self._plan.ptr, _native._vp(projections.data_ptr()),
            _native.i64(pst), rows, cols, _native._vp(volume.data_ptr()),
            _native.i32(world), _native.i64(vst), 1 if additive else 0,
            _native._vp(work.data_ptr()), _native.stream_ptr(stream)))
