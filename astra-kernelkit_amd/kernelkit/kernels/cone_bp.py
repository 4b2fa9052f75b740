"""Voxel-driven cone-beam backprojection kernel.

Drop-in for ``kernelkit/kernels/cone_bp.py`` of the reference
(``VoxelDrivenConeBP``, :23-421): same constructor, ``compile``,
``geoms2params``, ``set_params`` and ``__call__`` contract, backed by the
gfx950 kernel ``kk_cone_bp`` (one launch for all angles).
"""
from __future__ import annotations

import copy
from enum import Enum

import numpy as np

from kernelkit import _native
from kernelkit.geom import normalize_
from kernelkit.geom.proj import ProjectionGeometrySequence
from kernelkit.geom.vol import VolumeGeometry
from kernelkit import device as _dev
from kernelkit.kernel import (BaseKernel, KernelMemoryUnpreparedException,
                              KernelNotCompiledException, ProjectionTexture,
                              Texture, as_device_array, is_device_array)


def _world_strides(tensor, volume_axes) -> list:
    """Element strides of world axes (x, y, z) of an array whose axis k is
    world axis ``volume_axes[k]``."""
    st = [0, 0, 0]
    for k, ax in enumerate(volume_axes):
        st[ax] = tensor.stride(k)
    return st


class VoxelDrivenConeBP(BaseKernel):
    """Voxel-driven cone-beam backprojection kernel."""

    FLOAT_DTYPE = np.float32
    SUPPORTED_DTYPES = [np.float32]

    class InterpolationMethod(Enum):
        """Texture fetch variants of the reference (``kernels/cone_bp.py:
        29-41``).  All map to the same padded-texture gather here."""

        Tex3D = "tex3D"
        Tex2DLayered = "tex2DLayered"
        Surf2DLayered = "surf2DLayered"
        Tex2D = "tex2D"

    voxels_per_block = (6, 32, 16)
    projs_per_block = 32

    def __init__(self, voxels_per_block: tuple = None,
                 projs_per_block: int = None,
                 volume_axes: tuple = (0, 1, 2),
                 projection_axes: tuple = (0, 1, 2)):
        """Cone-beam backprojection kernel.

        ``voxels_per_block`` and ``projs_per_block`` are accepted for API
        compatibility (the reference's CUDA tiling); the gfx950 kernel runs
        all angles in one launch with its own tiling.
        """
        super().__init__()
        if voxels_per_block is not None and len(voxels_per_block) != 3:
            raise ValueError("`voxels_per_block` must have 3 elements.")
        for name, ax in (("volume_axes", volume_axes),
                         ("projection_axes", projection_axes)):
            if sorted(ax) != [0, 1, 2]:
                raise ValueError(f"`{name}` must be a permutation of "
                                 f"(0, 1, 2), got {ax}.")
        self._vox_block = tuple(voxels_per_block or self.voxels_per_block)
        self._projs_block = int(projs_per_block or self.projs_per_block)
        self._vol_axs = tuple(volume_axes)
        self._proj_axs = tuple(projection_axes)
        self._params_are_set = False
        self._n_params = 0

    @property
    def volume_axes(self):
        return self._vol_axs

    @property
    def projection_axes(self):
        return self._proj_axs

    def compile(self, max_projs: int | None = None,
                texture: "InterpolationMethod | str | None" =
                InterpolationMethod.Tex3D):
        """Load the gfx950 library and create this kernel's plan.

        Validates ``texture`` against ``projection_axes`` like the reference
        (``kernels/cone_bp.py:134-153``).
        """
        if texture is not None:
            if isinstance(texture, str):
                texture = self.InterpolationMethod(texture)
            if texture not in self.InterpolationMethod:
                raise ValueError(
                    f"`texture` must be one of {self.InterpolationMethod} or "
                    "`None`.")
            if (texture in (self.InterpolationMethod.Tex2D,
                            self.InterpolationMethod.Tex2DLayered)
                    and self._proj_axs[0] != 0):
                raise ValueError(
                    "If an array of texture pointers or a layered texture is"
                    " given, the first `projection_axes` axis must be angles,"
                    " i.e., must be 0.")
        self._compile({"max_projs": max_projs,
                       "texture": None if texture is None else texture.value,
                       "volume_axes": self._vol_axs,
                       "projection_axes": self._proj_axs})

    def set_params(self, params, projection_geometry=None,
                   volume_geometry=None):
        """Upload per-angle parameters (``kernels/cone_bp.py:168-190``).

        ``params`` is the flat ``12 N`` array of :meth:`geoms2params`; the
        extra arguments are accepted for symmetry with the FP kernel.
        """
        if not self.is_compiled:
            raise KernelNotCompiledException(
                f"Kernel {self.__class__.__name__} must be compiled before "
                "uploading parameters to the GPU.")
        arr = np.ascontiguousarray(np.asarray(params, dtype=np.float64).ravel())
        if arr.size % 12 != 0 or arr.size == 0:
            raise ValueError("Parameters must hold 12 values per angle.")
        n = arr.size // 12
        max_projs = self._compiled_template_kwargs.get("max_projs")
        if max_projs is not None and n > max_projs:
            raise ValueError(
                f"Number of projections, {n}, exceeds the maximum of the "
                f"compiled kernel, namely {max_projs}. Please recompile the "
                "kernel with a higher `max_projs`.")
        _native.check(_native.lib().kk_set_params_bp(
            self._plan.ptr, _native.f64_ptr(arr), n))
        self._n_params = n
        self._params_are_set = True

    def __call__(self, textures, volume, volume_geometry: VolumeGeometry, *,
                 additive: bool = True, stream=None):
        """Backproject the projection texture into ``volume`` (in place).

        ``textures`` is a :class:`~kernelkit.kernel.ProjectionTexture`, or a
        plain :class:`~kernelkit.kernel.Texture` snapshot of the projection
        array (in ``projection_axes`` order), which is converted on use.
        Like the reference kernel the result is accumulated
        (``cuda/cone_bp.cu:156-158``) unless ``additive=False``.
        """
        if not is_device_array(volume):
            raise TypeError("`volume` must be a device array.")
        if _dev.dtype_name(volume) != "float32":
            raise NotImplementedError(
                f"Volume has a dtype '{volume.dtype}'. However, currently"
                " there is only support for dtype=float32.")
        volume = _dev.asarray(volume)
        if not volume_geometry.has_isotropic_voxels():
            raise NotImplementedError(
                f"`{self.__class__.__name__}` is not tested with anisotropic "
                "voxels yet.")
        if not self.is_compiled:
            raise KernelNotCompiledException(
                "Please compile the kernel with `compile()`.")
        if not isinstance(textures, (Texture, ProjectionTexture)):
            raise ValueError(
                "Please give in a kernelkit texture of the projections, given "
                f"is a '{type(textures).__name__}'.")
        if not self._params_are_set:
            raise KernelMemoryUnpreparedException(
                "Please set the parameters with `set_params()`.")
        if isinstance(textures, Texture):
            textures = self._from_snapshot(textures)
        n, rows, cols = textures.shape
        expect = tuple(int(volume_geometry.shape[i]) for i in self._vol_axs)
        if tuple(volume.shape) != expect:
            raise ValueError(
                f"Volume shape {tuple(volume.shape)} does not match volume "
                f"geometry {tuple(volume_geometry.shape)} with axes "
                f"{self._vol_axs}.")
        _native.check(_native.lib().kk_cone_bp(
            self._plan.ptr, _native._vp(textures.ptr), n, rows, cols,
            _native._vp(volume.data_ptr()), _native.i32(volume_geometry.shape),
            _native.i64(_world_strides(volume, self._vol_axs)),
            float(volume_geometry.voxel_volume()), 1 if additive else 0,
            _native.stream_ptr(stream)))

    def projection_texture(self, projections,
                           texture: ProjectionTexture | None = None
                           ) -> ProjectionTexture:
        """Build (or refresh) the backprojection texture of projections
        given in ``projection_axes`` order."""
        projections = as_device_array(projections)
        logical = [0, 0, 0]
        for k, ax in enumerate(self._proj_axs):
            logical[ax] = projections.shape[k]
        if texture is None or texture.shape != tuple(logical):
            texture = ProjectionTexture(*logical, like=projections)
        texture.update(projections, self._proj_axs)
        return texture

    def _from_snapshot(self, snap: Texture) -> ProjectionTexture:
        key = (id(snap), snap.version)
        cached = getattr(self, "_snap_cache", None)
        if cached is not None and cached[0] == key:
            return cached[1]
        tex = self.projection_texture(snap.array)
        self._snap_cache = (key, tex)
        return tex

    @staticmethod
    def geoms2params(projection_geometry, volume_geometry: VolumeGeometry,
                     with_adjoint_scaling: bool = True):
        """Per-angle affine forms of the voxel-driven kernel, float64.

        Returns a flat ``12 N`` array, per angle ``numU.xyzw, numV.xyzw,
        den.xyzw`` (``kernels/cone_bp.py:304-421``).  With ``s`` the source,
        ``d`` the detector corner and ``U``/``V`` the pixel steps in the
        volume's voxel frame:

        * ``numU(X) = k (det[s, V, d] + X . (V x (s - d)))``,
        * ``numV(X) = k (-det[s, U, d] + X . ((s - d) x U))``,
        * ``den(X)  = k (det[U, V, s] - X . (U x V))``,

        so ``numU/den`` and ``numV/den`` are the detector coordinates of the
        ray through voxel centre X, and ``1/den^2`` the adjoint weight with
        ``k = sqrt(|(U x V) * (vy vz, vx vz, vx vy)|) / det[U, V, d - s]``
        (``k = 1`` without adjoint scaling).
        """
        if isinstance(projection_geometry, ProjectionGeometrySequence):
            seq = copy.deepcopy(projection_geometry)
        else:
            seq = ProjectionGeometrySequence.fromList(projection_geometry)
        normalize_(seq, volume_geometry)
        vs = np.asarray(volume_geometry.voxel_size, dtype=np.float64)
        U = seq.u * seq.detector.pixel_width[:, None]
        V = seq.v * seq.detector.pixel_height[:, None]
        s = seq.source_position
        d = seq.detector_extent_min
        uxv = np.cross(U, V)
        if with_adjoint_scaling:
            cr = uxv * np.array([vs[1] * vs[2], vs[0] * vs[2], vs[0] * vs[1]])
            k = np.sqrt(np.linalg.norm(cr, axis=1))
            k = k / np.linalg.det(np.stack([U, V, d - s], axis=1))
        else:
            k = np.ones(len(s))

        def det3(a, b, c):
            return np.einsum("ij,ij->i", a, np.cross(b, c))

        smd = s - d
        out = np.empty((len(s), 12))
        out[:, 0:3] = k[:, None] * np.cross(V, smd)
        out[:, 3] = k * det3(s, V, d)
        out[:, 4:7] = k[:, None] * np.cross(smd, U)
        out[:, 7] = -k * det3(s, U, d)
        out[:, 8:11] = -k[:, None] * uxv
        out[:, 11] = k * det3(U, V, s)
        return out.ravel()
