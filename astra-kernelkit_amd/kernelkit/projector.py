"""Projectors: stateful holders of geometry, volume, projections and kernel.

Drop-in for ``kernelkit/projector.py`` of the reference: ``BaseProjector``
(:41-123), ``Projector`` (:126-215), ``ForwardProjector`` (:218-346) and
``BackProjector`` (:349-588), with the same properties, setters, exceptions
and ``__call__(additive=False)``.

Arrays are device arrays: PyTorch-ROCm tensors, :class:`~kernelkit.DeviceArray`
or any ``__cuda_array_interface__`` array (the reference uses CuPy arrays).  Setting
``volume`` (FP) or ``projections`` (BP) copies the data into a padded
texture, like the reference's CUDA-array copy; later in-place changes of the
source are only seen after setting it again.
"""
from __future__ import annotations

from abc import ABC, abstractmethod
from typing import Any

import numpy as np

from kernelkit.geom.proj import ProjectionGeometry, ProjectionGeometrySequence
from kernelkit.geom.vol import VolumeGeometry
from kernelkit.kernel import (BaseKernel, VolumeTexture, as_device_array,
                              is_device_array)
from kernelkit.kernels.cone_bp import VoxelDrivenConeBP
from kernelkit.kernels.cone_fp import RayDrivenConeFP


class ProjectionGeometryNotSetError(AttributeError):
    """Raised when the projection geometry is not set in the projector."""


class VolumeGeometryNotSetError(AttributeError):
    """Raised when the volume geometry is not set in the projector."""


class VolumeNotSetError(AttributeError):
    """Raised when the volume is not set in the projector."""


class ProjectionsNotSetError(AttributeError):
    """Raised when the projections are not set in the projector."""


class BaseProjector(ABC):
    """Interface of projectors (``projector.py:41-123``)."""

    @property
    @abstractmethod
    def projection_axes(self) -> tuple: ...

    @property
    @abstractmethod
    def volume_axes(self) -> tuple: ...

    @property
    @abstractmethod
    def projection_geometry(self): ...

    @projection_geometry.setter
    @abstractmethod
    def projection_geometry(self, value): ...

    @projection_geometry.deleter
    @abstractmethod
    def projection_geometry(self): ...

    @property
    @abstractmethod
    def volume_geometry(self): ...

    @volume_geometry.setter
    @abstractmethod
    def volume_geometry(self, value): ...

    @volume_geometry.deleter
    @abstractmethod
    def volume_geometry(self): ...

    @property
    @abstractmethod
    def volume(self): ...

    @volume.setter
    @abstractmethod
    def volume(self, value): ...

    @volume.deleter
    @abstractmethod
    def volume(self): ...

    @property
    @abstractmethod
    def projections(self): ...

    @projections.setter
    @abstractmethod
    def projections(self, value): ...

    @projections.deleter
    @abstractmethod
    def projections(self): ...

    @abstractmethod
    def __call__(self, additive: bool = False): ...


class Projector(BaseProjector, ABC):
    """Geometry -> kernel parameter plumbing (``projector.py:126-215``)."""

    def __init__(self, kernel: BaseKernel):
        self._K = kernel

    @property
    def kernel(self) -> BaseKernel:
        return self._K

    @property
    def projection_axes(self):
        return self._K.projection_axes

    @property
    def volume_axes(self):
        return self._K.volume_axes

    @property
    def projection_geometry(self):
        try:
            return self._proj_geom
        except AttributeError:
            raise ProjectionGeometryNotSetError(
                f"Projection geometry not set in '{self.__class__.__name__}'."
                " Please set the projection geometry using "
                "`projection_geometry`.")

    @projection_geometry.setter
    def projection_geometry(self, value):
        if not isinstance(value, ProjectionGeometrySequence):
            for p in value:
                if p.beam != value[0].beam:
                    raise NotImplementedError(
                        "All geometries in the list must have the same beam "
                        "type.")
            seq = ProjectionGeometrySequence.fromList(value)
        else:
            seq = value
        self._proj_geom = value
        self._proj_geom_seq = seq
        self._update_params()

    @projection_geometry.deleter
    def projection_geometry(self):
        del self._proj_geom
        del self._proj_geom_seq

    @property
    def volume_geometry(self):
        try:
            return self._vol_geom
        except AttributeError:
            raise VolumeGeometryNotSetError(
                f"Volume geometry not set in '{self.__class__.__name__}'. "
                "Please set the volume geometry using `volume_geometry`.")

    @volume_geometry.setter
    def volume_geometry(self, value: VolumeGeometry):
        self._vol_geom = value
        self._update_params()

    @volume_geometry.deleter
    def volume_geometry(self):
        del self._vol_geom

    def _update_params(self):
        seq = getattr(self, "_proj_geom_seq", None)
        vg = getattr(self, "_vol_geom", None)
        if seq is None or vg is None:
            return
        self._params = self._K.geoms2params(seq, vg)
        self._K.set_params(self._params, seq, vg)

    def _require_geometry(self):
        self.projection_geometry  # noqa: B018 -- raises if missing
        self.volume_geometry  # noqa: B018

    def _n_rows_cols(self):
        seq = self._proj_geom_seq
        return (len(seq), int(seq.detector.rows[0]),
                int(seq.detector.cols[0]))


class ForwardProjector(Projector):
    """Cone-beam forward projector (``projector.py:218-346``)."""

    def __init__(self, kernel: RayDrivenConeFP = None, volume_axes=(0, 1, 2),
                 projection_axes=(0, 1, 2)):
        K = (RayDrivenConeFP(volume_axes=volume_axes,
                             projection_axes=projection_axes)
             if kernel is None else kernel)
        super().__init__(K)
        if not self._K.is_compiled:
            self._K.compile()

    @property
    def volume(self):
        try:
            return self._vol_ref
        except AttributeError:
            raise VolumeNotSetError(
                "Volume not set. Use the `volume` setter before calling the "
                "projector.")

    @volume.setter
    def volume(self, value):
        """Set the volume: copied into the projector's padded texture."""
        if not is_device_array(value):
            raise TypeError("Volume must be a device array.")
        self._vol_ref = value
        self._build_texture()

    def _build_texture(self):
        value = self._vol_ref
        world = [0, 0, 0]
        for k, ax in enumerate(self.volume_axes):
            world[ax] = int(value.shape[k])
        mask = self._K.frame_mask or 0b011
        tex = getattr(self, "_texture", None)
        if (tex is None or tex.world_shape != tuple(world)
                or not tex.covers(mask)):
            tex = VolumeTexture(world, mask, like=as_device_array(value))
            self._texture = tex
        tex.update(value, self.volume_axes)

    @volume.deleter
    def volume(self):
        del self._vol_ref
        del self._texture

    @property
    def projections(self):
        try:
            return self._projs
        except AttributeError:
            raise ProjectionsNotSetError(
                "Projections not set. Use the `projections` setter before "
                "calling the projector.")

    @projections.setter
    def projections(self, value):
        if not (is_device_array(value) or isinstance(value, list)):
            raise TypeError("Projections must be a device array or `list`.")
        if isinstance(value, list):
            # in place: other libraries' arrays become views of their memory
            value = [as_device_array(p) if is_device_array(p) else p
                     for p in value]
            ptrs = [p.data_ptr() for p in value]
            if len(ptrs) != len(set(ptrs)):
                raise ValueError(
                    "All projection files need to reside in own memory.")
        self._projs = value

    def __call__(self, additive: bool = False):
        """Project the volume into the projections (overwrite unless
        ``additive``).

        If the geometry was set after the volume and needs ray-axis frames
        the texture lacks, the texture is rebuilt from the volume first.
        """
        self._require_geometry()
        if not hasattr(self, "_texture"):
            raise VolumeNotSetError(
                "Volume not set. Use the `volume` setter before calling the "
                "projector.")
        if not self._texture.covers(self._K.frame_mask):
            self._build_texture()
        self._K(self._texture, self._proj_geom_seq, self._vol_geom,
                self.projections, additive=additive)


class BackProjector(Projector):
    """Cone-beam backprojector (``projector.py:349-588``)."""

    TEXTURE_TYPES = "array", "layered", "array2d", "pitch2d"

    def __init__(self, kernel: VoxelDrivenConeBP = None,
                 texture_type: str = "layered", **kernel_kwargs):
        if kernel is not None and len(kernel_kwargs) > 0:
            raise ValueError(
                "If `kernel` is given, other keyword arguments cannot be "
                "passed to the kernel constructor. You passed "
                f"`{kernel_kwargs}`.")
        kernel_kwargs.pop("beam", None)
        K = VoxelDrivenConeBP(**kernel_kwargs) if kernel is None else kernel
        super().__init__(K)
        self._texture_type = texture_type.lower()
        if self._texture_type not in self.TEXTURE_TYPES:
            raise ValueError(
                "`TEXTURE_TYPE` not supported. Please choose one of the "
                f"following: {self.TEXTURE_TYPES}.")
        if self._texture_type == "layered" and K.projection_axes[0] != 0:
            # A CUDA layered texture needs the angle axis first; the padded
            # texture here does not, so serve it as a 3-D texture.
            self._texture_type = "array"
        type2texture = {
            "pitch2d": VoxelDrivenConeBP.InterpolationMethod.Tex2D,
            "array2d": VoxelDrivenConeBP.InterpolationMethod.Tex2D,
            "array": VoxelDrivenConeBP.InterpolationMethod.Tex3D,
            "layered": VoxelDrivenConeBP.InterpolationMethod.Tex2DLayered,
        }
        if not self._K.is_compiled:
            self._K.compile(texture=type2texture[self._texture_type])

    @property
    def volume(self):
        try:
            return self._vol
        except AttributeError:
            raise VolumeNotSetError(
                "No volume set. Please set the volume before calling the "
                "backprojector.")

    @volume.setter
    def volume(self, value):
        if hasattr(self, "_vol_geom"):
            expected = tuple(int(self._vol_geom.shape[i])
                             for i in self._K.volume_axes)
            if tuple(value.shape) != expected:
                raise ValueError(
                    f"Volume shape {tuple(value.shape)} does not match volume"
                    f" geometry {tuple(self._vol_geom.shape)} with axes "
                    f"{self._K.volume_axes}.")
        self._vol = value

    @volume.deleter
    def volume(self):
        del self._vol

    @property
    def projections(self):
        """The stored projections as an (angle, row, col) view of the
        texture (``projector.py:427-446``: the reference returns its CUDA
        array(s))."""
        try:
            return self._textures.values()
        except AttributeError:
            raise ProjectionsNotSetError("No projections set in the "
                                         "projector.")

    @projections.setter
    def projections(self, value: Any):
        """Copy projections into the padded texture (``:448-556``)."""
        if not (is_device_array(value) or isinstance(value, (np.ndarray,
                                                             list))):
            raise TypeError("Projections must be a `ndarray`, device tensor "
                            "or `list`.")
        pa = self._K.projection_axes
        if hasattr(self, "_proj_geom_seq"):
            shp = self._n_rows_cols()
            expected = tuple(shp[i] for i in pa)
            if isinstance(value, list):
                if len(value) != expected[0]:
                    raise ValueError(
                        f"Projections shape {len(value)} does not match "
                        f"projection geometry {shp} with axes {pa}.")
                if tuple(value[0].shape) != expected[1:]:
                    raise ValueError(
                        f"Projections shape {tuple(value[0].shape)} does not "
                        f"match projection geometry {shp} with axes {pa}.")
            elif tuple(value.shape) != expected:
                raise ValueError(
                    f"Projections shape {tuple(value.shape)} does not match "
                    f"projection geometry {shp} with axes {pa}.")
        if isinstance(value, list):
            if pa[0] != 0:
                raise ValueError("Can only use a list of projections when "
                                 "projection axis is 0.")
            from kernelkit import device
            value = device.stack([as_device_array(p) for p in value])
        elif self._texture_type == "layered" and pa[0] != 0:
            raise ValueError("Can only use layered texture when the first "
                             "projection axis is 0.")
        value = as_device_array(value)
        self._textures = self._K.projection_texture(
            value, getattr(self, "_textures", None))

    @projections.deleter
    def projections(self):
        if hasattr(self, "_textures"):
            del self._textures

    def __call__(self, additive: bool = False):
        """Backproject the projections into the volume (overwrite unless
        ``additive``)."""
        self._require_geometry()
        if not hasattr(self, "_textures"):
            raise ProjectionsNotSetError("No projections set in the "
                                         "projector.")
        self._K(self._textures, self.volume, self.volume_geometry,
                additive=additive)
