"""Linear operators built on a forward projector and a backprojector.

Drop-in for ``kernelkit/operator.py`` of the reference: ``BaseOperator``
(:10-33), ``ProjectorOperator`` with its ``.T`` adjoint wrapper (:36-146) and
``XrayTransform`` (:149-233).  Inputs and outputs are device arrays (torch
tensors, :class:`~kernelkit.DeviceArray` or ``__cuda_array_interface__``
arrays); an output the operator allocates is of the input's kind.
"""
from __future__ import annotations

from abc import ABC, abstractmethod
from typing import Any

from kernelkit.kernels.cone_fp import detector_shape
from kernelkit.projector import (BackProjector, BaseProjector,  # noqa: F401
                                 ForwardProjector)


def _zeros(shape, like=None):
    """A zero output of the kind of the input ``like`` (a torch tensor for
    torch inputs, else a :class:`~kernelkit.DeviceArray`)."""
    from kernelkit import device
    from kernelkit.kernel import default_device
    default_device()          # raises without a HIP device
    return device.zeros(shape, like=like)


def _operands(input, out, domain_shape, range_shape):
    """``(x, y, out)``: the input and output as arrays the projectors take
    (``__cuda_array_interface__`` arrays through
    :class:`~kernelkit.DeviceArray` views, host input copied to the
    device), and the output to return (allocated, of the input's kind, if
    not given)."""
    from kernelkit import device
    from kernelkit.kernel import as_device_array, is_device_array
    x = as_device_array(input)
    if tuple(x.shape) != tuple(domain_shape):
        raise ValueError(
            f"Input shape {tuple(x.shape)} does not match domain shape "
            f"{tuple(domain_shape)} of operator.")
    if out is None:
        out = _zeros(range_shape, x)
        return x, out, out
    return x, (device.asarray(out) if is_device_array(out) else out), out


class BaseOperator(ABC):
    """A linear operator with domain/range shapes and an adjoint."""

    @property
    @abstractmethod
    def domain_shape(self) -> tuple: ...

    @property
    @abstractmethod
    def range_shape(self) -> tuple: ...

    @abstractmethod
    def __call__(self, input, out=None): ...

    @property
    def T(self):
        raise NotImplementedError


class ProjectorOperator(BaseOperator, ABC):
    """``A : X -> Y`` from a forward projector; ``A.T`` uses the
    backprojector (``operator.py:36-146``)."""

    def __init__(self, projector: BaseProjector,
                 backprojector: BaseProjector):
        self.projector = projector
        self.backprojector = backprojector

    @property
    def domain_shape(self) -> tuple:
        vg = self.projector.volume_geometry
        return tuple(int(vg.shape[i]) for i in self.projector.volume_axes)

    @property
    def range_shape(self) -> tuple:
        pg = self.projector.projection_geometry
        full = (len(pg), *detector_shape(pg))
        return tuple(full[i] for i in self.projector.projection_axes)

    def __call__(self, input, out=None, additive=False):
        """``out (+)= A(input)``; allocates a zero output if none is
        given."""
        x, y, out = _operands(input, out, self.domain_shape,
                              self.range_shape)
        self.projector.volume = x
        self.projector.projections = y
        self.projector(additive=additive)
        return out

    @property
    def T(self):
        outer = self

        class _Adjoint(BaseOperator):
            """``A^* : Y -> X`` (the unmatched voxel-driven BP)."""

            def __init__(self):
                self.projector = outer.backprojector
                self.backprojector = outer.projector

            @property
            def domain_shape(self) -> tuple:
                return outer.range_shape

            @property
            def range_shape(self) -> tuple:
                return outer.domain_shape

            def __call__(self, input, out=None, additive=False):
                x, y, out = _operands(input, out, self.domain_shape,
                                      self.range_shape)
                self.projector.projections = x
                self.projector.volume = y
                self.projector(additive)
                return out

            @property
            def T(self):
                return outer

        if not hasattr(self, "_T"):
            self._T = _Adjoint()
        return self._T

    @property
    def matched_T(self):
        """``A^T : Y -> X``, the exact transpose of the ray-driven forward
        projection (no counterpart in the reference, whose ``.T`` is the
        unmatched voxel-driven backprojector).  ``<A x, y> = <x, A^T y>``
        up to float32 rounding; used for exact gradients."""
        outer = self

        class _MatchedTranspose(BaseOperator):
            @property
            def domain_shape(self) -> tuple:
                return outer.range_shape

            @property
            def range_shape(self) -> tuple:
                return outer.domain_shape

            def __call__(self, input, out=None, additive=False):
                x, y, out = _operands(input, out, self.domain_shape,
                                      self.range_shape)
                p = outer.projector
                p._require_geometry()
                p.kernel.transpose(x, p._proj_geom_seq, p.volume_geometry, y,
                                   additive=additive)
                return out

            @property
            def T(self):
                return outer

        if not hasattr(self, "_matched_T"):
            self._matched_T = _MatchedTranspose()
        return self._matched_T


class XrayTransform(ProjectorOperator):
    """The cone-beam X-ray transform (``operator.py:149-233``)."""

    def __init__(self, projection_geometry: Any, volume_geometry: Any,
                 volume_axes: tuple = (0, 1, 2),
                 projection_axes: tuple = (0, 1, 2),
                 use_toolbox: bool = False, fp_kwargs=None, bp_kwargs=None):
        if use_toolbox:
            raise NotImplementedError(
                "ASTRA Toolbox projectors are not available in the MI355X "
                "build (`use_toolbox=True`).")
        fp_kwargs = fp_kwargs or {}
        bp_kwargs = bp_kwargs or {}
        projector = ForwardProjector(volume_axes=volume_axes,
                                     projection_axes=projection_axes,
                                     **fp_kwargs)
        backprojector = BackProjector(volume_axes=volume_axes,
                                      projection_axes=projection_axes,
                                      **bp_kwargs)
        projector.volume_geometry = volume_geometry
        projector.projection_geometry = projection_geometry
        backprojector.volume_geometry = volume_geometry
        backprojector.projection_geometry = projection_geometry
        super().__init__(projector, backprojector)
