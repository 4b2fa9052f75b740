"""kernelkit -- MI355X (gfx950) cone-beam CT operators.

Drop-in for the public namespace of the reference's ``kernelkit/__init__.py``
(:1-18): operators, projectors, algorithms (``fp``, ``bp``, ``fdk``,
``sirt``), geometry, data helpers and the ``kernel``/``kernels``/
``processing`` modules.  Device arrays are PyTorch tensors on the HIP device,
any ``__cuda_array_interface__`` array, or :class:`DeviceArray` (the
package's own, allocated by the library: PyTorch is optional); compute runs
in hand-written HIP kernels reached through a ctypes C ABI.
"""
__version__ = "1.0.0-alpha.2+mi355x"

KERNELKIT_VERSION = __version__

from .geom.proj import *  # noqa: E402,F401,F403
from .geom.vol import *  # noqa: E402,F401,F403
from .geom import resolve_volume_geometry, suggest_volume_extent  # noqa: E402,F401
from .operator import *  # noqa: E402,F401,F403
from .projector import (BackProjector, BaseProjector, ForwardProjector,  # noqa: E402,F401
                        Projector, ProjectionGeometryNotSetError,
                        ProjectionsNotSetError, VolumeGeometryNotSetError,
                        VolumeNotSetError)
from .algo import fp, bp, fbp, fdk, sirt  # noqa: E402,F401
from .data import *  # noqa: E402,F401,F403
from .device import DeviceArray  # noqa: E402,F401
from . import device  # noqa: E402,F401
from . import kernel  # noqa: E402,F401
from . import kernels  # noqa: E402,F401
from . import experimental  # noqa: E402,F401
from . import processing  # noqa: E402,F401
