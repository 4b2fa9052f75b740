from .cone_fp import RayDrivenConeFP  # noqa: F401
from .cone_bp import VoxelDrivenConeBP  # noqa: F401
