"""Experimental helpers (``kernelkit/experimental.py:1`` of the reference)."""
from kernelkit.geom import suggest_volume_extent  # noqa: F401
