"""Host-side argument checks of ``kernelkit.algo`` (no device calls).

``sirt``'s mask follows the reference's ``x *= mask`` (``algo.py:356-357``):
any shape that broadcasts to the volume is accepted, anything else is a
``ValueError`` raised before device work starts.
"""
import numpy as np
import pytest

kk = pytest.importorskip("kernelkit")
from kernelkit import algo, workloads  # noqa: E402


@pytest.mark.parametrize("mshape", [(16, 12, 10), (1, 1, 10), (12, 10),
                                    (10,), (16, 1, 1), (1, 12, 1), ()])
def test_mask_shapes_that_broadcast(mshape):
    algo._check_mask_shape(np.ones(mshape, dtype=np.float32), (16, 12, 10))


@pytest.mark.parametrize("mshape", [(16, 12), (2, 2), (16, 12, 11),
                                    (17, 12, 10), (2, 16, 12, 10),
                                    (16, 12, 10, 1)])
def test_mask_shapes_that_do_not_broadcast(mshape):
    with pytest.raises(ValueError, match="does not broadcast"):
        algo._check_mask_shape(np.ones(mshape, dtype=np.float32),
                               (16, 12, 10))


def test_sirt_rejects_a_bad_mask_before_any_device_work():
    # on a host without a device any device work would raise
    # NativeLibraryError; the mask check comes first
    pgs = workloads.circular_geometry(4, sod=50.0, odd=30.0, rows=6, cols=7,
                                      pixel=1.0)
    vg = kk.resolve_volume_geometry(shape=(8, 8, 6), voxel_size=1.0)
    y = np.zeros((4, 6, 7), dtype=np.float32)
    with pytest.raises(ValueError, match="does not broadcast"):
        kk.sirt(y, pgs, vg, iters=1, verbose=False,
                mask=np.ones((8, 8), dtype=np.float32))
