"""Pitched-row helpers on host arrays (``kernelkit/data.py:8-115`` of the
reference): rows padded to a 32-byte multiple, views in the original
shape, the padded base kept as the view's base."""
import numpy as np
import pytest

kk = pytest.importorskip("kernelkit")


@pytest.mark.parametrize("dtype,n,expect", [
    (np.float32, 37, 40), (np.float32, 40, 40), (np.float32, 1, 8),
    (np.float64, 37, 40), (np.float64, 4, 4), (np.int8, 33, 64),
    (np.uint16, 17, 32)])
def test_pitched_shape(dtype, n, expect):
    a = np.zeros((3, 5, n), dtype=dtype)
    assert kk.pitched_shape(a) == (3, 5, expect)


def test_aspitched_round_trip_and_base():
    rng = np.random.default_rng(0)
    a = rng.uniform(size=(4, 6, 37)).astype(np.float32)
    assert not kk.ispitched(a)
    p = kk.aspitched(a)
    assert p.shape == a.shape and p.base is not None
    assert p.base.shape == (4, 6, 40) and kk.ispitched(p)
    np.testing.assert_array_equal(p, a)
    assert np.all(p.base[..., 37:] == 0)
    # already pitched: returned as is
    assert kk.aspitched(p) is p
    assert p.strides == (6 * 40 * 4, 40 * 4, 4)


def test_pitched_like():
    p = kk.aspitched(np.ones((2, 3, 37), dtype=np.float32))
    q = kk.pitched_like(p, fill=2.5)
    assert q.shape == p.shape and kk.ispitched(q)
    assert q.base is not p.base
    assert np.all(q == 2.5)
    with pytest.raises(ValueError, match="pitched"):
        kk.pitched_like(np.ones((2, 3, 37), dtype=np.float32))
