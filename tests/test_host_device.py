"""Host-side arithmetic of :class:`kernelkit.DeviceArray` views.

Wrapping a ``__cuda_array_interface__`` array and slicing, reshaping or
broadcasting it only computes pointers and strides: nothing here touches a
device, so it runs without one.
"""
import numpy as np
import pytest

kk = pytest.importorskip("kernelkit")
from kernelkit import device  # noqa: E402


class _FakeCai:
    def __init__(self, shape, typestr="<f4", ptr=1 << 20, strides=None):
        self.__cuda_array_interface__ = {
            "shape": shape, "typestr": typestr, "data": (ptr, False),
            "strides": strides, "version": 3}


def _host_twin(shape, strides_elems, item=4):
    """A numpy view with the same element strides over a large buffer, to
    compare offsets with."""
    base = np.zeros(1 << 16, dtype=np.float32 if item == 4 else np.int64)
    return np.lib.stride_tricks.as_strided(
        base, shape=shape, strides=tuple(s * item for s in strides_elems))


def _offset(view_np, base_np):
    return (view_np.__array_interface__["data"][0]
            - base_np.__array_interface__["data"][0]) // view_np.itemsize


@pytest.mark.parametrize("idx", [
    (slice(1, 5),), (2,), (Ellipsis, slice(3, 9, 2)), (1, slice(None), 4),
    (slice(0, 6, 3), Ellipsis), (-1, -2), (slice(2, 2),)])
def test_getitem_matches_numpy(idx):
    shape = (6, 7, 11)
    a = device.asarray(_FakeCai(shape))
    assert a._device is None            # no device query on wrap
    v = a[idx]
    ref_base = _host_twin(shape, a.stride())
    ref = ref_base[idx]
    assert v.shape == ref.shape
    assert v.stride() == tuple(s // 4 for s in ref.strides)
    if ref.size:        # numpy leaves empty views at offset 0
        assert v.data_ptr() - a.data_ptr() == 4 * _offset(ref, ref_base)
    assert v._device is None and v.base is a


def test_strided_interface_and_views():
    # a foreign non-contiguous array: every other row of a (8, 10) block
    a = device.asarray(_FakeCai((4, 10), strides=(80, 4), ptr=4096))
    assert a.stride() == (20, 1) and not a.is_contiguous()
    cai = a.__cuda_array_interface__
    assert cai["strides"] == (80, 4) and cai["data"] == (4096, False)
    with pytest.raises(ValueError):
        a.view(40)
    b = device.asarray(_FakeCai((3, 4, 5)))
    assert b.is_contiguous()
    assert b.__cuda_array_interface__["strides"] is None
    r = b.view(12, -1)
    assert r.shape == (12, 5) and r.stride() == (5, 1)
    e = b[0, 0].expand(2, 3, 5)
    assert e.shape == (2, 3, 5) and e.stride() == (0, 0, 1)
    with pytest.raises(ValueError):
        b.expand(3, 4, 6)
    i = device.asarray(_FakeCai((5,), typestr="<i8"))
    assert i.element_size() == 8 and i[2:].data_ptr() == i.data_ptr() + 16
    with pytest.raises(NotImplementedError):
        device.asarray(_FakeCai((5,), typestr="<f8"))
    with pytest.raises(IndexError):
        b[3]
    with pytest.raises(IndexError):
        b[::-1]


def test_broadcast_helpers_and_kinds():
    m = device.asarray(_FakeCai((1, 4, 5)))
    bm = device.broadcast_to(m, (3, 4, 5))
    assert bm.stride() == (0, 5, 1)
    assert device.dtype_name(bm) == "float32"
    assert device._is_device(m) and not device._is_device(np.zeros(3))
    assert kk.kernel.is_device_array(_FakeCai((2,)))


def test_use_stream_sets_the_default_stream():
    """Graph capture records the library's launches by making its capturing
    stream the default of ``stream_ptr`` for the duration of the capture;
    explicit streams still win and the default comes back afterwards."""
    import ctypes
    import threading

    from kernelkit import _native
    before = _native.stream_ptr().value
    with _native.use_stream(0x1000):
        assert _native.stream_ptr().value == 0x1000
        assert _native.stream_ptr(0x2000).value == 0x2000
        with _native.use_stream(ctypes.c_void_p(0x3000)):
            assert _native.stream_ptr().value == 0x3000
        assert _native.stream_ptr().value == 0x1000
        seen = []
        t = threading.Thread(target=lambda: seen.append(
            _native.stream_ptr().value))
        t.start()
        t.join()
        assert seen == [before]          # other threads are not affected
    assert _native.stream_ptr().value == before
    with pytest.raises(RuntimeError):
        with _native.use_stream(0x1000):
            raise RuntimeError("inside")
    assert _native.stream_ptr().value == before
