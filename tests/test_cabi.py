"""The C ABI on the CPU: the library loads, exports every function the
header declares (and the ctypes binding names no other), and its host-only
entry points validate arguments and report errors through kk_last_error
and the Python exception classes.  No device calls.
"""
import ctypes
import os
import re

import numpy as np
import pytest

kk = pytest.importorskip("kernelkit")
from kernelkit import _native  # noqa: E402

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "kernelkit_hip.h")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)   # drop comments
    return set(re.findall(
        r"^\s*(?:const\s+)?(?:int|int32_t|int64_t|char|void)\s*\**\s*"
        r"(kk_\w+)\s*\(", text, flags=re.M))


@pytest.fixture(scope="module")
def lib():
    try:
        return _native.lib()
    except _native.NativeLibraryError as exc:
        pytest.skip(f"library not built: {exc}")


def test_header_and_binding_agree():
    decl = _declared()
    assert len(decl) >= 25
    assert decl == set(_native.exported_symbols())


def test_library_exports_every_declared_symbol(lib):
    raw = ctypes.CDLL(str(_native.LIB_PATH))
    missing = [n for n in sorted(_declared()) if not hasattr(raw, n)]
    assert not missing, missing


def test_version_and_last_error(lib):
    assert lib.kk_version().decode().startswith("kernelkit-hip")
    assert isinstance(lib.kk_last_error(), bytes)


def test_status_codes_map_to_reference_exceptions():
    from kernelkit.kernel import KernelMemoryUnpreparedException
    _native.check(_native.KK_OK)
    for code, exc in ((_native.KK_ERR_INVALID, ValueError),
                      (_native.KK_ERR_SHAPE, ValueError),
                      (_native.KK_ERR_UNSUPPORTED, NotImplementedError),
                      (_native.KK_ERR_NOT_SET,
                       KernelMemoryUnpreparedException),
                      (_native.KK_ERR_HIP, _native.KernelError)):
        try:
            _native.lib()
        except _native.NativeLibraryError:
            pytest.skip("library not built")
        with pytest.raises(exc):
            _native.check(code)


def test_texture_sizes(lib):
    H = _native.HALO
    shape = (4, 5, 6)
    # frame a holds n_a slices of (n_b + halo) x (n_c + halo) float2 texels
    fx = 4 * (5 + H) * (6 + H)
    fy = 5 * (4 + H) * (6 + H)
    fz = 6 * (5 + H) * (4 + H)
    sizes = {0b001: fx, 0b010: fy, 0b100: fz, 0b011: fx + fy,
             0b111: fx + fy + fz}
    for mask, texels in sizes.items():
        assert lib.kk_fp_texture_elems(_native.i32(shape), mask) == 2 * texels
    assert lib.kk_fp_texture_elems(_native.i32(shape), 0) == -1
    assert lib.kk_fp_texture_elems(_native.i32((0, 5, 6)), 1) == -1
    assert lib.kk_bp_texture_elems(3, 7, 9) == 2 * 3 * (7 + H) * (9 + H)
    assert lib.kk_bp_texture_elems(0, 7, 9) == -1


def _fp_params(n=3):
    from kernelkit import workloads
    from kernelkit.kernels.cone_fp import RayDrivenConeFP
    pgs = workloads.circular_geometry(n, sod=40.0, odd=20.0, rows=6, cols=7,
                                      pixel=1.0)
    vg = kk.resolve_volume_geometry(shape=(5, 6, 7), voxel_size=1.0)
    K = RayDrivenConeFP()
    params = K.geoms2params(kk.ProjectionGeometrySequence.fromList(pgs), vg)
    return np.ascontiguousarray(np.stack(params[:12]), dtype=np.float64), vg


def test_fp_table_validation(lib):
    P, vg = _fp_params()
    out = np.zeros(3, dtype=_native.FP_ANGLE_DTYPE)
    mask = ctypes.c_int32(0)
    shape = _native.i32(vg.shape)
    vs = _native.f64(vg.voxel_size)
    ok = lib.kk_fp_table(_native.f64_ptr(P), 3, None, 6, 7, shape, vs,
                         out.ctypes.data_as(ctypes.c_void_p),
                         ctypes.byref(mask))
    assert ok == _native.KK_OK
    assert mask.value in (1, 2, 3)
    assert set(out["axis"]) <= {0, 1}
    bad = [
        dict(n=0), dict(rows=0), dict(shape=(0, 6, 7)),
        dict(vs=(1.0, -1.0, 1.0)), dict(out=None),
    ]
    for b in bad:
        rc = lib.kk_fp_table(
            _native.f64_ptr(P), b.get("n", 3), None, b.get("rows", 6), 7,
            _native.i32(b.get("shape", vg.shape)),
            _native.f64(b.get("vs", vg.voxel_size)),
            None if "out" in b else out.ctypes.data_as(ctypes.c_void_p),
            None)
        assert rc == _native.KK_ERR_INVALID, b
        assert lib.kk_last_error().decode()
    axes = np.array([0, 5, 1], dtype=np.int32)
    rc = lib.kk_fp_table(_native.f64_ptr(P), 3, _native.i32_ptr(axes), 6, 7,
                         shape, vs, out.ctypes.data_as(ctypes.c_void_p), None)
    assert rc == _native.KK_ERR_INVALID
    assert "axis" in lib.kk_last_error().decode()


def test_bp_table_validation(lib):
    out = np.zeros(2, dtype=_native.BP_ANGLE_DTYPE)
    P = np.zeros((2, 12))
    assert lib.kk_bp_table(_native.f64_ptr(P), 2, None) == \
        _native.KK_ERR_INVALID
    assert lib.kk_bp_table(_native.f64_ptr(P), 0,
                           out.ctypes.data_as(ctypes.c_void_p)) == \
        _native.KK_ERR_INVALID


def test_null_handles_are_rejected(lib):
    assert lib.kk_plan_create(None) == _native.KK_ERR_INVALID
    assert lib.kk_filter_create(None) == _native.KK_ERR_INVALID
    assert lib.kk_plan_destroy(None) == _native.KK_OK
    assert lib.kk_filter_destroy(None) == _native.KK_OK
    assert lib.kk_fp_frame_mask(None) == 0
    rc = lib.kk_set_params_bp(None, None, 1)
    assert rc == _native.KK_ERR_INVALID
    rc = lib.kk_cone_fp(None, None, None, _native.i64((1, 1, 1)), 1, 1, 0,
                        None)
    assert rc in (_native.KK_ERR_INVALID, _native.KK_ERR_NOT_SET)


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    """No CPU fallback: without the built library the product raises."""
    monkeypatch.setattr(_native, "_lib", None)
    monkeypatch.setenv("KK_LIBRARY", str(tmp_path / "missing.so"))
    with pytest.raises(_native.NativeLibraryError):
        _native.lib()


def test_residual_texture_validation(lib):
    """kk_residual_to_bp_texture rejects null arguments and bad sizes on
    the host, before any launch."""
    st = _native.i64((1, 1, 1))
    buf = ctypes.c_void_p(16)   # never dereferenced: validation fails first
    rc = lib.kk_residual_to_bp_texture(None, st, buf, st, buf, st, 1, 1, 1,
                                       buf, None)
    assert rc == _native.KK_ERR_INVALID
    rc = lib.kk_residual_to_bp_texture(buf, st, buf, st, buf, st, 0, 4, 4,
                                       buf, None)
    assert rc == _native.KK_ERR_INVALID
    assert "bad sizes" in _native.last_error()
    rc = lib.kk_residual_to_bp_texture(buf, st, buf, st, buf, st, 70000, 4,
                                       4, buf, None)
    assert rc == _native.KK_ERR_UNSUPPORTED
