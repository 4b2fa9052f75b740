"""CPU emulation of the HIP kernels' float32 arithmetic, driven by the
exact per-angle tables the library uploads (``kk_fp_table`` /
``kk_bp_table``), checked against the FP64 oracle under the parity bar.

This pins the host-side table derivation (ray axes, b/c swap, slice
offsets, texel-centre shifts, fast-path detection) and the kernels' index
arithmetic (halo clamping, slice-range clipping) without a GPU.
"""
import numpy as np
import pytest

import oracle
from conftest import assert_parity

kk = pytest.importorskip("kernelkit")
from kernelkit import _native, workloads  # noqa: E402

f32 = np.float32


def _texture(arr):
    """Zero-halo padded copy (HALO_LO before, HALO_HI after each axis)."""
    lo, hi = _native.HALO_LO, _native.HALO_HI
    return np.pad(np.asarray(arr, f32), [(lo, hi)] * 3)


_FRAMES = {0: (0, 1, 2), 1: (1, 0, 2), 2: (2, 1, 0)}  # (a, b, c) per axis


def emulate_fp(vol_world, table, rows, cols):
    """cone_fp_kernel in float32 numpy: returns proj[angle, row, col].

    Reads the frame texture of each angle's ray axis, i.e. the volume in
    (a, b, c) order padded by the halo along b and c; coordinates are padded
    coordinates (halo folded into the intercepts) clamped to [0, n + 1]."""
    lo, hi = _native.HALO_LO, _native.HALO_HI
    frames = {a: np.pad(np.transpose(np.asarray(vol_world, f32), abc),
                        [(0, 0), (lo, hi), (lo, hi)])
              for a, abc in _FRAMES.items()}
    n = len(table)
    out = np.zeros((n, rows, cols), f32)
    r = np.arange(rows, dtype=f32)[:, None]
    c = np.arange(cols, dtype=f32)[None, :]
    for k, P in enumerate(table):
        T = frames[int(P["axis"])]
        Da = r * f32(P["va"]) + (c * f32(P["ua"]) + f32(P["qa"]))
        Db = r * f32(P["vb"]) + (c * f32(P["ub"]) + f32(P["qb"]))
        Dc = r * f32(P["vc"]) + (c * f32(P["uc"]) + f32(P["qc"]))
        with np.errstate(divide="ignore", invalid="ignore"):
            a1 = (Db / Da).astype(f32)
            a2 = (Dc / Da).astype(f32)
        b1 = (a1 * f32(P["sig"]) + f32(P["taub"])).astype(f32)
        b2 = (a2 * f32(P["sig"]) + f32(P["tauc"])).astype(f32)
        na, nb, nc = int(P["na"]), int(P["nb"]), int(P["nc"])
        assert T.shape == (na, nb + lo + hi, nc + lo + hi)
        assert int(P["sbb"]) == 8 * (nc + lo + hi)
        assert int(P["sab"]) == int(P["sbb"]) * (nb + lo + hi)
        acc = np.zeros((rows, cols), f32)
        for i in range(na):  # the kernel's range clipping only skips zeros
            xb = np.clip(a1 * f32(i) + b1, 0, nb + 1).astype(f32)
            xc = np.clip(a2 * f32(i) + b2, 0, nc + 1).astype(f32)
            jb = np.floor(xb).astype(np.int64)
            jc = np.floor(xc).astype(np.int64)
            fb = xb - jb
            fc = xc - jc
            Ti = T[i]
            l0 = fb * (Ti[jb + 1, jc] - Ti[jb, jc]) + Ti[jb, jc]
            l1 = fb * (Ti[jb + 1, jc + 1] - Ti[jb, jc + 1]) + Ti[jb, jc + 1]
            acc += fc * (l1 - l0) + l0
        scale = np.sqrt(a1 * a1 * f32(P["sc1"]) + a2 * a2 * f32(P["sc2"])
                        + f32(1)) * f32(P["oscale"])
        out[k] = acc * scale
    return out


def emulate_bp(proj, table, shape, scale):
    """cone_bp_kernel in float32 numpy (both paths): returns vol[x, y, z].

    The forms give padded texture coordinates, clamped to [0, n + 1]."""
    n, rows, cols = proj.shape
    tex = _texture(proj)
    X, Y, Z = np.meshgrid(*[np.arange(s, dtype=f32) - f32(0.5 * s)
                            + f32(0.5) for s in shape], indexing="ij")
    acc = np.zeros(shape, f32)
    lo = _native.HALO_LO
    for k, P in enumerate(table):
        den = Y * f32(P["dy"]) + (X * f32(P["dx"]) + f32(P["dw"]))
        nu = Y * f32(P["nuy"]) + (X * f32(P["nux"]) + f32(P["nuw"]))
        nv = Y * f32(P["nvy"]) + (X * f32(P["nvx"]) + f32(P["nvw"]))
        den = den + Z * f32(P["dz"])
        nu = nu + Z * f32(P["nuz"])
        nv = nv + Z * f32(P["nvz"])
        r = (f32(1) / den).astype(f32)
        u = np.clip(nu * r, 0, cols + 1)
        v = np.clip(nv * r, 0, rows + 1)
        uf, vf = np.floor(u), np.floor(v)
        fu, fv = u - uf, v - vf
        ui = uf.astype(np.int64)
        vi = vf.astype(np.int64)
        img = tex[k + lo]
        h0 = fu * (img[vi, ui + 1] - img[vi, ui]) + img[vi, ui]
        h1 = fu * (img[vi + 1, ui + 1] - img[vi + 1, ui]) + img[vi + 1, ui]
        acc += (r * r) * (fv * (h1 - h0) + h0)
    return acc * f32(scale)


def _case(n_angles=12, shape=(28, 32, 24), rows=36, cols=40, **pert):
    pgs = workloads.circular_geometry(n_angles, sod=120.0, odd=60.0,
                                      rows=rows, cols=cols, pixel=1.2, seed=4,
                                      **pert)
    vg = kk.resolve_volume_geometry(shape=shape, voxel_size=1.0)
    vol = workloads.ellipsoid_phantom(shape, n=5, seed=4)
    return pgs, vg, vol


def _fp_table(pgs, vg):
    P = kk.kernels.RayDrivenConeFP.geoms2params(pgs, vg)
    det = pgs[0].detector
    table, mask = _native.fp_table(np.stack(P), P.ray_axes, det.rows,
                                   det.cols, vg.shape, vg.voxel_size)
    assert mask == int(np.bitwise_or.reduce(1 << table["axis"]))
    return table


@pytest.mark.parametrize("pert", [{}, {"tilt_deg": 2.0, "drift_px": 4.0},
                                  {"wobble_deg": 1.0}])
def test_fp_emulation_matches_oracle(pert):
    pgs, vg, vol = _case(**pert)
    got = emulate_fp(vol, _fp_table(pgs, vg), 36, 40)
    assert_parity(got, oracle.cone_fp(vol, pgs, vg), label=f"fp-emu {pert}")


def test_fp_emulation_z_major():
    """A geometry whose rays run along z uses the z frame (b, c) = (y, x)."""
    g0 = kk.rotate(workloads.circular_geometry(
        1, sod=120.0, odd=60.0, rows=24, cols=28, pixel=1.2)[0],
        pitch=np.pi / 2)
    pgs = [kk.rotate(g0, yaw=t) for t in np.linspace(0, 2 * np.pi, 4, False)]
    vg = kk.resolve_volume_geometry(shape=(16, 20, 12), voxel_size=1.0)
    vol = workloads.ellipsoid_phantom(vg.shape, n=3, seed=1)
    table = _fp_table(pgs, vg)
    assert set(table["axis"].tolist()) == {2}
    got = emulate_fp(vol, table, 24, 28)
    assert_parity(got, oracle.cone_fp(vol, pgs, vg), label="z-major")


@pytest.mark.parametrize("pert", [{}, {"tilt_deg": 2.0, "drift_px": 4.0}])
def test_bp_emulation_matches_oracle(pert):
    pgs, vg, vol = _case(**pert)
    proj = oracle.cone_fp(vol, pgs, vg)
    B = kk.kernels.VoxelDrivenConeBP.geoms2params(pgs, vg)
    table = _native.bp_table(B)
    assert bool(table["fast"].all()) == (not pert)
    got = emulate_bp(proj, table, vg.shape, vg.voxel_volume())
    assert_parity(got, oracle.cone_bp(proj, pgs, vg), label=f"bp-emu {pert}")


def test_full_config_fp_table_precision():
    """At 512^3/768^2 the float32 table rounding moves sample coordinates by
    well under 1e-3 voxel against float64 (a few angles checked)."""
    pgs, vg = workloads.config_geometry("full")
    sub = [pgs[k] for k in (0, 90, 137, 400)]
    t = _fp_table(sub, vg)
    fr = oracle.normalized_frames(sub, vg)
    for k, P in enumerate(t):
        a, b, c = _FRAMES[int(P["axis"])]
        s, d0 = fr["s"][k], fr["d0"][k]
        U, V = fr["U"][k], fr["V"][k]
        for (r, col) in ((0, 0), (383, 400), (767, 767)):
            D = d0 + (col + 0.5) * U + (r + 0.5) * V
            a1 = (D[b] - s[b]) / (D[a] - s[a])
            for i in (0, 255, 511):
                exact = s[b] + a1 * (i - 255.5 - s[a]) + 255.5 + 1.0
                Df = [np.float32(P[f"q{x}"]) + np.float32(col) *
                      np.float32(P[f"u{x}"]) + np.float32(r) *
                      np.float32(P[f"v{x}"]) for x in "abc"]
                a1f = np.float32(Df[1] / Df[0])
                emu = a1f * np.float32(i) + (a1f * np.float32(P["sig"]) +
                                             np.float32(P["taub"]))
                assert abs(float(emu) - exact) < 1e-3
