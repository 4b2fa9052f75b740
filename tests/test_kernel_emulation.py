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


_TS, _TB, _TC = 8, 32, 32   # cone_fp_t_kernel tile (slices, b, c)


def _project_corner(P, xa, xb, xc):
    """cone_fp_t_kernel's project(): detector (col, row) of the ray through
    the source-relative point X, float32; ok iff finite and in front."""
    g = {k: f32(P[k]) for k in ("qa", "qb", "qc", "ua", "ub", "uc",
                                "va", "vb", "vc")}
    xa, xb, xc = f32(xa), f32(xb), f32(xc)
    with np.errstate(all="ignore"):
        vx = (g["vb"] * xc - g["vc"] * xb, g["vc"] * xa - g["va"] * xc,
              g["va"] * xb - g["vb"] * xa)
        qx = (g["qb"] * xc - g["qc"] * xb, g["qc"] * xa - g["qa"] * xc,
              g["qa"] * xb - g["qb"] * xa)
        vq = (g["vb"] * g["qc"] - g["vc"] * g["qb"],
              g["vc"] * g["qa"] - g["va"] * g["qc"],
              g["va"] * g["qb"] - g["vb"] * g["qa"])
        den = g["ua"] * vx[0] + g["ub"] * vx[1] + g["uc"] * vx[2]
        inv = f32(1) / den
        col = -(g["qa"] * vx[0] + g["qb"] * vx[1] + g["qc"] * vx[2]) * inv
        row = -(g["ua"] * qx[0] + g["ub"] * qx[1] + g["uc"] * qx[2]) * inv
        lam = (g["ua"] * vq[0] + g["ub"] * vq[1] + g["uc"] * vq[2]) * inv
    return col, row, bool(np.isfinite(col) and np.isfinite(row) and lam > 0)


def emulate_fp_t(proj, table, shape, use_box=True, tile=(_TS, _TB, _TC)):
    """cone_fp_t_kernel: per frame, per tile, per angle, the rays of the
    projected pixel rectangle add their in-tile taps.  Coordinates and
    weights in float32 as the kernel; sums in float64 (the kernel's LDS
    accumulation order is not defined).  use_box=False walks every ray of
    the detector instead (to check that the rectangle loses nothing)."""
    n, rows, cols = proj.shape
    lo = _native.HALO_LO
    TS, TB, TC = tile
    out = np.zeros(shape)
    for a, (_, b, c) in _FRAMES.items():
        na, nb, nc = shape[a], shape[b], shape[c]
        acc = np.zeros((na, nb, nc))
        angles = [k for k in range(n) if int(table[k]["axis"]) == a]
        for i0 in range(0, na, TS):
            ns = min(TS, na - i0)
            for j0 in range(0, nb, TB):
                nbt = min(TB, nb - j0)
                for k0 in range(0, nc, TC):
                    nct = min(TC, nc - k0)
                    for k in angles:
                        P = table[k]
                        c_lo, c_hi, r_lo, r_hi = 0, cols - 1, 0, rows - 1
                        if use_box:
                            pts = [_project_corner(
                                P, f32(ia) + f32(P["sig"]),
                                f32(jb) - f32(P["taub"]),
                                f32(jc) - f32(P["tauc"]))
                                for ia in (i0, i0 + ns - 1)
                                for jb in (j0, j0 + nbt + 1)
                                for jc in (k0, k0 + nct + 1)]
                            if all(p[2] for p in pts):
                                cs = [p[0] for p in pts]
                                rs = [p[1] for p in pts]
                                c_lo = max(0, int(np.floor(np.clip(
                                    min(cs), -4, 1e8))) - 1)
                                c_hi = min(cols - 1, int(np.ceil(np.clip(
                                    max(cs), -4, 1e8))) + 1)
                                r_lo = max(0, int(np.floor(np.clip(
                                    min(rs), -4, 1e8))) - 1)
                                r_hi = min(rows - 1, int(np.ceil(np.clip(
                                    max(rs), -4, 1e8))) + 1)
                        if c_lo > c_hi or r_lo > r_hi:
                            continue
                        R, C = np.meshgrid(np.arange(r_lo, r_hi + 1),
                                           np.arange(c_lo, c_hi + 1),
                                           indexing="ij")
                        R, C = R.ravel(), C.ravel()
                        fr, fcl = R.astype(f32), C.astype(f32)
                        Da = fr * f32(P["va"]) + (fcl * f32(P["ua"]) +
                                                  f32(P["qa"]))
                        Db = fr * f32(P["vb"]) + (fcl * f32(P["ub"]) +
                                                  f32(P["qb"]))
                        Dc = fr * f32(P["vc"]) + (fcl * f32(P["uc"]) +
                                                  f32(P["qc"]))
                        with np.errstate(all="ignore"):
                            a1 = (Db / Da).astype(f32)
                            a2 = (Dc / Da).astype(f32)
                        b1 = (a1 * f32(P["sig"]) + f32(P["taub"])).astype(f32)
                        b2 = (a2 * f32(P["sig"]) + f32(P["tauc"])).astype(f32)
                        w = (np.sqrt(a1 * a1 * f32(P["sc1"]) + a2 * a2 *
                                     f32(P["sc2"]) + f32(1)) *
                             f32(P["oscale"])).astype(f32)
                        val = (proj[k, R, C].astype(f32) * w).astype(
                            np.float64)
                        for s in range(ns):
                            i = i0 + s
                            xb = np.clip(a1 * f32(i) + b1, 0, nb + 1)
                            xc = np.clip(a2 * f32(i) + b2, 0, nc + 1)
                            ib = np.floor(xb).astype(np.int64) - lo
                            ic = np.floor(xc).astype(np.int64) - lo
                            fb = (xb - np.floor(xb)).astype(np.float64)
                            fc = (xc - np.floor(xc)).astype(np.float64)
                            for db, wb in ((0, 1 - fb), (1, fb)):
                                for dc, wc in ((0, 1 - fc), (1, fc)):
                                    jj, kk_ = ib + db, ic + dc
                                    m = ((jj >= j0) & (jj < j0 + nbt) &
                                         (kk_ >= k0) & (kk_ < k0 + nct))
                                    np.add.at(acc[i], (jj[m], kk_[m]),
                                              (val * wb * wc)[m])
        out += np.transpose(acc, np.argsort((a, b, c)))
    return out


@pytest.mark.parametrize("pert", [{}, {"tilt_deg": 2.0, "drift_px": 4.0}])
def test_fp_transpose_emulation(pert):
    """The tiled transpose equals the oracle's exact transpose, and its
    per-tile pixel rectangles lose no ray."""
    pgs, vg, _ = _case(n_angles=5, shape=(20, 40, 36), rows=30, cols=44,
                       **pert)
    pgs = list(pgs) + [kk.rotate(pgs[0], pitch=np.pi / 2)]   # + a z frame
    table = _fp_table(pgs, vg)
    assert set(table["axis"].tolist()) >= {2}
    y = np.random.default_rng(3).uniform(size=(len(pgs), 30, 44))
    got = emulate_fp_t(y, table, vg.shape)
    assert_parity(got, oracle.cone_fp_transpose(y, pgs, vg),
                  label=f"fp_t-emu {pert}")
    # small tiles make the pixel rectangles a small part of the detector:
    # walking every ray instead changes nothing
    small = emulate_fp_t(y, table, vg.shape, tile=(3, 5, 4))
    full = emulate_fp_t(y, table, vg.shape, use_box=False, tile=(3, 5, 4))
    np.testing.assert_allclose(small, full, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(small, got, rtol=1e-12, atol=1e-12)
