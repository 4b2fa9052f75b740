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
    """cone_fp_kernel in numpy: returns proj[angle, row, col].

    Reads the frame texture of each angle's ray axis, i.e. the volume in
    (a, b, c) order padded by the halo along b and c.  Ray slopes and sample
    coordinates are float64 (padded coordinates, halo folded into the
    intercepts, clamped to [0, n + 1]); weights and sums float32."""
    lo, hi = _native.HALO_LO, _native.HALO_HI
    frames = {a: np.pad(np.transpose(np.asarray(vol_world, f32), abc),
                        [(0, 0), (lo, hi), (lo, hi)])
              for a, abc in _FRAMES.items()}
    n = len(table)
    out = np.zeros((n, rows, cols), f32)
    r = np.arange(rows, dtype=np.float64)[:, None]
    c = np.arange(cols, dtype=np.float64)[None, :]
    for k, P in enumerate(table):
        T = frames[int(P["axis"])]
        a1, b1, a2, b2 = _fp_rays(P, r, c)
        na, nb, nc = int(P["na"]), int(P["nb"]), int(P["nc"])
        assert T.shape == (na, nb + lo + hi, nc + lo + hi)
        assert int(P["sbb"]) == 8 * (nc + lo + hi)
        assert int(P["sab"]) == int(P["sbb"]) * (nb + lo + hi)
        acc = np.zeros((rows, cols), f32)
        for i in range(na):  # the kernel's range clipping only skips zeros
            jb, fb = _coord(a1, b1, i, nb + 1)
            jc, fc = _coord(a2, b2, i, nc + 1)
            Ti = T[i]
            l0 = fb * (Ti[jb + 1, jc] - Ti[jb, jc]) + Ti[jb, jc]
            l1 = fb * (Ti[jb + 1, jc + 1] - Ti[jb, jc + 1]) + Ti[jb, jc + 1]
            acc += fc * (l1 - l0) + l0
        out[k] = acc * _fp_weight(P, a1, a2)
    return out


def _fp_rays(P, r, c):
    """kk_fp_ray: float64 slopes and intercepts of the rays (r, c)."""
    g = {x: float(P[x]) for x in ("qa", "qb", "qc", "ua", "ub", "uc", "va",
                                  "vb", "vc", "sig", "taub", "tauc")}
    Da = r * g["va"] + (c * g["ua"] + g["qa"])
    Db = r * g["vb"] + (c * g["ub"] + g["qb"])
    Dc = r * g["vc"] + (c * g["uc"] + g["qc"])
    with np.errstate(divide="ignore", invalid="ignore"):
        inv = 1.0 / Da
    a1, a2 = Db * inv, Dc * inv
    return a1, a1 * g["sig"] + g["taub"], a2, a2 * g["sig"] + g["tauc"]


def _fp_weight(P, a1, a2):
    """kk_fp_weight (float32)."""
    f1, f2 = a1.astype(f32), a2.astype(f32)
    return (np.sqrt(f1 * f1 * f32(P["sc1"]) + f2 * f2 * f32(P["sc2"])
                    + f32(1)) * f32(P["oscale"])).astype(f32)


def _coord(a, b, i, n1):
    """kk_fp_coord: float64 coordinate clamped to [0, n1], split into
    index and float32 weight."""
    x = np.clip(a * float(i) + b, 0.0, float(n1))
    fl = np.floor(x)
    return fl.astype(np.int64), (x - fl).astype(f32)


def _split_clamped(off, base, hi):
    """split_clamped of cone_bp.hip: base + off clamped into [0, hi]."""
    o = np.clip(off.astype(f32), f32(-1e7), f32(1e7))
    fl = np.floor(o)
    i = base + fl.astype(np.int64)
    out = (i < 0) | (i >= hi)
    return np.clip(i, 0, hi), np.where(out, f32(0), (o - fl).astype(f32))


def emulate_bp(proj, table, shape, scale):
    """cone_bp_kernel in numpy (both paths, z-runs of 8): returns
    vol[x, y, z].  Per (x, y, angle) the forms are float64; per voxel only
    offsets from integer bases are float32."""
    K = 8
    n, rows, cols = proj.shape
    tex = _texture(proj)
    nx, ny, nz = shape
    X, Y, IZ = np.meshgrid(np.arange(nx) - 0.5 * nx + 0.5,
                           np.arange(ny) - 0.5 * ny + 0.5,
                           np.arange(nz), indexing="ij")
    Z0 = (IZ - IZ % K) - 0.5 * nz + 0.5      # first voxel of the run
    kz = (IZ % K).astype(f32)
    acc = np.zeros(shape, f32)
    lo = _native.HALO_LO
    for k, P in enumerate(table):
        g = {x: float(P[x]) for x in ("nux", "nuy", "nuz", "nuw", "nvx",
                                      "nvy", "nvz", "nvw", "dx", "dy", "dz",
                                      "dw")}
        den = Y * g["dy"] + (X * g["dx"] + g["dw"])
        nu = Y * g["nuy"] + (X * g["nux"] + g["nuw"])
        nv = Y * g["nvy"] + (X * g["nvx"] + g["nvw"])
        with np.errstate(divide="ignore", invalid="ignore"):
            if P["fast"]:
                r = 1.0 / den
                u = nu * r
                v0 = (Z0 * g["nvz"] + nv) * r
                gv = (g["nvz"] * r).astype(f32)
                ok = np.isfinite(u) & np.isfinite(v0) & np.isfinite(gv)
                rf = (f32(1) / den.astype(f32)).astype(f32)
                w = np.where(ok, rf * rf, f32(0))
                uc = np.clip(np.where(ok, u, 0.0), 0.0, cols + 1)
                iu = np.floor(uc).astype(np.int64)
                fu = (uc - np.floor(uc)).astype(f32)
                vc0 = np.clip(np.where(ok, v0, 0.0), -1e6, 1e6)
                vfl = np.floor(vc0)
                fv0 = (vc0 - vfl).astype(f32)
                iv, fv = _split_clamped(kz * np.where(ok, gv, f32(0)) + fv0,
                                        vfl.astype(np.int64), rows + 1)
            else:
                den0 = Z0 * g["dz"] + den
                nu0 = Z0 * g["nuz"] + nu
                nv0 = Z0 * g["nvz"] + nv
                u0, v0 = nu0 / den0, nv0 / den0
                bok = np.isfinite(u0) & np.isfinite(v0)
                ub = np.where(bok, np.floor(np.clip(np.where(bok, u0, 0),
                                                    -1e6, 1e6)), 0.0)
                vb = np.where(bok, np.floor(np.clip(np.where(bok, v0, 0),
                                                    -1e6, 1e6)), 0.0)
                e0u = (nu0 - ub * den0).astype(f32)
                e1u = (g["nuz"] - ub * g["dz"]).astype(f32)
                e0v = (nv0 - vb * den0).astype(f32)
                e1v = (g["nvz"] - vb * g["dz"]).astype(f32)
                rk = (f32(1) / (kz * f32(g["dz"]) + den0.astype(f32))
                      ).astype(f32)
                du = ((kz * e1u + e0u) * rk).astype(f32)
                dv = ((kz * e1v + e0v) * rk).astype(f32)
                ok = np.isfinite(du) & np.isfinite(dv)
                iu, fu = _split_clamped(np.where(ok, du, f32(0)),
                                        np.where(ok, ub, 0).astype(np.int64),
                                        cols + 1)
                iv, fv = _split_clamped(np.where(ok, dv, f32(0)),
                                        np.where(ok, vb, 0).astype(np.int64),
                                        rows + 1)
                w = np.where(ok, rk * rk, f32(0))
        img = tex[k + lo]
        h0 = fv * (img[iv + 1, iu] - img[iv, iu]) + img[iv, iu]
        h1 = fv * (img[iv + 1, iu + 1] - img[iv, iu + 1]) + img[iv, iu + 1]
        acc += w * (fu * (h1 - h0) + h0)
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
    """At 512^3/768^2 the table's float64 geometry reproduces the sample
    coordinates of the float64 frames to 1e-9 voxel (a few angles, rays and
    slices checked); float32 tables would be off by up to ~1e-4."""
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
            ea1, eb1, _, _ = _fp_rays(P, np.float64(r), np.float64(col))
            for i in (0, 255, 511):
                exact = s[b] + a1 * (i - 255.5 - s[a]) + 255.5 + 1.0
                assert abs(float(ea1 * i + eb1) - exact) < 1e-9


_TS, _TB, _TC = 8, 32, 32   # cone_fp_t_kernel tile (slices, b, c)


def _project_corner(P, xa, xb, xc):
    """cone_fp_t_kernel's project(): detector (col, row) of the ray through
    the source-relative point X (float64); ok iff finite and in front."""
    g = {k: float(P[k]) for k in ("qa", "qb", "qc", "ua", "ub", "uc",
                                  "va", "vb", "vc")}
    xa, xb, xc = float(xa), float(xb), float(xc)
    vx = (g["vb"] * xc - g["vc"] * xb, g["vc"] * xa - g["va"] * xc,
          g["va"] * xb - g["vb"] * xa)
    qx = (g["qb"] * xc - g["qc"] * xb, g["qc"] * xa - g["qa"] * xc,
          g["qa"] * xb - g["qb"] * xa)
    vq = (g["vb"] * g["qc"] - g["vc"] * g["qb"],
          g["vc"] * g["qa"] - g["va"] * g["qc"],
          g["va"] * g["qb"] - g["vb"] * g["qa"])
    den = g["ua"] * vx[0] + g["ub"] * vx[1] + g["uc"] * vx[2]
    if den == 0.0:
        return 0.0, 0.0, False
    col = -(g["qa"] * vx[0] + g["qb"] * vx[1] + g["qc"] * vx[2]) / den
    row = -(g["ua"] * qx[0] + g["ub"] * qx[1] + g["uc"] * qx[2]) / den
    lam = (g["ua"] * vq[0] + g["ub"] * vq[1] + g["uc"] * vq[2]) / den
    ok = bool(np.isfinite(col) and np.isfinite(row) and lam > 0
              and abs(col) < 1e8 and abs(row) < 1e8)
    return (f32(col), f32(row), ok) if ok else (0.0, 0.0, False)


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
                                P, ia + float(P["sig"]),
                                jb - float(P["taub"]),
                                jc - float(P["tauc"]))
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
                        a1, b1, a2, b2 = _fp_rays(P, R.astype(np.float64),
                                                  C.astype(np.float64))
                        good = (np.isfinite(a1) & np.isfinite(a2) &
                                np.isfinite(b1) & np.isfinite(b2))
                        R, C = R[good], C[good]
                        a1, b1, a2, b2 = a1[good], b1[good], a2[good], \
                            b2[good]
                        w = _fp_weight(P, a1, a2)
                        val = (proj[k, R, C].astype(f32) * w).astype(
                            np.float64)
                        for s in range(ns):
                            i = i0 + s
                            ib, fb = _coord(a1, b1, i, nb + 1)
                            ic, fc = _coord(a2, b2, i, nc + 1)
                            ib, ic = ib - lo, ic - lo
                            fb = fb.astype(np.float64)
                            fc = fc.astype(np.float64)
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
