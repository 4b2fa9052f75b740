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
    """split_clamped of cone_bp.hip: off clamped into [-base, hi - base]
    (float32 med3), so base + off lies in [0, hi]; then floor and
    fraction."""
    base = np.asarray(base, np.int64)
    o = np.clip(off.astype(f32), (-base).astype(f32),
                (hi - base).astype(f32)).astype(f32)
    fl = np.floor(o)
    return base + fl.astype(np.int64), (o - fl).astype(f32)


def _split_clamped_flagged(off, base, hi):
    """The earlier form of split_clamped: clamp the float32 offset to
    +-1e7, split, then flag indices outside [0, hi) (weight 0, index
    clamped)."""
    o = np.clip(off.astype(f32), f32(-1e7), f32(1e7))
    fl = np.floor(o)
    i = base + fl.astype(np.int64)
    out = (i < 0) | (i >= hi)
    return np.clip(i, 0, hi), np.where(out, f32(0), (o - fl).astype(f32))


def test_split_clamped_equals_clamp_then_split():
    """The offset clamp into [-base, hi - base] gives the index and weight
    of clamping the exact coordinate base + off into [0, hi] (float64,
    exact for float32 offsets and integer bases) and of the flagged form,
    bit for bit, at the bounds, near integers and for huge offsets."""
    rng = np.random.default_rng(11)
    n = 200000
    hi = rng.integers(40, 1600, n)
    base = rng.integers(-3000, 3000, n)
    off = np.concatenate([
        rng.uniform(-4000, 4000, n // 4),
        rng.uniform(-3, 3, n // 4),
        np.round(rng.uniform(-3000, 3000, n // 4)),        # integers
        rng.choice([1e9, -1e9, 1e7, -1e7, 2.0 ** -30, -2.0 ** -30,
                    -1e-9, 1 - 2.0 ** -24], n // 4),
    ]).astype(f32)
    # offsets that land exactly on the bounds 0 and hi
    off[:1000] = (-base[:1000]).astype(f32)
    off[1000:2000] = (hi[1000:2000] - base[1000:2000]).astype(f32)
    idx, frac = _split_clamped(off, base, hi)
    x = np.clip(base + off.astype(np.float64), 0.0, hi.astype(np.float64))
    fl = np.floor(x)
    np.testing.assert_array_equal(idx, fl.astype(np.int64))
    np.testing.assert_array_equal(frac, (x - fl).astype(f32))
    idx2, frac2 = _split_clamped_flagged(off, base, hi)
    np.testing.assert_array_equal(idx, idx2)
    np.testing.assert_array_equal(frac, frac2)
    assert idx.min() >= 0 and (idx <= hi).all()


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
            if P["fast"] == 2:
                r = 1.0 / den
                u0 = (Z0 * g["nuz"] + nu) * r
                v0 = (Z0 * g["nvz"] + nv) * r
                gu = (g["nuz"] * r).astype(f32)
                gv = (g["nvz"] * r).astype(f32)
                ok = (np.isfinite(u0) & np.isfinite(v0) & np.isfinite(gu)
                      & np.isfinite(gv))
                rf = (f32(1) / den.astype(f32)).astype(f32)
                w = np.where(ok, rf * rf, f32(0))
                uc0 = np.clip(np.where(ok, u0, 0.0), -1e6, 1e6)
                ufl = np.floor(uc0)
                vc0 = np.clip(np.where(ok, v0, 0.0), -1e6, 1e6)
                vfl = np.floor(vc0)
                iu, fu = _split_clamped(
                    kz * np.where(ok, gu, f32(0)) + (uc0 - ufl).astype(f32),
                    ufl.astype(np.int64), cols + 1)
                iv, fv = _split_clamped(
                    kz * np.where(ok, gv, f32(0)) + (vc0 - vfl).astype(f32),
                    vfl.astype(np.int64), rows + 1)
            elif P["fast"]:
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
                # no finite projection: offset at the lower bound
                iu, fu = _split_clamped(np.where(ok, du, -ub).astype(f32),
                                        ub.astype(np.int64), cols + 1)
                iv, fv = _split_clamped(np.where(ok, dv, -vb).astype(f32),
                                        vb.astype(np.int64), rows + 1)
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


@pytest.mark.parametrize("pert,path", [
    ({}, 1), ({"tilt_deg": 2.0, "drift_px": 4.0}, 2),
    ({"wobble_deg": 1.0, "drift_px": 3.0}, 0)])
def test_bp_emulation_matches_oracle(pert, path):
    """All three paths: circular (shared u and 1/den), detector tilted in
    its plane (shared 1/den, affine u and v), wobbling orbit (per voxel)."""
    pgs, vg, vol = _case(**pert)
    proj = oracle.cone_fp(vol, pgs, vg)
    B = kk.kernels.VoxelDrivenConeBP.geoms2params(pgs, vg)
    table = _native.bp_table(B)
    assert (table["fast"] == path).all()
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


_KJB = 8        # cone_fp_t_kernel: voxels along b per lane
_KJC = 2        # and consecutive voxels along c
_KEPS = 1e-6    # its widening of candidate intervals (voxels)


def _project_corner(P, xa, xb, xc):
    """cone_fp_t_kernel's box corners: detector (col, row) of the ray
    through the source-relative point X (float64, vectorised) as the linear
    forms col = (V x Q).X / N.X, row = (Q x U).X / N.X with N = U x V; ok
    iff finite and in front of the source (U.(V x Q) / N.X > 0)."""
    g = {k: float(P[k]) for k in ("qa", "qb", "qc", "ua", "ub", "uc",
                                  "va", "vb", "vc")}
    n = (g["ub"] * g["vc"] - g["uc"] * g["vb"],
         g["uc"] * g["va"] - g["ua"] * g["vc"],
         g["ua"] * g["vb"] - g["ub"] * g["va"])
    m = (g["vb"] * g["qc"] - g["vc"] * g["qb"],
         g["vc"] * g["qa"] - g["va"] * g["qc"],
         g["va"] * g["qb"] - g["vb"] * g["qa"])
    r = (g["qb"] * g["uc"] - g["qc"] * g["ub"],
         g["qc"] * g["ua"] - g["qa"] * g["uc"],
         g["qa"] * g["ub"] - g["qb"] * g["ua"])
    lamn = g["ua"] * m[0] + g["ub"] * m[1] + g["uc"] * m[2]
    with np.errstate(divide="ignore", invalid="ignore"):
        inv = 1.0 / (n[0] * xa + n[1] * xb + n[2] * xc)
        col = (m[0] * xa + m[1] * xb + m[2] * xc) * inv
        row = (r[0] * xa + r[1] * xb + r[2] * xc) * inv
    return col, row, np.isfinite(col) & np.isfinite(row) & (lamn * inv > 0)


def _int_range(lo, hi, n):
    """int_range(): [ceil(lo), floor(hi)] clipped to [0, n - 1]."""
    lo = np.maximum(lo, -1)
    hi = np.minimum(hi, n)
    return (np.maximum(0, np.ceil(lo)).astype(np.int64),
            np.minimum(n - 1, np.floor(hi)).astype(np.int64))


def _fpt_taps(P, i, rows, cols, nb, nc):
    """Every tap of the forward kernel's samples at slice i: ray (r, col),
    voxel (j, k) along the frame's b and c, and its weight (float64)."""
    lo = _native.HALO_LO
    R, C = np.meshgrid(np.arange(rows), np.arange(cols), indexing="ij")
    R, C = R.ravel(), C.ravel()
    a1, b1, a2, b2 = _fp_rays(P, R.astype(np.float64), C.astype(np.float64))
    good = (np.isfinite(a1) & np.isfinite(b1) & np.isfinite(a2) &
            np.isfinite(b2))
    R, C, a1, b1, a2, b2 = (x[good] for x in (R, C, a1, b1, a2, b2))
    ib, fb = _coord(a1, b1, i, nb + 1)
    ic, fc = _coord(a2, b2, i, nc + 1)
    w = _fp_weight(P, a1, a2).astype(np.float64)
    out = []
    for db, wb in ((0, 1 - fb.astype(np.float64)), (1, fb.astype(np.float64))):
        for dc, wc in ((0, 1 - fc.astype(np.float64)),
                       (1, fc.astype(np.float64))):
            j, k = ib - lo + db, ic - lo + dc
            m = (j >= 0) & (j < nb) & (k >= 0) & (k < nc) & (wb * wc > 0)
            out.append((R[m], C[m], j[m], k[m], (w * wb * wc)[m]))
    return [np.concatenate(x) for x in zip(*out)]


def _circ_window(ia, betaf, kl, rows, pad=16, with_fast=False):
    """Circular path: the rows cone_fp_t_kernel reads for a lane (first c
    voxel kl) and a column (float32 1 / alpha and beta): nrow rows from the
    ceiling of the window's first row when the window is short, else the
    exact float32 range widened by 1e-2 row.  The first row is
    kl / alpha + origin, origin = fmaf(-beta, 1 / alpha, off)."""
    ia = np.asarray(ia, f32)
    with np.errstate(divide="ignore", over="ignore", invalid="ignore"):
        ln = (f32(_KJC + 1) * np.abs(ia)).astype(f32)
        fast = (ia != 0) & (ln + f32(3) <= f32(pad))
        nrow = np.floor(ln + f32(2e-3)).astype(np.int64) + 1
        off = (np.minimum(f32(0), f32(_KJC + 1) * ia) - f32(1e-3)).astype(f32)
        org = _f32fma(-np.asarray(betaf, f32), ia, np.where(fast, off,
                                                               f32(0)))
        r1 = _f32fma(kl.astype(f32), ia, org)
        rf = np.ceil(np.where(fast, r1, 0)).astype(np.int64)
        rf = np.minimum(np.maximum(rf, -pad), rows + pad - nrow)
        r2 = _f32fma(_KJC + 1.0, ia, r1)
    r_lo, r_hi = _int_range(np.minimum(r1, r2) - f32(1e-2),
                            np.maximum(r1, r2) + f32(1e-2), rows)
    steep = ia != 0
    r_lo, r_hi = np.where(steep, r_lo, 0), np.where(steep, r_hi, rows - 1)
    out = (np.where(fast, rf, r_lo), np.where(fast, rf + nrow - 1, r_hi))
    return out + (fast,) if with_fast else out


def _fpt_covered(P, i, R, C, J, K, rows, cols, eps=_KEPS):
    """Whether cone_fp_t_kernel's lane owning voxel (i, J, K) visits ray
    (R, C): the candidate enumeration of both of its paths, with the
    kernel's float64/float32 split."""
    jw = (J // _KJB) * _KJB
    K = (K // _KJC) * _KJC          # the lane's first c voxel
    arel = float(i) + float(P["sig"])
    taub, tauc = float(P["taub"]), float(P["tauc"])
    g = {x: float(P[x]) for x in ("qa", "qb", "qc", "ua", "ub", "uc", "vc")}
    if int(P["colc"]):
        t1 = (jw - eps - taub) / arel
        t2 = (jw + _KJB + 1 + eps - taub) / arel
        d1, d2 = t1 * g["ua"] - g["ub"], t2 * g["ua"] - g["ub"]
        with np.errstate(divide="ignore", invalid="ignore"):
            c1 = (-t1 * g["qa"] + g["qb"]) / d1
            c2 = (-t2 * g["qa"] + g["qb"]) / d2
        c_lo, c_hi = _int_range(np.minimum(c1, c2), np.maximum(c1, c2), cols)
        use = (d1 * d2 > 0) & np.isfinite(c1) & np.isfinite(c2)
        c_lo, c_hi = np.where(use, c_lo, 0), np.where(use, c_hi, cols - 1)
        fc = C.astype(np.float64)
        Da = fc * g["ua"] + g["qa"]
        inv = 1.0 / Da
        D0c = fc * g["uc"] + g["qc"]
        alpha = (g["vc"] * inv * arel).astype(f32)
        beta = (D0c * inv * arel + tauc).astype(f32)
        steep = np.abs(alpha) > f32(1e-9)
        with np.errstate(divide="ignore", over="ignore"):
            ia = np.where(steep, f32(1) / alpha, f32(0))
        r_lo, r_hi = _circ_window(ia, beta, K, rows)
    else:
        cs, rs, oks = [], [], []
        for corner in range(4):
            xb = (jw + _KJB + 1 + eps if corner & 1 else jw - eps) - taub
            xc = (K + _KJC + 1.0 + eps if corner & 2 else K - eps) - tauc
            pc, pr, ok = _project_corner(P, arel, xb, xc)
            cs.append(pc), rs.append(pr), oks.append(ok)
        ok = np.logical_and.reduce(oks)
        with np.errstate(invalid="ignore"):
            c_lo, c_hi = _int_range(np.min(cs, 0), np.max(cs, 0), cols)
            r_lo, r_hi = _int_range(np.min(rs, 0), np.max(rs, 0), rows)
        c_lo, c_hi = np.where(ok, c_lo, 0), np.where(ok, c_hi, cols - 1)
        r_lo, r_hi = np.where(ok, r_lo, 0), np.where(ok, r_hi, rows - 1)
    return ((C >= c_lo) & (C <= c_hi) & (R >= r_lo) & (R <= r_hi),
            (c_hi - c_lo + 1) * (r_hi - r_lo + 1))


def _f32fma(a, b, c):
    """fmaf: a * b + c rounded once to float32 (the float64 product of two
    float32 values is exact)."""
    return (np.asarray(a, np.float64) * np.asarray(b, np.float64)
            + np.asarray(c, np.float64)).astype(f32)


def _lane_boxes(P, i, jw, kl, rows, cols, eps=_KEPS):
    """General path: each lane's detector box (c_lo, c_hi, r_lo, r_hi) and
    whether its corners projected."""
    arel = float(i) + float(P["sig"])
    cs, rs, oks = [], [], []
    for corner in range(4):
        xb = (jw + _KJB + 1 + eps if corner & 1 else jw - eps) - \
            float(P["taub"])
        xc = (kl + _KJC + 1.0 + eps if corner & 2 else kl - eps) - \
            float(P["tauc"])
        pc, pr, ok = _project_corner(P, arel, xb, xc)
        cs.append(pc), rs.append(pr), oks.append(ok)
    ok = np.logical_and.reduce(oks)
    with np.errstate(invalid="ignore"):
        c_lo, c_hi = _int_range(np.min(cs, 0), np.max(cs, 0), cols)
        r_lo, r_hi = _int_range(np.min(rs, 0), np.max(rs, 0), rows)
    c_lo, c_hi = np.where(ok, c_lo, 0), np.where(ok, c_hi, cols - 1)
    r_lo, r_hi = np.where(ok, r_lo, 0), np.where(ok, r_hi, rows - 1)
    return c_lo, c_hi, r_lo, r_hi, ok


def _fpt_kernel_weights(P, i, R, C, J, K, rows, cols, nb):
    """The tap weights cone_fp_t_kernel forms for taps (R, C) -> (J, K):
    float64 anchors, float32 offsets and tents 1 - |x - centre| (without
    the ray weight), as its two paths compute them."""
    lo = _native.HALO_LO
    jw = (J // _KJB) * _KJB
    kl = (K // _KJC) * _KJC
    arel = float(i) + float(P["sig"])
    g = {x: float(P[x]) for x in ("qa", "qb", "qc", "ua", "ub", "uc", "va",
                                  "vb", "vc", "sig", "taub", "tauc")}
    if int(P["colc"]):
        fc = C.astype(np.float64)
        Da = fc * g["ua"] + g["qa"]
        Db = fc * g["ub"] + g["qb"]
        inv = 1.0 / Da
        a1 = Db * inv
        ib, fb = _coord(a1, a1 * g["sig"] + g["taub"], i, nb + 1)
        t0 = ib - lo - jw
        D0c = fc * g["uc"] + g["qc"]
        alpha = g["vc"] * inv * arel
        beta = D0c * inv * arel + g["tauc"]
        alphaf, betaf = alpha.astype(f32), beta.astype(f32)
        with np.errstate(divide="ignore", over="ignore"):
            ia = np.where(np.abs(alphaf) > f32(1e-9), f32(1) / alphaf, f32(0))
        r_lo, r_hi, fast = _circ_window(ia, betaf, kl, rows, with_fast=True)
        loc = (alpha * r_lo + beta - (kl + 1.0)).astype(f32)
        # short windows: groups of 4 rows from a float32 anchor
        # xg = fmaf(alpha, 4 g, loc), x = fmaf(alpha, t, xg) for t > 0, then
        # the remaining rows at fmaf(alpha, u, loc); long windows
        # fmaf(alpha, u, loc)
        u = R - r_lo
        grouped = fast & (u < ((r_hi - r_lo + 1) // 4) * 4)
        xg = _f32fma(alphaf, (u - u % 4).astype(f32), loc)
        x = np.where(grouped & (u % 4 > 0),
                     _f32fma(alphaf, (u % 4).astype(f32), xg),
                     np.where(grouped, xg,
                              _f32fma(alphaf, u.astype(f32), loc)))
        w0 = np.maximum(f32(0), (f32(1) - np.abs(x)).astype(f32))
        w1 = np.clip((f32(2) - x).astype(f32), f32(0), None)
        w1 = np.clip(np.minimum(x, w1), f32(0), None)  # med3(0, x, 2 - x)
        wc = np.where(K - kl == 0, w0, w1).astype(f32)
        fb64 = fb
    else:
        c_lo, _, r_lo, _, _ = _lane_boxes(P, i, jw, kl, rows, cols)
        # the wave's first column: the least c_lo over its 64 lanes
        wave = kl // (64 * _KJC)
        wc_lo = np.empty_like(C)
        for jw_u, w_u in set(zip(jw.tolist(), wave.tolist())):
            lanes = w_u * 64 * _KJC + _KJC * np.arange(64)
            cl, ch, _, _, _ = _lane_boxes(P, i, np.full(64, jw_u), lanes,
                                          rows, cols)
            m = (jw == jw_u) & (wave == w_u)
            wc_lo[m] = np.where(cl <= ch, cl, cols).min()
        Da0 = r_lo * g["va"] + (wc_lo * g["ua"] + g["qa"])
        Db0 = r_lo * g["vb"] + (wc_lo * g["ub"] + g["qb"])
        Dc0 = r_lo * g["vc"] + (wc_lo * g["uc"] + g["qc"])
        tb, tc = Db0 / Da0, Dc0 / Da0
        ob = (tb * arel + g["taub"] - (jw + lo)).astype(f32)
        oc = (tc * arel + g["tauc"] - (kl + lo)).astype(f32)
        F = {x: f32(g[x]) for x in ("ua", "ub", "uc", "va", "vb", "vc")}
        tbf, tcf = tb.astype(f32), tc.astype(f32)
        arf = f32(arel)
        dcf = (C - wc_lo).astype(f32)
        drf = (R - r_lo).astype(f32)
        # per column: numerators arel (dB - tb dA) affine in the row
        # offset, slopes arel (vb - tb va) and arel (vc - tc va);
        # denominator Da0 + dc ua + u va
        dAc = (dcf * F["ua"]).astype(f32)
        nb0 = (arf * _f32fma(-tbf, dAc, (dcf * F["ub"]).astype(f32))
               ).astype(f32)
        nc0 = (arf * _f32fma(-tcf, dAc, (dcf * F["uc"]).astype(f32))
               ).astype(f32)
        den0 = (Da0.astype(f32) + dAc).astype(f32)
        kvb = (arf * _f32fma(-tbf, F["va"], F["vb"])).astype(f32)
        kvc = (arf * _f32fma(-tcf, F["va"], F["vc"])).astype(f32)
        den = _f32fma(drf, F["va"], den0)
        rcp = (1.0 / den.astype(np.float64)).astype(f32)
        xb = _f32fma(_f32fma(drf, kvb, nb0), rcp, ob)
        xc = _f32fma(_f32fma(drf, kvc, nc0), rcp, oc)
        w0 = np.maximum(f32(0), (f32(1) - np.abs(xc)).astype(f32))
        w1 = np.clip(np.minimum(xc, (f32(2) - xc).astype(f32)), f32(0), None)
        wc = np.where(K - kl == 0, w0, w1).astype(f32)
        fl = np.floor(xb)
        t0 = fl.astype(np.int64)
        fb64 = (xb - fl).astype(f32)
    wb = np.where(J - jw == t0, f32(1) - fb64,
                  np.where(J - jw == t0 + 1, fb64, f32(0))).astype(f32)
    return (wb * wc).astype(f32)


def emulate_fp_t(proj, table, shape, eps=_KEPS, kernel_weights=False):
    """cone_fp_t_kernel: each voxel sums the taps of the samples its lane
    visits (float64 sums; the kernel's are float32 in its own order).
    With ``kernel_weights`` the tap weights are the kernel's own float32
    ones (:func:`_fpt_kernel_weights`), else the forward kernel's exact
    taps.  Returns (A^T proj, taps not visited, mean candidates per
    lane-angle / detector pixels per path, largest |kernel weight - exact
    weight|)."""
    n, rows, cols = proj.shape
    out = np.zeros(shape)
    missed = 0
    frac = {0: [], 1: []}
    werr = 0.0
    for a, (_, b, c) in _FRAMES.items():
        na, nb, nc = shape[a], shape[b], shape[c]
        acc = np.zeros((na, nb, nc))
        for k in [k for k in range(n) if int(table[k]["axis"]) == a]:
            P = table[k]
            for i in range(na):
                R, C, J, K, w = _fpt_taps(P, i, rows, cols, nb, nc)
                hit, cand = _fpt_covered(P, i, R, C, J, K, rows, cols, eps)
                missed += int((~hit).sum())
                if kernel_weights:
                    R, C, J, K, w = (x[hit] for x in (R, C, J, K, w))
                    wk = _fpt_kernel_weights(P, i, R, C, J, K, rows, cols,
                                             nb)
                    a1, _, a2, _ = _fp_rays(P, R.astype(np.float64),
                                            C.astype(np.float64))
                    wr = _fp_weight(P, a1, a2).astype(np.float64)
                    if len(w):
                        werr = max(werr, float(np.abs(wk * wr - w).max()
                                               / wr.max()))
                    np.add.at(acc[i], (J, K),
                              proj[k, R, C].astype(f32) * wk * wr)
                else:
                    np.add.at(acc[i], (J[hit], K[hit]),
                              proj[k, R[hit], C[hit]].astype(f32) * w[hit])
                frac[int(P["colc"])].append(cand.mean() / (rows * cols))
        out += np.transpose(acc, np.argsort((a, b, c)))
    return (out, missed, {p: float(np.mean(v)) for p, v in frac.items() if v},
            werr)


@pytest.mark.parametrize("pert", [{}, {"tilt_deg": 2.0, "drift_px": 4.0}])
def test_fp_transpose_emulation(pert):
    """The gather transpose's candidate rays cover every nonzero tap of the
    forward kernel, so it equals the oracle's exact transpose; the
    candidates are a small part of the detector (the enumeration prunes)."""
    pgs, vg, _ = _case(n_angles=5, shape=(20, 40, 36), rows=30, cols=44,
                       **pert)
    pgs = list(pgs) + [kk.rotate(pgs[0], pitch=np.pi / 2)]   # + a z frame
    table = _fp_table(pgs, vg)
    assert set(table["axis"].tolist()) >= {2}
    assert bool(table["colc"][:5].all()) == (not pert)
    y = np.random.default_rng(3).uniform(size=(len(pgs), 30, 44))
    got, missed, frac, _ = emulate_fp_t(y, table, vg.shape)
    assert missed == 0
    assert all(f < 0.25 for f in frac.values()), frac
    ref = oracle.cone_fp_transpose(y, pgs, vg)
    assert_parity(got, ref, label=f"fp_t-emu {pert}")
    # the kernel's float32 offsets from float64 anchors and its tent
    # weights: every tap weight within 5e-6 (of the largest ray weight) of
    # the forward kernel's float64-coordinate taps
    got, _, _, werr = emulate_fp_t(y, table, vg.shape, kernel_weights=True)
    assert werr < 5e-6, werr
    assert_parity(got, ref, label=f"fp_t-emu kernel weights {pert}")
    # windows a quarter voxel short on each side lose taps: the coverage
    # check can fail
    assert emulate_fp_t(y, table, vg.shape, eps=-0.25)[1] > 0


def _inner_range(a, beta, n1, lo, hi):
    """cone_fp.hip inner_range() in float32, vectorised over rays."""
    a, beta = a.astype(f32), beta.astype(f32)
    n1 = f32(n1)
    flat = np.abs(a) < f32(1e-3)
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        e0 = (a * lo + beta).astype(f32)
        e1 = (a * hi + beta).astype(f32)
        ok_flat = (np.minimum(e0, e1) >= 1) & (np.maximum(e0, e1) <= n1 - 1)
        inv = (f32(1) / a).astype(f32)
        t0 = ((f32(1) - beta) * inv).astype(f32)
        t1 = ((n1 - f32(1) - beta) * inv).astype(f32)
    lo2 = np.where(flat, np.where(ok_flat, lo, f32(3e38)),
                   np.maximum(lo, np.minimum(t0, t1)))
    hi2 = np.where(flat, np.where(ok_flat, hi, f32(-3e38)),
                   np.minimum(hi, np.maximum(t0, t1)))
    return lo2.astype(f32), hi2.astype(f32)


@pytest.mark.parametrize("name,angles", [
    ("full", (0, 45, 90, 137, 300, 719)), ("diff", (0, 13, 200)),
    ("large", (0, 611, 1000)), ("zmajor", (0, 1, 2, 3))])
def test_fp_unclamped_interior_is_inside(name, angles):
    """The forward kernel traces the slices where all of a wave's rays are
    inside the volume without clamping.  Their range comes from a float32
    estimate (inner_range) with one voxel of margin; here every slice it
    admits has both exact coordinates at least half a voxel inside the
    clamp range [0, n + 1], so the unclamped loads stay in bounds."""
    if name == "zmajor":
        g0 = kk.rotate(workloads.circular_geometry(
            1, sod=120.0, odd=60.0, rows=24, cols=28, pixel=1.2)[0],
            pitch=np.pi / 2)
        pgs = [kk.rotate(g0, yaw=t) for t in
               np.linspace(0, 2 * np.pi, 4, False)]
        vg = kk.resolve_volume_geometry(shape=(16, 20, 12), voxel_size=1.0)
    else:
        pgs, vg = workloads.config_geometry(name)
    sub = [pgs[k] for k in angles]
    t = _fp_table(sub, vg)
    rows, cols = int(sub[0].detector.rows), int(sub[0].detector.cols)
    R, C = np.meshgrid(np.arange(rows, dtype=np.float64),
                       np.arange(cols, dtype=np.float64), indexing="ij")
    checked = 0
    for P in t:
        na, nb, nc = int(P["na"]), int(P["nb"]), int(P["nc"])
        a1, b1, a2, b2 = (v.ravel() for v in _fp_rays(P, R, C))
        ok = (np.isfinite(a1) & np.isfinite(b1) & np.isfinite(a2)
              & np.isfinite(b2))
        a1, b1, a2, b2 = a1[ok], b1[ok], a2[ok], b2[ok]
        lo = np.zeros(a1.shape, f32)
        hi = np.full(a1.shape, f32(na - 1))
        lo, hi = _inner_range(a1, b1, nb + 1, lo, hi)
        lo, hi = _inner_range(a2, b2, nc + 1, lo, hi)
        with np.errstate(invalid="ignore", over="ignore"):
            jlo = np.maximum(0, np.ceil(lo).clip(-1, na).astype(np.int64))
            jhi = np.minimum(na - 1,
                             np.floor(hi).clip(-1, na).astype(np.int64))
        for i in (0, na // 3, na // 2, na - 1):
            m = (jlo <= i) & (i <= jhi)
            xb = a1[m] * i + b1[m]
            xc = a2[m] * i + b2[m]
            assert np.all((xb >= 0.5) & (xb <= nb + 0.5)), (name, i)
            assert np.all((xc >= 0.5) & (xc <= nc + 0.5)), (name, i)
            checked += int(m.sum())
        # the end points of each admitted range too
        for j in (jlo, jhi):
            m = jlo <= jhi
            xb = a1[m] * j[m] + b1[m]
            xc = a2[m] * j[m] + b2[m]
            assert np.all((xb >= 0.5) & (xb <= nb + 0.5)), name
            assert np.all((xc >= 0.5) & (xc <= nc + 0.5)), name
    assert checked > 0


def _emulate_filter_fft(x, he_bitrev):
    """filter_rows_kernel's transform in numpy (float64, same index
    arithmetic): radix-4 decimation-in-frequency passes (plus a span-1
    stage for odd log2 N), multiplication by the bit-reversed response,
    radix-4 decimation-in-time passes (plus a first span-1 stage)."""
    buf = np.array(x, dtype=np.complex128)
    N = buf.size
    half, quarter = N // 2, N // 4
    log2n = N.bit_length() - 1
    tw = np.exp(-2j * np.pi * np.arange(half) / N)
    g = np.arange(quarter)
    m, ts = half, 1
    while m >= 2:
        q = m // 2
        k = g & (q - 1)
        ia = ((g - k) << 2) + k
        xa, xb, xc, xd = buf[ia], buf[ia + q], buf[ia + m], buf[ia + m + q]
        a1, c1 = xa + xc, (xa - xc) * tw[k * ts]
        b1, d1 = xb + xd, (xb - xd) * tw[(k + q) * ts]
        w2 = tw[k * (ts << 1)]
        buf[ia], buf[ia + q] = a1 + b1, (a1 - b1) * w2
        buf[ia + m], buf[ia + m + q] = c1 + d1, (c1 - d1) * w2
        m, ts = m >> 2, ts << 2
    if m == 1:
        x0, x1 = buf[0::2].copy(), buf[1::2].copy()
        buf[0::2], buf[1::2] = x0 + x1, x0 - x1
    buf *= he_bitrev
    m, ts = 1, half
    if log2n & 1:
        x0, x1 = buf[0::2].copy(), buf[1::2].copy()
        buf[0::2], buf[1::2] = x0 + x1, x0 - x1
        m, ts = 2, half >> 1
    while m < N:
        k = g & (m - 1)
        ia = ((g - k) << 2) + k
        xa, xb = buf[ia], buf[ia + m]
        xc, xd = buf[ia + 2 * m], buf[ia + 3 * m]
        w1 = np.conj(tw[k * ts])
        tb, td = xb * w1, xd * w1
        a1, b1, c1, d1 = xa + tb, xa - tb, xc + td, xc - td
        tc = c1 * np.conj(tw[k * (ts >> 1)])
        tdd = d1 * np.conj(tw[(k + m) * (ts >> 1)])
        buf[ia], buf[ia + 2 * m] = a1 + tc, a1 - tc
        buf[ia + m], buf[ia + 3 * m] = b1 + tdd, b1 - tdd
        m, ts = m << 2, ts >> 2
    return buf


@pytest.mark.parametrize("N", [4, 8, 64, 512, 1024, 2048, 4096])
def test_filter_radix4_index_arithmetic(N):
    """The row filter's radix-4 passes equal circular convolution with the
    (even) response: IDFT(DFT(x) He) for a complex row (two real rows)."""
    rng = np.random.default_rng(N)
    x = rng.normal(size=N) + 1j * rng.normal(size=N)
    H = rng.uniform(size=N)
    he = 0.5 * (H + H[(-np.arange(N)) % N])
    bits = N.bit_length() - 1
    rev = np.array([int(format(i, f"0{bits}b")[::-1], 2) for i in range(N)])
    got = _emulate_filter_fft(x, he[rev] / N)
    ref = np.fft.ifft(np.fft.fft(x) * he)
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-12 * np.abs(x).max())


def _xcd_remap(bid, total, group, n_xcd=8):
    """kk_xcd_remap of kk_internal.h on an array of hardware ids."""
    bid = np.asarray(bid, np.int64)
    if group <= 1:
        return bid.copy()
    span = n_xcd * group
    chunk = bid // span
    r = bid - chunk * span
    out = chunk * span + (r % n_xcd) * group + r // n_xcd
    return np.where((chunk + 1) * span > total, bid, out)


@pytest.mark.parametrize("total,group", [(1, 64), (511, 64), (512, 64),
                                         (513, 64), (1000003, 64),
                                         (777, 30), (4096, 8), (90, 1)])
def test_xcd_remap_is_a_bijection_with_runs_per_xcd(total, group):
    """Every logical tile is run exactly once, and inside full chunks the
    ids dealt to one XCD (hardware id mod 8) are `group` consecutive
    logical ids."""
    hw = np.arange(total, dtype=np.int64)
    logical = _xcd_remap(hw, total, group)
    assert np.array_equal(np.sort(logical), hw)
    span = 8 * group
    full = (total // span) * span if group > 1 else 0
    for c0 in range(0, min(full, 4 * span), span):
        for x in range(8):
            ids = logical[c0 + x:c0 + span:8]
            assert np.array_equal(ids, c0 + x * group + np.arange(group))


def _fp_tile_decode(bid, n_ang, col_tiles, ga, gc):
    """fp_tile's logical id -> (angle, column tile, row band): groups of
    ga angles x gc column tiles (angle fastest inside), groups ordered
    angle group -> column-tile group -> row band; ids past the angles or
    the column tiles pad the grid to whole groups."""
    inner, outer = bid % (ga * gc), bid // (ga * gc)
    nag = -(-n_ang // ga)
    ncg = -(-col_tiles // gc)
    ang = (outer % nag) * ga + inner % ga
    t = outer // nag
    ct = (t % ncg) * gc + inner // ga
    return ang, ct, t // ncg


def _group_size(n, cap):
    """cone_fp.hip's group_size: the g in [ceil(cap/2), cap] whose whole
    groups pad n least (the larger g on ties), n itself if n <= cap."""
    if n <= cap:
        return n
    return min(range(cap, (cap + 1) // 2 - 1, -1),
               key=lambda g: -(-n // g) * g - n)


@pytest.mark.parametrize("n,cap,want", [
    (720, 16, 16), (1440, 16, 16), (360, 16, 15), (180, 16, 15),
    (181, 16, 14), (17, 16, 9), (7, 16, 7), (48, 8, 8), (32, 8, 8),
    (10, 8, 5), (9, 8, 5), (40, 16, 10), (3, 8, 3)])
def test_fp_group_size_pads_least(n, cap, want):
    """Group sizes of the forward kernel's workgroup order: the least
    padding over [ceil(cap/2), cap], brute-forced here against the closed
    rule, and the chosen sizes for the workloads' angle and column-tile
    counts."""
    g = _group_size(n, cap)
    assert g == want
    if n > cap:
        pads = {h: -(-n // h) * h - n for h in range((cap + 1) // 2, cap + 1)}
        best = min(pads.values())
        assert pads[g] == best and g == max(h for h in pads
                                            if pads[h] == best)


def _xcd_live_tiles(n_ang, col_tiles, row_tiles, ga, gc):
    """Live (angle, column tile, row band) workgroups per XCD when the
    hardware deals workgroup ids round-robin to the 8 XCDs."""
    total = (row_tiles * (-(-n_ang // ga)) * ga *
             (-(-col_tiles // gc)) * gc)
    hw = np.arange(total, dtype=np.int64)
    ang, ct, _ = _fp_tile_decode(_xcd_remap(hw, total, ga * gc), n_ang,
                                 col_tiles, ga, gc)
    live = (ang < n_ang) & (ct < col_tiles)
    return np.bincount(hw[live] % 8, minlength=8)


def test_fp_groups_balance_the_xcds_at_the_small_workload():
    """128^3 / 180 angles / 160^2 detector: 10 column tiles of 16 and one
    row band.  Groups of 16 x 8 pad to 192 angles and 16 column tiles and
    deal the XCDs 168 to 288 live workgroups (mean 225: the busiest XCD
    has 28% more than its share); the least-padding groups (15 x 5) deal
    every XCD 225."""
    old = _xcd_live_tiles(180, 10, 1, 16, 8)
    assert (old.min(), old.max(), old.sum()) == (168, 288, 1800)
    ga, gc = _group_size(180, 16), _group_size(10, 8)
    assert (ga, gc) == (15, 5)
    new = _xcd_live_tiles(180, 10, 1, ga, gc)
    assert new.sum() == 1800 and new.min() == new.max()


@pytest.mark.parametrize("n_ang,col_tiles,row_tiles,ga,gc", [
    (720, 48, 6, 16, 8),      # 512^3 / 720 angles, 768^2: no padding
    (180, 10, 1, 15, 5),      # 128^3 / 180, 160^2: least-padding groups
    (360, 24, 3, 16, 8),      # 256^3 / 360, 384^2
    (45, 10, 2, 16, 8),       # padding along angles and column tiles
    (7, 3, 1, 7, 3),          # groups clipped to the problem
    (90, 5, 2, 16, 5),        # a batch problem with 5 column tiles
    (90, 40, 2, 16, 16)])     # batch groups
def test_fp_tile_decode_covers_every_tile_once(n_ang, col_tiles, row_tiles,
                                               ga, gc):
    """The forward kernel's workgroup decode composed with the XCD remap
    (groups of ga x gc ids per XCD) visits every (angle, column tile, row
    band) once; the padding ids fall outside."""
    total = (row_tiles * (-(-n_ang // ga)) * ga *
             (-(-col_tiles // gc)) * gc)
    bid = _xcd_remap(np.arange(total, dtype=np.int64), total, ga * gc)
    ang, ct, rt = _fp_tile_decode(bid, n_ang, col_tiles, ga, gc)
    live = (ang < n_ang) & (ct < col_tiles)
    assert rt.max() == row_tiles - 1
    key = (rt[live] * col_tiles + ct[live]) * n_ang + ang[live]
    assert np.array_equal(np.sort(key),
                          np.arange(n_ang * col_tiles * row_tiles))


def _uniform_i2d_bits(i):
    """kk_uniform_i2d of kk_internal.h: the float64 bits of int32 i >= 0
    assembled from the leading-one position (0 maps to +0.0)."""
    u = np.asarray(i, dtype=np.uint64)
    clz = np.array([32 - int(v).bit_length() if v else 31 for v in u | 1],
                   dtype=np.uint64)   # __builtin_clz(u | 1)
    e = np.uint64(31) - clz
    bits = (u << (np.uint64(52) - e)) + ((np.uint64(1022) + e) << np.uint64(52))
    return np.where(u == 0, np.uint64(0), bits)


def test_uniform_i2d_matches_int_to_double():
    """The slice index's scalar float64 construction equals (double)i bit
    for bit over the int32 range (0, powers of two and their neighbours,
    the maximum, random values)."""
    rng = np.random.default_rng(3)
    p2 = np.array([1 << k for k in range(31)], dtype=np.int64)
    vals = np.unique(np.concatenate([
        np.arange(0, 4097), p2, p2 - 1, p2 + 1, [2 ** 31 - 1],
        rng.integers(0, 2 ** 31, 200000)]))
    vals = vals[(vals >= 0) & (vals < 2 ** 31)]
    got = _uniform_i2d_bits(vals.astype(np.uint64))
    want = vals.astype(np.float64).view(np.uint64)
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("name,wobble,angles,slices", [
    ("diff", 0.0, (0, 77), (3, 128, 250)),
    ("diff", 1.0, (5, 200), (0, 131, 255)),
    ("large", None, (9,), (517,))])
def test_fp_transpose_kernel_weights_at_scale(name, wobble, angles, slices):
    """The transpose kernel's float32 tap weights at the differentiation
    (256^3, 384^2, source-origin 1000) and large (1024^3, 1536^2, wobbling)
    workloads: within 5e-6 of the forward kernel's on every tap of the
    sampled slices (a sample of the taps for the large one)."""
    from kernelkit.workloads import CONFIGS, circular_geometry
    c = dict(CONFIGS[name])
    shape = c.pop("shape")
    c.pop("n_angles")
    if wobble is not None:
        c.update(wobble_deg=wobble, drift_px=2.0 if wobble else 0.0)
    n = 1440 if name == "large" else 360
    pgs = circular_geometry(n, seed=1, **c)
    vg = kk.resolve_volume_geometry(shape=shape, voxel_size=1.0)
    pgs = [pgs[k] for k in angles]
    table = _fp_table(pgs, vg)
    rows, cols = int(c["rows"]), int(c["cols"])
    rng = np.random.default_rng(0)
    for P in table:
        a = int(P["axis"])
        _, b, cc = _FRAMES[a]
        nb, nc = shape[b], shape[cc]
        for i in slices:
            R, C, J, K, w = _fpt_taps(P, i, rows, cols, nb, nc)
            if len(R) > 300000:
                pick = rng.choice(len(R), 300000, replace=False)
                R, C, J, K, w = (x[pick] for x in (R, C, J, K, w))
            hit, _ = _fpt_covered(P, i, R, C, J, K, rows, cols)
            assert hit.all()
            wk = _fpt_kernel_weights(P, i, R, C, J, K, rows, cols, nb)
            a1, _, a2, _ = _fp_rays(P, R.astype(np.float64),
                                    C.astype(np.float64))
            wr = _fp_weight(P, a1, a2).astype(np.float64)
            err = np.abs(wk * wr - w) / wr
            assert err.max() < 5e-6, (name, int(P["colc"]), i, err.max())


@pytest.mark.parametrize("tiles", [(32, 32, 64), (6, 5, 3), (1, 1, 2),
                                   (9, 17, 33)])
def test_bp_tile_decode_covers_every_tile_once(tiles):
    """The backprojector's workgroup decode (groups of up to 4 x 4 x 4
    tiles, x fastest inside; groups x -> y -> z), composed with the XCD
    remap of one group per run, visits every (x, y, z) tile once; padding
    ids fall outside."""
    tx_n, ty_n, tz_n = tiles
    gx, gy, gz = (min(4, t) for t in tiles)
    total = (-(-tx_n // gx) * gx) * (-(-ty_n // gy) * gy) * \
        (-(-tz_n // gz) * gz)
    bid = _xcd_remap(np.arange(total, dtype=np.int64), total, gx * gy * gz)
    inner, outer = bid % (gx * gy * gz), bid // (gx * gy * gz)
    ngx, ngy = -(-tx_n // gx), -(-ty_n // gy)
    tx = (outer % ngx) * gx + inner % gx
    o2 = outer // ngx
    ty = (o2 % ngy) * gy + (inner // gx) % gy
    tz = (o2 // ngy) * gz + inner // (gx * gy)
    live = (tx < tx_n) & (ty < ty_n) & (tz < tz_n)
    key = (tz[live] * ty_n + ty[live]) * tx_n + tx[live]
    assert np.array_equal(np.sort(key), np.arange(tx_n * ty_n * tz_n))
