"""The FP64 oracle over a pool of host processes (test infrastructure).

Forward projection and backprojection are exactly linear in the angles, so
a workload splits into angle subsets that run side by side, one process of
one BLAS thread each: FP results are concatenated, BP results summed.
Workers are spawned (not forked: the parent may hold a HIP context) and
import only numpy and the oracle; geometries cross as plain records.
"""
from __future__ import annotations

import multiprocessing as mp
import os
from types import SimpleNamespace

import numpy as np

_POOL = None


def processes() -> int:
    """Worker count: the host's cores, at most 16."""
    return max(1, min(16, os.cpu_count() or 1))


def _plain_pg(g):
    det = g.detector
    return SimpleNamespace(
        source_position=np.asarray(g.source_position, np.float64),
        detector_position=np.asarray(g.detector_position, np.float64),
        u=np.asarray(g.u, np.float64), v=np.asarray(g.v, np.float64),
        detector=SimpleNamespace(rows=int(det.rows), cols=int(det.cols),
                                 pixel_width=float(det.pixel_width),
                                 pixel_height=float(det.pixel_height)))


def _plain_vg(vg):
    return SimpleNamespace(
        shape=tuple(int(s) for s in vg.shape),
        voxel_size=tuple(float(s) for s in vg.voxel_size),
        extent_min=tuple(float(s) for s in vg.extent_min),
        extent_max=tuple(float(s) for s in vg.extent_max),
        rotation=tuple(float(s) for s in getattr(vg, "rotation",
                                                 (0.0, 0.0, 0.0))))


def _run(task):
    from threadpoolctl import threadpool_limits

    import oracle
    kind, data, pgs, vg, kw = task
    with threadpool_limits(limits=1):
        if kind == "fp":
            return oracle.cone_fp(data, pgs, vg)
        if kind == "bp":
            return oracle.cone_bp(data, pgs, vg, **kw)
        if kind == "t":
            return oracle.cone_fp_transpose(data, pgs, vg)
        raise ValueError(kind)


def _pool():
    global _POOL
    if _POOL is None:
        _POOL = mp.get_context("spawn").Pool(processes())
    return _POOL


def _chunks(n, parts):
    parts = max(1, min(parts, n))
    edges = np.linspace(0, n, parts + 1).round().astype(int)
    return [(int(a), int(b)) for a, b in zip(edges[:-1], edges[1:]) if b > a]


def cone_fp(vol, pgs, vg) -> np.ndarray:
    """``oracle.cone_fp`` over angle subsets in parallel."""
    pgs = [_plain_pg(g) for g in pgs]
    vgp = _plain_vg(vg)
    vol = np.asarray(vol, np.float64)
    tasks = [("fp", vol, pgs[a:b], vgp, {})
             for a, b in _chunks(len(pgs), 4 * processes())]
    return np.concatenate(_pool().map(_run, tasks), axis=0)


def cone_bp(proj, pgs, vg, **kw) -> np.ndarray:
    """``oracle.cone_bp`` over angle subsets in parallel (summed)."""
    pgs = [_plain_pg(g) for g in pgs]
    vgp = _plain_vg(vg)
    proj = np.asarray(proj, np.float64)
    tasks = [("bp", proj[a:b], pgs[a:b], vgp, kw)
             for a, b in _chunks(len(pgs), processes())]
    return np.sum(_pool().map(_run, tasks), axis=0)


def cone_fp_transpose(proj, pgs, vg) -> np.ndarray:
    """``oracle.cone_fp_transpose`` over angle subsets in parallel."""
    pgs = [_plain_pg(g) for g in pgs]
    vgp = _plain_vg(vg)
    proj = np.asarray(proj, np.float64)
    tasks = [("t", proj[a:b], pgs[a:b], vgp, {})
             for a, b in _chunks(len(pgs), processes())]
    return np.sum(_pool().map(_run, tasks), axis=0)


def fdk(proj, pgs, vg, filter="ramlak") -> np.ndarray:
    """``oracle.fdk``: pre-weight and row filter here, the backprojection in
    parallel."""
    import oracle
    pgs = list(pgs)
    y = oracle.filter_rows(oracle.preweight(proj, pgs), filter)
    return cone_bp(y, pgs, vg)


def close():
    global _POOL
    if _POOL is not None:
        _POOL.close()
        _POOL.join()
        _POOL = None
