"""Tile-shape autotuner for batched launches.

No counterpart in the reference (its kernels take fixed block shapes chosen
by the caller, ``kernels/cone_fp.py:57-66``, ``kernels/cone_bp.py:52-66``).
For a batch of problems the tuner times the candidate tile shapes of the
gfx950 kernels on the batch itself -- rows per thread of the forward
projector (1, 2, 4), z-run of the backprojector (4, 8, 16) and the size of
the grid (one workgroup per tile, or a persistent grid of a fixed number
of workgroups per CU) -- with HIP events on the launching stream, and keeps
the fastest.  Results are cached per batch signature (problem shapes), in
memory and optionally in the JSON file named by ``KK_TUNE_CACHE``.
"""
from __future__ import annotations

import json
import os
from dataclasses import dataclass
from itertools import product

from kernelkit import _native

FP_ROWS = (1, 2, 4)
BP_ZRUN = (4, 8, 16)
GRIDS_PER_CU = (0, 4, 8)   # workgroups per CU (0: one per tile)


def _grids():
    """Candidate grid sizes for the current device (0: one workgroup per
    tile, else a persistent grid)."""
    n_cu, _ = _native.device_info()
    return tuple(g * n_cu for g in GRIDS_PER_CU)


@dataclass(frozen=True)
class LaunchConfig:
    tile: int      # rows per thread (FP) or z-run (BP)
    grid: int      # workgroups (0: one per tile, else persistent)
    ms: float = float("nan")


class BatchTuner:
    """Pick tile shapes for ``kk_batch_*`` launches by timing them."""

    def __init__(self, reps: int = 2, cache_file: str | None = None):
        self.reps = reps
        self.cache_file = cache_file or os.environ.get("KK_TUNE_CACHE")
        self._cache: dict = {}
        if self.cache_file and os.path.exists(self.cache_file):
            try:
                with open(self.cache_file) as f:
                    for k, v in json.load(f).items():
                        self._cache[k] = tuple(LaunchConfig(**c) for c in v)
            except (OSError, ValueError, TypeError, AttributeError):
                self._cache = {}     # unreadable or foreign file: start empty

    def _store(self, key, cfgs):
        self._cache[key] = cfgs
        if self.cache_file:
            data = {k: [c.__dict__ for c in v] for k, v in self._cache.items()}
            tmp = f"{self.cache_file}.{os.getpid()}.tmp"
            with open(tmp, "w") as f:
                json.dump(data, f)
            os.replace(tmp, self.cache_file)   # readers never see a partial file

    @staticmethod
    def _key(signature) -> str:
        """Cache key: the batch signature, the library build and the device
        (timings do not carry over between builds or devices)."""
        return repr((_native.build_id(), _native.device_info(), signature))

    def _time(self, launch) -> float:
        launch()                       # warm-up
        stream = _native.stream_ptr()
        a, b = _native.Event(), _native.Event()
        ts = []
        for _ in range(self.reps):
            a.record(stream)
            launch()
            b.record(stream)
            b.synchronize()
            ts.append(a.elapsed_ms(b))
        return min(ts)

    def _tune(self, handle, which, other_tile):
        lib = _native.lib()
        best = None
        tiles = FP_ROWS if which == "fp" else BP_ZRUN
        fn = lib.kk_batch_fp if which == "fp" else lib.kk_batch_bp
        for tile, grid in product(tiles, _grids()):
            rpt, zrun = (tile, other_tile) if which == "fp" else \
                (other_tile, tile)
            _native.check(lib.kk_batch_configure(handle.ptr, rpt, zrun))
            ms = self._time(lambda: _native.check(
                fn(handle.ptr, grid, _native.stream_ptr())))
            if best is None or ms < best.ms:
                best = LaunchConfig(tile, grid, ms)
        return best

    def configure(self, handle, which: str, signature, tune: bool = True):
        """Configure ``handle`` for its forward (``which='fp'``) or backward
        (``'bp'``) problems and return the launch configuration."""
        key = self._key(signature)
        cfgs = self._cache.get(key)
        if cfgs is None:
            if tune:
                cfg = self._tune(handle, which, 8 if which == "fp" else 0)
            else:
                cfg = LaunchConfig(0 if which == "fp" else 8, 0)
            cfgs = (cfg,)
            if tune:
                self._store(key, cfgs)
        cfg = cfgs[0]
        rpt, zrun = (cfg.tile, 8) if which == "fp" else (0, cfg.tile)
        _native.check(_native.lib().kk_batch_configure(handle.ptr, rpt,
                                                       zrun))
        return cfg

    def configure_pair(self, handle, signature, tune: bool = True):
        """Configure a batch holding both forward and backward problems
        (a SIRT batch); returns ``(fp_config, bp_config)``.  Tuning runs the
        launches for real: the caller resets any state they update."""
        key = self._key(signature)
        cfgs = self._cache.get(key)
        if cfgs is None:
            if tune:
                fp = self._tune(handle, "fp", 8)
                bp = self._tune(handle, "bp", fp.tile)
            else:
                fp, bp = LaunchConfig(0, 0), LaunchConfig(8, 0)
            cfgs = (fp, bp)
            if tune:
                self._store(key, cfgs)
        fp, bp = cfgs
        _native.check(_native.lib().kk_batch_configure(handle.ptr, fp.tile,
                                                       bp.tile))
        return fp, bp
