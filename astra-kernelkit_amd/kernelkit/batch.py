"""Batches of independent cone-beam problems: one launch per operator.

No counterpart in the reference: it runs one geometry per projector
(``projector.py:160-183``) and loops over samples in Python
(``torch_support.py:19-20``).  ``ConeBatch`` keeps one ``XrayTransform`` per
problem (volumes, detectors, angle counts and geometries may all differ) and
runs all forward projections, all backprojections or a whole SIRT iteration
of the batch as single launches of the gfx950 kernels (``kk_batch_*``,
a grid-stride loop over all problems' tiles).  The tile shapes (rows per
thread of the forward kernel, z-run of the backprojector) and the grid
size (one workgroup per tile, or a persistent grid) come from
:class:`~kernelkit.autotune.BatchTuner`.  ``sirt`` can capture its iteration
loop in a HIP graph.
"""
from __future__ import annotations

from typing import Sequence


from kernelkit import _native, device
from kernelkit.autotune import BatchTuner
from kernelkit.kernel import ProjectionTexture, as_device_array
from kernelkit.kernels.cone_fp import _logical_strides, detector_shape
from kernelkit.operator import XrayTransform


class ConeBatch:
    """Independent cone-beam problems ``[(projection_geometry,
    volume_geometry), ...]`` run together.

    Volumes are ``(x, y, z)`` float32 device arrays (torch tensors,
    :class:`kernelkit.DeviceArray` or ``__cuda_array_interface__`` arrays)
    and projections ``(angle, row, col)`` ones (the default axes of
    ``XrayTransform``); new outputs are of the inputs' kind.
    """

    def __init__(self, problems: Sequence[tuple], tuner: BatchTuner = None):
        if len(problems) == 0:
            raise ValueError("empty batch")
        self.ops = [XrayTransform(pg, vg) for pg, vg in problems]
        self.vgs = [vg for _, vg in problems]
        self.range_shapes = [op.range_shape for op in self.ops]
        self.domain_shapes = [op.domain_shape for op in self.ops]
        self.tuner = tuner if tuner is not None else BatchTuner()
        self._fp_cache = None
        self._bp_cache = None

    def __len__(self):
        return len(self.ops)

    # ---------------------------------------------------------------- FP
    def fp(self, volumes: Sequence, projections: Sequence | None = None,
           tune: bool = True) -> list:
        """``projections[p] = A_p volumes[p]`` for every problem, one launch
        (after the per-problem texture conversions)."""
        vols = [as_device_array(v) for v in volumes]
        self._check(vols, self.domain_shapes, "volume")
        if projections is None:
            projections = [device.empty(s, like=vols[0])
                           for s in self.range_shapes]
        projections = [as_device_array(p) for p in projections]
        self._check(projections, self.range_shapes, "projection")
        for op, v in zip(self.ops, vols):
            op.projector.volume = v          # builds the frame textures
        key = tuple(p.data_ptr() for p in projections) + tuple(
            op.projector._texture.ptr for op in self.ops)
        if self._fp_cache is None or self._fp_cache[0] != key:
            b = _native.Batch()
            lib = _native.lib()
            for op, y in zip(self.ops, projections):
                K = op.projector.kernel
                rows, cols = detector_shape(op.projector._proj_geom_seq)
                _native.check(lib.kk_batch_add_fp(
                    b.ptr, K._plan.ptr, _native._vp(op.projector._texture.ptr),
                    _native._vp(y.data_ptr()),
                    _native.i64(_logical_strides(y, (0, 1, 2))), rows, cols,
                    0))
            self._fp_cache = (key, b, None)
        key, b, cfg = self._fp_cache
        if cfg is None:
            cfg = self.tuner.configure(b, "fp", self._signature("fp"),
                                       tune=tune)
            self._fp_cache = (key, b, cfg)
        _native.check(_native.lib().kk_batch_fp(b.ptr, cfg.grid,
                                                _native.stream_ptr()))
        return list(projections)

    # ---------------------------------------------------------------- BP
    def bp(self, projections: Sequence, volumes: Sequence | None = None,
           tune: bool = True) -> list:
        """``volumes[p] = A_p^* projections[p]`` (the voxel-driven
        backprojector) for every problem, one launch."""
        projs = [as_device_array(p) for p in projections]
        self._check(projs, self.range_shapes, "projection")
        if volumes is None:
            volumes = [device.empty(s, like=projs[0])
                       for s in self.domain_shapes]
        volumes = [as_device_array(v) for v in volumes]
        self._check(volumes, self.domain_shapes, "volume")
        for op, y in zip(self.ops, projs):
            op.backprojector.projections = y  # builds the BP textures
        key = tuple(v.data_ptr() for v in volumes) + tuple(
            op.backprojector._textures.ptr for op in self.ops)
        if self._bp_cache is None or self._bp_cache[0] != key:
            b = _native.Batch()
            lib = _native.lib()
            for op, v, vg in zip(self.ops, volumes, self.vgs):
                K = op.backprojector.kernel
                tex = op.backprojector._textures
                n, rows, cols = tex.shape
                _native.check(lib.kk_batch_add_bp(
                    b.ptr, K._plan.ptr, _native._vp(tex.ptr), n, rows, cols,
                    _native._vp(v.data_ptr()),
                    _native.i32(vg.shape),
                    _native.i64(_logical_strides(v, (0, 1, 2))),
                    float(vg.voxel_volume()), 0))
            self._bp_cache = (key, b, None)
        key, b, cfg = self._bp_cache
        if cfg is None:
            cfg = self.tuner.configure(b, "bp", self._signature("bp"),
                                       tune=tune)
            self._bp_cache = (key, b, cfg)
        _native.check(_native.lib().kk_batch_bp(b.ptr, cfg.grid,
                                                _native.stream_ptr()))
        return list(volumes)

    # -------------------------------------------------------------- SIRT
    def sirt_state(self, projections: Sequence, mask: Sequence | None = None,
                   min_constraint: float | None = None,
                   max_constraint: float | None = None) -> "BatchSirt":
        return BatchSirt(self, projections, mask, min_constraint,
                         max_constraint)

    def sirt(self, projections: Sequence, iters: int = 30,
             mask: Sequence | None = None, min_constraint: float = None,
             max_constraint: float = None, graph: bool = True,
             tune: bool = True) -> list:
        """``iters`` SIRT iterations of every problem (as ``algo.sirt``),
        four launches per iteration for the whole batch; with ``graph`` the
        iteration loop is captured once in a HIP graph and replayed.
        Returns the device volumes."""
        st = self.sirt_state(projections, mask, min_constraint,
                             max_constraint)
        st.configure(tune=tune)
        st.run(iters, graph=graph)
        return st.x

    # ----------------------------------------------------------- helpers
    def _signature(self, which):
        return (which,) + tuple(
            (s, r) for s, r in zip(self.domain_shapes, self.range_shapes))

    @staticmethod
    def _check(arrays, shapes, what):
        if len(arrays) != len(shapes):
            raise ValueError(f"{len(arrays)} {what}s for {len(shapes)} "
                             "problems")
        for k, (a, s) in enumerate(zip(arrays, shapes)):
            if tuple(a.shape) != tuple(s):
                raise ValueError(f"{what} {k} has shape {tuple(a.shape)}, "
                                 f"expected {tuple(s)}")


class BatchSirt:
    """Device state of a batched SIRT run (``algo.py:226-368`` per problem):
    ``x_p <- clip(mask_p * (x_p - C_p * BP_p(R_p * (FP_p(x_p) - y_p))))``
    from ``x_p = 0``, with ``C_p = 1/BP_p(1)`` and ``R_p = 1/FP_p(1)``."""

    def __init__(self, batch: ConeBatch, projections, mask, min_constraint,
                 max_constraint):
        self.batch = batch
        lib = _native.lib()
        st = _native.stream_ptr()
        ys = [as_device_array(p) for p in projections]
        batch._check(ys, batch.range_shapes, "projection")
        if mask is not None:
            masks = [as_device_array(m) for m in mask]
            batch._check(masks, batch.domain_shapes, "mask")
        else:
            masks = [None] * len(batch)
        self.y, self.x, self.C, self.R, self.p = ys, [], [], [], []
        self.masks = masks
        self.res_tex = []
        self._xtex = []
        b = _native.Batch()
        for op, y, m, vg in zip(batch.ops, ys, masks, batch.vgs):
            shape = tuple(int(s) for s in vg.shape)
            n, rows, cols = y.shape
            # C = 1/BP(1), R = 1/FP(1), inf -> 0 (plain single launches)
            C = op.T(device.full(y.shape, 1.0, like=y))
            R = op(device.full(shape, 1.0, like=y))
            _native.check(lib.kk_recip_or_zero(_native._vp(C.data_ptr()),
                                               C.numel(), st))
            _native.check(lib.kk_recip_or_zero(_native._vp(R.data_ptr()),
                                               R.numel(), st))
            x = device.zeros(shape, like=y)
            # the state's own FP texture of x: the operator's projector
            # texture is rebuilt by any later call of the operator
            fk = op.projector.kernel
            op.projector._require_geometry()
            xtex = fk.volume_texture(x, vg)
            rtex = ProjectionTexture(n, rows, cols, like=y)
            rtex.zero()
            bk = op.backprojector.kernel
            # residual in two steps: the batch FP into a buffer, then one
            # coalesced (p - y) R pass per problem into its BP texture (the
            # fused epilogue, kk_batch_add_sirt_fp, reads y and R one
            # detector row per lane, which costs more on MI355X)
            p = device.empty((n, rows, cols), like=y)
            _native.check(lib.kk_batch_add_fp(
                b.ptr, fk._plan.ptr, _native._vp(xtex.ptr),
                _native._vp(p.data_ptr()), _native.i64(p.stride()), rows,
                cols, 0))
            _native.check(lib.kk_batch_add_sirt_bp(
                b.ptr, bk._plan.ptr, _native._vp(rtex.ptr), n, rows, cols,
                _native._vp(x.data_ptr()), _native.i32(shape),
                _native.i64(x.stride()), _native._vp(C.data_ptr()),
                _native.i64(C.stride()),
                None if m is None else _native._vp(m.data_ptr()),
                None if m is None else _native.i64(m.stride()),
                0.0 if min_constraint is None else float(min_constraint),
                0.0 if max_constraint is None else float(max_constraint),
                0 if min_constraint is None else 1,
                0 if max_constraint is None else 1,
                None, 0,   # frames of x: rebuilt after the update
                float(vg.voxel_volume())))
            # the texture builds between the two operators, batched too
            _native.check(lib.kk_batch_add_residual(
                b.ptr, _native._vp(p.data_ptr()), _native.i64(p.stride()),
                _native._vp(y.data_ptr()), _native.i64(y.stride()),
                _native._vp(R.data_ptr()), _native.i64(R.stride()), n, rows,
                cols, _native._vp(rtex.ptr)))
            _native.check(lib.kk_batch_add_volume_texture(
                b.ptr, _native._vp(x.data_ptr()), _native.i32(shape),
                _native.i64(x.stride()), xtex.mask, _native._vp(xtex.ptr)))
            self.x.append(x)
            self.C.append(C)
            self.R.append(R)
            self.p.append(p)
            self.res_tex.append(rtex)
            self._xtex.append(xtex)
        self.handle = b
        self.cfg_fp = self.cfg_bp = None
        self._graph = None

    def reset(self):
        """Back to ``x = 0`` (volumes and their FP textures)."""
        for x, t in zip(self.x, self._xtex):
            x.zero_()
            t.data.zero_()
        self._graph = None

    def configure(self, tune: bool = True):
        """Choose tile shapes (tuning runs real iterations, then resets)."""
        sig = self.batch._signature("sirt")
        self.cfg_fp, self.cfg_bp = self.batch.tuner.configure_pair(
            self.handle, sig, tune=tune)
        self.reset()

    def iterate(self):
        """One SIRT iteration of every problem in four launches whatever
        the number of problems: all forward projections, all residual
        textures ``R (A x - y)``, all backprojections with the update
        epilogue, all rebuilds of the x's forward-projection frames
        (coalesced texture kernels; the update epilogue's scattered frame
        stores cost several times more)."""
        if self.cfg_fp is None:
            self.configure(tune=False)
        lib = _native.lib()
        st = _native.stream_ptr()
        h = self.handle.ptr
        _native.check(lib.kk_batch_fp(h, self.cfg_fp.grid, st))
        _native.check(lib.kk_batch_residual(h, st))
        _native.check(lib.kk_batch_bp(h, self.cfg_bp.grid, st))
        _native.check(lib.kk_batch_volume_textures(h, st))

    def capture(self, iters: int):
        """Capture ``iters`` iterations in a HIP graph (the library's own
        stream capture, :class:`kernelkit._native.Graph`); replaying it
        continues from the current ``x``."""
        if self.cfg_fp is None:
            self.configure(tune=False)

        def body():
            for _ in range(iters):
                self.iterate()
        g = _native.Graph.capture(body)
        self._graph = (iters, g)
        return g

    def run(self, iters: int, graph: bool = True):
        if graph:
            if self._graph is None or self._graph[0] != iters:
                self.capture(iters)
            self._graph[1].replay()
        else:
            for _ in range(iters):
                self.iterate()
