"""Cone-beam FP/BP throughput on one MI355X: 512^3 volume, 720 angles,
768x768 detector (SOD 1000, ODD 500, pixel 1.0), seeded 10-ellipsoid
phantom, FP32.

One step = one forward projection of the volume into all 720 projections
plus one voxel-driven backprojection of those projections into the volume,
each through the projector API with resident device inputs: the texture
conversion of the input (volume -> FP frame textures, projections -> BP
texture) and the kernel.  ``value`` is GUPS over the step:
``2 * 512^3 * 720 / t_step`` (an update is one (voxel, angle) pair; the step
does two passes of updates).

The JSON line also carries:

* ``roofline`` for the dominant kernel (the forward projection): achieved
  gather bytes per second = 16 B per bilinear sample x samples per launch /
  average kernel time (HIP events on the launching stream), against the
  gather ceiling: the vector-memory return path of 64 B/clk per CU x 256
  CUs x 2.4 GHz, next to this run's probe rates of 16-byte gathers in the
  projectors' access shape (``kk_probe_gather``) from L1-, L2- and
  HBM-resident footprints;
  ``traffic`` is the HBM bytes per launch from the rocprofv3 FETCH_SIZE +
  WRITE_SIZE counters (scaled by a plain-copy calibration), read from
  ``profiles/cone_fp_bp_512cube_720_traffic.json`` when present, and
  ``traffic_GBs`` / ``traffic_frac_of_hbm`` that amount at this run's
  kernel time against the 8 TB/s HBM peak.  The samples are the
  algorithm's (every ray samples every slice of its major axis, zeros
  outside the volume included); the kernel skips the samples outside its
  wavefronts' ray supports, so ``frac`` can exceed 1.  ``issued_bytes``
  (16 B per gather the kernel issues, from the SQ_INSTS_VMEM_RD counter in
  ``profiles/cone_fp_bp_512cube_720_gathers.json``) and ``issued_frac``
  give the rate of the gathers it does issue against the same ceiling;
* ``cpu_baseline``: the FP64 NumPy oracle in 16 host processes of one core
  each (``single_core``: one process), timed on a bounded sample of the
  same workload (one angle, a band of detector rows for FP and a slab of
  the volume for BP), in GUPS.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "astra-kernelkit_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = ("cone FP/BP GUPS at 512³/720 proj; batched + autograd vs "
          "plain, same GPU")
L2_PEAK_GBS = 8 * 16 * 128 * 2.4  # GB/s: 8 XCD x 16 ch x 128 B/clk x 2.4 GHz
# GB/s: vector-memory return path of 64 B/clk per CU x 256 CUs x 2.4 GHz,
# the rate at which 16-byte gathers reach the registers even from L1
L1_RETURN_GBS = 256 * 64 * 2.4
HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles",
                            "cone_fp_bp_512cube_720_traffic.json")
GATHERS_FILE = os.path.join(ROOT, "profiles",
                            "cone_fp_bp_512cube_720_gathers.json")


def _log(msg):
    print(msg, file=sys.stderr, flush=True)


def fp_samples(pgs, vg) -> int:
    """Bilinear samples of one forward projection: every ray samples each
    slice along its angle's major axis."""
    import numpy as np
    from kernelkit.kernels.cone_fp import RayDrivenConeFP
    axes = RayDrivenConeFP.geoms2params(list(pgs), vg).ray_axes
    g = pgs[0]
    rays = int(g.detector.rows) * int(g.detector.cols)
    shape = np.asarray(vg.shape, dtype=np.int64)
    return int(rays * shape[axes].sum())


def _cpu_sample_geometry(pgs, vg):
    """The bounded CPU sample: the middle half of angle 0's detector rows
    (FP) and the middle half-slab of the volume along x (BP)."""
    import numpy as np

    import kernelkit as kk
    from kernelkit.geom.proj import Detector, ProjectionGeometry

    g = pgs[0]
    rows, cols = int(g.detector.rows), int(g.detector.cols)
    band = rows // 2
    r0 = rows // 2 - band // 2
    shift = (r0 + band / 2.0 - rows / 2.0) * float(g.detector.pixel_height)
    g_band = ProjectionGeometry(
        g.source_position, np.asarray(g.detector_position) + shift *
        np.asarray(g.v), Detector(band, cols, g.detector.pixel_height,
                                  g.detector.pixel_width), u=g.u, v=g.v)
    nx = int(vg.shape[0])
    slab = nx // 2
    vs = np.asarray(vg.voxel_size, dtype=np.float64)
    emin = np.asarray(vg.extent_min, dtype=np.float64)
    emax = np.asarray(vg.extent_max, dtype=np.float64)
    x0 = (nx - slab) // 2
    smin = emin.copy()
    smax = emax.copy()
    smin[0] = emin[0] + x0 * vs[0]
    smax[0] = smin[0] + slab * vs[0]
    vg_slab = kk.VolumeGeometry((slab,) + tuple(vg.shape[1:]), vs, smin, smax)
    return g, g_band, vg_slab, rows, cols, band, r0, nx, slab


def _cpu_task(args):
    """One host process of the baseline: the FP64 oracle on one core, FP
    of the volume (shared memory) into a detector band, or BP of a band
    into a slab.  Returns the seconds it took."""
    import numpy as np
    from multiprocessing import shared_memory
    from threadpoolctl import threadpool_limits

    import oracle

    kind, shm_name, shape, geom, vg, arg = args
    with threadpool_limits(limits=1):
        if kind == "fp":
            shm = shared_memory.SharedMemory(name=shm_name)
            try:
                vol = np.ndarray(shape, dtype=np.float64, buffer=shm.buf)
                t0 = time.perf_counter()
                oracle.cone_fp(vol, [geom], vg)
                dt = time.perf_counter() - t0
                del vol
            finally:
                shm.close()
        else:
            t0 = time.perf_counter()
            oracle.cone_bp(arg, [geom], vg)
            dt = time.perf_counter() - t0
    return dt


def cpu_baseline(pgs, vg, vol, workers=16):
    """FP64 oracle over a bounded sample of the workload: on one core, and
    on ``workers`` host processes (one core each) running the same sample
    side by side (FP and BP are independent per angle, so this is how the
    oracle scales over cores)."""
    import multiprocessing as mp
    from concurrent.futures import ProcessPoolExecutor
    from multiprocessing import shared_memory

    import numpy as np
    from threadpoolctl import threadpool_limits

    import oracle

    g, g_band, vg_slab, rows, cols, band, r0, nx, slab = \
        _cpu_sample_geometry(pgs, vg)
    vol64 = vol.astype(np.float64)
    with threadpool_limits(limits=1):
        t0 = time.perf_counter()
        y = oracle.cone_fp(vol64, [g_band], vg)
        t_fp = time.perf_counter() - t0
        proj = np.zeros((1, rows, cols))
        proj[0, r0:r0 + band] = y[0]
        t0 = time.perf_counter()
        oracle.cone_bp(proj, [g], vg_slab)
        t_bp = time.perf_counter() - t0
    n_vox = int(np.prod(vg.shape))
    # per-angle time of the whole workload, by exact linearity in rays and
    # in voxels
    t_fp_angle = t_fp * rows / band
    t_bp_angle = t_bp * nx / slab
    gups1 = 2 * n_vox / (t_fp_angle + t_bp_angle) / 1e9
    single = {
        "value": round(gups1, 6), "cores": 1,
        "sample": (f"FP over {band} of {rows} detector rows "
                   f"({t_fp:.1f} s), BP over a {slab}-slice x-slab of the "
                   f"{nx}^3 volume ({t_bp:.1f} s)")}
    workers = max(1, min(workers, os.cpu_count() or 1))
    shm = shared_memory.SharedMemory(create=True, size=vol64.nbytes)
    try:
        np.ndarray(vol64.shape, np.float64, buffer=shm.buf)[...] = vol64
        # a worker that dies raises BrokenProcessPool here (no hang)
        with ProcessPoolExecutor(workers, mp_context=mp.get_context(
                "spawn")) as pool:
            list(pool.map(_cpu_task, [("bp", None, None, g, vg_slab,
                                       proj[:, :1, :1])] * workers,
                          timeout=120))   # start the workers
            t0 = time.perf_counter()
            list(pool.map(_cpu_task, [("fp", shm.name, vol64.shape, g_band,
                                       vg, None)] * workers, timeout=600))
            w_fp = time.perf_counter() - t0
            t0 = time.perf_counter()
            list(pool.map(_cpu_task, [("bp", None, None, g, vg_slab, proj)]
                          * workers, timeout=600))
            w_bp = time.perf_counter() - t0
    finally:
        shm.close()
        shm.unlink()
    # `workers` samples in each wall time
    t_angle = (w_fp * rows / band + w_bp * nx / slab) / workers
    gups = 2 * n_vox / t_angle / 1e9
    return {
        "value": round(gups, 6), "unit": "GUPS", "cores": workers,
        "kind": "port",
        "sample": (f"FP64 NumPy oracle, {workers} host processes of one "
                   f"core each, each running 1 angle's sample side by "
                   f"side: FP over {band} of {rows} detector rows "
                   f"({w_fp:.1f} s wall), BP over a {slab}-slice x-slab of "
                   f"the {nx}^3 volume ({w_bp:.1f} s wall); scaled to full "
                   f"angles of FP+BP by ray/voxel count"),
        "single_core": single,
        # wall seconds per angle of the full workload on `workers` cores
        "fp_s_per_angle": round(w_fp * rows / band / workers, 4),
        "bp_s_per_angle": round(w_bp * nx / slab / workers, 4),
    }


def _small_task(args):
    """One host process of the small-workload baseline (one core): ``fp``
    projects the shared volume into its angles of the shared projection
    array; ``bp`` backprojects its angles of the shared projections into
    its own slot of the shared partial volumes."""
    import numpy as np
    from multiprocessing import shared_memory
    from threadpoolctl import threadpool_limits

    import oracle

    kind, names, shapes, geoms, a0, slot, vg = args
    if kind == "start":        # imports done, process up
        return 0
    shms = [shared_memory.SharedMemory(name=n) for n in names]
    try:
        vol, proj, parts = (np.ndarray(sh, np.float64, buffer=m.buf)
                            for m, sh in zip(shms, shapes))
        with threadpool_limits(limits=1):
            if kind == "fp":
                proj[a0:a0 + len(geoms)] = oracle.cone_fp(vol, geoms, vg)
            else:
                parts[slot] = oracle.cone_bp(proj[a0:a0 + len(geoms)],
                                             geoms, vg)
        del vol, proj, parts
    finally:
        for m in shms:
            m.close()
    return 0


class _SmallPool:
    """FP64 oracle FP/BP over angle chunks on ``workers`` host processes
    (one core each); FP and BP are sums over angles, so the chunks are
    independent and their results add up exactly."""

    def __init__(self, pgs, vg, workers):
        import multiprocessing as mp
        from concurrent.futures import ProcessPoolExecutor
        from multiprocessing import shared_memory

        import numpy as np
        self.pgs, self.vg, self.workers = list(pgs), vg, workers
        n = len(self.pgs)
        det = (int(self.pgs[0].detector.rows), int(self.pgs[0].detector.cols))
        vshape = tuple(int(s) for s in vg.shape)
        self.shapes = [vshape, (n,) + det, (workers,) + vshape]
        self.shms = [shared_memory.SharedMemory(
            create=True, size=8 * int(np.prod(sh))) for sh in self.shapes]
        self.arrays = [np.ndarray(sh, np.float64, buffer=m.buf)
                       for m, sh in zip(self.shms, self.shapes)]
        bounds = np.linspace(0, n, workers + 1).round().astype(int)
        self.chunks = [(int(a), int(b)) for a, b in zip(bounds[:-1],
                                                        bounds[1:]) if b > a]
        self.pool = ProcessPoolExecutor(workers, mp_context=mp.get_context(
            "spawn"))

    def _run(self, kind):
        names = [m.name for m in self.shms]
        tasks = [(kind, names, self.shapes, self.pgs[a:b], a, k, self.vg)
                 for k, (a, b) in enumerate(self.chunks)]
        list(self.pool.map(_small_task, tasks, timeout=3600))

    def fp(self, x):
        self.arrays[0][...] = x
        self._run("fp")
        return self.arrays[1].copy()

    def bp(self, p):
        self.arrays[1][...] = p
        self.arrays[2][...] = 0.0
        self._run("bp")
        return self.arrays[2][:len(self.chunks)].sum(axis=0)

    def close(self):
        self.pool.shutdown()
        del self.arrays
        for m in self.shms:
            m.close()
            m.unlink()


def _small_run(pool, vol, y, sirt_iters):
    """FP, BP, FDK (Ram-Lak) and ``sirt_iters`` SIRT iterations through the
    pool, restating the oracle's FDK and SIRT with the pool's FP/BP.
    Returns seconds per part and the SIRT result."""
    import numpy as np

    import oracle
    from oracle.algorithms import _safe_reciprocal

    t = {}
    tag = f"cpu {pool.workers} core(s)"
    t0 = time.perf_counter()
    pool.fp(vol)
    t["fp"] = time.perf_counter() - t0
    _log(f"{tag}: FP {t['fp']:.1f} s")
    t0 = time.perf_counter()
    pool.bp(y)
    t["bp"] = time.perf_counter() - t0
    _log(f"{tag}: BP {t['bp']:.1f} s")
    t0 = time.perf_counter()
    pool.bp(oracle.filter_rows(oracle.preweight(y, pool.pgs), "ramlak"))
    t["fdk"] = time.perf_counter() - t0
    _log(f"{tag}: FDK {t['fdk']:.1f} s")
    x = None
    if sirt_iters:
        t0 = time.perf_counter()
        x = np.zeros(vol.shape)
        C = _safe_reciprocal(pool.bp(np.ones_like(y)))
        _log(f"{tag}: SIRT C at {time.perf_counter() - t0:.1f} s")
        R = _safe_reciprocal(pool.fp(np.ones(vol.shape)))
        _log(f"{tag}: SIRT R at {time.perf_counter() - t0:.1f} s")
        for k in range(sirt_iters):
            x = x - C * pool.bp((pool.fp(x) - y) * R)
            _log(f"{tag}: SIRT iteration {k + 1} at "
                 f"{time.perf_counter() - t0:.1f} s")
        t["sirt"] = time.perf_counter() - t0
    return t, x


def gpu_small(sirt_iters=20):
    """The same small-workload jobs on the GPU (inputs resident, HIP events
    on the current stream): FP, BP, Ram-Lak FDK (``fdk``'s pre-weight and
    filter fused into the BP texture, then the BP) and ``sirt_iters`` SIRT
    iterations with their setup, in milliseconds."""
    import torch

    import kernelkit as kk
    from kernelkit import workloads
    from kernelkit.kernel import ProjectionTexture
    from kernelkit.processing import filter_into

    pgs, vg = workloads.config_geometry("small")
    x = torch.from_numpy(workloads.ellipsoid_phantom(vg.shape, n=10,
                                                     seed=0)).cuda()
    A = kk.XrayTransform(pgs, vg)
    y = A(x)
    v = torch.empty_like(x)
    stream = torch.cuda.current_stream()
    tex = ProjectionTexture(*y.shape, like=y)
    tex.zero()
    bptor = kk.BackProjector()
    bptor.projection_geometry = pgs
    bptor.volume_geometry = vg
    bptor.volume = v

    def fdk():
        filter_into(y, tex, geoms=pgs, filter="ramlak")
        bptor._textures = tex
        bptor()

    def sirt():
        kk.sirt(y, pgs, vg, iters=sirt_iters, verbose=False, return_gpu=True)

    out = {}
    for name, fn, reps in (("fp", lambda: A(x, out=y), 5),
                           ("bp", lambda: A.T(y, out=v), 5),
                           ("fdk", fdk, 5), (f"sirt_{sirt_iters}", sirt, 2)):
        fn()
        out[f"{name}_ms"] = round(_event_ms(fn, stream, reps=reps), 3)
    return out


def cpu_small(workers=16, sirt_iters=20, single_sirt_iters=20, gpu=True):
    """CPU baseline at the small workload (128^3, 180 angles, 160^2
    detector) in full: FP, BP, Ram-Lak FDK and ``sirt_iters`` SIRT
    iterations of the FP64 oracle on ``workers`` host processes, and the
    same on one process (SIRT there for ``single_sirt_iters`` iterations);
    SIRT times include their setup (the two normalisation passes).  With
    ``gpu``, the same jobs on the GPU beside them."""
    import numpy as np

    from kernelkit import workloads

    pgs, vg = workloads.config_geometry("small")
    vol = workloads.ellipsoid_phantom(vg.shape, n=10, seed=0).astype(
        np.float64)
    import oracle
    y = oracle.cone_fp(vol, pgs[:1], vg)  # warm the imports
    del y
    out = {"workload": "128^3 volume, 180 angles, 160^2 detector",
           "kind": "port"}
    workers = max(1, min(workers, os.cpu_count() or 1))
    for cores, iters in ((workers, sirt_iters), (1, single_sirt_iters)):
        pool = _SmallPool(pgs, vg, cores)
        try:
            list(pool.pool.map(_small_task, [("start",) + (None,) * 6] *
                               cores))   # start the workers
            y = pool.fp(vol)
            t, _ = _small_run(pool, vol, y, iters)
        finally:
            pool.close()
        out[f"{cores}_cores"] = {
            "fp_s": round(t["fp"], 2), "bp_s": round(t["bp"], 2),
            "fdk_s": round(t["fdk"], 2),
            f"sirt_{iters}_iterations_s": round(t["sirt"], 2)}
        _log(f"cpu small, {cores} cores: {out[f'{cores}_cores']}")
    if gpu:
        out["gpu_ms"] = gpu_small(sirt_iters)
    return out


def l2_read_rate(stream, mib=2, blocks=2048, iters=1024):
    """Aggregate L2 read rate (GB/s) from ``kk_probe_l2_read``: float4
    streams over a ``mib`` MiB buffer, which fits each XCD's 4 MiB L2
    but not a CU's L1."""
    import torch

    from kernelkit import _native
    buf = torch.ones(mib * (1 << 20) // 4, dtype=torch.float32,
                     device="cuda")
    out = torch.empty(blocks * 256, dtype=torch.float32, device="cuda")
    lib = _native.lib()

    def run():
        _native.check(lib.kk_probe_l2_read(
            _native._vp(buf.data_ptr()), buf.numel(), blocks, iters,
            _native._vp(out.data_ptr()), _native.stream_ptr(stream)))
    run()
    ms = min(_event_ms(run, stream) for _ in range(5))
    return blocks * 256 * 8 * iters * 16 / (ms * 1e-3) / 1e9


# (name, footprint_log2 in float2, lane step across, lane step down, iters):
# the projectors' tile (neighbouring rays read neighbouring texels: lane
# step 1 float2 across, one frame row of 512 + 3 texels down) and a fully
# coalesced tile (64 contiguous 16-byte loads), each L1-resident (16 KiB),
# L2-resident (2 MiB) and HBM-resident (1 GiB)
GATHER_PROBES = [
    ("tile_l1", 11, 1, 515, 1024), ("coalesced_l1", 11, 2, 16, 1024),
    ("tile_l2", 18, 1, 515, 256), ("coalesced_l2", 18, 2, 16, 256),
    ("tile_hbm", 27, 1, 515, 8), ("coalesced_hbm", 27, 2, 16, 8),
]


def gather_ceilings(stream, blocks=4096):
    """Read rates (GB/s) of ``kk_probe_gather``: 16-byte loads in 64-lane
    8 x 8 tiles at pseudo-random bases, the access shape of the projectors'
    texture gathers, over footprints resident in L1, L2 and HBM."""
    import torch

    from kernelkit import _native
    n2 = (1 << 27) + 1
    buf = torch.ones(2 * n2, dtype=torch.float32, device="cuda")
    out = torch.empty(blocks * 256, dtype=torch.float32, device="cuda")
    lib = _native.lib()
    rates = {}
    for name, k, sx, sy, iters in GATHER_PROBES:
        def run():
            _native.check(lib.kk_probe_gather(
                _native._vp(buf.data_ptr()), buf.numel(), k, sx, sy, blocks,
                iters, _native._vp(out.data_ptr()),
                _native.stream_ptr(stream)))
        run()
        ms = min(_event_ms(run, stream) for _ in range(3))
        rates[name] = round(blocks * 256 * 8 * iters * 16
                            / (ms * 1e-3) / 1e9, 1)
    del buf, out
    return rates


def _event_ms(fn, stream, reps=1):
    """Average milliseconds of ``fn`` between HIP events on ``stream``."""
    import torch
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    a.record(stream)
    for _ in range(reps):
        fn()
    b.record(stream)
    b.synchronize()
    return a.elapsed_time(b) / reps


def _batch_kernel_ms(st, stream, reps=3):
    """Milliseconds of each of a batched SIRT iteration's four launches
    (HIP events around each, eager), averaged over ``reps`` iterations;
    the state is reset afterwards."""
    import torch

    from kernelkit import _native
    lib = _native.lib()
    h = st.handle.ptr
    sp = _native.stream_ptr(stream)
    steps = {
        "fp": lambda: lib.kk_batch_fp(h, st.cfg_fp.grid, sp),
        "residual": lambda: lib.kk_batch_residual(h, sp),
        "bp": lambda: lib.kk_batch_bp(h, st.cfg_bp.grid, sp),
        "volume_textures": lambda: lib.kk_batch_volume_textures(h, sp),
    }
    ms = {k: 0.0 for k in steps}
    for _ in range(reps):
        for k, fn in steps.items():
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record(stream)
            _native.check(fn())
            b.record(stream)
            b.synchronize()
            ms[k] += a.elapsed_time(b) / reps
    st.reset()
    return {k: round(v, 3) for k, v in ms.items()}


def batched_sirt(n_problems=64, iters=30, seed=0, scale=1.0, classes=True):
    """SIRT on 64 uneven problems: the whole batch per iteration in four
    launches (forward projections, residuals, backprojections with the
    update, frame rebuilds), the loop captured in a HIP graph, against the
    per-problem path (``algo.sirt``'s per-iteration launches, problem after
    problem).  Setup (textures, weights, tuning, capture) is not timed.
    Also: the four launches' own times (the graph's overhead over them),
    and with ``classes`` the same comparison per volume-size class."""
    import numpy as np
    import torch

    from kernelkit import algo, workloads
    from kernelkit.batch import ConeBatch

    dev = torch.device("cuda", 0)
    problems = workloads.batch_problems(n_problems, seed=seed, scale=scale)
    B = ConeBatch(problems)
    vols = [torch.from_numpy(workloads.ellipsoid_phantom(
        vg.shape, n=10, seed=seed + k)).to(dev)
        for k, (_, vg) in enumerate(problems)]
    ys = B.fp(vols, tune=False)
    torch.cuda.synchronize()
    updates = sum(int(np.prod(vg.shape)) * len(pg) for pg, vg in problems)
    stream = torch.cuda.current_stream(dev)
    st = B.sirt_state(ys)
    st.configure(tune=True)
    st.capture(iters)
    torch.cuda.synchronize()
    t_graph = _event_ms(lambda: st._graph[1].replay(), stream)
    st.reset()
    t_eager = _event_ms(lambda: [st.iterate() for _ in range(iters)], stream)
    kernel_ms = _batch_kernel_ms(st, stream)
    plain = [algo._SirtState(y, pg, vg, None, None, None)
             for (pg, vg), y in zip(problems, ys)]
    torch.cuda.synchronize()

    def loop():
        for _ in range(iters):
            for s in plain:
                s.iterate()
    t_plain = _event_ms(loop, stream)
    # the batch's x were reset after the kernel timing: run the iterations
    # again (graph)
    st.run(iters, graph=True)
    torch.cuda.synchronize()
    diff = max(float((a - b).abs().max() / b.abs().max().clamp_min(1e-30))
               for a, b in zip(st.x, [s.x for s in plain]))
    del plain, st
    per_class = None
    if classes:
        per_class = {}
        bounds = (("64-128", 0, 128), ("129-192", 129, 192),
                  ("193-256", 193, 256))
        for name, lo, hi in bounds:
            idx = [k for k, (_, vg) in enumerate(problems)
                   if lo * scale <= max(vg.shape) <= hi * scale]
            if not idx:
                continue
            sub = _batch_vs_plain([problems[k] for k in idx],
                                  [ys[k] for k in idx], iters, stream)
            per_class[name] = dict(problems=len(idx), **sub)
    lo_v, hi_v = (max(8, int(round(s * scale))) for s in (64, 256))
    lo_d, hi_d = int(96 * scale), int(384 * scale)
    kernel_sum = sum(kernel_ms.values())
    return {
        "workload": (f"{n_problems} problems, {lo_v}^3-{hi_v}^3 (cubic and "
                     f"anisotropic), 90-360 angles, {lo_d}^2-{hi_d}^2 "
                     f"detectors, tilted/shifted cone_vec geometry; {iters} "
                     "SIRT iterations"),
        "updates_per_iteration": 2 * updates,
        "batched_graph_ms": round(t_graph, 2),
        "batched_eager_ms": round(t_eager, 2),
        "per_problem_ms": round(t_plain, 2),
        "speedup_vs_per_problem": round(t_plain / t_graph, 3),
        "batched_gups": round(2 * updates * iters / (t_graph * 1e-3) / 1e9,
                              1),
        "max_rel_diff_vs_per_problem": diff,
        "tiles": {"fp": st_cfg_repr(B, "fp"), "bp": st_cfg_repr(B, "bp")},
        "launches_per_iteration": 4,
        "kernel_ms_per_iteration": kernel_ms,
        "graph_ms_per_iteration": round(t_graph / iters, 3),
        "graph_overhead_vs_kernels": round(
            t_graph / iters / kernel_sum - 1.0, 4),
        "per_class": per_class,
    }


def _batch_vs_plain(problems, ys, iters, stream):
    """GUPS of ``iters`` batched SIRT iterations (graph) and of the
    per-problem loop on a subset of problems."""
    import numpy as np

    from kernelkit import algo
    from kernelkit.batch import ConeBatch

    B = ConeBatch(problems)
    st = B.sirt_state(ys)
    st.configure(tune=True)
    st.capture(iters)
    t_graph = _event_ms(lambda: st._graph[1].replay(), stream)
    del st
    plain = [algo._SirtState(y, pg, vg, None, None, None)
             for (pg, vg), y in zip(problems, ys)]

    def loop():
        for _ in range(iters):
            for s in plain:
                s.iterate()
    t_plain = _event_ms(loop, stream)
    del plain
    upd = 2 * iters * sum(int(np.prod(vg.shape)) * len(pg)
                          for pg, vg in problems)
    return {"batched_gups": round(upd / (t_graph * 1e-3) / 1e9, 1),
            "per_problem_gups": round(upd / (t_plain * 1e-3) / 1e9, 1),
            "speedup": round(t_plain / t_graph, 3),
            # the tiles the tuner picks for this class alone
            "tiles": {"fp": st_cfg_repr(B, "fp"), "bp": st_cfg_repr(B, "bp")}}


def st_cfg_repr(B, which):
    cfg = [v for k, v in B.tuner._cache.items() if "sirt" in k]
    if not cfg:
        return None
    c = cfg[0][0 if which == "fp" else 1]
    return {"tile": c.tile, "grid": c.grid}


def algorithms_full(pgs, vg, x, iters=50):
    """FDK (Hann) and 50 SIRT iterations of the full workload, end to end
    on the device (inputs resident; FDK includes pre-weight + filter fused
    into the backprojection texture; SIRT includes its setup: C, R and the
    textures)."""
    import torch

    import kernelkit as kk

    from kernelkit.kernel import ProjectionTexture
    from kernelkit.processing import filter_into

    dev = x.device
    A = kk.XrayTransform(pgs, vg)
    y = A(x)
    out = torch.empty_like(x)
    stream = torch.cuda.current_stream(dev)
    # FDK on the device, as algo.fdk minus its host copy of the result:
    # pre-weight + Hann row filter fused into the BP texture, then the BP
    tex = ProjectionTexture(*y.shape, device=dev)
    tex.zero()
    bptor = kk.BackProjector()
    bptor.projection_geometry = pgs
    bptor.volume_geometry = vg
    bptor.volume = out

    def fdk():
        filter_into(y, tex, geoms=pgs, filter="hann")
        bptor._textures = tex
        bptor()

    fdk()
    t_fdk = _event_ms(fdk, stream, reps=3)
    t_sirt = _event_ms(lambda: kk.sirt(y, pgs, vg, iters=iters,
                                       verbose=False, return_gpu=True),
                       stream)
    return {"fdk_hann_ms": round(t_fdk, 2),
            f"sirt_{iters}_iterations_ms": round(t_sirt, 1),
            "sirt_ms_per_iteration": round(t_sirt / iters, 2)}


def autograd_diff(ceiling=None):
    """Forward + backward of ||A x - y||^2 through the exact transpose
    against two plain calls (forward projection + backprojection) at 256^3,
    360 angles, 384^2 detector; also through the reference's default
    autograd operator (voxel-driven backprojector in the backward pass).

    The transpose moves the forward projection's samples backwards: each
    sample's four tap weights are gathered by the voxels they touch, 16 B
    per sample as in the forward gather, so its rate against the gather
    ``ceiling`` (GB/s) compares directly with the forward projection's."""
    import numpy as np
    import torch

    import kernelkit as kk
    from kernelkit import workloads
    from kernelkit.torch_support import XrayForwardProjection

    dev = torch.device("cuda", 0)
    pgs, vg = workloads.config_geometry("diff")
    x = torch.from_numpy(workloads.ellipsoid_phantom(vg.shape, n=10,
                                                     seed=0)).to(dev)
    A = kk.XrayTransform(pgs, vg)
    y = A(x)
    v = torch.empty_like(x)
    fpm = XrayForwardProjection(pgs, vg, matched=True)
    fpu = XrayForwardProjection(pgs, vg)   # the reference's default
    xx = x[None, None].clone().requires_grad_(True)
    yy = (y * 1.01)[None, None]
    stream = torch.cuda.current_stream(dev)

    def plain():
        A(x, out=y)
        A.T(y, out=v)

    def fwd_bwd(module):
        def run():
            xx.grad = None
            ((module(xx) - yy) ** 2).sum().backward()
        return run

    plain(), fwd_bwd(fpm)(), fwd_bwd(fpu)()
    t_plain = _event_ms(plain, stream, reps=3)
    t_ag = _event_ms(fwd_bwd(fpm), stream, reps=3)
    t_agu = _event_ms(fwd_bwd(fpu), stream, reps=3)
    t_t = _event_ms(lambda: A.matched_T(y, out=v), stream, reps=3)
    t_fp = _event_ms(lambda: A(x, out=y), stream, reps=3)
    updates = int(np.prod(vg.shape)) * len(pgs)
    # the general-geometry path of the transpose: the same sizes on an
    # orbit with seeded +-1 degree wobble and +-2 pixel detector drift
    c = dict(workloads.CONFIGS["diff"])
    c.pop("shape"), c.pop("n_angles")
    pgw = workloads.circular_geometry(len(pgs), wobble_deg=1.0, drift_px=2.0,
                                      seed=1, **c)
    Aw = kk.XrayTransform(pgw, vg)
    yw = Aw(x)
    t_fpw = _event_ms(lambda: Aw(x, out=yw), stream, reps=3)
    t_tw = _event_ms(lambda: Aw.matched_T(yw, out=v), stream, reps=3)
    res = {
        "workload": "256^3 volume, 360 angles, 384^2 detector",
        "plain_fp_plus_bp_ms": round(t_plain, 3),
        "autograd_fwd_bwd_ms": round(t_ag, 3),
        "ratio_vs_plain": round(t_ag / t_plain, 3),
        "unmatched_autograd_fwd_bwd_ms": round(t_agu, 3),
        "unmatched_ratio_vs_plain": round(t_agu / t_plain, 3),
        "transpose_ms": round(t_t, 3),
        "transpose_gups": round(updates / (t_t * 1e-3) / 1e9, 1),
        "fp_ms": round(t_fp, 3),
        "transpose_vs_fp": round(t_t / t_fp, 3),
        "wobble_fp_ms": round(t_fpw, 3),
        "wobble_transpose_ms": round(t_tw, 3),
        "wobble_transpose_vs_fp": round(t_tw / t_fpw, 3),
    }
    if ceiling:
        b = 16 * fp_samples(pgs, vg)
        bw = 16 * fp_samples(pgw, vg)
        res.update({
            "gather_ceiling_GBs": ceiling,
            "fp_frac": round(b / (t_fp * 1e-3) / 1e9 / ceiling, 4),
            "transpose_frac": round(b / (t_t * 1e-3) / 1e9 / ceiling, 4),
            "wobble_fp_frac": round(bw / (t_fpw * 1e-3) / 1e9 / ceiling, 4),
            "wobble_transpose_frac": round(
                bw / (t_tw * 1e-3) / 1e9 / ceiling, 4),
        })
    return res


def large_transpose(reps=2):
    """The exact transpose against the forward projection at the largest
    workload (1024^3 volume, 1440 angles with seeded +-1 degree wobble and
    +-3 pixel drift, 1536^2 detector): the general-geometry path of the
    transpose, HIP-event times on the launching stream after a warm-up
    call of each."""
    import numpy as np
    import torch

    import kernelkit as kk
    from kernelkit import workloads

    dev = torch.device("cuda", 0)
    pgs, vg = workloads.config_geometry("large")
    x = workloads.ellipsoid_phantom_device(vg.shape, n=10, seed=0, device=dev)
    A = kk.XrayTransform(pgs, vg)
    stream = torch.cuda.current_stream(dev)
    y = A(x)
    t_fp = _event_ms(lambda: A(x, out=y), stream, reps=reps)
    _log(f"large: forward projection {t_fp:.1f} ms")
    v = torch.empty_like(x)
    A.matched_T(y, out=v)
    torch.cuda.synchronize()
    _log("large: transpose warm-up done")
    t_t = _event_ms(lambda: A.matched_T(y, out=v), stream, reps=reps)
    _log(f"large: transpose {t_t:.1f} ms")
    updates = int(np.prod(vg.shape)) * len(pgs)
    return {
        "workload": ("1024^3 volume, 1440 angles (seeded +-1 deg wobble, "
                     "+-3 px drift), 1536^2 detector"),
        "fp_ms": round(t_fp, 1),
        "transpose_ms": round(t_t, 1),
        "transpose_vs_fp": round(t_t / t_fp, 3),
        "transpose_gups": round(updates / (t_t * 1e-3) / 1e9, 1),
    }


def large_unrolled(iters=5):
    """Differentiation through ``iters`` unrolled SIRT iterations at the
    largest workload (1024^3 volume, 1440 angles with seeded +-1 degree
    wobble and +-3 pixel drift, 1536^2 detector): ``x_k`` from
    ``torch_support.unrolled_reconstruction`` (exact transpose, SIRT
    weights, one learnable step size per iteration, checkpointed
    iterations), loss ``||x_iters - x_true||^2``, gradient to the step
    sizes.  Reports the forward and backward times, the operator calls
    they make, and the peak device memory held by tensors."""
    import numpy as np
    import torch

    import kernelkit as kk
    from kernelkit import workloads
    from kernelkit.torch_support import unrolled_reconstruction

    dev = torch.device("cuda", 0)
    pgs, vg = workloads.config_geometry("large")
    x_true = workloads.ellipsoid_phantom_device(vg.shape, n=10, seed=0,
                                                device=dev)
    A = kk.XrayTransform(pgs, vg)
    stream = torch.cuda.current_stream(dev)
    y = A(x_true)
    t_fp = _event_ms(lambda: A(x_true, out=y), stream)
    _log(f"large: forward projection {t_fp:.1f} ms")
    # distinct steps: with x0 = 0 the iteration maps (I - s_k W A^T R A)
    # commute, so equal steps would give equal gradients for every k
    steps = torch.linspace(1.0, 0.8, iters, dtype=torch.float64,
                           device=dev).requires_grad_(True)
    torch.cuda.synchronize()
    torch.cuda.reset_peak_memory_stats(dev)
    base = torch.cuda.memory_allocated(dev)
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    c = torch.cuda.Event(enable_timing=True)
    t_start = time.perf_counter()

    def progress(k, xk):
        torch.cuda.synchronize()
        _log(f"large: iteration {k} forward done "
             f"({time.perf_counter() - t_start:.1f} s)")

        def back(g):
            # the gradient of x_k is ready once the GPU has run the
            # backward pass of the iterations after k
            torch.cuda.synchronize()
            _log(f"large: backward reached iteration {k} "
                 f"({time.perf_counter() - t_start:.1f} s)")
        xk.register_hook(back)

    a.record(stream)
    x = unrolled_reconstruction(y, A, iters=iters, step=steps,
                                weighting="sirt", checkpoint=True,
                                callback=progress)
    loss = ((x - x_true) ** 2).sum()
    b.record(stream)
    _log("large: forward done")
    loss.backward()
    c.record(stream)
    c.synchronize()
    t_f, t_b = a.elapsed_time(b), b.elapsed_time(c)
    peak = torch.cuda.max_memory_allocated(dev)
    updates = int(np.prod(vg.shape)) * len(pgs)
    # forward: R = 1/A1, W = 1/A^T 1, then A and A^T per iteration;
    # backward: each iteration recomputed (A, A^T) and differentiated
    # (A^T of A's output gradient, A of A^T's)
    calls_f = 2 + 2 * iters
    calls_b = 4 * iters
    return {
        "workload": ("1024^3 volume, 1440 angles (seeded +-1 deg wobble, "
                     "+-3 px drift), 1536^2 detector; "
                     f"{iters} unrolled SIRT iterations, checkpointed"),
        "forward_projection_ms": round(t_fp, 1),
        "unrolled_forward_ms": round(t_f, 1),
        "backward_ms": round(t_b, 1),
        "operator_calls": {"forward": calls_f, "backward": calls_b},
        "gups": round(updates * (calls_f + calls_b) / ((t_f + t_b) * 1e-3)
                      / 1e9, 1),
        "peak_tensor_bytes": int(peak),
        "peak_tensor_gb": round(peak / 1e9, 2),
        "resident_before_gb": round(base / 1e9, 2),
        "step_grad": [float(g) for g in steps.grad.cpu()],
    }


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="full",
                    help="workload: full (default), small, diff")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--quick", action="store_true",
                    help="only the FP/BP step (no batched SIRT or autograd "
                         "comparison)")
    ap.add_argument("--extras", default="batched,autograd,algorithms",
                    help="comma-separated extra measurements after the "
                         "FP/BP step at the full config: batched, "
                         "autograd, algorithms, large, large_transpose, "
                         "cpu_small (the last three not in the default "
                         "list: a minute or more each)")
    ap.add_argument("--copy-calibration", type=int, default=0,
                    metavar="MIB",
                    help="after the timed steps, run one plain device copy "
                         "of MIB MiB (the reference kernel for scaling the "
                         "FETCH_SIZE/WRITE_SIZE counters)")
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print("bench.py measures one GPU: the cone FP/BP path is built and "
              "timed on a single MI355X (--gpus 1).", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    import kernelkit as kk
    from kernelkit import workloads

    if not torch.cuda.is_available():
        print("bench.py needs a HIP device", file=sys.stderr)
        return 1
    dev = torch.device("cuda", 0)
    pgs, vg = workloads.config_geometry(args.config)
    n, rows, cols = len(pgs), int(pgs[0].detector.rows), \
        int(pgs[0].detector.cols)
    shape = tuple(int(s) for s in vg.shape)
    _log(f"workload {args.config}: volume {shape}, {n} angles, "
         f"detector {rows}x{cols}")
    vol = workloads.ellipsoid_phantom(shape, n=10, seed=0)
    x = torch.from_numpy(vol).to(dev)
    y = torch.empty((n, rows, cols), dtype=torch.float32, device=dev)
    v = torch.empty(shape, dtype=torch.float32, device=dev)

    fptor = kk.ForwardProjector()
    fptor.projection_geometry = pgs
    fptor.volume_geometry = vg
    fptor.projections = y
    bptor = kk.BackProjector()
    bptor.projection_geometry = pgs
    bptor.volume_geometry = vg
    bptor.volume = v

    stream = torch.cuda.current_stream(dev)

    def ev():
        e = torch.cuda.Event(enable_timing=True)
        e.record(stream)
        return e

    def step():
        e0 = ev()
        fptor.volume = x          # volume -> FP frame textures
        e1 = ev()
        fptor()                   # forward projection kernel
        e2 = ev()
        bptor.projections = y     # projections -> BP texture
        e3 = ev()
        bptor()                   # backprojection kernel
        e4 = ev()
        return e0, e1, e2, e3, e4

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    _log(f"warmup {args.warmup} done")

    events = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        events.append(step())
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    _log(f"timed {args.steps} steps in {wall:.3f} s")
    if args.copy_calibration > 0:
        n_el = args.copy_calibration * (1 << 20) // 4
        src = torch.ones(n_el, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        torch.mul(src, 1.0, out=dst)   # one read + one write kernel
        torch.cuda.synchronize()
        _log(f"copy calibration: {args.copy_calibration} MiB")
        del src, dst

    def mean(i, j):
        return float(np.mean([e[i].elapsed_time(e[j]) for e in events]))

    t_ftex, t_fp, t_btex, t_bp = mean(0, 1), mean(1, 2), mean(2, 3), \
        mean(3, 4)
    l2_rate = l2_read_rate(stream)
    _log(f"L2 read probe: {l2_rate:.0f} GB/s")
    gathers = gather_ceilings(stream)
    probe_best = max(v for k, v in gathers.items() if k.endswith("_l1"))
    # the ceiling: the return-path rate, or the probes if they beat it
    ceiling = round(max(L1_RETURN_GBS, probe_best), 1)
    _log(f"gather probes (GB/s): {gathers}")
    ms_step = wall * 1e3 / args.steps
    updates = int(np.prod(shape)) * n
    value = 2 * updates / (ms_step * 1e-3) / 1e9
    samples = fp_samples(pgs, vg)
    fp_bytes = 16 * samples
    bp_bytes = 16 * updates
    achieved = fp_bytes / (t_fp * 1e-3) / 1e9
    traffic = issued = None
    if os.path.exists(TRAFFIC_FILE) and args.config == "full":
        with open(TRAFFIC_FILE) as f:
            traffic = json.load(f).get("cone_fp_kernel_bytes")
    if os.path.exists(GATHERS_FILE) and args.config == "full":
        with open(GATHERS_FILE) as f:
            issued = json.load(f).get("cone_fp_kernel_issued_bytes")
    compulsory_fp = 4 * (int(np.prod(shape)) + n * rows * cols)
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "GUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded 10-ellipsoid phantom, circular orbit",
        "config": {
            "workload": (f"cone FP + voxel-driven BP, volume "
                         f"{shape[0]}x{shape[1]}x{shape[2]}, {n} angles, "
                         f"detector {rows}x{cols}"),
            "name": args.config,
            "fp_gups": round(updates / (t_fp * 1e-3) / 1e9, 3),
            "bp_gups": round(updates / (t_bp * 1e-3) / 1e9, 3),
            "fp_kernel_ms": round(t_fp, 3),
            "bp_kernel_ms": round(t_bp, 3),
            "fp_texture_ms": round(t_ftex, 3),
            "bp_texture_ms": round(t_btex, 3),
        },
        "roofline": {
            "bound": "gather",
            "kernel": "cone_fp_kernel",
            "achieved": round(achieved, 1),
            # the gather ceiling: 64 B/clk/CU x 256 CUs x 2.4 GHz, which
            # the L1-resident 16-byte gather probes of this run reach
            # within a few percent ("gather_probes_GBs")
            "peak": ceiling,
            "peak_source": "64 B/clk/CU return path; L1-resident "
                           "kk_probe_gather reaches "
                           f"{probe_best / L1_RETURN_GBS:.3f} of it",
            "unit": "GB/s",
            "frac": round(achieved / ceiling, 4),
            "frac_of_best_probe": round(achieved / probe_best, 4),
            # the gathers the kernel issues (it skips samples outside its
            # wavefronts' ray supports, which the algorithm counts)
            "issued_bytes": issued,
            "issued_frac": (None if issued is None else round(
                issued / (t_fp * 1e-3) / 1e9 / ceiling, 4)),
            "issued_over_algorithmic": (
                None if issued is None else round(issued / fp_bytes, 4)),
            "traffic": traffic,
            "algorithmic_bytes": fp_bytes,
            "bytes_per_update": round(fp_bytes / updates, 4),
            "gather_probes_GBs": gathers,
            "hbm_compulsory_GBs": round(
                compulsory_fp / (t_fp * 1e-3) / 1e9, 1),
            "l2_stream_GBs": round(l2_rate, 1),
            "l2_theoretical_GBs": round(L2_PEAK_GBS, 1),
            "frac_of_l2_theoretical": round(achieved / L2_PEAK_GBS, 4),
            "hbm_peak": HBM_PEAK_GBS,
            # measured beyond-L2 (fabric/HBM) bytes of the launch at this
            # launch's time, against the HBM peak
            "traffic_GBs": (None if traffic is None else
                            round(traffic / (t_fp * 1e-3) / 1e9, 1)),
            "traffic_frac_of_hbm": (
                None if traffic is None else
                round(traffic / (t_fp * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)),
            "bp_kernel": {
                "achieved": round(bp_bytes / (t_bp * 1e-3) / 1e9, 1),
                "frac": round(bp_bytes / (t_bp * 1e-3) / 1e9 / ceiling,
                              4),
                "bytes_per_update": 16,
            },
        },
        "cpu_baseline": None,
    }
    extras = set() if args.quick or args.config != "full" else {
        e.strip() for e in args.extras.split(",") if e.strip()}
    unknown = extras - {"batched", "autograd", "algorithms", "large",
                        "large_transpose", "cpu_small"}
    if unknown:
        print(f"unknown extras: {sorted(unknown)}", file=sys.stderr)
        return 2
    if "batched" in extras:
        _log("batched SIRT vs per-problem ...")
        out["batched"] = batched_sirt()
        # the same generator at a quarter of the sizes: problems small
        # enough for per-problem launches to be latency-bound
        out["batched_small"] = batched_sirt(scale=0.25, classes=False)
    if "autograd" in extras:
        _log("autograd (exact transpose) vs plain ...")
        out["autograd"] = autograd_diff(ceiling)
    if "algorithms" in extras:
        _log("FDK and SIRT at full size ...")
        out["algorithms"] = algorithms_full(pgs, vg, x)
    if "large" in extras or "large_transpose" in extras:
        del fptor, bptor, y, v
        torch.cuda.empty_cache()
    if "large_transpose" in extras:
        _log("transpose vs forward projection at 1024^3 ...")
        out["large_transpose"] = large_transpose()
        torch.cuda.empty_cache()
    if "large" in extras:
        _log("unrolled differentiation at 1024^3 ...")
        out["large"] = large_unrolled()
    if not args.no_cpu_baseline:
        _log("cpu baseline (FP64 oracle, 1 core and 16 processes) ...")
        cb = out["cpu_baseline"] = cpu_baseline(pgs, vg, vol)
        if "algorithms" in out:
            # the full-size FDK and SIRT on the same host cores, from the
            # sampled per-angle FP and BP times (the filter is left out)
            alg = out["algorithms"]
            t_bp_cpu = n * cb["bp_s_per_angle"]
            t_it_cpu = n * (cb["fp_s_per_angle"] + cb["bp_s_per_angle"])
            alg["cpu_estimate"] = {
                "cores": cb["cores"],
                "fdk_s": round(t_bp_cpu, 1),
                "sirt_s_per_iteration": round(t_it_cpu, 1),
                "fdk_gpu_speedup": round(t_bp_cpu * 1e3 /
                                         alg["fdk_hann_ms"], 1),
                "sirt_gpu_speedup": round(
                    t_it_cpu * 1e3 / alg["sirt_ms_per_iteration"], 1),
                "basis": "cpu_baseline's per-angle FP/BP wall times x "
                         f"{n} angles"}
    if "cpu_small" in extras:
        _log("cpu baseline at the small workload, in full ...")
        out["cpu_small"] = cpu_small()
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
