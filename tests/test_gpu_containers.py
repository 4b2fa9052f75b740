"""Input and output containers of the projectors against the oracle: lists
of per-angle (or per-row) arrays, pitched rows (``kernelkit/data.py``), and
the reference's int64 array of entry pointers at the kernel level
(``kernels/cone_fp.py:173``, ``projector.py:301-327``).  A pitched
projection whose padding holds garbage backprojects like a clean one
(``tests/kernels/test_cone_bp.py:228-263``)."""
import numpy as np
import pytest

import oracle
from conftest import assert_parity

pytestmark = pytest.mark.gpu

kk = pytest.importorskip("kernelkit")
from kernelkit import workloads  # noqa: E402


def _torch():
    import torch
    return torch


def _dev(a):
    torch = _torch()
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()


def _host(t):
    return t.detach().cpu().numpy()


ROWS, COLS = 30, 37          # 37 floats per row: pitched to 40


@pytest.fixture(scope="module")
def case():
    pgs = workloads.circular_geometry(14, sod=100.0, odd=60.0, rows=ROWS,
                                      cols=COLS, pixel=1.3, tilt_deg=1.5,
                                      drift_px=1.0, seed=6)
    vg = kk.resolve_volume_geometry(shape=(28, 30, 22), voxel_size=1.0)
    vol = workloads.ellipsoid_phantom(vg.shape, n=5, seed=6)
    return pgs, vg, vol, oracle.cone_fp(vol, pgs, vg)


def _fptor(pgs, vg, vol, projection_axes=(0, 1, 2)):
    ptor = kk.ForwardProjector(projection_axes=projection_axes)
    ptor.projection_geometry = pgs
    ptor.volume_geometry = vg
    ptor.volume = _dev(vol)
    return ptor


def test_fp_into_list_of_angles(case, device):
    pgs, vg, vol, ref = case
    ptor = _fptor(pgs, vg, vol)
    outs = [_dev(np.full((ROWS, COLS), np.nan)) for _ in pgs]
    ptor.projections = outs
    ptor()                                   # overwrite
    got = np.stack([_host(p) for p in outs])
    assert_parity(got, ref, label="fp list")
    ptor(additive=True)                      # accumulate into the list
    got = np.stack([_host(p) for p in outs])
    assert_parity(got, 2 * ref, label="fp list additive")


def test_fp_into_pitched_list(case, device):
    torch = _torch()
    pgs, vg, vol, ref = case
    base = torch.full((len(pgs), ROWS, 40), 7.0, device="cuda")
    outs = [base[k, :, :COLS] for k in range(len(pgs))]
    assert all(kk.ispitched(o) for o in outs)
    ptor = _fptor(pgs, vg, vol)
    ptor.projections = outs
    ptor()
    assert_parity(_host(base[:, :, :COLS]), ref, label="fp pitched list")
    # the padding is not written
    assert float(base[:, :, COLS:].min()) == float(
        base[:, :, COLS:].max()) == 7.0


def test_fp_into_list_of_rows(case, device):
    """projection_axes (1, 0, 2): the list holds one (angle, col) sinogram
    per detector row, the table is indexed by row."""
    pgs, vg, vol, ref = case
    ptor = _fptor(pgs, vg, vol, projection_axes=(1, 0, 2))
    outs = [_dev(np.zeros((len(pgs), COLS))) for _ in range(ROWS)]
    ptor.projections = outs
    ptor()
    got = np.stack([_host(p) for p in outs])        # (row, angle, col)
    assert_parity(got.transpose(1, 0, 2), ref, label="fp list of rows")


def test_fp_kernel_pointer_array(case, device):
    """The reference's kernel-level form: an int64 device array of entry
    addresses plus the pitched logical dimensions."""
    torch = _torch()
    from kernelkit.kernels.cone_fp import RayDrivenConeFP
    pgs, vg, vol, ref = case
    K = RayDrivenConeFP()
    K.compile()
    K.set_params(K.geoms2params(pgs, vg), pgs, vg)
    tex = K.volume_texture(_dev(vol), vg)
    base = torch.zeros((len(pgs), ROWS, 40), device="cuda")
    ptrs = torch.tensor([base[k].data_ptr() for k in range(len(pgs))],
                        dtype=torch.int64, device="cuda")
    K(tex, pgs, vg, ptrs, projections_pitch=(len(pgs), ROWS, 40))
    assert_parity(_host(base[:, :, :COLS]), ref, label="fp pointer array")
    assert float(base[:, :, COLS:].abs().max()) == 0.0
    # accumulate (the kernel's default, as the reference's +=)
    K(tex, pgs, vg, ptrs, projections_pitch=(len(pgs), ROWS, 40))
    assert_parity(_host(base[:, :, :COLS]), 2 * ref, label="fp ptrs +=")


def test_fp_list_call_allocates_once(case, device):
    """Repeated calls with the same list reuse the device pointer table."""
    pgs, vg, vol, _ = case
    ptor = _fptor(pgs, vg, vol)
    outs = [_dev(np.zeros((ROWS, COLS))) for _ in pgs]
    ptor.projections = outs
    ptor()
    table = ptor.kernel._ptr_cache[1]
    ptor()
    assert ptor.kernel._ptr_cache[1] is table


def test_bp_from_list_and_pitched(case, device):
    torch = _torch()
    pgs, vg, _, _ = case
    rng = np.random.default_rng(2)
    proj = rng.uniform(size=(len(pgs), ROWS, COLS))
    ref = oracle.cone_bp(proj, pgs, vg)
    btor = kk.BackProjector()
    btor.projection_geometry = pgs
    btor.volume_geometry = vg
    btor.volume = torch.zeros(vg.shape, device="cuda")
    btor.projections = [_dev(p) for p in proj]
    btor()
    assert_parity(_host(btor.volume), ref, label="bp list")
    pitched = kk.aspitched(_dev(proj))
    assert kk.ispitched(pitched) and pitched.stride(1) == 40
    pitched._base[..., COLS:] = 1e6            # garbage in the padding
    btor.projections = pitched
    btor()
    assert_parity(_host(btor.volume), ref, label="bp pitched")


def test_sirt_from_pitched_measurements(case, device):
    torch = _torch()
    pgs, vg, _, ref = case
    y = _dev(ref)
    yp = kk.aspitched(y)
    assert yp.stride(1) == 40
    a = kk.sirt(y, pgs, vg, iters=5, verbose=False, return_gpu=True)
    b = kk.sirt(yp, pgs, vg, iters=5, verbose=False, return_gpu=True)
    assert torch.equal(a, b)
    like = kk.pitched_like(yp, fill=0.0)
    assert kk.ispitched(like) and tuple(like.shape) == tuple(yp.shape)
    assert like.is_cuda and float(like.abs().max()) == 0.0
