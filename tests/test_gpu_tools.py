"""The grid-tool kernels (random ids, total variation, resampling) through the C
ABI against the oracle, and their Python entry points against PyTorch FP32."""
import numpy as np
import pytest
import torch

import rendervous_amd as rv
from rendervous_amd import _lib as L
from test_oracle_tools import _random_ids, _resample, _tv, _tv_bw

pytestmark = pytest.mark.gpu


def _stream():
    return L.stream_handle(torch.device("cuda", 0))


def test_random_ids_bit_exact(gpu):
    seeds = (1234, 5678, 91011, 121314)
    shape = (32, 512, 512)
    n = 100003
    out = torch.empty(n, 3, dtype=torch.int64, device="cuda")
    L.check(L.load().vr_random_ids(out.data_ptr(), n, L.i64s(shape), 3, L.u32s(seeds), _stream()), "vr_random_ids")
    torch.cuda.synchronize()
    np.testing.assert_array_equal(out.cpu().numpy(), _random_ids(n, shape, seeds))


def test_random_sensors_follow_the_seed(gpu):
    cam = rv.PerspectiveCameraSensor(40, 30, rv.functions.orbit_poses(6).cuda())
    rv.seed(5)
    a = cam.random_sensors(4096)
    rv.seed(5)
    b = cam.random_sensors(4096)
    assert a.dtype == torch.int64 and a.shape == (4096, 3) and torch.equal(a, b)
    mx = a.max(0).values.cpu().tolist()
    assert a.min() >= 0 and mx == [5, 29, 39]
    sigma = torch.rand(8, 8, 8, 1, device="cuda")
    bmin, bmax = rv.normalized_box(sigma)
    field = rv.GridDDATransmittance(rv.Grid3D(sigma, bmin, bmax), rv.RayBoxIntersection(bmin, bmax))
    with torch.no_grad():
        img = cam.capture(field, a)
        full = cam.capture(field)
    assert img.shape == (4096, 1)
    # a sensor batch renders the same rays as the full capture at those sensors
    idx = (a[:, 0] * 30 + a[:, 1]) * 40 + a[:, 2]
    torch.testing.assert_close(img, full.reshape(-1, 1)[idx])


@pytest.mark.parametrize("shape", [(4097, 1), (33, 300, 3), (17, 40, 300, 1), (9, 10, 11, 4), (5, 6, 7, 6),
                                   (16, 24, 64, 1), (3, 9, 4, 1), (2, 2, 1028, 1), (7, 5, 3, 1), (40, 33, 36, 1)])
def test_total_variation_bit_exact(gpu, shape):
    """Generic kernels and the single-channel 3-D x4 kernels (rows of whole float4s)."""
    rng = np.random.default_rng(len(shape) + shape[-1])
    x = rng.normal(size=shape).astype(np.float32)
    x[..., :1][x[..., :1] > 1.5] = 0.0  # some ties
    og = rng.normal(size=shape[:-1] + (1,)).astype(np.float32)
    xd, ogd = torch.from_numpy(x).cuda(), torch.from_numpy(og).cuda()
    out = torch.empty(shape[:-1] + (1,), device="cuda")
    grad = torch.full(shape, 9.0, device="cuda")
    lib = L.load()
    sh = L.i64s(shape)
    L.check(lib.vr_total_variation_fw(xd.data_ptr(), sh, len(shape) - 1, out.data_ptr(), _stream()), "tv_fw")
    L.check(lib.vr_total_variation_bw(xd.data_ptr(), ogd.data_ptr(), sh, len(shape) - 1, grad.data_ptr(), _stream()),
            "tv_bw")
    torch.cuda.synchronize()
    np.testing.assert_array_equal(out.cpu().numpy(), _tv(x))
    np.testing.assert_array_equal(grad.cpu().numpy(), _tv_bw(x, og))


def test_total_variation_autograd_matches_torch(gpu):
    x = torch.randn(24, 20, 28, 3, device="cuda", requires_grad=True)
    w = torch.randn(24, 20, 28, 1, device="cuda")
    (rv.functions.total_variation(x) * w).sum().backward()
    xr = x.detach().clone().requires_grad_(True)
    tv = torch.zeros(24, 20, 28, device="cuda")
    for ax in range(3):
        d = torch.diff(xr, dim=ax).abs().sum(-1)
        pa = [0] * 6
        pb = [0] * 6
        pa[2 * (2 - ax) + 1] = 1
        pb[2 * (2 - ax)] = 1
        tv = tv + torch.nn.functional.pad(d, pa) + torch.nn.functional.pad(d, pb)
    (tv[..., None] * w).sum().backward()
    torch.testing.assert_close(x.grad, xr.grad, rtol=1e-5, atol=1e-5)


@pytest.mark.parametrize("src,dst,C", [((300,), (599,), 1), ((33, 40), (65, 21), 3), ((16, 20, 24), (31, 39, 47), 1),
                                       ((64, 64, 64), (32, 33, 34), 2), ((16, 20, 24), (31, 39, 48), 1),
                                       ((8, 8, 8), (16, 16, 16), 1), ((40, 30, 20), (20, 15, 12), 1)])
def test_resample_step_matches_oracle(gpu, src, dst, C):
    x = np.random.default_rng(C + len(src)).normal(size=src + (C,)).astype(np.float32)
    xd = torch.from_numpy(x).cuda()
    out = torch.empty(dst + (C,), device="cuda")
    L.check(L.load().vr_resample_grid(xd.data_ptr(), L.i64s(src), L.i64s(dst), len(src), C, out.data_ptr(), _stream()),
            "vr_resample_grid")
    torch.cuda.synchronize()
    np.testing.assert_allclose(out.cpu().numpy(), _resample(x, dst), rtol=1e-6, atol=1e-6)


def test_resample_grid_chain_matches_torch(gpu):
    """8^3 -> 64^3 goes through 16^3 and 32^3 (at most 2x per step); each step is
    torch's align_corners=True trilinear interpolation."""
    x = torch.randn(8, 8, 8, 1, device="cuda")
    y = rv.functions.resample_grid(x, (64, 64, 64))
    ref = x.permute(3, 0, 1, 2)[None]
    for n in (16, 32, 64):
        ref = torch.nn.functional.interpolate(ref, size=(n, n, n), mode="trilinear", align_corners=True)
    torch.testing.assert_close(y, ref[0].permute(1, 2, 3, 0), rtol=1e-5, atol=1e-5)
    down = rv.functions.resample_grid(y, (10, 12, 14))
    assert down.shape == (10, 12, 14, 1)


@pytest.mark.parametrize("cell", [False, True])
@pytest.mark.parametrize("src,dst", [((32, 32, 32), (64, 64, 64)), ((20, 30, 40), (39, 61, 80)),
                                     ((64, 64, 64), (32, 32, 32)), ((40, 30, 600), (20, 15, 300)),
                                     ((3, 70, 9), (5, 141, 16)), ((16, 64, 1000), (16, 128, 2000))])
def test_lds_resample_is_bit_identical(gpu, src, dst, cell):
    """The LDS-staged resampler (a band of 8 output rows per workgroup, source rows
    copied to LDS once) equals the per-lane kernel (VR_RESAMPLE_LDS=0) bit for bit,
    for up- and downsampling, vertex-aligned (grid) and pixel-centred (img)."""
    import os
    print(f"[case] lds resample {src}->{dst} cell={cell}", flush=True)
    x = torch.from_numpy(np.random.default_rng(len(src)).normal(size=src + (1,)).astype(np.float32)).cuda()
    fn = L.load().vr_resample_img if cell else L.load().vr_resample_grid
    outs = []
    old = os.environ.get("VR_RESAMPLE_LDS")
    try:
        for lds in ("1", "0"):
            os.environ["VR_RESAMPLE_LDS"] = lds
            out = torch.full(dst + (1,), float("nan"), device="cuda")
            L.check(fn(x.data_ptr(), L.i64s(src), L.i64s(dst), 3, 1, out.data_ptr(), _stream()), "resample")
            torch.cuda.synchronize()
            outs.append(out)
    finally:
        if old is None:
            os.environ.pop("VR_RESAMPLE_LDS", None)
        else:
            os.environ["VR_RESAMPLE_LDS"] = old
    assert torch.isfinite(outs[0]).all()
    assert torch.equal(outs[0], outs[1])
