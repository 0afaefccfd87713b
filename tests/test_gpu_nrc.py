"""The neural radiance cache kernels (BF16 MFMA MLP, training, cached render)
against plain PyTorch FP32 references of the same operations, and the cached
render against the oracle's truncated NEE integrator."""
import math

import numpy as np
import pytest
import torch

import rendervous_amd as rv
from rendervous_amd import _lib as L
from rendervous_amd.nrc import LAYERS, NeuralRadianceCache
from helpers import Built, Camera, Scene, fbm_grid, match_fraction, oracle_capture_fw, pose

pytestmark = pytest.mark.gpu

BOX = ((-0.5, -0.5, -0.5), (0.5, 0.5, 0.5))


def _encode_ref(x, w, bmin, bmax):
    # u in FP32 as the kernels form it (x - bmin) * (1 / extent); the phases pi 2^f u in FP64
    ext = torch.tensor(bmax, dtype=torch.float32) - torch.tensor(bmin, dtype=torch.float32)
    u = ((x.float() - torch.tensor(bmin, dtype=torch.float32)) * (1.0 / ext)).double()
    cols = []
    for c in range(3):
        for f in range(12):
            a = math.pi * (2.0 ** f) * u[:, c]
            cols += [torch.sin(a), torch.cos(a)]
    n = x.shape[0]
    feats = torch.stack(cols, 1)
    return torch.cat([feats, w.double(), torch.ones(n, 1, dtype=torch.float64), torch.zeros(n, 20, dtype=torch.float64)], 1)


def _split(flat):
    out, o = [], 0
    for n, k in LAYERS:
        out.append(flat[o:o + n * k].reshape(n, k))
        o += n * k
    return out


def _mlp_ref(weights_bf16, feats, round_acts=True):
    """PyTorch FP32 forward with the kernel's BF16 weights; round_acts rounds the
    input and each hidden activation to BF16 as the kernels store them."""
    Ws = [w.float() for w in _split(weights_bf16.cpu())]
    a = feats.float()
    if round_acts:
        a = a.bfloat16().float()
    for W in Ws[:-1]:
        a = torch.relu(a @ W.t())
        if round_acts:
            a = a.bfloat16().float()
    return (a @ Ws[-1].t())[:, :3]


def _random_weights(seed, scale=1.0):
    c = NeuralRadianceCache(torch.device("cuda", 0), seed=seed)
    if scale != 1.0:
        c.params.mul_(scale)
        c.weights = c.params.to(torch.bfloat16)
    return c


def _box_ptrs():
    return L.f3(BOX[0]), L.f3(BOX[1])


def test_encoding_matches_reference(gpu):
    rng = np.random.default_rng(1)
    x = torch.from_numpy(rng.uniform(-0.5, 0.5, size=(500, 3)).astype(np.float32))
    w = torch.nn.functional.normalize(torch.from_numpy(rng.normal(size=(500, 3)).astype(np.float32)), dim=1)
    out = torch.zeros(500, L.NRC_FEATURES, device="cuda")
    xd, wd = x.cuda(), w.cuda()  # kept alive until the kernel has run
    L.check(L.load().vr_nrc_encode(xd.data_ptr(), wd.data_ptr(), 500, *_box_ptrs(), out.data_ptr(), None),
            "vr_nrc_encode")
    torch.cuda.synchronize()
    np.testing.assert_allclose(out.cpu().double().numpy(), _encode_ref(x, w, *BOX).numpy(), atol=2e-5)


@pytest.mark.parametrize("n", [1, 31, 32, 1000])
def test_mlp_forward_matches_torch(gpu, n):
    cache = _random_weights(2)
    feats = torch.from_numpy(np.random.default_rng(n).uniform(-1, 1, size=(n, 96)).astype(np.float32))
    feats[:, 76:] = 0.0
    out = torch.zeros(n, 3, device="cuda")
    xf = feats.cuda()
    L.check(L.load().vr_nrc_mlp_fw(xf.data_ptr(), n, cache.weights.data_ptr(), out.data_ptr(), None), "vr_nrc_mlp_fw")
    torch.cuda.synchronize()
    got = out.cpu()
    ref = _mlp_ref(cache.weights, feats)
    scale = float(ref.abs().max()) + 1e-6
    assert float((got - ref).abs().max()) <= 1e-2 * scale, (got[:3], ref[:3])
    # against the unrounded FP32 network: the BF16 activations cost ~1 %
    ref32 = _mlp_ref(cache.weights, feats, round_acts=False)
    assert float((got - ref32).norm() / ref32.norm()) <= 0.05


def _records(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-0.5, 0.5, size=(n, 3))
    w = rng.normal(size=(n, 3))
    w /= np.linalg.norm(w, axis=1, keepdims=True)
    t = np.stack([0.5 + 0.4 * np.sin(3 * x[:, 0]), 0.3 + 0.2 * x[:, 1] * w[:, 2], 0.2 * np.ones(n)], 1)
    return torch.from_numpy(np.concatenate([x, w, t], 1).astype(np.float32))


def test_training_gradient_matches_torch_autograd(gpu):
    """lr = 0: the kernel's gradient of the relative L2 loss equals torch's autograd
    of the same BF16-rounded network within the BF16 rounding of dZ."""
    cache = _random_weights(3)
    rec = _records(1000, 4)
    recd = rec.cuda()
    count = torch.tensor([1000], dtype=torch.int32, device="cuda")
    lib = L.load()
    L.check(lib.vr_nrc_train_step(recd.data_ptr(), count.data_ptr(), 0, 1024, *_box_ptrs(), cache.params.data_ptr(),
                                  cache.weights.data_ptr(), cache.grad_copies.data_ptr(), cache.m.data_ptr(), cache.v.data_ptr(),
                                  1, 0.0, 0.9, 0.99, 1e-8, cache.loss.data_ptr(), None), "vr_nrc_train_step")
    torch.cuda.synchronize()
    Ws = [W.float().clone().requires_grad_(True) for W in _split(cache.weights.cpu())]
    a = _encode_ref(rec[:, :3], rec[:, 3:6], *BOX).float().bfloat16().float()
    for W in Ws[:-1]:
        a = torch.relu(a @ W.t()).bfloat16().float()
    q = (a @ Ws[-1].t())[:, :3]
    loss = ((q - rec[:, 6:]) ** 2 / (q.detach() ** 2 + 0.01)).mean()
    loss.backward()
    assert abs(float(cache.loss.item()) - float(loss)) <= 2e-2 * float(loss)
    got = _split(cache.grad.cpu())
    for i, (W, g) in enumerate(zip(Ws, got)):
        ref = W.grad
        err = float((g - ref).norm() / (ref.norm() + 1e-12))
        assert err <= 0.05, (i, err)


def test_adam_step_matches_torch_optim(gpu):
    cache = _random_weights(5)
    rec = _records(256, 6).cuda()
    count = torch.tensor([256], dtype=torch.int32, device="cuda")
    lib = L.load()
    p0 = cache.params.clone()
    L.check(lib.vr_nrc_train_step(rec.data_ptr(), count.data_ptr(), 0, 256, *_box_ptrs(), cache.params.data_ptr(),
                                  cache.weights.data_ptr(), cache.grad_copies.data_ptr(), cache.m.data_ptr(), cache.v.data_ptr(),
                                  1, 1e-2, 0.9, 0.99, 1e-8, cache.loss.data_ptr(), None), "vr_nrc_train_step")
    torch.cuda.synchronize()
    p = torch.nn.Parameter(p0.cpu())
    opt = torch.optim.Adam([p], lr=1e-2, betas=(0.9, 0.99), eps=1e-8)
    p.grad = cache.grad.cpu()
    opt.step()
    np.testing.assert_allclose(cache.params.cpu().numpy(), p.detach().numpy(), rtol=1e-5, atol=1e-6)
    assert torch.equal(cache.weights.cpu(), cache.params.cpu().bfloat16())


@pytest.mark.parametrize("tkind", [L.VR_T_DT, L.VR_T_RT])
def test_zero_cache_render_is_the_truncated_nee_path(gpu, tkind):
    """With Q = 0 the cached render adds nothing at the (K+1)-th collision, so it
    equals the NEE integrator capped at K scatter events, sample by sample."""
    K = 2
    sc = Scene(sigma=fbm_grid(24, seed=9, scale=20.0), albedo_c=(0.9, 0.8, 0.7), phase_kind=L.VR_PHASE_HG,
               phase_g=0.6, env_c=(0.5, 0.6, 0.7), sun_dir=(0.3, 0.9, 0.3), sun_radiance=(3.0, 3.0, 3.0), sun_prob=0.5,
               majorant="grid", transmittance_kind=tkind, max_bounces=K)
    cam_args = dict(poses=pose((0.0, 0.3, 2.5))[None], width=40, height=40, jittered=True)
    seeds = (12345, 67890, 13579, 24680)
    ref = oracle_capture_fw(sc, L.VR_MAP_NEE_DT, cam_args, seeds, 2)
    b = Built(sc, "cuda")
    cam = Camera(device="cuda", **cam_args)
    out = torch.zeros(1600, 3, device="cuda")
    q = torch.zeros(3200, 10, device="cuda")
    qc = torch.zeros(1, dtype=torch.int32, device="cuda")
    wz = torch.zeros(L.NRC_PARAMS, dtype=torch.bfloat16, device="cuda")
    lau = L.launch(seeds, samples=2)
    L.check(L.load().vr_nrc_render(b.medium, cam.c, lau, K, wz.data_ptr(), out.data_ptr(), q.data_ptr(), 3200,
                                   qc.data_ptr(), None), "vr_nrc_render")
    torch.cuda.synchronize()
    assert 0 < int(qc.item()) <= 3200
    assert match_fraction(out.cpu().numpy(), ref, rtol=1e-3, atol=1e-4) >= 0.98


def test_fused_cache_inference_equals_encoding_then_mlp(gpu):
    """The render's fused inference (features formed per lane in registers, the
    MLP tile, the weighted add per pixel) equals the explicit encoder followed
    by the MLP kernel on the same queries: the cached render minus the Q = 0
    render of the same samples is sum W relu(Q(x, w)) per pixel (same BF16
    inputs and MLP; only the order of the float adds differs)."""
    K = 2
    sc = Scene(sigma=fbm_grid(24, seed=9, scale=20.0), albedo_c=(0.9, 0.8, 0.7), phase_kind=L.VR_PHASE_HG,
               phase_g=0.6, env_c=(0.5, 0.6, 0.7), sun_dir=(0.3, 0.9, 0.3), sun_radiance=(3.0, 3.0, 3.0), sun_prob=0.5,
               majorant="grid", max_bounces=8)
    b = Built(sc, "cuda")
    cam = Camera(device="cuda", poses=pose((0.0, 0.3, 2.5))[None], width=40, height=40, jittered=True)
    wz = torch.zeros(L.NRC_PARAMS, dtype=torch.bfloat16, device="cuda")
    wr = _random_weights(4).weights
    res = []
    for w in (wz, wr):
        out = torch.zeros(1600, 3, device="cuda")
        q = torch.zeros(3200, 10, device="cuda")
        qc = torch.zeros(1, dtype=torch.int32, device="cuda")
        L.check(L.load().vr_nrc_render(b.medium, cam.c, L.launch((7, 1, 3, 5), samples=2), K, w.data_ptr(),
                                       out.data_ptr(), q.data_ptr(), 3200, qc.data_ptr(), None), "vr_nrc_render")
        torch.cuda.synchronize()
        res.append((out, q[:int(qc.item())].clone()))
    (o0, _), (o1, q) = res
    n = q.shape[0]
    assert n > 200
    x = q[:, 0:3].contiguous()
    w = q[:, 3:6].contiguous()
    feats = torch.empty(n, 96, device="cuda")
    L.check(L.load().vr_nrc_encode(x.data_ptr(), w.data_ptr(), n, *_box_ptrs(), feats.data_ptr(), None), "encode")
    y = torch.empty(n, 3, device="cuda")
    L.check(L.load().vr_nrc_mlp_fw(feats.data_ptr(), n, wr.data_ptr(), y.data_ptr(), None), "mlp_fw")
    torch.cuda.synchronize()
    pix = q[:, 9].cpu().numpy().view(np.uint32).astype(np.int64)
    add = (q[:, 6:9] * torch.relu(y)).cpu().double().numpy()
    expect = o0.cpu().double().numpy().copy()
    np.add.at(expect, pix, add)
    np.testing.assert_allclose(o1.cpu().double().numpy(), expect, rtol=1e-5, atol=1e-7)
    assert np.abs(add).max() > 1e-4  # the cache term is far above the tolerance


@pytest.mark.parametrize("tkind,cells", [(L.VR_T_DT, True), (L.VR_T_DT, False), (L.VR_T_RT, False)])
def test_persistent_cached_render_equals_the_per_thread_one(gpu, tkind, cells):
    """The cached render as the persistent lane state machine (the default for
    constant environments) traces the same samples as the per-thread kernel
    (VR_NRC_PERSISTENT=0): at one sample per pixel the pre-cache images agree to
    float rounding (the two kernels contract a few products into FMAs
    differently) and so do the queries, matched by their pixel."""
    import os
    K = 2
    sc = Scene(sigma=fbm_grid(24, seed=9, scale=20.0), albedo_c=(0.9, 0.8, 0.7), phase_kind=L.VR_PHASE_HG,
               phase_g=0.6, env_c=(0.5, 0.6, 0.7), sun_dir=(0.3, 0.9, 0.3), sun_radiance=(3.0, 3.0, 3.0), sun_prob=0.5,
               majorant="grid", transmittance_kind=tkind, max_bounces=8)
    b = Built(sc, "cuda")
    if cells:
        table = torch.empty(23 ** 3 * 8, device="cuda")
        L.check(L.load().vr_grid_pack_cells(b.medium.sigma, table.data_ptr(), None), "pack")
        b.medium.sigma_cells = table.data_ptr()
    cam = Camera(device="cuda", poses=pose((0.0, 0.3, 2.5))[None], width=40, height=40, jittered=True)
    wz = torch.zeros(L.NRC_PARAMS, dtype=torch.bfloat16, device="cuda")
    res = []
    old = os.environ.get("VR_NRC_PERSISTENT")
    try:
        for flag in ("0", "1"):
            os.environ["VR_NRC_PERSISTENT"] = flag
            out = torch.zeros(1600, 3, device="cuda")
            q = torch.zeros(1600, 10, device="cuda")
            qc = torch.zeros(1, dtype=torch.int32, device="cuda")
            L.check(L.load().vr_nrc_render(b.medium, cam.c, L.launch((3, 1, 4, 1), samples=1), K, wz.data_ptr(),
                                           out.data_ptr(), q.data_ptr(), 1600, qc.data_ptr(), None), "vr_nrc_render")
            torch.cuda.synchronize()
            n = int(qc.item())
            rows = q[:n].cpu().numpy()
            pix = rows[:, 9].copy().view(np.uint32)
            assert len(np.unique(pix)) == n  # one query per pixel at most
            res.append((out.cpu().numpy(), rows[np.argsort(pix)], np.sort(pix)))
    finally:
        if old is None:
            os.environ.pop("VR_NRC_PERSISTENT", None)
        else:
            os.environ["VR_NRC_PERSISTENT"] = old
    (o0, q0, p0), (o1, q1, p1) = res
    assert len(p0) > 100 and abs(len(p0) - len(p1)) <= len(p0) // 200
    assert match_fraction(o1, o0, rtol=1e-5, atol=1e-6) >= 0.995
    both, i0, i1 = np.intersect1d(p0, p1, return_indices=True)
    assert len(both) >= 0.995 * len(p0)
    close = np.all(np.abs(q0[i0, :9] - q1[i1, :9]) <= 1e-5 * np.abs(q0[i0, :9]) + 1e-6, axis=1)
    assert close.mean() >= 0.995, close.mean()


def _nrc_scene():
    """48^3 fBm cloud, sun + environment, HG g = 0.5, NEE with DT shadow rays; 64x64 camera."""
    dev = torch.device("cuda", 0)
    sigma = rv.functions.fbm_value_noise(48, seed=3, scale=20.0, device=dev)
    ds = rv.objects.DependencySet()
    ds.add_parameters(sigma_tensor=sigma, scattering_albedo_tensor=torch.tensor([0.9, 0.9, 0.9]),
                      environment_tensor=torch.tensor([0.3, 0.3, 0.3]), phase_g_tensor=torch.tensor([0.5]))
    ds.requires(rv.objects.medium_majorant_coarse_grid, resolution=8)
    ds.requires(rv.objects.medium_environment_sampler_sun, direction=(0.4, 0.8, 0.45), radiance=(2.0, 2.0, 2.0),
                probability=0.5)
    ds.requires(rv.objects.medium_phase_HG, mixture=False)
    ds.requires(rv.objects.medium_phase_sampler_HG, mixture=False)
    ds.requires(rv.objects.medium_transmittance_DT)
    ds.requires(rv.objects.medium_radiance_path_integrator_NEE_DT, max_bounces=16)
    field = ds.radiance
    camera = rv.PerspectiveCameraSensor(64, 64, rv.functions.look_at_pose((0.0, 0.3, 2.5))[None].to(dev), jittered=True)
    return field, camera


def test_cache_learns_and_the_cached_render_approaches_the_plain_one(gpu):
    """Online training on the scene's own NEE paths lowers the loss, and at equal
    spp the cached render (paths ended after 2 bounces) is at least as close to
    a high-spp plain render as the plain render itself."""
    dev = torch.device("cuda", 0)
    field, camera = _nrc_scene()
    cache = NeuralRadianceCache(dev, seed=7, lr=1e-2)
    losses = []
    for it in range(300):
        rv.seed(1000 + it)
        losses.append(float(cache.train(field, camera, n_paths=4096, max_vertices=16).item()))
    print(f"[nrc] loss {losses[0]:.4f} -> {np.mean(losses[-20:]):.4f}", flush=True)
    assert np.mean(losses[-20:]) < 0.5 * losses[0]
    with torch.no_grad():
        rv.seed(1)
        ref = camera.capture(field, fw_samples=1024)
        rv.seed(2)
        plain = camera.capture(field, fw_samples=16)
        rv.seed(3)
        cached = cache.render(field, camera, spp=16)
    e_plain = float(((plain - ref) ** 2).mean().sqrt())
    e_cached = float(((cached - ref) ** 2).mean().sqrt())
    print(f"[nrc] rmse plain {e_plain:.5f} cached {e_cached:.5f} (mean {float(ref.mean()):.4f})", flush=True)
    assert e_cached <= 1.2 * e_plain


def test_paths_of_several_traces_in_one_launch_equal_separate_launches(gpu):
    """vr_nrc_paths_traces over 3 traces writes, per trace, the records that
    vr_nrc_paths writes with that trace's seeds (rows compared as sorted sets:
    the append order follows the atomics)."""
    import rendervous_amd as rv
    field, camera = _nrc_scene()
    lw = NeuralRadianceCache._lower(field)
    cam, keep = camera._camera(None)
    lib = L.load()
    n_paths, maxv, traces = 2048, 8, 3
    cap = n_paths * maxv
    seeds = [(1001 + 17 * t, 2002 + t, 3003 + 5 * t, 4004 + t) for t in range(traces)]
    rec = torch.zeros(traces * cap, 9, device="cuda")
    scratch = torch.empty(traces * cap * 12, device="cuda")
    counts = torch.zeros(traces, dtype=torch.int32, device="cuda")
    launches = (L.VrLaunch * traces)(*[L.launch(sd, bw_mode=L.VR_BW_CONSISTENT) for sd in seeds])
    L.check(lib.vr_nrc_paths_traces(lw.medium, cam, launches, traces, n_paths, maxv, scratch.data_ptr(), rec.data_ptr(),
                                    cap, counts.data_ptr(), None), "vr_nrc_paths_traces")
    torch.cuda.synchronize()
    for t in range(traces):
        one = torch.zeros(cap, 9, device="cuda")
        cnt = torch.zeros(1, dtype=torch.int32, device="cuda")
        L.check(lib.vr_nrc_paths(lw.medium, cam, L.launch(seeds[t], bw_mode=L.VR_BW_CONSISTENT), n_paths, maxv,
                                 scratch.data_ptr(), one.data_ptr(), cap, cnt.data_ptr(), None), "vr_nrc_paths")
        torch.cuda.synchronize()
        n = int(cnt.item())
        assert n == int(counts[t].item()) and n > n_paths // 2
        a = np.unique(rec[t * cap:t * cap + n].cpu().numpy(), axis=0)
        b = np.unique(one[:n].cpu().numpy(), axis=0)
        np.testing.assert_array_equal(a, b)
