"""Edge cases of the HIP path: sensor lists, batching, strided grids, empty
launches, degenerate media and rays that miss or start inside the box."""

import numpy as np
import pytest
import torch

import rendervous_amd as rv
from helpers import (L, Scene, Built, Camera, blob_grid, fbm_grid, gpu_capture_fw, gpu_map_fw, match_fraction,
                     oracle_capture_fw, oracle_map_fw, pose, random_rays)

pytestmark = pytest.mark.gpu

SEEDS = (12345, 67890, 13579, 24680)


def _nee_scene(**kw):
    args = dict(sigma=fbm_grid(24, seed=9, scale=20.0), albedo_c=(0.9, 0.8, 0.7), phase_kind=L.VR_PHASE_HG,
                phase_g=0.6, env_c=(0.5, 0.5, 0.5), sun_dir=(0.3, 0.9, 0.3), sun_radiance=(3.0, 3.0, 3.0),
                sun_prob=0.5, majorant="grid", transmittance_kind=L.VR_T_DT, max_bounces=16)
    args.update(kw)
    return Scene(**args)


@pytest.mark.parametrize("kind", [L.VR_MAP_NEE_DT, L.VR_MAP_MC_COLLISION])
def test_capture_of_a_sensor_list_is_seed_matched(gpu, kind):
    """Explicit (pose, row, col) sensors with duplicates, a row outside the image
    (used as given, as the reference does) and a pose outside the pose buffer
    (0): the persistent NEE kernel's sensor-list path and the generic one."""
    print(f"[case] sensor_list kind={kind}", flush=True)
    rng = np.random.default_rng(3)
    sensors = np.stack([rng.integers(0, 2, 300), rng.integers(0, 32, 300), rng.integers(0, 40, 300)], 1)
    sensors[7] = sensors[3]
    sensors[11] = (2, 3, 5)   # pose out of range
    sensors[12] = (0, 40, 5)  # row outside the image
    poses = np.stack([pose((0.0, 0.3, 2.5)), pose((1.5, 0.2, -1.8))])
    cam = dict(poses=poses, width=40, height=32, jittered=True, sensors=sensors)
    sc = _nee_scene(transmittance_kind=L.VR_T_RT)
    ref = oracle_capture_fw(sc, kind, cam, SEEDS, 3)
    out = gpu_capture_fw(sc, kind, cam, SEEDS, 3)
    assert np.all(out[11] == 0) and np.all(ref[11] == 0)
    assert match_fraction(out, ref, rtol=1e-3, atol=1e-4) >= 0.97


def _gpu_capture_range(sc, kind, cam_args, spp, start, count, out):
    b = Built(sc, "cuda")
    cam = Camera(device="cuda", **cam_args)
    lau = L.launch(SEEDS, samples=spp, start_index=start, count=count)
    L.check(L.load().vr_capture_fw(b.medium, kind, cam.c, lau, out.data_ptr(), None), "vr_capture_fw")


@pytest.mark.parametrize("kind", [L.VR_MAP_NEE_DT, L.VR_MAP_PATH_DT])
def test_batched_capture_equals_one_launch(gpu, kind):
    """start_index/count batches (_maps.py:807-810) draw the same per-sample
    streams as one launch, so the image is the same up to summation order."""
    print(f"[case] batched kind={kind}", flush=True)
    sc = _nee_scene()
    cam = dict(poses=np.stack([pose((0.0, 0.3, 2.5)), pose((-2.0, 0.5, 1.0))]), width=37, height=29, jittered=True)
    n = 2 * 37 * 29
    whole = torch.zeros(n, 3, device="cuda")
    _gpu_capture_range(sc, kind, cam, 8, 0, -1, whole)
    parts = torch.full((n, 3), -1.0, device="cuda")
    for start in range(0, n, 500):
        _gpu_capture_range(sc, kind, cam, 8, start, min(500, n - start), parts)
    torch.cuda.synchronize()
    np.testing.assert_allclose(parts.cpu().numpy(), whole.cpu().numpy(), rtol=2e-5, atol=1e-6)


def test_empty_launch_leaves_the_output_alone(gpu):
    sc = _nee_scene()
    cam = dict(poses=pose((0.0, 0.3, 2.5))[None], width=8, height=8)
    out = torch.full((64, 3), 7.0, device="cuda")
    _gpu_capture_range(sc, L.VR_MAP_NEE_DT, cam, 4, 10, 0, out)
    torch.cuda.synchronize()
    assert float(out.min()) == 7.0 and float(out.max()) == 7.0


def test_strided_density_view_equals_its_contiguous_copy(gpu):
    """A non-contiguous density (a transposed view, Parameter strides of
    common.h:88-94) is read through its strides: identical DDA images."""
    g = torch.from_numpy(blob_grid(20, seed=5, max_value=15.0)).cuda()
    view = g.permute(2, 1, 0, 3)  # strides (1, 20, 400, 1) in (z, y, x, c) order
    assert not view.is_contiguous()
    cam = rv.PerspectiveCameraSensor(48, 48, rv.functions.look_at_pose((0.4, 0.3, 2.2))[None].cuda())
    imgs = []
    for t in (view, view.contiguous()):
        bmin, bmax = rv.normalized_box(t)
        field = rv.GridDDATransmittance(rv.Grid3D(t, bmin, bmax), rv.RayBoxIntersection(bmin, bmax))
        with torch.no_grad():
            imgs.append(cam.capture(field).cpu())
    assert torch.equal(imgs[0], imgs[1])


@pytest.mark.parametrize("kind", [L.VR_MAP_T_DT, L.VR_MAP_T_RT, L.VR_MAP_T_DDA, L.VR_MAP_T_GDT, L.VR_MAP_T_GRT,
                                  L.VR_MAP_NEE_DT, L.VR_MAP_PATH_DT])
def test_empty_medium(gpu, kind):
    """sigma = 0: transmittance 1 and radiance = environment, exactly."""
    sc = Scene(sigma=np.zeros((6, 6, 6, 1), np.float32), albedo_c=(0.8, 0.8, 0.8), env_c=(0.25, 0.5, 1.0),
               phase_kind=L.VR_PHASE_HG, phase_g=0.3)
    rays = random_rays(1024, seed=31)
    out = gpu_map_fw(sc, kind, rays, SEEDS)
    want = 1.0 if L.MAP_DIMS[kind][1] == 1 else np.array([0.25, 0.5, 1.0], np.float32)
    np.testing.assert_allclose(out, np.broadcast_to(want, out.shape), rtol=1e-6)


@pytest.mark.parametrize("kind", [L.VR_MAP_T_DT, L.VR_MAP_T_RT, L.VR_MAP_NEE_DT, L.VR_MAP_MC_COLLISION])
def test_rays_that_miss_or_start_inside_the_box(gpu, kind):
    """Rays pointing away from the box (T = 1, radiance = env) and rays that
    start inside it, seed-matched with the oracle."""
    sc = _nee_scene(transmittance_kind=L.VR_T_RT, majorant="const")
    away = random_rays(512, seed=32)
    away[:, 3:] *= -1.0  # origins on a sphere of radius 1.2, pointing outwards
    inside = random_rays(512, seed=33, inside=True)
    rays = np.concatenate([away, inside]).astype(np.float32)
    ref = oracle_map_fw(sc, kind, rays, SEEDS)
    out = gpu_map_fw(sc, kind, rays, SEEDS)
    want = 1.0 if L.MAP_DIMS[kind][1] == 1 else np.array([0.5, 0.5, 0.5], np.float32)
    np.testing.assert_allclose(out[:512], np.broadcast_to(want, out[:512].shape), rtol=1e-6)
    assert match_fraction(out, ref, rtol=1e-3, atol=1e-4) >= 0.98


def test_cell_table_lookups_are_bit_identical(gpu):
    """Captures reading the sigma cell table (vr_grid_pack_cells) equal captures
    reading the grid bit for bit: the persistent NEE kernel and a generic capture
    kernel, on a strided density view; the table itself equals the oracle's."""
    import torch
    from helpers import O, np_grid, pose
    g = fbm_grid(24, seed=3, scale=25.0)
    wide = np.zeros(g.shape[:3] + (3,), np.float32)
    wide[..., 1:2] = g
    sc = Scene(sigma=g, albedo_c=(0.8, 0.8, 0.8), phase_g=0.5, env_c=(0.2, 0.3, 0.4), sun_dir=(0.3, 0.9, 0.3),
               sun_radiance=(2.0, 2.0, 2.0), sun_prob=0.5, majorant="grid", majorant_res=(4, 4, 4), max_bounces=8)
    b = Built(sc, "cuda")
    dev_wide = torch.from_numpy(wide).cuda()
    view = dev_wide[..., 1:2]  # strided: channel stride 3
    b.medium.sigma.t = L.tensor_view(view)
    n = 23 ** 3
    cells = torch.empty(n * 8, device="cuda")
    L.check(L.load().vr_grid_pack_cells(b.medium.sigma, cells.data_ptr(), torch.cuda.current_stream().cuda_stream),
            "pack")
    ref = np.zeros((n, 8), np.float32)
    O.check(O.load().oracle_grid_pack_cells(np_grid(np.ascontiguousarray(g)), O.nptr(ref)), "oracle pack")
    torch.cuda.synchronize()
    np.testing.assert_array_equal(cells.cpu().numpy().reshape(n, 8), ref)
    cam = Camera(np.stack([pose((0.2, 0.4, 2.2)), pose((-1.9, 0.1, 1.0))]), 24, 24, jittered=True, device="cuda")
    for kind in (L.VR_MAP_NEE_DT, L.VR_MAP_MC_COLLISION, L.VR_MAP_PATH_DT):
        outs = []
        for use in (False, True):
            b.medium.sigma_cells = cells.data_ptr() if use else None
            out = torch.zeros(cam.threads, 3, device="cuda")
            # one sample per pixel for the persistent kernel (its per-item partial
            # sums are added atomically, in no fixed order, when spp > 1)
            spp = 1 if kind == L.VR_MAP_NEE_DT else 4
            lau = L.launch((5, 6, 7, 8), samples=spp, bw_mode=L.VR_BW_CONSISTENT)
            L.check(L.load().vr_capture_fw(b.medium, kind, cam.c, lau, out.data_ptr(),
                                           torch.cuda.current_stream().cuda_stream), "capture")
            torch.cuda.synchronize()
            outs.append(out.cpu().numpy())
        np.testing.assert_array_equal(outs[0], outs[1])


@pytest.mark.parametrize("tkind,cells", [(L.VR_T_DT, True), (L.VR_T_DT, False), (L.VR_T_RT, False)])
def test_lds_majorant_equals_global(gpu, tkind, cells):
    """The persistent NEE kernel with its coarse majorant grid staged in LDS
    (VR_MAJ_LDS=1, grids of at most 4096 cells and constant environments)
    renders the same frame, bit for bit at one sample per pixel, as with the
    grid read from global memory (the default): lean (cell table) and generic
    forms, DT and RT shadow rays."""
    import os
    print(f"[case] lds majorant t_kind={tkind} cells={cells}", flush=True)
    sc = Scene(sigma=fbm_grid(24, seed=4, scale=25.0), albedo_c=(0.8, 0.8, 0.8), phase_g=0.5, env_c=(0.2, 0.3, 0.4),
               sun_dir=(0.3, 0.9, 0.3), sun_radiance=(2.0, 2.0, 2.0), sun_prob=0.5, majorant="grid",
               majorant_res=(16, 16, 16), max_bounces=8, transmittance_kind=tkind)
    b = Built(sc, "cuda")
    if cells:
        n = 23 ** 3
        table = torch.empty(n * 8, device="cuda")
        L.check(L.load().vr_grid_pack_cells(b.medium.sigma, table.data_ptr(), torch.cuda.current_stream().cuda_stream),
                "pack")
        b.medium.sigma_cells = table.data_ptr()
    cam = Camera(np.stack([pose((0.2, 0.4, 2.2)), pose((-1.9, 0.1, 1.0))]), 24, 24, jittered=True, device="cuda")
    outs = []
    old = os.environ.get("VR_MAJ_LDS")
    try:
        for flag in ("0", "1"):
            os.environ["VR_MAJ_LDS"] = flag
            out = torch.zeros(cam.threads, 3, device="cuda")
            lau = L.launch((5, 6, 7, 8), samples=1, bw_mode=L.VR_BW_CONSISTENT)
            L.check(L.load().vr_capture_fw(b.medium, L.VR_MAP_NEE_DT, cam.c, lau, out.data_ptr(),
                                           torch.cuda.current_stream().cuda_stream), "capture")
            torch.cuda.synchronize()
            outs.append(out.cpu().numpy())
    finally:
        if old is None:
            os.environ.pop("VR_MAJ_LDS", None)
        else:
            os.environ["VR_MAJ_LDS"] = old
    assert outs[0].mean() > 0.0
    np.testing.assert_array_equal(outs[0], outs[1])


@pytest.mark.parametrize("spp", [1, 8])
def test_compacted_samples_equal_the_per_wave_handlers(gpu, spp):
    """The persistent NEE kernel with its light and phase samples computed
    block-wide (VR_NEE_COMPACT=1, lean constant-environment form) draws the
    same numbers per sample as the per-wave handlers: the same frame to float
    rounding (the moved code contracts a few products into FMAs differently:
    ~1 % of the pixels differ by an ulp at one sample per pixel; at 8 the
    per-pixel sums also add in the order of the atomics)."""
    import os
    print(f"[case] compacted samples spp={spp}", flush=True)
    sc = Scene(sigma=fbm_grid(24, seed=4, scale=25.0), albedo_c=(0.8, 0.8, 0.8), phase_g=0.5, env_c=(0.2, 0.3, 0.4),
               sun_dir=(0.3, 0.9, 0.3), sun_radiance=(2.0, 2.0, 2.0), sun_prob=0.5, majorant="grid",
               majorant_res=(16, 16, 16), max_bounces=8, transmittance_kind=L.VR_T_DT)
    b = Built(sc, "cuda")
    table = torch.empty(23 ** 3 * 8, device="cuda")
    L.check(L.load().vr_grid_pack_cells(b.medium.sigma, table.data_ptr(), torch.cuda.current_stream().cuda_stream),
            "pack")
    b.medium.sigma_cells = table.data_ptr()
    cam = Camera(np.stack([pose((0.2, 0.4, 2.2)), pose((-1.9, 0.1, 1.0))]), 32, 32, jittered=True, device="cuda")
    outs = []
    old = os.environ.get("VR_NEE_COMPACT")
    try:
        for flag in ("0", "1"):
            os.environ["VR_NEE_COMPACT"] = flag
            out = torch.zeros(cam.threads, 3, device="cuda")
            lau = L.launch((5, 6, 7, 8), samples=spp, bw_mode=L.VR_BW_CONSISTENT)
            L.check(L.load().vr_capture_fw(b.medium, L.VR_MAP_NEE_DT, cam.c, lau, out.data_ptr(),
                                           torch.cuda.current_stream().cuda_stream), "capture")
            torch.cuda.synchronize()
            outs.append(out.cpu().numpy())
    finally:
        if old is None:
            os.environ.pop("VR_NEE_COMPACT", None)
        else:
            os.environ["VR_NEE_COMPACT"] = old
    assert outs[0].sum() > 0
    np.testing.assert_allclose(outs[1], outs[0], rtol=1e-5, atol=1e-6)


@pytest.mark.parametrize("bw_mode,majorant,seed_mode,env", [
    (L.VR_BW_CONSISTENT, "grid", L.VR_SEED_CAPTURE_FW, "sun"), (L.VR_BW_REFERENCE, "grid", L.VR_SEED_CAPTURE_FW, "sun"),
    (L.VR_BW_CONSISTENT, "const", L.VR_SEED_CAPTURE_FW, "sun"), (L.VR_BW_REFERENCE, "const", L.VR_SEED_CAPTURE_FW, "sun"),
    (L.VR_BW_CONSISTENT, "grid", L.VR_SEED_CAPTURE_BW, "sun"), (L.VR_BW_CONSISTENT, "grid", L.VR_SEED_CAPTURE_FW, "image")])
def test_flat_nee_backward_equals_the_nested_one(gpu, bw_mode, majorant, seed_mode, env):
    """The NEE capture backward as a flat lane state machine, with each thread's
    own samples (VR_NEE_BW_POOL=0) and with lanes pulling work items from a
    counter (the default for DT shadow rays and per-sample streams), scatters the
    same gradients, sample by sample, as the nested estimator
    (VR_NEE_BW_FLAT=0): same scatter counts, density and albedo gradients equal
    up to the order of the float atomics.  Sensors with a zero output gradient
    trace nothing.  Also with the reference's backward seeds (one stream over a
    sensor's samples: the pooled form does not apply) and an environment image
    sampled through its quadtree."""
    import os
    import torch
    from helpers import env_density, np_quadtree, sky_image
    print(f"[case] flat nee bw mode={bw_mode} majorant={majorant} seeds={seed_mode} env={env}", flush=True)
    g = fbm_grid(20, seed=5, scale=30.0)
    alb = (0.5 + 0.45 * fbm_grid(8, seed=6, scale=1.0)).astype(np.float32)
    light = dict(env_c=(0.3, 0.4, 0.5), sun_dir=(0.3, 0.9, 0.3), sun_radiance=(2.0, 2.0, 2.0), sun_prob=0.4)
    if env == "image":
        img = sky_image(16, seed=2)
        light = dict(env_image=img, env_quadtree=np_quadtree(env_density(img)), env_quadtree_levels=4)
    sc = Scene(sigma=g, albedo=np.repeat(alb, 3, -1), phase_g=0.4, majorant=majorant, majorant_res=(4, 4, 4),
               max_bounces=6, transmittance_kind=L.VR_T_DT, **light)
    cam = dict(poses=np.stack([pose((0.2, 0.4, 2.2)), pose((-1.9, 0.1, 1.0))]), width=20, height=20, jittered=True)
    og = (np.random.default_rng(3).uniform(size=(800, 3)) + 0.5).astype(np.float32)
    og[::7] = 0.0
    res = []
    old = {k: os.environ.get(k) for k in ("VR_NEE_BW_FLAT", "VR_NEE_BW_POOL")}
    try:
        for flat, pool in (("0", "0"), ("1", "0"), ("1", "1")):
            os.environ["VR_NEE_BW_FLAT"] = flat
            os.environ["VR_NEE_BW_POOL"] = pool
            b = Built(sc, "cuda", grad_sigma=True, grad_albedo=True)
            c = Camera(device="cuda", **cam)
            dbg = torch.zeros(c.threads, dtype=torch.int32, device="cuda")
            o = torch.from_numpy(og).cuda()
            lau = L.launch((9, 8, 7, 6), samples=8, bw_mode=bw_mode, seed_mode=seed_mode, debug_out=dbg)
            L.check(L.load().vr_capture_bw(b.medium, L.VR_MAP_NEE_DT, c.c, lau, o.data_ptr(),
                                           torch.cuda.current_stream().cuda_stream), "capture_bw")
            torch.cuda.synchronize()
            res.append((b.grads(), dbg.cpu().numpy()))
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    from helpers import rel_l1
    (gs0, ga0), d0 = res[0]
    assert np.abs(gs0).sum() > 0 and np.abs(ga0).sum() > 0
    assert (d0[::7] == 0).all() and (d0 > 0).mean() > 0.2
    for (gs1, ga1), d1 in res[1:]:
        np.testing.assert_array_equal(d0, d1)
        assert rel_l1(gs1, gs0) <= 1e-5 and rel_l1(ga1, ga0) <= 1e-5, (rel_l1(gs1, gs0), rel_l1(ga1, ga0))


@pytest.mark.parametrize("majorant,seed_mode,env,albedo_grid", [
    ("grid", L.VR_SEED_CAPTURE_FW, "sun", False), ("const", L.VR_SEED_CAPTURE_FW, "sun", True),
    ("grid", L.VR_SEED_CAPTURE_FW, "image", True), ("grid", L.VR_SEED_CAPTURE_BW, "sun", False)])
def test_pooled_single_scatter_backward_equals_the_nested_one(gpu, majorant, seed_mode, env, albedo_grid):
    """The collision-integrator backward with single-scatter exitance and RT
    shadow rays as a pooled lane state machine (VR_SS_BW_POOL=1; it applies to
    the consistent estimators, per-sample streams and no albedo gradient)
    scatters the same density gradients, sample by sample, as the nested
    estimator (VR_SS_BW_POOL=0): same scatter and lookup counts per sensor, gradients
    equal up to the order of the float atomics.  Sensors with a zero output
    gradient trace nothing.  With the reference's backward seeds (one stream
    over a sensor's samples) the pooled form does not apply and both runs are
    the nested kernel."""
    import os
    import torch
    from helpers import env_density, np_quadtree, rel_l1, sky_image
    print(f"[case] pooled ss bw majorant={majorant} seeds={seed_mode} env={env} albedo_grid={albedo_grid}", flush=True)
    g = fbm_grid(20, seed=7, scale=30.0)
    light = dict(env_c=(0.3, 0.4, 0.5), sun_dir=(0.3, 0.9, 0.3), sun_radiance=(2.0, 2.0, 2.0), sun_prob=0.4)
    if env == "image":
        img = sky_image(16, seed=2)
        light = dict(env_image=img, env_quadtree=np_quadtree(env_density(img)), env_quadtree_levels=4)
    alb = None
    if albedo_grid:
        alb = np.repeat((0.5 + 0.45 * fbm_grid(8, seed=6, scale=1.0)).astype(np.float32), 3, -1)
    sc = Scene(sigma=g, albedo=alb, phase_g=0.4, majorant=majorant, majorant_res=(4, 4, 4),
               transmittance_kind=L.VR_T_RT, **light)
    cam = dict(poses=np.stack([pose((0.2, 0.4, 2.2)), pose((-1.9, 0.1, 1.0))]), width=20, height=20, jittered=True)
    og = (np.random.default_rng(4).uniform(size=(800, 3)) + 0.5).astype(np.float32)
    og[::7] = 0.0
    res = []
    old = os.environ.get("VR_SS_BW_POOL")
    try:
        for pool in ("0", "1"):
            os.environ["VR_SS_BW_POOL"] = pool
            counts = []
            for kind in (0, 1):
                b = Built(sc, "cuda", grad_sigma=True)
                c = Camera(device="cuda", **cam)
                dbg = torch.zeros(c.threads, dtype=torch.int32, device="cuda")
                o = torch.from_numpy(og).cuda()
                lau = L.launch((9, 8, 7, 6), samples=8, bw_mode=L.VR_BW_CONSISTENT, seed_mode=seed_mode, debug_out=dbg,
                               debug_kind=kind)
                L.check(L.load().vr_capture_bw(b.medium, L.VR_MAP_MC_COLLISION, c.c, lau, o.data_ptr(),
                                               torch.cuda.current_stream().cuda_stream), "capture_bw")
                torch.cuda.synchronize()
                counts.append(dbg.cpu().numpy())
            res.append((b.grads()[0], counts))
    finally:
        if old is None:
            os.environ.pop("VR_SS_BW_POOL", None)
        else:
            os.environ["VR_SS_BW_POOL"] = old
    (gs0, (s0, l0)), (gs1, (s1, l1)) = res
    assert np.abs(gs0).sum() > 0
    assert (s0[::7] == 0).all() and (s0 > 0).mean() > 0.2 and (l0 > 0).mean() > 0.2
    np.testing.assert_array_equal(s1, s0)
    np.testing.assert_array_equal(l1, l0)
    assert rel_l1(gs1, gs0) <= 1e-5, rel_l1(gs1, gs0)
