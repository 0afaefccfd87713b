"""Environment images and quadtree importance sampling on the HIP path, seed
matched with the oracle; the image tools (resample_img, Morton copies,
create_density_quadtree) against the oracle and numpy."""
import numpy as np
import pytest
import torch

import rendervous_amd as rv
from helpers import (L, O, Scene, env_density, fbm_grid, gpu_capture_fw, gpu_map_fw, match_fraction, np_quadtree,
                     oracle_capture_fw, oracle_map_fw, pose, random_rays, sky_image)

pytestmark = pytest.mark.gpu

SEEDS = (12345, 67890, 13579, 24680)


def _stream():
    return L.stream_handle(torch.device("cuda", 0))


@pytest.mark.parametrize("src,dst,C", [((8, 8), (4, 4), 1), ((16, 32), (31, 17), 3), ((33,), (70,), 2),
                                       ((12, 20, 16), (6, 10, 32), 1), ((5, 6, 7), (9, 4, 13), 2)])
def test_resample_img_matches_oracle(gpu, src, dst, C):
    x = np.random.default_rng(len(src) * C).normal(size=src + (C,)).astype(np.float32)
    s_in, s_out = np.asarray(src, np.int64), np.asarray(dst, np.int64)
    ref = np.zeros(dst + (C,), np.float32)
    O.check(O.load().oracle_resample_img(O.nptr(x), O.nptr(s_in), O.nptr(s_out), len(src), C, O.nptr(ref)), "oracle")
    xd = torch.from_numpy(x).cuda()
    out = torch.empty(dst + (C,), device="cuda")
    L.check(L.load().vr_resample_img(xd.data_ptr(), L.i64s(src), L.i64s(dst), len(src), C, out.data_ptr(), _stream()),
            "vr_resample_img")
    torch.cuda.synchronize()
    np.testing.assert_allclose(out.cpu().numpy(), ref, rtol=1e-6, atol=1e-6)


def test_morton_copies_and_density_quadtree(gpu):
    img = sky_image(64, seed=3)
    x = torch.from_numpy(img).cuda()
    m = rv.functions.copy_img_to_morton(x)
    ref = np.zeros((64 * 64, 3), np.float32)
    O.check(O.load().oracle_copy_img_to_morton(O.nptr(img), 64, 3, O.nptr(ref)), "oracle")
    np.testing.assert_array_equal(m.cpu().numpy(), ref)
    assert torch.equal(rv.functions.copy_morton_to_img(m), x)
    d = env_density(img)
    q = rv.functions.create_density_quadtree(torch.from_numpy(d.astype(np.float32))[..., None].cuda())
    np.testing.assert_allclose(q[:, 0].cpu().numpy(), np_quadtree(d), rtol=2e-5, atol=1e-9)


def _env_scene(tkind=L.VR_T_DT, quadtree=True, R=32, **kw):
    img = sky_image(R, seed=4)
    args = dict(sigma=fbm_grid(24, seed=9, scale=20.0), albedo_c=(0.9, 0.8, 0.7), phase_kind=L.VR_PHASE_HG,
                phase_g=0.5, majorant="grid", transmittance_kind=tkind, max_bounces=16, env_image=img,
                env_quadtree=np_quadtree(env_density(img)) if quadtree else None,
                env_quadtree_levels=int(np.log2(R)) if quadtree else 0)
    args.update(kw)
    return Scene(**args)


def test_environment_lookup_and_trackers_with_an_image(gpu):
    """sigma = 0 maps return environment(w); a medium's transmittance-weighted
    lookups stay seed-matched."""
    empty = _env_scene(sigma=np.zeros((6, 6, 6, 1), np.float32), majorant="const")
    rays = random_rays(4096, seed=21)
    np.testing.assert_allclose(gpu_map_fw(empty, L.VR_MAP_NEE_DT, rays, SEEDS),
                               oracle_map_fw(empty, L.VR_MAP_NEE_DT, rays, SEEDS), rtol=2e-5, atol=1e-6)
    sc = _env_scene()
    for kind in (L.VR_MAP_MC_COLLISION, L.VR_MAP_PATH_DT, L.VR_MAP_NEE_DT):
        out = gpu_map_fw(sc, kind, rays, SEEDS)
        ref = oracle_map_fw(sc, kind, rays, SEEDS)
        assert match_fraction(out, ref, rtol=1e-3, atol=1e-4) >= 0.97, kind


@pytest.mark.parametrize("quadtree", [True, False])
@pytest.mark.parametrize("tkind", [L.VR_T_DT, L.VR_T_RT, L.VR_T_GRT])
def test_nee_capture_with_environment_image_seed_matched(gpu, tkind, quadtree):
    """DT / RT shadow rays go through the persistent NEE kernel, GRT through the
    generic capture kernel; quadtree or uniform environment sampling."""
    print(f"[case] env capture tkind={tkind} quadtree={quadtree}", flush=True)
    sc = _env_scene(tkind, quadtree)
    cam = dict(poses=pose((0.0, 0.3, 2.5))[None], width=40, height=40, jittered=True)
    ref = oracle_capture_fw(sc, L.VR_MAP_NEE_DT, cam, SEEDS, 2)
    out = gpu_capture_fw(sc, L.VR_MAP_NEE_DT, cam, SEEDS, 2)
    assert match_fraction(out, ref, rtol=1e-3, atol=1e-4) >= 0.97


def test_environment_recipes_build_the_quadtree_sampler(gpu):
    """environment_tensor [R, R, 3] -> Image2D after xr_projection, and
    medium_environment_sampler -> the quadtree of _objects.py:180-199; the render
    equals the oracle's with the numpy quadtree of the same image."""
    R = 32
    img = sky_image(R, seed=6)
    dev = torch.device("cuda", 0)
    sigma = torch.from_numpy(fbm_grid(24, seed=9, scale=20.0)).to(dev)
    ds = rv.objects.DependencySet()
    ds.add_parameters(sigma_tensor=sigma, scattering_albedo_tensor=torch.tensor([0.9, 0.8, 0.7]),
                      environment_tensor=torch.from_numpy(img).to(dev), phase_g_tensor=torch.tensor([0.5]))
    ds.requires(rv.objects.medium_environment_sampler_quadtree, levels=5)
    ds.requires(rv.objects.medium_phase_HG, mixture=False)
    ds.requires(rv.objects.medium_phase_sampler_HG, mixture=False)
    ds.requires(rv.objects.medium_transmittance_RT)
    ds.requires(rv.objects.medium_radiance_path_integrator_NEE_DT, max_bounces=16)
    q = ds.environment_sampler_quadtree
    np.testing.assert_allclose(q[:, 0].cpu().numpy(), np_quadtree(env_density(img)), rtol=2e-5, atol=1e-9)
    kind, lw = ds.radiance._lower_map()
    assert lw.medium.env_kind == L.VR_ENV_XR and lw.medium.env_sampler_kind == L.VR_ENVS_QUADTREE
    camera = rv.PerspectiveCameraSensor(32, 32, rv.functions.look_at_pose((0.0, 0.3, 2.5))[None].to(dev), jittered=True)
    rv.seed(3)
    seeds = rv.get_seeds()
    rv.seed(3)
    with torch.no_grad():
        img_gpu = camera.capture(ds.radiance, fw_samples=2).reshape(-1, 3).cpu().numpy()
    sc = Scene(sigma=fbm_grid(24, seed=9, scale=20.0), albedo_c=(0.9, 0.8, 0.7), phase_kind=L.VR_PHASE_HG,
               phase_g=0.5, transmittance_kind=L.VR_T_RT, max_bounces=16, env_image=img,
               env_quadtree=np_quadtree(env_density(img)), env_quadtree_levels=5,
               majorant_value=float(sigma.max()))
    cam = dict(poses=pose((0.0, 0.3, 2.5))[None], width=32, height=32, jittered=True)
    ref = oracle_capture_fw(sc, L.VR_MAP_NEE_DT, cam, seeds, 2)
    assert match_fraction(img_gpu, ref, rtol=1e-3, atol=1e-4) >= 0.97
