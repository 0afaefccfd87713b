"""Scene-level path tracing on the HIP path (BVH traversal in LDS, surfaces,
media between surface hits) against the oracle's brute-force closest hit and
path tracer with the same seeds; the Python scene API end to end."""
import math

import numpy as np
import pytest
import torch

import rendervous_amd as rv
from helpers import (L, SceneBuilt, blob_grid, box_mesh, match_fraction, patch, pose, quad_mesh, rel_l1,
                     scene_capture_bw, scene_capture_fw, scene_map_fw, scene_medium, scene_raycast, SEED_MATCHED_L1)

pytestmark = pytest.mark.gpu

SEEDS = (2718, 28182, 8459, 45235)


def _soup(n, seed):
    rng = np.random.default_rng(seed)
    c = rng.uniform(-1, 1, size=(n, 1, 3))
    tri = c + rng.normal(scale=0.1, size=(n, 3, 3))
    return tri.reshape(-1, 3).astype(np.float32), np.arange(3 * n, dtype=np.int32).reshape(n, 3)


def _rays(n, seed, radius=2.5):
    rng = np.random.default_rng(seed)
    u = rng.normal(size=(n, 3))
    o = u / np.linalg.norm(u, axis=1, keepdims=True) * radius
    d = rng.uniform(-0.8, 0.8, size=(n, 3)) - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    inside = rng.uniform(-0.9, 0.9, size=(n // 4, 3))
    di = rng.normal(size=(n // 4, 3))
    di /= np.linalg.norm(di, axis=1, keepdims=True)
    return np.concatenate([np.concatenate([o, d], 1), np.concatenate([inside, di], 1)]).astype(np.float32)


def _mixed_scene(device, grad=False, g=None):
    """Null box around a scattering medium, a Lambert floor with an emitting
    panel, a mirror wall and a Fresnel slab face."""
    g = blob_grid(12, seed=6, max_value=6.0) if g is None else g
    med = scene_medium(g, majorant=6.5, albedo_c=(0.85, 0.8, 0.75), phase_g=0.4)
    fl_p, fl_i = quad_mesh(-0.7, half=4.0)
    mw_p = np.array([[2.0, -0.7, -2.0], [2.0, 1.5, -2.0], [2.0, 1.5, 2.0], [2.0, -0.7, 2.0]], np.float32)
    mw_i = np.array([[0, 1, 2], [0, 2, 3]], np.int32)
    gl_p, gl_i = box_mesh(0.25, center=(-1.2, -0.3, 0.6))
    meshes = [box_mesh(0.5), (fl_p, fl_i), (mw_p, mw_i), (gl_p, gl_i)]
    patches = [patch(L.VR_SURF_NULL, inside=0), patch(L.VR_SURF_LAMBERT, L.VR_EMIT_ISOTROPIC, (0.3, 0.3, 0.3)),
               patch(L.VR_SURF_MIRROR), patch(L.VR_SURF_FRESNEL, refraction_index=1.4)]
    return SceneBuilt(meshes, patches, media=[med], env_c=(0.8, 0.9, 1.0), device=device, grad_sigma=grad)


CAM = dict(poses=np.stack([pose((0.4, 0.9, 3.0)), pose((-2.5, 0.6, 1.5))]), width=24, height=24, jittered=True)


def test_bvh_raycast_matches_brute_force(gpu):
    pos, idx = _soup(3000, 1)
    bp, bi = box_mesh(1.0)
    meshes = [(pos, idx), (bp, bi)]
    patches = [patch(), patch()]
    rays = _rays(20000, 2)
    ref = scene_raycast(SceneBuilt(meshes, patches, device="cpu"), rays)
    out = scene_raycast(SceneBuilt(meshes, patches, device="cuda"), rays)
    hit_ref, hit = ref[:, 15].view(np.int32), out[:, 15].view(np.int32)
    assert (hit_ref >= 0).sum() > 15000
    assert np.mean(hit_ref == hit) >= 0.999, np.mean(hit_ref == hit)
    both = (hit_ref >= 0) & (hit >= 0)
    np.testing.assert_allclose(out[both, 0], ref[both, 0], rtol=1e-4, atol=1e-5)
    same_face = both & (np.abs(out[:, 4:7] - ref[:, 4:7]).max(1) < 1e-4)
    assert same_face[both].mean() >= 0.999


def test_empty_scene_and_miss_rays(gpu):
    rays = _rays(256, 3)
    sb = SceneBuilt([], [patch()], env_c=(0.2, 0.4, 0.6), device="cuda")
    h = scene_raycast(sb, rays)
    assert (h[:, 15].view(np.int32) == -1).all() and (h[:, :15] == 0).all()
    np.testing.assert_allclose(scene_map_fw(sb, rays, SEEDS), np.tile([0.2, 0.4, 0.6], (len(rays), 1)), rtol=1e-6)


def test_scene_map_forward_is_seed_matched(gpu):
    rays = _rays(4096, 4)
    ref = scene_map_fw(_mixed_scene("cpu"), rays, SEEDS)
    out = scene_map_fw(_mixed_scene("cuda"), rays, SEEDS)
    assert np.isfinite(out).all()
    frac = match_fraction(out, ref, rtol=1e-3, atol=1e-4)
    assert frac >= 0.97, frac


@pytest.mark.parametrize("spp", [1, 16])
def test_scene_capture_forward_is_seed_matched(gpu, spp):
    ref = scene_capture_fw(_mixed_scene("cpu"), CAM, SEEDS, spp)
    out = scene_capture_fw(_mixed_scene("cuda"), CAM, SEEDS, spp)
    frac = match_fraction(out, ref, rtol=1e-3, atol=1e-4)
    assert frac >= 0.97, frac
    assert abs(out.mean() - ref.mean()) <= 0.02 * ref.mean()


def test_scene_capture_backward_matches_oracle(gpu):
    rng = np.random.default_rng(5)
    og = (rng.uniform(size=(2 * 24 * 24, 3)) + 0.5).astype(np.float32)
    ref = scene_capture_bw(_mixed_scene("cpu", grad=True), CAM, SEEDS, 8, og)[0][0]
    out = scene_capture_bw(_mixed_scene("cuda", grad=True), CAM, SEEDS, 8, og)[0][0]
    assert np.abs(ref).sum() > 0
    assert rel_l1(out, ref) <= SEED_MATCHED_L1


def test_scene_api_capture_and_gradient(gpu):
    """Scene.from_graph(...).pathtrace() through PerspectiveCameraSensor.capture:
    the image equals the C-ABI capture of the same arrays, and autograd returns
    the capture backward's gradient."""
    dev = gpu
    g = blob_grid(12, seed=6, max_value=6.0)
    sigma = torch.nn.Parameter(torch.from_numpy(g).to(dev), requires_grad=True)
    bmin, bmax = rv.normalized_box(sigma)
    grid = rv.Grid3D(sigma, bmin, bmax)
    medium = rv.HeterogeneousMediumMaterial(grid, scattering_albedo=rv.const[0.85, 0.8, 0.75],
                                            phase_sampler=rv.HGPhaseSampler(0.4), majorant=6.5)
    bp, bi = box_mesh(0.5)
    fp, fi = quad_mesh(-0.7, half=4.0)
    geoms = {"cloud": rv.MeshGeometry(torch.from_numpy(bp), indices=torch.from_numpy(bi), compute_normals=False),
             "floor": rv.MeshGeometry(torch.from_numpy(fp), indices=torch.from_numpy(fi), compute_normals=False)}
    mats = {"cloud": rv.SurfaceMaterial(rv.NullSurfaceScatteringSampler()),
            "floor": rv.SurfaceMaterial(rv.LambertSurfaceScatteringSampler(),
                                        rv.IsotropicSurfaceEmission(rv.const[0.3, 0.3, 0.3]))}
    scene = rv.Scene.from_graph(geoms, rv.UniformEnvironmentSampler(rv.const[0.8, 0.9, 1.0]), materials=mats,
                                inside_media={"cloud": medium})
    field = scene.pathtrace()
    camera = rv.PerspectiveCameraSensor(24, 24, torch.from_numpy(CAM["poses"]).to(dev), jittered=True)
    rv.seed(17)
    seeds = rv.get_seeds()
    rv.seed(17)
    img = camera.capture(field, fw_samples=4)
    med = scene_medium(g, majorant=6.5, albedo_c=(0.85, 0.8, 0.75), phase_g=0.4)
    direct = SceneBuilt([(bp, bi), (fp, fi)], [patch(L.VR_SURF_NULL, inside=0),
                                               patch(L.VR_SURF_LAMBERT, L.VR_EMIT_ISOTROPIC, (0.3, 0.3, 0.3))],
                        media=[med], env_c=(0.8, 0.9, 1.0), device="cuda", grad_sigma=True)
    ref = scene_capture_fw(direct, CAM, seeds, 4)
    np.testing.assert_allclose(img.detach().reshape(-1, 3).cpu().numpy(), ref, rtol=1e-5, atol=1e-6)
    w = torch.rand_like(img)
    with rv.backward_mode(seeds="replay"):
        (img * w).sum().backward()
    gref = scene_capture_bw(direct, CAM, seeds, 4, w.reshape(-1, 3).cpu().numpy())[0][0]
    assert sigma.grad is not None
    assert rel_l1(sigma.grad.cpu().numpy(), gref) <= 1e-4


def test_raycast_interpolates_normals_and_uvs(gpu):
    """Shading normals and coordinates interpolated with the hit's barycentrics
    (sample_surfel) equal the oracle's."""
    pos, idx = box_mesh(0.8)
    rng = np.random.default_rng(7)
    nrm = rng.normal(size=pos.shape).astype(np.float32)
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    uvs = rng.uniform(size=(pos.shape[0], 2)).astype(np.float32)
    rays = _rays(8192, 8)
    ref = scene_raycast(SceneBuilt([(pos, idx)], [patch()], normals=nrm, uvs=uvs, device="cpu"), rays)
    out = scene_raycast(SceneBuilt([(pos, idx)], [patch()], normals=nrm, uvs=uvs, device="cuda"), rays)
    hit = ref[:, 15].view(np.int32) >= 0
    assert hit.sum() > 4000 and (out[:, 15].view(np.int32) == ref[:, 15].view(np.int32)).mean() >= 0.999
    same = hit & (np.abs(out[:, 4:7] - ref[:, 4:7]).max(1) < 1e-4)  # same face
    assert same[hit].mean() >= 0.999
    np.testing.assert_allclose(out[same, 1:4], ref[same, 1:4], atol=2e-4)
    np.testing.assert_allclose(out[same, 7:9], ref[same, 7:9], atol=2e-4)


def test_nested_media_scene_is_seed_matched(gpu):
    """Two scattering media in nested null boxes (the medium stack pushes and pops)
    and a Fresnel box inside the inner one."""
    def build(device):
        ga = blob_grid(8, seed=1, max_value=2.0)
        gb = blob_grid(8, seed=2, max_value=5.0)
        ma = scene_medium(ga, majorant=2.5, albedo_c=(0.9, 0.9, 0.9), phase_g=0.2, bmin=(-1.0,) * 3, bmax=(1.0,) * 3)
        mb = scene_medium(gb, majorant=5.5, albedo_c=(0.7, 0.8, 0.9), phase_g=-0.3, bmin=(-0.5,) * 3, bmax=(0.5,) * 3)
        return SceneBuilt([box_mesh(1.0), box_mesh(0.5), box_mesh(0.15, center=(0.1, 0.1, 0.0))],
                          [patch(L.VR_SURF_NULL, inside=0), patch(L.VR_SURF_NULL, inside=1),
                           patch(L.VR_SURF_FRESNEL, refraction_index=1.3)],
                          media=[ma, mb], env_c=(1.0, 0.9, 0.8), device=device)
    ref = scene_capture_fw(build("cpu"), CAM, SEEDS, 4)
    out = scene_capture_fw(build("cuda"), CAM, SEEDS, 4)
    frac = match_fraction(out, ref, rtol=1e-3, atol=1e-4)
    assert frac >= 0.97, frac
