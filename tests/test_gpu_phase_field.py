"""A spatially varying HG asymmetry: phase_g as a Grid3D read at each
scattering position (the reference builds a Grid3D for any multi-dimensional
phase_g_tensor, _objects.py:49-57,83-89), against the oracle sample by sample
and through the recipe layer."""
import numpy as np
import pytest
import torch

import rendervous_amd as rv
from helpers import (L, Scene, fbm_grid, gpu_capture_bw, gpu_capture_fw, gpu_map_fw, match_fraction,
                     oracle_capture_bw, oracle_capture_fw, oracle_map_fw, pose, random_rays, rel_l1,
                     SEED_MATCHED_L1)

pytestmark = pytest.mark.gpu

SEEDS = (31, 41, 59, 26)


def _print_case(name, **kw):
    print(f"[case] {name} " + " ".join(f"{k}={v}" for k, v in kw.items()), flush=True)


def phase_field(n: int, seed: int) -> np.ndarray:
    """g in (-0.9, 0.9): forward scattering on one side, backward on the other."""
    rng = np.random.default_rng(seed)
    c = np.linspace(-1.0, 1.0, n)
    z, y, x = np.meshgrid(c, c, c, indexing="ij")
    g = 0.9 * np.tanh(1.5 * x + 0.7 * np.sin(3.0 * y + 2.0 * z) + 0.2 * rng.standard_normal(x.shape))
    return g[..., None].astype(np.float32)


def _scene(**kw):
    base = dict(sigma=fbm_grid(24, seed=9, scale=20.0), albedo_c=(0.9, 0.8, 0.7), phase_kind=L.VR_PHASE_HG_MIX,
                phase_g=0.6, env_c=(0.1, 0.1, 0.1), sun_dir=(0.3, 0.9, 0.3), sun_radiance=(3.0, 3.0, 3.0),
                sun_prob=0.5, max_bounces=16, phase_g_grid=phase_field(12, 1))
    base.update(kw)
    return Scene(**base)


@pytest.mark.parametrize("phase_kind", [L.VR_PHASE_HG, L.VR_PHASE_HG_MIX])
@pytest.mark.parametrize("tkind", [L.VR_T_DT, L.VR_T_RT])
def test_nee_capture_with_a_phase_field_matches_the_oracle(gpu, phase_kind, tkind):
    _print_case("nee_phase_field", phase_kind=phase_kind, tkind=tkind)
    sc = _scene(phase_kind=phase_kind, transmittance_kind=tkind, majorant="grid")
    cam = dict(poses=pose((0.0, 0.3, 2.5))[None], width=48, height=48, jittered=True)
    ref = oracle_capture_fw(sc, L.VR_MAP_NEE_DT, cam, SEEDS, 2)
    out = gpu_capture_fw(sc, L.VR_MAP_NEE_DT, cam, SEEDS, 2)
    assert match_fraction(out, ref, rtol=1e-3, atol=1e-4) >= 0.98
    # the field matters: the constant-g image differs
    const = oracle_capture_fw(_scene(phase_kind=phase_kind, transmittance_kind=tkind, majorant="grid",
                                     phase_g_grid=None), L.VR_MAP_NEE_DT, cam, SEEDS, 2)
    assert match_fraction(const, ref, rtol=1e-3, atol=1e-4) < 0.9


@pytest.mark.parametrize("kind", [L.VR_MAP_PATH_DT, L.VR_MAP_NEE_DRT])
def test_path_integrators_with_a_phase_field_match_the_oracle(gpu, kind):
    _print_case("path_phase_field", kind=kind)
    sc = _scene(env_c=(0.5, 0.6, 0.7), emission_c=(0.1, 0.0, 0.05), majorant="grid")
    cam = dict(poses=pose((0.0, 0.3, 2.5))[None], width=48, height=48, jittered=True)
    ref = oracle_capture_fw(sc, kind, cam, SEEDS, 1)
    out = gpu_capture_fw(sc, kind, cam, SEEDS, 1)
    assert match_fraction(out, ref, rtol=1e-3, atol=1e-4) >= 0.98


@pytest.mark.parametrize("kind", [L.VR_MAP_SS_RADIANCE, L.VR_MAP_MC_COLLISION])
def test_single_scatter_with_a_phase_field_matches_the_oracle(gpu, kind):
    _print_case("ss_phase_field", kind=kind)
    sc = _scene(env_c=(1.0, 0.5, 0.25), transmittance_kind=L.VR_T_RT, sun_prob=0.0)
    rays = random_rays(4096, seed=12, inside=(kind == L.VR_MAP_SS_RADIANCE))
    ref = oracle_map_fw(sc, kind, rays, SEEDS)
    out = gpu_map_fw(sc, kind, rays, SEEDS)
    assert match_fraction(out, ref, rtol=1e-3, atol=1e-4) >= 0.99


@pytest.mark.parametrize("kind", [L.VR_MAP_NEE_DT, L.VR_MAP_MC_COLLISION])
def test_backward_with_a_phase_field_matches_the_oracle(gpu, kind):
    _print_case("bw_phase_field", kind=kind)
    sc = _scene(sigma=fbm_grid(16, seed=23, scale=10.0), albedo=np.full((16, 16, 16, 1), 0.85, np.float32),
                phase_kind=L.VR_PHASE_HG, env_c=(1.0, 1.0, 1.0), transmittance_kind=L.VR_T_RT, sun_prob=0.0,
                max_bounces=4, phase_g_grid=phase_field(9, 2))
    cam = dict(poses=pose((0.2, 0.3, 2.5))[None], width=24, height=24, jittered=True)
    _, od = L.MAP_DIMS[kind]
    og = np.random.default_rng(24).uniform(0.5, 1.5, size=(24 * 24, od)).astype(np.float32)
    ref_s, ref_a = oracle_capture_bw(sc, kind, cam, SEEDS, 2, og, grad_albedo=True)
    out_s, out_a = gpu_capture_bw(sc, kind, cam, SEEDS, 2, og, grad_albedo=True)
    assert rel_l1(out_s, ref_s) <= SEED_MATCHED_L1
    assert rel_l1(out_a, ref_a) <= SEED_MATCHED_L1


def _recipe(phase_g_tensor, sigma):
    ds = rv.objects.DependencySet()
    ds.add_parameters(sigma_tensor=sigma, scattering_albedo_tensor=torch.tensor([0.9, 0.9, 0.9]),
                      environment_tensor=torch.tensor([0.2, 0.2, 0.2]), phase_g_tensor=phase_g_tensor)
    ds.requires(rv.objects.medium_majorant_coarse_grid, resolution=8)
    ds.requires(rv.objects.medium_environment_sampler_sun, direction=(0.3, 0.9, 0.3), radiance=(2.0, 2.0, 2.0),
                probability=0.5)
    ds.requires(rv.objects.medium_transmittance_DT)
    ds.requires(rv.objects.medium_radiance_path_integrator_NEE_DT, max_bounces=8)
    return ds.radiance


def test_phase_g_tensor_lowers_to_a_field_through_the_recipes(gpu):
    """A [D,H,W,1] phase_g_tensor becomes a Grid3D that the capture reads (the
    generic persistent kernel: the lean one assumes a constant g).  A field
    holding one value everywhere renders the constant-g image (up to the
    rounding of the trilinear weights); a varying one does not; gradients reach
    the density and not the phase field."""
    _print_case("recipe_phase_field")
    sigma = torch.from_numpy(fbm_grid(24, seed=6, scale=20.0)).cuda()
    camera = rv.PerspectiveCameraSensor(40, 40, rv.functions.look_at_pose((0.0, 0.3, 2.5))[None].cuda(),
                                        jittered=True)
    flat = _recipe(torch.full((7, 7, 7, 1), 0.55, device="cuda"), sigma)
    kind, lw = flat._lower_map()
    assert kind == L.VR_MAP_NEE_DT and lw.medium.phase_g_kind == L.VR_FIELD_GRID
    assert tuple(lw.medium.phase_g_grid.t.shape) == (7, 7, 7, 1)
    const = _recipe(torch.tensor([0.55], device="cuda"), sigma)
    assert const._lower_map()[1].medium.phase_g_kind == L.VR_FIELD_CONST
    images = []
    for field in (flat, const, _recipe(torch.from_numpy(phase_field(7, 3)).cuda(), sigma)):
        rv.seed(5)
        with torch.no_grad():
            images.append(camera.capture(field, fw_samples=2).reshape(-1, 3).cpu().numpy())
    assert match_fraction(images[0], images[1], rtol=1e-4, atol=1e-6) >= 0.99
    assert match_fraction(images[2], images[1], rtol=1e-3, atol=1e-4) < 0.9
    # backward: the density receives a gradient, the phase field none
    s = torch.nn.Parameter(sigma.clone())
    g = torch.nn.Parameter(torch.from_numpy(phase_field(7, 3)).cuda())
    img = camera.capture(_recipe(g, s), fw_samples=2)
    img.sum().backward()
    assert s.grad is not None and float(s.grad.abs().sum()) > 0.0
    assert g.grad is None or float(g.grad.abs().sum()) == 0.0
