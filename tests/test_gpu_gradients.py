"""Gradients of the HIP backward kernels against central finite differences.

Each check compares the capture backward (through the map API and torch
autograd, i.e. AutogradCaptureFunction -> vr_capture_bw) with a central
difference of the expected loss.  Both sides of a difference are rendered with
the same dispatch seeds (common random numbers) and the same fixed constant
majorant (a majorant taken from max(sigma) would move with the perturbation
and change every free flight): a perturbation of one voxel changes only the
paths that reach it, so the difference keeps little of the Monte Carlo noise
of either render.  The loss is a fixed linear functional of
the image, L = sum(w * I), so dL/dI = w exactly.
"""
import math

import numpy as np
import pytest
import torch

import rendervous_amd as rv
from rendervous_amd import inverse

pytestmark = pytest.mark.gpu


def _grid(n, seed, lo=1.0, hi=4.0):
    g = torch.Generator().manual_seed(seed)
    return (lo + (hi - lo) * torch.rand(n, n, n, 1, generator=g)).contiguous()


def _loss(camera, field, weights, spp, seed):
    rv.seed(seed)
    with torch.no_grad():
        img = camera.capture(field, fw_samples=spp)
    return float((img.double() * weights.double()).sum())


def _fd_check(make_field, sigma0, voxels, spp_grad, spp_fd, h, rel_tol, cam_res=16, jittered=True):
    dev = torch.device("cuda", 0)
    poses = torch.stack([rv.functions.look_at_pose((0.3, 0.4, 2.0)), rv.functions.look_at_pose((-1.8, 0.2, 0.9))]).to(dev)
    camera = rv.PerspectiveCameraSensor(cam_res, cam_res, poses, jittered=jittered)
    weights = torch.rand(2, cam_res, cam_res, 3, generator=torch.Generator().manual_seed(5)).to(dev) + 0.5
    sigma = torch.nn.Parameter(sigma0.clone().to(dev), requires_grad=True)
    field = make_field(sigma)
    rv.seed(123)
    img = camera.capture(field, fw_samples=spp_grad, bw_samples=spp_grad)
    (img * weights).sum().backward()
    grad = sigma.grad.detach().cpu()
    for k, v in enumerate(voxels):
        results = []
        for sgn in (+1, -1):
            s2 = sigma0.clone()
            s2[v] += sgn * h
            f2 = make_field(torch.nn.Parameter(s2.to(dev), requires_grad=False))
            results.append(_loss(camera, f2, weights, spp_fd, 1000 + k))
        fd = (results[0] - results[1]) / (2 * h)
        g = float(grad[v])
        print(f"[fd] voxel {v}: adjoint {g:.5f} fd {fd:.5f}", flush=True)
        assert abs(g - fd) <= rel_tol * abs(fd) + 0.05 * float(grad.abs().max()), (v, g, fd)


def test_mc_collision_single_scatter_gradient_matches_fd(gpu):
    """The inverse-rendering graph: DT collision sampler + single scatter with RT shadow rays."""
    sigma0 = _grid(6, seed=1)

    def make(sig):
        ds = inverse.single_scatter_scene(sig, albedo=(0.8, 0.8, 0.8), g=0.3, env=(1.0, 0.8, 0.6), majorant=5.0)
        return ds.radiance

    with rv.backward_mode(seeds="replay", estimator="consistent"):
        _fd_check(make, sigma0, [(2, 3, 3, 0), (3, 2, 2, 0), (1, 4, 2, 0)], spp_grad=4096, spp_fd=16384, h=0.4,
                  rel_tol=0.15)


def test_rt_transmittance_gradient_matches_fd(gpu):
    sigma0 = _grid(6, seed=2)

    def make(sig):
        b = rv.RayBoxIntersection(*rv.normalized_box(sig))
        g = rv.Grid3D(sig, *rv.normalized_box(sig))
        return rv.RatiotrackingTransmittance(g, b, rv.const[6.0, 1e6])

    with rv.backward_mode(seeds="replay", estimator="consistent"):
        _fd_check(make, sigma0, [(2, 3, 3, 0), (3, 2, 2, 0)], spp_grad=2048, spp_fd=8192, h=0.4, rel_tol=0.1)


def test_nee_path_gradient_matches_fd(gpu):
    """Multiple scattering (NEE + delta tracking), backward by the defensive
    replay estimator with matched primal/replay majorants."""
    sigma0 = _grid(6, seed=3, lo=1.0, hi=3.0)

    def make(sig):
        ds = rv.objects.DependencySet()
        ds.add_parameters(sigma_tensor=sig, scattering_albedo_tensor=torch.tensor([0.8, 0.8, 0.8]),
                          environment_tensor=torch.tensor([1.0, 1.0, 1.0]), phase_g_tensor=torch.tensor([0.3]),
                          majorant_tensor=torch.tensor([4.0, 1e6]))
        ds.requires(rv.objects.medium_phase_HG, mixture=False)
        ds.requires(rv.objects.medium_phase_sampler_HG, mixture=False)
        ds.requires(rv.objects.medium_transmittance_DT)
        ds.requires(rv.objects.medium_radiance_path_integrator_NEE_DT, max_bounces=8)
        return ds.radiance

    with rv.backward_mode(seeds="replay", estimator="consistent"):
        _fd_check(make, sigma0, [(2, 3, 3, 0), (3, 2, 2, 0)], spp_grad=4096, spp_fd=16384, h=0.4, rel_tol=0.2)


def test_drt_path_gradient_matches_fd(gpu):
    """NEE forward, reservoir (DRT) backward."""
    sigma0 = _grid(6, seed=3, lo=1.0, hi=3.0)

    def make(sig):
        ds = rv.objects.DependencySet()
        ds.add_parameters(sigma_tensor=sig, scattering_albedo_tensor=torch.tensor([0.8, 0.8, 0.8]),
                          environment_tensor=torch.tensor([1.0, 1.0, 1.0]), phase_g_tensor=torch.tensor([0.3]),
                          majorant_tensor=torch.tensor([4.0, 1e6]))
        ds.requires(rv.objects.medium_phase_HG, mixture=False)
        ds.requires(rv.objects.medium_phase_sampler_HG, mixture=False)
        ds.requires(rv.objects.medium_transmittance_RT)
        ds.requires(rv.objects.medium_radiance_path_integrator_NEE_DRT, max_bounces=8)
        return ds.radiance

    _fd_check(make, sigma0, [(2, 3, 3, 0), (3, 2, 2, 0)], spp_grad=8192, spp_fd=16384, h=0.4, rel_tol=0.2)


def test_dt_path_gradient_matches_fd(gpu):
    """Delta-tracking path tracer without NEE, defensive replay backward."""
    sigma0 = _grid(6, seed=5, lo=1.0, hi=3.0)

    def make(sig):
        ds = rv.objects.DependencySet()
        ds.add_parameters(sigma_tensor=sig, scattering_albedo_tensor=torch.tensor([0.8, 0.8, 0.8]),
                          environment_tensor=torch.tensor([1.0, 0.9, 0.8]), phase_g_tensor=torch.tensor([0.3]),
                          majorant_tensor=torch.tensor([4.0, 1e6]))
        ds.requires(rv.objects.medium_phase_sampler_HG, mixture=False)
        ds.requires(rv.objects.medium_radiance_path_integrator_DT, max_bounces=8)
        return ds.radiance

    _fd_check(make, sigma0, [(2, 3, 3, 0), (3, 2, 2, 0)], spp_grad=8192, spp_fd=16384, h=0.4, rel_tol=0.2)


def test_dda_gradient_matches_fd_exactly(gpu):
    """Deterministic DDA: adjoint vs FD to 1e-3 relative."""
    sigma0 = _grid(8, seed=4)

    def make(sig):
        b = rv.RayBoxIntersection(*rv.normalized_box(sig))
        return rv.GridDDATransmittance(rv.Grid3D(sig, *rv.normalized_box(sig)), b)

    _fd_check(make, sigma0, [(3, 3, 4, 0), (4, 2, 5, 0), (2, 5, 3, 0)], spp_grad=1, spp_fd=1, h=1e-2, rel_tol=1e-2,
              jittered=False)


def test_inverse_render_reduces_loss(gpu):
    """Adam on a 32^3 grid from constant 5 towards an fBm target: the image
    error, measured at 256 spp before and after 30 steps at 8 spp, drops."""
    dev = torch.device("cuda", 0)
    target_sigma = rv.functions.fbm_value_noise(32, seed=3, scale=20.0, device=dev)
    poses = rv.functions.orbit_poses(4, radius=2.5).to(dev)
    camera = rv.PerspectiveCameraSensor(48, 48, poses, jittered=True)
    target_field = inverse.single_scatter_scene(target_sigma).radiance
    rv.seed(1)
    with torch.no_grad():
        target = camera.capture(target_field, fw_samples=1024)
    sigma = torch.nn.Parameter(torch.full((32, 32, 32, 1), 5.0, device=dev), requires_grad=True)
    field = inverse.single_scatter_scene(sigma).radiance

    def image_error():
        rv.seed(2)
        with torch.no_grad():
            return float(((camera.capture(field, fw_samples=256) - target) ** 2).mean())

    before = image_error()
    opt = inverse.FusedAdam(sigma, lr=0.3)
    losses = [float(inverse.optimisation_step(camera, field, sigma, target, 8, opt)) for _ in range(30)]
    after = image_error()
    print(f"[inverse] image error {before:.5f} -> {after:.5f}; step losses {losses[0]:.5f} .. {losses[-1]:.5f}",
          flush=True)
    assert after < 0.7 * before


def test_minibatched_inverse_render_with_tv_reduces_loss(gpu):
    """Random-sensor minibatches (camera.random_sensors, the reference's
    _maps.py:1751-1762) and a total-variation term on the density: the full-image
    error still drops, and the TV of the estimate stays below the unregularised run's."""
    dev = torch.device("cuda", 0)
    target_sigma = rv.functions.fbm_value_noise(32, seed=3, scale=20.0, device=dev)
    poses = rv.functions.orbit_poses(4, radius=2.5).to(dev)
    camera = rv.PerspectiveCameraSensor(48, 48, poses, jittered=True)
    rv.seed(1)
    with torch.no_grad():
        target = camera.capture(inverse.single_scatter_scene(target_sigma).radiance, fw_samples=1024)
    tvs = []
    for tv_weight in (0.0, 0.005):
        sigma = torch.nn.Parameter(torch.full((32, 32, 32, 1), 5.0, device=dev), requires_grad=True)
        field = inverse.single_scatter_scene(sigma).radiance

        def image_error():
            rv.seed(2)
            with torch.no_grad():
                return float(((camera.capture(field, fw_samples=256) - target) ** 2).mean())

        before = image_error()
        opt = inverse.FusedAdam(sigma, lr=0.3)
        rv.seed(5)
        for _ in range(40):
            inverse.optimisation_step(camera, field, sigma, target, 8, opt, batch=2048, tv_weight=tv_weight)
        after = image_error()
        tvs.append(float(rv.functions.total_variation(sigma.detach()).mean()))
        print(f"[inverse-mb] tv_weight {tv_weight}: image error {before:.5f} -> {after:.5f}, mean TV {tvs[-1]:.3f}",
              flush=True)
        assert after < 0.85 * before
    assert tvs[1] < 0.5 * tvs[0]


def test_l2_loss_gradient_fresh_seeds_are_unbiased_replayed_seeds_are_not(gpu):
    """The L2 image loss L = sum((I_hat - T)^2) is quadratic in the estimate: a
    backward that replays the forward's samples estimates
    2 E[(I_hat - T) G] = d (E I_hat - T)^2 + 2 Cov(I_hat, G), a backward with
    fresh seeds (the reference's, the default) 2 (E I_hat - T) E[G] = d J with
    J = sum((E I_hat - T)^2).  At 1 spp, along the uniform direction of the
    density: the mean fresh gradient equals the central difference of J (E
    I_hat from 32768-spp renders with the same seeds on both sides), the mean
    replayed one does not."""
    dev = torch.device("cuda", 0)
    sigma0 = _grid(6, seed=4, lo=1.0, hi=3.0)
    poses = torch.stack([rv.functions.look_at_pose((0.3, 0.4, 2.0)), rv.functions.look_at_pose((-1.8, 0.2, 0.9))]).to(dev)
    camera = rv.PerspectiveCameraSensor(16, 16, poses, jittered=True)

    def make(sig):
        return inverse.single_scatter_scene(sig, albedo=(0.8, 0.8, 0.8), g=0.3, env=(1.0, 0.8, 0.6), majorant=5.0).radiance

    target = torch.full((2, 16, 16, 3), 0.1, device=dev, dtype=torch.float64)
    h = 0.05
    J = []
    for sgn in (1.0, -1.0):
        f = make(torch.nn.Parameter((sigma0 + sgn * h).to(dev), requires_grad=False))
        rv.seed(77)
        with torch.no_grad():
            img = camera.capture(f, fw_samples=32768).double()
        J.append(float(((img - target) ** 2).sum()))
    fd = (J[0] - J[1]) / (2 * h)
    sigma = torch.nn.Parameter(sigma0.clone().to(dev), requires_grad=True)
    field = make(sigma)
    means, ses = {}, {}
    for mode in ("fresh", "reference", "replay"):
        vals = []
        with rv.backward_mode(seeds=mode):
            for b in range(1500):
                rv.seed(5000 + b)
                img = camera.capture(field, fw_samples=1, bw_samples=1)
                ((img.double() - target) ** 2).sum().backward()
                vals.append(sigma.grad.sum())
                sigma.grad = None
        v = torch.stack(vals).double().cpu()
        means[mode], ses[mode] = float(v.mean()), float(v.std() / len(vals) ** 0.5)
    print(f"[l2 bias] fd of J {fd:.4f}; " + "; ".join(f"{m} {means[m]:.4f} +- {ses[m]:.4f}" for m in means), flush=True)
    # fresh seeds: per-sample streams (the default) or the reference's per-sensor stream
    for m in ("fresh", "reference"):
        assert abs(means[m] - fd) <= 4.0 * ses[m] + 0.02 * abs(fd), m
    assert abs(means["replay"] - fd) >= 4.0 * ses["replay"]
