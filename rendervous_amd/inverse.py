"""Differentiable inverse rendering of a density grid.

Renders a set of views of a single-scattering medium with the reference's
collision-integrator graph (MCCollisionIntegrator over a delta-tracking
collision sampler, exitance = SingleScatteredRadiance with ratio-tracking
shadow rays; _maps.py:2714-2808, _objects.py:247-254,445-452), compares them
with target images under an L2 loss and updates the density with Adam.  The
gradient comes from the backward kernels of the capture (replaying the
forward's random streams from its dispatch seeds); the Adam update and the
gradient reset are one fused kernel (vr_adam_step).
"""
from __future__ import annotations

import dataclasses
import math
import typing

import torch
from torch.autograd.graph import increment_version

from . import _internal
from . import _lib as L
from . import maps as M
from . import objects as O


def single_scatter_scene(sigma: torch.Tensor, *, albedo=(0.9, 0.9, 0.9), g: float = 0.0, env=(1.0, 1.0, 1.0),
                         majorant: typing.Union[str, float] = "grid", majorant_resolution: int = 16,
                         ds_epsilon: float = 0.1):
    """DependencySet with radiance = MCCollisionIntegrator(DT collision sampler,
    SingleScatteredRadiance(RT transmittance), constant environment).

    majorant: "grid" (coarse max grid), "const" (max of sigma) or a number (a
    fixed constant majorant, which must bound sigma)."""
    ds = O.DependencySet()
    ds.add_parameters(sigma_tensor=sigma, scattering_albedo_tensor=torch.tensor(list(albedo), dtype=torch.float32),
                      environment_tensor=torch.tensor(list(env), dtype=torch.float32),
                      phase_g_tensor=torch.tensor([g], dtype=torch.float32))
    if majorant == "grid":
        ds.requires(O.medium_majorant_coarse_grid, resolution=majorant_resolution)
    elif not isinstance(majorant, str):
        ds.add_parameters(majorant_tensor=torch.tensor([float(majorant), 1000000.0], dtype=torch.float32))
    ds.requires(O.medium_phase_HG, mixture=False)
    ds.requires(O.medium_transmittance_RT)
    ds.requires(O.medium_collision_sampler_DT, ds_epsilon=ds_epsilon)
    ds.requires(O.medium_radiance_collision_integrator_MC)
    return ds


class FusedAdam:
    """Adam over one float32 tensor with the gradient cleared in the same pass
    (torch.optim.Adam semantics: no weight decay, bias-corrected)."""

    def __init__(self, p: torch.Tensor, lr: float = 1e-2, betas=(0.9, 0.999), eps: float = 1e-8):
        assert p.dtype == torch.float32 and p.is_contiguous()
        self.p = p
        self.m = torch.zeros_like(p)
        self.v = torch.zeros_like(p)
        self.lr, self.b1, self.b2, self.eps = lr, betas[0], betas[1], eps
        self.t = 0

    def step(self, grad: torch.Tensor, zero_grad: bool = True, bounds: typing.Optional[typing.Tuple[float, float]] = None):
        """One update; bounds = (lo, hi) projects the parameters onto [lo, hi] in the same pass."""
        assert grad.is_contiguous() and grad.shape == self.p.shape
        self.t += 1
        lo, hi = (-math.inf, math.inf) if bounds is None else (float(bounds[0]), float(bounds[1]))
        L.check(L.load().vr_adam_step_bounded(self.p.data_ptr(), grad.data_ptr(), self.m.data_ptr(), self.v.data_ptr(),
                                              self.p.numel(), self.lr, self.b1, self.b2, self.eps, self.t, lo, hi,
                                              int(zero_grad), L.stream_handle(self.p.device)), "vr_adam_step_bounded")
        # the kernel writes p and grad behind torch's back: bump their version
        # counters so caches keyed on them (the majorant grid) are rebuilt
        increment_version(self.p)
        if zero_grad:
            increment_version(grad)


@dataclasses.dataclass
class InverseStats:
    loss: float
    scatters: float  # trilinear scatters of the backward pass (8 corner updates each)


def render(camera: M.PerspectiveCameraSensor, radiance: M.MapBase, spp: int) -> torch.Tensor:
    return camera.capture(radiance, fw_samples=spp, bw_samples=spp)


def optimisation_step(camera, radiance, sigma_param: torch.nn.Parameter, target: torch.Tensor, spp: int,
                      opt: FusedAdam, clamp_min: float = 0.0, batch: typing.Optional[int] = None,
                      tv_weight: float = 0.0) -> torch.Tensor:
    """One Adam step of the L2 image loss (plus tv_weight * mean total variation of
    the density); returns the loss as a 0-d tensor (no host sync).  batch = B
    renders B random sensors (camera.random_sensors) against the matching target
    pixels instead of every view, as the reference's minibatched inverse renders."""
    if batch is None:
        img = render(camera, radiance, spp)
        ref = target
    else:
        sensors = camera.random_sensors(batch)
        img = camera.capture(radiance, sensors, fw_samples=spp, bw_samples=spp)
        ref = target[sensors[:, 0], sensors[:, 1], sensors[:, 2]]
    loss = ((img - ref) ** 2).mean()
    if tv_weight:
        from .functions import total_variation
        loss = loss + tv_weight * total_variation(sigma_param).mean()
    loss.backward()
    opt.step(sigma_param.grad, zero_grad=False, bounds=None if clamp_min is None else (clamp_min, math.inf))
    sigma_param.grad = None  # the next backward's fresh gradient becomes .grad without an extra add
    return loss.detach()
