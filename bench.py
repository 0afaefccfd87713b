"""Benchmark of the participating-media hot path on one MI355X.

Default workload (one step = one full frame):
  NEE volume path tracer with delta tracking over a 256^3 FP32 grid of 4-octave
  fBm value noise (lacunarity 2, gain 0.5) clamped to [0,1]*40 inside a sphere,
  albedo 0.9, Henyey-Greenstein g=0.6, one directional light + constant
  environment 0.1, 16^3 majorant grid, 512x512 pixels, 64 spp, max 16 bounces,
  RGB FP32 output.

Prints ONE JSON line with the metric (camera samples per second), the roofline
of the dominant kernel (gather model: 32 B per trilinear lookup + 12 B per
output pixel, lookups counted by the kernel's own per-thread counter) and a CPU
baseline (the oracle on a bounded pixel sample of the same frame).
"""
from __future__ import annotations

import argparse
import gc
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


# HBM traffic of one launch of the persistent NEE kernel at the default workload
# (256^3, 512x512, 64 spp, density read through its cell table), from rocprofv3
# --pmc FETCH_SIZE and WRITE_SIZE in KiB per dispatch, each scaled by what the
# counter reads for a 256 MiB device copy (FETCH_SIZE reads 128 MiB for it,
# WRITE_SIZE 256 MiB: factors 2 and 1).  The raw figures and the commands are in
# profiles/nee_persistent_pmc_256cube_512x512x64spp.txt.
NEE_PROFILED = {"workload": (256, 512, 512, 64), "fetch_kib": 5.4066e6, "write_kib": 7.2189e5,
                "fetch_scale": 2.0, "write_scale": 1.0}


def profiled_traffic(res, width, height, spp):
    p = NEE_PROFILED
    if p["workload"] != (res, width, height, spp) or p["fetch_kib"] is None:
        return None
    return 1024.0 * (p["fetch_kib"] * p["fetch_scale"] + p["write_kib"] * p["write_scale"])


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_scene(res: int, width: int, height: int, device):
    import torch
    import rendervous_amd as rv

    sigma = rv.functions.fbm_value_noise(res, seed=7, octaves=4, lacunarity=2.0, gain=0.5, base_frequency=4.0,
                                         scale=40.0, sphere_radius=0.5, device=device)
    ds = rv.objects.DependencySet()
    ds.add_parameters(sigma_tensor=sigma, scattering_albedo_tensor=torch.tensor([0.9, 0.9, 0.9]),
                      environment_tensor=torch.tensor([0.1, 0.1, 0.1]), phase_g_tensor=torch.tensor([0.6]))
    ds.requires(rv.objects.medium_majorant_coarse_grid, resolution=16)
    ds.requires(rv.objects.medium_environment_sampler_sun, direction=(0.4, 0.8, 0.45), radiance=(3.0, 3.0, 3.0),
                probability=0.5)
    ds.requires(rv.objects.medium_phase_HG, mixture=False)
    ds.requires(rv.objects.medium_phase_sampler_HG, mixture=False)
    ds.requires(rv.objects.medium_transmittance_DT)
    ds.requires(rv.objects.medium_radiance_path_integrator_NEE_DT, max_bounces=16)
    poses = rv.functions.look_at_pose((0.0, 0.3, 2.5))[None].to(device)
    camera = rv.PerspectiveCameraSensor(width, height, poses, jittered=True)
    return sigma, ds.radiance, camera


def cpu_baseline(sigma_cpu, radiance, width, height, spp, budget_s: float = 15.0):
    """Oracle (C, OpenMP) on every k-th row of the same frame; returns samples/s."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from helpers import Scene, oracle_capture_fw, pose, L

    m = radiance._lower_map()[1].medium
    sc = Scene(sigma=sigma_cpu.numpy(), albedo_c=tuple(m.albedo_c), env_c=tuple(m.env_c), sun_dir=tuple(m.sun_dir),
               sun_radiance=tuple(m.sun_radiance), sun_prob=m.sun_prob, phase_kind=m.phase_kind, phase_g=m.phase_g,
               majorant="grid", majorant_res=(16, 16, 16), transmittance_kind=m.transmittance_kind,
               max_bounces=m.max_bounces)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    # bounded sample: full rows spread evenly over the frame, doubled until one
    # run takes at least a quarter of the budget (or the whole frame is used)
    rows = 1
    done_samples, elapsed = 0, 0.0
    p = pose((0.0, 0.3, 2.5))[None]
    while True:
        ys = np.unique(np.linspace(0, height - 1, rows).astype(np.int64))
        sensors = np.stack(np.meshgrid(np.zeros(1, np.int64), ys, np.arange(width, dtype=np.int64), indexing="ij"),
                           -1).reshape(-1, 3)
        t0 = time.perf_counter()
        oracle_capture_fw(sc, L.VR_MAP_NEE_DT, dict(poses=p, width=width, height=height, jittered=True,
                                                     sensors=sensors), (11, 22, 33, 44), spp)
        dt = time.perf_counter() - t0
        done_samples, elapsed, rows = sensors.shape[0] * spp, dt, len(ys)
        if dt > budget_s / 4 or rows >= height:
            break
        rows = min(height, rows * 2)
    return {"value": done_samples / elapsed, "unit": "camera samples/s", "cores": threads, "kind": "port",
            "sample": f"{done_samples // spp} pixels ({rows} rows x {width}) of the same frame at {spp} spp, "
                      f"{elapsed:.1f} s, C oracle with OpenMP"}


def _events(n):
    import torch
    return [torch.cuda.Event(enable_timing=True) for _ in range(n)]


def _inverse_step_timing(camera, field, target, spp, params, steps):
    """Times `steps` inverse-render steps (after one warm-up): capture forward +
    L2 loss, adjoint backward, fused bounded Adam on each (param, opt, bounds).
    An untimed extra step counts the forward's and the backward's trilinear
    lookups and the backward's scatters from the kernels' own counters.  Means
    over the timed steps, with each step's forward and backward times and the
    host time of the forward / backward calls (launch work, not completion)."""
    import torch
    import rendervous_amd as rv

    dev = target.device
    n_sensors = camera.num_sensors()
    dbg = torch.zeros(n_sensors, dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream(dev)

    def step(debug=None):
        img = camera.capture(field, fw_samples=spp, bw_samples=spp, debug_out=debug)
        loss = ((img - target) ** 2).mean()
        return img, loss

    def update():
        for p, opt, bounds in params:
            opt.step(p.grad, zero_grad=False, bounds=bounds)
            p.grad = None

    fw_ms = bw_ms = adam_ms = 0.0
    fw_steps, bw_steps = [], []
    fw_host = bw_host = 0.0  # host time of the forward / backward calls (launches, not completion)
    scatters = 0.0
    gc.collect()
    enabled = gc.isenabled()
    gc.disable()  # no collector pause inside a timed step (a stalled launch idles the GPU)
    for k in range(steps + 1):  # the first step is a warm-up
        e = _events(4)
        e[0].record(stream)
        h0 = time.perf_counter()
        img, loss = step(dbg)
        h1 = time.perf_counter()
        e[1].record(stream)
        loss.backward()
        h2 = time.perf_counter()
        e[2].record(stream)
        update()
        e[3].record(stream)
        torch.cuda.synchronize()
        if k == 0:
            continue
        fw_steps.append(e[0].elapsed_time(e[1]))
        bw_steps.append(e[1].elapsed_time(e[2]))
        fw_ms += fw_steps[-1]
        bw_ms += bw_steps[-1]
        adam_ms += e[2].elapsed_time(e[3])
        fw_host += (h1 - h0) * 1e3
        bw_host += (h2 - h1) * 1e3
        scatters += float(dbg.to(torch.float64).sum().item())
    if enabled:
        gc.enable()
    fw_ms, bw_ms, adam_ms, scatters = fw_ms / steps, bw_ms / steps, adam_ms / steps, scatters / steps
    # lookup counts (untimed; the gradients of this step are dropped)
    with rv.backward_mode(debug="lookups"):
        img, loss = step(dbg)
        torch.cuda.synchronize()
        fw_lookups = float(dbg.to(torch.float64).sum().item())
        loss.backward()
        torch.cuda.synchronize()
        bw_lookups = float(dbg.to(torch.float64).sum().item())
    for p, _, _ in params:
        p.grad = None
    voxels = float(sum(p.numel() for p, _, _ in params))
    out_dim = img.shape[-1]
    # gather model: 32 B per lookup, 64 B per scatter, the image written and its
    # gradient read (4 B x OUT per pixel each), 4 B per voxel to zero the fresh
    # gradient and 28 B per voxel for Adam (p, g, m, v read; p, m, v written)
    bytes_step = 32.0 * (fw_lookups + bw_lookups) + 64.0 * scatters + 8.0 * out_dim * n_sensors + 32.0 * voxels
    t = (fw_ms + bw_ms + adam_ms) * 1e-3
    return {
        "grad_voxel_updates_per_s": 8.0 * scatters / (bw_ms * 1e-3),
        "trilinear_scatters_per_step": scatters, "fw_lookups_per_step": fw_lookups, "bw_lookups_per_step": bw_lookups,
        "fw_ms": fw_ms, "bw_ms": bw_ms, "adam_ms": adam_ms,
        "fw_host_ms": fw_host / steps, "bw_host_ms": bw_host / steps, "fw_ms_steps": fw_steps, "bw_ms_steps": bw_steps,
        "roofline": {"bound": "hbm", "achieved": bytes_step / t / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": bytes_step / t / 1e9 / HBM_PEAK_GBS, "bytes_per_step": bytes_step},
        "loss": float(loss.item()),
    }


def gradient_leg(dev, res=256, views=8, hw=256, spp=16, steps=3):
    """Inverse-render step of the 256^3 single-scatter configuration: forward
    capture (DT collision + single scatter with RT shadow rays), L2 loss,
    adjoint backward replaying the forward's seeds, fused Adam.  Reports grad
    voxel-updates/s = 8 x trilinear scatters / backward time."""
    import torch
    import rendervous_amd as rv
    from rendervous_amd import inverse

    target_sigma = rv.functions.fbm_value_noise(res, seed=11, scale=40.0, device=dev)
    poses = rv.functions.orbit_poses(views, radius=2.5).to(dev)
    camera = rv.PerspectiveCameraSensor(hw, hw, poses, jittered=True)
    rv.seed(3)
    with torch.no_grad():
        target = camera.capture(inverse.single_scatter_scene(target_sigma).radiance, fw_samples=spp)
    del target_sigma
    sigma = torch.nn.Parameter(torch.full((res, res, res, 1), 5.0, device=dev), requires_grad=True)
    field = inverse.single_scatter_scene(sigma).radiance
    r = _inverse_step_timing(camera, field, target, spp, [(sigma, inverse.FusedAdam(sigma, lr=1e-2), (0.0, math.inf))],
                             steps)
    samples = views * hw * hw * spp
    r["workload"] = (f"inverse render step: {res}^3 density, {views} views {hw}x{hw}, {spp} spp, DT collision + "
                     f"single scatter (RT shadow rays), L2 loss, adjoint backward, fused Adam")
    r["fw_bw_samples_per_s"] = 2.0 * samples / ((r["fw_ms"] + r["bw_ms"]) * 1e-3)
    r["roofline"]["measured_bound"] = (
        "VALU issue: forward and backward kernels 75 % VALU-busy at 29 % / 17 % active lanes, ~780 / ~950 "
        "lane-instructions per sample for 0.5 / 1.1 lookups; at the device's VALU peak (3.9e13 lane-instructions/s) "
        "this work alone takes 0.37 ms, so the step stays <= 29 % of HBM "
        "(profiles/inverse_step_256cube_8views_256x256x16spp.txt)")
    return r


def _directional_check(camera, field_of, param, spp, n_dirs=4, h=0.25, batches=4, seed=0):
    """Adjoint gradient of sum(w * image) (replayed seeds, consistent estimator)
    against central differences of the same functional along smooth positive
    directions v = 0.5 + |a random 8^3 field resampled to the grid|, both sides
    of a difference rendered with the same dispatch seeds.  Each of `batches`
    independent seed batches gives one adjoint and one difference per
    direction; returns per direction the means, their standard errors and the
    discrepancy in standard errors."""
    import torch
    import rendervous_amd as rv

    dev = param.device
    gen = torch.Generator(device="cpu").manual_seed(seed)
    w = (torch.rand(*camera.index_shape, 3, generator=gen) + 0.5).to(dev)
    dirs = []
    for _ in range(n_dirs):
        coarse = torch.randn(8, 8, 8, param.shape[-1], generator=gen).abs().to(dev)
        dirs.append((0.5 + rv.functions.resample_grid(coarse, tuple(param.shape[:-1]))).double())
    ad = [[] for _ in dirs]
    fd = [[] for _ in dirs]
    fd2 = [[] for _ in dirs]  # step h / 2, for Richardson extrapolation
    for b in range(batches):
        rv.seed(1000 + b)
        img = camera.capture(field_of(param), fw_samples=spp, bw_samples=spp)
        (img * w).sum().backward()
        g = param.grad.detach().double()
        param.grad = None
        for k, v in enumerate(dirs):
            ad[k].append(float((g * v).sum()))
            for step, acc in ((h, fd[k]), (0.5 * h, fd2[k])):
                vals = []
                for sgn in (1.0, -1.0):
                    with torch.no_grad():
                        p2 = (param.detach() + sgn * step * v.float()).contiguous()
                        rv.seed(1000 + b)
                        vals.append(float((camera.capture(field_of(p2), fw_samples=spp) * w).double().sum()))
                acc.append((vals[0] - vals[1]) / (2 * step))
    out = []
    for k in range(n_dirs):
        a = torch.tensor(ad[k], dtype=torch.float64)
        f = torch.tensor(fd[k], dtype=torch.float64)
        f2 = torch.tensor(fd2[k], dtype=torch.float64)
        # central differences err by O(h^2): (4 D(h/2) - D(h)) / 3 removes that term
        fr = (4.0 * f2 - f) / 3.0
        sa, sf, sr = (float(x.std() / batches ** 0.5) for x in (a, f, fr))
        ma, mf, mr = float(a.mean()), float(f.mean()), float(fr.mean())
        out.append({"fd": mf, "fd_se": sf, "fd_half_step": float(f2.mean()), "fd_richardson": mr, "fd_richardson_se": sr,
                    "adjoint": ma, "adjoint_se": sa,
                    "rel_err": abs(ma - mf) / max(abs(mf), 1e-30),
                    "z": abs(ma - mf) / max((sa * sa + sf * sf) ** 0.5, 1e-30),
                    "rel_err_richardson": abs(ma - mr) / max(abs(mr), 1e-30),
                    "z_richardson": abs(ma - mr) / max((sa * sa + sr * sr) ** 0.5, 1e-30)})
    return out


def gradient_gate_leg(dev, res=256, views=8, hw=256, spp=16, voxels=64, h=1e-3, batches=256, ref_spp=1024):
    """The gradient check of the 256^3 inverse-render configuration: central
    finite differences (h = 1e-3) on 64 random voxels against the adjoint
    gradient, at the initial density 5 (8 views of the seeded fBm target,
    256x256, 16 spp, single scatter with RT shadow rays, fixed constant majorant
    7 so that both sides of a difference draw the same random numbers).

    The checked quantity is the L2 loss linearised at the current density,
    Lambda = sum(w * I_hat) with w = dL/dI = 2 (I - T) / N and I a 1024-spp
    render: its derivative is the loss gradient, and for a linear functional the
    differences (same samples on both sides) and the adjoint estimate the same
    thing.  For the L2 loss itself they do not: the differences of
    (I_hat - T)^2 estimate d E[(I_hat - T)^2] = d (E[I_hat] - T)^2 + d Var(I_hat),
    the fresh-seed adjoint (the reference's) d (E[I_hat] - T)^2 and the
    replayed-seed adjoint adds 2 Cov(I_hat, gradient estimate); those three are
    reported beside the gate as the measured variance terms.  Each of `batches`
    dispatch seeds gives one adjoint and one difference per voxel; z = (mean
    adjoint - mean difference) / the combined standard error.  Pixels are
    rendered by one thread each (summation order fixed) and differences are
    formed in float64 from the two images, so unchanged pixels cancel exactly."""
    import torch
    import rendervous_amd as rv
    from rendervous_amd import inverse

    old_chunk = os.environ.get("VR_CAPTURE_CHUNK")
    os.environ["VR_CAPTURE_CHUNK"] = str(spp)
    try:
        target_sigma = rv.functions.fbm_value_noise(res, seed=11, scale=40.0, device=dev)
        poses = rv.functions.orbit_poses(views, radius=2.5).to(dev)
        camera = rv.PerspectiveCameraSensor(hw, hw, poses, jittered=True)
        rv.seed(3)
        with torch.no_grad():
            target = camera.capture(inverse.single_scatter_scene(target_sigma).radiance, fw_samples=spp)
        del target_sigma
        target64 = target.double()
        n_out = target.numel()
        sigma = torch.nn.Parameter(torch.full((res, res, res, 1), 5.0, device=dev), requires_grad=True)
        field = inverse.single_scatter_scene(sigma, majorant=7.0).radiance
        work = torch.full((res, res, res, 1), 5.0, device=dev)
        wfield = inverse.single_scatter_scene(work, majorant=7.0).radiance
        os.environ.pop("VR_CAPTURE_CHUNK", None)
        rv.seed(77)
        with torch.no_grad():
            ref_img = camera.capture(wfield, fw_samples=ref_spp)
        os.environ["VR_CAPTURE_CHUNK"] = str(spp)
        w = (2.0 * (ref_img - target) / n_out).contiguous()
        w64 = w.double()
        gen = torch.Generator(device="cpu").manual_seed(17)
        idx = torch.randint(8, res - 8, (voxels, 3), generator=gen)
        vox = [tuple(int(a) for a in r) + (0,) for r in idx]
        base = torch.tensor(5.0, dtype=torch.float32)
        up, dn = float(base + h), float(base - h)  # float32 values of the perturbed voxel
        h_eff = 0.5 * (up - dn)
        adj = {"linear": [], "l2_replay": [], "l2_fresh": []}
        fd_lin, fd_l2 = [], []
        t0 = time.perf_counter()
        for b in range(batches):
            for name, mode, lin in (("linear", "fresh", True), ("l2_replay", "replay", False), ("l2_fresh", "fresh", False)):
                rv.seed(2000 + b)
                with rv.backward_mode(seeds=mode):
                    img = camera.capture(field, fw_samples=spp, bw_samples=spp)
                    ((img * w).sum() if lin else ((img - target) ** 2).mean()).backward()
                g = sigma.grad.detach()
                adj[name].append([float(g[v]) for v in vox])
                sigma.grad = None
            rl, r2 = [], []
            for v in vox:
                imgs = []
                for val in (up, dn):
                    with torch.no_grad():
                        work[v] = val
                        rv.seed(2000 + b)
                        imgs.append(camera.capture(wfield, fw_samples=spp).double())
                with torch.no_grad():
                    work[v] = 5.0
                a, c = imgs
                rl.append(float(((a - c) * w64).sum()) / (2.0 * h_eff))
                r2.append(float(((a - c) * (a + c - 2.0 * target64)).sum()) / n_out / (2.0 * h_eff))
            fd_lin.append(rl)
            fd_l2.append(r2)
        wall = time.perf_counter() - t0
    finally:
        if old_chunk is None:
            os.environ.pop("VR_CAPTURE_CHUNK", None)
        else:
            os.environ["VR_CAPTURE_CHUNK"] = old_chunk
    sq = batches ** 0.5
    A = {m: torch.tensor(x, dtype=torch.float64) for m, x in adj.items()}
    F = {"linear": torch.tensor(fd_lin, dtype=torch.float64), "l2": torch.tensor(fd_l2, dtype=torch.float64)}

    def compare(a, f):
        ma, sa, mf, sf = a.mean(0), a.std(0) / sq, f.mean(0), f.std(0) / sq
        z = (ma - mf) / (sa * sa + sf * sf).sqrt().clamp_min(1e-300)
        za = z.abs()
        return z, {"abs_z_max": float(za.max()), "abs_z_median": float(za.median()),
                   "frac_abs_z_below_3": float((za < 3).double().mean()), "chi2_per_voxel": float((z * z).mean()),
                   "corr_of_means": float(torch.corrcoef(torch.stack([ma, mf]))[0, 1]),
                   "rel_err_of_sum": float((ma.sum() - mf.sum()).abs() / mf.abs().sum()),
                   "median_rel_se_fd": float((sf / mf.abs().clamp_min(1e-300)).median())}

    z_lin, gate = compare(A["linear"], F["linear"])
    _, rep = compare(A["l2_replay"], F["l2"])
    _, fre = compare(A["l2_fresh"], F["l2"])
    per_voxel = [{"voxel": list(v[:3]), "fd": float(F["linear"][:, k].mean()), "fd_se": float(F["linear"][:, k].std() / sq),
                  "adjoint": float(A["linear"][:, k].mean()), "adjoint_se": float(A["linear"][:, k].std() / sq),
                  "z": float(z_lin[k])} for k, v in enumerate(vox)]
    return {"workload": f"{res}^3 density 5.0, {views} views {hw}x{hw}, {spp} spp, single scatter (RT shadow rays); "
                        f"central differences h = {h} on {voxels} random voxels of the L2 loss linearised at a "
                        f"{ref_spp}-spp render, fixed majorant 7, {batches} seed batches", "h_float32": h_eff,
            "wall_s": wall, "gate": gate,
            "l2_loss_fd_vs_replay_adjoint": rep, "l2_loss_fd_vs_fresh_adjoint": fre,
            "l2_replay_minus_fresh_over_sum_abs": float((A["l2_replay"].mean(0) - A["l2_fresh"].mean(0)).abs().sum()
                                                       / A["l2_fresh"].mean(0).abs().sum()),
            "per_voxel": per_voxel}


def convergence_leg(dev, res=256, views=8, hw=256, spp=16, steps=50):
    """The 256^3 single-scatter inverse render run to completion: 50 Adam steps
    (lr 1e-2) from the constant density 5 towards 8 views of a seeded fBm target;
    the loss at every step and the wall time of the loop; then the adjoint
    gradient against central differences (h = 0.25, same seeds) along 4 smooth
    directions at the full grid size, with a fixed majorant so that both sides
    of each difference draw the same random numbers."""
    import torch
    import rendervous_amd as rv
    from rendervous_amd import inverse

    target_sigma = rv.functions.fbm_value_noise(res, seed=11, scale=40.0, device=dev)
    poses = rv.functions.orbit_poses(views, radius=2.5).to(dev)
    camera = rv.PerspectiveCameraSensor(hw, hw, poses, jittered=True)
    rv.seed(3)
    with torch.no_grad():
        target = camera.capture(inverse.single_scatter_scene(target_sigma).radiance, fw_samples=spp)
    del target_sigma
    sigma = torch.nn.Parameter(torch.full((res, res, res, 1), 5.0, device=dev), requires_grad=True)
    # gradient check first, at the initial density (fixed majorant 7 >= 5 + 0.25 max v)
    checks = _directional_check(camera, lambda s: inverse.single_scatter_scene(s, majorant=7.0).radiance, sigma, 64)
    field = inverse.single_scatter_scene(sigma).radiance
    init = sigma.detach().clone()
    with torch.no_grad():  # low-noise loss: 256-spp renders of target and estimate with fixed seeds
        rv.seed(99)
        target_hi = camera.capture(inverse.single_scatter_scene(
            rv.functions.fbm_value_noise(res, seed=11, scale=40.0, device=dev)).radiance, fw_samples=256)

        def eval_loss(f=None):
            with torch.no_grad():
                rv.seed(98)
                return float(((camera.capture(field if f is None else f, fw_samples=256) - target_hi) ** 2).mean())
        eval_first = eval_loss()
    runs = {}
    for mode in ("replay", "fresh"):  # backward seeds: the forward's (replay) or new ones (the reference's)
        with torch.no_grad():
            sigma.copy_(init)
        opt = inverse.FusedAdam(sigma, lr=1e-2)
        losses = []
        rv.seed(4)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        with rv.backward_mode(seeds=mode):
            for _ in range(steps):
                losses.append(inverse.optimisation_step(camera, field, sigma, target, spp, opt))
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        losses = [float(x) for x in losses]
        runs[mode] = {"loss_curve": losses, "loss_256spp_last": eval_loss(), "wall_s": wall,
                      "ms_per_step": wall * 1e3 / steps}
    # the same 50 steps coarse to fine (32^3 -> 64^3 -> 128^3 -> 256^3 with
    # resample_grid between stages), as the reference's inverse-rendering
    # examples do: each coarse voxel collects many gradient samples per step
    stages = [(32, 13), (64, 13), (128, 12), (res, 12)]
    cur = torch.full((stages[0][0],) * 3 + (1,), 5.0, device=dev)
    losses = []
    rv.seed(4)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k, (r, n) in enumerate(stages):
        if cur.shape[0] != r:
            cur = rv.functions.resample_grid(cur, (r, r, r))
        p = torch.nn.Parameter(cur.contiguous(), requires_grad=True)
        f = inverse.single_scatter_scene(p).radiance
        opt = inverse.FusedAdam(p, lr=1e-2)
        for _ in range(n):
            losses.append(inverse.optimisation_step(camera, f, p, target, spp, opt))
        cur = p.detach()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    runs["coarse_to_fine"] = {"loss_curve": [float(x) for x in losses], "loss_256spp_last": eval_loss(f),
                              "wall_s": wall, "ms_per_step": wall * 1e3 / len(losses)}
    return {"workload": f"{res}^3 density from 5.0, {views} views {hw}x{hw}, {spp} spp, single scatter (RT shadow rays), "
                        f"L2 loss, Adam lr 1e-2, {steps} steps (backward seeds replayed or fresh); gradient check at "
                        f"64 spp, 4 seed batches",
            "loss_256spp_first": eval_first, "runs": runs, "ms_per_step": runs["replay"]["ms_per_step"],
            "directional_fd": checks,
            "fd_rel_err_max": max(c["rel_err"] for c in checks), "fd_z_max": max(c["z"] for c in checks)}


def nee_convergence_leg(dev, res=512, ares=128, views=32, hw=512, spp=32, steps=100, check_views=8, check_spp=32):
    """The multiple-scattering inverse render at its stated size: a 512^3 density
    and a 128^3 RGB albedo optimised from 32 views at 512x512, 32 spp, NEE with
    delta tracking (max 8 bounces), backward by defensive path replay from the
    dispatch seeds (fresh seeds, no stored vertices), targets rendered from
    seeded fBm grids, Adam (lr 1e-2) for 100 steps.  Reports the training loss
    of every step, a low-noise loss (4 views at 256 spp against a 256-spp
    target) before and after, the step time split into forward, backward and
    Adam, and the adjoint against central differences (h and h/2 with Richardson
    extrapolation, same seeds on both sides, fixed majorant) along smooth
    directions of the density and of the albedo on 8 of the views."""
    import torch
    import rendervous_amd as rv
    from rendervous_amd import inverse

    t_sigma = rv.functions.fbm_value_noise(res, seed=21, scale=60.0, sphere_radius=0.5, device=dev)
    t_albedo = (0.55 + 0.4 * rv.functions.fbm_value_noise(ares, seed=22, scale=1.0, device=dev)).expand(
        ares, ares, ares, 3).contiguous()
    poses = rv.functions.orbit_poses(32, radius=2.5)[:views].to(dev)
    camera = rv.PerspectiveCameraSensor(hw, hw, poses, jittered=True)
    eval_camera = rv.PerspectiveCameraSensor(hw, hw, poses[::max(1, views // 4)][:4].contiguous(), jittered=True)
    rv.seed(5)
    with torch.no_grad():
        target_field = _nee_scene(t_sigma, t_albedo, 8)
        target = camera.capture(target_field, fw_samples=spp)
        rv.seed(6)
        target_hi = eval_camera.capture(target_field, fw_samples=256)
    del t_sigma, t_albedo, target_field
    sigma = torch.nn.Parameter(torch.full((res, res, res, 1), 5.0, device=dev), requires_grad=True)
    albedo = torch.nn.Parameter(torch.full((ares, ares, ares, 3), 0.7, device=dev), requires_grad=True)
    check_camera = rv.PerspectiveCameraSensor(hw, hw, poses[:check_views].contiguous(), jittered=True)
    albedo.requires_grad_(False)
    checks = _directional_check(check_camera, lambda s: _nee_scene_fixed(s, albedo, 8, 7.0), sigma, check_spp, n_dirs=2,
                                h=0.25)
    albedo.requires_grad_(True)
    sigma.requires_grad_(False)
    checks_albedo = _directional_check(check_camera, lambda a: _nee_scene_fixed(sigma, a, 8, 7.0), albedo, check_spp,
                                       n_dirs=2, h=0.02)
    sigma.requires_grad_(True)
    field = _nee_scene(sigma, albedo, 8)

    def eval_loss():
        with torch.no_grad():
            rv.seed(7)
            return float(((eval_camera.capture(field, fw_samples=256) - target_hi) ** 2).mean())

    eval_first = eval_loss()
    opts = [(sigma, inverse.FusedAdam(sigma, lr=1e-2), (0.0, math.inf)),
            (albedo, inverse.FusedAdam(albedo, lr=1e-2), (0.0, 1.0))]
    stream = torch.cuda.current_stream(dev)
    losses, fw, bw, ad = [], [], [], []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        e = _events(4)
        e[0].record(stream)
        img = camera.capture(field, fw_samples=spp, bw_samples=spp)
        loss = ((img - target) ** 2).mean()
        e[1].record(stream)
        loss.backward()
        e[2].record(stream)
        for p, opt, bounds in opts:
            opt.step(p.grad, zero_grad=False, bounds=bounds)
            p.grad = None
        e[3].record(stream)
        losses.append(loss.detach())
        fw.append(e)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    fw_ms = sum(e[0].elapsed_time(e[1]) for e in fw) / steps
    bw_ms = sum(e[1].elapsed_time(e[2]) for e in fw) / steps
    adam_ms = sum(e[2].elapsed_time(e[3]) for e in fw) / steps
    losses = [float(x) for x in losses]
    samples = views * hw * hw * spp
    return {"workload": f"{res}^3 density + {ares}^3 RGB albedo from 5.0 / 0.7, {views} views {hw}x{hw}, {spp} spp, "
                        f"NEE + DT, max 8 bounces, defensive-replay backward (fresh seeds), Adam lr 1e-2, {steps} steps; "
                        f"gradient checks on {check_views} views at {check_spp} spp, 4 seed batches",
            "loss_curve": losses, "loss_first": losses[0], "loss_last": losses[-1],
            "eval_loss_256spp_first": eval_first, "eval_loss_256spp_last": eval_loss(),
            "wall_s": wall, "ms_per_step": wall * 1e3 / steps, "fw_ms": fw_ms, "bw_ms": bw_ms, "adam_ms": adam_ms,
            "fw_samples_per_s": samples / (fw_ms * 1e-3), "bw_samples_per_s": samples / (bw_ms * 1e-3),
            "directional_fd_density": checks, "directional_fd_albedo": checks_albedo,
            "fd_rel_err_max": max(c["rel_err"] for c in checks + checks_albedo),
            "fd_rel_err_richardson_max": max(c["rel_err_richardson"] for c in checks + checks_albedo),
            "fd_z_richardson_max": max(c["z_richardson"] for c in checks + checks_albedo)}


def _nee_scene_fixed(sigma, albedo, max_bounces, majorant):
    """_nee_scene with a fixed constant majorant (finite-difference checks)."""
    import torch
    import rendervous_amd as rv

    ds = rv.objects.DependencySet()
    ds.add_parameters(sigma_tensor=sigma, scattering_albedo_tensor=albedo,
                      environment_tensor=torch.tensor([0.1, 0.1, 0.1]), phase_g_tensor=torch.tensor([0.6]),
                      majorant_tensor=torch.tensor([float(majorant), 1000000.0]))
    ds.requires(rv.objects.medium_environment_sampler_sun, direction=(0.4, 0.8, 0.45), radiance=(3.0, 3.0, 3.0),
                probability=0.5)
    ds.requires(rv.objects.medium_phase_HG, mixture=False)
    ds.requires(rv.objects.medium_phase_sampler_HG, mixture=False)
    ds.requires(rv.objects.medium_transmittance_DT)
    ds.requires(rv.objects.medium_radiance_path_integrator_NEE_DT, max_bounces=max_bounces)
    return ds.radiance


def _nee_scene(sigma, albedo, max_bounces):
    import torch
    import rendervous_amd as rv

    ds = rv.objects.DependencySet()
    ds.add_parameters(sigma_tensor=sigma, scattering_albedo_tensor=albedo,
                      environment_tensor=torch.tensor([0.1, 0.1, 0.1]), phase_g_tensor=torch.tensor([0.6]))
    ds.requires(rv.objects.medium_majorant_coarse_grid, resolution=16)
    ds.requires(rv.objects.medium_environment_sampler_sun, direction=(0.4, 0.8, 0.45), radiance=(3.0, 3.0, 3.0),
                probability=0.5)
    ds.requires(rv.objects.medium_phase_HG, mixture=False)
    ds.requires(rv.objects.medium_phase_sampler_HG, mixture=False)
    ds.requires(rv.objects.medium_transmittance_DT)
    ds.requires(rv.objects.medium_radiance_path_integrator_NEE_DT, max_bounces=max_bounces)
    return ds.radiance


def nee_gradient_leg(dev, res=512, ares=128, views=4, hw=512, spp=32, steps=3):
    """One GPU's share (4 of 32 views) of the multiple-scattering inverse-render
    step: 512^3 density + 128^3 RGB albedo, 512x512, 32 spp, NEE + DT, max 8
    bounces; L2 loss, backward by the defensive replay estimator replaying the
    forward's seeds, fused Adam on both grids."""
    import torch
    import rendervous_amd as rv
    from rendervous_amd import inverse

    t_sigma = rv.functions.fbm_value_noise(res, seed=21, scale=60.0, sphere_radius=0.5, device=dev)
    t_albedo = (0.55 + 0.4 * rv.functions.fbm_value_noise(ares, seed=22, scale=1.0, device=dev)).expand(
        ares, ares, ares, 3).contiguous()
    poses = rv.functions.orbit_poses(32, radius=2.5)[:views].to(dev)
    camera = rv.PerspectiveCameraSensor(hw, hw, poses, jittered=True)
    rv.seed(5)
    with torch.no_grad():
        target = camera.capture(_nee_scene(t_sigma, t_albedo, 8), fw_samples=spp)
    del t_sigma, t_albedo
    sigma = torch.nn.Parameter(torch.full((res, res, res, 1), 5.0, device=dev), requires_grad=True)
    albedo = torch.nn.Parameter(torch.full((ares, ares, ares, 3), 0.7, device=dev), requires_grad=True)
    field = _nee_scene(sigma, albedo, 8)
    r = _inverse_step_timing(camera, field, target, spp,
                             [(sigma, inverse.FusedAdam(sigma, lr=1e-2), (0.0, math.inf)),
                              (albedo, inverse.FusedAdam(albedo, lr=1e-2), (0.0, 1.0))], steps)
    samples = views * hw * hw * spp
    r["workload"] = (f"NEE inverse render step, one GPU's share: {res}^3 density + {ares}^3 RGB albedo, {views} of 32 "
                     f"views {hw}x{hw}, {spp} spp, DT + NEE, max 8 bounces, replay backward, fused Adam")
    r["fw_samples_per_s"] = samples / (r["fw_ms"] * 1e-3)
    r["bw_samples_per_s"] = samples / (r["bw_ms"] * 1e-3)
    return r


def time_to_rmse_leg(dev, radiance, camera, ref_spp=4096, rel_eps=0.05):
    """Plain renderer (no radiance cache): wall time of the smallest single
    capture whose image RMSE against a 4096-spp render of the same frame is at
    most rel_eps x the reference's mean.  Captures of 8, 16, 32, ... spp."""
    import torch
    import rendervous_amd as rv

    stream = torch.cuda.current_stream(dev)
    with torch.no_grad():
        rv.seed(12345)
        ref = None
        for _ in range(ref_spp // 64):
            img = camera.capture(radiance, fw_samples=64)
            ref = img if ref is None else ref + img
        ref = ref / (ref_spp // 64)
        eps = rel_eps * float(ref.mean())
        rows = []
        spp = 8
        while spp <= ref_spp // 4:
            rv.seed(spp)
            e = _events(2)
            e[0].record(stream)
            img = camera.capture(radiance, fw_samples=spp)
            e[1].record(stream)
            torch.cuda.synchronize()
            rmse = float(((img - ref) ** 2).mean().sqrt())
            rows.append({"spp": spp, "ms": e[0].elapsed_time(e[1]), "rmse": rmse})
            if rmse <= eps:
                break
            spp *= 2
    hit = rows[-1] if rows[-1]["rmse"] <= eps else None
    return {"workload": f"time to RMSE <= {rel_eps:g} x mean vs a {ref_spp}-spp render, same NEE frame, plain renderer",
            "epsilon": eps, "time_to_rmse_ms": hit["ms"] if hit else None, "spp_needed": hit["spp"] if hit else None,
            "ladder": rows}


def nrc_leg(dev, res=512, hw=1024, ref_spp=4096, target_spp=256, max_bounces=32, traces=2, steps_per_trace=32,
            batch=24576, lr=(1e-2, 1e-3), n_paths=65536, full_ladder=False, terminate_after=2):
    """Neural radiance cache vs plain NEE on the 512^3 cloud (fBm x60, albedo
    0.99, HG g=0.8, sun + env 0.1), 1024x1024: epsilon = the RMSE of the plain
    renderer at 256 spp against a 4096-spp plain render; the cache is trained
    online (`traces` traces of `n_paths` paths, `steps_per_trace` Adam steps of
    `batch` records after each, learning rate decaying geometrically from lr[0]
    to lr[1]) and its renders (paths ended into the MLP after 2 bounces) are
    taken at doubling spp until the RMSE is <= epsilon.  Times: HIP events
    around the GPU work."""
    import torch
    import rendervous_amd as rv
    from rendervous_amd.nrc import NeuralRadianceCache

    sigma = rv.functions.fbm_value_noise(res, seed=13, scale=60.0, sphere_radius=0.5, device=dev)
    ds = rv.objects.DependencySet()
    ds.add_parameters(sigma_tensor=sigma, scattering_albedo_tensor=torch.tensor([0.99, 0.99, 0.99]),
                      environment_tensor=torch.tensor([0.1, 0.1, 0.1]), phase_g_tensor=torch.tensor([0.8]))
    ds.requires(rv.objects.medium_majorant_coarse_grid, resolution=16)
    ds.requires(rv.objects.medium_environment_sampler_sun, direction=(0.4, 0.8, 0.45), radiance=(3.0, 3.0, 3.0),
                probability=0.5)
    ds.requires(rv.objects.medium_phase_HG, mixture=False)
    ds.requires(rv.objects.medium_phase_sampler_HG, mixture=False)
    ds.requires(rv.objects.medium_transmittance_DT)
    ds.requires(rv.objects.medium_radiance_path_integrator_NEE_DT, max_bounces=max_bounces)
    field = ds.radiance
    camera = rv.PerspectiveCameraSensor(hw, hw, rv.functions.look_at_pose((0.0, 0.3, 2.5))[None].to(dev),
                                        jittered=True)
    stream = torch.cuda.current_stream(dev)

    def timed(fn):
        e = _events(2)
        e[0].record(stream)
        out = fn()
        e[1].record(stream)
        torch.cuda.synchronize()
        return out, e[0].elapsed_time(e[1])

    with torch.no_grad():
        ref = None
        for k in range(ref_spp // 64):
            rv.seed(100000 + k)
            img = camera.capture(field, fw_samples=64)
            ref = img if ref is None else ref + img
        ref = ref / (ref_spp // 64)

        def rmse(img):
            return float(((img - ref) ** 2).mean().sqrt())

        rv.seed(7)
        img, plain_ms = timed(lambda: camera.capture(field, fw_samples=target_spp))
        eps = rmse(img)
        cache = NeuralRadianceCache(dev, seed=7, lr=lr[0], terminate_after=terminate_after)
        n_steps = traces * steps_per_trace
        rates = [lr[0] * (lr[1] / lr[0]) ** (k / max(1, n_steps - 1)) for k in range(n_steps)]
        cache.render(field, camera, spp=32)  # untimed: sizes the query buffer once

        trace_seeds = []
        for t in range(traces):
            rv.seed(5000 + t)
            trace_seeds.append(rv._internal.get_seeds())

        def train():
            # the traces' paths (independent of the cache) in one launch, then
            # steps_per_trace Adam steps on each trace's records in turn
            cache.train(field, camera, n_paths=n_paths, max_vertices=32, steps=steps_per_trace, batch=batch,
                        lr=rates, traces=traces, seeds=trace_seeds)
        _, train_ms = timed(train)
        ladder, hit = [], None
        spp = 4
        while spp <= target_spp:
            rv.seed(77 + spp)
            img, ms = timed(lambda: cache.render(field, camera, spp=spp))
            ladder.append({"spp": spp, "ms": ms, "rmse": rmse(img)})
            if ladder[-1]["rmse"] <= eps and hit is None:
                hit = ladder[-1]
                if not full_ladder:
                    break
            spp *= 2
        # MLP throughput: inference over the last render's queries and one training step
        queries = cache.queries_last()
        feats = torch.rand(queries, 96, device=dev) * 2.0 - 1.0
        outq = torch.empty(queries, 3, device=dev)
        from rendervous_amd import _lib as L
        _, infer_ms = timed(lambda: L.check(L.load().vr_nrc_mlp_fw(feats.data_ptr(), queries, cache.weights.data_ptr(),
                                                                    outq.data_ptr(), L.stream_handle(dev)), "mlp_fw"))
    flops_q = 2.0 * (76 * 64 + 3 * 64 * 64 + 64 * 3)  # algorithmic FLOPs per query (no padding)
    tflops = flops_q * queries / (infer_ms * 1e-3) / 1e12
    # vr_nrc_mlp_fw reads each query's 96 FP32 features and writes its RGB: the
    # feature stream, not the MFMAs, bounds it (the render's fused inference
    # forms the features in registers instead)
    bytes_q = 96 * 4 + 3 * 4
    gbs = bytes_q * queries / (infer_ms * 1e-3) / 1e9
    nrc_ms = train_ms + hit["ms"] if hit else None
    return {
        "workload": f"time to RMSE <= plain {target_spp} spp RMSE, {res}^3 cloud fBm x60, albedo 0.99, g 0.8, "
                    f"{hw}x{hw}, max {max_bounces} bounces, vs a {ref_spp}-spp plain render; cache trained on "
                    f"{traces} traces of {n_paths} NEE training paths (<= 32 records each), {steps_per_trace} Adam "
                    f"steps of {batch} records per trace; cached render ends paths after {terminate_after} bounces",
        "epsilon": eps, "plain_ms": plain_ms, "plain_spp": target_spp,
        "nrc_train_ms": train_ms, "nrc_render": hit, "nrc_ms": nrc_ms,
        "speedup": plain_ms / nrc_ms if nrc_ms else None, "ladder": ladder,
        "mlp_inference": {"queries": queries, "ms": infer_ms,
                          "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                       "frac": gbs / HBM_PEAK_GBS, "bytes_per_query": bytes_q},
                          "mfma": {"achieved": tflops, "peak": 2500.0, "unit": "TFLOP/s", "frac": tflops / 2500.0,
                                   "flops_per_query": flops_q}},
    }


def dda_leg(dev, reps=20):
    """Deterministic DDA transmittance of the 64^3 Gaussian-blob grid, 128x128, 1 ray/pixel."""
    import torch
    import rendervous_amd as rv

    sigma = rv.functions.gaussian_blobs(64, seed=0, max_value=20.0, device=dev)
    bmin, bmax = rv.normalized_box(sigma)
    field = rv.GridDDATransmittance(rv.Grid3D(sigma, bmin, bmax), rv.RayBoxIntersection(bmin, bmax))
    camera = rv.PerspectiveCameraSensor(128, 128, rv.functions.look_at_pose((0.0, 0.0, 2.5))[None].to(dev))
    stream = torch.cuda.current_stream(dev)
    with torch.no_grad():
        camera.capture(field)
        torch.cuda.synchronize()
        e = _events(2)
        e[0].record(stream)
        for _ in range(reps):
            img = camera.capture(field)
        e[1].record(stream)
        torch.cuda.synchronize()
    ms = e[0].elapsed_time(e[1]) / reps
    return {"workload": "DDA transmittance, 64^3 blobs, 128x128, 1 ray/pixel", "rays_per_s": 128 * 128 / (ms * 1e-3),
            "ms_per_frame": ms, "mean_T": float(img.mean().item())}


def grid_tools_leg(dev, res=512, reps=10):
    """TV regulariser fw/bw over a res^3 density and a 2x resample (res/2)^3 -> res^3,
    the per-iteration grid passes of coarse-to-fine inverse rendering; HBM roofline
    with algorithmic bytes 8 B/voxel (TV fw), 12 B/voxel (TV bw), 4 B per input and
    output vertex (resample)."""
    import torch
    import rendervous_amd as rv
    from rendervous_amd import _lib as L

    V = res ** 3
    x = rv.functions.fbm_value_noise(res, seed=3, scale=20.0, device=dev).contiguous()
    og = torch.randn(res, res, res, 1, device=dev)
    tv = torch.empty_like(og)
    g = torch.empty_like(x)
    half = rv.functions.resample_grid(x, (res // 2,) * 3)
    up = torch.empty_like(x)
    lib, st = L.load(), L.stream_handle(dev)
    sh, hs = L.i64s(x.shape), L.i64s(half.shape[:-1])
    dst = L.i64s((res,) * 3)
    calls = {
        "tv_fw": (lambda: lib.vr_total_variation_fw(x.data_ptr(), sh, 3, tv.data_ptr(), st), 8.0 * V),
        "tv_bw": (lambda: lib.vr_total_variation_bw(x.data_ptr(), og.data_ptr(), sh, 3, g.data_ptr(), st), 12.0 * V),
        "resample_2x": (lambda: lib.vr_resample_grid(half.data_ptr(), hs, dst, 3, 1, up.data_ptr(), st),
                        4.0 * (V + half.numel())),
        # practical ceiling: a device copy of the same grid (reads and writes 4 B/voxel)
        "copy": (lambda: (up.copy_(x), 0)[1], 8.0 * V),
    }
    out = {"workload": f"{res}^3 x 1 float32 grid: TV forward, TV backward, {res // 2}^3 -> {res}^3 resample"}
    stream = torch.cuda.current_stream(dev)
    for name, (fn, nbytes) in calls.items():
        L.check(fn(), name)
        torch.cuda.synchronize()
        e = _events(2)
        e[0].record(stream)
        for _ in range(reps):
            fn()
        e[1].record(stream)
        torch.cuda.synchronize()
        ms = e[0].elapsed_time(e[1]) / reps
        gbs = nbytes / (ms * 1e-3) / 1e9
        out[name] = {"ms": ms, "achieved_GBs": gbs, "frac": gbs / HBM_PEAK_GBS, "bytes": nbytes}
    return out


def _icosphere(subdiv: int, radius: float, center):
    """Triangle sphere by repeated 4:1 subdivision of an icosahedron."""
    import numpy as np
    t = (1.0 + 5 ** 0.5) / 2
    v = [[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
         [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]]
    f = [[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6],
         [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5], [2, 4, 11], [6, 2, 10],
         [8, 6, 7], [9, 8, 1]]
    v = [list(np.asarray(p) / np.linalg.norm(p)) for p in v]
    for _ in range(subdiv):
        mid = {}

        def m(a, b):
            k = (min(a, b), max(a, b))
            if k not in mid:
                p = (np.asarray(v[a]) + np.asarray(v[b])) * 0.5
                v.append(list(p / np.linalg.norm(p)))
                mid[k] = len(v) - 1
            return mid[k]
        nf = []
        for a, b, c in f:
            ab, bc, ca = m(a, b), m(b, c), m(c, a)
            nf += [[a, ab, ca], [b, bc, ab], [c, ca, bc], [ab, bc, ca]]
        f = nf
    pos = np.asarray(v, np.float32) * radius + np.asarray(center, np.float32)
    return pos, np.asarray(f, np.int32)


def scene_leg(dev, res=256, hw=512, spp=16, reps=3):
    """Scene-level path tracing (PathtracedScene): the 256^3 fBm cloud inside a
    null box over a Lambert floor with an emitting panel, a mirror wall and a
    glass (Fresnel) icosphere of 20480 triangles, constant sky; forward capture
    and replay backward into the density; plus BVH ray casts against the
    sphere and a 1M-triangle soup."""
    import numpy as np
    import torch
    import rendervous_amd as rv

    sigma = torch.nn.Parameter(rv.functions.fbm_value_noise(res, seed=7, scale=40.0, sphere_radius=0.5, device=dev),
                               requires_grad=True)
    bmin, bmax = rv.normalized_box(sigma)
    medium = rv.HeterogeneousMediumMaterial(rv.Grid3D(sigma, bmin, bmax), scattering_albedo=rv.const[0.9, 0.9, 0.9],
                                            phase_sampler=rv.HGPhaseSampler(0.6), majorant=float(sigma.detach().max()))
    half = torch.diag(torch.tensor([0.5, 0.5, 0.5]))
    cloud = rv.TransformedGeometry(rv.MeshGeometry.box(), torch.cat([half, torch.zeros(1, 3)]))
    floor = rv.MeshGeometry(torch.tensor([[-4.0, -0.6, -4.0], [4.0, -0.6, -4.0], [4.0, -0.6, 4.0], [-4.0, -0.6, 4.0]]),
                            indices=torch.tensor([[0, 2, 1], [0, 3, 2]], dtype=torch.int32), compute_normals=False)
    mirror = rv.MeshGeometry(torch.tensor([[1.5, -0.6, -2.0], [1.5, 1.5, -2.0], [1.5, 1.5, 2.0], [1.5, -0.6, 2.0]]),
                             indices=torch.tensor([[0, 1, 2], [0, 2, 3]], dtype=torch.int32), compute_normals=False)
    sp, sf = _icosphere(5, 0.35, (-0.9, -0.25, 0.5))
    glass = rv.MeshGeometry(torch.from_numpy(sp), indices=torch.from_numpy(sf))
    geoms = {"cloud": cloud, "floor": floor, "mirror": mirror, "glass": glass}
    mats = {"cloud": rv.SurfaceMaterial(rv.NullSurfaceScatteringSampler()),
            "floor": rv.SurfaceMaterial(rv.LambertSurfaceScatteringSampler(),
                                        rv.IsotropicSurfaceEmission(rv.const[0.2, 0.2, 0.2])),
            "mirror": rv.SurfaceMaterial(rv.MirrorSurfaceScatteringSampler()),
            "glass": rv.SurfaceMaterial(rv.FresnelSurfaceScatteringSampler(1.5))}
    scene = rv.Scene.from_graph(geoms, rv.UniformEnvironmentSampler(rv.const[0.5, 0.5, 0.5]), materials=mats,
                                inside_media={"cloud": medium})
    field = scene.pathtrace()
    camera = rv.PerspectiveCameraSensor(hw, hw, rv.functions.look_at_pose((0.0, 0.4, 2.5))[None].to(dev),
                                        jittered=True)
    stream = torch.cuda.current_stream(dev)
    pixels = hw * hw
    dbg = torch.zeros(pixels, dtype=torch.int32, device=dev)
    with torch.no_grad():
        camera.capture(field, fw_samples=spp, debug_out=dbg)
    torch.cuda.synchronize()
    lookups = float(dbg.to(torch.float64).sum().item())
    fw_ms = bw_ms = 0.0
    scatters = 0.0
    for k in range(reps + 1):
        e = _events(3)
        e[0].record(stream)
        img = camera.capture(field, fw_samples=spp, bw_samples=spp, debug_out=dbg)
        e[1].record(stream)
        img.sum().backward()
        e[2].record(stream)
        torch.cuda.synchronize()
        if k:
            fw_ms += e[0].elapsed_time(e[1]) / reps
            bw_ms += e[1].elapsed_time(e[2]) / reps
            scatters += float(dbg.to(torch.float64).sum().item()) / reps
        sigma.grad = None
    samples = pixels * spp
    # BVH ray casts: 1M rays against the sphere alone and against a 1M-triangle soup
    from rendervous_amd import _lib as L
    rng = np.random.default_rng(0)
    casts = {}
    soup_c = rng.uniform(-1, 1, size=(1 << 20, 1, 3))
    soup = (soup_c + rng.normal(scale=0.004, size=(1 << 20, 3, 3))).reshape(-1, 3).astype(np.float32)
    for name, geom in (("sphere_20480", glass),
                       ("soup_1M", rv.MeshGeometry(torch.from_numpy(soup), compute_normals=False))):
        n = 1 << 20
        o = rng.normal(size=(n, 3))
        o = o / np.linalg.norm(o, axis=1, keepdims=True) * 2.5
        d = rng.uniform(-0.7, 0.7, size=(n, 3)) - o
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        rays = torch.from_numpy(np.concatenate([o, d], 1).astype(np.float32)).to(dev)
        geom(rays)  # builds the BVH and uploads it
        torch.cuda.synchronize()
        e = _events(2)
        e[0].record(stream)
        for _ in range(5):
            hits = geom(rays)
        e[1].record(stream)
        torch.cuda.synchronize()
        ms = e[0].elapsed_time(e[1]) / 5
        casts[name] = {"rays": n, "ms": ms, "rays_per_s": n / (ms * 1e-3),
                       "hit_fraction": float((hits[:, 15].view(torch.int32) >= 0).float().mean())}
    return {
        "workload": f"PathtracedScene: {res}^3 fBm cloud in a null box, Lambert floor with emission, mirror, glass "
                    f"icosphere (20480 triangles), {hw}x{hw}, {spp} spp; forward + replay backward into the density",
        "fw_ms": fw_ms, "bw_ms": bw_ms, "fw_samples_per_s": samples / (fw_ms * 1e-3),
        "bw_samples_per_s": samples / (bw_ms * 1e-3), "fw_lookups": lookups,
        "grad_voxel_updates_per_s": 8.0 * scatters / (bw_ms * 1e-3), "trilinear_scatters": scatters,
        "mean": float(img.mean()), "bvh_casts": casts,
    }


def _quiet(leg, *args):
    """Runs a leg with Python's cyclic collector paused (collected first), so
    that no collector pause lands inside its timed regions."""
    gc.collect()
    enabled = gc.isenabled()
    gc.disable()
    try:
        return leg(*args)
    finally:
        if enabled:
            gc.enable()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # the first frames of a run are slower while the GPU clocks ramp (4.48 ms
    # per headline frame over 5 frames after 1 warm-up, 4.37 ms over 20 after 5)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--res", type=int, default=256)
    ap.add_argument("--width", type=int, default=512)
    ap.add_argument("--height", type=int, default=512)
    ap.add_argument("--spp", type=int, default=64)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-secondary", action="store_true", help="skip the gradient and DDA legs")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--leg", choices=["gradient", "nee_gradient", "nrc", "dda", "grid_tools", "scene", "convergence",
                                      "nee_convergence", "gradient_gate"],
                    help="run only this secondary leg and print its JSON (for profiling)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run one 256 MiB device copy first (reference kernel for HBM byte counters)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported (run with --gpus 1)",
              file=sys.stderr)
        sys.exit(2)

    import torch
    import rendervous_amd as rv

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    if args.leg is not None:
        legs = {"gradient": gradient_leg, "nee_gradient": nee_gradient_leg, "nrc": nrc_leg, "dda": dda_leg,
                "grid_tools": grid_tools_leg, "scene": scene_leg, "convergence": convergence_leg,
                "nee_convergence": nee_convergence_leg, "gradient_gate": gradient_gate_leg}
        print(json.dumps({args.leg: _quiet(legs[args.leg], dev)}), flush=True)
        return
    if args.calibrate_copy:
        a = torch.empty(64 << 20, dtype=torch.float32, device=dev).fill_(1.0)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)  # reads and writes exactly 256 MiB
        torch.cuda.synchronize()
        del a, b
    log(f"building {args.res}^3 fBm scene on {torch.cuda.get_device_name(dev)}")
    sigma, radiance, camera = build_scene(args.res, args.width, args.height, dev)
    torch.cuda.synchronize()

    samples = args.width * args.height * args.spp
    pixels = args.width * args.height

    # lookups per launch from the kernel's per-thread counter (untimed run)
    dbg = torch.zeros(pixels, dtype=torch.int32, device=dev)
    rv.seed(0)
    with torch.no_grad():
        camera.capture(radiance, fw_samples=args.spp, debug_out=dbg)
    torch.cuda.synchronize()
    lookups = float(dbg.to(torch.float64).sum().item())
    log(f"lookups per frame: {lookups:.4e} ({lookups / samples:.1f} per sample)")

    with torch.no_grad():
        for _ in range(args.warmup):
            img = camera.capture(radiance, fw_samples=args.spp)
        torch.cuda.synchronize()
        stream = torch.cuda.current_stream(dev)
        starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
        ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
        gc.collect()
        gc.disable()  # no collector pause between the timed launches
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(args.steps):
            starts[k].record(stream)
            img = camera.capture(radiance, fw_samples=args.spp)
            ends[k].record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        gc.enable()
    kernel_ms = sum(s.elapsed_time(e) for s, e in zip(starts, ends)) / args.steps
    ms_per_step = wall * 1000.0 / args.steps
    assert torch.isfinite(img).all()
    log(f"frame mean {img.mean().item():.5f}; {kernel_ms:.2f} ms per frame (events), {ms_per_step:.2f} ms wall")

    bytes_per_launch = 32.0 * lookups + 12.0 * pixels
    achieved = bytes_per_launch / (kernel_ms * 1e-3) / 1e9
    result = {
        "metric": "rays/s (camera samples/s, NEE delta-tracking volume path tracer)",
        "value": samples / (ms_per_step * 1e-3),
        "unit": "rays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded fBm value-noise density)",
        "config": {"workload": f"NEE+DT volume path tracing, {args.res}^3 fBm x40 in sphere, albedo 0.9, HG g=0.6, "
                               f"sun + env 0.1, 16^3 majorant grid, {args.width}x{args.height}, {args.spp} spp, "
                               f"max 16 bounces", "samples_per_step": samples, "lookups_per_step": lookups},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": profiled_traffic(args.res, args.width, args.height, args.spp),
                     "measured_bound": "VALU issue: the persistent kernel's SIMDs issue VALU work ~100 % of "
                                       "the frame at 38 % active lanes (2.7e9 wave instructions, 161 per camera "
                                       "sample); HBM traffic 11.8 GB per frame is 34 % of its time at 8 TB/s "
                                       "(profiles/nee_persistent_pmc_256cube_512x512x64spp.txt)"},
    }
    if not args.no_secondary:
        rmse = _quiet(time_to_rmse_leg, dev, radiance, camera)
        del img, radiance, camera
        torch.cuda.empty_cache()
        result["secondary"] = {"time_to_rmse": rmse, "gradient": _quiet(gradient_leg, dev)}
        torch.cuda.empty_cache()
        result["secondary"]["nee_gradient"] = _quiet(nee_gradient_leg, dev)
        torch.cuda.empty_cache()
        result["secondary"]["nrc"] = _quiet(nrc_leg, dev)
        torch.cuda.empty_cache()
        result["secondary"]["dda"] = _quiet(dda_leg, dev)
        torch.cuda.empty_cache()
        result["secondary"]["grid_tools"] = _quiet(grid_tools_leg, dev)
        torch.cuda.empty_cache()
        result["secondary"]["scene"] = _quiet(scene_leg, dev)
        torch.cuda.empty_cache()
        result["secondary"]["convergence"] = _quiet(convergence_leg, dev)
        torch.cuda.empty_cache()
        result["secondary"]["gradient_gate"] = _quiet(gradient_gate_leg, dev)
        torch.cuda.empty_cache()
        result["secondary"]["nee_inverse_32_views"] = _quiet(nee_convergence_leg, dev)
        log(f"secondary: {json.dumps(result['secondary'])}")
    if not args.no_cpu_baseline:
        m = build_scene(args.res, args.width, args.height, dev)[1]
        result["cpu_baseline"] = cpu_baseline(sigma.cpu(), m, args.width, args.height, args.spp, args.cpu_budget)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
