"""Synthetic scene generators and camera helpers for the benchmark workloads.

The density generators use only integer hashing and polynomial interpolation
in torch, so CPU and GPU produce bit-identical grids for the same seed:
  * ``gaussian_blobs``: sum of isotropic Gaussians evaluated at vertex
    coordinates i/(n-1), rescaled to a given maximum;
  * ``fbm_value_noise``: fractal sum of lattice value noise (hashed lattice,
    smoothstep interpolation), normalised to [0,1], scaled and zeroed outside a
    sphere.
Poses follow the reference's camera layout [origin, direction, up]
(include/maps/sensor_perspective_camera.h:4-7).
"""
from __future__ import annotations

import math
import typing

import numpy as np
import torch

_M32 = 0xFFFFFFFF


def _hash3(ix: torch.Tensor, iy: torch.Tensor, iz: torch.Tensor, seed: int) -> torch.Tensor:
    """32-bit lattice hash -> float in [0,1) (integer mixing in int64 with 32-bit masks)."""
    h = (ix * 0x8DA6B343 + iy * 0xD8163841 + iz * 0xCB1AB31F + seed * 0x165667B1) & _M32
    h = h ^ (h >> 15)
    h = (h * 0x2C1B3C6D) & _M32
    h = h ^ (h >> 12)
    h = (h * 0x297A2D39) & _M32
    h = h ^ (h >> 15)
    return (h >> 8).to(torch.float32) * (1.0 / 16777216.0)


def _smooth(t: torch.Tensor) -> torch.Tensor:
    return t * t * (3.0 - 2.0 * t)


def value_noise(p: torch.Tensor, seed: int) -> torch.Tensor:
    """Lattice value noise at points p[...,3] (float32), values in [0,1)."""
    fl = torch.floor(p)
    i = fl.to(torch.int64)
    f = _smooth(p - fl)
    out = torch.zeros(p.shape[:-1], dtype=torch.float32, device=p.device)
    for dz in (0, 1):
        wz = f[..., 2] if dz else 1.0 - f[..., 2]
        for dy in (0, 1):
            wy = f[..., 1] if dy else 1.0 - f[..., 1]
            for dx in (0, 1):
                wx = f[..., 0] if dx else 1.0 - f[..., 0]
                out = out + _hash3(i[..., 0] + dx, i[..., 1] + dy, i[..., 2] + dz, seed) * (wx * wy * wz)
    return out


def fbm_value_noise(n: int, *, seed: int = 0, octaves: int = 4, lacunarity: float = 2.0, gain: float = 0.5,
                    base_frequency: float = 4.0, scale: float = 40.0, sphere_radius: float = 0.5,
                    device: typing.Union[str, torch.device] = "cpu", chunk: int = 32) -> torch.Tensor:
    """[n,n,n,1] fBm density over the normalised box [-0.5,0.5]^3.

    Sum over octaves o of gain^o * noise(base_frequency * lacunarity^o * p),
    min/max-normalised to [0,1], multiplied by `scale` and zeroed outside the
    sphere of radius `sphere_radius` centred at the origin.
    """
    dev = torch.device(device)
    coords = torch.linspace(-0.5, 0.5, n, dtype=torch.float32, device=dev)
    out = torch.empty(n, n, n, dtype=torch.float32, device=dev)
    for z0 in range(0, n, chunk):
        z1 = min(n, z0 + chunk)
        zz, yy, xx = torch.meshgrid(coords[z0:z1], coords, coords, indexing="ij")
        p = torch.stack([xx, yy, zz], dim=-1)
        acc = torch.zeros(p.shape[:-1], dtype=torch.float32, device=dev)
        amp, freq = 1.0, base_frequency
        for o in range(octaves):
            acc = acc + amp * value_noise(p * freq + 17.0 * (o + 1), seed * 131 + o)
            amp *= gain
            freq *= lacunarity
        out[z0:z1] = acc
    lo, hi = out.min(), out.max()
    out = (out - lo) / torch.clamp(hi - lo, min=1e-12)
    out = out.clamp(0.0, 1.0) * scale
    zz, yy, xx = torch.meshgrid(coords, coords, coords, indexing="ij")
    inside = (xx * xx + yy * yy + zz * zz) <= sphere_radius * sphere_radius
    out = torch.where(inside, out, torch.zeros_like(out))
    return out.unsqueeze(-1).contiguous()


def gaussian_blobs(n: int, *, seed: int = 0, count: int = 8, max_value: float = 20.0,
                   device: typing.Union[str, torch.device] = "cpu") -> torch.Tensor:
    """[n,n,n,1] sum of `count` isotropic Gaussians, centres U[0.2,0.8]^3, sigma U[0.05,0.15],
    evaluated at vertex coordinates i/(n-1) and rescaled to max `max_value`."""
    g = torch.Generator().manual_seed(seed)
    centers = 0.2 + 0.6 * torch.rand(count, 3, generator=g, dtype=torch.float64)
    sig = 0.05 + 0.10 * torch.rand(count, generator=g, dtype=torch.float64)
    c = torch.linspace(0.0, 1.0, n, dtype=torch.float64)
    zz, yy, xx = torch.meshgrid(c, c, c, indexing="ij")
    acc = torch.zeros(n, n, n, dtype=torch.float64)
    for k in range(count):
        d2 = (xx - centers[k, 0]) ** 2 + (yy - centers[k, 1]) ** 2 + (zz - centers[k, 2]) ** 2
        acc += torch.exp(-d2 / (2 * sig[k] ** 2))
    acc = acc * (max_value / acc.max())
    return acc.to(torch.float32).unsqueeze(-1).contiguous().to(device)


def look_at_pose(origin, target=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0)) -> torch.Tensor:
    o = np.asarray(origin, dtype=np.float64)
    d = np.asarray(target, dtype=np.float64) - o
    d = d / np.linalg.norm(d)
    return torch.tensor([*o, *d, *up], dtype=torch.float32)


def orbit_poses(n: int, radius: float = 2.5, elevation: float = 0.3, seed: int = 0) -> torch.Tensor:
    """n poses on a sphere of `radius` looking at the origin (deterministic golden-angle spiral)."""
    poses = []
    ga = math.pi * (3.0 - math.sqrt(5.0))
    for i in range(n):
        y = elevation * (1.0 - 2.0 * (i + 0.5) / n)
        r = math.sqrt(max(0.0, 1.0 - y * y))
        th = ga * i + seed
        o = (radius * r * math.cos(th), radius * y, radius * r * math.sin(th))
        poses.append(look_at_pose(o))
    return torch.stack(poses)


def random_ids(batch_size: int, shape: typing.Sequence[int], out: typing.Optional[torch.Tensor] = None) -> torch.Tensor:
    """Uniform random integer indices [batch_size, len(shape)] (_functions.py:114-142)."""
    dims = torch.tensor(list(shape), dtype=torch.float64)
    r = (torch.rand(batch_size, len(shape), dtype=torch.float64) * dims).long()
    r = torch.minimum(r, (dims - 1).long())
    if out is not None:
        out.copy_(r)
        return out
    return r
