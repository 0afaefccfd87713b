"""Grid tools of the inverse-rendering loop (HIP kernels through the C ABI) and
synthetic scene generators and camera helpers for the benchmark workloads.

Grid tools (``_functions.py`` of the reference):
  * ``random_ids``: uniform random integer indices drawn with the reference's
    compute-function RNG stream;
  * ``total_variation``: the per-voxel anisotropic TV regulariser, differentiable;
  * ``resample_grid``: vertex-aligned linear resampling, in steps of at most 2x.

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

from . import _lib as L

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


def _dev(t: typing.Optional[torch.Tensor] = None) -> torch.device:
    if t is not None and t.is_cuda:
        return t.device
    from . import _internal
    return _internal.device()


def random_ids(batch_size: int, shape: typing.Sequence[int], out: typing.Optional[torch.Tensor] = None) -> torch.Tensor:
    """Uniform random integer indices [batch_size, len(shape)], int64 on the device
    (_functions.py:114-142): out[i][k] = int(u * shape[k]) with u from thread i's
    stream of the dispatch seeds (one get_seeds() draw per call)."""
    from . import _internal
    shape = [int(v) for v in shape]
    dim = len(shape)
    dev = _dev(out)
    if out is None:
        out = torch.empty(batch_size, dim, dtype=torch.int64, device=dev)
    if not out.is_cuda:
        raise RuntimeError("random_ids writes a device tensor; no CPU fallback exists")
    assert out.dtype == torch.int64 and out.is_contiguous() and out.numel() == batch_size * dim
    seeds = _internal.get_seeds()
    L.check(L.load().vr_random_ids(out.data_ptr(), batch_size, L.i64s(shape), dim, L.u32s(seeds), L.stream_handle(dev)),
            "vr_random_ids")
    return out


def _spatial(t: torch.Tensor) -> typing.Tuple[typing.List[int], int]:
    assert t.dtype == torch.float32, "grid tools take float32 tensors"
    assert 2 <= t.dim() <= 4, "shape is (spatial dims..., channels) with 1 to 3 spatial dims"
    return list(t.shape), t.dim() - 1


class _TotalVariation(torch.autograd.Function):
    """_functions.py:416-431: forward tv [..., 1], backward through the adjoint kernel."""

    @staticmethod
    def forward(ctx, x):
        x = x.contiguous()
        shape, dim = _spatial(x)
        out = torch.empty(*shape[:-1], 1, dtype=torch.float32, device=x.device)
        L.check(L.load().vr_total_variation_fw(x.data_ptr(), L.i64s(shape), dim, out.data_ptr(),
                                               L.stream_handle(x.device)), "vr_total_variation_fw")
        ctx.save_for_backward(x)
        return out

    @staticmethod
    def backward(ctx, out_grad):
        x, = ctx.saved_tensors
        og = out_grad.contiguous()
        shape, dim = _spatial(x)
        grad = torch.empty_like(x)
        L.check(L.load().vr_total_variation_bw(x.data_ptr(), og.data_ptr(), L.i64s(shape), dim, grad.data_ptr(),
                                               L.stream_handle(x.device)), "vr_total_variation_bw")
        return grad


def total_variation(in_tensor: torch.Tensor) -> torch.Tensor:
    """Per-voxel total variation [..., 1] of a grid [..., C] (_functions.py:434-435):
    sum over the existing +-1 neighbours along every axis of sum_c |v - v_adj|."""
    if not in_tensor.is_cuda:
        raise RuntimeError("total_variation runs on the HIP device; no CPU fallback exists")
    return _TotalVariation.apply(in_tensor)


def _resample_step(x: torch.Tensor, dst: typing.Tuple[int, ...], out: typing.Optional[torch.Tensor],
                   cell: bool = False) -> torch.Tensor:
    shape, dim = _spatial(x)
    C = shape[-1]
    if out is None:
        out = torch.empty(*dst, C, dtype=torch.float32, device=x.device)
    assert tuple(out.shape) == tuple(dst) + (C,) and out.is_contiguous()
    fn = L.load().vr_resample_img if cell else L.load().vr_resample_grid
    L.check(fn(x.data_ptr(), L.i64s(shape[:-1]), L.i64s(dst), dim, C, out.data_ptr(), L.stream_handle(x.device)),
            "vr_resample_img" if cell else "vr_resample_grid")
    return out


def _resample_chain(in_tensor, dst_shape, out, cell):
    if not in_tensor.is_cuda:
        raise RuntimeError("resampling runs on the HIP device; no CPU fallback exists")
    dst = tuple(int(v) for v in dst_shape)
    x = in_tensor.contiguous()
    while True:
        cur = x.shape[:-1]
        assert len(cur) == len(dst)
        lo = tuple((d + 1) // 2 for d in cur)
        hi = tuple(d * 2 for d in cur)
        step = tuple(max(min(dst[i], hi[i]), lo[i]) for i in range(len(dst)))
        if step == dst:
            return _resample_step(x, dst, out, cell)
        x = _resample_step(x, step, None, cell)


def resample_img(in_tensor: torch.Tensor, dst_shape: typing.Sequence[int],
                 out: typing.Optional[torch.Tensor] = None) -> torch.Tensor:
    """Pixel-centred linear resampling of an image [..., C] (_functions.py:337-345),
    in steps of at most 2x per axis; `out`, if given, receives the final step."""
    return _resample_chain(in_tensor, dst_shape, out, True)


def _morton(in_tensor: torch.Tensor, resolution: int, to_morton: bool, out: typing.Optional[torch.Tensor]):
    if not in_tensor.is_cuda:
        raise RuntimeError("Morton copies run on the HIP device; no CPU fallback exists")
    x = in_tensor.contiguous()
    C = x.shape[-1]
    shape = (resolution * resolution, C) if to_morton else (resolution, resolution, C)
    if out is None:
        out = torch.empty(*shape, dtype=torch.float32, device=x.device)
    assert tuple(out.shape) == shape and out.is_contiguous() and out.dtype == torch.float32
    fn = L.load().vr_copy_img_to_morton if to_morton else L.load().vr_copy_morton_to_img
    L.check(fn(x.data_ptr(), resolution, C, out.data_ptr(), L.stream_handle(x.device)), "morton copy")
    return out


def copy_img_to_morton(in_tensor: torch.Tensor, out: typing.Optional[torch.Tensor] = None) -> torch.Tensor:
    """[R, R, C] square image (R a power of two) -> [R*R, C] in Morton order (_functions.py:434-468)."""
    assert in_tensor.dim() == 3 and in_tensor.shape[0] == in_tensor.shape[1], "in_tensor should be a square image"
    return _morton(in_tensor, int(in_tensor.shape[0]), True, out)


def copy_morton_to_img(in_tensor: torch.Tensor, out: typing.Optional[torch.Tensor] = None) -> torch.Tensor:
    """[R*R, C] in Morton order -> [R, R, C] (_functions.py:471-506)."""
    assert in_tensor.dim() == 2
    r = int(math.isqrt(int(in_tensor.shape[0])))
    assert r * r == in_tensor.shape[0], "linearization does not correspond to a square image"
    return _morton(in_tensor, r, False, out)


def create_density_quadtree(densities: torch.Tensor) -> torch.Tensor:
    """Quadtree of absolute pixel probabilities for XRQuadtreeRandomDirection
    (_functions.py:509-530): levels 2x2, 4x4, .. RxR concatenated coarse to fine,
    each in Morton order; a node's 4 children are consecutive one level below."""
    resolution = densities.shape[0]
    assert densities.dim() == 3 and densities.shape[-1] == 1
    assert densities.shape[1] == resolution, "Must be square"
    assert resolution >= 2 and resolution & (resolution - 1) == 0, "Resolution must be power of two"
    sizes = []
    size = resolution
    while size > 1:
        sizes.append(size * size)
        size //= 2
    sizes.reverse()
    offsets = [0]
    for n in sizes:
        offsets.append(offsets[-1] + n)
    pdfs = torch.zeros(offsets[-1], 1, dtype=torch.float32, device=densities.device)
    for o in reversed(offsets[:-1]):
        copy_img_to_morton(densities, out=pdfs[o:o + densities.numel()])
        densities = resample_img(densities, (int(densities.shape[0]) // 2, int(densities.shape[1]) // 2)) * 4
    return pdfs


def resample_grid(in_tensor: torch.Tensor, dst_shape: typing.Sequence[int],
                  out: typing.Optional[torch.Tensor] = None) -> torch.Tensor:
    """Vertex-aligned linear resampling of a grid [..., C] to dst_shape + (C,)
    (_functions.py:287-295): steps of at most 2x up or down per axis until the
    shape is reached; `out`, if given, receives the final step."""
    return _resample_chain(in_tensor, dst_shape, out, False)