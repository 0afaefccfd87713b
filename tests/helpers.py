"""Shared fixtures: build the same medium / camera / launch for the oracle
(host numpy buffers) and for the HIP library (device torch tensors)."""
from __future__ import annotations

import dataclasses
import math
import os
import sys
import typing

import numpy as np
import torch

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from rendervous_amd import _lib as L  # noqa: E402
from oracle import oracle as O  # noqa: E402


@dataclasses.dataclass
class Scene:
    sigma: np.ndarray                      # [D,H,W,1] float32
    bmin: tuple = (-0.5, -0.5, -0.5)
    bmax: tuple = (0.5, 0.5, 0.5)
    box_min: typing.Optional[tuple] = None  # boundary box (defaults to the grid box)
    box_max: typing.Optional[tuple] = None
    albedo_c: tuple = (0.9, 0.9, 0.9)
    albedo: typing.Optional[np.ndarray] = None  # [D,H,W,3] or [D,H,W,1]
    emission_c: tuple = (0.0, 0.0, 0.0)
    env_c: tuple = (1.0, 1.0, 1.0)
    sun_dir: tuple = (0.0, 1.0, 0.0)
    sun_radiance: tuple = (0.0, 0.0, 0.0)
    sun_prob: float = 0.0
    phase_kind: int = L.VR_PHASE_HG
    phase_g: float = 0.0
    majorant: str = "const"                # "const" or "grid"
    majorant_value: typing.Optional[float] = None
    majorant_distance: float = 1e6
    majorant_res: tuple = (4, 4, 4)
    transmittance_kind: int = L.VR_T_DT
    ds_epsilon: float = 0.1
    max_bounces: int = -1
    medium_filter: int = 3
    gathering_c: tuple = (0.0, 0.0, 0.0)
    medium_majorant: float = 1.0
    env_image: typing.Optional[np.ndarray] = None     # [H,W,3]: environment Image2D over xr coordinates
    env_quadtree: typing.Optional[np.ndarray] = None  # flat quadtree (np_quadtree) for importance sampling
    env_quadtree_levels: int = 0
    phase_g_grid: typing.Optional[np.ndarray] = None  # [D,H,W,1] over the sigma box: HG asymmetry per position


class Built:
    """A vr_medium plus the buffers it points to (and gradient buffers)."""

    def __init__(self, scene: Scene, device: str, grad_sigma: bool = False, grad_albedo: bool = False):
        self.scene = scene
        self.device = device
        self.keep = []
        m = L.VrMedium()
        sig = self._buf(scene.sigma)
        self.sigma = sig
        self.sigma_grad = self._zeros_like(sig) if grad_sigma else None
        m.sigma = L.VrGrid()
        m.sigma.t = self._view(sig, self.sigma_grad)
        m.sigma.bmin = L.f3(scene.bmin)
        m.sigma.bmax = L.f3(scene.bmax)
        m.box_min = L.f3(scene.box_min or scene.bmin)
        m.box_max = L.f3(scene.box_max or scene.bmax)
        self.albedo_grad = None
        if scene.albedo is not None:
            a = self._buf(scene.albedo)
            self.albedo = a
            self.albedo_grad = self._zeros_like(a) if grad_albedo else None
            m.albedo_kind = L.VR_FIELD_GRID
            m.albedo = L.VrGrid()
            m.albedo.t = self._view(a, self.albedo_grad)
            m.albedo.bmin = L.f3(scene.bmin)
            m.albedo.bmax = L.f3(scene.bmax)
        else:
            m.albedo_kind = L.VR_FIELD_CONST
        m.albedo_c = L.f3(scene.albedo_c)
        m.emission_c = L.f3(scene.emission_c)
        m.env_c = L.f3(scene.env_c)
        m.sun_dir = L.f3(scene.sun_dir)
        m.sun_radiance = L.f3(scene.sun_radiance)
        m.sun_prob = scene.sun_prob
        m.phase_kind = scene.phase_kind
        m.phase_g = scene.phase_g
        if scene.phase_g_grid is not None:
            pg = self._buf(scene.phase_g_grid)
            self.phase_g_grid = pg
            m.phase_g_kind = L.VR_FIELD_GRID
            m.phase_g_grid = L.VrGrid()
            m.phase_g_grid.t = self._view(pg, None)
            m.phase_g_grid.bmin = L.f3(scene.bmin)
            m.phase_g_grid.bmax = L.f3(scene.bmax)
        if scene.majorant == "grid":
            res = np.zeros(scene.majorant_res, dtype=np.float32)
            g = L.VrGrid()
            g.t = L.VrTensor()
            g.t = _np_view(scene.sigma)
            g.bmin = L.f3(scene.bmin)
            g.bmax = L.f3(scene.bmax)
            r = (L.ctypes.c_int32 * 3)(*scene.majorant_res)
            O.check(O.load().oracle_majorant_build(g, r, O.nptr(res)), "majorant_build")
            self.majorant_grid_np = res
            mg = self._buf(res)
            m.majorant_kind = L.VR_MAJ_GRID
            m.majorant_grid = self._addr(mg)
            for i in range(3):
                m.majorant_res[i] = scene.majorant_res[i]
        else:
            m.majorant_kind = L.VR_MAJ_CONST
            m.majorant_value = float(scene.sigma.max()) if scene.majorant_value is None else scene.majorant_value
            m.majorant_distance = scene.majorant_distance
        m.transmittance_kind = scene.transmittance_kind
        m.ds_epsilon = scene.ds_epsilon
        m.max_bounces = scene.max_bounces
        m.medium_filter = scene.medium_filter
        m.gathering_c = L.f3(scene.gathering_c)
        m.medium_majorant = scene.medium_majorant
        if scene.env_image is not None:
            img = self._buf(scene.env_image)
            m.env_kind = L.VR_ENV_XR
            v = L.VrTensor()
            v.data = self._addr(img)
            v.ndim = 3
            for i in range(3):
                v.shape[i] = img.shape[i]
                v.stride[i] = int(np.prod(img.shape[i + 1:]))
            v.shape[3] = 1
            v.stride[3] = 1
            m.env_image = v
            m.env_image_bmin[0] = m.env_image_bmin[1] = -1.0
            m.env_image_bmax[0] = m.env_image_bmax[1] = 1.0
        if scene.env_quadtree is not None:
            q = self._buf(scene.env_quadtree)
            m.env_sampler_kind = L.VR_ENVS_QUADTREE
            m.env_quadtree = self._addr(q)
            m.env_quadtree_levels = scene.env_quadtree_levels
        self.medium = m

    def _buf(self, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=np.float32)
        if self.device == "cpu":
            self.keep.append(a)
            return a
        t = torch.from_numpy(a).to(self.device)
        self.keep.append(t)
        return t

    def _zeros_like(self, b):
        z = np.zeros_like(b) if self.device == "cpu" else torch.zeros_like(b)
        self.keep.append(z)
        return z

    def _addr(self, b) -> int:
        return O.nptr(b) if self.device == "cpu" else b.data_ptr()

    def _view(self, b, g) -> L.VrTensor:
        v = L.VrTensor()
        v.data = self._addr(b)
        v.grad = None if g is None else self._addr(g)
        v.ndim = 4
        for i in range(4):
            v.shape[i] = b.shape[i]
            v.stride[i] = int(np.prod(b.shape[i + 1:]))
        return v

    def grads(self):
        def host(x):
            if x is None:
                return None
            return x.copy() if self.device == "cpu" else x.detach().cpu().numpy()
        return host(self.sigma_grad), host(self.albedo_grad)


def morton_pixels(n: int) -> typing.Tuple[np.ndarray, np.ndarray]:
    """(px, py) of Morton indices 0..n-1: the even and the odd bits."""
    i = np.arange(n, dtype=np.int64)
    px = np.zeros(n, np.int64)
    py = np.zeros(n, np.int64)
    for b in range(16):
        px |= ((i >> (2 * b)) & 1) << b
        py |= ((i >> (2 * b + 1)) & 1) << b
    return px, py


def np_quadtree(density: np.ndarray) -> np.ndarray:
    """The quadtree of create_density_quadtree (_functions.py:509-530) in numpy:
    a normalised [R, R] pixel density, its 2x2 block sums level by level, each
    level in Morton order, levels 2x2 .. RxR concatenated."""
    R = density.shape[0]
    levels = []
    d = density.astype(np.float64)
    while d.shape[0] >= 2:
        n = d.shape[0]
        px, py = morton_pixels(n * n)
        levels.append(d[py, px])
        d = d.reshape(n // 2, 2, n // 2, 2).sum(axis=(1, 3))
    return np.concatenate(levels[::-1]).astype(np.float32)


def env_density(img: np.ndarray) -> np.ndarray:
    """Solid-angle-weighted, normalised luminance sum of an [R, R, 3] image
    (the quadtree build of _objects.py:184-192 at the image's own resolution)."""
    R = img.shape[0]
    py = np.arange(R)
    w = np.cos(py * np.pi / R) - np.cos((py + 1) * np.pi / R)
    d = img.sum(-1).astype(np.float64) * w[:, None]
    return d / d.sum()


def sky_image(R: int, seed: int = 0) -> np.ndarray:
    """A smooth sky [R, R, 3] with a bright sun-like blob (equirectangular)."""
    rng = np.random.default_rng(seed)
    v, u = np.meshgrid((np.arange(R) + 0.5) / R, (np.arange(R) + 0.5) / R, indexing="ij")
    base = 0.2 + 0.3 * (1 - v)[..., None] * np.array([0.5, 0.7, 1.0])
    cu, cv = rng.uniform(0.2, 0.8), rng.uniform(0.2, 0.45)
    blob = 20.0 * np.exp(-((u - cu) ** 2 + (v - cv) ** 2) / 0.003)[..., None] * np.array([1.0, 0.9, 0.7])
    return (base + blob).astype(np.float32)


def _np_view(a: np.ndarray) -> L.VrTensor:
    v = L.VrTensor()
    v.data = O.nptr(a)
    v.grad = None
    v.ndim = 4
    for i in range(4):
        v.shape[i] = a.shape[i]
        v.stride[i] = int(np.prod(a.shape[i + 1:]))
    return v


def np_grid(a: np.ndarray, bmin=(-0.5, -0.5, -0.5), bmax=(0.5, 0.5, 0.5), grad: typing.Optional[np.ndarray] = None) -> L.VrGrid:
    g = L.VrGrid()
    g.t = _np_view(a)
    if grad is not None:
        g.t.grad = O.nptr(grad)
    g.bmin = L.f3(bmin)
    g.bmax = L.f3(bmax)
    return g


class Camera:
    def __init__(self, poses: np.ndarray, width: int, height: int, jittered: bool = False, fov: float = math.pi / 4,
                 sensors: typing.Optional[np.ndarray] = None, device: str = "cpu"):
        self.keep = []
        c = L.VrCamera()
        p = np.ascontiguousarray(poses, dtype=np.float32)
        if device == "cpu":
            self.keep.append(p)
            c.poses = O.nptr(p)
        else:
            t = torch.from_numpy(p).to(device)
            self.keep.append(t)
            c.poses = t.data_ptr()
        c.num_poses = p.shape[0]
        c.width = width
        c.height = height
        c.jittered = int(jittered)
        c.fov = fov
        c.znear = 0.001
        if sensors is not None:
            s = np.ascontiguousarray(sensors, dtype=np.int64)
            if device == "cpu":
                self.keep.append(s)
                c.sensors = O.nptr(s)
            else:
                t = torch.from_numpy(s).to(device)
                self.keep.append(t)
                c.sensors = t.data_ptr()
            c.num_sensors = s.shape[0]
        self.c = c
        self.threads = sensors.shape[0] if sensors is not None else p.shape[0] * width * height


def pose(origin, target=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0)) -> np.ndarray:
    o = np.asarray(origin, dtype=np.float64)
    d = np.asarray(target, dtype=np.float64) - o
    d /= np.linalg.norm(d)
    return np.array([*o, *d, *up], dtype=np.float32)


def blob_grid(n: int, seed: int = 0, max_value: float = 20.0) -> np.ndarray:
    from rendervous_amd.functions import gaussian_blobs
    return gaussian_blobs(n, seed=seed, max_value=max_value).numpy()


def fbm_grid(n: int, seed: int = 0, scale: float = 40.0) -> np.ndarray:
    from rendervous_amd.functions import fbm_value_noise
    return fbm_value_noise(n, seed=seed, scale=scale).numpy()


def random_rays(n: int, seed: int = 0, radius: float = 1.2, inside: bool = False) -> np.ndarray:
    """Rays [n,6]: origins on a sphere (or inside the box) aimed near the centre."""
    rng = np.random.default_rng(seed)
    if inside:
        o = rng.uniform(-0.45, 0.45, size=(n, 3))
        d = rng.normal(size=(n, 3))
    else:
        u = rng.normal(size=(n, 3))
        o = u / np.linalg.norm(u, axis=1, keepdims=True) * radius
        t = rng.uniform(-0.3, 0.3, size=(n, 3))
        d = t - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return np.concatenate([o, d], axis=1).astype(np.float32)


def oracle_map_fw(scene: Scene, kind: int, rays: np.ndarray, seeds, bw_mode=L.VR_BW_REFERENCE) -> np.ndarray:
    b = Built(scene, "cpu")
    ind, outd = L.MAP_DIMS[kind]
    x = np.ascontiguousarray(rays, dtype=np.float32)
    out = np.zeros((x.shape[0], outd), dtype=np.float32)
    lau = L.launch(seeds, bw_mode=bw_mode, seed_mode=L.VR_SEED_MAP)
    O.check(O.load().oracle_map_fw(b.medium, kind, O.nptr(x), x.shape[0], O.nptr(out), lau), "oracle_map_fw")
    return out


def oracle_capture_fw(scene: Scene, kind: int, cam_args: dict, seeds, spp: int, bw_mode=L.VR_BW_REFERENCE,
                      debug: bool = False):
    b = Built(scene, "cpu")
    cam = Camera(device="cpu", **cam_args)
    _, outd = L.MAP_DIMS[kind]
    out = np.zeros((cam.threads, outd), dtype=np.float32)
    dbg = np.zeros(cam.threads, dtype=np.int32) if debug else None
    lau = L.launch(seeds, samples=spp, bw_mode=bw_mode)
    if dbg is not None:
        lau.debug_out = O.nptr(dbg)
    O.check(O.load().oracle_capture_fw(b.medium, kind, cam.c, lau, O.nptr(out)), "oracle_capture_fw")
    return (out, dbg) if debug else out


def oracle_capture_bw(scene: Scene, kind: int, cam_args: dict, seeds, spp: int, out_grad: np.ndarray,
                      bw_mode=L.VR_BW_CONSISTENT, seed_mode=L.VR_SEED_CAPTURE_FW, grad_albedo=False):
    b = Built(scene, "cpu", grad_sigma=True, grad_albedo=grad_albedo)
    cam = Camera(device="cpu", **cam_args)
    og = np.ascontiguousarray(out_grad, dtype=np.float32)
    lau = L.launch(seeds, samples=spp, bw_mode=bw_mode, seed_mode=seed_mode)
    O.check(O.load().oracle_capture_bw(b.medium, kind, cam.c, lau, O.nptr(og)), "oracle_capture_bw")
    return b.grads()


# ---------------------------------------------------------------- GPU side --
def _stream():
    return torch.cuda.current_stream().cuda_stream


def gpu_map_fw(scene: Scene, kind: int, rays: np.ndarray, seeds, bw_mode=L.VR_BW_REFERENCE) -> np.ndarray:
    b = Built(scene, "cuda")
    _, outd = L.MAP_DIMS[kind]
    x = torch.from_numpy(np.ascontiguousarray(rays, dtype=np.float32)).cuda()
    out = torch.zeros(x.shape[0], outd, device="cuda")
    lau = L.launch(seeds, bw_mode=bw_mode, seed_mode=L.VR_SEED_MAP)
    L.check(L.load().vr_map_fw(b.medium, kind, x.data_ptr(), x.shape[0], out.data_ptr(), lau, _stream()), "vr_map_fw")
    torch.cuda.synchronize()
    return out.cpu().numpy()


def gpu_map_bw(scene: Scene, kind: int, rays: np.ndarray, out_grad: np.ndarray, seeds, bw_mode=L.VR_BW_REFERENCE,
               grad_albedo=False):
    b = Built(scene, "cuda", grad_sigma=True, grad_albedo=grad_albedo)
    x = torch.from_numpy(np.ascontiguousarray(rays, dtype=np.float32)).cuda()
    og = torch.from_numpy(np.ascontiguousarray(out_grad, dtype=np.float32)).cuda()
    lau = L.launch(seeds, bw_mode=bw_mode, seed_mode=L.VR_SEED_MAP)
    L.check(L.load().vr_map_bw(b.medium, kind, x.data_ptr(), og.data_ptr(), x.shape[0], lau, _stream()), "vr_map_bw")
    torch.cuda.synchronize()
    return b.grads()


def oracle_map_bw(scene: Scene, kind: int, rays: np.ndarray, out_grad: np.ndarray, seeds, bw_mode=L.VR_BW_REFERENCE,
                  grad_albedo=False):
    b = Built(scene, "cpu", grad_sigma=True, grad_albedo=grad_albedo)
    x = np.ascontiguousarray(rays, dtype=np.float32)
    og = np.ascontiguousarray(out_grad, dtype=np.float32)
    lau = L.launch(seeds, bw_mode=bw_mode, seed_mode=L.VR_SEED_MAP)
    O.check(O.load().oracle_map_bw(b.medium, kind, O.nptr(x), O.nptr(og), x.shape[0], lau), "oracle_map_bw")
    return b.grads()


def gpu_capture_fw(scene: Scene, kind: int, cam_args: dict, seeds, spp: int, bw_mode=L.VR_BW_REFERENCE, debug=False):
    b = Built(scene, "cuda")
    cam = Camera(device="cuda", **cam_args)
    _, outd = L.MAP_DIMS[kind]
    out = torch.zeros(cam.threads, outd, device="cuda")
    dbg = torch.zeros(cam.threads, dtype=torch.int32, device="cuda") if debug else None
    lau = L.launch(seeds, samples=spp, bw_mode=bw_mode, debug_out=dbg)
    L.check(L.load().vr_capture_fw(b.medium, kind, cam.c, lau, out.data_ptr(), _stream()), "vr_capture_fw")
    torch.cuda.synchronize()
    if debug:
        return out.cpu().numpy(), dbg.cpu().numpy()
    return out.cpu().numpy()


def gpu_capture_bw(scene: Scene, kind: int, cam_args: dict, seeds, spp: int, out_grad: np.ndarray,
                   bw_mode=L.VR_BW_CONSISTENT, seed_mode=L.VR_SEED_CAPTURE_FW, grad_albedo=False):
    b = Built(scene, "cuda", grad_sigma=True, grad_albedo=grad_albedo)
    cam = Camera(device="cuda", **cam_args)
    og = torch.from_numpy(np.ascontiguousarray(out_grad, dtype=np.float32)).cuda()
    lau = L.launch(seeds, samples=spp, bw_mode=bw_mode, seed_mode=seed_mode)
    L.check(L.load().vr_capture_bw(b.medium, kind, cam.c, lau, og.data_ptr(), _stream()), "vr_capture_bw")
    torch.cuda.synchronize()
    return b.grads()


def match_fraction(a: np.ndarray, b: np.ndarray, rtol=1e-3, atol=1e-5) -> float:
    """Fraction of rows whose entries all agree within tolerance."""
    ok = np.all(np.abs(a - b) <= atol + rtol * np.abs(b), axis=-1)
    return float(ok.mean())


# Relative L1 bound of seed-matched gradients (GPU against the oracle on the
# same random streams): the draws are identical, so the grids differ only by
# float rounding (contractions into FMAs, the order of the atomic adds) and the
# rare branch that rounding flips; measured 5e-8 .. 7e-6 over the parity cases.
SEED_MATCHED_L1 = 1e-4


def rel_l1(a: np.ndarray, b: np.ndarray) -> float:
    r = float(np.abs(a - b).sum() / max(np.abs(b).sum(), 1e-30))
    print(f"[rel_l1] {r:.3e}", flush=True)
    return r


# ------------------------------------------------------------------ scenes --
def box_mesh(half: float = 1.0, center=(0.0, 0.0, 0.0)):
    """positions [24,3], indices [12,3] of an axis-aligned cube with outward
    winding (cross(p1 - p0, p2 - p0) points out of the box)."""
    pos, idx = [], []
    c = np.asarray(center, np.float64)
    for axis in range(3):
        for side in (-1.0, 1.0):
            b = len(pos)
            for a, bb in ((-1, -1), (1, -1), (1, 1), (-1, 1)):
                p = np.zeros(3)
                p[axis] = side
                p[(axis + 1) % 3] = a
                p[(axis + 2) % 3] = bb
                pos.append(c + half * p)
            for t in ((b, b + 1, b + 2), (b, b + 2, b + 3)):
                p0, p1, p2 = (np.asarray(pos[i]) - c for i in t)
                n = np.cross(p1 - p0, p2 - p0)
                idx.append(t if np.dot(n, p0 + p1 + p2) > 0 else (t[0], t[2], t[1]))
    return np.asarray(pos, np.float32), np.asarray(idx, np.int32)


def quad_mesh(y: float = 0.0, half: float = 50.0, up: bool = True):
    """A large square at height y whose geometric normal points +y (up=True) or -y."""
    pos = np.array([[-half, y, -half], [half, y, -half], [half, y, half], [-half, y, half]], np.float32)
    idx = np.array([[0, 2, 1], [0, 3, 2]], np.int32) if up else np.array([[0, 1, 2], [0, 2, 3]], np.int32)
    return pos, idx


def patch(sampler=L.VR_SURF_PASS, emission=L.VR_EMIT_NONE, emission_c=(0.0, 0.0, 0.0), inside=-1,
          refraction_index=1.5) -> dict:
    return dict(sampler=sampler, emission=emission, emission_c=emission_c, inside=inside,
                refraction_index=refraction_index)


class SceneBuilt:
    """A vr_scene over host (device="cpu", for the oracle) or device buffers.

    meshes: list of (positions [V,3], indices [T,3]) -- one patch each; patches:
    one dict (patch()) per mesh; media: helpers.Scene media (DTMediumPathIntegrator
    fields: sigma grid, albedo, phase, medium_majorant, medium_filter)."""

    def __init__(self, meshes, patches, media=(), env_c=(1.0, 1.0, 1.0), env_image=None, normals=None, uvs=None,
                 device="cpu", grad_sigma=False, grad_albedo=False):
        self.device = device
        self.keep = []
        pos, idx, tp = [], [], []
        base = 0
        for k, (p, ix) in enumerate(meshes):
            pos.append(np.asarray(p, np.float32))
            idx.append(np.asarray(ix, np.int64) + base)
            tp.append(np.full(len(ix), k, np.int32))
            base += len(p)
        positions = np.ascontiguousarray(np.concatenate(pos) if pos else np.zeros((0, 3), np.float32))
        indices = np.ascontiguousarray((np.concatenate(idx) if idx else np.zeros((0, 3))).astype(np.int32))
        tri_patch = np.ascontiguousarray(np.concatenate(tp) if tp else np.zeros(0, np.int32))
        T = indices.shape[0]
        nodes = np.zeros((max(1, 2 * T), 8), np.float32)
        tri_ids = np.zeros(max(1, T), np.int32)
        tri_verts = np.zeros((max(1, T), 9), np.float32)
        nn = L.ctypes.c_int32(0)
        if T:
            L.check(L.load().vr_scene_build_bvh(positions.ctypes.data, positions.shape[0], indices.ctypes.data, T,
                                                 nodes.ctypes.data, nodes.shape[0], tri_ids.ctypes.data,
                                                 tri_verts.ctypes.data, L.ctypes.byref(nn)), "build_bvh")
        self.num_nodes = int(nn.value)
        self.nodes_np, self.tri_ids_np, self.tri_verts_np = nodes, tri_ids, tri_verts
        table = (L.VrPatch * max(1, len(patches)))()
        for i, p in enumerate(patches):
            table[i].sampler = p["sampler"]
            table[i].refraction_index = p["refraction_index"]
            table[i].emission = p["emission"]
            table[i].emission_c = L.f3(p["emission_c"])
            table[i].inside_medium = p["inside"]
            table[i].outside_medium = -1
        ptab = np.frombuffer(bytes(table), dtype=np.int32).copy()
        self.media = [Built(m, device, grad_sigma=grad_sigma, grad_albedo=grad_albedo) for m in media]
        sc = L.VrScene()
        sc.nodes = self._addr(self._buf(nodes[:max(1, self.num_nodes)], np.float32))
        sc.tri_verts = self._addr(self._buf(tri_verts, np.float32))
        sc.tri_ids = self._addr(self._buf(tri_ids, np.int32))
        sc.num_nodes = self.num_nodes
        sc.num_tris = T
        sc.positions = self._addr(self._buf(positions, np.float32)) if T else None
        sc.normals = None if normals is None else self._addr(self._buf(normals, np.float32))
        sc.uvs = None if uvs is None else self._addr(self._buf(uvs, np.float32))
        sc.indices = self._addr(self._buf(indices, np.int32)) if T else None
        sc.tri_patch = self._addr(self._buf(tri_patch, np.int32)) if T else None
        sc.patches = self._addr(self._buf(ptab, np.int32))
        sc.num_patches = len(patches)
        sc.num_vertices = positions.shape[0]
        self.media_array = (L.VrMedium * max(1, len(self.media)))(*[b.medium for b in self.media])
        self.keep.append(self.media_array)
        sc.media = L.ctypes.cast(self.media_array, L.ctypes.c_void_p).value if self.media else None
        sc.num_media = len(self.media)
        sc.env_c = L.f3(env_c)
        if env_image is not None:
            img = self._buf(env_image, np.float32)
            sc.env_kind = L.VR_ENV_XR
            v = L.VrTensor()
            v.data = self._addr(img)
            v.ndim = 3
            for i in range(3):
                v.shape[i] = img.shape[i]
                v.stride[i] = int(np.prod(img.shape[i + 1:]))
            v.shape[3] = v.stride[3] = 1
            sc.env_image = v
            sc.env_image_bmin[0] = sc.env_image_bmin[1] = -1.0
            sc.env_image_bmax[0] = sc.env_image_bmax[1] = 1.0
        else:
            sc.env_kind = L.VR_ENV_CONST
        self.scene = sc

    def _buf(self, a, dtype):
        a = np.ascontiguousarray(a, dtype=dtype)
        if self.device == "cpu":
            self.keep.append(a)
            return a
        t = torch.from_numpy(a).to(self.device)
        self.keep.append(t)
        return t

    def _addr(self, b) -> int:
        return O.nptr(b) if self.device == "cpu" else b.data_ptr()

    def grads(self):
        return [b.grads() for b in self.media]


def scene_medium(sigma: np.ndarray, majorant: float, albedo_c=(1.0, 1.0, 1.0), phase_g=0.0, medium_filter=3,
                 gathering_c=(0.0, 0.0, 0.0), albedo=None, bmin=(-0.5, -0.5, -0.5), bmax=(0.5, 0.5, 0.5)) -> Scene:
    """A DTMediumPathIntegrator medium for SceneBuilt."""
    return Scene(sigma=sigma, bmin=bmin, bmax=bmax, box_min=(-1e30,) * 3, box_max=(1e30,) * 3, albedo_c=albedo_c,
                 albedo=albedo, phase_kind=L.VR_PHASE_HG, phase_g=phase_g, medium_majorant=majorant,
                 medium_filter=medium_filter, gathering_c=gathering_c)


def scene_raycast(sb: SceneBuilt, rays: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(rays, np.float32)
    if sb.device == "cpu":
        out = np.zeros((x.shape[0], 16), np.float32)
        O.check(O.load().oracle_scene_raycast(sb.scene, O.nptr(x), x.shape[0], O.nptr(out)), "oracle_scene_raycast")
        return out
    xt = torch.from_numpy(x).cuda()
    out = torch.zeros(x.shape[0], 16, device="cuda")
    L.check(L.load().vr_scene_raycast(sb.scene, xt.data_ptr(), x.shape[0], out.data_ptr(), _stream()), "vr_scene_raycast")
    torch.cuda.synchronize()
    return out.cpu().numpy()


def scene_map_fw(sb: SceneBuilt, rays: np.ndarray, seeds) -> np.ndarray:
    x = np.ascontiguousarray(rays, np.float32)
    lau = L.launch(seeds, bw_mode=L.VR_BW_CONSISTENT, seed_mode=L.VR_SEED_MAP)
    if sb.device == "cpu":
        out = np.zeros((x.shape[0], 3), np.float32)
        O.check(O.load().oracle_scene_map_fw(sb.scene, O.nptr(x), x.shape[0], O.nptr(out), lau), "oracle_scene_map_fw")
        return out
    xt = torch.from_numpy(x).cuda()
    out = torch.zeros(x.shape[0], 3, device="cuda")
    L.check(L.load().vr_scene_map_fw(sb.scene, xt.data_ptr(), x.shape[0], out.data_ptr(), lau, _stream()), "vr_scene_map_fw")
    torch.cuda.synchronize()
    return out.cpu().numpy()


def scene_capture_fw(sb: SceneBuilt, cam_args: dict, seeds, spp: int) -> np.ndarray:
    cam = Camera(device=sb.device, **cam_args)
    lau = L.launch(seeds, samples=spp, bw_mode=L.VR_BW_CONSISTENT)
    if sb.device == "cpu":
        out = np.zeros((cam.threads, 3), np.float32)
        O.check(O.load().oracle_scene_capture_fw(sb.scene, cam.c, lau, O.nptr(out)), "oracle_scene_capture_fw")
        return out
    out = torch.zeros(cam.threads, 3, device="cuda")
    L.check(L.load().vr_scene_capture_fw(sb.scene, cam.c, lau, out.data_ptr(), _stream()), "vr_scene_capture_fw")
    torch.cuda.synchronize()
    return out.cpu().numpy()


def scene_capture_bw(sb: SceneBuilt, cam_args: dict, seeds, spp: int, out_grad: np.ndarray,
                     seed_mode=L.VR_SEED_CAPTURE_FW):
    cam = Camera(device=sb.device, **cam_args)
    lau = L.launch(seeds, samples=spp, bw_mode=L.VR_BW_CONSISTENT, seed_mode=seed_mode)
    og = np.ascontiguousarray(out_grad, np.float32)
    if sb.device == "cpu":
        O.check(O.load().oracle_scene_capture_bw(sb.scene, cam.c, lau, O.nptr(og)), "oracle_scene_capture_bw")
    else:
        ogt = torch.from_numpy(og).cuda()
        L.check(L.load().vr_scene_capture_bw(sb.scene, cam.c, lau, ogt.data_ptr(), _stream()), "vr_scene_capture_bw")
        torch.cuda.synchronize()
    return sb.grads()
