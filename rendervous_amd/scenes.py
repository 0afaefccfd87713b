"""Scene API: triangle-mesh geometry, surface materials, media and the
scene-level path tracer (rendervous _scenes.py).

* DTMediumPathIntegrator (_scenes.py:1614-1632) maps a segment (x, w, d) inside
  a medium to (xo, wo, W, A) by delta tracking with a scalar majorant
  (include/sc_medium_path_integrator_DT_fw.h / _bw.h).
* MeshGeometry / GroupGeometry / TransformedGeometry (_scenes.py:93-605) are
  the 6 -> 16 geometry maps (ray -> t, N, G, C, T, B, patch).  The reference
  builds Vulkan acceleration structures; here the library builds a BVH
  (vr_scene_build_bvh) that its kernels traverse.
* Scene.pathtrace() / PathtracedScene (_scenes.py:2251-2423,
  include/sc_path_tracing.h) trace paths through surfaces (Lambert, mirror,
  Fresnel, null and blocking samplers; isotropic / Lambert emission) and the
  media on the inner side of patches, with gradients into the media grids.
"""
from __future__ import annotations

import ctypes
import math
import typing

import numpy as np
import torch

from . import _internal
from . import _lib as L
from .maps import (CompositionMap, ConstantMap, Grid3D, Image2D, MapBase, VHGPhaseSampler, XRProjection, ZERO, _Lowered,
                   _RayMap, _functionsampler, UniformDirectionSampler)

__all__ = ["HGPhaseSampler", "DTMediumPathIntegrator", "Geometry", "MeshGeometry", "GroupGeometry", "TransformedGeometry",
           "MeshBoundary", "SurfaceScatteringSampler", "LambertSurfaceScatteringSampler", "MirrorSurfaceScatteringSampler",
           "FresnelSurfaceScatteringSampler", "NullSurfaceScatteringSampler", "NoSurfaceScatteringSampler",
           "SurfaceEmission", "IsotropicSurfaceEmission", "LambertSurfaceEmission", "NoSurfaceEmission", "SurfaceMaterial",
           "EnvironmentSampler", "UniformEnvironmentSampler", "MediumMaterial", "HeterogeneousMediumMaterial",
           "PathtracedScene", "Scene"]


class HGPhaseSampler(MapBase):
    """Single Henyey-Greenstein sampler win -> (wout, weight=1, pdf) (_scenes.py:1984-2004)."""

    def __init__(self, g: torch.Tensor):
        super().__init__(3, 5)
        self.g = g if isinstance(g, torch.nn.Parameter) else torch.nn.Parameter(torch.as_tensor(g, dtype=torch.float32).reshape(-1),
                                                                                requires_grad=False)

    def g_value(self) -> float:
        return float(self.g.detach().reshape(-1)[0].cpu())


class DTMediumPathIntegrator(_RayMap):
    """(x, w, d) -> (xo, wo, W, A) (_scenes.py:1614-1632, medium_path_integrator_interface.h:1-46).

    medium_filter bit 0 selects scattering, bit 1 accumulation (gathering); the
    gathering map must be a constant here.
    """

    def __init__(self, sigma: MapBase, majorant: torch.Tensor, scattering_albedo: typing.Optional[MapBase] = None,
                 phase_sampler: typing.Optional[MapBase] = None, gathering: typing.Optional[MapBase] = None,
                 medium_filter: int = 3):
        if gathering is None:
            medium_filter &= 1
        if scattering_albedo is None or phase_sampler is None:
            medium_filter &= 2
            scattering_albedo = ZERO
            phase_sampler = None
        super().__init__(7, 12)
        self.sigma = sigma.cast(3, 1)
        self.scattering_albedo = scattering_albedo.cast(3, 3)
        self.phase_sampler = phase_sampler
        self.gathering = gathering
        self.majorant = majorant if isinstance(majorant, torch.nn.Parameter) else torch.nn.Parameter(
            torch.as_tensor(majorant, dtype=torch.float32).reshape(-1), requires_grad=False)
        self.medium_filter = medium_filter

    def _lower_map(self):
        lw = _Lowered()
        lw.set_sigma(self.sigma)
        lw.set_albedo(self.scattering_albedo)
        m = lw.medium
        # the 7->12 integrator is unbounded by a box: it tracks over [0, d]
        m.box_min = L.f3([-1e30, -1e30, -1e30])
        m.box_max = L.f3([1e30, 1e30, 1e30])
        m.medium_majorant = float(self.majorant.detach().reshape(-1)[0].cpu())
        m.medium_filter = int(self.medium_filter)
        if self.gathering is not None:
            if not isinstance(self.gathering, ConstantMap):
                raise NotImplementedError("gathering must be a constant map")
            m.gathering_c = L.f3(self.gathering.vec(3))
        ps = self.phase_sampler
        if ps is None:
            m.phase_kind = L.VR_PHASE_ISO
        elif isinstance(ps, HGPhaseSampler):
            m.phase_kind = L.VR_PHASE_HG
            m.phase_g = ps.g_value()
        elif isinstance(ps, VHGPhaseSampler) and isinstance(ps.phase_g, ConstantMap):
            m.phase_kind = L.VR_PHASE_HG_MIX if ps.mixture else L.VR_PHASE_HG
            m.phase_g = ps.phase_g.vec(1)[0]
        else:
            raise NotImplementedError("phase sampler must be an HG sampler")
        return L.VR_MAP_MEDIUM_DT, lw


# ------------------------------------------------------------- geometry --
def _as_f32(t) -> torch.Tensor:
    return torch.as_tensor(t).detach().to(torch.float32)


class Geometry(MapBase):
    """Ray -> hit map (6 -> 16: t, N, G, C, T, B, patch index) over patches of
    triangles (_scenes.py:93-190)."""

    def __init__(self):
        super().__init__(6, 16)

    @staticmethod
    def signature():
        return (6, 16)

    def __len__(self) -> int:
        return len(self._patches())

    def _patches(self) -> typing.List[typing.Tuple[torch.Tensor, typing.Optional[torch.Tensor],
                                                   typing.Optional[torch.Tensor], torch.Tensor]]:
        """(positions [V,3], normals [V,3] | None, uvs [V,2] | None, indices [T,3]) per patch, world space."""
        raise NotImplementedError()

    def _version_key(self) -> tuple:
        raise NotImplementedError()

    def get_boundary(self) -> "MeshBoundary":
        return MeshBoundary(self)

    def transformed(self, T: torch.Tensor) -> "TransformedGeometry":
        return TransformedGeometry(self, T)

    def update_ads_if_necessary(self) -> bool:
        """The BVH is rebuilt lazily when a position tensor changes (TensorCheck semantics)."""
        return False

    def forward(self, *args):
        x, = args
        buffers = _GeometryBuffers.of(self, x.device)
        sc = buffers.scene_struct()
        xc = x.detach().to(torch.float32).contiguous()
        n = xc.numel() // 6
        out = torch.empty(*x.shape[:-1], 16, device=x.device, dtype=torch.float32)
        L.check(L.load().vr_scene_raycast(sc, L.ptr(xc), n, L.ptr(out), L.stream_handle(x.device)), "vr_scene_raycast")
        return out


class MeshGeometry(Geometry):
    """A triangle mesh as one patch (_scenes.py:229-505)."""

    def __init__(self, positions: torch.Tensor, normals: typing.Optional[torch.Tensor] = None,
                 uvs: typing.Optional[torch.Tensor] = None, indices: typing.Optional[torch.Tensor] = None,
                 normalize: bool = False, compute_normals: bool = True):
        super().__init__()
        positions = _as_f32(positions)
        if indices is None:
            assert positions.shape[0] % 3 == 0, "without indices the positions form consecutive triangles"
            indices = torch.arange(positions.shape[0], dtype=torch.int32).view(-1, 3)
        indices = torch.as_tensor(indices).detach().to(torch.int32)
        if normalize:
            bmin, _ = positions.min(dim=0)
            bmax, _ = positions.max(dim=0)
            positions = (positions - bmin - (bmax - bmin) * 0.5) * 2 / (bmax - bmin).max()
        if normals is None and compute_normals:
            normals = MeshGeometry.compute_normals(positions, indices)
        self.positions = torch.nn.Parameter(positions, requires_grad=False)
        self.normals = None if normals is None else torch.nn.Parameter(_as_f32(normals), requires_grad=False)
        self.uvs = None if uvs is None else torch.nn.Parameter(_as_f32(uvs), requires_grad=False)
        self.indices = indices

    @staticmethod
    def compute_normals(positions: torch.Tensor, indices: torch.Tensor) -> torch.Tensor:
        """Vertex normals: sums of the faces' cross(normalised edge 1, normalised edge 2) (_scenes.py:242-258)."""
        positions = _as_f32(positions)
        idx = torch.as_tensor(indices).long()
        p0, p1, p2 = positions[idx[:, 0]], positions[idx[:, 1]], positions[idx[:, 2]]
        v1 = torch.nn.functional.normalize(p1 - p0, dim=-1)
        v2 = torch.nn.functional.normalize(p2 - p0, dim=-1)
        n = torch.cross(v1, v2, dim=-1)
        normals = torch.zeros_like(positions)
        for c in range(3):
            normals.index_add_(0, idx[:, c], n)
        return normals / torch.sqrt((normals ** 2).sum(-1, keepdim=True))

    @staticmethod
    def box() -> "MeshGeometry":
        """The [-1, 1]^3 cube with outward faces (_scenes.py:270-418: the unit
        box normalised)."""
        pos = []
        for axis in range(3):
            for side in (-0.5, 0.5):
                quad = []
                for a in (-0.5, 0.5):
                    for b in (-0.5, 0.5):
                        p = [0.0, 0.0, 0.0]
                        p[axis] = side
                        p[(axis + 1) % 3] = a
                        p[(axis + 2) % 3] = b
                        quad.append(p)
                pos.extend(quad)
        positions = torch.tensor(pos, dtype=torch.float32)
        tris = []
        for f in range(6):
            base = 4 * f
            q = [base, base + 1, base + 3, base + 2]
            for t in ((q[0], q[1], q[2]), (q[0], q[2], q[3])):
                p0, p1, p2 = (positions[i] for i in t)
                n = torch.cross(p1 - p0, p2 - p0, dim=0)
                centre = (p0 + p1 + p2) / 3
                tris.append(t if float(torch.dot(n, centre)) > 0 else (t[0], t[2], t[1]))
        uvs = torch.tensor([[0.0, 0.0], [0.0, 1.0], [1.0, 0.0], [1.0, 1.0]] * 6)
        return MeshGeometry(positions, uvs=uvs, indices=torch.tensor(tris, dtype=torch.int32), normalize=True)

    @staticmethod
    def load_obj(path: str) -> "MeshGeometry":
        """Positions and faces of a Wavefront OBJ file (polygons fanned into
        triangles), normalised into [-1, 1] along the longest axis (_scenes.py:261-267)."""
        verts, faces = [], []
        with open(path) as f:
            for line in f:
                parts = line.split()
                if not parts:
                    continue
                if parts[0] == "v":
                    verts.append([float(v) for v in parts[1:4]])
                elif parts[0] == "f":
                    ids = [int(p.split("/")[0]) for p in parts[1:]]
                    ids = [i - 1 if i > 0 else len(verts) + i for i in ids]
                    for k in range(1, len(ids) - 1):
                        faces.append([ids[0], ids[k], ids[k + 1]])
        pos = torch.tensor(verts, dtype=torch.float32)
        bmin, _ = pos.min(dim=0)
        bmax, _ = pos.max(dim=0)
        pos = (pos - bmin - (bmax - bmin) * 0.5) * 2 / (bmax - bmin).max()
        return MeshGeometry(positions=pos, indices=torch.tensor(faces, dtype=torch.int32))

    def _patches(self):
        return [(self.positions.detach(), None if self.normals is None else self.normals.detach(),
                 None if self.uvs is None else self.uvs.detach(), self.indices)]

    def _version_key(self):
        return (id(self), self.positions._version, self.positions.data_ptr())


class TransformedGeometry(Geometry):
    """A geometry under a mat4x3 transform [4, 3]: p' = p @ T[:3] + T[3] (_scenes.py:194-226)."""

    def __init__(self, base_geometry: Geometry, initial_transform: typing.Optional[torch.Tensor] = None):
        super().__init__()
        self.base_geometry = base_geometry
        if initial_transform is None:
            initial_transform = torch.tensor([[1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0], [0.0, 0.0, 0.0]])
        self.transform = torch.nn.Parameter(_as_f32(initial_transform).reshape(4, 3), requires_grad=False)

    def _patches(self):
        T = self.transform.detach().cpu()
        R, t = T[:3], T[3]
        nR = torch.linalg.inv(R).T  # normals transform with the inverse transpose
        out = []
        for pos, nrm, uv, idx in self.base_geometry._patches():
            p = pos.cpu() @ R + t
            n = None if nrm is None else torch.nn.functional.normalize(nrm.cpu() @ nR, dim=-1)
            out.append((p, n, uv, idx))
        return out

    def _version_key(self):
        return (id(self), self.transform._version, self.base_geometry._version_key())


class GroupGeometry(Geometry):
    """Patches of several geometries, in order (_scenes.py:507-603)."""

    def __init__(self, *geometries: Geometry):
        super().__init__()
        self.geometries_module_list = torch.nn.ModuleList(geometries)
        self.geometries = geometries

    def _patches(self):
        out = []
        for g in self.geometries:
            out.extend(g._patches())
        return out

    def _version_key(self):
        return (id(self),) + tuple(g._version_key() for g in self.geometries)


class MeshBoundary(MapBase):
    """Boundary view of a geometry (_scenes.py:32-54); the path tracer uses the geometry itself."""

    def __init__(self, geometry: Geometry):
        super().__init__(6, 3)
        self.geometry = geometry

    def __len__(self):
        return len(self.geometry)


class _GeometryBuffers:
    """Device buffers and BVH of a geometry, rebuilt when a position tensor or
    transform changes (MeshGeometry.update_ads_if_necessary, _scenes.py:485-490)."""

    @staticmethod
    def of(geometry: Geometry, device: torch.device) -> "_GeometryBuffers":
        """The geometry's buffers on `device`, cached on the geometry object."""
        key = (geometry._version_key(), str(device))
        b = geometry.__dict__.get("_bvh_cache")
        if b is None or b.key != key:
            b = _GeometryBuffers(geometry, device)
            b.key = key
            object.__setattr__(geometry, "_bvh_cache", b)
        return b

    def __init__(self, geometry: Geometry, device: torch.device):
        patches = geometry._patches()
        pos, nrm, uvs, idx, tri_patch = [], [], [], [], []
        base = 0
        any_n = any(p[1] is not None for p in patches)
        any_uv = any(p[2] is not None for p in patches)
        for k, (p, n, uv, ix) in enumerate(patches):
            p = p.detach().cpu().to(torch.float32).reshape(-1, 3)
            ix = torch.as_tensor(ix).detach().cpu().to(torch.int64).reshape(-1, 3)
            if ix.numel() and (int(ix.min()) < 0 or int(ix.max()) >= p.shape[0]):
                raise ValueError(f"patch {k}: vertex index out of range")
            if any_n:
                if n is None:  # face normals stand in for a patch without normals
                    f = torch.cross(p[ix[:, 1]] - p[ix[:, 0]], p[ix[:, 2]] - p[ix[:, 0]], dim=-1)
                    n = torch.zeros_like(p)
                    for c in range(3):
                        n.index_add_(0, ix[:, c], f)
                    n = torch.nn.functional.normalize(n, dim=-1)
                nrm.append(n.detach().cpu().to(torch.float32).reshape(-1, 3))
            if any_uv:
                uvs.append(torch.zeros(p.shape[0], 2) if uv is None else uv.detach().cpu().to(torch.float32).reshape(-1, 2))
            pos.append(p)
            idx.append(ix + base)
            tri_patch.append(torch.full((ix.shape[0],), k, dtype=torch.int32))
            base += p.shape[0]
        self.num_patches = len(patches)
        positions = torch.cat(pos).contiguous() if pos else torch.zeros(0, 3)
        indices = torch.cat(idx).to(torch.int32).contiguous() if idx else torch.zeros(0, 3, dtype=torch.int32)
        T = indices.shape[0]
        nodes = np.zeros((max(1, 2 * T), 8), dtype=np.float32)
        tri_ids = np.zeros(max(1, T), dtype=np.int32)
        tri_verts = np.zeros((max(1, T), 9), dtype=np.float32)
        n_nodes = ctypes.c_int32(0)
        pn, ixn = positions.numpy(), indices.numpy()
        L.check(L.load().vr_scene_build_bvh(pn.ctypes.data, pn.shape[0], ixn.ctypes.data, T, nodes.ctypes.data,
                                             nodes.shape[0], tri_ids.ctypes.data, tri_verts.ctypes.data,
                                             ctypes.byref(n_nodes)), "vr_scene_build_bvh")
        self.num_nodes = int(n_nodes.value)
        self.num_tris = T
        self.num_vertices = positions.shape[0]
        dev = torch.device(device)
        self.nodes = torch.from_numpy(nodes[:max(1, self.num_nodes)].copy()).to(dev)
        self.tri_ids = torch.from_numpy(tri_ids).to(dev)
        self.tri_verts = torch.from_numpy(tri_verts).to(dev)
        self.positions = positions.to(dev)
        self.normals = torch.cat(nrm).contiguous().to(dev) if any_n else None
        self.uvs = torch.cat(uvs).contiguous().to(dev) if any_uv else None
        self.indices = indices.to(dev)
        self.tri_patch = (torch.cat(tri_patch) if tri_patch else torch.zeros(0, dtype=torch.int32)).to(dev)
        # a patch table with no material (pass-through) for the geometry map alone
        self._default_patches = torch.zeros(max(1, self.num_patches), ctypes.sizeof(L.VrPatch) // 4,
                                            dtype=torch.int32)
        self._default_patches[:, 6] = -1  # inside_medium
        self._default_patches[:, 7] = -1
        self._default_patches = self._default_patches.to(dev)

    def fill(self, sc: "L.VrScene", patches: torch.Tensor) -> None:
        sc.nodes = self.nodes.data_ptr()
        sc.tri_verts = self.tri_verts.data_ptr()
        sc.tri_ids = self.tri_ids.data_ptr()
        sc.num_nodes = self.num_nodes
        sc.num_tris = self.num_tris
        sc.positions = self.positions.data_ptr()
        sc.normals = None if self.normals is None else self.normals.data_ptr()
        sc.uvs = None if self.uvs is None else self.uvs.data_ptr()
        sc.indices = self.indices.data_ptr()
        sc.tri_patch = self.tri_patch.data_ptr()
        sc.patches = patches.data_ptr()
        sc.num_patches = self.num_patches
        sc.num_vertices = self.num_vertices

    def scene_struct(self) -> "L.VrScene":
        sc = L.VrScene()
        self.fill(sc, self._default_patches)
        sc.num_media = 0
        sc.env_kind = L.VR_ENV_CONST
        return sc


# ------------------------------------------------------ surface materials --
class SurfaceScatteringSampler(MapBase):
    """win, surfel -> wout, W/pdf, pdf (20 -> 7, _scenes.py:848-879)."""

    _kind = L.VR_SURF_PASS
    __instance__ = None

    def __init__(self):
        super().__init__(20, 7)

    @classmethod
    def get_instance(cls):
        if cls.__dict__.get("__instance__") is None:
            cls.__instance__ = cls()
        return cls.__instance__

    @staticmethod
    def signature():
        return (20, 7)

    def _patch(self, p: "L.VrPatch") -> None:
        p.sampler = self._kind


class LambertSurfaceScatteringSampler(SurfaceScatteringSampler):
    """Cosine-weighted hemisphere on the incoming side (_scenes.py:906-937)."""
    _kind = L.VR_SURF_LAMBERT


class MirrorSurfaceScatteringSampler(SurfaceScatteringSampler):
    """Perfect reflection (_scenes.py:1003-1022)."""
    _kind = L.VR_SURF_MIRROR


class FresnelSurfaceScatteringSampler(SurfaceScatteringSampler):
    """Dielectric interface: Schlick reflection / refraction (_scenes.py:964-1000)."""
    _kind = L.VR_SURF_FRESNEL

    def __init__(self, refraction_index: float = 1.5):
        super().__init__()
        self.refraction_index = float(refraction_index)

    def _patch(self, p):
        p.sampler = self._kind
        p.refraction_index = self.refraction_index


class NullSurfaceScatteringSampler(SurfaceScatteringSampler):
    """Passes the ray through unchanged (_scenes.py:1308-1322)."""
    _kind = L.VR_SURF_NULL


class NoSurfaceScatteringSampler(SurfaceScatteringSampler):
    """Absorbs: zero direction, weight and pdf (_scenes.py:1325-1340)."""
    _kind = L.VR_SURF_BLOCK


class SurfaceEmission(MapBase):
    """win, surfel -> E (20 -> 3, _scenes.py:1343-1389); the emission map must be a constant."""

    _kind = L.VR_EMIT_NONE

    def __init__(self, emission_map: typing.Optional[MapBase] = None):
        super().__init__(20, 3)
        if emission_map is not None:
            if not isinstance(emission_map, ConstantMap):
                raise NotImplementedError("surface emission maps must be constants here")
            emission_map = emission_map.cast(output_dim=3)
        self.emission_map = emission_map

    @staticmethod
    def signature():
        return (20, 3)

    def _patch(self, p):
        p.emission = self._kind
        p.emission_c = L.f3(self.emission_map.vec(3) if self.emission_map is not None else [0.0, 0.0, 0.0])


class IsotropicSurfaceEmission(SurfaceEmission):
    """E * 0.5/pi (_scenes.py:1392-1412)."""
    _kind = L.VR_EMIT_ISOTROPIC


class LambertSurfaceEmission(SurfaceEmission):
    """E * dot(-win, fN)/pi (_scenes.py:1415-1435)."""
    _kind = L.VR_EMIT_LAMBERT


class NoSurfaceEmission(SurfaceEmission):
    """Zero emission (_scenes.py:1438-1454)."""
    _kind = L.VR_EMIT_NONE
    __instance__ = None

    def __init__(self):
        super().__init__(None)

    @classmethod
    def get_instance(cls):
        if cls.__instance__ is None:
            cls.__instance__ = NoSurfaceEmission()
        return cls.__instance__


class SurfaceMaterial:
    """A surface's sampler and emission (_scenes.py:1555-1583)."""

    def __init__(self, sampler: typing.Optional[SurfaceScatteringSampler] = None,
                 emission: typing.Optional[SurfaceEmission] = None):
        self.sampler = sampler
        self.emission = emission

    def emission_gathering(self):
        return self.emission


# ------------------------------------------------------------ environment --
class EnvironmentSampler(MapBase):
    """x -> (w, E(w)/pdf(w), pdf(w)) over an environment (_scenes.py:664-697)."""

    def __init__(self, environment: MapBase):
        super().__init__(3, 7)
        self.environment = environment

    def get_environment(self) -> MapBase:
        return self.environment


class UniformEnvironmentSampler(EnvironmentSampler):
    """Uniform directions over the sphere (_scenes.py:758-783)."""


# ------------------------------------------------------------------ media --
class MediumMaterial:
    """A medium's integrator factory (_scenes.py:2007-2067)."""

    def __init__(self, path_integrator_factory, scattering_albedo=None, phase_sampler=None, emission=None,
                 enclosed: bool = False, custom_boundary=None, volume_track: bool = True):
        self.path_integrator_factory = path_integrator_factory
        self.scattering_albedo = scattering_albedo
        self.emission = emission
        self.phase_sampler = phase_sampler
        self.enclosed = enclosed
        self.custom_boundary = custom_boundary
        self.volume_track = volume_track

    def transmittance(self):
        return self.path_integrator_factory(None, None, 0)

    def default_gathering(self, boundaries=None):
        """Segment integrator with scattering and emission gathering (filter 3).
        The emission gathering (1 - albedo) * emission (VolumeEmissionGathering,
        _scenes.py:1814-1824) is formed here as a constant."""
        gathering = None
        if self.emission is not None:
            if not isinstance(self.emission, ConstantMap):
                raise NotImplementedError("medium emission must be a constant")
            a = self.scattering_albedo
            if a is None:
                av = [0.0, 0.0, 0.0]
            elif isinstance(a, ConstantMap):
                av = a.vec(3)
            else:
                raise NotImplementedError("an emitting medium needs a constant albedo here")
            e = self.emission.vec(3)
            gathering = ConstantMap(torch.tensor([(1 - av[i]) * e[i] for i in range(3)], dtype=torch.float32))
        return self.path_integrator_factory(gathering, None, 3)


class HeterogeneousMediumMaterial(MediumMaterial):
    """Grid medium tracked with DTMediumPathIntegrator (_scenes.py:2087-2165).

    The reference's volume-tracking variants (DTVolumePathIntegrator,
    DSVolumePathIntegrator) do not compile there (their backward takes 8
    arguments where the interface passes 6); every technique here integrates the
    segment between surface hits with delta tracking, as volume_track=False does.
    """

    def __init__(self, sigma: MapBase, scattering_albedo: typing.Optional[MapBase] = None,
                 emission: typing.Optional[MapBase] = None, phase_sampler=None, technique: str = "dt",
                 enclosed: bool = False, custom_boundary=None, volume_track: bool = True, **kwargs):
        if technique not in ("dt", "ds"):
            raise NotImplementedError("medium technique must be 'dt' (or 'ds', tracked the same way)")
        assert "majorant" in kwargs, "Delta tracking technique requires a majorant: torch.Tensor parameter"
        majorant = kwargs["majorant"]

        def factory(gathering, boundaries, medium_filter):
            return DTMediumPathIntegrator(sigma=sigma, majorant=majorant, scattering_albedo=scattering_albedo,
                                          phase_sampler=phase_sampler, gathering=gathering, medium_filter=medium_filter)

        super().__init__(factory, scattering_albedo=scattering_albedo, phase_sampler=phase_sampler, emission=emission,
                         enclosed=enclosed, custom_boundary=custom_boundary, volume_track=volume_track)


# ---------------------------------------------------------- path tracing --
class _LoweredScene:
    """A vr_scene under construction plus the tensors it points to."""

    def __init__(self, scene_map: "PathtracedScene", device: torch.device):
        self.buffers = _GeometryBuffers.of(scene_map.surfaces, device)
        self.keep = []
        # patches
        P = len(scene_map.surfaces)
        table = (L.VrPatch * max(1, P))()
        for i in range(P):
            p = table[i]
            p.sampler = L.VR_SURF_PASS
            p.emission = L.VR_EMIT_NONE
            p.refraction_index = 1.0
            p.inside_medium = -1
            p.outside_medium = -1
            ss = None if scene_map.surface_scatterers is None else scene_map.surface_scatterers[i]
            if ss is not None:
                ss._patch(p)
            sg = None if scene_map.surface_gatherers is None else scene_map.surface_gatherers[i]
            if sg is not None:
                if not isinstance(sg, SurfaceEmission):
                    raise NotImplementedError("surface gathering must be a surface emission")
                sg._patch(p)
            if scene_map.inside_medium_indices is not None:
                p.inside_medium = int(scene_map.inside_medium_indices[i])
            if scene_map.outside_medium_indices is not None:
                p.outside_medium = int(scene_map.outside_medium_indices[i])
        raw = np.frombuffer(bytes(table), dtype=np.int32).copy()
        self.patches = torch.from_numpy(raw).to(device)
        # media
        self.media_lw = []
        for m in (scene_map.media_modules or []):
            if m is None:
                raise NotImplementedError("every scene medium needs an integrator")
            kind, lw = m._lower_map()
            if kind != L.VR_MAP_MEDIUM_DT:
                raise NotImplementedError("scene media must be DTMediumPathIntegrator maps")
            self.media_lw.append(lw)
        if len(self.media_lw) > L.VR_SCENE_MAX_MEDIA:
            raise NotImplementedError(f"at most {L.VR_SCENE_MAX_MEDIA} media per scene")
        for i in range(P):
            mi = table[i].inside_medium
            if mi < -1 or mi >= len(self.media_lw):
                raise ValueError(f"patch {i}: inside medium index {mi} out of range")
        # environment
        env_lw = _Lowered()
        env_lw.set_environment(scene_map.environment)
        self.keep.extend(env_lw.keep)
        self.env = env_lw.medium
        self.scene = L.VrScene()
        self._refresh()

    def _refresh(self):
        sc = self.scene
        self.buffers.fill(sc, self.patches)
        self.media_array = (L.VrMedium * max(1, len(self.media_lw)))(*[lw.medium for lw in self.media_lw])
        sc.media = ctypes.cast(self.media_array, ctypes.c_void_p).value if self.media_lw else None
        sc.num_media = len(self.media_lw)
        sc.env_kind = self.env.env_kind
        sc.env_c = self.env.env_c
        sc.env_image = self.env.env_image
        for a in range(2):
            sc.env_image_bmin[a] = self.env.env_image_bmin[a]
            sc.env_image_bmax[a] = self.env.env_image_bmax[a]

    def grad_targets(self, params: typing.Sequence[torch.Tensor]):
        """Fresh zero gradient buffers for the media grids that require grad."""
        grads = {}
        for lw in self.media_lw:
            for name in ("sigma", "albedo"):
                g3 = getattr(lw, name)
                if g3 is None or not g3.grid.requires_grad:
                    continue
                key = id(g3.grid)
                if key not in grads:
                    grads[key] = torch.zeros_like(g3.grid)
                setattr(lw.medium, name, g3.view(grads[key]))
        self._refresh()
        return [grads.get(id(p)) for p in params]


class PathtracedScene(MapBase):
    """x, w -> radiance through surfaces and media (6 -> 3, _scenes.py:2251-2312)."""

    def __init__(self, surfaces: Geometry, environment: MapBase, direct_gathering: MapBase = ZERO,
                 surface_gathering: typing.Optional[typing.List[typing.Optional[SurfaceEmission]]] = None,
                 surface_scattering_sampler: typing.Optional[typing.List[typing.Optional[SurfaceScatteringSampler]]] = None,
                 medium_integrator: typing.Optional[typing.List[MapBase]] = None,
                 inside_medium_indices: typing.Optional[typing.List[int]] = None,
                 outside_medium_indices: typing.Optional[typing.List[int]] = None):
        super().__init__(6, 3)
        n = len(surfaces)
        if not (isinstance(direct_gathering, ConstantMap) and all(v == 0.0 for v in direct_gathering.host_values())):
            raise NotImplementedError("direct gathering must be ZERO (Scene.pathtrace)")
        for lst in (surface_gathering, surface_scattering_sampler, inside_medium_indices, outside_medium_indices):
            assert lst is None or len(lst) == n, "per-patch lists need one entry per patch"
        self.surfaces = surfaces
        self.environment = environment
        self.surface_gatherers = surface_gathering
        self.surface_scatterers = surface_scattering_sampler
        self.media_modules = torch.nn.ModuleList([m for m in (medium_integrator or [])])
        self.inside_medium_indices = inside_medium_indices
        self.outside_medium_indices = outside_medium_indices

    def _lower_map(self):
        dev = None
        for m in self.media_modules:
            if isinstance(m, DTMediumPathIntegrator) and isinstance(m.sigma, Grid3D):
                dev = m.sigma.grid.device
                break
        if dev is None:
            dev = _internal.device()
        return L.MAP_SCENE, _LoweredScene(self, dev)


class Scene:
    """Surfaces, environment, materials and media (_scenes.py:2315-2423)."""

    def __init__(self, surfaces: Geometry, environment_sampler: EnvironmentSampler,
                 materials: typing.Optional[typing.List[typing.Optional[SurfaceMaterial]]] = None,
                 inside_media: typing.Optional[typing.List[typing.Optional[MediumMaterial]]] = None,
                 outside_media: typing.Optional[typing.List[typing.Optional[MediumMaterial]]] = None):
        self.surfaces = surfaces
        self.environment_sampler = environment_sampler
        self.materials = materials
        self.media = []
        for lst in (inside_media, outside_media):
            for m in (lst or []):
                if m is not None and m not in self.media:
                    self.media.append(m)
        self.inside_media = None if inside_media is None else [-1 if m is None else self.media.index(m) for m in inside_media]
        self.outside_media = None if outside_media is None else [-1 if m is None else self.media.index(m)
                                                                 for m in outside_media]

    @staticmethod
    def from_graph(geometries: typing.Dict[str, Geometry], environment_sampler: EnvironmentSampler,
                   materials: typing.Optional[typing.Dict[str, SurfaceMaterial]] = None,
                   inside_media: typing.Optional[typing.Dict[str, MediumMaterial]] = None,
                   outside_media: typing.Optional[typing.Dict[str, MediumMaterial]] = None) -> "Scene":
        """Per-name geometries, materials and media flattened to patches (_scenes.py:2351-2398)."""
        surfaces = GroupGeometry(*geometries.values())
        ranges, offset = {}, 0
        for k, v in geometries.items():
            ranges[k] = (offset, offset + len(v))
            offset += len(v)

        def spread(d):
            if d is None:
                return None
            out = [None] * offset
            for k, m in d.items():
                if k == "__all__":
                    out = [m] * offset
                elif k in geometries:
                    for i in range(*ranges[k]):
                        out[i] = m
            return out
        mats = None if materials is None else [None] * offset
        if materials is not None:
            for k, m in materials.items():
                if k in geometries:
                    for i in range(*ranges[k]):
                        mats[i] = m
        return Scene(surfaces, environment_sampler, mats, spread(inside_media), spread(outside_media))

    def pathtrace(self) -> PathtracedScene:
        return PathtracedScene(
            surfaces=self.surfaces,
            environment=self.environment_sampler.get_environment(),
            direct_gathering=ZERO,
            surface_gathering=None if self.materials is None else [
                None if m is None else m.emission_gathering() for m in self.materials],
            surface_scattering_sampler=None if self.materials is None else [
                None if m is None else m.sampler for m in self.materials],
            medium_integrator=[m.default_gathering() for m in self.media],
            inside_medium_indices=self.inside_media,
            outside_medium_indices=self.outside_media)
