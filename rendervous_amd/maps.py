"""Map API of the participating-media hot path.

Same class names, constructor signatures and call semantics as the reference's
map library (rendervous _maps.py) for the maps on the volume path.  A map graph
is not compiled to a shader: it is *lowered* to a ``vr_medium`` descriptor plus
a map kind and evaluated by the hand-written HIP kernels behind the C ABI
(include/vr_api.h).  Graphs outside the supported set raise NotImplementedError;
there is no fallback evaluator.
"""
from __future__ import annotations

import math
import threading
import typing
import weakref

import numpy as np
import torch

from . import _internal
from . import _lib as L

__all__ = [
    "vec3", "MapBase", "ConstantMap", "const", "ZERO", "ONE", "Grid3D", "RayBoxIntersection", "ray_box_intersection",
    "MajorantGrid", "CompositionMap", "Image2D", "XRProjection", "xr_projection", "XRQuadtreeRandomDirection",
    "GridDDATransmittance", "GridDeltatrackingTransmittance", "GridRatiotrackingTransmittance",
    "DeltatrackingTransmittance", "RatiotrackingTransmittance",
    "DeltatrackingCollisionSampler", "HGPhase", "VHGPhaseSampler", "UniformDirectionSampler", "_functionsampler",
    "DirectionalLightEnvironmentSampler", "SingleScatteredRadiance", "MCCollisionIntegrator",
    "DeltatrackingNEEPathIntegrator", "DeltatrackingPathIntegrator", "DRTPathIntegrator", "SensorsBase", "PerspectiveCameraSensor", "normalized_box",
    "AutogradMapFunction", "AutogradCaptureFunction",
]

_ENGINE_LOCK = threading.Lock()  # serialises capture fw/bw and map bw (_maps.py:57,752,827,929)


def vec3(x: float, y: float, z: float) -> torch.Tensor:
    """A float32 3-vector (the reference's vulky vec3)."""
    return torch.tensor([x, y, z], dtype=torch.float32)


def _host3(v) -> typing.List[float]:
    if isinstance(v, torch.Tensor):
        return [float(a) for a in v.detach().cpu().tolist()]
    return [float(a) for a in v]


# ----------------------------------------------------------------- base --
class MapBase(torch.nn.Module):
    """A map R^INPUT_DIM -> R^OUTPUT_DIM (include/common.h:23-25)."""

    input_dim: typing.Optional[int] = None
    output_dim: typing.Optional[int] = None

    def __init__(self, input_dim: typing.Optional[int] = None, output_dim: typing.Optional[int] = None):
        super().__init__()
        object.__setattr__(self, "input_dim", input_dim)
        object.__setattr__(self, "output_dim", output_dim)

    @property
    def is_generic_input(self) -> bool:
        return self.input_dim is None

    def after(self, prev_map: "MapBase") -> "MapBase":
        """self(prev_map(x)) (_maps.py:1347-1348)."""
        return CompositionMap(prev_map, self)

    def then(self, next_map: "MapBase") -> "MapBase":
        """next_map(self(x)) (_maps.py:1350-1355)."""
        return CompositionMap(self, next_map)

    def cast(self, input_dim: typing.Optional[int] = None, output_dim: typing.Optional[int] = None) -> "MapBase":
        """Type a generic map (_maps.py:1260-1271)."""
        assert input_dim is None or self.input_dim is None or self.input_dim == input_dim, \
            f"Required input {input_dim} was not satisfied {self.input_dim}"
        assert output_dim is None or self.output_dim == output_dim or self.output_dim == 1, \
            f"Required output {output_dim} was not satisfied {self.output_dim}"
        return self

    def get_maximum(self) -> float:
        raise NotImplementedError()

    def support_maximum_query(self) -> bool:
        return False

    # Lowering to the C ABI: returns (kind, medium-builder) for ray maps.
    def _lower_map(self) -> typing.Tuple[int, "_Lowered"]:
        raise NotImplementedError(f"{type(self).__name__} cannot be evaluated by the HIP path")

    def forward(self, *args):
        x, = args
        assert self.output_dim is not None, "Can not eval a map with generic output"
        assert self.input_dim is None or self.input_dim == x.shape[-1], "Map input dimension and tensor last dim missmatch"
        params = [p for p in self.parameters() if p.requires_grad]
        return AutogradMapFunction.apply(x, self, *params)


class ConstantMap(MapBase):
    """const[v...] (_maps.py:1957-1985, include/maps/constant.h)."""

    def __init__(self, value: torch.Tensor, input_dim: typing.Optional[int] = None):
        assert len(value.shape) <= 1, f"Error {value.shape}"
        assert value.dtype == torch.float
        super().__init__(input_dim, value.numel())
        self.value = value if isinstance(value, torch.nn.Parameter) else torch.nn.Parameter(value, requires_grad=False)
        self._host = None
        self._host_version = -1

    def host_values(self) -> typing.List[float]:
        if self._host is None or self._host_version != self.value._version:
            self._host = [float(a) for a in self.value.detach().reshape(-1).cpu().tolist()]
            self._host_version = self.value._version
        return self._host

    def cast(self, input_dim=None, output_dim=None):
        assert output_dim is None or output_dim == self.output_dim or self.output_dim == 1, \
            f"Constant map cast to an invalid output dim {output_dim} from {self.output_dim}"
        c = self
        if input_dim is not None and self.input_dim is None:
            c = ConstantMap(self.value, input_dim)
        if output_dim is not None and self.output_dim == 1 and output_dim > 1:
            c = _PromotedConstant(c, output_dim)
        return c

    def vec(self, n: int) -> typing.List[float]:
        v = self.host_values()
        if len(v) == 1:
            return v * n
        assert len(v) == n, f"constant has {len(v)} values, {n} required"
        return v

    def get_maximum(self) -> float:
        return max(self.host_values())


class _PromotedConstant(ConstantMap):
    """A scalar constant promoted to `dim` outputs (ComposePromoteMap of a ConstantMap)."""

    def __init__(self, base: ConstantMap, dim: int):
        super().__init__(base.value, base.input_dim)
        object.__setattr__(self, "output_dim", dim)

    def host_values(self):
        v = super().host_values()
        return v * self.output_dim if len(v) == 1 else v


class _ConstFactory:
    """const[...] parametric constructor (_maps.py:4791-4806)."""

    def __getitem__(self, args):
        if not isinstance(args, tuple):
            args = (args,)
        input_dim = None
        if len(args) > 0 and isinstance(args[-1], dict):
            input_dim = args[-1].get("input_dim")
            args = args[:-1]
        if len(args) == 1 and isinstance(args[0], torch.Tensor):
            return ConstantMap(args[0], input_dim=input_dim)
        return ConstantMap(torch.tensor([float(a) for a in args], dtype=torch.float32), input_dim=input_dim)


const = _ConstFactory()
ZERO = const[0.0]
ONE = const[1.0]


class CompositionMap(MapBase):
    """outer(inner(x)) (_maps.py:1804-1830, include/maps/composition.h); lowered by
    pattern (Image2D after XRProjection as an environment), not evaluated alone."""

    def __init__(self, inner: MapBase, outer: MapBase):
        super().__init__(inner.input_dim, outer.output_dim)
        self.inner = inner
        self.outer = outer


class Image2D(MapBase):
    """Pixel-centred bilinear image [H,W,C] over [bmin, bmax] with border clamping
    (_maps.py:2288-2316, include/maps/image2d.h)."""

    def __init__(self, grid: torch.Tensor, bmin=(-1.0, -1.0), bmax=(1.0, 1.0)):
        assert len(grid.shape) == 3
        super().__init__(2, grid.shape[-1])
        self.grid = grid if isinstance(grid, torch.nn.Parameter) else torch.nn.Parameter(grid, requires_grad=False)
        self.bmin = [float(v) for v in bmin]
        self.bmax = [float(v) for v in bmax]


class XRProjection(MapBase):
    """Direction (or ray) -> equirectangular coordinates in [-1, 1]^2 (dir2xr,
    include/maths.h:172-178, include/maps/xr_projection.h)."""

    def __init__(self, ray_input: bool = False):
        super().__init__(6 if ray_input else 3, 2)
        self.ray_input = ray_input


xr_projection = XRProjection(False)


class XRQuadtreeRandomDirection(MapBase):
    """Importance-sampled direction from a quadtree of equirectangular pixel
    probabilities (functions.create_density_quadtree), weight 1/pdf (_maps.py:2494-2549)."""

    def __init__(self, input_dim: int, densities: torch.Tensor, levels: int):
        super().__init__(input_dim, 4)
        self.densities = densities
        self.levels = int(levels)


# ---------------------------------------------------------------- Grid3D --
class _TensorStamp:
    """Identity, storage address, version and layout of a tensor, for caches of
    tensors derived from it (cell tables, majorant grids): a cache entry is
    reused only while all of them match, as the reference's TensorCheck compares
    identity and version (_maps.py:1061-1073).  In-place writes through `.data`
    do not bump a tensor's version; follow them with
    torch.autograd.graph.increment_version(t)."""

    def __init__(self, t: torch.Tensor):
        self.ref = weakref.ref(t)
        self.key = (t.data_ptr(), t._version, tuple(t.shape), tuple(t.stride()))

    def matches(self, t: torch.Tensor) -> bool:
        return self.ref() is t and self.key == (t.data_ptr(), t._version, tuple(t.shape), tuple(t.stride()))


class Grid3D(MapBase):
    """Vertex-centred trilinear grid [D,H,W,C] over [bmin,bmax] (_maps.py:2319-2365)."""

    def __init__(self, grid: torch.Tensor, bmin: torch.Tensor = None, bmax: torch.Tensor = None):
        assert len(grid.shape) == 4
        super().__init__(3, grid.shape[-1])
        bmin = vec3(-1.0, -1.0, -1.0) if bmin is None else bmin
        bmax = vec3(1.0, 1.0, 1.0) if bmax is None else bmax
        self.grid = grid if isinstance(grid, torch.nn.Parameter) else torch.nn.Parameter(grid, requires_grad=False)
        self.bmin = _host3(bmin)
        self.bmax = _host3(bmax)
        self.width = grid.shape[2]
        self.height = grid.shape[1]
        self.depth = grid.shape[0]

    def view(self, grad: typing.Optional[torch.Tensor] = None) -> L.VrGrid:
        return L.grid_view(self.grid.data, self.bmin, self.bmax, grad)

    def cells_tensor(self) -> torch.Tensor:
        """The grid's cell table (vr_grid_pack_cells: 8 corners per cell, 32 bytes),
        rebuilt when the grid tensor changes (its pointer, version or layout)."""
        g = self.grid
        cache = self.__dict__.get("_cells_cache")
        if cache is None or not cache[0].matches(g):
            n = (g.shape[0] - 1) * (g.shape[1] - 1) * (g.shape[2] - 1)
            out = torch.empty(max(1, n) * 8, device=g.device, dtype=torch.float32)
            L.check(L.load().vr_grid_pack_cells(self.view(), L.ptr(out), L.stream_handle(g.device)), "vr_grid_pack_cells")
            cache = (_TensorStamp(g), out)
            object.__setattr__(self, "_cells_cache", cache)
        return cache[1]

    def get_maximum(self) -> float:
        return self.grid.max().item()

    def support_maximum_query(self) -> bool:
        return True

    def forward(self, *args):
        x, = args
        assert x.shape[-1] == 3
        params = [self.grid] if self.grid.requires_grad else []
        return _Grid3DFunction.apply(x, self, *params)


class _Grid3DFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, grid_map: Grid3D, *params):
        lib = L.load()
        xc = x.detach().contiguous()
        n = xc.numel() // 3
        out = torch.empty(*x.shape[:-1], grid_map.output_dim, device=x.device, dtype=torch.float32)
        g = grid_map.view()
        L.check(lib.vr_grid_fw(g, L.ptr(xc), n, L.ptr(out), L.stream_handle(x.device)), "vr_grid_fw")
        ctx.save_for_backward(xc)
        ctx.grid_map = grid_map
        return out

    @staticmethod
    def backward(ctx, og):
        xc, = ctx.saved_tensors
        grid_map: Grid3D = ctx.grid_map
        if not grid_map.grid.requires_grad:
            return (None, None)
        lib = L.load()
        grad = torch.zeros_like(grid_map.grid)
        og = og.contiguous()
        g = grid_map.view(grad)
        L.check(lib.vr_grid_bw(g, L.ptr(xc), L.ptr(og), xc.numel() // 3, L.stream_handle(xc.device)), "vr_grid_bw")
        return (None, None, grad)


# -------------------------------------------------------------- boundary --
class RayBoxIntersection(MapBase):
    """(x,w) -> (tMin, tMax) of the slab test (_maps.py:4128-4154, maths.h:28-38)."""

    def __init__(self, box_min, box_max, **kwargs):
        super().__init__(6, 2)
        self.box_min = _host3(box_min)
        self.box_max = _host3(box_max)

    def _lower_map(self):
        lw = _Lowered()
        lw.medium.box_min = L.f3(self.box_min)
        lw.medium.box_max = L.f3(self.box_max)
        return L.VR_MAP_RAYBOX, lw


class _RayBoxFactory:
    def __getitem__(self, args):
        return RayBoxIntersection(*args)


ray_box_intersection = _RayBoxFactory()


def normalized_box(s) -> typing.Tuple[torch.Tensor, torch.Tensor]:
    """Box of half-size (shape-1)/(2*max_dim) (_maps.py:4780-4788)."""
    shape = s.shape if isinstance(s, torch.Tensor) else s
    max_dim = max(shape[0], shape[1], shape[2]) - 1
    b_max = vec3(shape[2] - 1, shape[1] - 1, shape[0] - 1) * 0.5 / max_dim
    return -b_max, b_max


# -------------------------------------------------------------- majorant --
class MajorantGrid(MapBase):
    """(x,w) -> (majorant, distance) from a coarse max grid over a Grid3D.

    Implements the 6->2 majorant contract (include/vr_majorant.h:1-7) that the
    reference fills with const[max(sigma), 1e6] (_objects.py:96-101).  Each
    coarse cell stores the maximum of the vertices that can influence the
    trilinear field inside it; the trackers walk the cells with a DDA.  The
    coarse grid is rebuilt when the density tensor changes (TensorCheck,
    _maps.py:1061-1073).
    """

    def __init__(self, sigma: Grid3D, resolution: typing.Union[int, typing.Sequence[int]] = 16):
        super().__init__(6, 2)
        assert isinstance(sigma, Grid3D) and sigma.output_dim == 1
        self.sigma_grid = sigma
        r = [resolution] * 3 if isinstance(resolution, int) else list(resolution)
        self.resolution = [int(a) for a in r]  # rz, ry, rx
        self._cache = None
        self._cache_key = None

    def majorant_tensor(self) -> torch.Tensor:
        g = self.sigma_grid.grid
        if self._cache is None or not self._cache_key.matches(g):
            if self._cache_key is None or self._cache_key.ref() is not g:
                # a density first seen here: NaN or infinite values would leave the
                # coarse bounds NaN (cells crossed as empty) -- refuse them once
                # (one host synchronisation per new tensor, none per update)
                if not bool(torch.isfinite(g.detach()).all()):
                    raise ValueError("the density holds NaN or infinite values")
            out = torch.empty(*self.resolution, device=g.device, dtype=torch.float32)
            res = (L.ctypes.c_int32 * 3)(*self.resolution)
            L.check(L.load().vr_majorant_build(self.sigma_grid.view(), res, L.ptr(out), L.stream_handle(g.device)),
                    "vr_majorant_build")
            self._cache = out
            self._cache_key = _TensorStamp(g)
        return self._cache


# ---------------------------------------------------------- directions --
class UniformDirectionSampler(MapBase):
    """Uniform sphere direction with weight 4*pi (include/maps/uniform_direction_sampler.h)."""

    def __init__(self, input_dim):
        super().__init__(input_dim, 4)


class _functionsampler(MapBase):
    """f(w)*weight, w from a point sampler (_maps.py:2552-2585)."""

    def __init__(self, point_sampler: MapBase, function_map: MapBase):
        assert point_sampler.output_dim - 1 == function_map.input_dim or function_map.input_dim is None
        super().__init__(point_sampler.input_dim, point_sampler.output_dim - 1 + function_map.output_dim)
        self.point_sampler = point_sampler
        self.function_map = function_map


class DirectionalLightEnvironmentSampler(MapBase):
    """Environment sampler for a constant environment plus a directional light.

    The reference has no directional light; this sampler expresses one with the
    reference's MixtureSampler semantics (_maps.py:1541-1571): with probability
    `probability` it returns the light direction with weight radiance/probability,
    otherwise a uniform direction with weight env*4*pi/(1-probability).  The
    light's direction points from the medium towards the light.
    """

    def __init__(self, environment: MapBase, direction, radiance, probability: float = 0.5):
        super().__init__(6, 6)
        assert 0.0 < probability < 1.0
        self.environment = environment
        d = np.asarray(_host3(direction), dtype=np.float64)
        d = d / np.linalg.norm(d)
        self.direction = [float(a) for a in d]
        self.radiance = _host3(radiance)
        self.probability = float(probability)


class HGPhase(MapBase):
    """Henyey-Greenstein phase (x, w, wo) -> p (_maps.py:4436-4456).

    mixture=True evaluates the reference's 0.8*HG(g) + 0.2*HG(g/5) (:4453);
    mixture=False a single HG(g).
    """

    def __init__(self, phase_g: MapBase, mixture: bool = True):
        super().__init__(9, 1)
        self.phase_g = phase_g.cast(3, 1)
        self.mixture = mixture


class VHGPhaseSampler(MapBase):
    """HG direction sampler (x, w) -> (1, wo) (_maps.py:4463-4484).

    mixture=True samples HG(g/5) with probability 0.8 as the reference does (:4475);
    mixture=False samples HG(g).
    """

    def __init__(self, phase_g: MapBase, mixture: bool = True):
        super().__init__(6, 4)
        self.phase_g = phase_g.cast(3, 1)
        self.mixture = mixture


# ---------------------------------------------------------- lowering --
class _Lowered:
    """A vr_medium under construction plus the tensors it points to."""

    def __init__(self):
        self.medium = L.VrMedium()
        self.medium.albedo_kind = L.VR_FIELD_CONST
        self.medium.albedo_c = L.f3([0.0, 0.0, 0.0])
        self.medium.phase_kind = L.VR_PHASE_ISO
        self.medium.majorant_kind = L.VR_MAJ_CONST
        self.medium.majorant_value = 0.0
        self.medium.majorant_distance = 1e6
        self.medium.transmittance_kind = L.VR_T_DT
        self.medium.max_bounces = -1
        self.medium.medium_filter = 0
        self.keep: typing.List[torch.Tensor] = []
        self.sigma: typing.Optional[Grid3D] = None
        self.albedo: typing.Optional[Grid3D] = None
        self.majorant_set = False
        self.device: typing.Optional[torch.device] = None

    def set_sigma(self, sigma: MapBase, multiple_scattering: bool = False):
        """multiple_scattering: the integrator traces paths of many tracking steps
        per sample (NEE / DT path / DRT), where a cell table pays for its rebuild
        after every update of a density that requires a gradient."""
        if not isinstance(sigma, Grid3D) or sigma.output_dim != 1:
            raise NotImplementedError("sigma must be a single-channel Grid3D")
        if self.sigma is not None and self.sigma is not sigma:
            raise NotImplementedError("all sub-maps must share one sigma Grid3D")
        self.sigma = sigma
        self.device = sigma.grid.device
        self.medium.sigma = sigma.view()
        ct = _internal.settings.cell_tables
        g = sigma.grid
        n_cells = (g.shape[0] - 1) * (g.shape[1] - 1) * (g.shape[2] - 1)
        plane = (g.shape[1] - 1) * (g.shape[2] - 1)
        auto = not g.requires_grad or multiple_scattering
        if g.is_cuda and n_cells < (1 << 30) and plane < (1 << 24) and (ct is True or (ct == "auto" and auto)):
            table = sigma.cells_tensor()
            self.keep.append(table)
            self.medium.sigma_cells = table.data_ptr()

    def set_boundary(self, boundary: MapBase):
        if not isinstance(boundary, RayBoxIntersection):
            raise NotImplementedError("boundary must be a RayBoxIntersection")
        self.medium.box_min = L.f3(boundary.box_min)
        self.medium.box_max = L.f3(boundary.box_max)

    def set_majorant(self, majorant: MapBase):
        before = (self.medium.majorant_kind, self.medium.majorant_value, self.medium.majorant_distance,
                  self.medium.majorant_grid)
        self._set_majorant(majorant)
        after = (self.medium.majorant_kind, self.medium.majorant_value, self.medium.majorant_distance,
                 self.medium.majorant_grid)
        if self.majorant_set and before != after:
            raise NotImplementedError("all sub-maps must share one majorant")
        self.majorant_set = True

    def _set_majorant(self, majorant: MapBase):
        if isinstance(majorant, MajorantGrid):
            if majorant.sigma_grid is not self.sigma:
                raise NotImplementedError("MajorantGrid must bound the medium's sigma")
            t = majorant.majorant_tensor()
            self.keep.append(t)
            self.medium.majorant_kind = L.VR_MAJ_GRID
            self.medium.majorant_grid = t.data_ptr()
            for i in range(3):
                self.medium.majorant_res[i] = majorant.resolution[i]
        elif isinstance(majorant, ConstantMap):
            v = majorant.vec(2)
            self.medium.majorant_kind = L.VR_MAJ_CONST
            self.medium.majorant_value = v[0]
            self.medium.majorant_distance = v[1]
        else:
            raise NotImplementedError("majorant must be const[value, distance] or a MajorantGrid")

    def set_albedo(self, albedo: MapBase):
        if isinstance(albedo, Grid3D):
            if albedo.output_dim not in (1, 3):
                raise NotImplementedError("albedo grid must have 1 or 3 channels")
            self.albedo = albedo
            self.medium.albedo_kind = L.VR_FIELD_GRID
            self.medium.albedo = albedo.view()
        elif isinstance(albedo, ConstantMap):
            self.medium.albedo_kind = L.VR_FIELD_CONST
            self.medium.albedo_c = L.f3(albedo.vec(3))
        else:
            raise NotImplementedError("scattering albedo must be a constant or a Grid3D")

    def set_emission(self, emission: MapBase):
        if not isinstance(emission, ConstantMap):
            raise NotImplementedError("emission must be a constant")
        self.medium.emission_c = L.f3(emission.vec(3))

    def _environment_of(self, environment: MapBase):
        """('const', [r,g,b]) or ('xr', Image2D) for the supported environments."""
        if isinstance(environment, ConstantMap):
            return "const", environment.vec(3)
        if isinstance(environment, CompositionMap) and isinstance(environment.inner, XRProjection) \
                and isinstance(environment.outer, Image2D) and environment.outer.output_dim == 3:
            if environment.outer.grid.requires_grad:
                raise NotImplementedError("environment images are constants here (no gradient into the image)")
            return "xr", environment.outer
        raise NotImplementedError("environment must be a constant or Image2D(...).after(xr_projection)")

    def _apply_environment(self, kind, value):
        if kind == "const":
            key = ("const", tuple(float(a) for a in L.f3(value)))
        else:
            key = ("xr", id(value))
        prev = getattr(self, "_env_key", None)
        if prev is not None and prev != key:
            raise NotImplementedError("environment sampler and environment must share the environment")
        self._env_key = key
        if kind == "const":
            self.medium.env_kind = L.VR_ENV_CONST
            self.medium.env_c = L.f3(value)
            return
        img = value.grid.data
        if img.dtype != torch.float32:
            raise NotImplementedError("environment image must be float32")
        self.keep.append(img)
        self.medium.env_kind = L.VR_ENV_XR
        self.medium.env_image = L.tensor_view(img)
        for a in range(2):
            self.medium.env_image_bmin[a] = value.bmin[a]
            self.medium.env_image_bmax[a] = value.bmax[a]

    def set_environment(self, environment: MapBase):
        self._apply_environment(*self._environment_of(environment))

    def set_environment_sampler(self, sampler: MapBase):
        if isinstance(sampler, _functionsampler) and isinstance(sampler.point_sampler, UniformDirectionSampler):
            self._apply_environment(*self._environment_of(sampler.function_map))
            self.medium.env_sampler_kind = L.VR_ENVS_UNIFORM
            self.medium.sun_prob = 0.0
        elif isinstance(sampler, _functionsampler) and isinstance(sampler.point_sampler, XRQuadtreeRandomDirection):
            self._apply_environment(*self._environment_of(sampler.function_map))
            q = sampler.point_sampler
            t = q.densities
            if not (isinstance(t, torch.Tensor) and t.dtype == torch.float32 and t.is_contiguous()):
                raise NotImplementedError("quadtree densities must be a contiguous float32 tensor")
            if t.numel() < (4 ** (q.levels + 1) - 4) // 3:
                raise ValueError("quadtree densities hold fewer nodes than its levels need")
            self.keep.append(t)
            self.medium.env_sampler_kind = L.VR_ENVS_QUADTREE
            self.medium.env_quadtree = t.data_ptr()
            self.medium.env_quadtree_levels = q.levels
            self.medium.sun_prob = 0.0
        elif isinstance(sampler, DirectionalLightEnvironmentSampler):
            self._apply_environment(*self._environment_of(sampler.environment))
            self.medium.env_sampler_kind = L.VR_ENVS_UNIFORM
            self.medium.sun_dir = L.f3(sampler.direction)
            self.medium.sun_radiance = L.f3(sampler.radiance)
            self.medium.sun_prob = sampler.probability
        else:
            raise NotImplementedError("environment sampler must be uniform or quadtree over the environment, "
                                      "or a DirectionalLightEnvironmentSampler")

    def set_phase_g(self, *phases: MapBase):
        """The asymmetry shared by an HG phase and its sampler: a constant, or a
        single-channel Grid3D that the kernels read at each scattering position
        (HGPhase and VHGPhaseSampler evaluate phase_g at x, _maps.py:4449,4473).
        No gradient flows into it: the reference's HGPhase has no backward."""
        gs = [p.phase_g for p in phases]
        if isinstance(gs[0], Grid3D):
            grid = gs[0]
            if any(g is not grid for g in gs):
                raise NotImplementedError("phase and phase sampler must share one phase_g grid")
            if grid.output_dim != 1:
                raise NotImplementedError("phase_g grid must have one channel")
            self.medium.phase_g_kind = L.VR_FIELD_GRID
            self.medium.phase_g_grid = grid.view()
            self.medium.phase_g = 0.0
            self.keep.append(grid.grid)
            return
        if not all(isinstance(g, ConstantMap) for g in gs):
            raise NotImplementedError("phase g must be a constant or a Grid3D")
        values = [g.vec(1)[0] for g in gs]
        if any(v != values[0] for v in values):
            raise NotImplementedError("phase and phase sampler must use the same HG parameters")
        self.medium.phase_g_kind = L.VR_FIELD_CONST
        self.medium.phase_g = values[0]

    def set_phase(self, phase: MapBase, sampler: MapBase):
        if isinstance(phase, HGPhase) and isinstance(sampler, VHGPhaseSampler):
            if phase.mixture != sampler.mixture:
                raise NotImplementedError("phase and phase sampler must use the same HG parameters")
            self.set_phase_g(phase, sampler)
            self.medium.phase_kind = L.VR_PHASE_HG_MIX if phase.mixture else L.VR_PHASE_HG
        elif isinstance(phase, ConstantMap) and abs(phase.vec(1)[0] - 1.0 / (4 * math.pi)) < 1e-7 \
                and _is_uniform_phase_sampler(sampler):
            self.medium.phase_kind = L.VR_PHASE_ISO
        else:
            raise NotImplementedError("phase must be HGPhase + VHGPhaseSampler or isotropic")

    def set_phase_sampler(self, sampler: MapBase):
        """A phase sampler alone (integrators that never evaluate the phase)."""
        if isinstance(sampler, VHGPhaseSampler):
            self.set_phase_g(sampler)
            self.medium.phase_kind = L.VR_PHASE_HG_MIX if sampler.mixture else L.VR_PHASE_HG
        elif _is_uniform_phase_sampler(sampler):
            self.medium.phase_kind = L.VR_PHASE_ISO
        else:
            raise NotImplementedError("phase sampler must be a VHGPhaseSampler or uniform")

    def set_transmittance(self, t: MapBase):
        if isinstance(t, GridDDATransmittance):
            self.set_sigma(t.grid_model)
            self.medium.transmittance_kind = L.VR_T_DDA
            return
        if isinstance(t, _GridTrackingTransmittance):
            self.set_sigma(t.grid_model)
            self.set_boundary(t.boundary)
            self.medium.transmittance_kind = t._t_kind
            return
        if isinstance(t, (DeltatrackingTransmittance, RatiotrackingTransmittance)):
            self.set_sigma(t.sigma)
            self.set_boundary(t.boundary)
            self.set_majorant(t.majorant)
            self.medium.transmittance_kind = L.VR_T_DT if isinstance(t, DeltatrackingTransmittance) else L.VR_T_RT
            return
        raise NotImplementedError("transmittance must be a DT, RT, DDA, GDT or GRT transmittance map")


def _is_uniform_phase_sampler(s: MapBase) -> bool:
    return isinstance(s, UniformDirectionSampler) or getattr(s, "_uniform_phase", False)


class _RayMap(MapBase):
    """Base of maps evaluated by vr_map_fw / vr_map_bw."""

    def _lower_map(self):
        raise NotImplementedError

    def forward(self, *args):
        x, = args
        assert x.shape[-1] == self.input_dim, "Map input dimension and tensor last dim missmatch"
        params = [p for p in self.parameters() if p.requires_grad]
        return AutogradMapFunction.apply(x, self, *params)


class GridDDATransmittance(_RayMap):
    """Deterministic DDA transmittance over a Grid3D (_maps.py:2640-2667)."""

    def __init__(self, grid: Grid3D, boundary: MapBase):
        super().__init__(6, 1)
        self.boundary = boundary
        self.grid_model = grid

    def update_grid(self, grid: typing.Optional[Grid3D] = None):
        if grid is not None:
            self.grid_model = grid

    def _lower_map(self):
        lw = _Lowered()
        lw.set_sigma(self.grid_model)
        lw.set_boundary(self.boundary)
        lw.medium.transmittance_kind = L.VR_T_DDA
        return L.VR_MAP_T_DDA, lw


class _GridTrackingTransmittance(_RayMap):
    """Grid-decomposition tracking over a Grid3D: per DDA cell, the minorant of
    the cell's 8 vertices is integrated in closed form and the residual tracked
    with the cell's max - min as majorant.  No backward (the reference has none):
    the gradient with respect to the grid is 0."""
    _t_kind = None
    _map_kind = None

    def __init__(self, grid: Grid3D, boundary: MapBase):
        super().__init__(6, 1)
        self.grid_model = grid
        self.boundary = boundary

    def _lower_map(self):
        lw = _Lowered()
        lw.set_transmittance(self)
        return self._map_kind, lw


class GridDeltatrackingTransmittance(_GridTrackingTransmittance):
    """(_maps.py:2619-2637, vr_transmittance_GDT.h): 0 at the first residual collision."""
    _t_kind = L.VR_T_GDT
    _map_kind = L.VR_MAP_T_GDT


class GridRatiotrackingTransmittance(_GridTrackingTransmittance):
    """(_maps.py:2588-2616, vr_transmittance_GRT.h): the product of 1 - Pc over residual events."""
    _t_kind = L.VR_T_GRT
    _map_kind = L.VR_MAP_T_GRT


class DeltatrackingTransmittance(_RayMap):
    """Delta-tracking transmittance (_maps.py:2685-2697, vr_transmittance_DT.h)."""

    def __init__(self, sigma: MapBase, boundary: MapBase, majorant: MapBase):
        super().__init__(6, 1)
        self.sigma = sigma
        self.boundary = boundary
        self.majorant = majorant

    def _lower_map(self):
        lw = _Lowered()
        lw.set_transmittance(self)
        return L.VR_MAP_T_DT, lw


class RatiotrackingTransmittance(_RayMap):
    """Ratio-tracking transmittance (_maps.py:2670-2682, vr_transmittance_RT.h)."""

    def __init__(self, sigma: MapBase, boundary: MapBase, majorant: MapBase):
        super().__init__(6, 1)
        self.sigma = sigma.cast(3, 1)
        self.boundary = boundary.cast(6, 2)
        self.majorant = majorant.cast(6, 2)

    def _lower_map(self):
        lw = _Lowered()
        lw.set_transmittance(self)
        return L.VR_MAP_T_RT, lw


class DeltatrackingCollisionSampler(_RayMap):
    """(x,w) -> (C, t, T) collision sampling (_maps.py:2714-2727, vr_collision_sampler_DT.h)."""

    def __init__(self, sigma: MapBase, boundary: MapBase, majorant: MapBase, ds_epsilon: float):
        super().__init__(6, 3)
        self.sigma = sigma
        self.boundary = boundary
        self.majorant = majorant
        self.ds_epsilon = ds_epsilon

    def _lower_into(self, lw: _Lowered):
        lw.set_sigma(self.sigma)
        lw.set_boundary(self.boundary)
        lw.set_majorant(self.majorant)
        lw.medium.ds_epsilon = self.ds_epsilon

    def _lower_map(self):
        lw = _Lowered()
        self._lower_into(lw)
        return L.VR_MAP_COLLISION_DT, lw


class SingleScatteredRadiance(_RayMap):
    """env * albedo * phase * T(x, wo) (_maps.py:2761-2784, single_scattered_radiance_mc.h)."""

    def __init__(self, scattering_albedo: MapBase, phase: MapBase, environment_sampler: MapBase, transmittance: MapBase):
        super().__init__(6, 3)
        self.scattering_albedo = scattering_albedo.cast(3, 3)
        self.phase = phase.cast(9, 1)
        self.environment_sampler = environment_sampler.cast(6, 6)
        self.transmittance = transmittance.cast(6, 1)

    def _lower_into(self, lw: _Lowered):
        lw.set_transmittance(self.transmittance)
        lw.set_albedo(self.scattering_albedo)
        lw.set_environment_sampler(self.environment_sampler)
        if isinstance(self.phase, HGPhase):
            lw.set_phase_g(self.phase)
            lw.medium.phase_kind = L.VR_PHASE_HG_MIX if self.phase.mixture else L.VR_PHASE_HG
        elif isinstance(self.phase, ConstantMap):
            lw.medium.phase_kind = L.VR_PHASE_ISO
        else:
            raise NotImplementedError("phase must be HGPhase or isotropic")

    def _lower_map(self):
        lw = _Lowered()
        self._lower_into(lw)
        return L.VR_MAP_SS_RADIANCE, lw


class MCCollisionIntegrator(_RayMap):
    """C * exitance(x_t, -w) + T * env(w) (_maps.py:2788-2808, collision_integrator_mc.h).

    Lowered when the exitance is a SingleScatteredRadiance and the environment a constant.
    """

    def __init__(self, collision_sampler: MapBase, exitance_radiance: MapBase, environment: MapBase):
        super().__init__(6, 3)
        self.collision_sampler = collision_sampler
        self.exitance_radiance = exitance_radiance
        self.environment = environment

    def _lower_map(self):
        if not isinstance(self.collision_sampler, DeltatrackingCollisionSampler):
            raise NotImplementedError("collision sampler must be DeltatrackingCollisionSampler")
        if not isinstance(self.exitance_radiance, SingleScatteredRadiance):
            raise NotImplementedError("exitance radiance must be SingleScatteredRadiance")
        lw = _Lowered()
        self.collision_sampler._lower_into(lw)
        self.exitance_radiance._lower_into(lw)
        lw.set_environment(self.environment)
        return L.VR_MAP_MC_COLLISION, lw


class DeltatrackingNEEPathIntegrator(_RayMap):
    """NEE volume path tracer with delta tracking (_maps.py:2924-2978,
    vr_path_integrator_NEE.h; backward vr_path_integrator_NEE_DS.h).

    max_bounces < 0 keeps the reference's unbounded paths; otherwise the path
    ends after that many scattering events (their NEE included).
    """

    def __init__(self, sigma: MapBase, scattering_albedo: MapBase, emission: MapBase, environment: MapBase,
                 environment_sampler: MapBase, phase: MapBase, phase_sampler: MapBase, boundary: MapBase,
                 majorant: MapBase, transmittance: MapBase, ds_epsilon: float, max_bounces: int = -1):
        super().__init__(6, 3)
        self.sigma = sigma.cast(3, 1)
        self.scattering_albedo = scattering_albedo.cast(3, 3)
        self.emission = emission.cast(6, 3)
        self.environment = environment.cast(3, 3)
        self.environment_sampler = environment_sampler.cast(6, 6)
        self.phase = phase.cast(9, 1)
        self.phase_sampler = phase_sampler.cast(6, 4)
        self.boundary = boundary.cast(6, 2)
        self.majorant = majorant.cast(6, 2)
        self.transmittance = transmittance.cast(6, 1)
        self.ds_epsilon = ds_epsilon
        self.max_bounces = max_bounces

    def _lower_map(self):
        lw = _Lowered()
        lw.set_sigma(self.sigma, multiple_scattering=True)
        lw.set_boundary(self.boundary)
        lw.set_majorant(self.majorant)
        lw.set_albedo(self.scattering_albedo)
        lw.set_emission(self.emission)
        lw.set_environment(self.environment)
        lw.set_environment_sampler(self.environment_sampler)
        lw.set_phase(self.phase, self.phase_sampler)
        lw.set_transmittance(self.transmittance)
        lw.medium.ds_epsilon = self.ds_epsilon
        lw.medium.max_bounces = int(self.max_bounces)
        return L.VR_MAP_NEE_DT, lw


class DeltatrackingPathIntegrator(_RayMap):
    """Delta-tracking path tracer without NEE: radiance = W * environment at the
    path's escape plus emission (_maps.py:2886-2921, maps/radiance_DT.h,
    vr_path_integrator_DT.h:7-56); backward by defensive sampling and replay
    (vr_path_integrator_DT.h:59-235).  max_bounces as in
    DeltatrackingNEEPathIntegrator (a capped path has no escape term).
    """

    def __init__(self, sigma: MapBase, scattering_albedo: MapBase, emission: MapBase, environment: MapBase,
                 phase_sampler: MapBase, boundary: MapBase, majorant: MapBase, ds_epsilon: float,
                 max_bounces: int = -1):
        super().__init__(6, 3)
        self.sigma = sigma.cast(3, 1)
        self.scattering_albedo = scattering_albedo.cast(3, 3)
        self.emission = emission.cast(6, 3)
        self.environment = environment.cast(3, 3)
        self.phase_sampler = phase_sampler.cast(6, 4)
        self.boundary = boundary.cast(6, 2)
        self.majorant = majorant.cast(6, 2)
        self.ds_epsilon = ds_epsilon
        self.max_bounces = max_bounces

    def _lower_map(self):
        lw = _Lowered()
        lw.set_sigma(self.sigma, multiple_scattering=True)
        lw.set_boundary(self.boundary)
        lw.set_majorant(self.majorant)
        lw.set_albedo(self.scattering_albedo)
        lw.set_emission(self.emission)
        lw.set_environment(self.environment)
        lw.set_phase_sampler(self.phase_sampler)
        lw.medium.ds_epsilon = self.ds_epsilon
        lw.medium.max_bounces = int(self.max_bounces)
        return L.VR_MAP_PATH_DT, lw


class DRTPathIntegrator(_RayMap):
    """NEE path tracer whose backward is the reservoir estimator of
    vr_path_integrator_NEE_DRT.h:46-269 (_maps.py:2981-3022,
    maps/radiance_NEE_DRT.h).  The forward is the NEE integrator
    (vr_path_integrator_NEE.h).  max_bounces as in DeltatrackingNEEPathIntegrator.
    """

    def __init__(self, sigma: MapBase, scattering_albedo: MapBase, emission: MapBase, environment: MapBase,
                 environment_sampler: MapBase, phase: MapBase, phase_sampler: MapBase, boundary: MapBase,
                 majorant: MapBase, transmittance: MapBase, max_bounces: int = -1):
        super().__init__(6, 3)
        self.sigma = sigma.cast(3, 1)
        self.scattering_albedo = scattering_albedo.cast(3, 3)
        self.emission = emission.cast(6, 3)
        self.environment = environment.cast(3, 3)
        self.environment_sampler = environment_sampler.cast(6, 6)
        self.phase = phase.cast(9, 1)
        self.phase_sampler = phase_sampler.cast(6, 4)
        self.boundary = boundary.cast(6, 2)
        self.majorant = majorant.cast(6, 2)
        self.transmittance = transmittance.cast(6, 1)
        self.max_bounces = max_bounces

    def _lower_map(self):
        lw = _Lowered()
        lw.set_sigma(self.sigma, multiple_scattering=True)
        lw.set_boundary(self.boundary)
        lw.set_majorant(self.majorant)
        lw.set_albedo(self.scattering_albedo)
        lw.set_emission(self.emission)
        lw.set_environment(self.environment)
        lw.set_environment_sampler(self.environment_sampler)
        lw.set_phase(self.phase, self.phase_sampler)
        lw.set_transmittance(self.transmittance)
        lw.medium.max_bounces = int(self.max_bounces)
        return L.VR_MAP_NEE_DRT, lw


# ------------------------------------------------------------- autograd --
def _grad_targets(lw: _Lowered, params: typing.Sequence[torch.Tensor]):
    """Fresh zero gradient buffers for the parameters that require grad."""
    if hasattr(lw, "grad_targets"):  # a lowered scene: one buffer per media grid
        return lw.grad_targets(params)
    grads = {}
    if lw.sigma is not None and lw.sigma.grid.requires_grad:
        grads["sigma"] = torch.zeros_like(lw.sigma.grid)
        lw.medium.sigma = lw.sigma.view(grads["sigma"])
    if lw.albedo is not None and lw.albedo.grid.requires_grad:
        grads["albedo"] = torch.zeros_like(lw.albedo.grid)
        lw.medium.albedo = lw.albedo.view(grads["albedo"])
    out = []
    for p in params:
        if lw.sigma is not None and p is lw.sigma.grid:
            out.append(grads.get("sigma"))
        elif lw.albedo is not None and p is lw.albedo.grid:
            out.append(grads.get("albedo"))
        else:
            out.append(None)
    return out


def _bw_mode() -> int:
    return L.VR_BW_CONSISTENT if _internal.backward_settings.estimator == "consistent" else L.VR_BW_REFERENCE


class AutogradMapFunction(torch.autograd.Function):
    """Map evaluation on explicit inputs (_maps.py:1598-1613)."""

    @staticmethod
    def forward(ctx, x, map_object, *params):
        kind, lw = map_object._lower_map()
        in_dim, out_dim = L.MAP_DIMS[kind]
        xc = x.detach().to(torch.float32).contiguous()
        n = xc.numel() // in_dim
        out = torch.empty(*x.shape[:-1], out_dim, device=x.device, dtype=torch.float32)
        seeds = _internal.get_seeds()
        launch = L.launch(seeds, bw_mode=_bw_mode(), seed_mode=L.VR_SEED_MAP)
        if kind == L.MAP_SCENE:
            L.check(L.load().vr_scene_map_fw(lw.scene, L.ptr(xc), n, L.ptr(out), launch, L.stream_handle(x.device)),
                    "vr_scene_map_fw")
        else:
            L.check(L.load().vr_map_fw(lw.medium, kind, L.ptr(xc), n, L.ptr(out), launch, L.stream_handle(x.device)),
                    "vr_map_fw")
        ctx.save_for_backward(xc)
        ctx.map_object = map_object
        ctx.seeds = seeds
        return out

    @staticmethod
    def backward(ctx, og):
        xc, = ctx.saved_tensors
        map_object = ctx.map_object
        params = [p for p in map_object.parameters() if p.requires_grad]
        with _ENGINE_LOCK:
            kind, lw = map_object._lower_map()
            grads = _grad_targets(lw, params)
            in_dim, _ = L.MAP_DIMS[kind]
            seeds = ctx.seeds if _internal.backward_settings.seeds == "replay" else _internal.get_seeds()
            launch = L.launch(seeds, bw_mode=_bw_mode(), seed_mode=L.VR_SEED_MAP)
            og = og.to(torch.float32).contiguous()
            if kind == L.MAP_SCENE:
                L.check(L.load().vr_scene_map_bw(lw.scene, L.ptr(xc), L.ptr(og), xc.numel() // in_dim, launch,
                                                 L.stream_handle(xc.device)), "vr_scene_map_bw")
            else:
                L.check(L.load().vr_map_bw(lw.medium, kind, L.ptr(xc), L.ptr(og), xc.numel() // in_dim, launch,
                                           L.stream_handle(xc.device)), "vr_map_bw")
        return (None, None, *grads)


class AutogradCaptureFunction(torch.autograd.Function):
    """Camera capture of a field (_maps.py:1616-1642).

    The backward returns fresh gradients for the trigger parameters and torch
    adds them into .grad once (the reference's backward both accumulates into
    .grad in place and returns it, _maps.py:1049-1058,1642).  A `debug_out`
    tensor receives per-sensor trilinear lookups in the forward and trilinear
    gradient scatters (or lookups, backward_mode(debug="lookups")) in the
    backward.
    """

    @staticmethod
    def forward(ctx, sensors, batch_size, fw_samples, bw_samples, field, capture_object, debug_out, *parameters):
        with _ENGINE_LOCK:
            kind, lw = field._lower_map()
            in_dim, out_dim = L.MAP_DIMS[kind]
            if in_dim != 6:
                raise NotImplementedError("only 6-input fields can be captured")
            cam, keep = capture_object._camera(sensors)
            total = cam.num_sensors if sensors is not None else capture_object.num_sensors()
            shape = (*sensors.shape[:-1],) if sensors is not None else tuple(capture_object.index_shape)
            out = torch.empty(*shape, out_dim, device=capture_object.poses.device, dtype=torch.float32)
            seeds = _internal.get_seeds()
            bs = total if batch_size is None else int(batch_size)
            lib = L.load()
            workspace = L.capture_workspace(out.device)
            for start in range(0, total, max(1, bs)):
                launch = L.launch(seeds, samples=fw_samples, bw_mode=_bw_mode(), start_index=start,
                                  count=min(bs, total - start), debug_out=debug_out, workspace=workspace)
                if kind == L.MAP_SCENE:
                    L.check(lib.vr_scene_capture_fw(lw.scene, cam, launch, L.ptr(out), L.stream_handle(out.device)),
                            "vr_scene_capture_fw")
                else:
                    L.check(lib.vr_capture_fw(lw.medium, kind, cam, launch, L.ptr(out), L.stream_handle(out.device)),
                            "vr_capture_fw")
        ctx.field = field
        ctx.capture_object = capture_object
        ctx.sensors = sensors
        ctx.batch_size = batch_size
        ctx.bw_samples = bw_samples
        ctx.fw_samples = fw_samples
        ctx.seeds = seeds
        ctx.debug_out = debug_out
        ctx.n_params = len(parameters)
        ctx.save_for_backward(*parameters)
        return out

    @staticmethod
    def backward(ctx, og):
        params = list(ctx.saved_tensors)
        with _ENGINE_LOCK:
            kind, lw = ctx.field._lower_map()
            grads = _grad_targets(lw, params)
            cam, keep = ctx.capture_object._camera(ctx.sensors)
            total = cam.num_sensors if ctx.sensors is not None else ctx.capture_object.num_sensors()
            mode = _internal.backward_settings.seeds
            if mode == "replay":
                seeds, samples, seed_mode = ctx.seeds, ctx.fw_samples, L.VR_SEED_CAPTURE_FW
            elif mode == "fresh":
                seeds, samples, seed_mode = _internal.get_seeds(), ctx.bw_samples, L.VR_SEED_CAPTURE_FW
            else:
                seeds, samples, seed_mode = _internal.get_seeds(), ctx.bw_samples, L.VR_SEED_CAPTURE_BW
            og = og.to(torch.float32).contiguous()
            bs = total if ctx.batch_size is None else int(ctx.batch_size)
            lib = L.load()
            for start in range(0, total, max(1, bs)):
                launch = L.launch(seeds, samples=samples, bw_mode=_bw_mode(), seed_mode=seed_mode,
                                  start_index=start, count=min(bs, total - start), debug_out=ctx.debug_out,
                                  debug_kind=int(_internal.backward_settings.debug == "lookups"))
                if kind == L.MAP_SCENE:
                    L.check(lib.vr_scene_capture_bw(lw.scene, cam, launch, L.ptr(og), L.stream_handle(og.device)),
                            "vr_scene_capture_bw")
                else:
                    L.check(lib.vr_capture_bw(lw.medium, kind, cam, launch, L.ptr(og), L.stream_handle(og.device)),
                            "vr_capture_bw")
        return (None, None, None, None, None, None, None, *grads)


# ---------------------------------------------------------------- sensors --
class SensorsBase(MapBase):
    """Sensor collections with capture() (_maps.py:1690-1762)."""

    def __init__(self, index_shape: typing.List[int], output_dim: int):
        super().__init__(len(index_shape), output_dim)
        self.index_shape = list(index_shape)

    def num_sensors(self) -> int:
        return int(np.prod(self.index_shape))

    def capture(self, field: MapBase, sensors_batch: typing.Optional[torch.Tensor] = None,
                batch_size: typing.Optional[int] = None, fw_samples: int = 1, bw_samples: int = 1,
                debug_out: typing.Optional[torch.Tensor] = None) -> torch.Tensor:
        field = field.cast(input_dim=self.output_dim)
        triggers = [p for p in field.parameters() if p.requires_grad]
        return AutogradCaptureFunction.apply(sensors_batch, batch_size, fw_samples, bw_samples, field, self, debug_out,
                                             *triggers)

    def random_sensors(self, batch_size: int, out: typing.Optional[torch.Tensor] = None) -> torch.Tensor:
        """Uniform random sensor indices [batch_size, dims] (_maps.py:1751-1762),
        drawn on the device by functions.random_ids with the reference's stream."""
        from . import functions
        if out is None:
            out = torch.empty(batch_size, len(self.index_shape), dtype=torch.int64, device=self.poses.device)
        return functions.random_ids(batch_size, self.index_shape, out=out)


class PerspectiveCameraSensor(SensorsBase):
    """Pinhole cameras from [N,9] poses (origin, direction, up) (_maps.py:1765-1786)."""

    def __init__(self, width: int, height: int, poses: torch.Tensor, *, fov: float = np.pi / 4, jittered: bool = False):
        super().__init__([len(poses), height, width], 6)
        assert poses.shape[-1] == 9
        self.poses = poses.to(torch.float32).contiguous()
        self.fov = fov
        self.znear = 0.001
        self.width = width
        self.height = height
        self.generation_mode = 0 if not jittered else 1

    def _camera(self, sensors: typing.Optional[torch.Tensor]):
        c = L.VrCamera()
        c.poses = self.poses.data_ptr()
        c.num_poses = self.poses.shape[0]
        c.width = self.width
        c.height = self.height
        c.jittered = self.generation_mode
        c.fov = self.fov
        c.znear = self.znear
        keep = [self.poses]
        if sensors is not None:
            s = sensors.to(torch.int64).contiguous()
            assert s.shape[-1] == 3, "sensor indices are (pose, row, col)"
            c.sensors = s.data_ptr()
            c.num_sensors = s.numel() // 3
            keep.append(s)
        else:
            c.sensors = None
            c.num_sensors = 0
        return c, keep
