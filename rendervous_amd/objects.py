"""DependencySet recipes for the volume path (rendervous _objects.py, _internal.py:225-333).

Each builder adds the maps a volume renderer needs to a DependencySet, with the
reference's defaults: a normalised box boundary, Grid3D sigma from a tensor, the
constant majorant [max(sigma), 1e6], RT transmittance, HG phase, uniform
environment sampling.  ``medium_majorant_coarse_grid`` adds the coarse-grid
majorant of this library.
"""
from __future__ import annotations

import collections
import math
import typing
import warnings

import torch

from . import maps as _maps

_ABSENT = object()

__all__ = [
    "DependencySet", "medium_box_AABB", "medium_box_normalized", "medium_box", "medium_boundary_box", "medium_boundary",
    "build_map_from_tensor", "medium_sigma_tensor", "medium_sigma", "medium_scattering_albedo_tensor",
    "medium_scattering_albedo", "medium_emission", "medium_phase_g_tensor", "medium_phase_g", "medium_majorant_tensor",
    "medium_majorant_grid", "medium_majorant_coarse_grid", "medium_majorant", "medium_transmittance_DDA", "medium_transmittance_GDT", "medium_transmittance_GRT",
    "medium_transmittance_DT", "medium_transmittance_RT", "medium_transmittance", "medium_environment_tensor", "medium_environment_sampler_quadtree",
    "medium_environment", "medium_environment_sampler_uniform", "medium_environment_sampler_sun",
    "medium_environment_sampler", "medium_phase_iso", "medium_phase_HG", "medium_phase", "medium_phase_sampler_HG",
    "medium_phase_sampler_uniform", "medium_phase_sampler", "medium_collision_sampler_DT", "medium_collision_sampler",
    "medium_exitance_radiance_single_scatter", "medium_radiance_collision_integrator_MC",
    "medium_radiance_path_integrator_NEE_DT", "camera_sensors",
]


class _Derived:
    """A field given as a callable (the reference's Reparameter, _internal.py:224-250).

    The maps hold ``leaf``, a tensor of its own storage.  ``refresh`` re-runs the
    callable and copies its value into the leaf; when that value carries a graph,
    ``push_grad`` hands the gradient accumulated on the leaf back into it, so a
    parameter behind the callable receives its share."""

    def __init__(self, fn: typing.Callable[[], torch.Tensor]):
        self.fn = fn
        self.graph: typing.Optional[torch.Tensor] = None
        value = fn()
        self.leaf = value.detach().clone().requires_grad_(value.requires_grad)

    def refresh(self):
        value = self.fn()
        self.graph = value if value.requires_grad else None
        with torch.no_grad():
            self.leaf.copy_(value)
        self.leaf.requires_grad_(value.requires_grad)
        if self.leaf.grad is not None:
            self.leaf.grad.zero_()

    def push_grad(self):
        if self.graph is not None and self.leaf.grad is not None:
            self.graph.backward(self.leaf.grad)


class DependencySet:
    """Named fields that the recipes below read and add (the interface of
    _internal.py:256-333: add_parameters, requires, ensures, assert_ensures,
    branch, copy_parameters, forward_dependency, backward_dependency, attribute
    access).

    The fields of a set and of its ancestors form one ``ChainMap``: a branch
    pushes an empty layer of its own, so what it adds shadows its parent and
    stays invisible to it, while the parent's later additions show through.
    Fields given as callables become ``_Derived`` values kept in one registry
    shared by the whole tree; forward_dependency from any set re-evaluates all
    of them in the order they were added, and backward_dependency hands their
    gradients back."""

    def __init__(self, parent: typing.Optional["DependencySet"] = None):
        if parent is None:
            scope, registry = collections.ChainMap(), []
        else:
            scope, registry = parent._scope.new_child(), parent._registry
        object.__setattr__(self, "_scope", scope)
        object.__setattr__(self, "_registry", registry)

    def ensures(self, field_name: str, field_type) -> bool:
        """True when this set or an ancestor holds ``field_name`` as a ``field_type``."""
        return any(isinstance(layer.get(field_name, _ABSENT), field_type) for layer in self._scope.maps)

    def assert_ensures(self, field_name: str, field_type):
        if not self.ensures(field_name, field_type):
            raise AssertionError(f"the set holds no field {field_name!r} of type {field_type}")

    def branch(self) -> "DependencySet":
        return DependencySet(parent=self)

    def add_parameters(self, **kwargs):
        layer = self._scope.maps[0]
        for name, value in kwargs.items():
            if callable(value) and not isinstance(value, torch.nn.Module):
                derived = _Derived(value)
                self._registry.append(derived)
                value = derived.leaf
            layer[name] = value

    def copy_parameters(self, ds: "DependencySet"):
        """Adds every field visible from ``ds`` (its own layer wins over its ancestors')."""
        self.add_parameters(**dict(ds._scope))

    def requires(self, builder, **kwargs):
        builder(self, **kwargs)

    def forward_dependency(self):
        for derived in self._registry:
            derived.refresh()

    def backward_dependency(self):
        for derived in self._registry:
            derived.push_grad()

    def __getattr__(self, item):
        try:
            return object.__getattribute__(self, "_scope")[item]
        except KeyError:
            raise AttributeError(item) from None

    def __repr__(self):
        return "".join(f"{k}: {v!r}\n" for k, v in dict(self._scope).items())


def medium_box_AABB(ds: DependencySet, *, bmin=None, bmax=None):
    ds.add_parameters(box=(_maps.vec3(-1.0, -1.0, -1.0) if bmin is None else bmin,
                           _maps.vec3(1.0, 1.0, 1.0) if bmax is None else bmax))


def medium_box_normalized(ds: DependencySet, *, t):
    ds.add_parameters(box=_maps.normalized_box(t))


def medium_box(ds: DependencySet):
    if ds.ensures("box", tuple):
        return
    if ds.ensures("sigma_tensor", torch.Tensor):
        medium_box_normalized(ds, t=ds.sigma_tensor)
    elif ds.ensures("sigma", _maps.Grid3D):
        medium_box_normalized(ds, t=ds.sigma.grid)
    else:
        warnings.warn("no sigma grid or sigma tensor in the set: the medium box is [-1, 1]^3", stacklevel=2)
        medium_box_AABB(ds)


def medium_boundary_box(ds: DependencySet):
    ds.requires(medium_box)
    bmin, bmax = ds.box
    ds.add_parameters(boundary=_maps.RayBoxIntersection(bmin, bmax))


def medium_boundary(ds: DependencySet):
    if ds.ensures("boundary", _maps.MapBase):
        return
    medium_boundary_box(ds)


def build_map_from_tensor(ds: DependencySet, *, field_name: str):
    ds.requires(medium_box)
    bmin, bmax = ds.box
    ds.assert_ensures(field_name + "_tensor", torch.Tensor)
    tensor = getattr(ds, field_name + "_tensor")
    if len(tensor.shape) == 1:
        return _maps.const[tensor]
    return _maps.Grid3D(tensor, bmin, bmax)


def medium_sigma_tensor(ds: DependencySet):
    ds.add_parameters(sigma=build_map_from_tensor(ds, field_name="sigma"))


def medium_sigma(ds: DependencySet):
    if ds.ensures("sigma", _maps.MapBase):
        return
    medium_sigma_tensor(ds)


def medium_scattering_albedo_tensor(ds: DependencySet):
    ds.add_parameters(scattering_albedo=build_map_from_tensor(ds, field_name="scattering_albedo"))


def medium_scattering_albedo(ds: DependencySet):
    if ds.ensures("scattering_albedo", _maps.MapBase):
        return
    medium_scattering_albedo_tensor(ds)


def medium_emission(ds: DependencySet):
    if ds.ensures("emission", _maps.MapBase):
        return
    ds.assert_ensures("emission_tensor", torch.Tensor)
    ds.add_parameters(emission=_maps.const[ds.emission_tensor])


def medium_phase_g_tensor(ds: DependencySet):
    ds.add_parameters(phase_g=build_map_from_tensor(ds, field_name="phase_g"))


def medium_phase_g(ds: DependencySet):
    if ds.ensures("phase_g", _maps.MapBase):
        return
    medium_phase_g_tensor(ds)


def medium_majorant_tensor(ds: DependencySet):
    ds.assert_ensures("majorant_tensor", torch.Tensor)
    ds.add_parameters(majorant=_maps.const[ds.majorant_tensor])


def medium_majorant_grid(ds: DependencySet):
    """const[max(sigma), 1e6] re-evaluated on forward_dependency (_objects.py:96-101)."""
    ds.requires(medium_sigma)
    ds.assert_ensures("sigma", _maps.Grid3D)
    g: _maps.Grid3D = ds.sigma
    ds.add_parameters(majorant_tensor=lambda: torch.tensor([g.get_maximum(), 1000000.0], dtype=torch.float32))
    ds.add_parameters(majorant=_maps.const[ds.majorant_tensor])


def medium_majorant_coarse_grid(ds: DependencySet, *, resolution=16):
    """Coarse max-grid majorant traversed by DDA (this library's form of the 6->2 contract)."""
    ds.requires(medium_sigma)
    ds.assert_ensures("sigma", _maps.Grid3D)
    ds.add_parameters(majorant=_maps.MajorantGrid(ds.sigma, resolution))


def medium_majorant(ds: DependencySet):
    if ds.ensures("majorant", _maps.MapBase):
        return
    if ds.ensures("majorant_tensor", torch.Tensor):
        medium_majorant_tensor(ds)
    else:
        medium_majorant_grid(ds)


def medium_transmittance_DDA(ds: DependencySet):
    ds.requires(medium_sigma)
    ds.requires(medium_boundary)
    ds.assert_ensures("sigma", _maps.Grid3D)
    ds.add_parameters(transmittance=_maps.GridDDATransmittance(ds.sigma, ds.boundary))


def medium_transmittance_GDT(ds: DependencySet):
    """_objects.py:116-120"""
    ds.requires(medium_sigma)
    ds.requires(medium_boundary)
    ds.assert_ensures("sigma", _maps.Grid3D)
    ds.add_parameters(transmittance=_maps.GridDeltatrackingTransmittance(ds.sigma, ds.boundary))


def medium_transmittance_GRT(ds: DependencySet):
    """_objects.py:122-126"""
    ds.requires(medium_sigma)
    ds.requires(medium_boundary)
    ds.assert_ensures("sigma", _maps.Grid3D)
    ds.add_parameters(transmittance=_maps.GridRatiotrackingTransmittance(ds.sigma, ds.boundary))


def medium_transmittance_DT(ds: DependencySet):
    ds.requires(medium_sigma)
    ds.requires(medium_boundary)
    ds.requires(medium_majorant)
    ds.add_parameters(transmittance=_maps.DeltatrackingTransmittance(ds.sigma, ds.boundary, ds.majorant))


def medium_transmittance_RT(ds: DependencySet):
    ds.requires(medium_sigma)
    ds.requires(medium_boundary)
    ds.requires(medium_majorant)
    ds.add_parameters(transmittance=_maps.RatiotrackingTransmittance(ds.sigma, ds.boundary, ds.majorant))


def medium_transmittance(ds: DependencySet):
    """Default: ratio tracking (_objects.py:149-161)."""
    if ds.ensures("transmittance", _maps.MapBase):
        return
    try:
        medium_transmittance_RT(ds)
    except Exception:
        warnings.warn("ratio tracking needs a majorant the set cannot build: falling back to DDA transmittance", stacklevel=2)
        medium_transmittance_DDA(ds)


def medium_environment_tensor(ds: DependencySet, *, projection: str = "xr"):
    """_objects.py:164-173: a constant [3] or an equirectangular image [H, W, 3]."""
    ds.assert_ensures("environment_tensor", torch.Tensor)
    t = ds.environment_tensor
    if len(t.shape) == 1:
        ds.add_parameters(environment=_maps.const[t])
        return
    if projection != "xr":
        raise NotImplementedError(f"environment projection {projection} (only 'xr')")
    ds.add_parameters(environment=_maps.Image2D(t).after(_maps.xr_projection))


def medium_environment(ds: DependencySet):
    if ds.ensures("environment", _maps.MapBase):
        return
    medium_environment_tensor(ds)


def medium_environment_sampler_uniform(ds: DependencySet):
    ds.requires(medium_environment)
    dir_sampling = _maps.UniformDirectionSampler(6)
    ds.add_parameters(environment_sampler=_maps._functionsampler(point_sampler=dir_sampling, function_map=ds.environment))


def _xr_direction_density(image: torch.Tensor, levels: int) -> torch.Tensor:
    """Probability of each cell of an n x n equirectangular grid (n = 2^levels)
    for sampling directions towards an environment image: the sum of the image's
    channels, resampled to the grid, times the cell's solid angle.  Row r spans
    polar angles [r pi/n, (r+1) pi/n], so each of its cells subtends
    (2 pi/n)(cos(r pi/n) - cos((r+1) pi/n)); the factor 2 pi/n cancels in the
    normalisation.  The reference weights its rows the same way
    (_objects.py:185-191)."""
    from . import functions
    n = 1 << levels
    edges = torch.cos(torch.linspace(0.0, math.pi, n + 1, dtype=torch.float64))
    band = (edges[:-1] - edges[1:]).to(image.device, torch.float32)
    radiance = functions.resample_img(image.sum(-1, keepdim=True).contiguous(), (n, n))
    density = radiance * band.view(n, 1, 1)
    return density / max(float(density.sum()), 1e-8)


def medium_environment_sampler_quadtree(ds: DependencySet, *, projection: str = "xr", levels: int = 10):
    """Environment directions drawn from a density quadtree over an
    equirectangular image (the recipe of _objects.py:180-199: the quadtree is a
    derived field, rebuilt by forward_dependency)."""
    ds.requires(medium_environment_tensor, projection=projection)
    image = ds.environment_tensor
    if projection != "xr" or image.dim() != 3:
        raise NotImplementedError("the quadtree sampler needs an equirectangular ('xr') environment image [H, W, C]")

    def quadtree():
        from . import functions
        with torch.no_grad():
            return functions.create_density_quadtree(_xr_direction_density(image, levels))

    ds.add_parameters(environment_sampler_quadtree=quadtree)
    directions = _maps.XRQuadtreeRandomDirection(input_dim=6, densities=ds.environment_sampler_quadtree, levels=levels)
    ds.add_parameters(environment_sampler=_maps._functionsampler(point_sampler=directions, function_map=ds.environment))


def medium_environment_sampler_sun(ds: DependencySet, *, direction, radiance, probability: float = 0.5):
    ds.requires(medium_environment)
    ds.add_parameters(environment_sampler=_maps.DirectionalLightEnvironmentSampler(ds.environment, direction, radiance,
                                                                                   probability))


def medium_environment_sampler(ds: DependencySet):
    """_objects.py:204-212: the quadtree sampler for an environment image, uniform
    sampling for a constant environment (where the reference's quadtree build fails
    and its fallback sets nothing)."""
    if ds.ensures("environment_sampler", _maps.MapBase):
        return
    if ds.ensures("environment_tensor", torch.Tensor) and len(ds.environment_tensor.shape) == 3:
        medium_environment_sampler_quadtree(ds)
    else:
        medium_environment_sampler_uniform(ds)


def medium_phase_iso(ds: DependencySet):
    ds.add_parameters(phase=_maps.const[1.0 / (4 * math.pi)])


def medium_phase_HG(ds: DependencySet, *, mixture: bool = True):
    ds.requires(medium_phase_g)
    ds.add_parameters(phase=_maps.HGPhase(ds.phase_g, mixture=mixture))


def medium_phase(ds: DependencySet):
    if ds.ensures("phase", _maps.MapBase):
        return
    try:
        medium_phase_HG(ds)
    except Exception:
        warnings.warn("no phase_g in the set: the phase function is isotropic", stacklevel=2)
        medium_phase_iso(ds)


def medium_phase_sampler_HG(ds: DependencySet, *, mixture: bool = True):
    ds.requires(medium_phase_g)
    ds.add_parameters(phase_sampler=_maps.VHGPhaseSampler(ds.phase_g, mixture=mixture))


def medium_phase_sampler_uniform(ds: DependencySet):
    ds.requires(medium_phase)
    ds.add_parameters(phase_sampler=_maps.UniformDirectionSampler(6))


def medium_phase_sampler(ds: DependencySet):
    if ds.ensures("phase_sampler", _maps.MapBase):
        return
    try:
        medium_phase_sampler_HG(ds)
    except Exception:
        warnings.warn("no phase_g in the set: phase directions are sampled uniformly", stacklevel=2)
        medium_phase_sampler_uniform(ds)


def medium_collision_sampler_DT(ds: DependencySet, *, ds_epsilon: float = 0.1):
    ds.requires(medium_sigma)
    ds.requires(medium_boundary)
    ds.requires(medium_majorant)
    ds.add_parameters(collision_sampler=_maps.DeltatrackingCollisionSampler(ds.sigma, ds.boundary, ds.majorant, ds_epsilon))


def medium_collision_sampler(ds: DependencySet):
    if ds.ensures("collision_sampler", _maps.MapBase):
        return
    medium_collision_sampler_DT(ds)


def medium_exitance_radiance_single_scatter(ds: DependencySet):
    """exitance = albedo * phase * env_sampler * T (SingleScatteredRadiance, _maps.py:2761-2784)."""
    ds.requires(medium_scattering_albedo)
    ds.requires(medium_phase)
    ds.requires(medium_environment_sampler)
    ds.requires(medium_transmittance)
    ds.add_parameters(exitance_radiance=_maps.SingleScatteredRadiance(ds.scattering_albedo, ds.phase,
                                                                      ds.environment_sampler, ds.transmittance))


def medium_radiance_collision_integrator_MC(ds: DependencySet):
    ds.requires(medium_collision_sampler)
    ds.requires(medium_environment)
    if not ds.ensures("exitance_radiance", _maps.MapBase):
        medium_exitance_radiance_single_scatter(ds)
    ds.add_parameters(radiance=_maps.MCCollisionIntegrator(ds.collision_sampler, ds.exitance_radiance, ds.environment))


def medium_radiance_path_integrator_NEE_DT(ds: DependencySet, *, ds_epsilon: float = 0.1, max_bounces: int = -1):
    """_objects.py:334-363 (+ optional bounce cap)."""
    ds.requires(medium_sigma)
    if not ds.ensures("scattering_albedo", _maps.MapBase):
        try:
            ds.requires(medium_scattering_albedo)
        except AssertionError:
            ds.add_parameters(scattering_albedo=_maps.ZERO)
    if not ds.ensures("emission", _maps.MapBase):
        try:
            ds.requires(medium_emission)
        except AssertionError:
            ds.add_parameters(emission=_maps.ZERO)
    ds.requires(medium_boundary)
    ds.requires(medium_majorant)
    ds.requires(medium_environment)
    ds.requires(medium_environment_sampler)
    ds.requires(medium_phase)
    ds.requires(medium_phase_sampler)
    ds.requires(medium_transmittance)
    ds.add_parameters(radiance=_maps.DeltatrackingNEEPathIntegrator(
        ds.sigma, ds.scattering_albedo, ds.emission, ds.environment, ds.environment_sampler, ds.phase,
        ds.phase_sampler, ds.boundary, ds.majorant, ds.transmittance, ds_epsilon, max_bounces=max_bounces))


def medium_radiance_path_integrator_DT(ds: DependencySet, *, ds_epsilon: float = 0.1, max_bounces: int = -1):
    """_objects.py:309-331 (+ optional bounce cap)."""
    ds.requires(medium_sigma)
    if not ds.ensures("scattering_albedo", _maps.MapBase):
        try:
            ds.requires(medium_scattering_albedo)
        except AssertionError:
            ds.add_parameters(scattering_albedo=_maps.ZERO)
    if not ds.ensures("emission", _maps.MapBase):
        try:
            ds.requires(medium_emission)
        except AssertionError:
            ds.add_parameters(emission=_maps.ZERO)
    ds.requires(medium_boundary)
    ds.requires(medium_majorant)
    ds.requires(medium_environment)
    ds.requires(medium_phase_sampler)
    ds.add_parameters(radiance=_maps.DeltatrackingPathIntegrator(
        ds.sigma, ds.scattering_albedo, ds.emission, ds.environment, ds.phase_sampler, ds.boundary, ds.majorant,
        ds_epsilon, max_bounces=max_bounces))


def medium_radiance_path_integrator_NEE_DRT(ds: DependencySet, *, max_bounces: int = -1):
    """_objects.py:398-426 (+ optional bounce cap)."""
    ds.requires(medium_sigma)
    if not ds.ensures("scattering_albedo", _maps.MapBase):
        try:
            ds.requires(medium_scattering_albedo)
        except AssertionError:
            ds.add_parameters(scattering_albedo=_maps.ZERO)
    if not ds.ensures("emission", _maps.MapBase):
        try:
            ds.requires(medium_emission)
        except AssertionError:
            ds.add_parameters(emission=_maps.ZERO)
    ds.requires(medium_boundary)
    ds.requires(medium_majorant)
    ds.requires(medium_environment)
    ds.requires(medium_environment_sampler)
    ds.requires(medium_phase)
    ds.requires(medium_phase_sampler)
    ds.requires(medium_transmittance)
    ds.add_parameters(radiance=_maps.DRTPathIntegrator(
        ds.sigma, ds.scattering_albedo, ds.emission, ds.environment, ds.environment_sampler, ds.phase,
        ds.phase_sampler, ds.boundary, ds.majorant, ds.transmittance, max_bounces=max_bounces))


def medium_radiance(ds: DependencySet):
    """_objects.py:492-498: the delta-tracking path integrator unless a radiance is given."""
    if ds.ensures("radiance", _maps.MapBase):
        return
    ds.requires(medium_radiance_path_integrator_DT)


def camera_sensors(ds: DependencySet, *, width: int, height: int, jittered: bool = False, fov: float = math.pi / 4):
    ds.assert_ensures("camera_poses", torch.Tensor)
    ds.add_parameters(camera=_maps.PerspectiveCameraSensor(width, height, ds.camera_poses, fov=fov, jittered=jittered))
