"""DependencySet recipes for the volume path (rendervous _objects.py, _internal.py:225-333).

Each builder adds the maps a volume renderer needs to a DependencySet, with the
reference's defaults: a normalised box boundary, Grid3D sigma from a tensor, the
constant majorant [max(sigma), 1e6], RT transmittance, HG phase, uniform
environment sampling.  ``medium_majorant_coarse_grid`` adds the coarse-grid
majorant of this library.
"""
from __future__ import annotations

import math
import typing

import torch

from . import maps as _maps

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


class _Lazy:
    """A tensor recomputed from a callable on forward_dependency (Reparameter, _internal.py:223-254)."""

    def __init__(self, fn: typing.Callable[[], torch.Tensor]):
        self.fn = fn
        self.value = fn()

    def refresh(self):
        with torch.no_grad():
            self.value.copy_(self.fn())


class DependencySet:
    """Named fields filled by builder functions (_internal.py:256-333)."""

    def __init__(self, parent: typing.Optional["DependencySet"] = None):
        object.__setattr__(self, "_fields", {})
        object.__setattr__(self, "_lazy", [])
        object.__setattr__(self, "_parent", parent)

    def ensures(self, field_name: str, field_type) -> bool:
        return (field_name in self._fields and isinstance(self._fields[field_name], field_type)) or \
            (self._parent is not None and self._parent.ensures(field_name, field_type))

    def assert_ensures(self, field_name: str, field_type):
        assert self.ensures(field_name, field_type), f"Required field {field_name} of type {field_type} not found."

    def branch(self) -> "DependencySet":
        return DependencySet(parent=self)

    def add_parameters(self, **kwargs):
        for k, v in kwargs.items():
            if not isinstance(v, torch.nn.Module) and callable(v):
                lz = _Lazy(v)
                self._lazy.append(lz)
                v = lz.value
            self._fields[k] = v

    def requires(self, builder, **kwargs):
        builder(self, **kwargs)

    def forward_dependency(self):
        if self._parent is not None:
            self._parent.forward_dependency()
        for lz in self._lazy:
            lz.refresh()

    def __getattr__(self, item):
        fields = object.__getattribute__(self, "_fields")
        if item in fields:
            return fields[item]
        parent = object.__getattribute__(self, "_parent")
        if parent is not None:
            return getattr(parent, item)
        raise AttributeError(item)


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
        print("[WARNING] Neither Grid3D sigma nor tensor was found, -1..1 box used instead.")
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
        print("[WARNING] fields for ratio-tracking transmittance not found, trying DDA instead")
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


def medium_environment_sampler_quadtree(ds: DependencySet, *, projection: str = "xr", levels: int = 10):
    """_objects.py:180-199: importance sampling of an equirectangular environment
    image by a quadtree of its solid-angle-weighted luminance sum."""
    ds.requires(medium_environment_tensor, projection=projection)
    if projection != "xr":
        raise NotImplementedError(f"quadtree from projection {projection} (only 'xr')")
    skybox = ds.environment_tensor
    if len(skybox.shape) != 3:
        raise NotImplementedError("the quadtree sampler needs an environment image")

    def build_quadtree():
        from . import functions
        with torch.no_grad():
            resolution = 1 << levels
            densities = functions.resample_img(skybox.sum(-1, keepdim=True).contiguous(), (resolution, resolution))
            py = torch.arange(resolution, dtype=torch.float64)
            w = torch.cos(py * math.pi / resolution) - torch.cos((py + 1) * math.pi / resolution)
            densities *= w.to(densities.device, torch.float32)[:, None, None]
            densities /= max(float(densities.sum()), 0.00000001)
            return functions.create_density_quadtree(densities)

    ds.add_parameters(environment_sampler_quadtree=build_quadtree)
    dir_sampling = _maps.XRQuadtreeRandomDirection(input_dim=6, densities=ds.environment_sampler_quadtree, levels=levels)
    ds.add_parameters(environment_sampler=_maps._functionsampler(point_sampler=dir_sampling, function_map=ds.environment))


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
        print("[WARNING] a HG phase could not be created, assuming isotropic instead.")
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
        print("[WARNING] a HG phase sampler could not be created, assuming isotropic instead.")
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
