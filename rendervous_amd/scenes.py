"""Medium segment integrator (7 -> 12) of the scene API.

DTMediumPathIntegrator (rendervous _scenes.py:1614-1632) maps a segment
(x, w, d) inside a medium to (xo, wo, W, A) by delta tracking with a scalar
majorant (include/sc_medium_path_integrator_DT_fw.h / _bw.h).  The surface
scene around it (meshes, ray-query BVH) is not part of this library; the
integrator is evaluated on explicit segments.
"""
from __future__ import annotations

import typing

import torch

from . import _lib as L
from .maps import (ConstantMap, Grid3D, MapBase, VHGPhaseSampler, ZERO, _Lowered, _RayMap)

__all__ = ["HGPhaseSampler", "DTMediumPathIntegrator"]


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
