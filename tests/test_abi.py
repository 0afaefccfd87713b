"""CPU checks of the C ABI and the host-side lowering (no GPU calls)."""
import ctypes
import os
import re

import numpy as np
import pytest
import torch

from helpers import L, O, ROOT

import rendervous_amd as rv


def _declared(header: str, prefix: str):
    text = open(os.path.join(ROOT, "include", header)).read()
    return sorted(set(re.findall(r"\b(" + prefix + r"[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    names = _declared("vr_api.h", "vr_")
    assert "vr_capture_fw" in names and "vr_adam_step" in names
    lib = L.load()  # loads the HIP library; no device call is made
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(L.EXPORTED)
    assert lib.vr_abi_version() == 1


def test_oracle_exports_the_same_interface():
    """Every entry point of the path has an oracle twin; the radiance-cache MLP
    (vr_nrc_*) has no reference counterpart and is checked against PyTorch FP32."""
    lib = O.load()
    for n in _declared("vr_api.h", "vr_"):
        if n.startswith("vr_nrc_"):
            continue
        assert hasattr(lib, "oracle_" + n[3:]), n


def test_struct_layouts_match_the_header():
    # offsets that a C compiler gives the header's structures (x86-64 SysV)
    assert ctypes.sizeof(L.VrTensor) == 8 + 8 + 32 + 32 + 8
    assert ctypes.sizeof(L.VrGrid) == ctypes.sizeof(L.VrTensor) + 24
    assert L.VrLaunch.start_index.offset == 32
    assert ctypes.sizeof(L.VrLaunch) == 56
    assert L.VrCamera.sensors.offset == 32
    assert ctypes.sizeof(L.VrBvhNode) == 32 and ctypes.sizeof(L.VrPatch) == 32
    assert ctypes.sizeof(L.VrScene) == 224 and L.VrScene.media.offset == 88 and L.VrScene.env_image.offset == 136


def test_argument_errors_are_reported_without_a_device():
    lib = L.load()
    # null grid -> VR_ERR_ARG before any device work
    st = lib.vr_grid_fw(None, None, 0, None, None)
    assert st == L.VR_ERR_ARG
    assert b"NULL" in lib.vr_last_error()
    g = L.VrGrid()
    g.t.ndim = 3
    g.t.data = 1
    assert lib.vr_grid_fw(g, None, 1, None, None) == L.VR_ERR_ARG
    lau = L.launch((1, 2, 3, 4))
    assert lib.vr_map_fw(None, 99, None, 0, None, lau, None) == L.VR_ERR_UNSUPPORTED
    assert lib.vr_adam_step(None, None, None, None, 4, 1e-3, 0.9, 0.999, 1e-8, 0, 0, None) == L.VR_ERR_ARG
    assert lib.vr_adam_step_bounded(None, None, None, None, 4, 1e-3, 0.9, 0.999, 1e-8, 1, 1.0, 0.0, 0, None) == L.VR_ERR_ARG


def test_nee_integrator_lowers_to_a_medium_descriptor():
    sigma = torch.rand(9, 9, 9, 1)
    ds = rv.objects.DependencySet()
    ds.add_parameters(sigma_tensor=sigma, scattering_albedo_tensor=torch.tensor([0.9, 0.8, 0.7]),
                      environment_tensor=torch.tensor([0.1, 0.1, 0.1]), phase_g_tensor=torch.tensor([0.6]))
    ds.requires(rv.objects.medium_transmittance_DT)
    ds.requires(rv.objects.medium_radiance_path_integrator_NEE_DT, max_bounces=16)
    kind, lw = ds.radiance._lower_map()
    m = lw.medium
    assert kind == L.VR_MAP_NEE_DT
    assert list(m.box_max) == pytest.approx([0.5, 0.5, 0.5])
    assert m.majorant_kind == L.VR_MAJ_CONST
    assert m.majorant_value == pytest.approx(float(sigma.max()))
    assert m.majorant_distance == pytest.approx(1e6)
    assert m.phase_kind == L.VR_PHASE_HG_MIX and m.phase_g == pytest.approx(0.6)
    assert list(m.albedo_c) == pytest.approx([0.9, 0.8, 0.7])
    assert m.transmittance_kind == L.VR_T_DT
    assert m.max_bounces == 16
    assert m.sigma.t.shape[0] == 9 and m.sigma.t.stride[2] == 1


@pytest.mark.parametrize("recipe,tkind", [("medium_transmittance_GDT", L.VR_T_GDT), ("medium_transmittance_GRT", L.VR_T_GRT)])
def test_grid_tracking_recipes_lower_to_their_transmittance(recipe, tkind):
    sigma = torch.rand(7, 7, 7, 1)
    ds = rv.objects.DependencySet()
    ds.add_parameters(sigma_tensor=sigma, scattering_albedo_tensor=torch.tensor([0.9, 0.8, 0.7]),
                      environment_tensor=torch.tensor([0.1, 0.1, 0.1]), phase_g_tensor=torch.tensor([0.6]))
    ds.requires(getattr(rv.objects, recipe))
    kind, lw = ds.transmittance._lower_map()
    assert kind == (L.VR_MAP_T_GDT if tkind == L.VR_T_GDT else L.VR_MAP_T_GRT)
    assert lw.medium.transmittance_kind == tkind
    ds.requires(rv.objects.medium_radiance_path_integrator_NEE_DT, max_bounces=4)
    kind, lw = ds.radiance._lower_map()
    assert kind == L.VR_MAP_NEE_DT and lw.medium.transmittance_kind == tkind


def test_unsupported_graphs_raise():
    sigma = torch.rand(5, 5, 5, 1)
    g = rv.Grid3D(sigma, *rv.normalized_box(sigma))
    other = rv.Grid3D(sigma.clone(), *rv.normalized_box(sigma))
    b = rv.RayBoxIntersection(*rv.normalized_box(sigma))
    t = rv.DeltatrackingTransmittance(other, b, rv.const[1.0, 1e6])
    nee = rv.DeltatrackingNEEPathIntegrator(g, rv.const[0.9, 0.9, 0.9], rv.ZERO, rv.const[1.0, 1.0, 1.0],
                                            rv._functionsampler(rv.UniformDirectionSampler(6), rv.const[1.0, 1.0, 1.0]),
                                            rv.HGPhase(rv.const[0.5]), rv.VHGPhaseSampler(rv.const[0.5]), b,
                                            rv.const[1.0, 1e6], t, 0.1)
    with pytest.raises(NotImplementedError):
        nee._lower_map()


def test_seed_contract():
    rv.seed(7)
    a = rv.get_seeds()
    rv.seed(7)
    b = rv.get_seeds()
    assert a == b
    assert all(129 <= s < (1 << 30) for s in a)
