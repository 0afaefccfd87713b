"""CPU checks of the C ABI and the host-side lowering (no GPU calls)."""
import ctypes
import os
import re

import numpy as np
import pytest
import torch

from helpers import L, O, ROOT, Built, Camera, Scene, pose

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
    assert lib.vr_abi_version() == L.ABI_VERSION == 3


def test_oracle_exports_the_same_interface():
    """Every entry point of the path has an oracle twin; the radiance-cache MLP
    (vr_nrc_*) has no reference counterpart and is checked against PyTorch FP32,
    and the allocation-free capture is checked against vr_capture_fw."""
    lib = O.load()
    for n in _declared("vr_api.h", "vr_"):
        if n.startswith("vr_nrc_") or n == "vr_capture_fw_noalloc":
            continue
        assert hasattr(lib, "oracle_" + n[3:]), n


def test_struct_layouts_match_the_header():
    # offsets that a C compiler gives the header's structures (x86-64 SysV)
    assert ctypes.sizeof(L.VrTensor) == 8 + 8 + 32 + 32 + 8
    assert ctypes.sizeof(L.VrGrid) == ctypes.sizeof(L.VrTensor) + 24
    assert L.VrLaunch.start_index.offset == 32
    assert L.VrLaunch.workspace.offset == 56 and ctypes.sizeof(L.VrLaunch) == 72
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


# ------------------------------------------- host-side rejection of bad input --
# Each case is rejected while validating arguments, before any device work: the
# buffers are host arrays that the library never dereferences.
def _nee_args(**scene_kw):
    sc = Scene(sigma=np.ones((4, 4, 4, 1), np.float32), **scene_kw)
    b = Built(sc, "cpu")
    cam = Camera(pose((0.0, 0.0, 2.5))[None], 8, 8, device="cpu")
    out = np.zeros((64, 3), np.float32)
    return b, cam, out


@pytest.mark.parametrize("value,distance", [(float("inf"), 1e6), (float("nan"), 1e6), (-1.0, 1e6), (2.0, 0.0),
                                            (2.0, float("nan")), (2.0, -3.0)])
def test_bad_constant_majorants_are_rejected(value, distance):
    lib = L.load()
    b, cam, out = _nee_args(majorant_value=value, majorant_distance=distance)
    for kind in (L.VR_MAP_NEE_DT, L.VR_MAP_T_DT, L.VR_MAP_T_RT, L.VR_MAP_COLLISION_DT):
        st = lib.vr_capture_fw(b.medium, kind, cam.c, L.launch((1, 2, 3, 4), samples=1), O.nptr(out), None)
        assert st == L.VR_ERR_ARG, kind
        assert b"majorant" in lib.vr_last_error()
        st = lib.vr_map_fw(b.medium, kind, O.nptr(out), 4, O.nptr(out), L.launch((1, 2, 3, 4)), None)
        assert st == L.VR_ERR_ARG


def test_bad_medium_majorant_is_rejected():
    lib = L.load()
    b, cam, out = _nee_args(medium_majorant=float("inf"))
    st = lib.vr_map_fw(b.medium, L.VR_MAP_MEDIUM_DT, O.nptr(out), 2, O.nptr(out), L.launch((1, 2, 3, 4)), None)
    assert st == L.VR_ERR_ARG and b"medium majorant" in lib.vr_last_error()


def test_cell_table_plane_limit_is_enforced():
    """Cell-table lookups index with 24-bit multiplies: a table is refused for a
    grid whose xy planes hold 2^24 cells or more (the grid is only described,
    never read)."""
    lib = L.load()
    b, cam, out = _nee_args()
    table = np.zeros(64, np.float32)
    b.medium.sigma_cells = O.nptr(table) if O.nptr(table) % 16 == 0 else O.nptr(table[2:])
    t = b.medium.sigma.t
    for i, (n, st_) in enumerate(zip((2, 4097, 4097, 1), (4097 * 4097, 4097, 1, 1))):
        t.shape[i] = n
        t.stride[i] = st_
    st = lib.vr_capture_fw(b.medium, L.VR_MAP_NEE_DT, cam.c, L.launch((1, 2, 3, 4), samples=1), O.nptr(out), None)
    assert st == L.VR_ERR_ARG and b"cell table" in lib.vr_last_error()


def test_sample_indices_beyond_32_bits_are_rejected():
    """Sample s of sensor i draws the stream of flat index s * sensors + i (32
    bits): 64 x 2^26 = 2^32 samples is the largest capture, one more sensor's
    worth is rejected (replayed backward too; the fresh-seed backward seeds one
    stream per sensor)."""
    lib = L.load()
    b, _, out = _nee_args()
    p = np.ascontiguousarray(pose((0.0, 0.0, 2.5))[None], np.float32)
    cam = L.VrCamera()
    cam.poses = O.nptr(p)
    cam.num_poses = 1
    cam.width = 1 << 13
    cam.height = (1 << 13) + 1  # 2^26 + 2^13 sensors
    cam.fov = 0.7
    cam.znear = 0.001
    for kind in (L.VR_MAP_NEE_DT, L.VR_MAP_SS_RADIANCE, L.VR_MAP_T_DT):
        st = lib.vr_capture_fw(b.medium, kind, cam, L.launch((1, 2, 3, 4), samples=64), O.nptr(out), None)
        assert st == L.VR_ERR_UNSUPPORTED and b"2^32" in lib.vr_last_error(), kind
        st = lib.vr_capture_bw(b.medium, kind, cam, L.launch((1, 2, 3, 4), samples=64, seed_mode=L.VR_SEED_CAPTURE_FW),
                               O.nptr(out), None)
        assert st == L.VR_ERR_UNSUPPORTED, kind
    cam.height = 1 << 19  # 2^32 sensors + ... : too many sensors for any seed mode
    cam.width = (1 << 13) + 1
    st = lib.vr_capture_bw(b.medium, L.VR_MAP_NEE_DT, cam, L.launch((1, 2, 3, 4), samples=1, seed_mode=L.VR_SEED_CAPTURE_BW),
                           O.nptr(out), None)
    assert st == L.VR_ERR_UNSUPPORTED and b"2^32 sensors" in lib.vr_last_error()


@pytest.mark.parametrize("nbytes,offset", [(512, 0), (1024, 64)])
def test_capture_workspace_must_fit(nbytes, offset):
    lib = L.load()
    b, cam, out = _nee_args(majorant_value=1.0)
    ws = np.zeros(4096, np.uint8)
    base = (O.nptr(ws) + 255) // 256 * 256 + offset
    lau = L.launch((1, 2, 3, 4), samples=2)
    lau.workspace = base
    lau.workspace_bytes = nbytes
    st = lib.vr_capture_fw(b.medium, L.VR_MAP_NEE_DT, cam.c, lau, O.nptr(out), None)
    assert st == L.VR_ERR_ARG and b"workspace" in lib.vr_last_error()


@pytest.mark.parametrize("bad", [float("nan"), float("inf")])
def test_map_layer_refuses_a_non_finite_density(bad):
    """The map layer checks a density once, when it first builds its majorant
    grid, before any library call."""
    sigma = torch.rand(9, 9, 9, 1)
    sigma[4, 4, 4, 0] = bad
    g = rv.Grid3D(sigma, *rv.normalized_box(sigma))
    with pytest.raises(ValueError, match="NaN or infinite"):
        rv.MajorantGrid(g, 4).majorant_tensor()
