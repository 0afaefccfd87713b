"""ctypes binding of the C ABI declared in include/vr_api.h.

The structures below mirror the header field by field.  The shared library is
loaded from the package directory (built in-tree by ``rendervous_amd.build``);
there is no fallback: every entry point raises if the library is missing or a
call returns a non-zero status.
"""
from __future__ import annotations

import ctypes
import os
import threading

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librendervous_amd.so")

# enums (include/vr_api.h)
VR_OK, VR_ERR_ARG, VR_ERR_UNSUPPORTED, VR_ERR_DEVICE = 0, 1, 2, 3
VR_FIELD_CONST, VR_FIELD_GRID = 0, 1
VR_PHASE_HG_MIX, VR_PHASE_HG, VR_PHASE_ISO = 0, 1, 2
VR_MAJ_CONST, VR_MAJ_GRID = 0, 1
VR_T_DT, VR_T_RT, VR_T_DDA, VR_T_GDT, VR_T_GRT = 0, 1, 2, 3, 4
VR_ENV_CONST, VR_ENV_XR = 0, 1
VR_ENVS_UNIFORM, VR_ENVS_QUADTREE = 0, 1
VR_MAP_RAYBOX = 1
VR_MAP_T_DDA = 2
VR_MAP_T_DT = 3
VR_MAP_T_RT = 4
VR_MAP_COLLISION_DT = 5
VR_MAP_SS_RADIANCE = 6
VR_MAP_MC_COLLISION = 7
VR_MAP_NEE_DT = 8
VR_MAP_MEDIUM_DT = 9
VR_MAP_PATH_DT = 10
VR_MAP_NEE_DRT = 11
VR_MAP_T_GDT = 12
VR_MAP_T_GRT = 13
VR_SEED_MAP, VR_SEED_CAPTURE_FW, VR_SEED_CAPTURE_BW = 0, 1, 2
VR_BW_REFERENCE, VR_BW_CONSISTENT = 0, 1
VR_SURF_PASS, VR_SURF_LAMBERT, VR_SURF_MIRROR, VR_SURF_FRESNEL, VR_SURF_NULL, VR_SURF_BLOCK = 0, 1, 2, 3, 4, 5
VR_EMIT_NONE, VR_EMIT_ISOTROPIC, VR_EMIT_LAMBERT = 0, 1, 2
VR_SCENE_MAX_MEDIA = 4
# Python-side map kind of a PathtracedScene (evaluated by the vr_scene_* entry points)
MAP_SCENE = 100

MAP_DIMS = {
    VR_MAP_RAYBOX: (6, 2),
    VR_MAP_T_DDA: (6, 1),
    VR_MAP_T_DT: (6, 1),
    VR_MAP_T_RT: (6, 1),
    VR_MAP_COLLISION_DT: (6, 3),
    VR_MAP_SS_RADIANCE: (6, 3),
    VR_MAP_MC_COLLISION: (6, 3),
    VR_MAP_NEE_DT: (6, 3),
    VR_MAP_MEDIUM_DT: (7, 12),
    VR_MAP_PATH_DT: (6, 3),
    VR_MAP_NEE_DRT: (6, 3),
    VR_MAP_T_GDT: (6, 1),
    VR_MAP_T_GRT: (6, 1),
    MAP_SCENE: (6, 3),
}

F3 = ctypes.c_float * 3


class VrTensor(ctypes.Structure):
    _fields_ = [
        ("data", ctypes.c_void_p),
        ("grad", ctypes.c_void_p),
        ("shape", ctypes.c_int64 * 4),
        ("stride", ctypes.c_int64 * 4),
        ("ndim", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


class VrGrid(ctypes.Structure):
    _fields_ = [("t", VrTensor), ("bmin", F3), ("bmax", F3)]


class VrMedium(ctypes.Structure):
    _fields_ = [
        ("sigma", VrGrid),
        ("box_min", F3),
        ("box_max", F3),
        ("albedo_kind", ctypes.c_int32),
        ("albedo_c", F3),
        ("albedo", VrGrid),
        ("emission_c", F3),
        ("env_c", F3),
        ("sun_dir", F3),
        ("sun_radiance", F3),
        ("sun_prob", ctypes.c_float),
        ("phase_kind", ctypes.c_int32),
        ("phase_g", ctypes.c_float),
        ("majorant_kind", ctypes.c_int32),
        ("majorant_value", ctypes.c_float),
        ("majorant_distance", ctypes.c_float),
        ("majorant_grid", ctypes.c_void_p),
        ("majorant_res", ctypes.c_int32 * 3),
        ("transmittance_kind", ctypes.c_int32),
        ("ds_epsilon", ctypes.c_float),
        ("max_bounces", ctypes.c_int32),
        ("medium_filter", ctypes.c_int32),
        ("gathering_c", F3),
        ("medium_majorant", ctypes.c_float),
        ("env_kind", ctypes.c_int32),
        ("env_sampler_kind", ctypes.c_int32),
        ("env_image", VrTensor),
        ("env_image_bmin", ctypes.c_float * 2),
        ("env_image_bmax", ctypes.c_float * 2),
        ("env_quadtree", ctypes.c_void_p),
        ("env_quadtree_levels", ctypes.c_int32),
        ("env_reserved", ctypes.c_int32),
        ("sigma_cells", ctypes.c_void_p),
        ("phase_g_kind", ctypes.c_int32),
        ("phase_reserved", ctypes.c_int32),
        ("phase_g_grid", VrGrid),
    ]


class VrCamera(ctypes.Structure):
    _fields_ = [
        ("poses", ctypes.c_void_p),
        ("num_poses", ctypes.c_int32),
        ("width", ctypes.c_int32),
        ("height", ctypes.c_int32),
        ("jittered", ctypes.c_int32),
        ("fov", ctypes.c_float),
        ("znear", ctypes.c_float),
        ("sensors", ctypes.c_void_p),
        ("num_sensors", ctypes.c_int64),
    ]


class VrLaunch(ctypes.Structure):
    _fields_ = [
        ("seeds", ctypes.c_uint32 * 4),
        ("samples", ctypes.c_int32),
        ("bw_mode", ctypes.c_int32),
        ("seed_mode", ctypes.c_int32),
        ("debug_kind", ctypes.c_int32),
        ("start_index", ctypes.c_int64),
        ("count", ctypes.c_int64),
        ("debug_out", ctypes.c_void_p),
        ("workspace", ctypes.c_void_p),
        ("workspace_bytes", ctypes.c_int64),
    ]


# VR_ABI_VERSION of include/vr_api.h
ABI_VERSION = 3
# vr_launch.workspace of a capture: the persistent NEE kernel's queue counters
CAPTURE_WORKSPACE_BYTES = 1024


class VrBvhNode(ctypes.Structure):
    _fields_ = [("bmin", F3), ("first", ctypes.c_int32), ("bmax", F3), ("count", ctypes.c_int32)]


class VrPatch(ctypes.Structure):
    _fields_ = [
        ("sampler", ctypes.c_int32),
        ("refraction_index", ctypes.c_float),
        ("emission", ctypes.c_int32),
        ("emission_c", F3),
        ("inside_medium", ctypes.c_int32),
        ("outside_medium", ctypes.c_int32),
    ]


class VrScene(ctypes.Structure):
    _fields_ = [
        ("nodes", ctypes.c_void_p),
        ("tri_verts", ctypes.c_void_p),
        ("tri_ids", ctypes.c_void_p),
        ("num_nodes", ctypes.c_int32),
        ("num_tris", ctypes.c_int32),
        ("positions", ctypes.c_void_p),
        ("normals", ctypes.c_void_p),
        ("uvs", ctypes.c_void_p),
        ("indices", ctypes.c_void_p),
        ("tri_patch", ctypes.c_void_p),
        ("patches", ctypes.c_void_p),
        ("num_patches", ctypes.c_int32),
        ("num_vertices", ctypes.c_int32),
        ("media", ctypes.c_void_p),
        ("num_media", ctypes.c_int32),
        ("env_kind", ctypes.c_int32),
        ("env_c", F3),
        ("env_image_bmin", ctypes.c_float * 2),
        ("env_image_bmax", ctypes.c_float * 2),
        ("env_image", VrTensor),
    ]


P = ctypes.POINTER
_SIGNATURES = {
    "vr_abi_version": (ctypes.c_int, []),
    "vr_last_error": (ctypes.c_char_p, []),
    "vr_grid_fw": (ctypes.c_int, [P(VrGrid), ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p]),
    "vr_grid_bw": (ctypes.c_int, [P(VrGrid), ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p]),
    "vr_majorant_build": (ctypes.c_int, [P(VrGrid), P(ctypes.c_int32), ctypes.c_void_p, ctypes.c_void_p]),
    "vr_map_fw": (ctypes.c_int, [P(VrMedium), ctypes.c_int32, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, P(VrLaunch), ctypes.c_void_p]),
    "vr_map_bw": (ctypes.c_int, [P(VrMedium), ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, P(VrLaunch), ctypes.c_void_p]),
    "vr_capture_fw": (ctypes.c_int, [P(VrMedium), ctypes.c_int32, P(VrCamera), P(VrLaunch), ctypes.c_void_p, ctypes.c_void_p]),
    "vr_capture_fw_noalloc": (ctypes.c_int, [P(VrMedium), ctypes.c_int32, P(VrCamera), P(VrLaunch), ctypes.c_void_p,
                                             ctypes.c_void_p]),
    "vr_capture_bw": (ctypes.c_int, [P(VrMedium), ctypes.c_int32, P(VrCamera), P(VrLaunch), ctypes.c_void_p, ctypes.c_void_p]),
    "vr_adam_step": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                    ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_int32,
                                    ctypes.c_int32, ctypes.c_void_p]),
    "vr_adam_step_bounded": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                            ctypes.c_int64, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                            ctypes.c_int32, ctypes.c_float, ctypes.c_float, ctypes.c_int32,
                                            ctypes.c_void_p]),
    "vr_random_ids": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, P(ctypes.c_int64), ctypes.c_int32, P(ctypes.c_uint32),
                                     ctypes.c_void_p]),
    "vr_total_variation_fw": (ctypes.c_int, [ctypes.c_void_p, P(ctypes.c_int64), ctypes.c_int32, ctypes.c_void_p,
                                             ctypes.c_void_p]),
    "vr_total_variation_bw": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, P(ctypes.c_int64), ctypes.c_int32,
                                             ctypes.c_void_p, ctypes.c_void_p]),
    "vr_resample_grid": (ctypes.c_int, [ctypes.c_void_p, P(ctypes.c_int64), P(ctypes.c_int64), ctypes.c_int32, ctypes.c_int32,
                                        ctypes.c_void_p, ctypes.c_void_p]),
    "vr_resample_img": (ctypes.c_int, [ctypes.c_void_p, P(ctypes.c_int64), P(ctypes.c_int64), ctypes.c_int32, ctypes.c_int32,
                                       ctypes.c_void_p, ctypes.c_void_p]),
    "vr_copy_img_to_morton": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32, ctypes.c_void_p,
                                             ctypes.c_void_p]),
    "vr_copy_morton_to_img": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32, ctypes.c_void_p,
                                             ctypes.c_void_p]),
    "vr_nrc_encode": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, P(ctypes.c_float), P(ctypes.c_float),
                                     ctypes.c_void_p, ctypes.c_void_p]),
    "vr_nrc_mlp_fw": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "vr_nrc_train_step": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, P(ctypes.c_float),
                                         P(ctypes.c_float),
                                         ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.c_int32, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                         ctypes.c_void_p, ctypes.c_void_p]),
    "vr_nrc_paths": (ctypes.c_int, [P(VrMedium), P(VrCamera), P(VrLaunch), ctypes.c_int32, ctypes.c_int32, ctypes.c_void_p,
                                    ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p]),
    "vr_nrc_paths_traces": (ctypes.c_int, [P(VrMedium), P(VrCamera), P(VrLaunch), ctypes.c_int32, ctypes.c_int32,
                                           ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p,
                                           ctypes.c_void_p]),
    "vr_nrc_render": (ctypes.c_int, [P(VrMedium), P(VrCamera), P(VrLaunch), ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p,
                                     ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p]),
    "vr_grid_pack_cells": (ctypes.c_int, [P(VrGrid), ctypes.c_void_p, ctypes.c_void_p]),
    "vr_scene_build_bvh": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p,
                                          ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, P(ctypes.c_int32)]),
    "vr_scene_raycast": (ctypes.c_int, [P(VrScene), ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p]),
    "vr_scene_map_fw": (ctypes.c_int, [P(VrScene), ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, P(VrLaunch),
                                       ctypes.c_void_p]),
    "vr_scene_map_bw": (ctypes.c_int, [P(VrScene), ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, P(VrLaunch),
                                       ctypes.c_void_p]),
    "vr_scene_capture_fw": (ctypes.c_int, [P(VrScene), P(VrCamera), P(VrLaunch), ctypes.c_void_p, ctypes.c_void_p]),
    "vr_scene_capture_bw": (ctypes.c_int, [P(VrScene), P(VrCamera), P(VrLaunch), ctypes.c_void_p, ctypes.c_void_p]),
}

def i64s(values) -> ctypes.Array:
    """A host int64 array for the shape arguments."""
    values = [int(v) for v in values]
    return (ctypes.c_int64 * max(1, len(values)))(*values)


def u32s(values) -> ctypes.Array:
    values = [int(v) & 0xFFFFFFFF for v in values]
    return (ctypes.c_uint32 * len(values))(*values)


NRC_PARAMS = 19456
NRC_GRAD_COPIES = 16  # weight-gradient copies of a training step (VR_NRC_GRAD_COPIES)
NRC_LOSS_SLOTS = 64  # floats of a training step's loss buffer (VR_NRC_LOSS_SLOTS)
NRC_FEATURES = 96

# Symbols declared in include/vr_api.h (checked by the CPU test suite).
EXPORTED = tuple(_SIGNATURES)

_lock = threading.Lock()
_lib = None


class VrError(RuntimeError):
    """A non-zero status returned by the native library."""


def bind(lib: ctypes.CDLL, prefix_from: str = "vr_", prefix_to: str = "vr_") -> ctypes.CDLL:
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(lib, prefix_to + name[len(prefix_from):])
        fn.restype = res
        fn.argtypes = args
    return lib


def load() -> ctypes.CDLL:
    """Load the HIP library (built by rendervous_amd.build); raises if absent."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise ImportError(
                    f"{LIB_PATH} is missing: build it with `python -m rendervous_amd.build` "
                    "(there is no CPU fallback for the HIP path)")
            lib = ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_LOCAL)
            bind(lib)
            if lib.vr_abi_version() != ABI_VERSION:
                raise ImportError("librendervous_amd.so ABI version mismatch")
            _lib = lib
        return _lib


def check(status: int, what: str) -> None:
    if status != VR_OK:
        msg = load().vr_last_error()
        raise VrError(f"{what} failed with status {status}: {msg.decode() if msg else ''}")


def stream_handle(device: torch.device) -> int:
    return torch.cuda.current_stream(device).cuda_stream


def ptr(t: torch.Tensor | None) -> int | None:
    return None if t is None else t.data_ptr()


def f3(v) -> "ctypes.Array":
    a = [float(x) for x in v]
    assert len(a) == 3
    return F3(*a)


def tensor_view(t: torch.Tensor, grad: torch.Tensor | None = None) -> VrTensor:
    """A vr_tensor over a float32 tensor of up to 4 dims (Parameter, common.h:88-94)."""
    if t.dtype != torch.float32:
        raise TypeError("vr tensors must be float32")
    if t.dim() > 4:
        raise ValueError("vr tensors have at most 4 dims")
    v = VrTensor()
    v.data = t.data_ptr()
    v.grad = grad.data_ptr() if grad is not None else None
    v.ndim = t.dim()
    for i in range(4):
        v.shape[i] = t.shape[i] if i < t.dim() else 1
        v.stride[i] = t.stride(i) if i < t.dim() else 1
    if grad is not None and (grad.shape != t.shape or grad.stride() != t.stride()):
        raise ValueError("gradient buffer must share the parameter's shape and strides")
    return v


def grid_view(t: torch.Tensor, bmin, bmax, grad: torch.Tensor | None = None) -> VrGrid:
    g = VrGrid()
    g.t = tensor_view(t, grad)
    g.bmin = f3(bmin)
    g.bmax = f3(bmax)
    return g


def launch(seeds, samples: int = 1, bw_mode: int = VR_BW_REFERENCE, seed_mode: int = VR_SEED_CAPTURE_FW,
           start_index: int = 0, count: int = -1, debug_out: torch.Tensor | None = None,
           debug_kind: int = 0, workspace: torch.Tensor | None = None) -> VrLaunch:
    L = VrLaunch()
    for i in range(4):
        L.seeds[i] = int(seeds[i]) & 0xFFFFFFFF
    L.samples = int(samples)
    L.bw_mode = int(bw_mode)
    L.seed_mode = int(seed_mode)
    L.start_index = int(start_index)
    L.count = int(count)
    L.debug_out = None if debug_out is None else debug_out.data_ptr()
    L.debug_kind = int(debug_kind)
    if workspace is not None:
        assert workspace.is_contiguous() and workspace.numel() * workspace.element_size() >= CAPTURE_WORKSPACE_BYTES
        L.workspace = workspace.data_ptr()
        L.workspace_bytes = workspace.numel() * workspace.element_size()
    return L


def capture_workspace(device) -> torch.Tensor:
    """A capture's workspace from torch's stream-ordered allocator (freed when
    the caller drops it; later work on the same stream reuses it only after
    the capture)."""
    return torch.empty(CAPTURE_WORKSPACE_BYTES // 8, dtype=torch.int64, device=device)
