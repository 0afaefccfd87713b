"""ctypes binding of the CPU oracle (oracle/build/liboracle.so).

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker of the HIP library.  The oracle
exports the functions of include/vr_api.h with the prefix ``oracle_``, on host
buffers, plus RNG access for known-answer tests.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "build", "liboracle.so")

_lib = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB


def load():
    """Load (building first if needed) the oracle library with its signatures bound."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        import sys
        sys.path.insert(0, os.path.dirname(HERE))
        from rendervous_amd import _lib as L  # structures only; loads no device code

        lib = ctypes.CDLL(LIB)
        P = ctypes.POINTER
        sig = {
            "oracle_abi_version": (ctypes.c_int, []),
            "oracle_last_error": (ctypes.c_char_p, []),
            "oracle_grid_fw": (ctypes.c_int, [P(L.VrGrid), ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p]),
            "oracle_grid_fw_f64": (ctypes.c_int, [P(L.VrGrid), ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p]),
            "oracle_grid_bw": (ctypes.c_int, [P(L.VrGrid), ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64]),
            "oracle_majorant_build": (ctypes.c_int, [P(L.VrGrid), P(ctypes.c_int32), ctypes.c_void_p]),
            "oracle_map_fw": (ctypes.c_int, [P(L.VrMedium), ctypes.c_int32, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p,
                                             P(L.VrLaunch)]),
            "oracle_map_bw": (ctypes.c_int, [P(L.VrMedium), ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                             P(L.VrLaunch)]),
            "oracle_capture_fw": (ctypes.c_int, [P(L.VrMedium), ctypes.c_int32, P(L.VrCamera), P(L.VrLaunch), ctypes.c_void_p]),
            "oracle_capture_bw": (ctypes.c_int, [P(L.VrMedium), ctypes.c_int32, P(L.VrCamera), P(L.VrLaunch), ctypes.c_void_p]),
            "oracle_adam_step": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                                ctypes.c_int64, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                                ctypes.c_float, ctypes.c_int32, ctypes.c_int32]),
            "oracle_adam_step_bounded": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                                        ctypes.c_void_p, ctypes.c_int64, ctypes.c_float, ctypes.c_float,
                                                        ctypes.c_float, ctypes.c_float, ctypes.c_int32, ctypes.c_float,
                                                        ctypes.c_float, ctypes.c_int32]),
            "oracle_random_ids": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int32,
                                                 ctypes.c_void_p]),
            "oracle_total_variation_fw": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p]),
            "oracle_total_variation_bw": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32,
                                                         ctypes.c_void_p]),
            "oracle_resample_grid": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32,
                                                    ctypes.c_int32, ctypes.c_void_p]),
            "oracle_resample_img": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32,
                                                   ctypes.c_int32, ctypes.c_void_p]),
            "oracle_copy_img_to_morton": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32, ctypes.c_void_p]),
            "oracle_copy_morton_to_img": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32, ctypes.c_void_p]),
            "oracle_env_sample": (ctypes.c_int, [P(L.VrMedium), ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p]),
            "oracle_phase_sample": (ctypes.c_int, [P(L.VrMedium), ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                                   ctypes.c_void_p]),
            "oracle_phase_eval": (ctypes.c_int, [P(L.VrMedium), ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                                 ctypes.c_void_p]),
            "oracle_camera_rays": (ctypes.c_int, [P(L.VrCamera), ctypes.c_void_p, ctypes.c_void_p]),
            "oracle_grid_pack_cells": (ctypes.c_int, [P(L.VrGrid), ctypes.c_void_p]),
            "oracle_scene_build_bvh": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int64,
                                                      ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p,
                                                      P(ctypes.c_int32)]),
            "oracle_scene_raycast": (ctypes.c_int, [P(L.VrScene), ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p]),
            "oracle_scene_map_fw": (ctypes.c_int, [P(L.VrScene), ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p,
                                                   P(L.VrLaunch)]),
            "oracle_scene_map_bw": (ctypes.c_int, [P(L.VrScene), ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                                   P(L.VrLaunch)]),
            "oracle_scene_capture_fw": (ctypes.c_int, [P(L.VrScene), P(L.VrCamera), P(L.VrLaunch), ctypes.c_void_p]),
            "oracle_scene_capture_bw": (ctypes.c_int, [P(L.VrScene), P(L.VrCamera), P(L.VrLaunch), ctypes.c_void_p]),
            "oracle_rng_uniforms": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p]),
            "oracle_rng_seed_states": (ctypes.c_int, [ctypes.c_int32, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64,
                                                      ctypes.c_void_p]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def check(status: int, what: str):
    if status != 0:
        raise RuntimeError(f"{what}: oracle status {status}: {load().oracle_last_error().decode()}")


def nptr(a: np.ndarray) -> int:
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


def rng_uniforms(state, n: int):
    st = np.asarray(state, dtype=np.uint32)
    out = np.zeros(n, dtype=np.float32)
    fin = np.zeros(4, dtype=np.uint32)
    check(load().oracle_rng_uniforms(nptr(st), n, nptr(out), nptr(fin)), "rng_uniforms")
    return out, fin


def seed_states(mode: int, seeds, first: int, n: int) -> np.ndarray:
    sd = np.asarray(seeds, dtype=np.uint32)
    out = np.zeros((n, 4), dtype=np.uint32)
    check(load().oracle_rng_seed_states(mode, nptr(sd), first, n, nptr(out)), "seed_states")
    return out
