"""The headline frame against the oracle, sample-matched.

bench.py's default workload, built by bench.build_scene: a 256^3 fBm density
(x40, inside a sphere), albedo 0.9, single HG g = 0.6, sun + constant
environment 0.1, 16^3 majorant grid, 512x512 at 64 spp, max 16 bounces.  The
density needs no gradient, so the map layer gives it a cell table and the
capture runs the lean persistent NEE kernel (cell table, constant albedo,
constant environment, DT shadow rays).  The oracle renders the whole frame with
the same dispatch seeds; both draw the same HybridTaus streams, so pixels agree
unless a path's float comparison lands within rounding distance of the FP64
one.
"""
import math

import numpy as np
import pytest
import torch

from helpers import L, Scene, match_fraction, oracle_capture_fw


@pytest.mark.gpu
def test_headline_frame_matches_the_oracle(gpu):
    import bench
    import rendervous_amd as rv

    res, W, H, spp = 256, 512, 512, 64
    sigma, radiance, camera = bench.build_scene(res, W, H, gpu)
    rv.seed(2024)
    seeds = rv._internal.get_seeds()
    rv.seed(2024)
    with torch.no_grad():
        img = camera.capture(radiance, fw_samples=spp)
    torch.cuda.synchronize()
    kind, lw = radiance._lower_map()
    m = lw.medium
    # the lean kernel's conditions: cell table, constant albedo and environment, DT shadow rays
    assert kind == L.VR_MAP_NEE_DT
    assert m.sigma_cells and m.albedo_kind == L.VR_FIELD_CONST and m.env_kind == L.VR_ENV_CONST
    assert m.transmittance_kind == L.VR_T_DT and m.majorant_kind == L.VR_MAJ_GRID
    assert tuple(m.majorant_res) == (16, 16, 16) and m.max_bounces == 16
    sc = Scene(sigma=sigma.cpu().numpy(), albedo_c=tuple(m.albedo_c), env_c=tuple(m.env_c), sun_dir=tuple(m.sun_dir),
               sun_radiance=tuple(m.sun_radiance), sun_prob=m.sun_prob, phase_kind=m.phase_kind, phase_g=m.phase_g,
               majorant="grid", majorant_res=(16, 16, 16), transmittance_kind=m.transmittance_kind,
               max_bounces=m.max_bounces)
    assert sc.phase_kind == L.VR_PHASE_HG and abs(sc.phase_g - 0.6) < 1e-7 and sc.sun_prob == 0.5
    print(f"headline frame: oracle over {W}x{H}x{spp} with seeds {seeds}", flush=True)
    ref = oracle_capture_fw(sc, L.VR_MAP_NEE_DT, dict(poses=camera.poses.cpu().numpy(), width=W, height=H,
                                                      jittered=True), seeds, spp)
    out = img.reshape(-1, 3).cpu().numpy()
    assert np.isfinite(out).all()
    frac = match_fraction(out, ref, rtol=1e-3, atol=1e-5)
    d = (out.astype(np.float64) - ref).ravel()
    z = d.mean() / max(d.std() / math.sqrt(d.size), 1e-30)
    print(f"headline frame: {frac:.5f} of pixels within 1e-3, mean {out.mean():.6f} vs {ref.mean():.6f}, "
          f"z of the mean difference {z:.2f}", flush=True)
    assert frac >= 0.98, frac
    assert abs(z) < 4.0, z
    assert abs(out.mean() - ref.mean()) <= 1e-3 * ref.mean()
