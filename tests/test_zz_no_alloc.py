"""vr_capture_fw_noalloc (the capture entry point that allocates nothing: its
persistent-kernel work counters live in the library's static device array)
against vr_capture_fw, through the C ABI: two calls in a row, the second on a
smaller input, each equal to the allocating entry point's result."""
import numpy as np
import pytest
import torch

from helpers import L, Built, Camera, Scene, fbm_grid, pose

pytestmark = pytest.mark.gpu


def _scene():
    return Scene(sigma=fbm_grid(32, seed=8, scale=25.0), albedo_c=(0.9, 0.9, 0.9), phase_kind=L.VR_PHASE_HG,
                 phase_g=0.6, env_c=(0.1, 0.1, 0.1), sun_dir=(0.4, 0.8, 0.45), sun_radiance=(3.0, 3.0, 3.0),
                 sun_prob=0.5, majorant="grid", majorant_res=(8, 8, 8), max_bounces=16)


@pytest.mark.parametrize("kind,spp", [(L.VR_MAP_NEE_DT, 1), (L.VR_MAP_NEE_DT, 6), (L.VR_MAP_SS_RADIANCE, 3)])
def test_noalloc_capture_equals_the_allocating_one(gpu, kind, spp):
    print(f"[case] noalloc kind={kind} spp={spp}", flush=True)
    b = Built(_scene(), "cuda")
    lib = L.load()
    stream = torch.cuda.current_stream().cuda_stream
    for k, (w, h) in enumerate([(96, 80), (40, 24)]):  # the second call on a smaller image
        cam = Camera(np.stack([pose((0.0, 0.3, 2.5)), pose((1.6, 0.5, -1.4))]), w, h, jittered=True, device="cuda")
        seeds = (1000 + k, 2000, 3000, 4000)
        ref = torch.full((cam.threads, 3), -1.0, device="cuda")
        out = torch.full((cam.threads, 3), -1.0, device="cuda")
        L.check(lib.vr_capture_fw(b.medium, kind, cam.c, L.launch(seeds, samples=spp), ref.data_ptr(), stream),
                "vr_capture_fw")
        L.check(lib.vr_capture_fw_noalloc(b.medium, kind, cam.c, L.launch(seeds, samples=spp), out.data_ptr(), stream),
                "vr_capture_fw_noalloc")
        torch.cuda.synchronize()
        a, r = out.cpu().numpy(), ref.cpu().numpy()
        assert np.isfinite(a).all() and (a >= 0).all()
        if spp == 1 or kind != L.VR_MAP_NEE_DT:
            # one work item per pixel (or the generic kernel's fixed split): bit for bit
            np.testing.assert_array_equal(a, r)
        else:
            # the persistent kernel adds its per-item partial sums atomically, in no fixed order
            np.testing.assert_allclose(a, r, rtol=1e-5, atol=1e-7)
