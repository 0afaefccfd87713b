"""Robustness of the HIP path: non-finite densities rejected by the majorant
build, caches of a density (cell table, majorant grid) rebuilt after in-place
updates, and captures on concurrent streams that keep their own work
counters."""
import numpy as np
import pytest
import torch

import rendervous_amd as rv
from helpers import L, O, Built, Camera, Scene, fbm_grid, np_grid, pose

pytestmark = pytest.mark.gpu


def _stream():
    return torch.cuda.current_stream().cuda_stream


@pytest.mark.parametrize("bad", [float("nan"), float("inf"), -float("inf")])
def test_majorant_build_marks_non_finite_densities(gpu, bad):
    """An infinite majorant would give zero-length free flights: coarse cells
    whose footprint holds a NaN or infinite density get a NaN bound (which the
    trackers floor, crossing the cell as empty), the same cells as the oracle's;
    every other cell is the oracle's bound.  Only the build runs (it reads the
    grid); no tracker is launched on such a density."""
    print(f"[case] majorant_build bad={bad}", flush=True)
    g = fbm_grid(20, seed=4, scale=10.0)
    g[7, 3, 11, 0] = bad
    lib = L.load()
    res = (L.ctypes.c_int32 * 3)(4, 5, 6)
    out = torch.zeros(4, 5, 6, device="cuda")
    t = torch.from_numpy(g).cuda()
    L.check(lib.vr_majorant_build(L.grid_view(t, [-0.5] * 3, [0.5] * 3), res, out.data_ptr(), _stream()), "majorant")
    ref = np.zeros((4, 5, 6), np.float32)
    O.check(O.load().oracle_majorant_build(np_grid(g), res, O.nptr(ref)), "oracle majorant")
    got = out.cpu().numpy()
    assert np.isnan(ref).any() and (~np.isnan(ref)).any()
    np.testing.assert_array_equal(np.isnan(got), np.isnan(ref))
    np.testing.assert_array_equal(got[~np.isnan(ref)], ref[~np.isnan(ref)])


def _nee_field(sigma: torch.Tensor):
    ds = rv.objects.DependencySet()
    ds.add_parameters(sigma_tensor=sigma, scattering_albedo_tensor=torch.tensor([0.9, 0.9, 0.9]),
                      environment_tensor=torch.tensor([0.2, 0.2, 0.2]), phase_g_tensor=torch.tensor([0.5]))
    ds.requires(rv.objects.medium_majorant_coarse_grid, resolution=8)
    ds.requires(rv.objects.medium_environment_sampler_sun, direction=(0.3, 0.9, 0.3), radiance=(2.0, 2.0, 2.0),
                probability=0.5)
    ds.requires(rv.objects.medium_phase_HG, mixture=False)
    ds.requires(rv.objects.medium_phase_sampler_HG, mixture=False)
    ds.requires(rv.objects.medium_transmittance_DT)
    ds.requires(rv.objects.medium_radiance_path_integrator_NEE_DT, max_bounces=8)
    return ds.radiance


def _capture(field, camera, seed, cell_tables="auto"):
    old = rv._internal.settings.cell_tables
    rv._internal.settings.cell_tables = cell_tables
    try:
        rv.seed(seed)
        with torch.no_grad():
            img = camera.capture(field, fw_samples=1)
        torch.cuda.synchronize()
        return img.cpu().numpy()
    finally:
        rv._internal.settings.cell_tables = old


def test_in_place_density_updates_rebuild_the_cell_table_and_majorant(gpu):
    """A density updated in place under no_grad (version bump) or through
    .data followed by increment_version: the next capture rebuilds the cell
    table and the majorant grid, so it equals (bit for bit, one sample per
    pixel) a capture that reads the grid directly with a fresh majorant."""
    print("[case] in_place_update", flush=True)
    sigma = torch.from_numpy(fbm_grid(32, seed=6, scale=20.0)).cuda()
    field = _nee_field(sigma)
    camera = rv.PerspectiveCameraSensor(32, 32, rv.functions.look_at_pose((0.0, 0.3, 2.5))[None].cuda(), jittered=True)
    kind, lw = field._lower_map()
    grid = lw.sigma.grid
    assert lw.medium.sigma_cells  # the density needs no gradient: it reads a cell table
    before = _capture(field, camera, 11)
    updates = [lambda: grid.mul_(0.35).add_(1.5),
               lambda: (grid.data.mul_(2.0), torch.autograd.graph.increment_version(grid))]
    for k, update in enumerate(updates):
        with torch.no_grad():
            update()
        with_table = _capture(field, camera, 11)
        direct = _capture(field, camera, 11, cell_tables=False)
        np.testing.assert_array_equal(with_table, direct)
        assert not np.array_equal(with_table, before), k
        # the majorant grid the medium uses equals a fresh build of the updated density
        _, lw2 = field._lower_map()
        res = (L.ctypes.c_int32 * 3)(8, 8, 8)
        fresh = torch.zeros(8, 8, 8, device="cuda")
        L.check(L.load().vr_majorant_build(lw2.medium.sigma, res, fresh.data_ptr(), _stream()), "majorant")
        torch.cuda.synchronize()
        used = [t for t in lw2.keep if t.data_ptr() == lw2.medium.majorant_grid]
        assert used and all(torch.equal(u, fresh) for u in used)
        before = with_table


@pytest.mark.parametrize("own_workspace", [True, False])
def test_captures_on_two_streams_keep_their_own_work_counters(gpu, own_workspace):
    """Two persistent NEE captures launched on two streams without a sync in
    between: each keeps its own queue counters (the caller's workspace, or one
    the library allocates on the stream), so both images equal the serial ones
    (up to the order of the per-item float atomics)."""
    print(f"[case] two_streams own_workspace={own_workspace}", flush=True)
    sc = Scene(sigma=fbm_grid(32, seed=2, scale=25.0), albedo_c=(0.9, 0.9, 0.9), phase_kind=L.VR_PHASE_HG, phase_g=0.6,
               env_c=(0.1, 0.1, 0.1), sun_dir=(0.4, 0.8, 0.45), sun_radiance=(3.0, 3.0, 3.0), sun_prob=0.5,
               majorant="grid", majorant_res=(8, 8, 8), max_bounces=16)
    b = Built(sc, "cuda")
    cams = [Camera(pose((0.0, 0.3, 2.5))[None], 96, 80, jittered=True, device="cuda"),
            Camera(pose((1.8, 0.4, -1.2))[None], 64, 72, jittered=True, device="cuda")]
    lib = L.load()

    def run(k, stream, out, ws):
        lau = L.launch((100 + k, 200, 300, 400), samples=8, workspace=ws)
        L.check(lib.vr_capture_fw(b.medium, L.VR_MAP_NEE_DT, cams[k].c, lau, out.data_ptr(), stream.cuda_stream),
                "vr_capture_fw")

    main = torch.cuda.current_stream()
    serial = []
    for k in range(2):
        out = torch.zeros(cams[k].threads, 3, device="cuda")
        run(k, main, out, L.capture_workspace("cuda") if own_workspace else None)
        torch.cuda.synchronize()
        serial.append(out.cpu().numpy())
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    outs = [torch.zeros(cams[k].threads, 3, device="cuda") for k in range(2)]
    wss = [L.capture_workspace("cuda") if own_workspace else None for _ in range(2)]
    torch.cuda.synchronize()
    for rep in range(3):
        for k in range(2):
            with torch.cuda.stream(streams[k]):
                outs[k].zero_()
                run(k, streams[k], outs[k], wss[k])
        torch.cuda.synchronize()
        for k in range(2):
            np.testing.assert_allclose(outs[k].cpu().numpy(), serial[k], rtol=1e-5, atol=1e-7)
