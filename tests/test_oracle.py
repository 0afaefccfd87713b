"""Pins of the CPU oracle (no GPU): textbook values, closed forms, invariants,
brute force and finite differences, chosen so that a dropped term, a wrong
sign/index or a transposed operand in the restatement fails one of them."""
import json
import math
import os

import numpy as np
import pytest
import torch

from helpers import (L, O, Scene, blob_grid, fbm_grid, np_grid, oracle_capture_bw, oracle_capture_fw, oracle_map_bw, oracle_map_fw, pose,
                     random_rays)

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


# ---------------------------------------------------------------- RNG --
def test_lcg_component_matches_textbook_sequence():
    """The 4th state word is the Numerical Recipes quick generator
    x' = 1664525 x + 1013904223 (mod 2^32); from 0 it yields 1013904223,
    1196435762, 3519870697, 2868466484 (Press et al., Numerical Recipes ch. 7.1)."""
    expect = [1013904223, 1196435762, 3519870697, 2868466484]
    for n, e in enumerate(expect, start=1):
        _, fin = O.rng_uniforms([129, 129, 129, 0], n)
        assert int(fin[3]) == e


def test_tausworthe_components_are_gf2_linear_and_masked():
    """Each Tausworthe step is linear over GF(2) (shifts, masks, xors) and ignores
    the masked low bits (1, 3, 4 bits for the three components)."""
    rng = np.random.default_rng(0)
    for _ in range(20):
        a = rng.integers(1 << 20, 1 << 32, size=4, dtype=np.uint64).astype(np.uint32)
        b = rng.integers(1 << 20, 1 << 32, size=4, dtype=np.uint64).astype(np.uint32)
        _, fa = O.rng_uniforms(a, 1)
        _, fb = O.rng_uniforms(b, 1)
        _, fab = O.rng_uniforms(a ^ b, 1)
        assert np.array_equal(fab[:3], fa[:3] ^ fb[:3])
        lowbits = np.array([1, 7, 15, 0], dtype=np.uint32)
        _, fl = O.rng_uniforms(a ^ lowbits, 1)
        assert np.array_equal(fl[:3], fa[:3])


def test_uniforms_are_23_bit_and_uniform():
    u, _ = O.rng_uniforms([12345, 67890, 13579, 24680], 1 << 20)
    assert u.min() >= 0.0 and u.max() < 1.0
    assert np.all(np.floor(u.astype(np.float64) * 2 ** 23) == u.astype(np.float64) * 2 ** 23)
    hist, _ = np.histogram(u, bins=256, range=(0, 1))
    e = len(u) / 256
    chi2 = ((hist - e) ** 2 / e).sum()
    assert chi2 < 255 + 5 * math.sqrt(2 * 255)
    assert abs(np.corrcoef(u[:-1], u[1:])[0, 1]) < 0.005


def test_rng_golden_vectors():
    with open(os.path.join(GOLDEN, "rng.json")) as f:
        gold = json.load(f)
    for case in gold["streams"]:
        u, _ = O.rng_uniforms(case["state"], len(case["uniforms"]))
        assert u.tolist() == case["uniforms"]
    for case in gold["seed_states"]:
        st = O.seed_states(case["mode"], case["seeds"], case["first"], len(case["states"]))
        assert st.tolist() == case["states"]


# ------------------------------------------------------------ ray-box --
def test_ray_box_slab_cases():
    sc = Scene(sigma=np.zeros((2, 2, 2, 1), np.float32), box_min=(-1, -2, -3), box_max=(1, 2, 3))
    rays = np.array([
        [0, 0, -10, 0, 0, 1],     # along z: enters at 7, leaves at 13
        [5, 0, 0, -1, 0, 0],      # along -x: 4 .. 6
        [0, 0, 0, 0, 1, 0],       # from inside along y: -2 .. 2
        [0, 5, 0, 1, 0, 0],       # parallel outside the slab -> empty interval
    ], np.float32)
    out = oracle_map_fw(sc, L.VR_MAP_RAYBOX, rays, (1, 2, 3, 4))
    np.testing.assert_allclose(out[0], [7, 13], rtol=1e-6)
    np.testing.assert_allclose(out[1], [4, 6], rtol=1e-6)
    np.testing.assert_allclose(out[2], [-2, 2], rtol=1e-6)
    assert out[3, 1] < out[3, 0]


# ------------------------------------------------------------- Grid3D --
def _lookup(g, x, bmin=(-0.5,) * 3, bmax=(0.5,) * 3):
    out = np.zeros((x.shape[0], g.shape[3]), np.float64)
    O.check(O.load().oracle_grid_fw_f64(np_grid(g, bmin, bmax), O.nptr(np.ascontiguousarray(x, np.float32)), x.shape[0],
                                        O.nptr(out)), "grid")
    return out


def test_grid_reproduces_affine_fields():
    n = (5, 6, 7)
    zz, yy, xx = np.meshgrid(*[np.linspace(-0.5, 0.5, k) for k in n], indexing="ij")
    g = (1.5 + 2.0 * xx - 3.0 * yy + 0.5 * zz)[..., None].astype(np.float32)
    x = np.random.default_rng(0).uniform(-0.5, 0.4999, size=(1000, 3)).astype(np.float32)
    out = _lookup(g, x)[:, 0]
    ref = 1.5 + 2.0 * x[:, 0] - 3.0 * x[:, 1] + 0.5 * x[:, 2]
    np.testing.assert_allclose(out, ref, atol=1e-5)


def test_grid_matches_torch_grid_sample_and_zero_outside():
    rng = np.random.default_rng(1)
    g = rng.uniform(size=(4, 5, 6, 2)).astype(np.float32)
    x = rng.uniform(-0.49, 0.49, size=(500, 3)).astype(np.float32)
    out = _lookup(g, x)
    t = torch.nn.functional.grid_sample(torch.from_numpy(g).permute(3, 0, 1, 2)[None].double(),
                                        torch.from_numpy(2 * x).double().reshape(1, 500, 1, 1, 3),
                                        mode="bilinear", padding_mode="zeros", align_corners=True)
    ref = t.reshape(2, 500).T.numpy()
    np.testing.assert_allclose(out, ref, atol=1e-6)
    outside = np.array([[0.6, 0, 0], [0, -0.51, 0], [0, 0, 0.5]], np.float32)  # 0.5 maps to ncoord 1 -> outside
    assert np.all(_lookup(g, outside) == 0.0)


def test_grid_scatter_is_the_adjoint_of_lookup():
    rng = np.random.default_rng(2)
    g = rng.uniform(size=(5, 4, 3, 1)).astype(np.float32)
    x = rng.uniform(-0.5, 0.5, size=(300, 3)).astype(np.float32)
    og = rng.normal(size=(300, 1)).astype(np.float32)
    v = rng.normal(size=g.shape).astype(np.float32)
    grad = np.zeros_like(g)
    O.check(O.load().oracle_grid_bw(np_grid(g, grad=grad), O.nptr(x), O.nptr(og), 300), "grid_bw")
    lhs = float((grad.astype(np.float64) * v).sum())
    rhs = float((og[:, 0] * _lookup(v, x)[:, 0]).sum())
    assert abs(lhs - rhs) <= 1e-4 * max(1.0, abs(rhs))


# ---------------------------------------------------------------- DDA --
def test_dda_constant_grid_is_exact():
    sc = Scene(sigma=np.full((9, 9, 9, 1), 3.0, np.float32))
    rays = random_rays(200, seed=3)
    T = oracle_map_fw(sc, L.VR_MAP_T_DDA, rays, (1, 2, 3, 4))[:, 0]
    bx = oracle_map_fw(sc, L.VR_MAP_RAYBOX, rays, (1, 2, 3, 4))
    chord = np.clip(bx[:, 1], 0, None) - np.clip(bx[:, 0], 0, None)
    chord = np.where(bx[:, 1] > np.maximum(bx[:, 0], 0), chord, 0.0)
    # the traversal stops at d - 1e-4 (vr_transmittance_DDA.h:38): a final piece
    # shorter than 1e-4 may be dropped
    lo, hi = np.exp(-3.0 * chord), np.exp(-3.0 * np.maximum(chord - 1e-4, 0))
    assert np.all((T >= lo * (1 - 2e-6)) & (T <= hi * (1 + 2e-6)))


def test_dda_midpoint_exact_for_linear_fields_and_misses():
    n = 17
    z = np.linspace(-0.5, 0.5, n)
    g = np.broadcast_to((2.0 + 4.0 * (z + 0.5))[:, None, None, None], (n, n, n, 1)).astype(np.float32).copy()
    sc = Scene(sigma=g)
    rays = np.array([[0.1, -0.2, -3.0, 0, 0, 1], [0.0, 3.0, 0.0, 0, 1, 0], [0.2, 0.1, 0.0, 0, 0, -1]], np.float32)
    T = oracle_map_fw(sc, L.VR_MAP_T_DDA, rays, (1, 2, 3, 4))[:, 0]
    # integral of 2 + 4 s over s in [0,1] = 4; from z=0 down to -0.5: int_0^0.5 (2 + 4 s) ds = 1.5
    np.testing.assert_allclose(T, [math.exp(-4.0), 1.0, math.exp(-1.5)], rtol=1e-6)
    assert oracle_map_fw(Scene(sigma=np.zeros((4, 4, 4, 1), np.float32)), L.VR_MAP_T_DDA, rays, (1, 2, 3, 4)).min() == 1.0


def test_dda_backward_matches_central_differences():
    rng = np.random.default_rng(4)
    g = rng.uniform(0.5, 2.0, size=(5, 5, 5, 1)).astype(np.float32)
    rays = random_rays(64, seed=5)
    og = rng.normal(size=(64, 1)).astype(np.float32)
    grad, _ = oracle_map_bw(Scene(sigma=g), L.VR_MAP_T_DDA, rays, og, (1, 2, 3, 4))
    h = 1e-2
    for idx in [(1, 2, 3, 0), (2, 2, 2, 0), (4, 0, 1, 0), (0, 3, 3, 0)]:
        gp, gm = g.copy(), g.copy()
        gp[idx] += h
        gm[idx] -= h
        fp = (og * oracle_map_fw(Scene(sigma=gp), L.VR_MAP_T_DDA, rays, (1, 2, 3, 4))).sum()
        fm = (og * oracle_map_fw(Scene(sigma=gm), L.VR_MAP_T_DDA, rays, (1, 2, 3, 4))).sum()
        fd = (fp - fm) / (2 * h)
        assert abs(grad[idx] - fd) <= 2e-3 * max(1.0, abs(fd)), (idx, grad[idx], fd)


# --------------------------------------------------- stochastic trackers --
def _exact_T(g, ray, bmin=-0.5, bmax=0.5):
    """exp(-integral of the trilinear field) by fine quadrature (not the DDA midpoint rule)."""
    from scipy.integrate import quad
    o, d = ray[:3].astype(np.float64), ray[3:].astype(np.float64)
    w = np.where(np.abs(d) <= 1e-6, 1e-6, d)
    t0 = np.maximum(np.minimum((bmin - o) / w, (bmax - o) / w).max(), 0)
    t1 = np.maximum((bmin - o) / w, (bmax - o) / w).min()
    if t1 <= t0:
        return 1.0

    def f(t):
        return _lookup(g, (o + d * t)[None].astype(np.float32))[0, 0]

    val, _ = quad(f, t0, t1, limit=400, epsabs=1e-7)
    return math.exp(-val)


@pytest.mark.parametrize("majorant", ["const", "grid"])
@pytest.mark.parametrize("kind", [L.VR_MAP_T_DT, L.VR_MAP_T_RT])
def test_dt_rt_expectation_equals_exact_transmittance(kind, majorant):
    g = blob_grid(12, seed=2, max_value=6.0)
    sc = Scene(sigma=g, majorant=majorant, majorant_res=(3, 3, 3))
    ray = np.array([0.05, -0.1, -1.5, 0.1, 0.15, 1.0], np.float32)
    ray[3:] /= np.linalg.norm(ray[3:])
    n = 200000
    T = oracle_map_fw(sc, kind, np.tile(ray, (n, 1)), (11, 22, 33, 44))[:, 0]
    exact = _exact_T(g, ray)
    se = T.std() / math.sqrt(n)
    assert abs(T.mean() - exact) <= 4 * se + 1e-4, (T.mean(), exact, se)


@pytest.mark.parametrize("kind", [L.VR_MAP_T_GDT, L.VR_MAP_T_GRT])
def test_grid_tracking_constant_grid_is_exact(kind):
    """sigma constant: every cell's minorant is sigma and the residual majorant
    1e-6, so both trackers return exp(-sigma * chord) (up to the final 1e-4 piece
    the cell walk drops, vr_transmittance_GDT.h:39)."""
    sc = Scene(sigma=np.full((7, 7, 7, 1), 2.5, np.float32))
    rays = np.concatenate([random_rays(200, seed=6), random_rays(50, seed=7, inside=True)])
    T = oracle_map_fw(sc, kind, rays, (5, 6, 7, 8))[:, 0]
    bx = oracle_map_fw(sc, L.VR_MAP_RAYBOX, rays, (1, 2, 3, 4))
    chord = np.where(bx[:, 1] > np.maximum(bx[:, 0], 0), bx[:, 1] - np.clip(bx[:, 0], 0, None), 0.0)
    lo, hi = np.exp(-2.5 * chord), np.exp(-2.5 * np.maximum(chord - 1e-4, 0))
    assert np.all((T >= lo * (1 - 2e-6)) & (T <= hi * (1 + 2e-6)))


@pytest.mark.parametrize("kind", [L.VR_MAP_T_GDT, L.VR_MAP_T_GRT])
def test_grid_tracking_expectation_equals_exact_transmittance(kind):
    """E[T] = exp(-integral of the trilinear field) by adaptive quadrature; a
    closed-form part taken with the cell max, a wrong interpolation offset or a
    dropped cell would each move the mean by many standard errors."""
    g = blob_grid(10, seed=4, max_value=8.0)
    sc = Scene(sigma=g)
    ray = np.array([-0.1, 0.05, -1.5, 0.2, -0.1, 1.0], np.float32)
    ray[3:] /= np.linalg.norm(ray[3:])
    n = 100000
    T = oracle_map_fw(sc, kind, np.tile(ray, (n, 1)), (31, 32, 33, 34))[:, 0]
    exact = _exact_T(g, ray)
    se = max(T.std(), 1e-3) / math.sqrt(n)
    assert abs(T.mean() - exact) <= 4 * se + 1e-4, (T.mean(), exact, se)
    if kind == L.VR_MAP_T_GDT:
        # delta tracking returns 0 or exp(-integral of the cell minorants), one value per ray
        nz = np.unique(T[T > 0])
        assert np.ptp(nz) <= 1e-6 * nz.max()
        # ratio tracking multiplies that value by factors 1 - Pc in [0, 1] (minorant <= sigma <= majorant)
        Tr = oracle_map_fw(sc, L.VR_MAP_T_GRT, np.tile(ray, (1000, 1)), (1, 3, 5, 7))[:, 0]
        assert Tr.min() >= 0 and Tr.max() <= nz.max() * (1 + 1e-6) and Tr.mean() < nz.max() * 0.99


def test_grid_tracking_has_no_backward():
    """maps/transmittance_gdt.h and _grt.h define FORWARD only: no gradient."""
    g = blob_grid(6, seed=1, max_value=4.0)
    rays = random_rays(32, seed=2)
    for kind in (L.VR_MAP_T_GDT, L.VR_MAP_T_GRT):
        grad, _ = oracle_map_bw(Scene(sigma=g), kind, rays, np.ones((32, 1), np.float32), (1, 2, 3, 4))
        assert not grad.any()


def test_collision_sampler_homogeneous_closed_form():
    s = 2.5
    sc = Scene(sigma=np.full((4, 4, 4, 1), s, np.float32), majorant_value=4.0)
    n = 100000
    ray = np.array([0.1, 0.0, -1.0, 0, 0, 1], np.float32)
    out = oracle_map_fw(sc, L.VR_MAP_COLLISION_DT, np.tile(ray, (n, 1)), (5, 6, 7, 8))
    C, t, T = out[:, 0], out[:, 1], out[:, 2]
    p = 1 - math.exp(-s)
    assert abs(C.mean() - p) <= 4 * math.sqrt(p * (1 - p) / n)
    assert np.all((C == 1) == (T == 0))
    # t - tMin | C=1 ~ Exp(s) truncated to [0, 1]: KS test against the closed-form CDF
    tt = np.sort(t[C == 1] - 0.5)
    cdf = (1 - np.exp(-s * tt)) / (1 - math.exp(-s))
    ks = np.abs(cdf - np.arange(1, len(tt) + 1) / len(tt)).max()
    assert ks < 1.63 / math.sqrt(len(tt)) + 1e-3


# ------------------------------------------------------------------ NEE --
def test_nee_white_furnace():
    """Albedo 1, no emission, consistent HG phase, constant environment L: R == L."""
    sc = Scene(sigma=blob_grid(16, seed=3, max_value=15.0), albedo_c=(1, 1, 1), phase_kind=L.VR_PHASE_HG, phase_g=0.6,
               env_c=(0.7, 0.7, 0.7))
    rays = random_rays(20000, seed=6)
    R = oracle_map_fw(sc, L.VR_MAP_NEE_DT, rays, (3, 4, 5, 6))
    se = R[:, 0].std() / math.sqrt(len(R))
    assert abs(R[:, 0].mean() - 0.7) <= 4 * se, (R[:, 0].mean(), se)


def test_nee_absorbing_medium_is_transmittance_times_env():
    g = blob_grid(12, seed=4, max_value=8.0)
    sc = Scene(sigma=g, albedo_c=(0, 0, 0), env_c=(2.0, 1.0, 0.5))
    ray = np.array([0.0, 0.05, -1.5, 0.0, 0.0, 1.0], np.float32)
    n = 100000
    R = oracle_map_fw(sc, L.VR_MAP_NEE_DT, np.tile(ray, (n, 1)), (7, 8, 9, 10))
    exact = _exact_T(g, ray)
    se = R[:, 0].std() / math.sqrt(n)
    assert abs(R[:, 0].mean() - 2.0 * exact) <= 4 * se + 1e-4
    np.testing.assert_allclose(R[:, 1] * 2.0, R[:, 0], rtol=1e-6)


def test_majorant_grid_bounds_the_trilinear_field():
    g = fbm_grid(20, seed=5, scale=10.0)
    res = (3, 4, 5)
    maj = np.zeros(res, np.float32)
    O.check(O.load().oracle_majorant_build(np_grid(g), (L.ctypes.c_int32 * 3)(*res), O.nptr(maj)), "maj")
    x = np.random.default_rng(7).uniform(-0.5, 0.5, size=(50000, 3)).astype(np.float32)
    v = _lookup(g, x)[:, 0]
    cell = np.floor((x.astype(np.float64) + 0.5) * np.array(res[::-1])).astype(int)
    cell = np.minimum(cell, np.array(res[::-1]) - 1)
    bound = maj[cell[:, 2], cell[:, 1], cell[:, 0]]
    assert np.all(v <= bound + 1e-5)
    assert maj.max() == g.max()


def test_camera_principal_ray_follows_pose_direction():
    sc = Scene(sigma=np.zeros((2, 2, 2, 1), np.float32), box_min=(-10, -10, -10), box_max=(10, 10, 10))
    cam = dict(poses=pose((0.0, 0.0, 2.5))[None], width=2, height=2)
    # RAYBOX returns the raw slab interval; for the 4 pixels of a 2x2 image the
    # rays are symmetric about the -z principal axis.
    out = oracle_capture_fw(sc, L.VR_MAP_RAYBOX, cam, (1, 2, 3, 4), 1)
    assert np.allclose(out[:, 1], out[0, 1], rtol=1e-5)


# ----------------------------------------------------------- gradients --
def test_rt_backward_matches_finite_differences():
    """Central differences (h=1e-2) of E[T_RT] against the mean of the RT backward
    estimator (same estimator family the inverse-render configs use)."""
    rng = np.random.default_rng(8)
    g = rng.uniform(1.0, 3.0, size=(4, 4, 4, 1)).astype(np.float32)
    ray = np.array([0.02, 0.03, -1.5, 0.0, 0.0, 1.0], np.float32)
    n = 400000
    rays = np.tile(ray, (n, 1))
    og = np.ones((n, 1), np.float32) / n
    grad, _ = oracle_map_bw(Scene(sigma=g, transmittance_kind=L.VR_T_RT), L.VR_MAP_T_RT, rays, og, (1, 2, 3, 4),
                            bw_mode=L.VR_BW_CONSISTENT)
    h = 1e-2
    for idx in [(1, 1, 1, 0), (2, 2, 2, 0), (0, 1, 2, 0)]:
        gp, gm = g.copy(), g.copy()
        gp[idx] += h
        gm[idx] -= h
        fd = (math.log(_exact_T(gp, ray)) - math.log(_exact_T(gm, ray))) / (2 * h) * _exact_T(g, ray)
        assert abs(grad[idx] - fd) <= 0.05 * abs(fd) + 2e-4, (idx, grad[idx], fd)


def _capture_fd_check(kind, transmittance_kind, voxels, spp_bw, spp_fd, h=0.4, rel_tol=0.06, bw_kinds=None):
    """Mean of the oracle's capture backward (consistent estimator, replayed
    forward seeds) against a central difference of the oracle's forward, both
    sides of the difference rendered with the same seeds.  bw_kinds: several
    backward estimators {kind: (spp, rel_tol)} of the same forward."""
    rng = np.random.default_rng(1)
    g = rng.uniform(1.0, 4.0, size=(6, 6, 6, 1)).astype(np.float32)
    cam = dict(poses=np.stack([pose((0.3, 0.4, 2.0)), pose((-1.8, 0.2, 0.9))]), width=8, height=8, jittered=True)
    w = (rng.uniform(size=(128, 3)) + 0.5).astype(np.float32)

    def scene(s):  # a fixed majorant: max(sigma) would move with the perturbation
        return Scene(sigma=s, albedo_c=(0.8, 0.8, 0.8), env_c=(1.0, 0.8, 0.6), phase_g=0.3,
                     transmittance_kind=transmittance_kind, max_bounces=8, majorant_value=4.5)

    bw_kinds = bw_kinds or {kind: (spp_bw, rel_tol)}
    grads = {k: oracle_capture_bw(scene(g), k, cam, (1, 2, 3, 4), n, w, bw_mode=L.VR_BW_CONSISTENT)[0]
             for k, (n, _) in bw_kinds.items()}
    for v in voxels:
        r = []
        for sgn in (1, -1):
            s2 = g.copy()
            s2[v] += sgn * h
            r.append(float((oracle_capture_fw(scene(s2), kind, cam, (7, 8, 9, 10), spp_fd).astype(np.float64) * w).sum()))
        fd = (r[0] - r[1]) / (2 * h)
        for k, (_, tol) in bw_kinds.items():
            assert abs(grads[k][v] - fd) <= tol * abs(fd), (k, v, grads[k][v], fd)


def test_collision_single_scatter_backward_matches_finite_differences():
    """MC collision integrator over the DT collision sampler with single-scatter
    exitance and RT shadow rays (the inverse-rendering graph)."""
    _capture_fd_check(L.VR_MAP_MC_COLLISION, L.VR_T_RT, [(2, 3, 3, 0), (3, 2, 2, 0)], 16384, 65536)


def test_nee_path_backward_matches_finite_differences():
    """NEE delta-tracking path integrator, up to 8 bounces, DT shadow rays: the
    defensive replay estimator (vr_path_integrator_NEE_DS.h) and the reservoir
    estimator (vr_path_integrator_NEE_DRT.h) against one central difference."""
    _capture_fd_check(L.VR_MAP_NEE_DT, L.VR_T_DT, [(2, 3, 3, 0), (3, 2, 2, 0)], None, 262144,
                      bw_kinds={L.VR_MAP_NEE_DT: (32768, 0.08), L.VR_MAP_NEE_DRT: (65536, 0.08)})


def test_dt_path_backward_matches_finite_differences():
    """Delta-tracking path integrator (no NEE), defensive replay backward."""
    _capture_fd_check(L.VR_MAP_PATH_DT, L.VR_T_DT, [(2, 3, 3, 0), (3, 2, 2, 0)], 16384, 131072)


def test_dt_path_white_furnace_is_exact():
    """Albedo 1, HG sampling weight 1, constant environment L: every sample is L."""
    sc = Scene(sigma=blob_grid(16, seed=3, max_value=15.0), albedo_c=(1, 1, 1), phase_kind=L.VR_PHASE_HG, phase_g=0.6,
               env_c=(0.7, 0.7, 0.7))
    R = oracle_map_fw(sc, L.VR_MAP_PATH_DT, random_rays(4000, seed=6), (3, 4, 5, 6))
    np.testing.assert_allclose(R, 0.7, rtol=1e-6)


def test_dt_path_absorbing_medium_is_transmittance_times_env():
    g = blob_grid(12, seed=4, max_value=8.0)
    sc = Scene(sigma=g, albedo_c=(0, 0, 0), env_c=(2.0, 1.0, 0.5))
    ray = np.array([0.0, 0.05, -1.5, 0.0, 0.0, 1.0], np.float32)
    n = 100000
    R = oracle_map_fw(sc, L.VR_MAP_PATH_DT, np.tile(ray, (n, 1)), (7, 8, 9, 10))
    se = R[:, 0].std() / math.sqrt(n)
    assert abs(R[:, 0].mean() - 2.0 * _exact_T(g, ray)) <= 4 * se + 1e-4


def test_dt_path_and_nee_estimate_the_same_radiance():
    """Unbounded paths, constant environment, no sun: the path tracer without
    NEE and the NEE path tracer are unbiased for the same radiance."""
    sc = Scene(sigma=blob_grid(10, seed=8, max_value=10.0), albedo_c=(0.8, 0.7, 0.6), phase_kind=L.VR_PHASE_HG,
               phase_g=0.4, env_c=(1.0, 0.9, 0.8))
    ray = np.array([0.02, -0.03, -1.5, 0.0, 0.0, 1.0], np.float32)
    n = 200000
    a = oracle_map_fw(sc, L.VR_MAP_PATH_DT, np.tile(ray, (n, 1)), (1, 2, 3, 4)).astype(np.float64)
    b = oracle_map_fw(sc, L.VR_MAP_NEE_DT, np.tile(ray, (n, 1)), (5, 6, 7, 8)).astype(np.float64)
    se = np.sqrt(a.var(0) / n + b.var(0) / n)
    assert np.all(np.abs(a.mean(0) - b.mean(0)) <= 4 * se), (a.mean(0), b.mean(0), se)


# ---------------------------------------------------------------- Adam --
def test_adam_matches_torch_optim_adam():
    """Three steps of the oracle's Adam equal torch.optim.Adam (no weight decay)
    on the same gradients; the bounded form is the projection of that update."""
    rng = np.random.default_rng(26)
    n = 1000
    p0 = rng.normal(size=n).astype(np.float32)
    grads = [rng.normal(size=n).astype(np.float32) for _ in range(3)]
    tp = torch.nn.Parameter(torch.from_numpy(p0.copy()))
    opt = torch.optim.Adam([tp], lr=1e-2, betas=(0.9, 0.999), eps=1e-8)
    p, m, v = p0.copy(), np.zeros(n, np.float32), np.zeros(n, np.float32)
    pb, mb, vb = p0.copy(), np.zeros(n, np.float32), np.zeros(n, np.float32)
    for k, g in enumerate(grads):
        tp.grad = torch.from_numpy(g.copy())
        opt.step()
        gg = g.copy()
        O.check(O.load().oracle_adam_step(O.nptr(p), O.nptr(gg), O.nptr(m), O.nptr(v), n, 1e-2, 0.9, 0.999, 1e-8,
                                          k + 1, 1), "adam")
        assert np.all(gg == 0)
        gb = g.copy()
        O.check(O.load().oracle_adam_step_bounded(O.nptr(pb), O.nptr(gb), O.nptr(mb), O.nptr(vb), n, 1e-2, 0.9, 0.999,
                                                  1e-8, k + 1, 0.0, 0.5, 0), "adam_bounded")
        assert np.all(gb == g)
    np.testing.assert_allclose(p, tp.detach().numpy(), rtol=1e-5, atol=1e-6)
    assert pb.min() >= 0.0 and pb.max() <= 0.5


def test_cell_table_layout():
    """vr_grid_pack_cells' layout (oracle twin): cell (z, y, x) of a strided
    [D,H,W,1] view holds its 8 corners, x fastest, then y, then z."""
    rng = np.random.default_rng(4)
    full = rng.normal(size=(5, 7, 9, 2)).astype(np.float32)
    g = np.ascontiguousarray(full[..., :1])
    out = np.zeros((4 * 6 * 8, 8), np.float32)
    O.check(O.load().oracle_grid_pack_cells(np_grid(g), O.nptr(out)), "pack")
    v = g[..., 0]
    expect = np.stack([v[z:z + 4, y:y + 6, x:x + 8] for z in (0, 1) for y in (0, 1) for x in (0, 1)], -1)
    np.testing.assert_array_equal(out, expect.reshape(-1, 8))
