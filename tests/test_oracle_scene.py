"""Pins of the scene-level path tracer (PathtracedScene, include/sc_path_tracing.h)
in the oracle, and of the library's host-side BVH builder (no GPU): closed-form
hits on a cube, exact transparency, the white furnace, absorption = exp(-tau)
by quadrature, mirror reflection of an environment image (torch grid_sample),
Fresnel reflectance (Schlick), surface emission formulas, and the backward
against central finite differences."""
import math

import numpy as np
import pytest
import torch

from helpers import (L, O, SceneBuilt, blob_grid, box_mesh, patch, pose, quad_mesh, scene_capture_bw, scene_capture_fw,
                     scene_map_fw, scene_medium, scene_raycast)
from test_oracle import _exact_T
from test_oracle_env import _env_lookup

SEEDS = (31, 41, 59, 26)


def _unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


# ------------------------------------------------------------------- BVH --
def _soup(n, seed):
    rng = np.random.default_rng(seed)
    c = rng.uniform(-1, 1, size=(n, 1, 3))
    tri = c + rng.normal(scale=0.08, size=(n, 3, 3))
    return tri.reshape(-1, 3).astype(np.float32), np.arange(3 * n, dtype=np.int32).reshape(n, 3)


def test_bvh_builder_invariants():
    """Every triangle in exactly one leaf; leaf boxes hold their triangles; parents
    hold their children; depth <= 30; at most 2T - 1 nodes (host builder, no GPU)."""
    pos, idx = _soup(700, 3)
    bp, bi = box_mesh(1.0)
    pos = np.concatenate([pos, bp])
    idx = np.concatenate([idx, bi + 2100])
    sb = SceneBuilt([(pos, idx)], [patch()], device="cpu")
    nodes, ids, verts = sb.nodes_np[:sb.num_nodes], sb.tri_ids_np, sb.tri_verts_np
    T = idx.shape[0]
    assert 1 <= sb.num_nodes <= 2 * T - 1
    assert sorted(ids.tolist()) == list(range(T))
    np.testing.assert_array_equal(verts.reshape(T, 3, 3), pos[idx[ids]])
    first, count = nodes[:, 3].view(np.int32), nodes[:, 7].view(np.int32)
    seen = np.zeros(T, int)
    stack = [(0, 0)]
    while stack:
        n, depth = stack.pop()
        assert depth <= 30
        lo, hi = nodes[n, :3], nodes[n, 4:7]
        if count[n] > 0:
            tv = verts[first[n]:first[n] + count[n]].reshape(-1, 3)
            assert (tv >= lo - 1e-6).all() and (tv <= hi + 1e-6).all()
            seen[first[n]:first[n] + count[n]] += 1
        else:
            for c in (first[n], first[n] + 1):
                assert (nodes[c, :3] >= lo - 1e-6).all() and (nodes[c, 4:7] <= hi + 1e-6).all()
                stack.append((c, depth + 1))
    assert (seen == 1).all()


def test_bvh_builder_rejects_bad_indices():
    pos, idx = box_mesh(1.0)
    idx = idx.copy()
    idx[3, 1] = 99
    nodes = np.zeros((64, 8), np.float32)
    out = np.zeros(12, np.int32), np.zeros((12, 9), np.float32)
    nn = L.ctypes.c_int32(0)
    st = L.load().vr_scene_build_bvh(pos.ctypes.data, pos.shape[0], idx.ctypes.data, 12, nodes.ctypes.data, 64,
                                     out[0].ctypes.data, out[1].ctypes.data, L.ctypes.byref(nn))
    assert st == L.VR_ERR_ARG and b"out of range" in L.load().vr_last_error()


# ------------------------------------------------------------ raycasting --
def test_oracle_raycast_cube_closed_form():
    """Rays from outside hit the +z face at t = (z0 - 1)/|w_z| with G = +z; rays
    from the centre leave through the face of the largest |w| component at
    t = 1/max|w_i| with G along w; rays pointing away miss (patch -1)."""
    sb = SceneBuilt([box_mesh(1.0)], [patch()], device="cpu")
    rng = np.random.default_rng(0)
    tgt = np.concatenate([rng.uniform(-0.9, 0.9, size=(200, 2)), np.ones((200, 1))], 1)
    o = np.array([0.1, -0.2, 3.0])
    w = _unit(tgt - o)
    h = scene_raycast(sb, np.concatenate([np.tile(o, (200, 1)), w], 1))
    np.testing.assert_allclose(h[:, 0], (o[2] - 1) / -w[:, 2], rtol=1e-6)
    np.testing.assert_allclose(h[:, 4:7], np.tile([0, 0, 1], (200, 1)), atol=1e-6)
    assert (h[:, 15].view(np.int32) == 0).all()
    w2 = _unit(rng.normal(size=(300, 3)))
    h2 = scene_raycast(sb, np.concatenate([np.zeros((300, 3)), w2], 1))
    np.testing.assert_allclose(h2[:, 0], 1 / np.abs(w2).max(1), rtol=1e-6)
    assert (np.einsum("ij,ij->i", h2[:, 4:7], w2) > 0).all()  # leaving: G along w
    away = np.concatenate([np.tile([0.0, 0.0, 3.0], (50, 1)), _unit(rng.normal(size=(50, 3)) * 0.1 + [0, 0, 1])], 1)
    assert (scene_raycast(sb, away)[:, 15].view(np.int32) == -1).all()


# ------------------------------------------------------------ path tracer --
def _rays_through_centre(n, seed, radius=3.0, spread=0.4):
    rng = np.random.default_rng(seed)
    o = _unit(rng.normal(size=(n, 3))) * radius
    w = _unit(rng.uniform(-spread, spread, size=(n, 3)) - o)
    return np.concatenate([o, w], 1).astype(np.float32)


def test_oracle_empty_and_transparent_scenes_return_the_environment():
    rays = _rays_through_centre(64, 1)
    env = (0.3, 0.6, 0.9)
    empty = SceneBuilt([], [patch()], env_c=env, device="cpu")
    np.testing.assert_allclose(scene_map_fw(empty, rays, SEEDS), np.tile(env, (64, 1)), rtol=1e-6)
    null_box = SceneBuilt([box_mesh(1.0)], [patch(L.VR_SURF_NULL)], env_c=env, device="cpu")
    np.testing.assert_allclose(scene_map_fw(null_box, rays, SEEDS), np.tile(env, (64, 1)), rtol=1e-6)
    block = SceneBuilt([box_mesh(1.0)], [patch(L.VR_SURF_BLOCK)], env_c=env, device="cpu")
    np.testing.assert_allclose(scene_map_fw(block, rays, SEEDS), 0.0)


def test_oracle_white_furnace_through_a_null_boundary():
    """Albedo 1, consistent HG, constant environment L: every path returns L."""
    g = blob_grid(10, seed=4, max_value=8.0)
    med = scene_medium(g, majorant=8.5, albedo_c=(1, 1, 1), phase_g=0.4)
    sb = SceneBuilt([box_mesh(1.0)], [patch(L.VR_SURF_NULL, inside=0)], media=[med], env_c=(0.7, 0.7, 0.7),
                    device="cpu")
    out = scene_map_fw(sb, _rays_through_centre(2000, 2), SEEDS)
    np.testing.assert_allclose(out, 0.7, rtol=1e-5)


def test_oracle_absorbing_medium_is_transmittance_times_env():
    """medium_filter 0 (no scatter, no gathering): a path that collides ends with
    W = 0, so E[A] = env * exp(-integral of the trilinear sigma) (quadrature)."""
    g = blob_grid(8, seed=5, max_value=4.0)
    med = scene_medium(g, majorant=4.5, medium_filter=0)
    sb = SceneBuilt([box_mesh(1.0)], [patch(L.VR_SURF_NULL, inside=0)], media=[med], env_c=(1, 1, 1), device="cpu")
    for ray in _rays_through_centre(3, 7, spread=0.2):
        n = 60000
        A = scene_map_fw(sb, np.tile(ray, (n, 1)), SEEDS)[:, 0]
        exact = _exact_T(g, ray)
        se = A.std() / math.sqrt(n)
        assert abs(A.mean() - exact) <= 4 * se + 1e-4, (A.mean(), exact, se)


def test_oracle_mirror_plane_reflects_the_environment_image():
    rng = np.random.default_rng(3)
    img = rng.uniform(0.1, 2.0, size=(16, 32, 3)).astype(np.float32)
    sb = SceneBuilt([quad_mesh(0.0)], [patch(L.VR_SURF_MIRROR)], env_image=img, device="cpu")
    o = np.tile([0.0, 2.0, 0.0], (100, 1))
    w = _unit(np.concatenate([rng.uniform(-1, 1, (100, 1)), -rng.uniform(0.2, 1, (100, 1)),
                              rng.uniform(-1, 1, (100, 1))], 1))
    out = scene_map_fw(sb, np.concatenate([o, w], 1), SEEDS)
    refl = w * np.array([1, -1, 1])
    np.testing.assert_allclose(out, _env_lookup(img, refl), rtol=2e-5, atol=2e-6)


@pytest.mark.parametrize("elev_deg", [25.0, 50.0, 75.0])
def test_oracle_fresnel_plane_reflects_with_schlick_probability(elev_deg):
    """Environment 1 above the horizon, 0 below: reflections (Schlick probability R
    with eta = 1/n) see 1, refractions go down and see 0; every path carries the
    reference's weight 1/eta^2 (_scenes.py:964-1000), so E[A] = R / eta^2."""
    n_ior = 1.5
    img = np.zeros((64, 128, 3), np.float32)
    img[:32] = 1.0  # rows of the upper hemisphere (xr v < 0)
    sb = SceneBuilt([quad_mesh(0.0)], [patch(L.VR_SURF_FRESNEL, refraction_index=n_ior)], env_image=img,
                    device="cpu")
    e = math.radians(elev_deg)
    w = np.array([math.cos(e), -math.sin(e), 0.0])
    ray = np.concatenate([[0.0, 1.0, 0.0], w]).astype(np.float32)
    N = 40000
    A = scene_map_fw(sb, np.tile(ray, (N, 1)), SEEDS)[:, 0]
    eta = 1.0 / n_ior
    f = ((1 - eta) / (1 + eta)) ** 2
    R = f + (1 - f) * (1 - math.sin(e)) ** 5
    se = A.std() / math.sqrt(N)
    assert abs(A.mean() - R / eta ** 2) <= 4 * se + 1e-4, (A.mean(), R / eta ** 2, se)


def test_oracle_surface_emission_formulas():
    """Emitting plane with no sampler (the path passes through), environment 0:
    Lambert emission gives E dot(-w, fN)/pi, isotropic E/(2 pi), from either side."""
    rng = np.random.default_rng(9)
    w = _unit(np.concatenate([rng.uniform(-1, 1, (60, 1)), -rng.uniform(0.1, 1, (60, 1)), rng.uniform(-1, 1, (60, 1))], 1))
    rays = np.concatenate([np.tile([0.0, 1.0, 0.0], (60, 1)), w], 1).astype(np.float32)
    E = (2.0, 1.0, 0.5)
    for up in (True, False):
        lam = SceneBuilt([quad_mesh(0.0, up=up)], [patch(emission=L.VR_EMIT_LAMBERT, emission_c=E)], env_c=(0, 0, 0),
                         device="cpu")
        cos = -w[:, 1]  # dot(-w, fN) with fN facing the incoming ray
        np.testing.assert_allclose(scene_map_fw(lam, rays, SEEDS), np.outer(cos / math.pi, E), rtol=2e-5, atol=1e-7)
        iso = SceneBuilt([quad_mesh(0.0, up=up)], [patch(emission=L.VR_EMIT_ISOTROPIC, emission_c=E)], env_c=(0, 0, 0),
                         device="cpu")
        np.testing.assert_allclose(scene_map_fw(iso, rays, SEEDS), np.tile(np.asarray(E) / (2 * math.pi), (60, 1)),
                                   rtol=1e-6)


def test_oracle_lambert_floor_paths_escape_upwards():
    """A Lambert floor under a constant sky: every path that hits the floor leaves
    upwards with weight 1 (one bounce, no roulette) and returns the sky."""
    sb = SceneBuilt([quad_mesh(0.0)], [patch(L.VR_SURF_LAMBERT)], env_c=(0.5, 0.25, 1.0), device="cpu")
    rays = np.concatenate([np.tile([0.0, 2.0, 0.0], (500, 1)),
                           _unit(np.random.default_rng(2).normal(size=(500, 3)) * [1, 0.1, 1] + [0, -1, 0])], 1)
    np.testing.assert_allclose(scene_map_fw(sb, rays.astype(np.float32), SEEDS), np.tile([0.5, 0.25, 1.0], (500, 1)),
                               rtol=1e-6)


def test_oracle_scene_backward_matches_finite_differences():
    """A scattering medium behind a null boundary, over a Lambert floor, with an
    emitting panel: mean of the backward (consistent estimator, replayed
    forward seeds) against central differences of the forward (common random
    numbers on both sides, fixed majorant)."""
    rng = np.random.default_rng(11)
    g = rng.uniform(1.0, 4.0, size=(5, 5, 5, 1)).astype(np.float32)
    fl_p, fl_i = quad_mesh(-0.6)

    def build(s, grad=False):
        med = scene_medium(s, majorant=4.5, albedo_c=(0.8, 0.8, 0.8), phase_g=0.3)
        return SceneBuilt([box_mesh(0.5), (fl_p, fl_i)],
                          [patch(L.VR_SURF_NULL, inside=0), patch(L.VR_SURF_LAMBERT, L.VR_EMIT_ISOTROPIC, (0.5, 0.5, 0.5))],
                          media=[med], env_c=(1.0, 0.8, 0.6), device="cpu", grad_sigma=grad)

    cam = dict(poses=np.stack([pose((0.3, 0.8, 2.0)), pose((-1.8, 0.4, 0.9))]), width=6, height=6, jittered=True)
    w = (rng.uniform(size=(72, 3)) + 0.5).astype(np.float32)
    grad = scene_capture_bw(build(g, True), cam, (1, 2, 3, 4), 16384, w)[0][0]
    h = 0.4
    for v in [(2, 2, 2, 0), (1, 3, 2, 0)]:
        r = []
        for sgn in (1, -1):
            s2 = g.copy()
            s2[v] += sgn * h
            r.append(float((scene_capture_fw(build(s2), cam, (7, 8, 9, 10), 65536).astype(np.float64) * w).sum()))
        fd = (r[0] - r[1]) / (2 * h)
        assert abs(grad[v] - fd) <= 0.08 * abs(fd), (v, grad[v], fd)


def _chord(ray, half, center=(0.0, 0.0, 0.0)):
    o, d = ray[:3].astype(np.float64) - np.asarray(center), ray[3:].astype(np.float64)
    t0 = np.minimum((-half - o) / d, (half - o) / d).max()
    t1 = np.maximum((-half - o) / d, (half - o) / d).min()
    return max(0.0, t1 - max(t0, 0.0))


def test_oracle_nested_media_follow_the_medium_stack():
    """A null box (medium A, constant sigma_A) around a smaller null box (medium B,
    sigma_B), both absorbing: inside B only B's medium acts, and after leaving B
    the stack returns to A, so E[A] = exp(-sigma_A (L_A - L_B) - sigma_B L_B)."""
    sa, sb = 1.3, 3.1
    med_a = scene_medium(np.full((2, 2, 2, 1), sa, np.float32), majorant=sa + 0.5, medium_filter=0,
                         bmin=(-1.0,) * 3, bmax=(1.0,) * 3)
    med_b = scene_medium(np.full((2, 2, 2, 1), sb, np.float32), majorant=sb + 0.5, medium_filter=0,
                         bmin=(-0.4,) * 3, bmax=(0.4,) * 3)
    sbuilt = SceneBuilt([box_mesh(1.0), box_mesh(0.4)], [patch(L.VR_SURF_NULL, inside=0), patch(L.VR_SURF_NULL, inside=1)],
                        media=[med_a, med_b], env_c=(1, 1, 1), device="cpu")
    for ray in _rays_through_centre(4, 12, spread=0.3):
        n = 40000
        A = scene_map_fw(sbuilt, np.tile(ray, (n, 1)), SEEDS)[:, 0]
        la, lb = _chord(ray, 1.0), _chord(ray, 0.4)
        exact = math.exp(-sa * (la - lb) - sb * lb)
        se = A.std() / math.sqrt(n)
        assert abs(A.mean() - exact) <= 4 * se + 1e-4, (A.mean(), exact, se, la, lb)


def test_mesh_geometry_api_obj_group_and_transform(tmp_path):
    """load_obj normalises into [-1, 1] and fans polygons; GroupGeometry concatenates
    patches; TransformedGeometry moves positions by p @ R + t (host side, no GPU)."""
    import rendervous_amd as rv
    obj = tmp_path / "quad.obj"
    obj.write_text("v 0 0 0\nv 4 0 0\nv 4 2 0\nv 0 2 0\nf 1 2 3 4\n")
    m = rv.MeshGeometry.load_obj(str(obj))
    assert m.indices.tolist() == [[0, 1, 2], [0, 2, 3]]
    np.testing.assert_allclose(m.positions.detach().numpy(), [[-1, -0.5, 0], [1, -0.5, 0], [1, 0.5, 0], [-1, 0.5, 0]])
    box = rv.MeshGeometry.box()
    p = box.positions.detach().numpy()
    assert np.allclose(np.abs(p), 1.0)
    T = torch.tensor([[0.0, 1.0, 0.0], [-1.0, 0.0, 0.0], [0.0, 0.0, 2.0], [0.5, -1.0, 3.0]])
    moved = rv.TransformedGeometry(box, T)
    pm, nm, _, _ = moved._patches()[0]
    np.testing.assert_allclose(pm.numpy(), p @ T[:3].numpy() + T[3].numpy(), atol=1e-6)
    n0 = box.normals.detach().numpy()
    expect = n0 @ np.linalg.inv(T[:3].numpy()).T
    np.testing.assert_allclose(nm.numpy(), expect / np.linalg.norm(expect, axis=1, keepdims=True), atol=1e-6)
    g = rv.GroupGeometry(box, m, moved)
    assert len(g) == 3 and [len(pp[3]) for pp in g._patches()] == [12, 2, 12]
