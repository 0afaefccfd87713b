"""Pins of the oracle's environment image and quadtree sampler (no GPU):
torch grid_sample, histograms, closed-form weights, quadrature."""
import math

import numpy as np
import pytest
import torch

from helpers import (L, O, Built, Scene, env_density, morton_pixels, np_quadtree, oracle_map_fw, random_rays,
                     sky_image)

SEEDS = (101, 202, 303, 404)


def _xr(w):
    w = np.asarray(w, np.float64)
    x = np.where((w[:, 0] == 0) & (w[:, 2] == 0), 1e-7, w[:, 0])
    return np.stack([np.arctan2(x, w[:, 2]) / np.pi, -2 * np.arcsin(np.clip(w[:, 1], -1, 1)) / np.pi], 1)


def _env_lookup(img, w):
    """Image2D over xr(w): torch grid_sample(bilinear, border, align_corners=False),
    the equivalence the reference states (_maps.py:2307-2316)."""
    c = torch.from_numpy(_xr(w))
    t = torch.from_numpy(img.astype(np.float64)).permute(2, 0, 1)[None]
    out = torch.nn.functional.grid_sample(t, c.reshape(1, -1, 1, 2), mode="bilinear", padding_mode="border",
                                          align_corners=False)
    return out[0, :, :, 0].T.numpy()


def _scene(img, quadtree=None, levels=0, **kw):
    args = dict(sigma=np.zeros((4, 4, 4, 1), np.float32), env_image=img, env_quadtree=quadtree,
                env_quadtree_levels=levels)
    args.update(kw)
    return Scene(**args)


def _env_samples(sc, n, seeds=SEEDS):
    b = Built(sc, "cpu")
    out = np.zeros((n, 6), np.float32)
    sd = np.asarray(seeds, np.uint32)
    O.check(O.load().oracle_env_sample(b.medium, O.nptr(sd), n, O.nptr(out)), "env_sample")
    return out


def test_env_image_lookup_matches_grid_sample():
    """sigma = 0: the NEE radiance of every ray is environment(w) exactly."""
    img = sky_image(16, seed=1)
    rays = np.concatenate([random_rays(2000, seed=3), random_rays(500, seed=4, inside=True)])
    rays[0, 3:] = (0, 1, 0)   # poles and the seam
    rays[1, 3:] = (0, -1, 0)
    rays[2, 3:] = (0, 0, -1)
    out = oracle_map_fw(_scene(img), L.VR_MAP_NEE_DT, rays, SEEDS)
    np.testing.assert_allclose(out, _env_lookup(img, rays[:, 3:]), rtol=1e-5, atol=1e-6)


def test_np_quadtree_is_the_block_sum_pyramid():
    d = env_density(sky_image(16, seed=2))
    q = np_quadtree(d).astype(np.float64)
    assert q.shape == (4 + 16 + 64 + 256,)
    assert abs(q[:4].sum() - 1) < 1e-6
    # XRQuadtreeRandomDirection's walk: node n >= 1 (= 4 parent + 1 + child) has
    # probability q[n - 1] and its children at q[4n .. 4n + 3]
    for node in range(1, 1 + 4 + 16):
        assert abs(q[4 * node:4 * node + 4].sum() - q[node - 1]) < 1e-6, node
    px, py = morton_pixels(256)
    np.testing.assert_allclose(q[84:][np.argsort(py * 16 + px)], d.ravel(), rtol=1e-6)


def test_quadtree_sampler_follows_the_density_with_exact_weights():
    R, levels = 16, 4
    img = sky_image(R, seed=5)
    dens = env_density(img)
    sc = _scene(img, np_quadtree(dens), levels)
    n = 200000
    s = _env_samples(sc, n)
    wo, Lw = s[:, :3].astype(np.float64), s[:, 3:].astype(np.float64)
    np.testing.assert_allclose(np.linalg.norm(wo, axis=1), 1.0, atol=1e-6)
    xr = _xr(wo)
    px = np.clip(np.floor((xr[:, 0] + 1) / 2 * R), 0, R - 1).astype(int)
    py = np.clip(np.floor((xr[:, 1] + 1) / 2 * R), 0, R - 1).astype(int)
    # pixel frequencies follow the density (chi-square over pixels with enough mass)
    h = np.bincount(py * R + px, minlength=R * R)
    e = n * dens.ravel()
    m = e > 20
    chi2 = ((h[m] - e[m]) ** 2 / e[m]).sum()
    k = m.sum() - 1
    assert chi2 < k + 6 * math.sqrt(2 * k), (chi2, k)
    # L / pdf = env(wo) * solid angle(pixel) / probability(pixel)
    area = 2 * np.pi * (np.cos(py * np.pi / R) - np.cos((py + 1) * np.pi / R)) / R
    want = _env_lookup(img, wo) * (area / dens[py, px])[:, None]
    inner = (np.abs((xr[:, 0] + 1) / 2 * R - np.round((xr[:, 0] + 1) / 2 * R)) > 1e-3) & \
            (np.abs((xr[:, 1] + 1) / 2 * R - np.round((xr[:, 1] + 1) / 2 * R)) > 1e-3)  # off pixel edges
    np.testing.assert_allclose(Lw[inner], want[inner], rtol=2e-4)


def test_quadtree_estimate_equals_the_environment_integral():
    """E[L / pdf] = integral of env over the sphere (midpoint quadrature of the
    bilinear image on a fine theta-phi grid), and the uniform sampler agrees."""
    R = 16
    img = sky_image(R, seed=7)
    th = (np.arange(512) + 0.5) / 512 * np.pi
    ph = (np.arange(1024) + 0.5) / 1024 * 2 * np.pi - np.pi
    T, P = np.meshgrid(th, ph, indexing="ij")
    w = np.stack([np.sin(T) * np.sin(P), np.cos(T), np.sin(T) * np.cos(P)], -1).reshape(-1, 3)
    integral = (_env_lookup(img, w) * (np.sin(T).reshape(-1, 1))).sum(0) * (np.pi / 512) * (2 * np.pi / 1024)
    n = 200000
    for sc in (_scene(img, np_quadtree(env_density(img)), 4), _scene(img)):
        Lw = _env_samples(sc, n)[:, 3:].astype(np.float64)
        se = Lw.std(0) / math.sqrt(n)
        assert np.all(np.abs(Lw.mean(0) - integral) <= 4 * se + 1e-3 * integral), (Lw.mean(0), integral, se)


def test_nee_with_quadtree_and_uniform_environment_sampling_agree():
    """The same NEE radiance expectation whichever environment sampler is used."""
    img = sky_image(16, seed=9)
    g = np.full((6, 6, 6, 1), 2.0, np.float32)
    ray = np.array([[0.05, 0.1, -1.5, 0.0, 0.0, 1.0]], np.float32)
    n = 40000
    means = []
    for q in (np_quadtree(env_density(img)), None):
        sc = Scene(sigma=g, env_image=img, env_quadtree=q, env_quadtree_levels=4 if q is not None else 0,
                   albedo_c=(0.8, 0.8, 0.8), phase_kind=L.VR_PHASE_HG, phase_g=0.3, max_bounces=8)
        out = oracle_map_fw(sc, L.VR_MAP_NEE_DT, np.tile(ray, (n, 1)), SEEDS).astype(np.float64)
        means.append((out.mean(0), out.std(0) / math.sqrt(n)))
    (m1, s1), (m2, s2) = means
    assert np.all(np.abs(m1 - m2) <= 4 * np.sqrt(s1 ** 2 + s2 ** 2)), (m1, m2, s1, s2)
    assert np.all(s1 < s2)  # importance sampling lowers the variance for a peaked sky


def test_image_tools_resample_img_and_morton():
    rng = np.random.default_rng(3)
    x = rng.normal(size=(8, 8, 2)).astype(np.float32)
    lib = O.load()
    s8, s4, s13 = np.array([8, 8], np.int64), np.array([4, 4], np.int64), np.array([13, 11], np.int64)  # kept alive
    # pixel-centred 2x down = 2x2 block means (exact: power-of-two coordinates)
    out = np.zeros((4, 4, 2), np.float32)
    O.check(lib.oracle_resample_img(O.nptr(x), O.nptr(s8), O.nptr(s4), 2, 2, O.nptr(out)), "resample_img")
    np.testing.assert_allclose(out, x.reshape(4, 2, 4, 2, 2).mean(axis=(1, 3)), rtol=1e-6, atol=1e-6)
    # against torch interpolate(align_corners=False) for an up-sampling
    up = np.zeros((13, 11, 2), np.float32)
    O.check(lib.oracle_resample_img(O.nptr(x), O.nptr(s8), O.nptr(s13), 2, 2, O.nptr(up)), "resample_img")
    ref = torch.nn.functional.interpolate(torch.from_numpy(x).permute(2, 0, 1)[None].double(), size=(13, 11),
                                          mode="bilinear", align_corners=False)[0].permute(1, 2, 0).numpy()
    np.testing.assert_allclose(up, ref, rtol=1e-5, atol=1e-5)
    # Morton copies: a permutation and its inverse
    m = np.zeros((64, 2), np.float32)
    O.check(lib.oracle_copy_img_to_morton(O.nptr(x), 8, 2, O.nptr(m)), "to_morton")
    px, py = morton_pixels(64)
    np.testing.assert_array_equal(m, x[py, px])
    back = np.zeros_like(x)
    O.check(lib.oracle_copy_morton_to_img(O.nptr(m), 8, 2, O.nptr(back)), "from_morton")
    np.testing.assert_array_equal(back, x)
