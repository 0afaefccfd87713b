"""Independent pins of oracle functions that the furnace and finite-difference
tests leave open (no GPU): the Henyey-Greenstein phase function and sampler,
the perspective camera, the directional light of the environment sampler, and
the period of the Tausworthe components of the HybridTaus generator.

Each check comes from the mathematics or from a restatement the reference
itself holds, not from the oracle's own formula:

* HG (include/scattering_AD.h:4-47): the phase function integrates to 1 over
  the sphere and its mean cosine is g (quadrature); sampled cosines follow the
  analytic HG distribution (Kolmogorov-Smirnov) with mean g, azimuths are
  uniform, and the sampler's histogram equals the evaluated phase function.
* Camera (include/maps/sensor_perspective_camera.h:13-24): rays equal a numpy
  restatement of the torch camera the reference keeps beside it
  (_maps.py:1788-1801), and the angles of the edge pixels follow from the FOV
  in closed form (the vertical FOV is set by the image height).
* Directional light (MixtureSampler semantics, _maps.py:1541-1571): the single
  scattering of a homogeneous box lit by the sun alone equals a 1-D quadrature
  of sigma * T(cam -> x) * albedo * phase * E * T(x -> sun), for two selection
  probabilities.
* Tausworthe (include/randoms.h:6,41-47): each component's transition over
  GF(2) on its 31 / 29 / 28 significant bits has maximal order 2^k - 1.
"""
import math

import numpy as np
import pytest

from helpers import L, O, Built, Camera, Scene, oracle_map_fw, pose

KS_CRIT = 1.95  # sqrt(n) * D at p = 0.001


def _medium(**kw):
    sc = Scene(sigma=np.zeros((2, 2, 2, 1), np.float32), **kw)
    return Built(sc, "cpu")


def _phase_eval(b, w, wo):
    w = np.ascontiguousarray(w, np.float64)
    wo = np.ascontiguousarray(wo, np.float64)
    out = np.zeros(w.shape[0], np.float64)
    O.check(O.load().oracle_phase_eval(b.medium, O.nptr(w), O.nptr(wo), w.shape[0], O.nptr(out)), "phase_eval")
    return out


def _phase_sample(b, w, seeds):
    w = np.ascontiguousarray(w, np.float32)
    out = np.zeros((w.shape[0], 5), np.float32)
    O.check(O.load().oracle_phase_sample(b.medium, O.nptr(w), O.nptr(np.asarray(seeds, np.uint32)), w.shape[0],
                                         O.nptr(out)), "phase_sample")
    return out


def _unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def _hg_cdf(mu, g):
    """CDF of cos(theta) under HG(g): integral of 2 pi p over [-1, mu]."""
    if abs(g) < 1e-3:
        return 0.5 * (mu + 1.0)
    return (1.0 - g * g) / (2.0 * g) * (1.0 / np.sqrt(1.0 + g * g - 2.0 * g * mu) - 1.0 / (1.0 + g))


# ------------------------------------------------------------------- HG --
@pytest.mark.parametrize("g", [-0.6, 0.3, 0.8])
def test_hg_phase_integrates_to_one_with_mean_cosine_g(g):
    b = _medium(phase_kind=L.VR_PHASE_HG, phase_g=g)
    mu, wq = np.polynomial.legendre.leggauss(4000)
    rng = np.random.default_rng(1)
    # a random propagation direction and its orthonormal frame: the phase
    # function depends on the angle between w and wo only
    w = _unit(rng.normal(size=3))
    t = _unit(np.cross(w, [0.3, 0.5, 0.8]))
    phi = rng.uniform(0, 2 * math.pi, size=mu.shape)
    s = np.sqrt(1 - mu * mu)
    b2 = np.cross(w, t)
    wo = mu[:, None] * w + (s * np.cos(phi))[:, None] * t + (s * np.sin(phi))[:, None] * b2
    p = _phase_eval(b, np.tile(w, (len(mu), 1)), wo)
    assert abs(2 * math.pi * (wq * p).sum() - 1.0) < 1e-6
    assert abs(2 * math.pi * (wq * p * mu).sum() - g) < 1e-6
    # forward scattering for g > 0: the peak sits at mu = sign(g)
    assert (p[np.argmax(mu)] > p[np.argmin(mu)]) == (g > 0)


@pytest.mark.parametrize("g", [-0.6, 0.3, 0.8])
def test_hg_sampler_follows_the_analytic_distribution(g):
    b = _medium(phase_kind=L.VR_PHASE_HG, phase_g=g)
    n = 200000
    rng = np.random.default_rng(2)
    w = _unit(rng.normal(size=(n, 3))).astype(np.float32)
    out = _phase_sample(b, w, (11, 22, 33, 44)).astype(np.float64)
    wo, rho = out[:, :3], out[:, 3]
    np.testing.assert_allclose(np.linalg.norm(wo, axis=1), 1.0, atol=1e-5)
    assert np.all(rho == 1.0)
    mu = np.clip((wo * w).sum(1), -1, 1)
    # Kolmogorov-Smirnov against the closed-form CDF
    ms = np.sort(mu)
    F = _hg_cdf(ms, g)
    ks = max(np.abs(F - np.arange(1, n + 1) / n).max(), np.abs(F - np.arange(0, n) / n).max())
    assert ks * math.sqrt(n) < KS_CRIT, ks
    # mean cosine g (variance of mu under HG: E[mu^2] - g^2 from the quadrature)
    se = mu.std() / math.sqrt(n)
    assert abs(mu.mean() - g) < 4 * se
    # azimuth about w is uniform
    t = _unit(np.cross(w, np.array([0.3, 0.5, 0.8])))
    b2 = np.cross(w, t)
    az = np.sort(np.arctan2((wo * b2).sum(1), (wo * t).sum(1)) % (2 * math.pi)) / (2 * math.pi)
    ks_az = np.abs(az - np.arange(1, n + 1) / n).max()
    assert ks_az * math.sqrt(n) < KS_CRIT, ks_az


@pytest.mark.parametrize("g", [-0.6, 0.8])
def test_hg_sampler_histogram_equals_the_evaluated_phase(g):
    """The sampled cosines' density equals 2 pi p_eval(mu): eval and sampler
    describe the same distribution (the furnace tests need exactly this)."""
    b = _medium(phase_kind=L.VR_PHASE_HG, phase_g=g)
    n = 200000
    w = np.tile(np.array([[0.0, 0.0, 1.0]], np.float32), (n, 1))
    out = _phase_sample(b, w, (5, 6, 7, 8)).astype(np.float64)
    mu = np.clip(out[:, 2], -1, 1)
    edges = np.linspace(-1, 1, 41)
    hist, _ = np.histogram(mu, bins=edges)
    # expected counts from the evaluated phase function (Gauss-Legendre per bin)
    x, wq = np.polynomial.legendre.leggauss(64)
    exp = []
    for a, c in zip(edges[:-1], edges[1:]):
        m = 0.5 * (c - a) * x + 0.5 * (c + a)
        wo = np.stack([np.sqrt(1 - m * m), np.zeros_like(m), m], 1)
        p = _phase_eval(b, np.tile([0.0, 0.0, 1.0], (len(m), 1)), wo)
        exp.append(n * 2 * math.pi * 0.5 * (c - a) * (wq * p).sum())
    exp = np.array(exp)
    chi2 = ((hist - exp) ** 2 / exp).sum()
    assert abs(exp.sum() - n) < 1e-3 * n
    assert chi2 < 39 + 5 * math.sqrt(2 * 39), chi2
    # the per-sample phase value returned with the sample is the same function
    p_at = out[:, 4]
    p_ref = _phase_eval(b, w.astype(np.float64), out[:, :3])
    np.testing.assert_allclose(p_at, p_ref, rtol=2e-5)


def test_hg_mixture_reproduces_the_reference_phase_and_sampler():
    """HGPhase evaluates 0.8 HG(g) + 0.2 HG(g/5) (_maps.py:4453) while
    VHGPhaseSampler draws HG(g/5) with probability 0.8 (_maps.py:4475-4478):
    mean cosines 0.8 g + 0.04 g and 0.2 g + 0.16 g, and the sampler's CDF is the
    0.2 / 0.8 mixture of the HG CDFs."""
    g = 0.7
    b = _medium(phase_kind=L.VR_PHASE_HG_MIX, phase_g=g)
    mu, wq = np.polynomial.legendre.leggauss(4000)
    wo = np.stack([np.sqrt(1 - mu * mu), np.zeros_like(mu), mu], 1)
    p = _phase_eval(b, np.tile([0.0, 0.0, 1.0], (len(mu), 1)), wo)
    assert abs(2 * math.pi * (wq * p).sum() - 1.0) < 1e-6
    assert abs(2 * math.pi * (wq * p * mu).sum() - (0.8 * g + 0.2 * g / 5)) < 1e-6
    n = 200000
    out = _phase_sample(b, np.tile(np.array([[0.0, 0.0, 1.0]], np.float32), (n, 1)), (9, 8, 7, 6)).astype(np.float64)
    ms = np.sort(np.clip(out[:, 2], -1, 1))
    F = 0.2 * _hg_cdf(ms, g) + 0.8 * _hg_cdf(ms, g / 5)
    ks = np.abs(F - np.arange(1, n + 1) / n).max()
    assert ks * math.sqrt(n) < KS_CRIT, ks


# --------------------------------------------------------------- camera --
def _torch_camera_restatement(poses, rows, cols, width, height, fov, znear=0.001):
    """numpy form of the commented torch restatement at _maps.py:1788-1801
    (pixel centres: the non-jittered generation mode)."""
    o, d, n = poses[:, 0:3].astype(np.float64), poses[:, 3:6].astype(np.float64), poses[:, 6:9].astype(np.float64)
    s0 = ((rows + 0.5) * 2 - height) * znear / height
    s1 = ((cols + 0.5) * 2 - width) * znear / height
    t = znear / math.tan(fov * 0.5)
    z = _unit(d)
    x = _unit(np.cross(n, z))
    y = np.cross(z, x)
    w = x * s1[:, None] + y * s0[:, None] + z * t
    return np.concatenate([o + w, _unit(w)], axis=1)


def _camera_rays(poses, width, height, fov, jittered=False, seeds=(1, 2, 3, 4)):
    cam = Camera(poses, width, height, jittered=jittered, fov=fov, device="cpu")
    out = np.zeros((cam.threads, 6), np.float64)
    O.check(O.load().oracle_camera_rays(cam.c, O.nptr(np.asarray(seeds, np.uint32)), O.nptr(out)), "camera_rays")
    return out


@pytest.mark.parametrize("fov", [math.pi / 4, math.pi / 3])
def test_camera_rays_equal_the_reference_torch_restatement(fov):
    W, H = 37, 22  # non-square: the horizontal extent follows from the height
    poses = np.stack([pose((0.3, 0.4, 2.5)), pose((-1.5, 0.2, -1.0), target=(0.1, 0.0, 0.2), up=(0.2, 1.0, 0.1))])
    rays = _camera_rays(poses, W, H, fov)
    p_idx, r, c = np.meshgrid(np.arange(2), np.arange(H), np.arange(W), indexing="ij")
    # the camera descriptor holds fov and znear as float32 (as the reference's
    # parameter buffer does)
    ref = _torch_camera_restatement(poses[p_idx.ravel()], r.ravel().astype(np.float64), c.ravel().astype(np.float64),
                                    W, H, float(np.float32(fov)), znear=float(np.float32(0.001)))
    np.testing.assert_allclose(rays, ref, rtol=0, atol=1e-12)


def test_camera_edge_angles_follow_the_fov():
    """Non-jittered, odd W and H: the centre pixel looks along the pose
    direction; the centre column of row 0 sits atan((H-1)/H tan(fov/2)) off the
    axis (vertical FOV set by H), the centre row of column 0 atan((W-1)/H
    tan(fov/2)) (square pixels); rows grow along the up vector's image."""
    W, H, fov = 41, 21, math.pi / 3
    p = pose((0.0, 0.0, 3.0))
    rays = _camera_rays(p[None], W, H, fov).reshape(H, W, 6)
    axis = _unit(p[3:6])
    centre = rays[H // 2, W // 2, 3:]
    np.testing.assert_allclose(centre, axis, atol=1e-12)
    t = math.tan(float(np.float32(fov)) / 2)  # the descriptor's float32 fov
    ang_v = math.acos(np.clip(rays[0, W // 2, 3:] @ axis, -1, 1))
    ang_h = math.acos(np.clip(rays[H // 2, 0, 3:] @ axis, -1, 1))
    assert abs(ang_v - math.atan((H - 1) / H * t)) < 1e-9
    assert abs(ang_h - math.atan((W - 1) / H * t)) < 1e-9
    # the pose (origin (0,0,3), looking along -z, up +y): row index grows along +y
    # and column index along -x (xaxis = normalize(cross(up, z)))
    assert rays[H - 1, W // 2, 4] > 0 > rays[0, W // 2, 4]
    assert rays[H // 2, 0, 3] > 0 > rays[H // 2, W - 1, 3]
    # ray origins: the pose origin plus the unnormalised direction, whose depth
    # along the axis is znear / tan(fov / 2)
    dist = np.linalg.norm(rays[..., :3] - p[:3], axis=-1)
    np.testing.assert_allclose(dist * (rays[..., 3:] @ axis), float(np.float32(0.001)) / t, rtol=1e-9)


def test_jittered_camera_rays_stay_inside_their_pixel():
    W, H, fov = 16, 12, math.pi / 4
    p = pose((0.2, -0.1, 2.0))[None]
    rays = _camera_rays(p, W, H, fov, jittered=True, seeds=(7, 77, 777, 7777))
    z = _unit(p[0, 3:6])
    x = _unit(np.cross(p[0, 6:9], z))
    y = np.cross(z, x)
    w = rays[:, 3:]
    zn = 0.001
    sz = zn / math.tan(fov / 2)
    sx = (w @ x) / (w @ z) * sz
    sy = (w @ y) / (w @ z) * sz
    u = (sx * H / zn + W) / 2
    v = (sy * H / zn + H) / 2
    rows, cols = np.divmod(np.arange(W * H), W)
    assert np.all((u >= cols - 1e-9) & (u < cols + 1 + 1e-9))
    assert np.all((v >= rows - 1e-9) & (v < rows + 1 + 1e-9))
    frac = np.concatenate([u - cols, v - rows])
    assert 0.3 < frac.mean() < 0.7 and frac.std() > 0.2


# ------------------------------------------------------ directional light --
def _exit_distance(x, s, lo=-0.5, hi=0.5):
    t = np.full(x.shape[0], np.inf)
    for a in range(3):
        if s[a] > 0:
            t = np.minimum(t, (hi - x[:, a]) / s[a])
        elif s[a] < 0:
            t = np.minimum(t, (lo - x[:, a]) / s[a])
    return t


def _hg(mu, g):
    return (1 - g * g) / (4 * math.pi * (1 + g * g - 2 * g * mu) ** 1.5)


def _linear_phase_g(n: int) -> np.ndarray:
    """A phase_g grid [n, n, n, 1] over [-0.5, 0.5]^3 holding g(x) = 0.1 + 0.4 x -
    0.25 y + 0.15 z at its vertices: trilinear interpolation reproduces a linear
    field exactly, so the estimator's g at any point is g(x) up to rounding."""
    c = np.linspace(-0.5, 0.5, n)
    z, y, x = np.meshgrid(c, c, c, indexing="ij")
    return (0.1 + 0.4 * x - 0.25 * y + 0.15 * z)[..., None].astype(np.float32)


def _linear_g(x: np.ndarray) -> np.ndarray:
    return 0.1 + 0.4 * x[..., 0] - 0.25 * x[..., 1] + 0.15 * x[..., 2]


@pytest.mark.parametrize("sun_prob", [0.3, 0.9])
@pytest.mark.parametrize("phase_field", [False, True])
def test_directional_light_single_scatter_equals_quadrature(sun_prob, phase_field):
    """Homogeneous box (sigma 3), albedo 0.8, HG g = 0.6, environment 0, sun
    radiance E along s, one scattering event (max_bounces 1): the NEE estimate's
    mean is the 1-D integral over the ray of
    sigma e^{-sigma t} * albedo * E * e^{-sigma l(x_t)} * p(w . s; g),
    l = distance from x_t to the box along s (Simpson, 2e5 intervals).  With
    phase_field, g is the linear field of _linear_phase_g, read at the collision
    x_t (a read at another point, or with the grid's axes swapped, moves the
    mean by many standard errors)."""
    sig, alb, g = 3.0, 0.8, 0.6
    E = np.array([2.0, 1.0, 0.5])
    s_dir = _unit([0.35, 0.8, 0.3])
    x0 = np.array([0.12, -0.07, 1.6])
    w = _unit([-0.05, 0.03, -1.0])
    sc = Scene(sigma=np.full((4, 4, 4, 1), sig, np.float32), albedo_c=(alb, alb, alb), env_c=(0.0, 0.0, 0.0),
               sun_dir=tuple(s_dir), sun_radiance=tuple(E), sun_prob=sun_prob, phase_kind=L.VR_PHASE_HG, phase_g=g,
               max_bounces=1, phase_g_grid=_linear_phase_g(5) if phase_field else None)
    n = 200000
    ray = np.concatenate([x0, w]).astype(np.float32)
    R = oracle_map_fw(sc, L.VR_MAP_NEE_DT, np.tile(ray, (n, 1)), (21, 22, 23, 24)).astype(np.float64)
    # quadrature along the float32 ray the estimator sees
    x0f, wf = ray[:3].astype(np.float64), ray[3:].astype(np.float64)
    tb = [(np.array([-0.5, 0.5]) - x0f[a]) / wf[a] for a in range(3)]
    tmin = max(float(t.min()) for t in tb)
    tmax = min(float(t.max()) for t in tb)
    m = 200000
    t = np.linspace(tmin, tmax, m + 1)
    xt = x0f + t[:, None] * wf
    gt = _linear_g(xt) if phase_field else g
    f = sig * np.exp(-sig * (t - tmin)) * alb * np.exp(-sig * _exit_distance(xt, s_dir)) * _hg(wf @ s_dir, gt)
    h = (tmax - tmin) / m
    integral = h / 3 * (f[0] + f[-1] + 4 * f[1:-1:2].sum() + 2 * f[2:-1:2].sum())
    expect = integral * E
    se = R.std(axis=0) / math.sqrt(n)
    assert np.all(np.abs(R.mean(axis=0) - expect) <= 4 * se + 1e-6), (R.mean(axis=0), expect, se)
    # the channels carry E's ratios sample by sample
    np.testing.assert_allclose(R[:, 1] * 2.0, R[:, 0], rtol=1e-6)


# ------------------------------------------------------------ Tausworthe --
def _prime_factors(n):
    f, p = [], 2
    while p * p <= n:
        if n % p == 0:
            f.append(p)
            while n % p == 0:
                n //= p
        p += 1
    if n > 1:
        f.append(n)
    return f


def _gf2_mul(A, B):
    """Product of two GF(2) matrices stored as lists of column bitmasks."""
    out = []
    for col in B:
        acc, j = 0, 0
        while col:
            if col & 1:
                acc ^= A[j]
            col >>= 1
            j += 1
        out.append(acc)
    return out


def _gf2_pow(A, e):
    n = len(A)
    R = [1 << j for j in range(n)]
    while e:
        if e & 1:
            R = _gf2_mul(R, A)
        A = _gf2_mul(A, A)
        e >>= 1
    return R


@pytest.mark.parametrize("comp,k", [(0, 31), (1, 29), (2, 28)])
def test_tausworthe_component_has_maximal_period(comp, k):
    """The transition of component `comp` restricted to its k significant
    (top) bits is a primitive element of GL(k, GF(2)): its order is 2^k - 1.
    The step is probed through the oracle's generator on unit states."""
    low = 32 - k
    cols = []
    for j in range(low, 32):
        st = [129, 129, 129, 129]
        st[comp] = 1 << j
        _, fin = O.rng_uniforms(np.array(st, np.uint32), 1)
        cols.append(int(fin[comp]) >> low)  # project the new state onto the top k bits
    order = (1 << k) - 1
    I = [1 << j for j in range(k)]
    assert _gf2_pow(cols, order) == I
    for p in _prime_factors(order):
        assert _gf2_pow(cols, order // p) != I, p
