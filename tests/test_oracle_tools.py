"""Pins of the oracle's grid tools (random ids, total variation, resampling):
independent numpy / torch formulations, closed forms and brute force."""
import math

import numpy as np
import pytest
import torch

from helpers import L, O


def _random_ids(n, shape, seeds):
    out = np.zeros((n, len(shape)), np.int64)
    sh = np.asarray(shape, np.int64)
    sd = np.asarray(seeds, np.uint32)
    O.check(O.load().oracle_random_ids(O.nptr(out), n, O.nptr(sh), len(shape), O.nptr(sd)), "random_ids")
    return out


def _tv(x):
    x = np.ascontiguousarray(x, np.float32)
    sh = np.asarray(x.shape, np.int64)
    out = np.zeros(x.shape[:-1] + (1,), np.float32)
    O.check(O.load().oracle_total_variation_fw(O.nptr(x), O.nptr(sh), x.ndim - 1, O.nptr(out)), "tv")
    return out


def _tv_bw(x, og):
    x = np.ascontiguousarray(x, np.float32)
    og = np.ascontiguousarray(og, np.float32)
    sh = np.asarray(x.shape, np.int64)
    g = np.full(x.shape, 7.0, np.float32)  # overwritten, not accumulated
    O.check(O.load().oracle_total_variation_bw(O.nptr(x), O.nptr(og), O.nptr(sh), x.ndim - 1, O.nptr(g)), "tv_bw")
    return g


def _resample(x, dst):
    x = np.ascontiguousarray(x, np.float32)
    out = np.zeros(tuple(dst) + (x.shape[-1],), np.float32)
    ish = np.asarray(x.shape[:-1], np.int64)
    osh = np.asarray(dst, np.int64)
    O.check(O.load().oracle_resample_grid(O.nptr(x), O.nptr(ish), O.nptr(osh), x.ndim - 1, x.shape[-1], O.nptr(out)),
            "resample")
    return out


# ---------------------------------------------------------- random ids --
def test_random_ids_are_the_function_stream():
    """Thread i's k-th id is int(u_k * shape[k]) with u_k the k-th uniform after the
    compute-function seeding (functions/common_functions.h:36-45), composed here
    from the oracle's seed states and raw generator rather than retyped."""
    seeds = (1234, 5678, 91011, 121314)
    shape = (5, 300, 7)
    ids = _random_ids(64, shape, seeds)
    st = O.seed_states(L.VR_SEED_CAPTURE_BW, seeds, 0, 64)
    for i in range(64):
        u, _ = O.rng_uniforms(st[i], 3)
        want = [int(np.float32(u[k]) * np.float32(shape[k])) for k in range(3)]
        assert ids[i].tolist() == want
    # the seeding itself from the shader text: seeds + (0x23F1, 0x3137, 129, i + 129), 4 steps, float bits of 4 uniforms
    for i in (0, 1, 63):
        s0 = np.array([seeds[0] + 0x23F1, seeds[1] + 0x3137, seeds[2] + 129, seeds[3] + i + 129], np.uint32)
        _, s4 = O.rng_uniforms(s0, 4)
        u4, _ = O.rng_uniforms(s4, 4)
        assert st[i].tolist() == u4.view(np.uint32).tolist()


def test_random_ids_are_uniform_and_in_range():
    shape = (3, 64, 50)
    ids = _random_ids(200000, shape, (7, 8, 9, 10))
    assert ids.min() == 0 and all(ids[:, k].max() == shape[k] - 1 for k in range(3))
    for k, n in enumerate(shape):
        h = np.bincount(ids[:, k], minlength=n)
        e = ids.shape[0] / n
        chi2 = ((h - e) ** 2 / e).sum()
        assert chi2 < (n - 1) + 6 * math.sqrt(2 * (n - 1)), (k, chi2)
    assert abs(np.corrcoef(ids[:, 1], ids[:, 2])[0, 1]) < 0.01


# ------------------------------------------------------ total variation --
def _tv_numpy(x):
    """sum over axes of |v - v(+1)| + |v - v(-1)| where the neighbour exists, by slicing."""
    x = x.astype(np.float64)
    tv = np.zeros(x.shape[:-1], np.float64)
    for ax in range(x.ndim - 1):
        d = np.abs(np.diff(x, axis=ax)).sum(-1)
        lo = [slice(None)] * (x.ndim - 1)
        hi = [slice(None)] * (x.ndim - 1)
        lo[ax] = slice(0, -1)
        hi[ax] = slice(1, None)
        tv[tuple(lo)] += d
        tv[tuple(hi)] += d
    return tv[..., None]


@pytest.mark.parametrize("shape", [(17, 1), (6, 9, 3), (5, 4, 7, 1), (3, 5, 4, 2)])
def test_total_variation_matches_slicing_formula(shape):
    x = np.random.default_rng(len(shape)).normal(size=shape).astype(np.float32)
    np.testing.assert_allclose(_tv(x), _tv_numpy(x), rtol=1e-5, atol=1e-5)


def test_total_variation_closed_forms():
    assert not _tv(np.full((4, 5, 6, 2), 3.0, np.float32)).any()
    # v = 2 x + 3 y along the two inner axes of a 3-D grid: interior voxels see
    # 2 * (2 + 3) = 10; the corner (0,0,0) sees one x and one y neighbour: 5
    z, y, x = np.meshgrid(np.arange(4), np.arange(5), np.arange(6), indexing="ij")
    tv = _tv((2.0 * x + 3.0 * y)[..., None].astype(np.float32))[..., 0]
    assert tv[1, 2, 3] == 10.0 and tv[0, 0, 0] == 5.0 and tv[3, 4, 5] == 5.0


@pytest.mark.parametrize("shape", [(11, 1), (5, 6, 2), (4, 5, 3, 1)])
def test_total_variation_backward_is_the_gradient(shape):
    """Against torch autograd (float64) of the slicing formula; inputs without ties."""
    rng = np.random.default_rng(10 + len(shape))
    x = rng.permutation(np.prod(shape)).reshape(shape).astype(np.float32) * 0.25
    og = rng.normal(size=shape[:-1] + (1,)).astype(np.float32)
    xt = torch.from_numpy(x.astype(np.float64)).requires_grad_(True)
    tv = torch.zeros(shape[:-1], dtype=torch.float64)
    for ax in range(len(shape) - 1):
        d = torch.diff(xt, dim=ax).abs().sum(-1)
        tv = tv + torch.nn.functional.pad(d, _pad(len(shape) - 1, ax, after=True)) + \
            torch.nn.functional.pad(d, _pad(len(shape) - 1, ax, after=False))
    (tv * torch.from_numpy(og[..., 0].astype(np.float64))).sum().backward()
    np.testing.assert_allclose(_tv_bw(x, og), xt.grad.numpy(), rtol=1e-5, atol=1e-5)


def _pad(nd, ax, after):
    p = [0] * (2 * nd)
    k = 2 * (nd - 1 - ax)  # torch pads the last axis first
    p[k + (1 if after else 0)] = 1
    return p


def test_total_variation_backward_ties_use_sign_zero():
    x = np.zeros((3, 3, 1), np.float32)
    assert not _tv_bw(x, np.ones((3, 3, 1), np.float32)).any()


# ------------------------------------------------------------ resample --
@pytest.mark.parametrize("src,dst", [((9,), (17,)), ((6, 5), (11, 3)), ((4, 5, 6), (7, 9, 11)), ((8, 8, 8), (4, 5, 6))])
def test_resample_matches_align_corners_interpolation(src, dst):
    """Vertex-aligned linear resampling is torch's interpolate(align_corners=True)."""
    C = 2
    x = np.random.default_rng(len(src)).normal(size=src + (C,)).astype(np.float32)
    mode = {1: "linear", 2: "bilinear", 3: "trilinear"}[len(src)]
    t = torch.from_numpy(np.moveaxis(x, -1, 0)[None].astype(np.float64))
    ref = torch.nn.functional.interpolate(t, size=dst, mode=mode, align_corners=True)[0].numpy()
    np.testing.assert_allclose(_resample(x, dst), np.moveaxis(ref, 0, -1), rtol=1e-5, atol=1e-5)


def test_resample_identity_and_affine_fields():
    x = np.random.default_rng(3).normal(size=(5, 6, 7, 1)).astype(np.float32)
    assert np.array_equal(_resample(x, (5, 6, 7)), x)
    # an affine field is reproduced exactly (up to rounding) at any resolution
    def field(n):
        z, y, xx = np.meshgrid(*[np.linspace(0, 1, k) for k in n], indexing="ij")
        return (1.0 + 2.0 * xx - 0.5 * y + 3.0 * z)[..., None].astype(np.float32)
    np.testing.assert_allclose(_resample(field((4, 5, 6)), (7, 9, 11)), field((7, 9, 11)), atol=2e-6)


def test_tool_argument_errors():
    lib = O.load()
    sh = np.asarray([4, 4, 1], np.int64)
    x = np.zeros((4, 4, 1), np.float32)
    assert lib.oracle_total_variation_fw(O.nptr(x), O.nptr(sh), 0, O.nptr(x)) == L.VR_ERR_ARG
    one = np.asarray([1, 4], np.int64)
    assert lib.oracle_resample_grid(O.nptr(x), O.nptr(sh), O.nptr(one), 2, 1, O.nptr(x)) == L.VR_ERR_ARG
