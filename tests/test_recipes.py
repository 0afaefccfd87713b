"""Checks of the recipe layer: the DependencySet scope rules, derived
fields (refresh and gradient hand-back) on the CPU, and the environment
quadtree density (resampled on the device)."""
import math
import warnings

import numpy as np
import pytest
import torch

import rendervous_amd as rv
from rendervous_amd import objects


def test_branches_shadow_their_parent_and_see_its_later_fields():
    root = objects.DependencySet()
    root.add_parameters(a=1, b="x")
    child = root.branch()
    child.add_parameters(a=2.5)
    assert child.a == 2.5 and root.a == 1
    root.add_parameters(c=torch.zeros(2))
    assert torch.equal(child.c, torch.zeros(2))
    assert not hasattr(root, "d")
    # ensures looks through every layer for a field of the asked type
    assert child.ensures("a", float) and child.ensures("a", int) and not root.ensures("a", float)
    with pytest.raises(AssertionError, match="'b'"):
        child.assert_ensures("b", int)
    copy = objects.DependencySet()
    copy.copy_parameters(child)
    assert copy.a == 2.5 and copy.b == "x"


def test_derived_fields_refresh_and_hand_gradients_back():
    p = torch.tensor([0.5, 2.0], requires_grad=True)
    root = objects.DependencySet()
    root.add_parameters(sigma_scale=lambda: p * 3.0, module=torch.nn.Identity())
    child = root.branch()
    child.add_parameters(shift=lambda: p.detach() + 1.0)
    leaf, shift = root.sigma_scale, child.shift
    assert leaf.is_leaf and leaf.requires_grad and torch.allclose(leaf, torch.tensor([1.5, 6.0]))
    assert isinstance(root.module, torch.nn.Identity)
    with torch.no_grad():
        p.mul_(2.0)
    child.forward_dependency()  # a branch refreshes the whole tree
    assert root.sigma_scale is leaf and torch.allclose(leaf, torch.tensor([3.0, 12.0]))
    assert torch.allclose(shift, torch.tensor([2.0, 5.0])) and not shift.requires_grad
    (leaf * torch.tensor([1.0, -2.0])).sum().backward()
    root.backward_dependency()
    assert torch.allclose(p.grad, torch.tensor([3.0, -6.0]))
    root.forward_dependency()  # clears the leaf's gradient for the next step
    assert float(leaf.grad.abs().sum()) == 0.0


def test_fallback_recipes_warn():
    ds = objects.DependencySet()
    with warnings.catch_warnings(record=True) as caught:
        warnings.simplefilter("always")
        ds.requires(objects.medium_box)
        ds.requires(objects.medium_phase)
    assert ds.box[0].tolist() == [-1.0, -1.0, -1.0] and ds.box[1].tolist() == [1.0, 1.0, 1.0]
    assert len(caught) == 2 and "isotropic" in str(caught[1].message)


@pytest.mark.gpu
def test_quadtree_density_weights_rows_by_solid_angle(gpu):
    """A constant image gives every cell its share of the sphere's solid angle:
    2 pi (cos t0 - cos t1) / n / (4 pi) for a cell of row [t0, t1]."""
    levels = 4
    n = 1 << levels
    image = torch.full((8, 16, 3), 0.7, device="cuda")
    d = objects._xr_direction_density(image, levels)[..., 0].double().cpu().numpy()
    rows = np.arange(n)
    expect = (np.cos(rows * math.pi / n) - np.cos((rows + 1) * math.pi / n)) / (2.0 * n)
    np.testing.assert_allclose(d, np.repeat(expect[:, None], n, 1), rtol=1e-5)
    assert abs(d.sum() - 1.0) < 1e-5
    # a bright single column carries its luminance share
    image = torch.zeros(16, 16, 3, device="cuda")
    image[:, 5] = 1.0
    d = objects._xr_direction_density(image, levels)[..., 0].double().cpu().numpy()
    np.testing.assert_allclose(d.sum(1), expect * n, rtol=1e-5)
    assert np.count_nonzero(d.sum(0) > 1e-7) <= 3


@pytest.mark.gpu
def test_quadtree_recipe_builds_a_sampler_and_rebuilds_on_refresh(gpu):
    env = torch.rand(32, 64, 3, device="cuda")
    ds = objects.DependencySet()
    ds.add_parameters(environment_tensor=env)
    ds.requires(objects.medium_environment_sampler_quadtree, levels=5)
    q = ds.environment_sampler_quadtree
    assert isinstance(ds.environment_sampler, rv.MapBase)
    before = q.clone()
    env.mul_(torch.linspace(0.1, 1.0, 64, device="cuda")[None, :, None])
    ds.forward_dependency()
    assert ds.environment_sampler_quadtree is q and not torch.equal(before, q)
    with pytest.raises(NotImplementedError):
        objects.medium_environment_sampler_quadtree(objects.DependencySet(parent=ds), projection="cube")
