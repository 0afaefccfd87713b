"""Seeds, device and run-time settings of the map engine.

Mirrors rendervous _internal.py:21-34: ``seed(n)`` arms torch's generator and
``get_seeds()`` draws the 4 dispatch seeds in [129, 2^30) used by every kernel.
"""
from __future__ import annotations

import contextlib
import dataclasses
import os
import typing

import torch

__MANUAL_SEED__: typing.Optional[int] = None


def device() -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError("rendervous_amd needs a HIP device (torch.cuda on ROCm); no CPU fallback exists")
    return torch.device("cuda", torch.cuda.current_device())


def seed(manual_seed: typing.Optional[int] = None) -> None:
    """Arm the seed generator (_internal.py:21-23)."""
    global __MANUAL_SEED__
    __MANUAL_SEED__ = manual_seed


def get_seeds() -> typing.Tuple[int, int, int, int]:
    """Four dispatch seeds in [129, 2^30) drawn from torch's CPU generator (_internal.py:26-34)."""
    global __MANUAL_SEED__
    if __MANUAL_SEED__ is not None:
        torch.manual_seed(__MANUAL_SEED__)
        __MANUAL_SEED__ = None
    s = torch.randint(low=129, high=1 << 30, size=(4,), dtype=torch.int32)
    return tuple(int(v) for v in s.tolist())


@dataclasses.dataclass
class BackwardSettings:
    """How capture/map backward passes draw randoms and which estimators they use.

    seeds:     'fresh' (default) draws new dispatch seeds for the backward, as the
               reference does (_maps.py:864): its paths are independent of the
               forward's, so the gradient of a loss that is nonlinear in the image
               (the L2 loss) is unbiased; each of its bw_samples samples per sensor
               draws its own stream (the per-sample seeding of the forward), so a
               sensor's samples can run on several threads.  'reference' also draws
               new seeds but continues one stream per sensor over its samples, the
               reference's capture-backward init (_maps.py:676-685).  'replay' reuses
               the forward dispatch seeds (the backward replays the forward's
               samples and needs no new ones), which for such a loss adds
               2 Cov(image, gradient estimate) to the expected gradient.
    estimator: 'consistent' fixes the reference backward defects (matched
               primal/replay majorants, single-scatter transmittance gradient
               along the sampled direction, full path weights in the collision
               sampler); 'reference' reproduces the reference formulas.
    debug:     what a capture's debug_out counts in the backward: 'scatters'
               (trilinear gradient scatters) or 'lookups' (trilinear lookups).
    """
    seeds: str = "fresh"
    estimator: str = "consistent"
    debug: str = "scatters"


backward_settings = BackwardSettings()


@dataclasses.dataclass
class Settings:
    """cell_tables: give a medium a cell table of its density (8 corners per
    cell, 8x the grid's memory) so that each trilinear lookup reads one 32-byte
    cell; the values are the same bit for bit as reading the grid.  "auto"
    builds one for densities that do not require a gradient, and for densities
    that do when the integrator scatters many times per sample (NEE, DT path,
    DRT): a density an optimiser updates rebuilds its table at every update,
    which a single-scatter step does not win back; True / False force it on /
    off."""
    cell_tables: typing.Union[str, bool] = "auto"


settings = Settings(cell_tables={"0": False, "1": True}.get(os.environ.get("VR_CELL_TABLES", ""), "auto"))


@contextlib.contextmanager
def backward_mode(seeds: typing.Optional[str] = None, estimator: typing.Optional[str] = None,
                  debug: typing.Optional[str] = None):
    old = dataclasses.replace(backward_settings)
    if debug is not None:
        assert debug in ("scatters", "lookups")
        backward_settings.debug = debug
    if seeds is not None:
        assert seeds in ("replay", "fresh", "reference")
        backward_settings.seeds = seeds
    if estimator is not None:
        assert estimator in ("consistent", "reference")
        backward_settings.estimator = estimator
    try:
        yield
    finally:
        backward_settings.seeds = old.seeds
        backward_settings.estimator = old.estimator
        backward_settings.debug = old.debug
