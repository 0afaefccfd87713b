"""rendervous_amd -- MI355X (gfx950) participating-media hot path.

A drop-in for rendervous's volume-rendering maps (sigma grids, majorants,
DDA / delta / ratio tracking, the collision sampler, the NEE path integrator
and their backward integrators) evaluated by hand-written HIP kernels behind a
C ABI (include/vr_api.h).  The reference's package init picks a device and
starts its Vulkan engine (rendervous __init__.py:12-45); here the device is
torch's current HIP device and the library is loaded on first use.
"""
from ._internal import seed, get_seeds, device, backward_mode, backward_settings
from .maps import *  # noqa: F401,F403
from .maps import __all__ as _maps_all
from .scenes import *  # noqa: F401,F403
from .scenes import __all__ as _scenes_all
from . import functions
from . import objects
from . import scenes

__all__ = ["seed", "get_seeds", "device", "backward_mode", "backward_settings", "functions", "objects", "scenes",
           *_maps_all, *_scenes_all]
