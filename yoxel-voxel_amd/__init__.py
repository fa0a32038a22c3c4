"""MI355X-native sparse-voxel-octree ray caster (libyoxel over a C ABI).

The directory name is not a Python identifier; load it with load_package()
from __graft_entry__ or import it after registering it as `yoxel_voxel_amd`.
"""
from ._lib import (YX_EMPTY_NODE, YX_FULL_NODE, YX_LIGHT_HEADLIGHT, YX_LIGHT_SUN,
                   YX_LIGHT_SUN_SHADOW, YX_MISS, YX_PAYLOAD_C16N16, YX_PAYLOAD_RGBA8, LIB_PATH, YoxelError, lib)
from .renderer import SVOData, SVORenderer
from .batch import BatchRenderer
from .ntree import BrickTree
from . import workloads
from . import composite
from . import voxtree

__all__ = ["SVOData", "SVORenderer", "BatchRenderer", "BrickTree", "YoxelError", "lib", "workloads", "composite", "voxtree", "LIB_PATH",
           "YX_EMPTY_NODE", "YX_FULL_NODE", "YX_MISS", "YX_LIGHT_HEADLIGHT", "YX_LIGHT_SUN",
           "YX_LIGHT_SUN_SHADOW", "YX_PAYLOAD_RGBA8", "YX_PAYLOAD_C16N16"]
