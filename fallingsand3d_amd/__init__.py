"""fallingsand3d_amd -- MI355X-native (gfx950) falling-sand voxel tick.

The compute path is libfs3d.so (hand-written HIP kernels behind the C ABI in
include/fs3d.h); this package is its Python host mirror.  There is no CPU
fallback: without the built library or a GPU every compute call raises Fs3dError.
"""
from ._lib import Fs3dError, LIB_PATH, lib  # noqa: F401
from .world import (DENSE, SKIP_AUTO, SKIP_LISTED, SKIP_OFF, SKIP_SEGS, SKIP_WAVES, TRACK_AUTO, TRACK_OFF, TRACK_ON, EMPTY, FLUID, SAND, STATIC, BOXES_512, PILLARS_2048, World,  # noqa: F401
                    batch_specs, copy_probe, rowquad_probe, scatter_probe, device_count, file_info, file_read, file_write, init, rule_hash,
                    rule_lut, serve_specs, threshold)
from .batch import Batch  # noqa: F401,E402
from .serve import Server  # noqa: F401,E402
from . import render  # noqa: F401,E402
