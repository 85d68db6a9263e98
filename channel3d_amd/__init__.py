"""MI355X-native per-time-step core of the Channel3d channel-flow DNS solver.

The product is the HIP library libchannel3d.so (C ABI in include/channel3d.h);
this package builds it (``channel3d_amd.build``) and mirrors the reference's
routine interface on top of it (``channel3d_amd.channel.Channel``).  Importing
the package does not load the library; creating a ``Channel`` does, and fails
loudly when it has not been built.
"""
from . import synth  # noqa: F401
from .channel import Channel, channel_yp, make_params, mesh_coords  # noqa: F401
from .lib import (BC_FREESLIP, BC_NOSLIP, BC_PERIOD, FP32, FP64, HALO, INTERIOR,  # noqa: F401
                  WALL, C3DError)

__all__ = ["Channel", "channel_yp", "mesh_coords", "make_params", "synth", "C3DError"]
