"""sailor_amd: MI355X (gfx950) implementation of SAILOR's VoxelEncoding +
FastNerf rendering path.

Modules mirroring the reference's extension surface:
  sailor_amd.VoxelEncoding          -- c_lib/VoxelEncoding (pybind module)
  sailor_amd.voxel_encoding_helper  -- c_lib/VoxelEncoding/voxel_encoding_helper.py
  sailor_amd.freq_encoding          -- c_lib/VoxelEncoding/freq_encoding.py
  sailor_amd.InferNerf              -- c_lib/FastNerf (pybind module)
Everything runs through libsailor_hip.so (include/sailor_hip.h); importing a
kernel module on a machine where that library is missing raises ImportError.
"""
from . import _lib  # noqa: F401

__all__ = ["VoxelEncoding", "voxel_encoding_helper", "freq_encoding", "InferNerf", "scenes"]
