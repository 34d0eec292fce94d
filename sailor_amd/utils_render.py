"""Drop-ins for the preprocessing helpers of utils_render/utils_render.py
that sit on the frame path, on this library's HIP kernels.  (proj_persp,
z_normalizer and the grid samples that follow them are fused into
fastnerf.sample_features and fastnerf.forward_upsampling.)"""
from __future__ import annotations

import torch

from ._lib import as_input, call, ptr, stream_of

_f32 = torch.float32


def rgb_normalizer(rgb_tensors, mask_tensors, mid=0.5, dist=0.5):
    """utils_render.py:601-605: ((rgb - mid) / dist) * mask with rgb
    [B,C,H,W] and mask [B,1,H,W]."""
    dev = rgb_tensors.device
    rgb = as_input(rgb_tensors, "rgb_tensors", _f32, dev)
    mask = as_input(mask_tensors, "mask_tensors", _f32, dev)
    N, C, H, W = rgb.shape
    if mask.numel() != N * H * W:
        raise RuntimeError("mask_tensors must be [B, 1, H, W] for rgb [B, C, H, W]")
    out = torch.empty_like(rgb)
    call("sailor_rgb_normalize", ptr(rgb), ptr(mask), N, C, H, W, float(mid), float(dist),
         ptr(out), stream_of(dev))
    return out
