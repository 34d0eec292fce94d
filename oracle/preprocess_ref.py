"""numpy restatements of the preprocessing kernels -- TEST INFRASTRUCTURE
ONLY (imported by tests/).

depth_normalize: c_lib/VoxelEncoding/src/depth_normalizer.cu:98-189 (float32,
the reference's operation order; torch's tensor + scalar evaluates in the
tensor's float32).  Pinned in tests/test_preprocess_oracle.py by closed forms.
"""
import numpy as np


def depth_normalize(d, lo, hi, user_dist, multi_mask, divided2):
    """d [B, V, H, W] -> (normalized, mask, mid [B,V,1,1], distance [B,V,1,1])."""
    d = np.asarray(d, np.float32)
    valid = (d >= np.float32(lo)) & (d <= np.float32(hi))
    mn = np.where(valid, d, np.float32(99999.0)).min(axis=(2, 3), keepdims=True)
    mx = np.where(valid, d, np.float32(-99999.0)).max(axis=(2, 3), keepdims=True)
    dist = (mx - mn) + np.float32(1e-10)
    mid = (mx + mn) / np.float32(2)
    mask = np.where(valid, 1, 0).astype(np.float32) if multi_mask else np.ones_like(d)
    if user_dist < 0:
        r = ((d - mid) * mask) / dist if multi_mask else (d - mid) / dist
        if divided2:
            r = r * np.float32(2.0)
        r = np.clip(r, -1, 1)
    elif multi_mask:
        r = ((d - mid) * mask) / np.float32(user_dist)
    else:
        r = np.clip((d - mid) / np.float32(user_dist), -1, 1)
    return r.astype(np.float32), mask, mid, dist


def rgb_normalizer(rgb, mask, mid=0.5, dist=0.5):
    """utils_render.py:601-605 in float32: ((rgb - mid) / dist) * mask, the
    [N,1,H,W] mask broadcast over channels."""
    rgb = np.asarray(rgb, np.float32)
    m = np.asarray(mask, np.float32)
    t = (rgb - np.float32(mid)) / np.float32(dist)
    return (t * m).astype(np.float32)
