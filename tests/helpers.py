"""Shared inputs for the parity tests (seeded synthetic capsule scenes)."""
from __future__ import annotations

import functools

import numpy as np

from sailor_amd import scenes


@functools.lru_cache(maxsize=8)
def scene_inputs(frame_id=0, n_views=4, H=256, W=256, fx=300.0):
    """Scene as the reference's integrate() expects it: [1,V,...] arrays."""
    sc = scenes.make_scene(frame_id, n_views, H, W, fx)
    K = sc.K()[None]
    RT = sc.RT()[None]
    R = np.ascontiguousarray(RT[..., :3])
    T = np.ascontiguousarray(RT[..., 3])
    colors = scenes.pack_colors(sc.rgb)[None]
    depths = sc.depth[None]
    masks = sc.mask.astype(np.int32)[None]
    return dict(scene=sc, K=K, RT=RT, R=R, T=T, colors=colors, depths=depths, masks=masks)


def target_rays_np(H, W, fx, phase=np.pi / 4):
    """Target-view camera (K, R, T as [1,1,...])."""
    cam = scenes.target_camera(H, W, fx, phase)
    return cam.K[None, None], np.ascontiguousarray(cam.R)[None, None], \
        np.ascontiguousarray(cam.T)[None, None]


def max_abs(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b))) if a.size else 0.0
