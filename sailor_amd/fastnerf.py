"""The reference GUI's FastNerf inference step on this library's kernels:
per-view sample features (sailor_sample_features: proj_persp, z_normalizer,
grid_sample x4, t-psdf, camera-space ray directions) feeding
InferNerf.infer_nerf, for the rays whose samples are all valid
(gui/gui_render.py:713-751, forward_nerf).
"""
from __future__ import annotations

import torch

from . import InferNerf
from ._lib import as_input, call, ptr, stream_of

_f32 = torch.float32


def to_channels_last(x):
    """[N, C, H, W] float32 (cuda) -> [N, H, W, C] with the HIP transpose."""
    x = as_input(x, "map", _f32)
    N, C, H, W = x.shape
    out = torch.empty((N, H, W, C), dtype=_f32, device=x.device)
    call("sailor_nchw_to_nhwc", ptr(x), N, C, H, W, ptr(out), stream_of(x.device))
    return out


def sample_features(rays, dists, ks, rts, geo_cl, depths, rgbf_cl, rgb_cl, depths_mid,
                    depths_dist, res=512):
    """rays [B,R,6] (origin, direction), dists [B,R,S]; ks [B,V,3,3],
    rts [B,V,3,4]; channels-last maps geo_cl [B,V,Hg,Wg,16], rgbf_cl
    [B,V,Hr,Wr,16], rgb_cl [B,V,Hc,Wc,3]; depths [B,V,Hd,Wd]; depths_mid /
    depths_dist [B,V].  Returns (z_normalized, t_psdf [B,V,R,S], geo feats,
    rgb feats [B,V,R,S,16], rgbs, camera-space directions [B,V,R,S,3])."""
    dev = dists.device
    B, R, S = dists.shape
    V = ks.shape[1]
    rays = as_input(rays, "rays", _f32, dev)
    dists = as_input(dists, "dists", _f32, dev)
    ks = as_input(ks, "ks", _f32, dev)
    rts = as_input(rts, "rts", _f32, dev)
    geo_cl = as_input(geo_cl, "geo_cl", _f32, dev)
    depths = as_input(depths, "depths", _f32, dev)
    rgbf_cl = as_input(rgbf_cl, "rgbf_cl", _f32, dev)
    rgb_cl = as_input(rgb_cl, "rgb_cl", _f32, dev)
    mid = as_input(depths_mid, "depths_mid", _f32, dev).reshape(-1)
    dst = as_input(depths_dist, "depths_dist", _f32, dev).reshape(-1)
    if mid.numel() != B * V or dst.numel() != B * V:
        raise RuntimeError("depths_mid / depths_dist must hold one value per (batch, view)")
    o = dict(device=dev, dtype=_f32)
    z = torch.empty((B, V, R, S), **o)
    tps = torch.empty((B, V, R, S), **o)
    geo = torch.empty((B, V, R, S, 16), **o)
    rgbf = torch.empty((B, V, R, S, 16), **o)
    rgb = torch.empty((B, V, R, S, 3), **o)
    dirs = torch.empty((B, V, R, S, 3), **o)
    Hg, Wg = geo_cl.shape[-3:-1]
    Hd, Wd = depths.shape[-2:]
    Hr, Wr = rgbf_cl.shape[-3:-1]
    Hc, Wc = rgb_cl.shape[-3:-1]
    call("sailor_sample_features", ptr(rays), ptr(dists), B, R, S, ptr(ks), ptr(rts), V,
         ptr(geo_cl), Hg, Wg, ptr(depths), Hd, Wd, ptr(rgbf_cl), Hr, Wr, ptr(rgb_cl), Hc, Wc,
         ptr(mid), ptr(dst), int(res), ptr(z), ptr(tps), ptr(geo), ptr(rgbf), ptr(rgb),
         ptr(dirs), stream_of(dev))
    return z, tps, geo, rgbf, rgb, dirs


def forward_nerf(ray_ori, ray_dir, sorted_dists, ks, rts, geo_cl, depths, rgbf_cl, rgb_cl,
                 depths_mid, depths_dist, params_density, params_color, res=512):
    """gui_render.forward_nerf for one batch: keeps the rays whose samples
    are all valid, gathers their features and runs InferNerf.infer_nerf.
    Returns (colour [1,R,3], depth [1,R,1], colour features [1,R,30]) with
    zeros for the other rays."""
    dev = sorted_dists.device
    R, S = sorted_dists.shape[1], sorted_dists.shape[2]
    V = ks.shape[1]
    o_c = torch.zeros((1, R, 3), dtype=_f32, device=dev)
    o_d = torch.zeros((1, R, 1), dtype=_f32, device=dev)
    o_f = torch.zeros((1, R, 30), dtype=_f32, device=dev)
    valid = torch.where(torch.prod(sorted_dists != -1, dim=-1)[0])[0]
    if valid.numel() == 0:
        return o_c, o_d, o_f
    rays = torch.cat([ray_ori[:, valid], ray_dir[:, valid]], -1)
    dists = sorted_dists[:, valid].contiguous()
    z, tps, geo, rgbf, rgb, dirs = sample_features(rays, dists, ks, rts, geo_cl, depths, rgbf_cl,
                                                   rgb_cl, depths_mid, depths_dist, res)
    d, c, f = InferNerf.infer_nerf(dists, z, tps, geo, rgb, rgbf, dirs, params_density,
                                   params_color, 1, valid.numel(), S, V, V, dev.index or 0)
    o_c[:, valid] = c
    o_d[:, valid] = d
    o_f[:, valid] = f
    return o_c, o_d, o_f
