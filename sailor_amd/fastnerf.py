"""The reference GUI's FastNerf inference steps on this library's kernels:
  forward_nerf        per-view sample features (sailor_sample_features:
                      proj_persp, z_normalizer, grid_sample x4, t-psdf,
                      camera-space ray directions) feeding InferNerf.infer_nerf,
                      for the rays whose samples are all valid
                      (gui/gui_render.py:713-751);
  rays_interpolate    the 4 sub-pixel rays of each ray (gui_render.py:753-773);
  forward_upsampling  sub-pixel rays, neighbour-view sampling and the blending
                      MLP in one kernel (gui_render.py:775-838).
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
                 depths_mid, depths_dist, params_density, params_color, res=512,
                 precision="fp16"):
    """gui_render.forward_nerf for one batch: keeps the rays whose samples
    are all valid, gathers their features and runs InferNerf.infer_nerf
    (`precision` as there).
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
                                   params_color, 1, valid.numel(), S, V, V, dev.index or 0,
                                   precision=precision)
    o_c[:, valid] = c
    o_d[:, valid] = d
    o_f[:, valid] = f
    return o_c, o_d, o_f


def rays_interpolate(ray_dirs, num_rays, tar_K, tar_R):
    """gui_render.rays_interpolate: ray_dirs [.., num_rays, 3] of the cameras
    tar_K, tar_R ([.., 3, 3], one per camera) -> [1, cams * num_rays * 4, 3]
    sub-pixel directions (ray-major, 4 per ray: +0 / +0.5 in x, then y)."""
    dev = ray_dirs.device
    K = as_input(tar_K, "tar_K", _f32, dev).reshape(-1, 3, 3)
    R = as_input(tar_R, "tar_R", _f32, dev).reshape(-1, 3, 3)
    n = int(num_rays)
    d = as_input(ray_dirs, "ray_dirs", _f32, dev).reshape(-1, n, 3)
    cams = d.shape[0]
    if K.shape[0] != cams or R.shape[0] != cams:
        raise RuntimeError(f"rays_interpolate: {cams} ray sets for {K.shape[0]} / {R.shape[0]} "
                           "cameras")
    out = torch.empty((1, cams * n * 4, 3), dtype=_f32, device=dev)
    call("sailor_rays_interpolate", ptr(K), ptr(R), ptr(d), cams, n, ptr(out), stream_of(dev))
    return out


def _neighbours(nei_ids):
    ids = nei_ids.reshape(-1)[:2].tolist() if isinstance(nei_ids, torch.Tensor) \
        else list(nei_ids)[:2]
    if len(ids) != 2:
        raise RuntimeError("forward_upsampling blends exactly 2 neighbour views")
    return int(ids[0]), int(ids[1])


def forward_upsampling(ray_ori, ray_dir, o_d, o_c, o_c_feat, tar_K, tar_R, ks, rs, ts,
                       rgb_feats_cl, rgbs_cl, depths, nei_ids, params_blend,
                       valid_depth_min=0.2, valid_depth_max=4.0, res=512, precision="fp16"):
    """gui_render.forward_upsampling + sample_neighbor for one target camera.
    ray_ori, ray_dir [1,R,3]; rendered o_d [1,R,1], o_c [1,R,3], o_c_feat
    [1,R,30]; tar_K, tar_R [3,3]; source views ks, rs [V,3,3], ts [V,3(,1)];
    channels-last high-resolution colour features rgb_feats_cl [V,Hf,Wf,16]
    and colours rgbs_cl [V,Hc,Wc,3]; depths [V,(1,)Hd,Wd]; nei_ids the 2
    neighbour view indices; params_blend the 15075 blending halves.
    Returns the sub-pixel colours [1,R,4,3] (zeros where o_d is outside
    (valid_depth_min, valid_depth_max)).  Rays are not compacted: the depth
    test runs in the kernel, so there is no host synchronisation.
    precision "fp16" (default): the blending MLP on fp16 MFMA (half inputs,
    fp32 accumulation; the reference evaluates it in half); "fp32": on the
    vector ALUs in fp32."""
    dev = o_d.device
    R = o_d.reshape(-1).numel()
    ori = as_input(ray_ori, "ray_ori", _f32, dev).reshape(R, 3)
    dirs = as_input(ray_dir, "ray_dir", _f32, dev).reshape(R, 3)
    od = as_input(o_d, "o_d", _f32, dev).reshape(R)
    oc = as_input(o_c, "o_c", _f32, dev).reshape(R, 3)
    of = as_input(o_c_feat, "o_c_feat", _f32, dev).reshape(R, 30)
    tK = as_input(tar_K, "tar_K", _f32, dev).reshape(3, 3)
    tR = as_input(tar_R, "tar_R", _f32, dev).reshape(3, 3)
    K = as_input(ks, "ks", _f32, dev).reshape(-1, 3, 3)
    V = K.shape[0]
    Rm = as_input(rs, "rs", _f32, dev).reshape(V, 3, 3)
    T = as_input(ts, "ts", _f32, dev).reshape(V, 3)
    rf = as_input(rgb_feats_cl, "rgb_feats_cl", _f32, dev)
    rc = as_input(rgbs_cl, "rgbs_cl", _f32, dev)
    dp = as_input(depths, "depths", _f32, dev)
    if rf.dim() != 4 or rf.shape[0] != V or rf.shape[-1] != 16:
        raise RuntimeError("rgb_feats_cl must be [V, H, W, 16]")
    if rc.dim() != 4 or rc.shape[0] != V or rc.shape[-1] != 3:
        raise RuntimeError("rgbs_cl must be [V, H, W, 3]")
    Hd, Wd = dp.shape[-2:]
    if dp.numel() != V * Hd * Wd:
        raise RuntimeError("depths must be [V, (1,) H, W]")
    if not isinstance(params_blend, torch.Tensor) or not params_blend.is_cuda \
            or params_blend.numel() < InferNerf.BLEND_PARAMS:
        raise RuntimeError("params_blend must be a cuda tensor of 15075 halves")
    pb = params_blend.to(device=dev, dtype=torch.float16).contiguous()
    n0, n1 = _neighbours(nei_ids)
    out = torch.empty((1, R, 4, 3), dtype=_f32, device=dev)
    if precision not in ("fp16", "fp32"):
        raise RuntimeError(f"precision must be 'fp16' or 'fp32', got {precision!r}")
    name = "sailor_forward_upsampling" + ("" if precision == "fp16" else "_fp32")
    call(name, ptr(ori), ptr(dirs), ptr(od), ptr(oc), ptr(of), R,
         ptr(tK), ptr(tR), ptr(K), ptr(Rm), ptr(T), V, n0, n1, ptr(rf), rf.shape[1], rf.shape[2],
         ptr(rc), rc.shape[1], rc.shape[2], ptr(dp), Hd, Wd, int(res), float(valid_depth_min),
         float(valid_depth_max), ptr(pb), ptr(out), stream_of(dev))
    return out
