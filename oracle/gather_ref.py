"""PyTorch restatement of the per-view sample features fed to FastNerf --
TEST INFRASTRUCTURE ONLY (imported by tests/).

Follows gui/gui_render.py:713-751 (forward_nerf) and
SRONet/BasicRenNet.py:118-131,245-259,296-311 step by step:
  rays_points = ray_ori + ray_dir * dists
  proj_persp(points, K, R, T, p_depths=dists)      utils_render.py:380-410
  z_normalizer(proj_z, mid, dist)                   utils_render.py:656-663
  F.grid_sample(align_corners=True, bilinear)       geo feats, depths, rgb feats, rgbs
  t_psdf = clamp(proj_z - sampled_depth, -0.01, 0.01) * (1 / 0.01)
  ray_dir_cam = R @ ray_dir (repeated over samples)
F.grid_sample is used as the library routine the reference calls; the tests
pin it on pixel centres and planes (closed forms).
"""
from __future__ import annotations

import torch
import torch.nn.functional as Fn


def proj_persp(points, K, R, T, p_depths, depth_thres=0.01, min_d=0.1, max_d=4.0, res=512):
    """points [N, 3] (one batch), K/R [V,3,3], T [V,3], p_depths [N] ->
    (grids [V, N, 2], depth [V, N])."""
    invalid_d = (p_depths < min_d) | (p_depths > max_d)
    pts = points[None] @ R.transpose(1, 2) + T[:, None, :]          # [V, N, 3]
    depth = pts[..., 2].clone()
    invalid = depth < depth_thres
    depth_ = torch.where(invalid, torch.ones_like(depth), depth)
    depth = torch.where(invalid_d[None].expand_as(depth), torch.full_like(depth, -2.0), depth)
    q = pts / depth_[..., None]
    scr = (q @ K.transpose(1, 2))[..., :2]
    half = float(res // 2)
    grids = (scr - half) / half
    bad = (invalid | invalid_d[None])[..., None].expand_as(grids)
    grids = torch.where(bad, torch.full_like(grids, -2.0), grids)
    return grids, depth


def sample_features(rays, dists, K, RT, geo, depth, rgbf, rgb, mid, dist, res=512,
                    dtype=torch.float64):
    """rays [B,R,6], dists [B,R,S], K [B,V,3,3], RT [B,V,3,4]; maps
    channel-first [B,V,C,H,W] (geo C=16, depth C=1, rgbf C=16, rgb C=3);
    mid/dist [B,V].  Returns dict of z, tpsdf [B,V,R,S], geo, rgbf [..,16],
    rgb, dir [..,3]."""
    T_ = lambda a: torch.as_tensor(a).to(dtype)
    rays, dists, K, RT = map(T_, (rays, dists, K, RT))
    geo, depth, rgbf, rgb, mid, dist = map(T_, (geo, depth, rgbf, rgb, mid, dist))
    B, R, S = dists.shape
    V = K.shape[1]
    out = {k: [] for k in ("z", "tpsdf", "geo", "rgbf", "rgb", "dir")}
    for b in range(B):
        pts = rays[b, :, None, :3] + rays[b, :, None, 3:] * dists[b][..., None]   # [R, S, 3]
        grids, pz = proj_persp(pts.reshape(-1, 3), K[b], RT[b, :, :, :3], RT[b, :, :, 3],
                               dists[b].reshape(-1), res=res)
        zn = (pz - mid[b][:, None]) / (dist[b][:, None] / 2.0)
        g = grids[:, :, None, :]                                     # [V, N, 1, 2]

        def samp(m):
            return Fn.grid_sample(m, g, mode="bilinear", padding_mode="zeros",
                                  align_corners=True)[..., 0].transpose(1, 2)  # [V, N, C]
        sd = samp(depth[b])[..., 0]
        tps = torch.clamp(pz - sd, -0.01, 0.01) * (1.0 / 0.01)
        d_cam = rays[b, :, 3:] @ RT[b, :, :, :3].transpose(1, 2)     # [V, R, 3]
        out["z"].append(zn.reshape(V, R, S))
        out["tpsdf"].append(tps.reshape(V, R, S))
        out["geo"].append(samp(geo[b]).reshape(V, R, S, -1))
        out["rgbf"].append(samp(rgbf[b]).reshape(V, R, S, -1))
        out["rgb"].append(samp(rgb[b]).reshape(V, R, S, -1))
        out["dir"].append(d_cam[:, :, None, :].expand(V, R, S, 3))
    return {k: torch.stack(v) for k, v in out.items()}
