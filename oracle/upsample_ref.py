"""PyTorch restatement of the GUI's sub-pixel upsampling -- TEST
INFRASTRUCTURE ONLY (imported by tests/).

Follows gui/gui_render.py step by step, in float64:
  rays_interpolate      :753-773  pixel of K R d, shifted by 0 / 0.5 in x and
                                  y, back-projected with torch.inverse and
                                  F.normalize
  forward_upsampling    :775-810  rays with valid_min < o_d < valid_max;
                                  surface = ori + dir_k * o_d; proj_persp
                                  with p_depths = o_d (utils_render.py:380-410)
  sample_neighbor       :812-838  grid_sample of the neighbour features
                                  (bilinear), colours and depths (nearest),
                                  exp(-200 (z - d)^2), the blending MLP
                                  (fastnerf_ref.infer_blending), clamp of
                                  w0 c0 + w1 c1 + w2 c_coarse to [0, 1]
F.grid_sample and torch.inverse are the library routines the reference
calls; the tests pin the sampling on closed forms (affine maps, pixel-index
maps) and the interpolated directions on reprojection identities.
"""
from __future__ import annotations

import torch
import torch.nn.functional as Fn

from . import fastnerf_ref as fr
from . import gather_ref as gr


def _t(a, dtype=torch.float64):
    return torch.as_tensor(a).to(dtype)


def rays_interpolate(ray_dirs, K, R):
    """ray_dirs [cams, n, 3], K, R [cams, 3, 3] -> [cams, n * 4, 3]."""
    K, R, d = _t(K), _t(R), _t(ray_dirs)
    cams, n, _ = d.shape
    pts = K @ (R @ d.transpose(1, 2))                          # [cams, 3, n]
    pts = pts / pts[:, 2:3].clone()
    x, y = pts[:, 0], pts[:, 1]
    px = torch.stack([x, x + 0.5, x, x + 0.5], -1).reshape(cams, n * 4)
    py = torch.stack([y, y, y + 0.5, y + 0.5], -1).reshape(cams, n * 4)
    h = torch.stack([px, py, torch.ones_like(px)], 1)          # [cams, 3, n*4]
    dirs = (torch.inverse(R) @ (torch.inverse(K) @ h)).transpose(1, 2)
    return Fn.normalize(dirs, dim=-1)


def forward_upsampling(ray_ori, ray_dir, o_d, o_c, o_c_feat, tar_K, tar_R, ks, rs, ts,
                       rgb_feats, rgbs, depths, nei, pb, valid_min=0.2, valid_max=4.0,
                       res=512, trace=None, act_dtype=None):
    """ray_ori, ray_dir [R,3]; o_d [R]; o_c [R,3]; o_c_feat [R,30]; tar_K,
    tar_R [3,3]; ks, rs [V,3,3]; ts [V,3]; channel-first maps rgb_feats
    [V,16,Hf,Wf], rgbs [V,3,Hc,Wc], depths [V,1,Hd,Wd]; nei (n0, n1); pb the
    blending halves.  Returns [R, 4, 3].  With a dict `trace`, stores the
    valid ray indices and, per valid ray and sub-pixel, the pixel
    coordinates of the nearest-mode taps (for tie detection), the 40 blend
    inputs, the neighbour colours, the blend weights, the directions and the
    neighbour-camera depths.  act_dtype rounds the blending MLP's layer
    inputs (fastnerf_ref.infer_blending)."""
    o_d = _t(o_d).reshape(-1)
    Rn = o_d.numel()
    out = torch.zeros((Rn, 4, 3), dtype=torch.float64)
    valid = torch.nonzero((o_d > valid_min) & (o_d < valid_max))[:, 0]
    if trace is not None:
        trace["valid"] = valid
    if valid.numel() == 0:
        return out
    n = valid.numel()
    de = rays_interpolate(_t(ray_dir)[valid][None], _t(tar_K)[None], _t(tar_R)[None])[0]
    ori = _t(ray_ori)[valid].repeat_interleave(4, 0)
    od = o_d[valid].repeat_interleave(4, 0)
    oc = _t(o_c)[valid].repeat_interleave(4, 0)
    of = _t(o_c_feat)[valid].repeat_interleave(4, 0)
    pos = ori + de * od[:, None]
    ks, rs, ts = _t(ks), _t(rs), _t(ts).reshape(-1, 3)
    grids, pz = gr.proj_persp(pos, ks, rs, ts, od, res=res)     # [V, 4n, 2], [V, 4n]
    feats, cols = [], []
    pix = []
    for v in nei:
        g = grids[v][None, :, None, :]                         # [1, 4n, 1, 2]
        samp = lambda m, mode: Fn.grid_sample(_t(m)[v][None], g, mode=mode, padding_mode="zeros",
                                             align_corners=True)[0, :, :, 0].T
        f = samp(rgb_feats, "bilinear")
        c = samp(rgbs, "nearest")
        sd = samp(depths, "nearest")[:, 0]
        sdir = de @ rs[v].T
        coff = torch.exp(-1.0 * 200 * (pz[v] - sd) ** 2)
        feats.append(torch.cat([f, sdir, coff[:, None]], -1))
        cols.append(c)
        for m in (rgbs, depths):
            H, W = _t(m).shape[-2:]
            pix.append(torch.stack([(grids[v][:, 0] + 1) / 2 * (W - 1),
                                    (grids[v][:, 1] + 1) / 2 * (H - 1)], -1))
    w = fr.infer_blending(torch.cat(feats, -1), of, pb, act_dtype=act_dtype)
    c = torch.clamp(w[:, 0:1] * cols[0] + w[:, 1:2] * cols[1] + w[:, 2:3] * oc, 0, 1)
    out[valid] = c.reshape(n, 4, 3)
    if trace is not None:
        trace["pixels"] = torch.stack(pix, 0).reshape(len(pix), n, 4, 2)
        trace["feats"] = torch.cat(feats, -1).reshape(n, 4, 40)
        trace["cols"] = torch.stack(cols, 0).reshape(2, n, 4, 3)
        trace["weights"] = w.reshape(n, 4, 3)
        trace["dirs"] = de.reshape(n, 4, 3)
        trace["proj_z"] = pz[list(nei)].reshape(2, n, 4)
    return out
