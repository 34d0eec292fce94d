"""Drop-in for the reference's `InferNerf` extension module
(c_lib/FastNerf/src/binding.cpp, imported as
`from c_lib.FastNerf.dist import InferNerf` by gui/gui_render.py:60-80).

Same function names, positional arguments (including the integer `device`
ordinal) and return shapes; the parameter vectors are the reference's packed
half tensors (gui/gui_render.py:658-685: 22521 density, 15807 colour, 15075
blending halves).  Each call runs the HIP kernels of libsailor_hip.so on
torch's current stream with no host synchronisation (the reference creates
and destroys streams and calls cudaDeviceSynchronize per call,
infer_mlp.cu:677-680,771,783-785).

Numerics (`precision`, a keyword the reference does not have):
  "fp16" (default)  the density MLPs on fp16 MFMA -- layer inputs rounded to
                    half as in the reference's half arithmetic, products
                    accumulated in fp32 -- everything else in fp32;
  "fp32"            every activation in fp32 on the vector ALUs: the exact
                    value of the half parameters to fp32 rounding.
"""
from __future__ import annotations

import torch

from . import _lib
from ._lib import as_input, call, ptr, stream_of

_f32 = torch.float32
DENSITY_PARAMS, COLOR_PARAMS, BLEND_PARAMS = 22521, 15807, 15075


def _dev(device):
    return torch.device("cuda", int(device))


def _entry(name, precision):
    if precision not in ("fp16", "fp32"):
        raise RuntimeError(f"precision must be 'fp16' or 'fp32', got {precision!r}")
    return name if precision == "fp16" else name + "_fp32"


def _params(p, n, name, dev):
    if not isinstance(p, torch.Tensor) or not p.is_cuda:
        raise RuntimeError(f"{name} must be a cuda tensor")
    if p.numel() < n:
        raise RuntimeError(f"{name} holds {p.numel()} parameters, {n} needed")
    return p.to(device=dev, dtype=torch.float16).contiguous()


def infer_nerf(sorted_dists, points_z, points_tpsdf, sampled_feats_geo, sampled_rgbs,
               sampled_feats_rgb, rays_dir, net_params_density, net_params_color, batch_size,
               n_rays, n_sampled, geo_num_views, tex_num_views, device, precision="fp16"):
    """infer_fast_rendering (infer_mlp.cu:638-789).
    sorted_dists [B, n_rays, n_sampled]; per-view inputs flattened in (batch,
    view, ray, sample) order: points_z, points_tpsdf [B*Vg*R*S(, 1)],
    sampled_feats_geo [.., 16]; sampled_rgbs, rays_dir [B*Vt*R*S, 3],
    sampled_feats_rgb [.., 16].
    Returns (depth [B, n_rays, 1], rgb [B, n_rays, 3], colour features
    [B, n_rays, 30])."""
    dev = _dev(device)
    B, R, S = int(batch_size), int(n_rays), int(n_sampled)
    Vg, Vt = int(geo_num_views), int(tex_num_views)
    t = as_input(sorted_dists, "sorted_dists", _f32, dev)
    z = as_input(points_z, "points_z", _f32, dev)
    ps = as_input(points_tpsdf, "points_tpsdf", _f32, dev)
    geo = as_input(sampled_feats_geo, "sampled_feats_geo", _f32, dev)
    rgb = as_input(sampled_rgbs, "sampled_rgbs", _f32, dev)
    frgb = as_input(sampled_feats_rgb, "sampled_feats_rgb", _f32, dev)
    dirs = as_input(rays_dir, "rays_dir", _f32, dev)
    n_geo, n_tex = B * Vg * R * S, B * Vt * R * S
    for x, n, c, name in ((t, B * R * S, 1, "sorted_dists"), (z, n_geo, 1, "points_z"),
                          (ps, n_geo, 1, "points_tpsdf"), (geo, n_geo, 16, "sampled_feats_geo"),
                          (rgb, n_tex, 3, "sampled_rgbs"), (frgb, n_tex, 16, "sampled_feats_rgb"),
                          (dirs, n_tex, 3, "rays_dir")):
        if x.numel() != n * c:
            raise RuntimeError(f"{name} has {x.numel()} values, expected {n * c}")
    pd = _params(net_params_density, DENSITY_PARAMS, "net_params_density", dev)
    pc = _params(net_params_color, COLOR_PARAMS, "net_params_color", dev)
    depth = torch.empty((B, R, 1), dtype=_f32, device=dev)
    color = torch.empty((B, R, 3), dtype=_f32, device=dev)
    feats = torch.empty((B, R, 30), dtype=_f32, device=dev)
    wsb = _lib.lib().sailor_infer_nerf_workspace_bytes(B, R, S, Vt)
    ws = torch.empty(max(wsb, 1), dtype=torch.uint8, device=dev)
    call(_entry("sailor_infer_nerf", precision), ptr(t), ptr(z), ptr(ps), ptr(geo), ptr(rgb),
         ptr(frgb), ptr(dirs),
         ptr(pd), ptr(pc), B, R, S, Vg, Vt, ptr(depth), ptr(color), ptr(feats), ptr(ws),
         ws.numel(), stream_of(dev))
    return depth, color, feats


def infer_density(points_z, points_tpsdf, sampled_feats_geo, net_params_density, batch_size,
                  num_points, num_views, device, precision="fp16"):
    """infer_density_mlp (infer_mlp.cu:583-634): occupancy [B, 1, N] of points
    seen from `num_views` views (inputs in (batch, view, point) order)."""
    dev = _dev(device)
    B, N, V = int(batch_size), int(num_points), int(num_views)
    z = as_input(points_z, "points_z", _f32, dev)
    ps = as_input(points_tpsdf, "points_tpsdf", _f32, dev)
    geo = as_input(sampled_feats_geo, "sampled_feats_geo", _f32, dev)
    for x, c, name in ((z, 1, "points_z"), (ps, 1, "points_tpsdf"), (geo, 16, "sampled_feats_geo")):
        if x.numel() != B * V * N * c:
            raise RuntimeError(f"{name} has {x.numel()} values, expected {B * V * N * c}")
    pd = _params(net_params_density, DENSITY_PARAMS, "net_params_density", dev)
    out = torch.empty((B, 1, N), dtype=_f32, device=dev)
    call(_entry("sailor_infer_density", precision), ptr(z), ptr(ps), ptr(geo), ptr(pd), B, N, V,
         ptr(out),
         stream_of(dev))
    return out


def infer_blending(feats_neighbors, ray_rgb_feats, net_params, device):
    """infer_upsampling (infer_mlp.cu:793-821): [N, 40] neighbour features and
    [N, 30] ray colour features -> [N, 3] blend weights."""
    dev = _dev(device)
    nf = as_input(feats_neighbors, "feats_neighbors", _f32, dev)
    rf = as_input(ray_rgb_feats, "ray_rgb_feats", _f32, dev)
    N = nf.shape[0]
    if nf.numel() != N * 40 or rf.numel() != N * 30:
        raise RuntimeError("feats_neighbors must be [N, 40] and ray_rgb_feats [N, 30]")
    pb = _params(net_params, BLEND_PARAMS, "net_params", dev)
    out = torch.empty((N, 3), dtype=_f32, device=dev)
    call("sailor_infer_blending", ptr(nf), ptr(rf), ptr(pb), N, ptr(out), stream_of(dev))
    return out
