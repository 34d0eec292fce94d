"""Fused radiance evaluation (sailor_render_rays) and the fused trace+sample
path (sailor_trace_sample) as Python calls on torch tensors."""
from __future__ import annotations

import torch

from . import _lib
from ._lib import as_input, call, ptr, stream_of
from .model import PREC_IDS

_f32, _i32 = torch.float32, torch.int32


def trace_sample(K, R, T, img_hw, vol_coarse, node_pos, node_info, vol_bbox, voxel_size,
                 dim_fine, start_scale, noise, out=None, want_rays=True):
    """Rays of a target camera per batch -> DDA -> samples (unsorted, -1 padded).
    K, R [B,3,3], T [B,3]; noise [B, H*W, S].  Returns (rays [B,HW,6] or None,
    sampled_idx [B,HW,S], sampled_t [B,HW,S], valid_len [B,HW])."""
    dev = vol_coarse.device
    Hh, Ww = img_hw
    B = vol_coarse.shape[0]
    dc = vol_coarse.shape[1]
    S = noise.shape[-1]
    n = Hh * Ww
    if out is None:
        out = dict(rays=torch.empty((B, n, 6), dtype=_f32, device=dev) if want_rays else None,
                   sidx=torch.empty((B, n, S), dtype=_i32, device=dev),
                   st=torch.empty((B, n, S), dtype=_f32, device=dev),
                   vl=torch.empty((B, n), dtype=_f32, device=dev))
    call("sailor_trace_sample", ptr(K), ptr(R), ptr(T), B, Hh, Ww, ptr(vol_coarse),
         ptr(node_pos), ptr(node_info), node_pos.shape[0], ptr(vol_bbox), ptr(voxel_size),
         int(dim_fine), int(dc), int(start_scale), S, ptr(noise), ptr(out["rays"]),
         ptr(out["sidx"]), ptr(out["st"]), ptr(out["vl"]), stream_of(dev))
    return out["rays"], out["sidx"], out["st"], out["vl"]


class PackedModel:
    """Device-resident packed weights of a NerfWeights for one precision."""

    def __init__(self, weights, prec="fp32", device="cuda"):
        self.F, self.H, self.prec = weights.F, weights.H, prec
        self.w_dense, self.b_dense, self.w_small = weights.pack(prec, device)
        want = _lib.lib().sailor_render_small_weights_floats(self.F, self.H)
        if self.w_small.numel() != want:
            raise RuntimeError(f"small-weight buffer has {self.w_small.numel()} floats, "
                               f"kernel layout expects {want}")


def render_rays(t, rays, Kv, RTv, img_hw, fmap_cl, bbox, zdir, model: PackedModel,
                t_thres=1e-4, early_exit=False, want_occ=False, out=None, chunk_samples=0):
    """t [B,R,S] sorted; rays [B,R,6]; Kv [B,V,3,3]; RTv [B,V,3,4];
    fmap_cl [B,V,Hf,Wf,F] channels-last, float32 or the model's 16-bit dtype;
    bbox [B,3,2]; zdir [B,3].
    Returns dict(rgb [B,R,3], depth [B,R], aux [B,R,2], occ [B,R,S] or None)."""
    dev = t.device
    B, R, S = t.shape
    V = Kv.shape[1]
    Hf, Wf, F = fmap_cl.shape[2], fmap_cl.shape[3], fmap_cl.shape[4]
    if F != model.F:
        raise RuntimeError(f"feature maps have {F} channels, model expects {model.F}")
    map_prec = {torch.float32: 0, torch.bfloat16: PREC_IDS["bf16"],
                torch.float16: PREC_IDS["fp16"]}.get(fmap_cl.dtype)
    if map_prec is None:
        raise RuntimeError(f"feature maps of dtype {fmap_cl.dtype} are not supported")
    if out is None:
        out = dict(rgb=torch.empty((B, R, 3), dtype=_f32, device=dev),
                   depth=torch.empty((B, R), dtype=_f32, device=dev),
                   aux=torch.empty((B, R, 2), dtype=_f32, device=dev),
                   occ=torch.zeros((B, R, S), dtype=_f32, device=dev) if want_occ else None)
    ws_bytes = _lib.lib().sailor_render_workspace_bytes_views(B, R, F, V)
    if out.get("workspace") is None or out["workspace"].numel() < ws_bytes:
        out["workspace"] = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
    call("sailor_render_rays", ptr(t), ptr(rays), B, R, S, ptr(Kv), ptr(RTv), V, int(img_hw[0]),
         int(img_hw[1]), ptr(fmap_cl), map_prec, Hf, Wf, F, ptr(bbox), ptr(zdir), model.H,
         PREC_IDS[model.prec], ptr(model.w_dense), ptr(model.b_dense), ptr(model.w_small),
         float(t_thres), int(bool(early_exit)), int(chunk_samples), ptr(out["rgb"]),
         ptr(out["depth"]),
         ptr(out["aux"]), ptr(out["occ"]), ptr(out["workspace"]),
         out["workspace"].numel(), stream_of(dev))
    return out


def density_points(points, Kv, RTv, img_hw, fmap_cl, bbox, model: PackedModel, out=None,
                   chunk_samples=0):
    """Occupancy of the fused model at points [B,N,3] (infer_density path):
    per-view MLP on gathered features + encoding, view mean, occupancy head.
    Returns [B,N] float32."""
    B, N = points.shape[0], points.shape[1]
    V = Kv.shape[1]
    Hf, Wf, F = fmap_cl.shape[2], fmap_cl.shape[3], fmap_cl.shape[4]
    map_prec = {torch.float32: 0, torch.bfloat16: PREC_IDS["bf16"],
                torch.float16: PREC_IDS["fp16"]}.get(fmap_cl.dtype)
    if map_prec is None:
        raise RuntimeError(f"feature maps of dtype {fmap_cl.dtype} are not supported")
    if out is None:
        out = torch.empty((B, N), dtype=_f32, device=points.device)
    call("sailor_density_points", ptr(points), B, N, ptr(Kv), ptr(RTv), V, int(img_hw[0]),
         int(img_hw[1]), ptr(fmap_cl), map_prec, Hf, Wf, F, ptr(bbox), model.H,
         PREC_IDS[model.prec], ptr(model.w_dense), ptr(model.b_dense), ptr(model.w_small),
         int(chunk_samples), ptr(out), stream_of(points.device))
    return out
