"""Fused radiance evaluation (sailor_render_rays) and the fused trace+sample
path (sailor_trace_sample) as Python calls on torch tensors."""
from __future__ import annotations

import torch

from . import _lib
from ._lib import call, check_input, ptr, stream_of
from .model import PREC_IDS

_f32, _i32 = torch.float32, torch.int32


def _check(t, name, dtype, shape, device):
    """Input validation before any launch: a CUDA tensor of `dtype` on
    `device`, contiguous, of exactly `shape`."""
    check_input(t, name, dtype, device)
    if tuple(t.shape) != tuple(shape):
        raise RuntimeError(f"{name} must have shape {list(shape)}, got {list(t.shape)}")
    return t


def _check_out(out, shapes, device):
    """A reusable output dict matches the call: every listed entry present
    (or None where allowed) with the expected shape, float32/int32 dtype and
    device."""
    for key, (shape, dtype, required) in shapes.items():
        v = out.get(key)
        if v is None:
            if required:
                raise RuntimeError(f"out[{key!r}] is missing")
            continue
        _check(v, f"out[{key!r}]", dtype, shape, device)


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
    _check(vol_coarse, "vol_coarse", torch.int32, (B, dc, dc, dc), dev)
    _check(K, "K", _f32, (B, 3, 3), dev)
    _check(R, "R", _f32, (B, 3, 3), dev)
    _check(T, "T", _f32, (B, 3), dev)
    _check(noise, "noise", _f32, (B, n, S), dev)
    _check(vol_bbox, "vol_bbox", _f32, (B, 3, 2), dev)
    check_input(voxel_size, "voxel_size", _f32, dev)
    if voxel_size.numel() != B:
        raise RuntimeError(f"voxel_size must hold {B} values")
    check_input(node_pos, "node_pos", _f32, dev)
    if node_pos.dim() != 2 or node_pos.shape[1] != 4:
        raise RuntimeError("node_pos must be [n_nodes, 4]")
    _check(node_info, "node_info", torch.int32, tuple(node_pos.shape), dev)
    if out is not None:
        _check_out(out, dict(rays=((B, n, 6), _f32, False), sidx=((B, n, S), torch.int32, True),
                             st=((B, n, S), _f32, True), vl=((B, n), _f32, True)), dev)
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
        self.weights, self.device = weights, device
        self.w_dense, self.b_dense, self.w_small = weights.pack(prec, device)
        want = _lib.lib().sailor_render_small_weights_floats(self.F, self.H)
        if self.w_small.numel() != want:
            raise RuntimeError(f"small-weight buffer has {self.w_small.numel()} floats, "
                               f"kernel layout expects {want}")
        self._first = None

    def first_layer(self):
        """(packed first layer in fp32 [H * (F + 64)], int16 input index of
        every packed slot) for folding per-channel feature scales into it
        (sailor_scale_packed_layer); the first layer is the head of w_dense."""
        if self._first is None:
            import numpy as np
            from .model import kmap, natural_first_layer, pack_layer, unit_rows
            K0 = self.F + 64
            w0 = pack_layer(natural_first_layer(self.weights.dense_w[0], self.F), self.prec)
            rows, cols = np.broadcast_arrays(unit_rows(self.prec, self.H),
                                             kmap(self.prec, K0)[None, :, :, None, :])
            self._first = (torch.from_numpy(np.ascontiguousarray(w0, np.float32).reshape(-1))
                           .to(self.device),
                           torch.from_numpy(cols.reshape(-1).astype(np.int16)).to(self.device))
        return self._first


def render_rays(t, rays, Kv, RTv, img_hw, fmap_cl, bbox, zdir, model: PackedModel,
                t_thres=1e-4, early_exit=False, want_occ=False, out=None, chunk_samples=0,
                compensated=False, feature_scales=None, fp32_hydra=False):
    """t [B,R,S] sorted; rays [B,R,6]; Kv [B,V,3,3]; RTv [B,V,3,4];
    fmap_cl [B,V,Hf,Wf,F] channels-last, float32 or the model's 16-bit dtype;
    bbox [B,3,2]; zdir [B,3].  compensated: Kahan-compensated compositing
    sums (precision form).  feature_scales [F]: fmap_cl holds maps quantised
    by sailor_quantize_channels with these scales, and the model's first
    layer carries them (FramePipeline with quantize=True).  fp32_hydra: the
    16-bit forms run the Hydra encoder and RGB head on fp32 MFMA instead of
    16-bit MFMA (the differentiation form, whose backward recomputes them in
    fp32).
    Returns dict(rgb [B,R,3], depth [B,R], aux [B,R,2], occ [B,R,S] or None)."""
    dev = t.device
    check_input(t, "t", _f32)
    if t.dim() != 3:
        raise RuntimeError(f"t must be [B, R, S], got {list(t.shape)}")
    B, R, S = t.shape
    check_input(Kv, "Kv", _f32, dev)
    if Kv.dim() != 4 or Kv.shape[0] != B or tuple(Kv.shape[2:]) != (3, 3):
        raise RuntimeError(f"Kv must be [{B}, V, 3, 3], got {list(Kv.shape)}")
    V = Kv.shape[1]
    _check(rays, "rays", _f32, (B, R, 6), dev)
    _check(RTv, "RTv", _f32, (B, V, 3, 4), dev)
    _check(bbox, "bbox", _f32, (B, 3, 2), dev)
    _check(zdir, "zdir", _f32, (B, 3), dev)
    if fmap_cl.dim() != 5 or fmap_cl.shape[0] != B or fmap_cl.shape[1] != V:
        raise RuntimeError(f"fmap_cl must be [{B}, {V}, Hf, Wf, F], got {list(fmap_cl.shape)}")
    Hf, Wf, F = fmap_cl.shape[2], fmap_cl.shape[3], fmap_cl.shape[4]
    if F != model.F:
        raise RuntimeError(f"feature maps have {F} channels, model expects {model.F}")
    map_prec = {torch.float32: 0, torch.bfloat16: PREC_IDS["bf16"],
                torch.float16: PREC_IDS["fp16"]}.get(fmap_cl.dtype)
    if map_prec is None:
        raise RuntimeError(f"feature maps of dtype {fmap_cl.dtype} are not supported")
    if map_prec not in (0, PREC_IDS[model.prec]):
        raise RuntimeError(f"{fmap_cl.dtype} feature maps need a {fmap_cl.dtype} model, "
                           f"got {model.prec}")
    check_input(fmap_cl, "fmap_cl", fmap_cl.dtype, dev)
    if out is not None:
        _check_out(out, dict(rgb=((B, R, 3), _f32, True), depth=((B, R), _f32, True),
                             aux=((B, R, 2), _f32, False), occ=((B, R, S), _f32, False)), dev)
        if want_occ and out.get("occ") is None:
            raise RuntimeError("want_occ with an out dict that has no 'occ' buffer")
    if out is None:
        out = dict(rgb=torch.empty((B, R, 3), dtype=_f32, device=dev),
                   depth=torch.empty((B, R), dtype=_f32, device=dev),
                   aux=torch.empty((B, R, 2), dtype=_f32, device=dev),
                   occ=torch.zeros((B, R, S), dtype=_f32, device=dev) if want_occ else None)
    ws_bytes = _lib.lib().sailor_render_workspace_bytes_views(B, R, F, V)
    if out.get("workspace") is None or out["workspace"].numel() < ws_bytes:
        out["workspace"] = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
    flags = int(bool(early_exit)) | (2 if compensated else 0) | (4 if fp32_hydra else 0)
    if feature_scales is not None:
        _check(feature_scales, "feature_scales", _f32, (F,), dev)
        call("sailor_render_rays_quantized", 7, ptr(t), ptr(rays), B, R, S, ptr(Kv), ptr(RTv), V,
             int(img_hw[0]), int(img_hw[1]), ptr(fmap_cl), Hf, Wf, F, ptr(feature_scales),
             ptr(bbox), ptr(zdir), model.H, PREC_IDS[model.prec], ptr(model.w_dense),
             ptr(model.b_dense), ptr(model.w_small), float(t_thres), flags, int(chunk_samples),
             ptr(out["rgb"]), ptr(out["depth"]), ptr(out["aux"]), ptr(out["occ"]),
             ptr(out["workspace"]), out["workspace"].numel(), stream_of(dev))
        return out
    call("sailor_render_rays", ptr(t), ptr(rays), B, R, S, ptr(Kv), ptr(RTv), V, int(img_hw[0]),
         int(img_hw[1]), ptr(fmap_cl), map_prec, Hf, Wf, F, ptr(bbox), ptr(zdir), model.H,
         PREC_IDS[model.prec], ptr(model.w_dense), ptr(model.b_dense), ptr(model.w_small),
         float(t_thres), flags, int(chunk_samples), ptr(out["rgb"]), ptr(out["depth"]),
         ptr(out["aux"]), ptr(out["occ"]), ptr(out["workspace"]), out["workspace"].numel(),
         stream_of(dev))
    return out


def density_points(points, Kv, RTv, img_hw, fmap_cl, bbox, model: PackedModel, out=None,
                   chunk_samples=0):
    """Occupancy of the fused model at points [B,N,3] (infer_density path):
    per-view MLP on gathered features + encoding, view mean, occupancy head.
    Returns [B,N] float32."""
    dev = points.device
    check_input(points, "points", _f32)
    if points.dim() != 3 or points.shape[2] != 3:
        raise RuntimeError(f"points must be [B, N, 3], got {list(points.shape)}")
    B, N = points.shape[0], points.shape[1]
    check_input(Kv, "Kv", _f32, dev)
    if Kv.dim() != 4 or Kv.shape[0] != B or tuple(Kv.shape[2:]) != (3, 3):
        raise RuntimeError(f"Kv must be [{B}, V, 3, 3], got {list(Kv.shape)}")
    V = Kv.shape[1]
    _check(RTv, "RTv", _f32, (B, V, 3, 4), dev)
    _check(bbox, "bbox", _f32, (B, 3, 2), dev)
    if fmap_cl.dim() != 5 or fmap_cl.shape[0] != B or fmap_cl.shape[1] != V:
        raise RuntimeError(f"fmap_cl must be [{B}, {V}, Hf, Wf, F], got {list(fmap_cl.shape)}")
    check_input(fmap_cl, "fmap_cl", fmap_cl.dtype, dev)
    Hf, Wf, F = fmap_cl.shape[2], fmap_cl.shape[3], fmap_cl.shape[4]
    if F != model.F:
        raise RuntimeError(f"feature maps have {F} channels, model expects {model.F}")
    if out is not None:
        _check(out, "out", _f32, (B, N), dev)
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
