"""Differentiation of the fused render (the training form): backward of the
Hydra encoder + RGB head, of the occupancy compositing and of the per-view
MLP, on the HIP kernels of libsailor_hip.so.

Layout helpers read the forward's workspace (ray list, per-view tokens) the
way sailor_render_rays writes it: [counter (256 B)] [list: B*R int32,
256-byte aligned] [tokens: B*R x 4 x F float32].
"""
from __future__ import annotations

import torch

from . import _lib
from ._lib import call, ptr, stream_of


def workspace_views(ws: torch.Tensor, n_rays: int, F: int):
    """(n_valid, list [n_valid] int32, tokens [n_valid, 4, F] float32) views
    of a render workspace (synchronises to read the count)."""
    n_valid = int(ws[:4].view(torch.int32)[0].item())
    lst = ws[256:256 + 4 * n_rays].view(torch.int32)[:n_valid]
    toff = 256 + (n_rays * 4 + 255) // 256 * 256
    tok = ws[toff:toff + n_rays * 4 * F * 4].view(torch.float32).view(n_rays, 4, F)[:n_valid]
    return n_valid, lst, tok


class BackwardScratch:
    """Device buffers of the backward for B x R rays of a model."""

    def __init__(self, B, R, F, H, device):
        L = _lib.lib()
        self.F, self.H = F, H
        self.grad_tokens = torch.empty((B * R, 4, F), dtype=torch.float32, device=device)
        self.grad_small = torch.empty(L.sailor_render_small_weights_floats(F, H),
                                      dtype=torch.float32, device=device)
        self.partials = torch.empty(L.sailor_render_backward_partials_floats(F, H),
                                    dtype=torch.float32, device=device)


def hydra_backward(grad_rgb, B, R, V, model, workspace, scratch: BackwardScratch):
    """dL/d rgb [B*R, 3] -> (grad_tokens [B*R, 4, F] in list order, grad of
    the small-weight buffer)."""
    call("sailor_render_hydra_backward", ptr(grad_rgb), B, R, V, model.F, model.H,
         ptr(model.w_small), ptr(workspace), workspace.numel(), ptr(scratch.grad_tokens),
         ptr(scratch.grad_small), ptr(scratch.partials), scratch.partials.numel(),
         stream_of(grad_rgb.device))
    return scratch.grad_tokens, scratch.grad_small
