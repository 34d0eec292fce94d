"""Tensor restatement of the reference's PyTorch compositing -- TEST
INFRASTRUCTURE ONLY (imported by tests/).

Follows c_lib/VoxelEncoding/voxel_encoding_helper.py:
  _volume_render_torch.forward      :132-164  (density form)
  _volume_render_torch.forward_occ  :166-186  (occupancy form)
Its torch autograd is the "reference PyTorch autograd gradient" of the
differentiation form: the reference trains through these tensor ops.

Per ray, samples in the order given (no sort index, no early stop):
  density form:   e_i = (t_{i+1} - t_i) sigma_i with the last interval 0,
                  alpha_i = 1 - exp(-e_i), T_i = exp(-(e_0 + ... + e_{i-1}))
  occupancy form: alpha_i = sigma_i, T_i = prod_{j<i} (1 - alpha_j + 1e-6)
  w_i = alpha_i T_i; feats = sum w_i f_i, depth = sum w_i t_i,
  ws = sum w_i, alphas = sum w_i alpha_i; rays with valid == False give 0.

How it is pinned (tests/test_composite_oracle.py): constant-sigma rows on
equal spacing (T_k = exp(-k sigma delta), closed-form sums), constant
occupancy (geometric series with ratio 1 - a + 1e-6), the zero weight of
the last density-form sample, the valid mask, agreement with an explicit
per-sample Python loop, and its autograd against central finite
differences of the forward.
"""
from __future__ import annotations

import torch

OCC_EPS = 1e-6


def weights(t, sigma, occ_mode):
    """(w [..., S], alpha [..., S]) for rows t, sigma [..., S]."""
    if occ_mode:
        alpha = sigma
        keep = 1.0 - alpha + OCC_EPS
        ones = torch.ones_like(keep[..., :1])
        T = torch.cumprod(torch.cat([ones, keep[..., :-1]], -1), -1)
    else:
        gaps = torch.cat([t[..., 1:] - t[..., :-1], torch.zeros_like(t[..., :1])], -1)
        e = gaps * sigma
        alpha = 1.0 - torch.exp(-e)
        before = torch.cumsum(e, -1) - e            # exclusive prefix sum
        T = torch.exp(-before)
    return alpha * T, alpha


def render(t, sigma, feats, occ_mode, valid=None):
    """t, sigma [B, R, S] (sigma may carry a trailing 1), feats [B, R, S, F]
    -> (feats [B, R, F], depth, ws, alphas [B, R, 1])."""
    sigma = sigma.reshape(t.shape)
    w, alpha = weights(t, sigma, occ_mode)
    out = ((w[..., None] * feats).sum(-2), (w * t).sum(-1, keepdim=True),
           w.sum(-1, keepdim=True), (w * alpha).sum(-1, keepdim=True))
    if valid is not None:
        keep = valid.reshape(t.shape[:-1] + (1,)).to(t.dtype)
        out = tuple(o * keep for o in out)
    return out


def render_loop(t, sigma, feats, occ_mode):
    """The same sums with an explicit per-sample loop over one ray (float64
    scalars): t, sigma [S], feats [S, F] -> (feats [F], depth, ws, alphas)."""
    S = len(t)
    T, acc, d, ws, al = 1.0, [0.0] * len(feats[0]), 0.0, 0.0, 0.0
    for i in range(S):
        if occ_mode:
            a = float(sigma[i])
        else:
            delta = float(t[i + 1] - t[i]) if i + 1 < S else 0.0
            a = 1.0 - float(torch.exp(torch.tensor(-delta * float(sigma[i]), dtype=torch.float64)))
        w = a * T
        acc = [x + w * float(f) for x, f in zip(acc, feats[i])]
        d += w * float(t[i])
        ws += w
        al += w * a
        if occ_mode:
            T *= 1.0 - a + OCC_EPS
        else:
            T *= 1.0 - a
    return acc, d, ws, al
