"""PyTorch restatement of FastNerf's inference entry points -- TEST
INFRASTRUCTURE ONLY (imported by tests/ and bench.py's CPU baseline).

Written from c_lib/FastNerf/src/infer_mlp.cu (read, not run):
  infer_nerf      = infer_fast_rendering        :638-789
                    density_feat_kernel         :120-202
                    density_sigma_kernel        :205-240
                    volume_rendering_occ_kernel :508-581
                    hydra_attention_kernel      :243-445
  infer_density   = infer_density_mlp           :583-634
  infer_blending  = infer_upsampling            :793-821, upsampling_kernel :449-506

Parameter packing (gui/gui_render.py:658-685): the state-dict tensors of
each group concatenated in module order and cast to half; a linear layer is
its weight [out][in] row-major followed by its bias [out].  The counts the
kernels hard-code pin the layout: density 22521 = feature MLP 13696
(18->64->64->64->64) + occupancy MLP 8385 (64->64->64->1, starting at 13696)
+ geometry-feature MLP 440 (18->16->8, starting at 22081); colour 15807 =
two Hydra stacks of 5822 (q, k, v [32][30] and fc [30][32] without bias,
w1 [32][30] + b1, w2 [30][32] + b2) + RGB MLP 4163 (30->64->32->3, starting
at 11644); blending 15075 (40->64->64->32, then [32 | ray feats 30] ->
64->32->3).

The kernels compute in half precision (__hfma accumulation); this
restatement computes in `dtype` (float64 for the tests) from the same half
parameters, so it is the exact value the reference approximates.  With
act_dtype=torch.float16 the inputs of the density MLPs' matrix products
(feature MLP layers, the two hidden occupancy layers, both geometry-feature
layers) and of every Hydra and RGB-head product are rounded to half first:
the numerics of a half-input, wide-accumulation evaluation.  Kept
quirks: q and k are divided by their norm with no epsilon; compositing skips
samples with negative depth, returns zeros for a ray whose last sample depth
is negative, and stops after the sample that takes T below 1e-4; the depth
output is sum_i w_i t_i; ReLU is max(x, 0) with NaN -> 0, so a ray without
samples (0/0 in the normalisation) gets NaN colour features but a finite
colour.

How it is pinned: tests/test_fastnerf_oracle.py (layout offsets, compositing
vs the C oracle, Hydra view permutation and an explicit per-view loop,
density vs view mean) and tests/test_fastnerf_pins.py (every MLP against
torch.nn.Sequential modules loaded at the kernels' offsets, in the kernels'
input order [z, tpsdf, geo]; the occupancy and RGB heads' sigmoids; Hydra
stacks against the identity-projection and zero-value closed forms; copies
of this file with a ReLU, the sigmoid, either residual dropped or z / tpsdf
swapped fail those checks).
"""
from __future__ import annotations

import numpy as np
import torch

DENSITY_PARAMS, COLOR_PARAMS, BLEND_PARAMS = 22521, 15807, 15075
HYDRA_STACK_PARAMS, HYDRA_PARAMS, FEAT_PARAMS, GEO_OFFSET = 5822, 11644, 13696, 22081


class _Reader:
    """Sequential reader of a packed parameter vector."""

    def __init__(self, p, start=0):
        self.p, self.i = p, start

    def take(self, *shape):
        n = int(np.prod(shape))
        v = self.p[self.i:self.i + n].reshape(*shape)
        self.i += n
        return v

    def linear(self, n_in, n_out, bias=True):
        w = self.take(n_out, n_in)
        b = self.take(n_out) if bias else None
        return w, b


def unpack_density(p):
    """-> (feature layers [(W, b)] x4, occupancy layers x3, geometry layers x2)."""
    r = _Reader(p)
    feats = [r.linear(18, 64)] + [r.linear(64, 64) for _ in range(3)]
    assert r.i == FEAT_PARAMS
    sigma = [r.linear(64, 64), r.linear(64, 64), r.linear(64, 1)]
    assert r.i == GEO_OFFSET
    geo = [r.linear(18, 16), r.linear(16, 8)]
    assert r.i == DENSITY_PARAMS
    return feats, sigma, geo


def unpack_color(p):
    """-> (Hydra stacks [dict(q, k, v, fc, w1, b1, w2, b2)] x2, RGB layers x3)."""
    r = _Reader(p)
    stacks = []
    for s in range(2):
        st = {}
        st["q"], _ = r.linear(30, 32, False)
        st["k"], _ = r.linear(30, 32, False)
        st["v"], _ = r.linear(30, 32, False)
        st["fc"], _ = r.linear(32, 30, False)
        st["w1"], st["b1"] = r.linear(30, 32)
        st["w2"], st["b2"] = r.linear(32, 30)
        stacks.append(st)
        assert r.i == HYDRA_STACK_PARAMS * (s + 1)
    rgb = [r.linear(30, 64), r.linear(64, 32), r.linear(32, 3)]
    assert r.i == COLOR_PARAMS
    return stacks, rgb


def unpack_blend(p):
    r = _Reader(p)
    layers = [r.linear(40, 64), r.linear(64, 64), r.linear(64, 32), r.linear(62, 64),
              r.linear(64, 32), r.linear(32, 3)]
    assert r.i == BLEND_PARAMS
    return layers


def _relu(x):
    """max(x, 0.0f) as the kernels write it (infer_mlp.cu:42-45): like fmaxf it
    returns 0 for NaN, where torch.relu would propagate the NaN."""
    return torch.where(x > 0, x, torch.zeros_like(x))


def _t(a, dtype):
    return torch.as_tensor(np.asarray(a, dtype=np.float64)).to(dtype)


def _lin(x, wb):
    w, b = wb
    y = x @ w.T
    return y + b if b is not None else y


def _params(p, dtype):
    """half parameters (numpy float16 / torch half) -> exact values in dtype."""
    if torch.is_tensor(p):
        p = p.detach().cpu().float().numpy()
    return _t(np.asarray(p, np.float16).astype(np.float64), dtype)


def _round(x, act_dtype):
    return x if act_dtype is None else x.to(act_dtype).to(x.dtype)


def density_feats(z, tpsdf, geo, pd, dtype=torch.float64, act_dtype=None):
    """Per (view, point): 64 density features and 8 geometry features
    (density_feat_kernel).  z, tpsdf [N], geo [N, 16]."""
    feats, _, geo_l = unpack_density(_params(pd, dtype))
    x = torch.cat([_t(z, dtype)[:, None], _t(tpsdf, dtype)[:, None], _t(geo, dtype)], 1)
    h = x
    for wb in feats:
        h = _relu(_lin(_round(h, act_dtype), wb))
    g = _relu(_lin(_round(_relu(_lin(_round(x, act_dtype), geo_l[0])), act_dtype), geo_l[1]))
    return h, g


def occupancy(mean_feats, pd, dtype=torch.float64, act_dtype=None):
    """density_sigma_kernel on [N, 64] view-averaged features -> [N]."""
    _, sigma, _ = unpack_density(_params(pd, dtype))
    h = _relu(_lin(_round(mean_feats, act_dtype), sigma[0]))
    h = _relu(_lin(_round(h, act_dtype), sigma[1]))
    return torch.sigmoid(_lin(h, sigma[2]))[:, 0]


def infer_density(points_z, points_tpsdf, feats_geo, pd, B, N, V, dtype=torch.float64,
                  act_dtype=None):
    """infer_density_mlp: inputs [B*V*N] in (batch, view, point) order ->
    occupancy [B, 1, N]."""
    f, _ = density_feats(np.reshape(points_z, -1), np.reshape(points_tpsdf, -1),
                         np.reshape(feats_geo, (-1, 16)), pd, dtype, act_dtype)
    m = f.reshape(B, V, N, 64).mean(1).reshape(B * N, 64)
    return occupancy(m, pd, dtype, act_dtype).reshape(B, 1, N)


def composite(t, alpha, feats, t_thres=1e-4):
    """volume_rendering_occ_kernel: t, alpha [M, S], feats [M, S, C] ->
    (feats [M, C], depth [M])."""
    M, S = t.shape
    out = torch.zeros((M, feats.shape[-1]), dtype=feats.dtype)
    dep = torch.zeros(M, dtype=feats.dtype)
    for m in range(M):
        if t[m, S - 1] < 0:
            continue
        T = 1.0
        f = torch.zeros(feats.shape[-1], dtype=feats.dtype)
        d = 0.0
        for i in range(S):
            if t[m, i] < 0:
                continue
            a = alpha[m, i]
            w = a * T
            f = f + w * feats[m, i]
            d = d + w * t[m, i]
            T = T * (1.0 - a)
            if T < t_thres:
                break
        out[m] = f
        dep[m] = d
    return out, dep


def hydra_rgb(tokens, pc, dtype=torch.float64, act_dtype=None):
    """hydra_attention_kernel: tokens [M, V, 30] -> (colour features of
    token 0 after both stacks [M, 30], rgb [M, 3]); with act_dtype the input
    of every matrix product is rounded to it first."""
    stacks, rgb = unpack_color(_params(pc, dtype))
    r = lambda a: _round(a, act_dtype)
    x = tokens
    for st in stacks:
        q = r(x) @ st["q"].T
        k = r(x) @ st["k"].T
        v = r(x) @ st["v"].T
        q = q / torch.sqrt((q * q).sum(-1, keepdim=True))
        k = k / torch.sqrt((k * k).sum(-1, keepdim=True))
        kv = (k * v).sum(1, keepdim=True)
        out = r(q * kv) @ st["fc"].T
        res = x + out
        h = _relu(r(res) @ st["w1"].T + st["b1"])
        x = r(h) @ st["w2"].T + st["b2"] + res
    f = x[:, 0]
    h = _relu(_lin(r(f), rgb[0]))
    h = _relu(_lin(r(h), rgb[1]))
    return f, torch.sigmoid(_lin(r(h), rgb[2]))


def infer_nerf(sorted_dists, points_z, points_tpsdf, feats_geo, rgbs, feats_rgb, dirs, pd, pc,
               B, R, S, Vg, Vt, dtype=torch.float64, act_dtype=None):
    """infer_fast_rendering.  sorted_dists [B, R, S]; per-view inputs flattened
    in (batch, view, ray, sample) order: z, tpsdf [B*Vg*R*S], feats_geo
    [.., 16]; rgbs, dirs [B*Vt*R*S, 3], feats_rgb [.., 16].
    Returns (depth [B, R, 1], rgb [B, R, 3], colour features [B, R, 30])."""
    f, g = density_feats(np.reshape(points_z, -1), np.reshape(points_tpsdf, -1),
                         np.reshape(feats_geo, (-1, 16)), pd, dtype, act_dtype)
    m = f.reshape(B, Vg, R * S, 64).mean(1).reshape(B * R * S, 64)
    alpha = occupancy(m, pd, dtype, act_dtype).reshape(B * R, S)
    g = g.reshape(B, Vg, R, S, 8)[:, :Vt]
    t = _t(sorted_dists, dtype).reshape(B * R, S)
    fr = _t(feats_rgb, dtype).reshape(B, Vt, R, S, 16)
    rg = _t(rgbs, dtype).reshape(B, Vt, R, S, 3)
    dr = _t(dirs, dtype).reshape(B, Vt, R, S, 3)
    toks, dep = [], None
    for v in range(Vt):
        fv = torch.cat([fr[:, v], rg[:, v], g[:, v], dr[:, v]], -1).reshape(B * R, S, 30)
        tk, dep = composite(t, alpha, fv)
        toks.append(tk)
    toks = torch.stack(toks, 1)                      # [B*R, Vt, 30]
    feats, rgb = hydra_rgb(toks, pc, dtype, act_dtype)
    return dep.reshape(B, R, 1), rgb.reshape(B, R, 3), feats.reshape(B, R, 30)


def infer_blending(feats_neighbors, ray_feats, pb, dtype=torch.float64, act_dtype=None):
    """upsampling_kernel: [N, 40], [N, 30] -> blend weights [N, 3]; with
    act_dtype the inputs of the first five layers are rounded to it."""
    L = unpack_blend(_params(pb, dtype))
    r = lambda a: _round(a, act_dtype)
    x = _t(feats_neighbors, dtype)
    h = _relu(_lin(r(x), L[0]))
    h = _relu(_lin(r(h), L[1]))
    h = _relu(_lin(r(h), L[2]))
    h = torch.cat([h, _t(ray_feats, dtype)], 1)
    h = _relu(_lin(r(h), L[3]))
    h = _relu(_lin(r(h), L[4]))
    return torch.sigmoid(_lin(h, L[5]))
