"""PyTorch restatement of the fused radiance evaluation -- TEST
INFRASTRUCTURE ONLY (imported by tests/ and bench.py's CPU baseline).

Written from the reference modules, not from the HIP kernel:
  projection + grid_sample   utils_render/utils_render.py:380-410 (proj_persp),
  density query              c_lib/FastNerf/src/infer_mlp.cu:583-634
                             SRONet/BasicRenNet.py:245-259 (F.grid_sample,
                             align_corners=True, zero padding)
  frequency encoding         c_lib/VoxelEncoding/src/freq_encoding.cu:37-45
  per-view MLP, view mean    SRONet/nerf.py:139-152, FastNerf infer_mlp.cu:687-713
  occupancy compositing      c_lib/VoxelEncoding/src/volume_rendering.cu:666-745
  per-view token compositing c_lib/FastNerf/src/infer_mlp.cu:724-747
  Hydra encoder              SRONet/aggregation.py:85-128,155-171,288-307
  RGB head                   SRONet/nerf.py:117-124
Runs in any dtype/device; tests use float64 on the CPU as the accuracy oracle.
"""
from __future__ import annotations

import torch
import torch.nn.functional as Fn


def freq_encode(x, deg=10):
    outs = [x]
    for f in range(deg):
        outs.append(torch.sin(x * (2.0 ** f)))
        outs.append(torch.sin(x * (2.0 ** f) + torch.pi / 2))
    # reference layout: [x, sin f0 (3), cos f0 (3), sin f1, ...]
    return torch.cat(outs, -1)


def gather_features(x, Kv, RTv, fmaps, img_hw):
    """Per-view features at world points x [..., 3]: proj_persp + grid_sample
    (align_corners=True, zero padding); Kv [V,3,3], RTv [V,3,4], fmaps
    [V,F,Hf,Wf].  Returns a list over views of [..., F]."""
    img_h, img_w = img_hw
    feats = []
    shape = x.shape[:-1]
    xf = x.reshape(1, -1, 3)
    for v in range(Kv.shape[0]):
        Rv, Tv = RTv[v, :, :3], RTv[v, :, 3]
        p = xf @ Rv.T + Tv
        z = p[..., 2:3]
        inv = z < 0.01
        zz = torch.where(inv, torch.ones_like(z), z)
        q = p / zz
        scr = q @ Kv[v].T
        res = torch.tensor([img_w // 2, img_h // 2], dtype=x.dtype, device=x.device)
        grid = (scr[..., :2] - res) / res
        grid = torch.where(inv.expand_as(grid), torch.full_like(grid, -2.0), grid)
        smp = Fn.grid_sample(fmaps[v][None], grid[None], mode="bilinear",
                             padding_mode="zeros", align_corners=True)[0]  # [F, 1, n]
        feats.append(smp[:, 0].T.reshape(*shape, -1))
    return feats


def _mm(x, w, mm_dtype=None):
    """x @ w.T; with mm_dtype both operands are rounded to it first and the
    product accumulated in x's dtype (a 16-bit matrix unit).  float16 has a
    narrow range (max 65504), so each row of x is divided by the power of two
    2^e >= its max |x| (m = f 2^e, f in [0.5, 1)) before rounding and the
    row's results multiplied back by it: exact scalings, the rounding of
    every element that stays in fp16's normal range is unchanged."""
    if mm_dtype is None:
        return x @ w.T
    rd = lambda z: z.to(mm_dtype).to(x.dtype)
    if mm_dtype != torch.float16:
        return rd(x) @ rd(w).T
    m = x.detach().abs().amax(dim=-1, keepdim=True)
    _, e = torch.frexp(m)
    sig = torch.where(m > 0, torch.ldexp(torch.ones_like(m), e), torch.ones_like(m))
    return (rd(x / sig) @ rd(w).T) * sig


def hydra_encode(hydra, toks, mm_dtype=None):
    """Hydra encoder stacks over the view tokens toks [N, V, F]
    (aggregation.py:101-128 HydraAttention, :164-171 HydraForward,
    :299-307 HydraEncoder): q, k, v projections, q and k L2-normalised over
    features, kv = sum over views of k * v, out = fc(q * kv) + x, then
    x' = w2 relu(w1 out + b1) + b2 + out.  hydra: list of dicts of tensors
    (q, k, v, fc, w1, b1, w2, b2).  mm_dtype: products on rounded operands.
    Returns the tokens [N, V, F]."""
    x_tok = toks
    for st in hydra:
        q = Fn.normalize(_mm(x_tok, st["q"], mm_dtype), dim=-1)
        k = Fn.normalize(_mm(x_tok, st["k"], mm_dtype), dim=-1)
        vv = _mm(x_tok, st["v"], mm_dtype)
        kv = (k * vv).sum(dim=-2, keepdim=True)
        out = _mm(q * kv, st["fc"], mm_dtype) + x_tok
        hid = torch.relu(_mm(out, st["w1"], mm_dtype) + st["b1"])
        x_tok = _mm(hid, st["w2"], mm_dtype) + st["b2"] + out
    return x_tok


def hydra_rgb(hydra, rgbw, toks, mm_dtype=None):
    """hydra_encode, then the RGB head on the first view's token
    (nerf.py:117-124).  rgbw: [(w1, b1), (w2, b2), (w3, b3)].  Returns rgb
    [N, 3]."""
    x_tok = hydra_encode(hydra, toks, mm_dtype)
    h = x_tok[:, 0]
    (w1, b1), (w2, b2), (w3, b3) = rgbw
    h = torch.relu(_mm(h, w1, mm_dtype) + b1)
    h = torch.relu(_mm(h, w2, mm_dtype) + b2)
    return torch.sigmoid(_mm(h, w3, mm_dtype) + b3)


def density(weights, x, Kv, RTv, fmaps, bbox, img_hw, dtype=torch.float64, device="cpu"):
    """Occupancy at points x [B,N,3] (infer_density_mlp, infer_mlp.cu:583-634,
    for this model): per-view MLP on [PE | gathered features], view mean,
    occupancy head.  Returns [B,N]."""
    T_ = lambda a: torch.as_tensor(a).to(device=device, dtype=dtype)
    x, Kv, RTv, fmaps, bbox = map(T_, (x, Kv, RTv, fmaps, bbox))
    W = [T_(w) for w in weights.dense_w]
    Bs = [T_(b) for b in weights.dense_b]
    out = []
    for b in range(x.shape[0]):
        c = 0.5 * (bbox[b, :, 0] + bbox[b, :, 1])
        ih = 2.0 / (bbox[b, :, 1] - bbox[b, :, 0])
        pe = freq_encode((x[b] - c) * ih)
        feats = gather_features(x[b], Kv[b], RTv[b], fmaps[b], img_hw)
        h_sum = 0
        for f in feats:
            h = torch.cat([pe, f], -1)
            for li in range(4):
                h = torch.relu(h @ W[li].T + Bs[li])
            h_sum = h_sum + h
        out.append(torch.sigmoid((h_sum / len(feats)) @ T_(weights.sigma_w) + weights.sigma_b))
    return torch.stack(out)


def params_of(weights, dtype=torch.float64, device="cpu", requires_grad=False):
    """The model's parameters as tensors (leaves when requires_grad)."""
    T_ = lambda a: torch.tensor(a, dtype=dtype, device=device, requires_grad=requires_grad)
    return dict(dense_w=[T_(w) for w in weights.dense_w], dense_b=[T_(b) for b in weights.dense_b],
                sigma_w=T_(weights.sigma_w), sigma_b=T_(float(weights.sigma_b)),
                hydra=[{k: T_(v) for k, v in st.items()} for st in weights.hydra],
                rgb=[(T_(w), T_(b)) for w, b in weights.rgb])


def render_rays(weights, t, rays, Kv, RTv, fmaps, bbox, zdir, img_hw, t_thres=1e-4,
                dtype=torch.float64, device="cpu", params=None, trace=None, mm_dtype=None,
                hydra_mm_dtype=None):
    """weights: sailor_amd.model.NerfWeights; t [B,R,S] sorted; rays [B,R,6];
    Kv [B,V,3,3]; RTv [B,V,3,4]; fmaps [B,V,F,Hf,Wf] (channel-first);
    bbox [B,3,2]; zdir [B,3].  Returns rgb [B,R,3], depth [B,R], ws [B,R],
    asum [B,R], occ [B,R,S] (zeros for dropped rays).  With a dict `trace`
    (B == 1), trace["x0"][v] holds view v's first-layer input [N,S,63+F]
    and trace["pre"][v][l] its layer-l pre-activation [N,S,H] (gradients
    retained), trace["toks"] the per-view tokens [N,V,F] and trace["occ"]
    the occupancies [N,S], for the valid rays in order.  With mm_dtype (e.g. bfloat16)
    the dense layers' inputs and weights are rounded to that type and the
    products accumulated in `dtype` (mixed-precision matrix units; gradients
    are rounded where they pass the same casts).  hydra_mm_dtype does the
    same for the Hydra encoder's and the RGB head's products."""
    T_ = lambda a: torch.as_tensor(a).to(device=device, dtype=dtype)
    t, rays, Kv, RTv, fmaps, bbox, zdir = map(T_, (t, rays, Kv, RTv, fmaps, bbox, zdir))
    B, R, S = t.shape
    V = Kv.shape[1]
    P = params if params is not None else params_of(weights, dtype, device)
    W, Bs = P["dense_w"], P["dense_b"]
    rgb_out = torch.zeros((B, R, 3), dtype=dtype, device=device)
    dep_out = torch.zeros((B, R), dtype=dtype, device=device)
    ws_out = torch.zeros((B, R), dtype=dtype, device=device)
    as_out = torch.zeros((B, R), dtype=dtype, device=device)
    occ_out = torch.zeros((B, R, S), dtype=dtype, device=device)
    for b in range(B):
        valid = t[b, :, 0] >= 0
        idx = torch.nonzero(valid).flatten()
        if idx.numel() == 0:
            continue
        tb = t[b, idx]                                   # [N, S]
        o, d = rays[b, idx, :3], rays[b, idx, 3:]
        x = o[:, None, :] + tb[..., None] * d[:, None, :]  # [N, S, 3]
        c = 0.5 * (bbox[b, :, 0] + bbox[b, :, 1])
        ih = 2.0 / (bbox[b, :, 1] - bbox[b, :, 0])
        pe = freq_encode((x - c) * ih)                   # [N, S, 63]
        feats = gather_features(x, Kv[b], RTv[b], fmaps[b], img_hw)  # V x [N, S, F]
        h_sum = 0
        if trace is not None:
            trace["x0"], trace["pre"] = [], []
        for v in range(V):
            h = torch.cat([pe, feats[v]], -1)
            if trace is not None:
                trace["x0"].append(h)
                trace["pre"].append([])
            for li in range(4):
                if mm_dtype is None:
                    pre = h @ W[li].T + Bs[li]
                else:
                    rd = lambda z: z.to(mm_dtype).to(dtype)
                    pre = rd(h) @ rd(W[li]).T + Bs[li]
                if trace is not None:
                    if pre.requires_grad:
                        pre.retain_grad()
                    trace["pre"][v].append(pre)
                h = torch.relu(pre)
            h_sum = h_sum + h
        h_mean = h_sum / V
        alpha = torch.sigmoid(h_mean @ P["sigma_w"] + P["sigma_b"])  # [N, S]
        # sequential front-to-back compositing with early termination
        N = tb.shape[0]
        Tt = torch.ones(N, dtype=dtype, device=device)
        alive = torch.ones(N, dtype=torch.bool, device=device)
        wsum = torch.zeros(N, dtype=dtype, device=device)
        dsum = torch.zeros(N, dtype=dtype, device=device)
        asum = torch.zeros(N, dtype=dtype, device=device)
        toks = torch.zeros((N, V, feats[0].shape[-1]), dtype=dtype, device=device)
        for s in range(S):
            a = alpha[:, s]
            w = torch.where(alive, a * Tt, torch.zeros_like(a))
            dsum = dsum + w * tb[:, s]
            wsum = wsum + w
            asum = asum + w * a
            for v in range(V):
                toks[:, v] += w[:, None] * feats[v][:, s]
            Tt = torch.where(alive, Tt * (1 - a), Tt)
            alive = alive & ~(Tt < t_thres)
        if trace is not None:
            trace["toks"], trace["occ"] = toks, alpha
        rgb = hydra_rgb(P["hydra"], P["rgb"], toks, hydra_mm_dtype)
        rgb_out[b, idx] = rgb
        dep_out[b, idx] = dsum * (d @ zdir[b])
        ws_out[b, idx] = wsum
        as_out[b, idx] = asum
        occ_out[b, idx] = alpha
    return rgb_out, dep_out, ws_out, as_out, occ_out
