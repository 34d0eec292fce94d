"""Differentiation of the fused render (the training form): backward of the
Hydra encoder + RGB head, of the occupancy compositing and of the per-view
MLP, on the HIP kernels of libsailor_hip.so.

Layout helpers read the forward's workspace (ray list, per-view tokens) the
way sailor_render_rays writes it: [counter (256 B)] [list: B*R int32,
256-byte aligned] [tokens: B*R x 4 x F float32].
"""
from __future__ import annotations

import ctypes

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
        self.grad_dense = torch.empty(L.sailor_render_dense_grad_floats(F, H),
                                      dtype=torch.float32, device=device)
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


def pack_transposed(weights, prec, device):
    """Packed A fragments of W_3^T, W_2^T, W_1^T (the delta propagation's
    layers: input width H), in the forward's packing."""
    import numpy as np
    from .model import pack_layer
    frags = [pack_layer(np.ascontiguousarray(weights.dense_w[l].T), prec) for l in (3, 2, 1)]
    dense = np.concatenate([f.reshape(-1) for f in frags]).astype(np.float32)
    t = torch.from_numpy(dense)
    if prec != "fp32":
        t = t.to(torch.bfloat16 if prec == "bf16" else torch.float16)
    return t.to(device).contiguous()


def _mm_t(D, A, split=64):
    """D^T A with float32 output over a long column dimension: the columns
    are cut into `split` batches (one batched GEMM, then a sum) so the
    reduction spreads over the whole device; 16-bit inputs accumulate in
    fp32."""
    C = D.shape[0]
    s = split if C >= 64 * split else 1
    c = C // s * s
    Db = D[:c].reshape(s, c // s, D.shape[1]).transpose(1, 2)
    Ab = A[:c].reshape(s, c // s, A.shape[1])
    if D.dtype == torch.float32:
        out = torch.bmm(Db, Ab).sum(0)
    else:
        try:
            out = torch.bmm(Db, Ab, out_dtype=torch.float32).sum(0)
        except (TypeError, RuntimeError):
            out = torch.bmm(Db.float(), Ab.float()).sum(0)
    if c < C:
        out = out + _mm_t(D[c:], A[c:], 1)
    return out


def sample_backward(t, rays, Kv, RTv, img_hw, fmap_cl, bbox, model, wt, occ, workspace,
                    grad_tokens, grad_small, partials, n_valid, t_thres=1e-4,
                    batch_items=4096, columns=None):
    """Compositing + MLP backward over the listed rays in batches of
    batch_items; returns (dW [4] in kernel input order, db [4]) as float32
    and adds the occupancy-head gradient to grad_small.  A list `columns`
    receives (A [4], D [4]) column matrices of the last batch."""
    B, R, S = t.shape
    V = Kv.shape[1]
    Hf, Wf, F = fmap_cl.shape[2], fmap_cl.shape[3], fmap_cl.shape[4]
    H = model.H
    dev = t.device
    dt = {"fp32": torch.float32, "bf16": torch.bfloat16, "fp16": torch.float16}[model.prec]
    from .model import PREC_IDS
    cols = min(batch_items, max(n_valid, 1)) * S * 4
    bufs = [torch.empty((cols, H), dtype=dt, device=dev) for _ in range(8)]
    arr = (ctypes.c_void_p * 8)(*[b.data_ptr() for b in bufs])
    K0 = F + 64
    dW = [torch.zeros((H, K0 if l == 0 else H), dtype=torch.float32, device=dev)
          for l in range(4)]
    db = [torch.zeros(H, dtype=torch.float32, device=dev) for _ in range(4)]
    if F + 64 > H:
        raise RuntimeError(f"first-layer width {F + 64} exceeds the column stride {H}")
    ones = torch.ones(cols, dtype=dt, device=dev)
    for i0 in range(0, n_valid, batch_items):
        i1 = min(i0 + batch_items, n_valid)
        call("sailor_render_sample_backward", ptr(t), ptr(rays), B, R, S, ptr(Kv), ptr(RTv), V,
             int(img_hw[0]), int(img_hw[1]), ptr(fmap_cl), Hf, Wf, F, ptr(bbox), H,
             PREC_IDS[model.prec], ptr(model.w_dense), ptr(wt), ptr(model.b_dense),
             ptr(model.w_small), float(t_thres), ptr(occ), ptr(grad_tokens), ptr(workspace),
             workspace.numel(), i0, i1, arr, ptr(grad_small), ptr(partials), partials.numel(),
             stream_of(dev))
        c = (i1 - i0) * S * 4
        A = [bufs[0][:c, :K0], bufs[1][:c], bufs[2][:c], bufs[3][:c]]
        D = [bufs[4][:c], bufs[5][:c], bufs[6][:c], bufs[7][:c]]
        for l in range(4):
            dW[l] += _mm_t(D[l], A[l])
            db[l] += _mm_t(D[l], ones[:c, None])[:, 0]
        if columns is not None:
            columns[:] = [A, D]
    return dW, db


def sample_backward_fused(t, rays, Kv, RTv, img_hw, fmap_cl, bbox, model, wt, occ, workspace,
                          grad_tokens, grad_small, partials, grad_dense, t_thres=1e-4):
    """Compositing + MLP backward with the weight gradients accumulated
    inside the kernel (H = 128 models); returns (dW [4] in kernel input
    order, db [4]) as views of grad_dense and adds the occupancy-head
    gradient to grad_small."""
    B, R, S = t.shape
    V = Kv.shape[1]
    Hf, Wf, F = fmap_cl.shape[2], fmap_cl.shape[3], fmap_cl.shape[4]
    H = model.H
    from .model import PREC_IDS
    call("sailor_render_sample_backward_fused", ptr(t), ptr(rays), B, R, S, ptr(Kv), ptr(RTv),
         V, int(img_hw[0]), int(img_hw[1]), ptr(fmap_cl), Hf, Wf, F, ptr(bbox), H,
         PREC_IDS[model.prec], ptr(model.w_dense), ptr(wt), ptr(model.b_dense),
         ptr(model.w_small), float(t_thres), ptr(occ), ptr(grad_tokens), ptr(workspace),
         workspace.numel(), ptr(grad_dense), ptr(grad_small), ptr(partials), partials.numel(),
         stream_of(t.device))
    K0 = F + 64
    o1 = H * K0
    dW = [grad_dense[:o1].view(H, K0)] + [grad_dense[o1 + l * H * H:o1 + (l + 1) * H * H]
                                          .view(H, H) for l in range(3)]
    ob = o1 + 3 * H * H
    db = [grad_dense[ob + l * H:ob + (l + 1) * H] for l in range(4)]
    return dW, db


def natural_first_layer_grad(dW0, F):
    """Kernel-order first-layer gradient [H, F + 64] -> the model's [H, 63 + F]
    ([encoding | features])."""
    return torch.cat([dW0[:, F:F + 63], dW0[:, :F]], 1)


class RenderTrainer:
    """Forward and backward of the fused render for batches of rays, with
    parameter gradients in the layout of NerfWeights.

    forward() runs sailor_render_rays keeping the occupancies; backward(g)
    takes dL/d rgb and runs (1) the Hydra + RGB head backward, (2) the
    compositing + per-view MLP backward in item batches, each followed by
    the weight-gradient GEMMs D_l^T A_l on the GPU.  The reference's
    differentiation of this path is torch autograd through SRONet/nerf.py
    and _volume_render_torch.forward_occ (voxel_encoding_helper.py:166-186).
    """

    def __init__(self, weights, prec="fp32", device="cuda", batch_items=16384, fused=True):
        from .render import PackedModel
        self.weights, self.prec, self.dev = weights, prec, torch.device(device)
        self.model = PackedModel(weights, prec, self.dev)
        self.wt = pack_transposed(weights, prec, self.dev)
        self.batch_items = batch_items
        self.fused = fused
        self._out = None
        self._scratch = None
        self._ctx = None

    def forward(self, t, rays, Kv, RTv, img_hw, fmap_cl, bbox, zdir, t_thres=1e-4):
        from . import render
        B, R, S = t.shape
        # the output buffers are reused only for a batch of the same shape on
        # the same device (the kernel writes occ[ray * S + sample])
        o = self._out
        if o is not None and (tuple(o["occ"].shape) != (B, R, S) or o["occ"].device != t.device):
            self._out = None
        # the backward recomputes the Hydra encoder and RGB head in fp32 from
        # the tokens, so the forward evaluates them the same way
        self._out = render.render_rays(t, rays, Kv, RTv, img_hw, fmap_cl, bbox, zdir,
                                       self.model, t_thres, early_exit=False, want_occ=True,
                                       out=self._out, fp32_hydra=True)
        self._ctx = (t, rays, Kv, RTv, img_hw, fmap_cl, bbox, t_thres)
        return self._out

    def backward(self, grad_rgb, keep_columns=False):
        """dL/d rgb [B, R, 3] -> dict of float32 gradients shaped like the
        model's parameters (dense_w, dense_b, sigma_w, sigma_b, hydra, rgb).
        H = 128 models with `fused` accumulate the weight gradients inside the
        kernel (no host synchronisation); otherwise the column form runs in
        item batches followed by GEMMs (synchronises once to read the number
        of listed rays).  With keep_columns (column form), self.columns holds
        the last item batch's (A, D) column matrices (column (item, sample,
        view), item in list order)."""
        if self._ctx is None:
            raise RuntimeError("backward() before forward()")
        t, rays, Kv, RTv, img_hw, fmap_cl, bbox, t_thres = self._ctx
        if fmap_cl.dtype != torch.float32:
            raise RuntimeError("the backward reads float32 feature maps")
        B, R, _ = t.shape
        V, F, H = Kv.shape[1], self.model.F, self.model.H
        if self._scratch is None or self._scratch.grad_tokens.shape[0] != B * R:
            self._scratch = BackwardScratch(B, R, F, H, self.dev)
        ws = self._out["workspace"]
        g = grad_rgb.reshape(B * R, 3).contiguous().float()
        hydra_backward(g, B, R, V, self.model, ws, self._scratch)
        sc = self._scratch
        if self.fused and H == 128 and not keep_columns:
            dW, db = sample_backward_fused(t, rays, Kv, RTv, img_hw, fmap_cl, bbox, self.model,
                                           self.wt, self._out["occ"], ws, sc.grad_tokens,
                                           sc.grad_small, sc.partials, sc.grad_dense, t_thres)
            self.columns = None
        else:
            n_valid = int(ws[:4].view(torch.int32)[0].item())
            cols = [] if keep_columns else None
            dW, db = sample_backward(t, rays, Kv, RTv, img_hw, fmap_cl, bbox, self.model,
                                     self.wt, self._out["occ"], ws, sc.grad_tokens,
                                     sc.grad_small, sc.partials, n_valid, t_thres,
                                     self.batch_items, columns=cols)
            self.columns = cols
        # the scratch buffers are overwritten by the next backward(): hand out
        # copies, so gradients of several micro-batches can be accumulated
        small = self.weights.split_small(self._scratch.grad_small.clone())
        return dict(dense_w=[natural_first_layer_grad(dW[0], F)] + [w.clone() for w in dW[1:]],
                    dense_b=[b.clone() for b in db], sigma_w=small["sigma_w"],
                    sigma_b=small["sigma_b"], hydra=small["hydra"], rgb=small["rgb"])


class TrainingBatch:
    """Rays of one training batch: n_targets target views (ring cameras
    between the source views) of cfg.target^2 rays each, traced through the
    tree of one fused frame (TSDF -> tree -> DDA -> samples -> sort) and
    concatenated into one [1, n_targets * target^2] ray batch, as the
    reference's training step batches target views (SRONet/BasicRenNet.py
    forward_ray_intersect / forward_ray_nerf)."""

    def __init__(self, cfg, n_targets=8, device="cuda"):
        import numpy as np
        from . import scenes
        from .pipeline import FramePipeline
        self.cfg = cfg
        p = FramePipeline.synthetic(cfg, prec="fp32", device=device)
        p.build()
        nt = cfg.target * cfg.target
        n = n_targets * nt
        self.t = torch.empty((1, n, cfg.S), dtype=torch.float32, device=device)
        self.rays = torch.empty((1, n, 6), dtype=torch.float32, device=device)
        cams = scenes.ring_cameras(n_targets, cfg.target_fx, cfg.target, cfg.target,
                                   phase=np.pi / n_targets)
        cu = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float32))
        for i, cam in enumerate(cams):
            p.Kt.copy_(cu(cam.K[None]))
            p.Rt.copy_(cu(cam.R[None]))
            p.Tt.copy_(cu(cam.T[None]))
            p.trace()
            self.t[0, i * nt:(i + 1) * nt].copy_(p.sorted[0])
            self.rays[0, i * nt:(i + 1) * nt].copy_(p.rays[0])
        torch.cuda.synchronize()
        self.pipe = p
        self.n_rays = n
        self.n_valid = int((self.t[0, :, 0] >= 0).sum().item())
        self.img_hw = (cfg.img, cfg.img)

    def inputs(self):
        """Positional arguments of RenderTrainer.forward."""
        p = self.pipe
        return (self.t, self.rays, p.K, p.RT, self.img_hw, p.fmap, p.bbox, p.zdir)


def mse_grad(rgb, target, n_valid):
    """dL/d rgb of L = mean over the valid rays' channels of (rgb - target)^2
    (dropped rays render 0 in both)."""
    return (rgb - target) * (2.0 / (3.0 * max(n_valid, 1)))


# ---------------------------------------------------------------------------
# autograd form: parameters as torch leaves, packed on the device each step
def _pack_index(prec, H, K):
    """Flat gather index (into a natural-order [H, K] layer) of every slot of
    pack_layer's output, so that packing is one device gather."""
    import numpy as np
    from .model import kmap, unit_rows
    rows, cols = np.broadcast_arrays(unit_rows(prec, H), kmap(prec, K)[None, :, :, None, :])
    return torch.from_numpy((rows.astype(np.int64) * K + cols).reshape(-1))


class RadianceField(torch.nn.Module):
    """The fused render's parameters (NerfWeights layout) as nn.Parameters.

    forward() renders through FusedRender, a torch.autograd.Function whose
    backward is the HIP differentiation of the fused render (Hydra + RGB head
    backward, compositing reverse scan, per-view MLP with per-ray recompute;
    RenderTrainer), so a training step is
        rgb = field(*batch.inputs()); loss(rgb).backward(); optimizer.step()
    as in the reference's SRONet training step (SRONet/SRONet.py:238-258),
    where torch autograd differentiates nerf.py:139-186 and
    _volume_render_torch.forward_occ (voxel_encoding_helper.py:166-186).
    The packed device buffers are rebuilt (device gathers, no host copies)
    whenever a parameter changed since the last render."""

    def __init__(self, weights, prec="fp32", device="cuda", batch_items=16384, fused=True):
        super().__init__()
        import numpy as np
        dev = torch.device(device)
        P = lambda a: torch.nn.Parameter(torch.as_tensor(np.asarray(a, np.float32)).to(dev))
        self.weights, self.prec = weights, prec
        self.F, self.H = weights.F, weights.H
        self.dense_w = torch.nn.ParameterList([P(w) for w in weights.dense_w])
        self.dense_b = torch.nn.ParameterList([P(b) for b in weights.dense_b])
        self.sigma_w = P(weights.sigma_w)
        self.sigma_b = P(np.array([weights.sigma_b], np.float32))
        self.hydra = torch.nn.ParameterList(
            [P(st[k]) for st in weights.hydra for k in HYDRA_KEYS])
        self.rgb = torch.nn.ParameterList([P(x) for wb in weights.rgb for x in wb])
        K0 = self.F + 64
        self._idx = dict(first=_pack_index(prec, self.H, K0).to(dev),
                         hidden=_pack_index(prec, self.H, self.H).to(dev))
        self.trainer = RenderTrainer(weights, prec, dev, batch_items=batch_items, fused=fused)
        self._versions = None

    def flat_params(self):
        return [*self.dense_w, *self.dense_b, self.sigma_w, self.sigma_b, *self.hydra,
                *self.rgb]

    def _pack(self):
        """Repacks the device buffers the kernels read from the parameters."""
        vers = tuple(p._version for p in self.flat_params())
        if vers == self._versions:
            return
        dt = {"fp32": torch.float32, "bf16": torch.bfloat16, "fp16": torch.float16}[self.prec]
        F, H = self.F, self.H
        with torch.no_grad():
            W0 = self.dense_w[0]
            first = torch.cat([W0[:, 63:63 + F], W0[:, :63], torch.zeros_like(W0[:, :1])], 1)
            parts = [first.reshape(-1)[self._idx["first"]]]
            parts += [self.dense_w[i].reshape(-1)[self._idx["hidden"]] for i in (1, 2, 3)]
            m = self.trainer.model
            m.w_dense.copy_(torch.cat(parts).to(dt))
            m.b_dense.copy_(torch.stack(list(self.dense_b)))
            wt = [self.dense_w[i].t().contiguous().reshape(-1)[self._idx["hidden"]]
                  for i in (3, 2, 1)]
            self.trainer.wt.copy_(torch.cat(wt).to(dt))
            small = [self.sigma_w.reshape(-1),
                     torch.cat([self.sigma_b.reshape(1), self.sigma_b.new_zeros(3)])]
            small += [p.reshape(-1) for p in self.hydra]
            r = list(self.rgb)
            small += [r[0].reshape(-1), r[1], r[2].reshape(-1), r[3], r[4].reshape(-1),
                      torch.cat([r[5], r[5].new_zeros(1)])]
            m.w_small.copy_(torch.cat(small))
        self._versions = vers

    def forward(self, t, rays, Kv, RTv, img_hw, fmap_cl, bbox, zdir, t_thres=1e-4):
        """-> (rgb [B,R,3] differentiable, depth [B,R], aux [B,R,2])."""
        self._pack()
        return FusedRender.apply(self, t, rays, Kv, RTv, tuple(img_hw), fmap_cl, bbox, zdir,
                                 float(t_thres), *self.flat_params())


HYDRA_KEYS = ("q", "k", "v", "fc", "w1", "b1", "w2", "b2")


class FusedRender(torch.autograd.Function):
    """rgb = fused render(parameters); backward = the HIP kernels' gradient
    (no activations stored: the sample kernel's backward recomputes each
    ray's layers from its inputs)."""

    @staticmethod
    def forward(ctx, field, t, rays, Kv, RTv, img_hw, fmap_cl, bbox, zdir, t_thres, *params):
        tr = field.trainer
        out = tr.forward(t, rays, Kv, RTv, img_hw, fmap_cl, bbox, zdir, t_thres)
        # this render's occupancies and workspace belong to its backward: the
        # next forward allocates its own
        ctx.field, ctx.state = field, (tr._ctx, out)
        tr._out = None
        depth, aux = out["depth"].clone(), out["aux"].clone()
        ctx.mark_non_differentiable(depth, aux)
        return out["rgb"].clone(), depth, aux

    @staticmethod
    def backward(ctx, g_rgb, g_depth, g_aux):
        field = ctx.field
        if g_rgb is None:
            return (None,) * (10 + len(field.flat_params()))
        tr = field.trainer
        tr._ctx, tr._out = ctx.state
        gr = tr.backward(g_rgb.contiguous())
        tr._out = None
        grads = [*gr["dense_w"], *gr["dense_b"], gr["sigma_w"], gr["sigma_b"].reshape(1)]
        grads += [st[k] for st in gr["hydra"] for k in HYDRA_KEYS]
        grads += [x for wb in gr["rgb"] for x in wb]
        grads = [g.reshape(p.shape) for g, p in zip(grads, field.flat_params())]
        return (None,) * 10 + tuple(grads)
