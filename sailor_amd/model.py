"""Radiance model of the benchmark configurations and its weight packing.

Structure (reference: SRONet/nerf.py TransNerf.density / forward_color and the
FastNerf inference path c_lib/FastNerf/src/infer_mlp.cu:638-789):
  per (sample, view):  in = [pe(x) (63) | f_v(x) (F)]  ->  H -> H -> H -> H
                       (ReLU after every layer; "hidden H x 4")
  per sample:          alpha = sigmoid(w_s . mean_v(h_v) + b_s)   (occupancy)
  per ray:             token_v = sum_i w_i f_v(x_i) (front-to-back weights)
                       2-stack Hydra attention over the views (inner 32)
                       rgb = MLP(token_0): F -> 64 -> 32 -> 3, sigmoid
pe is the degree-10 frequency encoding of x normalised to the volume box, f_v
the bilinear sample of view v's feature map at the projection of x.

Weights are Xavier-uniform (seed 3130), biases zero.  `pack()` writes the
dense layers as MFMA A-fragments in the order the fused kernel consumes them
(layout in sailor_amd/csrc/mlp_layout.h): per layer and per wave (a quarter
of the output rows each), per MFMA step, per lane.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
import torch

PE_DIMS = 63
PE_SLOTS = 64
HYDRA_INNER = 32
HYDRA_STACKS = 2
RGB_H1, RGB_H2 = 64, 32

PREC_IDS = {"fp32": 0, "bf16": 1, "fp16": 2}


def _xavier(rng, out_f, in_f):
    a = np.sqrt(6.0 / (in_f + out_f))
    return rng.uniform(-a, a, (out_f, in_f)).astype(np.float32)


@dataclass
class NerfWeights:
    F: int
    H: int
    dense_w: list = field(default_factory=list)   # [H, 63+F], [H, H] x3
    dense_b: list = field(default_factory=list)   # [H] x4
    sigma_w: np.ndarray = None                    # [H]
    sigma_b: float = 0.0
    hydra: list = field(default_factory=list)     # per stack dict
    rgb: list = field(default_factory=list)       # [(W, b)] x3

    @staticmethod
    def init(F, H, seed=3130, perturb_std=0.0):
        rng = np.random.default_rng(seed)
        w = NerfWeights(F, H)
        ins = [PE_DIMS + F, H, H, H]
        for i in range(4):
            w.dense_w.append(_xavier(rng, H, ins[i]))
            w.dense_b.append(np.zeros(H, np.float32))
        w.sigma_w = _xavier(rng, 1, H)[0]
        w.sigma_b = 0.0
        for _ in range(HYDRA_STACKS):
            st = dict(q=_xavier(rng, HYDRA_INNER, F), k=_xavier(rng, HYDRA_INNER, F),
                      v=_xavier(rng, HYDRA_INNER, F), fc=_xavier(rng, F, HYDRA_INNER),
                      w1=_xavier(rng, HYDRA_INNER, F), b1=np.zeros(HYDRA_INNER, np.float32),
                      w2=_xavier(rng, F, HYDRA_INNER), b2=np.zeros(F, np.float32))
            w.hydra.append(st)
        w.rgb = [(_xavier(rng, RGB_H1, F), np.zeros(RGB_H1, np.float32)),
                 (_xavier(rng, RGB_H2, RGB_H1), np.zeros(RGB_H2, np.float32)),
                 (_xavier(rng, 3, RGB_H2), np.zeros(3, np.float32))]
        if perturb_std > 0:
            w.perturb(perturb_std, seed + 1)
        return w

    def perturb(self, std, seed):
        rng = np.random.default_rng(seed)
        add = lambda a: (a + rng.normal(0, std, a.shape)).astype(np.float32)
        self.dense_w = [add(a) for a in self.dense_w]
        self.sigma_w = add(self.sigma_w)
        self.hydra = [{k: add(v) for k, v in st.items()} for st in self.hydra]
        self.rgb = [(add(a), add(b)) for a, b in self.rgb]

    # ------------------------------------------------------------------
    def small_buffer(self):
        """Flat fp32 buffer: occupancy head, Hydra stacks, RGB head (the
        order of mlp_layout.h small_offsets)."""
        parts = [self.sigma_w, np.array([self.sigma_b, 0, 0, 0], np.float32)]
        for st in self.hydra:
            parts += [st["q"], st["k"], st["v"], st["fc"], st["w1"], st["b1"], st["w2"],
                      st["b2"]]
        (w1, b1), (w2, b2), (w3, b3) = self.rgb
        parts += [w1, b1, w2, b2, w3, np.concatenate([b3, np.zeros(1, np.float32)])]
        return np.concatenate([np.asarray(p, np.float32).reshape(-1) for p in parts])

    def split_small(self, buf):
        """Inverse of small_buffer: views of a flat buffer (numpy or torch) in
        the layout of this model's parameters, as dict(sigma_w, sigma_b,
        hydra=[dict per stack], rgb=[(W, b)] x3)."""
        pos = 0

        def take(shape):
            nonlocal pos
            n = int(np.prod(shape))
            v = buf[pos:pos + n].reshape(shape)
            pos += n
            return v
        out = dict(sigma_w=take(self.sigma_w.shape))
        out["sigma_b"] = take((4,))[0]
        out["hydra"] = [{k: take(st[k].shape) for k in ("q", "k", "v", "fc", "w1", "b1", "w2",
                                                        "b2")} for st in self.hydra]
        rgb = []
        for i, (w, b) in enumerate(self.rgb):
            wv = take(w.shape)
            bv = take((4,))[:3] if i == 2 else take(b.shape)
            rgb.append((wv, bv))
        out["rgb"] = rgb
        if pos != len(buf):
            raise RuntimeError(f"small buffer has {len(buf)} floats, layout uses {pos}")
        return out

    def pack(self, prec="fp32", device="cuda"):
        """-> (w_dense tensor, b_dense [4,H] fp32, w_small fp32) on `device`."""
        frags = [pack_layer(natural_first_layer(self.dense_w[0], self.F), prec)]
        for i in range(1, 4):
            frags.append(pack_layer(self.dense_w[i], prec))
        dense = np.concatenate([f.reshape(-1) for f in frags])
        if prec == "fp32":
            wd = torch.from_numpy(dense.astype(np.float32))
        else:
            dt = torch.bfloat16 if prec == "bf16" else torch.float16
            wd = torch.from_numpy(dense.astype(np.float32)).to(dt)
        bd = torch.from_numpy(np.stack(self.dense_b).astype(np.float32))
        ws = torch.from_numpy(self.small_buffer())
        return wd.to(device).contiguous(), bd.to(device).contiguous(), ws.to(device).contiguous()


def natural_first_layer(W0, F):
    """First-layer weights in the kernel's input order: feature channels,
    then the 63 PE dims, then one zero column (mlp_layout.h)."""
    H = W0.shape[0]
    return np.concatenate([W0[:, PE_DIMS:PE_DIMS + F], W0[:, :PE_DIMS],
                           np.zeros((H, 1), np.float32)], 1)


def kmap(prec, K):
    """Reduction index of every packed slot [steps, 64 lanes, EPL] for a layer
    with K inputs (mlp_layout.h).  fp32 (16x16x4 MFMA): one slot per lane and
    step; 16-bit (32x32x16 MFMA): 8 slots per lane and step."""
    if prec == "fp32":
        g = np.arange(64) // 16
        s = np.arange(K // 4)[:, None]
        return (16 * (s >> 2) + 4 * g[None, :] + (s & 3))[:, :, None]
    h = np.arange(64) // 32
    u = np.arange(K // 16)[:, None, None]
    j = np.arange(8)[None, None, :]
    return 16 * u + 8 * h[None, :, None] + j


def unit_rows(prec, H, waves=4):
    """Output unit computed by A row m (= lane % 16 or lane % 32) of m-tile mt
    in wave w, as [waves, 1, 64 lanes, m-tiles per wave, 1] (mlp.hip
    unit_of / unit_of32): each lane's accumulators then hold runs of
    consecutive units of one column."""
    lane = np.arange(64)
    w = np.arange(waves)[:, None, None, None, None]
    if prec == "fp32":
        MTW = H // (16 * waves)
        mt = np.arange(MTW)[None, None, None, :, None]
        m = (lane % 16)[None, None, :, None, None]
        return (H // waves) * w + 4 * MTW * (m // 4) + 4 * mt + m % 4
    M32 = H // (32 * waves)
    mt = np.arange(M32)[None, None, None, :, None]
    m = (lane % 32)[None, None, :, None, None]
    # D row m = 8q + 4h + i of a 32x32 tile sits in accumulator 4q + i of the
    # lanes with l/32 = h; unit 32 mt + 16 h + 4 q + i makes accumulator r of
    # those lanes unit 32 mt + 16 h + r
    return (H // waves) * w + 32 * mt + 16 * ((m // 4) % 2) + 4 * (m // 8) + m % 4


def pack_layer(Wn, prec, waves=4):
    """A fragments of a natural-order layer Wn [H, K] as
    [waves, steps, 64 lanes, m-tiles per wave, EPL]."""
    H, K = Wn.shape
    col = kmap(prec, K)                                 # [steps, 64, EPL]
    rows = unit_rows(prec, H, waves)
    cols = col[None, :, :, None, :]
    rows, cols = np.broadcast_arrays(rows, cols)
    return Wn[rows, cols]


@dataclass
class ModelConfig:
    F: int
    H: int
    Hf: int
    Wf: int


CONFIG_MODELS = {
    "c1": ModelConfig(F=32, H=128, Hf=128, Wf=128),
    "c2": ModelConfig(F=64, H=256, Hf=256, Wf=256),
}


def _packed_layers(rng, layers):
    """Packed [W (out x in) row-major | b (out)] blocks in order, float16."""
    parts = []
    for n_in, n_out, bias in layers:
        a = np.sqrt(6.0 / (n_in + n_out))
        parts.append(rng.uniform(-a, a, n_out * n_in))
        if bias:
            parts.append(rng.normal(0, 0.05, n_out))
    return np.concatenate(parts).astype(np.float16)


def fastnerf_params(seed=11):
    """Synthetic FastNerf parameter vectors in the reference's packed half
    format (gui/gui_render.py:658-685): density (22521), colour (15807) and
    blending (15075) halves, Xavier-uniform weights and N(0, 0.05) biases."""
    rng = np.random.default_rng(seed)
    pd = _packed_layers(rng, [(18, 64, 1)] + [(64, 64, 1)] * 3
                        + [(64, 64, 1), (64, 64, 1), (64, 1, 1), (18, 16, 1), (16, 8, 1)])
    stack = [(30, 32, 0)] * 3 + [(32, 30, 0), (30, 32, 1), (32, 30, 1)]
    pc = _packed_layers(rng, stack + stack + [(30, 64, 1), (64, 32, 1), (32, 3, 1)])
    pb = _packed_layers(rng, [(40, 64, 1), (64, 64, 1), (64, 32, 1), (62, 64, 1), (64, 32, 1),
                              (32, 3, 1)])
    return pd, pc, pb
