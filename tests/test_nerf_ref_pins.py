"""Pins the MLP and Hydra layers of the float64 model restatement
oracle/nerf_ref.py (the parts its encoding / gather / compositing tests in
test_oracle.py leave open):

  * hydra_encode, identity projections (q = k = v = fc = I) and a zero
    feed-forward: each stack is x + x_hat * sum over views of (x_hat * x)
    (x_hat = x / |x|), the closed form of HydraAttention
    (SRONet/aggregation.py:101-128) with HydraForward (:164-171) reduced to
    its residual -- catches a kv sum over the wrong axis, a missing
    normalisation or residual;
  * hydra_encode, zero projections (q_hat = 0, so the attention adds 0) and
    the feed-forward w1 = I, b1 = -0.1, w2 = 2 I, b2 = 0.05: each stack is
    x + 2 relu(x - 0.1) + 0.05 -- catches a dropped FFN residual or bias;
  * permuting the views permutes the encoder's output tokens the same way
    (random weights);
  * render_rays, layer by layer against torch.nn.Linear built from the same
    weights (the library routine the reference's modules are made of:
    SRONet/nerf.py:139-152): pre-activation l = Linear_l(relu(pre_{l-1})),
    occupancy = sigmoid(mean over views of relu(pre_3) . sigma_w + sigma_b),
    rgb = sigmoid(Linear(relu(Linear(relu(Linear(token_0)))))) of the encoded
    tokens -- catches a transposed weight, a misplaced bias or activation."""
import numpy as np
import torch
import torch.nn as nn

from oracle import nerf_ref

D = torch.float64


def _stack(F, q=None, k=None, v=None, fc=None, w1=None, b1=None, w2=None, b2=None, inner=None):
    inner = inner or F
    z = lambda *s: torch.zeros(*s, dtype=D)
    return dict(q=z(inner, F) if q is None else q, k=z(inner, F) if k is None else k,
                v=z(inner, F) if v is None else v, fc=z(F, inner) if fc is None else fc,
                w1=z(inner, F) if w1 is None else w1, b1=z(inner) if b1 is None else b1,
                w2=z(F, inner) if w2 is None else w2, b2=z(F) if b2 is None else b2)


def test_hydra_identity_attention_closed_form():
    rng = np.random.default_rng(0)
    N, V, F = 9, 4, 8
    x = rng.normal(size=(N, V, F))
    I = torch.eye(F, dtype=D)
    st = _stack(F, q=I, k=I, v=I, fc=I)
    got = nerf_ref.hydra_encode([st, st], torch.tensor(x)).numpy()
    want = x
    for _ in range(2):
        xh = want / np.linalg.norm(want, axis=-1, keepdims=True)
        kv = (xh * want).sum(axis=1, keepdims=True)          # sum over the views
        want = want + xh * kv
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12)


def test_hydra_feed_forward_closed_form():
    rng = np.random.default_rng(1)
    N, V, F = 7, 3, 6
    x = rng.normal(size=(N, V, F))
    I = torch.eye(F, dtype=D)
    st = _stack(F, w1=I, b1=torch.full((F,), -0.1, dtype=D), w2=2 * I,
                b2=torch.full((F,), 0.05, dtype=D))
    got = nerf_ref.hydra_encode([st], torch.tensor(x)).numpy()
    want = x + 2 * np.maximum(x - 0.1, 0) + 0.05
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12)


def test_hydra_view_permutation_equivariance():
    rng = np.random.default_rng(2)
    N, V, F, inner = 5, 4, 8, 6
    r = lambda *s: torch.tensor(rng.normal(0, 0.5, s))
    st = [_stack(F, q=r(inner, F), k=r(inner, F), v=r(inner, F), fc=r(F, inner),
                 w1=r(inner, F), b1=r(inner), w2=r(F, inner), b2=r(F), inner=inner)
          for _ in range(2)]
    x = r(N, V, F)
    perm = [2, 0, 3, 1]
    a = nerf_ref.hydra_encode(st, x)[:, perm]
    b = nerf_ref.hydra_encode(st, x[:, perm])
    np.testing.assert_allclose(a.numpy(), b.numpy(), rtol=1e-12, atol=1e-12)
    assert not torch.allclose(a[:, 0], a[:, 1])             # the views do differ


def _linear(w, b):
    lin = nn.Linear(w.shape[1], w.shape[0]).to(D)
    with torch.no_grad():
        lin.weight.copy_(torch.as_tensor(w, dtype=D))
        lin.bias.copy_(torch.as_tensor(b, dtype=D))
    return lin


def test_render_rays_layers_match_nn_linear():
    from sailor_amd import model as M
    w = M.NerfWeights.init(32, 128, perturb_std=0.05)
    rng = np.random.default_rng(3)
    B, R, S, V = 1, 5, 6, 4
    t = np.sort(rng.uniform(2.0, 3.0, (B, R, S)), -1)
    rays = np.zeros((B, R, 6))
    rays[..., :2] = rng.uniform(-0.2, 0.2, (B, R, 2))
    rays[..., 2] = 2.5
    rays[..., 5] = -1.0
    K = np.tile(np.array([[100., 0, 64], [0, 100., 64], [0, 0, 1]]), (B, V, 1, 1))
    RT = np.tile(np.concatenate([np.eye(3), np.array([[0], [0], [3.0]])], 1), (B, V, 1, 1))
    fm = rng.standard_normal((B, V, 32, 16, 16))
    bbox = np.array([[[-1, 1], [-1, 1], [-1, 1]]], np.float64)
    zdir = np.array([[0, 0, 1.0]])
    tr = {}
    rgb = nerf_ref.render_rays(w, t, rays, K, RT, fm, bbox, zdir, (128, 128), trace=tr)[0]
    layers = [_linear(wl, bl) for wl, bl in zip(w.dense_w, w.dense_b)]
    with torch.no_grad():
        h_sum = 0
        for v in range(V):
            h = tr["x0"][v]
            for li, lin in enumerate(layers):
                pre = lin(h)
                np.testing.assert_allclose(tr["pre"][v][li].detach().numpy(), pre.numpy(),
                                           rtol=1e-12, atol=1e-12)
                h = torch.relu(pre)
            h_sum = h_sum + h
        sig = torch.tensor(w.sigma_w[None], dtype=D)
        occ = torch.sigmoid(nn.functional.linear(h_sum / V, sig)[..., 0] + w.sigma_b)
        np.testing.assert_allclose(tr["occ"].numpy(), occ.numpy(), rtol=1e-12, atol=1e-12)
        hyd = [{k: torch.tensor(a, dtype=D) for k, a in st.items()} for st in w.hydra]
        tok0 = nerf_ref.hydra_encode(hyd, tr["toks"])[:, 0]
        (w1, b1), (w2, b2), (w3, b3) = w.rgb
        head = nn.Sequential(_linear(w1, b1), nn.ReLU(), _linear(w2, b2), nn.ReLU(),
                             _linear(w3, b3), nn.Sigmoid())
        np.testing.assert_allclose(rgb[0].numpy(), head(tok0).numpy(), rtol=1e-12, atol=1e-12)
    assert float(occ.std()) > 1e-3                            # a non-degenerate case


def test_sixteen_bit_product_emulation():
    """_mm's 16-bit emulation: bf16 rounds both operands as they are; fp16
    divides each row of x by the power of two >= its max |x|, which for
    rows inside fp16's normal range changes no rounding (checked against the
    plain rounded product) and keeps a row of magnitude 1e5 finite where
    plain rounding overflows to inf."""
    rng = np.random.default_rng(7)
    x = torch.tensor(rng.uniform(0.5, 1.0, (6, 32)) * rng.choice([-1, 1], (6, 32)), dtype=D)
    w = torch.tensor(rng.standard_normal((5, 32)) * 0.2, dtype=D)
    rd = lambda z, dt: z.to(dt).to(D)
    for dt in (torch.bfloat16, torch.float16):
        want = rd(x, dt) @ rd(w, dt).T
        np.testing.assert_allclose(nerf_ref._mm(x, w, dt).numpy(), want.numpy(), rtol=0, atol=0)
    big = x.clone()
    big[2] *= 1e5
    plain = rd(big, torch.float16) @ rd(w, torch.float16).T
    assert not bool(torch.isfinite(plain[2]).all())
    got = nerf_ref._mm(big, w, torch.float16)
    assert bool(torch.isfinite(got).all())
    np.testing.assert_allclose(got[2].numpy(), (big[2] @ w.T).numpy(), rtol=5e-3, atol=1e-3 * 1e5)
    np.testing.assert_allclose(got[[0, 1, 3, 4, 5]].numpy(), plain[[0, 1, 3, 4, 5]].numpy(),
                               rtol=0, atol=0)
