"""Pins the FastNerf restatement oracle/fastnerf_ref.py (CPU):
  * packed-parameter layout: the unpackers consume exactly the counts the
    reference kernels hard-code (13696 / 22081 / 440 / 22521, 5822 / 11644 /
    4163 / 15807, 15075) and place each layer where those offsets say;
  * its compositing equals the C oracle's volume_rendering_occ restatement
    (itself pinned in test_oracle.py) on the same inputs;
  * the Hydra encoder is invariant to permuting the non-first views and
    sees a view's token only through its own row (brute-force check of the
    cross-view sum against an explicit loop over views);
  * infer_density of a point equals the occupancy infer_nerf composites
    (the same feature MLP, view mean and occupancy MLP);
  * with all views identical, the view mean equals the single-view value.
"""
import numpy as np
import torch

from oracle import fastnerf_ref as fr
from oracle import oracle as orc
from tests import fastnerf_cases as fc


def test_packed_layout_offsets():
    pd, pc, pb = fc.params()
    p = fr._params(pd, torch.float64)
    feats, sigma, geo = fr.unpack_density(p)
    # the occupancy MLP starts at 13696 and the geometry MLP at 22081
    assert torch.equal(sigma[0][0].reshape(-1)[:5], p[fr.FEAT_PARAMS:fr.FEAT_PARAMS + 5])
    assert torch.equal(geo[0][0].reshape(-1)[:5], p[fr.GEO_OFFSET:fr.GEO_OFFSET + 5])
    assert feats[0][0].shape == (64, 18) and geo[1][0].shape == (8, 16)
    q = fr._params(pc, torch.float64)
    stacks, rgb = fr.unpack_color(q)
    assert torch.equal(stacks[1]["q"].reshape(-1)[:5], q[fr.HYDRA_STACK_PARAMS:fr.HYDRA_STACK_PARAMS + 5])
    assert torch.equal(rgb[0][0].reshape(-1)[:5], q[fr.HYDRA_PARAMS:fr.HYDRA_PARAMS + 5])
    assert stacks[0]["fc"].shape == (30, 32) and stacks[0]["w2"].shape == (30, 32)
    L = fr.unpack_blend(fr._params(pb, torch.float64))
    assert L[3][0].shape == (64, 62) and L[5][0].shape == (3, 32)


def test_composite_matches_c_oracle():
    rng = np.random.default_rng(3)
    t, *_ = fc.nerf_inputs(1, 40, 24, 1, 1)
    a = rng.uniform(0, 0.9, t.shape).astype(np.float32)
    f = rng.normal(0, 1, t.shape + (30,)).astype(np.float32)
    got_f, got_d = fr.composite(torch.tensor(t[0], dtype=torch.float64),
                                torch.tensor(a[0], dtype=torch.float64),
                                torch.tensor(f[0], dtype=torch.float64))
    of, od, _, _ = orc.render_occ(t, a, f)
    np.testing.assert_allclose(got_f.numpy(), of[0], rtol=1e-5, atol=1e-5)
    np.testing.assert_allclose(got_d.numpy(), od[0], rtol=1e-5, atol=1e-5)


def test_hydra_view_permutation_and_cross_view_sum():
    _, pc, _ = fc.params()
    rng = np.random.default_rng(8)
    tok = torch.tensor(rng.normal(0, 1, (6, 4, 30)))
    f0, rgb0 = fr.hydra_rgb(tok, pc)
    perm = tok[:, [0, 3, 1, 2]]
    f1, rgb1 = fr.hydra_rgb(perm, pc)
    np.testing.assert_allclose(f0.numpy(), f1.numpy(), rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(rgb0.numpy(), rgb1.numpy(), rtol=1e-12, atol=1e-12)
    # one stack written with explicit per-view loops
    stacks, _ = fr.unpack_color(fr._params(pc, torch.float64))
    st = stacks[0]
    x = tok[0]
    qs, ks, vs = [], [], []
    for v in range(4):
        q = st["q"] @ x[v]
        k = st["k"] @ x[v]
        qs.append(q / q.norm())
        ks.append(k / k.norm())
        vs.append(st["v"] @ x[v])
    kv = sum(ks[v] * vs[v] for v in range(4))
    out0 = st["fc"] @ (qs[0] * kv) + x[0]
    y0 = st["w2"] @ torch.relu(st["w1"] @ out0 + st["b1"]) + st["b2"] + out0
    x1 = tok[:1]
    q = x1 @ st["q"].T
    k = x1 @ st["k"].T
    q = q / torch.sqrt((q * q).sum(-1, keepdim=True))
    k = k / torch.sqrt((k * k).sum(-1, keepdim=True))
    kvb = (k * (x1 @ st["v"].T)).sum(1, keepdim=True)
    res = x1 + (q * kvb) @ st["fc"].T
    yb = torch.relu(res @ st["w1"].T + st["b1"]) @ st["w2"].T + st["b2"] + res
    np.testing.assert_allclose(yb[0, 0].numpy(), y0.numpy(), rtol=1e-12, atol=1e-12)


def test_density_consistent_with_nerf_and_view_mean():
    pd, pc, _ = fc.params()
    B, R, S, Vg, Vt = 1, 5, 6, 3, 2
    t, z, ps, geo, rgb, frgb, dirs = fc.nerf_inputs(B, R, S, Vg, Vt)
    occ = fr.infer_density(z, ps, geo, pd, B, R * S, Vg)          # [B, 1, R*S]
    f, _ = fr.density_feats(z, ps, geo, pd)
    m = f.reshape(B, Vg, R * S, 64).mean(1).reshape(-1, 64)
    np.testing.assert_allclose(occ.reshape(-1).numpy(), fr.occupancy(m, pd).numpy(), rtol=1e-12)
    # identical views: the mean is the single view's features
    z1 = np.tile(z[:R * S], Vg)
    p1 = np.tile(ps[:R * S], Vg)
    g1 = np.tile(geo[:R * S], (Vg, 1))
    occ1 = fr.infer_density(z1, p1, g1, pd, 1, R * S, Vg)
    occ_single = fr.infer_density(z[:R * S], ps[:R * S], geo[:R * S], pd, 1, R * S, 1)
    np.testing.assert_allclose(occ1.numpy(), occ_single.numpy(), rtol=1e-12)
