"""Pins the C oracle's training-only kernels (oracle/sailor_oracle.c:
orc_octree_traversal, orc_sampling_adaptive, orc_trilinear_fwd/bwd).

  * octree_traversal: its coarse list is the coarse-only traversal's (pinned
    by brute force in test_oracle.py); every coarse hit is followed by the
    children a float64 slab test says the ray crosses, in descending child
    order (the DFS stack pops the last child first), and nothing else;
    valid_len adds the lengths times coarse size / entry size; a hand-made
    tree where a child's interval equals the coarse interval exactly (a ray
    clipping a corner) replaces the coarse entry by the child, and a ray
    crossing two children lists coarse, child 2, child 0 with length 3.
  * adaptive sampling: hand-computed counts and depths on two intervals
    (exact fill, the further-sampling fill of the smallest node, the cap),
    the noise term's sign, and per-sample interval invariants on a tree.
  * trilinear: affine corner features are reproduced exactly, one-hot
    corners give the product weights, the backward is the adjoint of the
    forward and its weights sum to 1, index -1 gives zeros.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.test_oracle import _slab64, _small_tree


def _rays(n, seed):
    rng = np.random.default_rng(seed)
    o = rng.uniform(-2.5, 2.5, (n, 3))
    o[:, 2] = -3.0
    d = rng.uniform(-0.8, 0.8, (n, 3)) - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return o.astype(np.float32), d.astype(np.float32)


@pytest.mark.parametrize("rate", [2, 4])
def test_octree_traversal_groups_match_bruteforce(rate):
    tr, bbox, vs, df, rate = _small_tree(rate=rate, p=0.02)
    pos, info = tr["node_pos"].astype(np.float64), tr["node_info"]
    o, d = _rays(300, 9)
    res = O.octree_traversal(tr["vol_coarse"], tr["node_pos"], info, bbox, vs, df, rate,
                             o[None], d[None], 256, 512)
    coarse = O.traverse(tr["vol_coarse"], tr["node_pos"], info, bbox, vs, df, rate, o[None],
                        d[None], 256, is_training=True)
    np.testing.assert_array_equal(res["idx_c"], coarse["idx"])
    np.testing.assert_array_equal(res["tmin_c"], coarse["tmin"])
    np.testing.assert_array_equal(res["tmax_c"], coarse["tmax"])
    assert res["status"][2] == 0 and res["status"][0] == coarse["gmax"]
    o64, d64 = o.astype(np.float64), d.astype(np.float64)
    n_groups = n_kids = 0
    for r in range(o.shape[0]):
        hits = [int(h) for h in res["idx"][0, r] if h >= 0]
        k, vl = 0, 0.0
        for c in [int(h) for h in res["idx_c"][0, r] if h >= 0]:
            c0, nch = int(info[c, 2]), int(info[c, 3])
            kids = set(range(c0, c0 + nch)) if c0 >= 0 else set()
            g = []
            while k < len(hits) and (hits[k] == c or hits[k] in kids):
                g.append(hits[k])
                k += 1
            assert g, "every coarse hit opens a group"
            crossed = [x for x in sorted(kids)
                       if (lambda s: s is not None and s[1] - s[0] > 1e-4)(
                           _slab64(o64[r], d64[r], pos[x, :3], pos[x, 3]))]
            assert set(crossed) <= set(g)
            body = g[1:] if g[0] == c else g
            assert body == sorted(body, reverse=True) or g[0] != c
            for j in range(k - len(g), k):
                h = hits[j]
                vl += (float(res["tmax"][0, r, j]) - float(res["tmin"][0, r, j])) * \
                    pos[c, 3] / pos[h, 3]
            n_groups += 1
            n_kids += len(body)
        assert k == len(hits)
        assert abs(vl - res["valid_len"][0, r]) <= 1e-4 * max(1.0, vl)
    assert n_groups > 50 and n_kids > 100


def _hand_tree():
    """One coarse cell [-0.5, 0.5]^3 (gindex 8) with its 8 children (gindex
    0-7, raster order, centres +-0.25, size 0.5) in a 2^3 coarse grid over
    [-0.5, 1.5]^3 (fine size 0.5, bbox min -0.25)."""
    pos = np.zeros((9, 4), np.float32)
    info = np.full((9, 4), -1, np.int32)
    for ci in range(8):
        x, y, z = ci >> 2, (ci >> 1) & 1, ci & 1
        pos[ci] = (-0.25 + 0.5 * x, -0.25 + 0.5 * y, -0.25 + 0.5 * z, 0.5)
        info[ci] = (0, 8, -1, 0)
    pos[8] = (0, 0, 0, 1.0)
    info[8] = (0, -1, 0, 8)
    vc = np.full((1, 2, 2, 2), -1, np.int32)
    vc[0, 0, 0, 0] = 8
    bbox = np.array([[[-0.25, 1.25]] * 3], np.float32)
    vs = np.array([0.5], np.float32)
    return vc, pos, info, bbox, vs


def test_octree_traversal_hand_tree_dedup_and_order():
    vc, pos, info, bbox, vs = _hand_tree()
    dg = np.array([1.0, -1.0, 0.0]) / np.sqrt(2.0)
    o = np.array([[[-0.6, -0.3, -0.3], [-0.4, -1.0, -0.4]]], np.float32)
    d = np.array([[dg, [0.0, 1.0, 0.0]]], np.float32)
    res = O.octree_traversal(vc, pos, info, bbox, vs, 4, 2, o, d, 8, 8)
    # ray 0 clips the corner: the child's interval is the coarse interval
    assert res["idx_c"][0, 0, 0] == 8 and res["idx_c"][0, 0, 1] == -1
    assert res["idx"][0, 0, 0] == 0 and res["idx"][0, 0, 1] == -1
    assert res["tmin"][0, 0, 0] == res["tmin_c"][0, 0, 0]
    assert res["tmax"][0, 0, 0] == res["tmax_c"][0, 0, 0]
    np.testing.assert_allclose([res["tmin"][0, 0, 0], res["tmax"][0, 0, 0]],
                               [0.1 * np.sqrt(2), 0.2 * np.sqrt(2)], atol=1e-6)
    np.testing.assert_allclose(res["valid_len"][0, 0], 2 * 0.1 * np.sqrt(2), atol=1e-6)
    # ray 1 runs along +y through children 0 and 2
    assert list(res["idx"][0, 1, :4]) == [8, 2, 0, -1]
    np.testing.assert_allclose(res["tmin"][0, 1, :3], [0.5, 1.0, 0.5], atol=1e-6)
    np.testing.assert_allclose(res["tmax"][0, 1, :3], [1.5, 1.5, 1.0], atol=1e-6)
    np.testing.assert_allclose(res["valid_len"][0, 1], 3.0, atol=1e-6)
    assert list(res["status"]) == [1, 3, 0]


def _two_hits(valid_len, S, noise=0.5):
    pos = np.array([[0, 0, 0, 1.0], [0, 0, 0, 0.5]], np.float32)
    hit = np.array([[[0, 1, -1]]], np.int32)
    t0 = np.array([[[1.0, 2.0, -1]]], np.float32)
    t1 = np.array([[[2.0, 2.5, -1]]], np.float32)
    nz = np.full((1, 1, S), noise, np.float32)
    vl = np.array([[valid_len]], np.float32)
    return O.sampling_adaptive(hit, t0, t1, nz, vl, pos, np.array([0.5], np.float32), S, 2)


def test_adaptive_sampling_known_answers():
    idx, t = _two_hits(2.0, 8)              # 4 + 4 samples, no fill
    assert list(idx[0, 0]) == [0] * 4 + [1] * 4
    np.testing.assert_allclose(t[0, 0], [1.2, 1.4, 1.6, 1.8, 2.1, 2.2, 2.3, 2.4], atol=1e-6)
    idx, t = _two_hits(4.0, 8)              # 2 + 2, then 4 in the smaller node
    assert list(idx[0, 0]) == [0, 0, 1, 1, 1, 1, 1, 1]
    np.testing.assert_allclose(t[0, 0], [1 + 1 / 3, 1 + 2 / 3, 2 + 0.5 / 3, 2 + 1 / 3,
                                         2.1, 2.2, 2.3, 2.4], atol=1e-6)
    idx, t = _two_hits(0.5, 8)              # the first hit asks for 16: capped at 8
    assert list(idx[0, 0]) == [0] * 8
    np.testing.assert_allclose(t[0, 0], 1 + np.arange(1, 9) / 17, atol=1e-6)
    idx, t = _two_hits(2.0, 8, noise=0.0)   # u = 0: half an offset before each point
    np.testing.assert_allclose(t[0, 0], [1.1, 1.3, 1.5, 1.7, 2.05, 2.15, 2.25, 2.35], atol=1e-6)


def test_adaptive_sampling_invariants_on_tree():
    tr, bbox, vs, df, rate = _small_tree(rate=2, p=0.02)
    o, d = _rays(200, 4)
    res = O.octree_traversal(tr["vol_coarse"], tr["node_pos"], tr["node_info"], bbox, vs, df,
                             rate, o[None], d[None], 128, 256)
    S = 32
    nz = np.random.default_rng(1).random((1, 200, S)).astype(np.float32)
    sidx, st = O.sampling_adaptive(res["idx"], res["tmin"], res["tmax"], nz, res["valid_len"],
                                   tr["node_pos"], vs, S, 1)
    n_full = 0
    for r in range(200):
        hits = res["idx"][0, r]
        if hits[0] < 0:
            assert np.all(sidx[0, r] == -1) and np.all(st[0, r] == -1)
            continue
        assert np.all(sidx[0, r] >= 0)
        n_full += 1
        for j in range(S):
            ks = np.nonzero(hits == sidx[0, r, j])[0]
            assert any(res["tmin"][0, r, k] - 1e-6 <= st[0, r, j] <= res["tmax"][0, r, k] + 1e-6
                       for k in ks)
    assert n_full > 20


def _tri_case(B=2, R=30, S=8, C=5, seed=3):
    tr, bbox, vs, df, rate = _small_tree(rate=2, p=0.02)
    o, d = _rays(B * R, seed)
    res = O.octree_traversal(tr["vol_coarse"], tr["node_pos"], tr["node_info"], bbox,
                             np.repeat(vs, 1), df, rate, o.reshape(1, -1, 3), d.reshape(1, -1, 3),
                             128, 256)
    nz = np.random.default_rng(seed).random((1, B * R, S)).astype(np.float32)
    sidx, st = O.sampling_adaptive(res["idx"], res["tmin"], res["tmax"], nz, res["valid_len"],
                                   tr["node_pos"], vs, S, 1)
    sidx, st = sidx.reshape(B, R, S), st.reshape(B, R, S)
    sidx[0, 0, 3] = -1                                  # a dropped sample
    return tr["node_pos"], o.reshape(B, R, 3), d.reshape(B, R, 3), sidx, st, C


def test_trilinear_affine_and_one_hot_closed_forms():
    pos, o, d, sidx, st, C = _tri_case()
    B, R, S = sidx.shape
    p = o[:, :, None].astype(np.float64) + d[:, :, None].astype(np.float64) * st[..., None]
    node = pos[np.maximum(sidx, 0)].astype(np.float64)
    corner = np.array([[(k >> 2) & 1, (k >> 1) & 1, k & 1] for k in range(8)], np.float64)
    cpos = node[..., None, :3] + (corner - 0.5) * node[..., None, 3:4]      # [B,R,S,8,3]
    a = np.linspace(-1, 1, C)
    b = np.stack([np.linspace(0.5, -0.2, C), np.linspace(0.1, 0.9, C), np.linspace(-0.7, 0.3, C)], 1)
    feats = a[None, :, None, None, None] + np.einsum("cj,brskj->bcrsk", b, cpos)
    out = O.trilinear_fwd(pos, o, d, sidx, st, feats)
    want = a[None, :, None, None] + np.einsum("cj,brsj->bcrs", b, p)
    ok = sidx >= 0
    np.testing.assert_allclose(out.transpose(0, 2, 3, 1)[ok], want.transpose(0, 2, 3, 1)[ok],
                               atol=2e-5)
    assert np.all(out.transpose(0, 2, 3, 1)[~ok] == 0)
    frac = (p - (node[..., :3] - node[..., 3:4] / 2)) / node[..., 3:4]          # [B,R,S,3]
    for k in range(8):
        onehot = np.zeros((B, 1, R, S, 8))
        onehot[..., k] = 1.0
        got = O.trilinear_fwd(pos, o, d, sidx, st, onehot)[:, 0]
        w = np.prod(np.where(corner[k] == 1, frac, 1 - frac), -1)
        np.testing.assert_allclose(got[ok], w[ok], atol=1e-5)


def test_trilinear_backward_is_adjoint():
    pos, o, d, sidx, st, C = _tri_case(seed=5)
    B, R, S = sidx.shape
    rng = np.random.default_rng(2)
    F = rng.normal(size=(B, C, R, S, 8)).astype(np.float32)
    G = rng.normal(size=(B, C, R, S)).astype(np.float32)
    fwd = O.trilinear_fwd(pos, o, d, sidx, st, F).astype(np.float64)
    bwd = O.trilinear_bwd(pos, o, d, sidx, st, G).astype(np.float64)
    lhs = (fwd * G).sum()
    rhs = (F.astype(np.float64) * bwd).sum()
    assert abs(lhs - rhs) <= 1e-4 * max(1.0, abs(lhs))
    ones = O.trilinear_bwd(pos, o, d, sidx, st, np.ones((B, 1, R, S), np.float32))[:, 0]
    ok = sidx >= 0
    np.testing.assert_allclose(ones.sum(-1)[ok], 1.0, atol=1e-6)
    assert np.all(ones[~ok] == 0)


def test_project_closed_form():
    """K = [[w/2, 0, w/2], [0, h/2, h/2], [0, 0, 1]], RT = [I | (0, 0, 3)]:
    the grid value of a point is X / (Z + 3) and Y / (Z + 3); corners are the
    node centre +- size / 2 in corner order 4x + 2y + z; -1 for dropped
    samples."""
    pos, o, d, sidx, st, _ = _tri_case(B=1, R=20, S=6)
    w, h = 64, 48
    K = np.array([[w / 2, 0, w / 2], [0, h / 2, h / 2], [0, 0, 1]], np.float32)[None, None]
    RT = np.concatenate([np.eye(3), [[0], [0], [3.0]]], 1).astype(np.float32)[None, None]
    xy, z = O.project(pos, o, d, sidx, st, K, RT, w, h, corners=True)
    p = o[:, :, None].astype(np.float64) + d[:, :, None].astype(np.float64) * st[..., None]
    node = pos[np.maximum(sidx, 0)].astype(np.float64)
    corner = np.array([[(k >> 2) & 1, (k >> 1) & 1, k & 1] for k in range(8)], np.float64)
    pts = np.concatenate([p[..., None, :], node[..., None, :3] + (corner - 0.5) *
                          node[..., None, 3:4]], -2)                        # [B,R,S,9,3]
    ok = sidx >= 0
    zz = pts[..., 2] + 3.0
    np.testing.assert_allclose(xy[:, 0, ..., 0][ok], (pts[..., 0] / zz)[ok], atol=2e-6)
    np.testing.assert_allclose(xy[:, 0, ..., 1][ok], (pts[..., 1] / zz)[ok], atol=2e-6)
    np.testing.assert_allclose(z[:, 0, ..., 0][ok], zz[ok], atol=1e-6)
    assert np.all(xy[:, 0][~ok] == -1) and np.all(z[:, 0][~ok] == -1)
    xy1, z1 = O.project(pos, o, d, sidx, st, K, RT, w, h, corners=False)
    np.testing.assert_array_equal(xy1, xy[..., 0, :])
    np.testing.assert_array_equal(z1, z[..., 0, :])
