"""Pins oracle/gather_ref.py (the per-view sample features before FastNerf)
with closed forms (CPU):
  * a point that projects onto pixel (i, j) of a map (align_corners=True:
    grid (2i/(W-1) - 1)) samples exactly map[j, i]; its t-psdf and
    normalised depth follow from the formula;
  * a point behind the camera or with t outside [0.1, 4.0] samples zeros
    (grid -2); for t outside the range the projected depth is -2;
  * a point half-way between pixels samples their mean;
  * camera-space directions are R d.
"""
import numpy as np
import torch

from oracle import gather_ref as gr


def _setup(W=9, H=7, V=2, seed=0):
    rng = np.random.default_rng(seed)
    K = np.tile(np.eye(3), (1, V, 1, 1))
    RT = np.zeros((1, V, 3, 4))
    RT[0, :, :, :3] = np.eye(3)
    geo = rng.normal(size=(1, V, 16, H, W))
    dep = rng.uniform(0.5, 3.0, (1, V, 1, H, W))
    rgbf = rng.normal(size=(1, V, 16, H, W))
    rgb = rng.uniform(0, 1, (1, V, 3, H, W))
    mid = rng.uniform(1, 2, (1, V))
    dist = rng.uniform(1, 2, (1, V))
    return K, RT, geo, dep, rgbf, rgb, mid, dist


def test_pixel_centre_closed_form():
    W, H = 9, 7
    K, RT, geo, dep, rgbf, rgb, mid, dist = _setup(W, H)
    res = 2 * (W - 1)          # grid half-width W-1: screen x = 2 i hits pixel column i
    K[0, :, 1, 1] = (W - 1) / (H - 1)  # screen y = 2 j (W-1)/(H-1) hits row j
    i, j, z = 3, 2, 0.4        # t = z |d| / d_z = 2.9, inside [0.1, 4]
    # ray along +z from the origin of a point (x, y, z) with x/z = 2i, y/z = 2j
    d = np.array([2 * i, 2 * j * (H - 1) / (W - 1) * (W - 1) / (H - 1), 1.0])
    d = d / np.linalg.norm(d)
    t = z / d[2]
    rays = np.concatenate([np.zeros(3), d])[None, None]
    dists = np.full((1, 1, 1), t)
    out = gr.sample_features(rays, dists, K, RT, geo, dep, rgbf, rgb, mid, dist, res=res)
    for v in range(2):
        np.testing.assert_allclose(out["geo"][0, v, 0, 0].numpy(), geo[0, v, :, j, i], atol=1e-9)
        np.testing.assert_allclose(out["rgbf"][0, v, 0, 0].numpy(), rgbf[0, v, :, j, i], atol=1e-9)
        np.testing.assert_allclose(out["rgb"][0, v, 0, 0].numpy(), rgb[0, v, :, j, i], atol=1e-9)
        tps = np.clip(z - dep[0, v, 0, j, i], -0.01, 0.01) * 100.0
        np.testing.assert_allclose(float(out["tpsdf"][0, v, 0, 0]), tps, atol=1e-9)
        zn = (z - mid[0, v]) / (dist[0, v] / 2)
        np.testing.assert_allclose(float(out["z"][0, v, 0, 0]), zn, atol=1e-9)
        np.testing.assert_allclose(out["dir"][0, v, 0, 0].numpy(), d, atol=1e-12)


def test_invalid_points_and_midpoint():
    W, H = 9, 7
    K, RT, geo, dep, rgbf, rgb, mid, dist = _setup(W, H)
    res = 2 * (W - 1)
    K[0, :, 1, 1] = (W - 1) / (H - 1)
    rays = np.zeros((1, 3, 6))
    rays[0, :, 5] = 1.0                       # along +z
    rays[0, 0, 2] = -5.0                      # point behind the camera at t = 2
    rays[0, 2, 3:] = [2 * 2.5 / np.sqrt(1 + 25), 0.0, 1 / np.sqrt(1 + 25)]  # between columns 2, 3
    dists = np.array([[[2.0], [5.0], [1.5]]])  # ray 1: t beyond 4.0
    out = gr.sample_features(rays, dists, K, RT, geo, dep, rgbf, rgb, mid, dist, res=res)
    for v in range(2):
        assert float(out["geo"][0, v, 0, 0].abs().sum()) == 0.0
        assert float(out["geo"][0, v, 1, 0].abs().sum()) == 0.0
        np.testing.assert_allclose(float(out["z"][0, v, 1, 0]), (-2 - mid[0, v]) / (dist[0, v] / 2))
        np.testing.assert_allclose(float(out["tpsdf"][0, v, 1, 0]), -1.0)
        # column 2.5 of row 0: the mean of pixels (2, 0) and (3, 0)
        np.testing.assert_allclose(out["geo"][0, v, 2, 0].numpy(),
                                   0.5 * (geo[0, v, :, 0, 2] + geo[0, v, :, 0, 3]), atol=1e-9)
