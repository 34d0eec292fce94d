"""Pins the upsampling restatement (oracle/upsample_ref.py) on closed forms:
  * rays_interpolate: sub-pixel 0 reproduces the input direction, sub-pixel
    k reprojects to the pixel of the input plus (0.5 if k & 1, 0.5 if k & 2),
    unit norm (catches a swapped K / R order or offset order);
  * neighbour sampling on maps affine in the pixel coordinates (bilinear
    reproduces them exactly) and on pixel-index maps (nearest returns the
    rounded coordinate), a constant depth map (the Gaussian weight of the
    independently projected depth) and R_v dir (pins the res / 2 grid scale
    with align_corners, the neighbour order and the 40-vector layout);
  * the final blend with a blending MLP whose last layer is constant:
    clamp(s(10) c0 + s(-10) c1 + 0.5 c_coarse, 0, 1), zeros for rays outside
    the depth range."""
import numpy as np
import torch

from oracle import upsample_ref as ur
from tests import fastnerf_cases as fc
from tests import upsample_cases as uc

RES = 512


def test_rays_interpolate_reprojection():
    c = uc.case(R=50, seed=3)
    d = c["ray_dir"].astype(np.float64)
    K, R = c["tar_K"].astype(np.float64), c["tar_R"].astype(np.float64)
    out = ur.rays_interpolate(d[None], K[None], R[None])[0].numpy().reshape(50, 4, 3)
    np.testing.assert_allclose(np.linalg.norm(out, axis=-1), 1.0, atol=1e-12)
    np.testing.assert_allclose(out[:, 0], d / np.linalg.norm(d, axis=-1, keepdims=True), atol=1e-12)
    q = (K @ R @ d.T).T
    base = q[:, :2] / q[:, 2:]
    for k in range(4):
        p = (K @ R @ out[:, k].T).T
        pix = p[:, :2] / p[:, 2:]
        np.testing.assert_allclose(pix, base + np.array([0.5 * (k & 1), 0.5 * (k >> 1)]),
                                   atol=1e-9)


def _project(c, v, pts):
    """Camera-space z and pixel coordinates of world points in view v."""
    cam = pts @ c["rs"][v].astype(np.float64).T + c["ts"][v].astype(np.float64)
    scr = (cam / cam[:, 2:]) @ c["ks"][v].astype(np.float64).T
    return cam[:, 2], scr[:, :2]


def test_neighbour_sampling_closed_forms():
    c = uc.case(R=80, seed=4)
    V, Hf, Wf = c["rgb_feats"].shape[0], 256, 256
    Hc, Wc, Hd, Wd = 512, 512, 256, 256
    yy, xx = np.meshgrid(np.arange(Hf, dtype=np.float64), np.arange(Wf, dtype=np.float64),
                         indexing="ij")
    a = np.linspace(-1, 1, 16)
    b = np.linspace(0.5, -0.3, 16)
    feats = (a[:, None, None] + 0.01 * b[:, None, None] * xx + 0.02 * yy / (1 + np.arange(16))[:, None, None])
    feats = np.broadcast_to(feats, (V, 16, Hf, Wf)).copy()
    yc, xc = np.meshgrid(np.arange(Hc, dtype=np.float64), np.arange(Wc, dtype=np.float64),
                         indexing="ij")
    rgbs = np.broadcast_to(np.stack([xc, yc, np.full_like(xc, 0.25)]), (V, 3, Hc, Wc)).copy()
    depths = np.full((V, 1, Hd, Wd), 2.4)
    nei = (1, 0)
    tr = {}
    _, _, pb = fc.params()
    ur.forward_upsampling(c["ray_ori"], c["ray_dir"], c["o_d"], c["o_c"], c["o_c_feat"],
                          c["tar_K"], c["tar_R"], c["ks"], c["rs"], c["ts"], feats, rgbs, depths,
                          nei, pb, res=RES, trace=tr)
    valid = tr["valid"].numpy()
    od = c["o_d"].astype(np.float64)
    assert 0 < valid.size < od.size
    dirs = tr["dirs"].numpy()                                  # [n, 4, 3]
    pts = c["ray_ori"][valid, None].astype(np.float64) + dirs * od[valid, None, None]
    got = tr["feats"].numpy()
    checked = 0
    for j, v in enumerate(nei):
        z, pix = _project(c, v, pts.reshape(-1, 3))
        z, pix = z.reshape(-1, 4), pix.reshape(-1, 4, 2)
        blk = got[..., 20 * j:20 * j + 20]
        # bilinear of an affine map at pixel u (W - 1) / res
        fx, fy = pix[..., 0] * (Wf - 1) / RES, pix[..., 1] * (Hf - 1) / RES
        inside = (fx >= 0) & (fx <= Wf - 1) & (fy >= 0) & (fy <= Hf - 1) & (z > 0.01)
        want = (a[None, None] + 0.01 * b[None, None] * fx[..., None]
                + 0.02 * fy[..., None] / (1 + np.arange(16)))
        np.testing.assert_allclose(blk[..., :16][inside], want[inside], atol=1e-9)
        np.testing.assert_allclose(blk[..., 16:19], dirs @ c["rs"][v].astype(np.float64).T,
                                   atol=1e-12)
        np.testing.assert_allclose(blk[..., 19][inside], np.exp(-200 * (z - 2.4) ** 2)[inside],
                                   atol=1e-9)
        # nearest of the pixel-index map: the rounded coordinate
        cx, cy = pix[..., 0] * (Wc - 1) / RES, pix[..., 1] * (Hc - 1) / RES
        rx, ry = np.rint(cx), np.rint(cy)
        ok = (rx >= 0) & (rx < Wc) & (ry >= 0) & (ry < Hc) & (z > 0.01)
        ok &= (np.abs(cx - np.floor(cx) - 0.5) > 1e-6) & (np.abs(cy - np.floor(cy) - 0.5) > 1e-6)
        cols = tr["cols"][j].numpy()
        np.testing.assert_array_equal(cols[..., 0][ok], rx[ok])
        np.testing.assert_array_equal(cols[..., 1][ok], ry[ok])
        np.testing.assert_array_equal(cols[..., 2][ok], 0.25)
        checked += int(inside.sum())
    assert checked > valid.size * 4            # most sub-pixels land inside both views


def test_blend_closed_form_and_depth_mask():
    c = uc.case(R=60, seed=5)
    _, _, pb = fc.params()
    pb = pb.copy()
    last = pb.size - (32 * 3 + 3)
    pb[last:last + 96] = 0
    pb[last + 96:] = np.array([10.0, -10.0, 0.0], np.float16)
    tr = {}
    out = ur.forward_upsampling(c["ray_ori"], c["ray_dir"], c["o_d"], c["o_c"], c["o_c_feat"],
                                c["tar_K"], c["tar_R"], c["ks"], c["rs"], c["ts"], c["rgb_feats"],
                                c["rgbs"], c["depths"], (0, 1), pb, res=RES, trace=tr).numpy()
    valid = tr["valid"].numpy()
    s = lambda x: 1.0 / (1.0 + np.exp(-x))
    cols = tr["cols"].numpy()
    coarse = c["o_c"][valid, None].astype(np.float64)
    want = np.clip(s(10.0) * cols[0] + s(-10.0) * cols[1] + 0.5 * coarse, 0, 1)
    np.testing.assert_allclose(out[valid], want, atol=1e-12)
    od = c["o_d"]
    bad = ~((od > 0.2) & (od < 4.0))
    assert bad.sum() >= 3 and np.all(out[bad] == 0)
    assert set(valid.tolist()) == set(np.nonzero(~bad)[0].tolist())
