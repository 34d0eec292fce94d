"""GPU parity of the per-view sample features (sailor_sample_features) and
of the GUI's FastNerf step built on them (sailor_amd.fastnerf.forward_nerf)
against the float64 restatements oracle/gather_ref.py and
oracle/fastnerf_ref.py, on ring cameras with maps of different resolutions,
samples behind / outside the depth range and dropped rays.  Tolerance: fp32
vs float64, 1e-4 absolute; 5e-3 on t-psdf, which scales depth differences by
100: the random depth maps change by up to 2 per pixel, so the ~1e-5 pixel
difference of the fp32 projection moves it by up to ~3e-3."""
import numpy as np
import pytest
import torch

from oracle import fastnerf_ref as fr
from oracle import gather_ref as gr
from tests import fastnerf_cases as fc

pytestmark = pytest.mark.gpu


def _case(B=1, V=4, R=60, S=8, seed=2):
    from sailor_amd import scenes
    rng = np.random.default_rng(seed)
    cams = scenes.ring_cameras(V, 600.0, 512, 512)
    K = np.stack([c.K for c in cams])[None].repeat(B, 0)
    RT = np.stack([np.concatenate([c.R, c.T[:, None]], 1) for c in cams])[None].repeat(B, 0)
    tgt = scenes.target_camera(512, 512, 600.0)
    o = -tgt.R.T @ tgt.T
    look = np.asarray(scenes.LOOK_AT, np.float64)
    dirs = look + rng.normal(0, 0.3, (B, R, 3)) - o
    dirs /= np.linalg.norm(dirs, axis=-1, keepdims=True)
    rays = np.concatenate([np.broadcast_to(o, (B, R, 3)), dirs], -1).astype(np.float32)
    dists = np.sort(rng.uniform(0.05, 4.6, (B, R, S)), -1).astype(np.float32)
    maps = dict(geo=rng.normal(size=(B, V, 16, 128, 96)), dep=rng.uniform(1.5, 3.5, (B, V, 1, 200, 160)),
                rgbf=rng.normal(size=(B, V, 16, 64, 64)), rgb=rng.uniform(0, 1, (B, V, 3, 512, 512)))
    maps = {k: v.astype(np.float32) for k, v in maps.items()}
    mid = rng.uniform(2.0, 3.0, (B, V)).astype(np.float32)
    dist = rng.uniform(1.0, 2.0, (B, V)).astype(np.float32)
    return rays, dists, K.astype(np.float32), RT.astype(np.float32), maps, mid, dist


def _cl(x):
    """[B, V, C, H, W] numpy -> channels-last [B, V, H, W, C] via the HIP transpose."""
    from sailor_amd.fastnerf import to_channels_last
    B, V, C, H, W = x.shape
    t = to_channels_last(torch.from_numpy(x).cuda().reshape(B * V, C, H, W))
    return t.reshape(B, V, H, W, C)


@pytest.mark.parametrize("B", [1, 2])
def test_sample_features_match_float64(B):
    from sailor_amd.fastnerf import sample_features
    rays, dists, K, RT, maps, mid, dist = _case(B=B)
    cu = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    got = sample_features(cu(rays), cu(dists), cu(K), cu(RT), _cl(maps["geo"]),
                          cu(maps["dep"][:, :, 0]), _cl(maps["rgbf"]), _cl(maps["rgb"]), cu(mid),
                          cu(dist), res=512)
    torch.cuda.synchronize()
    ref = gr.sample_features(rays, dists, K, RT, maps["geo"], maps["dep"], maps["rgbf"],
                             maps["rgb"], mid, dist, res=512)
    names = ("z", "tpsdf", "geo", "rgbf", "rgb", "dir")
    for name, g in zip(names, got):
        tol = 5e-3 if name == "tpsdf" else 1e-4
        np.testing.assert_allclose(g.double().cpu().numpy(), ref[name].numpy(), atol=tol,
                                   err_msg=name)
    # the case covers samples inside the maps and dropped ones
    geo = ref["geo"].abs().sum(-1)
    assert (geo > 0).float().mean() > 0.2 and (geo == 0).any()


def test_forward_nerf_chain_matches_float64():
    from sailor_amd.fastnerf import forward_nerf
    rays, dists, K, RT, maps, mid, dist = _case(B=1, R=40, S=8, seed=4)
    dists[0, ::5] = -1.0                                   # rays without samples
    pd, pc, _ = fc.params()
    cu = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    c, d, f = forward_nerf(cu(rays[..., :3]), cu(rays[..., 3:]), cu(dists), cu(K), cu(RT),
                           _cl(maps["geo"]), cu(maps["dep"][:, :, 0]), _cl(maps["rgbf"]),
                           _cl(maps["rgb"]), cu(mid), cu(dist), cu(pd).half(), cu(pc).half(),
                           precision="fp32")
    torch.cuda.synchronize()
    valid = np.nonzero(np.all(dists[0] != -1, -1))[0]
    sf = gr.sample_features(rays[:, valid], dists[:, valid], K, RT, maps["geo"], maps["dep"],
                            maps["rgbf"], maps["rgb"], mid, dist)
    Rv, S, V = valid.size, dists.shape[-1], K.shape[1]
    rd, rc, rf = fr.infer_nerf(dists[:, valid], sf["z"].numpy(), sf["tpsdf"].numpy(),
                               sf["geo"].numpy(), sf["rgb"].numpy(), sf["rgbf"].numpy(),
                               sf["dir"].numpy(), pd, pc, 1, Rv, S, V, V)
    np.testing.assert_allclose(c[0, valid].double().cpu().numpy(), rc[0].numpy(), atol=2e-4)
    np.testing.assert_allclose(d[0, valid].double().cpu().numpy(), rd[0].numpy(), atol=2e-4)
    np.testing.assert_allclose(f[0, valid].double().cpu().numpy(), rf[0].numpy(), atol=2e-3,
                               equal_nan=True)
    other = np.setdiff1d(np.arange(dists.shape[1]), valid)
    assert float(c[0, other].abs().sum()) == 0.0
