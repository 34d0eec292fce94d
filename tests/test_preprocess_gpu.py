"""GPU parity of the preprocessing drop-ins against numpy restatements of
the reference (bit-exact: integer indexing, exact min/max, and the
reference's float32 operation order for the normalisation):
  * depth_normalize (depth_normalizer.cu:98-189), every (user_dist,
    multi_mask, divided2) branch, including a view without valid depths;
  * rays_selecting / rgbdv_selecting (sampling_kernel.cu:489-607,736-825).
"""
import numpy as np
import pytest
import torch

from oracle.preprocess_ref import depth_normalize as _depth_normalize_ref

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("user_dist,multi_mask,divided2",
                         [(-1, False, True), (-1, False, False), (-1, True, True),
                          (2.5, True, False), (2.5, False, False)])
def test_depth_normalize_bit_exact(user_dist, multi_mask, divided2):
    from sailor_amd import VoxelEncoding as VE
    rng = np.random.default_rng(1)
    d = rng.uniform(0.0, 6.0, (2, 3, 67, 91)).astype(np.float32)
    d[1, 2] = 9.0                                  # no valid depth in this view
    out = VE.depth_normalize(torch.from_numpy(d).cuda(), 0.2, 5.0, user_dist, multi_mask,
                             divided2, 0)
    torch.cuda.synchronize()
    ref = _depth_normalize_ref(d, 0.2, 5.0, user_dist, multi_mask, divided2)
    for g, r in zip(out, ref):
        np.testing.assert_array_equal(g.cpu().numpy(), r)


def test_rays_and_rgbdv_selecting_bit_exact():
    from sailor_amd import VoxelEncoding as VE
    rng = np.random.default_rng(2)
    B, V, H, W, N = 2, 3, 40, 56, 300
    rays = rng.normal(size=(B, V, H, W, 6)).astype(np.float32)
    rgbs = rng.uniform(size=(B, V, H, W, 12)).astype(np.float32)
    nml = rng.normal(size=(B, V, H, W, 3)).astype(np.float32)
    dep = np.where(rng.random((B, V, H, W, 1)) < 0.3, 0.0, rng.uniform(0.5, 3, (B, V, H, W, 1)))
    dep = dep.astype(np.float32)
    xy = np.stack([rng.integers(0, W, (B, V, N)), rng.integers(0, H, (B, V, N))], -1)
    xy = xy.astype(np.int32)
    cu = lambda a: torch.from_numpy(a).cuda()
    sel = VE.rays_selecting(cu(rays), cu(xy), 0)
    r, n, d, lab = VE.rgbdv_selecting(cu(rgbs), cu(nml), cu(dep), cu(xy), 0)
    torch.cuda.synchronize()
    b, v = np.meshgrid(np.arange(B), np.arange(V), indexing="ij")
    bi, vi = b[..., None], v[..., None]
    X, Y = xy[..., 0], xy[..., 1]
    np.testing.assert_array_equal(sel.cpu().numpy(), rays[bi, vi, Y, X])
    np.testing.assert_array_equal(r.cpu().numpy(), rgbs[bi, vi, Y, X])
    np.testing.assert_array_equal(n.cpu().numpy(), nml[bi, vi, Y, X])
    np.testing.assert_array_equal(d.cpu().numpy(), dep[bi, vi, Y, X])
    np.testing.assert_array_equal(lab.cpu().numpy(), (dep[bi, vi, Y, X] > 0).astype(np.float32))


def test_rgb_normalizer_bit_exact():
    from oracle.preprocess_ref import rgb_normalizer as ref
    from sailor_amd.utils_render import rgb_normalizer
    rng = np.random.default_rng(4)
    rgb = rng.uniform(0, 1, (3, 3, 37, 53)).astype(np.float32)
    mask = (rng.random((3, 1, 37, 53)) < 0.7).astype(np.float32)
    got = rgb_normalizer(torch.from_numpy(rgb).cuda(), torch.from_numpy(mask).cuda())
    np.testing.assert_array_equal(got.cpu().numpy(), ref(rgb, mask))
    got = rgb_normalizer(torch.from_numpy(rgb).cuda(), torch.from_numpy(mask).cuda(), 0.3, 0.7)
    np.testing.assert_array_equal(got.cpu().numpy(), ref(rgb, mask, 0.3, 0.7))
