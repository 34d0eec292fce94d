"""GPU parity of the training-only octree kernels against the C oracle,
through the Python surface (MultiLayerOctree.octree_traversal,
ray_voxels_points_sampling, trilinear_aggregate_features[_backward],
voxel_encoding_helper.voxel_trilinear_interpolation) on trees built on the
GPU from a fused capsule scene.  Everything is compared bit for bit (same
single-rounding operation sequence on both sides, canonical entry order);
the trilinear backward is also checked as the adjoint of the forward at a
larger size."""
import numpy as np
import pytest
import torch

from oracle import oracle as orc
from tests.helpers import scene_inputs, target_rays_np

pytestmark = pytest.mark.gpu

CASES = {"d64_r2": (64, 2), "d32_r4": (32, 4)}


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _tree(dim, rate):
    from sailor_amd import VoxelEncoding as VE
    print(f"[training case] dim={dim} rate={rate}", flush=True)
    inp = scene_inputs(0, 4, 128, 128, 150.0)
    occ0 = torch.full((1, dim, dim, dim), -1, dtype=torch.int32, device="cuda")
    dc = dim // rate
    occ1 = torch.full((1, dc, dc, dc), -1, dtype=torch.int32, device="cuda")
    num0 = torch.zeros((1, 1), dtype=torch.int32, device="cuda")
    num1 = torch.zeros((1, 1), dtype=torch.int32, device="cuda")
    _, _, _, _, bbox, vs = VE.integrate(occ0, num0, _t(inp["K"]), _t(inp["RT"]), _t(inp["R"]),
                                        _t(inp["T"]), _t(inp["colors"]), _t(inp["depths"]),
                                        _t(inp["masks"]), 1.0, 0.025, 3.0, 13.5, 13.5, 0)
    VE.build_multi_res_volumes([occ0, occ1], [num0, num1], 2, rate, 0)
    totals = [int(num0.sum()), int(num1.sum())]
    tree = VE.MultiLayerOctree(2, 1, rate, 0)
    tree.init_octree([occ0, occ1], totals, [num0, num1], [0, totals[0]], vs, bbox)
    tree.build_octree()
    Kt, Rt, Tt = target_rays_np(64, 64, 75.0)
    rays = VE.rays_calculating(_t(Kt), _t(Rt), _t(Tt), 64, 0)[0, 0].reshape(1, -1, 6)
    ori, dirs = rays[..., :3].contiguous(), rays[..., 3:].contiguous()
    host = dict(vc=occ1.cpu().numpy(), pos=tree.node_pos.cpu().numpy(),
                info=tree.node_info.cpu().numpy(), bbox=bbox.cpu().numpy(),
                vs=vs.cpu().numpy().reshape(-1), ori=ori.cpu().numpy(), dirs=dirs.cpu().numpy())
    return tree, ori, dirs, host


_cache = {}


@pytest.fixture(scope="module", params=list(CASES))
def case(request):
    if request.param not in _cache:
        dim, rate = CASES[request.param]
        _cache[request.param] = (dim, rate) + _tree(dim, rate)
    return _cache[request.param]


@pytest.mark.parametrize("early,opt", [(-1, -1), (2, -1), (-1, 3)])
def test_octree_traversal_bitexact(case, early, opt):
    dim, rate, tree, ori, dirs, h = case
    got = tree.octree_traversal(ori, dirs, 96, 256, 0, early, opt)
    ref = orc.octree_traversal(h["vc"], h["pos"], h["info"], h["bbox"], h["vs"], dim, rate,
                               h["ori"], h["dirs"], 96, 256, early, opt)
    names = ("idx", "tmin", "tmax", "valid_len", "idx_c", "tmin_c", "tmax_c")
    for name, g in zip(names, got):
        np.testing.assert_array_equal(g.cpu().numpy().reshape(ref[name].shape), ref[name],
                                      err_msg=name)
    assert tree.num_max_intersected_voxels == ref["status"][0]
    assert tree.num_max_intersected_voxels_fine == ref["status"][1]
    assert (ref["idx"][..., 0] >= 0).sum() > 50


def test_adaptive_sampling_bitexact(case):
    dim, rate, tree, ori, dirs, h = case
    idx, t0, t1, vl, *_ = tree.octree_traversal(ori, dirs, 96, 256, 0, -1, -1)
    S = 48
    for seed in (None, 3):
        if seed is None:
            nz = torch.full((1, ori.shape[1], S), 0.5, device="cuda")
        else:
            nz = torch.rand((1, ori.shape[1], S), generator=torch.Generator().manual_seed(seed))
            nz = nz.cuda()
        sidx, st = tree.ray_voxels_points_sampling(idx, t0, t1, vl, nz, 0, S)
        ri, rt = orc.sampling_adaptive(idx.cpu().numpy(), t0.cpu().numpy(), t1.cpu().numpy(),
                                       nz.cpu().numpy(), vl.cpu().numpy().reshape(1, -1),
                                       h["pos"], h["vs"], S, 2,  # 2^(levels - 0 - 1)
                                       tree.num_max_intersected_voxels_fine)
        np.testing.assert_array_equal(sidx.cpu().numpy(), ri)
        np.testing.assert_array_equal(st.cpu().numpy(), rt)
        assert (ri >= 0).mean() > 0.05


def _samples(tree, ori, dirs, S=24):
    idx, t0, t1, vl, *_ = tree.octree_traversal(ori, dirs, 96, 256, 0, -1, -1)
    nz = torch.rand((1, ori.shape[1], S), generator=torch.Generator().manual_seed(9)).cuda()
    sidx, st = tree.ray_voxels_points_sampling(idx, t0, t1, vl, nz, 0, S)
    sidx[:, ::7, 3] = -1                                   # dropped samples
    return sidx, st


@pytest.mark.parametrize("C", [1, 8])
def test_trilinear_bitexact(case, C):
    from sailor_amd.voxel_encoding_helper import voxel_trilinear_interpolation
    dim, rate, tree, ori, dirs, h = case
    sidx, st = _samples(tree, ori, dirs)
    B, R, S = sidx.shape
    g = torch.Generator().manual_seed(C)
    feats = torch.randn((B, C, R, S, 8), generator=g).cuda().requires_grad_(True)
    grad = torch.randn((B, C, R, S), generator=g).cuda()
    out = voxel_trilinear_interpolation(tree, ori, dirs, st, sidx, feats)
    out.backward(grad)
    args = (h["pos"], h["ori"], h["dirs"], sidx.cpu().numpy(), st.cpu().numpy())
    ref = orc.trilinear_fwd(*args, feats.detach().cpu().numpy())
    ref_g = orc.trilinear_bwd(*args, grad.cpu().numpy())
    np.testing.assert_array_equal(out.detach().cpu().numpy(), ref)
    np.testing.assert_array_equal(feats.grad.cpu().numpy(), ref_g)
    assert (sidx >= 0).float().mean() > 0.05 and (sidx < 0).any()


def test_trilinear_adjoint_at_size(case):
    """<fwd(F), G> = <F, bwd(G)> on 4096 rays x 48 samples x 16 channels."""
    dim, rate, tree, ori, dirs, h = case
    sidx, st = _samples(tree, ori, dirs, S=48)
    C = 16
    B, R, S = sidx.shape
    F = torch.randn((B, C, R, S, 8), device="cuda")
    G = torch.randn((B, C, R, S), device="cuda")
    fwd = tree.trilinear_aggregate_features(ori, dirs, st, sidx, F)
    bwd = tree.trilinear_aggregate_features_backward(ori, dirs, st, sidx, F, G)
    lhs = (fwd.double() * G.double()).sum().item()
    rhs = (F.double() * bwd.double()).sum().item()
    assert abs(lhs - rhs) <= 1e-5 * (fwd.double().abs() * G.double().abs()).sum().item()
    ones = tree.trilinear_aggregate_features_backward(ori, dirs, st, sidx, F[:, :1],
                                                      torch.ones_like(G[:, :1]))
    ok = (sidx >= 0)[:, None]
    np.testing.assert_allclose(ones.sum(-1)[ok].cpu().numpy(), 1.0, atol=1e-6)


@pytest.mark.parametrize("corners", [False, True])
def test_project_sampled_xyz_bitexact(case, corners):
    from sailor_amd import scenes
    dim, rate, tree, ori, dirs, h = case
    sidx, st = _samples(tree, ori, dirs, S=12)
    cams = scenes.ring_cameras(3, 150.0, 128, 128)
    K = np.stack([c.K for c in cams])[None].astype(np.float32)
    RT = np.stack([c.RT for c in cams])[None].astype(np.float32)
    xy, z = tree.project_sampled_xyz(ori, dirs, st, sidx, _t(K), _t(RT), 128, 128, 3, corners)
    rxy, rz = orc.project(h["pos"], h["ori"], h["dirs"], sidx.cpu().numpy(), st.cpu().numpy(),
                          K, RT, 128, 128, corners)
    np.testing.assert_array_equal(xy.cpu().numpy(), rxy)
    np.testing.assert_array_equal(z.cpu().numpy(), rz)
