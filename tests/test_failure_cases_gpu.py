"""Degenerate inputs and error paths of the drop-in, on the GPU:
  * a frame whose depth maps are all holes (nothing observed): the adaptive
    box, volumes and counters equal the oracle's, the tree is empty and the
    render writes zeros;
  * NaN depths: ignored exactly as the oracle ignores them (bit-exact
    volumes);
  * a target camera looking away from the volume: no ray hits, every sample
    slot -1, colours and depths 0;
  * hit lists too short for the rays: RuntimeError, rows not overrun;
  * tree node capacity too small: reported as -1 counts, RuntimeError from
    get_num_nodes_valid;
  * in/out tensors of the wrong dtype or non-contiguous inputs of the fused
    render: RuntimeError before any launch.
Every kernel here runs on valid, in-bounds inputs; the error cases are
rejected by host-side checks or reported through status words."""
import numpy as np
import pytest
import torch

from oracle import oracle as orc
from tests.helpers import scene_inputs

pytestmark = pytest.mark.gpu


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _integrate(inp, depths, dim=64, fixed=None):
    from sailor_amd import VoxelEncoding as VE
    occ = torch.full((1, dim, dim, dim), -1, dtype=torch.int32, device="cuda")
    num = torch.zeros((1, 1), dtype=torch.int32, device="cuda")
    out = VE.integrate(occ, num, _t(inp["K"]), _t(inp["RT"]), _t(inp["R"]), _t(inp["T"]),
                       _t(inp["colors"]), _t(depths), _t(inp["masks"]), 1.0, 0.025, 3.0, 13.5,
                       13.5, 0, fixed_bbox=None if fixed is None else _t(fixed))
    return occ, num, out


def _oracle(inp, depths, dim=64, fixed=None):
    return orc.integrate(inp["K"], inp["RT"], inp["R"], inp["T"], inp["colors"], depths,
                         (dim,) * 3, fixed_bbox=fixed)


def test_all_hole_depth_frame():
    from sailor_amd.pipeline import FrameConfig, FramePipeline
    inp = scene_inputs(0, 4, 128, 128, 150.0)
    holes = np.zeros_like(inp["depths"])
    occ, num, (tsdf, w, c, origin, bbox, vs) = _integrate(inp, holes)
    o = _oracle(inp, holes)
    assert int(num.item()) == 0 and int(o["num_occ"].sum()) == 0
    np.testing.assert_array_equal(origin.cpu().numpy(), o["vol_origin"])
    np.testing.assert_array_equal(bbox.cpu().numpy(), o["vol_bbox"])
    np.testing.assert_array_equal(tsdf.cpu().numpy(), o["tsdf"])
    np.testing.assert_array_equal(w.cpu().numpy(), o["weight"])
    # the whole frame on an empty scene: empty tree, no samples, zero image.
    # (128^3 over the 2.2 m cube: the occupancy test tsdf < trunc * 13.5 of
    # the reference (fusion_kernel.cu:333) holds for the unobserved tsdf = 1
    # once trunc = 3 voxels exceeds 1 / 13.5 m, i.e. on grids coarser than
    # about 90^3 -- there an empty scene is all "occupied")
    cfg = FrameConfig("holes", 4, 128, 150.0, 128, 8, 32, 37.5, 16, "c1")
    p = FramePipeline.synthetic(cfg, prec="fp32")
    p.depths.zero_()
    p.rgb.fill_(7.0)
    p.run()
    torch.cuda.synchronize()
    info = p.check()
    assert info["occupied"] == 0 and info["n_coarse"] == 0 and info["valid_rays"] == 0
    assert bool((p.sorted == -1).all())
    assert bool((p.rgb == 0).all()) and bool((p.depth == 0).all())


def test_nan_depths_are_ignored_like_the_oracle():
    inp = scene_inputs(0, 4, 128, 128, 150.0)
    d = inp["depths"].copy()
    rng = np.random.default_rng(3)
    d[rng.random(d.shape) < 0.05] = np.nan
    occ, num, (tsdf, w, c, origin, bbox, vs) = _integrate(inp, d)
    o = _oracle(inp, d)
    assert int(o["num_occ"].sum()) > 100
    np.testing.assert_array_equal(bbox.cpu().numpy(), o["vol_bbox"])
    np.testing.assert_array_equal(occ.cpu().numpy(), o["occ"])
    for name, x in (("tsdf", tsdf), ("weight", w), ("color", c)):
        np.testing.assert_array_equal(x.cpu().numpy(), o[name], err_msg=name)


def test_camera_looking_away_renders_nothing():
    from sailor_amd import scenes
    from sailor_amd.pipeline import FrameConfig, FramePipeline
    cfg = FrameConfig("away", 4, 128, 150.0, 64, 8, 32, 37.5, 16, "c1")
    p = FramePipeline.synthetic(cfg, prec="fp32")
    cam = scenes.target_camera(32, 32, 37.5)
    # turn the camera around its own centre: it now looks away from the box
    flip = np.diag([-1.0, 1.0, -1.0]).astype(np.float32)
    R = (flip @ cam.R).astype(np.float32)
    centre = -cam.R.T @ cam.T
    T = (-R @ centre).astype(np.float32)
    p.Rt.copy_(_t(R[None]))
    p.Tt.copy_(_t(T[None]))
    p.zdir.copy_(_t(R[2][None]))
    p.run()
    torch.cuda.synchronize()
    assert p.check()["occupied"] > 100
    assert bool((p.sorted == -1).all()) and bool((p.sidx == -1).all())
    assert bool((p.rgb == 0).all()) and bool((p.depth == 0).all())


def test_hit_list_overflow_raises():
    from sailor_amd import VoxelEncoding as VE
    inp = scene_inputs(0, 4, 128, 128, 150.0)
    occ0, num0, (_, _, _, _, bbox, vs) = _integrate(inp, inp["depths"], 32)
    occ1 = torch.full((1, 8, 8, 8), -1, dtype=torch.int32, device="cuda")
    num1 = torch.zeros((1, 1), dtype=torch.int32, device="cuda")
    VE.build_multi_res_volumes([occ0, occ1], [num0, num1], 2, 4, 0)
    totals = [int(num0.sum()), int(num1.sum())]
    tree = VE.MultiLayerOctree(2, 1, 4, 0)
    tree.init_octree([occ0, occ1], totals, [num0, num1], [0, totals[0]], vs, bbox)
    tree.build_octree()
    from tests.helpers import target_rays_np
    K, R, T = target_rays_np(64, 64, 75.0)
    rays = VE.rays_calculating(_t(K), _t(R), _t(T), 64, 0)
    ori = rays[0, 0].reshape(1, -1, 6)[..., :3].contiguous()
    dirs = rays[0, 0].reshape(1, -1, 6)[..., 3:].contiguous()
    idx, *_ = tree.voxel_traversal(ori, dirs, 1024, 0, -1, False, -1)
    assert tree.num_max_intersected_voxels > 2
    with pytest.raises(RuntimeError, match="more than n_max_hits"):
        tree.voxel_traversal(ori, dirs, 2, 0, -1, False, -1)


def test_node_capacity_exceeded_is_reported():
    from sailor_amd import VoxelEncoding as VE
    inp = scene_inputs(0, 4, 128, 128, 150.0)
    occ0, num0, (_, _, _, _, bbox, vs) = _integrate(inp, inp["depths"], 32)
    occ1 = torch.full((1, 8, 8, 8), -1, dtype=torch.int32, device="cuda")
    num1 = torch.zeros((1, 1), dtype=torch.int32, device="cuda")
    VE.build_multi_res_volumes([occ0, occ1], [num0, num1], 2, 4, 0)
    totals = [int(num0.sum()), int(num1.sum())]
    tree = VE.MultiLayerOctree(2, 1, 4, 0)
    # claim fewer occupied fine voxels than there are: the coarse range then
    # starts inside the fine nodes and the build must refuse to write
    tree.init_octree([occ0, occ1], [totals[0] // 2, totals[1]], [num0, num1],
                     [0, totals[0] // 2], vs, bbox)
    tree.build_octree()
    with pytest.raises(RuntimeError, match="capacity"):
        tree.get_num_nodes_valid()


def test_bad_inputs_rejected_before_launch():
    from sailor_amd import VoxelEncoding as VE
    from sailor_amd.model import NerfWeights
    from sailor_amd.render import PackedModel, render_rays
    inp = scene_inputs(0, 4, 128, 128, 150.0)
    occ = torch.full((1, 32, 32, 32), -1, dtype=torch.int64, device="cuda")
    num = torch.zeros((1, 1), dtype=torch.int32, device="cuda")
    with pytest.raises(RuntimeError, match="dtype"):
        VE.integrate(occ, num, _t(inp["K"]), _t(inp["RT"]), _t(inp["R"]), _t(inp["T"]),
                     _t(inp["colors"]), _t(inp["depths"]), _t(inp["masks"]), 1.0, 0.025, 3.0,
                     13.5, 13.5, 0)
    m = PackedModel(NerfWeights.init(32, 128), "fp32", "cuda")
    B, R, S, V = 1, 16, 8, 4
    t = torch.rand((B, S, R), device="cuda").transpose(1, 2)         # non-contiguous
    rays = torch.rand((B, R, 6), device="cuda")
    Kv = torch.rand((B, V, 3, 3), device="cuda")
    RTv = torch.rand((B, V, 3, 4), device="cuda")
    fmap = torch.rand((B, V, 8, 8, 32), device="cuda")
    bbox = torch.rand((B, 3, 2), device="cuda")
    zdir = torch.rand((B, 3), device="cuda")
    with pytest.raises(RuntimeError, match="contiguous"):
        render_rays(t, rays, Kv, RTv, (64, 64), fmap, bbox, zdir, m)
    t = t.contiguous()
    with pytest.raises(RuntimeError, match="rays must have shape"):
        render_rays(t, rays[:, :8], Kv, RTv, (64, 64), fmap, bbox, zdir, m)
    with pytest.raises(RuntimeError, match="channels"):
        render_rays(t, rays, Kv, RTv, (64, 64), fmap[..., :16].contiguous(), bbox, zdir, m)
    out = dict(rgb=torch.empty((B, R, 3), device="cuda"), depth=torch.empty((B, R), device="cuda"),
               aux=None, occ=torch.empty((B, R, S + 8), device="cuda"))
    with pytest.raises(RuntimeError, match="occ"):
        render_rays(t, rays, Kv, RTv, (64, 64), fmap, bbox, zdir, m, want_occ=True, out=out)
