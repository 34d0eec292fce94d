"""GPU parity of the geometry stages against the CPU oracle, through the
C ABI (sailor_amd.VoxelEncoding -> libsailor_hip.so):

  integrate (TSDF fusion + adaptive bbox), build_multi_res_volumes, the
  two-layer tree, rays_calculating, voxel_traversal,
  ray_voxels_points_sampling_coarse, sort_samplings.

Integer / index outputs must match bit for bit.  The float outputs of these
stages are computed with the same single-rounding operation sequence on both
sides, so they are compared bit for bit as well.
"""
import numpy as np
import pytest
import torch

from oracle import oracle as orc
from tests.helpers import scene_inputs, target_rays_np

pytestmark = pytest.mark.gpu

CASES = {
    # name: (H=W, fx, dim, rate, target H=W, target fx, fixed cube?)
    "small_r4": (128, 150.0, 32, 4, 64, 75.0, False),
    "c1_r8": (512, 600.0, 128, 8, 256, 300.0, True),
    "c1_r4_adaptive": (512, 600.0, 128, 4, 256, 300.0, False),
    # rate 2: 16 coarse cells per wave in the row-per-lane merge / tree count
    "small_r2": (128, 150.0, 32, 2, 64, 75.0, False),
}


def _t(a, dtype=None):
    t = torch.from_numpy(np.ascontiguousarray(a)).cuda()
    return t if dtype is None else t.to(dtype)


def run_case(name):
    from sailor_amd import VoxelEncoding as VE
    from sailor_amd import scenes
    HW, fx, dim, rate, tHW, tfx, fixed = CASES[name]
    inp = scene_inputs(0, 4, HW, HW, fx)
    print(f"[case {name}] H=W={HW} dim={dim} rate={rate}", flush=True)
    fb = scenes.fixed_cube()[None] if fixed else None
    # ---- GPU
    occ0 = torch.full((1, dim, dim, dim), -1, dtype=torch.int32, device="cuda")
    dc = dim // rate
    occ1 = torch.full((1, dc, dc, dc), -1, dtype=torch.int32, device="cuda")
    num0 = torch.zeros((1, 1), dtype=torch.int32, device="cuda")
    num1 = torch.zeros((1, 1), dtype=torch.int32, device="cuda")
    tsdf, weight, color, origin, bbox, vs = VE.integrate(
        occ0, num0, _t(inp["K"]), _t(inp["RT"]), _t(inp["R"]), _t(inp["T"]),
        _t(inp["colors"]), _t(inp["depths"]), _t(inp["masks"]), 1.0, 0.025, 3.0, 13.5, 13.5,
        0, fixed_bbox=None if fb is None else _t(fb))
    g = dict(occ=occ0.cpu().numpy(), num_occ=num0.cpu().numpy().reshape(-1),
             tsdf=tsdf.cpu().numpy(), weight=weight.cpu().numpy(), color=color.cpu().numpy(),
             vol_origin=origin.cpu().numpy(), vol_bbox=bbox.cpu().numpy(),
             voxel_size=vs.cpu().numpy().reshape(-1))
    VE.build_multi_res_volumes([occ0, occ1], [num0, num1], 2, rate, 0)
    g["coarse"] = occ1.cpu().numpy()
    g["coarse_cnt"] = num1.cpu().numpy().reshape(-1)
    totals = [int(num0.sum()), int(num1.sum())]
    index = [0, totals[0]]
    tree = VE.MultiLayerOctree(2, 1, rate, 0)
    tree.init_octree([occ0, occ1], totals, [num0, num1], index, vs, bbox)
    tree.build_octree()
    g["queue"] = tree.get_num_nodes_valid()
    g["node_pos"] = tree.node_pos.cpu().numpy()
    g["node_info"] = tree.node_info.cpu().numpy()
    g["vf"] = occ0.cpu().numpy()
    g["vc"] = occ1.cpu().numpy()
    g["nodes6"] = tree.get_nodes_tensor().cpu().numpy()
    Kt, Rt, Tt = target_rays_np(tHW, tHW, tfx)
    rays = VE.rays_calculating(_t(Kt), _t(Rt), _t(Tt), tHW, 0)
    g["rays"] = rays.cpu().numpy()
    ori = rays[0, 0].reshape(1, -1, 6)[..., :3].contiguous()
    dirs = rays[0, 0].reshape(1, -1, 6)[..., 3:].contiguous()
    idx, tmin, tmax, vl = tree.voxel_traversal(ori, dirs, 1024, 0, -1, False, -1)
    g.update(idx=idx.cpu().numpy(), tmin=tmin.cpu().numpy(), tmax=tmax.cpu().numpy(),
             vl=vl.cpu().numpy().reshape(1, -1), gmax=tree.num_max_intersected_voxels)
    S = 48
    noise = torch.full((1, ori.shape[1], S), 0.5, device="cuda")
    noise_r = torch.rand((1, ori.shape[1], S), generator=torch.Generator().manual_seed(7)).cuda()
    sidx, st = tree.ray_voxels_points_sampling_coarse(idx, tmin, tmax, noise, vl, 0, S)
    sidx_r, st_r = tree.ray_voxels_points_sampling_coarse(idx, tmin, tmax, noise_r, vl, 0, S)
    srt, order = tree.sort_samplings(st)
    g.update(sidx=sidx.cpu().numpy(), st=st.cpu().numpy(), sidx_r=sidx_r.cpu().numpy(),
             st_r=st_r.cpu().numpy(), sorted=srt.cpu().numpy(), order=order.cpu().numpy(),
             noise_r=noise_r.cpu().numpy())
    # ---- oracle
    o = orc.integrate(inp["K"], inp["RT"], inp["R"], inp["T"], inp["colors"], inp["depths"],
                      (dim, dim, dim), fixed_bbox=fb)
    oc, ocnt = orc.merge(o["occ"], rate)
    o["coarse"], o["coarse_cnt"] = oc, ocnt
    tr = orc.tree_build(o["occ"], oc, o["voxel_size"], o["vol_bbox"], int(o["num_occ"].sum()),
                        int(o["num_occ"].sum() + ocnt.sum()))
    o["tree"] = tr
    o["rays"] = orc.rays(Kt, Rt, Tt, tHW, tHW)
    r6 = o["rays"][0, 0].reshape(1, -1, 6)
    trv = orc.traverse(tr["vol_coarse"], tr["node_pos"], tr["node_info"], o["vol_bbox"],
                       o["voxel_size"], dim, rate, r6[..., :3], r6[..., 3:], 1024)
    o["trv"] = trv
    o["samp"] = orc.sampling(trv["idx"], trv["tmin"], trv["tmax"],
                             np.full((1, r6.shape[1], S), 0.5, np.float32), trv["valid_len"],
                             tr["node_pos"], o["voxel_size"], S, rate, trv["gmax"])
    o["samp_r"] = orc.sampling(trv["idx"], trv["tmin"], trv["tmax"], g["noise_r"],
                               trv["valid_len"], tr["node_pos"], o["voxel_size"], S, rate,
                               trv["gmax"])
    o["sort"] = orc.sort_samplings(o["samp"][1])
    return g, o


_cache = {}


@pytest.fixture(scope="module", params=list(CASES))
def case(request):
    if request.param not in _cache:
        _cache[request.param] = run_case(request.param)
    return _cache[request.param]


def test_integrate_bbox(case):
    g, o = case
    np.testing.assert_array_equal(g["vol_origin"], o["vol_origin"])
    np.testing.assert_array_equal(g["vol_bbox"], o["vol_bbox"])
    np.testing.assert_array_equal(g["voxel_size"], o["voxel_size"])


def test_integrate_occupancy_bitexact(case):
    g, o = case
    np.testing.assert_array_equal(g["occ"], o["occ"])
    np.testing.assert_array_equal(g["num_occ"], o["num_occ"])
    assert o["num_occ"].sum() > 100  # the scene is actually observed


def test_integrate_volumes(case):
    g, o = case
    for k in ("tsdf", "weight", "color"):
        np.testing.assert_array_equal(g[k], o[k], err_msg=k)


def test_multi_res_merge(case):
    g, o = case
    np.testing.assert_array_equal(g["coarse"], o["coarse"])
    np.testing.assert_array_equal(g["coarse_cnt"], o["coarse_cnt"])


@pytest.mark.parametrize("rate,dim", [(3, 36), (8, 64), (4, 32)])
def test_merge_sparse_volumes(rate, dim):
    """build_multi_res_volumes on seeded sparse occupancy (a few per cent of
    voxels, one isolated cell, one full block) vs the oracle's merge; rate 3
    (9 rows per cell do not divide a wave) takes the one-lane-per-cell
    kernel, 4 and 8 the row-per-lane one."""
    from sailor_amd import VoxelEncoding as VE
    rng = np.random.default_rng(rate)
    B = 2
    hit = rng.random((B, dim, dim, dim)) < 0.03
    idx = np.arange(dim ** 3, dtype=np.int32).reshape(dim, dim, dim)
    fine = np.where(hit, idx[None], -1).astype(np.int32)
    fine[1, :rate, :rate, :rate] = idx[:rate, :rate, :rate]          # one full block
    fine[0, -1, -1, -1] = idx[-1, -1, -1]                             # a lone corner voxel
    dc = dim // rate
    occ0 = torch.from_numpy(fine).cuda()
    occ1 = torch.full((B, dc, dc, dc), -7, dtype=torch.int32, device="cuda")
    num0 = torch.from_numpy((fine > -1).reshape(B, -1).sum(1).astype(np.int32)).cuda()
    num1 = torch.zeros(B, dtype=torch.int32, device="cuda")
    VE.build_multi_res_volumes([occ0, occ1], [num0, num1], 2, rate, "cuda")
    want, want_cnt = orc.merge(fine, rate)
    np.testing.assert_array_equal(occ1.cpu().numpy(), want)
    np.testing.assert_array_equal(num1.cpu().numpy(), want_cnt)


def test_tree_nodes_bitexact(case):
    g, o = case
    tr = o["tree"]
    assert list(g["queue"]) == list(tr["queue"])
    np.testing.assert_array_equal(g["node_info"], tr["node_info"])
    np.testing.assert_array_equal(g["node_pos"], tr["node_pos"])
    np.testing.assert_array_equal(g["vf"], tr["vol_fine"])
    np.testing.assert_array_equal(g["vc"], tr["vol_coarse"])
    valid = tr["node_info"][:, 0] >= 0
    np.testing.assert_array_equal(g["nodes6"][:, 4], valid.astype(np.float32))
    np.testing.assert_array_equal(g["nodes6"][valid, :4], tr["node_pos"][valid])


def test_rays_bitexact(case):
    g, o = case
    np.testing.assert_array_equal(g["rays"], o["rays"])


def test_traversal_bitexact(case):
    g, o = case
    t = o["trv"]
    np.testing.assert_array_equal(g["idx"], t["idx"])
    np.testing.assert_array_equal(g["tmin"], t["tmin"])
    np.testing.assert_array_equal(g["tmax"], t["tmax"])
    np.testing.assert_array_equal(g["vl"], t["valid_len"])
    assert g["gmax"] == t["gmax"]
    assert (t["idx"][..., 0] >= 0).sum() > 50  # rays do hit the performer


def test_sampling_bitexact(case):
    g, o = case
    np.testing.assert_array_equal(g["sidx"], o["samp"][0])
    np.testing.assert_array_equal(g["st"], o["samp"][1])
    np.testing.assert_array_equal(g["sidx_r"], o["samp_r"][0])
    np.testing.assert_array_equal(g["st_r"], o["samp_r"][1])


def test_sort_bitexact(case):
    g, o = case
    np.testing.assert_array_equal(g["sorted"], o["sort"][0])
    np.testing.assert_array_equal(g["order"], o["sort"][1])


def test_sort_rows_edge_cases():
    """sort_samplings on hand-made rows of S = 100 (two 64-lane passes) vs
    numpy's stable argsort: already ascending with ties (the in-order fast
    path), all -1, descending, ascending except one pair across lanes 63/64,
    ascending except the last pair, random with many duplicates."""
    from sailor_amd import VoxelEncoding as VE
    S = 100
    rng = np.random.default_rng(5)
    asc = np.sort(rng.integers(0, 20, S)).astype(np.float32)        # ties
    cross = np.arange(S, dtype=np.float32)
    cross[[63, 64]] = cross[[64, 63]]
    last = np.arange(S, dtype=np.float32)
    last[[S - 2, S - 1]] = last[[S - 1, S - 2]]
    rows = np.stack([asc, np.full(S, -1.0, np.float32), asc[::-1].copy(), cross, last,
                     rng.integers(0, 5, S).astype(np.float32)])
    srt, order = VE.sort_samplings(torch.from_numpy(rows).cuda())
    want = np.argsort(rows, axis=1, kind="stable")
    np.testing.assert_array_equal(order.cpu().numpy(), want)
    np.testing.assert_array_equal(srt.cpu().numpy(), np.take_along_axis(rows, want, 1))


def test_run_twice_identical():
    """Determinism: node order and hit lists do not depend on scheduling."""
    a = run_case("small_r4")[0]
    b = run_case("small_r4")[0]
    for k in ("node_info", "node_pos", "idx", "tmin", "st", "order"):
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)


def test_get_center_xyz_matches_oracle():
    """get_center_xyz (sum of the back-projected valid points and their
    count per batch) vs the oracle's float64 sum of the same float32 points;
    both sums are double, so they agree to float32 rounding."""
    from sailor_amd import VoxelEncoding as VE
    inp = scene_inputs(0, 4, 512, 512, 600.0)
    dep = np.concatenate([inp["depths"], inp["depths"][:, ::-1]], 0)   # B = 2
    K = np.concatenate([inp["K"]] * 2)
    R = np.concatenate([inp["R"]] * 2)
    T = np.concatenate([inp["T"]] * 2)
    c, n = VE.get_center_xyz(_t(dep), _t(K), _t(R), _t(T), 0.025, 0)
    rc, rn = orc.center_xyz(dep, K, R, T)
    np.testing.assert_array_equal(n.cpu().numpy().reshape(-1), rn.astype(np.float32))
    np.testing.assert_allclose(c.cpu().numpy(), rc, rtol=2e-7, atol=1e-6)
    assert rn.min() > 1e5


def test_get_nodes_records():
    """MultiLayerOctree.get_nodes: every valid record's gindex sits at its
    vindex in its level's rewritten volume, its centre equals node_pos,
    children point back at their parent, counts match queue_control."""
    from sailor_amd import VoxelEncoding as VE
    inp = scene_inputs(0, 4, 128, 128, 150.0)
    dim, rate = 32, 4
    occ0 = torch.full((1, dim, dim, dim), -1, dtype=torch.int32, device="cuda")
    dc = dim // rate
    occ1 = torch.full((1, dc, dc, dc), -1, dtype=torch.int32, device="cuda")
    num0 = torch.zeros((1, 1), dtype=torch.int32, device="cuda")
    num1 = torch.zeros((1, 1), dtype=torch.int32, device="cuda")
    *_, bbox, vs = VE.integrate(occ0, num0, _t(inp["K"]), _t(inp["RT"]), _t(inp["R"]),
                                _t(inp["T"]), _t(inp["colors"]), _t(inp["depths"]),
                                _t(inp["masks"]), 1.0, 0.025, 3.0, 13.5, 13.5, 0)
    VE.build_multi_res_volumes([occ0, occ1], [num0, num1], 2, rate, 0)
    totals = [int(num0.sum()), int(num1.sum())]
    tree = VE.MultiLayerOctree(2, 1, rate, 0)
    tree.init_octree([occ0, occ1], totals, [num0, num1], [0, totals[0]], vs, bbox)
    tree.build_octree()
    nodes = tree.get_nodes()
    assert len(nodes) == sum(totals)
    vols = {dim: occ0.cpu().numpy().reshape(-1), dim // rate: occ1.cpu().numpy().reshape(-1)}
    pos = tree.node_pos.cpu().numpy()
    valid = [nd for nd in nodes if nd.valid]
    n_coarse = sum(1 for nd in valid if nd.father_gindex < 0)
    assert [n_coarse, len(valid) - n_coarse] == tree.get_num_nodes_valid()
    for nd in valid:
        assert vols[nd.vol_dim][nd.vindex] == nd.gindex
        np.testing.assert_array_equal(np.array(nd.xyz_center_, np.float32), pos[nd.gindex, :3])
        kids = [k for k in nd.children_gindex if k >= 0]
        assert len(kids) == nd.valid_num_leaves
        for k in kids:
            assert nodes[k].father_gindex == nd.gindex and nodes[k].size == 1
        if nd.father_gindex < 0:
            assert nd.size == rate and nd.depth == 3
