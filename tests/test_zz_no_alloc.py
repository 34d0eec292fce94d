"""sailor_octree_build_noalloc (workspace from the library's built-in
scratch) against sailor_octree_build (caller workspace), through the C ABI:
two builds in a row on one GPU, the second on a smaller grid, each compared
bit for bit (node records, rewritten volumes, node counts); a grid whose
workspace exceeds the scratch is rejected before any launch."""
import ctypes

import numpy as np
import pytest
import torch

from tests.helpers import scene_inputs

pytestmark = pytest.mark.gpu


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _volumes(HW, fx, dim, rate):
    """Occupied fine + merged coarse volumes and counters of a seeded scene."""
    from sailor_amd import VoxelEncoding as VE
    inp = scene_inputs(0, 4, HW, HW, fx)
    occ0 = torch.full((1, dim, dim, dim), -1, dtype=torch.int32, device="cuda")
    dc = dim // rate
    occ1 = torch.full((1, dc, dc, dc), -1, dtype=torch.int32, device="cuda")
    num0 = torch.zeros((1, 1), dtype=torch.int32, device="cuda")
    num1 = torch.zeros((1, 1), dtype=torch.int32, device="cuda")
    _, _, _, _, bbox, vs = VE.integrate(
        occ0, num0, _t(inp["K"]), _t(inp["RT"]), _t(inp["R"]), _t(inp["T"]),
        _t(inp["colors"]), _t(inp["depths"]), _t(inp["masks"]), 1.0, 0.025, 3.0, 13.5, 13.5, 0)
    VE.build_multi_res_volumes([occ0, occ1], [num0, num1], 2, rate, 0)
    return occ0, occ1, num0, num1, bbox, vs


def _build(entry, vols, dim, rate, capacity):
    """Runs one build entry point on copies of the volumes; returns host arrays."""
    from sailor_amd import _lib
    from sailor_amd._lib import ptr, stream_of
    occ0, occ1, num0, num1, bbox, vs = vols
    vf, vc = occ0.clone(), occ1.clone()
    pos = torch.full((capacity, 4), 7.0, dtype=torch.float32, device="cuda")
    info = torch.full((capacity, 4), 7, dtype=torch.int32, device="cuda")
    queue = torch.full((2,), 7, dtype=torch.int32, device="cuda")
    dc = dim // rate
    args = [ptr(vf), ptr(vc), ptr(num0), ptr(num1), 1, dim, dc, ptr(vs), ptr(bbox), -1,
            capacity, ptr(pos), ptr(info), ptr(queue)]
    if entry == "sailor_octree_build":
        nb = _lib.lib().sailor_octree_workspace_bytes(1, dc)
        ws = torch.empty(nb, dtype=torch.uint8, device="cuda")
        args += [ptr(ws), ctypes.c_size_t(nb)]
    _lib.call(entry, *args, stream_of())
    torch.cuda.synchronize()
    return dict(pos=pos.cpu().numpy(), info=info.cpu().numpy(), queue=queue.cpu().numpy(),
                vf=vf.cpu().numpy(), vc=vc.cpu().numpy())


@pytest.mark.parametrize("first,second", [((512, 600.0, 128, 4), (128, 150.0, 32, 4))])
def test_noalloc_build_matches_workspace_build(first, second):
    for HW, fx, dim, rate in (first, second):
        print(f"[noalloc] H=W={HW} dim={dim} rate={rate}", flush=True)
        vols = _volumes(HW, fx, dim, rate)
        cap = int(vols[2].sum() + vols[3].sum())
        ref = _build("sailor_octree_build", vols, dim, rate, cap)
        got = _build("sailor_octree_build_noalloc", vols, dim, rate, cap)
        assert ref["queue"][0] > 0 and ref["queue"][1] > 0, ref["queue"]
        for k in ref:
            np.testing.assert_array_equal(got[k], ref[k], err_msg=k)


def test_noalloc_build_rejects_oversized_grid():
    """1 x 96^3 coarse cells need ~14 MiB of workspace: rejected by the host
    check (no launch, so the null pointers are never read)."""
    from sailor_amd import _lib
    from sailor_amd._lib import stream_of
    with pytest.raises(_lib.SailorError, match="built-in scratch"):
        _lib.call("sailor_octree_build_noalloc", None, None, None, None, 1, 384, 96, None, None,
                  -1, 1, None, None, None, stream_of())
