"""Drop-in for the reference's `VoxelEncoding` extension module
(c_lib/VoxelEncoding/src/binding.cpp:17-74, imported by the reference as
`from c_lib.VoxelEncoding.dist import VoxelEncoding`).

Same function names, positional signatures, return tuples, dtypes and in-place
effects; every call runs the HIP kernels of libsailor_hip.so on the tensors'
GPU, on torch's current stream.  `device` arguments are GPU ordinals, as in
the reference.

Deliberate differences (all documented in the kernels):
  * tree node indices follow one canonical order instead of atomicAdd order;
  * occupied coarse cells without occupied neighbours are set to -1 in the
    coarse volume (the reference leaves an index there that is later misread
    as a node id);
  * hit lists are bounded: a ray that needs more than n_max_hits entries
    raises an error instead of overwriting its neighbour's row;
  * compositing handles any feature width, not only 3 channels.
"""
from __future__ import annotations

import ctypes
from collections import namedtuple

import numpy as np
import torch

from . import _lib
from ._lib import as_input, call, check_input, ptr, stream_of

_f32, _i32 = torch.float32, torch.int32


def _dev(device):
    return torch.device("cuda", int(device))


# ---------------------------------------------------------------------------
# TSDF fusion
def integrate(occupied_vol, num_occ_voxel, cam_intr, cam_pose, cam_R, cam_T, color_ims,
              depth_ims, mask_ims, obs_weight, th_bbox, trunc_weight, tsdf_th_low,
              tsdf_th_high, device, fixed_bbox=None, compensated=False):
    """fusion_cuda_integrate (fusion_kernel.cu:600-727).

    Returns (tsdf, weight, color, vol_origin[B,3,2], vol_bbox[B,3,2],
    voxel_size[B,1]); fills `occupied_vol` with v_id/-1 and adds the number of
    occupied voxels to `num_occ_voxel`.  `fixed_bbox` ([B,3,2], optional,
    not in the reference) replaces the adaptive depth-derived box;
    `compensated` (not in the reference) forms the TSDF as a
    Kahan-compensated sum divided once by the weight.
    """
    dev = occupied_vol.device
    check_input(occupied_vol, "occupied_vol", _i32)
    check_input(num_occ_voxel, "num_occ_voxel", _i32, dev)
    B, X, Y, Z = occupied_vol.shape
    colors = as_input(color_ims, "color_ims", _f32, dev)
    depths = as_input(depth_ims, "depth_ims", _f32, dev)
    masks = as_input(mask_ims, "mask_ims", _i32, dev)
    if depths.dim() != 4 or depths.shape[0] != B:
        raise RuntimeError(f"depth_ims must be [B,V,H,W] with B={B}, got {tuple(depths.shape)}")
    _, V, H, W = depths.shape
    K = as_input(cam_intr, "cam_intr", _f32, dev).reshape(B, V, 3, 3)
    RT = as_input(cam_pose, "cam_pose", _f32, dev).reshape(B, V, 3, 4)
    R = as_input(cam_R, "cam_R", _f32, dev).reshape(B, V, 3, 3)
    T = as_input(cam_T, "cam_T", _f32, dev).reshape(B, V, 3)
    if num_occ_voxel.numel() != B:
        raise RuntimeError("num_occ_voxel must hold one counter per batch")
    fb = None
    if fixed_bbox is not None:
        fb = as_input(fixed_bbox, "fixed_bbox", _f32, dev).reshape(B, 3, 2)
    opts = dict(dtype=_f32, device=dev)
    tsdf = torch.empty((B, X, Y, Z), **opts)
    weight = torch.empty((B, X, Y, Z), **opts)
    color = torch.empty((B, X, Y, Z), **opts)
    vol_origin = torch.empty((B, 3, 2), **opts)
    vol_bbox = torch.empty((B, 3, 2), **opts)
    voxel_size = torch.empty((B, 1), **opts)
    with torch.cuda.device(dev):
        call("sailor_integrate_compensated" if compensated else "sailor_integrate",
             ptr(occupied_vol), ptr(num_occ_voxel), ptr(K), ptr(RT),
             ptr(R), ptr(T), ptr(colors), ptr(depths), ptr(masks), B, V, H, W, X, Y, Z,
             float(obs_weight), float(th_bbox), float(trunc_weight), float(tsdf_th_low),
             float(tsdf_th_high), ptr(fb), ptr(tsdf), ptr(weight), ptr(color),
             ptr(vol_origin), ptr(vol_bbox), ptr(voxel_size), stream_of(dev))
    return tsdf, weight, color, vol_origin, vol_bbox, voxel_size


def get_origin_bbox(depth_ims, cam_intr, cam_R, cam_T, th_bbox, device):
    """get_origin_bbox (fusion_kernel.cu:473-534) -> vol_origin [B,3,2]."""
    dev = depth_ims.device
    depths = as_input(depth_ims, "depth_ims", _f32)
    B, V, H, W = depths.shape
    K = as_input(cam_intr, "cam_intr", _f32, dev).reshape(B, V, 3, 3)
    R = as_input(cam_R, "cam_R", _f32, dev).reshape(B, V, 3, 3)
    T = as_input(cam_T, "cam_T", _f32, dev).reshape(B, V, 3)
    out = torch.empty((B, 3, 2), dtype=_f32, device=dev)
    with torch.cuda.device(dev):
        call("sailor_get_origin_bbox", ptr(depths), ptr(K), ptr(R), ptr(T), B, V, H, W,
             float(th_bbox), ptr(out), stream_of(dev))
    return out


def get_center_xyz(depth_ims, cam_intr, cam_R, cam_T, th_bbox, device):
    """get_center_xyz (fusion_kernel.cu:536-598) -> (vol_center [B,3] = sum
    of the back-projected valid points, vol_num_pts [B,1] = their count);
    th_bbox is unused, as in the reference."""
    dev = depth_ims.device
    depths = as_input(depth_ims, "depth_ims", _f32)
    B, V, H, W = depths.shape
    K = as_input(cam_intr, "cam_intr", _f32, dev).reshape(B, V, 3, 3)
    R = as_input(cam_R, "cam_R", _f32, dev).reshape(B, V, 3, 3)
    T = as_input(cam_T, "cam_T", _f32, dev).reshape(B, V, 3)
    center = torch.empty((B, 3), dtype=_f32, device=dev)
    num = torch.empty((B, 1), dtype=_f32, device=dev)
    ws = torch.empty(8192 * B, dtype=torch.uint8, device=dev)
    with torch.cuda.device(dev):
        call("sailor_get_center_xyz", ptr(depths), ptr(K), ptr(R), ptr(T), B, V, H, W,
             ptr(center), ptr(num), ptr(ws), ws.numel(), stream_of(dev))
    return center, num


# ---------------------------------------------------------------------------
# multi-resolution volumes and the tree
def build_multi_res_volumes(o_volumes, num_occupied_voxels, level, rate, device):
    """interface_build_m_res_volumes (octree.cpp:28-38): in place."""
    level = int(level)
    if len(o_volumes) < level or len(num_occupied_voxels) < level:
        raise RuntimeError("build_multi_res_volumes: need `level` volumes and counters")
    for i, v in enumerate(o_volumes[:level]):
        check_input(v, f"o_volumes[{i}]", _i32)
    for i, c in enumerate(num_occupied_voxels[:level]):
        check_input(c, f"num_occupied_voxels[{i}]", _i32)
    dev = o_volumes[0].device
    B, dim = o_volumes[0].shape[0], o_volumes[0].shape[1]
    d = dim
    for i in range(1, level):
        d //= int(rate)
        if tuple(o_volumes[i].shape) != (B, d, d, d):
            raise RuntimeError(f"o_volumes[{i}] must be [{B},{d},{d},{d}]")
    vols = (ctypes.c_void_p * level)(*[v.data_ptr() for v in o_volumes[:level]])
    cnts = (ctypes.c_void_p * level)(*[c.data_ptr() for c in num_occupied_voxels[:level]])
    with torch.cuda.device(dev):
        call("sailor_build_multi_res_volumes", vols, cnts, level, B, dim, int(rate),
             stream_of(dev))


OcTreeNode = namedtuple("OcTreeNode", "valid depth gindex vindex valid_num_leaves size vol_dim "
                                     "batch_id voxel_size xyz_center xyz_center_ father_gindex "
                                     "children_gindex")


class MultiLayerOctree:
    """MultiLayerOctree (octree.h:156-303; binding.cpp:56-74) for two levels.

    Holds the node arrays (structure of arrays, canonical order) and
    references to the caller's dense volumes, as the reference does.
    """

    def __init__(self, num_levels, batch_size, rate, device):
        self.num_levels = int(num_levels)
        self.batch_size = int(batch_size)
        self.rate = int(rate)
        self.device = int(device)
        self._vols = None
        self._counts = None
        self.node_pos = None
        self.node_info = None
        self.queue = None
        self.num_octree_nodes = 0

    # -- init_octree (octree_kernel.cu:569-644) --------------------------------
    def init_octree(self, dense_volumes, total_num_occupied_voxels, num_occ_voxels_batches,
                    octree_nodes_index, vol_size, vol_bbox):
        if len(dense_volumes) != 2:
            raise NotImplementedError(
                "MultiLayerOctree: only the two-layer tree (octree_level=2) is implemented")
        for i, v in enumerate(dense_volumes):
            check_input(v, f"dense_volumes[{i}]", _i32)
        self.num_levels = len(dense_volumes)
        self.batch_size = int(dense_volumes[0].shape[0])
        self._vols = list(dense_volumes)
        self._counts = [check_input(c, "num_occ_voxels", _i32) for c in num_occ_voxels_batches]
        self._totals = [int(t) for t in total_num_occupied_voxels]
        self._index = [int(i) for i in octree_nodes_index]
        self.num_octree_nodes = int(sum(self._totals))
        self.voxel_size = check_input(vol_size, "voxel_size", _f32).reshape(-1)
        self.volume_bbox = check_input(vol_bbox, "vol_bbox", _f32)
        dev = dense_volumes[0].device
        cap = max(self.num_octree_nodes, 1)
        self.node_pos = torch.zeros((cap, 4), dtype=_f32, device=dev)
        self.node_info = torch.full((cap, 4), -1, dtype=_i32, device=dev)
        self.queue = torch.zeros((2,), dtype=_i32, device=dev)
        self.dims = [int(v.shape[1]) for v in dense_volumes]
        self.rate = self.dims[0] // self.dims[1]

    def build_octree(self):
        """build_octree (octree_kernel.cu:656-697)."""
        dev = self._vols[0].device
        B = self.batch_size
        ws_bytes = _lib.lib().sailor_octree_workspace_bytes(B, self.dims[1])
        ws = torch.empty((ws_bytes + 3) // 4, dtype=_i32, device=dev)
        with torch.cuda.device(dev):
            call("sailor_octree_build", ptr(self._vols[0]), ptr(self._vols[1]),
                 ptr(self._counts[0]), ptr(self._counts[1]), B, self.dims[0], self.dims[1],
                 ptr(self.voxel_size), ptr(self.volume_bbox), self._index[1],
                 max(self.num_octree_nodes, 1), ptr(self.node_pos), ptr(self.node_info),
                 ptr(self.queue), ptr(ws), ws.numel() * 4, stream_of(dev))
        self._ws = ws

    def get_num_nodes(self):
        return self.num_octree_nodes

    def get_num_nodes_valid(self):
        """Allocated node count per level, coarsest first (queue_control)."""
        q = self.queue.tolist()
        if q[0] < 0:
            raise RuntimeError("octree build failed: node capacity exceeded")
        return q

    def get_nodes(self):
        """get_octree_nodes (octree.h:250-254, bound as get_nodes,
        binding.cpp:59): one OcTreeNode per node slot, in gindex order, with
        the fields of the reference's OcTree record (octree.h:29-55) that
        describe the node: valid, depth (log2 of the level's dim), gindex,
        vindex (raster index in its level's volume), valid_num_leaves,
        size (fine voxels per edge), vol_dim, batch_id, voxel_size (edge
        length), xyz_center (integer cell), xyz_center_ (centre),
        father_gindex and children_gindex (rate^3 slots: the occupied
        children in raster order, then -1).  Host copy; synchronises."""
        pos = self.node_pos.cpu().numpy()
        info = self.node_info.cpu().numpy()
        bbox = self.volume_bbox.reshape(-1, 3, 2).cpu().numpy()
        vs = self.voxel_size.reshape(-1).cpu().numpy()
        dims = self.dims
        rate = dims[0] // dims[1]
        out = []
        for g in range(self.num_octree_nodes):
            b, parent, child0, nchild = (int(x) for x in info[g])
            if b < 0:
                out.append(OcTreeNode(0, -1, g, -1, 0, 0, 0, -1, 0.0, (-1, -1, -1),
                                      (0.0, 0.0, 0.0), -1, [-1] * rate ** 3))
                continue
            size = float(pos[g, 3])
            coarse = parent < 0
            lvl_dim = dims[1] if coarse else dims[0]
            cell = size / vs[b]
            # centre = box min + (cell - 1) / 2 * vs + cell * vs * i (octree.h:110-151)
            idx = tuple(int(round((pos[g, a] - bbox[b, a, 0] - (cell - 1) / 2 * vs[b]) / size))
                        for a in range(3))
            vindex = (idx[0] * lvl_dim + idx[1]) * lvl_dim + idx[2]
            kids = list(range(child0, child0 + nchild)) if child0 >= 0 else []
            out.append(OcTreeNode(1, int(np.log2(lvl_dim)), g, vindex, nchild,
                                  int(round(cell)), lvl_dim, b, size, idx,
                                  tuple(float(x) for x in pos[g, :3]), parent,
                                  kids + [-1] * (rate ** 3 - len(kids))))
        return out

    def get_nodes_tensor(self):
        """output_all_nodes_tensor: [N,6] (x, y, z, size, valid, batch)."""
        dev = self.node_pos.device
        n = self.num_octree_nodes
        out = torch.empty((n, 6), dtype=_f32, device=dev)
        with torch.cuda.device(dev):
            call("sailor_octree_nodes_tensor", ptr(self.node_pos), ptr(self.node_info), n,
                 ptr(out), stream_of(dev))
        return out

    def _top(self, start_level):
        lvl = self.num_levels - int(start_level) - 1
        if lvl < 0 or lvl >= self.num_levels:
            raise RuntimeError(f"start_level {start_level} out of range")
        return lvl

    # -- voxel_traversal (intersection_kernel.cu:827-895) ---------------------
    def voxel_traversal(self, batch_rays_ori, batch_rays_dir, n_max_hits_coarse, start_level,
                        th_early_stop, is_training, opt_num_intersected_voxels):
        dev = self.node_pos.device
        ori = as_input(batch_rays_ori, "batch_rays_ori", _f32, dev)
        dirs = as_input(batch_rays_dir, "batch_rays_dir", _f32, dev)
        B, R = ori.shape[0], ori.shape[1]
        if B != self.batch_size:
            raise RuntimeError(f"rays batch {B} != tree batch {self.batch_size}")
        lvl = self._top(start_level)
        n_max = int(n_max_hits_coarse)
        opts = dict(device=dev)
        idx = torch.empty((B, R, n_max), dtype=_i32, **opts)
        tmin = torch.empty((B, R, n_max), dtype=_f32, **opts)
        tmax = torch.empty((B, R, n_max), dtype=_f32, **opts)
        vl = torch.empty((B, R, 1), dtype=_f32, **opts)
        status = torch.zeros((2,), dtype=_i32, **opts)
        with torch.cuda.device(dev):
            call("sailor_octree_voxel_traversal", ptr(self._vols[lvl]), ptr(self.node_pos),
                 ptr(self.node_info), self.node_pos.shape[0], ptr(self.volume_bbox), ptr(self.voxel_size), B,
                 self.dims[0], self.dims[lvl], self.rate, ptr(ori), ptr(dirs), R, n_max,
                 int(th_early_stop), int(bool(is_training)), int(opt_num_intersected_voxels),
                 ptr(idx), ptr(tmin), ptr(tmax), ptr(vl), ptr(status), stream_of(dev))
        self._last_status = status
        st = status.tolist()  # host read, as the reference's thrust reads
        self.num_max_intersected_voxels = st[0]
        if st[1]:
            raise RuntimeError(
                f"voxel_traversal: a ray hit {st[0]} voxels, more than n_max_hits={n_max}")
        return idx, tmin, tmax, vl

    # -- ray_voxels_points_sampling_coarse (sampling_kernel.cu:1251-1304) -----
    def ray_voxels_points_sampling_coarse(self, hits_idxs, min_depths, max_depths,
                                          uniform_rand_values, rays_valid_len, start_level,
                                          num_sampled):
        dev = self.node_pos.device
        hi = as_input(hits_idxs, "hits_idxs", _i32, dev)
        t0 = as_input(min_depths, "min_depths", _f32, dev)
        t1 = as_input(max_depths, "max_depths", _f32, dev)
        nz = as_input(uniform_rand_values, "uniform_rand_values", _f32, dev)
        vl = as_input(rays_valid_len, "rays_valid_len", _f32, dev)
        B, R, n_max = hi.shape
        S = int(num_sampled)
        if nz.numel() < B * R * S:
            raise RuntimeError("uniform_rand_values must be [B, N_rays, num_sampled]")
        lvl = self._top(start_level)
        start_scale = self.rate ** lvl
        sidx = torch.empty((B, R, S), dtype=_i32, device=dev)
        st = torch.empty((B, R, S), dtype=_f32, device=dev)
        with torch.cuda.device(dev):
            call("sailor_octree_sampling_coarse", ptr(hi), ptr(t0), ptr(t1), ptr(nz), ptr(vl),
                 ptr(self.node_pos), ptr(self.voxel_size), B, R, n_max, S, start_scale,
                 ptr(sidx), ptr(st), stream_of(dev))
        return sidx, st

    # -- octree_traversal (intersection_kernel.cu:738-825) --------------------
    def octree_traversal(self, batch_rays_ori, batch_rays_dir, n_max_hits_coarse, n_max_hits,
                         start_level, th_early_stop, opt_num_intersected_voxels):
        """Recursive traversal (training): returns (idx, min_depths, max_depths,
        ray_total_valid_len [B,R,1], idx_c, min_depths_c, max_depths_c)."""
        dev = self.node_pos.device
        ori = as_input(batch_rays_ori, "batch_rays_ori", _f32, dev)
        dirs = as_input(batch_rays_dir, "batch_rays_dir", _f32, dev)
        B, R = ori.shape[0], ori.shape[1]
        if B != self.batch_size:
            raise RuntimeError(f"rays batch {B} != tree batch {self.batch_size}")
        lvl = self._top(start_level)
        nc, nf = int(n_max_hits_coarse), int(n_max_hits)
        o = dict(device=dev)
        idx_c = torch.empty((B, R, nc), dtype=_i32, **o)
        t0_c = torch.empty((B, R, nc), dtype=_f32, **o)
        t1_c = torch.empty((B, R, nc), dtype=_f32, **o)
        idx = torch.empty((B, R, nf), dtype=_i32, **o)
        t0 = torch.empty((B, R, nf), dtype=_f32, **o)
        t1 = torch.empty((B, R, nf), dtype=_f32, **o)
        vl = torch.empty((B, R, 1), dtype=_f32, **o)
        status = torch.zeros((3,), dtype=_i32, **o)
        with torch.cuda.device(dev):
            call("sailor_octree_traversal", ptr(self._vols[lvl]), ptr(self.node_pos),
                 ptr(self.node_info), self.node_pos.shape[0], ptr(self.volume_bbox),
                 ptr(self.voxel_size), B, self.dims[0], self.dims[lvl], self.rate, ptr(ori),
                 ptr(dirs), R, nc, nf, int(th_early_stop), int(opt_num_intersected_voxels),
                 ptr(idx_c), ptr(t0_c), ptr(t1_c), ptr(idx), ptr(t0), ptr(t1), ptr(vl),
                 ptr(status), stream_of(dev))
        st = status.tolist()  # host read, as the reference's thrust reads
        self.num_max_intersected_voxels = st[0]
        self.num_max_intersected_voxels_fine = st[1]
        if st[2]:
            raise RuntimeError(f"octree_traversal: a ray needed more than n_max_hits_coarse="
                               f"{nc} / n_max_hits={nf} entries")
        return idx, t0, t1, vl, idx_c, t0_c, t1_c

    # -- ray_voxels_points_sampling (sampling_kernel.cu:1178-1248) ------------
    def ray_voxels_points_sampling(self, hits_idxs, min_depths, max_depths, rays_valid_len,
                                   uniform_rand_values, start_level, num_sampled):
        """Adaptive sampling (training): -> (sampled_idx [B,R,S], sampled_depths)."""
        dev = self.node_pos.device
        hi = as_input(hits_idxs, "hits_idxs", _i32, dev)
        t0 = as_input(min_depths, "min_depths", _f32, dev)
        t1 = as_input(max_depths, "max_depths", _f32, dev)
        vl = as_input(rays_valid_len, "rays_valid_len", _f32, dev)
        nz = as_input(uniform_rand_values, "uniform_rand_values", _f32, dev)
        B, R, n_max = hi.shape
        S = int(num_sampled)
        if nz.numel() < B * R * S:
            raise RuntimeError("uniform_rand_values must be [B, N_rays, num_sampled]")
        self._top(start_level)
        start_scale = 2 ** (self.num_levels - int(start_level) - 1)
        # slots the per-hit pass visits: the last traversal's largest count
        # (num_max_intersected_voxels_fine), else every slot
        max_hits = getattr(self, "num_max_intersected_voxels_fine", None)
        max_hits = n_max if max_hits is None else min(int(max_hits), n_max)
        sidx = torch.empty((B, R, S), dtype=_i32, device=dev)
        st = torch.empty((B, R, S), dtype=_f32, device=dev)
        with torch.cuda.device(dev):
            call("sailor_octree_sampling", ptr(hi), ptr(t0), ptr(t1), ptr(nz), ptr(vl),
                 ptr(self.node_pos), self.node_pos.shape[0], ptr(self.voxel_size), B, R, n_max,
                 max_hits, S, start_scale, ptr(sidx), ptr(st), stream_of(dev))
        return sidx, st

    # -- trilinear_aggregate_features[_backward] (sampling_kernel.cu:1384-1462)
    def _tri_args(self, ray_ori, ray_dir, sampled_z, sampled_idx):
        dev = self.node_pos.device
        ori = as_input(ray_ori, "ray_ori", _f32, dev)
        dirs = as_input(ray_dir, "ray_dir", _f32, dev)
        t = as_input(sampled_z, "sampled_z", _f32, dev)
        si = as_input(sampled_idx, "sampled_idx", _i32, dev)
        B, R, S = t.shape
        if si.shape != t.shape or ori.numel() != B * R * 3 or dirs.numel() != B * R * 3:
            raise RuntimeError("trilinear_aggregate_features: shapes of rays / samples differ")
        return dev, ori, dirs, t, si, B, R, S

    def trilinear_aggregate_features(self, ray_ori, ray_dir, sampled_z, sampled_idx,
                                     input_feats):
        """input_feats [B,C,R,S,8] -> [B,C,R,S]."""
        dev, ori, dirs, t, si, B, R, S = self._tri_args(ray_ori, ray_dir, sampled_z, sampled_idx)
        f = as_input(input_feats, "input_feats", _f32, dev)
        C = f.shape[1]
        if tuple(f.shape) != (B, C, R, S, 8):
            raise RuntimeError(f"input_feats must be [B, C, R, S, 8], got {tuple(f.shape)}")
        out = torch.empty((B, C, R, S), dtype=_f32, device=dev)
        with torch.cuda.device(dev):
            call("sailor_trilinear_aggregate", ptr(self.node_pos), self.node_pos.shape[0],
                 ptr(ori), ptr(dirs), ptr(si), ptr(t), ptr(f), B, R, S, C, ptr(out),
                 stream_of(dev))
        return out

    def trilinear_aggregate_features_backward(self, ray_ori, ray_dir, sampled_z, sampled_idx,
                                              input_feats, grad_output):
        """grad_output [B,C,R,S] -> gradient of input_feats [B,C,R,S,8]."""
        dev, ori, dirs, t, si, B, R, S = self._tri_args(ray_ori, ray_dir, sampled_z, sampled_idx)
        C = int(input_feats.shape[1])
        g = as_input(grad_output, "grad_output", _f32, dev)
        if tuple(g.shape) != (B, C, R, S):
            raise RuntimeError(f"grad_output must be [B, C, R, S], got {tuple(g.shape)}")
        out = torch.empty((B, C, R, S, 8), dtype=_f32, device=dev)
        with torch.cuda.device(dev):
            call("sailor_trilinear_aggregate_backward", ptr(self.node_pos),
                 self.node_pos.shape[0], ptr(ori), ptr(dirs), ptr(si), ptr(t), ptr(g), B, R, S,
                 C, ptr(out), stream_of(dev))
        return out

    # -- project_sampled_xyz (sampling_kernel.cu:1333-1381) --------------------
    def project_sampled_xyz(self, ray_ori, ray_dir, sampled_z, sampled_idx, Ks, RTs, ori_w,
                            ori_h, num_views, record_corners):
        """-> (proj_xy [B,V,R,S,(9,)2], proj_z [B,V,R,S,(9,)1]); -1 for dropped
        samples."""
        dev, ori, dirs, t, si, B, R, S = self._tri_args(ray_ori, ray_dir, sampled_z, sampled_idx)
        V = int(num_views)
        K = as_input(Ks, "Ks", _f32, dev)
        RT = as_input(RTs, "RTs", _f32, dev)
        if K.numel() != B * V * 9 or RT.numel() != B * V * 12:
            raise RuntimeError("Ks must be [B, num_views, 3, 3] and RTs [B, num_views, 3, 4]")
        shape = (B, V, R, S) + ((9,) if record_corners else ())
        xy = torch.empty(shape + (2,), dtype=_f32, device=dev)
        z = torch.empty(shape + (1,), dtype=_f32, device=dev)
        with torch.cuda.device(dev):
            call("sailor_project_sampled_xyz", ptr(self.node_pos), self.node_pos.shape[0],
                 ptr(ori), ptr(dirs), ptr(si), ptr(t), ptr(K), ptr(RT), B, V, R, S, int(ori_w),
                 int(ori_h), int(bool(record_corners)), ptr(xy), ptr(z), stream_of(dev))
        return xy, z

    def sort_samplings(self, sampled_depths):
        return sort_samplings(sampled_depths)


def sort_samplings(sampled_depths):
    """sort_samplings (sampling_kernel.cu:1308-1329): (sorted, int32 index);
    stable for equal values."""
    t = as_input(sampled_depths, "sampled_depths", _f32)
    S = t.shape[-1]
    rows = t.numel() // S if S else 0
    out = torch.empty_like(t)
    order = torch.empty(t.shape, dtype=_i32, device=t.device)
    with torch.cuda.device(t.device):
        call("sailor_sort_samplings", ptr(t), rows, S, ptr(out), ptr(order), stream_of(t.device))
    return out, order


# ---------------------------------------------------------------------------
# rays
def rays_calculating(cam_intr, cam_R, cam_T, load_size, device):
    """rays_calculating (sampling_kernel.cu:674-733) -> [B,V,H,W,6]."""
    return _rays(cam_intr, cam_R, cam_T, int(load_size), int(load_size), 1, 0)


def rays_calculating_parallel(cam_intr, cam_R, cam_T, load_size, device, num_gpus):
    """rays_calculating_parallel (sampling_kernel.cu:609-670): rows
    h*num_gpus + device of a load_size^2 image."""
    return _rays(cam_intr, cam_R, cam_T, int(load_size) // int(num_gpus), int(load_size),
                 int(num_gpus), int(device))


def _rays(K, R, T, H, W, stride, offset):
    dev = K.device
    K = as_input(K, "cam_intr", _f32)
    B, V = K.shape[0], K.shape[1]
    R = as_input(R, "cam_R", _f32, dev).reshape(B, V, 3, 3)
    T = as_input(T, "cam_T", _f32, dev).reshape(B, V, 3)
    out = torch.empty((B, V, H, W, 6), dtype=_f32, device=dev)
    with torch.cuda.device(dev):
        call("sailor_rays_calculating", ptr(K), ptr(R), ptr(T), B, V, H, W, stride, offset,
             ptr(out), stream_of(dev))
    return out


# ---------------------------------------------------------------------------
# compositing
def _render_shapes(sampled_depth, feats):
    B, R, S = sampled_depth.shape
    F = feats.shape[-1]
    if feats.numel() != B * R * S * F:
        raise RuntimeError("feats must be [B, N_rays, N_sampled, C]")
    return B, R, S, F


def volume_rendering_occ_forward(sampled_depth, occs, feats, device, t_thresh):
    """volume_rendering_occ_forward (volume_rendering.cu:749-787) ->
    (feats [B,R,C], depth [B,R,1], ws [B,R,1], alpha [B,R,1])."""
    dev = sampled_depth.device
    t = as_input(sampled_depth, "sampled_depth", _f32)
    o = as_input(occs, "occs", _f32, dev)
    f = as_input(feats, "feats", _f32, dev)
    B, R, S, F = _render_shapes(t, f)
    opts = dict(dtype=_f32, device=dev)
    of, od = torch.empty((B, R, F), **opts), torch.empty((B, R, 1), **opts)
    ow, oa = torch.empty((B, R, 1), **opts), torch.empty((B, R, 1), **opts)
    with torch.cuda.device(dev):
        call("sailor_volume_rendering_occ_forward", ptr(t), ptr(o), ptr(f), B, R, S, F,
             float(t_thresh), ptr(of), ptr(od), ptr(ow), ptr(oa), stream_of(dev))
    return of, od, ow, oa


def volume_rendering_training_forward(sampled_depth, sampled_sort_idx, sigmas, feats, device,
                                      t_thresh, support_occ):
    """volume_rendering_training_forward (volume_rendering.cu:525-583)."""
    dev = sampled_depth.device
    t = as_input(sampled_depth, "sampled_depth", _f32)
    si = as_input(sampled_sort_idx, "sampled_sort_idx", _i32, dev)
    s = as_input(sigmas, "sigmas", _f32, dev)
    f = as_input(feats, "feats", _f32, dev)
    B, R, S, F = _render_shapes(t, f)
    opts = dict(dtype=_f32, device=dev)
    of, od = torch.empty((B, R, F), **opts), torch.empty((B, R, 1), **opts)
    ow, oa = torch.empty((B, R, 1), **opts), torch.empty((B, R, 1), **opts)
    with torch.cuda.device(dev):
        call("sailor_volume_rendering_training_forward", ptr(t), ptr(si), ptr(s), ptr(f), B, R,
             S, F, float(t_thresh), int(bool(support_occ)), ptr(of), ptr(od), ptr(ow), ptr(oa),
             stream_of(dev))
    return of, od, ow, oa


def volume_rendering_training_backward(grad_output_feats, grad_output_depths, grad_weight_sums,
                                       grad_output_alphas, output_feats, output_depths,
                                       output_ws, output_alphas, sampled_depth,
                                       sampled_sort_idx, sigmas, feats, device, t_thresh,
                                       support_occ):
    """volume_rendering_training_backward (volume_rendering.cu:585-663) ->
    (grad_feats [B,R,S,C], grad_sigmas [B,R,S,1])."""
    dev = sampled_depth.device
    t = as_input(sampled_depth, "sampled_depth", _f32)
    B, R, S = t.shape
    args = [as_input(x, n, _f32, dev) for x, n in (
        (grad_output_feats, "grad_output_feats"), (grad_output_depths, "grad_output_depths"),
        (grad_weight_sums, "grad_weight_sums"), (grad_output_alphas, "grad_output_alphas"),
        (output_feats, "output_feats"), (output_depths, "output_depths"),
        (output_ws, "output_ws"), (output_alphas, "output_alphas"))]
    si = as_input(sampled_sort_idx, "sampled_sort_idx", _i32, dev)
    s = as_input(sigmas, "sigmas", _f32, dev)
    f = as_input(feats, "feats", _f32, dev)
    F = f.shape[-1]
    gf = torch.empty((B, R, S, F), dtype=_f32, device=dev)
    gs = torch.empty((B, R, S, 1), dtype=_f32, device=dev)
    with torch.cuda.device(dev):
        call("sailor_volume_rendering_training_backward", *[ptr(a) for a in args], ptr(t),
             ptr(si), ptr(s), ptr(f), B, R, S, F, float(t_thresh), int(bool(support_occ)),
             ptr(gf), ptr(gs), stream_of(dev))
    return gf, gs


def volume_render_torch_forward(t, sigma, feats, valid, occ_mode, compensated=False):
    """Compositing with the tensor formulation of _volume_render_torch
    (voxel_encoding_helper.py:132-186): t, sigma [B,R,S], feats [B,R,S,F],
    valid uint8 [B,R] or None -> (feats [B,R,F], depth, ws, alpha [B,R,1]).
    `compensated` selects Kahan-compensated sums (not in the reference)."""
    dev = t.device
    t = check_input(t, "sampled_depths", _f32)
    s = check_input(sigma, "sigma", _f32, dev)
    f = check_input(feats, "feats", _f32, dev)
    B, R, S = t.shape
    F = f.shape[-1]
    if s.numel() != B * R * S or f.numel() != B * R * S * F:
        raise RuntimeError("volume_render_torch: sigma must be [B, R, S(, 1)] and feats "
                           "[B, R, S, F] for sampled_depths [B, R, S]")
    if valid is not None:
        check_input(valid, "valid_rays", torch.uint8, dev)
        if valid.numel() != B * R:
            raise RuntimeError("volume_render_torch: valid_rays must be [B, R]")
    o = dict(dtype=_f32, device=dev)
    of, od = torch.empty((B, R, F), **o), torch.empty((B, R, 1), **o)
    ow, oa = torch.empty((B, R, 1), **o), torch.empty((B, R, 1), **o)
    with torch.cuda.device(dev):
        call("sailor_volume_render_torch_forward", ptr(t), ptr(s), ptr(f), ptr(valid), B * R, S,
             F, int(bool(occ_mode)), int(bool(compensated)), ptr(of), ptr(od), ptr(ow), ptr(oa),
             stream_of(dev))
    return of, od, ow, oa


def volume_render_torch_backward(g_feats, g_depth, g_ws, g_alpha, t, sigma, feats, valid,
                                 occ_mode):
    """Gradient of volume_render_torch_forward with respect to sigma and
    feats from the four output gradients (None = zero) ->
    (grad_sigma [B,R,S], grad_feats [B,R,S,F])."""
    dev = t.device
    B, R, S = t.shape
    F = feats.shape[-1]
    shapes = ((B, R, F), (B, R, 1), (B, R, 1), (B, R, 1))
    gs = []
    for g, shp, name in zip((g_feats, g_depth, g_ws, g_alpha), shapes,
                            ("feats", "depth", "ws", "alpha")):
        if g is None:
            g = torch.zeros(shp, dtype=_f32, device=dev)
        g = as_input(g, f"grad_{name}", _f32, dev)
        if g.numel() != int(torch.Size(shp).numel()):
            raise RuntimeError(f"volume_render_torch: grad_{name} must be {list(shp)}")
        gs.append(g)
    grad_sigma = torch.empty((B, R, S), dtype=_f32, device=dev)
    grad_feats = torch.empty((B, R, S, F), dtype=_f32, device=dev)
    with torch.cuda.device(dev):
        call("sailor_volume_render_torch_backward", *[ptr(g) for g in gs], ptr(t), ptr(sigma),
             ptr(feats), ptr(valid), B * R, S, F, int(bool(occ_mode)), ptr(grad_sigma),
             ptr(grad_feats), stream_of(dev))
    return grad_sigma, grad_feats


# ---------------------------------------------------------------------------
# frequency encoding
def freq_encode_forward(input_tensor, deg, device):
    """freq_encode_forward (freq_encoding.cu:89-118): [N,D] -> [N, D+2*D*deg]."""
    x = as_input(input_tensor, "input_tensor", _f32)
    N, D = x.shape
    y = torch.empty((N, D + 2 * D * int(deg)), dtype=_f32, device=x.device)
    with torch.cuda.device(x.device):
        call("sailor_freq_encode_forward", ptr(x), N, D, int(deg), ptr(y), stream_of(x.device))
    return y


def freq_encode_backward(grad_output, inputs, outputs, deg, device):
    """freq_encode_backward (freq_encoding.cu:120-151) -> [N, D]."""
    dev = grad_output.device
    g = as_input(grad_output, "grad_output", _f32)
    y = as_input(outputs, "outputs", _f32, dev)
    N, D = inputs.shape
    gx = torch.empty((N, D), dtype=_f32, device=dev)
    with torch.cuda.device(dev):
        call("sailor_freq_encode_backward", ptr(g), ptr(y), N, D, int(deg), ptr(gx),
             stream_of(dev))
    return gx


# ---------------------------------------------------------------------------
# preprocessing
def depth_normalize(depths, valid_min_depth, valid_max_depth, user_dist, multi_mask, divided2,
                    device):
    """depth_normalize (depth_normalizer.cu:98-189): depths [B, N_v, H, W] ->
    (normalized [B,N_v,H,W], mask [B,N_v,H,W], mid [B,N_v,1,1],
    distance [B,N_v,1,1])."""
    dev = _dev(device)
    d = as_input(depths, "depths", _f32, dev)
    B, V, H, W = d.shape
    out = torch.empty_like(d)
    mask = torch.empty_like(d)
    mid = torch.empty((B, V, 1, 1), dtype=_f32, device=dev)
    dist = torch.empty((B, V, 1, 1), dtype=_f32, device=dev)
    ws = torch.empty(8 * B * V, dtype=torch.uint8, device=dev)
    call("sailor_depth_normalize", ptr(d), B, V, H, W, float(valid_min_depth),
         float(valid_max_depth), float(user_dist), int(bool(multi_mask)), int(bool(divided2)),
         ptr(out), ptr(mask), ptr(mid), ptr(dist), ptr(ws), ws.numel(), stream_of(dev))
    return out, mask, mid, dist


def rays_selecting(ray_ori_dir, sampled_xy, device):
    """rays_selecting (sampling_kernel.cu:736-769): [B,N_v,H,W,6] at
    sampled_xy [B,N_v,N_rays,2] (x, y) -> [B,N_v,N_rays,6]."""
    dev = _dev(device)
    r = as_input(ray_ori_dir, "ray_ori_dir", _f32, dev)
    xy = as_input(sampled_xy, "sampled_xy", torch.int32, dev)
    B, V, H, W = r.shape[:4]
    N = xy.shape[2]
    out = torch.zeros((B, V, N, 6), dtype=_f32, device=dev)
    call("sailor_rays_selecting", ptr(r), B, V, H, W, ptr(xy), N, ptr(out), stream_of(dev))
    return out


def rgbdv_selecting(target_rgbs, target_normals, target_depths, sampled_xy, device):
    """rgbdv_selecting (sampling_kernel.cu:771-825): rgbs [B,N_v,H,W,C],
    normals [B,N_v,H,W,3], depths [B,N_v,H,W,1] at sampled_xy [B,N_v,N,2] ->
    (rgb [B,N_v,N,C], normals [B,N_v,N,3], depth [B,N_v,N,1], label [B,N_v,N,1])."""
    dev = _dev(device)
    rgb = as_input(target_rgbs, "target_rgbs", _f32, dev)
    nml = as_input(target_normals, "target_normals", _f32, dev)
    dep = as_input(target_depths, "target_depths", _f32, dev)
    xy = as_input(sampled_xy, "sampled_xy", torch.int32, dev)
    B, V, H, W, C = rgb.shape
    N = xy.shape[2]
    o = dict(dtype=_f32, device=dev)
    out_rgb = torch.zeros((B, V, N, C), **o)
    out_nml = torch.zeros((B, V, N, 3), **o)
    out_dep = torch.zeros((B, V, N, 1), **o)
    out_lab = torch.ones((B, V, N, 1), **o)
    call("sailor_rgbdv_selecting", ptr(rgb), C, ptr(nml), ptr(dep), B, V, H, W, ptr(xy), N,
         ptr(out_rgb), ptr(out_nml), ptr(out_dep), ptr(out_lab), stream_of(dev))
    return out_rgb, out_nml, out_dep, out_lab
