"""Drop-in for c_lib/VoxelEncoding/voxel_encoding_helper.py: the autograd
Functions and stateful helpers that SRONet / SRONetUp call.

  integrate_depths   _integrate_depths (:206-294)
  OcTreeOptions      (:535-662)
  volume_render_infer  _volume_rendering_inference (:62-80)
  volume_render        _volume_render_training (:83-127) -- forward and
                       backward run the HIP compositing kernels
  volume_render_torch  _volume_render_torch (:130-189) -- the compositing the
                       reference trains with (tensor-op semantics), forward
                       and backward on HIP kernels
  voxel_trilinear_interpolation  _voxel_tri_interpolation (:12-57) -- forward
                       and backward run the HIP trilinear kernels
"""
from __future__ import annotations

import torch
import torch.nn as nn
from torch.autograd import Function

from . import VoxelEncoding

MAX_VOL_DIST = 999999


def _gpu_id(t):
    return int(t.device.index or 0)


class _voxel_tri_interpolation(Function):
    """Trilinear aggregation of per-corner sample features inside the
    samples' octree nodes; only the features receive a gradient."""

    @staticmethod
    def forward(ctx, octree, ray_ori, ray_dir, sampled_depths, sampled_idx, sampled_feats):
        assert (not ray_ori.requires_grad) and (not ray_dir.requires_grad) \
            and (not sampled_depths.requires_grad) and (not sampled_idx.requires_grad), \
            "Not supported"
        ctx.octree = octree
        ctx.ray_ori, ctx.ray_dir = ray_ori, ray_dir
        ctx.sampled_depths, ctx.sampled_idx = sampled_depths, sampled_idx
        ctx.save_for_backward(sampled_feats)
        return octree.trilinear_aggregate_features(ray_ori, ray_dir, sampled_depths, sampled_idx,
                                                   sampled_feats)

    @staticmethod
    def backward(ctx, grad_output_feats):
        sampled_feats, = ctx.saved_tensors
        grad = None
        if ctx.needs_input_grad[5]:
            grad = ctx.octree.trilinear_aggregate_features_backward(
                ctx.ray_ori, ctx.ray_dir, ctx.sampled_depths, ctx.sampled_idx, sampled_feats,
                grad_output_feats.contiguous())
        return None, None, None, None, None, grad


voxel_trilinear_interpolation = _voxel_tri_interpolation.apply


class _volume_rendering_inference(Function):
    @staticmethod
    def forward(ctx, depths, sigmas, feats, t_thres=1e-4):
        return VoxelEncoding.volume_rendering_occ_forward(
            depths.float(), sigmas.float(), feats.float(), _gpu_id(depths), t_thres)

    @staticmethod
    def backward(ctx, *grads):
        return None, None, None, None


volume_render_infer = _volume_rendering_inference.apply


class _volume_render_training(Function):
    @staticmethod
    def forward(ctx, depths, depth_sort_idxs, sigmas, feats, t_thres=1e-4, using_occ=True):
        depths, sigmas, feats = depths.float(), sigmas.float(), feats.float()
        ctx.t_thres = t_thres
        ctx.using_occ = using_occ
        out = VoxelEncoding.volume_rendering_training_forward(
            depths, depth_sort_idxs, sigmas, feats, _gpu_id(depths), t_thres, using_occ)
        ctx.save_for_backward(depths, depth_sort_idxs, sigmas, feats, *out)
        return out

    @staticmethod
    def backward(ctx, g_feats, g_depths, g_ws, g_alphas):
        depths, sort_idx, sigmas, feats, o_feats, o_depths, o_ws, o_alphas = ctx.saved_tensors
        zeros = lambda x, g: torch.zeros_like(x) if g is None else g
        grad_feats, grad_sigmas = VoxelEncoding.volume_rendering_training_backward(
            zeros(o_feats, g_feats), zeros(o_depths, g_depths), zeros(o_ws, g_ws),
            zeros(o_alphas, g_alphas), o_feats, o_depths, o_ws, o_alphas, depths, sort_idx,
            sigmas, feats, _gpu_id(depths), ctx.t_thres, ctx.using_occ)
        grad_sigmas = grad_sigmas.view_as(sigmas) if ctx.needs_input_grad[2] else None
        grad_feats = grad_feats if ctx.needs_input_grad[3] else None
        return None, None, grad_sigmas, grad_feats, None, None


volume_render = _volume_render_training.apply


class _volume_render_scan(Function):
    """Compositing of the samples in the order given with the tensor
    formulation's semantics (last density interval 0, occupancy transmittance
    prod(1 - alpha + 1e-6), no early stop), forward and gradient on the HIP
    kernels sailor_volume_render_torch_forward / _backward.  Sample depths
    are constants (the reference's sampler output carries no gradient)."""

    @staticmethod
    def forward(ctx, depths, sigmas, feats, valid_rays, occ_mode):
        if depths.requires_grad:
            raise RuntimeError("volume_render_torch: sampled depths must not require grad")
        t = depths.float().contiguous()
        s = sigmas.float().reshape(t.shape).contiguous()
        f = feats.float().contiguous()
        vr = None
        if valid_rays is not None:
            vr = valid_rays.reshape(t.shape[:-1]).to(torch.uint8).contiguous()
        out = VoxelEncoding.volume_render_torch_forward(t, s, f, vr, occ_mode)
        ctx.occ_mode = occ_mode
        ctx.sigma_shape = sigmas.shape
        ctx.save_for_backward(t, s, f, vr)
        return out

    @staticmethod
    def backward(ctx, g_feats, g_depths, g_ws, g_alphas):
        t, s, f, vr = ctx.saved_tensors
        gs, gf = VoxelEncoding.volume_render_torch_backward(
            g_feats, g_depths, g_ws, g_alphas, t, s, f, vr, ctx.occ_mode)
        gs = gs.view(ctx.sigma_shape) if ctx.needs_input_grad[1] else None
        gf = gf if ctx.needs_input_grad[2] else None
        return None, gs, gf, None, None


class _volume_render_torch(nn.Module):
    """_volume_render_torch (voxel_encoding_helper.py:130-189): same methods
    and arguments, running the HIP compositing kernels with torch autograd
    through _volume_render_scan.  sampled_depths [B, R, S], sigma [B, R, S, 1],
    feats [B, R, S, F], valid_rays [B, R] or None -> (feats [B, R, F],
    depths, ws, alphas [B, R, 1]).  max_depths is unused, as in the
    reference."""

    def forward(self, sampled_depths, sigma, feats, max_depths, valid_rays=None):
        return _volume_render_scan.apply(sampled_depths, sigma, feats, valid_rays, False)

    def forward_occ(self, sampled_depths, sigma, feats, max_depths, valid_rays=None):
        return _volume_render_scan.apply(sampled_depths, sigma, feats, valid_rays, True)


volume_render_torch = _volume_render_torch()


class _integrate_depths(nn.Module):
    """TSDF integration + multi-resolution volumes (voxel_encoding_helper.py:206-294)."""

    def preprocess_data(self, batch_size, inputs_data):
        colors, depths, masks, Ks, RTs = inputs_data
        h, w = depths.shape[-2:]
        num_views = depths.shape[0] // batch_size
        colors = torch.floor(colors[:, 2] * 256.0 * 256.0 + colors[:, 1] * 256.0 + colors[:, 0])
        if len(depths.shape) == 4:
            depths = depths[:, 0]
        if len(masks.shape) == 4:
            masks = masks[:, 0]
        depths = depths.view(-1, num_views, h, w).contiguous()
        colors = colors.view(-1, num_views, h, w).contiguous()
        masks = masks.view(-1, num_views, h, w).contiguous().int()
        Ks = Ks.view(-1, num_views, 3, 3).contiguous()
        RTs = RTs.view(-1, num_views, 3, 4).contiguous()
        Rs = RTs[..., :3, :3].contiguous()
        Ts = RTs[..., :3, -1].contiguous()
        return colors, depths, masks, Ks, RTs, Rs, Ts

    def clear_vol_buf(self, volumes, _level_volumes):
        occ, num = volumes
        for i in range(_level_volumes):
            occ[i].fill_(-1)
            num[i].fill_(0)
        return occ, num

    @torch.no_grad()
    def get_target_center(self, inputs_data, batch_size=1, _th_boox=0.025):
        """Mean of the back-projected valid depth points per batch [B, 3]
        (voxel_encoding_helper.py:256-264)."""
        _, depths, _, Ks, _, Rs, Ts = self.preprocess_data(batch_size, inputs_data)
        center, num = VoxelEncoding.get_center_xyz(depths, Ks, Rs, Ts, _th_boox,
                                                   _gpu_id(depths))
        return center / num

    @torch.no_grad()
    def get_origin_bbox(self, inputs_data, batch_size=1, _th_bbox=0.025):
        _, depths, _, Ks, _, Rs, Ts = self.preprocess_data(batch_size, inputs_data)
        return VoxelEncoding.get_origin_bbox(depths, Ks, Rs, Ts, _th_bbox, _gpu_id(depths))

    @torch.no_grad()
    def forward(self, inputs_data, volumes, batch_size=1, _obs_weight=1.0, _th_bbox=0.025,
                _tsdf_th_low=8.0, _tsdf_th_high=8.0, _level_volumes=6, _volume_res_rate=2,
                build_octree=True, fixed_bbox=None):
        colors, depths, masks, Ks, RTs, Rs, Ts = self.preprocess_data(batch_size, inputs_data)
        occ, num = self.clear_vol_buf(volumes, _level_volumes)
        gid = _gpu_id(colors)
        tsdf_vol, _, color_vol, vol_origin, vol_bbox, voxel_size = VoxelEncoding.integrate(
            occ[0], num[0], Ks, RTs, Rs, Ts, colors, depths, masks, _obs_weight, _th_bbox,
            3.0, _tsdf_th_low, _tsdf_th_high, gid, fixed_bbox=fixed_bbox)
        if build_octree:
            VoxelEncoding.build_multi_res_volumes(occ, num, _level_volumes, _volume_res_rate, gid)
        return occ, num, tsdf_vol, color_vol, vol_origin, vol_bbox, voxel_size


integrate_depths = _integrate_depths()


class OcTreeOptions(nn.Module):
    """voxel_encoding_helper.py:535-662."""

    def __init__(self, levels=6, batch_size=1, rate=2, vol_dim=256, n_max_hits=512,
                 device="cuda:0"):
        super().__init__()
        self._levels, self._batch_size, self._rate = levels, batch_size, rate
        self._gpu_id = int(str(device).split(":")[-1])
        self._octree = VoxelEncoding.MultiLayerOctree(levels, batch_size, rate, self._gpu_id)
        self._vol_dim = (vol_dim, vol_dim, vol_dim)
        self._n_max_hits = n_max_hits
        self._n_max_hits_coarse = None

    @torch.no_grad()
    def _init_octree(self, _occupied_volumes, _num_occ_voxels, _voxel_size, _vol_bbox):
        self._vol_dim = _occupied_volumes[0].shape[-3:]
        assert len(_occupied_volumes) == self._levels and len(_num_occ_voxels) == self._levels
        total, index = [], []
        for i, nv in enumerate(_num_occ_voxels):
            total.append(int(torch.sum(nv)))
            index.append(sum(total[:i]))
        self._octree.init_octree(_occupied_volumes, total, _num_occ_voxels, index, _voxel_size,
                                 _vol_bbox)
        self._octree.build_octree()
        return total, index

    @torch.no_grad()
    def _ray_intersection(self, _ray_ori, _ray_dir, _instersect_start_level=1, th_early_stop=-1,
                          only_voxel_traversal=True, is_training=False,
                          opt_num_intersected_voxels=-1):
        # e.g. 256 // 2 ** (levels - start - 1) coarse cells per axis, 3 x that
        self._n_max_hits_coarse = self._vol_dim[0] * 3 // self._rate ** (
            self._levels - _instersect_start_level - 1)
        if not only_voxel_traversal:
            return self._octree.octree_traversal(_ray_ori, _ray_dir, self._n_max_hits_coarse,
                                                 self._n_max_hits, _instersect_start_level,
                                                 th_early_stop, opt_num_intersected_voxels)
        self._n_max_hits_coarse = 1024
        return self._octree.voxel_traversal(_ray_ori, _ray_dir, self._n_max_hits_coarse,
                                            _instersect_start_level, th_early_stop, is_training,
                                            opt_num_intersected_voxels)

    @torch.no_grad()
    def _sort_sampling(self, sampled_depths):
        return self._octree.sort_samplings(sampled_depths)

    @torch.no_grad()
    def _ray_voxels_sampling_points(self, _idx, _min_depths, _max_depths, _ray_total_valid_len,
                                    _uniform_noise, _instersect_start_level, _n_sampled_points):
        return self._octree.ray_voxels_points_sampling(_idx, _min_depths, _max_depths,
                                                       _ray_total_valid_len, _uniform_noise,
                                                       _instersect_start_level,
                                                       _n_sampled_points)

    @torch.no_grad()
    def _project_sampled_xyz(self, _ray_ori, _ray_dir, _sampled_depths, _sampled_idx, Ks, RTs,
                             _ori_w, _ori_h, _num_views, _record_corners=False):
        if len(Ks.shape) == 3:
            Ks = Ks.view(self._batch_size, -1, 3, 3).contiguous()
        if len(RTs.shape) == 3:
            RTs = RTs.view(self._batch_size, -1, 3, 4).contiguous()
        return self._octree.project_sampled_xyz(_ray_ori, _ray_dir, _sampled_depths,
                                                _sampled_idx, Ks, RTs, _ori_w, _ori_h,
                                                _num_views, _record_corners)

    @torch.no_grad()
    def _ray_voxels_points_sampling_coarse(self, _idx, _min_depths, _max_depths, _uniform_noise,
                                           _ray_total_valid_len, _instersect_start_level,
                                           _n_sampled_points):
        return self._octree.ray_voxels_points_sampling_coarse(
            _idx, _min_depths, _max_depths, _uniform_noise, _ray_total_valid_len,
            _instersect_start_level, _n_sampled_points)

    def init_octree(self, occupied_volumes, num_occ_voxels, voxel_size, vol_bbox):
        return self._init_octree(occupied_volumes, num_occ_voxels, voxel_size, vol_bbox)

    def ray_intersection(self, ray_ori, ray_dir, instersect_start_level=1, th_early_stop=-1,
                         only_voxel_traversal=True, is_training=False,
                         opt_num_intersected_voxels=-1):
        return self._ray_intersection(ray_ori, ray_dir, instersect_start_level, th_early_stop,
                                      only_voxel_traversal, is_training,
                                      opt_num_intersected_voxels)

    def ray_voxels_points_sampling_coarse(self, idx, min_depths, max_depths, uniform_noise,
                                          ray_total_valid_len, instersect_start_level=1,
                                          n_sampled_points=128):
        return self._ray_voxels_points_sampling_coarse(idx, min_depths, max_depths,
                                                       uniform_noise, ray_total_valid_len,
                                                       instersect_start_level, n_sampled_points)

    def sort_sampling(self, _sampled_depths):
        return self._sort_sampling(_sampled_depths)

    @torch.no_grad()
    def get_all_nodes(self):
        return self._octree.get_nodes_tensor()

    @torch.no_grad()
    def get_octree(self):
        return self._octree
