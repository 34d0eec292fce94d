/*
 * sailor_hip.h - C ABI of the MI355X (gfx950) implementation of SAILOR's
 * VoxelEncoding + FastNerf rendering path.
 *
 * Conventions
 *  - Every pointer argument is a device pointer to a dense, row-major buffer
 *    unless documented otherwise.  Shapes are given as [dim0, dim1, ...].
 *  - `stream` is a hipStream_t passed as void* (NULL = default stream).  All
 *    work is enqueued on it; no entry point synchronises the device unless
 *    stated.
 *  - Every entry point returns SAILOR_OK (0) or an error code; the message of
 *    the last error on the calling thread is returned by sailor_last_error().
 *    Nothing here terminates the process.
 *  - Outputs are written into caller-provided buffers (the Python layer
 *    allocates them with torch), so capacity is explicit and every call can be
 *    captured into a hipGraph.
 *
 * Reference citations are paths relative to the zdlarr/SAILOR repository.
 */
#ifndef SAILOR_HIP_H
#define SAILOR_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SAILOR_OK 0
#define SAILOR_ERR_ARG 1      /* invalid shape / argument */
#define SAILOR_ERR_CAPACITY 2 /* an output buffer is too small */
#define SAILOR_ERR_HIP 3      /* HIP runtime / launch failure */

const char* sailor_last_error(void);
int sailor_abi_version(void);

/* ------------------------------------------------------------------------ */
/* TSDF fusion                                                               */
/* ------------------------------------------------------------------------ */

/* Replaces fusion_cuda_integrate (c_lib/VoxelEncoding/src/fusion_kernel.cu:600-727),
 * bound as VoxelEncoding.integrate (src/binding.cpp:18).
 *  occ_vol    int32 [B,X,Y,Z]  out: v_id of occupied voxels, -1 otherwise
 *  num_occ    int32 [B]        in/out: += number of occupied voxels
 *  K          f32 [B,V,3,3], RT f32 [B,V,3,4], R f32 [B,V,3,3], T f32 [B,V,3]
 *  colors     f32 [B,V,H,W] packed b*65536+g*256+r; depths f32 [B,V,H,W] (m)
 *  masks      int32 [B,V,H,W] (unused by the algorithm, kept for the ABI)
 *  fixed_bbox f32 [B,3,2] or NULL.  NULL: adaptive cubic bbox from the depth
 *             maps (reference behaviour); otherwise the given box is used and
 *             voxel_size = (max-min along x)/(X-1).
 *  tsdf, weight, color  f32 [B,X,Y,Z] (fully overwritten)
 *  vol_origin f32 [B,3,2], vol_bbox f32 [B,3,2], voxel_size f32 [B]        */
int sailor_integrate(int32_t* occ_vol, int32_t* num_occ, const float* K,
                     const float* RT, const float* R, const float* T,
                     const float* colors, const float* depths,
                     const int32_t* masks, int B, int V, int H, int W, int X,
                     int Y, int Z, float obs_weight, float th_bbox,
                     float trunc_weight, float tsdf_th_low, float tsdf_th_high,
                     const float* fixed_bbox, float* tsdf, float* weight,
                     float* color, float* vol_origin, float* vol_bbox,
                     float* voxel_size, void* stream);

/* sailor_integrate with the TSDF formed as a Kahan-compensated sum of
 * obs_weight * dist over the observing views divided once by the weight sum
 * (precision form; no counterpart in the reference, whose running average
 * (tsdf w_old + obs_weight dist) / w_new has the same value in exact
 * arithmetic).  Same arguments; the occupancy test is applied to the
 * compensated value, so a voxel on the threshold may differ. */
int sailor_integrate_compensated(int32_t* occ_vol, int32_t* num_occ, const float* K,
                                 const float* RT, const float* R, const float* T,
                                 const float* colors, const float* depths,
                                 const int32_t* masks, int B, int V, int H, int W, int X,
                                 int Y, int Z, float obs_weight, float th_bbox,
                                 float trunc_weight, float tsdf_th_low, float tsdf_th_high,
                                 const float* fixed_bbox, float* tsdf, float* weight,
                                 float* color, float* vol_origin, float* vol_bbox,
                                 float* voxel_size, void* stream);

/* Replaces get_origin_bbox (fusion_kernel.cu:473-534).
 *  depths f32 [B,V,H,W], K/R f32 [B,V,3,3], T f32 [B,V,3] -> vol_origin [B,3,2] */
int sailor_get_origin_bbox(const float* depths, const float* K, const float* R,
                           const float* T, int B, int V, int H, int W,
                           float th_bbox, float* vol_origin, void* stream);

/* Replaces get_center_xyz (fusion_kernel.cu:536-598; kernel
 * obtain_center_points :186-233), reached from the GUI through
 * _integrate_depths.get_target_center (voxel_encoding_helper.py:256-264):
 * per batch, the sum of the world points back-projected from the pixels with
 * 0.1 m <= depth <= 4 m and their count.
 *  depths f32 [B,V,H,W], K/R f32 [B,V,3,3], T f32 [B,V,3] ->
 *  vol_center f32 [B,3] (sum of x, y, z), vol_num_pts f32 [B] (count).
 *  The sums are formed in double in a fixed order (the reference adds float
 *  atomics in scheduling order).  workspace: at least 8192 * B bytes. */
int sailor_get_center_xyz(const float* depths, const float* K, const float* R, const float* T,
                          int B, int V, int H, int W, float* vol_center, float* vol_num_pts,
                          void* workspace, size_t workspace_bytes, void* stream);

/* ------------------------------------------------------------------------ */
/* Two-layer occupancy tree                                                  */
/* ------------------------------------------------------------------------ */

/* Replaces build_multi_res_occupied_volumes (octree_kernel.cu:367-411), bound as
 * VoxelEncoding.build_multi_res_volumes (binding.cpp:24).
 *  vols   host array of `levels` device pointers; vols[l] is int32 [B,d_l,d_l,d_l]
 *         with d_l = dim / rate^l.  vols[l+1] is rewritten from vols[l].
 *  counts host array of `levels` device pointers to int32 [B]; counts[l+1] += */
int sailor_build_multi_res_volumes(int32_t* const* vols, int32_t* const* counts,
                                   int levels, int B, int dim, int rate,
                                   void* stream);

/* Node record layout (structure of arrays, capacity = sum of per-level totals):
 *  node_pos  f32 [cap,4]  (centre x, y, z, voxel size)
 *  node_info int32 [cap,4] (batch id or -1 when invalid, parent gindex or -1,
 *                           first child gindex or -1, number of children)
 * Node order is canonical: fine nodes occupy [fine_start, fine_start+n_fine)
 * ordered by (batch, parent coarse cell in x,y,z raster order, child in raster
 * order inside the block); coarse nodes occupy [coarse_start, coarse_start+n_c)
 * in (batch, raster) order.  Children of a coarse node are contiguous. */

/* Bytes of device workspace sailor_octree_build needs. */
size_t sailor_octree_workspace_bytes(int B, int dim_coarse);

/* Replaces MultiLayerOctree::init_octree + build_octree for two levels
 * (octree_kernel.cu:569-697, kernel gen_root_octree_nodes :126-239).
 *  vol_fine int32 [B,df,df,df] in/out, vol_coarse int32 [B,dc,dc,dc] in/out:
 *    occupied cells are rewritten with node gindex.  Occupied coarse cells
 *    with no occupied 26-neighbour are not allocated and are set to -1.
 *  fine_counts / coarse_counts int32 [B]: occupied cells per batch (the
 *    num_occ counters of integrate / build_multi_res_volumes).
 *  coarse_start < 0 means "sum of fine_counts[0..B)" read on the device.
 *  Entries [0, coarse_start + sum(coarse_counts)) that hold no node are
 *  marked invalid.
 *  queue_counts int32 [2] out: number of allocated coarse / fine nodes; both
 *  are set to -1 (and nothing is written) if the nodes do not fit into
 *  `capacity` or the fine nodes would overlap the coarse range.            */
int sailor_octree_build(int32_t* vol_fine, int32_t* vol_coarse,
                        const int32_t* fine_counts,
                        const int32_t* coarse_counts, int B, int dim_fine,
                        int dim_coarse, const float* voxel_size,
                        const float* vol_bbox, int coarse_start, int capacity,
                        float* node_pos, int32_t* node_info,
                        int32_t* queue_counts, void* workspace,
                        size_t workspace_bytes, void* stream);

/* No counterpart in the reference.  sailor_octree_build for callers that
 * cannot allocate device memory (code under hipGraph capture, serving loops,
 * harnesses that forbid hipMalloc): the same launches and results, with the
 * workspace taken from an 8 MiB array built into the library (one per GPU,
 * held until the process ends) instead of a caller buffer.  Sizes whose
 * workspace (sailor_octree_workspace_bytes(B, dim_coarse)) exceeds 8 MiB --
 * about 524,000 coarse cells over the batch -- are rejected with
 * SAILOR_ERR_ARG.  Nothing here calls hipMalloc or hipFree.  All calls on one
 * GPU share the array: calls that overlap on one GPU (different streams or
 * threads) race.  The device current at the call must be the stream's. */
int sailor_octree_build_noalloc(int32_t* vol_fine, int32_t* vol_coarse,
                                const int32_t* fine_counts,
                                const int32_t* coarse_counts, int B, int dim_fine,
                                int dim_coarse, const float* voxel_size,
                                const float* vol_bbox, int coarse_start, int capacity,
                                float* node_pos, int32_t* node_info,
                                int32_t* queue_counts, void* stream);

/* Replaces MultiLayerOctree::output_all_nodes_tensor (octree_kernel.cu:699-717):
 *  out f32 [n,6] = (x, y, z, voxel size, valid, batch).                     */
int sailor_octree_nodes_tensor(const float* node_pos, const int32_t* node_info,
                               int n, float* out, void* stream);

/* Replaces MultiLayerOctree::voxel_traversal (intersection_kernel.cu:827-895;
 * kernels ray_voxel_traversal :286-439, RayVoxelIntersectionNodeListNew
 * :128-284, RayVoxelIntersection :36-71).
 *  n_nodes: length of the node arrays; volume entries >= n_nodes are
 *  treated as empty cells.
 *  ray_ori/ray_dir f32 [B,R,3]; outputs hit_idx int32 [B,R,n_max_hits],
 *  tmin/tmax f32 [B,R,n_max_hits] (-1 padded), valid_len f32 [B,R].
 *  status int32 [2] device: [0] = max hits over all rays, [1] = 1 if any ray
 *  needed more than n_max_hits entries (those rays are truncated).          */
int sailor_octree_voxel_traversal(
    const int32_t* vol_coarse, const float* node_pos, const int32_t* node_info,
    int n_nodes, const float* vol_bbox, const float* voxel_size, int B, int dim_fine,
    int dim_coarse, int rate, const float* ray_ori, const float* ray_dir, int R,
    int n_max_hits, int th_early_stop, int is_training,
    int opt_num_intersected, int32_t* hit_idx, float* tmin, float* tmax,
    float* valid_len, int32_t* status, void* stream);

/* Replaces MultiLayerOctree::ray_voxels_points_sampling_coarse
 * (sampling_kernel.cu:1251-1304; kernel :1070-1177).
 *  noise f32 [B,R,S]; outputs sampled_idx int32 [B,R,S], sampled_t f32 [B,R,S]
 *  (-1 for unused slots).  start_scale = rate^(levels-start_level-1).      */
int sailor_octree_sampling_coarse(const int32_t* hit_idx, const float* tmin,
                                  const float* tmax, const float* noise,
                                  const float* valid_len, const float* node_pos,
                                  const float* voxel_size, int B, int R,
                                  int n_max_hits, int S, int start_scale,
                                  int32_t* sampled_idx, float* sampled_t,
                                  void* stream);

/* Replaces MultiLayerOctree::octree_traversal (intersection_kernel.cu:738-825;
 * kernels ray_voxel_traversal with RayVoxelIntersectionNodeList :72-126,
 * ray_octree_traversal :441-643), the recursive form used in training.
 *  Coarse lists (hit_idx_coarse, tmin_coarse, tmax_coarse [B,R,n_max_hits_coarse])
 *  hold the occupied coarse cells the DDA crosses; fine lists (hit_idx, tmin,
 *  tmax [B,R,n_max_hits]) hold, per coarse hit in order, the coarse entry and
 *  the children the ray crosses (last child first; a child whose interval
 *  equals an earlier entry's replaces it when smaller); valid_len [B,R] =
 *  sum of (tmax - tmin) * coarse size / entry size.  All lists -1 padded.
 *  status int32 [3], zeroed by the caller: max coarse count, max fine count,
 *  1 if a list needed more than its capacity (those rays are truncated).
 *  The reference's entry order depends on atomics; this order is one of its
 *  possible orders and is deterministic. */
int sailor_octree_traversal(const int32_t* vol_coarse, const float* node_pos,
                            const int32_t* node_info, int n_nodes, const float* vol_bbox,
                            const float* voxel_size, int B, int dim_fine, int dim_coarse,
                            int rate, const float* ray_ori, const float* ray_dir, int R,
                            int n_max_hits_coarse, int n_max_hits, int th_early_stop,
                            int opt_num_intersected, int32_t* hit_idx_coarse,
                            float* tmin_coarse, float* tmax_coarse, int32_t* hit_idx,
                            float* tmin, float* tmax, float* valid_len, int32_t* status,
                            void* stream);

/* Replaces MultiLayerOctree::ray_voxels_points_sampling (sampling_kernel.cu:
 * 1178-1248; kernels ray_voxels_sampling_kernel :911-982 and
 * ray_voxels_further_sampling_kernel :984-1068), the adaptive sampler:
 * per hit max(1, (int)((max-min) * start_size/size * S / valid_len)) samples
 * at min + off (i+1) - off/2 + off u, slots in hit order up to S, the rest
 * in the smallest hit node.  noise f32 [B,R,S]; outputs sampled_idx int32,
 * sampled_t f32 [B,R,S] (-1 for unused slots);
 * start_scale = 2^(levels - start_level - 1); max_hits (<= n_max_hits) is
 * the largest per-ray entry count of the traversal (status[1] of
 * sailor_octree_traversal): the per-hit pass visits slots [0, max_hits). */
int sailor_octree_sampling(const int32_t* hit_idx, const float* tmin, const float* tmax,
                           const float* noise, const float* valid_len, const float* node_pos,
                           int n_nodes, const float* voxel_size, int B, int R, int n_max_hits,
                           int max_hits, int S, int start_scale, int32_t* sampled_idx,
                           float* sampled_t, void* stream);

/* Replaces MultiLayerOctree::trilinear_aggregate_features (sampling_kernel.cu:
 * 1384-1421, kernel :162-234): feats f32 [B,C,R,S,8] (per sample the values
 * at the 8 corners of its node, corner id 4x+2y+z) -> out f32 [B,C,R,S],
 * trilinear in the point ray_ori + ray_dir * sampled_t; 0 where
 * sampled_idx is -1 (or outside [0, n_nodes)).  feats 16-byte aligned. */
int sailor_trilinear_aggregate(const float* node_pos, int n_nodes, const float* ray_ori,
                               const float* ray_dir, const int32_t* sampled_idx,
                               const float* sampled_t, const float* feats, int B, int R, int S,
                               int C, float* out, void* stream);

/* Replaces MultiLayerOctree::trilinear_aggregate_features_backward
 * (sampling_kernel.cu:1423-1462, kernel :237-327): grad_out f32 [B,C,R,S]
 * -> grad_feats f32 [B,C,R,S,8] (corner weight times grad; 0 where
 * sampled_idx is -1).  grad_feats 16-byte aligned. */
int sailor_trilinear_aggregate_backward(const float* node_pos, int n_nodes,
                                        const float* ray_ori, const float* ray_dir,
                                        const int32_t* sampled_idx, const float* sampled_t,
                                        const float* grad_out, int B, int R, int S, int C,
                                        float* grad_feats, void* stream);

/* Replaces MultiLayerOctree::project_sampled_xyz (sampling_kernel.cu:1333-1381,
 * kernel project_xyz :55-157): per (batch, view, ray, sample) the sample
 * point (and with record_corners its node's 8 corners, corner id 4x+2y+z)
 * projected by RTs [B,V,3,4] and Ks [B,V,3,3] to (p/z - w/2) / (w/2) ->
 * proj_xy f32 [B,V,R,S,(9,)2], camera z -> proj_z [B,V,R,S,(9,)1]; -1 where
 * sampled_idx is -1 (corners also -1 for an index outside [0, n_nodes)). */
int sailor_project_sampled_xyz(const float* node_pos, int n_nodes, const float* ray_ori,
                               const float* ray_dir, const int32_t* sampled_idx,
                               const float* sampled_t, const float* Ks, const float* RTs, int B,
                               int V, int R, int S, int ori_w, int ori_h, int record_corners,
                               float* proj_xy, float* proj_z, void* stream);

/* Replaces MultiLayerOctree::sort_samplings (sampling_kernel.cu:1308-1329):
 * ascending sort of each row of S values; ties keep their original order.
 *  t f32 [n_rows,S] -> sorted f32 [n_rows,S], order int32 [n_rows,S]       */
int sailor_sort_samplings(const float* t, int n_rows, int S, float* sorted,
                          int32_t* order, void* stream);

/* Fused form of the chain rays_calculating (sampling_kernel.cu:674-733) ->
 * MultiLayerOctree::voxel_traversal (intersection_kernel.cu:827-895, inference
 * mode, no early stop) -> ray_voxels_points_sampling_coarse
 * (sampling_kernel.cu:1251-1304) for one target view per batch, without the
 * [B,R,1024] hit lists: the DDA is walked again instead of storing hits.
 * Samples are bit-identical to the unfused chain.
 *  K, R f32 [B,3,3], T f32 [B,3]: target camera per batch; image H x W.
 *  noise f32 [B,H*W,S]; outputs sampled_idx int32 [B,H*W,S] and
 *  sampled_t f32 [B,H*W,S] (-1 padded, unsorted, as the reference returns
 *  them); rays_out f32 [B,H*W,6] and valid_len f32 [B,H*W] may be NULL.    */
int sailor_trace_sample(const float* K, const float* R, const float* T, int B,
                        int H, int W, const int32_t* vol_coarse,
                        const float* node_pos, const int32_t* node_info,
                        int n_nodes, const float* vol_bbox,
                        const float* voxel_size, int dim_fine, int dim_coarse,
                        int start_scale, int S, const float* noise,
                        float* rays_out, int32_t* sampled_idx, float* sampled_t,
                        float* valid_len, void* stream);

/* ------------------------------------------------------------------------ */
/* Rays                                                                      */
/* ------------------------------------------------------------------------ */

/* Replaces rays_calculating (sampling_kernel.cu:674-733) and
 * rays_calculating_parallel (:609-670, row interleave h*num_gpus+device).
 *  K,R f32 [B,V,3,3], T f32 [B,V,3] -> out f32 [B,V,H,W,6] (origin, dir),
 *  image row of output row h = h*row_stride + row_offset.                   */
int sailor_rays_calculating(const float* K, const float* R, const float* T,
                            int B, int V, int H, int W, int row_stride,
                            int row_offset, float* out, void* stream);

/* ------------------------------------------------------------------------ */
/* Volume compositing                                                        */
/* ------------------------------------------------------------------------ */

/* Replaces volume_rendering_occ_forward (volume_rendering.cu:749-787; kernel
 * :666-745), generalised from 3 to F feature channels.
 *  t f32 [B,R,S] (sorted), occ f32 [B,R,S], feats f32 [B,R,S,F] ->
 *  out_feats [B,R,F], out_depth/out_ws/out_alpha [B,R].                     */
int sailor_volume_rendering_occ_forward(const float* t, const float* occ,
                                        const float* feats, int B, int R, int S,
                                        int F, float t_thresh, float* out_feats,
                                        float* out_depth, float* out_ws,
                                        float* out_alpha, void* stream);

/* Replaces volume_rendering_training_forward (volume_rendering.cu:525-583;
 * kernels :17-126 density form, :129-240 occupancy form).                   */
int sailor_volume_rendering_training_forward(
    const float* t, const int32_t* sort_idx, const float* sigma,
    const float* feats, int B, int R, int S, int F, float t_thresh,
    int support_occ, float* out_feats, float* out_depth, float* out_ws,
    float* out_alpha, void* stream);

/* Replaces volume_rendering_training_backward (volume_rendering.cu:585-663;
 * kernels :244-381, :384-522).  grad_feats [B,R,S,F], grad_sigma [B,R,S]
 * are fully overwritten (zero where the reference leaves zero).             */
int sailor_volume_rendering_training_backward(
    const float* g_feats, const float* g_depth, const float* g_ws,
    const float* g_alpha, const float* out_feats, const float* out_depth,
    const float* out_ws, const float* out_alpha, const float* t,
    const int32_t* sort_idx, const float* sigma, const float* feats, int B,
    int R, int S, int F, float t_thresh, int support_occ, float* grad_feats,
    float* grad_sigma, void* stream);

/* Replaces _volume_render_torch.forward / .forward_occ
 * (c_lib/VoxelEncoding/voxel_encoding_helper.py:132-164 / :166-186), the
 * PyTorch compositing the reference trains with: samples in the order given,
 * no sort index and no early termination.  occ_mode 0 (density form):
 * delta_i = t_{i+1} - t_i with delta_{S-1} = 0, alpha_i = 1 - exp(-delta_i
 * sigma_i), T_i = exp(-sum_{j<i} delta_j sigma_j); occ_mode 1 (occupancy
 * form): alpha_i = sigma_i, T_i = prod_{j<i} (1 - alpha_j + 1e-6).  w_i =
 * alpha_i T_i; out_feats [n_rays,F] = sum w_i f_i, out_depth / out_ws /
 * out_alpha [n_rays] = sum w_i t_i / sum w_i / sum w_i alpha_i.
 *  t, sigma f32 [n_rays,S]; feats f32 [n_rays,S,F]; valid uint8 [n_rays] or
 *  NULL (rays with 0 output zeros, the reference's valid_rays mask).
 *  compensated != 0: Kahan-compensated exponent prefix sum and output sums
 *  (no counterpart in the reference). */
int sailor_volume_render_torch_forward(const float* t, const float* sigma, const float* feats,
                                       const unsigned char* valid, long long n_rays, int S,
                                       int F, int occ_mode, int compensated, float* out_feats,
                                       float* out_depth, float* out_ws, float* out_alpha,
                                       void* stream);

/* Gradient of sailor_volume_render_torch_forward (what torch autograd gives
 * the reference through the same tensor ops) with respect to sigma and
 * feats, from the gradients of its four outputs: grad_sigma [n_rays,S],
 * grad_feats [n_rays,S,F] (fully overwritten; zero for invalid rays).
 * Sample depths are constants, as in the reference's training step.
 * F <= 64. */
int sailor_volume_render_torch_backward(const float* g_feats, const float* g_depth,
                                        const float* g_ws, const float* g_alpha, const float* t,
                                        const float* sigma, const float* feats,
                                        const unsigned char* valid, long long n_rays, int S,
                                        int F, int occ_mode, float* grad_sigma,
                                        float* grad_feats, void* stream);

/* ------------------------------------------------------------------------ */
/* Fused radiance evaluation                                                 */
/* ------------------------------------------------------------------------ */

/* Number of floats of the small-weight buffer (occupancy head, Hydra
 * encoder, RGB head) for feature width F and hidden width H; the layout is
 * given in sailor_amd/csrc/mlp_layout.h and written by sailor_amd/model.py. */
size_t sailor_render_small_weights_floats(int F, int H);

/* Bytes of device workspace sailor_render_rays needs for B x R rays of
 * feature width F seen from up to 4 views (0 for invalid sizes). */
size_t sailor_render_workspace_bytes(int B, int R, int F);

/* The same for V views (1 <= V <= 8): 5 to 8 views use 8 view slots per
 * sample, so the per-ray token block doubles. */
size_t sailor_render_workspace_bytes_views(int B, int R, int F, int V);

/* Fused form of the FastNerf inference path (c_lib/FastNerf/src/infer_mlp.cu:
 * infer_fast_rendering :638-789 -- density_feat_kernel per (point, view),
 * mean over views, density_sigma_kernel, per-view feature compositing,
 * hydra_attention_kernel, RGB head) followed by the occupancy compositing of
 * volume_rendering_occ_forward (volume_rendering.cu:666-745), with feature
 * gather (proj_persp + grid_sample, utils_render.py:380-410,
 * SRONet/BasicRenNet.py:245-259) and frequency encoding fused in.  Widths are
 * template parameters: F (feature channels) in {32, 64}, H (hidden) in
 * {128, 256}; Hydra inner width 32, RGB head 64-32-3.
 *  t        f32 [B,R,S] sorted sample distances (-1 padded rows are dropped)
 *  rays     f32 [B,R,6] origin and unit direction
 *  Kv, RTv  f32 [B,V,3,3], [B,V,3,4] source cameras (1 <= V <= 8); img_h x
 *           img_w images
 *  fmap     [B,V,Hf,Wf,F] channels-last per-view feature maps, in fp32
 *           (fmap_precision 0) or in the form's 16-bit type (fmap_precision
 *           == precision; FastNerf keeps its features in half precision)
 *  bbox     f32 [B,3,2] volume box (frame of the positional encoding)
 *  zdir     f32 [B,3] third row of the target rotation (depth = t * zdir.dir)
 *  precision 0 = fp32 (MFMA f32), 1 = bf16, 2 = fp16 (MFMA, fp32 accumulate)
 *  w_dense  packed dense layers (sailor_amd/model.py), b_dense f32 [4,H]
 *  early_exit  flags: bit 0 -- stop evaluating a ray once its transmittance
 *              is below t_thresh (outputs are unchanged); bit 1 -- Kahan-
 *              compensated compositing sums (depth, weight and alpha sums,
 *              per-view tokens; precision form, no counterpart in the
 *              reference); bit 2 -- in the 16-bit forms, run the Hydra
 *              encoder and RGB head on fp32 MFMA instead of 16-bit MFMA
 *              (the differentiation form, whose backward recomputes them
 *              in fp32)
 *  chunk_samples  launch shape: bits 0-7 = samples evaluated together per
 *              workgroup step (up to 4 views: 8, 16, 24 or 48; 5-8 views:
 *              fp32 8, 12 or 16, 16-bit 4, 8, 12 or 16; 0 = up to 4
 *              views: 24 when S is a multiple of 24, else the largest of
 *              16 and 8 that divides S; more than 4 views: 12, or 8 when S
 *              is not a multiple of 12); S must be a
 *              multiple of it; a chunk whose activations do not fit the
 *              160 KiB of LDS of a workgroup (fp32 hidden-256 with 48
 *              samples) is rejected.  Bits 8-15 = percent of
 *              the device's resident workgroup slots the sample kernel
 *              launches (0 = all), for frames that share the chip with
 *              other streams' work
 *  out_rgb f32 [B,R,3], out_depth f32 [B,R]; out_aux f32 [B,R,2] (weight sum,
 *  alpha sum) and out_occ f32 [B,R,S] (per-sample occupancy) may be NULL.
 *  workspace  device buffer of at least sailor_render_workspace_bytes_views(B,
 *           R, F, V) bytes (list of rays with samples and their composited
 *           per-view features, handed from the sample kernel to the
 *           Hydra/RGB kernel); calls that overlap on one workspace race.    */
int sailor_render_rays(const float* t, const float* rays, int B, int R, int S,
                       const float* Kv, const float* RTv, int V, int img_h,
                       int img_w, const void* fmap, int fmap_precision, int Hf,
                       int Wf, int F,
                       const float* bbox, const float* zdir, int H,
                       int precision, const void* w_dense, const float* b_dense,
                       const float* w_small, float t_thresh, int early_exit,
                       int chunk_samples, float* out_rgb, float* out_depth,
                       float* out_aux, float* out_occ, void* workspace,
                       size_t workspace_bytes, void* stream);

/* sailor_render_rays split into its three launches, for profiling and for
 * callers that interleave other work: `stages` is a mask of 1 (reset the ray
 * list, list the rays with samples, write the outputs of dropped rays), 2
 * (the sample kernel: gather, encoding, MLP, occupancy, compositing; depth,
 * sums, occupancy and per-view features) and 4 (Hydra + RGB head: colour).
 * A stage reads the workspace left by the earlier stages of a call with the
 * same inputs; stages = 7 is sailor_render_rays.  Same arguments otherwise;
 * no counterpart in the reference (which runs the path as one call). */
int sailor_render_rays_stages(int stages, const float* t, const float* rays, int B, int R,
                              int S, const float* Kv, const float* RTv, int V, int img_h,
                              int img_w, const void* fmap, int fmap_precision, int Hf,
                              int Wf, int F,
                              const float* bbox, const float* zdir, int H, int precision,
                              const void* w_dense, const float* b_dense,
                              const float* w_small, float t_thresh, int early_exit,
                              int chunk_samples, float* out_rgb, float* out_depth,
                              float* out_aux, float* out_occ, void* workspace,
                              size_t workspace_bytes, void* stream);

/* Per-channel input quantisation of the precision form (no counterpart in
 * the reference): x f32 [rows, C] channels-last -> out [rows, C] in bf16
 * (precision 1) or fp16 (precision 2) holding x / s_c, scales f32 [C] with
 * s_c the smallest power of two >= max |x[:, c]| (1 for an all-zero
 * channel).  C must divide 256 or be a multiple of it; workspace: 4 C
 * bytes.  No memsets (graph-capturable). */
int sailor_quantize_channels(const float* x, long long rows, int C, int precision, void* out,
                             float* scales, void* workspace, size_t workspace_bytes,
                             void* stream);

/* Folds per-channel scales into a packed layer: out[i] = w[i] * scales[k]
 * if k = kidx[i] < F else w[i], converted to bf16 (precision 1) or fp16
 * (precision 2).  w f32 [n] (the packed layer in fp32), kidx int16 [n] (the
 * input index of every packed slot).  Used on the first dense layer, whose
 * inputs k < F are the feature channels quantised by
 * sailor_quantize_channels. */
int sailor_scale_packed_layer(const float* w, const int16_t* kidx, long long n, int F,
                              const float* scales, int precision, void* out, void* stream);

/* sailor_render_rays_stages on maps quantised by sailor_quantize_channels
 * (precision form, no counterpart in the reference): fmap holds x / s_c in
 * the form's 16-bit type, feature_scales f32 [F] the scales s_c, and
 * w_dense's first layer must have its feature inputs multiplied by s_c
 * (sailor_scale_packed_layer); the per-view tokens are scaled back by s_c.
 * flags as early_exit of sailor_render_rays (bit 0 early exit, bit 1
 * compensated sums).  With power-of-two scales every product and sum equals
 * the unquantised 16-bit form's for inputs within its range. */
int sailor_render_rays_quantized(int stages, const float* t, const float* rays, int B, int R,
                                 int S, const float* Kv, const float* RTv, int V, int img_h,
                                 int img_w, const void* fmap, int Hf, int Wf, int F,
                                 const float* feature_scales, const float* bbox,
                                 const float* zdir, int H, int precision, const void* w_dense,
                                 const float* b_dense, const float* w_small, float t_thresh,
                                 int flags, int chunk_samples, float* out_rgb, float* out_depth,
                                 float* out_aux, float* out_occ, void* workspace,
                                 size_t workspace_bytes, void* stream);

/* ------------------------------------------------------------------------ */
/* Differentiation of the fused render                                       */
/* ------------------------------------------------------------------------ */

/* Floats of the per-workgroup partial-gradient buffer the backward kernels
 * use (one small-weight gradient per workgroup of the device). */
size_t sailor_render_backward_partials_floats(int F, int H);

/* Backward of the Hydra encoder + RGB head of sailor_render_rays (the
 * autograd of aggregation.py:101-128,164-171,299-307 and nerf.py:117-124):
 * from dL/d rgb [B*R, 3] (ray order) and the forward's workspace (ray list
 * and per-view tokens) to dL/d token [n_valid, 4, F] (list order, in
 * grad_tokens of B*R*4*F floats) and the gradient of the Hydra and RGB-head
 * entries of the small-weight buffer (grad_small, small_weights_floats(F, H)
 * floats, overwritten; other entries zero).  partials: scratch of
 * partials_floats floats (sailor_render_backward_partials_floats). */
int sailor_render_hydra_backward(const float* grad_rgb, int B, int R, int V, int F, int H,
                                 const float* w_small, const void* workspace,
                                 size_t workspace_bytes, float* grad_tokens, float* grad_small,
                                 float* partials, size_t partials_floats, void* stream);

/* Backward of the sample kernel of sailor_render_rays (compositing reverse
 * scan + per-view MLP, per-ray recompute): for the list entries
 * [item_begin, item_end) of the forward's workspace (entries past the
 * listed-ray count are skipped, their columns left untouched) it writes, per column c = (item - item_begin) * S
 * * 4 + sample * 4 + view, the inputs A_l and deltas D_l of the four dense
 * layers (col_bufs = {A0, A1, A2, A3, D0, D1, D2, D3}, each [C, H] in the
 * form's type, C >= (item_end - item_begin) * S * 4, F + 64 <= H; A0 holds the first layer's inputs in kernel order
 * [features | encoding | 0] in its first F + 64 columns), so that
 * dW_l = D_l^T A_l and db_l = sum_c D_l; and adds the occupancy-head
 * gradient to grad_small.  Inputs as sailor_render_rays plus w_dense_t
 * (the packed W_3^T, W_2^T, W_1^T), occ (the forward's per-sample
 * occupancy, [B*R, S]) and grad_tokens (sailor_render_hydra_backward).
 * fmap is fp32; S <= 64, multiple of 4 (fp32) or 8 (16-bit; 16 for the
 * fused variant below). */
int sailor_render_sample_backward(const float* t, const float* rays, int B, int R, int S,
                                  const float* Kv, const float* RTv, int V, int img_h,
                                  int img_w, const float* fmap, int Hf, int Wf, int F,
                                  const float* bbox, int H, int precision, const void* w_dense,
                                  const void* w_dense_t, const float* b_dense,
                                  const float* w_small, float t_thresh, const float* occ,
                                  const float* grad_tokens, const void* workspace,
                                  size_t workspace_bytes, long long item_begin,
                                  long long item_end, void* const* col_bufs, float* grad_small,
                                  float* partials, size_t partials_floats, void* stream);

/* Floats of a dense-layer gradient buffer: dW_0 [H][F + 64] (first-layer
 * inputs in kernel order [features | encoding | 0]), dW_1..3 [H][H],
 * db [4][H]. */
size_t sailor_render_dense_grad_floats(int F, int H);

/* The same backward over all listed rays with the weight gradients summed
 * inside the kernel (MFMA accumulators held across the workgroup's rays,
 * one partial per workgroup, then a sum): writes grad_dense
 * (sailor_render_dense_grad_floats(F, H) floats, overwritten) and adds the
 * occupancy-head gradient to grad_small.  No column buffers and no host
 * synchronisation.  H = 128 models only (the H = 256 gradient does not fit
 * on chip; use the column form).  partials: at least
 * sailor_render_backward_partials_floats(F, H) floats.  Together with
 * sailor_render_hydra_backward this is the differentiation of the fused
 * render that the reference obtains from torch autograd through
 * SRONet/nerf.py:139-186 and _volume_render_torch.forward_occ
 * (voxel_encoding_helper.py:166-186). */
int sailor_render_sample_backward_fused(const float* t, const float* rays, int B, int R, int S,
                                        const float* Kv, const float* RTv, int V, int img_h,
                                        int img_w, const float* fmap, int Hf, int Wf, int F,
                                        const float* bbox, int H, int precision,
                                        const void* w_dense, const void* w_dense_t,
                                        const float* b_dense, const float* w_small,
                                        float t_thresh, const float* occ,
                                        const float* grad_tokens, const void* workspace,
                                        size_t workspace_bytes, float* grad_dense,
                                        float* grad_small, float* partials,
                                        size_t partials_floats, void* stream);

/* Density query at arbitrary points: replaces infer_density_mlp
 * (c_lib/FastNerf/src/infer_mlp.cu:583-634, the GUI's occupancy query after
 * fusion) for the fused model: per point, the per-view MLP on the gathered
 * features and the encoding, the view mean and the occupancy head.
 *  points f32 [B,N,3]; other inputs as sailor_render_rays (zdir, t and the
 *  compositing arguments do not apply); chunk_samples = points per group
 *  (0 = default).  out_density f32 [B,N]. */
int sailor_density_points(const float* points, int B, int N, const float* Kv,
                          const float* RTv, int V, int img_h, int img_w, const void* fmap,
                          int fmap_precision, int Hf, int Wf, int F, const float* bbox, int H,
                          int precision, const void* w_dense, const float* b_dense,
                          const float* w_small, int chunk_samples, float* out_density,
                          void* stream);

/* ------------------------------------------------------------------------ */
/* Frequency encoding                                                        */
/* ------------------------------------------------------------------------ */

/* Replaces freq_encode_forward/backward (freq_encoding.cu:89-151).
 *  x f32 [N,D] -> y f32 [N, D + 2*D*deg] = [x, sin(2^f x), sin(2^f x + pi/2) ...] */
int sailor_freq_encode_forward(const float* x, int N, int D, int deg, float* y,
                               void* stream);
int sailor_freq_encode_backward(const float* grad_y, const float* y, int N,
                                int D, int deg, float* grad_x, void* stream);

/* ------------------------------------------------------------------------ */
/* Preprocessing                                                            */
/* ------------------------------------------------------------------------ */

/* Replaces depth_normalize (c_lib/VoxelEncoding/src/depth_normalizer.cu:
 * 98-189), bound as VoxelEncoding.depth_normalize (binding.cpp:38).
 *  depths f32 [B,V,H,W]; per view: min / max over depths in [valid_min,
 *  valid_max] (99999 / -99999 when none), distance = max - min + 1e-10,
 *  mid = (max + min) / 2.  user_dist < 0: out = clamp((d - mid) [* mask] /
 *  distance [* 2 if divided2], -1, 1); else (d - mid) * mask / user_dist
 *  with multi_mask, clamp((d - mid) / user_dist, -1, 1) without.
 *  mask = 0 where d is outside the valid range (multi_mask), else 1.
 *  Outputs out_normalized, out_mask [B,V,H,W], out_mid, out_dist [B,V];
 *  workspace: 8 * B * V bytes. */
int sailor_depth_normalize(const float* depths, int B, int V, int H, int W, float valid_min,
                           float valid_max, float user_dist, int multi_mask, int divided2,
                           float* out_normalized, float* out_mask, float* out_mid,
                           float* out_dist, void* workspace, size_t workspace_bytes,
                           void* stream);

/* Replaces rgb_normalizer (utils_render/utils_render.py:601-605):
 * out = ((rgb - mid) / dist) * mask, rgb / out f32 [N,C,H,W], mask f32
 * [N,1,H,W] broadcast over channels. */
int sailor_rgb_normalize(const float* rgb, const float* mask, int N, int C, int H, int W,
                         float mid, float dist, float* out, void* stream);

/* Replaces rays_selecting (sampling_kernel.cu:489-527,736-769): rows of
 * ray_ori_dir [B,V,H,W,6] at sampled_xy int32 [B,V,N,2] (x, y) -> out
 * [B,V,N,6].  An (x, y) outside the image selects zeros (the reference
 * reads out of bounds). */
int sailor_rays_selecting(const float* ray_ori_dir, int B, int V, int H, int W,
                          const int* sampled_xy, int n_rays, float* out, void* stream);

/* Replaces rgbdv_selecting (sampling_kernel.cu:530-607,771-825): rgbs
 * [B,V,H,W,C], normals [B,V,H,W,3], depths [B,V,H,W] at sampled_xy
 * [B,V,N,2] -> out_rgb [B,V,N,C], out_normals [B,V,N,3], out_depth
 * [B,V,N], out_label [B,V,N] = 1 where the depth is > 0. */
int sailor_rgbdv_selecting(const float* rgbs, int rgb_channels, const float* normals,
                           const float* depths, int B, int V, int H, int W,
                           const int* sampled_xy, int n_rays, float* out_rgb,
                           float* out_normals, float* out_depth, float* out_label,
                           void* stream);

/* ------------------------------------------------------------------------ */
/* Per-view sample features for the FastNerf surface                        */
/* ------------------------------------------------------------------------ */

/* [N][C][H][W] -> [N][H][W][C] (the channels-last maps sailor_sample_features
 * reads). */
int sailor_nchw_to_nhwc(const float* src, int N, int C, int H, int W, float* dst, void* stream);

/* The chain the reference runs before InferNerf.infer_nerf
 * (gui/gui_render.py:713-751; SRONet/BasicRenNet.py:245-259,296-311):
 * points o + t d, proj_persp with the sample distances as p_depths
 * (utils_render.py:380-410: grid -2 where the camera depth < 0.01 or t is
 * outside [0.1, 4.0], projected depth -2 for the latter), z_normalizer
 * (:656-663), grid_sample (align_corners=True, bilinear, zero padding) of the
 * geometry features, depth map, colour features and normalised colours, and
 * t-psdf = clamp(z - sampled depth, -0.01, 0.01) / 0.01; plus the ray
 * direction in each camera (R d).
 *  rays [B,R,6], dists [B,R,S], K [B,V,3,3], RT [B,V,3,4];
 *  geo_cl [B,V,Hg,Wg,16], depth [B,V,Hd,Wd], rgbf_cl [B,V,Hr,Wr,16],
 *  rgb_cl [B,V,Hc,Wc,3] (channels-last); depth_mid, depth_dist [B,V];
 *  res: the image size of proj_persp's grid normalisation ((p - res/2) /
 *  (res/2), integer halves).
 *  Outputs in (batch, view, ray, sample) order: out_z, out_tpsdf [B,V,R,S],
 *  out_geo, out_rgbf [B,V,R,S,16], out_rgb, out_dir [B,V,R,S,3]. */
int sailor_sample_features(const float* rays, const float* dists, int B, int R, int S,
                           const float* K, const float* RT, int V, const float* geo_cl, int Hg,
                           int Wg, const float* depth, int Hd, int Wd, const float* rgbf_cl,
                           int Hr, int Wr, const float* rgb_cl, int Hc, int Wc,
                           const float* depth_mid, const float* depth_dist, int res,
                           float* out_z, float* out_tpsdf, float* out_geo, float* out_rgbf,
                           float* out_rgb, float* out_dir, void* stream);

/* ------------------------------------------------------------------------ */
/* FastNerf inference surface (InferNerf, c_lib/FastNerf/src/binding.cpp)   */
/* ------------------------------------------------------------------------ */
/* Parameters are the reference's packed half vectors (gui/gui_render.py:
 * 658-685): density 22521 (feature MLP 18->64x4 | occupancy MLP 64->64->64->1
 * at 13696 | geometry MLP 18->16->8 at 22081), colour 15807 (two Hydra stacks
 * of 5822 | RGB MLP 30->64->32->3 at 11644), blending 15075; each linear
 * layer is W [out][in] row-major then b [out].  Arithmetic is fp32 on the
 * half parameters (the reference accumulates in half). */

/* Workspace of sailor_infer_nerf: per-sample occupancy + the texture views'
 * 8 geometry features per sample. */
size_t sailor_infer_nerf_workspace_bytes(int B, int R, int S, int tex_views);

/* Replaces infer_fast_rendering (c_lib/FastNerf/src/infer_mlp.cu:638-789),
 * bound as InferNerf.infer_nerf (binding.cpp:8).
 *  sorted_dists f32 [B,R,S]; per view, in (batch, view, ray, sample) order:
 *  points_z, points_tpsdf f32 [B*Vg*R*S], feats_geo f32 [B*Vg*R*S,16];
 *  rgbs, dirs f32 [B*Vt*R*S,3], feats_rgb f32 [B*Vt*R*S,16];
 *  1 <= tex_views <= min(geo_views, 4).
 *  out_depth [B,R] = sum_i w_i t_i, out_rgb [B,R,3], out_feats [B,R,30]
 *  (the first view's token after the Hydra stacks).  Kept from the
 *  reference: q and k normalised without epsilon (a ray with no samples gets
 *  0/0), samples with t < 0 skipped, rays whose last t < 0 composite to 0,
 *  compositing stops after T < 1e-4. */
int sailor_infer_nerf(const float* sorted_dists, const float* points_z,
                      const float* points_tpsdf, const float* feats_geo, const float* rgbs,
                      const float* feats_rgb, const float* dirs, const void* params_density,
                      const void* params_color, int B, int R, int S, int geo_views,
                      int tex_views, float* out_depth, float* out_rgb, float* out_feats,
                      void* workspace, size_t workspace_bytes, void* stream);

/* sailor_infer_nerf computes the density MLPs (the reference's half
 * precision) on fp16 MFMA: layer inputs rounded to half, fp32 accumulation,
 * for up to 8 geometry views (more views use the fp32 kernels).  The _fp32
 * variant keeps every activation in fp32 on the vector ALUs (no counterpart
 * in the reference; for callers that want the exact value of the half
 * parameters). */
int sailor_infer_nerf_fp32(const float* sorted_dists, const float* points_z,
                           const float* points_tpsdf, const float* feats_geo, const float* rgbs,
                           const float* feats_rgb, const float* dirs, const void* params_density,
                           const void* params_color, int B, int R, int S, int geo_views,
                           int tex_views, float* out_depth, float* out_rgb, float* out_feats,
                           void* workspace, size_t workspace_bytes, void* stream);

/* Replaces infer_density_mlp (infer_mlp.cu:583-634), bound as
 * InferNerf.infer_density: points_z, points_tpsdf [B*V*N], feats_geo
 * [B*V*N,16] in (batch, view, point) order -> out [B,N] occupancy. */
int sailor_infer_density(const float* points_z, const float* points_tpsdf,
                         const float* feats_geo, const void* params_density, int B, int N,
                         int V, float* out, void* stream);

/* fp32 vector-ALU form of sailor_infer_density (no counterpart in the
 * reference; as sailor_infer_nerf_fp32). */
int sailor_infer_density_fp32(const float* points_z, const float* points_tpsdf,
                              const float* feats_geo, const void* params_density, int B, int N,
                              int V, float* out, void* stream);

/* Replaces infer_upsampling (infer_mlp.cu:793-821), bound as
 * InferNerf.infer_blending: feats_neighbors [N,40], ray_feats [N,30] ->
 * out [N,3] blend weights. */
int sailor_infer_blending(const float* feats_neighbors, const float* ray_feats,
                          const void* params_blend, int N, float* out, void* stream);

/* Replaces rays_interpolate (gui/gui_render.py:753-773,
 * upsampling/BasicRenNet.py:348-367): per camera c, K[c], R[c] [3,3] and
 * dirs [c][n_rays][3] -> out [c][n_rays*4][3], element ray*4+k the
 * normalised R^-1 K^-1 [x + 0.5 (k&1), y + 0.5 (k>>1), 1] of the pixel
 * (x, y) of K R dir. */
int sailor_rays_interpolate(const float* K, const float* R, const float* dirs, int n_cams,
                            int n_rays, float* out, void* stream);

/* Replaces the body of gui_render.forward_upsampling + sample_neighbor
 * (gui/gui_render.py:775-838) for one target camera: for every ray with
 * valid_min < depth < valid_max, its 4 sub-pixel rays (rays_interpolate),
 * surface points ray_ori + dir * depth projected into neighbour views
 * nei0, nei1 (proj_persp, utils_render.py:380-410, grid scale res/2),
 * bilinear colour features rgbf_cl [V,Hf,Wf,16], nearest colours rgb_cl
 * [V,Hc,Wc,3] and depths [V,Hd,Wd], exp(-200 (z - d)^2), the blending MLP
 * (15075 halves, as sailor_infer_blending) on [feats | R_v dir | weight]
 * x 2 and ray_feats [N,30], and clamp(w0 c0 + w1 c1 + w2 coarse_rgb, 0, 1)
 * -> out [N,4,3]; other rays get zeros.  K, R [V,3,3], T [V,3] are the
 * source views; tar_K, tar_R [3,3] the target camera. */
int sailor_forward_upsampling(const float* ray_ori, const float* ray_dir, const float* depth,
                              const float* coarse_rgb, const float* ray_feats, int N,
                              const float* tar_K, const float* tar_R, const float* K,
                              const float* R, const float* T, int V, int nei0, int nei1,
                              const float* rgbf_cl, int Hf, int Wf, const float* rgb_cl, int Hc,
                              int Wc, const float* depths, int Hd, int Wd, int res,
                              float valid_min, float valid_max, const void* params_blend,
                              float* out, void* stream);

/* sailor_forward_upsampling runs the blending MLP (the reference's half
 * precision) on fp16 MFMA: layer inputs rounded to half, fp32 accumulation.
 * The _fp32 variant keeps it in fp32 on the vector ALUs (no counterpart in
 * the reference). */
int sailor_forward_upsampling_fp32(const float* ray_ori, const float* ray_dir,
                                   const float* depth, const float* coarse_rgb,
                                   const float* ray_feats, int N, const float* tar_K,
                                   const float* tar_R, const float* K, const float* R,
                                   const float* T, int V, int nei0, int nei1,
                                   const float* rgbf_cl, int Hf, int Wf, const float* rgb_cl,
                                   int Hc, int Wc, const float* depths, int Hd, int Wd, int res,
                                   float valid_min, float valid_max, const void* params_blend,
                                   float* out, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* SAILOR_HIP_H */
