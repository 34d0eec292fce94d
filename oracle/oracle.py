"""Python face of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg import
this module.  It wraps the C restatement in sailor_oracle.c (bit-exact
reference for integer / index outputs, float32 op order of the reference) and
adds float64 NumPy / torch restatements used as accuracy oracles.

How each part is pinned (tests/test_oracle.py; test names in brackets):
  * inverse3x3: numpy.linalg.inv  [test_inverse3x3_matches_linalg]
  * center_xyz: fronto-parallel plane closed form, float64 numpy
    back-projection of the capsule scene  [test_center_xyz_*]
  * bbox / resized bbox: closed form on a fronto-parallel plane, float64
    brute force over posed views, centred-cube closed form  [test_bbox_*,
    test_resized_bbox_is_centred_cube]
  * integrate: closed-form TSDF / weights / colour / occupancy of a plane
    seen by identical views, rounded running colour mean, float64
    restatement on the capsule scene  [test_integrate_*]
  * merge: brute-force OR over blocks  [test_merge_bruteforce]
  * tree: brute-force 26-neighbour isolation filter, centre closed form
    min + vs*((size-1)/2 + size*i), raster-ordered contiguous children
    [test_tree_bruteforce_and_centres]
  * rays: float64 R^-1 K^-1 closed form, row interleave  [test_rays_*]
  * traversal: float64 slab test against every allocated cell (hit set,
    children, valid length), axis-aligned row of cells  [test_traverse_*]
  * sampling: noise-0.5 closed form, per-interval counts, leftovers and
    interval containment  [test_sampling_*]
  * sort: ordering, permutation, tie stability  [test_sort_is_stable_and_ordered]
  * compositing: geometric series for constant alpha, early termination,
    dropped rays, density-mode closed form, central finite differences of
    the backward in both modes  [test_render_*]
  * frequency encoding: float64 sin / cos closed form and its derivative
    [test_freq_closed_form_and_gradient]
The reference ships no golden outputs for this path (its test_data frames
carry camera parameters only as pickled objects, which are not loaded).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
c_int, c_float, c_ll = ctypes.c_int, ctypes.c_float, ctypes.c_longlong


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            from oracle import build as _b
            _b.build(verbose=False)
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_inverse3x3.argtypes = [f32p, c_int, f32p]
        L.orc_origin_bbox.argtypes = [f32p, f32p, f32p, f32p, c_int, c_int, c_int, c_int,
                                      c_float, f32p]
        L.orc_center_xyz.argtypes = [f32p, f32p, f32p, f32p, c_int, c_int, c_int, c_int,
                                     np.ctypeslib.ndpointer(dtype=np.float64,
                                                            flags="C_CONTIGUOUS"),
                                     np.ctypeslib.ndpointer(dtype=np.float64,
                                                            flags="C_CONTIGUOUS")]
        L.orc_resized_bbox.argtypes = [f32p, c_int, c_int, f32p, f32p]
        L.orc_integrate.argtypes = [i32p, i32p, f32p, f32p, f32p, f32p] + [c_int] * 7 + \
            [c_float] * 4 + [f32p, f32p, f32p, f32p, f32p]
        L.orc_merge.argtypes = [i32p, c_int, c_int, c_int, i32p, i32p]
        L.orc_tree_build.argtypes = [i32p, i32p, c_int, c_int, c_int, f32p, f32p, c_int,
                                     c_int, c_int, f32p, i32p, i32p]
        L.orc_tree_build.restype = c_int
        L.orc_rays.argtypes = [f32p, f32p, f32p] + [c_int] * 6 + [f32p]
        L.orc_traverse.argtypes = [i32p, f32p, i32p, f32p, f32p] + [c_int] * 4 + \
            [f32p, f32p] + [c_int] * 5 + [i32p, f32p, f32p, f32p,
                                          ctypes.POINTER(c_int)]
        L.orc_traverse.restype = c_int
        L.orc_sampling.argtypes = [i32p, f32p, f32p, f32p, f32p, f32p, f32p] + \
            [c_int] * 6 + [i32p, f32p]
        L.orc_render_occ.argtypes = [f32p, f32p, f32p, c_ll, c_int, c_int, c_float,
                                     f32p, f32p, f32p, f32p]
        L.orc_render_train_fwd.argtypes = [f32p, i32p, f32p, f32p, c_ll, c_int, c_int,
                                           c_float, c_int, f32p, f32p, f32p, f32p]
        L.orc_render_train_bwd.argtypes = [f32p] * 8 + [f32p, i32p, f32p, f32p, c_ll,
                                                        c_int, c_int, c_float, c_int,
                                                        f32p, f32p]
        L.orc_freq_fwd.argtypes = [f32p, c_ll, c_int, c_int, f32p]
        L.orc_freq_bwd.argtypes = [f32p, f32p, c_ll, c_int, c_int, f32p]
        L.orc_octree_traversal.argtypes = [i32p, f32p, i32p, f32p, f32p] + [c_int] * 4 + \
            [f32p, f32p] + [c_int] * 5 + [i32p, f32p, f32p, i32p, f32p, f32p, f32p, i32p]
        L.orc_sampling_adaptive.argtypes = [i32p, f32p, f32p, f32p, f32p, f32p, f32p] + \
            [c_int] * 6 + [i32p, f32p]
        L.orc_trilinear_fwd.argtypes = [f32p, f32p, f32p, i32p, f32p, f32p] + [c_int] * 4 + \
            [f32p]
        L.orc_trilinear_bwd.argtypes = [f32p, f32p, f32p, i32p, f32p, f32p] + [c_int] * 4 + \
            [f32p]
        L.orc_project.argtypes = [f32p, f32p, f32p, i32p, f32p, f32p, f32p] + [c_int] * 7 + \
            [f32p, f32p]
        _lib = L
    return _lib


def _f(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _i(a):
    return np.ascontiguousarray(a, dtype=np.int32)


# --------------------------------------------------------------------------
# fusion
def inverse3x3(M):
    M = _f(M)
    out = np.empty_like(M)
    lib().orc_inverse3x3(M, M.size // 9, out)
    return out


def origin_bbox(depths, K, R, T, th_bbox):
    depths, K, R, T = _f(depths), _f(K), _f(R), _f(T)
    B, V, H, W = depths.shape
    out = np.empty((B, 3, 2), np.float32)
    lib().orc_origin_bbox(depths, K, R, T, B, V, H, W, th_bbox, out)
    return out


def center_xyz(depths, K, R, T):
    """get_center_xyz: (sum of back-projected valid points [B,3], count [B])
    in float64 (float32 point arithmetic)."""
    depths, K, R, T = _f(depths), _f(K), _f(R), _f(T)
    B, V, H, W = depths.shape
    c = np.zeros((B, 3), np.float64)
    n = np.zeros((B,), np.float64)
    lib().orc_center_xyz(depths, K, R, T, B, V, H, W, c, n)
    return c, n


def resized_bbox(origin, dim):
    origin = _f(origin)
    B = origin.shape[0]
    bbox = np.empty((B, 3, 2), np.float32)
    vs = np.empty((B,), np.float32)
    lib().orc_resized_bbox(origin, B, dim, bbox, vs)
    return bbox, vs


def integrate(K, RT, R, T, colors, depths, dims, obs_weight=1.0, th_bbox=0.025,
              trunc_weight=3.0, th_low=13.5, th_high=13.5, fixed_bbox=None):
    """fusion_cuda_integrate: returns dict(occ, num_occ, tsdf, weight, color,
    vol_origin, vol_bbox, voxel_size)."""
    colors, depths = _f(colors), _f(depths)
    K, RT, R, T = _f(K), _f(RT), _f(R), _f(T)
    B, V, H, W = depths.shape
    X, Y, Z = dims
    if fixed_bbox is None:
        origin = origin_bbox(depths, K, R, T, th_bbox)
        bbox, vs = resized_bbox(origin, X)
    else:
        origin = _f(fixed_bbox).reshape(B, 3, 2).copy()
        bbox = origin.copy()
        vs = ((bbox[:, 0, 1] - bbox[:, 0, 0]) / np.float32(X - 1)).astype(np.float32)
    occ = np.empty((B, X, Y, Z), np.int32)
    num = np.zeros((B,), np.int32)
    tsdf = np.empty((B, X, Y, Z), np.float32)
    weight = np.empty_like(tsdf)
    color = np.empty_like(tsdf)
    lib().orc_integrate(occ, num, K, RT, colors, depths, B, V, H, W, X, Y, Z,
                        obs_weight, trunc_weight, th_low, th_high, _f(bbox), _f(vs),
                        tsdf, weight, color)
    return dict(occ=occ, num_occ=num, tsdf=tsdf, weight=weight, color=color,
                vol_origin=origin, vol_bbox=bbox, voxel_size=vs)


# --------------------------------------------------------------------------
# tree
def merge(fine, rate):
    fine = _i(fine)
    B, df = fine.shape[0], fine.shape[1]
    dc = df // rate
    coarse = np.empty((B, dc, dc, dc), np.int32)
    cnt = np.zeros((B,), np.int32)
    lib().orc_merge(fine, B, dc, rate, coarse, cnt)
    return coarse, cnt


def tree_build(vol_fine, vol_coarse, voxel_size, vol_bbox, coarse_start, capacity,
               fine_start=0):
    vf, vc = _i(vol_fine).copy(), _i(vol_coarse).copy()
    B, df, dc = vf.shape[0], vf.shape[1], vc.shape[1]
    pos = np.empty((capacity, 4), np.float32)
    info = np.empty((capacity, 4), np.int32)
    q = np.zeros((2,), np.int32)
    rc = lib().orc_tree_build(vf, vc, B, df, dc, _f(voxel_size), _f(vol_bbox), fine_start,
                              coarse_start, capacity, pos, info, q)
    if rc != 0:
        raise RuntimeError("oracle tree_build: capacity exceeded")
    return dict(vol_fine=vf, vol_coarse=vc, node_pos=pos, node_info=info, queue=q)


# --------------------------------------------------------------------------
# rays / traversal / sampling
def rays(K, R, T, H, W, row_stride=1, row_offset=0):
    K, R, T = _f(K), _f(R), _f(T)
    B, V = K.shape[:2]
    out = np.empty((B, V, H, W, 6), np.float32)
    lib().orc_rays(K, R, T, B, V, H, W, row_stride, row_offset, out)
    return out


def traverse(vol_coarse, node_pos, node_info, vol_bbox, voxel_size, dim_fine, rate,
             ray_ori, ray_dir, n_max_hits, th_early_stop=-1, is_training=False,
             opt_num=-1):
    vc = _i(vol_coarse)
    B, dc = vc.shape[0], vc.shape[1]
    ray_ori, ray_dir = _f(ray_ori), _f(ray_dir)
    R = ray_ori.shape[1]
    idx = np.empty((B, R, n_max_hits), np.int32)
    t0 = np.empty((B, R, n_max_hits), np.float32)
    t1 = np.empty_like(t0)
    vl = np.empty((B, R), np.float32)
    ovf = c_int(0)
    gmax = lib().orc_traverse(vc, _f(node_pos), _i(node_info), _f(vol_bbox), _f(voxel_size),
                              B, dim_fine, dc, rate, ray_ori, ray_dir, R, n_max_hits,
                              th_early_stop, int(is_training), opt_num, idx, t0, t1, vl,
                              ctypes.byref(ovf))
    return dict(idx=idx, tmin=t0, tmax=t1, valid_len=vl, gmax=gmax, overflow=ovf.value)


def sampling(hit_idx, tmin, tmax, noise, valid_len, node_pos, voxel_size, S, start_scale,
             gmax):
    hit_idx = _i(hit_idx)
    B, R, n_max = hit_idx.shape
    sidx = np.empty((B, R, S), np.int32)
    st = np.empty((B, R, S), np.float32)
    lib().orc_sampling(hit_idx, _f(tmin), _f(tmax), _f(noise), _f(valid_len), _f(node_pos),
                       _f(voxel_size), B, R, n_max, S, start_scale, gmax, sidx, st)
    return sidx, st


def octree_traversal(vol_coarse, node_pos, node_info, vol_bbox, voxel_size, dim_fine, rate,
                     ray_ori, ray_dir, n_max_hits_coarse, n_max_hits, th_early_stop=-1,
                     opt_num=-1):
    """MultiLayerOctree::octree_traversal (recursive form), canonical order."""
    vc = _i(vol_coarse)
    B, dc = vc.shape[0], vc.shape[1]
    ray_ori, ray_dir = _f(ray_ori), _f(ray_dir)
    R = ray_ori.shape[1]
    out = dict(idx_c=np.empty((B, R, n_max_hits_coarse), np.int32),
               tmin_c=np.empty((B, R, n_max_hits_coarse), np.float32),
               tmax_c=np.empty((B, R, n_max_hits_coarse), np.float32),
               idx=np.empty((B, R, n_max_hits), np.int32),
               tmin=np.empty((B, R, n_max_hits), np.float32),
               tmax=np.empty((B, R, n_max_hits), np.float32),
               valid_len=np.empty((B, R), np.float32), status=np.zeros(3, np.int32))
    lib().orc_octree_traversal(vc, _f(node_pos), _i(node_info), _f(vol_bbox), _f(voxel_size), B,
                               dim_fine, dc, rate, ray_ori, ray_dir, R, n_max_hits_coarse,
                               n_max_hits, th_early_stop, opt_num, out["idx_c"], out["tmin_c"],
                               out["tmax_c"], out["idx"], out["tmin"], out["tmax"],
                               out["valid_len"], out["status"])
    return out


def sampling_adaptive(hit_idx, tmin, tmax, noise, valid_len, node_pos, voxel_size, S,
                      start_scale, max_hits=None):
    """MultiLayerOctree::ray_voxels_points_sampling, canonical slot order;
    max_hits = slots the per-hit pass visits (default: all)."""
    hit_idx = _i(hit_idx)
    B, R, n_max = hit_idx.shape
    max_hits = n_max if max_hits is None else int(max_hits)
    sidx = np.empty((B, R, S), np.int32)
    st = np.empty((B, R, S), np.float32)
    lib().orc_sampling_adaptive(hit_idx, _f(tmin), _f(tmax), _f(noise), _f(valid_len),
                                _f(node_pos), _f(voxel_size), B, R, n_max, max_hits, S,
                                start_scale, sidx, st)
    return sidx, st


def trilinear_fwd(node_pos, ray_ori, ray_dir, sidx, st, feats):
    """trilinear_aggregate_features: feats [B,C,R,S,8] -> [B,C,R,S]."""
    feats = _f(feats)
    B, C, R, S, _ = feats.shape
    out = np.empty((B, C, R, S), np.float32)
    lib().orc_trilinear_fwd(_f(node_pos), _f(ray_ori), _f(ray_dir), _i(sidx), _f(st), feats,
                            B, R, S, C, out)
    return out


def trilinear_bwd(node_pos, ray_ori, ray_dir, sidx, st, grad):
    """trilinear_aggregate_features_backward: grad [B,C,R,S] -> [B,C,R,S,8]."""
    grad = _f(grad)
    B, C, R, S = grad.shape
    out = np.empty((B, C, R, S, 8), np.float32)
    lib().orc_trilinear_bwd(_f(node_pos), _f(ray_ori), _f(ray_dir), _i(sidx), _f(st), grad,
                            B, R, S, C, out)
    return out


def project(node_pos, ray_ori, ray_dir, sidx, st, Ks, RTs, ori_w, ori_h, corners=False):
    """project_xyz: -> (xy [B,V,R,S,(9,)2], z [B,V,R,S,(9,)1])."""
    sidx = _i(sidx)
    B, R, S = sidx.shape
    V = np.asarray(Ks).shape[1]
    np_ = 9 if corners else 1
    xy = np.empty((B, V, R, S, np_, 2), np.float32)
    z = np.empty((B, V, R, S, np_, 1), np.float32)
    lib().orc_project(_f(node_pos), _f(ray_ori), _f(ray_dir), sidx, _f(st), _f(Ks), _f(RTs), B,
                      V, R, S, ori_w, ori_h, int(corners), xy, z)
    if not corners:
        xy, z = xy[..., 0, :], z[..., 0, :]
    return xy, z


def sort_samplings(t):
    """torch::sort(t, dim=-1) with stable tie order (sampling_kernel.cu:1321)."""
    order = np.argsort(t, axis=-1, kind="stable").astype(np.int32)
    return np.take_along_axis(t, order, axis=-1), order


# --------------------------------------------------------------------------
# compositing (float32 restatement)
def render_occ(t, occ, feats, t_thres=1e-4):
    t, occ, feats = _f(t), _f(occ), _f(feats)
    B, R, S = t.shape
    F = feats.shape[-1]
    n = B * R
    of = np.empty((B, R, F), np.float32)
    od, ow, oa = (np.empty((B, R), np.float32) for _ in range(3))
    lib().orc_render_occ(t, occ, feats, n, S, F, t_thres, of, od, ow, oa)
    return of, od, ow, oa


def render_train_fwd(t, sidx, sigma, feats, t_thres=1e-4, occ_mode=True):
    t, sidx, sigma, feats = _f(t), _i(sidx), _f(sigma), _f(feats)
    B, R, S = t.shape
    F = feats.shape[-1]
    of = np.empty((B, R, F), np.float32)
    od, ow, oa = (np.empty((B, R), np.float32) for _ in range(3))
    lib().orc_render_train_fwd(t, sidx, sigma, feats, B * R, S, F, t_thres, int(occ_mode),
                               of, od, ow, oa)
    return of, od, ow, oa


def render_train_bwd(g_feats, g_depth, g_ws, g_alpha, o_feats, o_depth, o_ws, o_alpha,
                     t, sidx, sigma, feats, t_thres=1e-4, occ_mode=True):
    t, sidx, sigma, feats = _f(t), _i(sidx), _f(sigma), _f(feats)
    B, R, S = t.shape
    F = feats.shape[-1]
    gf = np.empty((B, R, S, F), np.float32)
    gs = np.empty((B, R, S), np.float32)
    lib().orc_render_train_bwd(_f(g_feats), _f(g_depth), _f(g_ws), _f(g_alpha), _f(o_feats),
                               _f(o_depth), _f(o_ws), _f(o_alpha), t, sidx, sigma, feats,
                               B * R, S, F, t_thres, int(occ_mode), gf, gs)
    return gf, gs


def freq_fwd(x, deg):
    x = _f(x)
    N, D = x.shape
    y = np.empty((N, D + 2 * D * deg), np.float32)
    lib().orc_freq_fwd(x, N, D, deg, y)
    return y


def freq_bwd(gy, y, D, deg):
    gy, y = _f(gy), _f(y)
    N = gy.shape[0]
    gx = np.empty((N, D), np.float32)
    lib().orc_freq_bwd(gy, y, N, D, deg, gx)
    return gx


# --------------------------------------------------------------------------
# float64 accuracy oracles
def integrate_f64(K, RT, colors, depths, bbox, vs, dims, obs_weight=1.0, trunc_weight=3.0,
                  voxels=None):
    """float64 restatement of integrate_cuda's TSDF/weight (colour omitted);
    returns (tsdf, weight) in float64 for error bounds.  With `voxels` (flat
    raster indices of one batch's grid) only those voxels are evaluated and
    the arrays are [B, len(voxels)]."""
    K = np.asarray(K, np.float64)
    RT = np.asarray(RT, np.float64)
    depths = np.asarray(depths, np.float64)
    B, V, H, W = depths.shape
    X, Y, Z = dims
    shape = (B, X, Y, Z) if voxels is None else (B, len(voxels))
    tsdf = np.ones(shape)
    weight = np.zeros(shape)
    for b in range(B):
        trunc = float(vs[b]) * trunc_weight
        if voxels is None:
            gx, gy, gz = np.meshgrid(np.arange(X), np.arange(Y), np.arange(Z), indexing="ij")
        else:
            v = np.asarray(voxels, np.int64)
            gx, gy, gz = v // (Y * Z), (v // Z) % Y, v % Z
        pts = np.stack([bbox[b, 0, 0] + gx * float(vs[b]), bbox[b, 1, 0] + gy * float(vs[b]),
                        bbox[b, 2, 0] + gz * float(vs[b])], -1).reshape(-1, 3)
        ts = tsdf[b].reshape(-1)
        ws = weight[b].reshape(-1)
        for k in range(V):
            cam = pts @ RT[b, k, :, :3].T + RT[b, k, :, 3]
            pix = cam @ K[b, k].T
            with np.errstate(divide="ignore", invalid="ignore"):
                px = pix[:, 0] / pix[:, 2]
                py = pix[:, 1] / pix[:, 2]
            ix = np.floor(np.abs(px) + 0.5) * np.sign(px)
            iy = np.floor(np.abs(py) + 0.5) * np.sign(py)
            ok = np.isfinite(ix) & np.isfinite(iy) & (ix >= 0) & (ix < W) & (iy >= 0) & \
                (iy < H) & (cam[:, 2] >= 0)
            d = np.zeros(len(pts))
            d[ok] = depths[b, k, iy[ok].astype(int), ix[ok].astype(int)]
            ok &= (d >= 0.1) & (d <= 4.0)
            diff = d - cam[:, 2]
            ok &= diff >= -trunc
            dist = np.minimum(1.0, diff / trunc)
            w_old = ws.copy()
            ws = np.where(ok, ws + obs_weight, ws)
            with np.errstate(divide="ignore", invalid="ignore"):
                ts = np.where(ok, (ts * w_old + obs_weight * dist) / ws, ts)
        tsdf[b] = ts.reshape(shape[1:])
        weight[b] = ws.reshape(shape[1:])
    return tsdf, weight


def render_occ_f64(t, occ, feats, t_thres=1e-4):
    """float64 sequential restatement of volume_rendering_occ_kernel."""
    t = np.asarray(t, np.float64)
    occ = np.asarray(occ, np.float64)
    feats = np.asarray(feats, np.float64)
    B, R, S = t.shape
    F = feats.shape[-1]
    of = np.zeros((B, R, F))
    od, ow, oa = np.zeros((B, R)), np.zeros((B, R)), np.zeros((B, R))
    for b in range(B):
        for r in range(R):
            if t[b, r, -1] < 0:
                continue
            T = 1.0
            for i in range(S):
                if t[b, r, i] < 0:
                    continue
                a = occ[b, r, i]
                w = a * T
                of[b, r] += w * feats[b, r, i]
                od[b, r] += w * t[b, r, i]
                ow[b, r] += w
                oa[b, r] += w * a
                T *= 1.0 - a
                if T < t_thres:
                    break
    return of, od, ow, oa


def freq_f64(x, deg):
    x = np.asarray(x, np.float64)
    outs = [x]
    for f in range(deg):
        outs.append(np.sin(x * 2.0 ** f))
        outs.append(np.cos(x * 2.0 ** f))
    return np.concatenate(outs, -1)
