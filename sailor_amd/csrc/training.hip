// Training-only kernels of the octree path: the recursive octree traversal,
// the adaptive (length- and size-weighted) point sampling and the trilinear
// aggregation of per-corner features with its backward.
//
// Follows c_lib/VoxelEncoding/src/intersection_kernel.cu (octree_traversal
// :738-825, ray_octree_traversal :441-643, RayVoxelIntersectionNodeList
// :72-126) and c_lib/VoxelEncoding/src/sampling_kernel.cu
// (ray_voxels_points_sampling :1178-1248 with ray_voxels_sampling_kernel
// :911-982 and ray_voxels_further_sampling_kernel :984-1068,
// trilinear_aggregate_features[_backward] :162-327, 1384-1462).
//
// The reference appends per-ray entries and samples through atomicAdd
// queues from one thread per (ray, coarse hit) and accumulates the ray's
// valid length with float atomics, so its entry order and the low bits of
// the valid length depend on scheduling.  Here one thread owns a ray and
// appends in coarse-hit order, which is one of the orders the reference can
// produce, and the result is reproducible bit for bit.  Lists are bounded by
// their capacity with an overflow flag (the reference writes past the row;
// its DFS buffers hold 8 entries).  Compiled with -ffp-contract=off.
#include "common.h"
#include "traversal_common.h"

namespace {

constexpr int kThreads = 256;

__global__ void fill_lists(int* __restrict__ idx, float* __restrict__ a, float* __restrict__ b,
                           long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    idx[i] = -1;
    a[i] = -1.0f;
    b[i] = -1.0f;
  }
}

// Receives the coarse hits of the DDA (RayVoxelIntersectionNodeList) and
// expands each into its fine entries right away: the coarse entry, then the
// children in stack order (last child first) that the slab test hits; an
// interval equal to one already listed for this coarse hit keeps the
// smaller node (ray_octree_traversal :518-567).
struct TreeSink {
  const float4* np;
  const int4* ni;
  int n_nodes;
  float3 o, inv;
  int* idx_c;
  float* t0_c;
  float* t1_c;
  int cap_c;
  int* idx;
  float* t0;
  float* t1;
  int cap;
  int cnt;
  float vl;
  bool overflow;

  __device__ __forceinline__ void put(int pos, int node, float a, float b) {
    if (pos < cap_c) {
      idx_c[pos] = node;
      t0_c[pos] = a;
      t1_c[pos] = b;
    } else {
      overflow = true;
    }
    const float start_vs = np[node].w;
    const int g0 = cnt;
    if (cnt < cap) {
      idx[cnt] = node;
      t0[cnt] = a;
      t1[cnt] = b;
    } else {
      overflow = true;
    }
    cnt += 1;
    const int4 info = ni[node];
    if (info.z >= 0) {
      const int last = min(info.z + info.w, n_nodes) - 1;
      for (int ch = last; ch >= info.z; --ch) {
        const float4 q = np[ch];
        const float2 cd = sailor::slab_test(o, inv, make_float3(q.x, q.y, q.z), q.w);
        if (cd.x == -1.0f || cd.y == -1.0f) continue;
        bool found = false;
        const int lim = cnt < cap ? cnt : cap;
        for (int j = g0; j < lim; ++j) {
          if (t0[j] == cd.x && t1[j] == cd.y) {
            if (np[idx[j]].w > q.w) idx[j] = ch;
            found = true;
            break;
          }
        }
        if (!found) {
          if (cnt < cap) {
            idx[cnt] = ch;
            t0[cnt] = cd.x;
            t1[cnt] = cd.y;
          } else {
            overflow = true;
          }
          cnt += 1;
        }
      }
    }
    const int lim = cnt < cap ? cnt : cap;
    for (int j = g0; j < lim; ++j) {
      const float scale = start_vs / np[idx[j]].w;
      vl += (t1[j] - t0[j]) * scale;
    }
  }
  __device__ __forceinline__ bool stop() const { return false; }
};

__global__ void __launch_bounds__(kThreads)
octree_traversal_kernel(const int* __restrict__ vc, const float4* __restrict__ node_pos,
                        const int4* __restrict__ node_info, int n_nodes,
                        const float* __restrict__ vol_bbox, const float* __restrict__ voxel_size,
                        int B, int dim_fine, int dim_coarse, int rate,
                        const float* __restrict__ ray_ori, const float* __restrict__ ray_dir,
                        int R, int n_max_c, int n_max, int th_early_stop, int opt_num,
                        int* __restrict__ idx_c, float* __restrict__ t0_c,
                        float* __restrict__ t1_c, int* __restrict__ idx,
                        float* __restrict__ t0, float* __restrict__ t1,
                        float* __restrict__ valid_len, int* __restrict__ status) {
  const long long ray = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  int n_coarse = 0, n_fine = 0;
  bool overflow = false;
  if (ray < (long long)B * R) {
    const int b = (int)(ray / R);
    sailor::RayCtx ctx;
    ctx.init(vc, node_pos, node_info, n_nodes, vol_bbox, voxel_size, b, dim_fine, dim_coarse,
             rate, ray_ori + ray * 3, ray_dir + ray * 3);
    TreeSink sink{node_pos, node_info, n_nodes, ctx.o, ctx.inv,
                  idx_c + ray * n_max_c, t0_c + ray * n_max_c, t1_c + ray * n_max_c, n_max_c,
                  idx + ray * n_max, t0 + ray * n_max, t1 + ray * n_max, n_max,
                  0, 0.0f, false};
    float vl_coarse = 0.0f;  // the coarse pass adds no valid length (:90-126)
    n_coarse = sailor::traverse(ctx, th_early_stop, true, opt_num, sink, &vl_coarse);
    valid_len[ray] = sink.vl;
    n_fine = sink.cnt;
    overflow = sink.overflow;
  }
  int mc = n_coarse, mf = n_fine;
  for (int o = 32; o > 0; o >>= 1) {
    mc = max(mc, __shfl_xor(mc, o, 64));
    mf = max(mf, __shfl_xor(mf, o, 64));
  }
  const unsigned long long ov = __ballot(overflow);
  if ((threadIdx.x & 63) == 0) {
    if (mc > 0) atomicMax(status, mc);
    if (mf > 0) atomicMax(status + 1, mf);
    if (ov) atomicOr(status + 2, 1);
  }
}

// One thread per ray: ray_voxels_sampling_kernel over the hits in order,
// then ray_voxels_further_sampling_kernel.
__global__ void __launch_bounds__(kThreads)
adaptive_sampling_kernel(const int* __restrict__ hit_idx, const float* __restrict__ tmin,
                         const float* __restrict__ tmax, const float* __restrict__ noise,
                         const float* __restrict__ valid_len,
                         const float4* __restrict__ node_pos, int n_nodes,
                         const float* __restrict__ voxel_size, int B, int R, int n_max,
                         int max_hits, int S, int start_scale, int* __restrict__ sidx,
                         float* __restrict__ st) {
  const long long ray = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (ray >= (long long)B * R) return;
  const int b = (int)(ray / R);
  const int* hi = hit_idx + ray * n_max;
  const float* mn_ = tmin + ray * n_max;
  const float* mx_ = tmax + ray * n_max;
  const float* rnd = noise + ray * S;
  int* oi = sidx + ray * S;
  float* ot = st + ray * S;
  for (int j = 0; j < S; ++j) {
    oi[j] = -1;
    ot[j] = -1.0f;
  }
  const float vlen = valid_len[ray];
  const float start_vs = voxel_size[b] * (float)start_scale;
  int cnt = 0;
  // ray_voxels_sampling_kernel visits the first max_hits slots of every ray
  // (num_max_intersected_voxels_fine, :944)
  for (int i = 0; i < max_hits; ++i) {
    const int vi = hi[i];
    if (vi == -1 || vi < 0 || vi >= n_nodes) continue;
    const float mn = mn_[i], mx = mx_[i];
    const float rate = start_vs / node_pos[vi].w;
    int n = sailor::sat_f2i((mx - mn) * rate * (float)S / vlen);
    n = n >= 1 ? n : 1;
    const float off = (mx - mn) / (float)(n + 1);
    for (int j = 0; j < n; ++j) {
      if (cnt >= S) break;
      const int rid = cnt++;
      const float u = rnd[rid];
      const float basic = mn + off * (float)(j + 1);
      ot[rid] = (float)((double)basic - (double)off / 2.0 + (double)(off * u));
      oi[rid] = vi;
    }
  }
  float min_mid = 99999.0f, min_vs = 99999.0f, z0 = -1.0f, z1 = -1.0f;
  int min_idx = -1;
  for (int i = 0; i < n_max; ++i) {
    const int vi = hi[i];
    if (vi == -1) break;
    if (vi < 0 || vi >= n_nodes) continue;
    const float vs = node_pos[vi].w;
    if (vs <= min_vs) {
      const float a = mn_[i], c = mx_[i];
      const float mid = (float)((double)(a + c) / 2.0);
      if (vs == min_vs) {
        if (mid < min_mid) {
          min_mid = mid;
          z0 = a;
          z1 = c;
          min_idx = vi;
        }
      } else {
        min_mid = mid;
        z0 = a;
        z1 = c;
        min_idx = vi;
      }
      min_vs = vs;
    }
  }
  if (min_idx < 0) return;
  const int n_new = S - cnt;
  if (n_new <= 0) return;
  const float off = (z1 - z0) / (float)(n_new + 1);
  int j = 0;
  for (int i = cnt; i < S; ++i) {
    const float basic = z0 + off * (float)(j + 1);
    const float u = rnd[i];
    ot[i] = (float)((double)basic - (double)off / 2.0 + (double)(off * u));
    oi[i] = min_idx;
    ++j;
  }
}

// Interpolation weights of the point ori + dir * t in node p (corners at
// centre +- size / 2; corner id 4 x + 2 y + z): tx, and ty / tz from the
// mean corner coordinates in double (sampling_kernel.cu:216-227).
__device__ __forceinline__ void tri_weights(float4 p, const float* __restrict__ o,
                                            const float* __restrict__ d, float t, float& tx,
                                            float& ty, float& tz) {
  const float px = o[0] + d[0] * t, py = o[1] + d[1] * t, pz = o[2] + d[2] * t;
  const float h_lo = ((float)(-1) * p.w) / 2.0f, h_hi = ((float)(1) * p.w) / 2.0f;
  const float xl = p.x + h_lo, xh = p.x + h_hi;
  const float yl = p.y + h_lo, yh = p.y + h_hi;
  const float zl = p.z + h_lo, zh = p.z + h_hi;
  tx = (px - xl) / (xh - xl);
  const double ym = (double)(yl + yl) / 2.0, yp = (double)(yh + yh) / 2.0;
  ty = (float)(((double)py - ym) / (yp - ym));
  const double zm = (double)(((zl + zl) + zl) + zl) / 4.0;
  const double zp = (double)(((zh + zh) + zh) + zh) / 4.0;
  tz = (float)(((double)pz - zm) / (zp - zm));
}

__device__ __forceinline__ float lerp1(float a, float b, float t) { return a * (1 - t) + b * t; }

// One thread per sample (b, r, s), looping over the C channels: the 8
// corner values of a channel are one 32-byte read, consecutive samples are
// consecutive in every channel plane.
__global__ void __launch_bounds__(kThreads)
trilinear_fwd_kernel(const float4* __restrict__ node_pos, int n_nodes,
                     const float* __restrict__ ray_ori, const float* __restrict__ ray_dir,
                     const int* __restrict__ sidx, const float* __restrict__ st,
                     const float4* __restrict__ feats, int B, int R, int S, int C,
                     float* __restrict__ out) {
  const long long RS = (long long)R * S;
  const long long total = (long long)B * RS;
  for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < total;
       k += (long long)gridDim.x * blockDim.x) {
    const long long b = k / RS, rs = k - b * RS, r = rs / S;
    const int vi = sidx[k];
    const bool ok = vi >= 0 && vi < n_nodes;
    float tx = 0.0f, ty = 0.0f, tz = 0.0f;
    if (ok) tri_weights(node_pos[vi], ray_ori + 3 * (b * R + r), ray_dir + 3 * (b * R + r), st[k],
                        tx, ty, tz);
    for (int c = 0; c < C; ++c) {
      const long long oi = (b * C + c) * RS + rs;
      float v = 0.0f;
      if (ok) {
        const float4 lo = feats[2 * oi], hi = feats[2 * oi + 1];
        const float x04 = lerp1(lo.x, hi.x, tx), x15 = lerp1(lo.y, hi.y, tx);
        const float x26 = lerp1(lo.z, hi.z, tx), x37 = lerp1(lo.w, hi.w, tx);
        const float y0 = lerp1(x04, x26, ty), y1 = lerp1(x15, x37, ty);
        v = lerp1(y0, y1, tz);
      }
      out[oi] = v;
    }
  }
}

__global__ void __launch_bounds__(kThreads)
trilinear_bwd_kernel(const float4* __restrict__ node_pos, int n_nodes,
                     const float* __restrict__ ray_ori, const float* __restrict__ ray_dir,
                     const int* __restrict__ sidx, const float* __restrict__ st,
                     const float* __restrict__ grad, int B, int R, int S, int C,
                     float4* __restrict__ grad_feats) {
  const long long RS = (long long)R * S;
  const long long total = (long long)B * RS;
  for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < total;
       k += (long long)gridDim.x * blockDim.x) {
    const long long b = k / RS, rs = k - b * RS, r = rs / S;
    const int vi = sidx[k];
    const bool ok = vi >= 0 && vi < n_nodes;
    float w[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    if (ok) {
      float tx, ty, tz;
      tri_weights(node_pos[vi], ray_ori + 3 * (b * R + r), ray_dir + 3 * (b * R + r), st[k], tx,
                  ty, tz);
      // sampling_kernel.cu:293-316
      const float z1 = tz, z0 = 1 - tz, y26 = ty, y04 = 1 - ty, y37 = ty, y15 = 1 - ty;
      const float x4 = tx, x0 = 1 - tx, x6 = tx, x2 = 1 - tx, x5 = tx, x1 = 1 - tx, x7 = tx,
                  x3 = 1 - tx;
      w[0] = x0 * y04 * z0;
      w[1] = x1 * y15 * z1;
      w[2] = x2 * y26 * z0;
      w[3] = x3 * y37 * z1;
      w[4] = x4 * y04 * z0;
      w[5] = x5 * y15 * z1;
      w[6] = x6 * y26 * z0;
      w[7] = x7 * y37 * z1;
    }
    for (int c = 0; c < C; ++c) {
      const long long gi = (b * C + c) * RS + rs;
      float4 lo = make_float4(0.f, 0.f, 0.f, 0.f), hi = lo;
      if (ok) {
        const float g = grad[gi];
        lo = make_float4(g * w[0], g * w[1], g * w[2], g * w[3]);
        hi = make_float4(g * w[4], g * w[5], g * w[6], g * w[7]);
      }
      grad_feats[2 * gi] = lo;
      grad_feats[2 * gi + 1] = hi;
    }
  }
}

// project_xyz (sampling_kernel.cu:55-157): one thread per (batch, view,
// ray, sample); the point ori + dir * t (and with corners, the node's 8
// corners, id 4 x + 2 y + z) through [R | T] and K, divided by the third
// coordinate and mapped to [-1, 1] as (p - w / 2.0) / (w / 2.0) in double.
// Samples with index -1 keep the -1 fill.
__global__ void __launch_bounds__(kThreads)
project_kernel(const float4* __restrict__ node_pos, int n_nodes,
               const float* __restrict__ ray_ori, const float* __restrict__ ray_dir,
               const float* __restrict__ Ks, const float* __restrict__ RTs,
               const int* __restrict__ sidx, const float* __restrict__ st, int B, int V, int R,
               int S, int ori_w, int ori_h, int corners, float* __restrict__ xy,
               float* __restrict__ z) {
  const long long total = (long long)B * V * R * S;
  const int np_ = corners ? 9 : 1;
  for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < total;
       k += (long long)gridDim.x * blockDim.x) {
    const long long rs = k % ((long long)R * S);
    const long long bv = k / ((long long)R * S);
    const long long b = bv / V;
    const long long r = rs / S;
    const long long si = b * R * S + rs;
    const int vi = sidx[si];
    if (vi == -1) continue;
    const float* RT = RTs + bv * 12;
    const float* K = Ks + bv * 9;
    const float* o = ray_ori + (b * R + r) * 3;
    const float* d = ray_dir + (b * R + r) * 3;
    const float t = st[si];
    const float pt[3] = {o[0] + d[0] * t, o[1] + d[1] * t, o[2] + d[2] * t};
    const float4 nd = (vi >= 0 && vi < n_nodes) ? node_pos[vi] : make_float4(0.f, 0.f, 0.f, 0.f);
    const bool corner_ok = corners && vi >= 0 && vi < n_nodes;
    const double hw = (double)ori_w / 2.0, hh = (double)ori_h / 2.0;
    for (int c = 0; c < np_; ++c) {
      float q[3];
      if (c == 0) {
        q[0] = pt[0];
        q[1] = pt[1];
        q[2] = pt[2];
      } else {
        if (!corner_ok) break;
        const int id = c - 1;
        q[0] = nd.x + ((float)((id >> 2) & 1 ? 1 : -1) * nd.w) / 2.0f;
        q[1] = nd.y + ((float)((id >> 1) & 1 ? 1 : -1) * nd.w) / 2.0f;
        q[2] = nd.z + ((float)(id & 1 ? 1 : -1) * nd.w) / 2.0f;
      }
      const float cx = RT[0] * q[0] + RT[1] * q[1] + RT[2] * q[2] + RT[3];
      const float cy = RT[4] * q[0] + RT[5] * q[1] + RT[6] * q[2] + RT[7];
      const float cz = RT[8] * q[0] + RT[9] * q[1] + RT[10] * q[2] + RT[11];
      float px = K[0] * cx + K[1] * cy + K[2] * cz;
      float py = K[3] * cx + K[4] * cy + K[5] * cz;
      const float pz = K[6] * cx + K[7] * cy + K[8] * cz;
      px = px / pz;
      py = py / pz;
      const long long o2 = (k * np_ + c) * 2;
      xy[o2] = (float)(((double)px - hw) / hw);
      xy[o2 + 1] = (float)(((double)py - hh) / hh);
      z[k * np_ + c] = cz;
    }
  }
}

__global__ void fill_value(float* __restrict__ a, long long n, float v) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    a[i] = v;
}

unsigned grid_of(long long work) {
  return (unsigned)std::max<long long>(
      1, std::min<long long>((work + kThreads - 1) / kThreads, 256 * 32));
}

}  // namespace

extern "C" int sailor_octree_traversal(
    const int32_t* vol_coarse, const float* node_pos, const int32_t* node_info, int n_nodes,
    const float* vol_bbox, const float* voxel_size, int B, int dim_fine, int dim_coarse,
    int rate, const float* ray_ori, const float* ray_dir, int R, int n_max_hits_coarse,
    int n_max_hits, int th_early_stop, int opt_num_intersected, int32_t* hit_idx_coarse,
    float* tmin_coarse, float* tmax_coarse, int32_t* hit_idx, float* tmin, float* tmax,
    float* valid_len, int32_t* status, void* stream) {
  SAILOR_REQUIRE(B > 0 && R >= 0 && n_max_hits_coarse > 0 && n_max_hits > 0,
                 "octree_traversal: bad sizes");
  SAILOR_REQUIRE(dim_coarse > 0 && rate >= 1 && n_nodes >= 0, "octree_traversal: bad tree dims");
  if (R == 0) return SAILOR_OK;
  hipStream_t st = sailor::as_stream(stream);
  const long long nc = (long long)B * R * n_max_hits_coarse, nf = (long long)B * R * n_max_hits;
  hipLaunchKernelGGL(fill_lists, dim3(grid_of(nc)), dim3(kThreads), 0, st, hit_idx_coarse,
                     tmin_coarse, tmax_coarse, nc);
  hipLaunchKernelGGL(fill_lists, dim3(grid_of(nf)), dim3(kThreads), 0, st, hit_idx, tmin, tmax,
                     nf);
  if (int e = sailor::check_launch("octree_traversal (fill)")) return e;
  hipLaunchKernelGGL(octree_traversal_kernel, dim3(sailor::ceil_div((long long)B * R, kThreads)),
                     dim3(kThreads), 0, st, vol_coarse,
                     reinterpret_cast<const float4*>(node_pos),
                     reinterpret_cast<const int4*>(node_info), n_nodes, vol_bbox, voxel_size, B,
                     dim_fine, dim_coarse, rate, ray_ori, ray_dir, R, n_max_hits_coarse,
                     n_max_hits, th_early_stop, opt_num_intersected, hit_idx_coarse,
                     tmin_coarse, tmax_coarse, hit_idx, tmin, tmax, valid_len, status);
  return sailor::check_launch("octree_traversal");
}

extern "C" int sailor_octree_sampling(const int32_t* hit_idx, const float* tmin,
                                      const float* tmax, const float* noise,
                                      const float* valid_len, const float* node_pos,
                                      int n_nodes, const float* voxel_size, int B, int R,
                                      int n_max_hits, int max_hits, int S, int start_scale,
                                      int32_t* sampled_idx, float* sampled_t, void* stream) {
  SAILOR_REQUIRE(B > 0 && R >= 0 && S > 0 && n_max_hits > 0 && n_nodes >= 0,
                 "ray_voxels_points_sampling: bad sizes");
  SAILOR_REQUIRE(max_hits >= 0 && max_hits <= n_max_hits,
                 "ray_voxels_points_sampling: max_hits %d outside [0, %d]", max_hits, n_max_hits);
  if (R == 0) return SAILOR_OK;
  hipLaunchKernelGGL(adaptive_sampling_kernel, dim3(sailor::ceil_div((long long)B * R, kThreads)),
                     dim3(kThreads), 0, sailor::as_stream(stream), hit_idx, tmin, tmax, noise,
                     valid_len, reinterpret_cast<const float4*>(node_pos), n_nodes, voxel_size,
                     B, R, n_max_hits, max_hits, S, start_scale, sampled_idx, sampled_t);
  return sailor::check_launch("ray_voxels_points_sampling");
}

extern "C" int sailor_trilinear_aggregate(const float* node_pos, int n_nodes,
                                          const float* ray_ori, const float* ray_dir,
                                          const int32_t* sampled_idx, const float* sampled_t,
                                          const float* feats, int B, int R, int S, int C,
                                          float* out, void* stream) {
  SAILOR_REQUIRE(B > 0 && R >= 0 && S > 0 && C > 0 && n_nodes >= 0,
                 "trilinear_aggregate_features: bad sizes");
  SAILOR_REQUIRE(((uintptr_t)feats & 15) == 0, "trilinear_aggregate_features: feats must be "
                 "16-byte aligned");
  if (R == 0) return SAILOR_OK;
  hipLaunchKernelGGL(trilinear_fwd_kernel, dim3(grid_of((long long)B * R * S)), dim3(kThreads),
                     0, sailor::as_stream(stream), reinterpret_cast<const float4*>(node_pos),
                     n_nodes, ray_ori, ray_dir, sampled_idx, sampled_t,
                     reinterpret_cast<const float4*>(feats), B, R, S, C, out);
  return sailor::check_launch("trilinear_aggregate_features");
}

extern "C" int sailor_trilinear_aggregate_backward(const float* node_pos, int n_nodes,
                                                   const float* ray_ori, const float* ray_dir,
                                                   const int32_t* sampled_idx,
                                                   const float* sampled_t,
                                                   const float* grad_out, int B, int R, int S,
                                                   int C, float* grad_feats, void* stream) {
  SAILOR_REQUIRE(B > 0 && R >= 0 && S > 0 && C > 0 && n_nodes >= 0,
                 "trilinear_aggregate_features_backward: bad sizes");
  SAILOR_REQUIRE(((uintptr_t)grad_feats & 15) == 0, "trilinear_aggregate_features_backward: "
                 "grad_feats must be 16-byte aligned");
  if (R == 0) return SAILOR_OK;
  hipLaunchKernelGGL(trilinear_bwd_kernel, dim3(grid_of((long long)B * R * S)), dim3(kThreads),
                     0, sailor::as_stream(stream), reinterpret_cast<const float4*>(node_pos),
                     n_nodes, ray_ori, ray_dir, sampled_idx, sampled_t, grad_out, B, R, S, C,
                     reinterpret_cast<float4*>(grad_feats));
  return sailor::check_launch("trilinear_aggregate_features_backward");
}

extern "C" int sailor_project_sampled_xyz(const float* node_pos, int n_nodes, const float* ray_ori,
                                          const float* ray_dir, const int32_t* sampled_idx,
                                          const float* sampled_t, const float* Ks,
                                          const float* RTs, int B, int V, int R, int S,
                                          int ori_w, int ori_h, int record_corners,
                                          float* proj_xy, float* proj_z, void* stream) {
  SAILOR_REQUIRE(B > 0 && V > 0 && R >= 0 && S > 0 && n_nodes >= 0 && ori_w > 0 && ori_h > 0,
                 "project_sampled_xyz: bad sizes");
  if (R == 0) return SAILOR_OK;
  hipStream_t st = sailor::as_stream(stream);
  const long long n = (long long)B * V * R * S * (record_corners ? 9 : 1);
  hipLaunchKernelGGL(fill_value, dim3(grid_of(2 * n)), dim3(kThreads), 0, st, proj_xy, 2 * n,
                     -1.0f);
  hipLaunchKernelGGL(fill_value, dim3(grid_of(n)), dim3(kThreads), 0, st, proj_z, n, -1.0f);
  if (int e = sailor::check_launch("project_sampled_xyz (fill)")) return e;
  hipLaunchKernelGGL(project_kernel, dim3(grid_of((long long)B * V * R * S)), dim3(kThreads), 0,
                     st, reinterpret_cast<const float4*>(node_pos), n_nodes, ray_ori, ray_dir,
                     Ks, RTs, sampled_idx, sampled_t, B, V, R, S, ori_w, ori_h,
                     record_corners ? 1 : 0, proj_xy, proj_z);
  return sailor::check_launch("project_sampled_xyz");
}
