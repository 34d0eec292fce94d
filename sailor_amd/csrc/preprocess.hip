// Preprocessing either side of the frame build: per-view depth
// normalisation (c_lib/VoxelEncoding/src/depth_normalizer.cu:98-189), colour
// normalisation (utils_render/utils_render.py:601-605) and the training-ray
// selectors rays_selecting / rgbdv_selecting (sampling_kernel.cu:489-607,
// 736-825).
//
// Depth normalisation: the reference spins one float CAS loop per valid
// pixel on the two per-view words; here each workgroup reduces its slice of
// one view in registers / LDS and issues one integer atomic per workgroup on
// order-preserving keys (min/max are exact, so the result does not depend on
// the reduction order), then one element-wise kernel writes the normalised
// depths and the mask with the reference's operation order.
#include "common.h"

namespace {

constexpr int kBlock = 256;
constexpr int kSlicesPerView = 64;  // workgroups per (batch, view) in the reduction

__global__ void minmax_init(int* keys, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    keys[2 * i] = sailor::fkey(99999.0f);
    keys[2 * i + 1] = sailor::fkey(-99999.0f);
  }
}

// keys[2*bv] = min, keys[2*bv+1] = max over the valid depths of view bv.
__global__ void __launch_bounds__(kBlock)
minmax_kernel(const float* __restrict__ depths, long long hw, float dmin, float dmax,
              int* __restrict__ keys) {
  const int bv = blockIdx.y;
  const float* d = depths + (long long)bv * hw;
  float lo = 99999.0f, hi = -99999.0f;
  for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < hw;
       i += (long long)gridDim.x * kBlock) {
    const float v = d[i];
    if (v < dmin || v > dmax) continue;
    lo = fminf(lo, v);
    hi = fmaxf(hi, v);
  }
  lo = sailor::wave_min_f(lo);
  hi = sailor::wave_max_f(hi);
  __shared__ float s_lo[kBlock / 64], s_hi[kBlock / 64];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0) {
    s_lo[wave] = lo;
    s_hi[wave] = hi;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kBlock / 64; ++w) {
      lo = fminf(lo, s_lo[w]);
      hi = fmaxf(hi, s_hi[w]);
    }
    atomicMin(keys + 2 * bv, sailor::fkey(lo));
    atomicMax(keys + 2 * bv + 1, sailor::fkey(hi));
  }
}

// mid = (max + min) / 2, distance = max - min + 1e-10 (float arithmetic of
// depth_normalizer.cu:155-156).
__global__ void mid_dist_kernel(const int* __restrict__ keys, int n, float* __restrict__ mid,
                                float* __restrict__ dist) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float lo = sailor::fkey_inv(keys[2 * i]), hi = sailor::fkey_inv(keys[2 * i + 1]);
  dist[i] = (hi - lo) + 1e-10f;
  mid[i] = (hi + lo) / 2.0f;
}

// normalised depth and mask (depth_normalizer.cu:159-188): mode 0 = computed
// distance, 1 = user distance.
__global__ void __launch_bounds__(kBlock)
normalize_kernel(const float* __restrict__ depths, long long hw, int n_views_total,
                 const float* __restrict__ mid, const float* __restrict__ dist, float dmin,
                 float dmax, float user_dist, int multi_mask, int divided2,
                 float* __restrict__ out, float* __restrict__ mask) {
  const long long total = hw * n_views_total;
  for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < total;
       i += (long long)gridDim.x * kBlock) {
    const int bv = (int)(i / hw);
    const float d = depths[i];
    const float m = (multi_mask && (d < dmin || d > dmax)) ? 0.0f : 1.0f;
    mask[i] = m;
    float r;
    if (user_dist < 0.0f) {
      r = multi_mask ? ((d - mid[bv]) * m) / dist[bv] : (d - mid[bv]) / dist[bv];
      if (divided2) r = r * 2.0f;
      r = fminf(fmaxf(r, -1.0f), 1.0f);
    } else if (multi_mask) {
      r = ((d - mid[bv]) * m) / user_dist;
    } else {
      r = fminf(fmaxf((d - mid[bv]) / user_dist, -1.0f), 1.0f);
    }
    out[i] = r;
  }
}

// One thread per selected ray: rows of `C` floats of [.., H, W, C] maps at
// (x, y) = sampled_xy; an index outside the image selects zeros.
__global__ void __launch_bounds__(kBlock)
select_kernel(const float* __restrict__ src, int C, int H, int W, const int* __restrict__ xy,
              long long n_views_total, int n_rays, float* __restrict__ dst,
              float* __restrict__ label) {
  const long long total = n_views_total * n_rays;
  for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < total;
       i += (long long)gridDim.x * kBlock) {
    const long long bv = i / n_rays;
    const int x = xy[2 * i], y = xy[2 * i + 1];
    const bool in = x >= 0 && x < W && y >= 0 && y < H;
    const float* s = src + ((bv * H + (in ? y : 0)) * W + (in ? x : 0)) * C;
    for (int c = 0; c < C; ++c) dst[i * C + c] = in ? s[c] : 0.0f;
    if (label) label[i] = (in && s[0] > 0.0f) ? 1.0f : 0.0f;
  }
}

// rgb_normalizer: ((rgb - mid) / dist) * mask, the mask [N,1,H,W]
// broadcast over the C channels of rgb [N,C,H,W].
__global__ void __launch_bounds__(kBlock)
rgb_normalize_kernel(const float* __restrict__ rgb, const float* __restrict__ mask, int C,
                     long long hw, long long total, float mid, float dist,
                     float* __restrict__ out) {
  for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < total;
       i += (long long)gridDim.x * kBlock) {
    const long long n = i / (C * hw), p = i % hw;
    out[i] = ((rgb[i] - mid) / dist) * mask[n * hw + p];
  }
}

unsigned grid_of(long long work) {
  return (unsigned)std::max<long long>(1, std::min<long long>((work + kBlock - 1) / kBlock, 4096));
}

}  // namespace

extern "C" int sailor_depth_normalize(const float* depths, int B, int V, int H, int W,
                                      float valid_min, float valid_max, float user_dist,
                                      int multi_mask, int divided2, float* out_normalized,
                                      float* out_mask, float* out_mid, float* out_dist,
                                      void* workspace, size_t workspace_bytes, void* stream) {
  SAILOR_REQUIRE(B > 0 && V > 0 && H > 0 && W > 0, "depth_normalize: bad sizes");
  const int n = B * V;
  SAILOR_REQUIRE(workspace != nullptr && workspace_bytes >= (size_t)n * 2 * sizeof(int),
                 "depth_normalize: workspace needs %zu bytes", (size_t)n * 2 * sizeof(int));
  hipStream_t st = sailor::as_stream(stream);
  int* keys = static_cast<int*>(workspace);
  const long long hw = (long long)H * W;
  hipLaunchKernelGGL(minmax_init, dim3((n + 255) / 256), dim3(256), 0, st, keys, n);
  if (int e = sailor::check_launch("depth_normalize (init)")) return e;
  const unsigned slices = (unsigned)std::min<long long>(kSlicesPerView, (hw + kBlock - 1) / kBlock);
  hipLaunchKernelGGL(minmax_kernel, dim3(slices, n), dim3(kBlock), 0, st, depths, hw, valid_min,
                     valid_max, keys);
  if (int e = sailor::check_launch("depth_normalize (min/max)")) return e;
  hipLaunchKernelGGL(mid_dist_kernel, dim3((n + 255) / 256), dim3(256), 0, st, keys, n, out_mid,
                     out_dist);
  if (int e = sailor::check_launch("depth_normalize (mid)")) return e;
  hipLaunchKernelGGL(normalize_kernel, dim3(grid_of(hw * n)), dim3(kBlock), 0, st, depths, hw, n,
                     out_mid, out_dist, valid_min, valid_max, user_dist, multi_mask, divided2,
                     out_normalized, out_mask);
  return sailor::check_launch("depth_normalize");
}

extern "C" int sailor_rays_selecting(const float* ray_ori_dir, int B, int V, int H, int W,
                                     const int* sampled_xy, int n_rays, float* out,
                                     void* stream) {
  SAILOR_REQUIRE(B > 0 && V > 0 && H > 0 && W > 0 && n_rays >= 0, "rays_selecting: bad sizes");
  if (n_rays == 0) return SAILOR_OK;
  const long long nv = (long long)B * V;
  hipLaunchKernelGGL(select_kernel, dim3(grid_of(nv * n_rays)), dim3(kBlock), 0,
                     sailor::as_stream(stream), ray_ori_dir, 6, H, W, sampled_xy, nv, n_rays,
                     out, nullptr);
  return sailor::check_launch("rays_selecting");
}

extern "C" int sailor_rgbdv_selecting(const float* rgbs, int rgb_channels, const float* normals,
                                      const float* depths, int B, int V, int H, int W,
                                      const int* sampled_xy, int n_rays, float* out_rgb,
                                      float* out_normals, float* out_depth, float* out_label,
                                      void* stream) {
  SAILOR_REQUIRE(B > 0 && V > 0 && H > 0 && W > 0 && n_rays >= 0 && rgb_channels > 0,
                 "rgbdv_selecting: bad sizes");
  if (n_rays == 0) return SAILOR_OK;
  hipStream_t st = sailor::as_stream(stream);
  const long long nv = (long long)B * V;
  const unsigned g = grid_of(nv * n_rays);
  hipLaunchKernelGGL(select_kernel, dim3(g), dim3(kBlock), 0, st, rgbs, rgb_channels, H, W,
                     sampled_xy, nv, n_rays, out_rgb, nullptr);
  if (int e = sailor::check_launch("rgbdv_selecting (rgb)")) return e;
  hipLaunchKernelGGL(select_kernel, dim3(g), dim3(kBlock), 0, st, normals, 3, H, W, sampled_xy,
                     nv, n_rays, out_normals, nullptr);
  if (int e = sailor::check_launch("rgbdv_selecting (normals)")) return e;
  hipLaunchKernelGGL(select_kernel, dim3(g), dim3(kBlock), 0, st, depths, 1, H, W, sampled_xy,
                     nv, n_rays, out_depth, out_label);
  return sailor::check_launch("rgbdv_selecting (depth)");
}

extern "C" int sailor_rgb_normalize(const float* rgb, const float* mask, int N, int C, int H,
                                    int W, float mid, float dist, float* out, void* stream) {
  SAILOR_REQUIRE(N >= 0 && C > 0 && H > 0 && W > 0, "rgb_normalizer: bad sizes");
  const long long hw = (long long)H * W, total = (long long)N * C * hw;
  if (total == 0) return SAILOR_OK;
  hipLaunchKernelGGL(rgb_normalize_kernel, dim3(grid_of(total)), dim3(kBlock), 0,
                     sailor::as_stream(stream), rgb, mask, C, hw, total, mid, dist, out);
  return sailor::check_launch("rgb_normalizer");
}
