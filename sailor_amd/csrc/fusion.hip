// TSDF fusion: adaptive bounding box + projective TSDF integration.
//
// Follows c_lib/VoxelEncoding/src/fusion_kernel.cu (integrate_cuda :236-340,
// obtain_origin_bbox :137-184, obtain_resized_bbox :51-103, inverse_mat3x3
// :105-135).  Compiled with -ffp-contract=off: every arithmetic expression is
// evaluated as written (one rounding per operation), which is what makes the
// occupied-voxel list reproducible bit for bit by the CPU oracle.
//
// MI355X mapping
//  * bbox: one pass over the depth maps; each workgroup reduces its pixels in
//    registers / LDS and issues 6 integer atomics (order-preserving float keys)
//    instead of one CAS loop per pixel and bound.
//  * integrate: one lane per voxel, z fastest, so the 16 B/voxel of output
//    (tsdf, weight, colour, occupancy) is written with fully coalesced stores;
//    the per-view camera rows are wave-uniform scalar loads; the occupied
//    count is aggregated with a wave ballot and one atomic per workgroup.
#include "common.h"

namespace {

constexpr int kThreads = 256;
constexpr int kMaxViews = 64;

// 3x3 inverse by adjugate / determinant, same expression order as the
// reference (fusion_kernel.cu:115-133).
__device__ __forceinline__ void inverse3x3(const float* M, float* out) {
  float a1 = M[0], b1 = M[1], c1 = M[2];
  float a2 = M[3], b2 = M[4], c2 = M[5];
  float a3 = M[6], b3 = M[7], c3 = M[8];
  float det = a1 * (b2 * c3 - c2 * b3) - a2 * (b1 * c3 - c1 * b3) +
              a3 * (b1 * c2 - c1 * b2);
  out[0] = (b2 * c3 - c2 * b3) / det;
  out[1] = (b3 * c1 - c3 * b1) / det;
  out[2] = (b1 * c2 - c1 * b2) / det;
  out[3] = (a3 * c2 - c3 * a2) / det;
  out[4] = (a1 * c3 - c1 * a3) / det;
  out[5] = (a2 * c1 - c2 * a1) / det;
  out[6] = (a2 * b3 - b2 * a3) / det;
  out[7] = (b1 * a3 - a1 * b3) / det;
  out[8] = (b2 * a1 - a2 * b1) / det;
}

__global__ void bbox_init_keys(int* keys, int B) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= B * 6) return;
  // [B,3,2]: (min, max) per axis; initial values of the reference (:655-656)
  keys[i] = (i & 1) ? sailor::fkey(-99999.0f) : sailor::fkey(99999.0f);
}

// Back-projects every valid pixel (0.1 m < d <= 5 m) and reduces min/max of
// (p - th) / (p + th) per batch.  grid = (blocks, B).
__global__ void __launch_bounds__(kThreads)
bbox_reduce(const float* __restrict__ depths, const float* __restrict__ K,
            const float* __restrict__ R, const float* __restrict__ T, int V,
            int H, int W, float thres, int* keys) {
  __shared__ float s_inv[kMaxViews * 18];
  __shared__ float s_red[6][kThreads / 64];
  const int b = blockIdx.y;
  for (int v = threadIdx.x; v < V; v += blockDim.x) {
    inverse3x3(K + ((size_t)b * V + v) * 9, s_inv + v * 18);
    inverse3x3(R + ((size_t)b * V + v) * 9, s_inv + v * 18 + 9);
  }
  __syncthreads();

  float lo[3] = {3.0e38f, 3.0e38f, 3.0e38f};
  float hi[3] = {-3.0e38f, -3.0e38f, -3.0e38f};
  const long long HW = (long long)H * W;
  const long long n = (long long)V * HW;
  const float* dep = depths + (size_t)b * n;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    float d = dep[i];
    if (d < 0.1f || d > 5.0f) continue;
    int v = (int)(i / HW);
    long long r = i - (long long)v * HW;
    int py = (int)(r / W);
    int px = (int)(r - (long long)py * W);
    const float* ki = s_inv + v * 18;
    const float* ri = ki + 9;
    const float* t = T + ((size_t)b * V + v) * 3;
    float fx = (float)px, fy = (float)py;
    float cx = (ki[0] * fx + ki[1] * fy + ki[2]) * d;
    float cy = (ki[3] * fx + ki[4] * fy + ki[5]) * d;
    float cz = (ki[6] * fx + ki[7] * fy + ki[8]) * d;
    cx -= t[0];
    cy -= t[1];
    cz -= t[2];
    float p[3];
    p[0] = ri[0] * cx + ri[1] * cy + ri[2] * cz;
    p[1] = ri[3] * cx + ri[4] * cy + ri[5] * cz;
    p[2] = ri[6] * cx + ri[7] * cy + ri[8] * cz;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      lo[a] = fminf(lo[a], p[a] - thres);
      hi[a] = fmaxf(hi[a], p[a] + thres);
    }
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    float l = sailor::wave_min_f(lo[a]);
    float h = sailor::wave_max_f(hi[a]);
    if (lane == 0) {
      s_red[2 * a][wid] = l;
      s_red[2 * a + 1][wid] = h;
    }
  }
  __syncthreads();
  if (threadIdx.x < 6) {
    const int j = threadIdx.x;
    float acc = s_red[j][0];
    for (int w = 1; w < (int)(blockDim.x / 64); ++w)
      acc = (j & 1) ? fmaxf(acc, s_red[j][w]) : fminf(acc, s_red[j][w]);
    // Only contribute if this block saw a valid pixel.
    bool seen = (j & 1) ? (acc > -3.0e38f) : (acc < 3.0e38f);
    if (seen) {
      if (j & 1)
        atomicMax(keys + b * 6 + j, sailor::fkey(acc));
      else
        atomicMin(keys + b * 6 + j, sailor::fkey(acc));
    }
  }
}

// keys (stored in-place in vol_bbox) -> vol_origin, cubic vol_bbox, voxel_size
// (obtain_resized_bbox, fusion_kernel.cu:51-103).  One thread per batch.
__global__ void bbox_finalize(float* vol_bbox_keys, int B, int dim,
                              const float* fixed_bbox, float* vol_origin,
                              float* voxel_size) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  float o[6];
  if (fixed_bbox) {
    for (int j = 0; j < 6; ++j) o[j] = fixed_bbox[b * 6 + j];
  } else {
    const int* keys = reinterpret_cast<const int*>(vol_bbox_keys) + b * 6;
    for (int j = 0; j < 6; ++j) o[j] = sailor::fkey_inv(keys[j]);
  }
  for (int j = 0; j < 6; ++j) vol_origin[b * 6 + j] = o[j];
  if (fixed_bbox) {
    for (int j = 0; j < 6; ++j) vol_bbox_keys[b * 6 + j] = o[j];
    voxel_size[b] = (o[1] - o[0]) / (float)(dim - 1);
    return;
  }
  float x_min = o[0], x_max = o[1], y_min = o[2], y_max = o[3], z_min = o[4],
        z_max = o[5];
  float x_len = x_max - x_min, y_len = y_max - y_min, z_len = z_max - z_min;
  float max_len = fmaxf(x_len, fmaxf(y_len, z_len));
  // (a + b) / 2.0 in the reference is a double division of a float sum; the
  // halving is exact, so the float division below gives identical bits.
  float x_c = (x_min + x_max) / 2.0f, y_c = (y_min + y_max) / 2.0f,
        z_c = (z_min + z_max) / 2.0f;
  voxel_size[b] = max_len / (float)(dim - 1);
  float half = max_len / 2.0f;
  if (max_len == x_len) {
    y_min = y_c - half; y_max = y_c + half;
    z_min = z_c - half; z_max = z_c + half;
  } else if (max_len == y_len) {
    x_min = x_c - half; x_max = x_c + half;
    z_min = z_c - half; z_max = z_c + half;
  } else {
    x_min = x_c - half; x_max = x_c + half;
    y_min = y_c - half; y_max = y_c + half;
  }
  float* bb = vol_bbox_keys + b * 6;
  bb[0] = x_min; bb[1] = x_max; bb[2] = y_min;
  bb[3] = y_max; bb[4] = z_min; bb[5] = z_max;
}

__global__ void keys_to_float(float* buf, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  buf[i] = sailor::fkey_inv(reinterpret_cast<const int*>(buf)[i]);
}

// One lane per voxel; grid = (ceil(X*Y*Z / 256), B).  COMP (precision form):
// the TSDF is formed as the Kahan-compensated sum of obs_weight * dist over
// the observing views divided once by the weight sum, instead of the
// reference's running average (tsdf * w_old + obs_weight * dist) / w_new --
// the same value in exact arithmetic, with one rounding of the sum instead
// of one per view.
template <bool COMP>
__global__ void __launch_bounds__(kThreads)
integrate_kernel(float* __restrict__ tsdf_vol, float* __restrict__ weight_vol,
                 float* __restrict__ color_vol, int* __restrict__ occ_vol,
                 int* __restrict__ num_occ, const float* __restrict__ vol_bbox,
                 const float* __restrict__ K, const float* __restrict__ RT,
                 const float* __restrict__ colors,
                 const float* __restrict__ depths,
                 const float* __restrict__ voxel_size, float obs_weight, int H,
                 int W, int V, int X, int Y, int Z, float trunc_weight,
                 float th_low, float th_high) {
  __shared__ int s_cnt[kThreads / 64];
  const int b = blockIdx.y;
  const int YZ = Y * Z;
  const int nvox = X * YZ;
  const int vid = blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = vid < nvox;
  const float vs = voxel_size[b];
  const float trunc = vs * trunc_weight;
  bool occupied = false;
  if (active) {
    int vx = vid / YZ;
    int vy = (vid - vx * YZ) / Z;
    int vz = vid - vx * YZ - vy * Z;
    const float* bb = vol_bbox + b * 6;
    float pt_x = bb[0] + (float)vx * vs;
    float pt_y = bb[2] + (float)vy * vs;
    float pt_z = bb[4] + (float)vz * vs;
    float tsdf = 1.0f, w = 0.0f, col = 0.0f;
    float ksum = 0.0f, kcomp = 0.0f;  // COMP: sum of obs_weight * dist, lost low bits
    const size_t img = (size_t)H * W;
    for (int k = 0; k < V; ++k) {
      const float* rt = RT + ((size_t)b * V + k) * 12;
      const float* kk = K + ((size_t)b * V + k) * 9;
      float cx = rt[0] * pt_x + rt[1] * pt_y + rt[2] * pt_z + rt[3];
      float cy = rt[4] * pt_x + rt[5] * pt_y + rt[6] * pt_z + rt[7];
      float cz = rt[8] * pt_x + rt[9] * pt_y + rt[10] * pt_z + rt[11];
      float px = kk[0] * cx + kk[1] * cy + kk[2] * cz;
      float py = kk[3] * cx + kk[4] * cy + kk[5] * cz;
      float pz = kk[6] * cx + kk[7] * cy + kk[8] * cz;
      px = px / pz;
      py = py / pz;
      int ix = sailor::sat_f2i(roundf(px));
      int iy = sailor::sat_f2i(roundf(py));
      if (ix < 0 || ix >= W || iy < 0 || iy >= H || cz < 0) continue;
      size_t pix = ((size_t)b * V + k) * img + (size_t)iy * W + ix;
      float d = depths[pix];
      if (d < 0.1f || d > 4.0f) continue;
      float diff = d - cz;
      if (diff < -trunc) continue;
      float dist = fminf(1.0f, diff / trunc);
      float w_old = w;
      float w_new = w_old + obs_weight;
      w = w_new;
      if constexpr (COMP) {
        const float y = obs_weight * dist - kcomp;
        const float t = ksum + y;
        kcomp = (t - ksum) - y;
        ksum = t;
      } else {
        tsdf = (tsdf * w_old + obs_weight * dist) / w_new;
      }
      // packed colour running average (fusion_kernel.cu:320-331)
      float old_b = floorf(col / 65536.0f);
      float old_g = floorf((col - old_b * 256.0f * 256.0f) / 256.0f);
      float old_r = col - old_b * 256.0f * 256.0f - old_g * 256.0f;
      float nc = colors[pix];
      float new_b = floorf(nc / 65536.0f);
      float new_g = floorf((nc - new_b * 256.0f * 256.0f) / 256.0f);
      float new_r = nc - new_b * 256.0f * 256.0f - new_g * 256.0f;
      new_b = fminf(roundf((old_b * w_old + obs_weight * new_b) / w_new), 255.0f);
      new_g = fminf(roundf((old_g * w_old + obs_weight * new_g) / w_new), 255.0f);
      new_r = fminf(roundf((old_r * w_old + obs_weight * new_r) / w_new), 255.0f);
      col = new_b * 256.0f * 256.0f + new_g * 256.0f + new_r;
    }
    if constexpr (COMP) {
      if (w > 0.0f) tsdf = (ksum - kcomp) / w;
    }
    occupied = tsdf < trunc * th_high && tsdf > -trunc * th_low;
    size_t o = (size_t)b * nvox + vid;
    tsdf_vol[o] = tsdf;
    weight_vol[o] = w;
    color_vol[o] = col;
    occ_vol[o] = occupied ? vid : -1;
  }
  // occupied count: ballot per wave, one atomic per workgroup
  unsigned long long m = __ballot(occupied);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) s_cnt[wid] = __popcll(m);
  __syncthreads();
  if (threadIdx.x == 0) {
    int c = 0;
    for (int i = 0; i < (int)(blockDim.x / 64); ++i) c += s_cnt[i];
    if (c) atomicAdd(num_occ + b, c);
  }
}

// Sum of the back-projected world points of the valid pixels (0.1 m <= d
// <= 4 m) and their count per batch (obtain_center_points,
// fusion_kernel.cu:186-233), with the same per-pixel arithmetic as the bbox
// pass.  The reference adds every point with float atomics (order-dependent
// low bits); here each workgroup sums its pixels in double and writes one
// partial [x, y, z, n]; center_finalize adds the partials in block order, so
// the result is reproducible and exact to float32 rounding of the double
// sum.  grid = (kCenterBlocks, B).
constexpr int kCenterBlocks = 256;

__global__ void __launch_bounds__(kThreads)
center_partials(const float* __restrict__ depths, const float* __restrict__ K,
                const float* __restrict__ R, const float* __restrict__ T, int V, int H, int W,
                double* __restrict__ partials) {
  __shared__ float s_inv[kMaxViews * 18];
  __shared__ double s_red[4][kThreads / 64];
  const int b = blockIdx.y;
  for (int v = threadIdx.x; v < V; v += blockDim.x) {
    inverse3x3(K + ((size_t)b * V + v) * 9, s_inv + v * 18);
    inverse3x3(R + ((size_t)b * V + v) * 9, s_inv + v * 18 + 9);
  }
  __syncthreads();
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  const long long HW = (long long)H * W;
  const long long n = (long long)V * HW;
  const float* dep = depths + (size_t)b * n;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    float d = dep[i];
    if (d < 0.1f || d > 4.0f) continue;
    int v = (int)(i / HW);
    long long r = i - (long long)v * HW;
    int py = (int)(r / W);
    int px = (int)(r - (long long)py * W);
    const float* ki = s_inv + v * 18;
    const float* ri = ki + 9;
    const float* t = T + ((size_t)b * V + v) * 3;
    float fx = (float)px, fy = (float)py;
    float cx = (ki[0] * fx + ki[1] * fy + ki[2]) * d;
    float cy = (ki[3] * fx + ki[4] * fy + ki[5]) * d;
    float cz = (ki[6] * fx + ki[7] * fy + ki[8]) * d;
    cx -= t[0];
    cy -= t[1];
    cz -= t[2];
    acc[0] += (double)(ri[0] * cx + ri[1] * cy + ri[2] * cz);
    acc[1] += (double)(ri[3] * cx + ri[4] * cy + ri[5] * cz);
    acc[2] += (double)(ri[6] * cx + ri[7] * cy + ri[8] * cz);
    acc[3] += 1.0;
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    double x = acc[a];
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    if (lane == 0) s_red[a][wid] = x;
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    double x = 0.0;
    for (int w = 0; w < kThreads / 64; ++w) x += s_red[threadIdx.x][w];
    partials[((size_t)b * gridDim.x + blockIdx.x) * 4 + threadIdx.x] = x;
  }
}

__global__ void center_finalize(const double* __restrict__ partials, int B, int blocks,
                                float* __restrict__ center, float* __restrict__ num_pts) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int k = 0; k < blocks; ++k)
    for (int a = 0; a < 4; ++a) acc[a] += partials[((size_t)b * blocks + k) * 4 + a];
  center[b * 3 + 0] = (float)acc[0];
  center[b * 3 + 1] = (float)acc[1];
  center[b * 3 + 2] = (float)acc[2];
  num_pts[b] = (float)acc[3];
}

int launch_bbox(const float* depths, const float* K, const float* R,
                const float* T, int B, int V, int H, int W, float th,
                int* keys, hipStream_t st) {
  hipLaunchKernelGGL(bbox_init_keys, dim3(sailor::ceil_div(B * 6, 64)),
                     dim3(64), 0, st, keys, B);
  long long n = (long long)V * H * W;
  int blocks = (int)std::min<long long>(2048, (n + kThreads - 1) / kThreads);
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(bbox_reduce, dim3(blocks, B), dim3(kThreads), 0, st,
                     depths, K, R, T, V, H, W, th, keys);
  return sailor::check_launch("bbox_reduce");
}

}  // namespace

namespace {

int integrate_impl(bool comp, int32_t* occ_vol, int32_t* num_occ, const float* K,
                   const float* RT, const float* R, const float* T, const float* colors,
                   const float* depths, int B, int V, int H, int W, int X, int Y, int Z,
                   float obs_weight, float th_bbox, float trunc_weight, float tsdf_th_low,
                   float tsdf_th_high, const float* fixed_bbox, float* tsdf, float* weight,
                   float* color, float* vol_origin, float* vol_bbox, float* voxel_size,
                   void* stream) {
  SAILOR_REQUIRE(B > 0 && V > 0 && H > 0 && W > 0, "integrate: empty input");
  SAILOR_REQUIRE(V <= kMaxViews, "integrate: at most %d views", kMaxViews);
  SAILOR_REQUIRE(X > 1 && Y > 0 && Z > 0, "integrate: bad volume dims");
  SAILOR_REQUIRE((long long)X * Y * Z < (1LL << 31), "integrate: volume too large");
  hipStream_t st = sailor::as_stream(stream);
  int* keys = reinterpret_cast<int*>(vol_bbox);
  if (!fixed_bbox) {
    int rc = launch_bbox(depths, K, R, T, B, V, H, W, th_bbox, keys, st);
    if (rc) return rc;
  }
  hipLaunchKernelGGL(bbox_finalize, dim3(sailor::ceil_div(B, 64)), dim3(64), 0,
                     st, vol_bbox, B, X, fixed_bbox, vol_origin, voxel_size);
  const int nvox = X * Y * Z;
  hipLaunchKernelGGL(comp ? integrate_kernel<true> : integrate_kernel<false>,
                     dim3(sailor::ceil_div(nvox, kThreads), B),
                     dim3(kThreads), 0, st, tsdf, weight, color, occ_vol,
                     num_occ, vol_bbox, K, RT, colors, depths, voxel_size,
                     obs_weight, H, W, V, X, Y, Z, trunc_weight, tsdf_th_low,
                     tsdf_th_high);
  return sailor::check_launch("integrate");
}

}  // namespace

extern "C" int sailor_integrate(
    int32_t* occ_vol, int32_t* num_occ, const float* K, const float* RT,
    const float* R, const float* T, const float* colors, const float* depths,
    const int32_t* masks, int B, int V, int H, int W, int X, int Y, int Z,
    float obs_weight, float th_bbox, float trunc_weight, float tsdf_th_low,
    float tsdf_th_high, const float* fixed_bbox, float* tsdf, float* weight,
    float* color, float* vol_origin, float* vol_bbox, float* voxel_size,
    void* stream) {
  (void)masks;  // read by nothing in the reference kernel either (:316-317)
  return integrate_impl(false, occ_vol, num_occ, K, RT, R, T, colors, depths, B, V, H, W, X, Y,
                        Z, obs_weight, th_bbox, trunc_weight, tsdf_th_low, tsdf_th_high,
                        fixed_bbox, tsdf, weight, color, vol_origin, vol_bbox, voxel_size,
                        stream);
}

extern "C" int sailor_integrate_compensated(
    int32_t* occ_vol, int32_t* num_occ, const float* K, const float* RT,
    const float* R, const float* T, const float* colors, const float* depths,
    const int32_t* masks, int B, int V, int H, int W, int X, int Y, int Z,
    float obs_weight, float th_bbox, float trunc_weight, float tsdf_th_low,
    float tsdf_th_high, const float* fixed_bbox, float* tsdf, float* weight,
    float* color, float* vol_origin, float* vol_bbox, float* voxel_size,
    void* stream) {
  (void)masks;
  return integrate_impl(true, occ_vol, num_occ, K, RT, R, T, colors, depths, B, V, H, W, X, Y,
                        Z, obs_weight, th_bbox, trunc_weight, tsdf_th_low, tsdf_th_high,
                        fixed_bbox, tsdf, weight, color, vol_origin, vol_bbox, voxel_size,
                        stream);
}

extern "C" int sailor_get_origin_bbox(const float* depths, const float* K,
                                      const float* R, const float* T, int B,
                                      int V, int H, int W, float th_bbox,
                                      float* vol_origin, void* stream) {
  SAILOR_REQUIRE(B > 0 && V > 0 && V <= kMaxViews, "get_origin_bbox: bad B/V");
  hipStream_t st = sailor::as_stream(stream);
  int rc = launch_bbox(depths, K, R, T, B, V, H, W, th_bbox,
                       reinterpret_cast<int*>(vol_origin), st);
  if (rc) return rc;
  hipLaunchKernelGGL(keys_to_float, dim3(sailor::ceil_div(B * 6, 64)), dim3(64),
                     0, st, vol_origin, B * 6);
  return sailor::check_launch("get_origin_bbox");
}

extern "C" int sailor_get_center_xyz(const float* depths, const float* K, const float* R,
                                     const float* T, int B, int V, int H, int W,
                                     float* vol_center, float* vol_num_pts, void* workspace,
                                     size_t workspace_bytes, void* stream) {
  SAILOR_REQUIRE(B > 0 && V > 0 && V <= kMaxViews && H > 0 && W > 0,
                 "get_center_xyz: bad sizes (B=%d, V=%d, H=%d, W=%d)", B, V, H, W);
  const size_t need = (size_t)B * kCenterBlocks * 4 * sizeof(double);
  SAILOR_REQUIRE(workspace_bytes >= need, "get_center_xyz: workspace too small (%zu < %zu)",
                 workspace_bytes, need);
  hipStream_t st = sailor::as_stream(stream);
  double* part = reinterpret_cast<double*>(workspace);
  hipLaunchKernelGGL(center_partials, dim3(kCenterBlocks, B), dim3(kThreads), 0, st, depths,
                     K, R, T, V, H, W, part);
  hipLaunchKernelGGL(center_finalize, dim3(sailor::ceil_div(B, 64)), dim3(64), 0, st, part, B,
                     kCenterBlocks, vol_center, vol_num_pts);
  return sailor::check_launch("get_center_xyz");
}
