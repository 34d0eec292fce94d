// Per-pixel camera rays: origin = -R^-1 T, direction = normalize(R^-1 K^-1 [w, h, 1]).
//
// Follows c_lib/VoxelEncoding/src/sampling_kernel.cu rays_calculating_kernel
// (:411-486) and rays_calculating_parallel_kernel (:330-407, image row
// h*num_gpus + device); the 3x3 inverses follow inverse_mat3x3 (:17-48).
// The normalisation divides by (|d| + 1e-10) in double precision exactly as
// the reference expression does (a float divided by a double).
// Compiled with -ffp-contract=off.
#include "common.h"

namespace {

constexpr int kThreads = 256;
constexpr int kMaxViews = 64;

__device__ __forceinline__ void inv3(const float* M, float* out) {
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

// grid = (ceil(H*W/256), B*V)
__global__ void __launch_bounds__(kThreads)
rays_kernel(const float* __restrict__ K, const float* __restrict__ R,
            const float* __restrict__ T, int H, int W, int row_stride,
            int row_offset, float* __restrict__ out) {
  __shared__ float s_m[18];
  const int bv = blockIdx.y;
  if (threadIdx.x == 0) {
    inv3(K + (size_t)bv * 9, s_m);
    inv3(R + (size_t)bv * 9, s_m + 9);
  }
  __syncthreads();
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= H * W) return;
  const int h = p / W, w = p - h * W;
  const int he = h * row_stride + row_offset;
  const float* ki = s_m;
  const float* ri = s_m + 9;
  const float* t = T + (size_t)bv * 3;
  float cc[3], cd[3], wd[3];
#pragma unroll
  for (int k = 0; k < 3; ++k)
    cc[k] = -ri[3 * k + 0] * t[0] - ri[3 * k + 1] * t[1] - ri[3 * k + 2] * t[2];
#pragma unroll
  for (int k = 0; k < 3; ++k)
    cd[k] = (float)w * ki[3 * k + 0] + (float)he * ki[3 * k + 1] + 1.0f * ki[3 * k + 2];
#pragma unroll
  for (int k = 0; k < 3; ++k)
    wd[k] = cd[0] * ri[3 * k + 0] + cd[1] * ri[3 * k + 1] + cd[2] * ri[3 * k + 2];
  const float nrm = sqrtf(wd[0] * wd[0] + wd[1] * wd[1] + wd[2] * wd[2]);
  const double den = (double)nrm + 1e-10;
  float* o = out + ((size_t)bv * H * W + p) * 6;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    o[k] = cc[k];
    o[3 + k] = (float)((double)wd[k] / den);
  }
}

}  // namespace

extern "C" int sailor_rays_calculating(const float* K, const float* R, const float* T,
                                       int B, int V, int H, int W, int row_stride,
                                       int row_offset, float* out, void* stream) {
  SAILOR_REQUIRE(B > 0 && V > 0 && H > 0 && W > 0, "rays_calculating: bad sizes");
  SAILOR_REQUIRE(row_stride >= 1 && row_offset >= 0, "rays_calculating: bad row mapping");
  hipLaunchKernelGGL(rays_kernel, dim3(sailor::ceil_div((long long)H * W, kThreads), B * V),
                     dim3(kThreads), 0, sailor::as_stream(stream), K, R, T, H, W,
                     row_stride, row_offset, out);
  return sailor::check_launch("rays_calculating");
}
