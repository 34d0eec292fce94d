// Image taps of F.grid_sample(align_corners=True, padding_mode='zeros') on
// channels-last maps [H][W][C], shared by the feature gathers.
#pragma once

#include "common.h"

namespace sailor {

// Pixel coordinate of a normalised grid value with align_corners=True.
__device__ __forceinline__ float grid_to_pixel(float g, int size) {
  return (g + 1.0f) * 0.5f * (float)(size - 1);
}

// mode='bilinear': the 4 taps weighted by the fractional position; taps
// outside the map contribute zero.
template <int C>
__device__ __forceinline__ void bilinear_cl(const float* __restrict__ map, int H, int W, float gx,
                                            float gy, float (&out)[C]) {
#pragma unroll
  for (int c = 0; c < C; ++c) out[c] = 0.0f;
  const float ix = grid_to_pixel(gx, W);
  const float iy = grid_to_pixel(gy, H);
  if (!(ix > -1.0f && ix < (float)W && iy > -1.0f && iy < (float)H)) return;
  const float fx0 = floorf(ix), fy0 = floorf(iy);
  const int x0 = (int)fx0, y0 = (int)fy0;
  const float ax = ix - fx0, ay = iy - fy0;
#pragma unroll
  for (int tap = 0; tap < 4; ++tap) {
    const int xx = x0 + (tap & 1), yy = y0 + (tap >> 1);
    if (xx < 0 || xx >= W || yy < 0 || yy >= H) continue;
    const float w = ((tap & 1) ? ax : 1.0f - ax) * ((tap >> 1) ? ay : 1.0f - ay);
    const float* src = map + ((size_t)yy * W + xx) * C;
    if constexpr (C % 4 == 0) {
#pragma unroll
      for (int c = 0; c < C; c += 4) {
        const float4 v = *reinterpret_cast<const float4*>(src + c);
        out[c] += w * v.x;
        out[c + 1] += w * v.y;
        out[c + 2] += w * v.z;
        out[c + 3] += w * v.w;
      }
    } else {
#pragma unroll
      for (int c = 0; c < C; ++c) out[c] += w * src[c];
    }
  }
}

// mode='nearest': the pixel at the coordinate rounded half to even (rintf,
// as nearbyint under the default rounding mode); zero outside the map.
template <int C>
__device__ __forceinline__ void nearest_cl(const float* __restrict__ map, int H, int W, float gx,
                                           float gy, float (&out)[C]) {
  const float ix = rintf(grid_to_pixel(gx, W));
  const float iy = rintf(grid_to_pixel(gy, H));
  const bool in = ix >= 0.0f && ix < (float)W && iy >= 0.0f && iy < (float)H;
  const float* src = map + (in ? ((size_t)iy * W + (size_t)ix) * C : 0);
#pragma unroll
  for (int c = 0; c < C; ++c) out[c] = in ? src[c] : 0.0f;
}

}  // namespace sailor
