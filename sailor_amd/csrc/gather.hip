// Per-view sample features for the FastNerf inference surface: the chain
// proj_persp -> z_normalizer -> 4 x grid_sample -> t-psdf that the reference
// runs as separate PyTorch ops before InferNerf.infer_nerf
// (gui/gui_render.py:713-751, SRONet/BasicRenNet.py:245-259,296-311,
// utils_render/utils_render.py:380-410,656-663), fused into one pass:
// one thread per (batch, view, ray, sample), channels-last maps so a bilinear
// tap of the 16 feature channels is one 64-byte read, outputs written in the
// (batch, view, ray, sample) order infer_nerf consumes.
#include "image_sample.h"

namespace {

constexpr int kGeoC = 16, kRgbC = 16;
using sailor::bilinear_cl;

struct GatherArgs {
  const float* rays;   // [B,R,6]
  const float* dists;  // [B,R,S]
  const float* K;      // [B,V,3,3]
  const float* RT;     // [B,V,3,4]
  const float* geo;    // [B,V,Hg,Wg,16]
  const float* depth;  // [B,V,Hd,Wd]
  const float* rgbf;   // [B,V,Hr,Wr,16]
  const float* rgb;    // [B,V,Hc,Wc,3]
  const float* mid;    // [B,V]
  const float* dist;   // [B,V]
  int B, R, S, V, Hg, Wg, Hd, Wd, Hr, Wr, Hc, Wc, res;
  float* out_z;        // [B,V,R,S]
  float* out_tpsdf;    // [B,V,R,S]
  float* out_geo;      // [B,V,R,S,16]
  float* out_rgbf;     // [B,V,R,S,16]
  float* out_rgb;      // [B,V,R,S,3]
  float* out_dir;      // [B,V,R,S,3]
};

__global__ void __launch_bounds__(256) sample_features_kernel(GatherArgs a) {
  const long long total = (long long)a.B * a.V * a.R * a.S;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const int s = (int)(e % a.S);
    const long long rv = e / a.S;
    const int r = (int)(rv % a.R);
    const long long bv = rv / a.R;
    const int v = (int)(bv % a.V);
    const int b = (int)(bv / a.V);
    const float* ray = a.rays + ((size_t)b * a.R + r) * 6;
    const float t = a.dists[((size_t)b * a.R + r) * a.S + s];
    // rays_points = ray_ori + ray_dir * dist (gui_render.py:723)
    const float px = ray[0] + ray[3] * t, py = ray[1] + ray[4] * t, pz = ray[2] + ray[5] * t;
    const float* Rm = a.RT + (size_t)bv * 12;
    const float* Km = a.K + (size_t)bv * 9;
    // proj_persp (utils_render.py:380-410): pts = R p + T
    const float x = Rm[0] * px + Rm[1] * py + Rm[2] * pz + Rm[3];
    const float y = Rm[4] * px + Rm[5] * py + Rm[6] * pz + Rm[7];
    const float z = Rm[8] * px + Rm[9] * py + Rm[10] * pz + Rm[11];
    const bool invalid = z < 0.01f;
    const bool invalid_d = t < 0.1f || t > 4.0f;  // p_depths outside [min_d, max_d]
    const float zd = invalid ? 1.0f : z;
    const float xn = x / zd, yn = y / zd, zn = z / zd;
    const float half = (float)(a.res / 2);
    float gx = (Km[0] * xn + Km[1] * yn + Km[2] * zn - half) / half;
    float gy = (Km[3] * xn + Km[4] * yn + Km[5] * zn - half) / half;
    if (invalid || invalid_d) gx = gy = -2.0f;
    const float proj_z = invalid_d ? -2.0f : z;
    // z_normalizer (utils_render.py:656-663)
    a.out_z[e] = (proj_z - a.mid[bv]) / (a.dist[bv] / 2.0f);
    float dep[1];
    bilinear_cl<1>(a.depth + (size_t)bv * a.Hd * a.Wd, a.Hd, a.Wd, gx, gy, dep);
    // t-psdf (gui_render.py:737)
    a.out_tpsdf[e] = fminf(fmaxf(proj_z - dep[0], -0.01f), 0.01f) * (1.0f / 0.01f);
    float g[kGeoC];
    bilinear_cl<kGeoC>(a.geo + (size_t)bv * a.Hg * a.Wg * kGeoC, a.Hg, a.Wg, gx, gy, g);
    float4* og = reinterpret_cast<float4*>(a.out_geo + e * kGeoC);
#pragma unroll
    for (int c = 0; c < kGeoC; c += 4) og[c / 4] = make_float4(g[c], g[c + 1], g[c + 2], g[c + 3]);
    float f[kRgbC];
    bilinear_cl<kRgbC>(a.rgbf + (size_t)bv * a.Hr * a.Wr * kRgbC, a.Hr, a.Wr, gx, gy, f);
    float4* of = reinterpret_cast<float4*>(a.out_rgbf + e * kRgbC);
#pragma unroll
    for (int c = 0; c < kRgbC; c += 4) of[c / 4] = make_float4(f[c], f[c + 1], f[c + 2], f[c + 3]);
    float c3[3];
    bilinear_cl<3>(a.rgb + (size_t)bv * a.Hc * a.Wc * 3, a.Hc, a.Wc, gx, gy, c3);
    a.out_rgb[e * 3 + 0] = c3[0];
    a.out_rgb[e * 3 + 1] = c3[1];
    a.out_rgb[e * 3 + 2] = c3[2];
    // ray direction in camera v (R @ dir, BasicRenNet.py:118-131)
    a.out_dir[e * 3 + 0] = Rm[0] * ray[3] + Rm[1] * ray[4] + Rm[2] * ray[5];
    a.out_dir[e * 3 + 1] = Rm[4] * ray[3] + Rm[5] * ray[4] + Rm[6] * ray[5];
    a.out_dir[e * 3 + 2] = Rm[8] * ray[3] + Rm[9] * ray[4] + Rm[10] * ray[5];
  }
}

// [N][C][H][W] -> [N][H][W][C]; one thread per pixel (coalesced plane reads
// and contiguous channel-block writes).
__global__ void __launch_bounds__(256) nchw_to_nhwc_kernel(const float* __restrict__ src, int N,
                                                           int C, int HW,
                                                           float* __restrict__ dst) {
  const long long total = (long long)N * HW;
  for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < total;
       p += (long long)gridDim.x * blockDim.x) {
    const long long n = p / HW, i = p - n * HW;
    const float* s = src + n * C * HW + i;
    float* d = dst + p * C;
    for (int c = 0; c < C; ++c) d[c] = s[(long long)c * HW];
  }
}

unsigned grid_of(long long work) {
  return (unsigned)std::max<long long>(1, std::min<long long>((work + 255) / 256, 256 * 16));
}

}  // namespace

extern "C" int sailor_nchw_to_nhwc(const float* src, int N, int C, int H, int W, float* dst,
                                   void* stream) {
  SAILOR_REQUIRE(N >= 0 && C > 0 && H > 0 && W > 0, "nchw_to_nhwc: bad sizes");
  if (N == 0) return SAILOR_OK;
  hipLaunchKernelGGL(nchw_to_nhwc_kernel, dim3(grid_of((long long)N * H * W)), dim3(256), 0,
                     sailor::as_stream(stream), src, N, C, H * W, dst);
  return sailor::check_launch("nchw_to_nhwc");
}

extern "C" int sailor_sample_features(
    const float* rays, const float* dists, int B, int R, int S, const float* K, const float* RT,
    int V, const float* geo_cl, int Hg, int Wg, const float* depth, int Hd, int Wd,
    const float* rgbf_cl, int Hr, int Wr, const float* rgb_cl, int Hc, int Wc,
    const float* depth_mid, const float* depth_dist, int res, float* out_z, float* out_tpsdf,
    float* out_geo, float* out_rgbf, float* out_rgb, float* out_dir, void* stream) {
  SAILOR_REQUIRE(B > 0 && R >= 0 && S > 0 && V > 0, "sample_features: bad sizes");
  SAILOR_REQUIRE(Hg > 1 && Wg > 1 && Hd > 1 && Wd > 1 && Hr > 1 && Wr > 1 && Hc > 1 && Wc > 1,
                 "sample_features: maps must be at least 2 x 2");
  SAILOR_REQUIRE(res >= 2, "sample_features: res must be at least 2");
  if (R == 0) return SAILOR_OK;
  GatherArgs a{rays, dists, K, RT, geo_cl, depth, rgbf_cl, rgb_cl, depth_mid, depth_dist,
               B, R, S, V, Hg, Wg, Hd, Wd, Hr, Wr, Hc, Wc, res,
               out_z, out_tpsdf, out_geo, out_rgbf, out_rgb, out_dir};
  hipLaunchKernelGGL(sample_features_kernel, dim3(grid_of((long long)B * V * R * S)), dim3(256),
                     0, sailor::as_stream(stream), a);
  return sailor::check_launch("sample_features");
}
