// FastNerf inference surface (InferNerf.infer_nerf / infer_density /
// infer_blending, c_lib/FastNerf/src/binding.cpp) on MI355X.
//
// The reference evaluates its small MLPs one point per thread in half
// precision with 128-entry per-thread half arrays and per-call stream
// creation (infer_mlp.cu:120-506, 638-789).  Here:
//  * every parameter group is staged once per workgroup into LDS as the
//    given half values; all arithmetic is fp32 (v_fma_mix: half weight x
//    fp32 activation), so the result is the exact value of the half
//    parameters to fp32 rounding rather than a half-accumulated one;
//  * one thread per (point, all views): the per-view feature MLP, the view
//    mean and the occupancy MLP run back to back, the geometry features of
//    the texture views go to a workspace, no [N, 64] intermediate;
//  * one thread per (ray, texture view): compositing of the 30-channel
//    features along the ray, then the two Hydra stacks with the cross-view
//    sums as 2 lane shuffles (4 lanes per ray) and the RGB head on the lane
//    of view 0; no host synchronisation, everything on the caller's stream.
#include "image_sample.h"

namespace {

typedef _Float16 half_t;
constexpr int kBlock = 256;
constexpr int kDensityParams = 22521, kFeatParams = 13696, kGeoOffset = 22081;
constexpr int kColorParams = 15807, kStackParams = 5822, kRgbOffset = 11644;
constexpr int kBlendParams = 15075;
constexpr int kTexSlots = 4;  // texture views per ray (the reference's NUM_VIEWS)

// out[o] = act(b[o] + sum_i W[o][i] in[i]); W [OUT][IN] then b [OUT] at w
// (half, LDS), in/out fp32 registers.  ACT: 0 none, 1 relu, 2 sigmoid.
template <int IN, int OUT, int ACT, bool BIAS = true>
__device__ __forceinline__ void dense(const half_t* __restrict__ w, const float (&in)[IN],
                                      float (&out)[OUT]) {
#pragma unroll
  for (int o = 0; o < OUT; ++o) {
    float acc = BIAS ? (float)w[OUT * IN + o] : 0.0f;
#pragma unroll
    for (int i = 0; i < IN; ++i) acc = fmaf((float)w[o * IN + i], in[i], acc);
    if (ACT == 1) acc = fmaxf(acc, 0.0f);
    if (ACT == 2) acc = 1.0f / (1.0f + expf(-acc));
    out[o] = acc;
  }
}
template <int IN, int OUT>
__device__ __forceinline__ constexpr int dense_size(bool bias = true) {
  return IN * OUT + (bias ? OUT : 0);
}

// The same with the outputs written to this thread's LDS row (o is a
// runtime loop, so only the input vector and one output live in registers;
// unrolling both loops makes the compiler hoist thousands of weight loads).
template <int IN, int OUT, int ACT, bool BIAS = true>
__device__ __forceinline__ void dense_row(const half_t* __restrict__ w, const float (&in)[IN],
                                          float* row) {
#pragma unroll 2
  for (int o = 0; o < OUT; ++o) {
    float acc = BIAS ? (float)w[OUT * IN + o] : 0.0f;
    const half_t* wr = w + o * IN;
#pragma unroll
    for (int i = 0; i < IN; ++i) acc = fmaf((float)wr[i], in[i], acc);
    if (ACT == 1) acc = fmaxf(acc, 0.0f);
    if (ACT == 2) acc = 1.0f / (1.0f + expf(-acc));
    row[o] = acc;
  }
}
template <int N>
__device__ __forceinline__ void from_row(const float* row, float (&out)[N]) {
#pragma unroll
  for (int i = 0; i < N; ++i) out[i] = row[i];
}

// Per-thread LDS rows: kRow floats each (odd stride: lane t's element o sits
// in bank (t + o) mod 64, so same-o accesses of a wave are conflict-free).
constexpr int kRow = 65;
constexpr size_t rows_bytes() { return (size_t)kBlock * kRow * sizeof(float); }

__device__ __forceinline__ void stage(half_t* dst, const half_t* __restrict__ src, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) dst[i] = src[i];
  __syncthreads();
}

// Per point p (flattened (b, point)): feature MLP of every view (inputs z,
// tpsdf, geo[16] at (b, v, point)), mean over views, occupancy MLP -> occ[p];
// with geo_out, the geometry MLP of views < n_geo_out -> geo_out[(b, v,
// point)][8] (infer_mlp.cu:120-240, 583-634, 681-712).
__global__ void __launch_bounds__(kBlock)
density_kernel(const half_t* __restrict__ params, const float* __restrict__ z,
               const float* __restrict__ tpsdf, const float* __restrict__ geo, int B,
               long long n_pts, int V, float* __restrict__ occ, float* __restrict__ geo_out,
               int n_geo_out) {
  extern __shared__ __attribute__((aligned(16))) char smem_i[];
  float* row = reinterpret_cast<float*>(smem_i) + threadIdx.x * kRow;
  half_t* w = reinterpret_cast<half_t*>(smem_i + rows_bytes());
  stage(w, params, kDensityParams);
  const long long total = (long long)B * n_pts;
  for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < total;
       p += (long long)gridDim.x * blockDim.x) {
    const long long b = p / n_pts, q = p - b * n_pts;
    float sum[64];
#pragma unroll
    for (int i = 0; i < 64; ++i) sum[i] = 0.0f;
    for (int v = 0; v < V; ++v) {
      const long long e = (b * V + v) * n_pts + q;
      float x[18];
      x[0] = z[e];
      x[1] = tpsdf[e];
      const float4* g4 = reinterpret_cast<const float4*>(geo + e * 16);
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const float4 t = g4[c];
        x[2 + 4 * c] = t.x;
        x[3 + 4 * c] = t.y;
        x[4 + 4 * c] = t.z;
        x[5 + 4 * c] = t.w;
      }
      float h[64];
      const half_t* wp = w;
      dense_row<18, 64, 1>(wp, x, row);
      from_row(row, h);
      wp += dense_size<18, 64>();
#pragma unroll 1
      for (int l = 0; l < 3; ++l) {
        dense_row<64, 64, 1>(wp, h, row);
        from_row(row, h);
        wp += dense_size<64, 64>();
      }
#pragma unroll
      for (int i = 0; i < 64; ++i) sum[i] += h[i];
      if (geo_out && v < n_geo_out) {
        float g16[16], g8[8];
        dense<18, 16, 1>(w + kGeoOffset, x, g16);
        dense<16, 8, 1>(w + kGeoOffset + dense_size<18, 16>(), g16, g8);
        float4* o4 = reinterpret_cast<float4*>(geo_out + ((b * n_geo_out + v) * n_pts + q) * 8);
        o4[0] = make_float4(g8[0], g8[1], g8[2], g8[3]);
        o4[1] = make_float4(g8[4], g8[5], g8[6], g8[7]);
      }
    }
    const float inv = 1.0f / (float)V;
#pragma unroll
    for (int i = 0; i < 64; ++i) sum[i] *= inv;
    float h[64], o[1];
    const half_t* wp = w + kFeatParams;
    dense_row<64, 64, 1>(wp, sum, row);
    from_row(row, h);
    wp += dense_size<64, 64>();
    dense_row<64, 64, 1>(wp, h, row);
    from_row(row, h);
    wp += dense_size<64, 64>();
    dense<64, 1, 2>(wp, h, o);
    occ[p] = o[0];
  }
}

// ---------------------------------------------------------------------------
// Density MLPs on MFMA (v_mfma_f32_32x32x16_f16): the reference evaluates
// them in half precision; here the layer inputs are rounded to half, the
// products are accumulated in fp32 (the half weights are exact).
//
// A wave takes 32 columns = 32 / VS points x VS view slots (VS = V rounded up
// to a power of two <= 8).  Weight rows are permuted so that, with the 32x32
// accumulator layout (lane l, accumulator r -> row 8 (r/4) + 4 (l/32) + r%4,
// column l%32), accumulator r of M-tile mt in lane half h holds output unit
// 32 mt + 16 (r/8) + 8 h + r%8: exactly the 8 inputs k = 16 s + 8 h + j that
// the same lane supplies as the B operand of K-step s = 2 mt + r/8 of the
// next layer.  Activations therefore stay in registers from layer to layer;
// the view mean is 1-3 lane shuffles; only the packed A fragments (45 KB) and
// biases live in LDS.
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// 4 feature layers, 2 hidden occupancy layers, the 2 geometry-feature layers
constexpr int kMfmaLayers = 8;
__host__ __device__ constexpr int mfma_steps(int l) { return l == 0 || l == 6 ? 2 : l == 7 ? 1 : 4; }
__host__ __device__ constexpr int mfma_mts(int l) { return l >= 6 ? 1 : 2; }
__host__ __device__ constexpr int mfma_frag_base(int l) {  // in f16x8 units
  return l <= 6 ? (l == 0 ? 0 : 256 + (l - 1) * 512) : 2816 + 128;
}
constexpr int kMfmaFrags = 2816 + 128 + 64;  // 3008 f16x8 = 48,128 B
// parameter offsets (halves) of the layers' weights, input and output widths
__host__ __device__ constexpr int mfma_w_off(int l) {
  return l == 0 ? 0
       : l < 4  ? 1216 + (l - 1) * 4160
       : l < 6  ? kFeatParams + (l - 4) * 4160
       : l == 6 ? kGeoOffset
                : kGeoOffset + 18 * 16 + 16;
}
__host__ __device__ constexpr int mfma_k_in(int l) { return l == 0 || l == 6 ? 18 : l == 7 ? 16 : 64; }
__host__ __device__ constexpr int mfma_n_out(int l) { return l == 6 ? 16 : l == 7 ? 8 : 64; }

__host__ __device__ constexpr int perm_unit(int mt, int m) {
  return 32 * mt + 16 * (m / 16) + 8 * ((m / 4) % 2) + 4 * ((m / 8) % 2) + m % 4;
}
__device__ __forceinline__ constexpr int acc_unit(int mt, int r, int h) {
  return 32 * mt + 16 * (r / 8) + 8 * h + r % 8;
}

struct DensityMfmaLds {
  f16x8 frag[kMfmaFrags];
  float bias[kMfmaLayers][64];
  float w_out[64];
  float b_out;
};

// acc = bias + W x for one layer (STEPS K-steps of 16) on this wave's 32 columns
// `zo` is an opaque zero redefined every tile: it keeps the bias reads in
// the tile loop, so the registers go to the weight fragments, which the
// compiler keeps resident across tiles
template <int STEPS>
__device__ __forceinline__ void mfma_layer(const DensityMfmaLds& L, int layer,
                                           const f16x8 (&bin)[4], f32x16 (&acc)[2], int lane,
                                           int zo) {
  const int h = lane >> 5;
#pragma unroll
  for (int mt = 0; mt < 2; ++mt)
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[mt][r] = L.bias[layer][acc_unit(mt, r, h) + zo];
  const f16x8* fr = L.frag + mfma_frag_base(layer);
#pragma unroll
  for (int s = 0; s < STEPS; ++s)
#pragma unroll
    for (int mt = 0; mt < 2; ++mt)
      acc[mt] = __builtin_amdgcn_mfma_f32_32x32x16_f16(fr[(mt * STEPS + s) * 64 + lane], bin[s],
                                                       acc[mt], 0, 0, 0);
}

// B operands of the next layer: relu(acc), rounded to half
__device__ __forceinline__ void relu_to_b(const f32x16 (&acc)[2], f16x8 (&b)[4]) {
#pragma unroll
  for (int s = 0; s < 4; ++s)
#pragma unroll
    for (int j = 0; j < 8; ++j) b[s][j] = (half_t)fmaxf(acc[s / 2][8 * (s % 2) + j], 0.0f);
}

// one M-tile (the geometry layers: <= 32 outputs)
template <int STEPS>
__device__ __forceinline__ void mfma_layer1(const DensityMfmaLds& L, int layer,
                                            const f16x8 (&bin)[4], f32x16& acc, int lane,
                                            int zo) {
  const int h = lane >> 5;
#pragma unroll
  for (int r = 0; r < 16; ++r) acc[r] = L.bias[layer][acc_unit(0, r, h) + zo];
  const f16x8* fr = L.frag + mfma_frag_base(layer);
#pragma unroll
  for (int s = 0; s < STEPS; ++s)
    acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(fr[s * 64 + lane], bin[s], acc, 0, 0, 0);
}

template <int VS>
__global__ void __launch_bounds__(kBlock)
density_mfma_kernel(const half_t* __restrict__ params, const float* __restrict__ z,
                    const float* __restrict__ tpsdf, const float* __restrict__ geo, int B,
                    long long n_pts, int V, float* __restrict__ occ,
                    float* __restrict__ geo_out, int n_geo_out) {
  __shared__ DensityMfmaLds L;
  // ---- stage the permuted A fragments, biases and the small weights
  for (int i = threadIdx.x; i < kMfmaFrags * 8; i += kBlock) {
    const int f = i >> 3, j = i & 7;
    int layer = kMfmaLayers - 1;
    while (f < mfma_frag_base(layer)) --layer;
    const int steps = mfma_steps(layer), local = f - mfma_frag_base(layer);
    const int ln = local & 63, s = (local >> 6) % steps, mt = (local >> 6) / steps;
    const int u = perm_unit(mt, ln & 31), k = 16 * s + 8 * (ln >> 5) + j;
    const int kin = mfma_k_in(layer);
    L.frag[f][j] = (k < kin && u < mfma_n_out(layer)) ? params[mfma_w_off(layer) + u * kin + k]
                                                      : (half_t)0.0f;
  }
  for (int i = threadIdx.x; i < kMfmaLayers * 64; i += kBlock) {
    const int layer = i / 64, u = i % 64, no = mfma_n_out(layer);
    L.bias[layer][u] =
        u < no ? (float)params[mfma_w_off(layer) + no * mfma_k_in(layer) + u] : 0.0f;
  }
  for (int i = threadIdx.x; i < 64; i += kBlock) L.w_out[i] = (float)params[kGeoOffset - 65 + i];
  if (threadIdx.x == 0) L.b_out = (float)params[kGeoOffset - 1];
  __syncthreads();

  constexpr int P = 32 / VS;
  const int lane = threadIdx.x & 63, h = lane >> 5, n = lane & 31;
  const int v = n % VS;
  const long long total = (long long)B * n_pts;
  const long long tiles = (total + P - 1) / P;
  const long long wave0 = (long long)blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
  const long long n_waves = (long long)gridDim.x * (kBlock / 64);
  const float inv = 1.0f / (float)V;
  // this lane's share of a column's inputs x = [z, tpsdf, geo 0..15]:
  // half 0 holds x 0..7 and x 16..17, half 1 holds x 8..15 (its layer-0 B
  // operand); the next tile's share is loaded before the current tile's
  // MFMA chain runs (one wave per SIMD: nothing else hides the latency)
  auto load_share = [&](long long tile, float (&x)[10]) {
#pragma unroll
    for (int i = 0; i < 10; ++i) x[i] = 0.0f;
    const long long p = tile * P + n / VS;
    if (tile < tiles && p < total && v < V) {
      const long long b = p / n_pts, q = p - b * n_pts;
      const long long e = (b * V + v) * n_pts + q;
      const float* g = geo + e * 16;
      if (h == 0) {
        x[0] = z[e];
        x[1] = tpsdf[e];
#pragma unroll
        for (int i = 0; i < 6; ++i) x[2 + i] = g[i];
        x[8] = g[14];
        x[9] = g[15];
      } else {
#pragma unroll
        for (int i = 0; i < 8; ++i) x[i] = g[6 + i];
      }
    }
  };
  float xc[10];
  load_share(wave0, xc);
  for (long long tile = wave0; tile < tiles; tile += n_waves) {
    float xn[10];
    load_share(tile + n_waves, xn);
    const long long p = tile * P + n / VS;
    const bool col_ok = p < total && v < V;
    const long long b = p / n_pts, q = p - b * n_pts;
    f16x8 bin[4];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      bin[0][j] = (half_t)xc[j];
      bin[1][j] = (half_t)(h == 0 && j < 2 ? xc[8 + j] : 0.0f);
      bin[2][j] = (half_t)0.0f;
      bin[3][j] = (half_t)0.0f;
    }
    int zo;
    asm volatile("v_mov_b32 %0, 0" : "=v"(zo));
    f16x8 bin0[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) bin0[t] = bin[t];
    f32x16 acc[2];
    mfma_layer<2>(L, 0, bin, acc, lane, zo);
#pragma unroll
    for (int layer = 1; layer < 4; ++layer) {
      relu_to_b(acc, bin);
      mfma_layer<4>(L, layer, bin, acc, lane, zo);
    }
    // mean of the ReLU outputs over the point's views (fp32)
    const float cm = col_ok ? 1.0f : 0.0f;
#pragma unroll
    for (int mt = 0; mt < 2; ++mt)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        float t = fmaxf(acc[mt][r], 0.0f) * cm;
#pragma unroll
        for (int off = 1; off < VS; off <<= 1) t += __shfl_xor(t, off, 64);
        acc[mt][r] = t * inv;
      }
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
      for (int j = 0; j < 8; ++j) bin[s][j] = (half_t)acc[s / 2][8 * (s % 2) + j];
    mfma_layer<4>(L, 4, bin, acc, lane, zo);
    relu_to_b(acc, bin);
    mfma_layer<4>(L, 5, bin, acc, lane, zo);
    float part = 0.0f;
#pragma unroll
    for (int mt = 0; mt < 2; ++mt)
#pragma unroll
      for (int r = 0; r < 16; ++r) part += L.w_out[acc_unit(mt, r, h)] * fmaxf(acc[mt][r], 0.0f);
    part += __shfl_xor(part, 32, 64);
    if (h == 0 && v == 0 && p < total) occ[p] = 1.0f / (1.0f + expf(-(part + L.b_out)));
    // geometry features of the texture views: 18 -> 16 -> 8 on MFMA from the
    // layer-0 operand; hidden unit 8 h + r sits in accumulator r < 8 of half
    // h, which is the next layer's B operand, and output r < 8 lands in
    // accumulator r of half 0.  The whole wave runs the MFMAs (an MFMA reads
    // the A and B registers of every lane, so a lane-dependent branch around
    // it would feed it stale rows); only the store is per lane.
    if (geo_out) {
      f32x16 g;
      mfma_layer1<2>(L, 6, bin0, g, lane, zo);
      f16x8 gb[4];
#pragma unroll
      for (int j = 0; j < 8; ++j) gb[0][j] = (half_t)fmaxf(g[j], 0.0f);
      mfma_layer1<1>(L, 7, gb, g, lane, zo);
      if (h == 0 && col_ok && v < n_geo_out) {
        float4* o4 =
            reinterpret_cast<float4*>(geo_out + ((b * n_geo_out + v) * n_pts + q) * 8);
        o4[0] = make_float4(fmaxf(g[0], 0.f), fmaxf(g[1], 0.f), fmaxf(g[2], 0.f),
                            fmaxf(g[3], 0.f));
        o4[1] = make_float4(fmaxf(g[4], 0.f), fmaxf(g[5], 0.f), fmaxf(g[6], 0.f),
                            fmaxf(g[7], 0.f));
      }
    }
#pragma unroll
    for (int i = 0; i < 10; ++i) xc[i] = xn[i];
  }
}

// One lane per (ray, texture view): composite the view's 30 features
// [rgb feats 16 | rgb 3 | geometry features 8 | direction 3] with the
// occupancies (volume_rendering_occ_kernel, infer_mlp.cu:508-581), then the
// Hydra stacks over the ray's views and the RGB head (hydra_attention_kernel,
// :243-445).  Lanes 4r .. 4r+3 of a wave hold ray r's views.
__global__ void __launch_bounds__(kBlock)
color_kernel(const half_t* __restrict__ params, const float* __restrict__ t,
             const float* __restrict__ occ, const float* __restrict__ feats_rgb,
             const float* __restrict__ rgbs, const float* __restrict__ geo8,
             const float* __restrict__ dirs, int B, int R, int S, int Vt, float t_thres,
             float* __restrict__ out_depth, float* __restrict__ out_rgb,
             float* __restrict__ out_feats) {
  extern __shared__ __attribute__((aligned(16))) char smem_i[];
  float* row = reinterpret_cast<float*>(smem_i) + threadIdx.x * kRow;
  half_t* w = reinterpret_cast<half_t*>(smem_i + rows_bytes());
  stage(w, params, kColorParams);
  const long long n_rays = (long long)B * R;
  const long long lanes = n_rays * kTexSlots;
  // whole waves iterate together (shuffles need every lane of a ray)
  for (long long base = (long long)blockIdx.x * blockDim.x; base < lanes;
       base += (long long)gridDim.x * blockDim.x) {
    const long long id = base + threadIdx.x;
    const long long ray = id / kTexSlots;
    const int v = (int)(id % kTexSlots);
    const bool live = ray < n_rays && v < Vt;
    float x[30];
#pragma unroll
    for (int c = 0; c < 30; ++c) x[c] = 0.0f;
    float depth = 0.0f;
    if (live) {
      const long long b = ray / R, r = ray - b * R;
      const float* tr = t + ray * S;
      const float* oc = occ + ray * S;
      if (!(tr[S - 1] < 0.0f)) {
        const long long e0 = ((b * Vt + v) * R + r) * S;
        float T = 1.0f;
        for (int s = 0; s < S; ++s) {
          const float ts = tr[s];
          if (ts < 0.0f) continue;
          const float a = oc[s];
          const float wgt = a * T;
          const long long e = e0 + s;
          const float4* f4 = reinterpret_cast<const float4*>(feats_rgb + e * 16);
#pragma unroll
          for (int c = 0; c < 4; ++c) {
            const float4 f = f4[c];
            x[4 * c + 0] += wgt * f.x;
            x[4 * c + 1] += wgt * f.y;
            x[4 * c + 2] += wgt * f.z;
            x[4 * c + 3] += wgt * f.w;
          }
#pragma unroll
          for (int c = 0; c < 3; ++c) x[16 + c] += wgt * rgbs[e * 3 + c];
          const float4* g4 = reinterpret_cast<const float4*>(geo8 + e * 8);
          const float4 ga = g4[0], gb = g4[1];
          x[19] += wgt * ga.x; x[20] += wgt * ga.y; x[21] += wgt * ga.z; x[22] += wgt * ga.w;
          x[23] += wgt * gb.x; x[24] += wgt * gb.y; x[25] += wgt * gb.z; x[26] += wgt * gb.w;
#pragma unroll
          for (int c = 0; c < 3; ++c) x[27 + c] += wgt * dirs[e * 3 + c];
          depth += wgt * ts;
          T *= 1.0f - a;
          if (T < t_thres) break;
        }
      }
    }
    // ---- Hydra stacks (all lanes take part in the shuffles) --------------
    for (int st = 0; st < 2; ++st) {
      const half_t* ws = w + st * kStackParams;
      float q[32], k[32], vv[32];
      dense_row<30, 32, 0, false>(ws, x, row);
      from_row(row, q);
      dense_row<30, 32, 0, false>(ws + 960, x, row);
      from_row(row, k);
      dense_row<30, 32, 0, false>(ws + 1920, x, row);
      from_row(row, vv);
      float qq = 0.0f, kk = 0.0f;
#pragma unroll
      for (int i = 0; i < 32; ++i) {
        qq += q[i] * q[i];
        kk += k[i] * k[i];
      }
      // no epsilon, as the reference (a zero token gives 0/0)
      const float rq = 1.0f / sqrtf(qq), rk = 1.0f / sqrtf(kk);
#pragma unroll
      for (int i = 0; i < 32; ++i) {
        float kv = live ? k[i] * rk * vv[i] : 0.0f;
        kv += __shfl_xor(kv, 1, 64);
        kv += __shfl_xor(kv, 2, 64);
        q[i] = q[i] * rq * kv;
      }
      float o[30], res[30], h[32];
      dense_row<32, 30, 0, false>(ws + 2880, q, row);
      from_row(row, o);
#pragma unroll
      for (int c = 0; c < 30; ++c) res[c] = x[c] + o[c];
      dense_row<30, 32, 1>(ws + 3840, res, row);
      from_row(row, h);
      dense_row<32, 30, 0>(ws + 3840 + dense_size<30, 32>(), h, row);
      from_row(row, x);
#pragma unroll
      for (int c = 0; c < 30; ++c) x[c] += res[c];
    }
    if (live && v == 0) {
      float h0[64], h1[32], rgb[3];
      const half_t* wr = w + kRgbOffset;
      dense_row<30, 64, 1>(wr, x, row);
      from_row(row, h0);
      dense_row<64, 32, 1>(wr + dense_size<30, 64>(), h0, row);
      from_row(row, h1);
      dense<32, 3, 2>(wr + dense_size<30, 64>() + dense_size<64, 32>(), h1, rgb);
      out_depth[ray] = depth;
#pragma unroll
      for (int c = 0; c < 3; ++c) out_rgb[ray * 3 + c] = rgb[c];
#pragma unroll
      for (int c = 0; c < 30; ++c) out_feats[ray * 30 + c] = x[c];
    }
  }
}

// ---- Colour path of the fp16 form ------------------------------------------
// color_composite_kernel: one lane per (ray, texture slot) composites the 30
// features exactly as color_kernel does and writes the token, padded to 32
// with zeros, to the workspace (slot 0's lane also writes the depth).
// color_mfma_kernel: Hydra stacks + RGB head with every product on
// v_mfma_f32_16x16x32_f16.  Each wave takes 16 tokens (4 rays x 4 slots): the
// MFMA's N = 16 columns are the tokens, M = 16 output units per tile, K = 32
// inputs per step.  The half parameters are exact in fp16; the layer inputs
// are rounded to fp16 for the products, which accumulate in fp32; the norms
// (no epsilon, as the reference), the cross-view sum, residuals, biases and
// activations stay in fp32.  The parameters are staged once per persistent
// workgroup as zero-padded fp16 rows of 32 inputs (64 for the RGB head's
// second layer) whose 16-byte chunks are XOR-swizzled by the row, so that the
// 8 lanes of an LDS pass (8 rows, one chunk) hit 8 different bank groups.
typedef float f32x4 __attribute__((ext_vector_type(4)));
constexpr int kTokW = 32;                  // token width in the workspace
constexpr int kCxLd = kTokW + 4;           // fp32 token rows in LDS
constexpr int kChLd = 64 + 4;              // fp32 rows of attention / hidden units
constexpr int kStackRows = 6 * 32;         // q | k | v | fc | w1 | w2 rows, 32 each

__global__ void __launch_bounds__(kBlock)
color_composite_kernel(const float* __restrict__ t, const float* __restrict__ occ,
                       const float* __restrict__ feats_rgb, const float* __restrict__ rgbs,
                       const float* __restrict__ geo8, const float* __restrict__ dirs, int B,
                       int R, int S, int Vt, float t_thres, float* __restrict__ out_depth,
                       float* __restrict__ tok) {
  const long long lanes = (long long)B * R * kTexSlots;
  for (long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x; id < lanes;
       id += (long long)gridDim.x * blockDim.x) {
    const long long ray = id / kTexSlots;
    const int v = (int)(id % kTexSlots);
    float x[kTokW];
#pragma unroll
    for (int c = 0; c < kTokW; ++c) x[c] = 0.0f;
    float depth = 0.0f;
    if (v < Vt) {
      const long long b = ray / R, r = ray - b * R;
      const float* tr = t + ray * S;
      const float* oc = occ + ray * S;
      if (!(tr[S - 1] < 0.0f)) {
        const long long e0 = ((b * Vt + v) * R + r) * S;
        float T = 1.0f;
        for (int s = 0; s < S; ++s) {
          const float ts = tr[s];
          if (ts < 0.0f) continue;
          const float a = oc[s];
          const float wgt = a * T;
          const long long e = e0 + s;
          const float4* f4 = reinterpret_cast<const float4*>(feats_rgb + e * 16);
#pragma unroll
          for (int c = 0; c < 4; ++c) {
            const float4 f = f4[c];
            x[4 * c + 0] += wgt * f.x;
            x[4 * c + 1] += wgt * f.y;
            x[4 * c + 2] += wgt * f.z;
            x[4 * c + 3] += wgt * f.w;
          }
#pragma unroll
          for (int c = 0; c < 3; ++c) x[16 + c] += wgt * rgbs[e * 3 + c];
          const float4* g4 = reinterpret_cast<const float4*>(geo8 + e * 8);
          const float4 ga = g4[0], gb = g4[1];
          x[19] += wgt * ga.x; x[20] += wgt * ga.y; x[21] += wgt * ga.z; x[22] += wgt * ga.w;
          x[23] += wgt * gb.x; x[24] += wgt * gb.y; x[25] += wgt * gb.z; x[26] += wgt * gb.w;
#pragma unroll
          for (int c = 0; c < 3; ++c) x[27 + c] += wgt * dirs[e * 3 + c];
          depth += wgt * ts;
          T *= 1.0f - a;
          if (T < t_thres) break;
        }
      }
    }
    float4* o4 = reinterpret_cast<float4*>(tok + id * kTokW);
#pragma unroll
    for (int c = 0; c < kTokW / 4; ++c)
      o4[c] = make_float4(x[4 * c], x[4 * c + 1], x[4 * c + 2], x[4 * c + 3]);
    if (v == 0) out_depth[ray] = depth;
  }
}

struct ColorMfmaLds {
  f16x8 wst[2][kStackRows][4];   // per stack, rows of 32 inputs (4 chunks)
  f16x8 wr1[64][4];              // RGB layer 1 [64][30 + 2]
  f16x8 wr2[32][8];              // RGB layer 2 [32][64]
  f16x8 wr3[16][4];              // RGB layer 3 [3 + 13][32]
  float b1[2][32], b2[2][32], rb1[64], rb2[32], rb3[4];
  float sx[kBlock / 64][16 * kCxLd];   // per wave: token rows
  float sh[kBlock / 64][16 * kChLd];   // per wave: attention / hidden rows
};

// position of chunk c of row r among a row's NCH chunks
template <int NCH>
__device__ __forceinline__ int cchunk(int r, int c) {
  return NCH == 4 ? (c ^ ((r >> 1) & 3)) : (c ^ (r & 7));
}

// acc[mt] = sum_k W[16 mt + n][k] x[token n][k], K = 32 KS; W: LDS rows of
// 4 KS swizzled f16x8 chunks; x: fp32 rows of pitch LD (rounded to fp16).
template <int MT, int KS, int LD>
__device__ __forceinline__ void cmm(const f16x8* W, const float* x, int lane, f32x4 (&acc)[MT]) {
  const int n = lane & 15, g = lane >> 4;
#pragma unroll
  for (int mt = 0; mt < MT; ++mt) acc[mt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int s = 0; s < KS; ++s) {
    const f32x4 x0 = *reinterpret_cast<const f32x4*>(x + n * LD + 32 * s + 8 * g);
    const f32x4 x1 = *reinterpret_cast<const f32x4*>(x + n * LD + 32 * s + 8 * g + 4);
    f16x8 b;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      b[j] = (half_t)x0[j];
      b[4 + j] = (half_t)x1[j];
    }
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) {
      const int r = 16 * mt + n;
      const f16x8 a = W[r * 4 * KS + cchunk<4 * KS>(r, 4 * s + g)];
      acc[mt] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, acc[mt], 0, 0, 0);
    }
  }
}

__global__ void __launch_bounds__(kBlock)
color_mfma_kernel(const half_t* __restrict__ params, const float* __restrict__ tok, int B, int R,
                  int Vt, float* __restrict__ out_rgb, float* __restrict__ out_feats) {
  __shared__ ColorMfmaLds L;
  // ---- stage the parameters (zero-padded, swizzled) -----------------------
  {
    half_t* wst = reinterpret_cast<half_t*>(&L.wst[0][0][0]);
    for (int i = threadIdx.x; i < 2 * kStackRows * 32; i += blockDim.x) {
      const int st = i / (kStackRows * 32), r = (i / 32) % kStackRows, c = i % 32;
      const half_t* ps = params + st * kStackParams;
      const int blk = r / 32, o = r % 32;
      float v = 0.0f;
      if (blk < 3) v = c < 30 ? (float)ps[blk * 960 + o * 30 + c] : 0.0f;      // q, k, v
      else if (blk == 3) v = o < 30 ? (float)ps[2880 + o * 32 + c] : 0.0f;    // fc
      else if (blk == 4) v = c < 30 ? (float)ps[3840 + o * 30 + c] : 0.0f;    // w1
      else v = o < 30 ? (float)ps[4832 + o * 32 + c] : 0.0f;                  // w2
      wst[((size_t)st * kStackRows + r) * 32 + cchunk<4>(r, c / 8) * 8 + c % 8] = (half_t)v;
    }
    half_t* w1 = reinterpret_cast<half_t*>(&L.wr1[0][0]);
    for (int i = threadIdx.x; i < 64 * 32; i += blockDim.x) {
      const int r = i / 32, c = i % 32;
      const float v = c < 30 ? (float)params[kRgbOffset + r * 30 + c] : 0.0f;
      w1[r * 32 + cchunk<4>(r, c / 8) * 8 + c % 8] = (half_t)v;
    }
    half_t* w2 = reinterpret_cast<half_t*>(&L.wr2[0][0]);
    for (int i = threadIdx.x; i < 32 * 64; i += blockDim.x) {
      const int r = i / 64, c = i % 64;
      w2[r * 64 + cchunk<8>(r, c / 8) * 8 + c % 8] = params[kRgbOffset + 1984 + r * 64 + c];
    }
    half_t* w3 = reinterpret_cast<half_t*>(&L.wr3[0][0]);
    for (int i = threadIdx.x; i < 16 * 32; i += blockDim.x) {
      const int r = i / 32, c = i % 32;
      const float v = r < 3 ? (float)params[kRgbOffset + 4064 + r * 32 + c] : 0.0f;
      w3[r * 32 + cchunk<4>(r, c / 8) * 8 + c % 8] = (half_t)v;
    }
    for (int i = threadIdx.x; i < 64; i += blockDim.x) {
      const int st = i / 32, o = i % 32;
      const half_t* ps = params + st * kStackParams;
      L.b1[st][o] = (float)ps[4800 + o];
      L.b2[st][o] = o < 30 ? (float)ps[5792 + o] : 0.0f;
    }
    for (int i = threadIdx.x; i < 64; i += blockDim.x) L.rb1[i] = (float)params[kRgbOffset + 1920 + i];
    for (int i = threadIdx.x; i < 32; i += blockDim.x) L.rb2[i] = (float)params[kRgbOffset + 4032 + i];
    for (int i = threadIdx.x; i < 4; i += blockDim.x)
      L.rb3[i] = i < 3 ? (float)params[kRgbOffset + 4160 + i] : 0.0f;
    __syncthreads();
  }
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int n = lane & 15, g = lane >> 4;
  float* xs = L.sx[wave];
  float* hs = L.sh[wave];
  const long long n_rays = (long long)B * R;
  const long long n_tiles = (n_rays * kTexSlots + 15) / 16;
  const int slot = n % kTexSlots;
  // the workgroup's 4 waves take 4 consecutive tiles per step (a wave past
  // the last tile runs on zero tokens and writes nothing), so every wave
  // reaches every barrier
  for (long long grp = blockIdx.x; grp * (kBlock / 64) < n_tiles; grp += gridDim.x) {
    const long long t0 = (grp * (kBlock / 64) + wave) * 16;
    const long long ray = (t0 + n) / kTexSlots;
    const bool live = ray < n_rays && slot < Vt;
    for (int i = lane; i < 16 * kTokW / 4; i += 64) {
      const int tk = i / (kTokW / 4), c = 4 * (i % (kTokW / 4));
      f32x4 v = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
      if (t0 + tk < n_rays * kTexSlots) v = *reinterpret_cast<const f32x4*>(tok + (t0 + tk) * kTokW + c);
      *reinterpret_cast<f32x4*>(xs + tk * kCxLd + c) = v;
    }
    __syncthreads();
    for (int st = 0; st < 2; ++st) {
      const f16x8* W = &L.wst[st][0][0];
      f32x4 a6[6];
      cmm<6, 1, kCxLd>(W, xs, lane, a6);
      float qq = 0.0f, kk = 0.0f;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        qq += a6[0][i] * a6[0][i] + a6[1][i] * a6[1][i];
        kk += a6[2][i] * a6[2][i] + a6[3][i] * a6[3][i];
      }
      qq += __shfl_xor(qq, 16, 64);
      qq += __shfl_xor(qq, 32, 64);
      kk += __shfl_xor(kk, 16, 64);
      kk += __shfl_xor(kk, 32, 64);
      // no epsilon, as the reference (a zero token gives 0/0)
      const float rq = 1.0f / sqrtf(qq), rk = 1.0f / sqrtf(kk);
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          float kv = live ? a6[2 + h][i] * rk * a6[4 + h][i] : 0.0f;
          kv += __shfl_xor(kv, 1, 64);
          kv += __shfl_xor(kv, 2, 64);
          hs[n * kChLd + 16 * h + 4 * g + i] = a6[h][i] * rq * kv;
        }
      __syncthreads();
      {
        f32x4 ao[2];   // res = x + fc(att) (units 30, 31 stay 0)
        cmm<2, 1, kChLd>(W + 3 * 32 * 4, hs, lane, ao);
        __syncthreads();
#pragma unroll
        for (int mt = 0; mt < 2; ++mt)
#pragma unroll
          for (int i = 0; i < 4; ++i) xs[n * kCxLd + 16 * mt + 4 * g + i] += ao[mt][i];
      }
      __syncthreads();
      {
        f32x4 ah[2];   // h = relu(W1 res + b1)
        cmm<2, 1, kCxLd>(W + 4 * 32 * 4, xs, lane, ah);
#pragma unroll
        for (int mt = 0; mt < 2; ++mt)
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const int u = 16 * mt + 4 * g + i;
            hs[n * kChLd + u] = fmaxf(ah[mt][i] + L.b1[st][u], 0.0f);
          }
      }
      __syncthreads();
      {
        f32x4 ax[2];   // x = W2 h + b2 + res
        cmm<2, 1, kChLd>(W + 5 * 32 * 4, hs, lane, ax);
#pragma unroll
        for (int mt = 0; mt < 2; ++mt)
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const int u = 16 * mt + 4 * g + i;
            xs[n * kCxLd + u] = (ax[mt][i] + L.b2[st][u]) + xs[n * kCxLd + u];
          }
      }
      __syncthreads();
    }
    // ---- RGB head: 30 -> 64 -> 32 -> 3 (used on slot 0's token) ------------
    {
      f32x4 a1[4];
      cmm<4, 1, kCxLd>(&L.wr1[0][0], xs, lane, a1);
#pragma unroll
      for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int u = 16 * mt + 4 * g + i;
          hs[n * kChLd + u] = fmaxf(a1[mt][i] + L.rb1[u], 0.0f);
        }
    }
    __syncthreads();
    {
      f32x4 a2[2];
      cmm<2, 2, kChLd>(&L.wr2[0][0], hs, lane, a2);
      __syncthreads();
#pragma unroll
      for (int mt = 0; mt < 2; ++mt)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int u = 16 * mt + 4 * g + i;
          hs[n * kChLd + u] = fmaxf(a2[mt][i] + L.rb2[u], 0.0f);
        }
    }
    __syncthreads();
    {
      f32x4 a3[1];
      cmm<1, 1, kChLd>(&L.wr3[0][0], hs, lane, a3);
      // lane (n, g = 0) holds the three outputs of token n
      if (g == 0 && slot == 0 && ray < n_rays) {
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          const float z = a3[0][j] + L.rb3[j];
          out_rgb[ray * 3 + j] = 1.0f / (1.0f + expf(-z));
        }
#pragma unroll
        for (int c = 0; c < 30; ++c) out_feats[ray * 30 + c] = xs[n * kCxLd + c];
      }
    }
    __syncthreads();
  }
}

// upsampling_kernel (infer_mlp.cu:449-506) for one point: [40] neighbour
// features and [30] ray features (read from rf) -> 3 blend weights.
__device__ __forceinline__ void blend_mlp(const half_t* __restrict__ w, const float (&x)[40],
                                          const float* __restrict__ rf, float* row,
                                          float (&o)[3]) {
  float h[64], c[62], h2[32];
  const half_t* wp = w;
  dense_row<40, 64, 1>(wp, x, row);
  from_row(row, h);
  wp += dense_size<40, 64>();
  dense_row<64, 64, 1>(wp, h, row);
  from_row(row, h);
  wp += dense_size<64, 64>();
  dense_row<64, 32, 1>(wp, h, row);
  from_row(row, h2);
  wp += dense_size<64, 32>();
#pragma unroll
  for (int i = 0; i < 32; ++i) c[i] = h2[i];
#pragma unroll
  for (int i = 0; i < 30; ++i) c[32 + i] = rf[i];
  dense_row<62, 64, 1>(wp, c, row);
  from_row(row, h);
  wp += dense_size<62, 64>();
  dense_row<64, 32, 1>(wp, h, row);
  from_row(row, h2);
  wp += dense_size<64, 32>();
  dense<32, 3, 2>(wp, h2, o);
}

// [N, 40] neighbour features and [N, 30] ray features -> [N, 3] blend weights.
__global__ void __launch_bounds__(kBlock)
blend_kernel(const half_t* __restrict__ params, const float* __restrict__ nf,
             const float* __restrict__ rf, long long n, float* __restrict__ out) {
  extern __shared__ __attribute__((aligned(16))) char smem_i[];
  float* row = reinterpret_cast<float*>(smem_i) + threadIdx.x * kRow;
  half_t* w = reinterpret_cast<half_t*>(smem_i + rows_bytes());
  stage(w, params, kBlendParams);
  for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < n;
       p += (long long)gridDim.x * blockDim.x) {
    float x[40], o[3];
#pragma unroll
    for (int i = 0; i < 40; ++i) x[i] = nf[p * 40 + i];
    blend_mlp(w, x, rf + p * 30, row, o);
#pragma unroll
    for (int i = 0; i < 3; ++i) out[p * 3 + i] = o[i];
  }
}

// ---------------------------------------------------------------------------
// Sub-pixel upsampling of the GUI (gui/gui_render.py:753-838): 4 rays per
// pixel re-derived from the target camera, each blended from 2 neighbour
// views and the coarse colour.

// Inverse of a row-major 3x3 matrix by the adjugate.
__device__ __forceinline__ void inv3(const float* __restrict__ m, float (&r)[9]) {
  const float c00 = m[4] * m[8] - m[5] * m[7];
  const float c01 = m[5] * m[6] - m[3] * m[8];
  const float c02 = m[3] * m[7] - m[4] * m[6];
  const float inv_det = 1.0f / (m[0] * c00 + m[1] * c01 + m[2] * c02);
  r[0] = c00 * inv_det;
  r[1] = (m[2] * m[7] - m[1] * m[8]) * inv_det;
  r[2] = (m[1] * m[5] - m[2] * m[4]) * inv_det;
  r[3] = c01 * inv_det;
  r[4] = (m[0] * m[8] - m[2] * m[6]) * inv_det;
  r[5] = (m[2] * m[3] - m[0] * m[5]) * inv_det;
  r[6] = c02 * inv_det;
  r[7] = (m[1] * m[6] - m[0] * m[7]) * inv_det;
  r[8] = (m[0] * m[4] - m[1] * m[3]) * inv_det;
}

__device__ __forceinline__ void mat3_vec(const float* __restrict__ m, const float (&v)[3],
                                         float (&o)[3]) {
  o[0] = m[0] * v[0] + m[1] * v[1] + m[2] * v[2];
  o[1] = m[3] * v[0] + m[4] * v[1] + m[5] * v[2];
  o[2] = m[6] * v[0] + m[7] * v[1] + m[8] * v[2];
}

// rays_interpolate (gui_render.py:753-773) for sub-pixel k of direction d:
// the pixel (x, y) of K R d, shifted by (+0.5 if k & 1, +0.5 if k & 2),
// back-projected as R^-1 K^-1 [x, y, 1] and normalised (F.normalize, eps
// 1e-12).  Kinv / Rinv are the inverses of K / R.
__device__ __forceinline__ void interp_dir(const float* __restrict__ K,
                                           const float* __restrict__ R, const float (&Kinv)[9],
                                           const float (&Rinv)[9], const float (&d)[3], int k,
                                           float (&out)[3]) {
  float rd[3], q[3];
  mat3_vec(R, d, rd);
  mat3_vec(K, rd, q);
  float px[3] = {q[0] / q[2], q[1] / q[2], 1.0f};
  if (k & 1) px[0] += 0.5f;
  if (k & 2) px[1] += 0.5f;
  float u[3], v[3];
  mat3_vec(Kinv, px, u);
  mat3_vec(Rinv, u, v);
  const float n = fmaxf(sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]), 1e-12f);
  out[0] = v[0] / n;
  out[1] = v[1] / n;
  out[2] = v[2] / n;
}

// dirs [cams][n][3] -> out [cams][n * 4][3] (element (ray, k) at ray * 4 + k).
__global__ void __launch_bounds__(kBlock)
rays_interp_kernel(const float* __restrict__ K, const float* __restrict__ R,
                   const float* __restrict__ dirs, int cams, int n, float* __restrict__ out) {
  const long long total = (long long)cams * n * 4;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const long long ray = e >> 2;
    const int cam = (int)(ray / n);
    float Kinv[9], Rinv[9], d[3], o[3];
    inv3(K + cam * 9, Kinv);
    inv3(R + cam * 9, Rinv);
#pragma unroll
    for (int i = 0; i < 3; ++i) d[i] = dirs[ray * 3 + i];
    interp_dir(K + cam * 9, R + cam * 9, Kinv, Rinv, d, (int)(e & 3), o);
#pragma unroll
    for (int i = 0; i < 3; ++i) out[e * 3 + i] = o[i];
  }
}

struct UpsampleArgs {
  const float* ray_ori;    // [N,3]
  const float* ray_dir;    // [N,3]
  const float* depth;      // [N]     rendered depth o_d
  const float* coarse;     // [N,3]   rendered colour o_c
  const float* feats;      // [N,30]  colour features o_c_feat
  const float* tar_K;      // [3,3]
  const float* tar_R;      // [3,3]
  const float* K;          // [V,3,3] source views
  const float* R;          // [V,3,3]
  const float* T;          // [V,3]
  const float* rgbf;       // [V,Hf,Wf,16] high-resolution colour features (channels last)
  const float* rgb;        // [V,Hc,Wc,3]  colours (channels last)
  const float* dep;        // [V,Hd,Wd]    depths
  long long N;
  int Hf, Wf, Hc, Wc, Hd, Wd, res, nei[2];
  float vmin, vmax;        // rays with vmin < o_d < vmax are upsampled
  float* out;              // [N,4,3]
};

// One thread per (ray, sub-pixel): gui_render.forward_upsampling +
// sample_neighbor (:775-838) with the neighbour projection of proj_persp
// (utils_render.py:380-410, depth range [0.1, 4]), F.grid_sample of the
// features (bilinear) and of colour and depth (nearest), the Gaussian depth
// agreement exp(-200 (z - d)^2), the blending MLP and the clamped blend
// w0 c0 + w1 c1 + w2 c_coarse.  Rays outside the depth range get zeros.
__global__ void __launch_bounds__(kBlock) upsample_kernel(const half_t* __restrict__ params,
                                                          UpsampleArgs a) {
  extern __shared__ __attribute__((aligned(16))) char smem_i[];
  float* row = reinterpret_cast<float*>(smem_i) + threadIdx.x * kRow;
  half_t* w = reinterpret_cast<half_t*>(smem_i + rows_bytes());
  stage(w, params, kBlendParams);
  float Kinv[9], Rinv[9];
  inv3(a.tar_K, Kinv);
  inv3(a.tar_R, Rinv);
  const long long total = a.N * 4;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const long long p = e >> 2;
    const float od = a.depth[p];
    float res_c[3] = {0.0f, 0.0f, 0.0f};
    if (od > a.vmin && od < a.vmax) {
      float d[3], de[3];
#pragma unroll
      for (int i = 0; i < 3; ++i) d[i] = a.ray_dir[p * 3 + i];
      interp_dir(a.tar_K, a.tar_R, Kinv, Rinv, d, (int)(e & 3), de);
      float pos[3];
#pragma unroll
      for (int i = 0; i < 3; ++i) pos[i] = a.ray_ori[p * 3 + i] + de[i] * od;
      const bool invalid_d = od < 0.1f || od > 4.0f;
      const float half = (float)(a.res / 2);
      float x[40], c_nei[2][3];
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int v = a.nei[j];
        const float* Rm = a.R + v * 9;
        const float* Km = a.K + v * 9;
        const float* Tv = a.T + v * 3;
        float cam[3];
        mat3_vec(Rm, pos, cam);
#pragma unroll
        for (int i = 0; i < 3; ++i) cam[i] += Tv[i];
        const bool invalid = cam[2] < 0.01f;
        const float zd = invalid ? 1.0f : cam[2];
        const float q[3] = {cam[0] / zd, cam[1] / zd, cam[2] / zd};
        float gx = (Km[0] * q[0] + Km[1] * q[1] + Km[2] * q[2] - half) / half;
        float gy = (Km[3] * q[0] + Km[4] * q[1] + Km[5] * q[2] - half) / half;
        if (invalid || invalid_d) gx = gy = -2.0f;
        const float pz = invalid_d ? -2.0f : cam[2];
        float f[16], sd[1], sdir[3];
        sailor::bilinear_cl<16>(a.rgbf + (size_t)v * a.Hf * a.Wf * 16, a.Hf, a.Wf, gx, gy, f);
        sailor::nearest_cl<3>(a.rgb + (size_t)v * a.Hc * a.Wc * 3, a.Hc, a.Wc, gx, gy, c_nei[j]);
        sailor::nearest_cl<1>(a.dep + (size_t)v * a.Hd * a.Wd, a.Hd, a.Wd, gx, gy, sd);
        mat3_vec(Rm, de, sdir);
        const float dz = pz - sd[0];
#pragma unroll
        for (int i = 0; i < 16; ++i) x[20 * j + i] = f[i];
#pragma unroll
        for (int i = 0; i < 3; ++i) x[20 * j + 16 + i] = sdir[i];
        x[20 * j + 19] = expf(-200.0f * (dz * dz));
      }
      float wts[3];
      blend_mlp(w, x, a.feats + p * 30, row, wts);
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        const float c = wts[0] * c_nei[0][i] + wts[1] * c_nei[1][i] + wts[2] * a.coarse[p * 3 + i];
        // torch.clamp: NaN stays NaN
        res_c[i] = c != c ? c : fminf(fmaxf(c, 0.0f), 1.0f);
      }
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) a.out[e * 3 + i] = res_c[i];
  }
}

// ---------------------------------------------------------------------------
// The same sub-pixel upsampling with the blending MLP on fp16 MFMA (layer
// inputs rounded to half, fp32 accumulation).  A wave takes 32 sub-pixel
// rays (8 rays x 4 sub-pixels); lane half h gathers neighbour view nei[h]
// (bilinear features, nearest colour and depth, the Gaussian weight), the
// halves exchange their shares with 32-lane shuffles to form the layer-0
// operand, and the layers 40 -> 64 -> 64 -> 32, [32 | ray features 30] ->
// 64 -> 32 run as a register chain with the row permutation of the density
// kernel; 32 -> 3 is a dot product and one shuffle.
struct BlendLayer {
  int w_off, k_in, n_out, steps, mts, frag_base;
};
__host__ __device__ constexpr BlendLayer blend_layer(int l) {
  return l == 0   ? BlendLayer{0, 40, 64, 3, 2, 0}
         : l == 1 ? BlendLayer{2624, 64, 64, 4, 2, 384}
         : l == 2 ? BlendLayer{6784, 64, 32, 4, 1, 896}
         : l == 3 ? BlendLayer{8864, 62, 64, 4, 2, 1152}
                  : BlendLayer{12896, 64, 32, 4, 1, 1664};
}
constexpr int kBlendMfmaLayers = 5, kBlendFrags = 1920, kBlendOutOff = 14976;

struct BlendMfmaLds {
  f16x8 frag[kBlendFrags];
  float bias[kBlendMfmaLayers][64];
  float w_out[3][32];
  float b_out[4];
};

template <int STEPS, int MTS>
__device__ __forceinline__ void blend_layer_mfma(const BlendMfmaLds& L, int layer,
                                                 const f16x8 (&bin)[4], f32x16 (&acc)[2],
                                                 int lane, int zo) {
  const int h = lane >> 5;
#pragma unroll
  for (int mt = 0; mt < MTS; ++mt)
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[mt][r] = L.bias[layer][acc_unit(mt, r, h) + zo];
  const f16x8* fr = L.frag + blend_layer(layer).frag_base;
#pragma unroll
  for (int s = 0; s < STEPS; ++s)
#pragma unroll
    for (int mt = 0; mt < MTS; ++mt)
      acc[mt] = __builtin_amdgcn_mfma_f32_32x32x16_f16(fr[(mt * STEPS + s) * 64 + lane], bin[s],
                                                       acc[mt], 0, 0, 0);
}

__global__ void __launch_bounds__(kBlock) upsample_mfma_kernel(const half_t* __restrict__ params,
                                                               UpsampleArgs a) {
  __shared__ BlendMfmaLds L;
  for (int i = threadIdx.x; i < kBlendFrags * 8; i += kBlock) {
    const int f = i >> 3, j = i & 7;
    int layer = kBlendMfmaLayers - 1;
    while (f < blend_layer(layer).frag_base) --layer;
    const BlendLayer d = blend_layer(layer);
    const int local = f - d.frag_base;
    const int ln = local & 63, s = (local >> 6) % d.steps, mt = (local >> 6) / d.steps;
    const int u = perm_unit(mt, ln & 31), k = 16 * s + 8 * (ln >> 5) + j;
    L.frag[f][j] = (k < d.k_in && u < d.n_out) ? params[d.w_off + u * d.k_in + k]
                                               : (half_t)0.0f;
  }
  for (int i = threadIdx.x; i < kBlendMfmaLayers * 64; i += kBlock) {
    const int layer = i / 64, u = i % 64;
    const BlendLayer d = blend_layer(layer);
    L.bias[layer][u] = u < d.n_out ? (float)params[d.w_off + d.n_out * d.k_in + u] : 0.0f;
  }
  for (int i = threadIdx.x; i < 96; i += kBlock) L.w_out[i / 32][i % 32] = (float)params[kBlendOutOff + i];
  if (threadIdx.x < 3) L.b_out[threadIdx.x] = (float)params[kBlendOutOff + 96 + threadIdx.x];
  __syncthreads();

  const int lane = threadIdx.x & 63, h = lane >> 5, n = lane & 31;
  float Kinv[9], Rinv[9];
  inv3(a.tar_K, Kinv);
  inv3(a.tar_R, Rinv);
  const int vview = a.nei[h];
  const float* Rm = a.R + vview * 9;
  const float* Km = a.K + vview * 9;
  const float* Tv = a.T + vview * 3;
  const float* rgbf_v = a.rgbf + (size_t)vview * a.Hf * a.Wf * 16;
  const float* rgb_v = a.rgb + (size_t)vview * a.Hc * a.Wc * 3;
  const float* dep_v = a.dep + (size_t)vview * a.Hd * a.Wd;
  const float half_res = (float)(a.res / 2);
  const long long tiles = (a.N * 4 + 31) / 32;
  const long long wave0 = (long long)blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
  const long long n_waves = (long long)gridDim.x * (kBlock / 64);
  for (long long tile = wave0; tile < tiles; tile += n_waves) {
    const long long e = tile * 32 + n;  // sub-pixel ray (ray e / 4, sub-pixel e % 4)
    const long long p = e >> 2;
    const bool in = p < a.N;
    const float od = in ? a.depth[p] : 0.0f;
    const bool ok = in && od > a.vmin && od < a.vmax;
    if (!__any(ok)) {
      if (h == 0 && in) {
#pragma unroll
        for (int i = 0; i < 3; ++i) a.out[e * 3 + i] = 0.0f;
      }
      continue;
    }
    // this half's neighbour view: 16 features, direction, Gaussian weight
    float f[16], dv[3], coff = 0.0f, col[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int i = 0; i < 16; ++i) f[i] = 0.0f;
    dv[0] = dv[1] = dv[2] = 0.0f;
    if (ok) {
      float d[3], de[3];
#pragma unroll
      for (int i = 0; i < 3; ++i) d[i] = a.ray_dir[p * 3 + i];
      interp_dir(a.tar_K, a.tar_R, Kinv, Rinv, d, (int)(e & 3), de);
      float pos[3];
#pragma unroll
      for (int i = 0; i < 3; ++i) pos[i] = a.ray_ori[p * 3 + i] + de[i] * od;
      const bool invalid_d = od < 0.1f || od > 4.0f;
      float cam[3];
      mat3_vec(Rm, pos, cam);
#pragma unroll
      for (int i = 0; i < 3; ++i) cam[i] += Tv[i];
      const bool invalid = cam[2] < 0.01f;
      const float zd = invalid ? 1.0f : cam[2];
      const float q[3] = {cam[0] / zd, cam[1] / zd, cam[2] / zd};
      float gx = (Km[0] * q[0] + Km[1] * q[1] + Km[2] * q[2] - half_res) / half_res;
      float gy = (Km[3] * q[0] + Km[4] * q[1] + Km[5] * q[2] - half_res) / half_res;
      if (invalid || invalid_d) gx = gy = -2.0f;
      const float pz = invalid_d ? -2.0f : cam[2];
      float sd[1];
      sailor::bilinear_cl<16>(rgbf_v, a.Hf, a.Wf, gx, gy, f);
      sailor::nearest_cl<3>(rgb_v, a.Hc, a.Wc, gx, gy, col);
      sailor::nearest_cl<1>(dep_v, a.Hd, a.Wd, gx, gy, sd);
      mat3_vec(Rm, de, dv);
      const float dz = pz - sd[0];
      coff = expf(-200.0f * (dz * dz));
    }
    // layer-0 operand x = [f(nei0) 16 | dir 3 | w | f(nei1) 16 | dir 3 | w],
    // k = 16 s + 8 h + j: half 0 supplies x 0..7, 16..23, 32..39 and half 1
    // x 8..15, 24..31
    float of[16], od3[3], ocf;
#pragma unroll
    for (int i = 0; i < 16; ++i) of[i] = __shfl_xor(f[i], 32, 64);
#pragma unroll
    for (int i = 0; i < 3; ++i) od3[i] = __shfl_xor(dv[i], 32, 64);
    ocf = __shfl_xor(coff, 32, 64);
    f16x8 bin[4];
    if (h == 0) {
#pragma unroll
      for (int j = 0; j < 8; ++j) bin[0][j] = (half_t)f[j];
      bin[1][0] = (half_t)dv[0];
      bin[1][1] = (half_t)dv[1];
      bin[1][2] = (half_t)dv[2];
      bin[1][3] = (half_t)coff;
#pragma unroll
      for (int j = 0; j < 4; ++j) bin[1][4 + j] = (half_t)of[j];
#pragma unroll
      for (int j = 0; j < 4; ++j) bin[2][j] = (half_t)of[12 + j];
      bin[2][4] = (half_t)od3[0];
      bin[2][5] = (half_t)od3[1];
      bin[2][6] = (half_t)od3[2];
      bin[2][7] = (half_t)ocf;
    } else {
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        bin[0][j] = (half_t)of[8 + j];
        bin[1][j] = (half_t)f[4 + j];
        bin[2][j] = (half_t)0.0f;
      }
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) bin[3][j] = (half_t)0.0f;
    int zo;
    asm volatile("v_mov_b32 %0, 0" : "=v"(zo));
    f32x16 acc[2];
    blend_layer_mfma<3, 2>(L, 0, bin, acc, lane, zo);
    relu_to_b(acc, bin);
    blend_layer_mfma<4, 2>(L, 1, bin, acc, lane, zo);
    relu_to_b(acc, bin);
    blend_layer_mfma<4, 1>(L, 2, bin, acc, lane, zo);
    // [relu(layer 2) 32 | ray colour features 30 | 0 0]
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
      for (int j = 0; j < 8; ++j) bin[s][j] = (half_t)fmaxf(acc[0][8 * s + j], 0.0f);
    {
      const float* rf = a.feats + (in ? p : 0) * 30;
#pragma unroll
      for (int s = 2; s < 4; ++s)
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          const int k = 16 * (s - 2) + 8 * h + j;
          bin[s][j] = (half_t)(k < 30 && in ? rf[k] : 0.0f);
        }
    }
    blend_layer_mfma<4, 2>(L, 3, bin, acc, lane, zo);
    relu_to_b(acc, bin);
    blend_layer_mfma<4, 1>(L, 4, bin, acc, lane, zo);
    float w[3];
#pragma unroll
    for (int o = 0; o < 3; ++o) {
      float part = 0.0f;
#pragma unroll
      for (int r = 0; r < 16; ++r) part += L.w_out[o][acc_unit(0, r, h)] * fmaxf(acc[0][r], 0.0f);
      part += __shfl_xor(part, 32, 64);
      w[o] = 1.0f / (1.0f + expf(-(part + L.b_out[o])));
    }
    float ocol[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) ocol[i] = __shfl_xor(col[i], 32, 64);
    if (h == 0 && in) {
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        float c = 0.0f;
        if (ok) {
          c = w[0] * col[i] + w[1] * ocol[i] + w[2] * a.coarse[p * 3 + i];
          c = c != c ? c : fminf(fmaxf(c, 0.0f), 1.0f);
        }
        a.out[e * 3 + i] = c;
      }
    }
  }
}

// Dynamic LDS of a kernel: per-thread rows + the parameter group (half).
template <typename K>
int lds_setup(K kern, size_t bytes) {
  SAILOR_HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
  return SAILOR_OK;
}
constexpr size_t kDensityLds = rows_bytes() + kDensityParams * sizeof(half_t);
constexpr size_t kColorLds = rows_bytes() + kColorParams * sizeof(half_t);
constexpr size_t kBlendLds = rows_bytes() + kBlendParams * sizeof(half_t);

int setup_all() {
  static std::atomic<bool> done[sailor::kMaxDevices];   // LDS limits set, per device
  int dev = 0;
  if (int e = sailor::device_slot(&dev)) return e;
  if (done[dev].load(std::memory_order_acquire)) return SAILOR_OK;
  if (int e = lds_setup(density_kernel, kDensityLds)) return e;
  if (int e = lds_setup(color_kernel, kColorLds)) return e;
  if (int e = lds_setup(blend_kernel, kBlendLds)) return e;
  if (int e = lds_setup(upsample_kernel, kBlendLds)) return e;
  done[dev].store(true, std::memory_order_release);
  return SAILOR_OK;
}

unsigned grid_for(long long work) {
  int dev = 0, cus = 256;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    cus = 256;
  const long long blocks = (work + kBlock - 1) / kBlock;
  return (unsigned)std::max<long long>(1, std::min<long long>(blocks, (long long)cus * 8));
}

}  // namespace

extern "C" size_t sailor_infer_nerf_workspace_bytes(int B, int R, int S, int tex_views) {
  if (B <= 0 || R < 0 || S <= 0 || tex_views <= 0) return 0;
  const size_t n = (size_t)B * R * S;
  // occupancies, geometry features of the texture views, colour tokens
  return (n * sizeof(float) + 255) / 256 * 256 +
         (n * tex_views * 8 * sizeof(float) + 255) / 256 * 256 +
         (size_t)B * R * kTexSlots * kTokW * sizeof(float);
}

namespace {

// Density pass: the MFMA kernel for up to 8 geometry views (half inputs,
// fp32 accumulation), or the fp32 vector kernel.
int launch_density(bool mfma, const half_t* pd, const float* z, const float* tpsdf,
                   const float* geo, int B, long long n_pts, int V, float* occ, float* geo_out,
                   int n_geo_out, hipStream_t st) {
  if (mfma && V <= 8) {
    const int VS = V <= 1 ? 1 : V <= 2 ? 2 : V <= 4 ? 4 : 8;
    const long long tiles = ((long long)B * n_pts + 32 / VS - 1) / (32 / VS);
    // persistent workgroups, 3 per CU (50 KB of LDS and ~150 registers per
    // lane each): the weights are staged once per workgroup
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
      cus = 256;
    const unsigned grid = (unsigned)std::max<long long>(
        1, std::min<long long>((tiles + kBlock / 64 - 1) / (kBlock / 64), 3LL * cus));
#define SAILOR_DENSITY_MFMA(N)                                                            \
  hipLaunchKernelGGL(density_mfma_kernel<N>, dim3(grid), dim3(kBlock), 0, st, pd, z, tpsdf, \
                     geo, B, n_pts, V, occ, geo_out, n_geo_out)
    if (VS == 1) SAILOR_DENSITY_MFMA(1);
    else if (VS == 2) SAILOR_DENSITY_MFMA(2);
    else if (VS == 4) SAILOR_DENSITY_MFMA(4);
    else SAILOR_DENSITY_MFMA(8);
#undef SAILOR_DENSITY_MFMA
    return sailor::check_launch("infer (density, mfma)");
  }
  hipLaunchKernelGGL(density_kernel, dim3(grid_for((long long)B * n_pts)), dim3(kBlock),
                     kDensityLds, st, pd, z, tpsdf, geo, B, n_pts, V, occ, geo_out, n_geo_out);
  return sailor::check_launch("infer (density)");
}

int infer_nerf_impl(bool mfma, const float* sorted_dists, const float* points_z,
                    const float* points_tpsdf, const float* feats_geo, const float* rgbs,
                    const float* feats_rgb, const float* dirs, const void* params_density,
                    const void* params_color, int B, int R, int S, int geo_views, int tex_views,
                    float* out_depth, float* out_rgb, float* out_feats, void* workspace,
                    size_t workspace_bytes, void* stream) {
  SAILOR_REQUIRE(B > 0 && R >= 0 && S > 0, "infer_nerf: bad sizes B=%d R=%d S=%d", B, R, S);
  SAILOR_REQUIRE(geo_views >= 1 && tex_views >= 1 && tex_views <= kTexSlots &&
                     tex_views <= geo_views,
                 "infer_nerf: need 1 <= tex_views <= min(geo_views, %d) (got %d, %d)", kTexSlots,
                 tex_views, geo_views);
  SAILOR_REQUIRE(workspace != nullptr &&
                     workspace_bytes >= sailor_infer_nerf_workspace_bytes(B, R, S, tex_views),
                 "infer_nerf: workspace too small");
  const long long n_pts = (long long)R * S;
  if (R == 0) return SAILOR_OK;
  hipStream_t st = sailor::as_stream(stream);
  float* occ = static_cast<float*>(workspace);
  float* geo8 = reinterpret_cast<float*>(static_cast<char*>(workspace) +
                                         ((size_t)B * n_pts * sizeof(float) + 255) / 256 * 256);
  if (int err = setup_all()) return err;
  if (int err = launch_density(mfma, static_cast<const half_t*>(params_density), points_z,
                               points_tpsdf, feats_geo, B, n_pts, geo_views, occ, geo8,
                               tex_views, st))
    return err;
  if (mfma && geo_views <= 8) {
    // fp16 form: compositing, then the Hydra stacks and RGB head on MFMA
    float* tok = reinterpret_cast<float*>(
        reinterpret_cast<char*>(geo8) +
        ((size_t)B * n_pts * tex_views * 8 * sizeof(float) + 255) / 256 * 256);
    hipLaunchKernelGGL(color_composite_kernel, dim3(grid_for((long long)B * R * kTexSlots)),
                       dim3(kBlock), 0, st, sorted_dists, occ, feats_rgb, rgbs, geo8, dirs, B, R,
                       S, tex_views, 1e-4f, out_depth, tok);
    if (int err = sailor::check_launch("infer_nerf (compositing)")) return err;
    const long long tiles = ((long long)B * R * kTexSlots + 15) / 16;
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
      cus = 256;
    const unsigned grid = (unsigned)std::max<long long>(
        1, std::min<long long>((tiles + kBlock / 64 - 1) / (kBlock / 64), 2LL * cus));
    hipLaunchKernelGGL(color_mfma_kernel, dim3(grid), dim3(kBlock), 0, st,
                       static_cast<const half_t*>(params_color), tok, B, R, tex_views, out_rgb,
                       out_feats);
    return sailor::check_launch("infer_nerf (colour, mfma)");
  }
  hipLaunchKernelGGL(color_kernel, dim3(grid_for((long long)B * R * kTexSlots)), dim3(kBlock),
                     kColorLds, st, static_cast<const half_t*>(params_color), sorted_dists, occ,
                     feats_rgb, rgbs, geo8, dirs, B, R, S, tex_views, 1e-4f, out_depth, out_rgb,
                     out_feats);
  return sailor::check_launch("infer_nerf (colour)");
}

int infer_density_impl(bool mfma, const float* points_z, const float* points_tpsdf,
                       const float* feats_geo, const void* params_density, int B, int N, int V,
                       float* out, void* stream) {
  SAILOR_REQUIRE(B > 0 && N >= 0 && V >= 1, "infer_density: bad sizes B=%d N=%d V=%d", B, N, V);
  if (N == 0) return SAILOR_OK;
  if (int err = setup_all()) return err;
  return launch_density(mfma, static_cast<const half_t*>(params_density), points_z, points_tpsdf,
                        feats_geo, B, (long long)N, V, out, nullptr, 0,
                        sailor::as_stream(stream));
}

}  // namespace

extern "C" int sailor_infer_nerf(const float* sorted_dists, const float* points_z,
                                 const float* points_tpsdf, const float* feats_geo,
                                 const float* rgbs, const float* feats_rgb, const float* dirs,
                                 const void* params_density, const void* params_color, int B,
                                 int R, int S, int geo_views, int tex_views, float* out_depth,
                                 float* out_rgb, float* out_feats, void* workspace,
                                 size_t workspace_bytes, void* stream) {
  return infer_nerf_impl(true, sorted_dists, points_z, points_tpsdf, feats_geo, rgbs, feats_rgb,
                         dirs, params_density, params_color, B, R, S, geo_views, tex_views,
                         out_depth, out_rgb, out_feats, workspace, workspace_bytes, stream);
}

extern "C" int sailor_infer_nerf_fp32(const float* sorted_dists, const float* points_z,
                                      const float* points_tpsdf, const float* feats_geo,
                                      const float* rgbs, const float* feats_rgb,
                                      const float* dirs, const void* params_density,
                                      const void* params_color, int B, int R, int S,
                                      int geo_views, int tex_views, float* out_depth,
                                      float* out_rgb, float* out_feats, void* workspace,
                                      size_t workspace_bytes, void* stream) {
  return infer_nerf_impl(false, sorted_dists, points_z, points_tpsdf, feats_geo, rgbs, feats_rgb,
                         dirs, params_density, params_color, B, R, S, geo_views, tex_views,
                         out_depth, out_rgb, out_feats, workspace, workspace_bytes, stream);
}

extern "C" int sailor_infer_density(const float* points_z, const float* points_tpsdf,
                                    const float* feats_geo, const void* params_density, int B,
                                    int N, int V, float* out, void* stream) {
  return infer_density_impl(true, points_z, points_tpsdf, feats_geo, params_density, B, N, V,
                            out, stream);
}

extern "C" int sailor_infer_density_fp32(const float* points_z, const float* points_tpsdf,
                                         const float* feats_geo, const void* params_density,
                                         int B, int N, int V, float* out, void* stream) {
  return infer_density_impl(false, points_z, points_tpsdf, feats_geo, params_density, B, N, V,
                            out, stream);
}

extern "C" int sailor_infer_blending(const float* feats_neighbors, const float* ray_feats,
                                     const void* params_blend, int N, float* out, void* stream) {
  SAILOR_REQUIRE(N >= 0, "infer_blending: bad size N=%d", N);
  if (N == 0) return SAILOR_OK;
  if (int err = setup_all()) return err;
  hipLaunchKernelGGL(blend_kernel, dim3(grid_for(N)), dim3(kBlock), kBlendLds,
                     sailor::as_stream(stream),
                     static_cast<const half_t*>(params_blend), feats_neighbors, ray_feats,
                     (long long)N, out);
  return sailor::check_launch("infer_blending");
}

extern "C" int sailor_rays_interpolate(const float* K, const float* R, const float* dirs,
                                       int n_cams, int n_rays, float* out, void* stream) {
  SAILOR_REQUIRE(n_cams > 0 && n_rays >= 0, "rays_interpolate: bad sizes cams=%d rays=%d",
                 n_cams, n_rays);
  if (n_rays == 0) return SAILOR_OK;
  hipLaunchKernelGGL(rays_interp_kernel, dim3(grid_for((long long)n_cams * n_rays * 4)),
                     dim3(kBlock), 0, sailor::as_stream(stream), K, R, dirs, n_cams, n_rays, out);
  return sailor::check_launch("rays_interpolate");
}

namespace {

int forward_upsampling_impl(bool mfma, const float* ray_ori, const float* ray_dir,
                            const float* depth, const float* coarse_rgb, const float* ray_feats,
                            int N, const float* tar_K, const float* tar_R, const float* K,
                            const float* R, const float* T, int V, int nei0, int nei1,
                            const float* rgbf_cl, int Hf, int Wf, const float* rgb_cl, int Hc,
                            int Wc, const float* depths, int Hd, int Wd, int res,
                            float valid_min, float valid_max, const void* params_blend,
                            float* out, void* stream) {
  SAILOR_REQUIRE(N >= 0 && V > 0, "forward_upsampling: bad sizes N=%d V=%d", N, V);
  SAILOR_REQUIRE(nei0 >= 0 && nei0 < V && nei1 >= 0 && nei1 < V,
                 "forward_upsampling: neighbour views (%d, %d) outside [0, %d)", nei0, nei1, V);
  SAILOR_REQUIRE(Hf > 1 && Wf > 1 && Hc > 1 && Wc > 1 && Hd > 1 && Wd > 1,
                 "forward_upsampling: maps must be at least 2 x 2");
  SAILOR_REQUIRE(res >= 2, "forward_upsampling: res must be at least 2");
  if (N == 0) return SAILOR_OK;
  if (int err = setup_all()) return err;
  UpsampleArgs a{ray_ori, ray_dir, depth, coarse_rgb, ray_feats, tar_K, tar_R, K, R, T,
                 rgbf_cl, rgb_cl, depths, (long long)N, Hf, Wf, Hc, Wc, Hd, Wd, res,
                 {nei0, nei1}, valid_min, valid_max, out};
  hipStream_t st = sailor::as_stream(stream);
  if (mfma) {
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
      cus = 256;
    const long long tiles = ((long long)N * 4 + 31) / 32;
    const unsigned grid = (unsigned)std::max<long long>(
        1, std::min<long long>((tiles + kBlock / 64 - 1) / (kBlock / 64), 3LL * cus));
    hipLaunchKernelGGL(upsample_mfma_kernel, dim3(grid), dim3(kBlock), 0, st,
                       static_cast<const half_t*>(params_blend), a);
  } else {
    hipLaunchKernelGGL(upsample_kernel, dim3(grid_for((long long)N * 4)), dim3(kBlock),
                       kBlendLds, st, static_cast<const half_t*>(params_blend), a);
  }
  return sailor::check_launch("forward_upsampling");
}

}  // namespace

extern "C" int sailor_forward_upsampling(
    const float* ray_ori, const float* ray_dir, const float* depth, const float* coarse_rgb,
    const float* ray_feats, int N, const float* tar_K, const float* tar_R, const float* K,
    const float* R, const float* T, int V, int nei0, int nei1, const float* rgbf_cl, int Hf,
    int Wf, const float* rgb_cl, int Hc, int Wc, const float* depths, int Hd, int Wd, int res,
    float valid_min, float valid_max, const void* params_blend, float* out, void* stream) {
  return forward_upsampling_impl(true, ray_ori, ray_dir, depth, coarse_rgb, ray_feats, N, tar_K,
                                 tar_R, K, R, T, V, nei0, nei1, rgbf_cl, Hf, Wf, rgb_cl, Hc, Wc,
                                 depths, Hd, Wd, res, valid_min, valid_max, params_blend, out,
                                 stream);
}

extern "C" int sailor_forward_upsampling_fp32(
    const float* ray_ori, const float* ray_dir, const float* depth, const float* coarse_rgb,
    const float* ray_feats, int N, const float* tar_K, const float* tar_R, const float* K,
    const float* R, const float* T, int V, int nei0, int nei1, const float* rgbf_cl, int Hf,
    int Wf, const float* rgb_cl, int Hc, int Wc, const float* depths, int Hd, int Wd, int res,
    float valid_min, float valid_max, const void* params_blend, float* out, void* stream) {
  return forward_upsampling_impl(false, ray_ori, ray_dir, depth, coarse_rgb, ray_feats, N, tar_K,
                                 tar_R, K, R, T, V, nei0, nei1, rgbf_cl, Hf, Wf, rgb_cl, Hc, Wc,
                                 depths, Hd, Wd, res, valid_min, valid_max, params_blend, out,
                                 stream);
}
