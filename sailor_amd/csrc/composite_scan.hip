// Differentiable compositing of samples taken in the order given (the
// PyTorch compositing the reference trains with), forward and gradient.
//
// Semantics of c_lib/VoxelEncoding/voxel_encoding_helper.py:
//  * _volume_render_torch.forward (:132-164), density form:
//      delta_i = t_{i+1} - t_i, delta_{S-1} = 0;  e_i = delta_i sigma_i
//      alpha_i = 1 - exp(-e_i);  T_i = exp(-sum_{j<i} e_j);  w_i = alpha_i T_i
//  * _volume_render_torch.forward_occ (:166-186), occupancy form:
//      alpha_i = sigma_i;  T_i = prod_{j<i} (1 - alpha_j + 1e-6);  w_i = alpha_i T_i
//  * outputs feats = sum w_i f_i, depth = sum w_i t_i, ws = sum w_i,
//    alphas = sum w_i alpha_i; rays whose `valid` byte is 0 get zeros (and
//    zero gradients), as the reference's masked assignment does.
// No sort index, no early termination and no skipping of negative depths:
// every sample of the row takes part, exactly as in the tensor formulation.
//
// Gradient: with G_i = dL/dw_i = g_f . f_i + g_d t_i + g_w + g_a alpha_i and
// the suffix sums S_k = sum_{i>k} G_i w_i (accumulated walking the row
// backwards),
//   density:   dL/dsigma_k = delta_k [ exp(-e_k) (G_k T_k + g_a w_k) - S_k ]
//   occupancy: dL/dalpha_k = G_k T_k + g_a w_k - S_k / (1 - alpha_k + 1e-6)
//   dL/df_k = w_k g_f.
// The backward's first walk parks T_k in the gradient row it later
// overwrites, so the reverse walk needs no division by transmittances.
//
// With `compensated` the forward keeps Kahan-compensated running sums (the
// density form's exponent prefix sum and the four output sums).
//
// One lane per ray; channels are processed in groups of up to FG per lane
// (grid.y = channel groups), each lane re-running the cheap scalar
// recurrence of its ray.
#include "common.h"

namespace {

constexpr int kThreads = 128;
constexpr float kOccEps = 1e-6f;

struct Kahan {
  float s = 0.0f, c = 0.0f;
  __device__ __forceinline__ void add(float x, bool comp) {
    if (!comp) {
      s += x;
      return;
    }
    const float y = x - c;
    const float t = s + y;
    c = (t - s) - y;
    s = t;
  }
  // the compensated value of the sum (c holds the negated lost low bits)
  __device__ __forceinline__ float value() const { return s - c; }
};

template <int FG>
__global__ void __launch_bounds__(kThreads)
scan_forward(const float* __restrict__ t, const float* __restrict__ sigma,
             const float* __restrict__ feats, const unsigned char* __restrict__ valid,
             long long n_rays, int S, int F, int occ_mode, int compensated,
             float* __restrict__ out_feats, float* __restrict__ out_depth,
             float* __restrict__ out_ws, float* __restrict__ out_alpha) {
  const long long ray = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (ray >= n_rays) return;
  const int f0 = blockIdx.y * FG;
  const int nf = min(FG, F - f0);
  const bool comp = compensated != 0;
  const float* tr = t + ray * S;
  const float* sr = sigma + ray * S;
  const float* fr = feats + ray * (long long)S * F + f0;
  Kahan acc[FG], dep, ws, al, pre;
  float Tprod = 1.0f;
  const bool live = valid == nullptr || valid[ray] != 0;
  if (live) {
    for (int i = 0; i < S; ++i) {
      const float z = tr[i];
      float alpha, T;
      if (occ_mode) {
        alpha = sr[i];
        T = Tprod;
        Tprod *= 1.0f - alpha + kOccEps;
      } else {
        const float delta = i + 1 < S ? tr[i + 1] - z : 0.0f;
        const float e = delta * sr[i];
        T = expf(-pre.value());
        alpha = 1.0f - expf(-e);
        pre.add(e, comp);
      }
      const float w = alpha * T;
#pragma unroll
      for (int j = 0; j < FG; ++j)
        if (j < nf) acc[j].add(w * fr[(size_t)i * F + j], comp);
      if (f0 == 0) {
        dep.add(w * z, comp);
        ws.add(w, comp);
        al.add(w * alpha, comp);
      }
    }
  }
  float* of = out_feats + ray * F + f0;
#pragma unroll
  for (int j = 0; j < FG; ++j)
    if (j < nf) of[j] = acc[j].value();
  if (f0 == 0) {
    out_depth[ray] = dep.value();
    out_ws[ray] = ws.value();
    out_alpha[ray] = al.value();
  }
}

// One lane per ray (all channels: G_i needs the full dot product g_f . f_i).
template <int FMAX>
__global__ void __launch_bounds__(kThreads)
scan_backward(const float* __restrict__ g_feats, const float* __restrict__ g_depth,
              const float* __restrict__ g_ws, const float* __restrict__ g_alpha,
              const float* __restrict__ t, const float* __restrict__ sigma,
              const float* __restrict__ feats, const unsigned char* __restrict__ valid,
              long long n_rays, int S, int F, int occ_mode, float* __restrict__ grad_sigma,
              float* __restrict__ grad_feats) {
  const long long ray = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (ray >= n_rays) return;
  const float* tr = t + ray * S;
  const float* sr = sigma + ray * S;
  const float* fr = feats + ray * (long long)S * F;
  float* gs = grad_sigma + ray * S;
  float* gf = grad_feats + ray * (long long)S * F;
  if (valid != nullptr && valid[ray] == 0) {
    for (int i = 0; i < S; ++i) gs[i] = 0.0f;
    for (long long i = 0; i < (long long)S * F; ++i) gf[i] = 0.0f;
    return;
  }
  float gF[FMAX];
#pragma unroll
  for (int j = 0; j < FMAX; ++j) gF[j] = j < F ? g_feats[ray * F + j] : 0.0f;
  const float gD = g_depth[ray], gW = g_ws[ray], gA = g_alpha[ray];
  // walk 1: transmittances (parked in the sigma-gradient row) and dL/df
  float pre = 0.0f, Tprod = 1.0f;
  for (int i = 0; i < S; ++i) {
    float alpha, T;
    if (occ_mode) {
      alpha = sr[i];
      T = Tprod;
      Tprod *= 1.0f - alpha + kOccEps;
    } else {
      const float delta = i + 1 < S ? tr[i + 1] - tr[i] : 0.0f;
      const float e = delta * sr[i];
      T = expf(-pre);
      alpha = 1.0f - expf(-e);
      pre += e;
    }
    const float w = alpha * T;
    gs[i] = T;
#pragma unroll
    for (int j = 0; j < FMAX; ++j)
      if (j < F) gf[(size_t)i * F + j] = w * gF[j];
  }
  // walk 2, last sample first: suffix sums and dL/dsigma
  float suffix = 0.0f;
  for (int i = S - 1; i >= 0; --i) {
    const float T = gs[i];
    const float z = tr[i];
    float G = gD * z + gW;
#pragma unroll
    for (int j = 0; j < FMAX; ++j)
      if (j < F) G += gF[j] * fr[(size_t)i * F + j];
    float grad;
    float w;
    if (occ_mode) {
      const float alpha = sr[i];
      w = alpha * T;
      G += gA * alpha;
      grad = G * T + gA * w - suffix / (1.0f - alpha + kOccEps);
    } else {
      const float delta = i + 1 < S ? tr[i + 1] - z : 0.0f;
      const float e = delta * sr[i];
      const float keep = expf(-e);
      const float alpha = 1.0f - keep;
      w = alpha * T;
      G += gA * alpha;
      grad = delta * (keep * (G * T + gA * w) - suffix);
    }
    gs[i] = grad;
    suffix += G * w;
  }
}

int check_args(long long n_rays, int S, int F) {
  SAILOR_REQUIRE(n_rays >= 0 && S >= 1 && F >= 1,
                 "volume_render_torch: need n_rays >= 0, S >= 1, F >= 1 (got %lld, %d, %d)",
                 n_rays, S, F);
  return SAILOR_OK;
}

}  // namespace

extern "C" int sailor_volume_render_torch_forward(const float* t, const float* sigma,
                                                  const float* feats,
                                                  const unsigned char* valid, long long n_rays,
                                                  int S, int F, int occ_mode, int compensated,
                                                  float* out_feats, float* out_depth,
                                                  float* out_ws, float* out_alpha,
                                                  void* stream) {
  int rc = check_args(n_rays, S, F);
  if (rc) return rc;
  if (n_rays == 0) return SAILOR_OK;
  constexpr int FG = 8;
  const int groups = (F + FG - 1) / FG;
  SAILOR_REQUIRE(groups <= 65535, "volume_render_torch: F = %d too large", F);
  hipLaunchKernelGGL(scan_forward<FG>, dim3(sailor::ceil_div(n_rays, kThreads), groups),
                     dim3(kThreads), 0, sailor::as_stream(stream), t, sigma, feats, valid,
                     n_rays, S, F, occ_mode, compensated, out_feats, out_depth, out_ws,
                     out_alpha);
  return sailor::check_launch("volume_render_torch_forward");
}

extern "C" int sailor_volume_render_torch_backward(
    const float* g_feats, const float* g_depth, const float* g_ws, const float* g_alpha,
    const float* t, const float* sigma, const float* feats, const unsigned char* valid,
    long long n_rays, int S, int F, int occ_mode, float* grad_sigma, float* grad_feats,
    void* stream) {
  int rc = check_args(n_rays, S, F);
  if (rc) return rc;
  SAILOR_REQUIRE(F <= 64, "volume_render_torch_backward: at most 64 channels (got %d)", F);
  if (n_rays == 0) return SAILOR_OK;
  const dim3 grid(sailor::ceil_div(n_rays, kThreads)), block(kThreads);
  hipStream_t st = sailor::as_stream(stream);
#define SAILOR_SCAN_BWD(FM)                                                               \
  hipLaunchKernelGGL(scan_backward<FM>, grid, block, 0, st, g_feats, g_depth, g_ws, g_alpha, \
                     t, sigma, feats, valid, n_rays, S, F, occ_mode, grad_sigma, grad_feats)
  if (F <= 4)
    SAILOR_SCAN_BWD(4);
  else if (F <= 16)
    SAILOR_SCAN_BWD(16);
  else
    SAILOR_SCAN_BWD(64);
#undef SAILOR_SCAN_BWD
  return sailor::check_launch("volume_render_torch_backward");
}
