// Fused per-ray radiance kernel: feature gather + positional encoding +
// per-view density MLP on MFMA + view mean + occupancy head + front-to-back
// compositing + multi-view Hydra attention + RGB head, one wavefront per ray.
//
// Semantics follow the reference's FastNerf path (c_lib/FastNerf/src/
// infer_mlp.cu: density_feat_kernel :120-199 per (point, view), torch::mean
// over views :698-699, density_sigma_kernel :202-239, per-view feature
// compositing along the ray :724-747 with volume_rendering_occ_kernel
// :508-581, hydra_attention_kernel :242-445 on the composited per-view
// tokens, RGB head :427-432) and the PyTorch modules it fuses
// (SRONet/nerf.py:139-186, SRONet/aggregation.py:85-128,155-171,288-307), with
// the widths of the benchmark model as template parameters instead of the
// reference's hard-coded 64/32/30.
//
// MI355X design
//  * columns (MFMA N) are (sample, view) pairs, rows (MFMA M) hidden units:
//    a layer's fp32 accumulator tiles are the next layer's B fragments after a
//    fixed permutation of the reduction index that the host applies to the
//    weights, so activations never leave registers (no LDS transposes);
//  * inputs are gathered straight into the first layer's B fragments: each
//    lane of a column fetches a contiguous quarter of the channels of the
//    four bilinear taps (channels-last feature maps -> coalesced 256 B rows);
//  * per-sample results are composited in registers in sample order, so the
//    per-sample density, colour features and the per-view tokens never reach
//    HBM; the per-ray Hydra attention and RGB head run from LDS.
//  * the view mean is folded into the occupancy head (mean of per-column dot
//    products) so it costs four cross-lane adds per column.
#include "common.h"
#include "mlp_layout.h"

namespace {

using namespace sailor::mlp;

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));

enum Prec { kF32 = 0, kBF16 = 1, kF16 = 2 };

constexpr int kWaves = 4;   // rays per workgroup
constexpr int kNT = 2;      // N-tiles (4 samples x 4 views each) per iteration
constexpr int kVP = 4;      // view slots per sample
constexpr float kHalfPi = 1.5707963267948966f;

struct RenderArgs {
  const float* t;          // [n_rays, S] sorted sample distances
  const float* rays;       // [n_rays, 6] origin, direction
  const float* Kv;         // [B, V, 3, 3] source intrinsics
  const float* RTv;        // [B, V, 3, 4] source extrinsics
  const float* fmap;       // [B, V, Hf, Wf, F] channels-last feature maps
  const float* bbox;       // [B, 3, 2] volume box (positional-encoding frame)
  const float* zdir;       // [B, 3] third row of the target rotation (depth = t * zdir.d)
  const void* wdense;      // packed dense layers
  const float* bdense;     // [4, H] biases
  const float* wsmall;     // occupancy head, Hydra, RGB head (see mlp_layout.h)
  float* out_rgb;          // [n_rays, 3]
  float* out_depth;        // [n_rays]
  float* out_aux;          // [n_rays, 2] (weight sum, alpha sum) or null
  float* out_occ;          // [n_rays, S] per-sample occupancy or null
  long long n_rays;
  int rays_per_batch;
  int S, V, img_w, img_h, Hf, Wf;
  float t_thres;
  int early_exit;
};

__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ float sigmoidf(float x) { return 1.0f / (1.0f + expf(-x)); }

template <int PREC>
struct ElemOf { typedef float type; };
template <>
struct ElemOf<kBF16> { typedef __bf16 type; };
template <>
struct ElemOf<kF16> { typedef _Float16 type; };

// Bilinear sample (grid_sample, align_corners=True, zero padding) of CPL
// channels starting at c0 for one (point, view); projection as proj_persp
// (utils_render/utils_render.py:380-410) with res = image size.
template <int F, int CPL>
__device__ __forceinline__ void gather_feats(const float* __restrict__ fmap, int Hf, int Wf,
                                             const float* Kv, const float* RT, int img_w,
                                             int img_h, bool vvalid, float x, float y, float z,
                                             int c0, float* out) {
  float px = RT[0] * x + RT[1] * y + RT[2] * z + RT[3];
  float py = RT[4] * x + RT[5] * y + RT[6] * z + RT[7];
  float pz = RT[8] * x + RT[9] * y + RT[10] * z + RT[11];
  const bool inv = !(pz >= 0.01f) || !vvalid;
  const float dz = inv ? 1.0f : pz;
  px /= dz;
  py /= dz;
  const float pzn = pz / dz;
  const float u = Kv[0] * px + Kv[1] * py + Kv[2] * pzn;
  const float w = Kv[3] * px + Kv[4] * py + Kv[5] * pzn;
  const float hw = (float)(img_w / 2), hh = (float)(img_h / 2);
  float gx = (u - hw) / hw, gy = (w - hh) / hh;
  if (inv) { gx = -2.0f; gy = -2.0f; }
  const float ix = (gx + 1.0f) * 0.5f * (float)(Wf - 1);
  const float iy = (gy + 1.0f) * 0.5f * (float)(Hf - 1);
#pragma unroll
  for (int c = 0; c < CPL; ++c) out[c] = 0.0f;
  if (!(ix > -1.0f && ix < (float)Wf && iy > -1.0f && iy < (float)Hf)) return;
  const float fx0 = floorf(ix), fy0 = floorf(iy);
  const int x0 = (int)fx0, y0 = (int)fy0;
  const float ax = ix - fx0, ay = iy - fy0;
#pragma unroll
  for (int tap = 0; tap < 4; ++tap) {
    const int xx = x0 + (tap & 1), yy = y0 + (tap >> 1);
    if (xx < 0 || xx >= Wf || yy < 0 || yy >= Hf) continue;
    const float wt = ((tap & 1) ? ax : 1.0f - ax) * ((tap >> 1) ? ay : 1.0f - ay);
    const f32x4* src = reinterpret_cast<const f32x4*>(fmap + ((size_t)yy * Wf + xx) * F + c0);
#pragma unroll
    for (int q = 0; q < CPL / 4; ++q) {
      f32x4 v4 = src[q];
      out[4 * q + 0] += wt * v4[0];
      out[4 * q + 1] += wt * v4[1];
      out[4 * q + 2] += wt * v4[2];
      out[4 * q + 3] += wt * v4[3];
    }
  }
}

// 16 positional-encoding dims [16g, 16g+16) of xn: [x, sin(2^f x), sin(2^f x + pi/2), ...]
// (freq_encoding.cu:37-45 layout, degree 10 -> 63 dims, dim 63 is padding).
__device__ __forceinline__ void pe16(const float* xn, int g, float* out) {
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const int q = 16 * g + k;
    float v;
    if (q < 3) {
      v = xn[q];
    } else if (q < 63) {
      const int d = q % 3, col = q / 3 - 1, f = col >> 1;
      v = sinf(ldexpf(xn[d], f) + (float)(col & 1) * kHalfPi);
    } else {
      v = 0.0f;
    }
    out[k] = v;
  }
}

// ---------------------------------------------------------------------------
// dense layers
template <int H, int PREC>
struct Dense {
  static constexpr int MT = H / 16;
  typedef typename ElemOf<PREC>::type E;
  static constexpr int EPL = PREC == kF32 ? 1 : 8;  // elements per lane per fragment

  static __device__ __forceinline__ f32x4 mfma(f32x4 acc, const E* a, const E* b) {
    if constexpr (PREC == kF32) {
      return __builtin_amdgcn_mfma_f32_16x16x4f32(a[0], b[0], acc, 0, 0, 0);
    } else if constexpr (PREC == kBF16) {
      bf16x8 av, bv;
#pragma unroll
      for (int j = 0; j < 8; ++j) { av[j] = a[j]; bv[j] = b[j]; }
      return __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc, 0, 0, 0);
    } else {
      f16x8 av, bv;
#pragma unroll
      for (int j = 0; j < 8; ++j) { av[j] = a[j]; bv[j] = b[j]; }
      return __builtin_amdgcn_mfma_f32_16x16x32_f16(av, bv, acc, 0, 0, 0);
    }
  }

  // acc[nt][mt] += A(step) * Bfrag(step) over KS steps; bfrag(nt, s, E* out)
  template <int KS, class BFrag>
  static __device__ __forceinline__ void layer(const E* __restrict__ W, int lane,
                                               f32x4 (&acc)[kNT][MT], BFrag bfrag) {
#pragma unroll
    for (int nt = 0; nt < kNT; ++nt)
#pragma unroll
      for (int mt = 0; mt < MT; ++mt) acc[nt][mt] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s = 0; s < KS; ++s) {
      E a[MT * EPL];
      const E* ap = W + ((size_t)s * 64 + lane) * MT * EPL;
      if constexpr (PREC == kF32) {
#pragma unroll
        for (int q = 0; q < MT / 4; ++q) {
          f32x4 v = reinterpret_cast<const f32x4*>(ap)[q];
          a[4 * q + 0] = v[0]; a[4 * q + 1] = v[1]; a[4 * q + 2] = v[2]; a[4 * q + 3] = v[3];
        }
      } else {
#pragma unroll
        for (int q = 0; q < MT; ++q) {
          f32x4 v = reinterpret_cast<const f32x4*>(ap)[q];  // 8 x 16-bit
          const E* pv = reinterpret_cast<const E*>(&v);
#pragma unroll
          for (int j = 0; j < 8; ++j) a[8 * q + j] = pv[j];
        }
      }
#pragma unroll
      for (int nt = 0; nt < kNT; ++nt) {
        E b[EPL];
        bfrag(nt, s, b);
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) acc[nt][mt] = mfma(acc[nt][mt], a + mt * EPL, b);
      }
    }
  }
};

template <int F, int H, int PREC>
__global__ void __launch_bounds__(kWaves * 64)
render_kernel(RenderArgs a) {
  constexpr int MT = H / 16;
  constexpr int CPL = F / 4;
  typedef Dense<H, PREC> D;
  typedef typename D::E E;
  constexpr int EPL = D::EPL;
  constexpr int KS0 = PREC == kF32 ? (F + kPeSlots) / 4 : (F + kPeSlots) / 32;
  constexpr int KSH = PREC == kF32 ? H / 4 : H / 32;
  constexpr int FS = PREC == kF32 ? CPL : F / 32;  // first-layer feature steps
  constexpr SmallOffsets so = small_offsets(F, H);

  __shared__ float s_x[kWaves][kVP][F];
  __shared__ float s_o[kWaves][kVP][F];
  __shared__ float s_q[kWaves][3][kVP][kHydraInner];
  __shared__ float s_h[kWaves][kVP][kHydraInner];
  __shared__ float s_r[kWaves][kRgbH1 + kRgbH2];
  __shared__ float s_nrm[kWaves][2][kVP];

  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const long long ray = (long long)blockIdx.x * kWaves + wid;
  if (ray >= a.n_rays) return;
  const int S = a.S;
  const float* trow = a.t + ray * S;
  if (!(trow[0] >= 0.0f)) {  // ray without a full set of samples: dropped
    if (lane < 3) a.out_rgb[ray * 3 + lane] = 0.0f;
    if (lane == 3) a.out_depth[ray] = 0.0f;
    if (a.out_aux && lane < 2) a.out_aux[ray * 2 + lane] = 0.0f;
    return;
  }
  const int b = (int)(ray / a.rays_per_batch);
  const int n = lane & 15, g = lane >> 4, vv = n & 3, sl = n >> 2;
  const bool vvalid = vv < a.V;
  const int vsel = vvalid ? vv : 0;
  float Kc[9], RTc[12];
#pragma unroll
  for (int k = 0; k < 9; ++k) Kc[k] = a.Kv[((size_t)b * a.V + vsel) * 9 + k];
#pragma unroll
  for (int k = 0; k < 12; ++k) RTc[k] = a.RTv[((size_t)b * a.V + vsel) * 12 + k];
  const float* fm = a.fmap + ((size_t)b * a.V + vsel) * a.Hf * a.Wf * F;
  const float ox = a.rays[ray * 6 + 0], oy = a.rays[ray * 6 + 1], oz = a.rays[ray * 6 + 2];
  const float dx = a.rays[ray * 6 + 3], dy = a.rays[ray * 6 + 4], dz = a.rays[ray * 6 + 5];
  const float* bb = a.bbox + b * 6;
  const float cx = 0.5f * (bb[0] + bb[1]), cy = 0.5f * (bb[2] + bb[3]),
              cz = 0.5f * (bb[4] + bb[5]);
  const float ihx = 2.0f / (bb[1] - bb[0]), ihy = 2.0f / (bb[3] - bb[2]),
              ihz = 2.0f / (bb[5] - bb[4]);
  const float inv_v = 1.0f / (float)a.V;
  const E* W0 = reinterpret_cast<const E*>(a.wdense);
  const size_t l0_elems = (size_t)KS0 * 64 * MT * EPL;
  const size_t lh_elems = (size_t)KSH * 64 * MT * EPL;
  const float* ws = a.wsmall;

  float T = 1.0f, dsum = 0.0f, wsum = 0.0f, asum = 0.0f;
  bool done = false;
  float tok[CPL];
#pragma unroll
  for (int c = 0; c < CPL; ++c) tok[c] = 0.0f;

  const int n_iter = S / (4 * kNT);
  for (int it = 0; it < n_iter; ++it) {
    if (a.early_exit && done) break;
    float feat[kNT][CPL], pe[kNT][16], ts[kNT];
#pragma unroll
    for (int nt = 0; nt < kNT; ++nt) {
      const int s = it * 4 * kNT + nt * 4 + sl;
      const float t = trow[s];
      ts[nt] = t;
      const float x = ox + t * dx, y = oy + t * dy, z = oz + t * dz;
      gather_feats<F, CPL>(fm, a.Hf, a.Wf, Kc, RTc, a.img_w, a.img_h, vvalid, x, y, z, CPL * g,
                           feat[nt]);
      const float xn[3] = {(x - cx) * ihx, (y - cy) * ihy, (z - cz) * ihz};
      pe16(xn, g, pe[nt]);
    }
    f32x4 acc[kNT][MT];
    // layer 0: [pe | feat] -> H
    D::template layer<KS0>(W0, lane, acc, [&](int nt, int s, E* out) {
      if constexpr (PREC == kF32) {
        out[0] = s < FS ? feat[nt][s] : pe[nt][s - FS];
      } else {
#pragma unroll
        for (int j = 0; j < 8; ++j)
          out[j] = (E)(s < FS ? feat[nt][8 * s + j] : pe[nt][8 * (s - FS) + j]);
      }
    });
    // hidden layers 1..3 (runtime loop keeps one copy of the unrolled body)
    for (int l = 1; l < 4; ++l) {
      const float* bias = a.bdense + (l - 1) * H;
      f32x4 h[kNT][MT];
#pragma unroll
      for (int nt = 0; nt < kNT; ++nt)
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) {
          const f32x4 bv = *reinterpret_cast<const f32x4*>(bias + 16 * mt + 4 * g);
#pragma unroll
          for (int i = 0; i < 4; ++i) h[nt][mt][i] = fmaxf(acc[nt][mt][i] + bv[i], 0.0f);
        }
      const E* Wl = W0 + l0_elems + (size_t)(l - 1) * lh_elems;
      D::template layer<KSH>(Wl, lane, acc, [&](int nt, int s, E* out) {
        if constexpr (PREC == kF32) {
          out[0] = h[nt][s >> 2][s & 3];
        } else {
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            out[j] = (E)h[nt][2 * s][j];
            out[4 + j] = (E)h[nt][2 * s + 1][j];
          }
        }
      });
    }
    // last ReLU, occupancy head on the view mean, compositing in sample order
    float alpha[kNT];
    {
      const float* bias = a.bdense + 3 * H;
#pragma unroll
      for (int nt = 0; nt < kNT; ++nt) {
        float part = 0.0f;
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) {
          const f32x4 bv = *reinterpret_cast<const f32x4*>(bias + 16 * mt + 4 * g);
          const f32x4 wv = *reinterpret_cast<const f32x4*>(ws + so.sigma_w + 16 * mt + 4 * g);
#pragma unroll
          for (int i = 0; i < 4; ++i) part += wv[i] * fmaxf(acc[nt][mt][i] + bv[i], 0.0f);
        }
        part = vvalid ? part : 0.0f;
        part += __shfl_xor(part, 16, 64);
        part += __shfl_xor(part, 32, 64);
        part += __shfl_xor(part, 1, 64);
        part += __shfl_xor(part, 2, 64);
        alpha[nt] = sigmoidf(part * inv_v + ws[so.sigma_b]);
      }
    }
#pragma unroll
    for (int nt = 0; nt < kNT; ++nt) {
      float wmine = 0.0f;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const float al = __shfl(alpha[nt], 4 * k, 64);
        const float tk = __shfl(ts[nt], 4 * k, 64);
        float w = 0.0f;
        if (!done) {
          w = al * T;
          dsum += w * tk;
          wsum += w;
          asum += w * al;
          T *= 1.0f - al;
          if (T < a.t_thres) done = true;
        }
        wmine = (sl == k) ? w : wmine;
      }
      if (a.out_occ && g == 0 && vv == 0)
        a.out_occ[ray * S + it * 4 * kNT + nt * 4 + sl] = alpha[nt];
#pragma unroll
      for (int c = 0; c < CPL; ++c) tok[c] += wmine * feat[nt][c];
    }
  }
  // per-view tokens: sum the four sample lanes (lane bits 2, 3)
#pragma unroll
  for (int c = 0; c < CPL; ++c) {
    tok[c] += __shfl_xor(tok[c], 4, 64);
    tok[c] += __shfl_xor(tok[c], 8, 64);
  }
  float(*x)[F] = s_x[wid];
  float(*o)[F] = s_o[wid];
  if (sl == 0) {
#pragma unroll
    for (int c = 0; c < CPL; ++c) x[vv][CPL * g + c] = vvalid ? tok[c] : 0.0f;
  }
  wave_lds_sync();
  const int V = a.V;
  // ---- Hydra encoder, 2 stacks (aggregation.py:101-128,164-171,299-307)
  for (int st = 0; st < kHydraStacks; ++st) {
    // q, k, v projections (no bias): 3 * 4 * 32 outputs
    for (int oi = lane; oi < 3 * kVP * kHydraInner; oi += 64) {
      const int which = oi / (kVP * kHydraInner), rem = oi % (kVP * kHydraInner);
      const int v = rem / kHydraInner, e = rem % kHydraInner;
      const int off = which == 0 ? so.q[st] : (which == 1 ? so.k[st] : so.v[st]);
      float accv = 0.0f;
      if (v < V) {
        const float* wr = ws + off + e * F;
#pragma unroll 8
        for (int c = 0; c < F; ++c) accv += wr[c] * x[v][c];
      }
      s_q[wid][which][v][e] = accv;
    }
    wave_lds_sync();
    // L2 norms of q and k (F.normalize, eps 1e-12)
    if (lane < 2 * kVP) {
      const int which = lane / kVP, v = lane % kVP;
      float ss = 0.0f;
      for (int e = 0; e < kHydraInner; ++e) ss += s_q[wid][which][v][e] * s_q[wid][which][v][e];
      s_nrm[wid][which][v] = 1.0f / fmaxf(sqrtf(ss), 1e-12f);
    }
    wave_lds_sync();
    // kv = sum_v khat * v ;  attn_v = qhat_v * kv
    if (lane < kHydraInner) {
      const int e = lane;
      float kv = 0.0f;
      for (int v = 0; v < V; ++v) kv += s_q[wid][1][v][e] * s_nrm[wid][1][v] * s_q[wid][2][v][e];
      for (int v = 0; v < kVP; ++v) s_h[wid][v][e] = s_q[wid][0][v][e] * s_nrm[wid][0][v] * kv;
    }
    wave_lds_sync();
    // out = fc(attn) + residual (the layer input)
    for (int oi = lane; oi < kVP * F; oi += 64) {
      const int v = oi / F, c = oi % F;
      float accv = 0.0f;
      const float* wr = ws + so.fc[st] + c * kHydraInner;
#pragma unroll 8
      for (int e = 0; e < kHydraInner; ++e) accv += wr[e] * s_h[wid][v][e];
      o[v][c] = accv + x[v][c];
    }
    wave_lds_sync();
    // FFN: relu(W1 out + b1)
    for (int oi = lane; oi < kVP * kHydraInner; oi += 64) {
      const int v = oi / kHydraInner, e = oi % kHydraInner;
      float accv = ws[so.b1[st] + e];
      const float* wr = ws + so.w1[st] + e * F;
#pragma unroll 8
      for (int c = 0; c < F; ++c) accv += wr[c] * o[v][c];
      s_h[wid][v][e] = fmaxf(accv, 0.0f);
    }
    wave_lds_sync();
    // W2 h + b2 + out -> next layer input
    for (int oi = lane; oi < kVP * F; oi += 64) {
      const int v = oi / F, c = oi % F;
      float accv = ws[so.b2[st] + c];
      const float* wr = ws + so.w2[st] + c * kHydraInner;
#pragma unroll 8
      for (int e = 0; e < kHydraInner; ++e) accv += wr[e] * s_h[wid][v][e];
      x[v][c] = accv + o[v][c];
    }
    wave_lds_sync();
  }
  // ---- RGB head on token 0: F -> 64 -> 32 -> 3 (nerf.py:117-124)
  float* r = s_r[wid];
  {
    const int e = lane;  // 64 outputs
    float accv = ws[so.r1b + e];
    const float* wr = ws + so.r1w + e * F;
#pragma unroll 8
    for (int c = 0; c < F; ++c) accv += wr[c] * x[0][c];
    r[e] = fmaxf(accv, 0.0f);
  }
  wave_lds_sync();
  if (lane < kRgbH2) {
    float accv = ws[so.r2b + lane];
    const float* wr = ws + so.r2w + lane * kRgbH1;
    for (int c = 0; c < kRgbH1; ++c) accv += wr[c] * r[c];
    r[kRgbH1 + lane] = fmaxf(accv, 0.0f);
  }
  wave_lds_sync();
  if (lane < 3) {
    float accv = ws[so.r3b + lane];
    const float* wr = ws + so.r3w + lane * kRgbH2;
    for (int c = 0; c < kRgbH2; ++c) accv += wr[c] * r[kRgbH1 + c];
    a.out_rgb[ray * 3 + lane] = sigmoidf(accv);
  }
  if (lane == 0) {
    const float* zd = a.zdir + b * 3;
    a.out_depth[ray] = dsum * (zd[0] * dx + zd[1] * dy + zd[2] * dz);
    if (a.out_aux) {
      a.out_aux[ray * 2 + 0] = wsum;
      a.out_aux[ray * 2 + 1] = asum;
    }
  }
}

template <int F, int H, int PREC>
int launch(const RenderArgs& a, hipStream_t st) {
  const int blocks = (int)((a.n_rays + kWaves - 1) / kWaves);
  hipLaunchKernelGGL((render_kernel<F, H, PREC>), dim3(blocks), dim3(kWaves * 64), 0, st, a);
  return sailor::check_launch("render_rays");
}

template <int F, int H>
int dispatch_prec(int prec, const RenderArgs& a, hipStream_t st) {
  switch (prec) {
    case kF32: return launch<F, H, kF32>(a, st);
    case kBF16: return launch<F, H, kBF16>(a, st);
    case kF16: return launch<F, H, kF16>(a, st);
  }
  return sailor::set_error(SAILOR_ERR_ARG, "render_rays: precision %d unsupported", prec);
}

}  // namespace

extern "C" size_t sailor_render_small_weights_floats(int F, int H) {
  return (size_t)small_offsets(F, H).total;
}

extern "C" int sailor_render_rays(const float* t, const float* rays, int B, int R, int S,
                                  const float* Kv, const float* RTv, int V, int img_h,
                                  int img_w, const float* fmap, int Hf, int Wf, int F,
                                  const float* bbox, const float* zdir, int H,
                                  int precision, const void* w_dense, const float* b_dense,
                                  const float* w_small, float t_thresh, int early_exit,
                                  float* out_rgb, float* out_depth, float* out_aux,
                                  float* out_occ, void* stream) {
  SAILOR_REQUIRE(B > 0 && R >= 0, "render_rays: bad sizes");
  SAILOR_REQUIRE(S > 0 && S % (4 * kNT) == 0, "render_rays: S must be a multiple of %d",
                 4 * kNT);
  SAILOR_REQUIRE(V >= 1 && V <= kVP, "render_rays: 1 to %d views supported", kVP);
  SAILOR_REQUIRE(Hf > 1 && Wf > 1, "render_rays: feature map too small");
  if (R == 0) return SAILOR_OK;
  RenderArgs a;
  a.t = t; a.rays = rays; a.Kv = Kv; a.RTv = RTv; a.fmap = fmap; a.bbox = bbox; a.zdir = zdir;
  a.wdense = w_dense; a.bdense = b_dense; a.wsmall = w_small;
  a.out_rgb = out_rgb; a.out_depth = out_depth; a.out_aux = out_aux; a.out_occ = out_occ;
  a.n_rays = (long long)B * R; a.rays_per_batch = R;
  a.S = S; a.V = V; a.img_w = img_w; a.img_h = img_h; a.Hf = Hf; a.Wf = Wf;
  a.t_thres = t_thresh; a.early_exit = early_exit;
  hipStream_t st = sailor::as_stream(stream);
  if (F == 64 && H == 256) return dispatch_prec<64, 256>(precision, a, st);
  if (F == 32 && H == 128) return dispatch_prec<32, 128>(precision, a, st);
  if (F == 64 && H == 128) return dispatch_prec<64, 128>(precision, a, st);
  if (F == 32 && H == 256) return dispatch_prec<32, 256>(precision, a, st);
  return sailor::set_error(SAILOR_ERR_ARG, "render_rays: (F=%d, H=%d) not instantiated "
                           "(supported: F in {32,64}, H in {128,256})", F, H);
}
