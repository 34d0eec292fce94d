// Front-to-back volume compositing (forward) and its reverse-scan gradient.
//
// Follows c_lib/VoxelEncoding/src/volume_rendering.cu:
//  * volume_rendering_occ_kernel (:666-745): inference, alpha = occupancy,
//    samples already sorted;
//  * volume_rendering_training_forward_kernel (:17-126, alpha = 1-exp(-s*d))
//    and volume_rendering_occ_training_forward_kernel (:129-240);
//  * the two backward kernels (:244-381, :384-522), which re-walk the prefix
//    and use the stored totals for the suffix sums.
// The reference hard-codes 3 feature channels (FEATS_DIM, :13); here the
// channel count F is a runtime argument.  exp uses the IEEE-accurate expf,
// not __expf.
//
// One lane per ray with the samples walked in order keeps the reference's
// exact floating-point recurrence (T *= 1 - alpha, running sums); the
// per-ray rows are contiguous so consecutive lanes stream consecutive rows.
#include <type_traits>

#include "common.h"

namespace {

constexpr int kThreads = 128;
constexpr int kMaxF = 64;  // channels held in registers; larger F loops twice

// Inference compositing over sorted samples.
template <int FMAX>
__global__ void __launch_bounds__(kThreads)
occ_forward_kernel(const float* __restrict__ t, const float* __restrict__ occ,
                   const float* __restrict__ feats, long long n_rays, int S, int F,
                   int f0, float t_thres, float* __restrict__ out_feats,
                   float* __restrict__ out_depth, float* __restrict__ out_ws,
                   float* __restrict__ out_alpha) {
  const long long ray = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (ray >= n_rays) return;
  const float* tr = t + ray * S;
  const float* orow = occ + ray * S;
  const float* frow = feats + ray * S * F;
  const int nf = min(FMAX, F - f0);
  float acc[FMAX];
#pragma unroll
  for (int j = 0; j < FMAX; ++j) acc[j] = 0.0f;
  float T = 1.0f, ws = 0.0f, d = 0.0f, alphas = 0.0f;
  if (!(tr[S - 1] < 0)) {
    for (int i = 0; i < S; ++i) {
      const float z = tr[i];
      if (z < 0) continue;
      const float alpha = orow[i];
      const float w = alpha * T;
#pragma unroll
      for (int j = 0; j < FMAX; ++j)
        if (j < nf) acc[j] += w * frow[(size_t)i * F + f0 + j];
      d += w * z;
      ws += w;
      alphas += w * alpha;
      T *= 1.0f - alpha;
      if (T < t_thres) break;
    }
  }
#pragma unroll
  for (int j = 0; j < FMAX; ++j)
    if (j < nf) out_feats[ray * F + f0 + j] = acc[j];
  if (f0 == 0) {
    out_depth[ray] = d;
    out_ws[ray] = ws;
    out_alpha[ray] = alphas;
  }
}

// The same recurrence with the rows staged through LDS: a wave owns 64 rays,
// each chunk of CS samples of their t / occupancy / feature rows is loaded
// with coalesced reads (16 or 8 lanes per row segment) into odd-stride LDS
// rows, then every lane walks its own ray's chunk in order.  Loads of the
// per-lane layout above touch 64 separate cache lines per instruction; here
// they are contiguous.  F <= 8 channels; both waves of a workgroup walk all
// chunks (one barrier per chunk), lanes whose ray has finished only skip work.
template <int CS>
__global__ void __launch_bounds__(kThreads)
occ_forward_tiled(const float* __restrict__ t, const float* __restrict__ occ,
                  const float* __restrict__ feats, long long n_rays, int S, int F,
                  float t_thres, float* __restrict__ out_feats, float* __restrict__ out_depth,
                  float* __restrict__ out_ws, float* __restrict__ out_alpha) {
  extern __shared__ float lds[];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int TS = CS + 1, FS = CS * F + 1;  // odd row strides: conflict-free lane reads
  float* lt = lds + wave * 64 * (2 * TS + FS);
  float* lo = lt + 64 * TS;
  float* lf = lo + 64 * TS;
  const long long ray0 = ((long long)blockIdx.x * (kThreads / 64) + wave) * 64;
  const long long ray = ray0 + lane;
  const bool live = ray < n_rays;
  float acc[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) acc[j] = 0.0f;
  float T = 1.0f, ws = 0.0f, d = 0.0f, alphas = 0.0f;
  bool done = !live || t[ray * S + S - 1] < 0;
  for (int c0 = 0; c0 < S; c0 += CS) {
    const int cs = min(CS, S - c0);
    __syncthreads();  // the previous chunk's rows are no longer read
    for (int k = lane; k < 64 * CS; k += 64) {
      const int r = k / CS, s = k % CS;
      const long long rr = ray0 + r;
      float tv = -1.0f, ov = 0.0f;
      if (s < cs && rr < n_rays) {
        tv = t[rr * S + c0 + s];
        ov = occ[rr * S + c0 + s];
      }
      lt[r * TS + s] = tv;
      lo[r * TS + s] = ov;
    }
    for (int k = lane; k < 64 * CS * F; k += 64) {
      const int r = k / (CS * F), q = k % (CS * F);
      const long long rr = ray0 + r;
      lf[r * FS + q] = (q < cs * F && rr < n_rays) ? feats[(rr * S + c0) * F + q] : 0.0f;
    }
    __syncthreads();
    if (!done) {
      for (int s = 0; s < cs; ++s) {
        const float z = lt[lane * TS + s];
        if (z < 0) continue;
        const float alpha = lo[lane * TS + s];
        const float w = alpha * T;
#pragma unroll
        for (int j = 0; j < 8; ++j)
          if (j < F) acc[j] += w * lf[lane * FS + s * F + j];
        d += w * z;
        ws += w;
        alphas += w * alpha;
        T *= 1.0f - alpha;
        if (T < t_thres) {
          done = true;
          break;
        }
      }
    }
  }
  if (live) {
#pragma unroll
    for (int j = 0; j < 8; ++j)
      if (j < F) out_feats[ray * F + j] = acc[j];
    out_depth[ray] = d;
    out_ws[ray] = ws;
    out_alpha[ray] = alphas;
  }
}

// Next valid depth after sorted position i (reference :87-97): the last value
// probed is used even if it is not positive, as in the reference loop.
__device__ __forceinline__ float next_depth(const float* tr, const int* idx, int i, int S) {
  float nxt = 0.0f;
  for (int m = i + 1; m < S; ++m) {
    const int k = idx[m];
    nxt = (k >= 0 && k < S) ? tr[k] : -1.0f;
    if (nxt > 0) break;
  }
  return nxt;
}

// One sample of a walk, read ahead of its step: the sort index has to be
// in range (a malformed entry is skipped), then its depth, sigma/occupancy
// and FMAX channel slots (slots past nf repeat the last channel, so every
// read stays in bounds and none needs a branch; for wide FMAX only the live
// channels are read).
template <int FMAX>
struct WalkSample {
  bool ok;
  int si;
  float z, sg, f[FMAX];
};

template <int FMAX>
__device__ __forceinline__ WalkSample<FMAX> read_sample(const float* tr, const float* sr,
                                                        const float* frow, int si, int S, int F,
                                                        int f0, int nf) {
  WalkSample<FMAX> x;
  x.ok = si >= 0 && si < S;
  x.si = x.ok ? si : 0;
  x.z = tr[x.si];
  x.sg = sr[x.si];
  const float* fs = frow + (size_t)x.si * F + f0;
#pragma unroll
  for (int j = 0; j < FMAX; ++j) {
    if constexpr (FMAX <= 8)
      x.f[j] = fs[min(j, nf - 1)];
    else
      x.f[j] = j < nf ? fs[j] : 0.0f;
  }
  return x;
}

// Forward walk of one ray over its sorted samples (reference :17-126 and
// :129-240).  The row pointers may address global memory or an LDS copy.
// The reads run ahead of the recurrence (a step's data one step early, the
// sort index two), so their latency overlaps the previous step's
// arithmetic; the arithmetic and its order are those of the reference.
// Without IDX the samples are in order already (step i is sample i) and
// alpha is the occupancy: that is the inference compositing (reference
// :666-745), whose early-out test on t[S-1] is the same one.
template <int FMAX, bool IDX = true>
__device__ __forceinline__ void forward_walk(const float* tr, const int* ir, const float* sr,
                                             const float* frow, int S, int F, int f0, int nf,
                                             float t_thres, int occ_mode, float (&acc)[FMAX],
                                             float& d, float& ws, float& alphas) {
#pragma unroll
  for (int j = 0; j < FMAX; ++j) acc[j] = 0.0f;
  float T = 1.0f;
  d = ws = alphas = 0.0f;
  auto idx = [&](int i) { if constexpr (IDX) return ir[i]; else return i; };
  const int last = idx(S - 1);
  if (last < 0 || last >= S || tr[last] < 0) return;
  WalkSample<FMAX> nxt = read_sample<FMAX>(tr, sr, frow, idx(0), S, F, f0, nf);
  int si2 = S > 1 ? idx(1) : 0;
  for (int i = 0; i < S; ++i) {
    const WalkSample<FMAX> c = nxt;
    if (i + 1 < S) nxt = read_sample<FMAX>(tr, sr, frow, si2, S, F, f0, nf);
    si2 = i + 2 < S ? idx(i + 2) : 0;
    if (!c.ok) continue;  // malformed permutation entry
    const float z = c.z;
    if (z < 0) continue;
    float alpha;
    if (!IDX || occ_mode) {
      alpha = c.sg;
    } else {
      const float delta = (i < S - 1) ? next_depth(tr, ir, i, S) - z : 1e10f;
      alpha = 1.0f - expf(-c.sg * delta);
    }
    const float w = alpha * T;
#pragma unroll
    for (int j = 0; j < FMAX; ++j)
      if (j < nf) acc[j] += w * c.f[j];
    d += w * z;
    ws += w;
    alphas += w * alpha;
    T *= 1.0f - alpha;
    if (T < t_thres) break;
  }
}

template <int FMAX>
__global__ void __launch_bounds__(kThreads)
train_forward_kernel(const float* __restrict__ t, const int* __restrict__ sidx,
                     const float* __restrict__ sigma, const float* __restrict__ feats,
                     long long n_rays, int S, int F, int f0, float t_thres, int occ_mode,
                     float* __restrict__ out_feats, float* __restrict__ out_depth,
                     float* __restrict__ out_ws, float* __restrict__ out_alpha) {
  const long long ray = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (ray >= n_rays) return;
  const int nf = min(FMAX, F - f0);
  float acc[FMAX], d, ws, alphas;
  forward_walk<FMAX>(t + ray * S, sidx + ray * S, sigma + ray * S, feats + ray * S * F, S, F,
                     f0, nf, t_thres, occ_mode, acc, d, ws, alphas);
#pragma unroll
  for (int j = 0; j < FMAX; ++j)
    if (j < nf) out_feats[ray * F + f0 + j] = acc[j];
  if (f0 == 0) {
    out_depth[ray] = d;
    out_ws[ray] = ws;
    out_alpha[ray] = alphas;
  }
}

// The reference's clip test 1.0 / ((double)(1 - alpha) + 1e-7) <= 15.0
// (:497) decided without the double division away from its boundary:
// x >= 0.0666667 gives 1/x < 14.9999993 and 0 < x <= 0.0666666 gives
// 1/x > 15.0000015, margins far above the division's rounding; the band
// between, x <= 0 and NaN take the division itself.  Agrees with the direct
// form for every float alpha (checked exhaustively on the CPU).
__device__ __forceinline__ bool clip_le15(float alpha) {
  const double x = (double)(1.0f - alpha) + 1e-7;
  if (x >= 0.0666667) return true;
  if (x > 0.0 && x <= 0.0666666) return false;
  return 1.0 / x <= 15.0;
}

// Gradient walk of one ray (reference :244-381, :384-522): re-walks the
// prefix and uses the stored totals for the suffix sums.  Only samples
// before early termination receive a gradient (the sample that ends the walk
// is skipped, :346); entries never written keep what the caller put there.
// Reads run ahead of the recurrence as in forward_walk.
template <int FMAX>
__device__ __forceinline__ void backward_walk(const float* tr, const int* ir, const float* sr,
                                              const float* frow, const float* gf,
                                              const float* of, float gd, float gw, float ga,
                                              float d_final, float ws_final, float a_final,
                                              int S, int F, float t_thres, int occ_mode,
                                              float* gfo, float* gso) {
  const int last = ir[S - 1];
  if (last < 0 || last >= S || tr[last] < 0) return;
  // per-ray output gradient and totals of the features: registers when
  // FMAX is small, read in place otherwise
  float gfr[FMAX <= 8 ? FMAX : 1], ofr[FMAX <= 8 ? FMAX : 1];
  if constexpr (FMAX <= 8) {
#pragma unroll
    for (int j = 0; j < FMAX; ++j) {
      gfr[j] = gf[min(j, F - 1)];
      ofr[j] = of[min(j, F - 1)];
    }
  }
  auto gfj = [&](int j) { if constexpr (FMAX <= 8) return gfr[j]; else return gf[j]; };
  auto ofj = [&](int j) { if constexpr (FMAX <= 8) return ofr[j]; else return of[j]; };
  // running prefix of the features; FMAX is a compile-time bound so the
  // prefix stays in registers
  float pf[FMAX];
#pragma unroll
  for (int j = 0; j < FMAX; ++j) pf[j] = 0.0f;
  float T = 1.0f, ws = 0.0f, d = 0.0f, alphas = 0.0f;
  WalkSample<FMAX> nxt = read_sample<FMAX>(tr, sr, frow, ir[0], S, F, 0, F);
  int si2 = S > 1 ? ir[1] : 0;
  for (int i = 0; i < S; ++i) {
    const WalkSample<FMAX> c = nxt;
    if (i + 1 < S) nxt = read_sample<FMAX>(tr, sr, frow, si2, S, F, 0, F);
    si2 = i + 2 < S ? ir[i + 2] : 0;
    if (!c.ok) continue;  // malformed permutation entry
    const int si = c.si;
    const float z = c.z;
    if (z < 0) continue;
    float alpha, delta = 0.0f;
    if (occ_mode) {
      alpha = c.sg;
    } else {
      delta = (i < S - 1) ? next_depth(tr, ir, i, S) - z : 1e10f;
      alpha = 1.0f - expf(-c.sg * delta);
    }
    const float w = alpha * T;
#pragma unroll
    for (int j = 0; j < FMAX; ++j)
      if (j < F) pf[j] += w * c.f[j];
    d += w * z;
    ws += w;
    alphas += w * alpha;
    T *= 1.0f - alpha;
    if (T < t_thres) break;
#pragma unroll
    for (int j = 0; j < FMAX; ++j)
      if (j < F) gfo[(size_t)si * F + j] = gfj(j) * w;
    float gsf = 0.0f;
#pragma unroll
    for (int j = 0; j < FMAX; ++j)
      if (j < F) gsf += gfj(j) * (T * c.f[j] - (ofj(j) - pf[j]));
    float g;
    if (occ_mode) {
      // clipped 1/(1-alpha) (:497); the test is evaluated in double there
      const float tmp = clip_le15(alpha) ? 1.0f / (1.0f - alpha) : 15.0f;
      const float g_ws_ = gw * (T * 1.0f - (ws_final - ws)) * tmp;
      const float g_d = gd * (T * z - (d_final - d)) * tmp;
      const float g_a = ga * (2.0f * T * alpha - (a_final - alphas)) * tmp;
      g = g_ws_ + g_d + gsf * tmp + g_a;
    } else {
      const float tmp = gw * (1.0f - ws_final);
      const float g_ws_ = delta * (gw * (T * 1.0f - (ws_final - ws)) + tmp);
      const float g_d = delta * (gd * (T * z - (d_final - d)) + tmp);
      const float g_a = delta * (ga * (2.0f * T * alpha - (a_final - alphas)) + tmp);
      const float g_f = delta * (gsf + tmp);
      g = g_ws_ + g_d + g_f + g_a;
    }
    gso[si] = g;
  }
}

// Per-lane global form of the gradient; the grad outputs are zero-filled
// beforehand by zero2.
template <int FMAX>
__global__ void __launch_bounds__(kThreads)
train_backward_kernel(const float* __restrict__ g_feats, const float* __restrict__ g_depth,
                      const float* __restrict__ g_ws, const float* __restrict__ g_alpha,
                      const float* __restrict__ o_feats, const float* __restrict__ o_depth,
                      const float* __restrict__ o_ws, const float* __restrict__ o_alpha,
                      const float* __restrict__ t, const int* __restrict__ sidx,
                      const float* __restrict__ sigma, const float* __restrict__ feats,
                      long long n_rays, int S, int F, float t_thres, int occ_mode,
                      float* __restrict__ grad_feats, float* __restrict__ grad_sigma) {
  const long long ray = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (ray >= n_rays) return;
  backward_walk<FMAX>(t + ray * S, sidx + ray * S, sigma + ray * S, feats + ray * S * F,
                      g_feats + ray * F, o_feats + ray * F, g_depth[ray], g_ws[ray],
                      g_alpha[ray], o_depth[ray], o_ws[ray], o_alpha[ray], S, F, t_thres,
                      occ_mode, grad_feats + ray * S * F, grad_sigma + ray * S);
}

__global__ void zero2(float* __restrict__ a, long long na, float* __restrict__ b,
                      long long nb) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < na + nb;
       i += (long long)gridDim.x * blockDim.x) {
    if (i < na) a[i] = 0.0f;
    else b[i - na] = 0.0f;
  }
}

// Value of lane Q of each aligned group of 4 lanes (DPP quad_perm).
template <int Q>
__device__ __forceinline__ float quad_lane(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), Q * 0x55, 0xf, 0xf, false));
}

// backward_walk with a ray's channels split over a group of LPR lanes (the
// lane at `base` carries channel 0): every lane of the group runs the same
// recurrence on the same samples, so T, w and the running sums are
// identical in all of them; lane j keeps channel j's prefix and gradient
// (lanes j >= F repeat channel F-1 and write nothing).  The channel sum of
// the sigma gradient is taken in channel order from the group's lanes, as
// the reference's loop adds it; lane 0 of the group writes it.  All lanes of
// a group take the same branches, so the shuffles always find their source
// lane active.
template <int LPR>
__device__ __forceinline__ void backward_walk_split(
    const float* tr, const int* ir, const float* sr, const float* frow, int j, int base,
    const float* gf, const float* of, float gd, float gw, float ga, float d_final,
    float ws_final, float a_final, int S, int F, float t_thres, int occ_mode, float* gfo,
    float* gso) {
  const int last = ir[S - 1];
  if (last < 0 || last >= S || tr[last] < 0) return;
  const int jc = min(j, F - 1);
  const bool own = j < F;
  const float gfj = gf[jc], ofj = of[jc];
  float pf = 0.0f;
  float T = 1.0f, ws = 0.0f, d = 0.0f, alphas = 0.0f;
  WalkSample<1> nxt = read_sample<1>(tr, sr, frow, ir[0], S, F, jc, 1);
  int si2 = S > 1 ? ir[1] : 0;
  for (int i = 0; i < S; ++i) {
    const WalkSample<1> c = nxt;
    if (i + 1 < S) nxt = read_sample<1>(tr, sr, frow, si2, S, F, jc, 1);
    si2 = i + 2 < S ? ir[i + 2] : 0;
    if (!c.ok) continue;  // malformed permutation entry
    const int si = c.si;
    const float z = c.z;
    if (z < 0) continue;
    float alpha, delta = 0.0f;
    if (occ_mode) {
      alpha = c.sg;
    } else {
      delta = (i < S - 1) ? next_depth(tr, ir, i, S) - z : 1e10f;
      alpha = 1.0f - expf(-c.sg * delta);
    }
    const float w = alpha * T;
    pf += w * c.f[0];
    d += w * z;
    ws += w;
    alphas += w * alpha;
    T *= 1.0f - alpha;
    if (T < t_thres) break;
    if (own) gfo[(size_t)si * F + j] = gfj * w;
    const float term = gfj * (T * c.f[0] - (ofj - pf));
    float gsf = 0.0f;
    if constexpr (LPR == 4) {
      // groups of 4 are the hardware's lane quads: broadcast by DPP
      const float t0 = quad_lane<0>(term), t1 = quad_lane<1>(term);
      const float t2 = quad_lane<2>(term), t3 = quad_lane<3>(term);
      gsf += t0;
      if (1 < F) gsf += t1;
      if (2 < F) gsf += t2;
      if (3 < F) gsf += t3;
    } else {
#pragma unroll
      for (int q = 0; q < LPR; ++q) {
        const float tq = __shfl(term, base + q, 64);
        if (q < F) gsf += tq;
      }
    }
    float g;
    if (occ_mode) {
      // clipped 1/(1-alpha) (:497); the test is evaluated in double there
      const float tmp = clip_le15(alpha) ? 1.0f / (1.0f - alpha) : 15.0f;
      const float g_ws_ = gw * (T * 1.0f - (ws_final - ws)) * tmp;
      const float g_d = gd * (T * z - (d_final - d)) * tmp;
      const float g_a = ga * (2.0f * T * alpha - (a_final - alphas)) * tmp;
      g = g_ws_ + g_d + gsf * tmp + g_a;
    } else {
      const float tmp = gw * (1.0f - ws_final);
      const float g_ws_ = delta * (gw * (T * 1.0f - (ws_final - ws)) + tmp);
      const float g_d = delta * (gd * (T * z - (d_final - d)) + tmp);
      const float g_a = delta * (ga * (2.0f * T * alpha - (a_final - alphas)) + tmp);
      const float g_f = delta * (gsf + tmp);
      g = g_ws_ + g_d + g_f + g_a;
    }
    if (j == 0) gso[si] = g;
  }
}

// ---- Streamed LDS forms (F <= 8) -------------------------------------------
// The walk visits a ray's samples in sort-index order, so a row cannot be
// consumed in chunks.  A persistent one-wave workgroup owns tiles of RB rays:
// a tile's contiguous rows (t | sort index | sigma, S words per ray each, then
// the features, S*F per ray) are read with coalesced loads into registers
// (a few words per lane and array, each array unrolled on its own so the
// loop body stays small), stored into odd-stride LDS rows (lanes reading the same
// sample slot hit distinct banks), and each ray is walked there by a group
// of 4 (F <= 4) or 8 lanes, one channel per lane, so a 16-ray tile fills the
// wave.  The
// loads of the wave's next tile are issued before the walk of the current
// one, so memory latency overlaps the walk.  The gradient form stages the
// per-ray inputs as well (its walk touches no global memory) and builds its
// output rows in LDS, zeroed again as they are written out, so it needs no
// separate zero-fill pass.

constexpr int kTileThreads = 64;

// q = k / L for 0 <= k < 2^24 via the float reciprocal: |k * invL - k/L| < 1,
// so one correction step gives the exact quotient.
__device__ __forceinline__ int div_small(int k, int L, float invL) {
  int q = (int)((float)k * invL);
  const int rem = k - q * L;
  q += (rem >= L) - (rem < 0);
  return q;
}

__host__ __device__ constexpr int odd_stride(int n) { return n | 1; }

struct TileShape {
  int S, SF, TS, FS;  // row lengths and LDS row strides
  float inv_s, inv_sf;
};

__device__ __forceinline__ TileShape tile_shape(int S, int F) {
  TileShape g;
  g.S = S;
  g.SF = S * F;
  g.TS = odd_stride(S);
  g.FS = odd_stride(S * F);
  g.inv_s = 1.0f / (float)S;
  g.inv_sf = 1.0f / (float)(S * F);
  return g;
}

// The lane index as a value the compiler cannot see through: the per-word
// tile map is then recomputed per tile instead of being hoisted out of the
// persistent loop, where it would hold several registers per word.
__device__ __forceinline__ int opaque_lane(int lane) {
  asm volatile("" : "+v"(lane));
  return lane;
}

// The next tile's words in registers: t, sort index and sigma (ETS words
// per lane each, word e = lane + 64 u of the tile's RB*S) and the features
// (ETF words per lane of RB*S*F).
template <int ETS, int ETF>
struct TileRegs {
  unsigned t[ETS], i[ETS], s[ETS], f[ETF];
};

// Issue the loads of the tile starting at ray ray0 with nr rays.  Words past
// the tile read its first word instead (always in bounds) and are not stored.
template <int ETS, int ETF, bool IDX>
__device__ __forceinline__ void tile_load(const TileShape& g, const unsigned* __restrict__ t,
                                          const unsigned* __restrict__ sidx,
                                          const unsigned* __restrict__ sigma,
                                          const unsigned* __restrict__ feats, long long ray0,
                                          int nr, int lane, TileRegs<ETS, ETF>& v) {
  const unsigned* bt = t + ray0 * g.S;
  const unsigned* bi = IDX ? sidx + ray0 * g.S : t;
  const unsigned* bs = sigma + ray0 * g.S;
  const unsigned* bf = feats + ray0 * g.SF;
  const int nrs = nr * g.S, nrf = nr * g.SF;
  const int ln = opaque_lane(lane);
#pragma unroll
  for (int u = 0; u < ETS; ++u) {
    const int e = ln + u * kTileThreads;
    const int k = e < nrs ? e : 0;
    v.t[u] = bt[k];
    if constexpr (IDX) v.i[u] = bi[k];
    v.s[u] = bs[k];
  }
#pragma unroll
  for (int u = 0; u < ETF; ++u) {
    const int e = ln + u * kTileThreads;
    v.f[u] = bf[e < nrf ? e : 0];
  }
}

// Store a loaded tile into the LDS rows: t | sort index | sigma (RB rows of
// stride TS each), then the features (RB rows of stride FS).
template <int ETS, int ETF, bool IDX>
__device__ __forceinline__ void tile_store(const TileShape& g, int RB, int nr, int lane,
                                           const TileRegs<ETS, ETF>& v, unsigned* lds) {
  const int nrs = nr * g.S, nrf = nr * g.SF;
  const int ln = opaque_lane(lane);
  unsigned* lt = lds;
  unsigned* li = lds + RB * g.TS;
  unsigned* ls = lds + 2 * RB * g.TS;
  unsigned* lf = lds + 3 * RB * g.TS;
#pragma unroll
  for (int u = 0; u < ETS; ++u) {
    const int e = ln + u * kTileThreads;
    if (e < nrs) {
      const int r = div_small(e, g.S, g.inv_s);
      const int w = r * g.TS + (e - r * g.S);
      lt[w] = v.t[u];
      if constexpr (IDX) li[w] = v.i[u];
      ls[w] = v.s[u];
    }
  }
#pragma unroll
  for (int u = 0; u < ETF; ++u) {
    const int e = ln + u * kTileThreads;
    if (e < nrf) {
      const int r = div_small(e, g.SF, g.inv_sf);
      lf[r * g.FS + (e - r * g.SF)] = v.f[u];
    }
  }
}

size_t train_forward_tile_lds(int RB, int S, int F) {
  return (size_t)RB * (3 * odd_stride(S) + odd_stride(S * F)) * sizeof(float);
}

size_t train_backward_tile_lds(int RB, int S, int F) {
  return (size_t)RB * (4 * odd_stride(S) + 2 * odd_stride(S * F) + 2 * F + 6) * sizeof(float);
}

// Waves per SIMD the streamed forward kernels are compiled for at least (a
// register budget of 512 / waves): what their LDS tiles leave resident, about
// two 16-ray or four 8-ray tiles per SIMD.  The gradient form's larger tiles
// leave two or three per SIMD; it is compiled for two.
constexpr int kStreamWaves(int RB) { return RB >= 16 ? 2 : 4; }

template <int RB, int LPR, int ETS, int ETF, bool IDX>
__global__ void __launch_bounds__(kTileThreads)
__attribute__((amdgpu_waves_per_eu(kStreamWaves(RB))))
train_forward_stream(const float* __restrict__ t, const int* __restrict__ sidx,
                     const float* __restrict__ sigma, const float* __restrict__ feats,
                     long long n_rays, int S, int F, float t_thres, int occ_mode,
                     float* __restrict__ out_feats, float* __restrict__ out_depth,
                     float* __restrict__ out_ws, float* __restrict__ out_alpha) {
  extern __shared__ unsigned lds_u[];
  const int lane = threadIdx.x;
  const TileShape g = tile_shape(S, F);
  const float* lt = reinterpret_cast<const float*>(lds_u);
  const int* li = reinterpret_cast<const int*>(lds_u + RB * g.TS);
  const float* ls = reinterpret_cast<const float*>(lds_u + 2 * RB * g.TS);
  const float* lf = reinterpret_cast<const float*>(lds_u + 3 * RB * g.TS);
  const long long n_tiles = (n_rays + RB - 1) / RB;
  long long tile = blockIdx.x;
  if (tile >= n_tiles) return;
  auto rays_in = [&](long long tl) { return (int)min((long long)RB, n_rays - tl * RB); };
  const unsigned* ut = reinterpret_cast<const unsigned*>(t);
  const unsigned* ui = reinterpret_cast<const unsigned*>(sidx);
  const unsigned* us = reinterpret_cast<const unsigned*>(sigma);
  const unsigned* uf = reinterpret_cast<const unsigned*>(feats);
  TileRegs<ETS, ETF> v;
  int nr = rays_in(tile);
  tile_load<ETS, ETF, IDX>(g, ut, ui, us, uf, tile * RB, nr, lane, v);
  for (;;) {
    tile_store<ETS, ETF, IDX>(g, RB, nr, lane, v, lds_u);
    __syncthreads();
    const int nr_cur = nr;
    const long long next = tile + gridDim.x;
    if (next < n_tiles) {
      nr = rays_in(next);
      tile_load<ETS, ETF, IDX>(g, ut, ui, us, uf, next * RB, nr, lane, v);
    }
    // LPR lanes per ray, lane j of a group accumulating channel j
    const int j = lane % LPR;
    for (int r = lane / LPR; r - lane / LPR < nr_cur; r += kTileThreads / LPR) {
      if (r >= nr_cur) continue;
      float acc[1], d, ws, alphas;
      forward_walk<1, IDX>(lt + r * g.TS, li + r * g.TS, ls + r * g.TS, lf + r * g.FS, S, F,
                      min(j, F - 1), 1, t_thres, occ_mode, acc, d, ws, alphas);
      const long long ray = tile * RB + r;
      if (j < F) out_feats[ray * F + j] = acc[0];
      if (j == 0) {
        out_depth[ray] = d;
        out_ws[ray] = ws;
        out_alpha[ray] = alphas;
      }
    }
    if (next >= n_tiles) break;
    __syncthreads();  // the walk's LDS reads are done before the next stores
    tile = next;
  }
}

template <int RB, int LPR, int ETS, int ETF>
__global__ void __launch_bounds__(kTileThreads)
__attribute__((amdgpu_waves_per_eu(2)))
train_backward_stream(const float* __restrict__ g_feats, const float* __restrict__ g_depth,
                      const float* __restrict__ g_ws, const float* __restrict__ g_alpha,
                      const float* __restrict__ o_feats, const float* __restrict__ o_depth,
                      const float* __restrict__ o_ws, const float* __restrict__ o_alpha,
                      const float* __restrict__ t, const int* __restrict__ sidx,
                      const float* __restrict__ sigma, const float* __restrict__ feats,
                      long long n_rays, int S, int F, float t_thres, int occ_mode,
                      float* __restrict__ grad_feats, float* __restrict__ grad_sigma) {
  extern __shared__ unsigned lds_u[];
  const int lane = threadIdx.x;
  const TileShape g = tile_shape(S, F);
  const float* lt = reinterpret_cast<const float*>(lds_u);
  const int* li = reinterpret_cast<const int*>(lds_u + RB * g.TS);
  const float* ls = reinterpret_cast<const float*>(lds_u + 2 * RB * g.TS);
  const float* lf = reinterpret_cast<const float*>(lds_u + 3 * RB * g.TS);
  float* lgs = reinterpret_cast<float*>(lds_u + 3 * RB * g.TS + RB * g.FS);
  float* lgf = lgs + RB * g.TS;
  // per ray: g_feats (F), out_feats (F), then g_depth, g_ws, g_alpha,
  // out_depth, out_ws, out_alpha
  const int PR = 2 * F + 6;
  float* lpr = lgf + RB * g.FS;
  const long long n_tiles = (n_rays + RB - 1) / RB;
  long long tile = blockIdx.x;
  if (tile >= n_tiles) return;
  auto rays_in = [&](long long tl) { return (int)min((long long)RB, n_rays - tl * RB); };
  const unsigned* ut = reinterpret_cast<const unsigned*>(t);
  const unsigned* ui = reinterpret_cast<const unsigned*>(sidx);
  const unsigned* us = reinterpret_cast<const unsigned*>(sigma);
  const unsigned* uf = reinterpret_cast<const unsigned*>(feats);
  for (int k = lane; k < RB * (g.TS + g.FS); k += kTileThreads) lgs[k] = 0.0f;
  TileRegs<ETS, ETF> v;
  float pr[22];
  auto load_ray = [&](long long tl, int n) {
    const long long ray = tl * RB + min(lane, n - 1);
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      pr[j] = j < F ? g_feats[ray * F + j] : 0.0f;
      pr[8 + j] = j < F ? o_feats[ray * F + j] : 0.0f;
    }
    pr[16] = g_depth[ray];
    pr[17] = g_ws[ray];
    pr[18] = g_alpha[ray];
    pr[19] = o_depth[ray];
    pr[20] = o_ws[ray];
    pr[21] = o_alpha[ray];
  };
  int nr = rays_in(tile);
  tile_load<ETS, ETF, true>(g, ut, ui, us, uf, tile * RB, nr, lane, v);
  load_ray(tile, nr);
  for (;;) {
    tile_store<ETS, ETF, true>(g, RB, nr, lane, v, lds_u);
    if (lane < nr) {
      float* p = lpr + lane * PR;
#pragma unroll
      for (int j = 0; j < 8; ++j)
        if (j < F) {
          p[j] = pr[j];
          p[F + j] = pr[8 + j];
        }
#pragma unroll
      for (int q = 0; q < 6; ++q) p[2 * F + q] = pr[16 + q];
    }
    __syncthreads();
    const long long ray0 = tile * RB;
    const int nr_cur = nr;
    const long long next = tile + gridDim.x;
    if (next < n_tiles) {
      nr = rays_in(next);
      tile_load<ETS, ETF, true>(g, ut, ui, us, uf, next * RB, nr, lane, v);
      load_ray(next, nr);
    }
    // LPR lanes per ray, lane j of a group carrying channel j
    const int j = lane % LPR;
    for (int r = lane / LPR; r - lane / LPR < nr_cur; r += kTileThreads / LPR) {
      if (r >= nr_cur) continue;
      const float* p = lpr + r * PR;
      backward_walk_split<LPR>(lt + r * g.TS, li + r * g.TS, ls + r * g.TS, lf + r * g.FS, j,
                               lane - j, p, p + F, p[2 * F], p[2 * F + 1], p[2 * F + 2],
                               p[2 * F + 3], p[2 * F + 4], p[2 * F + 5], S, F, t_thres,
                               occ_mode, lgf + r * g.FS, lgs + r * g.TS);
    }
    __syncthreads();
    // write the gradient rows out and leave them zeroed for the next tile
    for (int k = lane; k < nr_cur * g.S; k += kTileThreads) {
      const int r = div_small(k, g.S, g.inv_s);
      const int w = r * g.TS + (k - r * g.S);
      grad_sigma[ray0 * S + k] = lgs[w];
      lgs[w] = 0.0f;
    }
    for (int k = lane; k < nr_cur * g.SF; k += kTileThreads) {
      const int r = div_small(k, g.SF, g.inv_sf);
      const int w = r * g.FS + (k - r * g.SF);
      grad_feats[ray0 * g.SF + k] = lgf[w];
      lgf[w] = 0.0f;
    }
    if (next >= n_tiles) break;
    tile = next;
  }
}

// Streamed launch shapes: rays per tile RB with ETS / ETF register words per
// lane, taken in order, the first whose tile fits (RB * S <= 64 ETS,
// RB * S * F <= 64 ETF) and the LDS budget; none fits -> the per-lane global
// form.
constexpr size_t kTileLdsBudget = 64 * 1024;
struct StreamShape {
  int rb, ets, etf;
};
constexpr StreamShape kStreamShapes[] = {{16, 12, 36}, {8, 12, 36}, {4, 12, 36}};

int pick_stream(size_t (*lds_of)(int, int, int), int S, int F) {
  if (F > 8) return -1;
  for (int i = 0; i < (int)(sizeof(kStreamShapes) / sizeof(kStreamShapes[0])); ++i) {
    const StreamShape& sh = kStreamShapes[i];
    if ((long long)sh.rb * S <= 64LL * sh.ets && (long long)sh.rb * S * F <= 64LL * sh.etf &&
        lds_of(sh.rb, S, F) <= kTileLdsBudget)
      return i;
  }
  return -1;
}

// Calls launch(RB, ETS, ETF) with table entry i's shape as compile-time
// constants.
template <int I, typename L>
int dispatch_stream_from(int i, L& launch) {
  if constexpr (I < (int)(sizeof(kStreamShapes) / sizeof(kStreamShapes[0]))) {
    if (i == I)
      return launch(std::integral_constant<int, kStreamShapes[I].rb>{},
                    std::integral_constant<int, kStreamShapes[I].ets>{},
                    std::integral_constant<int, kStreamShapes[I].etf>{});
    return dispatch_stream_from<I + 1>(i, launch);
  }
  return SAILOR_ERR_ARG;
}

template <typename L>
int dispatch_stream(int i, L& launch) {
  return dispatch_stream_from<0>(i, launch);
}

// Persistent grid of a streamed kernel: its resident workgroups per CU at
// this LDS size times the CU count, at most one per tile.
template <typename K>
int stream_grid(K kern, size_t lds, long long n_tiles, int* grid) {
  SAILOR_HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                     hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)kTileLdsBudget));
  int dev = 0, cus = 0, per_cu = 0;
  SAILOR_HIP_TRY(hipGetDevice(&dev));
  SAILOR_HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  SAILOR_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, kTileThreads, lds));
  *grid = (int)std::min<long long>(n_tiles, (long long)std::max(per_cu, 1) * cus);
  return SAILOR_OK;
}

// IDX false: samples already in order, no sort index (inference compositing).
template <int RB, int ETS, int ETF, bool IDX>
int launch_forward_stream(const float* t, const int32_t* sort_idx, const float* sigma,
                          const float* feats, long long n, int S, int F, float t_thresh,
                          int support_occ, float* out_feats, float* out_depth, float* out_ws,
                          float* out_alpha, hipStream_t st, const char* what) {
  auto kern = F <= 4 ? train_forward_stream<RB, 4, ETS, ETF, IDX>
                     : train_forward_stream<RB, 8, ETS, ETF, IDX>;
  const size_t lds = train_forward_tile_lds(RB, S, F);
  int grid = 0;
  if (int e = stream_grid(kern, lds, (n + RB - 1) / RB, &grid)) return e;
  hipLaunchKernelGGL(kern, dim3(grid), dim3(kTileThreads), lds, st, t, sort_idx, sigma, feats, n,
                     S, F, t_thresh, support_occ, out_feats, out_depth, out_ws, out_alpha);
  return sailor::check_launch(what);
}

template <int RB, int ETS, int ETF>
int launch_backward_stream(const float* g_feats, const float* g_depth, const float* g_ws,
                           const float* g_alpha, const float* out_feats, const float* out_depth,
                           const float* out_ws, const float* out_alpha, const float* t,
                           const int32_t* sort_idx, const float* sigma, const float* feats,
                           long long n, int S, int F, float t_thresh, int support_occ,
                           float* grad_feats, float* grad_sigma, hipStream_t st) {
  auto kern = F <= 4 ? train_backward_stream<RB, 4, ETS, ETF>
                     : train_backward_stream<RB, 8, ETS, ETF>;
  const size_t lds = train_backward_tile_lds(RB, S, F);
  int grid = 0;
  if (int e = stream_grid(kern, lds, (n + RB - 1) / RB, &grid)) return e;
  hipLaunchKernelGGL(kern, dim3(grid), dim3(kTileThreads), lds, st, g_feats, g_depth, g_ws,
                     g_alpha, out_feats, out_depth, out_ws, out_alpha, t, sort_idx, sigma, feats,
                     n, S, F, t_thresh, support_occ, grad_feats, grad_sigma);
  return sailor::check_launch("volume_rendering_training_backward");
}

}  // namespace

extern "C" int sailor_volume_rendering_occ_forward(const float* t, const float* occ,
                                                   const float* feats, int B, int R, int S,
                                                   int F, float t_thresh, float* out_feats,
                                                   float* out_depth, float* out_ws,
                                                   float* out_alpha, void* stream) {
  SAILOR_REQUIRE(B > 0 && R >= 0 && S > 0 && F > 0, "volume_rendering_occ_forward: bad sizes");
  const long long n = (long long)B * R;
  if (n == 0) return SAILOR_OK;
  hipStream_t st = sailor::as_stream(stream);
  // rows that fit a tile: the streamed walk without a sort index
  const int sh = pick_stream(train_forward_tile_lds, S, F);
  if (sh >= 0) {
    auto launch = [&](auto rb, auto ets, auto etf) {
      return launch_forward_stream<decltype(rb)::value, decltype(ets)::value,
                                   decltype(etf)::value, false>(
          t, nullptr, occ, feats, n, S, F, t_thresh, 1, out_feats, out_depth, out_ws, out_alpha,
          st, "volume_rendering_occ_forward");
    };
    return dispatch_stream(sh, launch);
  }
  if (F <= 8) {
    // longer rows: chunks of 8 samples staged in LDS (0.21 ms vs 0.29 ms with
    // 16-sample chunks at 262,144 rays x 48 samples x 3 channels)
    constexpr int cs = 8;
    const size_t lds = (size_t)(kThreads / 64) * 64 * (2 * (cs + 1) + cs * F + 1) * sizeof(float);
    const dim3 grid((unsigned)sailor::ceil_div(n, kThreads));
    hipLaunchKernelGGL(occ_forward_tiled<cs>, grid, dim3(kThreads), lds, st, t, occ, feats, n, S,
                       F, t_thresh, out_feats, out_depth, out_ws, out_alpha);
    return sailor::check_launch("volume_rendering_occ_forward");
  }
  for (int f0 = 0; f0 < F; f0 += 8) {
    hipLaunchKernelGGL(occ_forward_kernel<8>, dim3(sailor::ceil_div(n, kThreads)),
                       dim3(kThreads), 0, st, t, occ, feats, n, S, F, f0, t_thresh,
                       out_feats, out_depth, out_ws, out_alpha);
  }
  return sailor::check_launch("volume_rendering_occ_forward");
}

extern "C" int sailor_volume_rendering_training_forward(
    const float* t, const int32_t* sort_idx, const float* sigma, const float* feats,
    int B, int R, int S, int F, float t_thresh, int support_occ, float* out_feats,
    float* out_depth, float* out_ws, float* out_alpha, void* stream) {
  SAILOR_REQUIRE(B > 0 && R >= 0 && S > 0 && F > 0, "volume_rendering_training_forward: bad sizes");
  const long long n = (long long)B * R;
  if (n == 0) return SAILOR_OK;
  hipStream_t st = sailor::as_stream(stream);
  const int sh = pick_stream(train_forward_tile_lds, S, F);
  if (sh >= 0) {
    auto launch = [&](auto rb, auto ets, auto etf) {
      return launch_forward_stream<decltype(rb)::value, decltype(ets)::value,
                                   decltype(etf)::value, true>(
          t, sort_idx, sigma, feats, n, S, F, t_thresh, support_occ, out_feats, out_depth,
          out_ws, out_alpha, st, "volume_rendering_training_forward");
    };
    return dispatch_stream(sh, launch);
  }
  for (int f0 = 0; f0 < F; f0 += 8) {
    hipLaunchKernelGGL(train_forward_kernel<8>, dim3(sailor::ceil_div(n, kThreads)),
                       dim3(kThreads), 0, st, t, sort_idx, sigma, feats, n, S, F, f0,
                       t_thresh, support_occ, out_feats, out_depth, out_ws, out_alpha);
  }
  return sailor::check_launch("volume_rendering_training_forward");
}

extern "C" int sailor_volume_rendering_training_backward(
    const float* g_feats, const float* g_depth, const float* g_ws, const float* g_alpha,
    const float* out_feats, const float* out_depth, const float* out_ws,
    const float* out_alpha, const float* t, const int32_t* sort_idx, const float* sigma,
    const float* feats, int B, int R, int S, int F, float t_thresh, int support_occ,
    float* grad_feats, float* grad_sigma, void* stream) {
  SAILOR_REQUIRE(B > 0 && R >= 0 && S > 0 && F > 0, "volume_rendering_training_backward: bad sizes");
  SAILOR_REQUIRE(F <= kMaxF, "volume_rendering_training_backward: at most %d channels", kMaxF);
  const long long n = (long long)B * R;
  if (n == 0) return SAILOR_OK;
  hipStream_t st = sailor::as_stream(stream);
  const int sh = pick_stream(train_backward_tile_lds, S, F);
  if (sh >= 0) {
    auto launch = [&](auto rb, auto ets, auto etf) {
      return launch_backward_stream<decltype(rb)::value, decltype(ets)::value,
                                    decltype(etf)::value>(
          g_feats, g_depth, g_ws, g_alpha, out_feats, out_depth, out_ws, out_alpha, t, sort_idx,
          sigma, feats, n, S, F, t_thresh, support_occ, grad_feats, grad_sigma, st);
    };
    return dispatch_stream(sh, launch);
  }
  const long long na = n * S * F, nb = n * S;
  int zb = (int)std::min<long long>(4096, (na + nb + 255) / 256);
  hipLaunchKernelGGL(zero2, dim3(zb), dim3(256), 0, st, grad_feats, na, grad_sigma, nb);
  if (F <= 8)
    hipLaunchKernelGGL(train_backward_kernel<8>, dim3(sailor::ceil_div(n, kThreads)),
                       dim3(kThreads), 0, st, g_feats, g_depth, g_ws, g_alpha, out_feats,
                       out_depth, out_ws, out_alpha, t, sort_idx, sigma, feats, n, S, F,
                       t_thresh, support_occ, grad_feats, grad_sigma);
  else
    hipLaunchKernelGGL(train_backward_kernel<kMaxF>, dim3(sailor::ceil_div(n, kThreads)),
                       dim3(kThreads), 0, st, g_feats, g_depth, g_ws, g_alpha, out_feats,
                       out_depth, out_ws, out_alpha, t, sort_idx, sigma, feats, n, S, F,
                       t_thresh, support_occ, grad_feats, grad_sigma);
  return sailor::check_launch("volume_rendering_training_backward");
}
