// Frequency (positional) encoding y = [x, sin(2^f x), sin(2^f x + pi/2), ...]
// and its gradient, following c_lib/VoxelEncoding/src/freq_encoding.cu:16-86
// (output layout o = D + col*D + d with col = 2f + {0: sin, 1: cos}).  Uses
// the accurate sinf instead of __sinf.
#include "common.h"

namespace {

constexpr int kThreads = 256;

__global__ void freq_fwd(const float* __restrict__ x, long long N, int D, int deg,
                         float* __restrict__ y) {
  const int C = D + D * deg * 2;
  const float half_pi = 3.141592653589793f / 2.0f;
  for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < N * C;
       k += (long long)gridDim.x * blockDim.x) {
    const long long n = k / C;
    const int o = (int)(k - n * C);
    const float* in = x + n * D;
    if (o < D) {
      y[k] = in[o];
    } else {
      const int d = o % D;
      const int col = o / D - 1;
      const int f = col / 2;
      const float phase = (float)(col % 2) * half_pi;
      y[k] = sinf(scalbnf(in[d], f) + phase);
    }
  }
}

// Tiled form: a block owns kFreqPts consecutive points, stages their inputs
// in LDS and writes its contiguous block of outputs with coalesced stores;
// the index splits use the float reciprocal (exact for k < 2^24) instead of
// 64-bit division.  Same values as freq_fwd: the phase is added as
// (col % 2) * pi/2 exactly as there.
constexpr int kFreqPts = 256;
constexpr int kFreqMaxD = 8;

__device__ __forceinline__ int div_small(int k, int L, float invL) {
  int q = (int)((float)k * invL);
  const int rem = k - q * L;
  q += (rem >= L) - (rem < 0);
  return q;
}

__global__ void __launch_bounds__(kThreads)
freq_fwd_tiled(const float* __restrict__ x, long long N, int D, int deg, float* __restrict__ y) {
  __shared__ float xs[kFreqPts * kFreqMaxD];
  const int C = D + D * deg * 2;
  const float half_pi = 3.141592653589793f / 2.0f;
  const long long n0 = (long long)blockIdx.x * kFreqPts;
  const int np = (int)min((long long)kFreqPts, N - n0);
  for (int i = threadIdx.x; i < np * D; i += kThreads) xs[i] = x[n0 * D + i];
  __syncthreads();
  const float invC = 1.0f / (float)C, invD = 1.0f / (float)D;
  float* out = y + n0 * C;
  const int tot = np * C;
  for (int k = threadIdx.x; k < tot; k += kThreads) {
    const int n = div_small(k, C, invC);
    const int o = k - n * C;
    const float* in = xs + n * D;
    float v;
    if (o < D) {
      v = in[o];
    } else {
      const int col = div_small(o - D, D, invD);
      const int d = o - D - col * D;
      const float phase = (float)(col % 2) * half_pi;
      v = sinf(scalbnf(in[d], col / 2) + phase);
    }
    out[k] = v;
  }
}

__global__ void freq_bwd(const float* __restrict__ gy, const float* __restrict__ y,
                         long long N, int D, int deg, float* __restrict__ gx) {
  const int C = D + D * deg * 2;
  for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < N * D;
       k += (long long)gridDim.x * blockDim.x) {
    const long long n = k / D;
    const int d = (int)(k - n * D);
    const float* g = gy + n * C;
    const float* o = y + n * C;
    float r = g[d];
    g += D;
    o += D;
    for (int f = 0; f < deg; ++f) {
      r += scalbnf(1.0f, f) * (g[d] * o[d + D] - o[d] * g[D + d]);
      g += 2 * D;
      o += 2 * D;
    }
    gx[k] = r;
  }
}

}  // namespace

extern "C" int sailor_freq_encode_forward(const float* x, int N, int D, int deg, float* y,
                                          void* stream) {
  SAILOR_REQUIRE(N >= 0 && D > 0 && deg >= 0, "freq_encode_forward: bad sizes");
  if (N == 0) return SAILOR_OK;
  const long long C = (long long)D + 2LL * D * deg;
  if (D <= kFreqMaxD && (long long)kFreqPts * C < (1 << 24)) {
    hipLaunchKernelGGL(freq_fwd_tiled, dim3((unsigned)sailor::ceil_div(N, kFreqPts)),
                       dim3(kThreads), 0, sailor::as_stream(stream), x, (long long)N, D, deg, y);
    return sailor::check_launch("freq_encode_forward");
  }
  long long tot = (long long)N * C;
  int blocks = (int)std::min<long long>(65536, (tot + kThreads - 1) / kThreads);
  hipLaunchKernelGGL(freq_fwd, dim3(blocks), dim3(kThreads), 0, sailor::as_stream(stream),
                     x, (long long)N, D, deg, y);
  return sailor::check_launch("freq_encode_forward");
}

extern "C" int sailor_freq_encode_backward(const float* grad_y, const float* y, int N,
                                           int D, int deg, float* grad_x, void* stream) {
  SAILOR_REQUIRE(N >= 0 && D > 0 && deg >= 0, "freq_encode_backward: bad sizes");
  if (N == 0) return SAILOR_OK;
  long long tot = (long long)N * D;
  int blocks = (int)std::min<long long>(65536, (tot + kThreads - 1) / kThreads);
  hipLaunchKernelGGL(freq_bwd, dim3(blocks), dim3(kThreads), 0, sailor::as_stream(stream),
                     grad_y, y, (long long)N, D, deg, grad_x);
  return sailor::check_launch("freq_encode_backward");
}
