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
  long long tot = (long long)N * (D + 2 * D * deg);
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
