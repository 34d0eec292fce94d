// Per-channel input quantisation of the precision form.
//
// The 16-bit forms feed the per-view MLP with feature maps stored in bf16 /
// fp16.  Each channel c of a channels-last map x [rows, C] is stored as
// q = x / s_c with s_c = 2^e_c the smallest power of two >= max |x[:, c]|
// (s_c = 1 for an all-zero channel), so every stored value lies in [-1, 1]
// whatever the channel's range (fp16 overflows above 65504 and loses
// precision below 6e-5; bf16 keeps 8 significant bits at any magnitude).
// The scale is folded into the first layer: the packed weights of feature
// input k are multiplied by s_k.  Both scalings are by powers of two, so
// they are exact and for in-range inputs the product q * (w s) equals
// x * w in every bit the 16-bit form computes.
//
// Kernels (no memsets, so a frame that uses them can be captured into a
// hipGraph): clear the per-channel max keys, reduce max |x| per channel
// (non-negative floats order like their bit patterns, so unsigned atomicMax
// on the bits is a float max), turn the maxima into scales, convert the map,
// and rescale the packed first layer from an fp32 copy of its fragments.
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

#include "common.h"

namespace {

constexpr int kThreads = 256;
typedef float f32x4 __attribute__((ext_vector_type(4)));

__global__ void clear_keys(unsigned* keys, int C) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < C) keys[c] = 0u;
}

// grid-stride over the rows; thread t handles channel t % C of rows
// t / C, t / C + kThreads / C, ... (C divides kThreads or C >= kThreads)
__global__ void __launch_bounds__(kThreads)
channel_absmax(const float* __restrict__ x, long long rows, int C, unsigned* __restrict__ keys) {
  __shared__ float s_max[kThreads];
  const int per_block = C >= kThreads ? 1 : kThreads / C;   // rows per block step
  const int c0 = threadIdx.x % C;
  const int r0 = threadIdx.x / C;
  for (int cb = 0; cb < C; cb += kThreads) {                // channel blocks (C > kThreads)
    const int c = C >= kThreads ? cb + threadIdx.x : c0;
    float m = 0.0f;
    if (c < C) {
      const long long step = (long long)gridDim.x * per_block;
      for (long long r = (long long)blockIdx.x * per_block + (C >= kThreads ? 0 : r0); r < rows;
           r += step) {
        const float v = fabsf(x[r * C + c]);
        m = v > m ? v : m;                                  // NaN never wins
      }
    }
    s_max[threadIdx.x] = m;
    __syncthreads();
    if (C < kThreads) {
      if (threadIdx.x < C) {
        float mm = 0.0f;
        for (int j = threadIdx.x; j < kThreads - kThreads % C; j += C)
          mm = s_max[j] > mm ? s_max[j] : mm;
        atomicMax(keys + threadIdx.x, __float_as_uint(mm));
      }
    } else if (c < C) {
      atomicMax(keys + c, __float_as_uint(m));
    }
    __syncthreads();
  }
}

// The same reduction with 16-byte loads for C % 4 == 0 and C <= kThreads:
// thread t takes channels 4 (t % G) .. 4 (t % G) + 3, G = C / 4, of rows
// t / G, + kThreads / G, ...; four rows are loaded before any is reduced.
// (A maximum does not depend on the order of its operands, so the keys are
// the scalar kernel's.)
__global__ void __launch_bounds__(kThreads)
channel_absmax4(const float* __restrict__ x, long long rows, int C, unsigned* __restrict__ keys) {
  __shared__ float s_max[kThreads * 4];
  const int G = C / 4;
  const int g = threadIdx.x % G;
  const long long per = kThreads / G;                       // rows per block step
  const long long step = (long long)gridDim.x * per;
  const f32x4* x4 = reinterpret_cast<const f32x4*>(x);
  f32x4 m = {0.f, 0.f, 0.f, 0.f};
  auto fold = [&](const f32x4& v) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const float a = fabsf(v[k]);
      m[k] = a > m[k] ? a : m[k];                           // NaN never wins
    }
  };
  long long r = (long long)blockIdx.x * per + threadIdx.x / G;
  for (; r + 3 * step < rows; r += 4 * step) {
    const f32x4 v0 = x4[r * G + g], v1 = x4[(r + step) * G + g];
    const f32x4 v2 = x4[(r + 2 * step) * G + g], v3 = x4[(r + 3 * step) * G + g];
    fold(v0);
    fold(v1);
    fold(v2);
    fold(v3);
  }
  for (; r < rows; r += step) fold(x4[r * G + g]);
#pragma unroll
  for (int k = 0; k < 4; ++k) s_max[threadIdx.x * 4 + k] = m[k];
  __syncthreads();
  if (threadIdx.x < C) {
    const int gg = threadIdx.x / 4, k = threadIdx.x % 4;
    float mm = 0.0f;
    for (int t = gg; t < kThreads; t += G) {
      const float v = s_max[t * 4 + k];
      mm = v > mm ? v : mm;
    }
    atomicMax(keys + threadIdx.x, __float_as_uint(mm));
  }
}

__global__ void keys_to_scales(const unsigned* __restrict__ keys, int C,
                               float* __restrict__ scales) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= C) return;
  const float m = __uint_as_float(keys[c]);
  float s = 1.0f;
  if (m > 0.0f && m <= 3.0e38f) {
    int e = 0;
    frexpf(m, &e);                   // m = f 2^e, f in [0.5, 1)
    s = ldexpf(1.0f, e);             // 2^e >= m
  }
  scales[c] = s;
}

template <typename E>
__device__ __forceinline__ E to_e(float v);
template <>
__device__ __forceinline__ __hip_bfloat16 to_e<__hip_bfloat16>(float v) {
  return __float2bfloat16(v);
}
template <>
__device__ __forceinline__ __half to_e<__half>(float v) {
  return __float2half(v);
}

template <typename E>
__global__ void __launch_bounds__(kThreads)
scale_convert(const float* __restrict__ x, long long n, int C, const float* __restrict__ scales,
              E* __restrict__ out) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    out[i] = to_e<E>(x[i] / scales[i % C]);
}

template <typename E>
__global__ void __launch_bounds__(kThreads)
scale_layer(const float* __restrict__ w, const short* __restrict__ kidx, long long n, int F,
            const float* __restrict__ scales, E* __restrict__ out) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const int k = kidx[i];
    out[i] = to_e<E>(k < F ? w[i] * scales[k] : w[i]);
  }
}

int grid_for(long long n) {
  long long g = (n + kThreads - 1) / kThreads;
  return (int)std::max<long long>(1, std::min<long long>(g, 4096));
}

}  // namespace

extern "C" int sailor_quantize_channels(const float* x, long long rows, int C, int precision,
                                        void* out, float* scales, void* workspace,
                                        size_t workspace_bytes, void* stream) {
  SAILOR_REQUIRE(rows >= 0 && C >= 1, "quantize_channels: bad sizes (%lld rows, %d channels)",
                 rows, C);
  SAILOR_REQUIRE(precision == 1 || precision == 2,
                 "quantize_channels: precision must be 1 (bf16) or 2 (fp16), got %d", precision);
  SAILOR_REQUIRE(C < kThreads ? kThreads % C == 0 : C % kThreads == 0,
                 "quantize_channels: C = %d must divide %d or be a multiple of it", C, kThreads);
  SAILOR_REQUIRE(workspace_bytes >= (size_t)C * 4, "quantize_channels: workspace too small");
  hipStream_t st = sailor::as_stream(stream);
  unsigned* keys = reinterpret_cast<unsigned*>(workspace);
  hipLaunchKernelGGL(clear_keys, dim3(sailor::ceil_div(C, kThreads)), dim3(kThreads), 0, st,
                     keys, C);
  if (rows > 0 && C % 4 == 0 && C <= kThreads) {
    const long long per = kThreads / (C / 4);
    // about 32 rows per thread: each block ends in C atomics on the same C
    // keys, so fewer, longer blocks pay (C2 maps, 67 MB: 24 us; 16 rows per
    // thread 33 us, 8 rows 46 us; the scalar kernel 86 us)
    const long long blocks = std::max<long long>(1, (rows + 32 * per - 1) / (32 * per));
    hipLaunchKernelGGL(channel_absmax4, dim3((unsigned)std::min<long long>(blocks, 8192)),
                       dim3(kThreads), 0, st, x, rows, C, keys);
  } else if (rows > 0) {
    const int per_block = C >= kThreads ? 1 : kThreads / C;
    hipLaunchKernelGGL(channel_absmax, dim3(grid_for((rows + per_block - 1) / per_block * 64)),
                       dim3(kThreads), 0, st, x, rows, C, keys);
  }
  hipLaunchKernelGGL(keys_to_scales, dim3(sailor::ceil_div(C, kThreads)), dim3(kThreads), 0, st,
                     keys, C, scales);
  const long long n = rows * C;
  if (n > 0) {
    if (precision == 1)
      hipLaunchKernelGGL(scale_convert<__hip_bfloat16>, dim3(grid_for(n)), dim3(kThreads), 0,
                         st, x, n, C, scales, reinterpret_cast<__hip_bfloat16*>(out));
    else
      hipLaunchKernelGGL(scale_convert<__half>, dim3(grid_for(n)), dim3(kThreads), 0, st, x, n,
                         C, scales, reinterpret_cast<__half*>(out));
  }
  return sailor::check_launch("quantize_channels");
}

extern "C" int sailor_scale_packed_layer(const float* w, const int16_t* kidx, long long n,
                                         int F, const float* scales, int precision, void* out,
                                         void* stream) {
  SAILOR_REQUIRE(n >= 0 && F >= 1, "scale_packed_layer: bad sizes");
  SAILOR_REQUIRE(precision == 1 || precision == 2,
                 "scale_packed_layer: precision must be 1 (bf16) or 2 (fp16), got %d",
                 precision);
  if (n == 0) return SAILOR_OK;
  hipStream_t st = sailor::as_stream(stream);
  const short* k = reinterpret_cast<const short*>(kidx);
  if (precision == 1)
    hipLaunchKernelGGL(scale_layer<__hip_bfloat16>, dim3(grid_for(n)), dim3(kThreads), 0, st, w,
                       k, n, F, scales, reinterpret_cast<__hip_bfloat16*>(out));
  else
    hipLaunchKernelGGL(scale_layer<__half>, dim3(grid_for(n)), dim3(kThreads), 0, st, w, k, n,
                       F, scales, reinterpret_cast<__half*>(out));
  return sailor::check_launch("scale_packed_layer");
}
