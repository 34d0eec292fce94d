// Fused per-ray trace: camera ray -> coarse-grid DDA -> length-weighted
// sampling, without materialising the reference's [B, R, 1024] hit lists.
//
// The reference chain is rays_calculating (sampling_kernel.cu:411-486) ->
// voxel_traversal (intersection_kernel.cu:827-895) ->
// ray_voxels_points_sampling_coarse (sampling_kernel.cu:1070-1177).  The
// sampler needs the total weighted valid length of a ray before it can place
// the first sample, so instead of storing every hit in HBM the walk is done
// in passes: pass 1 accumulates the valid length, pass 2 emits samples
// (stopping once all S slots are used), and a third pass places the leftover
// midpoint samples.  Pass 1 keeps up to kHitCache hit node ids per ray in
// LDS, and rays within that replay them for passes 2 and 3; longer rays walk
// again.  Every expression is the one of the
// separate kernels, so the samples are bit-identical to the unfused path.
// Compiled with -ffp-contract=off.
#include "common.h"
#include "traversal_common.h"

namespace {

constexpr int kThreads = 128;

struct NullSink {
  __device__ __forceinline__ void put(int, int, float, float) {}
  __device__ __forceinline__ bool stop() const { return false; }
};

struct SampleSink {
  sailor::SampleState ss;
  const float* nz;
  int* oi;
  float* ot;
  const float4* np;
  int has;
  __device__ __forceinline__ void put(int, int node, float t0, float t1) {
    if (has < ss.S) has = ss.interval(node, t0, t1, np[node].w, nz, oi, ot, has);
  }
  __device__ __forceinline__ bool stop() const { return has >= ss.S; }
};

// Hits of walk 1 kept in LDS (up to kHitCache node ids per ray, one column
// per lane): rays with at most that many hits replay them to the sample and
// leftover sinks instead of walking the DDA again.  A hit's interval is the
// slab test of the ray against that node, recomputed from the same inputs by
// the same code as in the walk, so only the node id is kept.  A sink receives
// the same (pos, node, t0, t1) sequence either way; stop() only cuts off puts
// that would change nothing, so the samples are identical.
constexpr int kHitCache = 44;

struct CacheSink {
  int* node;
  __device__ __forceinline__ void put(int pos, int nd, float, float) {
    if (pos < kHitCache) node[pos * kThreads] = nd;
  }
  __device__ __forceinline__ bool stop() const { return false; }
};

template <class Sink>
__device__ __forceinline__ void replay(const sailor::RayCtx& ctx, const CacheSink& c, int count,
                                       Sink& sink) {
  for (int k = 0; k < count && !sink.stop(); ++k) {
    const int nd = c.node[k * kThreads];
    const float4 q = ctx.np[nd];
    const float2 iv = sailor::slab_test(ctx.o, ctx.inv, make_float3(q.x, q.y, q.z), q.w);
    sink.put(k, nd, iv.x, iv.y);
  }
}

struct LeftoverSink {
  const float* nz;
  int* oi;
  float* ot;
  int has, left, S;
  __device__ __forceinline__ void put(int pos, int node, float t0, float t1) {
    if (pos < left && has < S) has = sailor::SampleState::leftover(node, t0, t1, nz, oi, ot, has);
  }
  __device__ __forceinline__ bool stop() const { return has >= S; }
};

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

// grid = (ceil(H*W/128), B); one lane per target pixel.
__global__ void __launch_bounds__(kThreads)
trace_kernel(const float* __restrict__ K, const float* __restrict__ Rm,
             const float* __restrict__ Tv, int H, int W,
             const int* __restrict__ vc, const float4* __restrict__ node_pos,
             const int4* __restrict__ node_info, int n_nodes,
             const float* __restrict__ vol_bbox, const float* __restrict__ voxel_size,
             int dim_fine, int dim_coarse, int rate, int start_scale, int S,
             const float* __restrict__ noise, float* __restrict__ rays_out,
             int* __restrict__ sidx, float* __restrict__ st, float* __restrict__ vl_out) {
  __shared__ float s_m[18];
  __shared__ int s_node[kHitCache * kThreads];
  const int b = blockIdx.y;
  if (threadIdx.x == 0) {
    inv3(K + (size_t)b * 9, s_m);
    inv3(Rm + (size_t)b * 9, s_m + 9);
  }
  __syncthreads();
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  // lanes past the image carry no ray but stay for the wave's tail fill
  const bool live = p < H * W;
  const long long ray = (long long)b * H * W + p;
  int has = S;   // samples placed; rows past the image have no tail
  if (live) {
    // ---- camera ray (rays_calculating_kernel, :443-484)
    const int h = p / W, w = p - h * W;
    const float* ki = s_m;
    const float* ri = s_m + 9;
    const float* t = Tv + (size_t)b * 3;
    float o[3], cd[3], wd[3], d[3];
  #pragma unroll
    for (int k = 0; k < 3; ++k)
      o[k] = -ri[3 * k + 0] * t[0] - ri[3 * k + 1] * t[1] - ri[3 * k + 2] * t[2];
  #pragma unroll
    for (int k = 0; k < 3; ++k)
      cd[k] = (float)w * ki[3 * k + 0] + (float)h * ki[3 * k + 1] + 1.0f * ki[3 * k + 2];
  #pragma unroll
    for (int k = 0; k < 3; ++k)
      wd[k] = cd[0] * ri[3 * k + 0] + cd[1] * ri[3 * k + 1] + cd[2] * ri[3 * k + 2];
    const float nrm = sqrtf(wd[0] * wd[0] + wd[1] * wd[1] + wd[2] * wd[2]);
    const double den = (double)nrm + 1e-10;
  #pragma unroll
    for (int k = 0; k < 3; ++k) d[k] = (float)((double)wd[k] / den);
    if (rays_out) {
      float* ro = rays_out + ray * 6;
      ro[0] = o[0]; ro[1] = o[1]; ro[2] = o[2];
      ro[3] = d[0]; ro[4] = d[1]; ro[5] = d[2];
    }
    // ---- walk 1: weighted valid length
    sailor::RayCtx ctx;
    ctx.init(vc, node_pos, node_info, n_nodes, vol_bbox, voxel_size, b, dim_fine, dim_coarse,
             rate, o, d);
    CacheSink cs{s_node + threadIdx.x};
    float vl = 0.0f;
    const int count = sailor::traverse(ctx, -1, false, -1, cs, &vl);
    const bool cached = count <= kHitCache;
    if (vl_out) vl_out[ray] = vl;
    int* oi = sidx + ray * S;
    float* ot = st + ray * S;
    const float* nz = noise + ray * S;
    has = 0;
    if (count > 0) {
      // ---- walk 2: samples (sampling_kernel.cu:1125-1154)
      SampleSink ss;
      ss.ss.init(vl, voxel_size[b] * (float)start_scale, S);
      ss.nz = nz; ss.oi = oi; ss.ot = ot; ss.np = node_pos; ss.has = 0;
      float dummy;
      if (cached)
        replay(ctx, cs, count, ss);
      else
        sailor::traverse(ctx, -1, false, -1, ss, &dummy);
      has = ss.has;
      // ---- walk 3: leftover midpoints (:1156-1174)
      if (has < S) {
        LeftoverSink ls{nz, oi, ot, has, S - has, S};
        if (cached)
          replay(ctx, cs, count, ls);
        else
          sailor::traverse(ctx, -1, false, -1, ls, &dummy);
        has = ls.has;
      }
    }
  }
  // ---- the rest of every row is -1, written row by row by the whole wave
  // (most rays have no sample: one lane per row would make each store
  // instruction touch 64 rows)
  const int lane = threadIdx.x & 63;
  const long long ray0 = ray - lane;
  for (int k = 0; k < 64; ++k) {
    const int hk = __builtin_amdgcn_readlane(has, k);
    if (hk >= S) continue;
    const long long rk = ray0 + k;
    for (int j = hk + lane; j < S; j += 64) {
      sidx[rk * S + j] = -1;
      st[rk * S + j] = -1.0f;
    }
  }
}

}  // namespace

extern "C" int sailor_trace_sample(const float* K, const float* R, const float* T, int B,
                                   int H, int W, const int32_t* vol_coarse,
                                   const float* node_pos, const int32_t* node_info,
                                   int n_nodes, const float* vol_bbox,
                                   const float* voxel_size, int dim_fine, int dim_coarse,
                                   int start_scale, int S, const float* noise,
                                   float* rays_out, int32_t* sampled_idx, float* sampled_t,
                                   float* valid_len, void* stream) {
  SAILOR_REQUIRE(B > 0 && H > 0 && W > 0 && S > 0, "trace_sample: bad sizes");
  SAILOR_REQUIRE(dim_coarse > 0 && dim_fine % dim_coarse == 0, "trace_sample: bad dims");
  const int rate = dim_fine / dim_coarse;
  hipLaunchKernelGGL(trace_kernel, dim3(sailor::ceil_div((long long)H * W, kThreads), B),
                     dim3(kThreads), 0, sailor::as_stream(stream), K, R, T, H, W, vol_coarse,
                     reinterpret_cast<const float4*>(node_pos),
                     reinterpret_cast<const int4*>(node_info), n_nodes, vol_bbox, voxel_size,
                     dim_fine, dim_coarse, rate, start_scale, S, noise, rays_out, sampled_idx,
                     sampled_t, valid_len);
  return sailor::check_launch("trace_sample");
}
