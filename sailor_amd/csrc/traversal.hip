// Ray / voxel intersection (Amanatides-Woo DDA over the coarse grid), per-ray
// adaptive sampling inside the hit intervals, and the per-ray sort.
//
// Follows c_lib/VoxelEncoding/src/intersection_kernel.cu (ray_voxel_traversal
// :286-439, RayVoxelIntersectionNodeListNew :128-284, RayVoxelIntersection
// :36-71) and c_lib/VoxelEncoding/src/sampling_kernel.cu
// (ray_voxels_sampling_uniform_kernel :1070-1177, sort_samplings :1308-1329).
//
// Differences from the reference (all deliberate):
//  * the slab test uses the IEEE reciprocal 1.0f/d instead of __fdividef;
//  * the child weight and the start-level voxel scale use the tree's `rate`
//    instead of the literal 4 (identical for rate 4);
//  * hit lists are bounded by n_max_hits with an overflow flag instead of
//    writing past the end of the row;
//  * the sort is stable (ties keep slot order) so the permutation is unique.
// Compiled with -ffp-contract=off: intervals, valid lengths and sample counts
// are reproduced bit for bit by the CPU oracle.
#include "common.h"
#include "traversal_common.h"

namespace {

constexpr int kThreads = 256;

__global__ void fill3(int* __restrict__ idx, float* __restrict__ a,
                      float* __restrict__ b, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    idx[i] = -1;
    a[i] = -1.0f;
    b[i] = -1.0f;
  }
}

struct ListSink {
  int* idx;
  float* tmin;
  float* tmax;
  int cap;
  bool overflow;
  __device__ __forceinline__ void put(int pos, int node, float t0, float t1) {
    if (pos < cap) {
      idx[pos] = node;
      tmin[pos] = t0;
      tmax[pos] = t1;
    } else {
      overflow = true;
    }
  }
  __device__ __forceinline__ bool stop() const { return false; }
};

__global__ void __launch_bounds__(kThreads)
traversal_kernel(const int* __restrict__ vc, const float4* __restrict__ node_pos,
                 const int4* __restrict__ node_info, int n_nodes,
                 const float* __restrict__ vol_bbox,
                 const float* __restrict__ voxel_size, int B, int dim_fine,
                 int dim_coarse, int rate, const float* __restrict__ ray_ori,
                 const float* __restrict__ ray_dir, int R, int n_max,
                 int th_early_stop, int is_training, int opt_num,
                 int* __restrict__ hit_idx, float* __restrict__ tmin,
                 float* __restrict__ tmax, float* __restrict__ valid_len,
                 int* __restrict__ status) {
  const long long ray = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  int count = 0;
  bool overflow = false;
  if (ray < (long long)B * R) {
    const int b = (int)(ray / R);
    sailor::RayCtx ctx;
    ctx.init(vc, node_pos, node_info, n_nodes, vol_bbox, voxel_size, b, dim_fine,
             dim_coarse, rate, ray_ori + ray * 3, ray_dir + ray * 3);
    ListSink sink{hit_idx + ray * n_max, tmin + ray * n_max, tmax + ray * n_max,
                  n_max, false};
    float vl = 0.0f;
    count = sailor::traverse(ctx, th_early_stop, is_training != 0, opt_num, sink, &vl);
    valid_len[ray] = vl;
    overflow = sink.overflow;
  }
  // global max hit count (reference: atomicMax per step, :405) and overflow
  int m = count;
  for (int o = 32; o > 0; o >>= 1) m = max(m, __shfl_xor(m, o, 64));
  unsigned long long ov = __ballot(overflow);
  if ((threadIdx.x & 63) == 0) {
    if (m > 0) atomicMax(status, m);
    if (ov) atomicOr(status + 1, 1);
  }
}

__global__ void __launch_bounds__(kThreads)
sampling_kernel(const int* __restrict__ hit_idx, const float* __restrict__ tmin,
                const float* __restrict__ tmax, const float* __restrict__ noise,
                const float* __restrict__ valid_len,
                const float4* __restrict__ node_pos,
                const float* __restrict__ voxel_size, int B, int R, int n_max,
                int S, int start_scale, int* __restrict__ sidx,
                float* __restrict__ st) {
  const long long ray = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (ray >= (long long)B * R) return;
  const int b = (int)(ray / R);
  int* oi = sidx + ray * S;
  float* ot = st + ray * S;
  const int* hi = hit_idx + ray * n_max;
  const float* h0 = tmin + ray * n_max;
  const float* h1 = tmax + ray * n_max;
  const float* nz = noise + ray * S;
  int has = 0;
  if (hi[0] >= 0) {
    sailor::SampleState ss;
    ss.init(valid_len[ray], voxel_size[b] * (float)start_scale, S);
    for (int i = 0; i < n_max; ++i) {
      int v = hi[i];
      float mn = h0[i], mx = h1[i];
      if (v < 0 || mn < 0 || mx < 0) break;
      has = ss.interval(v, mn, mx, node_pos[v].w, nz, oi, ot, has);
      if (has >= S) break;
    }
    // leftover slots: one midpoint sample per hit, in hit order (:1156-1174)
    int left = S - has;
    for (int i = 0; i < left && i < n_max; ++i) {
      int v = hi[i];
      float mn = h0[i], mx = h1[i];
      if (v < 0 || mn < 0 || mx < 0) break;
      has = sailor::SampleState::leftover(v, mn, mx, nz, oi, ot, has);
    }
  }
  for (int j = has; j < S; ++j) {
    oi[j] = -1;
    ot[j] = -1.0f;
  }
}

// Stable ascending sort of each row; one wave per row, rank counting in LDS.
constexpr int kSortMaxS = 1024;
__global__ void __launch_bounds__(kThreads)
sort_rows(const float* __restrict__ t, int n_rows, int S, float* __restrict__ out,
          int* __restrict__ order) {
  __shared__ float s_v[kThreads / 64][kSortMaxS];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const long long row = (long long)blockIdx.x * (kThreads / 64) + wid;
  const bool active = row < n_rows;
  float* sv = s_v[wid];
  if (active) {
    const float* src = t + row * S;
    for (int i = lane; i < S; i += 64) sv[i] = src[i];
  }
  __syncthreads();
  if (!active) return;
  // A row already in non-decreasing order (every row without samples, all
  // -1) is its own stable sort: rank of element i = i.  (A NaN fails the
  // test and takes the rank loop, as before.)
  bool out_of_order = false;
  for (int i = lane; i + 1 < S; i += 64) out_of_order |= !(sv[i] <= sv[i + 1]);
  if (__ballot(out_of_order) == 0) {
    for (int i = lane; i < S; i += 64) {
      out[row * S + i] = sv[i];
      order[row * S + i] = i;
    }
    return;
  }
  for (int i = lane; i < S; i += 64) {
    float v = sv[i];
    int r = 0;
    for (int j = 0; j < S; ++j) {
      float u = sv[j];
      r += (u < v) || (u == v && j < i);
    }
    out[row * S + r] = v;
    order[row * S + r] = i;
  }
}

}  // namespace

extern "C" int sailor_octree_voxel_traversal(
    const int32_t* vol_coarse, const float* node_pos, const int32_t* node_info,
    int n_nodes, const float* vol_bbox, const float* voxel_size, int B, int dim_fine,
    int dim_coarse, int rate, const float* ray_ori, const float* ray_dir, int R,
    int n_max_hits, int th_early_stop, int is_training, int opt_num_intersected,
    int32_t* hit_idx, float* tmin, float* tmax, float* valid_len,
    int32_t* status, void* stream) {
  SAILOR_REQUIRE(B > 0 && R >= 0 && n_max_hits > 0, "voxel_traversal: bad sizes");
  SAILOR_REQUIRE(dim_coarse > 0 && rate >= 1, "voxel_traversal: bad tree dims");
  hipStream_t st = sailor::as_stream(stream);
  long long n = (long long)B * R * n_max_hits;
  if (R == 0) return SAILOR_OK;
  int fb = (int)std::min<long long>(8192, (n + kThreads - 1) / kThreads);
  hipLaunchKernelGGL(fill3, dim3(fb), dim3(kThreads), 0, st, hit_idx, tmin, tmax, n);
  hipLaunchKernelGGL(traversal_kernel, dim3(sailor::ceil_div((long long)B * R, kThreads)),
                     dim3(kThreads), 0, st, vol_coarse,
                     reinterpret_cast<const float4*>(node_pos),
                     reinterpret_cast<const int4*>(node_info), n_nodes, vol_bbox,
                     voxel_size, B,
                     dim_fine, dim_coarse, rate, ray_ori, ray_dir, R, n_max_hits,
                     th_early_stop, is_training, opt_num_intersected, hit_idx, tmin,
                     tmax, valid_len, status);
  return sailor::check_launch("voxel_traversal");
}

extern "C" int sailor_octree_sampling_coarse(
    const int32_t* hit_idx, const float* tmin, const float* tmax,
    const float* noise, const float* valid_len, const float* node_pos,
    const float* voxel_size, int B, int R, int n_max_hits, int S, int start_scale,
    int32_t* sampled_idx, float* sampled_t, void* stream) {
  SAILOR_REQUIRE(B > 0 && R >= 0 && S > 0 && n_max_hits > 0, "sampling: bad sizes");
  if (R == 0) return SAILOR_OK;
  hipLaunchKernelGGL(sampling_kernel, dim3(sailor::ceil_div((long long)B * R, kThreads)),
                     dim3(kThreads), 0, sailor::as_stream(stream), hit_idx, tmin, tmax,
                     noise, valid_len, reinterpret_cast<const float4*>(node_pos),
                     voxel_size, B, R, n_max_hits, S, start_scale, sampled_idx,
                     sampled_t);
  return sailor::check_launch("sampling_coarse");
}

extern "C" int sailor_sort_samplings(const float* t, int n_rows, int S, float* sorted,
                                     int32_t* order, void* stream) {
  SAILOR_REQUIRE(S > 0 && S <= kSortMaxS, "sort_samplings: S must be in [1, %d]", kSortMaxS);
  if (n_rows <= 0) return SAILOR_OK;
  hipLaunchKernelGGL(sort_rows, dim3(sailor::ceil_div(n_rows, kThreads / 64)),
                     dim3(kThreads), 0, sailor::as_stream(stream), t, n_rows, S, sorted,
                     order);
  return sailor::check_launch("sort_samplings");
}
