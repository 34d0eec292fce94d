// Multi-resolution occupancy merge and the two-layer occupancy tree.
//
// Follows c_lib/VoxelEncoding/src/octree_kernel.cu: merge_non_empty_nodes
// (:242-303), gen_root_octree_nodes (:126-239) with the 26-neighbour isolation
// filter if_single_occ_in_neighbors (:101-122), node geometry from
// c_lib/VoxelEncoding/include/octree.h:110-151.
//
// The reference allocates node slots with atomicAdd, so its node order (and
// therefore every node index it returns) changes from run to run.  Here node
// slots are assigned by prefix sums over the coarse grid, giving one canonical
// order: coarse nodes in (batch, raster) order, and the fine children of each
// coarse node contiguous in raster order inside its block.  Node records are a
// structure of arrays (float4 centre+size, int4 batch/parent/first child/count)
// instead of the reference's ~1.1 KB array-of-structs OcTree record.
//
// Compiled with -ffp-contract=off (node centres feed bit-exact ray intervals).
#include "common.h"

#include <mutex>

namespace {

constexpr int kThreads = 256;
constexpr int kScanItems = 4;                         // per thread
constexpr int kScanBlock = kThreads * kScanItems;     // cells per scan block

// ---------------------------------------------------------------------------
// merge: coarse cell occupied iff any of its rate^3 fine cells is occupied.
// grid = (ceil(dc^3/256), B)
__global__ void __launch_bounds__(kThreads)
merge_kernel(int* __restrict__ coarse, int* __restrict__ count,
             const int* __restrict__ fine, int dc, int rate) {
  __shared__ int s_cnt[kThreads / 64];
  const int b = blockIdx.y;
  const int ncells = dc * dc * dc;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  bool occ = false;
  if (c < ncells) {
    const int df = dc * rate;
    const int cx = c / (dc * dc);
    const int cy = (c - cx * dc * dc) / dc;
    const int cz = c - cx * dc * dc - cy * dc;
    const int* f = fine + (size_t)b * df * df * df;
    for (int k = cx * rate; k < (cx + 1) * rate && !occ; ++k)
      for (int m = cy * rate; m < (cy + 1) * rate && !occ; ++m) {
        const int* row = f + ((size_t)k * df + m) * df + cz * rate;
        for (int n = 0; n < rate; ++n)
          if (row[n] > -1) { occ = true; break; }
      }
    coarse[(size_t)b * ncells + c] = occ ? c : -1;
  }
  unsigned long long msk = __ballot(occ);
  if ((threadIdx.x & 63) == 0) s_cnt[threadIdx.x >> 6] = __popcll(msk);
  __syncthreads();
  if (threadIdx.x == 0) {
    int s = 0;
    for (int i = 0; i < kThreads / 64; ++i) s += s_cnt[i];
    if (s) atomicAdd(count + b, s);
  }
}

// The fine rows of a coarse cell spread over lanes: rate^2 lanes per cell
// (a whole wave at rate 8), lane r reading fine row r (rate voxels along z).
// One lane per cell walking all rate^3 voxels leaves the grid at 32^3 lanes
// and each lane's 64 row reads in sequence.  Used when rate^2 divides 64.
__host__ __device__ constexpr bool rows_per_lane_ok(int rate) {
  return rate >= 2 && rate <= 8 && 64 % (rate * rate) == 0;
}

// fine row r of coarse cell (cx, cy, cz): any voxel occupied / how many
__device__ __forceinline__ int row_occupied_count(const int* __restrict__ fb, int df, int rate,
                                                  int cx, int cy, int cz, int r) {
  const int k = cx * rate + r / rate, m = cy * rate + r % rate;
  const int* row = fb + ((size_t)k * df + m) * df + cz * rate;
  int n = 0;
  if (rate % 4 == 0) {   // 16-byte aligned: rows start at multiples of rate ints
    for (int q = 0; q < rate; q += 4) {
      const int4 v = *reinterpret_cast<const int4*>(row + q);
      n += (v.x > -1) + (v.y > -1) + (v.z > -1) + (v.w > -1);
    }
  } else {
    for (int q = 0; q < rate; ++q) n += row[q] > -1;
  }
  return n;
}

// lanes of this lane's segment of rr lanes (rr divides 64, segments aligned)
__device__ __forceinline__ unsigned long long segment_mask(int lane, int rr) {
  return rr == 64 ? ~0ull : ((1ull << rr) - 1) << (lane - lane % rr);
}

// merge with rate^2 lanes per coarse cell.  grid = (ceil(dc^3 rate^2 / 256), B)
__global__ void __launch_bounds__(kThreads)
merge_rows_kernel(int* __restrict__ coarse, int* __restrict__ count,
                  const int* __restrict__ fine, int dc, int rate) {
  __shared__ int s_cnt[kThreads / 64];
  const int b = blockIdx.y;
  const int ncells = dc * dc * dc, rr = rate * rate;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int c = (int)(t / rr), r = (int)(t % rr);
  const int lane = threadIdx.x & 63;
  bool any = false;
  if (c < ncells) {
    const int df = dc * rate;
    const int cx = c / (dc * dc);
    const int cy = (c - cx * dc * dc) / dc;
    const int cz = c - cx * dc * dc - cy * dc;
    any = row_occupied_count(fine + (size_t)b * df * df * df, df, rate, cx, cy, cz, r) > 0;
  }
  const bool occ = (__ballot(any) & segment_mask(lane, rr)) != 0;
  const bool leader = r == 0 && c < ncells;
  if (leader) coarse[(size_t)b * ncells + c] = occ ? c : -1;
  const unsigned long long msk = __ballot(leader && occ);
  if (lane == 0) s_cnt[threadIdx.x >> 6] = __popcll(msk);
  __syncthreads();
  if (threadIdx.x == 0) {
    int s = 0;
    for (int i = 0; i < kThreads / 64; ++i) s += s_cnt[i];
    if (s) atomicAdd(count + b, s);
  }
}

// ---------------------------------------------------------------------------
// tree build, step 1: per coarse cell, is it allocated (occupied and not
// isolated) and how many occupied fine children does it have.
__global__ void __launch_bounds__(kThreads)
tree_count(const int* __restrict__ vc, const int* __restrict__ vf, int B,
           int dc, int rate, int* __restrict__ flag, int* __restrict__ nchild) {
  const int ncells = dc * dc * dc;
  const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= (long long)B * ncells) return;
  const int b = (int)(gid / ncells);
  const int c = (int)(gid - (long long)b * ncells);
  int f = 0, n = 0;
  if (vc[gid] > -1) {
    const int cx = c / (dc * dc);
    const int cy = (c - cx * dc * dc) / dc;
    const int cz = c - cx * dc * dc - cy * dc;
    const int* base = vc + (size_t)b * ncells;
    for (int k = cx - 1; k <= cx + 1 && !f; ++k)
      for (int m = cy - 1; m <= cy + 1 && !f; ++m)
        for (int q = cz - 1; q <= cz + 1; ++q) {
          if (k < 0 || k >= dc || m < 0 || m >= dc || q < 0 || q >= dc) continue;
          if (k == cx && m == cy && q == cz) continue;
          if (base[(k * dc + m) * dc + q] > -1) { f = 1; break; }
        }
    if (f) {
      const int df = dc * rate;
      const int* fb = vf + (size_t)b * df * df * df;
      for (int k = cx * rate; k < (cx + 1) * rate; ++k)
        for (int m = cy * rate; m < (cy + 1) * rate; ++m) {
          const int* row = fb + ((size_t)k * df + m) * df + cz * rate;
          for (int q = 0; q < rate; ++q) n += row[q] > -1;
        }
    }
  }
  flag[gid] = f;
  nchild[gid] = n;
}

// tree_count with rate^2 lanes per coarse cell: the lanes of a cell share
// its 26 neighbour tests and count one fine row each; the cell's first lane
// writes.  grid = ceil(B dc^3 rate^2 / 256).
__global__ void __launch_bounds__(kThreads)
tree_count_rows(const int* __restrict__ vc, const int* __restrict__ vf, int B,
                int dc, int rate, int* __restrict__ flag, int* __restrict__ nchild) {
  const int ncells = dc * dc * dc, rr = rate * rate;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long gid = t / rr;
  const int r = (int)(t % rr);
  const int lane = threadIdx.x & 63;
  const bool in = gid < (long long)B * ncells;
  int b = 0, c = 0, cx = 0, cy = 0, cz = 0;
  bool alloc = false, near = false;
  if (in) {
    b = (int)(gid / ncells);
    c = (int)(gid - (long long)b * ncells);
    alloc = vc[gid] > -1;
    if (alloc) {
      cx = c / (dc * dc);
      cy = (c - cx * dc * dc) / dc;
      cz = c - cx * dc * dc - cy * dc;
      const int* base = vc + (size_t)b * ncells;
      for (int j = r; j < 27; j += rr) {   // neighbour (j / 9, j / 3 % 3, j % 3) - 1
        const int k = cx + j / 9 - 1, m = cy + j / 3 % 3 - 1, q = cz + j % 3 - 1;
        if (j == 13 || k < 0 || k >= dc || m < 0 || m >= dc || q < 0 || q >= dc) continue;
        if (base[(k * dc + m) * dc + q] > -1) near = true;
      }
    }
  }
  const bool f = (__ballot(near) & segment_mask(lane, rr)) != 0;
  int n = 0;
  if (f) {
    const int df = dc * rate;
    n = row_occupied_count(vf + (size_t)b * df * df * df, df, rate, cx, cy, cz, r);
  }
  for (int off = 1; off < rr; off <<= 1) n += __shfl_xor(n, off, 64);   // segment sum
  if (in && r == 0) {
    flag[gid] = f ? 1 : 0;
    nchild[gid] = n;
  }
}

// Block-level exclusive scan of (a, b) pairs held as kScanItems per thread.
// Returns block totals in *ta, *tb.
__device__ void block_scan2(int (&a)[kScanItems], int (&b)[kScanItems], int* ta,
                            int* tb) {
  __shared__ int s_a[kThreads / 64], s_b[kThreads / 64];
  int sa = 0, sb = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    int x = a[i], y = b[i];
    a[i] = sa; b[i] = sb;
    sa += x; sb += y;
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int ia = sa, ib = sb;  // inclusive wave scan
  for (int o = 1; o < 64; o <<= 1) {
    int xa = __shfl_up(ia, o, 64), xb = __shfl_up(ib, o, 64);
    if (lane >= o) { ia += xa; ib += xb; }
  }
  if (lane == 63) { s_a[wid] = ia; s_b[wid] = ib; }
  __syncthreads();
  int wa = 0, wb = 0, tot_a = 0, tot_b = 0;
  for (int w = 0; w < kThreads / 64; ++w) {
    if (w < wid) { wa += s_a[w]; wb += s_b[w]; }
    tot_a += s_a[w]; tot_b += s_b[w];
  }
  const int ea = wa + ia - sa, eb = wb + ib - sb;  // exclusive thread offset
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) { a[i] += ea; b[i] += eb; }
  *ta = tot_a; *tb = tot_b;
  __syncthreads();
}

// step 2: per scan block totals
__global__ void __launch_bounds__(kThreads)
tree_block_sums(const int* __restrict__ flag, const int* __restrict__ nchild,
                long long n, int* __restrict__ bsum_a, int* __restrict__ bsum_b) {
  int a[kScanItems], b[kScanItems];
  const long long base = (long long)blockIdx.x * kScanBlock + threadIdx.x * kScanItems;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    long long j = base + i;
    a[i] = j < n ? flag[j] : 0;
    b[i] = j < n ? nchild[j] : 0;
  }
  int ta, tb;
  block_scan2(a, b, &ta, &tb);
  if (threadIdx.x == 0) { bsum_a[blockIdx.x] = ta; bsum_b[blockIdx.x] = tb; }
}

// step 3 (single workgroup): exclusive scan of the block totals, node counts,
// capacity check.  meta = [ok, coarse_start, n_coarse, n_fine, tail_end]
__global__ void __launch_bounds__(kThreads)
tree_scan_sums(int* bsum_a, int* bsum_b, int nblk, const int* fine_counts,
               const int* coarse_counts, int B, int coarse_start_arg,
               int capacity, int* meta, int* queue_counts) {
  int carry_a = 0, carry_b = 0;
  for (int base = 0; base < nblk; base += kScanBlock) {
    int a[kScanItems], b[kScanItems];
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
      int j = base + threadIdx.x * kScanItems + i;
      a[i] = j < nblk ? bsum_a[j] : 0;
      b[i] = j < nblk ? bsum_b[j] : 0;
    }
    int ta, tb;
    block_scan2(a, b, &ta, &tb);
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
      int j = base + threadIdx.x * kScanItems + i;
      if (j < nblk) { bsum_a[j] = a[i] + carry_a; bsum_b[j] = b[i] + carry_b; }
    }
    carry_a += ta;
    carry_b += tb;
  }
  if (threadIdx.x == 0) {
    int cs = coarse_start_arg;
    long long tot_fine = 0, tot_coarse = 0;
    for (int i = 0; i < B; ++i) { tot_fine += fine_counts[i]; tot_coarse += coarse_counts[i]; }
    if (cs < 0) cs = (int)tot_fine;
    long long tail = (long long)cs + tot_coarse;
    if (tail > capacity) tail = capacity;
    int ok = (carry_b <= cs) && ((long long)cs + carry_a <= capacity);
    meta[0] = ok;
    meta[1] = cs;
    meta[2] = carry_a;
    meta[3] = carry_b;
    meta[4] = (int)tail;
    queue_counts[0] = ok ? carry_a : -1;
    queue_counts[1] = ok ? carry_b : -1;
  }
}

__device__ __forceinline__ float coarse_offset(float vs_top, int size) {
  // octree.h:133-151: sum_{i < log2(size)} 2^i * (voxel_size/size) / 2
  float basic = vs_top / (float)size;
  int lg = sailor::floor_log2(size);
  float off = 0.0f;
  for (int i = 0; i < lg; ++i) off += (float)(1 << i) * basic / 2.0f;
  return off;
}

// step 4: coarse node records, coarse volume rewrite, per-cell child base.
__global__ void __launch_bounds__(kThreads)
tree_write_coarse(int* __restrict__ vc, const int* __restrict__ flag,
                  const int* __restrict__ nchild, const int* __restrict__ bsum_a,
                  const int* __restrict__ bsum_b, long long n, int dc, int df,
                  const float* __restrict__ voxel_size,
                  const float* __restrict__ vol_bbox, const int* __restrict__ meta,
                  float4* __restrict__ node_pos, int4* __restrict__ node_info,
                  int* __restrict__ cell_gidx, int* __restrict__ cell_child) {
  if (!meta[0]) return;
  const int cs = meta[1];
  int a[kScanItems], b[kScanItems], fl[kScanItems], nc[kScanItems];
  const long long base = (long long)blockIdx.x * kScanBlock + threadIdx.x * kScanItems;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    long long j = base + i;
    fl[i] = a[i] = j < n ? flag[j] : 0;
    nc[i] = b[i] = j < n ? nchild[j] : 0;
  }
  int ta, tb;
  block_scan2(a, b, &ta, &tb);
  const int oa = bsum_a[blockIdx.x], ob = bsum_b[blockIdx.x];
  const int ncells = dc * dc * dc;
  const int ratio = df / dc;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    long long j = base + i;
    if (j >= n) break;
    if (!fl[i]) {
      if (vc[j] > -1) vc[j] = -1;  // isolated coarse cell: not a node
      continue;
    }
    const int bb = (int)(j / ncells);
    const int c = (int)(j - (long long)bb * ncells);
    const int cx = c / (dc * dc);
    const int cy = (c - cx * dc * dc) / dc;
    const int cz = c - cx * dc * dc - cy * dc;
    const int g = cs + oa + a[i];
    const int child0 = ob + b[i];
    const float vs_top = voxel_size[bb] * (float)ratio;
    const float off = coarse_offset(vs_top, ratio);
    const float* box = vol_bbox + bb * 6;
    float sx = box[0] + off, sy = box[2] + off, sz = box[4] + off;
    node_pos[g] = make_float4(sx + vs_top * (float)cx, sy + vs_top * (float)cy,
                              sz + vs_top * (float)cz, vs_top);
    node_info[g] = make_int4(bb, -1, nc[i] > 0 ? child0 : -1, nc[i]);
    vc[j] = g;
    cell_gidx[j] = g;
    cell_child[j] = child0;
  }
}

// step 5: one wave per coarse cell writes its fine children in raster order.
__global__ void __launch_bounds__(kThreads)
tree_write_fine(int* __restrict__ vf, const int* __restrict__ flag,
                const int* __restrict__ cell_gidx,
                const int* __restrict__ cell_child, long long n, int dc, int rate,
                const float* __restrict__ voxel_size,
                const float* __restrict__ vol_bbox, const int* __restrict__ meta,
                float4* __restrict__ node_pos, int4* __restrict__ node_info) {
  if (!meta[0]) return;
  const long long j = (long long)blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (j >= n || !flag[j]) return;
  const int ncells = dc * dc * dc;
  const int bb = (int)(j / ncells);
  const int c = (int)(j - (long long)bb * ncells);
  const int cx = c / (dc * dc);
  const int cy = (c - cx * dc * dc) / dc;
  const int cz = c - cx * dc * dc - cy * dc;
  const int df = dc * rate;
  const int parent = cell_gidx[j];
  int next = cell_child[j];
  const float vs = voxel_size[bb] * (float)(df / df);
  const float* box = vol_bbox + bb * 6;
  const float sx = box[0] + 0.0f, sy = box[2] + 0.0f, sz = box[4] + 0.0f;
  int* fb = vf + (size_t)bb * df * df * df;
  const int nblock = rate * rate * rate;
  for (int i0 = 0; i0 < nblock; i0 += 64) {
    const int i = i0 + lane;
    bool occ = false;
    int k = 0, m = 0, q = 0;
    size_t fidx = 0;
    if (i < nblock) {
      k = cx * rate + i / (rate * rate);
      m = cy * rate + (i / rate) % rate;
      q = cz * rate + i % rate;
      fidx = ((size_t)k * df + m) * df + q;
      occ = fb[fidx] > -1;
    }
    unsigned long long msk = __ballot(occ);
    if (occ) {
      int g = next + __popcll(msk & ((1ull << lane) - 1ull));
      node_pos[g] = make_float4(sx + vs * (float)k, sy + vs * (float)m,
                                sz + vs * (float)q, vs);
      node_info[g] = make_int4(bb, parent, -1, 0);
      fb[fidx] = g;
    }
    next += __popcll(msk);
  }
}

// step 6: mark the slots that hold no node.
__global__ void tree_fill_invalid(const int* __restrict__ meta,
                                  float4* __restrict__ node_pos,
                                  int4* __restrict__ node_info) {
  if (!meta[0]) return;
  const int cs = meta[1], nco = meta[2], nfi = meta[3], tail = meta[4];
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < tail;
       i += (long long)gridDim.x * blockDim.x) {
    bool hole = (i >= nfi && i < cs) || (i >= (long long)cs + nco);
    if (hole) {
      node_pos[i] = make_float4(0.f, 0.f, 0.f, 0.f);
      node_info[i] = make_int4(-1, -1, -1, 0);
    }
  }
}

__global__ void nodes_tensor_kernel(const float4* __restrict__ pos,
                                    const int4* __restrict__ info, int n,
                                    float* __restrict__ out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float4 p = pos[i];
  int4 f = info[i];
  float* o = out + (size_t)i * 6;
  bool valid = f.x >= 0;
  o[0] = valid ? p.x : 0.f;
  o[1] = valid ? p.y : 0.f;
  o[2] = valid ? p.z : 0.f;
  o[3] = valid ? p.w : 0.f;
  o[4] = valid ? 1.f : 0.f;
  o[5] = valid ? (float)f.x : 0.f;
}

struct TreeWorkspace {
  int* flag;
  int* nchild;
  int* cell_gidx;
  int* cell_child;
  int* bsum_a;
  int* bsum_b;
  int* meta;
  size_t bytes;
};

TreeWorkspace layout(void* base, long long ncells) {
  TreeWorkspace w;
  long long nblk = (ncells + kScanBlock - 1) / kScanBlock;
  int* p = reinterpret_cast<int*>(base);
  w.flag = p;        p += ncells;
  w.nchild = p;      p += ncells;
  w.cell_gidx = p;   p += ncells;
  w.cell_child = p;  p += ncells;
  w.bsum_a = p;      p += nblk;
  w.bsum_b = p;      p += nblk;
  w.meta = p;        p += 8;
  w.bytes = (size_t)(4 * ncells + 2 * nblk + 8) * sizeof(int);
  return w;
}

}  // namespace

extern "C" int sailor_build_multi_res_volumes(int32_t* const* vols,
                                              int32_t* const* counts, int levels,
                                              int B, int dim, int rate,
                                              void* stream) {
  SAILOR_REQUIRE(levels >= 1 && B > 0 && rate >= 2, "build_multi_res_volumes: bad args");
  hipStream_t st = sailor::as_stream(stream);
  int d = dim;
  for (int i = 0; i < levels - 1; ++i) {
    SAILOR_REQUIRE(d % rate == 0,
                   "build_multi_res_volumes: dim %d of level %d not divisible by rate %d",
                   d, i, rate);
    int dc = d / rate;
    if (rows_per_lane_ok(rate))
      hipLaunchKernelGGL(merge_rows_kernel,
                         dim3(sailor::ceil_div((long long)dc * dc * dc * rate * rate, kThreads), B),
                         dim3(kThreads), 0, st, vols[i + 1], counts[i + 1], vols[i], dc, rate);
    else
      hipLaunchKernelGGL(merge_kernel,
                         dim3(sailor::ceil_div((long long)dc * dc * dc, kThreads), B),
                         dim3(kThreads), 0, st, vols[i + 1], counts[i + 1], vols[i], dc, rate);
    int rc = sailor::check_launch("merge_non_empty_nodes");
    if (rc) return rc;
    d = dc;
  }
  return SAILOR_OK;
}

extern "C" size_t sailor_octree_workspace_bytes(int B, int dim_coarse) {
  long long ncells = (long long)B * dim_coarse * dim_coarse * dim_coarse;
  return layout(nullptr, ncells).bytes;
}

// Workspace of sailor_octree_build_noalloc: one copy per GPU, held for the
// life of the process.  In host code `scratch` is only a symbol; its device
// address (per device) comes from hipGetSymbolAddress.
constexpr size_t kNoAllocScratchBytes = size_t(8) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

// The build's six launches on an already laid-out workspace.  Every
// workspace word a launch reads was written by an earlier launch of the same
// call (flags and child counts for every cell, block sums, meta; cell
// indices of the allocated cells only, read only for those), so the
// workspace needs no clearing between calls.
int octree_build_launch(int32_t* vol_fine, int32_t* vol_coarse, const int32_t* fine_counts,
                        const int32_t* coarse_counts, int B, int dim_fine, int dim_coarse,
                        const float* voxel_size, const float* vol_bbox, int coarse_start,
                        int capacity, float* node_pos, int32_t* node_info,
                        int32_t* queue_counts, const TreeWorkspace& w, long long n,
                        hipStream_t st) {
  const int rate = dim_fine / dim_coarse;
  const int nblk = (int)((n + kScanBlock - 1) / kScanBlock);
  float4* pos = reinterpret_cast<float4*>(node_pos);
  int4* info = reinterpret_cast<int4*>(node_info);

  if (rows_per_lane_ok(rate))
    hipLaunchKernelGGL(tree_count_rows, dim3(sailor::ceil_div(n * rate * rate, kThreads)),
                       dim3(kThreads), 0, st, vol_coarse, vol_fine, B, dim_coarse, rate, w.flag,
                       w.nchild);
  else
    hipLaunchKernelGGL(tree_count, dim3(sailor::ceil_div(n, kThreads)), dim3(kThreads), 0, st,
                       vol_coarse, vol_fine, B, dim_coarse, rate, w.flag, w.nchild);
  hipLaunchKernelGGL(tree_block_sums, dim3(nblk), dim3(kThreads), 0, st, w.flag, w.nchild,
                     n, w.bsum_a, w.bsum_b);
  hipLaunchKernelGGL(tree_scan_sums, dim3(1), dim3(kThreads), 0, st, w.bsum_a, w.bsum_b,
                     nblk, fine_counts, coarse_counts, B, coarse_start, capacity,
                     w.meta, queue_counts);
  hipLaunchKernelGGL(tree_write_coarse, dim3(nblk), dim3(kThreads), 0, st, vol_coarse,
                     w.flag, w.nchild, w.bsum_a, w.bsum_b, n, dim_coarse, dim_fine,
                     voxel_size, vol_bbox, w.meta, pos, info, w.cell_gidx, w.cell_child);
  hipLaunchKernelGGL(tree_write_fine, dim3(sailor::ceil_div(n, kThreads / 64)),
                     dim3(kThreads), 0, st, vol_fine, w.flag, w.cell_gidx, w.cell_child,
                     n, dim_coarse, rate, voxel_size, vol_bbox, w.meta, pos, info);
  int fill_blocks = std::max(1, std::min(4096, sailor::ceil_div(capacity, kThreads)));
  hipLaunchKernelGGL(tree_fill_invalid, dim3(fill_blocks), dim3(kThreads), 0, st,
                     w.meta, pos, info);
  return sailor::check_launch("octree_build");
}

// Shared argument checks of the two build entry points; sets *n to the
// number of coarse cells over the batch.
int octree_build_args(int B, int dim_fine, int dim_coarse, long long* n) {
  SAILOR_REQUIRE(B > 0 && dim_coarse > 0 && dim_fine % dim_coarse == 0,
                 "octree_build: fine dim %d is not a multiple of coarse dim %d",
                 dim_fine, dim_coarse);
  SAILOR_REQUIRE(dim_fine / dim_coarse >= 2, "octree_build: rate must be >= 2");
  *n = (long long)B * dim_coarse * dim_coarse * dim_coarse;
  return SAILOR_OK;
}

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up on the first
// call made with that device current.
int scratch_address(void** out) {
  int dev = -1;
  SAILOR_HIP_TRY(hipGetDevice(&dev));
  SAILOR_REQUIRE(dev >= 0 && dev < kMaxDevices, "octree_build_noalloc: device %d out of range",
                 dev);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    SAILOR_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    SAILOR_REQUIRE(p != nullptr, "octree_build_noalloc: no device address for the scratch");
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return SAILOR_OK;
}

}  // namespace

extern "C" int sailor_octree_build(int32_t* vol_fine, int32_t* vol_coarse,
                                   const int32_t* fine_counts,
                                   const int32_t* coarse_counts, int B,
                                   int dim_fine, int dim_coarse,
                                   const float* voxel_size, const float* vol_bbox,
                                   int coarse_start, int capacity,
                                   float* node_pos, int32_t* node_info,
                                   int32_t* queue_counts, void* workspace,
                                   size_t workspace_bytes, void* stream) {
  long long n = 0;
  int rc = octree_build_args(B, dim_fine, dim_coarse, &n);
  if (rc) return rc;
  TreeWorkspace w = layout(workspace, n);
  SAILOR_REQUIRE(workspace_bytes >= w.bytes, "octree_build: workspace too small (%zu < %zu)",
                 workspace_bytes, w.bytes);
  return octree_build_launch(vol_fine, vol_coarse, fine_counts, coarse_counts, B, dim_fine,
                             dim_coarse, voxel_size, vol_bbox, coarse_start, capacity, node_pos,
                             node_info, queue_counts, w, n, sailor::as_stream(stream));
}

extern "C" int sailor_octree_build_noalloc(int32_t* vol_fine, int32_t* vol_coarse,
                                           const int32_t* fine_counts,
                                           const int32_t* coarse_counts, int B, int dim_fine,
                                           int dim_coarse, const float* voxel_size,
                                           const float* vol_bbox, int coarse_start,
                                           int capacity, float* node_pos, int32_t* node_info,
                                           int32_t* queue_counts, void* stream) {
  long long n = 0;
  int rc = octree_build_args(B, dim_fine, dim_coarse, &n);
  if (rc) return rc;
  const size_t need = layout(nullptr, n).bytes;
  SAILOR_REQUIRE(need <= kNoAllocScratchBytes,
                 "octree_build_noalloc: %lld coarse cells need %zu bytes of workspace, more "
                 "than the %zu-byte built-in scratch (use sailor_octree_build)",
                 n, need, kNoAllocScratchBytes);
  void* base = nullptr;
  rc = scratch_address(&base);
  if (rc) return rc;
  TreeWorkspace w = layout(base, n);
  return octree_build_launch(vol_fine, vol_coarse, fine_counts, coarse_counts, B, dim_fine,
                             dim_coarse, voxel_size, vol_bbox, coarse_start, capacity, node_pos,
                             node_info, queue_counts, w, n, sailor::as_stream(stream));
}

extern "C" int sailor_octree_nodes_tensor(const float* node_pos,
                                          const int32_t* node_info, int n,
                                          float* out, void* stream) {
  if (n <= 0) return SAILOR_OK;
  hipLaunchKernelGGL(nodes_tensor_kernel, dim3(sailor::ceil_div(n, kThreads)),
                     dim3(kThreads), 0, sailor::as_stream(stream),
                     reinterpret_cast<const float4*>(node_pos),
                     reinterpret_cast<const int4*>(node_info), n, out);
  return sailor::check_launch("nodes_tensor");
}
