// Device-side building blocks shared by the ray/voxel traversal kernels: the
// slab test, the coarse-grid DDA with per-node child expansion, and the
// length-weighted sampler.  Every expression keeps the reference's operation
// order; callers compile with -ffp-contract=off.
#pragma once

#include "common.h"

namespace sailor {

constexpr float kNoHit = -1.0f;
constexpr float kDdaFar = 9999999.0f;  // DBL_MAX macro of intersection_kernel.cu:16

// RayVoxelIntersection (intersection_kernel.cu:36-71) with inv = 1/d
// precomputed (IEEE division; the reference uses __fdividef).
__device__ __forceinline__ float2 slab_test(float3 o, float3 inv, float3 c,
                                            float size) {
  float lo = 0.0f, hi = 100000.f;
  const float half = size / 2.0f;
#define SAILOR_SLAB_AXIS(OA, IA, CA)                                  \
  {                                                                   \
    float t0 = (CA - half - OA) * IA;                                 \
    float t1 = (CA + half - OA) * IA;                                 \
    if (t0 > t1) { float tmp = t0; t0 = t1; t1 = tmp; }               \
    if (t1 < lo) return make_float2(kNoHit, kNoHit);                  \
    if (t0 > hi) return make_float2(kNoHit, kNoHit);                  \
    lo = (t0 > lo) ? t0 : lo;                                         \
    hi = (t1 < hi) ? t1 : hi;                                         \
    if (lo > hi) return make_float2(kNoHit, kNoHit);                  \
  }
  SAILOR_SLAB_AXIS(o.x, inv.x, c.x)
  SAILOR_SLAB_AXIS(o.y, inv.y, c.y)
  SAILOR_SLAB_AXIS(o.z, inv.z, c.z)
#undef SAILOR_SLAB_AXIS
  return make_float2(lo, hi);
}

struct RayCtx {
  const int* vc;          // coarse volume of this batch (node gindex or -1)
  const float4* np;       // node centre + size
  const int4* ni;         // node batch / parent / first child / #children
  int n_nodes;            // entries >= n_nodes are treated as empty
  float3 o, d, inv;
  float3 center;          // centre of the volume bbox
  float3 bst;             // bbox min - half a fine voxel
  float vs_top;           // coarse voxel size
  int dc;
  float rate;

  __device__ __forceinline__ void init(const int* vol_coarse, const float4* node_pos,
                                       const int4* node_info, int num_nodes,
                                       const float* vol_bbox,
                                       const float* voxel_size, int b, int dim_fine,
                                       int dim_coarse, int rate_i, const float* ori,
                                       const float* dir) {
    dc = dim_coarse;
    vc = vol_coarse + (size_t)b * dc * dc * dc;
    np = node_pos;
    ni = node_info;
    n_nodes = num_nodes;
    rate = (float)rate_i;
    const float* bb = vol_bbox + b * 6;
    center = make_float3((bb[0] + bb[1]) / 2.0f, (bb[2] + bb[3]) / 2.0f,
                         (bb[4] + bb[5]) / 2.0f);
    const float vs = voxel_size[b];
    vs_top = vs * (float)(dim_fine / dim_coarse);
    bst = make_float3(bb[0] - vs / 2.0f, bb[2] - vs / 2.0f, bb[4] - vs / 2.0f);
    o = make_float3(ori[0], ori[1], ori[2]);
    d = make_float3(dir[0], dir[1], dir[2]);
    inv = make_float3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
  }
};

// Coarse-grid DDA of a ray (intersection_kernel.cu:128-439): the cell
// (x, y, z) and the ray parameters of the next x / y / z boundaries.
struct Dda {
  int x, y, z, ix, iy, iz;
  float tx, ty, tz, dtx, dty, dtz;
  // entry cell at ray parameter t_in (the volume's slab-test entry)
  __device__ __forceinline__ void start(const RayCtx& c, float t_in) {
    const float spx = c.o.x + t_in * c.d.x;
    const float spy = c.o.y + t_in * c.d.y;
    const float spz = c.o.z + t_in * c.d.z;
    const float sox = spx - c.bst.x, soy = spy - c.bst.y, soz = spz - c.bst.z;
    x = sat_f2i(sox / c.vs_top);
    y = sat_f2i(soy / c.vs_top);
    z = sat_f2i(soz / c.vs_top);
    x = x < 0 ? 0 : x;
    y = y < 0 ? 0 : y;
    z = z < 0 ? 0 : z;
    x = x < c.dc ? x : c.dc - 1;
    y = y < c.dc ? y : c.dc - 1;
    z = z < c.dc ? z : c.dc - 1;
    const float stx = (c.d.x >= 0) ? 1.f : -1.f;
    const float sty = (c.d.y >= 0) ? 1.f : -1.f;
    const float stz = (c.d.z >= 0) ? 1.f : -1.f;
    const float nbx = (stx > 0) ? ((float)x + stx) * c.vs_top : (float)x * c.vs_top;
    const float nby = (sty > 0) ? ((float)y + sty) * c.vs_top : (float)y * c.vs_top;
    const float nbz = (stz > 0) ? ((float)z + stz) * c.vs_top : (float)z * c.vs_top;
    tx = (c.d.x != 0) ? (nbx - sox) / c.d.x : kDdaFar;
    ty = (c.d.y != 0) ? (nby - soy) / c.d.y : kDdaFar;
    tz = (c.d.z != 0) ? (nbz - soz) / c.d.z : kDdaFar;
    dtx = (c.d.x != 0) ? c.vs_top / c.d.x * stx : kDdaFar;
    dty = (c.d.y != 0) ? c.vs_top / c.d.y * sty : kDdaFar;
    dtz = (c.d.z != 0) ? c.vs_top / c.d.z * stz : kDdaFar;
    ix = (int)stx;
    iy = (int)sty;
    iz = (int)stz;
  }
  __device__ __forceinline__ int cell(int dc) const { return (x * dc + y) * dc + z; }
  // one step; false once the cell leaves the grid
  __device__ __forceinline__ bool advance(int dc) {
    if (tx < ty) {
      if (tx < tz) { x += ix; tx += dtx; }
      else { z += iz; tz += dtz; }
    } else {
      if (ty < tz) { y += iy; ty += dty; }
      else { z += iz; tz += dtz; }
    }
    return !(x < 0 || x >= dc || y < 0 || y >= dc || z < 0 || z >= dc);
  }
};

// Sink interface: put(pos, node, t0, t1) receives hit number `pos`;
// stop() is polled after every coarse cell and may end the walk early (used
// by the fused kernel once all samples are placed; never changes hits that
// were already emitted).
//
// Returns the number of hits; *vl_out receives the weighted valid length
// (children weighted by `rate`, the coarse interval by 1).
template <class Sink>
__device__ __forceinline__ int traverse(const RayCtx& c, int th_early_stop,
                                        bool training, int opt_num, Sink& sink,
                                        float* vl_out) {
  float vl = 0.0f;
  int count = 0;
  float2 dep = slab_test(c.o, c.inv, c.center, c.vs_top * (float)c.dc);
  if (dep.x < 0.0f || dep.y < 0.0f) {
    *vl_out = vl;
    return 0;
  }
  Dda g;
  g.start(c, dep.x);
  const int J = 4 * c.dc;
  bool pre_empty = false, has_valid = false;
  int n_succ_empty = 0, n_inter = 0;
  // The DDA's path depends on the ray alone, so the coarse-volume entry of
  // the next cell is loaded while the current cell is visited.
  int node_next = c.vc[g.cell(c.dc)];
  for (int j = 0; j < J; ++j) {
    // ---- visit coarse cell (x, y, z): RayVoxelIntersectionNodeListNew
    int rid = -1;
    const int node = node_next;
    Dda nx = g;   // the cell after this one (the step below)
    const bool next_in = nx.advance(c.dc);
    if (next_in && j + 1 < J) node_next = c.vc[nx.cell(c.dc)];
    if (node > -1 && node < c.n_nodes) {
      // the node's record and its info are loaded together (the info is
      // used only when the slab test hits, which it nearly always does for
      // a cell the DDA visits)
      const float4 p = c.np[node];
      const int4 info = c.ni[node];
      const float2 vd = slab_test(c.o, c.inv, make_float3(p.x, p.y, p.z), p.w);
      if (!(vd.x < 0 || vd.y < 0)) {
        if (info.x >= 0) {
          int r = count;
          if (!training && info.z >= 0) {
            const int ch_end = min(info.z + info.w, c.n_nodes);
            // the next child's record is loaded while this one is tested
            float4 qn = info.z < ch_end ? c.np[info.z] : make_float4(0.f, 0.f, 0.f, 0.f);
            for (int ch = info.z; ch < ch_end; ++ch) {
              const float4 q = qn;
              if (ch + 1 < ch_end) qn = c.np[ch + 1];
              const float2 cd = slab_test(c.o, c.inv, make_float3(q.x, q.y, q.z), q.w);
              if (cd.x < 0 || cd.y < 0) continue;
              sink.put(r, ch, cd.x, cd.y);
              // float += double in the reference (:252): one rounding of the
              // exact double sum
              vl = (float)((double)vl + (double)(cd.y - cd.x) * (double)c.rate);
              ++r;
            }
          }
          sink.put(r, node, vd.x, vd.y);
          vl += vd.y - vd.x;
          ++r;
          count = r;
          rid = r - 1;
        }
      }
    }
    // ---- early termination options (intersection_kernel.cu:407-423)
    if (th_early_stop > -1) {
      if (rid == -1) {
        if (pre_empty && has_valid) n_succ_empty++;
        pre_empty = true;
      } else {
        n_succ_empty = 0;
        pre_empty = false;
        has_valid = true;
      }
      if ((n_succ_empty + 1) >= th_early_stop) break;
    }
    if (opt_num > -1) {
      if (rid + 1 > 0) n_inter++;
      if (n_inter >= opt_num) break;
    }
    if (sink.stop()) break;
    // ---- DDA step (computed above)
    g = nx;
    if (!next_in) break;
  }
  *vl_out = vl;
  return count;
}

// Length-weighted per-interval sampling (sampling_kernel.cu:1125-1154).
struct SampleState {
  float valid_len;
  float start_vs;
  int S;
  __device__ __forceinline__ void init(float vl, float svs, int s) {
    valid_len = vl;
    start_vs = svs;
    S = s;
  }
  // Number of samples interval (mn, mx) of a node of size `cvs` receives
  // when `has` slots are already used.
  __device__ __forceinline__ int count_for(float mn, float mx, float cvs, int has) const {
    const float rate = start_vs / cvs;
    int n = sat_f2i((mx - mn) * rate * (float)S / valid_len);
    n = n >= 1 ? n : 1;
    const int left = S - has;
    if (left < n) n = left;
    return n;
  }
  // Sample j (0-based) of an interval that receives n samples; r = noise.
  __device__ __forceinline__ static float sample_z(float mn, float mx, int n, int j,
                                                   float r) {
    const float off = (mx - mn) / (float)(n + 1);
    const float base = mn + off * (float)(j + 1);
    return (float)((double)base - (double)off / 2.0 + (double)(off * r));
  }
  __device__ __forceinline__ static float leftover_z(float mn, float mx, float r) {
    const float off = (float)((double)(mx - mn) / 2.0);
    const float base = mn + off;
    return (float)((double)base - (double)off / 2.0 + (double)(off * r));
  }
  // Writes the samples of one interval into (oi, ot) from slot `has`.
  __device__ __forceinline__ int interval(int v, float mn, float mx, float cvs,
                                          const float* nz, int* oi, float* ot,
                                          int has) const {
    const int n = count_for(mn, mx, cvs, has);
    for (int j = 0; j < n; ++j) {
      if (has >= S) break;
      ot[has] = sample_z(mn, mx, n, j, nz[has]);
      oi[has] = v;
      ++has;
    }
    return has;
  }
  __device__ __forceinline__ static int leftover(int v, float mn, float mx,
                                                 const float* nz, int* oi, float* ot,
                                                 int has) {
    ot[has] = leftover_z(mn, mx, nz[has]);
    oi[has] = v;
    return has + 1;
  }
};

}  // namespace sailor
