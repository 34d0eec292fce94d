/*
 * CPU oracle for the SAILOR VoxelEncoding path -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain C restatement of the reference CUDA kernels, written from the
 * reference sources (zdlarr/SAILOR, c_lib/VoxelEncoding/src/...), used by the
 * tests, by __graft_entry__.smoke() and by bench.py's CPU baseline as the
 * checker.  Nothing in the product path links or calls it.
 *
 * Arithmetic: IEEE single precision, one rounding per operation (the file is
 * compiled with -ffp-contract=off -fno-fast-math), correctly rounded division
 * and sqrt, double where the reference mixes in double literals.  Fast-math
 * intrinsics of the reference (__fdividef, __expf, __sinf) are replaced by
 * their accurate counterparts.  Float->int conversions use sat_f2i (NaN -> 0,
 * saturating), which is what the GPU conversion instruction does.
 *
 * Where the reference is nondeterministic (node slots taken with atomicAdd)
 * the oracle defines the canonical order: coarse nodes in (batch, raster)
 * order, fine children contiguous per parent in raster order.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static int sat_f2i(float x) {
  if (x != x) return 0;
  if (x >= 2147483648.0f) return 2147483647;
  if (x <= -2147483648.0f) return (int)(-2147483647 - 1);
  return (int)x;
}

/* fusion_kernel.cu:105-135 */
void orc_inverse3x3(const float* Ms, int n, float* inv) {
  for (int i = 0; i < n; ++i) {
    const float* M = Ms + 9 * i;
    float* o = inv + 9 * i;
    float a1 = M[0], b1 = M[1], c1 = M[2], a2 = M[3], b2 = M[4], c2 = M[5],
          a3 = M[6], b3 = M[7], c3 = M[8];
    float det = a1 * (b2 * c3 - c2 * b3) - a2 * (b1 * c3 - c1 * b3) + a3 * (b1 * c2 - c1 * b2);
    o[0] = (b2 * c3 - c2 * b3) / det;
    o[1] = (b3 * c1 - c3 * b1) / det;
    o[2] = (b1 * c2 - c1 * b2) / det;
    o[3] = (a3 * c2 - c3 * a2) / det;
    o[4] = (a1 * c3 - c1 * a3) / det;
    o[5] = (a2 * c1 - c2 * a1) / det;
    o[6] = (a2 * b3 - b2 * a3) / det;
    o[7] = (b1 * a3 - a1 * b3) / det;
    o[8] = (b2 * a1 - a2 * b1) / det;
  }
}

/* obtain_origin_bbox, fusion_kernel.cu:137-184 (min/max over valid pixels);
 * origin initialised to 99999 / -99999 as in :655-656. */
void orc_origin_bbox(const float* depths, const float* K, const float* R, const float* T,
                     int B, int V, int H, int W, float thres, float* origin) {
  float* Kinv = (float*)malloc(sizeof(float) * 9 * B * V);
  float* Rinv = (float*)malloc(sizeof(float) * 9 * B * V);
  orc_inverse3x3(K, B * V, Kinv);
  orc_inverse3x3(R, B * V, Rinv);
  for (int b = 0; b < B; ++b) {
    float* o = origin + 6 * b;
    o[0] = o[2] = o[4] = 99999.0f;
    o[1] = o[3] = o[5] = -99999.0f;
    for (int v = 0; v < V; ++v) {
      const float* ki = Kinv + 9 * (b * V + v);
      const float* ri = Rinv + 9 * (b * V + v);
      const float* t = T + 3 * (b * V + v);
      for (int py = 0; py < H; ++py)
        for (int px = 0; px < W; ++px) {
          float d = depths[(((size_t)b * V + v) * H + py) * W + px];
          if (d < 0.1f || d > 5.0f) continue;
          float fx = (float)px, fy = (float)py;
          float cx = (ki[0] * fx + ki[1] * fy + ki[2]) * d;
          float cy = (ki[3] * fx + ki[4] * fy + ki[5]) * d;
          float cz = (ki[6] * fx + ki[7] * fy + ki[8]) * d;
          cx -= t[0]; cy -= t[1]; cz -= t[2];
          float p[3];
          p[0] = ri[0] * cx + ri[1] * cy + ri[2] * cz;
          p[1] = ri[3] * cx + ri[4] * cy + ri[5] * cz;
          p[2] = ri[6] * cx + ri[7] * cy + ri[8] * cz;
          for (int a = 0; a < 3; ++a) {
            o[2 * a] = fminf(p[a] - thres, o[2 * a]);
            o[2 * a + 1] = fmaxf(p[a] + thres, o[2 * a + 1]);
          }
        }
    }
  }
  free(Kinv);
  free(Rinv);
}

/* get_center_xyz / obtain_center_points, fusion_kernel.cu:186-233,536-598:
 * per batch the sum of the back-projected points of the pixels with
 * 0.1 <= d <= 4 (float32 point arithmetic as the bbox pass) accumulated in
 * double in pixel order, and their count. */
void orc_center_xyz(const float* depths, const float* K, const float* R, const float* T,
                    int B, int V, int H, int W, double* center, double* num) {
  float* Kinv = (float*)malloc(sizeof(float) * 9 * B * V);
  float* Rinv = (float*)malloc(sizeof(float) * 9 * B * V);
  orc_inverse3x3(K, B * V, Kinv);
  orc_inverse3x3(R, B * V, Rinv);
  for (int b = 0; b < B; ++b) {
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int v = 0; v < V; ++v) {
      const float* ki = Kinv + 9 * (b * V + v);
      const float* ri = Rinv + 9 * (b * V + v);
      const float* t = T + 3 * (b * V + v);
      for (int py = 0; py < H; ++py)
        for (int px = 0; px < W; ++px) {
          float d = depths[(((size_t)b * V + v) * H + py) * W + px];
          if (d < 0.1f || d > 4.0f) continue;
          float fx = (float)px, fy = (float)py;
          float cx = (ki[0] * fx + ki[1] * fy + ki[2]) * d;
          float cy = (ki[3] * fx + ki[4] * fy + ki[5]) * d;
          float cz = (ki[6] * fx + ki[7] * fy + ki[8]) * d;
          cx -= t[0]; cy -= t[1]; cz -= t[2];
          acc[0] += (double)(ri[0] * cx + ri[1] * cy + ri[2] * cz);
          acc[1] += (double)(ri[3] * cx + ri[4] * cy + ri[5] * cz);
          acc[2] += (double)(ri[6] * cx + ri[7] * cy + ri[8] * cz);
          acc[3] += 1.0;
        }
    }
    for (int a = 0; a < 3; ++a) center[3 * b + a] = acc[a];
    num[b] = acc[3];
  }
  free(Kinv);
  free(Rinv);
}

/* obtain_resized_bbox, fusion_kernel.cu:51-103 */
void orc_resized_bbox(const float* origin, int B, int dim, float* bbox, float* vs) {
  for (int i = 0; i < B; ++i) {
    const float* o = origin + 6 * i;
    float x_min = o[0], x_max = o[1], y_min = o[2], y_max = o[3], z_min = o[4], z_max = o[5];
    float x_len = x_max - x_min, y_len = y_max - y_min, z_len = z_max - z_min;
    float max_len = fmaxf(x_len, fmaxf(y_len, z_len));
    float x_center = (float)((double)(x_min + x_max) / 2.0);
    float y_center = (float)((double)(y_min + y_max) / 2.0);
    float z_center = (float)((double)(z_min + z_max) / 2.0);
    vs[i] = max_len / (float)(dim - 1);
    if (max_len == x_len) {
      y_min = y_center - max_len / 2.0f; y_max = y_center + max_len / 2.0f;
      z_min = z_center - max_len / 2.0f; z_max = z_center + max_len / 2.0f;
    } else if (max_len == y_len) {
      x_min = x_center - max_len / 2.0f; x_max = x_center + max_len / 2.0f;
      z_min = z_center - max_len / 2.0f; z_max = z_center + max_len / 2.0f;
    } else {
      x_min = x_center - max_len / 2.0f; x_max = x_center + max_len / 2.0f;
      y_min = y_center - max_len / 2.0f; y_max = y_center + max_len / 2.0f;
    }
    float* bb = bbox + 6 * i;
    bb[0] = x_min; bb[1] = x_max; bb[2] = y_min; bb[3] = y_max; bb[4] = z_min; bb[5] = z_max;
  }
}

/* integrate_cuda, fusion_kernel.cu:236-340.  Volumes are fresh (tsdf=1,
 * weight=0, colour=0, :668-670); num_occ is accumulated. */
void orc_integrate(int* occ, int* num_occ, const float* K, const float* RT,
                   const float* colors, const float* depths, int B, int V, int H, int W,
                   int X, int Y, int Z, float obs_weight, float trunc_weight,
                   float th_low, float th_high, const float* bbox, const float* vsz,
                   float* tsdf_vol, float* weight_vol, float* color_vol) {
  const long long vol_size = (long long)X * Y * Z;
  for (int b = 0; b < B; ++b) {
    long long cnt = 0;
#pragma omp parallel for reduction(+ : cnt) schedule(static)
    for (long long v = 0; v < vol_size; ++v) {
      const long long voxel_id = (long long)b * vol_size + v;
      const float trunc_margin = vsz[b] * trunc_weight;
      int vx = (int)(v / ((long long)Y * Z));
      int vy = (int)((v - (long long)vx * Y * Z) / Z);
      int vz = (int)(v - (long long)vx * Y * Z - (long long)vy * Z);
      float pt_x = bbox[b * 6 + 0] + (float)vx * vsz[b];
      float pt_y = bbox[b * 6 + 2] + (float)vy * vsz[b];
      float pt_z = bbox[b * 6 + 4] + (float)vz * vsz[b];
      float tsdf = 1.0f, weight = 0.0f, color = 0.0f;
      for (int k = 0; k < V; ++k) {
        const float* cp = RT + 12 * (b * V + k);
        const float* ci = K + 9 * (b * V + k);
        float cam_x = cp[0] * pt_x + cp[1] * pt_y + cp[2] * pt_z + cp[3];
        float cam_y = cp[4] * pt_x + cp[5] * pt_y + cp[6] * pt_z + cp[7];
        float cam_z = cp[8] * pt_x + cp[9] * pt_y + cp[10] * pt_z + cp[11];
        float pxf = ci[0] * cam_x + ci[1] * cam_y + ci[2] * cam_z;
        float pyf = ci[3] * cam_x + ci[4] * cam_y + ci[5] * cam_z;
        float pzf = ci[6] * cam_x + ci[7] * cam_y + ci[8] * cam_z;
        pxf = pxf / pzf;
        pyf = pyf / pzf;
        int pixel_x = sat_f2i(roundf(pxf));
        int pixel_y = sat_f2i(roundf(pyf));
        if (pixel_x < 0 || pixel_x >= W || pixel_y < 0 || pixel_y >= H || cam_z < 0) continue;
        size_t pixel_idx = (((size_t)b * V + k) * H + pixel_y) * W + pixel_x;
        float depth_value = depths[pixel_idx];
        if (depth_value < 0.1f || depth_value > 4.0f) continue;
        float depth_diff = depth_value - cam_z;
        if (depth_diff < -trunc_margin) continue;
        float dist = fminf(1.0f, depth_diff / trunc_margin);
        float w_old = weight;
        float w_new = w_old + obs_weight;
        weight = w_new;
        tsdf = (tsdf * w_old + obs_weight * dist) / w_new;
        float old_color = color;
        float old_b = floorf(old_color / (float)(256 * 256));
        float old_g = floorf((old_color - old_b * 256 * 256) / 256);
        float old_r = old_color - old_b * 256 * 256 - old_g * 256;
        float new_color = colors[pixel_idx];
        float new_b = floorf(new_color / (float)(256 * 256));
        float new_g = floorf((new_color - new_b * 256 * 256) / 256);
        float new_r = new_color - new_b * 256 * 256 - new_g * 256;
        new_b = fminf(roundf((old_b * w_old + obs_weight * new_b) / w_new), 255.0f);
        new_g = fminf(roundf((old_g * w_old + obs_weight * new_g) / w_new), 255.0f);
        new_r = fminf(roundf((old_r * w_old + obs_weight * new_r) / w_new), 255.0f);
        color = new_b * 256 * 256 + new_g * 256 + new_r;
      }
      tsdf_vol[voxel_id] = tsdf;
      weight_vol[voxel_id] = weight;
      color_vol[voxel_id] = color;
      if (tsdf < trunc_margin * th_high && tsdf > -trunc_margin * th_low) {
        occ[voxel_id] = (int)v;
        cnt++;
      } else {
        occ[voxel_id] = -1;
      }
    }
    num_occ[b] += (int)cnt;
  }
}

/* merge_non_empty_nodes, octree_kernel.cu:242-303 */
void orc_merge(const int* fine, int B, int dc, int rate, int* coarse, int* count) {
  const int df = dc * rate;
  for (int b = 0; b < B; ++b)
    for (int c = 0; c < dc * dc * dc; ++c) {
      int cx = c / (dc * dc), cy = (c - cx * dc * dc) / dc, cz = c - cx * dc * dc - cy * dc;
      int occ = 0;
      for (int k = cx * rate; k < (cx + 1) * rate; ++k)
        for (int m = cy * rate; m < (cy + 1) * rate; ++m)
          for (int n = cz * rate; n < (cz + 1) * rate; ++n)
            occ += fine[(size_t)b * df * df * df + ((size_t)k * df + m) * df + n] > -1;
      if (occ) {
        coarse[(size_t)b * dc * dc * dc + c] = c;
        count[b] += 1;
      } else {
        coarse[(size_t)b * dc * dc * dc + c] = -1;
      }
    }
}

/* if_single_occ_in_neighbors, octree_kernel.cu:101-122 */
static int single_occ(const int* top, int b, int x, int y, int z, int dim) {
  for (int k = x - 1; k < x + 2; ++k)
    for (int m = y - 1; m < y + 2; ++m)
      for (int n = z - 1; n < z + 2; ++n)
        if (k >= 0 && k < dim && m >= 0 && m < dim && n >= 0 && n < dim &&
            !(k == x && m == y && n == z)) {
          size_t idx = (size_t)b * dim * dim * dim + ((size_t)k * dim + m) * dim + n;
          if (top[idx] > -1) return 0;
        }
  return 1;
}

/* gen_root_octree_nodes, octree_kernel.cu:126-239, geometry octree.h:110-151.
 * Canonical slot order; isolated coarse cells are not allocated and set to -1
 * (the reference leaves their dense index in place).  node_pos [cap,4]
 * (x,y,z,size), node_info [cap,4] (batch|-1, parent, first child, #children).
 * Returns 0, or -1 if the nodes do not fit. */
int orc_tree_build(int* vf, int* vc, int B, int df, int dc, const float* voxel_size,
                   const float* vol_bbox, int fine_start, int coarse_start, int cap,
                   float* node_pos, int* node_info, int* queue) {
  const int rate = df / dc;
  const int vol_dim_bottom = df, vol_dim_top = dc;
  for (int i = 0; i < cap; ++i) {
    node_pos[4 * i + 0] = node_pos[4 * i + 1] = node_pos[4 * i + 2] = node_pos[4 * i + 3] = 0.f;
    node_info[4 * i + 0] = -1; node_info[4 * i + 1] = -1; node_info[4 * i + 2] = -1;
    node_info[4 * i + 3] = 0;
  }
  /* isolation decided on the untouched coarse volume */
  const size_t nc = (size_t)dc * dc * dc, nf = (size_t)df * df * df;
  char* alloc = (char*)calloc((size_t)B * nc, 1);
  for (int b = 0; b < B; ++b)
    for (size_t c = 0; c < nc; ++c) {
      int x = (int)(c / (dc * dc)), y = (int)((c - (size_t)x * dc * dc) / dc),
          z = (int)(c - (size_t)x * dc * dc - (size_t)y * dc);
      if (vc[b * nc + c] > -1 && !single_occ(vc, b, x, y, z, dc)) alloc[b * nc + c] = 1;
    }
  int q0 = 0, q1 = 0;
  for (int b = 0; b < B; ++b) {
    float vs_bottom = voxel_size[b];
    float vs_top = vs_bottom * (float)(vol_dim_bottom / vol_dim_top);
    int size_top = vol_dim_bottom / vol_dim_top;
    /* calculate_bbox_start: offset = sum_{i<log2(size)} 2^i * (vs/size) / 2 */
    float basic = vs_top / (float)size_top;
    int rate_vol = 0;
    while ((size_top >> (rate_vol + 1)) > 0) ++rate_vol;
    float offset = 0.0f;
    for (int i = 0; i < rate_vol; ++i) offset += (float)(1 << i) * basic / 2.0f;
    float sx = vol_bbox[6 * b + 0] + offset, sy = vol_bbox[6 * b + 2] + offset,
          sz = vol_bbox[6 * b + 4] + offset;
    float vs_fine = vs_bottom * (float)(vol_dim_bottom / (vol_dim_top * rate));
    float fsx = vol_bbox[6 * b + 0] + 0.0f, fsy = vol_bbox[6 * b + 2] + 0.0f,
          fsz = vol_bbox[6 * b + 4] + 0.0f;
    for (size_t c = 0; c < nc; ++c) {
      if (vc[b * nc + c] < 0) continue;
      if (!alloc[b * nc + c]) { vc[b * nc + c] = -1; continue; }
      int x = (int)(c / (dc * dc)), y = (int)((c - (size_t)x * dc * dc) / dc),
          z = (int)(c - (size_t)x * dc * dc - (size_t)y * dc);
      int g = coarse_start + q0;
      if (g >= cap) { free(alloc); return -1; }
      q0++;
      node_pos[4 * g + 0] = sx + vs_top * (float)x;
      node_pos[4 * g + 1] = sy + vs_top * (float)y;
      node_pos[4 * g + 2] = sz + vs_top * (float)z;
      node_pos[4 * g + 3] = vs_top;
      node_info[4 * g + 0] = b;
      node_info[4 * g + 1] = -1;
      node_info[4 * g + 2] = -1;
      node_info[4 * g + 3] = 0;
      vc[b * nc + c] = g;
      for (int k = x * rate; k < (x + 1) * rate; ++k)
        for (int m = y * rate; m < (y + 1) * rate; ++m)
          for (int n = z * rate; n < (z + 1) * rate; ++n) {
            size_t fi = (size_t)b * nf + ((size_t)k * df + m) * df + n;
            if (vf[fi] < 0) continue;
            int h = fine_start + q1;
            if (h >= coarse_start) { free(alloc); return -1; }
            q1++;
            node_pos[4 * h + 0] = fsx + vs_fine * (float)k;
            node_pos[4 * h + 1] = fsy + vs_fine * (float)m;
            node_pos[4 * h + 2] = fsz + vs_fine * (float)n;
            node_pos[4 * h + 3] = vs_fine;
            node_info[4 * h + 0] = b;
            node_info[4 * h + 1] = g;
            node_info[4 * h + 2] = -1;
            node_info[4 * h + 3] = 0;
            if (node_info[4 * g + 3] == 0) node_info[4 * g + 2] = h;
            node_info[4 * g + 3] += 1;
            vf[fi] = h;
          }
    }
  }
  queue[0] = q0;
  queue[1] = q1;
  free(alloc);
  return 0;
}

/* rays_calculating_kernel, sampling_kernel.cu:411-486 (+ row interleave of
 * rays_calculating_parallel_kernel :348). */
void orc_rays(const float* K, const float* R, const float* T, int B, int V, int H, int W,
              int row_stride, int row_offset, float* out) {
  float* Kinv = (float*)malloc(sizeof(float) * 9 * B * V);
  float* Rinv = (float*)malloc(sizeof(float) * 9 * B * V);
  orc_inverse3x3(K, B * V, Kinv);
  orc_inverse3x3(R, B * V, Rinv);
  for (int bv = 0; bv < B * V; ++bv) {
    const float* ki = Kinv + 9 * bv;
    const float* ri = Rinv + 9 * bv;
    const float* t = T + 3 * bv;
    for (int h = 0; h < H; ++h)
      for (int w = 0; w < W; ++w) {
        int he = h * row_stride + row_offset;
        float cam_center[3], cam_dir[3], world_dir[3];
        for (int k = 0; k < 3; ++k)
          cam_center[k] = -ri[3 * k + 0] * t[0] - ri[3 * k + 1] * t[1] - ri[3 * k + 2] * t[2];
        for (int k = 0; k < 3; ++k)
          cam_dir[k] = (float)w * ki[3 * k + 0] + (float)he * ki[3 * k + 1] + (float)1 * ki[3 * k + 2];
        for (int k = 0; k < 3; ++k)
          world_dir[k] = cam_dir[0] * ri[3 * k + 0] + cam_dir[1] * ri[3 * k + 1] + cam_dir[2] * ri[3 * k + 2];
        float norm_dir = sqrtf(world_dir[0] * world_dir[0] + world_dir[1] * world_dir[1] +
                               world_dir[2] * world_dir[2]);
        float* o = out + (((size_t)bv * H + h) * W + w) * 6;
        for (int k = 0; k < 3; ++k) {
          o[k] = cam_center[k];
          o[3 + k] = (float)((double)world_dir[k] / ((double)norm_dir + 1e-10));
        }
      }
  }
  free(Kinv);
  free(Rinv);
}

/* RayVoxelIntersection, intersection_kernel.cu:36-71 (1.0f/d replaces __fdividef) */
static void ray_voxel(const float* o, const float* d, const float* c, float size, float* lo_hi) {
  float f_low = 0, f_high = 100000.f;
  float half = size / 2.0f;
  for (int i = 0; i < 3; ++i) {
    float inv = 1.0f / d[i];
    float lo = (c[i] - half - o[i]) * inv;
    float hi = (c[i] + half - o[i]) * inv;
    if (lo > hi) { float t = lo; lo = hi; hi = t; }
    if (hi < f_low || lo > f_high) { lo_hi[0] = lo_hi[1] = -1.0f; return; }
    f_low = (lo > f_low) ? lo : f_low;
    f_high = (hi < f_high) ? hi : f_high;
    if (f_low > f_high) { lo_hi[0] = lo_hi[1] = -1.0f; return; }
  }
  lo_hi[0] = f_low;
  lo_hi[1] = f_high;
}

/* ray_voxel_traversal + RayVoxelIntersectionNodeListNew,
 * intersection_kernel.cu:286-439 and :128-284.  Rows are -1 padded.
 * Returns the largest hit count over all rays; *overflow set if any ray
 * needed more than n_max entries. */
int orc_traverse(const int* vc, const float* node_pos, const int* node_info,
                 const float* vol_bbox, const float* voxel_size, int B, int dim_fine,
                 int dim_coarse, int rate, const float* ray_ori, const float* ray_dir, int R,
                 int n_max, int th_early_stop, int is_training, int opt_num, int* hit_idx,
                 float* tmin, float* tmax, float* valid_len, int* overflow) {
  int gmax = 0, ovf = 0;
#pragma omp parallel for reduction(max : gmax) reduction(| : ovf) schedule(dynamic, 256)
  for (long long k = 0; k < (long long)B * R; ++k) {
    int b = (int)(k / R);
    int* hi = hit_idx + k * n_max;
    float* t0 = tmin + k * n_max;
    float* t1 = tmax + k * n_max;
    for (int i = 0; i < n_max; ++i) { hi[i] = -1; t0[i] = -1.0f; t1[i] = -1.0f; }
    const int dtop = dim_coarse;
    const int* vol = vc + (size_t)b * dtop * dtop * dtop;
    const float* bb = vol_bbox + 6 * b;
    float center[3] = {(bb[0] + bb[1]) / 2.0f, (bb[2] + bb[3]) / 2.0f, (bb[4] + bb[5]) / 2.0f};
    float vs = voxel_size[b];
    float vs_top = vs * (float)(dim_fine / dim_coarse);
    float bst[3] = {bb[0] - vs / 2.0f, bb[2] - vs / 2.0f, bb[4] - vs / 2.0f};
    const float* o = ray_ori + 3 * k;
    const float* d = ray_dir + 3 * k;
    float vl = 0.0f;
    int count = 0;
    float dep[2];
    ray_voxel(o, d, center, vs_top * (float)dtop, dep);
    if (!(dep[0] < 0.0f || dep[1] < 0.0f)) {
      float off[3];
      int st[3];
      for (int a = 0; a < 3; ++a) {
        float sp = o[a] + dep[0] * d[a];
        off[a] = sp - bst[a];
        st[a] = sat_f2i(off[a] / vs_top);
        st[a] = st[a] < 0 ? 0 : st[a];
        st[a] = st[a] < dtop ? st[a] : dtop - 1;
      }
      float step[3], tm[3], td[3];
      for (int a = 0; a < 3; ++a) {
        step[a] = (d[a] >= 0) ? 1 : -1;
        float nb = (step[a] > 0) ? ((float)st[a] + step[a]) * vs_top : (float)st[a] * vs_top;
        tm[a] = (d[a] != 0) ? (nb - off[a]) / d[a] : 9999999.0f;
        td[a] = (d[a] != 0) ? vs_top / d[a] * step[a] : 9999999.0f;
      }
      int j = 0, J = 4 * dtop;
      int pre_empty = 0, has_valid = 0, n_succ = 0, n_inter = 0;
      while (j < J) {
        int rid = -1;
        int node = vol[((size_t)st[0] * dtop + st[1]) * dtop + st[2]];
        if (node > -1) {
          float vd[2];
          ray_voxel(o, d, node_pos + 4 * node, node_pos[4 * node + 3], vd);
          if (!(vd[0] < 0 || vd[1] < 0) && node_info[4 * node] >= 0) {
            int r = count;
            if (!is_training) {
              int c0 = node_info[4 * node + 2], nch = node_info[4 * node + 3];
              for (int ci = 0; ci < nch; ++ci) {
                int ch = c0 + ci;
                float cd[2];
                ray_voxel(o, d, node_pos + 4 * ch, node_pos[4 * ch + 3], cd);
                if (cd[0] < 0 || cd[1] < 0 || node_info[4 * ch] < 0) continue;
                if (r < n_max) { hi[r] = ch; t0[r] = cd[0]; t1[r] = cd[1]; } else ovf = 1;
                vl = (float)((double)vl + (double)(cd[1] - cd[0]) * (double)rate);
                r += 1;
              }
            }
            if (r < n_max) { hi[r] = node; t0[r] = vd[0]; t1[r] = vd[1]; } else ovf = 1;
            vl += vd[1] - vd[0];
            r += 1;
            count = r;
            rid = r - 1;
          }
        }
        if (th_early_stop > -1) {
          if (rid == -1) {
            if (pre_empty && has_valid) n_succ++;
            pre_empty = 1;
          } else {
            n_succ = 0;
            pre_empty = 0;
            has_valid = 1;
          }
          if ((n_succ + 1) >= th_early_stop) break;
        }
        if (opt_num > -1) {
          if (rid + 1 > 0) n_inter++;
          if (n_inter >= opt_num) break;
        }
        if (tm[0] < tm[1]) {
          if (tm[0] < tm[2]) { st[0] += (int)step[0]; tm[0] += td[0]; }
          else { st[2] += (int)step[2]; tm[2] += td[2]; }
        } else {
          if (tm[1] < tm[2]) { st[1] += (int)step[1]; tm[1] += td[1]; }
          else { st[2] += (int)step[2]; tm[2] += td[2]; }
        }
        if (st[0] < 0 || st[0] >= dtop || st[1] < 0 || st[1] >= dtop || st[2] < 0 || st[2] >= dtop) break;
        j += 1;
      }
    }
    valid_len[k] = vl;
    if (count > gmax) gmax = count;
  }
  *overflow = ovf;
  return gmax;
}

/* ray_voxels_sampling_uniform_kernel, sampling_kernel.cu:1070-1177.
 * start_scale = rate^(levels - start_level - 1) (the reference's powf(4, ...)).
 * gmax is the global maximum hit count returned by the traversal. */
void orc_sampling(const int* hit_idx, const float* min_depths, const float* max_depths,
                  const float* noise, const float* rays_valid_len, const float* node_pos,
                  const float* voxel_size, int B, int R, int n_max, int S, int start_scale,
                  int gmax, int* sampled_idx, float* sampled_t) {
#pragma omp parallel for schedule(static)
  for (long long k = 0; k < (long long)B * R; ++k) {
    int b = (int)(k / R);
    const int* hi = hit_idx + k * n_max;
    const float* mn_ = min_depths + k * n_max;
    const float* mx_ = max_depths + k * n_max;
    const float* rnd = noise + k * S;
    int* oi = sampled_idx + k * S;
    float* ot = sampled_t + k * S;
    for (int j = 0; j < S; ++j) { oi[j] = -1; ot[j] = -1.0f; }
    if (hi[0] < 0) continue;
    float valid_len = rays_valid_len[k];
    float start_voxel_size = voxel_size[b] * (float)start_scale;
    int has = 0, left;
    int lim = gmax < n_max ? gmax : n_max;
    for (int i = 0; i < lim; ++i) {
      int vi = hi[i];
      float min_depth = mn_[i], max_depth = mx_[i];
      if (vi < 0 || min_depth < 0 || max_depth < 0) break;
      float curr_voxel_size = node_pos[4 * vi + 3];
      float rate = start_voxel_size / curr_voxel_size;
      int n = sat_f2i((max_depth - min_depth) * rate * (float)S / valid_len);
      n = n >= 1 ? n : 1;
      left = S - has;
      if (left < n) n = left;
      float off = (max_depth - min_depth) / (float)(n + 1);
      for (int j = 0; j < n; ++j) {
        if (has >= S) break;
        float r = rnd[has];
        float basic = min_depth + off * (float)(j + 1);
        float z = (float)((double)basic - (double)off / 2.0 + (double)(off * r));
        ot[has] = z;
        oi[has] = vi;
        ++has;
      }
    }
    left = S - has;
    if (left > gmax) left = gmax;
    if (left > n_max) left = n_max;
    for (int i = 0; i < left; ++i) {
      if (has >= S) break;
      int vi = hi[i];
      float min_depth = mn_[i], max_depth = mx_[i];
      if (vi < 0 || min_depth < 0 || max_depth < 0) break;
      float off = (float)((double)(max_depth - min_depth) / 2.0);
      float r = rnd[has];
      float basic = min_depth + off;
      float z = (float)((double)basic - (double)off / 2.0 + (double)(off * r));
      ot[has] = z;
      oi[has] = vi;
      ++has;
    }
  }
}

/* volume_rendering_occ_kernel, volume_rendering.cu:666-745 (F channels) */
void orc_render_occ(const float* t, const float* occ, const float* feats, long long n_rays,
                    int S, int F, float t_thres, float* out_feats, float* out_depth,
                    float* out_ws, float* out_alpha) {
#pragma omp parallel for schedule(static)
  for (long long k = 0; k < n_rays; ++k) {
    const float* tr = t + k * S;
    float* of = out_feats + k * F;
    for (int j = 0; j < F; ++j) of[j] = 0.0f;
    if (tr[S - 1] < 0) { out_depth[k] = out_ws[k] = out_alpha[k] = 0.0f; continue; }
    float T = 1.0f, ws = 0, d = 0, alphas = 0;
    for (int i = 0; i < S; ++i) {
      float z = tr[i];
      if (z < 0) continue;
      float alpha = occ[k * S + i];
      float w = alpha * T;
      for (int j = 0; j < F; ++j) of[j] += w * feats[(k * S + i) * F + j];
      d += w * z;
      ws += w;
      alphas += w * alpha;
      T *= 1.0f - alpha;
      if (T < t_thres) break;
    }
    out_depth[k] = d;
    out_ws[k] = ws;
    out_alpha[k] = alphas;
  }
}

static float next_depth(const float* tr, const int* ir, int i, int S) {
  float nxt = 0.0f;
  for (int m = i + 1; m < S; ++m) {
    nxt = tr[ir[m]];
    if (nxt > 0) break;
  }
  return nxt;
}

/* volume_rendering_training_forward_kernel / _occ_ variant, volume_rendering.cu:17-240 */
void orc_render_train_fwd(const float* t, const int* sidx, const float* sigma,
                          const float* feats, long long n_rays, int S, int F, float t_thres,
                          int occ_mode, float* out_feats, float* out_depth, float* out_ws,
                          float* out_alpha) {
#pragma omp parallel for schedule(static)
  for (long long k = 0; k < n_rays; ++k) {
    const float* tr = t + k * S;
    const int* ir = sidx + k * S;
    float* of = out_feats + k * F;
    for (int j = 0; j < F; ++j) of[j] = 0.0f;
    if (tr[ir[S - 1]] < 0) { out_depth[k] = out_ws[k] = out_alpha[k] = 0.0f; continue; }
    float T = 1.0f, ws = 0, d = 0, alphas = 0;
    for (int i = 0; i < S; ++i) {
      int si = ir[i];
      float z = tr[si];
      if (z < 0) continue;
      float alpha;
      if (occ_mode) alpha = sigma[k * S + si];
      else {
        float delta = (i < S - 1) ? next_depth(tr, ir, i, S) - z : 1e10f;
        alpha = 1.0f - expf(-sigma[k * S + si] * delta);
      }
      float w = alpha * T;
      for (int j = 0; j < F; ++j) of[j] += w * feats[(k * S + si) * F + j];
      d += w * z;
      ws += w;
      alphas += w * alpha;
      T *= 1.0f - alpha;
      if (T < t_thres) break;
    }
    out_depth[k] = d;
    out_ws[k] = ws;
    out_alpha[k] = alphas;
  }
}

/* volume_rendering_training_backward_kernel / _occ_ variant, volume_rendering.cu:244-522 */
void orc_render_train_bwd(const float* g_feats, const float* g_depth, const float* g_ws,
                          const float* g_alpha, const float* o_feats, const float* o_depth,
                          const float* o_ws, const float* o_alpha, const float* t,
                          const int* sidx, const float* sigma, const float* feats,
                          long long n_rays, int S, int F, float t_thres, int occ_mode,
                          float* grad_feats, float* grad_sigma) {
#pragma omp parallel for schedule(static)
  for (long long k = 0; k < n_rays; ++k) {
    const float* tr = t + k * S;
    const int* ir = sidx + k * S;
    for (int i = 0; i < S; ++i) {
      grad_sigma[k * S + i] = 0.0f;
      for (int j = 0; j < F; ++j) grad_feats[(k * S + i) * F + j] = 0.0f;
    }
    if (tr[ir[S - 1]] < 0) continue;
    float* feat = (float*)calloc(F, sizeof(float));
    const float* gf = g_feats + k * F;
    const float* ff = o_feats + k * F;
    float T = 1.0f, ws = 0, d = 0, alphas = 0;
    for (int i = 0; i < S; ++i) {
      int si = ir[i];
      float z = tr[si];
      if (z < 0) continue;
      float alpha, delta = 0.0f;
      if (occ_mode) alpha = sigma[k * S + si];
      else {
        delta = (i < S - 1) ? next_depth(tr, ir, i, S) - z : 1e10f;
        alpha = 1.0f - expf(-sigma[k * S + si] * delta);
      }
      float w = alpha * T;
      const float* fs = feats + (k * S + si) * F;
      for (int j = 0; j < F; ++j) feat[j] += w * fs[j];
      d += w * z;
      ws += w;
      alphas += w * alpha;
      T *= 1.0f - alpha;
      if (T < t_thres) break;
      for (int j = 0; j < F; ++j) grad_feats[(k * S + si) * F + j] = gf[j] * w;
      float gsf = 0.0f;
      for (int j = 0; j < F; ++j) gsf += gf[j] * (T * fs[j] - (ff[j] - feat[j]));
      float g;
      if (occ_mode) {
        float tmp = (1.0 / ((double)(1 - alpha) + 1e-7) <= (double)15.0f) ? 1 / (1 - alpha) : 15.0f;
        float a_ = g_ws[k] * (T * 1 - (o_ws[k] - ws)) * tmp;
        float b_ = g_depth[k] * (T * z - (o_depth[k] - d)) * tmp;
        float c_ = g_alpha[k] * (2 * T * alpha - (o_alpha[k] - alphas)) * tmp;
        gsf *= tmp;
        g = a_ + b_ + gsf + c_;
      } else {
        float tmp = g_ws[k] * (1 - o_ws[k]);
        float a_ = delta * (g_ws[k] * (T * 1 - (o_ws[k] - ws)) + tmp);
        float b_ = delta * (g_depth[k] * (T * z - (o_depth[k] - d)) + tmp);
        float c_ = delta * (g_alpha[k] * (2 * T * alpha - (o_alpha[k] - alphas)) + tmp);
        gsf = delta * (gsf + tmp);
        g = a_ + b_ + gsf + c_;
      }
      grad_sigma[k * S + si] = g;
    }
    free(feat);
  }
}

/* freq_encoding_kernel / _backward_kernel, freq_encoding.cu:16-86 (sinf for __sinf) */
void orc_freq_fwd(const float* x, long long N, int D, int deg, float* y) {
  int C = D + D * deg * 2;
  const float PI = 3.141592653589793f;
  for (long long n = 0; n < N; ++n)
    for (int o = 0; o < C; ++o) {
      if (o < D) y[n * C + o] = x[n * D + o];
      else {
        int d = o % D, col = o / D - 1, f = col / 2;
        float phase = (float)(col % 2) * (PI / 2);
        y[n * C + o] = sinf(scalbnf(x[n * D + d], f) + phase);
      }
    }
}

void orc_freq_bwd(const float* gy, const float* y, long long N, int D, int deg, float* gx) {
  int C = D + D * deg * 2;
  for (long long n = 0; n < N; ++n)
    for (int d = 0; d < D; ++d) {
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
      gx[n * D + d] = r;
    }
}

/* ---------------------------------------------------------------------------
 * Training-only kernels.
 * ------------------------------------------------------------------------- */

/* MultiLayerOctree::octree_traversal, intersection_kernel.cu:738-825:
 * coarse DDA listing the coarse cells (ray_voxel_traversal with
 * RayVoxelIntersectionNodeList :72-126, no valid length), then per coarse hit
 * the DFS of ray_octree_traversal :441-643: the coarse entry with its coarse
 * interval, then the children popped from the stack (last child first) with
 * their slab intervals; an interval equal to an earlier entry of the same
 * coarse hit replaces that entry's node when the child is smaller.  Entries
 * are appended per ray in coarse-hit order (canonical order of the
 * reference's atomicAdd queue) and valid_len accumulates
 * (max - min) * (coarse size / entry size) in the same order.
 * status[0] = max coarse count, status[1] = max fine count, status[2] = 1 if
 * a fine list overflowed n_max (such rays are truncated). */
void orc_octree_traversal(const int* vc, const float* node_pos, const int* node_info,
                          const float* vol_bbox, const float* voxel_size, int B, int dim_fine,
                          int dim_coarse, int rate, const float* ray_ori, const float* ray_dir,
                          int R, int n_max_c, int n_max, int th_early_stop, int opt_num,
                          int* idx_c, float* tmin_c, float* tmax_c, int* idx, float* tmin,
                          float* tmax, float* valid_len, int* status) {
  int ovf_c = 0;
  float* vl_c = (float*)malloc(sizeof(float) * (size_t)B * R);
  int gmax_c = orc_traverse(vc, node_pos, node_info, vol_bbox, voxel_size, B, dim_fine,
                            dim_coarse, rate, ray_ori, ray_dir, R, n_max_c, th_early_stop, 1,
                            opt_num, idx_c, tmin_c, tmax_c, vl_c, &ovf_c);
  free(vl_c);
  int gmax = 0, ovf = 0;
#pragma omp parallel for reduction(max : gmax) reduction(| : ovf) schedule(dynamic, 256)
  for (long long k = 0; k < (long long)B * R; ++k) {
    const int* hc = idx_c + k * n_max_c;
    const float* c0 = tmin_c + k * n_max_c;
    const float* c1 = tmax_c + k * n_max_c;
    int* hi = idx + k * n_max;
    float* t0 = tmin + k * n_max;
    float* t1 = tmax + k * n_max;
    for (int i = 0; i < n_max; ++i) { hi[i] = -1; t0[i] = -1.0f; t1[i] = -1.0f; }
    const float* o = ray_ori + 3 * k;
    const float* d = ray_dir + 3 * k;
    float vl = 0.0f;
    int cnt = 0;
    for (int h = 0; h < n_max_c; ++h) {
      int node = hc[h];
      if (node == -1) break;
      float start_vs = node_pos[4 * node + 3];
      int g0 = cnt;
      if (cnt < n_max) { hi[cnt] = node; t0[cnt] = c0[h]; t1[cnt] = c1[h]; } else ovf = 1;
      cnt += 1;
      int first = node_info[4 * node + 2], nch = node_info[4 * node + 3];
      if (first >= 0) {
        for (int ci = nch - 1; ci >= 0; --ci) {
          int ch = first + ci;
          float cd[2];
          ray_voxel(o, d, node_pos + 4 * ch, node_pos[4 * ch + 3], cd);
          if (cd[0] == -1.0f || cd[1] == -1.0f) continue;
          float vsz = node_pos[4 * ch + 3];
          int found = 0, lim = cnt < n_max ? cnt : n_max;
          for (int j = g0; j < lim; ++j) {
            if (t0[j] == cd[0] && t1[j] == cd[1]) {
              if (node_pos[4 * hi[j] + 3] > vsz) hi[j] = ch;
              found = 1;
              break;
            }
          }
          if (!found) {
            if (cnt < n_max) { hi[cnt] = ch; t0[cnt] = cd[0]; t1[cnt] = cd[1]; } else ovf = 1;
            cnt += 1;
          }
        }
      }
      int lim = cnt < n_max ? cnt : n_max;
      for (int j = g0; j < lim; ++j) {
        float scale = start_vs / node_pos[4 * hi[j] + 3];
        vl += (t1[j] - t0[j]) * scale;
      }
    }
    valid_len[k] = vl;
    if (cnt > gmax) gmax = cnt;
  }
  status[0] = gmax_c;
  status[1] = gmax;
  status[2] = ovf | ovf_c;
}

/* MultiLayerOctree::ray_voxels_points_sampling, sampling_kernel.cu:1178-1248:
 * ray_voxels_sampling_kernel :911-982 places n = max(1, (int)((max - min) *
 * (start_size / size) * S / valid_len)) samples per hit at
 * min + off (i + 1) - off / 2 + off u (off = (max - min) / (n + 1), u the
 * noise of the slot), slots taken in hit order (canonical order of the
 * atomicAdd queue) until S are used, over the first max_hits slots
 * (num_max_intersected_voxels_fine); ray_voxels_further_sampling_kernel
 * :984-1068 fills the remaining slots the same way in the smallest hit node
 * (the one with the nearest interval mid-point among equal sizes).
 * start_scale = 2^(levels - start_level - 1) (powf(2.0, ...) of :952). */
void orc_sampling_adaptive(const int* hit_idx, const float* min_depths, const float* max_depths,
                           const float* noise, const float* rays_valid_len, const float* node_pos,
                           const float* voxel_size, int B, int R, int n_max, int max_hits,
                           int S, int start_scale, int* sampled_idx, float* sampled_t) {
#pragma omp parallel for schedule(static)
  for (long long k = 0; k < (long long)B * R; ++k) {
    int b = (int)(k / R);
    const int* hi = hit_idx + k * n_max;
    const float* mn_ = min_depths + k * n_max;
    const float* mx_ = max_depths + k * n_max;
    const float* rnd = noise + k * S;
    int* oi = sampled_idx + k * S;
    float* ot = sampled_t + k * S;
    for (int j = 0; j < S; ++j) { oi[j] = -1; ot[j] = -1.0f; }
    float valid_len = rays_valid_len[k];
    float start_vs = voxel_size[b] * (float)start_scale;
    int cnt = 0;
    for (int i = 0; i < max_hits; ++i) {
      int vi = hi[i];
      if (vi == -1) continue;
      float mn = mn_[i], mx = mx_[i];
      float cvs = node_pos[4 * vi + 3];
      float rate = start_vs / cvs;
      int n = sat_f2i((mx - mn) * rate * (float)S / valid_len);
      n = n >= 1 ? n : 1;
      float off = (mx - mn) / (float)(n + 1);
      for (int j = 0; j < n; ++j) {
        if (cnt >= S) break;
        int rid = cnt++;
        float u = rnd[rid];
        float basic = mn + off * (float)(j + 1);
        ot[rid] = (float)((double)basic - (double)off / 2.0 + (double)(off * u));
        oi[rid] = vi;
      }
    }
    float min_mid = 99999.0f, min_vs = 99999.0f, z0 = -1.0f, z1 = -1.0f;
    int min_idx = -1;
    for (int i = 0; i < n_max; ++i) {
      int vi = hi[i];
      if (vi == -1) break;
      float vs = node_pos[4 * vi + 3];
      if (vs <= min_vs) {
        float a = mn_[i], c = mx_[i];
        float mid = (float)((double)(a + c) / 2.0);
        if (vs == min_vs) {
          if (mid < min_mid) { min_mid = mid; z0 = a; z1 = c; min_idx = vi; }
        } else {
          min_mid = mid; z0 = a; z1 = c; min_idx = vi;
        }
        min_vs = vs;
      }
    }
    if (min_idx < 0) continue;
    int n_new = S - cnt;
    if (n_new <= 0) continue;
    float off = (z1 - z0) / (float)(n_new + 1);
    int j = 0;
    for (int i = cnt; i < S; ++i) {
      float basic = z0 + off * (float)(j + 1);
      float u = rnd[i];
      ot[i] = (float)((double)basic - (double)off / 2.0 + (double)(off * u));
      oi[i] = min_idx;
      ++j;
    }
  }
}

/* Interpolation weights of trilinear_aggregate_features_kernel,
 * sampling_kernel.cu:162-234: the point ori + dir * t inside node vi (corners
 * centre +- size / 2, corner id 4 x + 2 y + z), tx along x, ty along y from
 * the mean y of the corner pairs, tz along z from the mean z of 4 corners
 * (double literals of :225-227). */
static void tri_weights(const float* node_pos, int vi, const float* o, const float* d, float t,
                        float* tx, float* ty, float* tz) {
  float p[3];
  for (int a = 0; a < 3; ++a) p[a] = o[a] + d[a] * t;
  const float* c = node_pos + 4 * vi;
  float vs = c[3];
  float lo[3], hi[3];
  for (int a = 0; a < 3; ++a) {
    lo[a] = c[a] + ((float)(-1) * vs) / 2.0f;
    hi[a] = c[a] + ((float)(1) * vs) / 2.0f;
  }
  *tx = (p[0] - lo[0]) / (hi[0] - lo[0]);
  double ym = (double)(lo[1] + lo[1]) / 2.0, yp = (double)(hi[1] + hi[1]) / 2.0;
  *ty = (float)(((double)p[1] - ym) / (yp - ym));
  double zm = (double)(((lo[2] + lo[2]) + lo[2]) + lo[2]) / 4.0;
  double zp = (double)(((hi[2] + hi[2]) + hi[2]) + hi[2]) / 4.0;
  *tz = (float)(((double)p[2] - zm) / (zp - zm));
}

static float lerpf_(float a, float b, float t) { return a * (1.0f - t) + b * t; }

/* trilinear_aggregate_features (sampling_kernel.cu:162-234, 1384-1421):
 * feats [B,C,R,S,8] -> out [B,C,R,S]; samples with index -1 give 0. */
void orc_trilinear_fwd(const float* node_pos, const float* ray_ori, const float* ray_dir,
                       const int* sidx, const float* st, const float* feats, int B, int R,
                       int S, int C, float* out) {
#pragma omp parallel for schedule(static)
  for (long long k = 0; k < (long long)B * R * S; ++k) {
    long long b = k / ((long long)R * S), rs = k % ((long long)R * S), r = rs / S;
    int vi = sidx[k];
    float tx = 0, ty = 0, tz = 0;
    if (vi != -1)
      tri_weights(node_pos, vi, ray_ori + 3 * (b * R + r), ray_dir + 3 * (b * R + r), st[k],
                  &tx, &ty, &tz);
    for (int c = 0; c < C; ++c) {
      long long oidx = (b * C + c) * (long long)R * S + rs;
      if (vi == -1) { out[oidx] = 0.0f; continue; }
      const float* f = feats + oidx * 8;
      float x04 = lerpf_(f[0], f[4], tx), x15 = lerpf_(f[1], f[5], tx);
      float x26 = lerpf_(f[2], f[6], tx), x37 = lerpf_(f[3], f[7], tx);
      float y0 = lerpf_(x04, x26, ty), y1 = lerpf_(x15, x37, ty);
      out[oidx] = lerpf_(y0, y1, tz);
    }
  }
}

/* trilinear_aggregate_features_backward (sampling_kernel.cu:237-327):
 * grad [B,C,R,S] -> grad_feats [B,C,R,S,8] (corner weights times grad). */
void orc_trilinear_bwd(const float* node_pos, const float* ray_ori, const float* ray_dir,
                       const int* sidx, const float* st, const float* grad, int B, int R, int S,
                       int C, float* grad_feats) {
#pragma omp parallel for schedule(static)
  for (long long k = 0; k < (long long)B * R * S; ++k) {
    long long b = k / ((long long)R * S), rs = k % ((long long)R * S), r = rs / S;
    int vi = sidx[k];
    float w[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (vi != -1) {
      float tx, ty, tz;
      tri_weights(node_pos, vi, ray_ori + 3 * (b * R + r), ray_dir + 3 * (b * R + r), st[k],
                  &tx, &ty, &tz);
      float z1 = tz, z0 = 1 - tz, y26 = ty, y04 = 1 - ty, y37 = ty, y15 = 1 - ty;
      float x4 = tx, x0 = 1 - tx, x6 = tx, x2 = 1 - tx, x5 = tx, x1 = 1 - tx, x7 = tx,
            x3 = 1 - tx;
      w[0] = x0 * y04 * z0;
      w[1] = x1 * y15 * z1;
      w[2] = x2 * y26 * z0;
      w[3] = x3 * y37 * z1;
      w[4] = x4 * y04 * z0;
      w[5] = x5 * y15 * z1;
      w[6] = x6 * y26 * z0;
      w[7] = x7 * y37 * z1;
    }
    for (int c = 0; c < C; ++c) {
      long long gidx = (b * C + c) * (long long)R * S + rs;
      float g = grad[gidx];
      float* o = grad_feats + gidx * 8;
      for (int i = 0; i < 8; ++i) o[i] = (vi == -1) ? 0.0f : g * w[i];
    }
  }
}

/* project_xyz, sampling_kernel.cu:55-157 (MultiLayerOctree::project_sampled_xyz
 * :1333-1381): xy [B,V,R,S,np,2], z [B,V,R,S,np] with np = 9 (point + 8
 * corners) or 1; -1 where the sample index is -1. */
void orc_project(const float* node_pos, const float* ray_ori, const float* ray_dir,
                 const int* sidx, const float* st, const float* Ks, const float* RTs, int B,
                 int V, int R, int S, int ori_w, int ori_h, int corners, float* xy, float* z) {
  int np_ = corners ? 9 : 1;
  long long total = (long long)B * V * R * S;
#pragma omp parallel for schedule(static)
  for (long long k = 0; k < total; ++k) {
    for (int c = 0; c < np_; ++c) {
      xy[(k * np_ + c) * 2] = -1.0f;
      xy[(k * np_ + c) * 2 + 1] = -1.0f;
      z[k * np_ + c] = -1.0f;
    }
    long long rs = k % ((long long)R * S), bv = k / ((long long)R * S), b = bv / V, r = rs / S;
    long long si = b * R * S + rs;
    int vi = sidx[si];
    if (vi == -1) continue;
    const float* RT = RTs + bv * 12;
    const float* K = Ks + bv * 9;
    const float* o = ray_ori + (b * R + r) * 3;
    const float* d = ray_dir + (b * R + r) * 3;
    float t = st[si];
    for (int c = 0; c < np_; ++c) {
      float q[3];
      if (c == 0) {
        for (int a = 0; a < 3; ++a) q[a] = o[a] + d[a] * t;
      } else {
        const float* nd = node_pos + 4 * vi;
        int id = c - 1;
        int sg[3] = {(id >> 2) & 1 ? 1 : -1, (id >> 1) & 1 ? 1 : -1, id & 1 ? 1 : -1};
        for (int a = 0; a < 3; ++a) q[a] = nd[a] + ((float)sg[a] * nd[3]) / 2.0f;
      }
      float cx = RT[0] * q[0] + RT[1] * q[1] + RT[2] * q[2] + RT[3];
      float cy = RT[4] * q[0] + RT[5] * q[1] + RT[6] * q[2] + RT[7];
      float cz = RT[8] * q[0] + RT[9] * q[1] + RT[10] * q[2] + RT[11];
      float px = K[0] * cx + K[1] * cy + K[2] * cz;
      float py = K[3] * cx + K[4] * cy + K[5] * cz;
      float pz = K[6] * cx + K[7] * cy + K[8] * cz;
      px = px / pz;
      py = py / pz;
      xy[(k * np_ + c) * 2] = (float)(((double)px - ori_w / 2.0) / (ori_w / 2.0));
      xy[(k * np_ + c) * 2 + 1] = (float)(((double)py - ori_h / 2.0) / (ori_h / 2.0));
      z[k * np_ + c] = cz;
    }
  }
}
