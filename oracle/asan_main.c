/* Drives every function of the C oracle (sailor_oracle.c) once on a small
 * synthetic scene -- a sphere seen by three ring cameras -- so that the
 * oracle can be built with AddressSanitizer and UndefinedBehaviorSanitizer
 * (oracle/asan.sh) and every buffer access is checked on the CPU.  Test
 * infrastructure only; exits non-zero if a stage produces nothing. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

void orc_inverse3x3(const float* Ms, int n, float* inv);
void orc_origin_bbox(const float* depths, const float* K, const float* R, const float* T,
                     int B, int V, int H, int W, float thres, float* origin);
void orc_center_xyz(const float* depths, const float* K, const float* R, const float* T, int B,
                    int V, int H, int W, double* center, double* num);
void orc_resized_bbox(const float* origin, int B, int dim, float* bbox, float* vs);
void orc_integrate(int* occ, int* num_occ, const float* K, const float* RT, const float* colors,
                   const float* depths, int B, int V, int H, int W, int X, int Y, int Z,
                   float obs_weight, float trunc_weight, float th_low, float th_high,
                   const float* bbox, const float* vsz, float* tsdf_vol, float* weight_vol,
                   float* color_vol);
void orc_merge(const int* fine, int B, int dc, int rate, int* coarse, int* count);
int orc_tree_build(int* vf, int* vc, int B, int df, int dc, const float* voxel_size,
                   const float* vol_bbox, int fine_start, int coarse_start, int cap,
                   float* node_pos, int* node_info, int* queue);
void orc_rays(const float* K, const float* R, const float* T, int B, int V, int H, int W,
              int row_stride, int row_offset, float* out);
int orc_traverse(const int* vc, const float* node_pos, const int* node_info,
                 const float* vol_bbox, const float* voxel_size, int B, int dim_fine,
                 int dim_coarse, int rate, const float* ray_ori, const float* ray_dir, int R,
                 int n_max, int th_early_stop, int is_training, int opt_num, int* hit_idx,
                 float* tmin, float* tmax, float* valid_len, int* overflow);
void orc_sampling(const int* hit_idx, const float* min_depths, const float* max_depths,
                  const float* noise, const float* rays_valid_len, const float* node_pos,
                  const float* voxel_size, int B, int R, int n_max, int S, int start_scale,
                  int gmax, int* sampled_idx, float* sampled_t);
void orc_render_occ(const float* t, const float* occ, const float* feats, long long n_rays,
                    int S, int F, float t_thres, float* out_feats, float* out_depth,
                    float* out_ws, float* out_alpha);
void orc_render_train_fwd(const float* t, const int* sidx, const float* sigma,
                          const float* feats, long long n_rays, int S, int F, float t_thres,
                          int occ_mode, float* out_feats, float* out_depth, float* out_ws,
                          float* out_alpha);
void orc_render_train_bwd(const float* g_feats, const float* g_depth, const float* g_ws,
                          const float* g_alpha, const float* o_feats, const float* o_depth,
                          const float* o_ws, const float* o_alpha, const float* t,
                          const int* sidx, const float* sigma, const float* feats,
                          long long n_rays, int S, int F, float t_thres, int occ_mode,
                          float* grad_feats, float* grad_sigma);
void orc_freq_fwd(const float* x, long long N, int D, int deg, float* y);
void orc_freq_bwd(const float* gy, const float* y, long long N, int D, int deg, float* gx);
void orc_octree_traversal(const int* vc, const float* node_pos, const int* node_info,
                          const float* vol_bbox, const float* voxel_size, int B, int dim_fine,
                          int dim_coarse, int rate, const float* ray_ori, const float* ray_dir,
                          int R, int n_max_c, int n_max, int th_early_stop, int opt_num,
                          int* idx_c, float* tmin_c, float* tmax_c, int* idx, float* tmin,
                          float* tmax, float* valid_len, int* status);
void orc_sampling_adaptive(const int* hit_idx, const float* min_depths, const float* max_depths,
                           const float* noise, const float* rays_valid_len,
                           const float* node_pos, const float* voxel_size, int B, int R,
                           int n_max, int max_hits, int S, int start_scale, int* sampled_idx,
                           float* sampled_t);
void orc_trilinear_fwd(const float* node_pos, const float* ray_ori, const float* ray_dir,
                       const int* sidx, const float* st, const float* feats, int B, int R,
                       int S, int C, float* out);
void orc_trilinear_bwd(const float* node_pos, const float* ray_ori, const float* ray_dir,
                       const int* sidx, const float* st, const float* grad, int B, int R, int S,
                       int C, float* grad_feats);
void orc_project(const float* node_pos, const float* ray_ori, const float* ray_dir,
                 const int* sidx, const float* st, const float* Ks, const float* RTs, int B,
                 int V, int R, int S, int ori_w, int ori_h, int corners, float* xy, float* z);

#define NEW(T, n) ((T*)calloc((size_t)(n), sizeof(T)))

static float frand(unsigned* s) {
  *s = *s * 1664525u + 1013904223u;
  return (float)((*s >> 8) & 0xffffff) / 16777216.0f;
}

/* look-at camera at c towards a (y up): R rows = right, down, forward */
static void camera(const double c[3], const double a[3], double f, int W, int H, float* K,
                   float* R, float* T) {
  double z[3] = {a[0] - c[0], a[1] - c[1], a[2] - c[2]};
  double n = sqrt(z[0] * z[0] + z[1] * z[1] + z[2] * z[2]);
  for (int i = 0; i < 3; ++i) z[i] /= n;
  double up[3] = {0, 1, 0};
  double x[3] = {z[1] * up[2] - z[2] * up[1], z[2] * up[0] - z[0] * up[2],
                 z[0] * up[1] - z[1] * up[0]};
  n = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
  for (int i = 0; i < 3; ++i) x[i] /= n;
  double y[3] = {z[1] * x[2] - z[2] * x[1], z[2] * x[0] - z[0] * x[2], z[0] * x[1] - z[1] * x[0]};
  double Rm[9] = {x[0], x[1], x[2], y[0], y[1], y[2], z[0], z[1], z[2]};
  for (int i = 0; i < 9; ++i) R[i] = (float)Rm[i];
  for (int i = 0; i < 3; ++i)
    T[i] = (float)-(Rm[3 * i] * c[0] + Rm[3 * i + 1] * c[1] + Rm[3 * i + 2] * c[2]);
  float Km[9] = {(float)f, 0, W / 2.0f, 0, (float)f, H / 2.0f, 0, 0, 1};
  memcpy(K, Km, sizeof(Km));
}

int main(void) {
  const int V = 3, H = 48, W = 48, dim = 64, rate = 4, dc = dim / rate;
  const double ctr[3] = {0.0, 1.0, 0.0}, rad = 0.45;
  float *K = NEW(float, 9 * V), *R = NEW(float, 9 * V), *T = NEW(float, 3 * V);
  float *RT = NEW(float, 12 * V), *depth = NEW(float, V * H * W), *col = NEW(float, V * H * W);
  for (int v = 0; v < V; ++v) {
    double th = 2.0 * 3.14159265358979323846 * v / V;
    double c[3] = {2.0 * sin(th), 1.2, 2.0 * cos(th)};
    camera(c, ctr, 40.0, W, H, K + 9 * v, R + 9 * v, T + 3 * v);
    for (int i = 0; i < 3; ++i) {
      for (int j = 0; j < 3; ++j) RT[12 * v + 4 * i + j] = R[9 * v + 3 * i + j];
      RT[12 * v + 4 * i + 3] = T[3 * v + i];
    }
    for (int py = 0; py < H; ++py)
      for (int px = 0; px < W; ++px) {
        /* camera ray K^-1 [px, py, 1] (z = 1), world direction R^T d */
        double dc3[3] = {(px - W / 2.0) / 40.0, (py - H / 2.0) / 40.0, 1.0}, w[3];
        for (int i = 0; i < 3; ++i)
          w[i] = R[9 * v + i] * dc3[0] + R[9 * v + 3 + i] * dc3[1] + R[9 * v + 6 + i] * dc3[2];
        double oc[3] = {c[0] - ctr[0], c[1] - ctr[1], c[2] - ctr[2]};
        double A = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
        double Bq = 2 * (oc[0] * w[0] + oc[1] * w[1] + oc[2] * w[2]);
        double Cq = oc[0] * oc[0] + oc[1] * oc[1] + oc[2] * oc[2] - rad * rad;
        double disc = Bq * Bq - 4 * A * Cq;
        float d = 0.0f;
        if (disc > 0) d = (float)((-Bq - sqrt(disc)) / (2 * A));
        depth[(v * H + py) * W + px] = d;
        col[(v * H + py) * W + px] = (float)((px * 7 + py * 3) % 255);
      }
  }
  float origin[6], bbox[6], vs;
  double center[3], npts;
  orc_origin_bbox(depth, K, R, T, 1, V, H, W, 0.025f, origin);
  orc_center_xyz(depth, K, R, T, 1, V, H, W, center, &npts);
  orc_resized_bbox(origin, 1, dim, bbox, &vs);
  const int N = dim * dim * dim;
  int *occ = NEW(int, N), num = 0;
  float *tsdf = NEW(float, N), *wt = NEW(float, N), *cv = NEW(float, N);
  orc_integrate(occ, &num, K, RT, col, depth, 1, V, H, W, dim, dim, dim, 1.0f, 3.0f, 13.5f,
                13.5f, bbox, &vs, tsdf, wt, cv);
  int *coarse = NEW(int, dc * dc * dc), cnum = 0;
  orc_merge(occ, 1, dc, rate, coarse, &cnum);
  const int cap = num + cnum;
  float* node_pos = NEW(float, 4 * cap);
  int *node_info = NEW(int, 4 * cap), queue[2] = {0, 0};
  int rc = orc_tree_build(occ, coarse, 1, dim, dc, &vs, bbox, 0, num, cap, node_pos, node_info,
                          queue);
  /* target view between two sources */
  float Kt[9], Rt[9], Tt[3];
  const double ct[3] = {2.0 * sin(0.5), 1.0, 2.0 * cos(0.5)};
  camera(ct, ctr, 30.0, 32, 32, Kt, Rt, Tt);
  const int Rn = 32 * 32, nmax = 128, S = 16;
  float* rays = NEW(float, 6 * Rn);
  orc_rays(Kt, Rt, Tt, 1, 1, 32, 32, 1, 0, rays);
  float *ori = NEW(float, 3 * Rn), *dir = NEW(float, 3 * Rn);
  for (int i = 0; i < Rn; ++i)
    for (int a = 0; a < 3; ++a) {
      ori[3 * i + a] = rays[6 * i + a];
      dir[3 * i + a] = rays[6 * i + 3 + a];
    }
  int* hit = NEW(int, Rn * nmax);
  float *t0 = NEW(float, Rn * nmax), *t1 = NEW(float, Rn * nmax), *vl = NEW(float, Rn);
  int ovf = 0;
  int gmax = orc_traverse(coarse, node_pos, node_info, bbox, &vs, 1, dim, dc, rate, ori, dir, Rn,
                          nmax, -1, 0, -1, hit, t0, t1, vl, &ovf);
  float* noise = NEW(float, Rn * S);
  unsigned seed = 7;
  for (int i = 0; i < Rn * S; ++i) noise[i] = frand(&seed);
  int* sidx = NEW(int, Rn * S);
  float* st = NEW(float, Rn * S);
  orc_sampling(hit, t0, t1, noise, vl, node_pos, &vs, 1, Rn, nmax, S, rate, gmax, sidx, st);
  /* compositing forward / backward on the samples (identity order) */
  const int F = 3;
  float *feat = NEW(float, Rn * S * F), *sig = NEW(float, Rn * S);
  int* order = NEW(int, Rn * S);
  for (int i = 0; i < Rn * S; ++i) {
    sig[i] = frand(&seed);
    order[i] = i % S;
    for (int c = 0; c < F; ++c) feat[i * F + c] = frand(&seed);
  }
  float *of = NEW(float, Rn * F), *od = NEW(float, Rn), *ow = NEW(float, Rn), *oa = NEW(float, Rn);
  orc_render_occ(st, sig, feat, Rn, S, F, 1e-4f, of, od, ow, oa);
  float *gf = NEW(float, Rn * F), *gd = NEW(float, Rn), *gw = NEW(float, Rn), *ga = NEW(float, Rn);
  for (int i = 0; i < Rn; ++i) {
    gd[i] = gw[i] = ga[i] = 0.5f;
    for (int c = 0; c < F; ++c) gf[i * F + c] = 1.0f;
  }
  float *gfe = NEW(float, Rn * S * F), *gsi = NEW(float, Rn * S);
  for (int mode = 0; mode < 2; ++mode) {
    orc_render_train_fwd(st, order, sig, feat, Rn, S, F, 1e-4f, mode, of, od, ow, oa);
    orc_render_train_bwd(gf, gd, gw, ga, of, od, ow, oa, st, order, sig, feat, Rn, S, F, 1e-4f,
                         mode, gfe, gsi);
  }
  /* frequency encoding of the ray origins */
  float* y = NEW(float, Rn * 63);
  float* gx = NEW(float, Rn * 3);
  orc_freq_fwd(ori, Rn, 3, 10, y);
  orc_freq_bwd(y, y, Rn, 3, 10, gx);
  /* training-only path */
  const int nc = 3 * dc, nf = 128;
  int *idc = NEW(int, Rn * nc), *idx = NEW(int, Rn * nf), status[3] = {0, 0, 0};
  float *t0c = NEW(float, Rn * nc), *t1c = NEW(float, Rn * nc);
  float *t0f = NEW(float, Rn * nf), *t1f = NEW(float, Rn * nf), *vlf = NEW(float, Rn);
  orc_octree_traversal(coarse, node_pos, node_info, bbox, &vs, 1, dim, dc, rate, ori, dir, Rn, nc,
                       nf, -1, -1, idc, t0c, t1c, idx, t0f, t1f, vlf, status);
  orc_sampling_adaptive(idx, t0f, t1f, noise, vlf, node_pos, &vs, 1, Rn, nf,
                        status[1] < nf ? status[1] : nf, S, 2, sidx, st);
  const int C = 2;
  float *tf = NEW(float, C * Rn * S * 8), *to = NEW(float, C * Rn * S);
  for (int i = 0; i < C * Rn * S * 8; ++i) tf[i] = frand(&seed);
  orc_trilinear_fwd(node_pos, ori, dir, sidx, st, tf, 1, Rn, S, C, to);
  orc_trilinear_bwd(node_pos, ori, dir, sidx, st, to, 1, Rn, S, C, tf);
  float *pxy = NEW(float, V * Rn * S * 9 * 2), *pz = NEW(float, V * Rn * S * 9);
  orc_project(node_pos, ori, dir, sidx, st, K, RT, 1, V, Rn, S, W, H, 1, pxy, pz);
  float kinv[9];
  orc_inverse3x3(K, 1, kinv);
  printf("asan driver: %d occupied voxels, %d coarse, tree rc %d (%d, %d nodes), %d max hits, "
         "%g points, fine entries %d\n", num, cnum, rc, queue[0], queue[1], gmax, npts, status[1]);
  int ok = num > 0 && cnum > 0 && rc == 0 && gmax > 0 && npts > 0 && status[1] > 0 && !ovf;
  /* buffers are released by the process exit; LeakSanitizer is not the point here */
  return ok ? 0 : 1;
}
