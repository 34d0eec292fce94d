// Fused per-ray radiance kernel: feature gather + positional encoding +
// per-view density MLP on MFMA + view mean + occupancy head + front-to-back
// compositing + multi-view Hydra attention + RGB head.
//
// Semantics follow the reference's FastNerf path (c_lib/FastNerf/src/
// infer_mlp.cu: density_feat_kernel :120-199 per (point, view), torch::mean
// over views :698-699, density_sigma_kernel :202-239, per-view feature
// compositing along the ray :724-747 with volume_rendering_occ_kernel
// :508-581, hydra_attention_kernel :242-445 on the composited per-view
// tokens, RGB head :427-432) and the PyTorch modules it fuses
// (SRONet/nerf.py:139-186, SRONet/aggregation.py:85-128,155-171,288-307),
// with the widths of the benchmark model as template parameters instead of
// the reference's hard-coded 64/32/30.
//
// MI355X mapping (layouts in mlp_layout.h)
//  * one workgroup (4 waves) owns one ray at a time (persistent grid-stride
//    loop over rays); the ray's samples are processed in chunks of NS samples
//    x 4 views = NC columns, in sample order;
//  * the per-(sample, view) MLP runs transposed on MFMA: columns are the NC
//    (sample, view) pairs, each wave owns a quarter of the output rows, so a
//    weight fragment read from L2 is reused across all NC/16 column tiles;
//  * layer inputs / activations of the chunk live in LDS as act[col][k],
//    unpadded rows of H elements whose 16-byte chunks are XOR-swizzled by the
//    column (chunk c of column col is stored at c ^ p(col), p a permutation
//    of col & 15).  Measured on gfx950 with the SQ bank-conflict counter,
//    the XOR swizzle is the layout (among padded strides of 16/32 mod 256
//    bytes and partial swizzles) for which both the 16-byte B-fragment reads
//    and the 16-byte stores of the layer outputs are conflict-free; the
//    16-bit forms' permutation also keeps the backward's column-wise
//    (transposing) reads conflict-free;
//  * each lane's accumulators hold 4*MTW consecutive output units of one
//    column (the packer orders the weight rows so), so a layer's outputs are
//    written back with whole 16-byte stores;
//    the fp32 accumulators of a layer are written back as the next layer's
//    input after ReLU (bf16/fp16 forms store the activations in 16 bit);
//  * feature gather (channels-last maps: 4 lanes fetch the 4 quarters of a
//    tap's channel row) and positional encoding write the first layer's input
//    directly into LDS; per-sample occupancy and compositing weights never
//    reach HBM; the composited per-view tokens (4 x F floats per ray) do;
//  * the view mean is folded into the occupancy head (mean of per-column dot
//    products); compositing keeps the reference's sequential recurrence;
//  * Hydra + RGB head run in a second kernel over 16 rays (64 tokens) per
//    workgroup: lane = token, so the 4 views of a ray are 4 adjacent lanes
//    (cross-view sums are two shuffles), the weights are wave-uniform reads
//    shared by all 64 tokens, and the 4 waves split each layer's outputs.
#include <type_traits>

#include "common.h"
#include "mlp_layout.h"

namespace {

using namespace sailor::mlp;

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x4 __attribute__((ext_vector_type(4)));

enum Prec { kF32 = 0, kBF16 = 1, kF16 = 2 };
constexpr int kThreads = kWaves * 64;
constexpr int kVP = 4;       // view slots per sample (training form, point mode)
constexpr int kMaxViews = 8;  // view slots of the 8-view rendering instantiations
constexpr int kXcds = 8;      // accelerator dies of an MI355X (workgroups go round-robin)
constexpr float kHalfPi = 1.5707963267948966f;

struct RenderArgs {
  const float* t;          // [n_rays, S] sorted sample distances
  const float* rays;       // [n_rays, 6] origin, direction
  const float* Kv;         // [B, V, 3, 3] source intrinsics
  const float* RTv;        // [B, V, 3, 4] source extrinsics
  const void* fmap;        // [B, V, Hf, Wf, F] channels-last maps, fp32 or the form's 16 bit
  int fmap_half;           // maps stored in the form's 16-bit type
  const float* bbox;       // [B, 3, 2] volume box (positional-encoding frame)
  const float* zdir;       // [B, 3] third row of the target rotation
  const void* wdense;      // packed dense layers
  const float* bdense;     // [4, H] biases
  const float* wsmall;     // occupancy head, Hydra, RGB head (mlp_layout.h)
  float* out_rgb;          // [n_rays, 3]
  float* out_tok;          // [n_valid, 4, F] composited per-view features (workspace)
  const int* list;         // [n_valid] rays with samples, in list order (workspace)
  const unsigned* n_valid; // device count of list entries (workspace)
  float* out_depth;        // [n_rays]
  float* out_aux;          // [n_rays, 2] (weight sum, alpha sum) or null
  float* out_occ;          // [n_rays, S] per-sample occupancy or null
  long long n_rays;
  int rays_per_batch;
  int S, V, img_w, img_h, Hf, Wf;
  float t_thres;
  int early_exit;
  int compensated;         // Kahan-compensated compositing sums (precision form)
  const float* fscale;     // [F] per-channel scales of quantised maps (stored x / s) or null
  int grid_share;          // percent of the resident workgroup slots the sample kernel takes (0: all)
  int fp32_hydra;          // 16-bit forms: Hydra + RGB head on fp32 MFMA instead of 16-bit
  void* wsmall16;          // 16-bit copy of wsmall, element i at index i + 4 (workspace)
  int stages;              // 1: compaction, 2: sample kernel, 4: Hydra/RGB kernel
  // point mode (density query): points [B, n_points, 3]; groups of NS points
  // take the place of rays and out_occ receives one density per point
  const float* points;
  long long n_points;
};

template <int PREC> struct Tr;
template <> struct Tr<kF32> {
  typedef float E;
  static constexpr int EPL = 1, CE = 4;  // CE: elements per 16-byte chunk
};
template <> struct Tr<kBF16> {
  typedef __bf16 E;
  typedef bf16x8 V8;
  typedef bf16x4 V4;
  static constexpr int EPL = 8, CE = 8;
};
template <> struct Tr<kF16> {
  typedef _Float16 E;
  typedef f16x8 V8;
  typedef f16x4 V4;
  static constexpr int EPL = 8, CE = 8;
};

__device__ __forceinline__ float sigmoidf(float x) { return 1.0f / (1.0f + expf(-x)); }

// One Kahan step: sum += x with the rounding error kept in comp (negated);
// the compensated value is sum - comp.  (No fast-math here: the compiler
// does not reassociate (t - sum) - y, and it contains no product to fuse.)
__device__ __forceinline__ void kahan_add(float& sum, float& comp, float x) {
  const float y = x - comp;
  const float t = sum + y;
  comp = (t - sum) - y;
  sum = t;
}

// Element offset of 16-byte chunk `chunk` of column `col` in the swizzled
// activation layout: chunk ^ (col & 15) for fp32; chunk ^ q(col) for the
// 16-bit forms, q(col) = (col & 3) << 1 | (col >> 2) & 1 | (col & 8).  The
// LDS serves 128 bytes (8 chunk positions) per pass: 8 consecutive columns at
// one chunk (the forward's B reads, the layer stores) take 8 different
// positions under either map, and under q so do the 4 aligned consecutive
// columns x 2 consecutive chunks of a pass of the backward's transposing
// column reads (q's low bit is free for the chunk's).
template <int H, int PREC>
__device__ __forceinline__ int act_chunk(int col, int chunk) {
  if constexpr (PREC == kF32) {
    return col * H + ((chunk ^ (col & 15)) * Tr<PREC>::CE);
  } else {
    const int q = ((col & 3) << 1) | ((col >> 2) & 1) | (col & 8);
    return col * H + ((chunk ^ q) * Tr<PREC>::CE);
  }
}

// Output unit of accumulator element i of m-tile mt held by lane group g in
// wave w: lane (n, g) holds units [w*H/4 + 4*MTW*g, +4*MTW) of column n.
template <int H>
__device__ __forceinline__ int unit_of(int wave, int g, int mt, int i) {
  return (H / 4) * wave + 4 * (H / 64) * g + 4 * mt + i;
}

// act[col][unit] = relu(acc + bias) for this wave's units (16-byte stores).
template <int H, int PREC, int NTL>
__device__ __forceinline__ void write_act(typename Tr<PREC>::E* act, const float* bias,
                                          int wave, int lane, const f32x4 (&acc)[H / 64][NTL]) {
  typedef typename Tr<PREC>::E E;
  constexpr int MTW = H / 64, CE = Tr<PREC>::CE;
  constexpr int NV = 4 * MTW;  // consecutive units per lane and column
  const int n = lane & 15, g = lane >> 4;
  const int u0 = unit_of<H>(wave, g, 0, 0);
  float bv[NV];
#pragma unroll
  for (int q = 0; q < NV; q += 4) {
    const f32x4 b4 = *reinterpret_cast<const f32x4*>(bias + u0 + q);
    bv[q] = b4[0]; bv[q + 1] = b4[1]; bv[q + 2] = b4[2]; bv[q + 3] = b4[3];
  }
#pragma unroll
  for (int nt = 0; nt < NTL; ++nt) {
    const int col = 16 * nt + n;
#pragma unroll
    for (int c = 0; c < NV / CE; ++c) {
      E* dst = act + act_chunk<H, PREC>(col, u0 / CE + c);
      if constexpr (PREC == kF32) {
        f32x4 v;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int q = 4 * c + e;
          v[e] = fmaxf(acc[q / 4][nt][q % 4] + bv[q], 0.0f);
        }
        *reinterpret_cast<f32x4*>(dst) = v;
      } else {
        typename Tr<PREC>::V8 h;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          const int q = 8 * c + e;
          h[e] = (E)fmaxf(acc[q / 4][nt][q % 4] + bv[q], 0.0f);
        }
        *reinterpret_cast<typename Tr<PREC>::V8*>(dst) = h;
      }
    }
  }
}

// Per-precision shape of the MFMA work.  A "unit" is fp32: 4 MFMA steps
// that share one 16-byte B read; 16-bit: one step.  Weight (A) fragments of a
// unit are fetched PD units ahead (NB = PD + 1 rotating buffers, NB divides
// the units of a hidden layer so the rotation is the same in every hidden
// layer); activation (B) fragments are read from LDS one unit ahead.
template <int H, int PREC, int NTL>
struct Chain {
  static constexpr bool F32 = PREC == kF32;
  static constexpr int MTW = H / 64;
  static constexpr int KPU = F32 ? 16 : 32;   // reduction index per unit
  static constexpr int UE = F32 ? 4 : 8;      // elements per lane per m-tile per unit
  static constexpr int UH = H / KPU;
  // 16-bit: three units ahead where registers allow (one wave per SIMD, up
  // to 8 column tiles), else one
  static constexpr int PD = F32 ? 1 : ((NTL > 4 && NTL < 12) ? 3 : 1);
  static constexpr int NB = PD + 1;
  static constexpr int AN = F32 ? 4 : MTW;    // fp32: 4 steps x f32x4 over m-tiles
  typedef typename Tr<PREC>::E E;
  typedef typename std::conditional<F32, f32x4, typename Tr<F32 ? kBF16 : PREC>::V8>::type V;
  static constexpr size_t unit_bytes = (size_t)64 * MTW * UE * sizeof(E);
  static_assert(UH % NB == 0, "hidden-layer units must fill whole buffer rotations");
};

template <int H, int PREC, int NTL>
__device__ __forceinline__ void load_a(const char* base, int j, unsigned lane_off,
                                       typename Chain<H, PREC, NTL>::V (&dst)[Chain<H, PREC, NTL>::AN]) {
  typedef Chain<H, PREC, NTL> C;
  const char* ub = base + (size_t)j * C::unit_bytes + lane_off;
  if constexpr (C::F32) {
    // lane data of step i: MTW floats at (i * 64 + lane) * MTW
    static_assert(C::MTW == 4 || C::MTW == 2, "fp32 A unit layout");
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const char* ap = ub + (size_t)i * 64 * C::MTW * sizeof(float);
      if constexpr (C::MTW == 4) {
        dst[i] = *reinterpret_cast<const f32x4*>(ap);
      } else {
        const f32x2 v = *reinterpret_cast<const f32x2*>(ap);
        dst[i] = f32x4{v[0], v[1], 0.f, 0.f};
      }
    }
  } else {
    const typename C::V* ap = reinterpret_cast<const typename C::V*>(ub);
#pragma unroll
    for (int mt = 0; mt < C::MTW; ++mt) dst[mt] = ap[mt];
  }
}

template <int H, int PREC, int NTL>
__device__ __forceinline__ void read_b(const typename Tr<PREC>::E* act, int j, int n, int g,
                                       typename Chain<H, PREC, NTL>::V (&dst)[NTL]) {
  typedef Chain<H, PREC, NTL> C;
  // unit j covers chunks 4j..4j+3; lane group g reads chunk 4j + g
#pragma unroll
  for (int nt = 0; nt < NTL; ++nt)
    dst[nt] = *reinterpret_cast<const typename C::V*>(act + act_chunk<H, PREC>(16 * nt + n,
                                                                               4 * j + g));
}

// UL units of one layer whose first PD units are already in the A buffers
// (unit j in buffer (O + j) % NB) and whose first B unit is in bbuf[0]; the
// loads that run past the layer's end fetch the first PD units of the layer
// at `nxt`.  The scheduling barriers keep both prefetches where they are
// written: left alone, the scheduler sinks the loads next to their first use
// and exposes the full memory latency at every unit.
template <int H, int PREC, int NTL, int UL, int O>
__device__ __forceinline__ void run_layer(
    const char* cur, const char* nxt, unsigned lane_off, const typename Tr<PREC>::E* act, int n,
    int g, typename Chain<H, PREC, NTL>::V (&abuf)[Chain<H, PREC, NTL>::NB][Chain<H, PREC, NTL>::AN],
    typename Chain<H, PREC, NTL>::V (&bbuf)[2][NTL], f32x4 (&acc)[H / 64][NTL]) {
  typedef Chain<H, PREC, NTL> C;
  constexpr int NB = C::NB, PD = C::PD, MTW = C::MTW;
#pragma unroll
  for (int j = 0; j < UL; ++j) {
    if (j + PD < UL)
      load_a<H, PREC, NTL>(cur, j + PD, lane_off, abuf[(O + j + PD) % NB]);
    else
      load_a<H, PREC, NTL>(nxt, j + PD - UL, lane_off, abuf[(O + j + PD) % NB]);
    if (j + 1 < UL) read_b<H, PREC, NTL>(act, j + 1, n, g, bbuf[(j + 1) & 1]);
    __builtin_amdgcn_sched_barrier(0);
    const typename C::V (&A)[C::AN] = abuf[(O + j) % NB];
    const typename C::V (&Bv)[NTL] = bbuf[j & 1];
    if constexpr (C::F32) {
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int nt = 0; nt < NTL; ++nt)
#pragma unroll
          for (int mt = 0; mt < MTW; ++mt)
            acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x4f32(A[i][mt], Bv[nt][i], acc[mt][nt],
                                                               0, 0, 0);
    } else {
#pragma unroll
      for (int nt = 0; nt < NTL; ++nt)
#pragma unroll
        for (int mt = 0; mt < MTW; ++mt) {
          if constexpr (PREC == kBF16)
            acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[mt], Bv[nt], acc[mt][nt],
                                                                  0, 0, 0);
          else
            acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x32_f16(A[mt], Bv[nt], acc[mt][nt],
                                                                 0, 0, 0);
        }
    }
    __builtin_amdgcn_sched_barrier(0);
  }
}

// The four dense layers of one chunk: acc = W3 relu(W2 relu(W1 relu(W0 x +
// b0) + b1) + b2); bias and ReLU of the last layer are applied by the caller.
// Layer l's weights for wave w are [w][unit][lane][m-tile][slot] (see
// mlp_layout.h); the weight prefetch runs across layer boundaries so the
// loads of a layer's first units overlap the previous layer's epilogue.
template <int H, int PREC, int NTL, int K0>
__device__ __forceinline__ void dense_chain(const typename Tr<PREC>::E* __restrict__ W,
                                            typename Tr<PREC>::E* act,
                                            const float* __restrict__ s_bias, int wave,
                                            int lane, f32x4 (&acc)[H / 64][NTL]) {
  typedef Chain<H, PREC, NTL> C;
  typedef typename C::E E;
  constexpr int U0 = K0 / C::KPU, UH = C::UH, NB = C::NB, PD = C::PD, MTW = C::MTW;
  static_assert(K0 % C::KPU == 0, "layer widths must fill whole units");
  constexpr int O0 = (NB - U0 % NB) % NB;   // so that layer 1 starts at buffer 0
  const int n = lane & 15, g = lane >> 4;
  // Wave-uniform bases (scalar registers) plus a per-lane byte offset: unit
  // addresses are compile-time offsets from them.  The empty asm makes the
  // bases opaque so unit addresses are formed next to their loads instead of
  // being hoisted out of the ray loop as live vector registers.
  const char* wb0 = reinterpret_cast<const char*>(W) + (size_t)wave * U0 * C::unit_bytes;
  const char* wbh = reinterpret_cast<const char*>(W) + (size_t)kWaves * U0 * C::unit_bytes +
                    (size_t)wave * UH * C::unit_bytes;
  asm volatile("" : "+s"(wb0), "+s"(wbh));
  const size_t lh_bytes = (size_t)kWaves * UH * C::unit_bytes;
  const unsigned lane_off = (unsigned)lane * MTW * (C::F32 ? 1 : 8) * sizeof(E);

  typename C::V abuf[NB][C::AN];
  typename C::V bbuf[2][NTL];
  auto zero = [&]() {
#pragma unroll
    for (int mt = 0; mt < MTW; ++mt)
#pragma unroll
      for (int nt = 0; nt < NTL; ++nt) acc[mt][nt] = f32x4{0.f, 0.f, 0.f, 0.f};
  };
  zero();
#pragma unroll
  for (int j = 0; j < PD; ++j) load_a<H, PREC, NTL>(wb0, j, lane_off, abuf[(O0 + j) % NB]);
  read_b<H, PREC, NTL>(act, 0, n, g, bbuf[0]);
  run_layer<H, PREC, NTL, U0, O0>(wb0, wbh, lane_off, act, n, g, abuf, bbuf, acc);
#pragma unroll 1
  for (int l = 1; l < 4; ++l) {
    __syncthreads();  // every wave finished reading act
    write_act<H, PREC, NTL>(act, s_bias + (l - 1) * H, wave, lane, acc);
    zero();
    __syncthreads();
    read_b<H, PREC, NTL>(act, 0, n, g, bbuf[0]);
    const char* cur = wbh + (size_t)(l - 1) * lh_bytes;
    const char* nxt = l < 3 ? cur + lh_bytes : cur;  // the last layer re-reads its own head
    run_layer<H, PREC, NTL, UH, 0>(cur, nxt, lane_off, act, n, g, abuf, bbuf, acc);
  }
}

// ---- 16-bit forms: 32x32x16 MFMA ------------------------------------------
// (lane l holds A[l%32][8(l/32)+j] and B[8(l/32)+j][l%32]; accumulator r holds
// D[8(r/4) + 4(l/32) + r%4][l%32], checked on gfx950).  The packer orders the
// weight rows so that accumulator r of m-tile mt in lanes with l/32 = h is
// output unit unit_of32(w, mt, h, r): 16 consecutive units per lane and
// column.  A unit of work is one k-step of 16: M32 x NT MFMAs of 32x32x16,
// which sustain close to the dense peak with few independent accumulators
// (the 16x16x32 form needs 16 chains per wave for the same rate).
template <int H>
__device__ __forceinline__ int unit_of32(int wave, int mt, int h, int r) {
  return (H / 4) * wave + 32 * mt + 16 * h + r;
}

template <int H, int PREC, int NT>
struct Chain16 {
  static constexpr int M32 = H / 128;   // 32-row m-tiles per wave
  static constexpr int UH = H / 16;     // k-steps of a hidden layer
  static constexpr int PD = 3, NB = 4;  // weight prefetch distance / buffers
  typedef typename Tr<PREC>::V8 V;
  static constexpr size_t unit_bytes = (size_t)64 * M32 * 8 * 2;
  static_assert(UH % NB == 0, "hidden-layer steps must fill whole buffer rotations");
};

template <int H, int PREC, int NT>
__device__ __forceinline__ void load_a16(const char* base, int j, unsigned lane_off,
                                         typename Chain16<H, PREC, NT>::V (&dst)[H / 128]) {
  typedef Chain16<H, PREC, NT> C;
  const typename C::V* ap =
      reinterpret_cast<const typename C::V*>(base + (size_t)j * C::unit_bytes + lane_off);
#pragma unroll
  for (int mt = 0; mt < C::M32; ++mt) dst[mt] = ap[mt];
}

// k-step j covers chunks 2j, 2j+1 of a column; lanes l/32 = h read chunk 2j+h
template <int H, int PREC, int NT>
__device__ __forceinline__ void read_b16(const typename Tr<PREC>::E* act, int j, int lane,
                                         typename Chain16<H, PREC, NT>::V (&dst)[NT]) {
  typedef Chain16<H, PREC, NT> C;
#pragma unroll
  for (int nt = 0; nt < NT; ++nt)
    dst[nt] = *reinterpret_cast<const typename C::V*>(
        act + act_chunk<H, PREC>(32 * nt + (lane & 31), 2 * j + (lane >> 5)));
}

template <int H, int PREC, int NT, int UL, int O>
__device__ __forceinline__ void run_layer16(
    const char* cur, const char* nxt, unsigned lane_off, const typename Tr<PREC>::E* act,
    int lane, typename Chain16<H, PREC, NT>::V (&abuf)[Chain16<H, PREC, NT>::NB][H / 128],
    typename Chain16<H, PREC, NT>::V (&bbuf)[2][NT], f32x16 (&acc)[H / 128][NT]) {
  typedef Chain16<H, PREC, NT> C;
  constexpr int NB = C::NB, PD = C::PD;
#pragma unroll
  for (int j = 0; j < UL; ++j) {
    if (j + PD < UL)
      load_a16<H, PREC, NT>(cur, j + PD, lane_off, abuf[(O + j + PD) % NB]);
    else
      load_a16<H, PREC, NT>(nxt, j + PD - UL, lane_off, abuf[(O + j + PD) % NB]);
    if (j + 1 < UL) read_b16<H, PREC, NT>(act, j + 1, lane, bbuf[(j + 1) & 1]);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int nt = 0; nt < NT; ++nt)
#pragma unroll
      for (int mt = 0; mt < C::M32; ++mt) {
        if constexpr (PREC == kBF16)
          acc[mt][nt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(abuf[(O + j) % NB][mt],
                                                                bbuf[j & 1][nt], acc[mt][nt],
                                                                0, 0, 0);
        else
          acc[mt][nt] = __builtin_amdgcn_mfma_f32_32x32x16_f16(abuf[(O + j) % NB][mt],
                                                               bbuf[j & 1][nt], acc[mt][nt],
                                                               0, 0, 0);
      }
    __builtin_amdgcn_sched_barrier(0);
  }
}

// act[col][unit] = relu(acc + bias) for this wave's units (two 16-byte stores
// per m-tile and column).
template <int H, int PREC, int NT>
__device__ __forceinline__ void write_act16(typename Tr<PREC>::E* act, const float* bias,
                                            int wave, int lane, const f32x16 (&acc)[H / 128][NT]) {
  typedef typename Tr<PREC>::E E;
  const int c = lane & 31, h = lane >> 5;
#pragma unroll
  for (int mt = 0; mt < H / 128; ++mt) {
    const int u0 = unit_of32<H>(wave, mt, h, 0);
    float bv[16];
#pragma unroll
    for (int q = 0; q < 16; q += 4) {
      const f32x4 b4 = *reinterpret_cast<const f32x4*>(bias + u0 + q);
      bv[q] = b4[0]; bv[q + 1] = b4[1]; bv[q + 2] = b4[2]; bv[q + 3] = b4[3];
    }
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
#pragma unroll
      for (int e = 0; e < 2; ++e) {
        typename Tr<PREC>::V8 v;
#pragma unroll
        for (int k = 0; k < 8; ++k) v[k] = (E)fmaxf(acc[mt][nt][8 * e + k] + bv[8 * e + k], 0.0f);
        *reinterpret_cast<typename Tr<PREC>::V8*>(act + act_chunk<H, PREC>(32 * nt + c,
                                                                           u0 / 8 + e)) = v;
      }
    }
  }
}

template <int H, int PREC, int NT, int K0>
__device__ __forceinline__ void dense_chain16(const typename Tr<PREC>::E* __restrict__ W,
                                              typename Tr<PREC>::E* act,
                                              const float* __restrict__ s_bias, int wave,
                                              int lane, f32x16 (&acc)[H / 128][NT]) {
  typedef Chain16<H, PREC, NT> C;
  constexpr int U0 = K0 / 16, UH = C::UH, NB = C::NB, PD = C::PD, M32 = C::M32;
  static_assert(K0 % 16 == 0, "layer widths must fill whole k-steps");
  constexpr int O0 = (NB - U0 % NB) % NB;   // so that layer 1 starts at buffer 0
  const char* wb0 = reinterpret_cast<const char*>(W) + (size_t)wave * U0 * C::unit_bytes;
  const char* wbh = reinterpret_cast<const char*>(W) + (size_t)kWaves * U0 * C::unit_bytes +
                    (size_t)wave * UH * C::unit_bytes;
  asm volatile("" : "+s"(wb0), "+s"(wbh));  // see dense_chain
  const size_t lh_bytes = (size_t)kWaves * UH * C::unit_bytes;
  const unsigned lane_off = (unsigned)lane * M32 * 16;

  typename C::V abuf[NB][M32];
  typename C::V bbuf[2][NT];
  auto zero = [&]() {
#pragma unroll
    for (int mt = 0; mt < M32; ++mt)
#pragma unroll
      for (int nt = 0; nt < NT; ++nt)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[mt][nt][r] = 0.0f;
  };
  zero();
#pragma unroll
  for (int j = 0; j < PD; ++j) load_a16<H, PREC, NT>(wb0, j, lane_off, abuf[(O0 + j) % NB]);
  read_b16<H, PREC, NT>(act, 0, lane, bbuf[0]);
  run_layer16<H, PREC, NT, U0, O0>(wb0, wbh, lane_off, act, lane, abuf, bbuf, acc);
#pragma unroll 1
  for (int l = 1; l < 4; ++l) {
    __syncthreads();  // every wave finished reading act
    write_act16<H, PREC, NT>(act, s_bias + (l - 1) * H, wave, lane, acc);
    zero();
    __syncthreads();
    read_b16<H, PREC, NT>(act, 0, lane, bbuf[0]);
    const char* cur = wbh + (size_t)(l - 1) * lh_bytes;
    const char* nxt = l < 3 ? cur + lh_bytes : cur;  // the last layer re-reads its own head
    run_layer16<H, PREC, NT, UH, 0>(cur, nxt, lane_off, act, lane, abuf, bbuf, acc);
  }
}

template <int F, int H, int PREC, int NS, int VP = kVP>
struct Smem {
  static constexpr int NC = NS * VP;
  static constexpr size_t act = 0;
  static constexpr size_t act_bytes = (size_t)NC * H * sizeof(typename Tr<PREC>::E);
  // kept-feature rows of F + 4 floats: an odd number of 16-byte chunks, so
  // the 16-byte stores of 16 columns are conflict-free
  static constexpr int KF = F + 4;
  static constexpr size_t keep = (act + act_bytes + 15) / 16 * 16;      // f32 [NS][VP][KF]
  static constexpr size_t sig = keep + (size_t)NS * VP * KF * 4;       // f32 [4][NC]
  static constexpr size_t alpha = sig + (size_t)kWaves * NC * 4;       // f32 [NS]
  static constexpr size_t tval = alpha + (size_t)NS * 4;               // f32 [NS]
  static constexpr size_t wts = tval + (size_t)NS * 4;                 // f32 [NS]
  static constexpr size_t flag = wts + (size_t)NS * 4;                 // int [4]
  // f32 [4][H] layer biases, then (fp32 form) [H] the occupancy weights;
  // the 16-bit forms read those from L1, which keeps their 24-sample chunks
  // within two workgroups per CU
  static constexpr size_t bias = flag + 16;
  static constexpr size_t pos = bias + (size_t)(PREC == kF32 ? 5 : 4) * H * 4;  // f32 [NS][4]
  static constexpr size_t total = pos + (size_t)NS * 16;
};

constexpr size_t kLdsBytes = 160 * 1024;  // LDS per workgroup on gfx950

// Bilinear feature gather of one quarter of the F channels (CPL = F/4) of
// view v at world point (x, y, z): proj_persp (utils_render.py:380-410; depth
// < 0.01 reads nothing) + grid_sample(align_corners=True, zero padding)
// (SRONet/BasicRenNet.py:245-259) on channels-last maps, one 16-byte load
// per 4 fp32 or 8 16-bit channels of a tap.
template <int F, int PREC, bool HALF_MAP>
__device__ __forceinline__ void gather_quarter(const RenderArgs& a, int b, int v, float x,
                                               float y, float z, int quarter,
                                               float (&f)[F / 4]) {
  typedef typename Tr<PREC>::E E;
  constexpr int CPL = F / 4;
  const int V = a.V;
#pragma unroll
  for (int c = 0; c < CPL; ++c) f[c] = 0.0f;
  if (v >= V) return;
  const float* RT = a.RTv + ((size_t)b * V + v) * 12;
  const float* K = a.Kv + ((size_t)b * V + v) * 9;
  float px = RT[0] * x + RT[1] * y + RT[2] * z + RT[3];
  float py = RT[4] * x + RT[5] * y + RT[6] * z + RT[7];
  const float pz = RT[8] * x + RT[9] * y + RT[10] * z + RT[11];
  if (!(pz >= 0.01f)) return;
  px /= pz;
  py /= pz;
  const float u = K[0] * px + K[1] * py + K[2];
  const float w = K[3] * px + K[4] * py + K[5];
  const float hw = (float)(a.img_w / 2), hh2 = (float)(a.img_h / 2);
  const float gx = (u - hw) / hw, gy = (w - hh2) / hh2;
  const float ix = (gx + 1.0f) * 0.5f * (float)(a.Wf - 1);
  const float iy = (gy + 1.0f) * 0.5f * (float)(a.Hf - 1);
  if (!(ix > -1.0f && ix < (float)a.Wf && iy > -1.0f && iy < (float)a.Hf)) return;
  const float fx0 = floorf(ix), fy0 = floorf(iy);
  const int x0 = (int)fx0, y0 = (int)fy0;
  const float ax = ix - fx0, ay = iy - fy0;
  const size_t base = ((size_t)b * V + v) * a.Hf * a.Wf * F + quarter * CPL;
#pragma unroll
  for (int tap = 0; tap < 4; ++tap) {
    const int xx = x0 + (tap & 1), yy = y0 + (tap >> 1);
    if (xx < 0 || xx >= a.Wf || yy < 0 || yy >= a.Hf) continue;
    const float wt = ((tap & 1) ? ax : 1.0f - ax) * ((tap >> 1) ? ay : 1.0f - ay);
    const size_t off = base + ((size_t)yy * a.Wf + xx) * F;
    if constexpr (HALF_MAP) {
      typedef typename Tr<PREC>::V8 V8;
      const V8* src = reinterpret_cast<const V8*>(static_cast<const E*>(a.fmap) + off);
#pragma unroll
      for (int qq = 0; qq < CPL / 8; ++qq) {
        const V8 v8 = src[qq];
#pragma unroll
        for (int e = 0; e < 8; ++e) f[8 * qq + e] += wt * (float)v8[e];
      }
    } else {
      const f32x4* src = reinterpret_cast<const f32x4*>(static_cast<const float*>(a.fmap) + off);
#pragma unroll
      for (int qq = 0; qq < CPL / 4; ++qq) {
        const f32x4 v4 = src[qq];
        f[4 * qq + 0] += wt * v4[0];
        f[4 * qq + 1] += wt * v4[1];
        f[4 * qq + 2] += wt * v4[2];
        f[4 * qq + 3] += wt * v4[3];
      }
    }
  }
}

// Waves per SIMD the register allocation is fitted to: two workgroups per CU
// let one workgroup's gather / compositing overlap the other's MFMA work.
// (chunks of at most 96 columns -- up to 24 samples x 4 view slots -- of
// the 16-bit forms and of the fp32 hidden-128 model: 256 registers and 80 KB
// of LDS per workgroup; fp32 hidden-256 activations take 96 KB).
constexpr int render_occupancy(int PREC, int NC, int H) {
  return (NC <= 96 && (PREC != kF32 || H <= 128)) ? 2 : 1;
}

// VP view slots per sample: 4 (up to 4 views) or 8 (up to 8 views); the
// columns of a chunk are (sample, view slot) pairs, col = sample * VP + slot.
template <int F, int H, int PREC, int NS, bool HALF_MAP, bool POINTS, int VP>
__global__ void __launch_bounds__(kThreads)
__attribute__((amdgpu_waves_per_eu(render_occupancy(PREC, NS * VP, H),
                                   render_occupancy(PREC, NS * VP, H))))
render_kernel(RenderArgs a) {
  typedef typename Tr<PREC>::E E;
  typedef Smem<F, H, PREC, NS, VP> L;
  constexpr int NC = L::NC, KF = L::KF, NTL = NC / 16, MTW = H / 64;
  constexpr int CE = Tr<PREC>::CE;
  constexpr int CPL = F / 4;
  constexpr int K0 = F + kPeSlots;
  constexpr int STEP = PREC == kF32 ? 4 : 32;
  constexpr int KS0 = K0 / STEP, KSH = H / STEP;
  constexpr int EPL = Tr<PREC>::EPL;
  constexpr SmallOffsets so = small_offsets(F, H);
  static_assert(K0 <= H, "first-layer input must fit an activation row");
  static_assert(CPL % CE == 0 && F % CE == 0, "feature quarters must fill whole chunks");
  constexpr int TE = (VP * F + kThreads - 1) / kThreads;  // token entries per thread
  static_assert(VP == 4 || VP == 8, "4 or 8 view slots");

  extern __shared__ __attribute__((aligned(16))) char smem[];
  E* act = reinterpret_cast<E*>(smem + L::act);
  float* keep = reinterpret_cast<float*>(smem + L::keep);
  float* sig = reinterpret_cast<float*>(smem + L::sig);
  float* s_alpha = reinterpret_cast<float*>(smem + L::alpha);
  float* s_t = reinterpret_cast<float*>(smem + L::tval);
  float* s_w = reinterpret_cast<float*>(smem + L::wts);
  int* s_flag = reinterpret_cast<int*>(smem + L::flag);
  float* s_pos = reinterpret_cast<float*>(smem + L::pos);

  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int S = a.S, V = a.V;
  const float* ws = a.wsmall;
  const E* W0 = reinterpret_cast<const E*>(a.wdense);
  float* s_bias = reinterpret_cast<float*>(smem + L::bias);
  // layer biases and the occupancy-head weights stay in LDS for the whole
  // kernel (the first loop iteration's barrier publishes them)
  for (int i = tid; i < 4 * H; i += kThreads) s_bias[i] = a.bdense[i];
  if constexpr (PREC == kF32)
    for (int i = tid; i < H; i += kThreads) s_bias[4 * H + i] = ws[so.sigma_w + i];
  const float inv_v = 1.0f / (float)V;

  // every list entry costs about the same, so a static stride over the
  // compacted list balances the workgroups (dropped rays were finished by the
  // compaction kernel).  Point mode: item = a group of NS points.
  const long long n_items = POINTS ? a.n_rays : (long long)*a.n_valid;
  // Workgroups are dispatched to the 8 XCDs round-robin (blockIdx % 8), each
  // XCD with its own L2.  Every XCD takes one contiguous band of the list
  // (raster order), so the rays its workgroups hold at any time are image
  // neighbours whose bilinear taps overlap in that L2.
  const int bands = (int)gridDim.x < kXcds ? (int)gridDim.x : kXcds;
  const int xcd = blockIdx.x % bands;
  const int wg_in_xcd = blockIdx.x / bands;
  const int wgs_of_xcd = ((int)gridDim.x - xcd + bands - 1) / bands;
  const long long band = (n_items + bands - 1) / bands;
  const long long item_lo = band * xcd;
  const long long item_hi = item_lo + band < n_items ? item_lo + band : n_items;
  for (long long item = item_lo + wg_in_xcd; item < item_hi; item += wgs_of_xcd) {
    __syncthreads();  // LDS of the previous ray is no longer read
    const long long ray = POINTS ? item : (long long)a.list[item];
    const int b = (int)(ray / a.rays_per_batch);
    const long long p0 = (ray % a.rays_per_batch) * NS;  // point mode: first point
    const float* trow = POINTS ? nullptr : a.t + ray * S;
    const float* rr = POINTS ? a.points : a.rays + ray * 6;
    const float ox = rr[0], oy = rr[1], oz = rr[2];
    const float dx = POINTS ? 0.0f : rr[3], dy = POINTS ? 0.0f : rr[4],
                dz = POINTS ? 0.0f : rr[5];
    const float* bb = a.bbox + b * 6;
    const float cx = 0.5f * (bb[0] + bb[1]), cy = 0.5f * (bb[2] + bb[3]),
                cz = 0.5f * (bb[4] + bb[5]);
    const float ihx = 2.0f / (bb[1] - bb[0]), ihy = 2.0f / (bb[3] - bb[2]),
                ihz = 2.0f / (bb[5] - bb[4]);
    // compositing state lives in thread 0; tokens: thread (v, c) = (tid / F, tid % F)
    float T = 1.0f, dsum = 0.0f, wsum = 0.0f, asum = 0.0f;
    float dcmp = 0.0f, wcmp = 0.0f, acmp = 0.0f;  // compensated form: lost low bits
    bool done = false;
    float tok[TE], tcmp[TE];
#pragma unroll
    for (int j = 0; j < TE; ++j) tok[j] = tcmp[j] = 0.0f;
    if (tid == 0) s_flag[0] = 0;

    for (int c0 = 0; c0 < S; c0 += NS) {
      __syncthreads();
      if (a.early_exit && s_flag[0]) break;
      // sample positions of the chunk: o + t d, or the points themselves
      if (tid < NS) {
        float px, py, pz, tv = 0.0f;
        if constexpr (POINTS) {
          const long long pi = p0 + tid < a.n_points ? p0 + tid : a.n_points - 1;
          const float* q = a.points + ((long long)b * a.n_points + pi) * 3;
          px = q[0]; py = q[1]; pz = q[2];
        } else {
          tv = trow[c0 + tid];
          px = ox + tv * dx; py = oy + tv * dy; pz = oz + tv * dz;
        }
        s_pos[4 * tid + 0] = px; s_pos[4 * tid + 1] = py; s_pos[4 * tid + 2] = pz;
        s_t[tid] = tv;
      }
      __syncthreads();
      // ---- first-layer input: bilinear features (4 threads per column; the
      // 16 lanes of a quarter-wave take 16 columns of the same quarter) ------
      for (int item = tid; item < NC * 4; item += kThreads) {
        const int col = (item & 15) + 16 * (item >> 6), quarter = (item >> 4) & 3;
        const int sl = col / VP, v = col % VP;
        const float x = s_pos[4 * sl], y = s_pos[4 * sl + 1], z = s_pos[4 * sl + 2];
        float f[CPL];
        gather_quarter<F, PREC, HALF_MAP>(a, b, v, x, y, z, quarter, f);
        float* kp = keep + col * KF + quarter * CPL;
#pragma unroll
        for (int c = 0; c < CPL; c += 4)
          *reinterpret_cast<f32x4*>(kp + c) = f32x4{f[c], f[c + 1], f[c + 2], f[c + 3]};
        // layer-0 input k = channel (k < F); 16-byte chunks, swizzled
#pragma unroll
        for (int c = 0; c < CPL; c += CE) {
          E* dst = act + act_chunk<H, PREC>(col, (quarter * CPL + c) / CE);
          if constexpr (PREC == kF32) {
            *reinterpret_cast<f32x4*>(dst) = f32x4{f[c], f[c + 1], f[c + 2], f[c + 3]};
          } else {
            typename Tr<PREC>::V8 h;
#pragma unroll
            for (int i = 0; i < 8; ++i) h[i] = (E)f[c + i];
            *reinterpret_cast<typename Tr<PREC>::V8*>(dst) = h;
          }
        }
      }
      // ---- positional encoding (freq_encoding.cu:37-45 layout), shared by views;
      // a thread computes two adjacent slots and stores them as one pair
      for (int item = tid; item < NS * kPeSlots / 2; item += kThreads) {
        const int sl = item / (kPeSlots / 2), q0 = 2 * (item % (kPeSlots / 2));
        float val[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int q = q0 + h;
          val[h] = 0.0f;
          if (q < 63) {
            const int d = q % 3;
            const float xd = d == 0 ? (s_pos[4 * sl] - cx) * ihx
                                    : (d == 1 ? (s_pos[4 * sl + 1] - cy) * ihy
                                              : (s_pos[4 * sl + 2] - cz) * ihz);
            if (q < 3) {
              val[h] = xd;
            } else {
              const int col = q / 3 - 1;
              val[h] = sinf(ldexpf(xd, col >> 1) + (float)(col & 1) * kHalfPi);
            }
          }
        }
#pragma unroll
        for (int v = 0; v < VP; ++v) {
          const int col = sl * VP + v, k = F + q0;  // the pair never straddles a chunk
          E* d2 = act + act_chunk<H, PREC>(col, k / CE) + k % CE;
          if constexpr (PREC == kF32) {
            *reinterpret_cast<f32x2*>(d2) = f32x2{val[0], val[1]};
          } else {
            typedef E E2 __attribute__((ext_vector_type(2)));
            *reinterpret_cast<E2*>(d2) = E2{(E)val[0], (E)val[1]};
          }
        }
      }
      __syncthreads();
      // ---- dense layers on MFMA --------------------------------------------
      // ---- dense layers + occupancy head (per-column partial dot products) --
      if constexpr (PREC != kF32) {
        constexpr int NT = NC / 32;
        static_assert(NC % 32 == 0, "16-bit forms need chunks of a multiple of 8 samples");
        f32x16 acc[H / 128][NT];
        dense_chain16<H, PREC, NT, K0>(W0, act, s_bias, wave, lane, acc);
        const int c = lane & 31, h = lane >> 5;
        const float* bias = s_bias + 3 * H;
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
          float part = 0.0f;
#pragma unroll
          for (int mt = 0; mt < H / 128; ++mt) {
            const int u0 = unit_of32<H>(wave, mt, h, 0);
#pragma unroll
            for (int r = 0; r < 16; ++r)
              part += ws[so.sigma_w + u0 + r] * fmaxf(acc[mt][nt][r] + bias[u0 + r], 0.0f);
          }
          part += __shfl_xor(part, 32, 64);
          if (h == 0) sig[wave * NC + 32 * nt + c] = part;
        }
      } else {
        f32x4 acc[MTW][NTL];
        dense_chain<H, PREC, NTL, K0>(W0, act, s_bias, wave, lane, acc);
        const int n = lane & 15, g = lane >> 4;
        const float* bias = s_bias + 3 * H;
#pragma unroll
        for (int nt = 0; nt < NTL; ++nt) {
          float part = 0.0f;
#pragma unroll
          for (int mt = 0; mt < MTW; ++mt) {
            const int row0 = unit_of<H>(wave, g, mt, 0);
            const f32x4 bv = *reinterpret_cast<const f32x4*>(bias + row0);
            const f32x4 wv = *reinterpret_cast<const f32x4*>(s_bias + 4 * H + row0);
#pragma unroll
            for (int i = 0; i < 4; ++i) part += wv[i] * fmaxf(acc[mt][nt][i] + bv[i], 0.0f);
          }
          part += __shfl_xor(part, 16, 64);
          part += __shfl_xor(part, 32, 64);
          if (g == 0) sig[wave * NC + 16 * nt + n] = part;
        }
      }
      __syncthreads();
      if (tid < NS) {
        float sum = 0.0f;
        for (int v = 0; v < V; ++v)
          for (int w = 0; w < kWaves; ++w) sum += sig[w * NC + tid * VP + v];
        const float al = sigmoidf(sum * inv_v + ws[so.sigma_b]);
        if constexpr (POINTS) {
          if (p0 + tid < a.n_points) a.out_occ[(long long)b * a.n_points + p0 + tid] = al;
        } else {
          s_alpha[tid] = al;
          if (a.out_occ) a.out_occ[ray * S + c0 + tid] = al;
        }
      }
      if constexpr (POINTS) continue;  // density query: no compositing (S == NS)
      __syncthreads();
      // ---- front-to-back compositing (volume_rendering.cu:713-734) ---------
      if (tid == 0) {
        for (int sl = 0; sl < NS; ++sl) {
          float w = 0.0f;
          if (!done) {
            const float al = s_alpha[sl];
            w = al * T;
            if (a.compensated) {
              kahan_add(dsum, dcmp, w * s_t[sl]);
              kahan_add(wsum, wcmp, w);
              kahan_add(asum, acmp, w * al);
            } else {
              dsum += w * s_t[sl];
              wsum += w;
              asum += w * al;
            }
            T *= 1.0f - al;
            if (T < a.t_thres) done = true;
          }
          s_w[sl] = w;
        }
        s_flag[0] = done ? 1 : 0;
      }
      __syncthreads();
      // ---- per-view tokens: sum_i w_i f_v(x_i) -----------------------------
#pragma unroll
      for (int j = 0; j < TE; ++j) {
        const int e = tid + j * kThreads;
        if (e < VP * F) {
          const int v = e / F, c = e % F;
          if (a.compensated) {
            for (int sl = 0; sl < NS; ++sl)
              kahan_add(tok[j], tcmp[j], s_w[sl] * keep[(sl * VP + v) * KF + c]);
          } else {
#pragma unroll 4
            for (int sl = 0; sl < NS; ++sl) tok[j] += s_w[sl] * keep[(sl * VP + v) * KF + c];
          }
        }
      }
    }
    __syncthreads();
    if constexpr (!POINTS) {
      // ---- outputs: per-view tokens for the Hydra/RGB kernel, depth, sums --
#pragma unroll
      for (int j = 0; j < TE; ++j) {
        const int e = tid + j * kThreads;
        if (e < VP * F) {
          // quantised maps hold x / s_c: a power-of-two scale restores the
          // token exactly
          const float sc = a.fscale ? a.fscale[e % F] : 1.0f;
          a.out_tok[item * VP * F + e] = (e / F) < V ? (tok[j] - tcmp[j]) * sc : 0.0f;
        }
      }
      if (tid == 0) {
        const float* zd = a.zdir + b * 3;
        a.out_depth[ray] = (dsum - dcmp) * (zd[0] * dx + zd[1] * dy + zd[2] * dz);
        if (a.out_aux) {
          a.out_aux[ray * 2 + 0] = wsum - wcmp;
          a.out_aux[ray * 2 + 1] = asum - acmp;
        }
      }
    }
  }
}

// ---- Hydra encoder (2 stacks) + RGB head on the composited tokens ---------
// (aggregation.py:101-128,164-171,299-307; nerf.py:117-124; infer_mlp.cu
// hydra_attention_kernel :242-445).  A group is 16 rays x 4 view slots = 64
// tokens (token t: ray t/4, view t%4).  The VALU helpers below keep
// feature-major rows [row][token] padded to kLd floats (row reads and the
// transposing token load free of bank conflicts); they serve the backward
// of the F = 64 model.
constexpr int kRaysPerGroup = 16;          // rays per group with 4 view slots
constexpr int kTok = kRaysPerGroup * kVP;  // 64 = one token per lane
constexpr int kLd = kTok + 1;

// rows[o][lane] = act(bias[o] + sum_k W[o][k] in[k]) for this wave's NO
// outputs starting at o0; W rows are wave-uniform so they are fetched once
// per wave and shared by the 64 tokens.
template <int K, int NO, bool RELU, int LD = kLd>
__device__ __forceinline__ void token_dense(const float* __restrict__ W, const float* bias,
                                            const float (&in)[K], int o0, float* rows,
                                            int lane) {
#pragma unroll
  for (int i = 0; i < NO; ++i) {
    const int o = o0 + i;
    const float* wr = W + (size_t)o * K;
    float acc = bias ? bias[o] : 0.0f;
#pragma unroll
    for (int k = 0; k < K; ++k) acc += wr[k] * in[k];
    rows[o * LD + lane] = RELU ? fmaxf(acc, 0.0f) : acc;
  }
}

template <int N, int LD = kLd>
__device__ __forceinline__ void load_rows(const float* rows, int lane, float (&dst)[N]) {
#pragma unroll
  for (int k = 0; k < N; ++k) dst[k] = rows[k * LD + lane];
}

// ---- Hydra encoder + RGB head on MFMA (every form) ------------------------
// Every dense product on MFMA with fp32 accumulation: the 16-bit forms on
// v_mfma_f32_16x16x32_{bf16,f16} (operands rounded to the form's type), the
// fp32 form -- and every render that is differentiated -- on
// v_mfma_f32_16x16x4f32; normalisations, the cross-view sum, residuals,
// biases and activations stay in fp32 VALU.  Each wave owns 16 tokens
// (16 / VP rays): the MFMA's N = 16 columns are the tokens, M = 16 output
// units per tile.  A fragments are read straight from the fp32 small-weight
// buffer (lane l: row 16 mt + l%16), B fragments from the wave's LDS rows
// [token][input]; the result (lane l: units 16 mt + 4 (l/16) + i of token
// l%16) is post-processed in registers and written back to LDS for the next
// product.  The cross-view sum of k^ v is a shuffle over the lanes of one
// ray's tokens.
template <int PREC> struct HydraOp;
template <> struct HydraOp<kBF16> { typedef bf16x8 V; typedef __bf16 E; };
template <> struct HydraOp<kF16> { typedef f16x8 V; typedef _Float16 E; };

template <int PREC>
__device__ __forceinline__ f32x4 hmfma(typename HydraOp<PREC>::V a, typename HydraOp<PREC>::V b,
                                       f32x4 c) {
  if constexpr (PREC == kBF16) return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
  else return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}

// 8 consecutive floats -> one 16-bit operand (round to nearest even)
template <int PREC>
__device__ __forceinline__ typename HydraOp<PREC>::V hcvt8(const float* p) {
  typedef typename HydraOp<PREC>::E E;
  const f32x4 x0 = *reinterpret_cast<const f32x4*>(p);
  const f32x4 x1 = *reinterpret_cast<const f32x4*>(p + 4);
  typename HydraOp<PREC>::V v;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    v[j] = (E)x0[j];
    v[4 + j] = (E)x1[j];
  }
  return v;
}

// 8 consecutive weights as one 16-bit operand: fp32 rows converted here, or
// rows of the 16-bit copy loaded as they are (the same rounding, done once)
template <int PREC, typename WT>
__device__ __forceinline__ typename HydraOp<PREC>::V hload8(const WT* p) {
  if constexpr (std::is_same<WT, float>::value)
    return hcvt8<PREC>(p);
  else
    return *reinterpret_cast<const typename HydraOp<PREC>::V*>(p);
}

// acc[mt] = W[16 mt + l%16][0..K) . src[token l%16][0..K) for MT tiles (acc
// zero on entry); W row-major [rows][K] fp32 (rows >= n_rows read as zero),
// src fp32 LDS rows of pitch LD floats.  fp16 operands: each token's inputs
// are divided by a power of two sigma >= their max |x| before rounding and
// the token's results multiplied by sigma after (exact), so inputs beyond
// fp16's range (65504) cannot overflow; bf16 has fp32's range and needs none.
// fp32 form: lane group g takes inputs [g K/4, (g+1) K/4) (any assignment of
// inputs to (step, group) pairs sums the same products), so a lane's A and
// B values are K/4 contiguous floats: 16-byte loads from the weights and
// from LDS.
template <int MT, int K, int LD>
__device__ __forceinline__ void hydra_mm32(const float* __restrict__ W, int n_rows,
                                           const float* src, int lane, f32x4 (&acc)[MT]) {
  constexpr int KQ = K / 4;
  static_assert(KQ % 4 == 0 && LD % 4 == 0, "16-byte operand runs");
  const int n = lane & 15, g = lane >> 4;
  float b[KQ];
#pragma unroll
  for (int s = 0; s < KQ; s += 4) {
    const f32x4 v = *reinterpret_cast<const f32x4*>(src + n * LD + g * KQ + s);
#pragma unroll
    for (int j = 0; j < 4; ++j) b[s + j] = v[j];
  }
#pragma unroll
  for (int mt = 0; mt < MT; ++mt) {
    const int row = 16 * mt + n;
    float a[KQ];
#pragma unroll
    for (int s = 0; s < KQ; s += 4) {
      const f32x4 v = row < n_rows
                          ? *reinterpret_cast<const f32x4*>(W + (size_t)row * K + g * KQ + s)
                          : f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int j = 0; j < 4; ++j) a[s + j] = v[j];
    }
#pragma unroll
    for (int s = 0; s < KQ; ++s)
      acc[mt] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[s], b[s], acc[mt], 0, 0, 0);
  }
}

template <int PREC, int MT, int K, int LD, typename WT>
__device__ __forceinline__ void hydra_mm(const WT* __restrict__ W, int n_rows,
                                         const float* src, int lane, f32x4 (&acc)[MT]) {
  if constexpr (PREC == kF32) {
    hydra_mm32<MT, K, LD>(W, n_rows, src, lane, acc);
  } else {
  typedef typename HydraOp<PREC>::V V;
  const int n = lane & 15, g = lane >> 4;
  float sigma = 1.0f, inv_sigma = 1.0f;
  if constexpr (PREC == kF16) {
    float m = 0.0f;
#pragma unroll
    for (int s = 0; s < K / 32; ++s)
#pragma unroll
      for (int j = 0; j < 8; ++j) m = fmaxf(m, fabsf(src[n * LD + 32 * s + 8 * g + j]));
    m = fmaxf(m, __shfl_xor(m, 16, 64));
    m = fmaxf(m, __shfl_xor(m, 32, 64));
    if (m > 0.0f && m <= 3.0e38f) {
      int e = 0;
      frexpf(m, &e);
      sigma = ldexpf(1.0f, e);
      inv_sigma = ldexpf(1.0f, -e);
    }
  }
#pragma unroll
  for (int s = 0; s < K / 32; ++s) {
    V b;
    if constexpr (PREC == kF16) {
      float x[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) x[j] = src[n * LD + 32 * s + 8 * g + j] * inv_sigma;
      b = hcvt8<PREC>(x);
    } else {
      b = hcvt8<PREC>(src + n * LD + 32 * s + 8 * g);
    }
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) {
      const int row = 16 * mt + n;
      V a;
      if (row < n_rows) {
        a = hload8<PREC>(W + (size_t)row * K + 32 * s + 8 * g);
      } else {
#pragma unroll
        for (int j = 0; j < 8; ++j) a[j] = (typename HydraOp<PREC>::E)0.0f;
      }
      acc[mt] = hmfma<PREC>(a, b, acc[mt]);
    }
  }
  if constexpr (PREC == kF16) {
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) acc[mt] *= sigma;
  }
  }
}

template <int MT>
__device__ __forceinline__ void hzero(f32x4 (&acc)[MT]) {
#pragma unroll
  for (int mt = 0; mt < MT; ++mt) acc[mt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
}

template <int F, int H, int VP, int PREC>
__global__ void __launch_bounds__(kThreads)
hydra_rgb_mfma_kernel(const float* __restrict__ tok, const int* __restrict__ list,
                      const unsigned* __restrict__ n_valid_ptr, const float* __restrict__ ws,
                      const void* __restrict__ ws16, int V, float* __restrict__ out_rgb) {
  constexpr int RPG = kTok / VP;          // rays per workgroup (64 tokens)
  constexpr int HI = kHydraInner;         // 32
  constexpr int LX = F + 4;               // fp32 row pitch of token rows (16-byte aligned)
  constexpr int LH = 3 * HI + 4;          // q | k | v rows, then attention / hidden / RGB rows
  constexpr SmallOffsets so = small_offsets(F, H);
  constexpr int stack_stride = so.q[1] - so.q[0];
  static_assert(F % 32 == 0 && kRgbH1 % 32 == 0 && kRgbH2 % 32 == 0, "whole K steps");
  static_assert(kRgbH1 <= 3 * HI, "RGB layer 1 fits the q|k|v rows");
  __shared__ __attribute__((aligned(16))) float sx[kWaves][16 * LX];   // tokens (fp32)
  __shared__ __attribute__((aligned(16))) float sh[kWaves][16 * LH];   // products (fp32)

  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int n = lane & 15, g = lane >> 4;
  const long long n_valid = *n_valid_ptr;
  const long long item0 = (long long)blockIdx.x * RPG;
  if (item0 >= n_valid) return;
  float* xs = sx[wave];
  float* hs = sh[wave];
  // this wave's 16 tokens: a contiguous block of 16 x F floats (zero past the list)
  const long long t0 = item0 * VP + 16 * wave;
  const long long n_tok = (n_valid - item0) * VP - 16 * wave;   // tokens of this wave that exist
  for (int i = lane; i < 16 * F / 4; i += 64) {
    const int tk = (4 * i) / F, c = (4 * i) % F;
    f32x4 v = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    if (tk < n_tok) v = *reinterpret_cast<const f32x4*>(tok + (t0 + tk) * F + c);
    *reinterpret_cast<f32x4*>(xs + tk * LX + c) = v;
  }
  __syncthreads();
  const int view = n % VP;
  const float vmask = view < V ? 1.0f : 0.0f;
  // weight matrices: fp32 (converted at each use) or the 16-bit copy, at the
  // same element offsets; biases always from the fp32 buffer
  const auto* wm = [&]() {
    if constexpr (PREC == kF32)
      return ws;
    else
      return static_cast<const typename HydraOp<PREC>::E*>(ws16) + 4;
  }();

  for (int st = 0; st < kHydraStacks; ++st) {
    const auto* wqkv = wm + so.q[0] + st * stack_stride;   // q, k, v rows contiguous [96][F]
    f32x4 a6[6];
    hzero(a6);
    hydra_mm<PREC, 6, F, LX>(wqkv, 3 * HI, xs, lane, a6);
    // normalise q and k over their 32 units: this lane holds units 4g+i of
    // tiles 0,1 (q), 2,3 (k), 4,5 (v); the other 24 units sit in the lanes
    // with the same token and another g
    float qq = 0.0f, kk = 0.0f;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      qq += a6[0][i] * a6[0][i] + a6[1][i] * a6[1][i];
      kk += a6[2][i] * a6[2][i] + a6[3][i] * a6[3][i];
    }
    qq += __shfl_xor(qq, 16, 64);
    qq += __shfl_xor(qq, 32, 64);
    kk += __shfl_xor(kk, 16, 64);
    kk += __shfl_xor(kk, 32, 64);
    const float nq = 1.0f / fmaxf(sqrtf(qq), 1e-12f);
    const float nk = vmask / fmaxf(sqrtf(kk), 1e-12f);
    // attention: q^ * sum_views k^ v, written as rows [token][unit] for fc
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        float kv = a6[2 + h][i] * nk * a6[4 + h][i];
#pragma unroll
        for (int off = 1; off < VP; off <<= 1) kv += __shfl_xor(kv, off, 64);
        hs[n * LH + 16 * h + 4 * g + i] = a6[h][i] * nq * kv;
      }
    __syncthreads();
    // out = fc(attn) + x  (kept in the token rows as the new residual base)
    {
      constexpr int MT = F / 16;
      f32x4 ao[MT];
      hzero(ao);
      hydra_mm<PREC, MT, HI, LH>(wm + so.fc[0] + st * stack_stride, F, hs, lane, ao);
      __syncthreads();
#pragma unroll
      for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int i = 0; i < 4; ++i) xs[n * LX + 16 * mt + 4 * g + i] += ao[mt][i];
    }
    __syncthreads();
    // FFN hidden: relu(W1 out + b1)
    {
      f32x4 ah[2];
      hzero(ah);
      hydra_mm<PREC, 2, F, LX>(wm + so.w1[0] + st * stack_stride, HI, xs, lane, ah);
      const float* b1 = ws + so.b1[0] + st * stack_stride;
#pragma unroll
      for (int mt = 0; mt < 2; ++mt)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int u = 16 * mt + 4 * g + i;
          hs[n * LH + u] = fmaxf(ah[mt][i] + b1[u], 0.0f);
        }
    }
    __syncthreads();
    // x = W2 h + b2 + out
    {
      constexpr int MT = F / 16;
      f32x4 ax[MT];
      hzero(ax);
      hydra_mm<PREC, MT, HI, LH>(wm + so.w2[0] + st * stack_stride, F, hs, lane, ax);
      const float* b2 = ws + so.b2[0] + st * stack_stride;
#pragma unroll
      for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int u = 16 * mt + 4 * g + i;
          xs[n * LX + u] += ax[mt][i] + b2[u];
        }
    }
    __syncthreads();
  }
  // ---- RGB head on the first view's token: F -> 64 -> 32 -> 3 --------------
  {
    f32x4 a1[kRgbH1 / 16];
    hzero(a1);
    hydra_mm<PREC, kRgbH1 / 16, F, LX>(wm + so.r1w, kRgbH1, xs, lane, a1);
#pragma unroll
    for (int mt = 0; mt < kRgbH1 / 16; ++mt)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int u = 16 * mt + 4 * g + i;
        hs[n * LH + u] = fmaxf(a1[mt][i] + ws[so.r1b + u], 0.0f);
      }
  }
  __syncthreads();
  {
    f32x4 a2[kRgbH2 / 16];
    hzero(a2);
    hydra_mm<PREC, kRgbH2 / 16, kRgbH1, LH>(wm + so.r2w, kRgbH2, hs, lane, a2);
    __syncthreads();
#pragma unroll
    for (int mt = 0; mt < kRgbH2 / 16; ++mt)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int u = 16 * mt + 4 * g + i;
        hs[n * LH + u] = fmaxf(a2[mt][i] + ws[so.r2b + u], 0.0f);
      }
  }
  __syncthreads();
  {
    f32x4 a3[1];
    hzero(a3);
    hydra_mm<PREC, 1, kRgbH2, LH>(wm + so.r3w, 3, hs, lane, a3);
    const long long item = item0 + (16 * wave + n) / VP;
    if (g == 0 && view == 0 && item < n_valid) {
#pragma unroll
      for (int j = 0; j < 3; ++j)
        out_rgb[(long long)list[item] * 3 + j] = sigmoidf(a3[0][j] + ws[so.r3b + j]);
    }
  }
}

// ---- Backward of the sample kernel: compositing reverse scan + MLP ---------
// Per ray (list order), with per-ray recompute instead of stored activations:
//  1. the features of all S samples x 4 views are gathered again (LDS);
//  2. g_i = dL/dw_i = sum_v dL/dtoken_v . f_v(x_i) (the tokens are
//     sum_i w_i f_v(x_i));
//  3. the transmittance recurrence of the forward (same order, same early
//     termination) is replayed and walked back: with
//     U_i = sum_{k>i} g_k alpha_k prod_{i<j<k}(1 - alpha_j),
//     dL/dalpha_i = T_i (g_i - U_i) and U_{i-1} = g_i alpha_i + (1 - alpha_i) U_i
//     (no division by 1 - alpha);  dz_i = dL/dalpha_i alpha_i (1 - alpha_i);
//  4. per chunk of NS samples the MLP forward is recomputed keeping every
//     layer's output in LDS, delta_3 = dz / V sigma_w [pre_3 > 0] and
//     delta_{l-1} = W_l^T delta_l [h_{l-1} > 0] run on MFMA with the packed
//     transposed weights;
//  5. the layer inputs A_l and deltas D_l of every column go to HBM for the
//     weight-gradient GEMMs dW_l = D_l^T A_l (sums over all columns), and
//     the occupancy-head gradients go to this workgroup's partial buffer.
// Column c of a batch is (item - item_begin) * S * 4 + sample * 4 + view.
struct BackArgs {
  RenderArgs r;            // forward arguments (t, rays, cameras, maps, weights, ...)
  const void* wdense_t;    // packed W_3^T, W_2^T, W_1^T (layers of input width H)
  const float* occ;        // [n_rays, S] forward occupancies
  const float* dtok;       // [n_valid, 4, F] dL/d token (list order)
  long long item_begin, item_end;
  void* A0;                // [C, H] layer-0 inputs (kernel order, K0 used)
  void* A1;                // [C, H] h0
  void* A2;                // [C, H] h1
  void* A3;                // [C, H] h2
  void* D0;                // [C, H] delta of layer 0 .. 3
  void* D1;
  void* D2;
  void* D3;
  float* gpart;            // [gridDim.x, small_total] partial small-weight gradients
  float* gdense;           // [gridDim.x, dense_grad_total] partial dense gradients (DW form)
};


// dst[col][unit] = acc (times [mask[col][unit] > 0] when a mask buffer is
// given) for this wave's units, fp32 accumulator layout (16-byte stores).
template <int H, int PREC, int NTL>
__device__ __forceinline__ void write_vals(typename Tr<PREC>::E* dst, const f32x4 (&acc)[H / 64][NTL],
                                           const typename Tr<PREC>::E* mask, int wave, int lane) {
  typedef typename Tr<PREC>::E E;
  constexpr int MTW = H / 64, CE = Tr<PREC>::CE, NV = 4 * MTW;
  static_assert(CE == 4, "fp32 layout");
  const int n = lane & 15, g = lane >> 4;
  const int u0 = unit_of<H>(wave, g, 0, 0);
#pragma unroll
  for (int nt = 0; nt < NTL; ++nt) {
    const int col = 16 * nt + n;
#pragma unroll
    for (int c = 0; c < NV / CE; ++c) {
      const int off = act_chunk<H, PREC>(col, u0 / CE + c);
      f32x4 v;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int q = 4 * c + e;
        v[e] = acc[q / 4][nt][q % 4];
      }
      if (mask) {
        const f32x4 m = *reinterpret_cast<const f32x4*>(mask + off);
#pragma unroll
        for (int e = 0; e < 4; ++e) v[e] = m[e] > 0.0f ? v[e] : 0.0f;
      }
      *reinterpret_cast<f32x4*>(dst + off) = v;
    }
  }
}

// The same for the 32x32 accumulator layout of the 16-bit forms.
template <int H, int PREC, int NT>
__device__ __forceinline__ void write_vals16(typename Tr<PREC>::E* dst,
                                             const f32x16 (&acc)[H / 128][NT],
                                             const typename Tr<PREC>::E* mask, int wave,
                                             int lane) {
  typedef typename Tr<PREC>::E E;
  typedef typename Tr<PREC>::V8 V8;
  const int c = lane & 31, h = lane >> 5;
#pragma unroll
  for (int mt = 0; mt < H / 128; ++mt) {
    const int u0 = unit_of32<H>(wave, mt, h, 0);
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
#pragma unroll
      for (int e = 0; e < 2; ++e) {
        const int off = act_chunk<H, PREC>(32 * nt + c, u0 / 8 + e);
        V8 v;
        V8 m;
        if (mask) m = *reinterpret_cast<const V8*>(mask + off);
#pragma unroll
        for (int k = 0; k < 8; ++k) {
          const float x = acc[mt][nt][8 * e + k];
          v[k] = (E)((!mask || (float)m[k] > 0.0f) ? x : 0.0f);
        }
        *reinterpret_cast<V8*>(dst + off) = v;
      }
    }
  }
}

constexpr int kBackMaxS = 64;

// SMAX: samples per ray the buffers hold
template <int F, int H, int PREC, int NS, int SMAX = kBackMaxS>
struct SmemB {
  typedef typename Tr<PREC>::E E;
  static constexpr int NC = NS * kVP;
  static constexpr size_t buf = (size_t)NC * H * sizeof(E);   // one activation buffer
  static constexpr size_t x0 = 0;
  static constexpr size_t h0 = x0 + buf, h1 = h0 + buf, h2 = h1 + buf, dl = h2 + buf;
  static constexpr int KF = F + 4;
  static constexpr size_t keep = dl + buf;                         // f32 [SMAX][4][KF]
  static constexpr size_t alpha = keep + (size_t)SMAX * kVP * KF * 4;
  static constexpr size_t gw = alpha + SMAX * 4;                  // dL/dw
  static constexpr size_t dz = gw + SMAX * 4;                     // dL/d pre-sigmoid
  static constexpr size_t pos = dz + SMAX * 4;                    // f32 [SMAX][4]
  static constexpr size_t bias = pos + SMAX * 16;                 // f32 [5][H]
  static constexpr size_t total = bias + (size_t)5 * H * 4;
};

// D = W_layer x in[col] for one layer of input width K on MFMA, fp32 form
// (NTL column tiles of 16).  Wl: the layer's packed fragments for all waves.
template <int H, int PREC, int NTL, int K>
__device__ __forceinline__ void one_layer(const typename Tr<PREC>::E* Wl,
                                          const typename Tr<PREC>::E* in, int wave, int lane,
                                          f32x4 (&acc)[H / 64][NTL]) {
  typedef Chain<H, PREC, NTL> C;
  constexpr int UL = K / C::KPU, NB = C::NB, PD = C::PD;
  static_assert(K % C::KPU == 0, "layer width must fill whole units");
  const int n = lane & 15, g = lane >> 4;
  const char* wb = reinterpret_cast<const char*>(Wl) + (size_t)wave * UL * C::unit_bytes;
  asm volatile("" : "+s"(wb));
  const unsigned lane_off = (unsigned)lane * C::MTW * sizeof(typename C::E);
  typename C::V abuf[NB][C::AN];
  typename C::V bbuf[2][NTL];
#pragma unroll
  for (int mt = 0; mt < C::MTW; ++mt)
#pragma unroll
    for (int nt = 0; nt < NTL; ++nt) acc[mt][nt] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int j = 0; j < PD; ++j) load_a<H, PREC, NTL>(wb, j, lane_off, abuf[j % NB]);
  read_b<H, PREC, NTL>(in, 0, n, g, bbuf[0]);
  run_layer<H, PREC, NTL, UL, 0>(wb, wb, lane_off, in, n, g, abuf, bbuf, acc);
}

// The same for the 16-bit forms (32x32x16 MFMA, NT column tiles of 32).
template <int H, int PREC, int NT, int K>
__device__ __forceinline__ void one_layer16(const typename Tr<PREC>::E* Wl,
                                            const typename Tr<PREC>::E* in, int wave, int lane,
                                            f32x16 (&acc)[H / 128][NT]) {
  typedef Chain16<H, PREC, NT> C;
  constexpr int UL = K / 16, NB = C::NB, PD = C::PD;
  static_assert(K % 16 == 0, "layer width must fill whole k-steps");
  const char* wb = reinterpret_cast<const char*>(Wl) + (size_t)wave * UL * C::unit_bytes;
  asm volatile("" : "+s"(wb));
  const unsigned lane_off = (unsigned)lane * C::M32 * 16;
  typename C::V abuf[NB][C::M32];
  typename C::V bbuf[2][NT];
#pragma unroll
  for (int mt = 0; mt < C::M32; ++mt)
#pragma unroll
    for (int nt = 0; nt < NT; ++nt)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[mt][nt][r] = 0.0f;
#pragma unroll
  for (int j = 0; j < PD; ++j) load_a16<H, PREC, NT>(wb, j, lane_off, abuf[j % NB]);
  read_b16<H, PREC, NT>(in, 0, lane, bbuf[0]);
  run_layer16<H, PREC, NT, UL, 0>(wb, wb, lane_off, in, lane, abuf, bbuf, acc);
}

// Copies NC columns x K elements of a swizzled LDS buffer to rows of a
// [C, H]-strided matrix in global memory (16-byte chunks).
template <int H, int PREC, int NC, int K>
__device__ __forceinline__ void store_cols(const typename Tr<PREC>::E* buf, void* dst,
                                           long long col0, int tid) {
  typedef typename Tr<PREC>::E E;
  constexpr int CE = Tr<PREC>::CE, NCH = K / CE;
  E* out = static_cast<E*>(dst);
  for (int it = tid; it < NC * NCH; it += kThreads) {
    const int col = it / NCH, ch = it % NCH;
    *reinterpret_cast<f32x4*>(out + (col0 + col) * H + ch * CE) =
        *reinterpret_cast<const f32x4*>(buf + act_chunk<H, PREC>(col, ch));
  }
}

// ---- Weight gradients accumulated inside the backward kernel ---------------
// dW_l[u][k] = sum_c D_l[c][u] A_l[c][k] over the columns c a workgroup
// processes, db_l[u] = sum_c D_l[c][u].  Each wave owns 32 output units (H =
// 128) and keeps its rows of all four dW_l in MFMA accumulators for the whole
// kernel (240 registers per lane); per chunk, both operands are read from the
// LDS activation layout with the column index as the MFMA reduction index.
// Layout of a dense-gradient buffer (floats): dW_0 [H][F + 64] (kernel input
// order), dW_1..3 [H][H], db [4][H].
__host__ __device__ constexpr int dense_grad_total(int F, int H) {
  return H * (F + kPeSlots) + 3 * H * H + 4 * H;
}
__host__ __device__ constexpr int dense_grad_offset(int F, int H, int l) {
  return l == 0 ? 0 : H * (F + kPeSlots) + (l - 1) * H * H;
}

// 8 consecutive columns c0 + 8(l/32) + 0..7 of feature f0 + l%32 of a 16-bit
// activation buffer, as one lane's MFMA operand, with two transposing LDS
// reads (ds_read_b64_tr_b16).  That instruction works on groups of 16 lanes:
// lane L of a group supplies 4 consecutive 16-bit values, and lane i of the
// group receives element i%4 of the values supplied by lanes 4j + i/4,
// j = 0..3.  So lane L = 4j + q points at features f + 4q .. f + 4q + 3 of
// column c + j, and lane i receives feature f + i of columns c .. c + 3.
template <int H, int PREC>
__device__ __forceinline__ typename Tr<PREC>::V8 col_operand(const typename Tr<PREC>::E* buf,
                                                             int c0, int f0, int lane) {
  typedef short s4 __attribute__((ext_vector_type(4)));
  typedef short s8 __attribute__((ext_vector_type(8)));
  typedef __attribute__((address_space(3))) s4 lds_s4;
  const int L = lane & 15, g = lane >> 4;
  const int f = f0 + 16 * (g & 1) + 4 * (L & 3);
  const int c = c0 + 8 * (g >> 1) + (L >> 2);
  const typename Tr<PREC>::E* p0 = buf + act_chunk<H, PREC>(c, f >> 3) + (f & 7);
  const typename Tr<PREC>::E* p1 = buf + act_chunk<H, PREC>(c + 4, f >> 3) + (f & 7);
  const s4 lo = __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds_s4*)(p0));
  const s4 hi = __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds_s4*)(p1));
  const s8 v = __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7);
  return __builtin_bit_cast(typename Tr<PREC>::V8, v);
}

// 16-bit forms, v_mfma_f32_32x32x16: A operand = D^T (lane: unit 32w + l%32,
// columns c0 + 8(l/32) + j), B operand = A_l (columns c0 + 8(l/32) + j,
// input 32n + l%32); accumulator r of tile n holds
// dW[32w + 8(r/4) + 4(l/32) + r%4][32n + l%32].
template <int H, int PREC, int NC, int K>
__device__ __forceinline__ void accum_dw16(const typename Tr<PREC>::E* D,
                                           const typename Tr<PREC>::E* A,
                                           f32x16 (&dw)[K / 32], float& db, int wave, int lane) {
  typedef typename Tr<PREC>::V8 V8;
  static_assert(H == 128 && NC % 16 == 0 && K % 32 == 0, "32 units per wave, whole tiles");
#pragma unroll
  for (int s = 0; s < NC / 16; ++s) {
    const V8 a = col_operand<H, PREC>(D, 16 * s, 32 * wave, lane);
#pragma unroll
    for (int j = 0; j < 8; ++j) db += (float)a[j];
#pragma unroll
    for (int n = 0; n < K / 32; ++n) {
      const V8 b = col_operand<H, PREC>(A, 16 * s, 32 * n, lane);
      if constexpr (PREC == kBF16)
        dw[n] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, dw[n], 0, 0, 0);
      else
        dw[n] = __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, dw[n], 0, 0, 0);
    }
  }
}

// fp32 form, v_mfma_f32_16x16x4f32: two 16-unit tiles per wave; A operand =
// D[c0 + l/16][32w + 16mt + l%16], B operand = A_l[c0 + l/16][16n + l%16];
// accumulator i of tile (mt, n) holds dW[32w + 16mt + 4(l/16) + i][16n + l%16].
template <int H, int NC, int K>
__device__ __forceinline__ void accum_dw32(const float* D, const float* A,
                                           f32x4 (&dw)[2][K / 16], float (&db)[2], int wave,
                                           int lane) {
  static_assert(H == 128 && NC % 4 == 0 && K % 16 == 0, "32 units per wave, whole tiles");
  const int r = lane & 15, q = lane >> 4;
#pragma unroll
  for (int s = 0; s < NC / 4; ++s) {
    const int c = 4 * s + q;
    float a[2];
#pragma unroll
    for (int mt = 0; mt < 2; ++mt) {
      const int u = 32 * wave + 16 * mt + r;
      a[mt] = D[act_chunk<H, kF32>(c, u / 4) + u % 4];
      db[mt] += a[mt];
    }
#pragma unroll
    for (int n = 0; n < K / 16; ++n) {
      const int k = 16 * n + r;
      const float b = A[act_chunk<H, kF32>(c, k / 4) + k % 4];
#pragma unroll
      for (int mt = 0; mt < 2; ++mt)
        dw[mt][n] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[mt], b, dw[mt][n], 0, 0, 0);
    }
  }
}

// Per-wave register accumulators of the four layers' weight gradients.
template <int F, int H, int PREC> struct DenseAcc;
template <int F, int H> struct DenseAcc<F, H, kF32> {
  static constexpr int K0 = F + kPeSlots;
  f32x4 w0[2][K0 / 16], w1[2][H / 16], w2[2][H / 16], w3[2][H / 16];
  float b[4][2];
  __device__ __forceinline__ void zero() {
#pragma unroll
    for (int mt = 0; mt < 2; ++mt) {
#pragma unroll
      for (int n = 0; n < K0 / 16; ++n) w0[mt][n] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int n = 0; n < H / 16; ++n) {
        w1[mt][n] = f32x4{0.f, 0.f, 0.f, 0.f};
        w2[mt][n] = f32x4{0.f, 0.f, 0.f, 0.f};
        w3[mt][n] = f32x4{0.f, 0.f, 0.f, 0.f};
      }
#pragma unroll
      for (int l = 0; l < 4; ++l) b[l][mt] = 0.0f;
    }
  }
  template <int NTW>
  __device__ __forceinline__ void store_tiles(float* dst, const f32x4 (&w)[2][NTW], int K,
                                              int wave, int lane) {
    const int r = lane & 15, q = lane >> 4;
#pragma unroll
    for (int mt = 0; mt < 2; ++mt)
#pragma unroll
      for (int n = 0; n < NTW; ++n)
#pragma unroll
        for (int i = 0; i < 4; ++i)
          dst[(size_t)(32 * wave + 16 * mt + 4 * q + i) * K + 16 * n + r] = w[mt][n][i];
  }
  // writes this wave's rows of the workgroup's partial dense gradient
  __device__ __forceinline__ void store(float* gd, int wave, int lane) {
    store_tiles(gd + dense_grad_offset(F, H, 0), w0, K0, wave, lane);
    store_tiles(gd + dense_grad_offset(F, H, 1), w1, H, wave, lane);
    store_tiles(gd + dense_grad_offset(F, H, 2), w2, H, wave, lane);
    store_tiles(gd + dense_grad_offset(F, H, 3), w3, H, wave, lane);
    const int r = lane & 15, q = lane >> 4;
    float* gb = gd + dense_grad_offset(F, H, 4);
#pragma unroll
    for (int l = 0; l < 4; ++l)
#pragma unroll
      for (int mt = 0; mt < 2; ++mt) {
        float v = b[l][mt];
        v += __shfl_xor(v, 16, 64);
        v += __shfl_xor(v, 32, 64);
        if (q == 0) gb[l * H + 32 * wave + 16 * mt + r] = v;
      }
  }
};
template <int F, int H, int PREC> struct DenseAcc16 {
  static constexpr int K0 = F + kPeSlots;
  f32x16 w0[K0 / 32], w1[H / 32], w2[H / 32], w3[H / 32];
  float b[4];
  __device__ __forceinline__ void zero() {
#pragma unroll
    for (int i = 0; i < 16; ++i) {
#pragma unroll
      for (int n = 0; n < K0 / 32; ++n) w0[n][i] = 0.0f;
#pragma unroll
      for (int n = 0; n < H / 32; ++n) {
        w1[n][i] = 0.0f;
        w2[n][i] = 0.0f;
        w3[n][i] = 0.0f;
      }
    }
#pragma unroll
    for (int l = 0; l < 4; ++l) b[l] = 0.0f;
  }
  template <int NTW>
  __device__ __forceinline__ void store_tiles(float* dst, const f32x16 (&w)[NTW], int K,
                                              int wave, int lane) {
    const int r = lane & 31, half = lane >> 5;
#pragma unroll
    for (int n = 0; n < NTW; ++n)
#pragma unroll
      for (int i = 0; i < 16; ++i)
        dst[(size_t)(32 * wave + 8 * (i / 4) + 4 * half + i % 4) * K + 32 * n + r] = w[n][i];
  }
  __device__ __forceinline__ void store(float* gd, int wave, int lane) {
    store_tiles(gd + dense_grad_offset(F, H, 0), w0, K0, wave, lane);
    store_tiles(gd + dense_grad_offset(F, H, 1), w1, H, wave, lane);
    store_tiles(gd + dense_grad_offset(F, H, 2), w2, H, wave, lane);
    store_tiles(gd + dense_grad_offset(F, H, 3), w3, H, wave, lane);
    const int r = lane & 31, half = lane >> 5;
    float* gb = gd + dense_grad_offset(F, H, 4);
#pragma unroll
    for (int l = 0; l < 4; ++l) {
      const float v = b[l] + __shfl_xor(b[l], 32, 64);
      if (half == 0) gb[l * H + 32 * wave + r] = v;
    }
  }
};
template <int F, int H> struct DenseAcc<F, H, kBF16> : DenseAcc16<F, H, kBF16> {};
template <int F, int H> struct DenseAcc<F, H, kF16> : DenseAcc16<F, H, kF16> {};

// DW: weight gradients accumulated in registers (H = 128 forms) instead of
// column matrices written for separate GEMMs.
template <int F, int H, int PREC, int NS, bool DW, int SMAX>
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(1, 1)))
render_backward_kernel(BackArgs ba) {
  typedef typename Tr<PREC>::E E;
  typedef SmemB<F, H, PREC, NS, SMAX> L;
  const RenderArgs& a = ba.r;
  constexpr int NC = L::NC, KF = L::KF, MTW = H / 64, CE = Tr<PREC>::CE;
  constexpr int CPL = F / 4, K0 = F + kPeSlots;
  constexpr SmallOffsets so = small_offsets(F, H);
  constexpr bool F32 = PREC == kF32;
  constexpr int NTL = NC / 16, NT32 = NC / 32;
  static_assert(F32 || NC % 32 == 0, "16-bit forms need chunks of a multiple of 8 samples");
  extern __shared__ __attribute__((aligned(16))) char smem[];
  E* x0 = reinterpret_cast<E*>(smem + L::x0);
  auto hb = [&](int l) { return reinterpret_cast<E*>(smem + L::h0 + l * L::buf); };
  E* dl = reinterpret_cast<E*>(smem + L::dl);
  float* keep = reinterpret_cast<float*>(smem + L::keep);
  float* s_alpha = reinterpret_cast<float*>(smem + L::alpha);
  float* s_g = reinterpret_cast<float*>(smem + L::gw);
  float* s_dz = reinterpret_cast<float*>(smem + L::dz);
  float* s_pos = reinterpret_cast<float*>(smem + L::pos);
  float* s_bias = reinterpret_cast<float*>(smem + L::bias);

  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int S = a.S, V = a.V;
  const float* ws = a.wsmall;
  const E* W = reinterpret_cast<const E*>(a.wdense);
  const E* WT = reinterpret_cast<const E*>(ba.wdense_t);
  // packed layer sizes (elements, all waves): layer 0 has K0 inputs, the
  // others H; the transposed layers all have H inputs
  const size_t l0e = (size_t)H * K0, lhe = (size_t)H * H;
  for (int i = tid; i < 4 * H; i += kThreads) s_bias[i] = a.bdense[i];
  for (int i = tid; i < H; i += kThreads) s_bias[4 * H + i] = ws[so.sigma_w + i];
  const float inv_v = 1.0f / (float)V;
  // occupancy-head gradient partials held per lane: the units this lane's
  // accumulators cover, summed over its columns
  constexpr int NU = F32 ? 4 * MTW : 16 * (H / 128);
  float gsw[NU];
#pragma unroll
  for (int u = 0; u < NU; ++u) gsw[u] = 0.0f;
  double gsb = 0.0;   // occupancy-bias gradient: a sum of many cancelling terms
  [[maybe_unused]] DenseAcc<F, H, PREC> dacc;
  if constexpr (DW) dacc.zero();

  const long long item_end = std::min<long long>(ba.item_end, (long long)*a.n_valid);
  for (long long item = ba.item_begin + blockIdx.x; item < item_end; item += gridDim.x) {
    __syncthreads();
    const long long ray = a.list[item];
    const int b = (int)(ray / a.rays_per_batch);
    const float* trow = a.t + ray * S;
    const float* rr = a.rays + ray * 6;
    const float* bb = a.bbox + b * 6;
    const float cx = 0.5f * (bb[0] + bb[1]), cy = 0.5f * (bb[2] + bb[3]),
                cz = 0.5f * (bb[4] + bb[5]);
    const float ihx = 2.0f / (bb[1] - bb[0]), ihy = 2.0f / (bb[3] - bb[2]),
                ihz = 2.0f / (bb[5] - bb[4]);
    if (tid < S) {
      const float tv = trow[tid];
      s_pos[4 * tid + 0] = rr[0] + tv * rr[3];
      s_pos[4 * tid + 1] = rr[1] + tv * rr[4];
      s_pos[4 * tid + 2] = rr[2] + tv * rr[5];
      s_alpha[tid] = ba.occ[ray * S + tid];
    }
    __syncthreads();
    // 1. features of all samples and views, and on the way
    // 2. g_i = dL/dtoken . f(x_i): item `it` is (sample it / 16, view it / 4 %
    // 4, channel quarter it % 4), so a sample's 16 items are 16 consecutive
    // lanes; each dots its quarter of the channels with dL/dtoken, and four
    // shuffles add the quarters and the views
    for (int it = tid; it < S * kVP * 4; it += kThreads) {
      const int col = it >> 2, quarter = it & 3;
      const int sl = col >> 2, v = col & 3;
      float f[CPL];
      gather_quarter<F, PREC, false>(a, b, v, s_pos[4 * sl], s_pos[4 * sl + 1],
                                     s_pos[4 * sl + 2], quarter, f);
      float* kp = keep + col * KF + quarter * CPL;
#pragma unroll
      for (int c = 0; c < CPL; c += 4)
        *reinterpret_cast<f32x4*>(kp + c) = f32x4{f[c], f[c + 1], f[c + 2], f[c + 3]};
      const float* dtv = ba.dtok + item * kVP * F + v * F + quarter * CPL;
      float part = 0.0f;
#pragma unroll
      for (int c = 0; c < CPL; c += 4) {
        const f32x4 d4 = *reinterpret_cast<const f32x4*>(dtv + c);
        part += d4[0] * f[c] + d4[1] * f[c + 1] + d4[2] * f[c + 2] + d4[3] * f[c + 3];
      }
      part += __shfl_xor(part, 1, 64);
      part += __shfl_xor(part, 2, 64);
      part += __shfl_xor(part, 4, 64);
      part += __shfl_xor(part, 8, 64);
      if ((it & 15) == 0) s_g[sl] = part;
    }
    __syncthreads();
    // layer-0 input of the chunk at sample c0: features (from keep) and
    // positional encoding, by threads t0, t0 + tstep, ... (building the
    // first chunk's on waves 1-3 during the replay measured no faster: the
    // second copy of this code spills registers)
    auto build_input = [&](int c0, int t0, int tstep) {
      for (int it = t0; it < NC * (F / CE); it += tstep) {
        const int col = it % NC, ch = it / NC;   // consecutive lanes: consecutive columns
        const float* kp = keep + ((c0 + (col >> 2)) * kVP + (col & 3)) * KF + ch * CE;
        E* d = x0 + act_chunk<H, PREC>(col, ch);
        if constexpr (F32) {
          *reinterpret_cast<f32x4*>(d) = *reinterpret_cast<const f32x4*>(kp);
        } else {
          typename Tr<PREC>::V8 h;
#pragma unroll
          for (int e = 0; e < 8; ++e) h[e] = (E)kp[e];
          *reinterpret_cast<typename Tr<PREC>::V8*>(d) = h;
        }
      }
      for (int it = t0; it < NS * kPeSlots / 2; it += tstep) {
        const int sl = it / (kPeSlots / 2), q0 = 2 * (it % (kPeSlots / 2));
        const float* ps = s_pos + 4 * (c0 + sl);
        float val[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int q = q0 + h;
          val[h] = 0.0f;
          if (q < 63) {
            const int d = q % 3;
            const float xd = d == 0 ? (ps[0] - cx) * ihx
                                    : (d == 1 ? (ps[1] - cy) * ihy : (ps[2] - cz) * ihz);
            if (q < 3) {
              val[h] = xd;
            } else {
              const int col = q / 3 - 1;
              val[h] = sinf(ldexpf(xd, col >> 1) + (float)(col & 1) * kHalfPi);
            }
          }
        }
#pragma unroll
        for (int v = 0; v < kVP; ++v) {
          const int col = sl * kVP + v, k = F + q0;
          E* d2 = x0 + act_chunk<H, PREC>(col, k / CE) + k % CE;
          if constexpr (F32) {
            *reinterpret_cast<f32x2*>(d2) = f32x2{val[0], val[1]};
          } else {
            typedef E E2 __attribute__((ext_vector_type(2)));
            *reinterpret_cast<E2*>(d2) = E2{(E)val[0], (E)val[1]};
          }
        }
      }
    };
    // 3. compositing replayed and walked back by wave 0 from registers,
    // sample i in lane i: the recurrences run in sample order on values
    // broadcast with readlane, the same operations in the same order as the
    // forward's (so the early termination is the forward's)
    if (wave == 0) {
      const float al = lane < S ? s_alpha[lane] : 0.0f;
      const float gl = lane < S ? s_g[lane] : 0.0f;
      float T = 1.0f, myT = 0.0f;
      int last = S - 1;
      for (int i = 0; i < S; ++i) {
        const float ai = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(al), i));
        if (lane == i) myT = T;
        T *= 1.0f - ai;
        if (__builtin_amdgcn_readfirstlane((int)(T < a.t_thres))) {
          last = i;
          break;
        }
      }
      float U = 0.0f, myd = 0.0f;
      for (int i = last; i >= 0; --i) {
        const float ai = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(al), i));
        const float gi = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(gl), i));
        const float ti = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(myT), i));
        const float d = ti * (gi - U) * ai * (1.0f - ai);
        if (lane == i) myd = d;
        U = gi * ai + (1.0f - ai) * U;
      }
      if (lane < S) s_dz[lane] = myd;
      double dsum = myd;
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) dsum += __shfl_xor(dsum, off, 64);
      if (lane == 0) gsb += dsum;
    }
    __syncthreads();
    // 4./5. per chunk
    for (int c0 = 0; c0 < S; c0 += NS) {
      const long long col0 = (item - ba.item_begin) * S * kVP + (long long)c0 * kVP;
      build_input(c0, tid, kThreads);
      __syncthreads();
      if constexpr (!DW) store_cols<H, PREC, NC, K0>(x0, ba.A0, col0, tid);
      if constexpr (F32) {
        f32x4 acc[MTW][NTL];
        // forward, keeping every layer's output
        one_layer<H, PREC, NTL, K0>(W, x0, wave, lane, acc);
        write_act<H, PREC, NTL>(hb(0), s_bias, wave, lane, acc);
        __syncthreads();
        one_layer<H, PREC, NTL, H>(W + l0e, hb(0), wave, lane, acc);
        write_act<H, PREC, NTL>(hb(1), s_bias + H, wave, lane, acc);
        __syncthreads();
        one_layer<H, PREC, NTL, H>(W + l0e + lhe, hb(1), wave, lane, acc);
        write_act<H, PREC, NTL>(hb(2), s_bias + 2 * H, wave, lane, acc);
        __syncthreads();
        one_layer<H, PREC, NTL, H>(W + l0e + 2 * lhe, hb(2), wave, lane, acc);
        // delta_3 = dz / V sigma_w [pre_3 > 0] (real views); sigma_w gradient
        {
          const int n = lane & 15, g = lane >> 4;
#pragma unroll
          for (int nt = 0; nt < NTL; ++nt) {
            const int col = 16 * nt + n;
            const bool real = (col & 3) < V;
            const float dzc = real ? s_dz[c0 + (col >> 2)] * inv_v : 0.0f;
#pragma unroll
            for (int mt = 0; mt < MTW; ++mt)
#pragma unroll
              for (int i = 0; i < 4; ++i) {
                const int u = unit_of<H>(wave, g, mt, i);
                const float pre = acc[mt][nt][i] + s_bias[3 * H + u];
                gsw[4 * mt + i] += dzc * fmaxf(pre, 0.0f);
                acc[mt][nt][i] = pre > 0.0f ? dzc * s_bias[4 * H + u] : 0.0f;
              }
          }
        }
        write_vals<H, PREC, NTL>(dl, acc, nullptr, wave, lane);
        __syncthreads();
        if constexpr (DW) {
          accum_dw32<H, NC, H>(dl, hb(2), dacc.w3, dacc.b[3], wave, lane);
        } else {
          store_cols<H, PREC, NC, H>(dl, ba.D3, col0, tid);
          store_cols<H, PREC, NC, H>(hb(0), ba.A1, col0, tid);
          store_cols<H, PREC, NC, H>(hb(1), ba.A2, col0, tid);
          store_cols<H, PREC, NC, H>(hb(2), ba.A3, col0, tid);
        }
        void* Ds[3] = {ba.D2, ba.D1, ba.D0};
#pragma unroll
        for (int l = 0; l < 3; ++l) {   // delta_{2-l} = W_{3-l}^T delta [h_{2-l} > 0]
          one_layer<H, PREC, NTL, H>(WT + l * lhe, dl, wave, lane, acc);
          __syncthreads();
          write_vals<H, PREC, NTL>(dl, acc, hb(2 - l), wave, lane);
          __syncthreads();
          if constexpr (DW) {
            if (l == 0) accum_dw32<H, NC, H>(dl, hb(1), dacc.w2, dacc.b[2], wave, lane);
            else if (l == 1) accum_dw32<H, NC, H>(dl, hb(0), dacc.w1, dacc.b[1], wave, lane);
            else accum_dw32<H, NC, K0>(dl, x0, dacc.w0, dacc.b[0], wave, lane);
          } else {
            store_cols<H, PREC, NC, H>(dl, Ds[l], col0, tid);
          }
        }
      } else {
        f32x16 acc[H / 128][NT32];
        one_layer16<H, PREC, NT32, K0>(W, x0, wave, lane, acc);
        write_act16<H, PREC, NT32>(hb(0), s_bias, wave, lane, acc);
        __syncthreads();
        one_layer16<H, PREC, NT32, H>(W + l0e, hb(0), wave, lane, acc);
        write_act16<H, PREC, NT32>(hb(1), s_bias + H, wave, lane, acc);
        __syncthreads();
        one_layer16<H, PREC, NT32, H>(W + l0e + lhe, hb(1), wave, lane, acc);
        write_act16<H, PREC, NT32>(hb(2), s_bias + 2 * H, wave, lane, acc);
        __syncthreads();
        one_layer16<H, PREC, NT32, H>(W + l0e + 2 * lhe, hb(2), wave, lane, acc);
        {
          const int c = lane & 31, h = lane >> 5;
#pragma unroll
          for (int nt = 0; nt < NT32; ++nt) {
            const int col = 32 * nt + c;
            const bool real = (col & 3) < V;
            const float dzc = real ? s_dz[c0 + (col >> 2)] * inv_v : 0.0f;
#pragma unroll
            for (int mt = 0; mt < H / 128; ++mt)
#pragma unroll
              for (int r = 0; r < 16; ++r) {
                const int u = unit_of32<H>(wave, mt, h, r);
                const float pre = acc[mt][nt][r] + s_bias[3 * H + u];
                gsw[16 * mt + r] += dzc * fmaxf(pre, 0.0f);
                acc[mt][nt][r] = pre > 0.0f ? dzc * s_bias[4 * H + u] : 0.0f;
              }
          }
        }
        write_vals16<H, PREC, NT32>(dl, acc, nullptr, wave, lane);
        __syncthreads();
        if constexpr (DW) {
          accum_dw16<H, PREC, NC, H>(dl, hb(2), dacc.w3, dacc.b[3], wave, lane);
        } else {
          store_cols<H, PREC, NC, H>(dl, ba.D3, col0, tid);
          store_cols<H, PREC, NC, H>(hb(0), ba.A1, col0, tid);
          store_cols<H, PREC, NC, H>(hb(1), ba.A2, col0, tid);
          store_cols<H, PREC, NC, H>(hb(2), ba.A3, col0, tid);
        }
        void* Ds[3] = {ba.D2, ba.D1, ba.D0};
#pragma unroll
        for (int l = 0; l < 3; ++l) {
          one_layer16<H, PREC, NT32, H>(WT + l * lhe, dl, wave, lane, acc);
          __syncthreads();
          write_vals16<H, PREC, NT32>(dl, acc, hb(2 - l), wave, lane);
          __syncthreads();
          if constexpr (DW) {
            if (l == 0) accum_dw16<H, PREC, NC, H>(dl, hb(1), dacc.w2, dacc.b[2], wave, lane);
            else if (l == 1) accum_dw16<H, PREC, NC, H>(dl, hb(0), dacc.w1, dacc.b[1], wave, lane);
            else accum_dw16<H, PREC, NC, K0>(dl, x0, dacc.w0, dacc.b[0], wave, lane);
          } else {
            store_cols<H, PREC, NC, H>(dl, Ds[l], col0, tid);
          }
        }
      }
      __syncthreads();
    }
  }
  // occupancy-head gradient partials of this workgroup
  float* gp = ba.gpart + (size_t)blockIdx.x * so.total;
  if constexpr (F32) {
    const int n = lane & 15, g = lane >> 4;
#pragma unroll
    for (int k = 0; k < NU; ++k) {
      float v = gsw[k];
#pragma unroll
      for (int off = 1; off < 16; off <<= 1) v += __shfl_xor(v, off, 64);
      if (n == 0) gp[so.sigma_w + unit_of<H>(wave, g, k / 4, k % 4)] += v;
    }
  } else {
    const int c = lane & 31, h = lane >> 5;
#pragma unroll
    for (int k = 0; k < NU; ++k) {
      float v = gsw[k];
#pragma unroll
      for (int off = 1; off < 32; off <<= 1) v += __shfl_xor(v, off, 64);
      if (c == 0) gp[so.sigma_w + unit_of32<H>(wave, k / 16, h, k % 16)] += v;
    }
  }
  if (tid == 0) gp[so.sigma_b] += (float)gsb;
  if constexpr (DW)
    dacc.store(ba.gdense + (size_t)blockIdx.x * dense_grad_total(F, H), wave, lane);
}

// ---- Backward of the Hydra encoder + RGB head ------------------------------
// For each group of 16 rays (list order) the forward is recomputed from the
// stored tokens (stack inputs kept, stack internals recomputed during the
// backward walk), then dL/d rgb is propagated back to dL/d token.  Weight
// gradients are sums over the 64 tokens of outer products, computed from LDS
// rows with one output entry per thread, and accumulated in this workgroup's
// own partial buffer in global memory (no atomics); a second kernel adds the
// partials of all workgroups.
constexpr int kLdB = 68;  // 16-byte aligned rows of 64 tokens

// Row blocks are shared by values that are never live together, so that two
// workgroups fit in a CU's LDS (F = 32: 288 rows of 272 B):
//  * X0 / X2 / DHID: the input tokens (stack 0's input, dead once stack 0
//    has run), then stack 1's output (the final tokens, dead after the RGB
//    head's backward), then dL/d hidden and d attention while the stacks are
//    walked back;
//  * X1: stack 1's input, then the input tokens again (reloaded) for stack
//    0's backward;
//  * HID also holds the RGB head's dz3 [3] (the hidden rows are recomputed
//    before the stacks need them).
template <int F>
struct HydraBwdRows {
  static_assert(F >= kHydraInner, "DHID shares the X0 rows");
  static_assert(kHydraStacks == 2, "X0 / X1 reuse assumes two stacks");
  static constexpr int X0 = 0, X1 = F, X2 = 0;
  static constexpr int QKV = 2 * F;      // q, k, v then dq, dk, dv [96]; RGB r1 [64] + r2 [32]
  static constexpr int OUT = QKV + 3 * kHydraInner;   // attention output [F]
  static constexpr int HID = OUT + F;    // FFN hidden, then attention values [32]
  static constexpr int DZ3 = HID;        // dL/dz of the RGB output [3]
  static constexpr int DX = HID + kHydraInner;        // dL/d stack output, then input [F]
  static constexpr int DOUT = DX + F;                 // [F]
  static constexpr int DHID = X2;                     // dL/d hidden, then d attention [32]
  static constexpr int ROWS = DOUT + F;
  static constexpr size_t bytes = (size_t)ROWS * kLdB * sizeof(float);
};

// gp[o*K + c] += sum_t d[o][t] in[c][t];  gb[o] += sum_t d[o][t]
template <int O, int K>
__device__ __forceinline__ void accum_outer(const float* drows, const float* inrows,
                                            float* __restrict__ gp, float* __restrict__ gb,
                                            int tid) {
  for (int e = tid; e < O * K; e += kThreads) {
    const int o = e / K, c = e % K;
    const f32x4* dr = reinterpret_cast<const f32x4*>(drows + o * kLdB);
    const f32x4* ir = reinterpret_cast<const f32x4*>(inrows + c * kLdB);
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int q = 0; q < kTok / 4; ++q) acc += dr[q] * ir[q];
    // no-return atomic add: the workgroup's own slab, one writer per entry,
    // so it only saves the load-to-store round trip of a read-modify-write
    unsafeAtomicAdd(gp + e, (acc[0] + acc[1]) + (acc[2] + acc[3]));
  }
  if (gb) {
    for (int o = tid; o < O; o += kThreads) {
      const f32x4* dr = reinterpret_cast<const f32x4*>(drows + o * kLdB);
      f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int q = 0; q < kTok / 4; ++q) acc += dr[q];
      unsafeAtomicAdd(gb + o, (acc[0] + acc[1]) + (acc[2] + acc[3]));
    }
  }
}

// rows[k][lane] = (add ? rows[k][lane] : 0) + sum_o W[o][k] d[o] for this wave's
// NK outputs k0.. (W is [O][K], wave-uniform).  The o loop is outermost so
// each step reads the NK contiguous weights W[o][k0 .. k0 + NK) at once.
template <int O, int K, int NK>
__device__ __forceinline__ void token_dense_t(const float* __restrict__ W, const float (&d)[O],
                                              int k0, float* rows, const float* addrows,
                                              const float* maskrows, int lane) {
  float acc[NK];
#pragma unroll
  for (int i = 0; i < NK; ++i) acc[i] = addrows ? addrows[(k0 + i) * kLdB + lane] : 0.0f;
#pragma unroll
  for (int o = 0; o < O; ++o) {
    const float* wr = W + (size_t)o * K + k0;
#pragma unroll
    for (int i = 0; i < NK; ++i) acc[i] += wr[i] * d[o];
  }
#pragma unroll
  for (int i = 0; i < NK; ++i) {
    const int k = k0 + i;
    float v = acc[i];
    if (maskrows && !(maskrows[k * kLdB + lane] > 0.0f)) v = 0.0f;
    rows[k * kLdB + lane] = v;
  }
}

// Forward of one Hydra stack from rows xin into QKV / OUT / HID (and xout if
// given); all waves take part, synchronised inside.
template <int F, int H>
__device__ __forceinline__ void hydra_stack_fwd(const float* ws, int st, float* sm, int xin,
                                                int xout, int lane, int wave, int V) {
  typedef HydraBwdRows<F> R;
  constexpr SmallOffsets so = small_offsets(F, H);
  constexpr int stride = so.q[1] - so.q[0];
  const float vmask = (lane & 3) < V ? 1.0f : 0.0f;
  {
    float x[F];
    load_rows<F, kLdB>(sm + xin * kLdB, lane, x);
    constexpr int NO = 3 * kHydraInner / kWaves;
    token_dense<F, NO, false, kLdB>(ws + so.q[0] + st * stride, nullptr, x, wave * NO,
                                    sm + R::QKV * kLdB, lane);
  }
  __syncthreads();
  float attn[kHydraInner];
  {
    float qq = 0.0f, kk = 0.0f;
#pragma unroll
    for (int e = 0; e < kHydraInner; ++e) {
      const float q = sm[(R::QKV + e) * kLdB + lane];
      const float k = sm[(R::QKV + kHydraInner + e) * kLdB + lane];
      qq += q * q;
      kk += k * k;
    }
    const float nq = 1.0f / fmaxf(sqrtf(qq), 1e-12f);
    const float nk = vmask / fmaxf(sqrtf(kk), 1e-12f);
#pragma unroll
    for (int e = 0; e < kHydraInner; ++e) {
      const float q = sm[(R::QKV + e) * kLdB + lane];
      const float k = sm[(R::QKV + kHydraInner + e) * kLdB + lane];
      const float v = sm[(R::QKV + 2 * kHydraInner + e) * kLdB + lane];
      float kv = k * nk * v;
      kv += __shfl_xor(kv, 1, 64);
      kv += __shfl_xor(kv, 2, 64);
      attn[e] = q * nq * kv;
    }
  }
  {
    constexpr int NO = F / kWaves;
    const float* wf = ws + so.fc[0] + st * stride;
#pragma unroll
    for (int i = 0; i < NO; ++i) {
      const int c = wave * NO + i;
      float acc = 0.0f;
#pragma unroll
      for (int e = 0; e < kHydraInner; ++e) acc += wf[c * kHydraInner + e] * attn[e];
      sm[(R::OUT + c) * kLdB + lane] = acc + sm[(xin + c) * kLdB + lane];
    }
  }
  __syncthreads();
  {
    float o[F];
    load_rows<F, kLdB>(sm + R::OUT * kLdB, lane, o);
    constexpr int NO = kHydraInner / kWaves;
    token_dense<F, NO, true, kLdB>(ws + so.w1[0] + st * stride, ws + so.b1[0] + st * stride, o,
                                   wave * NO, sm + R::HID * kLdB, lane);
  }
  __syncthreads();
  if (xout >= 0) {
    float h[kHydraInner];
    load_rows<kHydraInner, kLdB>(sm + R::HID * kLdB, lane, h);
    constexpr int NO = F / kWaves;
    const float* w2 = ws + so.w2[0] + st * stride;
    const float* b2 = ws + so.b2[0] + st * stride;
#pragma unroll
    for (int i = 0; i < NO; ++i) {
      const int c = wave * NO + i;
      float acc = b2[c];
#pragma unroll
      for (int e = 0; e < kHydraInner; ++e) acc += w2[c * kHydraInner + e] * h[e];
      sm[(xout + c) * kLdB + lane] = acc + sm[(R::OUT + c) * kLdB + lane];
    }
    __syncthreads();
  }
}

template <int F, int H>
__global__ void __launch_bounds__(kThreads)
__attribute__((amdgpu_waves_per_eu(2)))
hydra_rgb_backward_kernel(const float* __restrict__ tok, const int* __restrict__ list,
                          const unsigned* __restrict__ n_valid_ptr, const float* __restrict__ ws,
                          int V, const float* __restrict__ grad_rgb, float* __restrict__ dtok,
                          float* __restrict__ gpart) {
  typedef HydraBwdRows<F> R;
  constexpr SmallOffsets so = small_offsets(F, H);
  constexpr int stride = so.q[1] - so.q[0];
  extern __shared__ __attribute__((aligned(16))) char smem_b[];
  float* sm = reinterpret_cast<float*>(smem_b);
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  float* g = gpart + (size_t)blockIdx.x * so.total;  // this workgroup's partial gradients
  const long long n_valid = *n_valid_ptr;
  for (long long item0 = (long long)blockIdx.x * kRaysPerGroup; item0 < n_valid;
       item0 += (long long)gridDim.x * kRaysPerGroup) {
    __syncthreads();
    const long long n_floats = (n_valid - item0) * kVP * F;
    for (int i = tid; i < kTok * F; i += kThreads) {
      const int tk = i / F, c = i % F;
      sm[(R::X0 + c) * kLdB + tk] = i < n_floats ? tok[item0 * kVP * F + i] : 0.0f;
    }
    __syncthreads();
    // ---- forward recompute: stacks (inputs kept), RGB head ----------------
    hydra_stack_fwd<F, H>(ws, 0, sm, R::X0, R::X1, lane, wave, V);
    hydra_stack_fwd<F, H>(ws, 1, sm, R::X1, R::X2, lane, wave, V);
    {
      float x[F];
      load_rows<F, kLdB>(sm + R::X2 * kLdB, lane, x);
      constexpr int NO = kRgbH1 / kWaves;
      token_dense<F, NO, true, kLdB>(ws + so.r1w, ws + so.r1b, x, wave * NO,
                                     sm + R::QKV * kLdB, lane);
    }
    __syncthreads();
    {
      float r1[kRgbH1];
      load_rows<kRgbH1, kLdB>(sm + R::QKV * kLdB, lane, r1);
      constexpr int NO = kRgbH2 / kWaves;
      token_dense<kRgbH1, NO, true, kLdB>(ws + so.r2w, ws + so.r2b, r1, wave * NO,
                                          sm + (R::QKV + kRgbH1) * kLdB, lane);
    }
    __syncthreads();
    // dz3 = dL/drgb * rgb (1 - rgb) on the first view's token of each valid ray
    if (wave == 0) {
      const long long item = item0 + (lane >> 2);
      const bool on = (lane & 3) == 0 && item < n_valid;
      float r2[kRgbH2];
      load_rows<kRgbH2, kLdB>(sm + (R::QKV + kRgbH1) * kLdB, lane, r2);
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        float z = ws[so.r3b + j];
#pragma unroll
        for (int k = 0; k < kRgbH2; ++k) z += ws[so.r3w + j * kRgbH2 + k] * r2[k];
        const float rgb = sigmoidf(z);
        const float gr = on ? grad_rgb[(long long)list[item] * 3 + j] : 0.0f;
        sm[(R::DZ3 + j) * kLdB + lane] = gr * rgb * (1.0f - rgb);
      }
    }
    __syncthreads();
    accum_outer<3, kRgbH2>(sm + R::DZ3 * kLdB, sm + (R::QKV + kRgbH1) * kLdB, g + so.r3w,
                           g + so.r3b, tid);
    {  // dr2 = (Wr3^T dz3) . [r2 > 0]  -> DOUT rows [0, 32)
      float dz[3];
      load_rows<3, kLdB>(sm + R::DZ3 * kLdB, lane, dz);
      constexpr int NK = kRgbH2 / kWaves;
      token_dense_t<3, kRgbH2, NK>(ws + so.r3w, dz, wave * NK, sm + R::DOUT * kLdB, nullptr,
                                   sm + (R::QKV + kRgbH1) * kLdB, lane);
    }
    __syncthreads();
    accum_outer<kRgbH2, kRgbH1>(sm + R::DOUT * kLdB, sm + R::QKV * kLdB, g + so.r2w,
                                g + so.r2b, tid);
    {  // dr1 = (Wr2^T dr2) . [r1 > 0]  -> DX rows [0, 64)
      float dr2[kRgbH2];
      load_rows<kRgbH2, kLdB>(sm + R::DOUT * kLdB, lane, dr2);
      constexpr int NK = kRgbH1 / kWaves;
      token_dense_t<kRgbH2, kRgbH1, NK>(ws + so.r2w, dr2, wave * NK, sm + R::DX * kLdB, nullptr,
                                        sm + R::QKV * kLdB, lane);
    }
    __syncthreads();
    accum_outer<kRgbH1, F>(sm + R::DX * kLdB, sm + R::X2 * kLdB, g + so.r1w, g + so.r1b, tid);
    {  // dx2 = Wr1^T dr1 -> DX rows
      float dr1[kRgbH1];
      load_rows<kRgbH1, kLdB>(sm + R::DX * kLdB, lane, dr1);
      __syncthreads();
      constexpr int NK = F / kWaves;
      token_dense_t<kRgbH1, F, NK>(ws + so.r1w, dr1, wave * NK, sm + R::DX * kLdB, nullptr,
                                   nullptr, lane);
    }
    __syncthreads();
    // ---- stacks, last to first ----------------------------------------------
    for (int st = kHydraStacks - 1; st >= 0; --st) {
      const int xin = R::X1;
      if (st == 0) {
        // stack 0's input: the tokens again, into the rows that held stack
        // 1's input (its backward has finished with them)
        for (int i = tid; i < kTok * F; i += kThreads) {
          const int tk = i / F, c = i % F;
          sm[(R::X1 + c) * kLdB + tk] = i < n_floats ? tok[item0 * kVP * F + i] : 0.0f;
        }
        __syncthreads();
      }
      hydra_stack_fwd<F, H>(ws, st, sm, xin, -1, lane, wave, V);
      const float* w1 = ws + so.w1[0] + st * stride;
      const float* w2 = ws + so.w2[0] + st * stride;
      const float* wf = ws + so.fc[0] + st * stride;
      const float* wq = ws + so.q[0] + st * stride;
      // x' = W2 hid + b2 + out
      accum_outer<F, kHydraInner>(sm + R::DX * kLdB, sm + R::HID * kLdB, g + so.w2[0] + st * stride,
                                  g + so.b2[0] + st * stride, tid);
      {
        float dxp[F];
        load_rows<F, kLdB>(sm + R::DX * kLdB, lane, dxp);
        constexpr int NK = kHydraInner / kWaves;
        token_dense_t<F, kHydraInner, NK>(w2, dxp, wave * NK, sm + R::DHID * kLdB, nullptr,
                                          sm + R::HID * kLdB, lane);
      }
      __syncthreads();
      // hid = relu(W1 out + b1)
      accum_outer<kHydraInner, F>(sm + R::DHID * kLdB, sm + R::OUT * kLdB,
                                  g + so.w1[0] + st * stride, g + so.b1[0] + st * stride, tid);
      {
        float dh[kHydraInner];
        load_rows<kHydraInner, kLdB>(sm + R::DHID * kLdB, lane, dh);
        constexpr int NK = F / kWaves;
        token_dense_t<kHydraInner, F, NK>(w1, dh, wave * NK, sm + R::DOUT * kLdB,
                                          sm + R::DX * kLdB, nullptr, lane);
      }
      __syncthreads();
      // out = fc(attn) + x: attention values into the HID rows (wave 0)
      const float vmask = (lane & 3) < V ? 1.0f : 0.0f;
      float nq, nk;
      {
        float qq = 0.0f, kk = 0.0f;
#pragma unroll
        for (int e = 0; e < kHydraInner; ++e) {
          const float q = sm[(R::QKV + e) * kLdB + lane];
          const float k = sm[(R::QKV + kHydraInner + e) * kLdB + lane];
          qq += q * q;
          kk += k * k;
        }
        nq = 1.0f / fmaxf(sqrtf(qq), 1e-12f);
        nk = vmask / fmaxf(sqrtf(kk), 1e-12f);
      }
      if (wave == 0) {
#pragma unroll
        for (int e = 0; e < kHydraInner; ++e) {
          const float q = sm[(R::QKV + e) * kLdB + lane];
          const float k = sm[(R::QKV + kHydraInner + e) * kLdB + lane];
          const float v = sm[(R::QKV + 2 * kHydraInner + e) * kLdB + lane];
          float kv = k * nk * v;
          kv += __shfl_xor(kv, 1, 64);
          kv += __shfl_xor(kv, 2, 64);
          sm[(R::HID + e) * kLdB + lane] = q * nq * kv;
        }
      }
      __syncthreads();
      accum_outer<F, kHydraInner>(sm + R::DOUT * kLdB, sm + R::HID * kLdB,
                                  g + so.fc[0] + st * stride, nullptr, tid);
      {  // d attn = Wfc^T dout -> DHID rows
        float dout[F];
        load_rows<F, kLdB>(sm + R::DOUT * kLdB, lane, dout);
        constexpr int NK = kHydraInner / kWaves;
        token_dense_t<F, kHydraInner, NK>(wf, dout, wave * NK, sm + R::DHID * kLdB, nullptr,
                                          nullptr, lane);
      }
      __syncthreads();
      // attention and normalisation backward, per token (wave 0): dq, dk, dv
      // overwrite q, k, v in the QKV rows
      if (wave == 0) {
        // two passes over the 32 attention units: the first accumulates
        // q_hat . dq_hat and k_hat . dk_hat, the second writes dq, dk, dv
        // (recomputing kv and dkv, which are cross-token shuffles)
        float qd = 0.0f, kd = 0.0f;
        for (int pass = 0; pass < 2; ++pass) {
          const float qn = 1.0f / nq, kn = vmask > 0.0f ? vmask / nk : 0.0f;
          const bool q_big = qn > 1e-12f, k_big = kn > 1e-12f;
#pragma unroll 4
          for (int e = 0; e < kHydraInner; ++e) {
            const float q = sm[(R::QKV + e) * kLdB + lane];
            const float k = sm[(R::QKV + kHydraInner + e) * kLdB + lane];
            const float v = sm[(R::QKV + 2 * kHydraInner + e) * kLdB + lane];
            const float da = sm[(R::DHID + e) * kLdB + lane];
            float kv = k * nk * v;
            kv += __shfl_xor(kv, 1, 64);
            kv += __shfl_xor(kv, 2, 64);
            // dq_hat = da * kv; dkv = sum over the ray's tokens of da * q_hat
            const float dqh = da * kv;
            float dkv = da * q * nq;
            dkv += __shfl_xor(dkv, 1, 64);
            dkv += __shfl_xor(dkv, 2, 64);
            const float dkh = dkv * v;
            if (pass == 0) {
              qd += q * nq * dqh;
              kd += k * nk * dkh;
            } else {
              // F.normalize backward: dx = n (dxhat - xhat (xhat . dxhat)) when
              // |x| > eps, n dxhat otherwise (n = 1/max(|x|, eps)); a masked
              // view's k_hat is 0 and gets no gradient
              const float dq = q_big ? nq * (dqh - q * nq * qd) : nq * dqh;
              const float dk =
                  vmask > 0.0f ? (k_big ? nk * (dkh - k * nk * kd) : nk * dkh) : 0.0f;
              const float dv = dkv * k * nk;
              __builtin_amdgcn_wave_barrier();
              sm[(R::QKV + e) * kLdB + lane] = dq;
              sm[(R::QKV + kHydraInner + e) * kLdB + lane] = dk;
              sm[(R::QKV + 2 * kHydraInner + e) * kLdB + lane] = dv;
            }
          }
        }
      }
      __syncthreads();
      accum_outer<3 * kHydraInner, F>(sm + R::QKV * kLdB, sm + xin * kLdB,
                                      g + so.q[0] + st * stride, nullptr, tid);
      {  // dx = dout + Wq^T dq + Wk^T dk + Wv^T dv -> DX rows
        float dqkv[3 * kHydraInner];
        load_rows<3 * kHydraInner, kLdB>(sm + R::QKV * kLdB, lane, dqkv);
        constexpr int NK = F / kWaves;
        token_dense_t<3 * kHydraInner, F, NK>(wq, dqkv, wave * NK, sm + R::DX * kLdB,
                                              sm + R::DOUT * kLdB, nullptr, lane);
      }
      __syncthreads();
    }
    // ---- dL/d token (list order) ------------------------------------------
    for (int i = tid; i < kTok * F; i += kThreads) {
      const int tk = i / F, c = i % F;
      if (i < n_floats) dtok[item0 * kVP * F + i] = sm[(R::DX + c) * kLdB + tk];
    }
  }
}

// ---- Backward of the Hydra encoder + RGB head on fp32 MFMA (F = 32) --------
// One wave per workgroup and 16 tokens (4 rays x 4 view slots) per step of a
// persistent loop over the list.  The three kinds of product all run on
// v_mfma_f32_16x16x4f32 from LDS rows [token][unit] of pitch LD:
//   y = W x        (forward recompute)   hydra_mm32: lane holds units of a token
//   dx = W^T dy    (input gradient)      bmm_wtd: lane holds tokens of a unit
//   dW += dy x^T   (weight gradient)     bmm_dw: reduction over the 16 tokens
// The weight gradients of every Hydra and RGB-head tensor stay in the wave's
// accumulator registers for the whole loop (66 tiles of 16 x 16 at F = 32)
// and are written once, to the workgroup's own partial slab; a second kernel
// adds the slabs.  Stack internals are recomputed before each stack's
// backward (only the stack inputs are kept), as in the forward's order, so
// the recomputed values equal the fp32 forward's bit for bit.
template <int F>
struct HydraBwdSlots {   // floats of one token's LDS row
  static constexpr int HI = kHydraInner;
  static constexpr int X0 = 0, X1 = F, X2 = 2 * F;  // stack inputs, final tokens
  static constexpr int QKV = 3 * F;     // q | k | v [96]; RGB r1 [64] | r2 [32]
  static constexpr int ATT = QKV + 3 * HI;          // attention values [32]; dz3 [4]
  static constexpr int OUT = ATT + HI;              // fc(att) + x [F]
  static constexpr int HID = OUT + F;               // FFN hidden [32]
  static constexpr int DX = HID + HI;               // dL/d stack output, then input [F]
  static constexpr int DOUT = DX + F;               // [F]
  static constexpr int DHID = DOUT + F;             // [32]
  static constexpr int DATT = DHID + HI;            // [32]
  static constexpr int DQKV = DATT + HI;            // [96]; RGB dr2 [32] | dr1 [64]
  static constexpr int NORM = DQKV + 3 * HI;        // 1/|q|, vmask/|k| of the current stack
  static constexpr int LD = NORM + 4;               // 516 at F = 32: rows 4 banks apart
  static_assert(3 * HI == kRgbH1 + kRgbH2, "RGB rows share the q|k|v rows");
};

// acc[nt] = sum_o d[token][o] W[o][16 nt + n] over o < O (dx = W^T d); W is
// [O][K] row-major in global memory, rows o >= n_o read as zero.  Lane (n =
// l%16, g = l/16) takes o in [g OQ, (g+1) OQ): A = d[n][o] (contiguous in
// LDS), B = W[o][16 nt + n].  The result lane holds tokens 4g + i of unit
// 16 nt + n.
template <int NT, int O, int K, int LD>
__device__ __forceinline__ void bmm_wtd(const float* __restrict__ W, int n_o, const float* d,
                                        int lane, f32x4 (&acc)[NT]) {
  constexpr int OQ = (O + 3) / 4;
  const int n = lane & 15, g = lane >> 4;
  float a[OQ];
  if constexpr (OQ % 4 == 0) {
#pragma unroll
    for (int s = 0; s < OQ; s += 4) {
      const f32x4 v = *reinterpret_cast<const f32x4*>(d + n * LD + g * OQ + s);
#pragma unroll
      for (int j = 0; j < 4; ++j) a[s + j] = v[j];
    }
  } else {
#pragma unroll
    for (int s = 0; s < OQ; ++s) a[s] = g * OQ + s < O ? d[n * LD + g * OQ + s] : 0.0f;
  }
#pragma unroll
  for (int nt = 0; nt < NT; ++nt) {
    acc[nt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int s = 0; s < OQ; ++s) {
      const int o = g * OQ + s;
      const float b = o < n_o ? W[(size_t)o * K + 16 * nt + n] : 0.0f;
      acc[nt] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[s], b, acc[nt], 0, 0, 0);
    }
  }
}

// acc[mt][nt] += sum over the 16 tokens of d[t][16 mt + o] x[t][16 nt + k]
// (dW += d x^T; units o >= n_o of d read as zero) and, when db is given,
// db[mt] += this lane's share of sum_t d[t][16 mt + n].  Step s, lane (n, g): A = d[4s + g][16 mt
// + n], B = x[4s + g][16 nt + n]; the result lane holds dW[16 mt + 4g + i]
// [16 nt + n].
template <int MT, int NT, int LD>
__device__ __forceinline__ void bmm_dw(const float* d, int n_o, const float* x, int lane,
                                       f32x4 (&acc)[MT][NT], float* db) {
  const int n = lane & 15, g = lane >> 4;
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const int t = 4 * s + g;
    float b[NT];
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) b[nt] = x[t * LD + 16 * nt + n];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) {
      const int o = 16 * mt + n;
      const float av = o < n_o ? d[t * LD + o] : 0.0f;
      if (db) db[mt] += av;
#pragma unroll
      for (int nt = 0; nt < NT; ++nt)
        acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, b[nt], acc[mt][nt], 0, 0, 0);
    }
  }
}

// rows[4g + i][col0 + 16 nt + n] = v (masked by [mask row > 0] and offset by
// add row when given) for the token-major layout of bmm_wtd's result.
template <int NT, int LD>
__device__ __forceinline__ void bput_t(float* rows, int col0, const f32x4 (&v)[NT],
                                       const float* mask, int mcol0, const float* add,
                                       int acol0, int lane) {
  const int n = lane & 15, g = lane >> 4;
#pragma unroll
  for (int nt = 0; nt < NT; ++nt)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int t = 4 * g + i, u = 16 * nt + n;
      float x = v[nt][i];
      if (add) x += add[t * LD + acol0 + u];
      if (mask && !(mask[t * LD + mcol0 + u] > 0.0f)) x = 0.0f;
      rows[t * LD + col0 + u] = x;
    }
}

// Forward of one Hydra stack (fp32 MFMA) on the wave's 16 token rows: q|k|v,
// the norms, attention values, out and hidden rows, and the stack output
// into row block xout when xout >= 0.  Same arithmetic, in the same order,
// as hydra_rgb_mfma_kernel<.., kF32>.
template <int F, int H>
__device__ __forceinline__ void bstack_fwd(const float* __restrict__ ws, int st, float* rw,
                                           int xin, int xout, int lane, int V) {
  typedef HydraBwdSlots<F> R;
  constexpr int LD = R::LD, HI = kHydraInner;
  constexpr SmallOffsets so = small_offsets(F, H);
  constexpr int stride = so.q[1] - so.q[0];
  const int n = lane & 15, g = lane >> 4;
  const float vmask = (n % kVP) < V ? 1.0f : 0.0f;
  f32x4 a6[6];
  hzero(a6);
  hydra_mm32<6, F, LD>(ws + so.q[0] + st * stride, 3 * HI, rw + xin, lane, a6);
#pragma unroll
  for (int mt = 0; mt < 6; ++mt)
    *reinterpret_cast<f32x4*>(rw + n * LD + R::QKV + 16 * mt + 4 * g) = a6[mt];
  float qq = 0.0f, kk = 0.0f;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    qq += a6[0][i] * a6[0][i] + a6[1][i] * a6[1][i];
    kk += a6[2][i] * a6[2][i] + a6[3][i] * a6[3][i];
  }
  qq += __shfl_xor(qq, 16, 64);
  qq += __shfl_xor(qq, 32, 64);
  kk += __shfl_xor(kk, 16, 64);
  kk += __shfl_xor(kk, 32, 64);
  const float nq = 1.0f / fmaxf(sqrtf(qq), 1e-12f);
  const float nk = vmask / fmaxf(sqrtf(kk), 1e-12f);
  if (g == 0) {
    rw[n * LD + R::NORM] = nq;
    rw[n * LD + R::NORM + 1] = nk;
  }
#pragma unroll
  for (int h = 0; h < 2; ++h)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      float kv = a6[2 + h][i] * nk * a6[4 + h][i];
#pragma unroll
      for (int off = 1; off < kVP; off <<= 1) kv += __shfl_xor(kv, off, 64);
      rw[n * LD + R::ATT + 16 * h + 4 * g + i] = a6[h][i] * nq * kv;
    }
  __syncthreads();
  {
    f32x4 ao[F / 16];
    hzero(ao);
    hydra_mm32<F / 16, HI, LD>(ws + so.fc[0] + st * stride, F, rw + R::ATT, lane, ao);
#pragma unroll
    for (int mt = 0; mt < F / 16; ++mt)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int u = 16 * mt + 4 * g + i;
        rw[n * LD + R::OUT + u] = rw[n * LD + xin + u] + ao[mt][i];
      }
  }
  __syncthreads();
  {
    f32x4 ah[2];
    hzero(ah);
    hydra_mm32<2, F, LD>(ws + so.w1[0] + st * stride, HI, rw + R::OUT, lane, ah);
    const float* b1 = ws + so.b1[0] + st * stride;
#pragma unroll
    for (int mt = 0; mt < 2; ++mt)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int u = 16 * mt + 4 * g + i;
        rw[n * LD + R::HID + u] = fmaxf(ah[mt][i] + b1[u], 0.0f);
      }
  }
  __syncthreads();
  if (xout >= 0) {
    f32x4 ax[F / 16];
    hzero(ax);
    hydra_mm32<F / 16, HI, LD>(ws + so.w2[0] + st * stride, F, rw + R::HID, lane, ax);
    const float* b2 = ws + so.b2[0] + st * stride;
#pragma unroll
    for (int mt = 0; mt < F / 16; ++mt)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int u = 16 * mt + 4 * g + i;
        rw[n * LD + xout + u] = rw[n * LD + R::OUT + u] + (ax[mt][i] + b2[u]);
      }
    __syncthreads();
  }
}

// Weight-gradient accumulators of one Hydra stack (rows x inputs in 16 x 16
// tiles) and its bias-gradient lane partials.
template <int F>
struct StackGrad {
  f32x4 qkv[6][F / 16];   // [96][F]
  f32x4 fc[F / 16][2];    // [F][32]
  f32x4 w1[2][F / 16];    // [32][F]
  f32x4 w2[F / 16][2];    // [F][32]
  float b1[2], b2[F / 16];
};

template <int MT, int NT>
__device__ __forceinline__ void zero_tiles(f32x4 (&t)[MT][NT]) {
#pragma unroll
  for (int mt = 0; mt < MT; ++mt)
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) t[mt][nt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
}

// gp[o K + 16 nt + n] = tile entries (o = 16 mt + 4g + i < n_o)
template <int MT, int NT>
__device__ __forceinline__ void put_tiles(float* gp, const f32x4 (&t)[MT][NT], int n_o, int K,
                                          int lane) {
  const int n = lane & 15, g = lane >> 4;
#pragma unroll
  for (int mt = 0; mt < MT; ++mt)
#pragma unroll
    for (int nt = 0; nt < NT; ++nt)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int o = 16 * mt + 4 * g + i;
        if (o < n_o) gp[(size_t)o * K + 16 * nt + n] = t[mt][nt][i];
      }
}

// gb[16 mt + n] = sum over the four lane groups of db[mt]
template <int MT>
__device__ __forceinline__ void put_bias(float* gb, const float (&db)[MT], int n_o, int lane) {
  const int n = lane & 15, g = lane >> 4;
#pragma unroll
  for (int mt = 0; mt < MT; ++mt) {
    float v = db[mt];
    v += __shfl_xor(v, 16, 64);
    v += __shfl_xor(v, 32, 64);
    if (g == 0 && 16 * mt + n < n_o) gb[16 * mt + n] = v;
  }
}

// Backward of Hydra stack st from dL/d output (DX rows) to dL/d input (DX
// rows again), accumulating the stack's weight gradients.
template <int F, int H>
__device__ __forceinline__ void bstack_bwd(const float* __restrict__ ws, int st, float* rw,
                                           int xin, int lane, int V, StackGrad<F>& gr) {
  typedef HydraBwdSlots<F> R;
  constexpr int LD = R::LD, HI = kHydraInner;
  constexpr SmallOffsets so = small_offsets(F, H);
  constexpr int stride = so.q[1] - so.q[0];
  const int n = lane & 15, g = lane >> 4;
  bstack_fwd<F, H>(ws, st, rw, xin, -1, lane, V);
  const float* w1 = ws + so.w1[0] + st * stride;
  const float* w2 = ws + so.w2[0] + st * stride;
  const float* wf = ws + so.fc[0] + st * stride;
  const float* wq = ws + so.q[0] + st * stride;
  // y = W2 hid + b2 + out
  bmm_dw<F / 16, 2, LD>(rw + R::DX, F, rw + R::HID, lane, gr.w2, gr.b2);
  {
    f32x4 t[2];
    bmm_wtd<2, F, HI, LD>(w2, F, rw + R::DX, lane, t);
    bput_t<2, LD>(rw, R::DHID, t, rw, R::HID, nullptr, 0, lane);
  }
  __syncthreads();
  // hid = relu(W1 out + b1); out also feeds y directly
  bmm_dw<2, F / 16, LD>(rw + R::DHID, HI, rw + R::OUT, lane, gr.w1, gr.b1);
  {
    f32x4 t[F / 16];
    bmm_wtd<F / 16, HI, F, LD>(w1, HI, rw + R::DHID, lane, t);
    bput_t<F / 16, LD>(rw, R::DOUT, t, nullptr, 0, rw, R::DX, lane);
  }
  __syncthreads();
  // out = fc(att) + x
  bmm_dw<F / 16, 2, LD>(rw + R::DOUT, F, rw + R::ATT, lane, gr.fc, nullptr);
  {
    f32x4 t[2];
    bmm_wtd<2, F, HI, LD>(wf, F, rw + R::DOUT, lane, t);
    bput_t<2, LD>(rw, R::DATT, t, nullptr, 0, nullptr, 0, lane);
  }
  __syncthreads();
  // att = q^ * sum_views k^ v with q^ = q nq, k^ = k nk (F.normalize, eps
  // 1e-12; nk = 0 on masked view slots); lane: token n, units 8g .. 8g + 7
  {
    const float nq = rw[n * LD + R::NORM], nk = rw[n * LD + R::NORM + 1];
    const float vmask = (n % kVP) < V ? 1.0f : 0.0f;
    float q[8], k[8], v[8], da[8];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const f32x4 q4 = *reinterpret_cast<const f32x4*>(rw + n * LD + R::QKV + 8 * g + 4 * h);
      const f32x4 k4 = *reinterpret_cast<const f32x4*>(rw + n * LD + R::QKV + HI + 8 * g + 4 * h);
      const f32x4 v4 =
          *reinterpret_cast<const f32x4*>(rw + n * LD + R::QKV + 2 * HI + 8 * g + 4 * h);
      const f32x4 d4 = *reinterpret_cast<const f32x4*>(rw + n * LD + R::DATT + 8 * g + 4 * h);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        q[4 * h + j] = q4[j];
        k[4 * h + j] = k4[j];
        v[4 * h + j] = v4[j];
        da[4 * h + j] = d4[j];
      }
    }
    float dqh[8], dkv[8], qd = 0.0f, kd = 0.0f;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      float kv = k[j] * nk * v[j];
      float dk = da[j] * q[j] * nq;
#pragma unroll
      for (int off = 1; off < kVP; off <<= 1) {
        kv += __shfl_xor(kv, off, 64);
        dk += __shfl_xor(dk, off, 64);
      }
      dqh[j] = da[j] * kv;     // dL/d q^
      dkv[j] = dk;             // dL/d (k^ v), summed over the ray's views
      qd += q[j] * nq * dqh[j];
      kd += k[j] * nk * (dkv[j] * v[j]);
    }
    qd += __shfl_xor(qd, 16, 64);
    qd += __shfl_xor(qd, 32, 64);
    kd += __shfl_xor(kd, 16, 64);
    kd += __shfl_xor(kd, 32, 64);
    // F.normalize backward: dx = n (dx^ - x^ (x^ . dx^)) when |x| > eps, n dx^
    // otherwise (n = 1 / max(|x|, eps)); a masked view's k^ is 0 and gets no
    // gradient
    const float qn = 1.0f / nq, kn = vmask > 0.0f ? vmask / nk : 0.0f;
    const bool q_big = qn > 1e-12f, k_big = kn > 1e-12f;
    float dq[8], dkk[8], dv[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const float dkh = dkv[j] * v[j];
      dq[j] = q_big ? nq * (dqh[j] - q[j] * nq * qd) : nq * dqh[j];
      dkk[j] = vmask > 0.0f ? (k_big ? nk * (dkh - k[j] * nk * kd) : nk * dkh) : 0.0f;
      dv[j] = dkv[j] * k[j] * nk;
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      float* base = rw + n * LD + R::DQKV + 8 * g + 4 * h;
      *reinterpret_cast<f32x4*>(base) =
          f32x4{dq[4 * h], dq[4 * h + 1], dq[4 * h + 2], dq[4 * h + 3]};
      *reinterpret_cast<f32x4*>(base + HI) =
          f32x4{dkk[4 * h], dkk[4 * h + 1], dkk[4 * h + 2], dkk[4 * h + 3]};
      *reinterpret_cast<f32x4*>(base + 2 * HI) =
          f32x4{dv[4 * h], dv[4 * h + 1], dv[4 * h + 2], dv[4 * h + 3]};
    }
  }
  __syncthreads();
  // q|k|v = Wqkv x; x also feeds out directly
  bmm_dw<6, F / 16, LD>(rw + R::DQKV, 3 * HI, rw + xin, lane, gr.qkv, nullptr);
  {
    f32x4 t[F / 16];
    bmm_wtd<F / 16, 3 * HI, F, LD>(wq, 3 * HI, rw + R::DQKV, lane, t);
    bput_t<F / 16, LD>(rw, R::DX, t, nullptr, 0, rw, R::DOUT, lane);
  }
  __syncthreads();
}

// Part PART of the backward, over every 16-token step of the workgroup:
//   0: forward of both stacks and the RGB head, dL/dz of the RGB output, the
//      head's backward -> dL/d final tokens (in dtok);
//   1: stack 0's forward, stack 1's backward from dtok -> dL/d stack-1 input
//      (in dtok);
//   2: stack 0's backward from dtok -> dL/d tokens (in dtok).
// Three launches keep at most 24 accumulator tiles per kernel in registers
// (all 66 at once do not fit beside the working set); dtok is read and
// rewritten in place by the wave that owns the tokens.
template <int F, int H, int PART>
__global__ void __launch_bounds__(64)
hydra_rgb_backward_mfma_kernel(const float* __restrict__ tok, const int* __restrict__ list,
                               const unsigned* __restrict__ n_valid_ptr,
                               const float* __restrict__ ws, int V,
                               const float* __restrict__ grad_rgb, float* __restrict__ dtok,
                               float* __restrict__ gpart) {
  typedef HydraBwdSlots<F> R;
  constexpr int LD = R::LD;
  constexpr SmallOffsets so = small_offsets(F, H);
  constexpr int stride = so.q[1] - so.q[0];
  constexpr int RPW = 16 / kVP;   // rays per step
  static_assert(F % 16 == 0 && kRgbH1 % 16 == 0 && kRgbH2 % 16 == 0, "whole tiles");
  __shared__ __attribute__((aligned(16))) float rw[16 * LD];
  const int lane = threadIdx.x;
  const int n = lane & 15, g = lane >> 4;

  [[maybe_unused]] StackGrad<F> gs;
  [[maybe_unused]] f32x4 g_r1[kRgbH1 / 16][F / 16], g_r2[kRgbH2 / 16][kRgbH1 / 16],
      g_r3[1][kRgbH2 / 16];
  [[maybe_unused]] float b_r1[kRgbH1 / 16], b_r2[kRgbH2 / 16], b_r3[1];
  if constexpr (PART == 0) {
    zero_tiles(g_r1);
    zero_tiles(g_r2);
    zero_tiles(g_r3);
#pragma unroll
    for (int i = 0; i < kRgbH1 / 16; ++i) b_r1[i] = 0.0f;
#pragma unroll
    for (int i = 0; i < kRgbH2 / 16; ++i) b_r2[i] = 0.0f;
    b_r3[0] = 0.0f;
  } else {
    zero_tiles(gs.qkv);
    zero_tiles(gs.fc);
    zero_tiles(gs.w1);
    zero_tiles(gs.w2);
#pragma unroll
    for (int i = 0; i < 2; ++i) gs.b1[i] = 0.0f;
#pragma unroll
    for (int i = 0; i < F / 16; ++i) gs.b2[i] = 0.0f;
  }

  const long long n_valid = *n_valid_ptr;
  for (long long item0 = (long long)blockIdx.x * RPW; item0 < n_valid;
       item0 += (long long)gridDim.x * RPW) {
    const long long n_tok = (n_valid - item0) * kVP;   // tokens that exist from item0 on
    // the weights are re-read from L1/L2 every step: an opaque copy of the
    // pointer keeps the compiler from hoisting thousands of loop-invariant
    // weight loads into registers
    const float* wsl = ws;
    asm volatile("" : "+s"(wsl));
    __syncthreads();
    for (int i = lane; i < 16 * F / 4; i += 64) {
      const int tk = (4 * i) / F, c = (4 * i) % F;
      f32x4 v = f32x4{0.0f, 0.0f, 0.0f, 0.0f}, d = v;
      if (tk < n_tok) {
        v = *reinterpret_cast<const f32x4*>(tok + (item0 * kVP + tk) * F + c);
        if (PART > 0) d = *reinterpret_cast<const f32x4*>(dtok + (item0 * kVP + tk) * F + c);
      }
      *reinterpret_cast<f32x4*>(rw + tk * LD + R::X0 + c) = v;
      if (PART > 0) *reinterpret_cast<f32x4*>(rw + tk * LD + R::DX + c) = d;
    }
    __syncthreads();
    if constexpr (PART == 0) {
      // ---- forward: both stacks, the RGB head, dL/dz of the RGB output ----
      bstack_fwd<F, H>(wsl, 0, rw, R::X0, R::X1, lane, V);
      bstack_fwd<F, H>(wsl, 1, rw, R::X1, R::X2, lane, V);
      {
        f32x4 a1[kRgbH1 / 16];
        hzero(a1);
        hydra_mm32<kRgbH1 / 16, F, LD>(wsl + so.r1w, kRgbH1, rw + R::X2, lane, a1);
#pragma unroll
        for (int mt = 0; mt < kRgbH1 / 16; ++mt)
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const int u = 16 * mt + 4 * g + i;
            rw[n * LD + R::QKV + u] = fmaxf(a1[mt][i] + wsl[so.r1b + u], 0.0f);
          }
      }
      __syncthreads();
      {
        f32x4 a2[kRgbH2 / 16];
        hzero(a2);
        hydra_mm32<kRgbH2 / 16, kRgbH1, LD>(wsl + so.r2w, kRgbH2, rw + R::QKV, lane, a2);
#pragma unroll
        for (int mt = 0; mt < kRgbH2 / 16; ++mt)
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const int u = 16 * mt + 4 * g + i;
            rw[n * LD + R::QKV + kRgbH1 + u] = fmaxf(a2[mt][i] + wsl[so.r2b + u], 0.0f);
          }
      }
      __syncthreads();
      {
        f32x4 a3[1];
        hzero(a3);
        hydra_mm32<1, kRgbH2, LD>(wsl + so.r3w, 3, rw + R::QKV + kRgbH1, lane, a3);
        // lane (n, g = 0) holds the three outputs of token n
        if (g == 0) {
          const long long item = item0 + n / kVP;
          const bool on = n % kVP == 0 && item < n_valid;
          const long long ray = on ? (long long)list[item] : 0;
#pragma unroll
          for (int j = 0; j < 3; ++j) {
            const float rgb = sigmoidf(a3[0][j] + wsl[so.r3b + j]);
            const float gr = on ? grad_rgb[ray * 3 + j] : 0.0f;
            rw[n * LD + R::ATT + j] = gr * rgb * (1.0f - rgb);
          }
          rw[n * LD + R::ATT + 3] = 0.0f;
        }
      }
      __syncthreads();
      // ---- RGB head backward -------------------------------------------------
      bmm_dw<1, kRgbH2 / 16, LD>(rw + R::ATT, 3, rw + R::QKV + kRgbH1, lane, g_r3, b_r3);
      {
        f32x4 t[kRgbH2 / 16];
        bmm_wtd<kRgbH2 / 16, 3, kRgbH2, LD>(wsl + so.r3w, 3, rw + R::ATT, lane, t);
        bput_t<kRgbH2 / 16, LD>(rw, R::DQKV, t, rw, R::QKV + kRgbH1, nullptr, 0, lane);
      }
      __syncthreads();
      bmm_dw<kRgbH2 / 16, kRgbH1 / 16, LD>(rw + R::DQKV, kRgbH2, rw + R::QKV, lane, g_r2, b_r2);
      {
        f32x4 t[kRgbH1 / 16];
        bmm_wtd<kRgbH1 / 16, kRgbH2, kRgbH1, LD>(wsl + so.r2w, kRgbH2, rw + R::DQKV, lane, t);
        bput_t<kRgbH1 / 16, LD>(rw, R::DQKV + kRgbH2, t, rw, R::QKV, nullptr, 0, lane);
      }
      __syncthreads();
      bmm_dw<kRgbH1 / 16, F / 16, LD>(rw + R::DQKV + kRgbH2, kRgbH1, rw + R::X2, lane, g_r1,
                                      b_r1);
      {
        f32x4 t[F / 16];
        bmm_wtd<F / 16, kRgbH1, F, LD>(wsl + so.r1w, kRgbH1, rw + R::DQKV + kRgbH2, lane, t);
        bput_t<F / 16, LD>(rw, R::DX, t, nullptr, 0, nullptr, 0, lane);
      }
      __syncthreads();
    } else if constexpr (PART == 1) {
      bstack_fwd<F, H>(wsl, 0, rw, R::X0, R::X1, lane, V);
      bstack_bwd<F, H>(wsl, 1, rw, R::X1, lane, V, gs);
    } else {
      bstack_bwd<F, H>(wsl, 0, rw, R::X0, lane, V, gs);
    }
    for (int i = lane; i < 16 * F / 4; i += 64) {
      const int tk = (4 * i) / F, c = (4 * i) % F;
      if (tk < n_tok)
        *reinterpret_cast<f32x4*>(dtok + (item0 * kVP + tk) * F + c) =
            *reinterpret_cast<const f32x4*>(rw + tk * LD + R::DX + c);
    }
  }
  // ---- this workgroup's partial gradients ---------------------------------
  float* gp = gpart + (size_t)blockIdx.x * so.total;
  if constexpr (PART == 0) {
    put_tiles(gp + so.r1w, g_r1, kRgbH1, F, lane);
    put_bias(gp + so.r1b, b_r1, kRgbH1, lane);
    put_tiles(gp + so.r2w, g_r2, kRgbH2, kRgbH1, lane);
    put_bias(gp + so.r2b, b_r2, kRgbH2, lane);
    put_tiles(gp + so.r3w, g_r3, 3, kRgbH2, lane);
    put_bias(gp + so.r3b, b_r3, 3, lane);
  } else {
    const int st = PART == 1 ? 1 : 0;
    put_tiles(gp + so.q[0] + st * stride, gs.qkv, 3 * kHydraInner, F, lane);
    put_tiles(gp + so.fc[0] + st * stride, gs.fc, F, kHydraInner, lane);
    put_tiles(gp + so.w1[0] + st * stride, gs.w1, kHydraInner, F, lane);
    put_bias(gp + so.b1[0] + st * stride, gs.b1, kHydraInner, lane);
    put_tiles(gp + so.w2[0] + st * stride, gs.w2, F, kHydraInner, lane);
    put_bias(gp + so.b2[0] + st * stride, gs.b2, F, lane);
  }
}

// out[e] = sum over parts of part[p][e]
// out[e] (+)= sum_p part[p][e], accumulated in double (the partial gradients
// cancel heavily).  A 256-thread block takes kSumEntries consecutive entries
// with 8 threads each, thread g summing partials g, g + 8, ...; the 8 sums
// are added in a fixed order, so the result does not depend on scheduling.
// (One thread per entry walking all partials left most CUs idle: a few
// hundred workgroups of dependent loads.)
constexpr int kSumEntries = 32;
__global__ void __launch_bounds__(256)
sum_partials(const float* __restrict__ part, int parts, int n, float* __restrict__ out,
             int accumulate) {
  __shared__ double s_sum[8][kSumEntries];
  const int o = threadIdx.x % kSumEntries, g = threadIdx.x / kSumEntries;
  const int e = blockIdx.x * kSumEntries + o;
  double s = 0.0;
  if (e < n)
    for (int p = g; p < parts; p += 8) s += (double)part[(size_t)p * n + e];
  s_sum[g][o] = s;
  __syncthreads();
  if (g == 0 && e < n) {
    double t = accumulate ? (double)out[e] : 0.0;
    for (int k = 0; k < 8; ++k) t += s_sum[k][o];
    out[e] = (float)t;
  }
}

void launch_sum_partials(const float* part, int parts, int n, float* out, int accumulate,
                         hipStream_t st) {
  hipLaunchKernelGGL(sum_partials, dim3((unsigned)((n + kSumEntries - 1) / kSumEntries)),
                     dim3(256), 0, st, part, parts, n, out, accumulate);
}

// zero a float buffer (graph-capturable, no memset)
__global__ void fill_zero(float* __restrict__ p, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    p[i] = 0.0f;
}

// Workspace: [counter (256 B)] [list: n_rays int32, 256-B aligned]
// [tokens: n_rays x 4 x F f32].
size_t workspace_list_offset() { return 256; }
size_t workspace_token_offset(long long n_rays) {
  return 256 + ((size_t)n_rays * 4 + 255) / 256 * 256;
}

// after the tokens: the 16-bit copy of the small weights (4 leading elements
// put every weight matrix on a 16-byte boundary); sized for H = 4 F, the
// instantiated models
size_t workspace_small16_offset(long long n_rays, int vp, int F) {
  return (workspace_token_offset(n_rays) + (size_t)n_rays * vp * F * sizeof(float) + 255) / 256 *
         256;
}
size_t workspace_small16_bytes(int F) {
  return ((size_t)small_offsets(F, 4 * F).total + 4) * 2;
}

// Lists every ray with a full set of samples (t[ray, 0] >= 0) in ray order
// and writes the outputs of dropped rays (zero colour, depth and sums).
// Three launches (no atomics, so the list is the same on every run and
// follows the image's raster order, which the sample kernel's XCD banding
// relies on): per block of kCompactRays rays the number of listed rays; one
// workgroup scans the block counts (and stores the total); each block writes
// its rays at its offset.  The block counts live at the start of the token
// area of the workspace, which the sample kernel overwrites afterwards.
constexpr int kCompactRays = 1024;   // rays per compaction block (4 per thread)

__device__ __forceinline__ int compact_valid(const float* t, int S, long long n_rays,
                                             long long ray) {
  return ray < n_rays && t[ray * S] >= 0.0f;
}

// inclusive scan of one int per thread over a 256-thread block; returns the
// exclusive prefix and stores the block total in *total
__device__ __forceinline__ int block_exclusive_scan(int v, int* total) {
  __shared__ int s_w[4];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int x = v;
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) s_w[wid] = x;
  __syncthreads();
  int before = 0, all = 0;
  for (int w = 0; w < 4; ++w) {
    if (w < wid) before += s_w[w];
    all += s_w[w];
  }
  __syncthreads();
  *total = all;
  return before + x - v;
}

__global__ void __launch_bounds__(256)
compact_count(const float* __restrict__ t, int S, long long n_rays, int* __restrict__ bcnt) {
  const long long base = (long long)blockIdx.x * kCompactRays + threadIdx.x * 4;
  int c = 0;
#pragma unroll
  for (int i = 0; i < 4; ++i) c += compact_valid(t, S, n_rays, base + i);
  int total;
  block_exclusive_scan(c, &total);
  if (threadIdx.x == 0) bcnt[blockIdx.x] = total;
}

__global__ void __launch_bounds__(256)
compact_scan(int* __restrict__ bcnt, int nb, unsigned* __restrict__ counter) {
  int carry = 0;
  for (int b0 = 0; b0 < nb; b0 += 256) {
    const int b = b0 + threadIdx.x;
    const int v = b < nb ? bcnt[b] : 0;
    int total;
    const int ex = block_exclusive_scan(v, &total);
    if (b < nb) bcnt[b] = carry + ex;
    carry += total;
  }
  if (threadIdx.x == 0) counter[0] = (unsigned)carry;
}

__global__ void __launch_bounds__(256)
compact_write(const float* __restrict__ t, int S, long long n_rays, const int* __restrict__ bcnt,
              int* __restrict__ list, float* __restrict__ out_rgb, float* __restrict__ out_depth,
              float* __restrict__ out_aux) {
  const long long base = (long long)blockIdx.x * kCompactRays + threadIdx.x * 4;
  int v[4], c = 0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    v[i] = compact_valid(t, S, n_rays, base + i);
    c += v[i];
  }
  int total;
  int pos = bcnt[blockIdx.x] + block_exclusive_scan(c, &total);
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const long long ray = base + i;
    if (v[i]) {
      list[pos++] = (int)ray;
    } else if (ray < n_rays) {
      out_rgb[ray * 3 + 0] = 0.0f;
      out_rgb[ray * 3 + 1] = 0.0f;
      out_rgb[ray * 3 + 2] = 0.0f;
      out_depth[ray] = 0.0f;
      if (out_aux) {
        out_aux[ray * 2 + 0] = 0.0f;
        out_aux[ray * 2 + 1] = 0.0f;
      }
    }
  }
}

template <int F, int H, int PREC, int NS, bool HALF_MAP, bool POINTS, int VP>
int launch_kernels(const RenderArgs& a, hipStream_t st);

template <int F, int H, int PREC, int NS, int VP = kVP>
int launch(const RenderArgs& a, hipStream_t st, bool dry) {
  typedef Smem<F, H, PREC, NS, VP> L;
  if constexpr (L::total > kLdsBytes) {
    return sailor::set_error(SAILOR_ERR_ARG, "render_rays: %d-sample chunks need %zu bytes of "
                             "LDS at this precision (limit %zu)", NS, L::total, kLdsBytes);
  } else {
  if (dry) return SAILOR_OK;  // configuration check only
  if constexpr (VP != kVP) {  // more than 4 views: rays only (no point mode)
    if (a.points)
      return sailor::set_error(SAILOR_ERR_ARG, "density_points: at most %d views", kVP);
    if constexpr (PREC != kF32) {
      if (a.fmap_half) return launch_kernels<F, H, PREC, NS, true, false, VP>(a, st);
    }
    return launch_kernels<F, H, PREC, NS, false, false, VP>(a, st);
  } else {
  if constexpr (PREC != kF32) {
    if (a.fmap_half) {
      return a.points ? launch_kernels<F, H, PREC, NS, true, true, VP>(a, st)
                      : launch_kernels<F, H, PREC, NS, true, false, VP>(a, st);
    }
  }
  return a.points ? launch_kernels<F, H, PREC, NS, false, true, VP>(a, st)
                  : launch_kernels<F, H, PREC, NS, false, false, VP>(a, st);
  }
  }
}

// out[i] = in[i] rounded to the 16-bit type (round to nearest even, as
// hcvt8 rounds)
template <typename E>
__global__ void small_to_16bit(const float* __restrict__ in, int n, E* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (E)in[i];
}

template <int F, int H, int PREC, int NS, bool HALF_MAP, bool POINTS, int VP>
int launch_kernels(const RenderArgs& a, hipStream_t st) {
  typedef Smem<F, H, PREC, NS, VP> L;
  auto kern = render_kernel<F, H, PREC, NS, HALF_MAP, POINTS, VP>;
  static std::atomic<int> grid_of[sailor::kMaxDevices];   // resident grid per device
  int dev = 0;
  if (int e = sailor::device_slot(&dev)) return e;
  int grid = grid_of[dev].load(std::memory_order_acquire);
  if (grid == 0) {
    SAILOR_HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                       hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)L::total));
    int cus = 0, per_cu = 0;
    SAILOR_HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    SAILOR_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, kThreads,
                                                                L::total));
    grid = cus * std::max(per_cu, 1);
    grid_of[dev].store(grid, std::memory_order_release);
  }
  // a frame batched with others on several streams may leave part of the
  // chip to their kernels (grid_share percent of the resident slots)
  const long long g = a.grid_share > 0 && a.grid_share < 100
                          ? std::max<long long>(1, (long long)grid * a.grid_share / 100)
                          : grid;
  const int blocks = (int)std::min<long long>(g, a.n_rays);
  if (a.stages & 2) {
    hipLaunchKernelGGL(kern, dim3(blocks), dim3(kThreads), L::total, st, a);
    if (int err = sailor::check_launch("render_rays")) return err;
  }
  if (POINTS || !(a.stages & 4)) return SAILOR_OK;
  constexpr int rpg = kTok / VP;
  const long long groups = (a.n_rays + rpg - 1) / rpg;
  if constexpr (PREC != kF32) {
    // 16-bit forms: Hydra + RGB head on 16-bit MFMA, unless the caller
    // differentiates the render (its backward recomputes them in fp32)
    if (!a.fp32_hydra) {
      // the small weights in the 16-bit type, once per launch (every wave
      // then reads half the bytes and converts nothing)
      typedef typename HydraOp<PREC>::E E;
      constexpr int n_small = small_offsets(F, H).total;
      hipLaunchKernelGGL((small_to_16bit<E>), dim3((n_small + 255) / 256), dim3(256), 0, st,
                         a.wsmall, n_small, static_cast<E*>(a.wsmall16) + 4);
      hipLaunchKernelGGL((hydra_rgb_mfma_kernel<F, H, VP, PREC>), dim3((unsigned)groups),
                         dim3(kThreads), 0, st, a.out_tok, a.list, a.n_valid, a.wsmall,
                         a.wsmall16, a.V, a.out_rgb);
      return sailor::check_launch("render_rays (hydra, mfma)");
    }
  }
  hipLaunchKernelGGL((hydra_rgb_mfma_kernel<F, H, VP, kF32>), dim3((unsigned)groups),
                     dim3(kThreads), 0, st, a.out_tok, a.list, a.n_valid, a.wsmall, nullptr,
                     a.V, a.out_rgb);
  return sailor::check_launch("render_rays (hydra, fp32 mfma)");
}

template <int F, int H, int NS>
int dispatch_prec(int prec, const RenderArgs& a, hipStream_t st, bool dry) {
  switch (prec) {
    case kF32: return launch<F, H, kF32, NS>(a, st, dry);
    case kBF16: return launch<F, H, kBF16, NS>(a, st, dry);
    case kF16: return launch<F, H, kF16, NS>(a, st, dry);
  }
  return sailor::set_error(SAILOR_ERR_ARG, "render_rays: precision %d unsupported", prec);
}

// 8 view slots: fp32 8, 12 or 16 samples per chunk, 16-bit 4, 8, 12 or 16
// (default 12, 8 when S is not a multiple of 12): the 4-slot defaults'
// column counts and their neighbours; chunks whose LDS does not fit are
// rejected by launch().
template <int F, int H>
int dispatch_vp8(int ns, int prec, const RenderArgs& a, hipStream_t st, bool dry) {
  if (prec == kF32) {
    switch (ns) {
      case 8: return launch<F, H, kF32, 8, 8>(a, st, dry);
      case 12: return launch<F, H, kF32, 12, 8>(a, st, dry);
      case 16: return launch<F, H, kF32, 16, 8>(a, st, dry);
    }
  } else if (prec == kBF16 || prec == kF16) {
    switch (ns) {
      case 4: return prec == kBF16 ? launch<F, H, kBF16, 4, 8>(a, st, dry)
                                   : launch<F, H, kF16, 4, 8>(a, st, dry);
      case 8: return prec == kBF16 ? launch<F, H, kBF16, 8, 8>(a, st, dry)
                                   : launch<F, H, kF16, 8, 8>(a, st, dry);
      case 12: return prec == kBF16 ? launch<F, H, kBF16, 12, 8>(a, st, dry)
                                    : launch<F, H, kF16, 12, 8>(a, st, dry);
      case 16: return prec == kBF16 ? launch<F, H, kBF16, 16, 8>(a, st, dry)
                                    : launch<F, H, kF16, 16, 8>(a, st, dry);
    }
  } else {
    return sailor::set_error(SAILOR_ERR_ARG, "render_rays: precision %d unsupported", prec);
  }
  return sailor::set_error(SAILOR_ERR_ARG, "render_rays: %d samples per chunk not "
                           "instantiated for 8 view slots (fp32: 8, 12, 16; 16-bit: 4, 8, 12, "
                           "16)", ns);
}

template <int F, int H>
int dispatch_ns(int ns, int prec, const RenderArgs& a, hipStream_t st, bool dry) {
  if (a.V > kVP) return dispatch_vp8<F, H>(ns, prec, a, st, dry);
  switch (ns) {
    case 8: return dispatch_prec<F, H, 8>(prec, a, st, dry);
    case 16: return dispatch_prec<F, H, 16>(prec, a, st, dry);
    case 24: return dispatch_prec<F, H, 24>(prec, a, st, dry);
    case 48: return dispatch_prec<F, H, 48>(prec, a, st, dry);
  }
  return sailor::set_error(SAILOR_ERR_ARG, "render_rays: samples per chunk %d not "
                           "instantiated (8, 16, 24, 48)", ns);
}

}  // namespace

extern "C" size_t sailor_render_small_weights_floats(int F, int H) {
  return (size_t)small_offsets(F, H).total;
}

extern "C" size_t sailor_render_workspace_bytes_views(int B, int R, int F, int V) {
  if (B <= 0 || R < 0 || F <= 0 || V < 1 || V > kMaxViews) return 0;
  const long long n = (long long)B * R;
  const int vp = V <= kVP ? kVP : kMaxViews;
  return workspace_small16_offset(n, vp, F) + workspace_small16_bytes(F);
}

extern "C" size_t sailor_render_workspace_bytes(int B, int R, int F) {
  return sailor_render_workspace_bytes_views(B, R, F, kVP);
}

namespace {
int render_stages(int stages, const float* t, const float* rays, int B, int R, int S,
                  const float* Kv, const float* RTv, int V, int img_h, int img_w,
                  const void* fmap, int fmap_precision, int Hf, int Wf, int F,
                  const float* bbox, const float* zdir, int H, int precision,
                  const void* w_dense,
                  const float* b_dense, const float* w_small, float t_thresh, int early_exit,
                  int chunk_samples, float* out_rgb, float* out_depth, float* out_aux,
                  float* out_occ, void* workspace, size_t workspace_bytes, void* stream,
                  const float* feature_scales = nullptr) {
  SAILOR_REQUIRE(B > 0 && R >= 0 && S > 0, "render_rays: bad sizes");
  SAILOR_REQUIRE((long long)B * R < (1ll << 31), "render_rays: at most 2^31-1 rays per call");
  SAILOR_REQUIRE(V >= 1 && V <= kMaxViews, "render_rays: 1 to %d views supported", kMaxViews);
  SAILOR_REQUIRE(workspace != nullptr &&
                     workspace_bytes >= sailor_render_workspace_bytes_views(B, R, F, V),
                 "render_rays: workspace of %zu bytes is smaller than the %zu required",
                 workspace_bytes, sailor_render_workspace_bytes_views(B, R, F, V));
  SAILOR_REQUIRE(Hf > 1 && Wf > 1, "render_rays: feature map too small");
  // launch shape: bits 0-7 samples per chunk, bits 8-15 grid share (percent)
  int ns = chunk_samples & 0xff;
  const int share = (chunk_samples >> 8) & 0xff;
  SAILOR_REQUIRE(chunk_samples >= 0 && share <= 100,
                 "render_rays: launch shape %#x: grid share %d%% above 100", chunk_samples, share);
  if (ns <= 0) {
    if (V > kVP) {
      ns = S % 12 == 0 ? 12 : 8;  // 8-slot instantiations (96 columns, as 24 x 4)
    } else {
      // measured on MI355X: 24-sample chunks -- fp32 hidden-256 at one
      // workgroup per CU (C2 frame; 48 do not fit the LDS), fp32 hidden-128
      // and the 16-bit forms at two (C2 shade 8.25 vs 8.67 ms for 16; C4
      // forward bf16 8.54 vs 11.0 ms, fp32 32.0 vs 35.4 ms for 48); 16 or 8
      // when S is not a multiple of 24
      if (S % 24 == 0) ns = 24;
      else ns = (S % 16 == 0) ? 16 : 8;
    }
  }
  SAILOR_REQUIRE(S % ns == 0, "render_rays: S=%d is not a multiple of the chunk (%d samples)",
                 S, ns);
  if (R == 0) return SAILOR_OK;
  RenderArgs a{};
  SAILOR_REQUIRE(fmap_precision == kF32 || fmap_precision == precision,
                 "render_rays: feature maps must be fp32 or the form's 16-bit type");
  a.t = t; a.rays = rays; a.Kv = Kv; a.RTv = RTv; a.fmap = fmap; a.bbox = bbox; a.zdir = zdir;
  a.fmap_half = fmap_precision != kF32;
  a.wdense = w_dense; a.bdense = b_dense; a.wsmall = w_small;
  a.out_rgb = out_rgb; a.out_depth = out_depth; a.out_aux = out_aux; a.out_occ = out_occ;
  a.n_rays = (long long)B * R; a.rays_per_batch = R;
  char* wsb = static_cast<char*>(workspace);
  unsigned* counter = reinterpret_cast<unsigned*>(wsb);
  int* list = reinterpret_cast<int*>(wsb + workspace_list_offset());
  a.out_tok = reinterpret_cast<float*>(wsb + workspace_token_offset(a.n_rays));
  a.wsmall16 = wsb + workspace_small16_offset(a.n_rays, V <= kVP ? kVP : kMaxViews, F);
  a.list = list;
  a.n_valid = counter;
  a.S = S; a.V = V; a.img_w = img_w; a.img_h = img_h; a.Hf = Hf; a.Wf = Wf;
  a.t_thres = t_thresh; a.early_exit = early_exit & 1; a.compensated = (early_exit >> 1) & 1;
  a.fp32_hydra = (early_exit >> 2) & 1;
  a.fscale = feature_scales;
  a.grid_share = share;
  SAILOR_REQUIRE(feature_scales == nullptr || fmap_precision == precision,
                 "render_rays: feature scales apply to quantised 16-bit maps only");
  a.stages = stages;
  hipStream_t st = sailor::as_stream(stream);
  SAILOR_REQUIRE(stages >= 1 && stages <= 7, "render_rays: stage mask %d out of range", stages);
  if (!((F == 64 && H == 256) || (F == 32 && H == 128))) {
    return sailor::set_error(SAILOR_ERR_ARG, "render_rays: (F=%d, H=%d) not instantiated "
                             "(supported: (32,128), (64,256))", F, H);
  }
  if (precision < kF32 || precision > kF16)
    return sailor::set_error(SAILOR_ERR_ARG, "render_rays: precision %d unsupported", precision);
  if (V <= kVP && ns != 8 && ns != 16 && ns != 24 && ns != 48)
    return sailor::set_error(SAILOR_ERR_ARG, "render_rays: samples per chunk %d not "
                             "instantiated (8, 16, 24, 48)", ns);
  // every configuration error is reported before anything is launched
  {
    const int err = (F == 64) ? dispatch_ns<64, 256>(ns, precision, a, st, true)
                              : dispatch_ns<32, 128>(ns, precision, a, st, true);
    if (err) return err;
  }
  // the counter is cleared by a kernel, not a memset, so the call can be
  // captured into a graph
  if (stages & 1) {
    const long long nb = (a.n_rays + kCompactRays - 1) / kCompactRays;
    int* bcnt = reinterpret_cast<int*>(a.out_tok);   // free until the sample kernel runs
    hipLaunchKernelGGL(compact_count, dim3((unsigned)nb), dim3(256), 0, st, t, S, a.n_rays,
                       bcnt);
    hipLaunchKernelGGL(compact_scan, dim3(1), dim3(256), 0, st, bcnt, (int)nb, counter);
    hipLaunchKernelGGL(compact_write, dim3((unsigned)nb), dim3(256), 0, st, t, S, a.n_rays,
                       bcnt, list, out_rgb, out_depth, out_aux);
    if (int err = sailor::check_launch("render_rays (compact)")) return err;
  }
  if (!(stages & 6)) return SAILOR_OK;
  if (F == 64 && H == 256) return dispatch_ns<64, 256>(ns, precision, a, st, false);
  if (F == 32 && H == 128) return dispatch_ns<32, 128>(ns, precision, a, st, false);
  return sailor::set_error(SAILOR_ERR_ARG, "render_rays: (F=%d, H=%d) not instantiated "
                           "(supported: (32,128), (64,256))", F, H);
}

template <int F, int H>
int hydra_backward_launch(const float* grad_rgb, const float* tok, const int* list,
                          const unsigned* counter, const float* w_small, int V,
                          float* grad_tokens, float* partials, int parts, float* grad_small,
                          hipStream_t st) {
  constexpr SmallOffsets so = small_offsets(F, H);
  const long long nz = (long long)parts * so.total;
  hipLaunchKernelGGL(fill_zero, dim3((unsigned)std::min<long long>((nz + 255) / 256, 4096)),
                     dim3(256), 0, st, partials, nz);
  if (int err = sailor::check_launch("render_backward (zero)")) return err;
  if constexpr (F == 32) {
    // fp32 MFMA, one wave per workgroup, gradients in registers: RGB head,
    // stack 1, stack 0, passing dL/dx on in grad_tokens
    hipLaunchKernelGGL((hydra_rgb_backward_mfma_kernel<F, H, 0>), dim3(parts), dim3(64), 0, st,
                       tok, list, counter, w_small, V, grad_rgb, grad_tokens, partials);
    hipLaunchKernelGGL((hydra_rgb_backward_mfma_kernel<F, H, 1>), dim3(parts), dim3(64), 0, st,
                       tok, list, counter, w_small, V, grad_rgb, grad_tokens, partials);
    hipLaunchKernelGGL((hydra_rgb_backward_mfma_kernel<F, H, 2>), dim3(parts), dim3(64), 0, st,
                       tok, list, counter, w_small, V, grad_rgb, grad_tokens, partials);
  } else {
    typedef HydraBwdRows<F> R;
    auto kern = hydra_rgb_backward_kernel<F, H>;
    static std::atomic<bool> init[sailor::kMaxDevices];   // LDS limit set, per device
    int dev = 0;
    if (int e = sailor::device_slot(&dev)) return e;
    if (!init[dev].load(std::memory_order_acquire)) {
      SAILOR_HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                         hipFuncAttributeMaxDynamicSharedMemorySize,
                                         (int)R::bytes));
      init[dev].store(true, std::memory_order_release);
    }
    hipLaunchKernelGGL(kern, dim3(parts), dim3(kThreads), R::bytes, st, tok, list, counter,
                       w_small, V, grad_rgb, grad_tokens, partials);
  }
  if (int err = sailor::check_launch("render_backward (hydra)")) return err;
  launch_sum_partials(partials, parts, so.total, grad_small, 0, st);
  return sailor::check_launch("render_backward (sum)");
}

}  // namespace

extern "C" size_t sailor_render_backward_partials_floats(int F, int H) {
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    cus = 256;
  return (size_t)cus * (small_offsets(F, H).total + dense_grad_total(F, H));
}

extern "C" int sailor_render_hydra_backward(const float* grad_rgb, int B, int R, int V, int F,
                                            int H, const float* w_small, const void* workspace,
                                            size_t workspace_bytes, float* grad_tokens,
                                            float* grad_small, float* partials,
                                            size_t partials_floats, void* stream) {
  SAILOR_REQUIRE(B > 0 && R >= 0, "render_hydra_backward: bad sizes");
  SAILOR_REQUIRE(V >= 1 && V <= kVP, "render_hydra_backward: 1 to %d views supported", kVP);
  SAILOR_REQUIRE(workspace != nullptr &&
                     workspace_bytes >= sailor_render_workspace_bytes(B, R, F),
                 "render_hydra_backward: workspace too small");
  if (!((F == 64 && H == 256) || (F == 32 && H == 128)))
    return sailor::set_error(SAILOR_ERR_ARG, "render_hydra_backward: (F=%d, H=%d) not "
                             "instantiated", F, H);
  const size_t per = (size_t)small_offsets(F, H).total;
  const int parts = (int)std::min<size_t>(partials_floats / per, 1024);
  SAILOR_REQUIRE(parts >= 1, "render_hydra_backward: partial buffer holds %zu floats, at least "
                 "%zu needed", partials_floats, per);
  const long long n = (long long)B * R;
  if (n == 0) return SAILOR_OK;
  const char* wsb = static_cast<const char*>(workspace);
  const unsigned* counter = reinterpret_cast<const unsigned*>(wsb);
  const int* list = reinterpret_cast<const int*>(wsb + workspace_list_offset());
  const float* tok = reinterpret_cast<const float*>(wsb + workspace_token_offset(n));
  hipStream_t st = sailor::as_stream(stream);
  return F == 64 ? hydra_backward_launch<64, 256>(grad_rgb, tok, list, counter, w_small, V,
                                                  grad_tokens, partials, parts, grad_small, st)
                 : hydra_backward_launch<32, 128>(grad_rgb, tok, list, counter, w_small, V,
                                                  grad_tokens, partials, parts, grad_small, st);
}

namespace {

template <int F, int H, int PREC, bool DW, int NS, int SMAX>
int sample_backward_launch_ns(BackArgs& ba, int parts, float* grad_small, float* grad_dense,
                              hipStream_t st) {
  constexpr SmallOffsets so = small_offsets(F, H);
  typedef SmemB<F, H, PREC, NS, SMAX> L;
  static_assert(L::total <= kLdsBytes, "backward LDS");
  SAILOR_REQUIRE(ba.r.S % NS == 0 && ba.r.S <= SMAX,
                 "render_sample_backward: S=%d must be a multiple of %d and at most %d",
                 ba.r.S, NS, SMAX);
  auto kern = render_backward_kernel<F, H, PREC, NS, DW, SMAX>;
  static std::atomic<bool> init[sailor::kMaxDevices];   // LDS limit set, per device
  int dev = 0;
  if (int e = sailor::device_slot(&dev)) return e;
  if (!init[dev].load(std::memory_order_acquire)) {
    SAILOR_HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                       hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)L::total));
    init[dev].store(true, std::memory_order_release);
  }
  const long long nz = (long long)parts * so.total;
  hipLaunchKernelGGL(fill_zero, dim3((unsigned)std::min<long long>((nz + 255) / 256, 4096)),
                     dim3(256), 0, st, ba.gpart, nz);
  if (int err = sailor::check_launch("render_sample_backward (zero)")) return err;
  hipLaunchKernelGGL(kern, dim3(parts), dim3(kThreads), L::total, st, ba);
  if (int err = sailor::check_launch("render_sample_backward")) return err;
  launch_sum_partials(ba.gpart, parts, so.total, grad_small, 1, st);
  if (int err = sailor::check_launch("render_sample_backward (sum)")) return err;
  if constexpr (DW) {
    constexpr int nd = dense_grad_total(F, H);
    launch_sum_partials(ba.gdense, parts, nd, grad_dense, 0, st);
    return sailor::check_launch("render_sample_backward (dense sum)");
  }
  return SAILOR_OK;
}

// Samples per chunk: fp32 with in-kernel weight gradients 8 (measured 9 %
// faster per C4 step than 4 despite ~50 spilled registers), fp32 column
// form 4; 16-bit with in-kernel weight gradients 16 (24 would cut the chunk
// phases by a third but spills ~120 registers per lane beside the
// weight-gradient accumulators and runs slower); 16-bit column form 8.
template <int F, int H, int PREC, bool DW>
int sample_backward_launch(BackArgs& ba, int parts, float* grad_small, float* grad_dense,
                           hipStream_t st) {
  constexpr int NS = PREC == kF32 ? (DW ? 8 : 4) : (DW ? 16 : 8);
  return sample_backward_launch_ns<F, H, PREC, DW, NS, kBackMaxS>(ba, parts, grad_small,
                                                                  grad_dense, st);
}

// Shared argument checks and dispatch of the two backward entry points:
// column form (col_bufs given) or in-kernel weight gradients (grad_dense).
int sample_backward_impl(const float* t, const float* rays, int B, int R, int S,
                         const float* Kv, const float* RTv, int V, int img_h, int img_w,
                         const float* fmap, int Hf, int Wf, int F, const float* bbox, int H,
                         int precision, const void* w_dense, const void* w_dense_t,
                         const float* b_dense, const float* w_small, float t_thresh,
                         const float* occ, const float* grad_tokens, const void* workspace,
                         size_t workspace_bytes, long long item_begin, long long item_end,
                         void* const* col_bufs, float* grad_dense, float* grad_small,
                         float* partials, size_t partials_floats, void* stream) {
  const bool dw = col_bufs == nullptr;
  SAILOR_REQUIRE(B > 0 && R >= 0 && S > 0, "render_sample_backward: bad sizes");
  SAILOR_REQUIRE(V >= 1 && V <= kVP, "render_sample_backward: 1 to %d views supported", kVP);
  SAILOR_REQUIRE(workspace != nullptr &&
                     workspace_bytes >= sailor_render_workspace_bytes(B, R, F),
                 "render_sample_backward: workspace too small");
  SAILOR_REQUIRE(item_begin >= 0 && item_end >= item_begin,
                 "render_sample_backward: bad item range");
  SAILOR_REQUIRE(!dw || grad_dense != nullptr, "render_sample_backward: no gradient output");
  if (!((F == 64 && H == 256) || (F == 32 && H == 128)))
    return sailor::set_error(SAILOR_ERR_ARG, "render_sample_backward: (F=%d, H=%d) not "
                             "instantiated", F, H);
  if (dw && H != 128)
    return sailor::set_error(SAILOR_ERR_ARG, "render_sample_backward_fused: in-kernel weight "
                             "gradients need H = 128 (got %d); use the column form", H);
  if (precision < kF32 || precision > kF16)
    return sailor::set_error(SAILOR_ERR_ARG, "render_sample_backward: precision %d unsupported",
                             precision);
  const size_t per = (size_t)small_offsets(F, H).total + (dw ? dense_grad_total(F, H) : 0);
  const int parts = (int)std::min<size_t>(partials_floats / per, 1024);
  SAILOR_REQUIRE(parts >= 1, "render_sample_backward: partial buffer too small");
  const long long n = (long long)B * R;
  BackArgs ba{};
  ba.gpart = partials;
  ba.gdense = partials + (size_t)parts * small_offsets(F, H).total;
  hipStream_t st = sailor::as_stream(stream);
  if (n == 0 || item_end == item_begin) {
    if (!dw) return SAILOR_OK;
    // no columns: the dense gradient is zero
    const long long nd = dense_grad_total(F, H);
    hipLaunchKernelGGL(fill_zero, dim3((unsigned)((nd + 255) / 256)), dim3(256), 0, st,
                       grad_dense, nd);
    return sailor::check_launch("render_sample_backward (zero)");
  }
  RenderArgs& a = ba.r;
  a.t = t; a.rays = rays; a.Kv = Kv; a.RTv = RTv; a.fmap = fmap; a.bbox = bbox;
  a.wdense = w_dense; a.bdense = b_dense; a.wsmall = w_small;
  a.n_rays = n; a.rays_per_batch = R;
  const char* wsb = static_cast<const char*>(workspace);
  a.n_valid = reinterpret_cast<const unsigned*>(wsb);
  a.list = reinterpret_cast<const int*>(wsb + workspace_list_offset());
  a.S = S; a.V = V; a.img_w = img_w; a.img_h = img_h; a.Hf = Hf; a.Wf = Wf;
  a.t_thres = t_thresh;
  ba.wdense_t = w_dense_t;
  ba.occ = occ;
  ba.dtok = grad_tokens;
  ba.item_begin = item_begin;
  ba.item_end = item_end;
  if (!dw) {
    ba.A0 = col_bufs[0]; ba.A1 = col_bufs[1]; ba.A2 = col_bufs[2]; ba.A3 = col_bufs[3];
    ba.D0 = col_bufs[4]; ba.D1 = col_bufs[5]; ba.D2 = col_bufs[6]; ba.D3 = col_bufs[7];
  }
  if (dw) {
    switch (precision) {
      case kF32: return sample_backward_launch<32, 128, kF32, true>(ba, parts, grad_small,
                                                                    grad_dense, st);
      case kBF16: return sample_backward_launch<32, 128, kBF16, true>(ba, parts, grad_small,
                                                                      grad_dense, st);
      default: return sample_backward_launch<32, 128, kF16, true>(ba, parts, grad_small,
                                                                  grad_dense, st);
    }
  }
  if (F == 64) {
    switch (precision) {
      case kF32: return sample_backward_launch<64, 256, kF32, false>(ba, parts, grad_small,
                                                                     nullptr, st);
      case kBF16: return sample_backward_launch<64, 256, kBF16, false>(ba, parts, grad_small,
                                                                       nullptr, st);
      default: return sample_backward_launch<64, 256, kF16, false>(ba, parts, grad_small,
                                                                   nullptr, st);
    }
  }
  switch (precision) {
    case kF32: return sample_backward_launch<32, 128, kF32, false>(ba, parts, grad_small,
                                                                   nullptr, st);
    case kBF16: return sample_backward_launch<32, 128, kBF16, false>(ba, parts, grad_small,
                                                                     nullptr, st);
    default: return sample_backward_launch<32, 128, kF16, false>(ba, parts, grad_small,
                                                                 nullptr, st);
  }
}

}  // namespace

extern "C" size_t sailor_render_dense_grad_floats(int F, int H) {
  return (size_t)dense_grad_total(F, H);
}

extern "C" int sailor_render_sample_backward(
    const float* t, const float* rays, int B, int R, int S, const float* Kv, const float* RTv,
    int V, int img_h, int img_w, const float* fmap, int Hf, int Wf, int F, const float* bbox,
    int H, int precision, const void* w_dense, const void* w_dense_t, const float* b_dense,
    const float* w_small, float t_thresh, const float* occ, const float* grad_tokens,
    const void* workspace, size_t workspace_bytes, long long item_begin, long long item_end,
    void* const* col_bufs, float* grad_small, float* partials, size_t partials_floats,
    void* stream) {
  SAILOR_REQUIRE(col_bufs != nullptr, "render_sample_backward: no column buffers");
  for (int i = 0; i < 8; ++i)
    SAILOR_REQUIRE(col_bufs[i] != nullptr, "render_sample_backward: column buffer %d is null", i);
  return sample_backward_impl(t, rays, B, R, S, Kv, RTv, V, img_h, img_w, fmap, Hf, Wf, F, bbox,
                              H, precision, w_dense, w_dense_t, b_dense, w_small, t_thresh, occ,
                              grad_tokens, workspace, workspace_bytes, item_begin, item_end,
                              col_bufs, nullptr, grad_small, partials, partials_floats, stream);
}

extern "C" int sailor_render_sample_backward_fused(
    const float* t, const float* rays, int B, int R, int S, const float* Kv, const float* RTv,
    int V, int img_h, int img_w, const float* fmap, int Hf, int Wf, int F, const float* bbox,
    int H, int precision, const void* w_dense, const void* w_dense_t, const float* b_dense,
    const float* w_small, float t_thresh, const float* occ, const float* grad_tokens,
    const void* workspace, size_t workspace_bytes, float* grad_dense, float* grad_small,
    float* partials, size_t partials_floats, void* stream) {
  return sample_backward_impl(t, rays, B, R, S, Kv, RTv, V, img_h, img_w, fmap, Hf, Wf, F, bbox,
                              H, precision, w_dense, w_dense_t, b_dense, w_small, t_thresh, occ,
                              grad_tokens, workspace, workspace_bytes, 0, (long long)B * R,
                              nullptr, grad_dense, grad_small, partials, partials_floats, stream);
}

extern "C" int sailor_density_points(const float* points, int B, int N, const float* Kv,
                                     const float* RTv, int V, int img_h, int img_w,
                                     const void* fmap, int fmap_precision, int Hf, int Wf,
                                     int F, const float* bbox, int H, int precision,
                                     const void* w_dense, const float* b_dense,
                                     const float* w_small, int chunk_samples,
                                     float* out_density, void* stream) {
  SAILOR_REQUIRE(B > 0 && N >= 0, "density_points: bad sizes");
  SAILOR_REQUIRE(V >= 1 && V <= kVP, "density_points: 1 to %d views supported", kVP);
  SAILOR_REQUIRE(Hf > 1 && Wf > 1, "density_points: feature map too small");
  SAILOR_REQUIRE(fmap_precision == kF32 || fmap_precision == precision,
                 "density_points: feature maps must be fp32 or the form's 16-bit type");
  if (!((F == 64 && H == 256) || (F == 32 && H == 128))) {
    return sailor::set_error(SAILOR_ERR_ARG, "density_points: (F=%d, H=%d) not instantiated "
                             "(supported: (32,128), (64,256))", F, H);
  }
  if (precision < kF32 || precision > kF16)
    return sailor::set_error(SAILOR_ERR_ARG, "density_points: precision %d unsupported",
                             precision);
  int ns = chunk_samples > 0 ? chunk_samples : (precision == kF32 ? 24 : 16);
  if (ns != 8 && ns != 16 && ns != 24 && ns != 48)
    return sailor::set_error(SAILOR_ERR_ARG, "density_points: points per group %d not "
                             "instantiated (8, 16, 24, 48)", ns);
  if (N == 0) return SAILOR_OK;
  RenderArgs a{};
  a.points = points;
  a.n_points = N;
  a.rays_per_batch = (N + ns - 1) / ns;
  a.n_rays = (long long)B * a.rays_per_batch;
  a.Kv = Kv; a.RTv = RTv; a.fmap = fmap; a.fmap_half = fmap_precision != kF32; a.bbox = bbox;
  a.wdense = w_dense; a.bdense = b_dense; a.wsmall = w_small;
  a.out_occ = out_density;
  a.S = ns; a.V = V; a.img_w = img_w; a.img_h = img_h; a.Hf = Hf; a.Wf = Wf;
  a.stages = 2;
  hipStream_t st = sailor::as_stream(stream);
  {
    const int err = (F == 64) ? dispatch_ns<64, 256>(ns, precision, a, st, true)
                              : dispatch_ns<32, 128>(ns, precision, a, st, true);
    if (err) return err;
  }
  return (F == 64) ? dispatch_ns<64, 256>(ns, precision, a, st, false)
                   : dispatch_ns<32, 128>(ns, precision, a, st, false);
}

extern "C" int sailor_render_rays(const float* t, const float* rays, int B, int R, int S,
                                  const float* Kv, const float* RTv, int V, int img_h,
                                  int img_w, const void* fmap, int fmap_precision, int Hf,
                                  int Wf, int F, const float* bbox, const float* zdir, int H,
                                  int precision, const void* w_dense, const float* b_dense,
                                  const float* w_small, float t_thresh, int early_exit,
                                  int chunk_samples, float* out_rgb, float* out_depth,
                                  float* out_aux, float* out_occ, void* workspace,
                                  size_t workspace_bytes, void* stream) {
  return render_stages(7, t, rays, B, R, S, Kv, RTv, V, img_h, img_w, fmap, fmap_precision, Hf,
                       Wf, F, bbox, zdir, H, precision, w_dense, b_dense, w_small, t_thresh,
                       early_exit, chunk_samples, out_rgb, out_depth, out_aux, out_occ,
                       workspace, workspace_bytes, stream);
}

extern "C" int sailor_render_rays_stages(int stages, const float* t, const float* rays, int B,
                                         int R, int S, const float* Kv, const float* RTv, int V,
                                         int img_h, int img_w, const void* fmap,
                                         int fmap_precision, int Hf, int Wf, int F,
                                         const float* bbox, const float* zdir, int H,
                                         int precision, const void* w_dense,
                                         const float* b_dense, const float* w_small,
                                         float t_thresh, int early_exit, int chunk_samples,
                                         float* out_rgb, float* out_depth, float* out_aux,
                                         float* out_occ, void* workspace, size_t workspace_bytes,
                                         void* stream) {
  return render_stages(stages, t, rays, B, R, S, Kv, RTv, V, img_h, img_w, fmap,
                       fmap_precision, Hf, Wf, F, bbox, zdir, H, precision, w_dense, b_dense,
                       w_small, t_thresh, early_exit, chunk_samples, out_rgb, out_depth,
                       out_aux, out_occ, workspace, workspace_bytes, stream);
}

extern "C" int sailor_render_rays_quantized(
    int stages, const float* t, const float* rays, int B, int R, int S, const float* Kv,
    const float* RTv, int V, int img_h, int img_w, const void* fmap, int Hf, int Wf, int F,
    const float* feature_scales, const float* bbox, const float* zdir, int H, int precision,
    const void* w_dense, const float* b_dense, const float* w_small, float t_thresh, int flags,
    int chunk_samples, float* out_rgb, float* out_depth, float* out_aux, float* out_occ,
    void* workspace, size_t workspace_bytes, void* stream) {
  SAILOR_REQUIRE(precision != kF32, "render_rays_quantized: a 16-bit form is required");
  SAILOR_REQUIRE(feature_scales != nullptr, "render_rays_quantized: feature scales missing");
  return render_stages(stages, t, rays, B, R, S, Kv, RTv, V, img_h, img_w, fmap, precision, Hf,
                       Wf, F, bbox, zdir, H, precision, w_dense, b_dense, w_small, t_thresh,
                       flags, chunk_samples, out_rgb, out_depth, out_aux, out_occ, workspace,
                       workspace_bytes, stream, feature_scales);
}
