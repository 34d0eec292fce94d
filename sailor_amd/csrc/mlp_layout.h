// Layout contract between the host-side weight packer (sailor_amd/model.py)
// and the fused render kernel (mlp.hip).  Shared constants only.
//
// MFMA operand layouts (gfx950, wave64):
//  fp32 v_mfma_f32_16x16x4f32   A: lane l holds A[l%16][l/16]
//                               B: lane l holds B[l/16][l%16]
//  bf16/fp16 v_mfma_f32_16x16x32_{bf16,f16}
//                               A: lane l holds A[l%16][8*(l/16)+j], j<8
//                               B: lane l holds B[8*(l/16)+j][l%16]
//  D (both): lane l holds D[4*(l/16)+i][l%16], i<4.
//
// The network runs transposed: MFMA columns (N) are (sample, view) pairs of
// one ray, rows (M) are output units.  A workgroup of 4 waves owns one ray;
// wave w computes output units [w*H/4, (w+1)*H/4) for every column of a
// chunk of samples.  Within a wave, A row m of m-tile mt computes unit
// w*H/4 + 4*MTW*(m/4) + 4*mt + m%4 (MTW = H/64 m-tiles per wave), so lane
// group g's accumulators hold 4*MTW consecutive units of its column.
// Layer inputs live in LDS as act[column][k] in natural k order (first
// layer: k < F feature channels, F <= k < F+63 positional encoding,
// k = F+63 zero padding; physically the 16-byte chunks of a column are
// XOR-swizzled, which the packer does not see); B fragments are read from
// there, so the reduction index of MFMA step s, lane group g (= l/16), slot
// j is
//   fp32:  k = 16*(s/4) + 4*g + s%4      (one 16-byte read serves 4 steps)
//   bf16:  k = 32*s + 8*g + j
// The packer stores, for every layer and wave w, the A fragments as
// [w][step][lane][m-tile][slot] so one lane's values of a step are
// contiguous.
#pragma once

namespace sailor {
namespace mlp {

constexpr int kPeSlots = 64;   // 63 PE dims + 1 zero pad
constexpr int kHydraInner = 32;
constexpr int kRgbH1 = 64;
constexpr int kRgbH2 = 32;
constexpr int kHydraStacks = 2;
constexpr int kWaves = 4;      // waves per workgroup (they split the rows)

// Offsets (in floats) into the small-weight buffer for feature width F and
// hidden width H.
struct SmallOffsets {
  int sigma_w, sigma_b;
  int q[kHydraStacks], k[kHydraStacks], v[kHydraStacks], fc[kHydraStacks];
  int w1[kHydraStacks], b1[kHydraStacks], w2[kHydraStacks], b2[kHydraStacks];
  int r1w, r1b, r2w, r2b, r3w, r3b;
  int total;
};

__host__ __device__ constexpr SmallOffsets small_offsets(int F, int H) {
  SmallOffsets o{};
  int p = 0;
  o.sigma_w = p; p += H;
  o.sigma_b = p; p += 4;
  for (int s = 0; s < kHydraStacks; ++s) {
    o.q[s] = p; p += kHydraInner * F;
    o.k[s] = p; p += kHydraInner * F;
    o.v[s] = p; p += kHydraInner * F;
    o.fc[s] = p; p += F * kHydraInner;
    o.w1[s] = p; p += kHydraInner * F;
    o.b1[s] = p; p += kHydraInner;
    o.w2[s] = p; p += F * kHydraInner;
    o.b2[s] = p; p += F;
  }
  o.r1w = p; p += kRgbH1 * F;
  o.r1b = p; p += kRgbH1;
  o.r2w = p; p += kRgbH2 * kRgbH1;
  o.r2b = p; p += kRgbH2;
  o.r3w = p; p += 3 * kRgbH2;
  o.r3b = p; p += 4;
  o.total = p;
  return o;
}

}  // namespace mlp
}  // namespace sailor
