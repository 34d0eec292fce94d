// Layout contract between the host-side weight packer and the fused render
// kernel (mlp.hip).  Shared constants only; no code.
//
// MFMA operand layouts used (gfx950, wave64):
//  fp32 v_mfma_f32_16x16x4f32   A: lane l holds A[l%16][l/16]
//                               B: lane l holds B[l/16][l%16]
//  bf16/fp16 v_mfma_f32_16x16x32_{bf16,f16}
//                               A: lane l holds A[l%16][8*(l/16)+j], j<8
//                               B: lane l holds B[8*(l/16)+j][l%16]
//  D (both): lane l holds D[4*(l/16)+i][l%16], i<4.
//
// The network is evaluated "transposed": columns (the MFMA N dimension) are
// (sample, view) pairs, rows (M) are output units.  Column n of an N-tile is
// sample n/4 (of four) and view n%4.  A layer's D tiles are then directly the
// next layer's B fragments once the reduction index is permuted: for hidden
// layers the k index of fragment slot (step s, lane group g, slot j) is
//   fp32:  s = 4t + i           -> k = 16t + 4g + i
//   bf16:  s, j<4 / j>=4        -> k = 32s + 4g + j  /  32s + 16 + 4g + (j-4)
// and for the first layer (inputs [pe(63) | feat(F)], padded to 64 + F):
//   lane group g owns feature channels [g*F/4, (g+1)*F/4) and PE dims
//   [16g, 16g+16);
//   fp32: s < F/4  -> feature channel g*F/4 + s;  else PE dim 16g + s - F/4
//   bf16: s < F/32 -> feature channel g*F/4 + 8s + j;
//         else PE dim 16g + 8(s - F/32) + j.
// The packer (sailor_amd/model.py) writes fragment (layer, step, lane, m-tile)
// so that one lane's A values for one step are contiguous.
#pragma once

namespace sailor {
namespace mlp {

constexpr int kPeSlots = 64;   // 63 PE dims + 1 zero pad
constexpr int kHydraInner = 32;
constexpr int kRgbH1 = 64;
constexpr int kRgbH2 = 32;
constexpr int kHydraStacks = 2;

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
