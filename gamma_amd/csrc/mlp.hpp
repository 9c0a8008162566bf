// mlp.hpp -- learned initial pressure for cons2prim: a 3 -> 64 -> 64 -> 1 MLP.
//
// Inputs (log D, |S|/D, tau/D) of a conserved state enter through fixed maps
// (log D, log1p(|S|/D), log of the non-kinetic part of tau/D; see features())
// and a per-feature affine normalisation;
// the output is the normalised log(p/D), so log p = log D + y/out_scale +
// out_mean (the map p/D <-> (|S|/D, tau/D) is scale free).  Hidden layers use
// ReLU.  Parameters are one float array (layout below); inference runs the
// 64x64 hidden layer as BF16 MFMA (v_mfma_f32_32x32x16_bf16, FP32 accumulate)
// over the 64 cells of a wavefront, training runs FP32 forward/backward.
#pragma once

namespace gh {
namespace mlp {

constexpr int H = 64;
// parameter layout (floats)
constexpr int OW1 = 0;                 // W1[64][3]
constexpr int OB1 = OW1 + H * 3;       // b1[64]
constexpr int OW2 = OB1 + H;           // W2[64 out][64 in]
constexpr int OB2 = OW2 + H * H;       // b2[64]
constexpr int OW3 = OB2 + H;           // w3[64]
constexpr int OB3 = OW3 + H;           // b3
constexpr int ONORM = OB3 + 1;         // in_mean[3], in_scale[3], out_mean, out_scale
constexpr int NPARAM = ONORM + 8;
constexpr int NTRAIN = ONORM;          // trained entries (the normalisation is fixed)

// features of a conserved state (D, |S|, tau), normalised: log D, log1p(q)
// and log e with q = |S|/D, t = tau/D and e = t + 1 - sqrt(1 + q^2), the part
// of tau/D that is not bulk kinetic energy.  For a cold gas e ~ (k - 1)(p/rho)/W
// with k = G/(G-1) and W the Lorentz factor, so log e carries log p where t and
// q alone hold it only in digits an FP32 input cannot keep.  e is formed in
// FP64 and floored at the rounding level of t.
GH_HD void features(const float* w, double D, double S, double tau, float* x) {
  double Dp = D > 1e-300 ? D : 1e-300;
  double q = S / Dp, t = tau / Dp;
  double e = t + 1. - sqrt(1. + q * q);
  double fl = 1e-16 * (1. + fabs(t));
  if (!(e > fl)) e = fl;
  double f0 = log(Dp), f1 = log1p(q), f2 = log(e);
  x[0] = (float)((f0 - w[ONORM + 0]) * w[ONORM + 3]);
  x[1] = (float)((f1 - w[ONORM + 1]) * w[ONORM + 4]);
  x[2] = (float)((f2 - w[ONORM + 2]) * w[ONORM + 5]);
}

// network output -> pressure guess
GH_HD double pressure_of(const float* w, double D, float y) {
  double lpd = (double)y / w[ONORM + 7] + w[ONORM + 6];
  return D * exp(lpd);
}

// FP32 forward of one sample; keeps the hidden activations for backprop
GH_HD float forward_f32(const float* w, const float* x, float* h1, float* h2) {
  for (int o = 0; o < H; ++o) {
    float a = w[OB1 + o];
    for (int k = 0; k < 3; ++k) a += w[OW1 + o * 3 + k] * x[k];
    h1[o] = a > 0.f ? a : 0.f;
  }
  for (int o = 0; o < H; ++o) {
    float a = w[OB2 + o];
    for (int k = 0; k < H; ++k) a += w[OW2 + o * H + k] * h1[k];
    h2[o] = a > 0.f ? a : 0.f;
  }
  float y = w[OB3];
  for (int k = 0; k < H; ++k) y += w[OW3 + k] * h2[k];
  return y;
}

}  // namespace mlp
}  // namespace gh

#ifndef GH_EMU
namespace gh {
namespace mlp {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

// Lane layout of v_mfma_f32_32x32x16_bf16 (checked on the device by
// tests/test_mlp.py::test_mfma_lane_layout):
//   A 32x16: lane l holds A[l%32][8*(l/32) + e], e = 0..7
//   B 16x32: lane l holds B[8*(l/32) + e][l%32]
//   D 32x32: lane l, register r holds D[8*(r/4) + 4*(l/32) + r%4][l%32]
__device__ __forceinline__ int d_row(int l, int r) { return 8 * (r / 4) + 4 * (l / 32) + (r % 4); }

// B operands of the hidden layer, pre-arranged per lane: for k-block kb (0..3)
// and n-block nb (0..1), lane l holds W2[nb*32 + l%32][kb*16 + 8*(l/32) + e]
// (B[k][n] = W2[n][k]).  Index: ((kb*2 + nb)*64 + l)*8 + e.
__host__ inline void pack_w2(const float* w, __bf16* out) {
  for (int kb = 0; kb < 4; ++kb)
    for (int nb = 0; nb < 2; ++nb)
      for (int l = 0; l < 64; ++l)
        for (int e = 0; e < 8; ++e) {
          int n = nb * 32 + l % 32, k = kb * 16 + 8 * (l / 32) + e;
          out[((kb * 2 + nb) * 64 + l) * 8 + e] = (__bf16)w[OW2 + n * H + k];
        }
}

// Forward of the 64 cells of a wavefront (one per lane).  Must be reached by
// all 64 lanes together (uniform control flow): lanes without a cell pass any
// finite features and ignore the result.  lds: 64*64 bf16 scratch per wave
// (H1 rows), rowsum: 64 floats per wave.
__device__ float forward_wave(const float* __restrict__ w, const bf16x8* __restrict__ w2b,
                              const float* x, __bf16* lds, float* rowsum) {
  int l = threadIdx.x & 63;
  // layer 1 on the vector ALU, written as bf16 rows of H1 (row = lane)
  for (int o = 0; o < H; ++o) {
    float a = w[OB1 + o] + w[OW1 + o * 3] * x[0] + w[OW1 + o * 3 + 1] * x[1] + w[OW1 + o * 3 + 2] * x[2];
    lds[l * H + o] = (__bf16)(a > 0.f ? a : 0.f);
  }
  __builtin_amdgcn_s_waitcnt(0);
  __builtin_amdgcn_wave_barrier();
  // hidden layer: D[rb][nb] = sum_kb A[rb][kb] B[kb][nb], 16 MFMA per wave
  floatx16 acc[2][2];
  for (int rb = 0; rb < 2; ++rb)
    for (int nb = 0; nb < 2; ++nb) acc[rb][nb] = floatx16{};
  for (int kb = 0; kb < 4; ++kb) {
    for (int rb = 0; rb < 2; ++rb) {
      bf16x8 a;
      const __bf16* src = lds + (rb * 32 + l % 32) * H + kb * 16 + 8 * (l / 32);
      for (int e = 0; e < 8; ++e) a[e] = src[e];
      for (int nb = 0; nb < 2; ++nb)
        acc[rb][nb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, w2b[(kb * 2 + nb) * 64 + l], acc[rb][nb], 0, 0, 0);
    }
  }
  // bias, ReLU and the output layer: per-lane partial row sums, then a reduction
  // over the 32 lanes that share the same rows
  for (int rb = 0; rb < 2; ++rb) {
    float part[16];
    for (int r = 0; r < 16; ++r) part[r] = 0.f;
    for (int nb = 0; nb < 2; ++nb) {
      int col = nb * 32 + l % 32;
      float b2 = w[OB2 + col], w3 = w[OW3 + col];
      for (int r = 0; r < 16; ++r) {
        float v = acc[rb][nb][r] + b2;
        part[r] += w3 * (v > 0.f ? v : 0.f);
      }
    }
    for (int r = 0; r < 16; ++r)
      for (int off = 16; off > 0; off >>= 1) part[r] += __shfl_xor(part[r], off, 64);
    if (l % 32 == 0)
      for (int r = 0; r < 16; ++r) rowsum[rb * 32 + d_row(l, r)] = part[r];
  }
  __builtin_amdgcn_s_waitcnt(0);
  __builtin_amdgcn_wave_barrier();
  return rowsum[l] + w[OB3];
}

}  // namespace mlp
}  // namespace gh
#endif
