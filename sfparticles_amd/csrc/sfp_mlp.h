/*
 * sfp_mlp.h -- the learned QED models, MLPs NIN -> 64 -> 64 -> 64 -> 1 with
 * SiLU:
 *  - the photon-energy sampler (NIN = 2): (log10 chi, r) -> log delta, which
 *    replaces the optical-depth inverse-CDF lookup (qed/optical_depth_tables.py:
 *    59-101) for photon emission;
 *  - the spin-resolved emission rate (NIN = 3): (log10 chi, r1, sigma) ->
 *    log(f + SFP_RATE_EPS), f = 3 r1^2 w(chi, r1^3, sigma) / (alpha m c^2 /
 *    hbar) the rejection-sampling acceptance density, in place of the rate
 *    table lookup.
 * The two 64 x 64 layers run on the matrix cores (v_mfma_f32_16x16x32_bf16:
 * BF16 inputs, FP32 accumulation) in tiles of 16 samples per wave; the
 * NIN -> 64 and 64 -> 1 layers are plain FP32.
 *
 * Parameters, FP32 master copy, flat (SFP_MLP_NPARAM floats for NIN = 2,
 * SFP_MLP_NPARAM_RATE for NIN = 3):
 *   W1[64][NIN] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] W4[64] b4
 * Inputs are normalised x0 = (log10 chi - 0.5 (lc_min + lc_max)) /
 * (0.5 (lc_max - lc_min)), x1 = 2 r - 1, x2 = sigma; the output is
 * (target - mu) / sd.
 *
 * MFMA operand layout (16 x 16 x 32, 64 lanes, lane = 16 g + c, g = 0..3):
 *   A (16 x 32): lane holds A[c][8 g .. 8 g + 7]
 *   B (32 x 16): lane holds B[8 g .. 8 g + 7][c]
 *   C (16 x 16): lane holds C[4 g .. 4 g + 3][c]
 * (checked against numpy by the forward-pass test).
 */
#ifndef SFP_MLP_H
#define SFP_MLP_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/sfp.h"
#include "sfp_device.h"

// offsets of the flat parameter blocks of a network with NIN inputs
template <int NIN>
struct MlpOff {
  static constexpr int W1 = 0, B1 = 64 * NIN, W2 = B1 + 64, B2 = W2 + 4096, W3 = B2 + 64,
                       B3 = W3 + 4096, W4 = B3 + 64, B4 = W4 + 64, N = B4 + 1;
};
static_assert(SFP_MLP_NPARAM == MlpOff<2>::N, "parameter count of the 2-64-64-64-1 network");
static_assert(SFP_MLP_NPARAM_RATE == MlpOff<3>::N, "parameter count of the 3-64-64-64-1 network");

typedef __bf16 sfp_bf16x8 __attribute__((ext_vector_type(8)));
typedef float sfp_f32x4 __attribute__((ext_vector_type(4)));

/* Device copy used by the kernels: the two square layers in BF16 (row-major
 * [out][in]), everything else FP32, plus the normalisation constants.  The
 * first layer is stored as W1[64][2] for the first two inputs and W1c[64]
 * for the third (0 for NIN = 2). */
typedef struct {
  __bf16 W2[64 * 64];
  __bf16 W3[64 * 64];
  __bf16 W2p[64 * 64];  // input columns permuted to the register order of sfp_mlp_eval16
  __bf16 W3p[64 * 64];
  float W1[128], W1c[64], b1[64], b2[64], b3[64], W4[64], b4;
  float x0_shift, x0_scale, out_mu, out_sd;
  float pad[3];
} sfp_mlp_dev;

/* Per-chi-bin table fallback: bins of log10 chi over [lc_min, lc_max]; a set
 * bit selects the inverse-CDF table for that bin. */
#define SFP_MLP_BINS 64
typedef struct {
  float lc_min, lc_max;
  uint64_t fallback;  // bit i: bin i uses the table
} sfp_mlp_bins;

/* SiLU z sigma(z) and its derivative, with the hardware exp2 and reciprocal
 * (about 1 ulp each; the products they feed are rounded to BF16) */
__device__ inline float sfp_sigmoid(float z) {
  return __builtin_amdgcn_rcpf(1.0f + __expf(-z));
}
__device__ inline float sfp_silu(float z) { return z * sfp_sigmoid(z); }
__device__ inline float sfp_dsilu(float z) {
  const float s = sfp_sigmoid(z);
  return s * (1.0f + z * (1.0f - s));
}

/* One 16 x 64 by 64 x 64 product on the matrix cores: out[r][n] =
 * sum_k act(in[r][k]) W[n][k] + b[n] for the wave's 16 rows.  `in` is a
 * 16 x 64 FP32 tile in LDS (pre-activations when SILU_IN, else values), W a
 * BF16 64 x 64 matrix in LDS; the result is written to `out` (16 x 64 FP32,
 * may alias `in`: all reads happen before the first write). */
template <bool SILU_IN>
__device__ inline void sfp_mlp_layer16(const float* in, const __bf16* W, const float* b,
                                       float* out) {
  const int l = __lane_id();
  const int c = l & 15, g = l >> 4;
  sfp_bf16x8 a[2];
#pragma unroll
  for (int s = 0; s < 2; ++s) {
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      float v = in[c * 64 + 32 * s + 8 * g + e];
      if (SILU_IN) v = sfp_silu(v);
      a[s][e] = (__bf16)v;
    }
  }
  sfp_f32x4 acc[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    acc[t] = (sfp_f32x4){0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      const sfp_bf16x8 bv = *(const sfp_bf16x8*)(W + (16 * t + c) * 64 + 32 * s + 8 * g);
      acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[s], bv, acc[t], 0, 0, 0);
    }
  }
  __builtin_amdgcn_wave_barrier();
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int e = 0; e < 4; ++e) out[(4 * g + e) * 64 + 16 * t + c] = acc[t][e] + b[16 * t + c];
  __builtin_amdgcn_wave_barrier();
}

/* Forward pass of 16 samples by one wave.  x0[r], x1[r] (and x2[r] for
 * NIN = 3) are the normalised inputs of row r (rows >= n are padding); z1, z2
 * and z3 (16 x 64 FP32 each, LDS) receive the pre-activations; returns the
 * normalised output of row (lane & 15) in every lane. */
template <int NIN = 2>
__device__ inline float sfp_mlp_forward16(const sfp_mlp_dev* M, const float* x0, const float* x1,
                                          const float* x2, float* z1, float* z2, float* z3) {
  const int l = __lane_id();
  // layer 1 (NIN -> 64) on the vector ALU: lane l computes units l of all rows
  for (int r = 0; r < 16; ++r) {
    float v = M->W1[2 * l] * x0[r] + M->W1[2 * l + 1] * x1[r];
    if (NIN == 3) v += M->W1c[l] * x2[r];
    z1[r * 64 + l] = v + M->b1[l];
  }
  __builtin_amdgcn_wave_barrier();
  sfp_mlp_layer16<true>(z1, M->W2, M->b2, z2);
  sfp_mlp_layer16<true>(z2, M->W3, M->b3, z3);
  // layer 4 (64 -> 1): 4 lanes per row, 16 units each
  const int r = l & 15, q = l >> 4;
  float s = 0.0f;
  for (int k = 16 * q; k < 16 * q + 16; ++k) s += sfp_silu(z3[r * 64 + k]) * M->W4[k];
  s += __shfl_xor(s, 16);
  s += __shfl_xor(s, 32);
  return s + M->b4;
}

/* Register order of the hidden units in sfp_mlp_eval16: lane (g, c) holds
 * units 16 t + 4 g + e (t, e = 0..3), and k-step s of the next product takes
 * position 8 g + e' of its 32 inputs from unit
 *   u(s, g, e') = 16 (2 s + e' / 4) + 4 g + e' % 4.
 * Wp[n][32 s + 8 g + e'] = W[n][u(s, g, e')]. */
SFP_HD int sfp_mlp_perm(int kk) {
  const int s = kk >> 5, g = (kk >> 3) & 3, e = kk & 7;
  return 16 * (2 * s + (e >> 2)) + 4 * g + (e & 3);
}

/* Forward pass of NG groups of 16 samples by one wave, in registers (the
 * groups share the weight loads and interleave their products).  Lane (g, c) =
 * (lane >> 4, lane & 15) works on sample c of each group (x0[q], x1[q]: its
 * normalised inputs; the four lanes of a sample pass the same values) and holds the hidden units
 * 16 t + 4 g + e.  The 64 x 64 layers are computed transposed, Out^T = W In^T,
 * with the activations as the B operand of v_mfma_f32_16x16x32_bf16: the
 * accumulator layout of one layer (units 4 g + e of tile t, sample c) is, with
 * the permuted weights, exactly the operand layout of the next.  W2p and W3p
 * may point to LDS or global memory.  out[q] receives the normalised output
 * of sample c of group q in its four lanes. */
/* Where the weights are: SFP_W_LDS (staged by the caller) or SFP_W_GLOBAL;
 * the loads then use that address space (ds_read / global_load), not
 * generic flat loads, which wait on both the memory and the LDS counters. */
#define SFP_W_GLOBAL 1
#define SFP_W_LDS 3
#ifdef __HIP_DEVICE_COMPILE__
#define SFP_WAS(n) __attribute__((address_space(n)))
#else
#define SFP_WAS(n)
#endif

/* The weight pointer, in its address space, passed through an empty asm
 * statement: the compiler can no longer treat the weight loads as
 * loop-invariant, so a caller's loop does not hoist all 32 operand loads into
 * registers (no memory access is made by the statement). */
template <int WAS>
__device__ inline SFP_WAS(WAS) const __bf16* sfp_mlp_opaque(const __bf16* p) {
  SFP_WAS(WAS) const __bf16* q = (SFP_WAS(WAS) const __bf16*)p;
  asm volatile("" : "+v"(q));
  return q;
}

template <int NG, int NIN = 2, int WAS = SFP_W_GLOBAL>
__device__ inline __attribute__((always_inline)) void sfp_mlp_evalN(const sfp_mlp_dev* M, const __bf16* W2g, const __bf16* W3g,
                                     const float* x0, const float* x1, float* out,
                                     const float* x2 = nullptr) {
  SFP_WAS(WAS) const __bf16* W2p = sfp_mlp_opaque<WAS>(W2g);
  SFP_WAS(WAS) const __bf16* W3p = sfp_mlp_opaque<WAS>(W3g);
  const int l = __lane_id();
  const int c = l & 15, g = l >> 4;
  float h[NG][16];
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int n = 16 * t + 4 * g + e;
      const float w0 = M->W1[2 * n], w1 = M->W1[2 * n + 1], bb = M->b1[n];
#pragma unroll
      for (int q = 0; q < NG; ++q) {
        float v = w0 * x0[q] + w1 * x1[q];
        if (NIN == 3) v += M->W1c[n] * x2[q];
        h[q][4 * t + e] = v + bb;
      }
    }
#pragma unroll
  for (int layer = 0; layer < 2; ++layer) {
    SFP_WAS(WAS) const __bf16* Wp = layer == 0 ? W2p : W3p;
    const float* b = layer == 0 ? M->b2 : M->b3;
    sfp_bf16x8 B[NG][2];
#pragma unroll
    for (int q = 0; q < NG; ++q)
#pragma unroll
      for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          B[q][s][e] = (__bf16)sfp_silu(h[q][4 * (2 * s) + e]);
          B[q][s][4 + e] = (__bf16)sfp_silu(h[q][4 * (2 * s + 1) + e]);
        }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const sfp_bf16x8 A0 = *(SFP_WAS(WAS) const sfp_bf16x8*)(Wp + (16 * t + c) * 64 + 8 * g);
      const sfp_bf16x8 A1 =
          *(SFP_WAS(WAS) const sfp_bf16x8*)(Wp + (16 * t + c) * 64 + 32 + 8 * g);
      float bb[4];
#pragma unroll
      for (int e = 0; e < 4; ++e) bb[e] = b[16 * t + 4 * g + e];
#pragma unroll
      for (int q = 0; q < NG; ++q) {
        sfp_f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A0, B[q][0], acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A1, B[q][1], acc, 0, 0, 0);
#pragma unroll
        for (int e = 0; e < 4; ++e) h[q][4 * t + e] = acc[e] + bb[e];
      }
    }
  }
#pragma unroll
  for (int q = 0; q < NG; ++q) {
    float o = 0.0f;
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int e = 0; e < 4; ++e) o += sfp_silu(h[q][4 * t + e]) * M->W4[16 * t + 4 * g + e];
    o += __shfl_xor(o, 16);
    o += __shfl_xor(o, 32);
    out[q] = o + M->b4;
  }
}

/* NG = 1: one group of 16 samples */
template <int WAS = SFP_W_GLOBAL>
__device__ inline float sfp_mlp_eval16(const sfp_mlp_dev* M, const __bf16* W2p,
                                       const __bf16* W3p, float x0, float x1) {
  float o;
  sfp_mlp_evalN<1, 2, WAS>(M, W2p, W3p, &x0, &x1, &o);
  return o;
}

/* Photon energies above this fraction are taken from the table at the same
 * uniform: pair creation by the photon depends exponentially on its energy
 * (about exp(-8 / (3 delta chi))), so the hard tail needs the table's
 * accuracy, while the network serves the bulk of the spectrum. */
#define SFP_MLP_DELTA_CUT 0.2

/* delta from the normalised output, floored like the table it learns
 * (sfp_od_inverse_cdf_from0) */
__device__ inline double sfp_mlp_delta(const sfp_mlp_dev* M, float o) {
  const double d = (double)__expf(o * M->out_sd + M->out_mu);
  return d > SFP_DELTA_FLOOR ? d : SFP_DELTA_FLOOR;
}

/* NG = 1 with three inputs (the emission-rate network) */
__device__ inline float sfp_mlp_eval16_3(const sfp_mlp_dev* M, float x0, float x1, float x2) {
  float o;
  sfp_mlp_evalN<1, 3>(M, M->W2p, M->W3p, &x0, &x1, &o, &x2);
  return o;
}

/* The emission-rate network learns log(f + SFP_RATE_EPS): a smooth floor
 * that keeps the target's range to about [-14, 3] while the densities that
 * matter (f above ~1e-4; f = 1e-6 is an emission probability below 1e-7 per
 * step) keep their relative accuracy; f = max(exp(out) - SFP_RATE_EPS, 0). */
#define SFP_RATE_EPS 1e-6f

__device__ inline float sfp_rate_from_net(const sfp_mlp_dev* M, float o) {
  const float f = __expf(o * M->out_sd + M->out_mu) - SFP_RATE_EPS;
  return f > 0.0f ? f : 0.0f;
}

/* log10 chi -> normalised input */
__device__ inline float sfp_mlp_x0(const sfp_mlp_dev* M, float lc) {
  return (lc - M->x0_shift) * M->x0_scale;
}

#endif /* SFP_MLP_H */
