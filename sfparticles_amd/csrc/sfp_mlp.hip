// sfp_mlp.hip -- libsfp_hip.so, part 3: the learned QED models.
//
// Training data, training (forward, backward, Adam) and batched inference of
// the MLPs defined in sfp_mlp.h, all on the GPU with BF16 matrix-core products
// and FP32 accumulation.  The reference has no learned component: the
// photon-energy sampler replaces the inverse-CDF table lookup of
// qed/optical_depth_tables.py:59-101 for photon emission, the emission-rate
// network the rejection-sampling rate lookup (qed/rejection_sampling_tables.py:
// 32-53, with the spin term); each is trained on its table, which stays as the
// reference for its error and as the fallback.

#include "sfp_internal.h"
#include "sfp_mlp.h"

namespace {

constexpr int TB = 32;  // training samples per workgroup (one wave)

struct Mlp {
  float* master = nullptr;  // [NPARAM] FP32
  float* m = nullptr;       // Adam moments
  float* v = nullptr;
  float* gpart = nullptr;   // [max_groups][NPARAM] per-workgroup partial gradients
  int64_t gpart_groups = 0;
  float* loss = nullptr;    // per-workgroup partial losses
  sfp_mlp_dev* dev = nullptr;  // device copy for the kernels
  __bf16* WT = nullptr;     // [2][64*64] transposed W2, W3 (backward)
  sfp_mlp_bins* bins = nullptr;  // device
  sfp_mlp_bins bins_host{};
  float lc_min = -3.0f, lc_max = 2.0f, mu = 0.0f, sd = 1.0f;
  int64_t step = 0;
  int nin = 2;       // 2: photon-energy sampler, 3: emission rate
  int nparam = SFP_MLP_NPARAM;
};

// device copy (BF16 square layers + transposes, FP32 rest) from the master
template <int NIN>
__global__ void k_mlp_publish(const float* master, sfp_mlp_dev* d, __bf16* WT, float x0_shift,
                              float x0_scale, float mu, float sd) {
  using O = MlpOff<NIN>;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < 4096; i += gridDim.x * blockDim.x) {
    const int n = i / 64, k = i % 64;
    d->W2[i] = (__bf16)master[O::W2 + i];
    d->W3[i] = (__bf16)master[O::W3 + i];
    d->W2p[i] = (__bf16)master[O::W2 + n * 64 + sfp_mlp_perm(k)];
    d->W3p[i] = (__bf16)master[O::W3 + n * 64 + sfp_mlp_perm(k)];
    WT[k * 64 + n] = (__bf16)master[O::W2 + i];
    WT[4096 + k * 64 + n] = (__bf16)master[O::W3 + i];
  }
  if (blockIdx.x == 0) {
    for (int i = threadIdx.x; i < 128; i += blockDim.x)
      d->W1[i] = master[O::W1 + NIN * (i / 2) + (i % 2)];
    for (int i = threadIdx.x; i < 64; i += blockDim.x)
      d->W1c[i] = NIN == 3 ? master[O::W1 + NIN * i + 2] : 0.0f;
    for (int i = threadIdx.x; i < 64; i += blockDim.x) {
      d->b1[i] = master[O::B1 + i];
      d->b2[i] = master[O::B2 + i];
      d->b3[i] = master[O::B3 + i];
      d->W4[i] = master[O::W4 + i];
    }
    if (threadIdx.x == 0) {
      d->b4 = master[O::B4];
      d->x0_shift = x0_shift;
      d->x0_scale = x0_scale;
      d->out_mu = mu;
      d->out_sd = sd;
    }
  }
}

// Training samples: log10 chi ~ U(lc_min, lc_max), r ~ U(0, 1) from Philox,
// target log delta from the FP64 inverse of the photon-emission cumulative
// rate measured from delta = 0 (sfp_od_inverse_cdf_from0).
__global__ void k_mlp_data(sfp_od_dev od, uint64_t seed, int64_t n, float lc_min, float lc_max,
                           float* x, float* y) {
  KLOOP(n) {
    sfp_u32x4 d = sfp_draw4(seed, (uint64_t)ip, 0, 0xffffu, 0x40u);
    const double u0 = sfp_u01(d.v[0], d.v[1]), r = sfp_u01(d.v[2], d.v[3]);
    const double lc = lc_min + (lc_max - lc_min) * u0;
    const double delta = sfp_od_inverse_cdf_from0(od.photon_cdf, &od, sfp_exp10(lc), r);
    x[2 * ip] = (float)lc;
    x[2 * ip + 1] = (float)r;
    y[ip] = (float)sfp_log(delta);
  }
}

// Training samples of the emission-rate network: log10 chi ~ U(lc_min,
// lc_max), r1 ~ U(0, 1), sigma ~ U(-1, 1) from Philox; target
// log(f + SFP_RATE_EPS), f = 3 r1^2 w(chi, r1^3, sigma) / (alpha m c^2 / hbar)
// from the FP64 spin-resolved rate table.
__global__ void k_rate_data(const sfp_rs_node* nodes, const double2* ai, int64_t z_N, double z_min,
                            double z_max, double z_delta, uint64_t seed, int64_t n, float lc_min,
                            float lc_max, float* x, float* y) {
  const double unit = SFP_ALPHA * SFP_ME * (SFP_C * SFP_C) / SFP_HBAR;
  KLOOP(n) {
    sfp_u32x4 d = sfp_draw4(seed, (uint64_t)ip, 0, 0xfffeu, 0x41u);
    const double u0 = sfp_u01(d.v[0], d.v[1]), r1 = sfp_u01(d.v[2], d.v[3]);
    sfp_u32x4 e = sfp_draw4(seed, (uint64_t)ip, 1, 0xfffeu, 0x41u);
    const double sig = 2.0 * sfp_u01(e.v[0], e.v[1]) - 1.0;
    const double lc = lc_min + (lc_max - lc_min) * u0;
    const double w = sfp_rs_rate_spin(nodes, ai, z_N, z_min, z_max, z_delta, sfp_exp10(lc),
                                      r1 * r1 * r1, sig);
    const double f = 3.0 * r1 * r1 * w / unit;
    const float lf = (float)sfp_log((f > 0.0 ? f : 0.0) + (double)SFP_RATE_EPS);
    x[3 * ip] = (float)lc;
    x[3 * ip + 1] = (float)r1;
    x[3 * ip + 2] = (float)sig;
    y[ip] = lf;
  }
}

// bf16x8 of act(src[r][8 consecutive columns]) (act = SiLU or identity)
template <bool SILU>
__device__ inline sfp_bf16x8 load8(const float* src, int row, int col) {
  sfp_bf16x8 v;
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    float f = src[row * 64 + col + e];
    v[e] = (__bf16)(SILU ? sfp_silu(f) : f);
  }
  return v;
}

// One training pass over TB samples per workgroup: forward, squared error
// against the normalised target, backward; the workgroup's gradient (sum over
// its samples of d(mean loss)/d params, the mean taken over `nb` samples) is
// written to gpart[blockIdx.x].  x holds raw (log10 chi, r), y raw log delta.
template <int NIN>
__global__ void __launch_bounds__(64) k_mlp_train(const sfp_mlp_dev* M, const __bf16* WT,
                                                  const float* x, const float* y, int64_t n,
                                                  int64_t off, int64_t nb, float* gpart,
                                                  float* loss) {
  using O = MlpOff<NIN>;
  __shared__ float z1[TB * 64], z2[TB * 64], z3[TB * 64], dz[TB * 64];
  __shared__ float xin[3][TB], dO[TB];
  const int l = __lane_id();
  const int c = l & 15, g = l >> 4;
  const int64_t s0 = off + (int64_t)blockIdx.x * TB;
  float* G = gpart + (int64_t)blockIdx.x * O::N;

  // inputs (samples wrap around the data set)
  if (l < TB) {
    const int64_t i = (s0 + l) % n;
    xin[0][l] = sfp_mlp_x0(M, x[NIN * i]);
    xin[1][l] = 2.0f * x[NIN * i + 1] - 1.0f;
    xin[2][l] = NIN == 3 ? x[NIN * i + 2] : 0.0f;
  }
  __builtin_amdgcn_wave_barrier();
  // forward, two tiles of 16 rows
  float o[2];
  for (int h = 0; h < 2; ++h)
    o[h] = sfp_mlp_forward16<NIN>(M, &xin[0][16 * h], &xin[1][16 * h], &xin[2][16 * h],
                                  z1 + 16 * h * 64, z2 + 16 * h * 64, z3 + 16 * h * 64);
  // output error: lanes 0..15 hold rows 0..15 of each tile
  float err2 = 0.0f;
  if (l < 16) {
    for (int h = 0; h < 2; ++h) {
      const int64_t i = (s0 + 16 * h + l) % n;
      const float t = (y[i] - M->out_mu) / M->out_sd;
      const float e = o[h] - t;
      dO[16 * h + l] = 2.0f * e / (float)nb;
      err2 += e * e;
    }
  }
  for (int k = 32; k > 0; k >>= 1) err2 += __shfl_xor(err2, k);
  if (l == 0) loss[blockIdx.x] = err2;
  __builtin_amdgcn_wave_barrier();

  // layer 4: dW4, db4 and dz3
  {
    float gw = 0.0f, gb = 0.0f;
    for (int r = 0; r < TB; ++r) {
      const float zz = z3[r * 64 + l];
      gw += dO[r] * sfp_silu(zz);
      dz[r * 64 + l] = dO[r] * M->W4[l] * sfp_dsilu(zz);
      gb += dO[r];
    }
    G[O::W4 + l] = gw;
    if (l == 0) G[O::B4] = gb;
  }
  __builtin_amdgcn_wave_barrier();

  // layers 3 and 2: dW = dz^T act(z_prev), db = sum dz, dz_prev = (dz W) silu'(z_prev)
  for (int layer = 3; layer >= 2; --layer) {
    const float* zp = (layer == 3) ? z2 : z1;
    const int wo = (layer == 3) ? O::W3 : O::W2, bo = (layer == 3) ? O::B3 : O::B2;
    const __bf16* WTl = WT + (layer == 3 ? 4096 : 0);
    // dW[n][k] = sum_r dz[r][n] silu(zp[r][k]): M = n, N = k, K = r (32)
    sfp_bf16x8 A[4], B[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        A[t][e] = (__bf16)dz[(8 * g + e) * 64 + 16 * t + c];
        B[t][e] = (__bf16)sfp_silu(zp[(8 * g + e) * 64 + 16 * t + c]);
      }
    }
#pragma unroll
    for (int tm = 0; tm < 4; ++tm)
#pragma unroll
      for (int tn = 0; tn < 4; ++tn) {
        sfp_f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[tm], B[tn], acc, 0, 0, 0);
#pragma unroll
        for (int e = 0; e < 4; ++e) G[wo + (16 * tm + 4 * g + e) * 64 + 16 * tn + c] = acc[e];
      }
    {
      float gb = 0.0f;
      for (int r = 0; r < TB; ++r) gb += dz[r * 64 + l];
      G[bo + l] = gb;
    }
    // dh[r][k] = sum_n dz[r][n] W[n][k]: M = r (2 tiles), N = k (4 tiles), K = n (2 steps)
    sfp_bf16x8 Ad[2][2];
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int s = 0; s < 2; ++s) Ad[h][s] = load8<false>(dz, 16 * h + c, 32 * s + 8 * g);
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        sfp_f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int s = 0; s < 2; ++s) {
          const sfp_bf16x8 bv = *(const sfp_bf16x8*)(WTl + (16 * t + c) * 64 + 32 * s + 8 * g);
          acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ad[h][s], bv, acc, 0, 0, 0);
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int r = 16 * h + 4 * g + e, k = 16 * t + c;
          dz[r * 64 + k] = acc[e] * sfp_dsilu(zp[r * 64 + k]);
        }
      }
    __builtin_amdgcn_wave_barrier();
  }
  // layer 1: dW1[n][j] = sum_r dz1[r][n] x_j[r], db1
  {
    float g0 = 0.0f, g1 = 0.0f, g2 = 0.0f, gb = 0.0f;
    for (int r = 0; r < TB; ++r) {
      const float d = dz[r * 64 + l];
      g0 += d * xin[0][r];
      g1 += d * xin[1][r];
      g2 += d * xin[2][r];
      gb += d;
    }
    G[O::W1 + NIN * l] = g0;
    G[O::W1 + NIN * l + 1] = g1;
    if (NIN == 3) G[O::W1 + NIN * l + 2] = g2;
    G[O::B1 + l] = gb;
  }
}

// sum of the workgroups' partial gradients (fixed order: deterministic)
__global__ void k_mlp_reduce(const float* gpart, int64_t groups, int np, float* grad) {
  KLOOP(np) {
    float s = 0.0f;
    for (int64_t b = 0; b < groups; ++b) s += gpart[b * np + ip];
    grad[ip] = s;
  }
}

// Adam (Kingma & Ba 2015) on the summed gradient
__global__ void k_mlp_adam(float* master, float* m, float* v, const float* gpart, int64_t groups,
                           int np, float lr, float b1, float b2, float eps, float c1, float c2) {
  KLOOP(np) {
    float gsum = 0.0f;
    for (int64_t b = 0; b < groups; ++b) gsum += gpart[b * np + ip];
    const float mm = b1 * m[ip] + (1.0f - b1) * gsum;
    const float vv = b2 * v[ip] + (1.0f - b2) * gsum * gsum;
    m[ip] = mm;
    v[ip] = vv;
    master[ip] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
  }
}

// Batched inference: two groups of 16 samples per wave and pass, in
// registers (sfp_mlp_evalN), weights read through the caches; x raw
// (log10 chi, r[, sigma]); out = normalised network output, delta = the
// sampler's photon energy, f = the rate network's acceptance density.
constexpr int NGB = 2;

template <int NIN>
__device__ inline __attribute__((always_inline)) void mlp_batch(const sfp_mlp_dev* M, const float* x, int64_t n, float* out,
                                 double* delta, double* f) {
  const int l = __lane_id(), c = l & 15;
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
#pragma unroll 1
  for (int64_t base = wave * (16 * NGB); base < n; base += nw * (16 * NGB)) {
    float x0[NGB], x1[NGB], x2[NGB], o[NGB];
#pragma unroll
    for (int q = 0; q < NGB; ++q) {
      int64_t i = base + 16 * q + c;
      if (i >= n) i = n - 1;
      x0[q] = sfp_mlp_x0(M, x[NIN * i]);
      x1[q] = 2.0f * x[NIN * i + 1] - 1.0f;
      x2[q] = NIN == 3 ? x[NIN * i + 2] : 0.0f;
    }
    sfp_mlp_evalN<NGB, NIN>(M, M->W2p, M->W3p, x0, x1, o, x2);
    if (l < 16) {
#pragma unroll
      for (int q = 0; q < NGB; ++q) {
        const int64_t i = base + 16 * q + c;
        if (i < n) {
          if (out) out[i] = o[q];
          if (delta) delta[i] = sfp_mlp_delta(M, o[q]);
          if (f) f[i] = (double)sfp_rate_from_net(M, o[q]);
        }
      }
    }
  }
}

template <int NIN>
__global__ void __launch_bounds__(256) k_mlp_forward(const sfp_mlp_dev* M, const float* x,
                                                     int64_t n, float* out) {
  mlp_batch<NIN>(M, x, n, out, nullptr, nullptr);
}

// delta of photon emission for (log10 chi, r): the network or the table
__global__ void __launch_bounds__(256) k_mlp_sample(const sfp_mlp_dev* M, sfp_od_dev od,
                                                    const float* x, int64_t n, double* delta) {
  mlp_batch<2>(M, x, n, nullptr, delta, nullptr);
}

// emission acceptance density for (log10 chi, r1, sigma): network or table
__global__ void __launch_bounds__(256) k_rate_net(const sfp_mlp_dev* M, const float* x, int64_t n,
                                                  double* f) {
  mlp_batch<3>(M, x, n, nullptr, nullptr, f);
}

__global__ void __launch_bounds__(256) k_rate_table(const sfp_rs_node* nodes, const double2* ai,
                                                    int64_t z_N, double z_min, double z_max,
                                                    double z_delta, const float* x, int64_t n,
                                                    double* f) {
  const double unit = SFP_ALPHA * SFP_ME * (SFP_C * SFP_C) / SFP_HBAR;
  KLOOP(n) {
    const double r1 = (double)x[3 * ip + 1];
    const double w = sfp_rs_rate_spin(nodes, ai, z_N, z_min, z_max, z_delta,
                                      sfp_exp10((double)x[3 * ip]), r1 * r1 * r1,
                                      (double)x[3 * ip + 2]);
    f[ip] = 3.0 * r1 * r1 * w / unit;
  }
}

__global__ void __launch_bounds__(256) k_table_sample(sfp_od_dev od, const float* x, int64_t n,
                                                      double* delta) {
  KLOOP(n) {
    delta[ip] = sfp_od_sample_from0(od.photon_cdf, &od, sfp_exp10((double)x[2 * ip]),
                                    (double)x[2 * ip + 1]);
  }
}

int publish(Mlp* m) {
  const float x0_shift = 0.5f * (m->lc_min + m->lc_max);
  const float x0_scale = 2.0f / (m->lc_max - m->lc_min);
  if (m->nin == 3)
    hipLaunchKernelGGL(k_mlp_publish<3>, dim3(16), dim3(256), 0, 0, (const float*)m->master,
                       m->dev, m->WT, x0_shift, x0_scale, m->mu, m->sd);
  else
    hipLaunchKernelGGL(k_mlp_publish<2>, dim3(16), dim3(256), 0, 0, (const float*)m->master,
                       m->dev, m->WT, x0_shift, x0_scale, m->mu, m->sd);
  HIPCHK(hipGetLastError());
  return SFP_OK;
}

int ensure_groups(Mlp* m, int64_t groups) {
  if (groups <= m->gpart_groups) return SFP_OK;
  if (m->gpart) HIPCHK(hipFree(m->gpart));
  if (m->loss) HIPCHK(hipFree(m->loss));
  m->gpart = nullptr;
  m->loss = nullptr;
  HIPCHK(hipMalloc(&m->gpart, sizeof(float) * m->nparam * groups));
  HIPCHK(hipMalloc(&m->loss, sizeof(float) * groups));
  m->gpart_groups = groups;
  return SFP_OK;
}

int launch_train(Mlp* m, const float* x, const float* y, int64_t n, int64_t off, int64_t nb,
                 int64_t groups) {
  if (m->nin == 3)
    hipLaunchKernelGGL(k_mlp_train<3>, dim3(groups), dim3(64), 0, 0, (const sfp_mlp_dev*)m->dev,
                       (const __bf16*)m->WT, x, y, n, off, nb, m->gpart, m->loss);
  else
    hipLaunchKernelGGL(k_mlp_train<2>, dim3(groups), dim3(64), 0, 0, (const sfp_mlp_dev*)m->dev,
                       (const __bf16*)m->WT, x, y, n, off, nb, m->gpart, m->loss);
  HIPCHK(hipGetLastError());
  return SFP_OK;
}

}  // namespace

// ------------------------------------------------------------------ C ABI --
static int mlp_create(int nin, void** out) {
  Mlp* m = new Mlp();
  m->nin = nin;
  m->nparam = nin == 3 ? SFP_MLP_NPARAM_RATE : SFP_MLP_NPARAM;
  const size_t pb = sizeof(float) * m->nparam;
  HIPCHK(hipMalloc(&m->master, pb));
  HIPCHK(hipMalloc(&m->m, pb));
  HIPCHK(hipMalloc(&m->v, pb));
  HIPCHK(hipMalloc(&m->dev, sizeof(sfp_mlp_dev)));
  HIPCHK(hipMalloc(&m->WT, sizeof(__bf16) * 2 * 4096));
  HIPCHK(hipMalloc(&m->bins, sizeof(sfp_mlp_bins)));
  HIPCHK(hipMemset(m->master, 0, pb));
  HIPCHK(hipMemset(m->m, 0, pb));
  HIPCHK(hipMemset(m->v, 0, pb));
  m->bins_host.lc_min = m->lc_min;
  m->bins_host.lc_max = m->lc_max;
  m->bins_host.fallback = 0;
  HIPCHK(hipMemcpy(m->bins, &m->bins_host, sizeof(sfp_mlp_bins), hipMemcpyHostToDevice));
  CHK(publish(m));
  *out = m;
  return SFP_OK;
}

extern "C" int sfp_mlp_create(void** out) { return mlp_create(2, out); }
extern "C" int sfp_mlp_create_rate(void** out) { return mlp_create(3, out); }

extern "C" int sfp_mlp_shape(void* p, int32_t* nin, int32_t* nparam) {
  Mlp* m = (Mlp*)p;
  if (!m) return set_err(SFP_E_ARG, "null network");
  *nin = m->nin;
  *nparam = m->nparam;
  return SFP_OK;
}

extern "C" int sfp_mlp_destroy(void* p) {
  Mlp* m = (Mlp*)p;
  if (!m) return SFP_OK;
  (void)hipDeviceSynchronize();
  for (void* q : {(void*)m->master, (void*)m->m, (void*)m->v, (void*)m->gpart, (void*)m->loss,
                  (void*)m->dev, (void*)m->WT, (void*)m->bins})
    if (q) (void)hipFree(q);
  delete m;
  return SFP_OK;
}

extern "C" int sfp_mlp_set_params(void* p, const float* params, const float* norm) {
  Mlp* m = (Mlp*)p;
  if (!m) return set_err(SFP_E_ARG, "null network");
  if (!(norm[1] > norm[0]) || !(norm[3] > 0.0f))
    return set_err(SFP_E_ARG, "normalisation needs lc_max > lc_min and sd > 0");
  HIPCHK(hipMemcpy(m->master, params, sizeof(float) * m->nparam, hipMemcpyHostToDevice));
  m->lc_min = norm[0];
  m->lc_max = norm[1];
  m->mu = norm[2];
  m->sd = norm[3];
  m->bins_host.lc_min = m->lc_min;
  m->bins_host.lc_max = m->lc_max;
  HIPCHK(hipMemcpy(m->bins, &m->bins_host, sizeof(sfp_mlp_bins), hipMemcpyHostToDevice));
  HIPCHK(hipMemset(m->m, 0, sizeof(float) * m->nparam));
  HIPCHK(hipMemset(m->v, 0, sizeof(float) * m->nparam));
  m->step = 0;
  CHK(publish(m));
  HIPCHK(hipDeviceSynchronize());
  return SFP_OK;
}

extern "C" int sfp_mlp_get_params(void* p, float* params, float* norm) {
  Mlp* m = (Mlp*)p;
  if (!m) return set_err(SFP_E_ARG, "null network");
  HIPCHK(hipMemcpy(params, m->master, sizeof(float) * m->nparam, hipMemcpyDeviceToHost));
  norm[0] = m->lc_min;
  norm[1] = m->lc_max;
  norm[2] = m->mu;
  norm[3] = m->sd;
  return SFP_OK;
}

extern "C" int sfp_mlp_set_fallback(void* p, uint64_t mask) {
  Mlp* m = (Mlp*)p;
  if (!m) return set_err(SFP_E_ARG, "null network");
  m->bins_host.fallback = mask;
  HIPCHK(hipMemcpy(m->bins, &m->bins_host, sizeof(sfp_mlp_bins), hipMemcpyHostToDevice));
  return SFP_OK;
}

extern "C" int sfp_mlp_make_data(void* p, void* tables, int64_t n, uint64_t seed, float* x,
                                 float* y) {
  Mlp* m = (Mlp*)p;
  Tables* T = (Tables*)tables;
  if (!m || !T) return set_err(SFP_E_ARG, "network and tables needed");
  if (m->nin == 3) {
    if (!T->rs_ai) return set_err(SFP_E_ARG, "the rate network needs the spin-resolved rate table");
    if (n <= 0) return SFP_OK;
    hipLaunchKernelGGL(k_rate_data, dim3(op_grid(n)), dim3(OPB), 0, 0,
                       (const sfp_rs_node*)T->rs_nodes, (const double2*)T->rs_ai, T->z_N, T->z_min,
                       T->z_max, T->z_delta, seed, n, m->lc_min, m->lc_max, x, y);
    HIPCHK(hipGetLastError());
    HIPCHK(hipDeviceSynchronize());
    return SFP_OK;
  }
  if (!T->has_od) return set_err(SFP_E_ARG, "the sampler needs the optical-depth tables");
  if (n <= 0) return SFP_OK;
  hipLaunchKernelGGL(k_mlp_data, dim3(op_grid(n)), dim3(OPB), 0, 0, T->od, seed, n, m->lc_min,
                     m->lc_max, x, y);
  HIPCHK(hipGetLastError());
  HIPCHK(hipDeviceSynchronize());
  return SFP_OK;
}

extern "C" int sfp_mlp_grad(void* p, const float* x, const float* y, int64_t n, float* grad,
                            float* loss) {
  Mlp* m = (Mlp*)p;
  if (!m) return set_err(SFP_E_ARG, "null network");
  if (n <= 0 || n % TB) return set_err(SFP_E_ARG, "sample count must be a positive multiple of %d", TB);
  const int64_t groups = n / TB;
  CHK(ensure_groups(m, groups));
  CHK(launch_train(m, x, y, n, 0, n, groups));
  float* gsum = nullptr;
  HIPCHK(hipMalloc(&gsum, sizeof(float) * m->nparam));
  hipLaunchKernelGGL(k_mlp_reduce, dim3(op_grid(m->nparam)), dim3(OPB), 0, 0,
                     (const float*)m->gpart, groups, m->nparam, gsum);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpy(grad, gsum, sizeof(float) * m->nparam, hipMemcpyDeviceToHost));
  HIPCHK(hipFree(gsum));
  std::vector<float> l(groups);
  HIPCHK(hipMemcpy(l.data(), m->loss, sizeof(float) * groups, hipMemcpyDeviceToHost));
  double s = 0.0;
  for (float v : l) s += v;
  *loss = (float)(s / (double)n);
  return SFP_OK;
}

extern "C" int sfp_mlp_train(void* p, const float* x, const float* y, int64_t n, int32_t steps,
                             int32_t batch, float lr, float* loss_curve) {
  Mlp* m = (Mlp*)p;
  if (!m) return set_err(SFP_E_ARG, "null network");
  if (batch <= 0 || batch % TB) return set_err(SFP_E_ARG, "batch must be a positive multiple of %d", TB);
  if (n < batch) return set_err(SFP_E_ARG, "fewer samples than one batch");
  const int64_t groups = batch / TB;
  CHK(ensure_groups(m, groups));
  std::vector<float> l(groups);
  const float b1 = 0.9f, b2 = 0.999f;
  for (int32_t it = 0; it < steps; ++it) {
    const int64_t off = ((int64_t)it * batch) % n;
    CHK(launch_train(m, x, y, n, off, (int64_t)batch, groups));
    m->step += 1;
    // cosine decay over this call's steps
    const float lr_t = lr * 0.5f * (1.0f + cosf(3.14159265f * (float)it / (float)steps));
    const float c1 = 1.0f - powf(b1, (float)m->step), c2 = 1.0f - powf(b2, (float)m->step);
    hipLaunchKernelGGL(k_mlp_adam, dim3(op_grid(m->nparam)), dim3(OPB), 0, 0, m->master, m->m,
                       m->v, (const float*)m->gpart, groups, m->nparam, lr_t, b1, b2, 1e-8f, c1,
                       c2);
    HIPCHK(hipGetLastError());
    CHK(publish(m));
    if (loss_curve) {
      HIPCHK(hipMemcpy(l.data(), m->loss, sizeof(float) * groups, hipMemcpyDeviceToHost));
      double s = 0.0;
      for (float v : l) s += v;
      loss_curve[it] = (float)(s / (double)batch);
    }
  }
  HIPCHK(hipDeviceSynchronize());
  return SFP_OK;
}

extern "C" int sfp_mlp_forward(void* p, const float* x, float* out, int64_t n) {
  Mlp* m = (Mlp*)p;
  if (!m) return set_err(SFP_E_ARG, "null network");
  if (n <= 0) return SFP_OK;
  int64_t blocks = std::min<int64_t>((n + 255) / 256, 2048);
  if (m->nin == 3)
    hipLaunchKernelGGL(k_mlp_forward<3>, dim3(blocks), dim3(256), 0, 0, (const sfp_mlp_dev*)m->dev,
                       x, n, out);
  else
    hipLaunchKernelGGL(k_mlp_forward<2>, dim3(blocks), dim3(256), 0, 0, (const sfp_mlp_dev*)m->dev,
                       x, n, out);
  HIPCHK(hipGetLastError());
  HIPCHK(hipDeviceSynchronize());
  return SFP_OK;
}

extern "C" int sfp_mlp_sample(void* p, void* tables, const float* x, int64_t n, int32_t method,
                              double* delta, float* ms) {
  Mlp* m = (Mlp*)p;
  Tables* T = (Tables*)tables;
  if (!m || !T || !T->has_od) return set_err(SFP_E_ARG, "network and optical-depth tables needed");
  if (m->nin != 2) return set_err(SFP_E_ARG, "not a photon-energy sampler");
  if (method != 0 && method != 1) return set_err(SFP_E_ARG, "method is 0 (network) or 1 (table)");
  if (n <= 0) return SFP_OK;
  int64_t blocks = std::min<int64_t>((n + 255) / 256, 2048);
  if (method == 1) blocks = op_grid(n);
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  HIPCHK(hipEventRecord(e0, 0));
  if (method == 0) {
    hipLaunchKernelGGL(k_mlp_sample, dim3(blocks), dim3(256), 0, 0, (const sfp_mlp_dev*)m->dev,
                       T->od, x, n, delta);
  } else {
    hipLaunchKernelGGL(k_table_sample, dim3(blocks), dim3(256), 0, 0, T->od, x, n, delta);
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(e1, 0));
  HIPCHK(hipEventSynchronize(e1));
  float t = 0.f;
  HIPCHK(hipEventElapsedTime(&t, e0, e1));
  if (ms) *ms = t;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return SFP_OK;
}

extern "C" int sfp_mlp_rate(void* p, void* tables, const float* x, int64_t n, int32_t method,
                            double* f, float* ms) {
  Mlp* m = (Mlp*)p;
  Tables* T = (Tables*)tables;
  if (!m || !T || !T->rs_ai) return set_err(SFP_E_ARG, "network and spin-resolved rate table needed");
  if (m->nin != 3) return set_err(SFP_E_ARG, "not an emission-rate network");
  if (method != 0 && method != 1) return set_err(SFP_E_ARG, "method is 0 (network) or 1 (table)");
  if (n <= 0) return SFP_OK;
  const int64_t blocks = method == 0 ? std::min<int64_t>((n + 255) / 256, 2048) : op_grid(n);
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  HIPCHK(hipEventRecord(e0, 0));
  if (method == 0)
    hipLaunchKernelGGL(k_rate_net, dim3(blocks), dim3(256), 0, 0, (const sfp_mlp_dev*)m->dev, x, n,
                       f);
  else
    hipLaunchKernelGGL(k_rate_table, dim3(blocks), dim3(256), 0, 0, (const sfp_rs_node*)T->rs_nodes,
                       (const double2*)T->rs_ai, T->z_N, T->z_min, T->z_max, T->z_delta, x, n, f);
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(e1, 0));
  HIPCHK(hipEventSynchronize(e1));
  float t = 0.f;
  HIPCHK(hipEventElapsedTime(&t, e0, e1));
  if (ms) *ms = t;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return SFP_OK;
}

// for the simulation: device copy and bins
extern "C" int sfp_mlp_device(void* p, const void** dev, const void** bins) {
  Mlp* m = (Mlp*)p;
  if (!m) return set_err(SFP_E_ARG, "null network");
  *dev = m->dev;
  *bins = m->bins;
  return SFP_OK;
}
