/*
 * sfp_device32.h -- the mixed-precision particle step: field evaluation, the
 * Boris / T-BMT / LL pushers and QED sampling in FP32 with the hardware's
 * transcendental instructions, while positions and time stay FP64 (the
 * callers accumulate x += c dt u / gamma in double) and chi is formed in FP64
 * from the FP32 fields and momenta, where E + v x B can nearly cancel.
 *
 * Carrier phases are formed in FP64 from the FP64 position and time and
 * reduced modulo 2 pi in FP64 before the FP32 cosine, so the phase error does
 * not grow with the distance travelled.
 *
 * Same formulas and operation order as the FP64 device code (sfp_device.h),
 * which follows fields.py, inline.py and qed/; only the precision differs,
 * and FP32 quotients are products with the hardware reciprocal (sfp32_div,
 * about 2 ulp: a tenth of the instructions of the IEEE division sequence).
 */
#ifndef SFP_DEVICE32_H
#define SFP_DEVICE32_H

#include "sfp_device.h"

#define SFP_TWO_PI 6.283185307179586
#define SFP_INV_TWO_PI 0.15915494309189535

/* phase reduced to [-pi, pi] in FP64, then rounded */
__device__ inline float sfp32_reduce(double ph) {
  return (float)(ph - SFP_TWO_PI * __builtin_rint(ph * SFP_INV_TWO_PI));
}

__device__ inline float sfp32_exp(float x) { return __expf(x); }

/* a / b as a times the hardware reciprocal of b (1 ulp), about 2 ulp */
__device__ inline float sfp32_div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }

/* fields.py:69-118 */
__device__ inline void sfp32_field_simple(const sfp_term_dev* d, double x, double y, double z,
                                          double ct, float* F) {
  const double* c = d->c;
  const float direction = (float)c[FS_DIR];
  const double phi_d = c[FS_K0] * (c[FS_DIR] * x - c[FS_DX0] - ct);
  const float phi = (float)phi_d;
  const float r2 = (float)(y * y + z * z);
  float pn = 1.0f;
  for (int k = 0; k < d->n; ++k) pn *= phi;
  const float E = (float)c[FS_E0] * __cosf(sfp32_reduce(phi_d + c[FS_CEP])) *
                  sfp32_exp(-r2 * (float)c[FS_RW0SQ]) * sfp32_exp(-pn * (float)c[FS_RKCTN]);
  const float Ey = E * (float)c[FS_COSP];
  const float Ez = E * (float)c[FS_SINP];
  const float ic = (float)(1.0 / SFP_C);
  F[0] = 0.0f;
  F[1] = Ey;
  F[2] = Ez;
  F[3] = 0.0f;
  F[4] = -Ez * ic * direction;
  F[5] = Ey * ic * direction;
}

/* fields.py:120-170; the curvature term is 0 on the focal plane x == 0 */
__device__ inline void sfp32_field_gaussian(const sfp_term_dev* d, double x, double y, double z,
                                            double t, float* F) {
  const double* c = d->c;
  const float k0 = (float)c[FG_K0], zR = (float)c[FG_ZR], w0 = (float)c[FG_W0];
  const float xf = (float)x;
  const float r2 = (float)(y * y + z * z);
  const float xz = xf * (float)c[FG_RZR];
  const float wx2 = w0 * w0 * (1.0f + xz * xz);
  const float gouy = atanf(xz);
  const float curv = (x == 0.0) ? 0.0f : sfp32_div(k0 * r2 * xf, 2.0f * (xf * xf + zR * zR));
  const double phl = c[FG_K0] * (x - c[FG_X0] - c[FG_DC] * t);
  const double ph_d = phl + (double)(curv - gouy);
  const float phi = (float)ph_d;
  const float E = (float)c[FG_E0] * __cosf(sfp32_reduce(ph_d + c[FG_CEP])) *
                  __frsqrt_rn(1.0f + xz * xz) * sfp32_exp(sfp32_div(-r2, wx2)) *
                  sfp32_exp(-(phi * phi) * (float)c[FG_RKCT2]);
  const float Ey = E * (float)c[FG_COSP];
  const float Ez = E * (float)c[FG_SINP];
  const float direction = (float)c[FG_DIR];
  const float ic = (float)(1.0 / SFP_C);
  F[0] = 0.0f;
  F[1] = Ey;
  F[2] = Ez;
  F[3] = 0.0f;
  F[4] = -Ez * ic * direction;
  F[5] = Ey * ic * direction;
}

/* fields.py:173-265 */
__device__ inline void sfp32_field_focused(const sfp_term_dev* d, double x, double y, double z,
                                           double ct, float* F) {
  const double* c = d->c;
  const float zR = (float)c[FF_ZR], k0 = (float)c[FF_K0], w0 = (float)c[FF_W0];
  const float eps = (float)c[FF_EPS], e2 = (float)c[FF_E2], e3 = (float)c[FF_E3],
              e4 = (float)c[FF_E4], e5 = (float)c[FF_E5];
  const float direction = (float)c[FF_DIR];
  const double xd = c[FF_DIR] * (x - c[FF_XFOC]);
  const float xf = (float)xd;
  const float r2 = (float)(y * y + z * z);
  const float xz = sfp32_div(xf, zR);
  const float wx = w0 * sqrtf(1.0f + xz * xz);
  const float psiG = atanf(xz);
  const float psiR = sfp32_div(k0 * xf * r2 / 2.0f, xf * xf + zR * zR);
  const double psiP = c[FF_K0] * (xd - c[FF_DX0] - ct);
  const double psi_d = (double)(psiG + psiR) + psiP + c[FF_CEP];
  const float pc = (float)(psi_d - c[FF_CEP]);
  const float psi = sfp32_reduce(psi_d);
  const float iw0 = sfp32_div(1.0f, w0);
  const float rho2 = r2 * (iw0 * iw0);
  const float xi = (float)y * iw0, nu = (float)z * iw0;
  const float eta = __frsqrt_rn(1.0f + xz * xz);
  const float rho4 = rho2 * rho2, rho6 = rho4 * rho2, rho8 = rho4 * rho4;
  const float iwx = sfp32_div(1.0f, wx);
  const float E = (float)c[FF_E0] * w0 * iwx * sfp32_exp(-r2 * (iwx * iwx)) *
                  sfp32_exp(sfp32_div(-(pc * pc), (float)c[FF_KCT2]));
  float ek[8];
  ek[0] = 1.0f;
  for (int k = 1; k < 8; ++k) ek[k] = ek[k - 1] * eta;
  const float S0 = __sinf(psi);
  const float S2 = ek[2] * __sinf(psi + 2 * psiG), S3 = ek[3] * __sinf(psi + 3 * psiG);
  const float S4 = ek[4] * __sinf(psi + 4 * psiG), S5 = ek[5] * __sinf(psi + 5 * psiG);
  const float S6 = ek[6] * __sinf(psi + 6 * psiG);
  const float C1 = ek[1] * __cosf(psi + psiG), C2 = ek[2] * __cosf(psi + 2 * psiG);
  const float C3 = ek[3] * __cosf(psi + 3 * psiG), C4 = ek[4] * __cosf(psi + 4 * psiG);
  const float C5 = ek[5] * __cosf(psi + 5 * psiG), C6 = ek[6] * __cosf(psi + 6 * psiG);
  const float C7 = ek[7] * __cosf(psi + 7 * psiG);
  const float xi2 = xi * xi;
  const float Ex_ = E * xi *
                    (eps * C1 + e3 * (-C2 / 2 + rho2 * C3 - rho4 * C4 / 4) +
                     e5 * (-3.0f / 8 * C3 - 3.0f / 8 * rho2 * C4 + 17.0f / 16 * rho4 * C5 -
                           3.0f / 8 * rho6 * C6 + rho8 * C7 / 32));
  const float Ey_ = E * (S0 + e2 * (xi2 * S2 - rho4 * S3 / 4) +
                         e4 * (S2 / 8 - rho2 * S3 / 4 - rho2 * (rho2 - 16 * xi2) * S4 / 16 -
                               rho4 * (rho2 + 2 * xi2) * S5 / 8 + rho8 * S6 / 32));
  const float Ez_ = E * xi * nu * (e2 * S2 + e4 * (rho2 * S4 - rho4 * S5 / 4));
  const float ic = (float)(1.0 / SFP_C);
  const float Bx_ = E * ic * nu *
                    (eps * C1 + e3 * (C2 / 2 + rho2 * C3 / 2 - rho4 * C4 / 4) +
                     e5 * (3.0f / 8 * C3 + 3.0f / 8 * rho2 * C4 + 3.0f / 16 * rho4 * C5 -
                           rho6 * C6 / 4 + rho8 * C7 / 32));
  const float Bz_ = E * ic *
                    (S0 + e2 * (rho2 * S2 / 2 - rho4 * S3 / 4) +
                     e4 * (-S2 / 8 + rho2 * S3 / 4 + 5.0f / 16 * rho4 * S4 - rho6 * S5 / 4 +
                           rho8 * S6 / 32));
  const float cp = (float)c[FF_COSP], sp = (float)c[FF_SINP];
  F[0] = Ex_;
  F[1] = Ey_ * cp - Ez_ * sp;
  F[2] = Ey_ * sp + Ez_ * cp;
  F[3] = Bx_;
  F[4] = (-Bz_ * sp) * direction;
  F[5] = (Bz_ * cp) * direction;
}

/* Fields.__add__ (fields.py:52-61): left fold over the terms */
template <int FM = SFP_FM_ANY>
__device__ inline void sfp32_field_eval(const sfp_fieldset* fs, double x, double y, double z,
                                        double t, float* F) {
  const double ct = SFP_C * t;
  for (int i = 0; i < 6; ++i) F[i] = 0.0f;
  for (int j = 0; j < fs->nterms; ++j) {
    float G[6];
    const sfp_term_dev* d = &fs->t[j];
    int kind = (FM == SFP_FM_GAUSSIAN) ? SFP_FIELD_GAUSSIAN : d->kind;
    if (FM == SFP_FM_PLANE && kind != SFP_FIELD_SIMPLE) kind = SFP_FIELD_STATIC;
    if (kind == SFP_FIELD_GAUSSIAN) {
      sfp32_field_gaussian(d, x, y, z, t, G);
    } else if (kind == SFP_FIELD_SIMPLE) {
      sfp32_field_simple(d, x, y, z, ct, G);
    } else if (kind == SFP_FIELD_STATIC) {
      for (int i = 0; i < 6; ++i) G[i] = (float)d->c[i];
    } else {
      sfp32_field_focused(d, x, y, z, ct, G);
    }
    for (int i = 0; i < 6; ++i) F[i] += G[i];
  }
}

/* inline.py:4-36 in FP32; returns 1/gamma */
__device__ inline float sfp32_boris(float& ux, float& uy, float& uz, const float* F, float q,
                                    float dt) {
  const float efactor = q * dt * (float)(1.0 / (2 * SFP_ME * SFP_C));
  const float bfactor = q * dt * (float)(1.0 / (2 * SFP_ME));
  const float uxm = ux + efactor * F[0], uym = uy + efactor * F[1], uzm = uz + efactor * F[2];
  const float igm = __frsqrt_rn(1.0f + uxm * uxm + uym * uym + uzm * uzm);
  const float Tx = bfactor * F[3] * igm, Ty = bfactor * F[4] * igm, Tz = bfactor * F[5] * igm;
  const float uxp = uxm + uym * Tz - uzm * Ty;
  const float uyp = uym + uzm * Tx - uxm * Tz;
  const float uzp = uzm + uxm * Ty - uym * Tx;
  const float Tf = sfp32_div(2.0f, 1.0f + Tx * Tx + Ty * Ty + Tz * Tz);
  const float Sx = Tf * Tx, Sy = Tf * Ty, Sz = Tf * Tz;
  ux = uxm + uyp * Sz - uzp * Sy + efactor * F[0];
  uy = uym + uzp * Sx - uxp * Sz + efactor * F[1];
  uz = uzm + uxp * Sy - uyp * Sx + efactor * F[2];
  return __frsqrt_rn(1.0f + ux * ux + uy * uy + uz * uz);
}

/* inline.py:39-100 in FP32 (the reference's B_eff coefficients); returns 1/gamma */
__device__ inline float sfp32_boris_tbmt(float& ux, float& uy, float& uz, float& sx, float& sy,
                                         float& sz, const float* F, float q, float ae, float dt) {
  const float efactor = q * dt * (float)(1.0 / (2 * SFP_ME * SFP_C));
  const float bfactor = q * dt * (float)(1.0 / (2 * SFP_ME));
  const float Ex = F[0], Ey = F[1], Ez = F[2], Bx = F[3], By = F[4], Bz = F[5];
  const float uxm = ux + efactor * Ex, uym = uy + efactor * Ey, uzm = uz + efactor * Ez;
  const float udB = uxm * Bx + uym * By + uzm * Bz;
  const float gm = sqrtf(1.0f + uxm * uxm + uym * uym + uzm * uzm);
  const float igm = sfp32_div(1.0f, gm);
  const float igp = sfp32_div(1.0f, gm + 1.0f);
  const float a1 = ae + igm, a2 = ae * gm * igp, a3 = ae + igp;
  const float Bxe = a1 * Bx - a2 * udB * uxm - a3 * (uym * Ez - uzm * Ey);
  const float Bye = a1 * By - a2 * udB * uym - a3 * (uzm * Ex - uxm * Ez);
  const float Bze = a1 * Bz - a2 * udB * uzm - a3 * (uxm * Ey - uym * Ex);
  float Tx = bfactor * Bxe, Ty = bfactor * Bye, Tz = bfactor * Bze;
  const float T2 = Tx * Tx + Ty * Ty + Tz * Tz;
  const float iT = sfp32_div(2.0f, 1.0f + T2);
  float Sx = Tx * iT, Sy = Ty * iT, Sz = Tz * iT;
  const float sxp = sx + sy * Tz - sz * Ty;
  const float syp = sy + sz * Tx - sx * Tz;
  const float szp = sz + sx * Ty - sy * Tx;
  const float sxn = sx + syp * Sz - szp * Sy;
  const float syn = sy + szp * Sx - sxp * Sz;
  const float szn = sz + sxp * Sy - syp * Sx;
  sx = sxn;
  sy = syn;
  sz = szn;
  Tx = bfactor * Bx * igm;
  Ty = bfactor * By * igm;
  Tz = bfactor * Bz * igm;
  const float uxp = uxm + uym * Tz - uzm * Ty;
  const float uyp = uym + uzm * Tx - uxm * Tz;
  const float uzp = uzm + uxm * Ty - uym * Tx;
  const float Tf = sfp32_div(2.0f, 1.0f + Tx * Tx + Ty * Ty + Tz * Tz);
  Sx = Tf * Tx;
  Sy = Tf * Ty;
  Sz = Tf * Tz;
  ux = uxm + uyp * Sz - uzp * Sy + efactor * Ex;
  uy = uym + uzp * Sx - uxp * Sz + efactor * Ey;
  uz = uzm + uxp * Sy - uyp * Sx + efactor * Ez;
  return __frsqrt_rn(1.0f + ux * ux + uy * uy + uz * uz);
}

/* chi in FP64 (inline.py:126-134) from FP32 fields and momentum, with gamma
 * (|u| for a photon) formed directly: no FP64 division */
__device__ inline double sfp32_chi_g(const float* F, double ux, double uy, double uz,
                                     double gamma) {
  const double factor = SFP_E * SFP_HBAR / ((SFP_ME * SFP_ME) * (SFP_C * SFP_C * SFP_C));
  const double Ex = F[0], Ey = F[1], Ez = F[2], Bx = F[3], By = F[4], Bz = F[5];
  const double a = gamma * Ex + (uy * Bz - uz * By) * SFP_C;
  const double b = gamma * Ey + (uz * Bx - ux * Bz) * SFP_C;
  const double c = gamma * Ez + (ux * By - uy * Bx) * SFP_C;
  const double d = ux * Ex + uy * Ey + uz * Ez;
  return factor * sfp_sqrt(a * a + b * b + c * c - d * d);
}
__device__ inline double sfp32_chi(const float* F, float ux, float uy, float uz) {
  const double dux = ux, duy = uy, duz = uz;
  return sfp32_chi_g(F, dux, duy, duz, sfp_sqrt(1.0 + dux * dux + duy * duy + duz * duz));
}
__device__ inline double sfp32_chi_photon(const float* F, float ux, float uy, float uz) {
  const double dux = ux, duy = uy, duz = uz;
  return sfp32_chi_g(F, dux, duy, duz, sfp_sqrt(dux * dux + duy * duy + duz * duz));
}

/* inline.py:104-123: LL (g = 1) or CLL (g = g(chi)) in FP32; returns 1/gamma */
__device__ inline float sfp32_ll(float& ux, float& uy, float& uz, float ig, float chi,
                                 float factor) {
  const float c2 = chi * chi;
  ux = ux + factor * c2 * ux * ig;
  uy = uy + factor * c2 * uy * ig;
  uz = uz + factor * c2 * uz * ig;
  return __frsqrt_rn(1.0f + ux * ux + uy * uy + uz * uz);
}
__device__ inline float sfp32_cll_g(float chi) {
  const float b = 1.0f + 4.8f * (1.0f + chi) * __logf(1.0f + 1.7f * chi) + 2.44f * (chi * chi);
  return exp2f(-2.0f / 3.0f * __log2f(b));
}

/* rejection-sampling table node in FP32 (sfp_rs_node) */
typedef struct {
  float ia, ka, ap, kp;
} sfp_rs_node32;

/* rejection_sampling_tables.py:32-78 in FP32; z > z_max -> 0 */
__device__ inline float sfp32_rs_rate(const sfp_rs_node32* nodes, int z_N, float z_min,
                                      float z_max, float z_delta, float inv_dz, float chi_parent,
                                      float delta, bool photon) {
  const float chi_a = delta * chi_parent;
  const float chi_b = chi_parent - chi_a;
  const float q = photon ? sfp32_div(chi_a, chi_parent * chi_b) : sfp32_div(chi_parent, chi_a * chi_b);
  const float z = exp2f(2.0f / 3.0f * __log2f(q));
  const float factor = (float)(SFP_ALPHA * SFP_ME * (SFP_C * SFP_C) / SFP_HBAR);
  if (!(z <= z_max)) return (z != z) ? z : 0.0f;
  int idx = (int)((z - z_min) * inv_dz);
  if (idx > z_N - 2) idx = z_N - 2;
  const float dz = z - (z_min + (float)idx * z_delta);
  const sfp_rs_node32 nd = nodes[idx];
  const float ia = nd.ia + nd.ka * dz;
  const float ap = nd.ap + nd.kp * dz;
  const float sz = sqrtf(z);
  const float tz = sfp32_div(2.0f, z);
  if (photon) return -factor * (ia + (tz + chi_a * sz) * ap);
  return factor * (ia + (tz - chi_parent * sz) * ap);
}

/* sfp_rs_rate_spin in FP32 (photon emission) */
__device__ inline float sfp32_rs_rate_spin(const sfp_rs_node32* nodes, const float2* ai, int z_N,
                                           float z_min, float z_max, float z_delta, float inv_dz,
                                           float chi_parent, float delta, float sigma) {
  const float chi_a = delta * chi_parent;
  const float chi_b = chi_parent - chi_a;
  const float q = sfp32_div(chi_a, chi_parent * chi_b);
  const float z = exp2f(2.0f / 3.0f * __log2f(q));
  const float factor = (float)(SFP_ALPHA * SFP_ME * (SFP_C * SFP_C) / SFP_HBAR);
  if (!(z <= z_max)) return (z != z) ? z : 0.0f;
  int idx = (int)((z - z_min) * inv_dz);
  if (idx > z_N - 2) idx = z_N - 2;
  const float dz = z - (z_min + (float)idx * z_delta);
  const sfp_rs_node32 nd = nodes[idx];
  const float2 an = ai[idx];
  const float ia = nd.ia + nd.ka * dz;
  const float ap = nd.ap + nd.kp * dz;
  const float a = an.x + an.y * dz;
  return -factor * (ia + (sfp32_div(2.0f, z) + chi_a * sqrtf(z)) * ap + sigma * chi_b * z * a);
}

/* sfp_spin_sigma in FP32 */
__device__ inline float sfp32_spin_sigma(double q, float ux, float uy, float uz, float ig,
                                         const float* F, float sx, float sy, float sz) {
  const float g = sfp32_div(1.0f, ig);
  const float ic = (float)(1.0 / SFP_C);
  const float ex = uy * F[2] - uz * F[1], ey = uz * F[0] - ux * F[2], ez = ux * F[1] - uy * F[0];
  const float bx = uy * F[5] - uz * F[4], by = uz * F[3] - ux * F[5], bz = ux * F[4] - uy * F[3];
  float zx = g * ex * ic + (uy * bz - uz * by);
  float zy = g * ey * ic + (uz * bx - ux * bz);
  float zz = g * ez * ic + (ux * by - uy * bx);
  if (q < 0.0) {
    zx = -zx;
    zy = -zy;
    zz = -zz;
  }
  const float n2 = zx * zx + zy * zy + zz * zz;
  if (!(n2 > 0.0f)) return 0.0f;
  return (sx * zx + sy * zy + sz * zz) * __frsqrt_rn(n2);
}

/* uniform in (0, 1) from 24 bits */
__device__ inline float sfp32_u01(uint32_t w) { return ((float)(w >> 8) + 0.5f) * 0x1p-24f; }

#endif /* SFP_DEVICE32_H */
