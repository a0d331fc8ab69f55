// sfp_lib.hip -- libsfp_hip.so, part 1: the per-operation kernels of the
// reference's three backend seams, QED table upload, device species stores and
// small device-memory helpers.  The fused simulation step is in sfp_step.hip.
//
// Layout in HBM: one struct-of-arrays store per species (x, y, z, ux, uy, uz
// as float64, a uint64 id and a uint8 pruned flag; tau and spin when the
// species needs them).  inv_gamma, chi, E, B, the event flag and delta are
// transients of the fused step; the per-operation entry points materialise
// them as auxiliary store attributes.
//
// All physics is compiled with -ffp-contract=off and uses the project's own
// elementary functions (sfp_math.h), so every FP64 result is bit-identical
// to the CPU restatement used for validation.

#include <mutex>

#include "sfp_internal.h"

// ----------------------------------------------------------------- errors --
static thread_local std::string g_err;

int sfp_set_err(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

// ------------------------------------------------------------ per-op API --

__global__ void k_eval_field(const sfp_fieldset* fs, const double* x, const double* y,
                             const double* z, double t, int64_t n, const uint8_t* pruned,
                             double* Ex, double* Ey, double* Ez, double* Bx, double* By,
                             double* Bz) {
  KLOOP(n) {
    if (pruned[ip]) continue;
    double F[6];
    sfp_field_eval(fs, x[ip], y[ip], z[ip], t, F);
    Ex[ip] = F[0]; Ey[ip] = F[1]; Ez[ip] = F[2];
    Bx[ip] = F[3]; By[ip] = F[4]; Bz[ip] = F[5];
  }
}

int sfp_compile_fields(const sfp_field* terms, int32_t nterms, sfp_fieldset* fs) {
  if (nterms < 0 || nterms > SFP_MAX_TERMS)
    return set_err(SFP_E_ARG, "field has %d terms; at most %d are supported", nterms,
                   SFP_MAX_TERMS);
  memset(fs, 0, sizeof(*fs));
  fs->nterms = nterms;
  for (int i = 0; i < nterms; ++i) {
    if (terms[i].kind < SFP_FIELD_STATIC || terms[i].kind > SFP_FIELD_FOCUSED)
      return set_err(SFP_E_ARG, "unknown field kind %d", terms[i].kind);
    sfp_compile_term(&terms[i], &fs->t[i]);
  }
  return SFP_OK;
}

// scratch for small per-op uploads (field sets); one per stream call, freed lazily
struct DevBuf {
  void* p = nullptr;
  size_t n = 0;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  int ensure(size_t bytes) {
    if (bytes <= n) return SFP_OK;
    if (p) HIPCHK(hipFree(p));
    p = nullptr;
    HIPCHK(hipMalloc(&p, bytes));
    n = bytes;
    return SFP_OK;
  }
};

static thread_local DevBuf g_fieldbuf;

extern "C" int sfp_eval_field(const sfp_field* terms, int32_t nterms, const double* x,
                              const double* y, const double* z, double t, int64_t n,
                              const uint8_t* pruned, double* Ex, double* Ey, double* Ez,
                              double* Bx, double* By, double* Bz, sfp_stream_t stream) {
  if (n <= 0) return SFP_OK;
  sfp_fieldset fs;
  CHK(sfp_compile_fields(terms, nterms, &fs));
  CHK(g_fieldbuf.ensure(sizeof(fs)));
  HIPCHK(hipMemcpyAsync(g_fieldbuf.p, &fs, sizeof(fs), hipMemcpyHostToDevice, S(stream)));
  hipLaunchKernelGGL(k_eval_field, dim3(op_grid(n)), dim3(OPB), 0, S(stream),
                     (const sfp_fieldset*)g_fieldbuf.p, x, y, z, t, n, pruned, Ex, Ey, Ez, Bx, By,
                     Bz);
  HIPCHK(hipGetLastError());
  // the field set buffer is reused by the next call: keep the copy ordered
  HIPCHK(hipStreamSynchronize(S(stream)));
  return SFP_OK;
}

__global__ void k_field_bound(const sfp_fieldset* fs, const double* x, const double* y,
                              const double* z, double t, const double* vx, const double* vy,
                              const double* vz, double T, int64_t n, float* out) {
  KLOOP(n) {
    out[ip] = vx ? sfp_field_bound_f32_path<SFP_FM_ANY>(fs, x[ip], y[ip], z[ip], t, vx[ip], vy[ip],
                                                       vz[ip], T)
                 : sfp_field_bound_f32<SFP_FM_ANY>(fs, x[ip], y[ip], z[ip], t);
  }
}

extern "C" int sfp_field_bound_eval(const sfp_field* terms, int32_t nterms, const double* x,
                                    const double* y, const double* z, double t, const double* vx,
                                    const double* vy, const double* vz, double T, int64_t n,
                                    float* out, sfp_stream_t stream) {
  if (n <= 0) return SFP_OK;
  if (!(T >= 0.0)) return set_err(SFP_E_ARG, "path duration must be >= 0");
  if ((vx == nullptr) != (vy == nullptr) || (vx == nullptr) != (vz == nullptr))
    return set_err(SFP_E_ARG, "give all three velocity components or none");
  sfp_fieldset fs;
  CHK(sfp_compile_fields(terms, nterms, &fs));
  CHK(g_fieldbuf.ensure(sizeof(fs)));
  HIPCHK(hipMemcpyAsync(g_fieldbuf.p, &fs, sizeof(fs), hipMemcpyHostToDevice, S(stream)));
  hipLaunchKernelGGL(k_field_bound, dim3(op_grid(n)), dim3(OPB), 0, S(stream),
                     (const sfp_fieldset*)g_fieldbuf.p, x, y, z, t, vx, vy, vz, T, n, out);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(S(stream)));
  return SFP_OK;
}

__global__ void k_push_position(double* x, double* y, double* z, const double* ux,
                                const double* uy, const double* uz, const double* ig, int64_t n,
                                const uint8_t* pruned, double cdt) {
  KLOOP(n) {
    if (pruned[ip]) continue;
    double xx = x[ip], yy = y[ip], zz = z[ip];
    sfp_push_position(&xx, &yy, &zz, ux[ip], uy[ip], uz[ip], ig[ip], cdt);
    x[ip] = xx; y[ip] = yy; z[ip] = zz;
  }
}

extern "C" int sfp_push_position(double* x, double* y, double* z, const double* ux,
                                 const double* uy, const double* uz, const double* inv_gamma,
                                 int64_t n, const uint8_t* pruned, double dt,
                                 sfp_stream_t stream) {
  if (n <= 0) return SFP_OK;
  hipLaunchKernelGGL(k_push_position, dim3(op_grid(n)), dim3(OPB), 0, S(stream), x, y, z, ux,
                     uy, uz, inv_gamma, n, pruned, SFP_C * dt);
  HIPCHK(hipGetLastError());
  return SFP_OK;
}

__global__ void k_boris(double* ux, double* uy, double* uz, double* ig, const double* Ex,
                        const double* Ey, const double* Ez, const double* Bx, const double* By,
                        const double* Bz, double q, int64_t n, const uint8_t* pruned, double dt) {
  KLOOP(n) {
    if (pruned[ip]) continue;
    double a = ux[ip], b = uy[ip], c = uz[ip], g;
    sfp_boris(&a, &b, &c, &g, Ex[ip], Ey[ip], Ez[ip], Bx[ip], By[ip], Bz[ip], q, dt);
    ux[ip] = a; uy[ip] = b; uz[ip] = c; ig[ip] = g;
  }
}

extern "C" int sfp_boris(double* ux, double* uy, double* uz, double* inv_gamma, const double* Ex,
                         const double* Ey, const double* Ez, const double* Bx, const double* By,
                         const double* Bz, double q, int64_t n, const uint8_t* pruned, double dt,
                         sfp_stream_t stream) {
  if (n <= 0) return SFP_OK;
  hipLaunchKernelGGL(k_boris, dim3(op_grid(n)), dim3(OPB), 0, S(stream), ux, uy, uz, inv_gamma,
                     Ex, Ey, Ez, Bx, By, Bz, q, n, pruned, dt);
  HIPCHK(hipGetLastError());
  return SFP_OK;
}

__global__ void k_boris_tbmt(double* ux, double* uy, double* uz, double* ig, double* sx,
                             double* sy, double* sz, const double* Ex, const double* Ey,
                             const double* Ez, const double* Bx, const double* By,
                             const double* Bz, double q, double ae, int64_t n,
                             const uint8_t* pruned, double dt) {
  KLOOP(n) {
    if (pruned[ip]) continue;
    double a = ux[ip], b = uy[ip], c = uz[ip], g, s0 = sx[ip], s1 = sy[ip], s2 = sz[ip];
    sfp_boris_tbmt(&a, &b, &c, &g, &s0, &s1, &s2, Ex[ip], Ey[ip], Ez[ip], Bx[ip], By[ip], Bz[ip],
                   q, ae, dt);
    ux[ip] = a; uy[ip] = b; uz[ip] = c; ig[ip] = g;
    sx[ip] = s0; sy[ip] = s1; sz[ip] = s2;
  }
}

extern "C" int sfp_boris_tbmt(double* ux, double* uy, double* uz, double* inv_gamma, double* sx,
                              double* sy, double* sz, const double* Ex, const double* Ey,
                              const double* Ez, const double* Bx, const double* By,
                              const double* Bz, double q, double ae, int64_t n,
                              const uint8_t* pruned, double dt, sfp_stream_t stream) {
  if (n <= 0) return SFP_OK;
  hipLaunchKernelGGL(k_boris_tbmt, dim3(op_grid(n)), dim3(OPB), 0, S(stream), ux, uy, uz,
                     inv_gamma, sx, sy, sz, Ex, Ey, Ez, Bx, By, Bz, q, ae, n, pruned, dt);
  HIPCHK(hipGetLastError());
  return SFP_OK;
}

__global__ void k_ll(double* ux, double* uy, double* uz, double* ig, const double* chi,
                     int64_t n, const uint8_t* pruned, double factor0, int cll) {
  KLOOP(n) {
    if (pruned[ip]) continue;
    double c = chi[ip];
    double factor = cll ? factor0 * sfp_cll_g(c) : factor0;
    double a = ux[ip], b = uy[ip], d = uz[ip], g = ig[ip];
    sfp_ll_apply(&a, &b, &d, &g, c, factor);
    ux[ip] = a; uy[ip] = b; uz[ip] = d; ig[ip] = g;
  }
}

extern "C" int sfp_LL_push(double* ux, double* uy, double* uz, double* inv_gamma,
                           const double* chi, int64_t n, const uint8_t* pruned, double dt,
                           sfp_stream_t stream) {
  if (n <= 0) return SFP_OK;
  hipLaunchKernelGGL(k_ll, dim3(op_grid(n)), dim3(OPB), 0, S(stream), ux, uy, uz, inv_gamma, chi,
                     n, pruned, sfp_ll_factor(dt), 0);
  HIPCHK(hipGetLastError());
  return SFP_OK;
}

extern "C" int sfp_CLL_push(double* ux, double* uy, double* uz, double* inv_gamma,
                            const double* chi, int64_t n, const uint8_t* pruned, double dt,
                            sfp_stream_t stream) {
  if (n <= 0) return SFP_OK;
  hipLaunchKernelGGL(k_ll, dim3(op_grid(n)), dim3(OPB), 0, S(stream), ux, uy, uz, inv_gamma, chi,
                     n, pruned, sfp_ll_factor(dt), 1);
  HIPCHK(hipGetLastError());
  return SFP_OK;
}

__global__ void k_update_chi(const double* Ex, const double* Ey, const double* Ez,
                             const double* Bx, const double* By, const double* Bz,
                             const double* ux, const double* uy, const double* uz,
                             const double* ig, double* chi, int64_t n, const uint8_t* pruned) {
  KLOOP(n) {
    if (pruned[ip]) continue;
    chi[ip] = sfp_chi(Ex[ip], Ey[ip], Ez[ip], Bx[ip], By[ip], Bz[ip], ux[ip], uy[ip], uz[ip],
                      ig[ip]);
  }
}

extern "C" int sfp_update_chi(const double* Ex, const double* Ey, const double* Ez,
                              const double* Bx, const double* By, const double* Bz,
                              const double* ux, const double* uy, const double* uz,
                              const double* inv_gamma, double* chi, int64_t n,
                              const uint8_t* pruned, sfp_stream_t stream) {
  if (n <= 0) return SFP_OK;
  hipLaunchKernelGGL(k_update_chi, dim3(op_grid(n)), dim3(OPB), 0, S(stream), Ex, Ey, Ez, Bx, By,
                     Bz, ux, uy, uz, inv_gamma, chi, n, pruned);
  HIPCHK(hipGetLastError());
  return SFP_OK;
}

__global__ void k_pick_hard(uint8_t* event, const double* delta, const double* ig, double thr,
                            int64_t n) {
  KLOOP(n) {
    if (event[ip]) {
      if (delta[ip] / ig[ip] < thr) event[ip] = 0;
    }
  }
}

extern "C" int sfp_pick_hard_photon(uint8_t* event, const double* delta, const double* inv_gamma,
                                    double threshold, int64_t n, sfp_stream_t stream) {
  if (n <= 0) return SFP_OK;
  hipLaunchKernelGGL(k_pick_hard, dim3(op_grid(n)), dim3(OPB), 0, S(stream), event, delta,
                     inv_gamma, threshold, n);
  HIPCHK(hipGetLastError());
  return SFP_OK;
}

__global__ void k_recoil(double* ux, double* uy, double* uz, double* ig, const uint8_t* event,
                         const double* delta, int64_t n, const uint8_t* pruned) {
  KLOOP(n) {
    if (pruned[ip]) continue;
    if (event[ip]) {
      double a = ux[ip] * (1 - delta[ip]);
      double b = uy[ip] * (1 - delta[ip]);
      double c = uz[ip] * (1 - delta[ip]);
      ux[ip] = a; uy[ip] = b; uz[ip] = c;
      ig[ip] = sfp_inv_gamma_lepton(a, b, c);
    }
  }
}

extern "C" int sfp_photon_recoil(double* ux, double* uy, double* uz, double* inv_gamma,
                                 const uint8_t* event, const double* delta, int64_t n,
                                 const uint8_t* pruned, sfp_stream_t stream) {
  if (n <= 0) return SFP_OK;
  hipLaunchKernelGGL(k_recoil, dim3(op_grid(n)), dim3(OPB), 0, S(stream), ux, uy, uz, inv_gamma,
                     event, delta, n, pruned);
  HIPCHK(hipGetLastError());
  return SFP_OK;
}

__global__ void k_create(const double* xs, const double* ys, const double* zs, const double* uxs,
                         const double* uys, const double* uzs, uint8_t* src_pruned, double* xd,
                         double* yd, double* zd, double* uxd, double* uyd, double* uzd,
                         double* igd, uint8_t* dst_pruned, const int64_t* eidx,
                         const double* delta_src, int64_t nbuf, int64_t nnew, int inverse,
                         int photon, uint64_t* idd, uint64_t next_id, double* chid, double* taud,
                         double* sxd, double* syd, double* szd) {
  KLOOP(nnew) {
    int64_t s = eidx[ip];
    int64_t d = nbuf + ip;
    xd[d] = xs[s];
    yd[d] = ys[s];
    zd[d] = zs[s];
    double dl = delta_src[s];
    if (inverse) dl = 1.0 - dl;
    double a = dl * uxs[s], b = dl * uys[s], c = dl * uzs[s];
    uxd[d] = a; uyd[d] = b; uzd[d] = c;
    igd[d] = photon ? sfp_inv_gamma_photon(a, b, c) : 1.0 / sfp_sqrt(1.0 + a * a + b * b + c * c);
    dst_pruned[d] = 0;
    if (src_pruned) src_pruned[s] = 1;
    if (idd) idd[d] = next_id + (uint64_t)ip;
    if (chid) chid[d] = 0.0;
    if (taud) taud[d] = 0.0;
    if (sxd) { sxd[d] = 0.0; syd[d] = 0.0; szd[d] = 0.0; }
  }
}

extern "C" int sfp_create_photon(const double* x_src, const double* y_src, const double* z_src,
                                 const double* ux_src, const double* uy_src, const double* uz_src,
                                 double* x_dst, double* y_dst, double* z_dst, double* ux_dst,
                                 double* uy_dst, double* uz_dst, double* inv_gamma_dst,
                                 uint8_t* pruned_dst, const int64_t* event_index,
                                 const double* delta, int64_t n_buffered, int64_t n_photon,
                                 uint64_t* id_dst, uint64_t next_id, double* chi_dst,
                                 double* tau_dst, sfp_stream_t stream) {
  if (n_photon <= 0) return SFP_OK;
  hipLaunchKernelGGL(k_create, dim3(op_grid(n_photon)), dim3(OPB), 0, S(stream), x_src, y_src,
                     z_src, ux_src, uy_src, uz_src, (uint8_t*)nullptr, x_dst, y_dst, z_dst, ux_dst,
                     uy_dst, uz_dst, inv_gamma_dst, pruned_dst, event_index, delta, n_buffered,
                     n_photon, 0, 1, id_dst, next_id, chi_dst, tau_dst, (double*)nullptr,
                     (double*)nullptr, (double*)nullptr);
  HIPCHK(hipGetLastError());
  return SFP_OK;
}

extern "C" int sfp_create_pair(const double* x_src, const double* y_src, const double* z_src,
                               const double* ux_src, const double* uy_src, const double* uz_src,
                               uint8_t* photon_pruned, double* x_dst, double* y_dst,
                               double* z_dst, double* ux_dst, double* uy_dst, double* uz_dst,
                               double* inv_gamma_dst, uint8_t* pruned_dst,
                               const int64_t* event_index, const double* delta, int64_t n_buffered,
                               int64_t n_pair, int32_t inverse_delta, uint64_t* id_dst,
                               uint64_t next_id, double* chi_dst, double* tau_dst, double* sx_dst,
                               double* sy_dst, double* sz_dst, sfp_stream_t stream) {
  if (n_pair <= 0) return SFP_OK;
  hipLaunchKernelGGL(k_create, dim3(op_grid(n_pair)), dim3(OPB), 0, S(stream), x_src, y_src,
                     z_src, ux_src, uy_src, uz_src, photon_pruned, x_dst, y_dst, z_dst, ux_dst,
                     uy_dst, uz_dst, inv_gamma_dst, pruned_dst, event_index, delta, n_buffered,
                     n_pair, inverse_delta, 0, id_dst, next_id, chi_dst, tau_dst, sx_dst, sy_dst,
                     sz_dst);
  HIPCHK(hipGetLastError());
  return SFP_OK;
}

__global__ void k_inv_gamma(const double* ux, const double* uy, const double* uz, double* ig,
                            int64_t n, const uint8_t* pruned, int photon) {
  KLOOP(n) {
    if (pruned[ip]) continue;
    ig[ip] = photon ? sfp_inv_gamma_photon(ux[ip], uy[ip], uz[ip])
                    : sfp_inv_gamma_lepton(ux[ip], uy[ip], uz[ip]);
  }
}

extern "C" int sfp_update_inv_gamma(const double* ux, const double* uy, const double* uz,
                                    double* inv_gamma, int64_t n, const uint8_t* pruned,
                                    int32_t photon, sfp_stream_t stream) {
  if (n <= 0) return SFP_OK;
  hipLaunchKernelGGL(k_inv_gamma, dim3(op_grid(n)), dim3(OPB), 0, S(stream), ux, uy, uz,
                     inv_gamma, n, pruned, photon);
  HIPCHK(hipGetLastError());
  return SFP_OK;
}

__global__ void k_math(int fn, const double* x, const double* y, double* out, int64_t n) {
  KLOOP(n) {
    double a = x[ip], b = y[ip], r;
    switch (fn) {
      case 0: r = sfp_exp(a); break;
      case 1: r = sfp_log(a); break;
      case 2: r = sfp_sin(a); break;
      case 3: r = sfp_cos(a); break;
      case 4: r = sfp_atan(a); break;
      case 5: r = sfp_pow(a, b); break;
      case 6: r = sfp_exp10(a); break;
      case 7: r = sfp_log10(a); break;
      case 8: r = sfp_sqrt(a); break;
      default: r = a / b; break;
    }
    out[ip] = r;
  }
}

extern "C" int sfp_math_eval(int32_t fn, const double* x, const double* y, double* out, int64_t n,
                             sfp_stream_t stream) {
  if (n <= 0) return SFP_OK;
  hipLaunchKernelGGL(k_math, dim3(op_grid(n)), dim3(OPB), 0, S(stream), (int)fn, x, y, out, n);
  HIPCHK(hipGetLastError());
  return SFP_OK;
}

// ---- flag compaction for the per-op API: block counts -> scan -> scatter --
static constexpr int CB = 1024;  // flags per block


__global__ void k_block_count(const uint8_t* f, int64_t n, int invert, int64_t* counts) {
  __shared__ int wsum[CB / 64];
  int64_t base = (int64_t)blockIdx.x * CB;
  int cnt = 0;
  for (int k = threadIdx.x; k < CB; k += blockDim.x) {
    int64_t i = base + k;
    if (i < n) cnt += ((f[i] != 0) ^ invert) ? 1 : 0;
  }
  for (int off = 32; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off);
  if (wave_lane() == 0) wsum[threadIdx.x / 64] = cnt;
  __syncthreads();
  if (threadIdx.x == 0) {
    int64_t s = 0;
    for (int w = 0; w < (int)(blockDim.x / 64); ++w) s += wsum[w];
    counts[blockIdx.x] = s;
  }
}

// exclusive scan of counts[0..nb) in place by one workgroup; total -> *total
__global__ void k_scan_counts(int64_t* counts, int64_t nb, int64_t* total) {
  __shared__ int64_t part[1024];
  int t = threadIdx.x;
  int64_t per = (nb + blockDim.x - 1) / blockDim.x;
  int64_t lo = t * per, hi = std::min(nb, lo + per);
  int64_t s = 0;
  for (int64_t i = lo; i < hi; ++i) s += counts[i];
  part[t] = s;
  __syncthreads();
  if (t == 0) {
    int64_t acc = 0;
    for (int i = 0; i < (int)blockDim.x; ++i) {
      int64_t v = part[i];
      part[i] = acc;
      acc += v;
    }
    *total = acc;
  }
  __syncthreads();
  int64_t acc = part[t];
  for (int64_t i = lo; i < hi; ++i) {
    int64_t v = counts[i];
    counts[i] = acc;
    acc += v;
  }
}

// one workgroup of 1024 lanes per block of flags, 1 flag per lane, stable
__global__ void __launch_bounds__(CB) k_scatter_idx(const uint8_t* f, int64_t n, int invert,
                                                    const int64_t* offsets, int64_t* out) {
  __shared__ int wsum[CB / 64];
  int64_t i = (int64_t)blockIdx.x * CB + threadIdx.x;
  bool on = (i < n) && (((f[i] != 0) ^ invert) != 0);
  uint64_t m = __ballot(on);
  int lane = wave_lane();
  int w = threadIdx.x / 64;
  int rank = __popcll(m & ((1ull << lane) - 1ull));
  if (lane == 0) wsum[w] = __popcll(m);
  __syncthreads();
  int woff = 0;
  for (int k = 0; k < w; ++k) woff += wsum[k];
  if (on) out[offsets[blockIdx.x] + woff + rank] = i;
}

static thread_local DevBuf g_cnt_buf;

// indices of flags (== 1 ^ invert) into out (device); count to *count_host
int sfp_compact_indices(const uint8_t* f, int64_t n, int invert, int64_t* out, int64_t* count_host,
                           hipStream_t st) {
  if (n <= 0) {
    *count_host = 0;
    return SFP_OK;
  }
  int64_t nb = (n + CB - 1) / CB;
  CHK(g_cnt_buf.ensure(sizeof(int64_t) * (nb + 1)));
  int64_t* counts = (int64_t*)g_cnt_buf.p;
  hipLaunchKernelGGL(k_block_count, dim3((unsigned)nb), dim3(256), 0, st, f, n, invert, counts);
  hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(1024), 0, st, counts, nb, counts + nb);
  if (out)
    hipLaunchKernelGGL(k_scatter_idx, dim3((unsigned)nb), dim3(CB), 0, st, f, n, invert, counts,
                       out);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(count_host, counts + nb, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  return SFP_OK;
}

extern "C" int sfp_find_event_index(const uint8_t* event, int64_t n, int64_t* event_index,
                                    int64_t* count_host, sfp_stream_t stream) {
  return sfp_compact_indices(event, n, 0, event_index, count_host, S(stream));
}

extern "C" int sfp_bool_sum(const uint8_t* flags, int64_t n, int64_t* count_host,
                            sfp_stream_t stream) {
  return sfp_compact_indices(flags, n, 0, nullptr, count_host, S(stream));
}

// ---- the same compaction without allocation --------------------------------
// For callers that cannot allocate (code under graph capture, tight serving
// loops): the block counts live in `scratch`, a static device array that each
// GPU holds for the life of the process.  Host code never passes the array's
// name to a kernel (on the host it is only a stand-in for the symbol): the
// device address comes from hipGetSymbolAddress, once per device, with that
// device current.  Calls that overlap on one GPU share the array and race.
static constexpr int64_t NOALLOC_MAX_BLOCKS = 1 << 16;  // flags up to 2^16 * CB = 2^26
__device__ __align__(256) int64_t scratch[NOALLOC_MAX_BLOCKS];

static constexpr int NOALLOC_MAX_DEVICES = 64;
static int64_t* g_scratch_dev[NOALLOC_MAX_DEVICES];
static std::mutex g_scratch_mu;

static int scratch_of_current_device(int64_t** out) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= NOALLOC_MAX_DEVICES)
    return set_err(SFP_E_ARG, "device %d beyond the %d supported", dev, NOALLOC_MAX_DEVICES);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_dev[dev] = (int64_t*)p;
  }
  *out = g_scratch_dev[dev];
  return SFP_OK;
}

extern "C" int sfp_find_event_index_noalloc(const uint8_t* event, int64_t n,
                                            int64_t* event_index, int64_t* count_dev,
                                            sfp_stream_t stream) {
  if (n < 0) return set_err(SFP_E_ARG, "n must be >= 0");
  if (n > NOALLOC_MAX_BLOCKS * CB)
    return set_err(SFP_E_ARG, "%lld flags: at most %lld without allocation", (long long)n,
                   (long long)(NOALLOC_MAX_BLOCKS * CB));
  if (!count_dev) return set_err(SFP_E_ARG, "count_dev must be a device pointer");
  int64_t* counts = nullptr;
  CHK(scratch_of_current_device(&counts));
  hipStream_t st = S(stream);
  const int64_t nb = (n + CB - 1) / CB;
  if (nb == 0) {  // no flags: the count is 0 (a kernel, not a memset, so that capture works)
    hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(1024), 0, st, counts, (int64_t)0, count_dev);
  } else {
    hipLaunchKernelGGL(k_block_count, dim3((unsigned)nb), dim3(256), 0, st, event, n, 0, counts);
    hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(1024), 0, st, counts, nb, count_dev);
    if (event_index)
      hipLaunchKernelGGL(k_scatter_idx, dim3((unsigned)nb), dim3(CB), 0, st, event, n, 0, counts,
                         event_index);
  }
  HIPCHK(hipGetLastError());
  return SFP_OK;
}

// ------------------------------------------------------ QED per-op kernels --
__global__ void k_rs(const double* ig, const double* chi, double dt, int64_t n,
                     const uint8_t* pruned, uint8_t* event, double* delta, const uint64_t* ids,
                     uint64_t seed, uint64_t step, uint32_t tag, uint32_t slot,
                     const sfp_rs_node* nodes, int64_t z_N, double z_min, double z_max,
                     double z_delta, int photon) {
  KLOOP(n) {
    double c = chi[ip];
    if (pruned[ip] || c == 0.0) {
      event[ip] = 0;
      continue;
    }
    sfp_u32x4 d = sfp_draw4(seed, ids[ip], step, tag, slot);
    double r1 = sfp_u01(d.v[0], d.v[1]);
    double r2 = sfp_u01(d.v[2], d.v[3]);
    double dtau = dt * ig[ip];
    if (photon) {
      double r1_3 = sfp_powi(r1, 3);
      double p = 3 * sfp_powi(r1, 2) *
                 sfp_rs_rate(nodes, z_N, z_min, z_max, z_delta, c, r1_3, 1) * dtau;
      if (r2 < p) { delta[ip] = r1_3; event[ip] = 1; }
      else { delta[ip] = 0.0; event[ip] = 0; }
    } else {
      double p = sfp_rs_rate(nodes, z_N, z_min, z_max, z_delta, c, r1, 0) * dtau;
      if (r2 < p) { delta[ip] = r1; event[ip] = 1; }
      else { delta[ip] = 0.0; event[ip] = 0; }
    }
  }
}

static int rs_launch(const double* ig, const double* chi, double dt, int64_t n,
                     const uint8_t* pruned, uint8_t* event, double* delta, const uint64_t* ids,
                     uint64_t seed, uint64_t step, uint32_t tag, uint32_t slot, void* tables,
                     sfp_stream_t stream, int photon) {
  Tables* T = (Tables*)tables;
  if (!T || !T->has_rs) return set_err(SFP_E_ARG, "rejection-sampling tables not loaded");
  if (n <= 0) return SFP_OK;
  hipLaunchKernelGGL(k_rs, dim3(op_grid(n)), dim3(OPB), 0, S(stream), ig, chi, dt, n, pruned,
                     event, delta, ids, seed, step, tag, slot, T->rs_nodes, T->z_N, T->z_min,
                     T->z_max, T->z_delta, photon);
  HIPCHK(hipGetLastError());
  return SFP_OK;
}

extern "C" int sfp_photon_from_rejection_sampling(const double* ig, const double* chi, double dt,
                                                  int64_t n, const uint8_t* pruned,
                                                  uint8_t* event, double* delta,
                                                  const uint64_t* ids, uint64_t seed,
                                                  uint64_t step, uint32_t tag, uint32_t slot,
                                                  void* tables, sfp_stream_t stream) {
  return rs_launch(ig, chi, dt, n, pruned, event, delta, ids, seed, step, tag, slot, tables,
                   stream, 1);
}

extern "C" int sfp_pair_from_rejection_sampling(const double* ig, const double* chi, double dt,
                                                int64_t n, const uint8_t* pruned, uint8_t* event,
                                                double* delta, const uint64_t* ids,
                                                uint64_t seed, uint64_t step, uint32_t tag,
                                                uint32_t slot, void* tables,
                                                sfp_stream_t stream) {
  return rs_launch(ig, chi, dt, n, pruned, event, delta, ids, seed, step, tag, slot, tables,
                   stream, 0);
}

__global__ void k_od(double* tau, const double* ig, const double* chi, double dt, int64_t n,
                     const uint8_t* pruned, uint8_t* event, double* delta, const uint64_t* ids,
                     uint64_t seed, uint64_t step, uint32_t tag, uint32_t slot,
                     const double2* nodes, const double* cdf, sfp_od_dev od) {
  KLOOP(n) {
    double c = chi[ip];
    if (pruned[ip] || !(c >= od.chi_min)) {  // NaN chi: no event
      event[ip] = 0;
      delta[ip] = 0.0;
      continue;
    }
    double t = tau[ip], dl;
    bool ev = od_update(&t, ig[ip], c, dt, seed, ids[ip], step, tag, slot, nodes, cdf, od, &dl);
    tau[ip] = t;
    event[ip] = ev ? 1 : 0;
    delta[ip] = dl;
  }
}

static int od_launch(double* tau, const double* ig, const double* chi, double dt, int64_t n,
                     const uint8_t* pruned, uint8_t* event, double* delta, const uint64_t* ids,
                     uint64_t seed, uint64_t step, uint32_t tag, uint32_t slot, void* tables,
                     sfp_stream_t stream, int photon) {
  Tables* T = (Tables*)tables;
  if (!T || !T->has_od) return set_err(SFP_E_ARG, "optical-depth tables not loaded");
  if (n <= 0) return SFP_OK;
  hipLaunchKernelGGL(k_od, dim3(op_grid(n)), dim3(OPB), 0, S(stream), tau, ig, chi, dt, n, pruned,
                     event, delta, ids, seed, step, tag, slot,
                     photon ? T->od_photon_nodes : T->od_pair_nodes,
                     photon ? T->od.photon_cdf : T->od.pair_cdf, T->od);
  HIPCHK(hipGetLastError());
  return SFP_OK;
}

extern "C" int sfp_update_tau_e(double* tau, const double* ig, const double* chi, double dt,
                                int64_t n, const uint8_t* pruned, uint8_t* event, double* delta,
                                const uint64_t* ids, uint64_t seed, uint64_t step, uint32_t tag,
                                uint32_t slot, void* tables, sfp_stream_t stream) {
  return od_launch(tau, ig, chi, dt, n, pruned, event, delta, ids, seed, step, tag, slot, tables,
                   stream, 1);
}

extern "C" int sfp_update_tau_gamma(double* tau, const double* ig, const double* chi, double dt,
                                    int64_t n, const uint8_t* pruned, uint8_t* event,
                                    double* delta, const uint64_t* ids, uint64_t seed,
                                    uint64_t step, uint32_t tag, uint32_t slot, void* tables,
                                    sfp_stream_t stream) {
  return od_launch(tau, ig, chi, dt, n, pruned, event, delta, ids, seed, step, tag, slot, tables,
                   stream, 0);
}

// ------------------------------------------------------------ tables API --
extern "C" int sfp_tables_create(const sfp_rs_tables* rs, const sfp_od_tables* od, void** out) {
  Tables* T = new Tables();
  if (rs) {
    if (rs->z_N < 2) return set_err(SFP_E_ARG, "rejection-sampling table needs >= 2 nodes");
    std::vector<sfp_rs_node> nodes(rs->z_N - 1);
    for (int64_t i = 0; i + 1 < rs->z_N; ++i) {
      nodes[i].ia = rs->int_Ai[i];
      nodes[i].ka = (rs->int_Ai[i + 1] - rs->int_Ai[i]) / rs->z_delta;
      nodes[i].ap = rs->Aip[i];
      nodes[i].kp = (rs->Aip[i + 1] - rs->Aip[i]) / rs->z_delta;
    }
    HIPCHK(hipMalloc(&T->rs_nodes, sizeof(sfp_rs_node) * nodes.size()));
    HIPCHK(hipMemcpy(T->rs_nodes, nodes.data(), sizeof(sfp_rs_node) * nodes.size(),
                     hipMemcpyHostToDevice));
    // the same nodes rounded to FP32 for the mixed-precision step
    std::vector<float4> n32(nodes.size());
    for (size_t i = 0; i < nodes.size(); ++i)
      n32[i] = make_float4((float)nodes[i].ia, (float)nodes[i].ka, (float)nodes[i].ap,
                           (float)nodes[i].kp);
    HIPCHK(hipMalloc(&T->rs_nodes32, sizeof(float4) * n32.size()));
    HIPCHK(hipMemcpy(T->rs_nodes32, n32.data(), sizeof(float4) * n32.size(),
                     hipMemcpyHostToDevice));
    // Envelopes used to prove that an emission/pair rate is negligible.  On
    // bracket j (z in [z_j, z_j+1], linear interpolation of the tables) the
    // interpolated int Ai and |Ai'| are bounded by their node maxima, so for
    // every z' >= z_j and c >= 0:
    //   |int Ai(z') + (2/z' +- c sqrt z') Ai'(z')| <= E1[j] + c E2[j]
    // with E1[j] = max_{k>=j} (max int Ai + (2/z_k) max|Ai'|) and
    //      E2[j] = max_{k>=j} sqrt(z_k+1) max|Ai'|  (suffix maxima).
    // Bracket 0 (z_0 = 0) has no bound.
    std::vector<double2> env(rs->z_N - 1);
    double e1 = 0.0, e2 = 0.0;
    for (int64_t j = rs->z_N - 2; j >= 0; --j) {
      double zj = rs->z_min + (double)j * rs->z_delta;
      double zj1 = rs->z_min + (double)(j + 1) * rs->z_delta;
      double ia = std::max(rs->int_Ai[j], rs->int_Ai[j + 1]);
      double ap = std::max(fabs(rs->Aip[j]), fabs(rs->Aip[j + 1]));
      double a = (zj > 0.0) ? ia + 2.0 / zj * ap : __builtin_inf();
      double b = sqrt(zj1) * ap;
      e1 = std::max(e1, a);
      e2 = std::max(e2, b);
      env[j] = make_double2(e1, e2);
    }
    HIPCHK(hipMalloc(&T->rs_env, sizeof(double2) * env.size()));
    HIPCHK(hipMemcpy(T->rs_env, env.data(), sizeof(double2) * env.size(), hipMemcpyHostToDevice));
    if (rs->Ai) {
      // Ai nodes for the spin-resolved rate; its spin term is bounded on
      // bracket j by chi * z_j+1 * max Ai (Ai > 0 and the interpolation stays
      // between node values), added to E2 as a suffix maximum
      std::vector<double2> ai(rs->z_N - 1);
      std::vector<float2> ai32(rs->z_N - 1);
      std::vector<double2> envs(rs->z_N - 1);
      double e3 = 0.0;
      for (int64_t j = rs->z_N - 2; j >= 0; --j) {
        ai[j] = make_double2(rs->Ai[j], (rs->Ai[j + 1] - rs->Ai[j]) / rs->z_delta);
        ai32[j] = make_float2((float)ai[j].x, (float)ai[j].y);
        double zj1 = rs->z_min + (double)(j + 1) * rs->z_delta;
        e3 = std::max(e3, zj1 * std::max(fabs(rs->Ai[j]), fabs(rs->Ai[j + 1])));
        envs[j] = make_double2(env[j].x, env[j].y + e3);
      }
      HIPCHK(hipMalloc(&T->rs_ai, sizeof(double2) * ai.size()));
      HIPCHK(hipMemcpy(T->rs_ai, ai.data(), sizeof(double2) * ai.size(), hipMemcpyHostToDevice));
      HIPCHK(hipMalloc(&T->rs_ai32, sizeof(float2) * ai32.size()));
      HIPCHK(hipMemcpy(T->rs_ai32, ai32.data(), sizeof(float2) * ai32.size(),
                       hipMemcpyHostToDevice));
      HIPCHK(hipMalloc(&T->rs_env_spin, sizeof(double2) * envs.size()));
      HIPCHK(hipMemcpy(T->rs_env_spin, envs.data(), sizeof(double2) * envs.size(),
                       hipMemcpyHostToDevice));
    }
    T->z_N = rs->z_N;
    T->z_min = rs->z_min;
    T->z_max = rs->z_max;
    T->z_delta = rs->z_delta;
    T->has_rs = true;
  }
  if (od) {
    int64_t cn = od->chi_N, dn = od->delta_N;
    if (cn < 2 || dn < 2) return set_err(SFP_E_ARG, "optical-depth tables need >= 2 nodes");
    size_t total = 2 * cn + 2 * cn * dn + cn + dn;
    HIPCHK(hipMalloc(&T->od_mem, sizeof(double) * total));
    double* p = T->od_mem;
    auto up = [&](const double* src, size_t count) -> double* {
      double* d = p;
      (void)hipMemcpy(d, src, sizeof(double) * count, hipMemcpyHostToDevice);
      p += count;
      return d;
    };
    T->od.photon_total = up(od->photon_total, cn);
    T->od.pair_total = up(od->pair_total, cn);
    T->od.photon_cdf = up(od->photon_cdf, cn * dn);
    T->od.pair_cdf = up(od->pair_cdf, cn * dn);
    T->od.chi_grid = up(od->chi_grid, cn);
    T->od.delta_grid = up(od->delta_grid, dn);
    HIPCHK(hipGetLastError());
    T->od.chi_N = cn;
    T->od.delta_N = dn;
    T->od.log_chi_min = od->log_chi_min;
    T->od.log_chi_max = od->log_chi_max;
    T->od.log_chi_delta = od->log_chi_delta;
    T->od.chi_min = od->chi_min;
    T->od.chi_max = od->chi_max;
    std::vector<double2> pn(cn - 1), qn(cn - 1);
    for (int64_t i = 0; i + 1 < cn; ++i) {
      pn[i] = make_double2(od->photon_total[i],
                           (od->photon_total[i + 1] - od->photon_total[i]) / od->log_chi_delta);
      qn[i] = make_double2(od->pair_total[i],
                           (od->pair_total[i + 1] - od->pair_total[i]) / od->log_chi_delta);
    }
    HIPCHK(hipMalloc(&T->od_photon_nodes, sizeof(double2) * (cn - 1)));
    HIPCHK(hipMalloc(&T->od_pair_nodes, sizeof(double2) * (cn - 1)));
    HIPCHK(hipMemcpy(T->od_photon_nodes, pn.data(), sizeof(double2) * (cn - 1),
                     hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(T->od_pair_nodes, qn.data(), sizeof(double2) * (cn - 1),
                     hipMemcpyHostToDevice));
    T->has_od = true;
  }
  *out = T;
  return SFP_OK;
}

extern "C" int sfp_tables_destroy(void* tables) {
  Tables* T = (Tables*)tables;
  if (!T) return SFP_OK;
  if (T->rs_nodes) (void)hipFree(T->rs_nodes);
  if (T->rs_nodes32) (void)hipFree(T->rs_nodes32);
  if (T->rs_env) (void)hipFree(T->rs_env);
  if (T->rs_ai) (void)hipFree(T->rs_ai);
  if (T->rs_ai32) (void)hipFree(T->rs_ai32);
  if (T->rs_env_spin) (void)hipFree(T->rs_env_spin);
  if (T->od_mem) (void)hipFree(T->od_mem);
  if (T->od_photon_nodes) (void)hipFree(T->od_photon_nodes);
  if (T->od_pair_nodes) (void)hipFree(T->od_pair_nodes);
  delete T;
  return SFP_OK;
}

// --------------------------------------------------------------- stores ---
const size_t kAttrSize[SFP_NATTR] = {8, 8, 8, 8, 8, 8, 8, 1, 8, 8, 8, 8,
                                     8, 8, 8, 8, 8, 8, 8, 8, 1, 8, 4};

static bool attr_core(int a) { return a <= SFP_PRUNED; }

static int store_alloc_attr(Store* s, int a) {
  if (s->p[a]) return SFP_OK;
  size_t bytes = kAttrSize[a] * (size_t)std::max<int64_t>(s->cap, 1);
  HIPCHK(hipMalloc(&s->p[a], bytes));
  if (a == SFP_PRUNED) {
    HIPCHK(hipMemset(s->p[a], 1, bytes));
  } else {
    HIPCHK(hipMemset(s->p[a], 0, bytes));
  }
  return SFP_OK;
}

extern "C" int sfp_store_create(int32_t flags, int64_t capacity, void** out) {
  if (capacity < 0) return set_err(SFP_E_ARG, "negative capacity");
  Store* s = new Store();
  s->flags = flags;
  s->cap = capacity;
  HIPCHK(hipMalloc(&s->cnt, sizeof(Counts)));
  HIPCHK(hipMemset(s->cnt, 0, sizeof(Counts)));
  for (int a = 0; a < SFP_NATTR; ++a) {
    bool want = attr_core(a) || (a == SFP_TAU && (flags & SFP_STORE_TAU)) ||
                ((a == SFP_SX || a == SFP_SY || a == SFP_SZ) && (flags & SFP_STORE_SPIN));
    if (want) CHK(store_alloc_attr(s, a));
  }
  *out = s;
  return SFP_OK;
}

extern "C" int sfp_store_destroy(void* store) {
  Store* s = (Store*)store;
  if (!s) return SFP_OK;
  (void)store_settle(s);
  (void)hipDeviceSynchronize();
  for (int a = 0; a < SFP_NATTR; ++a)
    if (s->p[a]) (void)hipFree(s->p[a]);
  if (s->cnt) (void)hipFree(s->cnt);
  delete s;
  return SFP_OK;
}

extern "C" int sfp_store_reserve(void* store, int64_t capacity) {
  Store* s = (Store*)store;
  if (!s) return set_err(SFP_E_ARG, "null store");
  if (capacity <= s->cap) return SFP_OK;
  CHK(store_settle(s));
  s->version++;
  HIPCHK(hipDeviceSynchronize());
  for (int a = 0; a < SFP_NATTR; ++a) {
    if (!s->p[a]) continue;
    void* np = nullptr;
    size_t nb = kAttrSize[a] * (size_t)capacity;
    size_t ob = kAttrSize[a] * (size_t)s->cap;
    HIPCHK(hipMalloc(&np, nb));
    if (ob) HIPCHK(hipMemcpy(np, s->p[a], ob, hipMemcpyDeviceToDevice));
    int fill = (a == SFP_PRUNED) ? 1 : 0;
    HIPCHK(hipMemset((char*)np + ob, fill, nb - ob));
    HIPCHK(hipFree(s->p[a]));
    s->p[a] = np;
  }
  s->cap = capacity;
  return SFP_OK;
}

extern "C" int sfp_store_capacity(void* store, int64_t* cap) {
  Store* s = (Store*)store;
  if (!s) return set_err(SFP_E_ARG, "null store");
  *cap = s->cap;
  return SFP_OK;
}

extern "C" int sfp_store_ptr(void* store, int32_t attr, void** dev) {
  Store* s = (Store*)store;
  if (!s || attr < 0 || attr >= SFP_NATTR) return set_err(SFP_E_ARG, "bad store/attribute");
  CHK(store_alloc_attr(s, attr));
  CHK(store_settle(s));
  s->version++;
  *dev = s->p[attr];
  return SFP_OK;
}

extern "C" int sfp_store_set_count(void* store, int64_t n, uint64_t next_id) {
  Store* s = (Store*)store;
  if (!s) return set_err(SFP_E_ARG, "null store");
  if (n < 0 || n > s->cap) return set_err(SFP_E_ARG, "count %lld outside capacity %lld",
                                           (long long)n, (long long)s->cap);
  Counts c{};
  HIPCHK(hipDeviceSynchronize());
  HIPCHK(hipMemcpy(&c, s->cnt, sizeof(c), hipMemcpyDeviceToHost));
  if (c.n != n || c.next_id != (int64_t)next_id) {
    CHK(store_settle(s));
    s->version++;
  }
  c.n = n;
  c.next_id = (int64_t)next_id;
  c.n_old = n;
  c.overflow = 0;
  HIPCHK(hipMemcpy(s->cnt, &c, sizeof(c), hipMemcpyHostToDevice));
  return SFP_OK;
}

extern "C" int sfp_store_get_count(void* store, int64_t* n, uint64_t* next_id) {
  Store* s = (Store*)store;
  if (!s) return set_err(SFP_E_ARG, "null store");
  Counts c{};
  HIPCHK(hipDeviceSynchronize());
  HIPCHK(hipMemcpy(&c, s->cnt, sizeof(c), hipMemcpyDeviceToHost));
  if (n) *n = c.n;
  if (next_id) *next_id = (uint64_t)c.next_id;
  return SFP_OK;
}

extern "C" int sfp_store_upload(void* store, int32_t attr, const void* host, int64_t off,
                                int64_t n) {
  Store* s = (Store*)store;
  if (!s || attr < 0 || attr >= SFP_NATTR) return set_err(SFP_E_ARG, "bad store/attribute");
  if (off < 0 || n < 0 || off + n > s->cap) return set_err(SFP_E_ARG, "upload out of range");
  CHK(store_alloc_attr(s, attr));
  if (n == 0) return SFP_OK;
  CHK(store_settle(s));
  s->version++;
  HIPCHK(hipDeviceSynchronize());
  HIPCHK(hipMemcpy((char*)s->p[attr] + kAttrSize[attr] * off, host, kAttrSize[attr] * n,
                   hipMemcpyHostToDevice));
  return SFP_OK;
}

extern "C" int sfp_store_download(void* store, int32_t attr, void* host, int64_t off,
                                  int64_t n) {
  Store* s = (Store*)store;
  if (!s || attr < 0 || attr >= SFP_NATTR) return set_err(SFP_E_ARG, "bad store/attribute");
  if (off < 0 || n < 0 || off + n > s->cap) return set_err(SFP_E_ARG, "download out of range");
  CHK(store_alloc_attr(s, attr));
  if (n == 0) return SFP_OK;
  CHK(store_settle(s));
  HIPCHK(hipDeviceSynchronize());
  HIPCHK(hipMemcpy(host, (char*)s->p[attr] + kAttrSize[attr] * off, kAttrSize[attr] * n,
                   hipMemcpyDeviceToHost));
  return SFP_OK;
}

template <typename T>
__global__ void k_gather(const T* src, T* dst, const int64_t* idx, int64_t n) {
  KLOOP(n) { dst[ip] = src[idx[ip]]; }
}

extern "C" int sfp_store_prune(void* store, int64_t* n_kept) {
  Store* s = (Store*)store;
  if (!s) return set_err(SFP_E_ARG, "null store");
  CHK(store_settle(s));
  s->version++;
  int64_t n = 0;
  uint64_t next_id = 0;
  CHK(sfp_store_get_count(s, &n, &next_id));
  int64_t kept = 0;
  int64_t* idx = nullptr;
  if (n > 0) {
    HIPCHK(hipMalloc(&idx, sizeof(int64_t) * n));
    CHK(sfp_compact_indices((const uint8_t*)s->p[SFP_PRUNED], n, 1, idx, &kept, nullptr));
    for (int a = 0; a < SFP_NATTR; ++a) {
      if (!s->p[a] || kept == 0) continue;
      void* tmp = nullptr;
      HIPCHK(hipMalloc(&tmp, kAttrSize[a] * (size_t)kept));
      if (kAttrSize[a] == 8)
        hipLaunchKernelGGL(k_gather<uint64_t>, dim3(op_grid(kept)), dim3(OPB), 0, 0,
                           (const uint64_t*)s->p[a], (uint64_t*)tmp, idx, kept);
      else if (kAttrSize[a] == 4)
        hipLaunchKernelGGL(k_gather<uint32_t>, dim3(op_grid(kept)), dim3(OPB), 0, 0,
                           (const uint32_t*)s->p[a], (uint32_t*)tmp, idx, kept);
      else
        hipLaunchKernelGGL(k_gather<uint8_t>, dim3(op_grid(kept)), dim3(OPB), 0, 0,
                           (const uint8_t*)s->p[a], (uint8_t*)tmp, idx, kept);
      HIPCHK(hipGetLastError());
      HIPCHK(hipMemcpy(s->p[a], tmp, kAttrSize[a] * (size_t)kept, hipMemcpyDeviceToDevice));
      HIPCHK(hipFree(tmp));
    }
    // slots past the kept range become holes
    HIPCHK(hipMemset((uint8_t*)s->p[SFP_PRUNED] + kept, 1, (size_t)(s->cap - kept)));
    HIPCHK(hipFree(idx));
  }
  Counts c{};
  HIPCHK(hipMemcpy(&c, s->cnt, sizeof(c), hipMemcpyDeviceToHost));
  c.n = kept;
  c.n_old = kept;
  HIPCHK(hipMemcpy(s->cnt, &c, sizeof(c), hipMemcpyHostToDevice));
  *n_kept = kept;
  return SFP_OK;
}

// --------------------------------------------------------- diagnostics ----
// Per-shot counts and energy histograms of live particles: row k of out holds
// the count of shot k, then nbins bins of log10(energy), energy = |u| for
// photons and sqrt(1 + u^2) otherwise (m c^2 units); values outside
// [lmin, lmax) are counted in the total only.  Rows are privatised in LDS when
// they fit.
static constexpr int HIST_LDS = 4096;

__global__ void __launch_bounds__(OPB)
k_histogram(const double* __restrict__ ux, const double* __restrict__ uy,
            const double* __restrict__ uz, const uint8_t* __restrict__ pruned,
            const uint32_t* __restrict__ shot, int64_t n, int32_t nshots, int32_t nbins,
            double lmin, double inv_w, int photon, unsigned long long* __restrict__ out) {
  __shared__ uint32_t sh[HIST_LDS];
  const int64_t cells = (int64_t)nshots * (nbins + 1);
  const bool lds = cells <= HIST_LDS;
  if (lds) {
    for (int64_t c = threadIdx.x; c < cells; c += blockDim.x) sh[c] = 0;
    __syncthreads();
  }
  KLOOP(n) {
    if (pruned[ip]) continue;
    const uint32_t k = shot ? min(shot[ip], (uint32_t)nshots - 1u) : 0u;
    const int64_t row = (int64_t)k * (nbins + 1);
    int64_t cell[2] = {row, -1};
    if (nbins > 0) {
      const double u2 = ux[ip] * ux[ip] + uy[ip] * uy[ip] + uz[ip] * uz[ip];
      const double en = photon ? sqrt(u2) : sqrt(1.0 + u2);
      const double b = (log10(en) - lmin) * inv_w;
      if (b >= 0.0 && b < (double)nbins) cell[1] = row + 1 + (int64_t)b;
    }
    for (int j = 0; j < 2; ++j) {
      if (cell[j] < 0) continue;
      if (lds)
        atomicAdd(&sh[cell[j]], 1u);
      else
        atomicAdd(&out[cell[j]], 1ull);
    }
  }
  if (lds) {
    __syncthreads();
    for (int64_t c = threadIdx.x; c < cells; c += blockDim.x)
      if (sh[c]) atomicAdd(&out[c], (unsigned long long)sh[c]);
  }
}

extern "C" int sfp_store_histogram(void* store, int32_t photon, int32_t nshots, int32_t nbins,
                                   double log10_min, double log10_max, int64_t* out) {
  Store* s = (Store*)store;
  if (!s || !out) return set_err(SFP_E_ARG, "null store/output");
  if (nshots < 1 || nbins < 0 || (int64_t)nshots * (nbins + 1) > ((int64_t)1 << 28))
    return set_err(SFP_E_ARG, "bad histogram shape");
  if (nbins > 0 && !(log10_max > log10_min)) return set_err(SFP_E_ARG, "empty energy range");
  int64_t n = 0;
  uint64_t next_id = 0;
  CHK(sfp_store_get_count(s, &n, &next_id));
  const size_t bytes = sizeof(uint64_t) * (size_t)nshots * (nbins + 1);
  unsigned long long* dev = nullptr;
  HIPCHK(hipMalloc(&dev, bytes));
  HIPCHK(hipMemset(dev, 0, bytes));
  if (n > 0) {
    const double inv_w = nbins > 0 ? nbins / (log10_max - log10_min) : 0.0;
    hipLaunchKernelGGL(k_histogram, dim3(op_grid(n)), dim3(OPB), 0, 0, (const double*)s->p[SFP_UX],
                       (const double*)s->p[SFP_UY], (const double*)s->p[SFP_UZ],
                       (const uint8_t*)s->p[SFP_PRUNED], (const uint32_t*)s->p[SFP_SHOT], n,
                       nshots, nbins, log10_min, inv_w, photon, dev);
    HIPCHK(hipGetLastError());
  }
  HIPCHK(hipMemcpy(out, dev, bytes, hipMemcpyDeviceToHost));
  HIPCHK(hipFree(dev));
  return SFP_OK;
}

// plain streaming copy of n doubles: the reference point for HBM byte counters
__global__ void k_copy(const double* __restrict__ src, double* __restrict__ dst, int64_t n) {
  KLOOP(n) { dst[ip] = src[ip]; }
}

extern "C" int sfp_copy_f64(const double* src, double* dst, int64_t n, sfp_stream_t stream) {
  if (n <= 0) return SFP_OK;
  int g = (int)std::min<int64_t>((n + OPB - 1) / OPB, 1 << 20);
  hipLaunchKernelGGL(k_copy, dim3(g), dim3(OPB), 0, S(stream), src, dst, n);
  HIPCHK(hipGetLastError());
  return SFP_OK;
}

// ---------------------------------------------------------------- misc ----
extern "C" int sfp_malloc(void** dev, int64_t bytes) {
  if (bytes < 0) return set_err(SFP_E_ARG, "negative size");
  HIPCHK(hipMalloc(dev, (size_t)std::max<int64_t>(bytes, 1)));
  return SFP_OK;
}

extern "C" int sfp_free(void* dev) {
  // work already queued may still read the buffer (a per-operation kernel
  // launched on it): wait for the device before the memory goes back
  if (dev) HIPCHK(hipDeviceSynchronize());
  if (dev) HIPCHK(hipFree(dev));
  return SFP_OK;
}

extern "C" int sfp_memcpy_htod(void* dst, const void* src, int64_t bytes) {
  if (bytes <= 0) return SFP_OK;
  HIPCHK(hipDeviceSynchronize());
  HIPCHK(hipMemcpy(dst, src, (size_t)bytes, hipMemcpyHostToDevice));
  return SFP_OK;
}

extern "C" int sfp_memcpy_dtoh(void* dst, const void* src, int64_t bytes) {
  if (bytes <= 0) return SFP_OK;
  HIPCHK(hipDeviceSynchronize());
  HIPCHK(hipMemcpy(dst, src, (size_t)bytes, hipMemcpyDeviceToHost));
  return SFP_OK;
}

extern "C" int sfp_synchronize(void) {
  HIPCHK(hipDeviceSynchronize());
  return SFP_OK;
}

extern "C" int sfp_mem_info(int64_t* free_bytes, int64_t* total_bytes) {
  size_t fr = 0, tot = 0;
  HIPCHK(hipMemGetInfo(&fr, &tot));
  if (free_bytes) *free_bytes = (int64_t)fr;
  if (total_bytes) *total_bytes = (int64_t)tot;
  return SFP_OK;
}

extern "C" const char* sfp_last_error(void) { return g_err.c_str(); }
extern "C" int sfp_version(void) { return 1; }

extern "C" int sfp_device_info(int32_t* num_cus, char* name, int32_t name_len) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  HIPCHK(hipGetDeviceProperties(&prop, dev));
  if (num_cus) *num_cus = prop.multiProcessorCount;
  if (name && name_len > 0) {
    strncpy(name, prop.gcnArchName, name_len - 1);
    name[name_len - 1] = 0;
  }
  return SFP_OK;
}
