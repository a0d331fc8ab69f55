// sfp_lib.hip -- libsfp_hip.so: the sfparticles per-particle step on MI355X (gfx950).
//
// Layout in HBM: one struct-of-arrays store per species (x, y, z, ux, uy, uz
// as float64, a uint64 id and a uint8 pruned flag; tau and spin when the
// species needs them).  inv_gamma, chi, E, B, the event flag and delta are
// transients of the fused step and live in registers; the per-operation
// entry points materialise them as auxiliary store attributes.
//
// The fused step launches one persistent kernel per species and step.  Each
// workgroup takes 256-particle tiles from an atomic tile counter; per tile it
// evaluates the QED front (field at t, chi, emission/pair draw, hard-photon
// cut), computes the stable rank of every event with a wave64 ballot +
// workgroup scan + decoupled look-back across tiles, appends the new
// particles (already pushed, as the reference pushes them in the step they
// are created), applies recoil and pushes the tile.
//
// All physics is compiled with -ffp-contract=off and uses the project's own
// elementary functions (sfp_math.h), so every FP64 result is bit-identical
// to the CPU restatement used for validation.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "sfp_device.h"

// ----------------------------------------------------------------- errors --
static thread_local std::string g_err;

static int set_err(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
static int set_err(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIPCHK(expr)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess)                                                              \
      return set_err(SFP_E_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                     __FILE__, __LINE__);                                              \
  } while (0)

#define CHK(expr)          \
  do {                     \
    int rc_ = (expr);      \
    if (rc_) return rc_;   \
  } while (0)

static inline hipStream_t S(sfp_stream_t s) { return (hipStream_t)s; }

static inline int grid_for(int64_t n, int block) {
  int64_t g = (n + block - 1) / block;
  if (g < 1) g = 1;
  if (g > (int64_t)INT32_MAX) g = INT32_MAX;
  return (int)g;
}

// --------------------------------------------------------------- tables ---
struct Tables {
  sfp_rs_node* rs_nodes = nullptr;  // [z_N-1]
  int64_t z_N = 0;
  double z_min = 0, z_max = 0, z_delta = 0;
  bool has_rs = false;
  // optical depth
  double* od_mem = nullptr;
  sfp_od_dev od{};
  double2* od_photon_nodes = nullptr;  // [chi_N-1] {value, slope}
  double2* od_pair_nodes = nullptr;
  bool has_od = false;
};

// ------------------------------------------------------------ per-op API --
#define KLOOP(n) for (int64_t ip = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; ip < (n); ip += (int64_t)gridDim.x * blockDim.x)

static constexpr int OPB = 256;

static int op_grid(int64_t n) {
  int64_t g = (n + OPB - 1) / OPB;
  if (g < 1) g = 1;
  if (g > 65536) g = 65536;
  return (int)g;
}

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

static int compile_fields(const sfp_field* terms, int32_t nterms, sfp_fieldset* fs) {
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
  CHK(compile_fields(terms, nterms, &fs));
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

__device__ inline int wave_lane() { return (int)__lane_id(); }

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

struct ScratchI64 {
  DevBuf buf;
};
static thread_local DevBuf g_cnt_buf;

// indices of flags (== 1 ^ invert) into out (device); count to *count_host
static int compact_indices(const uint8_t* f, int64_t n, int invert, int64_t* out, int64_t* count_host,
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
  return compact_indices(event, n, 0, event_index, count_host, S(stream));
}

extern "C" int sfp_bool_sum(const uint8_t* flags, int64_t n, int64_t* count_host,
                            sfp_stream_t stream) {
  return compact_indices(flags, n, 0, nullptr, count_host, S(stream));
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

// one optical-depth update (optical_depth.py:42-62); returns event
__device__ inline bool od_update(double* tau, double ig, double chi, double dt, uint64_t seed,
                                 uint64_t id, uint64_t step, uint32_t tag, uint32_t slot,
                                 const double2* nodes, const double* cdf, const sfp_od_dev& od,
                                 double* delta) {
  // total rate (optical_depth_tables.py:46-56) from {value, slope} nodes
  double rate;
  {
    double log_chi = sfp_log10(chi);
    if (log_chi > od.log_chi_max) log_chi = od.log_chi_max;
    if (log_chi < od.log_chi_min) {
      rate = 0.0;
    } else {
      double fidx = (log_chi - od.log_chi_min) / od.log_chi_delta;
      int64_t idx = (int64_t)sfp_floor(fidx);
      if (idx > od.chi_N - 2) idx = od.chi_N - 2;
      double left = od.log_chi_min + (double)idx * od.log_chi_delta;
      double2 nd = nodes[idx];
      rate = nd.x + nd.y * (log_chi - left);
    }
  }
  double dtau = dt * ig;
  double t = *tau;
  if (t == 0.0) {
    sfp_u32x4 d = sfp_draw4(seed, id, step, tag, slot + 1);
    t = -sfp_log(1 - sfp_u01(d.v[0], d.v[1]));
  }
  t -= rate * dtau;
  bool ev = false;
  if (t < 0) {
    sfp_u32x4 d = sfp_draw4(seed, id, step, tag, slot + 2);
    t = -sfp_log(1 - sfp_u01(d.v[0], d.v[1]));
    ev = true;
    *delta = sfp_od_inverse_cdf(cdf, &od, chi, sfp_u01(d.v[2], d.v[3]));
  } else {
    *delta = 0.0;
  }
  *tau = t;
  return ev;
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
  if (T->od_mem) (void)hipFree(T->od_mem);
  if (T->od_photon_nodes) (void)hipFree(T->od_photon_nodes);
  if (T->od_pair_nodes) (void)hipFree(T->od_pair_nodes);
  delete T;
  return SFP_OK;
}

// --------------------------------------------------------------- stores ---
struct Counts {
  int64_t n;         // buffered particles (live + pruned holes)
  int64_t next_id;   // id of the next appended particle
  int64_t n_old;     // n at the start of the current step
  int64_t overflow;  // number of appends that did not fit
  int64_t psteps;     // particle-steps advanced (live after each step)
  int64_t processed;  // particles read by this species' step kernel
  int64_t spawned;    // particles this species' step kernel appended (per destination)
  int64_t pad;
};

static const size_t kAttrSize[SFP_NATTR] = {8, 8, 8, 8, 8, 8, 8, 1, 8, 8, 8, 8,
                                            8, 8, 8, 8, 8, 8, 8, 8, 1, 8};

struct Store {
  int32_t flags = 0;
  int64_t cap = 0;
  void* p[SFP_NATTR] = {};
  Counts* cnt = nullptr;  // device
};

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
  int64_t n = 0;
  uint64_t next_id = 0;
  CHK(sfp_store_get_count(s, &n, &next_id));
  int64_t kept = 0;
  int64_t* idx = nullptr;
  if (n > 0) {
    HIPCHK(hipMalloc(&idx, sizeof(int64_t) * n));
    CHK(compact_indices((const uint8_t*)s->p[SFP_PRUNED], n, 1, idx, &kept, nullptr));
    for (int a = 0; a < SFP_NATTR; ++a) {
      if (!s->p[a] || kept == 0) continue;
      void* tmp = nullptr;
      HIPCHK(hipMalloc(&tmp, kAttrSize[a] * (size_t)kept));
      if (kAttrSize[a] == 8)
        hipLaunchKernelGGL(k_gather<uint64_t>, dim3(op_grid(kept)), dim3(OPB), 0, 0,
                           (const uint64_t*)s->p[a], (uint64_t*)tmp, idx, kept);
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

// ------------------------------------------------------- fused simulation --
static constexpr int TILE = 256;      // particles per tile = workgroup size
static constexpr int WAVES = TILE / 64;

struct SpeciesDev {
  double *x, *y, *z, *ux, *uy, *uz;
  uint64_t* id;
  uint8_t* pruned;
  double* tau;
  double *sx, *sy, *sz;
  Counts* cnt;
  int64_t cap;
  double q, ae;
  int32_t rr, push, spin, photon, tag;
  int32_t photon_dst, pair_e_dst, pair_p_dst;
};

struct ScanState {
  uint32_t tile_ctr;
  uint32_t done_ctr;
  uint32_t epoch;
  uint32_t pad;
};

#define SFP_MAX_SPECIES 8

struct Plan {
  SpeciesDev sp[SFP_MAX_SPECIES];
  ScanState* scan[SFP_MAX_SPECIES];
  uint64_t* flags[SFP_MAX_SPECIES];
  sfp_fieldset fields;
  const sfp_rs_node* rs_nodes;
  int64_t z_N;
  double z_min, z_max, z_delta;
  const double2* od_photon_nodes;
  const double2* od_pair_nodes;
  sfp_od_dev od;
  double threshold;
  uint64_t seed;
  int32_t nspecies;
  int32_t qed;
};

struct StepArgs {
  double t, t_half, dt, cdt_half, ll_factor;
  uint64_t step;
};

enum { KIND_PUSH = 0, KIND_RS = 1, KIND_OD = 2 };

__device__ inline uint64_t ld_acquire(const uint64_t* p) {
  return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void st_release(uint64_t* p, uint64_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// flag word: [epoch:24][status:2][value:38]; status 1 = aggregate, 2 = inclusive prefix
static constexpr int FLAG_VAL_BITS = 38;
static constexpr uint64_t FLAG_VAL_MASK = (1ull << FLAG_VAL_BITS) - 1ull;

__device__ inline uint64_t flag_word(uint32_t epoch, uint32_t status, uint64_t v) {
  return ((uint64_t)epoch << 40) | ((uint64_t)status << FLAG_VAL_BITS) | (v & FLAG_VAL_MASK);
}

// Decoupled look-back (one full wave): publish the tile aggregate, accumulate
// predecessors until an inclusive prefix is found, publish the inclusive
// prefix and return the exclusive prefix (uniform across the wave).
__device__ uint64_t lookback(uint64_t* flags, int64_t tile, uint64_t agg, uint32_t epoch) {
  int lane = wave_lane();
  if (tile == 0) {
    if (lane == 0) st_release(&flags[0], flag_word(epoch, 2, agg));
    return 0;
  }
  if (lane == 0) st_release(&flags[tile], flag_word(epoch, 1, agg));
  uint64_t prefix = 0;
  int64_t j = tile - 1;
  while (true) {
    int64_t idx = j - lane;
    uint64_t w;
    uint32_t st;
    while (true) {
      w = (idx < 0) ? flag_word(epoch, 2, 0) : ld_acquire(&flags[idx]);
      st = ((uint32_t)(w >> 40) == epoch) ? (uint32_t)((w >> FLAG_VAL_BITS) & 3u) : 0u;
      if (__all(st != 0)) break;
      __builtin_amdgcn_s_sleep(1);
    }
    uint64_t pm = __ballot(st == 2);
    uint64_t v = w & FLAG_VAL_MASK;
    int stop = pm ? (int)__builtin_ctzll(pm) : 63;
    uint64_t contrib = (lane <= stop) ? v : 0;
    for (int off = 32; off > 0; off >>= 1) contrib += __shfl_xor(contrib, off);
    prefix += contrib;
    if (pm) break;
    j -= 64;
  }
  if (lane == 0) st_release(&flags[tile], flag_word(epoch, 2, agg + prefix));
  return prefix;
}

// Push of a lepton from the first position half-step to the second
// (simulation.py:128-137 with particles.py:244-296); ig is 1/gamma on entry.
__device__ inline void lepton_push(const Plan& P, const SpeciesDev& sp, const StepArgs& a,
                                   double& x, double& y, double& z, double& ux, double& uy,
                                   double& uz, double& sx, double& sy, double& sz, double ig) {
  sfp_push_position(&x, &y, &z, ux, uy, uz, ig, a.cdt_half);
  double F[6];
  sfp_field_eval(&P.fields, x, y, z, a.t_half, F);
  if (sp.spin) {
    sfp_boris_tbmt(&ux, &uy, &uz, &ig, &sx, &sy, &sz, F[0], F[1], F[2], F[3], F[4], F[5], sp.q,
                   sp.ae, a.dt);
  } else {
    sfp_boris(&ux, &uy, &uz, &ig, F[0], F[1], F[2], F[3], F[4], F[5], sp.q, a.dt);
  }
  if (sp.rr == SFP_RR_LL || (sp.rr == SFP_RR_CLL && !sp.spin)) {
    double chi = sfp_chi(F[0], F[1], F[2], F[3], F[4], F[5], ux, uy, uz, ig);
    double factor = (sp.rr == SFP_RR_CLL) ? a.ll_factor * sfp_cll_g(chi) : a.ll_factor;
    sfp_ll_apply(&ux, &uy, &uz, &ig, chi, factor);
  }
  sfp_push_position(&x, &y, &z, ux, uy, uz, ig, a.cdt_half);
}

// Append one new lepton (pair member) at slot d of species `sp`, pushed.
__device__ inline void spawn_lepton(const Plan& P, const SpeciesDev& sp, const StepArgs& a,
                                    int64_t d, uint64_t id, double x, double y, double z,
                                    double ux, double uy, double uz) {
  double sx = 0.0, sy = 0.0, sz = 0.0;
  if (sp.push) {
    double ig = 1.0 / sfp_sqrt(1.0 + ux * ux + uy * uy + uz * uz);
    lepton_push(P, sp, a, x, y, z, ux, uy, uz, sx, sy, sz, ig);
  }
  sp.x[d] = x; sp.y[d] = y; sp.z[d] = z;
  sp.ux[d] = ux; sp.uy[d] = uy; sp.uz[d] = uz;
  sp.id[d] = id;
  sp.pruned[d] = 0;
  if (sp.tau) sp.tau[d] = 0.0;
  if (sp.spin) { sp.sx[d] = 0.0; sp.sy[d] = 0.0; sp.sz[d] = 0.0; }
}

// Append one new photon at slot d of species `sp`, pushed.
__device__ inline void spawn_photon(const SpeciesDev& sp, const StepArgs& a, int64_t d, uint64_t id,
                                    double x, double y, double z, double ux, double uy,
                                    double uz) {
  if (sp.push) {
    double ig = sfp_inv_gamma_photon(ux, uy, uz);
    sfp_push_position(&x, &y, &z, ux, uy, uz, ig, a.cdt_half);
    sfp_push_position(&x, &y, &z, ux, uy, uz, ig, a.cdt_half);
  }
  sp.x[d] = x; sp.y[d] = y; sp.z[d] = z;
  sp.ux[d] = ux; sp.uy[d] = uy; sp.uz[d] = uz;
  sp.id[d] = id;
  sp.pruned[d] = 0;
  if (sp.tau) sp.tau[d] = 0.0;
}

template <int KIND, bool PHOTON>
__global__ void __launch_bounds__(TILE) k_advance(const Plan* __restrict__ plan, int s, StepArgs a) {
  const Plan& P = *plan;
  const SpeciesDev& sp = P.sp[s];
  constexpr bool QED = KIND != KIND_PUSH;

  __shared__ sfp_rs_node lds_rs[KIND == KIND_RS ? 1023 : 1];
  __shared__ double2 lds_od[KIND == KIND_OD ? 255 : 1];
  __shared__ int64_t sh_tile;
  __shared__ uint32_t sh_wcnt[WAVES];
  __shared__ uint32_t sh_alive[WAVES];
  __shared__ uint64_t sh_prefix;

  const int64_t n_old = sp.cnt->n_old;
  const int64_t ntiles = (n_old + TILE - 1) / TILE;
  ScanState* scan = P.scan[s];
  uint32_t epoch = 0;
  if (QED) {
    epoch = (scan->epoch + 1u) & 0xffffffu;
    if (epoch == 0) epoch = 1;
  }

  // destinations of appended particles
  const SpeciesDev* dA = nullptr;  // photon dst, or electron dst
  const SpeciesDev* dB = nullptr;  // positron dst
  int64_t baseA = 0, baseB = 0, idA = 0, idB = 0;
  if (QED) {
    if (!PHOTON) {
      if (sp.photon_dst >= 0) dA = &P.sp[sp.photon_dst];
    } else {
      if (sp.pair_e_dst >= 0) dA = &P.sp[sp.pair_e_dst];
      if (sp.pair_p_dst >= 0) dB = &P.sp[sp.pair_p_dst];
    }
    if (dA) { baseA = dA->cnt->n; idA = dA->cnt->next_id; }
    if (dB) { baseB = dB->cnt->n; idB = dB->cnt->next_id; }
  }

  if (KIND == KIND_RS) {
    int64_t nn = P.z_N - 1;
    for (int64_t k = threadIdx.x; k < nn && k < 1023; k += TILE) lds_rs[k] = P.rs_nodes[k];
  }
  if (KIND == KIND_OD) {
    const double2* src = PHOTON ? P.od_pair_nodes : P.od_photon_nodes;
    int64_t nn = P.od.chi_N - 1;
    for (int64_t k = threadIdx.x; k < nn && k < 255; k += TILE) lds_od[k] = src[k];
  }
  __syncthreads();

  const int lane = wave_lane();
  const int wv = threadIdx.x / 64;
  uint32_t alive_acc = 0;
  uint32_t proc_acc = 0;

  while (true) {
    if (threadIdx.x == 0) sh_tile = (int64_t)atomicAdd(&scan->tile_ctr, 1u);
    __syncthreads();
    const int64_t tile = sh_tile;
    __syncthreads();
    if (tile >= ntiles) break;

    const int64_t i = tile * TILE + threadIdx.x;
    bool active = false;
    double x = 0, y = 0, z = 0, ux = 0, uy = 0, uz = 0, tau = 0, sx = 0, sy = 0, sz = 0;
    uint64_t pid = 0;
    if (i < n_old) {
      active = sp.pruned[i] == 0;
      if (active) {
        x = sp.x[i]; y = sp.y[i]; z = sp.z[i];
        ux = sp.ux[i]; uy = sp.uy[i]; uz = sp.uz[i];
        if (KIND == KIND_OD) tau = sp.tau[i];
        if (!PHOTON && sp.spin) { sx = sp.sx[i]; sy = sp.sy[i]; sz = sp.sz[i]; }
        if (QED) pid = sp.id[i];
      }
    }
    double ig = 0.0;
    if (active) {
      ig = PHOTON ? sfp_inv_gamma_photon(ux, uy, uz) : sfp_inv_gamma_lepton(ux, uy, uz);
      proc_acc += 1;
    }

    // ---- QED front at time t (simulation.py:103-114) ------------------
    bool event = false;
    double delta = 0.0;
    if (QED && active && sp.push) {
      double F[6];
      sfp_field_eval(&P.fields, x, y, z, a.t, F);
      double chi = sfp_chi(F[0], F[1], F[2], F[3], F[4], F[5], ux, uy, uz, ig);
      if (KIND == KIND_RS) {
        if (chi != 0.0) {
          sfp_u32x4 d = sfp_draw4(P.seed, pid, a.step, (uint32_t)sp.tag, 0u);
          double r1 = sfp_u01(d.v[0], d.v[1]);
          double r2 = sfp_u01(d.v[2], d.v[3]);
          double dtau = a.dt * ig;
          if (!PHOTON) {
            double r1_3 = sfp_powi(r1, 3);
            double p = 3 * sfp_powi(r1, 2) *
                       sfp_rs_rate(lds_rs, P.z_N, P.z_min, P.z_max, P.z_delta, chi, r1_3, 1) * dtau;
            if (r2 < p) { event = true; delta = r1_3; }
          } else {
            double p = sfp_rs_rate(lds_rs, P.z_N, P.z_min, P.z_max, P.z_delta, chi, r1, 0) * dtau;
            if (r2 < p) { event = true; delta = r1; }
          }
        }
      } else if (KIND == KIND_OD) {
        if (chi >= P.od.chi_min) {  // below the table or NaN: no event
          event = od_update(&tau, ig, chi, a.dt, P.seed, pid, a.step, (uint32_t)sp.tag, 0u, lds_od,
                            PHOTON ? P.od.pair_cdf : P.od.photon_cdf, P.od, &delta);
        }
      }
      // cpu.py:227-232 (radiating leptons only)
      if (!PHOTON && event) {
        if (delta / ig < P.threshold) event = false;
      }
    }
    if (!dA && !dB) event = false;  // nothing to append to: the event has no effect

    // ---- stable rank of the events (wave ballot + tile scan + look-back) --
    if (QED) {
      uint64_t m = __ballot(event);
      uint32_t lrank = (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
      if (lane == 0) sh_wcnt[wv] = (uint32_t)__popcll(m);
      __syncthreads();
      uint32_t woff = 0, agg = 0;
      for (int k = 0; k < WAVES; ++k) {
        uint32_t c = sh_wcnt[k];
        if (k < wv) woff += c;
        agg += c;
      }
      if (wv == 0) {
        uint64_t pre = lookback(P.flags[s], tile, agg, epoch);
        if (lane == 0) sh_prefix = pre;
      }
      __syncthreads();
      if (event) {
        int64_t rank = (int64_t)sh_prefix + woff + lrank;
        if (!PHOTON) {
          // cpu.py:90-118: photon at x(t) with momentum delta*u(t)
          int64_t d = baseA + rank;
          if (d < dA->cap)
            spawn_photon(*dA, a, d, (uint64_t)(idA + rank), x, y, z, delta * ux, delta * uy,
                         delta * uz);
          else
            atomicAdd((unsigned long long*)&dA->cnt->overflow, 1ull);
        } else {
          // cpu.py:120-191: electron takes delta, positron 1 - delta
          double dp = 1.0 - delta;
          if (dA) {
            int64_t d = baseA + rank;
            if (d < dA->cap)
              spawn_lepton(P, *dA, a, d, (uint64_t)(idA + rank), x, y, z, delta * ux, delta * uy,
                           delta * uz);
            else
              atomicAdd((unsigned long long*)&dA->cnt->overflow, 1ull);
          }
          if (dB) {
            int64_t d = baseB + rank;
            if (d < dB->cap)
              spawn_lepton(P, *dB, a, d, (uint64_t)(idB + rank), x, y, z, dp * ux, dp * uy,
                           dp * uz);
            else
              atomicAdd((unsigned long long*)&dB->cnt->overflow, 1ull);
          }
        }
      }
    }

    // ---- recoil (cpu.py:78-88) and push (simulation.py:128-137) -------
    if (active) {
      if (PHOTON) {
        if (event) {
          sp.pruned[i] = 1;  // the converted photon becomes a hole (cpu.py:189-190)
        } else {
          if (sp.push) {
            sfp_push_position(&x, &y, &z, ux, uy, uz, ig, a.cdt_half);
            sfp_push_position(&x, &y, &z, ux, uy, uz, ig, a.cdt_half);
            sp.x[i] = x; sp.y[i] = y; sp.z[i] = z;
          }
          if (KIND == KIND_OD) sp.tau[i] = tau;
          alive_acc += 1;
        }
      } else {
        if (event && sp.rr == SFP_RR_PHOTON) {
          ux *= 1 - delta;
          uy *= 1 - delta;
          uz *= 1 - delta;
          ig = sfp_inv_gamma_lepton(ux, uy, uz);
        }
        if (sp.push) {
          lepton_push(P, sp, a, x, y, z, ux, uy, uz, sx, sy, sz, ig);
          sp.x[i] = x; sp.y[i] = y; sp.z[i] = z;
          sp.ux[i] = ux; sp.uy[i] = uy; sp.uz[i] = uz;
          if (sp.spin) { sp.sx[i] = sx; sp.sy[i] = sy; sp.sz[i] = sz; }
        } else if (event && sp.rr == SFP_RR_PHOTON) {
          sp.ux[i] = ux; sp.uy[i] = uy; sp.uz[i] = uz;
        }
        if (KIND == KIND_OD) sp.tau[i] = tau;
        alive_acc += 1;
      }
    }
  }

  // ---- workgroup exit: live count; the last workgroup publishes totals ---
  {
    uint32_t c = alive_acc, pc = proc_acc;
    for (int off = 32; off > 0; off >>= 1) {
      c += __shfl_xor(c, off);
      pc += __shfl_xor(pc, off);
    }
    if (lane == 0) {
      sh_alive[wv] = c;
      sh_wcnt[wv] = pc;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      uint64_t tot = 0, ptot = 0;
      for (int k = 0; k < WAVES; ++k) {
        tot += sh_alive[k];
        ptot += sh_wcnt[k];
      }
      if (tot) atomicAdd((unsigned long long*)&sp.cnt->psteps, (unsigned long long)tot);
      if (ptot) atomicAdd((unsigned long long*)&sp.cnt->processed, (unsigned long long)ptot);
      if (QED) {
        __threadfence();
        uint32_t prev = atomicAdd(&scan->done_ctr, 1u);
        if (prev == gridDim.x - 1) {
          __threadfence();
          uint64_t total = 0;
          if (ntiles > 0) total = ld_acquire(&P.flags[s][ntiles - 1]) & FLAG_VAL_MASK;
          if (total) atomicAdd((unsigned long long*)&sp.cnt->spawned, (unsigned long long)total);
          if (dA) {
            dA->cnt->n = baseA + (int64_t)total < dA->cap ? baseA + (int64_t)total : dA->cap;
            dA->cnt->next_id = idA + (int64_t)total;
            atomicAdd((unsigned long long*)&dA->cnt->psteps, (unsigned long long)total);
          }
          if (dB) {
            dB->cnt->n = baseB + (int64_t)total < dB->cap ? baseB + (int64_t)total : dB->cap;
            dB->cnt->next_id = idB + (int64_t)total;
            atomicAdd((unsigned long long*)&dB->cnt->psteps, (unsigned long long)total);
          }
          scan->tile_ctr = 0;
          scan->done_ctr = 0;
          scan->epoch = epoch;
          __threadfence();
        }
      } else {
        __threadfence();
        uint32_t prev = atomicAdd(&scan->done_ctr, 1u);
        if (prev == gridDim.x - 1) {
          scan->tile_ctr = 0;
          scan->done_ctr = 0;
          __threadfence();
        }
      }
    }
  }
}

// prologue of a step: snapshot the counts the species kernels iterate over
__global__ void k_step_prologue(const Plan* plan) {
  int s = threadIdx.x;
  if (s < plan->nspecies) {
    Counts* c = plan->sp[s].cnt;
    c->n_old = c->n;
  }
}

static const void* pick_kernel(int kind, bool photon) {
  if (photon) {
    if (kind == KIND_RS) return (const void*)k_advance<KIND_RS, true>;
    if (kind == KIND_OD) return (const void*)k_advance<KIND_OD, true>;
    return (const void*)k_advance<KIND_PUSH, true>;
  }
  if (kind == KIND_RS) return (const void*)k_advance<KIND_RS, false>;
  if (kind == KIND_OD) return (const void*)k_advance<KIND_OD, false>;
  return (const void*)k_advance<KIND_PUSH, false>;
}

static void launch_advance(int kind, bool photon, int grid, hipStream_t st, const Plan* plan,
                           int s, StepArgs a) {
  dim3 g(grid), b(TILE);
  if (photon) {
    if (kind == KIND_RS) hipLaunchKernelGGL((k_advance<KIND_RS, true>), g, b, 0, st, plan, s, a);
    else if (kind == KIND_OD) hipLaunchKernelGGL((k_advance<KIND_OD, true>), g, b, 0, st, plan, s, a);
    else hipLaunchKernelGGL((k_advance<KIND_PUSH, true>), g, b, 0, st, plan, s, a);
  } else {
    if (kind == KIND_RS) hipLaunchKernelGGL((k_advance<KIND_RS, false>), g, b, 0, st, plan, s, a);
    else if (kind == KIND_OD) hipLaunchKernelGGL((k_advance<KIND_OD, false>), g, b, 0, st, plan, s, a);
    else hipLaunchKernelGGL((k_advance<KIND_PUSH, false>), g, b, 0, st, plan, s, a);
  }
}

struct Sim {
  int nsp = 0;
  std::vector<sfp_species_desc> desc;
  std::vector<Store*> stores;
  std::vector<int> kind;           // KIND_*
  std::vector<int> grid;           // persistent workgroups per species kernel
  std::vector<int64_t> flags_cap;  // tiles the flag array can hold
  std::vector<int64_t> n_ub;       // host-side upper bounds of the counts
  ScanState* scan_mem = nullptr;
  std::vector<uint64_t*> flags;
  Plan plan{};
  Plan* plan_dev = nullptr;
  Tables* tables = nullptr;
  sfp_sim_params params{};
  bool plan_dirty = true;
  int cus = 0;
  // timing
  std::vector<std::vector<hipEvent_t>> ev_pending;  // per species: start/stop pairs
  std::vector<double> ms_acc;
  std::vector<int64_t> launches;
  std::vector<hipEvent_t> ev_pool;
};

static int sim_sync_counts(Sim* sim) {
  HIPCHK(hipDeviceSynchronize());
  for (int s = 0; s < sim->nsp; ++s) {
    Counts c{};
    HIPCHK(hipMemcpy(&c, sim->stores[s]->cnt, sizeof(c), hipMemcpyDeviceToHost));
    if (c.overflow)
      return set_err(SFP_E_CAPACITY,
                     "species %d: %lld appended particles did not fit its capacity %lld", s,
                     (long long)c.overflow, (long long)sim->stores[s]->cap);
    sim->n_ub[s] = c.n;
  }
  return SFP_OK;
}

static int sim_refresh_plan(Sim* sim) {
  Plan& P = sim->plan;
  for (int s = 0; s < sim->nsp; ++s) {
    Store* st = sim->stores[s];
    SpeciesDev& d = P.sp[s];
    d.x = (double*)st->p[SFP_X];
    d.y = (double*)st->p[SFP_Y];
    d.z = (double*)st->p[SFP_Z];
    d.ux = (double*)st->p[SFP_UX];
    d.uy = (double*)st->p[SFP_UY];
    d.uz = (double*)st->p[SFP_UZ];
    d.id = (uint64_t*)st->p[SFP_ID];
    d.pruned = (uint8_t*)st->p[SFP_PRUNED];
    d.tau = (double*)st->p[SFP_TAU];
    d.sx = (double*)st->p[SFP_SX];
    d.sy = (double*)st->p[SFP_SY];
    d.sz = (double*)st->p[SFP_SZ];
    d.cnt = st->cnt;
    d.cap = st->cap;
    // look-back flags: one word per tile of the capacity
    int64_t need = st->cap / TILE + 2;
    if (need > sim->flags_cap[s]) {
      if (sim->flags[s]) HIPCHK(hipFree(sim->flags[s]));
      int64_t want = need + need / 2;
      HIPCHK(hipMalloc(&sim->flags[s], sizeof(uint64_t) * want));
      HIPCHK(hipMemset(sim->flags[s], 0, sizeof(uint64_t) * want));
      sim->flags_cap[s] = want;
    }
    P.flags[s] = sim->flags[s];
    P.scan[s] = sim->scan_mem + s;
  }
  HIPCHK(hipMemcpy(sim->plan_dev, &P, sizeof(Plan), hipMemcpyHostToDevice));
  sim->plan_dirty = false;
  return SFP_OK;
}

extern "C" int sfp_sim_create(const sfp_species_desc* species, int32_t nspecies,
                              const sfp_field* terms, int32_t nterms, void* tables,
                              const sfp_sim_params* params, void** out) {
  if (nspecies < 1 || nspecies > SFP_MAX_SPECIES)
    return set_err(SFP_E_ARG, "between 1 and %d species are supported", SFP_MAX_SPECIES);
  Sim* sim = new Sim();
  sim->nsp = nspecies;
  sim->params = *params;
  sim->tables = (Tables*)tables;
  sim->desc.assign(species, species + nspecies);
  Plan& P = sim->plan;
  memset(&P, 0, sizeof(P));
  CHK(compile_fields(terms, nterms, &P.fields));
  P.nspecies = nspecies;
  P.seed = params->seed;
  P.threshold = params->photon_threshold;
  P.qed = params->qed_model;
  bool need_rs = false, need_od = false;
  for (int s = 0; s < nspecies; ++s) {
    const sfp_species_desc& d = species[s];
    if (!d.store) return set_err(SFP_E_ARG, "species %d has no store", s);
    Store* st = (Store*)d.store;
    sim->stores.push_back(st);
    bool photon = d.m == 0.0;
    bool qed = photon ? (d.pair_e_dst >= 0 || d.pair_p_dst >= 0) : (d.photon_dst >= 0);
    int kind = !qed ? KIND_PUSH : (params->qed_model == SFP_QED_OD ? KIND_OD : KIND_RS);
    if (kind == KIND_RS) need_rs = true;
    if (kind == KIND_OD) {
      need_od = true;
      if (!(st->flags & SFP_STORE_TAU) || !st->p[SFP_TAU])
        return set_err(SFP_E_ARG, "species %d uses the optical-depth model but has no tau", s);
    }
    if (!photon && d.spin && !st->p[SFP_SX])
      return set_err(SFP_E_ARG, "species %d is a spin species without spin storage", s);
    for (int dst : {d.photon_dst, d.pair_e_dst, d.pair_p_dst})
      if (dst >= nspecies) return set_err(SFP_E_ARG, "species %d: destination out of range", s);
    if (!photon && (d.pair_e_dst >= 0 || d.pair_p_dst >= 0))
      return set_err(SFP_E_ARG, "species %d: only photons create pairs", s);
    if (photon && d.photon_dst >= 0)
      return set_err(SFP_E_ARG, "species %d: photons do not radiate photons", s);
    sim->kind.push_back(kind);
    SpeciesDev& sd = P.sp[s];
    sd.q = d.q;
    sd.ae = d.ae;
    sd.rr = d.rr;
    sd.push = d.push;
    sd.spin = photon ? 0 : d.spin;
    sd.photon = photon;
    sd.tag = s;
    sd.photon_dst = d.photon_dst;
    sd.pair_e_dst = d.pair_e_dst;
    sd.pair_p_dst = d.pair_p_dst;
  }
  if (need_rs) {
    if (!sim->tables || !sim->tables->has_rs)
      return set_err(SFP_E_ARG, "rejection-sampling tables required");
    if (sim->tables->z_N - 1 > 1023)
      return set_err(SFP_E_ARG, "rejection-sampling table larger than 1024 nodes");
    P.rs_nodes = sim->tables->rs_nodes;
    P.z_N = sim->tables->z_N;
    P.z_min = sim->tables->z_min;
    P.z_max = sim->tables->z_max;
    P.z_delta = sim->tables->z_delta;
  }
  if (need_od) {
    if (!sim->tables || !sim->tables->has_od)
      return set_err(SFP_E_ARG, "optical-depth tables required");
    if (sim->tables->od.chi_N - 1 > 255)
      return set_err(SFP_E_ARG, "optical-depth table larger than 256 chi nodes");
    P.od = sim->tables->od;
    P.od_photon_nodes = sim->tables->od_photon_nodes;
    P.od_pair_nodes = sim->tables->od_pair_nodes;
  }
  HIPCHK(hipMalloc(&sim->plan_dev, sizeof(Plan)));
  HIPCHK(hipMalloc(&sim->scan_mem, sizeof(ScanState) * nspecies));
  HIPCHK(hipMemset(sim->scan_mem, 0, sizeof(ScanState) * nspecies));
  sim->flags.assign(nspecies, nullptr);
  sim->flags_cap.assign(nspecies, 0);
  sim->n_ub.assign(nspecies, 0);
  sim->ms_acc.assign(nspecies, 0.0);
  sim->launches.assign(nspecies, 0);
  sim->ev_pending.resize(nspecies);
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  HIPCHK(hipGetDeviceProperties(&prop, dev));
  sim->cus = prop.multiProcessorCount;
  for (int s = 0; s < nspecies; ++s) {
    int per_cu = 0;
    HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu, pick_kernel(sim->kind[s], P.sp[s].photon), TILE, 0));
    if (per_cu < 1) per_cu = 1;
    sim->grid.push_back(per_cu * sim->cus);
  }
  CHK(sim_refresh_plan(sim));
  CHK(sim_sync_counts(sim));
  *out = sim;
  return SFP_OK;
}

extern "C" int sfp_sim_destroy(void* p) {
  Sim* sim = (Sim*)p;
  if (!sim) return SFP_OK;
  (void)hipDeviceSynchronize();
  for (auto f : sim->flags)
    if (f) (void)hipFree(f);
  if (sim->scan_mem) (void)hipFree(sim->scan_mem);
  if (sim->plan_dev) (void)hipFree(sim->plan_dev);
  for (auto& v : sim->ev_pending)
    for (auto e : v) (void)hipEventDestroy(e);
  for (auto e : sim->ev_pool) (void)hipEventDestroy(e);
  delete sim;
  return SFP_OK;
}

// make sure every append of the next step fits: worst case every source
// particle creates one particle in each of its destinations
static int sim_ensure_capacity(Sim* sim) {
  std::vector<int64_t> need(sim->n_ub);
  for (int s = 0; s < sim->nsp; ++s) {
    const sfp_species_desc& d = sim->desc[s];
    if (sim->kind[s] == KIND_PUSH) continue;
    for (int dst : {d.photon_dst, d.pair_e_dst, d.pair_p_dst})
      if (dst >= 0) need[dst] += sim->n_ub[s];
  }
  bool fits = true;
  for (int s = 0; s < sim->nsp; ++s)
    if (need[s] > sim->stores[s]->cap) fits = false;
  if (fits) {
    sim->n_ub = need;
    return SFP_OK;
  }
  // exact counts, then grow what still does not fit
  CHK(sim_sync_counts(sim));
  need = sim->n_ub;
  for (int s = 0; s < sim->nsp; ++s) {
    const sfp_species_desc& d = sim->desc[s];
    if (sim->kind[s] == KIND_PUSH) continue;
    for (int dst : {d.photon_dst, d.pair_e_dst, d.pair_p_dst})
      if (dst >= 0) need[dst] += sim->n_ub[s];
  }
  for (int s = 0; s < sim->nsp; ++s) {
    Store* st = sim->stores[s];
    if (need[s] > st->cap) {
      int64_t want = need[s] + need[s] / 2 + 1024;
      CHK(sfp_store_reserve(st, want));
      sim->plan_dirty = true;
    }
  }
  sim->n_ub = need;
  if (sim->plan_dirty) CHK(sim_refresh_plan(sim));
  return SFP_OK;
}

static hipEvent_t sim_event(Sim* sim) {
  if (!sim->ev_pool.empty()) {
    hipEvent_t e = sim->ev_pool.back();
    sim->ev_pool.pop_back();
    return e;
  }
  hipEvent_t e;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

extern "C" int sfp_sim_run(void* p, double* t, uint64_t step0, int64_t nsteps, int32_t flags) {
  Sim* sim = (Sim*)p;
  if (!sim) return set_err(SFP_E_ARG, "null simulation");
  // stores may have been resized or refilled by the caller since the last run
  for (int s = 0; s < sim->nsp; ++s)
    if (sim->plan.sp[s].cap != sim->stores[s]->cap || sim->plan.sp[s].x != sim->stores[s]->p[SFP_X])
      sim->plan_dirty = true;
  if (sim->plan_dirty) CHK(sim_refresh_plan(sim));
  CHK(sim_sync_counts(sim));
  const double dt = sim->params.dt;
  const bool timing = flags & SFP_RUN_TIMING;
  for (int64_t k = 0; k < nsteps; ++k) {
    CHK(sim_ensure_capacity(sim));
    StepArgs a;
    a.t = *t;
    a.t_half = *t + 0.5 * dt;
    a.dt = dt;
    a.cdt_half = SFP_C * (0.5 * dt);
    a.ll_factor = sfp_ll_factor(dt);
    a.step = step0 + (uint64_t)k;
    hipLaunchKernelGGL(k_step_prologue, dim3(1), dim3(64), 0, 0, (const Plan*)sim->plan_dev);
    for (int s = 0; s < sim->nsp; ++s) {
      int64_t tiles_ub = (sim->n_ub[s] + TILE - 1) / TILE;
      int g = (int)std::max<int64_t>(1, std::min<int64_t>(sim->grid[s], tiles_ub));
      hipEvent_t e0 = nullptr, e1 = nullptr;
      if (timing) {
        e0 = sim_event(sim);
        e1 = sim_event(sim);
        if (e0) HIPCHK(hipEventRecord(e0, 0));
      }
      launch_advance(sim->kind[s], sim->plan.sp[s].photon, g, 0, (const Plan*)sim->plan_dev, s, a);
      HIPCHK(hipGetLastError());
      if (timing && e1) {
        HIPCHK(hipEventRecord(e1, 0));
        sim->ev_pending[s].push_back(e0);
        sim->ev_pending[s].push_back(e1);
      }
    }
    *t = *t + dt;
  }
  CHK(sim_sync_counts(sim));
  return SFP_OK;
}

static int sim_collect_timing(Sim* sim) {
  HIPCHK(hipDeviceSynchronize());
  for (int s = 0; s < sim->nsp; ++s) {
    auto& v = sim->ev_pending[s];
    for (size_t k = 0; k + 1 < v.size(); k += 2) {
      float ms = 0.f;
      HIPCHK(hipEventElapsedTime(&ms, v[k], v[k + 1]));
      sim->ms_acc[s] += ms;
      sim->launches[s] += 1;
      sim->ev_pool.push_back(v[k]);
      sim->ev_pool.push_back(v[k + 1]);
    }
    v.clear();
  }
  return SFP_OK;
}

extern "C" int sfp_sim_kernel_time(void* p, int32_t species, double* ms, int64_t* launches) {
  Sim* sim = (Sim*)p;
  if (!sim || species < 0 || species >= sim->nsp) return set_err(SFP_E_ARG, "bad species");
  CHK(sim_collect_timing(sim));
  *ms = sim->ms_acc[species];
  *launches = sim->launches[species];
  return SFP_OK;
}

extern "C" int sfp_sim_reset_timing(void* p) {
  Sim* sim = (Sim*)p;
  if (!sim) return set_err(SFP_E_ARG, "null simulation");
  CHK(sim_collect_timing(sim));
  for (int s = 0; s < sim->nsp; ++s) {
    sim->ms_acc[s] = 0.0;
    sim->launches[s] = 0;
  }
  return SFP_OK;
}

extern "C" int sfp_sim_stats(void* p, int32_t species, int64_t* stats) {
  Sim* sim = (Sim*)p;
  if (!sim || species < 0 || species >= sim->nsp) return set_err(SFP_E_ARG, "bad species");
  Counts c{};
  HIPCHK(hipDeviceSynchronize());
  HIPCHK(hipMemcpy(&c, sim->stores[species]->cnt, sizeof(c), hipMemcpyDeviceToHost));
  stats[0] = c.psteps;
  stats[1] = c.processed;
  stats[2] = c.spawned;
  stats[3] = c.n;
  return SFP_OK;
}

extern "C" int sfp_sim_reset_stats(void* p) {
  Sim* sim = (Sim*)p;
  if (!sim) return set_err(SFP_E_ARG, "null simulation");
  HIPCHK(hipDeviceSynchronize());
  for (int s = 0; s < sim->nsp; ++s) {
    Counts c{};
    HIPCHK(hipMemcpy(&c, sim->stores[s]->cnt, sizeof(c), hipMemcpyDeviceToHost));
    c.psteps = c.processed = c.spawned = 0;
    HIPCHK(hipMemcpy(sim->stores[s]->cnt, &c, sizeof(c), hipMemcpyHostToDevice));
  }
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
