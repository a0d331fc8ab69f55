/*
 * oracle.c -- CPU restatement of sfparticles' per-particle step, used ONLY as
 * the checker for the HIP implementation (tests, the smoke check and the CPU
 * baseline leg of the benchmark).  It is never linked into or called by the
 * product path.
 *
 * Structure follows the reference pass by pass: one OpenMP-parallel loop per
 * reference function (the reference's numba `prange` loops), serial
 * bool_sum / find_event_index (the reference's are serial), FP64 throughout.
 * Each function cites the reference file:line it restates.  Expression order
 * follows the reference's Python source operator by operator so that the
 * floating-point result is the one the reference's formula defines.
 *
 * Two deliberate differences from the reference, both forced by the task:
 *  - random draws come from Philox4x32-10 keyed by (seed, particle id, step)
 *    instead of numba's unseeded per-thread numpy.random state, so that the
 *    GPU's decisions can be replayed;
 *  - the reference's undefined behaviour is defined: photon rejection sampling
 *    returns rate 0 for z > z_max (the reference reads past the table end),
 *    optical-depth lookups clamp chi to the table's top end, the inverse-CDF
 *    bisection resolves ties to the left and clamps the bracket index, and
 *    the paraxial Gaussian's curvature term is 0 on the focal plane x == 0.
 *
 * Transcendentals: by default the project's own elementary functions
 * (sfp_math.h), so that results are bit-comparable with the device code;
 * compiled with -DORC_LIBM the platform libm is used instead, which is the
 * closer restatement of the reference's numba/libm arithmetic and is used to
 * show that the two agree to far better than the 1e-10 trajectory tolerance.
 *
 * parity pinned by: tests/test_oracle_*.py (closed forms, known-answer
 * vectors, the reference's own statistical tests re-expressed).
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "../sfparticles_amd/csrc/sfp_math.h"

#ifdef ORC_LIBM
#define O_EXP exp
#define O_LOG log
#define O_LOG10 log10
#define O_COS cos
#define O_SIN sin
#define O_ATAN atan
#define O_POW pow
#define O_EXP10(x) pow(10.0, (x))
#else
#define O_EXP sfp_exp
#define O_LOG sfp_log
#define O_LOG10 sfp_log10
#define O_COS sfp_cos
#define O_SIN sfp_sin
#define O_ATAN sfp_atan
#define O_POW sfp_pow
#define O_EXP10 sfp_exp10
#endif
#define O_SQRT sqrt

/* Physical constants: CODATA 2018 (scipy.constants 1.7.x .. 1.14.x), the set
 * the reference's `from scipy.constants import ...` resolved to before scipy
 * 1.15 (sfparticles/particles.py:5).  Pinned so results do not depend on the
 * installed scipy. */
#define O_C 299792458.0
#define O_ME 9.1093837015e-31
#define O_E 1.602176634e-19
#define O_HBAR 1.0545718176461565e-34
#define O_EPS0 8.8541878128e-12
#define O_ALPHA 0.0072973525693
#define O_PI 3.141592653589793

/* integer power the way numba lowers float**int (binary exponentiation) */
static inline double o_powi(double a, int n) {
  double r = 1.0;
  while (1) {
    if (n & 1) r = r * a;
    n >>= 1;
    if (n == 0) break;
    a = a * a;
  }
  return r;
}

/* ================================================================ fields */
/* One analytic field term.  Parameters are the user-level arguments of the
 * reference's builders, in their order:
 *  kind 0 static  (fields.py:268-271): p = Ex, Ey, Ez, Bx, By, Bz
 *  kind 1 simple  (fields.py:69-118):  p = a0, w0, ctau, t_order, direction,
 *                                          x0, wavelength, pol_angle, cep
 *  kind 2 gaussian(fields.py:120-170): p = a0, w0, ctau, direction, x0, l0,
 *                                          pol_angle, cep
 *  kind 3 focused (fields.py:173-265): p = a0, w0, ctau, direction, xfoc, x0,
 *                                          l0, pol_angle, cep */
typedef struct {
  int32_t kind;
  int32_t pad;
  double p[12];
} orc_field;

static void o_field_static(const double* p, double* F) {
  for (int i = 0; i < 6; ++i) F[i] = p[i];
}

/* fields.py:69-118 */
static void o_field_simple(const double* p, double x, double y, double z, double t, double* F) {
  double a0 = p[0], w0 = p[1], ctau = p[2];
  int t_order = (int)p[3];
  double direction = p[4], x0 = p[5], wavelength = p[6], pol_angle = p[7], cep = p[8];
  double omega0 = 2 * O_PI * O_C / wavelength;
  double k0 = omega0 / O_C;
  double a0_norm = O_E / (O_ME * O_C * omega0);
  double E0 = a0 / a0_norm;

  x = direction * x;
  double r2 = y * y + z * z;
  double phi = k0 * (x - direction * x0 - O_C * t);
  double E = E0 * O_COS(phi + cep) * O_EXP(-r2 / (w0 * w0)) *
             O_EXP(-o_powi(phi, t_order) / o_powi(k0 * ctau, t_order));
  double Ey = E * O_COS(pol_angle);
  double Ez = E * O_SIN(pol_angle);
  F[0] = 0.0;
  F[1] = Ey;
  F[2] = Ez;
  F[3] = 0.0;
  F[4] = -Ez / O_C * direction;
  F[5] = Ey / O_C * direction;
}

/* fields.py:120-170 */
static void o_field_gaussian(const double* p, double x, double y, double z, double t, double* F) {
  double a0 = p[0], w0 = p[1], ctau = p[2], direction = p[3], x0 = p[4], l0 = p[5];
  double pol_angle = p[6], cep = p[7];
  double omega0 = 2 * O_PI * O_C / l0;
  double k0 = 2 * O_PI / l0;
  double zR = O_PI * (w0 * w0) / l0;
  double a0_norm = O_E / (O_ME * O_C * omega0);

  double r2 = y * y + z * z;
  double xz = x / zR;
  double wx = w0 * O_SQRT(1 + xz * xz);
  double gouy = O_ATAN(x / zR);
  double E0 = a0 / a0_norm;
  double curv;
  if (x == 0.0) {
    curv = 0.0; /* 1/R(x) -> 0 on the focal plane (reference divides by 0 there) */
  } else {
    double zx = zR / x;
    double Rx = x * (1 + zx * zx);
    curv = k0 * r2 / 2 / Rx;
  }
  double phi = k0 * (x - x0 - direction * O_C * t) + curv - gouy;
  double kct = k0 * ctau;
  double E = E0 * O_COS(phi + cep) * w0 / wx * O_EXP(-r2 / (wx * wx)) *
             O_EXP(-(phi * phi) / (kct * kct));
  double Ey = E * O_COS(pol_angle);
  double Ez = E * O_SIN(pol_angle);
  F[0] = 0.0;
  F[1] = Ey;
  F[2] = Ez;
  F[3] = 0.0;
  F[4] = -Ez / O_C * direction;
  F[5] = Ey / O_C * direction;
}

/* fields.py:173-265.
 * p[9] (not a reference argument; 0 in every product and parity use) selects
 * the sign convention.  0: the reference's, +psiG in the phase and the
 * longitudinal components as written.  1: the signs under which the
 * expansion (Salamin & Keitel, PRL 88, 095005 (2002), written there for the
 * phase omega t - k z) is a solution of Maxwell's equations in the phase
 * k x - omega t used here: the Gouy phase enters with -psiG, Ex and Bx change
 * sign, Ex is mirrored with the direction as By and Bz are, and the
 * transverse coordinates are rotated with the polarisation (the reference
 * rotates the field vectors by pol_angle but evaluates them at the unrotated
 * y, z, fields.py:257-263).  The
 * coefficients are the same code in both; mode 1 exists so that tests can pin
 * them against Maxwell's equations (the residuals fall as eps^6 only if every
 * coefficient up to eps^5 is right), which the reference's own signs do not
 * satisfy beyond O(eps). */
static void o_field_focused(const double* p, double x, double y, double z, double t, double* F) {
  double a0 = p[0], w0 = p[1], ctau = p[2], direction = p[3], xfoc = p[4], x0 = p[5];
  double l0 = p[6], pol_angle = p[7], cep = p[8];
  double gs = p[9] != 0.0 ? -1.0 : 1.0;
  double ex_dir = p[9] != 0.0 ? direction : 1.0;
  double omega0 = 2 * O_PI * O_C / l0;
  double k0 = 2 * O_PI / l0;
  double zR = O_PI * (w0 * w0) / l0;
  double eps = w0 / zR;
  double a0_norm = O_E / (O_ME * O_C * omega0);
  double E0 = a0 / a0_norm;

  x = direction * (x - xfoc);
  if (p[9] != 0.0) {
    double c_ = O_COS(pol_angle), s_ = O_SIN(pol_angle);
    double yl = y * c_ + z * s_;
    z = -y * s_ + z * c_;
    y = yl;
  }
  double r2 = y * y + z * z;
  double xz = x / zR;
  double wx = w0 * O_SQRT(1 + xz * xz);
  double psiG = gs * O_ATAN(x / zR);
  double psiR = k0 * x * r2 / 2 / (x * x + zR * zR);
  double psiP = k0 * (x - direction * x0 - O_C * t);
  double psi = psiG + psiR + psiP + cep;

  double rho = O_SQRT(r2) / w0;
  double xi = y / w0;
  double nu = z / w0;
  double eta = 1 / O_SQRT(1 + xz * xz);

  double rho2 = o_powi(rho, 2), rho4 = o_powi(rho, 4), rho6 = o_powi(rho, 6), rho8 = o_powi(rho, 8);
  double kct = k0 * ctau;
  double pc = psi - cep;
  double E = E0 * w0 / wx * O_EXP(-r2 / (wx * wx)) * O_EXP(-(pc * pc) / (kct * kct));

  double S0 = O_SIN(psi);
  double S2 = o_powi(eta, 2) * O_SIN(psi + 2 * psiG);
  double S3 = o_powi(eta, 3) * O_SIN(psi + 3 * psiG);
  double S4 = o_powi(eta, 4) * O_SIN(psi + 4 * psiG);
  double S5 = o_powi(eta, 5) * O_SIN(psi + 5 * psiG);
  double S6 = o_powi(eta, 6) * O_SIN(psi + 6 * psiG);
  double C1 = o_powi(eta, 1) * O_COS(psi + 1 * psiG);
  double C2 = o_powi(eta, 2) * O_COS(psi + 2 * psiG);
  double C3 = o_powi(eta, 3) * O_COS(psi + 3 * psiG);
  double C4 = o_powi(eta, 4) * O_COS(psi + 4 * psiG);
  double C5 = o_powi(eta, 5) * O_COS(psi + 5 * psiG);
  double C6 = o_powi(eta, 6) * O_COS(psi + 6 * psiG);
  double C7 = o_powi(eta, 7) * O_COS(psi + 7 * psiG);

  double e2 = o_powi(eps, 2), e3 = o_powi(eps, 3), e4 = o_powi(eps, 4), e5 = o_powi(eps, 5);
  double xi2 = xi * xi;
  double Ex_ = E * xi *
               (eps * C1 + e3 * (-C2 / 2 + rho2 * C3 - rho4 * C4 / 4) +
                e5 * (-3.0 / 8 * C3 - 3.0 / 8 * rho2 * C4 + 17.0 / 16 * rho4 * C5 - 3.0 / 8 * rho6 * C6 +
                      rho8 * C7 / 32));
  double Ey_ = E * (S0 + e2 * (xi2 * S2 - rho4 * S3 / 4) +
                    e4 * (S2 / 8 - rho2 * S3 / 4 - rho2 * (rho2 - 16 * xi2) * S4 / 16 -
                          rho4 * (rho2 + 2 * xi2) * S5 / 8 + rho8 * S6 / 32));
  double Ez_ = E * xi * nu * (e2 * S2 + e4 * (rho2 * S4 - rho4 * S5 / 4));
  double Bx_ = E / O_C * nu *
               (eps * C1 + e3 * (C2 / 2 + rho2 * C3 / 2 - rho4 * C4 / 4) +
                e5 * (3.0 / 8 * C3 + 3.0 / 8 * rho2 * C4 + 3.0 / 16 * rho4 * C5 - rho6 * C6 / 4 +
                      rho8 * C7 / 32));
  double By_ = 0.0;
  double Bz_ = E / O_C *
               (S0 + e2 * (rho2 * S2 / 2 - rho4 * S3 / 4) +
                e4 * (-S2 / 8 + rho2 * S3 / 4 + 5.0 / 16 * rho4 * S4 - rho6 * S5 / 4 + rho8 * S6 / 32));
  double cp = O_COS(pol_angle), sp = O_SIN(pol_angle);
  F[0] = Ex_ * gs * ex_dir;
  F[1] = Ey_ * cp - Ez_ * sp;
  F[2] = Ey_ * sp + Ez_ * cp;
  F[3] = Bx_ * gs;
  F[4] = (By_ * cp - Bz_ * sp) * direction;
  F[5] = (By_ * sp + Bz_ * cp) * direction;
}

static void o_field_point(const orc_field* f, int nf, double x, double y, double z, double t,
                          double* F) {
  /* Fields.__add__ (fields.py:52-61): left fold of the 6-tuples */
  for (int j = 0; j < nf; ++j) {
    double G[6];
    switch (f[j].kind) {
      case 0: o_field_static(f[j].p, G); break;
      case 1: o_field_simple(f[j].p, x, y, z, t, G); break;
      case 2: o_field_gaussian(f[j].p, x, y, z, t, G); break;
      default: o_field_focused(f[j].p, x, y, z, t, G); break;
    }
    if (j == 0) {
      for (int i = 0; i < 6; ++i) F[i] = G[i];
    } else {
      for (int i = 0; i < 6; ++i) F[i] = F[i] + G[i];
    }
  }
  if (nf == 0)
    for (int i = 0; i < 6; ++i) F[i] = 0.0;
}

/* Fields.field_func (fields.py:37-46) via Particles._eval_field (particles.py:299-323) */
void orc_eval_field(const orc_field* f, int32_t nf, const double* x, const double* y,
                    const double* z, double t, int64_t N, const uint8_t* pruned, double* Ex,
                    double* Ey, double* Ez, double* Bx, double* By, double* Bz) {
#pragma omp parallel for schedule(static)
  for (int64_t ip = 0; ip < N; ++ip) {
    if (pruned[ip]) continue;
    double F[6];
    o_field_point(f, nf, x[ip], y[ip], z[ip], t, F);
    Ex[ip] = F[0]; Ey[ip] = F[1]; Ez[ip] = F[2];
    Bx[ip] = F[3]; By[ip] = F[4]; Bz[ip] = F[5];
  }
}

/* single-point evaluation, for closed-form tests */
void orc_field_at(const orc_field* f, int32_t nf, double x, double y, double z, double t, double* F) {
  o_field_point(f, nf, x, y, z, t, F);
}

/* ================================================================ pusher */
/* cpu.py:5-19 */
void orc_push_position(double* x, double* y, double* z, const double* ux, const double* uy,
                       const double* uz, const double* inv_gamma, int64_t N,
                       const uint8_t* pruned, double dt) {
  double cdt = O_C * dt;
#pragma omp parallel for schedule(static)
  for (int64_t ip = 0; ip < N; ++ip) {
    if (pruned[ip]) continue;
    x[ip] += cdt * inv_gamma[ip] * ux[ip];
    y[ip] += cdt * inv_gamma[ip] * uy[ip];
    z[ip] += cdt * inv_gamma[ip] * uz[ip];
  }
}

/* inline.py:4-36 */
static inline void o_boris(double* pux, double* puy, double* puz, double* pig, double Ex, double Ey,
                           double Ez, double Bx, double By, double Bz, double q, double dt) {
  double efactor = q * dt / (2 * O_ME * O_C);
  double bfactor = q * dt / (2 * O_ME);
  double ux_minus = *pux + efactor * Ex;
  double uy_minus = *puy + efactor * Ey;
  double uz_minus = *puz + efactor * Ez;
  double inv_gamma_minus =
      1 / O_SQRT(1 + ux_minus * ux_minus + uy_minus * uy_minus + uz_minus * uz_minus);
  double Tx = bfactor * Bx * inv_gamma_minus;
  double Ty = bfactor * By * inv_gamma_minus;
  double Tz = bfactor * Bz * inv_gamma_minus;
  double ux_prime = ux_minus + uy_minus * Tz - uz_minus * Ty;
  double uy_prime = uy_minus + uz_minus * Tx - ux_minus * Tz;
  double uz_prime = uz_minus + ux_minus * Ty - uy_minus * Tx;
  double Tfactor = 2 / (1 + Tx * Tx + Ty * Ty + Tz * Tz);
  double Sx = Tfactor * Tx, Sy = Tfactor * Ty, Sz = Tfactor * Tz;
  double ux_plus = ux_minus + uy_prime * Sz - uz_prime * Sy;
  double uy_plus = uy_minus + uz_prime * Sx - ux_prime * Sz;
  double uz_plus = uz_minus + ux_prime * Sy - uy_prime * Sx;
  double ux_ = ux_plus + efactor * Ex;
  double uy_ = uy_plus + efactor * Ey;
  double uz_ = uz_plus + efactor * Ez;
  *pux = ux_; *puy = uy_; *puz = uz_;
  *pig = 1 / O_SQRT(1 + ux_ * ux_ + uy_ * uy_ + uz_ * uz_);
}

/* cpu.py:22-31 */
void orc_boris(double* ux, double* uy, double* uz, double* inv_gamma, const double* Ex,
               const double* Ey, const double* Ez, const double* Bx, const double* By,
               const double* Bz, double q, int64_t N, const uint8_t* pruned, double dt) {
#pragma omp parallel for schedule(static)
  for (int64_t ip = 0; ip < N; ++ip) {
    if (pruned[ip]) continue;
    o_boris(&ux[ip], &uy[ip], &uz[ip], &inv_gamma[ip], Ex[ip], Ey[ip], Ez[ip], Bx[ip], By[ip],
            Bz[ip], q, dt);
  }
}

/* inline.py:39-100 */
static inline void o_boris_tbmt(double* pux, double* puy, double* puz, double* pig, double* psx,
                                double* psy, double* psz, double Ex, double Ey, double Ez,
                                double Bx, double By, double Bz, double q, double ae, double dt) {
  double efactor = q * dt / (2 * O_ME * O_C);
  double bfactor = q * dt / (2 * O_ME);
  double ux_minus = *pux + efactor * Ex;
  double uy_minus = *puy + efactor * Ey;
  double uz_minus = *puz + efactor * Ez;
  double u_dot_B = ux_minus * Bx + uy_minus * By + uz_minus * Bz;
  double gamma_minus = O_SQRT(1.0 + ux_minus * ux_minus + uy_minus * uy_minus + uz_minus * uz_minus);
  double inv_gamma_minus = 1. / gamma_minus;
  double Bx_eff = (ae + inv_gamma_minus) * Bx - ae * gamma_minus / (gamma_minus + 1) * u_dot_B * ux_minus -
                  (ae + 1 / (gamma_minus + 1)) * (uy_minus * Ez - uz_minus * Ey);
  double By_eff = (ae + inv_gamma_minus) * By - ae * gamma_minus / (gamma_minus + 1) * u_dot_B * uy_minus -
                  (ae + 1 / (gamma_minus + 1)) * (uz_minus * Ex - ux_minus * Ez);
  double Bz_eff = (ae + inv_gamma_minus) * Bz - ae * gamma_minus / (gamma_minus + 1) * u_dot_B * uz_minus -
                  (ae + 1 / (gamma_minus + 1)) * (ux_minus * Ey - uy_minus * Ex);
  double Tx = bfactor * Bx_eff, Ty = bfactor * By_eff, Tz = bfactor * Bz_eff;
  double T2 = Tx * Tx + Ty * Ty + Tz * Tz;
  double Sx = 2 * Tx / (1. + T2), Sy = 2 * Ty / (1. + T2), Sz = 2 * Tz / (1. + T2);
  double sx = *psx, sy = *psy, sz = *psz;
  double sx_plus = sx + sy * Tz - sz * Ty;
  double sy_plus = sy + sz * Tx - sx * Tz;
  double sz_plus = sz + sx * Ty - sy * Tx;
  double sx_new = sx + sy_plus * Sz - sz_plus * Sy;
  double sy_new = sy + sz_plus * Sx - sx_plus * Sz;
  double sz_new = sz + sx_plus * Sy - sy_plus * Sx;

  Tx = bfactor * Bx * inv_gamma_minus;
  Ty = bfactor * By * inv_gamma_minus;
  Tz = bfactor * Bz * inv_gamma_minus;
  double ux_prime = ux_minus + uy_minus * Tz - uz_minus * Ty;
  double uy_prime = uy_minus + uz_minus * Tx - ux_minus * Tz;
  double uz_prime = uz_minus + ux_minus * Ty - uy_minus * Tx;
  double Tfactor = 2 / (1 + Tx * Tx + Ty * Ty + Tz * Tz);
  Sx = Tfactor * Tx; Sy = Tfactor * Ty; Sz = Tfactor * Tz;
  double ux_plus = ux_minus + uy_prime * Sz - uz_prime * Sy;
  double uy_plus = uy_minus + uz_prime * Sx - ux_prime * Sz;
  double uz_plus = uz_minus + ux_prime * Sy - uy_prime * Sx;
  double ux_new = ux_plus + efactor * Ex;
  double uy_new = uy_plus + efactor * Ey;
  double uz_new = uz_plus + efactor * Ez;
  *pux = ux_new; *puy = uy_new; *puz = uz_new;
  *pig = 1 / O_SQRT(1 + ux_new * ux_new + uy_new * uy_new + uz_new * uz_new);
  *psx = sx_new; *psy = sy_new; *psz = sz_new;
}

/* cpu.py:33-42 */
void orc_boris_tbmt(double* ux, double* uy, double* uz, double* inv_gamma, double* sx, double* sy,
                    double* sz, const double* Ex, const double* Ey, const double* Ez,
                    const double* Bx, const double* By, const double* Bz, double q, double ae,
                    int64_t N, const uint8_t* pruned, double dt) {
#pragma omp parallel for schedule(static)
  for (int64_t ip = 0; ip < N; ++ip) {
    if (pruned[ip]) continue;
    o_boris_tbmt(&ux[ip], &uy[ip], &uz[ip], &inv_gamma[ip], &sx[ip], &sy[ip], &sz[ip], Ex[ip],
                 Ey[ip], Ez[ip], Bx[ip], By[ip], Bz[ip], q, ae, dt);
  }
}

/* inline.py:104-111 (factor) */
static inline double o_ll_factor(double dt) {
  return -2.0 / 3 / (4 * O_PI * O_EPS0) * (O_E * O_E) * O_ME * O_C / (O_HBAR * O_HBAR) * dt;
}

/* inline.py:104-111 via cpu.py:44-51 */
void orc_LL_push(double* ux, double* uy, double* uz, double* inv_gamma, const double* chi_e,
                 int64_t N, const uint8_t* pruned, double dt) {
  double factor = o_ll_factor(dt);
#pragma omp parallel for schedule(static)
  for (int64_t ip = 0; ip < N; ++ip) {
    if (pruned[ip]) continue;
    double c2 = chi_e[ip] * chi_e[ip];
    double ig = inv_gamma[ip];
    double ux_ = ux[ip] + factor * c2 * ux[ip] * ig;
    double uy_ = uy[ip] + factor * c2 * uy[ip] * ig;
    double uz_ = uz[ip] + factor * c2 * uz[ip] * ig;
    ux[ip] = ux_; uy[ip] = uy_; uz[ip] = uz_;
    inv_gamma[ip] = 1 / O_SQRT(1 + ux_ * ux_ + uy_ * uy_ + uz_ * uz_);
  }
}

/* inline.py:114-123 via cpu.py:53-60 */
void orc_CLL_push(double* ux, double* uy, double* uz, double* inv_gamma, const double* chi_e,
                  int64_t N, const uint8_t* pruned, double dt) {
  double factor0 = o_ll_factor(dt);
#pragma omp parallel for schedule(static)
  for (int64_t ip = 0; ip < N; ++ip) {
    if (pruned[ip]) continue;
    double chi = chi_e[ip];
    double g = O_POW(1 + 4.8 * (1 + chi) * O_LOG(1 + 1.7 * chi) + 2.44 * (chi * chi), -2.0 / 3);
    double factor = factor0 * g;
    double c2 = chi * chi;
    double ig = inv_gamma[ip];
    double ux_ = ux[ip] + factor * c2 * ux[ip] * ig;
    double uy_ = uy[ip] + factor * c2 * uy[ip] * ig;
    double uz_ = uz[ip] + factor * c2 * uz[ip] * ig;
    ux[ip] = ux_; uy[ip] = uy_; uz[ip] = uz_;
    inv_gamma[ip] = 1 / O_SQRT(1 + ux_ * ux_ + uy_ * uy_ + uz_ * uz_);
  }
}

/* inline.py:126-134 */
static inline double o_chi(double Ex, double Ey, double Ez, double Bx, double By, double Bz,
                           double ux, double uy, double uz, double inv_gamma) {
  double factor = O_E * O_HBAR / ((O_ME * O_ME) * (O_C * O_C * O_C));
  double gamma = 1.0 / inv_gamma;
  double a = gamma * Ex + (uy * Bz - uz * By) * O_C;
  double b = gamma * Ey + (uz * Bx - ux * Bz) * O_C;
  double c = gamma * Ez + (ux * By - uy * Bx) * O_C;
  double d = ux * Ex + uy * Ey + uz * Ez;
  return factor * O_SQRT(a * a + b * b + c * c - d * d);
}

/* cpu.py:62-75 */
void orc_update_chi(const double* Ex, const double* Ey, const double* Ez, const double* Bx,
                    const double* By, const double* Bz, const double* ux, const double* uy,
                    const double* uz, const double* inv_gamma, double* chi, int64_t N,
                    const uint8_t* pruned) {
#pragma omp parallel for schedule(static)
  for (int64_t ip = 0; ip < N; ++ip) {
    if (pruned[ip]) continue;
    chi[ip] = o_chi(Ex[ip], Ey[ip], Ez[ip], Bx[ip], By[ip], Bz[ip], ux[ip], uy[ip], uz[ip],
                    inv_gamma[ip]);
  }
}

/* cpu.py:78-88 */
void orc_photon_recoil(double* ux, double* uy, double* uz, double* inv_gamma, const uint8_t* event,
                       const double* delta, int64_t N, const uint8_t* pruned) {
#pragma omp parallel for schedule(static)
  for (int64_t ip = 0; ip < N; ++ip) {
    if (pruned[ip]) continue;
    if (event[ip]) {
      ux[ip] *= 1 - delta[ip];
      uy[ip] *= 1 - delta[ip];
      uz[ip] *= 1 - delta[ip];
      inv_gamma[ip] = 1 / O_SQRT(1 + ux[ip] * ux[ip] + uy[ip] * uy[ip] + uz[ip] * uz[ip]);
    }
  }
}

/* cpu.py:90-118.  Additions over the reference: the appended photon gets id
 * next_id + k (k = its rank in the event list); chi and, when given, tau of
 * the new slot start at 0 (the reference leaves whatever a cyclic resize
 * copied there). */
void orc_create_photon(const double* x_src, const double* y_src, const double* z_src,
                       const double* ux_src, const double* uy_src, const double* uz_src,
                       double* x_dst, double* y_dst, double* z_dst, double* ux_dst, double* uy_dst,
                       double* uz_dst, double* inv_gamma_dst, uint8_t* photon_pruned,
                       const int64_t* event_index, const double* photon_delta, int64_t N_buffered,
                       int64_t N_photon, uint64_t* id_dst, uint64_t next_id, double* chi_dst,
                       double* tau_dst) {
#pragma omp parallel for schedule(static)
  for (int64_t ip = 0; ip < N_photon; ++ip) {
    int64_t idx_src = event_index[ip];
    int64_t idx_dst = N_buffered + ip;
    x_dst[idx_dst] = x_src[idx_src];
    y_dst[idx_dst] = y_src[idx_src];
    z_dst[idx_dst] = z_src[idx_src];
    ux_dst[idx_dst] = photon_delta[idx_src] * ux_src[idx_src];
    uy_dst[idx_dst] = photon_delta[idx_src] * uy_src[idx_src];
    uz_dst[idx_dst] = photon_delta[idx_src] * uz_src[idx_src];
    inv_gamma_dst[idx_dst] = 1.0 / O_SQRT(ux_dst[idx_dst] * ux_dst[idx_dst] +
                                          uy_dst[idx_dst] * uy_dst[idx_dst] +
                                          uz_dst[idx_dst] * uz_dst[idx_dst]);
    photon_pruned[idx_dst] = 0;
    id_dst[idx_dst] = next_id + (uint64_t)ip;
    if (chi_dst) chi_dst[idx_dst] = 0.0;
    if (tau_dst) tau_dst[idx_dst] = 0.0;
  }
}

/* cpu.py:120-191 (same additions as orc_create_photon; spin of the new
 * lepton, when given, starts at 0 -- the reference leaves it unset). */
void orc_create_pair(const double* x_src, const double* y_src, const double* z_src,
                     const double* ux_src, const double* uy_src, const double* uz_src,
                     uint8_t* photon_pruned, double* x_dst, double* y_dst, double* z_dst,
                     double* ux_dst, double* uy_dst, double* uz_dst, double* inv_gamma_dst,
                     uint8_t* pair_pruned, const int64_t* event_index, const double* pair_delta,
                     int64_t N_buffered, int64_t N_pair, int32_t inverse_delta, uint64_t* id_dst,
                     uint64_t next_id, double* chi_dst, double* tau_dst, double* sx_dst,
                     double* sy_dst, double* sz_dst) {
#pragma omp parallel for schedule(static)
  for (int64_t ip = 0; ip < N_pair; ++ip) {
    int64_t idx_src = event_index[ip];
    int64_t idx_dst = N_buffered + ip;
    x_dst[idx_dst] = x_src[idx_src];
    y_dst[idx_dst] = y_src[idx_src];
    z_dst[idx_dst] = z_src[idx_src];
    double delta = pair_delta[idx_src];
    if (inverse_delta) delta = 1.0 - delta;
    ux_dst[idx_dst] = delta * ux_src[idx_src];
    uy_dst[idx_dst] = delta * uy_src[idx_src];
    uz_dst[idx_dst] = delta * uz_src[idx_src];
    inv_gamma_dst[idx_dst] = 1.0 / O_SQRT(1.0 + ux_dst[idx_dst] * ux_dst[idx_dst] +
                                          uy_dst[idx_dst] * uy_dst[idx_dst] +
                                          uz_dst[idx_dst] * uz_dst[idx_dst]);
    pair_pruned[idx_dst] = 0;
    photon_pruned[idx_src] = 1;
    id_dst[idx_dst] = next_id + (uint64_t)ip;
    if (chi_dst) chi_dst[idx_dst] = 0.0;
    if (tau_dst) tau_dst[idx_dst] = 0.0;
    if (sx_dst) { sx_dst[idx_dst] = 0.0; sy_dst[idx_dst] = 0.0; sz_dst[idx_dst] = 0.0; }
  }
}

/* cpu.py:193-216 (serial, ascending order) */
int64_t orc_find_event_index(const uint8_t* event, int64_t N_event, int64_t N, int64_t* event_index) {
  int64_t idx = 0;
  for (int64_t i = 0; i < N; ++i) {
    if (event[i]) {
      if (idx < N_event) event_index[idx] = i;
      idx += 1;
    }
  }
  return idx;
}

/* cpu.py:218-225 (serial) */
int64_t orc_bool_sum(const uint8_t* event, int64_t N) {
  int64_t ntotal = 0;
  for (int64_t ip = 0; ip < N; ++ip)
    if (event[ip]) ntotal += 1;
  return ntotal;
}

/* cpu.py:227-232 */
void orc_pick_hard_photon(uint8_t* event, const double* delta, const double* inv_gamma,
                          double threshold, int64_t N) {
#pragma omp parallel for schedule(static)
  for (int64_t ip = 0; ip < N; ++ip) {
    if (event[ip]) {
      if (delta[ip] / inv_gamma[ip] < threshold) event[ip] = 0;
    }
  }
}

/* ========================================================= QED: tables */
typedef struct {
  const double* int_Ai;
  const double* Aip;
  int64_t z_N;
  double z_min, z_max, z_delta;
  const double* Ai; /* Ai(z) on the same grid, spin-resolved emission only */
} orc_rs_tables;

typedef struct {
  const double* photon_total; /* [chi_N] */
  const double* pair_total;   /* [chi_N] */
  const double* photon_cdf;   /* [chi_N][delta_N] */
  const double* pair_cdf;     /* [chi_N][delta_N] */
  int64_t chi_N;
  double log_chi_min, log_chi_max, log_chi_delta;
  int64_t delta_N;
  double log_delta_min, log_delta_max, log_delta_delta;
  const double* chi_grid;   /* [chi_N]  10**(log_chi_min + i*log_chi_delta), host libm pow */
  const double* delta_grid; /* [delta_N] 10**(log_delta_min + j*log_delta_delta) */
  double chi_min;           /* 10.0**log_chi_min (optical_depth.py:14) */
  double chi_max;           /* 10.0**log_chi_max */
} orc_od_tables;

/* linear interpolation of the two Airy tables at z (z_min <= z <= z_max) */
static inline void o_airy_lerp(const orc_rs_tables* T, double z, double* int_Ai_, double* Aip_) {
  int64_t idx = (int64_t)((z - T->z_min) / T->z_delta);
  if (idx > T->z_N - 2) idx = T->z_N - 2; /* z == z_max exactly */
  double z_left = T->z_min + (double)idx * T->z_delta;
  double k = (T->int_Ai[idx + 1] - T->int_Ai[idx]) / T->z_delta;
  *int_Ai_ = T->int_Ai[idx] + k * (z - z_left);
  k = (T->Aip[idx + 1] - T->Aip[idx]) / T->z_delta;
  *Aip_ = T->Aip[idx] + k * (z - z_left);
}

/* qed/rejection_sampling_tables.py:32-53; z > z_max -> 0 (reference: OOB read) */
double orc_photon_prob_rate(double chi_e, double delta, const orc_rs_tables* T) {
  double chi_gamma = delta * chi_e;
  double chi_ep = chi_e - chi_gamma;
  double z = O_POW(chi_gamma / chi_e / chi_ep, 2.0 / 3);
  double factor = -O_ALPHA * O_ME * (O_C * O_C) / O_HBAR;
  if (!(z <= T->z_max)) return (z != z) ? z : 0.0;
  double int_Ai_, Aip_;
  o_airy_lerp(T, z, &int_Ai_, &Aip_);
  return factor * (int_Ai_ + (2.0 / z + chi_gamma * O_SQRT(z)) * Aip_);
}

/* qed/rejection_sampling_tables.py:56-78 */
double orc_pair_prob_rate(double chi_gamma, double delta, const orc_rs_tables* T) {
  double chi_e = delta * chi_gamma;
  double chi_ep = chi_gamma - chi_e;
  double z = O_POW(chi_gamma / chi_e / chi_ep, 2.0 / 3);
  double factor = O_ALPHA * O_ME * (O_C * O_C) / O_HBAR;
  if (!(z <= T->z_max)) return (z != z) ? z : 0.0;
  double int_Ai_, Aip_;
  o_airy_lerp(T, z, &int_Ai_, &Aip_);
  return factor * (int_Ai_ + (2.0 / z - chi_gamma * O_SQRT(z)) * Aip_);
}

/* qed/rejection_sampling.py:68-85 with Philox draws (slot = slot_base):
 * r1 from words 0,1 and r2 from words 2,3. */
void orc_photon_from_rejection_sampling(const double* inv_gamma, const double* chi_e, double dt,
                                        int64_t N, const uint8_t* pruned, uint8_t* event,
                                        double* delta, const uint64_t* ids, uint64_t seed,
                                        uint64_t step, uint32_t tag, uint32_t slot_base,
                                        const orc_rs_tables* T) {
#pragma omp parallel for schedule(static)
  for (int64_t ip = 0; ip < N; ++ip) {
    if (pruned[ip] || chi_e[ip] == 0.0) {
      event[ip] = 0;
      continue;
    }
    sfp_u32x4 d = sfp_draw4(seed, ids[ip], step, tag, slot_base);
    double r1 = sfp_u01(d.v[0], d.v[1]);
    double r2 = sfp_u01(d.v[2], d.v[3]);
    double dtau = dt * inv_gamma[ip];
    double r1_3 = o_powi(r1, 3);
    double prob_rate = 3 * o_powi(r1, 2) * orc_photon_prob_rate(chi_e[ip], r1_3, T) * dtau;
    if (r2 < prob_rate) {
      delta[ip] = r1_3;
      event[ip] = 1;
    } else {
      delta[ip] = 0.0;
      event[ip] = 0;
    }
  }
}

/* Spin-resolved emission, an addition with no counterpart in the reference
 * (its emission ignores spin).  The LCFA emission probability summed over
 * photon polarisation and final spin carries the initial-spin term
 * -u (s.zeta) K_1/3(y) (u = delta, y = 2u/(3 chi (1-u))) beside the bracket
 * (1 - u + 1/(1-u)) K_2/3(y) - int_y^inf K_1/3 that the reference's table
 * formula restates in Airy form (divided by pi sqrt 3).  With
 * K_1/3(y) = pi sqrt(3/z) Ai(z) the term becomes -sigma delta Ai(z)/sqrt z,
 * written as -sigma chi_e' z Ai(z) since delta/sqrt z = chi_e' z. */
double orc_photon_prob_rate_spin(double chi_e, double delta, double sigma,
                                 const orc_rs_tables* T) {
  double chi_gamma = delta * chi_e;
  double chi_ep = chi_e - chi_gamma;
  double z = O_POW(chi_gamma / chi_e / chi_ep, 2.0 / 3);
  double factor = -O_ALPHA * O_ME * (O_C * O_C) / O_HBAR;
  if (!(z <= T->z_max)) return (z != z) ? z : 0.0;
  double int_Ai_, Aip_;
  o_airy_lerp(T, z, &int_Ai_, &Aip_);
  int64_t idx = (int64_t)((z - T->z_min) / T->z_delta);
  if (idx > T->z_N - 2) idx = T->z_N - 2;
  double z_left = T->z_min + (double)idx * T->z_delta;
  double k = (T->Ai[idx + 1] - T->Ai[idx]) / T->z_delta;
  double Ai_ = T->Ai[idx] + k * (z - z_left);
  return factor * (int_Ai_ + (2.0 / z + chi_gamma * O_SQRT(z)) * Aip_ + sigma * chi_ep * z * Ai_);
}

/* sigma = s.zeta, zeta the unit vector along beta x dbeta/dt under the
 * Lorentz force on a particle of charge sign q: since dbeta/dt is
 * q/(gamma m c) (E + c beta x B - beta (beta.E)), beta x dbeta/dt points along
 * sign(q) [gamma (u x E)/c + u x (u x B)]; 0 without a transverse force. */
double orc_spin_sigma(double q, double ux, double uy, double uz, double inv_gamma, double Ex,
                      double Ey, double Ez, double Bx, double By, double Bz, double sx, double sy,
                      double sz) {
  double gamma = 1.0 / inv_gamma;
  double ex = uy * Ez - uz * Ey, ey = uz * Ex - ux * Ez, ez = ux * Ey - uy * Ex;
  double bx = uy * Bz - uz * By, by = uz * Bx - ux * Bz, bz = ux * By - uy * Bx;
  double zx = gamma * ex / O_C + (uy * bz - uz * by);
  double zy = gamma * ey / O_C + (uz * bx - ux * bz);
  double zz = gamma * ez / O_C + (ux * by - uy * bx);
  if (q < 0.0) {
    zx = -zx;
    zy = -zy;
    zz = -zz;
  }
  double n = O_SQRT(zx * zx + zy * zy + zz * zz);
  if (!(n > 0.0)) return 0.0;
  return (sx * zx + sy * zy + sz * zz) / n;
}

/* qed/rejection_sampling.py:68-85 with the spin-resolved rate (the same
 * draws, the rate of orc_photon_prob_rate_spin at sigma = s.zeta from the
 * fields and momentum of this step's QED pass). */
void orc_photon_from_rejection_sampling_spin(
    const double* inv_gamma, const double* chi_e, double dt, int64_t N, const uint8_t* pruned,
    uint8_t* event, double* delta, const uint64_t* ids, uint64_t seed, uint64_t step,
    uint32_t tag, uint32_t slot_base, const orc_rs_tables* T, double q, const double* ux,
    const double* uy, const double* uz, const double* Ex, const double* Ey, const double* Ez,
    const double* Bx, const double* By, const double* Bz, const double* sx, const double* sy,
    const double* sz) {
#pragma omp parallel for schedule(static)
  for (int64_t ip = 0; ip < N; ++ip) {
    if (pruned[ip] || chi_e[ip] == 0.0) {
      event[ip] = 0;
      continue;
    }
    sfp_u32x4 d = sfp_draw4(seed, ids[ip], step, tag, slot_base);
    double r1 = sfp_u01(d.v[0], d.v[1]);
    double r2 = sfp_u01(d.v[2], d.v[3]);
    double dtau = dt * inv_gamma[ip];
    double r1_3 = o_powi(r1, 3);
    double sigma = orc_spin_sigma(q, ux[ip], uy[ip], uz[ip], inv_gamma[ip], Ex[ip], Ey[ip], Ez[ip],
                                  Bx[ip], By[ip], Bz[ip], sx[ip], sy[ip], sz[ip]);
    double prob_rate = 3 * o_powi(r1, 2) * orc_photon_prob_rate_spin(chi_e[ip], r1_3, sigma, T) * dtau;
    if (r2 < prob_rate) {
      delta[ip] = r1_3;
      event[ip] = 1;
    } else {
      delta[ip] = 0.0;
      event[ip] = 0;
    }
  }
}

/* qed/rejection_sampling.py:87-101 */
void orc_pair_from_rejection_sampling(const double* inv_gamma, const double* chi_gamma, double dt,
                                      int64_t N, const uint8_t* pruned, uint8_t* event,
                                      double* delta, const uint64_t* ids, uint64_t seed,
                                      uint64_t step, uint32_t tag, uint32_t slot_base,
                                      const orc_rs_tables* T) {
#pragma omp parallel for schedule(static)
  for (int64_t ip = 0; ip < N; ++ip) {
    if (pruned[ip] || chi_gamma[ip] == 0.0) {
      event[ip] = 0;
      continue;
    }
    sfp_u32x4 d = sfp_draw4(seed, ids[ip], step, tag, slot_base);
    double r1 = sfp_u01(d.v[0], d.v[1]);
    double r2 = sfp_u01(d.v[2], d.v[3]);
    double dtau = dt * inv_gamma[ip];
    double prob_rate = orc_pair_prob_rate(chi_gamma[ip], r1, T) * dtau;
    if (r2 < prob_rate) {
      delta[ip] = r1;
      event[ip] = 1;
    } else {
      delta[ip] = 0.0;
      event[ip] = 0;
    }
  }
}

/* ---------------------------------------------------- optical depth ---- */
/* qed/optical_depth_tables.py:34-44 with chi clamped to 10^log_chi_max.
 * Returns the fractional row index; -1 below the table. */
static inline double o_chi_idx(const orc_od_tables* T, double log_chi) {
  if (log_chi < T->log_chi_min) return -1.0;
  return (log_chi - T->log_chi_min) / T->log_chi_delta;
}

/* qed/optical_depth_tables.py:46-56 */
double orc_od_total_rate(const double* table1d, const orc_od_tables* T, double chi) {
  double log_chi = O_LOG10(chi);
  if (log_chi > T->log_chi_max) log_chi = T->log_chi_max;
  double fidx = o_chi_idx(T, log_chi);
  if (fidx < 0.0) return 0.0;
  int64_t idx = (int64_t)sfp_floor(fidx);
  if (idx > T->chi_N - 2) idx = T->chi_N - 2;
  double log_chi_left = T->log_chi_min + (double)idx * T->log_chi_delta;
  double k = (table1d[idx + 1] - table1d[idx]) / T->log_chi_delta;
  return table1d[idx] + k * (log_chi - log_chi_left);
}

/* qed/optical_depth_tables.py:59-101, with a lower-bound search (ties go
 * left) and the bracket index clamped into the table.  u is the uniform draw. */
double orc_od_inverse_cdf(const double* table2d, const orc_od_tables* T, double chi, double u) {
  int64_t DN = T->delta_N;
  if (chi > T->chi_max) chi = T->chi_max;
  double log_chi = O_LOG10(chi);
  if (log_chi > T->log_chi_max) log_chi = T->log_chi_max;
  double fidx = o_chi_idx(T, log_chi);
  if (fidx < 0.0) fidx = 0.0;
  int64_t lo_i = (int64_t)sfp_floor(fidx);
  if (lo_i > T->chi_N - 2) lo_i = T->chi_N - 2;
  int64_t hi_i = lo_i + 1;
  double chi_low = T->chi_grid[lo_i];
  double chi_high = T->chi_grid[hi_i];
  const double* rl = table2d + lo_i * DN;
  const double* rh = table2d + hi_i * DN;
  double dchi = chi_high - chi_low;
  double w = chi - chi_low;

  double k = (rh[0] - rl[0]) / dchi;
  double ymin = rl[0] + k * w;
  k = (rh[DN - 1] - rl[DN - 1]) / dchi;
  double ymax = rl[DN - 1] + k * w;
  double r = u * (ymax - ymin) + ymin;

  int64_t low = 0, high = DN - 1;
  while (low <= high) {
    int64_t mid = (low + high) / 2;
    k = (rh[mid] - rl[mid]) / dchi;
    double mid_delta = rl[mid] + k * w;
    if (mid_delta < r)
      low = mid + 1;
    else
      high = mid - 1;
  }
  int64_t di = high;
  if (di < 0) di = 0;
  if (di > DN - 2) di = DN - 2;
  k = (rh[di] - rl[di]) / dchi;
  double y1 = rl[di] + k * w;
  k = (rh[di + 1] - rl[di + 1]) / dchi;
  double y2 = rl[di + 1] + k * w;
  double delta_left = T->delta_grid[di];
  double delta_right = T->delta_grid[di + 1];
  k = (delta_right - delta_left) / (y2 - y1);
  return delta_left + k * (r - y1);
}

/* qed/optical_depth.py:15-62 (photon) and :64-110 (pair); draws: slot_base+1
 * word 0,1 for the initial tau, slot_base+2 words 0,1 for the re-drawn tau
 * and words 2,3 for the energy fraction. */
static void o_update_tau(double* tau, const double* inv_gamma, const double* chi, double dt,
                         int64_t N, const uint8_t* pruned, uint8_t* event, double* delta,
                         const uint64_t* ids, uint64_t seed, uint64_t step, uint32_t tag,
                         uint32_t slot_base, const orc_od_tables* T, int photon) {
  double chi_min = T->chi_min;
  const double* total = photon ? T->photon_total : T->pair_total;
  const double* cdf = photon ? T->photon_cdf : T->pair_cdf;
#pragma omp parallel for schedule(static)
  for (int64_t ip = 0; ip < N; ++ip) {
    if (pruned[ip] || !(chi[ip] >= chi_min)) { /* NaN chi (reference: undefined) -> no event */
      event[ip] = 0;
      delta[ip] = 0.0;
      continue;
    }
    double integ_prob_rate = orc_od_total_rate(total, T, chi[ip]);
    double dtau = dt * inv_gamma[ip];
    if (tau[ip] == 0.0) {
      sfp_u32x4 d = sfp_draw4(seed, ids[ip], step, tag, slot_base + 1);
      tau[ip] = -O_LOG(1 - sfp_u01(d.v[0], d.v[1]));
    }
    tau[ip] -= integ_prob_rate * dtau;
    if (tau[ip] < 0) {
      sfp_u32x4 d = sfp_draw4(seed, ids[ip], step, tag, slot_base + 2);
      tau[ip] = -O_LOG(1 - sfp_u01(d.v[0], d.v[1]));
      event[ip] = 1;
      delta[ip] = orc_od_inverse_cdf(cdf, T, chi[ip], sfp_u01(d.v[2], d.v[3]));
    } else {
      event[ip] = 0;
      delta[ip] = 0.0;
    }
  }
}

void orc_update_tau_e(double* tau, const double* inv_gamma, const double* chi, double dt, int64_t N,
                      const uint8_t* pruned, uint8_t* event, double* delta, const uint64_t* ids,
                      uint64_t seed, uint64_t step, uint32_t tag, uint32_t slot_base,
                      const orc_od_tables* T) {
  o_update_tau(tau, inv_gamma, chi, dt, N, pruned, event, delta, ids, seed, step, tag, slot_base, T, 1);
}

void orc_update_tau_gamma(double* tau, const double* inv_gamma, const double* chi, double dt,
                          int64_t N, const uint8_t* pruned, uint8_t* event, double* delta,
                          const uint64_t* ids, uint64_t seed, uint64_t step, uint32_t tag,
                          uint32_t slot_base, const orc_od_tables* T) {
  o_update_tau(tau, inv_gamma, chi, dt, N, pruned, event, delta, ids, seed, step, tag, slot_base, T, 0);
}

/* ================================================== elementary helpers */
/* Philox block and uniform conversion, exported for known-answer tests */
void orc_philox4x32_10(const uint32_t* ctr, const uint32_t* key, uint32_t* out) {
  sfp_u32x4 r = sfp_philox4x32_10(ctr[0], ctr[1], ctr[2], ctr[3], key[0], key[1]);
  memcpy(out, r.v, 16);
}

void orc_draw4(uint64_t seed, uint64_t id, uint64_t step, uint32_t tag, uint32_t slot, uint32_t* out) {
  sfp_u32x4 r = sfp_draw4(seed, id, step, tag, slot);
  memcpy(out, r.v, 16);
}

/* vectorised elementary functions (own implementation), for accuracy tests */
void orc_math(int32_t fn, const double* x, const double* y, double* out, int64_t n) {
  for (int64_t i = 0; i < n; ++i) {
    switch (fn) {
      case 0: out[i] = sfp_exp(x[i]); break;
      case 1: out[i] = sfp_log(x[i]); break;
      case 2: out[i] = sfp_sin(x[i]); break;
      case 3: out[i] = sfp_cos(x[i]); break;
      case 4: out[i] = sfp_atan(x[i]); break;
      case 5: out[i] = sfp_pow(x[i], y[i]); break;
      case 6: out[i] = sfp_exp10(x[i]); break;
      default: out[i] = sfp_log10(x[i]); break;
    }
  }
}

/* number of threads OpenMP will use (reported as the CPU baseline's cores) */
#ifdef _OPENMP
#include <omp.h>
int32_t orc_num_threads(void) { return omp_get_max_threads(); }
#else
int32_t orc_num_threads(void) { return 1; }
#endif
