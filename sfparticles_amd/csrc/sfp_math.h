/*
 * sfp_math.h -- double-precision elementary functions and the Philox4x32-10
 * counter-based generator, written once and compiled both for the gfx950
 * device (hipcc) and for the host (gcc / hipcc host pass).
 *
 * Why this exists: photon-emission and pair-creation decisions are
 * comparisons `r < p(chi)`.  Replaying the GPU's decisions bit for bit on the
 * CPU needs chi and p to be bit-identical on both sides, so every
 * transcendental used on the hot path comes from this file instead of the
 * platform libm (glibc on the host, OCML on the device), whose last-ulp
 * behaviour differs.  Only IEEE-754 correctly rounded operations are used:
 * +, -, *, /, sqrt, fma, rint, floor and exact bit manipulation.  Callers
 * must be compiled with floating-point contraction disabled
 * (-ffp-contract=off) so that `a*b+c` is never silently fused on one side
 * only; where a fused multiply-add is wanted it is written as sfp_fma().
 *
 * Methods (all written from the mathematics; coefficients are the exact
 * Taylor / arctanh-series coefficients 1/k! and 1/(2k+1) rounded to double):
 *   exp   : Cody & Waite (1980) additive reduction x = k ln2 + r, |r| <= ln2/2,
 *           two-part ln2 split, degree-14 Taylor polynomial, exact 2^k scaling.
 *   log   : x = 2^e m, m in [sqrt(1/2), sqrt(2)); log m = 2 artanh(s),
 *           s = (m-1)/(m+1), odd series to s^25.
 *   sin/cos: Cody-Waite three-part pi/2 reduction (33+33+53 bits), Taylor
 *           polynomials to r^17 / r^18 on |r| <= pi/4, quadrant selection.
 *   atan  : atan(x) = pi/2 - atan(1/x) for |x| > 1, then two half-angle steps
 *           t -> t / (1 + sqrt(1 + t^2)) to |t| <= tan(pi/16), odd series to t^27.
 *   pow   : exp(y log x) for x > 0.
 *   exp10 : exp(x ln10) with the product's rounding error recovered by fma.
 * Accuracy (checked in the test-suite against numpy): <= 2 ulp for exp, log,
 * sin and cos, <= 3 ulp for log10 and exp10, <= 4 ulp for atan, over the
 * argument ranges the particle step uses.
 *
 * Philox4x32-10: Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
 * easy as 1, 2, 3", SC'11.  Multipliers 0xD2511F53 / 0xCD9E8D57, Weyl key
 * increments 0x9E3779B9 / 0xBB67AE85, ten rounds.
 */
#ifndef SFP_MATH_H
#define SFP_MATH_H

#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define SFP_HD __host__ __device__ static inline __attribute__((always_inline))
#else
#define SFP_HD static inline __attribute__((always_inline))
#endif

#define sfp_fma(a, b, c) __builtin_fma((a), (b), (c))
#define sfp_sqrt(a) __builtin_sqrt(a)
#define sfp_floor(a) __builtin_floor(a)
#define sfp_rint(a) __builtin_rint(a)

SFP_HD uint64_t sfp_d2u(double x) { uint64_t u; __builtin_memcpy(&u, &x, 8); return u; }
SFP_HD double sfp_u2d(uint64_t u) { double x; __builtin_memcpy(&x, &u, 8); return x; }

/* 2^k for -1022 <= k <= 1023, exact. */
SFP_HD double sfp_pow2i(int k) { return sfp_u2d((uint64_t)(k + 1023) << 52); }

/* ---------------------------------------------------------------- exp -- */
#define SFP_LN2_HI 0x1.62e42fefa3800p-1 /* ln2 truncated to 42 bits */
#define SFP_LN2_LO 5.497923018708371e-14 /* ln2 - SFP_LN2_HI */
#define SFP_INV_LN2 1.4426950408889634

SFP_HD double sfp_exp(double x) {
  if (x != x) return x;
  if (x > 709.782712893384) return __builtin_inf();
  if (x < -745.1332191019412) return 0.0;
  double kd = sfp_rint(x * SFP_INV_LN2);
  int k = (int)kd;
  double r = sfp_fma(-kd, SFP_LN2_HI, x);
  r = sfp_fma(-kd, SFP_LN2_LO, r);
  /* exp(r), |r| <= 0.3466, Taylor to r^14 (Horner) */
  double p = 1.1470745597729725e-11;           /* 1/14! */
  p = sfp_fma(p, r, 1.6059043836821613e-10);   /* 1/13! */
  p = sfp_fma(p, r, 2.08767569878681e-09);     /* 1/12! */
  p = sfp_fma(p, r, 2.505210838544172e-08);    /* 1/11! */
  p = sfp_fma(p, r, 2.755731922398589e-07);    /* 1/10! */
  p = sfp_fma(p, r, 2.7557319223985893e-06);   /* 1/9!  */
  p = sfp_fma(p, r, 2.48015873015873e-05);     /* 1/8!  */
  p = sfp_fma(p, r, 0.0001984126984126984);    /* 1/7!  */
  p = sfp_fma(p, r, 0.001388888888888889);     /* 1/6!  */
  p = sfp_fma(p, r, 0.008333333333333333);     /* 1/5!  */
  p = sfp_fma(p, r, 0.041666666666666664);     /* 1/4!  */
  p = sfp_fma(p, r, 0.16666666666666666);      /* 1/3!  */
  p = sfp_fma(p, r, 0.5);                      /* 1/2!  */
  p = sfp_fma(p, r, 1.0);
  p = sfp_fma(p, r, 1.0);
  if (k > 1023) return p * sfp_pow2i(1023) * sfp_pow2i(k - 1023);
  if (k < -1021) return (p * sfp_pow2i(k + 600)) * sfp_pow2i(-600);
  return p * sfp_pow2i(k);
}

/* ---------------------------------------------------------------- log -- */
SFP_HD double sfp_log(double x) {
  if (x != x) return x;
  if (x < 0.0) return __builtin_nan("");
  if (x == 0.0) return -__builtin_inf();
  if (x == __builtin_inf()) return x;
  int e = 0;
  if (x < 0x1p-1022) { x *= 0x1p54; e = -54; }
  uint64_t u = sfp_d2u(x);
  e += (int)((u >> 52) & 0x7ff) - 1023;
  double m = sfp_u2d((u & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL); /* [1,2) */
  if (m > 1.4142135623730951) { m *= 0.5; e += 1; }
  double s = (m - 1.0) / (m + 1.0);
  double s2 = s * s;
  double p = 0.04; /* 1/25 */
  p = sfp_fma(p, s2, 0.043478260869565216);  /* 1/23 */
  p = sfp_fma(p, s2, 0.047619047619047616);  /* 1/21 */
  p = sfp_fma(p, s2, 0.05263157894736842);   /* 1/19 */
  p = sfp_fma(p, s2, 0.058823529411764705);  /* 1/17 */
  p = sfp_fma(p, s2, 0.06666666666666667);   /* 1/15 */
  p = sfp_fma(p, s2, 0.07692307692307693);   /* 1/13 */
  p = sfp_fma(p, s2, 0.09090909090909091);   /* 1/11 */
  p = sfp_fma(p, s2, 0.1111111111111111);    /* 1/9 */
  p = sfp_fma(p, s2, 0.14285714285714285);   /* 1/7 */
  p = sfp_fma(p, s2, 0.2);                   /* 1/5 */
  p = sfp_fma(p, s2, 0.3333333333333333);    /* 1/3 */
  double ls = 2.0 * s;                       /* exact */
  double lm = sfp_fma(ls * s2, p, ls);       /* 2s + 2s^3 (1/3 + ...) */
  double ed = (double)e;
  return sfp_fma(ed, SFP_LN2_HI, sfp_fma(ed, SFP_LN2_LO, lm));
}

#define SFP_INV_LN10 0.4342944819032518
#define SFP_LN10 2.302585092994046
#define SFP_LN10_LO -2.1707562233822494e-16

SFP_HD double sfp_log10(double x) { return sfp_log(x) * SFP_INV_LN10; }

/* 10^x */
SFP_HD double sfp_exp10(double x) {
  double t = x * SFP_LN10;
  double err = sfp_fma(x, SFP_LN10, -t) + x * SFP_LN10_LO;
  double r = sfp_exp(t);
  return sfp_fma(r, err, r);
}

/* x^y for x >= 0 (x == 0 handled; negative x returns NaN) */
SFP_HD double sfp_pow(double x, double y) {
  if (x == 0.0) return (y > 0.0) ? 0.0 : ((y == 0.0) ? 1.0 : __builtin_inf());
  return sfp_exp(y * sfp_log(x));
}

/* ------------------------------------------------------------ sin/cos -- */
#define SFP_PIO2_1 0x1.921fb54400000p+0   /* pi/2, first 33 bits  */
#define SFP_PIO2_2 0x1.0b4611a600000p-34  /* next 33 bits         */
#define SFP_PIO2_3 0x1.3198a2e037073p-69  /* remainder            */
#define SFP_2_OVER_PI 0.6366197723675814

SFP_HD double sfp_sin_poly(double r) {
  double r2 = r * r;
  double p = 2.8114572543455206e-15;          /* 1/17! */
  p = sfp_fma(p, r2, -7.647163731819816e-13);  /* -1/15! */
  p = sfp_fma(p, r2, 1.6059043836821613e-10);  /* 1/13! */
  p = sfp_fma(p, r2, -2.505210838544172e-08);  /* -1/11! */
  p = sfp_fma(p, r2, 2.7557319223985893e-06);  /* 1/9! */
  p = sfp_fma(p, r2, -0.0001984126984126984);  /* -1/7! */
  p = sfp_fma(p, r2, 0.008333333333333333);    /* 1/5! */
  p = sfp_fma(p, r2, -0.16666666666666666);    /* -1/3! */
  return sfp_fma(r * r2, p, r);
}

SFP_HD double sfp_cos_poly(double r) {
  double r2 = r * r;
  /* cos r = 1 - r^2/2! + r^4/4! - r^6/6! + ... - r^18/18! */
  double p = -1.5619206968586225e-16;         /* -1/18! */
  p = sfp_fma(p, r2, 4.779477332387385e-14);   /* 1/16! */
  p = sfp_fma(p, r2, -1.1470745597729725e-11); /* -1/14! */
  p = sfp_fma(p, r2, 2.08767569878681e-09);    /* 1/12! */
  p = sfp_fma(p, r2, -2.755731922398589e-07);  /* -1/10! */
  p = sfp_fma(p, r2, 2.48015873015873e-05);    /* 1/8! */
  p = sfp_fma(p, r2, -0.001388888888888889);   /* -1/6! */
  p = sfp_fma(p, r2, 0.041666666666666664);    /* 1/4! */
  /* 1 - r2/2 + r2^2 p, with the leading terms summed for accuracy */
  double h = 0.5 * r2;
  double w = 1.0 - h;
  double corr = ((1.0 - w) - h);
  return w + sfp_fma(r2 * r2, p, corr);
}

/* reduce x to r in [-pi/4, pi/4] and quadrant q (mod 4) */
SFP_HD double sfp_reduce_pio2(double x, int* q) {
  double kd = sfp_rint(x * SFP_2_OVER_PI);
  double r = sfp_fma(-kd, SFP_PIO2_1, x);
  r = sfp_fma(-kd, SFP_PIO2_2, r);
  r = sfp_fma(-kd, SFP_PIO2_3, r);
  /* kd can exceed int range for huge |x|; only its value mod 4 matters */
  double k4 = kd - 4.0 * sfp_floor(kd * 0.25);
  *q = (int)k4;
  return r;
}

/* Both polynomials are evaluated and the quadrant selects (no divergent
 * branches across a wavefront). */
SFP_HD double sfp_sin(double x) {
  if (x != x || x == __builtin_inf() || x == -__builtin_inf()) return __builtin_nan("");
  int q;
  double r = sfp_reduce_pio2(x, &q);
  double s = sfp_sin_poly(r), c = sfp_cos_poly(r);
  double v = (q & 1) ? c : s;
  return (q & 2) ? -v : v;
}

SFP_HD double sfp_cos(double x) {
  if (x != x || x == __builtin_inf() || x == -__builtin_inf()) return __builtin_nan("");
  int q;
  double r = sfp_reduce_pio2(x, &q);
  double s = sfp_sin_poly(r), c = sfp_cos_poly(r);
  double v = (q & 1) ? s : c;
  return ((q + 1) & 2) ? -v : v;
}

/* --------------------------------------------------------------- atan -- */
#define SFP_PIO2_HI 1.5707963267948966
#define SFP_PIO2_LO 6.123233995736766e-17

SFP_HD double sfp_atan_series(double t) {
  double t2 = t * t;
  double p = 0.037037037037037035;            /* 1/27 */
  p = sfp_fma(p, t2, -0.04);                  /* -1/25 */
  p = sfp_fma(p, t2, 0.043478260869565216);   /* 1/23 */
  p = sfp_fma(p, t2, -0.047619047619047616);  /* -1/21 */
  p = sfp_fma(p, t2, 0.05263157894736842);    /* 1/19 */
  p = sfp_fma(p, t2, -0.058823529411764705);  /* -1/17 */
  p = sfp_fma(p, t2, 0.06666666666666667);    /* 1/15 */
  p = sfp_fma(p, t2, -0.07692307692307693);   /* -1/13 */
  p = sfp_fma(p, t2, 0.09090909090909091);    /* 1/11 */
  p = sfp_fma(p, t2, -0.1111111111111111);    /* -1/9 */
  p = sfp_fma(p, t2, 0.14285714285714285);    /* 1/7 */
  p = sfp_fma(p, t2, -0.2);                   /* -1/5 */
  p = sfp_fma(p, t2, 0.3333333333333333);     /* 1/3 */
  return sfp_fma(-(t * t2), p, t);            /* t - t^3/3 + ... */
}

SFP_HD double sfp_atan(double x) {
  if (x != x) return x;
  double ax = x < 0.0 ? -x : x;
  int inv = ax > 1.0;
  double t = inv ? 1.0 / ax : ax;
  /* two half-angle steps: atan(t) = 2 atan(t / (1 + sqrt(1 + t^2))) */
  t = t / (1.0 + sfp_sqrt(sfp_fma(t, t, 1.0)));
  t = t / (1.0 + sfp_sqrt(sfp_fma(t, t, 1.0)));
  double a = 4.0 * sfp_atan_series(t);
  if (inv) a = (SFP_PIO2_HI - a) + SFP_PIO2_LO;
  return x < 0.0 ? -a : a;
}

/* x^n, n >= 0, by binary exponentiation -- the multiplication order of the
 * float**int lowering used by numba (LLVM powi). */
SFP_HD double sfp_powi(double a, int n) {
  double r = 1.0;
  while (1) {
    if (n & 1) r = r * a;
    n >>= 1;
    if (n == 0) break;
    a = a * a;
  }
  return r;
}

/* ------------------------------------------------------- Philox4x32-10 -- */
typedef struct { uint32_t v[4]; } sfp_u32x4;

SFP_HD void sfp_mulhilo32(uint32_t a, uint32_t b, uint32_t* hi, uint32_t* lo) {
  uint64_t p = (uint64_t)a * (uint64_t)b;
  *hi = (uint32_t)(p >> 32);
  *lo = (uint32_t)p;
}

SFP_HD sfp_u32x4 sfp_philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                   uint32_t k0, uint32_t k1) {
  for (int round = 0; round < 10; ++round) {
    uint32_t hi0, lo0, hi1, lo1;
    sfp_mulhilo32(0xD2511F53u, c0, &hi0, &lo0);
    sfp_mulhilo32(0xCD9E8D57u, c2, &hi1, &lo1);
    uint32_t n0 = hi1 ^ c1 ^ k0;
    uint32_t n1 = lo1;
    uint32_t n2 = hi0 ^ c3 ^ k1;
    uint32_t n3 = lo0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  sfp_u32x4 out;
  out.v[0] = c0; out.v[1] = c1; out.v[2] = c2; out.v[3] = c3;
  return out;
}

/* 52 random bits from two words -> uniform double strictly inside (0,1):
 * u = (k + 1/2) 2^-52, k in [0, 2^52); k + 1/2 is exact, so the extremes are
 * 2^-53 and 1 - 2^-53 (with 53 bits, k + 1/2 would round up to 2^53 and give 1). */
SFP_HD double sfp_u01(uint32_t hi, uint32_t lo) {
  uint64_t k = ((uint64_t)hi << 20) | (uint64_t)(lo >> 12);
  return ((double)k + 0.5) * 0x1p-52;
}

/* The draw stream of one particle in one step.  Counter words:
 *   c0,c1 = particle id (low, high 32 bits), c2 = step index (low 32 bits),
 *   c3 = (species tag << 16) | (step index high bits & 0xff) << 8 | slot.
 * Key = 64-bit seed. */
SFP_HD sfp_u32x4 sfp_draw4(uint64_t seed, uint64_t id, uint64_t step, uint32_t species_tag,
                           uint32_t slot) {
  uint32_t c3 = (species_tag << 16) | ((uint32_t)((step >> 32) & 0xffu) << 8) | (slot & 0xffu);
  return sfp_philox4x32_10((uint32_t)id, (uint32_t)(id >> 32), (uint32_t)step, c3,
                           (uint32_t)seed, (uint32_t)(seed >> 32));
}

/* ------------------------------------------------------- repeated add -- */
/* Exactly the result of n sequential additions x <- x + h (IEEE double,
 * round to nearest even), in a number of operations proportional to the
 * binades x crosses rather than to n.
 *
 * Inside a binade [2^e, 2^(e+1)) the doubles are the integer multiples k u
 * of u = 2^(e-52), k in [2^52, 2^53).  While the exact sum (k + h/u) u stays
 * inside the binade, x + h rounds to (k + q) u with q the integer nearest
 * r = h/u, the same every step.  A tie (r - floor(r) = 1/2) rounds to the
 * even neighbour: once k is even the increment is the even one of floor(r),
 * floor(r) + 1 and k stays even, so the increment is again constant after at
 * most one real addition.  s steps stay inside the binade when
 * 2^52 + 1 <= k + s q <= 2^53 - 1 (this covers |r - q| <= 1/2).  Binade
 * crossings, zeros, subnormal or huge values and non-finite operands take
 * single real additions.  Negative x uses the mirror image (rounding to
 * nearest is symmetric). */
/* The usual case in one piece: every one of the n additions stays inside
 * x's binade (and is no tie).  All in FP64 and exact: k = |x| / u and
 * r = h / u are scalings by a power of two, q = round(r) and the room
 * k - (2^52 + 1) or (2^53 - 1) - k are integers below 2^53, and so is the
 * product n |q| whenever it does not exceed the room (a product of integers
 * that is below 2^53 is exact; one at or above rounds to at least 2^53), so
 * the test n |q| <= room and the sum k + n q are exact.  Returns 0 when the
 * case does not apply (the caller then takes the loop below). */
SFP_HD int sfp_repeat_add_binade(double x, double h, uint32_t n, double* out) {
  const double ax = x < 0.0 ? -x : x;
  const int be = (int)(sfp_d2u(ax) >> 52);
  if (be < 54 || be > 2045) return 0;  /* zero, tiny, huge or non-finite x */
  const double sc = sfp_u2d((uint64_t)(2098 - be) << 52);  /* 2^(1075 - be) = 1 / u */
  const double r = (x < 0.0 ? -h : h) * sc;
  if (!(r < 4503599627370496.0 && r > -4503599627370496.0)) return 0;  /* also NaN h */
  const double fr = sfp_floor(r);
  const double f = r - fr;
  if (f == 0.5) return 0;  /* a tie: the parity of k decides */
  const double q = f < 0.5 ? fr : fr + 1.0;
  const double k = ax * sc;  /* the integer significand, in [2^52, 2^53) */
  if (q == 0.0) {  /* x does not move unless it sits at the binade's bottom */
    if (r >= 0.0 || k >= 4503599627370497.0) {
      *out = x;
      return 1;
    }
    return 0;
  }
  const double room = q > 0.0 ? 9007199254740991.0 - k : k - 4503599627370497.0;
  const double aq = q > 0.0 ? q : -q;
  if (!((double)n * aq <= room)) return 0;
  const double X = (k + (double)n * q) * sfp_u2d((uint64_t)(be - 52) << 52);
  *out = x < 0.0 ? -X : X;
  return 1;
}

SFP_HD double sfp_repeat_add_closed(double x, double h, uint32_t n) {
  if (n == 0) return x;
  double y1;
  if (sfp_repeat_add_binade(x, h, n, &y1)) return y1;
  if (h == 0.0 || !(h - h == 0.0) || !(x - x == 0.0)) {
    // the first addition fixes the sign of a zero / produces inf or NaN;
    // every later one returns its input
    return x + h;
  }
  const int64_t LO = (int64_t)1 << 52, HI = (int64_t)1 << 53;
  while (n > 0) {
    const int neg = x < 0.0;
    const double ax = neg ? -x : x;
    const uint64_t bits = sfp_d2u(ax);
    const int be = (int)(bits >> 52);
    if (be < 54 || be > 2045) {  // zero, tiny or huge: one real addition
      x = x + h;
      --n;
      continue;
    }
    const double u = sfp_u2d((uint64_t)(be - 52) << 52);
    const double H = neg ? -h : h;
    // r = H / u, as a product with the exact reciprocal 2^(1075 - be) of the
    // power of two u (the same rounding, no division sequence)
    const double r = H * sfp_u2d((uint64_t)(2098 - be) << 52);
    if (!(r < 4503599627370496.0 && r > -4503599627370496.0)) {
      x = x + h;
      --n;
      continue;
    }
    int64_t k = (int64_t)((bits & (((uint64_t)1 << 52) - 1)) | ((uint64_t)1 << 52));
    const double fr = sfp_floor(r);
    const double f = r - fr;  // exact
    int64_t q;
    if (f == 0.5) {
      if (k & 1) {  // the parity of k decides this one rounding
        x = x + h;
        --n;
        continue;
      }
      const int64_t fi = (int64_t)fr;
      q = (fi & 1) ? fi + 1 : fi;
    } else {
      q = (int64_t)(f < 0.5 ? fr : fr + 1.0);
    }
    int64_t m;  // steps that stay inside the binade
    if (q != 0) {
      // floor(room / |q|) with room, |q| < 2^53 exact in double; the rounded
      // quotient is off by at most one, which the integer checks correct
      const int64_t room = (q > 0) ? HI - 1 - k : k - LO - 1;
      const int64_t aq = (q > 0) ? q : -q;
      if (room < 0) {
        m = 0;
      } else {
        m = (int64_t)sfp_floor((double)room / (double)aq);
        if (m * aq > room) --m;
        else if ((m + 1) * aq <= room) ++m;
      }
    } else {
      m = (r >= 0.0 || k >= LO + 1) ? (int64_t)n : 0;  // x does not move
    }
    int64_t s = (m < (int64_t)n) ? m : (int64_t)n;
    if (s > 0) {
      k += s * q;
      const double X = (double)k * u;  // exact: integer below 2^53 times 2^j
      x = neg ? -X : X;
      n -= (uint32_t)s;
    }
    if (n > 0) {  // the next addition leaves the binade: a real one
      x = x + h;
      --n;
    }
  }
  return x;
}

/* n additions: up to SFP_REPEAT_LOOP of them one by one (the definition;
 * branch-free, and cheaper than the closed form's control flow for a short
 * sleep, such as the 2 x 20 of a 20-step run), more in closed form. */
#define SFP_REPEAT_LOOP 64u
SFP_HD double sfp_repeat_add(double x, double h, uint32_t n) {
  if (n <= SFP_REPEAT_LOOP) {
    for (uint32_t k = 0; k < n; ++k) x = x + h;
    return x;
  }
  return sfp_repeat_add_closed(x, h, n);
}

#endif /* SFP_MATH_H */
