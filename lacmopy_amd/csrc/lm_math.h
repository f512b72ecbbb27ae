/*
 * lm_math.h -- FP64 exp / log / pow for the super-droplet step, written once
 * and compiled twice: by hipcc for gfx950 device code and by gcc for the CPU
 * oracle.  Both builds use only IEEE-754 add/sub/mul/div, fma and rint, all of
 * which are correctly rounded on x86-64 (SSE2 + FMA3) and on CDNA4
 * (v_fma_f64, the correctly rounded v_div_scale/fmas/fixup division), so every
 * result is bit-identical on host and device.  Compile with
 * -ffp-contract=off on both sides: no a*b+c may be fused behind our back.
 *
 * Method (own implementation, standard table-driven techniques):
 *   log: x = 2^k m, m in [sqrt(1/2), sqrt(2)); i = rint(128 (m-1));
 *        c_i ~ 1/(1+i/128) (table);  r = m c_i - 1 exactly as a double-double
 *        (TwoProd); log x = k ln2 + (-log c_i) + log1p(r), with log1p(r)
 *        = r - r^2/2 + r^3 P(r), |r| < 2^-7, carried in double-double.
 *   exp: n = rint(x 128/ln2), x = n ln2/128 + r, |r| <= ln2/256 (Cody-Waite
 *        split, exact first step); exp x = 2^(n>>7) 2^((n&127)/128) e^r,
 *        2^(j/128) from a double-double table, e^r by Taylor to r^7.
 *   pow: x^y = exp(y log x) with log x and the product y log x kept in
 *        double-double, so the error of the exponent does not blow up.
 * Measured against glibc 2.35 and 200-bit mpmath
 * (tests/test_oracle_cpu.py::test_lm_math_against_glibc_and_mpmath, and
 * device against host in tests/test_gpu_math.py): max error below 1 ulp,
 * differing from glibc in under 0.2% of the samples, and where they differ
 * this library is the correctly rounded one (glibc itself is not correctly
 * rounded in every case).
 *
 * The reference (Numba -> LLVM -> libm) evaluates math.log, np.exp and the
 * ** operator with glibc: collision/all_or_nothing.py:62 (log),
 * microphysics.py:100 (pow(x, 1/3)), microphysics.py:1440 (exp),
 * material_properties.py:244,478,499,522,563 (pow / exp).
 */
#ifndef LM_MATH_H
#define LM_MATH_H

#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define LM_FN __device__ static inline
#define LM_TABLE __constant__ static const
#define LM_FMA(a, b, c) __builtin_fma((a), (b), (c))
#define LM_FMOD(a, b) fmod((a), (b))
#else
#include <math.h>
#define LM_FN static inline
#define LM_TABLE static const
#define LM_FMA(a, b, c) fma((a), (b), (c))
#define LM_FMOD(a, b) fmod((a), (b))
#endif

#include "lm_tables.h"

#define LM_LN2_HI 0x1.62e42fee00000p-1
#define LM_LN2_LO 0x1.a39ef35793c76p-33
#define LM_SQRT2 0x1.6a09e667f3bcdp+0
#define LM_EXP_INV 0x1.71547652b82fep+7 /* 128/ln2 */
#define LM_EXP_CHI 0x1.62e42fef80000p-8 /* ln2/128, 34 significant bits */
#define LM_EXP_CLO 0x1.1cf79abc9e3b4p-43
#define LM_INF (__builtin_inf())
#define LM_NAN (__builtin_nan(""))

LM_FN uint64_t lm_as_u64(double x) {
  uint64_t u;
  __builtin_memcpy(&u, &x, 8);
  return u;
}

LM_FN double lm_as_f64(uint64_t u) {
  double x;
  __builtin_memcpy(&x, &u, 8);
  return x;
}

LM_FN double lm_rint(double x) { return __builtin_rint(x); }

/* s + e == a + b exactly (Knuth TwoSum). */
LM_FN void lm_two_sum(double a, double b, double *s, double *e) {
  double ss = a + b;
  double bb = ss - a;
  *e = (a - (ss - bb)) + (b - bb);
  *s = ss;
}

/* p + e == a * b exactly. */
LM_FN void lm_two_prod(double a, double b, double *p, double *e) {
  double pp = a * b;
  *e = LM_FMA(a, b, -pp);
  *p = pp;
}

/* 2^k for k in [-1022, 1023]. */
LM_FN double lm_pow2i(int k) { return lm_as_f64((uint64_t)(k + 1023) << 52); }

/* log(x) as hi + lo for finite x > 0. */
LM_FN void lm_log_dd(double x, double *rh, double *rl) {
  uint64_t u = lm_as_u64(x);
  int k = 0;
  if ((u >> 52) == 0) { /* subnormal */
    x *= 0x1p54;
    u = lm_as_u64(x);
    k = -54;
  }
  k += (int)(u >> 52) - 1023;
  double m = lm_as_f64((u & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull);
  if (m > LM_SQRT2) {
    m *= 0.5;
    k += 1;
  }
  int i = (int)lm_rint((m - 1.0) * 128.0);
  const int t = 3 * (i - LM_LOG_IMIN);
  const double c = lm_log_tab[t];
  const double L_hi = lm_log_tab[t + 1];
  const double L_lo = lm_log_tab[t + 2];

  double p, pe;
  lm_two_prod(m, c, &p, &pe);
  double r_hi, r_lo;
  lm_two_sum(p - 1.0, pe, &r_hi, &r_lo); /* p - 1 is exact: p in [0.99, 1.01] */

  double q, qe;
  lm_two_prod(r_hi, r_hi, &q, &qe);
  const double r = r_hi;
  double P = -1.0 / 10.0;
  P = P * r + 1.0 / 9.0;
  P = P * r - 1.0 / 8.0;
  P = P * r + 1.0 / 7.0;
  P = P * r - 1.0 / 6.0;
  P = P * r + 1.0 / 5.0;
  P = P * r - 1.0 / 4.0;
  P = P * r + 1.0 / 3.0;
  const double small = ((r_lo - 0.5 * qe) - r * r_lo) + q * r_lo + (q * r) * P;

  const double kd = (double)k;
  double a, ae, b, be, s, se;
  lm_two_sum(kd * LM_LN2_HI, L_hi, &a, &ae);
  lm_two_sum(r_hi, -0.5 * q, &b, &be);
  lm_two_sum(a, b, &s, &se);
  const double lo = se + ae + be + (kd * LM_LN2_LO + L_lo) + small;
  double h, l;
  lm_two_sum(s, lo, &h, &l);
  *rh = h;
  *rl = l;
}

/* exp(xh + xl), |xl| <= ulp(xh). */
LM_FN double lm_exp_dd(double xh, double xl) {
  if (xh != xh) return xh;
  if (xh > 709.79) return LM_INF;
  if (xh < -745.2) return 0.0;
  const double nd = lm_rint(xh * LM_EXP_INV);
  const int n = (int)nd;
  const double r1 = xh - nd * LM_EXP_CHI; /* exact (Sterbenz) */
  double r_hi, r_lo;
  lm_two_sum(r1, -nd * LM_EXP_CLO, &r_hi, &r_lo);
  r_lo += xl;
  const double r = r_hi;
  double P = 1.0 / 5040.0;
  P = P * r + 1.0 / 720.0;
  P = P * r + 1.0 / 120.0;
  P = P * r + 1.0 / 24.0;
  P = P * r + 1.0 / 6.0;
  P = P * r + 0.5;
  const double p_hi = r_hi;
  const double p_lo = r_lo + r_hi * r_lo + (r * r) * P;

  const int j = n & (LM_EXP_N - 1);
  const int kk = (n - j) / LM_EXP_N;
  const double T_hi = lm_exp_tab[2 * j];
  const double T_lo = lm_exp_tab[2 * j + 1];
  double a, ae, s, se;
  lm_two_prod(T_hi, p_hi, &a, &ae);
  lm_two_sum(T_hi, a, &s, &se);
  const double lo = se + ae + T_hi * p_lo + T_lo + T_lo * p_hi;
  double res = s + lo;
  if (kk > -1022 && kk < 1024) return res * lm_pow2i(kk);
  /* far ranges: split the scaling */
  if (kk >= 1024) return (res * lm_pow2i(kk - 1000)) * lm_pow2i(1000);
  return (res * lm_pow2i(kk + 1000)) * lm_pow2i(-1000);
}

LM_FN double lm_exp(double x) { return lm_exp_dd(x, 0.0); }

LM_FN double lm_log(double x) {
  if (x != x) return x;
  if (x < 0.0) return LM_NAN;
  if (x == 0.0) return -LM_INF;
  if (x == LM_INF) return x;
  double h, l;
  lm_log_dd(x, &h, &l);
  return h + l;
}

LM_FN int lm_is_int(double y) { return lm_rint(y) == y; }

LM_FN double lm_pow(double x, double y) {
  if (y == 0.0) return 1.0;
  if (x == 1.0) return 1.0;
  if (x != x || y != y) return x + y;
  if (x == 0.0 || x == LM_INF || x == -LM_INF || x < 0.0) {
    const int yi = lm_is_int(y);
    const int yodd = yi && (lm_rint(0.5 * y) != 0.5 * y);
    if (x == 0.0) {
      const double r = (y > 0.0) ? 0.0 : LM_INF;
      return (yodd && lm_as_u64(x) >> 63) ? -r : r;
    }
    if (x == LM_INF) return (y > 0.0) ? LM_INF : 0.0;
    if (x == -LM_INF) {
      const double r = (y > 0.0) ? LM_INF : 0.0;
      return yodd ? -r : r;
    }
    /* x < 0, finite: only integral y is defined */
    if (!yi) return LM_NAN;
  }
  if (y == LM_INF || y == -LM_INF) {
    const double ax = x < 0.0 ? -x : x;
    if (ax == 1.0) return 1.0;
    return ((ax > 1.0) == (y > 0.0)) ? LM_INF : 0.0;
  }
  const double sgn = (x < 0.0 && lm_rint(0.5 * y) != 0.5 * y) ? -1.0 : 1.0;
  const double ax = x < 0.0 ? -x : x;
  double lh, ll;
  lm_log_dd(ax, &lh, &ll);
  double ph, pe;
  lm_two_prod(y, lh, &ph, &pe);
  if (ph > 800.0) return sgn * LM_INF;
  if (ph < -800.0) return sgn * 0.0;
  const double pl = pe + y * ll;
  double zh, zl;
  lm_two_sum(ph, pl, &zh, &zl);
  return sgn * lm_exp_dd(zh, zl);
}

/*
 * x^y for y = RN(1/3) = 1/3 - 2^-54/3, the exponent of the radius from the mass
 * (microphysics.py:83-100), without the log/exp pair of lm_pow:
 *   c ~ x^(1/3): reciprocal cube root y ~ x^(-1/3) from the exponent bits
 *     (Kahan's bit trick, a few percent), four Newton steps
 *     y <- y (4 - x y^3) / 3 (quadratic, no division), c = x y^2;
 *   c_lo: the residual x - c^3, computed exactly with two TwoProds, over 3c^2;
 *   x^y = (c + c_lo) x^(-2^-54/3) = (c + c_lo)(1 - eps), eps = ln(x) 2^-54/3,
 *     where ln x = e ln2 + ln(1+t) only needs ~1e-7 absolute accuracy
 *     (a degree-7 polynomial, own least-squares fit on [0, 1)).
 * The value before the final rounding is within ~1e-7 ulp of x^y, so the
 * result is the correctly rounded power except within that distance of a
 * rounding boundary.  Valid for 2^-900 <= x < 2^900; everything else
 * (x <= 0, tiny or huge x, inf, NaN) takes lm_pow.
 */
#define LM_CBRT_MAGIC 0x553EF0FF289DD796ull
#define LM_L1P_1 0x1.fffd2f11ff7f5p-1
#define LM_L1P_2 -0x1.ff68cfba18ccap-2
#define LM_L1P_3 0x1.4fceb28c71a45p-2
#define LM_L1P_4 -0x1.cccfff50f2358p-3
#define LM_L1P_5 0x1.1119b2a0def2cp-3
#define LM_L1P_6 -0x1.bbcb388b20f22p-5
#define LM_L1P_7 0x1.5797072ef79d9p-7
#define LM_LN2 0x1.62e42fefa39efp-1

#if defined(__HIPCC__)
__device__ __noinline__ static double lm_pow_third_rare(double x) { return lm_pow(x, 1.0 / 3.0); }
#else
static double lm_pow_third_rare(double x) { return lm_pow(x, 1.0 / 3.0); }
#endif

LM_FN double lm_pow_third(double x) {
  const uint64_t u = lm_as_u64(x);
  if (u - 0x07B0000000000000ull >= 0x7080000000000000ull) return lm_pow_third_rare(x);
  const double third = 1.0 / 3.0;
  double y = lm_as_f64(LM_CBRT_MAGIC - u / 3u);
  for (int it = 0; it < 4; ++it) {
    double t = x * ((y * y) * y);
    t = 4.0 - t;
    y = y * t;
    y = y * third;
  }
  const double y2 = y * y;
  const double c = x * y2;
  double p, pe, q, qe;
  lm_two_prod(c, c, &p, &pe);
  lm_two_prod(p, c, &q, &qe);
  const double r = (x - q) - (qe + pe * c);
  const double c_lo = (r * y2) * third;
  const int e = (int)(u >> 52) - 1023;
  const double t = lm_as_f64((u & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull) - 1.0;
  double P = LM_L1P_7;
  P = P * t + LM_L1P_6;
  P = P * t + LM_L1P_5;
  P = P * t + LM_L1P_4;
  P = P * t + LM_L1P_3;
  P = P * t + LM_L1P_2;
  P = P * t + LM_L1P_1;
  P = P * t;
  const double L = (double)e * LM_LN2 + P;
  const double eps = L * (0x1p-54 / 3.0);
  return c + (c_lo - c * eps);
}

/*
 * Quotients that share a denominator.  gfx950 divides a / b (IEEE, correctly rounded) by the sequence
 * v_div_scale (a and b), v_rcp_f64, two Newton steps on the reciprocal (y), q = a y, r = fma(-b, q, a),
 * v_div_fmas (fma(r, y, q)) and v_div_fixup.  With |a| and |b| in [2^-300, 2^300] (lm_div_range) the scale steps
 * leave a and b alone and the fix-up step leaves the normal quotient alone, so lm_div_nr(a, b, lm_rcp_nr(b))
 * is that sequence without them: a / b bit for bit, with one reciprocal for every numerator of b.  The host
 * build (the oracle) never takes this path and divides.
 */
LM_FN int lm_div_range(double x) {
  const double ax = __builtin_fabs(x);
  return ax >= 0x1p-300 && ax <= 0x1p300;
}
#if defined(__HIPCC__)
LM_FN double lm_rcp_nr(double b) {
  const double y0 = __builtin_amdgcn_rcp(b);
  const double y1 = LM_FMA(y0, LM_FMA(-b, y0, 1.0), y0);
  return LM_FMA(y1, LM_FMA(-b, y1, 1.0), y1);
}
LM_FN double lm_div_nr(double a, double b, double y) {
  const double q = a * y;
  return LM_FMA(LM_FMA(-b, q, a), y, q);
}
/* a / b, given y = lm_rcp_nr(b) for a b in lm_div_range */
LM_FN double lm_div_y(double a, double b, double y) { return lm_div_range(a) ? lm_div_nr(a, b, y) : a / b; }
#endif

/* Python / Numba float floor-modulo a % b (fmod, then shift into the sign of b). */
LM_FN double lm_pymod(double a, double b) {
  if (a > 0.0 && a < b) return a; /* fmod(a, b) = a, already of b's sign */
  double m = LM_FMOD(a, b);
  if (m != 0.0) {
    if ((b < 0.0) != (m < 0.0)) m += b;
  } else {
    m = __builtin_copysign(0.0, b);
  }
  return m;
}

/* np.maximum / np.minimum on scalars: NaN-propagating. */
LM_FN double lm_np_max(double a, double b) {
  if (a != a) return a;
  if (b != b) return b;
  return a >= b ? a : b;
}

LM_FN double lm_np_min(double a, double b) {
  if (a != a) return a;
  if (b != b) return b;
  return a <= b ? a : b;
}

/* np.fmin: NaN-ignoring. */
LM_FN double lm_np_fmin(double a, double b) {
  if (a != a) return b;
  if (b != b) return a;
  return a <= b ? a : b;
}

LM_FN double lm_np_sign(double a) {
  if (a != a) return a;
  return a > 0.0 ? 1.0 : (a < 0.0 ? -1.0 : 0.0);
}

#endif /* LM_MATH_H */
