/*
 * lm_dual.h -- forward-mode dual numbers for the local Jacobians of the adjoint.
 *
 * LmDual<N> carries a value and N directional derivatives.  The value part
 * of every operation is the plain double operation on the values, in the
 * same order as the double expression it replaces, so a template of the
 * step's physics instantiated on LmDual<N> reproduces the double results bit
 * for bit in .v while the .d parts carry the exact derivatives of the
 * computed map (branches and clamps taken as the values decide).  The
 * derivative parts use fused multiply-adds and a reciprocal; they are not
 * part of the reproduced state.
 */
#ifndef LM_DUAL_H
#define LM_DUAL_H

#include "lm_math.h"

#if defined(__HIPCC__)
#define LM_FN_CTOR __host__ __device__ inline
#else
#define LM_FN_CTOR inline
#endif

template <int N>
struct LmDual {
  double v;
  double d[N];
  LM_FN_CTOR LmDual() : v(0.0) {
#pragma unroll
    for (int k = 0; k < N; ++k) d[k] = 0.0;
  }
  LM_FN_CTOR LmDual(double x) : v(x) {
#pragma unroll
    for (int k = 0; k < N; ++k) d[k] = 0.0;
  }
  /* value x, derivative 1 in direction k */
  static LM_FN_CTOR LmDual seed(double x, int k) {
    LmDual r(x);
    r.d[k] = 1.0;
    return r;
  }
};

#define LMDUAL template <int N> LM_FN

/*
 * 1/x for the derivative parts only (never for a value): on the GPU the hardware reciprocal estimate
 * refined by one Newton step (three instructions instead of a correctly rounded division, error a few
 * units in the last place), on the host the division.
 */
LM_FN double lm_drcp(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  const double r = __builtin_amdgcn_rcp(x);
  return __builtin_fma(__builtin_fma(-x, r, 1.0), r, r);
#else
  return 1.0 / x;
#endif
}

/* expressions mixing doubles and LmDual<N> are LmDual<N> (LmCommonT of lm_physics.h) */
template <class A, class B>
struct LmCommon2;
template <int N>
struct LmCommon2<LmDual<N>, double> {
  using type = LmDual<N>;
};
template <int N>
struct LmCommon2<double, LmDual<N>> {
  using type = LmDual<N>;
};
template <int N>
struct LmCommon2<LmDual<N>, LmDual<N>> {
  using type = LmDual<N>;
};

LMDUAL LmDual<N> operator+(const LmDual<N> &a, const LmDual<N> &b) {
  LmDual<N> r;
  r.v = a.v + b.v;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = a.d[k] + b.d[k];
  return r;
}
LMDUAL LmDual<N> operator+(const LmDual<N> &a, double b) {
  LmDual<N> r = a;
  r.v = a.v + b;
  return r;
}
LMDUAL LmDual<N> operator+(double a, const LmDual<N> &b) {
  LmDual<N> r = b;
  r.v = a + b.v;
  return r;
}
LMDUAL LmDual<N> operator-(const LmDual<N> &a, const LmDual<N> &b) {
  LmDual<N> r;
  r.v = a.v - b.v;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = a.d[k] - b.d[k];
  return r;
}
LMDUAL LmDual<N> operator-(const LmDual<N> &a, double b) {
  LmDual<N> r = a;
  r.v = a.v - b;
  return r;
}
LMDUAL LmDual<N> operator-(double a, const LmDual<N> &b) {
  LmDual<N> r;
  r.v = a - b.v;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = -b.d[k];
  return r;
}
LMDUAL LmDual<N> operator-(const LmDual<N> &a) {
  LmDual<N> r;
  r.v = -a.v;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = -a.d[k];
  return r;
}
LMDUAL LmDual<N> operator*(const LmDual<N> &a, const LmDual<N> &b) {
  LmDual<N> r;
  r.v = a.v * b.v;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = LM_FMA(a.v, b.d[k], b.v * a.d[k]);
  return r;
}
LMDUAL LmDual<N> operator*(const LmDual<N> &a, double b) {
  LmDual<N> r;
  r.v = a.v * b;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * b;
  return r;
}
LMDUAL LmDual<N> operator*(double a, const LmDual<N> &b) {
  LmDual<N> r;
  r.v = a * b.v;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = a * b.d[k];
  return r;
}
LMDUAL LmDual<N> operator/(const LmDual<N> &a, const LmDual<N> &b) {
  LmDual<N> r;
  r.v = a.v / b.v;
  const double rb = lm_drcp(b.v);
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = (a.d[k] - r.v * b.d[k]) * rb;
  return r;
}
LMDUAL LmDual<N> operator/(const LmDual<N> &a, double b) {
  LmDual<N> r;
  r.v = a.v / b;
  const double rb = lm_drcp(b);
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * rb;
  return r;
}
LMDUAL LmDual<N> operator/(double a, const LmDual<N> &b) {
  LmDual<N> r;
  r.v = a / b.v;
  const double s = -r.v * lm_drcp(b.v);
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = s * b.d[k];
  return r;
}
LMDUAL LmDual<N> &operator+=(LmDual<N> &a, const LmDual<N> &b) { return a = a + b; }
LMDUAL LmDual<N> &operator-=(LmDual<N> &a, const LmDual<N> &b) { return a = a - b; }
LMDUAL LmDual<N> &operator*=(LmDual<N> &a, const LmDual<N> &b) { return a = a * b; }

/* comparisons look at values only */
LMDUAL bool operator<(const LmDual<N> &a, double b) { return a.v < b; }
LMDUAL bool operator>(const LmDual<N> &a, double b) { return a.v > b; }
LMDUAL bool operator<=(const LmDual<N> &a, double b) { return a.v <= b; }
LMDUAL bool operator>=(const LmDual<N> &a, double b) { return a.v >= b; }
LMDUAL bool operator<(const LmDual<N> &a, const LmDual<N> &b) { return a.v < b.v; }
LMDUAL bool operator>(const LmDual<N> &a, const LmDual<N> &b) { return a.v > b.v; }
LMDUAL bool operator<=(const LmDual<N> &a, const LmDual<N> &b) { return a.v <= b.v; }
LMDUAL bool operator>=(const LmDual<N> &a, const LmDual<N> &b) { return a.v >= b.v; }

/* value of a double or a dual */
LM_FN double lm_val(double x) { return x; }
LMDUAL double lm_val(const LmDual<N> &x) { return x.v; }

/* chain rule helper: f(x) with value fv and derivative df at x.v */
LMDUAL LmDual<N> lm_chain(const LmDual<N> &x, double fv, double df) {
  LmDual<N> r;
  r.v = fv;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = df * x.d[k];
  return r;
}

LMDUAL LmDual<N> lm_exp(const LmDual<N> &x) {
  const double e = lm_exp(x.v);
  return lm_chain(x, e, e);
}
LMDUAL LmDual<N> lm_log(const LmDual<N> &x) { return lm_chain(x, lm_log(x.v), lm_drcp(x.v)); }
/* x^y with a constant exponent */
LMDUAL LmDual<N> lm_pow(const LmDual<N> &x, double y) {
  const double p = lm_pow(x.v, y);
  return lm_chain(x, p, y * (p * lm_drcp(x.v)));
}
LMDUAL LmDual<N> lm_pow_third(const LmDual<N> &x) {
  const double p = lm_pow_third(x.v);
  return lm_chain(x, p, p * lm_drcp(3.0 * x.v));
}
/* np.maximum / np.minimum / np.fmin: the selected operand, with its derivatives */
LMDUAL LmDual<N> lm_np_max(const LmDual<N> &a, const LmDual<N> &b) {
  if (a.v != a.v) return a;
  if (b.v != b.v) return b;
  return a.v >= b.v ? a : b;
}
LMDUAL LmDual<N> lm_np_min(const LmDual<N> &a, const LmDual<N> &b) {
  if (a.v != a.v) return a;
  if (b.v != b.v) return b;
  return a.v <= b.v ? a : b;
}
LMDUAL LmDual<N> lm_np_fmin(const LmDual<N> &a, const LmDual<N> &b) {
  if (a.v != a.v) return b;
  if (b.v != b.v) return a;
  return a.v <= b.v ? a : b;
}

#endif /* LM_DUAL_H */
