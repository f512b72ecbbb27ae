/*
 * lm_physics.h -- gfx950 device physics of the SIP step (FP64).
 *
 * Each function evaluates the reference's expression in the reference's
 * operation order (Python evaluates a*b/c*d left to right; no contraction),
 * so the results are bit-identical to the reference's arithmetic given the
 * same transcendentals (lm_math.h).  Per-cell quantities that the reference
 * recomputes for every SIP (l_alpha, l_beta, c1, c2, rho_w(T), sigma_w(T))
 * are evaluated once per cell into the per-cell table by the grid kernel: same
 * expression, same inputs, same value.
 */
#ifndef LM_PHYSICS_H
#define LM_PHYSICS_H

#include "lm_math.h"

/* constants.py, atmosphere.py, material_properties.py, microphysics.py */
struct LmPhys {
  double pi, one_third, four_pi_over_three_inv;
  double R_v, R_d, c_pd, eps_gc, kappa_dry, kappa_factor, kappa_factor2, heat_factor_r_v;
  double c_alpha_1, c_alpha_2, c_beta_1, c_beta_2, T_alpha_0;
  double w_s_max_AS_inv, par_sigma_AS, par_sigma_NaCl;
  double par_rho_AS[4], par_rho_deriv_AS[3];
  double par_wat_act_AS[5], par_wat_act_deriv_AS[4];
  double par_wat_act_NaCl[5], par_wat_act_deriv_NaCl[4];
  double par_sol_dens_NaCl[6];
  double par_water_dens[3];
};

/* Per-cell constants read by the particle kernel (16 doubles, 128 B). */
enum {
  CC_T = 0, CC_P, CC_S, CC_ES, CC_LV, CC_K, CC_DV, CC_MU, CC_RHOF,
  CC_LALPHA, CC_LBETA, CC_C1, CC_C2, CC_RHOW, CC_SIGW,
  CC_TP, /* cell temperature before the last grid substep = T_p of its SIPs */
  CC_N = 16
};

#if defined(__HIPCC__)
#define LMD __device__ static inline
#else
#define LMD static inline
#endif
/* templates over the value type: double for the step, LmDual<N> (lm_dual.h) for its local Jacobians */
#define LMDT template <class R> LMD

/*
 * The condensation functions below take each argument's type on its own, so a caller can mix doubles
 * (no derivative) with dual numbers; LmCommonT<...> is the type of an expression over the arguments
 * (double if all are double; lm_dual.h adds the dual cases).  The value of every operation is the same
 * double operation in the same order whatever the types.
 */
template <class A, class B>
struct LmCommon2;
template <>
struct LmCommon2<double, double> {
  using type = double;
};
template <class... T>
struct LmCommon;
template <class A>
struct LmCommon<A> {
  using type = A;
};
template <class A, class B, class... T>
struct LmCommon<A, B, T...> {
  using type = typename LmCommon<typename LmCommon2<A, B>::type, T...>::type;
};
template <class... T>
using LmCommonT = typename LmCommon<T...>::type;

LMDT R lmp_poly(const double *par, int n, R x) { /* algebra.py:11-34 */
  R res = par[0] * x + par[1];
  for (int k = 2; k < n; ++k) res = res * x + par[k];
  return res;
}

/* material_properties.py:31-50 */
LMDT R lmp_density_water(const LmPhys &P, R T) {
  const R d = T - P.par_water_dens[2];
  return P.par_water_dens[0] + P.par_water_dens[1] * (d * d);
}
/* material_properties.py:60-84 */
template <class W, class TT>
LMD LmCommonT<W, TT> lmp_density_NaCl(const LmPhys &P, W w, TT T) {
  const double *p = P.par_sol_dens_NaCl;
  return p[0] + p[1] * w + p[2] * T + p[3] * w * w + p[4] * w * T + p[5] * T * T;
}
/* material_properties.py:224-244 */
LMDT R lmp_sigma_water(R T) {
  const R tau = 1 - T / 647.096;
  return 0.2358 * lm_pow(tau, 1.256) * (1 - 0.625 * tau);
}
/* q1 = a1 / b and q2 = a2 / b */
template <class A1, class A2, class B>
LMD void lmp_div2(A1 a1, A2 a2, B b, LmCommonT<A1, B> &q1, LmCommonT<A2, B> &q2) {
  q1 = a1 / b;
  q2 = a2 / b;
}
#if defined(__HIPCC__)
/* doubles on the device: one reciprocal for both quotients where lm_div_nr is exact (the same values) */
LMD void lmp_div2(double a1, double a2, double b, double &q1, double &q2) {
  if (lm_div_range(a1) & lm_div_range(a2) & lm_div_range(b)) {
    const double y = lm_rcp_nr(b);
    q1 = lm_div_nr(a1, b, y);
    q2 = lm_div_nr(a2, b, y);
  } else {
    q1 = a1 / b;
    q2 = a2 / b;
  }
}
#endif

/* microphysics.py:83-100 (the exponent is one_third = RN(1/3), lm_pow_third's) */
template <class M, class RH>
LMD LmCommonT<M, RH> lmp_radius(const LmPhys &P, M m, RH rho) {
  return lm_pow_third(P.four_pi_over_three_inv * m / rho);
}
/* microphysics.py:989-1006 */
LMDT R lmp_l_alpha(const LmPhys &P, R T, R p, R K) {
  return (P.c_alpha_1 + P.c_alpha_2 * (T - P.T_alpha_0)) * K / p;
}
/* microphysics.py:1026-1051 */
LMDT R lmp_l_beta(const LmPhys &P, R T, R D_v) {
  return P.c_beta_1 * (1.0 + P.c_beta_2 * (T - P.T_alpha_0)) * D_v;
}

/* Solution density at the cell temperature: AS (material_properties.py:89-115) or NaCl. */
template <class W, class TT, class RW>
LMD LmCommonT<W, TT, RW> lmp_density_solution(const LmPhys &P, int solute, W w, TT T, RW rho_w_T) {
  using RT = LmCommonT<W, TT, RW>;
  return solute == 0 ? RT(rho_w_T * lmp_poly(P.par_rho_AS, 4, w)) : RT(lmp_density_NaCl(P, w, T));
}

/* microphysics.py:179-234 */
template <class MW, class MS, class TT, class RW, class RR, class WS, class RH>
LMD void lmp_R_ws_rho(const LmPhys &P, int solute, MW m_w, MS m_s, TT T, RW rho_w_T, RR &R, WS &w_s, RH &rho) {
  const LmCommonT<MW, MS> m_p = m_w + m_s;
  w_s = m_s / m_p;
  rho = lmp_density_solution(P, solute, w_s, T, rho_w_T);
  R = lmp_radius(P, m_p, rho);
}

/* Surface tension of the solution: material_properties.py:274-292 (NaCl), 346-366 (AS). */
template <class W, class SW>
LMD LmCommonT<W, SW> lmp_sigma_solution(const LmPhys &P, int solute, W w, SW sigma_w_T) {
  return sigma_w_T * (1.0 + (solute == 0 ? P.par_sigma_AS : P.par_sigma_NaCl) * w / (1. - w));
}

/*
 * Condensation mass rate gamma and d gamma / d m_w (Fukuta-Walter):
 * microphysics.py:1359-1469 (AS), 1237-1353 (NaCl), given m_p_inv_SI = 1e18 / (m_w + m_s) and the first
 * quotient of drho_dm_over_rho, q_rho = (-rho_w(T) m_p_inv_SI) / rho_p (AS) or (-w_s m_p_inv_SI) / rho_p (NaCl).
 */
template <class MW, class MS, class WS, class RP, class TP, class RH, class SA, class CT, class SG, class MI, class QR,
          class G>
LMD void lmp_mass_rate_q(const LmPhys &P, int solute, MW m_w, MS m_s, WS w_s, RP R_p, TP T_p, RH rho_p, SA S_amb,
                         const CT *cc, SG sigma_p, MI m_p_inv_SI, QR q_rho, G &gamma, G &dgamma) {
  const LmCommonT<RP, CT> l_alpha_plus_R_p = 1.0E-6 * (R_p + cc[CC_LALPHA]);
  const LmCommonT<RP, CT> l_beta_plus_R_p = 1.0E-6 * (R_p + cc[CC_LBETA]);
  LmCommonT<CT, MW, MS, RH, WS, TP> drho_dm_over_rho;
  if (solute == 0) {
    drho_dm_over_rho = q_rho * w_s * lmp_poly(P.par_rho_deriv_AS, 3, w_s);
  } else {
    const double *p = P.par_sol_dens_NaCl;
    drho_dm_over_rho = q_rho * (p[1] + 2.0 * p[3] * w_s + p[4] * T_p);
  }
  const auto dR_p_dm_over_R_p = P.one_third * (m_p_inv_SI - drho_dm_over_rho);
  const auto dR_p_dm = 1.0E-6 * dR_p_dm_over_R_p * R_p;
  const auto eps_k = 2.0E6 * sigma_p / (P.R_v * T_p * rho_p * R_p);
  const auto kelvin_term = lm_exp(eps_k);
  const double *aw = solute == 0 ? P.par_wat_act_AS : P.par_wat_act_NaCl;
  const double *daw = solute == 0 ? P.par_wat_act_deriv_AS : P.par_wat_act_deriv_NaCl;
  const double psig = solute == 0 ? P.par_sigma_AS : P.par_sigma_NaCl;
  const auto a_w = lmp_poly(aw, 5, w_s);
  const auto da_w_dm = -m_p_inv_SI * w_s * lmp_poly(daw, 4, w_s);
  const auto dsigma_dm = -psig * cc[CC_SIGW] * m_p_inv_SI * (w_s / ((1. - w_s) * (1. - w_s)));
  const auto S_eq = a_w * kelvin_term;
  const auto dSeq_dm =
      da_w_dm * kelvin_term + S_eq * eps_k * (dsigma_dm / sigma_p - drho_dm_over_rho - dR_p_dm_over_R_p);
  const CT c1 = cc[CC_C1], c2 = cc[CC_C2];
  const auto f3 = 1.0 / ((l_alpha_plus_R_p)*S_eq * c1 + (l_beta_plus_R_p)*c2);
  const auto f1f3 = 4.0 * P.pi * R_p * R_p * f3;
  const auto dg1_dm = (dSeq_dm * (l_alpha_plus_R_p) + S_eq * dR_p_dm) * c1 + dR_p_dm * c2;
  const auto S_sup = S_amb - S_eq; /* the reference reuses the name S_eq for S - S_eq */
  gamma = 1.0E6 * f1f3 * S_sup;
  dgamma = 1.0E-12 * f1f3 * (S_sup * (2.0 * dR_p_dm_over_R_p - f3 * dg1_dm) - dSeq_dm);
}

template <class MW, class MS, class WS, class RP, class TP, class RH, class SA, class CT, class SG, class G>
LMD void lmp_mass_rate(const LmPhys &P, int solute, MW m_w, MS m_s, WS w_s, RP R_p, TP T_p, RH rho_p, SA S_amb,
                       const CT *cc, SG sigma_p, G &gamma, G &dgamma) {
  const LmCommonT<MW, MS> m_p_inv_SI = 1.0E18 / (m_w + m_s);
  using QR = LmCommonT<CT, MW, MS, RH, WS>;
  const QR q_rho = solute == 0 ? QR(-cc[CC_RHOW] * m_p_inv_SI / rho_p) : QR(-w_s * m_p_inv_SI / rho_p);
  lmp_mass_rate_q(P, solute, m_w, m_s, w_s, R_p, T_p, rho_p, S_amb, cc, sigma_p, m_p_inv_SI, q_rho, gamma, dgamma);
}

/*
 * Implicit Newton mass update, fixed no_iter iterations with the clamps of
 * integration.py:728-836 (the inner iterations use the AS density, radius
 * and surface tension for both solutes, as the reference does).
 * Returns dm = mass_new - m_w.
 */
/* the update after the first rate evaluation (gamma, dgamma) and the later iterations; returns dm */
template <class R, class MW, class MS, class TP, class CT>
LMD R lmp_newton_iterate(const LmPhys &P, int solute, double dt_sub, int no_iter, MW m_w, MS m_s, TP T_p,
                         const CT *cc, R gamma, R dgamma) {
  const R m_w_effl = m_s * (P.w_s_max_AS_inv - 1.0);
  const CT S_amb = cc[CC_S];
  R d = dt_sub * dgamma;
  R denom_inv = d < 0.9 ? R(1.0 / (1.0 - d)) : R(10.0);
  R mass_new = lm_np_max(m_w_effl, R(m_w + dt_sub * gamma * denom_inv));
  for (int it = 0; it < no_iter - 1; ++it) {
    const R m_p = mass_new + m_s;
    R ws, m_p_inv_SI; /* m_s / m_p and the rate's 1e18 / m_p */
    lmp_div2(m_s, 1.0E18, m_p, ws, m_p_inv_SI);
    const R rho = cc[CC_RHOW] * lmp_poly(P.par_rho_AS, 4, ws);
    R r_arg, q_rho; /* the radius's (m_p / (4 pi / 3)) / rho and the rate's drho quotient */
    lmp_div2(P.four_pi_over_three_inv * m_p, solute == 0 ? R(-cc[CC_RHOW] * m_p_inv_SI) : R(-ws * m_p_inv_SI), rho,
             r_arg, q_rho);
    const R Rr = lm_pow_third(r_arg);
    const R sigma = cc[CC_SIGW] * (1.0 + P.par_sigma_AS * ws / (1. - ws));
    lmp_mass_rate_q(P, solute, mass_new, m_s, ws, Rr, T_p, rho, S_amb, cc, sigma, m_p_inv_SI, q_rho, gamma, dgamma);
    d = dt_sub * dgamma;
    denom_inv = d < 0.9 ? R(1.0 / (1.0 - d)) : R(10.0);
    mass_new += (dt_sub * gamma + m_w - mass_new) * denom_inv;
    mass_new = lm_np_max(m_w_effl, mass_new);
  }
  return mass_new - m_w;
}

template <class MW, class MS, class WS, class RP, class TP, class RH, class CT, class SG>
LMD LmCommonT<MW, MS, WS, RP, TP, RH, CT, SG> lmp_newton_dm(const LmPhys &P, int solute, double dt_sub, int no_iter,
                                                            MW m_w, MS m_s, WS w_s, RP R_p, TP T_p, RH rho_p,
                                                            const CT *cc, SG sigma_p) {
  using R = LmCommonT<MW, MS, WS, RP, TP, RH, CT, SG>;
  R gamma, dgamma;
  lmp_mass_rate(P, solute, m_w, m_s, w_s, R_p, T_p, rho_p, cc[CC_S], cc, sigma_p, gamma, dgamma);
  return lmp_newton_iterate<R>(P, solute, dt_sub, no_iter, m_w, m_s, T_p, cc, gamma, dgamma);
}

/*
 * dm of the condensation substep of one SIP from its masses at the cell's temperature T_p: lmp_R_ws_rho,
 * lmp_sigma_solution and lmp_newton_dm in one, the quotients of one denominator taken together (lmp_div2).
 * The same values.
 */
LMDT R lmp_condensation_dm(const LmPhys &P, int solute, double dt_sub, int no_iter, R m_w, R m_s, R T_p,
                           const R *cc) {
  const R m_p = m_w + m_s;
  R w_s, m_p_inv_SI;
  lmp_div2(m_s, 1.0E18, m_p, w_s, m_p_inv_SI);
  const R rho = lmp_density_solution(P, solute, w_s, T_p, cc[CC_RHOW]);
  R r_arg, q_rho;
  lmp_div2(P.four_pi_over_three_inv * m_p, solute == 0 ? R(-cc[CC_RHOW] * m_p_inv_SI) : R(-w_s * m_p_inv_SI), rho,
           r_arg, q_rho);
  const R R_p = lm_pow_third(r_arg);
  const R sigma_p = lmp_sigma_solution(P, solute, w_s, cc[CC_SIGW]);
  R gamma, dgamma;
  lmp_mass_rate_q(P, solute, m_w, m_s, w_s, R_p, T_p, rho, cc[CC_S], cc, sigma_p, m_p_inv_SI, q_rho, gamma, dgamma);
  return lmp_newton_iterate<R>(P, solute, dt_sub, no_iter, m_w, m_s, T_p, cc, gamma, dgamma);
}

/* grid.py:225-254 with Python negative indexing of the first axis */
LMD double lmp_bilinear(int i, int j, double a, double b, const double *f, int nfx, int nfz) {
  const int i0 = i < 0 ? i + nfx : i, i1 = i + 1;
  const int j0 = j < 0 ? j + nfz : j, j1 = j + 1;
  return a * (b * f[i1 * nfz + j1] + (1 - b) * f[i1 * nfz + j0]) +
         (1 - a) * (b * f[i0 * nfz + j1] + (1 - b) * f[i0 * nfz + j0]);
}

/* grid.py:257-323 (one SIP, j >= 0), including the i-1 = -1 -> column nx wrap */
LMD void lmp_interp_vel(int i, int j, double wx, double wy, const double *U, const double *V, int nx, int nz,
                        double &u, double &v) {
  const int nfx = nx + 1, nfz = nz + 1;
  if ((j == 0 && wy <= 0.5) || (j == nz - 1 && wy >= 0.5))
    u = wx * U[(i + 1) * nfz + j] + (1 - wx) * U[i * nfz + j];
  else if (wy > 0.5)
    u = lmp_bilinear(i, j, wx, wy - 0.5, U, nfx, nfz);
  else
    u = lmp_bilinear(i, j - 1, wx, wy + 0.5, U, nfx, nfz);
  if (wx > 0.5)
    v = lmp_bilinear(i, j, wx - 0.5, wy, V, nfx, nfz);
  else
    v = lmp_bilinear(i - 1, j, wx + 0.5, wy, V, nfx, nfz);
}

/* all_or_nothing.py:35-65 */
LMD int lmp_kernel_index(double R, double lnR0, double lnf, int nb) {
  long long ind = (long long)((lm_log(R) - lnR0) / lnf + 0.5);
  if (ind < 0)
    ind = 0;
  else if (ind > nb - 1)
    ind = nb - 1;
  return (int)ind;
}

/* collision/kernel.py:389-409 */
LMD double lmp_kernel_hydro(const LmPhys &P, double R1, double R2, double E, double dv) {
  return P.pi * (R1 + R2) * (R1 + R2) * E * dv * 1.0E-12;
}

#endif /* LM_PHYSICS_H */
