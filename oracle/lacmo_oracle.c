/*
 * lacmo_oracle.c -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement (FP64, plain C, single thread) of the LacmoPy super-droplet
 * time step, written operation by operation from the reference sources so
 * that tests can check the gfx950 kernels against it.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the
 * product (lacmopy_amd) never calls it.
 *
 * Arrays use the reference's own layout (SIP arrays in ID order, pos/vel/cells
 * as (2,N), grid fields (Nx,Nz) row-major, face fields (2,Nx+1,Nz+1)).
 * Transcendentals come from lm_math.h (pinned against glibc separately);
 * compile with -ffp-contract=off so the operation order is the reference's.
 *
 * Each function cites the reference file:line it restates.  Pins (see
 * tests/test_oracle_cpu.py): Golovin moments, Wang 2007 data, conservation of
 * sum xi*m and of the water budget, sequential-vs-brute-force pair counts,
 * hand-enumerated 2- and 3-SIP cells, finite-difference derivative of the
 * mass rate, textbook Maxwell growth limit, Stokes drag limit, ideal-gas
 * identity of the grid update, Hundsdorfer limiter and upwind-flux
 * identities.  Numba/NumPy semantics restated here: floored float modulo,
 * np.fmin (NaN-ignoring), np.maximum (NaN-propagating), np.sign(0)=0,
 * int() truncation, Python negative indexing of the top row for inactive SIPs.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef ORACLE_LIBM
/* glibc transcendentals (the reference's own math) instead of lm_math */
#define OR_EXP exp
#define OR_LOG log
#define OR_POW pow
#define OR_POW_THIRD(x) pow((x), 1.0 / 3.0)
#include "../lacmopy_amd/csrc/lm_math.h" /* for lm_pymod etc. only */
#else
#include "../lacmopy_amd/csrc/lm_math.h"
#define OR_EXP lm_exp
#define OR_LOG lm_log
#define OR_POW lm_pow
#define OR_POW_THIRD lm_pow_third
#endif
#include "../lacmopy_amd/csrc/lm_philox.h"

#define EXPORT __attribute__((visibility("default")))

/* ------------------------------------------------------------------------ */
/* constants.py:27-117, atmosphere.py:74-101,361-371, material_properties.py, */
/* microphysics.py:968-1051, integration.py:723-724                          */
/* ------------------------------------------------------------------------ */
typedef struct {
  double pi, one_third, four_pi_over_three, four_pi_over_three_inv;
  double R_v, R_d, c_pd, c_pv, eps_gc, kappa_dry, kappa_factor, kappa_factor2, heat_factor_r_v;
  double c_alpha_1, c_alpha_2, c_beta_1, c_beta_2, T_alpha_0;
  double w_s_max_AS_inv;
  double par_sigma_AS, par_sigma_NaCl;
  double par_rho_AS[4], par_rho_deriv_AS[3];
  double par_wat_act_AS[5], par_wat_act_deriv_AS[4];
  double par_wat_act_NaCl[5], par_wat_act_deriv_NaCl[4];
  double par_sol_dens_NaCl[6];
  double par_water_dens[3];
} OrConst;

static OrConst K;
static int K_init = 0;

static void or_init_consts(void) {
  if (K_init) return;
  const double pi = 3.141592653589793; /* math.pi */
  K.pi = pi;
  K.one_third = 1.0 / 3.0;                  /* constants.py:27 */
  K.four_pi_over_three = 4.0 * pi / 3.0;    /* constants.py:28 */
  K.four_pi_over_three_inv = 0.75 / pi;     /* constants.py:29 */
  K.R_v = 461.53;                           /* constants.py:90 */
  K.R_d = 287.06;                           /* constants.py:95 */
  K.c_pd = 1004.71;                         /* constants.py:109 */
  K.c_pv = 1906.0;                          /* constants.py:114 */
  K.eps_gc = K.R_d / K.R_v;                 /* atmosphere.py:99 */
  K.kappa_dry = K.R_d / K.c_pd;             /* atmosphere.py:74 */
  K.kappa_factor = 1.0 / (1.0 - K.kappa_dry);          /* atmosphere.py:370 */
  K.kappa_factor2 = -K.kappa_dry * K.kappa_factor;     /* atmosphere.py:371 */
  K.heat_factor_r_v = (K.c_pv - K.R_v) / (K.c_pd - K.R_d); /* atmosphere.py:365 */
  /* microphysics.py:971-1026 */
  const double acc = 1.0, adi_inv = 0.7142857142857143, T0 = 289.0;
  K.T_alpha_0 = T0;
  K.c_alpha_1 = 1.0E6 * sqrt(2.0 * pi * K.R_d * T0) / (acc * (K.c_pd * adi_inv + 0.5 * K.R_d));
  K.c_alpha_2 = 0.5E6 * sqrt(2.0 * pi * K.R_d / T0) / (acc * (K.c_pd * adi_inv + 0.5 * K.R_d));
  K.c_beta_1 = 1.0E6 * sqrt(2.0 * pi * 18.015E-3 / (8.3144598 * T0)) / 0.0415;
  K.c_beta_2 = -0.5 / T0;
  /* integration.py:723-724 */
  K.w_s_max_AS_inv = 1. / 0.78;
  /* material_properties.py:272,344 */
  K.par_sigma_NaCl = 1.62E-3 / 58.4428E-3 / 0.072;
  K.par_sigma_AS = 2.362E-3 / 132.14E-3 / 0.072;
  /* material_properties.py:87 (reversed, each / 997.1) */
  const double rho_raw[4] = {1.024E1, -5.036E1, 592., 997.1};
  for (int i = 0; i < 4; ++i) K.par_rho_AS[i] = rho_raw[i] / 997.1;
  /* microphysics.py:1355-1356: par[:-1] * arange(1, len)[::-1] */
  for (int i = 0; i < 3; ++i) K.par_rho_deriv_AS[i] = K.par_rho_AS[i] * (double)(3 - i);
  /* material_properties.py:430 */
  const double aw_as[5] = {1.412, -2.336, 3.113E-1, -2.715E-1, 1.0};
  for (int i = 0; i < 5; ++i) K.par_wat_act_AS[i] = aw_as[i];
  for (int i = 0; i < 4; ++i) K.par_wat_act_deriv_AS[i] = aw_as[i] * (double)(4 - i);
  /* material_properties.py:400 */
  const double aw_nacl[5] = {1.518E1, -1.158E1, 8.642E-1, -6.366E-1, 1.0};
  for (int i = 0; i < 5; ++i) K.par_wat_act_NaCl[i] = aw_nacl[i];
  for (int i = 0; i < 4; ++i) K.par_wat_act_deriv_NaCl[i] = aw_nacl[i] * (double)(4 - i);
  /* material_properties.py:56-58 */
  const double sd[6] = {7.619443952135e+02, 1.021264281453e+03, 1.828970151543e+00,
                        2.405352122804e+02, -1.080547892416e+00, -3.492805028749e-03};
  for (int i = 0; i < 6; ++i) K.par_sol_dens_NaCl[i] = sd[i];
  /* material_properties.py:29 */
  K.par_water_dens[0] = 1.00013502e+03;
  K.par_water_dens[1] = -4.68112708e-03;
  K.par_water_dens[2] = 2.72389977e+02;
  K_init = 1;
}

EXPORT void or_get_consts(double *out, int n) {
  or_init_consts();
  const double *p = (const double *)&K;
  const int m = (int)(sizeof(OrConst) / sizeof(double));
  for (int i = 0; i < n && i < m; ++i) out[i] = p[i];
}

EXPORT int or_num_consts(void) { return (int)(sizeof(OrConst) / sizeof(double)); }

/* algebra.py:11-34 */
static double poly(const double *par, int n, double x) {
  double res = par[0] * x + par[1];
  for (int k = 2; k < n; ++k) res = res * x + par[k];
  return res;
}

/* ------------------------------------------------------------------------ */
/* material_properties.py                                                   */
/* ------------------------------------------------------------------------ */
static double density_water(double T) { /* :31-50 */
  const double d = T - K.par_water_dens[2];
  return K.par_water_dens[0] + K.par_water_dens[1] * (d * d);
}
static double density_NaCl_solution(double w, double T) { /* :60-84 */
  const double *p = K.par_sol_dens_NaCl;
  return p[0] + p[1] * w + p[2] * T + p[3] * w * w + p[4] * w * T + p[5] * T * T;
}
static double density_AS_solution(double w, double T) { /* :89-115 */
  return density_water(T) * poly(K.par_rho_AS, 4, w);
}
static double surface_tension_water(double T) { /* :224-244 */
  const double tau = 1 - T / 647.096;
  return 0.2358 * OR_POW(tau, 1.256) * (1 - 0.625 * tau);
}
static double surface_tension_NaCl(double w, double T) { /* :274-292 */
  return surface_tension_water(T) * (1.0 + K.par_sigma_NaCl * w / (1. - w));
}
static double surface_tension_AS(double w, double T) { /* :346-366 */
  return surface_tension_water(T) * (1.0 + K.par_sigma_AS * w / (1. - w));
}
static double water_activity_AS(double w) { return poly(K.par_wat_act_AS, 5, w); }     /* :432-456 */
static double water_activity_NaCl(double w) { return poly(K.par_wat_act_NaCl, 5, w); } /* :403-427 */
static double thermal_conductivity_air(double T) { return 4.1868E-3 * (5.69 + 0.017 * (T - 273.15)); } /* :461-478 */
static double viscosity_air(double T) { return 1.458E-6 * OR_POW(T, 1.5) / (T + 110.4); }              /* :482-499 */
static double diffusion_constant(double T, double p) { return 4.01218E-5 * OR_POW(T, 1.94) / p; }     /* :502-522 */
static double heat_of_vaporization(double T) { return 1.0E3 * (2500.82 - 2.358 * (T - 273.0)); }      /* :525-543 */
static double saturation_pressure_vapor_liquid(double T) { return 2.53E11 * OR_EXP(-5420.0 / T); }    /* :546-563 */

/* ------------------------------------------------------------------------ */
/* microphysics.py                                                          */
/* ------------------------------------------------------------------------ */
static double radius_from_mass(double m, double rho) { /* :83-100 */
  return OR_POW_THIRD(K.four_pi_over_three_inv * m / rho);
}
/* :179-234, dispatch :237-266.  solute 0 = AS, 1 = NaCl */
static void R_p_w_s_rho_p(double m_w, double m_s, double T, int solute, double *R, double *w_s, double *rho) {
  const double m_p = m_w + m_s;
  const double w = m_s / m_p;
  const double r = solute == 0 ? density_AS_solution(w, T) : density_NaCl_solution(w, T);
  *R = radius_from_mass(m_p, r);
  *w_s = w;
  *rho = r;
}
static double particle_reynolds_number(double R, double dv, double rho_f, double mu) { /* :297-324 */
  return 2.0E-6 * rho_f * R * dv / mu;
}
static double kelvin_argument(double R, double T, double rho, double sigma) { /* :329-362 */
  return 2.0E6 * sigma / (K.R_v * T * rho * R);
}
static double l_alpha_lin(double T, double p, double Kc) { /* :989-1006 */
  return (K.c_alpha_1 + K.c_alpha_2 * (T - K.T_alpha_0)) * Kc / p;
}
static double l_beta_lin(double T, double D_v) { /* :1026-1051 */
  return K.c_beta_1 * (1.0 + K.c_beta_2 * (T - K.T_alpha_0)) * D_v;
}

/* :1359-1469 (AS) and :1237-1353 (NaCl): gamma (fg/s) and d gamma / d m_w (1/s) */
EXPORT void or_mass_rate_and_derivative(int solute, double m_w, double m_s, double w_s, double R_p, double T_p,
                                        double rho_p, double T_amb, double p_amb, double S_amb, double e_s_amb,
                                        double L_v, double Kc, double D_v, double sigma_p, double *gamma,
                                        double *dgamma) {
  or_init_consts();
  const double l_alpha_plus_R_p = 1.0E-6 * (R_p + l_alpha_lin(T_amb, p_amb, Kc));
  const double l_beta_plus_R_p = 1.0E-6 * (R_p + l_beta_lin(T_amb, D_v));
  const double m_p_inv_SI = 1.0E18 / (m_w + m_s);
  double drho_dm_over_rho, a_w, da_w_dm, dsigma_dm;
  if (solute == 0) {
    drho_dm_over_rho = -density_water(T_p) * m_p_inv_SI / rho_p * w_s * poly(K.par_rho_deriv_AS, 3, w_s);
  } else {
    const double *p = K.par_sol_dens_NaCl;
    drho_dm_over_rho = -w_s * m_p_inv_SI / rho_p * (p[1] + 2.0 * p[3] * w_s + p[4] * T_p);
  }
  const double dR_p_dm_over_R_p = K.one_third * (m_p_inv_SI - drho_dm_over_rho);
  const double dR_p_dm = 1.0E-6 * dR_p_dm_over_R_p * R_p;
  const double eps_k = kelvin_argument(R_p, T_p, rho_p, sigma_p);
  const double kelvin_term = OR_EXP(eps_k);
  if (solute == 0) {
    a_w = water_activity_AS(w_s);
    da_w_dm = -m_p_inv_SI * w_s * poly(K.par_wat_act_deriv_AS, 4, w_s);
    dsigma_dm = -K.par_sigma_AS * surface_tension_water(T_p) * m_p_inv_SI * (w_s / ((1. - w_s) * (1. - w_s)));
  } else {
    a_w = water_activity_NaCl(w_s);
    da_w_dm = -m_p_inv_SI * w_s * poly(K.par_wat_act_deriv_NaCl, 4, w_s);
    dsigma_dm = -K.par_sigma_NaCl * surface_tension_water(T_p) * m_p_inv_SI * (w_s / ((1. - w_s) * (1. - w_s)));
  }
  double S_eq = a_w * kelvin_term;
  const double dSeq_dm =
      da_w_dm * kelvin_term + S_eq * eps_k * (dsigma_dm / sigma_p - drho_dm_over_rho - dR_p_dm_over_R_p);
  const double c1 = L_v * L_v / (K.R_v * Kc * T_amb * T_amb);
  const double c2 = K.R_v * T_amb / (D_v * e_s_amb);
  const double f3 = 1.0 / ((l_alpha_plus_R_p)*S_eq * c1 + (l_beta_plus_R_p)*c2);
  const double f1f3 = 4.0 * K.pi * R_p * R_p * f3;
  const double dg1_dm = (dSeq_dm * (l_alpha_plus_R_p) + S_eq * dR_p_dm) * c1 + dR_p_dm * c2;
  S_eq = S_amb - S_eq;
  *gamma = 1.0E6 * f1f3 * S_eq;
  *dgamma = 1.0E-12 * f1f3 * (S_eq * (2.0 * dR_p_dm_over_R_p - f3 * dg1_dm) - dSeq_dm);
}

/* integration.py:728-836 */
EXPORT double or_dml_impl_Newton(int solute, double dt_sub, int no_iter, double m_w, double m_s, double w_s,
                                 double R_p, double T_p, double rho_p, double T_amb, double p_amb, double S_amb,
                                 double e_s_amb, double L_v, double Kc, double D_v, double sigma_p, double *gamma0_out) {
  or_init_consts();
  const double m_w_effl = m_s * (K.w_s_max_AS_inv - 1.0);
  double gamma0, dgamma;
  or_mass_rate_and_derivative(solute, m_w, m_s, w_s, R_p, T_p, rho_p, T_amb, p_amb, S_amb, e_s_amb, L_v, Kc, D_v,
                              sigma_p, &gamma0, &dgamma);
  double d = dt_sub * dgamma;
  double denom_inv = d < 0.9 ? 1.0 / (1.0 - d) : 10.0;
  double mass_new = lm_np_max(m_w_effl, m_w + dt_sub * gamma0 * denom_inv);
  for (int cnt = 0; cnt < no_iter - 1; ++cnt) {
    const double m_p = mass_new + m_s;
    const double ws = m_s / m_p;
    const double rho = density_AS_solution(ws, T_p);
    const double R = radius_from_mass(m_p, rho);
    const double sigma = surface_tension_AS(ws, T_p);
    double gamma;
    or_mass_rate_and_derivative(solute, mass_new, m_s, ws, R, T_p, rho, T_amb, p_amb, S_amb, e_s_amb, L_v, Kc, D_v,
                                sigma, &gamma, &dgamma);
    d = dt_sub * dgamma;
    denom_inv = d < 0.9 ? 1.0 / (1.0 - d) : 10.0;
    mass_new += (dt_sub * gamma + m_w - mass_new) * denom_inv;
    mass_new = lm_np_max(m_w_effl, mass_new);
  }
  if (gamma0_out) *gamma0_out = gamma0;
  return mass_new - m_w;
}

/* ------------------------------------------------------------------------ */
/* collision kernels: collision/kernel.py:389-409, 461-480, 107-210          */
/* ------------------------------------------------------------------------ */
static double kernel_hydro(double R1, double R2, double E, double dv) {
  return K.pi * (R1 + R2) * (R1 + R2) * E * dv * 1.0E-12;
}
static double kernel_golovin(double mi, double mj) { return (mi + mj) * 1.5E-18; }

static const double p1_Beard[7] = {-0.327815e-5, 0.855176e-4, -0.578878e-3, -0.987059e-3,
                                   -0.153193e-2, 0.992696,    -0.318657e1};
static const double p2_Beard[6] = {0.238449e-2, -0.542819e-1, 0.475294, -0.204914e1, 0.523778e1, -0.500015e1};

EXPORT double or_terminal_velocity_Beard(double R) { /* kernel.py:107-167 */
  or_init_consts();
  const double rho_w = 1.0E3, rho_a = 1.225, visc = 1.818E-5, sigma_w = 73.0E-3;
  const double R_0 = 10.0, R_1 = 535.0, R_max = 3500.0;
  const double drho = rho_w - rho_a;
  const double g = 9.80665;
  double v;
  if (R < R_0) {
    const double l0 = 6.62E-2;
    const double C1 = drho * g / (4.5 * visc);
    const double C_sc = 1.0 + 1.257 * l0 / R;
    v = C1 * C_sc * R * R * 1.0E-12;
  } else if (R < R_1) {
    const double N_Da = 32.0E-18 * R * R * R * rho_a * drho * g / (3.0 * visc * visc);
    double Y = OR_LOG(N_Da);
    Y = poly(p1_Beard, 7, Y);
    const double l0 = 6.62E-2;
    const double C_sc = 1.0 + 1.257 * l0 / R;
    v = visc * C_sc * OR_EXP(Y) / (rho_a * R * 2.0E-6);
  } else if (R < R_max) {
    const double N_Bo = 16.0E-12 * R * R * drho * g / (3.0 * sigma_w);
    const double visc2 = visc * visc; /* viscosity**4: Numba integer power, by squaring */
    const double N_P16 = OR_POW(sigma_w * sigma_w * sigma_w * rho_a * rho_a / ((visc2 * visc2) * drho * g),
                                1.0 / 6.0);
    double Y = OR_LOG(N_Bo * N_P16);
    Y = poly(p2_Beard, 6, Y);
    v = visc * N_P16 * OR_EXP(Y) / (rho_a * R * 2.0E-6);
  } else {
    v = 9.04929248;
  }
  return v;
}

/* all_or_nothing.py:35-65 */
EXPORT int64_t or_kernel_index(double x, double lnx0, double lnf, int64_t nb) {
  int64_t ind = (int64_t)((OR_LOG(x) - lnx0) / lnf + 0.5);
  if (ind < 0)
    ind = 0;
  else if (ind > nb - 1)
    ind = nb - 1;
  return ind;
}

/* Uniform source: injected stream (ref order) or Philox keyed by pair number. */
typedef struct {
  const double *stream; /* non-NULL: consume sequentially */
  int64_t pos;
  uint64_t seed;
  uint32_t run, step, cell;
} OrRng;

static double or_draw(OrRng *g, int64_t cnt, int64_t i, int64_t j) {
  if (g->stream) return g->stream[g->pos + cnt];
  return lm_pair_uniform(g->seed, g->run, g->step, g->cell, (uint32_t)i, (uint32_t)j);
}

/*
 * all_or_nothing.py:498-635 (E table, vel 2D) and :637-751 (E const when E == NULL).
 * xis, m_w, m_s, radii are updated in place.  vel is (2, n) with row stride vstride.
 */
static void coll_cell_2D(int64_t n, double *xis, double *m_w, double *m_s, double *radii, const double *u,
                         const double *w, const double *rho, double dt_over_dV, const double *E, double E_const,
                         int64_t nb, double lnR0, double lnf, int64_t *no_cols, OrRng *g, int64_t *ind) {
  if (E)
    for (int64_t i = 0; i < n; ++i) ind[i] = or_kernel_index(radii[i], lnR0, lnf, nb);
  int64_t cnt = 0;
  for (int64_t i = 0; i < n - 1; ++i) {
    for (int64_t j = i + 1; j < n; ++j) {
      int64_t imin, imax;
      if (xis[i] <= xis[j]) {
        imin = i;
        imax = j;
      } else {
        imin = j;
        imax = i;
      }
      const double xi_min = xis[imin], xi_max = xis[imax];
      const double m_w_min = m_w[imin], m_w_max = m_w[imax];
      const double m_s_min = m_s[imin], m_s_max = m_s[imax];
      const double du = u[i] - u[j], dw = w[i] - w[j];
      const double Ec = E ? E[ind[i] * nb + ind[j]] : E_const;
      const double p_crit = xi_max * dt_over_dV * kernel_hydro(radii[i], radii[j], Ec, sqrt(du * du + dw * dw));
      if (p_crit > 1.0) {
        const double xi_col = p_crit * xi_min;
        m_w[imin] = m_w_min + p_crit * m_w_max;
        m_s[imin] = m_s_min + p_crit * m_s_max;
        radii[imin] = radius_from_mass((m_w[imin] + m_s[imin]), rho[imin]);
        xis[imax] -= xi_col;
        if (E) ind[imin] = or_kernel_index(radii[imin], lnR0, lnf, nb);
        no_cols[1] += 1;
      } else if (p_crit > or_draw(g, cnt, i, j)) {
        no_cols[0] += 1;
        const double xi_rel_dev = (xi_max - xi_min) / xi_max;
        if (xi_rel_dev < 1.0E-5) {
          const double xi_ges = xi_min + xi_max;
          m_w[imin] = 2.0 * (xi_min * m_w_min + xi_max * m_w_max) / xi_ges;
          m_s[imin] = 2.0 * (xi_min * m_s_min + xi_max * m_s_max) / xi_ges;
          m_w[imax] = m_w[imin];
          m_s[imax] = m_s[imin];
          radii[imin] = radius_from_mass((m_w[imin] + m_s[imin]), rho[imin]);
          radii[imax] = radii[imin];
          xis[imax] = 0.5 * 0.7 * xi_ges;
          xis[imin] = 0.5 * xi_ges - xis[imax];
          if (E) {
            ind[imin] = or_kernel_index(radii[imin], lnR0, lnf, nb);
            ind[imax] = ind[imin];
          }
        } else {
          m_w[imin] += m_w_max;
          m_s[imin] += m_s_max;
          radii[imin] = radius_from_mass((m_w[imin] + m_s[imin]), rho[imin]);
          xis[imax] -= xi_min;
          if (E) ind[imin] = or_kernel_index(radii[imin], lnR0, lnf, nb);
        }
      }
      cnt += 1;
    }
  }
}

/* Single cell, all_or_nothing.py:498-635.  rnd: n(n-1)/2 injected uniforms (or NULL -> Philox with key args). */
EXPORT void or_collision_cell_2D(int64_t n, double *xis, double *m_w, double *m_s, double *radii, const double *u,
                                 const double *w, const double *rho, double dt_over_dV, const double *E,
                                 double E_const, int64_t nb, double lnR0, double lnf, int64_t *no_cols,
                                 const double *rnd, uint64_t seed, uint32_t run, uint32_t step, uint32_t cell) {
  or_init_consts();
  OrRng g = {rnd, 0, seed, run, step, cell};
  int64_t *ind = (int64_t *)malloc(sizeof(int64_t) * (n > 0 ? n : 1));
  coll_cell_2D(n, xis, m_w, m_s, radii, u, w, rho, dt_over_dV, E, E_const, nb, lnR0, lnf, no_cols, &g, ind);
  free(ind);
}

/*
 * all_or_nothing.py:754-846.  cells (2,N) int64, vel (2,N), grid_T (nx,nz).
 * rnd_stream: if non-NULL, consumed cell by cell in x-major order exactly like
 * the reference's successive np.random.rand calls; else Philox(seed, run, step,
 * cell = i*nz + j, cnt).
 */
EXPORT void or_collision_all_cells(int64_t N, double *xis, double *m_w, double *m_s, const double *vel,
                                   const double *grid_T, const int64_t *cells, int64_t nx, int64_t nz,
                                   double dt_over_dV, const double *E, double E_const, int64_t nb, double lnR0,
                                   double lnf, int64_t *no_cols, int solute, const double *rnd_stream, uint64_t seed,
                                   uint32_t run, uint32_t step) {
  or_init_consts();
  int64_t *cnt_cell = (int64_t *)calloc((size_t)(nx * nz), sizeof(int64_t));
  for (int64_t id = 0; id < N; ++id) {
    const int64_t i = cells[id], j = cells[N + id];
    if (i >= 0 && i < nx && j >= 0 && j < nz) cnt_cell[i * nz + j]++;
  }
  int64_t *start = (int64_t *)malloc(sizeof(int64_t) * (size_t)(nx * nz + 1));
  start[0] = 0;
  for (int64_t c = 0; c < nx * nz; ++c) start[c + 1] = start[c] + cnt_cell[c];
  int64_t *members = (int64_t *)malloc(sizeof(int64_t) * (size_t)(start[nx * nz] > 0 ? start[nx * nz] : 1));
  for (int64_t c = 0; c < nx * nz; ++c) cnt_cell[c] = 0;
  for (int64_t id = 0; id < N; ++id) { /* ascending ID within each cell (boolean mask order) */
    const int64_t i = cells[id], j = cells[N + id];
    if (i >= 0 && i < nx && j >= 0 && j < nz) {
      const int64_t c = i * nz + j;
      members[start[c] + cnt_cell[c]++] = id;
    }
  }
  int64_t maxn = 1;
  for (int64_t c = 0; c < nx * nz; ++c)
    if (cnt_cell[c] > maxn) maxn = cnt_cell[c];
  double *buf = (double *)malloc(sizeof(double) * (size_t)(7 * maxn));
  int64_t *ind = (int64_t *)malloc(sizeof(int64_t) * (size_t)maxn);
  OrRng g = {rnd_stream, 0, seed, run, step, 0};
  for (int64_t i = 0; i < nx; ++i) {
    for (int64_t j = 0; j < nz; ++j) {
      const int64_t c = i * nz + j;
      const int64_t n = cnt_cell[c];
      double *x_c = buf, *mw_c = buf + maxn, *ms_c = buf + 2 * maxn, *R_c = buf + 3 * maxn;
      double *rho_c = buf + 4 * maxn, *u_c = buf + 5 * maxn, *w_c = buf + 6 * maxn;
      const double T = grid_T[i * nz + j];
      for (int64_t k = 0; k < n; ++k) {
        const int64_t id = members[start[c] + k];
        x_c[k] = xis[id];
        mw_c[k] = m_w[id];
        ms_c[k] = m_s[id];
        double ws;
        R_p_w_s_rho_p(mw_c[k], ms_c[k], T, solute, &R_c[k], &ws, &rho_c[k]);
        u_c[k] = vel[id];
        w_c[k] = vel[N + id];
      }
      g.cell = (uint32_t)c;
      coll_cell_2D(n, x_c, mw_c, ms_c, R_c, u_c, w_c, rho_c, dt_over_dV, E, E_const, nb, lnR0, lnf, no_cols, &g,
                   ind);
      if (rnd_stream) g.pos += n * (n - 1) / 2;
      for (int64_t k = 0; k < n; ++k) {
        const int64_t id = members[start[c] + k];
        xis[id] = x_c[k];
        m_w[id] = mw_c[k];
        m_s[id] = ms_c[k];
      }
    }
  }
  free(buf);
  free(ind);
  free(members);
  free(start);
  free(cnt_cell);
}

/* all_or_nothing.py:109-185 (Golovin kernel, box model) */
EXPORT void or_collision_golovin(int64_t n, double *xis, double *masses, double dt_over_dV, int64_t *no_cols,
                                 const double *rnd, uint64_t seed, uint32_t run, uint32_t step) {
  or_init_consts();
  OrRng g = {rnd, 0, seed, run, step, 0};
  int64_t cnt = 0;
  for (int64_t i = 0; i < n - 1; ++i) {
    for (int64_t j = i + 1; j < n; ++j) {
      int64_t imin, imax;
      if (xis[i] <= xis[j]) {
        imin = i;
        imax = j;
      } else {
        imin = j;
        imax = i;
      }
      const double xi_min = xis[imin], xi_max = xis[imax];
      const double m_min = masses[imin], m_max = masses[imax];
      const double p_crit = xi_max * kernel_golovin(m_min, m_max) * dt_over_dV;
      if (p_crit > 1.0) {
        no_cols[1] += 1;
        const double xi_col = p_crit * xi_min;
        masses[imin] = (xi_min * m_min + xi_col * m_max) / xi_min;
        xis[imax] -= xi_col;
      } else if (p_crit > or_draw(&g, cnt, i, j)) {
        no_cols[0] += 1;
        const double xi_rel_dev = (xi_max - xi_min) / xi_max;
        if (xi_rel_dev < 1.0E-5) {
          const double xi_ges = xi_min + xi_max;
          masses[imin] = 2.0 * (xi_min * m_min + xi_max * m_max) / xi_ges;
          masses[imax] = masses[imin];
          xis[imax] = 0.5 * 0.7 * xi_ges;
          xis[imin] = 0.5 * xi_ges - xis[imax];
        } else {
          masses[imin] += m_max;
          xis[imax] -= xi_min;
        }
      }
      cnt += 1;
    }
  }
}

/* all_or_nothing.py:392-491 (box model, kernel K(m1, m2) tabulated on a log-mass grid) */
EXPORT void or_collision_kernel_grid_m(int64_t n, double *xis, double *masses, double dt_over_dV,
                                       const double *kernel_grid, int64_t nb, double lnm0, double lnf,
                                       int64_t *no_cols, const double *rnd, uint64_t seed, uint32_t run,
                                       uint32_t step) {
  or_init_consts();
  OrRng g = {rnd, 0, seed, run, step, 0};
  int64_t *ind = (int64_t *)malloc(sizeof(int64_t) * (n > 0 ? n : 1));
  for (int64_t i = 0; i < n; ++i) ind[i] = or_kernel_index(masses[i], lnm0, lnf, nb);
  int64_t cnt = 0;
  for (int64_t i = 0; i < n - 1; ++i) {
    for (int64_t j = i + 1; j < n; ++j) {
      const int64_t imin = xis[i] <= xis[j] ? i : j, imax = xis[i] <= xis[j] ? j : i;
      const double xi_min = xis[imin], xi_max = xis[imax];
      const double m_min = masses[imin], m_max = masses[imax];
      const double p_crit = xi_max * kernel_grid[ind[i] * nb + ind[j]] * dt_over_dV;
      if (p_crit > 1.0) {
        const double xi_col = p_crit * xi_min;
        masses[imin] = (xi_min * m_min + xi_col * m_max) / xi_min;
        xis[imax] -= xi_col;
        ind[imin] = or_kernel_index(masses[imin], lnm0, lnf, nb);
        no_cols[1] += 1;
      } else if (p_crit > or_draw(&g, cnt, i, j)) {
        no_cols[0] += 1;
        if ((xi_max - xi_min) / xi_max < 1.0E-5) {
          const double xi_ges = xi_min + xi_max;
          masses[imin] = 2.0 * (xi_min * m_min + xi_max * m_max) / xi_ges;
          masses[imax] = masses[imin];
          xis[imax] = 0.5 * 0.7 * xi_ges;
          xis[imin] = 0.5 * xi_ges - xis[imax];
          ind[imin] = or_kernel_index(masses[imin], lnm0, lnf, nb);
          ind[imax] = ind[imin];
        } else {
          masses[imin] += m_max;
          xis[imax] -= xi_min;
          ind[imin] = or_kernel_index(masses[imin], lnm0, lnf, nb);
        }
      }
      cnt += 1;
    }
  }
  free(ind);
}

/* grid.py:370-412: r_l = 1e-18 m_dry^-1 sum over active ids (ascending) of xi m_w */
EXPORT void or_update_grid_r_l(int64_t N, const double *m_w, const double *xi, const int64_t *cells,
                               const uint8_t *active, int64_t nx, int64_t nz, double *r_l, const double *m_dry_inv) {
  for (int64_t c = 0; c < nx * nz; ++c) r_l[c] = 0.0;
  for (int64_t id = 0; id < N; ++id)
    if (active[id]) r_l[cells[id] * nz + cells[N + id]] += m_w[id] * xi[id];
  for (int64_t c = 0; c < nx * nz; ++c) r_l[c] *= 1.0E-18 * m_dry_inv[c];
}

/*
 * evaluation.py:363-474 for one seed and one frame: T_p = T of each SIP's cell (top row for inactive SIPs,
 * :176-200), R_p at T_p (:317-321), classes by np.digitize(R_p, [0.5, 25]) (:319, :383-385), then per cell,
 * in ascending id over the active SIPs: the mixing ratio (:93-133) and number per dry mass (:136-175) of
 * each class, and the radius moments (:1393-1444; R**n as repeated multiplication, R**3 = R*(R*R)).
 * out[f*nc + c], f = r_aero, r_c, r_r, n_aero, n_c, n_r, R_avg, R_2/1, R_eff (:404-474).
 */
EXPORT void or_derived_fields(int64_t N, const double *m_w, const double *m_s, const double *xi,
                              const int64_t *cells, const uint8_t *active, const double *T_grid,
                              const double *mass_dry_inv, int64_t nx, int64_t nz, double V0, int solute,
                              double *out) {
  or_init_consts();
  const int64_t nc = nx * nz;
  double *S = (double *)calloc((size_t)(11 * nc), sizeof(double));
  double *mass = S, *num = S + 3 * nc, *m0 = S + 6 * nc, *m1 = S + 7 * nc, *c1 = S + 8 * nc, *c2 = S + 9 * nc,
         *c3 = S + 10 * nc;
  for (int64_t id = 0; id < N; ++id) {
    const int64_t cx = cells[id];
    const int64_t cz = cells[N + id] < 0 ? cells[N + id] + nz : cells[N + id];
    double R, w, rho;
    R_p_w_s_rho_p(m_w[id], m_s[id], T_grid[cx * nz + cz], solute, &R, &w, &rho);
    if (!active[id]) continue;
    const int cls = R < 0.5 ? 0 : (R < 25. ? 1 : 2);
    const int64_t c = cx * nz + cz;
    mass[cls * nc + c] += m_w[id] * xi[id];
    num[cls * nc + c] += xi[id];
    m0[c] += xi[id];
    m1[c] += xi[id] * R;
    if (cls == 1) {
      c1[c] += xi[id] * R;
      c2[c] += xi[id] * (R * R);
      c3[c] += xi[id] * (R * (R * R));
    }
  }
  for (int64_t c = 0; c < nc; ++c) {
    for (int k = 0; k < 3; ++k) {
      out[k * nc + c] = mass[k * nc + c] * (1.0E-18 * mass_dry_inv[c]);
      out[(3 + k) * nc + c] = num[k * nc + c] * mass_dry_inv[c];
    }
    const double M0 = m0[c] / V0, M1 = m1[c] / V0, C1 = c1[c] / V0, C2 = c2[c] / V0, C3 = c3[c] / V0;
    out[6 * nc + c] = M0 == 0.0 ? 0.0 : M1 / M0;
    out[7 * nc + c] = C1 == 0.0 ? 0.0 : C2 / C1;
    out[8 * nc + c] = C2 == 0.0 ? 0.0 : C3 / C2;
  }
  free(S);
}

/* evaluation.py:443-474: mean over seeds (summed in seed order) and sqrt((sum f^2 - n mean^2) / (n (n-1))) */
EXPORT void or_ensemble_stats(int64_t n_seeds, int64_t n_vals, const double *f, double *mean, double *std) {
  for (int64_t k = 0; k < n_vals; ++k) {
    double s = 0.0, q = 0.0;
    for (int64_t r = 0; r < n_seeds; ++r) {
      s += f[r * n_vals + k];
      q += f[r * n_vals + k] * f[r * n_vals + k];
    }
    const double m = s / (double)n_seeds;
    mean[k] = m;
    if (n_seeds <= 1) {
      std[k] = 0.0;
    } else {
      double h = q - (double)n_seeds * m * m;
      if (h < 0) h = 0.0;
      std[k] = sqrt(h / (double)(n_seeds * (n_seeds - 1)));
    }
  }
}

/* all_or_nothing.py:269-390 (box model, E table, 1D velocity) */
EXPORT void or_collision_ecol_box(int64_t n, double *xis, double *masses, double *radii, const double *vel,
                                  const double *rho, double dt_over_dV, const double *E, int64_t nb, double lnR0,
                                  double lnf, int64_t *no_cols, const double *rnd, uint64_t seed, uint32_t run,
                                  uint32_t step) {
  or_init_consts();
  OrRng g = {rnd, 0, seed, run, step, 0};
  int64_t *ind = (int64_t *)malloc(sizeof(int64_t) * (n > 0 ? n : 1));
  for (int64_t i = 0; i < n; ++i) ind[i] = or_kernel_index(radii[i], lnR0, lnf, nb);
  int64_t cnt = 0;
  for (int64_t i = 0; i < n - 1; ++i) {
    for (int64_t j = i + 1; j < n; ++j) {
      int64_t imin, imax;
      if (xis[i] <= xis[j]) {
        imin = i;
        imax = j;
      } else {
        imin = j;
        imax = i;
      }
      const double xi_min = xis[imin], xi_max = xis[imax];
      const double m_min = masses[imin], m_max = masses[imax];
      const double p_crit =
          xi_max * dt_over_dV * kernel_hydro(radii[i], radii[j], E[ind[i] * nb + ind[j]], fabs(vel[i] - vel[j]));
      if (p_crit > 1.0) {
        const double xi_col = p_crit * xi_min;
        masses[imin] = (xi_min * m_min + xi_col * m_max) / xi_min;
        radii[imin] = radius_from_mass(masses[imin], rho[imin]);
        xis[imax] -= xi_col;
        ind[imin] = or_kernel_index(radii[imin], lnR0, lnf, nb);
        no_cols[1] += 1;
      } else if (p_crit > or_draw(&g, cnt, i, j)) {
        no_cols[0] += 1;
        const double xi_rel_dev = (xi_max - xi_min) / xi_max;
        if (xi_rel_dev < 1.0E-5) {
          const double xi_ges = xi_min + xi_max;
          masses[imin] = 2.0 * (xi_min * m_min + xi_max * m_max) / xi_ges;
          masses[imax] = masses[imin];
          radii[imin] = radius_from_mass(masses[imin], rho[imin]);
          radii[imax] = radii[imin];
          xis[imax] = 0.5 * 0.7 * xi_ges;
          xis[imin] = 0.5 * xi_ges - xis[imax];
          ind[imin] = or_kernel_index(radii[imin], lnR0, lnf, nb);
          ind[imax] = ind[imin];
        } else {
          masses[imin] += m_max;
          radii[imin] = radius_from_mass(masses[imin], rho[imin]);
          xis[imax] -= xi_min;
          ind[imin] = or_kernel_index(radii[imin], lnR0, lnf, nb);
        }
      }
      cnt += 1;
    }
  }
  free(ind);
}

EXPORT void or_update_velocity_Beard(int64_t n, double *vel, const double *R) { /* kernel.py:193-210 */
  for (int64_t i = 0; i < n; ++i) vel[i] = or_terminal_velocity_Beard(R[i]);
}

EXPORT double or_radius_from_mass(double m, double rho) {
  or_init_consts();
  return radius_from_mass(m, rho);
}

EXPORT void or_R_p_w_s_rho_p(int64_t n, const double *m_w, const double *m_s, const double *T, int solute,
                             double *R, double *w_s, double *rho) {
  or_init_consts();
  for (int64_t k = 0; k < n; ++k) R_p_w_s_rho_p(m_w[k], m_s[k], T[k], solute, &R[k], &w_s[k], &rho[k]);
}

EXPORT double or_surface_tension(int solute, double w, double T) {
  or_init_consts();
  return solute == 0 ? surface_tension_AS(w, T) : surface_tension_NaCl(w, T);
}

/* ------------------------------------------------------------------------ */
/* Grid: integration.py:359-483, atmosphere.py:373-443                      */
/* ------------------------------------------------------------------------ */
/* scal (10,nx,nz), mat (7,nx,nz); integration.py:359-400 (one cell) */
static void update_material_properties_cell(double *scal, double *mat, int64_t c, int64_t nc) {
  const double T = scal[0 * nc + c], p = scal[1 * nc + c];
  mat[0 * nc + c] = thermal_conductivity_air(T);
  mat[1 * nc + c] = diffusion_constant(T, p);
  mat[2 * nc + c] = heat_of_vaporization(T);
  mat[3 * nc + c] = surface_tension_water(T);
  mat[4 * nc + c] = K.c_pd * (1 + 0.897 * scal[4 * nc + c]); /* atmosphere.py:47-68 */
  mat[5 * nc + c] = viscosity_air(T);
  mat[6 * nc + c] = scal[3 * nc + c] * (1.0 + scal[4 * nc + c]);
}

EXPORT void or_update_material_properties(double *scal, double *mat, int64_t nc) {
  or_init_consts();
  for (int64_t c = 0; c < nc; ++c) update_material_properties_cell(scal, mat, c, nc);
}

/* integration.py:402-483 */
EXPORT void or_propagate_grid_subloop_step(double *scal, double *mat, double p_ref, double p_ref_inv,
                                           const double *dTheta_ad, const double *dr_v_ad, const double *delta_m_l,
                                           double V_cell, int64_t nc) {
  or_init_consts();
  for (int64_t c = 0; c < nc; ++c) {
    double *T = &scal[0 * nc + c], *p = &scal[1 * nc + c], *Th = &scal[2 * nc + c];
    const double rho_d = scal[3 * nc + c];
    double *r_v = &scal[4 * nc + c], *S = &scal[6 * nc + c], *e_s = &scal[7 * nc + c];
    const double m_dry_inv = scal[8 * nc + c];
    const double dml = delta_m_l[c];
    *Th += dTheta_ad[c];
    *r_v += dr_v_ad[c] - dml * m_dry_inv;
    const double Theta_over_T = OR_POW(p_ref_inv * K.R_d * rho_d * *Th, K.kappa_factor2); /* atm.py:408-443 */
    *Th += Theta_over_T * (mat[2 * nc + c] * dml) /
           (K.c_pd * rho_d * V_cell * (1.0 + *r_v * K.heat_factor_r_v));
    const double pdr = OR_POW(rho_d * *Th * K.R_d * p_ref_inv, K.kappa_factor); /* atm.py:373-405 */
    *T = *Th * OR_POW(pdr, K.kappa_dry);
    *p = pdr * p_ref * (1 + *r_v / K.eps_gc);
    *e_s = saturation_pressure_vapor_liquid(*T);
    *S = (rho_d * *r_v) * *T * K.R_v / *e_s; /* atm.py:314-333, ideal gas */
    update_material_properties_cell(scal, mat, c, nc);
  }
}

/* integration.py:80-102 */
static double limiter(double r) {
  const double Kv = (1.0 + 2.0 * r) * K.one_third;
  return lm_np_max(0.0, lm_np_fmin(2.0 * r, lm_np_fmin(Kv, 2.0)));
}
/* integration.py:113-148 (scalar form of the np.where) */
static double limiter_upwind(double a0, double a1, double a2) {
  double r;
  if (fabs(a1 - a2) > 1.0E-8 * fabs(a0 - a1))
    r = (a0 - a1) / (a1 - a2);
  else
    r = (a0 - a1) * 1.0E8 * lm_np_sign(a1 - a2);
  return limiter(r);
}

static int64_t wrap(int64_t i, int64_t n) { return ((i % n) + n) % n; }

/*
 * integration.py:150-286: div(field * flux).  field (nx,nz), flux (2,nx+1,nz+1);
 * bc_x/bc_z: 0 periodic, 1 solid.  out (nx,nz).
 */
EXPORT void or_divergence_upwind(const double *field, const double *flux, int64_t nx, int64_t nz, double dx,
                                 double dz, int bc_x, int bc_z, double *out) {
  or_init_consts();
  const int64_t fz = nz + 1;
  double *F = (double *)malloc(sizeof(double) * (size_t)((nx + 1) * (nz + 1)));
  /* x direction: faces i = 0..nx (left border of cell i), rows of length nz */
#define FX(ii, jj) field[(ii) * nz + (jj)]
  for (int64_t i = 0; i <= nx; ++i) {
    for (int64_t j = 0; j < nz; ++j) {
      /* ghost index g -> cell: periodic N-2,N-1 | 0..N-1 | 0,1 ; solid 0,0 | .. | N-1,N-1 */
      int64_t ia0 = i, ia1 = i - 1, ia2 = i - 2, ia3 = i + 1;
      if (bc_x == 0) {
        ia0 = wrap(ia0, nx);
        ia1 = wrap(ia1, nx);
        ia2 = wrap(ia2, nx);
        ia3 = wrap(ia3, nx);
      } else {
        ia0 = ia0 < 0 ? 0 : (ia0 > nx - 1 ? nx - 1 : ia0);
        ia1 = ia1 < 0 ? 0 : (ia1 > nx - 1 ? nx - 1 : ia1);
        ia2 = ia2 < 0 ? 0 : (ia2 > nx - 1 ? nx - 1 : ia2);
        ia3 = ia3 < 0 ? 0 : (ia3 > nx - 1 ? nx - 1 : ia3);
      }
      const double a0 = FX(ia0, j), a1 = FX(ia1, j), a2 = FX(ia2, j), a3 = FX(ia3, j);
      const double f_pos = a1 + 0.5 * limiter_upwind(a0, a1, a2) * (a1 - a2);
      const double f_neg = a0 + 0.5 * limiter_upwind(a1, a0, a3) * (a0 - a3);
      const double u = flux[0 * (nx + 1) * fz + i * fz + j];
      F[i * nz + j] = u >= 0.0 ? u * f_pos : u * f_neg;
    }
  }
  if (bc_x == 1)
    for (int64_t j = 0; j < nz; ++j) F[0 * nz + j] = F[nx * nz + j] = 0.0;
  for (int64_t i = 0; i < nx; ++i)
    for (int64_t j = 0; j < nz; ++j) out[i * nz + j] = (F[(i + 1) * nz + j] - F[i * nz + j]) / dx;
  /* z direction: faces j = 0..nz, for each column i */
  for (int64_t j = 0; j <= nz; ++j) {
    for (int64_t i = 0; i < nx; ++i) {
      int64_t ja0 = j, ja1 = j - 1, ja2 = j - 2, ja3 = j + 1;
      if (bc_z == 0) {
        ja0 = wrap(ja0, nz);
        ja1 = wrap(ja1, nz);
        ja2 = wrap(ja2, nz);
        ja3 = wrap(ja3, nz);
      } else {
        ja0 = ja0 < 0 ? 0 : (ja0 > nz - 1 ? nz - 1 : ja0);
        ja1 = ja1 < 0 ? 0 : (ja1 > nz - 1 ? nz - 1 : ja1);
        ja2 = ja2 < 0 ? 0 : (ja2 > nz - 1 ? nz - 1 : ja2);
        ja3 = ja3 < 0 ? 0 : (ja3 > nz - 1 ? nz - 1 : ja3);
      }
      const double a0 = FX(i, ja0), a1 = FX(i, ja1), a2 = FX(i, ja2), a3 = FX(i, ja3);
      const double f_pos = a1 + 0.5 * limiter_upwind(a0, a1, a2) * (a1 - a2);
      const double f_neg = a0 + 0.5 * limiter_upwind(a1, a0, a3) * (a0 - a3);
      const double w = flux[1 * (nx + 1) * fz + i * fz + j];
      F[j * nx + i] = w >= 0.0 ? w * f_pos : w * f_neg;
    }
  }
  if (bc_z == 1)
    for (int64_t i = 0; i < nx; ++i) F[0 * nx + i] = F[nz * nx + i] = 0.0;
  for (int64_t i = 0; i < nx; ++i)
    for (int64_t j = 0; j < nz; ++j) out[i * nz + j] = out[i * nz + j] + (F[(j + 1) * nx + i] - F[j * nx + i]) / dz;
#undef FX
  free(F);
}

/* integration.py:325-354: dt*(profile0 - column_mean)/t_relax, out (nz) */
EXPORT void or_relaxation_term(const double *field, const double *profile0, const double *t_relax, double dt,
                               int64_t nx, int64_t nz, double *out) {
  for (int64_t j = 0; j < nz; ++j) {
    double s = 0.0;
    for (int64_t i = 0; i < nx; ++i) s += field[i * nz + j];
    out[j] = dt * (profile0[j] - s / (double)nx) / t_relax[j];
  }
}

/* ------------------------------------------------------------------------ */
/* Particles: integration.py:489-719, 840-1089, grid.py:195-323              */
/* ------------------------------------------------------------------------ */
typedef struct {
  int64_t N, nx, nz;
  double x_min, x_max, z_min, z_max, dx, dz, V_cell, p_ref, p_ref_inv;
  double *pos, *vel, *rel_pos, *m_w, *m_s, *xi, *T_p, *water_removed;
  int64_t *cells, *no_cols;
  uint8_t *active;
  double *scal, *mat;           /* (10,nx,nz), (7,nx,nz) */
  const double *grid_vel, *flux; /* (2,nx+1,nz+1) */
  double *delta_m_l;             /* (nx,nz) */
  int solute, no_iter;
  double g_set;
  /* collisions */
  int act_collisions;
  const double *E;
  double E_const;
  int64_t nb;
  double lnR0, lnf, dt_col_over_dV;
  uint64_t seed;
  uint32_t run, col_step;
  const double *rnd_stream; /* optional: consumed by successive collision steps */
  int64_t rnd_pos;
} OrSim;

/* integration.py:509-550 */
static void update_cells_and_rel_pos(OrSim *s) {
  const int64_t N = s->N;
  for (int64_t id = 0; id < N; ++id) {
    if (s->active[id]) {
      const double x = s->pos[id], z = s->pos[N + id];
      s->rel_pos[id] = x - s->x_min;
      s->rel_pos[N + id] = z - s->z_min;
      s->cells[id] = (int64_t)floor(x / s->dx);
      s->cells[N + id] = (int64_t)floor(z / s->dz);
    }
  }
  for (int64_t id = 0; id < N; ++id) {
    s->rel_pos[id] = s->rel_pos[id] / s->dx - (double)s->cells[id];
    s->rel_pos[N + id] = s->rel_pos[N + id] / s->dz - (double)s->cells[N + id];
  }
}

/* integration.py:552-630 */
static void update_pos_from_vel_BC_PS(OrSim *s, double dt) {
  const int64_t N = s->N;
  const double zcap = s->z_max * 0.999999;
  for (int64_t id = 0; id < N; ++id) {
    s->pos[id] += dt * s->vel[id];
    s->pos[N + id] += dt * s->vel[N + id];
  }
  for (int64_t id = 0; id < N; ++id) {
    s->pos[id] = lm_pymod(s->pos[id], s->x_max);
    s->pos[N + id] = lm_np_min(s->pos[N + id], zcap);
  }
  for (int64_t id = 0; id < N; ++id) {
    if (!s->active[id]) continue;
    if (s->pos[N + id] <= s->z_min) {
      s->pos[N + id] = s->z_min - 0.01 * s->dz;
      s->vel[id] = 0.0;
      s->vel[N + id] = 0.0;
      s->active[id] = 0;
      s->water_removed[0] += s->xi[id] * s->m_w[id];
      s->cells[N + id] = -1;
    }
  }
}

/* grid.py:195-254 */
static double bilinear_weight(int64_t i, int64_t j, double a, double b, const double *f, int64_t nfx, int64_t nfz) {
  /* Python negative indexing */
  const int64_t i0 = i < 0 ? i + nfx : i, i1 = (i + 1) < 0 ? i + 1 + nfx : i + 1;
  const int64_t j0 = j < 0 ? j + nfz : j, j1 = (j + 1) < 0 ? j + 1 + nfz : j + 1;
  return a * (b * f[i1 * nfz + j1] + (1 - b) * f[i1 * nfz + j0]) +
         (1 - a) * (b * f[i0 * nfz + j1] + (1 - b) * f[i0 * nfz + j0]);
}

/* grid.py:257-323 for one SIP */
EXPORT void or_interpolate_velocity(int64_t i, int64_t j, double wx, double wy, const double *grid_vel, int64_t nx,
                                    int64_t nz, double *u_out, double *v_out) {
  const int64_t nfx = nx + 1, nfz = nz + 1;
  const double *U = grid_vel, *V = grid_vel + nfx * nfz;
  double u, v;
  if ((j == 0 && wy <= 0.5) || (j == nz - 1 && wy >= 0.5)) {
    u = wx * U[(i + 1) * nfz + j] + (1 - wx) * U[i * nfz + j]; /* weight_velocities_linear */
  } else if (wy > 0.5) {
    u = bilinear_weight(i, j, wx, wy - 0.5, U, nfx, nfz);
  } else {
    u = bilinear_weight(i, j - 1, wx, wy + 0.5, U, nfx, nfz);
  }
  if (wx > 0.5)
    v = bilinear_weight(i, j, wx - 0.5, wy, V, nfx, nfz);
  else
    v = bilinear_weight(i - 1, j, wx + 0.5, wy, V, nfx, nfz);
  *u_out = u;
  *v_out = v;
}

/* integration.py:632-719 (active SIPs, ascending ID) */
static void update_vel_impl(OrSim *s, const double *R_p, const double *rho_p, double dt) {
  const int64_t N = s->N, nc = s->nx * s->nz;
  const double *mu_g = s->mat + 5 * nc, *rhof_g = s->mat + 6 * nc;
  for (int64_t id = 0; id < N; ++id) {
    if (!s->active[id]) continue;
    const int64_t i = s->cells[id], j = s->cells[N + id];
    double uf, vf;
    or_interpolate_velocity(i, j, s->rel_pos[id], s->rel_pos[N + id], s->grid_vel, s->nx, s->nz, &uf, &vf);
    const double R = R_p[id];
    const double mu = mu_g[i * s->nz + j];
    const double rho_f = rhof_g[i * s->nz + j];
    const double dv0 = uf - s->vel[id], dv1 = vf - s->vel[N + id];
    const double vel_dev = sqrt(dv0 * dv0 + dv1 * dv1);
    const double Re = particle_reynolds_number(R, vel_dev, rho_f, mu);
    double k_dt = 4.5E12 * dt * mu / (rho_p[id] * R * R);
    if (Re > 0.5) {
      if (Re < 1000.0)
        k_dt *= (1 + 0.15 * OR_POW(Re, 0.687));
      else
        k_dt *= 0.0183333333333 * Re;
    }
    s->vel[id] = (s->vel[id] + k_dt * uf) / (1.0 + k_dt);
    s->vel[N + id] = (s->vel[N + id] + k_dt * vf - dt * s->g_set) / (1.0 + k_dt);
  }
}

/* integration.py:944-1089 */
static void propagate_particles_subloop_step(OrSim *s, double dt_sub, double dt_sub_pos, double *R_p, double *w_s,
                                             double *rho_p) {
  const int64_t N = s->N, nz = s->nz, nc = s->nx * s->nz;
  const double *T_g = s->scal, *p_g = s->scal + nc, *S_g = s->scal + 6 * nc, *es_g = s->scal + 7 * nc;
  const double *K_g = s->mat, *Dv_g = s->mat + nc, *Lv_g = s->mat + 2 * nc;
  /* 1. T_p = T_f (integration.py:489-507): inactive cells[1] = -1 -> top row */
  for (int64_t id = 0; id < N; ++id) {
    int64_t j = s->cells[N + id];
    if (j < 0) j += nz;
    s->T_p[id] = T_g[s->cells[id] * nz + j];
  }
  for (int64_t id = 0; id < N; ++id) R_p_w_s_rho_p(s->m_w[id], s->m_s[id], s->T_p[id], s->solute, &R_p[id], &w_s[id], &rho_p[id]);
  /* integration.py:840-940 */
  for (int64_t c = 0; c < nc; ++c) s->delta_m_l[c] = 0.0;
  for (int64_t id = 0; id < N; ++id) {
    if (!s->active[id]) continue;
    const int64_t c = s->cells[id] * nz + s->cells[N + id];
    const double sigma_p = s->solute == 0 ? surface_tension_AS(w_s[id], s->T_p[id]) : surface_tension_NaCl(w_s[id], s->T_p[id]);
    const double dm = or_dml_impl_Newton(s->solute, dt_sub, s->no_iter, s->m_w[id], s->m_s[id], w_s[id], R_p[id],
                                         s->T_p[id], rho_p[id], T_g[c], p_g[c], S_g[c], es_g[c], Lv_g[c], K_g[c],
                                         Dv_g[c], sigma_p, NULL);
    s->m_w[id] += dm;
    s->delta_m_l[c] += dm * s->xi[id];
  }
  for (int64_t c = 0; c < nc; ++c) s->delta_m_l[c] *= 1.0E-18;
  /* 9. */
  for (int64_t id = 0; id < N; ++id) R_p_w_s_rho_p(s->m_w[id], s->m_s[id], s->T_p[id], 0, &R_p[id], &w_s[id], &rho_p[id]);
  update_vel_impl(s, R_p, rho_p, dt_sub);
  /* 10., 11. */
  update_pos_from_vel_BC_PS(s, dt_sub_pos);
  update_cells_and_rel_pos(s);
}

static void collision_step(OrSim *s) {
  const double *stream = s->rnd_stream ? s->rnd_stream + s->rnd_pos : NULL;
  or_collision_all_cells(s->N, s->xi, s->m_w, s->m_s, s->vel, s->scal, s->cells, s->nx, s->nz, s->dt_col_over_dV,
                         s->E, s->E_const, s->nb, s->lnR0, s->lnf, s->no_cols, s->solute, stream, s->seed, s->run,
                         s->col_step);
  if (s->rnd_stream) {
    /* advance by the number of draws this step consumed */
    const int64_t N = s->N, nc = s->nx * s->nz;
    int64_t *cnt = (int64_t *)calloc((size_t)nc, sizeof(int64_t));
    for (int64_t id = 0; id < N; ++id) {
      const int64_t i = s->cells[id], j = s->cells[N + id];
      if (i >= 0 && i < s->nx && j >= 0 && j < s->nz) cnt[i * s->nz + j]++;
    }
    for (int64_t c = 0; c < nc; ++c) s->rnd_pos += cnt[c] * (cnt[c] - 1) / 2;
    free(cnt);
  }
  s->col_step += 1;
}

/* integration.py:1095-1294 */
static void integrate_subloop_n_steps(OrSim *s, const double *dTh, const double *dRv, double dt_sub,
                                      double dt_sub_pos, int no_cond_steps, int no_col_steps, double *R_p,
                                      double *w_s, double *rho_p) {
  if (s->act_collisions && no_col_steps == 1) collision_step(s);
  for (int n = 0; n < no_cond_steps; ++n) {
    if (s->act_collisions && no_col_steps > 1) collision_step(s);
    propagate_particles_subloop_step(s, dt_sub, dt_sub_pos, R_p, w_s, rho_p);
    or_propagate_grid_subloop_step(s->scal, s->mat, s->p_ref, s->p_ref_inv, dTh, dRv, s->delta_m_l, s->V_cell,
                                   s->nx * s->nz);
  }
  if (s->act_collisions && no_col_steps > no_cond_steps) collision_step(s);
}

typedef struct {
  int64_t N, nx, nz;
  double x_min, x_max, z_min, z_max, dx, dz, V_cell, p_ref, p_ref_inv;
  double dt, dt_sub, dt_sub_half, dt_col_over_dV;
  int scale_dt_cond, no_col_per_adv, act_collisions, act_relaxation, no_iter, solute;
  double g_set;
  int64_t nb;
  double lnR0, lnf, E_const;
  uint64_t seed;
  uint32_t run, col_step; /* col_step is in/out */
  int64_t rnd_pos;        /* in/out */
} OrAdvParams;

/*
 * integration.py:1298-1611: n_steps advection steps.  Arrays in reference
 * layout, all updated in place.  rel_pos and T_p are the caller's persistent
 * work arrays (as in simulate, integration.py:1932-1936).
 */
EXPORT void or_integrate_adv_steps(OrAdvParams *P, int64_t n_steps, double *pos, double *vel, int64_t *cells,
                                   double *rel_pos, double *m_w, double *m_s, double *xi, uint8_t *active,
                                   double *T_p, double *water_removed, int64_t *no_cols, double *scal, double *mat,
                                   const double *grid_vel, const double *flux, const double *E,
                                   const double *prof_rv, const double *prof_Th, const double *t_relax,
                                   const double *rnd_stream) {
  or_init_consts();
  OrSim s;
  memset(&s, 0, sizeof(s));
  s.N = P->N;
  s.nx = P->nx;
  s.nz = P->nz;
  s.x_min = P->x_min;
  s.x_max = P->x_max;
  s.z_min = P->z_min;
  s.z_max = P->z_max;
  s.dx = P->dx;
  s.dz = P->dz;
  s.V_cell = P->V_cell;
  s.p_ref = P->p_ref;
  s.p_ref_inv = P->p_ref_inv;
  s.pos = pos;
  s.vel = vel;
  s.rel_pos = rel_pos;
  s.m_w = m_w;
  s.m_s = m_s;
  s.xi = xi;
  s.T_p = T_p;
  s.water_removed = water_removed;
  s.cells = cells;
  s.no_cols = no_cols;
  s.active = active;
  s.scal = scal;
  s.mat = mat;
  s.grid_vel = grid_vel;
  s.flux = flux;
  s.solute = P->solute;
  s.no_iter = P->no_iter;
  s.g_set = P->g_set;
  s.act_collisions = P->act_collisions;
  s.E = E;
  s.E_const = P->E_const;
  s.nb = P->nb;
  s.lnR0 = P->lnR0;
  s.lnf = P->lnf;
  s.dt_col_over_dV = P->dt_col_over_dV;
  s.seed = P->seed;
  s.run = P->run;
  s.col_step = P->col_step;
  s.rnd_stream = rnd_stream;
  s.rnd_pos = P->rnd_pos;
  const int64_t N = P->N, nx = P->nx, nz = P->nz, nc = nx * nz;
  s.delta_m_l = (double *)calloc((size_t)nc, sizeof(double));
  double *R_p = (double *)malloc(sizeof(double) * (size_t)(N > 0 ? N : 1));
  double *w_s = (double *)malloc(sizeof(double) * (size_t)(N > 0 ? N : 1));
  double *rho_p = (double *)malloc(sizeof(double) * (size_t)(N > 0 ? N : 1));
  double *div = (double *)malloc(sizeof(double) * (size_t)nc);
  double *dRv = (double *)malloc(sizeof(double) * (size_t)nc);
  double *dTh = (double *)malloc(sizeof(double) * (size_t)nc);
  double *rel = (double *)malloc(sizeof(double) * (size_t)nz);
  const double dt_sub = P->dt_sub;
  for (int64_t step = 0; step < n_steps; ++step) {
    update_pos_from_vel_BC_PS(&s, P->dt_sub_half);
    update_cells_and_rel_pos(&s);
    or_divergence_upwind(scal + 4 * nc, flux, nx, nz, P->dx, P->dz, 0, 1, div);
    for (int64_t c = 0; c < nc; ++c) dRv[c] = -dt_sub * div[c] * scal[9 * nc + c];
    or_divergence_upwind(scal + 2 * nc, flux, nx, nz, P->dx, P->dz, 0, 1, div);
    for (int64_t c = 0; c < nc; ++c) dTh[c] = -dt_sub * div[c] * scal[9 * nc + c];
    if (P->act_relaxation) {
      or_relaxation_term(scal + 4 * nc, prof_rv, t_relax, dt_sub, nx, nz, rel);
      for (int64_t c = 0; c < nc; ++c) dRv[c] += rel[c % nz];
      or_relaxation_term(scal + 2 * nc, prof_Th, t_relax, dt_sub, nx, nz, rel);
      for (int64_t c = 0; c < nc; ++c) dTh[c] += rel[c % nz];
    }
    int no_col_steps = P->no_col_per_adv == 1 ? 0 : (P->no_col_per_adv == 2 ? 1 : P->scale_dt_cond);
    integrate_subloop_n_steps(&s, dTh, dRv, dt_sub, dt_sub, P->scale_dt_cond, no_col_steps, R_p, w_s, rho_p);
    or_divergence_upwind(scal + 4 * nc, flux, nx, nz, P->dx, P->dz, 0, 1, div);
    for (int64_t c = 0; c < nc; ++c) dRv[c] = -2.0 * dt_sub * div[c] * scal[9 * nc + c] - dRv[c];
    or_divergence_upwind(scal + 2 * nc, flux, nx, nz, P->dx, P->dz, 0, 1, div);
    for (int64_t c = 0; c < nc; ++c) dTh[c] = -2.0 * dt_sub * div[c] * scal[9 * nc + c] - dTh[c];
    if (P->act_relaxation) {
      or_relaxation_term(scal + 4 * nc, prof_rv, t_relax, dt_sub, nx, nz, rel);
      for (int64_t c = 0; c < nc; ++c) dRv[c] += 2. * rel[c % nz];
      or_relaxation_term(scal + 2 * nc, prof_Th, t_relax, dt_sub, nx, nz, rel);
      for (int64_t c = 0; c < nc; ++c) dTh[c] += 2. * rel[c % nz];
    }
    no_col_steps = (P->no_col_per_adv == 1 || P->no_col_per_adv == 2) ? 1 : P->scale_dt_cond;
    integrate_subloop_n_steps(&s, dTh, dRv, dt_sub, dt_sub, P->scale_dt_cond - 1, no_col_steps, R_p, w_s, rho_p);
    propagate_particles_subloop_step(&s, dt_sub, P->dt_sub_half, R_p, w_s, rho_p);
    or_propagate_grid_subloop_step(scal, mat, s.p_ref, s.p_ref_inv, dTh, dRv, s.delta_m_l, s.V_cell, nc);
  }
  P->col_step = s.col_step;
  P->rnd_pos = s.rnd_pos;
  free(s.delta_m_l);
  free(R_p);
  free(w_s);
  free(rho_p);
  free(div);
  free(dRv);
  free(dTh);
  free(rel);
}

/* Single particle substep / grid substep exposed for unit tests */
EXPORT void or_propagate_particles_substep(OrAdvParams *P, double dt_sub, double dt_sub_pos, double *pos,
                                           double *vel, int64_t *cells, double *rel_pos, double *m_w, double *m_s,
                                           double *xi, uint8_t *active, double *T_p, double *water_removed,
                                           double *scal, double *mat, const double *grid_vel, double *delta_m_l) {
  or_init_consts();
  OrSim s;
  memset(&s, 0, sizeof(s));
  s.N = P->N;
  s.nx = P->nx;
  s.nz = P->nz;
  s.x_min = P->x_min;
  s.x_max = P->x_max;
  s.z_min = P->z_min;
  s.z_max = P->z_max;
  s.dx = P->dx;
  s.dz = P->dz;
  s.pos = pos;
  s.vel = vel;
  s.rel_pos = rel_pos;
  s.m_w = m_w;
  s.m_s = m_s;
  s.xi = xi;
  s.T_p = T_p;
  s.water_removed = water_removed;
  s.cells = cells;
  s.active = active;
  s.scal = scal;
  s.mat = mat;
  s.grid_vel = grid_vel;
  s.delta_m_l = delta_m_l;
  s.solute = P->solute;
  s.no_iter = P->no_iter;
  s.g_set = P->g_set;
  const int64_t N = P->N > 0 ? P->N : 1;
  double *R_p = (double *)malloc(sizeof(double) * (size_t)N);
  double *w_s = (double *)malloc(sizeof(double) * (size_t)N);
  double *rho_p = (double *)malloc(sizeof(double) * (size_t)N);
  propagate_particles_subloop_step(&s, dt_sub, dt_sub_pos, R_p, w_s, rho_p);
  free(R_p);
  free(w_s);
  free(rho_p);
}

EXPORT void or_update_pos_and_cells(OrAdvParams *P, double dt, double *pos, double *vel, int64_t *cells,
                                    double *rel_pos, double *m_w, double *xi, uint8_t *active, double *water_removed) {
  OrSim s;
  memset(&s, 0, sizeof(s));
  s.N = P->N;
  s.nx = P->nx;
  s.nz = P->nz;
  s.x_min = P->x_min;
  s.x_max = P->x_max;
  s.z_min = P->z_min;
  s.z_max = P->z_max;
  s.dx = P->dx;
  s.dz = P->dz;
  s.pos = pos;
  s.vel = vel;
  s.rel_pos = rel_pos;
  s.m_w = m_w;
  s.xi = xi;
  s.water_removed = water_removed;
  s.cells = cells;
  s.active = active;
  if (dt != 0.0) update_pos_from_vel_BC_PS(&s, dt);
  update_cells_and_rel_pos(&s);
}

/* Philox known-answer access for tests */
EXPORT void or_philox(const uint32_t *ctr, const uint32_t *key, uint32_t *out) {
  lm_u32x4 c;
  for (int i = 0; i < 4; ++i) c.v[i] = ctr[i];
  lm_u32x4 o = lm_philox4x32_10(c, key[0], key[1]);
  for (int i = 0; i < 4; ++i) out[i] = o.v[i];
}

EXPORT double or_pair_uniform(uint64_t seed, uint32_t run, uint32_t step, uint32_t cell, uint32_t i, uint32_t j) {
  return lm_pair_uniform(seed, run, step, cell, i, j);
}

/* lm_math access for tests */
EXPORT void or_lm_vec(int which, int64_t n, const double *x, const double *y, double *out) {
  for (int64_t k = 0; k < n; ++k) {
    if (which == 0)
      out[k] = lm_exp(x[k]);
    else if (which == 1)
      out[k] = lm_log(x[k]);
    else if (which == 2)
      out[k] = lm_pow(x[k], y[k]);
    else if (which == 3)
      out[k] = lm_pymod(x[k], y[k]);
    else if (which == 7)
      out[k] = lm_pow_third(x[k]);
  }
}

/* ------------------------------------------------------------------------ */
/* initial state: init.py:439-1236 (one run), generation_SIP_ensemble.py:   */
/* 147-404, microphysics.py:531-610, 686-966, atmosphere.py:77-135           */
/* ------------------------------------------------------------------------ */
/*
 * Same layout as LmInitPar of the product's C ABI (the parameters init.py derives from the config).
 * Substitutions, shared with the product and documented there: Philox-4x32-10 uniforms keyed by the
 * run's seed (lm_init_uniform) for np.random.rand; 100 golden-section steps for fminbound and bisection
 * to the last bit for brentq; plain left-to-right sums for np.sum.
 */
typedef struct {
  int32_t n_modes, solute, weak_threshold, newton_iterations;
  int32_t iter_cnt_limit, nx, n_runs, pad;
  int32_t l_max[4];
  double mu_m_log[4], sigma_m_log[4], kappa[4], r_critmin[4], no_rpcm_0[4];
  double eta, mass_density_dry, S_init_max, dt_init;
  double p_ref, Theta_l, r_tot_0, rho_dry_0;
  double X, dx, dy, dz;
} OrInitPar;

#define OR_G 9.80665

/* init.py:384-437 with atmosphere.py:77-97 (kappa_moist) and :103-135 (beta) */
static void or_profile_dry(double z, double z_0, double p_0, double p_ref, double Theta_l, double r_tot, double *T,
                           double *p, double *rho_dry, double *S) {
  const double p_0_over_p_ref = p_0 / p_ref;
  const double kappa_tot = K.kappa_dry * (1 - 0.289 * r_tot);
  const double kappa_tot_inv = 1.0 / kappa_tot;
  const double p_0_over_p_ref_to_kappa_tot = OR_POW(p_0_over_p_ref, kappa_tot);
  const double beta_tot =
      OR_G * (K.kappa_dry * (1 - 0.289 * r_tot)) * (1 + r_tot) / (K.R_d * Theta_l * (1 + r_tot / K.eps_gc));
  const double T_over_Theta_l = p_0_over_p_ref_to_kappa_tot - beta_tot * (z - z_0);
  *T = T_over_Theta_l * Theta_l;
  *p = OR_POW(T_over_Theta_l, kappa_tot_inv) * p_ref;
  *rho_dry = *p * beta_tot / (T_over_Theta_l * OR_G * kappa_tot);
  /* atmosphere.py:314-335: e = rho_v T R_v; material_properties.py:546-563 */
  *S = *rho_dry * r_tot * *T * K.R_v / saturation_pressure_vapor_liquid(*T);
}

/* microphysics.py:400-446 (Kelvin term from the mass fraction), :570-610 (S_eq of AS / NaCl) */
static double or_S_eq_mf(int solute, double w_s, double T_p, double m_s) {
  const double rho_p = solute == 0 ? density_AS_solution(w_s, T_p) : density_NaCl_solution(w_s, T_p);
  const double sigma_p = solute == 0 ? surface_tension_AS(w_s, T_p) : surface_tension_NaCl(w_s, T_p);
  const double a_w = solute == 0 ? water_activity_AS(w_s) : water_activity_NaCl(w_s);
  const double volume_to_radius = OR_POW(K.four_pi_over_three_inv, K.one_third); /* constants.py:32 */
  const double kelvin = OR_EXP(2.0 * sigma_p * 1.0E6 * OR_POW(w_s / m_s, K.one_third) /
                               (K.R_v * T_p * OR_POW(rho_p, 0.66666667) * volume_to_radius));
  return a_w * kelvin;
}

/*
 * microphysics.py:866-966 (AS, w_s_max 0.78) and :686-785 (NaCl, w_s_max 0.45): the initial solute mass
 * fraction.  The maximum of S_eq on [1e-8, w_s_max] by golden-section search (100 steps, the interval
 * shrinks below one ulp long before), w_act = 1.000001 x its midpoint; the root of S_eq - S_amb on
 * [w_act, w_s_max] by bisection until the midpoint is an end point.
 */
EXPORT double or_initial_mass_fraction(int solute, double m_s, double S_amb, double T_amb) {
  or_init_consts();
  const double w_s_max = solute == 0 ? 0.78 : 0.45;
  if (S_amb <= or_S_eq_mf(solute, w_s_max, T_amb, m_s)) return w_s_max;
  const double inv_phi = 0.6180339887498949; /* (sqrt(5) - 1) / 2 */
  double lo = 1.0E-8, hi = w_s_max;
  double x1 = hi - inv_phi * (hi - lo), x2 = lo + inv_phi * (hi - lo);
  double f1 = or_S_eq_mf(solute, x1, T_amb, m_s), f2 = or_S_eq_mf(solute, x2, T_amb, m_s);
  for (int step = 0; step < 100; ++step) {
    if (f1 >= f2) {
      hi = x2;
      x2 = x1;
      f2 = f1;
      x1 = hi - inv_phi * (hi - lo);
      f1 = or_S_eq_mf(solute, x1, T_amb, m_s);
    } else {
      lo = x1;
      x1 = x2;
      f1 = f2;
      x2 = lo + inv_phi * (hi - lo);
      f2 = or_S_eq_mf(solute, x2, T_amb, m_s);
    }
  }
  const double w_act = 0.5 * (lo + hi) * 1.000001;
  if (S_amb > or_S_eq_mf(solute, w_act, T_amb, m_s)) return w_act;
  double a = w_act, b = w_s_max;
  for (int step = 0; step < 200; ++step) {
    const double c = 0.5 * (a + b);
    if (c <= a || c >= b) break;
    if (or_S_eq_mf(solute, c, T_amb, m_s) - S_amb > 0.0)
      a = c;
    else
      b = c;
  }
  const double c = 0.5 * (a + b);
  return c <= a ? a : c;
}

EXPORT double or_S_eq_mass_fraction(int solute, double w_s, double T_p, double m_s) {
  or_init_consts();
  return or_S_eq_mf(solute, w_s, T_p, m_s);
}

/*
 * generation_SIP_ensemble.py:147-244 (gen_mass_ensemble_weights_SinSIP_lognormal) for one cell and mode,
 * the draw of bin b being lm_init_uniform(seed, level, b, cell << 4 | mode, 0) (1 for the weak threshold's
 * second draw).  Writes the kept masses and weights, returns their number.
 */
EXPORT int or_sinsip_cell(const OrInitPar *q, uint64_t seed, uint32_t level, uint32_t cell, int mode, double *masses,
                          double *weights) {
  or_init_consts();
  const double r = q->r_critmin[mode];
  const double m_low = K.four_pi_over_three * q->mass_density_dry * r * r * r; /* microphysics.py:61-80 */
  const double bin_factor = OR_POW(10.0, 1.0 / q->kappa[mode]);
  const int l_max = q->l_max[mode];
  const double two_pi_sqrt = 2.5066282746310002; /* distributions.py:69 */
  double wbuf[4096], mbuf[4096];
  double m_left = m_low;
  for (int bin = 0; bin < l_max; ++bin) {
    const double m_right = m_left * bin_factor;
    const double bin_width = m_right - m_left;
    const double mu = m_left + lm_init_uniform(seed, level, (uint32_t)bin, (cell << 4) | (uint32_t)mode, 0) * bin_width;
    /* distributions.py:75-96 */
    const double t = (OR_LOG(mu) - q->mu_m_log[mode]) / q->sigma_m_log[mode];
    wbuf[bin] = OR_EXP(-0.5 * (t * t)) / (mu * two_pi_sqrt * q->sigma_m_log[mode]) * bin_width;
    mbuf[bin] = mu;
    m_left = m_right;
  }
  double weight_max = wbuf[0];
  for (int bin = 1; bin < l_max; ++bin)
    if (wbuf[bin] > weight_max) weight_max = wbuf[bin];
  const double weight_critmin = weight_max * q->eta;
  int n = 0;
  for (int bin = 0; bin < l_max; ++bin) {
    double w = wbuf[bin];
    int valid = 1;
    if (w < weight_critmin) {
      valid = 0;
      if (q->weak_threshold &&
          lm_init_uniform(seed, level, (uint32_t)bin, (cell << 4) | (uint32_t)mode, 1) < w / weight_critmin) {
        w = weight_critmin;
        valid = 1;
      }
    }
    if (valid) {
      masses[n] = mbuf[bin];
      weights[n] = w;
      ++n;
    }
  }
  return n;
}

/*
 * init.py:439-1236 for one run: the SIP arrays in the reference's order (level, cell, mode, bin) and
 * positions (init.py:322-381, draws lm_init_uniform(seed, 0, n, 0, 2 / 3) for the run's n-th SIP), the
 * level state after the adjustment (per level 21 doubles in the product's order: p_bot, p_top, T_bot,
 * T_top, T_avg, p_avg, rho_dry_avg, S_avg, r_v_avg, r_l_avg, m_air, m_vap, m_liq, m_part, scale factor,
 * iterations, last relative change, real particles per cell of each of 4 modes), and per cell r_v, r_l,
 * e_s, S (4 x nx x nz).  Returns the number of SIPs, or -1 if cap is too small.
 */
EXPORT int64_t or_init_SinSIP(const OrInitPar *q, uint64_t seed, int32_t nz, const double *levels_z, double p_0,
                              const double *corner_x, int64_t cap, double *m_s, double *xi, double *m_w, int32_t *mode,
                              int32_t *cell_x, int32_t *cell_z, double *x, double *z, double *rel, double *lv,
                              double *fields) {
  or_init_consts();
  const int nx = q->nx, no_modes = q->n_modes;
  const double r_tot = q->r_tot_0, dy = q->dy, dz = q->dz;
  const double kappa_tot_inv = 1.0 / (K.kappa_dry * (1 - 0.289 * r_tot));
  double *r_l_cell = fields + (size_t)nx * nz, *r_v_cell = fields, *e_s_cell = fields + (size_t)2 * nx * nz,
         *S_cell = fields + (size_t)3 * nx * nz;
  double p_bot = p_0;
  int64_t n = 0;
  double mb[4096], wb[4096];
  int64_t *cell_start = (int64_t *)malloc(sizeof(int64_t) * (size_t)(nx + 1));
  if (!cell_start) return -1;
  for (int j = 0; j < nz; ++j) {
    const double z_bot = levels_z[j], z_top = levels_z[j + 1];
    double T_bot, rho_dry_bot, S_bot, T_top, p_top, rho_dry_top, S_top;
    or_profile_dry(z_bot, z_bot, p_bot, q->p_ref, q->Theta_l, r_tot, &T_bot, &p_bot, &rho_dry_bot, &S_bot);
    or_profile_dry(z_top, z_bot, p_bot, q->p_ref, q->Theta_l, r_tot, &T_top, &p_top, &rho_dry_top, &S_top);
    double p_avg = (p_bot + p_top) * 0.5, T_avg = (T_bot + T_top) * 0.5;
    double rho_dry_avg = (rho_dry_bot + rho_dry_top) * 0.5, S_avg = (S_bot + S_top) * 0.5, r_v_avg = r_tot;
    double mass_air_dry_level = q->X * dy * (p_bot - p_top) / (OR_G * (1 + r_tot));
    double mass_water_vapor_level = r_v_avg * mass_air_dry_level;
    double mass_water_liquid_level = 0.0, mass_particles_level = 0.0, dm_l_level = 0.0, dm_p_level = 0.0;
    const double no_rpcm_scale_factor = rho_dry_avg / q->rho_dry_0;
    double no_rpcm[4] = {0, 0, 0, 0};
    for (int m = 0; m < no_modes; ++m) no_rpcm[m] = rint(q->no_rpcm_0[m] * no_rpcm_scale_factor);
    /* generation_SIP_ensemble.py:330-404: cells of the level, modes of the cell */
    const int64_t lvl0 = n;
    for (int i = 0; i < nx; ++i) {
      cell_start[i] = n;
      for (int m = 0; m < no_modes; ++m) {
        const int k = or_sinsip_cell(q, seed, (uint32_t)j, (uint32_t)i, m, mb, wb);
        if (n + k > cap) {
          free(cell_start);
          return -1;
        }
        for (int b = 0; b < k; ++b) {
          m_s[n] = mb[b];
          xi[n] = wb[b] * no_rpcm[m];
          mode[n] = m;
          cell_x[n] = i;
          cell_z[n] = j;
          ++n;
        }
      }
    }
    cell_start[nx] = n;
    /* init.py:787-800: initial water from the Köhler curve */
    for (int64_t k = lvl0; k < n; ++k) {
      const double w_s = or_initial_mass_fraction(q->solute, m_s[k], S_avg, T_avg);
      const double m_p = m_s[k] / w_s;
      m_w[k] = m_p - m_s[k];
      dm_l_level += m_w[k] * xi[k];
      dm_p_level += m_p * xi[k];
    }
    /* init.py:810-862 */
    dm_l_level *= 1.0E-18;
    dm_p_level *= 1.0E-18;
    mass_water_liquid_level += dm_l_level;
    mass_particles_level += dm_p_level;
    mass_water_vapor_level -= dm_l_level;
    double heat_capacity_level = mass_air_dry_level * K.c_pd + mass_water_vapor_level * 1906.0 +
                                 mass_particles_level * 4183.8; /* constants.py:101-111 */
    double dT_avg = dm_l_level * heat_of_vaporization(T_avg) / heat_capacity_level;
    T_avg += dT_avg;
    T_bot += dT_avg;
    T_top += dT_avg;
    r_v_avg = mass_water_vapor_level / mass_air_dry_level;
    p_top = p_bot * OR_POW(T_top / T_bot, ((1 + r_tot) * OR_G * dz) /
                                              ((1 + r_v_avg / K.eps_gc) * K.R_d * (T_bot - T_top)));
    p_avg = 0.5 * (p_bot + p_top);
    rho_dry_avg = (p_bot - p_top) / (dz * (1 + r_tot) * OR_G);
    mass_air_dry_level = q->X * dy * dz * rho_dry_avg;
    double r_l_avg = mass_water_liquid_level / mass_air_dry_level;
    r_v_avg = r_tot - r_l_avg;
    mass_water_vapor_level = r_v_avg * mass_air_dry_level;
    double e_s_avg = saturation_pressure_vapor_liquid(T_avg);
    S_avg = rho_dry_avg * r_v_avg * T_avg * K.R_v / e_s_avg;
    for (int i = 0; i < nx; ++i) {
      const size_t c = (size_t)i * nz + j;
      r_v_cell[c] = r_v_avg;
      r_l_cell[c] = r_l_avg;
      e_s_cell[c] = e_s_avg;
      S_cell[c] = S_avg;
    }
    /* init.py:893-1008: saturation adjustment */
    dm_l_level = mass_water_liquid_level;
    int iter_cnt = 0;
    while (fabs(dm_l_level / mass_water_liquid_level) > 1e-5 && iter_cnt < q->iter_cnt_limit) {
      dm_l_level = 0.0;
      dm_p_level = 0.0;
      const double D_v = diffusion_constant(T_avg, p_avg);
      const double Kc = thermal_conductivity_air(T_avg);
      const double L_v = heat_of_vaporization(T_avg);
      for (int i = 0; i < nx; ++i) {
        const size_t c = (size_t)i * nz + j;
        const double e_s = e_s_cell[c];
        const double S2 = S_cell[c] < q->S_init_max ? S_cell[c] : q->S_init_max;
        const double sigma_p_cell = surface_tension_water(T_avg);
        double sum = 0.0;
        for (int64_t k = cell_start[i]; k < cell_start[i + 1]; ++k) {
          double R_p, w_s, rho_p;
          R_p_w_s_rho_p(m_w[k], m_s[k], T_avg, q->solute, &R_p, &w_s, &rho_p);
          const double dm = or_dml_impl_Newton(q->solute, q->dt_init, q->newton_iterations, m_w[k], m_s[k], w_s, R_p,
                                               T_avg, rho_p, T_avg, p_avg, S2, e_s, L_v, Kc, D_v, sigma_p_cell,
                                               NULL);
          m_w[k] += dm;
          sum += dm * xi[k];
        }
        dm_l_level += sum;
        dm_p_level += sum;
      }
      dm_l_level *= 1.0E-18;
      dm_p_level *= 1.0E-18;
      mass_water_liquid_level += dm_l_level;
      mass_particles_level += dm_p_level;
      mass_water_vapor_level -= dm_l_level;
      heat_capacity_level =
          mass_air_dry_level * K.c_pd + mass_water_vapor_level * 1906.0 + mass_particles_level * 4183.8;
      dT_avg = dm_l_level * heat_of_vaporization(T_avg) / heat_capacity_level;
      T_avg += dT_avg;
      T_bot += dT_avg;
      T_top += dT_avg;
      r_v_avg = mass_water_vapor_level / mass_air_dry_level;
      p_top = p_bot * OR_POW(T_top / T_bot, kappa_tot_inv * (K.eps_gc + r_tot) / (K.eps_gc + r_v_avg));
      p_avg = 0.5 * (p_bot + p_top);
      rho_dry_avg = (p_bot - p_top) / (dz * (1 + r_tot) * OR_G);
      mass_air_dry_level = q->X * dy * dz * rho_dry_avg;
      const double mass_air_dry_cell = q->dx * dy * dz * rho_dry_avg;
      r_l_avg = mass_water_liquid_level / mass_air_dry_level;
      r_v_avg = r_tot - r_l_avg;
      const double e_s_T = saturation_pressure_vapor_liquid(T_avg);
      for (int i = 0; i < nx; ++i) {
        const size_t c = (size_t)i * nz + j;
        double s = 0.0;
        for (int64_t k = cell_start[i]; k < cell_start[i + 1]; ++k) s += m_w[k] * xi[k];
        r_l_cell[c] = s * (1.0E-18 / mass_air_dry_cell);
        r_v_cell[c] = r_tot - r_l_cell[c];
        e_s_cell[c] = e_s_T;
        S_cell[c] = rho_dry_avg * r_v_cell[c] * T_avg * K.R_v / e_s_cell[c];
      }
      mass_water_vapor_level = r_v_avg * mass_air_dry_level;
      e_s_avg = saturation_pressure_vapor_liquid(T_avg);
      S_avg = rho_dry_avg * r_v_avg * T_avg * K.R_v / e_s_avg;
      ++iter_cnt;
    }
    double *L = lv + (size_t)j * 21;
    const double st[17] = {p_bot, p_top, T_bot, T_top, T_avg, p_avg, rho_dry_avg, S_avg, r_v_avg, r_l_avg,
                           mass_air_dry_level, mass_water_vapor_level, mass_water_liquid_level, mass_particles_level,
                           no_rpcm_scale_factor, (double)iter_cnt, dm_l_level / mass_water_liquid_level};
    for (int e = 0; e < 17; ++e) L[e] = st[e];
    for (int m = 0; m < 4; ++m) L[17 + m] = no_rpcm[m];
    p_bot = p_top; /* init.py:1025 */
  }
  free(cell_start);
  /* init.py:322-381, 1224-1226 */
  for (int64_t k = 0; k < n; ++k) {
    const double u = lm_init_uniform(seed, 0, (uint32_t)k, 0, 2), w = lm_init_uniform(seed, 0, (uint32_t)k, 0, 3);
    x[k] = corner_x[cell_x[k]] + q->dx * u;
    z[k] = levels_z[cell_z[k]] + dz * w;
    rel[k] = u;
    rel[cap + k] = w;
  }
  return n;
}
