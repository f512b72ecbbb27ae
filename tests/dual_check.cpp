// Host check of the adjoint's local Jacobian: the condensation Newton of
// lm_physics.h on LmDual<5> (lm_dual.h) against the double evaluation (value
// part, bit for bit) and against central differences (derivative part).
// Built and run by tests/test_host_cpu.py::test_dual_newton_jacobian.
//
// stdin: n_const doubles of LmPhys, then n, then n rows (m_w, m_s, T, p, S).
// stdout per row: value_equal max_rel_err_of_5_derivatives
#include <cstdio>
#include <cstring>
#include <initializer_list>

#include "lm_dual.h"
#include "lm_physics.h"

template <class R>
static void cell_constants(const LmPhys &P, R T, R p, R S, R *cc) {
  const R K = 4.1868E-3 * (5.69 + 0.017 * (T - 273.15));
  const R D_v = 4.01218E-5 * lm_pow(T, 1.94) / p;
  const R L_v = 1.0E3 * (2500.82 - 2.358 * (T - 273.0));
  const R e_s = 2.53E11 * lm_exp(-5420.0 / T);
  cc[CC_T] = T;
  cc[CC_P] = p;
  cc[CC_S] = S;
  cc[CC_LALPHA] = lmp_l_alpha(P, T, p, K);
  cc[CC_LBETA] = lmp_l_beta(P, T, D_v);
  cc[CC_C1] = L_v * L_v / (P.R_v * K * T * T);
  cc[CC_C2] = P.R_v * T / (D_v * e_s);
  cc[CC_RHOW] = lmp_density_water(P, T);
  cc[CC_SIGW] = lmp_sigma_water(T);
}

template <class R>
static R newton(const LmPhys &P, double dt, R m_w, R m_s, R T, R p, R S) {
  R cc[CC_N];
  cell_constants(P, T, p, S, cc);
  R Rp, ws, rho;
  lmp_R_ws_rho(P, 0, m_w, m_s, cc[CC_T], cc[CC_RHOW], Rp, ws, rho);
  const R sig = lmp_sigma_solution(P, 0, ws, cc[CC_SIGW]);
  return lmp_newton_dm(P, 0, dt, 3, m_w, m_s, ws, Rp, cc[CC_T], rho, cc, sig);
}

int main() {
  LmPhys P;
  const int nconst = (int)(sizeof(LmPhys) / sizeof(double));
  double *pp = (double *)&P;
  for (int i = 0; i < nconst; ++i)
    if (scanf("%lf", &pp[i]) != 1) return 2;
  int n = 0;
  if (scanf("%d", &n) != 1) return 2;
  const double dt = 0.1;
  for (int r = 0; r < n; ++r) {
    double x[5];
    for (int q = 0; q < 5; ++q)
      if (scanf("%lf", &x[q]) != 1) return 2;
    const double d0 = newton<double>(P, dt, x[0], x[1], x[2], x[3], x[4]);
    using D5 = LmDual<5>;
    const D5 dd = newton<D5>(P, dt, D5::seed(x[0], 0), D5::seed(x[1], 1), D5::seed(x[2], 2), D5::seed(x[3], 3),
                             D5::seed(x[4], 4));
    double worst = 0.0;
    for (int q = 0; q < 5; ++q) {
      /* three difference quotients: fourth order with h = 1e-4 x (small truncation error), second order with
         h = 1e-6 x and 1e-8 x (where the Newton map bends sharply or a clamp kinks within 2e-4 x); the
         closest one counts */
      double fd4, fd2[2];
      {
        const double h = 1.0E-4 * x[q];
        double f[4];
        const double st[4] = {2.0, 1.0, -1.0, -2.0};
        for (int k = 0; k < 4; ++k) {
          double xs[5];
          memcpy(xs, x, sizeof x);
          xs[q] += st[k] * h;
          f[k] = newton<double>(P, dt, xs[0], xs[1], xs[2], xs[3], xs[4]);
        }
        fd4 = (8.0 * (f[1] - f[2]) - (f[0] - f[3])) / (12.0 * h);
      }
      for (int k = 0; k < 2; ++k) {
        const double h = (k == 0 ? 1.0E-6 : 1.0E-8) * x[q];
        double xp[5], xm[5];
        memcpy(xp, x, sizeof x);
        memcpy(xm, x, sizeof x);
        xp[q] += h;
        xm[q] -= h;
        fd2[k] = (newton<double>(P, dt, xp[0], xp[1], xp[2], xp[3], xp[4]) -
                  newton<double>(P, dt, xm[0], xm[1], xm[2], xm[3], xm[4])) /
                 (xp[q] - xm[q]);
      }
      double best = 1.0E300;
      for (const double fd : {fd4, fd2[0], fd2[1]}) {
        /* scale: the derivative of m_w' = m_w + dm, whose size sets what matters downstream */
        const double scale = (q == 0 ? 1.0 : 0.0) + (fd < 0 ? -fd : fd) + 1.0E-300;
        const double err = (dd.d[q] - fd) / scale;
        const double ae = err < 0 ? -err : err;
        if (ae < best) best = ae;
      }
      if (best > worst) worst = best;
    }
    printf("%d %.3e\n", dd.v == d0 ? 1 : 0, worst);
  }
  return 0;
}
