/*
 * lm_init.h -- initial SIP ensemble and saturation adjustment of the 2-D case on the GPU
 * (init.py:439-1236, generation_SIP_ensemble.py:147-404, microphysics.py:686-966), level by
 * level as the reference does (the pressure at a level's bottom is the adjusted pressure at the
 * top of the level below).  Included by lacmo_kernels.hip.
 *
 * Per run and level, the kernels run:
 *   k_init_profile  the level's dry profile (init.py:384-437), level masses and the SIP count scale;
 *   k_sinsip        the SinSIP masses and weights of every cell and mode (count pass, fill pass);
 *   k_init_mass     each SIP's initial solute mass fraction from the Köhler curve, its water mass;
 *   k_satadj        the level's saturation adjustment: implicit Newton condensation steps of all
 *                   its SIPs with the level's thermodynamics updated after each, until the
 *                   liquid water changes by less than 1e-5 relative (one workgroup per run).
 * Uniform draws come from Philox-4x32-10 (lm_init_uniform) instead of NumPy's MT19937.
 */
#ifndef LM_INIT_H
#define LM_INIT_H

/* per-run level state (LV_*), written by k_init_profile and k_satadj */
enum {
  LV_PBOT = 0, LV_PTOP, LV_TBOT, LV_TTOP, LV_TAVG, LV_PAVG, LV_RHO, LV_SAVG, LV_RV, LV_RL,
  LV_MAIR, LV_MVAP, LV_MLIQ, LV_MPART, LV_SCALE, LV_ITER, LV_DMLREL,
  LV_NRPCM, /* LV_NRPCM + m: SIPs' real-particle count per cell of mode m at this level */
  LV_N = LV_NRPCM + LM_INIT_MAX_MODES
};
static_assert(LV_N == LM_INIT_LV_N, "level state width of the C ABI");

#define LMI __device__ static inline

/* init.py:384-437 compute_profiles_T_p_rhod_S_without_liquid with atmosphere.py:77-166 */
LMI void lmi_profile(const LmPhys &P, double z, double z_0, double p_0, double p_ref, double Theta_l, double r_tot,
                     double &T, double &p, double &rho_dry, double &S) {
  const double g = 9.80665;
  const double p_0_over_p_ref = p_0 / p_ref;
  const double kappa_tot = P.kappa_dry * (1 - 0.289 * r_tot);
  const double kappa_tot_inv = 1.0 / kappa_tot;
  const double p0k = lm_pow(p_0_over_p_ref, kappa_tot);
  const double beta_tot = g * (P.kappa_dry * (1 - 0.289 * r_tot)) * (1 + r_tot) /
                          (P.R_d * Theta_l * (1 + r_tot / P.eps_gc));
  const double T_over_Theta_l = p0k - beta_tot * (z - z_0);
  T = T_over_Theta_l * Theta_l;
  const double p_over_p_ref = lm_pow(T_over_Theta_l, kappa_tot_inv);
  p = p_over_p_ref * p_ref;
  rho_dry = p * beta_tot / (T_over_Theta_l * g * kappa_tot);
  S = rho_dry * r_tot * T * P.R_v / (2.53E11 * lm_exp(-5420.0 / T));
}

/* Köhler equilibrium saturation from the mass fraction (microphysics.py:570-610, 531-569; kelvin term :400-430) */
template <int SOLUTE>
LMI double lmi_seq_mf(const LmPhys &P, double w, double T, double m_s) {
  double rho, sigma, a_w;
  const double sw = lmp_sigma_water(T);
  if (SOLUTE == 0) {
    rho = lmp_density_water(P, T) * lmp_poly(P.par_rho_AS, 4, w);
    sigma = sw * (1.0 + P.par_sigma_AS * w / (1. - w));
    a_w = lmp_poly(P.par_wat_act_AS, 5, w);
  } else {
    rho = lmp_density_NaCl(P, w, T);
    sigma = sw * (1.0 + P.par_sigma_NaCl * w / (1. - w));
    a_w = lmp_poly(P.par_wat_act_NaCl, 5, w);
  }
  const double volume_to_radius = lm_pow(P.four_pi_over_three_inv, P.one_third);
  return a_w * lm_exp(2.0 * sigma * 1.0E6 * lm_pow(w / m_s, P.one_third) /
                      (P.R_v * T * lm_pow(rho, 0.66666667) * volume_to_radius));
}

/*
 * Initial solute mass fraction (microphysics.py:686-785 NaCl, :866-966 AS): w_s_max if the ambient
 * saturation does not exceed S_eq(w_s_max); else the activation mass fraction (maximum of S_eq,
 * raised by 1e-6 relative) if the ambient saturation exceeds S_eq there; else the root of
 * S_eq(w_s) = S_amb between the two.  The maximum is found by 100 golden-section steps and the root
 * by bisection to the last bit (the reference calls scipy's fminbound and brentq).
 */
template <int SOLUTE>
LMI double lmi_initial_mass_fraction(const LmPhys &P, double m_s, double S_amb, double T) {
  const double w_max = SOLUTE == 0 ? 0.78 : 0.45;
  const double S_min = lmi_seq_mf<SOLUTE>(P, w_max, T, m_s);
  if (S_amb <= S_min) return w_max;
  const double phi = 0.6180339887498949;
  double a = 1.0E-8, b = w_max;
  double c = b - phi * (b - a), d = a + phi * (b - a);
  double fc = lmi_seq_mf<SOLUTE>(P, c, T, m_s), fd = lmi_seq_mf<SOLUTE>(P, d, T, m_s);
  for (int it = 0; it < 100; ++it) {
    if (fc >= fd) { /* maximum in [a, d] */
      b = d;
      d = c;
      fd = fc;
      c = b - phi * (b - a);
      fc = lmi_seq_mf<SOLUTE>(P, c, T, m_s);
    } else {
      a = c;
      c = d;
      fc = fd;
      d = a + phi * (b - a);
      fd = lmi_seq_mf<SOLUTE>(P, d, T, m_s);
    }
  }
  const double w_act = 0.5 * (a + b) * 1.000001;
  const double S_act = lmi_seq_mf<SOLUTE>(P, w_act, T, m_s);
  if (S_amb > S_act) return w_act;
  /* S_eq(w_act) >= S_amb >= S_eq(w_max), S_eq decreasing on [w_act, w_max] */
  double lo = w_act, hi = w_max;
  for (int it = 0; it < 200; ++it) {
    const double mid = 0.5 * (lo + hi);
    if (mid <= lo || mid >= hi) break;
    if (lmi_seq_mf<SOLUTE>(P, mid, T, m_s) - S_amb > 0.0)
      lo = mid;
    else
      hi = mid;
  }
  return 0.5 * (lo + hi) <= lo ? lo : 0.5 * (lo + hi);
}

/* generation_SIP_ensemble.py:225-226: the draw of bin b of the SinSIP scheme of (level, cell, mode) of a run */
LMI double lmi_bin_uniform(uint64_t seed, uint32_t level, uint32_t cell, uint32_t mode, uint32_t b, uint32_t stream) {
  return lm_init_uniform(seed, level, b, (cell << 4) | mode, stream);
}

/*
 * SinSIP masses and weights of one cell and mode (generation_SIP_ensemble.py:147-244): l_max
 * logarithmic bins from m_low = 4/3 pi rho_dry r_critmin^3 with factor 10^(1/kappa), one mass uniform
 * in each bin, weight = lognormal pdf(mass) * bin width, bins below eta * max weight dropped (or, with
 * the weak threshold, kept at eta * max weight with probability weight / (eta * max weight)).
 * fill == false: returns the number of SIPs; else writes them (mass, weight) and returns the number.
 */
LMI int lmi_sinsip_cell(const LmPhys &P, const LmInitPar &q, uint64_t seed, uint32_t level, uint32_t cell, int m,
                        bool fill, double *mass_out, double *weight_out) {
  const double r = q.r_critmin[m];
  const double m_low = 4.0 * P.pi / 3.0 * q.mass_density_dry * r * r * r;
  const double bin_factor = lm_pow(10.0, 1.0 / q.kappa[m]);
  const int l_max = q.l_max[m];
  const double two_pi_sqrt = 2.5066282746310002;
  /* pass 1: the largest weight */
  double w_max = 0.0;
  double m_left = m_low;
  for (int b = 0; b < l_max; ++b) {
    const double m_right = m_left * bin_factor;
    const double width = m_right - m_left;
    const double mu = m_left + lmi_bin_uniform(seed, level, cell, (uint32_t)m, (uint32_t)b, 0) * width;
    const double t = (lm_log(mu) - q.mu_m_log[m]) / q.sigma_m_log[m];
    const double wgt = lm_exp(-0.5 * (t * t)) / (mu * two_pi_sqrt * q.sigma_m_log[m]) * width;
    w_max = b == 0 ? wgt : (wgt > w_max ? wgt : w_max);
    m_left = m_right;
  }
  const double w_crit = w_max * q.eta;
  /* pass 2: the kept bins */
  int n = 0;
  m_left = m_low;
  for (int b = 0; b < l_max; ++b) {
    const double m_right = m_left * bin_factor;
    const double width = m_right - m_left;
    const double mu = m_left + lmi_bin_uniform(seed, level, cell, (uint32_t)m, (uint32_t)b, 0) * width;
    const double t = (lm_log(mu) - q.mu_m_log[m]) / q.sigma_m_log[m];
    double wgt = lm_exp(-0.5 * (t * t)) / (mu * two_pi_sqrt * q.sigma_m_log[m]) * width;
    bool keep = true;
    if (wgt < w_crit) {
      keep = false;
      if (q.weak_threshold &&
          lmi_bin_uniform(seed, level, cell, (uint32_t)m, (uint32_t)b, 1) < wgt / w_crit) {
        wgt = w_crit;
        keep = true;
      }
    }
    if (keep) {
      if (fill) {
        mass_out[n] = mu;
        weight_out[n] = wgt;
      }
      ++n;
    }
    m_left = m_right;
  }
  return n;
}

/*
 * The level's dry profile and masses before any water is on the particles (init.py:697-740): from
 * p_bot (the adjusted top pressure of the level below) the bottom/top values of T, p, rho_dry, S,
 * their averages, the dry-air and vapour masses of the level, and the real-particle counts per cell
 * and mode, rint(ceil(V N_m) rho_dry / rho_dry_0).
 */
LMI void lmi_level_profile(const LmPhys &P, const LmInitPar &q, double z_bot, double z_top, double p_bot, double *L) {
  double T_bot, rho_bot, S_bot, T_top, p_top, rho_top, S_top;
  lmi_profile(P, z_bot, z_bot, p_bot, q.p_ref, q.Theta_l, q.r_tot_0, T_bot, p_bot, rho_bot, S_bot);
  lmi_profile(P, z_top, z_bot, p_bot, q.p_ref, q.Theta_l, q.r_tot_0, T_top, p_top, rho_top, S_top);
  L[LV_PBOT] = p_bot;
  L[LV_PTOP] = p_top;
  L[LV_TBOT] = T_bot;
  L[LV_TTOP] = T_top;
  L[LV_PAVG] = (p_bot + p_top) * 0.5;
  L[LV_TAVG] = (T_bot + T_top) * 0.5;
  L[LV_RHO] = (rho_bot + rho_top) * 0.5;
  L[LV_SAVG] = (S_bot + S_top) * 0.5;
  L[LV_RV] = q.r_tot_0;
  L[LV_RL] = 0.0;
  L[LV_MAIR] = q.X * q.dy * (p_bot - p_top) / (9.80665 * (1 + q.r_tot_0));
  L[LV_MVAP] = q.r_tot_0 * L[LV_MAIR];
  L[LV_MLIQ] = 0.0;
  L[LV_MPART] = 0.0;
  const double scale = L[LV_RHO] / q.rho_dry_0;
  L[LV_SCALE] = scale;
  for (int m = 0; m < q.n_modes; ++m) L[LV_NRPCM + m] = lm_rint(q.no_rpcm_0[m] * scale);
  L[LV_ITER] = 0.0;
  L[LV_DMLREL] = 0.0;
}

/*
 * Level thermodynamics after the liquid water changed by dm_l (kg; dm_p for the particles):
 * init.py:769-800 (first == 1, after the SIPs got their initial water) and :930-976 (the adjustment
 * iterations).
 */
LMI void lmi_level_update(const LmPhys &P, const LmInitPar &q, double dm_l, double dm_p, int first, double *L) {
  const double g = 9.80665, c_pd = 1004.71, c_pv = 1906.0, c_w = 4183.8;
  const double r_tot = q.r_tot_0;
  L[LV_MLIQ] += dm_l;
  L[LV_MPART] += dm_p;
  L[LV_MVAP] -= dm_l;
  const double heat_capacity = L[LV_MAIR] * c_pd + L[LV_MVAP] * c_pv + L[LV_MPART] * c_w;
  const double L_v = 1.0E3 * (2500.82 - 2.358 * (L[LV_TAVG] - 273.0));
  const double dT = dm_l * L_v / heat_capacity;
  L[LV_TAVG] += dT;
  L[LV_TBOT] += dT;
  L[LV_TTOP] += dT;
  double r_v_avg = L[LV_MVAP] / L[LV_MAIR];
  if (first) {
    L[LV_PTOP] = L[LV_PBOT] * lm_pow(L[LV_TTOP] / L[LV_TBOT],
                                     ((1 + r_tot) * g * q.dz) /
                                         ((1 + r_v_avg / P.eps_gc) * P.R_d * (L[LV_TBOT] - L[LV_TTOP])));
  } else {
    const double kappa_tot_inv = 1.0 / (P.kappa_dry * (1 - 0.289 * r_tot));
    L[LV_PTOP] =
        L[LV_PBOT] * lm_pow(L[LV_TTOP] / L[LV_TBOT], kappa_tot_inv * (P.eps_gc + r_tot) / (P.eps_gc + r_v_avg));
  }
  L[LV_PAVG] = 0.5 * (L[LV_PBOT] + L[LV_PTOP]);
  L[LV_RHO] = (L[LV_PBOT] - L[LV_PTOP]) / (q.dz * (1 + r_tot) * g);
  L[LV_MAIR] = q.X * q.dy * q.dz * L[LV_RHO];
  L[LV_RL] = L[LV_MLIQ] / L[LV_MAIR];
  r_v_avg = r_tot - L[LV_RL];
  L[LV_RV] = r_v_avg;
  L[LV_MVAP] = r_v_avg * L[LV_MAIR];
  const double e_s = 2.53E11 * lm_exp(-5420.0 / L[LV_TAVG]);
  L[LV_SAVG] = L[LV_RHO] * r_v_avg * L[LV_TAVG] * P.R_v / e_s;
}

/* the cell constants of the adjustment's Newton step (integration.py:728-836 called from init.py:908-918) */
LMI void lmi_newton_constants(const LmPhys &P, double T, double p, double e_s, double S2, double *C) {
  const double K = 4.1868E-3 * (5.69 + 0.017 * (T - 273.15));
  const double D_v = 4.01218E-5 * lm_pow(T, 1.94) / p;
  const double L_v = 1.0E3 * (2500.82 - 2.358 * (T - 273.0));
  C[CC_T] = T;
  C[CC_P] = p;
  C[CC_S] = S2;
  C[CC_ES] = e_s;
  C[CC_LV] = L_v;
  C[CC_K] = K;
  C[CC_DV] = D_v;
  C[CC_LALPHA] = lmp_l_alpha(P, T, p, K);
  C[CC_LBETA] = lmp_l_beta(P, T, D_v);
  C[CC_C1] = L_v * L_v / (P.R_v * K * T * T);
  C[CC_C2] = P.R_v * T / (D_v * e_s);
  C[CC_RHOW] = lmp_density_water(P, T);
  C[CC_SIGW] = lmp_sigma_water(T);
}

/* ------------------------------------------------------------------------ */
/* kernels                                                                  */
/* ------------------------------------------------------------------------ */
/* one thread per run: the level's dry profile from the bottom pressure pbot[run] */
__global__ void k_init_profile(LmInitPar q, double z_bot, double z_top, const double *pbot, double *L) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= q.n_runs) return;
  lmi_level_profile(c_P, q, z_bot, z_top, pbot[r], L + (size_t)r * LV_N);
}

/*
 * one thread per (run, cell, mode), t = (run * nx + cell) * n_modes + mode: count (off == NULL) or
 * write the SinSIPs of the level at off[t] (m_s, xi = weight * real particles per cell, mode, cell)
 */
__global__ void k_sinsip(LmInitPar q, const uint64_t *seeds, int level, const double *L, const long long *off,
                         int32_t *count, double *m_s, double *xi, int32_t *mode, int32_t *cell, int32_t *run) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int nm = q.n_modes;
  if (t >= q.n_runs * q.nx * nm) return;
  const int m = t % nm, i = (t / nm) % q.nx, r = t / (nm * q.nx);
  if (!off) {
    count[t] = lmi_sinsip_cell(c_P, q, seeds[r], (uint32_t)level, (uint32_t)i, m, false, nullptr, nullptr);
    return;
  }
  const long long o = off[t];
  const int n = lmi_sinsip_cell(c_P, q, seeds[r], (uint32_t)level, (uint32_t)i, m, true, m_s + o, xi + o);
  const double nr = L[(size_t)r * LV_N + LV_NRPCM + m];
  for (int k = 0; k < n; ++k) {
    xi[o + k] = xi[o + k] * nr;
    mode[o + k] = m;
    cell[o + k] = i;
    run[o + k] = r;
  }
}

/* one thread per SIP of the level: initial water mass in equilibrium with the level's S and T (init.py:750-761) */
template <int SOLUTE>
__global__ void k_init_mass(long long n, const int32_t *run, const double *L, const double *m_s, double *m_w,
                            double *m_p) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const double *Lr = L + (size_t)run[k] * LV_N;
  const double w = lmi_initial_mass_fraction<SOLUTE>(c_P, m_s[k], Lr[LV_SAVG], Lr[LV_TAVG]);
  const double mp = m_s[k] / w;
  m_p[k] = mp;
  m_w[k] = mp - m_s[k];
}

#define SATADJ_BLOCK 512
#define SATADJ_MAX_NX 512
/*
 * Saturation adjustment of one level (init.py:763-1010), one workgroup per run.  SIPs of run r are
 * [cell_off[r*nx], cell_off[(r+1)*nx]) in cell order; per cell i the range cell_off[r*nx+i ..+1].
 * Sums run over each cell's SIPs in order, then over the cells.  Writes the level state L, the next
 * level's bottom pressure, and per (run, cell) r_v, r_l, e_s, S.
 */
template <int SOLUTE>
__global__ __launch_bounds__(SATADJ_BLOCK) void k_satadj(LmInitPar q, const long long *cell_off, const int32_t *cellx,
                                                       const double *m_s, const double *xi, const double *m_p,
                                                       double *m_w, double *dmxi, double *Lg, double *pbot_next,
                                                       double *fields) {
  const LmPhys &P = c_P;
  const int r = blockIdx.x, nx = q.nx, tid = threadIdx.x;
  /* per cell: saturation, e_s, min(S, S_init_max), c2 of the rate, sums, r_l */
  __shared__ double sL[LV_N], sC[CC_N], sS[SATADJ_MAX_NX], sES[SATADJ_MAX_NX], sSm[SATADJ_MAX_NX];
  __shared__ double sC2[SATADJ_MAX_NX], sSum[SATADJ_MAX_NX], sRl[SATADJ_MAX_NX];
  __shared__ int sGo;
  const long long k0 = cell_off[(size_t)r * nx], k1 = cell_off[(size_t)r * nx + nx];
  const long long *co = cell_off + (size_t)r * nx;
  if (tid < LV_N) sL[tid] = Lg[(size_t)r * LV_N + tid];
  __syncthreads();
  if (tid == 0) {
    /* the initial water on the particles (init.py:763-800) */
    double dm_l = 0.0, dm_p = 0.0;
    for (long long k = k0; k < k1; ++k) {
      dm_l += m_w[k] * xi[k];
      dm_p += m_p[k] * xi[k];
    }
    dm_l *= 1.0E-18;
    dm_p *= 1.0E-18;
    lmi_level_update(P, q, dm_l, dm_p, 1, sL);
    const double e_s = 2.53E11 * lm_exp(-5420.0 / sL[LV_TAVG]);
    for (int i = 0; i < nx; ++i) {
      sES[i] = e_s;
      sS[i] = sL[LV_SAVG];
      sRl[i] = sL[LV_RL];
    }
    /* dm_l_level = mass_water_liquid_level before the first iteration: 1, or NaN (no iteration) without liquid */
    sL[LV_DMLREL] = sL[LV_MLIQ] / sL[LV_MLIQ];
    sL[LV_ITER] = 0.0;
    sGo = (fabs(sL[LV_DMLREL]) > 1.0E-5) && (0 < q.iter_cnt_limit);
  }
  __syncthreads();
  int iter = 0;
  while (sGo) {
    /* the iteration's constants (init.py:885-918): level T, p; per cell e_s and min(S, S_init_max) */
    if (tid == 0) lmi_newton_constants(P, sL[LV_TAVG], sL[LV_PAVG], sES[0], 1.0, sC);
    __syncthreads();
    for (int i = tid; i < nx; i += blockDim.x) {
      const double S = sS[i];
      sC2[i] = P.R_v * sC[CC_T] / (sC[CC_DV] * sES[i]);
      sSm[i] = S < q.S_init_max ? S : q.S_init_max;
    }
    __syncthreads();
    for (long long k = k0 + tid; k < k1; k += blockDim.x) {
      const int i = cellx[k];
      double C[CC_N];
#pragma unroll
      for (int e = 0; e < CC_N; ++e) C[e] = sC[e];
      C[CC_S] = sSm[i];
      C[CC_ES] = sES[i];
      C[CC_C2] = sC2[i];
      double R, ws, rho;
      const double mw = m_w[k], ms = m_s[k];
      lmp_R_ws_rho(P, SOLUTE, mw, ms, C[CC_T], C[CC_RHOW], R, ws, rho);
      const double dm =
          lmp_newton_dm(P, SOLUTE, q.dt_init, q.newton_iterations, mw, ms, ws, R, C[CC_T], rho, C, C[CC_SIGW]);
      m_w[k] = mw + dm;
      dmxi[k] = dm * xi[k];
    }
    __syncthreads();
    for (int i = tid; i < nx; i += blockDim.x) {
      double s = 0.0;
      for (long long k = co[i]; k < co[i + 1]; ++k) s += dmxi[k];
      sSum[i] = s;
    }
    __syncthreads();
    if (tid == 0) {
      double dm_l = 0.0;
      for (int i = 0; i < nx; ++i) dm_l += sSum[i];
      dm_l *= 1.0E-18;
      lmi_level_update(P, q, dm_l, dm_l, 0, sL);
      sL[LV_DMLREL] = dm_l / sL[LV_MLIQ];
    }
    __syncthreads();
    /* per cell r_l, r_v, e_s, S of the level (init.py:957-972) */
    const double m_air_cell = q.dx * q.dy * q.dz * sL[LV_RHO];
    const double e_s = 2.53E11 * lm_exp(-5420.0 / sL[LV_TAVG]);
    for (int i = tid; i < nx; i += blockDim.x) {
      double s = 0.0;
      for (long long k = co[i]; k < co[i + 1]; ++k) s += m_w[k] * xi[k];
      const double r_l = s * (1.0E-18 / m_air_cell);
      const double r_v = q.r_tot_0 - r_l;
      sRl[i] = r_l;
      sES[i] = e_s;
      sS[i] = sL[LV_RHO] * r_v * sL[LV_TAVG] * P.R_v / e_s;
    }
    ++iter;
    __syncthreads();
    if (tid == 0) {
      sL[LV_ITER] = (double)iter;
      sGo = (fabs(sL[LV_DMLREL]) > 1.0E-5) && (iter < q.iter_cnt_limit);
    }
    __syncthreads();
  }
  for (int i = tid; i < nx; i += blockDim.x) {
    double *F = fields + (size_t)r * 4 * nx;
    F[i] = q.r_tot_0 - sRl[i];
    F[nx + i] = sRl[i];
    F[2 * nx + i] = sES[i];
    F[3 * nx + i] = sS[i];
  }
  if (tid < LV_N) Lg[(size_t)r * LV_N + tid] = sL[tid];
  if (tid == 0) pbot_next[r] = sL[LV_PTOP];
}

/*
 * positions uniform in each SIP's cell (init.py:322-381): x = x_corner + dx u, z = z_corner + dz u' with the
 * relative position (u, u') kept in rel (2 x n), from the run's seed and the SIP's index idx within its run
 */
__global__ void k_init_positions(long long n, const uint64_t *seeds, const int32_t *run, const int32_t *idx,
                                 const int32_t *ci, const int32_t *cj, const double *corner_x, const double *corner_z,
                                 double dx, double dz, double *x, double *z, double *rel) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const uint64_t seed = seeds[run[k]];
  const uint32_t a = (uint32_t)idx[k];
  const double u = lm_init_uniform(seed, 0u, a, 0u, 2u), w = lm_init_uniform(seed, 0u, a, 0u, 3u);
  x[k] = corner_x[ci[k]] + dx * u;
  z[k] = corner_z[cj[k]] + dz * w;
  rel[k] = u;
  rel[n + k] = w;
}

#endif /* LM_INIT_H */
