/*
 * lm_adjoint.h -- kernels of the adjoint window (included by lacmo_kernels.hip).
 *
 * Forward side: the event log of a window (record, order, replay) and the
 * tape of a segment (cell-constant derivatives, grid-substep Jacobians,
 * increment fields).  Reverse side: one kernel per forward operation of the
 * advection step (integration.py:1298-1611), each the transpose of the
 * operation's Jacobian at the taped state.
 *
 * Adjoint variables: per global SIP id lam_mw, lam_ms, lam_xi; per cell
 * lam_cell[8][n_runs*nc] = Theta, r_v, T, p, S (the state after the last grid
 * substep), dml (of the substep being reversed), dTheta_adv, dr_v_adv (the
 * increments of the subloop being reversed).
 */
#ifndef LM_ADJOINT_H
#define LM_ADJOINT_H

enum { LA_TH = 0, LA_RV, LA_T, LA_P, LA_S, LA_DML, LA_DTH, LA_DRV, LA_N };

/* ------------------------------------------------------------------------ */
/* event log                                                                */
/* ------------------------------------------------------------------------ */
/*
 * Offsets of the recorded collision step: the records of this step occupy
 * [b1 - total, b1) of the log (b1 = *ev_cnt after the step); off[gc] =
 * b0 + sum of the counts of the cells before gc.  One block.
 */
#define EV_SCAN_BLOCK 1024
__global__ __launch_bounds__(EV_SCAN_BLOCK) void k_ev_scan(const int32_t *count, int ncell, const int32_t *ev_cnt,
                                                         int32_t *off) {
  __shared__ long long part[EV_SCAN_BLOCK];
  __shared__ long long total;
  const int t = threadIdx.x;
  const int chunk = (ncell + EV_SCAN_BLOCK - 1) / EV_SCAN_BLOCK;
  const int lo = t * chunk < ncell ? t * chunk : ncell, hi = lo + chunk < ncell ? lo + chunk : ncell;
  long long s = 0;
  for (int g = lo; g < hi; ++g) s += count[g];
  part[t] = s;
  __syncthreads();
  if (t == 0) {
    long long acc = 0;
    for (int q = 0; q < EV_SCAN_BLOCK; ++q) {
      const long long x = part[q];
      part[q] = acc;
      acc += x;
    }
    total = acc;
  }
  __syncthreads();
  const long long b1 = *ev_cnt;
  long long acc = b1 - total + part[t];
  for (int g = lo; g < hi; ++g) {
    off[g] = (int32_t)acc;
    acc += count[g];
  }
  if (t == 0) off[ncell] = (int32_t)b1;
}

/* ev_order[off[cell] + seq] = record index, for the records of the step just recorded (grid-strided) */
__global__ void k_ev_order(const LmEvent *ev, int cap, const int32_t *off, int ncell, int32_t *order) {
  const int b0 = off[0], b1 = off[ncell] < cap ? off[ncell] : cap;
  for (int e = b0 + blockIdx.x * blockDim.x + threadIdx.x; e < b1; e += gridDim.x * blockDim.x) {
    const LmEvent &E = ev[e];
    const long long pos = (long long)off[E.cell] + E.seq;
    if (pos >= 0 && pos < cap) order[pos] = e;
  }
}

/* slot_of[gid] = slot, for every slot of the current buffer */
__global__ void k_slot_of(const int32_t *id, const long long *run_base, int n_runs, long long n_tot, int32_t *slot_of) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n_tot) return;
  const int r = slot_run(run_base, n_runs, k);
  slot_of[run_base[r] + id[k]] = (int32_t)k;
}

/*
 * Replays the recorded events of one collision step as fixed events: per
 * cell, in recorded order, the branch of all_or_nothing.py:577-632 the
 * event took, applied to the current state (p_crit of a multiple collection
 * as recorded).  Same expressions as apply_event.
 */
__global__ void k_ev_replay(KView v, const int32_t *off, const int32_t *order, int ncell) {
  const int gc = blockIdx.x * blockDim.x + threadIdx.x;
  if (gc >= ncell) return;
  const int q0 = off[gc], q1 = off[gc + 1];
  if (q0 >= q1) return;
  long long c_single = 0, c_multi = 0, c_eq = 0;
  for (int q = q0; q < q1; ++q) {
    const LmEvent E = v.ev[order[q]];
    const int kmin = v.slot_of[E.gid_min], kmax = v.slot_of[E.gid_max];
    const double xi_min = v.xi[kmin], xi_max = v.xi[kmax];
    const double mw_min = v.m_w[kmin], mw_max = v.m_w[kmax];
    const double ms_min = v.m_s[kmin], ms_max = v.m_s[kmax];
    if (E.kind == 1) {
      const double xi_col = E.p * xi_min;
      v.m_w[kmin] = mw_min + E.p * mw_max;
      v.m_s[kmin] = ms_min + E.p * ms_max;
      v.xi[kmax] = xi_max - xi_col;
      ++c_multi;
    } else if (E.kind == 3) {
      const double xi_ges = xi_min + xi_max;
      const double mwn = 2.0 * (xi_min * mw_min + xi_max * mw_max) / xi_ges;
      const double msn = 2.0 * (xi_min * ms_min + xi_max * ms_max) / xi_ges;
      v.m_w[kmin] = mwn;
      v.m_s[kmin] = msn;
      v.m_w[kmax] = mwn;
      v.m_s[kmax] = msn;
      const double xmax_new = 0.5 * 0.7 * xi_ges;
      v.xi[kmax] = xmax_new;
      v.xi[kmin] = 0.5 * xi_ges - xmax_new;
      ++c_single;
      ++c_eq;
    } else {
      v.m_w[kmin] = mw_min + mw_max;
      v.m_s[kmin] = ms_min + ms_max;
      v.xi[kmax] = xi_max - xi_min;
      ++c_single;
    }
  }
  const int r = gc / v.nc;
  if (c_single) atomicAdd((unsigned long long *)&v.no_cols[2 * r], (unsigned long long)c_single);
  if (c_multi) atomicAdd((unsigned long long *)&v.no_cols[2 * r + 1], (unsigned long long)c_multi);
  if (c_eq) atomicAdd((unsigned long long *)&v.n_equal_xi[r], (unsigned long long)c_eq);
}

/* ------------------------------------------------------------------------ */
/* tape                                                                     */
/* ------------------------------------------------------------------------ */
/*
 * d(l_alpha, l_beta, c1, c2, rho_w, sigma_w)/d(T, p) of each cell's constants
 * (the material laws of the grid kernel and fill_cc) at the cell's T, p.
 */
__global__ void k_ccd(KView v) {
  const LmPhys &P = c_P;
  const int gc = blockIdx.x * blockDim.x + threadIdx.x;
  if (gc >= v.n_runs * v.nc) return;
  using D2 = LmDual<2>;
  const double *C = v.cc + (size_t)gc * CC_N;
  const D2 T = D2::seed(C[CC_T], 0), p = D2::seed(C[CC_P], 1);
  const D2 K = 4.1868E-3 * (5.69 + 0.017 * (T - 273.15));
  const D2 D_v = 4.01218E-5 * lm_pow(T, 1.94) / p;
  const D2 L_v = 1.0E3 * (2500.82 - 2.358 * (T - 273.0));
  const D2 e_s = 2.53E11 * lm_exp(-5420.0 / T);
  D2 out[6];
  out[0] = lmp_l_alpha(P, T, p, K);
  out[1] = lmp_l_beta(P, T, D_v);
  out[2] = L_v * L_v / (P.R_v * K * T * T);
  out[3] = P.R_v * T / (D_v * e_s);
  out[4] = lmp_density_water(P, T);
  out[5] = lmp_sigma_water(T);
  double *G = v.ccd + (size_t)gc * 12;
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    G[2 * q] = out[q].d[0];
    G[2 * q + 1] = out[q].d[1];
  }
}

/* the grid substep of one cell (integration.py:402-483), as k_grid evaluates it, over any value type */
template <class R>
__device__ static void grid_map(const LmPhys &P, const KView &v, double rho_d, double mdinv, R Th0, R rv0, R L_v, R dml,
                                R dTh, R dRv, R &Th, R &r_v, R &T, R &p, R &Sat) {
  Th = Th0 + dTh;
  r_v = rv0;
  r_v += dRv - dml * mdinv;
  const R Theta_over_T = lm_pow(v.p_ref_inv * P.R_d * rho_d * Th, P.kappa_factor2);
  Th += Theta_over_T * (L_v * dml) / (P.c_pd * rho_d * v.V_cell * (1.0 + r_v * P.heat_factor_r_v));
  const R pdr = lm_pow(rho_d * Th * P.R_d * v.p_ref_inv, P.kappa_factor);
  T = Th * lm_pow(pdr, P.kappa_dry);
  p = pdr * v.p_ref * (1 + r_v / P.eps_gc);
  const R e_s = 2.53E11 * lm_exp(-5420.0 / T);
  Sat = (rho_d * r_v) * T * P.R_v / e_s;
}

/*
 * Jacobian of the grid substep of each cell at its current state, written
 * before k_grid runs: J[(o*6+i)*ncell + gc] = d out_o / d in_i, out = (Theta',
 * r_v', T', p', S'), in = (Theta, r_v, T (through L_v(T) of the previous
 * substep), dml, dTheta_adv, dr_v_adv).
 */
__global__ void k_grid_jac(KView v, double *J) {
  const LmPhys &P = c_P;
  const int ncell = v.n_runs * v.nc;
  const int gc = blockIdx.x * blockDim.x + threadIdx.x;
  if (gc >= ncell) return;
  using D6 = LmDual<6>;
  const int nc = v.nc, r = gc / nc, c = gc % nc;
  const double *S = v.scal + (size_t)r * 10 * nc;
  const double *M = v.mat + (size_t)r * 7 * nc;
  D6 L_v(M[2 * nc + c]);
  L_v.d[2] = 1.0E3 * -2.358; /* L_v = 1e3 (2500.82 - 2.358 (T - 273)) */
  D6 Th, r_v, T, p, Sat;
  grid_map(P, v, S[3 * nc + c], S[8 * nc + c], D6::seed(S[2 * nc + c], 0), D6::seed(S[4 * nc + c], 1), L_v,
           D6::seed(v.dml[gc], 3), D6::seed(v.dTh[gc], 4), D6::seed(v.dRv[gc], 5), Th, r_v, T, p, Sat);
  const D6 *o[5] = {&Th, &r_v, &T, &p, &Sat};
#pragma unroll
  for (int q = 0; q < 5; ++q)
#pragma unroll
    for (int i = 0; i < 6; ++i) J[(size_t)(q * 6 + i) * ncell + gc] = o[q]->d[i];
}

/* r_v and Theta of every cell at an increment evaluation */
__global__ void k_tape_fields(KView v, double *F) {
  const int ncell = v.n_runs * v.nc;
  const int gc = blockIdx.x * blockDim.x + threadIdx.x;
  if (gc >= ncell) return;
  const int nc = v.nc, r = gc / nc, c = gc % nc;
  const double *S = v.scal + (size_t)r * 10 * nc;
  F[gc] = S[4 * nc + c];
  F[ncell + gc] = S[2 * nc + c];
}

/* ------------------------------------------------------------------------ */
/* reverse sweep                                                            */
/* ------------------------------------------------------------------------ */
/* grid substep: lam(in) = J^T lam(out) */
__global__ void k_adj_grid(int ncell, const double *J, double *lam) {
  const int gc = blockIdx.x * blockDim.x + threadIdx.x;
  if (gc >= ncell) return;
  double lo[5];
#pragma unroll
  for (int q = 0; q < 5; ++q) lo[q] = lam[(size_t)q * ncell + gc];
  double g[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    double s = 0.0;
#pragma unroll
    for (int q = 0; q < 5; ++q) s += J[(size_t)(q * 6 + i) * ncell + gc] * lo[q];
    g[i] = s;
  }
  lam[(size_t)LA_TH * ncell + gc] = g[0];
  lam[(size_t)LA_RV * ncell + gc] = g[1];
  lam[(size_t)LA_T * ncell + gc] = g[2];
  lam[(size_t)LA_P * ncell + gc] = 0.0;
  lam[(size_t)LA_S * ncell + gc] = 0.0;
  lam[(size_t)LA_DML * ncell + gc] = g[3];
  lam[(size_t)LA_DTH * ncell + gc] += g[4];
  lam[(size_t)LA_DRV * ncell + gc] += g[5];
}

/* into the slot order of the step being reversed (and the slot of each global id) */
__global__ void k_adj_to_slots(long long n_tot, const int32_t *gid_of_slot, LmAdjoint a) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n_tot) return;
  const int g = gid_of_slot[k];
  a.sl_mw[k] = a.lam_mw[g];
  a.sl_ms[k] = a.lam_ms[g];
  a.sl_xi[k] = a.lam_xi[g];
  a.sl_of[g] = (int32_t)k;
}

/* back to global-id order once the step is reversed */
__global__ void k_adj_to_gids(long long n_tot, const int32_t *gid_of_slot, LmAdjoint a) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n_tot) return;
  const int g = gid_of_slot[k];
  a.lam_mw[g] = a.sl_mw[k];
  a.lam_ms[g] = a.sl_ms[k];
  a.lam_xi[g] = a.sl_xi[k];
}

/*
 * Particle substep of the condensation update (integration.py:840-940):
 * m_w' = m_w + dm(m_w, m_s, T, p, S of the cell), dml = 1e-18 sum xi dm.
 */
__global__ __launch_bounds__(256) void k_adj_part(long long n_tot, int ncell, const double *tp, const int32_t *ti,
                                                 LmAdjoint a) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int gc = k < n_tot ? ti[n_tot + k] : -1;
  double yT = 0.0, yp = 0.0, yS = 0.0;
  if (gc >= 0) {
    const double J0 = tp[k], J1 = tp[n_tot + k], J2 = tp[2 * n_tot + k], J3 = tp[3 * n_tot + k],
                 J4 = tp[4 * n_tot + k], dm = tp[5 * n_tot + k], xi = tp[6 * n_tot + k];
    const double ld = a.lam_cell[(size_t)LA_DML * ncell + gc] * 1.0E-18;
    const double y = a.sl_mw[k] + xi * ld;
    a.sl_mw[k] = y * J0 - xi * ld;
    a.sl_ms[k] += y * J1;
    a.sl_xi[k] += ld * dm;
    yT = y * J2;
    yp = y * J3;
    yS = y * J4;
  }
  /* cell sums: a segmented reduction over runs of equal cells in the wave (slots are sorted by cell), then one
     atomic per run */
  const int lane = (int)(threadIdx.x & (WAVE - 1));
  const int nxt = __shfl_down(gc, 1);
  const unsigned long long bnd = __ballot(lane == WAVE - 1 || nxt != gc);
  const int end = lane + __builtin_ctzll(bnd >> lane);
  for (int off = 1; off < WAVE; off <<= 1) {
    const double t0 = __shfl_down(yT, off), t1 = __shfl_down(yp, off), t2 = __shfl_down(yS, off);
    if (lane + off <= end) {
      yT += t0;
      yp += t1;
      yS += t2;
    }
  }
  const int prv = __shfl_up(gc, 1);
  if ((lane == 0 || prv != gc) && gc >= 0) {
    atomicAdd(&a.lam_cell[(size_t)LA_T * ncell + gc], yT);
    atomicAdd(&a.lam_cell[(size_t)LA_P * ncell + gc], yp);
    atomicAdd(&a.lam_cell[(size_t)LA_S * ncell + gc], yS);
  }
}

/*
 * Advective increments (integration.py:1483-1496, 1536-1549; divergence
 * :150-286, periodic x, fixed z): d = -mult dt_sub rho_inv div(field flux)
 * (+ the relaxation term, k_adj_inc_rows).  One thread per face: the face
 * flux enters the divergence of the cell on either side; its derivative in
 * the four stencil values (limiter included) scatters the adjoint to them.
 * Faces: x faces fi = 0..nx of every row, then z faces fj = 1..nz-1 of every
 * column (the outer z faces carry no flux).
 */
__global__ void k_adj_inc_faces(KView v, const double *F, double mult, double dt_sub, double *lam) {
  const LmPhys &P = c_P;
  const int nx = v.nx, nz = v.nz, nc = v.nc, nfz = nz + 1;
  const int ncell = v.n_runs * nc;
  const int nxf = (nx + 1) * nz, nzf = nx * (nz - 1);
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)v.n_runs * (nxf + nzf)) return;
  const int r = (int)(t / (nxf + nzf));
  const int f = (int)(t % (nxf + nzf));
  const double *S = v.scal + (size_t)r * 10 * nc;
  const double *FU = v.flux, *FW = v.flux + (size_t)(nx + 1) * nfz;
  int cl, cr, st[4];
  double u, h;
  if (f < nxf) {
    const int fi = f / nz, j = f % nz;
    cl = fi >= 1 ? (fi - 1) * nz + j : -1;
    cr = fi <= nx - 1 ? fi * nz + j : -1;
    st[0] = bc_index(fi, nx, 0) * nz + j;
    st[1] = bc_index(fi - 1, nx, 0) * nz + j;
    st[2] = bc_index(fi - 2, nx, 0) * nz + j;
    st[3] = bc_index(fi + 1, nx, 0) * nz + j;
    u = FU[fi * nfz + j];
    h = v.dx;
  } else {
    const int g = f - nxf;
    const int i = g / (nz - 1), fj = 1 + g % (nz - 1);
    cl = i * nz + fj - 1;
    cr = i * nz + fj;
    st[0] = i * nz + bc_index(fj, nz, 1);
    st[1] = i * nz + bc_index(fj - 1, nz, 1);
    st[2] = i * nz + bc_index(fj - 2, nz, 1);
    st[3] = i * nz + bc_index(fj + 1, nz, 1);
    u = FW[i * nfz + fj];
    h = v.dz;
  }
  const int base = r * nc;
  /* adjoint of each cell's divergence: a_c = -mult dt_sub rho_inv lam_d */
  double lf[2];
  const int ld_idx[2] = {LA_DRV, LA_DTH};
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    double g = 0.0;
    if (cl >= 0) g += -mult * dt_sub * S[9 * nc + cl] * lam[(size_t)ld_idx[q] * ncell + base + cl];
    if (cr >= 0) g -= -mult * dt_sub * S[9 * nc + cr] * lam[(size_t)ld_idx[q] * ncell + base + cr];
    lf[q] = g / h;
  }
  using D4 = LmDual<4>;
  const int fld_idx[2] = {LA_RV, LA_TH};
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    if (lf[q] == 0.0) continue;
    const double *fld = F + (size_t)q * ncell + base;
    const D4 Fx = face_flux(D4::seed(fld[st[0]], 0), D4::seed(fld[st[1]], 1), D4::seed(fld[st[2]], 2),
                            D4::seed(fld[st[3]], 3), u, P.one_third);
#pragma unroll
    for (int s = 0; s < 4; ++s)
      if (Fx.d[s] != 0.0) atomicAdd(&lam[(size_t)fld_idx[q] * ncell + base + st[s]], lf[q] * Fx.d[s]);
  }
}

/*
 * Relaxation part of the increments (integration.py:325-354 column means;
 * mult 1 or 2), then the increments' adjoint moves on: after the second
 * evaluation (d2 = ... - d1) it becomes -lam_d for the first subloop, after
 * the first it is cleared.  One wavefront per (run, level).
 */
__global__ __launch_bounds__(WAVE) void k_adj_inc_rows(KView v, double mult, double dt_sub, int act_relax,
                                                      int second, double *lam) {
  const int t = blockIdx.x;
  if (t >= v.n_runs * v.nz) return;
  const int r = t / v.nz, j = t % v.nz, nc = v.nc, nz = v.nz, nx = v.nx;
  const int ncell = v.n_runs * nc;
  double *lrv = lam + (size_t)LA_RV * ncell + r * nc, *lth = lam + (size_t)LA_TH * ncell + r * nc;
  double *drv = lam + (size_t)LA_DRV * ncell + r * nc, *dth = lam + (size_t)LA_DTH * ncell + r * nc;
  if (act_relax) {
    double srv = 0.0, sth = 0.0;
    for (int i = threadIdx.x; i < nx; i += WAVE) {
      srv += drv[i * nz + j];
      sth += dth[i * nz + j];
    }
    for (int off = WAVE / 2; off > 0; off >>= 1) {
      srv += __shfl_xor(srv, off);
      sth += __shfl_xor(sth, off);
    }
    const double cf = mult * (-dt_sub / (double)nx / v.t_relax[j]);
    for (int i = threadIdx.x; i < nx; i += WAVE) {
      lrv[i * nz + j] += cf * srv;
      lth[i * nz + j] += cf * sth;
    }
  }
  for (int i = threadIdx.x; i < nx; i += WAVE) {
    drv[i * nz + j] = second ? -drv[i * nz + j] : 0.0;
    dth[i * nz + j] = second ? -dth[i * nz + j] : 0.0;
  }
}

/* recorded events of one collision step, reversed per cell (all_or_nothing.py:577-632 as fixed events) */
__global__ void k_adj_events(const LmEvent *ev, const int32_t *off, const int32_t *order, int ncell, LmAdjoint a) {
  const int gc = blockIdx.x * blockDim.x + threadIdx.x;
  if (gc >= ncell) return;
  const int q0 = off[gc], q1 = off[gc + 1];
  for (int q = q1 - 1; q >= q0; --q) {
    const LmEvent E = ev[order[q]];
    const int i = a.sl_of[E.gid_min], j = a.sl_of[E.gid_max];
    if (E.kind == 1) {
      a.sl_xi[i] -= E.p * a.sl_xi[j];
      a.sl_mw[j] += E.p * a.sl_mw[i];
      a.sl_ms[j] += E.p * a.sl_ms[i];
    } else if (E.kind == 2) {
      a.sl_xi[i] -= a.sl_xi[j];
      a.sl_mw[j] += a.sl_mw[i];
      a.sl_ms[j] += a.sl_ms[i];
    } else {
      /* m' = 2 (xi_i m_i + xi_j m_j) / G for both, xi_j' = 0.35 G, xi_i' = 0.5 G - xi_j', G = xi_i + xi_j */
      const double G = E.xi_min + E.xi_max;
      const double mw = 2.0 * (E.xi_min * E.mw_min + E.xi_max * E.mw_max) / G;
      const double ms = 2.0 * (E.xi_min * E.ms_min + E.xi_max * E.ms_max) / G;
      const double A = a.sl_mw[i] + a.sl_mw[j], B = a.sl_ms[i] + a.sl_ms[j];
      const double LG = 0.5 * 0.7 * a.sl_xi[j] + (0.5 - 0.5 * 0.7) * a.sl_xi[i];
      a.sl_xi[i] = A * (2.0 * E.mw_min - mw) / G + B * (2.0 * E.ms_min - ms) / G + LG;
      a.sl_xi[j] = A * (2.0 * E.mw_max - mw) / G + B * (2.0 * E.ms_max - ms) / G + LG;
      a.sl_mw[i] = A * 2.0 * E.xi_min / G;
      a.sl_mw[j] = A * 2.0 * E.xi_max / G;
      a.sl_ms[i] = B * 2.0 * E.xi_min / G;
      a.sl_ms[j] = B * 2.0 * E.xi_max / G;
    }
  }
}

/* ------------------------------------------------------------------------ */
/* objectives                                                               */
/* ------------------------------------------------------------------------ */
/* smooth cloud-class indicator s((ln R - ln 0.5)/w) s((ln 25 - ln R)/w) */
template <class R>
__device__ static inline R cloud_indicator(R Rp, double w) {
  const double ln_half = -0x1.62e42fefa39efp-1, ln_25 = 3.2188758248682006;
  const R lr = lm_log(Rp);
  const R a = (lr - ln_half) / w;
  const R b = (ln_25 - lr) / w;
  return 1.0 / (1.0 + lm_exp(-a)) * (1.0 / (1.0 + lm_exp(-b)));
}

/*
 * per cell, ascending id: sum xi m_w and sum xi I(R) (radius at the cell temperature), each as a compensated
 * sum (value and the error term of Knuth's TwoSum carried along): out[q*ncell + gc] = sum q, out[(2+q)*ncell
 * + gc] = its correction.  Finite differences of the objectives then see the change of the state, not the
 * rounding of the sums.
 */
__global__ __launch_bounds__(WAVE) void k_objective(KView v, double width, double *out) {
  const LmPhys &P = c_P;
  const int nc = v.nc;
  const int gc = blockIdx.x, r = gc / nc;
  const int ncell = v.n_runs * nc;
  __shared__ int sSlot[DML_CAP], sId[DML_CAP], sRank[DML_CAP];
  __shared__ unsigned char sSub[DML_CAP];
  __shared__ double sA[DML_CAP], sB[DML_CAP];
  __shared__ NbrTab tb;
  nbr_table(gc, nc, v.nx, v.nz, v.seg + (size_t)r * (nc + 1), v.away + (size_t)r * nc, &tb);
  const int n = gather_members(gc, v.cell, v.id, &tb, sSlot, sId, sSub, DML_CAP);
  const double *C = v.cc + (size_t)gc * CC_N;
  if (n > DML_CAP) {
    /* over the LDS capacity: one lane runs both compensated sums straight from the segments, ascending id */
    if (threadIdx.x == 0) {
      double h0 = 0.0, l0 = 0.0, h1 = 0.0, l1 = 0.0;
      merge_members(gc, v.cell, v.id, &tb, [&](int k) {
        double R, ws, rho;
        lmp_R_ws_rho(P, v.solute, v.m_w[k], v.m_s[k], C[CC_T], C[CC_RHOW], R, ws, rho);
        double t, e;
        lm_two_sum(h0, v.xi[k] * v.m_w[k], &t, &e);
        h0 = t;
        l0 += e;
        lm_two_sum(h1, v.xi[k] * cloud_indicator(R, width), &t, &e);
        h1 = t;
        l1 += e;
      });
      out[gc] = h0;
      out[(size_t)ncell + gc] = h1;
      out[(size_t)2 * ncell + gc] = l0;
      out[(size_t)3 * ncell + gc] = l1;
    }
    return;
  }
  __syncthreads();
  rank_members(n, sId, sSub, &tb, sRank);
  __syncthreads();
  for (int m = threadIdx.x; m < n; m += WAVE) {
    const int k = sSlot[m];
    double R, ws, rho;
    lmp_R_ws_rho(P, v.solute, v.m_w[k], v.m_s[k], C[CC_T], C[CC_RHOW], R, ws, rho);
    sA[sRank[m]] = v.xi[k] * v.m_w[k];
    sB[sRank[m]] = v.xi[k] * cloud_indicator(R, width);
  }
  __syncthreads();
  if (threadIdx.x < 2) {
    const double *src = threadIdx.x == 0 ? sA : sB;
    double hi = 0.0, lo = 0.0;
    for (int q = 0; q < n; ++q) {
      double s, e;
      lm_two_sum(hi, src[q], &s, &e);
      hi = s;
      lo += e;
    }
    out[(size_t)threadIdx.x * ncell + gc] = hi;
    out[(size_t)(2 + threadIdx.x) * ncell + gc] = lo;
  }
}

/*
 * Seeds the adjoint at the end of the window: per active SIP the derivatives
 * of w_lwp c_lwp xi m_w + w_nc c_nc(cell) xi I(R(m_w, m_s, T)), the T part
 * summed into the cell's lam_T.  The adjoint arrays are cleared beforehand.
 */
__global__ __launch_bounds__(256) void k_adj_seed(KView v, double w_nc, double w_lwp, double width, LmAdjoint a) {
  const LmPhys &P = c_P;
  const long long n_tot = v.run_base[v.n_runs];
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n_tot) return;
  const int gc = v.cell[k];
  if (gc < 0) return;
  const int nc = v.nc, r = gc / nc, c = gc % nc;
  const int ncell = v.n_runs * nc;
  const int gid = (int)(v.run_base[r] + v.id[k]);
  using D3 = LmDual<3>;
  const double *C = v.cc + (size_t)gc * CC_N;
  const D3 mw = D3::seed(v.m_w[k], 0), ms = D3::seed(v.m_s[k], 1), T = D3::seed(C[CC_T], 2);
  D3 rw = lmp_density_water(P, T);
  rw.v = C[CC_RHOW];
  D3 R, ws, rho;
  lmp_R_ws_rho(P, v.solute, mw, ms, T, rw, R, ws, rho);
  const D3 I = cloud_indicator(R, width);
  const double xi = v.xi[k], m_w = v.m_w[k];
  const double c_lwp = 1.0E-18 / ((v.x_max - v.x_min) * (v.V_cell / (v.dx * v.dz)));
  const double c_nc = 1.0E-6 * v.scal[((size_t)r * 10 + 8) * nc + c] / (double)nc;
  a.lam_mw[gid] = w_lwp * c_lwp * xi + w_nc * c_nc * xi * I.d[0];
  a.lam_ms[gid] = w_nc * c_nc * xi * I.d[1];
  a.lam_xi[gid] = w_lwp * c_lwp * m_w + w_nc * c_nc * I.v;
  atomicAdd(&a.lam_cell[(size_t)LA_T * ncell + gc], w_nc * c_nc * xi * I.d[2]);
}

#endif /* LM_ADJOINT_H */
