/*
 * lacmo_kernels.hip -- gfx950 kernels and C ABI of the super-droplet step.
 *
 * Layout and schedule (one GPU, FP64 throughout):
 *  - SIPs of all runs live in one structure-of-arrays sorted by
 *    (run, home cell, id) at the last reindex.  A full reindex runs once per
 *    advection step; between reindexes a SIP moves at most one cell, so the
 *    members of cell c are the SIPs of the 3x3 home segments around c whose
 *    current cell is c ("neighbourhood gather").  Ordering by id inside a
 *    cell is recovered by ranking in LDS.
 *  - Per-SIP work (condensation Newton, drag, position) is one thread per SIP
 *    over the home-ordered arrays (coalesced).  Per-cell work (ordered
 *    Delta m_l sums, collisions) is one 64-lane wavefront per cell; cell
 *    thermodynamics is one thread per cell.
 *  - Collisions use the anti-diagonal schedule: pairs (i,j) of the
 *    row-major sequential loop grouped by d = i + j are mutually independent
 *    and every dependence points to a smaller d, so processing d = 1..2N-3
 *    in order reproduces the sequential result exactly.
 * All kernels are compiled with -ffp-contract=off.
 */
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <string.h>

#include "../../include/lacmo_amd.h"
#include "lm_dual.h"
#include "lm_philox.h"
#include "lm_physics.h"

#define WAVE 64
#define DML_CAP 512 /* member capacity of the gather kernels (Delta m_l, reindex) */
#define OVF_GRID 1024 /* blocks of the grid-strided overflow passes */
#define REF_BIG_BLOCKS 4 /* workgroups (and workspace regions) of the reference-layout step's pass over large cells */
#define PART_BLOCK_MAX 512 /* largest block of the particle kernel */
#define TAPE_BLOCK 256     /* block of the taped particle kernel (its dual Newton needs the registers of one wave per SIMD) */

/* ------------------------------------------------------------------------ */
/* host-side constants (same expressions as the reference modules)          */
/* ------------------------------------------------------------------------ */
static LmPhys make_phys() {
  LmPhys P;
  memset(&P, 0, sizeof(P));
  const double pi = 3.141592653589793;
  P.pi = pi;
  P.one_third = 1.0 / 3.0;                /* constants.py:27 */
  P.four_pi_over_three_inv = 0.75 / pi;   /* constants.py:29 */
  P.R_v = 461.53;                         /* constants.py:90 */
  P.R_d = 287.06;                         /* constants.py:95 */
  P.c_pd = 1004.71;                       /* constants.py:109 */
  const double c_pv = 1906.0;             /* constants.py:114 */
  P.eps_gc = P.R_d / P.R_v;               /* atmosphere.py:99 */
  P.kappa_dry = P.R_d / P.c_pd;           /* atmosphere.py:74 */
  P.kappa_factor = 1.0 / (1.0 - P.kappa_dry);      /* atmosphere.py:370 */
  P.kappa_factor2 = -P.kappa_dry * P.kappa_factor; /* atmosphere.py:371 */
  P.heat_factor_r_v = (c_pv - P.R_v) / (P.c_pd - P.R_d); /* atmosphere.py:365 */
  const double adi_inv = 0.7142857142857143, T0 = 289.0; /* microphysics.py:971-1026 */
  P.T_alpha_0 = T0;
  P.c_alpha_1 = 1.0E6 * sqrt(2.0 * pi * P.R_d * T0) / (1.0 * (P.c_pd * adi_inv + 0.5 * P.R_d));
  P.c_alpha_2 = 0.5E6 * sqrt(2.0 * pi * P.R_d / T0) / (1.0 * (P.c_pd * adi_inv + 0.5 * P.R_d));
  P.c_beta_1 = 1.0E6 * sqrt(2.0 * pi * 18.015E-3 / (8.3144598 * T0)) / 0.0415;
  P.c_beta_2 = -0.5 / T0;
  P.w_s_max_AS_inv = 1. / 0.78;                        /* integration.py:724 */
  P.par_sigma_AS = 2.362E-3 / 132.14E-3 / 0.072;       /* material_properties.py:344 */
  P.par_sigma_NaCl = 1.62E-3 / 58.4428E-3 / 0.072;     /* material_properties.py:272 */
  const double rho_raw[4] = {1.024E1, -5.036E1, 592., 997.1}; /* material_properties.py:87 */
  for (int i = 0; i < 4; ++i) P.par_rho_AS[i] = rho_raw[i] / 997.1;
  for (int i = 0; i < 3; ++i) P.par_rho_deriv_AS[i] = P.par_rho_AS[i] * (double)(3 - i); /* microphysics.py:1355 */
  const double aw_as[5] = {1.412, -2.336, 3.113E-1, -2.715E-1, 1.0}; /* material_properties.py:430 */
  const double aw_nacl[5] = {1.518E1, -1.158E1, 8.642E-1, -6.366E-1, 1.0}; /* material_properties.py:400 */
  for (int i = 0; i < 5; ++i) {
    P.par_wat_act_AS[i] = aw_as[i];
    P.par_wat_act_NaCl[i] = aw_nacl[i];
  }
  for (int i = 0; i < 4; ++i) {
    P.par_wat_act_deriv_AS[i] = aw_as[i] * (double)(4 - i);
    P.par_wat_act_deriv_NaCl[i] = aw_nacl[i] * (double)(4 - i);
  }
  const double sd[6] = {7.619443952135e+02, 1.021264281453e+03, 1.828970151543e+00,
                        2.405352122804e+02, -1.080547892416e+00, -3.492805028749e-03};
  for (int i = 0; i < 6; ++i) P.par_sol_dens_NaCl[i] = sd[i]; /* material_properties.py:56 */
  P.par_water_dens[0] = 1.00013502e+03;                        /* material_properties.py:29 */
  P.par_water_dens[1] = -4.68112708e-03;
  P.par_water_dens[2] = 2.72389977e+02;
  return P;
}

static const LmPhys &phys() {
  static LmPhys P = make_phys();
  return P;
}

/* physics constants in constant memory (scalar loads, no per-thread copies) */
__constant__ LmPhys c_P;

/* uploads c_P once per device (before the first kernel launch on it) */
static int ensure_phys() {
  static unsigned done_mask = 0;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (dev < 32 && (done_mask & (1u << dev))) return 0;
  const LmPhys &P = phys();
  if (hipMemcpyToSymbol(HIP_SYMBOL(c_P), &P, sizeof(LmPhys)) != hipSuccess) return -1;
  if (dev < 32) done_mask |= (1u << dev);
  return 0;
}

static thread_local char g_errmsg[512] = "";

static int set_err(int code, const char *msg) {
  snprintf(g_errmsg, sizeof(g_errmsg), "%s", msg);
  return code;
}

#define HIP_TRY(expr)                                                                        \
  do {                                                                                       \
    hipError_t _e = (expr);                                                                  \
    if (_e != hipSuccess) {                                                                  \
      snprintf(g_errmsg, sizeof(g_errmsg), "%s: %s", #expr, hipGetErrorString(_e));          \
      return (int)_e;                                                                        \
    }                                                                                        \
  } while (0)

#define LAUNCH_CHECK()                                                                       \
  do {                                                                                       \
    hipError_t _e = hipGetLastError();                                                       \
    if (_e != hipSuccess) {                                                                  \
      snprintf(g_errmsg, sizeof(g_errmsg), "kernel launch: %s", hipGetErrorString(_e));      \
      return (int)_e;                                                                        \
    }                                                                                        \
  } while (0)

/* ------------------------------------------------------------------------ */
/* device helpers                                                           */
/* ------------------------------------------------------------------------ */
__device__ static inline void raise_err(int32_t *err, int code, int detail) {
  if (atomicCAS(&err[0], 0, code) == 0) err[1] = detail;
}

__device__ static inline unsigned long long lane_mask_lt() {
  const int lane = threadIdx.x & (WAVE - 1);
  return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

/* Segment table of the 3x3 neighbourhood of a cell, in LDS (x periodic, z bounded). */
struct NbrTab {
  int start[9];
  int pre[10]; /* prefix of segment lengths */
  int off[10]; /* members before sub-list t (after the gather) */
};

/* Lanes 0..8 fill the table; the caller synchronises afterwards. */
/* bit of the move from home cell (hx, hz) to the adjacent cell (cx, cz), x periodic */
__device__ static inline int dir_bit(int hx, int hz, int cx, int cz, int nx) {
  const int dxm = ((cx - hx) % nx + nx) % nx;
  const int ddx = dxm == 0 ? 0 : (dxm == 1 ? 1 : -1);
  return 1 << ((ddx + 1) * 3 + (cz - hz + 1));
}

/*
 * Segments that can hold members of cell gc: its own home segment, and each
 * neighbouring home segment whose SIPs have moved towards gc since the last
 * reindex (away mask; a superset is harmless, so bits are never cleared
 * between reindexes).  Segments that cannot hold members get length 0.
 */
__device__ static inline void nbr_table(int gc, int nc, int nx, int nz, const int32_t *seg_run,
                                        const int32_t *away_run, NbrTab *tb) {
  const int lane = threadIdx.x;
  const int c = gc % nc;
  const int i = c / nz, j = c % nz;
  if (lane < 9) {
    const int di = lane / 3 - 1, dj = lane % 3 - 1;
    int ii = i + di;
    if (ii < 0) ii += nx;
    if (ii >= nx) ii -= nx;
    const int jj = j + dj;
    /* duplicates of the periodic wrap when nx < 3 are dropped */
    const bool valid = jj >= 0 && jj < nz && !(nx == 1 && di != 0) && !(nx == 2 && di == 1);
    const int cc = ii * nz + jj;
    const bool scan = valid && (lane == 4 || (away_run[valid ? cc : 0] & dir_bit(ii, jj, i, j, nx)) != 0);
    const int s0 = scan ? seg_run[cc] : 0;
    tb->start[lane] = s0;
    tb->pre[lane + 1] = scan ? seg_run[cc + 1] - s0 : 0;
  }
  __syncthreads();
  if (lane == 0) {
    tb->pre[0] = 0;
    for (int t = 0; t < 9; ++t) tb->pre[t + 1] += tb->pre[t];
  }
  __syncthreads();
}

/*
 * Wavefront gather of the members of global cell gc: slots k of the 3x3 home
 * segments with cell[k] == gc.  All segment loads are issued together
 * (flat index over the concatenated segments, 4 per lane per round).  Stores
 * up to cap (slot, id, sub-list) triples in LDS in segment order, so each
 * sub-list is ascending in id.  Returns the member count (all lanes).
 */
__device__ static int gather_members(int gc, const int32_t *cell, const int32_t *ids, const NbrTab *tb, int *s_slot,
                                     int *s_id, unsigned char *s_sub, int cap) {
  const int lane = threadIdx.x;
  int pre[10];
#pragma unroll
  for (int t = 0; t < 10; ++t) pre[t] = tb->pre[t];
  const int T = pre[9];
  int cnt = 0;
  for (int base = 0; base < T; base += 4 * WAVE) {
    int kk[4];
    int tt[4];
    bool m[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int f = base + q * WAVE + lane;
      int t = 0, pt = 0;
#pragma unroll
      for (int u = 1; u < 9; ++u) {
        const bool ge = f >= pre[u];
        t += ge;
        pt = ge ? pre[u] : pt;
      }
      tt[q] = t;
      kk[q] = f < T ? tb->start[t] + (f - pt) : -1;
    }
    int cv[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) cv[q] = kk[q] >= 0 ? cell[kk[q]] : -1;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      m[q] = kk[q] >= 0 && cv[q] == gc;
      const unsigned long long bal = __ballot(m[q]);
      if (m[q]) {
        const int pos = cnt + __popcll(bal & lane_mask_lt());
        if (pos < cap) {
          s_slot[pos] = kk[q];
          s_id[pos] = ids[kk[q]];
          s_sub[pos] = (unsigned char)tt[q];
        }
      }
      cnt += __popcll(bal);
    }
  }
  return cnt;
}

/* Member count only (no LDS output). */
__device__ static int count_members(int gc, const int32_t *cell, const NbrTab *tb) {
  const int lane = threadIdx.x;
  int pre[10];
#pragma unroll
  for (int t = 0; t < 10; ++t) pre[t] = tb->pre[t];
  const int T = pre[9];
  int cnt = 0;
  for (int base = 0; base < T; base += 4 * WAVE) {
    int cv[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int f = base + q * WAVE + lane;
      int t = 0, pt = 0;
#pragma unroll
      for (int u = 1; u < 9; ++u) {
        const bool ge = f >= pre[u];
        t += ge;
        pt = ge ? pre[u] : pt;
      }
      cv[q] = f < T ? cell[tb->start[t] + (f - pt)] : -1;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) cnt += __popcll(__ballot(cv[q] == gc));
  }
  return cnt;
}

/*
 * Rank of every member by id (ids unique within a run).  Each sub-list is
 * ascending, so rank = index in its own sub-list + lower bounds in the
 * others (binary search): O(n log n) instead of O(n^2).
 */
__device__ static void rank_members(int n, const int *s_id, const unsigned char *s_sub, NbrTab *tb, int *s_rank) {
  const int lane = threadIdx.x;
  if (lane < 10) {
    /* off[t] = first member whose sub-list >= t (s_sub is non-decreasing) */
    int lo = 0, hi = n;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if ((int)s_sub[mid] < lane)
        lo = mid + 1;
      else
        hi = mid;
    }
    tb->off[lane] = lane == 9 ? n : lo;
  }
  __syncthreads();
  int off[10];
#pragma unroll
  for (int t = 0; t < 10; ++t) off[t] = tb->off[t];
  for (int m = lane; m < n; m += WAVE) {
    const int idm = s_id[m];
    const int tm = s_sub[m];
    int rk = 0;
#pragma unroll
    for (int u = 0; u < 9; ++u) {
      int lo = off[u], hi = off[u + 1];
      if (u == tm) {
        rk += m - off[u];
        continue;
      }
      const int l0 = lo;
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (s_id[mid] < idm)
          lo = mid + 1;
        else
          hi = mid;
      }
      rk += lo - l0;
    }
    s_rank[m] = rk;
  }
}

/*
 * Members of global cell gc in ascending id for a cell of any size -- the path of the rare cells over a pass's
 * LDS capacity.  The calling lane merges the members of the neighbour segments of tb: after the reindex each
 * home segment holds its SIPs in ascending id, so its members of gc form an ascending sub-sequence, and ids
 * are unique within a run.  Calls f(slot) once per member, in ascending id; O(9 n) work on one lane.  Returns
 * the member count.
 */
template <class F>
__device__ static int merge_members(int gc, const int32_t *cell, const int32_t *ids, const NbrTab *tb, F &&f) {
  int pos[9], end[9], hid[9];
#pragma unroll
  for (int t = 0; t < 9; ++t) {
    pos[t] = tb->start[t];
    end[t] = pos[t] + (tb->pre[t + 1] - tb->pre[t]);
    while (pos[t] < end[t] && cell[pos[t]] != gc) ++pos[t];
    hid[t] = pos[t] < end[t] ? ids[pos[t]] : 0x7fffffff;
  }
  int n = 0;
  for (;;) {
    int best = 0x7fffffff, bt = -1, slot = 0;
#pragma unroll
    for (int t = 0; t < 9; ++t)
      if (hid[t] < best) {
        best = hid[t];
        bt = t;
        slot = pos[t];
      }
    if (bt < 0) break;
    f(slot);
    ++n;
#pragma unroll
    for (int t = 0; t < 9; ++t)
      if (t == bt) {
        ++pos[t];
        while (pos[t] < end[t] && cell[pos[t]] != gc) ++pos[t];
        hid[t] = pos[t] < end[t] ? ids[pos[t]] : 0x7fffffff;
      }
  }
  return n;
}

__global__ void k_fill_i32(int32_t *a, int32_t value, int n) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) a[t] = value;
}

__global__ void k_add_u32(uint32_t *a, uint32_t inc) {
  if (threadIdx.x == 0) a[0] += inc;
}

__global__ void k_set_u32(uint32_t *a, uint32_t value) {
  if (threadIdx.x == 0) a[0] = value;
}

/* ------------------------------------------------------------------------ */
/* kernel view: what the device needs from LmEngine                         */
/* ------------------------------------------------------------------------ */
struct KView {
  int n_runs, nx, nz, nc;
  double x_min, x_max, z_min, z_max, dx, dz, V_cell, p_ref, p_ref_inv;
  int solute, no_iter;
  double g_set, dt_col_over_dV;
  const double *E;
  int nb, e_const_mode;
  double E_const, lnR0, lnf, E_max;
  long long vel_stride;
  unsigned long long seed;
  const uint32_t *col_step_dev;
  double *x, *z, *u, *w, *m_w, *m_s, *xi;
  int32_t *id, *cell, *cell_next, *seg;
  double *x2, *z2, *u2, *w2, *m_w2, *m_s2, *xi2;
  int32_t *id2, *cell2, *seg2;
  double *dmxi, *scal, *mat, *cc, *dTh, *dRv, *dml;
  int32_t *away;
  const double *grid_vel, *flux, *prof_rv, *prof_Th, *t_relax;
  double *colsum;
  long long *no_cols;
  double *water_removed;
  long long *n_equal_xi;
  int32_t *rem_cnt, *rem_slot;
  int rem_cap;
  int32_t *scan_buf, *scan_part, *err, *ovf_list, *ovf_cnt;
  int32_t *mig_cnt, *mig_slot; /* migrant lists (NULL: off) */
  int mig_cap;
  int spec_events; /* events through the candidate queue before the anti-diagonal schedule */
  uint32_t *col_cost; /* per-cell cost of the last collision step (ticks), or null */
  /* member lists of the reference-layout collision step (NULL: engine layout): the SIP indices of cell c
     are mlist[moff[c] .. moff[c+1]), in any order */
  const int32_t *mlist, *moff;
  const int32_t *work; /* (run, cell) work list of the persistent passes */
  char *big_ws;        /* per-workgroup regions of the pass over cells of any size (NULL: off) */
  int big_cap;         /* members per region */
  int32_t *work_ctr;   /* its 2 counters: collision [0], Delta m_l [1] */
  const double *rnd_stream;
  long long *rnd_off;
  long long rnd_len;
  /* adjoint window: event log (record / replay) and the tape of the current substep */
  int ev_mode, ev_cap;
  LmEvent *ev;
  int32_t *ev_cnt, *ev_count, *slot_of;
  const long long *run_base;
  double *tape_part, *ccd;
  int32_t *tape_part_i;
  double ki_inv_lnf, ki_margin; /* kernel_index_fast: 1 / lnf and the bound on its error in t (set by ki_setup) */
};

#define SPEC_MAX_EVENTS 48 /* default of LmEngine.spec_events */

/*
 * Kernel-table index without the double log where the answer is certain.  The index is
 * clamp(trunc(t), 0, nb - 1) with t = (ln R - lnR0) / lnf + 0.5 (all_or_nothing.py:35-65, lmp_kernel_index).
 * kernel_index_fast takes ln R from the hardware log2 of R rounded to float: |ln(float(R)) - ln R| <= 2^-24
 * (1 + 2^-23), the hardware log2 errs by at most KI_LOG2_ERR (absolute, on every float of [KI_RMIN, KI_RMAX):
 * checked exhaustively on the GPU by lm_kernel_index_check), and the double operations of both evaluations
 * add below 1e-12 in ln R and 1e-9 in t.  So |t_approx - t| < ki_margin, and a t_approx farther than that from
 * every integer truncates (after the clamp) like t; otherwise, and outside [KI_RMIN, KI_RMAX) or for NaN, the
 * exact evaluation runs.
 */
#define KI_LOG2_ERR 4.0e-6
#define KI_RMIN 0x1p-14
#define KI_RMAX 0x1p16
static void ki_setup(KView &v) {
  const double e_ln = 0x1.000002p-24 + KI_LOG2_ERR * 0.6931471805599453 + 1e-12;
  v.ki_inv_lnf = 0.0;
  v.ki_margin = 1.0; /* fast path off */
  if (v.lnf > 0.0 && v.lnf < 1e300) {
    v.ki_inv_lnf = 1.0 / v.lnf;
    const double m = 2.0 * e_ln * v.ki_inv_lnf + 1e-9;
    if (m < 0.25) v.ki_margin = m;
  }
}

__device__ static inline int kernel_index_fast(double R, const KView &v) {
  if (R >= KI_RMIN && R < KI_RMAX) {
    const double L = (double)__builtin_amdgcn_logf((float)R) * 0.6931471805599453;
    const double ta = (L - v.lnR0) * v.ki_inv_lnf + 0.5;
    const double fl = floor(ta);
    const double fr = ta - fl;
    if (fr > v.ki_margin && fr < 1.0 - v.ki_margin)
      return fl < 0.0 ? 0 : (fl > (double)(v.nb - 1) ? v.nb - 1 : (int)fl);
  }
  return lmp_kernel_index(R, v.lnR0, v.lnf, v.nb);
}

static KView make_view(const LmEngine *e) {
  KView v;
  memset(&v, 0, sizeof(v));
  v.n_runs = e->n_runs;
  v.nx = e->nx;
  v.nz = e->nz;
  v.nc = e->nc;
  v.x_min = e->x_min;
  v.x_max = e->x_max;
  v.z_min = e->z_min;
  v.z_max = e->z_max;
  v.dx = e->dx;
  v.dz = e->dz;
  v.V_cell = e->V_cell;
  v.p_ref = e->p_ref;
  v.p_ref_inv = e->p_ref_inv;
  v.solute = e->solute;
  v.no_iter = e->no_iter;
  v.g_set = e->g_set;
  v.dt_col_over_dV = e->dt_col_over_dV;
  v.E = e->E;
  v.nb = e->nb;
  v.e_const_mode = e->e_const_mode;
  v.E_max = e->E_max;
  v.vel_stride = (long long)e->vel_run_stride;
  v.E_const = e->E_const;
  v.lnR0 = e->lnR0;
  v.lnf = e->lnf;
  ki_setup(v);
  v.seed = e->seed;
  v.col_step_dev = e->col_step_dev;
  const int c = e->cur, o = 1 - e->cur;
  v.x = e->x[c];
  v.z = e->z[c];
  v.u = e->u[c];
  v.w = e->w[c];
  v.m_w = e->m_w[c];
  v.m_s = e->m_s[c];
  v.xi = e->xi[c];
  v.id = e->id[c];
  v.seg = e->seg[c];
  v.x2 = e->x[o];
  v.z2 = e->z[o];
  v.u2 = e->u[o];
  v.w2 = e->w[o];
  v.m_w2 = e->m_w[o];
  v.m_s2 = e->m_s[o];
  v.xi2 = e->xi[o];
  v.id2 = e->id[o];
  v.seg2 = e->seg[o];
  v.cell = e->cell[e->ccur];
  v.cell_next = e->cell[1 - e->ccur];
  v.cell2 = e->cell[1 - e->ccur];
  v.dmxi = e->dmxi;
  v.away = e->away;
  v.scal = e->scal;
  v.mat = e->mat;
  v.cc = e->cc;
  v.dTh = e->dTh;
  v.dRv = e->dRv;
  v.dml = e->dml;
  v.grid_vel = e->grid_vel;
  v.flux = e->flux;
  v.prof_rv = e->prof_rv;
  v.prof_Th = e->prof_Th;
  v.t_relax = e->t_relax;
  v.colsum = e->colsum;
  v.no_cols = (long long *)e->no_cols;
  v.water_removed = e->water_removed;
  v.n_equal_xi = (long long *)e->n_equal_xi;
  v.rem_cnt = e->rem_cnt;
  v.rem_slot = e->rem_slot;
  v.rem_cap = e->rem_cap;
  v.scan_buf = e->scan_buf;
  v.scan_part = e->scan_part;
  v.err = e->err;
  v.ovf_list = e->ovf_list;
  v.ovf_cnt = e->ovf_cnt;
  v.mig_cnt = e->mig_cnt;
  v.mig_slot = e->mig_slot;
  v.mig_cap = e->mig_cap;
  v.spec_events = e->spec_events > 0 ? e->spec_events : SPEC_MAX_EVENTS;
  v.col_cost = e->col_cost;
  v.work = e->work;
  v.big_ws = (char *)e->big_ws;
  v.big_cap = e->big_cap;
  v.work_ctr = e->work ? e->work + (size_t)e->n_runs * e->nc + LM_WORK_BINS : nullptr;
  v.rnd_stream = e->rnd_stream;
  v.rnd_off = (long long *)e->rnd_off;
  v.rnd_len = e->rnd_len;
  v.ev_mode = e->ev_mode;
  v.ev_cap = e->ev_cap;
  v.ev = e->ev;
  v.ev_cnt = e->ev_cnt;
  v.ev_count = e->ev_count;
  v.slot_of = e->slot_of;
  v.run_base = (const long long *)e->run_base;
  if (e->tape_part) {
    v.tape_part = e->tape_part + (size_t)e->tape_sub * 7 * (size_t)e->n_tot;
    v.tape_part_i = e->tape_part_i + (size_t)e->tape_sub * 2 * (size_t)e->n_tot;
  }
  v.ccd = e->ccd;
  return v;
}

/*
 * Members of global cell gc from its home segment and its migrant list, valid while the list is complete
 * (mig_cnt[gc] <= mig_cap): the home segment [s0, s1) holds its SIPs in ascending id, the list the slots of
 * SIPs that entered from other segments since the reindex (some stale: moved on, or entries repeated).
 * Writes the members' slots in ascending id into sSL (up to cap) and returns their number (above cap: the
 * cell is over capacity), or -1 when the list is incomplete (the caller takes the 3x3 gather instead).
 * LDS scratch: sOwn, sOid (cap ints each), sM (3 * LM_MIG_CAP_MAX ints).  Ends with a barrier.
 */
__device__ static int fast_members(const KView &v, int gc, int *sSL, int cap, int *sOwn, int *sOid, int *sM) {
  const int lane = threadIdx.x;
  const int mc = v.mig_cnt[gc];
  if (mc > v.mig_cap) return -1;
  const int nc = v.nc, r = gc / nc, c = gc % nc;
  const int32_t *seg = v.seg + (size_t)r * (nc + 1);
  const int s0 = seg[c], T0 = seg[c + 1] - s0;
  /* home segment: members in slot order, which is id order */
  int n_own = 0;
  for (int base = 0; base < T0; base += WAVE) {
    const int t = base + lane;
    const bool mem = t < T0 && v.cell[s0 + t] == gc;
    const unsigned long long bal = __ballot(mem);
    const int pos = n_own + __popcll(bal & lane_mask_lt());
    if (mem && pos < cap) sOwn[pos] = s0 + t;
    n_own += __popcll(bal);
  }
  int *sMk = sM, *sMs = sM + LM_MIG_CAP_MAX, *sMid = sM + 2 * LM_MIG_CAP_MAX;
  int nm = 0;
  if (mc > 0) {
    /* the list's members, each once (an earlier entry with the same slot wins), then sorted by id */
    int k = -1;
    bool mem = false;
    if (lane < mc) {
      k = v.mig_slot[(size_t)gc * v.mig_cap + lane];
      mem = v.cell[k] == gc;
      sMk[lane] = k;
    }
    __syncthreads();
    if (mem)
      for (int e = 0; e < lane; ++e)
        if (sMk[e] == k) mem = false;
    const unsigned long long bal = __ballot(mem);
    nm = __popcll(bal);
    const int idk = mem ? v.id[k] : 0;
    int rk = 0;
    __syncthreads();
    if (mem) sMid[__popcll(bal & lane_mask_lt())] = idk;
    __syncthreads();
    if (mem)
      for (int e = 0; e < nm; ++e) rk += sMid[e] < idk;
    __syncthreads();
    if (mem) {
      sMs[rk] = k;
      sMid[rk] = idk;
    }
  }
  const int n = n_own + nm;
  __syncthreads();
  if (n > cap) return n;
  if (nm == 0) {
    for (int p = lane; p < n_own; p += WAVE) sSL[p] = sOwn[p];
  } else {
    for (int p = lane; p < n_own; p += WAVE) sOid[p] = v.id[sOwn[p]];
    __syncthreads();
    for (int p = lane; p < n_own; p += WAVE) {
      const int idp = sOid[p];
      int lo = 0;
      while (lo < nm && sMid[lo] < idp) ++lo; /* at most LM_MIG_CAP_MAX */
      sSL[p + lo] = sOwn[p];
    }
    for (int q = lane; q < nm; q += WAVE) {
      const int idq = sMid[q];
      int lo = 0, hi = n_own;
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (sOid[mid] < idq)
          lo = mid + 1;
        else
          hi = mid;
      }
      sSL[q + lo] = sMs[q];
    }
  }
  __syncthreads();
  return n;
}

/* ------------------------------------------------------------------------ */
/* per-cell constants                                                       */
/* ------------------------------------------------------------------------ */
/* sigma_w: the water surface tension at the cell's T when the caller has it (the same lmp_sigma_water), else NaN */
__device__ static void fill_cc(const LmPhys &P, const double *scal_r, const double *mat_r, int c, int nc, double *C,
                               double sigma_w = NAN) {
  const double T = scal_r[0 * nc + c], p = scal_r[1 * nc + c];
  const double K = mat_r[0 * nc + c], D_v = mat_r[1 * nc + c], L_v = mat_r[2 * nc + c];
  const double e_s = scal_r[7 * nc + c];
  C[CC_T] = T;
  C[CC_P] = p;
  C[CC_S] = scal_r[6 * nc + c];
  C[CC_ES] = e_s;
  C[CC_LV] = L_v;
  C[CC_K] = K;
  C[CC_DV] = D_v;
  C[CC_MU] = mat_r[5 * nc + c];
  C[CC_RHOF] = mat_r[6 * nc + c];
  C[CC_LALPHA] = lmp_l_alpha(P, T, p, K);
  C[CC_LBETA] = lmp_l_beta(P, T, D_v);
  C[CC_C1] = L_v * L_v / (P.R_v * K * T * T);  /* microphysics.py:1458 */
  C[CC_C2] = P.R_v * T / (D_v * e_s);          /* microphysics.py:1459 */
  C[CC_RHOW] = lmp_density_water(P, T);
  C[CC_SIGW] = sigma_w == sigma_w ? sigma_w : lmp_sigma_water(T);
}

__global__ void k_init_cc(KView v) {
  const LmPhys &P = c_P;
  const int gc = blockIdx.x * blockDim.x + threadIdx.x;
  if (gc >= v.n_runs * v.nc) return;
  const int r = gc / v.nc, c = gc % v.nc;
  double *C = v.cc + (size_t)gc * CC_N;
  fill_cc(P, v.scal + (size_t)r * 10 * v.nc, v.mat + (size_t)r * 7 * v.nc, c, v.nc, C);
  C[CC_TP] = C[CC_T];
}

/*
 * Cell source of the one-wavefront-per-cell passes.  PASS_CELLS: one
 * workgroup per global cell (blockIdx.x).  PASS_OVF: the cells queued in the
 * overflow list, grid-strided.  PASS_WORK: persistent workgroups over the
 * (run, cell) work list, either grid-strided (chunk 0: a static round robin
 * over the size-ordered list, no atomics) or taking `chunk` consecutive items
 * per atomic on the pass's counter until the list is exhausted (the counter
 * only grows, so every wave exits).  The state is wave-uniform.
 */
#define PASS_CELLS 0
#define PASS_OVF 1
#define PASS_WORK 2
struct CellSource {
  int mode, chunk, ctr; /* pass mode, items per grab, counter index */
  int it, next, end;    /* iteration; current chunk [next, end) */
};

__device__ static inline CellSource cell_source(int mode, int chunk, int ctr) {
  CellSource s = {mode, chunk, ctr, 0, 0, 0};
  return s;
}

/*
 * next global cell of this workgroup, or -1 (PASS_OVF: from ovf_list[0 .. *ovf_cnt), grid-strided, or, with
 * grab counters, claimed one entry at a time: the workgroups share GRAB_CTRS counters GRAB_STRIDE ints apart,
 * one cache line each (workgroup b the counter b mod G, G = min(GRAB_CTRS, gridDim.x), so consecutive
 * workgroups, which the dispatcher places on different XCDs, spread their atomics over several lines); counter
 * x hands out entries x, x + G, x + 2G, ...  Every counter has a workgroup, the counters only grow, so each
 * entry is taken exactly once and every workgroup exits.  The counters must be zero at the launch.)
 */
#define GRAB_CTRS 8
#define GRAB_STRIDE 32
__device__ static inline int next_cell(const KView &v, CellSource &s, const int32_t *ovf_list,
                                       const int32_t *ovf_cnt, int32_t *grab = nullptr) {
  const int it = s.it++;
  if (s.mode == PASS_CELLS) return it == 0 ? (int)blockIdx.x : -1;
  const int n = v.n_runs * v.nc;
  if (s.mode == PASS_OVF && grab) {
    const int G = (int)gridDim.x < GRAB_CTRS ? (int)gridDim.x : GRAB_CTRS;
    const int x = (int)blockIdx.x % G;
    int k = 0;
    if (threadIdx.x == 0) k = atomicAdd(grab + GRAB_STRIDE * x, 1);
    k = __builtin_amdgcn_readfirstlane(k);
    const long long idx = x + (long long)G * k;
    return idx < *ovf_cnt ? ovf_list[idx] : -1;
  }
  if (s.mode == PASS_OVF || s.chunk <= 0) {
    const int idx = (int)blockIdx.x + it * (int)gridDim.x;
    if (s.mode == PASS_OVF) return idx < *ovf_cnt ? ovf_list[idx] : -1;
    return idx < n ? v.work[idx] : -1;
  }
  if (s.next >= s.end) {
    int b = 0;
    if (threadIdx.x == 0) b = atomicAdd(v.work_ctr + s.ctr, s.chunk);
    b = __builtin_amdgcn_readfirstlane(b);
    if (b >= n) return -1;
    s.next = b;
    s.end = b + s.chunk < n ? b + s.chunk : n;
  }
  return v.work[s.next++];
}

/* ------------------------------------------------------------------------ */
/* collision step: one wavefront per cell, anti-diagonal pair schedule       */
/* ------------------------------------------------------------------------ */
/*
 * LDS holds what every pair reads (xi, R, u, w, kernel index) in id order;
 * m_w, m_s and the pre-step density live in global memory (slot-indexed) and
 * are touched only by collision events.  Lane l takes the rows i = l (mod 64)
 * of the pair loop, so a lane walks one row through consecutive diagonals
 * and consecutive pair numbers: one Philox block serves two diagonals.
 */
/* p > u(i, j), deciding from the guard byte G and the A word when possible (see lm_philox.h) */
__device__ static inline bool pair_fires(double p, uint32_t G, uint32_t aw, unsigned long long seed, uint32_t run,
                                         uint32_t step, uint32_t cell, uint32_t i, uint32_t d) {
  const double lo = lm_pair_lo40(G, aw); /* (G 2^32 + A) 2^-40 <= u < lo + 2^-40 */
  if (p < lo) return false;
  if (p >= lo + (1.0 / 1099511627776.0)) return true;
  const lm_u32x4 b = lm_pair_block(seed, run, step, cell, i, d, 1);
  return p > lm_pair_u53(G, aw, lm_word(b, d & 3));
}

/* State of one cell's collision step in LDS (id order) plus slot-indexed globals. */
struct CollCell {
  double *sX, *sR, *sU, *sW;
  int *sIND, *sSL;
  int n;
  int gc;   /* global cell */
  int *seq; /* LDS: events of this cell so far (event log) */
  const float *sB; /* LDS: row bounds of the pass filter */
  double T, rho_w_T; /* the cell's temperature and water density at it */
  uint32_t *chg;     /* one bit per member: its masses changed in this step (its pre-step density saved) */
};

/* global SIP id of slot k: runs occupy the slot ranges [run_base[r], run_base[r+1]) */
__device__ static inline int slot_run(const long long *run_base, int n_runs, long long k) {
  int lo = 0, hi = n_runs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (run_base[mid] <= k)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

/* appends one event to the log of an adjoint window (ev_mode 1) */
__device__ static void record_event(const KView &v, const CollCell &cc, int kmin, int kmax, int kind, double p,
                                   double xi_min, double xi_max, double mw_min, double mw_max, double ms_min,
                                   double ms_max) {
  const int seq = atomicAdd(cc.seq, 1);
  const int e = atomicAdd(v.ev_cnt, 1);
  if (e >= v.ev_cap) {
    raise_err(v.err, LM_ERR_EV_CAP, cc.gc);
    return;
  }
  const int r = cc.gc / v.nc;
  LmEvent E;
  E.gid_min = (int)(v.run_base[r] + v.id[kmin]);
  E.gid_max = (int)(v.run_base[r] + v.id[kmax]);
  E.cell = cc.gc;
  E.seq = seq;
  E.kind = kind;
  E.pad = 0;
  E.p = p;
  E.xi_min = xi_min;
  E.xi_max = xi_max;
  E.mw_min = mw_min;
  E.mw_max = mw_max;
  E.ms_min = ms_min;
  E.ms_max = ms_max;
  v.ev[e] = E;
}

/* injected uniform k; past the end of the stream: error, and a value no pair fires on */
__device__ static inline double rnd_at(const KView &v, long long k) {
  if (k < v.rnd_len) return v.rnd_stream[k];
  raise_err(v.err, LM_ERR_RND_SHORT, 0);
  return 2.0;
}

/* p_crit of pair (i, j) with the current state (all_or_nothing.py:569-575) */
__device__ static inline double pair_p(const LmPhys &P, const KView &v, const CollCell &cc, int i, int j, double Ec) {
  const double xi_i = cc.sX[i], xi_j = cc.sX[j];
  const double xi_max = xi_i <= xi_j ? xi_j : xi_i;
  const double du = cc.sU[i] - cc.sU[j], dw = cc.sW[i] - cc.sW[j];
  return xi_max * v.dt_col_over_dV * lmp_kernel_hydro(P, cc.sR[i], cc.sR[j], Ec, sqrt(du * du + dw * dw));
}

/* member m's bit of the changed set (plain read; set by chg_set, atomically, any number of lanes at once) */
__device__ static inline bool chg_get(const uint32_t *chg, int m) { return (chg[m >> 5] >> (m & 31)) & 1u; }
__device__ static inline void chg_set(uint32_t *chg, int m) { atomicOr(chg + (m >> 5), 1u << (m & 31)); }

/* the density of all_or_nothing.py:822-827 (compute_R_p_w_s_rho_p of the cell before its loop) from a member's
   masses */
__device__ static inline double pre_density(const LmPhys &P, const KView &v, const CollCell &cc, double mw, double ms) {
  const double m_p = mw + ms;
  const double w_s = ms / m_p;
  return lmp_density_solution(P, v.solute, w_s, cc.T, cc.rho_w_T);
}

/*
 * Global state of the two members of a pair: masses and pre-step densities (one round of loads).  A member
 * no event of this step has changed still holds its pre-step masses, and its density is computed from them;
 * a changed member's pre-step density was saved in rho_g (slot-indexed) by the event that first changed it.
 */
struct PairMass {
  double mw_i, mw_j, ms_i, ms_j, rho_i, rho_j;
};
__device__ static inline PairMass load_pair(const LmPhys &P, const KView &v, const CollCell &cc, const double *rho_g,
                                            int i, int j) {
  const int ki = cc.sSL[i], kj = cc.sSL[j];
  const bool ci = chg_get(cc.chg, i), cj = chg_get(cc.chg, j);
  PairMass m = {v.m_w[ki], v.m_w[kj], v.m_s[ki], v.m_s[kj], ci ? rho_g[ki] : 0.0, cj ? rho_g[kj] : 0.0};
  if (!ci) m.rho_i = pre_density(P, v, cc, m.mw_i, m.ms_i);
  if (!cj) m.rho_j = pre_density(P, v, cc, m.mw_j, m.ms_j);
  return m;
}

/* before member m (slot k) first takes new masses: keep its pre-step density for later events of the step */
__device__ static inline void keep_pre_density(const KView &v, const CollCell &cc, int m, int k, double rho) {
  if (!chg_get(cc.chg, m)) {
    v.dmxi[k] = rho;
    chg_set(cc.chg, m);
  }
}

/*
 * Apply the event of a firing pair (all_or_nothing.py:577-632): multiple
 * collection if p_crit > 1, else a single collection (with the near-equal xi
 * split).  m_w, m_s live in global memory (slot-indexed), the pre-step
 * density in rho_g; M holds the pair's values of both, loaded before.
 * Returns 1 for a multiple, 2 for a single, 3 for a near-equal single event.
 */
__device__ static inline int apply_event(const LmPhys &P, const KView &v, const CollCell &cc, const PairMass &M,
                                         int i, int j, double p_crit) {
  int imin, imax;
  const bool i_min = cc.sX[i] <= cc.sX[j];
  if (i_min) {
    imin = i;
    imax = j;
  } else {
    imin = j;
    imax = i;
  }
  const double xi_min = cc.sX[imin], xi_max = cc.sX[imax];
  const int kmin = cc.sSL[imin], kmax = cc.sSL[imax];
  const double mw_min = i_min ? M.mw_i : M.mw_j, mw_max = i_min ? M.mw_j : M.mw_i;
  const double ms_min = i_min ? M.ms_i : M.ms_j, ms_max = i_min ? M.ms_j : M.ms_i;
  const double rho_min = i_min ? M.rho_i : M.rho_j, rho_max = i_min ? M.rho_j : M.rho_i;
  const bool table = !v.e_const_mode;
  if (v.ev_mode == 1) {
    const int kind = p_crit > 1.0 ? 1 : ((xi_max - xi_min) / xi_max < 1.0E-5 ? 3 : 2);
    record_event(v, cc, kmin, kmax, kind, p_crit, xi_min, xi_max, mw_min, mw_max, ms_min, ms_max);
  }
  if (p_crit > 1.0) {
    const double xi_col = p_crit * xi_min;
    const double mwn = mw_min + p_crit * mw_max, msn = ms_min + p_crit * ms_max;
    keep_pre_density(v, cc, imin, kmin, rho_min);
    v.m_w[kmin] = mwn;
    v.m_s[kmin] = msn;
    const double Rn = lmp_radius(P, (mwn + msn), rho_min);
    cc.sR[imin] = Rn;
    cc.sX[imax] = xi_max - xi_col;
    if (table) cc.sIND[imin] = kernel_index_fast(Rn, v);
    return 1;
  }
  const double xi_rel_dev = (xi_max - xi_min) / xi_max;
  if (xi_rel_dev < 1.0E-5) {
    const double xi_ges = xi_min + xi_max;
    const double mwn = 2.0 * (xi_min * mw_min + xi_max * mw_max) / xi_ges;
    const double msn = 2.0 * (xi_min * ms_min + xi_max * ms_max) / xi_ges;
    keep_pre_density(v, cc, imin, kmin, rho_min);
    keep_pre_density(v, cc, imax, kmax, rho_max);
    v.m_w[kmin] = mwn;
    v.m_s[kmin] = msn;
    v.m_w[kmax] = mwn;
    v.m_s[kmax] = msn;
    const double Rn = lmp_radius(P, (mwn + msn), rho_min);
    cc.sR[imin] = Rn;
    cc.sR[imax] = Rn;
    const double xmax_new = 0.5 * 0.7 * xi_ges;
    cc.sX[imax] = xmax_new;
    cc.sX[imin] = 0.5 * xi_ges - xmax_new;
    if (table) {
      const int ind = kernel_index_fast(Rn, v);
      cc.sIND[imin] = ind;
      cc.sIND[imax] = ind;
    }
    return 3;
  }
  const double mwn = mw_min + mw_max, msn = ms_min + ms_max;
  keep_pre_density(v, cc, imin, kmin, rho_min);
  v.m_w[kmin] = mwn;
  v.m_s[kmin] = msn;
  const double Rn = lmp_radius(P, (mwn + msn), rho_min);
  cc.sR[imin] = Rn;
  cc.sX[imax] = xi_max - xi_min;
  if (table) cc.sIND[imin] = kernel_index_fast(Rn, v);
  return 2;
}

/*
 * Row bounds of the collision pass filter.  With the cell-wide X = max xi,
 * Rc = max |R| and velocity ranges [u_lo, u_hi], [w_lo, w_hi] (NaN entries
 * skipped: a pair with a NaN in xi_i, R, u or w never fires, and xi_max is
 * xi_i when xi_j is NaN), every rounded factor of a pair's Pl and d2 in row i
 * is at most the same expression evaluated on the bounds, because IEEE
 * rounding is monotone: xi_max <= X, |fl(R_i + R_j)| <= fl(|R_i| + Rc),
 * E <= Eb, |fl(u_i - u_j)| <= max(fl(u_i - u_lo), fl(u_hi - u_i)).  So every
 * pair of the row has Pl^2 d2 <= B_i = (P_i^2) d2_i, and a quad whose smallest
 * threshold thr has thr^2 (1 - 2^-40) > B_i holds no pair with p >= thr (the
 * per-pair filter's argument).  P_i <= 1e-140 with d2_i < 1e100 bounds p below
 * 1e-89 and is stored as B = 1e-178.  B is kept as a float rounded up (a
 * larger B only rejects less); NaN or inf never rejects.  X only decreases
 * with events; Rc can grow, so the bounds are recomputed after each event.
 */
/*
 * Quads the candidate queue of a pass of member capacity CAP holds (entries lane + e * WAVE of the evaluating
 * wave, e < qcap_of(CAP) / WAVE).  A queue must take at least the quads the guard pass can never reject --
 * those with a zero byte, 1 - (255/256)^4 = 1.55% of a cell's n^2/8 quads, 127 at 256 members -- or every
 * cell of that size overflows into the anti-diagonal schedule.
 */
__host__ __device__ constexpr int qcap_of(int cap) {
  return cap <= 0 ? 8 * WAVE : (cap <= 160 ? 2 * WAVE : (cap <= 256 ? 4 * WAVE : 8 * WAVE));
}

/*
 * Region of the pass over cells of any size (capacity 0): the per-member arrays the LDS passes keep in LDS, in
 * global memory, cap members.
 */
struct BigWs {
  double *X, *R, *U, *W;
  int32_t *IND, *SL, *Q;
  float *B;
  unsigned char *TG, *IsBig;
  uint32_t *Chg;
};
__host__ __device__ constexpr size_t big_up(size_t b) { return (b + 255) & ~(size_t)255; }
__host__ __device__ constexpr size_t big_stride(int cap) {
  return 4 * big_up(8 * (size_t)cap) + 2 * big_up(4 * (size_t)cap) + big_up(4 * ((size_t)cap + 1)) +
         big_up(4 * (size_t)cap) + 2 * big_up((size_t)cap) + big_up(4 * (((size_t)cap + 31) / 32));
}
__device__ static inline BigWs big_carve(char *b, int cap) {
  BigWs w;
  size_t o = 0;
  w.X = (double *)(b + o);
  o += big_up(8 * (size_t)cap);
  w.R = (double *)(b + o);
  o += big_up(8 * (size_t)cap);
  w.U = (double *)(b + o);
  o += big_up(8 * (size_t)cap);
  w.W = (double *)(b + o);
  o += big_up(8 * (size_t)cap);
  w.IND = (int32_t *)(b + o);
  o += big_up(4 * (size_t)cap);
  w.SL = (int32_t *)(b + o);
  o += big_up(4 * (size_t)cap);
  w.Q = (int32_t *)(b + o);
  o += big_up(4 * ((size_t)cap + 1));
  w.B = (float *)(b + o);
  o += big_up(4 * (size_t)cap);
  w.TG = (unsigned char *)(b + o);
  o += big_up((size_t)cap);
  w.IsBig = (unsigned char *)(b + o);
  o += big_up((size_t)cap);
  w.Chg = (uint32_t *)(b + o);
  return w;
}

/* Largest multiplicity and radius and the velocity ranges over the members outside big[0..m) (wave-wide). */
struct CellMax {
  double X, Rc, ulo, uhi, wlo, whi;
};

#define BIG_MAX 8 /* members the guard pass can leave to the per-pair pass */

/* Largest multiplicity and radius and the velocity ranges over the members whose rank byte is 0 or above m
   (rank = 1 + position in the cell's list of largest radii, 0 if not listed; wave-wide). */
__device__ static CellMax cell_max(int n, const double *sX, const double *sR, const double *sU, const double *sW,
                                   const unsigned char *sRank, int m) {
  const int lane = threadIdx.x;
  CellMax M = {-INFINITY, 0.0, INFINITY, -INFINITY, INFINITY, -INFINITY};
  for (int p = lane; p < n; p += WAVE) {
    const int rk = sRank[p];
    if (rk != 0 && rk <= m) continue;
    M.X = fmax(M.X, sX[p]);
    M.Rc = fmax(M.Rc, fabs(sR[p]));
    M.ulo = fmin(M.ulo, sU[p]);
    M.uhi = fmax(M.uhi, sU[p]);
    M.wlo = fmin(M.wlo, sW[p]);
    M.whi = fmax(M.whi, sW[p]);
  }
  for (int off = WAVE / 2; off > 0; off >>= 1) {
    M.X = fmax(M.X, __shfl_xor(M.X, off));
    M.Rc = fmax(M.Rc, __shfl_xor(M.Rc, off));
    M.ulo = fmin(M.ulo, __shfl_xor(M.ulo, off));
    M.uhi = fmax(M.uhi, __shfl_xor(M.uhi, off));
    M.wlo = fmin(M.wlo, __shfl_xor(M.wlo, off));
    M.whi = fmax(M.whi, __shfl_xor(M.whi, off));
  }
  return M;
}

/*
 * Bound B of row p against the members counted in M: for every such j, p(p, j)^2 <= B within the roundings of
 * the pair filter (xi_max <= X, R_p + R_j <= |R_p| + Rc, |v_p - v_j| within the velocity ranges), rounded up
 * to a float; and the row's guard-byte bound st: a pair whose leading byte G exceeds st has
 * (G/256)^2 (1 - 2^-40) > B, the quad filter's own test, so it cannot fire (G/256 is a lower bound of its
 * uniform).  st = 255 (NaN, or B too large for any byte) never rejects.
 */
__device__ static inline float row_bound(const LmPhys &P, double dtdV, double Eb, const CellMax &M, double R,
                                         double u, double w, int &st) {
  const double du = fmax(u - M.ulo, M.uhi - u), dw = fmax(w - M.wlo, M.whi - w);
  const double d2 = du * du + dw * dw;
  const double sr = fabs(R) + M.Rc;
  const double Pb = (((M.X * dtdV) * ((P.pi * sr) * sr)) * Eb) * 1.0E-12;
  double B = INFINITY;
  if (Pb > 1.0E-140)
    B = (Pb * Pb) * d2;
  else if ((Pb <= 1.0E-140) & (d2 < 1.0E100))
    B = 1.0E-178;
  float f = (float)B; /* B >= 0 or NaN: rounded down, the next float up is above B */
  if ((double)f < B) f = __int_as_float(__float_as_int(f) + 1);
  /* tg = the smallest byte G with (G/256)^2 (1 - 2^-40) > B; st = min(tg, 256) - 1 */
  const double Bf = (double)f;
  int tg = 257;
  if (Bf == Bf && Bf < 1.0) {
    tg = (int)(256.0 * sqrt(Bf));
    if (tg > 256) tg = 256;
    for (int k = 0; k < 4 && tg <= 256 && !(((tg * (1.0 / 256.0)) * (tg * (1.0 / 256.0))) * (1.0 - 0x1p-40) > Bf);
         ++k)
      ++tg;
    if (!(((tg * (1.0 / 256.0)) * (tg * (1.0 / 256.0))) * (1.0 - 0x1p-40) > Bf)) tg = 257;
    while (tg > 1 && (((tg - 1) * (1.0 / 256.0)) * ((tg - 1) * (1.0 / 256.0))) * (1.0 - 0x1p-40) > Bf) --tg;
  }
  st = (tg < 256 ? tg : 256) - 1;
  return f;
}

/* quads of row p (diagonals 2p+1 .. p+n-1 in blocks of four) */
__device__ static inline int row_quads(int p, int n) {
  return p < n - 1 ? ((p + n - 1) >> 2) - ((2 * p + 1) >> 2) + 1 : 0;
}

/* the guard byte bound of row_bound without its float rounding and exact search: only for estimating how many
   quads a choice of the exclusion list keeps (the bounds the passes use come from row_bound) */
__device__ static inline int row_bound_est(const LmPhys &P, double dtdV, double Eb, const CellMax &M, double R,
                                           double u, double w) {
  const double du = fmax(u - M.ulo, M.uhi - u), dw = fmax(w - M.wlo, M.whi - w);
  const double sr = fabs(R) + M.Rc;
  const double Pb = (((M.X * dtdV) * ((P.pi * sr) * sr)) * Eb) * 1.0E-12;
  const double B = (Pb * Pb) * (du * du + dw * dw);
  if (!(B < 1.0)) return 255;
  const int tg = (int)(256.0 * sqrt(B)) + 1;
  return (tg < 256 ? tg : 256) - 1;
}

/* expected quads of row p that the guard pass keeps with bound st (uniform bytes, four per quad) */
__device__ static inline double row_kept(int p, int n, int st) {
  const double q = 1.0 - (st + 1) * (1.0 / 256.0);
  return row_quads(p, n) * (1.0 - (q * q) * (q * q));
}

/*
 * Row bounds of the pass filters.  sB[p]: the row bound over all members (eval_quad's row filter, valid for
 * every pair of row p).  sTG[p]: the guard pass's byte bound, over the members outside the cell's exclusion
 * list sBig[0..*sNbig): a few members of the largest radii (drops among cloud droplets), whose pairs the
 * per-pair pass tests instead.  One large drop widens the radius and velocity ranges of the whole cell and
 * with them every row's bound; the list is used when the expected work of the guard pass's queue drops by
 * more than the per-pair tests cost (quads of nq_all in total; a queue over qcap_of(CAP) quads costs a pass over all).
 * Any list is exact: every pair is covered by its row's bound or by the per-pair pass.
 */
template <int CAP>
__device__ static void row_bounds(const LmPhys &P, double dtdV, double Eb, int n, const double *sX, const double *sR,
                                  const double *sU, const double *sW, float *sB, unsigned char *sTG, int *sBig,
                                  int *sNbig, unsigned char *sRank, int nq_all) {
  const int lane = threadIdx.x;
  for (int p = lane; p < n; p += WAVE) sRank[p] = 0;
  __syncthreads();
  const CellMax M0 = cell_max(n, sX, sR, sU, sW, sRank, 0);
  double kept0 = 0.0;
  for (int p = lane; p < n; p += WAVE) {
    int st;
    sB[p] = row_bound(P, dtdV, Eb, M0, sR[p], sU[p], sW[p], st);
    sTG[p] = (unsigned char)st;
    kept0 += row_kept(p, n, st);
  }
  for (int off = WAVE / 2; off > 0; off >>= 1) kept0 += __shfl_xor(kept0, off);
  const double over = 1.25 * nq_all; /* a queue overflow: the anti-diagonal schedule over all pairs */
  const double q_full = 0.8 * qcap_of(CAP); /* expected kept quads that risk an overflow */
  /* the floor: quads holding a zero byte, which no bound rejects (1 - (255/256)^4 of all) */
  const double floor_kept = 0.0155 * nq_all;
  int best_m = 0;
  if (kept0 > floor_kept + 24.0 && n >= 3 * BIG_MAX) {
    /* the BIG_MAX largest radii, in decreasing order (ties by position): sRank[p] = 1 + place.  Which members
       the list holds only matters for speed (any list is exact); a radius that is NaN is never listed, so the
       list can come out shorter (nsel) */
    int nsel = 0;
    constexpr int RPL = (CAP > 0 && CAP <= 512) ? (CAP + WAVE - 1) / WAVE : 0;
    if constexpr (RPL > 0) {
      /* the lane's radii in registers: a pass is a local scan and a wave reduction, no LDS reads or barrier */
      double rr[RPL];
#pragma unroll
      for (int q = 0; q < RPL; ++q) {
        const int p = lane + q * WAVE;
        rr[q] = p < n ? fabs(sR[p]) : -1.0;
      }
      for (int k = 0; k < BIG_MAX; ++k) {
        double br = -1.0;
        int bp = n;
#pragma unroll
        for (int q = 0; q < RPL; ++q)
          if (rr[q] > br) { /* positions grow with q: a tie keeps the smaller */
            br = rr[q];
            bp = lane + q * WAVE;
          }
        for (int off = WAVE / 2; off > 0; off >>= 1) {
          const double orr = __shfl_xor(br, off);
          const int op = __shfl_xor(bp, off);
          if (orr > br || (orr == br && op < bp)) {
            br = orr;
            bp = op;
          }
        }
        if (bp >= n) break;
#pragma unroll
        for (int q = 0; q < RPL; ++q)
          if (lane + q * WAVE == bp) rr[q] = -1.0;
        if (lane == 0) {
          sRank[bp] = (unsigned char)(k + 1);
          sBig[k] = bp;
        }
        ++nsel;
      }
      __syncthreads();
    } else {
      for (int k = 0; k < BIG_MAX; ++k) {
        double br = -1.0;
        int bp = n;
        for (int p = lane; p < n; p += WAVE) {
          const double r = fabs(sR[p]);
          if (sRank[p] == 0 && (r > br || (r == br && p < bp))) {
            br = r;
            bp = p;
          }
        }
        for (int off = WAVE / 2; off > 0; off >>= 1) {
          const double orr = __shfl_xor(br, off);
          const int op = __shfl_xor(bp, off);
          if (orr > br || (orr == br && op < bp)) {
            br = orr;
            bp = op;
          }
        }
        if (bp >= n) break;
        if (lane == 0) {
          sRank[bp] = (unsigned char)(k + 1);
          sBig[k] = bp;
        }
        ++nsel;
        __syncthreads();
      }
    }
    double best = kept0 + (kept0 > q_full ? over : 0.0);
    for (int m = 2; m <= BIG_MAX && m <= nsel; m *= 2) {
      const CellMax Mm = cell_max(n, sX, sR, sU, sW, sRank, m);
      double kept = 0.0;
      for (int p = lane; p < n; p += WAVE) {
        const int rk = sRank[p];
        if (rk != 0 && rk <= m) continue;
        kept += row_kept(p, n, row_bound_est(P, dtdV, Eb, Mm, sR[p], sU[p], sW[p]));
      }
      for (int off = WAVE / 2; off > 0; off >>= 1) kept += __shfl_xor(kept, off);
      const double cost = kept + 0.18 * n * m + (kept > q_full ? over : 0.0);
      if (cost < best) {
        best = cost;
        best_m = m;
      }
    }
    if (best_m > 0) {
      const CellMax Mm = cell_max(n, sX, sR, sU, sW, sRank, best_m);
      for (int p = lane; p < n; p += WAVE) {
        int st;
        (void)row_bound(P, dtdV, Eb, Mm, sR[p], sU[p], sW[p], st);
        sTG[p] = (unsigned char)st; /* rows of the list are left to the per-pair pass */
      }
    }
    __syncthreads();
    /* only the chosen part of the list stays marked */
    for (int p = lane; p < n; p += WAVE)
      if (sRank[p] > best_m) sRank[p] = 0;
  }
  if (lane == 0) *sNbig = best_m;
}

/*
 * Whether pair (i, j) can fire given only its guard byte G (u >= G/256): the pair filter of eval_quad with
 * the E bound Eb and the threshold G/256 in place of the pair's 40-bit lower bound.  false proves p < G/256
 * <= u: no event and no multiple collection.
 */
__device__ static inline bool pair_may_fire8(const LmPhys &P, double dtdV, double Eb, const double *sX,
                                             const double *sR, const double *sU, const double *sW, int i, int j,
                                             uint32_t G) {
  if (G == 0u) return true;
  const double xi_max = sX[i] <= sX[j] ? sX[j] : sX[i];
  const double du = sU[i] - sU[j], dw = sW[i] - sW[j];
  const double d2 = du * du + dw * dw;
  const double sR2 = sR[i] + sR[j];
  const double thr = (double)G * (1.0 / 256.0);
  const double thr2 = (thr * thr) * (1.0 - 0x1p-40);
  const double Pm = (((xi_max * dtdV) * ((P.pi * sR2) * sR2)) * Eb) * 1.0E-12;
  const bool big = Pm > 1.0E-140;
  const bool no = (big & ((Pm * Pm) * d2 < thr2)) | ((!big) & (d2 < 1.0E100) & (thr2 > 1.0E-178));
  return !no;
}

/* pair number of (i, j) in the reference's row-major loop */
__device__ static inline int pair_cnt(int i, int j, int n) { return i * n - i * (i + 1) / 2 + (j - i - 1); }

#define KEY_STALE 0xFFFFFFFFFFFFFFFEull /* queue entry to evaluate (keys of real pairs are below 2^63) */

/* State of one cell's collision step in LDS (id order) plus slot-indexed globals. */

/*
 * Quad (i, b) of the speculative pass -- four consecutive pairs (i, 4b + t - i) of row i sharing one Philox
 * block -- under the current state: the row-major key (cnt << 32 | i << 16 | j) of its first pair with
 * cnt >= k_next that fires, or ~0.  gw is the quad's guard word (Philox draws), unused with an injected
 * stream.
 */
template <int CAP, bool TAB, bool INJ>
__device__ static inline unsigned long long eval_quad(const LmPhys &P, const KView &v, const CollCell &cc, int i,
                                                      int b, int k_next, uint32_t gw, long long roff,
                                                      uint32_t key_run, uint32_t key_step, uint32_t key_cell,
                                                      bool row_filter) {
  const int n = cc.n;
  const double *sX = cc.sX, *sR = cc.sR, *sU = cc.sU, *sW = cc.sW;
  const int *sIND = cc.sIND;
  const float *sB = cc.sB;
  const double dtdV = v.dt_col_over_dV, pi = P.pi;
  unsigned long long best = ~0ull;
  lm_u32x4 blk = {{0u, 0u, 0u, 0u}};
  if (!INJ) blk = lm_pair_block(v.seed, key_run, key_step, key_cell, (uint32_t)i, (uint32_t)(4 * b), 0);
  {
    /* row filter: the quad's smallest threshold against the row bound (row_bounds) */
    double thq = 1.0;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int jr = 4 * b + t - i;
      const bool valid = (jr > i) & (jr < n) & (pair_cnt(i, jr, n) >= k_next);
      double th;
      if (INJ)
        th = valid ? rnd_at(v, roff + pair_cnt(i, jr, n)) : 1.0;
      else
        th = lm_pair_lo40(lm_guard_byte(gw, (uint32_t)t), lm_word(blk, (uint32_t)t));
      if (valid) thq = fmin(thq, th);
    }
    if (row_filter & (thq > 0.0) & ((thq * thq) * (1.0 - 0x1p-40) > (double)sB[i])) return ~0ull;
  }
  const double xi_i = sX[i], R_i = sR[i], u_i = sU[i], w_i = sW[i];
  const double *Erow = TAB ? v.E + (size_t)sIND[i] * v.nb : nullptr;
  /*
   * Filter.  p = (xi_max dt/dV) * ((((pi s) s) E) sqrt(d2)) 1e-12 with s = R_i + R_j.  Without the
   * square root, Pl^2 d2 (Pl = (xi_max dt/dV) (((pi s) s) E) 1e-12) equals p^2 within 10 roundings,
   * so Pl^2 d2 < thr^2 (1 - 2^-40) proves p < thr, thr = the 40-bit lower bound of the pair's
   * uniform (or the injected uniform): the pair does not fire and is no multiple collection.  For
   * Pl <= 1e-140 (and d2 < 1e100) p < 1e-89 < thr whenever thr^2 > 1e-178.  Everything else --
   * thr = 0, overflow, NaN -- takes the exact path below.
   */
  unsigned maybe = 0;
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int d = 4 * b + t;
    const int jr = d - i;
    const bool valid = (jr > i) & (jr < n) & (pair_cnt(i, jr, n) >= k_next);
    /* clamped into [0, n-1] (jr is -1 at the start of some rows, >= n at the end): every lane runs the
       same straight-line code on in-range LDS entries, and invalid pairs are masked out below */
    const int j = jr < 0 ? 0 : (jr < n ? jr : n - 1);
    const double xi_j = sX[j];
    const double xi_max = xi_i <= xi_j ? xi_j : xi_i;
    const double du = u_i - sU[j], dw = w_i - sW[j];
    const double d2 = du * du + dw * dw;
    const double sR2 = R_i + sR[j];
    double thr;
    if (INJ)
      thr = valid ? rnd_at(v, roff + pair_cnt(i, jr, n)) : 0.0;
    else
      thr = lm_pair_lo40(lm_guard_byte(gw, (uint32_t)t), lm_word(blk, (uint32_t)t));
    const double thr2 = (thr * thr) * (1.0 - 0x1p-40);
    const double geo = (xi_max * dtdV) * ((pi * sR2) * sR2);
    /* first with the table's largest entry: every rounding step is monotone in E, so a pair rejected
       with E_max is rejected with its own entry; only the others read the table */
    bool no = false;
    if (TAB && v.E_max > 0.0) {
      const double Pm = (geo * v.E_max) * 1.0E-12;
      no = (Pm > 1.0E-140) & ((Pm * Pm) * d2 < thr2);
    }
    if (!no) {
      const double Ec = TAB ? Erow[sIND[j]] : v.E_const;
      const double Pl = (geo * Ec) * 1.0E-12;
      const bool big = Pl > 1.0E-140;
      no = (big & ((Pl * Pl) * d2 < thr2)) | ((!big) & (d2 < 1.0E100) & (thr2 > 1.0E-178));
    }
    maybe |= (unsigned)(valid & !no) << t;
  }
  if (maybe) {
    /* rare: the reference's exact p_crit and decision, first firing pair of the quad */
    for (int t = 0; t < 4; ++t) {
      if (!((maybe >> t) & 1u)) continue;
      const int d = 4 * b + t;
      const int j = d - i;
      const int cnt = pair_cnt(i, j, n);
      const double Ec = TAB ? Erow[sIND[j]] : v.E_const;
      const double xi_j = sX[j];
      const double xi_max = xi_i <= xi_j ? xi_j : xi_i;
      const double du = u_i - sU[j], dw = w_i - sW[j];
      const double p_crit = xi_max * dtdV * lmp_kernel_hydro(P, R_i, sR[j], Ec, sqrt(du * du + dw * dw));
      bool fires;
      if (p_crit > 1.0)
        fires = true;
      else if (INJ)
        fires = p_crit > rnd_at(v, roff + cnt);
      else
        fires = pair_fires(p_crit, lm_guard_byte(gw, (uint32_t)t), lm_word(blk, (uint32_t)t), v.seed, key_run,
                           key_step, key_cell, (uint32_t)i, (uint32_t)d);
      if (fires) {
        const unsigned long long key = ((unsigned long long)cnt << 32) | ((unsigned)i << 16) | (unsigned)j;
        best = key < best ? key : best;
        break;
      }
    }
  }
  return best;
}

/*
 * One all-cells collision step, one wavefront per cell.
 *
 * Exactness argument.  The reference walks the pairs of a cell in row-major
 * order and changes the state only when a pair fires.  So, given the state
 * after the last applied event, every later pair up to the first one that
 * fires under that state is evaluated by the reference on that same state.
 * The kernel therefore evaluates all remaining pairs in parallel on the
 * current state (full lane use, no barriers), takes the first firing pair in
 * row-major order (wave minimum), applies it and repeats from the next pair.
 * Cells with more than v.spec_events events finish with the anti-diagonal
 * schedule: pairs grouped by d = i + j are mutually independent and every
 * dependence points to a smaller d, so processing the remaining pairs by
 * increasing d reproduces the sequential loop too.
 *
 * Work items of the parallel pass are quads: row i and one Philox block
 * (diagonals 4b..4b+3), i.e. four consecutive pairs of the row sharing one
 * block of A words.
 */
/* the lists of a collision pass: the cells it reads (PASS_OVF) and where it queues cells over its capacity
   (out_list null: such a cell is an error) */
struct ColLists {
  const int32_t *in_list, *in_cnt;
  int32_t *out_list, *out_cnt;
  int32_t *in_grab; /* GRAB_CTRS zeroed counters: claim in_list's entries one at a time (NULL: grid-strided) */
};

/*
 * Members of a cell with a complete migrant list, staged straight into the pass's arrays (the home segment's
 * loads issued together with the list's, the state of each member loaded with its cell and id rather than after
 * the ranking): every lane loads the cell, id, xi, u, w, m_w and m_s of its home slots (up to HOME_CHUNKS * 64
 * of them) and of its list entry, the members keep slot order (= id order) merged by id with the list's
 * members (each slot once), and each member's xi, u, w, radius, kernel index and slot go to its position.
 * Returns the member count (nothing staged above CAP), or -1 when the list is incomplete or the home segment
 * longer than the lanes hold (the caller takes the general path).  sM: 3 * LM_MIG_CAP_MAX ints, sH: CAP ints of
 * scratch.
 */
template <int CAP, bool TAB>
__device__ static int fast_stage(const LmPhys &P, const KView &v, int gc, double T, double rho_w_T, double *sX,
                                 double *sU, double *sW, double *sR, int *sIND, int *sSL, int *sM, int *sH) {
  constexpr int HC = CAP / WAVE + 1; /* home chunks a lane holds */
  const int lane = threadIdx.x;
  const int nc = v.nc, r = gc / nc, c = gc % nc;
  const int mc = v.mig_cnt[gc];
  const int32_t *seg = v.seg + (size_t)r * (nc + 1);
  const int s0 = seg[c], T0 = seg[c + 1] - s0;
  if (mc > v.mig_cap || T0 > HC * WAVE) return -1;
  /* one round: the home slots' and the list entry's state (entries past T0 / mc are not read) */
  int hcell[HC], hid[HC];
  double hx[HC], hu[HC], hw[HC], hmw[HC], hms[HC];
#pragma unroll
  for (int q = 0; q < HC; ++q) {
    const int t = lane + q * WAVE;
    hcell[q] = -1;
    hid[q] = 0;
    hx[q] = hu[q] = hw[q] = hmw[q] = hms[q] = 0.0;
    if (t < T0) {
      const int k = s0 + t;
      hcell[q] = v.cell[k];
      if (mc > 0) hid[q] = v.id[k]; /* the ids only place the list's members */
      hx[q] = v.xi[k];
      hu[q] = v.u[k];
      hw[q] = v.w[k];
      hmw[q] = v.m_w[k];
      hms[q] = v.m_s[k];
    }
  }
  int mk = -1;
  if (lane < mc) mk = v.mig_slot[(size_t)gc * v.mig_cap + lane];
  int mcell = -1, mid_ = 0;
  double mx = 0.0, mu = 0.0, mw = 0.0, mmw = 0.0, mms = 0.0;
  if (mk >= 0) {
    mcell = v.cell[mk];
    mid_ = v.id[mk];
    mx = v.xi[mk];
    mu = v.u[mk];
    mw = v.w[mk];
    mmw = v.m_w[mk];
    mms = v.m_s[mk];
  }
  /* home members: rank among them by slot order */
  int hpos[HC];
  int n_own = 0;
#pragma unroll
  for (int q = 0; q < HC; ++q) {
    const bool mem = hcell[q] == gc;
    const unsigned long long bal = __ballot(mem);
    hpos[q] = mem ? n_own + __popcll(bal & lane_mask_lt()) : -1;
    n_own += __popcll(bal);
  }
  /* list members, each slot once (an earlier entry with the same slot wins), ranked by id */
  int *sMk = sM, *sMid = sM + LM_MIG_CAP_MAX;
  if (lane < mc) sMk[lane] = mk;
  __syncthreads();
  bool mmem = mk >= 0 && mcell == gc;
  if (mmem)
    for (int e = 0; e < lane; ++e)
      if (sMk[e] == mk) mmem = false;
  const unsigned long long mbal = __ballot(mmem);
  const int nm = __popcll(mbal);
  const int n = n_own + nm;
  if (n > CAP) {
    __syncthreads();
    return n;
  }
  if (mmem) sMid[__popcll(mbal & lane_mask_lt())] = mid_;
  /* home members' ids in rank order, for the list members' positions */
#pragma unroll
  for (int q = 0; q < HC; ++q)
    if (hpos[q] >= 0) sH[hpos[q]] = hid[q];
  __syncthreads();
  /* position in id order: a home member's rank plus the list members with smaller ids; a list member's rank
     among them plus the home members with smaller ids (ascending in sH).  The masses go to sU / sW at that
     position until the radii are known: the radius and kernel index then take ceil(n / 64) rounds instead of
     one per home chunk plus one for the list */
  int hp[HC];
#pragma unroll
  for (int q = 0; q < HC; ++q) {
    hp[q] = -1;
    if (hpos[q] < 0) continue;
    int lo = 0;
    for (int e = 0; e < nm; ++e) lo += sMid[e] < hid[q];
    const int p = hpos[q] + lo;
    hp[q] = p;
    sX[p] = hx[q];
    sU[p] = hmw[q];
    sW[p] = hms[q];
    sSL[p] = s0 + lane + q * WAVE;
  }
  int mp = -1;
  if (mmem) {
    int rk = 0;
    for (int e = 0; e < nm; ++e) rk += sMid[e] < mid_;
    int lo = 0, hi = n_own;
    while (lo < hi) {
      const int m2 = (lo + hi) >> 1;
      if (sH[m2] < mid_)
        lo = m2 + 1;
      else
        hi = m2;
    }
    const int p = rk + lo;
    mp = p;
    sX[p] = mx;
    sU[p] = mmw;
    sW[p] = mms;
    sSL[p] = mk;
  }
  __syncthreads();
  for (int p = lane; p < n; p += WAVE) {
    double R, ws, rho;
    lmp_R_ws_rho(P, v.solute, sU[p], sW[p], T, rho_w_T, R, ws, rho);
    sR[p] = R;
    sIND[p] = TAB ? kernel_index_fast(R, v) : 0;
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < HC; ++q)
    if (hp[q] >= 0) {
      sU[hp[q]] = hu[q];
      sW[hp[q]] = hw[q];
    }
  if (mp >= 0) {
    sU[mp] = mu;
    sW[mp] = mw;
  }
  __syncthreads();
  return n;
}

template <int CAP, bool TAB, bool INJ>
__device__ static void collide_cell(const KView &v, int gc, const ColLists &io) {
  const LmPhys &P = c_P;
  __syncthreads(); /* the LDS of the previous cell of this block is free */
  const unsigned long long t_cell = v.col_cost ? wall_clock64() : 0ull; /* the cell's cost: gather included */
  const int nc = v.nc;
  const int r = gc / nc, c = gc % nc;
  const int lane = threadIdx.x;

  /* xi, R, u, w, kernel index and slot of each member, in id order: in LDS, where the gather scratch lives
     inside the u/w arrays until the ranks are known, or (CAP 0: cells of any size) in the workgroup's region
     of the global workspace big_ws */
  constexpr bool BIG = CAP == 0;
  constexpr int LC = BIG ? 1 : CAP;
  __shared__ double lX[LC], lR[LC], lU[LC], lW[LC];
  __shared__ int lIND[LC], lSL[LC], lQ[LC + 1];
  __shared__ float lB[LC];
  __shared__ unsigned char lTG[LC], lIsBig[LC];
  __shared__ uint32_t lChg[(LC + 31) / 32];
  double *sX = lX, *sR = lR, *sU = lU, *sW = lW;
  int *sIND = lIND, *sSL = lSL, *sQ = lQ;
  float *sB = lB;
  unsigned char *sTG = lTG, *sIsBig = lIsBig; /* sIsBig: rank in the list of largest radii the guard pass
                                                  leaves out, or 0 */
  uint32_t *sChg = lChg;
  if constexpr (BIG) {
    const BigWs b = big_carve(v.big_ws + (size_t)blockIdx.x * big_stride(v.big_cap), v.big_cap);
    sX = b.X;
    sR = b.R;
    sU = b.U;
    sW = b.W;
    sIND = b.IND;
    sSL = b.SL;
    sQ = b.Q;
    sB = b.B;
    sTG = b.TG;
    sIsBig = b.IsBig;
    sChg = b.Chg;
  }
  constexpr int QC = qcap_of(CAP), NQ = QC / WAVE;
  __shared__ int sQueue[QC];
  __shared__ unsigned long long sQkey[QC]; /* each entry's first firing pair (key), ~0: none, KEY_STALE */
  __shared__ int sQn;
  __shared__ int sBig[BIG_MAX], sNbig;
  __shared__ NbrTab tb;
  __shared__ int sSeq;
  int *sSlot = (int *)lU;
  int *sId = sSlot + CAP;
  int *sRank = (int *)lW;
  unsigned char *sSub = (unsigned char *)(sRank + CAP);

  /* members: from the home segment and the migrant list (slots in id order straight into sSL), or, with
     an incomplete list, gathered from the 3x3 home segments and ranked by id; cells of any size: merged in
     id order from the 3x3 home segments by one lane */
  int n;
  bool fast;
  bool staged = false; /* members' state already in the arrays (fast_stage) */
  if constexpr (BIG) {
    fast = true;
    if (v.mlist) {
      /* the cell's member list, sorted by SIP index (rank by counting: the rare cells of this pass) */
      const int o0 = v.moff[gc];
      n = v.moff[gc + 1] - o0;
      if (n <= v.big_cap)
        for (int p = lane; p < n; p += WAVE) {
          const int k = v.mlist[o0 + p];
          int rk = 0;
          for (int q = 0; q < n; ++q) rk += v.mlist[o0 + q] < k;
          sSL[rk] = k;
        }
    } else {
      nbr_table(gc, nc, v.nx, v.nz, v.seg + (size_t)r * (nc + 1), v.away + (size_t)r * nc, &tb);
      int m = 0;
      if (lane == 0) {
        const int cap = v.big_cap;
        int cnt = 0;
        merge_members(gc, v.cell, v.id, &tb, [&](int k) {
          if (cnt < cap) sSL[cnt] = k;
          ++cnt;
        });
        m = cnt;
      }
      n = __shfl(m, 0);
    }
    __syncthreads();
  } else if (v.mlist) {
    /* the cell's member list, sorted by SIP index (the reference's mask order, all_or_nothing.py:812-829) */
    const int o0 = v.moff[gc];
    n = v.moff[gc + 1] - o0;
    fast = true;
    if (n <= CAP) {
      for (int p = lane; p < n; p += WAVE) sSlot[p] = v.mlist[o0 + p];
      __syncthreads();
      for (int p = lane; p < n; p += WAVE) {
        const int k = sSlot[p];
        int rk = 0;
        for (int q = 0; q < n; ++q) rk += sSlot[q] < k;
        sSL[rk] = k;
      }
      __syncthreads();
    }
  } else {
    n = -1;
    if constexpr (CAP > 0 && CAP <= 256) {
      /* the queue's storage is free until the first pass */
      if (v.mig_cnt)
        n = fast_stage<CAP, TAB>(P, v, gc, v.cc[(size_t)gc * CC_N + CC_T], v.cc[(size_t)gc * CC_N + CC_RHOW], sX,
                                 sU, sW, sR, sIND, sSL, (int *)sQkey, sQ);
    }
    staged = n >= 0;
    if (!staged) n = v.mig_cnt ? fast_members(v, gc, sSL, CAP, sSlot, sId, (int *)sW) : -1;
    fast = n >= 0;
  }
  if (!fast) {
    nbr_table(gc, nc, v.nx, v.nz, v.seg + (size_t)r * (nc + 1), v.away + (size_t)r * nc, &tb);
    n = gather_members(gc, v.cell, v.id, &tb, sSlot, sId, sSub, CAP);
  }
  if (n > (BIG ? v.big_cap : CAP)) {
    if (lane == 0) {
      if (!io.out_list) {
        raise_err(v.err, LM_ERR_CELL_CAP, gc);
      } else {
        const int q = atomicAdd(io.out_cnt, 1);
        io.out_list[q] = gc;
      }
    }
    return;
  }
  if (n < 2) return;
  if (lane == 0) sSeq = 0;
  for (int w = lane; w < (n + 31) / 32; w += WAVE) sChg[w] = 0u;
  __syncthreads();
  if (!fast) {
    rank_members(n, sId, sSub, &tb, sRank);
    __syncthreads();
    for (int m = lane; m < n; m += WAVE) sSL[sRank[m]] = sSlot[m];
    __syncthreads();
  }

  const double *C = v.cc + (size_t)gc * CC_N;
  const double T = C[CC_T], rho_w_T = C[CC_RHOW];
  for (int p = lane; p < n && !staged; p += WAVE) {
    const int k = sSL[p];
    sX[p] = v.xi[k];
    sU[p] = v.u[k];
    sW[p] = v.w[k];
    double R, ws, rho;
    lmp_R_ws_rho(P, v.solute, v.m_w[k], v.m_s[k], T, rho_w_T, R, ws, rho);
    sR[p] = R;
    sIND[p] = TAB ? kernel_index_fast(R, v) : 0;
  }
  /* quads per row: row i holds diagonals d = 2i+1 .. i+n-1, blocks (2i+1)>>2 .. (i+n-1)>>2; sQ[i] = quads
     of the rows before i (wave-wide exclusive scan, 64 rows at a time), sQ[n-1] = all quads */
  {
    int carry = 0;
    for (int r0 = 0; r0 < n; r0 += WAVE) {
      const int r = r0 + lane;
      const int c = r < n - 1 ? ((r + n - 1) >> 2) - ((2 * r + 1) >> 2) + 1 : 0;
      int incl = c;
      for (int off = 1; off < WAVE; off <<= 1) {
        const int o = __shfl_up(incl, off);
        if (lane >= off) incl += o;
      }
      if (r < n) sQ[r] = carry + incl - c;
      carry += __shfl(incl, WAVE - 1);
    }
  }
  __syncthreads();

  const CollCell cc = {sX, sR, sU, sW, sIND, sSL, n, gc, &sSeq, sB, T, rho_w_T, sChg};
  const bool use_tab = TAB;
  const long long roff = INJ ? v.rnd_off[gc] : 0;
  const double dtdV = v.dt_col_over_dV, pi = P.pi;
  /* the E bound of the row filter: the table's largest entry (unknown: no row filter) or the constant */
  const double Eb = TAB ? (v.E_max > 0.0 ? v.E_max : INFINITY) : v.E_const;
  row_bounds<CAP>(P, dtdV, Eb, n, sX, sR, sU, sW, sB, sTG, sBig, &sNbig, sIsBig, sQ[n - 1]);
  __syncthreads();
  const uint32_t key_run = (uint32_t)r, key_step = *v.col_step_dev, key_cell = (uint32_t)c;
  const uint32_t gkey = INJ ? 0u : lm_guard_key(v.seed, key_run, key_step, key_cell);
  const double *rho_g = v.dmxi; /* slot-indexed scratch: pre-step densities of the changed members */
  const int nQ = sQ[n - 1];
  long long c_single = 0, c_multi = 0, c_eq = 0;
  int k_next = 0;  /* first pair (row-major number) not yet decided */
  unsigned long long prev_event = 0; /* key of the last applied event */
  int n_events = 0;
  bool finished = false;

  bool build = true; /* the candidate queue is built from the row bounds by the first pass */
  bool fresh = true; /* row bounds describe the current state (no event yet) */
  const int spec_events = v.spec_events > 0 ? v.spec_events : SPEC_MAX_EVENTS;
  while (n_events < spec_events) {
    unsigned long long kq[NQ]; /* this pass's keys of queue entries lane + e * WAVE */
#pragma unroll
    for (int e = 0; e < NQ; ++e) kq[e] = ~0ull;
    int nq = 0;
    /* parallel pass: first firing pair with cnt >= k_next under the current state.  The first quad holding
       pair k_next: the pair after the last event (or pair 0); quad order is row-major pair order. */
    unsigned long long best = ~0ull;
    int i0 = 0, q_first = 0;
    if (k_next > 0) {
      int ni = (int)((prev_event >> 16) & 0xffff), nj = (int)(prev_event & 0xffff) + 1;
      if (nj >= n) {
        ++ni;
        nj = ni + 1;
      }
      if (ni >= n - 1) {
        finished = true;
        break;
      }
      i0 = ni;
      q_first = sQ[ni] + ((ni + nj) >> 2) - ((2 * ni + 1) >> 2);
    }
    if (INJ) {
      int i = i0; /* row of quad q: rows only grow with q, so each lane walks forward */
      for (int q = q_first + lane; q < nQ; q += WAVE) {
        while (sQ[i + 1] <= q) ++i;
        const int b = ((2 * i + 1) >> 2) + (q - sQ[i]);
        const unsigned long long kq = eval_quad<CAP, TAB, INJ>(P, v, cc, i, b, k_next, 0u, roff, key_run,
                                                               key_step, key_cell, true);
        best = kq < best ? kq : best;
      }
    } else {
      /* The candidate queue: quads that may hold a firing pair with cnt >= k_next.  Built by the first pass
         with the guard pass, row-parallel (lane l takes rows l, l + 64, ...; per quad the guard word and an
         integer compare of its valid pairs' smallest leading byte with the row's bound sTG: no Philox, no
         memory access) and the per-pair pass over the pairs of the listed members; then kept across events
         (an event changes two members only, see below).  A queue over QC quads hands the cell to the
         anti-diagonal schedule. */
      if (build) {
        if (lane == 0) sQn = 0;
        __syncthreads();
        const int nbig = sNbig;
        /* rows i0 .. n-2 in pairs (i0 + p, n - 2 - p): every lane gets about the same number of quads.  The
           first and last quad of each row (those with pairs outside [jmin, n-1]) are tested on their own; the
           interior quads of both rows form one run of quads per lane, so that the wave walks max(interior
           quads) and not max(first row) + max(second row) */
        const int n_rows = n - 1 - i0;
        for (int pr0 = 0; 2 * pr0 < n_rows; pr0 += WAVE) {
          const int pr = pr0 + lane;
          int ir[2], b0r[2], b1r[2], jr[2];
          uint32_t str[2];
#pragma unroll
          for (int side = 0; side < 2; ++side) {
            const int i = side == 0 ? i0 + pr : n - 2 - pr;
            ir[side] = i;
            b0r[side] = 0;
            b1r[side] = -1; /* no quads */
            jr[side] = n;
            str[side] = 0u;
            if (2 * pr >= n_rows || (side == 1 && i <= i0 + pr)) continue;
            if (sIsBig[i]) continue; /* the per-pair pass takes the rows of the list */
            const int base = i * n - i * (i + 1) / 2;
            const int jmin = (k_next - base + i + 1 > i + 1) ? k_next - base + i + 1 : i + 1;
            if (jmin >= n) continue;
            jr[side] = jmin;
            str[side] = sTG[i]; /* keep a quad when a valid pair's byte is <= st */
            b0r[side] = (i + jmin) >> 2;
            b1r[side] = (i + n - 1) >> 2;
          }
          /* the first and last quad of each row: bytes of pairs outside [jmin, n-1] set to 0xFF, which no bound
             below 255 keeps (each such quad holds a valid pair: j = jmin in the first, j = n-1 in the last) */
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const int side = e >> 1, i = ir[side];
            const int b = (e & 1) ? b1r[side] : b0r[side];
            if (b1r[side] < b0r[side] || ((e & 1) && b1r[side] == b0r[side])) continue;
            uint32_t g = lm_guard_word(gkey, (uint32_t)i, (uint32_t)(4 * b));
#pragma unroll
            for (int t = 0; t < 4; ++t) {
              const int j = 4 * b + t - i;
              if ((j < jr[side]) | (j >= n)) g |= 0xFFu << (8 * t);
            }
            const uint32_t st = str[side];
            /* a byte below st + 1 (exact word test, valid for bounds up to 128); st >= 127 rejects (almost)
               nothing: every quad holds a valid pair */
            const bool keep = (st >= 127u) | (((g - 0x01010101u * (st + 1u)) & ~g & 0x80808080u) != 0u);
            if (keep) {
              const int pos = atomicAdd(&sQn, 1);
              if (pos < QC) {
                sQueue[pos] = (i << 16) | b;
                sQkey[pos] = KEY_STALE;
              }
            }
          }
          /* interior quads (every pair valid): b0 + 1 .. b1 - 1 of the first row, then of the second; the guard
             word's argument key + ((i << 16) | b) * 0x9E3779B9 is linear in the run index t */
          const int na = b1r[0] - b0r[0] - 1 > 0 ? b1r[0] - b0r[0] - 1 : 0;
          const int nb2 = b1r[1] - b0r[1] - 1 > 0 ? b1r[1] - b0r[1] - 1 : 0;
          const int ntot = na + nb2;
          const uint32_t GR = 0x9E3779B9u;
          const uint32_t hA = gkey + (((uint32_t)ir[0] << 16) + (uint32_t)(b0r[0] + 1)) * GR;
          const uint32_t hB = gkey + (((uint32_t)ir[1] << 16) + (uint32_t)(b0r[1] + 1)) * GR - (uint32_t)na * GR;
          int tmax = ntot;
          for (int off = WAVE / 2; off > 0; off >>= 1) {
            const int o = __shfl_xor(tmax, off);
            tmax = o > tmax ? o : tmax;
          }
          /* four quads per iteration: their hash chains are independent and overlap in the pipeline */
          for (int t0 = 0; t0 < tmax; t0 += 4) {
            uint32_t gq[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
              const int t = t0 + u;
              gq[u] = lm_fmix32((t < na ? hA : hB) + (uint32_t)t * GR);
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
              const int t = t0 + u;
              const bool in_a = t < na;
              const uint32_t st = in_a ? str[0] : str[1], g = gq[u];
              const bool keep =
                  (t < ntot) & ((st >= 127u) | (((g - 0x01010101u * (st + 1u)) & ~g & 0x80808080u) != 0u));
              if (keep) {
                const int pos = atomicAdd(&sQn, 1);
                if (pos < QC) {
                  sQueue[pos] = in_a ? ((ir[0] << 16) | (b0r[0] + 1 + t)) : ((ir[1] << 16) | (b0r[1] + 1 + t - na));
                  sQkey[pos] = KEY_STALE;
                }
              }
            }
          }
        }
        /* per-pair pass: every pair of a listed member (its row and its column) with cnt >= k_next, tested with
           its guard byte alone; the quads of the pairs that can fire join the queue (a pair of two listed
           members may be queued twice: the pass takes the minimum) */
        for (int kb = 0; kb < nbig; ++kb) {
          const int bm = sBig[kb];
          for (int o = lane; o < n; o += WAVE) {
            if (o == bm) continue;
            const int i = o < bm ? o : bm, j = o < bm ? bm : o;
            if (pair_cnt(i, j, n) < k_next) continue;
            const int d = i + j;
            const uint32_t G = lm_guard_byte(lm_guard_word(gkey, (uint32_t)i, (uint32_t)d), (uint32_t)d);
            if (pair_may_fire8(P, dtdV, Eb, sX, sR, sU, sW, i, j, G)) {
              const int pos = atomicAdd(&sQn, 1);
              if (pos < QC) {
                sQueue[pos] = (i << 16) | (d >> 2);
                sQkey[pos] = KEY_STALE;
              }
            }
          }
        }
        __syncthreads();
        build = false;
      }
      nq = sQn;
      /* too many candidates (the bounds reject little and events are many): the anti-diagonal schedule, exact
         from any k_next, finishes the cell */
      if (nq > QC) break;
      /* each lane evaluates queue entries lane + e * WAVE; the row filter of eval_quad needs bounds of the
         current state (fresh: no event since the queue was built) */
#pragma unroll
      for (int e = 0; e < NQ; ++e) {
        const int q = lane + e * WAVE;
        if (q < nq) {
          kq[e] = sQkey[q];
          if (kq[e] == KEY_STALE) {
            const int pk = sQueue[q], qi = pk >> 16, qb = pk & 0xffff;
            kq[e] = eval_quad<CAP, TAB, INJ>(P, v, cc, qi, qb, k_next,
                                             lm_guard_word(gkey, (uint32_t)qi, (uint32_t)(4 * qb)), roff, key_run,
                                             key_step, key_cell, fresh);
            sQkey[q] = kq[e];
          }
          best = kq[e] < best ? kq[e] : best;
        }
      }
      __syncthreads();
    }
    for (int off = WAVE / 2; off > 0; off >>= 1) {
      const unsigned long long o = __shfl_xor(best, off);
      best = o < best ? o : best;
    }
    if (best == ~0ull) {
      finished = true;
      break;
    }
    const int ei = (int)((best >> 16) & 0xffff), ej = (int)(best & 0xffff);
    const int ecnt = (int)(best >> 32);
    if (lane == 0) {
      /* the table entry and the two members' masses and densities in one round of loads */
      const double Ec = use_tab ? v.E[sIND[ei] * v.nb + sIND[ej]] : v.E_const;
      const PairMass M = load_pair(P, v, cc, rho_g, ei, ej);
      const int kind = apply_event(P, v, cc, M, ei, ej, pair_p(P, v, cc, ei, ej, Ec));
      if (kind == 1)
        ++c_multi;
      else
        ++c_single;
      if (kind == 3) ++c_eq;
    }
    __threadfence_block();
    __syncthreads();
    ++n_events;
    k_next = ecnt + 1;
    prev_event = best;
    if (n_events < spec_events) {
      if (INJ) {
        /* the event may have grown a radius: fresh row bounds for the next pass */
        row_bounds<CAP>(P, dtdV, Eb, n, sX, sR, sU, sW, sB, sTG, sBig, &sNbig, sIsBig, sQ[n - 1]);
        __syncthreads();
      } else {
        /*
         * Keep the queue.  The event changed members ei and ej only, so every pair without them keeps its p and
         * its uniform: a quad that held no firing pair (key ~0) holds none now except through pairs of ei or
         * ej, and a quad with a firing pair is kept with its key (evaluated again only if it holds a pair of ei
         * or ej).  The undecided pairs of ei and ej -- (ei, y) and (ej, y) for y > ej, (x, ej) for ei < x < ej --
         * are tested again with their guard bytes on the new state and their quads queued.
         */
        int pk[NQ];
        bool al[NQ];
        unsigned long long bal[NQ];
        const unsigned long long below = (1ull << lane) - 1ull;
#pragma unroll
        for (int e = 0; e < NQ; ++e) {
          const int q = lane + e * WAVE;
          pk[e] = q < nq ? sQueue[q] : 0;
          al[e] = (q < nq) & (kq[e] != ~0ull);
          /* a kept quad's key stays exact unless the quad holds a pair of ei or ej (pairs (i, 4b - i .. 4b + 3
             - i)): its pairs before the key still do not fire and the key's pair still fires */
          const int qi = pk[e] >> 16, qj = 4 * (pk[e] & 0xffff) - qi;
          const bool t = (qi == ei) | (qi == ej) | ((unsigned)(ei - qj) < 4u) | ((unsigned)(ej - qj) < 4u);
          if (t) kq[e] = KEY_STALE;
          bal[e] = __ballot(al[e]);
        }
        __syncthreads();
        int kept_n = 0;
#pragma unroll
        for (int e = 0; e < NQ; ++e) {
          if (al[e]) {
            const int q = kept_n + __popcll(bal[e] & below);
            sQueue[q] = pk[e];
            sQkey[q] = kq[e];
          }
          kept_n += __popcll(bal[e]);
        }
        if (lane == 0) sQn = kept_n;
        __syncthreads();
        const int la = n - 1 - ej, lb = ej - ei - 1;
        for (int t = lane; t < 2 * la + lb; t += WAVE) {
          int i, j;
          if (t < la) {
            i = ei;
            j = ej + 1 + t;
          } else if (t < la + lb) {
            i = ei + 1 + (t - la);
            j = ej;
          } else {
            i = ej;
            j = ej + 1 + (t - la - lb);
          }
          const int d = i + j;
          const uint32_t G = lm_guard_byte(lm_guard_word(gkey, (uint32_t)i, (uint32_t)d), (uint32_t)d);
          if (pair_may_fire8(P, dtdV, Eb, sX, sR, sU, sW, i, j, G)) {
            const int pos = atomicAdd(&sQn, 1);
            if (pos < QC) {
              sQueue[pos] = (i << 16) | (d >> 2);
              sQkey[pos] = KEY_STALE;
            }
          }
        }
        __syncthreads();
        fresh = false;
        if (sQn > QC) break; /* the anti-diagonal schedule takes over */
      }
    }
  }

  if (!finished) {
    /* many events or candidates: anti-diagonal schedule over the pairs with cnt >= k_next */
    int tag0 = -1, tag1 = -1;
    lm_u32x4 blk0, blk1;
    blk0.v[0] = blk0.v[1] = blk0.v[2] = blk0.v[3] = 0;
    blk1 = blk0;
    for (int d = 1; d <= 2 * n - 3; ++d) {
      const int i_lo = d - (n - 1) > 0 ? d - (n - 1) : 0;
      const int i_hi = (d - 1) >> 1;
      for (int i = i_lo + ((lane - i_lo) & (WAVE - 1)); i <= i_hi; i += WAVE) {
        const int j = d - i;
        const int cnt = pair_cnt(i, j, n);
        if (cnt < k_next) continue;
        uint32_t aw = 0, G = 0;
        double rnd_inj = 0.0;
        if (INJ) {
          rnd_inj = rnd_at(v, roff + cnt);
        } else {
          G = lm_guard_byte(lm_guard_word(gkey, (uint32_t)i, (uint32_t)d), (uint32_t)d);
          /* the pair's bound on the current state against its guard byte: most pairs end here */
          if (!pair_may_fire8(P, dtdV, Eb, sX, sR, sU, sW, i, j, G)) continue;
          const int tag = (i << 16) | (d >> 2);
          const bool cls = (i >> 6) & 1;
          if (tag != (cls ? tag1 : tag0)) {
            const lm_u32x4 o = lm_pair_block(v.seed, key_run, key_step, key_cell, (uint32_t)i, (uint32_t)d, 0);
            if (cls) {
              tag1 = tag;
              blk1 = o;
            } else {
              tag0 = tag;
              blk0 = o;
            }
          }
          aw = lm_word(cls ? blk1 : blk0, (uint32_t)d & 3);
        }
        const double Ec = use_tab ? v.E[sIND[i] * v.nb + sIND[j]] : v.E_const;
        const double p_crit = pair_p(P, v, cc, i, j, Ec);
        bool fires;
        if (p_crit > 1.0)
          fires = true;
        else if (INJ)
          fires = p_crit > rnd_inj;
        else
          fires = pair_fires(p_crit, G, aw, v.seed, key_run, key_step, key_cell, (uint32_t)i, (uint32_t)d);
        if (fires) {
          const int kind = apply_event(P, v, cc, load_pair(P, v, cc, rho_g, i, j), i, j, p_crit);
          if (kind == 1)
            ++c_multi;
          else
            ++c_single;
          if (kind == 3) ++c_eq;
        }
      }
      __threadfence_block();
      __syncthreads();
    }
  }

  if (n_events > 0 || !finished)
    for (int m = lane; m < n; m += WAVE) v.xi[sSL[m]] = sX[m];
  for (int off = WAVE / 2; off > 0; off >>= 1) {
    c_single += __shfl_down(c_single, off);
    c_multi += __shfl_down(c_multi, off);
    c_eq += __shfl_down(c_eq, off);
  }
  if (lane == 0) {
    if (v.ev_mode == 1) v.ev_count[gc] = sSeq;
    if (c_single) atomicAdd((unsigned long long *)&v.no_cols[2 * r], (unsigned long long)c_single);
    if (c_multi) atomicAdd((unsigned long long *)&v.no_cols[2 * r + 1], (unsigned long long)c_multi);
    if (c_eq) atomicAdd((unsigned long long *)&v.n_equal_xi[r], (unsigned long long)c_eq);
    if (v.col_cost) {
      const unsigned long long dt = wall_clock64() - t_cell;
      v.col_cost[gc] = dt < 0xFFFFFFFFull ? (uint32_t)dt : 0xFFFFFFFFu;
    }
  }
}

/* main pass: one block per cell or persistent blocks over the work list; overflow passes: the queued cells,
   grid-strided */
/* waves per SIMD the collision passes are compiled for: the 256-member pass's LDS holds 10 workgroups per compute
   unit, so 3 waves per SIMD (168 VGPRs) use it where 2 (its register count unbounded) would not */
__host__ __device__ constexpr int col_wpe(int cap) { return cap == 256 ? 3 : 4; }

template <int CAP, bool TAB, bool INJ>
__global__ __launch_bounds__(WAVE) __attribute__((amdgpu_waves_per_eu(col_wpe(CAP)))) void k_collision(
    KView v, int mode, int chunk, ColLists io) {
  CellSource src = cell_source(mode, chunk, 0);
  for (;;) {
    const int gc = next_cell(v, src, io.in_list, io.in_cnt, io.in_grab);
    if (gc < 0) return;
    collide_cell<CAP, TAB, INJ>(v, gc, io);
  }
}

/* member counts -> pair counts for the injected stream offsets */
__global__ __launch_bounds__(WAVE) void k_count_pairs(KView v) {
  const int gc = blockIdx.x, nc = v.nc, r = gc / nc;
  __shared__ NbrTab tb;
  nbr_table(gc, nc, v.nx, v.nz, v.seg + (size_t)r * (nc + 1), v.away + (size_t)r * nc, &tb);
  const int n = count_members(gc, v.cell, &tb);
  if (threadIdx.x == 0) v.rnd_off[gc] = (long long)n * (n - 1) / 2;
}

/* exclusive scan of rnd_off[0..L) in one block, offset by the running stream position rnd_off[L] */
__global__ __launch_bounds__(1024) void k_scan_pairs(long long *a, int L) {
  __shared__ long long part[1024];
  const int t = threadIdx.x;
  const int per = (L + 1023) / 1024;
  const int b = t * per, e = b + per < L ? b + per : L;
  long long s = 0;
  for (int k = b; k < e; ++k) s += a[k];
  part[t] = s;
  __syncthreads();
  if (t == 0) {
    long long acc = a[L];
    for (int q = 0; q < 1024; ++q) {
      const long long x = part[q];
      part[q] = acc;
      acc += x;
    }
    a[L] = acc;
  }
  __syncthreads();
  long long acc = part[t];
  for (int k = b; k < e; ++k) {
    const long long x = a[k];
    a[k] = acc;
    acc += x;
  }
}

/* ------------------------------------------------------------------------ */
/* particle substep: one thread per SIP slot                                */
/* ------------------------------------------------------------------------ */
/*
 * Between two reindexes the members of a cell are looked up in the 3x3 home
 * segments around it, so a SIP may drift at most one cell (periodic in x) from
 * the home cell of its slot k; the home cell is the segment that holds k.
 */
__device__ static inline bool near_home(const KView &v, int r, long long k, int ii, int jj, int *home) {
  const int32_t *s = v.seg + (size_t)r * (v.nc + 1);
  int lo = 0, hi = v.nc - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (s[mid] <= k)
      lo = mid;
    else
      hi = mid - 1;
  }
  *home = lo;
  const int hx = lo / v.nz, hz = lo % v.nz;
  const int ddx = ((ii - hx) % v.nx + v.nx) % v.nx;
  const int ddz = jj - hz;
  return (ddx <= 1 || ddx == v.nx - 1) && ddz >= -1 && ddz <= 1;
}

/*
 * Condensation update with its local Jacobian (adjoint tape): the Newton of
 * lmp_newton_dm on dual values seeded in (m_w, m_s, T, p, S); the cell
 * constants carry their derivatives in T and p from ccd.  The value part is
 * the double computation itself, so dm is the step's dm bit for bit.
 * PART 0 carries all five directions (LmDual<5>); PART 1 only (m_w, m_s) and
 * PART 2 only (T, p, S), so the tape can be written by two lighter passes
 * (each direction's derivative arithmetic is independent of the others, so
 * both ways write the same tape bit for bit).  PART 2 writes only its three
 * columns; the others write the rest of the slot's tape entry.
 */
template <int SOLUTE, int PART>
__device__ static double newton_dm_tape(const LmPhys &P, const KView &v, long long k, long long n_tot, int gc,
                                        double dt_sub, double m_w, double m_s, double xi, const double *C) {
  constexpr int N = PART == 0 ? 5 : (PART == 1 ? 2 : 3);
  constexpr int oT = PART == 0 ? 2 : 0; /* direction of T (then p, S) when carried */
  using DN = LmDual<N>;
  DN dm;
  if constexpr (PART == 1) {
    /* (m_w, m_s) columns: the cell constants enter as doubles (no derivative arithmetic on them) */
    const DN mw = DN::seed(m_w, 0), ms = DN::seed(m_s, 1);
    DN R, w_s, rho;
    lmp_R_ws_rho(P, SOLUTE, mw, ms, C[CC_T], C[CC_RHOW], R, w_s, rho);
    const DN sigma_p = lmp_sigma_solution(P, SOLUTE, w_s, C[CC_SIGW]);
    dm = lmp_newton_dm(P, SOLUTE, dt_sub, v.no_iter, mw, ms, w_s, R, C[CC_T], rho, C, sigma_p);
  } else {
    const double *G = v.ccd + (size_t)gc * 12;
    DN cx[CC_N];
    const int ent[6] = {CC_LALPHA, CC_LBETA, CC_C1, CC_C2, CC_RHOW, CC_SIGW};
    cx[CC_T] = DN::seed(C[CC_T], oT);
    cx[CC_S] = DN::seed(C[CC_S], oT + 2);
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      DN x(C[ent[q]]);
      x.d[oT] = G[2 * q];
      x.d[oT + 1] = G[2 * q + 1];
      cx[ent[q]] = x;
    }
    if constexpr (PART == 2) {
      /* (T, p, S) columns: the masses enter as doubles */
      DN R, rho;
      double w_s;
      lmp_R_ws_rho(P, SOLUTE, m_w, m_s, cx[CC_T], cx[CC_RHOW], R, w_s, rho);
      const DN sigma_p = lmp_sigma_solution(P, SOLUTE, w_s, cx[CC_SIGW]);
      dm = lmp_newton_dm(P, SOLUTE, dt_sub, v.no_iter, m_w, m_s, w_s, R, cx[CC_T], rho, cx, sigma_p);
    } else {
      const DN mw = DN::seed(m_w, 0), ms = DN::seed(m_s, 1);
      DN R, w_s, rho;
      lmp_R_ws_rho(P, SOLUTE, mw, ms, cx[CC_T], cx[CC_RHOW], R, w_s, rho);
      const DN sigma_p = lmp_sigma_solution(P, SOLUTE, w_s, cx[CC_SIGW]);
      dm = lmp_newton_dm(P, SOLUTE, dt_sub, v.no_iter, mw, ms, w_s, R, cx[CC_T], rho, cx, sigma_p);
    }
  }
  const size_t n = (size_t)n_tot;
  double *t = v.tape_part + k;
  if constexpr (PART != 2) {
    t[0] = 1.0 + dm.d[0];
    t[n] = dm.d[1];
  }
  if constexpr (PART != 1) {
#pragma unroll
    for (int q = 0; q < 3; ++q) t[(2 + q) * n] = dm.d[oT + q];
  }
  if constexpr (PART != 2) {
    t[5 * n] = dm.v;
    t[6 * n] = xi;
    v.tape_part_i[k] = (int)(v.run_base[gc / v.nc] + v.id[k]);
    v.tape_part_i[n + k] = gc;
  }
  return dm.v;
}

/* one SIP slot of the particle substep */
template <int SOLUTE, bool POS_ONLY, bool TAPE, int TPART = 0>
__device__ __forceinline__ static void particle_update(const KView &v, long long k, long long n_tot, double dt_sub,
                                                       double dt_pos) {
  const LmPhys &P = c_P;
  const int gc = v.cell[k];
  if (gc < 0) {
    v.cell_next[k] = gc;
    if (TAPE) {
      v.tape_part_i[k] = (int)(v.run_base[slot_run(v.run_base, v.n_runs, k)] + v.id[k]);
      v.tape_part_i[n_tot + k] = -1;
    }
    return;
  }
  const int nc = v.nc;
  const int r = gc / nc, c = gc % nc;
  const int i = c / v.nz, j = c % v.nz;
  double x = v.x[k], z = v.z[k], u = v.u[k], w = v.w[k];
  /* quotients by the cell sizes through their reciprocals (lm_div_y: the same values) */
  const bool fd = lm_div_range(v.dx) & lm_div_range(v.dz);
  const double ydx = lm_rcp_nr(v.dx), ydz = lm_rcp_nr(v.dz);
  const auto over_dx = [&](double a) { return fd ? lm_div_y(a, v.dx, ydx) : a / v.dx; };
  const auto over_dz = [&](double a) { return fd ? lm_div_y(a, v.dz, ydz) : a / v.dz; };
  if (!POS_ONLY) {
    const double *C = v.cc + (size_t)gc * CC_N;
    const double T = C[CC_T];
    double m_w = v.m_w[k];
    const double m_s = v.m_s[k], xi = v.xi[k];
    double R, w_s, rho;
    double dm;
    if (TAPE) {
      dm = newton_dm_tape<SOLUTE, TPART>(P, v, k, n_tot, gc, dt_sub, m_w, m_s, xi, C);
    } else {
      dm = lmp_condensation_dm(P, SOLUTE, dt_sub, v.no_iter, m_w, m_s, T, C);
    }
    m_w += dm;
    v.m_w[k] = m_w;
    v.dmxi[k] = dm * xi;
    /* integration.py:1075: AS radius and density after the mass update */
    lmp_R_ws_rho(P, 0, m_w, m_s, T, C[CC_RHOW], R, w_s, rho);
    /* grid.py:257-323 with rel_pos of integration.py:545-549 */
    const double wx = over_dx(x - v.x_min) - (double)i;
    const double wy = over_dz(z - v.z_min) - (double)j;
    const double *U = v.grid_vel + (size_t)r * v.vel_stride, *V = U + (size_t)(v.nx + 1) * (v.nz + 1);
    double uf, vf;
    lmp_interp_vel(i, j, wx, wy, U, V, v.nx, v.nz, uf, vf);
    /* integration.py:699-718 */
    const double mu = C[CC_MU], rho_f = C[CC_RHOF];
    const double dv0 = uf - u, dv1 = vf - w;
    const double vel_dev = sqrt(dv0 * dv0 + dv1 * dv1);
    const double Re = 2.0E-6 * rho_f * R * vel_dev / mu;
    double k_dt = 4.5E12 * dt_sub * mu / (rho * R * R);
    if (Re > 0.5) {
      if (Re < 1000.0)
        k_dt *= (1 + 0.15 * lm_pow(Re, 0.687));
      else
        k_dt *= 0.0183333333333 * Re;
    }
    lmp_div2(u + k_dt * uf, w + k_dt * vf - dt_sub * v.g_set, 1.0 + k_dt, u, w);
  }
  /* integration.py:552-630 */
  x += dt_pos * u;
  z += dt_pos * w;
  x = lm_pymod(x, v.x_max);
  z = lm_np_min(z, v.z_max * 0.999999);
  int gc_new;
  if (z <= v.z_min) {
    z = v.z_min - 0.01 * v.dz;
    u = 0.0;
    w = 0.0;
    gc_new = -(2 + i); /* inactive; keeps cells[0] as the reference does */
    const int q = atomicAdd(&v.rem_cnt[r], 1);
    if (q < v.rem_cap)
      v.rem_slot[(size_t)r * v.rem_cap + q] = (int)k;
    else
      raise_err(v.err, LM_ERR_REM_CAP, r);
  } else {
    /* integration.py:545-546 */
    const long long ii = (long long)floor(over_dx(x));
    const long long jj = (long long)floor(over_dz(z));
    if (ii < 0 || ii >= v.nx || jj < 0 || jj >= v.nz) {
      raise_err(v.err, LM_ERR_OUT_OF_DOMAIN, (int)k);
      gc_new = gc;
    } else {
      gc_new = r * nc + (int)ii * v.nz + (int)jj;
      if (gc_new != gc) { /* rare: the SIP changes cell */
        int home;
        if (!near_home(v, r, k, (int)ii, (int)jj, &home)) {
          raise_err(v.err, LM_ERR_LOST_SIP, (int)k);
        } else if (home != (int)ii * v.nz + (int)jj) {
          atomicOr(&v.away[(size_t)r * nc + home], dir_bit(home / v.nz, home % v.nz, (int)ii, (int)jj, v.nx));
          if (v.mig_cnt) {
            const int q = atomicAdd(&v.mig_cnt[gc_new], 1);
            if (q < v.mig_cap) v.mig_slot[(size_t)gc_new * v.mig_cap + q] = (int)k;
          }
        }
      }
    }
  }
  v.x[k] = x;
  v.z[k] = z;
  v.u[k] = u;
  v.w[k] = w;
  v.cell_next[k] = gc_new;
}

/*
 * A block covers blockDim.x * SPT consecutive slots, SPT SIPs per thread (the block size and SPT are launch
 * parameters an autotuner picks; SPT is a template parameter because a run-time loop costs the kernel a
 * quarter of its occupancy).
 */
template <int SOLUTE, bool POS_ONLY, bool TAPE, int SPT>
__global__ __launch_bounds__(TAPE ? TAPE_BLOCK : PART_BLOCK_MAX) void k_particles(KView v, long long n_tot,
                                                                                  double dt_sub, double dt_pos) {
  const long long base = (long long)blockIdx.x * blockDim.x * SPT + threadIdx.x;
#pragma unroll
  for (int s = 0; s < SPT; ++s) {
    const long long k = base + (long long)s * blockDim.x;
    if (k < n_tot) particle_update<SOLUTE, POS_ONLY, TAPE>(v, k, n_tot, dt_sub, dt_pos);
  }
}

/*
 * The particle substep (one SIP per thread) compiled for WPE waves per SIMD: the register budget drops
 * from the 101 VGPRs of k_particles (4 waves) to 96, 80 or 64 at the price of a few spills; an autotuner
 * picks (LmEngine.part_wpe).
 */
template <int SOLUTE, int WPE>
__global__ __launch_bounds__(PART_BLOCK_MAX) __attribute__((amdgpu_waves_per_eu(WPE, WPE))) void k_particles_occ(
    KView v, long long n_tot, double dt_sub, double dt_pos) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n_tot) particle_update<SOLUTE, false, false>(v, k, n_tot, dt_sub, dt_pos);
}

template <int SOLUTE>
static void launch_particles_occ(int wpe, dim3 grid, dim3 block, hipStream_t st, const KView &v, long long n,
                                 double dt_sub, double dt_pos) {
  if (wpe == 5)
    hipLaunchKernelGGL((k_particles_occ<SOLUTE, 5>), grid, block, 0, st, v, n, dt_sub, dt_pos);
  else if (wpe == 6)
    hipLaunchKernelGGL((k_particles_occ<SOLUTE, 6>), grid, block, 0, st, v, n, dt_sub, dt_pos);
  else
    hipLaunchKernelGGL((k_particles_occ<SOLUTE, 8>), grid, block, 0, st, v, n, dt_sub, dt_pos);
}

/*
 * Taped substep in two passes (LmEngine.tape_mode 0): k_tape_tps writes the (T, p, S) columns of each
 * slot's condensation Jacobian from the pre-update state, then k_particles_tape<S, 1> runs the substep
 * with the (m_w, m_s) columns.  Two narrower dual types need far fewer registers than one LmDual<5>.
 */
template <int SOLUTE, int WPE>
__global__ __launch_bounds__(TAPE_BLOCK) __attribute__((amdgpu_waves_per_eu(WPE, WPE))) void k_tape_tps(
    KView v, long long n_tot, double dt_sub) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n_tot) return;
  const int gc = v.cell[k];
  if (gc < 0) return;
  newton_dm_tape<SOLUTE, 2>(c_P, v, k, n_tot, gc, dt_sub, v.m_w[k], v.m_s[k], v.xi[k], v.cc + (size_t)gc * CC_N);
}

template <int SOLUTE, int PART, int WPE>
__global__ __launch_bounds__(TAPE_BLOCK) __attribute__((amdgpu_waves_per_eu(WPE, WPE))) void k_particles_tape(
    KView v, long long n_tot, double dt_sub, double dt_pos) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n_tot) particle_update<SOLUTE, false, true, PART>(v, k, n_tot, dt_sub, dt_pos);
}

/* water_removed[r] += xi*m_w of this substep's removed SIPs, ascending id (integration.py:620-628) */
#define REM_LDS LM_REM_CAP_MAX
__global__ __launch_bounds__(WAVE) void k_water(KView v) {
  const int r = blockIdx.x;
  __shared__ int sId[REM_LDS], sSlot[REM_LDS], sRank[REM_LDS], sOrd[REM_LDS];
  const int n0 = v.rem_cnt[r];
  if (n0 == 0) return;
  /* rem_slot holds rem_cap slots per run (1 <= rem_cap <= REM_LDS, checked on the host); particle_update
     flagged LM_ERR_REM_CAP for removals beyond it, and only the recorded ones are read */
  const int lim = v.rem_cap < REM_LDS ? v.rem_cap : REM_LDS;
  const int n = n0 < lim ? n0 : lim;
  if (n0 > lim && threadIdx.x == 0) raise_err(v.err, LM_ERR_REM_CAP, r);
  for (int t = threadIdx.x; t < n; t += WAVE) {
    const int k = v.rem_slot[(size_t)r * v.rem_cap + t];
    sSlot[t] = k;
    sId[t] = v.id[k];
  }
  __syncthreads();
  for (int m = threadIdx.x; m < n; m += WAVE) {
    const int idm = sId[m];
    int rk = 0;
    for (int q = 0; q < n; ++q) rk += (sId[q] < idm);
    sRank[m] = rk;
  }
  __syncthreads();
  for (int t = threadIdx.x; t < n; t += WAVE) sOrd[sRank[t]] = sSlot[t];
  __syncthreads();
  if (threadIdx.x == 0) {
    double acc = v.water_removed[r];
    for (int p = 0; p < n; ++p) {
      const int k = sOrd[p];
      acc += v.xi[k] * v.m_w[k];
    }
    v.water_removed[r] = acc;
    v.rem_cnt[r] = 0;
  }
}

/*
 * Delta m_l per cell: sum of dm*xi over the cell's members in ascending id
 * (integration.py:915-938), then *1e-18.  CAP bounds the members held in LDS;
 * a larger cell is appended to out_list (counter *out_cnt) for a pass of
 * larger capacity, or, with no out_list, summed by one lane straight from the
 * segments.
 */
template <int CAP>
__device__ static void dml_cell(const KView &v, int gc, int32_t *out_list, int32_t *out_cnt) {
  const int nc = v.nc;
  __syncthreads(); /* the LDS of the previous cell of this block is free */
  const int r = gc / nc;
  __shared__ int sSlot[CAP], sIdRank[2 * CAP];
  __shared__ unsigned char sSub[CAP];
  __shared__ NbrTab tb;
  __shared__ int sMig[3 * LM_MIG_CAP_MAX];
  int *sId = sIdRank, *sRank = sIdRank + CAP;
  double *sVal = (double *)sIdRank; /* reuses the id/rank storage once the ranks are in registers */
  constexpr int PER = CAP / WAVE;
  if (v.mig_cnt) {
    /*
     * Complete migrant list: the sum over the home segment in slot order (= id order), non-members adding
     * +0.0 (exact: a sum that starts at +0.0 is never -0.0, and s + 0.0 = s), merged by id with the list's
     * members (each slot once).  Each home slot's cell, value and id load in one round, the migrants' in
     * the next.
     */
    const int mc = v.mig_cnt[gc];
    const int32_t *seg = v.seg + (size_t)r * (nc + 1);
    const int c = gc % nc, s0 = seg[c], T0 = seg[c + 1] - s0;
    if (mc <= v.mig_cap && T0 + mc <= CAP) {
      const int lane = threadIdx.x;
      int *sMk = sMig, *sMid = sMig + LM_MIG_CAP_MAX;
      double ov[PER];
      int oid[PER];
#pragma unroll
      for (int q = 0; q < PER; ++q) {
        const int t = lane + q * WAVE;
        ov[q] = 0.0;
        oid[q] = 0;
        if (t < T0) {
          const int k = s0 + t;
          const double d = v.dmxi[k];
          ov[q] = v.cell[k] == gc ? d : 0.0;
          oid[q] = v.id[k];
        }
      }
      int mk = -1, mid_ = 0;
      double mval = 0.0;
      bool mem = false;
      if (lane < mc) {
        mk = v.mig_slot[(size_t)gc * v.mig_cap + lane];
        mem = v.cell[mk] == gc;
        mval = v.dmxi[mk];
        mid_ = v.id[mk];
        sMk[lane] = mk;
      }
      __syncthreads();
      if (mem)
        for (int e = 0; e < lane; ++e)
          if (sMk[e] == mk) mem = false;
      const unsigned long long bal = __ballot(mem);
      const int nm = __popcll(bal);
      int opos[PER];
      if (nm > 0) {
        if (mem) sMid[__popcll(bal & lane_mask_lt())] = mid_;
        __syncthreads();
        /* home entry: its index plus the migrants with smaller ids; migrant: its rank among them plus the
           home entries with smaller ids (ascending in the segment) */
#pragma unroll
        for (int q = 0; q < PER; ++q) {
          int lo = 0; /* sMid is in list order: count all */
          if (lane + q * WAVE < T0)
            for (int e = 0; e < nm; ++e) lo += sMid[e] < oid[q];
          opos[q] = lane + q * WAVE + lo;
        }
        int rk = 0;
        if (mem)
          for (int e = 0; e < nm; ++e) rk += sMid[e] < mid_;
        __syncthreads();
        /* home ids to LDS for the migrants' lower bounds */
#pragma unroll
        for (int q = 0; q < PER; ++q)
          if (lane + q * WAVE < T0) sSlot[lane + q * WAVE] = oid[q];
        __syncthreads();
        if (mem) {
          int lo = 0, hi = T0;
          while (lo < hi) {
            const int m2 = (lo + hi) >> 1;
            if (sSlot[m2] < mid_)
              lo = m2 + 1;
            else
              hi = m2;
          }
          rk += lo;
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < PER; ++q)
          if (lane + q * WAVE < T0) sVal[opos[q]] = ov[q];
        if (mem) sVal[rk] = mval;
      } else {
#pragma unroll
        for (int q = 0; q < PER; ++q)
          if (lane + q * WAVE < T0) sVal[lane + q * WAVE] = ov[q];
      }
      __syncthreads();
      if (lane == 0) {
        const int tot = T0 + nm;
        double sum = 0.0;
        for (int p = 0; p < tot; ++p) sum += sVal[p];
        v.dml[gc] = sum * 1.0E-18;
      }
      return;
    }
  }
  nbr_table(gc, nc, v.nx, v.nz, v.seg + (size_t)r * (nc + 1), v.away + (size_t)r * nc, &tb);
  const int n = gather_members(gc, v.cell, v.id, &tb, sSlot, sId, sSub, CAP);
  if (n > CAP) {
    if (threadIdx.x == 0) {
      if (!out_list) {
        /* over every LDS capacity: one lane sums in ascending id straight from the segments */
        double sum = 0.0;
        merge_members(gc, v.cell, v.id, &tb, [&](int k) { sum += v.dmxi[k]; });
        v.dml[gc] = sum * 1.0E-18;
      } else {
        const int q = atomicAdd(out_cnt, 1);
        out_list[q] = gc;
      }
    }
    return;
  }
  __syncthreads();
  rank_members(n, sId, sSub, &tb, sRank);
  __syncthreads();
  int rk[PER];
  double val[PER];
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int m = threadIdx.x + q * WAVE;
    rk[q] = m < n ? sRank[m] : 0;
    val[q] = m < n ? v.dmxi[sSlot[m]] : 0.0;
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < PER; ++q)
    if (threadIdx.x + q * WAVE < n) sVal[rk[q]] = val[q];
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int p = 0; p < n; ++p) s += sVal[p];
    v.dml[gc] = s * 1.0E-18;
  }
}

/* cells from the pass's source (PASS_OVF: in_list[0 .. *in_cnt)); larger cells on to out_list (NULL: summed by
   one lane) */
template <int CAP>
__global__ __launch_bounds__(WAVE) __attribute__((amdgpu_waves_per_eu(8))) void k_dml(KView v, int mode, int chunk,
                                                                                     const int32_t *in_list,
                                                                                     const int32_t *in_cnt,
                                                                                     int32_t *out_list,
                                                                                     int32_t *out_cnt) {
  CellSource src = cell_source(mode, chunk, 1);
  for (;;) {
    const int gc = next_cell(v, src, in_list, in_cnt);
    if (gc < 0) return;
    dml_cell<CAP>(v, gc, out_list, out_cnt);
  }
}

/*
 * Delta m_l of two cells per wavefront, 32 lanes each, for cells whose migrant list is complete and whose home
 * segment and list hold at most DML2_CAP entries: dml_cell's sum over the home segment in slot order (= id
 * order, non-members adding +0.0) merged by id with the list's members (each slot once), with both cells'
 * loads, ranks and sequential sums in the same instructions.  Other cells go to out_list (counter *out_cnt)
 * for dml_cell.  Needs the migrant lists (mig_cnt) and mig_cap <= 32.
 */
#define DML2_CAP 224
#define DML2_L 32
__global__ __launch_bounds__(WAVE) __attribute__((amdgpu_waves_per_eu(8))) void k_dml2(KView v, int32_t *out_list,
                                                                                      int32_t *out_cnt) {
  constexpr int L = DML2_L, G = WAVE / DML2_L, PER = DML2_CAP / DML2_L;
  __shared__ double sVal[G][DML2_CAP]; /* home ids (as int) until the ranks are known, then the ordered values */
  __shared__ int sMk[G][L], sMid[G][L];
  const int lane = threadIdx.x, g = lane / L, gl = lane % L;
  int *sHid = (int *)sVal[g];
  const int nc = v.nc, ncell = v.n_runs * nc;
  const int gc = (int)blockIdx.x * G + g;
  int mc = 0, s0 = 0, T0 = 0;
  bool fast = false;
  if (gc < ncell) {
    const int r = gc / nc, c = gc % nc;
    const int32_t *seg = v.seg + (size_t)r * (nc + 1);
    mc = v.mig_cnt[gc];
    s0 = seg[c];
    T0 = seg[c + 1] - s0;
    fast = mc <= v.mig_cap && T0 + mc <= DML2_CAP;
    if (!fast && gl == 0) {
      const int q = atomicAdd(out_cnt, 1);
      out_list[q] = gc;
    }
  }
  /* home slots: value (+0.0 for a slot whose SIP left) and id; the list's entries: one per lane */
  double ov[PER];
  int oid[PER];
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int t = gl + q * L;
    ov[q] = 0.0;
    oid[q] = 0;
    if (fast && t < T0) {
      const int k = s0 + t;
      const double d = v.dmxi[k];
      ov[q] = v.cell[k] == gc ? d : 0.0;
      if (mc > 0) oid[q] = v.id[k]; /* the ids only place the list's members */
    }
  }
  int mk = -1, mid_ = 0;
  double mval = 0.0;
  bool mem = false;
  if (fast && gl < mc) {
    mk = v.mig_slot[(size_t)gc * v.mig_cap + gl];
    mem = v.cell[mk] == gc;
    mval = v.dmxi[mk];
    mid_ = v.id[mk];
    sMk[g][gl] = mk;
  }
  __syncthreads();
  if (mem)
    for (int e = 0; e < gl; ++e)
      if (sMk[g][e] == mk) mem = false;
  const unsigned gb = (unsigned)(__ballot(mem) >> (g * L));
  const int nm = __popc(gb);
  if (mem) sMid[g][__popc(gb & ((1u << gl) - 1u))] = mid_;
#pragma unroll
  for (int q = 0; q < PER; ++q)
    if (fast && gl + q * L < T0) sHid[gl + q * L] = oid[q];
  __syncthreads();
  /* home entry: its index plus the members of the list with smaller ids; list member: its rank among them plus
     the home entries with smaller ids (ascending in the segment) */
  int opos[PER];
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    int lo = 0;
    if (fast && gl + q * L < T0)
      for (int e = 0; e < nm; ++e) lo += sMid[g][e] < oid[q];
    opos[q] = gl + q * L + lo;
  }
  int rk = 0;
  if (mem) {
    for (int e = 0; e < nm; ++e) rk += sMid[g][e] < mid_;
    int lo = 0, hi = T0;
    while (lo < hi) {
      const int m2 = (lo + hi) >> 1;
      if (sHid[m2] < mid_)
        lo = m2 + 1;
      else
        hi = m2;
    }
    rk += lo;
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < PER; ++q)
    if (fast && gl + q * L < T0) sVal[g][opos[q]] = ov[q];
  if (mem) sVal[g][rk] = mval;
  __syncthreads();
  if (fast && gl == 0) {
    const int tot = T0 + nm;
    double sum = 0.0;
    for (int p = 0; p < tot; ++p) sum += sVal[g][p];
    v.dml[gc] = sum * 1.0E-18;
  }
}

/* grid.py:370-412 update_grid_r_l_np: r_l = 1e-18 m_dry^-1 sum(xi m_w), summed over the cell in ascending id */
__global__ __launch_bounds__(WAVE) void k_rl(KView v) {
  const int nc = v.nc;
  const int gc = blockIdx.x, r = gc / nc, c = gc % nc;
  __shared__ int sSlot[DML_CAP], sId[DML_CAP], sRank[DML_CAP];
  __shared__ unsigned char sSub[DML_CAP];
  __shared__ double sVal[DML_CAP];
  __shared__ NbrTab tb;
  nbr_table(gc, nc, v.nx, v.nz, v.seg + (size_t)r * (nc + 1), v.away + (size_t)r * nc, &tb);
  const int n = gather_members(gc, v.cell, v.id, &tb, sSlot, sId, sSub, DML_CAP);
  if (n > DML_CAP) {
    /* over the LDS capacity: one lane sums in ascending id straight from the segments */
    if (threadIdx.x == 0) {
      double acc = 0.0;
      merge_members(gc, v.cell, v.id, &tb, [&](int k) { acc += v.m_w[k] * v.xi[k]; });
      double *S = v.scal + (size_t)r * 10 * nc;
      S[5 * nc + c] = acc * (1.0E-18 * S[8 * nc + c]);
    }
    return;
  }
  __syncthreads();
  rank_members(n, sId, sSub, &tb, sRank);
  __syncthreads();
  for (int m = threadIdx.x; m < n; m += WAVE) sVal[sRank[m]] = v.m_w[sSlot[m]] * v.xi[sSlot[m]];
  __syncthreads();
  if (threadIdx.x == 0) {
    double acc = 0.0;
    for (int p = 0; p < n; ++p) acc += sVal[p];
    double *S = v.scal + (size_t)r * 10 * nc;
    S[5 * nc + c] = acc * (1.0E-18 * S[8 * nc + c]);
  }
}

/*
 * Derived fields of one frame (evaluation.py:363-474) per (run, cell): R_p of every member at the cell
 * temperature (the solute's law, microphysics.py:179-266), classes by R_p against 0.5 and 25 um
 * (np.digitize, :319), then, over the members in ascending id, per class sum(xi m_w) and sum(xi), and the
 * radius moments sum(xi R^n) of all members (n = 0, 1) and of the cloud class (n = 1, 2, 3; R^n by repeated
 * multiplication).  Eleven lanes run the eleven ordered sums side by side.  out[(run*9 + f)*nc + c], f =
 * r_aero, r_c, r_r, n_aero, n_c, n_r, R_avg, R_2/1, R_eff.
 */
#define DERIV_CAP 1024
__global__ __launch_bounds__(WAVE) void k_derived(KView v, double *out) {
  const LmPhys &P = c_P;
  const int nc = v.nc;
  const int gc = blockIdx.x, r = gc / nc, c = gc % nc;
  __shared__ int sSlot[DERIV_CAP], sId[DERIV_CAP], sRank[DERIV_CAP];
  __shared__ unsigned char sSub[DERIV_CAP], sCls[DERIV_CAP];
  __shared__ double sMW[DERIV_CAP], sXI[DERIV_CAP], sRp[DERIV_CAP], sSum[11];
  __shared__ NbrTab tb;
  nbr_table(gc, nc, v.nx, v.nz, v.seg + (size_t)r * (nc + 1), v.away + (size_t)r * nc, &tb);
  const int n = gather_members(gc, v.cell, v.id, &tb, sSlot, sId, sSub, DERIV_CAP);
  const double *C = v.cc + (size_t)gc * CC_N;
  if (n > DERIV_CAP) {
    /* over the LDS capacity: one lane runs the eleven ordered sums straight from the segments */
    if (threadIdx.x == 0) {
      double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0, a4 = 0.0, a5 = 0.0, a6 = 0.0, a7 = 0.0, a8 = 0.0, a9 = 0.0,
             a10 = 0.0;
      merge_members(gc, v.cell, v.id, &tb, [&](int k) {
        double R, ws, rho;
        lmp_R_ws_rho(P, v.solute, v.m_w[k], v.m_s[k], C[CC_T], C[CC_RHOW], R, ws, rho);
        const double mw = v.m_w[k] * v.xi[k], x = v.xi[k];
        if (R < 0.5) {
          a0 += mw;
          a3 += x;
        } else if (R < 25.) {
          a1 += mw;
          a4 += x;
          a8 += x * R;
          a9 += x * (R * R);
          a10 += x * (R * (R * R));
        } else {
          a2 += mw;
          a5 += x;
        }
        a6 += x;
        a7 += x * R;
      });
      sSum[0] = a0;
      sSum[1] = a1;
      sSum[2] = a2;
      sSum[3] = a3;
      sSum[4] = a4;
      sSum[5] = a5;
      sSum[6] = a6;
      sSum[7] = a7;
      sSum[8] = a8;
      sSum[9] = a9;
      sSum[10] = a10;
    }
  } else {
  __syncthreads();
  rank_members(n, sId, sSub, &tb, sRank);
  __syncthreads();
  for (int m = threadIdx.x; m < n; m += WAVE) {
    const int k = sSlot[m], p = sRank[m];
    double R, ws, rho;
    lmp_R_ws_rho(P, v.solute, v.m_w[k], v.m_s[k], C[CC_T], C[CC_RHOW], R, ws, rho);
    sMW[p] = v.m_w[k] * v.xi[k];
    sXI[p] = v.xi[k];
    sRp[p] = R;
    sCls[p] = (unsigned char)(R < 0.5 ? 0 : (R < 25. ? 1 : 2));
  }
  __syncthreads();
  const int q = threadIdx.x;
  if (q < 11) {
    double acc = 0.0;
    for (int p = 0; p < n; ++p) {
      const int cl = sCls[p];
      const double x = sXI[p], R = sRp[p];
      if (q < 3) {
        if (cl == q) acc += sMW[p];
      } else if (q < 6) {
        if (cl == q - 3) acc += x;
      } else if (q == 6) {
        acc += x;
      } else if (q == 7) {
        acc += x * R;
      } else if (cl == 1) {
        acc += q == 8 ? x * R : (q == 9 ? x * (R * R) : x * (R * (R * R)));
      }
    }
    sSum[q] = acc;
  }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const double mdi = v.scal[((size_t)r * 10 + 8) * nc + c], V0 = v.V_cell;
    double *o = out + (size_t)r * 9 * nc + c;
    for (int k = 0; k < 3; ++k) {
      o[(size_t)k * nc] = sSum[k] * (1.0E-18 * mdi);
      o[(size_t)(3 + k) * nc] = sSum[3 + k] * mdi;
    }
    const double M0 = sSum[6] / V0, M1 = sSum[7] / V0, C1 = sSum[8] / V0, C2 = sSum[9] / V0, C3 = sSum[10] / V0;
    o[(size_t)6 * nc] = M0 == 0.0 ? 0.0 : M1 / M0;
    o[(size_t)7 * nc] = C1 == 0.0 ? 0.0 : C2 / C1;
    o[(size_t)8 * nc] = C2 == 0.0 ? 0.0 : C3 / C2;
  }
}

/* evaluation.py:443-474: per value, the mean over seeds (summed in seed order) and the reference's spread */
__global__ void k_ensemble(const double *f, int n_seeds, long long n_vals, double *mean, double *std) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n_vals) return;
  double s = 0.0, q = 0.0;
  for (int r = 0; r < n_seeds; ++r) {
    const double x = f[(size_t)r * n_vals + k];
    s += x;
    q += x * x;
  }
  const double m = s / (double)n_seeds;
  mean[k] = m;
  if (n_seeds <= 1) {
    std[k] = 0.0;
  } else {
    double h = q - (double)n_seeds * m * m;
    if (h < 0) h = 0.0;
    std[k] = sqrt(h / (double)((long long)n_seeds * (n_seeds - 1)));
  }
}

/* integration.py:402-483 + 359-400 for one cell; refreshes cc */
__global__ __launch_bounds__(256) void k_grid(KView v) {
  const LmPhys &P = c_P;
  const int gc = blockIdx.x * blockDim.x + threadIdx.x;
  if (gc >= v.n_runs * v.nc) return;
  const int nc = v.nc, r = gc / nc, c = gc % nc;
  double *S = v.scal + (size_t)r * 10 * nc;
  double *M = v.mat + (size_t)r * 7 * nc;
  const double rho_d = S[3 * nc + c];
  const double dml = v.dml[gc];
  v.cc[(size_t)gc * CC_N + CC_TP] = S[0 * nc + c]; /* particle temperature of this substep (integration.py:489-507) */
  double Th = S[2 * nc + c] + v.dTh[gc];
  double r_v = S[4 * nc + c];
  r_v += v.dRv[gc] - dml * S[8 * nc + c];
  const double Theta_over_T = lm_pow(v.p_ref_inv * P.R_d * rho_d * Th, P.kappa_factor2);
  Th += Theta_over_T * (M[2 * nc + c] * dml) / (P.c_pd * rho_d * v.V_cell * (1.0 + r_v * P.heat_factor_r_v));
  const double pdr = lm_pow(rho_d * Th * P.R_d * v.p_ref_inv, P.kappa_factor);
  const double T = Th * lm_pow(pdr, P.kappa_dry);
  const double p = pdr * v.p_ref * (1 + r_v / P.eps_gc);
  const double e_s = 2.53E11 * lm_exp(-5420.0 / T);
  const double Sat = (rho_d * r_v) * T * P.R_v / e_s;
  S[0 * nc + c] = T;
  S[1 * nc + c] = p;
  S[2 * nc + c] = Th;
  S[4 * nc + c] = r_v;
  S[6 * nc + c] = Sat;
  S[7 * nc + c] = e_s;
  M[0 * nc + c] = 4.1868E-3 * (5.69 + 0.017 * (T - 273.15));
  M[1 * nc + c] = 4.01218E-5 * lm_pow(T, 1.94) / p;
  M[2 * nc + c] = 1.0E3 * (2500.82 - 2.358 * (T - 273.0));
  const double sigma_w = lmp_sigma_water(T);
  M[3 * nc + c] = sigma_w;
  M[4 * nc + c] = P.c_pd * (1 + 0.897 * r_v);
  M[5 * nc + c] = 1.458E-6 * lm_pow(T, 1.5) / (T + 110.4);
  M[6 * nc + c] = rho_d * (1 + r_v);
  fill_cc(P, S, M, c, nc, v.cc + (size_t)gc * CC_N, sigma_w);
}

/* ------------------------------------------------------------------------ */
/* advection increments                                                     */
/* ------------------------------------------------------------------------ */
/* column sums over x in index order -> relaxation terms (integration.py:325-354) */
__global__ void k_colsum(KView v, double dt_sub, int act_relax) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= v.n_runs * v.nz) return;
  const int r = t / v.nz, j = t % v.nz, nc = v.nc;
  const double *S = v.scal + (size_t)r * 10 * nc;
  double srv = 0.0, sth = 0.0;
  for (int i = 0; i < v.nx; ++i) {
    srv += S[4 * nc + i * v.nz + j];
    sth += S[2 * nc + i * v.nz + j];
  }
  if (act_relax) {
    v.colsum[2 * t + 0] = dt_sub * (v.prof_rv[t] - srv / (double)v.nx) / v.t_relax[j];
    v.colsum[2 * t + 1] = dt_sub * (v.prof_Th[t] - sth / (double)v.nx) / v.t_relax[j];
  }
}

/* templates over the value type (double for the step, LmDual<4> for the adjoint's face derivatives) */
template <class R>
__device__ static inline R lim_fn(R r, double one_third) { /* integration.py:80-102 */
  const R Kv = (1.0 + 2.0 * r) * one_third;
  return lm_np_max(R(0.0), lm_np_fmin(2.0 * r, lm_np_fmin(Kv, R(2.0))));
}
template <class R>
__device__ static inline R lim_up(R a0, R a1, R a2, double one_third) { /* :113-148 */
  R r;
  if (fabs(lm_val(a1 - a2)) > 1.0E-8 * fabs(lm_val(a0 - a1)))
    r = (a0 - a1) / (a1 - a2);
  else
    r = (a0 - a1) * 1.0E8 * lm_np_sign(lm_val(a1 - a2));
  return lim_fn(r, one_third);
}
template <class R>
__device__ static inline R face_flux(R a0, R a1, R a2, R a3, double u, double ot) {
  const R f_pos = a1 + 0.5 * lim_up(a0, a1, a2, ot) * (a1 - a2);
  const R f_neg = a0 + 0.5 * lim_up(a1, a0, a3, ot) * (a0 - a3);
  return u >= 0.0 ? u * f_pos : u * f_neg;
}

/* ghost-cell index of the extended field (integration.py:202-212): periodic (bc 0) wraps, fixed (bc 1) clamps */
__device__ static inline int bc_index(int k, int n, int bc) {
  if (bc == 0) return ((k % n) + n) % n;
  return k < 0 ? 0 : (k > n - 1 ? n - 1 : k);
}

/*
 * integration.py:150-286 at cell (i, j): upwind flux through the faces
 * i, i+1 (x) and j, j+1 (z); a fixed boundary (bc 1) has zero flux through
 * the outer faces (:239-241, :283-285).  The default of the time step is
 * periodic x, fixed z.
 */
__device__ static double divergence_at(const double *f, const double *flux, int i, int j, int nx, int nz, double dx,
                                       double dz, double ot, int bcx = 0, int bcz = 1) {
  const int nfz = nz + 1;
  const double *FU = flux, *FW = flux + (size_t)(nx + 1) * nfz;
  double Fx[2];
  for (int s = 0; s < 2; ++s) {
    const int fi = i + s;
    if (bcx == 1 && (fi == 0 || fi == nx)) {
      Fx[s] = 0.0;
      continue;
    }
    const int i0 = bc_index(fi, nx, bcx), i1 = bc_index(fi - 1, nx, bcx), i2 = bc_index(fi - 2, nx, bcx),
              i3 = bc_index(fi + 1, nx, bcx);
    Fx[s] = face_flux(f[i0 * nz + j], f[i1 * nz + j], f[i2 * nz + j], f[i3 * nz + j], FU[fi * nfz + j], ot);
  }
  double Fz[2];
  for (int s = 0; s < 2; ++s) {
    const int fj = j + s;
    if (bcz == 1 && (fj == 0 || fj == nz)) {
      Fz[s] = 0.0;
      continue;
    }
    const int j0 = bc_index(fj, nz, bcz), j1 = bc_index(fj - 1, nz, bcz), j2 = bc_index(fj - 2, nz, bcz),
              j3 = bc_index(fj + 1, nz, bcz);
    Fz[s] = face_flux(f[i * nz + j0], f[i * nz + j1], f[i * nz + j2], f[i * nz + j3], FW[i * nfz + fj], ot);
  }
  double out = (Fx[1] - Fx[0]) / dx;
  out = out + (Fz[1] - Fz[0]) / dz;
  return out;
}

/* divergence of n_fields stacked (nx, nz) fields with one flux field (integration.py:150-286) */
__global__ void k_divergence(const double *fields, const double *flux, double *out, int n_fields, int nx, int nz,
                             double dx, double dz, int bcx, int bcz) {
  const LmPhys &P = c_P;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int nc = nx * nz;
  if (t >= n_fields * nc) return;
  const int q = t / nc, c = t % nc;
  out[t] = divergence_at(fields + (size_t)q * nc, flux, c / nz, c % nz, nx, nz, dx, dz, P.one_third, bcx, bcz);
}

__global__ void k_adv_incr(KView v, double dt_sub, int second, int act_relax) {
  const LmPhys &P = c_P;
  const int gc = blockIdx.x * blockDim.x + threadIdx.x;
  if (gc >= v.n_runs * v.nc) return;
  const int nc = v.nc, r = gc / nc, c = gc % nc, i = c / v.nz, j = c % v.nz;
  const double *S = v.scal + (size_t)r * 10 * nc;
  const double rho_inv = S[9 * nc + c];
  const double div_rv = divergence_at(S + 4 * nc, v.flux, i, j, v.nx, v.nz, v.dx, v.dz, P.one_third);
  const double div_th = divergence_at(S + 2 * nc, v.flux, i, j, v.nx, v.nz, v.dx, v.dz, P.one_third);
  double drv, dth;
  if (!second) {
    drv = -dt_sub * div_rv * rho_inv;
    dth = -dt_sub * div_th * rho_inv;
  } else {
    drv = -2.0 * dt_sub * div_rv * rho_inv - v.dRv[gc];
    dth = -2.0 * dt_sub * div_th * rho_inv - v.dTh[gc];
  }
  if (act_relax) {
    const double *cs = v.colsum + (size_t)2 * (r * v.nz + j);
    if (!second) {
      drv += cs[0];
      dth += cs[1];
    } else {
      drv += 2. * cs[0];
      dth += 2. * cs[1];
    }
  }
  v.dRv[gc] = drv;
  v.dTh[gc] = dth;
}

/* ------------------------------------------------------------------------ */
/* reindex: counting sort by (run, cell, id)                                */
/* ------------------------------------------------------------------------ */

/* scan layout per run: [cnt_0..cnt_{nc-1}, graveyard_old, dead_0..dead_{nc-1}] */
__global__ __launch_bounds__(WAVE) void k_ri_count(KView v) {
  const int nc = v.nc, L = 2 * nc + 1;
  const int lane = threadIdx.x;
  if ((int)blockIdx.x >= v.n_runs * nc) {
    const int r = blockIdx.x - v.n_runs * nc;
    if (lane == 0) {
      const int32_t *sr = v.seg + (size_t)r * (nc + 1);
      v.scan_buf[(size_t)r * L + nc] = sr[nc + 1] - sr[nc];
    }
    return;
  }
  const int gc = blockIdx.x, r = gc / nc, c = gc % nc;
  const int32_t *sr = v.seg + (size_t)r * (nc + 1);
  int n = 0, dead = 0;
  const int mc = v.mig_cnt ? v.mig_cnt[gc] : 0;
  if (v.mig_cnt && mc <= v.mig_cap) {
    /* a complete migrant list: the home segment's members and dead SIPs in one pass, then the list's
       members, each slot once (a SIP that entered the cell twice is listed twice) */
    for (int base = sr[c]; base < sr[c + 1]; base += WAVE) {
      const int k = base + lane;
      const int cv = k < sr[c + 1] ? v.cell[k] : gc + 1;
      n += __popcll(__ballot(cv == gc));
      dead += __popcll(__ballot(cv < 0));
    }
    const int k = lane < mc ? v.mig_slot[(size_t)gc * v.mig_cap + lane] : -1;
    bool mem = k >= 0 && v.cell[k] == gc;
    for (int e = 0; e < mc; ++e) {
      const int ke = __shfl(k, e);
      if (e < lane && ke == k) mem = false;
    }
    n += __popcll(__ballot(mem));
  } else {
    __shared__ NbrTab tb;
    nbr_table(gc, nc, v.nx, v.nz, sr, v.away + (size_t)r * nc, &tb);
    n = count_members(gc, v.cell, &tb);
    for (int base = sr[c]; base < sr[c + 1]; base += WAVE) {
      const int k = base + lane;
      dead += __popcll(__ballot(k < sr[c + 1] && v.cell[k] < 0));
    }
  }
  if (lane == 0) {
    v.scan_buf[(size_t)r * L + c] = n;
    v.scan_buf[(size_t)r * L + nc + 1 + c] = dead;
  }
}

#define SCAN_BLOCK 256
#define SCAN_PER 16
#define SCAN_CHUNK (SCAN_BLOCK * SCAN_PER)

__device__ static int block_exclusive_scan(int val, int *tmp) {
  const int t = threadIdx.x;
  tmp[t] = val;
  __syncthreads();
  for (int off = 1; off < SCAN_BLOCK; off <<= 1) {
    const int add = t >= off ? tmp[t - off] : 0;
    __syncthreads();
    tmp[t] += add;
    __syncthreads();
  }
  const int incl = tmp[t];
  __syncthreads();
  return incl - val;
}

__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_local(int32_t *a, int32_t *part, int L) {
  __shared__ int tmp[SCAN_BLOCK];
  const int base = blockIdx.x * SCAN_CHUNK + threadIdx.x * SCAN_PER;
  int loc[SCAN_PER];
  int s = 0;
  for (int q = 0; q < SCAN_PER; ++q) {
    const int k = base + q;
    loc[q] = k < L ? a[k] : 0;
    s += loc[q];
  }
  const int ex = block_exclusive_scan(s, tmp);
  int acc = ex;
  for (int q = 0; q < SCAN_PER; ++q) {
    const int k = base + q;
    if (k < L) a[k] = acc;
    acc += loc[q];
  }
  if (threadIdx.x == SCAN_BLOCK - 1) part[blockIdx.x] = acc;
}

__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_part(int32_t *part, int nparts, int32_t *a, int L) {
  __shared__ int tmp[SCAN_BLOCK];
  int carry = 0;
  for (int b0 = 0; b0 < nparts; b0 += SCAN_BLOCK) {
    const int k = b0 + threadIdx.x;
    const int val = k < nparts ? part[k] : 0;
    const int ex = block_exclusive_scan(val, tmp);
    if (k < nparts) part[k] = carry + ex;
    __syncthreads();
    if (threadIdx.x == SCAN_BLOCK - 1) tmp[0] = ex + val;
    __syncthreads();
    carry += tmp[0];
    __syncthreads();
  }
  if (threadIdx.x == 0) a[L] = carry;
}

__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_add(int32_t *a, const int32_t *part, int L) {
  const int base = blockIdx.x * SCAN_CHUNK;
  const int add = part[blockIdx.x];
  for (int q = threadIdx.x; q < SCAN_CHUNK; q += SCAN_BLOCK) {
    const int k = base + q;
    if (k < L) a[k] += add;
  }
}

#define WL_BLOCK 1024
#define WL_PER 4

/*
 * Work list of the persistent per-cell passes: the global cells ordered by decreasing home-segment size (the
 * member count up to drift), so the larger cells start first and the small ones fill the tail.  With col_cost
 * kept, the cells whose last collision step took more than half the expected length of a persistent pass
 * (sum of all cells' costs over the workers, thr) come first, by decreasing cost (log-scale bins, 16 per
 * octave): a cell that long has to start at once, while cells well below it keep the SIMDs issuing best when
 * mixed (they are latency-bound when many events, throughput-bound otherwise).  Counting sort over
 * LM_WORK_BINS bins (the lower half for costs, the upper half for sizes): histogram, exclusive scan, scatter.
 * The order within a bin follows the atomics; any order is a valid list.
 */
__device__ static inline int wl_bin(const KView &v, int gc, unsigned long long thr) {
  if (v.col_cost) {
    const uint32_t t = v.col_cost[gc];
    if (t > thr) {
      const int lg = 31 - __clz((int)t); /* t > thr >= 16: 4 .. 31 */
      const int b = (lg - 3) * 16 + (int)((t >> (lg - 4)) & 15u); /* 16 .. 463 */
      return LM_WORK_BINS / 2 - 1 - (b < LM_WORK_BINS / 2 - 1 ? b : LM_WORK_BINS / 2 - 1);
    }
  }
  const int r = gc / v.nc, c = gc - r * v.nc;
  const int32_t *s = v.seg + (size_t)r * (v.nc + 1);
  const int len = s[c + 1] - s[c];
  const int span = LM_WORK_BINS - 1 - (v.col_cost ? LM_WORK_BINS / 2 : 0); /* size bins: [1023 - span, 1023] */
  return LM_WORK_BINS - 1 - (len < span ? len : span);
}

/* sum of the cells' last collision costs (ticks) into *sum */
__global__ __launch_bounds__(WL_BLOCK) void k_wl_cost_sum(KView v, unsigned long long *sum) {
  __shared__ unsigned long long part[WL_BLOCK / WAVE];
  const int n = v.n_runs * v.nc;
  unsigned long long a = 0;
  for (int gc = (int)(blockIdx.x * WL_BLOCK + threadIdx.x); gc < n; gc += (int)(gridDim.x * WL_BLOCK))
    a += v.col_cost[gc];
  for (int off = WAVE / 2; off > 0; off >>= 1) a += __shfl_down(a, off);
  if ((threadIdx.x & (WAVE - 1)) == 0) part[threadIdx.x / WAVE] = a;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int w = 0; w < WL_BLOCK / WAVE; ++w) t += part[w];
    atomicAdd(sum, t);
  }
}

/* per-block histograms in LDS: one global atomic per (block, bin) instead of one per cell */
/* the threshold of the cost-ordered head of the list: half the expected length of a persistent collision pass
   over `workers` workgroups (no head: ~0) */
__device__ static inline unsigned long long wl_thr(const unsigned long long *sum, long long workers) {
  return (sum && workers > 0) ? *sum / (unsigned long long)(2 * workers) : ~0ull;
}

__global__ __launch_bounds__(WL_BLOCK) void k_wl_hist(KView v, int32_t *hist, const unsigned long long *sum,
                                                      long long workers) {
  const unsigned long long thr = wl_thr(sum, workers);
  __shared__ int h[LM_WORK_BINS];
  for (int b = threadIdx.x; b < LM_WORK_BINS; b += WL_BLOCK) h[b] = 0;
  __syncthreads();
  const int n = v.n_runs * v.nc;
  for (int q = 0; q < WL_PER; ++q) {
    const int gc = (int)blockIdx.x * WL_BLOCK * WL_PER + q * WL_BLOCK + (int)threadIdx.x;
    if (gc < n) atomicAdd(&h[wl_bin(v, gc, thr)], 1);
  }
  __syncthreads();
  for (int b = threadIdx.x; b < LM_WORK_BINS; b += WL_BLOCK)
    if (h[b]) atomicAdd(&hist[b], h[b]);
}

__global__ __launch_bounds__(SCAN_BLOCK) void k_wl_scan(int32_t *hist) {
  __shared__ int tmp[SCAN_BLOCK];
  int carry = 0;
  for (int b0 = 0; b0 < LM_WORK_BINS; b0 += SCAN_BLOCK) {
    const int k = b0 + threadIdx.x;
    const int val = hist[k];
    const int ex = block_exclusive_scan(val, tmp);
    hist[k] = carry + ex;
    __syncthreads();
    if (threadIdx.x == SCAN_BLOCK - 1) tmp[0] = ex + val;
    __syncthreads();
    carry += tmp[0];
    __syncthreads();
  }
}

/* hist holds each bin's first slot; a block reserves one range per bin and places its cells by local rank */
__global__ __launch_bounds__(WL_BLOCK) void k_wl_scatter(KView v, int32_t *hist, int32_t *work,
                                                         const unsigned long long *sum, long long workers) {
  const unsigned long long thr = wl_thr(sum, workers);
  __shared__ int h[LM_WORK_BINS];
  for (int b = threadIdx.x; b < LM_WORK_BINS; b += WL_BLOCK) h[b] = 0;
  __syncthreads();
  const int n = v.n_runs * v.nc;
  int bin[WL_PER], rk[WL_PER];
  for (int q = 0; q < WL_PER; ++q) {
    const int gc = (int)blockIdx.x * WL_BLOCK * WL_PER + q * WL_BLOCK + (int)threadIdx.x;
    bin[q] = gc < n ? wl_bin(v, gc, thr) : 0;
    rk[q] = gc < n ? atomicAdd(&h[bin[q]], 1) : 0;
  }
  __syncthreads();
  for (int b = threadIdx.x; b < LM_WORK_BINS; b += WL_BLOCK)
    if (h[b]) h[b] = atomicAdd(&hist[b], h[b]);
  __syncthreads();
  for (int q = 0; q < WL_PER; ++q) {
    const int gc = (int)blockIdx.x * WL_BLOCK * WL_PER + q * WL_BLOCK + (int)threadIdx.x;
    if (gc < n) {
      work[h[bin[q]] + rk[q]] = gc;
    }
  }
}

__device__ static inline void copy_sip(const KView &v, int src, int dst, int cellv) {
  v.x2[dst] = v.x[src];
  v.z2[dst] = v.z[src];
  v.u2[dst] = v.u[src];
  v.w2[dst] = v.w[src];
  v.m_w2[dst] = v.m_w[src];
  v.m_s2[dst] = v.m_s[src];
  v.xi2[dst] = v.xi[src];
  v.id2[dst] = v.id[src];
  v.cell2[dst] = cellv;
}

/* CAP: the LDS member capacity (larger cells merged by one lane); 256 keeps the kernel at eight waves per SIMD,
   512 at about five */
template <int CAP>
__global__ __launch_bounds__(WAVE) void k_ri_scatter(KView v) {
  const int nc = v.nc, L = 2 * nc + 1;
  const int lane = threadIdx.x;
  if ((int)blockIdx.x >= v.n_runs * nc) {
    /* graveyard of run r: old graveyard first, in slot order */
    const int r = blockIdx.x - v.n_runs * nc;
    const int32_t *sr = v.seg + (size_t)r * (nc + 1);
    const int32_t *sc = v.scan_buf + (size_t)r * L;
    const int g0 = sr[nc], g1 = sr[nc + 1];
    const int dst0 = sc[nc];
    for (int k = g0 + lane; k < g1; k += WAVE) copy_sip(v, k, dst0 + (k - g0), v.cell[k]);
    if (lane == 0) {
      v.seg2[(size_t)r * (nc + 1) + nc] = dst0;
      const int new_size = v.scan_buf[(size_t)(r + 1) * L] - sc[0];
      const int old_size = sr[nc + 1] - sr[0];
      if (new_size != old_size) raise_err(v.err, LM_ERR_LOST_SIP, r);
      if (r == v.n_runs - 1) v.seg2[(size_t)v.n_runs * (nc + 1)] = v.scan_buf[(size_t)v.n_runs * L];
    }
    return;
  }
  const int gc = blockIdx.x, r = gc / nc, c = gc % nc;
  const int32_t *sr = v.seg + (size_t)r * (nc + 1);
  const int32_t *sc = v.scan_buf + (size_t)r * L;
  __shared__ int sSlot[CAP], sId[CAP], sRank[CAP];
  __shared__ unsigned char sSub[CAP];
  __shared__ int sMig[3 * LM_MIG_CAP_MAX];
  __shared__ NbrTab tb;
  const int dst0 = sc[c];
  /* a complete migrant list: the members in id order from the home segment and the list (no 3x3 gather) */
  const int nf = v.mig_cnt ? fast_members(v, gc, sRank, CAP, sSlot, sId, sMig) : -1;
  if (nf >= 0 && nf <= CAP) {
    for (int m = lane; m < nf; m += WAVE) copy_sip(v, sRank[m], dst0 + m, gc);
  } else {
    nbr_table(gc, nc, v.nx, v.nz, sr, v.away + (size_t)r * nc, &tb);
    const int n = gather_members(gc, v.cell, v.id, &tb, sSlot, sId, sSub, CAP);
    if (n > CAP) {
      /* over the LDS capacity: one lane merges the members in id order, parking each source slot in the cell
         array of its destination (this cell's range of the new buffer), then the wave copies */
      if (lane == 0) {
        int idx = 0;
        merge_members(gc, v.cell, v.id, &tb, [&](int k) { v.cell2[dst0 + idx++] = k; });
      }
      __syncthreads();
      for (int m = lane; m < n; m += WAVE) copy_sip(v, v.cell2[dst0 + m], dst0 + m, gc);
    } else {
      __syncthreads();
      rank_members(n, sId, sSub, &tb, sRank);
      __syncthreads();
      for (int m = lane; m < n; m += WAVE) copy_sip(v, sSlot[m], dst0 + sRank[m], gc);
    }
  }
  /* SIPs that died in this home segment since the last reindex, in slot order */
  int dead = 0;
  const int ddst = sc[nc + 1 + c];
  for (int base = sr[c]; base < sr[c + 1]; base += WAVE) {
    const int k = base + lane;
    const bool d = k < sr[c + 1] && v.cell[k] < 0;
    const unsigned long long bal = __ballot(d);
    if (d) copy_sip(v, k, ddst + dead + __popcll(bal & lane_mask_lt()), v.cell[k]);
    dead += __popcll(bal);
  }
  if (lane == 0) v.seg2[(size_t)r * (nc + 1) + c] = dst0;
}

/* ------------------------------------------------------------------------ */
/* reference layout <-> engine layout on the device                          */
/* ------------------------------------------------------------------------ */
/*
 * One run's SIPs from the reference's arrays (pos, vel (2, N), cells (2, N) int64, m_w, m_s, xi (N),
 * active_ids (N) bytes) into the engine's order: active SIPs by (cell, index), then the inactive ones by
 * index (the graveyard).  A counting scatter into per-cell lists, each list ranked by index; the graveyard
 * by an exclusive scan of the inactive flags.
 */
__global__ void k_lay_count(long long N, const int64_t *cells, const uint8_t *active, int nx, int nz, int32_t *cnt,
                            int32_t *dead, int32_t *err) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= N) return;
  const long long cx = cells[k], cz = cells[N + k];
  if (active[k]) {
    dead[k] = 0;
    if (cx < 0 || cx >= nx || cz < 0 || cz >= nz) {
      raise_err(err, LM_ERR_ARG, (int)k);
      return;
    }
    atomicAdd(&cnt[cx * nz + cz], 1);
  } else {
    dead[k] = 1;
    if (cx < 0 || cx >= nx) raise_err(err, LM_ERR_ARG, (int)k);
  }
}

__global__ void k_lay_scatter(long long N, const int64_t *cells, const uint8_t *active, int nx, int nz,
                              const int32_t *off, int32_t *fill, int32_t *list) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= N || !active[k]) return;
  const long long cx = cells[k], cz = cells[N + k];
  if (cx < 0 || cx >= nx || cz < 0 || cz >= nz) return;
  const int c = (int)(cx * nz + cz);
  list[off[c] + atomicAdd(&fill[c], 1)] = (int32_t)k;
}

/* the engine's slot s of SIP k of the reference arrays */
__device__ static inline void lay_put(const KView &v, long long N, long long s, int k, int cellv, const double *pos,
                                      const double *vel, const double *m_w, const double *m_s, const double *xi) {
  v.x[s] = pos[k];
  v.z[s] = pos[N + k];
  v.u[s] = vel[k];
  v.w[s] = vel[N + k];
  v.m_w[s] = m_w[k];
  v.m_s[s] = m_s[k];
  v.xi[s] = xi[k];
  v.id[s] = k;
  v.cell[s] = cellv;
}

#define LAY_CAP 1024
/* one wave per cell of run r: its members in ascending index (ranked by counting; from LDS up to LAY_CAP) */
__global__ __launch_bounds__(WAVE) void k_lay_cell(KView v, int r, long long N, long long base, const double *pos,
                                                   const double *vel, const double *m_w, const double *m_s,
                                                   const double *xi, const int32_t *off, const int32_t *list) {
  __shared__ int sk[LAY_CAP];
  const int c = blockIdx.x, lane = threadIdx.x;
  const int o0 = off[c], n = off[c + 1] - o0;
  const bool lds = n <= LAY_CAP;
  if (lds)
    for (int p = lane; p < n; p += WAVE) sk[p] = list[o0 + p];
  __syncthreads();
  for (int p = lane; p < n; p += WAVE) {
    const int k = lds ? sk[p] : list[o0 + p];
    int rk = 0;
    if (lds)
      for (int q = 0; q < n; ++q) rk += sk[q] < k;
    else
      for (int q = 0; q < n; ++q) rk += list[o0 + q] < k;
    lay_put(v, N, base + o0 + rk, k, r * v.nc + c, pos, vel, m_w, m_s, xi);
  }
  if (lane == 0) v.seg[(size_t)r * (v.nc + 1) + c] = (int32_t)(base + o0);
}

/* the inactive SIPs after the cells, ascending index (dead: exclusive scan of the inactive flags) */
__global__ void k_lay_dead(KView v, int r, long long N, long long base, long long n_tot, const double *pos,
                           const double *vel, const int64_t *cells, const uint8_t *active, const double *m_w,
                           const double *m_s, const double *xi, const int32_t *off, const int32_t *dead) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int nc = v.nc;
  if (k == 0) {
    v.seg[(size_t)r * (nc + 1) + nc] = (int32_t)(base + off[nc]);
    v.seg[(size_t)v.n_runs * (nc + 1)] = (int32_t)n_tot;
  }
  if (k >= N || active[k]) return;
  lay_put(v, N, base + off[nc] + dead[k], (int)k, -(2 + (int)cells[k]), pos, vel, m_w, m_s, xi);
}

/* engine layout of run r -> the reference's arrays (scatter by id); rel_pos and T_p as integration.py:489-550
   leaves them (either may be NULL) */
__global__ void k_lay_store(KView v, int r, long long N, long long base, double *pos, double *vel, int64_t *cells,
                            double *rel_pos, double *m_w, double *m_s, double *xi, uint8_t *active, double *T_p,
                            const int32_t *cell_prev) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= N) return;
  const long long s = base + t;
  const int k = v.id[s], c = v.cell[s], nc = v.nc, nz = v.nz;
  if (k < 0 || k >= N) { /* not this run's ids: the engine holds another layout */
    raise_err(v.err, LM_ERR_ARG, (int)t);
    return;
  }
  const double x = v.x[s], z = v.z[s];
  pos[k] = x;
  pos[N + k] = z;
  vel[k] = v.u[s];
  vel[N + k] = v.w[s];
  m_w[k] = v.m_w[s];
  m_s[k] = v.m_s[s];
  xi[k] = v.xi[s];
  if (c >= 0) {
    const int loc = c - r * nc, cx = loc / nz, cz = loc % nz; /* a code of this run: 0 <= loc < nc */
    cells[k] = cx;
    cells[N + k] = cz;
    active[k] = 1;
    if (rel_pos) {
      rel_pos[k] = (x - v.x_min) / v.dx - (double)cx;
      rel_pos[N + k] = (z - v.z_min) / v.dz - (double)cz;
    }
  } else {
    cells[k] = -c - 2;
    cells[N + k] = -1;
    active[k] = 0;
  }
  if (T_p) {
    const int prev = cell_prev[s];
    const int loc = prev >= 0 ? prev - r * nc : (-prev - 2) * nz + (nz - 1);
    if (loc >= 0 && loc < nc)
      T_p[k] = v.cc[((size_t)r * nc + loc) * CC_N + CC_TP];
    else
      raise_err(v.err, LM_ERR_ARG, (int)t);
  }
}

/* ------------------------------------------------------------------------ */
/* all-cells collision step on the reference's arrays                       */
/* ------------------------------------------------------------------------ */
/* cell temperature and water density of the per-cell table (the slots the collision step reads) */
__global__ void k_ref_cc(int nc, const double *grid_T, double *cc) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  const double T = grid_T[c];
  cc[(size_t)c * CC_N + CC_T] = T;
  cc[(size_t)c * CC_N + CC_RHOW] = lmp_density_water(c_P, T);
}

/* cell of SIP i from cells (2, N) int64, -1 outside the grid (inactive SIPs carry cells[1] = -1) */
__device__ static inline int ref_cell(const int64_t *cells, long long N, long long i, int nx, int nz) {
  const long long a = cells[i], b = cells[N + i];
  return (a >= 0 && a < nx && b >= 0 && b < nz) ? (int)(a * nz + b) : -1;
}

__global__ void k_ref_count(long long N, const int64_t *cells, int nx, int nz, int32_t *cnt) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int c = ref_cell(cells, N, i, nx, nz);
  if (c >= 0) atomicAdd(&cnt[c], 1);
}

/* in-place exclusive scan of a[0..L) by one workgroup, total into a[L] */
__global__ __launch_bounds__(1024) void k_scan1_i32(int32_t *a, int L) {
  __shared__ int part[1024];
  const int t = threadIdx.x;
  const int per = (L + 1023) / 1024;
  const int b = t * per, e = b + per < L ? b + per : L;
  int s = 0;
  for (int k = b; k < e; ++k) s += a[k];
  part[t] = s;
  __syncthreads();
  if (t == 0) {
    int acc = 0;
    for (int q = 0; q < 1024; ++q) {
      const int x = part[q];
      part[q] = acc;
      acc += x;
    }
    a[L] = acc;
  }
  __syncthreads();
  int acc = part[t];
  for (int k = b; k < e; ++k) {
    const int x = a[k];
    a[k] = acc;
    acc += x;
  }
}

__global__ void k_ref_scatter(long long N, const int64_t *cells, int nx, int nz, const int32_t *off, int32_t *fill,
                              int32_t *list) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int c = ref_cell(cells, N, i, nx, nz);
  if (c >= 0) list[off[c] + atomicAdd(&fill[c], 1)] = (int)i;
}

/* pairs per cell for the offsets of an injected stream (reference order: cells x-major) */
__global__ void k_ref_count_pairs(int nc, const int32_t *off, long long *rnd_off) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  const long long n = off[c + 1] - off[c];
  rnd_off[c] = n * (n - 1) / 2;
}

/* ------------------------------------------------------------------------ */
/* box-model ensemble statistics (collision/box_model.py:1354-1585)          */
/* ------------------------------------------------------------------------ */
/*
 * Snapshots of all seeds live as (time, SIP) arrays: time t holds the SIPs of seed s in the columns
 * [off[s], off[s+1]).  k_box_minmax: per time the smallest and largest mass and multiplicity over all seeds
 * (exact: fmin/fmax do not round).
 */
__global__ __launch_bounds__(256) void k_box_minmax(long long n_tot, const double *xis, const double *masses,
                                                    double *out /* (n_times, 4): m_min, m_max, xi_min, xi_max */) {
  const int t = blockIdx.x;
  const double *X = xis + (size_t)t * n_tot, *M = masses + (size_t)t * n_tot;
  double a = INFINITY, b = -INFINITY, c = INFINITY, d = -INFINITY;
  for (long long k = threadIdx.x; k < n_tot; k += blockDim.x) {
    a = fmin(a, M[k]);
    b = fmax(b, M[k]);
    c = fmin(c, X[k]);
    d = fmax(d, X[k]);
  }
  __shared__ double s[4][256];
  s[0][threadIdx.x] = a;
  s[1][threadIdx.x] = b;
  s[2][threadIdx.x] = c;
  s[3][threadIdx.x] = d;
  __syncthreads();
  for (int h = blockDim.x / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      s[0][threadIdx.x] = fmin(s[0][threadIdx.x], s[0][threadIdx.x + h]);
      s[1][threadIdx.x] = fmax(s[1][threadIdx.x], s[1][threadIdx.x + h]);
      s[2][threadIdx.x] = fmin(s[2][threadIdx.x], s[2][threadIdx.x + h]);
      s[3][threadIdx.x] = fmax(s[3][threadIdx.x], s[3][threadIdx.x + h]);
    }
    __syncthreads();
  }
  if (threadIdx.x < 4) out[(size_t)t * 4 + threadIdx.x] = s[threadIdx.x][0];
}

/*
 * One wavefront per (seed, time): the moments sum(xi m^n)/dV, n = 0..3 (box_model.py:1521-1524), and the
 * histograms of box_model.py:1508-1519 over the time's nb mass bins (np.histogram's bins: [e_b, e_b+1),
 * the last one closed; masses outside are not counted): f = sum xi / (width_m dV), g = sum xi m /
 * (width_m dV), g_ln_r = sum xi m / (width_lnr dV) (binning the radius on the radius edges of the mass
 * edges is binning the mass: the radius is increasing in the mass).  Lane b sums bin b over the seed's SIPs
 * in ascending order, staged through LDS, so the sums are reproducible.  Outputs are (n_sims, n_times, .).
 */
#define BOXSTAT_CHUNK 1024
#define BOXSTAT_BINS 64
__global__ __launch_bounds__(WAVE) void k_box_stats(int n_times, int n_sims, const long long *off, long long n_tot,
                                                    const double *xis, const double *masses, const double *edges,
                                                    const double *width_m, const double *width_lnr, int nb,
                                                    double dV, double *mom, double *f_m, double *g_m,
                                                    double *g_lnr) {
  const int s = blockIdx.x / n_times, t = blockIdx.x % n_times, lane = threadIdx.x;
  const long long k0 = off[s], k1 = off[s + 1];
  const double *X = xis + (size_t)t * n_tot, *M = masses + (size_t)t * n_tot, *E = edges + (size_t)t * (nb + 1);
  __shared__ double sX[BOXSTAT_CHUNK], sM[BOXSTAT_CHUNK];
  __shared__ int sB[BOXSTAT_CHUNK];
  double m0 = 0.0, m1 = 0.0, m2 = 0.0, m3 = 0.0, hn = 0.0, hm = 0.0;
  for (long long c0 = k0; c0 < k1; c0 += BOXSTAT_CHUNK) {
    const int cn = (int)(k1 - c0 < BOXSTAT_CHUNK ? k1 - c0 : BOXSTAT_CHUNK);
    __syncthreads();
    for (int q = lane; q < cn; q += WAVE) {
      const double x = X[c0 + q], m = M[c0 + q];
      sX[q] = x;
      sM[q] = m;
      /* searchsorted(edges, m, 'right') - 1, the closed last bin, -1 outside */
      int lo = 0, hi = nb + 1;
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (E[mid] <= m)
          lo = mid + 1;
        else
          hi = mid;
      }
      int b = lo - 1;
      if (m == E[nb]) b = nb - 1;
      sB[q] = (b >= 0 && b < nb) ? b : -1;
    }
    __syncthreads();
    if (lane == 0) {
      /* moments in ascending SIP order */
      for (int q = 0; q < cn; ++q) {
        const double x = sX[q], m = sM[q];
        m0 += x;
        m1 += x * m;
        m2 += x * (m * m);
        m3 += x * (m * m * m);
      }
    }
    if (lane < nb) {
      for (int q = 0; q < cn; ++q) {
        if (sB[q] == lane) {
          hn += sX[q];
          hm += sX[q] * sM[q];
        }
      }
    }
  }
  const size_t base = (size_t)s * n_times + t;
  if (lane == 0) {
    mom[base * 4 + 0] = m0 / dV;
    mom[base * 4 + 1] = m1 / dV;
    mom[base * 4 + 2] = m2 / dV;
    mom[base * 4 + 3] = m3 / dV;
  }
  if (lane < nb) {
    const double wm = width_m[(size_t)t * nb + lane] * dV, wr = width_lnr[(size_t)t * nb + lane] * dV;
    f_m[base * nb + lane] = hn / wm;
    g_m[base * nb + lane] = hm / wm;
    g_lnr[base * nb + lane] = hm / wr;
  }
}

/* mean over seeds and its standard error std(ddof = 1) / sqrt(n) (box_model.py:1526-1541), two passes */
__global__ void k_mean_sem(const double *f, int n_sims, long long n_vals, double *mean, double *sem) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n_vals) return;
  double s = 0.0;
  for (int r = 0; r < n_sims; ++r) s += f[(size_t)r * n_vals + k];
  const double m = s / (double)n_sims;
  double q = 0.0;
  for (int r = 0; r < n_sims; ++r) {
    const double d = f[(size_t)r * n_vals + k] - m;
    q += d * d;
  }
  mean[k] = m;
  sem[k] = n_sims > 1 ? sqrt(q / (double)(n_sims - 1)) / sqrt((double)n_sims) : 0.0;
}

/* ------------------------------------------------------------------------ */
/* box models: one wavefront per box                                        */
/* ------------------------------------------------------------------------ */
#define BOX_CAP 1024

__device__ static long long box_stream_offset(const int32_t *off, int b) {
  long long s = 0;
  for (int q = 0; q < b; ++q) {
    const long long n = off[q + 1] - off[q];
    s += n * (n - 1) / 2;
  }
  return s;
}

/*
 * Single-component box model: all_or_nothing.py:109-185 (Golovin kernel,
 * K == NULL) and :392-491 (kernel K(m1, m2) tabulated on a log-mass grid,
 * index of kernel_index(m) as :35-65).
 */
__global__ __launch_bounds__(WAVE) void k_golovin(int nbox, const int32_t *off, double *xis, double *masses,
                                                  double dt_over_dV, const double *K, int nb, double lnm0,
                                                  double lnf, long long *no_cols, unsigned long long seed,
                                                  unsigned step, const double *rnd_stream, int32_t *err) {
  const int b = blockIdx.x, lane = threadIdx.x;
  const int s0 = off[b], n = off[b + 1] - off[b];
  if (n > BOX_CAP) {
    if (lane == 0) raise_err(err, LM_ERR_CELL_CAP, b);
    return;
  }
  __shared__ double sX[BOX_CAP], sM[BOX_CAP];
  __shared__ int sI[BOX_CAP];
  for (int m = lane; m < n; m += WAVE) {
    sX[m] = xis[s0 + m];
    sM[m] = masses[s0 + m];
    sI[m] = K ? lmp_kernel_index(sM[m], lnm0, lnf, nb) : 0;
  }
  __syncthreads();
  const long long roff = rnd_stream ? box_stream_offset(off, b) : 0;
  long long cs = 0, cm = 0;
  for (int d = 1; d <= 2 * n - 3; ++d) {
    const int i_lo = d - (n - 1) > 0 ? d - (n - 1) : 0;
    const int i_hi = (d - 1) >> 1;
    for (int i = i_lo + lane; i <= i_hi; i += WAVE) {
      const int j = d - i;
      int imin, imax;
      if (sX[i] <= sX[j]) {
        imin = i;
        imax = j;
      } else {
        imin = j;
        imax = i;
      }
      const double xi_min = sX[imin], xi_max = sX[imax];
      const double m_min = sM[imin], m_max = sM[imax];
      /* kernel.py:461-480 or the table */
      const double p_crit = xi_max * (K ? K[sI[i] * nb + sI[j]] : (m_min + m_max) * 1.5E-18) * dt_over_dV;
      if (p_crit > 1.0) {
        ++cm;
        const double xi_col = p_crit * xi_min;
        sM[imin] = (xi_min * m_min + xi_col * m_max) / xi_min;
        sX[imax] = xi_max - xi_col;
        if (K) sI[imin] = lmp_kernel_index(sM[imin], lnm0, lnf, nb);
      } else if (p_crit > 0.0) {
        const long long cnt = (long long)i * n - (long long)i * (i + 1) / 2 + (j - i - 1);
        const double rnd = rnd_stream ? rnd_stream[roff + cnt]
                                      : lm_pair_uniform(seed, (uint32_t)b, step, 0u, (uint32_t)i, (uint32_t)j);
        if (p_crit > rnd) {
          ++cs;
          const double xi_rel_dev = (xi_max - xi_min) / xi_max;
          if (xi_rel_dev < 1.0E-5) {
            const double xi_ges = xi_min + xi_max;
            const double mn = 2.0 * (xi_min * m_min + xi_max * m_max) / xi_ges;
            sM[imin] = mn;
            sM[imax] = mn;
            const double xmax_new = 0.5 * 0.7 * xi_ges;
            sX[imax] = xmax_new;
            sX[imin] = 0.5 * xi_ges - xmax_new;
            if (K) {
              const int ind = lmp_kernel_index(mn, lnm0, lnf, nb);
              sI[imin] = ind;
              sI[imax] = ind;
            }
          } else {
            sM[imin] = m_min + m_max;
            sX[imax] = xi_max - xi_min;
            if (K) sI[imin] = lmp_kernel_index(sM[imin], lnm0, lnf, nb);
          }
        }
      }
    }
    __syncthreads();
  }
  for (int m = lane; m < n; m += WAVE) {
    xis[s0 + m] = sX[m];
    masses[s0 + m] = sM[m];
  }
  for (int o = WAVE / 2; o > 0; o >>= 1) {
    cs += __shfl_down(cs, o);
    cm += __shfl_down(cm, o);
  }
  if (lane == 0) {
    no_cols[2 * b] += cs;
    no_cols[2 * b + 1] += cm;
  }
}

/* collision/kernel.py:107-167 */
__device__ static double beard_velocity(double R) {
  const double rho_w = 1.0E3, rho_a = 1.225, visc = 1.818E-5, sigma_w = 73.0E-3;
  const double drho = rho_w - rho_a, g = 9.80665;
  double v;
  if (R < 10.0) {
    const double l0 = 6.62E-2;
    const double C1 = drho * g / (4.5 * visc);
    const double C_sc = 1.0 + 1.257 * l0 / R;
    v = C1 * C_sc * R * R * 1.0E-12;
  } else if (R < 535.0) {
    const double p1[7] = {-0.327815e-5, 0.855176e-4, -0.578878e-3, -0.987059e-3, -0.153193e-2, 0.992696, -0.318657e1};
    const double N_Da = 32.0E-18 * R * R * R * rho_a * drho * g / (3.0 * visc * visc);
    const double Y = lmp_poly(p1, 7, lm_log(N_Da));
    const double C_sc = 1.0 + 1.257 * 6.62E-2 / R;
    v = visc * C_sc * lm_exp(Y) / (rho_a * R * 2.0E-6);
  } else if (R < 3500.0) {
    const double p2[6] = {0.238449e-2, -0.542819e-1, 0.475294, -0.204914e1, 0.523778e1, -0.500015e1};
    const double N_Bo = 16.0E-12 * R * R * drho * g / (3.0 * sigma_w);
    const double visc2 = visc * visc;
    const double N_P16 =
        lm_pow(sigma_w * sigma_w * sigma_w * rho_a * rho_a / ((visc2 * visc2) * drho * g), 1.0 / 6.0);
    const double Y = lmp_poly(p2, 6, lm_log(N_Bo * N_P16));
    v = visc * N_P16 * lm_exp(Y) / (rho_a * R * 2.0E-6);
  } else {
    v = 9.04929248;
  }
  return v;
}

/* all_or_nothing.py:269-390 followed by kernel.py:193-210 */
__global__ __launch_bounds__(WAVE) void k_ecol_box(int nbox, const int32_t *off, double *xis, double *masses,
                                                   double *radii, double *vel, const double *rho, double dt_over_dV,
                                                   const double *E, int nb, double lnR0, double lnf,
                                                   long long *no_cols, unsigned long long seed, unsigned step,
                                                   const double *rnd_stream, int32_t *err, int update_beard) {
  const LmPhys &P = c_P;
  const int b = blockIdx.x, lane = threadIdx.x;
  const int s0 = off[b], n = off[b + 1] - off[b];
  if (n > BOX_CAP) {
    if (lane == 0) raise_err(err, LM_ERR_CELL_CAP, b);
    return;
  }
  __shared__ double sX[BOX_CAP], sM[BOX_CAP], sR[BOX_CAP], sV[BOX_CAP], sRho[BOX_CAP];
  __shared__ int sI[BOX_CAP];
  for (int m = lane; m < n; m += WAVE) {
    sX[m] = xis[s0 + m];
    sM[m] = masses[s0 + m];
    sR[m] = radii[s0 + m];
    sV[m] = vel[s0 + m];
    sRho[m] = rho[s0 + m];
    sI[m] = lmp_kernel_index(sR[m], lnR0, lnf, nb);
  }
  __syncthreads();
  const long long roff = rnd_stream ? box_stream_offset(off, b) : 0;
  long long cs = 0, cm = 0;
  for (int d = 1; d <= 2 * n - 3; ++d) {
    const int i_lo = d - (n - 1) > 0 ? d - (n - 1) : 0;
    const int i_hi = (d - 1) >> 1;
    for (int i = i_lo + lane; i <= i_hi; i += WAVE) {
      const int j = d - i;
      int imin, imax;
      if (sX[i] <= sX[j]) {
        imin = i;
        imax = j;
      } else {
        imin = j;
        imax = i;
      }
      const double xi_min = sX[imin], xi_max = sX[imax];
      const double m_min = sM[imin], m_max = sM[imax];
      const double p_crit =
          xi_max * dt_over_dV * lmp_kernel_hydro(P, sR[i], sR[j], E[sI[i] * nb + sI[j]], fabs(sV[i] - sV[j]));
      if (p_crit > 1.0) {
        ++cm;
        const double xi_col = p_crit * xi_min;
        const double mn = (xi_min * m_min + xi_col * m_max) / xi_min;
        sM[imin] = mn;
        sR[imin] = lmp_radius(P, mn, sRho[imin]);
        sX[imax] = xi_max - xi_col;
        sI[imin] = lmp_kernel_index(sR[imin], lnR0, lnf, nb);
      } else if (p_crit > 0.0) {
        const long long cnt = (long long)i * n - (long long)i * (i + 1) / 2 + (j - i - 1);
        const double rnd = rnd_stream ? rnd_stream[roff + cnt]
                                      : lm_pair_uniform(seed, (uint32_t)b, step, 0u, (uint32_t)i, (uint32_t)j);
        if (p_crit > rnd) {
          ++cs;
          const double xi_rel_dev = (xi_max - xi_min) / xi_max;
          if (xi_rel_dev < 1.0E-5) {
            const double xi_ges = xi_min + xi_max;
            const double mn = 2.0 * (xi_min * m_min + xi_max * m_max) / xi_ges;
            sM[imin] = mn;
            sM[imax] = mn;
            const double Rn = lmp_radius(P, mn, sRho[imin]);
            sR[imin] = Rn;
            sR[imax] = Rn;
            const double xmax_new = 0.5 * 0.7 * xi_ges;
            sX[imax] = xmax_new;
            sX[imin] = 0.5 * xi_ges - xmax_new;
            const int ind = lmp_kernel_index(Rn, lnR0, lnf, nb);
            sI[imin] = ind;
            sI[imax] = ind;
          } else {
            const double mn = m_min + m_max;
            sM[imin] = mn;
            sR[imin] = lmp_radius(P, mn, sRho[imin]);
            sX[imax] = xi_max - xi_min;
            sI[imin] = lmp_kernel_index(sR[imin], lnR0, lnf, nb);
          }
        }
      }
    }
    __syncthreads();
  }
  for (int m = lane; m < n; m += WAVE) {
    xis[s0 + m] = sX[m];
    masses[s0 + m] = sM[m];
    radii[s0 + m] = sR[m];
    if (update_beard) vel[s0 + m] = beard_velocity(sR[m]);
  }
  for (int o = WAVE / 2; o > 0; o >>= 1) {
    cs += __shfl_down(cs, o);
    cm += __shfl_down(cm, o);
  }
  if (lane == 0) {
    no_cols[2 * b] += cs;
    no_cols[2 * b + 1] += cm;
  }
}

/* elementwise FP64 math of the kernels, for the host/device parity test */
/*
 * One collision step in each of nbox independent well-mixed boxes of
 * two-component SIPs with caller-supplied radii, 2-D velocities and densities:
 * all_or_nothing.py:498-635 (E table; E == NULL selects the constant E_const of
 * :637-751).  Pairs are visited by anti-diagonals d = i + j, which preserves the
 * row-major sequential result; pair draws are keyed (seed, box, step, 0, i, j)
 * or read from the injected stream in row-major order, box after box.
 */
__global__ __launch_bounds__(WAVE) void k_box2d(int nbox, const int32_t *off, double *xis, double *m_w,
                                                double *m_s, double *radii, const double *vx, const double *vz,
                                                const double *rho, double dt_over_dV, const double *E,
                                                double E_const, int nb, double lnR0, double lnf, long long *no_cols,
                                                unsigned long long seed, unsigned step, const double *rnd_stream,
                                                int32_t *err) {
  const LmPhys &P = c_P;
  const int b = blockIdx.x, lane = threadIdx.x;
  const int s0 = off[b], n = off[b + 1] - off[b];
  if (n > BOX_CAP) {
    if (lane == 0) raise_err(err, LM_ERR_CELL_CAP, b);
    return;
  }
  __shared__ double sX[BOX_CAP], sMw[BOX_CAP], sMs[BOX_CAP], sR[BOX_CAP], sU[BOX_CAP], sW[BOX_CAP], sRho[BOX_CAP];
  __shared__ int sI[BOX_CAP];
  for (int m = lane; m < n; m += WAVE) {
    sX[m] = xis[s0 + m];
    sMw[m] = m_w[s0 + m];
    sMs[m] = m_s[s0 + m];
    sR[m] = radii[s0 + m];
    sU[m] = vx[s0 + m];
    sW[m] = vz[s0 + m];
    sRho[m] = rho[s0 + m];
    sI[m] = E ? lmp_kernel_index(sR[m], lnR0, lnf, nb) : 0;
  }
  __syncthreads();
  const long long roff = rnd_stream ? box_stream_offset(off, b) : 0;
  long long cs = 0, cm = 0;
  for (int d = 1; d <= 2 * n - 3; ++d) {
    const int i_lo = d - (n - 1) > 0 ? d - (n - 1) : 0;
    const int i_hi = (d - 1) >> 1;
    for (int i = i_lo + lane; i <= i_hi; i += WAVE) {
      const int j = d - i;
      int imin, imax;
      if (sX[i] <= sX[j]) {
        imin = i;
        imax = j;
      } else {
        imin = j;
        imax = i;
      }
      const double xi_min = sX[imin], xi_max = sX[imax];
      const double mw_min = sMw[imin], mw_max = sMw[imax];
      const double ms_min = sMs[imin], ms_max = sMs[imax];
      const double du = sU[i] - sU[j], dw = sW[i] - sW[j];
      const double Ec = E ? E[sI[i] * nb + sI[j]] : E_const;
      const double p_crit = xi_max * dt_over_dV * lmp_kernel_hydro(P, sR[i], sR[j], Ec, sqrt(du * du + dw * dw));
      if (p_crit > 1.0) {
        ++cm;
        const double xi_col = p_crit * xi_min;
        sMw[imin] = mw_min + p_crit * mw_max;
        sMs[imin] = ms_min + p_crit * ms_max;
        sR[imin] = lmp_radius(P, (sMw[imin] + sMs[imin]), sRho[imin]);
        sX[imax] = xi_max - xi_col;
        if (E) sI[imin] = lmp_kernel_index(sR[imin], lnR0, lnf, nb);
      } else {
        const long long cnt = (long long)i * n - (long long)i * (i + 1) / 2 + (j - i - 1);
        const double rnd = rnd_stream ? rnd_stream[roff + cnt]
                                      : lm_pair_uniform(seed, (uint32_t)b, step, 0u, (uint32_t)i, (uint32_t)j);
        if (p_crit > rnd) {
          ++cs;
          const double xi_rel_dev = (xi_max - xi_min) / xi_max;
          if (xi_rel_dev < 1.0E-5) {
            const double xi_ges = xi_min + xi_max;
            const double mw = 2.0 * (xi_min * mw_min + xi_max * mw_max) / xi_ges;
            const double ms = 2.0 * (xi_min * ms_min + xi_max * ms_max) / xi_ges;
            sMw[imin] = mw;
            sMs[imin] = ms;
            sMw[imax] = mw;
            sMs[imax] = ms;
            const double Rn = lmp_radius(P, (mw + ms), sRho[imin]);
            sR[imin] = Rn;
            sR[imax] = Rn;
            const double xmax_new = 0.5 * 0.7 * xi_ges;
            sX[imax] = xmax_new;
            sX[imin] = 0.5 * xi_ges - xmax_new;
            if (E) {
              const int ind = lmp_kernel_index(Rn, lnR0, lnf, nb);
              sI[imin] = ind;
              sI[imax] = ind;
            }
          } else {
            sMw[imin] = mw_min + mw_max;
            sMs[imin] = ms_min + ms_max;
            sR[imin] = lmp_radius(P, (sMw[imin] + sMs[imin]), sRho[imin]);
            sX[imax] = xi_max - xi_min;
            if (E) sI[imin] = lmp_kernel_index(sR[imin], lnR0, lnf, nb);
          }
        }
      }
    }
    __syncthreads();
  }
  for (int m = lane; m < n; m += WAVE) {
    xis[s0 + m] = sX[m];
    m_w[s0 + m] = sMw[m];
    m_s[s0 + m] = sMs[m];
    radii[s0 + m] = sR[m];
  }
  for (int o = WAVE / 2; o > 0; o >>= 1) {
    cs += __shfl_down(cs, o);
    cm += __shfl_down(cm, o);
  }
  if (lane == 0) {
    no_cols[2 * b] += cs;
    no_cols[2 * b + 1] += cm;
  }
}

__global__ void k_math_probe(int which, long long n, const double *x, const double *y, double *out) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const double a = x[k], b = y[k];
  double r;
  switch (which) {
    case 0: r = lm_exp(a); break;
    case 1: r = lm_log(a); break;
    case 2: r = lm_pow(a, b); break;
    case 3: r = lm_pymod(a, b); break;
    case 4: r = sqrt(a); break;
    case 5: r = a / b; break;
    case 7: r = lm_pow_third(a); break;
    case 8: {
      double q1, q2;
      lmp_div2(a, 1.0, b, q1, q2); /* the shared-reciprocal quotient a / b */
      r = q1;
      break;
    }
    case 9: r = lm_div_range(b) ? lm_div_y(a, b, lm_rcp_nr(b)) : a / b; break;
    default: r = floor(a / b); break;
  }
  out[k] = r;
}

/* fast and exact kernel-table indices of the radii R[0..n) (lm_kernel_index_check) */
__global__ void k_ki_probe(KView v, long long n, const double *R, int32_t *fast, int32_t *exact) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  fast[k] = kernel_index_fast(R[k], v);
  exact[k] = lmp_kernel_index(R[k], v.lnR0, v.lnf, v.nb);
}

/* largest |hardware log2(x) - log2(x)| over every float x of [KI_RMIN, KI_RMAX) into out[0] (as ordered bits) */
__global__ __launch_bounds__(256) void k_log2_sweep(uint32_t lo, uint32_t count, unsigned long long *out) {
  double e = 0.0;
  for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < count; i += gridDim.x * 256u) {
    const float x = __uint_as_float(lo + i);
    const double hw = (double)__builtin_amdgcn_logf(x);
    const double ex = lm_log((double)x) * 1.4426950408889634;
    const double d = fabs(hw - ex);
    e = d > e ? d : e;
  }
  for (int off = WAVE / 2; off > 0; off >>= 1) {
    const double o = __shfl_xor(e, off);
    e = o > e ? o : e;
  }
  if ((threadIdx.x & (WAVE - 1)) == 0) atomicMax(out, (unsigned long long)__double_as_longlong(e));
}

__global__ void k_ki_consts(KView v, double *out, uint32_t count) {
  if (threadIdx.x == 0) {
    out[0] = 0.0;
    out[1] = KI_LOG2_ERR;
    out[2] = v.ki_margin;
    out[3] = (double)count;
  }
}

/* ------------------------------------------------------------------------ */
/* C ABI                                                                    */
/* ------------------------------------------------------------------------ */
/* event-pair timing of selected kernels (bench.py) */
struct LmProf {
  int max_ev, n;
  hipEvent_t *a, *b;
  int *kind;
};

struct ProfScope {
  LmProf *p;
  int idx;
  hipStream_t s;
  ProfScope(const LmEngine *e, int kind, hipStream_t st) : p((LmProf *)e->prof), idx(-1), s(st) {
    if (p && p->n < p->max_ev) {
      idx = p->n++;
      p->kind[idx] = kind;
      (void)hipEventRecord(p->a[idx], s);
    }
  }
  ~ProfScope() {
    if (idx >= 0) (void)hipEventRecord(p->b[idx], s);
  }
};

static int check_engine(const LmEngine *e) {
  if (!e) return set_err(LM_ERR_ARG, "engine is NULL");
  if (ensure_phys()) return set_err(LM_ERR_ARG, "cannot upload constants to the device");
  if (e->n_runs <= 0 || e->nx <= 0 || e->nz <= 0 || e->nc != e->nx * e->nz)
    return set_err(LM_ERR_ARG, "bad grid shape (need nc == nx*nz > 0, n_runs > 0)");
  if (e->cur < 0 || e->cur > 1 || e->ccur < 0 || e->ccur > 1) return set_err(LM_ERR_ARG, "bad buffer index");
  if (!e->x[0] || !e->x[1] || !e->cell[0] || !e->cell[1] || !e->seg[0] || !e->seg[1] || !e->cc || !e->err ||
      !e->away || !e->col_step_dev)
    return set_err(LM_ERR_ARG, "engine buffer pointer is NULL");
  if ((int64_t)e->n_runs * e->nc > 0x7fffffff / 2) return set_err(LM_ERR_ARG, "too many cells");
  if (e->n_tot > 0x7ffffff0) return set_err(LM_ERR_ARG, "too many SIPs for 32-bit slots");
  if (e->vel_run_stride != 0 && e->vel_run_stride != (int64_t)2 * (e->nx + 1) * (e->nz + 1))
    return set_err(LM_ERR_ARG, "vel_run_stride must be 0 or 2*(nx+1)*(nz+1)");
  if (e->ev_mode < 0 || e->ev_mode > 2) return set_err(LM_ERR_ARG, "bad ev_mode (0 off, 1 record, 2 replay)");
  if (e->mig_cnt && (!e->mig_slot || e->mig_cap < 1 || e->mig_cap > LM_MIG_CAP_MAX))
    return set_err(LM_ERR_ARG, "migrant lists need mig_slot and mig_cap in 1..LM_MIG_CAP_MAX (64)");
  if (e->big_ws && (e->big_cap < 2 || e->big_cap > LM_BIG_CELL_MAX || e->big_blocks < 1 || e->big_blocks > 4096))
    return set_err(LM_ERR_ARG, "big_ws needs big_cap in 2..LM_BIG_CELL_MAX (46340) and big_blocks in 1..4096");
  if (e->spec_events < 0 || e->spec_events > LM_SPEC_EVENTS_MAX)
    return set_err(LM_ERR_ARG, "spec_events must be 0 (default) or 1..LM_SPEC_EVENTS_MAX (1024)");
  if (e->rem_cap < 1 || e->rem_cap > LM_REM_CAP_MAX || !e->rem_slot || !e->rem_cnt)
    return set_err(LM_ERR_ARG, "rem_cap must be in 1..LM_REM_CAP_MAX (1024) with rem_slot and rem_cnt set");
  if (e->part_block < 0 || e->part_block > PART_BLOCK_MAX || (e->part_block % WAVE) || e->part_spt < 0 ||
      e->part_spt > 2 ||
      (e->col_cap && e->col_cap != 160 && e->col_cap != 256 && e->col_cap != 320 && e->col_cap != 512))
    return set_err(LM_ERR_ARG, "bad launch shape (block 64..512 in steps of 64, 1 or 2 SIPs per thread, "
                               "collision capacity 160, 256, 320 or 512)");
  if (e->col_work_per_cu < 0 || e->dml_work_per_cu < 0 || e->col_work_chunk < 0 || e->dml_work_chunk < 0 ||
      e->col_work_per_cu > 256 || e->dml_work_per_cu > 256 || e->col_work_chunk > 4096 || e->dml_work_chunk > 4096 ||
      (e->part_wpe && e->part_wpe != 5 && e->part_wpe != 6 && e->part_wpe != 8) || e->tape_mode < 0 ||
      e->tape_mode > 1)
    return set_err(LM_ERR_ARG, "bad launch shape (work-list shape outside 0..256 workgroups per CU or 0..4096 "
                               "items per grab, particle waves per SIMD not 0, 5, 6 or 8, or tape mode not 0 or 1)");
  if (e->ev_mode && (!e->ev || !e->ev_off || !e->ev_order || !e->run_base || e->ev_cap <= 0))
    return set_err(LM_ERR_ARG, "event log buffers missing");
  if (e->ev_mode == 1 && (!e->ev_cnt || !e->ev_count)) return set_err(LM_ERR_ARG, "event counters missing");
  if (e->ev_mode == 2 && !e->slot_of) return set_err(LM_ERR_ARG, "slot_of buffer missing");
  if (e->tape_part && (!e->tape_part_i || !e->tape_grid || !e->tape_fields || !e->ccd || !e->run_base))
    return set_err(LM_ERR_ARG, "tape buffers missing");
  return 0;
}

static inline hipStream_t S(void *s) { return (hipStream_t)s; }

/* mean + 2.5 sqrt(mean) SIPs per cell: the member capacity the main per-cell passes are sized for */
/* launch shape of the particle kernel: tuned values or 256 threads, one SIP per thread */
static int part_block(const LmEngine *e) { return e->part_block > 0 ? e->part_block : 256; }
static int part_spt(const LmEngine *e) { return e->part_spt > 0 ? e->part_spt : 1; }
static dim3 part_grid(const LmEngine *e, int spt) {
  const long long per = (long long)part_block(e) * spt;
  return dim3((unsigned)((e->n_tot + per - 1) / per));
}

static int cell_capacity(const LmEngine *e) {
  const double mean = (double)e->n_tot / ((double)e->n_runs * (double)e->nc);
  return (int)(mean + 2.5 * sqrt(mean)) + 1;
}

/* compute units of the current device, cached per device */
static int device_cus() {
  static int cus[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  if (cus[dev] <= 0) {
    int n = 0;
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) n = 256;
    cus[dev] = n;
  }
  return cus[dev];
}

/* Main pass of a per-cell kernel: wpc persistent workgroups per compute unit over the work list (taking
   chunk items per atomic on counter ctr, cleared by a fill kernel first; chunk 0: grid-strided), or one
   workgroup per global cell.  Returns the pass mode. */
static int main_pass(const LmEngine *e, int wpc, int chunk, int ctr, dim3 *grid, hipStream_t st) {
  const int ncell = e->n_runs * e->nc;
  if (wpc > 0 && e->work) {
    const long long g = (long long)wpc * device_cus();
    *grid = dim3((unsigned)(g < ncell ? g : ncell));
    if (chunk > 0)
      hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, st, e->work + (size_t)ncell + LM_WORK_BINS + ctr, 0, 1);
    return PASS_WORK;
  }
  *grid = dim3(ncell);
  return PASS_CELLS;
}

#include "lm_adjoint.h"
#include "lm_init.h"

template <int CAP>
static void launch_collision(const KView &v, bool tab, bool inj, dim3 grid, hipStream_t st, int mode, int chunk,
                             const ColLists &io) {
  if (tab && inj)
    hipLaunchKernelGGL((k_collision<CAP, true, true>), grid, dim3(WAVE), 0, st, v, mode, chunk, io);
  else if (tab)
    hipLaunchKernelGGL((k_collision<CAP, true, false>), grid, dim3(WAVE), 0, st, v, mode, chunk, io);
  else if (inj)
    hipLaunchKernelGGL((k_collision<CAP, false, true>), grid, dim3(WAVE), 0, st, v, mode, chunk, io);
  else
    hipLaunchKernelGGL((k_collision<CAP, false, false>), grid, dim3(WAVE), 0, st, v, mode, chunk, io);
}

extern "C" {

const char *lm_version(void) { return "lacmopy_amd 0.1 gfx950"; }

const char *lm_last_error_message(void) { return g_errmsg; }

int64_t lm_engine_sizeof(void) { return (int64_t)sizeof(LmEngine); }

int64_t lm_big_ws_bytes(int32_t cap) { return cap < 2 || cap > LM_BIG_CELL_MAX ? -1 : (int64_t)big_stride(cap); }

void lm_engine_offsets(int64_t *out, int32_t n) {
  const int64_t o[] = {
      (int64_t)offsetof(LmEngine, n_tot),     (int64_t)offsetof(LmEngine, solute),
      (int64_t)offsetof(LmEngine, dt),        (int64_t)offsetof(LmEngine, E),
      (int64_t)offsetof(LmEngine, seed),      (int64_t)offsetof(LmEngine, x),
      (int64_t)offsetof(LmEngine, id),        (int64_t)offsetof(LmEngine, dmxi),
      (int64_t)offsetof(LmEngine, scal),      (int64_t)offsetof(LmEngine, grid_vel),
      (int64_t)offsetof(LmEngine, no_cols),   (int64_t)offsetof(LmEngine, rem_cap),
      (int64_t)offsetof(LmEngine, scan_buf),  (int64_t)offsetof(LmEngine, rnd_stream),
      (int64_t)offsetof(LmEngine, rnd_len),     (int64_t)offsetof(LmEngine, away),
      (int64_t)offsetof(LmEngine, col_step_dev), (int64_t)offsetof(LmEngine, ev_mode),
      (int64_t)offsetof(LmEngine, ev),        (int64_t)offsetof(LmEngine, run_base),
      (int64_t)offsetof(LmEngine, tape_part), (int64_t)offsetof(LmEngine, ccd),
      (int64_t)offsetof(LmEngine, part_block), (int64_t)offsetof(LmEngine, E_max),
      (int64_t)offsetof(LmEngine, vel_run_stride), (int64_t)offsetof(LmEngine, col_work_per_cu),
      (int64_t)offsetof(LmEngine, work), (int64_t)offsetof(LmEngine, dml_work_chunk),
      (int64_t)offsetof(LmEngine, part_wpe), (int64_t)offsetof(LmEngine, mig_cnt),
      (int64_t)offsetof(LmEngine, mig_cap),
      (int64_t)offsetof(LmEngine, spec_events),
      (int64_t)offsetof(LmEngine, col_cost),
      (int64_t)offsetof(LmEngine, big_ws), (int64_t)offsetof(LmEngine, big_cap),
  };
  const int m = (int)(sizeof(o) / sizeof(o[0]));
  for (int i = 0; i < n && i < m; ++i) out[i] = o[i];
}

int32_t lm_num_constants(void) { return (int32_t)(sizeof(LmPhys) / sizeof(double)); }

void lm_get_constants(double *out, int32_t n) {
  const double *p = (const double *)&phys();
  const int m = lm_num_constants();
  for (int i = 0; i < n && i < m; ++i) out[i] = p[i];
}

int lm_engine_init_cells(LmEngine *e, void *stream) {
  int rc = check_engine(e);
  if (rc) return rc;
  const int ncell = e->n_runs * e->nc;
  hipLaunchKernelGGL(k_init_cc, dim3((ncell + 255) / 256), dim3(256), 0, S(stream), make_view(e));
  LAUNCH_CHECK();
  return 0;
}

int lm_collision_step(LmEngine *e, void *stream) {
  int rc = check_engine(e);
  if (rc) return rc;
  const int ncell = e->n_runs * e->nc;
  if (e->ev_mode && (e->ev_step < 0 || e->ev_step >= e->ev_nstep))
    return set_err(LM_ERR_ARG, "collision step outside the adjoint window's event log");
  if (e->ev_mode == 2) {
    /* fixed events: the recorded events of this collision step, no draws */
    const int32_t *off = e->ev_off + (size_t)e->ev_step * (ncell + 1);
    hipLaunchKernelGGL(k_ev_replay, dim3((ncell + 255) / 256), dim3(256), 0, S(stream), make_view(e), off, e->ev_order,
                       ncell);
    hipLaunchKernelGGL(k_add_u32, dim3(1), dim3(64), 0, S(stream), e->col_step_dev, 1u);
    LAUNCH_CHECK();
    e->col_step += 1;
    e->ev_step += 1;
    return 0;
  }
  if (!e->e_const_mode && (!e->E || e->nb <= 0)) return set_err(LM_ERR_ARG, "collision table missing");
  if (e->ev_mode == 1)
    hipLaunchKernelGGL(k_fill_i32, dim3((ncell + 255) / 256), dim3(256), 0, S(stream), e->ev_count, 0, ncell);
  if (e->rnd_stream) {
    if (!e->rnd_off) return set_err(LM_ERR_ARG, "rnd_off scratch missing");
    KView v = make_view(e);
    hipLaunchKernelGGL(k_count_pairs, dim3(ncell), dim3(WAVE), 0, S(stream), v);
    hipLaunchKernelGGL(k_scan_pairs, dim3(1), dim3(1024), 0, S(stream), (long long *)e->rnd_off, ncell);
    LAUNCH_CHECK();
  }
  /* list A: cells over the main pass's capacity; list B (second half of ovf_list, counter ovf_cnt[2]): cells
     over 256 members, after a main pass below 256 */
  hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, S(stream), e->ovf_cnt, 0, 1);
  hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, S(stream), e->ovf_cnt + 2, 0, e->big_ws ? 2 : 1);
  hipLaunchKernelGGL(k_fill_i32, dim3(GRAB_CTRS * GRAB_STRIDE / 64), dim3(64), 0, S(stream), e->ovf_cnt + GRAB_STRIDE,
                     0, GRAB_CTRS * GRAB_STRIDE);
  KView v = make_view(e);
  const bool tab = !e->e_const_mode, inj = e->rnd_stream != nullptr;
  /* member capacity of the main pass from the mean SIPs per cell (mean + 2.5 sigma of a Poisson count): a
     small capacity keeps the LDS footprint, and so the occupancy, low; larger cells go to the overflow passes */
  const int cap = e->col_cap > 0 ? e->col_cap : cell_capacity(e);
  const ColLists main_io = {nullptr, nullptr, e->ovf_list, e->ovf_cnt};
  {
    ProfScope ps(e, 1, S(stream));
    dim3 grid;
    const int mode = main_pass(e, e->col_work_per_cu, e->col_work_chunk, 0, &grid, S(stream));
    if (cap <= 160)
      launch_collision<160>(v, tab, inj, grid, S(stream), mode, e->col_work_chunk, main_io);
    else if (cap <= 256)
      launch_collision<256>(v, tab, inj, grid, S(stream), mode, e->col_work_chunk, main_io);
    else if (cap <= 320)
      launch_collision<320>(v, tab, inj, grid, S(stream), mode, e->col_work_chunk, main_io);
    else
      launch_collision<512>(v, tab, inj, grid, S(stream), mode, e->col_work_chunk, main_io);
  }
  {
    /* the overflow passes, timed apart from the main pass */
    ProfScope ps(e, 4, S(stream));
    ColLists last_io = {e->ovf_list, e->ovf_cnt, nullptr, nullptr};
    if (cap <= 160) {
      /* the 256-member pass over list A at its own occupancy (a few resident workgroups per compute unit,
         each claiming the next cell of the list: the costs of these cells spread widely), the larger cells on
         to list B */
      const ColLists mid_io = {e->ovf_list, e->ovf_cnt, e->ovf_list + ncell, e->ovf_cnt + 2,
                               e->ovf_cnt + GRAB_STRIDE};
      const long long gm = 12LL * device_cus();
      launch_collision<256>(v, tab, inj, dim3((unsigned)(gm < ncell ? gm : ncell)), S(stream), PASS_OVF, 0, mid_io);
      last_io.in_list = e->ovf_list + ncell;
      last_io.in_cnt = e->ovf_cnt + 2;
    }
    if (e->big_ws) {
      /* cells over 1024 members on to list C, then the pass over cells of any size */
      last_io.out_list = e->ovf_list + 2 * (size_t)ncell;
      last_io.out_cnt = e->ovf_cnt + 3;
    }
    launch_collision<1024>(v, tab, inj, dim3(ncell < OVF_GRID ? ncell : OVF_GRID), S(stream), PASS_OVF, 0, last_io);
    if (e->big_ws) {
      const ColLists big_io = {e->ovf_list + 2 * (size_t)ncell, e->ovf_cnt + 3, nullptr, nullptr};
      launch_collision<0>(v, tab, inj, dim3(e->big_blocks), S(stream), PASS_OVF, 0, big_io);
    }
  }
  LAUNCH_CHECK();
  if (e->ev_mode == 1) {
    /* the step's records in (cell, seq) order for the replay and the reverse sweep */
    int32_t *off = e->ev_off + (size_t)e->ev_step * (ncell + 1);
    hipLaunchKernelGGL(k_ev_scan, dim3(1), dim3(EV_SCAN_BLOCK), 0, S(stream), e->ev_count, ncell, e->ev_cnt, off);
    hipLaunchKernelGGL(k_ev_order, dim3(256), dim3(256), 0, S(stream), e->ev, e->ev_cap, off, ncell, e->ev_order);
    LAUNCH_CHECK();
    e->ev_step += 1;
  }
  hipLaunchKernelGGL(k_add_u32, dim3(1), dim3(64), 0, S(stream), e->col_step_dev, 1u);
  LAUNCH_CHECK();
  e->col_step += 1;
  return 0;
}

int lm_position_step(LmEngine *e, double dt, void *stream) {
  int rc = check_engine(e);
  if (rc) return rc;
  KView v = make_view(e);
  const long long n = e->n_tot;
  if (n > 0)
    hipLaunchKernelGGL((k_particles<0, true, false, 1>), part_grid(e, 1), dim3(part_block(e)), 0, S(stream), v, n, 0.0,
                       dt);
  hipLaunchKernelGGL(k_water, dim3(e->n_runs), dim3(WAVE), 0, S(stream), v);
  LAUNCH_CHECK();
  e->ccur = 1 - e->ccur;
  return 0;
}

int lm_build_work_list(LmEngine *e, void *stream) {
  int rc = check_engine(e);
  if (rc) return rc;
  if (!e->work) return set_err(LM_ERR_ARG, "work list buffer missing");
  const KView w = make_view(e);
  const int ncell = e->n_runs * e->nc;
  int32_t *hist = e->work + ncell;
  const dim3 g((unsigned)((ncell + WL_BLOCK * WL_PER - 1) / (WL_BLOCK * WL_PER)));
  /* the cost sum: an 8-byte aligned slot after the two work counters */
  unsigned long long *sum = nullptr;
  long long workers = 0;
  if (e->col_cost && e->col_work_per_cu > 0) {
    sum = (unsigned long long *)(e->work + (((size_t)ncell + LM_WORK_BINS + 2 + 1) & ~(size_t)1));
    const long long gw = (long long)e->col_work_per_cu * device_cus();
    workers = gw < ncell ? gw : ncell;
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, S(stream), (int32_t *)sum, 0, 2);
    hipLaunchKernelGGL(k_wl_cost_sum, dim3(g.x), dim3(WL_BLOCK), 0, S(stream), w, sum);
  }
  hipLaunchKernelGGL(k_fill_i32, dim3(LM_WORK_BINS / 256), dim3(256), 0, S(stream), hist, 0, LM_WORK_BINS);
  hipLaunchKernelGGL(k_wl_hist, g, dim3(WL_BLOCK), 0, S(stream), w, hist, sum, workers);
  hipLaunchKernelGGL(k_wl_scan, dim3(1), dim3(SCAN_BLOCK), 0, S(stream), hist);
  hipLaunchKernelGGL(k_wl_scatter, g, dim3(WL_BLOCK), 0, S(stream), w, hist, e->work, sum, workers);
  LAUNCH_CHECK();
  return 0;
}

int lm_reindex(LmEngine *e, void *stream) {
  int rc = check_engine(e);
  if (rc) return rc;
  KView v = make_view(e);
  const int nc = e->nc, R = e->n_runs;
  const int L = R * (2 * nc + 1);
  hipLaunchKernelGGL(k_ri_count, dim3(R * nc + R), dim3(WAVE), 0, S(stream), v);
  const int nchunk = (L + SCAN_CHUNK - 1) / SCAN_CHUNK;
  hipLaunchKernelGGL(k_scan_local, dim3(nchunk), dim3(SCAN_BLOCK), 0, S(stream), e->scan_buf, e->scan_part, L);
  hipLaunchKernelGGL(k_scan_part, dim3(1), dim3(SCAN_BLOCK), 0, S(stream), e->scan_part, nchunk, e->scan_buf, L);
  hipLaunchKernelGGL(k_scan_add, dim3(nchunk), dim3(SCAN_BLOCK), 0, S(stream), e->scan_buf, e->scan_part, L);
  {
    ProfScope ps(e, 3, S(stream));
    /* the 256-member build while cells average about 128 members (its occupancy: 1.66 against 1.92 ms for 46 M
       SIPs in 360,000 cells); the 512-member build for denser grids, where the one-lane merge would take many */
    if (cell_capacity(e) <= 192)
      hipLaunchKernelGGL(k_ri_scatter<256>, dim3(R * nc + R), dim3(WAVE), 0, S(stream), v);
    else
      hipLaunchKernelGGL(k_ri_scatter<DML_CAP>, dim3(R * nc + R), dim3(WAVE), 0, S(stream), v);
  }
  /* every SIP is in its home cell again (a fill kernel, not a memset: the step stays graph-capturable) */
  hipLaunchKernelGGL(k_fill_i32, dim3((R * nc + 255) / 256), dim3(256), 0, S(stream), e->away, 0, R * nc);
  if (e->mig_cnt)
    hipLaunchKernelGGL(k_fill_i32, dim3((R * nc + 255) / 256), dim3(256), 0, S(stream), e->mig_cnt, 0, R * nc);
  LAUNCH_CHECK();
  e->cur = 1 - e->cur;
  e->ccur = 1 - e->ccur;
  /* the work list of the persistent passes from the new segments (only when a persistent pass is on; the
     list is a permutation of the cells at all times, a stale one only balances worse) */
  if (e->work && (e->col_work_per_cu > 0 || e->dml_work_per_cu > 0) && (rc = lm_build_work_list(e, stream)))
    return rc;
  if (e->ev_mode == 2) return lm_build_slot_of(e, stream);
  return 0;
}

int lm_advection_increments(LmEngine *e, int32_t second, void *stream) {
  int rc = check_engine(e);
  if (rc) return rc;
  KView v = make_view(e);
  const int ncell = e->n_runs * e->nc;
  if (e->tape_fields) {
    if (e->tape_inc < 0 || e->tape_inc >= e->tape_ninc) return set_err(LM_ERR_ARG, "tape full (increments)");
    hipLaunchKernelGGL(k_tape_fields, dim3((ncell + 255) / 256), dim3(256), 0, S(stream), v,
                       e->tape_fields + (size_t)e->tape_inc * 2 * ncell);
    e->tape_inc += 1;
  }
  if (e->act_relaxation) {
    const int nt = e->n_runs * e->nz;
    hipLaunchKernelGGL(k_colsum, dim3((nt + 127) / 128), dim3(128), 0, S(stream), v, e->dt_sub, 1);
  }
  hipLaunchKernelGGL(k_adv_incr, dim3((ncell + 127) / 128), dim3(128), 0, S(stream), v, e->dt_sub, second,
                     e->act_relaxation);
  LAUNCH_CHECK();
  return 0;
}

int lm_particle_substep(LmEngine *e, double dt_sub, double dt_sub_pos, void *stream) {
  int rc = check_engine(e);
  if (rc) return rc;
  if (e->tape_part && (e->tape_sub < 0 || e->tape_sub >= e->tape_nsub))
    return set_err(LM_ERR_ARG, "tape full (substeps)");
  KView v = make_view(e);
  const long long n = e->n_tot;
  const int ncell = e->n_runs * e->nc;
  if (n > 0) {
    ProfScope ps(e, 0, S(stream));
    const dim3 b(part_block(e));
    const int spt = part_spt(e);
    if (e->tape_part) {
      hipLaunchKernelGGL(k_ccd, dim3((ncell + 255) / 256), dim3(256), 0, S(stream), v);
      const dim3 gt((unsigned)((n + TAPE_BLOCK - 1) / TAPE_BLOCK));
      if (e->tape_mode == 1) {
        /* one pass on LmDual<5> */
        if (e->solute == 0)
          hipLaunchKernelGGL((k_particles_tape<0, 0, 1>), gt, dim3(TAPE_BLOCK), 0, S(stream), v, n, dt_sub,
                             dt_sub_pos);
        else
          hipLaunchKernelGGL((k_particles_tape<1, 0, 1>), gt, dim3(TAPE_BLOCK), 0, S(stream), v, n, dt_sub,
                             dt_sub_pos);
      } else if (e->solute == 0) {
        hipLaunchKernelGGL((k_tape_tps<0, 2>), gt, dim3(TAPE_BLOCK), 0, S(stream), v, n, dt_sub);
        hipLaunchKernelGGL((k_particles_tape<0, 1, 2>), gt, dim3(TAPE_BLOCK), 0, S(stream), v, n, dt_sub,
                           dt_sub_pos);
      } else {
        hipLaunchKernelGGL((k_tape_tps<1, 2>), gt, dim3(TAPE_BLOCK), 0, S(stream), v, n, dt_sub);
        hipLaunchKernelGGL((k_particles_tape<1, 1, 2>), gt, dim3(TAPE_BLOCK), 0, S(stream), v, n, dt_sub,
                           dt_sub_pos);
      }
    } else if (spt != 2 && e->part_wpe > 0) {
      if (e->solute == 0)
        launch_particles_occ<0>(e->part_wpe, part_grid(e, 1), b, S(stream), v, n, dt_sub, dt_sub_pos);
      else
        launch_particles_occ<1>(e->part_wpe, part_grid(e, 1), b, S(stream), v, n, dt_sub, dt_sub_pos);
    } else if (spt == 2) {
      if (e->solute == 0)
        hipLaunchKernelGGL((k_particles<0, false, false, 2>), part_grid(e, 2), b, 0, S(stream), v, n, dt_sub,
                           dt_sub_pos);
      else
        hipLaunchKernelGGL((k_particles<1, false, false, 2>), part_grid(e, 2), b, 0, S(stream), v, n, dt_sub,
                           dt_sub_pos);
    } else if (e->solute == 0) {
      hipLaunchKernelGGL((k_particles<0, false, false, 1>), part_grid(e, 1), b, 0, S(stream), v, n, dt_sub, dt_sub_pos);
    } else {
      hipLaunchKernelGGL((k_particles<1, false, false, 1>), part_grid(e, 1), b, 0, S(stream), v, n, dt_sub, dt_sub_pos);
    }
  }
  {
    ProfScope ps(e, 2, S(stream));
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, S(stream), e->ovf_cnt + 1, 0, 1);
    const dim3 ogrid(ncell < OVF_GRID ? ncell : OVF_GRID);
    int32_t *list1 = e->ovf_list, *cnt1 = e->ovf_cnt + 1;
    if (e->mig_cnt && e->mig_cap <= DML2_L && e->dml_work_per_cu <= 0 && cell_capacity(e) <= DML2_CAP) {
      /* two cells per wavefront; the others through the 256- and 1024-member passes (second list: ovf_list +
         ncell, counter ovf_cnt[4]) */
      int32_t *list2 = e->ovf_list + ncell, *cnt2 = e->ovf_cnt + 4;
      hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, S(stream), cnt2, 0, 1);
      hipLaunchKernelGGL(k_dml2, dim3((unsigned)((ncell + WAVE / DML2_L - 1) / (WAVE / DML2_L))), dim3(WAVE), 0,
                         S(stream), v, list1, cnt1);
      hipLaunchKernelGGL(k_dml<256>, ogrid, dim3(WAVE), 0, S(stream), v, PASS_OVF, 0, list1, cnt1, list2, cnt2);
      hipLaunchKernelGGL(k_dml<1024>, ogrid, dim3(WAVE), 0, S(stream), v, PASS_OVF, 0, list2, cnt2, nullptr,
                         nullptr);
    } else {
      dim3 grid;
      const int mode = main_pass(e, e->dml_work_per_cu, e->dml_work_chunk, 1, &grid, S(stream));
      if (cell_capacity(e) <= 256)
        hipLaunchKernelGGL(k_dml<256>, grid, dim3(WAVE), 0, S(stream), v, mode, e->dml_work_chunk, list1, cnt1,
                           list1, cnt1);
      else
        hipLaunchKernelGGL(k_dml<512>, grid, dim3(WAVE), 0, S(stream), v, mode, e->dml_work_chunk, list1, cnt1,
                           list1, cnt1);
      hipLaunchKernelGGL(k_dml<1024>, ogrid, dim3(WAVE), 0, S(stream), v, PASS_OVF, 0, list1, cnt1, nullptr,
                         nullptr);
    }
  }
  hipLaunchKernelGGL(k_water, dim3(e->n_runs), dim3(WAVE), 0, S(stream), v);
  LAUNCH_CHECK();
  e->ccur = 1 - e->ccur;
  return 0;
}

int lm_grid_substep(LmEngine *e, void *stream) {
  int rc = check_engine(e);
  if (rc) return rc;
  KView v = make_view(e);
  const int ncell = e->n_runs * e->nc;
  if (e->tape_grid && (e->tape_sub < 0 || e->tape_sub >= e->tape_nsub))
    return set_err(LM_ERR_ARG, "tape full (substeps)");
  if (e->tape_grid)
    hipLaunchKernelGGL(k_grid_jac, dim3((ncell + 255) / 256), dim3(256), 0, S(stream), v,
                       e->tape_grid + (size_t)e->tape_sub * 30 * ncell);
  hipLaunchKernelGGL(k_grid, dim3((ncell + 255) / 256), dim3(256), 0, S(stream), v);
  LAUNCH_CHECK();
  if (e->tape_part || e->tape_grid) e->tape_sub += 1;
  return 0;
}

int lm_update_r_l(LmEngine *e, void *stream) {
  int rc = check_engine(e);
  if (rc) return rc;
  KView v = make_view(e);
  hipLaunchKernelGGL(k_rl, dim3(e->n_runs * e->nc), dim3(WAVE), 0, S(stream), v);
  LAUNCH_CHECK();
  return 0;
}

int lm_load_reference_layout(LmEngine *e, int32_t run, int64_t base, int64_t n, const double *pos,
                             const double *vel, const int64_t *cells, const double *m_w, const double *m_s,
                             const double *xi, const uint8_t *active_ids, void *stream) {
  int rc = check_engine(e);
  if (rc) return rc;
  if (run < 0 || run >= e->n_runs) return set_err(LM_ERR_ARG, "run index out of range");
  if (base < 0 || n < 0 || base + n > e->n_tot) return set_err(LM_ERR_ARG, "run slots outside the engine");
  if (n > 0 && (!pos || !vel || !cells || !m_w || !m_s || !xi || !active_ids))
    return set_err(LM_ERR_ARG, "array is NULL");
  const int nc = e->nc;
  KView v = make_view(e);
  /* scratch: counts (scan_buf, nc + 1 with the total), cursors (ovf_list), the cell lists (the idle id
     buffer), the inactive prefix (the idle x buffer, N + 1 ints) */
  int32_t *cnt = e->scan_buf, *fill = e->ovf_list, *list = e->id[1 - e->cur];
  int32_t *dead = (int32_t *)e->x[1 - e->cur];
  const unsigned gN = (unsigned)((n + 255) / 256), gC = (unsigned)((nc + 1 + 255) / 256);
  hipLaunchKernelGGL(k_fill_i32, dim3(gC), dim3(256), 0, S(stream), cnt, 0, nc + 1);
  hipLaunchKernelGGL(k_fill_i32, dim3(gC), dim3(256), 0, S(stream), fill, 0, nc);
  if (n > 0)
    hipLaunchKernelGGL(k_lay_count, dim3(gN), dim3(256), 0, S(stream), (long long)n, cells, active_ids, e->nx, e->nz,
                       cnt, dead, e->err);
  hipLaunchKernelGGL(k_scan1_i32, dim3(1), dim3(1024), 0, S(stream), cnt, nc);
  hipLaunchKernelGGL(k_scan1_i32, dim3(1), dim3(1024), 0, S(stream), dead, (int)n);
  if (n > 0)
    hipLaunchKernelGGL(k_lay_scatter, dim3(gN), dim3(256), 0, S(stream), (long long)n, cells, active_ids, e->nx,
                       e->nz, cnt, fill, list);
  hipLaunchKernelGGL(k_lay_cell, dim3(nc), dim3(WAVE), 0, S(stream), v, (int)run, (long long)n, (long long)base, pos,
                     vel, m_w, m_s, xi, cnt, list);
  hipLaunchKernelGGL(k_lay_dead, dim3(gN > 0 ? gN : 1), dim3(256), 0, S(stream), v, (int)run, (long long)n,
                     (long long)base, (long long)e->n_tot, pos, vel, cells, active_ids, m_w, m_s, xi, cnt, dead);
  /* every SIP sits in its home cell: no direction bits, no migrants */
  hipLaunchKernelGGL(k_fill_i32, dim3((nc + 255) / 256), dim3(256), 0, S(stream), e->away + (size_t)run * nc, 0, nc);
  if (e->mig_cnt)
    hipLaunchKernelGGL(k_fill_i32, dim3((nc + 255) / 256), dim3(256), 0, S(stream), e->mig_cnt + (size_t)run * nc, 0,
                       nc);
  LAUNCH_CHECK();
  return 0;
}

int lm_store_reference_layout(LmEngine *e, int32_t run, int64_t base, int64_t n, double *pos, double *vel,
                              int64_t *cells, double *rel_pos, double *m_w, double *m_s, double *xi,
                              uint8_t *active_ids, double *T_p, void *stream) {
  int rc = check_engine(e);
  if (rc) return rc;
  if (run < 0 || run >= e->n_runs) return set_err(LM_ERR_ARG, "run index out of range");
  if (base < 0 || n < 0 || base + n > e->n_tot) return set_err(LM_ERR_ARG, "run slots outside the engine");
  if (n > 0 && (!pos || !vel || !cells || !m_w || !m_s || !xi || !active_ids))
    return set_err(LM_ERR_ARG, "array is NULL");
  if (n == 0) return 0;
  KView v = make_view(e);
  hipLaunchKernelGGL(k_lay_store, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, S(stream), v, (int)run,
                     (long long)n, (long long)base, pos, vel, cells, rel_pos, m_w, m_s, xi, active_ids, T_p,
                     (const int32_t *)e->cell[1 - e->ccur]);
  LAUNCH_CHECK();
  return 0;
}

int lm_derived_fields(LmEngine *e, double *out, void *stream) {
  int rc = check_engine(e);
  if (rc) return rc;
  if (!out) return set_err(LM_ERR_ARG, "output is NULL");
  KView v = make_view(e);
  hipLaunchKernelGGL(k_derived, dim3(e->n_runs * e->nc), dim3(WAVE), 0, S(stream), v, out);
  LAUNCH_CHECK();
  return 0;
}

int lm_ensemble_stats(const double *fields, int32_t n_seeds, int64_t n_vals, double *mean, double *std,
                      void *stream) {
  if (!fields || !mean || !std || n_seeds <= 0 || n_vals <= 0) return set_err(LM_ERR_ARG, "bad ensemble arguments");
  hipLaunchKernelGGL(k_ensemble, dim3((unsigned)((n_vals + 255) / 256)), dim3(256), 0, S(stream), fields, n_seeds,
                     (long long)n_vals, mean, std);
  LAUNCH_CHECK();
  return 0;
}

int lm_substep(LmEngine *e, double dt_sub, double dt_sub_pos, void *stream) {
  int rc = lm_particle_substep(e, dt_sub, dt_sub_pos, stream);
  if (rc) return rc;
  return lm_grid_substep(e, stream);
}

static int subloop(LmEngine *e, int no_cond_steps, int no_col_steps, void *stream) {
  int rc;
  if (e->act_collisions && no_col_steps == 1)
    if ((rc = lm_collision_step(e, stream))) return rc;
  for (int n = 0; n < no_cond_steps; ++n) {
    if (e->act_collisions && no_col_steps > 1)
      if ((rc = lm_collision_step(e, stream))) return rc;
    if ((rc = lm_substep(e, e->dt_sub, e->dt_sub, stream))) return rc;
  }
  if (e->act_collisions && no_col_steps > no_cond_steps)
    if ((rc = lm_collision_step(e, stream))) return rc;
  return 0;
}

int lm_adv_steps(LmEngine *e, int32_t n_steps, void *stream) {
  int rc = check_engine(e);
  if (rc) return rc;
  for (int s = 0; s < n_steps; ++s) {
    if ((rc = lm_position_step(e, e->dt_sub_half, stream))) return rc;
    if ((rc = lm_reindex(e, stream))) return rc;
    if ((rc = lm_advection_increments(e, 0, stream))) return rc;
    int ncol = e->no_col_per_adv == 1 ? 0 : (e->no_col_per_adv == 2 ? 1 : e->scale_dt_cond);
    if ((rc = subloop(e, e->scale_dt_cond, ncol, stream))) return rc;
    if ((rc = lm_advection_increments(e, 1, stream))) return rc;
    ncol = (e->no_col_per_adv == 1 || e->no_col_per_adv == 2) ? 1 : e->scale_dt_cond;
    if ((rc = subloop(e, e->scale_dt_cond - 1, ncol, stream))) return rc;
    if ((rc = lm_substep(e, e->dt_sub, e->dt_sub_half, stream))) return rc;
  }
  return 0;
}

int lm_set_col_step(LmEngine *e, uint32_t step, void *stream) {
  int rc = check_engine(e);
  if (rc) return rc;
  hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(64), 0, S(stream), e->col_step_dev, step);
  LAUNCH_CHECK();
  e->col_step = step;
  return 0;
}

/* A captured sequence of advection steps: the executable graph and what one launch advances. */
struct LmGraph {
  hipGraphExec_t exec;
  uint32_t col_steps;
  int cur, ccur; /* buffer indices the captured kernels were bound to */
};

int lm_graph_capture(LmEngine *e, int32_t n_steps, void *stream, void **graph) {
  int rc = check_engine(e);
  if (rc) return rc;
  if (!graph || n_steps <= 0 || (n_steps & 1))
    return set_err(LM_ERR_ARG, "graph capture needs an even, positive number of advection steps "
                               "(one step swaps the SIP buffers once)");
  if (e->prof) return set_err(LM_ERR_ARG, "disable kernel timing (lm_prof_disable) before capturing a graph");
  if (!stream) return set_err(LM_ERR_ARG, "graph capture needs a non-default stream");
  const uint32_t step0 = e->col_step;
  const int cur0 = e->cur, ccur0 = e->ccur;
  HIP_TRY(hipStreamBeginCapture(S(stream), hipStreamCaptureModeThreadLocal));
  rc = lm_adv_steps(e, n_steps, stream);
  hipGraph_t g = nullptr;
  const hipError_t ec = hipStreamEndCapture(S(stream), &g);
  const uint32_t col_steps = e->col_step - step0;
  e->col_step = step0; /* nothing ran yet */
  if (rc || ec != hipSuccess || e->cur != cur0 || e->ccur != ccur0) {
    if (g) (void)hipGraphDestroy(g);
    e->cur = cur0;
    e->ccur = ccur0;
    if (rc) return rc;
    return set_err(LM_ERR_ARG, "stream capture of the advection steps failed");
  }
  hipGraphExec_t ex;
  const hipError_t ei = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  HIP_TRY(ei);
  LmGraph *h = new LmGraph;
  h->exec = ex;
  h->col_steps = col_steps;
  h->cur = cur0;
  h->ccur = ccur0;
  *graph = h;
  return 0;
}

int lm_graph_launch(LmEngine *e, void *graph, void *stream) {
  if (!e || !graph) return set_err(LM_ERR_ARG, "engine/graph is NULL");
  LmGraph *h = (LmGraph *)graph;
  if (e->cur != h->cur || e->ccur != h->ccur)
    return set_err(LM_ERR_ARG, "the engine's buffer indices differ from the graph's (an odd number of eager "
                               "steps ran since the capture)");
  HIP_TRY(hipGraphLaunch(h->exec, S(stream)));
  e->col_step += h->col_steps;
  return 0;
}

int lm_graph_destroy(void *graph) {
  if (!graph) return 0;
  LmGraph *h = (LmGraph *)graph;
  (void)hipGraphExecDestroy(h->exec);
  delete h;
  return 0;
}

int lm_prof_enable(LmEngine *e, int32_t max_events) {
  if (!e || max_events <= 0) return set_err(LM_ERR_ARG, "bad prof arguments");
  lm_prof_disable(e);
  LmProf *p = new LmProf;
  p->max_ev = max_events;
  p->n = 0;
  p->a = new hipEvent_t[max_events];
  p->b = new hipEvent_t[max_events];
  p->kind = new int[max_events];
  for (int i = 0; i < max_events; ++i) {
    HIP_TRY(hipEventCreate(&p->a[i]));
    HIP_TRY(hipEventCreate(&p->b[i]));
  }
  e->prof = p;
  return 0;
}

int lm_prof_collect(LmEngine *e, double *ms, int32_t *count, int32_t nkinds) {
  if (!e || !e->prof) return set_err(LM_ERR_ARG, "profiling not enabled");
  LmProf *p = (LmProf *)e->prof;
  for (int k = 0; k < nkinds; ++k) {
    ms[k] = 0.0;
    count[k] = 0;
  }
  for (int i = 0; i < p->n; ++i) {
    HIP_TRY(hipEventSynchronize(p->b[i]));
    float t = 0.f;
    HIP_TRY(hipEventElapsedTime(&t, p->a[i], p->b[i]));
    if (p->kind[i] < nkinds) {
      ms[p->kind[i]] += t;
      count[p->kind[i]] += 1;
    }
  }
  p->n = 0;
  return 0;
}

int lm_prof_disable(LmEngine *e) {
  if (!e || !e->prof) return 0;
  LmProf *p = (LmProf *)e->prof;
  for (int i = 0; i < p->max_ev; ++i) {
    (void)hipEventDestroy(p->a[i]);
    (void)hipEventDestroy(p->b[i]);
  }
  delete[] p->a;
  delete[] p->b;
  delete[] p->kind;
  delete p;
  e->prof = nullptr;
  return 0;
}

int lm_engine_check(LmEngine *e, void *stream) {
  if (!e || !e->err) return set_err(LM_ERR_ARG, "engine/err NULL");
  int32_t h[2] = {0, 0};
  HIP_TRY(hipMemcpyAsync(h, e->err, sizeof(h), hipMemcpyDeviceToHost, S(stream)));
  HIP_TRY(hipStreamSynchronize(S(stream)));
  if (h[0]) {
    const char *what = h[0] == LM_ERR_CELL_CAP   ? "cell holds more SIPs than the kernel capacity"
                       : h[0] == LM_ERR_LOST_SIP ? "SIP moved more than one cell between reindexes"
                       : h[0] == LM_ERR_REM_CAP  ? "too many SIPs removed in one substep"
                       : h[0] == LM_ERR_OUT_OF_DOMAIN ? "SIP position outside the grid"
                       : h[0] == LM_ERR_RND_SHORT ? "injected uniform stream shorter than the pairs it must cover"
                       : h[0] == LM_ERR_EV_CAP    ? "collision-event log of the adjoint window is full"
                                                       : "device error";
    snprintf(g_errmsg, sizeof(g_errmsg), "%s (code %d, detail %d)", what, h[0], h[1]);
    return h[0];
  }
  return 0;
}

int lm_math_probe(int32_t which, int64_t n, const double *x, const double *y, double *out, void *stream) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_math_probe, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, S(stream), which, (long long)n,
                     x, y, out);
  LAUNCH_CHECK();
  return 0;
}

int lm_kernel_index_check(int64_t n, const double *R, double lnR0, double lnf, int32_t nb, int32_t *out_fast,
                          int32_t *out_exact, double *sweep, void *stream) {
  if (n < 0 || nb < 1 || (n > 0 && (!R || !out_fast || !out_exact))) return set_err(LM_ERR_ARG, "bad arguments");
  KView v;
  memset(&v, 0, sizeof(v));
  v.lnR0 = lnR0;
  v.lnf = lnf;
  v.nb = nb;
  ki_setup(v);
  if (n > 0)
    hipLaunchKernelGGL(k_ki_probe, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, S(stream), v, (long long)n, R,
                       out_fast, out_exact);
  if (sweep) {
    float a = (float)KI_RMIN, b = (float)KI_RMAX;
    uint32_t lo, hi;
    memcpy(&lo, &a, 4);
    memcpy(&hi, &b, 4);
    hipLaunchKernelGGL(k_ki_consts, dim3(1), dim3(64), 0, S(stream), v, sweep, hi - lo);
    hipLaunchKernelGGL(k_log2_sweep, dim3(4096), dim3(256), 0, S(stream), lo, hi - lo, (unsigned long long *)sweep);
  }
  LAUNCH_CHECK();
  return 0;
}

static int32_t *g_box_err = nullptr;

static int box_prepare() {
  if (ensure_phys()) return set_err(LM_ERR_ARG, "cannot upload constants to the device");
  if (!g_box_err) {
    HIP_TRY(hipMalloc(&g_box_err, 16));
    HIP_TRY(hipMemset(g_box_err, 0, 16));
  }
  return 0;
}

int lm_golovin_step(int32_t n_boxes, const int32_t *off, double *xis, double *masses, double dt_over_dV,
                    int64_t *no_cols, uint64_t seed, uint32_t step, const double *rnd_stream, void *stream) {
  if (n_boxes <= 0 || !off || !xis || !masses || !no_cols) return set_err(LM_ERR_ARG, "bad box arguments");
  int rc = box_prepare();
  if (rc) return rc;
  hipLaunchKernelGGL(k_golovin, dim3(n_boxes), dim3(WAVE), 0, S(stream), n_boxes, off, xis, masses, dt_over_dV,
                     (const double *)nullptr, 0, 0.0, 1.0, (long long *)no_cols, (unsigned long long)seed, step,
                     rnd_stream, g_box_err);
  LAUNCH_CHECK();
  return 0;
}

int lm_kernel_grid_m_step(int32_t n_boxes, const int32_t *off, double *xis, double *masses, double dt_over_dV,
                          const double *kernel_grid, int32_t nb, double m_kernel_low_log, double bin_factor_m_log,
                          int64_t *no_cols, uint64_t seed, uint32_t step, const double *rnd_stream, void *stream) {
  if (n_boxes <= 0 || !off || !xis || !masses || !kernel_grid || nb <= 0 || !no_cols)
    return set_err(LM_ERR_ARG, "bad box arguments");
  int rc = box_prepare();
  if (rc) return rc;
  hipLaunchKernelGGL(k_golovin, dim3(n_boxes), dim3(WAVE), 0, S(stream), n_boxes, off, xis, masses, dt_over_dV,
                     kernel_grid, nb, m_kernel_low_log, bin_factor_m_log, (long long *)no_cols,
                     (unsigned long long)seed, step, rnd_stream, g_box_err);
  LAUNCH_CHECK();
  return 0;
}

int lm_ecol_box_step(int32_t n_boxes, const int32_t *off, double *xis, double *masses, double *radii, double *vel,
                     const double *rho, double dt_over_dV, const double *E, int32_t nb, double lnR0, double lnf,
                     int64_t *no_cols, uint64_t seed, uint32_t step, const double *rnd_stream, int32_t update_beard,
                     void *stream) {
  if (n_boxes <= 0 || !off || !xis || !masses || !radii || !vel || !rho || !E || nb <= 0 || !no_cols)
    return set_err(LM_ERR_ARG, "bad box arguments");
  int rc = box_prepare();
  if (rc) return rc;
  hipLaunchKernelGGL(k_ecol_box, dim3(n_boxes), dim3(WAVE), 0, S(stream), n_boxes, off, xis, masses, radii, vel,
                     rho, dt_over_dV, E, nb, lnR0, lnf, (long long *)no_cols, (unsigned long long)seed, step,
                     rnd_stream, g_box_err, update_beard);
  LAUNCH_CHECK();
  return 0;
}

int lm_box2d_step(int32_t n_boxes, const int32_t *off, double *xis, double *m_w, double *m_s, double *radii,
                  const double *vel_x, const double *vel_z, const double *rho, double dt_over_dV, const double *E,
                  double E_const, int32_t nb, double lnR0, double lnf, int64_t *no_cols, uint64_t seed, uint32_t step,
                  const double *rnd_stream, void *stream) {
  if (n_boxes <= 0 || !off || !xis || !m_w || !m_s || !radii || !vel_x || !vel_z || !rho || !no_cols ||
      (E && nb <= 0))
    return set_err(LM_ERR_ARG, "bad box arguments");
  int rc = box_prepare();
  if (rc) return rc;
  hipLaunchKernelGGL(k_box2d, dim3(n_boxes), dim3(WAVE), 0, S(stream), n_boxes, off, xis, m_w, m_s, radii, vel_x,
                     vel_z, rho, dt_over_dV, E, E_const, nb, lnR0, lnf, (long long *)no_cols,
                     (unsigned long long)seed, step, rnd_stream, g_box_err);
  LAUNCH_CHECK();
  return 0;
}

int lm_box_check(void *stream) {
  if (!g_box_err) return 0;
  int32_t h[4];
  HIP_TRY(hipMemcpyAsync(h, g_box_err, sizeof(h), hipMemcpyDeviceToHost, S(stream)));
  HIP_TRY(hipStreamSynchronize(S(stream)));
  if (h[0]) {
    HIP_TRY(hipMemsetAsync(g_box_err, 0, 16, S(stream)));
    HIP_TRY(hipStreamSynchronize(S(stream)));
    char msg[160];
    snprintf(msg, sizeof msg, "box %d holds more than %d SIPs (box kernels keep a box in LDS)", h[1], BOX_CAP);
    return set_err(h[0], msg);
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* all-cells collision step on the reference's arrays (all_or_nothing.py:754-846) */
/* ------------------------------------------------------------------------ */
/* device scratch of one call: member lists, pre-step densities, the per-cell table and counters */
struct RefWork {
  int32_t *cnt;      /* nc + 1: counts, then offsets (exclusive scan, total at nc) */
  int32_t *fill;     /* nc: scatter cursors */
  int32_t *list;     /* n: SIP indices grouped by cell */
  double *rho;       /* n: pre-step density of each SIP (events only) */
  double *cc;        /* nc * CC_N: cell temperature and water density */
  int32_t *ovf_list; /* 2 nc: cells over 256 members, then cells over 1024 */
  int32_t *small;    /* 8: ovf_cnt[2], err[4], step counter, count of cells over 1024 members */
  long long *n_eq;   /* 1: near-equal collections */
  long long *rnd_off; /* nc + 2: injected-stream offsets */
};

static constexpr size_t ref_up(size_t b) { return (b + 255) & ~(size_t)255; }
static constexpr size_t ref_bytes(long long n, long long nc) {
  return ref_up(4 * (nc + 1)) + ref_up(4 * nc) + ref_up(4 * n) + ref_up(8 * n) + ref_up(8 * CC_N * nc) +
         ref_up(8 * nc) + ref_up(4 * 8) + ref_up(8) + ref_up(8 * (nc + 2));
}

static RefWork ref_carve(char *base, long long n, long long nc) {
  auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
  RefWork w;
  size_t o = 0;
  w.cnt = (int32_t *)(base + o);
  o += up(4 * (nc + 1));
  w.fill = (int32_t *)(base + o);
  o += up(4 * nc);
  w.list = (int32_t *)(base + o);
  o += up(4 * n);
  w.rho = (double *)(base + o);
  o += up(8 * n);
  w.cc = (double *)(base + o);
  o += up(8 * CC_N * nc);
  w.ovf_list = (int32_t *)(base + o);
  o += up(8 * nc);
  w.small = (int32_t *)(base + o);
  o += up(4 * 8);
  w.n_eq = (long long *)(base + o);
  o += up(8);
  w.rnd_off = (long long *)(base + o);
  return w;
}

/* error words of the last reference-layout calls (device pointers), per device: allocating, no-alloc */
static int32_t *g_ref_err[64][2];

static int ref_args(long long n_sips, const double *xis, const double *m_w, const double *m_s, const double *vel,
                    const double *grid_T, const int64_t *cells, int nx, int nz, const double *E, int nb,
                    const int64_t *no_cols, int solute, const double *rnd_stream, long long rnd_len) {
  if (n_sips < 0 || nx <= 0 || nz <= 0 || (n_sips > 0 && (!xis || !m_w || !m_s || !vel || !cells)) || !grid_T ||
      !no_cols || solute < 0 || solute > 1 || (E && nb <= 0) || (rnd_stream && rnd_len < 0))
    return set_err(LM_ERR_ARG, "bad all-cells collision arguments");
  if ((long long)nx * nz > 0x3fffffff || n_sips > 0x7ffffff0) return set_err(LM_ERR_ARG, "grid or SIP count too large");
  return 0;
}

/* the step on caller-owned arrays with the scratch w; no host synchronisation, no allocation */
static int ref_step(const RefWork &w, long long n_sips, double *xis, double *m_w, double *m_s, const double *vel,
                    const double *grid_T, const int64_t *cells, int nx, int nz, double dt_over_dV, const double *E,
                    int nb, double lnR0, double lnf, double E_const, double E_max, int64_t *no_cols, int solute,
                    uint64_t seed, uint32_t step, const double *rnd_stream, long long rnd_len, hipStream_t st,
                    char *big_ws = nullptr, int big_cap = 0) {
  const int nc = nx * nz;
  const long long N = n_sips;
  const unsigned gN = (unsigned)((N + 255) / 256), gC = (unsigned)((nc + 255) / 256);
  /* what must start at zero: counts, cursors, overflow counters, the error word (so the check reports this
     call's errors), the near-equal counter */
  hipLaunchKernelGGL(k_fill_i32, dim3((nc + 1 + 255) / 256), dim3(256), 0, st, w.cnt, 0, nc + 1);
  hipLaunchKernelGGL(k_fill_i32, dim3(gC), dim3(256), 0, st, w.fill, 0, nc);
  hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, st, w.small, 0, 6); /* overflow counters, error word */
  hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, st, w.small + 7, 0, 1);
  hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, st, (int32_t *)w.n_eq, 0, 2);
  hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(64), 0, st, (uint32_t *)(w.small + 6), step);
  hipLaunchKernelGGL(k_ref_cc, dim3(gC), dim3(256), 0, st, nc, grid_T, w.cc);
  if (N > 0) hipLaunchKernelGGL(k_ref_count, dim3(gN), dim3(256), 0, st, N, cells, nx, nz, w.cnt);
  hipLaunchKernelGGL(k_scan1_i32, dim3(1), dim3(1024), 0, st, w.cnt, nc);
  if (N > 0) hipLaunchKernelGGL(k_ref_scatter, dim3(gN), dim3(256), 0, st, N, cells, nx, nz, w.cnt, w.fill, w.list);
  KView v;
  memset(&v, 0, sizeof(v));
  v.n_runs = 1;
  v.nx = nx;
  v.nz = nz;
  v.nc = nc;
  v.solute = solute;
  v.dt_col_over_dV = dt_over_dV;
  v.E = E;
  v.nb = nb;
  v.e_const_mode = E ? 0 : 1;
  v.E_const = E_const;
  v.lnR0 = lnR0;
  v.lnf = lnf;
  ki_setup(v);
  v.E_max = E ? E_max : 0.0;
  v.seed = seed;
  v.col_step_dev = (const uint32_t *)(w.small + 6);
  v.xi = xis;
  v.m_w = m_w;
  v.m_s = m_s;
  v.u = (double *)vel;
  v.w = (double *)vel + N;
  v.cc = w.cc;
  v.dmxi = w.rho;
  v.no_cols = (long long *)no_cols;
  v.n_equal_xi = w.n_eq;
  v.err = w.small + 2;
  v.ovf_list = w.ovf_list;
  v.ovf_cnt = w.small;
  v.mlist = w.list;
  v.moff = w.cnt;
  v.rnd_stream = rnd_stream;
  v.rnd_len = rnd_len;
  v.rnd_off = w.rnd_off;
  v.big_ws = big_ws;
  v.big_cap = big_cap;
  if (rnd_stream) {
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, st, (int32_t *)(w.rnd_off + nc), 0, 4);
    hipLaunchKernelGGL(k_ref_count_pairs, dim3(gC), dim3(256), 0, st, nc, w.cnt, w.rnd_off);
    hipLaunchKernelGGL(k_scan_pairs, dim3(1), dim3(1024), 0, st, w.rnd_off, nc);
  }
  const bool tab = E != nullptr, inj = rnd_stream != nullptr;
  /* main pass at the 256-member capacity, larger cells queued for the 1024-member pass, and with a workspace
     cells over 1024 members for the pass over cells of any size */
  const ColLists main_io = {nullptr, nullptr, w.ovf_list, w.small};
  const ColLists ovf_io = {w.ovf_list, w.small, big_ws ? w.ovf_list + nc : nullptr, big_ws ? w.small + 7 : nullptr};
  launch_collision<256>(v, tab, inj, dim3(nc), st, PASS_CELLS, 0, main_io);
  launch_collision<1024>(v, tab, inj, dim3(nc < OVF_GRID ? nc : OVF_GRID), st, PASS_OVF, 0, ovf_io);
  if (big_ws) {
    const ColLists big_io = {w.ovf_list + nc, w.small + 7, nullptr, nullptr};
    launch_collision<0>(v, tab, inj, dim3(REF_BIG_BLOCKS), st, PASS_OVF, 0, big_io);
  }
  LAUNCH_CHECK();
  return 0;
}

/* the allocating form's workspace, per device, grown on demand: the step's scratch, then REF_BIG_BLOCKS regions
   for cells of over 1024 members */
static char *g_ref_ws[64];
static size_t g_ref_ws_bytes[64];

int lm_collision_all_cells(int64_t n_sips, double *xis, double *m_w, double *m_s, const double *vel,
                           const double *grid_T, const int64_t *cells, int32_t nx, int32_t nz, double dt_over_dV,
                           const double *E_col, int32_t nb, double lnR0, double lnf, double E_const, double E_max,
                           int64_t *no_cols, int32_t solute, uint64_t seed, uint32_t step, const double *rnd_stream,
                           int64_t rnd_len, void *stream) {
  int rc = ref_args(n_sips, xis, m_w, m_s, vel, grid_T, cells, nx, nz, E_col, nb, no_cols, solute, rnd_stream,
                    rnd_len);
  if (rc) return rc;
  if (ensure_phys()) return set_err(LM_ERR_ARG, "cannot upload constants to the device");
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) return set_err(LM_ERR_ARG, "device index out of range");
  const int big_cap = n_sips < 2 ? 2 : (n_sips > LM_BIG_CELL_MAX ? LM_BIG_CELL_MAX : (int)n_sips);
  const size_t base_bytes = ref_bytes(n_sips, (long long)nx * nz);
  const size_t need = base_bytes + (size_t)REF_BIG_BLOCKS * big_stride(big_cap);
  if (g_ref_ws_bytes[dev] < need) {
    /* a larger workspace: the previous one may still be in use by queued work on this stream */
    HIP_TRY(hipStreamSynchronize(S(stream)));
    if (g_ref_ws[dev]) HIP_TRY(hipFree(g_ref_ws[dev]));
    g_ref_ws[dev] = nullptr;
    g_ref_ws_bytes[dev] = 0;
    HIP_TRY(hipMalloc((void **)&g_ref_ws[dev], need));
    g_ref_ws_bytes[dev] = need;
    HIP_TRY(hipMemset(g_ref_ws[dev], 0, need));
  }
  const RefWork w = ref_carve(g_ref_ws[dev], n_sips, (long long)nx * nz);
  g_ref_err[dev][0] = w.small + 2;
  return ref_step(w, n_sips, xis, m_w, m_s, vel, grid_T, cells, nx, nz, dt_over_dV, E_col, nb, lnR0, lnf, E_const,
                  E_max, no_cols, solute, seed, step, rnd_stream, rnd_len, S(stream), g_ref_ws[dev] + base_bytes,
                  big_cap);
}

/*
 * The same step with its scratch in `scratch`, a device array of fixed size that lives as long as the process
 * (one copy per GPU): for callers that cannot allocate (graph capture, serving loops).  Host code never
 * passes the array's name to a kernel: it is a host-side stand-in, and its device address (per GPU) comes
 * from hipGetSymbolAddress on the first call with that GPU current.  Calls that overlap on one GPU share
 * the array and race.
 */
static __device__ __align__(256) char scratch[ref_bytes(LM_NOALLOC_MAX_SIPS, LM_NOALLOC_MAX_CELLS)];
static char *g_scratch_ptr[64];

int lm_collision_all_cells_noalloc(int64_t n_sips, double *xis, double *m_w, double *m_s, const double *vel,
                                   const double *grid_T, const int64_t *cells, int32_t nx, int32_t nz,
                                   double dt_over_dV, const double *E_col, int32_t nb, double lnR0, double lnf,
                                   double E_const, double E_max, int64_t *no_cols, int32_t solute, uint64_t seed,
                                   uint32_t step, const double *rnd_stream, int64_t rnd_len, void *stream) {
  int rc = ref_args(n_sips, xis, m_w, m_s, vel, grid_T, cells, nx, nz, E_col, nb, no_cols, solute, rnd_stream,
                    rnd_len);
  if (rc) return rc;
  if (n_sips > LM_NOALLOC_MAX_SIPS || (long long)nx * nz > LM_NOALLOC_MAX_CELLS) {
    char msg[200];
    snprintf(msg, sizeof msg,
             "lm_collision_all_cells_noalloc: %lld SIPs on %lld cells exceed its fixed scratch (at most %d SIPs and "
             "%d cells); use lm_collision_all_cells",
             (long long)n_sips, (long long)nx * nz, LM_NOALLOC_MAX_SIPS, LM_NOALLOC_MAX_CELLS);
    return set_err(LM_ERR_ARG, msg);
  }
  if (ensure_phys()) return set_err(LM_ERR_ARG, "cannot upload constants to the device");
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) return set_err(LM_ERR_ARG, "device index out of range");
  if (!g_scratch_ptr[dev]) {
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      char msg[160];
      snprintf(msg, sizeof msg, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
      return set_err(e != hipSuccess ? (int)e : LM_ERR_ARG, msg);
    }
    g_scratch_ptr[dev] = (char *)ptr;
  }
  const RefWork w = ref_carve(g_scratch_ptr[dev], n_sips, (long long)nx * nz);
  g_ref_err[dev][1] = w.small + 2;
  return ref_step(w, n_sips, xis, m_w, m_s, vel, grid_T, cells, nx, nz, dt_over_dV, E_col, nb, lnR0, lnf, E_const,
                  E_max, no_cols, solute, seed, step, rnd_stream, rnd_len, S(stream));
}

int lm_collision_all_cells_check(int32_t no_alloc, void *stream) {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64 || no_alloc < 0 || no_alloc > 1) return set_err(LM_ERR_ARG, "bad check arguments");
  int32_t *e = g_ref_err[dev][no_alloc];
  if (!e) return 0;
  int32_t h[4];
  HIP_TRY(hipMemcpyAsync(h, e, sizeof(h), hipMemcpyDeviceToHost, S(stream)));
  HIP_TRY(hipStreamSynchronize(S(stream)));
  if (!h[0]) return 0;
  hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, S(stream), e, 0, 4);
  HIP_TRY(hipStreamSynchronize(S(stream)));
  char msg[200];
  if (h[0] == LM_ERR_CELL_CAP)
    snprintf(msg, sizeof msg,
             "cell %d holds more SIPs than the all-cells collision step takes (1024 without allocation, else "
             "LM_BIG_CELL_MAX)", h[1]);
  else if (h[0] == LM_ERR_RND_SHORT)
    snprintf(msg, sizeof msg, "the injected uniform stream is shorter than the step's pairs");
  else
    snprintf(msg, sizeof msg, "all-cells collision step error %d (detail %d)", h[0], h[1]);
  return set_err(h[0], msg);
}

int lm_box_minmax(int32_t n_times, int64_t n_tot, const double *xis, const double *masses, double *out,
                  void *stream) {
  if (n_times <= 0 || n_tot <= 0 || !xis || !masses || !out) return set_err(LM_ERR_ARG, "bad box min/max arguments");
  hipLaunchKernelGGL(k_box_minmax, dim3(n_times), dim3(256), 0, S(stream), (long long)n_tot, xis, masses, out);
  LAUNCH_CHECK();
  return 0;
}

int lm_box_stats(int32_t n_times, int32_t n_sims, const int64_t *off, int64_t n_tot, const double *xis,
                 const double *masses, const double *edges, const double *width_m, const double *width_lnr,
                 int32_t no_bins, double dV, double *moments, double *f_m, double *g_m, double *g_ln_r,
                 void *stream) {
  if (n_times <= 0 || n_sims <= 0 || n_tot <= 0 || !off || !xis || !masses || !edges || !width_m || !width_lnr ||
      !moments || !f_m || !g_m || !g_ln_r)
    return set_err(LM_ERR_ARG, "bad box statistics arguments");
  if (no_bins < 1 || no_bins > BOXSTAT_BINS) return set_err(LM_ERR_ARG, "no_bins must be in 1..64");
  if ((int64_t)n_times * n_sims > 0x7fffffff) return set_err(LM_ERR_ARG, "too many (time, seed) pairs");
  hipLaunchKernelGGL(k_box_stats, dim3((unsigned)(n_times * n_sims)), dim3(WAVE), 0, S(stream), n_times, n_sims,
                     (const long long *)off, (long long)n_tot, xis, masses, edges, width_m, width_lnr, no_bins, dV,
                     moments, f_m, g_m, g_ln_r);
  LAUNCH_CHECK();
  return 0;
}

int lm_mean_sem(const double *fields, int32_t n_sims, int64_t n_vals, double *mean, double *sem, void *stream) {
  if (!fields || !mean || !sem || n_sims <= 0 || n_vals <= 0) return set_err(LM_ERR_ARG, "bad mean/sem arguments");
  hipLaunchKernelGGL(k_mean_sem, dim3((unsigned)((n_vals + 255) / 256)), dim3(256), 0, S(stream), fields, n_sims,
                     (long long)n_vals, mean, sem);
  LAUNCH_CHECK();
  return 0;
}

int lm_divergence_upwind(const double *fields, const double *flux, double *out, int32_t n_fields, int32_t nx,
                         int32_t nz, double dx, double dz, int32_t bc_x, int32_t bc_z, void *stream) {
  if (!fields || !flux || !out || n_fields <= 0 || nx < 2 || nz < 2 || bc_x < 0 || bc_x > 1 || bc_z < 0 ||
      bc_z > 1)
    return set_err(LM_ERR_ARG, "bad divergence arguments (boundary types are 0 = periodic, 1 = fixed)");
  if (ensure_phys()) return set_err(LM_ERR_ARG, "cannot upload constants to the device");
  const long long n = (long long)n_fields * nx * nz;
  hipLaunchKernelGGL(k_divergence, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, S(stream), fields, flux, out,
                     n_fields, nx, nz, dx, dz, bc_x, bc_z);
  LAUNCH_CHECK();
  return 0;
}

/* ------------------------------------------------------------------------ */
/* adjoint window                                                           */
/* ------------------------------------------------------------------------ */
int lm_build_slot_of(LmEngine *e, void *stream) {
  int rc = check_engine(e);
  if (rc) return rc;
  if (!e->slot_of || !e->run_base) return set_err(LM_ERR_ARG, "slot_of / run_base missing");
  const long long n = e->n_tot;
  if (n > 0)
    hipLaunchKernelGGL(k_slot_of, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, S(stream), e->id[e->cur],
                       (const long long *)e->run_base, e->n_runs, n, e->slot_of);
  LAUNCH_CHECK();
  return 0;
}

int lm_objective(LmEngine *e, double nc_width, double *out, void *stream) {
  int rc = check_engine(e);
  if (rc) return rc;
  if (!out || !(nc_width > 0.0)) return set_err(LM_ERR_ARG, "objective: out NULL or width <= 0");
  hipLaunchKernelGGL(k_objective, dim3(e->n_runs * e->nc), dim3(WAVE), 0, S(stream), make_view(e), nc_width, out);
  LAUNCH_CHECK();
  return 0;
}

static int check_adjoint(const LmEngine *e, const LmAdjoint *a) {
  if (!a || !a->lam_mw || !a->lam_ms || !a->lam_xi || !a->lam_cell || !a->sl_mw || !a->sl_ms || !a->sl_xi ||
      !a->sl_of)
    return set_err(LM_ERR_ARG, "adjoint buffers missing");
  if (!e->run_base) return set_err(LM_ERR_ARG, "run_base missing");
  return 0;
}

int lm_adj_seed(LmEngine *e, LmAdjoint *a, double w_nc, double w_lwp, double nc_width, void *stream) {
  int rc = check_engine(e);
  if (rc || (rc = check_adjoint(e, a))) return rc;
  if (!(nc_width > 0.0)) return set_err(LM_ERR_ARG, "nc_width must be positive");
  const long long n = e->n_tot;
  const int ncell = e->n_runs * e->nc;
  HIP_TRY(hipMemsetAsync(a->lam_mw, 0, sizeof(double) * (size_t)n, S(stream)));
  HIP_TRY(hipMemsetAsync(a->lam_ms, 0, sizeof(double) * (size_t)n, S(stream)));
  HIP_TRY(hipMemsetAsync(a->lam_xi, 0, sizeof(double) * (size_t)n, S(stream)));
  HIP_TRY(hipMemsetAsync(a->lam_cell, 0, sizeof(double) * (size_t)LA_N * ncell, S(stream)));
  if (n > 0)
    hipLaunchKernelGGL(k_adj_seed, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, S(stream), make_view(e), w_nc,
                       w_lwp, nc_width, *a);
  LAUNCH_CHECK();
  return 0;
}

/* the operations of one advection step in the order lm_adv_steps runs them */
enum { OP_INC1 = 0, OP_INC2, OP_COL, OP_SUB };
static int step_ops(const LmEngine *e, int *ops, int cap) {
  int n = 0;
  auto push = [&](int op) {
    if (n < cap) ops[n] = op;
    ++n;
  };
  auto subloop = [&](int no_cond, int no_col) {
    if (e->act_collisions && no_col == 1) push(OP_COL);
    for (int q = 0; q < no_cond; ++q) {
      if (e->act_collisions && no_col > 1) push(OP_COL);
      push(OP_SUB);
    }
    if (e->act_collisions && no_col > no_cond) push(OP_COL);
  };
  push(OP_INC1);
  subloop(e->scale_dt_cond, e->no_col_per_adv == 1 ? 0 : (e->no_col_per_adv == 2 ? 1 : e->scale_dt_cond));
  push(OP_INC2);
  subloop(e->scale_dt_cond - 1, (e->no_col_per_adv == 1 || e->no_col_per_adv == 2) ? 1 : e->scale_dt_cond);
  push(OP_SUB);
  return n;
}

int lm_adj_adv_steps(LmEngine *e, LmAdjoint *a, int32_t n_steps, int32_t col_index0, void *stream) {
  int rc = check_engine(e);
  if (rc || (rc = check_adjoint(e, a))) return rc;
  if (!e->tape_part || !e->tape_part_i || !e->tape_grid || !e->tape_fields)
    return set_err(LM_ERR_ARG, "reverse sweep needs the segment's tape");
  if (e->act_collisions && (!e->ev || !e->ev_off || !e->ev_order))
    return set_err(LM_ERR_ARG, "reverse sweep needs the window's event log");
  int ops[256];
  const int nops = step_ops(e, ops, 256);
  if (nops > 256) return set_err(LM_ERR_ARG, "too many operations per advection step");
  int nsub = 0, ncol = 0;
  for (int q = 0; q < nops; ++q) {
    nsub += ops[q] == OP_SUB;
    ncol += ops[q] == OP_COL;
  }
  if (n_steps < 0 || (long long)n_steps * nsub > e->tape_nsub || 2LL * n_steps > e->tape_ninc)
    return set_err(LM_ERR_ARG, "reverse sweep longer than the tape");
  if (ncol && (col_index0 < 0 || col_index0 + (long long)n_steps * ncol > e->ev_nstep))
    return set_err(LM_ERR_ARG, "reverse sweep outside the window's event log");
  const long long n = e->n_tot;
  const int ncell = e->n_runs * e->nc;
  const KView v = make_view(e);
  const int nfaces = e->n_runs * ((e->nx + 1) * e->nz + e->nx * (e->nz - 1));
  const unsigned gn = (unsigned)((n + 255) / 256);
  for (int s = n_steps - 1; s >= 0; --s) {
    int sub = (s + 1) * nsub, col = col_index0 + (s + 1) * ncol, inc = 2 * (s + 1);
    /* the step's slot order (global id of each slot, the same in all its substeps) */
    const int32_t *gid_of_slot = e->tape_part_i + (size_t)(s * nsub) * 2 * (size_t)n;
    if (n > 0) hipLaunchKernelGGL(k_adj_to_slots, dim3(gn), dim3(256), 0, S(stream), n, gid_of_slot, *a);
    for (int q = nops - 1; q >= 0; --q) {
      const int op = ops[q];
      if (op == OP_SUB) {
        --sub;
        hipLaunchKernelGGL(k_adj_grid, dim3((ncell + 255) / 256), dim3(256), 0, S(stream), ncell,
                           e->tape_grid + (size_t)sub * 30 * ncell, a->lam_cell);
        if (n > 0)
          hipLaunchKernelGGL(k_adj_part, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, S(stream), n, ncell,
                             e->tape_part + (size_t)sub * 7 * (size_t)n, e->tape_part_i + (size_t)sub * 2 * (size_t)n,
                             *a);
      } else if (op == OP_COL) {
        --col;
        hipLaunchKernelGGL(k_adj_events, dim3((ncell + 255) / 256), dim3(256), 0, S(stream), e->ev,
                           e->ev_off + (size_t)col * (ncell + 1), e->ev_order, ncell, *a);
      } else {
        --inc;
        const int second = op == OP_INC2;
        const double mult = second ? 2.0 : 1.0;
        hipLaunchKernelGGL(k_adj_inc_faces, dim3((nfaces + 255) / 256), dim3(256), 0, S(stream), v,
                           e->tape_fields + (size_t)inc * 2 * ncell, mult, e->dt_sub, a->lam_cell);
        const int nt = e->n_runs * e->nz;
        hipLaunchKernelGGL(k_adj_inc_rows, dim3(nt), dim3(WAVE), 0, S(stream), v, mult, e->dt_sub,
                           e->act_relaxation, second, a->lam_cell);
      }
    }
    if (n > 0) hipLaunchKernelGGL(k_adj_to_gids, dim3(gn), dim3(256), 0, S(stream), n, gid_of_slot, *a);
  }
  LAUNCH_CHECK();
  return 0;
}

/* ------------------------------------------------------------------------ */
/* initial state                                                            */
/* ------------------------------------------------------------------------ */
int64_t lm_init_par_sizeof(void) { return (int64_t)sizeof(LmInitPar); }

static int check_init(const LmInitPar *q) {
  if (!q || q->n_modes < 1 || q->n_modes > LM_INIT_MAX_MODES || q->n_runs < 1 || q->nx < 1 ||
      q->nx > SATADJ_MAX_NX || q->solute < 0 || q->solute > 1 || q->iter_cnt_limit < 0 ||
      q->iter_cnt_limit > 10000000 || q->newton_iterations < 1 || q->newton_iterations > 100)
    return set_err(LM_ERR_ARG, "bad initialisation parameters");
  for (int m = 0; m < q->n_modes; ++m)
    if (q->l_max[m] < 1 || q->l_max[m] > 4096 || !(q->kappa[m] > 0.0))
      return set_err(LM_ERR_ARG, "bad SinSIP bins");
  if (ensure_phys()) return set_err(LM_ERR_ARG, "cannot upload constants to the device");
  return 0;
}

int lm_init_level_profile(const LmInitPar *q, const uint64_t *seeds, int32_t level, double z_bot, double z_top,
                          const double *pbot, double *L, int32_t *count, void *stream) {
  int rc = check_init(q);
  if (rc) return rc;
  if (!seeds || !pbot || !L || !count) return set_err(LM_ERR_ARG, "initialisation buffers missing");
  if (level < 0) return set_err(LM_ERR_ARG, "negative level");
  hipLaunchKernelGGL(k_init_profile, dim3((q->n_runs + 63) / 64), dim3(64), 0, S(stream), *q, z_bot, z_top, pbot, L);
  const int nt = q->n_runs * q->nx * q->n_modes;
  hipLaunchKernelGGL(k_sinsip, dim3((nt + 127) / 128), dim3(128), 0, S(stream), *q, seeds, level, L,
                     (const long long *)nullptr, count, nullptr, nullptr, nullptr, nullptr, nullptr);
  LAUNCH_CHECK();
  return 0;
}

int lm_init_level_sips(const LmInitPar *q, const uint64_t *seeds, int32_t level, double *L, const int64_t *off,
                       const int64_t *cell_off, int64_t n_level, double *m_s, double *xi, int32_t *mode,
                       int32_t *cell, int32_t *run, double *m_w, double *m_p, double *dmxi, double *pbot_next,
                       double *fields, void *stream) {
  int rc = check_init(q);
  if (rc) return rc;
  if (!seeds || !L || !off || !cell_off || !pbot_next || !fields ||
      (n_level > 0 && (!m_s || !xi || !mode || !cell || !run || !m_w || !m_p || !dmxi)))
    return set_err(LM_ERR_ARG, "initialisation buffers missing");
  if (level < 0 || n_level < 0) return set_err(LM_ERR_ARG, "negative level or SIP count");
  const int nt = q->n_runs * q->nx * q->n_modes;
  hipLaunchKernelGGL(k_sinsip, dim3((nt + 127) / 128), dim3(128), 0, S(stream), *q, seeds, level, L,
                     (const long long *)off, nullptr, m_s, xi, mode, cell, run);
  if (n_level > 0) {
    const dim3 g((unsigned)((n_level + 127) / 128));
    if (q->solute == 0)
      hipLaunchKernelGGL(k_init_mass<0>, g, dim3(128), 0, S(stream), (long long)n_level, run, L, m_s, m_w, m_p);
    else
      hipLaunchKernelGGL(k_init_mass<1>, g, dim3(128), 0, S(stream), (long long)n_level, run, L, m_s, m_w, m_p);
  }
  if (q->solute == 0)
    hipLaunchKernelGGL(k_satadj<0>, dim3(q->n_runs), dim3(SATADJ_BLOCK), 0, S(stream), *q,
                       (const long long *)cell_off, cell, m_s, xi, m_p, m_w, dmxi, L, pbot_next, fields);
  else
    hipLaunchKernelGGL(k_satadj<1>, dim3(q->n_runs), dim3(SATADJ_BLOCK), 0, S(stream), *q,
                       (const long long *)cell_off, cell, m_s, xi, m_p, m_w, dmxi, L, pbot_next, fields);
  LAUNCH_CHECK();
  return 0;
}

int lm_init_positions(int64_t n, const uint64_t *seeds, const int32_t *run, const int32_t *idx, const int32_t *ci,
                      const int32_t *cj, const double *corner_x, const double *corner_z, double dx, double dz,
                      double *x, double *z, double *rel, void *stream) {
  if (n <= 0) return 0;
  if (!seeds || !run || !idx || !ci || !cj || !corner_x || !corner_z || !x || !z || !rel)
    return set_err(LM_ERR_ARG, "position buffers missing");
  hipLaunchKernelGGL(k_init_positions, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, S(stream), (long long)n,
                     seeds, run, idx, ci, cj, corner_x, corner_z, dx, dz, x, z, rel);
  LAUNCH_CHECK();
  return 0;
}

} /* extern "C" */
