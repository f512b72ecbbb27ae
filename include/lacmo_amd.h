/*
 * lacmo_amd.h -- C ABI of the gfx950 super-droplet (SIP) time step.
 *
 * Plain C: device pointers, sizes and a hipStream_t passed as void*.  The
 * Python mirror of the reference interface (lacmopy_amd.integration,
 * lacmopy_amd.collision.all_or_nothing) fills LmEngine with pointers to
 * device buffers it allocates and calls these entry points through ctypes.
 * Every entry point returns 0 on success or a nonzero LM_ERR_* / hipError_t
 * code; lm_last_error_message() describes the last failure.  Calls are
 * stream-ordered and make no host synchronisation of their own (except
 * lm_engine_check, which copies the device error word back).
 *
 * Engine layout: SIPs of all runs in one structure-of-arrays, sorted by
 * (run, home cell, SIP id) at the last reindex; seg[] holds, per run, the
 * nc cell segments followed by one "graveyard" segment of inactive SIPs.
 * Cells are linear i*nz + j (x-major, like the reference's loops); the
 * current cell of a SIP is the global cell r*nc + c, or -(2 + i) when the SIP
 * is inactive (i = the x index the reference keeps in cells[0]).
 */
#ifndef LACMO_AMD_H
#define LACMO_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LM_ERR_OK 0
#define LM_ERR_ARG 1001        /* invalid argument / shape */
#define LM_ERR_CELL_CAP 1002   /* a cell holds more SIPs than the kernels' LDS capacity */
#define LM_ERR_LOST_SIP 1003   /* a SIP moved more than one cell between reindexes */
#define LM_ERR_REM_CAP 1004    /* too many SIPs reached the ground in one substep */
#define LM_ERR_OUT_OF_DOMAIN 1005 /* a SIP landed on x == x_max (cell index nx) */
#define LM_ERR_RND_SHORT 1006  /* the injected uniform stream ran out */
#define LM_ERR_EV_CAP 1007     /* the collision-event log of an adjoint window is full */

/* SIPs that may reach the ground per run in one particle substep (LmEngine.rem_cap is 1..this; more raise
   LM_ERR_REM_CAP): water_removed sums them in ascending id in LDS */
#define LM_REM_CAP_MAX 1024

/*
 * One collision event of an adjoint window, recorded by lm_collision_step
 * (ev_mode 1) and replayed as a fixed event (ev_mode 2): the pair, the branch
 * of all_or_nothing.py:577-632 it took, p_crit of a multiple collection, and
 * the state of the pair before the event.
 */
typedef struct LmEvent {
  int32_t gid_min, gid_max; /* global SIP ids of the smaller- and the larger-multiplicity SIP */
  int32_t cell, seq;        /* global cell, event number within the cell's collision step */
  int32_t kind, pad;        /* 1 multiple, 2 single, 3 near-equal single collection */
  double p;                 /* p_crit of a multiple collection (kind 1) */
  double xi_min, xi_max, mw_min, mw_max, ms_min, ms_max;
} LmEvent;

typedef struct LmEngine {
  /* geometry */
  int32_t n_runs, nx, nz, nc;
  int64_t n_tot;
  double x_min, x_max, z_min, z_max, dx, dz, V_cell, p_ref, p_ref_inv;
  /* physics parameters (config keys of integration.py:1908-1923) */
  int32_t solute;         /* 0 = 'AS', 1 = 'NaCl' */
  int32_t no_iter;        /* no_iter_impl_mass */
  int32_t scale_dt_cond;  /* no_cond_per_adv // 2 */
  int32_t no_col_per_adv;
  int32_t act_collisions, act_relaxation;
  double dt, dt_sub, dt_sub_half, dt_col_over_dV, g_set;
  /* collection efficiency: table E[nb*nb] (Ecol_grid_R) or constant */
  const double *E;
  int32_t nb, e_const_mode;
  double E_const, lnR0, lnf;
  uint64_t seed;
  uint32_t col_step; /* host mirror of *col_step_dev, the collision-step counter (Philox counter word) */
  int32_t cur;       /* current ping-pong buffer of the SIP state and seg */
  int32_t ccur;      /* current ping-pong buffer of cell[] */
  int32_t pad1;
  /* SIP state, ping-pong buffers [2] */
  double *x[2], *z[2], *u[2], *w[2], *m_w[2], *m_s[2], *xi[2];
  int32_t *id[2];
  int32_t *cell[2]; /* cell[ccur] = current cell; cell[1-ccur] = next */
  int32_t *seg[2];  /* n_runs*(nc+1)+1 segment starts */
  double *dmxi;     /* n_tot scratch: dm * xi per SIP */
  uint32_t *col_step_dev; /* 1: the collision-step counter the kernels read (set with lm_set_col_step) */
  int32_t *away;    /* n_runs*nc: per home cell, 9-bit mask of the neighbour directions its SIPs moved to
                       since the last reindex (zero after a reindex or a load) */
  /* grid, per run */
  double *scal; /* n_runs*10*nc: T, p, Theta, rho_d, r_v, r_l, S, e_s, m_dry_inv, rho_d_inv */
  double *mat;  /* n_runs*7*nc: K, D_v, L_v, sigma_w, c_p, mu, rho_f */
  double *cc;   /* n_runs*nc*16 per-cell constants for the particle kernel */
  double *dTh, *dRv; /* n_runs*nc advective increments of Theta and r_v */
  double *dml;       /* n_runs*nc Delta m_l of the last substep (kg) */
  const double *grid_vel; /* 2*(nx+1)*(nz+1) face velocities, per run at vel_run_stride */
  const double *flux;     /* 2*(nx+1)*(nz+1) dry-air mass flux */
  const double *prof_rv, *prof_Th; /* n_runs*nz initial profiles (relaxation) */
  const double *t_relax;           /* nz */
  double *colsum;                  /* n_runs*nz*2 scratch */
  /* per-run accumulators */
  int64_t *no_cols;      /* n_runs*2 */
  double *water_removed; /* n_runs */
  int64_t *n_equal_xi;   /* n_runs: near-equal-xi collision events (the reference prints them) */
  int32_t *rem_cnt;      /* n_runs */
  int32_t *rem_slot;     /* n_runs*rem_cap */
  int32_t rem_cap, pad0; /* removal slots per run and substep, 1..LM_REM_CAP_MAX */
  /* scratch */
  int32_t *scan_buf;  /* n_runs*(2nc+1)+1 */
  int32_t *scan_part; /* scan partials */
  int32_t *err;       /* 4 error words on the device */
  int32_t *ovf_list;  /* 3*n_runs*nc (2*n_runs*nc without big_ws): cells over the fast-path LDS capacity
                         (collision: list A in the first third, its cells over 256 members in the second, cells
                         over 1024 members in the third) */
  int32_t *ovf_cnt;   /* 288: cells queued for the large-capacity collision [0] and Delta m_l [1] passes, for
                         the collision's 1024-member pass after its 256-member pass [2] and for its pass over
                         cells of any size [3], for the Delta m_l 1024-member pass after its 256-member pass
                         [4]; [32..287]: the collision's 256-member pass's claim counters, one per 128 B */
  /* injected uniform stream (parity mode; NULL = Philox) */
  const double *rnd_stream;
  int64_t *rnd_off; /* n_runs*nc+1 scratch for stream offsets */
  int64_t rnd_len;  /* number of values in rnd_stream */
  /* optional kernel timing (lm_prof_*); NULL = off */
  void *prof;
  /*
   * Adjoint window (all NULL/0 outside one).  ev_mode 1: lm_collision_step
   * records every event into ev; ev_mode 2: it applies the recorded events of
   * the collision step instead of drawing (fixed events).  The host counters
   * ev_step, tape_sub and tape_inc advance with each collision step, substep
   * and increment evaluation.
   */
  int32_t ev_mode, ev_cap;
  int32_t ev_step, ev_nstep;   /* host: collision step within the window (selects ev_off), steps ev_off holds */
  int32_t tape_sub, tape_nsub; /* host: particle/grid substep within the tape, substeps the tape holds */
  int32_t tape_inc, tape_ninc; /* host: increment evaluation within the tape, evaluations it holds */
  LmEvent *ev;              /* ev_cap records */
  int32_t *ev_cnt;          /* 1: records written so far */
  int32_t *ev_count;        /* n_runs*nc scratch: events per cell of the current collision step */
  int32_t *ev_off;          /* per collision step of the window: n_runs*nc+1 offsets into ev_order */
  int32_t *ev_order;        /* ev_cap: record indices ordered by (collision step, cell, seq) */
  const int64_t *run_base;  /* n_runs+1: global id of SIP 0 of each run */
  int32_t *slot_of;         /* n_tot: slot of each global id in the current buffer (ev_mode 2) */
  /* tape of an adjoint segment (NULL: no tape is written) */
  double *tape_part;    /* per substep 7*n_tot, slot order: d m_w'/d(m_w, m_s, T, p, S), dm, xi */
  int32_t *tape_part_i; /* per substep 2*n_tot: global id, cell of the condensation update (-1: none) */
  double *tape_grid;    /* per substep 30*n_runs*nc: d(Theta', r_v', T', p', S')/d(Theta, r_v, T, dml,
                           dTheta_adv, dr_v_adv) of the grid substep */
  double *tape_fields;  /* per increment evaluation 2*n_runs*nc: r_v, Theta */
  double *ccd;          /* n_runs*nc*12 scratch: d(l_alpha, l_beta, c1, c2, rho_w, sigma_w)/d(T, p) */
  /*
   * Launch shape (0 = default; set by an autotuner, e.g. lacmopy_amd.Engine.autotune): threads per
   * block (64..512, a multiple of 64) and SIPs per thread (1 or 2) of the particle kernel, and the
   * member capacity of the main collision pass (160, 256, 320 or 512; larger cells take the overflow pass).
   */
  int32_t part_block, part_spt, col_cap;
  int32_t tape_mode; /* taped substep of the adjoint: 0 = two passes, (T, p, S) then (m_w, m_s); 1 = one pass */
  /* largest entry of the E_col table (0: unknown); lets the collision kernel reject most pairs before it
     reads their table entry */
  double E_max;
  /* elements between the face-velocity fields of consecutive runs in grid_vel: 0 (one field for all
     runs) or 2*(nx+1)*(nz+1) (one per run: runs initialised apart have their own dry-density profile,
     and the velocity is the mass flux over it) */
  int64_t vel_run_stride;
  /*
   * Persistent per-cell passes (0 = one workgroup per (run, cell); k = 1..256 workgroups per compute unit
   * that pull (run, cell) items from the work list until it is empty), for the main collision pass and
   * the main Delta m_l pass, each taking *_work_chunk (0..4096) consecutive items per atomic grab (0: a
   * static round robin over the list, no atomics); set by an autotuner.  The work list holds the n_runs*nc
   * global cells ordered by decreasing home-segment size (largest first), rebuilt by every reindex while a
   * persistent pass is on; any permutation of the cells gives the same results, only the balance changes.
   */
  int32_t col_work_per_cu, dml_work_per_cu;
  int32_t *work; /* n_runs*nc + LM_WORK_BINS + 6: work list, histogram scratch, 2 work counters, the int64 cost
                    sum of the list build at the next 8-byte boundary (NULL: off) */
  int32_t col_work_chunk, dml_work_chunk;
  /* particle substep compiled for this many waves per SIMD (0: the default build; 5, 6 or 8: a smaller
     register budget), one SIP per thread; set by an autotuner */
  int32_t part_wpe, pad4;
  /*
   * Migrant lists (NULL: off): per (run, cell) the slots of the SIPs that entered the cell from another home
   * segment since the last reindex, appended by the position updates (mig_cnt counts every entry, also
   * past mig_cap; an entry may be stale or repeated -- readers keep the slots whose cell is the cell, once).
   * Per-cell passes take a cell's members as its home segment plus this list while the list is complete
   * (mig_cnt <= mig_cap), else from the 3x3 home segments the away mask names.  Cleared by every reindex.
   */
  int32_t *mig_cnt;  /* n_runs*nc */
  int32_t *mig_slot; /* n_runs*nc*mig_cap */
  int32_t mig_cap; /* 1..LM_MIG_CAP_MAX */
  /* events a cell's collision step takes one at a time through its candidate queue before the anti-diagonal
     schedule finishes it: 0 = the default (48), else 1..LM_SPEC_EVENTS_MAX; results do not depend on it */
  int32_t spec_events;
  /* n_runs*nc or NULL: wall-clock ticks (100 MHz) each cell's last collision step took, written by the
     collision passes; with it the work list starts with the cells whose cost exceeds half the expected
     length of a persistent collision pass (their cost order), then the others by home-segment size */
  uint32_t *col_cost;
  /*
   * Collision pass over cells of more than 1024 members (NULL: such a cell is an error, LM_ERR_CELL_CAP):
   * big_blocks workgroups, each with its own region of lm_big_ws_bytes(big_cap) bytes of big_ws holding the
   * per-member state of one cell at a time; cells of up to big_cap (2..LM_BIG_CELL_MAX) members.
   */
  void *big_ws;
  int32_t big_cap, big_blocks;
} LmEngine;

#define LM_MIG_CAP_MAX 64       /* largest migrant list per cell */
#define LM_SPEC_EVENTS_MAX 1024 /* largest spec_events */
#define LM_BIG_CELL_MAX 46340   /* members of a cell the collision passes take at most (32-bit pair numbers) */

#define LM_WORK_BINS 1024 /* home-segment sizes are binned up to LM_WORK_BINS - 1 for the work-list order */

/*
 * Parameters of the initial SIP ensemble and saturation adjustment (config.py:84-132 as
 * init.py:43-206 derives them).  Up to LM_INIT_MAX_MODES lognormal dry-mass modes.
 */
#define LM_INIT_MAX_MODES 4
typedef struct LmInitPar {
  int32_t n_modes, solute, weak_threshold, newton_iterations;
  int32_t iter_cnt_limit, nx, n_runs, pad;
  int32_t l_max[LM_INIT_MAX_MODES];     /* bins per mode: ceil(kappa log10(m_high_over_m_low)) */
  double mu_m_log[LM_INIT_MAX_MODES];   /* ln of the geometric mean dry mass (1e-18 kg) */
  double sigma_m_log[LM_INIT_MAX_MODES];
  double kappa[LM_INIT_MAX_MODES];      /* SinSIP bins per decade */
  double r_critmin[LM_INIT_MAX_MODES];  /* smallest dry radius (um) */
  double no_rpcm_0[LM_INIT_MAX_MODES];  /* ceil(V_cell DNC0) */
  double eta, mass_density_dry, S_init_max, dt_init;
  double p_ref, Theta_l, r_tot_0, rho_dry_0;
  double X, dx, dy, dz;
} LmInitPar;

/* sizeof(LmInitPar), for bindings that mirror the struct */
int64_t lm_init_par_sizeof(void);

/*
 * Per-run level state of the initialisation, LM_INIT_LV_N doubles: p_bot, p_top, T_bot, T_top,
 * T_avg, p_avg, rho_dry_avg, S_avg, r_v_avg, r_l_avg, dry-air mass, vapour mass, liquid mass,
 * particle mass (kg, level), rho_dry scale of the particle counts, adjustment iterations, last
 * relative liquid change, then the real particles per cell of each mode.
 */
#define LM_INIT_LV_N 21

/*
 * One level of the initial state (init.py:697-740, generation_SIP_ensemble.py:330-404): every run's
 * dry profile from its bottom pressure pbot[run] (device), the level state L (device, n_runs x
 * LM_INIT_LV_N) and the number of SinSIPs of each (run, cell, mode) into count (device,
 * n_runs*nx*n_modes).  Uniform draws: Philox-4x32-10 keyed by each run's seed seeds[run] (device,
 * the reference's seed_SIP_gen of that run) instead of NumPy's generator.
 */
int lm_init_level_profile(const LmInitPar *q, const uint64_t *seeds, int32_t level, double z_bot, double z_top,
                          const double *pbot, double *L, int32_t *count, void *stream);

/*
 * The level's SIPs and its saturation adjustment (init.py:741-1010): SinSIPs written at off[(run, cell,
 * mode)] (device int64) into m_s, xi, mode, cell, run; the initial water mass in equilibrium with the
 * level's S and T (microphysics.py:686-966) into m_w (m_p: scratch); then one workgroup per run repeats
 * implicit Newton condensation steps (dt_init) of the level's SIPs with the level thermodynamics
 * updated after each until the liquid water changes by less than 1e-5 relative or iter_cnt_limit
 * steps.  cell_off (device int64, n_runs*nx+1): first SIP of each (run, cell).  Writes L, the next
 * level's bottom pressures pbot_next and per (run, cell) r_v, r_l, e_s, S into fields (n_runs x 4 x nx).
 */
int lm_init_level_sips(const LmInitPar *q, const uint64_t *seeds, int32_t level, double *L, const int64_t *off,
                       const int64_t *cell_off, int64_t n_level, double *m_s, double *xi, int32_t *mode,
                       int32_t *cell, int32_t *run, double *m_w, double *m_p, double *dmxi, double *pbot_next,
                       double *fields, void *stream);

/*
 * Positions uniform in each SIP's cell (init.py:322-381): x = corner_x[ci] + dx u, z = corner_z[cj] + dz u',
 * (u, u') from seeds[run[k]] and idx[k] (the SIP's index within its run) into rel (2 x n, device).
 */
int lm_init_positions(int64_t n, const uint64_t *seeds, const int32_t *run, const int32_t *idx, const int32_t *ci,
                      const int32_t *cj, const double *corner_x, const double *corner_z, double dx, double dz,
                      double *x, double *z, double *rel, void *stream);

/*
 * Adjoint variables of a window (lm_adj_*): per global SIP id and per cell,
 * plus the scratch the reverse sweep keeps them in while it reverses one
 * advection step (the step's slot order, and the slot of each global id).
 */
typedef struct LmAdjoint {
  double *lam_mw, *lam_ms, *lam_xi; /* n_tot each, by global id */
  double *lam_cell; /* 8*n_runs*nc: Theta, r_v, T, p, S, dml, dTheta_adv, dr_v_adv */
  double *sl_mw, *sl_ms, *sl_xi;    /* n_tot each, by slot of the step being reversed */
  int32_t *sl_of;                   /* n_tot: slot of each global id in that step */
} LmAdjoint;

/* Library identity: returns a static version string. */
const char *lm_version(void);
/* Message of the last error on this thread. */
const char *lm_last_error_message(void);
/* sizeof(LmEngine) and byte offsets of selected fields (for binding checks). */
int64_t lm_engine_sizeof(void);
/*
 * One run's SIPs between the reference's layout on the device and the engine's: the device form of the
 * drop-in boundary of integrate_adv_step_np and the subloop functions (integration.py:944-1294, 1298-1611),
 * whose arrays are pos, vel (2, n) float64, cells (2, n) int64, m_w, m_s, xi (n) float64 and active_ids (n)
 * bytes (0/1) in SIP-index order.  run occupies the engine's slots [base, base + n).  Load: the engine's
 * order (active SIPs by cell, then index; inactive SIPs after them by index), its segment table, no
 * direction bits or migrants; an active SIP outside the grid is an error (LM_ERR_ARG through
 * lm_engine_check).  Store: the arrays scattered back by index, cells from the engine's cell codes
 * (inactive: cells[1] = -1), and when not NULL rel_pos (2, n) of the active SIPs as integration.py:545-549
 * computes it and T_p (n), the cell temperature each SIP saw in the last particle substep
 * (integration.py:489-507).  Stream-ordered; the load uses the engine's idle buffers as scratch.
 */
int lm_load_reference_layout(LmEngine *e, int32_t run, int64_t base, int64_t n, const double *pos,
                             const double *vel, const int64_t *cells, const double *m_w, const double *m_s,
                             const double *xi, const uint8_t *active_ids, void *stream);
int lm_store_reference_layout(LmEngine *e, int32_t run, int64_t base, int64_t n, double *pos, double *vel,
                              int64_t *cells, double *rel_pos, double *m_w, double *m_s, double *xi,
                              uint8_t *active_ids, double *T_p, void *stream);
/* Bytes of one workgroup's region of LmEngine.big_ws for cells of up to cap members (no counterpart in the
   reference, which keeps a cell's members in NumPy arrays of any size, all_or_nothing.py:811-846). */
int64_t lm_big_ws_bytes(int32_t cap);
void lm_engine_offsets(int64_t *out, int32_t n);
/* Physical constants as the kernels use them (for the constant-parity test). */
void lm_get_constants(double *out, int32_t n);
int32_t lm_num_constants(void);

/* Builds cc[] from scal/mat (after initialisation or a host-side edit of the grid). */
int lm_engine_init_cells(LmEngine *e, void *stream);

/*
 * One all-cells collision step (collision/all_or_nothing.py:754-846,
 * collision_step_Ecol_grid_R_all_cells_2D_multicomp_np): every cell of every
 * run, pairs in the reference's row-major order over ascending SIP id, kernel
 * E_col(R_i,R_j) from the table or constant, radii at the cell temperature.
 * Draws come from rnd_stream (consumed in x-major cell order) or Philox keyed
 * by (seed, run, col_step, cell, pair).  Increments col_step.
 */
int lm_collision_step(LmEngine *e, void *stream);

/*
 * Position/cell update of all active SIPs by dt (integration.py:552-630 and
 * 509-550: update_pos_from_vel_BC_PS + update_cells_and_rel_pos), with the
 * water_removed bookkeeping in ascending SIP id.
 */
int lm_position_step(LmEngine *e, double dt, void *stream);

/* Full reindex: counting sort of the SIPs by (run, current cell, id). */
int lm_reindex(LmEngine *e, void *stream);

/*
 * (Re)builds the (run, cell) work list of the persistent per-cell passes from the current segments:
 * cells by decreasing home-segment size.  lm_reindex calls it while a persistent pass is on; an
 * autotuner calls it before timing persistent launch shapes.  The reference visits the cells of its
 * all-cells loops in x-major order (collision/all_or_nothing.py:806-815, integration.py:915-938); the
 * per-cell results do not depend on the visiting order, so the list only changes the load balance.
 */
int lm_build_work_list(LmEngine *e, void *stream);

/*
 * Advective increments of r_v and Theta (integration.py:1483-1508 with
 * second=0; 1536-1561 with second=1): Hundsdorfer-limited upwind divergence
 * of the dry-air mass flux (integration.py:150-286) plus relaxation
 * (integration.py:325-354).
 */
int lm_advection_increments(LmEngine *e, int32_t second, void *stream);

/*
 * One particle substep followed by one grid substep
 * (integration.py:944-1089 propagate_particles_subloop_step_np and
 * 402-483 propagate_grid_subloop_step_np): implicit Newton condensation,
 * per-cell Delta m_l summed in ascending id, drag update, position and cell
 * update, then the cell thermodynamics with latent heating.
 */
int lm_substep(LmEngine *e, double dt_sub, double dt_sub_pos, void *stream);

/*
 * The particle half of lm_substep alone (integration.py:944-1089,
 * propagate_particles_subloop_step_np): leaves the per-cell Delta m_l of the
 * substep in dml for a following lm_grid_substep.
 */
int lm_particle_substep(LmEngine *e, double dt_sub, double dt_sub_pos, void *stream);

/*
 * The grid half of lm_substep alone (integration.py:402-483,
 * propagate_grid_subloop_step_np): applies dTh, dRv (advection increments)
 * and dml (condensation) to every cell, then the material properties.
 */
int lm_grid_substep(LmEngine *e, void *stream);

/*
 * grid.py:370-412 update_grid_r_l_np: liquid water mixing ratio of every
 * cell, r_l = 1e-18 * m_dry^-1 * sum(xi * m_w) over the cell's active SIPs in
 * ascending id, written into grid_scalar_fields[5] (scal).
 */
int lm_update_r_l(LmEngine *e, void *stream);

/*
 * Derived fields of the current state of every run (evaluation.py:202-509,
 * generate_field_frame_data_avg, per seed and frame): out[(run*9 + f)*nc + c],
 * f = r_aero, r_c, r_r (mixing ratios of R_p < 0.5, < 25, >= 25 um),
 * n_aero, n_c, n_r (numbers per dry-air mass), R_avg, R_2/1, R_eff (radius
 * moment ratios, the last two of the cloud class).  Sums run over each cell's
 * active SIPs in ascending id, R_p at the cell temperature.
 */
int lm_derived_fields(LmEngine *e, double *out, void *stream);

/*
 * Ensemble mean and spread over seeds (evaluation.py:443-474): fields is
 * (n_seeds, n_vals); mean = sum / n (in seed order), std = sqrt(max(0,
 * sum f^2 - n mean^2) / (n (n-1))) (0 for one seed).
 */
int lm_ensemble_stats(const double *fields, int32_t n_seeds, int64_t n_vals, double *mean, double *std,
                      void *stream);

/* n_steps advection steps (integration.py:1298-1611, integrate_adv_step_np). */
int lm_adv_steps(LmEngine *e, int32_t n_steps, void *stream);

/* Sets the collision-step counter (device value and host mirror). */
int lm_set_col_step(LmEngine *e, uint32_t step, void *stream);

/*
 * lm_adv_steps(e, n_steps) captured into an executable HIP graph on a
 * non-default stream (n_steps even: a step swaps the SIP buffers once, so an
 * even count leaves the buffer indices where the capture found them).  Each
 * lm_graph_launch advances the engine by n_steps advection steps, with the
 * collision-step counter read from device memory.  The graph holds the engine's
 * buffers and parameters as they were at capture time.  No counterpart in the
 * reference (its steps are Python calls).
 */
int lm_graph_capture(LmEngine *e, int32_t n_steps, void *stream, void **graph);
int lm_graph_launch(LmEngine *e, void *graph, void *stream);
int lm_graph_destroy(void *graph);

/*
 * Kernel timing with HIP events on the launch stream (no reference
 * counterpart; used by bench.py).  While enabled, every launch of the
 * particle kernel (kind 0), the collision kernel's main pass (1), the Delta
 * m_l kernel (2), the reindex scatter (3) and the collision kernel's overflow
 * passes (4) is bracketed by events; lm_prof_collect
 * synchronises and returns the summed milliseconds and launch counts per
 * kind, then resets.  At most max_events launches are recorded.
 */
int lm_prof_enable(LmEngine *e, int32_t max_events);
int lm_prof_collect(LmEngine *e, double *ms, int32_t *count, int32_t nkinds);
int lm_prof_disable(LmEngine *e);

/* Synchronises the stream and reports the device error word (0 = fine). */
int lm_engine_check(LmEngine *e, void *stream);

/*
 * Test hook (no reference counterpart): elementwise device math of the
 * kernels -- which = 0 exp, 1 log, 2 pow(x,y), 3 x % y (Python float modulo),
 * 4 sqrt, 5 x / y, 6 floor(x / y), 7 x^(1/3 rounded), 8 and 9 x / y through the shared-reciprocal
 * quotients of the particle kernel (lmp_div2, lm_div_y) -- for the host/device bit-parity test.
 */
int lm_math_probe(int32_t which, int64_t n, const double *x, const double *y, double *out, void *stream);

/*
 * Check of the collision kernels' fast kernel-table index (the bin of all_or_nothing.py:35-65 from the hardware
 * float log2 where the bin is certain, else the exact double evaluation): out_fast[k] / out_exact[k] = the fast and
 * the exact index of R[k] for the table (lnR0, lnf, nb).  If sweep is not NULL (4 doubles): sweep[0] = the largest
 * |hardware log2(x) - log2(x)| over every float x of the fast path's range [2^-14, 2^16), sweep[1] = the bound the
 * fast path assumes for it, sweep[2] = the resulting margin in t, sweep[3] = the floats swept.  Test probe; no
 * counterpart in the reference.
 */
int lm_kernel_index_check(int64_t n, const double *R, double lnR0, double lnf, int32_t nb, int32_t *out_fast,
                          int32_t *out_exact, double *sweep, void *stream);

/*
 * Batched box model: one collision step in each of n_boxes independent boxes
 * (collision/all_or_nothing.py:109-185 collision_step_Golovin_np).  Box b
 * holds SIPs [off[b], off[b+1]) of xis/masses (at most 1024); counters
 * no_cols[2*b..].  Draws: Philox keyed (seed, run = b, step, cell = 0, i, j),
 * or, if rnd_stream is not NULL, its values in row-major pair order, box
 * after box; the caller guarantees it holds sum_b n_b (n_b - 1) / 2 values.
 * A box over capacity is left unchanged and reported by lm_box_check.
 */
int lm_golovin_step(int32_t n_boxes, const int32_t *off, double *xis, double *masses, double dt_over_dV,
                    int64_t *no_cols, uint64_t seed, uint32_t step, const double *rnd_stream, void *stream);

/*
 * Batched box model with a kernel K(m1, m2) tabulated on a logarithmic mass
 * grid (collision/all_or_nothing.py:392-491 collision_step_kernel_grid_m_np;
 * bin index as :35-65 compute_kernel_index of the mass).  Boxes, draws,
 * stream length and capacity as for lm_golovin_step.
 */
int lm_kernel_grid_m_step(int32_t n_boxes, const int32_t *off, double *xis, double *masses, double dt_over_dV,
                          const double *kernel_grid, int32_t nb, double m_kernel_low_log, double bin_factor_m_log,
                          int64_t *no_cols, uint64_t seed, uint32_t step, const double *rnd_stream, void *stream);

/*
 * Batched box model with the E_col(R1,R2) table and 1-D Beard velocities
 * (collision/all_or_nothing.py:269-390 collision_step_Ecol_grid_R_np followed
 * by collision/kernel.py:193-210 update_velocity_Beard when update_beard != 0,
 * as collision/box_model.py:164-176 does after every step).  Draws, stream
 * length and capacity as for lm_golovin_step.
 */
int lm_ecol_box_step(int32_t n_boxes, const int32_t *off, double *xis, double *masses, double *radii, double *vel,
                     const double *rho, double dt_over_dV, const double *E, int32_t nb, double lnR0, double lnf,
                     int64_t *no_cols, uint64_t seed, uint32_t step, const double *rnd_stream, int32_t update_beard,
                     void *stream);

/*
 * One collision step in each of n_boxes boxes of two-component SIPs with
 * caller-supplied radii, densities and 2-D velocities:
 * collision/all_or_nothing.py:498-635 collision_step_Ecol_grid_R_2D_multicomp_np
 * (table E, nb > 0) and :637-751 collision_step_Ecol_const_2D_multicomp_np
 * (E == NULL, constant E_const).  Box b holds SIPs off[b] .. off[b+1]-1 (at
 * most 1024).  xis, m_w, m_s, radii are updated in place; no_cols is
 * (n_boxes, 2).  Draws, stream length and capacity as for lm_golovin_step.
 */
int lm_box2d_step(int32_t n_boxes, const int32_t *off, double *xis, double *m_w, double *m_s, double *radii,
                  const double *vel_x, const double *vel_z, const double *rho, double dt_over_dV, const double *E,
                  double E_const, int32_t nb, double lnR0, double lnf, int64_t *no_cols, uint64_t seed, uint32_t step,
                  const double *rnd_stream, void *stream);

/*
 * Synchronises the stream and reports a box that exceeded the 1024-SIP
 * capacity of the box kernels since the last check (LM_ERR_CELL_CAP; such a
 * box is left unchanged).  No counterpart in the reference, which has no
 * capacity limit.
 */
int lm_box_check(void *stream);

/*
 * One all-cells collision step on the reference's own device arrays, replacing
 * collision_step_Ecol_grid_R_all_cells_2D_multicomp_np (collision/all_or_nothing.py:754-846): xis, m_w, m_s
 * (n_sips, updated in place), vel (2, n_sips), grid_T (nx, nz), cells (2, n_sips) int64; the members of cell
 * (i, j) are the SIPs with cells == (i, j) in ascending index (inactive SIPs, cells[1] = -1, are in no cell).
 * E_col (nb x nb, the Ecol_grid_R table with lnR0 = R_kernel_low_log, lnf = bin_factor_R_log) or NULL for the
 * constant efficiency E_const (Ecol_const); E_max the table's largest entry (0: unknown, the pair filter then
 * evaluates every pair).  no_cols (2) int64 += (single, multiple) collections.  Draws keyed (seed, run 0, step,
 * cell) as lm_collision_step's, or rnd_stream (rnd_len values, consumed in the reference's order).  Cells of
 * up to LM_BIG_CELL_MAX members (those over 1024 in a global workspace).  Stream-ordered, no host
 * synchronisation; the library keeps a device workspace (grown on demand).  Errors of the kernels (a cell over
 * capacity, a short stream) are reported by lm_collision_all_cells_check(0, stream).
 */
int lm_collision_all_cells(int64_t n_sips, double *xis, double *m_w, double *m_s, const double *vel,
                           const double *grid_T, const int64_t *cells, int32_t nx, int32_t nz, double dt_over_dV,
                           const double *E_col, int32_t nb, double lnR0, double lnf, double E_const, double E_max,
                           int64_t *no_cols, int32_t solute, uint64_t seed, uint32_t step, const double *rnd_stream,
                           int64_t rnd_len, void *stream);
/*
 * lm_collision_all_cells without any allocation: no counterpart in the reference; for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops, harnesses that forbid hipMalloc).  Same
 * arguments and results; its scratch (member lists, densities, the per-cell table, counters) is one fixed
 * device array of the library (about 17 MB per GPU, held until the process ends), so it takes at most
 * LM_NOALLOC_MAX_SIPS SIPs on at most LM_NOALLOC_MAX_CELLS cells (larger calls fail with LM_ERR_ARG), cells
 * of at most 1024 members (larger ones: LM_ERR_CELL_CAP; there is no workspace for the any-size pass), and
 * calls that overlap on one GPU race on it.  Kernel errors: lm_collision_all_cells_check(1, stream).
 */
#define LM_NOALLOC_MAX_SIPS (1 << 20)
#define LM_NOALLOC_MAX_CELLS (1 << 15)
int lm_collision_all_cells_noalloc(int64_t n_sips, double *xis, double *m_w, double *m_s, const double *vel,
                                   const double *grid_T, const int64_t *cells, int32_t nx, int32_t nz,
                                   double dt_over_dV, const double *E_col, int32_t nb, double lnR0, double lnf,
                                   double E_const, double E_max, int64_t *no_cols, int32_t solute, uint64_t seed,
                                   uint32_t step, const double *rnd_stream, int64_t rnd_len, void *stream);
/* Error word of the last all-cells calls on this device (no_alloc 0: lm_collision_all_cells, 1: its
   no-allocation form); synchronises the stream, clears the word. */
int lm_collision_all_cells_check(int32_t no_alloc, void *stream);

/*
 * Ensemble statistics of box-model snapshots, replacing the sums of analyze_sim_data
 * (collision/box_model.py:1354-1585).  The snapshots of all seeds are (n_times, n_tot) arrays of
 * multiplicities and masses (kg), seed s in the columns [off[s], off[s+1]).
 * lm_box_minmax: per time (m_min, m_max, xi_min, xi_max) into out (n_times, 4).
 * lm_box_stats: per (seed, time) the moments sum(xi m^n)/dV, n = 0..3 (moments (n_sims, n_times, 4)), and
 * over the no_bins (<= 64) mass bins edges (n_times, no_bins+1) of np.histogram's convention the spectra
 * f_m = sum xi / (width_m dV), g_m = sum xi m / (width_m dV), g_ln_r = sum xi m / (width_lnr dV)
 * (each (n_sims, n_times, no_bins)); per-bin sums in ascending SIP order.
 * lm_mean_sem: over the first axis of fields (n_sims, n_vals), the mean and std(ddof=1)/sqrt(n_sims).
 */
int lm_box_minmax(int32_t n_times, int64_t n_tot, const double *xis, const double *masses, double *out,
                  void *stream);
int lm_box_stats(int32_t n_times, int32_t n_sims, const int64_t *off, int64_t n_tot, const double *xis,
                 const double *masses, const double *edges, const double *width_m, const double *width_lnr,
                 int32_t no_bins, double dV, double *moments, double *f_m, double *g_m, double *g_ln_r,
                 void *stream);
int lm_mean_sem(const double *fields, int32_t n_sims, int64_t n_vals, double *mean, double *sem, void *stream);

/*
 * integration.py:150-286 compute_divergence_upwind_np for n_fields stacked
 * (nx, nz) fields with one (2, nx+1, nz+1) flux field: limited third-order
 * upwind flux divergence, boundary type 0 = periodic, 1 = fixed per axis.
 * out is (n_fields, nx, nz).
 */
int lm_divergence_upwind(const double *fields, const double *flux, double *out, int32_t n_fields, int32_t nx,
                         int32_t nz, double dx, double dz, int32_t bc_x, int32_t bc_z, void *stream);

/*
 * Adjoint of a window of advection steps (the reference has no adjoint; these
 * entry points differentiate its step, integration.py:1298-1611, as this
 * library computes it).  Collisions are fixed events: a window is first run
 * with ev_mode 1 (lm_adv_steps records every event with its pair, branch and
 * p_crit), then re-run from checkpoints with ev_mode 2 (the recorded events are
 * applied to the current state instead of drawing) and a tape, and swept
 * backwards segment by segment with lm_adj_adv_steps.  The gradients are those
 * of the computed map: the fixed 3-iteration Newton with its clamps, the grid
 * feedback, advection with the limiter, relaxation and the fixed events.
 * Positions and velocities carry no sensitivity: with fixed events and cells,
 * the objectives depend on them only through cell membership (piecewise
 * constant) and through p_crit of multiple collections, which a fixed event
 * keeps as recorded.
 */

/* gid -> slot map of the current buffer (ev_mode 2 builds it after every reindex). */
int lm_build_slot_of(LmEngine *e, void *stream);

/*
 * Objective sums per cell, over the active members in ascending id, as
 * compensated sums (out[(2+q)*n_runs*nc + gc] holds the correction of sum q):
 * out[gc] = sum xi m_w, out[n_runs*nc + gc] = sum xi I(R) with the smooth
 * cloud-class indicator I(R) = s((ln R - ln 0.5)/w) s((ln 25 - ln R)/w),
 * s the logistic function, w = nc_width, R the radius at the cell temperature
 * (microphysics.py:179-266).  The hard class of evaluation.py:319 has zero
 * derivative almost everywhere; I(R) is its smooth stand-in.
 */
int lm_objective(LmEngine *e, double nc_width, double *out, void *stream);

/*
 * Seeds the adjoint with d J / d(state) at the end of the window, J = w_lwp
 * J_lwp + w_nc J_nc, J_lwp = 1e-18 sum xi m_w / (X dy) (domain-mean liquid
 * water path, kg m^-2), J_nc = 1e-6 sum_cells m_dry_inv sum xi I(R) / (n_runs
 * nc) (domain-mean cloud droplet number, mg^-1).  Clears every other adjoint
 * variable.
 */
int lm_adj_seed(LmEngine *e, LmAdjoint *a, double w_nc, double w_lwp, double nc_width, void *stream);

/*
 * Reverse sweep over n_steps advection steps whose tape (substeps, increment
 * fields) starts at index 0 of e->tape_*, and whose first collision step is
 * window collision step col_index0 (events in e->ev, e->ev_off).  Moves the
 * adjoint from the end of the last step to the start of the first.
 */
int lm_adj_adv_steps(LmEngine *e, LmAdjoint *a, int32_t n_steps, int32_t col_index0, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* LACMO_AMD_H */
