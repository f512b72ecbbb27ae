"""Device-resident SIP state of one or many independent runs on one GPU.

The Engine owns torch tensors on the GPU (plumbing only) and an LmEngine
struct that points into them; every numerical operation is a HIP kernel of
liblacmo_amd.so.  SIP arrays are kept as structure-of-arrays sorted by
(run, home cell, SIP id); the reference's ID-ordered layout (pos (2,N),
vel (2,N), cells (2,N) int64, m_w, m_s, xi, active_ids) is only produced on
export.
"""
import ctypes as C
import math
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import torch

from ._native import LmEngine, check, lib

SOLUTES = {"AS": 0, "NaCl": 1}
CC_N, CC_TP = 16, 15  # per-cell constant table: width, slot of the pre-update cell temperature (lm_physics.h)
WORK_BINS = 1024  # LM_WORK_BINS: size bins of the work-list order (include/lacmo_amd.h)
REM_CAP_MAX = 1024  # LM_REM_CAP_MAX: ground removals per run and substep (include/lacmo_amd.h)
MIG_CAP = 32  # migrant-list entries per cell (LmEngine.mig_cap, at most LM_MIG_CAP_MAX = 64)
BIG_CAP = 46340  # members of a cell the collision step takes (LmEngine.big_cap = LM_BIG_CELL_MAX)
BIG_BLOCKS = 4  # workgroups of the collision pass over cells of more than 1024 members, one workspace region each


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else C.c_void_p(0)


class Engine:
    """Batched SIP engine: n_runs runs of the same grid geometry."""

    def __init__(self, nx, nz, n_per_run, ranges, steps, volume_cell, p_ref=1.0E5, p_ref_inv=1.0E-5,
                 device="cuda", rem_cap=1024):
        if not 1 <= int(rem_cap) <= REM_CAP_MAX:
            raise ValueError(f"rem_cap must be in 1..{REM_CAP_MAX} (removals per run and substep)")
        if not torch.cuda.is_available():
            raise RuntimeError("lacmopy_amd.Engine needs a GPU (HIP device); there is no CPU fallback")
        self.L = lib()
        self.device = torch.device(device)
        self.nx, self.nz = int(nx), int(nz)
        self.nc = self.nx * self.nz
        self.n_per_run = [int(n) for n in n_per_run]
        self.n_runs = len(self.n_per_run)
        self.base = np.concatenate([[0], np.cumsum(self.n_per_run)]).astype(np.int64)
        self.n_tot = int(self.base[-1])
        self.ranges = np.asarray(ranges, dtype=np.float64)
        self.steps = np.asarray(steps, dtype=np.float64)
        nr, nc, dev = self.n_runs, self.nc, self.device
        f64 = dict(dtype=torch.float64, device=dev)
        i32 = dict(dtype=torch.int32, device=dev)
        i64 = dict(dtype=torch.int64, device=dev)
        n = max(self.n_tot, 1)
        self.buf = {}

        def mk(name, shape, kw):
            t = torch.zeros(shape, **kw)
            self.buf[name] = t
            return t

        for b in (0, 1):
            for f in ("x", "z", "u", "w", "m_w", "m_s", "xi"):
                mk(f"{f}{b}", (n,), f64)
            mk(f"id{b}", (n,), i32)
            mk(f"cell{b}", (n,), i32)
            mk(f"seg{b}", (nr * (nc + 1) + 1,), i32)
        mk("dmxi", (n,), f64)
        mk("away", (nr * nc,), i32)
        mk("mig_cnt", (nr * nc,), i32)
        mk("mig_slot", (nr * nc * MIG_CAP,), i32)
        mk("col_step_dev", (1,), i32)
        mk("scal", (nr, 10, nc), f64)
        mk("mat", (nr, 7, nc), f64)
        mk("cc", (nr * nc, CC_N), f64)
        mk("dTh", (nr * nc,), f64)
        mk("dRv", (nr * nc,), f64)
        mk("dml", (nr * nc,), f64)
        mk("grid_vel", (nr, 2, self.nx + 1, self.nz + 1), f64)  # face velocities of each run
        mk("flux", (2, self.nx + 1, self.nz + 1), f64)
        mk("prof_rv", (nr, self.nz), f64)
        mk("prof_Th", (nr, self.nz), f64)
        mk("t_relax", (self.nz,), f64)
        mk("colsum", (nr * self.nz * 2,), f64)
        mk("no_cols", (nr, 2), i64)
        mk("water_removed", (nr,), f64)
        mk("n_equal_xi", (nr,), i64)
        mk("rem_cnt", (nr,), i32)
        mk("rem_slot", (nr * rem_cap,), i32)
        L = nr * (2 * nc + 1)
        mk("scan_buf", (L + 1,), i32)
        mk("scan_part", ((L + 4095) // 4096 + 1,), i32)
        mk("err", (4,), i32)
        mk("ovf_list", (3 * nr * nc,), i32)
        mk("ovf_cnt", (288,), i32)
        mk("rnd_off", (nr * nc + 2,), i64)
        # (run, cell) work list of the persistent passes (rebuilt by every reindex while one is on; starts as the
        # identity order), its size-bin scratch and two work counters
        mk("work", (nr * nc + WORK_BINS + 6,), i32)
        mk("col_cost", (nr * nc,), i32)  # last collision step per cell (ticks): the work list's order
        self.buf["work"][:nr * nc].copy_(torch.arange(nr * nc, dtype=torch.int32))
        self.E_tab = None
        self.rnd_stream = None

        e = self.e = LmEngine()
        e.n_runs, e.nx, e.nz, e.nc = nr, self.nx, self.nz, nc
        e.n_tot = self.n_tot
        e.x_min, e.x_max = float(self.ranges[0, 0]), float(self.ranges[0, 1])
        e.z_min, e.z_max = float(self.ranges[1, 0]), float(self.ranges[1, 1])
        e.dx, e.dz = float(self.steps[0]), float(self.steps[1])
        e.V_cell, e.p_ref, e.p_ref_inv = float(volume_cell), float(p_ref), float(p_ref_inv)
        e.cur, e.ccur = 0, 0
        for b in (0, 1):
            for f in ("x", "z", "u", "w", "m_w", "m_s", "xi", "id", "cell", "seg"):
                getattr(e, f)[b] = self.buf[f"{f}{b}"].data_ptr()
        for f in ("dmxi", "away", "col_step_dev", "scal", "mat", "cc", "dTh", "dRv", "dml", "grid_vel", "flux", "prof_rv",
                  "prof_Th", "t_relax", "colsum", "no_cols", "water_removed", "n_equal_xi", "rem_cnt",
                  "rem_slot", "scan_buf", "scan_part", "err", "ovf_list", "ovf_cnt", "rnd_off", "work", "mig_cnt", "col_cost",
                  "mig_slot"):
            setattr(e, f, self.buf[f].data_ptr())
        e.rem_cap = int(rem_cap)
        e.mig_cap = MIG_CAP
        # the collision pass over cells of more than 1024 members: BIG_BLOCKS regions of its member state
        ws = int(self.L.lm_big_ws_bytes(BIG_CAP))
        self.buf["big_ws"] = torch.empty(BIG_BLOCKS * ws, dtype=torch.uint8, device=self.device)
        e.big_ws = self.buf["big_ws"].data_ptr()
        e.big_cap, e.big_blocks = BIG_CAP, BIG_BLOCKS
        e.vel_run_stride = 2 * (self.nx + 1) * (self.nz + 1)
        e.e_const_mode = 1
        e.E_const = 0.5
        self.set_params()

    @classmethod
    def from_initial_states(cls, states, config, E_col_grid, device="cuda"):
        """One run per initial state (lacmopy_amd.initialize.initialize_batch: grid and SIPs of one
        seed_SIP_gen each), all with the step parameters of config as simulate reads them
        (integration.py:1908-1923): solute_type, dt_adv, dt_col, scale_dt_cond, no_col_per_adv,
        no_iter_impl_mass, g_set, act_collisions, act_relaxation, seed_sim and, for a table E_col_grid,
        R_kernel_low_log and bin_factor_R_log.  Each run keeps its own grid fields, relaxation profiles
        (integration.py:1948-1955, from its own initial grid) and face velocities."""
        from .integration import update_cells_and_rel_pos
        from .setup_case import init_profiles
        g0 = states[0].grid
        eng = cls(int(g0.no_cells[0]), int(g0.no_cells[1]), [len(s.xi) for s in states], g0.ranges, g0.steps,
                  g0.volume_cell, device=device)
        eng.set_params(solute=config["solute_type"], no_iter_impl_mass=config["no_iter_impl_mass"],
                       scale_dt_cond=config["scale_dt_cond"], no_col_per_adv=config["no_col_per_adv"],
                       act_collisions=config["act_collisions"], act_relaxation=config["act_relaxation"],
                       dt=config["dt_adv"], dt_col=config["dt_col"], g_set=config["g_set"], seed=config["seed_sim"])
        if isinstance(E_col_grid, float):
            eng.set_collision_kernel(E_col_grid)
        else:
            eng.set_collision_kernel(E_col_grid, config["R_kernel_low_log"], config["bin_factor_R_log"])
        eng.set_flow(g0.velocity, g0.mass_flux_air_dry)
        for r, s in enumerate(states):
            if s.grid.velocity.shape != g0.velocity.shape or not np.array_equal(s.grid.ranges, g0.ranges):
                raise ValueError("the runs of one engine share the grid geometry")
            eng.set_flow(s.grid.velocity, g0.mass_flux_air_dry, run=r)
            eng.set_relaxation(r, *init_profiles(s.grid))
            eng.set_grid(r, s.grid.scalar_fields(), s.grid.mat_prop())
        eng.init_cells()
        runs = []
        for s in states:
            # simulate() starts by deriving the cells from the positions (integration.py:1932-1933); a position
            # drawn at the far edge of its cell can round onto the next cell's corner
            cells = np.array(s.cells, copy=True)
            update_cells_and_rel_pos(s.pos, cells, np.zeros_like(s.pos), s.active_ids, s.grid.ranges, s.grid.steps)
            runs.append((s.pos, s.vel, cells, s.m_w, s.m_s, s.xi, s.active_ids))
        eng.load_particles(runs)
        return eng

    # ------------------------------------------------------------------ setup
    @property
    def stream(self):
        return C.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def set_params(self, solute="AS", no_iter_impl_mass=3, scale_dt_cond=5, no_col_per_adv=2,
                   act_collisions=True, act_relaxation=True, dt=1.0, dt_col=None, g_set=9.80665, seed=0):
        """Step parameters as config.py / init.set_config derive them (init.py:181-206, integration.py:1992)."""
        self.drop_graph()
        e = self.e
        e.solute = SOLUTES[solute]
        e.no_iter = int(no_iter_impl_mass)
        e.scale_dt_cond = int(scale_dt_cond)
        e.no_col_per_adv = int(no_col_per_adv)
        e.act_collisions = int(bool(act_collisions))
        e.act_relaxation = int(bool(act_relaxation))
        e.dt = float(dt)
        e.dt_sub = dt / (2 * scale_dt_cond)
        e.dt_sub_half = 0.5 * e.dt_sub
        if dt_col is None:
            dt_col = dt / no_col_per_adv
        e.dt_col_over_dV = dt_col / e.V_cell
        e.g_set = float(g_set)
        e.seed = int(seed) & 0xFFFFFFFFFFFFFFFF

    def set_collision_kernel(self, E_col_grid, R_kernel_low_log=None, bin_factor_R_log=None):
        """Table E_col(R_i,R_j) (Ecol_grid_R) or a float (Ecol_const), as passed to the reference."""
        self.drop_graph()
        e = self.e
        if isinstance(E_col_grid, float):
            e.e_const_mode, e.E_const, e.nb = 1, float(E_col_grid), 0
            self.E_tab = None
            e.E = None
            e.E_max = 0.0
        else:
            if isinstance(E_col_grid, torch.Tensor):
                self.E_tab = E_col_grid.to(device=self.device, dtype=torch.float64).contiguous()
            else:
                self.E_tab = torch.from_numpy(np.ascontiguousarray(E_col_grid, dtype=np.float64)).to(self.device)
            E = self.E_tab
            e.E = self.E_tab.data_ptr()
            # bound for the collision kernel's first filter (disabled for non-finite or negative tables)
            e.E_max = float(E.max()) if bool(torch.isfinite(E).all()) and float(E.min()) >= 0.0 else 0.0
            e.nb = int(E.shape[0])
            e.e_const_mode = 0
            e.lnR0 = float(R_kernel_low_log)
            e.lnf = float(bin_factor_R_log)

    def set_flow(self, velocity, mass_flux_air_dry, run=None):
        """The dry-air mass flux (one field, the same for all runs) and the face velocities derived from it:
        velocity (2, nx+1, nz+1) for every run, or for run `run` only (runs initialised apart divide the flux
        by their own dry-density profile, init.py:1179-1182)."""
        v = _as_f64(velocity)
        if run is None:
            self.buf["grid_vel"].copy_(v.expand_as(self.buf["grid_vel"]))
        else:
            self.buf["grid_vel"][run].copy_(v)
        self.buf["flux"].copy_(_as_f64(mass_flux_air_dry))

    def set_relaxation(self, run, init_profile_r_v, init_profile_Theta, relaxation_time_profile):
        self.buf["prof_rv"][run].copy_(_as_f64(init_profile_r_v))
        self.buf["prof_Th"][run].copy_(_as_f64(init_profile_Theta))
        self.buf["t_relax"].copy_(_as_f64(relaxation_time_profile))

    def set_grid(self, run, grid_scalar_fields, grid_mat_prop):
        """(10,nx,nz) and (7,nx,nz) in the index order of integration.py:366-386 (arrays or tensors)."""
        self.buf["scal"][run].copy_(_as_f64(grid_scalar_fields).reshape(10, self.nc))
        self.buf["mat"][run].copy_(_as_f64(grid_mat_prop).reshape(7, self.nc))

    def init_cells(self):
        check(self.L.lm_engine_init_cells(C.byref(self.e), self.stream))

    def load_particles(self, runs):
        """runs: list of (pos, vel, cells, m_w, m_s, xi, active_ids) in the reference layout."""
        assert len(runs) == self.n_runs
        host, hid, hcell, seg = host_layout(runs, self.nx, self.nz, self.n_per_run)
        cur = self.e.cur
        for f, a in host.items():
            self.buf[f"{f}{cur}"].copy_(torch.from_numpy(a))
        self.buf[f"id{cur}"].copy_(torch.from_numpy(hid))
        self.buf[f"cell{self.e.ccur}"].copy_(torch.from_numpy(hcell))
        self.buf[f"seg{cur}"].copy_(torch.from_numpy(seg))
        self.buf["away"].zero_()  # every SIP sits in its home cell
        self.buf["mig_cnt"].zero_()

    def load_particles_device(self, run, pos, vel, cells, m_w, m_s, xi, active_ids):
        """One run from the reference's layout in device memory -- CUDA tensors pos, vel (2, N) float64, cells
        (2, N) int64, m_w, m_s, xi (N) float64, active_ids (N) bool -- reordered into the engine's layout on the
        device (lm_load_reference_layout); no host copy."""
        N = self.n_per_run[run]
        t = _ref_tensors(N, pos=pos, vel=vel, cells=cells, m_w=m_w, m_s=m_s, xi=xi, active_ids=active_ids)
        check(self.L.lm_load_reference_layout(C.byref(self.e), int(run), int(self.base[run]), N, *t, self.stream))

    def store_particles_device(self, run, pos, vel, cells, m_w, m_s, xi, active_ids, rel_pos=None, T_p=None):
        """The inverse of load_particles_device into the caller's CUDA tensors (updated in place), with rel_pos
        (2, N) of the active SIPs and T_p (N) when given (lm_store_reference_layout)."""
        N = self.n_per_run[run]
        t = _ref_tensors(N, pos=pos, vel=vel, cells=cells, rel_pos=rel_pos, m_w=m_w, m_s=m_s, xi=xi,
                         active_ids=active_ids, T_p=T_p)
        check(self.L.lm_store_reference_layout(C.byref(self.e), int(run), int(self.base[run]), N, *t, self.stream))

    def export_grid_into(self, run, grid_scalar_fields, grid_mat_prop):
        """The grid fields of one run copied into the caller's arrays or tensors (in place)."""
        for name, dst in (("scal", grid_scalar_fields), ("mat", grid_mat_prop)):
            if dst is None:
                continue
            src = self.buf[name][run].reshape(tuple(dst.shape))
            if isinstance(dst, torch.Tensor):
                dst.copy_(src)
            else:
                dst[...] = src.cpu().numpy()

    def set_counters(self, run, no_cols=None, water_removed=None):
        if no_cols is not None:
            self.buf["no_cols"][run].copy_(no_cols.reshape(2) if isinstance(no_cols, torch.Tensor) else
                                           torch.as_tensor(np.asarray(no_cols, dtype=np.int64)))
        if water_removed is not None:
            if isinstance(water_removed, torch.Tensor):
                self.buf["water_removed"][run].copy_(water_removed.reshape(-1)[0])
            else:
                self.buf["water_removed"][run] = float(np.asarray(water_removed).reshape(-1)[0])

    def set_rnd_stream(self, stream_values):
        """Inject the uniform stream (1-D float64, consumed in the reference's order) or None for Philox."""
        self.drop_graph()
        if stream_values is None:
            self.rnd_stream = None
            self.e.rnd_stream = None
            self.e.rnd_len = 0
            return
        self.rnd_stream = torch.as_tensor(np.asarray(stream_values, dtype=np.float64).reshape(-1),
                                          device=self.device)
        self.e.rnd_stream = self.rnd_stream.data_ptr()
        self.e.rnd_len = int(self.rnd_stream.numel())
        self.buf["rnd_off"].zero_()

    # ------------------------------------------------------------- stepping
    def adv_steps(self, n):
        check(self.L.lm_adv_steps(C.byref(self.e), int(n), self.stream))

    def collision_step(self):
        check(self.L.lm_collision_step(C.byref(self.e), self.stream))

    def position_step(self, dt):
        check(self.L.lm_position_step(C.byref(self.e), float(dt), self.stream))

    def reindex(self):
        check(self.L.lm_reindex(C.byref(self.e), self.stream))

    def build_work_list(self):
        """Orders the (run, cell) work list of the persistent passes by decreasing cost: each cell's last
        collision step (the col_cost buffer, kept unless LmEngine.col_cost is set to None) or else its current
        home-segment size."""
        check(self.L.lm_build_work_list(C.byref(self.e), self.stream))

    def advection_increments(self, second):
        check(self.L.lm_advection_increments(C.byref(self.e), int(second), self.stream))

    def substep(self, dt_sub, dt_sub_pos):
        check(self.L.lm_substep(C.byref(self.e), float(dt_sub), float(dt_sub_pos), self.stream))

    def particle_substep(self, dt_sub, dt_sub_pos):
        check(self.L.lm_particle_substep(C.byref(self.e), float(dt_sub), float(dt_sub_pos), self.stream))

    def grid_substep(self):
        check(self.L.lm_grid_substep(C.byref(self.e), self.stream))

    def update_r_l(self):
        check(self.L.lm_update_r_l(C.byref(self.e), self.stream))

    def set_increments(self, run, delta_Theta_ad=None, delta_r_v_ad=None, delta_m_l=None):
        """Per-cell advection increments / condensed mass of one run ((nx, nz) arrays)."""
        sl = slice(run * self.nc, (run + 1) * self.nc)
        for name, a in (("dTh", delta_Theta_ad), ("dRv", delta_r_v_ad), ("dml", delta_m_l)):
            if a is not None:
                self.buf[name][sl].copy_(_as_f64(a).reshape(-1))

    def export_increments(self, run):
        """(delta_Theta_ad, delta_r_v_ad, delta_m_l) of one run as (nx, nz) arrays."""
        self.check()
        sl = slice(run * self.nc, (run + 1) * self.nc)
        return tuple(self.buf[n][sl].cpu().numpy().reshape(self.nx, self.nz).copy() for n in ("dTh", "dRv", "dml"))

    @property
    def col_step(self):
        return int(self.e.col_step)

    @col_step.setter
    def col_step(self, step):
        check(self.L.lm_set_col_step(C.byref(self.e), int(step) & 0xFFFFFFFF, self.stream))

    # ------------------------------------------------------------ graphs
    def _graph_stream(self):
        if getattr(self, "_gstream", None) is None:
            self._gstream = torch.cuda.Stream(device=self.device)
        return self._gstream

    def graph_capture(self, n_steps=2):
        """Capture n_steps (even) advection steps into a HIP graph; returns its handle (see graph_launch)."""
        gs = self._graph_stream()
        gs.wait_stream(torch.cuda.current_stream(self.device))
        h = C.c_void_p()
        check(self.L.lm_graph_capture(C.byref(self.e), int(n_steps), C.c_void_p(gs.cuda_stream), C.byref(h)))
        return h

    def graph_launch(self, h):
        gs = self._graph_stream()
        gs.wait_stream(torch.cuda.current_stream(self.device))
        check(self.L.lm_graph_launch(C.byref(self.e), h, C.c_void_p(gs.cuda_stream)))
        torch.cuda.current_stream(self.device).wait_stream(gs)

    def graph_destroy(self, h):
        check(self.L.lm_graph_destroy(h))

    def adv_steps_graph(self, n):
        """n advection steps as replays of a cached two-step graph (+ one eager step if n is odd).

        One graph is kept per buffer parity (an odd number of eager steps swaps the SIP buffers); the
        setters that change what the kernels are bound to (parameters, kernel table, injected stream)
        drop the cached graphs."""
        n = int(n)
        if n >= 2:
            key = (int(self.e.cur), int(self.e.ccur))
            graphs = self.__dict__.setdefault("_graphs", {})
            if key not in graphs:
                graphs[key] = self.graph_capture(2)
            for _ in range(n // 2):
                self.graph_launch(graphs[key])
        if n % 2:
            self.adv_steps(1)

    def drop_graph(self):
        for h in self.__dict__.pop("_graphs", {}).values():
            self.graph_destroy(h)

    def check(self):
        check(self.L.lm_engine_check(C.byref(self.e), self.stream))

    # ------------------------------------------------------------ autotune
    SHAPE_KEYS = ("part_block", "part_spt", "part_wpe", "col_cap", "col_work_per_cu", "col_work_chunk",
                  "dml_work_per_cu", "dml_work_chunk")

    def set_launch_shape(self, part_block=0, part_spt=0, col_cap=0, col_work_per_cu=None, dml_work_per_cu=None,
                         col_work_chunk=None, dml_work_chunk=None, part_wpe=None):
        """Launch shape of the particle kernel (threads per block, SIPs per thread, and the waves per SIMD
        it is compiled for: 0 the default build, 5, 6 or 8 with one SIP per thread), the member capacity
        of the main collision pass, and the persistent main collision and Delta m_l passes over the
        (run, cell) work list: workgroups per compute unit (0: one workgroup per cell) and items taken per
        atomic grab (0: static round robin); None keeps the current value.  0 restores the defaults (256,
        1, from the mean SIPs per cell, one workgroup per cell)."""
        self.drop_graph()
        e = self.e
        e.part_block, e.part_spt, e.col_cap = int(part_block), int(part_spt), int(col_cap)
        if part_wpe is not None:
            if int(part_wpe) not in (0, 5, 6, 8):
                raise ValueError("part_wpe must be 0, 5, 6 or 8")
            e.part_wpe = int(part_wpe)
        for k, val in (("col_work_per_cu", col_work_per_cu), ("dml_work_per_cu", dml_work_per_cu),
                       ("col_work_chunk", col_work_chunk), ("dml_work_chunk", dml_work_chunk)):
            if val is not None:
                hi = 256 if k.endswith("per_cu") else 4096
                if not 0 <= int(val) <= hi:
                    raise ValueError(f"{k} must be in 0..{hi}")
                setattr(e, k, int(val))

    def launch_shape(self):
        return {k: int(getattr(self.e, k)) for k in self.SHAPE_KEYS}

    def autotune(self, blocks=(64, 128, 256, 512), spts=(1, 2), wpes=(0, 5, 6, 8), col_caps=(160, 256, 320, 512),
                 col_work=((0, 0), (16, 0), (16, 1), (16, 4), (32, 0), (32, 4)),
                 dml_work=((0, 0), (32, 0), (32, 16), (64, 0), (64, 16)), reps=3):
        """Times candidate launch shapes on the engine's current state with HIP events on its stream and
        keeps the fastest: the particle substep (with its Delta m_l pass) for every (block, SIPs per
        thread, waves per SIMD of the build; the non-default builds with one SIP per thread), then the Delta m_l pass's persistent shapes dml_work = (workgroups per CU, chunk) on the
        substep, the collision step for every main-pass capacity, then its persistent shapes col_work.
        (0, 0) is one workgroup per cell.  The state is restored before every candidate and at the end;
        results do not depend on the launch shape.  Returns {"particles": {(block, spt, wpe): ms}, "dml":
        {(wpc, chunk): ms}, "collision": {cap: ms}, "collision_work": {(wpc, chunk): ms}, "chosen":
        launch_shape()}."""
        snap = self.snapshot()
        e = self.e
        stream = torch.cuda.current_stream(self.device)

        def timed(run):
            self.restore(snap)
            run()  # first launch of the candidate, untimed
            self.restore(snap)
            t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            t0.record(stream)
            for _ in range(reps):
                run()
            t1.record(stream)
            t1.synchronize()
            self.check()
            return t0.elapsed_time(t1) / reps

        def substep():
            self.particle_substep(e.dt_sub, e.dt_sub)

        res = {"particles": {}, "dml": {}, "collision": {}, "collision_work": {}}
        shape = self.launch_shape()
        self.build_work_list()
        for b in blocks:
            for spt in spts:
                for wpe in wpes:
                    if wpe and spt != 1:
                        continue
                    shape.update(part_block=b, part_spt=spt, part_wpe=wpe, dml_work_per_cu=0, dml_work_chunk=0)
                    self.set_launch_shape(**shape)
                    res["particles"][(b, spt, wpe)] = timed(substep)
        shape["part_block"], shape["part_spt"], shape["part_wpe"] = min(res["particles"], key=res["particles"].get)
        for wpc, chunk in dml_work:
            shape.update(dml_work_per_cu=wpc, dml_work_chunk=chunk)
            self.set_launch_shape(**shape)
            res["dml"][(wpc, chunk)] = timed(substep)
        shape["dml_work_per_cu"], shape["dml_work_chunk"] = min(res["dml"], key=res["dml"].get)
        if e.act_collisions:
            for cap in col_caps:
                shape.update(col_cap=cap, col_work_per_cu=0, col_work_chunk=0)
                self.set_launch_shape(**shape)
                res["collision"][cap] = timed(self.collision_step)
            shape["col_cap"] = min(res["collision"], key=res["collision"].get)
            for wpc, chunk in col_work:
                shape.update(col_work_per_cu=wpc, col_work_chunk=chunk)
                self.set_launch_shape(**shape)
                res["collision_work"][(wpc, chunk)] = timed(self.collision_step)
            shape["col_work_per_cu"], shape["col_work_chunk"] = min(res["collision_work"],
                                                                    key=res["collision_work"].get)
        self.set_launch_shape(**shape)
        self.restore(snap)
        res["chosen"] = self.launch_shape()
        return res

    # ------------------------------------------------------- checkpoints
    _SNAP_SIP = ("x", "z", "u", "w", "m_w", "m_s", "xi", "id", "seg")
    _SNAP_GRID = ("away", "mig_cnt", "mig_slot", "col_step_dev", "scal", "mat", "cc", "dTh", "dRv", "dml", "no_cols",
                  "water_removed", "n_equal_xi", "rem_cnt", "err")

    def snapshot(self):
        """Device copy of the evolving state (current SIP buffer and cells, grid, counters)."""
        cur, ccur = int(self.e.cur), int(self.e.ccur)
        names = [f"{f}{cur}" for f in self._SNAP_SIP] + [f"cell{ccur}"] + list(self._SNAP_GRID)
        return {"buf": {n: self.buf[n].clone() for n in names}, "cur": cur, "ccur": ccur,
                "col_step": int(self.e.col_step)}

    def restore(self, snap):
        """Returns the engine to a snapshot (buffers are copied, the snapshot stays valid)."""
        for n, t in snap["buf"].items():
            self.buf[n].copy_(t)
        self.e.cur, self.e.ccur, self.e.col_step = snap["cur"], snap["ccur"], snap["col_step"]

    KINDS = ("particles", "collision", "dml", "reindex_scatter", "collision_overflow")

    def prof_enable(self, max_events=4096):
        check(self.L.lm_prof_enable(C.byref(self.e), int(max_events)))

    def prof_collect(self):
        nk = len(self.KINDS)
        ms = (C.c_double * nk)()
        cnt = (C.c_int32 * nk)()
        check(self.L.lm_prof_collect(C.byref(self.e), ms, cnt, nk))
        return {k: (ms[i], cnt[i]) for i, k in enumerate(self.KINDS)}

    def prof_disable(self):
        check(self.L.lm_prof_disable(C.byref(self.e)))

    # --------------------------------------------------------------- export
    def export_particles(self, run):
        """Reference layout arrays of one run (pos, vel, cells, m_w, m_s, xi, active_ids)."""
        self.check()
        cur, ccur = self.e.cur, self.e.ccur
        b, N = int(self.base[run]), self.n_per_run[run]
        g = {f: self.buf[f"{f}{cur}"][b:b + N].cpu().numpy() for f in ("x", "z", "u", "w", "m_w", "m_s", "xi")}
        ids = self.buf[f"id{cur}"][b:b + N].cpu().numpy()
        cell = self.buf[f"cell{ccur}"][b:b + N].cpu().numpy()
        return host_unlayout(g, ids, cell, run * self.nc, self.nz)

    def export_T_p(self, run):
        """SIP temperatures in ID order: the cell temperature each SIP saw in the last particle substep.

        integration.py:489-507 sets T_p = T[cells] at the start of every particle
        substep, before positions move; inactive SIPs (cells[1] = -1) read the
        top row of their column.  The engine keeps the cell of that substep in
        the other cell buffer and its temperature in the per-cell table.
        """
        self.check()
        cur, ccur = self.e.cur, self.e.ccur
        b, N = int(self.base[run]), self.n_per_run[run]
        ids = self.buf[f"id{cur}"][b:b + N].cpu().numpy().astype(np.int64)
        prev = self.buf[f"cell{1 - ccur}"][b:b + N].cpu().numpy().astype(np.int64)
        tp = self.buf["cc"][run * self.nc:(run + 1) * self.nc, CC_TP].cpu().numpy()
        loc = np.where(prev >= 0, prev - run * self.nc, (-prev - 2) * self.nz + (self.nz - 1))
        out = np.empty(N)
        out[ids] = tp[loc]
        return out

    def export_grid(self, run):
        self.check()
        s = self.buf["scal"][run].cpu().numpy().reshape(10, self.nx, self.nz).copy()
        m = self.buf["mat"][run].cpu().numpy().reshape(7, self.nx, self.nz).copy()
        return s, m

    def counters(self, run):
        self.check()
        return (self.buf["no_cols"][run].cpu().numpy().copy(),
                np.array([float(self.buf["water_removed"][run].item())]))


def _as_f64(a):
    """A float64 tensor of an array (host) or of a tensor (kept where it is)."""
    if isinstance(a, torch.Tensor):
        return a.to(torch.float64)
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64))


_REF_SPEC = {"pos": (torch.float64, 2), "vel": (torch.float64, 2), "cells": (torch.int64, 2), "m_w": (torch.float64, 1),
             "m_s": (torch.float64, 1), "xi": (torch.float64, 1), "active_ids": (torch.bool, 1),
             "rel_pos": (torch.float64, 2), "T_p": (torch.float64, 1)}


def _ref_tensors(N, **arrays):
    """Device pointers of the reference-layout tensors, in argument order, after checking device, dtype, shape
    and contiguity (None -> a null pointer)."""
    out = []
    for name, t in arrays.items():
        if t is None:
            out.append(C.c_void_p(0))
            continue
        dtype, rows = _REF_SPEC[name]
        shape = (2, N) if rows == 2 else (N,)
        if not isinstance(t, torch.Tensor) or t.device.type != "cuda" or t.dtype != dtype or not t.is_contiguous() \
                or tuple(t.shape) != shape:
            raise TypeError(f"{name}: a contiguous CUDA {dtype} tensor of shape {shape} is required")
        out.append(C.c_void_p(t.data_ptr()))
    return out


def host_material_properties(T, p, r_v, rho_d):
    """integration.py:359-400 on the host with the reference's libm (Python math), elementwise."""
    T = np.asarray(T, dtype=np.float64)
    out = np.zeros((7,) + T.shape)
    flatT, flatp, flatr, flatd = T.ravel(), np.asarray(p).ravel(), np.asarray(r_v).ravel(), np.asarray(rho_d).ravel()
    o = out.reshape(7, -1)
    for k in range(flatT.size):
        t, pp, rv, rd = float(flatT[k]), float(flatp[k]), float(flatr[k]), float(flatd[k])
        o[0, k] = 4.1868E-3 * (5.69 + 0.017 * (t - 273.15))
        o[1, k] = 4.01218E-5 * math.pow(t, 1.94) / pp
        o[2, k] = 1.0E3 * (2500.82 - 2.358 * (t - 273.0))
        tau = 1 - t / 647.096
        o[3, k] = 0.2358 * math.pow(tau, 1.256) * (1 - 0.625 * tau)
        o[4, k] = 1004.71 * (1 + 0.897 * rv)
        o[5, k] = 1.458E-6 * math.pow(t, 1.5) / (t + 110.4)
        o[6, k] = rd * (1 + rv)
    return out


def host_layout(runs, nx, nz, n_per_run):
    """Engine layout of reference-layout runs (host side, pure numpy).

    SIPs of each run are ordered by (current cell, id) -- the reference visits
    cells x-major and the SIPs of a cell in ascending id -- with the inactive
    SIPs (cells[1] == -1) after the cells in ascending id.  Returns the SoA
    arrays, ids, engine cell codes (global cell, or -(2 + cells[0]) when
    inactive) and the segment table (per run: nc cell starts + graveyard start,
    then the total).
    """
    nc = nx * nz
    n_tot = int(sum(n_per_run))
    n = max(n_tot, 1)
    host = {f: np.zeros(n, dtype=np.float64) for f in ("x", "z", "u", "w", "m_w", "m_s", "xi")}
    hid = np.zeros(n, dtype=np.int32)
    hcell = np.zeros(n, dtype=np.int32)
    seg = np.zeros(len(runs) * (nc + 1) + 1, dtype=np.int32)
    bases = np.concatenate([[0], np.cumsum(np.asarray(n_per_run, dtype=np.int64))])

    def one(r):
        pos, vel, cells, m_w, m_s, xi, active = runs[r]
        base = int(bases[r])
        N = int(n_per_run[r])
        if len(xi) != N:
            raise ValueError("SIP count differs from the engine's allocation")
        active = np.asarray(active, dtype=bool)
        cells = np.asarray(cells, dtype=np.int64)
        ids = np.arange(N)
        key = cells[0] * nz + cells[1]
        ca = cells[:, active]
        if np.any((ca[0] < 0) | (ca[0] >= nx) | (ca[1] < 0) | (ca[1] >= nz)):
            raise ValueError("active SIP outside the grid")
        cd = cells[0, ~active]
        if np.any((cd < 0) | (cd >= nx)):
            raise ValueError("inactive SIP with a horizontal cell index outside the grid")
        act = ids[active]
        order = act[np.lexsort((act, key[active]))]
        dead = ids[~active]
        perm = np.concatenate([order, dead])
        sl = slice(base, base + N)
        host["x"][sl], host["z"][sl] = pos[0][perm], pos[1][perm]
        host["u"][sl], host["w"][sl] = vel[0][perm], vel[1][perm]
        host["m_w"][sl], host["m_s"][sl], host["xi"][sl] = m_w[perm], m_s[perm], xi[perm]
        hid[sl] = perm
        hcell[base:base + len(order)] = r * nc + key[order]
        hcell[base + len(order):base + N] = -(2 + cells[0][dead])
        counts = np.bincount(key[active], minlength=nc)
        seg[r * (nc + 1):r * (nc + 1) + nc + 1] = base + np.concatenate([[0], np.cumsum(counts)])

    # runs fill disjoint slices: laid out on host threads
    with ThreadPoolExecutor(max(1, min(16, len(runs), os.cpu_count() or 1))) as ex:
        for f in [ex.submit(one, r) for r in range(len(runs))]:
            f.result()
    seg[-1] = n_tot
    return host, hid, hcell, seg


def host_unlayout(g, ids, cell, cell0, nz):
    """Inverse of host_layout for one run: engine arrays -> reference layout (ID order)."""
    N = len(ids)
    ids = ids.astype(np.int64)
    cell = cell.astype(np.int64)
    pos = np.zeros((2, N))
    vel = np.zeros((2, N))
    cells = np.zeros((2, N), dtype=np.int64)
    m_w, m_s, xi = np.zeros(N), np.zeros(N), np.zeros(N)
    active = np.zeros(N, dtype=bool)
    pos[0, ids], pos[1, ids] = g["x"], g["z"]
    vel[0, ids], vel[1, ids] = g["u"], g["w"]
    m_w[ids], m_s[ids], xi[ids] = g["m_w"], g["m_s"], g["xi"]
    act = cell >= 0
    loc = np.where(act, cell - cell0, 0)
    cells[0, ids] = np.where(act, loc // nz, -cell - 2)
    cells[1, ids] = np.where(act, loc % nz, -1)
    active[ids] = act
    return pos, vel, cells, m_w, m_s, xi, active
