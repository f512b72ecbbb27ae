"""Drop-in replacements of the reference's time-step entry points (integration.py).

Same names, positional arguments and in-place effects as integration.py of
LacmoPy: the caller's numpy arrays (reference layout: pos/vel (2,N), cells
(2,N) int64, m_w, m_s, xi, T_p (N), active_ids (N) bool, grid_scalar_fields
(10,Nx,Nz), grid_mat_prop (7,Nx,Nz), no_cols (2) int64, water_removed (1))
are uploaded to one GPU, advanced by liblacmo_amd's HIP kernels and written
back in place.  ``simulate`` keeps the state resident in HBM for the whole
run and only copies it out for the frame and tracer dumps.

Differences a caller can see, all outside the reference's numerics:
  * collision draws come from the counter-based stream of lacmopy_amd.rng
    (the reference's np.random draws inside jitted code are not reproducible
    from any seed);
  * rel_pos is an output only: the kernels derive each SIP's relative
    position from its position and cell, which is the value the reference
    holds in rel_pos after every position update (integration.py:545-549);
    rel_pos of inactive SIPs is left as passed, where the reference keeps
    rescaling a stale value (:548-549) that nothing reads;
  * error conditions the reference would run through silently (a cell with
    more SIPs than the kernels hold, a SIP that left the domain sideways)
    raise lacmopy_amd._native.LacmoError.
"""
import math
import os
from datetime import datetime

import numpy as np

from . import file_handling as fh
from . import rng
from .engine import CC_TP, Engine

__all__ = [
    "compute_relaxation_time_profile", "compute_divergence_upwind_np", "compute_divergence_upwind",
    "propagate_grid_subloop_step_np", "propagate_grid_subloop_step", "propagate_particles_subloop_step_np",
    "propagate_particles_subloop_step", "integrate_subloop_n_steps_np", "integrate_subloop_n_steps",
    "integrate_adv_step_np", "integrate_adv_step", "simulate_interval", "simulate", "update_cells_and_rel_pos",
]

_cache = {"key": None, "eng": None}


def _engine(grid_no_cells, grid_ranges, grid_steps, grid_volume_cell, p_ref, p_ref_inv, N):
    """A one-run Engine of this geometry and SIP count (kept for the next call with the same shape)."""
    nx, nz = int(grid_no_cells[0]), int(grid_no_cells[1])
    ranges = np.asarray(grid_ranges, dtype=np.float64)
    steps = np.asarray(grid_steps, dtype=np.float64)
    key = (nx, nz, int(N), tuple(ranges.ravel()), tuple(steps.ravel()), float(grid_volume_cell), float(p_ref),
           float(p_ref_inv))
    if _cache["key"] != key:
        _cache["eng"], _cache["key"] = None, None
        _cache["eng"] = Engine(nx, nz, [int(N)], ranges, steps, float(grid_volume_cell), float(p_ref),
                               float(p_ref_inv))
        _cache["key"] = key
    return _cache["eng"]


def _check_ids(id_list, N):
    if id_list is None:
        return
    if _on_device(id_list):
        import torch
        ok = id_list.shape[0] == N and bool(torch.equal(id_list.long(), torch.arange(N, device=id_list.device)))
    else:
        ok = len(id_list) == N and np.array_equal(id_list, np.arange(N))
    if not ok:
        raise ValueError("id_list must be np.arange(N), as integration.py:1937 builds it")


def _set_step_params(eng, solute_type, dt, dt_sub, dt_sub_half, dt_col_over_dV, scale_dt_cond, no_col_per_adv,
                     act_collisions, act_relaxation, no_iter_impl_mass, g_set):
    eng.set_params(solute=solute_type, no_iter_impl_mass=no_iter_impl_mass, scale_dt_cond=scale_dt_cond,
                   no_col_per_adv=no_col_per_adv, act_collisions=act_collisions, act_relaxation=act_relaxation,
                   dt=dt, g_set=g_set, seed=rng.get_seed())
    e = eng.e
    e.dt_sub, e.dt_sub_half, e.dt_col_over_dV = float(dt_sub), float(dt_sub_half), float(dt_col_over_dV)


def _set_kernel(eng, E_col_grid, R_kernel_low_log, bin_factor_R_log):
    if isinstance(E_col_grid, float):
        eng.set_collision_kernel(E_col_grid)
    elif E_col_grid is None:
        eng.set_collision_kernel(0.5)
    else:
        eng.set_collision_kernel(E_col_grid, R_kernel_low_log, bin_factor_R_log)


def _on_device(a):
    """Whether a drop-in argument is a CUDA tensor (the device path: no host copies of the SIP arrays)."""
    import torch
    return isinstance(a, torch.Tensor) and a.device.type == "cuda"


def _upload(eng, grid_scalar_fields, grid_mat_prop, grid_velocity, grid_mass_flux_air_dry, pos, vel, cells, m_w,
            m_s, xi, active_ids, no_cols, water_removed):
    """The step's inputs into the engine: NumPy arrays through the host layout, or CUDA tensors of the
    reference's shapes and dtypes reordered on the device (Engine.load_particles_device)."""
    shape = (2, eng.nx + 1, eng.nz + 1)
    eng.set_flow(grid_velocity if grid_velocity is not None else np.zeros(shape),
                 grid_mass_flux_air_dry if grid_mass_flux_air_dry is not None else np.zeros(shape))
    eng.set_grid(0, grid_scalar_fields, grid_mat_prop)
    eng.init_cells()
    if _on_device(xi):
        eng.load_particles_device(0, pos, vel, cells, m_w, m_s, xi, active_ids)
    else:
        eng.load_particles([(pos, vel, cells, m_w, m_s, xi, active_ids)])
    eng.set_counters(0, no_cols, water_removed)
    eng.e.seed = rng.get_seed()
    eng.col_step = rng.get_step()


def _rel_pos(pos, cells, rel_pos, active_ids, grid_ranges, grid_steps):
    """integration.py:509-550 for the active SIPs."""
    a = np.asarray(active_ids, dtype=bool)
    for d in (0, 1):
        rel_pos[d][a] = (pos[d][a] - grid_ranges[d][0]) / grid_steps[d] - cells[d][a]


def _download(eng, grid_scalar_fields, grid_mat_prop, pos, vel, cells, rel_pos, m_w, m_s, xi, active_ids, T_p,
              delta_m_l, no_cols, water_removed, grid_ranges, grid_steps, particles_moved=True):
    if _on_device(xi):
        _download_device(eng, grid_scalar_fields, grid_mat_prop, pos, vel, cells, rel_pos, m_w, m_s, xi,
                         active_ids, T_p, delta_m_l, no_cols, water_removed, particles_moved)
        return
    p, v, c, mw, ms, x, a = eng.export_particles(0)
    m_w[...] = mw
    xi[...] = x
    m_s[...] = ms
    if particles_moved:
        pos[...] = p
        vel[...] = v
        cells[...] = c
        active_ids[...] = a
        if T_p is not None:
            T_p[...] = eng.export_T_p(0)
        if rel_pos is not None:
            _rel_pos(pos, cells, rel_pos, active_ids, grid_ranges, grid_steps)
    if grid_scalar_fields is not None:
        s, m = eng.export_grid(0)
        grid_scalar_fields[...] = s
        grid_mat_prop[...] = m
    if delta_m_l is not None:
        delta_m_l[...] = eng.export_increments(0)[2]
    nc, wr = eng.counters(0)
    no_cols[...] = nc
    if water_removed is not None:
        water_removed[0] = wr[0]
    rng.set_step(eng.col_step)


def _download_device(eng, grid_scalar_fields, grid_mat_prop, pos, vel, cells, rel_pos, m_w, m_s, xi, active_ids, T_p,
                     delta_m_l, no_cols, water_removed, particles_moved):
    """_download into CUDA tensors: the SIP arrays scattered back on the device (Engine.store_particles_device),
    the grid, increments and counters copied device to device (a host array among them is filled from the
    device)."""
    import torch
    eng.check()
    moved = particles_moved
    eng.store_particles_device(0, pos, vel, cells, m_w, m_s, xi, active_ids,
                               rel_pos=rel_pos if moved and rel_pos is not None else None,
                               T_p=T_p if moved and T_p is not None else None)
    if grid_scalar_fields is not None:
        eng.export_grid_into(0, grid_scalar_fields, grid_mat_prop)
    if delta_m_l is not None:
        src = eng.buf["dml"][:eng.nc].reshape(eng.nx, eng.nz)
        if isinstance(delta_m_l, torch.Tensor):
            delta_m_l.copy_(src)
        else:
            delta_m_l[...] = src.cpu().numpy()
    for dst, src in ((no_cols, eng.buf["no_cols"][0]), (water_removed, eng.buf["water_removed"][0:1])):
        if dst is None:
            continue
        if isinstance(dst, torch.Tensor):
            dst.copy_(src.reshape(dst.shape))
        else:
            dst[...] = src.cpu().numpy().reshape(np.shape(dst))
    eng.check()
    rng.set_step(eng.col_step)


# ------------------------------------------------------------------ helpers
def compute_relaxation_time_profile(z):
    """integration.py:291-311: relaxation time 300 s * exp(z / 200 m) (ICMW 2012 case 1)."""
    return 300 * np.exp(z / 200)


def update_cells_and_rel_pos(pos, cells, rel_pos, active_ids, grid_ranges, grid_steps):
    """integration.py:509-550 (set-up use by simulate; the time step does this on the GPU).

    cells and rel_pos of the active SIPs from their positions; rel_pos of
    inactive SIPs is left unchanged.
    """
    a = np.asarray(active_ids, dtype=bool)
    for d in (0, 1):
        cells[d][a] = np.floor((pos[d][a]) / grid_steps[d]).astype(np.int64)
    _rel_pos(pos, cells, rel_pos, active_ids, grid_ranges, grid_steps)


def compute_divergence_upwind_np(field, flux_field, grid_no_cells, grid_steps,
                                 boundary_conditions=np.array([0, 1])):
    """integration.py:150-286 on the GPU (lm_divergence_upwind).

    Limited third-order upwind divergence of field * flux; boundary type per
    axis 0 = periodic, 1 = fixed.  Returns a new (Nx, Nz) array (a CUDA tensor for a CUDA-tensor field).
    """
    import ctypes as C
    import torch
    from ._native import check, lib
    nx, nz = int(grid_no_cells[0]), int(grid_no_cells[1])
    on_dev = _on_device(field)

    def dev(a):
        if _on_device(a):
            return a.to(torch.float64).contiguous()
        return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()

    f, F = dev(field), dev(flux_field)
    if tuple(f.shape) != (nx, nz) or tuple(F.shape) != (2, nx + 1, nz + 1):
        raise ValueError("field must be (Nx, Nz) and flux_field (2, Nx+1, Nz+1)")
    out = torch.empty_like(f)
    bc = [int(b) for b in boundary_conditions]
    st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    check(lib().lm_divergence_upwind(C.c_void_p(f.data_ptr()), C.c_void_p(F.data_ptr()), C.c_void_p(out.data_ptr()),
                                     1, nx, nz, float(grid_steps[0]), float(grid_steps[1]), bc[0], bc[1], st))
    return out if on_dev else out.cpu().numpy()


compute_divergence_upwind = compute_divergence_upwind_np


# ------------------------------------------------------------ substeps
def propagate_grid_subloop_step_np(grid_scalar_fields, grid_mat_prop, p_ref, p_ref_inv, delta_Theta_ad,
                                   delta_r_v_ad, delta_m_l, grid_volume_cell):
    """integration.py:402-483: one grid substep (advection increments, latent heat, material properties)."""
    nx, nz = grid_scalar_fields.shape[1], grid_scalar_fields.shape[2]
    eng = _engine((nx, nz), [[0.0, float(nx)], [0.0, float(nz)]], [1.0, 1.0], grid_volume_cell, p_ref, p_ref_inv, 0)
    eng.set_grid(0, grid_scalar_fields, grid_mat_prop)
    eng.init_cells()
    eng.set_increments(0, delta_Theta_ad, delta_r_v_ad, delta_m_l)
    eng.grid_substep()
    eng.check()
    eng.export_grid_into(0, grid_scalar_fields, grid_mat_prop)


propagate_grid_subloop_step = propagate_grid_subloop_step_np


def propagate_particles_subloop_step_np(grid_scalar_fields, grid_mat_prop, grid_velocity, grid_no_cells,
                                        grid_ranges, grid_steps, pos, vel, cells, rel_pos, m_w, m_s, xi,
                                        water_removed, id_list, active_ids, T_p, solute_type, delta_m_l, dt_sub,
                                        dt_sub_pos, no_iter_impl_mass, g_set):
    """integration.py:944-1089: T_p, implicit condensation with Delta m_l, drag, positions and cells."""
    N = int(xi.shape[0])
    _check_ids(id_list, N)
    ranges, steps = np.asarray(grid_ranges), np.asarray(grid_steps)
    eng = _engine(grid_no_cells, ranges, steps, steps[0] * steps[1], 1.0E5, 1.0E-5, N)
    _set_step_params(eng, solute_type, dt_sub, dt_sub, 0.5 * dt_sub, 0.0, 1, 1, False, False, no_iter_impl_mass,
                     g_set)
    no_cols = np.zeros(2, dtype=np.int64)
    _upload(eng, grid_scalar_fields, grid_mat_prop, grid_velocity, None, pos, vel, cells, m_w, m_s, xi, active_ids,
            no_cols, water_removed)
    eng.particle_substep(dt_sub, dt_sub_pos)
    _download(eng, None, None, pos, vel, cells, rel_pos, m_w, m_s, xi, active_ids, T_p, delta_m_l, no_cols,
              water_removed, ranges, steps)


propagate_particles_subloop_step = propagate_particles_subloop_step_np


def integrate_subloop_n_steps_np(grid_scalar_fields, grid_mat_prop, grid_velocity, grid_no_cells, grid_ranges,
                                 grid_steps, grid_volume_cell, p_ref, p_ref_inv, pos, vel, cells, rel_pos, m_w, m_s,
                                 xi, solute_type, dt_col_over_dV, E_col_grid, no_kernel_bins, R_kernel_low_log,
                                 bin_factor_R_log, no_cols, water_removed, id_list, active_ids, T_p, delta_m_l,
                                 delta_Theta_ad, delta_r_v_ad, dt_sub, dt_sub_pos, no_cond_steps, no_col_steps,
                                 no_iter_impl_mass, g_set, act_collisions):
    """integration.py:1095-1294: one subloop of no_cond_steps substeps with its collision steps."""
    N = int(xi.shape[0])
    _check_ids(id_list, N)
    eng = _engine(grid_no_cells, grid_ranges, grid_steps, grid_volume_cell, p_ref, p_ref_inv, N)
    _set_step_params(eng, solute_type, 2 * dt_sub, dt_sub, 0.5 * dt_sub, dt_col_over_dV, 1, 1, act_collisions,
                     False, no_iter_impl_mass, g_set)
    _set_kernel(eng, E_col_grid, R_kernel_low_log, bin_factor_R_log)
    _upload(eng, grid_scalar_fields, grid_mat_prop, grid_velocity, None, pos, vel, cells, m_w, m_s, xi, active_ids,
            no_cols, water_removed)
    eng.set_increments(0, delta_Theta_ad, delta_r_v_ad)
    moved = no_cond_steps > 0
    if act_collisions and no_col_steps == 1:
        eng.collision_step()
    for _ in range(no_cond_steps):
        if act_collisions and no_col_steps > 1:
            eng.collision_step()
        eng.substep(dt_sub, dt_sub_pos)
    if act_collisions and no_col_steps > no_cond_steps:
        eng.collision_step()
    _download(eng, grid_scalar_fields, grid_mat_prop, pos, vel, cells, rel_pos, m_w, m_s, xi, active_ids,
              T_p if moved else None, delta_m_l if moved else None, no_cols, water_removed, grid_ranges, grid_steps,
              particles_moved=moved)


integrate_subloop_n_steps = integrate_subloop_n_steps_np


def integrate_adv_step_np(grid_scalar_fields, grid_mat_prop, grid_velocity, grid_mass_flux_air_dry, p_ref,
                          p_ref_inv, grid_no_cells, grid_ranges, grid_steps, grid_volume_cell, pos, vel, cells,
                          rel_pos, m_w, m_s, xi, solute_type, water_removed, id_list, active_ids, T_p, delta_m_l, dt,
                          dt_sub, dt_sub_half, dt_col_over_dV, scale_dt_cond, no_col_per_adv, act_collisions,
                          act_relaxation, init_profile_r_v, init_profile_Theta, relaxation_time_profile,
                          no_iter_impl_mass, g_set, E_col_grid, no_kernel_bins, R_kernel_low_log, bin_factor_R_log,
                          no_cols):
    """integration.py:1298-1611: one advection step of particles and grid (lm_adv_steps)."""
    N = int(xi.shape[0])
    _check_ids(id_list, N)
    eng = _engine(grid_no_cells, grid_ranges, grid_steps, grid_volume_cell, p_ref, p_ref_inv, N)
    _set_step_params(eng, solute_type, dt, dt_sub, dt_sub_half, dt_col_over_dV, scale_dt_cond, no_col_per_adv,
                     act_collisions, act_relaxation, no_iter_impl_mass, g_set)
    _set_kernel(eng, E_col_grid, R_kernel_low_log, bin_factor_R_log)
    if act_relaxation:
        eng.set_relaxation(0, init_profile_r_v, init_profile_Theta, relaxation_time_profile)
    _upload(eng, grid_scalar_fields, grid_mat_prop, grid_velocity, grid_mass_flux_air_dry, pos, vel, cells, m_w,
            m_s, xi, active_ids, no_cols, water_removed)
    eng.adv_steps(1)
    _download(eng, grid_scalar_fields, grid_mat_prop, pos, vel, cells, rel_pos, m_w, m_s, xi, active_ids, T_p,
              delta_m_l, no_cols, water_removed, grid_ranges, grid_steps)


integrate_adv_step = integrate_adv_step_np


# ------------------------------------------------------------ simulation
class _Resident:
    """One run held in HBM across many advection steps (used by simulate / simulate_interval)."""

    def __init__(self, eng):
        self.eng = eng
        self.stepped = False
        self._inv = None

    def steps(self, n):
        if n > 0:
            self.eng.adv_steps(n)
            self.stepped = True
            self._inv = None

    def traced(self, trace_ids):
        """(pos (2,n), vel (2,n), m_w, m_s, T_p, xi) of the traced IDs, gathered on the device."""
        import torch
        eng = self.eng
        eng.check()
        cur = eng.e.cur
        N = eng.n_tot
        ids = eng.buf[f"id{cur}"][:N].long()
        if self._inv is None:
            inv = torch.empty(N, dtype=torch.long, device=eng.device)
            inv[ids] = torch.arange(N, device=eng.device)
            self._inv = inv
        tid = np.asarray(trace_ids, dtype=np.int64)
        if np.any((tid < 0) | (tid >= N)):
            raise IndexError("trace_ids outside 0 .. N-1")
        t = torch.as_tensor(tid, device=eng.device)
        sl = self._inv[t]
        g = {f: eng.buf[f"{f}{cur}"][sl].cpu().numpy() for f in ("x", "z", "u", "w", "m_w", "m_s", "xi")}
        if self.stepped:
            ccur = eng.e.ccur
            prev = eng.buf[f"cell{1 - ccur}"][sl].long()
            loc = torch.where(prev >= 0, prev, (-prev - 2) * eng.nz + (eng.nz - 1))
            T_p = eng.buf["cc"][loc, CC_TP].cpu().numpy()
        else:
            T_p = np.ones(len(sl))
        return (np.array([g["x"], g["z"]]), np.array([g["u"], g["w"]]), g["m_w"], g["m_s"], T_p, g["xi"])


def _run_interval(res, no_adv_steps, dump_every, trace_ids, traced_vectors, traced_scalars, traced_xi,
                  traced_water):
    """The loop of integration.py:1813-1842 on a resident run."""
    dump_N = 0
    cnt = 0
    while cnt < no_adv_steps:
        if cnt % dump_every == 0:
            p, v, mw, ms, tp, x = res.traced(trace_ids)
            traced_vectors[dump_N, 0] = p
            traced_vectors[dump_N, 1] = v
            traced_scalars[dump_N, 0] = mw
            traced_scalars[dump_N, 1] = ms
            traced_scalars[dump_N, 2] = tp
            traced_xi[dump_N] = x
            traced_water[dump_N] = res.eng.counters(0)[1][0]
            dump_N += 1
        n = min(dump_every - cnt % dump_every, no_adv_steps - cnt)
        res.steps(n)
        cnt += n


def col_steps_per_adv(no_col_per_adv, scale_dt_cond, act_collisions=True):
    """Collision steps of one advection step as integrate_adv_step_np places them.

    Subloop 1 runs scale_dt_cond substeps with 0, 1 or scale_dt_cond collision steps
    (integration.py:1515-1519), subloop 2 runs scale_dt_cond - 1 substeps with 1 or scale_dt_cond
    (:1570-1572); a subloop with no_col collision steps over no_cond substeps performs one at its start if
    no_col == 1, one before each substep if no_col > 1, and one more at its end if no_col > no_cond
    (:1247-1292).  With scale_dt_cond == 1 the second subloop has no substep, so its single collision step
    is counted twice (start and end), as the reference runs it.
    """
    if not act_collisions:
        return 0

    def subloop(no_cond, no_col):
        return int(no_col == 1) + (no_cond if no_col > 1 else 0) + int(no_col > no_cond)

    sdc, npa = int(scale_dt_cond), int(no_col_per_adv)
    n1 = 0 if npa == 1 else (1 if npa == 2 else sdc)
    n2 = 1 if npa in (1, 2) else sdc
    return subloop(sdc, n1) + subloop(sdc - 1, n2)


def simulate_interval(grid_scalar_fields, grid_mat_prop, grid_velocity, grid_mass_flux_air_dry, p_ref, p_ref_inv,
                      grid_no_cells, grid_ranges, grid_steps, grid_volume_cell, pos, vel, cells, rel_pos, m_w, m_s,
                      xi, solute_type, water_removed, id_list, active_ids, T_p, delta_m_l, dt, dt_sub, dt_sub_half,
                      dt_col, scale_dt_cond, no_col_per_adv, no_adv_steps, act_collisions, act_relaxation,
                      init_profile_r_v, init_profile_Theta, relaxation_time_profile, no_iter_impl_mass, g_set,
                      dump_every, trace_ids, traced_vectors, traced_scalars, traced_xi, traced_water, E_col_grid,
                      no_kernel_bins, R_kernel_low_log, bin_factor_R_log, no_cols):
    """integration.py:1615-1842: no_adv_steps advection steps with tracer dumps every dump_every steps.

    The state stays on the GPU for the whole interval; all arguments are
    updated in place at its end.  T_p of the first dump is the caller's T_p
    if no step has run yet.
    """
    N = int(xi.shape[0])
    _check_ids(id_list, N)
    eng = _engine(grid_no_cells, grid_ranges, grid_steps, grid_volume_cell, p_ref, p_ref_inv, N)
    _set_step_params(eng, solute_type, dt, dt_sub, dt_sub_half, dt_col / grid_volume_cell, scale_dt_cond,
                     no_col_per_adv, act_collisions, act_relaxation, no_iter_impl_mass, g_set)
    _set_kernel(eng, E_col_grid, R_kernel_low_log, bin_factor_R_log)
    if act_relaxation:
        eng.set_relaxation(0, init_profile_r_v, init_profile_Theta, relaxation_time_profile)
    _upload(eng, grid_scalar_fields, grid_mat_prop, grid_velocity, grid_mass_flux_air_dry, pos, vel, cells, m_w,
            m_s, xi, active_ids, no_cols, water_removed)
    res = _Resident(eng)
    T_p0 = np.array(T_p, dtype=np.float64)
    _run_interval(res, no_adv_steps, dump_every, trace_ids, traced_vectors, traced_scalars, traced_xi, traced_water)
    if no_adv_steps > 0 and len(trace_ids) > 0:
        traced_scalars[0, 2] = T_p0[trace_ids]
    _download(eng, grid_scalar_fields, grid_mat_prop, pos, vel, cells, rel_pos, m_w, m_s, xi, active_ids,
              T_p if res.stepped else None, delta_m_l if res.stepped else None, no_cols, water_removed,
              grid_ranges, grid_steps, particles_moved=res.stepped)


def _update_grid_r_l(eng, grid_scalar_fields):
    """grid.py:370-412 on the GPU; refreshes r_l in grid_scalar_fields from the resident SIPs."""
    eng.update_r_l()
    s, _ = eng.export_grid(0)
    grid_scalar_fields[5] = s[5]


def simulate(grid, pos, vel, cells, m_w, m_s, xi, active_ids, water_removed, no_cols, config, E_col_grid):
    """integration.py:1846-2147: integrate from config['t_start'] to config['t_end'] with frame and tracer output.

    Writes the reference's files into config['paths']['output'] (grid scalar
    fields, all-particle dumps and collision counters every frame_every
    steps, tracer blocks every dump_every steps, the full grid + particle
    state and water_removed at t_end) and updates the arguments in place.
    The collision draws are keyed by config['seed_sim']; the step counter
    starts at the number of collision steps before t_start, so a run resumed
    from a saved state continues the stream of the uninterrupted run.
    """
    par_keys = ['solute_type', 'dt_adv', 'dt_col', 'scale_dt_cond', 'no_col_per_adv', 't_start', 't_end',
                'no_iter_impl_mass', 'g_set', 'act_collisions', 'act_relaxation', 'frame_every', 'dump_every',
                'trace_ids', 'no_kernel_bins', 'R_kernel_low_log', 'bin_factor_R_log', 'kernel_type',
                'kernel_method', 'seed_sim', 'simulation_mode']
    (solute_type, dt, dt_col, scale_dt_cond, no_col_per_adv, t_start, t_end, no_iter_impl_mass, g_set,
     act_collisions, act_relaxation, frame_every, dump_every, trace_ids, no_kernel_bins, R_kernel_low_log,
     bin_factor_R_log, kernel_type, kernel_method, seed_sim, simulation_mode) = [config.get(k) for k in par_keys]
    output_path = config['paths']['output']
    log_file = output_path + f"log_sim_t_{int(t_start)}_{int(t_end)}.txt"
    start_time = datetime.now()

    rel_pos = np.zeros_like(pos)
    update_cells_and_rel_pos(pos, cells, rel_pos, active_ids, grid.ranges, grid.steps)
    T_p = np.ones_like(m_w)
    id_list = np.arange(xi.shape[0])

    grid.update_material_properties()
    V0_inv = 1.0 / grid.volume_cell
    grid.rho_dry_inv = np.ones_like(grid.mass_density_air_dry) / grid.mass_density_air_dry
    grid.mass_dry_inv = V0_inv * grid.rho_dry_inv
    delta_m_l = np.zeros_like(grid.temperature)

    if act_relaxation:
        scalars0 = np.load(config['paths']['init'] + "arr_file1_" + str(config['t_init']) + ".npy")
        init_profile_r_v = np.average(scalars0[3], axis=0)
        init_profile_Theta = np.average(scalars0[7], axis=0)
        relaxation_time_profile = compute_relaxation_time_profile(grid.centers[1][0])
    else:
        init_profile_r_v = init_profile_Theta = relaxation_time_profile = None

    grid_scalar_fields = np.array((grid.temperature, grid.pressure, grid.potential_temperature,
                                   grid.mass_density_air_dry, grid.mixing_ratio_water_vapor,
                                   grid.mixing_ratio_water_liquid, grid.saturation, grid.saturation_pressure,
                                   grid.mass_dry_inv, grid.rho_dry_inv))
    grid_mat_prop = np.array((grid.thermal_conductivity, grid.diffusion_constant, grid.heat_of_vaporization,
                              grid.surface_tension, grid.specific_heat_capacity, grid.viscosity,
                              grid.mass_density_fluid))

    dt_sub = dt / (2 * scale_dt_cond)
    dt_sub_half = 0.5 * dt_sub
    print("dt = ", dt)
    print("dt_col = ", dt_col)
    print("dt_sub = ", dt_sub)
    with open(log_file, "w+") as f:
        f.write(f"simulation mode = {simulation_mode}\n")
        f.write(f"gravitation const = {g_set}\n")
        f.write(f"collisions activated = {act_collisions}\n")
        f.write(f"relaxation activated = {act_relaxation}\n")
        f.write(f"kernel_type = {kernel_type}\n")
        f.write(f"kernel_method = {kernel_method}")
        if kernel_method == "Ecol_const":
            f.write(f", E_col = {E_col_grid}")
        f.write(f"\nsolute material = {solute_type}\n")
        f.write(f"dt = {dt}\n")
        f.write(f"dt_col = {dt_col}\n")
        f.write(f"dt_sub = {dt_sub}\n")
    cnt_max = (t_end - t_start) / dt
    no_grid_frames = int(math.ceil(cnt_max / frame_every))
    np.save(output_path + "data_saving_paras.npy", (frame_every, no_grid_frames, dump_every))
    dt_save = int(frame_every * dt)
    grid_save_times = np.arange(t_start, t_end + 0.5 * dt_save, dt_save).astype(int)
    np.save(output_path + "grid_save_times.npy", grid_save_times)
    if isinstance(trace_ids, int):
        trace_id_dist = int(math.floor(len(xi) / trace_ids))
        trace_ids = np.arange(int(trace_id_dist * 0.5), len(xi), trace_id_dist)
    np.save(output_path + "trace_ids.npy", trace_ids)
    no_trace_ids = len(trace_ids)
    dump_factor = frame_every // dump_every
    with open(log_file, "a") as f:
        f.write("frame_every, no_grid_frames\n")
        f.write(f"{frame_every} {no_grid_frames}\n")
        f.write("dump_every, dump_factor\n")
        f.write(f"{dump_every} {dump_factor}\n")
    traced_vectors = np.zeros((dump_factor, 2, 2, no_trace_ids))
    traced_scalars = np.zeros((dump_factor, 3, no_trace_ids))
    traced_xi = np.zeros((dump_factor, no_trace_ids))
    traced_water = np.zeros(dump_factor)
    fh.save_sim_paras_to_file([dt, dt_sub, no_iter_impl_mass, seed_sim], "dt dt_sub no_iter_impl_mass rnd_seed_sim",
                              t_start, output_path)
    date = datetime.now()
    print("### simulation starts ###")
    with open(log_file, "a") as f:
        f.write("### simulation starts ###\n")
        f.write(f"start date and time = {date}\n")
        f.write(f"sim time = {date - start_time}\n")

    N = len(xi)
    eng = _engine(grid.no_cells, grid.ranges, grid.steps, grid.volume_cell, grid.p_ref, grid.p_ref_inv, N)
    if act_collisions:
        rng.seed(seed_sim, int(round(t_start / dt)) * col_steps_per_adv(no_col_per_adv, scale_dt_cond))
    _set_step_params(eng, solute_type, dt, dt_sub, dt_sub_half, dt_col / grid.volume_cell, scale_dt_cond,
                     no_col_per_adv, act_collisions, act_relaxation, no_iter_impl_mass, g_set)
    _set_kernel(eng, E_col_grid, R_kernel_low_log, bin_factor_R_log)
    if act_relaxation:
        eng.set_relaxation(0, init_profile_r_v, init_profile_Theta, relaxation_time_profile)
    _upload(eng, grid_scalar_fields, grid_mat_prop, grid.velocity, grid.mass_flux_air_dry, pos, vel, cells, m_w,
            m_s, xi, active_ids, no_cols, water_removed)
    res = _Resident(eng)

    def pull():
        _download(eng, grid_scalar_fields, grid_mat_prop, pos, vel, cells, rel_pos, m_w, m_s, xi, active_ids,
                  T_p if res.stepped else None, delta_m_l if res.stepped else None, no_cols, water_removed,
                  grid.ranges, grid.steps, particles_moved=True)

    for frame_N in range(no_grid_frames):
        t = t_start + frame_N * frame_every * dt
        eng.update_r_l()
        pull()
        fh.save_grid_scalar_fields(t, grid_scalar_fields, output_path, start_time)
        fh.dump_particle_data_all(t, pos, vel, cells, m_w, m_s, xi, active_ids, output_path)
        np.save(output_path + f"no_cols_{int(t)}.npy", no_cols)
        _run_interval(res, frame_every, dump_every, trace_ids, traced_vectors, traced_scalars, traced_xi,
                      traced_water)
        time_block = np.arange(t, t + frame_every * dt, dump_every * dt).astype(int)
        fh.dump_particle_tracer_data_block(time_block, traced_vectors, traced_scalars, traced_xi, traced_water,
                                           output_path)

    t = t_start + no_grid_frames * frame_every * dt
    eng.update_r_l()
    pull()
    fh.save_grid_scalar_fields(t, grid_scalar_fields, output_path, start_time)
    fh.dump_particle_data_all(t, pos, vel, cells, m_w, m_s, xi, active_ids, output_path)
    np.save(output_path + f"no_cols_{int(t)}.npy", no_cols)
    fh.dump_particle_data(t, pos[:, trace_ids], vel[:, trace_ids], m_w[trace_ids], m_s[trace_ids], xi[trace_ids],
                          grid_scalar_fields[0], grid_scalar_fields[4], output_path)
    np.save(output_path + f"water_removed_{int(t)}", water_removed)

    grid.temperature = grid_scalar_fields[0]
    grid.pressure = grid_scalar_fields[1]
    grid.potential_temperature = grid_scalar_fields[2]
    grid.mass_density_air_dry = grid_scalar_fields[3]
    grid.mixing_ratio_water_vapor = grid_scalar_fields[4]
    grid.mixing_ratio_water_liquid = grid_scalar_fields[5]
    grid.saturation = grid_scalar_fields[6]
    grid.saturation_pressure = grid_scalar_fields[7]
    grid.thermal_conductivity = grid_mat_prop[0]
    grid.diffusion_constant = grid_mat_prop[1]
    grid.heat_of_vaporization = grid_mat_prop[2]
    grid.surface_tension = grid_mat_prop[3]
    grid.specific_heat_capacity = grid_mat_prop[4]
    grid.viscosity = grid_mat_prop[5]
    grid.mass_density_fluid = grid_mat_prop[6]
    fh.save_grid_and_particles_full(t, grid, pos, cells, vel, m_w, m_s, xi, active_ids, output_path)

    end_time = datetime.now()
    print("### simulation ended ###")
    print("simulation time:", end_time - start_time)
    with open(log_file, "a") as f:
        f.write("### simulation ended ###\n")
        f.write(f"t_start = {t_start}\n")
        f.write(f"t_end = {t_end}\n")
        f.write(f"dt = {dt}; dt_sub = {dt_sub}\n")
        f.write(f"simulation time = {end_time - start_time}\n")
