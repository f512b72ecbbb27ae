"""Ensemble diagnostics on the GPU (evaluation.py of LacmoPy, the parts the SIP ensemble feeds).

``generate_field_frame_data_avg`` is the drop-in of evaluation.py:202-509: it reads each seed's frame
files, computes the derived fields of every frame on the GPU (lm_derived_fields: class mixing ratios and
numbers per dry mass, radius-moment ratios; ordered per-cell sums as in the reference) and the ensemble
mean and spread (lm_ensemble_stats).  ``Engine``-level helpers do the same for runs held in HBM without
files: ``derived_fields(eng)`` and ``ensemble_profiles(eng)`` (horizontal means per level, ensemble
mean and spread over the runs, and the surface precipitation flux).

The reference computes no precipitation flux; ``precipitation_flux`` defines it from the water that
left the domain through the ground (water_removed, integration.py:620-628): kg m^-2 s^-1 over the domain
width and the grid's y-depth.
"""
import ctypes as C

import numpy as np

from ._native import check, lib
from .file_handling import load_grid_scalar_fields, load_particle_data_all

FIELD_NAMES_ORIG = [r"r_\mathrm{v}", r"r_\mathrm{l}", r"\Theta", "T", "p", "S"]
SCALES_ORIG = [1000., 1000., 1, 1, 0.01, 1]
UNITS_ORIG = ["g/kg", "g/kg", "K", "K", "hPa", "-"]
FIELD_NAMES_DERI = [r"r_\mathrm{aero}", r"r_\mathrm{c}", r"r_\mathrm{r}", r"n_\mathrm{aero}", r"n_\mathrm{c}",
                    r"n_\mathrm{r}", r"R_\mathrm{avg}", r"R_{2/1}", r"R_\mathrm{eff}"]
UNITS_DERI = ["g/kg", "g/kg", "g/kg", "1/mg", "1/mg", "1/mg", r"$\mathrm{\mu m}$", r"$\mathrm{\mu m}$",
              r"$\mathrm{\mu m}$"]
SCALES_DERI = [1000., 1000., 1000., 1E-6, 1E-6, 1E-6, 1., 1., 1.]


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("lacmopy_amd diagnostics need a GPU (HIP device); there is no CPU fallback")
    return torch


def derived_fields(eng):
    """(n_runs, 9, nx, nz) derived fields of the engine's current state (lm_derived_fields)."""
    torch = _torch()
    out = torch.empty((eng.n_runs, 9, eng.nx, eng.nz), dtype=torch.float64, device=eng.device)
    check(eng.L.lm_derived_fields(C.byref(eng.e), C.c_void_p(out.data_ptr()), eng.stream))
    eng.check()
    return out.cpu().numpy()


def ensemble_stats(fields):
    """Mean and spread over the first axis (seeds) as evaluation.py:443-474, on the GPU."""
    torch = _torch()
    f = torch.from_numpy(np.ascontiguousarray(fields, dtype=np.float64)).cuda()
    n = f.shape[0]
    flat = f.reshape(n, -1)
    mean = torch.empty(flat.shape[1], dtype=torch.float64, device=f.device)
    std = torch.empty_like(mean)
    st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    check(lib().lm_ensemble_stats(C.c_void_p(flat.data_ptr()), n, flat.shape[1], C.c_void_p(mean.data_ptr()),
                                  C.c_void_p(std.data_ptr()), st))
    shape = tuple(fields.shape[1:])
    return mean.cpu().numpy().reshape(shape), std.cpu().numpy().reshape(shape)


def precipitation_flux(water_removed_t0, water_removed_t1, dt, domain_width, dy=1.0):
    """Surface precipitation flux (kg m^-2 s^-1) between two times from the removed water (1e-18 kg units)."""
    return 1.0E-18 * (np.asarray(water_removed_t1) - np.asarray(water_removed_t0)) / (dt * domain_width * dy)


def ensemble_profiles(eng, water_removed_prev=None, dt=None):
    """Horizontal means per level of the derived fields and of r_v, r_l, Theta, T (from the grid), ensemble
    mean and spread over the engine's runs; plus the precipitation flux per run if the previous
    water_removed and the elapsed time are given.

    Returns dict: 'derived_mean', 'derived_std' (9, nz); 'grid_mean', 'grid_std' (4, nz) for
    (r_v, r_l, Theta, T); optionally 'precipitation' (n_runs,) and its mean and spread."""
    d = derived_fields(eng)                       # (R, 9, nx, nz)
    prof_d = d.mean(axis=2)                       # (R, 9, nz)
    g = np.stack([eng.export_grid(r)[0] for r in range(eng.n_runs)])   # (R, 10, nx, nz)
    prof_g = g[:, [4, 5, 2, 0]].mean(axis=2)      # (R, 4, nz)
    out = {}
    out["derived_mean"], out["derived_std"] = ensemble_stats(prof_d)
    out["grid_mean"], out["grid_std"] = ensemble_stats(prof_g)
    if water_removed_prev is not None and dt is not None:
        wr = np.array([eng.counters(r)[1][0] for r in range(eng.n_runs)])
        width = float(eng.ranges[0, 1] - eng.ranges[0, 0])
        dy = float(eng.e.V_cell / (eng.e.dx * eng.e.dz))
        flux = precipitation_flux(np.asarray(water_removed_prev), wr, dt, width, dy)
        out["precipitation"] = flux
        m, s = ensemble_stats(flux[:, None])
        out["precipitation_mean"], out["precipitation_std"] = m[0], s[0]
    return out


def _frame_engine(nx, nz, N, grid_volume_cell, solute_type):
    from .engine import Engine
    eng = Engine(nx, nz, [N], [[0.0, float(nx)], [0.0, float(nz)]], [1.0, 1.0], float(grid_volume_cell))
    eng.set_params(solute=solute_type, act_collisions=False, act_relaxation=False)
    return eng


def generate_field_frame_data_avg(load_path_list, field_indices, time_indices, derived_indices, mass_dry_inv,
                                  grid_volume_cell, no_cells, solute_type):
    """evaluation.py:202-509: ensemble mean and spread of grid fields and derived fields per frame.

    Same arguments, return values and file inputs as the reference; the derived fields of every
    seed and frame and the ensemble statistics run on the GPU."""
    nx, nz = int(no_cells[0]), int(no_cells[1])
    field_indices = np.asarray(field_indices, dtype=np.int64)
    time_indices = np.asarray(time_indices, dtype=np.int64)
    derived_indices = np.asarray(derived_indices, dtype=np.int64)
    no_seeds, no_times = len(load_path_list), len(time_indices)
    no_orig, no_deri = len(field_indices), len(derived_indices)
    load_path = load_path_list[0]
    grid_save_times = np.load(load_path + "grid_save_times.npy")
    save_times_out = np.array([grid_save_times[t] for t in time_indices], dtype=np.int64)
    field_names_out = [FIELD_NAMES_ORIG[k] for k in field_indices] + [FIELD_NAMES_DERI[k] for k in derived_indices]
    units_out = [UNITS_ORIG[k] for k in field_indices] + [UNITS_DERI[k] for k in derived_indices]
    scales_out = [SCALES_ORIG[k] for k in field_indices] + [SCALES_DERI[k] for k in derived_indices]
    per_seed = np.zeros((no_seeds, no_times, no_orig + no_deri, nx, nz))
    for s_n, path in enumerate(load_path_list):
        fields = load_grid_scalar_fields(path, grid_save_times)
        vec, cells_t, scal_t, xi_t, act_t = load_particle_data_all(path, grid_save_times)
        per_seed[s_n, :, :no_orig] = fields[time_indices][:, field_indices]
        for t_n, it in enumerate(time_indices):
            N = len(xi_t[it])
            eng = _frame_engine(nx, nz, N, grid_volume_cell, solute_type)
            scal = np.ones((10, nx, nz))
            scal[0] = fields[it, 3]
            scal[1] = 1.0E5
            scal[8] = mass_dry_inv
            eng.set_grid(0, scal, np.ones((7, nx, nz)))
            eng.init_cells()
            eng.load_particles([(vec[it, 0], vec[it, 1], cells_t[it], scal_t[it, 0], scal_t[it, 1], xi_t[it],
                                 act_t[it])])
            d = derived_fields(eng)[0]
            per_seed[s_n, t_n, no_orig:] = d[derived_indices]
    mean, std = ensemble_stats(per_seed)
    return mean, std, save_times_out, field_names_out, units_out, scales_out
