"""TEST INFRASTRUCTURE ONLY: ctypes wrapper of the CPU oracle (lacmo_oracle.c).

Used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the
checker; the product never imports it.  All arrays are in the reference's
layout and are updated in place, like the reference functions.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_libs = {}

_d, _i64, _i32, _u32, _u64, _p = C.c_double, C.c_int64, C.c_int32, C.c_uint32, C.c_uint64, C.c_void_p


class OrAdvParams(C.Structure):
    _fields_ = [
        ("N", _i64), ("nx", _i64), ("nz", _i64),
        ("x_min", _d), ("x_max", _d), ("z_min", _d), ("z_max", _d), ("dx", _d), ("dz", _d),
        ("V_cell", _d), ("p_ref", _d), ("p_ref_inv", _d),
        ("dt", _d), ("dt_sub", _d), ("dt_sub_half", _d), ("dt_col_over_dV", _d),
        ("scale_dt_cond", C.c_int), ("no_col_per_adv", C.c_int), ("act_collisions", C.c_int),
        ("act_relaxation", C.c_int), ("no_iter", C.c_int), ("solute", C.c_int),
        ("g_set", _d), ("nb", _i64), ("lnR0", _d), ("lnf", _d), ("E_const", _d),
        ("seed", _u64), ("run", _u32), ("col_step", _u32), ("rnd_pos", _i64),
    ]


def lib(libm=False):
    name = "liblacmo_oracle_libm.so" if libm else "liblacmo_oracle.so"
    if name in _libs:
        return _libs[name]
    path = os.path.join(HERE, name)
    if not os.path.exists(path):
        raise RuntimeError(f"{path} missing: run `make -C oracle`")
    L = C.CDLL(path)
    L.or_kernel_index.restype = _i64
    L.or_kernel_index.argtypes = [_d, _d, _d, _i64]
    L.or_dml_impl_Newton.restype = _d
    L.or_dml_impl_Newton.argtypes = [C.c_int, _d, C.c_int] + [_d] * 14 + [_p]
    L.or_mass_rate_and_derivative.argtypes = [C.c_int] + [_d] * 14 + [_p, _p]
    L.or_terminal_velocity_Beard.restype = _d
    L.or_terminal_velocity_Beard.argtypes = [_d]
    L.or_radius_from_mass.restype = _d
    L.or_radius_from_mass.argtypes = [_d, _d]
    L.or_surface_tension.restype = _d
    L.or_surface_tension.argtypes = [C.c_int, _d, _d]
    L.or_pair_uniform.restype = _d
    L.or_pair_uniform.argtypes = [_u64, _u32, _u32, _u32, _u32, _u32]
    L.or_num_consts.restype = C.c_int
    L.or_interpolate_velocity.argtypes = [_i64, _i64, _d, _d, _p, _i64, _i64, _p, _p]
    L.or_collision_cell_2D.argtypes = [_i64, _p, _p, _p, _p, _p, _p, _p, _d, _p, _d, _i64, _d, _d, _p, _p,
                                       _u64, _u32, _u32, _u32]
    L.or_collision_all_cells.argtypes = [_i64, _p, _p, _p, _p, _p, _p, _i64, _i64, _d, _p, _d, _i64, _d, _d, _p,
                                         C.c_int, _p, _u64, _u32, _u32]
    L.or_collision_golovin.argtypes = [_i64, _p, _p, _d, _p, _p, _u64, _u32, _u32]
    L.or_collision_ecol_box.argtypes = [_i64, _p, _p, _p, _p, _p, _d, _p, _i64, _d, _d, _p, _p, _u64, _u32, _u32]
    L.or_update_velocity_Beard.argtypes = [_i64, _p, _p]
    L.or_collision_kernel_grid_m.argtypes = [_i64, _p, _p, _d, _p, _i64, _d, _d, _p, _p, _u64, _u32, _u32]
    L.or_update_grid_r_l.argtypes = [_i64, _p, _p, _p, _p, _i64, _i64, _p, _p]
    L.or_derived_fields.argtypes = [_i64, _p, _p, _p, _p, _p, _p, _p, _i64, _i64, _d, C.c_int, _p]
    L.or_ensemble_stats.argtypes = [_i64, _i64, _p, _p, _p]
    L.or_divergence_upwind.argtypes = [_p, _p, _i64, _i64, _d, _d, C.c_int, C.c_int, _p]
    L.or_relaxation_term.argtypes = [_p, _p, _p, _d, _i64, _i64, _p]
    L.or_propagate_grid_subloop_step.argtypes = [_p, _p, _d, _d, _p, _p, _p, _d, _i64]
    L.or_update_material_properties.argtypes = [_p, _p, _i64]
    L.or_integrate_adv_steps.argtypes = [C.POINTER(OrAdvParams), _i64] + [_p] * 20
    L.or_propagate_particles_substep.argtypes = [C.POINTER(OrAdvParams), _d, _d] + [_p] * 14
    L.or_update_pos_and_cells.argtypes = [C.POINTER(OrAdvParams), _d] + [_p] * 8
    L.or_lm_vec.argtypes = [C.c_int, _i64, _p, _p, _p]
    L.or_philox.argtypes = [_p, _p, _p]
    L.or_R_p_w_s_rho_p.argtypes = [_i64, _p, _p, _p, C.c_int, _p, _p, _p]
    _libs[name] = L
    return L


def P(a):
    """Pointer of a contiguous numpy array (or NULL for None)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "oracle needs contiguous arrays"
    return a.ctypes.data_as(C.c_void_p)


def consts(libm=False):
    L = lib(libm)
    n = L.or_num_consts()
    out = np.zeros(n)
    L.or_get_consts(P(out), C.c_int(n))
    return out


def lm_vec(which, x, y=None):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y if y is not None else np.zeros_like(x), dtype=np.float64)
    out = np.zeros_like(x)
    lib().or_lm_vec(which, x.size, P(x), P(y), P(out))
    return out


def adv_params(grid, N, config, seed=0, run=0, col_step=0, rnd_pos=0, E_col_grid=None):
    """OrAdvParams from a Grid and a config dict (keys as integration.py:1908-1923)."""
    p = OrAdvParams()
    p.N = N
    p.nx, p.nz = int(grid.no_cells[0]), int(grid.no_cells[1])
    p.x_min, p.x_max = grid.ranges[0]
    p.z_min, p.z_max = grid.ranges[1]
    p.dx, p.dz = grid.steps
    p.V_cell, p.p_ref, p.p_ref_inv = grid.volume_cell, grid.p_ref, grid.p_ref_inv
    dt = config["dt_adv"]
    sdc = config["scale_dt_cond"]
    p.dt = dt
    p.dt_sub = dt / (2 * sdc)
    p.dt_sub_half = 0.5 * p.dt_sub
    p.dt_col_over_dV = config["dt_col"] / grid.volume_cell
    p.scale_dt_cond = sdc
    p.no_col_per_adv = config["no_col_per_adv"]
    p.act_collisions = int(config["act_collisions"])
    p.act_relaxation = int(config["act_relaxation"])
    p.no_iter = config["no_iter_impl_mass"]
    p.solute = 0 if config["solute_type"] == "AS" else 1
    p.g_set = config["g_set"]
    if isinstance(E_col_grid, float) or E_col_grid is None:
        p.nb, p.E_const = 0, (E_col_grid if E_col_grid is not None else 0.5)
    else:
        p.nb = config["no_kernel_bins"]
        p.lnR0 = config["R_kernel_low_log"]
        p.lnf = config["bin_factor_R_log"]
    p.seed, p.run, p.col_step, p.rnd_pos = seed, run, col_step, rnd_pos
    return p


def integrate_adv_steps(p, n_steps, pos, vel, cells, rel_pos, m_w, m_s, xi, active, T_p, water_removed, no_cols,
                        scal, mat, grid_vel, flux, E, prof_rv, prof_Th, t_relax, rnd_stream=None, libm=False):
    """integration.py:1298-1611 repeated n_steps times, in place."""
    act_u8 = active.view(np.uint8)
    Eptr = None if (E is None or isinstance(E, float)) else P(E)
    lib(libm).or_integrate_adv_steps(C.byref(p), n_steps, P(pos), P(vel), P(cells), P(rel_pos), P(m_w), P(m_s),
                                     P(xi), P(act_u8), P(T_p), P(water_removed), P(no_cols), P(scal), P(mat),
                                     P(grid_vel), P(flux), Eptr, P(prof_rv), P(prof_Th), P(t_relax),
                                     P(rnd_stream))


def collision_all_cells(xis, m_w, m_s, vel, grid_T, cells, no_cells, dt_over_dV, E_col_grid, no_kernel_bins,
                        R_kernel_low_log, bin_factor_R_log, no_cols, solute_type="AS", rnd_stream=None, seed=0,
                        run=0, step=0, libm=False):
    """all_or_nothing.py:754-846 in place."""
    if isinstance(E_col_grid, float):
        Eptr, Ec, nb, l0, lf = None, E_col_grid, 0, 0.0, 1.0
    else:
        Eptr, Ec, nb, l0, lf = P(E_col_grid), 0.0, no_kernel_bins, R_kernel_low_log, bin_factor_R_log
    lib(libm).or_collision_all_cells(len(xis), P(xis), P(m_w), P(m_s), P(vel), P(grid_T), P(cells),
                                     int(no_cells[0]), int(no_cells[1]), dt_over_dV, Eptr, Ec, nb, l0, lf,
                                     P(no_cols), 0 if solute_type == "AS" else 1, P(rnd_stream), seed, run, step)


def golovin_step(xis, masses, dt_over_dV, no_cols, rnd=None, seed=0, run=0, step=0):
    lib().or_collision_golovin(len(xis), P(xis), P(masses), dt_over_dV, P(no_cols), P(rnd), seed, run, step)


def derived_fields(m_w, m_s, xi, cells, active, T_grid, mass_dry_inv, V0, solute_type="AS"):
    """evaluation.py:363-474 for one seed and frame: (9, nx, nz) derived fields."""
    nx, nz = T_grid.shape
    out = np.zeros((9, nx, nz))
    lib().or_derived_fields(len(xi), P(np.ascontiguousarray(m_w)), P(np.ascontiguousarray(m_s)),
                            P(np.ascontiguousarray(xi)), P(np.ascontiguousarray(cells, dtype=np.int64)),
                            P(np.ascontiguousarray(active, dtype=bool).view(np.uint8)),
                            P(np.ascontiguousarray(T_grid)), P(np.ascontiguousarray(mass_dry_inv)), nx, nz,
                            float(V0), 0 if solute_type == "AS" else 1, P(out))
    return out


def ensemble_stats(f):
    """evaluation.py:443-474 over the first axis (seeds)."""
    f = np.ascontiguousarray(f, dtype=np.float64)
    mean = np.zeros(f.shape[1:])
    std = np.zeros(f.shape[1:])
    lib().or_ensemble_stats(f.shape[0], int(np.prod(f.shape[1:])), P(f), P(mean), P(std))
    return mean, std
