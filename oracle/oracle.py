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


class OrInitPar(C.Structure):
    """Parameters of the initial state (same layout as the product's LmInitPar)."""

    _fields_ = [
        ("n_modes", _i32), ("solute", _i32), ("weak_threshold", _i32), ("newton_iterations", _i32),
        ("iter_cnt_limit", _i32), ("nx", _i32), ("n_runs", _i32), ("pad", _i32),
        ("l_max", _i32 * 4), ("mu_m_log", _d * 4), ("sigma_m_log", _d * 4), ("kappa", _d * 4),
        ("r_critmin", _d * 4), ("no_rpcm_0", _d * 4),
        ("eta", _d), ("mass_density_dry", _d), ("S_init_max", _d), ("dt_init", _d),
        ("p_ref", _d), ("Theta_l", _d), ("r_tot_0", _d), ("rho_dry_0", _d),
        ("X", _d), ("dx", _d), ("dy", _d), ("dz", _d),
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
    Q = C.POINTER(OrInitPar)
    L.or_initial_mass_fraction.restype = _d
    L.or_initial_mass_fraction.argtypes = [C.c_int, _d, _d, _d]
    L.or_S_eq_mass_fraction.restype = _d
    L.or_S_eq_mass_fraction.argtypes = [C.c_int, _d, _d, _d]
    L.or_sinsip_cell.restype = C.c_int
    L.or_sinsip_cell.argtypes = [Q, _u64, _u32, _u32, C.c_int, _p, _p]
    L.or_init_SinSIP.restype = _i64
    L.or_init_SinSIP.argtypes = [Q, _u64, _i32, _p, _d, _p, _i64] + [_p] * 11
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


def init_SinSIP(q, seed, levels_z, p_0, corner_x, cap=None, libm=False):
    """init.py:439-1236 for one run (seed): dict of the SIP arrays in the reference's order, positions, relative
    positions (2, N), the level state (nz, 21) and the per-cell fields r_v, r_l, e_s, S (4, nx, nz)."""
    levels_z = np.ascontiguousarray(levels_z, dtype=np.float64)
    corner_x = np.ascontiguousarray(corner_x, dtype=np.float64)
    nz = len(levels_z) - 1
    nx = q.nx
    if cap is None:
        cap = nx * nz * sum(q.l_max[m] for m in range(q.n_modes))
    a = {k: np.zeros(cap) for k in ("m_s", "xi", "m_w", "x", "z")}
    a.update({k: np.zeros(cap, dtype=np.int32) for k in ("mode", "cell_x", "cell_z")})
    rel = np.zeros((2, cap))
    lv = np.zeros((nz, 21))
    fields = np.zeros((4, nx, nz))
    n = lib(libm).or_init_SinSIP(C.byref(q), int(seed) & 0xFFFFFFFFFFFFFFFF, nz, P(levels_z), float(p_0),
                                 P(corner_x), cap, P(a["m_s"]), P(a["xi"]), P(a["m_w"]), P(a["mode"]),
                                 P(a["cell_x"]), P(a["cell_z"]), P(a["x"]), P(a["z"]), P(rel), P(lv), P(fields))
    if n < 0:
        raise RuntimeError("or_init_SinSIP: capacity too small")
    out = {k: v[:n].copy() for k, v in a.items()}
    out.update(rel=rel[:, :n].copy(), lv=lv, fields=fields, n=int(n))
    return out


def sinsip_cell(q, seed, level, cell, mode, libm=False):
    """generation_SIP_ensemble.py:147-244 for one cell and mode: (masses, weights)."""
    m = np.zeros(4096)
    w = np.zeros(4096)
    k = lib(libm).or_sinsip_cell(C.byref(q), int(seed) & 0xFFFFFFFFFFFFFFFF, level, cell, mode, P(m), P(w))
    return m[:k].copy(), w[:k].copy()


_tlm = None


def tlm_window(p, n_steps, width, pos, vel, cells, m_w, m_s, xi, active, scal, mat, grid_vel, flux, prof_rv, prof_Th,
               t_relax, dir_xi, dir_ms, events=None):
    """Tangent-linear reference of an advection window (tlm_window.cpp): returns (J, dJ) with J = (J_nc,
    J_lwp) at the end and dJ (2, nd) their derivatives along the nd <= 8 directions dir_xi, dir_ms (nd, N) of
    the initial multiplicities and dry masses.  With collisions on (p.act_collisions), events = (step_off,
    gid_min, gid_max, kind, p_crit) are the fixed events of the window's collision steps.  Inputs are not
    modified."""
    global _tlm
    if _tlm is None:
        path = os.path.join(HERE, "liblacmo_tlm.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} missing: run `make -C oracle`")
        _tlm = C.CDLL(path)
        _tlm.or_tlm_window.restype = C.c_int
        _tlm.or_tlm_window.argtypes = ([C.POINTER(OrAdvParams), _i64, _d] + [_p] * 14 + [C.c_int, _p, _p, _i64] +
                                       [_p] * 5 + [_p, _p])
    c = lambda a, dt=np.float64: np.ascontiguousarray(a, dtype=dt)  # noqa: E731
    dir_xi, dir_ms = c(np.atleast_2d(dir_xi)), c(np.atleast_2d(dir_ms))
    nd = dir_xi.shape[0]
    J = np.zeros(2)
    dJ = np.zeros((2, nd))
    args = [c(pos), c(vel), c(cells, np.int64), c(m_w), c(m_s), c(xi), c(active, np.uint8), c(scal), c(mat),
            c(grid_vel), c(flux), c(prof_rv), c(prof_Th), c(t_relax)]
    if events is None:
        ev = [None] * 5
        n_col = 0
    else:
        ev = [c(events[0], np.int64), c(events[1], np.int32), c(events[2], np.int32), c(events[3], np.int32),
              c(events[4])]
        n_col = len(ev[0]) - 1
    rc = _tlm.or_tlm_window(C.byref(p), int(n_steps), float(width), *[P(a) for a in args], nd, P(dir_xi), P(dir_ms),
                            n_col, *[P(a) for a in ev], P(J), P(dJ))
    if rc != 0:
        raise RuntimeError(f"or_tlm_window: error {rc} (-1: more than 8 directions; -2: events do not match the "
                           "window's collision steps)")
    return J, dJ
