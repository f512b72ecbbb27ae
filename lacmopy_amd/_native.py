"""ctypes binding of the C ABI declared in include/lacmo_amd.h.

The shared library liblacmo_amd.so is built in-tree by lacmopy_amd.build
(hipcc, gfx950).  Loading fails loudly: there is no CPU fallback anywhere in
the product.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblacmo_amd.so")

LM_ERR_NAMES = {
    1001: "LM_ERR_ARG",
    1002: "LM_ERR_CELL_CAP",
    1003: "LM_ERR_LOST_SIP",
    1004: "LM_ERR_REM_CAP",
    1005: "LM_ERR_OUT_OF_DOMAIN",
    1006: "LM_ERR_RND_SHORT",
    1007: "LM_ERR_EV_CAP",
}

_p = C.c_void_p
_d = C.c_double
_i32 = C.c_int32
_i64 = C.c_int64


class LmEngine(C.Structure):
    """Mirror of struct LmEngine (include/lacmo_amd.h), same field order."""

    _fields_ = [
        ("n_runs", _i32), ("nx", _i32), ("nz", _i32), ("nc", _i32),
        ("n_tot", _i64),
        ("x_min", _d), ("x_max", _d), ("z_min", _d), ("z_max", _d),
        ("dx", _d), ("dz", _d), ("V_cell", _d), ("p_ref", _d), ("p_ref_inv", _d),
        ("solute", _i32), ("no_iter", _i32), ("scale_dt_cond", _i32),
        ("no_col_per_adv", _i32), ("act_collisions", _i32), ("act_relaxation", _i32),
        ("dt", _d), ("dt_sub", _d), ("dt_sub_half", _d), ("dt_col_over_dV", _d), ("g_set", _d),
        ("E", _p), ("nb", _i32), ("e_const_mode", _i32),
        ("E_const", _d), ("lnR0", _d), ("lnf", _d),
        ("seed", C.c_uint64),
        ("col_step", C.c_uint32), ("cur", _i32), ("ccur", _i32), ("pad1", _i32),
        ("x", _p * 2), ("z", _p * 2), ("u", _p * 2), ("w", _p * 2),
        ("m_w", _p * 2), ("m_s", _p * 2), ("xi", _p * 2),
        ("id", _p * 2), ("cell", _p * 2), ("seg", _p * 2),
        ("dmxi", _p), ("col_step_dev", _p), ("away", _p),
        ("scal", _p), ("mat", _p), ("cc", _p), ("dTh", _p), ("dRv", _p), ("dml", _p),
        ("grid_vel", _p), ("flux", _p), ("prof_rv", _p), ("prof_Th", _p), ("t_relax", _p),
        ("colsum", _p),
        ("no_cols", _p), ("water_removed", _p), ("n_equal_xi", _p),
        ("rem_cnt", _p), ("rem_slot", _p),
        ("rem_cap", _i32), ("pad0", _i32),
        ("scan_buf", _p), ("scan_part", _p), ("err", _p), ("ovf_list", _p), ("ovf_cnt", _p),
        ("rnd_stream", _p), ("rnd_off", _p), ("rnd_len", _i64),
        ("prof", _p),
        ("ev_mode", _i32), ("ev_cap", _i32), ("ev_step", _i32), ("ev_nstep", _i32), ("tape_sub", _i32),
        ("tape_nsub", _i32), ("tape_inc", _i32), ("tape_ninc", _i32),
        ("ev", _p), ("ev_cnt", _p), ("ev_count", _p), ("ev_off", _p), ("ev_order", _p), ("run_base", _p),
        ("slot_of", _p),
        ("tape_part", _p), ("tape_part_i", _p), ("tape_grid", _p), ("tape_fields", _p), ("ccd", _p),
        ("part_block", _i32), ("part_spt", _i32), ("col_cap", _i32), ("tape_mode", _i32),
        ("E_max", _d), ("vel_run_stride", _i64),
        ("col_work_per_cu", _i32), ("dml_work_per_cu", _i32), ("work", _p),
        ("col_work_chunk", _i32), ("dml_work_chunk", _i32), ("part_wpe", _i32), ("pad4", _i32),
        ("mig_cnt", _p), ("mig_slot", _p), ("mig_cap", _i32), ("spec_events", _i32), ("col_cost", _p),
        ("big_ws", _p), ("big_cap", _i32), ("big_blocks", _i32),
    ]


class LmAdjoint(C.Structure):
    """Mirror of struct LmAdjoint (include/lacmo_amd.h)."""

    _fields_ = [("lam_mw", _p), ("lam_ms", _p), ("lam_xi", _p), ("lam_cell", _p),
                ("sl_mw", _p), ("sl_ms", _p), ("sl_xi", _p), ("sl_of", _p)]


class LmInitPar(C.Structure):
    """Mirror of struct LmInitPar (include/lacmo_amd.h)."""

    _fields_ = [
        ("n_modes", _i32), ("solute", _i32), ("weak_threshold", _i32), ("newton_iterations", _i32),
        ("iter_cnt_limit", _i32), ("nx", _i32), ("n_runs", _i32), ("pad", _i32),
        ("l_max", _i32 * 4), ("mu_m_log", _d * 4), ("sigma_m_log", _d * 4), ("kappa", _d * 4),
        ("r_critmin", _d * 4), ("no_rpcm_0", _d * 4),
        ("eta", _d), ("mass_density_dry", _d), ("S_init_max", _d), ("dt_init", _d),
        ("p_ref", _d), ("Theta_l", _d), ("r_tot_0", _d), ("rho_dry_0", _d),
        ("X", _d), ("dx", _d), ("dy", _d), ("dz", _d),
    ]


LM_EVENT_BYTES = 80  # sizeof(LmEvent): 6 int32 + 7 double


OFFSET_FIELDS = ["n_tot", "solute", "dt", "E", "seed", "x", "id", "dmxi", "scal",
                 "grid_vel", "no_cols", "rem_cap", "scan_buf", "rnd_stream", "rnd_len", "away", "col_step_dev",
                 "ev_mode", "ev", "run_base", "tape_part", "ccd", "part_block", "E_max", "vel_run_stride",
                 "col_work_per_cu", "work", "dml_work_chunk", "part_wpe", "mig_cnt", "mig_cap", "spec_events", "col_cost",
                 "big_ws", "big_cap"]

EXPORTED = [
    "lm_version", "lm_last_error_message", "lm_engine_sizeof", "lm_engine_offsets",
    "lm_get_constants", "lm_num_constants", "lm_engine_init_cells", "lm_collision_step",
    "lm_position_step", "lm_reindex", "lm_build_work_list", "lm_advection_increments", "lm_substep",
    "lm_adv_steps", "lm_engine_check", "lm_golovin_step", "lm_ecol_box_step", "lm_math_probe",
    "lm_prof_enable", "lm_prof_collect", "lm_prof_disable", "lm_particle_substep", "lm_grid_substep",
    "lm_box2d_step", "lm_box_check", "lm_divergence_upwind", "lm_update_r_l",
    "lm_kernel_grid_m_step", "lm_set_col_step", "lm_graph_capture", "lm_graph_launch", "lm_graph_destroy",
    "lm_derived_fields", "lm_ensemble_stats", "lm_build_slot_of", "lm_objective", "lm_adj_seed",
    "lm_adj_adv_steps", "lm_init_par_sizeof", "lm_init_level_profile", "lm_init_level_sips", "lm_init_positions",
    "lm_box_minmax", "lm_box_stats", "lm_mean_sem", "lm_collision_all_cells", "lm_collision_all_cells_check",
    "lm_collision_all_cells_noalloc", "lm_big_ws_bytes", "lm_load_reference_layout", "lm_store_reference_layout",
    "lm_kernel_index_check",
]

_lib = None


class LacmoError(RuntimeError):
    pass


def lib():
    """Load liblacmo_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise LacmoError(
            f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
            "(hipcc --offload-arch=gfx950); there is no CPU fallback")
    L = C.CDLL(LIB_PATH)
    E = C.POINTER(LmEngine)
    L.lm_version.restype = C.c_char_p
    L.lm_last_error_message.restype = C.c_char_p
    L.lm_engine_sizeof.restype = _i64
    L.lm_big_ws_bytes.argtypes = [_i32]
    L.lm_big_ws_bytes.restype = _i64
    L.lm_load_reference_layout.argtypes = [E, _i32, _i64, _i64] + [_p] * 7 + [_p]
    L.lm_store_reference_layout.argtypes = [E, _i32, _i64, _i64] + [_p] * 9 + [_p]
    L.lm_engine_offsets.argtypes = [C.POINTER(_i64), _i32]
    L.lm_get_constants.argtypes = [C.POINTER(_d), _i32]
    L.lm_num_constants.restype = _i32
    for name in ["lm_engine_init_cells", "lm_collision_step", "lm_reindex", "lm_build_work_list", "lm_engine_check",
                 "lm_grid_substep", "lm_update_r_l"]:
        getattr(L, name).argtypes = [E, _p]
        getattr(L, name).restype = C.c_int
    L.lm_position_step.argtypes = [E, _d, _p]
    L.lm_advection_increments.argtypes = [E, _i32, _p]
    L.lm_substep.argtypes = [E, _d, _d, _p]
    L.lm_particle_substep.argtypes = [E, _d, _d, _p]
    L.lm_box2d_step.argtypes = [_i32, _p, _p, _p, _p, _p, _p, _p, _p, _d, _p, _d, _i32, _d, _d, _p,
                                C.c_uint64, C.c_uint32, _p, _p]
    L.lm_box_check.argtypes = [_p]
    L.lm_set_col_step.argtypes = [E, C.c_uint32, _p]
    L.lm_derived_fields.argtypes = [E, _p, _p]
    L.lm_ensemble_stats.argtypes = [_p, _i32, _i64, _p, _p, _p]
    L.lm_graph_capture.argtypes = [E, _i32, _p, C.POINTER(_p)]
    L.lm_graph_launch.argtypes = [E, _p, _p]
    L.lm_graph_destroy.argtypes = [_p]
    L.lm_kernel_grid_m_step.argtypes = [_i32, _p, _p, _p, _d, _p, _i32, _d, _d, _p, C.c_uint64, C.c_uint32, _p, _p]
    L.lm_divergence_upwind.argtypes = [_p, _p, _p, _i32, _i32, _i32, _d, _d, _i32, _i32, _p]
    L.lm_adv_steps.argtypes = [E, _i32, _p]
    A = C.POINTER(LmAdjoint)
    L.lm_build_slot_of.argtypes = [E, _p]
    L.lm_objective.argtypes = [E, _d, _p, _p]
    L.lm_adj_seed.argtypes = [E, A, _d, _d, _d, _p]
    L.lm_adj_adv_steps.argtypes = [E, A, _i32, _i32, _p]
    Q = C.POINTER(LmInitPar)
    L.lm_init_par_sizeof.restype = _i64
    L.lm_init_level_profile.argtypes = [Q, _p, _i32, _d, _d, _p, _p, _p, _p]
    L.lm_init_level_sips.argtypes = [Q, _p, _i32, _p, _p, _p, _i64] + [_p] * 10 + [_p]
    L.lm_init_positions.argtypes = [_i64, _p, _p, _p, _p, _p, _p, _p, _d, _d, _p, _p, _p, _p]
    L.lm_golovin_step.argtypes = [_i32, _p, _p, _p, _d, _p, C.c_uint64, C.c_uint32, _p, _p]
    L.lm_ecol_box_step.argtypes = [_i32, _p, _p, _p, _p, _p, _p, _d, _p, _i32, _d, _d, _p,
                                   C.c_uint64, C.c_uint32, _p, _i32, _p]
    L.lm_math_probe.argtypes = [_i32, _i64, _p, _p, _p, _p]
    L.lm_kernel_index_check.argtypes = [_i64, _p, _d, _d, _i32, _p, _p, _p, _p]
    L.lm_box_minmax.argtypes = [_i32, _i64, _p, _p, _p, _p]
    L.lm_box_stats.argtypes = [_i32, _i32, _p, _i64, _p, _p, _p, _p, _p, _i32, _d, _p, _p, _p, _p, _p]
    L.lm_mean_sem.argtypes = [_p, _i32, _i64, _p, _p, _p]
    L.lm_collision_all_cells.argtypes = [_i64, _p, _p, _p, _p, _p, _p, _i32, _i32, _d, _p, _i32, _d, _d, _d, _d, _p,
                                         _i32, C.c_uint64, C.c_uint32, _p, _i64, _p]
    L.lm_collision_all_cells_check.argtypes = [_i32, _p]
    L.lm_collision_all_cells_noalloc.argtypes = L.lm_collision_all_cells.argtypes
    L.lm_prof_enable.argtypes = [E, _i32]
    L.lm_prof_collect.argtypes = [E, C.POINTER(_d), C.POINTER(_i32), _i32]
    L.lm_prof_disable.argtypes = [E]
    for name in ["lm_position_step", "lm_advection_increments", "lm_substep", "lm_adv_steps",
                 "lm_golovin_step", "lm_ecol_box_step", "lm_math_probe", "lm_prof_enable", "lm_prof_collect",
                 "lm_prof_disable", "lm_particle_substep", "lm_grid_substep", "lm_box2d_step", "lm_box_check",
                 "lm_divergence_upwind", "lm_kernel_grid_m_step", "lm_set_col_step", "lm_graph_capture",
                 "lm_graph_launch", "lm_graph_destroy", "lm_derived_fields", "lm_ensemble_stats",
                 "lm_build_slot_of", "lm_objective", "lm_adj_seed", "lm_adj_adv_steps", "lm_init_level_profile",
                 "lm_init_level_sips", "lm_init_positions", "lm_box_minmax", "lm_box_stats", "lm_mean_sem",
                 "lm_collision_all_cells", "lm_collision_all_cells_check", "lm_collision_all_cells_noalloc",
                 "lm_kernel_index_check"]:
        getattr(L, name).restype = C.c_int
    _lib = L
    return L


def check(rc):
    """Raise LacmoError for a nonzero return code of an lm_* entry point."""
    if rc != 0:
        msg = lib().lm_last_error_message().decode(errors="replace")
        raise LacmoError(f"lacmo_amd error {rc} ({LM_ERR_NAMES.get(rc, 'hip')}): {msg}")
