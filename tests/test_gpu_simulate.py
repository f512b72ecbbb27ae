"""GPU: the drop-in simulate() (integration.py:1846-2147) and its files vs the oracle stepped alongside.

simulate() keeps the state on the GPU between frames; the oracle runs the
same advection steps one at a time on copies, and every file simulate writes
(grid fields, all-particle dumps, collision counters, tracer blocks, the full
state at t_end) is compared bit for bit with what the oracle holds at that
time.
"""
import os

import numpy as np
import pytest

from lacmopy_amd import file_handling as fh
from lacmopy_amd import load_kernel_data_Ecol
from lacmopy_amd.setup_case import init_profiles
from oracle import oracle as orc
from test_gpu_parity import _assert_same, _case, _config, _prep

pytestmark = pytest.mark.gpu


def _setup(tmp_path, seed_case):
    grid, *parts = _case(nx=8, nz=8, spc=16, seed=seed_case)
    parts = _prep(grid, parts)
    grid.update_material_properties()
    init = str(tmp_path / "init") + os.sep
    out = str(tmp_path / "out") + os.sep
    os.makedirs(init)
    os.makedirs(out)
    fh.save_grid_arrays_to_npy_file(grid, init + "arr_file1_0.npy", init + "arr_file2_0.npy")
    return grid, parts, init, out


def test_simulate_files_match_oracle(gpu, tmp_path):
    from lacmopy_amd.integration import simulate
    E, lnR0, lnf, nb = load_kernel_data_Ecol("Long_Bott")
    grid, parts, init, out = _setup(tmp_path, 41)
    n_steps, frame_every, dump_every = 20, 10, 5
    config = dict(solute_type="AS", dt_adv=1.0, dt_col=0.5, scale_dt_cond=5, no_col_per_adv=2, t_start=0.0,
                  t_end=float(n_steps), no_iter_impl_mass=3, g_set=9.80665, act_collisions=True,
                  act_relaxation=True, frame_every=frame_every, dump_every=dump_every, trace_ids=6,
                  no_kernel_bins=nb, R_kernel_low_log=lnR0, bin_factor_R_log=lnf, kernel_type="Long_Bott",
                  kernel_method="Ecol_grid_R", seed_sim=2024, simulation_mode="with_collision",
                  paths={"output": out, "init": init}, t_init=0)
    # oracle copies before simulate mutates the grid and particles
    o_scal, o_mat = grid.scalar_fields(), grid.mat_prop()
    o_vel_field, o_flux = grid.velocity.copy(), grid.mass_flux_air_dry.copy()
    prof = init_profiles(grid)
    o = [np.array(a, copy=True) for a in parts]
    pos, vel, cells, m_w, m_s, xi, active = [np.array(a, copy=True) for a in parts]
    N = len(xi)
    water_removed = np.zeros(1)
    no_cols = np.zeros(2, dtype=np.int64)
    simulate(grid, pos, vel, cells, m_w, m_s, xi, active, water_removed, no_cols, config, E)

    trace_ids = np.load(out + "trace_ids.npy")
    assert len(trace_ids) == 6
    cfg = _config(E, lnR0, lnf, nb)
    o_rel, o_Tp, o_wr, o_nc = np.zeros((2, N)), np.ones(N), np.zeros(1), np.zeros(2, dtype=np.int64)
    r_l = np.zeros((8, 8))
    dumps = []
    frames = {}
    for k in range(n_steps + 1):
        if k % frame_every == 0:
            orc.lib().or_update_grid_r_l(N, orc.P(o[3]), orc.P(o[5]), orc.P(o[2]), orc.P(o[6].view(np.uint8)), 8, 8,
                                         orc.P(r_l), orc.P(o_scal[8].copy()))
            o_scal[5] = r_l
            frames[k] = ([a.copy() for a in o], o_scal.copy(), o_nc.copy())
        if k == n_steps:
            break
        if k % dump_every == 0:
            dumps.append((o[0][:, trace_ids].copy(), o[1][:, trace_ids].copy(), o[3][trace_ids].copy(),
                          o[4][trace_ids].copy(), o_Tp[trace_ids].copy(), o[5][trace_ids].copy(), o_wr[0]))
        p = orc.adv_params(grid, N, cfg, seed=2024, col_step=2 * k, E_col_grid=E)
        orc.integrate_adv_steps(p, 1, o[0], o[1], o[2], o_rel, o[3], o[4], o[5], o[6], o_Tp, o_wr, o_nc, o_scal,
                                o_mat, o_vel_field, o_flux, E, *prof)
    assert o_nc.sum() > 0

    for t, (st, sc, nc) in frames.items():
        g = np.load(out + f"grid_scalar_fields_t_{t}.npy")
        _assert_same(g, np.array((sc[4], sc[5], sc[2], sc[0], sc[1], sc[6])), f"grid fields t={t}")
        pv = np.load(out + f"particle_vector_data_all_{t}.npy")
        _assert_same(pv[0], st[0], f"pos t={t}")
        _assert_same(pv[1], st[1], f"vel t={t}")
        _assert_same(np.load(out + f"particle_cells_data_all_{t}.npy"), st[2], f"cells t={t}")
        ps = np.load(out + f"particle_scalar_data_all_{t}.npy")
        _assert_same(ps[0], st[3], f"m_w t={t}")
        _assert_same(ps[1], st[4], f"m_s t={t}")
        _assert_same(np.load(out + f"particle_xi_data_all_{t}.npy"), st[5], f"xi t={t}")
        _assert_same(np.load(out + f"particle_active_ids_data_all_{t}.npy"), st[6], f"active t={t}")
        _assert_same(np.load(out + f"no_cols_{t}.npy"), nc, f"no_cols t={t}")

    per_block = frame_every // dump_every
    for b, t0 in enumerate(range(0, n_steps, frame_every)):
        tv = np.load(out + f"particle_vector_data_{t0}.npy")
        ts = np.load(out + f"particle_scalar_data_{t0}.npy")
        tx = np.load(out + f"particle_xi_data_{t0}.npy")
        tw = np.load(out + f"water_removed_{t0}.npy")
        _assert_same(np.load(out + f"particle_time_block_{t0}.npy"), np.arange(t0, t0 + frame_every, dump_every),
                     "time block")
        for d in range(per_block):
            p_, v_, mw_, ms_, tp_, x_, w_ = dumps[b * per_block + d]
            _assert_same(tv[d, 0], p_, f"traced pos {t0}+{d}")
            _assert_same(tv[d, 1], v_, f"traced vel {t0}+{d}")
            _assert_same(ts[d, 0], mw_, f"traced m_w {t0}+{d}")
            _assert_same(ts[d, 1], ms_, f"traced m_s {t0}+{d}")
            _assert_same(ts[d, 2], tp_, f"traced T_p {t0}+{d}")
            _assert_same(tx[d], x_, f"traced xi {t0}+{d}")
            assert tw[d] == w_
    # final state in memory, full save and its reload
    for nm, a, b in zip(["pos", "vel", "cells", "m_w", "m_s", "xi", "active"], (pos, vel, cells, m_w, m_s, xi, active),
                        o):
        _assert_same(a, b, nm)
    _assert_same(water_removed, o_wr, "water_removed")
    _assert_same(no_cols, o_nc, "no_cols")
    _assert_same(grid.temperature, o_scal[0], "grid.temperature")
    g2, pos2, cells2, vel2, m_w2, m_s2, xi2, act2 = fh.load_grid_and_particles_full(n_steps, out)
    _assert_same(pos2, pos, "reloaded pos")
    _assert_same(cells2, cells, "reloaded cells")
    _assert_same(xi2, xi, "reloaded xi")
    _assert_same(g2.potential_temperature, o_scal[2], "reloaded Theta")
    _assert_same(g2.mixing_ratio_water_liquid, o_scal[5], "reloaded r_l")
    _assert_same(np.load(out + f"water_removed_{n_steps}.npy"), o_wr, "water_removed file")


def test_simulate_interval_matches_oracle(gpu):
    """simulate_interval (integration.py:1615-1842): 7 advection steps with tracer dumps every 3 steps (the
    last block ragged); the first dump holds the caller's T_p (no step has run), later dumps the T_p of the
    steps; traced arrays and the final state bit for bit against the oracle stepped alongside."""
    from lacmopy_amd import rng
    from lacmopy_amd.integration import simulate_interval
    E, lnR0, lnf, nb = load_kernel_data_Ecol("Hall_Bott")
    cfg = _config(E, lnR0, lnf, nb)
    grid, *parts = _case(nx=8, nz=8, spc=16, seed=43)
    parts = _prep(grid, parts)
    pos, vel, cells, m_w, m_s, xi, active = [np.array(a, copy=True) for a in parts]
    N = len(xi)
    scal, mat = grid.scalar_fields(), grid.mat_prop()
    o_scal, o_mat = scal.copy(), mat.copy()
    prof = init_profiles(grid)
    rel_pos = np.zeros((2, N))
    T_p = np.full(N, 281.25)
    dml = np.zeros((8, 8))
    wr, nc = np.zeros(1), np.zeros(2, dtype=np.int64)
    n_adv, every = 7, 3
    tid = np.arange(3, N, N // 9)[:9]
    nd = -(-n_adv // every)
    tv, ts, tx, tw = np.zeros((nd, 2, 2, len(tid))), np.zeros((nd, 3, len(tid))), np.zeros((nd, len(tid))), np.zeros(nd)
    rng.seed(99, 0)
    dt = 1.0
    dt_sub = dt / 10
    simulate_interval(scal, mat, grid.velocity, grid.mass_flux_air_dry, grid.p_ref, grid.p_ref_inv, grid.no_cells,
                      grid.ranges, grid.steps, grid.volume_cell, pos, vel, cells, rel_pos, m_w, m_s, xi, "AS", wr,
                      np.arange(N), active, T_p, dml, dt, dt_sub, 0.5 * dt_sub, 0.5, 5, 2, n_adv, True, True,
                      prof[0], prof[1], prof[2], 3, 9.80665, every, tid, tv, ts, tx, tw, E, nb, lnR0, lnf, nc)
    assert rng.get_step() == 2 * n_adv
    o = [np.array(a, copy=True) for a in parts]
    o_rel, o_Tp, o_wr, o_nc = np.zeros((2, N)), np.full(N, 281.25), np.zeros(1), np.zeros(2, dtype=np.int64)
    d = 0
    for k in range(n_adv):
        if k % every == 0:
            _assert_same(tv[d, 0], o[0][:, tid], f"traced pos dump {d}")
            _assert_same(tv[d, 1], o[1][:, tid], f"traced vel dump {d}")
            _assert_same(ts[d, 0], o[3][tid], f"traced m_w dump {d}")
            _assert_same(ts[d, 1], o[4][tid], f"traced m_s dump {d}")
            _assert_same(ts[d, 2], o_Tp[tid], f"traced T_p dump {d}")
            _assert_same(tx[d], o[5][tid], f"traced xi dump {d}")
            assert tw[d] == o_wr[0]
            d += 1
        p = orc.adv_params(grid, N, cfg, seed=99, col_step=2 * k, E_col_grid=E)
        orc.integrate_adv_steps(p, 1, o[0], o[1], o[2], o_rel, o[3], o[4], o[5], o[6], o_Tp, o_wr, o_nc, o_scal,
                                o_mat, grid.velocity, grid.mass_flux_air_dry, E, *prof)
    assert d == nd and o_nc.sum() > 0
    for nm, a, b in zip(["pos", "vel", "cells", "m_w", "m_s", "xi", "active"], (pos, vel, cells, m_w, m_s, xi, active),
                        o):
        _assert_same(a, b, nm)
    _assert_same(T_p, o_Tp, "T_p")
    _assert_same(rel_pos[:, active], o_rel[:, o[6]], "rel_pos (active)")
    _assert_same(scal[:8], o_scal[:8], "grid_scalar_fields")
    _assert_same(mat, o_mat, "grid_mat_prop")
    _assert_same(wr, o_wr, "water_removed")
    _assert_same(nc, o_nc, "no_cols")


@pytest.mark.parametrize("scale_dt_cond", [5, 1])
def test_resumed_simulate_continues_the_uninterrupted_run(gpu, tmp_path, scale_dt_cond):
    """simulate 0 -> 20 s against simulate 0 -> 10 s, full state saved (integration.py:2132-2134), reloaded with
    load_grid_and_particles_full (file_handling.py:672-760) and simulated 10 -> 20 s: the files of t = 15, 20 s
    and the final state agree bit for bit, including the collision counters (the resumed run's draws continue
    the stream; with scale_dt_cond = 1 an advection step takes 3 collision steps)."""
    from lacmopy_amd.integration import simulate
    E, lnR0, lnf, nb = load_kernel_data_Ecol("Long_Bott")
    grid, parts, init, _ = _setup(tmp_path, 47)
    n_steps = 20

    def config(t0, t1, out):
        return dict(solute_type="AS", dt_adv=1.0, dt_col=0.5, scale_dt_cond=scale_dt_cond, no_col_per_adv=2,
                    t_start=float(t0), t_end=float(t1), no_iter_impl_mass=3, g_set=9.80665, act_collisions=True,
                    act_relaxation=True, frame_every=5, dump_every=5, trace_ids=4, no_kernel_bins=nb,
                    R_kernel_low_log=lnR0, bin_factor_R_log=lnf, kernel_type="Long_Bott", kernel_method="Ecol_grid_R",
                    seed_sim=77, simulation_mode="with_collision", paths={"output": out, "init": init}, t_init=0)

    outs = {}
    for nm in ("full", "first", "second"):
        outs[nm] = str(tmp_path / nm) + os.sep
        os.makedirs(outs[nm])
    import copy
    g_full = copy.deepcopy(grid)
    full = [np.array(a, copy=True) for a in parts]
    wr_f, nc_f = np.zeros(1), np.zeros(2, dtype=np.int64)
    simulate(g_full, *full, wr_f, nc_f, config(0, n_steps, outs["full"]), E)
    half = [np.array(a, copy=True) for a in parts]
    wr_h, nc_h = np.zeros(1), np.zeros(2, dtype=np.int64)
    simulate(copy.deepcopy(grid), *half, wr_h, nc_h, config(0, n_steps // 2, outs["first"]), E)
    g2, pos2, cells2, vel2, m_w2, m_s2, xi2, act2 = fh.load_grid_and_particles_full(n_steps // 2, outs["first"])
    wr2 = np.load(outs["first"] + f"water_removed_{n_steps // 2}.npy")
    nc2 = np.load(outs["first"] + f"no_cols_{n_steps // 2}.npy")
    assert nc2.sum() > 0
    simulate(g2, pos2, vel2, cells2, m_w2, m_s2, xi2, act2, wr2, nc2, config(n_steps // 2, n_steps, outs["second"]), E)
    for nm, a, b in zip(["pos", "vel", "cells", "m_w", "m_s", "xi", "active"],
                        (pos2, vel2, cells2, m_w2, m_s2, xi2, act2), full):
        _assert_same(a, b, nm)
    _assert_same(wr2, wr_f, "water_removed")
    _assert_same(nc2, nc_f, "no_cols")
    assert nc_f.sum() > nc2.sum() // 2
    for t in (10, 15, 20):
        for f in (f"grid_scalar_fields_t_{t}.npy", f"particle_vector_data_all_{t}.npy",
                  f"particle_scalar_data_all_{t}.npy", f"particle_xi_data_all_{t}.npy", f"no_cols_{t}.npy"):
            _assert_same(np.load(outs["second"] + f), np.load(outs["full"] + f), f)
    for t in (10, 15):
        for f in (f"particle_vector_data_{t}.npy", f"particle_xi_data_{t}.npy", f"water_removed_{t}.npy"):
            _assert_same(np.load(outs["second"] + f), np.load(outs["full"] + f), f)
        a, b = np.load(outs["second"] + f"particle_scalar_data_{t}.npy"), np.load(outs["full"] + f"particle_scalar_data_{t}.npy")
        if t == 10:
            # a resumed simulate starts T_p at ones (integration.py:1935): its first tracer dump holds T_p = 1
            assert np.all(a[0, 2] == 1.0)
            a, b = a.copy(), b.copy()
            a[0, 2] = b[0, 2]
        _assert_same(a, b, f"particle_scalar_data_{t}.npy")
    _assert_same(g2.temperature, g_full.temperature, "grid.temperature")
    _assert_same(g2.mixing_ratio_water_vapor, g_full.mixing_ratio_water_vapor, "grid.r_v")
