"""GPU: the initial state (lacmopy_amd.initialize) against the oracle's restatement of init.py, bit for bit.

Grids of the reference's full height (75 levels of 20 m, so that the cloud levels above ~900 m take
hundreds of adjustment iterations and one level runs into iter_cnt_limit) and a few columns.  Every SIP
array, position, level state and per-cell field must be identical; a batch of seeds must equal the same
seeds initialised one at a time.
"""
import os

import numpy as np
import pytest

from init_cases import init_config, oracle_par
from lacmopy_amd.grid import Grid
from lacmopy_amd.initialize import LV_ITER, init_parameters, initialize_batch, initialize_grid_and_particles_SinSIP
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _oracle_run(cfg, seed):
    g = Grid(cfg["grid_ranges"], cfg["grid_steps"], cfg["dy"])
    q = init_parameters(cfg, g)
    q.n_runs = 1
    return orc.init_SinSIP(oracle_par(q), seed, g.corners[1][0], cfg["p_0"], g.corners[0][:, 0])


def _compare(s, o, what):
    assert len(s.m_s) == o["n"], f"{what}: {len(s.m_s)} SIPs vs {o['n']}"
    for k_gpu, k_or in (("m_s", "m_s"), ("xi", "xi"), ("m_w", "m_w"), ("modes", "mode")):
        a, b = getattr(s, k_gpu), o[k_or]
        bad = np.flatnonzero(a != b)
        assert bad.size == 0, f"{what}: {k_gpu} differs at {bad[:5]} ({a[bad[:3]]} vs {b[bad[:3]]})"
    assert np.array_equal(s.cells[0], o["cell_x"]) and np.array_equal(s.cells[1], o["cell_z"]), what
    assert np.array_equal(s.pos[0], o["x"]) and np.array_equal(s.pos[1], o["z"]), what
    assert np.array_equal(s.rel_pos, o["rel"]), what
    lv = o["lv"]
    assert np.array_equal(s.iterations, lv[:, LV_ITER].astype(np.int64)), (what, s.iterations, lv[:, LV_ITER])
    assert np.array_equal(s.no_rpcm_scale_factors, lv[:, 14]), what
    f = o["fields"]
    for k, name in enumerate(("mixing_ratio_water_vapor", "mixing_ratio_water_liquid")):
        assert np.array_equal(getattr(s.grid, name), f[k]), (what, name)
    assert np.array_equal(s.grid.saturation, f[3]), what
    assert np.array_equal(s.grid.pressure[0], lv[:, 5]) and np.array_equal(s.grid.temperature[0], lv[:, 4]), what
    assert np.array_equal(s.grid.mass_density_air_dry[0], lv[:, 6]), what


@pytest.mark.parametrize("solute,thr,nx,seeds", [("AS", "fix", 6, (3711, 3713)), ("NaCl", "weak", 3, (17,))])
def test_initial_state_matches_oracle(gpu, solute, thr, nx, seeds):
    cfg = init_config(nx=nx, solute=solute, eta_threshold=thr)
    states = initialize_batch(cfg, list(seeds))
    for seed, s in zip(seeds, states):
        o = _oracle_run(cfg, seed)
        _compare(s, o, f"{solute}/{thr} seed {seed}")
        assert s.iterations.max() > 100, "the cloud levels must take many adjustment iterations"
    print(f"{solute}/{thr}: {[len(s.m_s) for s in states]} SIPs, iterations per level max "
          f"{[int(s.iterations.max()) for s in states]}")


def test_batch_equals_single_runs(gpu):
    cfg = init_config(nx=4, nz=50)
    both = initialize_batch(cfg, [5, 9])
    for k, seed in enumerate((5, 9)):
        one = initialize_batch(cfg, [seed])[0]
        for name in ("m_s", "xi", "m_w", "pos", "cells", "vel", "iterations", "rel_pos"):
            assert np.array_equal(getattr(one, name), getattr(both[k], name)), (seed, name)
        assert np.array_equal(one.grid.saturation, both[k].grid.saturation)
    assert not np.array_equal(both[0].m_s[:10], both[1].m_s[:10]), "different seeds must give different SIPs"


def test_dropin_writes_reference_files(gpu, tmp_path):
    from lacmopy_amd import file_handling as fh
    cfg = init_config(nx=4, nz=30, paths={"grid": str(tmp_path) + "/"})
    grid, pos, cells, vel, m_w, m_s, xi, act = initialize_grid_and_particles_SinSIP(cfg)
    g2, pos2, cells2, vel2, m_w2, m_s2, xi2, act2 = fh.load_grid_and_particles_full(0, str(tmp_path) + "/")
    for a, b in ((pos, pos2), (cells, cells2), (vel, vel2), (m_w, m_w2), (m_s, m_s2), (xi, xi2), (act, act2)):
        assert np.array_equal(a, b)
    assert np.array_equal(g2.mixing_ratio_water_liquid, grid.mixing_ratio_water_liquid)
    assert np.array_equal(g2.velocity, grid.velocity)
    assert os.path.exists(os.path.join(tmp_path, "modes_0.npy"))
    scale = np.load(os.path.join(tmp_path, "no_rpcm_scale_factors_lvl_wise.npy"))
    assert scale.shape == (30,) and np.all(np.diff(scale) < 0), "dry density falls with height"


def test_step_from_gpu_initial_states_matches_oracle(gpu):
    """Initial states of two seeds -> one engine (each run with its own grid, relaxation profiles and face
    velocities) -> two advection steps, each run bit-exact against the oracle's step of that run."""
    from lacmopy_amd import load_kernel_data_Ecol
    from lacmopy_amd.engine import Engine
    from lacmopy_amd.setup_case import init_profiles
    from test_gpu_parity import _assert_same, _config
    E, lnR0, lnf, nb = load_kernel_data_Ecol("Long_Bott")
    cfg = dict(_config(E, lnR0, lnf, nb), seed_sim=41)
    states = initialize_batch(init_config(nx=4, nz=75), [101, 103])
    assert not np.array_equal(states[0].grid.velocity, states[1].grid.velocity), "velocities differ per seed"
    eng = Engine.from_initial_states(states, cfg, E)
    eng.adv_steps(2)
    for r, s in enumerate(states):
        g = s.grid
        pos, vel, cells, m_w, m_s, xi, act = [np.array(a, copy=True) for a in
                                              (s.pos, s.vel, s.cells, s.m_w, s.m_s, s.xi, s.active_ids)]
        N = len(xi)
        # simulate() derives the cells from the positions first (integration.py:1932-1933)
        p0 = orc.adv_params(g, N, cfg, E_col_grid=E)
        orc.lib().or_update_pos_and_cells(orc.C.byref(p0), 0.0, orc.P(pos), orc.P(vel), orc.P(cells),
                                          orc.P(np.zeros((2, N))), orc.P(m_w), orc.P(xi), orc.P(act.view(np.uint8)),
                                          orc.P(np.zeros(1)))
        scal, mat = g.scalar_fields(), g.mat_prop()
        prof_rv, prof_Th, t_relax = init_profiles(g)
        wr, no_cols = np.zeros(1), np.zeros(2, dtype=np.int64)
        p = orc.adv_params(g, N, cfg, seed=41, run=r, E_col_grid=E)
        orc.integrate_adv_steps(p, 2, pos, vel, cells, np.zeros((2, N)), m_w, m_s, xi, act, np.ones(N), wr, no_cols,
                                scal, mat, g.velocity, g.mass_flux_air_dry, E, prof_rv, prof_Th, t_relax)
        got = eng.export_particles(r)
        for nm, a, b in zip(["pos", "vel", "cells", "m_w", "m_s", "xi", "active"], got,
                            (pos, vel, cells, m_w, m_s, xi, act)):
            _assert_same(a, b, f"run {r} {nm}")
        sg, mg = eng.export_grid(r)
        _assert_same(sg[:8], scal[:8], f"run {r} grid_scalar_fields")
        _assert_same(mg, mat, f"run {r} grid_mat_prop")
        _assert_same(eng.counters(r)[0], no_cols, f"run {r} no_cols")
