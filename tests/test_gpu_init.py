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
