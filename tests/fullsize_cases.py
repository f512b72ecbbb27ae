"""Oracle runs of the benchmark workloads at their full sizes, for the GPU parity tests.

Each function runs in a worker process (multiprocessing, spawn) so that the oracle's single-threaded
steps of several runs, or of one large run, proceed beside the GPU work; results go to an .npz file.
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

PART_NAMES = ("pos", "vel", "cells", "m_w", "m_s", "xi", "active")


def workload_kernel(case):
    from lacmopy_amd import load_kernel_data_Ecol
    from lacmopy_amd.setup_case import WORKLOADS
    return load_kernel_data_Ecol(WORKLOADS[case]["kernel"])


def oracle_workload_run(case, seed0, run, n_steps, out_path):
    """n_steps advection steps of run `run` of a batch of `case` (seeded seed0 + 2 run, draws keyed (seed0,
    run)), with the oracle; writes the particles, grid, counters and water_removed to out_path."""
    from lacmopy_amd.setup_case import init_profiles, workload_case, workload_config
    from oracle import oracle as orc
    E, lnR0, lnf, nb = workload_kernel(case)
    cfg = workload_config(case, E, lnR0, lnf)
    grid, *parts = workload_case(case, seed0 + 2 * run)
    pos, vel, cells, m_w, m_s, xi, active = [np.array(a, copy=True) for a in parts]
    N = len(xi)
    scal, mat = grid.scalar_fields(), grid.mat_prop()
    prof_rv, prof_Th, t_relax = init_profiles(grid)
    rel_pos = np.zeros((2, N))
    T_p = np.ones(N)
    wr = np.zeros(1)
    no_cols = np.zeros(2, dtype=np.int64)
    p = orc.adv_params(grid, N, cfg, seed=seed0, run=run, E_col_grid=E)
    orc.integrate_adv_steps(p, n_steps, pos, vel, cells, rel_pos, m_w, m_s, xi, active, T_p, wr, no_cols, scal,
                            mat, grid.velocity, grid.mass_flux_air_dry, E, prof_rv, prof_Th, t_relax)
    np.savez(out_path, pos=pos, vel=vel, cells=cells, m_w=m_w, m_s=m_s, xi=xi, active=active, scal=scal, mat=mat,
             wr=wr, no_cols=no_cols)
    return out_path


def oracle_golovin_box(seed, box, box_seed, t_end, dt_save, out_path):
    """The Golovin box model of one seed of the reference's expo ensemble (run_box_model.py:133-153,
    306-328) with the oracle's collision step, draws keyed (seed, box, step); snapshots every dt_save."""
    from lacmopy_amd.collision.box_model import golovin_expo_ensemble, _save_schedule
    from oracle import oracle as orc
    xi, m, _ = golovin_expo_ensemble(box_seed)
    no_steps, no_saves, dn_save = _save_schedule(1.0, t_end, dt_save)
    xs, ms = [], []
    no_cols = np.zeros(2, dtype=np.int64)
    for step in range(no_steps):
        if step % dn_save == 0:
            xs.append(xi.copy())
            ms.append(m.copy())
        orc.golovin_step(xi, m, 1.0, no_cols, seed=seed, run=box, step=step)
    xs.append(xi.copy())
    ms.append(m.copy())
    np.savez(out_path, xs=np.array(xs), ms=np.array(ms), no_cols=no_cols)
    return out_path
