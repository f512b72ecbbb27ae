"""The benchmark configurations at their stated sizes on the GPU, against the oracle.

  * box model (Golovin, 50 seeds of the reference's expo SinSIP ensemble, kappa = 40, dt 1 s, 3600 s): every
    snapshot of every box bit-exact against the oracle, and the ensemble moments against Golovin's closed
    forms (golovin.py:48-58) within stated statistical tolerances;
  * icmw75 plain (75x75 cells, Poisson(128) SIPs per cell, 720k SIPs, Long table), autotuned launch shapes:
    one advection step bit-exact;
  * icmw75 batch (64 seeds, 46M SIPs, autotuned shapes, steps replayed from a captured HIP graph): two
    advection steps; runs 0, 31 and 63 bit-exact against the oracle with their own draw keys;
  * stress150 (150x150 cells of 10 m, Poisson(256), 5.8M SIPs, Hall table, dt 0.5 s, 20 substeps): one
    advection step bit-exact.
The oracle's steps run in worker processes beside the GPU work.
"""
import math
import multiprocessing as mp
import os
import sys
from concurrent.futures import ProcessPoolExecutor

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import fullsize_cases as fc  # noqa: E402

pytestmark = pytest.mark.gpu

SEED0 = 1001


def _pool(n):
    return ProcessPoolExecutor(max_workers=n, mp_context=mp.get_context("spawn"))


def _assert_same(a, b, name):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape, (name, a.shape, b.shape)
    if a.dtype.kind == "f":
        bad = ~((a == b) | (np.isnan(a) & np.isnan(b)))
        if bad.any():
            k = np.argwhere(bad)[0]
            raise AssertionError(f"{name}: {bad.sum()} of {a.size} values differ (first at {k}: {a[tuple(k)]!r} "
                                 f"vs {b[tuple(k)]!r})")
    else:
        assert np.array_equal(a, b), f"{name}: integer arrays differ"


def _workload_engine(case, n_runs):
    from lacmopy_amd.engine import Engine
    from lacmopy_amd.setup_case import WORKLOADS, init_profiles, workload_case
    c = WORKLOADS[case]
    E, lnR0, lnf, nb = fc.workload_kernel(case)
    cases = [workload_case(case, SEED0 + 2 * r) for r in range(n_runs)]
    g0 = cases[0][0]
    eng = Engine(c["nx"], c["nz"], [len(k[6]) for k in cases], g0.ranges, g0.steps, g0.volume_cell)
    eng.set_params(solute="AS", no_iter_impl_mass=3, scale_dt_cond=c["sdc"], no_col_per_adv=2, act_collisions=True,
                   act_relaxation=True, dt=c["dt"], dt_col=c["dt"] / 2, g_set=9.80665, seed=SEED0)
    eng.set_collision_kernel(E, lnR0, lnf)
    eng.set_flow(g0.velocity, g0.mass_flux_air_dry)
    for r, k in enumerate(cases):
        eng.set_relaxation(r, *init_profiles(k[0]))
        eng.set_grid(r, k[0].scalar_fields(), k[0].mat_prop())
    eng.init_cells()
    eng.load_particles([k[1:] for k in cases])
    del cases
    return eng


def _compare_run(eng, r, path):
    d = np.load(path)
    got = eng.export_particles(r)
    for nm, a in zip(fc.PART_NAMES, got):
        _assert_same(a, d[nm], f"run {r} {nm}")
    s, m = eng.export_grid(r)
    _assert_same(s[:8], d["scal"][:8], f"run {r} grid_scalar_fields")
    _assert_same(m, d["mat"], f"run {r} grid_mat_prop")
    nc, wr = eng.counters(r)
    _assert_same(nc, d["no_cols"], f"run {r} no_cols")
    _assert_same(wr, d["wr"], f"run {r} water_removed")
    return int(d["no_cols"].sum())


def test_golovin_box_model_50_seeds_3600_s(gpu, tmp_path):
    from lacmopy_amd.collision.box_model import golovin_expo_ensemble, simulate_collisions_batched
    n_seeds, t_end, dt_save, b, LWC0 = 50, 3600.0, 150.0, 1.5, 1.0E-3
    seeds = [SEED0 + 2 * k for k in range(n_seeds)]
    print(f"C1: {n_seeds} Golovin boxes x {int(t_end)} steps", flush=True)
    with _pool(16) as ex:
        futs = [ex.submit(fc.oracle_golovin_box, SEED0, k, s, t_end, dt_save, str(tmp_path / f"box{k}.npz"))
                for k, s in enumerate(seeds)]
        boxes = []
        for s in seeds:
            xi, m, DNC0 = golovin_expo_ensemble(s)
            boxes.append((xi, m))
        assert 190 <= np.mean([len(x) for x, _ in boxes]) <= 240  # about 5 kappa SIPs per box
        xs, ms, times, nc = simulate_collisions_batched(boxes, None, "Golovin", "analytic", 1.0, 1.0, t_end, dt_save,
                                                        SEED0)
        paths = [f.result() for f in futs]
    np.testing.assert_array_equal(times, np.arange(0.0, t_end + 1.0, dt_save))
    for k, p in enumerate(paths):
        d = np.load(p)
        _assert_same(xs[k], d["xs"], f"box {k} xi snapshots")
        _assert_same(ms[k], d["ms"], f"box {k} mass snapshots")
        _assert_same(nc[k], d["no_cols"], f"box {k} no_cols")
    # ensemble moments (kg, m^-3) vs Golovin's closed forms, DNC0 = LWC0 / m(9.3 um)
    mom = np.zeros((len(times), 4, n_seeds))
    for k in range(n_seeds):
        mk = 1.0E-18 * ms[k]
        for n in range(4):
            mom[:, n, k] = (xs[k] * mk ** n).sum(axis=1)
    avg = mom.mean(axis=2)
    an = np.array([[DNC0 * math.exp(-b * LWC0 * t), LWC0, 2 * LWC0 ** 2 / DNC0 * math.exp(2 * b * LWC0 * t),
                    LWC0 ** 3 / DNC0 ** 2 * (12 * math.exp(4 * b * LWC0 * t) - 6 * math.exp(3 * b * LWC0 * t))]
                   for t in times])
    rel = avg / an - 1.0
    # lambda_0 within 6% at every time: three standard errors of the 50-seed mean (about 2%)
    assert np.max(np.abs(rel[:, 0])) < 0.06, rel[:, 0]
    # lambda_1 conserved by every event (to rounding), and the sampled LWC within 0.5% of LWC0
    np.testing.assert_allclose(mom[:, 1, :], np.broadcast_to(mom[0, 1, :], mom[:, 1, :].shape), rtol=1e-12)
    assert np.max(np.abs(rel[:, 1])) < 0.005
    # lambda_2 within 10% to 600 s (standard error 3%); all-or-nothing with ~200 SIPs (kappa = 40) then
    # underestimates it more and more, by 30-45% at 3600 s (Unterstrasser et al. 2017 find the same
    # convergence in kappa), so only the sign and size of that bias are checked there
    assert np.max(np.abs(rel[times <= 600.0, 2])) < 0.10, rel[:, 2]
    assert -0.6 < rel[-1, 2] < -0.1
    assert nc.sum() > 0
    _check_box_statistics(xs, ms)


def _check_box_statistics(xs, ms, no_bins=50):
    """analyze_sim_data's statistics on the GPU (lm_box_stats, lm_mean_sem) against NumPy's histogram, mean
    and std of the same snapshots: the per-bin and moment sums differ only in summation order (tolerance
    1e-12 of the largest value of each array)."""
    from lacmopy_amd.collision.box_model import ensemble_statistics
    mk = [1.0E-18 * m for m in ms]
    got = ensemble_statistics(xs, mk, 1.0E3, 1.0, no_bins)
    ref = fc.numpy_box_statistics(xs, mk, 1.0E3, 1.0, no_bins)
    for key, r in ref.items():
        g = got[key]
        assert g.shape == r.shape, key
        if key.startswith("moments"):
            np.testing.assert_allclose(g, r, rtol=1e-10, atol=0, err_msg=key)
        else:  # NumPy's weighted histogram differences cumulative sums: absolute errors on the row's scale
            scale = np.max(np.abs(r), axis=-1, keepdims=True)
            assert np.all(np.abs(g - r) <= 1e-12 * scale), (key, float(np.max(np.abs(g - r) / scale)))


@pytest.mark.parametrize("kernel", ["Long_Bott", "Hall_Bott"])
def test_hydrodynamic_box_models_50_seeds_3600_s(gpu, tmp_path, kernel):
    """Long and Hall box models (Ecol_grid_R on the box model's table, 0.6-6000 um at 200 bins per decade,
    Beard velocities updated after every step) of the 50 seeds: every snapshot, radius and velocity of every
    box bit-exact against the oracle; ensemble moments against the Wang et al. (2007) moments the reference
    holds (lambda_0 within 4% to 1800 s, lambda_2 within 5% to 1200 s, lambda_1 within 0.5%)."""
    from lacmopy_amd.collision.box_model import ecol_expo_ensemble, simulate_collisions_batched
    from lacmopy_amd.collision.kernel import generate_E_col_grid_R_np
    n_seeds, t_end, dt_save = 50, 3600.0, 200.0
    seeds = [SEED0 + 2 * k for k in range(n_seeds)]
    E, rg = generate_E_col_grid_R_np(0.6, 6.0E3, 200, kernel)
    E_path = str(tmp_path / "E.npz")
    np.savez(E_path, E=E, rg=rg)
    print(f"C1 {kernel}: {n_seeds} boxes x {int(t_end)} steps", flush=True)
    with _pool(16) as ex:
        futs = [ex.submit(fc.oracle_ecol_box, kernel, SEED0, k, s, t_end, dt_save, E_path,
                          str(tmp_path / f"box{k}.npz")) for k, s in enumerate(seeds)]
        boxes = [ecol_expo_ensemble(s) for s in seeds]
        kq = (E, len(rg), float(np.log(rg[0])), float(np.log(rg[1] / rg[0])))
        xs, ms, times, nc = simulate_collisions_batched(boxes, kq, kernel, "Ecol_grid_R", 1.0, 1.0, t_end, dt_save,
                                                        SEED0)
        paths = [f.result() for f in futs]
    for k, p in enumerate(paths):
        d = np.load(p)
        _assert_same(xs[k], d["xs"], f"box {k} xi snapshots")
        _assert_same(ms[k], d["ms"], f"box {k} mass snapshots")
        _assert_same(boxes[k][2], d["R"], f"box {k} radii")
        _assert_same(boxes[k][3], d["vel"], f"box {k} velocities")
        _assert_same(nc[k], d["no_cols"], f"box {k} no_cols")
    ref = np.loadtxt(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_data", kernel,
                                  "Wang_2007_moments.txt"))
    mom = np.array([[(xs[k] * (1.0E-18 * ms[k]) ** n).sum(axis=1) for n in range(4)] for k in range(n_seeds)])
    rel = mom.mean(axis=0) / ref - 1.0
    assert np.max(np.abs(rel[0, times <= 1800.0])) < 0.04, rel[0]
    assert np.max(np.abs(rel[1])) < 0.005, rel[1]
    assert np.max(np.abs(rel[2, times <= 1200.0])) < 0.05, rel[2]
    _check_box_statistics(xs, ms)


def test_icmw75_plain_run_one_step_autotuned(gpu, tmp_path):
    print("C2: icmw75 plain run, one step, autotuned shapes", flush=True)
    with _pool(1) as ex:
        fut = ex.submit(fc.oracle_workload_run, "icmw75", SEED0, 0, 1, str(tmp_path / "c2.npz"))
        eng = _workload_engine("icmw75", 1)
        assert 7.0e5 < eng.n_tot < 7.4e5
        shape = eng.autotune()["chosen"]
        print("  launch shape", shape, flush=True)
        eng.adv_steps(1)
        path = fut.result()
    assert _compare_run(eng, 0, path) > 0


def test_icmw75_batch_of_64_seeds_graph_replay(gpu, tmp_path):
    runs = (0, 31, 63)
    print(f"C3: icmw75 batch of 64 seeds, two graph-replayed steps, runs {runs} vs the oracle", flush=True)
    with _pool(len(runs)) as ex:
        futs = {r: ex.submit(fc.oracle_workload_run, "icmw75", SEED0, r, 2, str(tmp_path / f"c3_{r}.npz"))
                for r in runs}
        eng = _workload_engine("icmw75", 64)
        assert 4.5e7 < eng.n_tot < 4.7e7
        shape = eng.autotune()["chosen"]
        print("  launch shape", shape, flush=True)
        eng.adv_steps_graph(2)
        paths = {r: f.result() for r, f in futs.items()}
    for r in runs:
        assert _compare_run(eng, r, paths[r]) > 0


def test_stress150_one_step(gpu, tmp_path):
    print("C5: stress150 run, one step, autotuned shapes", flush=True)
    with _pool(1) as ex:
        fut = ex.submit(fc.oracle_workload_run, "stress150", SEED0, 0, 1, str(tmp_path / "c5.npz"))
        eng = _workload_engine("stress150", 1)
        assert 5.6e6 < eng.n_tot < 5.9e6
        shape = eng.autotune()["chosen"]
        print("  launch shape", shape, flush=True)
        eng.adv_steps(1)
        path = fut.result()
    assert _compare_run(eng, 0, path) > 0
