"""GPU parity: gfx950 kernels (through the C ABI) vs the CPU oracle, bit-exact.

Inputs are seeded synthetic ICMW cases small enough for the oracle to finish
in seconds.  Integer / index results (cells, active flags, collision counts,
ID order) and every FP64 state value must be identical.
"""
import numpy as np
import pytest

from lacmopy_amd import load_kernel_data_Ecol, rng
from lacmopy_amd.setup_case import init_profiles, make_case
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _config(E, lnR0, lnf, nb, **kw):
    c = dict(solute_type="AS", dt_adv=1.0, dt_col=0.5, scale_dt_cond=5, no_col_per_adv=2, no_iter_impl_mass=3,
             g_set=9.80665, act_collisions=True, act_relaxation=True, no_kernel_bins=nb, R_kernel_low_log=lnR0,
             bin_factor_R_log=lnf)
    c.update(kw)
    return c


def _case(nx=10, nz=10, spc=24, seed=5, **kw):
    kw.setdefault("R_drop", 22.0)
    kw.setdefault("sigma_drop", 1.6)
    return make_case(nx=nx, nz=nz, spc=spc, seed=seed, cloud_base=nz * 20.0 * 0.4, cloud_top=nz * 20.0 * 0.7, **kw)


def _engine(grid, parts, config, E, lnR0, lnf, seed, n_runs=1, rnd=None):
    from lacmopy_amd.engine import Engine
    prof = init_profiles(grid)
    N = len(parts[3])
    eng = Engine(int(grid.no_cells[0]), int(grid.no_cells[1]), [N] * n_runs, grid.ranges, grid.steps,
                 grid.volume_cell)
    eng.set_params(solute=config["solute_type"], no_iter_impl_mass=config["no_iter_impl_mass"],
                   scale_dt_cond=config["scale_dt_cond"], no_col_per_adv=config["no_col_per_adv"],
                   act_collisions=config["act_collisions"], act_relaxation=config["act_relaxation"],
                   dt=config["dt_adv"], dt_col=config["dt_col"], g_set=config["g_set"], seed=seed)
    eng.set_collision_kernel(E, lnR0, lnf)
    eng.set_flow(grid.velocity, grid.mass_flux_air_dry)
    for r in range(n_runs):
        eng.set_relaxation(r, *prof)
        eng.set_grid(r, grid.scalar_fields(), grid.mat_prop())
    eng.init_cells()
    eng.load_particles([parts] * n_runs)
    eng.set_rnd_stream(rnd)
    return eng


def _oracle_run(grid, parts, config, E, seed, n_steps, rnd=None, libm=False, run=0):
    pos, vel, cells, m_w, m_s, xi, active = [np.array(a, copy=True) for a in parts]
    N = len(xi)
    scal, mat = grid.scalar_fields(), grid.mat_prop()
    prof_rv, prof_Th, t_relax = init_profiles(grid)
    rel_pos = np.zeros((2, N))
    T_p = np.ones(N)
    wr = np.zeros(1)
    no_cols = np.zeros(2, dtype=np.int64)
    p = orc.adv_params(grid, N, config, seed=seed, run=run, E_col_grid=E)
    orc.integrate_adv_steps(p, n_steps, pos, vel, cells, rel_pos, m_w, m_s, xi, active, T_p, wr, no_cols, scal,
                            mat, grid.velocity, grid.mass_flux_air_dry, E, prof_rv, prof_Th, t_relax,
                            rnd_stream=rnd, libm=libm)
    return (pos, vel, cells, m_w, m_s, xi, active), scal, mat, wr, no_cols


def _prep(grid, parts):
    """simulate() starts with update_cells_and_rel_pos (integration.py:1932-1933)."""
    pos, vel, cells, m_w, m_s, xi, active = [np.array(a, copy=True) for a in parts]
    N = len(xi)
    p = orc.adv_params(grid, N, _config(0.5, 0, 1, 0))
    rel = np.zeros((2, N))
    orc.lib().or_update_pos_and_cells(orc.C.byref(p), 0.0, orc.P(pos), orc.P(vel), orc.P(cells), orc.P(rel),
                                      orc.P(m_w), orc.P(xi), orc.P(active.view(np.uint8)), orc.P(np.zeros(1)))
    return pos, vel, cells, m_w, m_s, xi, active


def _assert_same(a, b, name):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape, name
    if a.dtype.kind == "f":
        bad = ~((a == b) | (np.isnan(a) & np.isnan(b)))
        if bad.any():
            k = np.argwhere(bad)[0]
            rel = np.max(np.abs(a[bad] - b[bad]) / np.maximum(np.abs(b[bad]), 1e-300))
            raise AssertionError(f"{name}: {bad.sum()} values differ (first at {k}: {a[tuple(k)]!r} vs "
                                 f"{b[tuple(k)]!r}; max rel {rel:.3e})")
    else:
        assert np.array_equal(a, b), f"{name}: integer arrays differ"


@pytest.mark.parametrize("n_steps", [1, 3])
def test_adv_steps_bit_exact(gpu, n_steps):
    E, lnR0, lnf, nb = load_kernel_data_Ecol("Long_Bott")
    cfg = _config(E, lnR0, lnf, nb)
    grid, *parts = _case()
    parts = _prep(grid, parts)
    ref_parts, scal, mat, wr, no_cols = _oracle_run(grid, parts, cfg, E, 11, n_steps)
    eng = _engine(grid, parts, cfg, E, lnR0, lnf, 11)
    eng.adv_steps(n_steps)
    got = eng.export_particles(0)
    for nm, a, b in zip(["pos", "vel", "cells", "m_w", "m_s", "xi", "active"], got, ref_parts):
        _assert_same(a, b, nm)
    s, m = eng.export_grid(0)
    _assert_same(s[:8], scal[:8], "grid_scalar_fields")
    _assert_same(m, mat, "grid_mat_prop")
    nc_got, wr_got = eng.counters(0)
    _assert_same(nc_got, no_cols, "no_cols")
    _assert_same(wr_got, wr, "water_removed")
    assert no_cols.sum() > 0, "case should produce collisions"


@pytest.mark.parametrize("lists", ["capacity_1", "off"])
def test_migrant_list_fallbacks_match_oracle(gpu, lists):
    """The members of a cell come from its home segment and migrant list while the list is complete, else from
    the 3x3 neighbourhood gather (collision passes, Delta m_l, the reindex's count and scatter).  With lists of
    one entry most cells that received a SIP overflow theirs and mix both paths; with the lists switched off
    every kernel gathers.  Three advection steps with collisions stay bit-identical to the oracle."""
    E, lnR0, lnf, nb = load_kernel_data_Ecol("Long_Bott")
    cfg = _config(E, lnR0, lnf, nb)
    grid, *parts = _case(seed=41)
    parts = _prep(grid, parts)
    ref_parts, scal, mat, wr, no_cols = _oracle_run(grid, parts, cfg, E, 19, 3)
    eng = _engine(grid, parts, cfg, E, lnR0, lnf, 19)
    if lists == "capacity_1":
        eng.e.mig_cap = 1
    else:
        eng.e.mig_cnt = None
    eng.adv_steps(3)
    for nm, a, b in zip(["pos", "vel", "cells", "m_w", "m_s", "xi", "active"], eng.export_particles(0), ref_parts):
        _assert_same(a, b, f"{lists} {nm}")
    s_, m_ = eng.export_grid(0)
    _assert_same(s_[:8], scal[:8], f"{lists} grid_scalar_fields")
    _assert_same(m_, mat, f"{lists} grid_mat_prop")
    nc_got, wr_got = eng.counters(0)
    _assert_same(nc_got, no_cols, f"{lists} no_cols")
    _assert_same(wr_got, wr, f"{lists} water_removed")
    assert no_cols.sum() > 0


@pytest.mark.parametrize("variant", ["NaCl", "stress150", "one_collision_step", "four_collision_steps_no_relaxation"])
def test_adv_steps_configurations(gpu, variant):
    """Two advection steps, bit-exact vs the oracle, for the other configurations the step supports: NaCl
    solute (its mass rate and density laws, integration.py:728-836 keeping the AS constants inside the
    iterations); the stress case's parameters (10 m cells, Hall table, dt 0.5 s, 20 substeps, mass-flux
    amplitude 1.0); one collision step per advection step (placed at each subloop start,
    integration.py:1247-1251); four per step (one per substep, integration.py:1257-1262) without
    relaxation."""
    kernel = "Hall_Bott" if variant == "stress150" else "Long_Bott"
    E, lnR0, lnf, nb = load_kernel_data_Ecol(kernel)
    kw, case_kw = {}, {}
    if variant == "NaCl":
        kw = dict(solute_type="NaCl")
    elif variant == "stress150":
        kw = dict(dt_adv=0.5, dt_col=0.25, scale_dt_cond=10)
        case_kw = dict(dx=10.0, dz=10.0, spc=48, j_max_ref=1.0)
    elif variant == "one_collision_step":
        kw = dict(no_col_per_adv=1, dt_col=1.0)
    else:
        kw = dict(no_col_per_adv=4, dt_col=0.25, act_relaxation=False)
    cfg = _config(E, lnR0, lnf, nb, **kw)
    grid, *parts = _case(seed=29, **case_kw)
    parts = _prep(grid, parts)
    ref_parts, scal, mat, wr, no_cols = _oracle_run(grid, parts, cfg, E, 13, 2)
    eng = _engine(grid, parts, cfg, E, lnR0, lnf, 13)
    eng.adv_steps(2)
    for nm, a, b in zip(["pos", "vel", "cells", "m_w", "m_s", "xi", "active"], eng.export_particles(0), ref_parts):
        _assert_same(a, b, f"{variant} {nm}")
    s, m = eng.export_grid(0)
    _assert_same(s[:8], scal[:8], f"{variant} grid_scalar_fields")
    _assert_same(m, mat, f"{variant} grid_mat_prop")
    nc_got, wr_got = eng.counters(0)
    _assert_same(nc_got, no_cols, f"{variant} no_cols")
    _assert_same(wr_got, wr, f"{variant} water_removed")
    assert no_cols.sum() > 0, "case should produce collisions"


def test_collision_step_injected_stream(gpu):
    """All-cells collision step with the reference's draw stream injected (x-major cells, row-major pairs)."""
    from lacmopy_amd.engine import Engine  # noqa: F401
    E, lnR0, lnf, nb = load_kernel_data_Ecol("Hall_Bott")
    cfg = _config(E, lnR0, lnf, nb)
    grid, *parts = _case(nx=6, nz=12, spc=40, seed=9)
    parts = _prep(grid, parts)
    pos, vel, cells, m_w, m_s, xi, active = [np.array(a, copy=True) for a in parts]
    rng = np.random.default_rng(3)
    rnd = rng.random(2_000_000)
    no_cols = np.zeros(2, dtype=np.int64)
    orc.collision_all_cells(xi, m_w, m_s, vel, grid.temperature.copy(), cells, grid.no_cells, 0.5 / grid.volume_cell,
                            E, nb, lnR0, lnf, no_cols, rnd_stream=rnd)
    eng = _engine(grid, parts, cfg, E, lnR0, lnf, 0, rnd=rnd)
    eng.collision_step()
    got = eng.export_particles(0)
    _assert_same(got[5], xi, "xi")
    _assert_same(got[3], m_w, "m_w")
    _assert_same(got[4], m_s, "m_s")
    _assert_same(eng.counters(0)[0], no_cols, "no_cols")
    assert no_cols.sum() > 0


def test_batched_runs_match_single(gpu):
    """Runs of a batch are independent: run r of a batch == a single run with the same run index."""
    E, lnR0, lnf, nb = load_kernel_data_Ecol("Long_Bott")
    cfg = _config(E, lnR0, lnf, nb)
    grid, *parts = _case(nx=8, nz=8, spc=20, seed=21)
    parts = _prep(grid, parts)
    eng = _engine(grid, parts, cfg, E, lnR0, lnf, 7, n_runs=3)
    eng.adv_steps(2)
    ref_parts, scal, mat, wr, no_cols = _oracle_run(grid, parts, cfg, E, 7, 2)
    # run 0 of the batch uses Philox run word 0, like the oracle
    got = eng.export_particles(0)
    for nm, a, b in zip(["pos", "vel", "cells", "m_w", "m_s", "xi", "active"], got, ref_parts):
        _assert_same(a, b, nm)
    # other runs differ only by their draws
    nc1 = eng.counters(1)[0]
    assert nc1.sum() > 0


@pytest.mark.parametrize("spec_events", [0, 2])
@pytest.mark.parametrize("kernel", ["Long_Bott", "const"])
def test_collision_many_events(gpu, kernel, spec_events):
    """Dense rain-like cells (many events per cell): exercises the candidate-queue passes and the
    anti-diagonal continuation of the collision kernel (spec_events = 2: after the second event of a cell);
    Philox draws, bit-exact vs the oracle."""
    if kernel == "const":
        E, lnR0, lnf, nb = 0.9, 0.0, 1.0, 0
    else:
        E, lnR0, lnf, nb = load_kernel_data_Ecol("Long_Bott")
    cfg = _config(E, lnR0, lnf, nb)
    grid, *parts = make_case(nx=6, nz=6, spc=60, seed=17, cloud_base=0.0, cloud_top=120.0, R_drop=25.0,
                             sigma_drop=1.6)
    parts = list(_prep(grid, parts))
    rng = np.random.default_rng(5)
    parts[1] = np.array([parts[1][0], -rng.uniform(0.0, 3.0, parts[1].shape[1])])  # fall speeds
    parts[5] = parts[5] * rng.uniform(3.0, 8.0, parts[5].size)                     # dense multiplicities
    parts = tuple(parts)
    pos, vel, cells, m_w, m_s, xi, active = [np.array(a, copy=True) for a in parts]
    eng = _engine(grid, parts, cfg, E, lnR0, lnf, 99)
    eng.e.spec_events = spec_events
    no_cols = np.zeros(2, dtype=np.int64)
    for step in range(3):
        orc.collision_all_cells(xi, m_w, m_s, vel, grid.temperature.copy(), cells, grid.no_cells,
                                cfg["dt_col"] / grid.volume_cell, E, nb, lnR0, lnf, no_cols, seed=99, run=0,
                                step=step)
        eng.collision_step()
    got = eng.export_particles(0)
    _assert_same(got[5], xi, "xi")
    _assert_same(got[3], m_w, "m_w")
    _assert_same(got[4], m_s, "m_s")
    _assert_same(eng.counters(0)[0], no_cols, "no_cols")
    assert no_cols[0] > 36 * 3 * 20, f"expected many events per cell, got {no_cols}"
    assert np.isfinite(xi).all()


@pytest.mark.parametrize("kernel", ["Long_Bott", "const"])
def test_equal_multiplicities_follow_the_tie_rule(gpu, kernel):
    """Cells whose SIPs all share one multiplicity: every event is a near-equal collection
    (all_or_nothing.py:597-614) on an exact tie, where xi_i <= xi_j makes i the smaller SIP
    (all_or_nothing.py:565-568) and decides which SIP keeps 0.35 and which 0.15 of the summed
    multiplicity; three collision steps bit-exact vs the oracle, row filter included."""
    if kernel == "const":
        E, lnR0, lnf, nb = 0.9, 0.0, 1.0, 0
    else:
        E, lnR0, lnf, nb = load_kernel_data_Ecol("Long_Bott")
    cfg = _config(E, lnR0, lnf, nb)
    grid, *parts = make_case(nx=6, nz=6, spc=50, seed=31, cloud_base=0.0, cloud_top=120.0, R_drop=30.0,
                             sigma_drop=1.6)
    parts = list(_prep(grid, parts))
    rng = np.random.default_rng(8)
    parts[1] = np.array([parts[1][0], -rng.uniform(0.0, 3.0, parts[1].shape[1])])  # fall speeds
    parts[5] = np.full(parts[5].size, 4.0E9)                                       # one multiplicity
    parts = tuple(parts)
    pos, vel, cells, m_w, m_s, xi, active = [np.array(a, copy=True) for a in parts]
    eng = _engine(grid, parts, cfg, E, lnR0, lnf, 41)
    no_cols = np.zeros(2, dtype=np.int64)
    for step in range(3):
        orc.collision_all_cells(xi, m_w, m_s, vel, grid.temperature.copy(), cells, grid.no_cells,
                                cfg["dt_col"] / grid.volume_cell, E, nb, lnR0, lnf, no_cols, seed=41, run=0,
                                step=step)
        eng.collision_step()
    got = eng.export_particles(0)
    _assert_same(got[5], xi, "xi")
    _assert_same(got[3], m_w, "m_w")
    _assert_same(got[4], m_s, "m_s")
    _assert_same(eng.counters(0)[0], no_cols, "no_cols")
    assert no_cols[0] > 36, f"expected many events, got {no_cols}"
    assert int(eng.buf["n_equal_xi"][0]) > 36, "expected near-equal collections on the ties"


def test_dense_cells_take_the_large_capacity_passes(gpu):
    """Cells of ~300 SIPs exceed the 256-member fast passes of the collision and Delta m_l kernels and go
    through their 1024-member overflow passes; results stay bit-identical to the oracle."""
    E, lnR0, lnf, nb = load_kernel_data_Ecol("Long_Bott")
    cfg = _config(E, lnR0, lnf, nb)
    grid, *parts = _case(nx=4, nz=4, spc=300, seed=13)
    parts = _prep(grid, parts)
    counts = np.bincount(parts[2][0] * 4 + parts[2][1], minlength=16)
    assert counts.min() > 256
    ref_parts, scal, mat, wr, no_cols = _oracle_run(grid, parts, cfg, E, 5, 1)
    eng = _engine(grid, parts, cfg, E, lnR0, lnf, 5)
    eng.adv_steps(1)
    for nm, a, b in zip(["pos", "vel", "cells", "m_w", "m_s", "xi", "active"], eng.export_particles(0), ref_parts):
        _assert_same(a, b, nm)
    s, m = eng.export_grid(0)
    _assert_same(s[:8], scal[:8], "grid_scalar_fields")
    _assert_same(eng.counters(0)[0], no_cols, "no_cols")


def test_one_crowded_cell_takes_the_overflow_passes(gpu):
    """A cell with ~400 SIPs among cells of ~30: the main collision and Delta m_l passes (sized for the mean)
    queue it for their 1024-member overflow passes; results stay bit-identical to the oracle."""
    E, lnR0, lnf, nb = load_kernel_data_Ecol("Long_Bott")
    cfg = _config(E, lnR0, lnf, nb)
    grid, pos, vel, cells, m_w, m_s, xi, act = _case(nx=4, nz=4, spc=30, seed=19)
    rs = np.random.default_rng(4)
    k = 380
    src = rs.integers(0, len(xi), k)
    extra_pos = np.array([rs.uniform(20.0, 40.0, k), rs.uniform(20.0, 40.0, k)])
    parts = (np.concatenate([pos, extra_pos], 1), np.concatenate([vel, vel[:, src]], 1),
             np.concatenate([cells, np.ones((2, k), dtype=np.int64)], 1), np.concatenate([m_w, m_w[src]]),
             np.concatenate([m_s, m_s[src]]), np.concatenate([xi, xi[src]]), np.concatenate([act, np.ones(k, bool)]))
    parts = _prep(grid, parts)
    counts = np.bincount(parts[2][0] * 4 + parts[2][1], minlength=16)
    assert counts[5] > 400 and np.median(counts) < 60
    ref_parts, scal, mat, wr, no_cols = _oracle_run(grid, parts, cfg, E, 8, 1)
    eng = _engine(grid, parts, cfg, E, lnR0, lnf, 8)
    eng.adv_steps(1)
    for nm, a, b in zip(["pos", "vel", "cells", "m_w", "m_s", "xi", "active"], eng.export_particles(0), ref_parts):
        _assert_same(a, b, nm)
    s, m = eng.export_grid(0)
    _assert_same(s[:8], scal[:8], "grid_scalar_fields")
    _assert_same(eng.counters(0)[0], no_cols, "no_cols")


@pytest.mark.parametrize("k", [1400, 2600])
def test_cells_over_1024_members_take_the_any_size_passes(gpu, k):
    """One cell holding about k SIPs among cells of ~30 -- over every LDS capacity: the collision step's
    1024-member pass queues it for the pass over cells of any size (its members merged in id order from the
    3x3 home segments into a global workspace), and the Delta m_l, r_l and derived-field passes sum it on one
    lane straight from the segments.  Two advection steps, then r_l and the derived fields: bit-identical to
    the oracle."""
    E, lnR0, lnf, nb = load_kernel_data_Ecol("Long_Bott")
    cfg = _config(E, lnR0, lnf, nb)
    grid, pos, vel, cells, m_w, m_s, xi, act = _case(nx=4, nz=4, spc=30, seed=19)
    rs = np.random.default_rng(5)
    src = rs.integers(0, len(xi), k)
    extra_pos = np.array([rs.uniform(20.0, 40.0, k), rs.uniform(20.0, 40.0, k)])
    parts = (np.concatenate([pos, extra_pos], 1), np.concatenate([vel, vel[:, src]], 1),
             np.concatenate([cells, np.ones((2, k), dtype=np.int64)], 1), np.concatenate([m_w, m_w[src]]),
             np.concatenate([m_s, m_s[src]]), np.concatenate([xi, xi[src]]), np.concatenate([act, np.ones(k, bool)]))
    parts = _prep(grid, parts)
    counts = np.bincount(parts[2][0] * 4 + parts[2][1], minlength=16)
    assert counts[5] > max(1024, k) and np.median(counts) < 60
    ref_parts, scal, mat, wr, no_cols = _oracle_run(grid, parts, cfg, E, 8, 2)
    eng = _engine(grid, parts, cfg, E, lnR0, lnf, 8)
    eng.adv_steps(2)
    got = eng.export_particles(0)
    for nm, a, b in zip(["pos", "vel", "cells", "m_w", "m_s", "xi", "active"], got, ref_parts):
        _assert_same(a, b, nm)
    s, m = eng.export_grid(0)
    _assert_same(s[:8], scal[:8], "grid_scalar_fields")
    _assert_same(eng.counters(0)[0], no_cols, "no_cols")
    assert no_cols.sum() > 0
    pos2, vel2, cells2, m_w2, m_s2, xi2, act2 = got
    assert np.bincount(cells2[0, act2] * 4 + cells2[1, act2], minlength=16).max() > 1024
    eng.update_r_l()
    s, _ = eng.export_grid(0)
    ref = np.zeros((4, 4))
    orc.lib().or_update_grid_r_l(len(xi2), orc.P(m_w2), orc.P(xi2), orc.P(cells2), orc.P(act2.view(np.uint8)), 4, 4,
                                 orc.P(ref), orc.P(s[8].copy()))
    _assert_same(s[5], ref, "r_l")
    from lacmopy_amd.evaluation import derived_fields
    d = derived_fields(eng)
    _assert_same(d[0], orc.derived_fields(m_w2, m_s2, xi2, cells2, act2, s[0].copy(), s[8].copy(),
                                          grid.volume_cell), "derived fields")
    # the adjoint's objective sums (compensated, per cell in ascending id; the crowded cell on one lane):
    # J_lwp is the correctly rounded total of xi m_w times its constant
    import math
    from lacmopy_amd.adjoint import AdjointWindow
    J_nc, J_lwp = AdjointWindow(eng, 1).objectives()
    e = eng.e
    c_lwp = 1.0E-18 / ((e.x_max - e.x_min) * (e.V_cell / (e.dx * e.dz)))
    want = c_lwp * math.fsum((xi2 * m_w2)[act2])
    assert abs(J_lwp[0] - want) <= 2e-16 * abs(want), (J_lwp[0], want)


@pytest.mark.parametrize("order", ["cost", "size"])
@pytest.mark.parametrize("work_per_cu,chunk", [(1, 0), (1, 1), (4, 3)])
def test_persistent_work_list_passes_match_oracle(gpu, work_per_cu, chunk, order):
    """Persistent collision and Delta m_l passes over the (run, cell) work list: 6 runs of 8x8 cells (384
    work items, more than one per workgroup at one workgroup per compute unit); run 0 stays bit-exact with
    the oracle over three advection steps, the other runs equal the one-workgroup-per-cell launch, and the
    work list is a permutation of the cells ordered by decreasing cost: the last collision step's ticks
    per cell for the cells above half the expected pass length (col_cost kept; monotone up to the bins of 16
    per octave), then the home-segment size."""
    E, lnR0, lnf, nb = load_kernel_data_Ecol("Long_Bott")
    cfg = _config(E, lnR0, lnf, nb)
    grid, *parts = _case(nx=8, nz=8, spc=24, seed=23)
    parts = _prep(grid, parts)
    ref_parts, scal, mat, wr, no_cols = _oracle_run(grid, parts, cfg, E, 3, 3)
    eng = _engine(grid, parts, cfg, E, lnR0, lnf, 3, n_runs=6)
    if order == "size":
        eng.e.col_cost = None
    eng.set_launch_shape(0, 0, 0, work_per_cu, work_per_cu, chunk, chunk)
    base = _engine(grid, parts, cfg, E, lnR0, lnf, 3, n_runs=6)
    ncell = 6 * 64
    for step in range(3):
        eng.adv_steps(1)
        base.adv_steps(1)
        if order == "cost":
            eng.build_work_list()  # from the costs of the step's collision passes, as they are now
        work = eng.buf["work"][:ncell].cpu().numpy()
        assert np.array_equal(np.sort(work), np.arange(ncell)), "work list is not a permutation of the cells"
        seg = eng.buf[f"seg{int(eng.e.cur)}"].cpu().numpy()
        r, c = work // 64, work % 64
        size = seg[r * 65 + c + 1] - seg[r * 65 + c]
        if order == "size":
            assert np.all(np.diff(size) <= 0), "work list not ordered by decreasing home-segment size"
        else:
            cost = eng.buf["col_cost"].cpu().numpy().view(np.uint32).astype(np.float64)[work]
            assert cost.max() > 0, "no collision costs recorded"
            # a head of the cells above a cost threshold, by decreasing cost (bins of 16 per octave), then the
            # rest by decreasing size: some split k has every head cost above every tail cost
            head_min = np.minimum.accumulate(cost)
            tail_max = np.maximum.accumulate(cost[::-1])[::-1]
            ok = [k for k in range(len(cost) + 1)
                  if (k == 0 or k == len(cost) or head_min[k - 1] > tail_max[k])
                  and np.all(np.diff(size[k:]) <= 0)
                  and np.all(cost[1:k] <= cost[:k - 1] * (17.0 / 16.0) + 16.0)]
            assert ok, "work list is not a cost-ordered head followed by a size-ordered tail"
    for nm, a, b in zip(["pos", "vel", "cells", "m_w", "m_s", "xi", "active"], eng.export_particles(0), ref_parts):
        _assert_same(a, b, nm)
    s, m = eng.export_grid(0)
    _assert_same(s[:8], scal[:8], "grid_scalar_fields")
    _assert_same(eng.counters(0)[0], no_cols, "no_cols")
    for r in range(1, 6):
        for nm, a, b in zip(["pos", "vel", "cells", "m_w", "m_s", "xi", "active"], eng.export_particles(r),
                            base.export_particles(r)):
            _assert_same(a, b, f"run {r} {nm}")
        _assert_same(eng.counters(r)[0], base.counters(r)[0], f"run {r} no_cols")
    assert no_cols.sum() > 0


def test_launch_shapes_and_autotune_leave_results_unchanged(gpu):
    """The particle kernel's block size and SIPs per thread and the collision main-pass capacity change
    only how the work is launched: the state after two advection steps is the same bit for bit, and
    Engine.autotune restores the state it measured on."""
    E, lnR0, lnf, nb = load_kernel_data_Ecol("Long_Bott")
    cfg = _config(E, lnR0, lnf, nb)
    grid, *parts = _case()
    parts = _prep(grid, parts)
    ref = None
    for shape in [(0, 0, 0, 0, 0), (64, 2, 512, 1, 2, 2, 0), (512, 1, 160, 0, 8, 0, 5, 5), (192, 2, 256, 16, 0, 1),
                  (128, 1, 0, 0, 0, 0, 0, 6), (256, 1, 0, 0, 0, 0, 0, 8)]:
        eng = _engine(grid, parts, cfg, E, lnR0, lnf, 11)
        eng.set_launch_shape(*shape)
        eng.adv_steps(2)
        got = eng.export_particles(0) + tuple(eng.export_grid(0))
        if ref is None:
            ref = got
            continue
        for nm, a, b in zip(["pos", "vel", "cells", "m_w", "m_s", "xi", "active", "scal", "mat"], got, ref):
            _assert_same(a, b, f"{nm} with launch shape {shape}")
    eng = _engine(grid, parts, cfg, E, lnR0, lnf, 11)
    eng.adv_steps(1)
    before = {k: v.cpu().numpy() for k, v in eng.snapshot()["buf"].items()}
    res = eng.autotune(reps=1)
    after = {k: v.cpu().numpy() for k, v in eng.snapshot()["buf"].items()}
    for k in before:
        _assert_same(after[k], before[k], f"{k} after autotune")
    ch = res["chosen"]
    assert ch["part_block"] in (64, 128, 256, 512) and ch["part_spt"] in (1, 2)
    assert ch["col_work_per_cu"] in (0, 16, 32) and ch["dml_work_per_cu"] in (0, 32, 64)


def _with_large_drops(grid, parts, seed, per_cell=(0, 7), R_um=(60.0, 320.0)):
    """Turns a few SIPs of every cell into drizzle drops (radius R_um, multiplicity / 200, falling at roughly
    their terminal speed): every row bound of the cell widens, as in the collision-active state."""
    pos, vel, cells, m_w, m_s, xi, act = [np.array(a, copy=True) for a in parts]
    rs = np.random.default_rng(seed)
    key = cells[0] * int(grid.no_cells[1]) + cells[1]
    for c in np.unique(key[act]):
        mem = np.flatnonzero((key == c) & act)
        k = min(int(rs.integers(per_cell[0], per_cell[1] + 1)), len(mem) // 4)
        pick = rs.choice(mem, k, replace=False)
        R = rs.uniform(R_um[0], R_um[1], k)
        m_w[pick] = 4.0 / 3.0 * np.pi * 1000.0 * (R * 1.0E-6) ** 3 * 1.0E18
        xi[pick] = xi[pick] / 200.0
        vel[1, pick] -= 0.008 * R
    return pos, vel, cells, m_w, m_s, xi, act


@pytest.mark.parametrize("spec_events", [0, 3])
@pytest.mark.parametrize("kernel", ["Long_Bott", "Hall_Bott"])
def test_cells_with_large_drops_match_oracle(gpu, kernel, spec_events):
    """Collision-active cells (~90 members, up to 7 drops of 60-320 um among cloud droplets): the guard pass
    leaves the drops' pairs to the per-pair pass, the candidate queue is carried from event to event, and
    cells past spec_events events (3: most event cells) finish with the filtered anti-diagonal schedule.  The
    all-cells drop-in (reference layout, member lists) for one collision step and the engine for two advection
    steps (autotuned capacities) stay bit-identical to the oracle."""
    from lacmopy_amd.collision.all_or_nothing import collision_step_Ecol_grid_R_all_cells_2D_multicomp_np as step
    E, lnR0, lnf, nb = load_kernel_data_Ecol(kernel)
    grid, *parts = _case(nx=6, nz=6, spc=90, seed=23, R_drop=12.0, sigma_drop=1.3)
    parts = _prep(grid, parts)
    parts = _with_large_drops(grid, parts, 5)
    pos, vel, cells, m_w, m_s, xi, act = parts
    T = grid.temperature.copy()
    o_xi, o_mw, o_ms, o_nc = xi.copy(), m_w.copy(), m_s.copy(), np.zeros(2, dtype=np.int64)
    orc.collision_all_cells(o_xi, o_mw, o_ms, vel, T, cells, grid.no_cells, 0.5 / grid.volume_cell, E, nb, lnR0, lnf,
                            o_nc, seed=3, run=0, step=1)
    assert o_nc[0] > 200  # many events: several cells take the anti-diagonal schedule
    g_xi, g_mw, g_ms, g_nc = xi.copy(), m_w.copy(), m_s.copy(), np.zeros(2, dtype=np.int64)
    rng.seed(3, 1)
    step(g_xi, g_mw, g_ms, vel, T, cells, grid.no_cells, 0.5 / grid.volume_cell, E, nb, lnR0, lnf, g_nc, "AS")
    _assert_same(g_xi, o_xi, "xi (one collision step)")
    _assert_same(g_mw, o_mw, "m_w (one collision step)")
    _assert_same(g_ms, o_ms, "m_s (one collision step)")
    _assert_same(g_nc, o_nc, "no_cols (one collision step)")
    cfg = _config(E, lnR0, lnf, nb)
    ref_parts, scal, mat, wr, no_cols = _oracle_run(grid, parts, cfg, E, 6, 2)
    eng = _engine(grid, parts, cfg, E, lnR0, lnf, 6)
    eng.e.spec_events = spec_events
    eng.autotune()
    eng.adv_steps(2)
    for nm, a, b in zip(["pos", "vel", "cells", "m_w", "m_s", "xi", "active"], eng.export_particles(0), ref_parts):
        _assert_same(a, b, nm)
    _assert_same(eng.counters(0)[0], no_cols, "no_cols")
    assert no_cols[0] > 400


@pytest.mark.parametrize("spc", [200, 300])
def test_collision_capacity_tiers_match_oracle(gpu, spc):
    """Main collision pass at 160 members: cells of ~200 members go to the 256-member pass, cells of ~300 on to
    the 1024-member pass; one advection step stays bit-identical to the oracle."""
    E, lnR0, lnf, nb = load_kernel_data_Ecol("Long_Bott")
    cfg = _config(E, lnR0, lnf, nb)
    grid, *parts = _case(nx=4, nz=4, spc=spc, seed=31)
    parts = _prep(grid, parts)
    counts = np.bincount(parts[2][0] * 4 + parts[2][1], minlength=16)
    assert counts.min() > 160 and (counts.max() <= 256 if spc == 200 else counts.min() > 256)
    ref_parts, scal, mat, wr, no_cols = _oracle_run(grid, parts, cfg, E, 17, 1)
    eng = _engine(grid, parts, cfg, E, lnR0, lnf, 17)
    eng.set_launch_shape(col_cap=160)
    eng.adv_steps(1)
    for nm, a, b in zip(["pos", "vel", "cells", "m_w", "m_s", "xi", "active"], eng.export_particles(0), ref_parts):
        _assert_same(a, b, nm)
    _assert_same(eng.counters(0)[0], no_cols, "no_cols")
    assert no_cols.sum() > 0
