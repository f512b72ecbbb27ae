"""GPU parity of the drop-in entry points (reference names and signatures) vs the CPU oracle.

Each drop-in function takes the reference's numpy arrays, runs the HIP
kernels and writes the results back in place; the oracle runs the same
function of the reference on copies.  Everything is compared bit for bit.
"""
import numpy as np
import pytest

from lacmopy_amd import load_kernel_data_Ecol, rng
from lacmopy_amd.setup_case import init_profiles
from oracle import oracle as orc
from test_gpu_parity import _assert_same, _case, _config, _oracle_run, _prep

pytestmark = pytest.mark.gpu

NAMES = ["pos", "vel", "cells", "m_w", "m_s", "xi", "active"]


def _grid_arrays(grid):
    g = grid
    g.rho_dry_inv = np.ones_like(g.mass_density_air_dry) / g.mass_density_air_dry
    g.mass_dry_inv = (1.0 / g.volume_cell) * g.rho_dry_inv
    return g.scalar_fields(), g.mat_prop()


def _rel_pos(grid, pos, cells, active):
    """integration.py:545-549: the relative positions a consistent caller holds."""
    rel = np.zeros(pos.shape)
    for d in (0, 1):
        rel[d][active] = (pos[d][active] - grid.ranges[d][0]) / grid.steps[d] - cells[d][active]
    return rel


def _oracle_params(grid, N, cfg, E, seed, step):
    return orc.adv_params(grid, N, cfg, seed=seed, col_step=step, E_col_grid=E)


@pytest.mark.parametrize("kernel", ["Long_Bott", "const"])
def test_integrate_adv_step_np(gpu, kernel):
    from lacmopy_amd.integration import integrate_adv_step_np
    if kernel == "const":
        E, lnR0, lnf, nb = 0.5, None, None, None
    else:
        E, lnR0, lnf, nb = load_kernel_data_Ecol(kernel)
    cfg = _config(E, lnR0, lnf, nb)
    grid, *parts = _case(nx=8, nz=8, spc=24, seed=21)
    parts = _prep(grid, parts)
    ref_parts, ref_scal, ref_mat, ref_wr, ref_nc = _oracle_run(grid, parts, cfg, E, 77, 2)
    # oracle T_p / rel_pos of the same run
    pos, vel, cells, m_w, m_s, xi, active = [np.array(a, copy=True) for a in parts]
    N = len(xi)
    scal, mat = grid.scalar_fields(), grid.mat_prop()
    rel_pos = np.zeros((2, N))
    for d in (0, 1):
        rel_pos[d][active] = (pos[d][active] - grid.ranges[d][0]) / grid.steps[d] - cells[d][active]
    T_p = np.ones(N)
    wr = np.zeros(1)
    no_cols = np.zeros(2, dtype=np.int64)
    dml = np.zeros((8, 8))
    prof = init_profiles(grid)
    rng.seed(77)
    dt = 1.0
    dt_sub = dt / 10
    for _ in range(2):
        integrate_adv_step_np(scal, mat, grid.velocity, grid.mass_flux_air_dry, grid.p_ref, grid.p_ref_inv,
                              grid.no_cells, grid.ranges, grid.steps, grid.volume_cell, pos, vel, cells, rel_pos, m_w,
                              m_s, xi, "AS", wr, np.arange(N), active, T_p, dml, dt, dt_sub, 0.5 * dt_sub,
                              0.5 / grid.volume_cell, 5, 2, True, True, prof[0], prof[1], prof[2], 3, 9.80665, E, nb,
                              lnR0, lnf, no_cols)
    for nm, a, b in zip(NAMES, (pos, vel, cells, m_w, m_s, xi, active), ref_parts):
        _assert_same(a, b, nm)
    _assert_same(scal[:8], ref_scal[:8], "grid_scalar_fields")
    _assert_same(mat, ref_mat, "grid_mat_prop")
    _assert_same(wr, ref_wr, "water_removed")
    _assert_same(no_cols, ref_nc, "no_cols")
    assert rng.get_step() == 4
    # T_p and rel_pos against the oracle's own arrays
    o_pos, o_vel, o_cells, o_mw, o_ms, o_xi, o_act = [np.array(a, copy=True) for a in parts]
    o_rel = np.zeros((2, N))
    o_Tp = np.ones(N)
    p = _oracle_params(grid, N, cfg, E, 77, 0)
    orc.integrate_adv_steps(p, 2, o_pos, o_vel, o_cells, o_rel, o_mw, o_ms, o_xi, o_act, o_Tp, np.zeros(1),
                            np.zeros(2, dtype=np.int64), grid.scalar_fields(), grid.mat_prop(), grid.velocity,
                            grid.mass_flux_air_dry, E, *prof)
    _assert_same(T_p, o_Tp, "T_p")
    _assert_same(rel_pos[:, active], o_rel[:, o_act], "rel_pos (active)")


def test_propagate_particles_and_grid_substep(gpu):
    from lacmopy_amd.integration import propagate_grid_subloop_step_np, propagate_particles_subloop_step_np
    grid, *parts = _case(nx=8, nz=8, spc=24, seed=23)
    parts = _prep(grid, parts)
    scal, mat = _grid_arrays(grid)
    N = len(parts[5])
    cfg = _config(0.5, 0, 1, 0)
    # GPU
    pos, vel, cells, m_w, m_s, xi, active = [np.array(a, copy=True) for a in parts]
    rel = _rel_pos(grid, pos, cells, active)
    T_p = np.ones(N)
    wr = np.zeros(1)
    dml = np.zeros((8, 8))
    propagate_particles_subloop_step_np(scal, mat, grid.velocity, grid.no_cells, grid.ranges, grid.steps, pos, vel,
                                        cells, rel, m_w, m_s, xi, wr, np.arange(N), active, T_p, "AS", dml, 0.1,
                                        0.1, 3, 9.80665)
    # oracle
    o = [np.array(a, copy=True) for a in parts]
    o_rel, o_Tp, o_wr, o_dml = _rel_pos(grid, o[0], o[2], o[6]), np.ones(N), np.zeros(1), np.zeros((8, 8))
    p = _oracle_params(grid, N, cfg, 0.5, 0, 0)
    s0, m0 = scal.copy(), mat.copy()
    orc.lib().or_propagate_particles_substep(orc.C.byref(p), 0.1, 0.1, *[orc.P(a) for a in (
        o[0], o[1], o[2], o_rel, o[3], o[4], o[5])], orc.P(o[6].view(np.uint8)), orc.P(o_Tp), orc.P(o_wr),
        orc.P(s0), orc.P(m0), orc.P(grid.velocity), orc.P(o_dml))
    for nm, a, b in zip(NAMES, (pos, vel, cells, m_w, m_s, xi, active), o):
        _assert_same(a, b, nm)
    _assert_same(T_p, o_Tp, "T_p")
    _assert_same(dml, o_dml, "delta_m_l")
    _assert_same(wr, o_wr, "water_removed")
    assert np.abs(dml).sum() > 0
    # grid substep with these increments
    rs = np.random.default_rng(4)
    dTh = rs.normal(0, 1e-3, (8, 8))
    dRv = rs.normal(0, 1e-6, (8, 8))
    g_s, g_m = scal.copy(), mat.copy()
    propagate_grid_subloop_step_np(g_s, g_m, grid.p_ref, grid.p_ref_inv, dTh, dRv, dml, grid.volume_cell)
    o_s, o_m = scal.copy(), mat.copy()
    orc.lib().or_propagate_grid_subloop_step(orc.P(o_s), orc.P(o_m), grid.p_ref, grid.p_ref_inv, orc.P(dTh),
                                             orc.P(dRv), orc.P(o_dml), grid.volume_cell, 64)
    _assert_same(g_s, o_s, "grid_scalar_fields")
    _assert_same(g_m, o_m, "grid_mat_prop")


def test_integrate_subloop_n_steps_np(gpu):
    from lacmopy_amd.integration import integrate_subloop_n_steps_np
    E, lnR0, lnf, nb = load_kernel_data_Ecol("Long_Bott")
    grid, *parts = _case(nx=8, nz=8, spc=24, seed=25)
    parts = _prep(grid, parts)
    scal, mat = _grid_arrays(grid)
    N = len(parts[5])
    cfg = _config(E, lnR0, lnf, nb)
    rs = np.random.default_rng(5)
    dTh = rs.normal(0, 1e-3, (8, 8))
    dRv = rs.normal(0, 1e-6, (8, 8))
    dt_sub, n_cond, n_col = 0.1, 3, 3
    pos, vel, cells, m_w, m_s, xi, active = [np.array(a, copy=True) for a in parts]
    rel, T_p, wr, nc, dml = (_rel_pos(grid, pos, cells, active), np.ones(N), np.zeros(1), np.zeros(2, dtype=np.int64),
                             np.zeros((8, 8)))
    g_s, g_m = scal.copy(), mat.copy()
    rng.seed(31, 10)
    integrate_subloop_n_steps_np(g_s, g_m, grid.velocity, grid.no_cells, grid.ranges, grid.steps, grid.volume_cell,
                                 grid.p_ref, grid.p_ref_inv, pos, vel, cells, rel, m_w, m_s, xi, "AS",
                                 0.5 / grid.volume_cell, E, nb, lnR0, lnf, nc, wr, np.arange(N), active, T_p, dml,
                                 dTh, dRv, dt_sub, dt_sub, n_cond, n_col, 3, 9.80665, True)
    assert rng.get_step() == 13
    # oracle: integration.py:1247-1292 composed from its pieces
    o = [np.array(a, copy=True) for a in parts]
    o_rel, o_Tp, o_wr, o_nc, o_dml = (_rel_pos(grid, o[0], o[2], o[6]), np.ones(N), np.zeros(1),
                                      np.zeros(2, np.int64), np.zeros((8, 8)))
    o_s, o_m = scal.copy(), mat.copy()
    p = _oracle_params(grid, N, cfg, E, 31, 0)
    for k in range(n_cond):
        orc.collision_all_cells(o[5], o[3], o[4], o[1], o_s[0].copy(), o[2], grid.no_cells, 0.5 / grid.volume_cell, E,
                                nb, lnR0, lnf, o_nc, seed=31, run=0, step=10 + k)
        orc.lib().or_propagate_particles_substep(orc.C.byref(p), dt_sub, dt_sub, *[orc.P(a) for a in (
            o[0], o[1], o[2], o_rel, o[3], o[4], o[5])], orc.P(o[6].view(np.uint8)), orc.P(o_Tp), orc.P(o_wr),
            orc.P(o_s), orc.P(o_m), orc.P(grid.velocity), orc.P(o_dml))
        orc.lib().or_propagate_grid_subloop_step(orc.P(o_s), orc.P(o_m), grid.p_ref, grid.p_ref_inv, orc.P(dTh),
                                                 orc.P(dRv), orc.P(o_dml), grid.volume_cell, 64)
    for nm, a, b in zip(NAMES, (pos, vel, cells, m_w, m_s, xi, active), o):
        _assert_same(a, b, nm)
    _assert_same(g_s, o_s, "grid_scalar_fields")
    _assert_same(g_m, o_m, "grid_mat_prop")
    _assert_same(T_p, o_Tp, "T_p")
    _assert_same(dml, o_dml, "delta_m_l")
    _assert_same(nc, o_nc, "no_cols")
    assert o_nc.sum() > 0


@pytest.mark.parametrize("kernel", ["Hall_Bott", "const"])
def test_collision_all_cells_np(gpu, kernel):
    from lacmopy_amd.collision.all_or_nothing import collision_step_Ecol_grid_R_all_cells_2D_multicomp_np as step
    if kernel == "const":
        E, lnR0, lnf, nb = 0.5, None, None, None
    else:
        E, lnR0, lnf, nb = load_kernel_data_Ecol(kernel)
    grid, *parts = _case(nx=6, nz=12, spc=40, seed=9)
    pos, vel, cells, m_w, m_s, xi, active = _prep(grid, parts)
    N = len(xi)
    cells = cells.copy()
    cells[1, ::17] = -1  # inactive SIPs take no part
    vel = vel.copy()
    vel[1] -= np.random.default_rng(1).uniform(0.0, 3.0, N)  # fall speeds, so that pairs collide
    T = grid.temperature.copy()
    o_xi, o_mw, o_ms, o_nc = xi.copy(), m_w.copy(), m_s.copy(), np.zeros(2, dtype=np.int64)
    orc.collision_all_cells(o_xi, o_mw, o_ms, vel, T, cells, grid.no_cells, 0.5 / grid.volume_cell, E, nb or 0,
                            lnR0 or 0.0, lnf or 1.0, o_nc, seed=5, run=0, step=3)
    rng.seed(5, 3)
    nc = np.zeros(2, dtype=np.int64)
    step(xi, m_w, m_s, vel, T, cells, grid.no_cells, 0.5 / grid.volume_cell, E, nb, lnR0, lnf, nc, "AS")
    _assert_same(xi, o_xi, "xi")
    _assert_same(m_w, o_mw, "m_w")
    _assert_same(m_s, o_ms, "m_s")
    _assert_same(nc, o_nc, "no_cols")
    assert o_nc.sum() > 0 and N > 0


@pytest.mark.parametrize("variant", ["device_tensors", "injected_stream", "crowded_cell", "cell_of_2000",
                                     "cell_of_2000_injected"])
def test_collision_all_cells_on_device_tensors(gpu, variant):
    """The all-cells drop-in on CUDA tensors of the reference's shapes updates them where they are (same
    storage, no copies) and matches the oracle bit for bit; with an injected uniform stream (consumed in the
    reference's order); with one cell of 700 SIPs (the 1024-member pass); with one cell of 2000 SIPs (the pass
    over cells of any size, its members ranked by SIP index in global memory), with Philox draws and with an
    injected stream."""
    import torch
    from lacmopy_amd.collision.all_or_nothing import collision_step_Ecol_grid_R_all_cells_2D_multicomp_np as step
    E, lnR0, lnf, nb = load_kernel_data_Ecol("Long_Bott")
    grid, *parts = _case(nx=6, nz=12, spc=40, seed=19)
    pos, vel, cells, m_w, m_s, xi, active = _prep(grid, parts)
    cells = cells.copy()
    if variant == "crowded_cell":
        cells[:, :700] = np.array([[2], [5]])
    if variant.startswith("cell_of_2000"):
        cells[:, 300:2300] = np.array([[2], [5]])
    cells[1, ::23] = -1
    N = len(xi)
    vel = vel.copy()
    vel[1] -= np.random.default_rng(4).uniform(0.0, 3.0, N)
    T = grid.temperature.copy()
    rnd = None
    if variant in ("injected_stream", "cell_of_2000_injected"):
        act = cells[1] >= 0
        cnt = np.bincount(cells[0, act] * 12 + cells[1, act], minlength=72)
        rnd = np.random.default_rng(8).random(int((cnt * (cnt - 1) // 2).sum()))
    o_xi, o_mw, o_ms, o_nc = xi.copy(), m_w.copy(), m_s.copy(), np.array([3, 1], dtype=np.int64)
    orc.collision_all_cells(o_xi, o_mw, o_ms, vel, T, cells, grid.no_cells, 0.5 / grid.volume_cell, E, nb, lnR0, lnf,
                            o_nc, rnd_stream=rnd, seed=7, run=0, step=4)
    dev = {k: torch.from_numpy(np.ascontiguousarray(a)).cuda() for k, a in
           (("xi", xi), ("m_w", m_w), ("m_s", m_s), ("vel", vel), ("cells", cells))}
    nc = torch.tensor([3, 1], dtype=torch.int64, device="cuda")
    ptrs = {k: t.data_ptr() for k, t in dev.items()}
    rng.seed(7, 4)
    step(dev["xi"], dev["m_w"], dev["m_s"], dev["vel"], T, dev["cells"], grid.no_cells, 0.5 / grid.volume_cell, E, nb,
         lnR0, lnf, nc, "AS", rnd=rnd)
    assert {k: t.data_ptr() for k, t in dev.items()} == ptrs
    _assert_same(dev["xi"].cpu().numpy(), o_xi, "xi")
    _assert_same(dev["m_w"].cpu().numpy(), o_mw, "m_w")
    _assert_same(dev["m_s"].cpu().numpy(), o_ms, "m_s")
    _assert_same(nc.cpu().numpy(), o_nc, "no_cols")
    assert o_nc.sum() > 4


def _cell_sips(n, seed):
    rs = np.random.default_rng(seed)
    xi = rs.uniform(1e4, 1e7, n)
    xi[5] = xi[6] * (1 + 1e-7)  # a near-equal pair
    m_s = rs.uniform(1e-3, 1.0, n)
    m_w = rs.lognormal(np.log(1e6), 1.5, n)
    rho = rs.uniform(1000, 1100, n)
    R = (0.75 / np.pi * (m_w + m_s) / rho) ** (1 / 3)
    vel = rs.normal(0, 2.0, (2, n))
    return xi, m_w, m_s, R, vel, rho


@pytest.mark.parametrize("kernel", ["Long_Bott", "const"])
@pytest.mark.parametrize("injected", [False, True])
def test_collision_single_cell_2D(gpu, kernel, injected):
    from lacmopy_amd.collision import all_or_nothing as aon
    n = 150
    xi, m_w, m_s, R, vel, rho = _cell_sips(n, 8)
    rnd = np.random.default_rng(2).random(n * (n - 1) // 2) if injected else None
    dt_over_dV = 20.0
    o = [a.copy() for a in (xi, m_w, m_s, R)]
    o_nc = np.zeros(2, dtype=np.int64)
    nc = np.zeros(2, dtype=np.int64)
    rng.seed(99, 7)
    if kernel == "const":
        orc.lib().or_collision_cell_2D(n, *[orc.P(a) for a in o], orc.P(vel[0].copy()), orc.P(vel[1].copy()),
                                       orc.P(rho), dt_over_dV, None, 0.5, 0, 0.0, 1.0, orc.P(o_nc), orc.P(rnd), 99, 0,
                                       7, 0)
        aon.collision_step_Ecol_const_2D_multicomp_np(xi, m_w, m_s, R, vel, rho, dt_over_dV, 0.5, nc, rnd=rnd)
    else:
        E, lnR0, lnf, nb = load_kernel_data_Ecol(kernel)
        orc.lib().or_collision_cell_2D(n, *[orc.P(a) for a in o], orc.P(vel[0].copy()), orc.P(vel[1].copy()),
                                       orc.P(rho), dt_over_dV, orc.P(E), 0.0, nb, lnR0, lnf, orc.P(o_nc), orc.P(rnd),
                                       99, 0, 7, 0)
        aon.collision_step_Ecol_grid_R_2D_multicomp_np(xi, m_w, m_s, R, vel, rho, dt_over_dV, E, nb, lnR0, lnf, nc,
                                                       rnd=rnd)
    for nm, a, b in zip(["xi", "m_w", "m_s", "radii"], (xi, m_w, m_s, R), o):
        _assert_same(a, b, nm)
    _assert_same(nc, o_nc, "no_cols")
    assert o_nc[0] > 0 and o_nc[1] > 0


def test_box_steps_golovin_ecol_kernel_grid(gpu):
    from lacmopy_amd.collision import all_or_nothing as aon
    n = 200
    rs = np.random.default_rng(12)
    xi = rs.uniform(1e5, 1e8, n)
    m = rs.lognormal(np.log(1e7), 1.5, n)
    # Golovin
    x1, m1, nc1 = xi.copy(), m.copy(), np.zeros(2, dtype=np.int64)
    ox, om, onc = xi.copy(), m.copy(), np.zeros(2, dtype=np.int64)
    rng.seed(4, 2)
    aon.collision_step_Golovin_np(x1, m1, 1.0, nc1)
    orc.golovin_step(ox, om, 1.0, onc, seed=4, run=0, step=2)
    _assert_same(x1, ox, "golovin xi")
    _assert_same(m1, om, "golovin m")
    _assert_same(nc1, onc, "golovin no_cols")
    assert onc.sum() > 0
    # Ecol table, 1-D velocities
    E, lnR0, lnf, nb = load_kernel_data_Ecol("Long_Bott")
    rho = np.full(n, 1000.0)
    R = (0.75 / np.pi * m / rho) ** (1 / 3)
    vel = np.array([orc.lib().or_terminal_velocity_Beard(r) for r in R])
    x2, m2, R2, nc2 = xi.copy(), m.copy(), R.copy(), np.zeros(2, dtype=np.int64)
    ox, om, oR, onc = xi.copy(), m.copy(), R.copy(), np.zeros(2, dtype=np.int64)
    rng.seed(4, 3)
    aon.collision_step_Ecol_grid_R_np(x2, m2, R2, vel, rho, 1.0, E, nb, lnR0, lnf, nc2)
    orc.lib().or_collision_ecol_box(n, orc.P(ox), orc.P(om), orc.P(oR), orc.P(vel), orc.P(rho), 1.0, orc.P(E), nb,
                                    lnR0, lnf, orc.P(onc), None, 4, 0, 3)
    _assert_same(x2, ox, "ecol xi")
    _assert_same(m2, om, "ecol m")
    _assert_same(R2, oR, "ecol R")
    _assert_same(nc2, onc, "ecol no_cols")
    assert onc.sum() > 0
    # kernel K(m1, m2) on a log-mass grid: Golovin values tabulated
    nbm = 300
    lnm0, lnfm = np.log(1e2), np.log(1.05)
    mg = np.exp(lnm0 + lnfm * np.arange(nbm))
    Kg = (mg[:, None] + mg[None, :]) * 1.5E-18
    x3, m3, nc3 = xi.copy(), m.copy(), np.zeros(2, dtype=np.int64)
    ox, om, onc = xi.copy(), m.copy(), np.zeros(2, dtype=np.int64)
    from lacmopy_amd.collision.box_model import simulate_collisions_batched
    simulate_collisions_batched([(x3, m3)], (Kg, nbm, lnm0, lnfm), "Golovin", "kernel_grid_m", 1.0, 1.0, 1.0, 1.0,
                                8, step0=0)
    orc.lib().or_collision_kernel_grid_m(n, orc.P(ox), orc.P(om), 1.0, orc.P(Kg), nbm, lnm0, lnfm, orc.P(onc), None,
                                         8, 0, 0)
    _assert_same(x3, ox, "kernel_grid_m xi")
    _assert_same(m3, om, "kernel_grid_m m")


def test_box_model_batched_matches_oracle(gpu):
    """50 boxes x 20 steps of the Ecol table box model with Beard velocities, every snapshot exact."""
    from lacmopy_amd.collision.box_model import simulate_collisions_batched
    E, lnR0, lnf, nb = load_kernel_data_Ecol("Long_Bott")
    rs = np.random.default_rng(40)
    boxes, ref = [], []
    for b in range(50):
        n = int(rs.integers(60, 160))
        xi = rs.uniform(1e5, 1e7, n)
        m = rs.lognormal(np.log(1e7), 1.0, n)
        rho = np.full(n, 1000.0)
        R = (0.75 / np.pi * m / rho) ** (1 / 3)
        vel = np.array([orc.lib().or_terminal_velocity_Beard(r) for r in R])
        boxes.append((xi, m, R, vel, rho))
        ref.append([a.copy() for a in (xi, m, R, vel, rho)])
    xs, ms, times, nc = simulate_collisions_batched(boxes, (E, nb, lnR0, lnf), "Long", "Ecol_grid_R", 1.0, 1.0,
                                                    20.0, 5.0, 1234)
    assert len(times) == 5 and times[-1] == 20.0
    for b in range(50):
        xi, m, R, vel, rho = ref[b]
        onc = np.zeros(2, dtype=np.int64)
        snaps = [(xi.copy(), m.copy())]
        for k in range(20):
            orc.lib().or_collision_ecol_box(len(xi), orc.P(xi), orc.P(m), orc.P(R), orc.P(vel), orc.P(rho), 1.0,
                                            orc.P(E), nb, lnR0, lnf, orc.P(onc), None, 1234, b, k)
            orc.lib().or_update_velocity_Beard(len(xi), orc.P(vel), orc.P(R))
            if (k + 1) % 5 == 0:
                snaps.append((xi.copy(), m.copy()))
        for s in range(5):
            _assert_same(xs[b][s], snaps[s][0], f"box {b} xi snapshot {s}")
            _assert_same(ms[b][s], snaps[s][1], f"box {b} m snapshot {s}")
        _assert_same(boxes[b][3], vel, f"box {b} vel")
        _assert_same(nc[b], onc, f"box {b} no_cols")
    assert nc.sum() > 0


@pytest.mark.parametrize("bc", [(0, 1), (0, 0), (1, 1), (1, 0)])
def test_divergence_upwind(gpu, bc):
    from lacmopy_amd.integration import compute_divergence_upwind_np
    rs = np.random.default_rng(bc[0] * 2 + bc[1])
    nx, nz = 9, 7
    f = rs.normal(0.0, 1.0, (nx, nz))
    f[3, :] = f[4, :]  # flat stretches exercise the limiter's degenerate branch
    flux = rs.normal(0.0, 1.0, (2, nx + 1, nz + 1))
    got = compute_divergence_upwind_np(f, flux, (nx, nz), (20.0, 15.0), np.array(bc))
    ref = np.zeros((nx, nz))
    orc.lib().or_divergence_upwind(orc.P(f), orc.P(flux), nx, nz, 20.0, 15.0, bc[0], bc[1], orc.P(ref))
    _assert_same(got, ref, f"divergence bc={bc}")


def test_update_r_l_after_steps(gpu):
    from lacmopy_amd.engine import Engine  # noqa: F401
    from test_gpu_parity import _engine
    E, lnR0, lnf, nb = load_kernel_data_Ecol("Long_Bott")
    cfg = _config(E, lnR0, lnf, nb)
    grid, *parts = _case(nx=8, nz=8, spc=24, seed=29)
    parts = _prep(grid, parts)
    grid.rho_dry_inv = 1.0 / grid.mass_density_air_dry
    grid.mass_dry_inv = grid.rho_dry_inv / grid.volume_cell
    eng = _engine(grid, parts, cfg, E, lnR0, lnf, 3)
    eng.adv_steps(2)
    eng.update_r_l()
    s, _ = eng.export_grid(0)
    pos, vel, cells, m_w, m_s, xi, active = eng.export_particles(0)
    ref = np.zeros((8, 8))
    orc.lib().or_update_grid_r_l(len(xi), orc.P(m_w), orc.P(xi), orc.P(cells), orc.P(active.view(np.uint8)), 8, 8,
                                 orc.P(ref), orc.P(s[8].copy()))
    _assert_same(s[5], ref, "r_l")
    assert ref.sum() > 0


def test_derived_fields_and_ensemble_of_a_batch(gpu):
    """lm_derived_fields on a 3-run batch after two steps vs the oracle's evaluation.py:363-474 per run,
    and lm_ensemble_stats vs the oracle's ensemble statistics."""
    from lacmopy_amd.evaluation import derived_fields, ensemble_stats
    from test_gpu_parity import _engine
    E, lnR0, lnf, nb = load_kernel_data_Ecol("Long_Bott")
    cfg = _config(E, lnR0, lnf, nb)
    grid, *parts = _case(nx=8, nz=8, spc=24, seed=33, R_drop=30.0)
    parts = _prep(grid, parts)
    eng = _engine(grid, parts, cfg, E, lnR0, lnf, 4, n_runs=3)
    eng.adv_steps(2)
    d = derived_fields(eng)
    ref = []
    for r in range(3):
        pos, vel, cells, m_w, m_s, xi, act = eng.export_particles(r)
        s, _ = eng.export_grid(r)
        ref.append(orc.derived_fields(m_w, m_s, xi, cells, act, s[0].copy(), s[8].copy(), grid.volume_cell))
        _assert_same(d[r], ref[r], f"derived fields run {r}")
    assert (ref[0][2] > 0).any() and (ref[0][1] > 0).any()
    m, sd = ensemble_stats(d)
    om, osd = orc.ensemble_stats(np.array(ref))
    _assert_same(m, om, "ensemble mean")
    _assert_same(sd, osd, "ensemble spread")


def test_generate_field_frame_data_avg_from_files(gpu, tmp_path):
    """The file-based drop-in over two seeds and two frames vs the oracle applied to the same files."""
    import os
    from lacmopy_amd import file_handling as fh
    from lacmopy_amd.evaluation import generate_field_frame_data_avg
    from lacmopy_amd.setup_case import make_case
    paths = []
    for seed in (3, 4):
        path = str(tmp_path / f"seed_{seed}") + os.sep
        os.makedirs(path)
        grid, pos, vel, cells, m_w, m_s, xi, act = make_case(nx=6, nz=6, spc=20, seed=seed, cloud_base=30.0,
                                                             cloud_top=90.0, R_drop=25.0, sigma_drop=1.8)
        rs = np.random.default_rng(seed)
        for t in (0, 300):
            if t:   # a later frame of the same SIPs: grown masses, fewer in number, warmer, some removed
                m_w = m_w * rs.uniform(0.5, 3.0, m_w.size)
                xi = xi * rs.uniform(0.8, 1.0, xi.size)
                grid.temperature = grid.temperature + 1.0
                act = act.copy()
                act[::7] = False
                cells = cells.copy()
                cells[1, ~act] = -1
            fh.save_grid_scalar_fields(t, grid.scalar_fields(), path)
            fh.dump_particle_data_all(t, pos, vel, cells, m_w, m_s, xi, act, path)
        np.save(path + "grid_save_times.npy", np.array([0, 300]))
        paths.append(path)
    mdi = grid.mass_dry_inv
    got = generate_field_frame_data_avg(paths, np.array([0, 1, 3]), np.array([0, 1]), np.arange(9), mdi,
                                        grid.volume_cell, grid.no_cells, "AS")
    per_seed = np.zeros((2, 2, 12, 6, 6))
    for s_n, path in enumerate(paths):
        fields = fh.load_grid_scalar_fields(path, [0, 300])
        vec, cells_t, scal_t, xi_t, act_t = fh.load_particle_data_all(path, [0, 300])
        for t_n in range(2):
            per_seed[s_n, t_n, :3] = fields[t_n][[0, 1, 3]]
            per_seed[s_n, t_n, 3:] = orc.derived_fields(scal_t[t_n, 0], scal_t[t_n, 1], xi_t[t_n], cells_t[t_n],
                                                        act_t[t_n], fields[t_n, 3].copy(), mdi, grid.volume_cell)
    om, osd = orc.ensemble_stats(per_seed)
    _assert_same(got[0], om, "mean")
    _assert_same(got[1], osd, "spread")
    assert list(got[2]) == [0, 300] and len(got[3]) == 12


def test_step_dropins_on_device_tensors(gpu):
    """integrate_adv_step_np, integrate_subloop_n_steps_np, propagate_particles_subloop_step_np,
    propagate_grid_subloop_step_np and compute_divergence_upwind_np called with CUDA tensors of the
    reference's shapes and dtypes: the SIP arrays are reordered into the engine's layout and back on the
    device (no host copy of them) and updated where they are (same storage); every output equals the NumPy
    call's bit for bit (the NumPy path is checked against the oracle above)."""
    import torch
    from lacmopy_amd import integration as itg
    E, lnR0, lnf, nb = load_kernel_data_Ecol("Long_Bott")
    grid, *parts = _case(nx=8, nz=8, spc=24, seed=23)
    parts = _prep(grid, parts)
    N = len(parts[5])
    prof = init_profiles(grid)

    def state():
        pos, vel, cells, m_w, m_s, xi, active = [np.array(a, copy=True) for a in parts]
        rel = np.zeros((2, N))
        for d in (0, 1):
            rel[d][active] = (pos[d][active] - grid.ranges[d][0]) / grid.steps[d] - cells[d][active]
        return dict(scal=grid.scalar_fields(), mat=grid.mat_prop(), pos=pos, vel=vel, cells=cells, rel=rel, m_w=m_w,
                    m_s=m_s, xi=xi, active=active, T_p=np.ones(N), wr=np.zeros(1), nc=np.zeros(2, dtype=np.int64),
                    dml=np.zeros((8, 8)), dTh=np.full((8, 8), 1e-4), dRv=np.full((8, 8), -1e-7), ids=np.arange(N))

    def to_dev(h):
        return {k: torch.from_numpy(np.ascontiguousarray(v)).cuda() for k, v in h.items()}

    def calls(a):
        dt, dt_sub = 1.0, 0.1
        rng.seed(31)
        itg.integrate_adv_step_np(a["scal"], a["mat"], grid.velocity, grid.mass_flux_air_dry, grid.p_ref,
                                  grid.p_ref_inv, grid.no_cells, grid.ranges, grid.steps, grid.volume_cell, a["pos"],
                                  a["vel"], a["cells"], a["rel"], a["m_w"], a["m_s"], a["xi"], "AS", a["wr"], a["ids"],
                                  a["active"], a["T_p"], a["dml"], dt, dt_sub, 0.5 * dt_sub, 0.5 / grid.volume_cell, 5, 2,
                                  True, True, prof[0], prof[1], prof[2], 3, 9.80665, E, nb, lnR0, lnf, a["nc"])
        itg.integrate_subloop_n_steps_np(a["scal"], a["mat"], grid.velocity, grid.no_cells, grid.ranges, grid.steps,
                                         grid.volume_cell, grid.p_ref, grid.p_ref_inv, a["pos"], a["vel"], a["cells"],
                                         a["rel"], a["m_w"], a["m_s"], a["xi"], "AS", 0.5 / grid.volume_cell, E, nb,
                                         lnR0, lnf, a["nc"], a["wr"], a["ids"], a["active"], a["T_p"], a["dml"],
                                         a["dTh"], a["dRv"], dt_sub, dt_sub, 3, 4, 3, 9.80665, True)
        itg.propagate_particles_subloop_step_np(a["scal"], a["mat"], grid.velocity, grid.no_cells, grid.ranges,
                                                grid.steps, a["pos"], a["vel"], a["cells"], a["rel"], a["m_w"],
                                                a["m_s"], a["xi"], a["wr"], a["ids"], a["active"], a["T_p"], "AS",
                                                a["dml"], dt_sub, dt_sub, 3, 9.80665)
        itg.propagate_grid_subloop_step_np(a["scal"], a["mat"], grid.p_ref, grid.p_ref_inv, a["dTh"], a["dRv"],
                                           a["dml"], grid.volume_cell)
        return itg.compute_divergence_upwind_np(a["scal"][2], grid.mass_flux_air_dry, grid.no_cells, grid.steps)

    h = state()
    d = to_dev(state())
    ptrs = {k: t.data_ptr() for k, t in d.items()}
    div_h = calls(h)
    div_d = calls(d)
    assert isinstance(div_d, torch.Tensor) and div_d.is_cuda
    assert {k: t.data_ptr() for k, t in d.items()} == ptrs, "a tensor was replaced instead of updated"
    for k in h:
        _assert_same(d[k].cpu().numpy(), h[k], k)
    _assert_same(div_d.cpu().numpy(), div_h, "divergence")
    assert h["nc"].sum() > 0


def test_box_steps_on_device_tensors(gpu):
    """The Golovin, Ecol_grid_R and two-component cell steps on CUDA tensors: updated where they are (same
    storage), no_cols a device tensor added to in place, results equal the NumPy calls' bit for bit."""
    import torch
    from lacmopy_amd.collision import all_or_nothing as aon
    n = 180
    rs = np.random.default_rng(21)
    xi = rs.uniform(1e5, 1e8, n)
    m = rs.lognormal(np.log(1e7), 1.5, n)
    E, lnR0, lnf, nb = load_kernel_data_Ecol("Long_Bott")
    rho = np.full(n, 1000.0)
    R = (0.75 / np.pi * m / rho) ** (1 / 3)
    vel1 = np.array([orc.lib().or_terminal_velocity_Beard(r) for r in R])
    vel2 = np.array([rs.normal(0, 0.5, n), -vel1])
    m_s = rs.uniform(1e-3, 1.0, n)

    def run(dev):
        t = (lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()) if dev else (lambda a: np.array(a, copy=True))
        a = dict(x1=t(xi), m1=t(m), x2=t(xi), m2=t(m), R2=t(R), x3=t(xi), mw3=t(m), ms3=t(m_s), R3=t(R),
                 nc=t(np.zeros(2, dtype=np.int64)))
        rng.seed(6, 1)
        aon.collision_step_Golovin_np(a["x1"], a["m1"], 1.0, a["nc"])
        aon.collision_step_Ecol_grid_R_np(a["x2"], a["m2"], a["R2"], t(vel1), t(rho), 1.0, E, nb, lnR0, lnf, a["nc"])
        aon.collision_step_Ecol_grid_R_2D_multicomp_np(a["x3"], a["mw3"], a["ms3"], a["R3"], t(vel2), t(rho), 1.0,
                                                       E, nb, lnR0, lnf, a["nc"])
        return a

    h = run(False)
    d = run(True)
    for k in h:
        _assert_same(d[k].cpu().numpy(), h[k], k)
    assert h["nc"].sum() > 0
