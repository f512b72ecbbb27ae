"""Host side of the drop-in package without a GPU: state-file formats, the draw-key module, and that
every GPU entry point fails loudly instead of falling back to the CPU."""
import os

import numpy as np
import pytest

from lacmopy_amd import file_handling as fh
from lacmopy_amd import rng
from lacmopy_amd.setup_case import make_case


def test_state_files_round_trip(tmp_path):
    """file_handling.py:622-762: full save + load returns the same grid and particles."""
    grid, pos, vel, cells, m_w, m_s, xi, act = make_case(nx=6, nz=5, spc=8, seed=3)
    out = str(tmp_path) + os.sep
    fh.save_grid_and_particles_full(7, grid, pos, cells, vel, m_w, m_s, xi, act, out)
    names = sorted(os.listdir(out))
    assert names == sorted(["grid_basics_7.txt", "arr_file1_7.npy", "arr_file2_7.npy", "particle_vectors_7.npy",
                            "particle_cells_7.npy", "particle_scalars_7.npy", "multiplicity_7.npy",
                            "active_ids_7.npy"])
    g2, pos2, cells2, vel2, m_w2, m_s2, xi2, act2 = fh.load_grid_and_particles_full(7, out)
    for a, b in ((pos2, pos), (cells2, cells), (vel2, vel), (m_w2, m_w), (m_s2, m_s), (xi2, xi), (act2, act)):
        np.testing.assert_array_equal(a, b)
    for f in ("temperature", "pressure", "potential_temperature", "mass_density_air_dry",
              "mixing_ratio_water_vapor", "saturation", "saturation_pressure", "velocity", "mass_flux_air_dry"):
        np.testing.assert_array_equal(getattr(g2, f), getattr(grid, f))
    np.testing.assert_array_equal(g2.ranges, grid.ranges)
    np.testing.assert_array_equal(g2.steps, grid.steps)
    # r_l written by the save: 1e-18 m_dry^-1 sum(xi m_w) per cell
    tot = (np.load(out + "arr_file1_7.npy")[4] / (1e-18 * grid.mass_dry_inv)).sum()
    np.testing.assert_allclose(tot, (xi * m_w)[act].sum(), rtol=1e-12)
    # arr_file1 order (p, T, rho_d, r_v, r_l, e_s, S, Theta)
    a1 = np.load(out + "arr_file1_7.npy")
    np.testing.assert_array_equal(a1[0], grid.pressure)
    np.testing.assert_array_equal(a1[7], grid.potential_temperature)


def test_frame_and_tracer_files(tmp_path):
    out = str(tmp_path) + os.sep
    grid, pos, vel, cells, m_w, m_s, xi, act = make_case(nx=4, nz=4, spc=4, seed=2)
    scal = grid.scalar_fields()
    fh.save_grid_scalar_fields(30, scal, out)
    g = fh.load_grid_scalar_fields(out, [30])
    np.testing.assert_array_equal(g[0], np.array((scal[4], scal[5], scal[2], scal[0], scal[1], scal[6])))
    fh.dump_particle_data_all(30, pos, vel, cells, m_w, m_s, xi, act, out)
    v, c, s, x, a = fh.load_particle_data_all(out, [30])
    np.testing.assert_array_equal(v[0, 0], pos)
    np.testing.assert_array_equal(c[0], cells)
    np.testing.assert_array_equal(s[0, 1], m_s)
    tv = np.zeros((3, 2, 2, 5))
    tv[1, 0, 1, 2] = 4.0
    fh.dump_particle_tracer_data_block(np.array([0, 10, 20]), tv, np.zeros((3, 3, 5)), np.zeros((3, 5)),
                                       np.zeros(3), out)
    fh.dump_particle_data(30, pos[:, :5], vel[:, :5], m_w[:5], m_s[:5], xi[:5], scal[0], scal[4], out)
    vec, sc, xs = fh.load_particle_data_from_blocks(out, [0], 3)
    assert vec.shape == (3, 2, 2, 5) and vec[1, 0, 1, 2] == 4.0 and sc.shape == (3, 3, 5)
    vec, sc, xs = fh.load_particle_data_from_blocks(out, [30], 3)
    np.testing.assert_array_equal(vec[0, 0], pos[:, :5])
    np.testing.assert_array_equal(sc[0, 1], m_s[:5])
    fh.save_sim_paras_to_file([1.0, 0.1, 3, 17], "dt dt_sub no_iter_impl_mass rnd_seed_sim", 0, out)
    with open(out + "sim_paras_t_0.txt") as f:
        assert f.read() == "dt dt_sub no_iter_impl_mass rnd_seed_sim\n1.0 0.1 3 17 "


def test_kernel_data_loader():
    from lacmopy_amd import DATA
    d = os.path.join(DATA, "Ecol_grid_data", "Long_Bott")
    E, r, R0, f, lnR0, lnf, nb = fh.load_kernel_data_Ecol("Ecol_grid_R", d)
    assert E.shape == (nb, nb) and nb == len(r)
    assert lnR0 == np.log(r[0]) and lnf == np.log(r[1] / r[0])
    assert fh.load_kernel_data_Ecol("Ecol_const", d, 0.7)[0] == 0.7


def test_rng_counter():
    rng.seed(5)
    assert rng.get_seed() == 5 and rng.get_step() == 0
    assert rng.take_steps(3) == 0 and rng.get_step() == 3
    rng.seed(2 ** 64 + 9, 7)
    assert rng.get_seed() == 9 and rng.get_step() == 7


def test_gpu_entry_points_fail_loudly_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from lacmopy_amd.collision.all_or_nothing import collision_step_Golovin_np
    from lacmopy_amd.integration import propagate_grid_subloop_step_np
    with pytest.raises(RuntimeError):
        collision_step_Golovin_np(np.ones(4), np.ones(4), 1.0, np.zeros(2, dtype=np.int64))
    with pytest.raises(RuntimeError):
        propagate_grid_subloop_step_np(np.ones((10, 3, 3)), np.ones((7, 3, 3)), 1e5, 1e-5, np.zeros((3, 3)),
                                       np.zeros((3, 3)), np.zeros((3, 3)), 1.0)


def test_dual_newton_jacobian(tmp_path):
    """The adjoint's local Jacobian (lm_physics.h on lm_dual.h, compiled for the host with the kernels'
    -ffp-contract=off): the dual value part equals the double Newton bit for bit, and its derivatives in
    (m_w, m_s, T, p, S) match central finite differences of the double Newton."""
    import ctypes as C
    import shutil
    import subprocess

    from lacmopy_amd._native import lib
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    here = os.path.dirname(os.path.abspath(__file__))
    csrc = os.path.join(here, "..", "lacmopy_amd", "csrc")
    exe = str(tmp_path / "dual_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-I", csrc, os.path.join(here, "dual_check.cpp"),
                    "-o", exe], check=True)
    L = lib()
    n_const = L.lm_num_constants()
    consts = (C.c_double * n_const)()
    L.lm_get_constants(consts, n_const)
    rng = np.random.default_rng(3)
    n = 400
    m_s = 10.0 ** rng.uniform(-3, 2, n)
    m_w = m_s * 10.0 ** rng.uniform(0.3, 4, n)
    T = rng.uniform(272.0, 292.0, n)
    p = rng.uniform(8.5E4, 1.0E5, n)
    S = rng.uniform(0.95, 1.005, n)
    lines = [" ".join(repr(float(c)) for c in consts), str(n)]
    lines += [" ".join(repr(float(v)) for v in row) for row in zip(m_w, m_s, T, p, S)]
    out = subprocess.run([exe], input="\n".join(lines) + "\n", capture_output=True, text=True, check=True).stdout
    rows = [ln.split() for ln in out.strip().splitlines()]
    assert len(rows) == n
    same = np.array([int(r[0]) for r in rows])
    err = np.array([float(r[1]) for r in rows])
    assert same.all(), f"{(same == 0).sum()} dual values differ from the double Newton"
    assert np.median(err) < 1.0E-8 and err.max() < 1.0E-4, (np.median(err), err.max())


@pytest.mark.parametrize("scale_dt_cond", [1, 2, 3])
@pytest.mark.parametrize("no_col_per_adv", [1, 2, 5])
def test_collision_steps_per_adv_step_match_the_subloop_schedule(scale_dt_cond, no_col_per_adv):
    """The collision-step counter simulate() resumes from (integration.col_steps_per_adv, also the adjoint's
    count) equals the number of collision steps the oracle's restatement of integrate_adv_step_np
    (integration.py:1515-1585 with the subloop placement of :1247-1292) takes per advection step."""
    from lacmopy_amd.integration import col_steps_per_adv
    from oracle import oracle as orc
    from lacmopy_amd.setup_case import init_profiles
    grid, pos, vel, cells, m_w, m_s, xi, act = make_case(nx=3, nz=3, spc=2, seed=1)
    cfg = dict(dt_adv=1.0, dt_col=1.0 / no_col_per_adv, scale_dt_cond=scale_dt_cond, no_col_per_adv=no_col_per_adv,
               act_collisions=True, act_relaxation=False, no_iter_impl_mass=3, solute_type="AS", g_set=9.80665)
    N = len(xi)
    p = orc.adv_params(grid, N, cfg, seed=1, col_step=7, E_col_grid=0.5)
    rel = np.zeros_like(pos)
    scal, mat = grid.scalar_fields(), grid.mat_prop()
    prof = init_profiles(grid)
    n_steps = 2
    orc.integrate_adv_steps(p, n_steps, pos, vel, cells, rel, m_w, m_s, xi, act, np.ones(N), np.zeros(1),
                            np.zeros(2, dtype=np.int64), scal, mat, grid.velocity, grid.mass_flux_air_dry, 0.5,
                            *prof)
    assert p.col_step - 7 == n_steps * col_steps_per_adv(no_col_per_adv, scale_dt_cond)
    assert col_steps_per_adv(no_col_per_adv, scale_dt_cond, act_collisions=False) == 0
