"""CPU: the tangent-linear reference of the adjoint window (oracle/tlm_window.cpp).

Its value part must be the oracle's own trajectory bit for bit: the objectives it returns equal
those of the oracle's window (or_integrate_adv_steps without collisions) evaluated from the final
state with the same sums.  Its derivative part is checked against central finite differences of
that oracle window along the same directions (relative multiplicity and dry-mass perturbations of
all droplets and of single SIPs).  The GPU adjoint is compared with it in test_gpu_adjoint.py.
"""
import math

import numpy as np

from lacmopy_amd.setup_case import init_profiles, make_case
from oracle import oracle as orc

CFG = dict(solute_type="AS", dt_adv=1.0, dt_col=0.5, scale_dt_cond=5, no_col_per_adv=2, no_iter_impl_mass=3,
           g_set=9.80665, act_collisions=False, act_relaxation=True)
WIDTH = 0.1
LN_HALF, LN_25 = -float.fromhex("0x1.62e42fefa39efp-1"), 3.2188758248682006


def _case():
    grid, pos, vel, cells, m_w, m_s, xi, act = make_case(nx=6, nz=10, spc=16, seed=3, cloud_base=60.0,
                                                         cloud_top=160.0, R_drop=14.0, sigma_drop=1.4)
    N = len(xi)
    p = orc.adv_params(grid, N, CFG, E_col_grid=0.5)
    rel = np.zeros((2, N))
    orc.lib().or_update_pos_and_cells(orc.C.byref(p), 0.0, orc.P(pos), orc.P(vel), orc.P(cells), orc.P(rel),
                                      orc.P(m_w), orc.P(xi), orc.P(act.view(np.uint8)), orc.P(np.zeros(1)))
    return grid, (pos, vel, cells, m_w, m_s, xi, act)


def _objectives(grid, parts, scal):
    """J_nc, J_lwp of a state: per cell in ascending id, then over the cells, as tlm_window.cpp sums them."""
    pos, vel, cells, m_w, m_s, xi, act = parts
    nx, nz = grid.no_cells
    nc = nx * nz
    T = scal[0].reshape(-1)
    ids = np.flatnonzero(act)
    c = cells[0, ids] * nz + cells[1, ids]
    R, ws, rho = (np.zeros(len(ids)) for _ in range(3))
    orc.lib().or_R_p_w_s_rho_p(len(ids), orc.P(np.ascontiguousarray(m_w[ids])), orc.P(np.ascontiguousarray(m_s[ids])),
                               orc.P(np.ascontiguousarray(T[c])), 0, orc.P(R), orc.P(ws), orc.P(rho))
    lr = orc.lm_vec(1, R)
    a, b = (lr - LN_HALF) / WIDTH, (LN_25 - lr) / WIDTH
    I = 1.0 / (1.0 + orc.lm_vec(0, -a)) * (1.0 / (1.0 + orc.lm_vec(0, -b)))
    A, B = np.zeros(nc), np.zeros(nc)
    for k, cc in enumerate(c):
        A[cc] += xi[ids[k]] * m_w[ids[k]]
        B[cc] += xi[ids[k]] * I[k]
    mdinv = scal[8].reshape(-1)
    J_lwp = J_nc = 0.0
    for cc in range(nc):
        J_lwp += A[cc]
        J_nc += mdinv[cc] * B[cc]
    c_lwp = 1.0E-18 / ((grid.ranges[0, 1] - grid.ranges[0, 0]) * (grid.volume_cell / (grid.steps[0] * grid.steps[1])))
    return 1.0E-6 * J_nc / nc, c_lwp * J_lwp


def _oracle_J(grid, parts, n_steps, xi=None, m_s=None):
    pos, vel, cells, m_w, m_s0, xi0, act = [np.array(a, copy=True) for a in parts]
    xi = xi0 if xi is None else np.array(xi, copy=True)
    m_s = m_s0 if m_s is None else np.array(m_s, copy=True)
    N = len(xi)
    scal, mat = grid.scalar_fields(), grid.mat_prop()
    prof_rv, prof_Th, t_relax = init_profiles(grid)
    p = orc.adv_params(grid, N, CFG, E_col_grid=0.5)
    orc.integrate_adv_steps(p, n_steps, pos, vel, cells, np.zeros((2, N)), m_w, m_s, xi, act, np.ones(N), np.zeros(1),
                            np.zeros(2, dtype=np.int64), scal, mat, grid.velocity, grid.mass_flux_air_dry, 0.5,
                            prof_rv, prof_Th, t_relax)
    return np.array(_objectives(grid, (pos, vel, cells, m_w, m_s, xi, act), scal))


def _tlm(grid, parts, n_steps, dir_xi, dir_ms):
    pos, vel, cells, m_w, m_s, xi, act = parts
    p = orc.adv_params(grid, len(xi), CFG, E_col_grid=0.5)
    prof_rv, prof_Th, t_relax = init_profiles(grid)
    return orc.tlm_window(p, n_steps, WIDTH, pos, vel, cells, m_w, m_s, xi, act, grid.scalar_fields(),
                          grid.mat_prop(), grid.velocity, grid.mass_flux_air_dry, prof_rv, prof_Th, t_relax,
                          dir_xi, dir_ms)


def test_tlm_values_are_the_oracle_window():
    grid, parts = _case()
    N = len(parts[5])
    J, _ = _tlm(grid, parts, 3, np.zeros((1, N)), np.zeros((1, N)))
    J_or = _oracle_J(grid, parts, 3)
    assert np.array_equal(J, J_or), (J, J_or)
    assert J[0] > 0 and J[1] > 0


def test_tlm_derivatives_against_finite_differences():
    grid, parts = _case()
    xi, m_s, m_w = parts[5], parts[4], parts[3]
    N = len(xi)
    rng = np.random.default_rng(11)
    drops = np.flatnonzero(m_w > 4.19E3)   # R > 1 um
    k1, k2 = rng.choice(drops, 2, replace=False)
    r = rng.uniform(-1.0, 1.0, N)
    dirs_xi = np.zeros((4, N))
    dirs_ms = np.zeros((4, N))
    dirs_xi[0] = r * xi                       # relative perturbation of every multiplicity
    dirs_ms[1] = r[::-1] * m_s                # of every dry mass
    dirs_xi[2, k1] = xi[k1]                   # one droplet's multiplicity
    dirs_ms[3, k2] = m_s[k2]                  # one droplet's dry mass
    n_steps = 3
    J, dJ = _tlm(grid, parts, n_steps, dirs_xi, dirs_ms)
    eps = np.finfo(np.float64).eps
    for d in range(4):
        for h in (1e-4, 1e-5):
            Jp = _oracle_J(grid, parts, n_steps, xi + h * dirs_xi[d], m_s + h * dirs_ms[d])
            Jm = _oracle_J(grid, parts, n_steps, xi - h * dirs_xi[d], m_s - h * dirs_ms[d])
            fd = (Jp - Jm) / (2 * h)
            for q in range(2):
                floor = 256 * eps * abs(J[q]) / h
                assert abs(fd[q] - dJ[q, d]) <= 1e-5 * abs(dJ[q, d]) + floor, (d, h, q, fd[q], dJ[q, d], floor)
    assert np.all(np.abs(dJ[:, :2]) > 0) and math.isfinite(float(np.sum(dJ)))
