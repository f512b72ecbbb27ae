"""More pins of the CPU oracle: the NaCl mass rate, the drag law's Reynolds-number branches, the relaxation
term and its use in the advection step, and Beard's terminal velocity.

Each check comes from the mathematics (finite differences, closed forms, continuity of piecewise fits), a
textbook law written here in its own form, or published measurements -- not from the oracle's expressions.
"""
import ctypes as C
import math
import os

import numpy as np
import pytest

from lacmopy_amd.integration import compute_relaxation_time_profile
from lacmopy_amd.setup_case import init_profiles, make_case
from oracle import oracle as orc

L = orc.lib()
R_V = 461.53


def _cell_state(T=283.0, p=9.0e4, S=1.002):
    Kc = 4.1868E-3 * (5.69 + 0.017 * (T - 273.15))
    D_v = 4.01218E-5 * T ** 1.94 / p
    L_v = 1.0E3 * (2500.82 - 2.358 * (T - 273.0))
    e_s = 2.53E11 * math.exp(-5420.0 / T)
    return T, p, S, e_s, L_v, Kc, D_v


def _rate(solute, m_w, m_s, T, p, S, e_s, L_v, Kc, D_v):
    """Mass rate and derivative with the radius, mass fraction, density and surface tension of the solute's
    own laws (microphysics.py:179-266, material_properties.py:272-366)."""
    Ra, wsa, rhoa = np.zeros(1), np.zeros(1), np.zeros(1)
    L.or_R_p_w_s_rho_p(1, orc.P(np.array([m_w])), orc.P(np.array([m_s])), orc.P(np.array([T])), solute,
                       orc.P(Ra), orc.P(wsa), orc.P(rhoa))
    sigma = L.or_surface_tension(solute, wsa[0], T)
    g, dg = C.c_double(), C.c_double()
    L.or_mass_rate_and_derivative(solute, m_w, m_s, wsa[0], Ra[0], T, rhoa[0], T, p, S, e_s, L_v, Kc, D_v, sigma,
                                  C.byref(g), C.byref(dg))
    return g.value, dg.value, Ra[0], wsa[0]


# ---------------------------------------------------------------- NaCl mass rate (microphysics.py:1237-1353)
@pytest.mark.parametrize("S", [0.98, 1.0005, 1.01])
def test_nacl_mass_rate_derivative_matches_finite_difference(S):
    """d gamma / d m_w of the NaCl rate against a central difference of gamma, the radius, mass fraction,
    density and surface tension all following m_w through the NaCl laws."""
    T, p, _, e_s, L_v, Kc, D_v = _cell_state()
    for m_w, m_s in [(2.0, 1.0), (40.0, 1.0), (3e3, 0.5), (1e6, 2.0), (0.2, 0.01)]:
        g, dg, _, _ = _rate(1, m_w, m_s, T, p, S, e_s, L_v, Kc, D_v)
        h = 1e-6 * m_w
        gp = _rate(1, m_w + h, m_s, T, p, S, e_s, L_v, Kc, D_v)[0]
        gm = _rate(1, m_w - h, m_s, T, p, S, e_s, L_v, Kc, D_v)[0]
        fd = (gp - gm) / (2 * h)
        assert abs(dg - fd) <= 1e-5 * max(abs(dg), abs(fd)) + 1e-14, (m_w, m_s, dg, fd)


def test_nacl_mass_rate_vanishes_at_the_koehler_equilibrium():
    """gamma = 0 when the ambient saturation is the drop's equilibrium saturation a_w(w_s) exp(Kelvin), taken
    from the initialisation's S_eq (pinned on its own against the radius form of Koehler's law); the only
    difference is the reference's rho**0.66666667 there, a relative 1e-9 in the Kelvin exponent."""
    T, p, _, e_s, L_v, Kc, D_v = _cell_state()
    for m_w, m_s in [(5.0, 1.0), (80.0, 1.0), (4e3, 0.3)]:
        w_s = m_s / (m_w + m_s)
        S_eq = L.or_S_eq_mass_fraction(1, w_s, T, m_s)
        g_eq = _rate(1, m_w, m_s, T, p, S_eq, e_s, L_v, Kc, D_v)[0]
        g_ref = _rate(1, m_w, m_s, T, p, S_eq + 0.01, e_s, L_v, Kc, D_v)[0]
        assert abs(g_eq) < 1e-6 * abs(g_ref), (m_w, g_eq, g_ref)
        # below / above equilibrium the drop evaporates / grows
        assert _rate(1, m_w, m_s, T, p, S_eq - 1e-4, e_s, L_v, Kc, D_v)[0] < 0 < g_ref


def test_nacl_mass_rate_large_drop_maxwell_limit():
    """R >> Knudsen lengths and a dilute drop: the Maxwell-Mason growth law (as for the AS rate)."""
    T, p, S, e_s, L_v, Kc, D_v = _cell_state(S=1.01)
    m_w, m_s = 4.0 / 3.0 * math.pi * 1000.0 * 2000.0 ** 3, 1.0
    g, _, R, _ = _rate(1, m_w, m_s, T, p, S, e_s, L_v, Kc, D_v)
    S_eq = 1.0 + 2 * 0.0745 / (R_V * T * 1000.0 * R * 1e-6)
    maxwell = 4 * math.pi * R * 1e-6 * (S - S_eq) / (S_eq * L_v ** 2 / (R_V * Kc * T * T) + R_V * T / (D_v * e_s))
    assert abs(g * 1e-18 / maxwell - 1) < 3e-3


def test_nacl_water_activity_dilute_limit_is_raoult():
    """A dilute NaCl drop: 1 - a_w -> 2 (M_w / M_NaCl) m_s / m_w (Raoult's law, two ions, osmotic coefficient
    -> 1), read from the equilibrium saturation of a drop so large that the Kelvin term is negligible."""
    T = 290.0
    m_s = 1.0
    for w in (1e-3, 3e-3):
        m_w = m_s / w - m_s
        S_eq = L.or_S_eq_mass_fraction(1, w, T, m_s)
        R = (0.75 / math.pi * (m_w + m_s) / 1000.0) ** (1 / 3)
        kelvin = math.exp(2 * 0.073 / (R_V * T * 1000.0 * R * 1e-6))
        raoult = 2 * 18.015 / 58.4428 * m_s / m_w
        assert abs((1 - S_eq / kelvin) / raoult - 1) < 0.05


# ---------------------------------------------------------------- drag (integration.py:699-718)
def _drag_substep(R_um, u0, dt=0.1):
    """One particle substep of the oracle in still air without gravity for SIPs of radius R_um (um) moving
    horizontally at u0 (m/s); returns u after the step and what the textbook drag needs."""
    grid, pos, vel, cells, m_w, m_s, xi, act = make_case(nx=3, nz=3, spc=4, seed=2)
    N = xi.size
    k = np.arange(N) % len(R_um)
    R0 = np.asarray(R_um, dtype=np.float64)[k]
    m_w[:] = 4.0 / 3.0 * math.pi * 1000.0 * R0 ** 3
    m_s[:] = 1e-3
    grid.velocity[:] = 0.0
    scal, mat = grid.scalar_fields(), grid.mat_prop()
    if callable(u0):  # speeds from the viscosity and fluid density of each SIP's cell
        vel[0] = u0(mat[5][cells[0], cells[1]], mat[6][cells[0], cells[1]], R0, np.arange(N))
    else:
        vel[0] = np.asarray(u0, dtype=np.float64)[k]
    vel[1] = 0.0
    cfg = dict(dt_adv=1.0, scale_dt_cond=5, dt_col=0.5, no_col_per_adv=2, act_collisions=False,
               act_relaxation=False, no_iter_impl_mass=3, solute_type="AS", g_set=0.0)
    p = orc.adv_params(grid, N, cfg)
    rel = np.zeros((2, N))
    wr = np.zeros(1)
    L.or_update_pos_and_cells(C.byref(p), 0.0, orc.P(pos), orc.P(vel), orc.P(cells), orc.P(rel), orc.P(m_w),
                              orc.P(xi), orc.P(act.view(np.uint8)), orc.P(wr))
    u_before = vel[0].copy()
    ci, cj = cells[0].copy(), cells[1].copy()
    T_p = np.ones(N)
    L.or_propagate_particles_substep(C.byref(p), dt, 0.0, orc.P(pos), orc.P(vel), orc.P(cells), orc.P(rel),
                                     orc.P(m_w), orc.P(m_s), orc.P(xi), orc.P(act.view(np.uint8)), orc.P(T_p),
                                     orc.P(wr), orc.P(scal), orc.P(mat), orc.P(grid.velocity), orc.P(np.zeros(9)))
    T = scal[0][ci, cj]
    R, ws, rho = np.zeros(N), np.zeros(N), np.zeros(N)
    # the drag uses the AS radius and density after the mass update (integration.py:1075)
    L.or_R_p_w_s_rho_p(N, orc.P(m_w), orc.P(m_s), orc.P(np.ascontiguousarray(T)), 0, orc.P(R), orc.P(ws),
                       orc.P(rho))
    return u_before, vel, R, rho, mat[5][ci, cj], mat[6][ci, cj], dt


def _textbook_drag_factor(R_um, rho_p, v_rel, mu, rho_f, dt):
    """Implicit-Euler factor dt/tau_p * f(Re) of the drag on a sphere: tau_p = 2 rho_p r^2 / (9 mu) (Stokes),
    Re = rho_f |v| 2r / mu; f = 1 (Stokes), Schiller-Naumann 1 + 0.15 Re^0.687, and Newton's regime
    C_D = 0.44, i.e. f = C_D Re / 24."""
    r = R_um * 1e-6
    Re = rho_f * np.abs(v_rel) * 2.0 * r / mu
    f = np.where(Re <= 0.5, 1.0, np.where(Re < 1000.0, 1.0 + 0.15 * Re ** 0.687, 0.44 * Re / 24.0))
    tau = 2.0 * rho_p * r * r / (9.0 * mu)
    return dt * f / tau, Re


def test_drag_schiller_naumann_and_newton_branches():
    """The implicit velocity update v' = (v + k v_f)/(1 + k) in still air (v_f = 0, no gravity) gives
    k = v/v' - 1, which must equal the textbook factor across the Stokes, Schiller-Naumann (0.5 < Re < 1000)
    and Newton (Re >= 1000) regimes."""
    R_um = [20.0, 60.0, 150.0, 400.0, 800.0, 1500.0]
    u0 = [0.05, 1.5, 4.0, 12.0, 25.0, 40.0]
    u_b, vel, R, rho, mu, rho_f, dt = _drag_substep(R_um, u0)
    k_oracle = u_b / vel[0] - 1.0
    k_text, Re = _textbook_drag_factor(R, rho, u_b, mu, rho_f, dt)
    assert (Re <= 0.5).any() and ((Re > 0.5) & (Re < 1000)).any() and (Re >= 1000).any(), Re
    np.testing.assert_allclose(k_oracle, k_text, rtol=1e-10)
    np.testing.assert_array_equal(vel[1], 0.0)


def test_drag_regimes_meet_at_re_1000():
    """Schiller-Naumann and Newton's regime nearly coincide at Re = 1000 (18.26 against 18.33): the factor
    jumps by under 0.5% across the branch point."""
    def speeds(mu, rho_f, R0, idx):  # Re = 1000 (1 -+ 1e-3) at the drop's radius
        return 1000.0 * mu / (rho_f * 2.0 * R0 * 1e-6) * np.where(idx % 2 == 0, 1.0 - 1e-3, 1.0 + 1e-3)

    u_b, vel, R, rho, mu_c, rhof_c, dt = _drag_substep([1000.0], speeds)
    k = u_b / vel[0] - 1.0
    _, Re = _textbook_drag_factor(R, rho, u_b, mu_c, rhof_c, dt)
    lo, hi = (Re < 1000.0), (Re >= 1000.0)
    assert lo.any() and hi.any()
    # k / Re removes the linear dependence on the speed common to both sides
    jump = np.mean((k / Re)[hi]) / np.mean((k / Re)[lo]) - 1.0
    assert 0.0 < jump < 5e-3, jump


# ---------------------------------------------------------------- relaxation (integration.py:291-354)
def test_relaxation_term_closed_form():
    """dt (profile0 - mean over x of the field) / tau(z), tau = 300 exp(z / 200): for a field whose x-mean
    per level is g_j the term is dt (p_j - g_j)/tau_j, whatever the x variation."""
    nx, nz, dt = 7, 5, 0.1
    z = 10.0 + 20.0 * np.arange(nz)
    tau = compute_relaxation_time_profile(z)
    np.testing.assert_allclose(tau, [300.0 * math.exp(zz / 200.0) for zz in z], rtol=1e-15)
    g = 288.0 + 0.3 * np.arange(nz)
    prof = g + np.linspace(-0.4, 0.6, nz)
    wig = np.outer(np.arange(nx) - (nx - 1) / 2.0, np.ones(nz)) * 0.25  # zero mean in x
    field = np.ascontiguousarray(g[None, :] + wig)
    out = np.zeros(nz)
    L.or_relaxation_term(orc.P(field), orc.P(prof), orc.P(np.ascontiguousarray(tau)), dt, nx, nz, orc.P(out))
    np.testing.assert_allclose(out, dt * (prof - g) / tau, rtol=1e-12, atol=1e-16)


def test_relaxation_in_the_advection_step_is_second_order():
    """With no flux and SIPs too thin to condense noticeably, one advection step changes Theta and r_v only by
    relaxation.  The step adds dt_sub * r(f_0) over the first subloop's scale_dt_cond substeps, then
    2 dt_sub r(f_mid) - that first increment over the second subloop's (integration.py:1483-1609), which for
    d f / dt = (p - f) / tau is the midpoint rule: f_1 - f_0 = dt/tau (p - f_0) (1 - dt / (2 tau))."""
    grid, pos, vel, cells, m_w, m_s, xi, act = make_case(nx=4, nz=6, spc=3, seed=7)
    N = xi.size
    xi[:] = 1e-30
    grid.mass_flux_air_dry[:] = 0.0
    grid.velocity[:] = 0.0
    vel[:] = 0.0
    prof_rv, prof_Th, _ = init_profiles(grid)
    prof_Th = prof_Th + 0.5
    prof_rv = prof_rv * 1.02
    tau = compute_relaxation_time_profile(grid.centers[1][0]) * 1e-2  # a fast relaxation, visible in one step
    cfg = dict(dt_adv=1.0, scale_dt_cond=5, dt_col=0.5, no_col_per_adv=2, act_collisions=False,
               act_relaxation=True, no_iter_impl_mass=3, solute_type="AS", g_set=0.0)
    p = orc.adv_params(grid, N, cfg)
    scal, mat = grid.scalar_fields(), grid.mat_prop()
    Th0, rv0 = scal[2].copy(), scal[4].copy()
    orc.integrate_adv_steps(p, 1, pos, vel, cells, np.zeros((2, N)), m_w, m_s, xi, act, np.ones(N), np.zeros(1),
                            np.zeros(2, dtype=np.int64), scal, mat, grid.velocity, grid.mass_flux_air_dry, 0.5,
                            np.ascontiguousarray(prof_rv), np.ascontiguousarray(prof_Th), np.ascontiguousarray(tau))
    dt = 1.0
    for f0, f1, prof in ((Th0, scal[2], prof_Th), (rv0, scal[4], prof_rv)):
        mean0 = f0.mean(axis=0)
        expect = dt / tau * (prof - mean0) * (1.0 - dt / (2.0 * tau))
        np.testing.assert_allclose(f1 - f0, np.broadcast_to(expect, f0.shape), rtol=1e-9, atol=1e-14)


# ---------------------------------------------------------------- Beard's velocity (kernel.py:107-167)
def test_beard_stokes_cunningham_regime():
    """R < 10 um: Stokes' law with the Cunningham slip correction, v = 2/9 (rho_w - rho_a) g r^2 / mu
    (1 + 1.257 l0 / r), l0 = 0.0662 um, with Bott's constants (rho_a 1.225, mu 1.818e-5)."""
    for R in (0.5, 2.0, 6.0, 9.9):
        r = R * 1e-6
        v = 2.0 / 9.0 * (1000.0 - 1.225) * 9.80665 * r * r / 1.818E-5 * (1.0 + 1.257 * 0.0662 / R)
        assert math.isclose(L.or_terminal_velocity_Beard(R), v, rel_tol=1e-12)


def test_beard_regimes_join_continuously():
    """The three fits of Beard (1976) meet at 10 um, 535 um and the 3.5 mm cap within their accuracy."""
    for R, tol in ((10.0, 5e-3), (535.0, 2e-3), (3500.0, 1e-6)):
        below = L.or_terminal_velocity_Beard(R * (1 - 1e-12))
        at = L.or_terminal_velocity_Beard(R)
        assert abs(below / at - 1) < tol, (R, below, at)


def test_beard_against_gunn_kinzer_measurements():
    """Terminal velocities of water drops in air at 20 C and 1013 hPa measured by Gunn & Kinzer (1949,
    J. Meteor. 6: 243), diameter (mm) -> m/s; Beard's fit reproduces them within a few percent (the air
    density here is Bott's 1.225 kg m^-3, slightly above that of the measurements)."""
    gunn_kinzer = {0.2: 0.72, 0.4: 1.62, 0.6: 2.47, 0.8: 3.27, 1.0: 4.03, 1.4: 5.17, 2.0: 6.49, 3.0: 8.06,
                   4.0: 8.83, 5.0: 9.09}
    for d, v in gunn_kinzer.items():
        got = L.or_terminal_velocity_Beard(d * 500.0)
        assert abs(got / v - 1) < 0.05, (d, got, v)
    # monotone in R up to the cap
    R = np.geomspace(0.5, 3400.0, 400)
    v = np.array([L.or_terminal_velocity_Beard(r) for r in R])
    assert np.all(np.diff(v) > -1e-3 * v[1:])


# ---------------------------------------------------------------- material laws (material_properties.py)
def _rho(solute, w, T):
    m_s = 1.0 if w > 0 else 0.0
    m_w = m_s / w - m_s if w > 0 else 1.0
    R, ws, rho = np.zeros(1), np.zeros(1), np.zeros(1)
    L.or_R_p_w_s_rho_p(1, orc.P(np.array([m_w])), orc.P(np.array([m_s])), orc.P(np.array([T])), solute,
                       orc.P(R), orc.P(ws), orc.P(rho))
    return rho[0]


def test_water_and_solution_densities_against_tables():
    """Liquid water (material_properties.py:29-50) against the CRC Handbook's table (kg m^-3 at 0-40 C), and
    the solution densities (:56-115) at 20 C against CRC's concentrative-property tables of NaCl and
    (NH4)2SO4 solutions, mass fractions 0.05-0.30."""
    water = {273.15: 999.84, 283.15: 999.70, 293.15: 998.21, 303.15: 995.65, 313.15: 992.22}
    for T, rho in water.items():
        assert abs(_rho(0, 0.0, T) / rho - 1) < 5e-4, T
    nacl = {0.05: 1034.0, 0.10: 1070.7, 0.20: 1147.8, 0.26: 1197.2}
    amsulf = {0.04: 1022.8, 0.10: 1057.5, 0.20: 1114.9, 0.30: 1172.4}
    for w, rho in nacl.items():
        assert abs(_rho(1, w, 293.15) / rho - 1) < 3e-3, ("NaCl", w)
    for w, rho in amsulf.items():
        assert abs(_rho(0, w, 293.15) / rho - 1) < 3e-3, ("AS", w)


def test_surface_tensions_against_measurements():
    """Water (material_properties.py:224-244) against IAPWS values (75.65, 72.74, 69.60 mN m^-1 at 0, 20,
    40 C); solutions (:272-366): the increase over water is linear in molality b = w / ((1 - w) M_solute),
    1.6 mN m^-1 per mol kg^-1 for NaCl and 2.4 for (NH4)2SO4 (Pruppacher & Klett, ch. 5)."""
    for T, s in ((273.15, 75.65e-3), (293.15, 72.74e-3), (313.15, 69.60e-3)):
        assert abs(L.or_surface_tension(0, 0.0, T) / s - 1) < 5e-3, T
    for solute, M, slope in ((1, 58.4428e-3, 1.62e-3), (0, 132.14e-3, 2.36e-3)):
        sw = L.or_surface_tension(solute, 0.0, 293.15)
        for w in (0.05, 0.2, 0.35):
            b = w / ((1.0 - w) * M)
            assert abs((L.or_surface_tension(solute, w, 293.15) - sw) / b / slope - 1) < 0.05, (solute, w)


# ---------------------------------------------------------------- collection-efficiency tables (kernel.py:485-742)
@pytest.mark.parametrize("kernel,R_high", [("Long_Bott", 51.0), ("Hall_Bott", 301.0)])
def test_regenerated_efficiency_tables_equal_the_reference_tables(kernel, R_high):
    """The table generator (Long's closed form; Hall's raw table interpolated in collector radius and radius
    ratio) on 100 bins per decade from 0.01 um reproduces the reference's saved Ecol_grid_data tables, which
    every hydrodynamic collision step reads, bit for bit."""
    from lacmopy_amd.collision.kernel import generate_E_col_grid_R_np
    from lacmopy_amd.file_handling import load_kernel_data_Ecol
    import lacmopy_amd
    E, rg = generate_E_col_grid_R_np(0.01, R_high, 100, kernel)
    d = os.path.join(lacmopy_amd.DATA, "Ecol_grid_data", kernel) + os.sep
    E0, r0 = load_kernel_data_Ecol("Ecol_grid_R", d)[:2]
    np.testing.assert_array_equal(rg, r0)
    np.testing.assert_array_equal(E, E0)


def _box_ecol_moments(args):
    """Moments 0-3 (SI) every 200 s of one Ecol_grid_R box of the reference's expo ensemble (run_box_model.py:
    306-341: radii and Beard velocities of the masses, velocities updated after every step)."""
    kernel, seed, box, E, rg = args
    from lacmopy_amd.collision.box_model import golovin_expo_ensemble
    lib = orc.lib()
    lnR0, lnf, nb = math.log(rg[0]), math.log(rg[1] / rg[0]), len(rg)
    xi, m, _ = golovin_expo_ensemble(seed)
    rho = np.full(len(xi), 1000.0)
    R = np.array([lib.or_radius_from_mass(mm, 1000.0) for mm in m])
    vel = np.array([lib.or_terminal_velocity_Beard(r) for r in R])
    nc = np.zeros(2, dtype=np.int64)
    out = []
    for step in range(3601):
        if step % 200 == 0:
            out.append([np.sum(xi * (m * 1e-18) ** n) for n in range(4)])
        if step == 3600:
            break
        lib.or_collision_ecol_box(len(xi), orc.P(xi), orc.P(m), orc.P(R), orc.P(vel), orc.P(rho), 1.0, orc.P(E), nb,
                                  lnR0, lnf, orc.P(nc), None, 1001, box, step)
        lib.or_update_velocity_Beard(len(xi), orc.P(vel), orc.P(R))
    return np.array(out)


@pytest.mark.parametrize("kernel", ["Long_Bott", "Hall_Bott"])
def test_hydrodynamic_box_model_against_wang_2007(kernel):
    """Long and Hall box models (all_or_nothing.py:269-390 with kernel.py:107-210, 389-409 and the tables of
    kernel.py:512-694 on the box model's grid, 0.6-6000 um at 200 bins per decade, run_box_model.py:204)
    against the bin-model moments of Wang et al. (2007) that the reference holds
    (collision/ref_data/*/Wang_2007_moments.txt, every 200 s to 3600 s).  8 seeds of the expo ensemble at
    kappa = 40: lambda_0 within 5% and lambda_2 within 6% while the ensemble still resolves the spectrum
    (t <= 1800 s and t <= 1200 s; later, runaway growth at this SIP count sets in at seed-dependent times),
    lambda_1 within 0.5% throughout."""
    from concurrent.futures import ProcessPoolExecutor
    import multiprocessing as mp
    from lacmopy_amd.collision.kernel import generate_E_col_grid_R_np
    E, rg = generate_E_col_grid_R_np(0.6, 6.0E3, 200, kernel)
    ref = np.loadtxt(os.path.join(os.path.dirname(__file__), "golden", "ref_data", kernel, "Wang_2007_moments.txt"))
    times = np.loadtxt(os.path.join(os.path.dirname(__file__), "golden", "ref_data", kernel, "Wang_2007_times.txt"))
    np.testing.assert_array_equal(times, np.arange(0.0, 3601.0, 200.0))
    n_seeds = 8
    with ProcessPoolExecutor(max_workers=min(8, os.cpu_count() or 1), mp_context=mp.get_context("spawn")) as ex:
        mom = np.array(list(ex.map(_box_ecol_moments, [(kernel, 1001 + 2 * k, k, E, rg) for k in range(n_seeds)])))
    rel = mom.mean(axis=0).T / ref - 1.0
    assert np.max(np.abs(rel[0, times <= 1800.0])) < 0.05, rel[0]
    assert np.max(np.abs(rel[1])) < 0.005, rel[1]
    assert np.max(np.abs(rel[2, times <= 1200.0])) < 0.06, rel[2]
    # drops grow: the number falls below 70% of its start by 3600 s in every seed
    assert np.all(mom[:, -1, 0] < 0.7 * mom[:, 0, 0])
