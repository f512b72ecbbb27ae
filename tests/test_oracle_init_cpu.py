"""CPU pins of the oracle's restatement of the initial state (init.py:439-1236 and what it calls).

  * parameters derived from config.py's values against their closed forms;
  * the mass-fraction form of the Köhler equilibrium saturation against the radius form
    (a_w exp(2 sigma / (R_v T rho R))) with R from the particle mass, up to the reference's
    rho**0.66666667 for rho^(2/3);
  * the initial mass fraction against the reference's own algorithm run with scipy's fminbound and
    brentq (microphysics.py:866-966, 686-785);
  * SinSIP cells against the mathematics of the scheme: one mass in each kept bin, the weight threshold,
    and the mean over many cells of sum(w) and sum(w m), which estimate the integrals of the lognormal
    pdf and of m times it over the sampled mass range without bias;
  * the whole initial state of a full-height column: pressure chaining between levels, water and liquid
    budgets against the particles, convergence of the adjustment, S close to 1 in cloud, latent heating
    against the condensed water, multiplicities against the real particle counts, positions in cells;
  * the adjustment's saturation cap S_init_max: one iteration with and without it.
"""
import math

import numpy as np
import pytest
from scipy import optimize, stats

from init_cases import init_config, oracle_par
from lacmopy_amd.grid import Grid
from lacmopy_amd.initialize import init_parameters
from oracle import oracle as orc

# material_properties.py:400, 430: water-activity polynomials (highest degree first)
AW = {0: [1.412, -2.336, 3.113E-1, -2.715E-1, 1.0], 1: [1.518E1, -1.158E1, 8.642E-1, -6.366E-1, 1.0]}
W_MAX = {0: 0.78, 1: 0.45}
R_V = 461.53


def _par(cfg):
    g = Grid(cfg["grid_ranges"], cfg["grid_steps"], cfg["dy"])
    q = init_parameters(cfg, g)
    q.n_runs = 1
    return g, q


def test_init_parameters_closed_forms():
    g, q = _par(init_config(nx=75, nz=75))
    assert q.n_modes == 2 and q.solute == 0 and q.weak_threshold == 0
    assert list(q.l_max)[:2] == [35, 39]                      # ceil(4.3 * 8), ceil(4.8 * 8)
    assert abs(q.kappa[0] - 4.3) < 1e-12 and abs(q.kappa[1] - 4.8) < 1e-12
    assert list(q.no_rpcm_0)[:2] == [2.4E10, 1.6E10]            # ceil(400 m^3 * DNC0)
    for m, (r, s) in enumerate(((0.02, 1.4), (0.075, 1.6))):
        assert math.isclose(q.mu_m_log[m], math.log(4.0 / 3.0 * math.pi * 2163.0 * r ** 3), rel_tol=1e-14)
        assert math.isclose(q.sigma_m_log[m], 3.0 * math.log(s), rel_tol=1e-15)
    assert q.mass_density_dry == 1774.0 and q.rho_dry_0 == 101500.0 / (287.06 * 293.0)
    assert q.X == 1500.0 and q.dx == 20.0 and q.dz == 20.0 and q.dy == 1.0
    g1, q1 = _par(init_config(spcm=(0, 38), solute="NaCl", eta_threshold="weak"))
    assert q1.n_modes == 1 and q1.solute == 1 and q1.weak_threshold == 1 and q1.mass_density_dry == 2163.0
    assert q1.l_max[0] == 44 and abs(q1.kappa[0] - 5.4) < 1e-12      # ceil(38 / 20 * 28) / 10, ceil(5.4 * 8)
    assert q1.r_critmin[0] == 3E-3 and q1.no_rpcm_0[0] == 1.6E10
    assert math.isclose(q1.mu_m_log[0], math.log(4.0 / 3.0 * math.pi * 2163.0 * 0.075 ** 3), rel_tol=1e-14)


@pytest.mark.parametrize("solute", [0, 1])
def test_S_eq_mass_fraction_form_equals_radius_form(solute):
    L = orc.lib()
    for T in (275.0, 289.0, 300.0):
        for m_s in (1E-4, 3E-2, 5.0):
            for w in (0.02, 0.1, 0.3, W_MAX[solute]):
                m_w = m_s / w - m_s
                R, ws, rho = (np.zeros(1) for _ in range(3))
                L.or_R_p_w_s_rho_p(1, orc.P(np.array([m_w])), orc.P(np.array([m_s])), orc.P(np.array([T])),
                                   solute, orc.P(R), orc.P(ws), orc.P(rho))
                sigma = L.or_surface_tension(solute, w, T)
                # the reference writes rho^(2/3) as rho**0.66666667 (microphysics.py:444)
                arg = 2.0E6 * sigma / (R_V * T * rho[0] * R[0]) * rho[0] ** (2.0 / 3.0 - 0.66666667)
                expect = np.polyval(AW[solute], w) * math.exp(arg)
                got = L.or_S_eq_mass_fraction(solute, w, T, m_s)
                assert math.isclose(got, expect, rel_tol=1e-13), (solute, T, m_s, w, got, expect)


def _scipy_initial_mass_fraction(solute, m_s, S_amb, T):
    """microphysics.py:866-966 / 686-785 with scipy's routines, on the oracle's S_eq."""
    L = orc.lib()
    S = lambda w: L.or_S_eq_mass_fraction(solute, w, T, m_s)  # noqa: E731
    w_max = W_MAX[solute]
    if S_amb <= S(w_max):
        return w_max, "dry"
    w_act, _, flag, _ = optimize.fminbound(lambda w: -S(w), 1E-8, w_max, xtol=1.0E-12, full_output=True)
    assert flag == 0
    w_act *= 1.000001
    if S_amb > S(w_act):
        return w_act, "act"
    return optimize.brentq(lambda w: S(w) - S_amb, w_act, w_max, xtol=1e-15), "root"


@pytest.mark.parametrize("solute", [0, 1])
def test_initial_mass_fraction_against_scipy(solute):
    L = orc.lib()
    seen = set()
    for T in (280.0, 290.0):
        for m_s in (2E-5, 1E-3, 4E-2, 1.0, 30.0):
            S_act = max(L.or_S_eq_mass_fraction(solute, w, T, m_s) for w in np.geomspace(1E-8, W_MAX[solute], 4000))
            for S_amb in (0.3, 0.7, 0.95, 0.999, 1.0 + 0.5 * (S_act - 1.0), S_act + 1E-3):
                w_ref, branch = _scipy_initial_mass_fraction(solute, m_s, S_amb, T)
                w = L.or_initial_mass_fraction(solute, m_s, S_amb, T)
                seen.add(branch)
                if branch == "dry":
                    assert w == W_MAX[solute]
                elif branch == "act":
                    # the maximum is flat: its abscissa is fixed to ~sqrt(eps) by either search
                    assert math.isclose(w, w_ref, rel_tol=1e-5), (solute, T, m_s, S_amb, w, w_ref)
                    assert S_amb > L.or_S_eq_mass_fraction(solute, w, T, m_s)
                else:
                    assert math.isclose(w, w_ref, rel_tol=1e-12, abs_tol=1e-15), (solute, T, m_s, S_amb, w, w_ref)
                    resid = L.or_S_eq_mass_fraction(solute, w, T, m_s) - S_amb
                    assert abs(resid) < 1e-13, resid
    assert seen == {"dry", "act", "root"}


def test_sinsip_cells_bins_threshold_and_unbiased_integrals():
    g, q = _par(init_config())
    oq = oracle_par(q)
    for m in range(2):
        r = q.r_critmin[m]
        m_low = 4.0 * math.pi / 3.0 * q.mass_density_dry * r ** 3
        f = 10.0 ** (1.0 / q.kappa[m])
        l_max = q.l_max[m]
        m_high = m_low * f ** l_max
        dist = stats.lognorm(s=q.sigma_m_log[m], scale=math.exp(q.mu_m_log[m]))
        I0 = dist.cdf(m_high) - dist.cdf(m_low)
        I1 = dist.expect(lambda x: x, lb=m_low, ub=m_high)
        s0, s1 = [], []
        for cell in range(40):
            for level in range(50):
                mass, w = orc.sinsip_cell(oq, 99, level, cell, m)
                b = np.floor(np.log(mass / m_low) / np.log(f) + 1e-9).astype(int)
                assert np.all(np.diff(b) > 0) and b.min() >= 0 and b.max() < l_max, "one mass per bin, in its bin"
                lo = m_low * f ** b
                assert np.all(mass >= lo * (1 - 1e-12)) and np.all(mass < lo * f * (1 + 1e-12))
                assert np.all(w >= q.eta * w.max()), "fixed threshold"
                s0.append(w.sum())
                s1.append((w * mass).sum())
        s0, s1 = np.array(s0), np.array(s1)
        for s, I in ((s0, I0), (s1, I1)):
            se = s.std(ddof=1) / math.sqrt(len(s))
            assert abs(s.mean() - I) < 4.0 * se + 1e-6 * I, (m, s.mean(), I, se)
    # weak threshold: a bin below eta * max is kept with weight exactly eta * max, or dropped
    cfg = init_config(eta_threshold="weak", eta=1E-3)
    g, qw = _par(cfg)
    raised = 0
    for cell in range(30):
        mass, w = orc.sinsip_cell(oracle_par(qw), 7, 3, cell, 0)
        crit = qw.eta * w.max()
        assert np.all(w >= crit * (1 - 1e-15))
        raised += int(np.sum(w == crit))
    assert raised > 0


def test_initial_column_budgets_and_adjustment():
    cfg = init_config(nx=4, nz=75)
    g, q = _par(cfg)
    o = orc.init_SinSIP(oracle_par(q), 3711, g.corners[1][0], cfg["p_0"], g.corners[0][:, 0])
    lv, F = o["lv"], o["fields"]
    nx, nz = 4, 75
    r_tot = cfg["r_tot_0"]
    dx = dz = 20.0
    # pressure chaining and hydrostatics of every level
    # (each level re-evaluates its bottom pressure through the profile at z = z_0: (p/p_ref)^k^(1/k) p_ref)
    assert math.isclose(lv[0, 0], cfg["p_0"], rel_tol=1e-14)
    assert np.allclose(lv[1:, 0], lv[:-1, 1], rtol=1e-14, atol=0)
    assert np.allclose(lv[:, 6], (lv[:, 0] - lv[:, 1]) / (dz * (1 + r_tot) * 9.80665), rtol=1e-15)
    # level budgets: vapour + liquid = r_tot dry air; liquid = water on the level's particles
    assert np.allclose(lv[:, 11] + lv[:, 12], r_tot * lv[:, 10], rtol=1e-12)
    for j in range(nz):
        k = o["cell_z"] == j
        m_l = np.sum(o["m_w"][k] * o["xi"][k]) * 1E-18
        assert math.isclose(m_l, lv[j, 12], rel_tol=1e-9), (j, m_l, lv[j, 12])
        for i in range(nx):
            kc = k & (o["cell_x"] == i)
            r_l = np.sum(o["m_w"][kc] * o["xi"][kc]) * 1E-18 / (dx * 1.0 * dz * lv[j, 6])
            assert math.isclose(F[1, i, j], r_l, rel_tol=1e-12, abs_tol=1e-25)
        # multiplicities: the SinSIP weights of a cell sum to about one mode's real particles
        for m in range(2):
            km = k & (o["mode"] == m)
            ratio = o["xi"][km].sum() / (nx * lv[j, 17 + m])
            assert 0.85 < ratio < 1.15, (j, m, ratio)
    assert np.allclose(F[0] + F[1], r_tot, rtol=0, atol=1e-17)
    it = lv[:, 15]
    conv = it < cfg["iter_cnt_limit"]
    assert np.all(np.abs(lv[conv, 16]) <= 1e-5) and np.all(it >= 1)
    cloud = lv[:, 9] > 1E-5
    assert cloud.sum() > 20 and np.all(np.abs(lv[cloud, 7] - 1.0) < 2E-3), lv[cloud, 7]
    assert np.all(lv[~cloud, 7] < 1.0)
    # latent heat: the level's temperature rise over the dry adiabat at its bottom (Poisson's relation,
    # T = Theta_l (p / p_ref)^kappa with kappa of moist air) is L_v m_liq / C_p to first order
    kappa = 287.06 / 1004.71 * (1 - 0.289 * r_tot)
    for j in list(np.flatnonzero(cloud)[:4]) + [0, 10]:
        T_bot_dry = cfg["Theta_l"] * (lv[j, 0] / cfg["p_ref"]) ** kappa
        dT = lv[j, 2] - T_bot_dry
        L_v = 1.0E3 * (2500.82 - 2.358 * (T_bot_dry - 273.0))
        C_p = lv[j, 10] * 1004.71 + lv[j, 11] * 1906.0 + lv[j, 13] * 4183.8
        assert math.isclose(dT, lv[j, 12] * L_v / C_p, rel_tol=1e-2), (j, dT, lv[j, 12] * L_v / C_p)
        assert math.isclose(lv[j, 4], 0.5 * (lv[j, 2] + lv[j, 3]), rel_tol=1e-13)
    heat = lv[cloud, 2] - cfg["Theta_l"] * (lv[cloud, 0] / cfg["p_ref"]) ** kappa
    assert np.all(heat > 0) and heat.max() > 1.0 and np.all(np.diff(heat) > 0), "warming grows with liquid water"
    # positions inside their cells
    assert np.all((o["rel"] >= 0) & (o["rel"] < 1))
    assert np.allclose(o["x"], o["cell_x"] * dx + dx * o["rel"][0], rtol=0, atol=1e-12)
    assert np.allclose(o["z"], o["cell_z"] * dz + dz * o["rel"][1], rtol=0, atol=1e-12)


def test_adjustment_caps_the_saturation_of_its_condensation_steps():
    """init.py:903-905: the saturation adjustment condenses at min(S, S_init_max).  One adjustment iteration of a
    full-height column with the cap out of reach (S_init_max = 100) and with it at 1.0: levels whose saturation
    stays below the cap condense exactly alike, and every level where the cap acts condenses less (a smaller
    ambient saturation in the implicit mass update), in the supersaturated upper column.  A dropped cap makes the
    two runs equal everywhere; the cap taken as a lower bound (max) makes the capped run condense more."""
    runs = {}
    for cap in (100.0, 1.0):
        cfg = init_config(nx=4, nz=75, S_init_max=cap, iter_cnt_limit=1)
        g, q = _par(cfg)
        o = orc.init_SinSIP(oracle_par(q), 3711, g.corners[1][0], cfg["p_0"], g.corners[0][:, 0])
        runs[cap] = o["lv"]
    free, capped = runs[100.0], runs[1.0]
    assert np.all(free[:, 15] == 1) and np.all(capped[:, 15] == 1)
    m_free, m_cap = free[:, 12], capped[:, 12]
    same = m_free == m_cap
    # levels below the cap: bit-identical levels up to the first level the cap changes (the levels above start
    # from the changed state of the levels below)
    first = int(np.argmin(same)) if not same.all() else len(same)
    assert first > 0 and same[:first].all()
    assert not same.all(), "the cap never acted"
    # the cap acts where the column is supersaturated: there every level condenses less under it
    assert free[first, 7] > 1.0 and np.all(free[~same, 7] > 1.0)
    assert np.all(m_cap[~same] < m_free[~same]), np.flatnonzero(m_cap[~same] >= m_free[~same])
