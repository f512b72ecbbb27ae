"""CPU: the aerosol-mode map of a SinSIP initial state (lacmopy_amd.initialize.sinsip_mode_map) against the
SinSIP generator itself.

The oracle's restatement of the initialisation (init.py:439-1236, generation_SIP_ensemble.py:147-404) is run
at the configuration's mode parameters and again with one parameter changed by a relative h; the generator's
multiplicities of the second run must equal the map's prediction from the first state, SIP by SIP.  The map
holds what the generator's second run changes only through rounding or through the levels below (the real
particles per cell, rint(ceil(V N) rho_d / rho_d0), follow the adjusted level densities), so the agreement is
limited by those: measured 1e-9 relative or better.  The dry masses must come out identical (the bins and
draws do not depend on the mode parameters).  A planted mistake in the map -- the pdf without its 1/sigma
factor, or r_med entering the mass as r_med^2 -- fails by orders of magnitude.
"""
import types

import numpy as np
import pytest
import torch

from init_cases import init_config, oracle_par
from lacmopy_amd.grid import Grid
from lacmopy_amd.initialize import init_parameters, mode_parameters, sinsip_mode_map
from oracle import oracle as orc

KEYS = ("DNC0", "mu_R", "sigma_R")


def _oracle_state(cfg):
    g = Grid(cfg["grid_ranges"], cfg["grid_steps"], cfg["dy"])
    q = init_parameters(cfg, g)
    q.n_runs = 1
    s = orc.init_SinSIP(oracle_par(q), cfg["seed_SIP_gen"], g.corners[1][0], cfg["p_0"], g.corners[0][:, 0])
    return types.SimpleNamespace(xi=s["xi"], m_s=s["m_s"], modes=s["mode"].astype(np.int64))


@pytest.fixture(scope="module")
def base():
    cfg = init_config(nx=2, nz=75)
    return cfg, _oracle_state(cfg)


def test_map_is_the_identity_at_the_configuration(base):
    cfg, s0 = base
    theta = torch.tensor(mode_parameters(cfg), dtype=torch.float64)
    xi, m_s = sinsip_mode_map(theta, s0, cfg)
    assert np.array_equal(xi.numpy(), s0.xi) and np.array_equal(m_s.numpy(), s0.m_s)


@pytest.mark.parametrize("k", range(6))
def test_map_predicts_the_generator(base, k):
    cfg, s0 = base
    h = 1.0E-3
    theta = mode_parameters(cfg).copy()
    theta[k] *= 1.0 + h
    cfg1 = dict(cfg)
    mode, key = divmod(k, 3)
    arr = np.array(cfg[KEYS[key]], dtype=np.float64)
    arr[mode] = theta[k]
    cfg1[KEYS[key]] = arr
    s1 = _oracle_state(cfg1)
    # the same draws: every mass of one run is a mass of the other, except the few bins whose weight crossed
    # the threshold eta * max weight (the map holds the kept set)
    both0 = np.isin(s0.m_s, s1.m_s)
    both1 = np.isin(s1.m_s, s0.m_s)
    assert both0.sum() == both1.sum() and (~both0).sum() + (~both1).sum() <= 1.0E-2 * len(s0.xi)
    assert np.array_equal(s0.m_s[both0], s1.m_s[both1]) and np.array_equal(s0.modes[both0], s1.modes[both1])
    xi, _ = sinsip_mode_map(torch.tensor(theta, dtype=torch.float64), s0, cfg)
    rel = np.abs(xi.numpy()[both0] / s1.xi[both1] - 1.0)
    # the change the map has to predict is at least h-sized for the parameter's own mode
    moved = np.abs(s1.xi[both1] / s0.xi[both0] - 1.0)[s0.modes[both0] == mode]
    assert np.median(moved) > 1.0E-5
    assert rel.max() < 1.0E-9, f"map vs generator: max relative deviation {rel.max():.3e}"


def test_map_gradient_matches_its_difference_quotient(base):
    """autograd through the map against central differences of the map (h = 1e-6 relative)."""
    cfg, s0 = base
    theta0 = torch.tensor(mode_parameters(cfg), dtype=torch.float64)
    rng = np.random.default_rng(3)
    lam = torch.tensor(rng.standard_normal(len(s0.xi)) / s0.xi, dtype=torch.float64)
    th = theta0.clone().requires_grad_(True)
    xi, _ = sinsip_mode_map(th, s0, cfg)
    (g,) = torch.autograd.grad((xi * lam).sum(), th)
    for k in range(6):
        h = 1.0E-6 * float(theta0[k])
        f = []
        for sgn in (1.0, -1.0):
            t = theta0.clone()
            t[k] += sgn * h
            f.append(float((sinsip_mode_map(t, s0, cfg)[0] * lam).sum()))
        fd = (f[0] - f[1]) / (2 * h)
        assert abs(fd - float(g[k])) <= 1e-6 * abs(fd) + 1e-12, (k, fd, float(g[k]))
