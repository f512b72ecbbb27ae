"""Configurations of the initial-state tests: the reference's config.py:76-132 values on narrower grids."""
import ctypes as C

import numpy as np


def init_config(nx=8, nz=75, solute="AS", spcm=(26, 38), eta_threshold="fix", seed=3711, **kw):
    """config.py:76-132 (ICMW 2012 case 1 profile, two lognormal modes), nx x nz cells of 20 m x 20 m."""
    c = dict(grid_ranges=[[0.0, 20.0 * nx], [0.0, 20.0 * nz]], grid_steps=[20.0, 20.0], dy=1.0, p_0=101500.0,
             p_ref=1.0E5, r_tot_0=7.5E-3, Theta_l=289.0, solute_type=solute, no_spcm=np.array(spcm),
             dist="lognormal", DNC0=np.array([60.0E6, 40.0E6]), mu_R=0.5 * np.array([0.04, 0.15]),
             sigma_R=np.array([1.4, 1.6]), eta=1E-10, eta_threshold=eta_threshold,
             r_critmin=np.array([1., 3.]) * 1E-3, m_high_over_m_low=1.0E8, S_init_max=1.04, dt_init=0.1,
             Newton_iterations=2, iter_cnt_limit=4000, seed_SIP_gen=seed)
    c.update(kw)
    return c


def oracle_par(q):
    """The oracle's copy of the product's LmInitPar (same layout)."""
    from oracle import oracle as orc
    oq = orc.OrInitPar()
    assert C.sizeof(oq) == C.sizeof(q)
    C.memmove(C.addressof(oq), C.addressof(q), C.sizeof(q))
    return oq
