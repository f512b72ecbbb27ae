"""Synthetic inputs of the 2-D kinematic drizzling stratocumulus case.

Host-side set-up (not part of the step).  The thermodynamic profile,
the dry-air mass-flux field and the velocity staggering follow the
reference's initialisation (init.py:384-437, 239-273, 1083-1182); the SIP
ensemble is a seeded Poisson sampler, a quick stand-in for the reference's
SinSIP + saturation adjustment (init.py:439-1236, which
lacmopy_amd.initialize computes on the GPU) with collisions firing at once:

  * per-cell SIP count ~ Poisson(spc), positions uniform in the cell;
  * dry radii from the bimodal lognormal of config.py:106-114
    (N = 60 / 40 cm^-3, r_med = 0.02 / 0.075 um, sigma = 1.4 / 1.6),
    ammonium sulfate, m_s = 4/3 pi 1774 r^3 (fg, r in um);
  * xi = N_mode * V_cell * (rho_d / rho_d0) / expected SIPs of that mode;
  * water: haze with w_s = 0.5 outside the cloud layer, lognormal droplets
    (median 9 um, sigma 1.35) between 600 m and 1000 m so that collisions
    fire; r_v = r_tot - r_l per cell.
"""
import math

import numpy as np

from .grid import Grid

R_D = 287.06
R_V = 461.53
C_PD = 1004.71
G = 9.80665
KAPPA_DRY = R_D / C_PD
EPS_GC = R_D / R_V
RHO_AS_DRY = 1774.0


def _profiles_without_liquid(z, z0, p0, p_ref, Theta_l, r_tot):
    """init.py:384-437 (with atmosphere.py:77-166)."""
    p_0_over_p_ref = p0 / p_ref
    kappa_tot = KAPPA_DRY * (1 - 0.289 * r_tot)
    p_0_over_p_ref_to_kappa_tot = p_0_over_p_ref ** kappa_tot
    beta_tot = G * kappa_tot * (1 + r_tot) / (R_D * Theta_l * (1 + r_tot / EPS_GC))
    T_over_Theta_l = p_0_over_p_ref_to_kappa_tot - beta_tot * (z - z0)
    T = T_over_Theta_l * Theta_l
    p = T_over_Theta_l ** (1.0 / kappa_tot) * p_ref
    rho_dry = p * beta_tot / (T_over_Theta_l * G * kappa_tot)
    return T, p, rho_dry


# Benchmark workloads.  icmw75: the 2-D ICMW kinematic stratocumulus on 75x75 cells of 20 m, Poisson(128) SIPs
# per cell, Long E_col, dt 1 s, 10 substeps, 64 seeds.  stress150: the same 1500 m domain on 150x150 cells of
# 10 m, Poisson(256) SIPs per cell, Hall E_col, mass-flux amplitude 1.0 kg m^-2 s^-1, dt 0.5 s, 20 substeps,
# 16 seeds.  Both: 2 collision steps per advection step, relaxation on.  Run r of a batch is seeded
# seed0 + 2 r (the reference's run_lacmo.sh:7-16 seed spacing).
WORKLOADS = {
    "icmw75": dict(nx=75, nz=75, dx=20.0, dz=20.0, spc=128.0, runs=64, kernel="Long_Bott", dt=1.0, sdc=5,
                   j_max=0.6),
    "stress150": dict(nx=150, nz=150, dx=10.0, dz=10.0, spc=256.0, runs=16, kernel="Hall_Bott", dt=0.5, sdc=10,
                      j_max=1.0),
}


def workload_config(case, E_col_grid=None, R_kernel_low_log=None, bin_factor_R_log=None):
    """Step configuration of a workload, with the keys simulate() reads (integration.py:1908-1923)."""
    c = WORKLOADS[case]
    cfg = dict(solute_type="AS", dt_adv=c["dt"], dt_col=c["dt"] / 2, scale_dt_cond=c["sdc"], no_col_per_adv=2,
               no_iter_impl_mass=3, g_set=9.80665, act_collisions=True, act_relaxation=True)
    if E_col_grid is not None and not isinstance(E_col_grid, float):
        cfg.update(no_kernel_bins=int(np.asarray(E_col_grid).shape[0]), R_kernel_low_log=R_kernel_low_log,
                   bin_factor_R_log=bin_factor_R_log)
    return cfg


def workload_case(case, seed, spc=None, return_draws=False):
    """make_case for one run of a workload (spc: mean SIPs per cell, default the workload's)."""
    c = WORKLOADS[case]
    return make_case(nx=c["nx"], nz=c["nz"], dx=c["dx"], dz=c["dz"], spc=c["spc"] if spc is None else spc,
                     seed=seed, j_max_ref=c["j_max"], return_draws=return_draws)


def make_case(nx=75, nz=75, dx=20.0, dz=20.0, spc=128.0, seed=1001, p_0=101500.0, p_ref=1.0E5,
              Theta_l=289.0, r_tot_0=7.5E-3, j_max_ref=0.6, cloud_base=600.0, cloud_top=1000.0,
              DNC0=(60.0E6, 40.0E6), mu_R=(0.02, 0.075), sigma_R=(1.4, 1.6), R_drop=9.0, sigma_drop=1.35,
              return_draws=False):
    """Returns (grid, pos, vel, cells, m_w, m_s, xi, active_ids) in the reference's layout.

    return_draws: also return the draws that map the aerosol-mode parameters to the SIPs' dry masses
    and multiplicities (see aerosol_sips)."""
    rng = np.random.default_rng(seed)
    grid = Grid([[0.0, nx * dx], [0.0, nz * dz]], [dx, dz], 1.0)
    nx, nz = int(grid.no_cells[0]), int(grid.no_cells[1])
    levels = grid.corners[1][0]
    p_bot = p_0
    T_c, p_c, rho_c = np.zeros(nz), np.zeros(nz), np.zeros(nz)
    for j in range(nz):
        Tb, pb, rb = _profiles_without_liquid(levels[j], levels[j], p_bot, p_ref, Theta_l, r_tot_0)
        Tt, pt, rt = _profiles_without_liquid(levels[j + 1], levels[j], p_bot, p_ref, Theta_l, r_tot_0)
        T_c[j], p_c[j], rho_c[j] = 0.5 * (Tb + Tt), 0.5 * (pb + pt), 0.5 * (rb + rt)
        p_bot = pt

    # SIP ensemble
    counts = rng.poisson(spc, size=(nx, nz)).astype(np.int64)
    N = int(counts.sum())
    cx = np.repeat(np.repeat(np.arange(nx), nz), counts.ravel())
    cz = np.repeat(np.tile(np.arange(nz), nx), counts.ravel())
    x = (cx + rng.random(N)) * dx
    z = (cz + rng.random(N)) * dz
    frac0 = DNC0[0] / (DNC0[0] + DNC0[1])
    mode = (rng.random(N) >= frac0).astype(np.int64)
    mu = np.array(mu_R)[mode]
    sig = np.array(sigma_R)[mode]
    z_draw = rng.standard_normal(N)
    r_dry = mu * np.exp(np.log(sig) * z_draw)
    m_s = 4.0 * math.pi / 3.0 * RHO_AS_DRY * r_dry ** 3
    rho_scale = rho_c[cz] / rho_c[0]
    exp_per_mode = np.where(mode == 0, spc * frac0, spc * (1 - frac0))
    xi = np.array(DNC0)[mode] * grid.volume_cell * rho_scale / exp_per_mode
    in_cloud = (z >= cloud_base) & (z < cloud_top)
    R_w = R_drop * np.exp(np.log(sigma_drop) * rng.standard_normal(N))
    m_drop = 4.0 * math.pi / 3.0 * 1000.0 * R_w ** 3
    m_w = np.where(in_cloud, np.maximum(m_drop - m_s, m_s), m_s)   # haze: w_s = 0.5

    # grid fields
    m_dry_cell = rho_c * grid.volume_cell  # kg per cell, by level
    r_l = np.zeros((nx, nz))
    np.add.at(r_l, (cx, cz), m_w * xi)
    r_l *= 1.0E-18 / m_dry_cell[None, :]
    r_v = np.maximum(r_tot_0 - r_l, 0.5 * r_tot_0)
    T = np.repeat(T_c[None, :], nx, axis=0)
    p = np.repeat(p_c[None, :], nx, axis=0)
    rho_d = np.repeat(rho_c[None, :], nx, axis=0)
    grid.temperature, grid.pressure, grid.mass_density_air_dry = T, p, rho_d
    grid.mixing_ratio_water_vapor, grid.mixing_ratio_water_liquid = r_v, r_l
    p_dry = rho_d * R_D * T
    grid.potential_temperature = T * (1.0E5 / p_dry) ** KAPPA_DRY           # init.py:1143
    e_s = 2.53E11 * np.exp(-5420.0 / T)
    grid.saturation_pressure = e_s
    grid.saturation = rho_d * r_v * T * R_V / e_s
    grid.update_material_properties()
    grid.rho_dry_inv = np.ones_like(rho_d) / rho_d
    grid.mass_dry_inv = (1.0 / grid.volume_cell) * grid.rho_dry_inv

    kinematic_flow(grid, rho_c, j_max_ref)

    pos = np.array([x, z])
    cells = np.array([cx, cz], dtype=np.int64)
    # initial SIP velocity: the flow velocity at the SIP (host bilinear interpolation)
    vel = np.array(interpolate_flow(grid, cells, pos))
    active = np.ones(N, dtype=bool)
    if return_draws:
        draws = dict(mode=mode, z=z_draw, rho_scale=rho_scale, exp_per_mode=exp_per_mode,
                     volume_cell=float(grid.volume_cell))
        return grid, pos, vel, cells, m_w, m_s, xi, active, draws
    return grid, pos, vel, cells, m_w, m_s, xi, active


def aerosol_sips(theta, draws):
    """Dry masses and multiplicities of the SIPs from the six aerosol-mode parameters.

    theta = (N_0 [m^-3], r_med_0 [um], sigma_0, N_1, r_med_1, sigma_1) as a float64 torch tensor (may
    require grad); draws from make_case(..., return_draws=True).  Same map as make_case:
    r_dry = r_med exp(ln(sigma) z), m_s = 4/3 pi 1774 r_dry^3 (1e-18 kg), xi = N V (rho_d / rho_d0) /
    expected SIPs of the mode.  Returns (xi, m_s) in ID order, differentiable in theta.
    """
    import torch
    mode = torch.as_tensor(draws["mode"], device=theta.device)
    z = torch.as_tensor(draws["z"], dtype=torch.float64, device=theta.device)
    scale = torch.as_tensor(draws["rho_scale"] / draws["exp_per_mode"] * draws["volume_cell"], dtype=torch.float64,
                            device=theta.device)
    N = theta[0::3][mode]
    r_med = theta[1::3][mode]
    sig = theta[2::3][mode]
    r_dry = r_med * torch.exp(torch.log(sig) * z)
    m_s = 4.0 * math.pi / 3.0 * RHO_AS_DRY * r_dry ** 3
    xi = N * scale
    return xi, m_s


def kinematic_flow(grid, rho_c, j_max_ref=0.6):
    """Dry-air mass flux of the ICMW 2012 case 1 stream function and the velocity from it (init.py:239-273,
    1146-1182; the velocity divides by the dry density at the faces built from the level centres rho_c)."""
    dx, dz = grid.steps
    X, Z = grid.sizes
    j_max = j_max_ref * math.sqrt(X * X + Z * Z) / math.sqrt(2.0 * 1500.0 * 1500.0)
    k_x, k_z = 2.0 * np.pi / X, np.pi / Z
    ux, uz = grid.corners[0], grid.corners[1] + 0.5 * dz
    wx, wz = grid.corners[0] + 0.5 * dx, grid.corners[1]
    j_x = j_max * (X / Z) * np.cos(k_x * ux) * np.cos(k_z * uz)
    j_z = 2 * j_max * np.sin(k_x * wx) * np.sin(k_z * wz)
    grid.mass_flux_air_dry = np.array([j_x, j_z])
    rho_bot = 0.5 * (rho_c[0:-1] + rho_c[1:])
    rho_bot = np.insert(rho_bot, 0, 0.5 * (3.0 * rho_c[0] - rho_c[1]))
    rho_bot = np.append(rho_bot, 0.5 * (3.0 * rho_c[-1] - rho_c[-2]))
    grid.velocity = np.array([j_x / rho_bot, j_z / np.append(rho_c, rho_c[-1])])


def interpolate_flow(grid, cells, pos, rel=None):
    """Host interpolation of the face velocities at SIP positions (grid.py:257-323); rel: the positions relative
    to the cell (2, N) where known (init.py:1228-1230 passes the uniform draws), else from pos."""
    nz = int(grid.no_cells[1])
    U, V = grid.velocity
    i, j = cells
    if rel is None:
        a = pos[0] / grid.steps[0] - i
        b = pos[1] / grid.steps[1] - j
    else:
        a, b = rel

    def bil(ii, jj, aa, bb, f):
        return aa * (bb * f[ii + 1, jj + 1] + (1 - bb) * f[ii + 1, jj]) + \
            (1 - aa) * (bb * f[ii, jj + 1] + (1 - bb) * f[ii, jj])

    lin = ((j == 0) & (b <= 0.5)) | ((j == nz - 1) & (b >= 0.5))
    u_lin = a * U[i + 1, j] + (1 - a) * U[i, j]
    jb = np.where(b > 0.5, j, j - 1)
    bb = np.where(b > 0.5, b - 0.5, b + 0.5)
    u_bil = bil(i, np.clip(jb, 0, nz - 1), a, bb, U)
    u = np.where(lin, u_lin, u_bil)
    ib = np.where(a > 0.5, i, i - 1)
    aa = np.where(a > 0.5, a - 0.5, a + 0.5)
    v = bil(ib, j, aa, b, V)
    return u, v


def init_profiles(grid):
    """Relaxation profiles and time scale as simulate() derives them (integration.py:1948-1955)."""
    prof_rv = np.average(grid.mixing_ratio_water_vapor, axis=0)
    prof_Th = np.average(grid.potential_temperature, axis=0)
    t_relax = 300 * np.exp(grid.centers[1][0] / 200)
    return prof_rv, prof_Th, t_relax
