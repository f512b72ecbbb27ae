"""Collision box model (collision/box_model.py of LacmoPy), batched on one GPU.

``simulate_collisions`` is the drop-in of box_model.py:44-204 for one box
and one seed (same arguments, same files).  ``simulate_collisions_batched``
runs many independent boxes (an ensemble of seeds) in one launch per time
step: box b is one wavefront, its SIPs live in LDS for the step, its draws
are keyed (seed, b, step).  Snapshots of xi and m are kept in HBM and copied
out once at the end.
"""
import math

import numpy as np

from .._native import check, lib
from .all_or_nothing import _p, _stream, _torch


FOUR_PI_OVER_THREE = 4.0 * math.pi / 3.0  # constants.py:28


def gen_mass_ensemble_weights_SinSIP_expo(m_mean, mass_density, dV, kappa, eta, weak_threshold, r_critmin,
                                          m_high_over_m_low=1.0E6, seed=3711, setseed=True, rng=None):
    """generation_SIP_ensemble.py:31-146: SinSIP ensemble of the exponential mass distribution
    f(m) = exp(-m/m_mean)/m_mean (Unterstrasser 2017): one mass drawn uniformly in each of the
    l_max = int(kappa log10(m_high/m_low)) logarithmic bins above m_low = 4/3 pi rho r_critmin^3, weight
    f(mu) * bin width; weights below eta * max are dropped (fixed threshold) or, weak threshold, raised to
    eta * max with probability w / (eta * max).  Returns (masses (fg), weights, m_low, bins).

    The draws are the reference's own: its jitted generator calls np.random.seed(seed) inside Numba code
    (:98), which seeds Numba's MT19937 by Knuth's init_genrand, and np.random.rand gives the 53-bit
    genrand_res53 doubles -- the same generator, seeding and construction as NumPy's legacy
    RandomState(seed).random_sample.  setseed=False continues `rng` (a RandomState), as the reference
    continues Numba's state.  Scalar exp is the C library's (Numba lowers np.exp of a float to libm).
    """
    if setseed:
        rng = np.random.RandomState(seed)
    elif rng is None:
        raise ValueError("setseed=False needs the RandomState to continue (rng=)")
    m_low = FOUR_PI_OVER_THREE * mass_density * r_critmin * r_critmin * r_critmin
    m_mean_inv = 1.0 / m_mean
    bin_factor = 10 ** (1.0 / kappa)
    m_high = m_low * m_high_over_m_low
    l_max = int(kappa * np.log10(m_high_over_m_low))
    rnd = rng.random_sample(l_max)
    rnd2 = rng.random_sample(l_max) if weak_threshold else None
    weights = np.zeros(l_max)
    masses = np.zeros(l_max)
    bins = np.zeros(l_max + 1)
    m_left = m_low
    bins[0] = m_left
    bin_n = 0
    while m_left < m_high:
        if bin_n >= l_max:
            # the reference's loop runs one bin past its arrays here (unchecked in Numba): no defined result
            raise ValueError(f"kappa = {kappa}: the bin loop of the reference overruns its {l_max} bins")
        m_right = m_left * bin_factor
        bin_width = m_right - m_left
        mu = m_left + rnd[bin_n] * bin_width
        weights[bin_n] = m_mean_inv * math.exp(-mu * m_mean_inv) * bin_width  # distributions.py:26-46
        masses[bin_n] = mu
        m_left = m_right
        bin_n += 1
        bins[bin_n] = m_left
    weight_critmin = weights.max() * eta
    valid = np.full(l_max, True)
    for b in range(l_max):
        if weights[b] < weight_critmin:
            if weak_threshold and rnd2[b] < weights[b] / weight_critmin:
                weights[b] = weight_critmin
            else:
                valid[b] = False
    return masses[valid], weights[valid], m_low, bins


def golovin_expo_ensemble(seed, kappa=40, eta=1.0E-9, weak_threshold=False, LWC0=1.0E-3, R_mean=9.3, r_critmin=0.6,
                          m_high_over_m_low=1.0E6, mass_density=1.0E3, dV=1.0):
    """(xis, masses in fg, DNC0) of one seed of the Golovin box-model case as run_box_model.py sets it up
    (:133-146, 236-245, 306-328): LWC0 = 1 g m^-3, R_mean = 9.3 um, DNC0 = LWC0 / m_mean, xi = DNC0 dV w."""
    c_radius_to_mass = 4.0E-18 * math.pi * mass_density / 3.0
    m_mean = c_radius_to_mass * R_mean ** 3
    DNC0 = LWC0 / m_mean
    masses, weights, _, _ = gen_mass_ensemble_weights_SinSIP_expo(1.0E18 * m_mean, mass_density, dV, kappa, eta,
                                                                   weak_threshold, r_critmin, m_high_over_m_low, seed)
    # run_box_model.py:321-323 stores 1e-18 * masses and simulates with 1e18 * (the stored kg masses)
    return (DNC0 * dV) * weights, 1.0E18 * (1.0E-18 * masses), DNC0


def compute_radius_from_mass_vec(mass, density):
    """microphysics.py:83-122: (3/(4 pi) m / rho)^(1/3) (um from fg), with the C library's pow as Numba's
    vectorized form calls it."""
    m = np.broadcast_to(np.asarray(mass, dtype=np.float64), np.broadcast(mass, density).shape)
    d = np.broadcast_to(np.asarray(density, dtype=np.float64), m.shape)
    c, third = 0.75 / math.pi, 1.0 / 3.0
    return np.array([math.pow(c * float(a) / float(b), third) for a, b in zip(m.ravel(), d.ravel())]).reshape(m.shape)


def ecol_expo_ensemble(seed, mass_density=1.0E3, **kw):
    """SIP quantities (xis, masses, radii, vel, mass_densities) of one seed of the expo ensemble for the
    hydrodynamic kernels, as run_box_model.py:325-333 builds them: radii from the masses, Beard velocities."""
    from .kernel import compute_terminal_velocity_Beard_vec
    xis, masses, _ = golovin_expo_ensemble(seed, mass_density=mass_density, **kw)
    mass_densities = np.ones_like(masses) * mass_density
    radii = compute_radius_from_mass_vec(masses, mass_densities)
    vel = compute_terminal_velocity_Beard_vec(radii)
    return xis, masses, radii, vel, mass_densities


def ensemble_statistics(xis_vs_time, masses_vs_time, mass_density, dV, no_bins):
    """The statistics of analyze_sim_data (box_model.py:1354-1585) for snapshots held in memory:
    xis_vs_time[s], masses_vs_time[s] of seed s with shape (no_times, N_s), masses in kg.

    Per time the mass bins span the smallest to the largest mass of all seeds (no_bins logarithmic bins, the
    outer edges widened by 1e-4), and per seed and time the moments 0-3 and the spectra f_m, g_m, g_ln_r are
    summed on the GPU (lm_box_stats), then averaged over the seeds with their standard errors
    (lm_mean_sem).  Returns a dict with the reference's arrays: moments_vs_time_avg/_std (no_times, 4),
    f_m_num_avg/std_vs_time, g_m_num_..., g_ln_r_num_... (no_times, no_bins), bins_mass_centers,
    bins_rad_centers (no_times, 3, no_bins), moments_vs_time_Unt, and the bins themselves.
    """
    torch = _torch()
    L = lib()
    no_sims = len(xis_vs_time)
    no_times = int(np.asarray(xis_vs_time[0]).shape[0])
    sizes = [int(np.asarray(x).shape[1]) for x in xis_vs_time]
    off_h = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    n_tot = int(off_h[-1])
    X = np.concatenate([np.asarray(x, dtype=np.float64) for x in xis_vs_time], axis=1)
    M = np.concatenate([np.asarray(m, dtype=np.float64) for m in masses_vs_time], axis=1)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()
    st = _stream(torch)
    dX, dM = dev(X), dev(M)
    mm = torch.empty((no_times, 4), dtype=torch.float64, device="cuda")
    check(L.lm_box_minmax(no_times, n_tot, _p(dX), _p(dM), _p(mm), st))
    mm = mm.cpu().numpy()
    # box_model.py:1460-1506, per time: bins, widths and the three kinds of bin centres
    bins_mass_vs_time = np.zeros((no_times, no_bins + 1))
    bins_mass_width_vs_time = np.zeros((no_times, no_bins))
    bins_rad_width_log_vs_time = np.zeros((no_times, no_bins))
    bin_factors_vs_time = np.zeros(no_times)
    bins_mass_centers, bins_rad_centers = [], []
    for t in range(no_times):
        m_min, m_max = mm[t, 0], mm[t, 1]
        bin_factor = (m_max / m_min) ** (1.0 / no_bins)
        bin_factors_vs_time[t] = bin_factor
        bins_mass = np.zeros(no_bins + 1)
        bins_mass[0] = m_min
        for b in range(1, no_bins + 1):
            bins_mass[b] = bins_mass[b - 1] * bin_factor
        bins_mass[0] *= 0.9999
        bins_mass[-1] *= 1.0001
        bins_mass_vs_time[t] = bins_mass
        bins_rad = compute_radius_from_mass_vec(1E18 * bins_mass, mass_density)
        bins_mass_log, bins_rad_log = np.log(bins_mass), np.log(bins_rad)
        width = bins_mass[1:] - bins_mass[:-1]
        bins_mass_width_vs_time[t] = width
        bins_rad_width_log_vs_time[t] = bins_rad_log[1:] - bins_rad_log[:-1]
        c_lin = 0.5 * (bins_mass[:-1] + bins_mass[1:])
        r_lin = 0.5 * (bins_rad[:-1] + bins_rad[1:])
        c_log = np.exp(0.5 * (bins_mass_log[:-1] + bins_mass_log[1:]))
        r_log = np.exp(0.5 * (bins_rad_log[:-1] + bins_rad_log[1:]))
        m_avg = M[t].sum() / X[t].sum()
        c_ex = bins_mass[:-1] + m_avg * np.log(width / (m_avg * (1 - np.exp(-width / m_avg))))
        r_ex = compute_radius_from_mass_vec(1E18 * c_ex, mass_density)
        bins_mass_centers.append(np.array((c_lin, c_log, c_ex)))
        bins_rad_centers.append(np.array((r_lin, r_log, r_ex)))
    off = torch.from_numpy(off_h).cuda()
    ed, wm, wr = dev(bins_mass_vs_time), dev(bins_mass_width_vs_time), dev(bins_rad_width_log_vs_time)
    mom = torch.empty((no_sims, no_times, 4), dtype=torch.float64, device="cuda")
    fm = torch.empty((no_sims, no_times, no_bins), dtype=torch.float64, device="cuda")
    gm, gr = torch.empty_like(fm), torch.empty_like(fm)
    check(L.lm_box_stats(no_times, no_sims, _p(off), n_tot, _p(dX), _p(dM), _p(ed), _p(wm), _p(wr), no_bins,
                         float(dV), _p(mom), _p(fm), _p(gm), _p(gr), st))
    out = {}

    def avg_sem(a, name):
        mean, sem = torch.empty(a.shape[1:], dtype=torch.float64, device="cuda"), torch.empty(
            a.shape[1:], dtype=torch.float64, device="cuda")
        check(L.lm_mean_sem(_p(a), no_sims, int(a[0].numel()), _p(mean), _p(sem), st))
        out[name + "_avg"], out[name + "_std"] = mean.cpu().numpy(), sem.cpu().numpy()

    avg_sem(mom, "moments_vs_time")
    avg_sem(fm, "f_m_num_vs_time")
    avg_sem(gm, "g_m_num_vs_time")
    avg_sem(gr, "g_ln_r_num_vs_time")
    out["moments_vs_time"] = np.transpose(mom.cpu().numpy(), (1, 2, 0))  # (no_times, 4, no_sims)
    # box_model.py:1543-1551
    unt = np.zeros((no_times, 4))
    for t in range(no_times):
        for n in range(4):
            unt[t, n] = math.log(bin_factors_vs_time[t]) / 3.0 * np.sum(
                out["g_ln_r_num_vs_time_avg"][t] * bins_mass_centers[t][1] ** (n - 1))
    out.update(moments_vs_time_Unt=unt, bins_mass_vs_time=bins_mass_vs_time,
               bins_mass_width_vs_time=bins_mass_width_vs_time, bins_rad_width_log_vs_time=bins_rad_width_log_vs_time,
               bins_mass_centers=np.array(bins_mass_centers), bins_rad_centers=np.array(bins_rad_centers),
               m_min_vs_time=mm[:, 0], m_max_vs_time=mm[:, 1], xi_min_vs_time=mm[:, 2], xi_max_vs_time=mm[:, 3],
               bin_factors_vs_time=bin_factors_vs_time)
    return out


def analyze_sim_data(kappa, mass_density, dV, no_sims, start_seed, no_bins, load_dir):
    """box_model.py:1354-1585: loads the snapshots xis_vs_time_{seed}.npy / masses_vs_time_{seed}.npy (fg) of
    the seeds start_seed, start_seed + 2, ... that simulate_collisions wrote, computes their statistics
    (ensemble_statistics) and writes the reference's files into load_dir."""
    seeds = np.arange(start_seed, start_seed + no_sims * 2, 2)
    xs = [np.load(load_dir + f"xis_vs_time_{seed}.npy") for seed in seeds]
    ms = [1E-18 * np.load(load_dir + f"masses_vs_time_{seed}.npy") for seed in seeds]
    r = ensemble_statistics(xs, ms, mass_density, dV, no_bins)
    end = f"no_sims_{no_sims}_no_bins_{no_bins}"
    for key, name in (("moments_vs_time_avg", "moments_vs_time_avg"), ("moments_vs_time_std", "moments_vs_time_std"),
                      ("f_m_num_vs_time_avg", "f_m_num_avg_vs_time"), ("f_m_num_vs_time_std", "f_m_num_std_vs_time"),
                      ("g_m_num_vs_time_avg", "g_m_num_avg_vs_time"), ("g_m_num_vs_time_std", "g_m_num_std_vs_time"),
                      ("g_ln_r_num_vs_time_avg", "g_ln_r_num_avg_vs_time"),
                      ("g_ln_r_num_vs_time_std", "g_ln_r_num_std_vs_time"),
                      ("bins_mass_centers", "bins_mass_centers"), ("bins_rad_centers", "bins_rad_centers")):
        np.save(load_dir + f"{name}_" + end + ".npy", r[key])
    return r


def _save_schedule(dt, t_end, dt_save):
    """box_model.py:152-157: steps, snapshots (t = 0, every dt_save, and t_end) and steps between snapshots."""
    no_steps = int(math.ceil(t_end / dt))
    no_saves = int(t_end / dt_save - 0.001) + 2
    dn_save = int(math.ceil(dt_save / dt))
    return no_steps, no_saves, dn_save


def simulate_collisions_batched(boxes, kernel_quantities, kernel_name, kernel_method, dV, dt, t_end, dt_save,
                                seed, step0=0):
    """Collision steps of many independent boxes.

    boxes: list of SIP_quantities tuples as box_model.py:44-204 takes them
    ((xis, masses) for Golovin and kernel_grid_m, (xis, masses, radii, vel,
    mass_densities) for Ecol_grid_R); updated in place.
    Returns (xis_vs_time, masses_vs_time, save_times, no_cols) with
    xis_vs_time[b] of shape (no_saves, N_b) and no_cols (n_boxes, 2).
    """
    torch = _torch()
    L = lib()
    n_boxes = len(boxes)
    sizes = [len(b[0]) for b in boxes]
    off_h = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
    n_tot = int(off_h[-1])
    if max(sizes) > 1024:
        raise ValueError("the box kernels hold at most 1024 SIPs per box")
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()
    cat = lambda k: np.concatenate([np.asarray(b[k], dtype=np.float64) for b in boxes])
    off = torch.from_numpy(off_h).cuda()
    xis, masses = dev(cat(0)), dev(cat(1))
    ecol = kernel_method == "Ecol_grid_R"
    if ecol:
        radii, vel, rho = dev(cat(2)), dev(cat(3)), dev(cat(4))
        E_col_grid, no_kernel_bins, R_kernel_low_log, bin_factor_R_log = kernel_quantities
        E = dev(E_col_grid)
    elif kernel_method == "kernel_grid_m":
        kernel_grid, no_kernel_bins, m_kernel_low_log, bin_factor_m_log = kernel_quantities
        Kg = dev(kernel_grid)
    elif not (kernel_method == "analytic" and kernel_name == "Golovin"):
        raise ValueError(f"unsupported kernel {kernel_name!r} / {kernel_method!r}")
    no_cols = torch.zeros((n_boxes, 2), dtype=torch.int64, device="cuda")
    no_steps, no_saves, dn_save = _save_schedule(dt, t_end, dt_save)
    snap_x = torch.zeros((no_saves, n_tot), dtype=torch.float64, device="cuda")
    snap_m = torch.zeros_like(snap_x)
    save_times = np.zeros(no_saves)
    dt_over_dV = dt / dV
    st = _stream(torch)
    seed = int(seed) & 0xFFFFFFFFFFFFFFFF
    save_n = 0
    for step_n in range(no_steps):
        if step_n % dn_save == 0:
            snap_x[save_n].copy_(xis)
            snap_m[save_n].copy_(masses)
            save_times[save_n] = step_n * dt
            save_n += 1
        step = (step0 + step_n) & 0xFFFFFFFF
        if ecol:
            rc = L.lm_ecol_box_step(n_boxes, _p(off), _p(xis), _p(masses), _p(radii), _p(vel), _p(rho),
                                    float(dt_over_dV), _p(E), int(no_kernel_bins), float(R_kernel_low_log),
                                    float(bin_factor_R_log), _p(no_cols), seed, step, None, 1, st)
        elif kernel_method == "kernel_grid_m":
            rc = L.lm_kernel_grid_m_step(n_boxes, _p(off), _p(xis), _p(masses), float(dt_over_dV), _p(Kg),
                                         int(no_kernel_bins), float(m_kernel_low_log), float(bin_factor_m_log),
                                         _p(no_cols), seed, step, None, st)
        else:
            rc = L.lm_golovin_step(n_boxes, _p(off), _p(xis), _p(masses), float(dt_over_dV), _p(no_cols), seed,
                                   step, None, st)
        check(rc)
    snap_x[save_n].copy_(xis)
    snap_m[save_n].copy_(masses)
    save_times[save_n] = no_steps * dt
    check(L.lm_box_check(st))
    hx, hm = snap_x.cpu().numpy(), snap_m.cpu().numpy()
    fin = {"x": xis.cpu().numpy(), "m": masses.cpu().numpy()}
    if ecol:
        fin["R"], fin["v"] = radii.cpu().numpy(), vel.cpu().numpy()
    xs, ms = [], []
    for b, box in enumerate(boxes):
        sl = slice(off_h[b], off_h[b + 1])
        xs.append(hx[:, sl].copy())
        ms.append(hm[:, sl].copy())
        box[0][...] = fin["x"][sl]
        box[1][...] = fin["m"][sl]
        if ecol:
            box[2][...] = fin["R"][sl]
            box[3][...] = fin["v"][sl]
    return xs, ms, save_times, no_cols.cpu().numpy()


def simulate_collisions(SIP_quantities, kernel_quantities, kernel_name, kernel_method, dV, dt, t_end, dt_save,
                        no_cols, seed, save_dir):
    """box_model.py:44-204 for one box: in-place update, snapshots saved as xis/masses/save_times_{seed}.npy.

    The draws are keyed by seed (box 0); no_cols accumulates the collision counts.
    """
    xs, ms, save_times, nc = simulate_collisions_batched([SIP_quantities], kernel_quantities, kernel_name,
                                                         kernel_method, dV, dt, t_end, dt_save, seed)
    no_cols[...] = no_cols + nc[0]
    np.save(save_dir + f"xis_vs_time_{seed}", xs[0])
    np.save(save_dir + f"masses_vs_time_{seed}", ms[0])
    np.save(save_dir + f"save_times_{seed}", save_times)
