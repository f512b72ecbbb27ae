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
