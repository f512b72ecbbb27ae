"""Drop-in all-or-nothing collision steps (collision/all_or_nothing.py of LacmoPy).

Same names, positional arguments and in-place effects as the reference; each
call runs one HIP launch of liblacmo_amd (one wavefront per box or cell,
pairs visited by anti-diagonals, which reproduces the sequential row-major
pair loop exactly).  The draws come from lacmopy_amd.rng: every call consumes
one collision-step counter value.  ``rnd`` (keyword only, not in the
reference) injects the uniforms instead, in the reference's consumption
order (pair cnt = i*N - i(i+1)/2 + (j-i-1); cells x-major for the all-cells
step), which is how a stream drawn elsewhere is replayed.
"""
import ctypes as C

import numpy as np

from .. import rng
from .._native import check, lib

__all__ = [
    "collision_step_Golovin_np", "collision_step_Golovin", "collision_step_Ecol_grid_R_np",
    "collision_step_Ecol_grid_R", "collision_step_Ecol_grid_R_2D_multicomp_np",
    "collision_step_Ecol_grid_R_2D_multicomp", "collision_step_Ecol_const_2D_multicomp_np",
    "collision_step_Ecol_const_2D_multicomp", "collision_step_Ecol_grid_R_all_cells_2D_multicomp_np",
    "collision_step_Ecol_grid_R_all_cells_2D_multicomp",
]


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("lacmopy_amd collision steps need a GPU (HIP device); there is no CPU fallback")
    return torch


def _dev(torch, a, dtype=None):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=dtype or np.float64)).cuda()


def _p(t):
    return C.c_void_p(t.data_ptr()) if t is not None else C.c_void_p(0)


def _stream(torch):
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _rnd(torch, rnd, n):
    if rnd is None:
        return None
    r = np.ascontiguousarray(rnd, dtype=np.float64).reshape(-1)
    if len(r) < n * (n - 1) // 2:
        raise ValueError(f"rnd holds {len(r)} uniforms, the step needs {n * (n - 1) // 2}")
    return _dev(torch, r)


def _finish(torch, L, no_cols, nc_dev):
    check(L.lm_box_check(_stream(torch)))
    no_cols[...] = no_cols + nc_dev.cpu().numpy().reshape(2)


def collision_step_Golovin_np(xis, masses, dt_over_dV, no_cols, *, rnd=None):
    """all_or_nothing.py:109-185: Golovin kernel K = b (m_i + m_j), box model."""
    torch = _torch()
    L = lib()
    n = len(xis)
    x, m = _dev(torch, xis), _dev(torch, masses)
    off = _dev(torch, [0, n], np.int32)
    nc = torch.zeros(2, dtype=torch.int64, device="cuda")
    r = _rnd(torch, rnd, n)
    check(L.lm_golovin_step(1, _p(off), _p(x), _p(m), float(dt_over_dV), _p(nc), rng.get_seed(), rng.take_steps(1),
                            _p(r), _stream(torch)))
    _finish(torch, L, no_cols, nc)
    xis[...] = x.cpu().numpy()
    masses[...] = m.cpu().numpy()


collision_step_Golovin = collision_step_Golovin_np


def collision_step_Ecol_grid_R_np(xis, masses, radii, vel, mass_densities, dt_over_dV, E_col_grid, no_kernel_bins,
                                  R_kernel_low_log, bin_factor_R_log, no_cols, *, rnd=None):
    """all_or_nothing.py:269-390: hydrodynamic kernel with tabulated E_col(R1, R2), box model, |v| 1-D."""
    torch = _torch()
    L = lib()
    n = len(xis)
    x, m, R, v, rho = (_dev(torch, a) for a in (xis, masses, radii, vel, mass_densities))
    E = _dev(torch, E_col_grid)
    off = _dev(torch, [0, n], np.int32)
    nc = torch.zeros(2, dtype=torch.int64, device="cuda")
    r = _rnd(torch, rnd, n)
    check(L.lm_ecol_box_step(1, _p(off), _p(x), _p(m), _p(R), _p(v), _p(rho), float(dt_over_dV), _p(E),
                             int(no_kernel_bins), float(R_kernel_low_log), float(bin_factor_R_log), _p(nc),
                             rng.get_seed(), rng.take_steps(1), _p(r), 0, _stream(torch)))
    _finish(torch, L, no_cols, nc)
    xis[...] = x.cpu().numpy()
    masses[...] = m.cpu().numpy()
    radii[...] = R.cpu().numpy()


collision_step_Ecol_grid_R = collision_step_Ecol_grid_R_np


def _box2d(xis, m_w, m_s, radii, vel, mass_densities, dt_over_dV, E_col_grid, E_const, nb, lnR0, lnf, no_cols, rnd):
    torch = _torch()
    L = lib()
    n = len(xis)
    x, mw, ms, R, rho = (_dev(torch, a) for a in (xis, m_w, m_s, radii, mass_densities))
    vx, vz = _dev(torch, vel[0]), _dev(torch, vel[1])
    E = _dev(torch, E_col_grid) if E_col_grid is not None else None
    off = _dev(torch, [0, n], np.int32)
    nc = torch.zeros(2, dtype=torch.int64, device="cuda")
    r = _rnd(torch, rnd, n)
    check(L.lm_box2d_step(1, _p(off), _p(x), _p(mw), _p(ms), _p(R), _p(vx), _p(vz), _p(rho), float(dt_over_dV),
                          _p(E), float(E_const), int(nb), float(lnR0), float(lnf), _p(nc), rng.get_seed(),
                          rng.take_steps(1), _p(r), _stream(torch)))
    _finish(torch, L, no_cols, nc)
    xis[...] = x.cpu().numpy()
    m_w[...] = mw.cpu().numpy()
    m_s[...] = ms.cpu().numpy()
    radii[...] = R.cpu().numpy()


def collision_step_Ecol_grid_R_2D_multicomp_np(xis, m_w, m_s, radii, vel, mass_densities, dt_over_dV, E_col_grid,
                                               no_kernel_bins, R_kernel_low_log, bin_factor_R_log, no_cols, *,
                                               rnd=None):
    """all_or_nothing.py:498-635: one cell, two mass components, tabulated E_col, 2-D velocities."""
    _box2d(xis, m_w, m_s, radii, vel, mass_densities, dt_over_dV, E_col_grid, 0.0, no_kernel_bins, R_kernel_low_log,
           bin_factor_R_log, no_cols, rnd)


collision_step_Ecol_grid_R_2D_multicomp = collision_step_Ecol_grid_R_2D_multicomp_np


def collision_step_Ecol_const_2D_multicomp_np(xis, m_w, m_s, radii, vel, mass_densities, dt_over_dV, E_col_grid,
                                              no_cols, *, rnd=None):
    """all_or_nothing.py:637-751: as the tabulated version with a constant E_col (E_col_grid is a float)."""
    _box2d(xis, m_w, m_s, radii, vel, mass_densities, dt_over_dV, None, float(E_col_grid), 0, 0.0, 1.0, no_cols,
           rnd)


collision_step_Ecol_const_2D_multicomp = collision_step_Ecol_const_2D_multicomp_np


def collision_step_Ecol_grid_R_all_cells_2D_multicomp_np(xis, m_w, m_s, vel, grid_temperature, cells, no_cells,
                                                         dt_over_dV, E_col_grid, no_kernel_bins, R_kernel_low_log,
                                                         bin_factor_R_log, no_cols, solute_type, *, rnd=None):
    """all_or_nothing.py:754-846: one collision step in every cell of the 2-D grid.

    Members of cell (i, j) are the SIPs with cells == (i, j) in ascending
    index; radii and densities from m_w, m_s at the cell temperature; a float
    E_col_grid selects the constant-efficiency kernel.  Updates xis, m_w, m_s
    and no_cols in place.
    """
    from ..engine import Engine
    from ..integration import _set_kernel
    nx, nz = int(no_cells[0]), int(no_cells[1])
    N = len(xis)
    cells = np.asarray(cells, dtype=np.int64)
    active = (cells[0] >= 0) & (cells[0] < nx) & (cells[1] >= 0) & (cells[1] < nz)
    eng = Engine(nx, nz, [N], [[0.0, float(nx)], [0.0, float(nz)]], [1.0, 1.0], 1.0)
    eng.set_params(solute=solute_type, act_collisions=True, act_relaxation=False, seed=rng.get_seed())
    eng.e.dt_col_over_dV = float(dt_over_dV)
    _set_kernel(eng, E_col_grid, R_kernel_low_log, bin_factor_R_log)
    # only the temperature matters to the collision step; the other fields keep the per-cell table finite
    scal = np.ones((10, nx, nz))
    scal[0] = grid_temperature
    scal[1] = 1.0E5
    eng.set_grid(0, scal, np.ones((7, nx, nz)))
    eng.init_cells()
    zeros = np.zeros((2, N))
    parked = np.where(active, cells, np.clip(cells, 0, [[nx - 1], [nz - 1]]))
    eng.load_particles([(zeros, vel, parked, m_w, m_s, xis, active)])
    eng.set_counters(0, no_cols)
    if rnd is not None:
        counts = np.bincount(cells[0, active] * nz + cells[1, active], minlength=nx * nz)
        need = int((counts * (counts - 1) // 2).sum())
        if np.asarray(rnd).size < need:
            raise ValueError(f"rnd holds {np.asarray(rnd).size} uniforms, the step needs {need}")
        eng.set_rnd_stream(rnd)
    eng.col_step = rng.take_steps(1)
    eng.collision_step()
    _, _, _, mw, ms, x, _ = eng.export_particles(0)
    no_cols[...] = eng.counters(0)[0]
    xis[...] = x
    m_w[...] = mw
    m_s[...] = ms


collision_step_Ecol_grid_R_all_cells_2D_multicomp = collision_step_Ecol_grid_R_all_cells_2D_multicomp_np
