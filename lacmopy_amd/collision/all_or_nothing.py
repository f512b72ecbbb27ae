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
    if isinstance(rnd, torch.Tensor):
        r = rnd.reshape(-1).to(device="cuda", dtype=torch.float64).contiguous()
    else:
        r = _dev(torch, np.ascontiguousarray(rnd, dtype=np.float64).reshape(-1))
    if r.numel() < n * (n - 1) // 2:
        raise ValueError(f"rnd holds {r.numel()} uniforms, the step needs {n * (n - 1) // 2}")
    return r


def _finish(torch, L, no_cols, nc_dev, pairs=()):
    """Error check of the box kernels, the collision counters added to no_cols, and the device results copied
    back into the caller's NumPy arrays (CUDA tensors were updated where they are)."""
    check(L.lm_box_check(_stream(torch)))
    if isinstance(no_cols, torch.Tensor):
        no_cols += nc_dev.reshape(no_cols.shape).to(no_cols.device)
    else:
        no_cols[...] = no_cols + nc_dev.cpu().numpy().reshape(2)
    for host, dev in pairs:
        if not isinstance(host, torch.Tensor):
            host[...] = dev.cpu().numpy()


def collision_step_Golovin_np(xis, masses, dt_over_dV, no_cols, *, rnd=None):
    """all_or_nothing.py:109-185: Golovin kernel K = b (m_i + m_j), box model.  NumPy arrays (copied to the GPU
    and back) or CUDA float64 tensors (updated where they are)."""
    torch = _torch()
    L = lib()
    n = int(xis.shape[0])
    f64 = torch.float64
    x, m = _as_device(torch, xis, f64, "xis", (n,)), _as_device(torch, masses, f64, "masses", (n,))
    off = _dev(torch, [0, n], np.int32)
    nc = torch.zeros(2, dtype=torch.int64, device="cuda")
    r = _rnd(torch, rnd, n)
    check(L.lm_golovin_step(1, _p(off), _p(x), _p(m), float(dt_over_dV), _p(nc), rng.get_seed(), rng.take_steps(1),
                            _p(r), _stream(torch)))
    _finish(torch, L, no_cols, nc, ((xis, x), (masses, m)))


collision_step_Golovin = collision_step_Golovin_np


def collision_step_Ecol_grid_R_np(xis, masses, radii, vel, mass_densities, dt_over_dV, E_col_grid, no_kernel_bins,
                                  R_kernel_low_log, bin_factor_R_log, no_cols, *, rnd=None):
    """all_or_nothing.py:269-390: hydrodynamic kernel with tabulated E_col(R1, R2), box model, |v| 1-D.  NumPy
    arrays or CUDA float64 tensors (updated where they are)."""
    torch = _torch()
    L = lib()
    n = int(xis.shape[0])
    x, m, R, v, rho = (_as_device(torch, a, torch.float64, nm, (n,)) for a, nm in
                       ((xis, "xis"), (masses, "masses"), (radii, "radii"), (vel, "vel"),
                        (mass_densities, "mass_densities")))
    E, _ = _device_table(torch, E_col_grid)
    off = _dev(torch, [0, n], np.int32)
    nc = torch.zeros(2, dtype=torch.int64, device="cuda")
    r = _rnd(torch, rnd, n)
    check(L.lm_ecol_box_step(1, _p(off), _p(x), _p(m), _p(R), _p(v), _p(rho), float(dt_over_dV), _p(E),
                             int(no_kernel_bins), float(R_kernel_low_log), float(bin_factor_R_log), _p(nc),
                             rng.get_seed(), rng.take_steps(1), _p(r), 0, _stream(torch)))
    _finish(torch, L, no_cols, nc, ((xis, x), (masses, m), (radii, R)))


collision_step_Ecol_grid_R = collision_step_Ecol_grid_R_np


def _box2d(xis, m_w, m_s, radii, vel, mass_densities, dt_over_dV, E_col_grid, E_const, nb, lnR0, lnf, no_cols, rnd):
    torch = _torch()
    L = lib()
    n = int(xis.shape[0])
    f64 = torch.float64
    x, mw, ms, R, rho = (_as_device(torch, a, f64, nm, (n,)) for a, nm in
                         ((xis, "xis"), (m_w, "m_w"), (m_s, "m_s"), (radii, "radii"),
                          (mass_densities, "mass_densities")))
    v = _as_device(torch, vel, f64, "vel", (2, n))
    vx, vz = v[0], v[1]  # contiguous rows of the (2, n) array
    E = _device_table(torch, E_col_grid)[0] if E_col_grid is not None else None
    off = _dev(torch, [0, n], np.int32)
    nc = torch.zeros(2, dtype=torch.int64, device="cuda")
    r = _rnd(torch, rnd, n)
    check(L.lm_box2d_step(1, _p(off), _p(x), _p(mw), _p(ms), _p(R), _p(vx), _p(vz), _p(rho), float(dt_over_dV),
                          _p(E), float(E_const), int(nb), float(lnR0), float(lnf), _p(nc), rng.get_seed(),
                          rng.take_steps(1), _p(r), _stream(torch)))
    _finish(torch, L, no_cols, nc, ((xis, x), (m_w, mw), (m_s, ms), (radii, R)))


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


_tables = {}


def _device_table(torch, E):
    """Device copy and largest entry of an E_col table, kept while the caller's array is alive and unchanged."""
    if isinstance(E, torch.Tensor):
        t = E.to(device="cuda", dtype=torch.float64).contiguous()
        return t, float(t.max()) if bool(torch.isfinite(t).all()) and float(t.min()) >= 0.0 else 0.0
    a = np.ascontiguousarray(E, dtype=np.float64)
    key = (a.__array_interface__["data"][0], a.shape, hash(a[::max(1, a.shape[0] // 8)].tobytes()))
    hit = _tables.get(key)
    if hit is None or not np.array_equal(hit[2], a):
        t = torch.from_numpy(a).cuda()
        emax = float(a.max()) if np.isfinite(a).all() and a.min() >= 0.0 else 0.0
        _tables.clear()
        hit = _tables[key] = (t, emax, a.copy())
    return hit[0], hit[1]


def _as_device(torch, a, dtype, name, shape=None):
    """A caller's CUDA tensor as it is (zero-copy; it is updated in place), or a device copy of an array."""
    if isinstance(a, torch.Tensor):
        if a.device.type != "cuda" or a.dtype != dtype or not a.is_contiguous():
            raise TypeError(f"{name}: a CUDA tensor must be contiguous {dtype} (got {a.device}, {a.dtype})")
        t = a
    else:
        t = torch.from_numpy(np.ascontiguousarray(a, dtype={torch.float64: np.float64, torch.int64: np.int64}[dtype]))
        t = t.cuda()
    if shape is not None and tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name}: shape {tuple(t.shape)}, expected {tuple(shape)}")
    return t


def collision_step_Ecol_grid_R_all_cells_2D_multicomp_np(xis, m_w, m_s, vel, grid_temperature, cells, no_cells,
                                                         dt_over_dV, E_col_grid, no_kernel_bins, R_kernel_low_log,
                                                         bin_factor_R_log, no_cols, solute_type, *, rnd=None):
    """all_or_nothing.py:754-846: one collision step in every cell of the 2-D grid.

    Members of cell (i, j) are the SIPs with cells == (i, j) in ascending
    index; radii and densities from m_w, m_s at the cell temperature; a float
    E_col_grid selects the constant-efficiency kernel.  Updates xis, m_w, m_s
    and no_cols in place.  The arrays may be NumPy arrays (copied to the GPU
    and back) or CUDA tensors of the reference's shapes and dtypes (float64;
    cells and no_cols int64), which the step reads and updates where they are
    (lm_collision_all_cells: no copy, no per-call engine).
    """
    torch = _torch()
    L = lib()
    nx, nz = int(no_cells[0]), int(no_cells[1])
    N = int(xis.shape[0])
    f64, i64 = torch.float64, torch.int64
    x = _as_device(torch, xis, f64, "xis", (N,))
    mw = _as_device(torch, m_w, f64, "m_w", (N,))
    ms = _as_device(torch, m_s, f64, "m_s", (N,))
    v = _as_device(torch, vel, f64, "vel", (2, N))
    c = _as_device(torch, cells, i64, "cells", (2, N))
    T = _as_device(torch, grid_temperature, f64, "grid_temperature", (nx, nz))
    nc = _as_device(torch, no_cols, i64, "no_cols", (2,))
    if isinstance(E_col_grid, float):
        E, emax, E_const, nb, l0, lf = None, 0.0, float(E_col_grid), 0, 0.0, 1.0
    else:
        E, emax = _device_table(torch, E_col_grid)
        E_const, nb, l0, lf = 0.0, int(no_kernel_bins), float(R_kernel_low_log), float(bin_factor_R_log)
        if tuple(E.shape) != (nb, nb):
            raise ValueError(f"E_col_grid has shape {tuple(E.shape)}, no_kernel_bins = {nb}")
    r = None
    if rnd is not None:
        r = rnd if isinstance(rnd, torch.Tensor) and rnd.device.type == "cuda" else _dev(torch, np.asarray(rnd))
        r = r.reshape(-1).to(f64).contiguous()
    st = _stream(torch)
    check(L.lm_collision_all_cells(N, _p(x), _p(mw), _p(ms), _p(v), _p(T), _p(c), nx, nz, float(dt_over_dV), _p(E), nb,
                                   l0, lf, E_const, emax, _p(nc), 0 if solute_type == "AS" else 1, rng.get_seed(),
                                   rng.take_steps(1), _p(r), int(r.numel()) if r is not None else 0, st))
    check(L.lm_collision_all_cells_check(0, st))
    for host, dev in ((xis, x), (m_w, mw), (m_s, ms), (no_cols, nc)):
        if not isinstance(host, torch.Tensor):
            host[...] = dev.cpu().numpy()


collision_step_Ecol_grid_R_all_cells_2D_multicomp = collision_step_Ecol_grid_R_all_cells_2D_multicomp_np
