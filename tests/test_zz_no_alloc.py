"""lm_collision_all_cells_noalloc (fixed device scratch, no allocation) against lm_collision_all_cells.

Both entry points run through the C-ABI on copies of the same device arrays with the same seed and step;
the results must agree bit for bit.  Two calls in a row, the second on a smaller input, so the second call
works on the scratch the first one left; sizes beyond the fixed scratch are refused before any launch.
"""
import ctypes as C

import numpy as np
import pytest

from lacmopy_amd import load_kernel_data_Ecol
from lacmopy_amd._native import LacmoError, check, lib
from test_gpu_parity import _case, _prep

pytestmark = pytest.mark.gpu

MAX_SIPS, MAX_CELLS = 1 << 20, 1 << 15


def _p(t):
    return C.c_void_p(t.data_ptr()) if t is not None else C.c_void_p(0)


def _inputs(nx, nz, spc, seed, crowd):
    grid, *parts = _case(nx=nx, nz=nz, spc=spc, seed=seed)
    pos, vel, cells, m_w, m_s, xi, active = _prep(grid, parts)
    N = len(xi)
    cells = cells.copy()
    if crowd:
        cells[:, :crowd] = np.array([[1], [2]])  # one cell over 256 members: the large-cell pass
    cells[1, ::19] = -1
    vel = vel.copy()
    vel[1] -= np.random.default_rng(seed).uniform(0.0, 3.0, N)
    return dict(xi=xi, m_w=m_w, m_s=m_s, vel=vel, cells=cells, T=grid.temperature.copy(),
                dt_over_dV=0.5 / grid.volume_cell, nx=nx, nz=nz)


def _run(entry, d, E, nb, lnR0, lnf, emax, seed, step):
    import torch
    L = lib()
    dev = {k: torch.from_numpy(np.ascontiguousarray(d[k])).cuda() for k in ("xi", "m_w", "m_s", "vel", "cells", "T")}
    nc = torch.zeros(2, dtype=torch.int64, device="cuda")
    st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    N = int(dev["xi"].numel())
    check(getattr(L, entry)(N, _p(dev["xi"]), _p(dev["m_w"]), _p(dev["m_s"]), _p(dev["vel"]), _p(dev["T"]),
                            _p(dev["cells"]), d["nx"], d["nz"], d["dt_over_dV"], _p(E), nb, lnR0, lnf, 0.0, emax,
                            _p(nc), 0, seed, step, None, 0, st))
    check(L.lm_collision_all_cells_check(1 if entry.endswith("noalloc") else 0, st))
    torch.cuda.synchronize()
    return {k: dev[k].cpu().numpy() for k in ("xi", "m_w", "m_s")} | {"no_cols": nc.cpu().numpy()}


def test_noalloc_matches_allocating_entry_twice(gpu):
    import torch
    Eh, lnR0, lnf, nb = load_kernel_data_Ecol("Long_Bott")
    E = torch.from_numpy(np.ascontiguousarray(Eh)).cuda()
    emax = float(Eh.max())
    for i, (nx, nz, spc, crowd) in enumerate(((16, 24, 60, 600), (6, 12, 40, 0))):
        d = _inputs(nx, nz, spc, 31 + i, crowd)
        print(f"noalloc call {i}: {len(d['xi'])} SIPs on {nx}x{nz} cells", flush=True)
        ref = _run("lm_collision_all_cells", d, E, nb, lnR0, lnf, emax, 11, 5 + i)
        got = _run("lm_collision_all_cells_noalloc", d, E, nb, lnR0, lnf, emax, 11, 5 + i)
        for k in ref:
            np.testing.assert_array_equal(got[k].view(np.int64), ref[k].view(np.int64), err_msg=f"call {i}: {k}")
        assert ref["no_cols"].sum() > 0


def test_noalloc_rejects_sizes_beyond_its_scratch(gpu):
    import torch
    L = lib()
    n = MAX_SIPS + 1
    f = torch.zeros(n, dtype=torch.float64, device="cuda")
    v = torch.zeros(2, n, dtype=torch.float64, device="cuda")
    c = torch.zeros(2, n, dtype=torch.int64, device="cuda")
    T = torch.full((4, 4), 280.0, dtype=torch.float64, device="cuda")
    nc = torch.zeros(2, dtype=torch.int64, device="cuda")
    with pytest.raises(LacmoError, match="fixed scratch"):
        check(L.lm_collision_all_cells_noalloc(n, _p(f), _p(f), _p(f), _p(v), _p(T), _p(c), 4, 4, 1.0, None, 0, 0.0,
                                               1.0, 0.5, 0.0, _p(nc), 0, 1, 0, None, 0, None))
    T2 = torch.full((MAX_CELLS + 1, 1), 280.0, dtype=torch.float64, device="cuda")
    with pytest.raises(LacmoError, match="fixed scratch"):
        check(L.lm_collision_all_cells_noalloc(16, _p(f), _p(f), _p(f), _p(v), _p(T2), _p(c), MAX_CELLS + 1, 1, 1.0,
                                               None, 0, 0.0, 1.0, 0.5, 0.0, _p(nc), 0, 1, 0, None, 0, None))
