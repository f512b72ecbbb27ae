"""C ABI and host logic on the CPU: the library loads, exports every entry
point declared in include/lacmo_amd.h, the ctypes mirror matches the C struct,
the kernels' constants equal the oracle's, and the host-side layout is the
reference's (cell, id) order with an exact inverse."""
import ctypes as C
import os
import re

import numpy as np

from lacmopy_amd import _native as N
from lacmopy_amd.engine import host_layout, host_unlayout
from lacmopy_amd.setup_case import make_case
from oracle import oracle as orc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    L = N.lib()
    hdr = open(os.path.join(ROOT, "include", "lacmo_amd.h")).read()
    declared = set(re.findall(r"\b(lm_[a-z0-9_]+)\s*\(", hdr))
    assert declared == set(N.EXPORTED), declared ^ set(N.EXPORTED)
    for name in declared:
        assert hasattr(L, name), name
    assert L.lm_version().startswith(b"lacmopy_amd")


def test_struct_layout_matches():
    L = N.lib()
    assert L.lm_engine_sizeof() == C.sizeof(N.LmEngine)
    off = (C.c_int64 * len(N.OFFSET_FIELDS))()
    L.lm_engine_offsets(off, len(N.OFFSET_FIELDS))
    for i, f in enumerate(N.OFFSET_FIELDS):
        assert off[i] == getattr(N.LmEngine, f).offset, f
    from lacmopy_amd.engine import WORK_BINS
    hdr = open(os.path.join(ROOT, "include", "lacmo_amd.h")).read()
    assert int(re.search(r"#define LM_WORK_BINS (\d+)", hdr).group(1)) == WORK_BINS


def test_kernel_constants_equal_oracle_and_reference_expressions():
    L = N.lib()
    n = L.lm_num_constants()
    a = (C.c_double * n)()
    L.lm_get_constants(a, n)
    k = np.array(a[:])
    o = orc.consts()
    # product order: pi, 1/3, 0.75/pi, R_v, R_d, c_pd, eps, kappa, kf, kf2, hf, ca1, ca2, cb1, cb2, T0, wsmax,
    #                psig_AS, psig_NaCl, 4+3+5+4+5+4+6+3 polynomial coefficients
    # oracle order:  pi, 1/3, 4pi/3, 0.75/pi, R_v, R_d, c_pd, c_pv, eps, kappa, kf, kf2, hf, ca1, ..., same tail
    assert k[0] == o[0] and k[1] == o[1] and k[2] == o[3]
    np.testing.assert_array_equal(k[3:6], o[4:7])
    np.testing.assert_array_equal(k[6:], o[8:8 + n - 6])
    # values the reference's Python module-level expressions give
    assert k[7] == 287.06 / 1004.71
    assert k[6] == 287.06 / 461.53
    assert k[17] == 2.362E-3 / 132.14E-3 / 0.072
    assert k[16] == 1. / 0.78


def test_host_layout_order_and_inverse():
    grid, pos, vel, cells, m_w, m_s, xi, act = make_case(nx=5, nz=4, spc=6, seed=2)
    act = act.copy()
    act[[3, 17, 40]] = False
    cells = cells.copy()
    cells[1, ~act] = -1
    runs = [(pos, vel, cells, m_w, m_s, xi, act)] * 2
    Np = [len(xi)] * 2
    host, hid, hcell, seg = host_layout(runs, 5, 4, Np)
    nc = 20
    for r in range(2):
        s = seg[r * (nc + 1):(r + 1) * (nc + 1) + 1]
        for c in range(nc):
            ids = hid[s[c]:s[c + 1]]
            assert np.all(np.diff(ids) > 0)                        # ascending id in a cell
            assert np.all(hcell[s[c]:s[c + 1]] == r * nc + c)
            assert np.all(cells[0, ids] * 4 + cells[1, ids] == c)
        gy = hid[s[nc]:s[nc + 1]]
        assert sorted(gy) == [3, 17, 40]
        assert np.all(hcell[s[nc]:s[nc + 1]] == -(2 + cells[0, gy]))
        b = r * len(xi)
        g = {f: host[f][b:b + len(xi)] for f in host}
        back = host_unlayout(g, hid[b:b + len(xi)], hcell[b:b + len(xi)], r * nc, 4)
        for x, y in zip(back, (pos, vel, cells, m_w, m_s, xi, act)):
            assert np.array_equal(x, y)
    assert seg[-1] == 2 * len(xi)


def _checked_engine(**over):
    """An LmEngine that passes every shape check, with placeholder (never dereferenced) buffer pointers."""
    e = N.LmEngine()
    e.n_runs, e.nx, e.nz, e.nc, e.n_tot = 1, 4, 4, 16, 10
    fake = 0x1000
    for f in ("x", "z", "u", "w", "m_w", "m_s", "xi", "id", "cell", "seg"):
        getattr(e, f)[0] = getattr(e, f)[1] = fake
    for f in ("cc", "err", "away", "col_step_dev", "rem_slot", "rem_cnt"):
        setattr(e, f, fake)
    e.rem_cap = 1024
    for k, v in over.items():
        setattr(e, k, v)
    return e


def test_engine_check_rejects_bad_removal_capacity():
    """rem_cap sizes rem_slot per run and bounds what k_water reads: outside 1..LM_REM_CAP_MAX, or without
    its buffers, every entry point refuses the engine before any device work."""
    L = N.lib()
    hdr = open(os.path.join(ROOT, "include", "lacmo_amd.h")).read()
    cap_max = int(re.search(r"#define LM_REM_CAP_MAX (\d+)", hdr).group(1))
    from lacmopy_amd.engine import REM_CAP_MAX
    assert cap_max == REM_CAP_MAX == 1024
    for bad in (dict(rem_cap=0), dict(rem_cap=-5), dict(rem_cap=cap_max + 1), dict(rem_slot=None)):
        e = _checked_engine(**bad)
        rc = L.lm_reindex(C.byref(e), None)
        assert rc == 1001, (bad, rc)
        assert b"rem_cap" in L.lm_last_error_message()
    import pytest
    from lacmopy_amd.engine import Engine
    with pytest.raises(ValueError, match="rem_cap"):
        Engine(4, 4, [10], [[0, 80], [0, 80]], [20, 20], 400.0, rem_cap=0)
