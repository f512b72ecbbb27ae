"""GPU: the advection step captured in a HIP graph.  Every replay, with eager work between replays, is
compared bit for bit with the same steps run eagerly (which the parity tests compare with the oracle)."""
import numpy as np
import pytest

from lacmopy_amd import load_kernel_data_Ecol
from test_gpu_parity import _assert_same, _case, _config, _engine, _prep

pytestmark = pytest.mark.gpu

NAMES = ["pos", "vel", "cells", "m_w", "m_s", "xi", "active"]


def _same_state(a, b, what):
    for nm, x, y in zip(NAMES, a.export_particles(0), b.export_particles(0)):
        _assert_same(x, y, f"{what}: {nm}")
    sa, ma = a.export_grid(0)
    sb, mb = b.export_grid(0)
    _assert_same(sa, sb, f"{what}: grid_scalar_fields")
    _assert_same(ma, mb, f"{what}: grid_mat_prop")
    for x, y in zip(a.counters(0), b.counters(0)):
        _assert_same(x, y, f"{what}: counters")
    assert a.col_step == b.col_step


@pytest.mark.parametrize("kernel", ["Long_Bott", "const"])
def test_graph_replays_match_eager_steps(gpu, kernel):
    if kernel == "const":
        E, lnR0, lnf, nb = 0.5, None, None, None
    else:
        E, lnR0, lnf, nb = load_kernel_data_Ecol(kernel)
    cfg = _config(E, lnR0, lnf, nb)
    grid, *parts = _case(nx=8, nz=8, spc=24, seed=31)
    parts = _prep(grid, parts)
    eager = _engine(grid, parts, cfg, E, lnR0, lnf, 21)
    graph = _engine(grid, parts, cfg, E, lnR0, lnf, 21)
    h = graph.graph_capture(2)
    for rep in range(3):
        eager.adv_steps(2)
        graph.graph_launch(h)
        _same_state(eager, graph, f"replay {rep}")
    # an odd eager step swaps the buffers: the captured graph refuses to run, the cache captures the other parity
    eager.adv_steps(1)
    graph.adv_steps(1)
    with pytest.raises(Exception):
        graph.graph_launch(h)
    eager.adv_steps(5)
    graph.adv_steps_graph(5)
    _same_state(eager, graph, "mixed eager and graph steps")
    assert eager.counters(0)[0].sum() > 0
    graph.graph_destroy(h)
    graph.drop_graph()
