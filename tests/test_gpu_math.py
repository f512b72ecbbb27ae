"""Device FP64 math (lm_math on gfx950, sqrt, division, floor, Python modulo) is
bit-identical to the host build used by the oracle."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _dev(which, x, y):
    import torch
    from lacmopy_amd._native import check, lib
    xt = torch.as_tensor(x, device="cuda")
    yt = torch.as_tensor(y, device="cuda")
    out = torch.empty_like(xt)
    s = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    check(lib().lm_math_probe(which, x.size, C.c_void_p(xt.data_ptr()), C.c_void_p(yt.data_ptr()),
                              C.c_void_p(out.data_ptr()), s))
    torch.cuda.synchronize()
    return out.cpu().numpy()


def _inputs(seed=1, n=400_000):
    rng = np.random.default_rng(seed)
    x = np.exp(rng.uniform(-40, 40, n))
    y = rng.choice([1.0 / 3.0, 1.256, 1.5, 1.94, 0.687, -0.4092, 3.46, 2.0, 0.5], n)
    return x, y


@pytest.mark.parametrize("which", [0, 1, 2, 7])
def test_lm_math_device_equals_host(gpu, which):
    x, y = _inputs(which + 1)
    if which == 0:
        x = np.random.default_rng(9).uniform(-700, 700, x.size)
    host = orc.lm_vec(which, x, y)
    dev = _dev(which, x, y)
    assert np.array_equal(host.view(np.int64), dev.view(np.int64))


def test_ieee_ops_device_equals_host(gpu):
    rng = np.random.default_rng(4)
    a = rng.uniform(-3000, 3000, 400_000) * np.exp(rng.uniform(-20, 20, 400_000))
    b = rng.uniform(0.5, 2000, 400_000)
    assert np.array_equal(_dev(4, np.abs(a), b), np.sqrt(np.abs(a)))
    assert np.array_equal(_dev(5, a, b), a / b)
    assert np.array_equal(_dev(6, a, b), np.floor(a / b))
    x = rng.uniform(-10, 3010, 400_000)
    L = np.full_like(x, 1500.0)
    assert np.array_equal(_dev(3, x, L), np.mod(x, L))
    assert np.array_equal(_dev(3, x, L), orc.lm_vec(3, x, L))
