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


def test_fast_kernel_index_equals_exact(gpu):
    """The collision kernels' fast kernel-table index (the bin of a radius from the hardware float log2 where the
    bin is certain) equals the exact double evaluation: the hardware log2 errs by no more than the bound the fast
    path assumes on every float of its range, and the two indices agree on random radii, on radii at and within a
    few ulps and a few margins of every bin edge of both tables, and outside the fast path's range."""
    import torch
    from lacmopy_amd import load_kernel_data_Ecol
    from lacmopy_amd._native import check, lib
    s = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    sweep = torch.zeros(4, dtype=torch.float64, device="cuda")
    rng = np.random.default_rng(17)
    for name in ["Long_Bott", "Hall_Bott"]:
        _, lnR0, lnf, nb = load_kernel_data_Ecol(name)
        parts = [np.exp(rng.uniform(np.log(1e-6), np.log(1e6), 2_000_000))]
        edges = np.exp(lnR0 + (np.arange(-3, nb + 3) - 0.5) * lnf)
        for d in range(-8, 9):  # the edges and their neighbours ulp by ulp
            parts.append(edges + d * np.spacing(edges))
        for rel in [1e-9, 1e-7, 1e-6, 3e-6, 1e-5, 3e-5, 1e-4]:  # around the fast path's margin (~2.5e-4 * lnf)
            parts += [edges * (1 + rel), edges * (1 - rel)]
        lim = np.array([2.0 ** -14, 2.0 ** 16])
        parts += [lim, np.nextafter(lim, 0), np.nextafter(lim, np.inf),
                  np.array([0.0, 5e-324, 1e-300, 1e300, np.inf, 1.0, np.exp(lnR0)])]
        R = torch.as_tensor(np.concatenate(parts), device="cuda")
        fast = torch.empty(R.numel(), dtype=torch.int32, device="cuda")
        exact = torch.empty_like(fast)
        check(lib().lm_kernel_index_check(R.numel(), C.c_void_p(R.data_ptr()), lnR0, lnf, nb,
                                          C.c_void_p(fast.data_ptr()), C.c_void_p(exact.data_ptr()),
                                          C.c_void_p(sweep.data_ptr()), s))
        torch.cuda.synchronize()
        f, e = fast.cpu().numpy(), exact.cpu().numpy()
        bad = np.flatnonzero(f != e)
        assert bad.size == 0, (name, R.cpu().numpy()[bad[:5]], f[bad[:5]], e[bad[:5]])
        assert e.min() == 0 and e.max() == nb - 1
        max_err, bound, margin, swept = sweep.cpu().numpy()
        assert swept == 30 * 2 ** 23
        assert 0.0 < max_err <= bound, (max_err, bound)
        assert 0.0 < margin < 1e-3, margin
        print(f"{name}: hardware log2 error {max_err:.3g} (bound {bound:.3g}), margin in t {margin:.3g}")


def test_shared_reciprocal_quotients_equal_ieee_division(gpu):
    """The particle kernel's quotients through one refined reciprocal per denominator (the division sequence
    without its scale and fix-up steps, taken for operands of exponent within +-300, else the full division)
    equal IEEE a / b bit for bit: dense in the shortcut's range, across the whole exponent range (both paths),
    and at zeros of either sign, subnormals, infinities and NaNs."""
    rng = np.random.default_rng(23)
    n = 2_000_000
    sign = rng.choice([-1.0, 1.0], (2, n))
    inside = sign * np.ldexp(rng.uniform(1.0, 2.0, (2, n)), rng.integers(-300, 300, (2, n)))
    anywhere = sign * np.ldexp(rng.uniform(1.0, 2.0, (2, n)), rng.integers(-1074, 1024, (2, n)))
    mantissa_edges = sign[:, :8] * np.array([1.0, np.nextafter(2.0, 0), 1.5, np.nextafter(1.0, 2), 1.25, 1.75,
                                            np.nextafter(1.5, 0), np.nextafter(1.5, 2)])
    special = np.array([0.0, -0.0, 5e-324, -5e-324, 2.2250738585072014e-308, 1.7976931348623157e308, np.inf,
                        -np.inf, np.nan, 1.0, -1.0, 2.0 ** -300, 2.0 ** 300, np.nextafter(2.0 ** -300, 0), 3.0])
    sa, sb = np.meshgrid(special, special)
    a = np.concatenate([inside[0], anywhere[0], mantissa_edges[0], sa.ravel()])
    b = np.concatenate([inside[1], anywhere[1], mantissa_edges[1], sb.ravel()])
    with np.errstate(all="ignore"):
        want = a / b
    plain = _dev(5, a, b)  # the device's own division

    def check(got, ref, what):
        same = (got.view(np.int64) == ref.view(np.int64)) | (np.isnan(got) & np.isnan(ref))
        bad = np.flatnonzero(~same)
        assert bad.size == 0, (what, a[bad[:4]], b[bad[:4]], got[bad[:4]], ref[bad[:4]])

    check(plain, want, "device division against IEEE")
    for which in (8, 9):
        check(_dev(which, a, b), plain, f"probe {which} against the device division")
