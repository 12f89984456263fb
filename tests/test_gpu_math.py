"""The default FP64 build's device reciprocal and square roots (v_rcp_f64 / v_rsq_f64 refined
in FP64, mhd_device.hpp), through ppk_mhd_math_selfcheck: over log-uniform inputs in
[2^-40, 2^40] (the range of densities, pressures and wave-speed differences) plus powers of
two and values next to them, 1/x and sqrt(x) equal the IEEE results bit for bit, and
1/sqrt(x) is within one unit in the last place of the correctly rounded value (80-bit
extended arithmetic on the host, then rounded)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def inputs():
    rng = np.random.default_rng(0)
    x = np.exp2(rng.uniform(-40.0, 40.0, 2_000_000))
    p2 = np.exp2(np.arange(-40, 41, dtype=np.float64))
    edge = np.concatenate([p2, np.nextafter(p2, 0.0), np.nextafter(p2, np.inf),
                           [1.0, 2.0, 3.0, 0.1, 1e-12, 1e12]])
    return np.ascontiguousarray(np.concatenate([x, edge]))


def selfcheck(L, x):
    out = np.empty(3 * x.size)
    dp = C.POINTER(C.c_double)
    rc = L.ppk_mhd_math_selfcheck(x.ctypes.data_as(dp), x.size, out.ctypes.data_as(dp))
    assert rc == 0, L.ppk_mhd_last_error(None)
    return out[0::3], out[1::3], out[2::3]


def test_reciprocal_and_sqrt_equal_ieee(gpu):
    x = inputs()
    rcp, sqrt, rsq = selfcheck(gpu, x)
    bad = rcp != 1.0 / x
    assert not bad.any(), f"{bad.sum()} reciprocals differ, e.g. x = {x[bad][:3]}"
    bad = sqrt != np.sqrt(x)
    assert not bad.any(), f"{bad.sum()} square roots differ, e.g. x = {x[bad][:3]}"
    ref = (1.0 / np.sqrt(x.astype(np.longdouble))).astype(np.float64)
    ulps = np.abs(rsq - ref) / np.spacing(ref)
    assert ulps.max() <= 1.0, f"inverse square root off by {ulps.max()} ulp"


def test_sqrt_of_zero(gpu):
    _, sqrt, _ = selfcheck(gpu, np.array([0.0, 4.0]))
    assert sqrt[0] == 0.0 and sqrt[1] == 2.0
