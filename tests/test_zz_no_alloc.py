"""ppk_mhd_compute_dt_noalloc (the CFL dt of a caller-owned device state, no device allocation)
against ppk_mhd_compute_dt on the same state: the same dt bit for bit and the same ghost fill,
on two calls in a row, the second on a smaller grid.  The test's own device buffers come from
the HIP runtime the library is linked with (plumbing only); the entry point is called through
the C ABI on the null stream.
"""
import ctypes as C

import numpy as np
import pytest

from ppkmhd_amd import _abi as A
from ppkmhd_amd import default_problem, make_params

pytestmark = pytest.mark.gpu


def blast(n, precision):
    p = make_params(3, n, n, n, gamma0=5 / 3, cfl=0.8, slope_type=1.0, tEnd=1.0,
                    precision=precision, kernel_path=A.PATH_FUSED_RECORDS)
    pr = default_problem(A.PROBLEM_BLAST, p)
    pr.blast_radius = 0.2
    pr.blast_density_out = 1.0
    pr.blast_bx = pr.blast_by = pr.blast_bz = 1.0 / np.sqrt(3.0)
    return p, pr


def hip():
    """The HIP runtime already loaded with the library (same soname)."""
    from ppkmhd_amd._lib import lib

    lib()
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return h


def noalloc_dt(p, U):
    """(dt, state after the call) from ppk_mhd_compute_dt_noalloc on a device copy of U."""
    from ppkmhd_amd._lib import check, lib

    H = hip()
    U = np.ascontiguousarray(U, dtype=np.float64)
    dU, ddt = C.c_void_p(), C.c_void_p()
    assert H.hipMalloc(C.byref(dU), U.nbytes) == 0
    assert H.hipMalloc(C.byref(ddt), 8) == 0
    try:
        assert H.hipMemcpy(dU, U.ctypes.data, U.nbytes, 1) == 0  # host to device
        check(lib().ppk_mhd_compute_dt_noalloc(C.byref(p), dU, ddt, None))
        assert H.hipDeviceSynchronize() == 0
        out = np.empty_like(U)
        dt = np.empty(1)
        assert H.hipMemcpy(out.ctypes.data, dU, U.nbytes, 2) == 0  # device to host
        assert H.hipMemcpy(dt.ctypes.data, ddt, 8, 2) == 0
    finally:
        H.hipFree(dU)
        H.hipFree(ddt)
    return float(dt[0]), out


@pytest.mark.parametrize("precision", [A.FP64_STRICT, A.FP64])
def test_noalloc_dt_matches_compute_dt(gpu, precision):
    from ppkmhd_amd.solver import SolverMHDMuscl

    for n, steps in ((40, 3), (24, 2)):  # the second call on a smaller grid
        p, pr = blast(n, precision)
        with SolverMHDMuscl(p, pr) as g:
            g.run(steps)
            U = g.download()  # ghosts as the last step left them
            ref = g.compute_dt_local()  # fills the ghosts of the handle's state
            Uref = g.download()
        got, Ugot = noalloc_dt(p, U)
        print(f"{n}^3 after {steps} steps, precision {precision}: dt {got!r} vs {ref!r}",
              flush=True)
        assert got == ref
        assert np.array_equal(Ugot, Uref)


def test_noalloc_refuses_oversized_grid(gpu):
    from ppkmhd_amd._lib import lib

    p = make_params(3, 700, 700, 700)
    rc = lib().ppk_mhd_compute_dt_noalloc(C.byref(p), C.c_void_p(8), C.c_void_p(8), None)
    assert rc == A.PPK_ERR_INVALID_ARGUMENT
    assert b"too large" in lib().ppk_mhd_last_error(None)
