"""TEST INFRASTRUCTURE -- ctypes wrapper of the CPU oracle (oracle/liboracle.so).

The oracle is a C++ restatement of the reference's implementation-0 MHD step; only tests/,
``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg may use it, and only as the
checker.  Arrays are numpy float64 in the reference layout (8, ksize, jsize, isize).
"""
import ctypes as C
import os
import subprocess

import numpy as np

from ppkmhd_amd import _abi as A

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBS = {}


def build(force=False):
    """Compile oracle/liboracle.so and oracle/liboracle_count.so with the committed Makefile."""
    out = os.path.join(_HERE, "liboracle.so")
    if force or not os.path.exists(out):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "all"], check=True)
    return out


def build_native():
    """The CPU baseline's build (-march=native for the host it runs on); returns its path."""
    subprocess.run(["make", "-s", "-B", "-C", _HERE, "liboracle_native.so"], check=True)
    return os.path.join(_HERE, "liboracle_native.so")


def _lib(counting=False, native=False):
    name = "liboracle_count.so" if counting else ("liboracle_native.so" if native else "liboracle.so")
    if name in _LIBS:
        return _LIBS[name]
    path = os.path.join(_HERE, name)
    if native:  # always for the host at hand: a copy built elsewhere may not run here
        build_native()
    elif not os.path.exists(path):
        build()
    lib = C.CDLL(path)
    lib.ora_threads.restype = C.c_int
    P = C.POINTER(A.Params)
    PR = C.POINTER(A.Problem)
    dp = C.POINTER(C.c_double)
    lib.ora_params_size.restype = C.c_int
    lib.ora_problem_size.restype = C.c_int
    lib.ora_fill_boundaries.argtypes = [P, dp]
    lib.ora_primitives.argtypes = [P, dp, dp]
    lib.ora_inv_dt.argtypes = [P, dp]
    lib.ora_inv_dt.restype = C.c_double
    lib.ora_init_problem_array.argtypes = [P, PR, dp]
    lib.ora_riemann.argtypes = [P, dp, dp, dp]
    lib.ora_compute_emf.argtypes = [P, C.c_int, dp]
    lib.ora_compute_emf.restype = C.c_double
    lib.ora_mag_riemann2d_hlld.argtypes = [P, dp, dp]
    lib.ora_mag_riemann2d_hlld.restype = C.c_double
    lib.ora_find_speed_fast.argtypes = [P, dp, C.c_int]
    lib.ora_find_speed_fast.restype = C.c_double
    lib.ora_slope.argtypes = [P, C.c_double, C.c_double, C.c_double]
    lib.ora_slope.restype = C.c_double
    lib.ora_create.argtypes = [P]
    lib.ora_create.restype = C.c_void_p
    lib.ora_destroy.argtypes = [C.c_void_p]
    lib.ora_init_problem.argtypes = [C.c_void_p, PR]
    lib.ora_set_state.argtypes = [C.c_void_p, dp]
    lib.ora_get_state.argtypes = [C.c_void_p, dp]
    lib.ora_step.argtypes = [C.c_void_p, C.c_double, C.c_int]
    lib.ora_step.restype = C.c_double
    lib.ora_get_time.argtypes = [C.c_void_p, dp, dp, C.POINTER(C.c_longlong)]
    lib.ora_set_time.argtypes = [C.c_void_p, C.c_double, C.c_longlong]
    lib.ora_finished.argtypes = [C.c_void_p]
    lib.ora_finished.restype = C.c_int
    lib.ora_scratch.argtypes = [C.c_void_p, C.c_int]
    lib.ora_scratch.restype = dp
    lib.ora_op_counts.argtypes = [C.POINTER(C.c_longlong)]
    lib.ora_op_counts_reset.argtypes = []
    lib.ora_op_counts_stage.argtypes = [C.c_int, C.POINTER(C.c_longlong)]
    lib.ora_is_counting_build.restype = C.c_int
    _LIBS[name] = lib
    return lib


def _dp(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


def zeros(p):
    return np.zeros(p.shape, dtype=np.float64)


def fill_boundaries(p, U):
    _lib().ora_fill_boundaries(C.byref(p), _dp(U))
    return U


def primitives(p, U):
    Q = zeros(p)
    _lib().ora_primitives(C.byref(p), _dp(U), _dp(Q))
    return Q


def inv_dt(p, Q):
    return _lib().ora_inv_dt(C.byref(p), _dp(Q))


def init_problem_array(p, prob):
    U = zeros(p)
    _lib().ora_init_problem_array(C.byref(p), C.byref(prob), _dp(U))
    return U


def riemann(p, ql, qr):
    ql = np.ascontiguousarray(ql, dtype=np.float64)
    qr = np.ascontiguousarray(qr, dtype=np.float64)
    f = np.zeros(8)
    _lib().ora_riemann(C.byref(p), _dp(ql), _dp(qr), _dp(f))
    return f


def compute_emf(p, emf_dir, qedge):
    qedge = np.ascontiguousarray(qedge, dtype=np.float64).reshape(4, 8)
    return _lib().ora_compute_emf(C.byref(p), int(emf_dir), _dp(qedge))


def mag_riemann2d_hlld(p, qLLRR, eLLRR):
    q = np.ascontiguousarray(qLLRR, dtype=np.float64).reshape(4, 8)
    e = np.ascontiguousarray(eLLRR, dtype=np.float64).reshape(4)
    return _lib().ora_mag_riemann2d_hlld(C.byref(p), _dp(q), _dp(e))


def find_speed_fast(p, q, d):
    q = np.ascontiguousarray(q, dtype=np.float64)
    return _lib().ora_find_speed_fast(C.byref(p), _dp(q), int(d))


def slope(p, q, qp, qm):
    return _lib().ora_slope(C.byref(p), q, qp, qm)


class Solver:
    """SolverMHDMuscl<dim> time loop of the oracle (U / U2 ping-pong, t, dt, iteration)."""

    def __init__(self, p, counting=False, native=False):
        self.p = p
        self._lib = _lib(counting, native)
        self._h = self._lib.ora_create(C.byref(p))

    def close(self):
        if self._h:
            self._lib.ora_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def init_problem(self, prob):
        self._lib.ora_init_problem(self._h, C.byref(prob))

    def set_state(self, U):
        U = np.ascontiguousarray(U, dtype=np.float64)
        assert U.size == self.p.count
        self._lib.ora_set_state(self._h, _dp(U))

    def get_state(self):
        U = zeros(self.p)
        self._lib.ora_get_state(self._h, _dp(U))
        return U

    def step(self, dt=0.0, dt_mode=A.DT_FIXED):
        return self._lib.ora_step(self._h, float(dt), int(dt_mode))

    def time(self):
        t, dt, it = C.c_double(), C.c_double(), C.c_longlong()
        self._lib.ora_get_time(self._h, C.byref(t), C.byref(dt), C.byref(it))
        return t.value, dt.value, it.value

    def set_time(self, t, iteration):
        self._lib.ora_set_time(self._h, float(t), int(iteration))

    def finished(self):
        return bool(self._lib.ora_finished(self._h))

    def scratch(self, which, ncomp=8):
        ptr = self._lib.ora_scratch(self._h, int(which))
        if not ptr:
            return None
        n = self.p.ksize * self.p.jsize * self.p.isize
        a = np.ctypeslib.as_array(ptr, shape=(ncomp * n,)).copy()
        return a.reshape(ncomp, self.p.ksize, self.p.jsize, self.p.isize)


STAGES = ("boundaries", "primitives_dt", "elec_field", "mag_slopes", "trace", "fluxes", "emf",
          "update_ct")


def threads(native=False):
    """OpenMP threads of the oracle's parallel loops."""
    return int(_lib(native=native).ora_threads())


def op_counts(stage=None):
    """Operation counts accumulated by the instrumented build (all stages, or one by name)."""
    lib = _lib(counting=True)
    c = (C.c_longlong * 5)()
    lib.ora_op_counts_stage(-1 if stage is None else STAGES.index(stage), c)
    return dict(zip(("add", "mul", "div", "sqrt", "select"), list(c)))


def op_counts_reset():
    _lib(counting=True).ora_op_counts_reset()
