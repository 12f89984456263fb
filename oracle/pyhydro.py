"""TEST INFRASTRUCTURE -- ctypes wrapper of the CPU hydro oracle (oracle/libhydro_oracle.so).

The hydro oracle restates the reference's implementation-0 Euler MUSCL-Hancock step
(src/muscl/SolverHydroMuscl.cpp:257-405); only tests/ and ``__graft_entry__.smoke()`` may use it,
and only as the checker.  Arrays are numpy float64 in the reference layout (nbvar, ksize,
jsize, isize) with nbvar = 4 in 2D (ID, IP, IU, IV) and 5 in 3D (+ IW).
"""
import ctypes as C
import os
import subprocess

import numpy as np

from ppkmhd_amd import _abi as A

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBS = {}
_NAMES = ("libhydro_oracle.so", "libhydro_oracle_count.so")


def build(force=False):
    """Compile oracle/libhydro_oracle.so (and the operation-counting build) with the committed
    Makefile."""
    for name in _NAMES:
        if force or not os.path.exists(os.path.join(_HERE, name)):
            cmd = ["make", "-s", "-C", _HERE] + (["-B"] if force else []) + [name]
            subprocess.run(cmd, check=True)
    return os.path.join(_HERE, _NAMES[0])


def build_native():
    """The CPU baseline's build (-march=native for the host it runs on)."""
    subprocess.run(["make", "-s", "-B", "-C", _HERE, "libhydro_oracle_native.so"], check=True)
    return os.path.join(_HERE, "libhydro_oracle_native.so")


def _lib(counting=False, native=False):
    name = _NAMES[1] if counting else ("libhydro_oracle_native.so" if native else _NAMES[0])
    if name in _LIBS:
        return _LIBS[name]
    path = os.path.join(_HERE, name)
    if native:  # always for the host at hand: a copy built elsewhere may not run here
        build_native()
    elif not os.path.exists(path):
        build()
    lib = C.CDLL(path)
    lib.hora_threads.restype = C.c_int
    P = C.POINTER(A.Params)
    PR = C.POINTER(A.Problem)
    dp = C.POINTER(C.c_double)
    lib.hora_create.argtypes = [P]
    lib.hora_create.restype = C.c_void_p
    lib.hora_destroy.argtypes = [C.c_void_p]
    lib.hora_init_problem.argtypes = [C.c_void_p, PR]
    lib.hora_init_array.argtypes = [P, PR, dp]
    lib.hora_gravity_field.argtypes = [P, PR, dp]
    lib.hora_set_state.argtypes = [C.c_void_p, dp]
    lib.hora_get_state.argtypes = [C.c_void_p, dp]
    lib.hora_step.argtypes = [C.c_void_p, C.c_double, C.c_int]
    lib.hora_get_time.argtypes = [C.c_void_p, dp, dp, C.POINTER(C.c_longlong)]
    lib.hora_set_time.argtypes = [C.c_void_p, C.c_double, C.c_longlong]
    lib.hora_boundaries.argtypes = [P, dp]
    lib.hora_riemann.argtypes = [P, dp, dp, dp]
    lib.hora_inv_dt.argtypes = [C.c_void_p]
    lib.hora_inv_dt.restype = C.c_double
    lib.hora_op_stages.restype = C.c_int
    lib.hora_op_counts.argtypes = [C.POINTER(C.c_longlong)]
    lib.hora_op_counts_reset.argtypes = []
    _LIBS[name] = lib
    return lib


# stages of the counting build (hydro_oracle.cpp)
OP_STAGES = ("boundaries", "primitives", "slopes", "slopes_neighbour", "sources",
             "faces_riemann", "update", "cfl")


def op_counts():
    """{stage: [add, mul, div, sqrt, minmax]} of the counting build since the last reset."""
    L = _lib(True)
    n = L.hora_op_stages()
    buf = (C.c_longlong * (5 * n))()
    L.hora_op_counts(buf)
    return {OP_STAGES[s]: list(buf[5 * s:5 * s + 5]) for s in range(n)}


def op_counts_reset():
    _lib(True).hora_op_counts_reset()


def _dp(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


def nbvar(p):
    return 5 if p.dim == 3 else 4


def shape(p):
    g = p.ghost_width
    ks = p.nz + 2 * g if p.dim == 3 else 1
    return (nbvar(p), ks, p.ny + 2 * g, p.nx + 2 * g)


def zeros(p):
    return np.zeros(shape(p), dtype=np.float64)


def init_array(p, prob):
    """Initial condition of the interior and ghosts, before any ghost fill."""
    U = zeros(p)
    _lib().hora_init_array(C.byref(p), C.byref(prob), _dp(U))
    return U


def gravity_field(p, prob):
    """The gravity field the problem's initial condition writes, shape (dim, ksize, jsize,
    isize)."""
    G = np.zeros((p.dim,) + shape(p)[1:], dtype=np.float64)
    _lib().hora_gravity_field(C.byref(p), C.byref(prob), _dp(G))
    return G


def fill_boundaries(p, U):
    _lib().hora_boundaries(C.byref(p), _dp(U))
    return U


def riemann(p, ql, qr):
    """Flux of the configured solver for primitive states (rho, p, u_normal, v[, w])."""
    n = nbvar(p)
    a = np.ascontiguousarray(ql, dtype=np.float64)
    b = np.ascontiguousarray(qr, dtype=np.float64)
    assert a.shape == (n,) and b.shape == (n,)
    f = np.zeros(n)
    _lib().hora_riemann(C.byref(p), _dp(a), _dp(b), _dp(f))
    return f


def threads(native=False):
    """OpenMP threads of the hydro oracle's parallel loops."""
    return int(_lib(native=native).hora_threads())


class Solver:
    """The reference's Hydro_Muscl time loop (U/U2 ping-pong, dt from the CFL or fixed)."""

    def __init__(self, p, counting=False, native=False):
        self.p = p
        self._L = _lib(counting, native)
        self._h = self._L.hora_create(C.byref(p))

    def close(self):
        if self._h:
            self._L.hora_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def init_problem(self, prob):
        self._L.hora_init_problem(self._h, C.byref(prob))

    def set_state(self, U):
        U = np.ascontiguousarray(U, dtype=np.float64)
        assert U.shape == shape(self.p)
        self._L.hora_set_state(self._h, _dp(U))

    def get_state(self):
        U = zeros(self.p)
        self._L.hora_get_state(self._h, _dp(U))
        return U

    def step(self, dt=0.0, dt_mode=A.DT_FIXED):
        self._L.hora_step(self._h, float(dt), int(dt_mode))

    def time(self):
        t, dt, it = C.c_double(), C.c_double(), C.c_longlong()
        self._L.hora_get_time(self._h, C.byref(t), C.byref(dt), C.byref(it))
        return t.value, dt.value, it.value

    def set_time(self, t, iteration):
        self._L.hora_set_time(self._h, float(t), int(iteration))

    def inv_dt(self):
        return self._L.hora_inv_dt(self._h)
