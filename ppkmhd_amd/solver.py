"""Python face of the solver, shaped like the reference's ``SolverBase`` / ``SolverFactory``.

``SolverFactory.Instance().create("MHD_Muscl_3D", params, configMap)`` returns a solver with
``next_iteration()``, ``compute_dt()``, ``finished()``, ``m_t``, ``m_dt``, ``m_iteration``,
``m_nCells`` (reference src/shared/SolverBase.h:53-157, src/shared/SolverFactory.h:43-116).
Every compute call goes through the HIP library's C ABI.
"""
import ctypes as C

import numpy as np

from . import _abi as A
from ._lib import PPKError, check, check_batch, check_with, lib
from .params import default_problem


def params_from_ini(path=None, text=None):
    """(params, problem, solver_name) from an .ini file or text, with the reference's rules."""
    p, pr = A.Params(), A.Problem()
    name = C.create_string_buffer(128)
    if text is not None:
        rc = lib().ppk_mhd_params_from_ini_text(text.encode(), C.byref(p), C.byref(pr), name, 128)
    else:
        rc = lib().ppk_mhd_params_from_ini(str(path).encode(), C.byref(p), C.byref(pr), name, 128)
    check(rc)
    return p, pr, name.value.decode()


class SolverMHDMuscl:
    """MHD MUSCL-Hancock solver on one GPU (one C-ABI handle)."""

    def __init__(self, params, problem=None):
        self.params = params
        h = C.c_void_p()
        check(lib().ppk_mhd_create(C.byref(params), C.byref(h)))
        self._h = h
        self.m_nCells = params.isize * params.jsize * params.ksize  # ghost cells included (H10)
        self.m_nCells_interior = params.nx * params.ny * (params.nz if params.dim == 3 else 1)
        if problem is not None:
            self.init_problem(problem)

    # ---- lifetime -------------------------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            lib().ppk_mhd_destroy(self._h)
            self._h = None
        # an attached surrogate network may be freed only once the handle is gone
        self._sur_net = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc):
        return check(rc, self._h)

    # ---- state ----------------------------------------------------------------------------
    def init_problem(self, problem, on_device=False):
        """Initial conditions (host-computed and uploaded, or computed on the GPU)."""
        fn = lib().ppk_mhd_init_problem_device if on_device else lib().ppk_mhd_init_problem
        self._check(fn(self._h, C.byref(problem)))

    def upload(self, U):
        U = np.ascontiguousarray(U, dtype=np.float64)
        if U.size != self.params.count:
            raise ValueError(f"state must have {self.params.count} doubles, got {U.size}")
        self._check(lib().ppk_mhd_upload(self._h, U.ctypes.data_as(C.POINTER(C.c_double)), U.size))

    def download(self):
        U = np.empty(self.params.shape, dtype=np.float64)
        self._check(lib().ppk_mhd_download(self._h, U.ctypes.data_as(C.POINTER(C.c_double)),
                                           U.size))
        return U

    def device_state_ptr(self):
        ptr = C.POINTER(C.c_double)()
        self._check(lib().ppk_mhd_device_state(self._h, C.byref(ptr)))
        return C.cast(ptr, C.c_void_p).value

    def stream_ptr(self):
        s = C.c_void_p()
        self._check(lib().ppk_mhd_stream(self._h, C.byref(s)))
        return s.value

    # ---- time loop (SolverBase) -------------------------------------------------------------
    def make_boundaries(self):
        self._check(lib().ppk_mhd_make_boundaries(self._h))

    def compute_dt_local(self):
        dt = C.c_double()
        self._check(lib().ppk_mhd_compute_dt(self._h, C.byref(dt)))
        return dt.value

    def step(self, dt=0.0, dt_mode=A.DT_CFL):
        self._check(lib().ppk_mhd_step(self._h, float(dt), int(dt_mode)))

    def next_iteration(self):
        """SolverBase::next_iteration with reference dt semantics."""
        self.step(0.0, A.DT_CFL)

    def run(self, nsteps, dt=0.0, dt_mode=A.DT_CFL):
        self._check(lib().ppk_mhd_run(self._h, int(nsteps), float(dt), int(dt_mode)))

    def time(self):
        t, dt, it = C.c_double(), C.c_double(), C.c_int64()
        self._check(lib().ppk_mhd_get_time(self._h, C.byref(t), C.byref(dt), C.byref(it)))
        return t.value, dt.value, it.value

    def set_time(self, t, iteration):
        self._check(lib().ppk_mhd_set_time(self._h, float(t), int(iteration)))

    # ---- restart (SolverMHDMuscl::init_restart, IO_HDF5 "time step" / "total time") --------
    def save_state(self, path, time_step=0):
        """Write the state, t and the iteration to a restart file."""
        self._check(lib().ppk_mhd_save_state(self._h, str(path).encode(), int(time_step)))

    def load_state(self, path):
        """Restore state, t and iteration from a restart file; returns its "time step"."""
        ts = C.c_int64()
        self._check(lib().ppk_mhd_load_state(self._h, str(path).encode(), C.byref(ts)))
        return ts.value

    @property
    def m_t(self):
        return self.time()[0]

    @property
    def m_dt(self):
        return self.time()[1]

    @property
    def m_iteration(self):
        return self.time()[2]

    def finished(self):
        f = C.c_int()
        self._check(lib().ppk_mhd_finished(self._h, C.byref(f)))
        return bool(f.value)

    def synchronize(self):
        self._check(lib().ppk_mhd_synchronize(self._h))

    # ---- per-stage GPU timers -------------------------------------------------------------
    @staticmethod
    def timer_names():
        L = lib()
        return [L.ppk_mhd_timer_name(i).decode() for i in range(L.ppk_mhd_timer_count())]

    def timers_enable(self, stages=None):
        """Bracket the given stages (names or indexes; None = all, [] = off) with hipEvents."""
        names = self.timer_names()
        if stages is None:
            mask = (1 << len(names)) - 1
        else:
            mask = 0
            for st in stages:
                mask |= 1 << (names.index(st) if isinstance(st, str) else int(st))
        self._check(lib().ppk_mhd_timers_enable(self._h, mask))

    def timers_read(self):
        """{stage name: (total ms, launches)} for stages with at least one launch."""
        out = {}
        for i, n in enumerate(self.timer_names()):
            ms, cnt = C.c_double(), C.c_int64()
            self._check(lib().ppk_mhd_timers_read(self._h, i, C.byref(ms), C.byref(cnt)))
            if cnt.value:
                out[n] = (ms.value, cnt.value)
        return out

    # ---- diagnostics ------------------------------------------------------------------------
    def divb_max(self):
        d = C.c_double()
        self._check(lib().ppk_mhd_divb_max(self._h, C.byref(d)))
        return d.value

    def conserved_sums(self):
        s = (C.c_double * 8)()
        self._check(lib().ppk_mhd_conserved_sums(self._h, s))
        return np.array(list(s))

    def save_vtk(self, directory=".", prefix="output", istep=None, ascii=False):
        """Write the current state as <directory>/<prefix>_<istep>.vti in the reference's VTK
        format (istep defaults to the iteration count)."""
        if istep is None:
            istep = self.m_iteration
        self._check(lib().ppk_mhd_save_vtk(self._h, str(directory).encode(), prefix.encode(),
                                           int(istep), 1 if ascii else 0))


class SolverHydroMuscl:
    """Euler MUSCL-Hancock solver on one GPU (the reference's Hydro_Muscl_2D/3D,
    src/muscl/SolverHydroMuscl.{h,cpp}); state arrays are (nbvar, ksize, jsize, isize) with
    nbvar 4 (2D) or 5 (3D)."""

    def __init__(self, params, problem=None):
        self.params = params
        h = C.c_void_p()
        L = lib()
        check_with(L.ppk_hydro_create(C.byref(params), C.byref(h)), None, L.ppk_hydro_last_error)
        self._h = h
        self.nbvar = L.ppk_hydro_nbvar(h)
        self.m_nCells = params.isize * params.jsize * params.ksize
        if problem is not None:
            self.init_problem(problem)

    def close(self):
        if getattr(self, "_h", None):
            lib().ppk_hydro_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc):
        return check_with(rc, self._h, lib().ppk_hydro_last_error)

    def init_problem(self, problem):
        self._check(lib().ppk_hydro_init_problem(self._h, C.byref(problem)))

    def upload(self, U):
        U = np.ascontiguousarray(U, dtype=np.float64)
        if U.size != self.params.count:
            raise ValueError(f"state must have {self.params.count} doubles, got {U.size}")
        self._check(lib().ppk_hydro_upload(self._h, U.ctypes.data_as(C.POINTER(C.c_double)),
                                           U.size))

    def download(self):
        U = np.empty(self.params.shape, dtype=np.float64)
        self._check(lib().ppk_hydro_download(self._h, U.ctypes.data_as(C.POINTER(C.c_double)),
                                             U.size))
        return U

    def device_state_ptr(self):
        ptr = C.POINTER(C.c_double)()
        self._check(lib().ppk_hydro_device_state(self._h, C.byref(ptr)))
        return C.cast(ptr, C.c_void_p).value

    def stream_ptr(self):
        s = C.c_void_p()
        self._check(lib().ppk_hydro_stream(self._h, C.byref(s)))
        return s.value

    def make_boundaries(self):
        self._check(lib().ppk_hydro_make_boundaries(self._h))

    def compute_dt_local(self):
        dt = C.c_double()
        self._check(lib().ppk_hydro_compute_dt(self._h, C.byref(dt)))
        return dt.value

    def step(self, dt=0.0, dt_mode=A.DT_CFL):
        self._check(lib().ppk_hydro_step(self._h, float(dt), int(dt_mode)))

    def next_iteration(self):
        """SolverBase::next_iteration with reference dt semantics."""
        self.step(0.0, A.DT_CFL)

    def run(self, nsteps, dt=0.0, dt_mode=A.DT_CFL):
        self._check(lib().ppk_hydro_run(self._h, int(nsteps), float(dt), int(dt_mode)))

    def time(self):
        t, dt, it = C.c_double(), C.c_double(), C.c_int64()
        self._check(lib().ppk_hydro_get_time(self._h, C.byref(t), C.byref(dt), C.byref(it)))
        return t.value, dt.value, it.value

    def set_time(self, t, iteration):
        self._check(lib().ppk_hydro_set_time(self._h, float(t), int(iteration)))

    @property
    def m_t(self):
        return self.time()[0]

    @property
    def m_dt(self):
        return self.time()[1]

    @property
    def m_iteration(self):
        return self.time()[2]

    def finished(self):
        f = C.c_int()
        self._check(lib().ppk_hydro_finished(self._h, C.byref(f)))
        return bool(f.value)

    def synchronize(self):
        self._check(lib().ppk_hydro_synchronize(self._h))

    def save_vtk(self, directory=".", prefix="output", istep=None, ascii=False):
        """Write the current state as <directory>/<prefix>_<istep>.vti in the reference's VTK
        format (rho, energy, rho_vx, rho_vy[, rho_vz]; istep defaults to the iteration)."""
        if istep is None:
            istep = self.m_iteration
        self._check(lib().ppk_hydro_save_vtk(self._h, str(directory).encode(), prefix.encode(),
                                             int(istep), 1 if ascii else 0))

    @staticmethod
    def timer_names():
        L = lib()
        return [L.ppk_hydro_timer_name(i).decode() for i in range(L.ppk_hydro_timer_count())]

    def timers_enable(self, stages=None):
        """Bracket the given stages (names or indexes; None = all, [] = off) with hipEvents."""
        names = self.timer_names()
        if stages is None:
            mask = (1 << len(names)) - 1
        else:
            mask = 0
            for st in stages:
                mask |= 1 << (names.index(st) if isinstance(st, str) else int(st))
        self._check(lib().ppk_hydro_timers_enable(self._h, mask))

    def timers_read(self):
        """{stage name: (total ms, launches)} for stages with at least one launch."""
        out = {}
        for i, n in enumerate(self.timer_names()):
            ms, cnt = C.c_double(), C.c_int64()
            self._check(lib().ppk_hydro_timers_read(self._h, i, C.byref(ms), C.byref(cnt)))
            if cnt.value:
                out[n] = (ms.value, cnt.value)
        return out


def hydro_initial_state(params, problem):
    """The initial state of a hydro problem (implode, blast) computed on the host, no GPU."""
    U = np.empty(params.shape, dtype=np.float64)
    L = lib()
    check_with(L.ppk_hydro_initial_state(C.byref(params), C.byref(problem),
                                         U.ctypes.data_as(C.POINTER(C.c_double)), U.size),
               None, L.ppk_hydro_last_error)
    return U


def initial_state(params, problem):
    """The initial state of a problem computed on the host (reference layout), no GPU."""
    U = np.empty(params.shape, dtype=np.float64)
    check(lib().ppk_mhd_initial_state(C.byref(params), C.byref(problem),
                                      U.ctypes.data_as(C.POINTER(C.c_double)), U.size))
    return U


def write_vtk(params, U, directory=".", prefix="output", istep=0, ascii=False):
    """Write a host state array (reference layout) in the reference's VTK format (no GPU)."""
    U = np.ascontiguousarray(U, dtype=np.float64)
    if U.size != params.count:
        raise ValueError("state must hold nbvar*isize*jsize*ksize doubles")
    check(lib().ppk_mhd_write_vtk(C.byref(params), U.ctypes.data_as(C.POINTER(C.c_double)),
                                  str(directory).encode(), prefix.encode(), int(istep),
                                  1 if ascii else 0))


class BatchSolver:
    """An ensemble of independent MHD simulations advanced together on one GPU (one C-ABI batch
    handle): each member has its own grid size, parameters, time, CFL dt and finished() test,
    and evolves exactly as a ``SolverMHDMuscl`` with the same parameters would in DT_CFL mode.
    """

    def __init__(self, params_list, problems=None):
        self.params = list(params_list)
        arr = (A.Params * len(self.params))(*self.params)
        h = C.c_void_p()
        check_batch(lib().ppk_mhd_batch_create(arr, len(self.params), C.byref(h)))
        self._h = h
        if problems is not None:
            self.init_problem(problems)

    def close(self):
        if getattr(self, "_h", None):
            lib().ppk_mhd_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __len__(self):
        return len(self.params)

    def _check(self, rc):
        return check_batch(rc, self._h)

    def init_problem(self, problems):
        if len(problems) != len(self.params):
            raise ValueError("one problem per member")
        arr = (A.Problem * len(problems))(*problems)
        self._check(lib().ppk_mhd_batch_init_problem(self._h, arr))

    def upload(self, member, U):
        U = np.ascontiguousarray(U, dtype=np.float64)
        if U.size != self.params[member].count:
            raise ValueError(f"state must have {self.params[member].count} doubles, got {U.size}")
        self._check(lib().ppk_mhd_batch_upload(self._h, member,
                                               U.ctypes.data_as(C.POINTER(C.c_double)), U.size))

    def download(self, member):
        U = np.empty(self.params[member].shape, dtype=np.float64)
        self._check(lib().ppk_mhd_batch_download(self._h, member,
                                                 U.ctypes.data_as(C.POINTER(C.c_double)), U.size))
        return U

    def run(self, nsteps):
        """nsteps steps of every member not yet finished (synchronizes)."""
        self._check(lib().ppk_mhd_batch_run(self._h, int(nsteps)))

    def time(self, member):
        """(t, dt of the last step, iteration) of one member."""
        t, dt, it = C.c_double(), C.c_double(), C.c_int64()
        self._check(lib().ppk_mhd_batch_get_time(self._h, member, C.byref(t), C.byref(dt),
                                                 C.byref(it)))
        return t.value, dt.value, it.value

    def synchronize(self):
        self._check(lib().ppk_mhd_batch_synchronize(self._h))

    def set_tiles(self, tile, brick):
        """Half-step tile shape and cell order of the batched kernels (results are unchanged)."""
        self._check(lib().ppk_mhd_batch_set_tiles(self._h, int(tile), int(brick)))

    def tiles(self):
        t, b = C.c_int(), C.c_int()
        self._check(lib().ppk_mhd_batch_get_tiles(self._h, C.byref(t), C.byref(b)))
        return t.value, b.value

    def autotune(self, trial_steps=5):
        """Time every tile x brick candidate on the members themselves (their states and times
        are restored), keep the fastest; returns (tile, brick, ms per step)."""
        t, b, ms = C.c_int(), C.c_int(), C.c_double()
        self._check(lib().ppk_mhd_batch_autotune(self._h, int(trial_steps), C.byref(t), C.byref(b),
                                                 C.byref(ms)))
        return t.value, b.value, ms.value


class SolverFactory:
    """Registry of solver constructors keyed by the reference's solver names
    (src/shared/SolverFactory.cpp:26-30)."""

    _instance = None

    def __init__(self):
        self._map = {}
        self.registerSolver("Hydro_Muscl_2D", SolverHydroMuscl)
        self.registerSolver("Hydro_Muscl_3D", SolverHydroMuscl)
        self.registerSolver("MHD_Muscl_2D", SolverMHDMuscl)
        self.registerSolver("MHD_Muscl_3D", SolverMHDMuscl)

    @classmethod
    def Instance(cls):
        if cls._instance is None:
            cls._instance = SolverFactory()
        return cls._instance

    def registerSolver(self, key, fn):
        self._map[key] = fn

    def create(self, solver_name, params, problem=None):
        fn = self._map.get(solver_name)
        if fn is None:
            raise PPKError(A.PPK_OK + 5, f"{solver_name}: is not recognized as a valid application "
                           f"name key; valid names: {sorted(self._map)}")
        if problem is None:  # the solvers' fallbacks: Orszag-Tang (MHD), implode (hydro)
            kind = A.PROBLEM_IMPLODE if solver_name.startswith("Hydro") else A.PROBLEM_ORSZAG_TANG
            problem = default_problem(kind, params)
        return fn(params, problem)


def write_restart(path, params, U, total_time=0.0, time_step=0, iteration=0):
    """Host-side restart writer (ppk_mhd_write_restart): no device involved."""
    U = np.ascontiguousarray(U, dtype=np.float64)
    if U.size != params.count:
        raise ValueError(f"state must have {params.count} doubles, got {U.size}")
    check(lib().ppk_mhd_write_restart(C.byref(params), U.ctypes.data_as(C.POINTER(C.c_double)),
                                      float(total_time), int(time_step), int(iteration),
                                      str(path).encode()))


def read_restart(path, params):
    """Host-side restart reader: (state, total time, time step, iteration)."""
    U = np.empty(params.shape, dtype=np.float64)
    t, ts, it = C.c_double(), C.c_int64(), C.c_int64()
    check(lib().ppk_mhd_read_restart(C.byref(params), str(path).encode(),
                                     U.ctypes.data_as(C.POINTER(C.c_double)), U.size,
                                     C.byref(t), C.byref(ts), C.byref(it)))
    return U, t.value, ts.value, it.value
