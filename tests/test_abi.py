"""The C-ABI library (CPU checks: no GPU compute).

The library must load, export every function declared in include/ppkmhd_amd.h, agree with
the ctypes mirror of its structs, and parse .ini files with the reference's rules.
"""
import ctypes as C
import os
import re

import numpy as np
import pytest

from ppkmhd_amd import _abi as A
from ppkmhd_amd import f32, make_params
from ppkmhd_amd._lib import LIB_PATH, SIGNATURES, PPKError, lib
from ppkmhd_amd.solver import params_from_ini

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include",
                      "ppkmhd_amd.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(ppk_(?:mhd|mlp|faceset|hydro)_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB_PATH), "build the library first (python -m ppkmhd_amd.build)"
    L = C.CDLL(LIB_PATH)
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # and the Python binding covers the whole header
    assert set(names) == set(SIGNATURES), set(names) ^ set(SIGNATURES)


def test_struct_layouts_match():
    L = lib()
    assert L.ppk_mhd_params_size() == C.sizeof(A.Params)
    assert L.ppk_mhd_problem_size() == C.sizeof(A.Problem)
    assert L.ppk_mhd_version().decode().startswith("ppkmhd_amd")


def test_params_init_matches_reference_formulas():
    p = make_params(3, 16, 8, 4, xmax=2.0, ymax=1.0, zmax=0.5, gamma0=5 / 3, smallc=1e-7,
                    smallr=1e-8)
    q = A.Params.from_buffer_copy(p)
    q.dx = q.dy = q.dz = q.smallp = q.smallpp = q.gamma6 = 0.0
    assert lib().ppk_mhd_params_init(C.byref(q)) == 0
    for f in ("dx", "dy", "dz", "smallp", "smallpp", "gamma6"):
        assert getattr(q, f) == getattr(p, f), f


OT_INI = """
[run]
solver_name=MHD_Muscl_2D
tEnd=0.5
nStepmax=1000
nOutput=0

[mesh]
nx=256
ny=256
boundary_type_xmin=3
boundary_type_xmax=3
boundary_type_ymin=3
boundary_type_ymax=3

[hydro]
gamma0=1.666666666667
cfl=0.8
slope_type=1
problem=orszag_tang
riemann=hlld
smallr=1e-7
smallc=1e-7

[OTHER]
implementationVersion=0
"""


def test_ini_reference_semantics():
    p, pr, name = params_from_ini(text=OT_INI)
    assert name == "MHD_Muscl_2D"
    assert p.dim == 2 and p.nx == 256 and p.ny == 256 and p.nz == 1
    assert p.ghost_width == 3
    assert p.riemann == A.RIEMANN_HLLD
    assert list(p.bc)[:4] == [3, 3, 3, 3]
    # reals come back float-rounded (ConfigMap::getFloat)
    assert p.gamma0 == f32(1.666666666667) and p.gamma0 != 5 / 3
    assert p.cfl == f32(0.8) and p.tEnd == f32(0.5)
    assert p.smallc == f32(1e-7) and p.smallr == f32(1e-7)
    ref = make_params(2, 256, 256, gamma0=1.666666666667, cfl=0.8, slope_type=1, smallr=1e-7,
                      smallc=1e-7, tEnd=0.5, riemann=A.RIEMANN_HLLD)
    for f in ("dx", "dy", "smallp", "smallpp", "gamma6", "xmin", "xmax"):
        assert getattr(p, f) == getattr(ref, f), f
    assert pr.kind == A.PROBLEM_ORSZAG_TANG


def test_ini_case_insensitive_keys_and_comments():
    txt = OT_INI.replace("[run]", "[RUN]").replace("tEnd=0.5", "TEND = 0.25 ; inline comment")
    txt = txt.replace("[mesh]", "# a comment line\n[Mesh]").replace("nx=256", "NX=128")
    p, _, _ = params_from_ini(text=txt)
    assert p.tEnd == f32(0.25)
    assert p.nx == 128


def test_ini_defaults_and_problem_params():
    txt = """[run]
solver_name=MHD_Muscl_3D
[mesh]
nx=8
ny=8
nz=8
[hydro]
problem=blast
riemann=hll
[blast]
density_out=1.0
"""
    p, pr, _ = params_from_ini(text=txt)
    assert p.dim == 3
    # HydroParams::setup defaults (src/shared/HydroParams.cpp)
    assert p.gamma0 == f32(1.4) and p.cfl == f32(0.5) and p.slope_type == 1.0
    assert p.smallc == f32(1e-10) and p.n_step_max == 1000
    assert list(p.bc) == [A.BC_DIRICHLET] * 6
    assert p.riemann == A.RIEMANN_HLL
    assert pr.kind == A.PROBLEM_BLAST
    assert pr.blast_density_out == 1.0 and pr.blast_density_in == 1.0
    assert pr.blast_radius == f32(0.05) and pr.blast_pressure_in == 10.0
    assert pr.blast_pressure_out == f32(0.1)
    assert pr.blast_bx == 0.5


def test_ini_rotor_and_field_loop():
    """settings/test_mhd_rotor.ini's [rotor] section; "field_loop" and "field loop" select the
    field loop, while "FieldLoop" (the spelling of settings/mhd_fieldloop2d.ini) matches none of
    the reference's names and falls back to Orszag-Tang as the reference does; the hydro solver
    falls back to implode for both."""
    import math
    base = """[run]
solver_name=MHD_Muscl_2D
[mesh]
nx=8
ny=8
[hydro]
problem=%s
[rotor]
r0=0.1
r1=0.115
u0=1.0
p0=1.5
[FieldLoop]
radius=0.3
amplitude=0.001
vflow=3
"""
    p, pr, _ = params_from_ini(text=base % "rotor")
    assert pr.kind == A.PROBLEM_ROTOR
    assert (pr.rotor_r1, pr.rotor_u0, pr.rotor_p0) == (f32(0.115), 1.0, 1.5)
    assert pr.rotor_b0 == f32(5.0 / math.sqrt(4 * math.pi))
    for name in ("field_loop", "field loop"):
        _, pr2, _ = params_from_ini(text=base % name)
        assert pr2.kind == A.PROBLEM_FIELD_LOOP
        assert (pr2.loop_radius, pr2.loop_amplitude, pr2.loop_vflow) == (f32(0.3), f32(0.001), 3.0)
    assert params_from_ini(text=base % "FieldLoop")[1].kind == A.PROBLEM_ORSZAG_TANG
    hydro = (base % "rotor").replace("MHD_Muscl_2D", "Hydro_Muscl_2D")
    assert params_from_ini(text=hydro)[1].kind == A.PROBLEM_IMPLODE


def test_ini_unknown_solver_is_an_error():
    with pytest.raises(PPKError) as e:
        params_from_ini(text="[run]\nsolver_name=Hydro_SDM_2D_degree2\n")
    assert e.value.code == 5
    assert "Hydro_SDM_2D_degree2" in str(e.value)


def test_ini_missing_file_is_an_error(tmp_path):
    with pytest.raises(PPKError) as e:
        params_from_ini(path=tmp_path / "nope.ini")
    assert e.value.code == 5


def test_reference_settings_files_parse():
    """The MHD .ini files shipped with the reference parse (their keys exist)."""
    txt = """[run]
solver_name=MHD_Muscl_2D
tEnd=1.0
nStepmax=1000
nOutput=10
[mesh]
nx=256
ny=256
boundary_type_xmin=3
boundary_type_xmax=3
boundary_type_ymin=3
boundary_type_ymax=3
[hydro]
gamma0=1.666
cfl=0.8
niter_riemann=10
iorder=2
slope_type=2
problem=orszag_tang
riemann=hlld
smallr=1e-7
smallc=1e-7
[output]
outputPrefix=orszag_tang_2d
outputVtkAscii=false
hdf5_enabled=yes
[other]
implementationVersion=2
"""
    p, pr, name = params_from_ini(text=txt)
    assert p.slope_type == 2.0 and p.gamma0 == f32(1.666)
    # settings/test_mhd_orszag_tang_2D.ini:32 selects implementation version 2, a different
    # scheme (SolverMHDMuscl.cpp:430-456) that this library does not provide: creating the
    # solver is refused (parameter checks run before any device call), never silently run as
    # version 0
    assert p.implementation_version == 2
    h = C.c_void_p()
    assert lib().ppk_mhd_create(C.byref(p), C.byref(h)) == A.PPK_ERR_UNSUPPORTED
    assert not h.value
    assert "implementationVersion 2" in lib().ppk_mhd_last_error(None).decode()
    p.implementation_version = 3  # outside 0..2: the reference exits (HydroParams.cpp:446-451)
    assert lib().ppk_mhd_create(C.byref(p), C.byref(h)) == A.PPK_ERR_INVALID_ARGUMENT
    assert params_from_ini(text=OT_INI)[0].implementation_version == 0


def test_batch_create_rejects_bad_members_before_any_device_call():
    """Parameter checks of ppk_mhd_batch_create run on the host: mixed dimensions, mixed
    precisions, bad sizes and an empty batch are refused with a message naming the member."""
    L = lib()
    a = make_params(2, 16, 16)
    b = make_params(3, 8, 8, 8)
    c = make_params(2, 16, 16, precision=A.FP32_BULK)
    d = make_params(2, 16, 16)
    d.nx = 0
    h = C.c_void_p()
    for members, what in (([a, b], "share dim"), ([a, c], "share dim"), ([a, d], "member 1")):
        arr = (A.Params * len(members))(*members)
        rc = L.ppk_mhd_batch_create(arr, len(members), C.byref(h))
        assert rc == A.PPK_ERR_INVALID_ARGUMENT
        assert not h.value
        assert what in L.ppk_mhd_batch_last_error(None).decode()
    arr = (A.Params * 1)(a)
    assert L.ppk_mhd_batch_create(arr, 0, C.byref(h)) == A.PPK_ERR_INVALID_ARGUMENT
