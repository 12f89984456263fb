"""A compiled C++ caller of the C ABI (tests/cxx/abi_caller.cpp): the reference-side binding's
call sequence (parameters from HydroParams fields, problem parameters through a ConfigMap
lookup, the reference's main loop with its timers and "Perf" line), built with g++ against
include/ppkmhd_amd.h and the in-tree library.  The error paths run without a GPU; the runs
need one."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cxx", "abi_caller.cpp")
LIBDIR = os.path.join(ROOT, "ppkmhd_amd")


@pytest.fixture(scope="module")
def caller(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("cxx") / "abi_caller")
    cmd = ["g++", "-std=c++17", "-O1", "-Wall", "-Wextra", "-Werror", "-I",
           os.path.join(ROOT, "include"), SRC, "-o", exe, "-L", LIBDIR, "-lppkmhd_amd",
           f"-Wl,-rpath,{LIBDIR}", "-Wl,-rpath-link,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def run(exe, mode):
    r = subprocess.run([exe, mode], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def test_cxx_caller_error_paths(caller):
    out = run(caller, "errors")
    assert "errors: 0 failed checks" in out


@pytest.mark.gpu
def test_cxx_caller_runs(caller):
    out = run(caller, "run")
    assert "Perf             :" in out and "run: 0 failed checks" in out
