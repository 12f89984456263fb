"""The library's ghost fill on the host (tests/cxx/ghost_fill.cpp, built with
AddressSanitizer and UndefinedBehaviorSanitizer): the one-pass fill that composes the
per-direction source maps and Dirichlet signs for corner and edge ghosts equals the oracle's
per-direction passes (x, then y, then z, as src/shared/SolverBase.cpp:339-537 and
src/shared/BoundariesFunctors.h:534-1057 run them) bit for bit, ghosts of every corner and
edge included, for every mix of boundary types in 2D and 3D; its copy target receives exactly
the ghost values and nothing else."""
import itertools
import os
import subprocess

import numpy as np
import pytest

from ppkmhd_amd import _abi as A
from ppkmhd_amd import make_params

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cxx", "ghost_fill.cpp")
BCS = (A.BC_DIRICHLET, A.BC_NEUMANN, A.BC_PERIODIC)


@pytest.fixture(scope="module")
def exe(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("ghost") / "ghost_fill")
    cmd = ["/opt/rocm/bin/hipcc", "--offload-host-only", "-std=c++17", "-O1", "-g",
           "-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
           "-fno-sanitize-recover=all", "-DPPK_NS=strict", "-I", os.path.join(ROOT, "include"),
           "-I", os.path.join(ROOT, "ppkmhd_amd", "csrc"), SRC, "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    return out


def cases():
    # 2D: every pair of (x, y) face types on both sides (lower and upper of a direction may differ)
    for bx, by in itertools.product(itertools.product(BCS, repeat=2), repeat=2):
        yield 2, (7, 5, 1), (bx[0], bx[1], by[0], by[1], A.BC_PERIODIC, A.BC_PERIODIC)
    # 3D: each type on all faces, and mixed sets
    for b in BCS:
        yield 3, (6, 5, 4), (b,) * 6
    yield 3, (5, 6, 7), (A.BC_DIRICHLET, A.BC_NEUMANN, A.BC_PERIODIC, A.BC_PERIODIC,
                         A.BC_NEUMANN, A.BC_DIRICHLET)
    yield 3, (4, 4, 5), (A.BC_NEUMANN, A.BC_PERIODIC, A.BC_DIRICHLET, A.BC_DIRICHLET,
                         A.BC_PERIODIC, A.BC_NEUMANN)


@pytest.mark.parametrize("case", list(cases()), ids=lambda c: f"{c[0]}d-{''.join(map(str, c[2]))}")
def test_ghost_fills_equal_the_oracle(exe, oracle, tmp_path, case):
    dim, (nx, ny, nz), bc = case
    p = make_params(dim, nx, ny, nz, bc=list(bc))
    rng = np.random.default_rng(sum(bc) + 10 * dim)
    U0 = rng.standard_normal((8, p.ksize, p.jsize, p.isize))
    d = str(tmp_path)
    with open(os.path.join(d, "params.bin"), "wb") as f:
        f.write(bytes(p))
    U0.tofile(os.path.join(d, "u0.bin"))
    r = subprocess.run([exe, d], capture_output=True, text=True, timeout=120,
                       env=dict(os.environ, ASAN_OPTIONS="detect_leaks=1"))
    assert r.returncode == 0, (r.stdout + r.stderr)[-4000:]
    ref = oracle.fill_boundaries(p, U0.copy())
    shell = np.fromfile(os.path.join(d, "u_shell.bin")).reshape(U0.shape)
    copy = np.fromfile(os.path.join(d, "u_copy.bin")).reshape(U0.shape)
    assert np.array_equal(shell, ref)
    ghost = np.ones(U0.shape, dtype=bool)
    ghost[p.interior()] = False
    assert np.array_equal(copy[ghost], ref[ghost])
    assert np.isnan(copy[~ghost]).all()
