"""The single-pass 3D kernel's block program run on the CPU (tests/cxx/march_emu.cpp): the
kernel bodies of march_body.hpp built for the host in the strict arithmetic, with
AddressSanitizer and UndefinedBehaviorSanitizer, every LDS slot and register poisoned with NaN
before each block.  One step at a fixed dt must equal the oracle's step bit for bit (ghost
cells included), and the folded CFL maximum must equal the oracle's CFL maximum of the new
state.  Cases cover partial tiles, block ranges that start and end inside a tile's planes
and span several tiles, the three Riemann solvers, both slope limiters and every boundary
type."""
import os
import subprocess

import numpy as np
import pytest

from ppkmhd_amd import _abi as A
from ppkmhd_amd import default_problem, make_params

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cxx", "march_emu.cpp")


@pytest.fixture(scope="module")
def emu(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("march") / "march_emu")
    cmd = ["/opt/rocm/bin/hipcc", "--offload-host-only", "-std=c++17", "-O1", "-g",
           "-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
           "-fno-sanitize-recover=all", "-DPPK_NS=strict", "-DPPK_EXACT_DIV",
           "-ffp-contract=off", "-I", os.path.join(ROOT, "include"),
           "-I", os.path.join(ROOT, "ppkmhd_amd", "csrc"), SRC, "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    return exe


def emulate(exe, tmp, p, U0, dt, blocks):
    d = str(tmp)
    with open(os.path.join(d, "params.bin"), "wb") as f:
        f.write(bytes(p))
    np.ascontiguousarray(U0).tofile(os.path.join(d, "u0.bin"))
    with open(os.path.join(d, "meta.txt"), "w") as f:
        f.write(f"{dt!r} {blocks}\n")
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([exe, d], capture_output=True, text=True, timeout=600, env=env)
    assert r.returncode == 0, r.stdout + r.stderr[-4000:]
    U1 = np.fromfile(os.path.join(d, "u1.bin")).reshape(U0.shape)
    cells = np.fromfile(os.path.join(d, "cfl_cells.bin")).reshape(U0.shape[1:])
    cfl = float(open(os.path.join(d, "cfl.txt")).read())
    return U1, cells, cfl


def cfl_terms(p, Q):
    """The CFL term sum_d (c_f,d + |v_d|) / dx_d of every cell of the primitive state Q, in the
    operation order of src/shared/mhd_utils.h:319-366 (numpy rounds every operation, as the
    strict build does)."""
    d, pp, u, v, w, a, b, c = Q
    b2 = a * a + b * b + c * c
    c2 = p.gamma0 * pp / d
    d2 = 0.5 * (b2 / d + c2)
    out = 0.0
    for vel, bn, dx in ((u, a, p.dx), (v, b, p.dy), (w, c, p.dz)):
        cf = np.sqrt(d2 + np.sqrt(d2 * d2 - c2 * bn * bn / d))
        out = out + (cf + np.abs(vel)) / dx
    return out


def blast(nx, ny, nz, **kw):
    p = make_params(3, nx, ny, nz, gamma0=5 / 3, cfl=0.8, tEnd=1.0, **kw)
    pr = default_problem(A.PROBLEM_BLAST, p)
    pr.blast_radius = 0.25
    pr.blast_density_out = 1.0
    pr.blast_bx, pr.blast_by, pr.blast_bz = 0.4, -0.7, 0.55
    return p, pr


def kh(nx, ny, nz, **kw):
    p = make_params(3, nx, ny, nz, xmin=-0.5, xmax=0.5, ymin=-0.5, ymax=0.5, zmin=-0.5,
                    zmax=0.5, gamma0=1.4, cfl=0.8, tEnd=1.0, **kw)
    pr = default_problem(A.PROBLEM_KELVIN_HELMHOLTZ, p)
    pr.kh_p_rand, pr.kh_p_sine, pr.kh_p_sine_rob = 1, 0, 0
    pr.kh_axis = 1
    pr.kh_amplitude = 0.05
    pr.kh_bx, pr.kh_by, pr.kh_bz = 0.1, 0.05, -0.08
    return p, pr


# (problem, size, parameters, blocks): tiles of 16 x 12 output columns; each block first
# marches whole tiles (one per block per round), then the left-over tiles' tile-planes are
# cut into `blocks` equal ranges, which start and end inside a tile's z range and may span
# two tiles
CASES = {
    # periodic, tiles cut at 40 / 29: 9 tiles x 20 planes; 7 whole tiles, then 2 tiles in 7
    # ranges of 6 planes
    "blast_hlld_minmod": (blast, (40, 29, 20), dict(slope_type=1.0, riemann=A.RIEMANN_HLLD), 7),
    # more blocks than tiles: every tile shared out in ranges of 17 planes
    "blast_hlld_ranges_only": (blast, (40, 29, 20), dict(slope_type=1.0,
                                                         riemann=A.RIEMANN_HLLD), 11),
    # every boundary type, MC slopes, HLL, odd sizes: 6 tiles x 13 planes; 5 whole tiles,
    # then one tile in 5 ranges of 3 planes
    "kh_hll_mc_bcs": (kh, (17, 31, 13), dict(slope_type=2.0, riemann=A.RIEMANN_HLL,
                                             bc=(A.BC_DIRICHLET, A.BC_NEUMANN, A.BC_NEUMANN,
                                                 A.BC_DIRICHLET, A.BC_PERIODIC,
                                                 A.BC_PERIODIC)), 5),
    # exact multiples of the output tile, LLF, non-periodic z: one block runs both tiles
    "kh_llf_whole": (kh, (32, 12, 9), dict(slope_type=1.0, riemann=A.RIEMANN_LLF,
                                           bc=(A.BC_PERIODIC, A.BC_PERIODIC, A.BC_PERIODIC,
                                               A.BC_PERIODIC, A.BC_DIRICHLET,
                                               A.BC_NEUMANN)), 1),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_march_block_program_matches_oracle(emu, oracle, tmp_path, case):
    make, size, kw, blocks = CASES[case]
    p, pr = make(*size, precision=A.FP64_STRICT, **kw)
    o = oracle.Solver(p)
    o.init_problem(pr)
    o.step(0, A.DT_CFL)  # a state with structure in every variable
    U0 = oracle.fill_boundaries(p, o.get_state().copy())
    dt = 0.5 * p.cfl / oracle.inv_dt(p, oracle.primitives(p, U0))
    o.set_state(U0)
    o.step(dt, A.DT_FIXED)
    ref = o.get_state()
    U1, cells, cfl = emulate(emu, tmp_path, p, U0, dt, blocks)
    bad = ~((U1 == ref) | (np.isnan(U1) & np.isnan(ref)))
    assert not bad.any(), f"{bad.sum()} values differ, first at {np.argwhere(bad)[:5]}"
    # the CFL term of every new interior cell, and their maximum, against the oracle's
    Q = oracle.primitives(p, oracle.fill_boundaries(p, ref.copy()))
    inner = p.interior()[1:]
    terms = cfl_terms(p, Q[p.interior()])
    assert terms.max() == oracle.inv_dt(p, Q)  # the restatement of the term is the oracle's
    assert np.array_equal(cells[inner], terms)
    assert cfl == terms.max()
