"""GPU parity of the hydro (Euler) solvers, Hydro_Muscl_2D/3D, against the CPU oracle.

PPK_FP64_STRICT: every step bit-identical to the oracle (state incl. ghosts, t, dt), for the
staged and the fused kernels, every Riemann solver, slope types 0/1/2, all boundary types,
sizes that leave partial strips / blocks.  PPK_FP64 (FMA contraction): relative L-inf
difference <= 1e-12 after a few steps.  Larger grids: fused and staged strict results equal.
The committed golden fixtures (tests/golden/hydro_*.npz); at the benchmark sizes, mass and
energy conserved by reflecting walls and the state physical.
"""
import numpy as np
import pytest

from ppkmhd_amd import _abi as A
from ppkmhd_amd.params import default_problem, make_hydro_params
from ppkmhd_amd.solver import SolverFactory, SolverHydroMuscl

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def H():
    from oracle import pyhydro

    pyhydro.build()
    return pyhydro


@pytest.fixture
def strip(monkeypatch):
    def set_rows(n):
        monkeypatch.setenv("PPKMHD_HYDRO_STRIP", str(n))
    return set_rows


def hparams(dim, n, **kw):
    kw.setdefault("gamma0", 1.666)
    kw.setdefault("cfl", 0.8)
    kw.setdefault("slope_type", 2.0)
    kw.setdefault("bc", A.BC_DIRICHLET)
    kw.setdefault("tEnd", 10.0)
    kw.setdefault("xmax", 2.0)
    return make_hydro_params(dim, *n, **kw)


def problem(p, kind):
    pr = default_problem(kind, p)
    if kind == A.PROBLEM_BLAST:  # off-centre blast: no symmetry to hide index mistakes
        pr.blast_center_x = 0.7
        pr.blast_center_y = 0.45
        pr.blast_center_z = 0.4
        pr.blast_radius = 0.3
    return pr


def compare_steps(H, p, pr, steps, exact=True, tol=1e-12, tag="", oracle_state=False):
    """Step the GPU solver and the oracle side by side.  oracle_state: start the GPU solver from
    the oracle's initial state (problems whose initial condition calls sin / cos / exp: the
    host math libraries may round those differently by an ulp)."""
    print(f"hydro parity {tag}: dim={p.dim} n=({p.nx},{p.ny},{p.nz}) riemann={p.riemann} "
          f"slope={p.slope_type} bc={list(p.bc)} prec={p.precision} path={p.kernel_path}",
          flush=True)
    o = H.Solver(p)
    o.init_problem(pr)
    g = SolverHydroMuscl(p, pr)
    try:
        if oracle_state:
            g.upload(o.get_state())
        np.testing.assert_array_equal(g.download(), o.get_state())
        for s in range(steps):
            o.step(0, A.DT_CFL)
            g.step(0, A.DT_CFL)
            a, b = g.download(), o.get_state()
            to, dto, ito = o.time()
            tg, dtg, itg = g.time()
            assert ito == itg == s + 1
            if exact:
                assert (tg, dtg) == (to, dto), (s, tg - to, dtg - dto)
                if not np.array_equal(a, b):
                    bad = np.argwhere(a != b)
                    raise AssertionError(f"step {s}: {len(bad)} values differ, first {bad[:5]}")
            else:
                err = np.max(np.abs(a - b)) / np.max(np.abs(b))
                assert err <= tol, (s, err)
                assert abs(dtg - dto) <= tol * dto
    finally:
        g.close()
        o.close()


PATHS = [A.PATH_STAGED, A.PATH_FUSED]


@pytest.mark.parametrize("path", PATHS)
@pytest.mark.parametrize("rs", [A.RIEMANN_APPROX, A.RIEMANN_LLF, A.RIEMANN_HLL, A.RIEMANN_HLLC])
def test_strict_bitwise_2d(gpu, H, strip, path, rs):
    strip(16)
    p = hparams(2, (130, 70), riemann=rs, precision=A.FP64_STRICT, kernel_path=path)
    compare_steps(H, p, problem(p, A.PROBLEM_IMPLODE), 5, tag="2d")


@pytest.mark.parametrize("path", PATHS)
@pytest.mark.parametrize("rs", [A.RIEMANN_APPROX, A.RIEMANN_LLF, A.RIEMANN_HLLC])
def test_strict_bitwise_3d(gpu, H, strip, path, rs):
    strip(4)
    p = hparams(3, (70, 13, 11), riemann=rs, precision=A.FP64_STRICT, kernel_path=path,
                ymax=1.0, zmax=1.0)
    compare_steps(H, p, problem(p, A.PROBLEM_BLAST), 4, tag="3d")


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("slope_type", [0.0, 1.0])
@pytest.mark.parametrize("bc", [A.BC_PERIODIC, A.BC_NEUMANN, "mixed"])
def test_strict_bitwise_slopes_and_boundaries(gpu, H, strip, dim, slope_type, bc):
    strip(8 if dim == 2 else 3)
    bcs = [1, 2, 3, 3, 2, 1] if bc == "mixed" else bc
    n = (67, 21) if dim == 2 else (63, 9, 10)
    p = hparams(dim, n, riemann=A.RIEMANN_HLLC, slope_type=slope_type, bc=bcs,
                precision=A.FP64_STRICT, kernel_path=A.PATH_FUSED, ymax=1.0, zmax=1.0)
    compare_steps(H, p, problem(p, A.PROBLEM_IMPLODE if dim == 2 else A.PROBLEM_BLAST), 3,
                  tag="slopes/bc")


@pytest.mark.parametrize("dim", [2, 3])
def test_strict_bitwise_tiny_grid(gpu, H, dim):
    n = (2, 3) if dim == 2 else (2, 2, 3)
    p = hparams(dim, n, riemann=A.RIEMANN_APPROX, precision=A.FP64_STRICT,
                kernel_path=A.PATH_FUSED, bc=A.BC_PERIODIC)
    compare_steps(H, p, problem(p, A.PROBLEM_IMPLODE), 3, tag="tiny")


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("path", PATHS)
def test_fast_fp64_within_tolerance(gpu, H, dim, path):
    n = (150, 77) if dim == 2 else (66, 20, 12)
    p = hparams(dim, n, riemann=A.RIEMANN_HLLC, precision=A.FP64, kernel_path=path,
                ymax=1.0, zmax=1.0)
    compare_steps(H, p, problem(p, A.PROBLEM_IMPLODE), 3, exact=False, tol=1e-12, tag="fast")


@pytest.mark.parametrize("dim", [2, 3])
def test_run_to_tend_matches_reference_time_loop(gpu, H, dim):
    """run() with the CFL dt: the last dt is clamped to tEnd, later steps are no-ops, the
    iteration count and the final state equal the oracle's (strict)."""
    n = (40, 30) if dim == 2 else (24, 16, 12)
    p = hparams(dim, n, riemann=A.RIEMANN_HLLC, precision=A.FP64_STRICT,
                kernel_path=A.PATH_FUSED, tEnd=0.01, ymax=1.0, zmax=1.0)
    pr = problem(p, A.PROBLEM_BLAST)
    o = H.Solver(p)
    o.init_problem(pr)
    steps = 0
    while o.time()[0] < p.tEnd - 1e-14:
        o.step(0, A.DT_CFL)
        steps += 1
    with SolverHydroMuscl(p, pr) as g:
        g.run(steps + 5)
        t, dt, it = g.time()
        assert it == steps and t == o.time()[0] and g.finished()
        np.testing.assert_array_equal(g.download(), o.get_state())


@pytest.mark.parametrize("dim", [2, 3])
def test_fused_equals_staged_large(gpu, strip, dim):
    """At the reference's own test sizes (settings/test_implode_2D.ini 1024 x 512,
    test_blast_3D.ini 64 x 48 x 32, and a larger 3D box) the fused kernel equals the staged one
    bit for bit (strict)."""
    strip(64)
    sizes = [(1024, 512)] if dim == 2 else [(64, 48, 32), (130, 70, 66)]
    for n in sizes:
        outs = []
        for path in PATHS:
            p = hparams(dim, n, riemann=A.RIEMANN_HLLC, precision=A.FP64_STRICT,
                        kernel_path=path, ymax=1.5 if dim == 3 else 1.0, zmax=1.0)
            with SolverHydroMuscl(p, problem(p, A.PROBLEM_IMPLODE if dim == 2
                                             else A.PROBLEM_BLAST)) as g:
                g.run(4)
                outs.append((g.download(), g.time()))
        assert outs[0][1] == outs[1][1]
        np.testing.assert_array_equal(outs[0][0], outs[1][0])


def test_solver_factory_keys(gpu):
    p = hparams(2, (32, 16))
    s = SolverFactory.Instance().create("Hydro_Muscl_2D", p)  # default problem: implode
    try:
        s.next_iteration()
        assert s.m_iteration == 1 and s.m_t > 0
    finally:
        s.close()


@pytest.mark.parametrize("name", ["hydro_implode2d_32x16", "hydro_blast3d_16x12x8"])
@pytest.mark.parametrize("path", PATHS)
@pytest.mark.parametrize("precision", [A.FP64_STRICT, A.FP64])
def test_golden_fixtures(gpu, name, path, precision):
    """The committed fixtures (made by tests/golden/make_golden_hydro.py from the oracle)."""
    from test_golden_cpu import load_case

    p, z = load_case(name)
    p.precision = precision
    p.kernel_path = path
    print(f"hydro golden {name} path={path} precision={precision}", flush=True)
    with SolverHydroMuscl(p) as g:
        g.upload(z["U0"])
        g.step(0, A.DT_CFL)
        U1 = g.download()
        for _ in range(9):
            g.step(0, A.DT_CFL)
        U10 = g.download()
        t10 = g.time()[:2]
    if precision == A.FP64_STRICT:
        np.testing.assert_array_equal(U1, z["U1"])
        np.testing.assert_array_equal(U10, z["U10"])
        assert tuple(t10) == tuple(z["time10"])
    else:
        for a, b, tol in ((U1, z["U1"], 1e-12), (U10, z["U10"], 1e-11)):
            assert np.max(np.abs(a - b)) <= tol * np.max(np.abs(b))


@pytest.mark.parametrize("dim", [2, 3])
def test_bench_size_walls_conserve_mass_and_energy(gpu, dim):
    """At the benchmark sizes (2D 8192 x 4096, 3D 512 x 384 x 256, the reference's hydro tests
    at 8x their resolution) reflecting walls keep total mass and energy to roundoff; the state
    stays physical (positive density and pressure)."""
    import bench

    p, pr = bench.hydro_workload(dim, 8, A.PATH_AUTO, A.FP64)
    print(f"hydro bench-size conservation dim={dim} n=({p.nx},{p.ny},{p.nz})", flush=True)
    g = 2
    ks = slice(g, g + p.nz) if dim == 3 else slice(0, 1)
    inner = (slice(None), ks, slice(g, g + p.ny), slice(g, g + p.nx))
    with SolverHydroMuscl(p, pr) as s:
        U = s.download()[inner]
        m0, e0 = U[0].sum(dtype=np.float64), U[1].sum(dtype=np.float64)
        del U
        s.run(5)
        U = s.download()[inner]
    assert abs(U[0].sum(dtype=np.float64) - m0) <= 1e-12 * m0
    assert abs(U[1].sum(dtype=np.float64) - e0) <= 1e-12 * e0
    assert U[0].min() > 0
    nv = U.shape[0]
    kin = 0.5 * sum(U[v] ** 2 for v in range(2, nv)) / U[0]
    assert (U[1] - kin).min() > 0


def test_save_vtk_matches_host_writer(gpu, tmp_path):
    """ppk_hydro_save_vtk writes the same bytes as the host writer on the downloaded state."""
    from ppkmhd_amd.solver import write_vtk

    p = hparams(3, (12, 10, 8), ymax=1.0, zmax=1.0)
    with SolverHydroMuscl(p, problem(p, A.PROBLEM_BLAST)) as g:
        g.run(3)
        g.save_vtk(tmp_path, "dev", 3)
        write_vtk(p, g.download(), tmp_path, "host", 3)
    assert (tmp_path / "dev_0000003.vti").read_bytes() == (tmp_path / "host_0000003.vti").read_bytes()


@pytest.mark.parametrize("case", ["kh2d_sine", "kh2d_rand", "kh3d_sine", "kh3d_rand", "gresho2d",
                                  "gresho3d", "isentropic2d"])
def test_strict_bitwise_more_problems(gpu, H, strip, case):
    """The other hydro problems of the reference without gravity: Kelvin-Helmholtz (Robertson
    sine or random layer), Gresho vortex, isentropic vortex; strict, bit for bit."""
    strip(8 if "2d" in case else 4)
    dim = 2 if "2d" in case else 3
    n = (70, 40) if dim == 2 else (66, 12, 20)
    kind = {"kh": A.PROBLEM_KELVIN_HELMHOLTZ, "gr": A.PROBLEM_GRESHO_VORTEX,
            "is": A.PROBLEM_ISENTROPIC_VORTEX}[case[:2]]
    box = dict(xmin=-0.5, xmax=0.5, ymin=-0.5, ymax=0.5, zmin=0.0, zmax=1.0) \
        if kind == A.PROBLEM_GRESHO_VORTEX else dict(xmax=10.0, ymax=10.0) \
        if kind == A.PROBLEM_ISENTROPIC_VORTEX else dict(xmax=1.0)
    p = make_hydro_params(dim, *n, gamma0=1.4, cfl=0.8, slope_type=2.0, riemann=A.RIEMANN_HLLC,
                          bc=A.BC_PERIODIC, tEnd=10.0, precision=A.FP64_STRICT,
                          kernel_path=A.PATH_FUSED, **box)
    pr = default_problem(kind, p)
    pr.kh_p_rand = 1 if case.endswith("rand") else 0
    compare_steps(H, p, pr, 3, tag=case, oracle_state=not case.endswith("rand"))


@pytest.mark.parametrize("config", [2, 5, 13])
@pytest.mark.parametrize("path", [A.PATH_FUSED, A.PATH_STAGED])
def test_strict_bitwise_four_quadrant(gpu, H, strip, config, path):
    """Four-quadrant 2D Riemann problems on settings/test_four_quadrant_2D.ini's settings
    (gamma 1.666, HLLC, slope type 2, absorbing walls; corner at (0.8, 0.8)) on a grid that
    leaves partial strips: configurations 3, 6 (the one with the reference's p = 0.5
    quadrant) and 14; strict, bit for bit over 4 steps."""
    strip(8)
    p = make_hydro_params(2, 90, 70, gamma0=1.666, cfl=0.8, slope_type=2.0,
                          riemann=A.RIEMANN_HLLC, bc=A.BC_NEUMANN, tEnd=10.0,
                          precision=A.FP64_STRICT, kernel_path=path)
    pr = default_problem(A.PROBLEM_FOUR_QUADRANT, p)
    pr.fq_config = config
    compare_steps(H, p, pr, 4, tag=f"four_quadrant config_number={config}")


def gravity_case(case, **kw):
    """Rayleigh-Taylor / rising bubble with an oblique uniform gravity (every component
    non-zero), periodic sides and reflecting walls across the vertical axis, the reference's
    settings/test_rayleigh_taylor_2d.ini scheme (approx solver, slope type 2, CFL 0.7)."""
    dim = 2 if "2d" in case else 3
    n = (70, 40) if dim == 2 else (66, 12, 20)
    bc = [A.BC_PERIODIC] * 6
    up = 1 if dim == 2 else 2
    bc[2 * up] = bc[2 * up + 1] = A.BC_DIRICHLET
    p = make_hydro_params(dim, *n, gamma0=1.4, cfl=0.7, slope_type=2.0, bc=bc, tEnd=10.0,
                          xmin=-0.5, xmax=0.5, ymin=-0.5, ymax=0.5, zmin=-0.5, zmax=0.5, **kw)
    p.gravity_static = 1
    if case.startswith("disk"):  # settings/test_muscl_disk_2d.ini: walls, point gravity
        p = make_hydro_params(dim, *n, gamma0=1.667, cfl=0.5, slope_type=1.0, bc=A.BC_NEUMANN,
                              riemann=A.RIEMANN_HLLC, tEnd=10.0, xmin=-0.5, xmax=0.5,
                              ymin=-0.5, ymax=0.5, zmin=-0.5, zmax=0.5, **kw)
        p.gravity_static = 1
        pr = default_problem(A.PROBLEM_DISK, p)
        pr.disk_radius, pr.disk_ref_density, pr.disk_ref_sound_speed = 0.2, 100.0, 0.316227766
        pr.disk_radius_inner = 0.02
        pr.disk_xc, pr.disk_yc = 0.03, -0.02
        pr.pg_x, pr.pg_y = 0.01, 0.0
        return p, pr
    kind = A.PROBLEM_RAYLEIGH_TAYLOR if case.startswith("rti") else A.PROBLEM_RISING_BUBBLE
    pr = default_problem(kind, p)
    g = (0.02, -0.1, 0.0) if dim == 2 else (0.02, 0.015, -0.1)
    pr.rti_gx, pr.rti_gy, pr.rti_gz = g
    pr.rb_gx, pr.rb_gy, pr.rb_gz = g
    pr.rb_x0, pr.rb_y0, pr.rb_z0, pr.rb_R = 0.1, -0.15, -0.2, 0.2
    pr.rti_amplitude = 0.05
    return p, pr


@pytest.mark.parametrize("path", PATHS)
@pytest.mark.parametrize("case", ["rti2d", "rti3d", "rb2d", "rb3d", "disk2d", "disk3d"])
def test_strict_bitwise_gravity(gpu, H, strip, path, case):
    """Gravity: the half-step predictor on every face state (each with its own cell's field),
    the source term after the update (2D with the new density floored, 3D without) and the
    gravity CFL dt; uniform fields and the disk's point mass; strict, bit for bit over 4 CFL
    steps (t and dt included), both kernel forms, partial strips."""
    strip(8 if "2d" in case else 4)
    p, pr = gravity_case(case, precision=A.FP64_STRICT, kernel_path=path)
    compare_steps(H, p, pr, 4, tag=case, oracle_state=not case.startswith("rb"))


@pytest.mark.parametrize("case", ["rti2d", "rb3d", "disk2d"])
def test_fast_gravity_within_tolerance(gpu, H, case):
    p, pr = gravity_case(case, kernel_path=A.PATH_FUSED)
    compare_steps(H, p, pr, 4, exact=False, tol=1e-12, tag=case,
                  oracle_state=not case.startswith("rb"))


def test_gravity_problems_need_static_gravity(gpu):
    """The reference writes the gravity field of rayleigh_taylor / rising_bubble into a view that
    exists only with [gravity] static (or point): without it the library refuses the problem."""
    p, pr = gravity_case("rb2d")
    p.gravity_static = 0
    with pytest.raises(Exception, match="gravity"):
        SolverHydroMuscl(p, pr)


def test_isentropic_vortex_second_order_at_scale(gpu):
    """The isentropic vortex moves with the ambient flow; on the GPU at 128^2 .. 512^2 (periodic
    [0,10]^2, t = 1) the density error against the translated initial field falls at second
    order."""
    from ppkmhd_amd.solver import hydro_initial_state

    errs = []
    for n in (128, 256, 512):
        p = make_hydro_params(2, n, n, xmax=10.0, ymax=10.0, gamma0=1.4, cfl=0.5,
                              slope_type=2.0, riemann=A.RIEMANN_HLLC, bc=A.BC_PERIODIC,
                              tEnd=1.0, n_step_max=100000, float_round=False)
        pr = default_problem(A.PROBLEM_ISENTROPIC_VORTEX, p, float_round=False)
        with SolverHydroMuscl(p, pr) as g:
            g.run(100000)
            assert g.finished() and g.time()[0] == p.tEnd
            rho = g.download()[0, 0, 2:2 + n, 2:2 + n]
        pt = default_problem(A.PROBLEM_ISENTROPIC_VORTEX, p, float_round=False)
        pt.iv_x += pt.iv_u_a * p.tEnd
        pt.iv_y += pt.iv_v_a * p.tEnd
        exact = hydro_initial_state(p, pt)[0, 0, 2:2 + n, 2:2 + n]
        errs.append(np.mean(np.abs(rho - exact)))
    orders = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    print(f"isentropic vortex L1 {errs}, orders {orders}", flush=True)
    assert np.all(orders > 1.8), (errs, orders)
