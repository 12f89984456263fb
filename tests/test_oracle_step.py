"""Pin the oracle's step-level functions against invariants of the mathematics (CPU only).

  fill_boundaries      textbook ghost fill (periodic / Neumann / Dirichlet, x then y then z)
  primitives           closed form of the conservative -> primitive map
  inv_dt               closed-form CFL sum on uniform states
  full step (2D, 3D)   uniform state preserved; conservation of totals and of the face-field
                       sums with periodic boundaries; div B kept at roundoff (CT); transpose
                       (2D), cyclic-permutation (3D) and mirror symmetry; second-order
                       convergence on linear MHD waves (the reference's convergence study)
"""
import numpy as np
import pytest

import textbook as T
from ppkmhd_amd import _abi as A
from ppkmhd_amd import default_problem, make_params


def rand_state(p, rng, smooth=True):
    """Random physical state over the padded box (density, pressure > 0)."""
    U = np.zeros(p.shape)
    sh = p.shape[1:]
    if smooth:
        k, j, i = np.meshgrid(*[np.arange(s) for s in sh], indexing="ij")
        def f(a, b, c, ph):
            return np.sin(2 * np.pi * (a * i / p.isize + b * j / p.jsize + c * k / max(p.ksize, 1)) + ph)
        r = 1.0 + 0.3 * f(1, 2, 1, 0.3)
        pr = 1.0 + 0.2 * f(2, 1, 1, 1.1)
        vel = [0.3 * f(1, 1, 2, 0.7 + m) for m in range(3)]
        B = [0.5 + 0.3 * f(2, 1, 1, 2.1 + m) for m in range(3)]
    else:
        r = rng.uniform(0.5, 2.0, sh)
        pr = rng.uniform(0.5, 2.0, sh)
        vel = [rng.normal(0, 0.3, sh) for _ in range(3)]
        B = [rng.normal(0, 0.5, sh) for _ in range(3)]
    U[0] = r
    for m in range(3):
        U[2 + m] = r * vel[m]
        U[5 + m] = B[m]
    U[1] = pr / (p.gamma0 - 1) + 0.5 * r * sum(v * v for v in vel) + 0.5 * sum(b * b for b in B)
    return U


def solenoidal_state(p, rng):
    """Smooth state whose face-centred field is the discrete curl of a random periodic edge
    vector potential (discrete div B = 0 up to roundoff, also across the periodic seam)."""
    U = rand_state(p, rng)
    nz = p.nz if p.dim == 3 else 1
    sh = (nz, p.ny, p.nx)
    Ax, Ay, Az = (rng.normal(0, 0.05, sh) for _ in range(3))
    def d(a, axis):  # forward periodic difference
        return np.roll(a, -1, axis=axis) - a
    if p.dim == 3:
        bx = d(Az, 1) / p.dy - d(Ay, 0) / p.dz
        by = d(Ax, 0) / p.dz - d(Az, 2) / p.dx
        bz = d(Ay, 2) / p.dx - d(Ax, 1) / p.dy
    else:
        bx = d(Az, 1) / p.dy
        by = -d(Az, 2) / p.dx
        bz = np.zeros(sh)
    ii = p.interior()
    for c, b in zip((A.IA, A.IB, A.IC), (bx, by, bz)):
        U[(c,) + ii[1:]] = b
    r = U[A.ID]
    v = U[A.IU:A.IW + 1] / r
    U[A.IP] = 1.0 / (p.gamma0 - 1) + 0.5 * r * (v ** 2).sum(0) + 0.5 * (U[A.IA:] ** 2).sum(0)
    return U


def divb(p, U):
    g = p.gw
    I = slice(g, g + p.nx)
    J = slice(g, g + p.ny)
    if p.dim == 3:
        K = slice(g, g + p.nz)
        d = ((U[A.IA, K, J, g + 1:g + p.nx + 1] - U[A.IA, K, J, I]) / p.dx +
             (U[A.IB, K, g + 1:g + p.ny + 1, I] - U[A.IB, K, J, I]) / p.dy +
             (U[A.IC, g + 1:g + p.nz + 1, J, I] - U[A.IC, K, J, I]) / p.dz)
    else:
        d = ((U[A.IA, 0, J, g + 1:g + p.nx + 1] - U[A.IA, 0, J, I]) / p.dx +
             (U[A.IB, 0, g + 1:g + p.ny + 1, I] - U[A.IB, 0, J, I]) / p.dy)
    return float(np.abs(d).max())


# ---------------------------------------------------------------------------------------------
@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("bcs", [(3,) * 6, (1, 2, 3, 3, 2, 1), (2, 1, 1, 2, 3, 3), (1,) * 6])
def test_boundaries_textbook(oracle, dim, bcs):
    rng = np.random.default_rng(1)
    p = make_params(dim, 7, 5, 6 if dim == 3 else 1, bc=list(bcs))
    U = rng.normal(size=p.shape)
    got = oracle.fill_boundaries(p, U.copy())
    ref = T.fill_ghosts_textbook(U, p.nx, p.ny, p.nz, p.gw, list(bcs), dim)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("dim", [2, 3])
def test_primitives_closed_form(oracle, dim):
    rng = np.random.default_rng(2)
    p = make_params(dim, 6, 5, 4 if dim == 3 else 1, gamma0=1.4)
    U = rand_state(p, rng, smooth=False)
    Q = oracle.primitives(p, U)
    n = p.shape[1:]
    sl = tuple(slice(0, s - 1) if s > 1 else slice(0, 1) for s in n)
    r = np.maximum(U[A.ID], p.smallr)
    bx = 0.5 * (U[A.IA] + np.roll(U[A.IA], -1, axis=2))
    by = 0.5 * (U[A.IB] + np.roll(U[A.IB], -1, axis=1))
    bz = 0.5 * (U[A.IC] + (np.roll(U[A.IC], -1, axis=0) if dim == 3 else 0.0))
    v = U[A.IU:A.IW + 1] / r
    eint = (U[A.IP] - 0.5 * (bx ** 2 + by ** 2 + bz ** 2)) / r - 0.5 * (v ** 2).sum(0)
    pres = np.maximum((p.gamma0 - 1) * r * eint, r * p.smallp)
    ref = np.array([r, pres, *v, bx, by, bz])
    np.testing.assert_allclose(Q[(slice(None),) + sl], ref[(slice(None),) + sl], rtol=1e-13, atol=1e-14)


@pytest.mark.parametrize("dim", [2, 3])
def test_inv_dt_closed_form(oracle, dim):
    rng = np.random.default_rng(3)
    p = make_params(dim, 8, 6, 5 if dim == 3 else 1, gamma0=5 / 3, xmax=2.0, ymax=1.0, zmax=0.5)
    q = T.random_state(rng)
    U = np.zeros(p.shape)
    r, pr, u, v, w, bx, by, bz = q
    U[0] = r
    U[2:5] = np.array([r * u, r * v, r * w])[:, None, None, None]
    U[5:8] = np.array([bx, by, bz])[:, None, None, None]
    U[1] = pr / (p.gamma0 - 1) + 0.5 * r * (u * u + v * v + w * w) + 0.5 * (bx * bx + by * by + bz * bz)
    Q = oracle.primitives(p, U)
    got = oracle.inv_dt(p, Q)
    dxs = (p.dx, p.dy, p.dz)
    qc = q.copy()
    if dim == 2:
        # the 2D primitive averages the cell-centred Bz with 0 (MHDRunFunctors2D.h:139)
        # and the pressure recovered from the energy then carries the difference
        qc[T.IC] *= 0.5
        qc[T.IP] = pr + (p.gamma0 - 1) * 0.5 * (bz * bz - qc[T.IC] ** 2)
    ref = sum((T.fast_speed(qc, p.gamma0, d) + abs(q[2 + d])) / dxs[d] for d in range(dim))
    assert got == pytest.approx(ref, rel=1e-13)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("slope_type", [1.0, 2.0])
def test_uniform_state_preserved(oracle, dim, slope_type):
    p = make_params(dim, 8, 8, 8 if dim == 3 else 1, gamma0=5 / 3, slope_type=slope_type,
                    tEnd=1.0)
    q = np.array([1.3, 0.7, 0.4, -0.3, 0.2, 0.6, -0.5, 0.3])
    U = np.zeros(p.shape)
    U[0] = q[0]
    U[2:5] = (q[0] * q[2:5])[:, None, None, None]
    U[5:8] = q[5:8, None, None, None]
    U[1] = q[1] / (p.gamma0 - 1) + 0.5 * q[0] * (q[2:5] @ q[2:5]) + 0.5 * (q[5:8] @ q[5:8])
    s = oracle.Solver(p)
    s.set_state(U)
    for _ in range(3):
        s.step(0.01, A.DT_FIXED)
    V = s.get_state()
    ii = p.interior()
    np.testing.assert_allclose(V[ii], U[ii], rtol=2e-14, atol=1e-14)


@pytest.mark.parametrize("dim", [2, 3])
def test_conservation_and_divb_periodic(oracle, dim):
    rng = np.random.default_rng(4)
    n3 = 12 if dim == 3 else 1
    p = make_params(dim, 16, 12, n3, gamma0=1.4, slope_type=2.0, tEnd=10.0, cfl=0.5)
    U = solenoidal_state(p, rng)
    U = oracle.fill_boundaries(p, U)
    d0 = divb(p, U)
    assert d0 < 1e-12
    s = oracle.Solver(p)
    s.set_state(U)
    ii = p.interior()
    tot0 = U[ii].sum(axis=(1, 2, 3))
    for _ in range(8):
        s.step(0, A.DT_CFL)
    V = oracle.fill_boundaries(p, s.get_state())
    tot = V[ii].sum(axis=(1, 2, 3))
    # mass, energy, momentum and the face-field sums telescope with periodic boundaries
    np.testing.assert_allclose(tot, tot0, rtol=1e-12, atol=1e-11)
    assert divb(p, V) < 1e-11 * max(1.0, np.abs(V[A.IA:]).max() / p.dx)


def test_transpose_symmetry_2d(oracle):
    rng = np.random.default_rng(5)
    p = make_params(2, 12, 12, gamma0=5 / 3, slope_type=2.0, tEnd=1.0)
    U = oracle.fill_boundaries(p, rand_state(p, rng))
    def tr(X):
        Y = X.transpose(0, 1, 3, 2).copy()
        Y[[A.IU, A.IV]] = Y[[A.IV, A.IU]]
        Y[[A.IA, A.IB]] = Y[[A.IB, A.IA]]
        return Y
    a, b = oracle.Solver(p), oracle.Solver(p)
    a.set_state(U)
    b.set_state(tr(U))
    for _ in range(3):
        a.step(2e-3, A.DT_FIXED)
        b.step(2e-3, A.DT_FIXED)
    ii = p.interior()
    np.testing.assert_allclose(b.get_state()[ii], tr(a.get_state())[ii], rtol=1e-12, atol=1e-13)


def test_cyclic_symmetry_3d(oracle):
    rng = np.random.default_rng(6)
    p = make_params(3, 8, 8, 8, gamma0=5 / 3, slope_type=1.0, tEnd=1.0)
    U = oracle.fill_boundaries(p, rand_state(p, rng))
    def perm(X):
        Y = np.transpose(X, (0, 3, 1, 2)).copy()
        Y[[A.IU, A.IV, A.IW]] = Y[[A.IV, A.IW, A.IU]]
        Y[[A.IA, A.IB, A.IC]] = Y[[A.IB, A.IC, A.IA]]
        return Y
    a, b = oracle.Solver(p), oracle.Solver(p)
    a.set_state(U)
    b.set_state(perm(U))
    for _ in range(2):
        a.step(2e-3, A.DT_FIXED)
        b.step(2e-3, A.DT_FIXED)
    ii = p.interior()
    np.testing.assert_allclose(b.get_state()[ii], perm(a.get_state())[ii], rtol=1e-12, atol=1e-13)


@pytest.mark.parametrize("dim", [2, 3])
def test_mirror_symmetry_x(oracle, dim):
    """x -> -x on a periodic box: cells reverse, left faces become right faces, u and Bx flip."""
    rng = np.random.default_rng(7)
    p = make_params(dim, 10, 8, 6 if dim == 3 else 1, gamma0=1.4, slope_type=2.0, tEnd=1.0)
    U = oracle.fill_boundaries(p, rand_state(p, rng))
    g, n = p.gw, p.nx
    def mir(X):
        Y = X.copy()
        ii = np.arange(g, g + n)
        src = g + n - 1 - (ii - g)
        for v in range(8):
            if v == A.IA:
                Y[v][..., ii] = -X[v][..., src + 1]  # left face of the image = right face
            else:
                Y[v][..., ii] = X[v][..., src] * (-1 if v == A.IU else 1)
        return oracle.fill_boundaries(p, Y)
    a, b = oracle.Solver(p), oracle.Solver(p)
    a.set_state(U)
    b.set_state(mir(U))
    for _ in range(3):
        a.step(2e-3, A.DT_FIXED)
        b.step(2e-3, A.DT_FIXED)
    ii = p.interior()
    np.testing.assert_allclose(b.get_state()[ii], mir(oracle.fill_boundaries(p, a.get_state()))[ii],
                               rtol=1e-12, atol=1e-13)


def _wave_l1(oracle, nx, wave_type, tend):
    """One period of a linear MHD wave as in the reference's convergence study
    (test/convergence/mhd_wave_3D.ini.skel, plot-results.py:73-82)."""
    p = make_params(3, nx, nx // 2, nx // 2, xmax=3.0, ymax=1.5, zmax=1.5, gamma0=1.6666666667,
                    cfl=0.4, slope_type=2.0, smallr=1e-8, smallc=1e-8, tEnd=tend,
                    n_step_max=10000)
    pr = default_problem(A.PROBLEM_WAVE, p)
    pr.wave_type = wave_type
    pr.wave_gamma0 = p.gamma0
    pr.wave_lx, pr.wave_ly, pr.wave_lz = p.xmax - p.xmin, p.ymax - p.ymin, p.zmax - p.zmin
    s = oracle.Solver(p)
    s.init_problem(pr)
    U0 = s.get_state()
    while not s.finished():
        s.step(0, A.DT_CFL)
    U1 = s.get_state()
    ii = p.interior()
    d = np.abs(U1[ii] - U0[ii]).reshape(8, -1).mean(axis=1)
    return float(np.sqrt((d ** 2).sum()))


# one period of each wave, the periods of test/convergence/fulltest.sh:34 (fast, Alfven,
# slow, and the entropy wave advected at V0 = 1)
@pytest.mark.parametrize("wave_type,tend", [(0, 0.5), (1, 1.0), (2, 2.0), (3, 1.0)])
def test_linear_wave_second_order(oracle, wave_type, tend):
    e16 = _wave_l1(oracle, 16, wave_type, tend)
    e32 = _wave_l1(oracle, 32, wave_type, tend)
    order = np.log2(e16 / e32)
    assert order > 1.7, (e16, e32, order)


def test_fixed_dt_is_clamped_to_tend(oracle):
    """A fixed dt is clamped to tEnd - t as the CFL dt is (SolverBase.cpp:173-177): four steps
    of 0.003 towards tEnd = 0.01 take 0.003, 0.003, 0.003, 0.001."""
    p = make_params(2, 16, 16, gamma0=5 / 3, cfl=0.8, slope_type=1.0, tEnd=0.01)
    o = oracle.Solver(p)
    o.init_problem(default_problem(A.PROBLEM_ORSZAG_TANG, p))
    for _ in range(3):
        o.step(0.003, A.DT_FIXED)
    assert not o.finished()
    o.step(0.003, A.DT_FIXED)
    t, dt, it = o.time()
    # tEnd holds the float-rounded 0.01 of an .ini real (ConfigMap.cpp:37-46)
    assert it == 4 and dt == pytest.approx(p.tEnd - 0.009, abs=1e-15) and t == p.tEnd
    assert o.finished()


@pytest.mark.parametrize("n", [8, 12])
def test_counted_flops_match_bench_constants(oracle, n):
    """The instrumented oracle's operation counts (add + mul + div + sqrt) of one CFL step of
    the 3D blast, per stage, are the per-item figures bench.py prices the rooflines with, over
    the reference functors' item ranges: primitives of [0, size-1)^3 (28 each) and the CFL
    term of the interior (42 each), E and face-field slopes over (n+4)^3, the trace over
    (n+3)^3, fluxes and EMFs over (n+1)^3, the update over n^3.  Two grid sizes separate the
    per-cell figures of stages whose ranges differ."""
    import bench

    p = make_params(3, n, n, n, gamma0=5 / 3, cfl=0.8, tEnd=1.0)
    pr = default_problem(A.PROBLEM_BLAST, p)
    pr.blast_radius = 0.1
    pr.blast_density_out = 1.0
    pr.blast_pressure_in, pr.blast_pressure_out = 10.0, 0.1
    pr.blast_bx = pr.blast_by = pr.blast_bz = 1.0 / np.sqrt(3.0)
    s = oracle.Solver(p, counting=True)
    s.init_problem(pr)
    s.step(0, A.DT_CFL)
    oracle.op_counts_reset()
    s.step(0, A.DT_CFL)
    for stage in ("primitives_dt", "elec_field", "mag_slopes", "trace", "fluxes", "emf",
                  "update_ct"):
        c = oracle.op_counts(stage)
        counted = c["add"] + c["mul"] + c["div"] + c["sqrt"]
        assert counted == pytest.approx(bench.stage_flops(stage, n), rel=1e-12), stage


@pytest.mark.parametrize("nx,ny", [(16, 16), (24, 24), (20, 12)])
def test_counted_flops_2d_match_bench_constants(oracle, nx, ny):
    """The same for one CFL step of the 2D Orszag-Tang (the batched ensemble's workload): the
    instrumented oracle's operation counts per stage equal bench.py's 2D per-item figures over
    their (nx + k) x (ny + k) ranges (a non-square grid separates the two extents); the E and
    face-field slope stages are part of the 2D trace."""
    import bench

    p = make_params(2, nx, ny, gamma0=5 / 3, cfl=0.8, tEnd=1.0)
    pr = default_problem(A.PROBLEM_ORSZAG_TANG, p)
    s = oracle.Solver(p, counting=True)
    s.init_problem(pr)
    s.step(0, A.DT_CFL)
    oracle.op_counts_reset()
    s.step(0, A.DT_CFL)

    def counted(stage):
        c = oracle.op_counts(stage)
        return c["add"] + c["mul"] + c["div"] + c["sqrt"]

    F = bench.FLOP_ITEMS_2D
    items = lambda key: F[key][0] * (nx + F[key][1]) * (ny + F[key][1])
    assert counted("primitives_dt") == items("primitives") + items("cfl")
    assert counted("elec_field") == 0 and counted("mag_slopes") == 0
    assert counted("trace") == items("trace")
    assert counted("fluxes") == items("fluxes")
    assert counted("emf") == items("emf")
    assert counted("update_ct") == items("update_ct")
    total = sum(counted(k) for k in ("primitives_dt", "trace", "fluxes", "emf", "update_ct"))
    assert total == bench.step_flops_2d(nx, ny)
