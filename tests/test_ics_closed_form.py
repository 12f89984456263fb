"""The oracle's initial conditions against closed forms written from the problem statements
(CPU only).  The host initial states of the library equal the oracle's bit for bit
(test_ics_cpu.py), so these checks pin both: a wrong constant (B0 = 1/sqrt(2 pi) for
1/sqrt(4 pi), p0 = gamma/(2 pi)), a swapped sign, a missing factor in the energy, or a shear
layer on the wrong side fails here even though both restatements would agree with each other.

  Orszag-Tang 2D   src/muscl/MHDInitFunctors2D.h:317-343: rho = 25/(36 pi), p = 5/(12 pi) for
                   gamma = 5/3, v = (-sin 2 pi y, sin 2 pi x), B = B0 (-sin 2 pi y, sin 4 pi x),
                   B0 = 1/sqrt(4 pi), E = p/(gamma-1) + rho v^2/2 + B^2/2
  blast 3D         src/muscl/MHDInitFunctors3D.h:155-259 with the uniform field as a parameter:
                   rho 1, p 10 inside r < 0.1 and 0.1 outside, v = 0, B = (1,1,1)/sqrt(3)
  Kelvin-Helmholtz src/muscl/MHDInitFunctors3D.h:420-613 (random branch, shear layer normal to
  3D, y-normal      y): rho 2 for |y| <= 1/4, 1 outside, vx = -/+ 0.5 with uniform perturbations
                   of amplitude 0.01 on vx and vy, vz = 0, p 2.5, B = (0.1, 0, 0), gamma 1.4
  implode 2D / 3D  src/muscl/MHDInitFunctors{2D,3D}.h: outer state where x + y (+ z) exceeds
                   the diagonal through the domain centre, the velocity stored as given in the
                   momentum slots (the reference's functor does not multiply by the density)
"""
import math

import numpy as np
import pytest

from ppkmhd_amd import _abi as A
from ppkmhd_amd import default_problem, make_params

ID, IP, IU, IV, IW, IA, IB, IC = range(8)


def centres(p):
    """Cell-centre coordinates of every padded cell, arrays shaped (ksize, jsize, isize)."""
    g = p.ghost_width
    i = np.arange(p.isize)
    j = np.arange(p.jsize)
    k = np.arange(p.ksize)
    x = p.xmin + p.dx / 2 + (i - g) * p.dx
    y = p.ymin + p.dy / 2 + (j - g) * p.dy
    z = p.zmin + p.dz / 2 + (k - g) * p.dz if p.dim == 3 else np.zeros(1)
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    return X, Y, Z


def close(a, b, rtol=2e-14, atol=1e-15):
    np.testing.assert_allclose(a, b, rtol=rtol, atol=atol)


@pytest.mark.parametrize("n", [16, 40])
def test_orszag_tang_closed_form(oracle, n):
    gamma = 5.0 / 3.0
    p = make_params(2, n, n, gamma0=gamma, cfl=0.8, tEnd=0.5, float_round=False)
    U = oracle.init_problem_array(p, default_problem(A.PROBLEM_ORSZAG_TANG, p))
    X, Y, _ = centres(p)
    rho = 25.0 / (36.0 * math.pi)      # config 0's density
    p0 = 5.0 / (12.0 * math.pi)        # config 0's pressure
    B0 = 1.0 / math.sqrt(4.0 * math.pi)
    vx, vy = -np.sin(2 * math.pi * Y), np.sin(2 * math.pi * X)
    bx, by = -B0 * np.sin(2 * math.pi * Y), B0 * np.sin(4 * math.pi * X)
    close(U[ID], rho)
    close(U[IU], rho * vx)
    close(U[IV], rho * vy)
    assert (U[IW] == 0).all() and (U[IC] == 0).all()
    close(U[IA], bx)
    close(U[IB], by)
    # energy on the cells whose +1 faces exist (the functor's range, :365); face fields that
    # depend only on the transverse coordinate average to the centre value exactly
    E = p0 / (gamma - 1) + 0.5 * rho * (vx**2 + vy**2) + 0.5 * (bx**2 + by**2)
    close(U[IP][:, :-1, :-1], E[:, :-1, :-1])
    # derived: the sound and Alfven speeds of the classic problem, c^2 = gamma p / rho = 1
    assert gamma * p0 / rho == pytest.approx(1.0, rel=1e-15)


def test_orszag_tang_b0_scale_extension(oracle):
    """The ensemble's field scaling multiplies B0 only (the density, pressure and velocity stay
    the problem's)."""
    gamma = 1.7
    p = make_params(2, 16, 16, gamma0=gamma, float_round=False)
    pr = default_problem(A.PROBLEM_ORSZAG_TANG, p)
    base = oracle.init_problem_array(p, pr)
    pr.ot_b0_scale = 1.75
    U = oracle.init_problem_array(p, pr)
    close(U[IA], 1.75 * base[IA])
    close(U[IB], 1.75 * base[IB])
    for v in (ID, IU, IV):
        assert np.array_equal(U[v], base[v])
    close(U[ID], gamma * gamma / (4 * math.pi))


def test_blast_3d_closed_form(oracle):
    gamma = 5.0 / 3.0
    n = 20
    p = make_params(3, n, n, n, gamma0=gamma, float_round=False)
    pr = default_problem(A.PROBLEM_BLAST, p)
    pr.blast_radius = 0.1
    pr.blast_center_x = pr.blast_center_y = pr.blast_center_z = 0.5
    pr.blast_density_in = pr.blast_density_out = 1.0
    pr.blast_pressure_in, pr.blast_pressure_out = 10.0, 0.1
    b = 1.0 / math.sqrt(3.0)
    pr.blast_bx = pr.blast_by = pr.blast_bz = b
    U = oracle.init_problem_array(p, pr)
    X, Y, Z = centres(p)
    inside = (X - 0.5) ** 2 + (Y - 0.5) ** 2 + (Z - 0.5) ** 2 < 0.1**2
    # 20^3 cells of 0.05, centres at odd multiples of 0.025 from the blast centre: the 8 at
    # (1,1,1) x 0.025 (r = 0.043) and the 24 at permutations of (1,1,3) x 0.025 (r = 0.083)
    assert inside.sum() == 32
    pres = np.where(inside, 10.0, 0.1)
    close(U[ID], 1.0)
    for v in (IU, IV, IW):
        assert (U[v] == 0).all()
    for v in (IA, IB, IC):
        close(U[v], b)
    close(U[IP], pres / (gamma - 1) + 0.5)  # |B|^2 / 2 = 1/2
    assert np.array_equal(U[IP] > 10.0, inside)


def test_kelvin_helmholtz_3d_y_layer_closed_form(oracle):
    gamma = 1.4
    n = 16
    p = make_params(3, n, n, n, xmin=-0.5, xmax=0.5, ymin=-0.5, ymax=0.5, zmin=-0.5, zmax=0.5,
                    gamma0=gamma, float_round=False)
    pr = default_problem(A.PROBLEM_KELVIN_HELMHOLTZ, p)
    pr.kh_p_rand, pr.kh_p_sine, pr.kh_p_sine_rob = 1, 0, 0
    pr.kh_axis = 1
    pr.kh_d_in, pr.kh_d_out = 2.0, 1.0
    pr.kh_vflow_in, pr.kh_vflow_out = -0.5, 0.5
    pr.kh_pressure = 2.5
    pr.kh_amplitude = 0.01
    pr.kh_seed = 3
    pr.kh_bx, pr.kh_by, pr.kh_bz = 0.1, 0.0, 0.0
    U = oracle.init_problem_array(p, pr)
    _, Y, _ = centres(p)
    layer = np.abs(Y) <= 0.25
    rho = np.where(layer, 2.0, 1.0)
    assert np.array_equal(U[ID], rho)
    u = U[IU] / rho
    v = U[IV] / rho
    du = u - np.where(layer, -0.5, 0.5)
    assert np.abs(du).max() <= 0.005 and np.abs(v).max() <= 0.005
    # uniform draws: spread over the whole amplitude, mean near 0, the two draws independent
    assert np.abs(du).max() > 0.0045 and np.abs(v).max() > 0.0045
    assert abs(du.mean()) < 5e-4 and abs(v.mean()) < 5e-4
    assert abs(np.corrcoef(du.ravel(), v.ravel())[0, 1]) < 0.1
    assert (U[IW] == 0).all()
    close(U[IA], 0.1)
    assert (U[IB] == 0).all() and (U[IC] == 0).all()
    close(U[IP], 2.5 / (gamma - 1) + 0.5 * rho * (u**2 + v**2) + 0.5 * 0.01)
    # the seed selects the field: another seed, other perturbations, same layers
    pr.kh_seed = 4
    U2 = oracle.init_problem_array(p, pr)
    assert np.array_equal(U2[ID], U[ID]) and not np.array_equal(U2[IU], U[IU])


@pytest.mark.parametrize("dim", [2, 3])
def test_implode_closed_form(oracle, dim):
    gamma = 1.4
    n = (18, 14, 10)[:dim] if dim == 3 else (18, 14)
    p = make_params(dim, *n, gamma0=gamma, float_round=False)
    pr = default_problem(A.PROBLEM_IMPLODE, p)
    out = (1.0, 1.0, 0.3, -0.2, 0.1, 0.25, 0.5, 0.75)
    inn = (0.125, 0.14, -0.1, 0.05, 0.2, 0.6, -0.3, 0.1)
    for v in range(8):
        pr.implode_out[v], pr.implode_in[v] = out[v], inn[v]
    U = oracle.init_problem_array(p, pr)
    X, Y, Z = centres(p)
    if dim == 2:
        outer = X + Y > 0.5 + 0.0
    else:
        outer = X + Y + Z > 0.5 + 0.0 + 0.0
    assert 0 < outer.sum() < outer.size

    def pick(v):
        return np.where(outer, out[v], inn[v])

    nv = 3 if dim == 3 else 2
    close(U[ID], pick(0))
    for m in range(nv):  # velocities as given, not rho * v
        close(U[IU + m], pick(2 + m))
        close(U[IA + m], pick(5 + m))
    if dim == 2:
        assert (U[IW] == 0).all() and (U[IC] == 0).all()
    kin = sum(pick(2 + m) ** 2 for m in range(nv))
    mag = sum(pick(5 + m) ** 2 for m in range(nv))
    close(U[IP], pick(1) / (gamma - 1) + 0.5 * pick(0) * kin + 0.5 * mag)
