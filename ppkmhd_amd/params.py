"""Build solver and problem parameters the way the reference's .ini path does.

Every real read from an .ini by the reference goes through ``ConfigMap::getFloat``, which
parses with ``strtof`` and returns a ``float`` (reference src/utils/config/ConfigMap.cpp:37-46);
``HydroParams`` then widens it to double.  ``f32`` reproduces that rounding so parameters built
here equal those the C++ ``ppk_mhd_params_from_ini`` produces from the same .ini text.
"""
import math

import numpy as np

from . import _abi as A


def f32(x):
    """Round a Python float through IEEE single precision and back (strtof semantics)."""
    return float(np.float32(x))


def make_params(dim, nx, ny, nz=1, *, xmin=0.0, xmax=1.0, ymin=0.0, ymax=1.0, zmin=0.0,
                zmax=1.0, gamma0=1.4, cfl=0.5, slope_type=1.0, smallr=1e-10, smallc=1e-10,
                riemann=A.RIEMANN_HLLD, bc=A.BC_PERIODIC, tEnd=0.0, n_step_max=1000,
                precision=A.FP64, kernel_path=A.PATH_AUTO, float_round=True):
    """ppk_mhd_params as HydroParams::setup + init would build them.

    Defaults are the reference's (src/shared/HydroParams.cpp:28-217).  ``bc`` is one value for
    all faces or a sequence of 6.  With ``float_round`` every real goes through float32 first
    (the .ini semantics); pass False to inject exact doubles on both sides of a comparison.
    """
    r = f32 if float_round else float
    p = A.Params()
    p.dim = dim
    p.nx, p.ny, p.nz = nx, ny, (nz if dim == 3 else 1)
    p.ghost_width = 3  # MHD, src/shared/HydroParams.cpp:56-71
    p.riemann = riemann
    bcs = [bc] * 6 if np.isscalar(bc) else list(bc)
    for f in range(6):
        p.bc[f] = int(bcs[f])
    p.n_step_max = n_step_max
    p.precision = precision
    p.kernel_path = kernel_path
    p.implementation_version = 0
    p.xmin, p.xmax, p.ymin, p.ymax, p.zmin, p.zmax = (r(v) for v in (xmin, xmax, ymin, ymax,
                                                                      zmin, zmax))
    p.gamma0 = r(gamma0)
    p.cfl = r(cfl)
    p.slope_type = r(slope_type)
    p.smallr = r(smallr)
    p.smallc = r(smallc)
    p.cIso = 0.0  # never read from the config (src/shared/HydroParams.h:58)
    p.tEnd = r(tEnd)
    init_params(p)
    return p


def make_hydro_params(dim, nx, ny, nz=1, *, riemann=A.RIEMANN_APPROX, **kw):
    """Parameters of the Euler solvers Hydro_Muscl_2D/3D: ghostWidth 2 and the reference's
    default Riemann solver "approx" (src/shared/HydroParams.cpp:42-55, 145-160)."""
    p = make_params(dim, nx, ny, nz, riemann=riemann, **kw)
    p.ghost_width = 2
    return p


def init_params(p):
    """Derived fields, HydroParams::init (src/shared/HydroParams.cpp:421-453)."""
    p.dx = (p.xmax - p.xmin) / p.nx
    p.dy = (p.ymax - p.ymin) / p.ny
    p.dz = (p.zmax - p.zmin) / p.nz
    p.smallp = p.smallc * p.smallc / p.gamma0
    p.smallpp = p.smallr * p.smallp
    p.gamma6 = (p.gamma0 + 1.0) / (2.0 * p.gamma0)
    return p


def default_problem(kind, params=None, float_round=True):
    """ppk_mhd_problem with the reference's per-problem defaults
    (src/shared/problems/{OrszagTang,Blast,KH,Wave,Implode}Params.h)."""
    r = f32 if float_round else float
    pr = A.Problem()
    pr.kind = kind
    pr.ot_kt = 0.0
    pr.ot_b0_scale = 1.0
    xmin = params.xmin if params is not None else 0.0
    xmax = params.xmax if params is not None else 1.0
    ymin = params.ymin if params is not None else 0.0
    ymax = params.ymax if params is not None else 1.0
    zmin = params.zmin if params is not None else 0.0
    zmax = params.zmax if params is not None else 1.0
    # BlastParams defaults are computed in double then passed as float defaults (:31-45)
    pr.blast_radius = r((xmin + xmax) / 2.0 / 10)
    pr.blast_center_x = r((xmin + xmax) / 2)
    pr.blast_center_y = r((ymin + ymax) / 2)
    pr.blast_center_z = r((zmin + zmax) / 2)
    pr.blast_density_in = r(1.0)
    pr.blast_density_out = r(1.2)
    pr.blast_pressure_in = r(10.0)
    pr.blast_pressure_out = r(0.1)
    pr.blast_bx = pr.blast_by = pr.blast_bz = 0.5  # hard-coded in the reference
    # KHParams (:60-104)
    pr.kh_d_in = r(1.0)
    pr.kh_d_out = r(1.0)
    pr.kh_pressure = r(10.0)
    pr.kh_p_sine = 0
    pr.kh_p_sine_rob = 1
    pr.kh_p_rand = 0
    pr.kh_vflow_in = r(-0.5)
    pr.kh_vflow_out = r(0.5)
    pr.kh_seed = 12
    pr.kh_amplitude = r(0.1)
    pr.kh_mode = 2
    pr.kh_w0 = r(0.1)
    pr.kh_delta = r(0.03)
    pr.kh_axis = 2
    pr.kh_bx = pr.kh_by = pr.kh_bz = 0.0
    # WaveParams (:31-146)
    pr.wave_type = 0
    pr.wave_amplitude = r(1.0e-6)
    pr.wave_gamma0 = r(1.66667)
    pr.wave_lx, pr.wave_ly, pr.wave_lz = 3.0 - 0.0, 1.5 - 0.0, 1.5 - 0.0
    # ImplodeParams (:40-58)
    pr.implode_shape = 0
    for v, x in enumerate((1.0, 1.0, 0, 0, 0, 0, 0, 0)):
        pr.implode_out[v] = r(x)
    for v, x in enumerate((0.125, 0.14, 0, 0, 0, 0, 0, 0)):
        pr.implode_in[v] = r(x)
    # GreshoParams, IsentropicVortexParams (centre defaults computed in double, passed as
    # float defaults)
    pr.gresho_rho0 = r(1.0)
    pr.gresho_ma = r(0.1)
    pr.iv_rho_a = pr.iv_p_a = pr.iv_t_a = r(1.0)
    pr.iv_u_a = pr.iv_v_a = pr.iv_w_a = r(1.0)
    pr.iv_x = r((xmin + xmax) / 2)
    pr.iv_y = r((ymin + ymax) / 2)
    pr.iv_z = r((zmin + zmax) / 2)
    pr.iv_beta = r(5.0)
    pr.fq_config = 0
    pr.fq_x = r(0.8)
    pr.fq_y = r(0.8)
    # RayleighTaylorInstabilityParams, RisingBubbleParams
    pr.rti_amplitude, pr.rti_d0, pr.rti_d1 = r(0.01), r(1.0), r(2.0)
    pr.rti_gx, pr.rti_gy, pr.rti_gz = r(0.0), r(-0.1), r(0.0)
    pr.rb_R, pr.rb_x0, pr.rb_y0, pr.rb_z0 = r(0.1), r(0.5), r(0.25), r(0.5)
    pr.rb_din, pr.rb_dout = r(1.0), r(2.0)
    pr.rb_gx, pr.rb_gy, pr.rb_gz = r(0.0), r(-0.1), r(0.0)
    # DiskParams (its own mesh defaults -0.5 .. 0.5), PointSourceGravity
    dx0, dx1 = (xmin, xmax) if params is not None else (-0.5, 0.5)
    dy0, dy1 = (ymin, ymax) if params is not None else (-0.5, 0.5)
    dz0, dz1 = (zmin, zmax) if params is not None else (-0.5, 0.5)
    pr.disk_radius = r((dx1 - dx0) * 0.125)
    pr.disk_radius_inner = r(pr.disk_radius / 10)
    pr.disk_xc, pr.disk_yc, pr.disk_zc = r((dx0 + dx1) / 2), r((dy0 + dy1) / 2), r((dz0 + dz1) / 2)
    pr.disk_ref_density, pr.disk_contrast_density = r(1.0), r(100.0)
    pr.disk_contrast_width, pr.disk_ref_sound_speed = r(0.01), r(0.2)
    pr.pg_x = pr.pg_y = pr.pg_z = r(0.0)
    pr.pg_gm = r(1.0) * r(1.0)
    pr.pg_eps = r(0.01)
    # RotorParams, FieldLoopParams
    pr.rotor_r0, pr.rotor_r1, pr.rotor_u0, pr.rotor_p0 = r(0.1), r(0.115), r(2.0), r(1.0)
    pr.rotor_b0 = r(5.0 / math.sqrt(4 * math.pi))
    pr.loop_radius, pr.loop_density_in, pr.loop_amplitude, pr.loop_vflow = r(1.0), r(1.0), r(1.0), r(1.0)
    return pr
