"""ctypes mirror of the plain-data structs in include/ppkmhd_amd.h.

The field order and types must match the C header exactly; tests compare ``ctypes.sizeof``
against ``ppk_mhd_params_size()`` / ``ppk_mhd_problem_size()`` exported by the library.
"""
import ctypes as C

# variable indexes, reference src/shared/enums.h:17-34
ID, IP, IU, IV, IW, IA, IB, IC = range(8)
VAR_NAMES = ("rho", "energy", "rho_vx", "rho_vy", "rho_vz", "bx", "by", "bz")

# enum values, reference src/shared/enums.h:141-159
RIEMANN_APPROX, RIEMANN_LLF, RIEMANN_HLL, RIEMANN_HLLC, RIEMANN_HLLD = range(5)
BC_UNDEFINED, BC_DIRICHLET, BC_NEUMANN, BC_PERIODIC = range(4)
FP64, FP32_BULK, FP64_STRICT = 0, 1, 2
PATH_AUTO, PATH_STAGED, PATH_FUSED, PATH_FUSED_RECORDS = 0, 1, 2, 3
DT_FIXED, DT_CFL = 0, 1
(PROBLEM_ORSZAG_TANG, PROBLEM_BLAST, PROBLEM_KELVIN_HELMHOLTZ, PROBLEM_WAVE, PROBLEM_IMPLODE,
 PROBLEM_GRESHO_VORTEX, PROBLEM_ISENTROPIC_VORTEX, PROBLEM_FOUR_QUADRANT, PROBLEM_RAYLEIGH_TAYLOR,
 PROBLEM_RISING_BUBBLE, PROBLEM_DISK, PROBLEM_ROTOR, PROBLEM_FIELD_LOOP) = range(13)

PPK_OK = 0
PPK_ERR_INVALID_ARGUMENT, PPK_ERR_HIP, PPK_ERR_OUT_OF_MEMORY, PPK_ERR_UNSUPPORTED, PPK_ERR_CONFIG = 1, 2, 3, 4, 5


class Params(C.Structure):
    """ppk_mhd_params (mirror of the HydroParams fields the MHD MUSCL path reads)."""

    _fields_ = [
        ("dim", C.c_int32),
        ("nx", C.c_int32),
        ("ny", C.c_int32),
        ("nz", C.c_int32),
        ("ghost_width", C.c_int32),
        ("riemann", C.c_int32),
        ("bc", C.c_int32 * 6),
        ("n_step_max", C.c_int32),
        ("precision", C.c_int32),
        ("kernel_path", C.c_int32),
        ("implementation_version", C.c_int32),
        ("xmin", C.c_double),
        ("xmax", C.c_double),
        ("ymin", C.c_double),
        ("ymax", C.c_double),
        ("zmin", C.c_double),
        ("zmax", C.c_double),
        ("dx", C.c_double),
        ("dy", C.c_double),
        ("dz", C.c_double),
        ("gamma0", C.c_double),
        ("gamma6", C.c_double),
        ("cfl", C.c_double),
        ("slope_type", C.c_double),
        ("smallr", C.c_double),
        ("smallc", C.c_double),
        ("smallp", C.c_double),
        ("smallpp", C.c_double),
        ("cIso", C.c_double),
        ("tEnd", C.c_double),
        ("gravity_static", C.c_int32),
        ("reserved2", C.c_int32),
    ]

    @property
    def gw(self):
        return self.ghost_width

    @property
    def isize(self):
        return self.nx + 2 * self.ghost_width

    @property
    def jsize(self):
        return self.ny + 2 * self.ghost_width

    @property
    def ksize(self):
        return self.nz + 2 * self.ghost_width if self.dim == 3 else 1

    @property
    def nbvar(self):
        """8 for MHD (ghost width 3); 4 (2D) or 5 (3D) for the hydro solvers (ghost width 2),
        as HydroParams::setup assigns them (src/shared/HydroParams.cpp:42-71)."""
        if self.ghost_width == 2:
            return 5 if self.dim == 3 else 4
        return 8

    @property
    def shape(self):
        """numpy shape of a state array in the reference layout: (nbvar, ksize, jsize, isize)."""
        return (self.nbvar, self.ksize, self.jsize, self.isize)

    @property
    def count(self):
        return self.nbvar * self.ksize * self.jsize * self.isize

    def interior(self):
        """numpy index of the interior cells of a (8, ksize, jsize, isize) array."""
        g = self.ghost_width
        ks = slice(g, g + self.nz) if self.dim == 3 else slice(0, 1)
        return (slice(None), ks, slice(g, g + self.ny), slice(g, g + self.nx))


class Problem(C.Structure):
    """ppk_mhd_problem (initial-condition parameters)."""

    _fields_ = [
        ("kind", C.c_int32),
        ("kh_mode", C.c_int32),
        ("kh_p_sine", C.c_int32),
        ("kh_p_sine_rob", C.c_int32),
        ("kh_p_rand", C.c_int32),
        ("kh_seed", C.c_int32),
        ("kh_axis", C.c_int32),
        ("wave_type", C.c_int32),
        ("implode_shape", C.c_int32),
        ("pad", C.c_int32),
        ("ot_kt", C.c_double),
        ("ot_b0_scale", C.c_double),
        ("blast_radius", C.c_double),
        ("blast_center_x", C.c_double),
        ("blast_center_y", C.c_double),
        ("blast_center_z", C.c_double),
        ("blast_density_in", C.c_double),
        ("blast_density_out", C.c_double),
        ("blast_pressure_in", C.c_double),
        ("blast_pressure_out", C.c_double),
        ("blast_bx", C.c_double),
        ("blast_by", C.c_double),
        ("blast_bz", C.c_double),
        ("kh_d_in", C.c_double),
        ("kh_d_out", C.c_double),
        ("kh_pressure", C.c_double),
        ("kh_vflow_in", C.c_double),
        ("kh_vflow_out", C.c_double),
        ("kh_amplitude", C.c_double),
        ("kh_w0", C.c_double),
        ("kh_delta", C.c_double),
        ("kh_bx", C.c_double),
        ("kh_by", C.c_double),
        ("kh_bz", C.c_double),
        ("wave_amplitude", C.c_double),
        ("wave_gamma0", C.c_double),
        ("wave_lx", C.c_double),
        ("wave_ly", C.c_double),
        ("wave_lz", C.c_double),
        ("implode_in", C.c_double * 8),
        ("implode_out", C.c_double * 8),
        ("gresho_rho0", C.c_double),
        ("gresho_ma", C.c_double),
        ("iv_rho_a", C.c_double),
        ("iv_p_a", C.c_double),
        ("iv_t_a", C.c_double),
        ("iv_u_a", C.c_double),
        ("iv_v_a", C.c_double),
        ("iv_w_a", C.c_double),
        ("iv_x", C.c_double),
        ("iv_y", C.c_double),
        ("iv_z", C.c_double),
        ("iv_beta", C.c_double),
        ("fq_config", C.c_int),
        ("fq_reserved", C.c_int),
        ("fq_x", C.c_double),
        ("fq_y", C.c_double),
        ("rti_amplitude", C.c_double),
        ("rti_d0", C.c_double),
        ("rti_d1", C.c_double),
        ("rti_gx", C.c_double),
        ("rti_gy", C.c_double),
        ("rti_gz", C.c_double),
        ("rb_R", C.c_double),
        ("rb_x0", C.c_double),
        ("rb_y0", C.c_double),
        ("rb_z0", C.c_double),
        ("rb_din", C.c_double),
        ("rb_dout", C.c_double),
        ("rb_gx", C.c_double),
        ("rb_gy", C.c_double),
        ("rb_gz", C.c_double),
        ("disk_radius", C.c_double),
        ("disk_radius_inner", C.c_double),
        ("disk_xc", C.c_double),
        ("disk_yc", C.c_double),
        ("disk_zc", C.c_double),
        ("disk_ref_density", C.c_double),
        ("disk_contrast_density", C.c_double),
        ("disk_contrast_width", C.c_double),
        ("disk_ref_sound_speed", C.c_double),
        ("pg_x", C.c_double),
        ("pg_y", C.c_double),
        ("pg_z", C.c_double),
        ("pg_gm", C.c_double),
        ("pg_eps", C.c_double),
        ("rotor_r0", C.c_double),
        ("rotor_r1", C.c_double),
        ("rotor_u0", C.c_double),
        ("rotor_p0", C.c_double),
        ("rotor_b0", C.c_double),
        ("loop_radius", C.c_double),
        ("loop_density_in", C.c_double),
        ("loop_amplitude", C.c_double),
        ("loop_vflow", C.c_double),
    ]
