"""ctypes mirror of ``include/kapsel_gpu.h`` and loader of the HIP library.

The library ``icep_kapsel_amd/libkapsel_gpu.so`` is the product: a C ABI over
hand-written gfx950 kernels.  This module only declares its types and
prototypes; there is no fallback when the library is missing -- loading raises.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libkapsel_gpu.so")

KAPSEL_OK = 0
KAPSEL_FP64 = 0
KAPSEL_FP32 = 1

FIELDS = [
    "ZETA0", "ZETA1", "U0", "U1", "U2", "UP0", "UP1", "UP2", "F0", "F1", "F2",
    "PHI", "PHI_SUM", "EPSILON", "POTENTIAL", "RHO_FREE", "RHS",
    "WORK0", "WORK1", "WORK2", "WORK3", "WORK4",
    "CONC0", "CONC1",
]
FIELD = {name: i for i, name in enumerate(FIELDS)}
NFIELDS = len(FIELDS)

PFIELDS = [
    "X", "X_PREVIOUS", "X_NOPBC", "V", "V_OLD", "OMEGA", "OMEGA_OLD",
    "F_HYDRO", "F_HYDRO_PREVIOUS", "FR", "FR_PREVIOUS",
    "TORQUE_HYDRO", "TORQUE_HYDRO_PREVIOUS", "TORQUE_R", "TORQUE_R_PREVIOUS",
]
PFIELD = {name: i for i, name in enumerate(PFIELDS)}
P_NFIELDS = len(PFIELDS)

PRECOND_NONE = 0
PRECOND_SPECTRAL = 1

KERNEL_CLASSES = ["FFT_Z_R2C", "FFT_Z_C2R", "FFT_LINES_FWD", "FFT_LINES_INV", "POINTWISE", "PARTICLE",
                  "NETWORK", "REFINE64", "FFT_PLANE_FACE", "FFT_X_GRAD", "FFT_X_DIV"]
KCLASS = {name: i for i, name in enumerate(KERNEL_CLASSES)}


class Params(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("dx", ctypes.c_double),
        ("rho", ctypes.c_double), ("eta", ctypes.c_double),
        ("n_particles", ctypes.c_int),
        ("radius", ctypes.c_double), ("xi", ctypes.c_double),
        ("mass_ratio", ctypes.c_double),
        ("rotation", ctypes.c_int),
        ("lj_truncate", ctypes.c_int), ("lj_powers", ctypes.c_int),
        ("lj_epsilon", ctypes.c_double), ("lj_sigma", ctypes.c_double),
        ("body_force", ctypes.c_double * 3),
        ("fix_cell", ctypes.c_int * 3),
        ("precision", ctypes.c_int),
        ("eps_fluid", ctypes.c_double), ("eps_particle", ctypes.c_double),
        ("dielectric_cst", ctypes.c_double),
    ]


class Time(ctypes.Structure):
    _fields_ = [
        ("ts", ctypes.c_int), ("time", ctypes.c_double),
        ("dt_fluid", ctypes.c_double), ("hdt_fluid", ctypes.c_double),
        ("dt_md", ctypes.c_double), ("hdt_md", ctypes.c_double),
    ]


class Electrolyte(ctypes.Structure):
    _fields_ = [
        ("kbt", ctypes.c_double), ("elementary_charge", ctypes.c_double),
        ("dielectric_cst", ctypes.c_double), ("surface_charge", ctypes.c_double),
        ("valency", ctypes.c_double * 2), ("onsager_coeff", ctypes.c_double * 2),
        ("debye_length", ctypes.c_double), ("e_ext", ctypes.c_double * 3),
        ("ac", ctypes.c_int), ("angular_frequency", ctypes.c_double),
        ("dielectric", ctypes.c_int),
    ]


def make_electrolyte(kbt=1.0, elementary_charge=1.0, dielectric_cst=1.0, surface_charge=100.0,
                     valency=(1.0, -1.0), onsager_coeff=(7.24, 7.24), debye_length=10.0,
                     e_ext=(0.0, 0.0, 0.0), ac=0, angular_frequency=0.0, dielectric=0):
    """Defaults are the salt section of the reference's UDF/input.udf:53-100."""
    e = Electrolyte()
    e.kbt, e.elementary_charge, e.dielectric_cst = kbt, elementary_charge, dielectric_cst
    e.surface_charge = surface_charge
    e.valency[:] = list(valency)
    e.onsager_coeff[:] = list(onsager_coeff)
    e.debye_length = debye_length
    e.e_ext[:] = list(e_ext)
    e.ac, e.angular_frequency = ac, angular_frequency
    e.dielectric = dielectric
    return e


class RestartData(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("n_particles", ctypes.c_int), ("n_spec", ctypes.c_int),
        ("ts", ctypes.c_int), ("time", ctypes.c_double), ("uk_dc", ctypes.c_double * 3),
        ("zeta0", ctypes.POINTER(ctypes.c_double)), ("zeta1", ctypes.POINTER(ctypes.c_double)),
        ("particles", ctypes.POINTER(ctypes.c_double)), ("orientations", ctypes.POINTER(ctypes.c_double)),
        ("conc0", ctypes.POINTER(ctypes.c_double)), ("conc1", ctypes.POINTER(ctypes.c_double)),
    ]


def make_time(dt, ts=0, time=0.0):
    """CTime as main() initialises it (sp_3d_ns.cxx:412)."""
    return Time(ts, time, dt, 0.5 * dt, dt, 0.5 * dt)


def make_params(n, n_particles=0, dx=1.0, rho=1.0, eta=1.0, radius=5.0, xi=2.0,
                mass_ratio=1.0, rotation=1, lj_truncate=1, lj_powers=0,
                lj_epsilon=1.0, lj_sigma=None, body_force=(0.0, 0.0, 0.0),
                fix_cell=(0, 0, 0), precision=KAPSEL_FP64, eps_fluid=1.0,
                eps_particle=1.0, dielectric_cst=1.0):
    if isinstance(n, int):
        n = (n, n, n)
    if lj_sigma is None:
        lj_sigma = 2.0 * radius  # SIGMA = 2 RADIUS (input.cxx:1404)
    p = Params()
    p.nx, p.ny, p.nz = n
    p.dx, p.rho, p.eta = dx, rho, eta
    p.n_particles = n_particles
    p.radius, p.xi, p.mass_ratio = radius, xi, mass_ratio
    p.rotation = rotation
    p.lj_truncate, p.lj_powers = lj_truncate, lj_powers
    p.lj_epsilon, p.lj_sigma = lj_epsilon, lj_sigma
    p.body_force[:] = list(body_force)
    p.fix_cell[:] = list(fix_cell)
    p.precision = precision
    p.eps_fluid, p.eps_particle, p.dielectric_cst = eps_fluid, eps_particle, dielectric_cst
    return p


_VP = ctypes.c_void_p
_DP = ctypes.POINTER(ctypes.c_double)
_IP = ctypes.POINTER(ctypes.c_int)
_I = ctypes.c_int
_D = ctypes.c_double

# name -> argtypes (after ctx) ; all return int
_PROTOS = {
    "kapsel_destroy": [],
    "kapsel_synchronize": [],
    "kapsel_get_sekibun": [_IP, _IP],
    "kapsel_set_field": [_I, _DP],
    "kapsel_get_field": [_I, _DP],
    "kapsel_set_uk_dc": [_DP],
    "kapsel_get_uk_dc": [_DP],
    "kapsel_set_particles": [_DP],
    "kapsel_get_particles": [_DP],
    "kapsel_set_orientations": [_DP],
    "kapsel_get_orientations": [_DP],
    "kapsel_copy_field": [_I, _I],
    "kapsel_fft_r2c": [_I],
    "kapsel_fft_c2r": [_I],
    "kapsel_truncate": [_I],
    "kapsel_zeta_k2u_k": [],
    "kapsel_u_k2zeta_k": [],
    "kapsel_solenoidal_u": [_I],
    "kapsel_ns_slaved_euler": [ctypes.POINTER(Time)],
    "kapsel_make_phi_particle_sum": [],
    "kapsel_make_u_particle_sum": [],
    "kapsel_make_phi_particle": [_I],
    "kapsel_calc_f_hydro": [ctypes.POINTER(Time)],
    "kapsel_force": [],
    "kapsel_md_position": [ctypes.POINTER(Time)],
    "kapsel_md_velocity": [ctypes.POINTER(Time)],
    "kapsel_make_f_particle_dt_sole": [],
    "kapsel_init_rigidity": [],
    "kapsel_spm_step": [ctypes.POINTER(Time)],
    "kapsel_dielectric_field": [],
    "kapsel_poisson_apply": [_I, _I],
    "kapsel_poisson_apply_host": [_DP, _DP],
    "kapsel_poisson_rhs": [_DP],
    "kapsel_poisson_solve": [_DP, _D, _I, _I, _IP, _DP],
    "kapsel_set_dielectric": [_D, _D, _D],
    "kapsel_maxwell_force": [_DP],
    "kapsel_icep_step": [ctypes.POINTER(Time), _DP, _D, _I, _I, _IP, _DP],
    "kapsel_set_concurrency": [_I],
    "kapsel_set_thermostat": [_D, _D, _D, ctypes.c_ulong],
    "kapsel_thermostat": [ctypes.POINTER(Time)],
    "kapsel_save_restart": [ctypes.c_char_p, ctypes.POINTER(Time), ctypes.c_char_p],
    "kapsel_load_restart": [ctypes.c_char_p, ctypes.POINTER(Time)],
    "kapsel_set_electrolyte": [ctypes.POINTER(Electrolyte)],
    "kapsel_init_rho_ion": [],
    "kapsel_init_rho_ion_pb": [ctypes.POINTER(Time), _I, _IP, _DP],
    "kapsel_free_energy_pb": [ctypes.POINTER(Time), _DP],
    "kapsel_get_total_solute": [_DP],
    "kapsel_set_total_solute": [_DP],
    "kapsel_make_phi_qq": [_I, _I],
    "kapsel_make_surface_normal": [_I],
    "kapsel_charge_field": [_I],
    "kapsel_nssolute_euler": [ctypes.POINTER(Time)],
    "kapsel_rescale_solute": [_DP],
    "kapsel_coulomb_force": [ctypes.POINTER(Time)],
    "kapsel_electrolyte_step": [ctypes.POINTER(Time)],
    "kapsel_init_rho_ion_dielectric": [],
    "kapsel_dielectric_potential": [ctypes.POINTER(Time), _D, _I, _I, _IP, _DP],
    "kapsel_icep_electrolyte_step": [ctypes.POINTER(Time), _D, _I, _I, _IP, _DP],
    "kapsel_pnet_init": [ctypes.POINTER(ctypes.c_long)],
    "kapsel_pnet_set_params": [ctypes.POINTER(ctypes.c_float)],
    "kapsel_pnet_get_params": [ctypes.POINTER(ctypes.c_float)],
    "kapsel_pnet_loss_grad": [_D, _DP, _D, _DP, ctypes.POINTER(ctypes.c_float)],
    "kapsel_pnet_train_step": [_D, _DP, _D, _D, _DP],
    "kapsel_pnet_warm_start": [_D, _DP, _D],
    "kapsel_poisson_galerkin_start": [_DP, _DP],
    "kapsel_kernel_timing": [_I],
    "kapsel_kernel_timing_get": [_I, _DP, _IP, _DP],
    "kapsel_kernel_timing_every": [_I],
}

_lib = None


def load_library(path=None):
    """Load the HIP library; raises OSError when it is not built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    path = path or LIB_PATH
    if not os.path.exists(path):
        raise OSError(f"HIP library not built: {path} (run __graft_entry__.build())")
    lib = ctypes.CDLL(path)
    lib.kapsel_last_error.restype = ctypes.c_char_p
    lib.kapsel_last_error.argtypes = []
    lib.kapsel_create.restype = _I
    lib.kapsel_create.argtypes = [ctypes.POINTER(Params), ctypes.POINTER(_VP)]
    for name, args in _PROTOS.items():
        fn = getattr(lib, name)
        fn.restype = _I
        fn.argtypes = [_VP] + args
    # host-only entry points (no context)
    lib.kapsel_restart_write.restype = _I
    lib.kapsel_restart_write.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(RestartData)]
    lib.kapsel_restart_read.restype = _I
    lib.kapsel_restart_read.argtypes = [ctypes.c_char_p, ctypes.POINTER(RestartData)]
    # device-pointer entry point with no context (allocates nothing)
    lib.kapsel_calc_f_hydro_noalloc.restype = _I
    lib.kapsel_calc_f_hydro_noalloc.argtypes = [ctypes.POINTER(Params), ctypes.POINTER(Time), _VP, _VP, _VP, _VP,
                                                _VP, _VP]
    if path == LIB_PATH:
        _lib = lib
    return lib


def declared_symbols(header=None):
    """Names of the functions declared in include/kapsel_gpu.h."""
    import re
    header = header or os.path.join(os.path.dirname(_HERE), "include", "kapsel_gpu.h")
    text = open(header).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(kapsel_\w+)\s*\(", text, re.M)))
