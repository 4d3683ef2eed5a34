"""ctypes front end of the CPU FP64 oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py.  The method names mirror icep_kapsel_amd.KapselGPU
so a parity test drives both with the same calls.
"""
import ctypes
import os

import numpy as np

from icep_kapsel_amd import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(_HERE, "liboracle.so")
REF_LIB = os.path.join(_HERE, "_ref", "libooura_ref.so")

_lib = None


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise OSError(f"oracle not built: {LIB} (make -C oracle)")
        lib = ctypes.CDLL(LIB)
        vp, dp, ip, i = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int), ctypes.c_int
        tp = ctypes.POINTER(abi.Time)
        lib.oracle_create.restype = vp
        lib.oracle_create.argtypes = [ctypes.POINTER(abi.Params)]
        lib.oracle_destroy.argtypes = [vp]
        lib.oracle_destroy.restype = None
        lib.oracle_last_error.restype = ctypes.c_char_p
        lib.oracle_num_threads.restype = i
        lib.oracle_set_num_threads.argtypes = [i]
        lib.oracle_set_num_threads.restype = None
        protos = {
            "oracle_get_sekibun": [ip, ip],
            "oracle_set_field": [i, dp], "oracle_get_field": [i, dp],
            "oracle_set_uk_dc": [dp], "oracle_get_uk_dc": [dp],
            "oracle_set_particles": [dp], "oracle_get_particles": [dp],
            "oracle_fft_r2c": [i], "oracle_fft_c2r": [i], "oracle_truncate": [i],
            "oracle_zeta_k2u_k": [], "oracle_u_k2zeta_k": [], "oracle_u2advection_k": [],
            "oracle_solenoidal_u": [i], "oracle_ns_slaved_euler": [tp],
            "oracle_make_phi_particle_sum": [], "oracle_make_u_particle_sum": [],
            "oracle_make_phi_particle": [i], "oracle_make_phi_sin": [i],
            "oracle_calc_f_hydro": [tp], "oracle_force": [],
            "oracle_md_position": [tp], "oracle_md_velocity": [tp],
            "oracle_make_f_particle_dt_sole": [], "oracle_init_rigidity": [], "oracle_set_plant": [i],
            "oracle_spm_step": [tp],
            "oracle_make_dielectric_field": [], "oracle_poisson_apply": [i, i],
            "oracle_ax_unpadded": [dp, dp], "oracle_poisson_rhs": [dp],
            "oracle_poisson_solve": [dp, ctypes.c_double, i, i, ip, dp],
            "oracle_maxwell_force": [dp], "oracle_add_force_to_fluid": [tp],
            "oracle_icep_step": [tp, dp, ctypes.c_double, i, i, ip, dp],
            "oracle_set_orientations": [dp], "oracle_get_orientations": [dp],
            "oracle_set_thermostat": [ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_ulong],
            "oracle_thermostat": [tp], "oracle_mt_real3": [i, dp],
            "oracle_set_electrolyte": [ctypes.POINTER(abi.Electrolyte)], "oracle_init_rho_ion": [],
            "oracle_get_total_solute": [dp], "oracle_set_total_solute": [dp],
            "oracle_make_phi_qq": [i, i], "oracle_make_surface_normal": [i], "oracle_charge_field": [i],
            "oracle_nssolute_euler": [tp], "oracle_rescale_solute": [dp], "oracle_coulomb_force": [tp],
            "oracle_electrolyte_step": [tp], "oracle_free_energy_pb": [tp, dp],
            "oracle_init_rho_ion_pb": [tp, i, ip, dp], "oracle_init_rho_ion_dielectric": [], "oracle_dielectric_grad": [tp],
            "oracle_dielectric_potential": [tp, ctypes.c_double, i, i, ip, dp],
            "oracle_icep_electrolyte_step": [tp, ctypes.c_double, i, i, ip, dp],
        }
        for name, args in protos.items():
            fn = getattr(lib, name)
            fn.restype = i
            fn.argtypes = [vp] + args
        _lib = lib
    return _lib


def set_threads(n):
    _load().oracle_set_num_threads(int(n))


def num_threads():
    return _load().oracle_num_threads()


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Oracle:
    """One parameter set per instance (the reference latches statics per process)."""

    def __init__(self, params):
        self.lib = _load()
        self.params = params
        self.h = self.lib.oracle_create(ctypes.byref(params))
        if not self.h:
            raise ValueError(self.lib.oracle_last_error().decode())
        self.nx, self.ny, self.nz = params.nx, params.ny, params.nz
        self.nz_ = 2 * (self.nz // 2 + 1)
        self.shape = (self.nx, self.ny, self.nz_)
        self.np = params.n_particles

    def close(self):
        if self.h:
            self.lib.oracle_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _call(self, name, *args):
        rc = getattr(self.lib, "oracle_" + name)(self.h, *args)
        if rc != 0:
            raise RuntimeError(f"oracle_{name} failed")

    # fields -----------------------------------------------------------
    def set_field(self, name, a):
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(self.shape)
        self._call("set_field", abi.FIELD[name], _dp(a))

    def get_field(self, name):
        a = np.empty(self.shape, dtype=np.float64)
        self._call("get_field", abi.FIELD[name], _dp(a))
        return a

    def set_uk_dc(self, v):
        a = np.ascontiguousarray(v, dtype=np.float64)
        self._call("set_uk_dc", _dp(a))

    def get_uk_dc(self):
        a = np.zeros(3)
        self._call("get_uk_dc", _dp(a))
        return a

    def set_particles(self, state):
        a = np.ascontiguousarray(state, dtype=np.float64).reshape(abi.P_NFIELDS, 3, self.np)
        self._call("set_particles", _dp(a))

    def get_particles(self):
        a = np.zeros((abi.P_NFIELDS, 3, self.np))
        self._call("get_particles", _dp(a))
        return a

    def set_thermostat(self, kbt, alpha_v, alpha_o, seed=0):
        self._call("set_thermostat", float(kbt), float(alpha_v), float(alpha_o), int(seed))

    def thermostat(self, t):
        self._call("thermostat", ctypes.byref(t))

    def mt_real3(self, n):
        a = np.zeros(int(n))
        self._call("mt_real3", int(n), _dp(a))
        return a

    def set_orientations(self, q):
        """q: (8, Np) -- q0..q3 then q_old0..q_old3 (scalar part first)."""
        a = np.ascontiguousarray(q, dtype=np.float64).reshape(8, self.np)
        self._call("set_orientations", _dp(a))

    def get_orientations(self):
        a = np.zeros((8, self.np))
        self._call("get_orientations", _dp(a))
        return a

    def sekibun(self):
        n = ctypes.c_int(0)
        self._call("get_sekibun", ctypes.byref(n), None)
        cells = np.zeros((n.value, 3), dtype=np.int32)
        self._call("get_sekibun", ctypes.byref(n), cells.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
        return cells

    # operations --------------------------------------------------------
    def fft_r2c(self, name):
        self._call("fft_r2c", abi.FIELD[name])

    def fft_c2r(self, name):
        self._call("fft_c2r", abi.FIELD[name])

    def truncate(self, name):
        self._call("truncate", abi.FIELD[name])

    def zeta_k2u_k(self):
        self._call("zeta_k2u_k")

    def u_k2zeta_k(self):
        self._call("u_k2zeta_k")

    def u2advection_k(self):
        self._call("u2advection_k")

    def solenoidal_u(self, name0):
        self._call("solenoidal_u", abi.FIELD[name0])

    def ns_slaved_euler(self, t):
        self._call("ns_slaved_euler", ctypes.byref(t))

    def make_phi_particle_sum(self):
        self._call("make_phi_particle_sum")

    def make_u_particle_sum(self):
        self._call("make_u_particle_sum")

    def make_phi_particle(self, name):
        self._call("make_phi_particle", abi.FIELD[name])

    def make_phi_sin(self, name):
        self._call("make_phi_sin", abi.FIELD[name])

    def calc_f_hydro(self, t):
        self._call("calc_f_hydro", ctypes.byref(t))

    def force(self):
        self._call("force")

    def md_position(self, t):
        self._call("md_position", ctypes.byref(t))

    def md_velocity(self, t):
        self._call("md_velocity", ctypes.byref(t))

    def make_f_particle_dt_sole(self):
        self._call("make_f_particle_dt_sole")

    def init_rigidity(self):
        self._call("init_rigidity")

    def set_plant(self, code):
        """Plants a known mistake in the step (0 removes it); see Oracle::plant."""
        self._call("set_plant", int(code))

    def spm_step(self, t):
        self._call("spm_step", ctypes.byref(t))

    # dielectric Poisson ---------------------------------------------------
    def make_dielectric_field(self):
        self._call("make_dielectric_field")

    def poisson_apply(self, x_name, ax_name):
        self._call("poisson_apply", abi.FIELD[x_name], abi.FIELD[ax_name])

    def ax_unpadded(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        out = np.zeros_like(x)
        self._call("ax_unpadded", _dp(x), _dp(out))
        return out

    def poisson_rhs(self, e0):
        e = np.ascontiguousarray(e0, dtype=np.float64)
        self._call("poisson_rhs", _dp(e))

    def maxwell_force(self, e0):
        e = np.ascontiguousarray(e0, dtype=np.float64)
        self._call("maxwell_force", _dp(e))

    def add_force_to_fluid(self, t):
        self._call("add_force_to_fluid", ctypes.byref(t))

    def icep_step(self, t, e0, tol, maxit, solver=0):
        e = np.ascontiguousarray(e0, dtype=np.float64)
        it = ctypes.c_int(0)
        rr = ctypes.c_double(0.0)
        self._call("icep_step", ctypes.byref(t), _dp(e), tol, int(maxit), int(solver), ctypes.byref(it),
                   ctypes.byref(rr))
        return it.value, rr.value

    def poisson_solve(self, e0, tol, maxit, solver=0):
        """solver 0 = CG, 1 = spectral-preconditioned CG, 2 = BiCGSTAB; returns (iters, relres)."""
        e = np.ascontiguousarray(e0, dtype=np.float64)
        it = ctypes.c_int(0)
        rr = ctypes.c_double(0.0)
        self._call("poisson_solve", _dp(e), tol, int(maxit), int(solver), ctypes.byref(it), ctypes.byref(rr))
        return it.value, rr.value


    # electrolyte -------------------------------------------------------
    def set_electrolyte(self, el):
        self._call("set_electrolyte", ctypes.byref(el))

    def init_rho_ion(self):
        self._call("init_rho_ion")

    def get_total_solute(self):
        a = np.zeros(2)
        self._call("get_total_solute", _dp(a))
        return a

    def set_total_solute(self, v):
        a = np.ascontiguousarray(v, dtype=np.float64)
        self._call("set_total_solute", _dp(a))

    def make_phi_qq(self, phi_name, surface_name):
        self._call("make_phi_qq", abi.FIELD[phi_name], abi.FIELD[surface_name])

    def make_surface_normal(self, name0):
        self._call("make_surface_normal", abi.FIELD[name0])

    def charge_field(self, name):
        self._call("charge_field", abi.FIELD[name])

    def nssolute_euler(self, t):
        self._call("nssolute_euler", ctypes.byref(t))

    def rescale_solute(self):
        a = np.zeros(2)
        self._call("rescale_solute", _dp(a))
        return a

    def coulomb_force(self, t):
        self._call("coulomb_force", ctypes.byref(t))

    def electrolyte_step(self, t):
        self._call("electrolyte_step", ctypes.byref(t))

    def init_rho_ion_dielectric(self):
        self._call("init_rho_ion_dielectric")

    def dielectric_grad(self, t):
        self._call("dielectric_grad", ctypes.byref(t))

    def dielectric_potential(self, t, tol, maxit, solver=0):
        it = ctypes.c_int(0)
        rr = ctypes.c_double(0.0)
        self._call("dielectric_potential", ctypes.byref(t), tol, int(maxit), int(solver), ctypes.byref(it),
                   ctypes.byref(rr))
        return it.value, rr.value

    def icep_electrolyte_step(self, t, tol, maxit, solver=0):
        it = ctypes.c_int(0)
        rr = ctypes.c_double(0.0)
        self._call("icep_electrolyte_step", ctypes.byref(t), tol, int(maxit), int(solver), ctypes.byref(it),
                   ctypes.byref(rr))
        return it.value, rr.value

    def free_energy_pb(self, t):
        fe = np.zeros(3)
        self._call("free_energy_pb", ctypes.byref(t), _dp(fe))
        return fe

    def init_rho_ion_pb(self, t, max_iter=10000):
        it = ctypes.c_int(0)
        fe = np.zeros(3)
        self._call("init_rho_ion_pb", ctypes.byref(t), int(max_iter), ctypes.byref(it), _dp(fe))
        return it.value, fe


class MTRef:
    """The reference's own mt19937ar.c (built from source into _ref/)."""

    def __init__(self):
        path = os.path.join(_HERE, "_ref", "libmt_ref.so")
        if not os.path.exists(path):
            raise OSError(f"reference generator not built: {path}")
        self.lib = ctypes.CDLL(path)
        self.lib.init_genrand.argtypes = [ctypes.c_ulong]
        self.lib.init_genrand.restype = None
        self.lib.genrand_real3.argtypes = []
        self.lib.genrand_real3.restype = ctypes.c_double

    def seed(self, s):
        self.lib.init_genrand(int(s))

    def real3(self, n):
        return np.array([self.lib.genrand_real3() for _ in range(int(n))])


class OouraRef:
    """The reference's own Ooura rdft3d/rdft3dsort (built from source into _ref/)."""

    def __init__(self):
        if not os.path.exists(REF_LIB):
            raise OSError(f"reference FFT not built: {REF_LIB}")
        self.lib = ctypes.CDLL(REF_LIB)
        for fn in (self.lib.ref_a2a_k, self.lib.ref_a_k2a):
            fn.restype = None
            fn.argtypes = [ctypes.c_int] * 3 + [ctypes.POINTER(ctypes.c_double)]

    def a2a_k(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64).copy()
        nx, ny, nz_ = a.shape
        self.lib.ref_a2a_k(nx, ny, nz_ - 2, _dp(a))
        return a

    def a_k2a(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64).copy()
        nx, ny, nz_ = a.shape
        self.lib.ref_a_k2a(nx, ny, nz_ - 2, _dp(a))
        return a


# ---------------------------------------------------------------- network
def _pnet():
    lib = _load()
    if not hasattr(lib, "_pnet_ready"):
        dp = ctypes.POINTER(ctypes.c_double)
        lib.pnet_param_count.restype = ctypes.c_long
        lib.pnet_layers.argtypes = [ctypes.POINTER(ctypes.c_long)]
        lib.pnet_layers.restype = None
        lib.pnet_forward.argtypes = [ctypes.c_int, dp, dp, dp]
        lib.pnet_forward.restype = None
        lib.pnet_loss_grad.argtypes = [ctypes.c_int, dp, dp, dp, dp, dp]
        lib.pnet_loss_grad.restype = ctypes.c_double
        lib._pnet_ready = True
    return lib


def pnet_param_count():
    return int(_pnet().pnet_param_count())


def pnet_layers():
    out = (ctypes.c_long * (9 * 6))()
    _pnet().pnet_layers(out)
    return [tuple(out[6 * i:6 * i + 6]) for i in range(9)]


def pnet_forward(n, params, inp):
    params = np.ascontiguousarray(params, dtype=np.float64)
    inp = np.ascontiguousarray(inp, dtype=np.float64).reshape(n ** 3, 2)
    y = np.zeros((n ** 3, 3))
    _pnet().pnet_forward(n, _dp(params), _dp(inp), _dp(y))
    return y


def pnet_loss_grad(n, params, inp, e, target):
    params = np.ascontiguousarray(params, dtype=np.float64)
    inp = np.ascontiguousarray(inp, dtype=np.float64).reshape(n ** 3, 2)
    e = np.ascontiguousarray(e, dtype=np.float64)
    t = np.ascontiguousarray(target, dtype=np.float64).reshape(n ** 3)
    g = np.zeros(pnet_param_count())
    loss = _pnet().pnet_loss_grad(n, _dp(params), _dp(inp), _dp(e), _dp(t), _dp(g))
    return loss, g
