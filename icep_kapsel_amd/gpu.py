"""Python front end of the C ABI (ctypes), used by the tests and bench.py.

``KapselGPU`` owns one ``kapsel_ctx``.  Its methods are the C entry points
with host arrays in the reference's padded Ooura layout; every call raises
``KapselError`` on a non-zero status.  There is no CPU fallback: without the
built library or a gfx950 device the constructor raises.
"""
import ctypes

import numpy as np

from . import abi


class KapselError(RuntimeError):
    pass


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class KapselGPU:
    def __init__(self, params):
        self.lib = abi.load_library()
        self.params = params
        ctx = ctypes.c_void_p()
        rc = self.lib.kapsel_create(ctypes.byref(params), ctypes.byref(ctx))
        if rc != 0:
            raise KapselError(f"kapsel_create: {rc}: {self.lib.kapsel_last_error().decode()}")
        self.ctx = ctx
        self.nx, self.ny, self.nz = params.nx, params.ny, params.nz
        self.nz_ = 2 * (self.nz // 2 + 1)
        self.shape = (self.nx, self.ny, self.nz_)
        self.np = params.n_particles

    def close(self):
        if getattr(self, "ctx", None):
            self.lib.kapsel_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _call(self, name, *args):
        rc = getattr(self.lib, "kapsel_" + name)(self.ctx, *args)
        if rc != 0:
            raise KapselError(f"kapsel_{name}: {rc}: {self.lib.kapsel_last_error().decode()}")

    # host <-> device ----------------------------------------------------
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

    def set_orientations(self, q):
        """q: (8, Np) -- q0..q3 then q_old0..q_old3 (scalar part first)."""
        a = np.ascontiguousarray(q, dtype=np.float64).reshape(8, self.np)
        self._call("set_orientations", _dp(a))

    def get_orientations(self):
        a = np.zeros((8, self.np))
        self._call("get_orientations", _dp(a))
        return a

    def set_particles(self, state):
        a = np.ascontiguousarray(state, dtype=np.float64).reshape(abi.P_NFIELDS, 3, self.np)
        self._call("set_particles", _dp(a))

    def get_particles(self):
        a = np.zeros((abi.P_NFIELDS, 3, self.np))
        self._call("get_particles", _dp(a))
        return a

    def sekibun(self):
        n = ctypes.c_int(0)
        self._call("get_sekibun", ctypes.byref(n), None)
        cells = np.zeros((n.value, 3), dtype=np.int32)
        self._call("get_sekibun", ctypes.byref(n), cells.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
        return cells

    def synchronize(self):
        self._call("synchronize")

    # operations ----------------------------------------------------------
    def copy_field(self, dst, src):
        self._call("copy_field", abi.FIELD[dst], abi.FIELD[src])

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

    def spm_step(self, t):
        self._call("spm_step", ctypes.byref(t))

    # dielectric Poisson -----------------------------------------------------
    def make_dielectric_field(self):
        self._call("dielectric_field")

    def poisson_apply(self, x_name, ax_name):
        self._call("poisson_apply", abi.FIELD[x_name], abi.FIELD[ax_name])

    def ax_unpadded(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        out = np.zeros_like(x)
        self._call("poisson_apply_host", _dp(x), _dp(out))
        return out

    def poisson_rhs(self, e0):
        e = np.ascontiguousarray(e0, dtype=np.float64)
        self._call("poisson_rhs", _dp(e))

    def poisson_solve(self, e0, tol, maxit, precond=0, check=True):
        """Returns (iters, relres); raises on non-convergence when check."""
        e = np.ascontiguousarray(e0, dtype=np.float64)
        it = ctypes.c_int(0)
        rr = ctypes.c_double(0.0)
        rc = self.lib.kapsel_poisson_solve(self.ctx, _dp(e), tol, int(maxit), int(precond), ctypes.byref(it),
                                           ctypes.byref(rr))
        if rc != 0 and (check or rc != -4):
            raise KapselError(f"kapsel_poisson_solve: {rc}: {self.lib.kapsel_last_error().decode()}")
        return it.value, rr.value

    def set_dielectric(self, eps_fluid, eps_particle, dielectric_cst=1.0):
        self._call("set_dielectric", float(eps_fluid), float(eps_particle), float(dielectric_cst))

    def maxwell_force(self, e0):
        e = np.ascontiguousarray(e0, dtype=np.float64)
        self._call("maxwell_force", _dp(e))

    def set_concurrency(self, on):
        self._call("set_concurrency", 1 if on else 0)

    def icep_step(self, t, e0, tol, maxit, precond=0, check=True):
        e = np.ascontiguousarray(e0, dtype=np.float64)
        it = ctypes.c_int(0)
        rr = ctypes.c_double(0.0)
        rc = self.lib.kapsel_icep_step(self.ctx, ctypes.byref(t), _dp(e), tol, int(maxit), int(precond),
                                       ctypes.byref(it), ctypes.byref(rr))
        if rc != 0 and (check or rc != -4):
            raise KapselError(f"kapsel_icep_step: {rc}: {self.lib.kapsel_last_error().decode()}")
        return it.value, rr.value

    # learned warm start -----------------------------------------------------
    def pnet_init(self):
        n = ctypes.c_long(0)
        self._call("pnet_init", ctypes.byref(n))
        return n.value

    def pnet_set_params(self, params):
        a = np.ascontiguousarray(params, dtype=np.float32)
        self._call("pnet_set_params", a.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))

    def pnet_get_params(self, n):
        a = np.zeros(n, dtype=np.float32)
        self._call("pnet_get_params", a.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
        return a

    def pnet_loss_grad(self, beta, e_hat, scale, n_params=None):
        e = np.ascontiguousarray(e_hat, dtype=np.float64)
        loss = ctypes.c_double(0.0)
        g = np.zeros(n_params, dtype=np.float32) if n_params else None
        gp = g.ctypes.data_as(ctypes.POINTER(ctypes.c_float)) if g is not None else None
        self._call("pnet_loss_grad", float(beta), _dp(e), float(scale), ctypes.byref(loss), gp)
        return loss.value, g

    def pnet_train_step(self, beta, e_hat, scale, lr):
        e = np.ascontiguousarray(e_hat, dtype=np.float64)
        loss = ctypes.c_double(0.0)
        self._call("pnet_train_step", float(beta), _dp(e), float(scale), float(lr), ctypes.byref(loss))
        return loss.value

    def pnet_warm_start(self, beta, e0, scale):
        e = np.ascontiguousarray(e0, dtype=np.float64)
        self._call("pnet_warm_start", float(beta), _dp(e), float(scale))

    def poisson_galerkin_start(self, e0):
        """Rescales the warm start in POTENTIAL (energy-norm optimal); returns gamma."""
        e = np.ascontiguousarray(e0, dtype=np.float64)
        g = ctypes.c_double(0.0)
        self._call("poisson_galerkin_start", _dp(e), ctypes.byref(g))
        return g.value

    # thermal fluctuations -------------------------------------------------------
    def set_thermostat(self, kbt, alpha_v, alpha_o, seed=0):
        self._call("set_thermostat", float(kbt), float(alpha_v), float(alpha_o), int(seed))

    def thermostat(self, t):
        self._call("thermostat", ctypes.byref(t))

    # restart files -------------------------------------------------------------
    def save_restart(self, path, t, template=None):
        """Save_Restart_udf: the context state as a UDF resume record."""
        tp = template.encode() if template else None
        self._call("save_restart", str(path).encode(), ctypes.byref(t), tp)

    def load_restart(self, path, t):
        """Force_restore_parameters: state from a UDF resume record; sets t.ts, t.time."""
        self._call("load_restart", str(path).encode(), ctypes.byref(t))

    # electrolyte -------------------------------------------------------------
    def set_electrolyte(self, el):
        self._call("set_electrolyte", ctypes.byref(el))

    def init_rho_ion(self):
        self._call("init_rho_ion")

    def init_rho_ion_dielectric(self):
        self._call("init_rho_ion_dielectric")

    def _solve_call(self, name, t, tol, maxit, precond, check):
        it = ctypes.c_int(0)
        rr = ctypes.c_double(0.0)
        rc = getattr(self.lib, "kapsel_" + name)(self.ctx, ctypes.byref(t), tol, int(maxit), int(precond),
                                                 ctypes.byref(it), ctypes.byref(rr))
        if rc != 0 and (check or rc != -4):
            raise KapselError(f"kapsel_{name}: {rc}: {self.lib.kapsel_last_error().decode()}")
        return it.value, rr.value

    def dielectric_potential(self, t, tol, maxit, precond=0, check=True):
        """The ions' (and surface charge's) potential in the permittivity; (iters, relres)."""
        return self._solve_call("dielectric_potential", t, tol, maxit, precond, check)

    def icep_electrolyte_step(self, t, tol, maxit, precond=0, check=True):
        """One Electrolyte + Dielectric step; advances t; (iters, relres)."""
        return self._solve_call("icep_electrolyte_step", t, tol, maxit, precond, check)

    def init_rho_ion_pb(self, t, max_iter=10000):
        it = ctypes.c_int(0)
        fe = np.zeros(3)
        self._call("init_rho_ion_pb", ctypes.byref(t), int(max_iter), ctypes.byref(it), _dp(fe))
        return it.value, fe

    def free_energy_pb(self, t):
        fe = np.zeros(3)
        self._call("free_energy_pb", ctypes.byref(t), _dp(fe))
        return fe

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

    # timing ---------------------------------------------------------------
    def kernel_timing(self, classes):
        mask = 0
        for c in classes:
            mask |= 1 << abi.KCLASS[c]
        self._call("kernel_timing", mask)

    def kernel_timing_every(self, every):
        """Time one launch in `every` of each selected class."""
        self._call("kernel_timing_every", int(every))

    def kernel_timing_get(self, cls):
        ms = ctypes.c_double(0.0)
        n = ctypes.c_int(0)
        b = ctypes.c_double(0.0)
        self._call("kernel_timing_get", abi.KCLASS[cls], ctypes.byref(ms), ctypes.byref(n), ctypes.byref(b))
        return ms.value, n.value, b.value
