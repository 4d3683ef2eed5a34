"""UDF restart records (the reference's resume.CONTINUE.Saved_Data) on the host.

Thin numpy front end of kapsel_restart_write / kapsel_restart_read: no GPU is
needed.  Field arrays use the padded k-space layout (NX, NY, NZ_), particle
state the (P_NFIELDS, 3, Np) layout of KapselGPU.get_particles, orientations
(8, Np) with q then q_old, scalar part first.
"""
import ctypes

import numpy as np

from . import abi


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _check(lib, rc, name):
    if rc != 0:
        raise OSError(f"{name}: {rc}: {lib.kapsel_last_error().decode()}")


def write_restart(path, zeta0, zeta1, uk_dc, particles, orientations, ts, time, conc=None, template=None):
    lib = abi.load_library()
    nx, ny, nz_ = zeta0.shape
    z0 = np.ascontiguousarray(zeta0, dtype=np.float64)
    z1 = np.ascontiguousarray(zeta1, dtype=np.float64)
    ps = np.ascontiguousarray(particles, dtype=np.float64)
    q = np.ascontiguousarray(orientations, dtype=np.float64)
    d = abi.RestartData()
    d.nx, d.ny, d.nz = nx, ny, nz_ - 2
    d.n_particles = ps.shape[-1] if ps.size else 0
    d.ts, d.time = int(ts), float(time)
    d.uk_dc[:] = [float(v) for v in uk_dc]
    d.zeta0, d.zeta1, d.particles, d.orientations = _dp(z0), _dp(z1), _dp(ps), _dp(q)
    keep = [z0, z1, ps, q]
    if conc is not None:
        c0 = np.ascontiguousarray(conc[0], dtype=np.float64)
        c1 = np.ascontiguousarray(conc[1], dtype=np.float64)
        d.n_spec, d.conc0, d.conc1 = 2, _dp(c0), _dp(c1)
        keep += [c0, c1]
    rc = lib.kapsel_restart_write(str(path).encode(), template.encode() if template else None, ctypes.byref(d))
    _check(lib, rc, "kapsel_restart_write")


def read_restart(path, nx, ny, nz, n_particles, n_spec=0):
    """Returns a dict: zeta0, zeta1, uk_dc, particles, orientations, ts, time (+ conc)."""
    lib = abi.load_library()
    nz_ = 2 * (nz // 2 + 1)
    out = {"zeta0": np.zeros((nx, ny, nz_)), "zeta1": np.zeros((nx, ny, nz_)),
           "particles": np.zeros((abi.P_NFIELDS, 3, max(n_particles, 1))),
           "orientations": np.zeros((8, max(n_particles, 1)))}
    d = abi.RestartData()
    d.nx, d.ny, d.nz, d.n_particles, d.n_spec = nx, ny, nz, n_particles, n_spec
    d.zeta0, d.zeta1 = _dp(out["zeta0"]), _dp(out["zeta1"])
    d.particles, d.orientations = _dp(out["particles"]), _dp(out["orientations"])
    if n_spec:
        out["conc"] = [np.zeros((nx, ny, nz_)), np.zeros((nx, ny, nz_))]
        d.conc0, d.conc1 = _dp(out["conc"][0]), _dp(out["conc"][1])
    rc = lib.kapsel_restart_read(str(path).encode(), ctypes.byref(d))
    _check(lib, rc, "kapsel_restart_read")
    out["uk_dc"] = np.array(d.uk_dc[:])
    out["ts"], out["time"] = d.ts, d.time
    out["particles"] = out["particles"][:, :, :n_particles]
    out["orientations"] = out["orientations"][:, :n_particles]
    return out
