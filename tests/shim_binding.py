"""ctypes access to icep_kapsel_amd/libkapsel_shims.so, the reference-signature
shims (C++ functions, so their symbols are found by their demangled names)."""
import ctypes
import os
import subprocess

import numpy as np

from icep_kapsel_amd import abi

SHIM_LIB = os.path.join(os.path.dirname(abi.LIB_PATH), "libkapsel_shims.so")
DIM = 3


class CTime(ctypes.Structure):
    _fields_ = [("ts", ctypes.c_int), ("time", ctypes.c_double), ("dt_fluid", ctypes.c_double),
                ("hdt_fluid", ctypes.c_double), ("dt_md", ctypes.c_double), ("hdt_md", ctypes.c_double)]


class Quaternion(ctypes.Structure):
    _fields_ = [("s", ctypes.c_double), ("v", ctypes.c_double * DIM)]


_V = ctypes.c_double * DIM
_M = (ctypes.c_double * DIM) * DIM
_PARTICLE_MEMBERS = [
    ("spec", ctypes.c_int), ("x", _V), ("x_previous", _V), ("x_nopbc", _V),
    ("v", _V), ("v_old", _V), ("v_slip", _V),
    ("f_hydro", _V), ("f_hydro_previous", _V), ("f_hydro1", _V), ("f_slip", _V), ("f_slip_previous", _V),
    ("fr", _V), ("fr_previous", _V), ("torque_r", _V), ("torque_r_previous", _V),
    ("omega", _V), ("omega_old", _V), ("omega_slip", _V),
    ("torque_hydro", _V), ("torque_hydro_previous", _V), ("torque_hydro1", _V), ("torque_slip", _V),
    ("torque_slip_previous", _V),
    ("momentum_depend_fr", _V), ("QR", _M), ("QR_old", _M), ("q", Quaternion), ("q_old", Quaternion),
    ("mass", ctypes.c_double), ("mass_center", _V), ("inertia", _M),
    ("surface_mass", ctypes.c_double), ("surface_mass_center", _V), ("surface_inertia", _M),
    ("surface_dv", _V), ("surface_dw", _V),
]


class Particle(ctypes.Structure):
    _fields_ = _PARTICLE_MEMBERS


# library particle field -> Particle member
MEMBER_OF = {
    "X": "x", "X_PREVIOUS": "x_previous", "X_NOPBC": "x_nopbc", "V": "v", "V_OLD": "v_old",
    "OMEGA": "omega", "OMEGA_OLD": "omega_old", "F_HYDRO": "f_hydro", "F_HYDRO_PREVIOUS": "f_hydro_previous",
    "FR": "fr", "FR_PREVIOUS": "fr_previous", "TORQUE_HYDRO": "torque_hydro",
    "TORQUE_HYDRO_PREVIOUS": "torque_hydro_previous", "TORQUE_R": "torque_r",
    "TORQUE_R_PREVIOUS": "torque_r_previous",
}


def particles_from_state(state):
    """Particle array from the library's SoA state (P_NFIELDS, 3, n)."""
    n = state.shape[2]
    arr = (Particle * n)()
    for f, m in MEMBER_OF.items():
        for i in range(n):
            getattr(arr[i], m)[:] = list(state[abi.PFIELD[f], :, i])
    return arr


def member_array(arr, name):
    return np.array([list(getattr(p, name)) for p in arr])


def _symbols():
    out = subprocess.run(["nm", "-D", "--defined-only", SHIM_LIB], capture_output=True, text=True,
                         check=True).stdout
    mangled = [line.split()[-1] for line in out.splitlines() if " T " in line]
    demangled = subprocess.run(["c++filt"], input="\n".join(mangled), capture_output=True, text=True,
                               check=True).stdout.splitlines()
    return dict(zip(demangled, mangled))


def load():
    """(library, {function name: ctypes function}) with argtypes set."""
    if not os.path.exists(SHIM_LIB):
        raise OSError(f"shim library not built: {SHIM_LIB}")
    abi.load_library()
    lib = ctypes.CDLL(SHIM_LIB)
    syms = _symbols()
    dp, vp, i, d = ctypes.POINTER(ctypes.c_double), ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    pp = ctypes.POINTER(Particle)
    dpp = ctypes.POINTER(dp)
    protos = {
        "kapsel_shims_attach": [vp, i, i, d, d, i, dp, i],
        "kapsel_shims_pack": [pp, i, dp],
        "kapsel_shims_unpack": [dp, i, pp],
        "A2a_k": [dp], "A_k2a": [dp],
        "Make_phi_particle_sum": [dp, dp, pp, d],
        "Make_u_particle_sum": [dpp, dp, pp, d],
        "Calc_f_hydro_correct_precision": [pp, dp, dpp, ctypes.POINTER(CTime)],
        "Make_dielectric_field": [pp, dp, dp, d, d, d],
        "Ax": [dp, dp],
        "Build_rhs": [dp, dp, dp, dp],
        "Charge_field2potential_dielectric_LO": [dp, dp, dp, dp, dp],
    }
    fns = {}
    for name, args in protos.items():
        matches = [m for dm, m in syms.items() if dm.startswith(name + "(")]
        if len(matches) != 1:
            raise OSError(f"{SHIM_LIB}: {name} not found once ({matches})")
        fn = getattr(lib, matches[0])
        fn.restype = None
        fn.argtypes = args
        fns[name] = fn
    return lib, fns


def dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def dptrs(arrays):
    return (ctypes.POINTER(ctypes.c_double) * len(arrays))(*[dptr(a) for a in arrays])
