"""ctypes front end of libkapsel_h5.so: HDF5 field and particle dumps in the
reference's hdf5_writer layout (include/kapsel_h5.h).  Host only."""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libkapsel_h5.so")
_lib = None


class Crop(ctypes.Structure):
    _fields_ = [("start", ctypes.c_int * 3), ("count", ctypes.c_int * 3), ("stride", ctypes.c_int * 3)]


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"HDF5 dump library not built: {LIB_PATH}")
        lib = ctypes.CDLL(LIB_PATH)
        vp, i, d, dp, ip, cp = (ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.POINTER(ctypes.c_double),
                                ctypes.POINTER(ctypes.c_int), ctypes.c_char_p)
        lib.kapsel_h5_last_error.restype = cp
        lib.kapsel_h5_create.argtypes = [cp, cp, i, i, i, d, d, i, i, ip, ip, ctypes.POINTER(Crop), i,
                                         ctypes.POINTER(vp)]
        lib.kapsel_h5_frame_begin.argtypes = [vp]
        lib.kapsel_h5_write_field.argtypes = [vp, cp, dp]
        lib.kapsel_h5_write_particles.argtypes = [vp, dp, dp]
        lib.kapsel_h5_frame_end.argtypes = [vp]
        lib.kapsel_h5_close.argtypes = [vp]
        lib.kapsel_h5_read.argtypes = [cp, cp, dp, ctypes.c_longlong, ctypes.POINTER(ctypes.c_longlong)]
        for fn in ("kapsel_h5_create", "kapsel_h5_frame_begin", "kapsel_h5_write_field", "kapsel_h5_write_particles",
                   "kapsel_h5_frame_end", "kapsel_h5_close", "kapsel_h5_read"):
            getattr(lib, fn).restype = i
        _lib = lib
    return _lib


def _check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what}: {rc}: {load().kapsel_h5_last_error().decode()}")


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Writer:
    """hdf5_writer: create, then frames of fields and particles, then close."""

    def __init__(self, out_dir, out_name, shape, dx, dt, nump_total, print_list=None, spec=None, crop=None,
                 fields=True):
        lib = load()
        self.lib = lib
        pl = None if print_list is None else np.ascontiguousarray(print_list, dtype=np.int32)
        sp = None if spec is None else np.ascontiguousarray(spec, dtype=np.int32)
        npr = nump_total if pl is None else len(pl)
        c = None
        if crop is not None:
            c = Crop()
            c.start[:], c.count[:], c.stride[:] = crop
        h = ctypes.c_void_p()
        ip = ctypes.POINTER(ctypes.c_int)
        _check(lib.kapsel_h5_create(out_dir.encode(), out_name.encode(), shape[0], shape[1], shape[2], dx, dt,
                                    nump_total, npr, None if pl is None else pl.ctypes.data_as(ip),
                                    None if sp is None else sp.ctypes.data_as(ip),
                                    None if c is None else ctypes.byref(c), 1 if fields else 0, ctypes.byref(h)),
               "kapsel_h5_create")
        self.h = h
        self.nump = nump_total

    def frame(self, fields=None, state=None, orientations=None):
        _check(self.lib.kapsel_h5_frame_begin(self.h), "kapsel_h5_frame_begin")
        for name, a in (fields or {}).items():
            a = np.ascontiguousarray(a, dtype=np.float64)
            _check(self.lib.kapsel_h5_write_field(self.h, name.encode(), _dp(a)), "kapsel_h5_write_field")
        if state is not None:
            s = np.ascontiguousarray(state, dtype=np.float64)
            q = None if orientations is None else np.ascontiguousarray(orientations, dtype=np.float64)
            _check(self.lib.kapsel_h5_write_particles(self.h, _dp(s), None if q is None else _dp(q)),
                   "kapsel_h5_write_particles")
        _check(self.lib.kapsel_h5_frame_end(self.h), "kapsel_h5_frame_end")

    def close(self):
        if self.h:
            _check(self.lib.kapsel_h5_close(self.h), "kapsel_h5_close")
            self.h = None


def read(path, obj):
    """A dataset or attribute (owner@name) as a flat float64 array."""
    lib = load()
    n = ctypes.c_longlong(0)
    _check(lib.kapsel_h5_read(path.encode(), obj.encode(), None, 0, ctypes.byref(n)), f"kapsel_h5_read {obj}")
    out = np.zeros(n.value)
    _check(lib.kapsel_h5_read(path.encode(), obj.encode(), _dp(out), n.value, ctypes.byref(n)), f"read {obj}")
    return out
