"""kapsel_calc_f_hydro_noalloc (no reference counterpart; for callers that
cannot allocate device memory) against kapsel_calc_f_hydro on the same state:
two calls in a row through the C ABI, the second with a smaller grid and
fewer particles, each compared with the context entry point's F_HYDRO and
TORQUE_HYDRO.  The caller's device buffers come from hipMalloc of the HIP
runtime the library itself links (test plumbing).  The bar: the same kernels
in the same order, so the results are identical; asserted to 1e-15 of the
largest value."""
import ctypes
import os

import numpy as np
import pytest

from icep_kapsel_amd import abi, workloads

pytestmark = pytest.mark.gpu


class _Hip:
    """hipMalloc / hipMemcpy / hipFree of the runtime libkapsel_gpu.so links."""

    def __init__(self):
        abi.load_library()
        self.lib = ctypes.CDLL(os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libamdhip64.so.7"))
        self.lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.lib.hipFree.argtypes = [ctypes.c_void_p]
        self.lib.hipDeviceSynchronize.argtypes = []
        self.ptrs = []

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = ctypes.c_void_p()
        assert self.lib.hipMalloc(ctypes.byref(p), a.nbytes) == 0
        self.ptrs.append(p)
        assert self.lib.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return p

    def download(self, p, like):
        out = np.empty_like(like)
        assert self.lib.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0  # device to host
        return out

    def close(self):
        for p in self.ptrs:
            self.lib.hipFree(p)
        self.ptrs = []


def _case(n, npart, prec, seed):
    from icep_kapsel_amd.gpu import KapselGPU

    rng = np.random.default_rng(seed)
    p = abi.make_params(n, n_particles=npart, radius=4.0, xi=2.0, rotation=1, precision=prec)
    pos = workloads.random_positions(workloads.Xoshiro256(seed), npart, (n, n, n), 8.0)
    st = workloads.particle_state(pos, rng.normal(0, 1e-2, (npart, 3)), rng.normal(0, 1e-3, (npart, 3)))
    k = KapselGPU(p)
    k.set_particles(st)
    k.make_phi_particle_sum()
    shape = k.shape
    u = [rng.normal(0, 1e-2, shape) for _ in range(3)]
    for d in range(3):
        k.set_field(f"U{d}", u[d])
    t = abi.make_time(0.05)
    k.calc_f_hydro(t)
    ref = k.get_particles()
    phi_sum = k.get_field("PHI_SUM")
    k.close()
    ft = np.float64 if prec == abi.KAPSEL_FP64 else np.float32
    hip = _Hip()
    try:
        du = [hip.upload(x.astype(ft)) for x in u]
        dphi = hip.upload(phi_sum.astype(ft))
        dps = hip.upload(st)
        lib = abi.load_library()
        rc = lib.kapsel_calc_f_hydro_noalloc(ctypes.byref(p), ctypes.byref(t), dps, du[0], du[1], du[2], dphi, None)
        assert rc == 0, lib.kapsel_last_error()
        assert hip.lib.hipDeviceSynchronize() == 0
        got = hip.download(dps, st)
    finally:
        hip.close()
    for f in ("F_HYDRO", "TORQUE_HYDRO"):
        a, b = got[abi.PFIELD[f]], ref[abi.PFIELD[f]]
        err = np.abs(a - b).max() / np.abs(b).max()
        print(f"noalloc n={n} np={npart} prec={prec} {f}: {err:.3e}", flush=True)
        assert err <= 1e-15, (f, err)


@pytest.mark.parametrize("prec", [abi.KAPSEL_FP64, abi.KAPSEL_FP32])
def test_noalloc_gather_matches_context_entry_point(prec):
    print(f"case noalloc gather prec={prec}", flush=True)
    _case(128, 200, prec, 11)   # first call
    _case(64, 20, prec, 12)     # second call, smaller grid and fewer particles


def test_noalloc_rejects_oversized_input():
    lib = abi.load_library()
    p = abi.make_params(256, n_particles=40000, radius=4.0, xi=2.0)
    t = abi.make_time(0.05)
    dummy = ctypes.c_void_p(16)  # never dereferenced: the size check comes first
    rc = lib.kapsel_calc_f_hydro_noalloc(ctypes.byref(p), ctypes.byref(t), dummy, dummy, dummy, dummy, dummy, None)
    assert rc == -1 and b"partial-sum slots" in lib.kapsel_last_error()
