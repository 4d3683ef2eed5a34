"""FP32 fields against FP64, per field, and long-run FP64 drift against the oracle.

Precision workload: 128^3, 200 spheres a=4 xi=2, Gaussian particle velocities
std 1e-3.  The FP32 context runs the FFTs, advection and projection in FP32
and keeps the particle reductions (hydrodynamic force and torque, TwoSum
compensated) and the particle update in FP64.  Each field's error is
max|f32 - f64| / max|f64| after 1 and after 20 steps; the particle and fluid
velocities must stay within 1e-5 of FP64.  The ICEP box (FP32 + spectral PCG
with FP64 refinement vs FP64) is held to the same bound on the particle
velocities, the rigid-body surface velocities |v| + a |omega| and the fluid
velocity.

Drift: the 64^3, 8-sphere sedimentation workload, FP64 GPU against the FP64
oracle over 1000 steps (bound written in the test).
"""
import numpy as np
import pytest

from icep_kapsel_amd import abi, workloads

pytestmark = pytest.mark.gpu

FIELDS = ["ZETA0", "ZETA1", "U0", "U1", "U2", "PHI", "PHI_SUM", "UP0", "UP1", "UP2", "F0", "F1", "F2"]
PFIELDS = ["X", "V", "OMEGA", "F_HYDRO_PREVIOUS", "TORQUE_HYDRO_PREVIOUS"]
# the step's scratch: an FP32 step ends with the particle force as a spectrum
# (it enters U2zeta_k in k-space), an FP64 step with it in real space
SCRATCH = ("F0", "F1", "F2")


def rel(a, b):
    return float(np.abs(a - b).max() / max(1e-300, np.abs(b).max()))


def _pair(cfg):
    from icep_kapsel_amd.gpu import KapselGPU
    p64, st, dt = workloads.spm_config(cfg, precision=abi.KAPSEL_FP64)
    p32, _, _ = workloads.spm_config(cfg, precision=abi.KAPSEL_FP32)
    ks = [KapselGPU(p) for p in (p64, p32)]
    for k in ks:
        k.set_particles(st)
        k.init_rigidity()
    return ks, dt


def _errors(k64, k32):
    e = {f: rel(k32.get_field(f), k64.get_field(f)) for f in FIELDS}
    p64, p32 = k64.get_particles(), k32.get_particles()
    for f in PFIELDS:
        i = abi.PFIELD[f]
        e[f] = rel(p32[i], p64[i])
    return e


def test_fp32_per_field_error_c2():
    print("case precision c2 fp32 vs fp64", flush=True)
    (k64, k32), dt = _pair("c2")
    t64, t32 = abi.make_time(dt), abi.make_time(dt)
    k64.spm_step(t64)
    k32.spm_step(t32)
    e1 = _errors(k64, k32)
    print("after 1 step:", {f: f"{v:.1e}" for f, v in e1.items() if f not in SCRATCH}, flush=True)
    for _ in range(19):
        k64.spm_step(t64)
        k32.spm_step(t32)
    e20 = _errors(k64, k32)
    print("after 20 steps:", {f: f"{v:.1e}" for f, v in e20.items() if f not in SCRATCH}, flush=True)
    for e in (e1, e20):
        assert e["V"] < 1e-5 and e["OMEGA"] < 1e-5
        assert max(e["U0"], e["U1"], e["U2"]) < 1e-5
        assert max(e["ZETA0"], e["ZETA1"]) < 1e-5
        assert e["X"] < 1e-9  # positions move by v dt from an FP64 start
        assert max(e[f] for f in FIELDS if f not in SCRATCH) < 1e-4


def test_icep_fp32_velocity_error():
    """ICEP box, 5 steps: FP32 context (spectral PCG, FP64 refinement, tol 1e-8)
    against the FP64 context (plain CG, tol 1e-8)."""
    from icep_kapsel_amd.gpu import KapselGPU
    print("case precision icep fp32 vs fp64", flush=True)
    ks = []
    for prec, pc in ((abi.KAPSEL_FP64, abi.PRECOND_NONE), (abi.KAPSEL_FP32, abi.PRECOND_SPECTRAL)):
        p, st, e0, dt = workloads.icep_config(precision=prec)
        k = KapselGPU(p)
        k.set_particles(st)
        k.init_rigidity()
        t = abi.make_time(dt)
        for _ in range(5):
            it, rr = k.icep_step(t, e0, 1e-8, 20000, pc)
            assert rr < 1e-8
        ks.append(k)
    k64, k32 = ks
    e = _errors(k64, k32)
    for f in SCRATCH:
        del e[f]
    e["POTENTIAL"] = rel(k32.get_field("POTENTIAL") - k32.get_field("POTENTIAL")[:, :, :128].mean(),
                         k64.get_field("POTENTIAL") - k64.get_field("POTENTIAL")[:, :, :128].mean())
    # rigid-body velocity: surface velocity |v| + a |omega| of each sphere
    # (the angular velocities alone are small here: spheres in a uniform
    # field carry almost no torque, so their own relative error is larger)
    a = 6.0
    p64, p32 = k64.get_particles(), k32.get_particles()
    v64, v32 = p64[abi.PFIELD["V"]], p32[abi.PFIELD["V"]]
    w64, w32 = p64[abi.PFIELD["OMEGA"]], p32[abi.PFIELD["OMEGA"]]
    surf = np.linalg.norm(v64, axis=0) + a * np.linalg.norm(w64, axis=0)
    dsurf = np.linalg.norm(v32 - v64, axis=0) + a * np.linalg.norm(w32 - w64, axis=0)
    e["rigid_body_velocity"] = float(dsurf.max() / surf.max())
    print("icep after 5 steps:", {f: f"{v:.1e}" for f, v in e.items()}, flush=True)
    assert e["V"] < 1e-5 and e["rigid_body_velocity"] < 1e-5
    assert max(e["U0"], e["U1"], e["U2"]) < 1e-5
    assert e["OMEGA"] < 1e-4 and max(e["ZETA0"], e["ZETA1"]) < 1e-4


def test_fp64_drift_1000_steps_c1():
    """64^3, 8 spheres a=5, body force 1e-3 along z: 1000 FP64 steps on the
    GPU and in the oracle; every compared field within 1e-12 (relative to its
    maximum) at step 1000."""
    from oracle.oracle import Oracle
    from icep_kapsel_amd.gpu import KapselGPU
    print("case drift c1 1000 steps", flush=True)
    p, st, dt = workloads.spm_config("c1")
    k, o = KapselGPU(p), Oracle(p)
    for x in (k, o):
        x.set_particles(st)
        x.init_rigidity()
    tk, to = abi.make_time(dt), abi.make_time(dt)
    worst = {}
    for step in range(1, 1001):
        k.spm_step(tk)
        o.spm_step(to)
        if step in (1, 10, 100, 1000):
            e = {f: rel(k.get_field(f), o.get_field(f)) for f in ("ZETA0", "ZETA1", "PHI")}
            pk, po = k.get_particles(), o.get_particles()
            for f in ("X", "V", "OMEGA"):
                e[f] = rel(pk[abi.PFIELD[f]], po[abi.PFIELD[f]])
            worst[step] = max(e.values())
            print(f"step {step}: " + ", ".join(f"{a} {b:.1e}" for a, b in e.items()), flush=True)
    assert worst[1] < 1e-12
    assert worst[1000] < 1e-12
