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


def test_fp32_step_spectrum_reuse_matches_recompute():
    """An FP32 slaved step right after a step starts from the velocity
    spectrum that step left in U; with an entry point in between it forms the
    spectrum from zeta again (Zeta_k2u_k).  The two are equal but for FP32
    rounding: 10 steps of c1 both ways agree to 1e-5 (relative, velocities
    and vorticity) and each stays as close to the FP64 run as the other."""
    print("case fp32 spectrum reuse c1", flush=True)
    from icep_kapsel_amd.gpu import KapselGPU
    p32, st, dt = workloads.spm_config("c1", precision=abi.KAPSEL_FP32)
    p64, _, _ = workloads.spm_config("c1", precision=abi.KAPSEL_FP64)
    ks = [KapselGPU(p) for p in (p32, p32, p64)]
    for k in ks:
        k.set_particles(st)
        k.init_rigidity()
    ts = [abi.make_time(dt) for _ in ks]
    for _ in range(10):
        for i, (k, t) in enumerate(zip(ks, ts)):
            if i == 1:
                k.set_concurrency(True)  # an entry point between the steps: no reuse
            k.spm_step(t)
    reuse, again, k64 = ks
    e = _errors(again, reuse)
    print("reuse vs recompute:", {f: f"{v:.1e}" for f, v in e.items() if f not in SCRATCH}, flush=True)
    assert e["V"] < 1e-5 and e["OMEGA"] < 1e-5
    assert max(e["U0"], e["U1"], e["U2"], e["ZETA0"], e["ZETA1"]) < 1e-5
    er, ea = _errors(k64, reuse), _errors(k64, again)
    for f in ("V", "OMEGA", "ZETA0", "ZETA1"):
        assert er[f] < 2 * ea[f] + 1e-7 and ea[f] < 2 * er[f] + 1e-7, f
    for k in ks:
        k.close()


def _icep_run(prec, precond, tol, steps=5):
    from icep_kapsel_amd.gpu import KapselGPU
    p, st, e0, dt = workloads.icep_config(precision=prec)
    k = KapselGPU(p)
    k.set_particles(st)
    k.init_rigidity()
    t = abi.make_time(dt)
    for _ in range(steps):
        it, rr = k.icep_step(t, e0, tol, 20000, precond)
        assert rr < tol
    return k


def _icep_errors(k, ref):
    e = _errors(ref, k)
    for f in SCRATCH:
        del e[f]
    n = k.nz
    e["POTENTIAL"] = rel(k.get_field("POTENTIAL") - k.get_field("POTENTIAL")[:, :, :n].mean(),
                         ref.get_field("POTENTIAL") - ref.get_field("POTENTIAL")[:, :, :n].mean())
    # a |omega - omega_64| against the largest particle speed: the spheres'
    # rotation is small beside their translation (a dielectric sphere in a
    # uniform field carries no torque; here only the neighbours' fields turn it)
    a = 6.0
    pk, pr = k.get_particles(), ref.get_particles()
    e["a_domega_over_vmax"] = float(a * np.abs(pk[abi.PFIELD["OMEGA"]] - pr[abi.PFIELD["OMEGA"]]).max() /
                                    np.abs(pr[abi.PFIELD["V"]]).max())
    e["a_omega_over_vmax"] = float(a * np.abs(pr[abi.PFIELD["OMEGA"]]).max() / np.abs(pr[abi.PFIELD["V"]]).max())
    return e


def test_icep_fp32_velocity_error():
    """ICEP box, 5 steps, FP32 context (spectral PCG, FP64 refinement) against
    the FP64 context (plain CG).

    Two sources of difference: the solve's own tolerance (two different
    iterations stopped at the same relative residual leave different
    potentials) and FP32 rounding of the fields.  The first is measured by two
    FP64 runs at 1e-8 (plain CG against spectral PCG) and by repeating the
    comparison at 1e-11: it is below 1e-7 everywhere, so the FP32-FP64
    difference is FP32 rounding.  Bounds: particle and fluid velocities and
    zeta 1e-5 (the target).  omega and the hydrodynamic torque 1e-4 relative:
    a dielectric sphere in a uniform field carries almost no torque, a |omega|
    is about 2 % of the largest particle speed, and the torque -- a moment of
    (u_p - u) phi over the sphere -- turns the ~1e-6 FP32 error of u into a
    relative error some 50 times larger; in absolute terms a |d omega| stays
    within 1e-5 of the largest particle speed (asserted)."""
    print("case precision icep fp32 vs fp64", flush=True)
    k64 = _icep_run(abi.KAPSEL_FP64, abi.PRECOND_NONE, 1e-8)
    k64s = _icep_run(abi.KAPSEL_FP64, abi.PRECOND_SPECTRAL, 1e-8)
    k32 = _icep_run(abi.KAPSEL_FP32, abi.PRECOND_SPECTRAL, 1e-8)
    e32 = _icep_errors(k32, k64)
    esol = _icep_errors(k64s, k64)
    for k in (k64, k64s, k32):
        k.close()
    print("tol 1e-8, fp32 vs fp64:", {f: f"{v:.1e}" for f, v in e32.items()}, flush=True)
    print("tol 1e-8, fp64 spectral vs fp64 plain:", {f: f"{v:.1e}" for f, v in esol.items()}, flush=True)
    k64 = _icep_run(abi.KAPSEL_FP64, abi.PRECOND_SPECTRAL, 1e-11)
    k32 = _icep_run(abi.KAPSEL_FP32, abi.PRECOND_SPECTRAL, 1e-11)
    etight = _icep_errors(k32, k64)
    k64.close()
    k32.close()
    print("tol 1e-11, fp32 vs fp64:", {f: f"{v:.1e}" for f, v in etight.items()}, flush=True)
    solver_fields = [f for f in esol if f != "a_omega_over_vmax"]
    assert max(esol[f] for f in solver_fields) < 1e-7, esol
    for e in (e32, etight):
        assert e["V"] < 1e-5 and max(e["U0"], e["U1"], e["U2"]) < 1e-5
        assert max(e["ZETA0"], e["ZETA1"]) < 1e-5
        assert e["a_domega_over_vmax"] < 1e-5
        assert e["OMEGA"] < 1e-4 and e["TORQUE_HYDRO_PREVIOUS"] < 1e-4


def test_electrolyte_fp32_drift_500_steps_256():
    """The scale + electrolyte workload at its stated length: 256^3, 2000
    charged spheres, 500 steps, FP32 against FP64 from the same start.
    Reported: the relative difference of particle velocities, the ion
    spectra, zeta and positions at steps 1, 50, 100, 250, 500.  The velocity
    difference grows about linearly, 2e-7 per step (measured: 1.1e-5 at step
    50, 9.2e-5 at step 500): FP32 rounding of the fluid fields enters every
    step's gather and accumulates in the particle momenta.  Bounds: particle
    velocity 2e-4, ion spectra 1e-5, positions 1e-6 of the box."""
    from icep_kapsel_amd.gpu import KapselGPU
    print("case electrolyte 256^3 x 2000 fp32 vs fp64, 500 steps", flush=True)
    ks = []
    for prec in (abi.KAPSEL_FP64, abi.KAPSEL_FP32):
        p, st, el, dt = workloads.electrolyte_config(precision=prec)
        k = KapselGPU(p)
        k.set_particles(st)
        k.set_electrolyte(el)
        k.init_rigidity()
        k.init_rho_ion()
        ks.append((k, abi.make_time(dt)))
    (k64, t64), (k32, t32) = ks
    worst = {}
    for step in range(1, 501):
        k64.electrolyte_step(t64)
        k32.electrolyte_step(t32)
        if step in (1, 50, 100, 250, 500):
            p64, p32 = k64.get_particles(), k32.get_particles()
            e = {f: rel(p32[abi.PFIELD[f]], p64[abi.PFIELD[f]]) for f in ("V", "OMEGA")}
            e["X_over_box"] = float(np.abs(p32[abi.PFIELD["X_NOPBC"]] - p64[abi.PFIELD["X_NOPBC"]]).max() / 256.0)
            for f in ("CONC0", "CONC1", "ZETA0", "ZETA1"):
                e[f] = rel(k32.get_field(f), k64.get_field(f))
            worst = e
            print(f"step {step}: " + ", ".join(f"{a} {b:.2e}" for a, b in e.items()), flush=True)
    k64.close()
    k32.close()
    assert worst["V"] < 2e-4
    assert max(worst["CONC0"], worst["CONC1"]) < 1e-5
    assert worst["X_over_box"] < 1e-6


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
