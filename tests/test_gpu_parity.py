"""GPU parity: the HIP library (through the C ABI) against the CPU oracle and
against golden vectors of the reference's own Ooura FFT.

Tolerances (relative to the field's max magnitude):
  FP64  1e-13 per FFT / pointwise op, 1e-12 after one full SPM step
  FP32  4e-6 per FFT, see test_gpu_precision.py for the per-field step errors
"""
import glob
import os

import numpy as np
import pytest

from icep_kapsel_amd import abi, workloads
from tests import spectral as sp

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def gpu(p):
    from icep_kapsel_amd.gpu import KapselGPU
    return KapselGPU(p)


def oracle(p):
    from oracle.oracle import Oracle
    return Oracle(p)


def rel(a, b):
    s = np.abs(b).max()
    return np.abs(a - b).max() / (s if s > 0 else 1.0)


@pytest.mark.parametrize("prec,tol", [(abi.KAPSEL_FP64, 1e-13), (abi.KAPSEL_FP32, 4e-6)])
@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(HERE, "golden", "ooura_fft_*.npz"))))
def test_fft_matches_reference_golden(path, prec, tol):
    g = np.load(path)
    nx, ny, nz = g["real_in"].shape
    print(f"case fft golden {nx}x{ny}x{nz} prec={prec}", flush=True)
    k = gpu(abi.make_params((nx, ny, nz), precision=prec))
    k.set_field("U0", sp.pad_real(g["real_in"]))
    k.fft_r2c("U0")
    assert rel(k.get_field("U0"), g["spec_out"]) < tol
    k.set_field("U1", g["spec_in"])
    k.fft_c2r("U1")
    assert rel(k.get_field("U1")[:, :, :nz], g["real_out"]) < tol


@pytest.mark.parametrize("shape", [(16, 16, 16), (32, 32, 32), (64, 64, 64), (128, 128, 128), (32, 64, 256),
                                   (256, 16, 32)])
def test_fft_matches_oracle(shape):
    print(f"case fft oracle {shape}", flush=True)
    nx, ny, nz = shape
    rng = np.random.default_rng(sum(shape))
    real = rng.standard_normal(shape)
    p = abi.make_params(shape)
    k, o = gpu(p), oracle(p)
    for x in (k, o):
        x.set_field("U0", sp.pad_real(real))
        x.fft_r2c("U0")
    spec_g, spec_o = k.get_field("U0"), o.get_field("U0")
    assert rel(spec_g, spec_o) < 1e-13
    # inverse of a Hermitian-consistent spectrum
    spec = spec_o.copy()
    for x in (k, o):
        x.set_field("U1", spec)
        x.fft_c2r("U1")
    assert rel(k.get_field("U1")[:, :, :nz], o.get_field("U1")[:, :, :nz]) < 1e-13
    assert np.abs(k.get_field("U1")[:, :, :nz] - real).max() < 1e-12


@pytest.mark.parametrize("shape", [(32, 32, 32), (64, 64, 64), (128, 128, 128), (64, 32, 32)])
def test_fft_fp32_matches_oracle(shape):
    """FP32 transforms (the fused z+y plane pass where a plane fits in LDS:
    NY = NZ = 32, 64, 128) against the FP64 oracle, relative 3e-6."""
    print(f"case fft fp32 oracle {shape}", flush=True)
    nx, ny, nz = shape
    rng = np.random.default_rng(7 + sum(shape))
    real = rng.standard_normal(shape)
    k = gpu(abi.make_params(shape, precision=abi.KAPSEL_FP32))
    o = oracle(abi.make_params(shape))
    for x in (k, o):
        x.set_field("U0", sp.pad_real(real))
        x.fft_r2c("U0")
    spec_o = o.get_field("U0")
    assert rel(k.get_field("U0"), spec_o) < 3e-6
    for x in (k, o):
        x.set_field("U1", spec_o)
        x.fft_c2r("U1")
    assert rel(k.get_field("U1")[:, :, :nz], o.get_field("U1")[:, :, :nz]) < 3e-6
    # three fields in one launch
    fields = ["U0", "U1", "U2"]
    reals = [rng.standard_normal(shape) for _ in fields]
    for f, r in zip(fields, reals):
        k.set_field(f, sp.pad_real(r))
    k.solenoidal_u("U0")  # r2c of 3 fields, projection, c2r of 3 fields
    for f, r in zip(fields, reals):
        o.set_field(f, sp.pad_real(r))
    o.solenoidal_u("U0")
    for f in fields:
        assert rel(k.get_field(f)[:, :, :nz], o.get_field(f)[:, :, :nz]) < 3e-6, f


@pytest.mark.parametrize("prec,tol", [(abi.KAPSEL_FP64, 1e-13), (abi.KAPSEL_FP32, 3e-6)])
def test_fft_round_trip_256(prec, tol):
    n = 256
    print(f"case fft round trip 256^3 prec={prec}", flush=True)
    rng = np.random.default_rng(99)
    real = rng.standard_normal((n, n, n))
    k = gpu(abi.make_params(n, precision=prec))
    k.set_field("U0", sp.pad_real(real))
    k.fft_r2c("U0")
    spec = sp.to_complex(k.get_field("U0"))
    # DC and Parseval are size-independent checks of the forward transform
    assert abs(spec[0, 0, 0] - real.sum()) < 1e-9 * n ** 3 if prec == 0 else True
    w = np.full(n // 2 + 1, 2.0)
    w[0] = w[-1] = 1.0
    lhs = np.sum(real ** 2)
    rhs = np.sum(w * np.abs(spec) ** 2) / n ** 3
    assert abs(lhs - rhs) < (1e-11 if prec == 0 else 1e-5) * lhs
    k.fft_c2r("U0")
    assert np.abs(k.get_field("U0")[:, :, :n] - real).max() < tol * 10


def _random_zeta(shape, seed):
    rng = np.random.default_rng(seed)
    u = sp.random_solenoidal(shape, rng, amp=0.05, dc=(1e-3, -2e-3, 5e-4))
    uk = np.fft.rfftn(u, axes=(1, 2, 3))
    # a full (untruncated) spectrum as U2zeta_k leaves it: add energy above the 2/3 band
    extra = rng.standard_normal(uk.shape) * 1e-3
    return u, uk, extra


def test_kspace_ops_match_oracle():
    shape = (64, 32, 64)
    print(f"case kspace ops {shape}", flush=True)
    p = abi.make_params(shape, eta=0.8)
    k, o = gpu(p), oracle(p)
    rng = np.random.default_rng(21)
    u = rng.standard_normal((3,) + shape) * 0.05
    for x in (k, o):
        for d in range(3):
            x.set_field(f"U{d}", sp.pad_real(u[d]))
            x.fft_r2c(f"U{d}")
        x.u_k2zeta_k()
    assert rel(k.get_uk_dc(), o.get_uk_dc()) < 1e-13
    for f in ("ZETA0", "ZETA1"):
        assert rel(k.get_field(f), o.get_field(f)) < 1e-13
    for x in (k, o):
        x.zeta_k2u_k()
    for d in range(3):
        assert rel(k.get_field(f"U{d}"), o.get_field(f"U{d}")) < 1e-13
    for x in (k, o):
        x.truncate("ZETA0")
    assert rel(k.get_field("ZETA0"), o.get_field("ZETA0")) < 1e-13
    # slaved Euler step on the full (untruncated) zeta
    t = abi.make_time(0.05)
    for x in (k, o):
        x.ns_slaved_euler(t)
    for f in ("ZETA0", "ZETA1", "WORK0", "WORK1"):
        assert rel(k.get_field(f), o.get_field(f)) < 1e-12, f
    # projection of a random real vector field
    v = rng.standard_normal((3,) + shape)
    for x in (k, o):
        for d in range(3):
            x.set_field(f"F{d}", sp.pad_real(v[d]))
        x.solenoidal_u("F0")
    for d in range(3):
        assert rel(k.get_field(f"F{d}")[:, :, :shape[2]], o.get_field(f"F{d}")[:, :, :shape[2]]) < 1e-13


def _particles_case(n, npart, a, seed):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(0, n, size=(npart, 3))
    pos[1] = (pos[0] + np.array([a * 1.3, 0.2, -0.4])) % n  # overlapping profiles
    pos[2] = np.array([0.3, n - 0.2, 0.9])                  # across the boundary
    vel = rng.normal(0, 1e-2, size=(npart, 3))
    om = rng.normal(0, 1e-3, size=(npart, 3))
    st = workloads.particle_state(pos, vel, om)
    st[abi.PFIELD["V_OLD"]] = (vel * 0.9).T
    st[abi.PFIELD["FR_PREVIOUS"]] = rng.normal(0, 1e-3, size=(3, npart))
    return st


@pytest.mark.parametrize("n,npart,a", [(32, 5, 3.0), (64, 40, 5.0), (128, 200, 4.0)])
def test_particle_ops_match_oracle(n, npart, a):
    print(f"case particle ops n={n} np={npart} a={a}", flush=True)
    p = abi.make_params(n, n_particles=npart, radius=a, xi=2.0, body_force=(0, 0, 1e-3))
    k, o = gpu(p), oracle(p)
    st = _particles_case(n, npart, a, n + npart)
    rng = np.random.default_rng(5)
    u = rng.normal(0, 1e-2, size=(3, n, n, n))
    for x in (k, o):
        x.set_particles(st)
        for d in range(3):
            x.set_field(f"U{d}", sp.pad_real(u[d]))
        x.make_phi_particle_sum()
        x.make_u_particle_sum()
        x.make_phi_particle("WORK2")
    for f in ("PHI_SUM", "PHI", "UP0", "UP1", "UP2", "WORK2"):
        assert rel(k.get_field(f)[:, :, :n], o.get_field(f)[:, :, :n]) < 1e-13, f
    t = abi.make_time(0.05)
    for x in (k, o):
        x.calc_f_hydro(t)
    pk, po = k.get_particles(), o.get_particles()
    for f in ("F_HYDRO", "TORQUE_HYDRO"):
        i = abi.PFIELD[f]
        assert rel(pk[i], po[i]) < 1e-12, f
    for x in (k, o):
        x.force()
    i = abi.PFIELD["FR"]
    assert rel(k.get_particles()[i], o.get_particles()[i]) < 1e-12
    # random unit orientations (q and q_old) for the quaternion update
    q = rng.standard_normal((8, npart))
    q[:4] /= np.linalg.norm(q[:4], axis=0)
    q[4:] /= np.linalg.norm(q[4:], axis=0)
    for ts in (0, 3):
        tt = abi.make_time(0.05, ts=ts)
        for x in (k, o):
            x.set_particles(st)
            x.set_orientations(q)
            x.md_position(tt)
            x.md_velocity(tt)
        pk, po = k.get_particles(), o.get_particles()
        for f in abi.PFIELDS:
            i = abi.PFIELD[f]
            assert np.abs(pk[i] - po[i]).max() <= 1e-12 * max(1e-30, np.abs(po[i]).max()), (ts, f)
        assert np.abs(k.get_orientations() - o.get_orientations()).max() < 1e-14, ts


def _compare_step_state(k, o, n, tol):
    worst = {}
    for f in ("ZETA0", "ZETA1"):
        worst[f] = rel(k.get_field(f), o.get_field(f))
    for f in ("U0", "U1", "U2", "PHI", "PHI_SUM", "UP0", "UP1", "UP2", "F0", "F1", "F2"):
        worst[f] = rel(k.get_field(f)[:, :, :n], o.get_field(f)[:, :, :n])
    worst["uk_dc"] = np.abs(k.get_uk_dc() - o.get_uk_dc()).max() / max(1e-30, np.abs(o.get_uk_dc()).max())
    pk, po = k.get_particles(), o.get_particles()
    for f in ("X", "V", "OMEGA", "F_HYDRO_PREVIOUS", "TORQUE_HYDRO_PREVIOUS", "FR_PREVIOUS"):
        i = abi.PFIELD[f]
        worst[f] = rel(pk[i], po[i])
    worst["orientation q, q_old"] = rel(k.get_orientations(), o.get_orientations())
    bad = {f: v for f, v in worst.items() if not v < tol}
    return worst, bad


def test_spm_step_fp64_matches_oracle_c1():
    """64^3, 8 spheres a=5 (configs[0] shape): init projection + one step at 1e-12,
    then 20 more steps, max relative deviation reported and bounded."""
    p, st, dt = workloads.spm_config("c1")
    n = p.nx
    print("case spm step c1 fp64", flush=True)
    k, o = gpu(p), oracle(p)
    for x in (k, o):
        x.set_particles(st)
        x.init_rigidity()
    worst, bad = _compare_step_state(k, o, n, 1e-12)
    assert not bad, bad
    tk, to = abi.make_time(dt), abi.make_time(dt)
    k.spm_step(tk)
    o.spm_step(to)
    assert tk.ts == to.ts == 1
    worst, bad = _compare_step_state(k, o, n, 1e-12)
    print("one-step worst", {f: f"{v:.1e}" for f, v in worst.items()}, flush=True)
    assert not bad, bad
    for _ in range(20):
        k.spm_step(tk)
        o.spm_step(to)
    worst, bad = _compare_step_state(k, o, n, 1e-9)
    print("21-step worst", {f: f"{v:.1e}" for f, v in worst.items()}, flush=True)
    assert not bad, bad


def test_spm_step_fp64_matches_oracle_c2():
    """128^3, 200 spheres a=4: one step at 1e-12."""
    p, st, dt = workloads.spm_config("c2")
    n = p.nx
    print("case spm step c2 fp64", flush=True)
    k, o = gpu(p), oracle(p)
    for x in (k, o):
        x.set_particles(st)
        x.init_rigidity()
    tk, to = abi.make_time(dt), abi.make_time(dt)
    k.spm_step(tk)
    o.spm_step(to)
    worst, bad = _compare_step_state(k, o, n, 1e-12)
    print("c2 one-step worst", {f: f"{v:.1e}" for f, v in worst.items()}, flush=True)
    assert not bad, bad


def test_spm_step_thermostat_matches_oracle():
    """64^3, 8 spheres, kBT = 0.4 (alpha_v = alpha_o = 1, seed 11): the same
    MT19937 stream drives both; 5 steps at 1e-12."""
    p, st, dt = workloads.spm_config("c1")
    n = p.nx
    print("case spm step thermostat c1 fp64", flush=True)
    k, o = gpu(p), oracle(p)
    for x in (k, o):
        x.set_particles(st)
        x.set_thermostat(0.4, 1.0, 1.0, 11)
        x.init_rigidity()
    tk, to = abi.make_time(dt), abi.make_time(dt)
    for _ in range(5):
        k.spm_step(tk)
        o.spm_step(to)
    worst, bad = _compare_step_state(k, o, n, 1e-12)
    print("thermostat 5-step worst", {f: f"{v:.1e}" for f, v in worst.items()}, flush=True)
    assert not bad, bad
    v = k.get_particles()[abi.PFIELD["V"]]
    assert np.abs(v).max() > 1e-3  # the kicks moved the spheres
