"""Restart parity on the GPU: a run saved to a UDF restart record and resumed
in a fresh context continues as the uninterrupted run does.

SPM step (64^3, 8 spheres, rotation on): FP64, deviation below 1e-14 after
5 resumed steps.  Electrolyte step (64^3, 8 charged spheres): FP64, below
1e-12 (the resumed run recounts the ion totals from the saved
concentrations and recomputes what an uninterrupted run reuses).
"""
import numpy as np
import pytest

from icep_kapsel_amd import abi, workloads

pytestmark = pytest.mark.gpu


def gpu(p):
    from icep_kapsel_amd.gpu import KapselGPU
    return KapselGPU(p)


def rel(a, b):
    s = np.abs(b).max()
    return float(np.abs(a - b).max() / (s if s > 0 else 1.0))


def _compare(a, b, fields):
    e = {f: rel(a.get_field(f), b.get_field(f)) for f in fields}
    pa, pb = a.get_particles(), b.get_particles()
    for f in ("X", "X_NOPBC", "V", "OMEGA", "F_HYDRO_PREVIOUS", "TORQUE_HYDRO_PREVIOUS", "FR_PREVIOUS"):
        e[f] = rel(pa[abi.PFIELD[f]], pb[abi.PFIELD[f]])
    e["orientation"] = rel(a.get_orientations(), b.get_orientations())
    e["uk_dc"] = float(np.abs(a.get_uk_dc() - b.get_uk_dc()).max())
    return e


def test_spm_restart_continues_exactly(tmp_path):
    print("case restart spm c1 64^3", flush=True)
    p, st, dt = workloads.spm_config("c1")
    k = gpu(p)
    k.set_particles(st)
    k.init_rigidity()
    t = abi.make_time(dt)
    for _ in range(5):
        k.spm_step(t)
    path = tmp_path / "spm_restart.udf"
    k.save_restart(path, t)
    r = gpu(p)
    tr = abi.make_time(dt)
    r.load_restart(path, tr)
    assert (tr.ts, tr.time) == (t.ts, t.time)
    for _ in range(5):
        k.spm_step(t)
        r.spm_step(tr)
    e = _compare(k, r, ("ZETA0", "ZETA1"))
    print({f: f"{v:.1e}" for f, v in e.items()}, flush=True)
    assert max(e.values()) < 1e-14, e


def test_electrolyte_restart_continues(tmp_path):
    print("case restart electrolyte 64^3", flush=True)
    rng = workloads.Xoshiro256(23)
    n, npart = 64, 8
    p = abi.make_params(n, n_particles=npart, radius=4.0, xi=2.0)
    pos = workloads.random_positions(rng, npart, (n, n, n), 8.0)
    el = abi.make_electrolyte(surface_charge=60.0, debye_length=4.0, e_ext=(0.05, 0.0, 0.02))
    dt = workloads.electrolyte_auto_dt(p, el)
    k = gpu(p)
    k.set_particles(workloads.particle_state(pos))
    k.set_electrolyte(el)
    k.init_rigidity()
    k.init_rho_ion()
    t = abi.make_time(dt)
    for _ in range(4):
        k.electrolyte_step(t)
    path = tmp_path / "el_restart.udf"
    k.save_restart(path, t)
    r = gpu(p)
    r.set_electrolyte(el)
    tr = abi.make_time(dt)
    r.load_restart(path, tr)
    tot_k, tot_r = k.get_total_solute(), r.get_total_solute()
    assert np.abs(tot_r - tot_k).max() < 1e-12 * tot_k.max()
    for _ in range(4):
        k.electrolyte_step(t)
        r.electrolyte_step(tr)
    e = _compare(k, r, ("ZETA0", "ZETA1", "CONC0", "CONC1"))
    print({f: f"{v:.1e}" for f, v in e.items()}, flush=True)
    assert max(e.values()) < 1e-12, e
