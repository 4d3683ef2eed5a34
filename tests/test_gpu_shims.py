"""The reference-signature shims (icep_kapsel_amd/shims) on the GPU against
the CPU oracle: each shim is called exactly as the reference's call sites call
the function it replaces (AoS Particle array, padded host fields, double**
vector fields, CTime by reference, unpadded Eigen-side vectors for Ax), on
64^3 with 6 spheres a=4 (two overlapping) and eps ratio 0.1.  Bars: profile,
particle velocity field, force/torque, permittivity 1e-13 (relative to the
largest value); A x and the right-hand side 1e-12; the solved potential 1e-8
(same CG, same tolerance 1e-10, constants removed)."""
import numpy as np
import pytest

from icep_kapsel_amd import abi, workloads
from tests import shim_binding as sb

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.abs(a - b).max() / max(1e-300, np.abs(b).max()))


def test_shims_match_oracle():
    from icep_kapsel_amd.gpu import KapselGPU
    from oracle.oracle import Oracle

    print("case reference-signature shims 64^3, 6 spheres", flush=True)
    n, a = 64, 4.0
    rng = np.random.default_rng(31)
    pos = workloads.random_positions(workloads.Xoshiro256(31), 6, (n, n, n), 2 * a)
    pos[1] = (pos[0] + np.array([7.0, 0.4, -0.3])) % n
    st = workloads.particle_state(pos, rng.normal(0, 1e-2, (6, 3)), rng.normal(0, 1e-3, (6, 3)))
    p = abi.make_params(n, n_particles=6, radius=a, xi=2.0, rotation=1, eps_fluid=1.0, eps_particle=0.1)
    k, o = KapselGPU(p), Oracle(p)
    lib, fn = sb.load()
    shape = k.shape
    eps = np.zeros(shape)  # the reference's global epsilon, filled by Make_dielectric_field
    fn["kapsel_shims_attach"](k.ctx, 6, 1, 1.0, 1e-10, 5000, sb.dptr(eps), abi.PRECOND_NONE)
    P = sb.particles_from_state(st)
    for pt in P:
        pt.v_slip[:] = [7.0, 7.0, 7.0]
        pt.f_slip[:] = [5.0, 5.0, 5.0]
    o.set_particles(st)

    # Make_phi_particle_sum (make_phi.h:51)
    phi, phi_sum = np.zeros(shape), np.zeros(shape)
    fn["Make_phi_particle_sum"](sb.dptr(phi), sb.dptr(phi_sum), P, a)
    o.make_phi_particle_sum()
    e = {"phi_sum": rel(phi_sum[:, :, :n], o.get_field("PHI_SUM")[:, :, :n]),
         "phi": rel(phi[:, :, :n], o.get_field("PHI")[:, :, :n])}
    assert phi_sum.max() > 1.0  # the overlapping pair

    # Make_u_particle_sum (make_phi.h:56)
    up = [np.zeros(shape) for _ in range(3)]
    fn["Make_u_particle_sum"](sb.dptrs(up), sb.dptr(phi_sum), P, a)
    o.make_u_particle_sum()
    e["up"] = max(rel(up[d][:, :, :n], o.get_field(f"UP{d}")[:, :, :n]) for d in range(3))

    # Calc_f_hydro_correct_precision (md_force.h:55)
    u = [np.zeros(shape) for _ in range(3)]
    for d in range(3):
        u[d][:, :, :n] = rng.normal(0, 1e-2, (n, n, n))
        o.set_field(f"U{d}", u[d])
    jikan = sb.CTime(3, 0.15, 0.05, 0.025, 0.05, 0.025)
    fn["Calc_f_hydro_correct_precision"](P, sb.dptr(phi_sum), sb.dptrs(u), jikan)
    o.calc_f_hydro(abi.make_time(0.05, ts=3, time=0.15))
    po = o.get_particles()
    e["f_hydro"] = rel(sb.member_array(P, "f_hydro"), po[abi.PFIELD["F_HYDRO"]].T)
    e["torque_hydro"] = rel(sb.member_array(P, "torque_hydro"), po[abi.PFIELD["TORQUE_HYDRO"]].T)
    assert np.all(sb.member_array(P, "f_slip") == 0.0) and np.all(sb.member_array(P, "torque_slip") == 0.0)
    assert np.all(sb.member_array(P, "v_slip") == 7.0)

    # Make_dielectric_field (operate_dielectric.h:126) into the global epsilon
    phi_p = np.zeros(shape)
    fn["Make_dielectric_field"](P, sb.dptr(eps), sb.dptr(phi_p), 0.1, 0.1, 1.0)
    o.make_dielectric_field()
    e["epsilon"] = rel(eps[:, :, :n], o.get_field("EPSILON")[:, :, :n])
    e["phi_sin"] = rel(phi_p[:, :, :n], o.get_field("WORK4")[:, :, :n])

    # Ax (operate_dielectric.h:193; MatrixReplacement mv_function, LinearOperator.h:32)
    x = rng.standard_normal(n ** 3)
    ax = np.zeros(n ** 3)
    fn["Ax"](sb.dptr(x), sb.dptr(ax))
    e["Ax"] = rel(ax, o.ax_unpadded(x))

    # Build_rhs (operate_dielectric.h:207)
    rho = np.zeros(shape)
    rho[:, :, :n] = rng.normal(0, 1e-3, (n, n, n))
    rho[:, :, :n] -= rho[:, :, :n].mean()
    e0 = np.array([0.02, -0.05, 0.1])
    rhs = np.zeros(shape)
    fn["Build_rhs"](sb.dptr(rhs), sb.dptr(rho), sb.dptr(eps), sb.dptr(e0.copy()))
    o.set_field("RHO_FREE", rho)
    o.poisson_rhs(e0)
    e["rhs"] = rel(rhs[:, :, :n], o.get_field("RHS")[:, :, :n])

    # Charge_field2potential_dielectric_LO (operate_dielectric.h:221)
    pot = np.zeros(shape)
    rhs2 = np.zeros(shape)
    fn["Charge_field2potential_dielectric_LO"](sb.dptr(rho), sb.dptr(pot), sb.dptr(eps), sb.dptr(rhs2),
                                               sb.dptr(e0.copy()))
    o.set_field("POTENTIAL", np.zeros(shape))
    it, rr = o.poisson_solve(e0, 1e-10, 5000, 0)
    xo = o.get_field("POTENTIAL")[:, :, :n]
    xs = pot[:, :, :n]
    e["potential"] = rel(xs - xs.mean(), xo - xo.mean())
    print({f: f"{v:.1e}" for f, v in e.items()}, flush=True)
    for f in ("phi_sum", "phi", "up", "f_hydro", "torque_hydro", "epsilon", "phi_sin"):
        assert e[f] < 1e-13, f
    assert e["Ax"] < 1e-12 and e["rhs"] < 1e-12
    assert e["potential"] < 1e-8
    k.close()
