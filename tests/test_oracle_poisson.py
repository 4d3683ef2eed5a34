"""Pins the oracle's dielectric Poisson pieces (operate_dielectric.cxx:57-572).

operator  -> constant eps: A cos(k.x) = -eps |k|^2 cos(k.x) (the staggered
             gradient/divergence pair has symbol -k^2 exactly); symmetric;
             negative semidefinite; constants in the null space
rhs       -> eps = e0 + d cos(2 pi m z / N), E0 along z: the face average at
             z + 1/2 followed by the staggered divergence back to z gives
             -E d cos(pi m / N) (2 pi m / N) sin(2 pi m z / N) exactly
eps field -> compact-sine profile (profile.h:139-152) summed by brute force
solve     -> residual below tol; CG, preconditioned CG and the reference's
             BiCGSTAB agree on the solution (up to the null-space constant)
"""
import numpy as np

from icep_kapsel_amd import abi, workloads
from tests import spectral as sp


def _oracle(p):
    from oracle.oracle import Oracle
    return Oracle(p)


def _cos_mode(n, m):
    x = np.arange(n)
    return np.cos(2 * np.pi * (m[0] * x[:, None, None] + m[1] * x[None, :, None] + m[2] * x[None, None, :]) / n)


def test_operator_constant_eps_symbol():
    n = 32
    o = _oracle(abi.make_params(n, eps_fluid=2.5, eps_particle=2.5))
    o.make_dielectric_field()
    for m in [(1, 2, 3), (0, 0, 5), (15, -7, 4), (16, 0, 0)]:
        f = _cos_mode(n, m)
        o.set_field("WORK0", sp.pad_real(f))
        o.poisson_apply("WORK0", "WORK1")
        got = o.get_field("WORK1")[:, :, :n]
        k2 = sum((2 * np.pi * mm / n) ** 2 for mm in m)
        assert np.abs(got + 2.5 * k2 * f).max() < 1e-12 * max(1.0, 2.5 * k2), m


def _sphere_setup(n=32, npart=3, a=4.0, ratio=0.1, seed=3):
    p = abi.make_params(n, n_particles=npart, radius=a, xi=2.0, eps_fluid=1.0, eps_particle=ratio)
    pos = workloads.random_positions(workloads.Xoshiro256(seed), npart, (n, n, n), 2 * a)
    return p, workloads.particle_state(pos), pos


def phi_sin(r, a, xi):
    d = a - r
    out = np.where(d > xi / 2, 1.0, 0.0)
    band = np.abs(d) <= xi / 2
    return np.where(band, 0.5 * np.sin(np.pi * d / xi) + 0.5, out)


def test_dielectric_field_brute_force():
    n, a = 32, 4.0
    p, st, pos = _sphere_setup(n, 3, a, 0.1)
    o = _oracle(p)
    o.set_particles(st)
    o.make_dielectric_field()
    g = np.arange(n, dtype=float)
    G = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1)
    acc = np.zeros((n, n, n))
    for x in pos:
        d = G - x
        d -= n * np.rint(d / n)
        acc += phi_sin(np.sqrt((d * d).sum(-1)), a, 2.0)
    eps = 1.0 + (0.1 - 1.0) * acc
    assert np.abs(o.get_field("EPSILON")[:, :, :n] - eps).max() < 1e-13


def test_operator_symmetric_semidefinite_nullspace():
    n = 32
    p, st, _ = _sphere_setup(n)
    o = _oracle(p)
    o.set_particles(st)
    o.make_dielectric_field()
    rng = np.random.default_rng(1)
    a, b = rng.standard_normal(n ** 3), rng.standard_normal(n ** 3)
    Aa, Ab = o.ax_unpadded(a), o.ax_unpadded(b)
    assert abs(a @ Ab - b @ Aa) < 1e-13 * abs(a @ Ab)
    assert a @ Aa < 0 and b @ Ab < 0
    assert np.abs(o.ax_unpadded(np.ones(n ** 3))).max() < 1e-12


def test_rhs_closed_form():
    n, m, e0, d, E = 32, 3, 1.3, 0.4, 0.7
    o = _oracle(abi.make_params(n))
    z = np.arange(n)
    eps = np.broadcast_to(e0 + d * np.cos(2 * np.pi * m * z / n), (n, n, n))
    o.set_field("EPSILON", sp.pad_real(eps))
    rng = np.random.default_rng(2)
    rho = rng.standard_normal((n, n, n)) * 1e-3
    o.set_field("RHO_FREE", sp.pad_real(rho))
    o.poisson_rhs([0.0, 0.0, E])
    got = o.get_field("RHS")[:, :, :n]
    expect = -E * d * np.cos(np.pi * m / n) * (2 * np.pi * m / n) * np.sin(2 * np.pi * m * z / n) - rho
    assert np.abs(got - np.broadcast_to(expect, got.shape)).max() < 1e-13


def test_solvers_agree_and_converge():
    n = 32
    p, st, _ = _sphere_setup(n, 4, 4.0, 0.1, seed=5)
    sols = {}
    for solver in (0, 1, 2):
        o = _oracle(p)
        o.set_particles(st)
        o.make_dielectric_field()
        it, rr = o.poisson_solve([0.0, 0.03, 0.1], 1e-10, 5000, solver)
        assert rr < 1e-10 and it > 0
        x = o.get_field("POTENTIAL")[:, :, :n]
        sols[solver] = (x - x.mean(), it)
        # the solution satisfies A x = b
        o.poisson_apply("POTENTIAL", "WORK0")
        ax = o.get_field("WORK0")[:, :, :n]
        b = o.get_field("RHS")[:, :, :n]
        assert np.linalg.norm(ax - b) <= 1.01e-10 * np.linalg.norm(b)
    scale = np.abs(sols[0][0]).max()
    assert np.abs(sols[1][0] - sols[0][0]).max() < 1e-8 * scale
    assert np.abs(sols[2][0] - sols[0][0]).max() < 1e-8 * scale
    assert sols[1][1] < sols[0][1] / 4  # the spectral preconditioner cuts iterations


def test_maxwell_force_closed_forms():
    """f = rho E - E^2 grad(eps) / 2 on the staggered grid, then averaged to the
    cell centres (operate_dielectric.cxx:646-763), in two closed forms."""
    n, m, d, e0z = 32, 3, 0.4, 0.7
    z = np.arange(n)
    o = _oracle(abi.make_params(n))
    # (1) zero potential, eps = 1.3 + d cos(2 pi m z / N), E0 along z:
    #     f_z = (E0^2 / 2) d (2 pi m / N) sin(2 pi m z / N) cos(pi m / N)
    eps = np.broadcast_to(1.3 + d * np.cos(2 * np.pi * m * z / n), (n, n, n))
    o.set_field("EPSILON", sp.pad_real(eps))
    o.set_field("POTENTIAL", np.zeros(o.shape))
    o.set_field("RHO_FREE", np.zeros(o.shape))
    o.maxwell_force([0.0, 0.0, e0z])
    fz = o.get_field("F2")[:, :, :n]
    expect = 0.5 * e0z ** 2 * d * (2 * np.pi * m / n) * np.sin(2 * np.pi * m * z / n) * np.cos(np.pi * m / n)
    assert np.abs(fz - np.broadcast_to(expect, fz.shape)).max() < 1e-13
    assert np.abs(o.get_field("F0")[:, :, :n]).max() < 1e-13
    # (2) uniform eps, charge rho(z): f_z = (rho(z+1) + 2 rho(z) + rho(z-1)) / 4 * E0
    rng = np.random.default_rng(4)
    rho1 = rng.standard_normal(n)
    o.set_field("EPSILON", np.full(o.shape, 2.0))
    o.set_field("RHO_FREE", sp.pad_real(np.broadcast_to(rho1, (n, n, n))))
    o.maxwell_force([0.0, 0.0, e0z])
    fz = o.get_field("F2")[:, :, :n]
    expect = 0.25 * (np.roll(rho1, -1) + 2 * rho1 + np.roll(rho1, 1)) * e0z
    assert np.abs(fz - np.broadcast_to(expect, fz.shape)).max() < 1e-13


def test_icep_step_runs_and_keeps_solenoidal_velocity():
    n = 32
    p = abi.make_params(n, n_particles=3, radius=4.0, xi=2.0, eps_fluid=1.0, eps_particle=0.1)
    pos = workloads.random_positions(workloads.Xoshiro256(9), 3, (n, n, n), 8.0)
    o = _oracle(p)
    o.set_particles(workloads.particle_state(pos))
    o.init_rigidity()
    t = abi.make_time(0.05)
    for _ in range(2):
        it, rr = o.icep_step(t, [0.0, 0.0, 0.1], 1e-10, 5000, 1)
        assert rr < 1e-10 and it > 0
    # U2zeta_k leaves u in reciprocal space (operate_omega.h:203-215)
    uk = np.stack([sp.to_complex(o.get_field(f"U{d}")) for d in range(3)])
    u = uk
    _, (kx, ky, kz) = sp.wavevectors((n, n, n))
    div = kx * uk[0] + ky * uk[1] + kz * uk[2]
    assert np.abs(div).max() < 1e-10 * max(1e-30, np.abs(uk).max())
    # the dielectric force moved the fluid (it is zero without a field gradient)
    assert np.abs(u).max() > 0
