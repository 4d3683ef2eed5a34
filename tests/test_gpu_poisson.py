"""GPU parity of the dielectric Poisson path against the CPU oracle.

FP64: eps field 1e-13, operator and RHS 1e-12 (relative to the max), solution
of the same solver 1e-8 after removing the null-space constant, true FP64
residual below tol.  FP32 context: the FP64-refined residual reaches 1e-8 and
1e-10 and the potential agrees with the FP64 solution to 1e-5.
"""
import numpy as np
import pytest

from icep_kapsel_amd import abi, workloads
from tests import spectral as sp

pytestmark = pytest.mark.gpu


def gpu(p):
    from icep_kapsel_amd.gpu import KapselGPU
    return KapselGPU(p)


def oracle(p):
    from oracle.oracle import Oracle
    return Oracle(p)


def rel(a, b):
    return np.abs(a - b).max() / max(1e-300, np.abs(b).max())


def _case(n, npart, a, ratio, seed, precision=abi.KAPSEL_FP64):
    p = abi.make_params(n, n_particles=npart, radius=a, xi=2.0, eps_fluid=1.0, eps_particle=ratio,
                        precision=precision)
    pos = workloads.random_positions(workloads.Xoshiro256(seed), npart, (n, n, n), 2 * a)
    return p, workloads.particle_state(pos)


def test_poisson_pieces_match_oracle():
    n = 64
    print("case poisson pieces 64^3", flush=True)
    p, st = _case(n, 6, 6.0, 0.1, 11)
    k, o = gpu(p), oracle(p)
    rng = np.random.default_rng(3)
    x = rng.standard_normal((n, n, n))
    rho = rng.standard_normal((n, n, n)) * 1e-3
    rho -= rho.mean()
    for s in (k, o):
        s.set_particles(st)
        s.make_dielectric_field()
        s.set_field("WORK3", sp.pad_real(x))
        s.poisson_apply("WORK3", "WORK2")
        s.set_field("RHO_FREE", sp.pad_real(rho))
        s.poisson_rhs([0.02, -0.05, 0.1])
    assert rel(k.get_field("EPSILON")[:, :, :n], o.get_field("EPSILON")[:, :, :n]) < 1e-13
    assert rel(k.get_field("WORK2")[:, :, :n], o.get_field("WORK2")[:, :, :n]) < 1e-12
    assert rel(k.get_field("RHS")[:, :, :n], o.get_field("RHS")[:, :, :n]) < 1e-12
    xu = rng.standard_normal(n ** 3)
    assert rel(k.ax_unpadded(xu), o.ax_unpadded(xu)) < 1e-12


@pytest.mark.parametrize("precond", [0, 1])
def test_poisson_solve_fp64_matches_oracle(precond):
    n = 64
    print(f"case poisson solve fp64 precond={precond}", flush=True)
    p, st = _case(n, 6, 6.0, 0.1, 12)
    k, o = gpu(p), oracle(p)
    e0 = [0.0, 0.0, 0.1]
    for s in (k, o):
        s.set_particles(st)
        s.make_dielectric_field()
    itk, rrk = k.poisson_solve(e0, 1e-10, 5000, precond)
    ito, rro = o.poisson_solve(e0, 1e-10, 5000, precond)
    print(f"iterations gpu {itk} oracle {ito}; relres gpu {rrk:.2e} oracle {rro:.2e}", flush=True)
    assert rrk < 1e-10 and abs(itk - ito) <= max(3, ito // 50)
    xk, xo = k.get_field("POTENTIAL")[:, :, :n], o.get_field("POTENTIAL")[:, :, :n]
    assert rel(xk - xk.mean(), xo - xo.mean()) < 1e-8


@pytest.mark.parametrize("n,tol", [(64, 1e-8), (64, 1e-10), (32, 1e-10)])
def test_poisson_solve_fp32_refined(n, tol):
    """64^3 runs the fused x passes of the correction and the residual, 32^3
    the separate gradient, face and divergence passes."""
    print(f"case poisson solve fp32 refined {n}^3 tol={tol}", flush=True)
    npart, a = (6, 6.0) if n == 64 else (2, 4.0)
    p64, st = _case(n, npart, a, 0.1, 13)
    p32, _ = _case(n, npart, a, 0.1, 13, precision=abi.KAPSEL_FP32)
    k64, k32 = gpu(p64), gpu(p32)
    e0 = [0.05, 0.0, 0.1]
    for s in (k64, k32):
        s.set_particles(st)
        s.make_dielectric_field()
    it64, rr64 = k64.poisson_solve(e0, 1e-12, 5000, 1)
    it32, rr32 = k32.poisson_solve(e0, tol, 5000, 1)
    print(f"fp32 refined: {it32} inner iterations, fp64 relres {rr32:.2e}", flush=True)
    assert rr32 < tol
    x64, x32 = k64.get_field("POTENTIAL")[:, :, :n], k32.get_field("POTENTIAL")[:, :, :n]
    assert rel(x32 - x32.mean(), x64 - x64.mean()) < 1e-5
    # solves in a row from the kept FP64 potential: the first residual already
    # meets the bound (no correction), then a changed field needs one
    it2, rr2 = k32.poisson_solve(e0, tol, 5000, 1)
    assert rr2 < tol and it2 == 0
    it3, rr3 = k32.poisson_solve([0.0, 0.08, 0.1], tol, 5000, 1)
    print(f"again: {it2} and {it3} inner iterations, relres {rr2:.2e}, {rr3:.2e}", flush=True)
    assert rr3 < tol and it3 > 0


@pytest.mark.parametrize("shape", [(128, 64, 32), (64, 32, 32)])
def test_poisson_solve_non_cubic_grid(shape):
    """128 x 64 x 32 grid: the fused x passes of the correction and of the FP64
    residual (chosen by NX) meet the separate y line and z face passes (no
    plane pass for NY != NZ).  64 x 32 x 32: the fused x passes with the
    y-split gradient fields and the half-plane round trip (NX != NY).  FP64
    solves (plain and spectral) against the oracle; the FP32 refined solve
    against the FP64 solution."""
    print(f"case poisson non-cubic {shape}", flush=True)
    nz = shape[2]
    pos = workloads.random_positions(workloads.Xoshiro256(17), 5, shape, 8.0)
    e0 = [0.03, -0.04, 0.1]
    sols = {}
    for prec in (abi.KAPSEL_FP64, abi.KAPSEL_FP32):
        p = abi.make_params(shape, n_particles=5, radius=4.0, xi=2.0, eps_fluid=1.0, eps_particle=0.1,
                            precision=prec)
        k = gpu(p)
        k.set_particles(workloads.particle_state(pos))
        k.make_dielectric_field()
        if prec == abi.KAPSEL_FP64:
            o = oracle(p)
            o.set_particles(workloads.particle_state(pos))
            o.make_dielectric_field()
            for precond in (0, 1):
                for s in (k, o):
                    s.set_field("POTENTIAL", np.zeros(k.shape))
                itk, rrk = k.poisson_solve(e0, 1e-10, 5000, precond)
                ito, rro = o.poisson_solve(e0, 1e-10, 5000, precond)
                print(f"fp64 precond={precond}: gpu {itk} it ({rrk:.1e}), oracle {ito} it ({rro:.1e})", flush=True)
                assert rrk < 1e-10 and abs(itk - ito) <= max(3, ito // 50)
                xk, xo = k.get_field("POTENTIAL")[:, :, :nz], o.get_field("POTENTIAL")[:, :, :nz]
                assert rel(xk - xk.mean(), xo - xo.mean()) < 1e-8
            sols[prec] = xk
        else:
            it, rr = k.poisson_solve(e0, 1e-10, 5000, 1)
            print(f"fp32 refined: {it} inner iterations, fp64 relres {rr:.1e}", flush=True)
            assert rr < 1e-10
            x = k.get_field("POTENTIAL")[:, :, :nz]
            x64 = sols[abi.KAPSEL_FP64]
            assert rel(x - x.mean(), x64 - x64.mean()) < 1e-5
        k.close()


def test_poisson_solve_128_dielectric_spheres():
    """128^3, 64 spheres a=6, eps ratio 0.1 (the ICEP configuration): FP64
    plain CG and spectral PCG both reach 1e-10 and agree."""
    n = 128
    print("case poisson 128^3 64 spheres", flush=True)
    p, st = _case(n, 64, 6.0, 0.1, 3003)
    k = gpu(p)
    k.set_particles(st)
    k.make_dielectric_field()
    e0 = [0.0, 0.0, 0.1]
    it1, rr1 = k.poisson_solve(e0, 1e-10, 20000, 1)
    x1 = k.get_field("POTENTIAL")[:, :, :n]
    k.set_field("POTENTIAL", np.zeros(k.shape))
    it0, rr0 = k.poisson_solve(e0, 1e-10, 20000, 0)
    x0 = k.get_field("POTENTIAL")[:, :, :n]
    print(f"plain CG {it0} it ({rr0:.1e}), spectral PCG {it1} it ({rr1:.1e})", flush=True)
    assert rr0 < 1e-10 and rr1 < 1e-10
    assert it1 * 5 < it0
    assert rel(x1 - x1.mean(), x0 - x0.mean()) < 1e-8


def test_maxwell_force_matches_oracle():
    n = 64
    print("case maxwell force 64^3", flush=True)
    p, st = _case(n, 6, 6.0, 0.1, 21)
    k, o = gpu(p), oracle(p)
    e0 = [0.02, 0.0, 0.1]
    o.set_particles(st)
    o.make_dielectric_field()
    o.poisson_solve(e0, 1e-10, 5000, 1)
    rng = np.random.default_rng(8)
    rho = rng.standard_normal((n, n, n)) * 1e-3
    for s in (k, o):
        s.set_particles(st)
        s.make_dielectric_field()
        s.set_field("POTENTIAL", o.get_field("POTENTIAL"))
        s.set_field("RHO_FREE", sp.pad_real(rho))
        s.maxwell_force(e0)
    for f in ("F0", "F1", "F2"):
        assert rel(k.get_field(f)[:, :, :n], o.get_field(f)[:, :, :n]) < 1e-12, f


def test_icep_step_fp64_matches_oracle():
    """Two ICEP steps (64^3, 6 dielectric spheres a=6): fields and particles
    agree to 1e-9 (the solves stop at 1e-12, not at bit-identical iterates)."""
    n = 64
    print("case icep step 64^3 fp64", flush=True)
    p, st = _case(n, 6, 6.0, 0.1, 22)
    k, o = gpu(p), oracle(p)
    e0 = [0.0, 0.05, 0.1]
    for s in (k, o):
        s.set_particles(st)
        s.init_rigidity()
    tk, to = abi.make_time(0.05), abi.make_time(0.05)
    for _ in range(2):
        itk, rrk = k.icep_step(tk, e0, 1e-12, 5000, 1)
        ito, rro = o.icep_step(to, e0, 1e-12, 5000, 1)
        print(f"icep step: gpu {itk} it {rrk:.1e}, oracle {ito} it {rro:.1e}", flush=True)
        assert rrk < 1e-12
    for f in ("ZETA0", "ZETA1", "EPSILON", "F0", "F1", "F2"):
        assert rel(k.get_field(f), o.get_field(f)) < 1e-9, f
    xk, xo = k.get_field("POTENTIAL")[:, :, :n], o.get_field("POTENTIAL")[:, :, :n]
    assert rel(xk - xk.mean(), xo - xo.mean()) < 1e-9
    pk, po = k.get_particles(), o.get_particles()
    for f in ("X", "V", "OMEGA"):
        i = abi.PFIELD[f]
        assert rel(pk[i], po[i]) < 1e-9, f


@pytest.mark.parametrize("prec,n,npart", [(abi.KAPSEL_FP64, 64, 6), (abi.KAPSEL_FP32, 64, 6),
                                           (abi.KAPSEL_FP32, 128, 64)])
def test_icep_step_two_streams_bit_identical(prec, n, npart):
    """kapsel_icep_step with its Navier-Stokes part on the second stream
    (default) against the one-stream order: four steps, every field and the
    particle state bit-identical (same kernels on the same inputs; a race
    between the two chains would show as a difference)."""
    print(f"case icep two streams vs one, {n}^3 prec={prec}", flush=True)
    runs = []
    for concurrent in (True, False):
        p, st = _case(n, npart, 6.0, 0.1, 23, precision=prec)
        k = gpu(p)
        k.set_concurrency(concurrent)
        k.set_particles(st)
        k.init_rigidity()
        t = abi.make_time(0.05)
        its = []
        for _ in range(4):
            it, rr = k.icep_step(t, [0.02, 0.05, 0.1], 1e-8, 20000, 1)
            its.append((it, rr))
        runs.append((its, {f: k.get_field(f) for f in ("ZETA0", "ZETA1", "U0", "U1", "U2", "POTENTIAL", "PHI")},
                     k.get_particles(), k.get_uk_dc()))
        k.close()
    (i1, f1, p1, u1), (i2, f2, p2, u2) = runs
    print("iterations / residuals", i1, i2, flush=True)
    assert i1 == i2
    for f in f1:
        assert np.array_equal(f1[f], f2[f]), f
    assert np.array_equal(p1, p2) and np.array_equal(u1, u2)


def test_icep_step_fp32_outputs_made_on_read():
    """An FP32 ICEP step leaves the real POTENTIAL and RHS to be made when the
    API next reads or uses them (its own later parts read the FP64 spectra):
    after three steps the RHS read back equals Build_rhs recomputed in FP32
    from the same permittivity, and the potential read back is the solution:
    the FP64 operator leaves a residual of 1e-5 relative on it, what rounding
    the FP64 solution to FP32 (6e-8 relative) does under A, whose symbol grows
    as |k|^2 up to the grid scale (a wrong or stale potential would leave
    order 1)."""
    n = 64
    print("case icep fp32 outputs on read", flush=True)
    p, st = _case(n, 6, 6.0, 0.1, 24, precision=abi.KAPSEL_FP32)
    k = gpu(p)
    k.set_particles(st)
    k.init_rigidity()
    t = abi.make_time(0.05)
    e0 = [0.02, 0.05, 0.1]
    for _ in range(3):
        k.icep_step(t, e0, 1e-8, 20000, 1)
    rhs_made = k.get_field("RHS")[:, :, :n]
    pot = k.get_field("POTENTIAL")
    eps = k.get_field("EPSILON")
    k.poisson_rhs(e0)
    rhs_fresh = k.get_field("RHS")[:, :, :n]
    assert rel(rhs_made, rhs_fresh) < 1e-5
    p64, _ = _case(n, 6, 6.0, 0.1, 24)
    k64 = gpu(p64)
    k64.set_field("EPSILON", eps)
    k64.set_field("POTENTIAL", pot)
    k64.poisson_apply("POTENTIAL", "WORK0")
    k64.poisson_rhs(e0)
    r = k64.get_field("WORK0")[:, :, :n] - k64.get_field("RHS")[:, :, :n]
    relres = np.linalg.norm(r) / np.linalg.norm(k64.get_field("RHS")[:, :, :n])
    print(f"FP32 potential, FP64 residual {relres:.2e}", flush=True)
    assert relres < 1e-4


def test_fp32_solve_repeatable_with_and_without_timing():
    """Repeated FP32-context solves from the same start give bit-identical
    iterates (fixed-order reductions), with kernel timing events on or off
    (the events bracket launches; they must not change what runs)."""
    n = 64
    print("case fp32 solve repeated, timing on and off 64^3", flush=True)
    p, st = _case(n, 6, 6.0, 0.1, 14, precision=abi.KAPSEL_FP32)
    k = gpu(p)
    k.set_particles(st)
    k.make_dielectric_field()
    e0 = [0.0, 0.03, 0.1]
    out = []
    for mode in ("plain", "timed", "plain", "timed", "plain"):
        k.kernel_timing(["POINTWISE"] if mode == "timed" else [])
        k.set_field("POTENTIAL", np.zeros(k.shape))
        it, rr = k.poisson_solve(e0, 1e-10, 5000, 1)
        out.append((it, rr, k.get_field("POTENTIAL")[:, :, :n]))
        print(f"{mode}: {it} iterations, relres {rr:.3e}", flush=True)
    k.kernel_timing([])
    for it, rr, x in out[1:]:
        assert it == out[0][0] and rr == out[0][1]
        assert np.array_equal(x, out[0][2])
    assert out[0][1] < 1e-10
