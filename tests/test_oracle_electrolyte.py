"""Pins the oracle's electrolyte step (two ion species, Coulomb force,
Nernst-Planck) against closed forms, invariants and independent numpy
mathematics.  The reference's electrolyte sources do not build here (they
include Eigen through fluid_solver.h), and its tests hold no fixtures for this
path.

make_phi_qq        -> brute-force profile sum; the surface density is minus the
                      radial derivative of the compact-sine profile (numerical
                      derivative) scaled so that sum |surface| dx^3 = N |Q|
surface normal     -> unit radial vector on the band |r - a| < xi/2, 0 elsewhere
init_rho_ion       -> the box is electroneutral; the salt densities give back
                      the Debye length (kappa^2 = e^2 sum z^2 c / (eps kBT));
                      the totals are the (1 - phi)-weighted amounts
Coulomb force      -> numpy Poisson solve (psi = rho / (eps k^2)) and
                      f = rho (-grad psi + E); the self-force sums to zero
nssolute_euler     -> numpy restatement of the Nernst-Planck equation in
                      flux form (diffusion, electromigration, impermeable
                      band, advection) and of the vorticity equation
                      (curl of the advective form) with a forward Euler step
                      on the kept modes; a shear wave decays by 1 - nu k^2 dt
rescale_solute     -> the (1 - phi)-weighted amounts equal the totals after it
electrolyte_step   -> electroneutrality and the totals hold over steps
Dielectric branch  -> the gradient of psi = cos(k x) on the faces averaged onto
                      the cells is -k sin(k x) cos(k dx / 2) - E_x; the bulk
                      densities give back the Debye length with eps_fluid
                      (with and without surface charge); the free charge is
                      sum z e c (1 - phi_sin) over the brute-force profile;
                      the solve meets A psi = b + mean(rho); a neutral box
                      stays at rest; the ion amounts are step invariants
"""
import numpy as np
import pytest

from icep_kapsel_amd import abi, workloads
from tests import spectral as sp
from tests.test_oracle_kspace import _numpy_curl_advection, _reduce_rule
from tests.test_oracle_particles import grid_vectors, phi_profile


def _oracle(p, el):
    from oracle.oracle import Oracle
    o = Oracle(p)
    o.set_electrolyte(el)
    return o


def _setup(n=32, npart=3, a=4.0, xi=2.0, q=40.0, seed=11, valency=(1.0, -1.0), e_ext=(0.0, 0.0, 0.0),
           debye=4.0, dielectric=0, **kw):
    rng = workloads.Xoshiro256(seed)
    p = abi.make_params(n, n_particles=npart, radius=a, xi=xi, **kw)
    pos = workloads.random_positions(rng, npart, (n, n, n), 2.0 * a)
    el = abi.make_electrolyte(surface_charge=q, valency=valency, debye_length=debye, e_ext=e_ext,
                              dielectric=dielectric)
    return p, pos, el


def _phi_sin(r, a, xi):
    x = a - r
    return np.where(np.abs(x) <= xi / 2, 0.5 * np.sin(np.pi * x / xi) + 0.5, np.where(x > xi / 2, 1.0, 0.0))


def _brute(pos, n, a, xi):
    """phi (unclamped), raw -d phi_sin / dr (numerical), band normals; (n, n, n) arrays."""
    phi = np.zeros((n, n, n))
    dphi = np.zeros((n, n, n))
    nrm = np.zeros((3, n, n, n))
    h = 1e-5
    for xp in pos:
        r = grid_vectors(n, xp)
        d = np.sqrt((r * r).sum(-1))
        phi += phi_profile(d, a, xi)
        inside = np.abs(d - a) < xi / 2
        dphi += np.where(inside, -(_phi_sin(d + h, a, xi) - _phi_sin(d - h, a, xi)) / (2 * h), 0.0)
        safe = np.where(d > 0, d, 1.0)
        for c in range(3):
            nrm[c] += np.where(inside, r[..., c] / safe, 0.0)
    return phi, dphi, nrm


def _real(o, name, n):
    return o.get_field(name)[:, :, :n]


def test_phi_qq_profile_and_surface_normalisation():
    n, a, xi, q = 32, 4.0, 2.0, 40.0
    p, pos, el = _setup(n=n, a=a, xi=xi, q=q)
    o = _oracle(p, el)
    o.set_particles(workloads.particle_state(pos))
    o.make_phi_qq("WORK0", "WORK1")
    phi, dphi, _ = _brute(pos, n, a, xi)
    surf = _real(o, "WORK1", n)
    assert np.abs(_real(o, "WORK0", n) - phi).max() < 1e-13
    # sum |surface| dx^3 = N |Q| e
    assert abs(np.abs(surf).sum() - len(pos) * abs(q)) < 1e-11 * len(pos) * abs(q)
    expect = q * dphi * (len(pos) * abs(q) / np.abs(q * dphi).sum())
    assert np.abs(surf - expect).max() < 1e-8 * np.abs(expect).max()  # central-difference error


def test_surface_normal_unit_radial_band():
    n, a, xi = 32, 5.0, 2.0
    p, pos, el = _setup(n=n, npart=2, a=a, xi=xi)
    o = _oracle(p, el)
    o.set_particles(workloads.particle_state(pos))
    o.make_surface_normal("WORK0")
    got = np.stack([_real(o, f"WORK{d}", n) for d in range(3)])
    _, _, nrm = _brute(pos, n, a, xi)
    assert np.abs(got - nrm).max() < 1e-13
    mag = np.sqrt((got * got).sum(0))
    band = mag > 0
    assert band.sum() > 100
    # particles 2a * 1.05 apart at least: bands do not overlap, unit length
    assert np.abs(mag[band] - 1.0).max() < 1e-13


@pytest.mark.parametrize("q,valency", [(40.0, (1.0, -1.0)), (-25.0, (2.0, -1.0)), (30.0, (1.0, -2.0))])
def test_init_rho_ion_neutral_and_debye_length(q, valency):
    n, a, xi, debye = 32, 4.0, 2.0, 4.0
    p, pos, el = _setup(n=n, a=a, xi=xi, q=q, valency=valency, debye=debye)
    o = _oracle(p, el)
    o.set_particles(workloads.particle_state(pos))
    o.init_rho_ion()
    phi, _, _ = _brute(pos, n, a, xi)
    c = []
    for s in (0, 1):
        ck = o.get_field(f"CONC{s}")
        cc = sp.to_complex(ck)
        # uniform fields: only the k = 0 mode
        cc0 = cc[0, 0, 0].real
        rest = cc.copy()
        rest[0, 0, 0] = 0
        assert np.abs(rest).max() < 1e-12 * abs(cc0)
        c.append(cc0 / n ** 3)
    vol = (1.0 - phi).sum()
    counter = len(pos) * abs(q) / abs(valency[0] if q < 0 else valency[1]) / vol
    salt = [c[0] - (counter if q < 0 else 0.0), c[1] - (counter if q > 0 else 0.0)]
    kappa2 = (valency[0] ** 2 * salt[0] + valency[1] ** 2 * salt[1])  # e = eps = kBT = 1
    assert abs(kappa2 - 1.0 / debye ** 2) < 1e-13
    assert abs(valency[0] * salt[0] + valency[1] * salt[1]) < 1e-15  # salt is neutral
    tot = o.get_total_solute()
    for s in (0, 1):
        assert abs(tot[s] - c[s] * vol) < 1e-12 * tot[s]
    # the whole box is electroneutral: surface charge + ions
    o.charge_field("WORK0")
    rho = _real(o, "WORK0", n)
    assert abs(rho.sum()) < 1e-11 * len(pos) * abs(q)


def _random_conc(n, rng, base, amp, kmax=5):
    """Positive smooth field band-limited to |k_i| <= kmax (below the 2/3 cut)."""
    f = rng.standard_normal((n, n, n))
    fk = np.fft.rfftn(f) * sp.band_mask((n, n, n), kmax)
    f = np.fft.irfftn(fk, s=(n, n, n), axes=(0, 1, 2))
    f = base + amp * f / np.abs(f).max()
    return f


def _kspace_conc(c):
    return sp.from_complex(np.fft.rfftn(c))


def _numpy_charge(o, n, el, conc_real):
    o.make_phi_qq("WORK3", "WORK4")
    phi_qq, surf = _real(o, "WORK3", n), _real(o, "WORK4", n)
    rho = surf.copy()
    for s in (0, 1):
        rho += el.valency[s] * el.elementary_charge * conc_real[s] * (1.0 - phi_qq)
    return rho


def _grad_psi(rho, shape, eps, trunc):
    rk = np.fft.rfftn(rho)
    _, (kx, ky, kz) = sp.wavevectors(shape)
    k2 = kx * kx + ky * ky + kz * kz
    pk = np.where(k2 > 0, rk / (eps * np.where(k2 > 0, k2, 1.0)), 0.0)
    if trunc:
        pk = pk * sp.kept_mask(shape)
    return [np.fft.irfftn(1j * k * pk, s=shape, axes=(0, 1, 2)) for k in (kx, ky, kz)]


@pytest.mark.parametrize("ac", [0, 1])
def test_coulomb_force_numpy_poisson(ac):
    n = 32
    e_ext = (0.03, -0.01, 0.02)
    p, pos, el = _setup(n=n, e_ext=e_ext)
    el.dielectric_cst = 1.7
    el.ac, el.angular_frequency = ac, 0.9
    o = _oracle(p, el)
    o.set_particles(workloads.particle_state(pos))
    rng = np.random.default_rng(3)
    conc = [_random_conc(n, rng, 0.05, 0.02), _random_conc(n, rng, 0.04, 0.02)]
    for s in (0, 1):
        o.set_field(f"CONC{s}", _kspace_conc(conc[s]))
    t = abi.make_time(0.01, ts=3, time=0.7)
    o.coulomb_force(t)
    rho = _numpy_charge(o, n, el, conc)
    g = _grad_psi(rho, (n, n, n), el.dielectric_cst, trunc=False)
    s = np.sin(0.9 * 0.7) if ac else 1.0
    f = np.stack([_real(o, f"F{d}", n) for d in range(3)])
    expect = np.stack([rho * (-g[d] + e_ext[d] * s) for d in range(3)])
    assert np.abs(f - expect).max() < 1e-12 * np.abs(expect).max()
    # the self-field exerts no net force: sum f = E sum rho
    for d in range(3):
        assert abs(f[d].sum() - e_ext[d] * s * rho.sum()) < 1e-11 * np.abs(expect).max() * n ** 1.5


def _shear_zeta(o, n, m, amp):
    y = np.arange(n)
    u = np.zeros((3, n, n, n))
    u[0] = amp * np.sin(2 * np.pi * m * y / n)[None, :, None]
    for d in range(3):
        o.set_field(f"U{d}", sp.pad_real(u[d]))
        o.fft_r2c(f"U{d}")
    o.u_k2zeta_k()
    return u


def test_nssolute_euler_shear_wave_decay():
    n, eta, dt, m = 32, 0.8, 0.01, 2
    p, pos, el = _setup(n=n, eta=eta)
    o = _oracle(p, el)
    o.set_particles(workloads.particle_state(pos))
    for s in (0, 1):
        c = np.full((n, n, n), 0.03)
        o.set_field(f"CONC{s}", _kspace_conc(c))
    u = _shear_zeta(o, n, m, 0.1)
    o.nssolute_euler(abi.make_time(dt))
    o.zeta_k2u_k()
    got = sp.to_complex(o.get_field("U0"))
    k2 = (2 * np.pi * m / n) ** 2
    expect = np.fft.rfftn(u[0]) * (1.0 - eta * k2 * dt)
    assert np.abs(got - expect).max() < 1e-12 * np.abs(expect).max()


def test_nssolute_euler_matches_numpy_nernst_planck():
    n = 32
    shape = (n, n, n)
    e_ext = (0.02, 0.0, -0.01)
    p, pos, el = _setup(n=n, npart=3, e_ext=e_ext, eta=0.9)
    el.onsager_coeff[:] = [2.0, 3.5]
    el.kbt = 0.8
    o = _oracle(p, el)
    o.set_particles(workloads.particle_state(pos))
    rng = np.random.default_rng(21)
    conc = [_random_conc(n, rng, 0.05, 0.02), _random_conc(n, rng, 0.04, 0.02)]
    ck = [np.fft.rfftn(c) for c in conc]
    # a mode above the 2/3 cut as well: the step truncates it
    for s in (0, 1):
        ck[s][12, 3, 13] += 0.5
        o.set_field(f"CONC{s}", sp.from_complex(ck[s]))
    u = sp.random_solenoidal(shape, rng, kmax=5, amp=0.01)
    for d in range(3):
        o.set_field(f"U{d}", sp.pad_real(u[d]))
        o.fft_r2c(f"U{d}")
    o.u_k2zeta_k()
    z0 = sp.to_complex(o.get_field("ZETA0")).copy()
    z1 = sp.to_complex(o.get_field("ZETA1")).copy()
    dt = 0.01
    o.nssolute_euler(abi.make_time(dt))

    keep = sp.kept_mask(shape)
    _, (kx, ky, kz) = sp.wavevectors(shape)
    ks = (kx, ky, kz)
    k2 = kx * kx + ky * ky + kz * kz
    # charge from the untruncated concentrations, truncated potential, field
    conc_full = [np.fft.irfftn(c, s=shape, axes=(0, 1, 2)) for c in ck]
    rho = _numpy_charge(o, n, el, conc_full)
    grad = _grad_psi(rho, shape, el.dielectric_cst, trunc=True)
    grad = [grad[d] - e_ext[d] for d in range(3)]
    _, _, nrm = _brute(pos, n, p.radius, p.xi)
    band = (nrm * nrm).sum(0) > 0
    nn = np.where(band, (nrm * nrm).sum(0), 1.0)
    for s in (0, 1):
        ct = ck[s] * keep
        c = np.fft.irfftn(ct, s=shape, axes=(0, 1, 2))
        kl = el.kbt * el.onsager_coeff[s]
        ze = el.valency[s] * el.elementary_charge
        flux = np.stack([np.fft.irfftn(1j * k * kl * ct, s=shape, axes=(0, 1, 2)) + c * el.onsager_coeff[s] * ze * grad[d]
                         for d, k in enumerate(ks)])
        dot = (flux * nrm).sum(0) / nn
        flux = flux - np.where(band, nrm * dot, 0.0)
        flux = flux - c * u
        fk = np.fft.rfftn(flux, axes=(1, 2, 3))
        rhs = 1j * (kx * fk[0] + ky * fk[1] + kz * fk[2])
        expect = np.where(keep, ct + dt * rhs, 0.0)
        got = sp.to_complex(o.get_field(f"CONC{s}"))
        assert np.abs(got - expect).max() < 1e-12 * np.abs(expect).max(), s
        # the k = 0 amount is untouched (a divergence has no mean)
        assert abs(got[0, 0, 0] - ck[s][0, 0, 0]) < 1e-14 * abs(ck[s][0, 0, 0])
    om = _numpy_curl_advection(u, shape)
    f0, f1 = _reduce_rule(om, shape)
    nu = p.eta / p.rho
    for z, f, name in ((z0, f0, "ZETA0"), (z1, f1, "ZETA1")):
        expect = np.where(keep, z + dt * (f - nu * k2 * z), 0.0)
        got = sp.to_complex(o.get_field(name))
        assert np.abs(got - expect).max() < 1e-11 * np.abs(z).max(), name


def test_rescale_solute_restores_totals():
    n = 32
    p, pos, el = _setup(n=n)
    o = _oracle(p, el)
    o.set_particles(workloads.particle_state(pos))
    o.init_rho_ion()
    tot = o.get_total_solute()
    rng = np.random.default_rng(5)
    conc = [_random_conc(n, rng, 0.07, 0.02), _random_conc(n, rng, 0.02, 0.01)]
    for s in (0, 1):
        o.set_field(f"CONC{s}", _kspace_conc(conc[s]))
    fac = o.rescale_solute()
    phi, _, _ = _brute(pos, n, p.radius, p.xi)
    assert np.abs(_real(o, "PHI", n) - phi).max() < 1e-13  # PHI is left unclamped
    for s in (0, 1):
        amount = ((1.0 - phi) * conc[s]).sum()
        assert abs(fac[s] - tot[s] / amount) < 1e-12 * fac[s]
        c = np.fft.irfftn(sp.to_complex(o.get_field(f"CONC{s}")), s=(n, n, n), axes=(0, 1, 2))
        assert abs(((1.0 - phi) * c).sum() - tot[s]) < 1e-12 * tot[s]


def test_electrolyte_step_neutral_and_conserving():
    n = 32
    p, pos, el = _setup(n=n, npart=3, e_ext=(0.05, 0.0, 0.0))
    o = _oracle(p, el)
    o.set_particles(workloads.particle_state(pos))
    o.init_rigidity()
    o.init_rho_ion()
    tot = o.get_total_solute()
    dt = workloads.electrolyte_auto_dt(p, el)
    t = abi.make_time(dt)
    for _ in range(3):
        o.electrolyte_step(t)
    assert t.ts == 3
    x = o.get_particles()[abi.PFIELD["X"]].T
    phi, _, _ = _brute(x, n, p.radius, p.xi)
    for s in (0, 1):
        c = np.fft.irfftn(sp.to_complex(o.get_field(f"CONC{s}")), s=(n, n, n), axes=(0, 1, 2))
        assert abs(((1.0 - phi) * c).sum() - tot[s]) < 1e-11 * tot[s]
    o.charge_field("WORK0")
    assert abs(_real(o, "WORK0", n).sum()) < 1e-10 * len(pos) * abs(el.surface_charge)
    # the field drags the positively charged spheres along +x
    v = o.get_particles()[abi.PFIELD["V"]]
    assert np.all(v[0] > 0)


def test_free_energy_pb_numpy():
    """Calc_free_energy_PB: ideal part kBT dx^3 sum (1-phi) c (ln c - 1) over
    both species, electrostatic part dx^3 sum rho (psi/2 - E.x) with the
    numpy Poisson solve."""
    n = 32
    e_ext = (0.02, -0.01, 0.03)
    p, pos, el = _setup(n=n, e_ext=e_ext)
    el.kbt = 0.7
    el.dielectric_cst = 1.3
    o = _oracle(p, el)
    o.set_particles(workloads.particle_state(pos))
    rng = np.random.default_rng(8)
    conc = [_random_conc(n, rng, 0.05, 0.02), _random_conc(n, rng, 0.04, 0.02)]
    for s in (0, 1):
        o.set_field(f"CONC{s}", _kspace_conc(conc[s]))
    fe = o.free_energy_pb(abi.make_time(0.01))
    phi, _, _ = _brute(pos, n, p.radius, p.xi)
    ideal = el.kbt * sum(((1 - phi) * c * (np.log(c) - 1)).sum() for c in conc)
    rho = _numpy_charge(o, n, el, conc)
    rk = np.fft.rfftn(rho)
    _, (kx, ky, kz) = sp.wavevectors((n, n, n))
    k2 = kx * kx + ky * ky + kz * kz
    psi = np.fft.irfftn(np.where(k2 > 0, rk / (el.dielectric_cst * np.where(k2 > 0, k2, 1.0)), 0.0),
                        s=(n, n, n), axes=(0, 1, 2))
    g = np.arange(n, dtype=float)
    X = np.stack(np.meshgrid(g, g, g, indexing="ij"))
    ext = sum(e_ext[d] * X[d] for d in range(3))
    electro = (rho * (0.5 * psi - ext)).sum()
    assert abs(fe[1] - ideal) < 1e-12 * abs(ideal)
    assert abs(fe[2] - electro) < 1e-11 * np.abs(rho * (0.5 * psi - ext)).sum()
    assert fe[0] == fe[1] + fe[2]


def test_steady_state_ions_relax_toward_boltzmann():
    """Set_steadystate_ion_density from the uniform salt start (32^3, three
    spheres, no field): the free energy falls, the amounts stay, and away
    from the spheres the electrochemical potential ln c_n + z_n e psi / kBT
    is flat to a tenth of the variation of z_n e psi / kBT (the relaxation
    stops at a free-energy change of 1e-4, and the 2/3 cut limits psi)."""
    n = 32
    p, pos, el = _setup(n=n, q=30.0)
    o = _oracle(p, el)
    o.set_particles(workloads.particle_state(pos))
    t = abi.make_time(workloads.electrolyte_auto_dt(p, el))
    o.init_rho_ion()
    fe0, tot0 = o.free_energy_pb(t), o.get_total_solute()
    it, fe = o.init_rho_ion_pb(t)
    assert 1 < it < 10000 and fe[0] < fe0[0]
    assert np.abs(o.get_total_solute() - tot0).max() < 1e-12 * tot0.max()
    c = [np.fft.irfftn(sp.to_complex(o.get_field(f"CONC{s}")), s=(n, n, n), axes=(0, 1, 2)) for s in (0, 1)]
    o.charge_field("WORK0")
    rho = _real(o, "WORK0", n)
    rk = np.fft.rfftn(rho)
    _, (kx, ky, kz) = sp.wavevectors((n, n, n))
    k2 = kx * kx + ky * ky + kz * kz
    psi = np.fft.irfftn(np.where(k2 > 0, rk / np.where(k2 > 0, k2, 1.0), 0.0), s=(n, n, n), axes=(0, 1, 2))
    g = np.arange(n, dtype=float)
    G = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1)
    dmin = np.full((n, n, n), 1e9)
    for xp in pos:
        d = G - xp
        d -= n * np.rint(d / n)
        dmin = np.minimum(dmin, np.sqrt((d * d).sum(-1)))
    far = dmin > p.radius + p.xi
    for s, z in ((0, 1.0), (1, -1.0)):
        mu = np.log(c[s]) + z * psi
        assert mu[far].std() < 0.15 * (z * psi)[far].std(), s


def test_pb_one_particle_fixed_point():
    """Set_Poisson_Boltzmann_ion_charge_density_salt (one sphere, 1:1 salt):
    c+ = rho+ exp(-e psi/kBT) nu and c- = rho- exp(e psi/kBT)/nu, so c+ c- is
    the bulk product everywhere; the box is neutral; and at the fixed point
    psi is the Coulomb potential of the charge it makes (to the stopping
    tolerance of the damped iteration)."""
    n, q, debye = 32, 40.0, 4.0
    p = abi.make_params(n, n_particles=1, radius=4.0, xi=2.0)
    el = abi.make_electrolyte(surface_charge=q, debye_length=debye)
    o = _oracle(p, el)
    o.set_particles(workloads.particle_state(np.array([[15.3, 16.1, 14.7]])))
    it, fe = o.init_rho_ion_pb(abi.make_time(0.01), 0)
    assert 1 < it < 100000
    c = [np.fft.irfftn(sp.to_complex(o.get_field(f"CONC{s}")), s=(n, n, n), axes=(0, 1, 2)) for s in (0, 1)]
    rp = rn = 1.0 / debye ** 2 / 2.0
    assert np.abs(c[0] * c[1] / (rp * rn) - 1.0).max() < 1e-12
    tot = o.get_total_solute()
    assert abs(tot[1] - tot[0] - q) < 1e-9 * q  # counter-ions balance the surface charge
    o.charge_field("WORK0")
    rho = _real(o, "WORK0", n)
    assert abs(rho.sum()) < 1e-10 * q
    rk = np.fft.rfftn(rho)
    _, (kx, ky, kz) = sp.wavevectors((n, n, n))
    k2 = kx * kx + ky * ky + kz * kz
    psi = np.fft.irfftn(np.where(k2 > 0, rk / np.where(k2 > 0, k2, 1.0), 0.0), s=(n, n, n), axes=(0, 1, 2))
    psi_e = -np.log(c[0] / c[1]) / 2.0  # c+/c- = (rp/rn) nu^2 exp(-2 psi_e): psi_e up to a constant
    d = (psi - psi.mean()) - (psi_e - psi_e.mean())
    assert np.abs(d).max() < 0.05 * np.abs(psi - psi.mean()).max()


def test_dielectric_grad_closed_form():
    """Dielectric branch of Rhs_NS_Nernst_Planck: the potential's gradient at
    the staggered faces, shifted by the applied field, averaged back onto the
    cells.  For psi = cos(k x): -k sin(k x) cos(k dx/2) - E_x."""
    n, m, ex = 32, 3, 0.07
    p, pos, el = _setup(n=n, e_ext=(ex, 0.0, 0.0), dielectric=1)
    o = _oracle(p, el)
    o.set_particles(workloads.particle_state(pos))
    k = 2 * np.pi * m / n
    x = np.arange(n, dtype=float)
    psi = np.broadcast_to(np.cos(k * x)[:, None, None], (n, n, n))
    o.set_field("POTENTIAL", sp.pad_real(psi))
    o.dielectric_grad(abi.make_time(0.01))
    gx = _real(o, "F0", n)
    expect = (-k * np.sin(k * x) * np.cos(k / 2) - ex)[:, None, None]
    assert np.abs(gx - expect).max() < 1e-13
    assert np.abs(_real(o, "F1", n)).max() < 1e-14 and np.abs(_real(o, "F2", n)).max() < 1e-14


def test_dielectric_ions_init_potential_and_rest_state():
    """Init_rho_ion_dielectric without surface charge: bulk density with
    kappa^2 = e^2 sum z^2 c / (kBT Dielectric_cst eps_f) = 1 / (lambda_D a)^2,
    a neutral box, so no potential and no flux: with no field the state stays
    at rest.  With a field, the potential of the ions in the permittivity is
    the solve's answer, and the ion amounts (k = 0 modes) are exact invariants
    of the step."""
    n, debye, epsf = 32, 0.5, 2.0
    p, pos, el = _setup(n=n, q=0.0, dielectric=1, debye=debye, eps_fluid=epsf, eps_particle=0.2)
    o = _oracle(p, el)
    o.set_particles(workloads.particle_state(pos))
    o.init_rho_ion_dielectric()
    c = [sp.to_complex(o.get_field(f"CONC{s}"))[0, 0, 0].real / n ** 3 for s in (0, 1)]
    kappa2 = (c[0] + c[1]) / (el.kbt * el.dielectric_cst * epsf)
    assert abs(kappa2 * (debye * p.radius) ** 2 - 1.0) < 1e-13
    t = abi.make_time(0.01)
    it, rr = o.dielectric_potential(t, 1e-10, 2000)
    assert np.abs(o.get_field("POTENTIAL")).max() < 1e-12  # neutral everywhere: zero charge
    o.init_rigidity()
    o.icep_electrolyte_step(t, 1e-10, 2000)
    assert np.abs(o.get_particles()[abi.PFIELD["V"]]).max() < 1e-14
    # with a field: ions move, amounts stay
    p2, pos2, el2 = _setup(n=n, q=0.0, dielectric=1, debye=debye, eps_fluid=epsf, eps_particle=0.2,
                           e_ext=(0.05, 0.0, 0.0))
    o2 = _oracle(p2, el2)
    o2.set_particles(workloads.particle_state(pos2))
    o2.init_rho_ion_dielectric()
    o2.init_rigidity()
    it, rr = o2.dielectric_potential(t, 1e-10, 2000)
    assert rr < 1e-10
    c0 = [sp.to_complex(o2.get_field(f"CONC{s}"))[0, 0, 0] for s in (0, 1)]
    for _ in range(3):
        it, rr = o2.icep_electrolyte_step(t, 1e-10, 2000)
        assert rr < 1e-10
    for s in (0, 1):
        assert abs(sp.to_complex(o2.get_field(f"CONC{s}"))[0, 0, 0] - c0[s]) < 1e-13 * abs(c0[s])
    assert np.abs(o2.get_field("ZETA0")).max() > 0  # the induced charge drives a flow


def test_dielectric_free_charge_and_neutralised_solve():
    """Conc_k2charge_field_no_surfase_charge: rho = sum_n z_n e c_n (1 - phi_sin)
    with phi_sin the brute-force compact-sine profile sum; ions spread into
    the particles make its total nonzero, which the periodic problem cannot
    carry, and the solve meets A psi = b + mean(rho) (the uniform background
    that cancels it) to its tolerance."""
    n, a, xi = 32, 4.0, 2.0
    p, pos, el = _setup(n=n, q=0.0, dielectric=1, debye=0.5, eps_fluid=2.0, eps_particle=0.2,
                        e_ext=(0.03, 0.01, 0.0))
    o = _oracle(p, el)
    o.set_particles(workloads.particle_state(pos))
    rng = np.random.default_rng(7)
    c = []
    for base in (0.05, 0.045):
        f = rng.standard_normal((n, n, n))
        f = np.fft.irfftn(np.fft.rfftn(f) * sp.band_mask((n, n, n), 4), s=(n, n, n), axes=(0, 1, 2))
        c.append(base + 0.02 * f / np.abs(f).max())
    for s in (0, 1):
        o.set_field(f"CONC{s}", sp.from_complex(np.fft.rfftn(c[s])))
    t = abi.make_time(0.01)
    it, rr = o.dielectric_potential(t, 1e-12, 4000, 1)
    assert rr < 1e-12
    phi_sin = np.zeros((n, n, n))
    for xp in pos:
        r = grid_vectors(n, xp)
        phi_sin += _phi_sin(np.sqrt((r * r).sum(-1)), a, xi)
    rho = (el.valency[0] * c[0] + el.valency[1] * c[1]) * el.elementary_charge * (1.0 - phi_sin)
    got = o.get_field("RHO_FREE")[:, :, :n]
    assert np.abs(got - rho).max() < 1e-13 * np.abs(rho).max()
    mean = rho.mean()
    psi = o.get_field("POTENTIAL")[:, :, :n]
    o.poisson_rhs([0.03, 0.01, 0.0])
    b = o.get_field("RHS")[:, :, :n]
    ax = o.ax_unpadded(np.ascontiguousarray(psi))
    assert abs(mean) > 1e-4 * np.abs(rho).max()
    assert np.linalg.norm(ax - (b + mean)) < 1e-11 * np.linalg.norm(b + mean)
    assert abs(ax.mean()) < 1e-13 * np.abs(ax).max()  # the operator's range has no mean


def test_dielectric_bulk_density_with_surface_charge():
    """Init_rho_ion under Dielectric: the bulk salt density carries eps_fluid,
    kappa^2 = e^2 sum z^2 c_bulk / (kBT Dielectric_cst eps_fluid) = 1 / lambda_D^2,
    and the counter-ions of the surface charge are added uniformly."""
    n, debye, epsf, q = 32, 4.0, 3.0, 40.0
    p, pos, el = _setup(n=n, q=q, dielectric=1, debye=debye, eps_fluid=epsf, eps_particle=0.2)
    o = _oracle(p, el)
    o.set_particles(workloads.particle_state(pos))
    o.init_rho_ion_dielectric()
    c = [sp.to_complex(o.get_field(f"CONC{s}"))[0, 0, 0].real / n ** 3 for s in (0, 1)]
    # q > 0: the negative species holds the counter-ions, the positive one is the bulk alone
    assert abs(c[0] * 2.0 * debye ** 2 / (el.kbt * el.dielectric_cst * epsf) - 1.0) < 1e-13
    o.make_phi_particle("WORK0")
    vol = (1.0 - o.get_field("WORK0")[:, :, :n]).sum()
    counter = q * p.n_particles / abs(el.valency[1]) / vol
    assert abs((c[1] - c[0]) / counter - 1.0) < 1e-12
