"""Seeded synthetic inputs for the SPM step (no datasets exist for this path).

Generator: splitmix64 seeds xoshiro256** (Blackman & Vigna); uniform doubles
from the top 53 bits, normals by Box-Muller.  Particle positions are uniform
in [0, L) with the reference's non-overlap rule |x_i - x_j| > 1.05 sigma
under periodic images (init_particle.cxx:56-73).
"""
import math

import numpy as np

from . import abi

_M64 = (1 << 64) - 1


class Xoshiro256:
    def __init__(self, seed):
        s = seed & _M64
        self.s = []
        for _ in range(4):
            s = (s + 0x9E3779B97F4A7C15) & _M64
            z = s
            z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
            z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
            self.s.append(z ^ (z >> 31))

    @staticmethod
    def _rotl(x, k):
        return ((x << k) | (x >> (64 - k))) & _M64

    def next(self):
        s = self.s
        result = (self._rotl((s[1] * 5) & _M64, 7) * 9) & _M64
        t = (s[1] << 17) & _M64
        s[2] ^= s[0]
        s[3] ^= s[1]
        s[1] ^= s[2]
        s[0] ^= s[3]
        s[2] ^= t
        s[3] = self._rotl(s[3], 45)
        return result

    def uniform(self):
        return (self.next() >> 11) * (1.0 / (1 << 53))

    def normal(self):
        u1 = self.uniform()
        u2 = self.uniform()
        u1 = max(u1, 1e-300)
        return math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)


def random_positions(rng, n, box, sigma, max_tries=200000):
    """n positions in [0, box) (3-vector), pairwise periodic distance > 1.05 sigma."""
    box = np.asarray(box, dtype=np.float64)
    pos = np.zeros((n, 3))
    placed = 0
    tries = 0
    while placed < n:
        tries += 1
        if tries > max_tries:
            raise RuntimeError(f"could not place {n} non-overlapping particles")
        x = np.array([rng.uniform() * box[d] for d in range(3)])
        if placed:
            d = pos[:placed] - x
            d -= box * np.rint(d / box)
            if np.min(np.sum(d * d, axis=1)) <= (1.05 * sigma) ** 2:
                continue
        pos[placed] = x
        placed += 1
    return pos


def particle_state(positions, velocities=None, omegas=None):
    """Particle state array (P_NFIELDS, 3, Np) with x, x_nopbc, v, omega set."""
    n = positions.shape[0]
    st = np.zeros((abi.P_NFIELDS, 3, n))
    st[abi.PFIELD["X"]] = positions.T
    st[abi.PFIELD["X_PREVIOUS"]] = positions.T
    st[abi.PFIELD["X_NOPBC"]] = positions.T
    if velocities is not None:
        st[abi.PFIELD["V"]] = velocities.T
    if omegas is not None:
        st[abi.PFIELD["OMEGA"]] = omegas.T
    return st


def spm_config(name, precision=abi.KAPSEL_FP64, seed=None, n_particles=None, grid=None):
    """Parameters, initial particle state and dt of the SPM workloads.

    c1: 64^3, 8 spheres a=5 xi=2, eta=rho=1, dt=0.05, WCA LJ sigma=2a eps=1,
        body force 1e-3 along +z, rotation on.
    c2: 128^3, 200 spheres a=4 xi=2, Gaussian particle velocities std 1e-3.
    small: 32^3, 4 spheres a=3 xi=2 (parity-test size).
    """
    if name == "c1":
        n, npart, a, xi, seed0 = 64, 8, 5.0, 2.0, 1001
        body = (0.0, 0.0, 1e-3)
        vstd = 0.0
    elif name == "c2":
        n, npart, a, xi, seed0 = 128, 200, 4.0, 2.0, 2002
        body = (0.0, 0.0, 0.0)
        vstd = 1e-3
    elif name == "small":
        n, npart, a, xi, seed0 = 32, 4, 3.0, 2.0, 7
        body = (0.0, 0.0, 1e-3)
        vstd = 1e-3
    else:
        raise ValueError(name)
    if grid is not None:
        n = grid
    if n_particles is not None:
        npart = n_particles
    seed = seed0 if seed is None else seed
    rng = Xoshiro256(seed)
    p = abi.make_params(n, n_particles=npart, radius=a, xi=xi, rotation=1, lj_truncate=1,
                        lj_powers=0, lj_epsilon=1.0, lj_sigma=2.0 * a, body_force=body,
                        precision=precision)
    pos = random_positions(rng, npart, (n, n, n), 2.0 * a)
    vel = np.array([[rng.normal() * vstd for _ in range(3)] for _ in range(npart)]) if vstd else None
    st = particle_state(pos, vel)
    return p, st, 0.05


def icep_config(run=0, precision=abi.KAPSEL_FP64, ratio=0.1, grid=128, n_particles=64):
    """The ICEP box: 128^3, 64 dielectric spheres a=6 xi=2 at rest in a fluid
    at rest, eps_f=1, eps_p=ratio (0.1), no free charge, eta=rho=1, dt=0.05,
    WCA LJ sigma=2a.  E0 = R (0, 0, 0.1) with R a seeded uniform random
    rotation, i.e. |E0| = 0.1 along a uniformly random direction (normal
    deviates, normalised).  Seed 3003 + run; positions are drawn first.
    Returns (params, particle state, e0, dt)."""
    n, a = grid, 6.0
    rng = Xoshiro256(3003 + run)
    p = abi.make_params(n, n_particles=n_particles, radius=a, xi=2.0, rotation=1, lj_truncate=1,
                        lj_powers=0, lj_epsilon=1.0, lj_sigma=2.0 * a, eps_fluid=1.0, eps_particle=ratio,
                        precision=precision)
    pos = random_positions(rng, n_particles, (n, n, n), 2.0 * a)
    v = np.array([rng.normal() for _ in range(3)])
    e0 = 0.1 * v / np.linalg.norm(v)
    return p, particle_state(pos), e0, 0.05


def electrolyte_auto_dt(p, el):
    """Time_increment "auto" with factor 1 for the Electrolyte equation
    (input.cxx:329-379): min(1 / (nu KMAX2), 1 / (kBT max(L_+, L_-) KMAX2)),
    KMAX2 = sum_d (WAVE_d TRN_d)^2; with AC, also at most 1e-2 / Frequency."""
    n = (p.nx, p.ny, p.nz)
    kmax2 = sum((2.0 * math.pi / (m * p.dx) * ((m + 2) // 3)) ** 2 for m in n)
    nu = p.eta / p.rho
    dt = min(1.0 / (nu * kmax2), 1.0 / (el.kbt * max(el.onsager_coeff[0], el.onsager_coeff[1]) * kmax2))
    if el.ac and el.angular_frequency > 0:
        dt = min(dt, 1e-2 * 2.0 * math.pi / el.angular_frequency)
    return dt


def electrolyte_config(precision=abi.KAPSEL_FP64, grid=256, n_particles=2000, radius=4.0, xi=2.0,
                       surface_charge=100.0, debye_length=4.0, e_ext=(0.01, 0.0, 0.0), seed=5005):
    """Scale + electrolyte workload: 256^3, 2000 spheres a=4 xi=2 carrying
    surface charge 100 e each, two monovalent ion species (Onsager
    coefficients 7.24, kBT = e = Dielectric_cst = 1), Debye length 4, DC field
    E_ext = (0.01, 0, 0) (the reference UDF's salt and Electric_field
    sections, UDF/input.udf:53-100), eta = rho = 1, WCA LJ sigma = 2a, spheres
    at rest from seeded non-overlapping positions, auto time step.
    Returns (params, particle state, electrolyte parameters, dt)."""
    rng = Xoshiro256(seed)
    n, a = grid, radius
    p = abi.make_params(n, n_particles=n_particles, radius=a, xi=xi, rotation=1, lj_truncate=1,
                        lj_powers=0, lj_epsilon=1.0, lj_sigma=2.0 * a, precision=precision)
    pos = random_positions(rng, n_particles, (n, n, n), 2.0 * a)
    el = abi.make_electrolyte(surface_charge=surface_charge, debye_length=debye_length, e_ext=e_ext)
    return p, particle_state(pos), el, electrolyte_auto_dt(p, el)
