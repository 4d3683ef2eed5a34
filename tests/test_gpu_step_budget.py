"""GPU property check at the 128^3 x 200-sphere size: the momentum budget of
the SPM step (Time_evolution_hydro, sp_3d_ns.cxx:24-186) holds for the HIP
library's own steps, one Euler and two AB2.

The budget (derived in tests/test_oracle_step.py) is
    rho dx^3 d(uk_dc) + sum_i (M_i - rho dx^3 V_i) dv_i
        - rho dx^3 sum_i d(omega_i) x c_i = dt N_p F_body,
with V_i = sum_x w_i, c_i = sum_x r w_i, w_i = phi_i / max(phi_sum, 1) at the
step's positions.  V_i and c_i are computed here in numpy from the closed-form
profile (profile.h:29-49) over each sphere's neighbourhood, independently of
the library.  Bound: 1e-11 of the budget's scale in FP64; in FP32 the fluid
momentum is summed from FP32 fields, bound 2e-5.
"""
import numpy as np
import pytest

from icep_kapsel_amd import abi, workloads
from tests.test_oracle_particles import phi_profile

pytestmark = pytest.mark.gpu


def _weights_local(pos, n, a, xi):
    """V_i, c_i with w_i = phi_i / max(phi_sum, 1) over the box of integer
    points around each sphere that holds its support (r < a + xi/2)."""
    reach = a + xi / 2.0
    h = int(np.ceil(reach)) + 1
    o = np.arange(-h, h + 1)
    O = np.stack(np.meshgrid(o, o, o, indexing="ij"), -1).reshape(-1, 3)
    L = float(n)
    V, C = np.zeros(len(pos)), np.zeros((len(pos), 3))
    for i, x in enumerate(pos):
        pts = np.floor(x) + O                   # integer grid points (unwrapped)
        r = pts - x
        phi_i = phi_profile(np.sqrt((r ** 2).sum(-1)), a, xi)
        # neighbours whose support can reach these points
        d = pos - x
        d -= L * np.rint(d / L)
        nb = np.where(np.sqrt((d ** 2).sum(-1)) < 2 * reach + 2)[0]
        phisum = np.zeros(len(pts))
        for j in nb:
            rj = pts - pos[j]
            rj -= L * np.rint(rj / L)
            phisum += phi_profile(np.sqrt((rj ** 2).sum(-1)), a, xi)
        w = phi_i / np.maximum(phisum, 1.0)
        V[i] = w.sum()
        C[i] = (r * w[:, None]).sum(0)
    return V, C


@pytest.mark.parametrize("prec,bound", [(abi.KAPSEL_FP64, 1e-11), (abi.KAPSEL_FP32, 2e-5)])
def test_spm_step_momentum_budget_128_200_spheres(prec, bound):
    from icep_kapsel_amd.gpu import KapselGPU

    print(f"case spm momentum budget 128^3 x 200 prec={prec}", flush=True)
    p, st, dt = workloads.spm_config("c2", precision=prec)
    p.mass_ratio = 2.0
    p.body_force[:] = [2e-4, -1e-4, 1e-3]
    st[abi.PFIELD["OMEGA"]] = np.random.default_rng(7).normal(0, 1e-4, size=(3, p.n_particles))
    k = KapselGPU(p)
    k.set_particles(st)
    k.init_rigidity()
    t = abi.make_time(dt)
    dx3 = p.dx ** 3
    mass = p.mass_ratio * p.rho * 4.0 / 3.0 * np.pi * p.radius ** 3
    worst = 0.0
    for step in range(3):
        uk0, s0 = k.get_uk_dc(), k.get_particles()
        k.spm_step(t)
        uk1, s1 = k.get_uk_dc(), k.get_particles()
        V, C = _weights_local(s1[abi.PFIELD["X"]].T, p.nx, p.radius, p.xi)
        dv = (s1[abi.PFIELD["V"]] - s0[abi.PFIELD["V"]]).T
        dom = (s1[abi.PFIELD["OMEGA"]] - s0[abi.PFIELD["OMEGA"]]).T
        fluid = p.rho * dx3 * (uk1 - uk0)
        lhs = fluid + ((mass - p.rho * dx3 * V)[:, None] * dv).sum(0) - p.rho * dx3 * np.cross(dom, C).sum(0)
        rhs = dt * p.n_particles * np.array(p.body_force[:])
        scale = np.abs(fluid).max() + np.abs(mass * dv).sum() + np.abs(rhs).max()
        err = np.abs(lhs - rhs).max() / scale
        print(f"step {step}: budget error {err:.3e} of scale {scale:.3e}", flush=True)
        worst = max(worst, err)
    k.close()
    assert worst < bound
