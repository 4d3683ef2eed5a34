"""GPU BF16/MFMA network kernels against the FP64 CPU oracle.

The GPU stores weights, inputs and activations in BF16 (8-bit mantissa) and
accumulates in FP32; the oracle gets the same BF16-rounded weights and
inputs in FP64.  Bar (written here): loss within 3 %, gradient direction
cosine > 0.995 and relative L2 error < 8 % overall; per layer cosine > 0.98
and relative error < 20 % (the gradients of the deep layers pass through
more BF16 roundings); the forward output within 3 % (relative L2).
Training lowers the loss.
"""
import numpy as np
import pytest

from icep_kapsel_amd import abi, workloads
from tests import spectral as sp

pytestmark = pytest.mark.gpu


def bf16_round(x):
    a = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    a = (a + 0x7FFF + ((a >> 16) & 1)) >> 16 << 16
    return a.astype(np.uint32).view(np.float32).astype(np.float64)


def _init(rng):
    from oracle import oracle as om
    params = np.zeros(om.pnet_param_count())
    for cin, cout, level, relu, woff, boff in om.pnet_layers():
        params[woff:boff] = rng.standard_normal(boff - woff) * np.sqrt(2.0 / (27 * cin))
        params[boff:boff + cout] = rng.uniform(0.0, 0.05, cout)
    return params


def _setup(n=16, seed=0):
    from icep_kapsel_amd.gpu import KapselGPU
    p = abi.make_params(n, n_particles=2, radius=3.0, xi=2.0, eps_fluid=1.0, eps_particle=0.1)
    pos = np.array([[4.3, 5.1, 6.7], [11.2, 10.4, 3.3]]) * (n / 16)
    k = KapselGPU(p)
    k.set_particles(workloads.particle_state(pos))
    k.make_dielectric_field()
    rng = np.random.default_rng(seed)
    x = np.arange(n)
    t = 0.3 * np.sin(2 * np.pi * x / n)[:, None, None] * np.cos(2 * np.pi * 2 * x / n)[None, :, None] \
        + 0.1 * rng.standard_normal((n, n, n))
    k.set_field("POTENTIAL", sp.pad_real(t))
    return k, rng, t


@pytest.mark.parametrize("n", [16, 32])
def test_pnet_gradient_matches_oracle(n):
    """16^3: one weight-gradient chunk per layer; 32^3: the full-resolution
    layers reduce 8 chunk partials."""
    from oracle import oracle as om
    beta, scale = -0.4285714, 0.5
    print(f"case pnet gradient {n}^3", flush=True)
    k, rng, t = _setup(n)
    P = k.pnet_init()
    assert P == om.pnet_param_count()
    params = _init(rng).astype(np.float32)
    k.pnet_set_params(params)
    e = np.array([0.2, -0.4, 0.8])
    e /= np.linalg.norm(e)
    loss_g, grad_g = k.pnet_loss_grad(beta, e, scale, P)
    phi = k.get_field("WORK4")[:, :, :n].reshape(-1)
    inp = np.stack([bf16_round(phi), bf16_round(beta * phi)], -1)
    target = (t.reshape(-1) - t.mean()) / scale
    loss_o, grad_o = om.pnet_loss_grad(n, bf16_round(params), inp, e, target)
    print(f"loss gpu {loss_g:.6e} oracle {loss_o:.6e}", flush=True)
    assert abs(loss_g - loss_o) < 0.03 * loss_o
    g = grad_g.astype(np.float64)
    cos = g @ grad_o / (np.linalg.norm(g) * np.linalg.norm(grad_o))
    rel = np.linalg.norm(g - grad_o) / np.linalg.norm(grad_o)
    print(f"gradient cosine {cos:.5f} rel {rel:.3e}", flush=True)
    assert cos > 0.995 and rel < 0.08
    for cin, cout, level, relu, woff, boff in om.pnet_layers():
        sl = slice(woff, boff + cout)
        r = np.linalg.norm(g[sl] - grad_o[sl]) / max(1e-30, np.linalg.norm(grad_o[sl]))
        c = g[sl] @ grad_o[sl] / max(1e-30, np.linalg.norm(g[sl]) * np.linalg.norm(grad_o[sl]))
        print(f"layer {cin}->{cout}: cosine {c:.5f} rel {r:.3e}", flush=True)
        assert r < 0.2 and c > 0.98, (cin, cout, r, c)


@pytest.mark.parametrize("n", [16, 32])
def test_pnet_forward_matches_oracle_and_training_reduces_loss(n):
    """16^3: the decoder's convolutions read the concatenated inputs; 32^3:
    they read [up(coarse) | skip] straight from the two tensors (the warm
    start's fused form)."""
    from oracle import oracle as om
    beta, scale = -0.4285714, 0.5
    print(f"case pnet forward + training {n}^3", flush=True)
    k, rng, t = _setup(n, seed=1)
    P = k.pnet_init()
    params = _init(rng).astype(np.float32)
    k.pnet_set_params(params)
    e0 = np.array([0.0, 0.6, 0.8])
    k.pnet_warm_start(beta, e0, scale)
    got = k.get_field("POTENTIAL")[:, :, :n].reshape(-1)
    phi = k.get_field("WORK4")[:, :, :n].reshape(-1)
    inp = np.stack([bf16_round(phi), bf16_round(beta * phi)], -1)
    y = om.pnet_forward(n, bf16_round(params), inp)
    ref = scale * (y @ e0)
    assert np.linalg.norm(got - ref) < 0.03 * np.linalg.norm(ref)
    k.set_field("POTENTIAL", sp.pad_real(t))
    losses = [k.pnet_train_step(beta, e0, scale, 1e-3) for _ in range(30)]
    print("losses", [f"{v:.3e}" for v in losses[::5]], flush=True)
    assert losses[-1] < 0.7 * losses[0]
    p2 = k.pnet_get_params(P)
    assert np.abs(p2 - params).max() > 0


def test_trained_network_gain_on_held_out_solves():
    """What the trained network (icep_kapsel_amd/weights) gains on held-out
    configurations of the ICEP box (seed 4104): the energy-norm error of its
    start, (e, S e) with e = x0 - x* and S = -A, against that of a zero start
    (= (x*, S x*)), and the spectral-PCG iterations to 1e-8 from each start.
    The start must remove most of the energy-norm error (ratio < 0.35; about
    0.1-0.2 is measured) and must not cost more than 2 iterations."""
    import os
    from icep_kapsel_amd import train_pnet
    n = 128
    path = os.path.join(os.path.dirname(__file__), "..", "icep_kapsel_amd", "weights", "pnet_icep_128.f32")
    k = train_pnet.make_context(n)
    w = np.fromfile(path, dtype=np.float32)
    assert k.pnet_init() == w.size
    k.pnet_set_params(w)

    def energy(x):
        k.set_field("WORK3", x)
        k.poisson_apply("WORK3", "WORK2")
        return -float(np.sum(x[:, :, :n].astype(np.float64) * k.get_field("WORK2")[:, :, :n]))

    for s in range(3):
        print(f"case network gain sample {s}", flush=True)
        pos, ratio, e_hat = train_pnet.sample(4104 * 100000 + 50 + s, n)
        e0 = train_pnet.E_MAG * e_hat
        train_pnet.prepare(k, pos, ratio)
        k.set_field("POTENTIAL", np.zeros(k.shape))
        k.poisson_solve(e0, 1e-10, 20000, abi.PRECOND_SPECTRAL)
        xs = k.get_field("POTENTIAL")
        xs[:, :, :n] -= xs[:, :, :n].mean()
        k.pnet_warm_start(train_pnet.beta_of(ratio), e0, train_pnet.A)
        gamma = k.poisson_galerkin_start(e0)
        ratio_e = energy(k.get_field("POTENTIAL") - xs) / energy(xs)
        it_net, rr_net = k.poisson_solve(e0, 1e-8, 20000, abi.PRECOND_SPECTRAL)
        k.set_field("POTENTIAL", np.zeros(k.shape))
        it_zero, rr_zero = k.poisson_solve(e0, 1e-8, 20000, abi.PRECOND_SPECTRAL)
        print(f"eps ratio {ratio:.3f}: energy-norm error of the start {ratio_e:.3f} of a zero start "
              f"(gamma {gamma:.3f}); PCG iterations {it_net} from the network, {it_zero} from zero", flush=True)
        assert rr_net < 1e-8 and rr_zero < 1e-8
        assert ratio_e < 0.35
        assert it_net <= it_zero + 2
