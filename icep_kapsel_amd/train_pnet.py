"""Trains the Poisson warm-start network on solves of the plain solver, on the GPU.

python -m icep_kapsel_amd.train_pnet --out WEIGHTS [--n 128] [--samples 2000]

Each sample is a fresh dielectric configuration of the ICEP box: 64 spheres
(a = 6, xi = 2) at seeded random non-overlapping positions, permittivity
ratio eps_p / eps_f log-uniform in [0.05, 20], a field E0 of magnitude 0.1
in a random direction.  The target potential comes from the library's own
solver (FP32 spectral-preconditioned CG refined to an FP64 relative residual
of 1e-6); the network then takes Adam steps on it through the library's
BF16 MFMA forward/backward kernels.  Inputs: the compact-sine sum phi and
beta phi with beta = (r - 1) / (r + 2); target: (potential - mean) / (|E0| a).
Writes the FP32 parameters (oracle layout) and a JSON record beside them.
"""
import argparse
import json
import math
import sys
import time

import numpy as np

from . import abi, workloads
from .gpu import KapselGPU

A, XI, NPART, E_MAG = 6.0, 2.0, 64, 0.1


def he_init(seed, layers, nparam):
    rng = np.random.default_rng(seed)
    p = np.zeros(nparam, dtype=np.float32)
    for cin, cout, woff, boff in layers:
        p[woff:boff] = rng.standard_normal(boff - woff) * math.sqrt(2.0 / (27 * cin))
    # the linear output layer starts small
    cin, cout, woff, boff = layers[-1]
    p[woff:boff] *= 0.1
    return p


LAYERS = [(2, 16), (16, 16), (16, 32), (32, 32), (32, 32), (32, 32), (64, 32), (48, 16), (16, 3)]


def layer_offsets():
    out, off = [], 0
    for cin, cout in LAYERS:
        w = 27 * cin * cout
        out.append((cin, cout, off, off + w))
        off += w + cout
    return out, off


def sample(seed, n):
    rng = workloads.Xoshiro256(seed)
    pos = workloads.random_positions(rng, NPART, (n, n, n), 2 * A)
    ratio = math.exp(math.log(0.05) + rng.uniform() * (math.log(20.0) - math.log(0.05)))
    v = np.array([rng.normal() for _ in range(3)])
    e_hat = v / np.linalg.norm(v)
    return pos, ratio, e_hat


def beta_of(ratio):
    return (ratio - 1.0) / (ratio + 2.0)


def make_context(n):
    p = abi.make_params(n, n_particles=NPART, radius=A, xi=XI, eps_fluid=1.0, eps_particle=0.1,
                        precision=abi.KAPSEL_FP32)
    return KapselGPU(p)


def prepare(k, pos, ratio):
    k.set_particles(workloads.particle_state(pos))
    k.set_dielectric(1.0, ratio)
    k.make_dielectric_field()


def load_weights(path):
    return np.fromfile(path, dtype=np.float32)


def evaluate(k, params, n_samples, seed=4104, tol=1e-8, n=128):
    """Cold solves of held-out configurations, warm-started from zero, from
    the network, and from the network rescaled by the energy-norm (Galerkin)
    factor: FP32 spectral PCG refined to an FP64 residual tol.  Returns a
    list of dicts (iterations, seconds, final relative residual)."""
    import time as _t
    k.pnet_init()
    k.pnet_set_params(params)
    out = []
    for s in range(n_samples):
        pos, ratio, e_hat = sample(seed * 100000 + s, n)
        prepare(k, pos, ratio)
        e0 = E_MAG * e_hat
        rec = {"ratio": ratio}
        for mode in ("zero", "network", "network_galerkin"):
            k.set_field("POTENTIAL", np.zeros(k.shape))
            k.synchronize()
            t0 = _t.perf_counter()
            if mode != "zero":
                k.pnet_warm_start(beta_of(ratio), e0, A)
            if mode == "network_galerkin":
                rec["gamma"] = k.poisson_galerkin_start(e0)
            it, rr = k.poisson_solve(e0, tol, 8000, abi.PRECOND_SPECTRAL)
            k.synchronize()
            rec[mode + "_iters"] = it
            rec[mode + "_seconds"] = _t.perf_counter() - t0
            rec[mode + "_relres"] = rr
        out.append(rec)
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--samples", type=int, default=2000)
    ap.add_argument("--steps-per-sample", type=int, default=3)
    ap.add_argument("--lr", type=float, default=5e-4)
    ap.add_argument("--seed", type=int, default=4004)
    ap.add_argument("--out", help="weights file to write")
    ap.add_argument("--eval", type=int, default=0, help="after training, evaluate on this many held-out samples")
    ap.add_argument("--eval-only", metavar="WEIGHTS", help="evaluate these weights (no training)")
    args = ap.parse_args(argv)
    offs, nparam = layer_offsets()
    k = make_context(args.n)
    if args.eval_only:
        ev = evaluate(k, load_weights(args.eval_only), args.eval or 8, n=args.n)
        for e in ev:
            print("eval:", json.dumps(e), flush=True)
        for mode in ("zero", "network", "network_galerkin"):
            print(f"eval: mean {mode} iterations {np.mean([e[mode + '_iters'] for e in ev]):.2f} "
                  f"seconds {np.mean([e[mode + '_seconds'] for e in ev]):.4f}", flush=True)
        return 0
    if not args.out:
        ap.error("--out is required for training")
    assert k.pnet_init() == nparam
    k.pnet_set_params(he_init(args.seed, offs, nparam))
    t0 = time.time()
    losses, iters = [], []
    for s in range(args.samples):
        pos, ratio, e_hat = sample(args.seed * 100000 + s, args.n)
        prepare(k, pos, ratio)
        k.set_field("POTENTIAL", np.zeros(k.shape))
        it, rr = k.poisson_solve(E_MAG * e_hat, 1e-6, 4000, abi.PRECOND_SPECTRAL)
        iters.append(it)
        lr = args.lr * (0.1 + 0.9 * 0.5 * (1 + math.cos(math.pi * s / args.samples)))
        for _ in range(args.steps_per_sample):
            loss = k.pnet_train_step(beta_of(ratio), e_hat, E_MAG * A, lr)
        losses.append(loss)
        if s % 50 == 0 or s == args.samples - 1:
            x = k.get_field("POTENTIAL")[:, :, :args.n]
            t2 = float(np.mean(((x - x.mean()) / (E_MAG * A)) ** 2))
            print(f"train: sample {s} ratio {ratio:.3f} solve {it} it loss {loss:.4e} "
                  f"(relative {loss / t2:.3f}; mean last 50: {np.mean(losses[-50:]):.4e}) "
                  f"{time.time() - t0:.0f}s", flush=True)
    params = k.pnet_get_params(nparam)
    params.astype(np.float32).tofile(args.out)
    rec = {"samples": args.samples, "steps_per_sample": args.steps_per_sample, "lr": args.lr,
           "seed": args.seed, "grid": args.n, "particles": NPART, "radius": A, "xi": XI, "e_mag": E_MAG,
           "final_loss_mean_last_100": float(np.mean(losses[-100:])),
           "first_loss_mean_100": float(np.mean(losses[:100])),
           "mean_solver_iterations": float(np.mean(iters)), "train_seconds": time.time() - t0,
           "n_params": nparam}
    if args.eval:
        ev = evaluate(k, params, args.eval, n=args.n)
        rec["eval"] = ev
        rec["eval_mean_iters_zero"] = float(np.mean([e["zero_iters"] for e in ev]))
        rec["eval_mean_iters_network"] = float(np.mean([e["network_iters"] for e in ev]))
        rec["eval_mean_iters_network_galerkin"] = float(np.mean([e["network_galerkin_iters"] for e in ev]))
    with open(args.out + ".json", "w") as f:
        json.dump(rec, f, indent=1)
    print("train: done", json.dumps(rec), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
