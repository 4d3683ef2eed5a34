"""ICEP step throughput on one MI355X: FP32 + learned warm start vs FP64 plain.

python bench.py [--gpus 1] [--steps K] [--warmup W]

Workload: the ICEP box -- 128^3 grid, 64 dielectric spheres a=6 xi=2
(eps_f = 1, eps_p = 0.1, no free charge), E0 of magnitude 0.1 along a seeded
random direction, eta = rho = 1, dt = 0.05, WCA LJ.  One step is the
reference's Navier-Stokes step with the dielectric force (sp_3d_ns.cxx:24-186
with the Electrolyte+Dielectric insert :51-95 and no ions): slaved-Euler NS
update, profile scatter, permittivity field, variable-permittivity Poisson
solve warm-started from the previous potential, Maxwell force, projection,
hydrodynamic gather, LJ, particle update and rigidity constraint.  Both
paths stop the Poisson solve at the same bound, relative residual 1e-8
checked in FP64.

  value : FP32 context (FP32 fields and FFTs, FP64 particle reductions),
          spectral-preconditioned CG in FP32 inside FP64 iterative
          refinement, first (cold) solve warm-started by the trained network
  base  : FP64 context, plain CG (no preconditioner), cold solve from zero

Both runs start from the same state and do the same W + K steps; the line
reports their particle-velocity and vorticity differences.  Beside them:
cold solves of held-out configurations (zero vs network warm start), and the
FP64 CPU oracle (the reference's BiCGSTAB) timed on one step from the GPU's
state after the warm-up.  Inputs are resident in HBM before timing.  Prints
one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "SPM steps/s at 128^3: FP32+learned-precond vs FP64 plain, same error bound"
# the timed steps carry HIP events on one launch of the dominant class in this many
TIMING_EVERY = 8
HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s
TOL = 1e-8
MAXIT = 20000
WEIGHTS = os.path.join(ROOT, "icep_kapsel_amd", "weights", "pnet_icep_128.f32")
PMC_SUMMARY = os.path.join(ROOT, "profiles", "icep_step_128_fp32_pmc_traffic.md")
SINGLE_KERNEL_CLASSES = ["FFT_Z_R2C", "FFT_Z_C2R", "FFT_LINES_FWD", "FFT_LINES_INV", "FFT_PLANE_FACE"]
KERNEL_OF_CLASS = {  # the kernel symbol of each single-kernel class in an FP32 128^3 context
    "FFT_Z_R2C": "fft_plane_r2c<float, 128, 128, 512>",  # fused z + y plane pass
    "FFT_Z_C2R": "fft_plane_c2r<float, 128, 128, 512>",
    "FFT_LINES_FWD": "fft_c2c_lines_io<float, 128, -1, 32>",
    "FFT_LINES_INV": "fft_c2c_lines_io<float, 128, 1, 32>",
    "FFT_PLANE_FACE": "fft_half_face<float, 128, 128, 512>",  # inverse z+y, face eps, forward z+y (half planes)
}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def beta_of(ratio):
    return (ratio - 1.0) / (ratio + 2.0)


def load_network(k):
    import numpy as np
    if not os.path.exists(WEIGHTS):
        return False
    n = k.pnet_init()
    w = np.fromfile(WEIGHTS, dtype=np.float32)
    if w.size != n:
        raise RuntimeError(f"{WEIGHTS}: {w.size} parameters, the network has {n}")
    k.pnet_set_params(w)
    return True


def run_icep(precision, precond, steps, warmup, learned, timed_class=None, snapshot=False):
    import numpy as np
    from icep_kapsel_amd import abi, workloads
    from icep_kapsel_amd.gpu import KapselGPU

    p, st, e0, dt = workloads.icep_config(precision=precision)
    ratio = p.eps_particle / p.eps_fluid
    k = KapselGPU(p)
    k.set_particles(st)
    k.init_rigidity()
    t = abi.make_time(dt)
    used_net = load_network(k) if learned else False
    k.synchronize()
    iters = []
    t_cold = time.perf_counter()
    if used_net:
        # cold start: the network's potential for the initial configuration
        k.make_dielectric_field()
        k.pnet_warm_start(beta_of(ratio), e0, p.radius)
        k.poisson_galerkin_start(e0)
    for i in range(warmup):
        it, rr = k.icep_step(t, e0, TOL, MAXIT, precond)
        iters.append(it)
        if i == 0:
            k.synchronize()
            t_cold = time.perf_counter() - t_cold
    k.synchronize()
    snap = None
    if snapshot:
        snap = {"ZETA0": k.get_field("ZETA0"), "ZETA1": k.get_field("ZETA1"), "uk_dc": k.get_uk_dc(),
                "particles": k.get_particles(), "POTENTIAL": k.get_field("POTENTIAL"), "ts": t.ts,
                "time": t.time}
    # device time per kernel class over one step; the dominant kernel is the
    # single-kernel class (the FFT passes: one kernel symbol each) with the
    # largest share
    k.kernel_timing(abi.KERNEL_CLASSES)
    it, rr = k.icep_step(t, e0, TOL, MAXIT, precond)
    iters.append(it)
    shares = {c: k.kernel_timing_get(c) for c in abi.KERNEL_CLASSES}
    k.kernel_timing([])
    dominant = timed_class or max(SINGLE_KERNEL_CLASSES, key=lambda c: shares[c][0])
    # timed region: only the dominant class carries HIP events, on one launch
    # in TIMING_EVERY; the half-plane face launch takes them itself
    # (hipExtLaunchKernelGGL), so they hold the kernel's own start and end
    k.kernel_timing([dominant])
    k.kernel_timing_every(TIMING_EVERY)
    k.synchronize()
    timed_iters, relres = [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        it, rr = k.icep_step(t, e0, TOL, MAXIT, precond)
        timed_iters.append(it)
        relres.append(rr)
    k.synchronize()
    t1 = time.perf_counter()
    ms, launches, nbytes = k.kernel_timing_get(dominant)
    k.kernel_timing([])
    k.kernel_timing_every(1)
    v = k.get_particles()[abi.PFIELD["V"]]
    zeta = np.stack([k.get_field("ZETA0"), k.get_field("ZETA1")])
    k.close()
    return {
        "steps_per_s": steps / (t1 - t0),
        "ms_per_step": 1e3 * (t1 - t0) / steps,
        "first_step_s": t_cold,
        "warmup_iters": iters,
        "iters_mean": float(np.mean(timed_iters)),
        "relres_max": float(np.max(relres)),
        "dominant": dominant,
        "dominant_ms": ms,
        "dominant_launches": launches,
        "dominant_bytes": nbytes,
        "shares_ms": {c: v_[0] for c, v_ in shares.items()},
        "used_network": used_net,
        "v": v,
        "zeta": zeta,
        "snapshot": snap,
    }


def cold_solves(n_samples):
    """Held-out configurations of the training distribution (seed 4104): cold
    solves to 1e-8 from zero and from the network, FP32 spectral PCG with
    FP64 refinement; FP64 plain CG from zero beside them."""
    import numpy as np
    from icep_kapsel_amd import abi, train_pnet
    from icep_kapsel_amd.gpu import KapselGPU

    k32 = train_pnet.make_context(128)
    if not load_network(k32):
        k32.close()
        return None
    p64 = abi.make_params(128, n_particles=train_pnet.NPART, radius=train_pnet.A, xi=train_pnet.XI,
                          eps_fluid=1.0, eps_particle=0.1, precision=abi.KAPSEL_FP64)
    k64 = KapselGPU(p64)
    rec = []
    for s in range(n_samples):
        pos, ratio, e_hat = train_pnet.sample(4104 * 100000 + s, 128)
        e0 = train_pnet.E_MAG * e_hat
        r = {"ratio": ratio}
        train_pnet.prepare(k32, pos, ratio)
        for mode in ("zero", "network", "network_galerkin"):
            k32.set_field("POTENTIAL", np.zeros(k32.shape))
            k32.synchronize()
            t0 = time.perf_counter()
            if mode != "zero":
                k32.pnet_warm_start(beta_of(ratio), e0, train_pnet.A)
            if mode == "network_galerkin":
                k32.poisson_galerkin_start(e0)
            it, rr = k32.poisson_solve(e0, TOL, MAXIT, abi.PRECOND_SPECTRAL)
            k32.synchronize()
            r[f"fp32_pcg_{mode}_iters"] = it
            r[f"fp32_pcg_{mode}_s"] = time.perf_counter() - t0
        train_pnet.prepare(k64, pos, ratio)
        k64.set_field("POTENTIAL", np.zeros(k64.shape))
        k64.synchronize()
        t0 = time.perf_counter()
        it, rr = k64.poisson_solve(e0, TOL, MAXIT, abi.PRECOND_NONE)
        k64.synchronize()
        r["fp64_cg_zero_iters"] = it
        r["fp64_cg_zero_s"] = time.perf_counter() - t0
        rec.append(r)
        log(f"bench: cold solve {s}: " + ", ".join(f"{a} {b:.4g}" for a, b in r.items()))
    k32.close()
    k64.close()
    keys = [k_ for k_ in rec[0] if k_ != "ratio"]
    return {"samples": n_samples, "seed": 4104, "tol": TOL,
            "mean": {k_: float(np.mean([r[k_] for r in rec])) for k_ in keys}, "per_sample": rec}


def run_spm(precision, steps, warmup):
    """The 128^3 x 200-sphere SPM Navier-Stokes step without the Poisson solve."""
    from icep_kapsel_amd import abi, workloads
    from icep_kapsel_amd.gpu import KapselGPU

    p, st, dt = workloads.spm_config("c2", precision=precision)
    k = KapselGPU(p)
    k.set_particles(st)
    k.init_rigidity()
    t = abi.make_time(dt)
    for _ in range(warmup):
        k.spm_step(t)
    k.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        k.spm_step(t)
    k.synchronize()
    t1 = time.perf_counter()
    k.close()
    return steps / (t1 - t0)


def run_electrolyte(precision, steps, warmup, snapshot=False):
    """The scale + electrolyte workload: 256^3, 2000 spheres a=4 xi=2 with
    surface charge 100 e, two monovalent ion species (Debye length 4), DC
    field 0.01 along x, auto time step; uniform salt start, rigidity
    projection, W warm-up steps, then K timed steps of the Electrolyte branch
    of Time_evolution_hydro (Nernst-Planck Euler step, profile scatter,
    solute rescaling, Coulomb force, projection, gather, particle update)."""
    import numpy as np
    from icep_kapsel_amd import abi, workloads
    from icep_kapsel_amd.gpu import KapselGPU

    p, st, el, dt = workloads.electrolyte_config(precision=precision)
    k = KapselGPU(p)
    k.set_particles(st)
    k.set_electrolyte(el)
    k.init_rigidity()
    k.init_rho_ion()
    t = abi.make_time(dt)
    for _ in range(warmup):
        k.electrolyte_step(t)
    k.synchronize()
    snap = None
    if snapshot:
        snap = {f: k.get_field(f) for f in ("ZETA0", "ZETA1", "CONC0", "CONC1")}
        snap.update({"uk_dc": k.get_uk_dc(), "particles": k.get_particles(), "totals": k.get_total_solute(),
                     "ts": t.ts, "time": t.time})
    k.kernel_timing(abi.KERNEL_CLASSES)
    k.electrolyte_step(t)
    shares = {c: k.kernel_timing_get(c) for c in abi.KERNEL_CLASSES}
    k.kernel_timing([])
    k.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        k.electrolyte_step(t)
    k.synchronize()
    t1 = time.perf_counter()
    v = k.get_particles()[abi.PFIELD["V"]]
    k.close()
    return {"steps_per_s": steps / (t1 - t0), "ms_per_step": 1e3 * (t1 - t0) / steps, "dt": dt,
            "shares_ms": {c: v_[0] for c, v_ in shares.items() if v_[1]},
            "shares_launches": {c: v_[1] for c, v_ in shares.items() if v_[1]},
            "v": v, "snapshot": snap}


def run_cpu_electrolyte(snap, steps):
    """FP64 oracle electrolyte steps from the GPU FP64 state after the warm-up."""
    from icep_kapsel_amd import abi, workloads
    from oracle import oracle as omod

    p, st, el, dt = workloads.electrolyte_config()
    o = omod.Oracle(p)
    o.set_particles(snap["particles"])
    o.set_electrolyte(el)
    for f in ("ZETA0", "ZETA1", "CONC0", "CONC1"):
        o.set_field(f, snap[f])
    o.set_uk_dc(snap["uk_dc"])
    o.set_total_solute(snap["totals"])
    t = abi.make_time(dt, ts=snap["ts"], time=snap["time"])
    t0 = time.perf_counter()
    for _ in range(steps):
        o.electrolyte_step(t)
    t1 = time.perf_counter()
    return steps / (t1 - t0), omod.num_threads()


def run_cpu(snap, steps):
    """FP64 oracle ICEP steps (reference BiCGSTAB, identity preconditioner)
    from the GPU FP64 state after the warm-up steps."""
    from icep_kapsel_amd import abi, workloads
    from oracle import oracle as omod

    p, st, e0, dt = workloads.icep_config()
    o = omod.Oracle(p)
    o.set_particles(snap["particles"])
    o.set_field("ZETA0", snap["ZETA0"])
    o.set_field("ZETA1", snap["ZETA1"])
    o.set_uk_dc(snap["uk_dc"])
    o.set_field("POTENTIAL", snap["POTENTIAL"])
    t = abi.make_time(dt, ts=snap["ts"], time=snap["time"])
    its = []
    t0 = time.perf_counter()
    for _ in range(steps):
        it, rr = o.icep_step(t, e0, TOL, MAXIT, 2)
        its.append(it)
    t1 = time.perf_counter()
    return steps / (t1 - t0), omod.num_threads(), float(sum(its)) / len(its)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def fft_calibration(n=128, reps=5):
    """The oracle's radix-2 FFT against the reference's own Ooura rdft3d +
    rdft3dsort (built from the reference sources into oracle/_ref when they
    are present), one forward and one inverse 3-D transform of the same
    n^3 field, one thread each (fftsg3d.c is built without its thread
    option).  Returns None when the reference build is absent."""
    import numpy as np
    from icep_kapsel_amd import abi
    from oracle import oracle as omod

    try:
        ref = omod.OouraRef()
    except OSError:
        return None
    a = np.zeros((n, n, n + 2))
    a[:, :, :n] = np.random.default_rng(1).standard_normal((n, n, n))
    threads = omod.num_threads()
    omod.set_threads(1)
    o = omod.Oracle(abi.make_params(n))
    try:
        best_o = best_r = float("inf")
        for _ in range(reps):
            o.set_field("WORK0", a)
            t0 = time.perf_counter()
            o.fft_r2c("WORK0")
            o.fft_c2r("WORK0")
            best_o = min(best_o, time.perf_counter() - t0)
            t0 = time.perf_counter()
            ref.a_k2a(ref.a2a_k(a))
            best_r = min(best_r, time.perf_counter() - t0)
    finally:
        omod.set_threads(threads)
    best_all = float("inf")
    for _ in range(reps):
        o.set_field("WORK0", a)
        t0 = time.perf_counter()
        o.fft_r2c("WORK0")
        o.fft_c2r("WORK0")
        best_all = min(best_all, time.perf_counter() - t0)
    return {"grid": n, "threads": 1, "oracle_fft_pair_ms": 1e3 * best_o, "ooura_rdft3d_pair_ms": 1e3 * best_r,
            "oracle_over_ooura": best_o / best_r, "oracle_fft_pair_ms_all_threads": 1e3 * best_all,
            "all_threads": threads}


def pmc_traffic(kernel_class):
    """HBM bytes per launch of the dominant class from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (lines 'traffic <class> <bytes>')."""
    try:
        with open(PMC_SUMMARY) as f:
            for line in f:
                parts = line.split()
                if len(parts) == 3 and parts[0] == "traffic" and parts[1] == kernel_class:
                    return float(parts[2])
    except OSError:
        pass
    return None


def copy_calibration(k):
    from icep_kapsel_amd import abi
    from icep_kapsel_amd.gpu import KapselGPU

    g = KapselGPU(abi.make_params(128, precision=abi.KAPSEL_FP32))
    for _ in range(k):
        g.copy_field("WORK1", "WORK0")
    g.synchronize()
    nbytes = 128 * 128 * 130 * 4
    print(f"copy calibration: {k} copies of {nbytes} bytes read + {nbytes} bytes written", flush=True)
    g.close()


def rel_diff(a, b):
    import numpy as np
    return float(np.abs(a - b).max() / max(1e-300, np.abs(b).max()))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--cold", type=int, default=3, help="held-out cold solves (zero vs network warm start)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=8, help="FP64 oracle ICEP steps of the CPU baseline")
    ap.add_argument("--cpu-el-steps", type=int, default=4, help="FP64 oracle electrolyte steps")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 plain run")
    ap.add_argument("--no-spm", action="store_true", help="skip the SPM-only step timings")
    ap.add_argument("--el-steps", type=int, default=50, help="timed steps of the 256^3 electrolyte workload")
    ap.add_argument("--no-electrolyte", action="store_true", help="skip the 256^3 electrolyte workload")
    ap.add_argument("--only-electrolyte", action="store_true", help="print the electrolyte record only")
    ap.add_argument("--only-spm", action="store_true",
                    help="print the SPM-only record (128^3 x 200 spheres, FP32 and FP64) only")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="K",
                    help="only run K plain field copies (128^3 FP32, 8.52 MB each way) to calibrate "
                         "rocprofv3 FETCH_SIZE/WRITE_SIZE, then exit")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the 128^3 ICEP step runs on a single MI355X).", file=sys.stderr)
        sys.exit(2)
    from icep_kapsel_amd import abi

    if args.only_spm:
        log(f"bench: SPM step (128^3 x 200 spheres, no Poisson) FP32 / FP64, warmup 5, steps {args.steps}")
        print(json.dumps({"spm_ns_step_128_200_spheres": {
            "fp32_steps_per_s": run_spm(abi.KAPSEL_FP32, args.steps, 5),
            "fp64_steps_per_s": run_spm(abi.KAPSEL_FP64, args.steps, 5)}}), flush=True)
        return
    electrolyte = None
    if not args.no_electrolyte:
        electrolyte = electrolyte_record(args)
        if args.only_electrolyte:
            print(json.dumps(electrolyte), flush=True)
            return
    log(f"bench: FP32 + learned ICEP step 128^3 x 64 spheres, warmup {args.warmup}, steps {args.steps}")
    r32 = run_icep(abi.KAPSEL_FP32, abi.PRECOND_SPECTRAL, args.steps, args.warmup, learned=True)
    log(f"bench: fp32 {r32['steps_per_s']:.2f} steps/s, {r32['iters_mean']:.1f} inner it/step, "
        f"warm-up iterations {r32['warmup_iters']}, dominant {r32['dominant']}")
    r64 = None
    split = None
    if not args.no_fp64:
        log("bench: FP64 plain CG ICEP step")
        r64 = run_icep(abi.KAPSEL_FP64, abi.PRECOND_NONE, args.steps, args.warmup, learned=False,
                       timed_class=r32["dominant"], snapshot=not args.no_cpu)
        log(f"bench: fp64 plain {r64['steps_per_s']:.2f} steps/s, {r64['iters_mean']:.1f} it/step, "
            f"warm-up iterations {r64['warmup_iters']}")
        # the same step with one ingredient of the chosen form at a time: the
        # spectral preconditioner in FP64, then FP32 + refinement without the
        # network's cold start
        log("bench: FP64 spectral PCG ICEP step")
        r64s = run_icep(abi.KAPSEL_FP64, abi.PRECOND_SPECTRAL, args.steps, args.warmup, learned=False,
                        timed_class=r32["dominant"])
        log(f"bench: fp64 spectral {r64s['steps_per_s']:.2f} steps/s, {r64s['iters_mean']:.1f} it/step")
        log("bench: FP32 spectral PCG ICEP step without the network")
        r32n = run_icep(abi.KAPSEL_FP32, abi.PRECOND_SPECTRAL, args.steps, args.warmup, learned=False)
        log(f"bench: fp32 without network {r32n['steps_per_s']:.2f} steps/s, {r32n['iters_mean']:.1f} it/step")
        split = {
            "fp64_plain_cg_steps_per_s": r64["steps_per_s"],
            "fp64_spectral_pcg_steps_per_s": r64s["steps_per_s"],
            "fp64_spectral_pcg_iters_per_step": r64s["iters_mean"],
            "fp32_spectral_pcg_refined_no_network_steps_per_s": r32n["steps_per_s"],
            "fp32_spectral_pcg_refined_no_network_iters_per_step": r32n["iters_mean"],
            "fp32_spectral_pcg_refined_network_cold_start_steps_per_s": r32["steps_per_s"],
            "factor_preconditioner": r64s["steps_per_s"] / r64["steps_per_s"],
            "factor_precision": r32n["steps_per_s"] / r64s["steps_per_s"],
            "factor_network": r32["steps_per_s"] / r32n["steps_per_s"],
            "first_step_s_network": r32["first_step_s"],
            "first_step_s_no_network": r32n["first_step_s"],
            "first_step_iters_network": r32["warmup_iters"][0] if r32["warmup_iters"] else None,
            "first_step_iters_no_network": r32n["warmup_iters"][0] if r32n["warmup_iters"] else None,
        }
    cold = cold_solves(args.cold) if args.cold else None
    spm = None
    if not args.no_spm:
        log("bench: SPM step (128^3 x 200 spheres, no Poisson) FP32 / FP64")
        spm = {"fp32_steps_per_s": run_spm(abi.KAPSEL_FP32, 50, 5),
               "fp64_steps_per_s": run_spm(abi.KAPSEL_FP64, 50, 5)}
    cpu = None
    if not args.no_cpu and r64 is not None:
        log(f"bench: CPU oracle, {args.cpu_steps} ICEP steps from the GPU FP64 state after the warm-up")
        cps, cores, cit = run_cpu(r64["snapshot"], args.cpu_steps)
        cpu = {"value": cps, "unit": "steps/s", "cores": cores, "kind": "port",
               "sample": f"{args.cpu_steps} FP64 ICEP steps of the same 128^3 x 64-sphere workload from the "
                         f"state after {args.warmup} warm-up steps, reference BiCGSTAB to 1e-8 ({cit:.1f} "
                         "iterations per step), OpenMP C++ oracle",
               "cpu_model": cpu_model(), "nproc": os.cpu_count()}
        log(f"bench: cpu {cps:.4f} steps/s on {cores} threads ({cit:.1f} BiCGSTAB iterations per step)")
        cal = fft_calibration()
        if cal:
            cpu["fft_calibration"] = cal
            log(f"bench: oracle FFT / reference Ooura rdft3d at 128^3, 1 thread: {cal['oracle_over_ooura']:.3f}")
    avg_ms = r32["dominant_ms"] / max(1, r32["dominant_launches"])
    achieved = (r32["dominant_bytes"] / 1e9) / (r32["dominant_ms"] / 1e3) if r32["dominant_ms"] > 0 else 0.0
    traffic = pmc_traffic(r32["dominant"])
    out = {
        "metric": METRIC,
        "value": r32["steps_per_s"],
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r32["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": "icep_step_128^3_64_dielectric_spheres_a6_xi2_eps0.1_E0.1", "grid": 128,
                   "particles": 64, "radius": 6.0, "xi": 2.0, "eps_ratio": 0.1, "e0_magnitude": 0.1,
                   "dt": 0.05, "poisson_tol": TOL, "seed": 3003,
                   "solver": "fp32 spectral PCG in fp64 refinement, learned cold start (first solve only)",
                   "baseline_solver": "fp64 plain CG", "parallelism": "single_gpu"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic,
                     "kernel": KERNEL_OF_CLASS[r32["dominant"]], "kernel_class": r32["dominant"],
                     "avg_launch_ms": avg_ms,
                     "algorithmic_bytes_per_launch": r32["dominant_bytes"] / max(1, r32["dominant_launches"]),
                     "launches": r32["dominant_launches"], "sampled_every": TIMING_EVERY},
        "cpu_baseline": cpu,
        "fp32_inner_iters_per_step": r32["iters_mean"],
        "fp32_relres_max": r32["relres_max"],
        "fp32_first_step_s": r32["first_step_s"],
        "learned_warm_start_used": r32["used_network"],
        "kernel_class_ms_per_step_fp32": r32["shares_ms"],
    }
    if r64:
        out.update({
            "fp64_plain_steps_per_s": r64["steps_per_s"],
            "fp64_plain_ms_per_step": r64["ms_per_step"],
            "fp64_plain_iters_per_step": r64["iters_mean"],
            "fp64_plain_relres_max": r64["relres_max"],
            "fp64_plain_first_step_s": r64["first_step_s"],
            "speedup_vs_fp64_plain": r32["steps_per_s"] / r64["steps_per_s"],
            "particle_velocity_rel_err_vs_fp64": rel_diff(r32["v"], r64["v"]),
            "vorticity_rel_err_vs_fp64": rel_diff(r32["zeta"], r64["zeta"]),
            "kernel_class_ms_per_step_fp64": r64["shares_ms"],
        })
    if split:
        out["solver_split"] = split
    if cold:
        out["cold_solve_1e-8"] = cold["mean"]
    if spm:
        out["spm_ns_step_128_200_spheres"] = spm
    if electrolyte:
        out["electrolyte_256_2000_spheres"] = electrolyte
    print(json.dumps(out), flush=True)


def electrolyte_record(args):
    """The 256^3 x 2000-sphere electrolyte workload, FP32 and FP64, and one
    FP64 oracle step from the same state."""
    from icep_kapsel_amd import abi

    log(f"bench: electrolyte step 256^3 x 2000 charged spheres, warmup 3, steps {args.el_steps}")
    e32 = run_electrolyte(abi.KAPSEL_FP32, args.el_steps, 3)
    log(f"bench: electrolyte fp32 {e32['steps_per_s']:.2f} steps/s")
    e64 = run_electrolyte(abi.KAPSEL_FP64, args.el_steps, 3, snapshot=not args.no_cpu)
    log(f"bench: electrolyte fp64 {e64['steps_per_s']:.2f} steps/s")
    rec = {"workload": "electrolyte_step_256^3_2000_spheres_a4_xi2_q100_debye4_E0.01",
           "fp32_steps_per_s": e32["steps_per_s"], "fp32_ms_per_step": e32["ms_per_step"],
           "fp64_steps_per_s": e64["steps_per_s"], "fp64_ms_per_step": e64["ms_per_step"],
           "dt": e32["dt"], "steps": args.el_steps, "warmup": 3,
           "particle_velocity_rel_err_fp32_vs_fp64": rel_diff(e32["v"], e64["v"]),
           "kernel_class_ms_per_step_fp32": e32["shares_ms"],
           "kernel_class_launches_per_step_fp32": e32["shares_launches"],
           "kernel_class_ms_per_step_fp64": e64["shares_ms"]}
    if not args.no_cpu:
        log(f"bench: CPU oracle, {args.cpu_el_steps} electrolyte steps from the GPU FP64 state after the warm-up")
        cps, cores = run_cpu_electrolyte(e64["snapshot"], args.cpu_el_steps)
        rec["cpu_baseline"] = {"value": cps, "unit": "steps/s", "cores": cores, "kind": "port",
                               "sample": f"{args.cpu_el_steps} FP64 electrolyte steps of the same 256^3 x "
                                         "2000-sphere workload from the state after 3 warm-up steps, OpenMP C++ "
                                         "oracle"}
        log(f"bench: electrolyte cpu {cps:.4f} steps/s on {cores} threads")
    return rec


if __name__ == "__main__":
    main()
