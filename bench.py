"""SPM step throughput on one MI355X.

python bench.py [--gpus 1] [--steps K] [--warmup W]

Workload: the 128^3 SPM Navier-Stokes step with 200 spheres (a=4, xi=2,
Gaussian particle velocities std 1e-3, random non-overlapping positions):
slaved-Euler NS update, profile scatter, hydrodynamic gather, LJ, particle
update, rigidity constraint and projection -- Time_evolution_hydro of the
reference (sp_3d_ns.cxx:24-186).  Inputs are resident in HBM before timing.
The FP32 path (float fields, FP64 particle reductions) is the headline value;
the FP64 path is timed the same way beside it.  Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "SPM steps/s at 128^3: FP32+learned-precond vs FP64 plain, same error bound"
HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def run_gpu(precision, steps, warmup, timed_class=None):
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
    # find the kernel class with the largest share of device time
    k.kernel_timing(abi.KERNEL_CLASSES)
    k.spm_step(t)
    shares = {c: k.kernel_timing_get(c) for c in abi.KERNEL_CLASSES}
    k.kernel_timing([])
    dominant = timed_class or max(shares, key=lambda c: shares[c][0])
    # timed region: only the dominant class carries HIP events
    k.kernel_timing([dominant])
    k.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        k.spm_step(t)
    k.synchronize()
    t1 = time.perf_counter()
    ms, launches, nbytes = k.kernel_timing_get(dominant)
    k.kernel_timing([])
    k.close()
    return {
        "steps_per_s": steps / (t1 - t0),
        "ms_per_step": 1e3 * (t1 - t0) / steps,
        "dominant": dominant,
        "dominant_ms": ms,
        "dominant_launches": launches,
        "dominant_bytes": nbytes,
        "shares_ms": {c: v[0] for c, v in shares.items()},
    }


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


def run_cpu(steps):
    from icep_kapsel_amd import abi, workloads
    from oracle import oracle as omod

    p, st, dt = workloads.spm_config("c2")
    o = omod.Oracle(p)
    o.set_particles(st)
    o.init_rigidity()
    t = abi.make_time(dt)
    o.spm_step(t)  # warm-up
    t0 = time.perf_counter()
    for _ in range(steps):
        o.spm_step(t)
    t1 = time.perf_counter()
    return steps / (t1 - t0), omod.num_threads()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 comparison run")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="K",
                    help="only run K plain field copies (128^3 FP32, 8.52 MB each way) to calibrate "
                         "rocprofv3 FETCH_SIZE/WRITE_SIZE, then exit")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the SPM step at these sizes runs on a single MI355X).", file=sys.stderr)
        sys.exit(2)
    from icep_kapsel_amd import abi

    log(f"bench: fp32 SPM step 128^3 x 200 spheres, warmup {args.warmup}, steps {args.steps}")
    r32 = run_gpu(abi.KAPSEL_FP32, args.steps, args.warmup)
    log(f"bench: fp32 {r32['steps_per_s']:.1f} steps/s; dominant {r32['dominant']}")
    r64 = None
    if not args.no_fp64:
        log("bench: fp64 SPM step")
        r64 = run_gpu(abi.KAPSEL_FP64, args.steps, args.warmup, timed_class=r32["dominant"])
        log(f"bench: fp64 {r64['steps_per_s']:.1f} steps/s")
    cpu = None
    if not args.no_cpu:
        log(f"bench: CPU oracle baseline, {args.cpu_steps} steps after 1 warm-up")
        cps, cores = run_cpu(args.cpu_steps)
        cpu = {"value": cps, "unit": "steps/s", "cores": cores, "kind": "port",
               "sample": f"{args.cpu_steps} FP64 SPM steps of the same 128^3 x 200-sphere workload "
                         "(after 1 warm-up step), OpenMP C++ oracle"}
        log(f"bench: cpu {cps:.3f} steps/s on {cores} threads")
    avg_ms = r32["dominant_ms"] / max(1, r32["dominant_launches"])
    achieved = (r32["dominant_bytes"] / 1e9) / (r32["dominant_ms"] / 1e3) if r32["dominant_ms"] > 0 else 0.0
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
        "config": {"workload": "spm_ns_step_128^3_200_spheres_a4_xi2", "grid": 128, "particles": 200,
                   "radius": 4.0, "xi": 2.0, "dt": 0.05, "poisson": "not in this step (NS branch)",
                   "parallelism": "single_gpu"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": f"class {r32['dominant']}", "avg_launch_ms": avg_ms,
                     "launches": r32["dominant_launches"]},
        "cpu_baseline": cpu,
        "fp64_steps_per_s": r64["steps_per_s"] if r64 else None,
        "fp64_ms_per_step": r64["ms_per_step"] if r64 else None,
        "fp32_over_fp64": r32["steps_per_s"] / r64["steps_per_s"] if r64 else None,
        "kernel_class_ms_per_step_fp32": r32["shares_ms"],
        "kernel_class_ms_per_step_fp64": r64["shares_ms"] if r64 else None,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
