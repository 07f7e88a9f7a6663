"""Throughput of the SAMoS particle pipeline on one MI355X.

Workload (one "step" = one iteration of the reference's run loop,
src/samos.cpp:878-970): Lennard-Jones Langevin BAOAB dynamics in a periodic 2-D
box, N = 1024^2 = 1,048,576 particles on a jittered square lattice at number
density 0.7 (jitter +-0.1 sigma), Maxwell velocities at T = 1; LJ sigma = eps = 1,
rc = 2.5 shifted; gamma = 1, T = 1, dt = 0.002; neighbour list rcut 2.5, pad 0.5
(the reference's default skin).  All arithmetic FP64.  Inputs are synthetic and
resident in HBM before timing.

Prints one JSON line: particle-steps/s, the roofline of the dominant kernel
(the fused force + BAOAB step kernel, timed with HIP events on its stream) and
the single-threaded CPU baseline (the oracle built with the reference's
Release flags) timed on this host over a bounded sample.
"""
import argparse
import json
import math
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
ALG_BYTES_PER_PARTICLE_STEP = 64  # read + write x, y, vx, vy (FP64)
NSIDE = 1024


def build_system(samos_amd, sc, nside, pad=0.3, rebuild="coop", batch=4):
    st = sc.lj_lattice(nside, rho=0.7, jitter=0.1, T=1.0, seed=2024)
    s = sc.load(samos_amd.System, st)
    s.set_option("rebuild_coop", 1 if rebuild == "coop" else 0)
    s.set_option("gather_batch", batch)
    s.nlist(2.5, pad)
    s.pair_lj(1.0, 1.0, 2.5, shifted=True)
    s.constraint("plane")
    s.set_timestep(0.002)
    s.integrator_langevin(gamma=1.0, seed=1, temp=(0, 1.0, 1.0, 1))
    return s, st


def timed_run(s, steps, flags):
    """run(steps - 1) executes `steps` loop iterations."""
    s.sync()
    t0 = time.perf_counter()
    s.run(steps - 1, flags)
    s.sync()
    return time.perf_counter() - t0


def cpu_baseline(nside, steps, pad):
    """Reference-equivalent CPU path (oracle/liboracle_fast.so: AoS particle
    records, half neighbour list, -O3 -ffast-math), one thread pinned to one
    core, timed over `steps` loop iterations after the initial nlist build."""
    code = r"""
import os, sys, time, json
sys.path.insert(0, %r)
os.sched_setaffinity(0, {sorted(os.sched_getaffinity(0))[0]})
import oracle
from tests import scenarios as sc
st = sc.lj_lattice(%d, rho=0.7, jitter=0.1, T=1.0, seed=2024)
s = sc.load(lambda n, b, p: oracle.OracleSystem(n, b, p, fast=True), st)
s.nlist(2.5, %r)
s.pair_lj(1.0, 1.0, 2.5, shifted=True)
s.constraint("plane")
s.set_timestep(0.002)
s.integrator_langevin(gamma=1.0, seed=1, temp=(0, 1.0, 1.0, 1))
t0 = time.perf_counter()
s.run(%d - 1)
dt = time.perf_counter() - t0
print(json.dumps({"seconds": dt, "n": st["n"], "steps": %d}))
""" % (ROOT, nside, pad, steps, steps)
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, check=True)
    r = json.loads(out.stdout.strip().splitlines()[-1])
    cpu = ""
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                cpu = ln.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {
        "value": r["n"] * r["steps"] / r["seconds"],
        "unit": "particle-steps/s",
        "cores": 1,
        "kind": "port",
        "sample": "same workload, N=%d, %d run-loop iterations after the initial neighbour-list build "
                  "(%.1f s), one thread; host: %s, %d logical CPUs" % (r["n"], r["steps"], r["seconds"], cpu,
                                                                        os.cpu_count() or 0),
    }


def copy_probe(mib, reps=10):
    """Plain device-to-device copies of a known size: the reference point that
    calibrates rocprofv3's FETCH_SIZE / WRITE_SIZE counters."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    nbytes = mib << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    for p in (a, b):
        if hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 1, ctypes.c_size_t(nbytes))
    for _ in range(reps):
        if hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) != 0:  # hipMemcpyDeviceToDevice
            raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_bytes": nbytes, "reps": reps}))


def traffic_from_profiles():
    """HBM bytes per launch of the fused step kernel from the committed
    rocprofv3 --pmc summary (FETCH_SIZE + WRITE_SIZE, copy-kernel calibrated)."""
    p = os.path.join(ROOT, "profiles", "fused_baoab_lj_N1048576_traffic.json")
    try:
        return json.load(open(p))["bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--nside", type=int, default=NSIDE)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=20)
    ap.add_argument("--variants", action="store_true", help="also time the FP32-force and faithful schedules")
    # skin: the reference's default padding (src/system/defaults.hpp:37); the
    # fastest of 0.3-0.8 in a sweep on MI355X (0.4-0.6 within 1%)
    ap.add_argument("--pad", type=float, default=0.5, help="neighbour-list skin")
    ap.add_argument("--rebuild", choices=["coop", "stages"], default="stages")
    ap.add_argument("--batch", type=int, choices=[4, 8], default=4)
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run device-to-device copies of this size (HBM counter calibration) and exit")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: the pipeline runs one particle system per GPU and this build has no "
              "multi-GPU path (the sharding design is described in DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import samos_amd
    from tests import scenarios as sc

    s, st = build_system(samos_amd, sc, args.nside, args.pad, args.rebuild, args.batch)
    n = st["n"]
    print("bench: N=%d box %.3f, warmup %d, steps %d" % (n, st["box"][0], args.warmup, args.steps),
          file=sys.stderr, flush=True)
    if args.warmup > 0:
        timed_run(s, args.warmup, 0)
    secs = timed_run(s, args.steps, 0)
    value = n * args.steps / secs
    print("bench: timed %.3f s -> %.4e particle-steps/s" % (secs, value), file=sys.stderr, flush=True)

    # Dominant kernel: the fused LJ force + BAOAB step.  HIP events bracket each
    # launch on the context's stream (step-by-step launches, same kernels).
    s.profile(True)
    prof_steps = min(args.steps, 500)
    timed_run(s, prof_steps, 0)
    ms, launches = s.profile_read("fused")
    rb_ms, rb_launches = s.profile_read("rebuild")
    s.profile(False)
    avg_s = ms / 1e3 / max(launches, 1)
    alg_bytes = ALG_BYTES_PER_PARTICLE_STEP * n
    achieved = alg_bytes / avg_s / 1e9
    info = s.nlist_info()

    extra = {}
    if args.variants:
        s32, _ = build_system(samos_amd, sc, args.nside, args.pad, args.rebuild, args.batch)
        timed_run(s32, 50, samos_amd.RUN_FP32_FORCE)
        t32 = timed_run(s32, args.steps, samos_amd.RUN_FP32_FORCE)
        extra["fp32_force_value"] = n * args.steps / t32
        s32.close()
        tf = timed_run(s, min(args.steps, 500), samos_amd.RUN_FAITHFUL)
        extra["faithful_two_force_evals_value"] = n * min(args.steps, 500) / tf

    out = {
        "metric": "particle-steps/s at N=1M",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": secs * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "lj_langevin_plane: N=%d LJ (sigma=eps=1, rc=2.5 shifted) Langevin BAOAB gamma=1 T=1 "
                        "dt=0.002, periodic 2-D box L=%.4f (density 0.7), nlist rcut=2.5 pad=%g" % (n, st["box"][0], args.pad),
            "particles": n,
            "nlist_pad": args.pad,
            "rebuild": args.rebuild,
            "gather_batch": args.batch,
            "schedule": "fused: one force evaluation per step (the loop-head re-evaluation of identical "
                        "positions is not repeated)",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic_from_profiles(),
            "kernel": "k_baoab_fused<LJ>",
            "avg_launch_us": avg_s * 1e6,
            "launches_timed": launches,
            "alg_bytes_per_launch": alg_bytes,
        },
        "nlist": {"builds": info["builds"], "stride": info["stride"], "rebuild_ms_total": rb_ms,
                  "rebuild_stage_launches": rb_launches},
    }
    out.update(extra)
    s.close()
    if not args.no_cpu:
        cpu_side = min(args.nside, NSIDE)
        out["cpu_baseline"] = cpu_baseline(cpu_side, args.cpu_steps, args.pad)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
