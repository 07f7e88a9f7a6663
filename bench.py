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

import numpy as np

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


def host_cpu():
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return ""


def oracle_timed(body):
    """Runs `body` (Python source using `oracle`, `sc`, `time`, `np`) in a child
    process pinned to one core with the reference-flag oracle build; the body
    prints one JSON line, which is returned."""
    code = ("import os, sys, time, json\nimport numpy as np\nsys.path.insert(0, %r)\n"
            "os.sched_setaffinity(0, {sorted(os.sched_getaffinity(0))[0]})\n"
            "import oracle\nfrom tests import scenarios as sc\n" % ROOT) + body
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, check=True)
    return json.loads(out.stdout.strip().splitlines()[-1])


def cpu_baseline_ensemble(sample_every=32, steps=400):
    """The ensemble's replicas as the reference runs them, one `samos` per
    replica: the oracle's reference-flag build on one core, over every
    `sample_every`-th replica of the same workload for `steps` run-loop
    iterations each."""
    r = oracle_timed(r"""
states, p = sc.c4_ensemble_workload()
tot_ps, tot_t, picked = 0, 0.0, []
for r in range(0, len(states), %d):
    st = states[r]
    o = sc.load(lambda n, b, q: oracle.OracleSystem(n, b, q, fast=True), st)
    o.nlist(2.4, 0.3)
    o.pair_soft(p["k"], p["a"])
    if p["J"][r] != 0.0:
        o.align_polar(float(p["J"][r]), p["rcut_align"])
    o.constraint("sphere", st["R"])
    o.set_timestep(p["dt"])
    o.integrator_brownian(v0=float(p["v0"][r]), nu=float(p["nu"][r]), mu=p["mu"], mur=p["mur"], seed=int(p["seed"][r]))
    t0 = time.perf_counter()
    o.run(%d - 1)
    tot_t += time.perf_counter() - t0
    tot_ps += st["n"] * %d
    picked.append(st["n"])
    o.close()
print(json.dumps({"ps": tot_ps, "seconds": tot_t, "counts": picked}))
""" % (sample_every, steps, steps))
    return {"value": r["ps"] / r["seconds"], "unit": "particle-steps/s", "cores": 1, "kind": "port",
            "sample": "every %dth replica of the same ensemble (%d replicas, N = %s), each run alone for %d run-loop "
                      "iterations (%.1f s), one thread; host: %s" % (sample_every, len(r["counts"]),
                                                                     r["counts"], steps, r["seconds"], host_cpu())}


def cpu_baseline_fire(iterations=40000, n=4096):
    """One configuration of the FIRE workload relaxed by the oracle's
    reference-flag build on one core (IntegratorFIRE, which never stops by
    itself): `iterations` run-loop iterations."""
    r = oracle_timed(r"""
st = sc.plane_replicas([%d], seed=5)[0]
o = sc.load(lambda n, b, q: oracle.OracleSystem(n, b, q, fast=True), st)
o.nlist(2.4, 0.3)
o.pair_soft(1.0, 2.0, use_radii=True)
o.constraint("plane")
o.set_timestep(0.01)
o.integrator_fire(alpha=0.1, dt_max=0.1, f_tol=0.0, e_tol=0.0)
t0 = time.perf_counter()
o.run(%d - 1)
print(json.dumps({"seconds": time.perf_counter() - t0, "n": st["n"], "iters": %d}))
""" % (n, iterations, iterations))
    return {"value": r["n"] * r["iters"] / r["seconds"], "unit": "particle-iterations/s", "cores": 1, "kind": "port",
            "sample": "the first configuration of the same batch (N = %d), %d FIRE run-loop iterations (%.1f s), one "
                      "thread; host: %s" % (r["n"], r["iters"], r["seconds"], host_cpu())}


ENS_BYTES_PER_PARTICLE_STEP = 120   # read x, y, z, nx, ny, nz; write those and vx, vy, vz (FP64)
FIRE_BYTES_PER_PARTICLE_ITER = 72   # read x, y, vx, vy, r; write x, y, vx, vy (FP64)


def plain_replica_system(samos_amd, sc, st, p, r):
    """Replica r as its own single-system run (the drop-in path, one conf per replica)."""
    s = sc.load(samos_amd.System, st)
    s.nlist(2.4, 0.3)
    s.pair_soft(p["k"], p["a"])
    if p["J"][r] != 0.0:
        s.align_polar(float(p["J"][r]), p["rcut_align"])
    s.constraint("sphere", st["R"])
    s.set_timestep(p["dt"])
    s.integrator_brownian(v0=float(p["v0"][r]), nu=float(p["nu"][r]), mu=p["mu"], mur=p["mur"],
                          seed=int(p["seed"][r]))
    return s


def bench_ensemble(samos_amd, sc, steps, plain_steps, nrep=256, pad=None, cpu=True):
    """Batched active-sphere ensemble: every replica advanced by one fused launch
    per step, against the same replicas run one at a time."""
    states, p = sc.c4_ensemble_workload(nrep=nrep)
    counts = np.array([st["n"] for st in states])
    n_tot = int(counts.sum())
    e = sc.load_ensemble(samos_amd.Ensemble, states, "sphere")
    e.set_params(**p)
    if pad is None:
        e.nlist(2.4, 0.3)
        tune = e.autotune(trial_steps=128, pads=(0.3, 0.45, 0.6, 0.8, 1.0))
    else:
        e.nlist(2.4, pad)
        tune = None
    print("bench: ensemble %d replicas, %d particles, autotune %s" % (nrep, n_tot, tune), file=sys.stderr, flush=True)
    e.run_brownian(63)
    e.sync()
    t0 = time.perf_counter()
    e.run_brownian(steps - 1)
    e.sync()
    secs = time.perf_counter() - t0
    value = n_tot * steps / secs
    print("bench: ensemble %d steps in %.3f s -> %.4e particle-steps/s" % (steps, secs, value), file=sys.stderr,
          flush=True)
    e.profile(True)
    e.run_brownian(199)
    ms, launches = e.profile_read("step")
    rb_ms, rb_launches = e.profile_read("rebuild")
    e.profile(False)
    info = e.nlist_info()
    e.close()
    avg_s = ms / 1e3 / max(launches, 1)
    achieved = ENS_BYTES_PER_PARTICLE_STEP * n_tot / avg_s / 1e9
    plain_value = one_value = float("nan")
    if plain_steps > 0:
        # the same replicas as plain single-system runs, one after another
        plain_t, plain_ps = 0.0, 0
        for r, st in enumerate(states):
            s = plain_replica_system(samos_amd, sc, st, p, r)
            s.run(9)
            s.sync()
            t0 = time.perf_counter()
            s.run(plain_steps - 1)
            s.sync()
            plain_t += time.perf_counter() - t0
            plain_ps += st["n"] * plain_steps
            s.close()
        plain_value = plain_ps / plain_t
        # the ensemble engine on one replica at a time (batching gain alone)
        one_t, one_ps = 0.0, 0
        for r, st in enumerate(states):
            q = {k: (v[r:r + 1] if isinstance(v, np.ndarray) else v) for k, v in p.items()}
            e1 = sc.load_ensemble(samos_amd.Ensemble, [st], "sphere")
            e1.set_params(**q)
            e1.nlist(2.4, tune["pad"] if tune else pad)
            if tune:
                e1.set_option("gather_batch", tune["gather_batch"])
            e1.run_brownian(63)
            e1.sync()
            t0 = time.perf_counter()
            e1.run_brownian(plain_steps - 1)
            e1.sync()
            one_t += time.perf_counter() - t0
            one_ps += st["n"] * plain_steps
            e1.close()
        one_value = one_ps / one_t
        print("bench: plain per-replica %.4e, ensemble engine one replica at a time %.4e particle-steps/s"
              % (plain_value, one_value), file=sys.stderr, flush=True)
    return {
        "workload": "active_sphere_ensemble: %d replicas, N_r log-uniform in [1024, 65536] (total %d), "
                    "phi_r U[0.5, 1.2], R_r = sqrt(N_r / (4 phi_r)); soft k=10 a=2, polar J U[0, 1] rcut 2.4, "
                    "Brownian v0 log-U[0.01, 1], nu log-U[0.01, 1], dt 0.005; nlist rcut 2.4" % (nrep, n_tot),
        "value": value,
        "unit": "particle-steps/s",
        "steps": steps,
        "ms_per_step": secs * 1e3 / steps,
        "autotune": tune,
        "plain_per_replica_value": plain_value,
        "plain_per_replica_steps": plain_steps,
        "vs_plain_per_replica": value / plain_value,
        "ensemble_engine_one_replica_at_a_time_value": one_value,
        "vs_one_replica_at_a_time": value / one_value,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "kernel": "k_ens_brownian",
                     "avg_launch_us": avg_s * 1e6, "launches_timed": launches,
                     "alg_bytes_per_launch": ENS_BYTES_PER_PARTICLE_STEP * n_tot},
        "nlist": {"builds": info["builds"], "stride": info["stride"], "buckets": info["buckets"],
                  "rebuild_ms_total": rb_ms, "rebuild_sequences": rb_launches},
        "cpu_baseline": cpu_baseline_ensemble() if cpu and nrep == 256 else None,
    }


def c1_system(samos_amd, sc):
    """Active polar soft disks in a periodic 2-D box (N = 4096, packing 0.8)."""
    st = sc.plane_disks(4096, phi=0.8, seed=101)
    s = sc.load(samos_amd.System, st)
    s.nlist(2.4, 0.3)
    s.pair_soft(1.0, 2.0)
    s.align_polar(0.1, 2.4)
    s.constraint("plane")
    s.set_timestep(0.01)
    s.integrator_brownian(v0=0.1, nu=0.1, mu=1.0, seed=11)
    return s, st


def c2_system(samos_amd, sc):
    """Active nematic soft disks on a sphere (N = 100000, R = sqrt(N / 4))."""
    st = sc.sphere_disks(100000, phi=1.0, seed=102)
    s = sc.load(samos_amd.System, st)
    s.nlist(2.4, 0.3)
    s.pair_soft(10.0, 2.0)
    s.align_nematic(0.5, 2.4)
    s.constraint("sphere", st["R"])
    s.set_timestep(0.005)
    s.integrator_brownian(v0=0.5, nu=0.05, mu=1.0, seed=12)
    return s, st


def c1_large_system(samos_amd, sc):
    """Active polar soft disks in a periodic 2-D box (N = 1,048,576, packing 0.8): the physics of the
    4096-particle case at the headline size."""
    st = sc.plane_disks(1 << 20, phi=0.8, seed=103)
    s = sc.load(samos_amd.System, st)
    s.nlist(2.4, 0.3)
    s.pair_soft(1.0, 2.0)
    s.align_polar(0.1, 2.4)
    s.constraint("plane")
    s.set_timestep(0.01)
    s.integrator_brownian(v0=0.1, nu=0.1, mu=1.0, seed=13)
    return s, st


def bench_tiled_brownian(samos_amd, sc, steps):
    """The fused Brownian step with polar alignment on the cell-tiled engine (positions and directors
    of each tile's halo staged in LDS) against the neighbour-list kernels, same system and steps."""
    out = {}
    for label, flags, kernel in (("tiles", samos_amd.RUN_TILES, "k_tile_brownian<soft, polar>"),
                                 ("list", 0, "k_brownian_fused")):
        s, st = c1_large_system(samos_amd, sc)
        n = st["n"]
        timed_run(s, 100, flags)
        secs = timed_run(s, steps, flags)
        s.profile(True)
        timed_run(s, 200, flags)
        ms, launches = s.profile_read("fused")
        s.profile(False)
        avg = ms / 1e3 / max(launches, 1)
        achieved = 80 * n / avg / 1e9
        out[label] = {"value": n * steps / secs, "ms_per_step": secs * 1e3 / steps,
                      "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                   "frac": achieved / HBM_PEAK_GBS, "kernel": kernel, "avg_launch_us": avg * 1e6,
                                   "launches_timed": launches, "alg_bytes_per_launch": 80 * n}}
        if label == "tiles":
            out["tiles_used"] = s.tile_info()
        s.close()
    print("bench: tiled Brownian %s" % out, file=sys.stderr, flush=True)
    return {"workload": c1_large_system.__doc__.strip(), "particles": 1 << 20, "steps": steps,
            "unit": "particle-steps/s", **out}


def bench_single(samos_amd, sc, which, steps, cpu_steps, extra_flags=0):
    """One system through the drop-in single-system path (the `run` loop).
    extra_flags: e.g. RUN_NO_GRAPH for a profiler run (step-by-step launches)."""
    build = c1_system if which == "c1" else c2_system
    out = {}
    variants = [("f64", 0)] + ([("f32_force", samos_amd.RUN_FP32_FORCE)] if which == "c2" else [])
    alg = 80 if which == "c1" else 120   # algorithmic bytes per particle-step
    for label, flags in variants:
        flags |= extra_flags
        s, st = build(samos_amd, sc)
        s.run(99, flags)
        secs = timed_run(s, steps, flags)
        out[label] = {"value": st["n"] * steps / secs, "ms_per_step": secs * 1e3 / steps}
        # step kernel (fused loop-head forces + Brownian step): HIP events on the
        # context's stream around each launch of a step-by-step pass
        s.profile(True)
        timed_run(s, 200, flags)
        ms, launches = s.profile_read("fused")
        s.profile(False)
        if launches:
            avg = ms / 1e3 / launches
            out[label]["roofline"] = {"bound": "hbm", "achieved": alg * st["n"] / avg / 1e9, "peak": HBM_PEAK_GBS,
                                      "unit": "GB/s", "frac": alg * st["n"] / avg / 1e9 / HBM_PEAK_GBS,
                                      "kernel": "k_brownian_fused", "avg_launch_us": avg * 1e6,
                                      "launches_timed": launches, "alg_bytes_per_launch": alg * st["n"]}
        s.close()
    print("bench: %s %s" % (which, out), file=sys.stderr, flush=True)
    res = {"workload": build.__doc__.strip(), "particles": st["n"], "steps": steps, "unit": "particle-steps/s",
           **out}
    if cpu_steps > 0:
        code = r"""
import os, sys, time, json
sys.path.insert(0, %r)
os.sched_setaffinity(0, {sorted(os.sched_getaffinity(0))[0]})
import oracle, bench
from tests import scenarios as sc
class O:
    System = staticmethod(lambda n, b, p: oracle.OracleSystem(n, b, p, fast=True))
    RUN_FP32_FORCE = 4
s, st = bench.%s_system(O, sc)
t0 = time.perf_counter()
s.run(%d - 1)
print(json.dumps({"seconds": time.perf_counter() - t0, "n": st["n"], "steps": %d}))
""" % (ROOT, which, cpu_steps, cpu_steps)
        r = json.loads(subprocess.run([sys.executable, "-c", code], capture_output=True, text=True,
                                      check=True).stdout.strip().splitlines()[-1])
        res["cpu_baseline"] = {"value": r["n"] * r["steps"] / r["seconds"], "unit": "particle-steps/s", "cores": 1,
                               "kind": "port", "sample": "same system, %d run-loop iterations (%.1f s), one thread"
                                                         % (r["steps"], r["seconds"])}
    return res


WARM_WEIGHTS = os.path.join(ROOT, "samos_amd", "data", "warm_start_mpnn_n4096_phi088.npz")


def fire_batch(samos_amd, sc, states):
    e = sc.load_ensemble(samos_amd.Ensemble, states, "plane")
    e.set_params(k=1.0, a=2.0, use_radii=1.0, dt=0.01)
    e.nlist(2.4, 0.3)
    return e


def warm_start_ab(samos_amd, sc, states, fs_plain, secs_plain, max_iter, weights, scales=(1.0,)):
    """The same configurations relaxed from the network's predicted positions:
    inference, displacement, list rebuild and FIRE are all timed.  An answer
    counts only when its replica stopped at max|F| < 1e-10 with an energy equal
    to plain FIRE's to 1e-12 (relative); every other configuration is charged
    the fallback, plain FIRE from its raw start, run and timed as one batch."""
    z = np.load(weights, allow_pickle=False)
    meta = json.loads(str(z["meta"]))
    mp = samos_amd.Mpnn(seed=0)
    mp.set_params(z["params"])
    out = {"weights": os.path.relpath(weights, ROOT), "trained_on_configs": meta.get("train_configs"),
           "val_loss": meta["history"][-1]["val_loss"] if meta.get("history") else None,
           "val_loss_zero_prediction": meta.get("val_loss_zero_prediction")}
    for scale in scales:
        e = fire_batch(samos_amd, sc, states)
        e.sync()
        t0 = time.perf_counter()
        mp.warm_start(e, scale)
        e.sync()
        t_inf = time.perf_counter() - t0
        e.fire_setup(dt=0.01, alpha=0.1, dt_max=0.1, f_tol=0.0, stop_fmax=1e-10)
        rem, iters = e.fire_run(max_iter, check=100)
        e.sync()
        secs = time.perf_counter() - t0
        fs = e.fire_state()
        e.close()
        stopped = fs["stopped"] == 1
        same = stopped & (np.abs(fs["energy"] - fs_plain["energy"]) <= 1e-12 * np.abs(fs_plain["energy"]))
        fallback_s = 0.0
        failed = np.nonzero(~same)[0]
        if len(failed):
            f = fire_batch(samos_amd, sc, [states[i] for i in failed])
            f.fire_setup(dt=0.01, alpha=0.1, dt_max=0.1, f_tol=0.0, stop_fmax=1e-10)
            f.sync()
            t1 = time.perf_counter()
            f.fire_run(max_iter, check=100)
            f.sync()
            fallback_s = time.perf_counter() - t1
            f.close()
        rel = (fs["energy"] - fs_plain["energy"]) / fs_plain["energy"]
        total = secs + fallback_s
        print("bench: warm start x%.2f: %.3f s warm (inference + rebuild %.1f ms) + %.3f s fallback for %d/%d, "
              "same minimum %d/%d" % (scale, secs, t_inf * 1e3, fallback_s, len(failed), len(same), int(same.sum()),
                                     len(same)), file=sys.stderr, flush=True)
        out["scale_%g" % scale] = {
            "time_to_converge_s": total,
            "warm_run_s": secs,
            "inference_and_rebuild_ms": t_inf * 1e3,
            "fallback_s": fallback_s,
            "fallback_configs": int(len(failed)),
            "iterations": iters,
            "iterations_per_replica": {"median": float(np.median(fs["iters"])), "max": float(fs["iters"].max())},
            "speedup_vs_plain": secs_plain / total,
            "speedup_vs_plain_without_fallback": secs_plain / secs,
            "same_minimum_fraction": float(np.mean(same)),
            "energy_rel_diff_vs_plain": {"median": float(np.median(rel)), "mean": float(np.mean(rel)),
                                         "lower_fraction": float(np.mean(rel < -1e-12))},
        }
    mp.close()
    return out


def bench_fire(samos_amd, sc, nrep=512, n=4096, max_iter=400000, weights=WARM_WEIGHTS, cpu=True):
    """Plain batched FIRE to max|F| < 1e-10 on every replica, then (when trained
    weights exist) the same configurations warm-started by the network."""
    states = sc.plane_replicas([n] * nrep, seed=5)
    e = fire_batch(samos_amd, sc, states)
    e.fire_setup(dt=0.01, alpha=0.1, dt_max=0.1, f_tol=0.0, stop_fmax=1e-10)
    e.sync()
    t0 = time.perf_counter()
    rem, iters = e.fire_run(max_iter, check=100)
    e.sync()
    secs = time.perf_counter() - t0
    fs = e.fire_state()
    particle_iters = float(np.sum(fs["iters"]) * n)
    print("bench: FIRE %d x %d: %d iterations, %.3f s, %d not converged" % (nrep, n, iters, secs, rem),
          file=sys.stderr, flush=True)
    # force-pass kernel timing on a fresh copy of the same start
    e.close()
    e = fire_batch(samos_amd, sc, states)
    e.fire_setup(dt=0.01, alpha=0.1, dt_max=0.1, f_tol=0.0, stop_fmax=1e-10)
    e.profile(True)
    e.fire_run(200, check=200)
    ms, launches = e.profile_read("step")
    e.profile(False)
    e.close()
    avg_s = ms / 1e3 / max(launches, 1)
    achieved = FIRE_BYTES_PER_PARTICLE_ITER * nrep * n / avg_s / 1e9
    # Control for any warm start: FIRE restarted (v = 0, fresh dt and alpha)
    # from the exact positions of plain FIRE's own path after one iteration.
    e = fire_batch(samos_amd, sc, states)
    e.fire_setup(dt=0.01, alpha=0.1, dt_max=0.1, f_tol=0.0, stop_fmax=1e-10)
    e.fire_run(1, check=1)
    zeros = np.zeros(nrep * n)
    e.set("vx", zeros)
    e.set("vy", zeros)
    e.fire_setup(dt=0.01, alpha=0.1, dt_max=0.1, f_tol=0.0, stop_fmax=1e-10)
    e.fire_run(max_iter, check=100)
    fr = e.fire_state()
    e.close()
    restart_same = float(np.mean((fr["stopped"] == 1) & (np.abs(fr["energy"] - fs["energy"]) <= 1e-12 * np.abs(fs["energy"]))))
    print("bench: FIRE restarted from its own positions after one iteration: same minimum %.3f" % restart_same,
          file=sys.stderr, flush=True)
    warm = None
    if weights and os.path.exists(weights):
        warm = warm_start_ab(samos_amd, sc, states, fs, secs, max_iter, weights)
    return {
        "warm_start": warm,
        "restart_control": {
            "what": "FIRE restarted (v=0, dt0, alpha0) from plain FIRE's own positions after 1 iteration",
            "same_minimum_fraction": restart_same,
            "iterations_per_replica_median": float(np.median(fr["iters"]) + 1),
        },
        "workload": "fire_polydisperse: %d replicas x N=%d, radii U[0.8, 1.2], periodic, packing 0.88; soft k=1 "
                    "use_particle_radii; FIRE dt0 0.01 dt_max 0.1 alpha 0.1, stop at max|F| < 1e-10" % (nrep, n),
        "time_to_converge_s": secs,
        "iterations": iters,
        "not_converged": rem,
        "iterations_per_replica": {"min": float(fs["iters"].min()), "median": float(np.median(fs["iters"])),
                                   "max": float(fs["iters"].max())},
        "particle_iterations_per_s": particle_iters / secs,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "kernel": "k_ens_fire_b",
                     "avg_launch_us": avg_s * 1e6, "launches_timed": launches,
                     "alg_bytes_per_launch": FIRE_BYTES_PER_PARTICLE_ITER * nrep * n},
        "cpu_baseline": cpu_baseline_fire(n=n) if cpu else None,
    }


def profile_run(samos_amd, sc, name):
    """Workloads for rocprofv3, launched step by step (no graph replay, so every
    kernel appears in a kernel trace):
      ensemble  300 Brownian steps of the active-sphere ensemble (skin 1.0), then
                300 FIRE iterations of 128 x 4096 polydisperse disks
      mpnn      64 training pairs (FIRE, 200 iterations), 5 + 40 Adam steps on
                32-configuration minibatches, one prediction on 64 configurations"""
    if name == "ensemble":
        states, p = sc.c4_ensemble_workload()
        e = sc.load_ensemble(samos_amd.Ensemble, states, "sphere")
        e.set_params(**p)
        e.nlist(2.4, 1.0)
        e.run_brownian(299, samos_amd.RUN_NO_GRAPH)
        e.sync()
        print(json.dumps({"ensemble": e.nlist_info()}), flush=True)
        e.close()
        f = fire_batch(samos_amd, sc, sc.plane_replicas([4096] * 128, seed=5))
        f.fire_setup(dt=0.01, alpha=0.1, dt_max=0.1, f_tol=0.0, stop_fmax=1e-10)
        print(json.dumps({"fire": f.fire_run(300, check=2)}), flush=True)
        f.close()
    elif name == "mpnn":
        mp = samos_amd.Mpnn(seed=3)
        e = fire_batch(samos_amd, sc, sc.plane_replicas([4096] * 64, seed=4242))
        e.fire_setup(dt=0.01, alpha=0.1, dt_max=0.1, f_tol=0.0, stop_fmax=1e-10)
        mp.dataset_reserve(64, 4096)
        mp.dataset_begin(e)
        e.fire_run(200, check=2)
        mp.dataset_end(e)
        work = fire_batch(samos_amd, sc, sc.plane_replicas([4096] * 32, seed=4243))
        mp.train(work, 5, lr=1e-3)
        mp.profile(True)
        mp.train(work, 40, lr=1e-3)
        ms, fl, cnt = mp.profile_read()
        mp.profile(False)
        mp.profile(True)
        mp.predict(e)
        ims, ifl, icnt = mp.profile_read()
        mp.profile(False)
        print(json.dumps({"training_products": {"launches": cnt, "ms": ms, "flop": fl, "tflops": fl / ms / 1e9},
                          "inference_products_64_configs": {"launches": icnt, "ms": ims, "flop": ifl,
                                                            "tflops": ifl / ims / 1e9}}), flush=True)
    else:
        raise SystemExit("unknown profile run " + name)


def traffic_from_profiles():
    """HBM bytes per launch of the fused step kernel from the committed
    rocprofv3 --pmc summary (FETCH_SIZE + WRITE_SIZE, copy-kernel calibrated)."""
    p = os.path.join(ROOT, "profiles", "fused_baoab_lj_N1048576_traffic.json")
    try:
        return json.load(open(p))["bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def bench_tiled_line(samos_amd, sc, args):
    """The same workload on the cell-tiled engine (state in cell order, LDS-staged
    halos, lists of 16-bit stage indices): whole-step rate and its step kernel."""
    s, st = build_system(samos_amd, sc, args.nside, args.pad, args.rebuild, args.batch)
    s.set_option("tile_cells", args.tile_cells)
    s.set_option("chunk_slots", args.chunk_slots)
    n = st["n"]
    flags = samos_amd.RUN_TILES
    timed_run(s, max(args.warmup, 20), flags)
    secs = timed_run(s, args.steps, flags)
    s.profile(True)
    timed_run(s, min(max(args.steps, 200), 500), flags)
    ms, launches = s.profile_read("fused")
    s.profile(False)
    avg_s = ms / 1e3 / max(launches, 1)
    achieved = ALG_BYTES_PER_PARTICLE_STEP * n / avg_s / 1e9
    t = s.tile_info()
    s.close()
    print("bench: tiled engine %.4e particle-steps/s, step kernel %.1f us" % (n * args.steps / secs, avg_s * 1e6),
          file=sys.stderr, flush=True)
    return {"value": n * args.steps / secs, "unit": "particle-steps/s", "steps": args.steps,
            "ms_per_step": secs * 1e3 / args.steps,
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": achieved / HBM_PEAK_GBS, "kernel": "k_tile_step<LJ>", "avg_launch_us": avg_s * 1e6,
                         "launches_timed": launches},
            "cells_per_tile": t["tile_cells"], "list_capacity": t["list_capacity"], "max_neighbours": t["max_neighbours"],
            "max_halo": t["max_halo"]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--nside", type=int, default=NSIDE)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=60)
    ap.add_argument("--variants", action="store_true", help="also time the FP32-force and faithful schedules")
    # skin: the reference's default padding (src/system/defaults.hpp:37); the
    # fastest of 0.3-0.8 in a sweep on MI355X (0.4-0.6 within 1%)
    ap.add_argument("--pad", type=float, default=0.5, help="neighbour-list skin")
    ap.add_argument("--rebuild", choices=["coop", "stages"], default="stages")
    ap.add_argument("--batch", type=int, choices=[4, 8], default=4)
    ap.add_argument("--no-graph", action="store_true",
                    help="launch the c1 and c2 runs step by step (for kernel-trace profiling)")
    ap.add_argument("--chunk-slots", type=int, default=32,
                    help="fused steps per captured chunk (each chunk ends in one conditional rebuild)")
    ap.add_argument("--tiles", action="store_true",
                    help="headline line on the cell-tiled engine instead of the neighbour-list kernels")
    ap.add_argument("--no-tiled-line", action="store_true", help="skip the cell-tiled engine's extra line")
    ap.add_argument("--tile-cells", type=int, default=0, help="cells per tile (0: automatic)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run device-to-device copies of this size (HBM counter calibration) and exit")
    ap.add_argument("--parts", default="main,ensemble,fire,c1,c2,tiles_brownian",
                    help="comma list: main (the N=1M metric), ensemble (batched active spheres), "
                         "fire (batched FIRE relaxation and the warm start), c1 (polar disks, N=4096), "
                         "c2 (nematic sphere, N=100000), tiles_brownian (polar disks, N=1M, cell-tiled "
                         "Brownian step against the list kernels)")
    ap.add_argument("--profile-run", choices=["ensemble", "mpnn"], default=None,
                    help="only run a fixed workload for rocprofv3 and exit")
    ap.add_argument("--ensemble-steps", type=int, default=5000)
    ap.add_argument("--ensemble-replicas", type=int, default=256)
    ap.add_argument("--ensemble-pad", type=float, default=None, help="fixed skin instead of the autotuner")
    ap.add_argument("--plain-steps", type=int, default=200,
                    help="steps per replica when timing the replicas as separate runs")
    ap.add_argument("--fire-replicas", type=int, default=512)
    ap.add_argument("--warm-weights", default=WARM_WEIGHTS, help="trained warm-start network (.npz)")
    args = ap.parse_args()
    parts = set(args.parts.split(","))
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: each particle system runs on one GPU, and there is no multi-GPU "
              "path to time", file=sys.stderr)
        sys.exit(2)

    import samos_amd
    from tests import scenarios as sc

    if args.profile_run:
        profile_run(samos_amd, sc, args.profile_run)
        return

    extras = {}
    if "ensemble" in parts:
        extras["ensemble"] = bench_ensemble(samos_amd, sc, args.ensemble_steps, args.plain_steps,
                                            nrep=args.ensemble_replicas, pad=args.ensemble_pad, cpu=not args.no_cpu)
    if "fire" in parts:
        extras["fire"] = bench_fire(samos_amd, sc, nrep=args.fire_replicas, weights=args.warm_weights,
                                    cpu=not args.no_cpu)
    ng = samos_amd.RUN_NO_GRAPH if args.no_graph else 0
    if "c1" in parts:
        extras["polar_disks_plane_n4096"] = bench_single(samos_amd, sc, "c1", 10000, 0 if args.no_cpu else 25000, ng)
    if "c2" in parts:
        extras["nematic_sphere_n100000"] = bench_single(samos_amd, sc, "c2", 5000, 0 if args.no_cpu else 280, ng)
    if "tiles_brownian" in parts:
        extras["polar_disks_plane_n1048576_tiled"] = bench_tiled_brownian(samos_amd, sc, 500)
    if "main" not in parts:
        print(json.dumps(extras))
        return

    s, st = build_system(samos_amd, sc, args.nside, args.pad, args.rebuild, args.batch)
    s.set_option("tile_cells", args.tile_cells)
    s.set_option("chunk_slots", args.chunk_slots)
    n = st["n"]
    print("bench: N=%d box %.3f, warmup %d, steps %d" % (n, st["box"][0], args.warmup, args.steps),
          file=sys.stderr, flush=True)
    run_flags = samos_amd.RUN_TILES if args.tiles else 0
    if args.warmup > 0:
        timed_run(s, args.warmup, run_flags)
    secs = timed_run(s, args.steps, run_flags)
    value = n * args.steps / secs
    print("bench: timed %.3f s -> %.4e particle-steps/s" % (secs, value), file=sys.stderr, flush=True)

    # Dominant kernel: the fused LJ force + BAOAB step.  HIP events bracket each
    # launch on the context's stream (step-by-step launches, same kernels), over
    # at least 200 launches so that a short timed run still gets a steady
    # average (each run's first launch is the heavier phase-0 step).
    s.profile(True)
    prof_steps = min(max(args.steps, 200), 500)
    timed_run(s, prof_steps, run_flags)
    ms, launches = s.profile_read("fused")
    rb_ms, rb_launches = s.profile_read("rebuild")
    s.profile(False)
    avg_s = ms / 1e3 / max(launches, 1)
    alg_bytes = ALG_BYTES_PER_PARTICLE_STEP * n
    achieved = alg_bytes / avg_s / 1e9
    info = s.nlist_info()
    tiles = s.tile_info()

    extra = {}
    if not args.tiles:
        # the FP32-force variant of the same workload (pair arithmetic in FP32,
        # sums, positions and the integrator in FP64)
        s32, _ = build_system(samos_amd, sc, args.nside, args.pad, args.rebuild, args.batch)
        s32.set_option("chunk_slots", args.chunk_slots)
        timed_run(s32, args.warmup, samos_amd.RUN_FP32_FORCE)
        t32 = timed_run(s32, args.steps, samos_amd.RUN_FP32_FORCE)
        extra["fp32_force_value"] = n * args.steps / t32
        s32.close()
    if args.variants:
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
            "chunk_slots": args.chunk_slots,
            "schedule": "fused: one force evaluation per step (the loop-head re-evaluation of identical "
                        "positions is not repeated); captured chunks of fused steps, each ending in one "
                        "device-conditional rebuild",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic_from_profiles(),
            "kernel": "k_tile_step<LJ>" if args.tiles else "k_baoab_fused<LJ>",
            "avg_launch_us": avg_s * 1e6,
            "launches_timed": launches,
            "alg_bytes_per_launch": alg_bytes,
        },
        "nlist": {"builds": info["builds"], "stride": info["stride"], "rebuild_ms_total": rb_ms,
                  "rebuild_stage_launches": rb_launches},
        "tiles": None if not args.tiles else {"cells_per_tile": tiles["tile_cells"],
                                              "list_capacity": tiles["list_capacity"],
                                              "max_neighbours": tiles["max_neighbours"],
                                              "max_halo": tiles["max_halo"], "rebuilds": tiles["rebuilds"]},
    }
    out.update(extra)
    out.update(extras)
    s.close()
    if not args.tiles and not args.no_tiled_line:
        out["tiled_engine"] = bench_tiled_line(samos_amd, sc, args)
    if not args.no_cpu:
        cpu_side = min(args.nside, NSIDE)
        out["cpu_baseline"] = cpu_baseline(cpu_side, args.cpu_steps, args.pad)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
