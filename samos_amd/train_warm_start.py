"""Training of the FIRE warm-start network on the GPU.

Generates (start, relaxed) pairs with batched plain FIRE (2-D polydisperse
soft disks, radii uniform in [0.8, 1.2], packing fraction 0.88, periodic box,
soft k = 1 with particle radii, FIRE dt0 0.01, dt_max 0.1, alpha 0.1, stopped
at max|F| < 1e-10), keeps them in device memory, trains the network with Adam
on minibatches of configurations, and writes the parameters to an .npz file
(arrays `params` float32 and `meta` JSON text).

    python -m samos_amd.train_warm_start --configs 20480 --out weights.npz

Configurations for training use seeds 100000 + batch index; the benchmark's
test configurations use other seeds.
"""
import argparse
import json
import math
import sys
import time

import numpy as np

import samos_amd
from samos_amd import workloads as wl

RCUT, PAD = 2.4, 0.3


def fire_ensemble(states):
    e = wl.load_ensemble(samos_amd.Ensemble, states, "plane")
    e.set_params(k=1.0, a=2.0, use_radii=1.0, dt=0.01)
    e.nlist(RCUT, PAD)
    return e


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--configs", type=int, default=20480, help="training pairs to generate")
    ap.add_argument("--per", type=int, default=4096, help="particles per configuration")
    ap.add_argument("--gen-batch", type=int, default=512, help="configurations relaxed together")
    ap.add_argument("--train-batch", type=int, default=32, help="configurations per Adam step")
    ap.add_argument("--val", type=int, default=480, help="held-out configurations")
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--train-seconds", type=float, default=600.0, help="wall-clock budget of the training loop")
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--lr-min", type=float, default=5e-5)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--out", required=True)
    args = ap.parse_args()

    t_start = time.time()
    mp = samos_amd.Mpnn(seed=args.seed)
    mp.dataset_reserve(args.configs, args.per)
    nb = args.configs // args.gen_batch
    iters = []
    for b in range(nb):
        t0 = time.time()
        states = wl.plane_replicas([args.per] * args.gen_batch, seed=100000 + b)
        e = fire_ensemble(states)
        e.fire_setup(dt=0.01, alpha=0.1, dt_max=0.1, f_tol=0.0, stop_fmax=1e-10)
        mp.dataset_begin(e)
        rem, it = e.fire_run(400000, check=200)
        if rem:
            raise SystemExit(f"batch {b}: {rem} configurations did not reach max|F| < 1e-10")
        mp.dataset_end(e)
        iters.append(e.fire_state()["iters"])
        e.close()
        print(f"pairs: batch {b + 1}/{nb}, {mp.dataset_size()} configurations, {it} iterations, "
              f"{time.time() - t0:.1f} s", flush=True)
    n_train = mp.dataset_size() - args.val
    work = fire_ensemble(wl.plane_replicas([args.per] * args.train_batch, seed=99))
    steps_per_epoch = max(1, n_train // args.train_batch)
    history = []
    t_train = time.time()
    for ep in range(args.epochs):
        lr = args.lr_min + 0.5 * (args.lr - args.lr_min) * (1.0 + math.cos(math.pi * ep / max(args.epochs, 1)))
        t0 = time.time()
        losses = mp.train(work, steps_per_epoch, lr=lr, seed=args.seed * 1000 + ep, first=0, count=n_train)
        val = mp.evaluate(work, n_train, args.val - args.val % args.train_batch) if args.val >= args.train_batch \
            else float("nan")
        history.append({"epoch": ep, "lr": lr, "train_loss": float(np.mean(losses)), "val_loss": val})
        print(f"train: epoch {ep + 1}/{args.epochs} lr {lr:.2e} loss {np.mean(losses):.5f} val {val:.5f} "
              f"({time.time() - t0:.1f} s)", flush=True)
        if time.time() - t_train > args.train_seconds:
            print("train: wall-clock budget reached", flush=True)
            break
    # reference point: the loss of predicting no displacement at all
    trained = mp.get_params()
    mp.set_params(np.zeros_like(trained))
    val_zero = mp.evaluate(work, n_train, args.val - args.val % args.train_batch) if args.val >= args.train_batch \
        else float("nan")
    mp.set_params(trained)
    meta = {
        "configs": mp.dataset_size(), "per_config": args.per, "train_configs": n_train, "val_configs": args.val,
        "train_batch": args.train_batch, "epochs_run": len(history), "history": history,
        "val_loss_zero_prediction": val_zero,
        "fire_iterations_median": float(np.median(np.concatenate(iters))),
        "seconds": time.time() - t_start,
        "workload": "2-D polydisperse soft disks, radii U[0.8, 1.2], packing 0.88, periodic; soft k=1 "
                    "use_particle_radii; FIRE dt0 0.01 dt_max 0.1 alpha 0.1 to max|F| < 1e-10; nlist 2.4 + 0.3",
    }
    np.savez(args.out, params=mp.get_params(), meta=np.array(json.dumps(meta)))
    print(json.dumps({k: v for k, v in meta.items() if k != "history"}), flush=True)


if __name__ == "__main__":
    sys.exit(main())
