"""Train the learned initial density (mmd.learned) and evaluate it.

    python make_learned_guess.py --train 20000 --heldout 2000 --out DIR

1. Training set: (H2O)n, n uniform in 2..6 (rng seed 1), perturbed monomers,
   O in a (2 + 1.4 n) A cube; converged 6-31G direct RHF densities from the
   GPU batch driver (worker processes), reduced to per-monomer features and
   local-frame targets.
2. The perceptron (15 -> 64 -> 64 -> 91, tanh) is trained with libmmd's HIP
   kernels (Adam, mean-squared error on normalised targets).
3. Held-out set (rng seed 2, 2000 geometries): every geometry is solved
   from the core, SAD and learned starts with the batched device driver;
   iterations (overall and per cluster size), wall time and energies against
   the core start are reported.

Writes DIR/learned_guess_631g.npz (the model) and DIR/learned_guess_631g.json
(the training and evaluation record).  Run it on a machine with the GPU; the
process starts its SCF worker processes before it uses the GPU itself.
"""
import argparse
import json
import os
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import numpy as np  # noqa: E402


def geometries(n, seed):
    from mmd.workloads import water_cluster
    rng = np.random.default_rng(seed)
    ns = rng.integers(2, 7, size=n)
    return [water_cluster(int(k), 2.0 + 1.4 * int(k), seed=seed * 1000003 + i, perturb=True)
            for i, k in enumerate(ns)]


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--train", type=int, default=20000)
    ap.add_argument("--heldout", type=int, default=2000)
    ap.add_argument("--evaluate", type=int, default=2000, help="held-out geometries solved three ways")
    ap.add_argument("--conver", type=float, default=1e-8, help="density convergence of the evaluation runs")
    ap.add_argument("--model", default=None, help="model to evaluate (default: DIR/learned_guess_631g.npz)")
    ap.add_argument("--record", default=None, help="evaluation record to write (default: DIR/learned_guess_631g.json)")
    ap.add_argument("--epochs", type=int, default=300)
    ap.add_argument("--procs", type=int, default=8)
    ap.add_argument("--threads", type=int, default=2)
    ap.add_argument("--out", default=HERE)
    ap.add_argument("--evaluate-only", action="store_true",
                    help="skip data generation and training; evaluate the model already in --out")
    args = ap.parse_args()
    from mmd.batch import run_batch_procs
    path = args.model or os.path.join(args.out, "learned_guess_631g.npz")
    rec_path = args.record or os.path.join(args.out, "learned_guess_631g.json")
    rec = {"train_geometries": args.train, "heldout_geometries": args.heldout, "basis": "6-31g"}
    if args.evaluate_only:
        old = os.path.join(args.out, "learned_guess_631g.json")
        if os.path.exists(old):
            with open(old) as f:
                rec = json.load(f)
        evaluate(args, path, rec, rec_path)
        return

    t0 = time.perf_counter()
    train = geometries(args.train, 1)
    res, X, Y = [], [], []
    for c0 in range(0, len(train), 2000):   # rounds of 2000 geometries, with a progress line each
        with tempfile.TemporaryDirectory() as td:
            res += run_batch_procs(train[c0:c0 + 2000], "6-31g", procs=args.procs, threads=args.threads,
                                   pairs_dir=td)
            for f in sorted(os.listdir(td)):
                with np.load(os.path.join(td, f)) as part:
                    X.append(part["X"])
                    Y.append(part["Y"])
        print("training set: %d / %d geometries (%.1f s)" % (len(res), len(train), time.perf_counter() - t0),
              flush=True)
    X, Y = np.concatenate(X), np.concatenate(Y)
    rec["dataset_seconds"] = time.perf_counter() - t0
    rec["train_converged"] = int(sum(r["converged"] for r in res))
    rec["train_monomers"] = int(X.shape[0])
    print("dataset: %d monomers in %.1f s" % (X.shape[0], rec["dataset_seconds"]), flush=True)

    # the GPU is used by this process from here on
    from mmd.learned import Model, train_model
    t2 = time.perf_counter()
    model, losses = train_model(X, Y, hidden=(64, 64), epochs=args.epochs, batch=512, lr=2e-3, seed=0)
    rec["train_seconds"] = time.perf_counter() - t2
    rec["loss_first_last"] = [float(losses[0]), float(losses[-1])]
    print("trained: loss %.4g -> %.4g in %.1f s" % (losses[0], losses[-1], rec["train_seconds"]), flush=True)
    model.save(path)
    evaluate(args, path, rec, rec_path)


def evaluate(args, path, rec, rec_path):
    """Solve the held-out geometries from the core, SAD and learned starts
    with the batched device driver (run_batch_device, FP64, groups of 64);
    iterations, wall time and energies against the core start."""
    from mmd import learned
    learned.DEFAULT_MODEL = path
    from mmd.batch import run_batch_device
    heldout = geometries(args.heldout, 2)[:args.evaluate]
    run_batch_device(heldout[:128], "6-31g", guess="learned")   # contexts, atomic densities, network: once
    runs = {}
    for guess in ("core", "sad", "learned"):
        t3 = time.perf_counter()
        runs[guess] = run_batch_device(heldout, "6-31g", guess=guess, conver=args.conver)
        runs[guess + "_s"] = time.perf_counter() - t3
        print("held-out %s done (%.2f s)" % (guess, runs[guess + "_s"]), flush=True)
    sizes = np.array([len([ln for ln in g.split("\n")[1:] if ln.strip()]) // 3 for g in heldout])
    ev = {}
    e_core = np.array([r["energy"] for r in runs["core"]])
    for g in ("core", "sad", "learned"):
        its = np.array([r["iterations"] for r in runs[g]])
        e = np.array([r["energy"] for r in runs[g]])
        ev[g] = {"mean_iterations": float(its.mean()), "median_iterations": float(np.median(its)),
                 "max_iterations": int(its.max()),
                 "mean_iterations_by_size": {str(n): float(its[sizes == n].mean()) for n in sorted(set(sizes))},
                 "all_converged": bool(all(r["converged"] for r in runs[g])), "seconds": runs[g + "_s"]}
        if g != "core":
            d = np.abs(e - e_core)
            ev[g]["max_abs_dE_vs_core"] = float(d.max())
            ev[g]["within_1e-8_of_core"] = int((d <= 1e-8).sum())
    ev["learned"]["iterations_saved_vs_sad"] = ev["sad"]["mean_iterations"] - ev["learned"]["mean_iterations"]
    rec["heldout_evaluated"] = len(heldout)
    rec["evaluation_driver"] = "run_batch_device (FP64, groups of 64), conver %g" % args.conver
    rec["evaluation"] = ev
    with open(rec_path, "w") as f:
        json.dump(rec, f, indent=1)
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
