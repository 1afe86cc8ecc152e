"""Train the atom-pair initial density (mmd.pairguess) and compare it with
the SAD, monomer-network (mmd.learned) and core starts.

    python make_pair_guess.py --train 8000 --heldout 2000 --out DIR

1. Training set: the clusters of make_learned_guess.py (rng seed 1),
   converged 6-31G direct RHF densities from the batched device driver
   (SAD start), reduced to the bond-frame invariants of every atom pair
   within the cutoff.
2. One network per ordered element pair (32 radial features -> 3 x 128
   SiLU -> invariants), trained with libmmd's HIP kernels (Adam,
   mean-squared error on normalised targets).
3. Held-out set (rng seed 2): every geometry solved from the core, SAD,
   monomer-network and atom-pair starts with the batched device driver;
   iterations (overall and per cluster size) and energies against the core
   start.

Writes DIR/pair_guess_631g.npz (the model) and DIR/pair_guess_631g.json (the
training and evaluation record).  Run it on a machine with the GPU.
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, HERE)

import numpy as np  # noqa: E402


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--train", type=int, default=8000)
    ap.add_argument("--heldout", type=int, default=2000)
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--conver", type=float, nargs="+", default=[1e-8, 1e-10],
                    help="density convergence levels of the evaluation runs")
    ap.add_argument("--out", default=HERE)
    args = ap.parse_args()
    from make_learned_guess import geometries
    from mmd.batch import _molecule_arrays, run_batch_device
    from mmd import pairguess as pg

    rec = {"train_geometries": args.train, "heldout_geometries": args.heldout, "basis": "6-31g",
           "features": "%d Gaussian radial functions on [%.1f, %.1f] bohr, cosine cutoff %.1f bohr"
                       % (pg.NRBF, pg.R_MIN, pg.R_MAX, pg.R_CUT),
           "network": "32 -> 128 -> 128 -> 128 (SiLU) -> invariants, one per ordered element pair"}
    t0 = time.perf_counter()
    train = geometries(args.train, 1)
    res = run_batch_device(train, "6-31g", guess="sad", keep_density=True, conver=1e-8)
    layouts = pg.element_layouts("6-31g")
    data = {}
    for g, r in zip(train, res):
        if not r["converged"]:
            continue
        a = _molecule_arrays(g, "6-31g")
        for k, (F, Y) in pg.training_pairs(a[0], a[1], r["P"], layouts).items():
            data.setdefault(k, ([], []))
            data[k][0].append(F)
            data[k][1].append(Y)
    data = {k: (np.concatenate(v[0]), np.concatenate(v[1])) for k, v in data.items()}
    rec["dataset_seconds"] = time.perf_counter() - t0
    rec["train_converged"] = int(sum(r["converged"] for r in res))
    rec["train_pairs"] = {"%d-%d" % k: int(v[0].shape[0]) for k, v in data.items()}
    print("dataset: %s in %.1f s" % (rec["train_pairs"], rec["dataset_seconds"]), flush=True)

    t1 = time.perf_counter()
    model, losses = pg.train(data, layouts, epochs=args.epochs)
    rec["train_seconds"] = time.perf_counter() - t1
    rec["loss_first_last"] = {"%d-%d" % k: [float(l[0]), float(l[-1])] for k, l in losses.items()}
    print("trained in %.1f s: %s" % (rec["train_seconds"], rec["loss_first_last"]), flush=True)
    path = os.path.join(args.out, "pair_guess_631g.npz")
    model.save(path)

    heldout = geometries(args.heldout, 2)
    sizes = np.array([len([ln for ln in g.split("\n")[1:] if ln.strip()]) // 3 for g in heldout])
    run_batch_device(heldout[:128], "6-31g", guess="learned")   # contexts, atomic densities, networks: once
    run_batch_device(heldout[:128], "6-31g", guess="pair", guess_model=model)
    rec["evaluation"] = {}
    for conver in args.conver:
        runs, ev = {}, {}
        for guess in ("core", "sad", "learned", "pair"):
            t3 = time.perf_counter()
            runs[guess] = run_batch_device(heldout, "6-31g", guess=guess, conver=conver,
                                           guess_model=model if guess == "pair" else None)
            runs[guess + "_s"] = time.perf_counter() - t3
        e_core = np.array([r["energy"] for r in runs["core"]])
        for g in ("core", "sad", "learned", "pair"):
            its = np.array([r["iterations"] for r in runs[g]])
            e = np.array([r["energy"] for r in runs[g]])
            ev[g] = {"mean_iterations": float(its.mean()), "median_iterations": float(np.median(its)),
                     "max_iterations": int(its.max()),
                     "mean_iterations_by_size": {str(n): float(its[sizes == n].mean()) for n in sorted(set(sizes))},
                     "all_converged": bool(all(r["converged"] for r in runs[g])), "seconds": runs[g + "_s"]}
            if g != "core":
                d = np.abs(e - e_core)
                ev[g]["max_abs_dE_vs_core"] = float(d.max())
        rec["evaluation"]["conver_%g" % conver] = ev
        print("conver %g: %s" % (conver, {g: ev[g]["mean_iterations"] for g in ev}), flush=True)
    rec["evaluation_driver"] = "run_batch_device (FP64, groups of 64)"
    with open(os.path.join(args.out, "pair_guess_631g.json"), "w") as f:
        json.dump(rec, f, indent=1)
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
