#!/usr/bin/env python
"""Benchmark: converged direct FP64 RHF of (H2O)8 / cc-pVDZ on one MI355X.

Workload: 8 waters at the README monomer geometry, each rotated uniformly at
random and placed in a 7 A cube with O-O >= 2.7 A (seed 0); cc-pVDZ with
Cartesian d (200 functions); direct SCF with Schwarz threshold 1e-12, all
FP64 on the GPU.  One step = one `Molecule(geometry, 'cc-pvdz').RHF(direct=True)`
call from the geometry string to a converged energy (one-electron integrals,
S^-1/2 and the Schwarz pass included).

Printed (one JSON line): converged SCF/s (value), median Fock-build ms,
quartets/s, the roofline of the dominant Fock-build class kernel (model
FLOPs / HIP-event time) and a CPU baseline from the oracle on a bounded
sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "mcmurchie-davidson_amd")
for _p in (ROOT, PKG):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = "converged SCF/sec at |ΔE|≤1 µEh vs plain FP64 GPU; Fock-build ms; quartets/s"
FP64_PEAK_TFLOPS = 78.6


def ncart(l):
    return (l + 1) * (l + 2) // 2


def class_flops(la, lb, lc, ld):
    """Model FLOPs per primitive shell quartet (la lb|lc ld):
    2 H_ab H_cd C_cd + 2 C_ab H_ab C_cd + 3 N_R(L) + 40."""
    H = lambda L: (L + 1) * (L + 2) * (L + 3) // 6
    NR = lambda L: (L + 1) * (L + 2) * (L + 3) * (L + 4) // 24
    Lab, Lcd = la + lb, lc + ld
    Cab, Ccd = ncart(la) * ncart(lb), ncart(lc) * ncart(ld)
    return 2 * H(Lab) * H(Lcd) * Ccd + 2 * Cab * H(Lab) * Ccd + 3 * NR(Lab + Lcd) + 40


def run_scf(geom, basis, profile=False):
    from mmd.molecule import Molecule
    mol = Molecule(geom, basis)
    if profile:
        mol.native.set_profiling(True)
        mol._profile_stats = []
        orig = mol.native.fock_delta

        def wrapped(*a, **k):
            r = orig(*a, **k)
            mol._profile_stats.append(mol.native.fock_stats())
            return r
        mol.native.fock_delta = wrapped
    mol.RHF(doPrint=False, direct=True, conver=1e-8, acc2e=1e-12)
    return mol


def roofline_from_stats(stats):
    """Per class: (model flops, kernel ms, launches, shell quartets, primitive
    quartets) summed over builds."""
    per = {}
    for lc, nq, npr, ms in stats:
        for c in range(len(nq)):
            key = tuple(int(x) for x in lc[c])
            if nq[c] == 0 and ms[c] == 0:
                continue
            la, lb, lcc, ld = key
            fl = npr[c] * class_flops(*key) + 12.0 * nq[c] * ncart(la) * ncart(lb) * ncart(lcc) * ncart(ld)
            f, t, n, sq, sp = per.get(key, (0.0, 0.0, 0, 0, 0))
            per[key] = (f + fl, t + ms[c], n + 1, sq + int(nq[c]), sp + int(npr[c]))
    return per


def cpu_baseline(geom, basis, n_builds, seconds):
    """Oracle (C restatement of the reference's per-quartet ERI, 1 core) timed
    on a random sample of canonical basis-function quartets, scaled to one
    direct SCF (Schwarz pass + n_builds unscreened Fock builds)."""
    import numpy as np
    import oracle
    om = oracle.OracleMolecule(geom, os.path.join(PKG, "mmd", "basis", basis + ".gbs"))
    N = om.nbasis
    npair = N * (N + 1) // 2
    nquart = npair * (npair + 1) // 2
    rng = np.random.default_rng(1234)
    done, t_used, n = 0, 0.0, 2000
    while t_used < seconds:
        ij = rng.integers(0, npair, size=n)
        kl = (rng.random(n) * (ij + 1)).astype(np.int64)
        i = ((np.sqrt(8 * ij + 1) - 1) // 2).astype(np.int64)
        j = ij - i * (i + 1) // 2
        k = ((np.sqrt(8 * kl + 1) - 1) // 2).astype(np.int64)
        l = kl - k * (k + 1) // 2
        idx = np.stack([i, j, k, l], axis=1)
        t0 = time.perf_counter()
        om.eri_list(idx)
        t_used += time.perf_counter() - t0
        done += n
        n = min(200000, n * 2)
    per_q = t_used / done
    t_scf = per_q * (npair + n_builds * nquart)
    return {"value": 1.0 / t_scf, "unit": "SCF/s", "cores": 1, "kind": "port",
            "sample": "oracle ERI() on %d random canonical quartets of the same (H2O)8/cc-pVDZ basis "
                      "(%.1f s, %.1f us/quartet), scaled to %d Fock builds x %d quartets + Schwarz "
                      "(reference loop structure, unscreened, 1 core; extrapolated)" % (
                          done, t_used, per_q * 1e6, n_builds, nquart),
            "fock_build_s": per_q * nquart}


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-profile", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    from mmd.workloads import water_cluster
    from mmd import _native
    if _native.device_count() < 1:
        print("no HIP device visible", file=sys.stderr)
        sys.exit(1)

    geom = water_cluster(8, 7.0, seed=0)
    basis = "cc-pvdz"
    for _ in range(args.warmup):
        run_scf(geom, basis)
    print("warmup done", file=sys.stderr, flush=True)

    mols = []
    t0 = time.perf_counter()
    for s in range(args.steps):
        mols.append(run_scf(geom, basis))
        print("step %d done" % s, file=sys.stderr, flush=True)
    elapsed = time.perf_counter() - t0
    scf_per_s = args.steps / elapsed

    m0 = mols[-1]
    N = m0.nbasis
    npair = N * (N + 1) // 2
    nquart = npair * (npair + 1) // 2
    fock_ms = [ms for m in mols for ms in m.fock_ms[1:]] or [ms for m in mols for ms in m.fock_ms]
    fock_med = float(np.median(fock_ms))
    energies = [m.energy.real for m in mols]
    out = {
        "metric": METRIC, "value": scf_per_s, "unit": "SCF/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * elapsed / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "(H2O)8 cc-pVDZ direct FP64 RHF", "n_basis": N, "seed": 0, "box_A": 7.0,
                   "min_OO_A": 2.7, "acc2e": 1e-12, "conver": 1e-8},
        "fock_build_ms": fock_med, "quartets_per_s": nquart / (fock_med * 1e-3),
        "unique_bf_quartets": nquart, "scf_iterations": int(m0.iterations), "energy": energies[-1],
        "delta_energy": float(abs(m0.delta_energy.real)),
    }

    if not args.no_profile:
        pm = run_scf(geom, basis, profile=True)
        per = roofline_from_stats(pm._profile_stats)
        tot_f = sum(v[0] for v in per.values())
        tot_t = sum(v[1] for v in per.values())
        dom = max(per, key=lambda k: per[k][1])
        f, t, n = per[dom][:3]
        achieved = f / (t * 1e-3) / 1e12
        out["roofline"] = {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                           "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None,
                           "kernel": "eri_fock_class_kernel<%d,%d,%d,%d>" % dom,
                           "kernel_avg_ms": t / n, "launches": n,
                           "all_classes_achieved": tot_f / (tot_t * 1e-3) / 1e12,
                           "all_classes_ms_per_build": tot_t / max(1, len(pm._profile_stats))}
        nb = max(1, len(pm._profile_stats))
        out["class_breakdown"] = {
            "(%d%d|%d%d)" % k: {"ms_per_build": v[1] / nb, "tflops": v[0] / (v[1] * 1e-3) / 1e12 if v[1] else 0,
                                "shell_quartets_per_build": v[3] // nb, "prim_quartets_per_build": v[4] // nb}
            for k, v in sorted(per.items(), key=lambda kv: -kv[1][1])}
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(geom, basis, len(m0.fock_ms), args.cpu_seconds)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
