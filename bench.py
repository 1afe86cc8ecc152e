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

# Eight hardware queues for this process (HIP's default is four), so that
# the Fock build's eight class streams run side by side; set before HIP
# initialises.  The value is recorded in the printed config.
HW_QUEUES = "8"
os.environ["GPU_MAX_HW_QUEUES"] = HW_QUEUES

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "mcmurchie-davidson_amd")
for _p in (ROOT, PKG):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = "converged SCF/sec at |ΔE|≤1 µEh vs plain FP64 GPU; Fock-build ms; quartets/s"
FP64_PEAK_TFLOPS = 78.6
# HBM bytes per launch of the class kernels on this workload, from rocprofv3
# --pmc FETCH_SIZE and WRITE_SIZE runs of this benchmark, scaled by the same
# counters read for a copy of known size (recorded with the command lines in
# profiles/); a class absent here reports null.
TRAFFIC = {
    (0, 0, 0, 0): 2.242e+07, (1, 0, 0, 0): 4.229e+07, (1, 0, 1, 0): 4.499e+07, (1, 1, 0, 0): 2.063e+07,
    (1, 1, 1, 0): 5.513e+07, (1, 1, 1, 1): 7.862e+06, (2, 0, 0, 0): 1.916e+07, (2, 0, 1, 0): 3.828e+07,
    (2, 0, 1, 1): 1.154e+07, (2, 0, 2, 0): 5.946e+06, (2, 1, 0, 0): 1.251e+07, (2, 1, 1, 0): 1.652e+07,
    (2, 1, 1, 1): 7.645e+06, (2, 1, 2, 0): 5.551e+06, (2, 1, 2, 1): 6.124e+06, (2, 2, 0, 0): 3.453e+06,
    (2, 2, 1, 0): 5.270e+06, (2, 2, 1, 1): 2.754e+06, (2, 2, 2, 0): 2.329e+06, (2, 2, 2, 1): 5.559e+06,
}


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


def run_scf(geom, basis, profile=False, serial=False):
    from mmd.molecule import Molecule
    mol = Molecule(geom, basis)
    if profile:
        ctx = mol.native
        ctx.set_profiling(True, serial=serial)
        mol._profile_stats = []
        orig_delta, orig_step = ctx.fock_delta, ctx.scf_step

        def delta(*a, **k):
            r = orig_delta(*a, **k)
            mol._profile_stats.append(ctx.fock_stats())
            return r

        def step(tol, build, diis, FO):   # the GPU-resident loop builds inside scf_step
            r = orig_step(tol, build, diis, FO)
            if build:
                mol._profile_stats.append(ctx.fock_stats())
            return r
        ctx.fock_delta, ctx.scf_step = delta, step
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


def bench_tier(args):
    """(H2O)16 / cc-pVDZ (400 functions, 9 A cube, seed 0): direct RHF in
    FP64 and with the FP32 tier (quartets with bound < 1e-4 in FP32)."""
    import numpy as np
    from mmd.molecule import Molecule
    from mmd.workloads import water_cluster
    geom = water_cluster(16, 9.0, seed=0)

    def run(fp32_below):
        mol = Molecule(geom, "cc-pvdz")
        ctx = mol.native
        ctx.set_precision(fp32_below)
        tiers = []
        orig_delta, orig_step = ctx.fock_delta, ctx.scf_step

        def delta(*a, **k):
            r = orig_delta(*a, **k)
            tiers.append(ctx.fock_tiers())
            return r

        def step(tol, build, diis, FO):
            r = orig_step(tol, build, diis, FO)
            if build:
                tiers.append(ctx.fock_tiers())
            return r
        ctx.fock_delta, ctx.scf_step = delta, step
        mol.RHF(doPrint=False, direct=True, conver=1e-8, acc2e=1e-12)
        mol._tiers = tiers
        return mol

    for _ in range(args.warmup):
        run(1e-4)
    res = {}
    for name, thr in (("fp64", 0.0), ("tier", 1e-4)):
        t0 = time.perf_counter()
        mols = [run(thr) for _ in range(args.steps)]
        el = time.perf_counter() - t0
        print("%s done" % name, file=sys.stderr, flush=True)
        m = mols[-1]
        nq = sum(a for a, b in m._tiers)
        n32 = sum(b for a, b in m._tiers)
        res[name] = {"scf_per_s": args.steps / el, "ms_per_scf": 1e3 * el / args.steps,
                     "energy": float(m.energy.real), "iterations": int(m.iterations),
                     "fock_build_ms": float(np.median(m.fock_ms[1:])), "fp32_quartet_fraction": n32 / max(1, nq)}
    dE = abs(res["tier"]["energy"] - res["fp64"]["energy"])
    N = Molecule(geom, "cc-pvdz").nbasis
    return {
        "metric": METRIC, "value": res["tier"]["scf_per_s"] if dE <= 1e-6 else 0.0, "unit": "SCF/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": res["tier"]["ms_per_scf"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64+f32", "data": "synthetic",
        "config": {"workload": "(H2O)16 cc-pVDZ direct RHF, FP32 tier below 1e-4", "n_basis": N, "seed": 0,
                   "box_A": 9.0, "min_OO_A": 2.7, "acc2e": 1e-12, "conver": 1e-8, "hw_queues": int(HW_QUEUES)},
        "delta_energy_vs_fp64": dE, "fp64": res["fp64"], "tier": res["tier"],
    }


def bench_batch(args):
    """512 x (H2O)n, n uniform in 2..6, perturbed monomers (seed 1), 6-31G:
    all direct FP64 RHF calls on one GPU through a warm pool of batch worker
    processes (pool shape autotuned on a 64-molecule calibration batch); the
    timed region is the 512-molecule batch."""
    import numpy as np
    from mmd.batch import BatchPool
    from mmd.molecule import Molecule
    from mmd.workloads import water_cluster
    rng = np.random.default_rng(1)
    ns = rng.integers(2, 7, size=512)
    geoms = [water_cluster(int(n), 2.0 + 1.4 * int(n), seed=1000 + i, perturb=True) for i, n in enumerate(ns)]
    # autotune the pool shape on a warm pool and a 64-molecule calibration
    # batch; worker processes start before this process uses the GPU
    rates = {}
    for shape in ((4, 2), (6, 2), (8, 2), (8, 1), (12, 1)):
        with BatchPool(*shape) as pool:
            pool.run(geoms[:2 * shape[0] * shape[1]], "6-31g")        # warm-up: contexts, caches
            t0 = time.perf_counter()
            pool.run(geoms[:64], "6-31g")
            rates[shape] = 64 / (time.perf_counter() - t0)
        print("pool %dx%d: %.1f SCF/s" % (shape[0], shape[1], rates[shape]), file=sys.stderr, flush=True)
    best = max(rates, key=rates.get)
    with BatchPool(*best) as pool:
        pool.run(geoms[:2 * best[0] * best[1]], "6-31g")
        t0 = time.perf_counter()
        res = pool.run(geoms, "6-31g")
        el = time.perf_counter() - t0
        print("batch FP64 done", file=sys.stderr, flush=True)
        # the full form: batched + learned start + FP32 tier below 1e-4
        pool.run(geoms[:2 * best[0] * best[1]], "6-31g", guess="learned", fp32_below=1e-4)
        t1 = time.perf_counter()
        res_full = pool.run(geoms, "6-31g", guess="learned", fp32_below=1e-4)
        el_full = time.perf_counter() - t1
        print("batch full form done", file=sys.stderr, flush=True)
    # per-molecule energies against their own single-molecule runs, and the
    # plain path (one molecule at a time, FP64, core guess) on a sample
    dmax = 0.0
    sample = list(range(0, 512, 8))
    for i in sample[:4]:   # warm-up
        Molecule(geoms[i], "6-31g").RHF(doPrint=False, direct=True)
    t2 = time.perf_counter()
    for i in sample:
        m = Molecule(geoms[i], "6-31g")
        m.RHF(doPrint=False, direct=True)
        dmax = max(dmax, abs(m.energy.real - res[i]["energy"]))
    plain_rate = len(sample) / (time.perf_counter() - t2)
    e64 = np.array([r["energy"] for r in res])
    efull = np.array([r["energy"] for r in res_full])
    return {
        "metric": METRIC, "value": len(geoms) / el, "unit": "SCF/s", "n_gpus": 1, "steps": 1, "warmup": 0,
        "ms_per_step": 1e3 * el, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "512 x (H2O)n n=2..6 6-31G direct RHF, batched", "seed": 1, "basis": "6-31g",
                   "procs": best[0], "threads_per_proc": best[1],
                   "autotune_scf_per_s": {"%dx%d" % k: v for k, v in rates.items()}},
        "all_converged": all(r["converged"] for r in res),
        "mean_iterations": float(np.mean([r["iterations"] for r in res])),
        "max_abs_dE_vs_single_runs": dmax,
        "plain_fp64_one_at_a_time_scf_per_s": plain_rate,
        "full_form": {"what": "batched + learned start + FP32 tier below 1e-4", "scf_per_s": len(geoms) / el_full,
                      "all_converged": all(r["converged"] for r in res_full),
                      "mean_iterations": float(np.mean([r["iterations"] for r in res_full])),
                      "max_abs_dE_vs_fp64_batch": float(np.max(np.abs(efull - e64))),
                      "speedup_vs_plain": (len(geoms) / el_full) / plain_rate},
    }


def batch_geometries(n=512, seed=1):
    """512 x (H2O)n, n uniform in 2..6, perturbed monomers (O-H N(0, 0.02 A),
    H-O-H N(0, 2 deg)), random orientations, O in a (2 + 1.4 n) A cube."""
    import numpy as np
    from mmd.workloads import water_cluster
    rng = np.random.default_rng(seed)
    ns = rng.integers(2, 7, size=n)
    return [water_cluster(int(k), 2.0 + 1.4 * int(k), seed=1000 + i, perturb=True) for i, k in enumerate(ns)]


def bench_batch_device(args):
    """The headline workload: 512 x (H2O)n, n = 2..6, 6-31G.  Full form =
    batched on the device (mmd_batch_rhf, groups of `--group` molecules in one
    block-diagonal basis, graph-replayed iterations) + learned start + FP32
    tier below 1e-4, from the geometry strings to converged energies.  Checked
    per molecule against the batched FP64 core-guess run (|dE| <= 1 uEh), and
    compared with the plain FP64 path one molecule at a time
    (Molecule(g).RHF(direct=True)) on every 8th geometry."""
    import numpy as np
    from mmd.batch import run_batch_device
    from mmd.molecule import Molecule
    geoms = batch_geometries()
    G = args.group
    full_kw = dict(group=G, guess="learned", fp32_below=1e-4)
    for _ in range(max(1, args.warmup)):
        run_batch_device(geoms[:2 * G], "6-31g", **full_kw)
        run_batch_device(geoms[:G], "6-31g", group=G)
    print("warmup done", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    for s in range(args.steps):
        full = run_batch_device(geoms, "6-31g", **full_kw)
        print("full-form step %d done" % s, file=sys.stderr, flush=True)
    el_full = (time.perf_counter() - t0) / args.steps
    t1 = time.perf_counter()
    ref = run_batch_device(geoms, "6-31g", group=G)
    el_64 = time.perf_counter() - t1
    print("batched FP64 done", file=sys.stderr, flush=True)
    e_full = np.array([r["energy"] for r in full])
    e_64 = np.array([r["energy"] for r in ref])
    dE = np.abs(e_full - e_64)
    sample = list(range(0, len(geoms), 8))
    for i in sample[:2]:
        Molecule(geoms[i], "6-31g").RHF(doPrint=False, direct=True)
    t2 = time.perf_counter()
    d_plain = 0.0
    for i in sample:
        m = Molecule(geoms[i], "6-31g")
        m.RHF(doPrint=False, direct=True)
        d_plain = max(d_plain, abs(m.energy.real - e_64[i]))
    plain_rate = len(sample) / (time.perf_counter() - t2)
    print("plain one-at-a-time done", file=sys.stderr, flush=True)
    value = len(geoms) / el_full if dE.max() <= 1e-6 and all(r["converged"] for r in full) else 0.0
    return {
        "value": value, "ms_per_step": 1e3 * el_full,
        "config": {"workload": "512 x (H2O)n n=2..6 6-31G direct RHF, full form: batched on the device "
                               "(groups of %d, HIP graph) + learned start + FP32 tier below 1e-4" % G,
                   "seed": 1, "basis": "6-31g", "molecules": len(geoms), "group": G, "conver": 1e-8,
                   "acc2e": 1e-12},
        "max_abs_dE_full_vs_fp64_uEh": float(dE.max() * 1e6),
        "all_converged": bool(all(r["converged"] for r in full)),
        "mean_iterations_full": float(np.mean([r["iterations"] for r in full])),
        "mean_iterations_fp64": float(np.mean([r["iterations"] for r in ref])),
        "batched_fp64_scf_per_s": len(geoms) / el_64,
        "plain_fp64_one_at_a_time_scf_per_s": plain_rate,
        "plain_fp64_sample": "every 8th geometry (%d molecules), Molecule(g, '6-31g').RHF(direct=True)" % len(sample),
        "max_abs_dE_plain_vs_batched_fp64": float(d_plain),
        "speedup_vs_plain_fp64": (len(geoms) / el_full) / plain_rate,
    }


def bench_ao2mo(args):
    """Four-index transformation of the stored ERI tensor of (H2O)4 / cc-pVDZ
    (N = 100, 7 A cube, seed 0) on the GPU: four FP64 GEMM quarter
    transforms, 8 N^5 FLOPs; FP64 peak 78.6 TFLOP/s."""
    import numpy as np
    from mmd.molecule import Molecule
    from mmd.workloads import water_cluster
    mol = Molecule(water_cluster(4, 7.0, seed=0), "cc-pvdz")
    mol.RHF(doPrint=False, direct=False)
    N = mol.nbasis
    ctx = mol.native
    C = mol.C.real
    for _ in range(max(1, args.warmup)):
        ctx.ao2mo(mol.TwoE, C)
    ms = []
    for _ in range(args.steps):
        ctx.ao2mo(mol.TwoE, C)
        ms.append(ctx.last_ao2mo_ms)
    t = float(np.median(ms))
    flops = 8.0 * N ** 5
    return {"metric": "four-index transformation (ao2mo), device time of the four GEMM passes", "value": flops / (t * 1e-3) / 1e12,
            "unit": "TFLOP/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": t,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": "(H2O)4 cc-pVDZ ao2mo", "n_basis": N, "seed": 0},
            "roofline": {"bound": "mfma", "achieved": flops / (t * 1e-3) / 1e12, "peak": FP64_PEAK_TFLOPS,
                         "unit": "TFLOP/s", "frac": flops / (t * 1e-3) / 1e12 / FP64_PEAK_TFLOPS, "traffic": None}}


def copy_calibration():
    """A device-to-device copy of 256 MiB (five launches), the reference
    point for scaling the HBM byte counters of the profiles."""
    import torch
    n = 256 << 20
    src = torch.ones(n // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    return {"metric": "device copy", "bytes_read_per_launch": n, "bytes_written_per_launch": n,
            "launches": 5, "GB_per_s": 2 * n * 5 / el / 1e9}


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-profile", action="store_true")
    ap.add_argument("--serial-classes", action="store_true",
                    help="run the Fock classes one after the other in the profiled step (for counter runs)")
    ap.add_argument("--group", type=int, default=64, help="molecules per block-diagonal batch (batch workloads)")
    ap.add_argument("--workload", choices=("h2o8", "h2o16-tier", "batch512", "batch-device", "ao2mo",
                                           "copy-calibration"), default="h2o8",
                    help="h2o8: (H2O)8 cc-pVDZ FP64 (default); h2o16-tier: (H2O)16 cc-pVDZ FP64 vs FP32 tier; "
                         "batch512: 512 small 6-31G clusters")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    from mmd.workloads import water_cluster
    from mmd import _native
    if args.workload == "copy-calibration":
        print(json.dumps(copy_calibration()))
        return
    if args.workload == "batch512":   # starts its worker processes before any GPU use here
        print(json.dumps(bench_batch(args)))
        return
    if _native.device_count() < 1:
        print("no HIP device visible", file=sys.stderr)
        sys.exit(1)
    if args.workload == "batch-device":
        print(json.dumps(bench_batch_device(args)))
        return
    if args.workload == "h2o16-tier":
        print(json.dumps(bench_tier(args)))
        return
    if args.workload == "ao2mo":
        print(json.dumps(bench_ao2mo(args)))
        return

    geom = water_cluster(8, 7.0, seed=0)
    basis = "cc-pvdz"
    for _ in range(args.warmup):
        run_scf(geom, basis)
    print("warmup done", file=sys.stderr, flush=True)

    mols = []
    t0 = time.perf_counter()
    for s in range(args.steps):
        # per-class HIP events on the classes' own streams, inside the timed
        # region (last step)
        mols.append(run_scf(geom, basis, profile=(not args.no_profile) and s == args.steps - 1,
                            serial=args.serial_classes))
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
                   "min_OO_A": 2.7, "acc2e": 1e-12, "conver": 1e-8, "hw_queues": int(HW_QUEUES)},
        "fock_build_ms": fock_med, "quartets_per_s": nquart / (fock_med * 1e-3),
        "unique_bf_quartets": nquart, "scf_iterations": int(m0.iterations), "energy": energies[-1],
        "delta_energy": float(abs(m0.delta_energy.real)),
    }

    if not args.no_profile:
        stats = [st for m in mols for st in getattr(m, "_profile_stats", [])]
        per = roofline_from_stats(stats)
        tot_f = sum(v[0] for v in per.values())
        tot_t = sum(v[1] for v in per.values())
        dom = max(per, key=lambda k: per[k][1])
        f, t, n = per[dom][:3]
        achieved = f / (t * 1e-3) / 1e12
        out["roofline"] = {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                           "frac": achieved / FP64_PEAK_TFLOPS, "traffic": TRAFFIC.get(dom),
                           "kernel": "eri_fock_class_kernel<%d,%d,%d,%d,double>" % dom,
                           "kernel_avg_ms": t / n, "launches": n,
                           "all_classes_achieved": tot_f / (tot_t * 1e-3) / 1e12,
                           "all_classes_event_ms_per_build": tot_t / max(1, len(stats)),
                           # the whole build: model FLOPs of all classes per
                           # build over the median Fock-build time (the
                           # classes overlap on the streams)
                           "build_achieved": tot_f / max(1, len(stats)) / (fock_med * 1e-3) / 1e12,
                           "build_frac": tot_f / max(1, len(stats)) / (fock_med * 1e-3) / 1e12 / FP64_PEAK_TFLOPS}
        nb = max(1, len(stats))
        out["class_breakdown"] = {
            "(%d%d|%d%d)" % k: {"ms_per_build": v[1] / nb, "tflops": v[0] / (v[1] * 1e-3) / 1e12 if v[1] else 0,
                                "shell_quartets_per_build": v[3] // nb, "prim_quartets_per_build": v[4] // nb}
            for k, v in sorted(per.items(), key=lambda kv: -kv[1][1])}
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(geom, basis, len(m0.fock_ms), args.cpu_seconds)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
