#!/usr/bin/env python
"""Benchmark: converged SCF/s of 512 small water clusters on one MI355X.

Default workload (the headline): 512 x (H2O)n with n uniform in 2..6
(perturbed monomers, random orientations, seed 1), 6-31G, direct RHF to
||dP|| < 1e-8.  One step = all 512 molecules from geometry strings to
converged energies in the full form: batched on the device (groups of 128
molecules in one block-diagonal basis, four groups in flight, each SCF
iteration replayed as a HIP graph), learned initial density, FP64.  Every molecule's energy is checked
against the batched FP64 core-guess run (|dE| <= 1 uEh), which is checked
against the plain FP64 path one molecule at a time; the plain path's rate is
measured beside it.

Printed (one JSON line): SCF/s (value), the roofline of the dominant class
kernel (model FLOPs per launch over its HIP-event time, from the profiled
last timed step), (H2O)8 / cc-pVDZ Fock-build ms and quartets/s as extra
keys, and a CPU baseline from the oracle on a bounded sample of the same
workload, rebased on formPT's survivors per iteration.  Other workloads:
--workload h2o8 | h2o16-tier | batch512 | ao2mo.
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
    (0, 0, 0, 0): 2.203e+07, (1, 0, 0, 0): 4.100e+07, (1, 0, 1, 0): 4.481e+07, (1, 1, 0, 0): 2.065e+07,
    (1, 1, 1, 0): 5.334e+07, (1, 1, 1, 1): 7.910e+06, (2, 0, 0, 0): 1.781e+07, (2, 0, 1, 0): 3.797e+07,
    (2, 0, 1, 1): 1.138e+07, (2, 0, 2, 0): 6.591e+06, (2, 1, 0, 0): 1.420e+07, (2, 1, 1, 0): 1.657e+07,
    (2, 1, 1, 1): 7.255e+06, (2, 1, 2, 0): 5.512e+06, (2, 1, 2, 1): 5.707e+06, (2, 2, 0, 0): 3.591e+06,
    (2, 2, 1, 0): 4.837e+06, (2, 2, 1, 1): 2.786e+06, (2, 2, 2, 0): 2.254e+06, (2, 2, 2, 1): 5.863e+06,
    (2, 2, 2, 2): 1.207e+06,
}


# The same for the dominant class kernel of the headline workload.
TRAFFIC_HEADLINE = {
    (0, 0, 0, 0): 1.001e+09, (1, 0, 0, 0): 8.355e+08, (1, 0, 1, 0): 3.344e+08, (1, 1, 0, 0): 3.327e+08,
    (1, 1, 1, 0): 2.322e+08, (1, 1, 1, 1): 3.710e+07,
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


def run_scf(geom, basis, profile=False, serial=False, class_impl=None):
    from mmd.molecule import Molecule
    mol = Molecule(geom, basis)
    if class_impl is not None:
        mol.native.set_class_impl(class_impl)
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
        if args.class_impl is not None:
            ctx.set_class_impl(args.class_impl)
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


def _profile_sums(ctxs):
    """Per-class sums of the profiled builds of all contexts: {class: [nq64,
    nq32, np64, np32, ms64, ms32, launches]}."""
    per = {}
    for c in ctxs:
        lc, nq, npr, ms, nb = c.fock_profile_sums()
        for k in range(len(nb)):
            if nb[k] == 0:
                continue
            key = tuple(int(x) for x in lc[k])
            acc = per.setdefault(key, [0, 0, 0, 0, 0.0, 0.0, 0])
            for t in range(2):
                acc[t] += int(nq[k, t])
                acc[2 + t] += int(npr[k, t])
                acc[4 + t] += float(ms[k, t])
            acc[6] += int(nb[k])
    return per


def headline_roofline(per):
    """Roofline of the dominant class kernel (most FP64 event time): model
    FLOPs of its surviving primitive quartets (plus 12 per surviving basis-
    function quartet for the digestion) per launch over the launch's event
    time; FP64 peak."""
    def flops(key, t):
        la, lb, lc, ld = key
        return per[key][2 + t] * class_flops(*key) + 12.0 * per[key][t] * ncart(la) * ncart(lb) * ncart(lc) * ncart(ld)
    dom = max(per, key=lambda k: per[k][4])
    v = per[dom]
    achieved = flops(dom, 0) / (v[4] * 1e-3) / 1e12
    tot_f = sum(flops(k, t) for k in per for t in range(2))
    tot_t = sum(per[k][4] + per[k][5] for k in per)
    return {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved / FP64_PEAK_TFLOPS, "traffic": TRAFFIC_HEADLINE.get(dom),
            "kernel": "eri_fock_class_kernel<%d,%d,%d,%d,double>" % dom,
            "kernel_avg_ms": v[4] / v[6], "launches": v[6],
            "model_flops_per_launch": flops(dom, 0) / v[6],
            "prim_quartets_per_launch": v[2] / v[6], "shell_quartets_per_launch": v[0] / v[6],
            "all_classes_achieved": tot_f / (tot_t * 1e-3) / 1e12,
            "timing": "HIP events on the class's stream around its FP64 launch, profiled timed step"}


def _record_builds(mol):
    """Wrap mol._direct_G so that every Fock build's real dP = P - P_old is kept."""
    import numpy as np
    mol._dps = []
    orig = mol._direct_G

    def rec(P, P_old):
        mol._dps.append(np.real(np.asarray(P) - np.asarray(P_old)).copy())
        return orig(P, P_old)
    mol._direct_G = rec
    return mol


def cpu_baseline_batch(geoms, seconds):
    """The oracle (C restatement of formPT, cython/fock.pyx:10-87, with the
    reference's per-quartet ERI) timed on the host cores, rebased on formPT's
    survivors per iteration.

    One molecule of each size n = 2..6 of the workload runs the plain GPU SCF
    (reference iteration semantics) with every build's dP recorded.  For each
    build the exact number of quartets formPT keeps is counted (no integrals)
    and windows of formPT (limit / start_ij: L consecutive canonical quartets
    from a random ij row, survivors computed with their ERIs) are timed; a
    least-squares fit t = a visited + c survivors over the windows prices the
    builds.  An SCF = one-electron integrals + Schwarz pass (timed whole) +
    its builds.  The 512-molecule time is the sum over the size mix.  The
    all-cores leg times the same windows on every allowed core at once
    (threads; the oracle's C calls release the GIL) and divides the total
    over the cores, as independent molecules would run."""
    import concurrent.futures as cf
    import numpy as np
    import oracle
    from mmd.molecule import Molecule
    gbs = os.path.join(PKG, "mmd", "basis", "6-31g.gbs")
    n_of = [len([ln for ln in g.split("\n")[1:] if ln.strip()]) // 3 for g in geoms]
    sizes = sorted(set(n_of))
    rng = np.random.default_rng(77)
    jobs = []      # (size, om, screen, dP, start_ij, L)
    mols = {}
    for n in sizes:
        i = n_of.index(n)
        m = _record_builds(Molecule(geoms[i], "6-31g"))
        m.RHF(doPrint=False, direct=True)
        om = oracle.OracleMolecule(geoms[i], gbs)
        t0 = time.perf_counter()
        om.one_electron()
        scr = om.schwarz()
        t_setup = time.perf_counter() - t0
        N = om.nbasis
        npair = N * (N + 1) // 2
        visited = npair * (npair + 1) // 2
        surv = [oracle.form_pt_survivors(dp, scr, 1e-12) for dp in m._dps]
        mols[n] = {"N": N, "builds": len(m._dps), "visited": visited, "survivors": surv, "setup_s": t_setup,
                   "iterations": int(m.iterations)}
        for b, dp in enumerate(m._dps):
            for _ in range(3):
                jobs.append((n, b, om, scr, dp, int(rng.integers(0, npair)), 1500))

    def run_window(job):
        n, b, om, scr, dp, start, L = job
        t0 = time.perf_counter()
        _, sv = om.form_pt(dp.astype(complex), np.zeros_like(dp, dtype=complex), scr, 1e-12, limit=L, start_ij=start)
        el = time.perf_counter() - t0
        # visited = quartets from row `start` up to L (fewer at the last rows)
        N = mols[n]["N"]
        npair = N * (N + 1) // 2
        avail = sum(ij + 1 for ij in range(start, npair)) if npair - start < 4 * L else L
        return n, min(L, avail), sv, el

    def fit(res):
        cost = {}
        for n in sizes:
            r = [(v, sv, el) for (nn, v, sv, el) in res if nn == n]
            A = np.array([[v, sv] for v, sv, _ in r], dtype=float)
            y = np.array([el for _, _, el in r])
            coef, *_ = np.linalg.lstsq(A, y, rcond=None)
            a, c = float(coef[0]), float(coef[1])
            if a < 0 or c <= 0:   # degenerate sample: price every quartet by its survivors only
                a, c = 0.0, float(y.sum() / max(1.0, A[:, 1].sum()))
            info = mols[n]
            t_scf = info["setup_s"] + sum(a * info["visited"] + c * s for s in info["survivors"])
            cost[n] = (t_scf, a, c)
        return cost

    # 1 core
    t0 = time.perf_counter()
    res1 = []
    for job in jobs:
        res1.append(run_window(job))
        if time.perf_counter() - t0 > 3 * seconds:
            break
    single_wall = time.perf_counter() - t0
    cost1 = fit(res1)
    total1 = sum(cost1[n][0] for n in n_of)
    # all allowed cores at once
    try:
        ncores = len(os.sched_getaffinity(0))
    except AttributeError:
        ncores = os.cpu_count() or 1
    ncores = max(1, min(16, ncores))
    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(max_workers=ncores) as ex:
        resP = list(ex.map(run_window, jobs[:len(res1)]))
    par_wall = time.perf_counter() - t0
    costP = fit(resP)
    totalP = sum(costP[n][0] for n in n_of) / ncores
    return {
        "value": len(geoms) / total1, "unit": "SCF/s", "cores": 1, "kind": "port",
        "sample": ("oracle formPT windows (%d windows of <= 1500 canonical quartets at random ij rows, with ERIs) "
                   "on the recorded dP of every build of one molecule per size n=%s (plain GPU SCF, reference "
                   "iteration semantics), priced by t = a visited + c survivors with the exact survivor count of "
                   "each build; + one-electron and Schwarz timed whole; summed over the 512-molecule size mix "
                   "(%.1f s of 1-core CPU)") % (len(res1), ",".join(map(str, sizes)), single_wall),
        "nproc": os.cpu_count(),
        "all_cores": {"value": len(geoms) / totalP, "cores": ncores, "wall_s": par_wall,
                      "how": "the same windows on %d threads at once; total / cores" % ncores},
        "per_size": {str(n): {"nbasis": mols[n]["N"], "builds": mols[n]["builds"],
                              "iterations": mols[n]["iterations"],
                              "survivors_per_build": mols[n]["survivors"],
                              "visited_per_build": mols[n]["visited"],
                              "cpu_s_per_scf_1core": cost1[n][0], "us_per_survivor_1core": 1e6 * cost1[n][2]}
                     for n in sizes},
    }


def cfg2_extras(args):
    """(H2O)8 / cc-pVDZ FP64 direct RHF (seed 0, 7 A cube): median Fock-build
    ms over iterations 2..n and quartets/s (unique basis-function quartets
    per build over the build time)."""
    import numpy as np
    from mmd.workloads import water_cluster
    geom = water_cluster(8, 7.0, seed=0)
    run_scf(geom, "cc-pvdz")
    t0 = time.perf_counter()
    m = run_scf(geom, "cc-pvdz")
    el = time.perf_counter() - t0
    N = m.nbasis
    npair = N * (N + 1) // 2
    nquart = npair * (npair + 1) // 2
    fms = float(np.median(m.fock_ms[1:]))
    # the screened work of the same builds, from one more SCF with the
    # per-class survivor counters on (not timed)
    mp = run_scf(geom, "cc-pvdz", profile=True)
    shells, comps = [], []
    for lc, nq, npr, ms in mp._profile_stats[1:]:
        shells.append(int(nq.sum()))
        comps.append(int(sum(int(q) * ncart(a) * ncart(b) * ncart(c) * ncart(d) for (a, b, c, d), q in zip(lc, nq))))
    return {"workload": "(H2O)8 cc-pVDZ direct FP64 RHF", "n_basis": N, "fock_build_ms": fms,
            "quartets_per_s": nquart / (fms * 1e-3), "unique_bf_quartets": nquart, "scf_s": el,
            "energy": float(m.energy.real), "iterations": int(m.iterations),
            "surviving_shell_quartets_per_build": float(np.mean(shells)),
            "computed_component_quartets_per_build": float(np.mean(comps)),
            "computed_quartets_per_s": float(np.mean(comps)) / (fms * 1e-3),
            "quartets_note": "quartets_per_s counts every unique basis-function quartet of a build (the unit of "
                             "the reference's ERI() loop); computed_quartets_per_s counts the component quartets "
                             "of the shell quartets that survive the density-weighted screen (mean over builds "
                             "2..n)"}


def bench_headline(args):
    """The headline workload: 512 x (H2O)n, n = 2..6, 6-31G.  Full form =
    batched on the device (mmd_batch_rhf, groups of `--group` molecules in one
    block-diagonal basis, graph-replayed iterations, `--batch-streams` groups
    in flight, per-class ket slices tuned on one group before the timed
    steps) + learned start, FP64, from the geometry strings to converged
    energies.  Checked per molecule against the batched FP64
    core-guess run (|dE| <= 1 uEh), which is checked against the plain FP64
    path one molecule at a time (Molecule(g).RHF(direct=True)) on every 8th
    geometry; the plain path's rate is measured on those."""
    import numpy as np
    from mmd.batch import run_batch_device, _device_contexts
    from mmd.molecule import Molecule
    geoms = batch_geometries(args.molecules)
    G = args.group
    full_kw = dict(group=G, guess="learned", fp32_below=args.fp32_below, streams=args.batch_streams,
                   class_streams=args.class_streams)
    ctxs = _device_contexts(0, max(1, args.batch_streams))
    slices = None
    if not args.no_tune:
        # per-class ket slices of the row kernels, measured on one group of
        # the workload's molecules (not timed)
        from mmd.tune import tune_ket_slices
        idx = np.linspace(0, len(geoms) - 1, min(G, len(geoms))).astype(int)
        slices, _ = tune_ket_slices([geoms[i] for i in idx], "6-31g", ctx=ctxs[0], others=ctxs[1:])
        print("ket slices tuned: %s" % list(slices), file=sys.stderr, flush=True)
    for _ in range(max(1, args.warmup)):
        run_batch_device(geoms[:4 * G], "6-31g", **full_kw)
        run_batch_device(geoms[:2 * G], "6-31g", group=G, streams=args.batch_streams)
    print("warmup done", file=sys.stderr, flush=True)
    step_s = []
    full = None
    per = None
    t0 = time.perf_counter()
    for s in range(args.steps):
        prof = (not args.no_profile) and s == args.steps - 1
        if prof:
            for c in ctxs:
                c.set_profiling(True)
        ts = time.perf_counter()
        full = run_batch_device(geoms, "6-31g", **full_kw)
        step_s.append(time.perf_counter() - ts)
        if prof:
            per = _profile_sums(ctxs)
            for c in ctxs:
                c.set_profiling(False)
        print("full-form step %d done (%.3f s%s)" % (s, step_s[-1], ", profiled" if prof else ""),
              file=sys.stderr, flush=True)
    el_full = (time.perf_counter() - t0) / args.steps
    if args.steps_only:   # counter and trace runs: the full-form steps alone, unchecked
        out = {"metric": METRIC, "value": None, "unit": "SCF/s", "checked": False,
               "note": "--steps-only: full-form steps without the FP64 and plain checks",
               "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * el_full, "step_s": step_s,
               "molecules": len(geoms), "ket_slices": None if slices is None else [int(x) for x in slices]}
        if per:
            out["roofline"] = headline_roofline(per)
        return out
    t1 = time.perf_counter()
    ref = run_batch_device(geoms, "6-31g", group=G, streams=args.batch_streams)
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
    ok = dE.max() <= 1e-6 and all(r["converged"] for r in full) and d_plain <= 1e-6
    value = len(geoms) / el_full if ok else 0.0
    ns = [len([ln for ln in g.split("\n")[1:] if ln.strip()]) // 3 for g in geoms]
    out = {
        "metric": METRIC, "value": value, "unit": "SCF/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * el_full, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64+f32", "data": "synthetic",
        "config": {"workload": "%d x (H2O)n, n=2..6, 6-31G direct RHF, full form (batched on the device in "
                               "groups of %d, %d groups in flight, HIP-graph iterations, learned start, %s) vs "
                               "plain FP64 one at a time" % (len(geoms), G, args.batch_streams,
                                                            "FP32 tier below %g" % args.fp32_below
                                                            if args.fp32_below > 0 else "FP64"),
                   "molecules": len(geoms), "seed": 1, "basis": "6-31g", "group": G,
                   "groups_in_flight": args.batch_streams, "class_streams": args.class_streams,
                   "fp32_below": args.fp32_below, "conver": 1e-8,
                   "acc2e": 1e-12,
                   "size_mix": {str(k): ns.count(k) for k in sorted(set(ns))}, "hw_queues": int(HW_QUEUES)},
        "step": "one pass over the %d molecules, geometry strings to converged energies" % len(geoms),
        "step_s": step_s,
        "ket_slices": None if slices is None else [int(x) for x in slices],
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
    if per:
        out["roofline"] = headline_roofline(per)
        nb = {k: max(1, v[6]) for k, v in per.items()}
        out["class_breakdown"] = {
            "(%d%d|%d%d)" % k: {"fp64_ms_per_launch": v[4] / nb[k], "fp32_ms_per_launch": v[5] / nb[k],
                                "shell_quartets_per_launch": (v[0] + v[1]) / nb[k],
                                "prim_quartets_per_launch": (v[2] + v[3]) / nb[k], "launches": v[6]}
            for k, v in sorted(per.items(), key=lambda kv: -kv[1][4])}
    out["cfg2"] = cfg2_extras(args)
    print("cfg2 extras done", file=sys.stderr, flush=True)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline_batch(geoms, args.cpu_seconds)
        print("cpu baseline done", file=sys.stderr, flush=True)
    return out


def bench_gradient(args):
    """RHF nuclear gradient (Molecule.forces, mmd_rhf_gradient) of (H2O)4 /
    6-31G (N = 52, 5.5 A cube, seed 0, perturbed monomers) after a converged
    direct SCF; one step = one forces() call (derivative integrals of all
    atoms and directions, contraction, host transfer).  CPU baseline: the
    oracle's gradient (forces.py restated, one core) of one water in the
    same basis, scaled by the canonical basis-function quartet count."""
    import numpy as np
    import oracle
    from mmd.molecule import Molecule
    from mmd.workloads import water_cluster
    geom = water_cluster(4, 5.5, seed=0, perturb=True)
    mol = Molecule(geom, "6-31g")
    mol.RHF(doPrint=False, direct=True, conver=1e-10)
    for _ in range(max(1, args.warmup)):
        mol.forces()
    t0 = time.perf_counter()
    dev = []
    for _ in range(args.steps):
        f = mol.forces()
        dev.append(mol.gradient_ms)
    el = (time.perf_counter() - t0) / args.steps
    N = mol.nbasis
    nq = (N * (N + 1) // 2) * (N * (N + 1) // 2 + 1) // 2
    out = {"metric": "RHF nuclear gradients/s", "value": 1.0 / el, "unit": "gradients/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * el, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "(H2O)4 6-31G RHF nuclear gradient", "n_basis": N, "seed": 0, "box_A": 5.5,
                      "canonical_quartets": nq},
           "device_ms_median": float(np.median(dev)), "max_abs_force": float(np.abs(f).max()),
           "sum_of_forces": float(np.abs(f.sum(0)).max())}
    if not args.no_cpu_baseline:
        w = water_cluster(1, 3.0, seed=0, perturb=True)
        om = oracle.OracleMolecule(w, os.path.join(PKG, "mmd", "basis", "6-31g.gbs"))
        r = oracle.rhf(om, conver=1e-10)
        t1 = time.perf_counter()
        om.rhf_gradient(r["P"], r["F"])
        tw = time.perf_counter() - t1
        n1 = om.nbasis * (om.nbasis + 1) // 2
        nq1 = n1 * (n1 + 1) // 2
        out["cpu_baseline"] = {"value": 1.0 / (tw * nq / nq1), "unit": "gradients/s", "cores": 1, "kind": "port",
                               "sample": "oracle gradient of one water / 6-31G (%d quartets, %.2f s), scaled by the "
                                         "canonical quartet count to %d (extrapolated)" % (nq1, tw, nq)}
    return out


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


def pmc_summary(paths):
    """Per-kernel averages per dispatch of rocprofv3 --pmc counter_collection
    CSV files (one line per kernel, most GRBM_GUI_ACTIVE first): what the
    committed counter profiles are made from."""
    import csv
    import collections
    agg = collections.defaultdict(lambda: collections.defaultdict(float))
    disp = collections.defaultdict(lambda: collections.defaultdict(set))
    meta = {}
    for path in paths:
        with open(path) as f:
            for r in csv.DictReader(f):
                k = r["Kernel_Name"]
                agg[k][r["Counter_Name"]] += float(r["Counter_Value"])
                disp[k][r["Counter_Name"]].add(r["Dispatch_Id"])
                meta[k] = ",".join(r.get(x, "") for x in ("VGPR_Count", "Accum_VGPR_Count", "SGPR_Count",
                                                          "LDS_Block_Size", "Workgroup_Size", "Grid_Size",
                                                          "Scratch_Size"))
    cols = sorted({c for v in agg.values() for c in v})
    lines = ["kernel|dispatches|" + "|".join(cols) + "|vgpr,agpr,sgpr,lds,wg,grid,scratch"]
    key = "GRBM_GUI_ACTIVE" if "GRBM_GUI_ACTIVE" in cols else cols[0]
    for k, v in sorted(agg.items(), key=lambda kv: -kv[1].get(key, 0)):
        n = max(len(d) for d in disp[k].values())
        vals = "|".join("%.4g" % (v[c] / max(1, len(disp[k][c]))) if c in v else "" for c in cols)
        lines.append("%s|%d|%s|%s" % (k.replace("mmd::", "").replace("(anonymous namespace)::", "")[:90], n, vals,
                                      meta[k]))
    return "\n".join(lines)


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-profile", action="store_true")
    ap.add_argument("--no-tune", action="store_true",
                    help="headline: keep the default ket slices instead of tuning them on one group")
    ap.add_argument("--serial-classes", action="store_true",
                    help="run the Fock classes one after the other in the profiled step (for counter runs)")
    ap.add_argument("--group", type=int, default=128, help="molecules per block-diagonal batch (batch workloads)")
    ap.add_argument("--batch-streams", type=int, default=4, help="groups in flight at once (headline)")
    ap.add_argument("--molecules", type=int, default=512,
                    help="headline molecules (512 is the workload; fewer only for counter runs)")
    ap.add_argument("--class-streams", type=int, default=2,
                    help="headline: streams per context for the classes of a Fock build (with three groups in "
                         "flight, two each fill the hardware queues without sharing them)")
    ap.add_argument("--fp32-below", type=float, default=0.0,
                    help="FP32 tier threshold of the headline's full form (0: all FP64; at these sizes the "
                         "tier's second pass over the units costs more than its FP32 arithmetic saves)")
    ap.add_argument("--workload", choices=("headline", "h2o8", "h2o16-tier", "batch512", "ao2mo", "gradient",
                                           "copy-calibration"), default="headline",
                    help="headline: 512 x (H2O)2-6 6-31G, full form vs plain FP64 (default); "
                         "h2o8: (H2O)8 cc-pVDZ FP64; h2o16-tier: (H2O)16 cc-pVDZ FP64 vs FP32 tier; "
                         "batch512: 512 small 6-31G clusters through worker processes")
    ap.add_argument("--class-impl", type=int, choices=(0, 1, 2), default=None,
                    help="h2o8 / h2o16-tier: kernels of the classes with a row and a matrix-core kernel "
                         "(0 row, 1 matrix cores, 2 measured per class, the default)")
    ap.add_argument("--steps-only", action="store_true",
                    help="headline: run the full-form steps only (for counter and trace runs; no value)")
    ap.add_argument("--pmc-summary", nargs="+", metavar="CSV",
                    help="print per-kernel counter averages of rocprofv3 --pmc CSV files and exit")
    args = ap.parse_args()
    if args.pmc_summary:
        print(pmc_summary(args.pmc_summary))
        return
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
    if args.workload == "headline":
        print(json.dumps(bench_headline(args)))
        return
    if args.workload == "h2o16-tier":
        print(json.dumps(bench_tier(args)))
        return
    if args.workload == "ao2mo":
        print(json.dumps(bench_ao2mo(args)))
        return
    if args.workload == "gradient":
        print(json.dumps(bench_gradient(args)))
        return

    geom = water_cluster(8, 7.0, seed=0)
    basis = "cc-pvdz"
    for _ in range(args.warmup):
        run_scf(geom, basis, class_impl=args.class_impl)
    print("warmup done", file=sys.stderr, flush=True)

    mols = []
    t0 = time.perf_counter()
    for s in range(args.steps):
        # per-class HIP events on the classes' own streams, inside the timed
        # region (last step)
        mols.append(run_scf(geom, basis, profile=(not args.no_profile) and s == args.steps - 1,
                            serial=args.serial_classes, class_impl=args.class_impl))
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
                   "min_OO_A": 2.7, "acc2e": 1e-12, "conver": 1e-8, "hw_queues": int(HW_QUEUES),
                   "class_impl": 2 if args.class_impl is None else args.class_impl},
        "class_choice": [int(x) for x in m0.native.set_class_impl()],
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
