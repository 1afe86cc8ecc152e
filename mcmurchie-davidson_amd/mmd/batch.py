"""Many independent RHF calculations on one GPU.

``run_batch`` runs ``Molecule(geometry, basis).RHF(...)`` for a list of
geometries with a pool of worker threads.  Each worker owns one libmmd
context (its own streams and device buffers) and reuses it for every
molecule it takes, so that the Fock builds of different molecules overlap
on the GPU while other workers diagonalise on the host cores (the library
calls and LAPACK release the GIL).  For small molecules the SCF driver's
Python code itself is the bottleneck, so ``run_batch_procs`` spreads the
batch over several worker processes (each with its own threads and GPU
contexts); ``autotune_workers`` / ``autotune_procs`` pick the pool shape by
timing a short calibration batch.

The SCF of every molecule is exactly ``SCF.RHF`` (mmd/scf.py:10-80 of the
reference); batching changes only what runs concurrently.
"""
import concurrent.futures as cf
import json
import os
import subprocess
import sys
import threading
import time

import numpy as np

from mmd import _native
from mmd.molecule import Molecule
from mmd.scf import _blas_threads


def _worker_ctx(local, device):
    ctx = getattr(local, "ctx", None)
    if ctx is None:
        ctx = _native.Context(device)
        ctx.set_streams(1)   # the other workers' contexts supply the concurrency
        local.ctx = ctx
    return ctx


def _run_one(geom, basis, kwargs, local, device, keep_density):
    mol = Molecule(geom, basis)
    ctx = mol.attach_context(_worker_ctx(local, device))
    ctx.set_precision(kwargs.get("fp32_below", 0.0))
    if not kwargs.get("device_loop", False):
        # small molecules side by side: the host loop (the GPU-resident one
        # measured 179 against 247 SCF/s on 512 (H2O)2-6 / 6-31G)
        mol.device_loop_min_nbasis = None
    t0 = time.perf_counter()
    mol._rhf(False, kwargs.get("DIIS", True), kwargs.get("direct", True), kwargs.get("conver", 1e-8),
             kwargs.get("acc2e", 1e-12), kwargs.get("guess", "core"))
    out = {"energy": float(mol.energy.real), "iterations": int(mol.iterations),
           "converged": bool(mol.is_converged), "nbasis": int(mol.nbasis),
           "seconds": time.perf_counter() - t0}
    if keep_density:
        out["P"] = np.real(mol.P).copy()
    return out


def run_batch(geometries, basis="6-31g", workers=8, device=0, keep_density=False, executor=None, **rhf_kwargs):
    """RHF energies of all geometries.  rhf_kwargs: DIIS, direct (default
    True), conver, acc2e, guess, fp32_below (FP32 tier of the Fock build,
    0 = all FP64), device_loop (default False: the host SCF loop; True: the
    GPU-resident one).
    Returns one dict per geometry, in order:
    energy, iterations, converged, nbasis, seconds (and the converged density
    P with keep_density).  ``executor``: a thread pool to reuse (its
    threads keep their GPU contexts between calls)."""
    if workers < 1:
        raise ValueError("workers must be >= 1")
    with _blas_threads(1):
        if executor is not None:
            futs = [executor.submit(_run_one, g, basis, rhf_kwargs, _TLS, device, keep_density) for g in geometries]
            return [f.result() for f in futs]
        local = threading.local()
        with cf.ThreadPoolExecutor(max_workers=workers) as ex:
            futs = [ex.submit(_run_one, g, basis, rhf_kwargs, local, device, keep_density) for g in geometries]
            return [f.result() for f in futs]


_TLS = threading.local()   # per-thread contexts of a reused executor


class BatchPool:
    """Persistent worker processes (``threads`` threads each, with their GPU
    contexts kept between calls) that solve batches of geometries.  Start it
    from a process that has not used the GPU yet."""

    def __init__(self, procs=4, threads=2, device=0):
        if procs < 1 or threads < 1:
            raise ValueError("procs and threads must be >= 1")
        pkg = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
        env = dict(os.environ)
        env["PYTHONPATH"] = pkg + (os.pathsep + env["PYTHONPATH"] if env.get("PYTHONPATH") else "")
        # one hardware queue per worker thread (its context's stream), so that
        # the pool does not oversubscribe the GPU's queues
        env["GPU_MAX_HW_QUEUES"] = str(min(threads, 4))
        self.procs, self.threads = procs, threads
        self._ps = [subprocess.Popen([sys.executable, "-m", "mmd.batch", "--serve", str(threads), str(device)],
                                     stdin=subprocess.PIPE, stdout=subprocess.PIPE, env=env) for _ in range(procs)]

    def run(self, geometries, basis="6-31g", **rhf_kwargs):
        chunks = [list(range(i, len(geometries), self.procs)) for i in range(self.procs)]
        for c, p in zip(chunks, self._ps):
            p.stdin.write((json.dumps({"geoms": [geometries[j] for j in c], "basis": basis,
                                       "kwargs": rhf_kwargs}) + "\n").encode())
            p.stdin.flush()
        out = [None] * len(geometries)
        for c, p in zip(chunks, self._ps):
            line = p.stdout.readline()
            if not line:
                raise RuntimeError("batch worker exited (status %s)" % p.poll())
            for j, r in zip(c, json.loads(line.decode())):
                out[j] = r
        return out

    def close(self):
        for p in self._ps:
            try:
                p.stdin.close()
            except Exception:
                pass
        for p in self._ps:
            p.wait()
        self._ps = []

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def _serve_main(threads, device):
    ex = cf.ThreadPoolExecutor(max_workers=threads)
    for line in sys.stdin:
        job = json.loads(line)
        res = run_batch(job["geoms"], job["basis"], workers=threads, device=device, executor=ex, **job["kwargs"])
        sys.stdout.write(json.dumps(res) + "\n")
        sys.stdout.flush()
    ex.shutdown()


def autotune_workers(geometries, basis="6-31g", candidates=(4, 8, 16), device=0, **rhf_kwargs):
    """Pool size with the highest SCF/s on the calibration geometries, and
    the measured rates {workers: SCF/s}."""
    rates = {}
    for w in candidates:
        t0 = time.perf_counter()
        run_batch(geometries, basis, workers=w, device=device, **rhf_kwargs)
        rates[w] = len(geometries) / (time.perf_counter() - t0)
    best = max(rates, key=rates.get)
    return best, rates


def _spawn(chunk, basis, threads, device, rhf_kwargs, pairs_path):
    pkg = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    env = dict(os.environ)
    env["PYTHONPATH"] = pkg + (os.pathsep + env["PYTHONPATH"] if env.get("PYTHONPATH") else "")
    p = subprocess.Popen([sys.executable, "-m", "mmd.batch", "--worker"], stdin=subprocess.PIPE,
                         stdout=subprocess.PIPE, env=env)
    p.stdin.write(json.dumps({"geoms": chunk, "basis": basis, "threads": threads, "device": device,
                              "kwargs": rhf_kwargs, "pairs": pairs_path}).encode())
    p.stdin.close()
    return p


def run_batch_procs(geometries, basis="6-31g", procs=8, threads=2, device=0, pairs_dir=None, **rhf_kwargs):
    """``run_batch`` spread over ``procs`` worker processes of ``threads``
    threads each (molecules dealt round-robin).  Call it from a process that
    has not used the GPU yet: the workers are started as new programs.
    With pairs_dir, every worker also writes the learned guess's training
    pairs of its converged molecules to pairs_dir/pairs_<k>.npz (arrays X, Y
    and mol, the molecule index of each row)."""
    if procs < 1 or threads < 1:
        raise ValueError("procs and threads must be >= 1")
    procs = min(procs, len(geometries)) or 1
    chunks = [list(range(i, len(geometries), procs)) for i in range(procs)]
    ps = [_spawn([geometries[j] for j in c], basis, threads, device, rhf_kwargs,
                 os.path.join(pairs_dir, "pairs_%d.npz" % k) if pairs_dir else None) for k, c in enumerate(chunks)]
    out = [None] * len(geometries)
    err = None
    for c, p in zip(chunks, ps):
        data = p.stdout.read()
        rc = p.wait()
        if rc != 0:
            err = err or RuntimeError("batch worker exited with status %d" % rc)
            continue
        for j, r in zip(c, json.loads(data.decode())):
            out[j] = r
    if err:
        raise err
    return out


def autotune_procs(geometries, basis="6-31g", candidates=((4, 2), (8, 2), (12, 1)), device=0, **rhf_kwargs):
    """(procs, threads) with the highest SCF/s on the calibration geometries,
    and the measured rates."""
    rates = {}
    for pt in candidates:
        t0 = time.perf_counter()
        run_batch_procs(geometries, basis, procs=pt[0], threads=pt[1], device=device, **rhf_kwargs)
        rates[pt] = len(geometries) / (time.perf_counter() - t0)
    best = max(rates, key=rates.get)
    return best, rates


def _worker_main():
    job = json.loads(sys.stdin.read())
    pairs = job.get("pairs")
    res = run_batch(job["geoms"], job["basis"], workers=job["threads"], device=job["device"],
                    keep_density=bool(pairs), **job["kwargs"])
    if pairs:
        from mmd.learned import training_pair
        X, Y, M = [], [], []
        for i, (g, r) in enumerate(zip(job["geoms"], res)):
            P = r.pop("P")
            if not r["converged"]:
                continue
            f, t = training_pair(Molecule(g, job["basis"]), P)
            X.append(f)
            Y.append(t)
            M.append(np.full(len(f), i))
        np.savez(pairs, X=np.concatenate(X), Y=np.concatenate(Y), mol=np.concatenate(M))
    sys.stdout.write(json.dumps(res))
    sys.stdout.flush()


def batch_energies(results):
    return np.array([r["energy"] for r in results])


if __name__ == "__main__" and "--worker" in sys.argv:
    _worker_main()
elif __name__ == "__main__" and "--serve" in sys.argv:
    _serve_main(int(sys.argv[sys.argv.index("--serve") + 1]), int(sys.argv[sys.argv.index("--serve") + 2]))


# ---------------------------------------------------------------------------
# Batched RHF on the device (mmd_batch_rhf): groups of molecules in one
# block-diagonal basis, one Fock build per iteration for the whole group and
# the rest of the loop in one workgroup per molecule, replayed as a HIP graph.
# ---------------------------------------------------------------------------
_TEMPLATES = {}


def _element_template(basis, Z):
    """Basis functions of one atom of element Z at the origin: (lmn, nprim,
    exps, weights) in the reference's order (mmd/molecule.py:57-88), with the
    normalised weights norm * coef of cython/basis.pxi:87-120."""
    key = (basis, Z)
    t = _TEMPLATES.get(key)
    if t is None:
        from mmd.molecule import _SYMBOLS
        mol = Molecule("\n0 1\n%s 0.0 0.0 0.0\n" % _SYMBOLS[Z], basis)
        lmn = np.array([b.shell for b in mol.bfs], dtype=np.int32).reshape(-1, 3)
        nprim = np.array([b.num_exps for b in mol.bfs], dtype=np.int32)
        exps = np.concatenate([np.asarray(b.exps, dtype=np.float64) for b in mol.bfs])
        w = np.concatenate([np.asarray(b.norm, dtype=np.float64) * np.asarray(b.coefs, dtype=np.float64)
                            for b in mol.bfs])
        t = _TEMPLATES[key] = (lmn, nprim, exps, w)
    return t


def _group_arrays(geometries, basis):
    """The arrays of many geometry strings at once, as Molecule(geometry,
    basis) would build them: per molecule (Z, xyz bohr, nocc, E_nuc, first
    function, function count), and the basis of all of them concatenated
    (origin, lmn, nprim, exps, weights), each built in one pass over the
    group's atoms from the per-element templates."""
    from mmd.molecule import BOHR_ANGSTROM, _SYMBOLS
    mols, Zs, Xs = [], [], []
    kinds = {}
    for k, g in enumerate(geometries):
        rows = [ln.split() for ln in g.split("\n") if ln.strip()]
        Z = np.array([_SYMBOLS.index(r[0]) for r in rows[1:]], dtype=np.int64)
        xyz = np.array([r[1:4] for r in rows[1:]], dtype=np.float64) / BOHR_ANGSTROM
        mols.append([Z, xyz, (int(Z.sum()) - int(rows[0][0])) // 2, 0.0])
        kinds.setdefault(tuple(Z), []).append(k)
        Zs.append(Z)
        Xs.append(xyz)
    # nuclear repulsion, molecules of one composition at once
    for Zt, ks in kinds.items():
        Z = np.array(Zt, dtype=np.float64)
        X = np.stack([mols[k][1] for k in ks])
        iu = np.triu_indices(len(Z), 1)
        d = np.linalg.norm(X[:, iu[0]] - X[:, iu[1]], axis=-1)
        e = ((Z[iu[0]] * Z[iu[1]])[None] / d).sum(axis=1)
        for k, v in zip(ks, e):
            mols[k][3] = float(v)
    Zall = np.concatenate(Zs)
    T = {int(z): _element_template(basis, int(z)) for z in np.unique(Zall)}
    parts = [T[int(z)] for z in Zall]
    nbf_atom = np.array([len(p[1]) for p in parts])
    lmn = np.concatenate([p[0] for p in parts])
    nprim = np.concatenate([p[1] for p in parts])
    exps = np.concatenate([p[2] for p in parts])
    w = np.concatenate([p[3] for p in parts])
    origin = np.repeat(np.concatenate(Xs), nbf_atom, axis=0)
    f0, a0 = 0, 0
    for m in mols:
        n = int(nbf_atom[a0:a0 + len(m[0])].sum())
        m += [f0, n]
        f0 += n
        a0 += len(m[0])
    return mols, (origin, lmn, nprim, exps, w)


def _molecule_arrays(geometry, basis):
    """(Z, xyz bohr, origin, lmn, nprim, exps, weights, nocc, E_nuc) of one
    geometry string, as Molecule(geometry, basis) would build them."""
    (m,), (origin, lmn, nprim, exps, w) = _group_arrays([geometry], basis)
    return m[0], m[1], origin, lmn, nprim, exps, w, m[2], m[3]


def _device_group(ctx, geometries, idx, basis, guess, fp32_below, conver, acc2e, use_graph, keep_density,
                  guess_model, out):
    """One group of run_batch_device on ctx: results into out[i] for i in idx."""
    mols, (origin, lmn, nprim, exps, w) = _group_arrays([geometries[i] for i in idx], basis)
    ctx.set_basis(origin, lmn, nprim, exps, w)
    ctx.set_atoms(np.concatenate([m[0] for m in mols]).astype(np.float64), np.concatenate([m[1] for m in mols]))
    ctx.set_molecules([m[5] for m in mols], [len(m[0]) for m in mols])
    ctx.set_precision(fp32_below)
    P0 = None
    # per molecule (Z, xyz, lmn of its functions, electrons) for the guesses
    gm = [(m[0], m[1], lmn[m[4]:m[4] + m[5]], 2 * m[2]) for m in mols]
    if guess == "learned":
        from mmd.learned import batch_learned_densities
        dens = batch_learned_densities(gm, basis, guess_model, ctx)
        P0 = np.concatenate([d.ravel() for d in dens])
    elif guess == "pair":
        from mmd.pairguess import batch_pair_densities
        dens = batch_pair_densities(gm, basis, guess_model, ctx)
        P0 = np.concatenate([d.ravel() for d in dens])
    elif guess == "sad":
        from mmd.learned import _element_density
        blocks = []
        for Z, _, _, nelec in gm:
            parts = [_element_density(int(z), basis) for z in Z]
            n = sum(len(b) for b in parts)
            P = np.zeros((n, n))
            o = 0
            for b in parts:
                P[o:o + len(b), o:o + len(b)] = b
                o += len(b)
            if Z.sum() != nelec:
                P *= nelec / Z.sum()
            blocks.append(P.ravel())
        P0 = np.concatenate(blocks)
    res = ctx.batch_rhf([m[2] for m in mols], [m[3] for m in mols], P0=P0, conver=conver, acc2e=acc2e,
                        use_graph=use_graph, keep_density=keep_density)
    o = 0
    for k, i in enumerate(idx):
        n = mols[k][5]
        r = {"energy": float(res[0][k]), "iterations": int(res[1][k]), "converged": bool(res[2][k]),
             "nbasis": n}
        if keep_density:
            r["P"] = res[3][o:o + n * n].reshape(n, n).copy()
        o += n * n
        out[i] = r


_DEVICE_CTXS = {}
_DEVICE_CTX_LOCK = threading.Lock()


def _device_contexts(device, count):
    """`count` contexts on `device`, kept for the process (each holds its own
    stream and device buffers)."""
    with _DEVICE_CTX_LOCK:
        lst = _DEVICE_CTXS.setdefault(device, [])
        while len(lst) < count:
            lst.append(_native.Context(device))
        return lst[:count]


def run_batch_device(geometries, basis="6-31g", group=64, guess="core", fp32_below=0.0, conver=1e-8,
                     acc2e=1e-12, device=0, use_graph=True, keep_density=False, guess_model=None, ctx=None,
                     streams=2, class_streams=None, sort_by_size=True):
    """Direct RHF energies of all geometries, `group` molecules at a time in
    one block-diagonal basis on the GPU (mmd_batch_rhf).  guess: 'core' (the
    reference's), 'sad', 'learned' (the monomer network) or 'pair' (the
    atom-pair networks); initial densities formed for the whole group on the
    host side, the networks on the GPU.  fp32_below: the FP32
    tier of the Fock builds.  streams: groups in flight at once, each on its
    own context (stream), so that one group's host-side set-up overlaps
    another's device work; with ctx given, that context alone.
    class_streams: streams over which each context spreads the classes of
    a Fock build (default: the context's setting).  sort_by_size: group the
    molecules in order of their atom counts (a group iterates until its
    slowest molecule has converged and steps at the pace of its largest,
    so groups of similar molecules waste less).  Returns one
    dict per geometry: energy, iterations, converged, nbasis (and the density
    P with keep_density)."""
    if group < 1:
        raise ValueError("group must be >= 1")
    if guess not in ("core", "sad", "learned", "pair"):
        raise ValueError("guess must be 'core', 'sad', 'learned' or 'pair'")
    if guess == "learned":   # the network's weights once, before any worker needs them
        from mmd.learned import default_model
        guess_model = guess_model or default_model()
    elif guess == "pair":
        from mmd.pairguess import default_model as default_pair_model
        guess_model = guess_model or default_pair_model()
    order = list(range(len(geometries)))
    if sort_by_size:
        natoms = [sum(1 for ln in g.split("\n")[1:] if ln.strip()) for g in geometries]
        order.sort(key=lambda i: natoms[i])
    groups = [order[g0:g0 + group] for g0 in range(0, len(order), group)]
    ctxs = [ctx] if ctx is not None else _device_contexts(device, max(1, min(int(streams), len(groups))))
    if class_streams is not None:
        for c in ctxs:
            c.set_streams(int(class_streams))
    out = [None] * len(geometries)
    args = (basis, guess, fp32_below, conver, acc2e, use_graph, keep_density, guess_model, out)
    if len(ctxs) == 1:
        for idx in groups:
            _device_group(ctxs[0], geometries, idx, *args)
        return out
    # groups are taken from the end of the list: with sort_by_size the largest
    # molecules first, so that the groups that finish last are the short ones
    todo = list(groups) if sort_by_size else list(reversed(groups))
    lock = threading.Lock()
    errors = []

    def worker(c):
        while True:
            with lock:
                if not todo or errors:
                    return
                idx = todo.pop()
            try:
                _device_group(c, geometries, idx, *args)
            except BaseException as e:   # re-raised in the caller
                with lock:
                    errors.append(e)
                return

    threads = [threading.Thread(target=worker, args=(c,)) for c in ctxs]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    if errors:
        raise errors[0]
    return out
