"""Per-class tile tuner of the Fock-build row kernels (no reference
counterpart).

The row kernels' sliced classes give each lane item a slice of a ket pair's
primitive pairs; the slice length trades lane balance (short slices even out
batches that mix 1- and 81-primitive kets) against per-item work (each item
adds its partial integrals to the batch buffer or digests them).  The best
length depends on the class and on the molecules.  ``tune_ket_slices`` times
one representative batch with per-class HIP events once per candidate length
(every class set to it), in ascending and then descending order, and keeps
for each class the length with the smallest summed event time per build.
"""
import numpy as np

CANDIDATES = (1, 2, 4, 8)


def tune_ket_slices(geometries, basis="6-31g", candidates=CANDIDATES, ctx=None, guess="learned", others=()):
    """Choose the ket slice of every quartet class on one batch of
    ``geometries`` (solved as one group on ``ctx``, the first device context
    by default).  The choice is set on ctx and on every context in
    ``others``.  Returns (slices (21,), {candidate: per-class ms per build})."""
    from mmd.batch import _device_contexts, run_batch_device
    ctx = ctx or _device_contexts(0, 1)[0]
    before = ctx.set_ket_slices()
    kw = dict(group=len(geometries), guess=guess, ctx=ctx)
    run_batch_device(geometries, basis, **kw)           # contexts, tables and graphs warm
    total = {k: np.zeros(21) for k in candidates}
    builds = np.zeros(21)
    order = list(candidates) + list(reversed(candidates))
    try:
        for k in order:
            ctx.set_ket_slices(np.full(21, k, dtype=np.int32))
            ctx.set_profiling(True)
            run_batch_device(geometries, basis, **kw)
            _, _, _, ms, nb = ctx.fock_profile_sums()
            ctx.set_profiling(False)
            total[k][:len(nb)] += ms[:, 0]
            builds[:len(nb)] = np.maximum(builds[:len(nb)], nb)
    except BaseException:
        ctx.set_profiling(False)
        ctx.set_ket_slices(before)
        raise
    per_build = {k: v / np.maximum(builds, 1) / 2.0 for k, v in total.items()}
    stack = np.stack([per_build[k] for k in candidates])
    best = np.array([candidates[int(np.argmin(stack[:, qc]))] if builds[qc] > 0 else before[qc]
                     for qc in range(21)], dtype=np.int32)
    for c in [ctx] + list(others):
        c.set_ket_slices(best)
    return best, per_build
