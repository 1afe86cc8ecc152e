"""Batched RHF on the device (mmd_batch_rhf): molecules of a group share one
block-diagonal basis, one Fock build per iteration and one workgroup per
molecule for the rest of the loop (DIIS, Jacobi eigensolver, density),
replayed as a HIP graph.  Through the C ABI.

Every molecule's result is its own single-molecule run's (Molecule.RHF with
direct=True, mmd/scf.py:10-80): energy within 1e-9 Eh and the same number
of iterations, for 6-31G water clusters of the batching workload (13-78
functions, odd sizes included) and for cc-pVDZ water clusters, whose d
shells run on the matrix-core kernels with survivor lists.  The FP32 tier
and the learned start converge within 1 microhartree of the FP64 batch.
"""
import numpy as np
import pytest

from geometries import water_cluster

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def native():
    from mmd import _native
    if _native.device_count() < 1:
        pytest.fail("no HIP device: the GPU tests need an MI355X")
    return _native


def _geoms(ns, seed0):
    return [water_cluster(int(n), 2.0 + 1.4 * int(n), seed=seed0 + i, perturb=True) for i, n in enumerate(ns)]


def _single(geom, basis):
    from mmd.molecule import Molecule
    m = Molecule(geom, basis)
    m.RHF(doPrint=False, direct=True)
    return m


@pytest.mark.parametrize("basis,ns,group", [("6-31g", [1, 2, 3, 2, 4, 6, 5, 1, 3], 4),
                                             ("cc-pvdz", [1, 2, 1], 3)])
def test_batch_matches_single_runs(native, basis, ns, group):
    from mmd.batch import run_batch_device
    geoms = _geoms(ns, 500)
    res = run_batch_device(geoms, basis, group=group)
    for g, r in zip(geoms, res):
        m = _single(g, basis)
        print("%s n=%d: batch %.10f (%d it)  single %.10f (%d it)" % (basis, r["nbasis"], r["energy"],
                                                                     r["iterations"], m.energy.real, m.iterations))
        assert r["converged"] and m.is_converged
        assert abs(r["energy"] - m.energy.real) < 1e-9
        assert r["iterations"] == m.iterations


def test_batch_graph_equals_stream_launches(native):
    from mmd.batch import run_batch_device
    geoms = _geoms([2, 3, 4, 2, 5], 700)
    a = run_batch_device(geoms, "6-31g", group=5, use_graph=True, keep_density=True)
    b = run_batch_device(geoms, "6-31g", group=5, use_graph=False, keep_density=True)
    for x, y in zip(a, b):
        assert x["iterations"] == y["iterations"]
        assert abs(x["energy"] - y["energy"]) < 1e-11
        assert np.max(np.abs(x["P"] - y["P"])) < 1e-10


def test_batch_full_form_within_microhartree(native):
    """FP32 tier below 1e-4 and the learned start against the FP64 core-guess
    batch: every energy within 1 microhartree."""
    from mmd.batch import run_batch_device
    geoms = _geoms([2, 3, 4, 5, 6, 3], 900)
    ref = run_batch_device(geoms, "6-31g", group=6)
    full = run_batch_device(geoms, "6-31g", group=6, guess="learned", fp32_below=1e-4)
    for r, f in zip(ref, full):
        print("FP64 %.10f (%d it)  full form %.10f (%d it)" % (r["energy"], r["iterations"], f["energy"],
                                                                f["iterations"]))
        assert f["converged"]
        assert abs(f["energy"] - r["energy"]) < 1e-6


def test_batch_sad_start_and_uneven_groups(native):
    """SAD start, groups of 2 over 5 molecules (the last one alone), two
    groups in flight and sorting off: every energy within 1e-8 Eh of the
    core-guess batch and no more iterations; the density of each molecule
    idempotent in its metric (P S P = P) and of trace n_occ."""
    from mmd.batch import run_batch_device
    from mmd.molecule import Molecule
    geoms = _geoms([3, 1, 4, 2, 5], 1300)
    ref = run_batch_device(geoms, "6-31g", group=5, conver=1e-10)
    sad = run_batch_device(geoms, "6-31g", group=2, streams=2, guess="sad", conver=1e-10, keep_density=True,
                           sort_by_size=False)
    for g, r, s in zip(geoms, ref, sad):
        print("core %.10f (%d it)  SAD %.10f (%d it)" % (r["energy"], r["iterations"], s["energy"], s["iterations"]))
        assert s["converged"]
        assert abs(s["energy"] - r["energy"]) < 1e-8
        assert s["iterations"] <= r["iterations"]
        m = Molecule(g, "6-31g")
        m.build(direct=True)
        S = np.real(m.S)
        P = s["P"]
        assert abs(np.trace(P @ S) - m.nocc) < 1e-8
        assert np.max(np.abs(P @ S @ P - P)) < 1e-6


def test_batch_input_edges(native):
    """An empty list gives no results; a molecule above the engine's basis
    limit (88 functions per molecule) is refused with a clear error."""
    from mmd.batch import run_batch_device
    assert run_batch_device([], "6-31g") == []
    with pytest.raises(Exception, match="basis functions"):
        run_batch_device([water_cluster(4, 7.0, seed=3)], "cc-pvdz", group=1)


def test_ket_slices_keep_energies_and_tuner_sets_valid_slices(native):
    """The ket slice of the row kernels only regroups the primitive sums:
    batches with slices of 1 and of 8 give the same energies (1e-10 Eh) and
    iterations as the default; the tuner sets a valid slice on every class
    and the batch with its choice gives the same energies again."""
    from mmd import _native
    from mmd.batch import run_batch_device
    from mmd.tune import tune_ket_slices
    geoms = [water_cluster(k, 2.0 + 1.4 * k, seed=300 + k, perturb=True) for k in (2, 3, 4, 5, 2, 6)]
    ctx = _native.Context(0)
    base = run_batch_device(geoms, "6-31g", guess="sad", ctx=ctx)
    for k in (1, 8):
        ctx.set_ket_slices(np.full(21, k, dtype=np.int32))
        assert list(ctx.set_ket_slices()) == [k] * 21
        got = run_batch_device(geoms, "6-31g", guess="sad", ctx=ctx)
        for a, b in zip(base, got):
            assert a["iterations"] == b["iterations"]
            assert abs(a["energy"] - b["energy"]) < 1e-10
    with pytest.raises(ValueError):
        ctx.set_ket_slices(np.full(21, 17, dtype=np.int32))
    slices, per = tune_ket_slices(geoms, "6-31g", ctx=ctx, guess="sad")
    assert list(ctx.set_ket_slices()) == list(slices)
    assert all(s in (1, 2, 4, 8) for s in slices) and set(per) == {1, 2, 4, 8}
    got = run_batch_device(geoms, "6-31g", guess="sad", ctx=ctx)
    for a, b in zip(base, got):
        assert abs(a["energy"] - b["energy"]) < 1e-10
