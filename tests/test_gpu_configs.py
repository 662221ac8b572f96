"""Full-size parity of the benchmark configurations with the FP64 oracle, per system and per call.

* 200,001-atom water box: 10 successive E+grad calls after seeded cumulative N(0, 0.05 A) moves, list rebuilt by
  the half-skin rule (NBModelABFS_Update, NBModelABFS.c:508-623,691-746); both GPU paths.
* 64 water droplets of 500-20,000 atoms evaluated as one batch, each droplet against the oracle.
* 16 perturbed images of the 30,000-atom solute-in-water system evaluated as one batch, each image against the
  oracle, then again after a move that makes only some images cross the half skin (as a chain-of-states
  optimizer evaluates its images through SGOFProcessPool.FunctionGradients, ChainOfStatesOptimizer.py:168-184).

Per call and system: the rebuild decision equals the oracle's, the pair set is bit for bit the oracle's
(sorted canonical keys max*n+min) on every call that built a list, every energy term is within 1e-6 (mixed) /
1e-10 (tiled full FP64: the fixed-point sums round each work item's energy by up to 2^-33 kJ/mol) / 1e-12
(plain FP64) of its own magnitude, or within the FP64 reordering bound 4u sum|e_pair| of the oracle's pair
terms where that is larger, and the relative RMS gradient within 1e-5 / 1e-10 / 1e-12 (parity_checks.check).
All three GPU paths run on every configuration.
"""

import os

import numpy as np
import pytest

from parity_checks import check

pytestmark = pytest.mark.gpu

NTHREADS = max(1, min(16, len(os.sched_getaffinity(0))))
TOL = {"mixed": (1e-6, 1e-5), "tiled_fp64": (1e-10, 1e-10), "fp64": (1e-12, 1e-12)}
PRECISIONS = ("mixed", "tiled_fp64", "fp64")
LABELS = ("MM/MM Elect.", "MM/MM LJ", "MM/MM 1-4 Elect.", "MM/MM 1-4 LJ")


def _terms(e, has_main, has_14):
    """pDynamo's labels: an empty list contributes no terms (NBModelABFSState.pyx:41-49)."""
    out = {}
    if has_main:
        out[LABELS[0]], out[LABELS[1]] = float(e[0]), float(e[1])
    if has_14:
        out[LABELS[2]], out[LABELS[3]] = float(e[2]), float(e[3])
    return out


def _state(mm, precision):
    from pdynamo_amd import _native
    from pdynamo_amd.NBModelABFS import NBModelABFS
    return _native.NativeState(NBModelABFS(precision=precision).NativeOptions(), mm.charges, mm.ljtypes, mm.ntypes,
                               mm.tableA, mm.tableB, exclusions=mm.exclusions, interactions14=mm.interactions14)


@pytest.mark.timeout(900)
def test_water200001_moves(oracle_mod):
    from pdynamo_amd import _native, synthetic
    mm = synthetic.config2()
    seq = [mm.coordinates3] + synthetic.displacement_sequence(mm, 10, 0.05, seed=7)
    o = oracle_mod.OracleABFS(mm, nthreads=NTHREADS)
    states = {p: _state(mm, p) for p in PRECISIONS}
    g = _native.DeviceArray(mm.coordinates3.shape)
    rebuilds = 0
    for call, xyz in enumerate(seq):
        rterms, rgrad, rb = o.energy(xyz)
        scales = o.abs_sums(xyz)
        rebuilds += int(rb)
        x = _native.DeviceArray.from_numpy(xyz)
        for precision, st in states.items():
            label = "water200001 call %d %s" % (call, precision)
            print(label, flush=True)
            rebuilt = st.update(x.ptr)
            g.zero(st)
            st.energy_async(x.ptr, g.ptr)
            e, present = st.energies()
            assert rebuilt == rb, (label, rebuilt, rb)
            if rb:
                assert np.array_equal(st.export_pairs(0), o.records.canonical()), label
            check(_terms(e, present[0], present[1]), g.numpy(st), rterms, rgrad, *TOL[precision], label=label,
                  scales=scales)
        x.free()
    assert rebuilds >= 2, rebuilds         # the first build and at least one rebuild by the half-skin rule
    for st in states.values():
        st.close()


class _Batches:
    """One batch per precision over the same systems, with device coordinate and gradient arrays that persist
    across calls (the graph of a batch stays valid while its arrays and lists do)."""

    def __init__(self, systems):
        from pdynamo_amd import _native
        self.systems = systems
        self.offs = np.concatenate([[0], np.cumsum([m.natoms for m in systems])]).astype(np.int64)
        self.states, self.batches, self.X, self.G = {}, {}, {}, {}
        for precision in PRECISIONS:
            sts = [_state(m, precision) for m in systems]
            for s in sts:
                s.set_gradient_mode(_native.GRADIENT_ASSIGN)     # the batch writes every gradient row
            self.states[precision] = sts
            self.batches[precision] = _native.NativeBatch(sts, use_graph=True)
            self.X[precision] = _native.DeviceArray((int(self.offs[-1]), 3))
            self.G[precision] = _native.DeviceArray((int(self.offs[-1]), 3))

    def call(self, precision, xyzs):
        X, G = self.X[precision], self.G[precision]
        X.upload(np.concatenate(xyzs))
        e, rebuilt = self.batches[precision].energy([X.ptr + 24 * int(a) for a in self.offs[:-1]],
                                                    [G.ptr + 24 * int(a) for a in self.offs[:-1]])
        return e, rebuilt, G.numpy()

    def close(self):
        for precision in self.batches:
            self.batches[precision].close()
            for s in self.states[precision]:
                s.close()
            self.X[precision].free()
            self.G[precision].free()


def _batch_step(B, oracles, xyzs, name):
    """One oracle call per system, then one batched E+grad per precision checked against it system by system;
    returns the rebuilt flags (the same for both precisions)."""
    refs = []
    for o, xyz in zip(oracles, xyzs):
        rterms, rgrad, rb = o.energy(xyz)
        refs.append((rterms, rgrad, rb, o.records.canonical() if rb else None, o.abs_sums(xyz)))
    flags = None
    for precision in PRECISIONS:
        e, rebuilt, grads = B.call(precision, xyzs)
        for k, (mm, (rterms, rgrad, rb, keys, scales)) in enumerate(zip(B.systems, refs)):
            label = "%s system %d (%d atoms) %s" % (name, k, mm.natoms, precision)
            assert bool(rebuilt[k]) == rb, (label, rebuilt[k], rb)
            st = B.states[precision][k]
            if rb:
                assert np.array_equal(st.export_pairs(0), keys), label
            stats = st.stats()
            check(_terms(e[k], stats["npairs"] > 0, stats["npairs14"] > 0), grads[B.offs[k]:B.offs[k + 1]], rterms,
                  rgrad, *TOL[precision], label=label, scales=scales)
        flags = rebuilt
    return flags


@pytest.mark.timeout(900)
def test_droplets64_batch(oracle_mod):
    from pdynamo_amd import synthetic
    systems = synthetic.config3()
    oracles = [oracle_mod.OracleABFS(m, nthreads=NTHREADS) for m in systems]
    xyzs = [m.coordinates3 for m in systems]
    B = _Batches(systems)
    try:
        assert _batch_step(B, oracles, xyzs, "droplets64").all()
        assert not _batch_step(B, oracles, xyzs, "droplets64 repeat").any()
        for precision in PRECISIONS:
            info = B.batches[precision].info()
            assert info["graph_captures"] == 1 and info["launches"] == 2, info
    finally:
        B.close()


@pytest.mark.timeout(900)
def test_images16_batch_partial_rebuild(oracle_mod):
    from pdynamo_amd import synthetic
    systems = synthetic.config4()
    oracles = [oracle_mod.OracleABFS(m, nthreads=NTHREADS) for m in systems]
    rng = np.random.default_rng(11)
    movers = {3, 11}
    moved = [m.coordinates3 + rng.normal(0.0, 0.5 if k in movers else 0.02, m.coordinates3.shape)
             for k, m in enumerate(systems)]
    B = _Batches(systems)
    try:
        assert _batch_step(B, oracles, [m.coordinates3 for m in systems], "images16").all()
        for precision in PRECISIONS:
            info = B.batches[precision].info()
            assert info["graph_captures"] == 1 and info["launches"] == 1, info
        rebuilt = _batch_step(B, oracles, moved, "images16 moved")
        assert set(np.flatnonzero(rebuilt)) == movers, np.flatnonzero(rebuilt)
        for precision in ("mixed", "fp64"):
            # the rebuilt images changed the batch's layout: a new graph was captured and replayed
            info = B.batches[precision].info()
            assert info["graph_captures"] == 2 and info["launches"] == 2 and info["layouts"] == 2, info
    finally:
        B.close()


def test_uneven_batch_partial_rebuild(oracle_mod):
    """Five uneven systems (one small enough for the direct method) through one batch: all built at the first
    call; then two of them cross the half skin while the others move within it; then a repeat with no move.
    Every system is checked against its own oracle state at every call, and the batch re-captures its graph
    exactly when its layout changed."""
    from pdynamo_amd import synthetic
    systems = [synthetic.random_small_system(n, b, sd) for n, b, sd in
               [(400, 25.0, 1), (2500, 55.0, 2), (1200, 40.0, 3), (3000, 60.0, 4), (900, 38.0, 5)]]
    oracles = [oracle_mod.OracleABFS(m, nthreads=NTHREADS) for m in systems]
    rng = np.random.default_rng(5)
    movers = {1, 4}
    moved = [m.coordinates3 + rng.normal(0.0, 0.6 if k in movers else 0.01, m.coordinates3.shape)
             for k, m in enumerate(systems)]
    B = _Batches(systems)
    try:
        assert _batch_step(B, oracles, [m.coordinates3 for m in systems], "uneven5").all()
        rebuilt = _batch_step(B, oracles, moved, "uneven5 moved")
        assert set(np.flatnonzero(rebuilt)) == movers, np.flatnonzero(rebuilt)
        assert not _batch_step(B, oracles, moved, "uneven5 repeat").any()
        for precision in PRECISIONS:
            info = B.batches[precision].info()
            assert info == {"graph_captures": 2, "launches": 3, "layouts": 2}, info
    finally:
        B.close()


@pytest.mark.timeout(600)
def test_batch_autotune(oracle_mod):
    """The batch work-item autotuner (nbabfs_batch_autotune) re-cuts every state of the batch alike and keeps their
    lists: the pair sets are unchanged, no gradient array is written by it, the states' call counts and the batch's
    launch count are as before, and the next batched call matches every system's oracle (both tiled paths)."""
    from pdynamo_amd import synthetic
    systems = [synthetic.random_small_system(n, b, sd) for n, b, sd in [(600, 35.0, 7), (2500, 55.0, 8), (1200, 40.0, 9)]]
    oracles = [oracle_mod.OracleABFS(m, nthreads=NTHREADS) for m in systems]
    B = _Batches(systems)
    try:
        xyzs = [m.coordinates3 for m in systems]
        assert _batch_step(B, oracles, xyzs, "before autotune").all()
        for precision in ("mixed", "tiled_fp64"):
            batch, sts, G = B.batches[precision], B.states[precision], B.G[precision]
            pairs = [s.export_pairs(0) for s in sts]
            calls = [s.stats()["numberOfCalls"] for s in sts]
            launches = batch.info()["launches"]
            G.upload(np.full((int(B.offs[-1]), 3), 7.0))
            epi, ms = batch.autotune([B.X[precision].ptr + 24 * int(a) for a in B.offs[:-1]])
            print("batch autotune", precision, epi, ms, flush=True)
            assert epi in (8, 12, 16, 20, 24, 32, 40, 48) and ms > 0.0, (epi, ms)
            assert np.all(G.numpy() == 7.0)
            assert [s.stats()["numberOfCalls"] for s in sts] == calls
            assert batch.info()["launches"] == launches
            for s, p in zip(sts, pairs):
                assert np.array_equal(s.export_pairs(0), p)
        assert not _batch_step(B, oracles, xyzs, "after autotune").any()
    finally:
        B.close()
