"""Parity of the HIP path (through the C ABI) with the FP64 oracle.

Tolerances (parity_checks.check): mixed precision -- every energy term within
1e-6 of its own magnitude (|dE_k| <= 1e-6 * max(|E_k|, 1 kJ/mol)), the total
likewise, relative RMS gradient <= 1e-5 over all atoms and over the atoms in no
overlapping pair; tiled full FP64 -- 1e-10; plain FP64 -- the same with 1e-12.  Pair sets are compared
bit for bit as sorted canonical keys max*n+min.
"""

import numpy as np
import pytest

from parity_checks import check

pytestmark = pytest.mark.gpu

# The three GPU paths and their bounds: mixed (FP64 pair core with FP32 parts), the tiled kernel in full FP64
# (fixed-point sums round each work item's energy by up to 2^-33 kJ/mol), the plain FP64 CSR kernel.
PATHS = [("mixed", 1e-6, 1e-5), ("tiled_fp64", 1e-10, 1e-10), ("fp64", 1e-12, 1e-12)]


def _system(mm, **opts):
    from pdynamo_amd import NBModelABFS, System
    s = System.FromSynthetic(mm)
    s.DefineNBModel(NBModelABFS(**opts))
    return s


def _run(s, xyz=None):
    if xyz is not None:
        s.coordinates3 = xyz
    e = s.Energy(log=None, doGradients=True)
    terms = s.configuration.energyTerms.Terms(asDictionary=True)
    return e, terms, s.configuration.gradients3.copy()


def _check(terms, grad, rterms, rgrad, etol, gtol, label="", scales=None):
    check(terms, grad, rterms, rgrad, etol, gtol, label, scales)


@pytest.mark.parametrize("precision,etol,gtol", PATHS)
def test_water648(oracle_mod, precision, etol, gtol):
    from pdynamo_amd import synthetic
    mm = synthetic.config0()
    s = _system(mm, precision=precision)
    _, terms, grad = _run(s)
    o = oracle_mod.OracleABFS(mm)
    rterms, rgrad, _ = o.energy(mm.coordinates3)
    assert np.array_equal(s.configuration.nbState.Pairs(), o.records.canonical())
    _check(terms, grad, rterms, rgrad, etol, gtol, "", o.abs_sums(mm.coordinates3))


@pytest.mark.parametrize("natoms,box,seed", [(300, 30.0, 1), (900, 45.0, 2), (2000, 60.0, 3), (7, 10.0, 4)])
@pytest.mark.parametrize("precision,etol,gtol", PATHS)
def test_random_mixtures(oracle_mod, natoms, box, seed, precision, etol, gtol):
    from pdynamo_amd import synthetic
    mm = synthetic.random_small_system(natoms, box, seed)
    s = _system(mm, precision=precision)
    _, terms, grad = _run(s)
    o = oracle_mod.OracleABFS(mm)
    rterms, rgrad, _ = o.energy(mm.coordinates3)
    assert np.array_equal(s.configuration.nbState.Pairs(), o.records.canonical())
    _check(terms, grad, rterms, rgrad, etol, gtol, "", o.abs_sums(mm.coordinates3))


def _shaped(shape):
    """Atoms on a jittered 2.8 A lattice filling a slab, a rod, a 0.5 A thin plane, or an exact line."""
    from pdynamo_amd import synthetic
    n, ext = {"slab": (900, (60.0, 60.0, 4.0)), "rod": (600, (9.0, 9.0, 170.0)),
              "thin_plane": (700, (80.0, 0.5, 80.0)), "line": (600, (1700.0, 0.0, 0.0))}[shape]
    mm = synthetic.random_small_system(n, 40.0, 17)
    rng = np.random.default_rng(3)
    axes = [np.arange(0.0, max(e, 1e-9), 2.8) if e > 0 else np.zeros(1) for e in ext]
    pts = np.stack(np.meshgrid(*axes, indexing="ij"), -1).reshape(-1, 3)
    pts = pts[rng.permutation(len(pts))[:n]]
    e = np.array(ext)
    jit = np.where(e >= 1.0, rng.uniform(-0.3, 0.3, pts.shape), rng.uniform(0.0, 1.0, pts.shape) * e)
    mm.coordinates3 = np.ascontiguousarray(pts + jit)
    assert len(mm.coordinates3) == n
    return mm


@pytest.mark.parametrize("shape", ["slab", "rod", "thin_plane", "line"])
@pytest.mark.parametrize("precision,etol,gtol", PATHS)
def test_extreme_shapes(oracle_mod, shape, precision, etol, gtol):
    """Flat and thin systems: super-cluster columns much wider than the system in some direction,
    columns padded to whole super-clusters; same pair set and terms as the oracle.  An exact line of
    more than 500 atoms (grid method, two zero extents) is an invalid argument, as in the reference,
    whose grid then has no cells (RegularGridOccupancy.c:29,54)."""
    from pdynamo_amd import CLibraryError
    mm = _shaped(shape)
    s = _system(mm, precision=precision)
    o = oracle_mod.OracleABFS(mm)
    if shape == "line":
        with pytest.raises(CLibraryError):
            _run(s)
        with pytest.raises(ValueError):
            o.energy(mm.coordinates3)
        return
    _, terms, grad = _run(s)
    rterms, rgrad, _ = o.energy(mm.coordinates3)
    assert np.array_equal(s.configuration.nbState.Pairs(), o.records.canonical())
    _check(terms, grad, rterms, rgrad, etol, gtol, "", o.abs_sums(mm.coordinates3))


def test_options_14_dielectric(oracle_mod):
    """1-4 list with its own LJ table, scale14 and a dielectric (NBModelABFS.c:228-301)."""
    from pdynamo_amd import (LJParameterContainer_FromTableCoefficients, MMAtomContainer, NBModelABFS, System,
                             synthetic)
    mm = synthetic.random_small_system(1200, 50.0, 11)
    rng = np.random.default_rng(5)
    i = rng.integers(10, mm.natoms, 300)
    p14 = np.stack([i, i - rng.integers(3, 9, 300)], 1)
    mm.interactions14 = p14.astype(np.int32)
    a14, b14 = mm.tableA * 0.5, mm.tableB * 0.5
    lj = LJParameterContainer_FromTableCoefficients(mm.ntypes, mm.tableA, mm.tableB)
    lj14 = LJParameterContainer_FromTableCoefficients(mm.ntypes, a14, b14)
    s = System.FromMMParameters(mm.coordinates3, MMAtomContainer(mm.charges, mm.ljtypes), lj, mm.exclusions,
                                mm.interactions14, lj14)
    s.DefineNBModel(NBModelABFS(dielectric=2.5, electrostaticScale14=0.5, innerCutoff=7.0, outerCutoff=11.0,
                                listCutoff=12.5, dampingCutoff=0.75))
    _, terms, grad = _run(s)
    o = oracle_mod.OracleABFS(mm, damping=0.75, inner=7.0, outer=11.0, list_cutoff=12.5, dielectric=2.5,
                              scale14=0.5, tableA14=a14, tableB14=b14)
    rterms, rgrad, _ = o.energy(mm.coordinates3)
    assert "MM/MM 1-4 Elect." in terms
    _check(terms, grad, rterms, rgrad, 1e-6, 1e-5, "", o.abs_sums(mm.coordinates3))


def _water_with_14(nwater, box, seed):
    """Waters on a jittered lattice (no overlapping atoms: the Coulomb and 1-4 terms are not small beside the
    LJ term) plus intermolecular "1-4" pairs between atoms 2.5-5 A apart, excluded from the main list as a
    force field's 1-4 pairs are, and a 1-4 LJ table of its own (H types with non-zero 1-4 parameters)."""
    from pdynamo_amd import synthetic
    mm = synthetic.water_lattice(nwater, box, seed=seed)
    rng = np.random.default_rng(seed + 1)
    x = mm.coordinates3
    i = rng.integers(0, mm.natoms, 40 * nwater)
    j = rng.integers(0, mm.natoms, 40 * nwater)
    d = np.linalg.norm(x[i] - x[j], axis=1)
    keep = (i // 3 != j // 3) & (d > 2.5) & (d < 5.0)
    p14 = np.unique(np.stack([np.maximum(i, j), np.minimum(i, j)], 1)[keep], axis=0)[: nwater // 2]
    mm.interactions14 = p14.astype(np.int32)
    mm.exclusions = np.concatenate([mm.exclusions, mm.interactions14]).astype(np.int32)
    a14 = mm.tableA * 0.5
    b14 = mm.tableB * 0.75
    a14[1:] = [2.0e3, 1.5e2]            # (H,O) and (H,H) 1-4 rows only
    b14[1:] = [6.0e1, 8.0e0]
    return mm, a14, b14


@pytest.mark.parametrize("precision,etol,gtol", [("mixed", 1e-6, 1e-5), ("tiled_fp64", 1e-10, 1e-10),
                                                 ("fp64", 1e-12, 1e-11)])
def test_options_14_dielectric_water(oracle_mod, precision, etol, gtol):
    """Dielectric, electrostaticScale14, a 1-4 LJ table of its own and non-default cutoffs on an overlap-free
    water lattice, so that the Coulomb, 1-4 Coulomb and 1-4 LJ terms are each checked against their own
    magnitude (NBModelABFS.c:228-301).  The plain path's gradient bound is 1e-11 here: its 1-4 gradients go
    through the fixed-point sums, which round each contribution by up to 2^-33 kJ/mol/A (1.2e-10), a few
    1e-12 of this system's gradients (38 kJ/mol/A RMS per atom at dielectric 2.5)."""
    from pdynamo_amd import LJParameterContainer_FromTableCoefficients, MMAtomContainer, NBModelABFS, System
    mm, a14, b14 = _water_with_14(1000, 31.0, 41)
    lj = LJParameterContainer_FromTableCoefficients(mm.ntypes, mm.tableA, mm.tableB)
    lj14 = LJParameterContainer_FromTableCoefficients(mm.ntypes, a14, b14)
    s = System.FromMMParameters(mm.coordinates3, MMAtomContainer(mm.charges, mm.ljtypes), lj, mm.exclusions,
                                mm.interactions14, lj14, electrostaticScale14=0.5)
    s.DefineNBModel(NBModelABFS(dielectric=2.5, electrostaticScale14=0.5, innerCutoff=7.0, outerCutoff=11.0,
                                listCutoff=12.5, dampingCutoff=0.75, precision=precision))
    _, terms, grad = _run(s)
    o = oracle_mod.OracleABFS(mm, damping=0.75, inner=7.0, outer=11.0, list_cutoff=12.5, dielectric=2.5,
                              scale14=0.5, tableA14=a14, tableB14=b14)
    rterms, rgrad, _ = o.energy(mm.coordinates3)
    r = dict(rterms)
    assert all(abs(r[k]) > 1.0 for k in ("MM/MM Elect.", "MM/MM 1-4 Elect.", "MM/MM 1-4 LJ")), r
    assert np.array_equal(s.configuration.nbState.Pairs(), o.records.canonical())
    _check(terms, grad, rterms, rgrad, etol, gtol, "water1000 options " + precision,
           o.abs_sums(mm.coordinates3))


def test_rebuild_sequence(oracle_mod):
    """Rebuild decisions of the half-skin rule and energies over moving coordinates."""
    from pdynamo_amd import synthetic
    mm = synthetic.water_lattice(1000, 31.0, seed=7)
    seq = synthetic.displacement_sequence(mm, 8, 0.08, seed=3)
    s = _system(mm)
    o = oracle_mod.OracleABFS(mm)
    flags = []
    for xyz in [mm.coordinates3] + seq:
        _, terms, grad = _run(s, xyz)
        rterms, rgrad, rb = o.energy(xyz)
        flags.append((s.configuration.nbState.lastRebuilt, rb))
        _check(terms, grad, rterms, rgrad, 1e-6, 1e-5, "", o.abs_sums(xyz))
        assert np.array_equal(s.configuration.nbState.Pairs(), o.records.canonical())
    assert all(a == b for a, b in flags), flags
    assert 2 <= sum(b for _, b in flags) < len(flags)


def test_deterministic_repeat():
    """Fixed-point accumulation: identical inputs give bit-identical results."""
    from pdynamo_amd import synthetic
    mm = synthetic.random_small_system(3000, 60.0, 21)
    s = _system(mm)
    e1, t1, g1 = _run(s)
    e2, t2, g2 = _run(s)
    assert e1 == e2
    assert np.array_equal(g1, g2)


@pytest.mark.parametrize("factor", [0.25, 0.5, 1.0, 2.75])
def test_cell_size_invariance(oracle_mod, factor):
    """GridUpdating.py analogue: any cell size gives the reference set for that generator."""
    from pdynamo_amd import NBModelABFS, PairListGenerator, System, synthetic
    mm = synthetic.water_lattice(900, 28.0, seed=9)
    gen = PairListGenerator(cutoff=13.5, cutoffCellSizeFactor=factor, useGridByCell=True, sortIndices=False,
                            minimumPoints=500, minimumCellExtent=2)
    s = System.FromSynthetic(mm)
    s.DefineNBModel(NBModelABFS(generator=gen))
    _, terms, grad = _run(s)
    o = oracle_mod.OracleABFS(mm, cell_factor=factor)
    rterms, rgrad, _ = o.energy(mm.coordinates3)
    assert np.array_equal(s.configuration.nbState.Pairs(), o.records.canonical())
    _check(terms, grad, rterms, rgrad, 1e-6, 1e-5, "", o.abs_sums(mm.coordinates3))


@pytest.mark.parametrize("precision,etol,gtol", PATHS)
def test_solute30000(oracle_mod, precision, etol, gtol):
    """The 30,000-atom solute-in-water system (damped-region pairs included)."""
    from pdynamo_amd import synthetic
    mm = synthetic.config1()
    s = _system(mm, precision=precision)
    _, terms, grad = _run(s)
    o = oracle_mod.OracleABFS(mm, nthreads=8)
    rterms, rgrad, _ = o.energy(mm.coordinates3)
    assert np.array_equal(s.configuration.nbState.Pairs(), o.records.canonical())
    _check(terms, grad, rterms, rgrad, etol, gtol, "", o.abs_sums(mm.coordinates3))


def test_batch_matches_single():
    from pdynamo_amd import _native, synthetic
    from pdynamo_amd.NBModelABFS import NBModelABFS
    systems = [synthetic.random_small_system(n, b, sd) for n, b, sd in [(600, 35.0, 1), (2500, 55.0, 2), (1200, 40.0, 3)]]
    model = NBModelABFS()
    singles, batch_states = [], []
    for mm in systems:
        for bucket in (singles, batch_states):
            bucket.append(_native.NativeState(model.NativeOptions(), mm.charges, mm.ljtypes, mm.ntypes, mm.tableA,
                                              mm.tableB, exclusions=mm.exclusions))
    ref = []
    for st, mm in zip(singles, systems):
        g = np.zeros((mm.natoms, 3))
        e, _, _ = st.energy_host(mm.coordinates3, g)
        ref.append((e, g))
    xs = [_native.DeviceArray.from_numpy(mm.coordinates3) for mm in systems]
    gs = [_native.DeviceArray(x.shape) for x in xs]
    batch = _native.NativeBatch(batch_states, use_graph=True)
    for rep in range(3):
        for g in gs:
            g.zero()
        e, rebuilt = batch.energy([x.ptr for x in xs], [g.ptr for g in gs])
        assert rebuilt.all() == (rep == 0)
        for k, (er, gr) in enumerate(ref):
            assert np.array_equal(e[k], er)
            assert np.array_equal(gs[k].numpy(), gr)
    batch.close()


@pytest.mark.parametrize("precision,periodic", [("mixed", False), ("fp64", False), ("mixed", True), ("fp64", True)])
def test_kernel_timing_passes(precision, periodic):
    """nbabfs_set_timing: one timed pass per evaluated pair pass (a guarded pass that the
    update discards is not counted), with a positive mean, for both paths, with and without
    an image launch beside the main one."""
    from pdynamo_amd import _native, synthetic
    from pdynamo_amd.NBModelABFS import NBModelABFS
    if periodic:
        mm, (rot, tra), cell = synthetic.periodic_water_box(300, 21.0, 3)
    else:
        mm = synthetic.random_small_system(1500, 45.0, 8)
    st = _native.NativeState(NBModelABFS(precision=precision).NativeOptions(), mm.charges, mm.ljtypes, mm.ntypes,
                             mm.tableA, mm.tableB, exclusions=mm.exclusions)
    if periodic:
        st.set_transformations(rot, tra)
        st.set_symmetry_parameters(cell)
    x = _native.DeviceArray.from_numpy(mm.coordinates3)
    g = _native.DeviceArray(mm.coordinates3.shape)
    st.update(x.ptr)
    st.energy_async(x.ptr, g.ptr)
    st.set_timing(True)
    for _ in range(5):
        st.update(x.ptr)
        st.energy_async(x.ptr, g.ptr)
    moved = _native.DeviceArray.from_numpy(mm.coordinates3 + 1.0)    # every atom beyond the half skin
    assert st.update(moved.ptr)
    st.energy_async(moved.ptr, g.ptr)
    _native.synchronize()
    s = st.stats()
    st.set_timing(False)
    assert s["forceKernelLaunches"] == 6
    assert s["forceKernelMs"] > 0.0
    st.close()


@pytest.mark.parametrize("precision", ["mixed", "tiled_fp64", "fp64"])
def test_gradient_assign_mode(precision):
    """nbabfs_set_gradient_mode(ASSIGN): every gradient row is overwritten with the model's
    gradient, bit for bit what ADD gives on a zeroed array, for the device and host entry points
    and for a batch; ADD again adds."""
    from pdynamo_amd import _native, synthetic
    from pdynamo_amd.NBModelABFS import NBModelABFS
    systems = [synthetic.random_small_system(n, b, sd) for n, b, sd in [(900, 40.0, 5), (2500, 55.0, 6)]]
    opts = NBModelABFS(precision=precision).NativeOptions()

    def state(mm):
        return _native.NativeState(opts, mm.charges, mm.ljtypes, mm.ntypes, mm.tableA, mm.tableB,
                                   exclusions=mm.exclusions)
    ref = []
    for mm in systems:
        st = state(mm)
        g = np.zeros((mm.natoms, 3))
        e, _, _ = st.energy_host(mm.coordinates3, g)
        ref.append((e, g))
        st.close()
    for mm, (er, gr) in zip(systems, ref):
        st = state(mm)
        st.set_gradient_mode(_native.GRADIENT_ASSIGN)
        x = _native.DeviceArray.from_numpy(mm.coordinates3)
        g = _native.DeviceArray.from_numpy(np.full((mm.natoms, 3), 123.25))
        for _ in range(2):
            st.update(x.ptr)
            st.energy_async(x.ptr, g.ptr)
            e, _ = st.energies()
            assert np.array_equal(e, er)
            assert np.array_equal(g.numpy(), gr)
        h = np.full((mm.natoms, 3), -7.5)
        e, _, _ = st.energy_host(mm.coordinates3, h)
        assert np.array_equal(h, gr)
        st.set_gradient_mode(_native.GRADIENT_ADD)
        st.update(x.ptr)
        st.energy_async(x.ptr, g.ptr)
        assert np.array_equal(g.numpy(), gr + gr)
        with pytest.raises(_native.CLibraryError):
            st.set_gradient_mode(2)
        st.close()
    states = [state(mm) for mm in systems]
    for st in states:
        st.set_gradient_mode(_native.GRADIENT_ASSIGN)
    xs = [_native.DeviceArray.from_numpy(mm.coordinates3) for mm in systems]
    gs = [_native.DeviceArray.from_numpy(np.full((mm.natoms, 3), 9.0)) for mm in systems]
    batch = _native.NativeBatch(states, use_graph=True)
    for _ in range(3):
        e, _ = batch.energy([x.ptr for x in xs], [g.ptr for g in gs])
        for k, (er, gr) in enumerate(ref):
            assert np.array_equal(e[k], er)
            assert np.array_equal(gs[k].numpy(), gr)
    batch.close()


@pytest.mark.parametrize("precision", ["mixed", "tiled_fp64", "fp64"])
def test_call_orders_match(precision):
    """Update and energy calls in other orders give the same bits as update-then-energy.

    The update queues the pair kernel for its coordinates before it knows whether
    the list must be rebuilt; the cases below make that early work stale or
    unused: an energy call on another coordinate array, two updates in a row, an
    energy call with no update, and a move that forces a rebuild.
    """
    from pdynamo_amd import _native, synthetic
    from pdynamo_amd.NBModelABFS import NBModelABFS
    mm = synthetic.random_small_system(2500, 55.0, 21)
    opts = NBModelABFS(precision=precision).NativeOptions()

    def state():
        return _native.NativeState(opts, mm.charges, mm.ljtypes, mm.ntypes, mm.tableA, mm.tableB,
                                   exclusions=mm.exclusions)

    x0 = mm.coordinates3
    x1 = x0 + np.random.default_rng(4).normal(0.0, 0.05, x0.shape)      # within the half skin
    x2 = x0 + np.random.default_rng(5).normal(0.0, 1.0, x0.shape)       # forces a rebuild

    # References: the plain update-then-energy sequence through the same list
    # history (list built at x0, reused at x1, rebuilt at x2), on a state of its own.
    rs = state()
    g_ref = _native.DeviceArray(x0.shape)
    refs = {}
    for key, x in (("x0", x0), ("x1", x1), ("x2", x2)):
        d = _native.DeviceArray.from_numpy(x, rs)
        rs.update(d.ptr)
        g_ref.zero(rs)
        rs.energy_async(d.ptr, g_ref.ptr)
        refs[key] = (rs.energies()[0], g_ref.numpy(rs))
        d.free()

    st = state()
    d0 = _native.DeviceArray.from_numpy(x0, st)
    d1 = _native.DeviceArray.from_numpy(x1, st)
    d1b = _native.DeviceArray.from_numpy(x1, st)
    d2 = _native.DeviceArray.from_numpy(x2, st)
    g = _native.DeviceArray(x0.shape)

    def energy(d, key):
        g.zero(st)
        st.energy_async(d.ptr, g.ptr)
        e, _ = st.energies()
        assert np.array_equal(e, refs[key][0]), key
        assert np.array_equal(g.numpy(st), refs[key][1]), key

    assert st.update(d0.ptr)
    energy(d0, "x0")
    assert not st.update(d1.ptr)            # speculative pass for d1 ...
    energy(d1b, "x1")                       # ... energy asked on another array
    assert not st.update(d1.ptr)
    assert not st.update(d1.ptr)            # two updates, one energy
    energy(d1, "x1")
    energy(d1, "x1")                        # energy with no update
    assert st.update(d2.ptr)                # guarded pass discarded, list rebuilt
    energy(d2, "x2")


def test_autotune_keeps_the_list_and_parity(oracle_mod):
    """The work-item autotuner changes only how the list is cut: same pairs, same tolerances."""
    from pdynamo_amd import NBModelABFS, System, synthetic
    mm = synthetic.random_small_system(3000, 60.0, 31)
    model = NBModelABFS()
    s = System.FromSynthetic(mm)
    s.DefineNBModel(model)
    _, _, _ = _run(s)
    pairs0 = s.configuration.nbState.Pairs()
    choice = model.Autotune(s, log=None)
    assert choice["entriesPerItem"] in (8, 12, 16, 24, 32, 48) and choice["wavesPerCU"] in (0, 12)
    assert choice["kernelMs"] > 0.0
    _, terms, grad = _run(s)
    assert np.array_equal(s.configuration.nbState.Pairs(), pairs0)
    o = oracle_mod.OracleABFS(mm)
    rterms, rgrad, _ = o.energy(mm.coordinates3)
    _check(terms, grad, rterms, rgrad, 1e-6, 1e-5, "", o.abs_sums(mm.coordinates3))


@pytest.mark.parametrize("precision,tol", [("mixed", 2e-5), ("fp64", 1e-7)])
def test_virial_matches_strain_derivative(oracle_mod, precision, tol):
    """W_ab = sum_i x_i,a g_i,b against central differences of the FP64 oracle energy
    under homogeneous strains x -> (I + eps S) x:  dE/deps = tr(S W).  Also the atom-sum
    identity with the returned gradients (W is taken about the state's grid origin,
    which moves it by o (x) sum_i g_i ~ 0)."""
    from pdynamo_amd import synthetic
    mm = synthetic.config0()
    s = _system(mm, precision=precision)
    _, terms, grad = _run(s)
    W = s.configuration.nbState.Virial()
    x = mm.coordinates3
    Wsum = x.T @ grad
    assert np.allclose(W, Wsum, rtol=0.0, atol=tol * np.abs(Wsum).max())
    rng = np.random.default_rng(12)
    eps = 1e-5
    for S in [np.eye(3), rng.normal(size=(3, 3)), rng.normal(size=(3, 3))]:
        def energy(e):
            o = oracle_mod.OracleABFS(mm)
            t, _, _ = o.energy(x @ (np.eye(3) + e * S).T, gradients=False)
            return sum(v for _, v in t)
        fd = (energy(eps) - energy(-eps)) / (2.0 * eps)
        an = float(np.trace(S @ W))
        assert abs(an - fd) <= tol * max(abs(fd), np.abs(W).max()), (S, an, fd)


@pytest.mark.parametrize("precision,etol,gtol", [("mixed", 1e-6, 1e-5), ("tiled_fp64", None, 1e-10),
                                                 ("fp64", None, 1e-12)])
def test_dhfr(oracle_mod, dhfr, precision, etol, gtol):
    """The reference's DHFR benchmark system (CHARMM22, 23,558 atoms, 35 LJ types so the LJ rows are read from
    global memory): the published pair counts exactly; the published four terms to their 4 printed decimals
    (both FP64 paths) or to 1e-6 of the total (mixed); gradients against the oracle."""
    s = dhfr.system(precision=precision)
    _, terms, grad = _run(s)
    st = s.configuration.nbState.Statistics()
    assert st["npairs"] == dhfr.published_pairs and st["npairs14"] == dhfr.published_pairs14
    etot = sum(dhfr.published_terms.values())
    for k, v in dhfr.published_terms.items():
        tol = 5.01e-5 if etol is None else etol * abs(etot)
        assert abs(terms[k] - v) <= tol, (k, terms[k], v)
    o = oracle_mod.OracleABFS(dhfr.mm, tableA14=dhfr.tableA14, tableB14=dhfr.tableB14, scale14=dhfr.scale14,
                              nthreads=8)
    _, rgrad, _ = o.energy(dhfr.mm.coordinates3)
    assert np.linalg.norm(grad - rgrad) <= gtol * np.linalg.norm(rgrad)


def test_sgof_pool_matches_system_energy():
    """The GPU image pool (pDynamo's SGOFProcessPool interface) gives each image the energy and gradients of a
    System.Energy call on that image, bit for bit, and keeps each image's list between calls."""
    from pdynamo_amd import NBModelABFS, SGOFProcessPoolGPU, System, synthetic
    mm = synthetic.random_small_system(1500, 45.0, 5)
    s = System.FromSynthetic(mm)
    s.DefineNBModel(NBModelABFS())
    rng = np.random.default_rng(3)
    images = [mm.coordinates3 + rng.normal(0.0, 0.1, mm.coordinates3.shape) for _ in range(5)]
    pool = SGOFProcessPoolGPU.FromSystem(s)
    for rep in range(2):
        g = [np.zeros((mm.natoms, 3)) for _ in images]
        f = pool.FunctionGradients(images, g)
        assert pool.rebuilt.all() == (rep == 0)
        for k, x in enumerate(images):
            t = System.FromSynthetic(mm.with_coordinates(x))
            t.DefineNBModel(NBModelABFS())
            e = t.Energy(log=None, doGradients=True)
            assert f[k] == e, (k, f[k], e)
            assert np.array_equal(g[k], t.configuration.gradients3)
    pool.Finalize()


def test_sgof_pool_autotune():
    """SGOFProcessPoolGPU.Autotune (the batch work-item autotuner) writes no gradient, keeps the images' lists, and
    the pool's next call still gives each image the energy and gradients of a System.Energy call, up to the
    rounding of the other cut (energies 1e-9, gradients 1e-6 relative RMS)."""
    from pdynamo_amd import NBModelABFS, SGOFProcessPoolGPU, System, synthetic
    mm = synthetic.random_small_system(1500, 45.0, 6)
    s = System.FromSynthetic(mm)
    s.DefineNBModel(NBModelABFS())
    rng = np.random.default_rng(4)
    images = [mm.coordinates3 + rng.normal(0.0, 0.1, mm.coordinates3.shape) for _ in range(4)]
    pool = SGOFProcessPoolGPU.FromSystem(s)
    pool.FunctionGradients(images, None)
    choice = pool.Autotune(images)
    assert choice["entriesPerItem"] in (8, 12, 16, 20, 24, 32, 40, 48) and choice["kernelMs"] > 0.0, choice
    g = [np.zeros((mm.natoms, 3)) for _ in images]
    f = pool.FunctionGradients(images, g)
    assert not pool.rebuilt.any()
    for k, x in enumerate(images):
        t = System.FromSynthetic(mm.with_coordinates(x))
        t.DefineNBModel(NBModelABFS())
        e = t.Energy(log=None, doGradients=True)
        assert abs(f[k] - e) <= 1e-9 * abs(e), (k, f[k], e)
        ref = t.configuration.gradients3
        assert np.linalg.norm(g[k] - ref) <= 1e-6 * np.linalg.norm(ref)
    pool.Finalize()


@pytest.mark.parametrize("seed", range(10))
def test_randomized_configurations(oracle_mod, seed):
    """Seeded random systems, box shapes and options (cutoffs, dielectric, 1-4 scaling), both paths: the
    pair set bit for bit and the terms within tolerance of the oracle."""
    from pdynamo_amd import (LJParameterContainer_FromTableCoefficients, MMAtomContainer, NBModelABFS, System,
                             synthetic)
    rng = np.random.default_rng(1000 + seed)
    n = int(rng.integers(60, 3000))
    mm = synthetic.random_small_system(n, 40.0, 500 + seed)
    aspect = rng.uniform(0.3, 2.5, 3)
    mm.coordinates3 = np.ascontiguousarray(mm.coordinates3 * aspect)
    k = int(rng.integers(0, n // 4 + 1))
    i = rng.integers(8, n, k) if n > 8 else np.zeros(0, np.int64)
    mm.interactions14 = np.stack([i, i - rng.integers(3, 8, i.size)], 1).astype(np.int32).reshape(-1, 2)
    inner = float(rng.uniform(6.0, 9.0))
    outer = inner + float(rng.uniform(2.0, 4.0))
    lst = outer + float(rng.uniform(0.5, 2.0))
    damp = float(rng.uniform(0.3, 1.0))
    diel = float(rng.choice([1.0, 2.0, 4.0]))
    sc14 = float(rng.choice([1.0, 0.5, 0.8333]))
    lj = LJParameterContainer_FromTableCoefficients(mm.ntypes, mm.tableA, mm.tableB)
    o = oracle_mod.OracleABFS(mm, damping=damp, inner=inner, outer=outer, list_cutoff=lst, dielectric=diel,
                              scale14=sc14)
    rterms, rgrad, _ = o.energy(mm.coordinates3)
    for precision, etol, gtol in PATHS:
        s = System.FromMMParameters(mm.coordinates3, MMAtomContainer(mm.charges, mm.ljtypes), lj, mm.exclusions,
                                    mm.interactions14)
        s.DefineNBModel(NBModelABFS(dielectric=diel, electrostaticScale14=sc14, innerCutoff=inner, outerCutoff=outer,
                                    listCutoff=lst, dampingCutoff=damp, precision=precision))
        print("seed", seed, "n", n, precision, flush=True)
        _, terms, grad = _run(s)
        assert np.array_equal(s.configuration.nbState.Pairs(), o.records.canonical())
        _check(terms, grad, rterms, rgrad, etol, gtol, "", o.abs_sums(mm.coordinates3))
