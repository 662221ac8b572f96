"""Host-side logic of the drop-in API and the C ABI surface (CPU; no GPU needed)."""

import copy
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_nbmodel_defaults_and_generator():
    from pdynamo_amd import NBModelABFS
    m = NBModelABFS()
    assert (m.dampingCutoff, m.innerCutoff, m.outerCutoff, m.listCutoff) == (0.5, 8.0, 12.0, 13.5)
    assert m.dielectric == 1.0 and m.electrostaticScale14 == 1.0
    g = m.generator
    assert g.cutoff == 13.5 and g.cellSize == 6.75 and g.useGridByCell and not g.sortIndices
    assert g.minimumPoints == 500 and g.minimumCellExtent == 2
    m.SetOptions(listCutoff=14.0)
    assert m.generator.cutoff == 14.0 and m.generator.cellSize == 7.0


@pytest.mark.parametrize("bad", [dict(dampingCutoff=-1.0), dict(innerCutoff=0.2), dict(outerCutoff=7.0),
                                 dict(listCutoff=11.0)])
def test_nbmodel_invalid_cutoffs_raise_typeerror(bad):
    from pdynamo_amd import NBModelABFS
    with pytest.raises(TypeError):
        NBModelABFS(**bad)


def test_nbmodel_invalid_option_raises_valueerror():
    from pdynamo_amd import NBModelABFS, PairListGenerator
    with pytest.raises(ValueError):
        NBModelABFS(notAnOption=1)
    with pytest.raises(ValueError):
        PairListGenerator(cutof=1.0)
    with pytest.raises(TypeError):
        NBModelABFS(qcmmCoupling="XX")


def test_nbmodel_state_roundtrip():
    from pdynamo_amd import NBModelABFS
    m = NBModelABFS(innerCutoff=7.5, dielectric=2.0, precision="fp64")
    m2 = copy.deepcopy(m)
    assert m2.__getstate__() == m.__getstate__() or m2.innerCutoff == 7.5
    assert m2.innerCutoff == 7.5 and m2.dielectric == 2.0 and m2.precision == "fp64"
    assert m2.generator.cellSize == 6.75


def test_selfpairlist_semantics():
    from pdynamo_amd import SelfPairList
    pl = SelfPairList([[1, 5], [5, 1], [3, 3], [2, 0], [7, 6]])
    assert pl.pairs.tolist() == [[2, 0], [5, 1], [7, 6]]
    assert pl.upperBound == 8


def test_lj_table_roundtrip():
    from pdynamo_amd import (LJParameterContainer, LJParameterContainer_FromEpsilonSigma,
                             LJParameterContainer_FromTableCoefficients)
    lj = LJParameterContainer_FromEpsilonSigma(2, [0.5, 0.2], [3.0, 2.0], analyticForm="OPLS")
    t = LJParameterContainer_FromTableCoefficients(2, lj.tableA, lj.tableB, analyticForm="OPLS")
    assert np.allclose(t.epsilon, [0.5, 0.2]) and np.allclose(t.sigma, [3.0, 2.0])
    st = lj.__getstate__()
    lj2 = LJParameterContainer.__new__(LJParameterContainer)
    lj2.__setstate__(st)
    assert np.allclose(lj2.tableA, lj.tableA)
    with pytest.raises(ValueError):
        LJParameterContainer("BOGUS")


def test_energy_terms():
    from pdynamo_amd import EnergyTerms
    et = EnergyTerms()
    et.Extend([("MM/MM Elect.", -3.0), ("MM/MM LJ", 1.5), ("bad", 1)])
    et.Total()
    assert et.potentialEnergy == -1.5 and len(et.Terms()) == 2


def test_system_validates_shapes():
    from pdynamo_amd import MMAtomContainer, System, synthetic
    mm = synthetic.config0()
    with pytest.raises(ValueError):
        System.FromMMParameters(mm.coordinates3[:-1], MMAtomContainer(mm.charges, mm.ljtypes), None)


def _gpu_present():
    try:
        from pdynamo_amd import _native
        return _native.load().nbabfs_init() == _native.STATUS_CONTINUE
    except Exception:
        return False


def test_no_cpu_fallback_without_gpu():
    """The product path fails loudly when no gfx950 device is usable."""
    from pdynamo_amd import CLibraryError, NBModelABFS, System, synthetic
    if _gpu_present():
        pytest.skip("a GPU is present")
    s = System.FromSynthetic(synthetic.config0())
    s.DefineNBModel(NBModelABFS())
    with pytest.raises(CLibraryError):
        s.Energy(log=None, doGradients=True)


def _header_symbols():
    text = open(os.path.join(ROOT, "include", "nbabfs.h")).read()
    return sorted(set(re.findall(r"\b(nbabfs_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from pdynamo_amd import _native
    lib = _native.load()
    declared = _header_symbols()
    assert declared == sorted(_native.EXPORTED)
    for name in declared:
        assert hasattr(lib, name), name


def test_every_entry_point_with_arguments_has_a_prototype():
    """ctypes passes undeclared pointer arguments as C int: each exported function
    that takes arguments must have argtypes set (a 64-bit handle would be cut)."""
    from pdynamo_amd import _native
    lib = _native.load()
    text = open(os.path.join(ROOT, "include", "nbabfs.h")).read()
    for name in _native.EXPORTED:
        m = re.search(r"\b" + name + r"\s*\(([^)]*)\)", text)
        assert m, name
        args = m.group(1).strip()
        if args and args != "void":
            assert getattr(lib, name).argtypes is not None, name


def test_options_struct_matches_header():
    from pdynamo_amd import _native
    text = open(os.path.join(ROOT, "include", "nbabfs.h")).read()
    body = re.search(r"typedef struct \{(.*?)\} nbabfs_options;", text, re.S).group(1)
    fields = re.findall(r"\b(double|int)\s+(\w+);", body)
    assert [f for _, f in fields] == [f for f, _ in _native.Options._fields_]
    o = _native.default_options()
    assert (o.dampingCutoff, o.innerCutoff, o.outerCutoff, o.listCutoff) == (0.5, 8.0, 12.0, 13.5)
    assert o.cutoffCellSizeFactor == 0.5 and o.minimumPoints == 500


def test_reference_library_exports():
    import oracle
    if not oracle.ref_available():
        pytest.skip("reference library not built here")
    lib = ctypes.CDLL(oracle.ref_path())
    for name in ("ref_create", "ref_energy", "ref_pairs", "ref_use_grid", "ref_destroy"):
        assert hasattr(lib, name)
