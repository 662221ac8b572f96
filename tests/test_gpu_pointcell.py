"""Point/cell grid search (PairListGenerator useGridByCell False) on the HIP path, through the C ABI.

The pair sets are compared bit for bit with the reference library's (tests/golden/pointcell_*.npz, written by
make_golden.py) and with the oracle's point/cell restatement: a water lattice at two cell sizes, a box whose
extent is a whole number of cells so that atoms on its upper faces fall off the reference grid (they still search
as points, around the cell decoded before them), and a moving sequence with the half-skin rebuilds.  Energies and
gradients are checked against the oracle with the tolerances of the other parity tests (parity_checks.check).
The GridUpdating.py analogue runs eleven point/cell generators (cell size factors 0.25 ... 2.75) on one system and
requires the oracle's set for each.
"""

import os

import numpy as np
import pytest

from parity_checks import check
from test_oracle import POINTCELL_FIXTURES, _Sys, _keys

pytestmark = pytest.mark.gpu

PATHS = [("mixed", 1e-6, 1e-5), ("tiled_fp64", 1e-10, 1e-10), ("fp64", 1e-12, 1e-12)]


def _system(mm, factor, precision="mixed", cutoff=13.5, min_points=500, min_extent=2):
    from pdynamo_amd import NBModelABFS, PairListGenerator, System
    gen = PairListGenerator(cutoff=cutoff, cutoffCellSizeFactor=factor, useGridByCell=False, sortIndices=False,
                            minimumPoints=min_points, minimumCellExtent=min_extent)
    s = System.FromSynthetic(mm)
    s.DefineNBModel(NBModelABFS(generator=gen, precision=precision))
    return s


def _run(s, xyz):
    s.coordinates3 = xyz
    s.Energy(log=None, doGradients=True)
    return s.configuration.energyTerms.Terms(asDictionary=True), s.configuration.gradients3.copy()


@pytest.mark.parametrize("precision,etol,gtol", PATHS)
@pytest.mark.parametrize("path", POINTCELL_FIXTURES, ids=lambda p: os.path.basename(p)[:-4])
def test_point_cell_fixtures(oracle_mod, path, precision, etol, gtol):
    fx = dict(np.load(path, allow_pickle=False))
    print("point/cell fixture", os.path.basename(path), precision, flush=True)
    mm = _Sys(fx)
    o = fx["options"]
    assert not bool(fx["grid_by_cell"])
    s = _system(mm, float(o[6]), precision, cutoff=float(o[3]))
    orc = oracle_mod.OracleABFS(mm, cell_factor=float(o[6]), use_grid_by_cell=False, nthreads=2)
    for k, xyz in enumerate(fx["coords"]):
        terms, grad = _run(s, xyz)
        rterms, rgrad, rebuilt = orc.energy(xyz)
        assert s.configuration.nbState.lastRebuilt == bool(fx["rebuilt"][k]) == rebuilt
        keys = s.configuration.nbState.Pairs()
        assert keys.size == fx["npairs"][k]
        assert np.array_equal(keys, orc.records.canonical())
        if k == 0:
            assert np.array_equal(keys, _keys(fx))
        check(terms, grad, rterms, rgrad, etol, gtol, "call %d" % k, orc.abs_sums(xyz))


def test_point_cell_generators(oracle_mod):
    """GridUpdating.py:76-86 analogue: point/cell generators with cell size factors 0.25, 0.5, ..., 2.75
    (minimumCellExtent 0, minimumPoints 100) on one water lattice each give the oracle's pair set and energies."""
    from pdynamo_amd import synthetic
    mm = synthetic.water_lattice(700, 27.0, seed=12)
    for i in range(11):
        f = 0.25 * (i + 1)
        print("point/cell factor", f, flush=True)
        s = _system(mm, f, min_points=100, min_extent=0)
        terms, grad = _run(s, mm.coordinates3)
        orc = oracle_mod.OracleABFS(mm, cell_factor=f, min_extent=0, min_points=100, use_grid_by_cell=False)
        rterms, rgrad, _ = orc.energy(mm.coordinates3)
        assert s.configuration.nbState.Pairs().size == orc.records.npairs
        assert np.array_equal(s.configuration.nbState.Pairs(), orc.records.canonical())
        check(terms, grad, rterms, rgrad, 1e-6, 1e-5, "factor %.2f" % f, orc.abs_sums(mm.coordinates3))
