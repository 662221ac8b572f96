"""Parity of the periodic-image path of the HIP library (through the C ABI).

Against the reference library's own outputs (tests/golden/crystal_*.npz,
waterbox4500_moves_cells.npz and edge900_P1_*.npz, written by
make_golden_images.py; the last pair is a box whose extent is a whole number of
cells, searched cell by cell and point by cell) and the FP64
oracle: the image list (order, a b c, transformation, scale, pair count per
image) and every image's pair set exactly, the rebuild decisions of
CheckForUpdate + CheckForImageUpdate exactly, and within the mixed-precision
tolerances the six energies (|dE| <= 1e-6 |E_total|), the gradients (relative
RMS <= 1e-5) and dE/dM (largest element error <= 1e-5 of the largest
element).  The plain FP64 path (precision="fp64") meets 1e-11 relative to the
total energy, 1e-10 relative RMS gradient and 1e-9 on dE/dM: its image energies
are FP64 sums per work item added in item order, its image gradients and dE/dM
go through the fixed-point sums (resolution 2^-32 per addition).  DHFR with
its cubic cell reproduces the reference's published image numbers.
"""

import os

import numpy as np
import pytest

from test_oracle import IMAGE_FIXTURES, IMAGE_LABELS, image_fixture_keys

pytestmark = pytest.mark.gpu


TOL = {"mixed": (1e-6, 1e-5, 1e-5), "fp64": (1e-11, 1e-10, 1e-9)}


def _state(fx, precision="mixed"):
    from pdynamo_amd import _native
    o = _native.default_options()
    o.precision = _native.MIXED if precision == "mixed" else _native.PLAIN_FP64
    o.cutoffCellSizeFactor = float(fx["options"][2])
    o.useGridByCell = 1 if bool(fx.get("grid_by_cell", True)) else 0
    st = _native.NativeState(o, fx["charges"], fx["ljtypes"], int(fx["ntypes"]), fx["tableA"], fx["tableB"],
                             exclusions=fx["exclusions"], interactions14=fx["interactions14"])
    st.set_transformations(fx["rotations"], fx["translations"], bool(fx["options"][0]), int(fx["options"][1]))
    return st


def _check_call(e6, grad, dEdM, ref_e6, ref_grad, ref_dEdM, tol):
    etol, gtol, mtol = tol
    total = float(np.sum(ref_e6))
    scale = max(abs(total), 1.0)
    for q, l in enumerate(IMAGE_LABELS):
        assert abs(e6[q] - ref_e6[q]) <= etol * scale, (l, e6[q], ref_e6[q])
    rel = np.linalg.norm(grad - ref_grad) / max(np.linalg.norm(ref_grad), 1e-300)
    assert rel <= gtol, rel
    dm = np.abs(ref_dEdM).max()
    assert np.abs(dEdM - ref_dEdM).max() <= mtol * max(dm, 1.0), (dEdM, ref_dEdM)


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
@pytest.mark.parametrize("path", IMAGE_FIXTURES, ids=lambda p: os.path.basename(p)[:-4])
def test_image_fixtures(path, precision):
    fx = dict(np.load(path, allow_pickle=False))
    print("image fixture", os.path.basename(path), precision, flush=True)
    st = _state(fx, precision)
    off = 0
    keys0 = image_fixture_keys(fx)
    for k, (xyz, sp) in enumerate(zip(fx["coords"], fx["cells"])):
        st.set_symmetry_parameters(sp)
        g = np.zeros_like(xyz)
        e, present, rebuilt = st.energy_host(xyz, g)
        dm = np.zeros((3, 3))
        ie, ipresent = st.image_terms(dm)
        assert rebuilt == bool(fx["rebuilt"][k]), k
        assert [bool(present[0]), bool(present[1]), ipresent] == [bool(v) for v in fx["present"][k]]
        m = int(fx["images_per_call"][k])
        ims = fx["images"][off:off + m]
        off += m
        got = st.images()
        assert len(got) == m
        for q, (a, b, c, t, scale, npr) in enumerate(got):
            assert (a, b, c, t) == tuple(int(v) for v in ims[q, :4])
            assert scale == ims[q, 4] and npr == int(ims[q, 5])
        if k == 0:
            for q, keys in enumerate(keys0):
                assert np.array_equal(st.export_image_pairs(q), keys), q
        _check_call(list(e) + list(ie), g, dm, fx["energies"][k], fx["grads"][k], fx["dEdM"][k], TOL[precision])
    st.close()


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_dhfr_periodic(oracle_mod, dhfr, precision):
    """DHFR in its cubic cell through System.DefineSymmetry: the reference's published image list and terms
    (benchmarks/log/systemBenchmarks_Serial_1ps.log:392-404), the oracle's image pair sets, gradients and
    dE/dM."""
    etol, gtol, mtol = TOL[precision]
    m = dhfr.mm
    s = dhfr.system(periodic=True, precision=precision)
    s.Energy(log=None, doGradients=True)
    terms = s.configuration.energyTerms.Terms(asDictionary=True)
    nb = s.configuration.nbState
    ims = nb.Images()
    assert len(ims) == dhfr.published_images == 13
    assert sum(x[5] for x in ims) == dhfr.published_image_pairs == 2942151
    orc = oracle_mod.OracleABFS(m, tableA14=dhfr.tableA14, tableB14=dhfr.tableB14, scale14=dhfr.scale14,
                                transformations=dhfr.transformations)
    rterms, rgrad, rdm, _ = orc.energy_sym(m.coordinates3, dhfr.cell)
    rterms = dict(rterms)
    total = sum(rterms.values())
    assert set(terms) == set(rterms)
    for label, value in list(dhfr.published_image_terms.items()) + list(dhfr.published_terms.items()):
        assert abs(terms[label] - value) <= etol * abs(total) + 5e-5, (label, terms[label], value)
        assert abs(terms[label] - rterms[label]) <= etol * abs(total), (label, terms[label], rterms[label])
    for q in range(orc.images.nimages):
        assert tuple(ims[q][:4]) == tuple(int(v) for v in orc.images.abct[q])
        assert np.array_equal(nb.ImagePairs(q), orc.images.image_keys(q)), q
    g = s.configuration.gradients3
    rel = np.linalg.norm(g - rgrad) / np.linalg.norm(rgrad)
    assert rel <= gtol, rel
    dm = s.configuration.symmetryParameterGradients.dEdM
    assert np.abs(dm - rdm).max() <= mtol * np.abs(rdm).max(), (dm, rdm)


def test_image_rebuild_on_cell_change_only(oracle_mod):
    """A cell change alone rebuilds only the image lists, and only when CheckForImageUpdate says so; the
    main list and its reference coordinates stay (NBModelABFS_Update, NBModelABFS.c:508-623)."""
    from pdynamo_amd import synthetic
    w, tr, sp = synthetic.periodic_water_box(800, 29.0, seed=5)
    from pdynamo_amd import _native
    st = _native.NativeState(_native.default_options(), w.charges, w.ljtypes, w.ntypes, w.tableA, w.tableB,
                             exclusions=w.exclusions)
    st.set_transformations(tr[0], tr[1])
    orc = oracle_mod.OracleABFS(w, transformations=tr)
    cells = [sp, tuple([29.3] * 3 + [90.0] * 3), tuple([31.0] * 3 + [90.0] * 3), tuple([31.0] * 3 + [90.0] * 3)]
    for k, cell in enumerate(cells):
        st.set_symmetry_parameters(cell)
        g = np.zeros_like(w.coordinates3)
        e, present, rebuilt = st.energy_host(w.coordinates3, g)
        rterms, rgrad, rdm, rup = orc.energy_sym(w.coordinates3, cell)
        assert rebuilt == rup, k
        dm = np.zeros((3, 3))
        ie, _ = st.image_terms(dm)
        t = dict(rterms)
        total = sum(t.values())
        assert abs(ie[0] - t["MM/MM Image Elect."]) <= 1e-6 * abs(total)
        assert abs(ie[1] - t["MM/MM Image LJ"]) <= 1e-6 * abs(total)
        assert np.linalg.norm(g - rgrad) / np.linalg.norm(rgrad) <= 1e-5
        assert np.abs(dm - rdm).max() <= 1e-5 * np.abs(rdm).max()
    st.close()
