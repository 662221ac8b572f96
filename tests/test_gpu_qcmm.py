"""Parity of the QC/MM path of the HIP library (through the C ABI) with the reference.

Against the reference pDynamo C library's outputs (tests/golden/qcmm_*.npz, written by
make_golden_qcmm.py: MM, RC and RD coupling, link distances and ratios, fixed atoms, grid and
direct search, and boxes whose extent is a whole number of cells, so that atoms fall off the
grids, under both grid searches), for both precisions:

* exactly: the QC/MM electrostatic, LJ and 1-4 lists, the MM/MM list of the MM selection (size
  and key checksum) and its 1-4 list, the QC atom coordinates, the MM sites and their charges, the
  rebuild decisions;
* the six energies (MM/MM Elect., LJ, 1-4 Elect., 1-4 LJ, QC/MM LJ, QC/MM 1-4 LJ) within 1e-6 of
  the total energy (mixed) or 1e-11 (fp64), the energy gradients within 1e-5 / 1e-10 relative RMS;
* the QC/MM potentials (FP64) within 1e-12 absolute (atomic units), the QC/MM electrostatic
  gradients within 1e-10 relative RMS (fixed-point sums, 2^-32 per addition).

A System with QC atoms and the point-charge QC model reproduces the reference's terms and the sum
of both gradients.
"""

import os

import numpy as np
import pytest

from test_oracle import QCMM_FIXTURES

pytestmark = pytest.mark.gpu

QCMM_LABELS = ("MM/MM Elect.", "MM/MM LJ", "MM/MM 1-4 Elect.", "MM/MM 1-4 LJ", "QC/MM LJ", "QC/MM 1-4 LJ")
TOL = {"mixed": (1e-6, 1e-5), "fp64": (1e-11, 1e-10)}


def _state(fx, precision):
    from pdynamo_amd import _native
    o = _native.default_options()
    o.precision = _native.MIXED if precision == "mixed" else _native.PLAIN_FP64
    o.useGridByCell = 1 if bool(fx.get("grid_by_cell", True)) else 0
    fixed = fx["fixed"] if fx["fixed"].size else None
    st = _native.NativeState(o, fx["charges"], fx["ljtypes"], int(fx["ntypes"]), fx["tableA"], fx["tableB"],
                             exclusions=fx["exclusions"], interactions14=fx["interactions14"], fixed=fixed)
    st.set_qc_atoms(fx["qc_index"], fx["qc_boundary"], fx["qc_partner"], fx["qc_linkfactor"], bool(fx["link_ratio"]),
                    fx["qc_mmp_off"], fx["qc_mmp"], int(fx["coupling"]))
    st.set_qcmm_spline(fx["spline_x"], fx["spline_y"])
    return st


def _keys20(keys):
    return (keys >> 32) * (1 << 20) + (keys & 0xFFFFFFFF)


def rel_rms(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
@pytest.mark.parametrize("path", QCMM_FIXTURES, ids=lambda p: os.path.basename(p)[:-4])
def test_qcmm_fixtures(path, precision):
    fx = dict(np.load(path, allow_pickle=False))
    print("qcmm fixture", os.path.basename(path), precision, flush=True)
    etol, gtol = TOL[precision]
    st = _state(fx, precision)
    n = fx["charges"].size
    for k, xyz in enumerate(fx["coords"]):
        g = np.zeros_like(xyz)
        e, present, rebuilt = st.energy_host(xyz, g)
        qe, qp = st.qcmm_terms()
        assert rebuilt == bool(fx["rebuilt"][k]), k
        ref = fx["energies"][k]
        got = list(e) + list(qe)
        have = [present[0], present[0], present[1], present[1], qp[0], qp[1]]
        scale = max(abs(float(np.nansum(ref))), 1.0)
        for q, label in enumerate(QCMM_LABELS):
            assert bool(have[q]) == (not np.isnan(ref[q])), label
            if have[q]:
                assert abs(got[q] - ref[q]) <= etol * scale, (label, got[q], ref[q])
        assert rel_rms(g, fx["grads_e"][k]) <= gtol
        pot = st.qcmm_potentials()
        assert np.abs(pot - fx["potentials"][k]).max() <= 1e-12, (pot, fx["potentials"][k])
        gq = np.zeros_like(xyz)
        st.qcmm_gradients(fx["qc_charges"], gq)
        assert rel_rms(gq, fx["grads_q"][k]) <= 1e-10
        if k == 0:
            for w in ("el", "lj", "el14", "lj14"):
                assert np.array_equal(_keys20(st.export_qcmm_pairs(w)), fx["keys_" + w]), w
            keys = st.export_pairs(0).astype(np.uint64)
            wts = np.arange(1, keys.size + 1, dtype=np.uint64)
            assert int(fx["mmmm_checksum"][0]) == keys.size
            assert int(fx["mmmm_checksum"][1]) == int(np.sum(keys * wts, dtype=np.uint64))
            assert np.array_equal(st.export_pairs(1), fx["keys_mmmm14"])
            qx, sx, sq = st.qcmm_sites()
            assert np.array_equal(qx, fx["qc_xyz"])
            assert np.array_equal(sx, fx["site_xyz"])
            assert np.array_equal(sq, fx["site_q"])
            s = st.stats()
            assert s["qcmmElPairs"] == fx["keys_el"].size and s["qcmmLJPairs"] == fx["keys_lj"].size
    assert n == fx["grads_e"].shape[1]
    st.close()


def test_qcmm_system_point_charges():
    """System.Energy with QC atoms, NBModelABFS(qcmmCoupling="MM") and the point-charge QC model: the
    reference's NB terms, "QC/MM Elect." = 2625.5 sum_q Q_q phi_q, and the summed gradients."""
    from pdynamo_amd import (LJParameterContainer_FromTableCoefficients, MMAtomContainer, NBModelABFS,
                             QCAtomContainer, QCModelPointCharges, System)
    path = [p for p in QCMM_FIXTURES if p.endswith("qcmm_MM_abs.npz")][0]
    fx = dict(np.load(path, allow_pickle=False))
    off = fx["qc_mmp_off"]
    qc = QCAtomContainer(fx["qc_index"], fx["qc_boundary"].astype(bool), fx["qc_partner"], fx["qc_linkfactor"],
                         [fx["qc_mmp"][off[q]:off[q + 1]] for q in range(fx["qc_index"].size)],
                         bool(fx["link_ratio"]))
    lj = LJParameterContainer_FromTableCoefficients(int(fx["ntypes"]), fx["tableA"], fx["tableB"])
    s = System.FromMMParameters(fx["coords"][0], MMAtomContainer(fx["charges"], fx["ljtypes"]), lj,
                                exclusions=fx["exclusions"], interactions14=fx["interactions14"], qcAtoms=qc)
    s.DefineNBModel(NBModelABFS(qcmmCoupling="MM", precision="fp64"))
    s.DefineQCModel(QCModelPointCharges(fx["qc_charges"]))
    s.Energy(log=None, doGradients=True)
    terms = s.configuration.energyTerms.Terms(asDictionary=True)
    for q, label in enumerate(QCMM_LABELS):
        assert abs(terms[label] - fx["energies"][0][q]) <= 1e-10 * max(abs(fx["energies"][0][q]), 1.0), label
    eq = 2625.5 * float(np.dot(fx["qc_charges"], fx["potentials"][0]))
    assert abs(terms["QC/MM Elect."] - eq) <= 1e-9 * abs(eq)
    g = s.configuration.gradients3
    assert rel_rms(g, fx["grads_e"][0] + fx["grads_q"][0]) <= 1e-10
    assert np.allclose(s.configuration.qcmmstate.qcmmPotentials, fx["potentials"][0], rtol=0, atol=1e-12)


@pytest.mark.parametrize("path", QCMM_FIXTURES, ids=lambda p: os.path.basename(p)[:-4])
def test_qcmm_device_variants(path):
    """nbabfs_qcmm_potentials_async / nbabfs_qcmm_gradients_async add, into device arrays, bit for bit
    what the host calls add into host arrays."""
    from pdynamo_amd import _native
    fx = dict(np.load(path, allow_pickle=False))
    st = _state(fx, "mixed")
    xyz = fx["coords"][0]
    st.energy_host(xyz, np.zeros_like(xyz))
    nqc = fx["qc_index"].size
    p0 = np.linspace(-1.0, 1.0, nqc)
    g0 = np.full_like(xyz, 0.25)
    ph = p0.copy()
    st.qcmm_potentials(ph)
    gh = g0.copy()
    st.qcmm_gradients(fx["qc_charges"], gh)
    dp = _native.DeviceArray.from_numpy(p0)
    dq = _native.DeviceArray.from_numpy(np.ascontiguousarray(fx["qc_charges"], dtype=np.float64))
    dg = _native.DeviceArray.from_numpy(g0)
    st.qcmm_potentials_async(dp.ptr)
    st.qcmm_gradients_async(dq.ptr, dg.ptr)
    _native.synchronize()
    assert np.array_equal(dp.numpy(), ph)
    assert np.array_equal(dg.numpy(), gh)
    st.close()


@pytest.mark.parametrize("coupling", ["MM", "RC"])
def test_qcmm_many_columns(oracle_mod, coupling):
    """A QC/MM system with more columns than one workgroup of the cross-list passes takes (the chain and 1,500
    waters, 4,517 atoms: three chunks of 2,048 columns per row): the four QC/MM lists bit for bit, the six energy
    terms, the energy gradient, the potentials and the electrostatic gradient against the QC/MM oracle."""
    import qcmm_oracle as Q
    from pdynamo_amd import _native, synthetic
    from pdynamo_amd.NBModelABFS import NBModelABFS
    mm, qd = synthetic.qcmm_chain_in_water(1500, 35.8, qc_waters=8)
    assert mm.natoms > 2 * 2048
    lf = np.where(qd["boundary"], 1.09, 0.0)
    model = NBModelABFS(qcmmCoupling=coupling)
    model.CheckPairwiseInteractions()
    cp = {"MM": _native.QCMM_MM, "RC": _native.QCMM_RC}[coupling]
    st = _native.NativeState(model.NativeOptions(), mm.charges, mm.ljtypes, mm.ntypes, mm.tableA, mm.tableB,
                             exclusions=mm.exclusions, interactions14=mm.interactions14)
    st.set_qc_atoms(qd["index"], qd["boundary"], qd["partner"], lf, False, qd["mmp_off"], qd["mmp"], cp)
    qi = model.qcmmPairwiseInteraction
    st.set_qcmm_spline(qi.splineX, qi.splineY)
    qc = Q.QCAtoms(mm.natoms, qd["index"], qd["boundary"], qd["partner"], lf, qd["mmp_off"], qd["mmp"], False)
    orc = Q.QCMMOracle(mm, qc, cp, Q.CubicSpline(qi.splineX, qi.splineY))
    xyz = mm.coordinates3
    nqc = qd["index"].size
    qcq = np.linspace(-0.8, 0.8, nqc)
    rterms, rpot, rge, rgq, _ = orc.call(xyz, qcq)
    print("qcmm many columns", coupling, mm.natoms, nqc, flush=True)
    g = np.zeros_like(xyz)
    e, present, rebuilt = st.energy_host(xyz, g)
    qe, qp = st.qcmm_terms()
    L = orc.lists
    key = 1 << 20
    assert np.array_equal(_keys20(st.export_qcmm_pairs("el")), Q.QCMMLists.keys(L.el, range(nqc), key))
    assert np.array_equal(_keys20(st.export_qcmm_pairs("el14")), Q.QCMMLists.keys(L.el14, range(nqc), key))
    assert np.array_equal(_keys20(st.export_qcmm_pairs("lj")), Q.QCMMLists.keys(L.lj, L.lj_rows, key))
    assert np.array_equal(_keys20(st.export_qcmm_pairs("lj14")), Q.QCMMLists.keys(L.lj14, L.lj_rows, key))
    got = dict(zip(QCMM_LABELS, list(e) + list(qe)))
    have = dict(zip(QCMM_LABELS, [present[0], present[0], present[1], present[1], qp[0], qp[1]]))
    etol, gtol = TOL["mixed"]
    scale = max(abs(sum(rterms.values())), 1.0)
    for label in QCMM_LABELS:
        assert bool(have[label]) == (label in rterms), label
        if label in rterms:
            assert abs(got[label] - rterms[label]) <= etol * scale, (label, got[label], rterms[label])
    assert rel_rms(g, rge) <= gtol
    assert np.abs(st.qcmm_potentials() - rpot).max() <= 1e-12
    gq = np.zeros_like(xyz)
    st.qcmm_gradients(qcq, gq)
    assert rel_rms(gq, rgq) <= 1e-10
    st.close()
