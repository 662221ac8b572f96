"""Write golden vectors of the QC/MM part of NBModelABFS from the REFERENCE pDynamo C library.

Like make_golden.py (build oracle/_ref first with ``make -C oracle ref``).  The
systems are a branched chain in water (pdynamo_amd.synthetic.qcmm_chain_in_water)
with six pure QC atoms and two link (boundary) atoms.  The reference state is
set up with the QC atoms (NBModelABFSState_SetUp with a QCAtomContainer), and
each call runs NBModelABFS_Update, NBModelABFS_MMMMEnergy,
NBModelABFS_QCMMEnergyLJ, NBModelABFS_QCMMPotentials and, with fixed QC
charges, NBModelABFS_QCMMGradients (ref_driver.c: ref_qcmm_energy).

The QC/MM electrostatic spline (the "Delta/Gaussian" interaction of
NBModelABFS.CheckPairwiseInteractions, width 2 bohr, atomic units) is made
from the points of oracle/qcmm_oracle.electrostatic_spline_points (the
restatement of the reference's Python integration) and built into the
reference's own CubicSpline; the fixture stores the points and the
reference's second derivatives.

Stored per case: inputs (coordinates per call, charges, LJ types and tables,
exclusions, 1-4 pairs, the QC atoms, coupling, link mode, fixed atoms, spline
points, QC charges) and per call: the six energies (MM/MM Elect., LJ, 1-4
Elect., 1-4 LJ, QC/MM LJ, QC/MM 1-4 LJ; NaN where the list is absent), the
rebuild flag, the QC/MM potentials, the gradients of the energy part and of
the QC/MM electrostatics.  For the first call: the QC/MM lists as sorted keys
(row * 2^20 + column: QC atom position or atom, MM site or atom), the MM/MM
list size and a key checksum, the QC atom coordinates and the MM sites.

Usage (in a container holding the reference tree):
    python tests/golden/make_golden_qcmm.py
"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "pdynamo-mirror_amd"))

import oracle as O  # noqa: E402
import qcmm_oracle as Q  # noqa: E402
from pdynamo_amd import synthetic as S  # noqa: E402

KEY = 1 << 20


def key_checksum(i, j, n):
    k = np.sort(np.maximum(i, j) * n + np.minimum(i, j)).astype(np.uint64)
    w = np.arange(1, k.size + 1, dtype=np.uint64)
    return np.array([k.size, int(np.sum(k * w, dtype=np.uint64))], np.uint64)


def edge_coordinates(s, first_mm, after_qc, cell=6.75):
    """Shift the system to the positive octant and move, per axis, one MM atom to the next whole number of cells
    above the largest coordinate: that atom lies on the upper face of the reference grid and so off it.  On x it is
    the MM atom after_qc that follows a pure QC atom (the chain, atoms below first_mm, is first moved next to that
    face so that the atom has pairs with it); on y and z the MM atom (index first_mm up) with the largest
    coordinate."""
    x = s.coordinates3 - s.coordinates3.min(axis=0)
    x[:first_mm, 0] += x[:, 0].max() - 4.0 - x[:first_mm, 0].max()
    x = x - x.min(axis=0)
    for d in range(3):
        top = float(np.ceil(x[:, d].max() / cell)) * cell
        k = after_qc if d == 0 else first_mm + int(np.argmax(x[first_mm:, d]))
        x[k, d] = top
    s.coordinates3 = np.ascontiguousarray(x)


def run_case(name, nwater, box, coupling, link_ratio, fixed_fn=None, seed=20261207, grid_by_cell=True, edge=False):
    s, qd = S.qcmm_chain_in_water(nwater, box, seed)
    if edge:
        last_qc = int(np.max(qd["index"]))
        edge_coordinates(s, last_qc + 1, last_qc + 1)
        # The pure QC atom before the off-grid MM atom sits in a far cell, so that its cell differs from the one
        # the reference searches around (that of the latest earlier atom of the MM selection).
        s.coordinates3[last_qc] = [1.0, 1.0, 1.0]
    lf = np.where(qd["boundary"], 1.09 / 1.53 if link_ratio else 1.09, 0.0)
    qc = Q.QCAtoms(s.natoms, qd["index"], qd["boundary"], qd["partner"], lf, qd["mmp_off"], qd["mmp"], link_ratio)
    fixed = fixed_fn(s.natoms) if fixed_fn else None
    sx, sy = Q.electrostatic_spline_points("Delta/Gaussian", 0.0, 2.0 * Q.BOHRS_TO_ANGSTROMS, 0.0, 8.0, 12.0)
    ref = Q.ReferenceQCMM(s, qc, coupling, sx, sy, fixed=fixed, use_grid_by_cell=grid_by_cell)
    qcq = np.round(np.linspace(-0.6, 0.7, qc.nqc), 3)
    rng = np.random.default_rng(seed + 1)
    x0 = s.coordinates3
    x1 = x0 + rng.normal(0.0, 0.05, x0.shape)
    x2 = x1 + rng.normal(0.0, 0.4, x0.shape)
    if fixed is not None:
        x1[fixed] = x0[fixed]
        x2[fixed] = x0[fixed]
    calls = [x0, x1, x2]
    energies, rebuilt, pots, ges, gqs = [], [], [], [], []
    first = {}
    for k, xyz in enumerate(calls):
        terms, pot, ge, gq, up = ref.call(xyz, qcq)
        energies.append([terms.get(l, np.nan) for l in Q.ReferenceQCMM.LABELS])
        rebuilt.append(up)
        pots.append(pot)
        ges.append(ge)
        gqs.append(gq)
        if k == 0:
            for w in ("el", "lj", "el14", "lj14"):
                i, j = ref.pairs(w)
                first["keys_" + w] = np.sort(i * KEY + j).astype(np.int64)
            i, j = ref.pairs("mmmm")
            first["mmmm_checksum"] = key_checksum(i, j, s.natoms)
            i, j = ref.pairs("mmmm14")
            first["keys_mmmm14"] = np.sort(np.maximum(i, j) * s.natoms + np.minimum(i, j)).astype(np.int64)
            qx, mx, mq = ref.sites()
            first.update(qc_xyz=qx, site_xyz=mx, site_q=mq)
    h = ref.spline_h()
    use_grid = O.determine_method(x0)
    ref.close()
    out = dict(name=name, coords=np.array(calls), charges=s.charges, ljtypes=s.ljtypes, ntypes=s.ntypes,
               tableA=s.tableA, tableB=s.tableB, exclusions=s.exclusions, interactions14=s.interactions14,
               qc_index=qd["index"], qc_boundary=qd["boundary"].astype(np.int32), qc_partner=qd["partner"],
               qc_linkfactor=lf, qc_mmp_off=qd["mmp_off"], qc_mmp=qd["mmp"], link_ratio=int(link_ratio),
               coupling=int(coupling), fixed=(fixed if fixed is not None else np.zeros(0, bool)).astype(np.uint8),
               spline_x=sx, spline_y=sy, spline_h=h, qc_charges=qcq, energies=np.array(energies),
               rebuilt=np.array(rebuilt), potentials=np.array(pots), grads_e=np.array(ges), grads_q=np.array(gqs),
               use_grid=use_grid, grid_by_cell=grid_by_cell, **first)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
    print(name, s.natoms, "grid" if use_grid else "direct", "rebuilt", rebuilt,
          {w: first["keys_" + w].size for w in ("el", "lj", "el14", "lj14")}, "mmmm", int(first["mmmm_checksum"][0]))


def fixed_waters(n):
    f = np.zeros(n, bool)
    f[17:17 + 3 * 150] = True         # 150 waters
    f[[4, 9, 15]] = True              # a pure QC atom and two MM atoms of the chain
    return f


def main():
    if not O.ref_available():
        O.build()
    run_case("qcmm_MM_abs", 500, 25.0, 0, False)
    run_case("qcmm_RC_ratio", 500, 25.0, 1, True)
    run_case("qcmm_RD_fixed", 500, 25.0, 2, False, fixed_waters)
    run_case("qcmm_MM_direct", 100, 15.0, 0, True)
    pointcell_cases()


def pointcell_cases():
    """Boxes whose extent is a whole number of cells, with QC atoms: under the point/cell generator, MM atoms on the
    grid's upper faces search around the cell of the latest earlier atom of the MM selection (the pure QC atoms
    decode no cell); under both generators such atoms are in no cell of the cross lists' grids."""
    run_case("qcmm_MM_pointcell_edge", 500, 25.0, 0, False, grid_by_cell=False, edge=True)
    run_case("qcmm_MM_cellcell_edge", 500, 25.0, 0, False, grid_by_cell=True, edge=True)
    run_case("qcmm_RC_pointcell_edge", 500, 25.0, 1, True, grid_by_cell=False, edge=True)


if __name__ == "__main__":
    if sys.argv[1:] == ["pointcell"]:
        pointcell_cases()
    else:
        main()
