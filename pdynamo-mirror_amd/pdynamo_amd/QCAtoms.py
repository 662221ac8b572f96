"""QC atoms of a QC/MM system, as the ABFS model sees them.

* ``QCAtomContainer``      -- the QC atom data the NB model consumes (pMolecule-1.9.0/extensions/
                              cinclude/QCAtomContainer.h:24-60): QC atoms in ascending atom order;
                              each boundary (link) atom with its QC partner, link factor and MM
                              partners.  ``QCAtomContainer_FromAtomContainer`` mirrors
                              pMolecule.QCAtomContainer.pyx:430-496 with the link distances given
                              by the caller (pDynamo takes them from its periodic table).
* ``QCMMInteractionState`` -- the QC charges and QC/MM potentials exchanged with the QC model
                              (pMolecule-1.9.0/extensions/csource/QCMMInteractionState.c:16-80).
* ``QCModelPointCharges``  -- a minimal QC model: fixed charges on the QC atoms, energy
                              2625.5 sum_q Q_q phi_q.  It stands in for a QC program (pDynamo's QC
                              models are outside this build) so a System's QC/MM flow runs end to end.
"""

import numpy as np

UNITS_ENERGY_HARTREES_TO_KILOJOULES_PER_MOLE = 2625.5      # pCore-1.9.0/pCore/Units.py:34


class QCAtomContainer:
    """QC atoms: ``index`` (ascending), ``boundary`` flags, ``qcPartner`` and ``linkFactor`` of each
    boundary atom, its ``mmPartners``, and ``useLinkRatio``."""

    def __init__(self, index, boundary=None, qcPartner=None, linkFactor=None, mmPartners=None, useLinkRatio=False):
        self.index = np.asarray(index, np.int32)
        n = self.index.size
        if n == 0 or np.any(np.diff(self.index) <= 0):
            raise ValueError("QC atoms must be a non-empty ascending list of atom indices.")
        self.boundary = np.zeros(n, bool) if boundary is None else np.asarray(boundary, bool)
        self.qcPartner = -np.ones(n, np.int32) if qcPartner is None else np.asarray(qcPartner, np.int32)
        self.linkFactor = np.zeros(n) if linkFactor is None else np.asarray(linkFactor, np.float64)
        self.mmPartners = [[] for _ in range(n)] if mmPartners is None else [sorted(int(p) for p in m) for m in mmPartners]
        self.useLinkRatio = bool(useLinkRatio)
        if not (self.boundary.size == self.qcPartner.size == self.linkFactor.size == len(self.mmPartners) == n):
            raise ValueError("QC atom arrays differ in length.")

    @property
    def size(self):
        return int(self.index.size)

    @property
    def nboundary(self):
        return int(self.boundary.sum())

    def PureSelection(self):
        """QCAtomContainer_MakePureSelection: the QC atoms that are not boundary atoms."""
        return self.index[~self.boundary]

    def FullSelection(self):
        return self.index.copy()

    def Arrays(self):
        """(index, boundary, qcPartner, linkFactor, mmPartner offsets, mmPartners) for the native library."""
        off = np.concatenate([[0], np.cumsum([len(m) for m in self.mmPartners])]).astype(np.int32)
        flat = np.array([p for m in self.mmPartners for p in m], np.int32)
        return (self.index, self.boundary.astype(np.int32), self.qcPartner, self.linkFactor, off, flat)


def QCAtomContainer_FromAtomContainer(qcSelection, boundaryAtoms=None, linkFactors=None, useLinkRatio=False):
    """QC atoms from a QC selection and boundary atoms (pMolecule.QCAtomContainer.pyx:430-496).

    boundaryAtoms maps a boundary (MM) atom to (boundary partners, MM partners, QC partners) as pDynamo's
    connectivity analysis makes it; a boundary atom may have one QC partner only.  linkFactors maps each
    boundary atom to its link factor: the link-atom distance, or the ratio of distances with
    useLinkRatio."""
    boundaryAtoms = {} if boundaryAtoms is None else boundaryAtoms
    linkFactors = {} if linkFactors is None else linkFactors
    atoms = sorted(set(int(i) for i in qcSelection) | set(int(b) for b in boundaryAtoms))
    bnd, part, lf, mmp = [], [], [], []
    for a in atoms:
        if a in boundaryAtoms:
            bp, mp, qp = boundaryAtoms[a]
            qp = list(qp)
            if len(qp) != 1:
                raise ValueError("A QC/MM boundary atom - {:d} - has multiple QC partners - {!r}.".format(a, qp))
            if a not in linkFactors:
                raise ValueError("No link factor for boundary atom {:d}.".format(a))
            bnd.append(True)
            part.append(int(qp[0]))
            lf.append(float(linkFactors[a]))
            mmp.append(sorted(set(int(p) for p in bp) | set(int(p) for p in mp)))
        else:
            bnd.append(False)
            part.append(-1)
            lf.append(0.0)
            mmp.append([])
    return QCAtomContainer(atoms, bnd, part, lf, mmp, useLinkRatio)


class QCMMInteractionState:
    """QC charges (set by the QC model) and QC/MM potentials (set by the NB model), atomic units."""

    def __init__(self, n):
        self.qcCharges = np.zeros(n)
        self.qcmmPotentials = np.zeros(n)

    @classmethod
    def WithExtent(cls, n):
        return cls(n)

    def Initialize(self):
        self.qcCharges[:] = 0.0
        self.qcmmPotentials[:] = 0.0


class QCModelPointCharges:
    """Fixed point charges on the QC atoms in place of a QC program: sets the QC charges and returns
    2625.5 sum_q Q_q phi_q as "QC/MM Elect."."""

    label = "QC Point Charges"

    def __init__(self, charges):
        self.charges = np.asarray(charges, np.float64)

    def Energy(self, qcAtoms, configuration):
        state = configuration.qcmmstate
        if self.charges.size != state.qcCharges.size:
            raise ValueError("one charge per QC atom is needed")
        state.qcCharges[:] = self.charges
        e = UNITS_ENERGY_HARTREES_TO_KILOJOULES_PER_MOLE * float(np.dot(self.charges, state.qcmmPotentials))
        return [("QC/MM Elect.", e)]
