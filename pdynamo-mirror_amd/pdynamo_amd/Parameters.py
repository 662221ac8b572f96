"""MM parameter containers on the ABFS path.

* ``LJParameterContainer`` -- packed lower-triangular A/B tables
  (pMolecule-1.9.0/extensions/cinclude/LJParameterContainer.h:16-23) built
  from epsilon/sigma in the AMBER or OPLS analytic form or taken as a TABLE
  (pMolecule-1.9.0/extensions/pyrex/pMolecule.LJParameterContainer.pyx:1-293;
  MakeTable, pMolecule-1.9.0/extensions/csource/LJParameterContainer.c:184-210).
* ``MMAtomContainer`` -- per-atom charge and LJ type
  (pMolecule-1.9.0/extensions/cinclude/MMAtomContainer.h:24-35), stored SoA.
* ``SelfPairList`` -- exclusion and 1-4 lists built from integer pairs with the
  semantics of PairList_FromIntegerPairArray
  (pCore-1.9.0/extensions/csource/PairList.c:107-189): i > j, sorted,
  duplicates and self pairs dropped.
* Units: the Coulomb factor of pCore-1.9.0/extensions/cinclude/Units.h:48-49.
"""

import math

import numpy as np

_AMBER, _OPLS, _TABLE = "AMBER", "OPLS", "TABLE"
_FORMS = (_AMBER, _OPLS, _TABLE)
UNITS_ENERGY_KILOCALORIES_PER_MOLE_TO_KILOJOULES_PER_MOLE = 4.184


def coulomb_factor():
    """1e7 N_A e^2 / (4 pi eps0) in kJ mol^-1 A e^-2 (Units.h:18-30,48-49)."""
    na, e, eps0 = 6.0221415e+23, 1.60217653e-19, 8.854187817e-12
    return (1.0e+7 * na * e * e) / (4.0e+00 * math.pi * eps0)


UNITS_ENERGY_E2ANGSTROMS_TO_KILOJOULES_PER_MOLE = coulomb_factor()


class LJParameterContainer:
    """Lennard-Jones parameters as packed A (r^-12) and B (r^-6) tables."""

    def __init__(self, analyticForm=_OPLS, label="LJ Parameters"):
        form = analyticForm.upper()
        if form not in _FORMS:
            raise ValueError("Unrecognized analytic form: {:s}.".format(analyticForm))
        self.analyticForm = form
        self.label = label
        self.parameterKeys = None
        self.ntypes = 0
        self.epsilon = np.zeros(0)
        self.sigma = np.zeros(0)
        self.tableA = np.zeros(0)
        self.tableB = np.zeros(0)

    @property
    def size(self):
        return self.ntypes

    def _allocate(self, ntypes):
        self.ntypes = int(ntypes)
        m = ntypes * (ntypes + 1) // 2
        self.epsilon = np.zeros(ntypes)
        self.sigma = np.zeros(ntypes)
        self.tableA = np.zeros(m)
        self.tableB = np.zeros(m)

    def MakeTable(self):
        """AMBER: arithmetic sigma, A = e s^12, B = 2 e s^6; OPLS: geometric sigma, A = 4 e s^12, B = 4 e s^6."""
        if self.analyticForm == _TABLE:
            return
        geometric = self.analyticForm == _OPLS
        n = 0
        for i in range(self.ntypes):
            for j in range(i + 1):
                eij = math.sqrt(self.epsilon[i] * self.epsilon[j])
                sij = math.sqrt(self.sigma[i] * self.sigma[j]) if geometric else 0.5 * (self.sigma[i] + self.sigma[j])
                s6 = sij ** 6
                s12 = s6 * s6
                if geometric:
                    eij *= 4.0
                else:
                    s6 *= 2.0
                self.tableA[n] = eij * s12
                self.tableB[n] = eij * s6
                n += 1

    def MakeEpsilonSigma(self):
        """Diagonal table entries back to epsilon/sigma (LJParameterContainer.c:155-182)."""
        if self.analyticForm == _TABLE:
            return
        four = self.analyticForm == _OPLS
        for i in range(self.ntypes):
            n = i * (i + 1) // 2 + i
            es12, es6 = abs(self.tableA[n]), abs(self.tableB[n])
            if four:
                es12 /= 4.0
                es6 /= 4.0
            else:
                es6 /= 2.0
            if es6 < 1.0e-20 or es12 < 1.0e-20:
                self.epsilon[i] = self.sigma[i] = 0.0
            else:
                self.epsilon[i] = es6 * es6 / es12
                self.sigma[i] = math.exp(math.log(es12 / es6) / 6.0)

    def Scale(self, scale):
        self.epsilon = self.epsilon * scale
        self.tableA = self.tableA * scale
        self.tableB = self.tableB * scale

    def __getstate__(self):
        state = {"analyticForm": self.analyticForm, "numberOfTypes": self.ntypes, "label": self.label}
        if self.parameterKeys is not None:
            state["parameterKeys"] = list(self.parameterKeys)
        if self.analyticForm == _TABLE:
            state["aCoefficients"] = [float(x) for x in self.tableA]
            state["bCoefficients"] = [float(x) for x in self.tableB]
        else:
            state["epsilons"] = [float(x) for x in self.epsilon]
            state["sigmas"] = [float(x) for x in self.sigma]
        return state

    def __setstate__(self, state):
        self.__init__(state["analyticForm"], state.get("label", "LJ Parameters"))
        self._allocate(state["numberOfTypes"])
        self.parameterKeys = state.get("parameterKeys", None)
        if self.analyticForm == _TABLE:
            self.tableA[:] = state["aCoefficients"]
            self.tableB[:] = state["bCoefficients"]
            self.MakeEpsilonSigma()
        else:
            self.epsilon[:] = state["epsilons"]
            self.sigma[:] = state["sigmas"]
            self.MakeTable()

    def SummaryEntry(self, summary):
        if summary is not None:
            summary.Entry(self.label + " Form", self.analyticForm)
            summary.Entry(self.label + " Types", "{:d}".format(self.ntypes))


def LJParameterContainer_FromEpsilonSigma(ntypes, epsilons, sigmas, analyticForm=_OPLS, parameterKeys=None):
    """Constructor from epsilon and sigma (pMolecule.LJParameterContainer.pyx:254-270)."""
    if ntypes <= 0 or len(epsilons) != ntypes or len(sigmas) != ntypes:
        return None
    self = LJParameterContainer(analyticForm)
    self._allocate(ntypes)
    self.epsilon[:] = epsilons
    self.sigma[:] = sigmas
    self.MakeTable()
    self.parameterKeys = parameterKeys
    return self


def LJParameterContainer_FromTableCoefficients(ntypes, acoefficients, bcoefficients, analyticForm=_TABLE,
                                               parameterKeys=None):
    """Constructor from packed table coefficients (pMolecule.LJParameterContainer.pyx:272-290)."""
    nsize = ntypes * (ntypes + 1) // 2
    if ntypes <= 0 or len(acoefficients) != nsize or len(bcoefficients) != nsize:
        return None
    self = LJParameterContainer(analyticForm)
    self._allocate(ntypes)
    self.tableA[:] = acoefficients
    self.tableB[:] = bcoefficients
    self.MakeEpsilonSigma()
    self.parameterKeys = parameterKeys
    return self


class MMAtomContainer:
    """Charges (e) and LJ types of the MM atoms, structure of arrays."""

    def __init__(self, charges, ljtypes, atomtypes=None, active=None):
        self.charges = np.ascontiguousarray(charges, dtype=np.float64)
        self.ljtypes = np.ascontiguousarray(ljtypes, dtype=np.int32)
        if self.charges.shape != self.ljtypes.shape:
            raise ValueError("charges and ljtypes must have the same length")
        self.atomtypes = self.ljtypes.copy() if atomtypes is None else np.ascontiguousarray(atomtypes, np.int32)
        self.active = np.ones(self.charges.shape[0], bool) if active is None else np.asarray(active, bool)

    @property
    def size(self):
        return int(self.charges.shape[0])

    def __len__(self):
        return self.size

    def TotalCharge(self):
        return float(self.charges.sum())


class SelfPairList:
    """A self pair list (i > j records) from integer pairs."""

    def __init__(self, pairs=None, natoms=None):
        p = np.zeros((0, 2), np.int64) if pairs is None else np.asarray(pairs, dtype=np.int64).reshape(-1, 2)
        if p.size and p.min() < 0:
            raise ValueError("negative atom index in pair list")
        p = p[p[:, 0] != p[:, 1]]
        hi = np.maximum(p[:, 0], p[:, 1])
        lo = np.minimum(p[:, 0], p[:, 1])
        base = int(hi.max()) + 1 if hi.size else 1
        key = np.unique(hi * base + lo)
        self.pairs = np.stack([key // base, key % base], 1).astype(np.int32) if key.size else np.zeros((0, 2), np.int32)
        self.upperBound = int(self.pairs[:, 0].max()) + 1 if key.size else 0
        if natoms is not None and self.upperBound > natoms:
            raise ValueError("pair index beyond the number of atoms")

    @property
    def numberOfPairs(self):
        return int(self.pairs.shape[0])

    def __len__(self):
        return self.numberOfPairs


def SelfPairList_FromIntegerPairs(pairs, natoms=None):
    return SelfPairList(pairs, natoms)
