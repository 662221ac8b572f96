"""CHARMM PSF, CHARMM parameter and XYZ readers, for the non-bonding path only.

What the NBModelABFS path needs from a CHARMM system: the atoms' types and
charges, the exclusions and 1-4 pairs generated from the bond list, the
Lennard-Jones parameters of the NONBONDED section (normal and 1-4) in the
AMBER analytic form, and the coordinates.  Names and rules follow
pBabel-1.9.0/pBabel/CHARMMPSFFileReader.py, CHARMMParameterFileReader.py and
XYZFileReader.py; the bonded MM terms are outside this build.
"""

import numpy as np

from .Parameters import LJParameterContainer_FromEpsilonSigma, MMAtomContainer, SelfPairList

# pCore-1.9.0/pCore/Units.py:35.
UNITS_ENERGY_KILOCALORIES_PER_MOLE_TO_KILOJOULES_PER_MOLE = 4.184

# CHARMMParameterFileReader.py:24-34: section keys (first four letters) and wildcards.
_PRMSections = {"ANGL": "Angle", "ATOM": "Atom", "BOND": "Bond", "CMAP": "CMap", "DIHE": "Dihedral", "END": "End",
                "EQUI": "Equivalence", "HBON": "HBond", "IMPH": "Improper", "IMPR": "Improper", "NBFI": "NBFix",
                "NBON": "NonBond", "NONB": "NonBond", "PHI": "Dihedral", "PRIN": "Print", "SPAS": "SPAS",
                "THET": "Angle"}
_NonBondWildCards = ("*", "%", "#", "+")

# CHARMMPSFFileReader.py:41-51: PSF section names and the number of integers per record.
_PSFSections = {"NATOM": "Atoms", "NBOND": "Bonds", "NTHETA": "Angles", "NPHI": "Dihedrals", "NIMPHI": "Impropers",
                "NDON": "Donors", "NACC": "Acceptors", "NNB": "Exclusions", "NGRP": "Groups", "NTITLE": "Title",
                "MOLNT": "CheqLabels", "NUMLP": "LonePairs", "NCRTERM": "CrossTerms"}
_PSFIntegersPerRecord = {"Bonds": 2, "Angles": 3, "Dihedrals": 4, "Impropers": 4, "Donors": 2, "Acceptors": 2,
                         "CrossTerms": 8}


class CHARMMParameterFileReader:
    """The NONBONDED section of a CHARMM parameter file (the behaviour of CHARMMParameterFileReader.py:76-160,
    315-345)."""

    def __init__(self, path):
        self.path = path
        self.nonbonds, self.nonbondwilds = {}, {}
        self.nonbond14s, self.nonbond14wilds = {}, {}
        self.warnings = []
        self.QPARSED = False

    @staticmethod
    def _lines(path):
        """Non-empty lines without "!" comments, continuation lines ("-" at the end) joined."""
        found = ""
        with open(path, "r", errors="replace") as f:
            for raw in f:
                line = raw.strip()
                k = line.find("!")
                if k >= 0:
                    line = line[:k].strip()
                if line.endswith("-"):
                    found += line[:-1].strip()
                    continue
                found += line
                if found:
                    yield found
                found = ""
        if found:
            yield found

    @staticmethod
    def _section(line):
        if line.startswith("*"):
            return "Title", line[1:].strip()
        tokens = line.split(" ", 1)
        head = tokens[0].upper()
        for key, name in _PRMSections.items():
            if head.startswith(key):
                return name, (tokens[1] if len(tokens) > 1 else "")
        return None, line

    @staticmethod
    def _triple(tokens, first):
        """The three numbers (ignored, emin, rmin/2) of a record starting at tokens[first], or None."""
        if len(tokens) < first + 3:
            return None
        try:
            return tuple(float(v) for v in tokens[first:first + 3])
        except ValueError:
            return None

    def _store_nonbond(self, tokens, conv):
        """One NONBONDED record, "type ignored emin rmin/2 [ignored emin14 rmin14/2]" (kcal/mol, A): the type's
        Lennard-Jones epsilon |emin| (in kJ/mol when converting) and sigma = 2 rmin/2 of the AMBER form, for the
        normal and, when given, the 1-4 interaction (CHARMMParameterFileReader.py:315-345).  Types with a
        wildcard character are kept apart; the first record of a type wins; a non-zero first column or a
        positive emin is not a Lennard-Jones record of this form and is skipped with a warning."""
        key = tokens[0].upper()
        wild = any(c in key for c in _NonBondWildCards)
        parts = [("nonbond", self._triple(tokens, 1), self.nonbondwilds if wild else self.nonbonds)]
        if len(tokens) > 4:
            parts.append(("1-4 nonbond", self._triple(tokens, 4), self.nonbond14wilds if wild else self.nonbond14s))
        for what, values, table in parts:
            if values is None:
                self.warnings.append("Malformed %s record: %s." % (what, " ".join(tokens)))
                break
            ignored, emin, rhalf = values
            if key in table:
                self.warnings.append("Repeated %s record for type %s (the first is kept)." % (what, key))
            elif ignored != 0.0 or emin > 0.0:
                self.warnings.append("The %s record of type %s is not of the epsilon/Rmin form." % (what, key))
            else:
                table[key] = (abs(emin) * conv, 2.0 * rhalf)

    def Parse(self, convertUnits=True):
        if self.QPARSED:
            return self
        conv = UNITS_ENERGY_KILOCALORIES_PER_MOLE_TO_KILOJOULES_PER_MOLE if convertUnits else 1.0
        current, body = None, []
        for line in self._lines(self.path):
            section, rest = self._section(line)
            if section is None:
                if current == "NonBond":
                    body.append(rest.split())
            else:
                if section == "End":
                    break
                current = section
        for tokens in body:
            self._store_nonbond(tokens, conv)
        self.QPARSED = True
        return self


class CHARMMPSFFileReader:
    """Atoms, bonds and explicit exclusions of a CHARMM PSF file (CHARMMPSFFileReader.py:62-160)."""

    def __init__(self, path):
        self.path = path
        self.atoms = []          # (type, charge)
        self.bonds = []          # 1-based, flattened
        self.exclusionsI, self.exclusionsJ = [], []
        self.counters = {}
        self.QPARSED = False

    def Parse(self):
        if self.QPARSED:
            return self
        with open(self.path, "r", errors="replace") as f:
            lines = f.read().splitlines()
        k = 0
        while k < len(lines):
            line = lines[k]
            k += 1
            if "!" not in line:
                continue
            head, tail = line.split("!", 1)
            name = tail.split(":", 1)[0].strip().split()[0] if tail.strip() else ""
            if name not in _PSFSections:
                continue
            try:
                count = int(head.split()[0])
            except (IndexError, ValueError):
                continue
            section = _PSFSections[name]
            self.counters[section] = count
            if section == "Title":
                k += count
            elif section == "Atoms":
                for n in range(count):
                    tokens = lines[k + n].split()
                    self.atoms.append((tokens[5].upper(), float(tokens[6])))
                k += count
            elif section in _PSFIntegersPerRecord or section == "Exclusions":
                need = count * _PSFIntegersPerRecord.get(section, 1)
                if section == "Exclusions":
                    need = count + self.counters.get("Atoms", 0)
                values = []
                while len(values) < need and k < len(lines):
                    values.extend(int(v) for v in lines[k].split())
                    k += 1
                if section == "Bonds":
                    self.bonds = values[:need]
                elif section == "Exclusions":
                    self.exclusionsJ = [v - 1 for v in values[:count]]
                    self.exclusionsI = values[count:need]
        self.QPARSED = True
        return self

    def ToMMAtomContainer(self):
        """Sorted unique atom types and the atoms' type indices and charges (CHARMMPSFFileReader.py:355-380)."""
        types = [t for t, _ in self.atoms]
        self.uniqueAtomTypes = sorted(set(types))
        index = {t: i for i, t in enumerate(self.uniqueAtomTypes)}
        return MMAtomContainer(np.array([q for _, q in self.atoms]), np.array([index[t] for t in types], np.int32))

    def _explicit_exclusions(self, natoms):
        """The PSF's NNB section as (i > j) pairs.  Atom i takes the first INB[i] entries of the excluded-atom
        list, the reading pBabel's PSF reader gives the INB array (CHARMMPSFFileReader.py:367-372)."""
        pairs = set()
        if self.counters.get("Exclusions", 0) > 0:
            for i, end in enumerate(self.exclusionsI[:natoms]):
                for j in self.exclusionsJ[:end]:
                    if j != i:
                        pairs.add((max(i, j), min(i, j)))
        return pairs

    def ToExclusionPairLists(self):
        """Exclusions and 1-4 pairs from the bond graph, by CHARMM's rule (the behaviour of
        CHARMMPSFFileReader.py:334-396): two atoms joined by a chain of at most three bonds are excluded from
        the main list, and so are the PSF's explicit exclusions; the pairs whose shortest chain has exactly
        three bonds interact through the 1-4 list, unless explicitly excluded.  Shortest chains come from a
        breadth-first search of depth three from every atom."""
        natoms = self.counters.get("Atoms", 0)
        nbonds = self.counters.get("Bonds", 0)
        neighbours = [set() for _ in range(natoms)]
        for n in range(nbonds):
            i, j = self.bonds[2 * n] - 1, self.bonds[2 * n + 1] - 1
            if i != j:
                neighbours[i].add(j)
                neighbours[j].add(i)
        near, third = set(), set()
        for a in range(natoms):
            seen = {a}
            shell = [a]
            for depth in (1, 2, 3):
                nxt = []
                for u in shell:
                    for v in neighbours[u]:
                        if v not in seen:
                            seen.add(v)
                            nxt.append(v)
                shell = nxt
                for b in shell:
                    if b < a:                      # each pair once, from its larger atom
                        near.add((a, b))
                        if depth == 3:
                            third.add((a, b))
        explicit = self._explicit_exclusions(natoms)
        excluded = near | explicit
        onefour = third - explicit
        exclusions = SelfPairList([k for pair in sorted(excluded) for k in pair]) if excluded else None
        interactions14 = SelfPairList([k for pair in sorted(onefour) for k in pair]) if onefour else None
        return exclusions, interactions14

    def ToLJParameterContainers(self, parameters):
        """Normal and 1-4 LJ containers in the AMBER form (CHARMMPSFFileReader.py:346-372)."""
        if not hasattr(self, "uniqueAtomTypes"):
            self.ToMMAtomContainer()
        eps, sig, eps14, sig14 = [], [], [], []
        for t in self.uniqueAtomTypes:
            e, r = parameters.nonbonds[t]
            d14 = parameters.nonbond14s.get(t)
            eps.append(e)
            sig.append(r)
            eps14.append(e if d14 is None else d14[0])
            sig14.append(r if d14 is None else d14[1])
        n = len(self.uniqueAtomTypes)
        lj = LJParameterContainer_FromEpsilonSigma(n, eps, sig, analyticForm="AMBER",
                                                   parameterKeys=list(self.uniqueAtomTypes))
        lj14 = LJParameterContainer_FromEpsilonSigma(n, eps14, sig14, analyticForm="AMBER",
                                                     parameterKeys=list(self.uniqueAtomTypes))
        return lj, lj14


def XYZFile_ToCoordinates3(path):
    """Coordinates of an XYZ file: atom count, title, then "symbol x y z" lines (XYZFileReader.py:20-75)."""
    with open(path, "r", errors="replace") as f:
        natoms = int(f.readline().split()[0])
        f.readline()
        xyz = np.empty((natoms, 3))
        for n in range(natoms):
            tokens = f.readline().split()
            xyz[n] = [float(tokens[1]), float(tokens[2]), float(tokens[3])]
    return xyz


def CHARMMPSFFile_ToNBParameters(psfPath, parameterPaths):
    """Everything the non-bonding model needs from a CHARMM PSF and parameter files: an MMAtomContainer, the LJ
    containers (normal, 1-4), the exclusions and 1-4 SelfPairLists, and electrostaticScale14 = 1.0 (MMModelCHARMM,
    pMolecule-1.9.0/pMolecule/MMModel.py:31-34,113-118)."""
    psf = CHARMMPSFFileReader(psfPath).Parse()
    params = None
    for path in parameterPaths:
        p = CHARMMParameterFileReader(path).Parse()
        if params is None:
            params = p
        else:
            params.nonbonds.update(p.nonbonds)
            params.nonbond14s.update(p.nonbond14s)
    atoms = psf.ToMMAtomContainer()
    exclusions, interactions14 = psf.ToExclusionPairLists()
    lj, lj14 = psf.ToLJParameterContainers(params)
    return {"atoms": atoms, "ljParameters": lj, "ljParameters14": lj14, "exclusions": exclusions,
            "interactions14": interactions14, "electrostaticScale14": 1.0, "atomTypes": psf.uniqueAtomTypes}
