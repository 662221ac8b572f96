"""CHARMM PSF / parameter / XYZ readers of the non-bonding path (CPU)."""

import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REFERENCE = "/root/reference"

PSF = """PSF EXT

         1 !NTITLE
* toy chain: 0-1-2-3-4 with a branch 2-5

         6 !NATOM
         1 TOY      1        ALA      C1       CT1      -0.100000       12.0110           0
         2 TOY      1        ALA      C2       CT2       0.200000       12.0110           0
         3 TOY      1        ALA      C3       CT1      -0.300000       12.0110           0
         4 TOY      1        ALA      C4       CT2       0.100000       12.0110           0
         5 TOY      1        ALA      C5       ct1       0.050000       12.0110           0
         6 TOY      1        ALA      H6       HA        0.050000        1.0080           0

         5 !NBOND: bonds
         1         2         2         3         3         4         4         5
         3         6

         0 !NNB

"""

PRM = """* toy parameters
*
NONBONDED nbxmod  5 atom cdiel shift vatom vdistance vswitch -
cutnb 14.0 ctofnb 12.0 ctonnb 10.0 eps 1.0 e14fac 1.0 wmin 1.5
!atom  ignored    epsilon      Rmin/2   ignored   eps,1-4       Rmin/2,1-4
CT1    0.0       -0.0200     2.2750   0.0  -0.01 1.9 ! comment
CT2    0.0       -0.0550     2.1750   0.0  -0.01 1.9
HA     0.0       -0.0220     1.3200
C*     0.0       -0.1100     2.0000
END
"""


def _readers(tmp_path):
    from pdynamo_amd import CHARMM
    psf = tmp_path / "toy.psf"
    prm = tmp_path / "toy.prm"
    psf.write_text(PSF)
    prm.write_text(PRM)
    return CHARMM, str(psf), str(prm)


def test_exclusions_and_14_from_bonds(tmp_path):
    """CHARMM rule: exclusions = 1-2 + 1-3 + every 1-4 candidate; 1-4 list = candidates not 1-2/1-3."""
    CHARMM, psf, prm = _readers(tmp_path)
    r = CHARMM.CHARMMPSFFileReader(psf).Parse()
    exc, i14 = r.ToExclusionPairLists()
    bonds = {(1, 0), (2, 1), (3, 2), (4, 3), (5, 2)}
    angles = {(2, 0), (3, 1), (5, 1), (4, 2), (5, 3)}
    dihed = {(3, 0), (5, 0), (4, 1), (5, 4)}
    assert {tuple(p) for p in exc.pairs.tolist()} == bonds | angles | dihed
    assert {tuple(p) for p in i14.pairs.tolist()} == dihed


def test_types_charges_and_lj_tables(tmp_path):
    CHARMM, psf, prm = _readers(tmp_path)
    nb = CHARMM.CHARMMPSFFile_ToNBParameters(psf, [prm])
    assert nb["atomTypes"] == ["CT1", "CT2", "HA"]          # sorted, upper-cased
    assert nb["atoms"].ljtypes.tolist() == [0, 1, 0, 1, 0, 2]
    assert np.allclose(nb["atoms"].charges, [-0.1, 0.2, -0.3, 0.1, 0.05, 0.05])
    lj, lj14 = nb["ljParameters"], nb["ljParameters14"]
    # AMBER form: e_ij = sqrt(e_i e_j), R_ij = (R_i + R_j)/2 with R = 2 Rmin/2, A = e R^12, B = 2 e R^6 (kJ/mol).
    e = {"CT1": 0.02 * 4.184, "CT2": 0.055 * 4.184, "HA": 0.022 * 4.184}
    R = {"CT1": 4.55, "CT2": 4.35, "HA": 2.64}
    eij, rij = np.sqrt(e["CT2"] * e["CT1"]), 0.5 * (R["CT2"] + R["CT1"])
    assert np.isclose(lj.tableA[1], eij * rij ** 12, rtol=1e-14)   # packed index (1,0)
    assert np.isclose(lj.tableB[1], 2.0 * eij * rij ** 6, rtol=1e-14)
    e14, r14 = 0.01 * 4.184, 3.8
    assert np.isclose(lj14.tableA[0], e14 * r14 ** 12, rtol=1e-14)
    assert np.isclose(lj14.tableA[5], e["HA"] * R["HA"] ** 12, rtol=1e-14)   # no 1-4 columns: normal values
    assert nb["electrostaticScale14"] == 1.0


def test_xyz_reader(tmp_path):
    from pdynamo_amd import CHARMM
    p = tmp_path / "toy.xyz"
    p.write_text("2\ntitle\nC 1.0 2.0 3.0\nH -1.5 0.25 1e-3\n")
    assert np.allclose(CHARMM.XYZFile_ToCoordinates3(str(p)), [[1.0, 2.0, 3.0], [-1.5, 0.25, 1e-3]])


@pytest.mark.skipif(not os.path.isdir(os.path.join(REFERENCE, "benchmarks", "data", "dhfr")),
                    reason="reference benchmark data not in this container")
def test_dhfr_fixture_is_the_readers_output(dhfr):
    """The committed DHFR fixture equals what the readers make of the reference's files."""
    from pdynamo_amd import CHARMM
    d = os.path.join(REFERENCE, "benchmarks", "data", "dhfr")
    nb = CHARMM.CHARMMPSFFile_ToNBParameters(os.path.join(d, "dhfr.psfx"), [os.path.join(d, "par_all22_prot.prm")])
    assert np.array_equal(nb["atoms"].charges, dhfr.mm.charges)
    assert np.array_equal(nb["atoms"].ljtypes, dhfr.mm.ljtypes)
    assert np.array_equal(nb["ljParameters"].tableA, dhfr.mm.tableA)
    assert np.array_equal(nb["ljParameters14"].tableB, dhfr.tableB14)
    assert np.array_equal(nb["exclusions"].pairs, dhfr.mm.exclusions)
    assert np.array_equal(nb["interactions14"].pairs, dhfr.mm.interactions14)
    assert np.array_equal(CHARMM.XYZFile_ToCoordinates3(os.path.join(d, "dhfr.xyz")), dhfr.mm.coordinates3)


def test_exclusion_rule_on_rings_matches_path_enumeration():
    """The shortest-chain rule (excluded within three bonds, 1-4 at exactly three) against a direct enumeration
    of bond chains: 1-2 bonds, 1-3 pairs sharing a neighbour, 1-4 candidates the ends of every chain of three
    distinct atoms' bonds; exclusions = all of them, 1-4 pairs = candidates that are neither 1-2 nor 1-3.  Random
    graphs with 3-, 4- and 5-membered rings, where a pair can be both a 1-3 and the end of a three-bond chain."""
    from pdynamo_amd import CHARMM
    rng = np.random.default_rng(8)
    for trial in range(30):
        n = int(rng.integers(6, 40))
        bonds = {(max(i, i + 1), min(i, i + 1)) for i in range(n - 1)}           # a chain ...
        for _ in range(int(rng.integers(1, 8))):                                # ... closed into small rings
            i = int(rng.integers(0, n))
            j = i + int(rng.integers(2, 5))
            if j < n:
                bonds.add((j, i))
        r = CHARMM.CHARMMPSFFileReader.__new__(CHARMM.CHARMMPSFFileReader)
        r.counters = {"Atoms": n, "Bonds": len(bonds)}
        r.bonds = [k + 1 for b in sorted(bonds) for k in b]
        r.exclusionsI, r.exclusionsJ = [], []
        exc, i14 = r.ToExclusionPairLists()
        nb = {a: set() for a in range(n)}
        for i, j in bonds:
            nb[i].add(j)
            nb[j].add(i)
        key = lambda a, b: (max(a, b), min(a, b))
        i12 = {key(i, j) for i, j in bonds}
        i13 = {key(a, b) for c in range(n) for a in nb[c] for b in nb[c] if a != b}
        i14p = {key(a, d) for (b, c) in i12 for a in nb[b] for d in nb[c] if a != c and d != b and a != d}
        i14p |= {key(a, d) for (c, b) in i12 for a in nb[b] for d in nb[c] if a != c and d != b and a != d}
        assert {tuple(p) for p in exc.pairs.tolist()} == i12 | i13 | i14p, trial
        got14 = set() if i14 is None else {tuple(p) for p in i14.pairs.tolist()}
        assert got14 == i14p - i12 - i13, trial


def test_explicit_exclusions_as_the_reference_reads_them():
    """The NNB section: atom i takes the first INB[i] entries of the excluded-atom list (the reading of the
    reference's PSF reader); explicit exclusions join the exclusions and leave the 1-4 list."""
    from pdynamo_amd import CHARMM
    r = CHARMM.CHARMMPSFFileReader.__new__(CHARMM.CHARMMPSFFileReader)
    r.counters = {"Atoms": 6, "Bonds": 5, "Exclusions": 2}
    r.bonds = [2, 1, 3, 2, 4, 3, 5, 4, 6, 5]              # chain 0-1-2-3-4-5 (1-based in the PSF)
    r.exclusionsJ = [3, 5]                                # 0-based excluded atoms
    r.exclusionsI = [2, 2, 2, 2, 2, 2]                    # INB per atom: each takes J[0:2]
    exc, i14 = r.ToExclusionPairLists()
    within3 = {(i, j) for i in range(6) for j in range(i) if i - j <= 3}
    assert {tuple(p) for p in exc.pairs.tolist()} == within3 | {(5, 0), (5, 1)}
    # 1-4 candidates (3, 0), (4, 1), (5, 2): (3, 0) and (5, 2) are explicitly excluded.
    assert {tuple(p) for p in i14.pairs.tolist()} == {(4, 1)}
