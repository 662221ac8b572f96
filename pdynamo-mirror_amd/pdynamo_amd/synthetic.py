"""Seeded synthetic MM systems for the NBModelABFS hot path.

Five workloads, all non-periodic, in pDynamo units (A, e, kJ/mol):

* ``water_lattice``      -- rigid 3-site waters on a jittered cubic lattice
                            (648 atoms in an 18.6 A cube; 200,001 atoms in a
                            126 A cube).
* ``solute_in_water``    -- 9,667 waters in a 67 A cube plus 999 random solute
                            atoms with 8 Lorentz-Berthelot LJ types and 1-2
                            exclusions along a random chain (30,000 atoms).
* ``water_droplets``     -- spheres of water at 0.0334 molecules/A^3 with atom
                            counts log-uniform in [500, 20000].
* ``perturbed_copies``   -- independent Gaussian perturbations of one system.

Water geometry: O-H 0.9572 A, H-O-H 104.52 deg; charges O -0.834, H +0.417;
O LJ epsilon 0.1521 kcal/mol, sigma 3.1507 A; H has no LJ.  The LJ table for
water alone uses the OPLS form (A = 4 eps s^12, B = 4 eps s^6, geometric sigma,
pMolecule-1.9.0/extensions/csource/LJParameterContainer.c:184-210); the
solute mixture uses Lorentz-Berthelot coefficients supplied as a table
(LJParameterContainer_FromTableCoefficients,
pMolecule-1.9.0/extensions/pyrex/pMolecule.LJParameterContainer.pyx:272-290).
Tables are packed lower-triangular: index i*(i+1)/2 + j for j <= i.
"""

from dataclasses import dataclass, field
import math

import numpy as np

KCAL_TO_KJ = 4.184
O_H = 0.9572
HOH_DEG = 104.52
Q_O, Q_H = -0.834, 0.417
EPS_O_KCAL, SIGMA_O = 0.1521, 3.1507
WATER_DENSITY = 0.0334  # molecules / A^3


@dataclass
class MMSystem:
    """Coordinates plus the MM parameters the ABFS model consumes."""
    name: str
    coordinates3: np.ndarray            # (n, 3) float64
    charges: np.ndarray                 # (n,) float64
    ljtypes: np.ndarray                 # (n,) int32
    ntypes: int
    tableA: np.ndarray                  # packed triangular, float64
    tableB: np.ndarray
    exclusions: np.ndarray              # (m, 2) int32
    interactions14: np.ndarray = field(default_factory=lambda: np.zeros((0, 2), np.int32))

    @property
    def natoms(self):
        return int(self.coordinates3.shape[0])

    def with_coordinates(self, xyz, name=None):
        return MMSystem(name or self.name, np.ascontiguousarray(xyz, dtype=np.float64), self.charges,
                        self.ljtypes, self.ntypes, self.tableA, self.tableB, self.exclusions,
                        self.interactions14)


def packed_index(i, j):
    if i < j:
        i, j = j, i
    return i * (i + 1) // 2 + j


def _random_rotations(rng, m):
    """Uniform random rotation matrices from unit quaternions."""
    q = rng.normal(size=(m, 4))
    q /= np.linalg.norm(q, axis=1)[:, None]
    w, x, y, z = q.T
    r = np.empty((m, 3, 3))
    r[:, 0, 0] = 1 - 2 * (y * y + z * z); r[:, 0, 1] = 2 * (x * y - z * w); r[:, 0, 2] = 2 * (x * z + y * w)
    r[:, 1, 0] = 2 * (x * y + z * w); r[:, 1, 1] = 1 - 2 * (x * x + z * z); r[:, 1, 2] = 2 * (y * z - x * w)
    r[:, 2, 0] = 2 * (x * z - y * w); r[:, 2, 1] = 2 * (y * z + x * w); r[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return r


def _water_sites(rng, centres):
    """O at each centre, the two H at random orientations (rigid geometry)."""
    m = centres.shape[0]
    half = math.radians(HOH_DEG) / 2.0
    body = np.array([[0.0, 0.0, 0.0],
                     [O_H * math.sin(half), 0.0, O_H * math.cos(half)],
                     [-O_H * math.sin(half), 0.0, O_H * math.cos(half)]])
    rot = _random_rotations(rng, m)
    sites = centres[:, None, :] + np.einsum("mij,kj->mki", rot, body)
    return sites.reshape(-1, 3)


def _water_params(nwater):
    n = 3 * nwater
    q = np.tile(np.array([Q_O, Q_H, Q_H]), nwater)
    t = np.tile(np.array([0, 1, 1], np.int32), nwater)
    base = 3 * np.arange(nwater, dtype=np.int32)
    excl = np.stack([np.stack([base + 1, base], 1), np.stack([base + 2, base], 1),
                     np.stack([base + 2, base + 1], 1)], 1).reshape(-1, 2)
    return n, q, t, excl


def opls_table(eps_kj, sigma):
    """OPLS-form packed table (geometric sigma, A = 4 eps s^12, B = 4 eps s^6)."""
    nt = len(eps_kj)
    a = np.zeros(nt * (nt + 1) // 2)
    b = np.zeros_like(a)
    for i in range(nt):
        for j in range(i + 1):
            e = math.sqrt(eps_kj[i] * eps_kj[j]) * 4.0
            s = math.sqrt(sigma[i] * sigma[j])
            s6 = s ** 6
            a[packed_index(i, j)] = e * s6 * s6
            b[packed_index(i, j)] = e * s6
    return a, b


def lorentz_berthelot_table(eps_kj, sigma):
    """Packed table with arithmetic sigma and geometric epsilon, 4 eps form."""
    nt = len(eps_kj)
    a = np.zeros(nt * (nt + 1) // 2)
    b = np.zeros_like(a)
    for i in range(nt):
        for j in range(i + 1):
            e = 4.0 * math.sqrt(eps_kj[i] * eps_kj[j])
            s6 = (0.5 * (sigma[i] + sigma[j])) ** 6
            a[packed_index(i, j)] = e * s6 * s6
            b[packed_index(i, j)] = e * s6
    return a, b


def _lattice_centres(rng, box, per_side, count, jitter):
    spacing = box / per_side
    g = (np.arange(per_side) + 0.5) * spacing
    pts = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    if count < pts.shape[0]:
        pts = pts[np.sort(rng.choice(pts.shape[0], size=count, replace=False))]
    return pts + rng.uniform(-jitter, jitter, size=pts.shape)


def water_lattice(nwater=216, box=18.6, seed=20261202, jitter=0.3, name="water_lattice"):
    """Waters on a (subset of a) jittered cubic lattice filling a cube."""
    rng = np.random.default_rng(seed)
    per_side = int(math.ceil(round(nwater ** (1.0 / 3.0), 9)))
    centres = _lattice_centres(rng, box, per_side, nwater, jitter)
    xyz = _water_sites(rng, centres)
    n, q, t, excl = _water_params(nwater)
    a, b = opls_table([EPS_O_KCAL * KCAL_TO_KJ, 0.0], [SIGMA_O, 0.0])
    return MMSystem(name, xyz, q, t, 2, a, b, excl.astype(np.int32))


def config0(seed=20261202):
    """648 atoms: 216 waters, 6x6x6 lattice in an 18.6 A cube."""
    return water_lattice(216, 18.6, seed, name="C0_water648")


def config2(seed=20261204):
    """200,001 atoms: 66,667 waters in a 126 A cube."""
    return water_lattice(66667, 126.0, seed, name="C2_water200001")


def solute_in_water(nwater=9667, nsolute=999, box=67.0, ntypes_solute=8, seed=20261203,
                    name="solute_in_water"):
    """Waters on a lattice subset plus random solute atoms (Lorentz-Berthelot)."""
    rng = np.random.default_rng(seed)
    per_side = int(math.ceil(nwater ** (1.0 / 3.0)))
    centres = _lattice_centres(rng, box, per_side, nwater, 0.3)
    wxyz = _water_sites(rng, centres)
    nw, qw, tw, exw = _water_params(nwater)
    sxyz = rng.uniform(0.0, box, size=(nsolute, 3))
    sq = rng.uniform(-0.5, 0.5, size=nsolute)
    st = rng.integers(0, ntypes_solute, size=nsolute).astype(np.int32) + 2
    eps = [EPS_O_KCAL * KCAL_TO_KJ, 0.0] + list(rng.uniform(0.05, 0.2, ntypes_solute) * KCAL_TO_KJ)
    sig = [SIGMA_O, 0.0] + list(rng.uniform(2.5, 4.0, ntypes_solute))
    a, b = lorentz_berthelot_table(eps, sig)
    chain = rng.permutation(nsolute) + nw
    exs = np.stack([chain[1:], chain[:-1]], 1)
    xyz = np.concatenate([wxyz, sxyz])
    q = np.concatenate([qw, sq])
    t = np.concatenate([tw, st]).astype(np.int32)
    excl = np.concatenate([exw, exs]).astype(np.int32)
    return MMSystem(name, xyz, q, t, 2 + ntypes_solute, a, b, excl)


def config1(seed=20261203):
    """30,000 atoms: 9,667 waters plus 999 solute atoms in a 67 A cube."""
    return solute_in_water(seed=seed, name="C1_solute30000")


def water_droplet(natoms, rng, name="droplet"):
    """A sphere of waters at 0.0334 molecules/A^3 (closest lattice sites)."""
    nwater = max(1, natoms // 3)
    spacing = (1.0 / WATER_DENSITY) ** (1.0 / 3.0)
    radius = (3.0 * nwater / (4.0 * math.pi * WATER_DENSITY)) ** (1.0 / 3.0)
    k = int(math.ceil(radius / spacing)) + 2
    g = np.arange(-k, k + 1) * spacing
    pts = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    d = np.linalg.norm(pts, axis=1)
    centres = pts[np.argsort(d, kind="stable")[:nwater]]
    centres = centres + rng.uniform(-0.3, 0.3, size=centres.shape)
    xyz = _water_sites(rng, centres)
    n, q, t, excl = _water_params(nwater)
    a, b = opls_table([EPS_O_KCAL * KCAL_TO_KJ, 0.0], [SIGMA_O, 0.0])
    return MMSystem(name, xyz, q, t, 2, a, b, excl.astype(np.int32))


def config3(count=64, lo=500, hi=20000, seed=20261205):
    """64 droplets with atom counts log-uniform in [500, 20000] (multiples of 3)."""
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(math.log(lo), math.log(hi), size=count))
    out = []
    for k, s in enumerate(sizes):
        out.append(water_droplet(int(round(s / 3.0)) * 3, rng, name="C3_droplet%02d" % k))
    return out


def perturbed_copies(system, count, sigma, seed):
    rng = np.random.default_rng(seed)
    return [system.with_coordinates(system.coordinates3 + rng.normal(0.0, sigma, system.coordinates3.shape),
                                    name="%s_image%02d" % (system.name, k)) for k in range(count)]


def config4(count=16, sigma=0.1, seed=20261206):
    """16 copies of the 30,000-atom solute system with 0.1 A Gaussian moves."""
    return perturbed_copies(config1(), count, sigma, seed)


def displacement_sequence(system, steps, sigma, seed):
    """Successive coordinates after cumulative Gaussian moves (sigma per call)."""
    rng = np.random.default_rng(seed)
    xyz = system.coordinates3.copy()
    out = []
    for _ in range(steps):
        xyz = xyz + rng.normal(0.0, sigma, xyz.shape)
        out.append(xyz.copy())
    return out


def random_small_system(natoms, box, seed, ntypes=3, exclusion_fraction=0.3):
    """Small random mixture for brute-force and edge-case tests."""
    rng = np.random.default_rng(seed)
    xyz = rng.uniform(0.0, box, size=(natoms, 3))
    q = rng.uniform(-0.8, 0.8, size=natoms)
    t = rng.integers(0, ntypes, size=natoms).astype(np.int32)
    eps = list(rng.uniform(0.1, 0.9, ntypes))
    sig = list(rng.uniform(2.5, 3.8, ntypes))
    a, b = lorentz_berthelot_table(eps, sig)
    m = int(exclusion_fraction * natoms)
    if natoms > 1 and m > 0:
        i = rng.integers(1, natoms, size=m)
        j = (i - rng.integers(1, 4, size=m)).clip(0)
        keep = i != j
        excl = np.unique(np.stack([i[keep], j[keep]], 1), axis=0).astype(np.int32)
    else:
        excl = np.zeros((0, 2), np.int32)
    return MMSystem("random%d" % natoms, xyz, q, t, ntypes, a, b, excl)


# Space groups as operation strings; the sets are those of
# pMolecule-1.9.0/tests/CrystalMMEnergies.py:31-40 plus P-1.
SPACE_GROUPS = {
    "P1": ["(x,y,z)"],
    "P-1": ["(x,y,z)", "(-x,-y,-z)"],
    "C2": ["(x,y,z)", "(-x,y,-z)"],
    "P21c": ["(x,y,z)", "(-x,y+1/2,-z+1/2)", "(-x,-y,-z)", "(x,-y+1/2,z+1/2)"],
    "P212121": ["(x,y,z)", "(-x+1/2,-y,z+1/2)", "(-x,y+1/2,-z+1/2)", "(x+1/2,-y+1/2,-z)"],
    "P61": ["(x,y,z)", "(-y,x-y,z+1/3)", "(-x+y,-x,z+2/3)", "(-x,-y,z+1/2)", "(y,-x+y,z+5/6)", "(x-y,x,z+1/6)"],
    "I4": ["(x,y,z)", "(-x,-y,z)", "(-y,x,z)", "(y,-x,z)", "(x+1/2,y+1/2,z+1/2)", "(-x+1/2,-y+1/2,z+1/2)",
           "(-y+1/2,x+1/2,z+1/2)", "(y+1/2,-x+1/2,z+1/2)"],
    "R3": ["(x,y,z)", "(z,x,y)", "(y,z,x)"],
}


def molecular_crystal(group, cell, nwater, seed, min_dist=2.7, name=None):
    """Asymmetric unit of ``nwater`` rigid waters in a unit cell with the given space group.

    cell = (a, b, c, alpha, beta, gamma).  Oxygens are drawn at random
    fractional positions and kept when at least ``min_dist`` from every
    symmetry and lattice copy of the oxygens already kept (and of themselves).
    Returns (MMSystem, (rotations (k,3,3), translations (k,3)), cell).
    """
    from .Symmetry import SymmetryParameters, Transformation3_FromSymmetryOperationString
    rng = np.random.default_rng(seed)
    ops = [Transformation3_FromSymmetryOperationString(s) for s in SPACE_GROUPS[group]]
    rot = np.array([o.rotation for o in ops])
    tra = np.array([o.translation for o in ops])
    sp = SymmetryParameters(*cell)
    M = sp.M
    r2 = (-2, -1, 0, 1, 2)
    shifts = np.array([[i, j, k] for i in r2 for j in r2 for k in r2], np.float64)

    def copies(f):
        fr = np.einsum("tij,j->ti", rot, f) + tra                     # (t, 3)
        fr = fr[:, None, :] + shifts[None, :, :]                      # (t, 125, 3)
        return fr.reshape(-1, 3) @ M.T

    kept, kept_copies = [], np.zeros((0, 3))
    tries = 0
    while len(kept) < nwater and tries < 20000:
        tries += 1
        f = rng.uniform(0.0, 1.0, 3)
        x0 = M @ f
        own = copies(f)
        own = own[np.linalg.norm(own - x0, axis=1) > 1.0e-9]
        pool = np.concatenate([kept_copies, own]) if kept_copies.size else own
        if pool.size and np.min(np.linalg.norm(pool - x0, axis=1)) < min_dist:
            continue
        kept.append(x0)
        kept_copies = np.concatenate([kept_copies, copies(f)])
    if len(kept) < nwater:
        raise ValueError("could not place %d waters in the cell" % nwater)
    xyz = _water_sites(rng, np.array(kept))
    n, q, t, excl = _water_params(nwater)
    a, b = opls_table([EPS_O_KCAL * KCAL_TO_KJ, 0.0], [SIGMA_O, 0.0])
    sysm = MMSystem(name or ("crystal_%s_%d" % (group, nwater)), xyz, q, t, 2, a, b, excl.astype(np.int32))
    return sysm, (rot, tra), tuple(float(v) for v in cell)


def periodic_water_box(nwater, box, seed, jitter=0.3, name=None):
    """Waters on a jittered lattice filling a cubic box, for P1 periodic images (cell = box)."""
    s = water_lattice(nwater, box, seed, jitter, name or "water_box%d" % (3 * nwater))
    rot = np.eye(3)[None]
    tra = np.zeros((1, 3))
    return s, (rot, tra), (float(box), float(box), float(box), 90.0, 90.0, 90.0)


# ------------------------------------------------------------------------------------------------
# A QC/MM system: a branched 17-atom chain in water with a QC region of six atoms and two link
# (boundary) atoms, the layout QCAtomContainer_FromAtomContainer
# (pMolecule-1.9.0/extensions/pyrex/pMolecule.QCAtomContainer.pyx:430-496) produces for it.
# ------------------------------------------------------------------------------------------------
CHAIN_BONDS = [(k, k + 1) for k in range(13)] + [(4, 14), (8, 15), (8, 16)]
CHAIN_QC_PURE = (3, 4, 5, 6, 7, 14)
CHAIN_BOUNDARY = {2: 3, 8: 7}        # boundary (link) atom -> its QC partner


def _chain_geometry(rng, nat, bonds, bond=1.53):
    """Heavy-atom chain with bond length 1.53 A, tetrahedral-like angles, no two atoms closer than 2.4 A
    unless bonded or sharing a neighbour."""
    nbr = [[] for _ in range(nat)]
    for i, j in bonds:
        nbr[i].append(j)
        nbr[j].append(i)
    x = np.zeros((nat, 3))
    placed = [0]
    order = []
    seen = {0}
    stack = [0]
    while stack:
        a = stack.pop(0)
        for b in sorted(nbr[a]):
            if b not in seen:
                seen.add(b)
                order.append((a, b))
                stack.append(b)
    for a, b in order:
        for _ in range(2000):
            v = rng.normal(size=3)
            v /= np.linalg.norm(v)
            prev = [p for p in nbr[a] if p in placed]
            ok = all(np.dot(v, (x[p] - x[a]) / np.linalg.norm(x[p] - x[a])) < -0.2 for p in prev)
            cand = x[a] + bond * v
            close = [p for p in placed if p != a and p not in nbr[a] and np.linalg.norm(cand - x[p]) < 2.4]
            if ok and not close:
                x[b] = cand
                break
        else:
            raise RuntimeError("chain geometry failed")
        placed.append(b)
    return x


def _bond_paths(nat, bonds):
    """1-2 and 1-3 pairs (the exclusions) and 1-4 pairs (not excluded) of a bond graph."""
    nbr = [set() for _ in range(nat)]
    for i, j in bonds:
        nbr[i].add(j)
        nbr[j].add(i)
    e12 = {tuple(sorted(b)) for b in bonds}
    e13 = set()
    for a in range(nat):
        for i in nbr[a]:
            for j in nbr[a]:
                if i < j:
                    e13.add((i, j))
    excl = e12 | e13
    p14 = set()
    for i, j in e12:
        for a in nbr[i] - {j}:
            for b in nbr[j] - {i}:
                if a != b:
                    key = tuple(sorted((a, b)))
                    if key not in excl:
                        p14.add(key)
    return sorted(excl), sorted(p14)


def qcmm_chain_in_water(nwater=1000, box=31.0, seed=20261207, name=None, qc_waters=0):
    """Returns (MMSystem, qc): the chain (atoms 0-16) followed by the waters; qc describes the QC atoms as
    pDynamo's QCAtomContainer holds them (index order, boundary flags, QC partners, MM partners of each
    boundary atom).  qc_waters > 0 adds that many whole waters nearest the chain to the QC region."""
    rng = np.random.default_rng(seed)
    nch = 17
    xc = _chain_geometry(rng, nch, CHAIN_BONDS)
    xc += 0.5 * box - xc.mean(0)
    excl_c, p14_c = _bond_paths(nch, CHAIN_BONDS)
    per = int(math.ceil((nwater * 1.6) ** (1.0 / 3.0)))
    cen = _lattice_centres(rng, box, per, per ** 3, 0.25)
    d = np.min(np.linalg.norm(cen[:, None, :] - xc[None, :, :], axis=2), axis=1)
    cen = cen[d > 3.0]
    cen = cen[np.sort(rng.choice(cen.shape[0], size=min(nwater, cen.shape[0]), replace=False))]
    nw = cen.shape[0]
    w = _water_sites(rng, cen)
    _, qw, tw, ew = _water_params(nw)
    xyz = np.concatenate([xc, w])
    qch = rng.uniform(-0.45, 0.45, nch)
    qch -= qch.mean()
    charges = np.concatenate([qch, qw])
    # LJ types: water O (0), water H (1), chain heavy atoms (2, 3).
    tch = np.where(np.arange(nch) % 2 == 0, 2, 3).astype(np.int32)
    ljtypes = np.concatenate([tch, tw]).astype(np.int32)
    eps = [EPS_O_KCAL * KCAL_TO_KJ, 0.0, 0.35, 0.28]
    sig = [SIGMA_O, 1.0, 3.5, 3.3]
    tA, tB = lorentz_berthelot_table(eps, sig)
    excl = np.concatenate([np.array(excl_c, np.int32).reshape(-1, 2), ew + nch]).astype(np.int32)
    p14 = np.array(p14_c, np.int32).reshape(-1, 2)
    s = MMSystem(name or f"qcmm_chain_{nw}w", np.ascontiguousarray(xyz), charges, ljtypes, 4, tA, tB, excl, p14)
    nbr = [set() for _ in range(nch)]
    for i, j in CHAIN_BONDS:
        nbr[i].add(j)
        nbr[j].add(i)
    qc_set = set(CHAIN_QC_PURE) | set(CHAIN_BOUNDARY)
    if qc_waters > 0:
        dw = np.linalg.norm(w[0::3] - xc.mean(0), axis=1)
        for k in np.argsort(dw)[:qc_waters]:
            qc_set |= {nch + 3 * int(k), nch + 3 * int(k) + 1, nch + 3 * int(k) + 2}
    qc_atoms = sorted(qc_set)
    boundary = np.array([a in CHAIN_BOUNDARY for a in qc_atoms])
    partner = np.array([CHAIN_BOUNDARY.get(a, -1) for a in qc_atoms], np.int32)
    mmp = [sorted(nbr[a] - set(qc_atoms)) if a in CHAIN_BOUNDARY else [] for a in qc_atoms]
    off = np.concatenate([[0], np.cumsum([len(m) for m in mmp])]).astype(np.int32)
    qc = dict(index=np.array(qc_atoms, np.int32), boundary=boundary, partner=partner,
              mmp_off=off, mmp=np.array(sum(mmp, []), np.int32))
    return s, qc
