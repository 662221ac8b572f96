"""ctypes front end of the CPU checker libraries.

TEST INFRASTRUCTURE ONLY.  Imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py -- never by the product package.

* ``liboracle.so``            -- the FP64 C restatement (abfs_oracle.c).
* ``_ref/libpdynamo_ref.so``  -- the reference pDynamo C library compiled in
                                 place plus ref_driver.c (present only where it
                                 was built; it travels with the tree).

``OracleABFS`` replays NBModelABFS_Update + NBModelABFS_MMMMEnergy
(pMolecule-1.9.0/extensions/csource/NBModelABFS.c:228-301,508-623) on the
restatement so tests can compare whole System.Energy calls.
"""

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_lp = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_u8 = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")


def build():
    """Compile liboracle.so (and the reference library where the tree exists)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)
    if os.path.isdir("/root/reference/pCore-1.9.0/extensions/csource"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        L.orc_coulomb_factor.restype = C.c_double
        L.orc_make_factors.argtypes = [C.c_double, C.c_double, C.c_double, _dp]
        L.orc_pair_terms.argtypes = [_dp, C.c_double, C.c_double, C.c_double, C.c_double,
                                     C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.orc_determine_method.argtypes = [C.c_int, _dp, C.c_double, C.c_double, C.c_int, C.c_int]
        L.orc_make_grid.argtypes = [C.c_int, _dp, C.c_double, _ip, _dp, _ip]
        L.orc_make_grid.restype = C.c_long
        L.orc_search_range.argtypes = [C.c_double, C.c_double, C.c_int, C.c_void_p, C.c_void_p]
        L.orc_self_pair_list.argtypes = [C.c_int, _dp, C.c_void_p, C.c_void_p, C.c_double, C.c_double, C.c_int,
                                         C.c_void_p, C.c_long, C.c_long, _ip, _lp, _ip, C.POINTER(C.c_long)]
        L.orc_self_pair_list.restype = C.c_long
        L.orc_check_for_update.argtypes = [C.c_int, _dp, _dp, C.c_void_p, C.c_double, C.c_double,
                                           C.POINTER(C.c_double)]
        L.orc_mmmm_energy.argtypes = [_dp, C.c_int, _dp, _dp, _ip, C.c_int, _dp, _dp, _ip, C.c_double, C.c_double,
                                      C.c_long, _ip, _lp, _ip, C.c_int, C.POINTER(C.c_double),
                                      C.POINTER(C.c_double), C.c_void_p]
        L.orc_mmmm_abs_sums.argtypes = [_dp, C.c_int, _dp, _dp, _ip, C.c_int, _dp, _dp, _ip, C.c_double, C.c_double,
                                        C.c_long, _ip, _lp, _ip, C.c_int, C.POINTER(C.c_double),
                                        C.POINTER(C.c_double)]
        L.orc_lj_make_table.argtypes = [C.c_int, _dp, _dp, C.c_int, C.c_int, _dp, _dp, _ip]
        L.orc_symmetry_make_m.argtypes = [_dp, _dp, _dp]
        L.orc_orthogonalize.argtypes = [_dp, _dp, _dp, _dp, _dp, _dp]
        L.orc_find_inverses.argtypes = [C.c_int, _dp, _dp, _ip]
        L.orc_image_lists.restype = C.c_long
        L.orc_image_lists.argtypes = [C.c_int, _dp, C.c_int, C.c_double, C.c_int, _dp, _dp, _dp, C.c_double, C.c_int,
                                      C.c_int,
                                      C.c_long, C.c_void_p, C.c_void_p, C.c_void_p, C.c_long, C.c_long, C.c_void_p,
                                      C.c_void_p, C.c_void_p, _lp]
        L.orc_image_energy.argtypes = [_dp, C.c_int, _dp, _dp, _ip, C.c_int, _dp, _dp, _ip, C.c_double, _dp, _dp, _dp,
                                       C.c_long, _ip, _dp, _lp, _ip, _lp, _ip, C.POINTER(C.c_double),
                                       C.POINTER(C.c_double), C.c_void_p, C.c_void_p]
        L.orc_crystal_derivatives.argtypes = [_dp, _dp, _dp]
        _LIB = L
    return _LIB


COULOMB = None


def coulomb_factor():
    return lib().orc_coulomb_factor()


def factors(damp=0.5, inner=8.0, outer=12.0):
    f = np.zeros(21)
    lib().orc_make_factors(damp, inner, outer, f)
    return f


FACTOR_NAMES = ["r2Damp", "r2Off", "r2On", "a", "b", "c", "d", "qShift1", "qShift2", "qF0", "qAlpha",
                "aF6", "aK12", "aShift12", "aF0", "aAlpha", "bF3", "bK6", "bShift6", "bF0", "bAlpha"]


def pair_terms(f, r2, qij, aij, bij):
    e1, e2, df = C.c_double(), C.c_double(), C.c_double()
    ok = lib().orc_pair_terms(f, r2, qij, aij, bij, C.byref(e1), C.byref(e2), C.byref(df))
    if not ok:
        return None
    return e1.value, e2.value, df.value


def exclusion_csr(n, pairs):
    """Symmetric CSR with ascending rows, duplicates and self pairs dropped
    (PairList_FromIntegerPairArray + SelfPairList_MakeConnections,
    pCore-1.9.0/extensions/csource/PairList.c:107-189,458-526)."""
    pairs = np.asarray(pairs, dtype=np.int64).reshape(-1, 2)
    pairs = pairs[pairs[:, 0] != pairs[:, 1]]
    hi = np.maximum(pairs[:, 0], pairs[:, 1])
    lo = np.minimum(pairs[:, 0], pairs[:, 1])
    key = np.unique(hi * max(n, 1) + lo)
    hi, lo = key // max(n, 1), key % max(n, 1)
    rows = np.concatenate([hi, lo])
    cols = np.concatenate([lo, hi])
    order = np.lexsort((cols, rows))
    rows, cols = rows[order], cols[order]
    ptr = np.zeros(n + 1, np.int32)
    np.add.at(ptr, rows + 1, 1)
    return np.cumsum(ptr).astype(np.int32), cols.astype(np.int32)


ORC_INVALID_ARGUMENT = -(1 << 62)      # abfs_oracle.h


def determine_method(xyz, cutoff=13.5, factor=0.5, min_extent=2, min_points=500):
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    return bool(lib().orc_determine_method(xyz.shape[0], xyz, cutoff, factor, min_extent, min_points))


def make_grid(xyz, cell_size):
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    bins = np.zeros(3, np.int32)
    lower = np.zeros(3)
    cell = np.zeros(max(xyz.shape[0], 1), np.int32)
    lib().orc_make_grid(xyz.shape[0], xyz, cell_size, bins, lower, cell)
    return bins, lower, cell[:xyz.shape[0]]


def search_range(cell_size, cutoff):
    n = lib().orc_search_range(cell_size, cutoff, 0, None, None)
    off = np.zeros((n, 3), np.int32)
    flg = np.zeros(n, np.int32)
    lib().orc_search_range(cell_size, cutoff, n, off.ctypes.data, flg.ctypes.data)
    return off, flg


class PairRecords:
    """Reference-ordered records (i, [j ...]) of a self pair list."""

    def __init__(self, rec_i, rec_ptr, js, n):
        self.rec_i, self.rec_ptr, self.js, self.n = rec_i, rec_ptr, js, n

    @property
    def npairs(self):
        return int(self.js.shape[0])

    def pairs(self):
        counts = np.diff(self.rec_ptr)
        return np.repeat(self.rec_i, counts), self.js

    def canonical(self):
        """Sorted int64 keys max*n + min (orientation-free pair set)."""
        i, j = self.pairs()
        i = i.astype(np.int64)
        j = j.astype(np.int64)
        return np.sort(np.maximum(i, j) * self.n + np.minimum(i, j))


def self_pair_list(xyz, xptr, xidx, cutoff=13.5, cell_size=6.75, method=1, qand=None):
    """method: 0 direct, 1 grid cell/cell, 2 grid point/cell; qand: the point/cell AND selection (bool per atom)."""
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    n = xyz.shape[0]
    qa = None if qand is None else np.ascontiguousarray(qand, dtype=np.uint8)
    cap_r, cap_p = max(n, 1), max(16 * n, 1024)
    while True:
        rec_i = np.zeros(cap_r, np.int32)
        rec_ptr = np.zeros(cap_r + 1, np.int64)
        js = np.zeros(cap_p, np.int32)
        nrec = C.c_long()
        npairs = lib().orc_self_pair_list(n, xyz, None if xptr is None else xptr.ctypes.data,
                                          None if xidx is None else xidx.ctypes.data,
                                          cutoff, cell_size, int(method), None if qa is None else qa.ctypes.data,
                                          cap_r, cap_p, rec_i, rec_ptr, js,
                                          C.byref(nrec))
        if npairs >= 0:
            r = nrec.value
            return PairRecords(rec_i[:r].copy(), rec_ptr[:r + 1].copy(), js[:npairs].copy(), n)
        if npairs == ORC_INVALID_ARGUMENT:
            raise ValueError("Status_InvalidArgument: the cell grid has an axis with no cells (zero extent)")
        cap_p = max(cap_p, -npairs) + 16


def check_for_update(xyz, ref, list_cutoff=13.5, outer_cutoff=12.0, fixed=None):
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    ref = np.ascontiguousarray(ref, dtype=np.float64)
    md = C.c_double()
    fx = None if fixed is None else np.ascontiguousarray(fixed, dtype=np.uint8)
    up = lib().orc_check_for_update(xyz.shape[0], xyz, ref, None if fx is None else fx.ctypes.data,
                                    list_cutoff, outer_cutoff, C.byref(md))
    return bool(up), md.value


def full_table(ntypes, tableA, tableB):
    """Expand packed triangular tables to ntypes x ntypes plus tableindex."""
    idx = np.zeros((ntypes, ntypes), np.int32)
    for i in range(ntypes):
        for j in range(ntypes):
            a, b = max(i, j), min(i, j)
            idx[i, j] = a * (a + 1) // 2 + b
    return idx.reshape(-1)


def mmmm_energy(f, xyz, charges, ljtypes, ntypes, tableA, tableB, records, electrostatic_scale=1.0,
                lj_scale=1.0, nthreads=0, grad=None):
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    n = xyz.shape[0]
    tidx = full_table(ntypes, tableA, tableB)
    eel, elj = C.c_double(), C.c_double()
    g = None
    if grad is not None:
        g = grad
        assert g.dtype == np.float64 and g.flags.c_contiguous and g.shape == (n, 3)
    lib().orc_mmmm_energy(f, n, xyz, np.ascontiguousarray(charges, np.float64),
                          np.ascontiguousarray(ljtypes, np.int32), ntypes,
                          np.ascontiguousarray(tableA, np.float64), np.ascontiguousarray(tableB, np.float64),
                          tidx, electrostatic_scale, lj_scale, records.rec_i.shape[0], records.rec_i,
                          records.rec_ptr, records.js, nthreads, C.byref(eel), C.byref(elj),
                          None if g is None else g.ctypes.data)
    return eel.value, elj.value


def mmmm_abs_sums(f, xyz, charges, ljtypes, ntypes, tableA, tableB, records, electrostatic_scale=1.0, nthreads=0):
    """(sum |e_el|, sum |e_lj|) of the pair terms of `records`: the scale of the rounding error of an FP64 sum
    of the energies in any order."""
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    tidx = full_table(ntypes, tableA, tableB)
    a1, a2 = C.c_double(), C.c_double()
    lib().orc_mmmm_abs_sums(f, xyz.shape[0], xyz, np.ascontiguousarray(charges, np.float64),
                            np.ascontiguousarray(ljtypes, np.int32), ntypes,
                            np.ascontiguousarray(tableA, np.float64), np.ascontiguousarray(tableB, np.float64), tidx,
                            electrostatic_scale, 1.0, records.rec_i.shape[0], records.rec_i, records.rec_ptr,
                            records.js, nthreads, C.byref(a1), C.byref(a2))
    return a1.value, a2.value


def records_from_pairs(n, pairs):
    """Records (i > j) from an integer pair array, as PairList_FromIntegerPairArray."""
    pairs = np.asarray(pairs, dtype=np.int64).reshape(-1, 2)
    if pairs.shape[0] == 0:
        return PairRecords(np.zeros(0, np.int32), np.zeros(1, np.int64), np.zeros(0, np.int32), n)
    hi = np.maximum(pairs[:, 0], pairs[:, 1])
    lo = np.minimum(pairs[:, 0], pairs[:, 1])
    keep = hi != lo
    key = np.unique(hi[keep] * max(n, 1) + lo[keep])
    hi, lo = key // max(n, 1), key % max(n, 1)
    ri, counts = np.unique(hi, return_counts=True)
    ptr = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    return PairRecords(ri.astype(np.int32), ptr, lo.astype(np.int32), n)


# ----------------------------------------------------------------------------------------------------------------
# Crystal symmetry: image lists and image energy.
# ----------------------------------------------------------------------------------------------------------------
def symmetry_make_m(sp):
    """M (columns = lattice vectors) and M^-1 (SymmetryParameters_MakeM)."""
    M, Mi = np.zeros(9), np.zeros(9)
    lib().orc_symmetry_make_m(np.asarray(sp, np.float64), M, Mi)
    return M.reshape(3, 3), Mi.reshape(3, 3)


def orthogonalize(sp, rot, tr):
    """Real-space R' = M R M^-1 and T' = M t of a fractional transformation."""
    M, Mi = symmetry_make_m(sp)
    Ro, To = np.zeros(9), np.zeros(3)
    lib().orc_orthogonalize(M.reshape(-1).copy(), Mi.reshape(-1).copy(), np.ascontiguousarray(rot, np.float64).reshape(-1),
                            np.ascontiguousarray(tr, np.float64), Ro, To)
    return Ro.reshape(3, 3), To


def find_inverses(rot, tr):
    """(identity index, inverses) of a transformation set (rotations only)."""
    rot = np.ascontiguousarray(rot, np.float64).reshape(-1)
    tr = np.ascontiguousarray(tr, np.float64).reshape(-1)
    inv = np.zeros(tr.shape[0] // 3, np.int32)
    ident = lib().orc_find_inverses(inv.shape[0], rot, tr, inv)
    return ident, inv


def crystal_derivatives(sp, dEdM):
    """dE/d(a, b, c, alpha, beta, gamma) from dE/dM (SymmetryParameterGradients_CrystalDerivatives)."""
    out = np.zeros(6)
    lib().orc_crystal_derivatives(np.asarray(sp, np.float64), np.ascontiguousarray(dEdM, np.float64).reshape(-1), out)
    return out


class ImageLists:
    """MM/MM image list: abct (k, 4) = a, b, c, t; scale (k); records of every image back to back,
    image k owning records img_rec[k]..img_rec[k+1]."""

    def __init__(self, n, abct, scale, img_rec, records):
        self.n, self.abct, self.scale, self.img_rec, self.records = n, abct, scale, img_rec, records

    @property
    def nimages(self):
        return self.abct.shape[0]

    @property
    def npairs(self):
        return self.records.npairs

    def image_pairs(self, k):
        r = self.records
        r0, r1 = self.img_rec[k], self.img_rec[k + 1]
        cnt = np.diff(r.rec_ptr[r0:r1 + 1])
        i = np.repeat(r.rec_i[r0:r1], cnt)
        j = r.js[r.rec_ptr[r0]:r.rec_ptr[r1]]
        return i.astype(np.int64), j.astype(np.int64)

    def image_keys(self, k):
        """Sorted i * n + j keys of image k (i primary, j image)."""
        i, j = self.image_pairs(k)
        return np.sort(i * self.n + j)


def image_lists(xyz, transformations, sp, cutoff=13.5, use_grid=True, cell_size=6.75, check_for_inverses=True,
                expand_factor=0, use_grid_by_cell=True):
    """GenerateImageLists (MM/MM) on the restatement."""
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    n = xyz.shape[0]
    rot = np.ascontiguousarray(transformations[0], np.float64).reshape(-1)
    tr = np.ascontiguousarray(transformations[1], np.float64).reshape(-1)
    method = (1 if use_grid_by_cell else 2) if use_grid else 0
    sp = np.asarray(sp, np.float64)
    sizes = np.zeros(3, np.int64)
    L = lib()
    args = [n, xyz, method, cell_size, tr.shape[0] // 3, rot, tr, sp, cutoff, 1 if check_for_inverses else 0,
            int(expand_factor)]
    L.orc_image_lists(*args, 0, None, None, None, 0, 0, None, None, None, sizes)
    ni, nr, npr = (int(v) for v in sizes)
    abct = np.zeros(4 * max(ni, 1), np.int32)
    scale = np.zeros(max(ni, 1))
    img_rec = np.zeros(ni + 1, np.int64)
    rec_i = np.zeros(max(nr, 1), np.int32)
    rec_ptr = np.zeros(nr + 1, np.int64)
    js = np.zeros(max(npr, 1), np.int32)
    got = L.orc_image_lists(*args, ni, abct.ctypes.data, scale.ctypes.data, img_rec.ctypes.data, nr, npr,
                            rec_i.ctypes.data, rec_ptr.ctypes.data, js.ctypes.data, sizes)
    assert got == ni
    return ImageLists(n, abct[:4 * ni].reshape(-1, 4), scale[:ni], img_rec,
                      PairRecords(rec_i[:nr], rec_ptr, js[:npr], n))


def image_energy(f, xyz, charges, ljtypes, ntypes, tableA, tableB, transformations, sp, images, electrostatic_scale=1.0,
                 grad=None, dEdM=None):
    """MMMMImageEnergy on the restatement; grad (n,3) and dEdM (3,3) are incremented when given."""
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    n = xyz.shape[0]
    tidx = full_table(ntypes, tableA, tableB)
    rot = np.ascontiguousarray(transformations[0], np.float64).reshape(-1)
    tr = np.ascontiguousarray(transformations[1], np.float64).reshape(-1)
    eel, elj = C.c_double(), C.c_double()
    if grad is not None:
        assert grad.dtype == np.float64 and grad.flags.c_contiguous and grad.shape == (n, 3)
    dm = None
    if dEdM is not None:
        dm = np.ascontiguousarray(dEdM, np.float64).reshape(9).copy()
    r = images.records
    lib().orc_image_energy(f, n, xyz, np.ascontiguousarray(charges, np.float64), np.ascontiguousarray(ljtypes, np.int32),
                           ntypes, np.ascontiguousarray(tableA, np.float64), np.ascontiguousarray(tableB, np.float64),
                           tidx, electrostatic_scale, rot, tr, np.asarray(sp, np.float64), images.nimages,
                           np.ascontiguousarray(images.abct, np.int32).reshape(-1), np.ascontiguousarray(images.scale),
                           np.ascontiguousarray(images.img_rec, np.int64), r.rec_i, r.rec_ptr, r.js,
                           C.byref(eel), C.byref(elj), None if grad is None else grad.ctypes.data,
                           None if dm is None else dm.ctypes.data)
    if dEdM is not None:
        dEdM[...] = dm.reshape(dEdM.shape)
    return eel.value, elj.value


class OracleABFS:
    """NBModelABFS state machine on the FP64 restatement.

    Mirrors NBModelABFS_Update (NBModelABFS.c:508-623): the first call is a
    rebuild (and decides grid vs direct once, NBModelABFSState.c:207-226);
    later calls rebuild when any free atom moved more than half the skin.
    Energies follow NBModelABFS_MMMMEnergy (NBModelABFS.c:228-301): the main
    list with 1/dielectric, the 1-4 list with scale14/dielectric and lj14.
    """

    def __init__(self, system, damping=0.5, inner=8.0, outer=12.0, list_cutoff=13.5, dielectric=1.0,
                 scale14=1.0, cell_factor=0.5, min_extent=2, min_points=500, tableA14=None, tableB14=None,
                 fixed=None, nthreads=0, transformations=None, check_for_inverses=True, image_expand_factor=0,
                 use_grid_by_cell=True):
        self.s = system
        self.use_grid_by_cell = use_grid_by_cell
        self.transformations = transformations
        self.check_for_inverses = check_for_inverses
        self.image_expand_factor = image_expand_factor
        self.images = None
        self.ref_sp = None
        self.max_displacement = 0.0
        self.f = factors(damping, inner, outer)
        self.outer, self.list_cutoff = outer, list_cutoff
        self.dielectric, self.scale14 = dielectric, scale14
        self.cell_factor, self.min_extent, self.min_points = cell_factor, min_extent, min_points
        self.xptr, self.xidx = exclusion_csr(system.natoms, system.exclusions)
        self.rec14 = records_from_pairs(system.natoms, system.interactions14)
        self.tA14 = system.tableA if tableA14 is None else tableA14
        self.tB14 = system.tableB if tableB14 is None else tableB14
        self.fixed = fixed
        self.nthreads = nthreads
        self.ref = None
        self.records = None
        self.use_grid = None
        self.calls = 0
        self.updates = 0

    def update(self, xyz):
        rebuild = self.ref is None
        self.max_displacement = 0.0
        if not rebuild:
            rebuild, self.max_displacement = check_for_update(xyz, self.ref, self.list_cutoff, self.outer, self.fixed)
        if rebuild:
            if self.use_grid is None:
                self.use_grid = determine_method(xyz, self.list_cutoff, self.cell_factor, self.min_extent,
                                                 self.min_points)
            if self.fixed is not None:
                raise NotImplementedError("fixed atoms are checked against the reference driver instead")
            method = (1 if self.use_grid_by_cell else 2) if self.use_grid else 0
            self.records = self_pair_list(xyz, self.xptr, self.xidx, self.list_cutoff,
                                          self.cell_factor * self.list_cutoff, method)
            self.ref = np.array(xyz, dtype=np.float64, copy=True)
            self.updates += 1
        return rebuild

    def check_for_image_update(self, sp):
        """CheckForImageUpdate (pMolecule-1.9.0/extensions/csource/NBModelABFS.c:635-684): some image's lattice
        translation moved by more than listCutoff - outerCutoff - the largest atom displacement."""
        if self.images is None or self.images.nimages == 0:
            return False
        buffac = self.list_cutoff - self.outer - self.max_displacement
        M, _ = symmetry_make_m(sp)
        Mr, _ = symmetry_make_m(self.ref_sp)
        dM = M - Mr
        tr = np.asarray(self.transformations[1], np.float64).reshape(-1, 3)
        for a, b, c, t in self.images.abct:
            d = tr[t] + np.array([a, b, c], np.float64)
            if np.linalg.norm(dM @ d) > buffac:
                return True
        return False

    def energy_sym(self, xyz, sp, gradients=True):
        """A call with crystal parameters sp = (a, b, c, alpha, beta, gamma): the update of NBModelABFS_Update
        (NBModelABFS.c:508-623, image lists rebuilt with the main list or when CheckForImageUpdate holds), then
        the image energy before the main and 1-4 lists (NBModelABFS.c:228-301).
        Returns (terms, gradient, dEdM (3x3), rebuilt)."""
        xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        self.calls += 1
        rebuilt = self.update(xyz)
        if rebuilt or self.check_for_image_update(sp):
            self.images = image_lists(xyz, self.transformations, sp, self.list_cutoff, bool(self.use_grid),
                                      self.cell_factor * self.list_cutoff, self.check_for_inverses,
                                      self.image_expand_factor, self.use_grid_by_cell)
            self.ref_sp = np.array(sp, np.float64)
            rebuilt = True
        s = self.s
        g = np.zeros((s.natoms, 3)) if gradients else None
        dm = np.zeros((3, 3)) if gradients else None
        iterms = []
        if self.images.nimages > 0:
            e1, e2 = image_energy(self.f, xyz, s.charges, s.ljtypes, s.ntypes, s.tableA, s.tableB, self.transformations,
                                  sp, self.images, 1.0 / self.dielectric, g, dm)
            iterms = [("MM/MM Image Elect.", e1), ("MM/MM Image LJ", e2)]
        self.calls -= 1
        terms, g2, _ = self.energy(xyz, gradients, _rebuilt=rebuilt)
        if gradients:
            g += g2
        return terms + iterms, g, dm, rebuilt

    def energy(self, xyz, gradients=True, _rebuilt=None):
        """Returns (terms, gradient, rebuilt) with pDynamo's labels."""
        xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        self.calls += 1
        rebuilt = self.update(xyz) if _rebuilt is None else _rebuilt
        s = self.s
        g = np.zeros((s.natoms, 3)) if gradients else None
        terms = []
        escale = 1.0 / self.dielectric
        if self.records.npairs > 0:
            e1, e2 = mmmm_energy(self.f, xyz, s.charges, s.ljtypes, s.ntypes, s.tableA, s.tableB, self.records,
                                 escale, 1.0, self.nthreads, g)
            terms += [("MM/MM Elect.", e1), ("MM/MM LJ", e2)]
        if self.rec14.npairs > 0:
            e1, e2 = mmmm_energy(self.f, xyz, s.charges, s.ljtypes, s.ntypes, self.tA14, self.tB14, self.rec14,
                                 escale * self.scale14, 1.0, 1, g)
            terms += [("MM/MM 1-4 Elect.", e1), ("MM/MM 1-4 LJ", e2)]
        return terms, g, rebuilt

    def abs_sums(self, xyz):
        """{label: sum of |pair term|} of the current lists at xyz (the rounding scale of each term)."""
        s = self.s
        out = {}
        escale = 1.0 / self.dielectric
        if self.records is not None and self.records.npairs > 0:
            a1, a2 = mmmm_abs_sums(self.f, xyz, s.charges, s.ljtypes, s.ntypes, s.tableA, s.tableB, self.records,
                                   escale, self.nthreads)
            out["MM/MM Elect."], out["MM/MM LJ"] = a1, a2
        if self.rec14.npairs > 0:
            a1, a2 = mmmm_abs_sums(self.f, xyz, s.charges, s.ljtypes, s.ntypes, self.tA14, self.tB14, self.rec14,
                                   escale * self.scale14, 1)
            out["MM/MM 1-4 Elect."], out["MM/MM 1-4 LJ"] = a1, a2
        return out


# ----------------------------------------------------------------------------------------------------------------
# The compiled reference (oracle/_ref).
# ----------------------------------------------------------------------------------------------------------------
def ref_path():
    return os.path.join(HERE, "_ref", "libpdynamo_ref.so")


def ref_available():
    return os.path.exists(ref_path())


def ref_lib():
    global _REF
    if _REF is None:
        L = C.CDLL(ref_path())
        L.ref_create.restype = C.c_void_p
        L.ref_create.argtypes = [C.c_int, _dp, _ip, C.c_int, _dp, _dp, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p,
                                 C.c_int, C.c_void_p, _dp, _dp, C.c_int]
        L.ref_energy.argtypes = [C.c_void_p, _dp, C.c_void_p, _dp, _ip]
        L.ref_pairs.restype = C.c_long
        L.ref_pairs.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_long]
        L.ref_use_grid.argtypes = [C.c_void_p]
        L.ref_destroy.argtypes = [C.c_void_p]
        L.ref_create_sym.restype = C.c_void_p
        L.ref_create_sym.argtypes = L.ref_create.argtypes + [C.c_int, _dp, _dp, C.c_int, C.c_int]
        L.ref_energy_sym.argtypes = [C.c_void_p, _dp, _dp, C.c_void_p, _dp, _dp, _ip]
        L.ref_images.restype = C.c_long
        L.ref_images.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_long]
        L.ref_image_pairs.restype = C.c_long
        L.ref_image_pairs.argtypes = [C.c_void_p, C.c_long, C.c_void_p, C.c_void_p, C.c_long]
        _REF = L
    return _REF


class ReferenceABFS:
    """The reference C library driven call by call (see ref_driver.c)."""

    def __init__(self, system, damping=0.5, inner=8.0, outer=12.0, list_cutoff=13.5, dielectric=1.0,
                 scale14=1.0, cell_factor=0.5, min_extent=2, min_points=500, tableA14=None, tableB14=None,
                 nthreads=1, transformations=None, check_for_inverses=True, image_expand_factor=0,
                 use_grid_by_cell=True):
        """transformations: None, or (rotations (k,3,3), translations (k,3)) in fractional form.
        use_grid_by_cell: the generator's useGridByCell (cell/cell or point/cell grid search)."""
        L = ref_lib()
        s = system
        self.n = s.natoms
        ex = np.ascontiguousarray(s.exclusions, np.int32).reshape(-1)
        p14 = np.ascontiguousarray(s.interactions14, np.int32).reshape(-1)
        self._keep = [ex, p14]
        tA14 = None if tableA14 is None else np.ascontiguousarray(tableA14, np.float64)
        tB14 = None if tableB14 is None else np.ascontiguousarray(tableB14, np.float64)
        self._keep += [tA14, tB14]
        opts = np.array([damping, inner, outer, list_cutoff, dielectric, scale14])
        gen = np.array([cell_factor, min_extent, min_points, 1.0 if use_grid_by_cell else 0.0], dtype=np.float64)
        args = [self.n, np.ascontiguousarray(s.charges, np.float64), np.ascontiguousarray(s.ljtypes, np.int32),
                s.ntypes, np.ascontiguousarray(s.tableA, np.float64), np.ascontiguousarray(s.tableB, np.float64),
                None if tA14 is None else tA14.ctypes.data, None if tB14 is None else tB14.ctypes.data,
                ex.shape[0] // 2, ex.ctypes.data if ex.size else None,
                p14.shape[0] // 2, p14.ctypes.data if p14.size else None, opts, gen, nthreads]
        self.symmetric = transformations is not None
        if self.symmetric:
            rot = np.ascontiguousarray(transformations[0], np.float64).reshape(-1)
            tra = np.ascontiguousarray(transformations[1], np.float64).reshape(-1)
            self._keep += [rot, tra]
            self.h = L.ref_create_sym(*args, tra.shape[0] // 3, rot, tra, 1 if check_for_inverses else 0,
                                      int(image_expand_factor))
        else:
            self.h = L.ref_create(*args)

    def energy_sym(self, xyz, symmetry_parameters, gradients=True):
        """Returns (terms, gradient, dEdM (3x3), rebuilt) for a call with crystal parameters
        (a, b, c, alpha, beta, gamma)."""
        xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        e = np.zeros(6)
        present = np.zeros(3, np.int32)
        dm = np.zeros(9)
        g = np.zeros((self.n, 3)) if gradients else None
        up = ref_lib().ref_energy_sym(self.h, xyz, np.asarray(symmetry_parameters, np.float64),
                                      None if g is None else g.ctypes.data, e, dm, present)
        if up < 0:
            raise RuntimeError("reference library error")
        terms = []
        if present[0]:
            terms += [("MM/MM Elect.", e[0]), ("MM/MM LJ", e[1])]
        if present[1]:
            terms += [("MM/MM 1-4 Elect.", e[2]), ("MM/MM 1-4 LJ", e[3])]
        if present[2]:
            terms += [("MM/MM Image Elect.", e[4]), ("MM/MM Image LJ", e[5])]
        return terms, g, dm.reshape(3, 3), bool(up)

    def images(self):
        """[(a, b, c, t, scale, npairs)] of the MM/MM image list, in list order."""
        L = ref_lib()
        m = L.ref_images(self.h, None, None, None, 0)
        abct = np.zeros(4 * max(m, 1), np.int32)
        sc = np.zeros(max(m, 1))
        npr = np.zeros(max(m, 1), np.int64)
        L.ref_images(self.h, abct.ctypes.data, sc.ctypes.data, npr.ctypes.data, m)
        return [(int(abct[4 * k]), int(abct[4 * k + 1]), int(abct[4 * k + 2]), int(abct[4 * k + 3]), float(sc[k]),
                 int(npr[k])) for k in range(m)]

    def image_pairs(self, k):
        """(i, j) pairs of image k: i in the primary image, j in image k."""
        L = ref_lib()
        m = L.ref_image_pairs(self.h, k, None, None, 0)
        oi = np.zeros(max(m, 1), np.int32)
        oj = np.zeros(max(m, 1), np.int32)
        L.ref_image_pairs(self.h, k, oi.ctypes.data, oj.ctypes.data, m)
        return oi[:m], oj[:m]

    def energy(self, xyz, gradients=True):
        xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        e = np.zeros(4)
        present = np.zeros(2, np.int32)
        g = np.zeros((self.n, 3)) if gradients else None
        up = ref_lib().ref_energy(self.h, xyz, None if g is None else g.ctypes.data, e, present)
        if up < 0:
            raise RuntimeError("reference library error")
        terms = []
        if present[0]:
            terms += [("MM/MM Elect.", e[0]), ("MM/MM LJ", e[1])]
        if present[1]:
            terms += [("MM/MM 1-4 Elect.", e[2]), ("MM/MM 1-4 LJ", e[3])]
        return terms, g, bool(up)

    def pairs(self, which=0):
        L = ref_lib()
        m = L.ref_pairs(self.h, which, None, None, 0)
        oi = np.zeros(max(m, 1), np.int32)
        oj = np.zeros(max(m, 1), np.int32)
        L.ref_pairs(self.h, which, oi.ctypes.data, oj.ctypes.data, m)
        return oi[:m], oj[:m]

    def canonical(self, which=0):
        i, j = self.pairs(which)
        i = i.astype(np.int64)
        j = j.astype(np.int64)
        return np.sort(np.maximum(i, j) * self.n + np.minimum(i, j))

    @property
    def use_grid(self):
        return bool(ref_lib().ref_use_grid(self.h))

    def close(self):
        if self.h:
            ref_lib().ref_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
