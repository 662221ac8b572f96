"""QC/MM part of the ABFS model, restated in numpy.

TEST INFRASTRUCTURE ONLY.  Imported by tests/ (and the fixture scripts under
tests/golden/) as the checker of the QC/MM path -- never by the product.

Restates, in FP64 with the reference's expression order where it matters:

* the QC/MM electrostatic spline of PairwiseInteractionABFS.MakeSplines for
  the "Delta/Gaussian" and "Gaussian/Gaussian" models
  (pMolecule-1.9.0/extensions/pyrex/pMolecule.PairwiseInteraction.pyx:24-105,182-194),
  the adaptive Simpson rule it integrates with
  (pCore-1.9.0/pCore/NumericalIntegration.py:21-70) and the cubic spline built
  and evaluated by pCore-1.9.0/extensions/csource/CubicSpline.c:86-126,306-405;
* the QC atom coordinates and gradients of link (boundary) atoms
  (pMolecule-1.9.0/extensions/csource/QCAtomContainer.c:415-432,583-630), the
  MM charge sites of the RC/RD couplings (QCAtomContainer.c:465-559,674-706);
* the selections and lists NBModelABFSState_SetUp / GenerateLists /
  GenerateLists14 make (NBModelABFSState.c:323-361, NBModelABFS.c:1050-1155,
  PairList.c:342-453,534-672, PairListGenerator.c:28-85,414-493): the MM/MM
  list without pure QC atoms, the QC/MM electrostatic list (QC atom, MM site),
  the QC/MM LJ list (pure QC atom, MM atom) and their 1-4 counterparts;
* NBModelABFS_QCMMEnergyLJ, _QCMMPotentials and _QCMMGradients without images
  (NBModelABFS.c:306-498) over PairwiseInteractionABFS_QCMMPotentials /
  _QCMMGradients (PairwiseInteraction.c:536-620) and the analytic MM/MM pair
  terms for the QC/MM LJ lists.

Pair lists are compared as sets; the reference's grid search includes a cell
pair "fully within range" without a distance test, which can differ from the
distance test only for a pair at the cutoff to the last bit.
"""

import heapq
import math

import numpy as np

import oracle as O

HARTREES_TO_KJ = 2625.5                       # pCore-1.9.0/extensions/cinclude/Units.h:52
BOHR_M = 5.291772083e-11                      # pCore-1.9.0/pCore/Units.py:43-44
ANGSTROMS_TO_BOHRS = 1.0e-10 / BOHR_M
BOHRS_TO_ANGSTROMS = BOHR_M / 1.0e-10


# ----------------------------------------------------------------------------------------------------------------
# Splines.
# ----------------------------------------------------------------------------------------------------------------
def f_gaussian(p):
    def F(r):
        x = p * r
        x2 = x * x
        if x > 1.0e-2:
            f = math.erf(x) / x
        else:
            f = (2.0 - x2 * (2.0 / 3.0 - x2 * (0.2 - x2 * (1.0 / 21.0 - x2 * (1.0 / 108.0 - x2 / 660.0))))) / \
                math.sqrt(math.pi)
        return f * p
    return F


def g_gaussian(p):
    def G(r):
        sp = math.sqrt(math.pi)
        x = p * r
        x2 = x * x
        if x > 1.0e-2:
            g = 2.0 * math.exp(-x2) / (sp * x) - math.erf(x) / x2
        else:
            g = -x * (4.0 / 3.0 - x2 * (0.8 - x2 * (2.0 / 7.0 - x2 * (2.0 / 27.0 - x2 / 66.0)))) / sp
        return g * p * p
    return G


def switch(r, r_on, r_off):
    return (r_off ** 2 - r ** 2) ** 2 * (r_off ** 2 + 2.0 * r ** 2 - 3.0 * r_on ** 2) / (r_off ** 2 - r_on ** 2) ** 3


def adaptive_simpson(F, a, b, tolerance, maximum_iterations):
    """Heap-driven adaptive Simpson rule (NumericalIntegration.py:21-70)."""
    h0 = abs(a - b)
    eps = tolerance / h0
    integral = 0.0
    heap = []
    it = 0
    p1 = 0.5 * (a + b)
    f0, f1, f2 = F(a), F(p1), F(b)
    whole = (f0 + 4.0 * f1 + f2) * h0 / 6.0
    heapq.heappush(heap, (0.0, h0, a, p1, b, f0, f1, f2, whole))
    while it < maximum_iterations and heap:
        _, h, p0, p2, p4, f0, f2, f4, whole = heapq.heappop(heap)
        p1 = 0.5 * (p0 + p2)
        p3 = 0.5 * (p2 + p4)
        f1, f3 = F(p1), F(p3)
        left = (f0 + 4.0 * f1 + f2) * h / 12.0
        right = (f2 + 4.0 * f3 + f4) * h / 12.0
        lrw = left + right - whole
        err = abs(lrw)
        if err < 15.0 * h * eps:
            integral += left + right + lrw / 15.0
        else:
            h *= 0.5
            heapq.heappush(heap, (-err, h, p0, p1, p2, f0, f1, f2, left))
            heapq.heappush(heap, (-err, h, p2, p3, p4, f2, f3, f4, right))
        it += 1
    for item in heap:
        integral += item[8]
    return -integral if a > b else integral


def electrostatic_spline_points(model, width1, width2, damping, inner, outer, density=50, atomic_units=True):
    """(x = r^2, y) of PairwiseInteractionABFS.MakeElectrostaticSplines (pyx:182-194, Integrate pyx:57-105)."""
    factor = 2.0 / math.sqrt(math.pi)
    if model == "Delta/Gaussian":
        p = factor / width2
    elif model == "Gaussian/Gaussian":
        p = factor / math.sqrt(width1 ** 2 + width2 ** 2)
    else:
        raise ValueError(model)
    F, G = f_gaussian(p), g_gaussian(p)
    npts = int(math.ceil(float(density) * outer + 1.0))
    x = np.zeros(npts)
    y = np.zeros(npts)
    inc = outer / float(npts - 1)
    for i in range(1, npts - 1):
        x[i] = inc * float(i)
    x[-1] = outer
    integral = 0.0
    iu = npts - 2
    y[-1] = 0.0
    GS = lambda r: G(r) * switch(r, inner, outer)
    for i in range(npts - 2, -1, -1):
        lo, hi = x[i], x[i + 1]
        if lo < inner:
            iu = i
            if hi > inner:
                integral -= adaptive_simpson(GS, inner, hi, 1.0e-15, 1000)
            break
        integral -= adaptive_simpson(GS, lo, hi, 1.0e-15, 1000)
        y[i] = integral
    off = integral - F(inner)
    if damping > 0:
        fd, gd = F(damping) + off, G(damping)
        alpha, beta = 0.5 * gd / damping, -0.5 * gd * damping + fd
    else:
        alpha = beta = 0.0
    for i in range(iu + 1):
        r = x[i]
        y[i] = alpha * r ** 2 + beta if r < damping else F(r) + off
    x = x ** 2
    if atomic_units:
        y = y * (1.0 / ANGSTROMS_TO_BOHRS)
    else:
        y = y * O.coulomb_factor()
    return x, y


class CubicSpline:
    """Clamped cubic spline through (x, y) with end first derivatives d0, d1 (CubicSpline.c:306-405)."""

    def __init__(self, x, y, d0=0.0, d1=0.0):
        x = np.asarray(x, np.float64)
        y = np.asarray(y, np.float64)
        n = x.size
        ab = np.zeros((3, n))
        rhs = np.zeros(n)
        dl = x[1] - x[0]
        ab[1, 0], ab[0, 1] = dl / 3.0, dl / 6.0
        rhs[0] = (y[1] - y[0]) / dl - d0
        for i in range(1, n - 1):
            dl, du = x[i] - x[i - 1], x[i + 1] - x[i]
            ab[2, i - 1] = dl / 6.0
            ab[1, i] = (dl + du) / 3.0
            ab[0, i + 1] = du / 6.0
            rhs[i] = (y[i + 1] - y[i]) / du + (y[i - 1] - y[i]) / dl
        du = x[n - 1] - x[n - 2]
        ab[2, n - 2] = du / 6.0
        ab[1, n - 1] = du / 3.0
        rhs[n - 1] = d1 - (y[n - 1] - y[n - 2]) / du
        from scipy.linalg import solve_banded
        self.x, self.y, self.h = x, y, solve_banded((1, 1), ab, rhs)

    def evaluate(self, r2):
        """(f, df/dx) at x = r2 (CubicSpline_Evaluate, CubicSpline.c:86-126); r2 array."""
        x, y, h = self.x, self.y, self.h
        r2 = np.asarray(r2, np.float64)
        u = np.searchsorted(x, r2, side="right")
        u = np.clip(u, 1, x.size - 1)
        l = u - 1
        d = x[u] - x[l]
        s = (r2 - x[l]) / d
        t = (x[u] - r2) / d
        hl = h[l] * d / 6.0
        hu = h[u] * d / 6.0
        yl, yu = y[l], y[u]
        f = t * yl + s * yu + d * (t * (t * t - 1.0) * hl + s * (s * s - 1.0) * hu)
        g = (yu - yl) / d + (-(3.0 * t * t - 1.0) * hl + (3.0 * s * s - 1.0) * hu)
        return f, g


# ----------------------------------------------------------------------------------------------------------------
# QC atoms, selections and MM sites.
# ----------------------------------------------------------------------------------------------------------------
class QCAtoms:
    """QC atoms as QCAtomContainer holds them: index (ascending), boundary flags, QC partner and link factor of
    each boundary atom, its MM partners (CSR), the link-ratio flag."""

    def __init__(self, n, index, boundary, partner, linkfactor, mmp_off, mmp, link_ratio=False):
        self.n = n
        self.index = np.asarray(index, np.int64)
        self.boundary = np.asarray(boundary, bool)
        self.partner = np.asarray(partner, np.int64)
        self.linkfactor = np.asarray(linkfactor, np.float64)
        self.mmp_off = np.asarray(mmp_off, np.int64)
        self.mmp = np.asarray(mmp, np.int64)
        self.link_ratio = bool(link_ratio)
        self.nqc = self.index.size
        self.full = np.zeros(n, bool)
        self.full[self.index] = True
        self.pure = np.zeros(n, bool)
        self.pure[self.index[~self.boundary]] = True
        self.mm = ~self.pure                       # mmSelection = complement of the pure QC selection

    def coordinates(self, xyz):
        """QCAtomContainer_GetAtomCoordinates3 for every QC atom (link atoms between partner and boundary)."""
        out = np.array(xyz[self.index], np.float64)
        for q in np.nonzero(self.boundary)[0]:
            b, p = self.index[q], self.partner[q]
            dv = xyz[b] - xyz[p]
            d = self.linkfactor[q]
            if not self.link_ratio:
                d = d / math.sqrt((dv[0] * dv[0] + dv[1] * dv[1]) + dv[2] * dv[2])
            out[q] = xyz[p] + d * dv
        return out

    def set_atom_gradients(self, xyz, q, gq, grad):
        """QCAtomContainer_SetAtomGradients3: the gradient of QC atom q onto the atoms."""
        i = self.index[q]
        if not self.boundary[q]:
            grad[i] += gq
            return
        p = self.partner[q]
        grad[p] += gq
        if self.link_ratio:
            t = gq * self.linkfactor[q]
        else:
            dv = xyz[i] - xyz[p]
            d = math.sqrt((dv[0] * dv[0] + dv[1] * dv[1]) + dv[2] * dv[2])
            u = dv / d
            t = gq * (self.linkfactor[q] / d)
            t = t - np.dot(u, t) * u
        grad[i] += t
        grad[p] -= t

    def mm_sites(self, xyz, charges, coupling):
        """(site coordinates, site charges, site -> (atom, partner or -1)) of the QC/MM electrostatics:
        the atoms with the pure QC charges zeroed (MM coupling, MMAtomContainer_GetCharges of active atoms), or
        the pure MM atoms followed by one site per boundary atom and MM partner (RC/RD)."""
        if coupling == 0:
            q = np.where(self.pure, 0.0, charges)
            return np.array(xyz, np.float64), q, np.stack([np.arange(self.n), -np.ones(self.n, np.int64)], 1)
        pure_mm = np.nonzero(~self.full)[0]
        pos = -np.ones(self.n, np.int64)
        pos[pure_mm] = np.arange(pure_mm.size)
        sites = [(a, -1) for a in pure_mm]
        q = list(charges[pure_mm])
        for k in np.nonzero(self.boundary)[0]:
            b = self.index[k]
            parts = self.mmp[self.mmp_off[k]:self.mmp_off[k + 1]]
            q0 = charges[b] / float(parts.size)
            for p in parts:
                sites.append((b, p))
                if self.full[p] and self.boundary[np.searchsorted(self.index, p)] or coupling == 1:
                    q.append(q0)
                else:
                    q.append(2.0 * q0)
                    q[pos[p]] -= q0
        sites = np.array(sites, np.int64)
        sx = np.array(xyz[sites[:, 0]], np.float64)
        half = sites[:, 1] >= 0
        sx[half] = 0.5 * (xyz[sites[half, 0]] + xyz[sites[half, 1]])
        return sx, np.array(q), sites


# ----------------------------------------------------------------------------------------------------------------
# Lists.
# ----------------------------------------------------------------------------------------------------------------
def _csr_rows(n, pairs):
    ptr, idx = O.exclusion_csr(n, pairs)
    return [set(idx[ptr[i]:ptr[i + 1]].tolist()) for i in range(n)]


def _within(xi, xj, cutoff):
    d = xi[None, :] - xj
    r2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
    return r2 <= cutoff * cutoff


def _cells(x, lower, h, bins):
    """RegularGrid_FindCellIDOfPoint as indices: floor((x - lower)/h) per axis and whether all are on the grid."""
    k = np.floor((x - lower) / h).astype(np.int64)
    return k, np.all((k >= 0) & (k < bins), axis=1)


def _point_cell_r2(p, k, lower, h):
    """RegularGrid_DistancesSquaredBetweenCellAndPoint (non-periodic), point p against cells k (rows)."""
    lo = k * h + lower
    hi = lo + h
    dl, du = p - lo, hi - p
    mx = np.where(p > hi, dl, np.where(p < lo, du, np.maximum(dl, du)))
    mn = np.where(p > hi, p - hi, np.where(p < lo, lo - p, 0.0))
    return (mn[:, 0] * mn[:, 0] + mn[:, 1] * mn[:, 1]) + mn[:, 2] * mn[:, 2], \
           (mx[:, 0] * mx[:, 0] + mx[:, 1] * mx[:, 1]) + mx[:, 2] * mx[:, 2]


class CrossGrid:
    """The grid search of MakeCrossPairListFromCoordinates3 (pCore-1.9.0/extensions/csource/PairListGenerator.c:
    611-940): row points on grid 1 (made from rows_grid_xyz), column points on a grid conforming to it made from the
    box of the columns in col_box (Coordinates3_MakeConformingGrid, Coordinates3.c:998-1072), or on grid 1 itself
    (same=True, PairListGenerator_CrossPairListFromSingleCoordinates3, :450-493).  select(r, x_r, cols) is the
    mask of the candidate columns the search lists for row r: the column on its grid, in a cell I + o - lowN over
    the stencil offsets o around the row's search cell I, and then whole cells (cell/cell: stencil flag 2; point/
    cell: largest point/cell distance) or the FP64 distance test; point/cell cells also need the smallest
    point/cell distance below the cutoff.  Cell/cell rows must be on grid 1; point/cell rows off it search around
    the cell of the latest earlier row on it ((-u, -u, -u) before any)."""

    def __init__(self, point_cell, h, cutoff, grid1_xyz, rows_xyz, cols_xyz, col_box, same=False):
        self.pc, self.h, self.c2 = point_cell, h, cutoff * cutoff
        bins1, lower1, _ = O.make_grid(grid1_xyz, h)
        self.u = int(np.ceil(cutoff / h))
        off, flg = O.search_range(h, cutoff)
        w = 2 * self.u + 1
        self.sten = np.zeros((w, w, w), np.int32)
        for (a, b, c), f in zip(off, flg):
            self.sten[a + self.u, b + self.u, c + self.u] = f
        if same:
            self.low = np.zeros(3, np.int64)
            bins2, lower2 = bins1, lower1
        else:
            lo = col_box.min(axis=0)
            up = (col_box.max(axis=0) - lo) + lo
            self.low = np.floor((lo - lower1) / h).astype(np.int64)
            bins2 = np.floor((up - lower1) / h).astype(np.int64) - self.low + 1
            lower2 = lower1 + self.low * h
        self.lower2 = lower2
        self.kc, self.onc = _cells(cols_xyz, lower2, h, bins2)
        kr, onr = _cells(rows_xyz, lower1, h, bins1)
        self.rows_on = onr
        self.scell = kr.copy()
        last = np.array([-self.u] * 3, np.int64)
        for r in range(rows_xyz.shape[0]):
            if onr[r]:
                last = kr[r]
            else:
                self.scell[r] = last

    def select(self, r, xr, cols_xyz, cand):
        if not self.pc and not self.rows_on[r]:
            return np.zeros(cand.size, bool)
        o = self.kc[cand] + self.low - self.scell[r] + self.u
        w = self.sten.shape[0]
        inb = np.all((o >= 0) & (o < w), axis=1)
        oc = np.clip(o, 0, w - 1)
        flag = np.where(inb, self.sten[oc[:, 0], oc[:, 1], oc[:, 2]], 0)
        d = xr[None, :] - cols_xyz[cand]
        near = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2] <= self.c2
        ok = self.onc[cand] & (flag != 0)
        if self.pc:
            mn, mx = _point_cell_r2(xr[None, :], self.kc[cand], self.lower2, self.h)
            return ok & (mn < self.c2) & ((mx <= self.c2) | near)
        return ok & ((flag == 2) | near)


class QCMMLists:
    """The QC/MM lists of one update, as sorted arrays per row.  grid: None (direct method) or (point_cell, h):
    the grid search of both cross lists (CrossGrid; the electrostatic list on the grid of the QC coordinates when
    there is more than one QC atom, NBModelABFSState_GridInitialize, pMolecule-1.9.0/extensions/csource/
    NBModelABFSState.c:207-226, the LJ list on the grid of all atoms)."""

    def __init__(self, qc, xyz, exclusions, interactions14, coupling, free=None, cutoff=13.5, charges=None, grid=None):
        n = qc.n
        self.free = np.ones(n, bool) if free is None else np.asarray(free, bool)
        excl = _csr_rows(n, exclusions)
        p14 = _csr_rows(n, interactions14)
        qx = qc.coordinates(xyz)
        sx, self.site_q, self.sites = qc.mm_sites(xyz, charges if charges is not None else np.zeros(n), coupling)
        self.el, self.el14, self.lj, self.lj14 = [], [], [], []
        gel = None
        if grid is not None and qc.nqc > 1:
            # Columns: all atoms with the MM selection's box (MM coupling) or the MM sites (RC, RD).
            gel = CrossGrid(grid[0], grid[1], cutoff, qx, qx, sx, sx[qc.mm] if coupling == 0 else sx)
        for q in range(qc.nqc):
            i = qc.index[q]
            if coupling == 0:
                ex = {j for j in excl[i] if qc.mm[j]} | ({i} if qc.mm[i] else set())
                ok = qc.mm.copy()
                ok[list(ex)] = False
                cand = np.nonzero(ok)[0]
            else:
                cand = np.arange(sx.shape[0])
            sel = _within(qx[q], sx[cand], cutoff) if gel is None else gel.select(q, qx[q], sx, cand)
            self.el.append(cand[sel])
            if coupling == 0:
                self.el14.append(np.array(sorted(j for j in p14[i] if qc.mm[j] and j != i), np.int64))
            else:
                self.el14.append(np.zeros(0, np.int64))
        self.lj_rows = np.nonzero(qc.pure)[0]
        glj = None
        if grid is not None:
            glj = CrossGrid(grid[0], grid[1], cutoff, xyz, xyz[self.lj_rows], xyz, None, same=True)
        for r, i in enumerate(self.lj_rows):
            ok = qc.mm & (self.free | self.free[i])
            ok[list(excl[i])] = False
            cand = np.nonzero(ok)[0]
            sel = _within(xyz[i], xyz[cand], cutoff) if glj is None else glj.select(r, xyz[i], xyz, cand)
            self.lj.append(cand[sel])
            self.lj14.append(np.array(sorted(j for j in p14[i] if qc.mm[j] and (self.free[i] or self.free[j])),
                                      np.int64))
        self.site_xyz = sx
        self.qc_xyz = qx

    @staticmethod
    def keys(rows, labels, width):
        """Canonical keys row_label * width + column, sorted."""
        out = [np.asarray(r, np.int64) + int(l) * width for r, l in zip(rows, labels)]
        return np.sort(np.concatenate(out)) if out else np.zeros(0, np.int64)

    def npairs(self, which):
        return int(sum(len(r) for r in getattr(self, which)))


def records(rows_i, rows_j, n):
    """PairRecords of a cross list (rows with at least one pair)."""
    ri, ptr, js = [], [0], []
    for i, r in zip(rows_i, rows_j):
        if len(r):
            ri.append(int(i))
            js.extend(int(j) for j in r)
            ptr.append(len(js))
    return O.PairRecords(np.array(ri, np.int32), np.array(ptr, np.int64), np.array(js, np.int32), n)


# ----------------------------------------------------------------------------------------------------------------
# Potentials, gradients, LJ.
# ----------------------------------------------------------------------------------------------------------------
def potentials(spline, qc_xyz, site_xyz, site_q, rows, escale, outer=12.0):
    """PairwiseInteractionABFS_QCMMPotentials: pot[q] = sum eScale q_m f(r^2) over r^2 <= outer^2."""
    out = np.zeros(len(rows))
    r2off = outer * outer
    for q, r in enumerate(rows):
        if len(r) == 0:
            continue
        d = qc_xyz[q][None, :] - site_xyz[r]
        r2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
        m = r2 <= r2off
        f, _ = spline.evaluate(r2[m])
        out[q] = np.sum((escale * site_q[r][m]) * f)
    return out


def gradients(spline, qc, xyz, qc_xyz, site_xyz, site_q, rows, escale, qcq, site_grad, grad, outer=12.0):
    """PairwiseInteractionABFS_QCMMGradients: MM site gradients decremented, QC atom gradients through the
    link-atom rule."""
    r2off = outer * outer
    scale = HARTREES_TO_KJ * escale
    for q, r in enumerate(rows):
        if len(r) == 0:
            continue
        r = np.asarray(r)
        d = qc_xyz[q][None, :] - site_xyz[r]
        r2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
        m = r2 <= r2off
        _, g = spline.evaluate(r2[m])
        qqm = (scale * qcq[q]) * site_q[r][m]
        v = d[m] * (g * (2.0 * qqm))[:, None]
        np.add.at(site_grad, r[m], -v)
        qc.set_atom_gradients(xyz, q, v.sum(0), grad)


class QCMMOracle:
    """One QC/MM system: the MM/MM part (no pure QC atoms), QC/MM LJ, potentials and gradients."""

    def __init__(self, system, qc, coupling, spline, fixed=None, damping=0.5, inner=8.0, outer=12.0,
                 list_cutoff=13.5, dielectric=1.0, scale14=1.0, tableA14=None, tableB14=None, use_grid_by_cell=True):
        self.s, self.qc, self.coupling, self.spline = system, qc, coupling, spline
        self.use_grid_by_cell = use_grid_by_cell
        self.free = None if fixed is None else ~np.asarray(fixed, bool)
        self.f = O.factors(damping, inner, outer)
        self.outer, self.list_cutoff, self.dielectric, self.scale14 = outer, list_cutoff, dielectric, scale14
        self.tA14 = system.tableA if tableA14 is None else tableA14
        self.tB14 = system.tableB if tableB14 is None else tableB14
        self.xptr, self.xidx = O.exclusion_csr(system.natoms, system.exclusions)
        self.ref = None
        self.lists = None
        self.use_grid = None           # decided once, on the first build (NBModelABFSState.c:212)

    def mmmm_records(self, xyz):
        """Self list of the MM selection with the free-atom OR test (the full list filtered)."""
        s = self.s
        use_grid = O.determine_method(xyz, self.list_cutoff)
        method = (1 if self.use_grid_by_cell else 2) if use_grid else 0
        # The MM selection is the AND selection of GenerateLists (NBModelABFS.c:1080); the point/cell search
        # applies it itself (its points outside it decode no cell), the other methods through the filter.
        rec = O.self_pair_list(xyz, self.xptr, self.xidx, self.list_cutoff, 0.5 * self.list_cutoff, method,
                               qand=self.qc.mm if method == 2 else None)
        return self._filter(rec)

    def _filter(self, rec):
        mm = self.qc.mm
        free = np.ones(self.s.natoms, bool) if self.free is None else self.free
        rows_i, rows_j = [], []
        for k, i in enumerate(rec.rec_i):
            js = rec.js[rec.rec_ptr[k]:rec.rec_ptr[k + 1]]
            if mm[i]:
                keep = js[mm[js] & (free[js] | free[i])]
                rows_i.append(i)
                rows_j.append(keep)
        return records(rows_i, rows_j, self.s.natoms)

    def call(self, xyz, qcq):
        """One System.Energy call: (energies dict, potentials, gradient of the energy part, gradient of
        NBModelABFS_QCMMGradients, rebuilt)."""
        s, qc = self.s, self.qc
        xyz = np.ascontiguousarray(xyz, np.float64)
        rebuilt = self.ref is None
        if not rebuilt:
            fixed = None if self.free is None else ~self.free
            rebuilt, _ = O.check_for_update(xyz, self.ref, self.list_cutoff, self.outer, fixed)
        if rebuilt:
            self.rec = self.mmmm_records(xyz)
            if self.use_grid is None:
                self.use_grid = O.determine_method(xyz, self.list_cutoff)
            grid = (not self.use_grid_by_cell, 0.5 * self.list_cutoff) if self.use_grid else None
            self.lists = QCMMLists(qc, xyz, s.exclusions, s.interactions14, self.coupling, self.free,
                                   self.list_cutoff, s.charges, grid)
            p14 = O.records_from_pairs(s.natoms, s.interactions14)
            self.rec14 = self._filter(p14)
            self.ref = xyz.copy()
        L = self.lists
        g = np.zeros((s.natoms, 3))
        e = {}
        escale = 1.0 / self.dielectric
        if self.rec.js.size:
            e["MM/MM Elect."], e["MM/MM LJ"] = O.mmmm_energy(self.f, xyz, s.charges, s.ljtypes, s.ntypes, s.tableA,
                                                             s.tableB, self.rec, escale, 1.0, 0, g)
        if self.rec14.js.size:
            e["MM/MM 1-4 Elect."], e["MM/MM 1-4 LJ"] = O.mmmm_energy(self.f, xyz, s.charges, s.ljtypes, s.ntypes,
                                                                     self.tA14, self.tB14, self.rec14,
                                                                     escale * self.scale14, 1.0, 1, g)
        rlj = records(L.lj_rows, L.lj, s.natoms)
        if rlj.js.size:
            e["QC/MM LJ"] = O.mmmm_energy(self.f, xyz, s.charges, s.ljtypes, s.ntypes, s.tableA, s.tableB, rlj,
                                          0.0, 1.0, 1, g)[1]
        rlj14 = records(L.lj_rows, L.lj14, s.natoms)
        if rlj14.js.size:
            e["QC/MM 1-4 LJ"] = O.mmmm_energy(self.f, xyz, s.charges, s.ljtypes, s.ntypes, self.tA14, self.tB14,
                                              rlj14, 0.0, 1.0, 1, g)[1]
        # Potentials (atomic units): the main list, then the 1-4 list with the 1-4 scale.
        qx = qc.coordinates(xyz)
        sx, sq, sites = qc.mm_sites(xyz, s.charges, self.coupling)
        pot = potentials(self.spline, qx, sx, sq, L.el, escale, self.outer)
        if self.coupling == 0:
            pot = pot + potentials(self.spline, qx, sx, sq, L.el14, escale * self.scale14, self.outer)
        # Gradients with the QC charges.
        gq = np.zeros((s.natoms, 3))
        sg = np.zeros((sx.shape[0], 3))
        gradients(self.spline, qc, xyz, qx, sx, sq, L.el, escale, qcq, sg, gq, self.outer)
        if self.coupling == 0:
            gradients(self.spline, qc, xyz, qx, sx, sq, L.el14, escale * self.scale14, qcq, sg, gq, self.outer)
            gq += sg
        else:
            for k, (a, p) in enumerate(sites):
                if p < 0:
                    gq[a] += sg[k]
                else:
                    gq[a] += 0.5 * sg[k]
                    gq[p] += 0.5 * sg[k]
        return e, pot, g, gq, rebuilt


# ----------------------------------------------------------------------------------------------------------------
# The compiled reference with QC atoms (oracle/_ref, ref_driver.c: ref_create_qc and friends).
# ----------------------------------------------------------------------------------------------------------------
def _ref():
    import ctypes as C
    L = O.ref_lib()
    if not hasattr(L, "_qc_ready"):
        L.ref_create_qc.restype = C.c_void_p
        L.ref_create_qc.argtypes = L.ref_create.argtypes + [C.c_int, O._ip, O._ip, O._ip, O._dp, C.c_int, O._ip, O._ip,
                                                              C.c_int, C.c_int, C.c_void_p, C.c_int, O._dp, O._dp]
        L.ref_qcmm_energy.argtypes = [C.c_void_p, O._dp, O._dp, C.c_void_p, C.c_void_p, O._dp, O._ip, O._dp]
        L.ref_qcmm_pairs.restype = C.c_long
        L.ref_qcmm_pairs.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_long]
        L.ref_qcmm_sites.restype = C.c_long
        L.ref_qcmm_sites.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_long]
        L.ref_qcmm_spline.argtypes = [C.c_void_p, C.c_int, O._dp]
        L._qc_ready = True
    return L


class ReferenceQCMM:
    """The reference C library with QC atoms, driven call by call."""

    LABELS = ("MM/MM Elect.", "MM/MM LJ", "MM/MM 1-4 Elect.", "MM/MM 1-4 LJ", "QC/MM LJ", "QC/MM 1-4 LJ")

    def __init__(self, system, qc, coupling, spline_x, spline_y, fixed=None, damping=0.5, inner=8.0, outer=12.0,
                 list_cutoff=13.5, dielectric=1.0, scale14=1.0, cell_factor=0.5, min_extent=2, min_points=500,
                 nthreads=1, use_grid_by_cell=True):
        L = _ref()
        s = system
        self.n, self.nqc = s.natoms, qc.nqc
        ex = np.ascontiguousarray(s.exclusions, np.int32).reshape(-1)
        p14 = np.ascontiguousarray(s.interactions14, np.int32).reshape(-1)
        fx = np.nonzero(np.asarray(fixed, bool))[0].astype(np.int32) if fixed is not None else np.zeros(0, np.int32)
        self.sx = np.ascontiguousarray(spline_x, np.float64)
        self.sy = np.ascontiguousarray(spline_y, np.float64)
        arrs = [np.ascontiguousarray(qc.index, np.int32), np.ascontiguousarray(qc.boundary, np.int32),
                np.ascontiguousarray(qc.partner, np.int32), np.ascontiguousarray(qc.linkfactor, np.float64),
                np.ascontiguousarray(qc.mmp_off, np.int32),
                np.ascontiguousarray(qc.mmp if qc.mmp.size else np.zeros(1), np.int32)]
        self._keep = [ex, p14, fx, self.sx, self.sy] + arrs
        opts = np.array([damping, inner, outer, list_cutoff, dielectric, scale14])
        gen = np.array([cell_factor, min_extent, min_points, 1.0 if use_grid_by_cell else 0.0], dtype=np.float64)
        self.h = L.ref_create_qc(self.n, np.ascontiguousarray(s.charges, np.float64),
                                 np.ascontiguousarray(s.ljtypes, np.int32), s.ntypes,
                                 np.ascontiguousarray(s.tableA, np.float64), np.ascontiguousarray(s.tableB, np.float64),
                                 None, None, ex.size // 2, ex.ctypes.data if ex.size else None, p14.size // 2,
                                 p14.ctypes.data if p14.size else None, opts, gen, nthreads,
                                 qc.nqc, arrs[0], arrs[1], arrs[2], arrs[3], 1 if qc.link_ratio else 0, arrs[4],
                                 arrs[5], int(coupling), fx.size, fx.ctypes.data if fx.size else None,
                                 self.sx.size, self.sx, self.sy)

    def call(self, xyz, qcq):
        """(energies dict, potentials, gradient of the energy part, gradient of the QC/MM electrostatics,
        rebuilt)."""
        L = _ref()
        xyz = np.ascontiguousarray(xyz, np.float64)
        e = np.zeros(6)
        present = np.zeros(4, np.int32)
        pot = np.zeros(self.nqc)
        ge = np.zeros((self.n, 3))
        gq = np.zeros((self.n, 3))
        up = L.ref_qcmm_energy(self.h, xyz, np.ascontiguousarray(qcq, np.float64), ge.ctypes.data, gq.ctypes.data,
                               e, present, pot)
        if up < 0:
            raise RuntimeError("reference library error")
        terms = {}
        for k, (a, b) in enumerate(((0, 1), (2, 3))):
            if present[k]:
                terms[self.LABELS[a]] = e[a]
                terms[self.LABELS[b]] = e[b]
        if present[2]:
            terms["QC/MM LJ"] = e[4]
        if present[3]:
            terms["QC/MM 1-4 LJ"] = e[5]
        return terms, pot, ge, gq, bool(up)

    def pairs(self, which):
        """which: 'el', 'lj', 'el14', 'lj14' (QC/MM lists) or 'mmmm', 'mmmm14' -> (i, j) in reference order."""
        L = _ref()
        k = {"el": 0, "lj": 1, "el14": 2, "lj14": 3}.get(which)
        if k is None:
            f = lambda oi, oj, cap: L.ref_pairs(self.h, 0 if which == "mmmm" else 1, oi, oj, cap)
        else:
            f = lambda oi, oj, cap: L.ref_qcmm_pairs(self.h, k, oi, oj, cap)
        m = f(None, None, 0)
        oi = np.zeros(max(m, 1), np.int32)
        oj = np.zeros(max(m, 1), np.int32)
        f(oi.ctypes.data, oj.ctypes.data, m)
        return oi[:m].astype(np.int64), oj[:m].astype(np.int64)

    def sites(self):
        L = _ref()
        m = L.ref_qcmm_sites(self.h, None, None, None, 0)
        qx = np.zeros((self.nqc, 3))
        sx = np.zeros((max(m, 1), 3))
        sq = np.zeros(max(m, 1))
        L.ref_qcmm_sites(self.h, qx.ctypes.data, sx.ctypes.data, sq.ctypes.data, m)
        return qx, sx[:m], sq[:m]

    def spline_h(self):
        h = np.zeros(self.sx.size)
        _ref().ref_qcmm_spline(self.h, h.size, h)
        return h

    def close(self):
        if self.h:
            _ref().ref_destroy(self.h)
            self.h = None
