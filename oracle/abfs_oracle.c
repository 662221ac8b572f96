/*
 * abfs_oracle.c -- CPU FP64 restatement of pDynamo's NBModelABFS MM/MM
 * energy/gradient and its PairListGenerator cell-grid self pair list.
 *
 * TEST INFRASTRUCTURE ONLY (the checker for the HIP path; see abfs_oracle.h).
 *
 * Arithmetic follows the reference expression order so that, compiled with
 * gcc -O3 on x86-64 without FMA contraction, each pair term matches the
 * reference bit for bit.  Function headers cite the reference lines restated.
 */
#include "abfs_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* pCx/cinclude/Units.h:18,27,30,48-49: 1e7 N_A e^2 / (4 pi eps0). */
double orc_coulomb_factor(void)
{
    const double avogadro = 6.0221415e+23;
    const double echarge = 1.60217653e-19;
    const double eps0 = 8.854187817e-12;
    return (1.0e+7 * avogadro * echarge * echarge) / (4.0e+00 * M_PI * eps0);
}

/* pMx/csource/PairwiseInteraction.c:89-141.  The damped-region constants are
 * the value and slope matching of a quadratic in r at the damping radius. */
void orc_make_factors(double damp, double inner, double outer, double *f)
{
    double r2d = damp * damp, r2off = outer * outer, r2on = inner * inner;
    f[ORC_R2DAMP] = r2d;
    f[ORC_R2OFF] = r2off;
    f[ORC_R2ON] = r2on;

    double gamma = pow(r2off - r2on, 3);
    double a = r2off * r2off * (r2off - 3.0 * r2on) / gamma;
    double b = 6.0 * r2off * r2on / gamma;
    double c = -(r2off + r2on) / gamma;
    double d = 0.4 / gamma;
    double qs1 = 8.0 * (r2off * r2on * (outer - inner) -
                        0.2 * (outer * r2off * r2off - inner * r2on * r2on)) / gamma;
    double qs2 = -(a / outer) + b * outer + c * outer * r2off + d * outer * r2off * r2off;
    f[ORC_A] = a; f[ORC_B] = b; f[ORC_C] = c; f[ORC_D] = d;
    f[ORC_QSHIFT1] = qs1; f[ORC_QSHIFT2] = qs2;
    {
        double v = 1.0 / damp + qs1, g = -1.0 / r2d;
        f[ORC_QF0] = v - 0.5 * damp * g;
        f[ORC_QALPHA] = -0.5 * g / damp;
    }
    f[ORC_AF6] = 1.0 / (r2off * r2off * r2off);
    f[ORC_AK12] = pow(r2off, 3) / (pow(r2off, 3) - pow(r2on, 3));
    f[ORC_ASHIFT12] = 1.0 / pow(inner * outer, 6);
    {
        double s12 = 1.0 / pow(damp, 12);
        double v = s12 - f[ORC_ASHIFT12], g = -12.0 * s12 / damp;
        f[ORC_AF0] = v - 0.5 * damp * g;
        f[ORC_AALPHA] = -0.5 * g / damp;
    }
    f[ORC_BF3] = 1.0 / (outer * r2off);
    f[ORC_BK6] = (outer * r2off) / (outer * r2off - inner * r2on);
    f[ORC_BSHIFT6] = 1.0 / pow(inner * outer, 3);
    {
        double s6 = 1.0 / pow(damp, 6);
        double v = -s6 + f[ORC_BSHIFT6], g = 6.0 * s6 / damp;
        f[ORC_BF0] = v - 0.5 * damp * g;
        f[ORC_BALPHA] = -0.5 * g / damp;
    }
}

/* pMx/cinclude/PairwiseInteraction.h:72-119 (CheckDistances, ElectrostaticTerm,
 * LennardJonesTerm).  pow(x, 2) in the reference is x*x (exact). */
int orc_pair_terms(const double *f, double r2, double qij, double aij, double bij,
                   double *eel, double *elj, double *df)
{
    double s, s2, dF = 0.0;
    if (r2 > f[ORC_R2OFF]) return 0;
    if (r2 < f[ORC_R2DAMP]) { s2 = 0.0; s = 0.0; }
    else { s2 = 1.0 / r2; s = sqrt(s2); }
    /* Coulomb. */
    if (r2 > f[ORC_R2ON]) {
        *eel = qij * (s * (f[ORC_A] - r2 * (f[ORC_B] + r2 * (f[ORC_C] + f[ORC_D] * r2))) + f[ORC_QSHIFT2]);
        dF += -qij * 0.5 * s * (f[ORC_A] + r2 * (f[ORC_B] + r2 * (3.0 * f[ORC_C] + 5.0 * f[ORC_D] * r2))) / r2;
    } else if (r2 > f[ORC_R2DAMP]) {
        *eel = qij * (s + f[ORC_QSHIFT1]);
        dF += -qij * 0.5 * s / r2;
    } else {
        *eel = qij * (f[ORC_QF0] - f[ORC_QALPHA] * r2);
        dF += -qij * f[ORC_QALPHA];
    }
    /* Lennard-Jones. */
    {
        double s6 = s2 * s2 * s2;
        if (r2 > f[ORC_R2ON]) {
            double l1 = s6 - f[ORC_AF6], l2 = (s / r2) - f[ORC_BF3];
            *elj = aij * f[ORC_AK12] * (l1 * l1) - bij * f[ORC_BK6] * (l2 * l2);
            dF += -3.0 * s6 * (2.0 * aij * f[ORC_AK12] * l1 / r2 - bij * f[ORC_BK6] * l2 / s);
        } else if (r2 > f[ORC_R2DAMP]) {
            *elj = aij * (s6 * s6 - f[ORC_ASHIFT12]) - bij * (s6 - f[ORC_BSHIFT6]);
            dF += -3.0 * s6 * (2.0 * aij * s6 - bij) / r2;
        } else {
            *elj = aij * (f[ORC_AF0] - f[ORC_AALPHA] * r2) - bij * (f[ORC_BF0] - f[ORC_BALPHA] * r2);
            dF += -aij * f[ORC_AALPHA] + bij * f[ORC_BALPHA];
        }
    }
    *df = dF;
    return 1;
}

/* pCx/csource/Coordinates3.c:498-605 (no selection, no radii). */
static void enclosing_box(int n, const double *xyz, double *lo, double *ext)
{
    double mn[3], mx[3];
    if (n <= 0) { for (int d = 0; d < 3; d++) { lo[d] = 0.0; ext[d] = 0.0; } return; }
    for (int d = 0; d < 3; d++) { mn[d] = xyz[d]; mx[d] = xyz[d]; }
    for (int i = 1; i < n; i++)
        for (int d = 0; d < 3; d++) {
            double v = xyz[3 * i + d];
            if (v > mx[d]) mx[d] = v;
            if (v < mn[d]) mn[d] = v;
        }
    for (int d = 0; d < 3; d++) { lo[d] = mn[d]; ext[d] = mx[d] - mn[d]; }
}

/* pCx/csource/PairListGenerator.c:498-525. cellSize = factor * cutoff as set
 * by pCx/pyrex/pCore.PairListGenerator.pyx:104. */
int orc_determine_method(int n, const double *xyz, double cutoff, double cutoffCellSizeFactor,
                         int minimumCellExtent, int minimumPoints)
{
    double lo[3], ext[3];
    enclosing_box(n, xyz, lo, ext);
    double mext = fabs(ext[0]);
    if (fabs(ext[1]) > mext) mext = fabs(ext[1]);
    if (fabs(ext[2]) > mext) mext = fabs(ext[2]);
    double cellSize = cutoffCellSizeFactor * cutoff;
    int mcells = (int)ceil(mext / cellSize);
    return (mext > cutoffCellSizeFactor * cutoff) && (mcells > minimumCellExtent) && (n > minimumPoints);
}

/* pCx/csource/Coordinates3.c:1111-1158 (grid) and pCx/cinclude/RegularGrid.h:
 * 73-85 with pCx/csource/RegularGrid.c:240-251 (cell of a point). */
long orc_make_grid(int n, const double *xyz, double cellSize, int *bins, double *lower, int *cell)
{
    double lo[3], ext[3];
    long stride[3];
    enclosing_box(n, xyz, lo, ext);
    for (int d = 0; d < 3; d++) {
        double extent = ext[d], l0 = lo[d];
        double upper = l0 + extent;
        double mid = 0.5 * (l0 + upper);
        double cells = ceil(extent / cellSize);
        bins[d] = (int)cells;
        lower[d] = mid - 0.5 * (cells * cellSize);
    }
    stride[2] = 1;
    stride[1] = bins[2];
    stride[0] = (long)bins[1] * bins[2];
    for (int i = 0; i < n; i++) {
        long c = 0;
        int off = 0;
        for (int d = 0; d < 3; d++) {
            double fr = (xyz[3 * i + d] - lower[d]) / cellSize;
            int k = (int)floor(fr);
            if (k < 0 || k >= bins[d]) off = 1;
            c += stride[d] * k;
        }
        cell[i] = off ? -1 : (int)c;
    }
    return (long)bins[0] * bins[1] * bins[2];
}

/* pCx/csource/RegularGrid.c:526-575 with DistancesSquaredBetweenCells
 * (:188-234, non-periodic branch) and BinsWithinCutoff (RegularGrid.h:65-70). */
int orc_search_range(double cellSize, double cutoff, int cap, int *offsets, int *flags)
{
    int u = (int)ceil(cutoff / cellSize);
    double c2 = cutoff * cutoff;
    int n = 0;
    for (int a = -u; a <= u; a++)
        for (int b = -u; b <= u; b++)
            for (int c = -u; c <= u; c++) {
                int idx[3] = {a, b, c};
                double maxR2 = 0.0, minR2 = 0.0;
                for (int d = 0; d < 3; d++) {
                    double delta = fabs((double)(0 - idx[d]));
                    double mx = (delta + 1.0) * cellSize;
                    double mn = (delta - 1.0) * cellSize;
                    if (mn < 0.0) mn = 0.0;
                    maxR2 += mx * mx;
                    minR2 += mn * mn;
                }
                if (minR2 < c2) {
                    if (n < cap) {
                        offsets[3 * n] = a; offsets[3 * n + 1] = b; offsets[3 * n + 2] = c;
                        flags[n] = (maxR2 <= c2) ? 2 : 1;
                    }
                    n++;
                }
            }
    return n;
}

typedef struct {
    int *rec_i; long *rec_ptr; int *js;
    long cap_rec, cap_pairs, nrec, npairs;
    int overflow;
} PairSink;

/* pCx/csource/PairListGenerator.c:1223-1243 (SaveInteractions, unsorted). */
static void save_record(PairSink *s, int i, int n, const int *ind)
{
    if (n <= 0) return;
    if (s->nrec + 1 > s->cap_rec || s->npairs + n > s->cap_pairs) { s->overflow = 1; s->nrec++; s->npairs += n; return; }
    s->rec_i[s->nrec] = i;
    s->rec_ptr[s->nrec] = s->npairs;
    memcpy(s->js + s->npairs, ind, sizeof(int) * (size_t)n);
    s->nrec++;
    s->npairs += n;
    s->rec_ptr[s->nrec] = s->npairs;
}

/* Point/cell self list (pCore-1.9.0/extensions/csource/PairListGenerator.c:1155-1218): atom i, ascending from 1,
 * is the point.  The cells searched are the stencil offsets around i's cell that lie on the grid and pass the
 * point/cell range test of RegularGrid_FindCellsWithinRangeOfPoint (RegularGrid.c:305-336): the smallest squared
 * distance from the point to the cell below cutoff^2; a cell whose largest squared distance is at most cutoff^2 is
 * taken whole, otherwise each atom by the distance test.  Atoms j of a cell come in ascending order and the scan of
 * a cell stops at j >= i.  Distances are RegularGrid_DistancesSquaredBetweenCellAndPoint's (RegularGrid.c:139-180,
 * non-periodic branch): per axis the cell [n h + lower, n h + lower + h] against the coordinate, summed in axis
 * order.  A point off the grid has no cell (RegularGrid_CellIDToIndices, RegularGrid.c:69-85, returns without
 * writing): its search is centred on the cell indices last decoded, those of the latest earlier point on the grid,
 * or, before any, the lower corner (-u, -u, -u) where the search-range construction (RegularGrid.c:526-575) leaves
 * the same work array. */
static void point_cell_records(PairSink *s, int n, const double *xyz, const int *xptr, const int *xidx, double c2,
                               double cellSize, double cutoff, const int *bins, const double *lower, const int *cell,
                               const long *first, const long *total, const int *points, int noff, const int *off,
                               const unsigned char *qand, unsigned char *incl, int *ind)
{
    const long stride0 = (long)bins[1] * bins[2], stride1 = bins[2];
    const int u = (int)ceil(cutoff / cellSize);
    int pc[3] = {-u, -u, -u};
    for (int i = 1; i < n; i++) {
        if (qand && !qand[i]) continue;      /* outside the AND selection: no search, no cell decoded */
        if (cell[i] >= 0) {
            long c = cell[i];
            pc[0] = (int)(c / stride0);
            pc[1] = (int)((c % stride0) / stride1);
            pc[2] = (int)(c % stride1);
        }
        int xf = xptr ? xptr[i] : 0, xl = xptr ? xptr[i + 1] : 0, m = 0;
        for (int x = xf; x < xl; x++) { int j = xidx[x]; if (j >= i) break; incl[j] = 0; }
        for (int k = 0; k < noff; k++) {
            int idx[3] = {pc[0] + off[3 * k], pc[1] + off[3 * k + 1], pc[2] + off[3 * k + 2]};
            if (idx[0] < 0 || idx[0] >= bins[0] || idx[1] < 0 || idx[1] >= bins[1] || idx[2] < 0 || idx[2] >= bins[2])
                continue;
            double maxR2 = 0.0, minR2 = 0.0;
            for (int d = 0; d < 3; d++) {
                double lo = (double)idx[d] * cellSize + lower[d], hi = lo + cellSize, x = xyz[3 * i + d];
                double mx, mn;
                if (x > hi) { mx = x - lo; mn = x - hi; }
                else if (x < lo) { mx = hi - x; mn = lo - x; }
                else { mx = (x - lo) > (hi - x) ? (x - lo) : (hi - x); mn = 0.0; }
                maxR2 += mx * mx;
                minR2 += mn * mn;
            }
            if (!(minR2 < c2)) continue;
            const int all = maxR2 <= c2;
            const long cid = idx[0] * stride0 + idx[1] * stride1 + idx[2];
            for (long p = first[cid], pe = first[cid] + total[cid]; p < pe; p++) {
                int j = points[p];
                if (j >= i) break;
                if (!incl[j] || (qand && !qand[j])) continue;
                if (all) { ind[m++] = j; continue; }
                double dx = xyz[3 * i] - xyz[3 * j], dy = xyz[3 * i + 1] - xyz[3 * j + 1], dz = xyz[3 * i + 2] - xyz[3 * j + 2];
                if ((dx * dx) + (dy * dy) + (dz * dz) <= c2) ind[m++] = j;
            }
        }
        save_record(s, i, m, ind);
        for (int x = xf; x < xl; x++) { int j = xidx[x]; if (j >= i) break; incl[j] = 1; }
    }
}

long orc_self_pair_list(int n, const double *xyz, const int *xptr, const int *xidx,
                        double cutoff, double cellSize, int method, const unsigned char *qand,
                        long cap_rec, long cap_pairs, int *rec_i, long *rec_ptr, int *js,
                        long *nrec_out)
{
    PairSink s = {rec_i, rec_ptr, js, cap_rec, cap_pairs, 0, 0, 0};
    double c2 = cutoff * cutoff;
    unsigned char *incl = (unsigned char *)malloc((size_t)(n > 0 ? n : 1));
    int *ind = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    if (cap_rec > 0 && rec_ptr) rec_ptr[0] = 0;
    for (int i = 0; i < n; i++) incl[i] = 1;

    if (method == 0) {
        /* Direct (PairListGenerator.c:1104-1150): i > j, j ascending; self
         * exclusion flags stop at j >= i (rows are sorted). */
        for (int i = 1; i < n; i++) {
            int xf = xptr ? xptr[i] : 0, xl = xptr ? xptr[i + 1] : 0, m = 0;
            for (int x = xf; x < xl; x++) { int j = xidx[x]; if (j >= i) break; incl[j] = 0; }
            for (int j = 0; j < i; j++) {
                if (!incl[j]) continue;
                double dx = xyz[3 * i] - xyz[3 * j], dy = xyz[3 * i + 1] - xyz[3 * j + 1], dz = xyz[3 * i + 2] - xyz[3 * j + 2];
                if ((dx * dx) + (dy * dy) + (dz * dz) <= c2) ind[m++] = j;
            }
            save_record(&s, i, m, ind);
            for (int x = xf; x < xl; x++) { int j = xidx[x]; if (j >= i) break; incl[j] = 1; }
        }
    } else {
        /* Cell/cell (PairListGenerator.c:1020-1099); method 2, point/cell, shares the grid and occupancy. */
        int bins[3];
        double lower[3];
        int *cell = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
        long ncells = orc_make_grid(n, xyz, cellSize, bins, lower, cell);
        /* A grid axis with no cells (an exactly zero extent) has no occupancy:
         * RegularGridOccupancy_Allocate rejects zero cells with
         * Status_InvalidArgument (pCx/csource/RegularGridOccupancy.c:29,54). */
        if (ncells <= 0) {
            free(cell);
            free(ind);
            free(incl);
            return ORC_INVALID_ARGUMENT;
        }
        /* Occupancy by counting sort (pCx/csource/RegularGridOccupancy.c:95-126). */
        long *first = (long *)calloc((size_t)ncells + 1, sizeof(long));
        long *total = (long *)calloc((size_t)ncells + 1, sizeof(long));
        int *points = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
        for (int i = 0; i < n; i++) if (cell[i] >= 0) total[cell[i]]++;
        for (long c = 0, acc = 0; c < ncells; c++) { first[c] = acc; acc += total[c]; }
        for (int i = 0; i < n; i++) if (cell[i] >= 0) points[first[cell[i]]++] = i;
        for (long c = 0; c < ncells; c++) first[c] -= total[c];
        int noff = orc_search_range(cellSize, cutoff, 0, NULL, NULL);
        int *off = (int *)malloc(sizeof(int) * 3 * (size_t)noff);
        int *flg = (int *)malloc(sizeof(int) * (size_t)noff);
        orc_search_range(cellSize, cutoff, noff, off, flg);
        long *ncell_id = (long *)malloc(sizeof(long) * (size_t)noff);
        int *ncell_all = (int *)malloc(sizeof(int) * (size_t)noff);
        long stride0 = (long)bins[1] * bins[2], stride1 = bins[2];
        if (method == 2)
            point_cell_records(&s, n, xyz, xptr, xidx, c2, cellSize, cutoff, bins, lower, cell, first, total, points,
                               noff, off, qand, incl, ind);
        else for (long c1 = 0; c1 < ncells; c1++) {
            long p1s = first[c1], p1t = total[c1];
            int ci[3] = {(int)(c1 / stride0), (int)((c1 % stride0) / stride1), (int)(c1 % stride1)};
            int nn = 0;
            /* FindCellsWithinRangeOfCell (RegularGrid.c:274-300). */
            for (int k = 0; k < noff; k++) {
                int a = ci[0] + off[3 * k], b = ci[1] + off[3 * k + 1], c = ci[2] + off[3 * k + 2];
                if (a < 0 || a >= bins[0] || b < 0 || b >= bins[1] || c < 0 || c >= bins[2]) continue;
                ncell_id[nn] = a * stride0 + b * stride1 + c;
                ncell_all[nn] = (flg[k] == 2);
                nn++;
            }
            for (long p1 = 0; p1 < p1t; p1++) {
                int i = points[p1s + p1], m = 0;
                int xf = xptr ? xptr[i] : 0, xl = xptr ? xptr[i + 1] : 0;
                for (int x = xf; x < xl; x++) incl[xidx[x]] = 0;
                for (int k = 0; k < nn; k++) {
                    long c2i = ncell_id[k];
                    if (c2i > c1) continue;
                    long p2s = first[c2i];
                    long up = (c2i == c1) ? p1 : total[c2i];
                    for (long p2 = 0; p2 < up; p2++) {
                        int j = points[p2s + p2];
                        if (!incl[j]) continue;
                        if (ncell_all[k]) { ind[m++] = j; continue; }
                        double dx = xyz[3 * i] - xyz[3 * j], dy = xyz[3 * i + 1] - xyz[3 * j + 1], dz = xyz[3 * i + 2] - xyz[3 * j + 2];
                        if ((dx * dx) + (dy * dy) + (dz * dz) <= c2) ind[m++] = j;
                    }
                }
                save_record(&s, i, m, ind);
                for (int x = xf; x < xl; x++) incl[xidx[x]] = 1;
            }
        }
        free(cell); free(first); free(total); free(points); free(off); free(flg); free(ncell_id); free(ncell_all);
    }
    free(incl);
    free(ind);
    *nrec_out = s.nrec;
    if (s.overflow) return -1 - s.npairs;
    return s.npairs;
}

/* pMx/csource/NBModelABFS.c:691-746. */
int orc_check_for_update(int n, const double *xyz, const double *ref, const uint8_t *fixed,
                         double listCutoff, double outerCutoff, double *maxdisp)
{
    double buf = 0.5 * (listCutoff - outerCutoff), bufsq = buf * buf, maxr2 = 0.0;
    int up = 0;
    for (int i = 0; i < n; i++) {
        if (fixed && fixed[i]) continue;
        double dx = xyz[3 * i] - ref[3 * i], dy = xyz[3 * i + 1] - ref[3 * i + 1], dz = xyz[3 * i + 2] - ref[3 * i + 2];
        double r2 = dx * dx + dy * dy + dz * dz;
        if (r2 > maxr2) maxr2 = r2;
        if (r2 > bufsq) { up = 1; break; }
    }
    if (maxdisp) *maxdisp = sqrt(maxr2);
    return up;
}

/* pMx/csource/PairwiseInteraction.c:292-429 (analytic branch) with the
 * per-thread gradient copies reduced afterwards as in
 * pMx/csource/NBModelABFSState.c:19-38. */
void orc_mmmm_energy(const double *f, int n, const double *xyz, const double *charge,
                     const int *ljtype, int ntypes, const double *tableA, const double *tableB,
                     const int *tableindex, double electrostaticScale, double ljScale,
                     long nrec, const int *rec_i, const long *rec_ptr, const int *js,
                     int nthreads, double *eel, double *elj, double *grad)
{
    double eQQ = 0.0, eLJ = 0.0;
    double escale = electrostaticScale * orc_coulomb_factor();
    int nt = 1;
#ifdef _OPENMP
    nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#endif
    double **tg = (double **)calloc((size_t)nt, sizeof(double *));
    for (int t = 0; t < nt; t++) {
        if (!grad) break;
        if (t == 0) tg[t] = grad;
        else tg[t] = (double *)calloc((size_t)n * 3, sizeof(double));
    }
#ifdef _OPENMP
#pragma omp parallel num_threads(nt)
#endif
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        double *g = grad ? tg[tid] : NULL;
#ifdef _OPENMP
#pragma omp for reduction(+ : eLJ, eQQ) schedule(dynamic)
#endif
        for (long r = 0; r < nrec; r++) {
            int i = rec_i[r];
            double qi = escale * charge[i];
            int ti = ntypes * ljtype[i];
            double xi = xyz[3 * i], yi = xyz[3 * i + 1], zi = xyz[3 * i + 2];
            for (long k = rec_ptr[r]; k < rec_ptr[r + 1]; k++) {
                int j = js[k];
                double xij = xi - xyz[3 * j], yij = yi - xyz[3 * j + 1], zij = zi - xyz[3 * j + 2];
                double r2 = (xij * xij + yij * yij + zij * zij);
                double qij = qi * charge[j];
                int tij = tableindex[ti + ljtype[j]];
                double a = tableA[tij] * ljScale, b = tableB[tij] * ljScale;
                double e1, e2, dF;
                if (!orc_pair_terms(f, r2, qij, a, b, &e1, &e2, &dF)) continue;
                eQQ += e1;
                eLJ += e2;
                if (g) {
                    xij *= (2.0 * dF); yij *= (2.0 * dF); zij *= (2.0 * dF);
                    g[3 * i] += xij; g[3 * i + 1] += yij; g[3 * i + 2] += zij;
                    g[3 * j] -= xij; g[3 * j + 1] -= yij; g[3 * j + 2] -= zij;
                }
            }
        }
    }
    if (grad) {
        for (int t = 1; t < nt; t++) {
            for (long k = 0; k < (long)n * 3; k++) grad[k] += tg[t][k];
            free(tg[t]);
        }
    }
    free(tg);
    *eel = eQQ;
    *elj = eLJ;
}

/* Sums of the magnitudes of the pair terms, sum |e_el| and sum |e_lj| over the
 * same records as orc_mmmm_energy: the scale of the rounding error of any FP64
 * summation order of a cancelling total (checker tolerances; no counterpart in
 * the reference). */
void orc_mmmm_abs_sums(const double *f, int n, const double *xyz, const double *charge,
                       const int *ljtype, int ntypes, const double *tableA, const double *tableB,
                       const int *tableindex, double electrostaticScale, double ljScale,
                       long nrec, const int *rec_i, const long *rec_ptr, const int *js,
                       int nthreads, double *sel, double *slj)
{
    double aQQ = 0.0, aLJ = 0.0;
    double escale = electrostaticScale * orc_coulomb_factor();
    int nt = 1;
#ifdef _OPENMP
    nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#endif
    (void)n;
#ifdef _OPENMP
#pragma omp parallel for num_threads(nt) reduction(+ : aQQ, aLJ) schedule(dynamic)
#endif
    for (long r = 0; r < nrec; r++) {
        int i = rec_i[r];
        double qi = escale * charge[i];
        int ti = ntypes * ljtype[i];
        for (long k = rec_ptr[r]; k < rec_ptr[r + 1]; k++) {
            int j = js[k];
            double xij = xyz[3 * i] - xyz[3 * j], yij = xyz[3 * i + 1] - xyz[3 * j + 1], zij = xyz[3 * i + 2] - xyz[3 * j + 2];
            double r2 = (xij * xij + yij * yij + zij * zij);
            int tij = tableindex[ti + ljtype[j]];
            double e1, e2, dF;
            if (!orc_pair_terms(f, r2, qi * charge[j], tableA[tij] * ljScale, tableB[tij] * ljScale, &e1, &e2, &dF))
                continue;
            aQQ += fabs(e1);
            aLJ += fabs(e2);
        }
    }
    *sel = aQQ;
    *slj = aLJ;
}

/* pMx/csource/LJParameterContainer.c:184-210 (MakeTable). */
void orc_lj_make_table(int ntypes, const double *eps, const double *sigma, int geometric,
                       int four_eps, double *tableA, double *tableB, int *tableindex)
{
    int n = 0;
    for (int i = 0; i < ntypes; i++)
        for (int j = 0; j <= i; j++) {
            double eij = sqrt(eps[i] * eps[j]);
            double sij = geometric ? sqrt(sigma[i] * sigma[j]) : 0.5 * (sigma[i] + sigma[j]);
            double s6 = pow(sij, 6), s12 = s6 * s6;
            if (four_eps) eij *= 4.0;
            else s6 *= 2.0;
            tableA[n] = eij * s12;
            tableB[n] = eij * s6;
            tableindex[j + i * ntypes] = n;
            tableindex[i + j * ntypes] = n;
            n++;
        }
}

/* ==========================================================================
 * Crystal symmetry: the MM/MM image list and image energy.
 * 3x3 matrices are row-major double[9]; every product keeps the reference's
 * operand order (pCx/csource/Matrix33.c), so that, compiled without FMA
 * contraction, the matrices and image coordinates match the reference bit for
 * bit.
 * ========================================================================== */

/* Matrix33_Determinant (pCx/csource/Matrix33.c:96-111). */
static double m33_det(const double *a)
{
    return a[0] * (a[4] * a[8] - a[7] * a[5]) - a[1] * (a[3] * a[8] - a[6] * a[5]) +
           a[2] * (a[3] * a[7] - a[6] * a[4]);
}

/* Matrix33_Invert (pCx/csource/Matrix33.c:165-195): cofactors, then a scale
 * by 1/det. */
static void m33_invert(const double *a, double *r)
{
    double det = m33_det(a), s;
    r[0] = (a[4] * a[8] - a[5] * a[7]);
    r[1] = (a[2] * a[7] - a[8] * a[1]);
    r[2] = (a[1] * a[5] - a[4] * a[2]);
    r[3] = (a[5] * a[6] - a[3] * a[8]);
    r[4] = (a[0] * a[8] - a[2] * a[6]);
    r[5] = (a[2] * a[3] - a[0] * a[5]);
    r[6] = (a[3] * a[7] - a[4] * a[6]);
    r[7] = (a[1] * a[6] - a[0] * a[7]);
    r[8] = (a[0] * a[4] - a[3] * a[1]);
    s = 1.0 / det;
    for (int k = 0; k < 9; k++) r[k] *= s;
}

/* Matrix33_PreMultiplyBy (:330-356): self = other * self. */
static void m33_premul(double *self, const double *o)
{
    double s[9], r[9];
    memcpy(s, self, sizeof(s));
    for (int i = 0; i < 3; i++) {
        double o0 = o[3 * i], o1 = o[3 * i + 1], o2 = o[3 * i + 2];
        r[3 * i] = s[0] * o0 + s[3] * o1 + s[6] * o2;
        r[3 * i + 1] = s[1] * o0 + s[4] * o1 + s[7] * o2;
        r[3 * i + 2] = s[2] * o0 + s[5] * o1 + s[8] * o2;
    }
    memcpy(self, r, sizeof(r));
}

/* Matrix33_PostMultiplyBy (:298-325): self = self * other. */
static void m33_postmul(double *self, const double *o)
{
    double s[9], r[9];
    memcpy(s, self, sizeof(s));
    for (int i = 0; i < 3; i++) {
        double o0 = o[i], o1 = o[3 + i], o2 = o[6 + i];
        r[i] = s[0] * o0 + s[1] * o1 + s[2] * o2;
        r[3 + i] = s[3] * o0 + s[4] * o1 + s[5] * o2;
        r[6 + i] = s[6] * o0 + s[7] * o1 + s[8] * o2;
    }
    memcpy(self, r, sizeof(r));
}

/* Matrix33_ApplyToVector3 (:17-36). */
static void m33_apply(const double *m, double *v)
{
    double x = v[0], y = v[1], z = v[2];
    v[0] = x * m[0] + y * m[1] + z * m[2];
    v[1] = x * m[3] + y * m[4] + z * m[5];
    v[2] = x * m[6] + y * m[7] + z * m[8];
}

/* SymmetryParameters_MakeM (pMx/csource/SymmetryParameters.c:392-425): the
 * columns of M are the lattice vectors in the standard orientation. */
void orc_symmetry_make_m(const double *sp, double *M, double *Minv)
{
    const double dr = M_PI / 180.0e+00;
    double alpha = sp[3] * dr, beta = sp[4] * dr, gamma = sp[5] * dr;
    double ca = cos(alpha), cb = cos(beta), cg = cos(gamma), sg = sin(gamma);
    for (int k = 0; k < 9; k++) M[k] = 0.0;
    M[0] = sp[0];
    M[1] = sp[1] * cg;
    M[4] = sp[1] * sg;
    M[2] = sp[2] * cb;
    M[5] = sp[2] * (ca - cb * cg) / sg;
    M[8] = sp[2] * sqrt(1.0e+00 - ca * ca - cb * cb - cg * cg + 2.0e+00 * ca * cb * cg) / sg;
    m33_invert(M, Minv);
}

/* Transformation3_Orthogonalize (pCx/csource/Transformation3.c:91-99):
 * R' = M R M^-1, T' = M t. */
void orc_orthogonalize(const double *M, const double *Minv, const double *rot, const double *tr, double *Ro, double *To)
{
    memcpy(Ro, rot, 9 * sizeof(double));
    m33_premul(Ro, M);
    m33_postmul(Ro, Minv);
    To[0] = tr[0]; To[1] = tr[1]; To[2] = tr[2];
    m33_apply(M, To);
}

/* Matrix33_IsIdentity / Matrix33_IsEqual with their 1e-6 tolerance
 * (pCx/csource/Matrix33.c:200-253); Transformation3_IsIdentity (Transformation3.c:79-86). */
static int m33_is_equal(const double *a, const double *b)
{
    for (int k = 0; k < 9; k++) if (fabs(a[k] - b[k]) > 1.0e-6) return 0;
    return 1;
}

/* Transformation3Container_FindIdentity and _FindInverses (rotations only;
 * pCx/csource/Transformation3Container.c:52-66,107-141).  Returns the identity's
 * index (-1 if none); inverses[ntrans]. */
int orc_find_inverses(int ntrans, const double *rot, const double *tr, int *inverses)
{
    int identity = -1;
    static const double I[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    for (int t = 0; t < ntrans; t++) {
        if (m33_is_equal(rot + 9 * t, I) && fabs(tr[3 * t]) <= 1.0e-6 && fabs(tr[3 * t + 1]) <= 1.0e-6 &&
            fabs(tr[3 * t + 2]) <= 1.0e-6) { identity = t; break; }
    }
    for (int t = 0; t < ntrans; t++) inverses[t] = -1;
    for (int i = 0; i < ntrans; i++) {
        if (inverses[i] >= 0) continue;
        double m[9];
        m33_invert(rot + 9 * i, m);
        for (int j = 0; j <= i; j++)
            if (inverses[j] < 0 && m33_is_equal(m, rot + 9 * j)) { inverses[i] = j; inverses[j] = i; break; }
    }
    return identity;
}

/* GetLimits of SymmetryParameters_FindBoxSearchLimits (SymmetryParameters.c:182-205). */
static int get_limits(double bl, double bu, double il, double iu, double t, int *low, int *high)
{
    int n = 0;
    while (iu >= bl) { il -= t; iu -= t; n--; }
    while (iu < bl) { il += t; iu += t; n++; }
    if (il <= bu) {
        *low = n;
        while (il <= bu) { il += t; iu += t; n++; }
        *high = n - 1;
        return 1;
    }
    return 0;
}

/* SymmetryParameters_FindBoxSearchLimits (SymmetryParameters.c:207-246). */
static void box_search_limits(const double *M, const double *lower, const double *upper, const double *ilower,
                              const double *iupper, int *lim)
{
    lim[0] = lim[2] = lim[4] = 0;
    lim[1] = lim[3] = lim[5] = -1;
    if (!get_limits(lower[2], upper[2], ilower[2], iupper[2], M[8], &lim[4], &lim[5])) return;
    double d1 = lim[4] * M[5], d2 = lim[5] * M[5];
    double bl = lower[1] - (d1 > d2 ? d1 : d2), bu = upper[1] - (d1 < d2 ? d1 : d2);
    if (!get_limits(bl, bu, ilower[1], iupper[1], M[4], &lim[2], &lim[3])) return;
    d1 = lim[4] * M[2]; d2 = lim[5] * M[2];
    bl = lower[0] - (d1 > d2 ? d1 : d2); bu = upper[0] - (d1 < d2 ? d1 : d2);
    d1 = lim[2] * M[1]; d2 = lim[3] * M[1];
    bl -= (d1 > d2 ? d1 : d2); bu -= (d1 < d2 ? d1 : d2);
    get_limits(bl, bu, ilower[0], iupper[0], M[0], &lim[0], &lim[1]);
}

/* Round (pCx/cinclude/Macros.h:34). */
static int ref_round(double a) { return a >= 0 ? (int)(a + 0.5) : (int)(a - 0.5); }

/* Transformation3Container_FindInverseIntegerTranslation
 * (Transformation3Container.c:70-101). */
static void inverse_integer_translation(const double *rot, const double *tr, int t, int tinv, int a, int b, int c,
                                        int *inv)
{
    const int BIG = -999999;
    double v[3] = {tr[3 * t] + (double)a, tr[3 * t + 1] + (double)b, tr[3 * t + 2] + (double)c};
    m33_apply(rot + 9 * tinv, v);
    for (int i = 0; i < 3; i++) {
        v[i] *= -1.0e+00;
        v[i] += -1.0e+00 * tr[3 * tinv + i];
        int r = ref_round(v[i]);
        inv[i] = (fabs(v[i] - (double)r) < 1.0e-4) ? r : BIG;
    }
}

/* Coordinates3_Rotate then Coordinates3_Translate (pCx/csource/Coordinates3.c:1473-1500,1760-1800). */
static void transform_coordinates(int n, const double *xyz, const double *R, const double *T, double *out)
{
    for (int i = 0; i < n; i++) {
        double x0 = xyz[3 * i], y0 = xyz[3 * i + 1], z0 = xyz[3 * i + 2];
        double x1 = R[0] * x0 + R[1] * y0 + R[2] * z0;
        double y1 = R[3] * x0 + R[4] * y0 + R[5] * z0;
        double z1 = R[6] * x0 + R[7] * y0 + R[8] * z0;
        out[3 * i] = x1 + T[0]; out[3 * i + 1] = y1 + T[1]; out[3 * i + 2] = z1 + T[2];
    }
}

static void translate_coordinates(int n, double *xyz, const double *d)
{
    for (int i = 0; i < n; i++) { xyz[3 * i] += d[0]; xyz[3 * i + 1] += d[1]; xyz[3 * i + 2] += d[2]; }
}

/* Cross pair list between xyz (i) and the image coordinates ixyz (j), direct
 * method (MakeCrossPairListFromCoordinates3Direct, pCore-1.9.0/extensions/
 * csource/PairListGenerator.c:795-858, without exclusions or selections):
 * every pair with the FP64 squared distance (x, y, z order) <= cutoff^2.
 * Records: i ascending, j ascending. */
static long cross_pair_list_direct(int n, const double *xyz, const double *ixyz, double cutoff, PairSink *s)
{
    double c2 = cutoff * cutoff;
    int *ind = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    /* Buckets of the image points on a cutoff-sized grid bound the scan. */
    double lo[3], ext[3];
    enclosing_box(n, ixyz, lo, ext);
    int nb[3];
    for (int d = 0; d < 3; d++) nb[d] = (int)(ext[d] / cutoff) + 1;
    long nbk = (long)nb[0] * nb[1] * nb[2];
    long *first = (long *)calloc((size_t)nbk + 1, sizeof(long));
    int *pts = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    int *bk = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    for (int j = 0; j < n; j++) {
        long b = 0;
        for (int d = 0; d < 3; d++) {
            int k = (int)((ixyz[3 * j + d] - lo[d]) / cutoff);
            if (k >= nb[d]) k = nb[d] - 1;
            b = b * nb[d] + k;
        }
        bk[j] = (int)b;
        first[b + 1]++;
    }
    for (long b = 0; b < nbk; b++) first[b + 1] += first[b];
    {
        long *cur = (long *)malloc(sizeof(long) * (size_t)nbk);
        memcpy(cur, first, sizeof(long) * (size_t)nbk);
        for (int j = 0; j < n; j++) pts[cur[bk[j]]++] = j;   /* ascending j within a bucket */
        free(cur);
    }
    for (int i = 0; i < n; i++) {
        int m = 0, k0[3], k1[3];
        for (int d = 0; d < 3; d++) {
            double f = (xyz[3 * i + d] - lo[d]) / cutoff;
            k0[d] = (int)floor(f) - 1;
            k1[d] = (int)floor(f) + 1;
            if (k0[d] < 0) k0[d] = 0;
            if (k1[d] > nb[d] - 1) k1[d] = nb[d] - 1;
        }
        for (int a = k0[0]; a <= k1[0]; a++)
            for (int b = k0[1]; b <= k1[1]; b++)
                for (int c = k0[2]; c <= k1[2]; c++) {
                    long q = ((long)a * nb[1] + b) * nb[2] + c;
                    for (long p = first[q]; p < first[q + 1]; p++) {
                        int j = pts[p];
                        double dx = xyz[3 * i] - ixyz[3 * j], dy = xyz[3 * i + 1] - ixyz[3 * j + 1],
                               dz = xyz[3 * i + 2] - ixyz[3 * j + 2];
                        if ((dx * dx) + (dy * dy) + (dz * dz) <= c2) ind[m++] = j;
                    }
                }
        /* ascending j */
        for (int x = 1; x < m; x++) {
            int v = ind[x], y = x - 1;
            while (y >= 0 && ind[y] > v) { ind[y + 1] = ind[y]; y--; }
            ind[y + 1] = v;
        }
        save_record(s, i, m, ind);
    }
    free(ind); free(first); free(pts); free(bk);
    return s->npairs;
}

/* Point/cell squared distances between point x and cell idx of a grid with lower corner lower and cell size h
 * (RegularGrid_DistancesSquaredBetweenCellAndPoint, RegularGrid.c:139-180, non-periodic branch). */
static void point_cell_distances(const double *x, const int *idx, const double *lower, double h, double *minR2,
                                 double *maxR2)
{
    double mxs = 0.0, mns = 0.0;
    for (int d = 0; d < 3; d++) {
        double lo = (double)idx[d] * h + lower[d], hi = lo + h, v = x[d], mx, mn;
        if (v > hi) { mx = v - lo; mn = v - hi; }
        else if (v < lo) { mx = hi - v; mn = lo - v; }
        else { mx = (v - lo) > (hi - v) ? (v - lo) : (hi - v); mn = 0.0; }
        mxs += mx * mx;
        mns += mn * mn;
    }
    *minR2 = mns;
    *maxR2 = mxs;
}

/* Cross pair list between xyz (i, on the primary grid: bins1, lower1, cell size h, cell1[i] or -1) and the image
 * coordinates ixyz (j), grid method, without exclusions or selections
 * (PairListGenerator_CrossPairListFromDoubleCoordinates3, PairListGenerator.c:414-445, with
 * MakeCrossPairListFromCoordinates3CellCell / PointCell, :703-794 and :861-940).
 *
 * The image points get a grid conforming to the primary one (Coordinates3_MakeConformingGrid,
 * Coordinates3.c:998-1072): its lower cell indices lowerN are those of the image points' box corner on the
 * primary grid (floor, no regularisation), its bins reach the cell of the box's upper corner, and its lower corner
 * is lower1 + lowerN h; each image point takes its cell on that grid or none (RegularGridOccupancy_Fill,
 * RegularGridOccupancy.c:77-130).  The cells searched for a primary cell I are I + o - lowerN over the stencil
 * offsets o, where on the image grid (RegularGrid.c:341-405).
 *
 * method 1 (cell/cell): i in the primary cells in cell order; a cell is taken whole where the stencil flag of o is
 * 2 (fully within range), otherwise each j by the distance test.
 * method 2 (point/cell): i ascending over all points; I is i's cell or, for a point off the grid, the cell indices
 * decoded last (the latest earlier point on the grid, else (-u, -u, -u), as in the self list); a cell is searched
 * when its smallest squared distance to x_i on the image grid is below cutoff^2 and taken whole when its largest
 * is at most cutoff^2.  Records in the reference's order; j ascending within a cell. */
static long cross_pair_list_grid(int n, const double *xyz, const double *ixyz, int method, double cutoff, double h,
                                 const int *bins1, const double *lower1, const int *cell1, int noff, const int *off,
                                 const int *flg, PairSink *s)
{
    double c2 = cutoff * cutoff;
    double lo[3], ext[3], up[3], lower2[3];
    int lowerN[3], bins2[3];
    enclosing_box(n, ixyz, lo, ext);
    for (int d = 0; d < 3; d++) {
        up[d] = ext[d] + lo[d];
        lowerN[d] = (int)floor((lo[d] - lower1[d]) / h);
        int upperN = (int)floor((up[d] - lower1[d]) / h);
        bins2[d] = upperN - lowerN[d] + 1;
        lower2[d] = lower1[d] + (double)lowerN[d] * h;
    }
    long ncell2 = (long)bins2[0] * bins2[1] * bins2[2];
    long st2[3] = {(long)bins2[1] * bins2[2], bins2[2], 1};
    long *first = (long *)calloc((size_t)ncell2 + 1, sizeof(long));
    long *total = (long *)calloc((size_t)ncell2 + 1, sizeof(long));
    int *points = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    int *c2id = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    int *ind = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    for (int j = 0; j < n; j++) {
        long c = 0;
        int offg = 0;
        for (int d = 0; d < 3; d++) {
            int k = (int)floor((ixyz[3 * j + d] - lower2[d]) / h);
            if (k < 0 || k >= bins2[d]) offg = 1;
            c += st2[d] * k;
        }
        c2id[j] = offg ? -1 : (int)c;
        if (!offg) total[c]++;
    }
    for (long c = 0, acc = 0; c < ncell2; c++) { first[c] = acc; acc += total[c]; }
    for (int j = 0; j < n; j++) if (c2id[j] >= 0) points[first[c2id[j]]++] = j;
    for (long c = 0; c < ncell2; c++) first[c] -= total[c];
    const long st1[3] = {(long)bins1[1] * bins1[2], bins1[2], 1};
    const int u = (int)ceil(cutoff / h);
    int pc[3] = {-u, -u, -u};
    /* Outer order: cell/cell walks the primary cells, point/cell the points. */
    long ncell1 = (long)bins1[0] * bins1[1] * bins1[2];
    int *order = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    int norder = 0;
    if (method == 1) {
        long *cnt = (long *)calloc((size_t)ncell1 + 1, sizeof(long));
        for (int i = 0; i < n; i++) if (cell1[i] >= 0) cnt[cell1[i] + 1]++;
        for (long c = 0; c < ncell1; c++) cnt[c + 1] += cnt[c];
        for (int i = 0; i < n; i++) if (cell1[i] >= 0) order[cnt[cell1[i]]++] = i;
        norder = (int)cnt[ncell1 - 1];
        free(cnt);
    } else {
        for (int i = 0; i < n; i++) order[i] = i;
        norder = n;
    }
    for (int q = 0; q < norder; q++) {
        int i = order[q], m = 0;
        if (cell1[i] >= 0) {
            long c = cell1[i];
            pc[0] = (int)(c / st1[0]);
            pc[1] = (int)((c % st1[0]) / st1[1]);
            pc[2] = (int)(c % st1[1]);
        }
        for (int k = 0; k < noff; k++) {
            int idx[3] = {pc[0] + off[3 * k] - lowerN[0], pc[1] + off[3 * k + 1] - lowerN[1],
                          pc[2] + off[3 * k + 2] - lowerN[2]};
            if (idx[0] < 0 || idx[0] >= bins2[0] || idx[1] < 0 || idx[1] >= bins2[1] || idx[2] < 0 ||
                idx[2] >= bins2[2])
                continue;
            int all;
            if (method == 1) {
                all = flg[k] == 2;
            } else {
                double minR2, maxR2;
                point_cell_distances(xyz + 3 * i, idx, lower2, h, &minR2, &maxR2);
                if (!(minR2 < c2)) continue;
                all = maxR2 <= c2;
            }
            const long cid = idx[0] * st2[0] + idx[1] * st2[1] + idx[2];
            for (long p = first[cid], pe = first[cid] + total[cid]; p < pe; p++) {
                int j = points[p];
                if (all) { ind[m++] = j; continue; }
                double dx = xyz[3 * i] - ixyz[3 * j], dy = xyz[3 * i + 1] - ixyz[3 * j + 1],
                       dz = xyz[3 * i + 2] - ixyz[3 * j + 2];
                if ((dx * dx) + (dy * dy) + (dz * dz) <= c2) ind[m++] = j;
            }
        }
        save_record(s, i, m, ind);
    }
    free(order); free(first); free(total); free(points); free(c2id); free(ind);
    return s->npairs;
}

/* GenerateImageLists for the MM/MM list of a pure MM system
 * (pMx/csource/NBModelABFS.c:753-1045): transformations t (fractional
 * rotation rot[9t..], translation tr[3t..]) with their inverses, box search
 * limits, inverse skipping and scales, the box-overlap test and one cross
 * pair list per image.  The image coordinates are generated as the reference
 * does: transform once per t, then translate by +d and back by -d for every
 * (a, b, c) that is not skipped.
 *
 * Outputs per image (at most cap_img, in reference order, images with no
 * pairs dropped as ImageList_CreateImage does, pMx/csource/ImageList.c:31-62):
 * abct[4k..] = a, b, c, t; scale[k]; records appended to the sink with
 * img_rec[k] = first record of image k (img_rec[nimg] = total records).
 * method: 0 direct, 1 grid cell/cell, 2 grid point/cell; the primary grid is
 * made from xyz with cellSize.  Returns the number of images, or -1 on a
 * capacity overflow; sizes[3] = images, records, pairs needed. */
long orc_image_lists(int n, const double *xyz, int method, double cellSize, int ntrans, const double *rot,
                     const double *tr, const double *sp, double cutoff, int checkForInverses, int expandFactor,
                     long cap_img, int *abct, double *scale, long *img_rec,
                     long cap_rec, long cap_pairs, int *rec_i, long *rec_ptr, int *js, long *sizes)
{
    double M[9], Minv[9], lower[3], upper[3], ext[3];
    int *inverses = (int *)malloc(sizeof(int) * (size_t)(ntrans > 0 ? ntrans : 1));
    int identity = orc_find_inverses(ntrans, rot, tr, inverses);
    double *ixyz = (double *)malloc(sizeof(double) * 3 * (size_t)(n > 0 ? n : 1));
    PairSink s = {rec_i, rec_ptr, js, cap_rec, cap_pairs, 0, 0, 0};
    long nimg = 0;
    if (cap_rec > 0 && rec_ptr) rec_ptr[0] = 0;
    int bins1[3] = {0, 0, 0}, noff = 0, *cell1 = NULL, *off = NULL, *flg = NULL;
    double lower1[3] = {0.0, 0.0, 0.0};
    if (method != 0 && n > 0) {
        cell1 = (int *)malloc(sizeof(int) * (size_t)n);
        orc_make_grid(n, xyz, cellSize, bins1, lower1, cell1);
        noff = orc_search_range(cellSize, cutoff, 0, NULL, NULL);
        off = (int *)malloc(sizeof(int) * 3 * (size_t)noff);
        flg = (int *)malloc(sizeof(int) * (size_t)noff);
        orc_search_range(cellSize, cutoff, noff, off, flg);
    }
    orc_symmetry_make_m(sp, M, Minv);
    enclosing_box(n, xyz, lower, ext);
    for (int d = 0; d < 3; d++) { upper[d] = ext[d] + lower[d]; lower[d] += -cutoff; upper[d] += cutoff; }
    for (int t = 0; t < ntrans && n > 0; t++) {
        int tinv = checkForInverses ? inverses[t] : -1, qtskip = 0;
        double defscale = 0.5;
        if (tinv >= 0) {
            if (t < tinv) continue;          /* pure MM: skip, the inverse comes later */
            defscale = 1.0;
        }
        double Ro[9], To[3], ilower[3], iupper[3], iext[3];
        orc_orthogonalize(M, Minv, rot + 9 * t, tr + 3 * t, Ro, To);
        transform_coordinates(n, xyz, Ro, To, ixyz);
        enclosing_box(n, ixyz, ilower, iext);
        for (int d = 0; d < 3; d++) iupper[d] = iext[d] + ilower[d];
        int lim[6];
        box_search_limits(M, lower, upper, ilower, iupper, lim);
        if (expandFactor > 0) {
            lim[0] -= expandFactor; lim[1] += expandFactor;
            lim[2] -= expandFactor; lim[3] += expandFactor;
            lim[4] -= expandFactor; lim[5] += expandFactor;
        }
        for (int a = lim[0]; a <= lim[1]; a++)
            for (int b = lim[2]; b <= lim[3]; b++)
                for (int c = lim[4]; c <= lim[5]; c++) {
                    if (a == 0 && b == 0 && c == 0 && t == identity) continue;
                    double sc = defscale;
                    if (tinv >= 0 && tinv == t) {
                        int inv[3];
                        inverse_integer_translation(rot, tr, t, tinv, a, b, c, inv);
                        if (inv[0] >= lim[0] && inv[0] <= lim[1] && inv[1] >= lim[2] && inv[1] <= lim[3] &&
                            inv[2] >= lim[4] && inv[2] <= lim[5]) {
                            if (a == inv[0] && b == inv[1] && c == inv[2]) sc = 0.5;
                            else {
                                sc = 1.0;
                                if (a < inv[0] || (a == inv[0] && b < inv[1]) || (a == inv[0] && b == inv[1] && c < inv[2]))
                                    continue;
                            }
                        } else sc = 1.0;
                    }
                    (void)qtskip;
                    /* SymmetryParameters_Displacement (SymmetryParameters.c:166-178). */
                    double dsp[3], na = (double)a, nb = (double)b, nc = (double)c;
                    for (int i = 0; i < 3; i++) dsp[i] = na * M[3 * i] + nb * M[3 * i + 1] + nc * M[3 * i + 2];
                    translate_coordinates(n, ixyz, dsp);
                    for (int d = 0; d < 3; d++) { ilower[d] += dsp[d]; iupper[d] += dsp[d]; }
                    if (ilower[0] <= upper[0] && ilower[1] <= upper[1] && ilower[2] <= upper[2] &&
                        iupper[0] >= lower[0] && iupper[1] >= lower[1] && iupper[2] >= lower[2]) {
                        long r0 = s.nrec, p0 = s.npairs;
                        if (method == 0)
                            cross_pair_list_direct(n, xyz, ixyz, cutoff, &s);
                        else
                            cross_pair_list_grid(n, xyz, ixyz, method, cutoff, cellSize, bins1, lower1, cell1, noff,
                                                 off, flg, &s);
                        if (s.npairs > p0) {
                            if (nimg < cap_img) {
                                abct[4 * nimg] = a; abct[4 * nimg + 1] = b; abct[4 * nimg + 2] = c;
                                abct[4 * nimg + 3] = t;
                                scale[nimg] = sc;
                                img_rec[nimg] = r0;
                            }
                            nimg++;
                        }
                    }
                    for (int d = 0; d < 3; d++) dsp[d] *= -1.0e+00;
                    translate_coordinates(n, ixyz, dsp);
                    for (int d = 0; d < 3; d++) { ilower[d] += dsp[d]; iupper[d] += dsp[d]; }
                }
    }
    if (nimg <= cap_img && img_rec) img_rec[nimg] = s.nrec;
    free(inverses);
    free(ixyz);
    free(cell1);
    free(off);
    free(flg);
    sizes[0] = nimg;
    sizes[1] = s.nrec;
    sizes[2] = s.npairs;
    if (s.overflow || nimg > cap_img) return -1;
    return nimg;
}

/* Matrix33_InverseDerivative (pCx/csource/Matrix33.c:116-160): d(M^-1)/dM_ij. */
static void m33_inverse_derivative(const double *m, int i, int j, double *o)
{
    double ddet = 0.0;
    double m00 = m[0], m01 = m[1], m02 = m[2], m10 = m[3], m11 = m[4], m12 = m[5], m20 = m[6], m21 = m[7], m22 = m[8];
    m33_invert(m, o);
    switch (3 * i + j) {
    case 0: ddet = -m12 * m21 + m11 * m22; break;
    case 1: ddet = m12 * m20 - m10 * m22; break;
    case 2: ddet = -m11 * m20 + m10 * m21; break;
    case 3: ddet = m02 * m21 - m01 * m22; break;
    case 4: ddet = -m02 * m20 + m00 * m22; break;
    case 5: ddet = m01 * m20 - m00 * m21; break;
    case 6: ddet = -m02 * m11 + m01 * m12; break;
    case 7: ddet = m02 * m10 - m00 * m12; break;
    case 8: ddet = -m01 * m10 + m00 * m11; break;
    }
    for (int k = 0; k < 9; k++) o[k] *= -ddet;
    switch (3 * i + j) {
    case 0: o[4] += m22; o[5] -= m12; o[7] -= m21; o[8] += m11; break;
    case 1: o[1] -= m22; o[2] += m12; o[7] += m20; o[8] -= m10; break;
    case 2: o[1] += m21; o[2] -= m11; o[4] -= m20; o[5] += m10; break;
    case 3: o[3] -= m22; o[5] += m02; o[6] += m21; o[8] -= m01; break;
    case 4: o[0] += m22; o[2] -= m02; o[6] -= m20; o[8] += m00; break;
    case 5: o[0] -= m21; o[2] += m01; o[3] += m20; o[5] -= m00; break;
    case 6: o[3] += m12; o[4] -= m02; o[6] -= m11; o[7] += m01; break;
    case 7: o[0] -= m12; o[1] += m02; o[6] += m10; o[7] -= m00; break;
    case 8: o[0] += m11; o[1] -= m01; o[3] -= m10; o[4] += m00; break;
    }
    double det = m33_det(m), s = 1.0 / det;
    for (int k = 0; k < 9; k++) o[k] *= s;
}

/* SymmetryParameterGradients_ImageDerivatives (pMx/csource/
 * SymmetryParameterGradients.c:142-216): dE/dM of one image from its
 * unrotated image gradients; xrot/xtr is the fractional transformation with
 * the integer translation added. */
static void image_derivatives(const double *M, const double *Minv, const double *xrot, const double *xtr, int n,
                              const double *xyz, const double *ig, double *dEdM)
{
    double ms[9], si[9];
    memcpy(ms, xrot, sizeof(ms)); m33_premul(ms, M);
    memcpy(si, xrot, sizeof(si)); m33_postmul(si, Minv);
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) {
            double di[9];
            m33_inverse_derivative(M, a, b, di);
            m33_premul(di, ms);
            di[3 * a] += si[3 * b]; di[3 * a + 1] += si[3 * b + 1]; di[3 * a + 2] += si[3 * b + 2];
            double t = xtr[b], sum = 0.0;
            for (int i = 0; i < n; i++) {
                double x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
                double dx = di[0] * x + di[1] * y + di[2] * z;
                double dy = di[3] * x + di[4] * y + di[5] * z;
                double dz = di[6] * x + di[7] * y + di[8] * z;
                if (a == 0) dx += t; else if (a == 1) dy += t; else dz += t;
                sum += dx * ig[3 * i] + dy * ig[3 * i + 1] + dz * ig[3 * i + 2];
            }
            dEdM[3 * a + b] += sum;
        }
}

/* MMMMImageEnergy (pMx/csource/NBModelABFS.c:1161-1290), serially over the
 * images: for image k the fractional transformation gets (a, b, c) added,
 * is orthogonalized with the current M, the image coordinates are generated,
 * PairwiseInteractionABFS_MMMMEnergy runs over the image's records with
 * electrostatic and LJ scales times the image scale (i gradients into grad,
 * j gradients into the image gradient), then the lattice derivative, and the
 * image gradient rotated by R'^T is added to grad.  grad and dEdM (row-major,
 * incremented) may be NULL. */
void orc_image_energy(const double *f, int n, const double *xyz, const double *charge, const int *ljtype, int ntypes,
                      const double *tableA, const double *tableB, const int *tableindex, double electrostaticScale,
                      const double *rot, const double *tr, const double *sp, long nimg, const int *abct,
                      const double *scale, const long *img_rec, const int *rec_i, const long *rec_ptr, const int *js,
                      double *eel, double *elj, double *grad, double *dEdM)
{
    double M[9], Minv[9];
    double *ixyz = (double *)malloc(sizeof(double) * 3 * (size_t)(n > 0 ? n : 1));
    double *ig = grad ? (double *)malloc(sizeof(double) * 3 * (size_t)(n > 0 ? n : 1)) : NULL;
    double eQQ = 0.0, eLJ = 0.0;
    orc_symmetry_make_m(sp, M, Minv);
    for (long k = 0; k < nimg; k++) {
        int t = abct[4 * k + 3];
        double xtr[3] = {tr[3 * t] + (double)abct[4 * k], tr[3 * t + 1] + (double)abct[4 * k + 1],
                         tr[3 * t + 2] + (double)abct[4 * k + 2]};
        double Ro[9], To[3];
        orc_orthogonalize(M, Minv, rot + 9 * t, xtr, Ro, To);
        transform_coordinates(n, xyz, Ro, To, ixyz);
        if (ig) memset(ig, 0, sizeof(double) * 3 * (size_t)n);
        double escale = electrostaticScale * scale[k] * orc_coulomb_factor(), ljs = scale[k];
        for (long r = img_rec[k]; r < img_rec[k + 1]; r++) {
            int i = rec_i[r];
            double qi = escale * charge[i];
            int ti = ntypes * ljtype[i];
            double xi = xyz[3 * i], yi = xyz[3 * i + 1], zi = xyz[3 * i + 2];
            for (long q = rec_ptr[r]; q < rec_ptr[r + 1]; q++) {
                int j = js[q];
                double xij = xi - ixyz[3 * j], yij = yi - ixyz[3 * j + 1], zij = zi - ixyz[3 * j + 2];
                double r2 = (xij * xij + yij * yij + zij * zij);
                double qij = qi * charge[j];
                int tij = tableindex[ti + ljtype[j]];
                double e1, e2, dF;
                if (!orc_pair_terms(f, r2, qij, tableA[tij] * ljs, tableB[tij] * ljs, &e1, &e2, &dF)) continue;
                eQQ += e1;
                eLJ += e2;
                if (grad) {
                    xij *= (2.0 * dF); yij *= (2.0 * dF); zij *= (2.0 * dF);
                    grad[3 * i] += xij; grad[3 * i + 1] += yij; grad[3 * i + 2] += zij;
                    ig[3 * j] -= xij; ig[3 * j + 1] -= yij; ig[3 * j + 2] -= zij;
                }
            }
        }
        if (grad) {
            if (dEdM) image_derivatives(M, Minv, rot + 9 * t, xtr, n, xyz, ig, dEdM);
            /* Coordinates3_Rotate by R'^T, then add. */
            for (int i = 0; i < n; i++) {
                double x0 = ig[3 * i], y0 = ig[3 * i + 1], z0 = ig[3 * i + 2];
                grad[3 * i] += Ro[0] * x0 + Ro[3] * y0 + Ro[6] * z0;
                grad[3 * i + 1] += Ro[1] * x0 + Ro[4] * y0 + Ro[7] * z0;
                grad[3 * i + 2] += Ro[2] * x0 + Ro[5] * y0 + Ro[8] * z0;
            }
        }
    }
    free(ixyz);
    free(ig);
    *eel = eQQ;
    *elj = eLJ;
}

/* SymmetryParameterGradients_CrystalDerivatives (SymmetryParameterGradients.c:50-82):
 * out = dE/da, dE/db, dE/dc, dE/dalpha, dE/dbeta, dE/dgamma (angles in degrees). */
void orc_crystal_derivatives(const double *sp, const double *dEdM, double *out)
{
    const double dr = M_PI / 180.0e+00;
    double alpha = sp[3] * dr, beta = sp[4] * dr, gamma = sp[5] * dr;
    double ca = cos(alpha), cb = cos(beta), cg = cos(gamma), sg = sin(gamma);
    double f12 = (ca - cb * cg);
    double f22 = sqrt(1.0e+00 - ca * ca - cb * cb - cg * cg + 2.0e+00 * ca * cb * cg);
    const double *d = dEdM;
    out[0] = d[0];
    out[1] = cg * d[1] + sg * d[4];
    out[2] = cb * d[2] + f12 * d[5] / sg + f22 * d[8] / sg;
    out[3] = -(d[5] - f12 * d[8] / f22) * (sp[2] * sin(alpha) * dr) / sg;
    out[4] = -(sg * d[2] - cg * d[5] + (ca * cg - cb) * d[8] / f22) * (sp[2] * sin(beta) * dr) / sg;
    out[5] = (sp[1] * (-sg * d[1] + cg * d[4]) +
              sp[2] * ((cb - f12 * cg / (sg * sg)) * d[5] + ((cg - ca * cb) / f22 - f22 * cg / (sg * sg)) * d[8])) * dr;
}
