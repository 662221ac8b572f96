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

long orc_self_pair_list(int n, const double *xyz, const int *xptr, const int *xidx,
                        double cutoff, double cellSize, int method,
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
        /* Cell/cell (PairListGenerator.c:1020-1099). */
        int bins[3];
        double lower[3];
        int *cell = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
        long ncells = orc_make_grid(n, xyz, cellSize, bins, lower, cell);
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
        for (long c1 = 0; c1 < ncells; c1++) {
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
