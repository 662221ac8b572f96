/*
 * abfs_oracle.h -- CPU FP64 restatement of pDynamo's NBModelABFS MM/MM path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product
 * path: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * load it.  The product library never links or calls it.
 *
 * Every function restates one piece of the reference algorithm and cites the
 * reference file:line it follows (paths relative to the pDynamo-1.9.0 tree,
 * pCx = pCore-1.9.0/extensions, pMx = pMolecule-1.9.0/extensions).
 */
#ifndef ABFS_ORACLE_H
#define ABFS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Index of each switching factor inside the 21-double factor block.
 * Mirrors the field set of PairwiseInteractionABFSFactors
 * (pMx/cinclude/PairwiseInteraction.h:40-66). */
enum {
    ORC_R2DAMP = 0, ORC_R2OFF, ORC_R2ON,
    ORC_A, ORC_B, ORC_C, ORC_D, ORC_QSHIFT1, ORC_QSHIFT2, ORC_QF0, ORC_QALPHA,
    ORC_AF6, ORC_AK12, ORC_ASHIFT12, ORC_AF0, ORC_AALPHA,
    ORC_BF3, ORC_BK6, ORC_BSHIFT6, ORC_BF0, ORC_BALPHA,
    ORC_NFACTORS
};

/* Coulomb factor in kJ mol^-1 A e^-2 (pCx/cinclude/Units.h:18-30,48-49). */
double orc_coulomb_factor(void);

/* Switching factors (pMx/csource/PairwiseInteraction.c:89-141). */
void orc_make_factors(double damp, double inner, double outer, double *f);

/* One pair: energies and dF = dE/d(r^2) for r2 <= r2Off
 * (pMx/cinclude/PairwiseInteraction.h:72-119). Returns 0 if r2 > r2Off. */
int orc_pair_terms(const double *f, double r2, double qij, double aij, double bij,
                   double *eel, double *elj, double *df);

/* Grid search decision (pCx/csource/PairListGenerator.c:498-525). */
int orc_determine_method(int n, const double *xyz, double cutoff, double cutoffCellSizeFactor,
                         int minimumCellExtent, int minimumPoints);

/* Reference cell grid (pCx/csource/Coordinates3.c:1111-1158) and cell of each
 * point (pCx/csource/RegularGrid.c:240-251); off-grid points get -1.
 * bins[3], lower[3] out; cell[n] out. Returns the number of cells. */
long orc_make_grid(int n, const double *xyz, double cellSize, int *bins, double *lower, int *cell);

/* Search stencil (pCx/csource/RegularGrid.c:526-575): offsets[k*3+d], flags[k]
 * (1 = test distance, 2 = include all).  Returns the number of offsets; cap is
 * the capacity of offsets/flags in entries. */
int orc_search_range(double cellSize, double cutoff, int cap, int *offsets, int *flags);

/* Self pair list (pCx/csource/PairListGenerator.c:946-1218).  method: 1 grid
 * cell/cell, 2 grid point/cell, 0 direct.  qand (n, or NULL for every atom):
 * the AND selection of the point/cell search, whose points outside it neither
 * search nor decode a cell (the other methods list every atom; callers filter
 * their rows).  Exclusions as symmetric CSR sorted ascending per row.
 * Output records in reference order: rec_i[r], rec_ptr[r..r+1] into js.
 * Returns npairs, or -1 if capacities are too small (need_* report sizes), or
 * ORC_INVALID_ARGUMENT for a cell grid with no cells (a zero extent). */
#define ORC_INVALID_ARGUMENT (-(1L << 62))
long orc_self_pair_list(int n, const double *xyz, const int *xptr, const int *xidx,
                        double cutoff, double cellSize, int method, const unsigned char *qand,
                        long cap_rec, long cap_pairs, int *rec_i, long *rec_ptr, int *js,
                        long *nrec_out);

/* Half-skin rebuild rule (pMx/csource/NBModelABFS.c:691-746); fixed may be NULL. */
int orc_check_for_update(int n, const double *xyz, const double *ref, const uint8_t *fixed,
                         double listCutoff, double outerCutoff, double *maxdisp);

/* Analytic MM/MM energy over a record list with OpenMP per-thread gradient
 * copies (pMx/csource/PairwiseInteraction.c:292-429,532-535;
 * pMx/csource/NBModelABFSState.c:19-38).  grad (n*3) is incremented and may be
 * NULL.  nthreads <= 0 means the OpenMP default. */
void orc_mmmm_energy(const double *f, int n, const double *xyz, const double *charge,
                     const int *ljtype, int ntypes, const double *tableA, const double *tableB,
                     const int *tableindex, double electrostaticScale, double ljScale,
                     long nrec, const int *rec_i, const long *rec_ptr, const int *js,
                     int nthreads, double *eel, double *elj, double *grad);

/* Sum |e_el| and sum |e_lj| of the pair terms over the same records (the
 * rounding scale of a cancelling FP64 total; checker tolerances only). */
void orc_mmmm_abs_sums(const double *f, int n, const double *xyz, const double *charge,
                       const int *ljtype, int ntypes, const double *tableA, const double *tableB,
                       const int *tableindex, double electrostaticScale, double ljScale,
                       long nrec, const int *rec_i, const long *rec_ptr, const int *js,
                       int nthreads, double *sel, double *slj);

/* LJ tables from epsilon/sigma (pMx/csource/LJParameterContainer.c:184-210).
 * geometric: 1 = geometric sigma (OPLS), 0 = arithmetic (AMBER); four_eps: 1 =
 * A=4 eps s^12 form (OPLS), 0 = B = eps 2 s^6 form (AMBER). */
void orc_lj_make_table(int ntypes, const double *eps, const double *sigma, int geometric,
                       int four_eps, double *tableA, double *tableB, int *tableindex);

/* ---- Crystal symmetry and the MM/MM image list (pMx/csource/NBModelABFS.c:
 * 625-684,753-1045,1161-1290; pMx/csource/SymmetryParameters.c;
 * pMx/csource/SymmetryParameterGradients.c; pCx/csource/Transformation3*.c).
 * sp = a, b, c, alpha, beta, gamma (degrees); rot 9 per transformation
 * (fractional, row-major), tr 3 per transformation; matrices row-major. */
void orc_symmetry_make_m(const double *sp, double *M, double *Minv);
void orc_orthogonalize(const double *M, const double *Minv, const double *rot, const double *tr, double *Ro,
                       double *To);
int orc_find_inverses(int ntrans, const double *rot, const double *tr, int *inverses);
long orc_image_lists(int n, const double *xyz, int method, double cellSize, int ntrans, const double *rot,
                     const double *tr, const double *sp, double cutoff, int checkForInverses, int expandFactor,
                     long cap_img, int *abct, double *scale, long *img_rec,
                     long cap_rec, long cap_pairs, int *rec_i, long *rec_ptr, int *js, long *sizes);
void orc_image_energy(const double *f, int n, const double *xyz, const double *charge, const int *ljtype, int ntypes,
                      const double *tableA, const double *tableB, const int *tableindex, double electrostaticScale,
                      const double *rot, const double *tr, const double *sp, long nimg, const int *abct,
                      const double *scale, const long *img_rec, const int *rec_i, const long *rec_ptr, const int *js,
                      double *eel, double *elj, double *grad, double *dEdM);
void orc_crystal_derivatives(const double *sp, const double *dEdM, double *out);

#ifdef __cplusplus
}
#endif
#endif
