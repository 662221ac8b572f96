/*
 * ref_driver.c -- thin C driver over the REAL pDynamo 1.9.0 C library, built
 * from the reference sources in place (see oracle/Makefile, target `ref`).
 *
 * TEST INFRASTRUCTURE ONLY: it produces golden vectors (tests/golden/) and can
 * serve as the timed CPU baseline.  It calls the reference's own entry points
 * in the order pMolecule.NBModelABFS.pyx uses them:
 *   NBModelABFSState_SetUp  (pyx:242)        -> state
 *   NBModelABFSState_Initialize (pyx:268)    -> per call
 *   NBModelABFS_Update (pyx:271)             -> rebuild decision + lists
 *   NBModelABFS_MMMMEnergy (pyx:119)         -> energies and gradients
 * No reference source is copied here; the reference headers are included from
 * the reference tree at build time.
 */
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "Coordinates3.h"
#include "CubicSpline.h"
#include "LJParameterContainer.h"
#include "MMAtomContainer.h"
#include "Memory.h"
#include "NBModelABFS.h"
#include "NBModelABFSState.h"
#include "PairList.h"
#include "PairListGenerator.h"
#include "PairwiseInteraction.h"
#include "QCAtomContainer.h"
#include "Real1DArray.h"
#include "Selection.h"
#include "SymmetryParameterGradients.h"
#include "SymmetryParameters.h"
#include "Transformation3Container.h"

typedef struct {
    int n;
    NBModelABFS *model;
    PairListGenerator *gen;
    PairwiseInteractionABFS *pw;
    MMAtomContainer *atoms;
    LJParameterContainer *lj, *lj14;
    PairList *excl, *i14;
    NBModelABFSState *state;
    Coordinates3 *crd, *grd;
    Transformation3Container *trans;     /* NULL: no symmetry */
    SymmetryParameters *sp;
    SymmetryParameterGradients *spg;
    /* QC/MM (NULL without QC atoms). */
    QCAtomContainer *qc;
    Selection *fixed;
    Real1DArray *qcCharges, *qcmmPotentials;
    PairwiseInteractionABFS *pwqcmm, *pwqcqc;
} RefHandle;

static long dump_list(PairList *pl, int *oi, int *oj, long cap);

static PairList *make_self_list(int npairs, const int *pairs)
{
    if (npairs <= 0) return NULL;
    Integer *tmp = (Integer *)malloc(sizeof(Integer) * 2 * (size_t)npairs);
    for (int k = 0; k < 2 * npairs; k++) tmp[k] = pairs[k];
    PairList *pl = PairList_FromIntegerPairArray(True, npairs, tmp);
    free(tmp);
    return pl;
}

static LJParameterContainer *make_lj(int ntypes, const double *tA, const double *tB)
{
    LJParameterContainer *lj = LJParameterContainer_Allocate(ntypes);
    int size = ntypes * (ntypes + 1) / 2;
    for (int k = 0; k < size; k++) { lj->tableA[k] = tA[k]; lj->tableB[k] = tB[k]; }
    return lj;
}

/* opts: damping, inner, outer, list, dielectric, scale14.
 * gen: cutoffCellSizeFactor, minimumCellExtent, minimumPoints, useGridByCell
 * (the ABFS default generator of pMolecule.NBModelABFS.pyx:59-71 otherwise). */
static Transformation3Container *make_transformations(int ntrans, const double *rot, const double *trans)
{
    Transformation3Container *tc = Transformation3Container_Allocate(ntrans);
    tc->QOWNER = True;
    for (int t = 0; t < ntrans; t++) {
        Transformation3 *x = Transformation3_Allocate();
        for (int r = 0; r < 3; r++) {
            for (int c = 0; c < 3; c++) Matrix33_Item(x->rotation, r, c) = rot[9 * t + 3 * r + c];
            Vector3_Item(x->translation, r) = trans[3 * t + r];
        }
        tc->items[t] = x;
    }
    /* As pCore.Transformation3Container.FillCObject (pyx:93-104). */
    Transformation3Container_FindIdentity(tc);
    Transformation3Container_FindInverses(tc);
    return tc;
}

static void *create_common(int n, const double *charges, const int *ljtypes, int ntypes,
                           const double *tA, const double *tB, const double *tA14, const double *tB14,
                           int nexcl, const int *excl, int n14, const int *p14,
                           const double *opts, const double *gen, int nthreads,
                           Transformation3Container *tc, int checkForInverses, int expandFactor,
                           QCAtomContainer *qc, Selection *fixed, int coupling)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
    RefHandle *h = (RefHandle *)calloc(1, sizeof(RefHandle));
    h->n = n;
    h->model = NBModelABFS_Allocate();
    h->model->dampingCutoff = opts[0];
    h->model->innerCutoff = opts[1];
    h->model->outerCutoff = opts[2];
    h->model->listCutoff = opts[3];
    h->model->dielectric = opts[4];
    h->model->electrostaticScale14 = opts[5];
    h->gen = PairListGenerator_Allocate();
    h->gen->cutoff = opts[3];
    h->gen->cutoffCellSizeFactor = gen[0];
    h->gen->minimumCellExtent = (Integer)gen[1];
    h->gen->minimumPoints = (Integer)gen[2];
    h->gen->minimumCellSize = 3.0;
    h->gen->minimumExtentFactor = 1.5;
    h->gen->sortIndices = False;
    h->gen->useGridByCell = gen[3] != 0.0 ? True : False;
    h->gen->cellSize = h->gen->cutoffCellSizeFactor * h->gen->cutoff;
    h->pw = PairwiseInteractionABFS_Allocate();
    h->pw->dampingCutoff = opts[0];
    h->pw->innerCutoff = opts[1];
    h->pw->outerCutoff = opts[2];
    h->atoms = MMAtomContainer_Allocate(n);
    for (int i = 0; i < n; i++) {
        h->atoms->data[i].QACTIVE = True;
        h->atoms->data[i].atomtype = ljtypes[i];
        h->atoms->data[i].ljtype = ljtypes[i];
        h->atoms->data[i].charge = charges[i];
    }
    h->lj = make_lj(ntypes, tA, tB);
    h->lj14 = make_lj(ntypes, tA14 ? tA14 : tA, tB14 ? tB14 : tB);
    h->excl = make_self_list(nexcl, excl);
    h->i14 = make_self_list(n14, p14);
    h->trans = tc;
    if (tc) {
        h->model->checkForInverses = checkForInverses ? True : False;
        h->model->imageExpandFactor = expandFactor;
        h->sp = SymmetryParameters_Allocate();
        h->spg = SymmetryParameterGradients_Allocate();
    }
    h->qc = qc;
    h->fixed = fixed;
    if (qc) {
        /* MMAtomContainer.DeactivateQCAtoms (pMolecule.MMAtomContainer.pyx:116-123): all QC atoms
         * inactive, boundary atoms active again. */
        for (int q = 0; q < qc->natoms; q++) h->atoms->data[qc->data[q].index].QACTIVE = False;
        for (int q = 0; q < qc->natoms; q++)
            if (qc->data[q].QBOUNDARY) h->atoms->data[qc->data[q].index].QACTIVE = True;
        h->qcCharges = Real1DArray_Allocate(qc->natoms, NULL);
        h->qcmmPotentials = Real1DArray_Allocate(qc->natoms, NULL);
        h->model->qcmmCoupling = (QCMMLinkAtomCoupling)coupling;
    }
    h->state = NBModelABFSState_SetUp(h->atoms, qc, fixed, h->excl, h->i14, h->lj, h->lj14,
                                      h->qcCharges, h->qcmmPotentials, NULL, tc,
                                      qc ? (QCMMLinkAtomCoupling)coupling : QCMMLinkAtomCoupling_MM);
    h->crd = Coordinates3_Allocate(n);
    h->grd = Coordinates3_Allocate(n);
    return h;
}

void *ref_create(int n, const double *charges, const int *ljtypes, int ntypes,
                 const double *tA, const double *tB, const double *tA14, const double *tB14,
                 int nexcl, const int *excl, int n14, const int *p14,
                 const double *opts, const double *gen, int nthreads)
{
    return create_common(n, charges, ljtypes, ntypes, tA, tB, tA14, tB14, nexcl, excl, n14, p14, opts, gen,
                         nthreads, NULL, 1, 0, NULL, NULL, 0);
}

/* With crystal symmetry: ntrans fractional transformations (rot 9 per item,
 * row-major; trans 3 per item), as System.DefineSymmetry's transformations. */
void *ref_create_sym(int n, const double *charges, const int *ljtypes, int ntypes,
                     const double *tA, const double *tB, const double *tA14, const double *tB14,
                     int nexcl, const int *excl, int n14, const int *p14,
                     const double *opts, const double *gen, int nthreads,
                     int ntrans, const double *rot, const double *trans, int checkForInverses, int expandFactor)
{
    return create_common(n, charges, ljtypes, ntypes, tA, tB, tA14, tB14, nexcl, excl, n14, p14, opts, gen,
                         nthreads, make_transformations(ntrans, rot, trans), checkForInverses, expandFactor, NULL, NULL,
                         0);
}

/* A QC/MM system (no symmetry).  QC atoms in ascending atom order, as
 * QCAtomContainer_FromAtomContainer (pMolecule.QCAtomContainer.pyx:430-496)
 * makes them: qc_index[q]; boundary[q] != 0 for a link atom replacing MM atom
 * qc_index[q], bonded to QC atom qc_partner[q], with link factor linkfactor[q]
 * and MM partners mmp[mmp_off[q] .. mmp_off[q+1]).  coupling: 0 MM, 1 RC,
 * 2 RD (QCMMLinkAtomCouplingOptions.h).  fixed: nfixed atom indices (may be
 * 0).  The QC/MM electrostatic interaction gets the cubic spline through
 * (spline_x, spline_y) with zero end derivatives, as
 * PairwiseInteractionABFS.MakeElectrostaticSplines makes it
 * (pMolecule.PairwiseInteraction.pyx:182-194). */
void *ref_create_qc(int n, const double *charges, const int *ljtypes, int ntypes,
                    const double *tA, const double *tB, const double *tA14, const double *tB14,
                    int nexcl, const int *excl, int n14, const int *p14,
                    const double *opts, const double *gen, int nthreads,
                    int nqc, const int *qc_index, const int *boundary, const int *qc_partner, const double *linkfactor,
                    int linkRatio, const int *mmp_off, const int *mmp, int coupling, int nfixed, const int *fixed,
                    int nspline, const double *spline_x, const double *spline_y)
{
    QCAtomContainer *qc = QCAtomContainer_Allocate(nqc);
    int nb = 0;
    qc->QLINKRATIO = linkRatio ? True : False;
    for (int q = 0; q < nqc; q++) {
        qc->data[q].index = qc_index[q];
        qc->data[q].atomicNumber = 1;
        if (boundary[q]) {
            int np = mmp_off[q + 1] - mmp_off[q];
            nb++;
            qc->data[q].QBOUNDARY = True;
            qc->data[q].qcpartner = qc_partner[q];
            qc->data[q].linkfactor = linkfactor[q];
            if (np > 0) {
                qc->data[q].nmmpartners = np;
                qc->data[q].mmpartners = Memory_Allocate_Array_Integer(np);
                for (int t = 0; t < np; t++) qc->data[q].mmpartners[t] = mmp[mmp_off[q] + t];
            }
        }
    }
    qc->nboundary = nb;
    Selection *fx = NULL;
    if (nfixed > 0) {
        Integer *tmp = (Integer *)malloc(sizeof(Integer) * (size_t)nfixed);
        for (int k = 0; k < nfixed; k++) tmp[k] = fixed[k];
        fx = Selection_FromIntegerArray(nfixed, tmp);
        free(tmp);
    }
    RefHandle *h = (RefHandle *)create_common(n, charges, ljtypes, ntypes, tA, tB, tA14, tB14, nexcl, excl, n14, p14,
                                              opts, gen, nthreads, NULL, 1, 0, qc, fx, coupling);
    /* The QC/MM and QC/QC interactions of NBModelABFS.CheckPairwiseInteractions
     * (pMolecule.NBModelABFS.pyx:72-97): damping 0, the model's inner and outer cutoffs. */
    for (int k = 0; k < 2; k++) {
        PairwiseInteractionABFS *pw = PairwiseInteractionABFS_Allocate();
        pw->dampingCutoff = 0.0;
        pw->innerCutoff = opts[1];
        pw->outerCutoff = opts[2];
        Real1DArray *X = Real1DArray_Allocate(nspline, NULL), *Y = Real1DArray_Allocate(nspline, NULL);
        for (int i = 0; i < nspline; i++) {
            Real1DArray_Item(X, i) = spline_x[i];
            Real1DArray_Item(Y, i) = spline_y[i];
        }
        CubicSpline_MakeFromReal1DArrays(&pw->electrostaticSpline, &X, &Y, 1, 0.0, 1, 0.0);
        if (k == 0) h->pwqcmm = pw; else h->pwqcqc = pw;
    }
    return h;
}

/* The second derivatives h of the QC/MM spline as the reference builds them
 * (CubicSpline_MakeFromReal1DArrays, dgtsv). */
int ref_qcmm_spline(void *vh, int cap, double *h2)
{
    RefHandle *h = (RefHandle *)vh;
    if (!h->pwqcmm || !h->pwqcmm->electrostaticSpline) return -1;
    CubicSpline *sp = h->pwqcmm->electrostaticSpline;
    for (int i = 0; i < sp->length && i < cap; i++) h2[i] = Real1DArray_Item(sp->h, i);
    return sp->length;
}

static void copy_rows(Coordinates3 *c, double *out, int n)
{
    for (int i = 0; i < n; i++) {
        double x, y, z;
        Coordinates3_GetRow(c, i, x, y, z);
        out[3 * i] = x; out[3 * i + 1] = y; out[3 * i + 2] = z;
    }
}

/* One System.Energy call of a QC/MM system (pMolecule/System.py:297-304):
 * NBModelABFSState_Initialize, NBModelABFS_Update, NBModelABFS_MMMMEnergy +
 * NBModelABFS_QCMMEnergyLJ (NBModelABFS.Energy, pyx:108-122), then
 * NBModelABFS_QCMMPotentials; the QC charges qcq are then set (as the QC model
 * does) and NBModelABFS_QCMMGradients runs.  energies[6]: MM/MM Elect., LJ,
 * 1-4 Elect., 1-4 LJ, QC/MM LJ, QC/MM 1-4 LJ; present[4]: nbmmmm, nbmmmm14,
 * nbqcmmlj, nbqcmmlj14.  potentials[nqc] (atomic units); grads_e: gradients of
 * the energy part, grads_q: those of NBModelABFS_QCMMGradients alone (either
 * may be NULL; grads_q needs grads_e).  Returns 1 if the lists were rebuilt. */
int ref_qcmm_energy(void *vh, const double *coords, const double *qcq, double *grads_e, double *grads_q,
                    double *energies, int *present, double *potentials)
{
    RefHandle *h = (RefHandle *)vh;
    Status status = Status_Continue;
    if (!h->qc) return -1;
    for (int i = 0; i < h->n; i++) {
        Coordinates3_SetRow(h->crd, i, coords[3 * i], coords[3 * i + 1], coords[3 * i + 2]);
        Coordinates3_SetRow(h->grd, i, 0.0, 0.0, 0.0);
    }
    NBModelABFSState_Initialize(h->state, h->crd, NULL, grads_e ? h->grd : NULL, NULL);
    Boolean up = NBModelABFS_Update(h->model, h->gen, h->state, &status);
    if (status != Status_Continue) return -1;
    NBModelABFS_MMMMEnergy(h->model, h->pw, h->state);
    NBModelABFS_QCMMEnergyLJ(h->model, h->pw, h->state);
    energies[0] = h->state->emmel;
    energies[1] = h->state->emmlj;
    energies[2] = h->state->emmel14;
    energies[3] = h->state->emmlj14;
    energies[4] = h->state->eqcmmlj;
    energies[5] = h->state->eqcmmlj14;
    present[0] = (h->state->nbmmmm != NULL);
    present[1] = (h->state->nbmmmm14 != NULL);
    present[2] = (h->state->nbqcmmlj != NULL);
    present[3] = (h->state->nbqcmmlj14 != NULL);
    NBModelABFS_QCMMPotentials(h->model, h->pwqcmm, h->pwqcqc, h->state);
    for (int q = 0; q < h->qc->natoms; q++) potentials[q] = Real1DArray_Item(h->qcmmPotentials, q);
    if (grads_e) {
        copy_rows(h->grd, grads_e, h->n);
        if (grads_q) {
            Coordinates3_Set(h->grd, 0.0);
            for (int q = 0; q < h->qc->natoms; q++) Real1DArray_Item(h->qcCharges, q) = qcq[q];
            NBModelABFS_QCMMGradients(h->model, h->pwqcmm, h->pwqcqc, h->state);
            copy_rows(h->grd, grads_q, h->n);
        }
    }
    return up == True ? 1 : 0;
}

/* QC/MM lists in reference order: which = 0 nbqcmmel (QC atom position, MM
 * site), 1 nbqcmmlj (atom, atom), 2 nbqcmmel14 (QC position, atom),
 * 3 nbqcmmlj14 (atom, atom). */
long ref_qcmm_pairs(void *vh, int which, int *oi, int *oj, long cap)
{
    RefHandle *h = (RefHandle *)vh;
    PairList *pl = which == 0 ? h->state->nbqcmmel : which == 1 ? h->state->nbqcmmlj
                 : which == 2 ? h->state->nbqcmmel14 : h->state->nbqcmmlj14;
    return dump_list(pl, oi, oj, cap);
}

/* The QC atom coordinates and the MM sites of the QC/MM electrostatics (the
 * atoms for MM coupling, the pure MM atoms plus the redistributed sites for
 * RC/RD) with their charges, after the last call.  Returns the site count. */
long ref_qcmm_sites(void *vh, double *qcxyz, double *mmxyz, double *mmq, long cap)
{
    RefHandle *h = (RefHandle *)vh;
    NBModelABFSState *st = h->state;
    long m = st->mmCharges ? st->mmCharges->length : 0;
    if (qcxyz) copy_rows(st->qcCoordinates3, qcxyz, h->qc->natoms);
    for (long k = 0; k < m && k < cap; k++) {
        double x, y, z;
        Coordinates3_GetRow(st->mmCoordinates3, k, x, y, z);
        if (mmxyz) { mmxyz[3 * k] = x; mmxyz[3 * k + 1] = y; mmxyz[3 * k + 2] = z; }
        if (mmq) mmq[k] = Real1DArray_Item(st->mmCharges, k);
    }
    return m;
}

/* One System.Energy-style call.  energies[4] = MM/MM Elect., LJ, 1-4 Elect.,
 * 1-4 LJ; present[2] = main list / 1-4 list non-empty.  grads may be NULL.
 * Returns 1 if the lists were rebuilt, -1 on a library error. */
int ref_energy(void *vh, const double *coords, double *grads, double *energies, int *present)
{
    RefHandle *h = (RefHandle *)vh;
    Status status = Status_Continue;
    for (int i = 0; i < h->n; i++) {
        Coordinates3_SetRow(h->crd, i, coords[3 * i], coords[3 * i + 1], coords[3 * i + 2]);
        if (grads) Coordinates3_SetRow(h->grd, i, 0.0, 0.0, 0.0);
    }
    NBModelABFSState_Initialize(h->state, h->crd, NULL, grads ? h->grd : NULL, NULL);
    Boolean up = NBModelABFS_Update(h->model, h->gen, h->state, &status);
    if (status != Status_Continue) return -1;
    NBModelABFS_MMMMEnergy(h->model, h->pw, h->state);
    energies[0] = h->state->emmel;
    energies[1] = h->state->emmlj;
    energies[2] = h->state->emmel14;
    energies[3] = h->state->emmlj14;
    present[0] = (h->state->nbmmmm != NULL);
    present[1] = (h->state->nbmmmm14 != NULL);
    if (grads) {
        for (int i = 0; i < h->n; i++) {
            double x, y, z;
            Coordinates3_GetRow(h->grd, i, x, y, z);
            grads[3 * i] = x; grads[3 * i + 1] = y; grads[3 * i + 2] = z;
        }
    }
    return up == True ? 1 : 0;
}

/* As ref_energy with symmetry parameters sp6 = a, b, c, alpha, beta, gamma
 * (System.Energy creates fresh SymmetryParameterGradients per call,
 * pMolecule-1.9.0/pMolecule/System.py:284-285).  energies[6] adds MM/MM Image
 * Elect. and LJ; present[3] adds the image list; dEdM[9] row-major. */
int ref_energy_sym(void *vh, const double *coords, const double *sp6, double *grads, double *energies, double *dEdM,
                   int *present)
{
    RefHandle *h = (RefHandle *)vh;
    Status status = Status_Continue;
    if (!h->trans) return -1;
    for (int i = 0; i < h->n; i++) {
        Coordinates3_SetRow(h->crd, i, coords[3 * i], coords[3 * i + 1], coords[3 * i + 2]);
        if (grads) Coordinates3_SetRow(h->grd, i, 0.0, 0.0, 0.0);
    }
    SymmetryParameters_SetCrystalParameters(h->sp, sp6[0], sp6[1], sp6[2], sp6[3], sp6[4], sp6[5]);
    Matrix33_Set(h->spg->dEdM, 0.0);
    NBModelABFSState_Initialize(h->state, h->crd, h->sp, grads ? h->grd : NULL, grads ? h->spg : NULL);
    Boolean up = NBModelABFS_Update(h->model, h->gen, h->state, &status);
    if (status != Status_Continue) return -1;
    NBModelABFS_MMMMEnergy(h->model, h->pw, h->state);
    energies[0] = h->state->emmel;
    energies[1] = h->state->emmlj;
    energies[2] = h->state->emmel14;
    energies[3] = h->state->emmlj14;
    energies[4] = h->state->eimmmel;
    energies[5] = h->state->eimmmlj;
    present[0] = (h->state->nbmmmm != NULL);
    present[1] = (h->state->nbmmmm14 != NULL);
    present[2] = (h->state->inbmmmm != NULL);
    if (grads) {
        for (int i = 0; i < h->n; i++) {
            double x, y, z;
            Coordinates3_GetRow(h->grd, i, x, y, z);
            grads[3 * i] = x; grads[3 * i + 1] = y; grads[3 * i + 2] = z;
        }
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 3; c++) dEdM[3 * r + c] = Matrix33_Item(h->spg->dEdM, r, c);
    }
    return up == True ? 1 : 0;
}

/* Images of the MM/MM image list in list order: abct[4k..4k+3] = a, b, c and
 * the transformation's index; scale[k]; npairs[k].  Returns the image count
 * (writes at most cap). */
long ref_images(void *vh, int *abct, double *scale, long *npairs, long cap)
{
    RefHandle *h = (RefHandle *)vh;
    ImageList *il = h->state->inbmmmm;
    if (il == NULL) return 0;
    long k = 0;
    Image *im;
    List_Iterate_Initialize(il->images);
    while ((im = ImageList_Iterate(il)) != NULL) {
        if (k < cap) {
            int t = -1;
            for (int q = 0; q < h->trans->nitems; q++)
                if (h->trans->items[q] == im->transformation3) t = q;
            abct[4 * k] = im->a; abct[4 * k + 1] = im->b; abct[4 * k + 2] = im->c; abct[4 * k + 3] = t;
            scale[k] = im->scale;
            npairs[k] = PairList_Length(im->pairlist);
        }
        k++;
    }
    return k;
}

static long dump_list(PairList *pl, int *oi, int *oj, long cap);

/* Pairs (i in the primary image, j in image k) of image k of the list. */
long ref_image_pairs(void *vh, long image, int *oi, int *oj, long cap)
{
    RefHandle *h = (RefHandle *)vh;
    ImageList *il = h->state->inbmmmm;
    if (il == NULL) return -1;
    long k = 0;
    Image *im;
    List_Iterate_Initialize(il->images);
    while ((im = ImageList_Iterate(il)) != NULL) {
        if (k == image) return dump_list(im->pairlist, oi, oj, cap);
        k++;
    }
    return -1;
}

static long dump_list(PairList *pl, int *oi, int *oj, long cap)
{
    if (pl == NULL) return 0;
    long m = 0;
    IndexedSelection *rec;
    List_Iterate_Initialize(pl->pairs);
    while ((rec = PairList_Iterate(pl)) != NULL) {
        for (int k = 0; k < rec->nindices; k++) {
            if (m < cap) { oi[m] = rec->index; oj[m] = rec->indices[k]; }
            m++;
        }
    }
    return m;
}

/* Main MM/MM list (which = 0) or 1-4 list (which = 1) in reference order. */
long ref_pairs(void *vh, int which, int *oi, int *oj, long cap)
{
    RefHandle *h = (RefHandle *)vh;
    return dump_list(which == 0 ? h->state->nbmmmm : h->state->nbmmmm14, oi, oj, cap);
}

int ref_use_grid(void *vh) { return ((RefHandle *)vh)->state->useGridSearch == True; }

void ref_destroy(void *vh)
{
    RefHandle *h = (RefHandle *)vh;
    if (!h) return;
    NBModelABFSState_Deallocate(&h->state);
    PairList_Deallocate(&h->excl);
    PairList_Deallocate(&h->i14);
    LJParameterContainer_Deallocate(&h->lj);
    LJParameterContainer_Deallocate(&h->lj14);
    MMAtomContainer_Deallocate(&h->atoms);
    PairwiseInteractionABFS_Deallocate(&h->pw);
    PairListGenerator_Deallocate(&h->gen);
    NBModelABFS_Deallocate(&h->model);
    Coordinates3_Deallocate(&h->crd);
    Coordinates3_Deallocate(&h->grd);
    Transformation3Container_Deallocate(&h->trans);
    SymmetryParameters_Deallocate(&h->sp);
    SymmetryParameterGradients_Deallocate(&h->spg);
    QCAtomContainer_Deallocate(&h->qc);
    Selection_Deallocate(&h->fixed);
    Real1DArray_Deallocate(&h->qcCharges);
    Real1DArray_Deallocate(&h->qcmmPotentials);
    PairwiseInteractionABFS_Deallocate(&h->pwqcmm);
    PairwiseInteractionABFS_Deallocate(&h->pwqcqc);
    free(h);
}
