/*
 * nbabfs.h -- C ABI of the gfx950 NBModelABFS MM/MM energy + gradient library
 * (libnbabfs.so).  Plain pointers and sizes only; no exceptions cross it.
 *
 * It stands in for the reference's C seam of this path (pDynamo 1.9.0; paths
 * relative to the pDynamo tree, pMx = pMolecule-1.9.0/extensions,
 * pCx = pCore-1.9.0/extensions):
 *
 *   NBModelABFSState_SetUp       pMx/cinclude/NBModelABFSState.h:145-156
 *   NBModelABFSState_Initialize  pMx/cinclude/NBModelABFSState.h:138-142
 *   NBModelABFS_Update           pMx/cinclude/NBModelABFS.h:46
 *   NBModelABFS_MMMMEnergy       pMx/cinclude/NBModelABFS.h:42
 *   NBModelABFSState_Deallocate  pMx/cinclude/NBModelABFSState.h:133
 *   PairListGenerator_SelfPairListFromCoordinates3
 *                                pCx/cinclude/PairListGenerator.h:71-79
 *
 * Conventions kept from the reference: energies in kJ/mol, lengths in A,
 * charges in e; gradients (dE/dx) are INCREMENTED, never zeroed; status codes
 * use the reference's Status enum values (pCx/cinclude/Status.h); one state
 * per configuration, not re-entrant; calls are synchronous unless named
 * *_async.  Host arrays are borrowed, device buffers are owned by the state.
 */
#ifndef NBABFS_H
#define NBABFS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status values (pCx/cinclude/Status.h). */
#define NBABFS_STATUS_CONTINUE          16
#define NBABFS_STATUS_INVALID_ARGUMENT  22
#define NBABFS_STATUS_OUT_OF_MEMORY     41
#define NBABFS_STATUS_VALUE_ERROR       47
#define NBABFS_STATUS_LOGIC_ERROR       10

/* Arithmetic of the pair kernel. */
#define NBABFS_MIXED      0   /* tiled kernel: FP64 pair core with 1/r from the FP32
                                 rsqrt seed plus one FP64 Newton step, FP32 force
                                 vectors, 64-bit fixed-point sums, FP64 totals      */
#define NBABFS_PLAIN_FP64 1   /* one thread per atom over a full CSR list, FP64 in
                                 the reference expression order (divide, sqrt)      */
#define NBABFS_TILED_FP64 2   /* the tiled kernel's list, tiles and fixed-point sums
                                 with FP64 pair arithmetic in the reference
                                 expression order (divide, sqrt) and FP64 force
                                 vectors: the mixed path without its FP32 parts
                                 (no image lists)                                    */

/* Upper bound of nbabfs_options.entriesPerItem (per-item j-force staging in LDS). */
#define NBABFS_MAX_ENTRIES_PER_ITEM 128

typedef struct nbabfs_state nbabfs_state;
typedef struct nbabfs_batch nbabfs_batch;

/* Options: the NBModelABFS fields (pMx/cinclude/NBModelABFS.h:22-33, defaults
 * NBModelABFS.c:28-37) and the ABFS default PairListGenerator
 * (pMx/pyrex/pMolecule.NBModelABFS.pyx:59-71), plus kernel tuning knobs. */
typedef struct {
    double dampingCutoff;          /* 0.5  */
    double innerCutoff;            /* 8.0  */
    double outerCutoff;            /* 12.0 */
    double listCutoff;             /* 13.5 */
    double dielectric;             /* 1.0  */
    double electrostaticScale14;   /* 1.0  */
    double cutoffCellSizeFactor;   /* 0.5: cellSize = factor * listCutoff */
    int    minimumCellExtent;      /* 2    */
    int    minimumPoints;          /* 500  */
    int    precision;              /* NBABFS_MIXED, NBABFS_PLAIN_FP64 or NBABFS_TILED_FP64 */
    int    entriesPerItem;         /* cluster-pair entries per work item (0 = 16, at most
                                      NBABFS_MAX_ENTRIES_PER_ITEM) */
    int    wavesPerCU;             /* persistent-kernel waves per CU; 0 = one workgroup per work item */
    int    imageEntriesPerItem;    /* entries per work item of the image lists (0 = 16) */
    int    useGridByCell;          /* 1 (the ABFS default): grid search cell by cell
                                      (PairListGenerator.c:1020-1099); 0: point by cell
                                      (PairListGenerator.c:1155-1218) for the MM/MM list */
} nbabfs_options;

typedef struct {
    long long numberOfCalls;       /* NBModelABFSState.numberOfCalls   */
    long long numberOfUpdates;     /* NBModelABFSState.numberOfUpdates */
    double    numberOfMMMMPairs;   /* summed over updates (StatisticsAccumulate) */
    long long npairs;              /* current main list */
    long long npairs14;            /* 1-4 list */
    int       useGridSearch;
    int       clusters;            /* 8-atom clusters (with padding) */
    long long entries;             /* (super-cluster, j-cluster) entries */
    long long workItems;
    double    forceKernelMs;       /* mean duration of the pair kernel (timing on) */
    long long forceKernelLaunches;
    int       images;              /* MM/MM image list: images with pairs */
    int       reserved2;
    long long imagePairs;          /* pairs of the image list */
    double    numberOfMMMMImageImages;  /* summed over updates (StatisticsAccumulate) */
    double    numberOfMMMMImagePairs;
    long long qcmmElPairs;         /* current QC/MM electrostatic list (QC atom, MM site) */
    long long qcmmLJPairs;         /* current QC/MM LJ list (pure QC atom, MM atom) */
    double    numberOfQCMMElPairs; /* summed over updates (StatisticsAccumulate) */
    double    numberOfQCMMLJPairs;
} nbabfs_stats;

/* Fill *opt with the reference defaults. */
void nbabfs_default_options(nbabfs_options *opt);

/* Library / device information.  Returns NBABFS_STATUS_CONTINUE when a gfx950
 * device is usable; the product path never falls back to the CPU. */
int nbabfs_init(void);
const char *nbabfs_last_error(void);
int nbabfs_device_cu_count(void);

/* NBModelABFSState_SetUp: LJ tables are the packed lower triangles of
 * LJParameterContainer (index i*(i+1)/2+j, j<=i, pMx/cinclude/
 * LJParameterContainer.h:16-23); tableA14/B14 may be NULL (= main tables).
 * exclusions / interactions14 are (i,j) integer pairs in any orientation
 * (PairList_FromIntegerPairArray semantics, pCx/csource/PairList.c:107-189).
 * fixedAtoms (n bytes, 1 = fixed) may be NULL. */
int nbabfs_state_create(const nbabfs_options *opt, int natoms, const double *charges, const int *ljtypes,
                        int ntypes, const double *tableA, const double *tableB, const double *tableA14,
                        const double *tableB14, int nexclusions, const int *exclusions, int ninteractions14,
                        const int *interactions14, const unsigned char *fixedAtoms, nbabfs_state **out);
void nbabfs_state_destroy(nbabfs_state *state);

/* Stream used by every launch of this state (a hipStream_t; NULL = the
 * library's own stream). */
int nbabfs_set_stream(nbabfs_state *state, void *hip_stream);

/* Errors of an energy call.  Where the reference would return NaN or an
 * overflowed value with Status_Continue, these calls return
 * NBABFS_STATUS_VALUE_ERROR: a NaN or infinite coordinate (found by the list
 * build, or by the per-call pass that reads every coordinate), a gradient or
 * energy contribution that is NaN or beyond the +-2^66 range of the
 * fixed-point sums, coordinates spanning more than 1e4 A in one axis, and, at
 * state creation, non-finite charges or LJ parameters.  The energy outputs are
 * still written; the gradients of such a call are not defined.  The error of
 * an asynchronous energy call is returned by nbabfs_energies (and
 * nbabfs_image_terms); a batch returns it from nbabfs_batch_energy. */

/* NBModelABFSState_Initialize + NBModelABFS_Update + NBModelABFS_MMMMEnergy on
 * host arrays (coordinates n*3 FP64 row-major).  grad_inout (n*3) is
 * incremented and may be NULL.  energies[4] = MM/MM Elect., MM/MM LJ,
 * MM/MM 1-4 Elect., MM/MM 1-4 LJ; present[2] says whether the main / 1-4 list
 * is non-empty (the reference drops the labels of an empty list,
 * pMx/pyrex/pMolecule.NBModelABFSState.pyx:41-49).  *rebuilt = list rebuilt. */
int nbabfs_energy_host(nbabfs_state *state, const double *xyz, double *grad_inout, double *energies,
                       int *present, int *rebuilt);

/* Device-resident variants.  d_xyz / d_grad are device pointers (n*3 FP64);
 * d_grad may be NULL.  nbabfs_update: NBModelABFS_Update; when the list is
 * kept it has already queued the pair evaluation for d_xyz on the state's
 * stream, and an nbabfs_mmmm_energy_async on the same array only finishes it
 * (so the array must not change between the two calls; any other array, or a
 * second energy call, is evaluated from scratch).
 * nbabfs_mmmm_energy_async: enqueue the energy/gradient on the state's stream;
 * the four energies land in state-owned device memory readable with
 * nbabfs_energies (which synchronises). */
int nbabfs_update(nbabfs_state *state, const double *d_xyz, int *rebuilt);
int nbabfs_mmmm_energy_async(nbabfs_state *state, const double *d_xyz, double *d_grad);
int nbabfs_energies(nbabfs_state *state, double *energies, int *present);

/* How the energy calls of a state hand over its gradient:
 * NBABFS_GRADIENT_ADD (the default, and the reference's convention: the model
 * increments the caller's array, NBModelABFSState_AccumulateGradients,
 * pMx/csource/NBModelABFSState.c:19-38) or NBABFS_GRADIENT_ASSIGN (every
 * atom's row is overwritten with this model's gradient).  No counterpart in
 * the reference: for callers whose first gradient term is this model, so the
 * zero fill they would issue before it is folded into the model's last pass.
 * Applies to nbabfs_energy_host, nbabfs_mmmm_energy_async and to batches
 * (per state).  Synchronises the state's stream. */
#define NBABFS_GRADIENT_ADD 0
#define NBABFS_GRADIENT_ASSIGN 1
int nbabfs_set_gradient_mode(nbabfs_state *state, int mode);

/* Virial of the last energy call's gradient contribution, row-major 3x3:
 * W[3a+b] = sum_i (x_i - o)_a * dE/dx_i,b over the atoms (main list, 1-4
 * list and, for the plain path, its rows), o being the state's cluster-grid
 * origin; for the non-periodic pair model this equals sum over pairs of
 * r_ij,a * g_ij,b.  No counterpart in the reference, which computes no virial
 * for non-periodic systems; computed on request from the call's fixed-point
 * gradient sums, so it is valid until the state's next update or energy call;
 * synchronises the state's stream. */
int nbabfs_virial(nbabfs_state *state, double *virial);

/* Work-item autotuner (no counterpart in the reference, which has no work
 * items): updates the list for d_xyz (as nbabfs_update), then times the pair
 * kernel for each candidate (entriesPerItem, wavesPerCU) on that list and
 * keeps the fastest in the state's options.  The pair list is not changed;
 * only its cut into work items.  Outputs may be NULL. */
int nbabfs_autotune(nbabfs_state *state, const double *d_xyz, int *best_entries_per_item, int *best_waves_per_cu,
                    double *best_kernel_ms);

/* ---- Crystal symmetry: the MM/MM image list (pMx/csource/NBModelABFS.c:
 * 625-684 CheckForImageUpdate, 753-1045 GenerateImageLists, 1161-1290
 * MMMMImageEnergy). */

/* The transformations of NBModelABFSState_SetUp (pMx/cinclude/NBModelABFSState.h:
 * 145-156, argument Transformation3Container): ntransformations fractional
 * rotations (9 each, row-major) and translations (3 each); their identity and
 * rotational inverses are found as Transformation3Container_FindIdentity /
 * _FindInverses do (pCx/csource/Transformation3Container.c:52-141).
 * checkForInverses and imageExpandFactor are the NBModelABFS options
 * (pMx/cinclude/NBModelABFS.h:22-33).  Call before the first update; the image
 * terms are evaluated by the NBABFS_MIXED path. */
int nbabfs_set_transformations(nbabfs_state *state, int ntransformations, const double *rotations,
                               const double *translations, int checkForInverses, int imageExpandFactor);

/* The symmetry parameters of NBModelABFSState_Initialize (pMx/cinclude/
 * NBModelABFSState.h:138-142): crystal = a, b, c (A), alpha, beta, gamma
 * (degrees), from which M is formed as SymmetryParameters_MakeM does.  Applies
 * to the following updates and energy calls. */
int nbabfs_set_symmetry_parameters(nbabfs_state *state, const double *crystal);

/* The last energy call's image terms: energies[2] = MM/MM Image Elect. and LJ,
 * *present = the image list is non-empty, and dE/dM (row-major 3x3,
 * INCREMENTED, may be NULL) as MMMMImageEnergy adds it into
 * SymmetryParameterGradients.dEdM.  Synchronises the state's stream. */
int nbabfs_image_terms(nbabfs_state *state, double *energies, double *dEdM, int *present);

/* The image list (ImageList of inbmmmm, pMx/cinclude/ImageList.h): per image
 * abct[4k..] = a, b, c, transformation index, scale[k], npairs[k] (outputs may
 * be NULL, at most cap written).  Returns the number of images. */
int nbabfs_images(nbabfs_state *state, int cap, int *abct, double *scale, long long *npairs);

/* Pairs of image k as sorted keys i*n + j (i in the primary cell, j in the
 * image).  Returns the pair count (keys may be NULL to query it). */
long long nbabfs_export_image_pairs(nbabfs_state *state, int image, long long *keys, long long cap);

/* QC/MM systems ------------------------------------------------------------------------------
 *
 * QC atoms (NBModelABFSState_SetUp's QCAtomContainer argument and the model's
 * qcmmCoupling, pMolecule-1.9.0/extensions/cinclude/NBModelABFSState.h:145-156,
 * QCAtomContainer.h:24-60): nqc atoms in ascending atom order index[q]; a
 * boundary (link) atom has boundary[q] != 0, its QC partner qcpartner[q], link
 * factor linkfactor[q] (a distance, or a ratio when linkRatio != 0) and MM
 * partners mmpartners[mmpartnerOffsets[q] .. mmpartnerOffsets[q+1]) (arrays
 * of the other atoms are ignored; qcpartner, linkfactor and the offsets may
 * be NULL when no atom is a boundary atom).  coupling: NBABFS_QCMM_MM, _RC or
 * _RD (QCMMLinkAtomCouplingOptions.h).  Call once, before the first update.
 * Pure QC atoms leave the MM/MM lists; the QC/MM lists are built with them. */
#define NBABFS_QCMM_MM 0
#define NBABFS_QCMM_RC 1
#define NBABFS_QCMM_RD 2
int nbabfs_set_qc_atoms(nbabfs_state *state, int nqc, const int *index, const int *boundary, const int *qcpartner,
                        const double *linkfactor, int linkRatio, const int *mmpartnerOffsets, const int *mmpartners,
                        int coupling);

/* The QC/MM electrostatic interaction: the cubic spline through (x[k] = r^2,
 * y[k]) with zero first derivatives at both ends, as
 * CubicSpline_MakeFromReal1DArrays builds PairwiseInteractionABFS's
 * electrostaticSpline (pCore-1.9.0/extensions/csource/CubicSpline.c:306-405,
 * pMolecule.PairwiseInteraction.pyx:182-194), in atomic units. */
int nbabfs_set_qcmm_spline(nbabfs_state *state, int npoints, const double *x, const double *y);

/* After an energy call: QC/MM LJ and QC/MM 1-4 LJ (NBModelABFS_QCMMEnergyLJ,
 * NBModelABFS.h:43, NBModelABFSState eqcmmlj / eqcmmlj14) and whether each list
 * has pairs.  Their gradients are part of the energy call's. */
int nbabfs_qcmm_terms(nbabfs_state *state, double *energies, int *present);

/* NBModelABFS_QCMMPotentials (NBModelABFS.h:45) on the coordinates of the last
 * update: potentials[nqc] (atomic units) are incremented. */
int nbabfs_qcmm_potentials(nbabfs_state *state, double *potentials);

/* NBModelABFS_QCMMGradients (NBModelABFS.h:44) for the QC charges
 * qcCharges[nqc] on the coordinates of the last update: grad_inout (n x 3,
 * host, kJ/mol/A) is incremented. */
int nbabfs_qcmm_gradients_host(nbabfs_state *state, const double *qcCharges, double *grad_inout);

/* Device-array variants of the two calls above (the same NBModelABFS.h:44-45
 * entry points, for a QC program whose charges and potentials live in GPU
 * memory): enqueued on the state's stream, no host copy and no
 * synchronisation; d_potentials[nqc] and d_grad (n x 3) are incremented with
 * bit for bit the values the host calls add. */
int nbabfs_qcmm_potentials_async(nbabfs_state *state, double *d_potentials);
int nbabfs_qcmm_gradients_async(nbabfs_state *state, const double *d_qcCharges, double *d_grad);

/* The QC/MM lists of the last build as sorted keys row * 2^32 + column.
 * which: 0 electrostatic (row = QC atom position, column = MM site: the atom
 * for MM coupling, else the site of nbabfs_qcmm_sites), 1 LJ (atom, atom),
 * 2 electrostatic 1-4 (QC position, atom), 3 LJ 1-4 (atom, atom).  Returns the
 * pair count (keys may be NULL). */
long long nbabfs_export_qcmm_pairs(nbabfs_state *state, int which, long long *keys, long long cap);

/* QC atom coordinates (nqc x 3) and MM sites (coordinates, charges) of the
 * last update's coordinates (any output may be NULL).  Returns the site count. */
long long nbabfs_qcmm_sites(nbabfs_state *state, double *qcxyz, double *sitexyz, double *siteq, long long cap);

/* Statistics (NBModelABFSState.pyx:68-97) and kernel timing with HIP events on
 * the launch stream (enable != 0 turns it on). */
int nbabfs_stats_get(nbabfs_state *state, nbabfs_stats *out);
int nbabfs_set_timing(nbabfs_state *state, int enable);

/* The current main (which=0) or 1-4 (which=1) list as canonical int64 keys
 * max(i,j)*n + min(i,j), sorted ascending.  Returns the pair count; writes at
 * most cap keys (keys may be NULL to query the count). */
long long nbabfs_export_pairs(nbabfs_state *state, int which, long long *keys, long long cap);

/* Device memory helpers, so that callers need no second HIP runtime (a
 * framework that bundles its own libamdhip64 cannot share streams with this
 * library).  No counterpart in the reference, whose arrays live in host
 * memory.  Copies and the zero fill are ordered on the state's stream when a
 * state is given; with NULL they first wait for all work on the device (every
 * state's stream), then run synchronously. */
int nbabfs_device_alloc(size_t bytes, void **out);
int nbabfs_device_free(void *ptr);
int nbabfs_memcpy_htod(nbabfs_state *state, void *dst, const void *src, size_t bytes);
int nbabfs_memcpy_dtoh(nbabfs_state *state, void *dst, const void *src, size_t bytes);
int nbabfs_zero_async(nbabfs_state *state, void *dst, size_t bytes);
int nbabfs_synchronize(nbabfs_state *state);

/* Batches of independent systems (reaction-path images, replicas): one
 * persistent launch over the union of their work items; replays a captured
 * hipGraph while no list is rebuilt.  states stay owned by the caller. */
int nbabfs_batch_create(nbabfs_state *const *states, int nsystems, int use_graph, nbabfs_batch **out);
void nbabfs_batch_destroy(nbabfs_batch *batch);
/* d_xyz[k], d_grad[k] device pointers per system (d_grad entries may be
 * NULL); energies (4*nsystems) and rebuilt (nsystems, may be NULL) on host.
 * The update test of every system and the batched passes are queued together
 * (the passes do nothing when the test asks for a rebuild, and run again after
 * it), so a call with no rebuild waits for the GPU once. */
int nbabfs_batch_energy(nbabfs_batch *batch, const double *const *d_xyz, double *const *d_grad, double *energies,
                        int *rebuilt);
/* Bookkeeping of a batch (no counterpart in the reference): hipGraph captures,
 * batched passes whose results a call returned (a pass queued behind the
 * update test and abandoned because some system had to be rebuilt is not
 * counted) and re-layouts (concatenations of the systems' work items, after a
 * rebuild of any system or a change of its arrays) so far, and the mean
 * duration (ms) of the batched pair launch over the timed launches.
 * Outputs may be NULL. */
int nbabfs_batch_info(nbabfs_batch *batch, long long *graph_captures, long long *launches, long long *layouts,
                      double *force_kernel_ms, long long *timed_launches);
/* Batch work-item autotuner (no counterpart in the reference, which has no
 * work items): after one batched call for d_xyz (lists built or updated as by
 * nbabfs_batch_energy), times the batched pair launch for each candidate
 * entriesPerItem (every state re-cut alike, one work item per wavefront) and
 * keeps the fastest in every state's options.  No gradient array is written;
 * the pair lists are unchanged, only their cut into work items.  Outputs may
 * be NULL. */
int nbabfs_batch_autotune(nbabfs_batch *batch, const double *const *d_xyz, int *best_entries_per_item,
                          double *best_kernel_ms);
/* Timing of the batched pair launch (no counterpart in the reference): while
 * enable != 0 the batch's calls enqueue their passes directly instead of
 * replaying the graph, with the runtime's dispatch start/stop events around
 * the pair kernel (the plain path: from the first system's pair kernel to the
 * last's); enabling resets the mean. */
int nbabfs_batch_set_timing(nbabfs_batch *batch, int enable);

#ifdef __cplusplus
}
#endif
#endif
