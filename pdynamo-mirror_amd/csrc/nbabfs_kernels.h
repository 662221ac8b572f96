// nbabfs_kernels.h -- parameter blocks and launch wrappers shared by the host
// driver (nbabfs_host.hip) and the kernel files.
#pragma once

#include "nbabfs_internal.h"

// Reference cell grid (Coordinates3_MakeGrid, pCore-1.9.0/extensions/csource/
// Coordinates3.c:1111-1158) plus the cluster-column grid used for sorting.
struct BuildGrid {
    double lower[3];
    double h;
    int bins[3];
    int use_grid;
    double corg[3];   // cluster-column grid origin (= bounding-box minimum)
    double cw, dz;    // column width and z-bin height
    int ncx, ncy, nzb;
};

// Grid conforming to the primary grid that the reference makes for one image's points
// (Coordinates3_MakeConformingGrid, pCore-1.9.0/extensions/csource/Coordinates3.c:998-1072): lower corner,
// lower cell indices on the primary grid and number of cells per axis.
struct ImageGrid {
    double lower[3];
    int lowN[3];
    int bins[3];
};

struct SearchParams {
    double cutoff2;                       // listCutoff^2 (FP64 test)
    const unsigned char *stencil;         // (2u+1)^3 flags
    int u;
    int use_grid;
    const int *xptr, *xidx;               // exclusions CSR
    const unsigned char *free_flags;      // or null (no fixed atoms)
    float reach;                          // cull distance for FP32 boxes
    float cw;
    int ncx, ncy;
    int cap;                              // candidate slots per super-cluster
    // Image lists (cross search; blockIdx.y = image): the image's exact
    // coordinates (n*3 FP64 each), its cluster boxes (ncl each, cluster frame)
    // and its bounding box (6 floats, cluster frame, padded).
    const double *ixyz;
    const float4 *iblo, *ibhi;
    const float *ibox;
    // Grid method: each i atom's search cell on the primary grid with its on-grid flag in w (k_image_cells),
    // and each image's conforming grid (k_image_grid).
    const int4 *icell;
    const ImageGrid *igrid;
    const int *soff;                      // scratch offset and capacity of each (image, super-cluster)
    const int *scap;
    int n, ncl;
    // Self lists: slot-ordered search records (k_slot_records).
    const double4 *srec;
    const int4 *scell;
    // Self lists: the counting sort's offsets per (column, z-bin) key, the z-bin height and count, so that each
    // column is scanned only over the super-clusters whose z-bins come within reach.
    const int *zoff;
    float dz;
    int nzb, ncol;
    // Point/cell grid search (useGridByCell false): the reference grid's lower corner and cell size.
    int pointcell;
    double glower[3];
    double gh;
};

// Search record flag bits (the high word of srec[s].w; the low word is the atom, -1 for an empty slot).
#define NB_SF_FREE 1        // free atom (every atom when no atom is fixed)
#define NB_SF_QC 2          // outside the MM selection (pure QC atom)
#define NB_SF_OFFGRID 4     // grid method: the atom lies outside the reference grid
#define NB_SF_EMPTY 8       // padding slot

// One step of a transformation's (a, b, c) sequence in GenerateImageLists:
// the coordinates are displaced by d, stored as image `out` when that image is
// searched (out >= 0), and displaced back.
struct ImageStep {
    double d[3];
    int out;
    int pad;
};

// List construction (nbabfs_build.hip).
void launch_bbox(const double *xyz, int n, double *part, int nparts, double *out, hipStream_t s);
void launch_bin(const double *xyz, int n, const BuildGrid &g, int4 *cellv, int *key, int *count, hipStream_t s);
size_t scan_tmp_size(int m);
void launch_scan(const int *in, int *out, int m, int *tmp, hipStream_t s);
void launch_fill_int(int *p, size_t m, int v, hipStream_t s);
void launch_sort_slots(const int *key, int n, int nkeys, int ncol, int nzb, int *off, int *cursor, int *unsorted,
                       int *sorted, int *padded, hipStream_t s);
void launch_slots(const int *sorted, const int *key, int n, int nzb, const int *off, const int *colstart,
                  const int *padded, int ncol, const double *xyz, int nsc, int *slot_atom, int *atom_slot, int2 *colcl,
                  hipStream_t s);
void launch_cluster_geom(const double *xyz, const int *slot_atom, int ncl, const double *origin, float4 *blo,
                         float4 *bhi, float4 *centre, hipStream_t s);
// Slot-ordered search records: {x, y, z, atom | flags << 32} and the atom's reference cell.
void launch_slot_records(const double *xyz, const int *slot_atom, int n_pad, const int4 *cellv,
                         const unsigned char *free_flags, int use_grid, int pointcell, int u, double4 *srec,
                         int4 *scell, hipStream_t s);
// Self-list search (the records of P.srec / P.scell must describe the current coordinates and cells).
void launch_search(const SearchParams &P, int nsc, const double *xyz, const int *slot_atom, const float4 *blo,
                   const float4 *bhi, const int2 *colcl, Entry *scratch, int *ncand, int *counts, int *overflow,
                   hipStream_t s);
void launch_cut_items(const int *counts, const int *eoff, const int *ioff, int nblocks, int nsc, int per_item,
                      int sys, WorkItem *items, hipStream_t s, bool images = false);
// Image lists: R'x + T' per transformation (12 doubles each), the stepped image
// coordinates, the on-grid flags of the current coordinates, the cross search.
void launch_image_transform(const double *xyz, int n, const double *rt, int ntrans, double *out, hipStream_t s);
void launch_image_coords(const double *tcoords, int n, int ntrans, const int *step_off, const ImageStep *steps,
                         double *out, hipStream_t s);
void launch_image_cells(const double *xyz, int n, const BuildGrid &g, int pointcell, int u,
                        const unsigned char *free_flags, int4 *icell, hipStream_t s);
void launch_image_grid(const double *ixyz, int n, int nimg, const BuildGrid &g, ImageGrid *igrid, hipStream_t s);
void launch_search_images(const SearchParams &P, int nsc, int nimg, const double *xyz, const int *slot_atom,
                          const float4 *blo, const float4 *bhi, Entry *scratch, int *ncand, int *counts, int *overflow,
                          hipStream_t s);
// Upper bound of each (image, super-cluster) block's candidates (its scratch capacity).
void launch_image_bounds(const SearchParams &P, int nsc, int nimg, const float4 *blo, const float4 *bhi, int *bound,
                         hipStream_t s);
void launch_add_int(int *p, int m, int v, hipStream_t s);
void launch_export_image_pairs(const Entry *entries, int nent, const int *slot_atom, long long n, int img,
                               unsigned long long *cursor, long long cap, long long *keys, hipStream_t s);
void launch_order_items(const WorkItem *items, int nitems, const Entry *entries, int maxcost, int interleave, int *cost,
                        int *hist, int *cursor, int *scantmp, WorkItem *out, hipStream_t s);
void launch_items_per_sc(const int *counts, int nsc, int per_item, int *nitems, hipStream_t s);
// Entry and item offsets (exclusive scans of counts and of ceil(counts / per_item)) and the totals
// {entries, items, *need} into out, in one launch; false (nothing launched) when nsc is beyond its single workgroup.
bool launch_list_offsets(const int *counts, int nsc, int per_item, int *eoff, int *nits, int *ioff, const int *need,
                         int *out, hipStream_t s);
// nblocks compaction blocks (super-cluster = block % nsc); the pairs of each
// block are stored into npairs_block and, when npairs is given, their sum into *npairs.
// soff (or null: block * cap) gives each block's scratch offset.
void launch_compact(const Entry *scratch, int cap, const int *soff, const int *ncand, const int *counts,
                    const int *eoff, const int *ioff, int nblocks, int nsc, int per_item, int sys, Entry *entries,
                    WorkItem *items, unsigned long long *npairs, unsigned long long *npairs_block, hipStream_t s);
void launch_export_pairs(const Entry *entries, int nent, const int *slot_atom, long long n,
                         unsigned long long *cursor, long long cap, long long *keys, hipStream_t s);
void launch_csr_degree(const Entry *entries, int nent, const int *slot_atom, int *deg, hipStream_t s);
void launch_csr_sort_rows(const int *rowptr, int n, int *cols, hipStream_t s);
void launch_csr_fill(const Entry *entries, int nent, const int *slot_atom, int *cursor, int *cols, hipStream_t s);

// Per-call kernels (nbabfs_force.hip).  Every launcher takes the batch's
// descriptor array (device) and the number of systems.
struct QueueParams {
    int *heads;            // 8 counters 128 bytes apart (heads[32*q]), reset by launch_prepare;
                           // null = one work item per workgroup
    const int *starts;     // 9 item offsets: queue q owns [starts[q], starts[q+1])
};

// Optional list-update test fused into the prepare pass (single system).
struct PrepareCheck {
    const double *xyz;                 // coordinates to prepare from (null: the descriptor's)
    const double *ref;                 // coordinates of the last list build (null: no test)
    const unsigned char *free_flags;   // null: every atom is free
    double bufsq;                      // (half the list buffer)^2
    int *flag;                         // host-mapped: receives epoch when some free atom moved too far
    int *dflag;                        // device copy of the same flag (read by ForceGuard)
    int epoch;
    unsigned long long *counter;       // device: blocks finished << 32 | blocks with a moved atom (reset by the last)
    int *done;                         // host-mapped: receives epoch once every block has finished
};

// A force launch queued before the update decision is known: every
// workgroup returns at once when *flag == epoch (the list must be rebuilt).
// With epoch_ptr, the epoch is read from device memory (*epoch_ptr), so a
// captured graph compares against the epoch of the check that ran before it.
struct ForceGuard {
    const int *flag;
    int epoch;
    const int *epoch_ptr;
};
void launch_prepare(const SysDesc *d_desc, int nsys, int max_npad, int tiled, int *heads, const PrepareCheck &chk,
                    hipStream_t s);
// nblocks: work items (one per wave) or, with q.heads, the persistent waves.
// images: the launch over image-list items (one workgroup per item).  The pair
// launchers take optional timing events (start, stop: recorded by the runtime
// at the dispatch's own begin and end; null = not timed).
// full_fp64: the NBABFS_TILED_FP64 instance (FP64 pair arithmetic and force vectors; not for image items).
// sys1 >= 0: every item belongs to system sys1 (a single-state launch), so the kernel reads its descriptor without
// waiting for the items.
void launch_force_tiled(const SysDesc *d_desc, const WorkItem *items, QueueParams q, const AbfsFactors &F,
                        int nblocks, int max_ntypes, int max_count, ForceGuard g, hipStream_t s, bool images = false,
                        hipEvent_t ev_start = nullptr, hipEvent_t ev_stop = nullptr, int n14_blocks = 0,
                        bool full_fp64 = false, int sys1 = -1);
// The 1-4 list in blocks of NB_P14_BLOCK pairs (k_pairs14, or extra waves of a single-state pair launch).
#define NB_P14_BLOCK 64
// Waves per workgroup of the one-item-per-wave pair launches; a batch pads each system's items to a multiple of it
// (items of one workgroup share a system and its LJ row table).
#define NB_TILED_WPG 4
// The plain FP64 CSR pass of nsys systems in one launch (blockIdx.y = system; rows from the descriptors).
void launch_force_plain(const SysDesc *d_desc, int nsys, int max_n, const AbfsFactors &F, ForceGuard g, hipStream_t s,
                        hipEvent_t ev_start = nullptr, hipEvent_t ev_stop = nullptr);
// Plain FP64 image pass over the image lists' work items.
void launch_force_image_fp64(const SysDesc *d_desc, const WorkItem *items, int nitems, const AbfsFactors &F,
                             ForceGuard g, hipStream_t s, hipEvent_t ev_start = nullptr, hipEvent_t ev_stop = nullptr);
void launch_pairs14(const SysDesc *d_desc, int nsys, int max_n14_blocks, const AbfsFactors &F, hipStream_t s,
                    ForceGuard g = ForceGuard{nullptr, 0, nullptr});
// max_npad: the largest padded slot count of the systems (one thread per slot).
void launch_finalize(const SysDesc *d_desc, int nsys, int max_npad, int plain, hipStream_t s,
                     ForceGuard g = ForceGuard{nullptr, 0, nullptr});
// Virial of the single system's last call (NB_VACC_WORDS in D.vacc, zeroed by the caller).
void launch_virial(const SysDesc *d_desc, int n, int plain, hipStream_t s);

// QC/MM electrostatics (nbabfs_qcmm.hip).  QC atoms in ascending atom order;
// qc_partner[q] >= 0 marks a link atom (the QC partner of boundary atom
// qc_index[q]) with link factor qc_lf[q].  MM sites: the atoms (site_a null)
// or site k at atom site_a[k], or the midpoint of site_a[k] and site_p[k].
struct QcmmGeom {
    const double *xyz;
    int n;
    const int *qc_index, *qc_partner;
    const double *qc_lf;
    int nqc, linkRatio;
    const int *site_a, *site_p;
    int nsites;
};

// Grid search of one QC/MM cross list (MakeCrossPairListFromCoordinates3, pCore-1.9.0/extensions/csource/
// PairListGenerator.c:611-940): the rows' grid (lower1, bins1) and the columns' grid conforming to it (lower2,
// its lower cell indices lowN on the rows' grid, bins2), made on the device by k_qcmm_grids.
struct QcmmGrid {
    double lower1[3], lower2[3];
    int bins1[3], bins2[3], lowN[3];
    int bad;                            // the rows' grid has an axis with no cells
};

struct CrossListParams {
    QcmmGeom G;
    int rows_are_qc;                    // rows: QC atoms (link coordinates), else atoms row_atom[r]
    const int *row_atom;
    int nrows;
    const int *ex_ptr, *ex_idx;         // sorted excluded columns per row (QC rows) or per atom, or null
    const unsigned char *flags;         // atom flags (bit 0 free, bit 1 pure QC) of atom columns, or null
    int or_test;                        // apply the free-atom OR test
    int col_atoms;                      // columns: atoms, else MM sites
    // Grid method (grid non-null): stencil flags over (2u+1)^3 offsets, cell size, point/cell search.
    const QcmmGrid *grid;
    const int4 *colcell;                // each column's cell on the columns' grid, w = 1 on it (launch_qcmm_colcells)
    const unsigned char *stencil;
    int u, pointcell;
    double h;
    double cutoff2;
    int *count;                         // count pass: accepted columns per (row, column chunk), row-major
    const int *off;                     // fill pass: first list position of each (row, column chunk)
    int *idx;
};

// Columns per workgroup of the cross-list passes: each row is searched by ceil(columns / NB_XL_CHUNK) workgroups.
#define NB_XL_CHUNK 2048
inline int cross_list_chunks(const CrossListParams &P)
{
    const int ncol = P.col_atoms ? P.G.n : P.G.nsites;
    return ncol > 0 ? (ncol + NB_XL_CHUNK - 1) / NB_XL_CHUNK : 1;
}

struct SplineDev {                      // CubicSpline: abscissae x (= r^2), ordinates y, second derivatives h
    const double *x, *y, *h;
    int n;
};

struct QcmmRows {                       // main and 1-4 lists of the QC atoms (CSR; the 1-4 pointers may be null)
    const int *ptr, *idx, *ptr14, *idx14;
};

void launch_cross_list(const CrossListParams &P, int pass, hipStream_t s);
// Row pointers ptr[r] = off[r * nchunk] (r = 0 .. nrows) of the exclusive scan of the per-(row, chunk) counts.
void launch_row_pointers(const int *off, int nrows, int nchunk, int *ptr, hipStream_t s);
// The grids of the two QC/MM cross lists into out[0] (electrostatic: the QC coordinates' grid and the grid of the MM
// sites, or of the MM selection's atoms with MM coupling, conforming to it) and out[1] (LJ: the grid of all atoms,
// lower/bins of the main list, for rows and columns).
// part: scratch of 12 * NB_QGRID_PARTS doubles (the per-workgroup boxes).
#define NB_QGRID_PARTS 64
void launch_qcmm_grids(const QcmmGeom &G, const unsigned char *flags, const double *lower, const int *bins, double h,
                       double *part, QcmmGrid *out, hipStream_t s);
// The cells of the electrostatic list's columns (MM sites, or atoms with MM coupling) on their grid (grid->lower2).
void launch_qcmm_colcells(const QcmmGeom &G, const QcmmGrid *grid, double h, int4 *out, hipStream_t s);
// chunks workgroups per QC atom; part: 2 per (QC atom, chunk), main and 1-4 sums.
void launch_qcmm_potentials(const QcmmGeom &G, const SplineDev &S, const QcmmRows &L, const double *site_q,
                            double escale, double escale14, double r2off, int chunks, double *part, hipStream_t s);
// gacc: six int64 arrays of n (x, y, z low words, then high words), fixed point as facc.
void launch_qcmm_gradients(const QcmmGeom &G, const SplineDev &S, const QcmmRows &L, const double *site_q,
                           const double *qcq, double scale, double scale14, double r2off, int chunks, long long *gacc,
                           hipStream_t s);
void launch_qcmm_grad_out(const long long *gacc, int n, double *grad, bool add, hipStream_t s);
void launch_qcmm_pot_sum(const double *part, int nqc, int chunks, double *pot, hipStream_t s);
void launch_qcmm_coords(const QcmmGeom &G, double *qcxyz, double *sitexyz, hipStream_t s);
// QC/MM LJ (nbabfs_force.hip): the QC/MM LJ and 1-4 LJ lists of the descriptor's pure QC atoms,
// NB_QCMM_CHUNKS workgroups per row.
#define NB_QCMM_CHUNKS 4
void launch_qclj(const SysDesc *d_desc, int nrows, const AbfsFactors &F, hipStream_t s);
