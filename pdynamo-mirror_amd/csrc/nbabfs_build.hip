// nbabfs_build.hip -- pair-list construction on gfx950.
//
// Replaces PairListGenerator_SelfPairListFromCoordinates3 (CellCell and
// Direct, pCore-1.9.0/extensions/csource/PairListGenerator.c:946-1150),
// RegularGridOccupancy_Fill (pCore-1.9.0/extensions/csource/
// RegularGridOccupancy.c:77-147) and, for the periodic image lists,
// PairListGenerator_CrossPairListFromDoubleCoordinates3 (PairListGenerator.c:
// 414-450,611-860) as GenerateImageLists calls it (pMolecule-1.9.0/extensions/
// csource/NBModelABFS.c:753-1045).  The pair SET is the reference's exactly:
// every candidate that survives an FP32 bounding-box cull is re-tested with
// the reference predicate in FP64 (same cell grid, same stencil, same
// x,y,z-ordered sum of squares, no FMA contraction).  Only the storage
// differs: 8x8 cluster tiles with 64-bit masks instead of linked records.
#include "nbabfs_internal.h"
#include "nbabfs_kernels.h"

#include <float.h>

// ------------------------------------------------------------------------------------------------
// Bounding box (Coordinates3_EnclosingOrthorhombicBox, pCore-1.9.0/extensions/csource/
// Coordinates3.c:498-605).  Min/max are order independent, so a tree is exact.
// ------------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_bbox_partial(const double *__restrict__ xyz, int n, double *__restrict__ part)
{
    __shared__ double s[6][256];
    double mn[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, mx[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        for (int d = 0; d < 3; d++) {
            double v = xyz[3 * (size_t)i + d];
            // fmin/fmax skip NaN: a non-finite coordinate makes the box infinite
            // instead, which the host rejects before any cell index is formed.
            if (!isfinite(v)) {
                mn[d] = -INFINITY;
                mx[d] = INFINITY;
            }
            mn[d] = fmin(mn[d], v);
            mx[d] = fmax(mx[d], v);
        }
    }
    for (int d = 0; d < 3; d++) { s[d][threadIdx.x] = mn[d]; s[d + 3][threadIdx.x] = mx[d]; }
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o)
            for (int d = 0; d < 3; d++) {
                s[d][threadIdx.x] = fmin(s[d][threadIdx.x], s[d][threadIdx.x + o]);
                s[d + 3][threadIdx.x] = fmax(s[d + 3][threadIdx.x], s[d + 3][threadIdx.x + o]);
            }
        __syncthreads();
    }
    if (threadIdx.x < 6) part[6 * blockIdx.x + threadIdx.x] = s[threadIdx.x][0];
}

__global__ __launch_bounds__(256) void k_bbox_final(const double *__restrict__ part, int nblocks, double *__restrict__ out)
{
    __shared__ double s[6][256];
    double v[6] = {DBL_MAX, DBL_MAX, DBL_MAX, -DBL_MAX, -DBL_MAX, -DBL_MAX};
    for (int b = threadIdx.x; b < nblocks; b += 256)
        for (int d = 0; d < 6; d++) v[d] = (d < 3) ? fmin(v[d], part[6 * b + d]) : fmax(v[d], part[6 * b + d]);
    for (int d = 0; d < 6; d++) s[d][threadIdx.x] = v[d];
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o)
            for (int d = 0; d < 6; d++)
                s[d][threadIdx.x] = (d < 3) ? fmin(s[d][threadIdx.x], s[d][threadIdx.x + o])
                                            : fmax(s[d][threadIdx.x], s[d][threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x < 6) out[threadIdx.x] = s[threadIdx.x][0];
}

// ------------------------------------------------------------------------------------------------
// Binning.  Reference cell of each atom (RegularGridDimension_FindBinIndex,
// pCore-1.9.0/extensions/cinclude/RegularGrid.h:73-85: floor((x-lower)/h),
// off-grid atoms dropped) and the cluster-column sort key.
// ------------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_bin(const double *__restrict__ xyz, int n, BuildGrid g,
                                             int4 *__restrict__ cellv, int *__restrict__ key, int *__restrict__ count)
{
#pragma clang fp contract(off)
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    double x[3] = {xyz[3 * (size_t)i], xyz[3 * (size_t)i + 1], xyz[3 * (size_t)i + 2]};
    int4 cv = make_int4(0, 0, 0, 1);
    if (g.use_grid) {
        int k[3], on = 1;
        for (int d = 0; d < 3; d++) {
            double fr = (x[d] - g.lower[d]) / g.h;
            k[d] = (int)floor(fr);
            if (k[d] < 0 || k[d] >= g.bins[d]) on = 0;
        }
        cv = make_int4(k[0], k[1], k[2], on);
    }
    cellv[i] = cv;
    int cx = (int)floor((x[0] - g.corg[0]) / g.cw);
    int cy = (int)floor((x[1] - g.corg[1]) / g.cw);
    int cz = (int)floor((x[2] - g.corg[2]) / g.dz);
    cx = min(max(cx, 0), g.ncx - 1);
    cy = min(max(cy, 0), g.ncy - 1);
    cz = min(max(cz, 0), g.nzb - 1);
    int k = (cx * g.ncy + cy) * g.nzb + cz;
    key[i] = k;
    atomicAdd(&count[k], 1);
}

// ------------------------------------------------------------------------------------------------
// Exclusive scan of int32 (three-phase; blocks of 1024 elements).
// ------------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_scan_block(const int *__restrict__ in, int *__restrict__ out, int m,
                                                   int *__restrict__ bsum)
{
    __shared__ int s[1024];
    int base = blockIdx.x * 1024;
    for (int k = threadIdx.x; k < 1024; k += 256) s[k] = (base + k < m) ? in[base + k] : 0;
    __syncthreads();
    // Thread t scans elements 4t..4t+3 serially, then a Hillis-Steele over 256 partials.
    int t = threadIdx.x, v0 = s[4 * t], v1 = s[4 * t + 1], v2 = s[4 * t + 2], v3 = s[4 * t + 3];
    int tot = v0 + v1 + v2 + v3;
    __shared__ int p[256];
    p[t] = tot;
    __syncthreads();
    for (int o = 1; o < 256; o <<= 1) {
        int add = (t >= o) ? p[t - o] : 0;
        __syncthreads();
        p[t] += add;
        __syncthreads();
    }
    int excl = p[t] - tot;
    int r0 = excl, r1 = r0 + v0, r2 = r1 + v1, r3 = r2 + v2;
    if (base + 4 * t < m) out[base + 4 * t] = r0;
    if (base + 4 * t + 1 < m) out[base + 4 * t + 1] = r1;
    if (base + 4 * t + 2 < m) out[base + 4 * t + 2] = r2;
    if (base + 4 * t + 3 < m) out[base + 4 * t + 3] = r3;
    if (t == 255 && bsum) bsum[blockIdx.x] = p[255];
}

__global__ __launch_bounds__(256) void k_scan_add(int *__restrict__ out, int m, const int *__restrict__ boff)
{
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i < m) out[i] += boff[i / 1024];
}

// ------------------------------------------------------------------------------------------------
// Counting-sort scatter; then each key bucket is ordered by atom index so the
// slot order (and everything downstream) is deterministic.
// ------------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_scatter(const int *__restrict__ key, int n, int *__restrict__ cursor,
                                                 int *__restrict__ unsorted)
{
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    int pos = atomicAdd(&cursor[key[i]], 1);
    unsorted[pos] = i;
}

// Deterministic bucket order: atom i goes to its bucket's start plus the
// number of the bucket's atoms with a smaller index (ascending atom order,
// whatever order the scatter's atomics produced).
__global__ __launch_bounds__(256) void k_bucket_rank(const int *__restrict__ key, const int *__restrict__ off,
                                                     int nkeys, int n, const int *__restrict__ unsorted,
                                                     int *__restrict__ sorted)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int k = key[i];
    const int b = off[k], e = (k + 1 < nkeys) ? off[k + 1] : n;
    int r = 0;
    for (int p = b; p < e; p++) r += unsorted[p] < i ? 1 : 0;
    sorted[b + r] = i;
}

// Column sizes padded to whole super-clusters.
__global__ __launch_bounds__(256) void k_colpad(const int *__restrict__ off, int ncol, int nzb, int n,
                                                int *__restrict__ padded)
{
    int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= ncol) return;
    int b = off[c * nzb];
    int e = (c + 1 < ncol) ? off[(c + 1) * nzb] : n;
    int cnt = e - b;
    padded[c] = (cnt + NB_SC - 1) / NB_SC * NB_SC;
}

__global__ __launch_bounds__(256) void k_fill_int(int *__restrict__ p, size_t m, int v)
{
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < m) p[i] = v;
}

__global__ __launch_bounds__(256) void k_slots(const int *__restrict__ sorted, const int *__restrict__ key, int n,
                                               int nzb, const int *__restrict__ off, const int *__restrict__ colstart,
                                               int *__restrict__ slot_atom, int *__restrict__ atom_slot)
{
    int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= n) return;
    int a = sorted[p];
    int col = key[a] / nzb;
    int slot = colstart[col] + (p - off[col * nzb]);
    slot_atom[slot] = a;
    atom_slot[a] = slot;
}

// Super-cluster shape.  Columns are one super-cluster wide and padded to whole
// super-clusters, so the 64 slots of a super-cluster hold atoms that are
// consecutive in z within one column: a roughly cubic block.  One wavefront
// per super-cluster splits it into 2 x 2 x 2 clusters: its m atoms are ranked
// by x and cut into halves of min(32, m) and the rest, each half ranked by y
// and cut at 16, each quarter ranked by z and cut at 8; full clusters come
// first.  Ties are broken by atom index, so the order is deterministic.  The
// pair set does not depend on it (every pair is tested atom by atom).
// Order of the split: coordinate, then atom index; empty slots (atom -1, key
// +infinity) after every atom.
__device__ __forceinline__ bool before(double u, int a, double v, int b)
{
    const unsigned ua = (unsigned)a, ub = (unsigned)b;   // -1 (empty) compares last
    return u < v || (u == v && ua < ub);
}

// One bitonic sort of the wave's lanes in blocks of n (a power of two, up to
// 64), ascending in every block: lane l ends up holding the element of rank
// l % n of its block.  An element is its key and the lane it came from.
__device__ __forceinline__ void wave_bitonic(double &key, int &src, int &atom, int n)
{
    const int l = threadIdx.x & 63;
    for (int size = 2; size <= n; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            const int pl = l ^ stride;
            const double pk = __hiloint2double(__shfl(__double2hiint(key), pl, 64), __shfl(__double2loint(key), pl, 64));
            const int ps = __shfl(src, pl, 64);
            const int pa = __shfl(atom, pl, 64);
            const bool up = size == n || (l & size) == 0;
            const bool keep_min = ((l & stride) == 0) == up;
            const bool pless = before(pk, pa, key, atom);
            if (pless == keep_min) {
                key = pk;
                src = ps;
                atom = pa;
            }
        }
    }
}

// 2 x 2 x 2 split of each super-cluster: the 64 atoms are sorted by x, each
// half of 32 by y, each quarter of 16 by z, and lane l's atom after the three
// sorts takes slot l (cluster l / 8).  Empty slots sort last at every level, so
// the atoms of a part-filled super-cluster fill its first slots of each part.
__global__ __launch_bounds__(64) void k_sc_split(const double *__restrict__ xyz, int *__restrict__ slot_atom,
                                                 int *__restrict__ atom_slot)
{
    const int l = threadIdx.x;
    const size_t base = (size_t)blockIdx.x * NB_SC;
    const int a0 = slot_atom[base + l];
    double c[3];
    for (int d = 0; d < 3; d++) c[d] = a0 >= 0 ? xyz[3 * (size_t)a0 + d] : __builtin_inf();
    int src = l, atom = a0;
    double key = c[0];
    wave_bitonic(key, src, atom, 64);
    // Next key: y (then z) of the element each lane now holds.
    key = __hiloint2double(__shfl(__double2hiint(c[1]), src, 64), __shfl(__double2loint(c[1]), src, 64));
    wave_bitonic(key, src, atom, 32);
    key = __hiloint2double(__shfl(__double2hiint(c[2]), src, 64), __shfl(__double2loint(c[2]), src, 64));
    wave_bitonic(key, src, atom, 16);
    slot_atom[base + l] = atom;
    if (atom >= 0) atom_slot[atom] = (int)(base + l);
}

// Per-column cluster range [first, last) for the search.
__global__ __launch_bounds__(256) void k_colclusters(const int *__restrict__ colstart, const int *__restrict__ padded,
                                                     int ncol, int2 *__restrict__ colcl)
{
    int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= ncol) return;
    colcl[c] = make_int2(colstart[c] / NB_CL, (colstart[c] + padded[c]) / NB_CL);
}

// ------------------------------------------------------------------------------------------------
// Cluster geometry: FP32 box (outward padded) and an FP32-exact centre.
// ------------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_cluster_geom(const double *__restrict__ xyz, const int *__restrict__ slot_atom,
                                                      int ncl, double ox, double oy, double oz,
                                                      float4 *__restrict__ blo, float4 *__restrict__ bhi,
                                                      float4 *__restrict__ centre)
{
    int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= ncl) return;
    double mn[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, mx[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    double o[3] = {ox, oy, oz};
    int cnt = 0;
    for (int k = 0; k < NB_CL; k++) {
        int a = slot_atom[c * NB_CL + k];
        if (a < 0) continue;
        cnt++;
        for (int d = 0; d < 3; d++) {
            double v = xyz[3 * (size_t)a + d] - o[d];
            mn[d] = fmin(mn[d], v);
            mx[d] = fmax(mx[d], v);
        }
    }
    if (cnt == 0) {
        blo[c] = make_float4(3.0e30f, 3.0e30f, 3.0e30f, 0.0f);
        bhi[c] = make_float4(-3.0e30f, -3.0e30f, -3.0e30f, 0.0f);
        centre[c] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        return;
    }
    float lo[3], hi[3], ce[3];
    for (int d = 0; d < 3; d++) {
        // 1e-3 A outward pad covers the FP64->FP32 rounding of the box.
        lo[d] = (float)(mn[d] - 1.0e-3);
        hi[d] = (float)(mx[d] + 1.0e-3);
        ce[d] = (float)(rint(0.5 * (mn[d] + mx[d]) * 256.0) / 256.0);
    }
    blo[c] = make_float4(lo[0], lo[1], lo[2], (float)cnt);
    bhi[c] = make_float4(hi[0], hi[1], hi[2], 0.0f);
    centre[c] = make_float4(ce[0], ce[1], ce[2], 0.0f);
}

// Slot-ordered search records of the self-list search: the atom's FP64
// position with its index and flag bits (NB_SF_*), and its reference cell.
//
// Point/cell search (pointcell): an atom off the grid still searches as the
// point of its pairs with lower-index atoms.  The reference takes its search
// cell from the cell indices it decoded last (RegularGrid_CellIDToIndices,
// pCore-1.9.0/extensions/csource/RegularGrid.c:69-85, leaves them unchanged
// for a point with no cell): those of the latest earlier atom (index 1 up, in
// the MM selection) on the grid, else the corner (-u, -u, -u) at which the
// search-range construction leaves the same array (RegularGrid.c:526-575).
// That cell is stored for such an atom; its OFFGRID flag stays set.
__global__ __launch_bounds__(256) void k_slot_records(const double *__restrict__ xyz, const int *__restrict__ slot_atom,
                                                      int n_pad, const int4 *__restrict__ cellv,
                                                      const unsigned char *__restrict__ free_flags, int use_grid,
                                                      int pointcell, int u, double4 *__restrict__ srec,
                                                      int4 *__restrict__ scell)
{
    const int s = blockIdx.x * 256 + threadIdx.x;
    if (s >= n_pad) return;
    const int a = slot_atom[s];
    double4 p = make_double4(0.0, 0.0, 0.0, 0.0);
    int4 c = make_int4(0, 0, 0, 0);
    int fl = NB_SF_EMPTY;
    if (a >= 0) {
        p.x = xyz[3 * (size_t)a];
        p.y = xyz[3 * (size_t)a + 1];
        p.z = xyz[3 * (size_t)a + 2];
        c = cellv[a];
        fl = free_flags ? (free_flags[a] & (NB_SF_FREE | NB_SF_QC)) : NB_SF_FREE;
        if (use_grid && c.w == 0) {
            fl |= NB_SF_OFFGRID;
            if (pointcell) {
                int4 sc = make_int4(-u, -u, -u, 0);
                for (int k = a - 1; k >= 1; k--) {
                    const int4 ck = cellv[k];
                    if (ck.w != 0 && !(free_flags && (free_flags[k] & NB_SF_QC))) {
                        sc = make_int4(ck.x, ck.y, ck.z, 0);
                        break;
                    }
                }
                c = sc;
            }
        }
    }
    p.w = __hiloint2double(fl, a);
    srec[s] = p;
    scell[s] = c;
}

// The reference's squared distance, in its operation order
// (PairListGenerator.c:131-140: no contraction into fused multiply-adds).
__device__ __forceinline__ double ref_dist2(double xi, double yi, double zi, double xj, double yj, double zj)
{
#pragma clang fp contract(off)
    const double xij = xi - xj, yij = yi - yj, zij = zi - zj;
    return ((xij * xij) + (yij * yij)) + (zij * zij);
}

__device__ __forceinline__ float box_gap2(float4 alo, float4 ahi, float4 blo, float4 bhi)
{
    float gx = fmaxf(0.0f, fmaxf(alo.x - bhi.x, blo.x - ahi.x));
    float gy = fmaxf(0.0f, fmaxf(alo.y - bhi.y, blo.y - ahi.y));
    float gz = fmaxf(0.0f, fmaxf(alo.z - bhi.z, blo.z - ahi.z));
    return gx * gx + gy * gy + gz * gz;
}

// ------------------------------------------------------------------------------------------------
// Cluster-pair search: one workgroup of NB_SEARCH_WAVES wavefronts per
// super-cluster.  The super-cluster's atoms and the grid stencil are staged
// in LDS.  The candidate j-clusters of all x-y columns within reach are
// scanned as one flattened range and appended to an LDS list in scan order;
// the waves then share out the candidates round-robin and evaluate each with
// the reference pair rule (the j-atom's record in registers, fetched one
// candidate ahead), writing the result into the candidate's own scratch slot.
// k_compact keeps the slots with any listed pair, in candidate order, so the
// entries of each super-cluster come out in a deterministic order.
//
// The reference pair predicate for atoms i, j (PairListGenerator.c:103-143,
// 1020-1150): both atoms in the MM selection, not excluded, OR selection (free atoms), and either the grid
// rule (both on grid; stencil flag 2 = whole cell pair within the cutoff,
// 1 = FP64 distance test, 0 = out of range) or, for the direct method, the
// FP64 distance test alone.  The self-list evaluation computes every part of
// it without branches (a tile's lanes all take the same path); only the
// exclusion scan, needed when j falls inside i's excluded index range, is a
// branch.
//
// STEN (self lists): 0 direct method, 1 grid with the stencil in LDS, 2 grid
// with the stencil read from global memory (more than NB_STENCIL_LDS cells).
// ------------------------------------------------------------------------------------------------
#define NB_SEARCH_WAVES 8
#define NB_SEARCH_THREADS (64 * NB_SEARCH_WAVES)
#define NB_XPOOL 1024        // exclusion indices of one super-cluster held in LDS
#define NB_STENCIL_LDS 4096  // stencil bytes held in LDS (u <= 7)
#define NB_CAND 1024         // candidate list: flushed once it holds more than NB_CAND - NB_SEARCH_THREADS
#define NB_MAXCOL 256        // x-y columns handled per flattened scan

struct JAtom {
    double x, y, z;
    int4 cell;
    int atom, free_;
};

__device__ __forceinline__ JAtom load_jatom_cross(int J, int b, const double *__restrict__ xyz,
                                                  const int *__restrict__ slot_atom,
                                                  const unsigned char *__restrict__ free_flags)
{
    JAtom r;
    r.atom = slot_atom[J * NB_CL + b];
    r.x = r.y = r.z = 0.0;
    r.cell = make_int4(0, 0, 0, 0);
    r.free_ = 1;
    if (r.atom >= 0) {
        r.x = xyz[3 * (size_t)r.atom];
        r.y = xyz[3 * (size_t)r.atom + 1];
        r.z = xyz[3 * (size_t)r.atom + 2];
        if (free_flags) r.free_ = free_flags[r.atom];
    }
    return r;
}

// Write candidate k's entry (scratch slot base + k); returns 1 if it lists a pair.
__device__ __forceinline__ int search_store(Entry *__restrict__ scratch, size_t sbase, int slot, int scap, int lane,
                                            int Jc, int sc, int img, int lbits, int cbits)
{
    if (slot < scap) {
        Entry *en = &scratch[sbase + slot];
        en->bits[lane] = (unsigned char)lbits;
        if (lane == 0) {
            en->J = Jc;
            en->sc = sc;
            en->cbits = cbits;
            en->img = img;
        }
    }
    return cbits ? 1 : 0;
}

// Image lists: the candidates sCand[0..npend) of one wave (k = w, w + NB_SEARCH_WAVES, ...).
//
// Grid method (PairListGenerator.c:703-794,861-940 over the conforming grid of
// the image, RegularGrid.c:341-405): j must lie on the image's grid, in cell
// k2; the offset o = k2 + lowN - I from i's primary search cell I must be in
// the stencil.  Cell/cell (i on the primary grid): the pair is listed when the
// stencil flag of o is 2 (the cell pair is fully within range) or the FP64
// distance test passes.  Point/cell (any i): when the smallest squared distance
// from x_i to cell k2 of the image grid is below cutoff^2 and either the
// largest is at most cutoff^2 or the distance test passes.
template <bool PC>
__device__ __forceinline__ int search_eval_cross(const SearchParams &P, int npend, int w, int lane, int I0, int sc,
                                                 int img, int base, int scap, size_t sbase, const int *sCand,
                                                 const unsigned char *sCm, const int *sIatom, const int *sIfree,
                                                 const int4 *sIcell, const double *sIx, const double *sIy,
                                                 const double *sIz, const ImageGrid &G, const double *__restrict__ jxyz,
                                                 const int *__restrict__ slot_atom, Entry *__restrict__ scratch)
{
    const int a = lane >> 3, b = lane & 7;
    const int wst = 2 * P.u + 1;
    int nvalid = 0;
    JAtom jn;
    if (w < npend) jn = load_jatom_cross(sCand[w], b, jxyz, slot_atom, P.free_flags);
    for (int k = w; k < npend; k += NB_SEARCH_WAVES) {
        const JAtom jd = jn;
        const int Jc = sCand[k];
        const int cmc = sCm[k];
        if (k + NB_SEARCH_WAVES < npend) jn = load_jatom_cross(sCand[k + NB_SEARCH_WAVES], b, jxyz, slot_atom, P.free_flags);
        // j's cell on the image grid (RegularGrid_FindCellIDOfPoint, RegularGrid.c:240-251).
        int jk[3] = {0, 0, 0};
        bool jon = true;
        if (P.use_grid) {
#pragma clang fp contract(off)
            const double jx[3] = {jd.x, jd.y, jd.z};
#pragma unroll
            for (int d = 0; d < 3; d++) {
                jk[d] = (int)floor((jx[d] - G.lower[d]) / P.gh);
                jon = jon && jk[d] >= 0 && jk[d] < G.bins[d];
            }
        }
        int lbits = 0, cbits = 0;
        for (int c = 0; c < NB_SCC; c++) {
            if (!(cmc & (1 << c))) continue;
            const int li = c * NB_CL + a;
            bool ok = (sIatom[li] >= 0) && (jd.atom >= 0);
            if (ok && P.free_flags) {
                // Atom flags: bit 0 free (the OR selection), bit 1 outside the MM
                // selection (pure QC atoms, NBModelABFSState.c:339-341).
                const int fl = sIfree[li] | jd.free_;
                ok = (fl & 1) && !(fl & 2);
            }
            bool near = false;
            if (ok) near = ref_dist2(sIx[li], sIy[li], sIz[li], jd.x, jd.y, jd.z) <= P.cutoff2;
            if (ok && P.use_grid) {
                const int4 ic = sIcell[li];
                const unsigned ux = (unsigned)(jk[0] + G.lowN[0] - ic.x + P.u);
                const unsigned uy = (unsigned)(jk[1] + G.lowN[1] - ic.y + P.u);
                const unsigned uz = (unsigned)(jk[2] + G.lowN[2] - ic.z + P.u);
                const bool inb = max(max(ux, uy), uz) < (unsigned)wst;
                const int flag = inb ? (int)P.stencil[(ux * wst + uy) * wst + uz] : 0;
                ok = jon && flag != 0;
                if (PC) {
                    if (ok) {
                        double minR2, maxR2;
                        point_cell_r2(G.lower, P.gh, sIx[li], sIy[li], sIz[li], jk[0], jk[1], jk[2], minR2, maxR2);
                        ok = (minR2 < P.cutoff2) && (maxR2 <= P.cutoff2 || near);
                    }
                } else {
                    ok = ok && ic.w != 0 && (flag == 2 || near);
                }
            } else {
                ok = ok && near;
            }
            if (ok) lbits |= 1 << c;
            if (__ballot(ok)) cbits |= 1 << c;
        }
        nvalid += search_store(scratch, sbase, base + k, scap, lane, Jc, sc, img, lbits, cbits);
    }
    return nvalid;
}

// Self lists: the candidates sCand[0..npend) of one wave.  Lane 8a + b
// evaluates i-atom (c, a) of each listed tile c against j-atom b; lanes with
// b < a hold the lower triangle of the self tile (J = I0 + c).  Every part of
// the rule is computed without branches; only the exclusion scan (j inside
// i's excluded index range) branches.
//
// PC (point/cell grid search, PairListGenerator.c:1155-1218): the atom with
// the higher index is the point and the other atom's cell is searched.  The
// pair is listed when that cell is in the stencil around the point's search
// cell, the point/cell smallest squared distance is below cutoff^2, and either
// the largest is at most cutoff^2 (the cell is taken whole) or the FP64
// distance test passes.  The lower-index atom must be on the grid; the point
// may be off it (its record then holds its search cell, see k_slot_records).
template <int STEN, bool PC>
__device__ __forceinline__ int search_eval_self(const SearchParams &P, int npend, int w, int lane, int I0, int sc,
                                                int base, int scap, size_t sbase, const int *sCand,
                                                const unsigned char *sCm, const double4 *sIp, const int4 *sIc,
                                                const int2 *sIxr, const int *sXoff, const int *sXcnt,
                                                const int *sXpool, const unsigned char *sSten,
                                                Entry *__restrict__ scratch)
{
    const int a = lane >> 3, b = lane & 7;
    const unsigned wst = 2 * P.u + 1, nst = wst * wst * wst;
    int nvalid = 0;
    double4 jpn = make_double4(0.0, 0.0, 0.0, 0.0);
    int4 jcn = make_int4(0, 0, 0, 0);
    if (w < npend) {
        const int Jn = sCand[w];
        jpn = P.srec[Jn * NB_CL + b];
        if (STEN) jcn = P.scell[Jn * NB_CL + b];
    }
    for (int k = w; k < npend; k += NB_SEARCH_WAVES) {
        const double4 jp = jpn;
        const int4 jc = jcn;
        const int Jc = __builtin_amdgcn_readfirstlane(sCand[k]);
        const int cmc = __builtin_amdgcn_readfirstlane(sCm[k]);
        if (k + NB_SEARCH_WAVES < npend) {
            const int Jn = sCand[k + NB_SEARCH_WAVES];
            jpn = P.srec[Jn * NB_CL + b];
            if (STEN) jcn = P.scell[Jn * NB_CL + b];
        }
        const int jatom = __double2loint(jp.w), jfl = __double2hiint(jp.w);
        int lbits = 0, cbits = 0;
#pragma unroll
        for (int c = 0; c < NB_SCC; c++) {
            if (!(cmc & (1 << c))) continue;
            const int li = c * NB_CL + a;
            const double4 ip = sIp[li];
            const int4 ic = sIc[li];
            const int2 xr = sIxr[li];
            // Both atoms real, in the MM selection and (grid) on the grid, one of them free.
            bool ok;
            if (PC) {
                const bool ipt = __double2loint(ip.w) > jatom;      // i is the point
                ok = (((ic.w | jfl) & (NB_SF_FREE | NB_SF_QC | NB_SF_EMPTY)) == NB_SF_FREE) &
                     !((ipt ? jfl : ic.w) & NB_SF_OFFGRID);
            } else {
                ok = ((ic.w | jfl) & (NB_SF_FREE | NB_SF_QC | NB_SF_OFFGRID | NB_SF_EMPTY)) == NB_SF_FREE;
            }
            ok = ok & ((Jc < I0 + c) | (b < a));
            const bool near = ref_dist2(ip.x, ip.y, ip.z, jp.x, jp.y, jp.z) <= P.cutoff2;
            if (PC) {
                // The stencil is symmetric, so the offset between the two search cells decides membership in
                // either direction.
                const unsigned ux = (unsigned)(jc.x - ic.x), uy = (unsigned)(jc.y - ic.y), uz = (unsigned)(jc.z - ic.z);
                const bool inb = max(max(ux, uy), uz) < wst;
                const unsigned si = min((ux * wst + uy) * wst + uz, nst - 1u);
                const int flag = STEN == 1 ? (int)sSten[si] : (int)P.stencil[si];
                const bool ipt = __double2loint(ip.w) > jatom;
                double minR2, maxR2;
                point_cell_r2(P.glower, P.gh, ipt ? ip.x : jp.x, ipt ? ip.y : jp.y, ipt ? ip.z : jp.z, ipt ? jc.x : ic.x + P.u,
                              ipt ? jc.y : ic.y + P.u, ipt ? jc.z : ic.z + P.u, minR2, maxR2);
                ok = ok & inb & (flag != 0) & (minR2 < P.cutoff2) & ((maxR2 <= P.cutoff2) | near);
            } else if (STEN) {
                // Cell offset + u per axis (sIc holds the cell - u), in [0, 2u] within the stencil.
                const unsigned ux = (unsigned)(jc.x - ic.x), uy = (unsigned)(jc.y - ic.y), uz = (unsigned)(jc.z - ic.z);
                const bool inb = max(max(ux, uy), uz) < wst;
                const unsigned si = min((ux * wst + uy) * wst + uz, nst - 1u);
                const int flag = STEN == 1 ? (int)sSten[si] : (int)P.stencil[si];
                ok = ok & inb & ((flag == 2) | ((flag == 1) & near));
            } else {
                ok = ok & near;
            }
            const bool inx = ok & (jatom >= xr.x) & (jatom <= xr.y);
            if (__ballot(inx)) {
                if (inx) {
                    const int xo = sXoff[li], xn = sXcnt[li];
                    const int *xl = xo >= 0 ? sXpool + xo : P.xidx + (-1 - xo);
                    for (int x = 0; x < xn; x++)
                        if (xl[x] == jatom) ok = false;
                }
            }
            lbits |= ok ? (1 << c) : 0;
            if (__ballot(ok)) cbits |= 1 << c;
        }
        nvalid += search_store(scratch, sbase, base + k, scap, lane, Jc, sc, -1, lbits, cbits);
    }
    return nvalid;
}

//
// CROSS (image lists, blockIdx.y = image): the j-clusters are the image's
// copies of all clusters, scanned as one flat range after a test of the
// super-cluster against the image's bounding box; the predicate is the cross
// list's (pCore-1.9.0/extensions/csource/PairListGenerator.c:28-85,703-794):
// no exclusions, every (i, j) pair including an atom with its own image, the
// OR selection, i on the reference grid (grid method), and the FP64 distance
// test against the image coordinates GenerateImageLists produced.
template <bool CROSS, int STEN, bool PC = false>
__global__ __launch_bounds__(NB_SEARCH_THREADS) __attribute__((amdgpu_waves_per_eu((CROSS || PC) ? 1 : 8))) void k_search(SearchParams P, const double *__restrict__ xyz,
                                                              const int *__restrict__ slot_atom,
                                                              const float4 *__restrict__ blo,
                                                              const float4 *__restrict__ bhi,
                                                              const int2 *__restrict__ colcl,
                                                              Entry *__restrict__ scratch, int *__restrict__ ncand,
                                                              int *__restrict__ counts, int *__restrict__ overflow)
{
    const int sc = blockIdx.x;
    const int img = CROSS ? (int)blockIdx.y : -1;
    const int blk = CROSS ? img * (int)gridDim.x + sc : sc;        // scratch / count slot of this block
    const double *__restrict__ jxyz = CROSS ? P.ixyz + (size_t)img * 3 * P.n : xyz;
    const float4 *__restrict__ jblo = CROSS ? P.iblo + (size_t)img * P.ncl : blo;
    const float4 *__restrict__ jbhi = CROSS ? P.ibhi + (size_t)img * P.ncl : bhi;
    ImageGrid G = {};
    if (CROSS && P.igrid) G = P.igrid[img];
    const size_t sbase = CROSS ? (size_t)P.soff[blk] : (size_t)blk * P.cap;
    const int scap = CROSS ? P.scap[blk] : P.cap;
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    const int a = lane >> 3, b = lane & 7;
    const int I0 = sc * NB_SCC;
    __shared__ float4 sIlo[NB_SCC], sIhi[NB_SCC];
    // Cross lists: the i-atoms' positions, on-grid and free flags.
    __shared__ double sIx[CROSS ? NB_SC : 1], sIy[CROSS ? NB_SC : 1], sIz[CROSS ? NB_SC : 1];
    __shared__ int sIatom[CROSS ? NB_SC : 1], sIfree[CROSS ? NB_SC : 1];
    __shared__ int4 sIcell[CROSS ? NB_SC : 1];
    // Self lists: the i-atoms' records, their cells minus u with the flag bits, exclusion index ranges.
    __shared__ double4 sIp[CROSS ? 1 : NB_SC];
    __shared__ int4 sIc[CROSS ? 1 : NB_SC];
    __shared__ int2 sIxr[CROSS ? 1 : NB_SC];
    __shared__ int sXoff[CROSS ? 1 : NB_SC], sXcnt[CROSS ? 1 : NB_SC];
    __shared__ int sXpool[CROSS ? 1 : NB_XPOOL];
    __shared__ unsigned char sSten[STEN == 1 ? NB_STENCIL_LDS : 1];
    __shared__ int sCand[NB_CAND];
    __shared__ unsigned char sCm[NB_CAND];
    __shared__ int sColStart[NB_MAXCOL], sColPre[NB_MAXCOL + 1];
    __shared__ int sWcnt[NB_SEARCH_WAVES];

    // Stage the super-cluster (wave 0) and the stencil (all waves).
    if (w == 0) {
        if (lane < NB_SCC) {
            sIlo[lane] = blo[I0 + lane];
            sIhi[lane] = bhi[I0 + lane];
        }
        if (CROSS) {
            const int ia = slot_atom[I0 * NB_CL + lane];
            sIatom[lane] = ia;
            if (ia >= 0) {
                sIx[lane] = xyz[3 * (size_t)ia];
                sIy[lane] = xyz[3 * (size_t)ia + 1];
                sIz[lane] = xyz[3 * (size_t)ia + 2];
                sIcell[lane] = P.icell ? P.icell[ia] : make_int4(0, 0, 0, 1);
                sIfree[lane] = P.free_flags ? P.free_flags[ia] : 1;
            }
        } else {
            const double4 ip = P.srec[I0 * NB_CL + lane];
            const int4 ic = P.scell[I0 * NB_CL + lane];
            const int ia = __double2loint(ip.w);
            sIp[lane] = ip;
            sIc[lane] = make_int4(ic.x - P.u, ic.y - P.u, ic.z - P.u, __double2hiint(ip.w));
            int xs = 0, xc = 0;
            if (ia >= 0) {
                xs = P.xptr[ia];
                xc = P.xptr[ia + 1] - xs;
            }
            int pre = xc;   // inclusive scan of exclusion counts over the wave
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const int v = __shfl_up(pre, o, 64);
                if (lane >= o) pre += v;
            }
            const int off = pre - xc;
            if (off + xc <= NB_XPOOL) {
                for (int k = 0; k < xc; k++) sXpool[off + k] = P.xidx[xs + k];
                sXoff[lane] = off;
            } else {
                sXoff[lane] = -1 - xs;      // list read from global memory
            }
            sXcnt[lane] = xc;
            // Rows are ascending: j outside [first, last] is not excluded.
            sIxr[lane] = xc > 0 ? make_int2(P.xidx[xs], P.xidx[xs + xc - 1]) : make_int2(0, -1);
        }
    }
    const int wst = 2 * P.u + 1;
    const int nst = wst * wst * wst;
    if (STEN == 1)
        for (int k = tid; k < nst; k += NB_SEARCH_THREADS) sSten[k] = P.stencil[k];
    __syncthreads();

    float4 lo = sIlo[0], hi = sIhi[0];
    for (int c = 1; c < NB_SCC; c++) {
        lo.x = fminf(lo.x, sIlo[c].x); lo.y = fminf(lo.y, sIlo[c].y); lo.z = fminf(lo.z, sIlo[c].z);
        hi.x = fmaxf(hi.x, sIhi[c].x); hi.y = fmaxf(hi.y, sIhi[c].y); hi.z = fmaxf(hi.z, sIhi[c].z);
    }
    int nvalid = 0;      // entries written by this wave
    int base = 0;        // candidates handed out before the current list
    if (lo.x <= hi.x) {  // super-cluster holds at least one atom (uniform over the block)
        const float reach = P.reach;
        const float reach2 = reach * reach;
        int cx0 = 0, cy0 = 0, nyc = 1, ncol = 0;
        const int Jmax = I0 + NB_SCC - 1;
        // z-bins that can hold an atom within reach of the super-cluster (one bin of margin for the FP32 box).
        int zb_lo = 0, zb_hi = 0;
        if (!CROSS) {
            zb_lo = min(max(0, (int)floorf((lo.z - reach) / P.dz) - 1), P.nzb - 1);
            zb_hi = min(max(0, (int)floorf((hi.z + reach) / P.dz) + 1), P.nzb - 1);
            cx0 = (int)floorf((lo.x - reach) / P.cw);
            cy0 = (int)floorf((lo.y - reach) / P.cw);
            int cx1 = (int)floorf((hi.x + reach) / P.cw), cy1 = (int)floorf((hi.y + reach) / P.cw);
            cx0 = max(cx0, 0); cy0 = max(cy0, 0);
            cx1 = min(cx1, P.ncx - 1); cy1 = min(cy1, P.ncy - 1);
            nyc = cy1 - cy0 + 1;
            ncol = (cx1 >= cx0 && cy1 >= cy0) ? (cx1 - cx0 + 1) * nyc : 0;
        } else {
            // One flat range over the image's clusters, if its box is in reach.
            const float4 ilo = make_float4(P.ibox[6 * img], P.ibox[6 * img + 1], P.ibox[6 * img + 2], 0.0f);
            const float4 ihi = make_float4(P.ibox[6 * img + 3], P.ibox[6 * img + 4], P.ibox[6 * img + 5], 0.0f);
            ncol = box_gap2(lo, hi, ilo, ihi) < reach2 ? 1 : 0;
        }

        // Evaluate the listed candidates sCand[0..npend): this wave takes
        // k = w, w + NB_SEARCH_WAVES, ...
        auto evaluate = [&](int npend) {
            if (CROSS)
                nvalid += search_eval_cross<PC>(P, npend, w, lane, I0, sc, img, base, scap, sbase, sCand, sCm, sIatom,
                                                sIfree, sIcell, sIx, sIy, sIz, G, jxyz, slot_atom, scratch);
            else
                nvalid += search_eval_self<STEN, PC>(P, npend, w, lane, I0, sc, base, scap, sbase, sCand, sCm, sIp, sIc,
                                                 sIxr, sXoff, sXcnt, sXpool, sSten, scratch);
        };

        int npend = 0;
        for (int c0 = 0; c0 < ncol; c0 += NB_MAXCOL) {
            const int nc = min(NB_MAXCOL, ncol - c0);
            // Column ranges [start, min(end, Jmax+1)) and their prefix sums (wave 0).
            if (!CROSS && w == 0) {
                int carry = 0;
                for (int cb0 = 0; cb0 < nc; cb0 += 64) {
                    const int k = cb0 + lane;
                    int len = 0;
                    if (k < nc) {
                        const int col = c0 + k;
                        const int ci = (cx0 + col / nyc) * P.ncy + cy0 + col % nyc;
                        const int2 r = colcl[ci];
                        // The column's atoms are in z-bin order from its first slot: the super-clusters holding
                        // atoms of bins zb_lo..zb_hi (slots plo..phi of the column) are the only ones in reach.
                        const int base = P.zoff[ci * P.nzb];
                        const int plo = P.zoff[ci * P.nzb + zb_lo] - base;
                        const int phi = (zb_hi + 1 < P.nzb ? P.zoff[ci * P.nzb + zb_hi + 1]
                                                            : (ci + 1 < P.ncol ? P.zoff[(ci + 1) * P.nzb] : P.n)) -
                                        base;
                        const int js = max(r.x, r.x + (plo / NB_SC) * NB_SCC);
                        const int je = min(min(r.y, Jmax + 1), r.x + (phi + NB_SC - 1) / NB_SC * NB_SCC);
                        len = max(0, je - js);
                        sColStart[k] = js;
                    }
                    int incl = len;
#pragma unroll
                    for (int o = 1; o < 64; o <<= 1) {
                        const int v = __shfl_up(incl, o, 64);
                        if (lane >= o) incl += v;
                    }
                    if (k < nc) sColPre[k] = carry + incl - len;
                    carry += __shfl(incl, 63, 64);
                }
                if (lane == 0) sColPre[nc] = carry;
            }
            __syncthreads();
            const int total = CROSS ? P.ncl : sColPre[nc];
            for (int f0 = 0; f0 < total; f0 += NB_SEARCH_THREADS) {
                const int f = f0 + tid;
                int cm = 0, J = 0;
                if (f < total) {
                    if (CROSS) {
                        J = f;
                    } else {
                        int l = 0, h = nc;   // last column with sColPre[col] <= f
                        while (h - l > 1) {
                            const int m = (l + h) >> 1;
                            if (sColPre[m] <= f) l = m; else h = m;
                        }
                        J = sColStart[l] + (f - sColPre[l]);
                    }
                    const float4 jl = jblo[J], jh = jbhi[J];
                    if (box_gap2(lo, hi, jl, jh) < reach2) {
                        for (int c = 0; c < NB_SCC; c++)
                            if ((CROSS || J <= I0 + c) && box_gap2(sIlo[c], sIhi[c], jl, jh) < reach2) cm |= 1 << c;
                    }
                }
                const unsigned long long m = __ballot(cm != 0);
                if (lane == 0) sWcnt[w] = __popcll(m);
                __syncthreads();
                int off = npend, tot = 0;
                for (int v = 0; v < NB_SEARCH_WAVES; v++) {
                    const int cv = sWcnt[v];
                    if (v < w) off += cv;
                    tot += cv;
                }
                const int pos = off + __popcll(m & ((1ull << lane) - 1ull));
                if (cm) {
                    sCand[pos] = J;
                    sCm[pos] = (unsigned char)cm;
                }
                npend += tot;
                __syncthreads();
                if (npend > NB_CAND - NB_SEARCH_THREADS) {
                    evaluate(npend);
                    base += npend;
                    npend = 0;
                    __syncthreads();
                }
            }
            __syncthreads();
        }
        evaluate(npend);
        base += npend;
    }
    if (lane == 0) sWcnt[w] = nvalid;
    __syncthreads();
    if (tid == 0) {
        int tot = 0;
        for (int v = 0; v < NB_SEARCH_WAVES; v++) tot += sWcnt[v];
        counts[blk] = tot;
        ncand[blk] = base;
        if (base > scap) atomicMax(overflow, base);
    }
}

// Compact per-super-cluster entries (candidate slots with any listed pair, in
// candidate order), count their pairs into *npairs (or per block), and cut
// them into work items.  Block b compacts search block b, whose
// super-cluster is b % nsc.
#define NB_COMPACT_THREADS 256
__global__ __launch_bounds__(NB_COMPACT_THREADS) void k_compact(const Entry *__restrict__ scratch, int cap,
                                                                const int *__restrict__ soff,
                                                                const int *__restrict__ ncand,
                                                                const int *__restrict__ counts,
                                                                const int *__restrict__ eoff,
                                                                const int *__restrict__ ioff, int nsc, int per_item,
                                                                int sys, Entry *__restrict__ entries,
                                                                WorkItem *__restrict__ items,
                                                                unsigned long long *__restrict__ npairs_block)
{
    const int sc = blockIdx.x, tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int nc = soff ? ncand[sc] : min(ncand[sc], cap), cnt = counts[sc], e0 = eoff[sc], i0 = ioff[sc];
    const Entry *__restrict__ src = scratch + (soff ? (size_t)soff[sc] : (size_t)sc * cap);
    __shared__ int sW[NB_COMPACT_THREADS / 64];
    __shared__ unsigned long long sP[NB_COMPACT_THREADS / 64];
    int pos = 0;
    unsigned long long pc = 0;
    // Chunks of NB_COMPACT_THREADS candidates: the valid ones keep their order.
    for (int k0 = 0; k0 < nc; k0 += NB_COMPACT_THREADS) {
        const int k = k0 + tid;
        const bool valid = k < nc && src[k].cbits != 0;
        const unsigned long long m = __ballot(valid);
        if (lane == 0) sW[w] = __popcll(m);
        __syncthreads();
        int before = 0, total = 0;
        for (int v = 0; v < NB_COMPACT_THREADS / 64; v++) {
            before += v < w ? sW[v] : 0;
            total += sW[v];
        }
        if (valid) {
            // 80-byte entry as five 16-byte words.
            const int4 *s4 = (const int4 *)&src[k];
            int4 *d4 = (int4 *)&entries[e0 + pos + before + __popcll(m & ((1ull << lane) - 1ull))];
            int4 q[5];
#pragma unroll
            for (int t = 0; t < 5; t++) q[t] = s4[t];
#pragma unroll
            for (int t = 0; t < 5; t++) d4[t] = q[t];
            // The lane masks are words 1..4.
#pragma unroll
            for (int t = 1; t < 5; t++)
                pc += (unsigned long long)(__popc(q[t].x) + __popc(q[t].y) + __popc(q[t].z) + __popc(q[t].w));
        }
        pos += total;
        __syncthreads();
    }
    for (int o = 32; o >= 1; o >>= 1) pc += __shfl_xor(pc, o, 64);
    if (lane == 0) sP[w] = pc;
    __syncthreads();
    if (tid == 0) {
        unsigned long long t = 0;
        for (int v = 0; v < NB_COMPACT_THREADS / 64; v++) t += sP[v];
        npairs_block[sc] = t;
    }
    const int nit = (cnt + per_item - 1) / per_item;
    for (int k = tid; k < nit; k += NB_COMPACT_THREADS) {
        WorkItem wi;
        wi.sys = sys;
        if (soff) wi.sys |= (sc / nsc + 1) << 16;   // image lists: image of the block, plus one
        wi.sc = sc % nsc;
        wi.first = e0 + k * per_item;
        wi.count = min(per_item, cnt - k * per_item);
        items[i0 + k] = wi;
    }
}

// Longest-first order of the work items (by the number of i-cluster tiles of
// their entries), so that the items still running when the GPU drains are the
// short ones.  Results do not depend on the order: every sum is fixed point.
__global__ __launch_bounds__(256) void k_item_cost(const WorkItem *__restrict__ items, int nitems,
                                                   const Entry *__restrict__ entries, int maxcost,
                                                   int *__restrict__ cost, int *__restrict__ hist)
{
    // One wavefront per item: its lanes read the entries' i-cluster masks.
    const int k = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (k >= nitems) return;
    const WorkItem w = items[k];
    int c = 0;
    for (int e = w.first + lane; e < w.first + w.count; e += 64) c += __popc(entries[e].cbits & 0xff);
    for (int o = 32; o >= 1; o >>= 1) c += __shfl_xor(c, o, 64);
    c = min(c, maxcost);
    if (lane == 0) {
        cost[k] = c;
        if (hist) atomicAdd(&hist[maxcost - c], 1);
    }
}

// Rank r in the longest-first order goes to position r, or, for the
// persistent queues (interleave != 0), to queue r % 8 at place r / 8, so every
// per-XCD queue is itself longest-first and the queues are balanced.
__device__ __forceinline__ int queue_start(int nitems, int q)
{
    int s = 0;
    for (int v = 0; v < q; v++) s += (nitems - v + 7) / 8;
    return s;
}

// Longest-first order in one workgroup: histogram of the costs, its scan and
// the scatter all in LDS.  Most items share a few costs, so the histogram is
// kept in NB_ORDER_COPIES copies (wave w adds into copy w % NB_ORDER_COPIES)
// to spread the atomics on a hot bin; the scan runs over (bin, copy) in that
// order.  Items of equal cost keep no particular order.
#define NB_ORDER_THREADS 1024
#define NB_ORDER_BINS 2048
#define NB_ORDER_COPIES 4
#define NB_ORDER_UNROLL 8
__global__ __launch_bounds__(NB_ORDER_THREADS) void k_item_order(const WorkItem *__restrict__ items, int nitems,
                                                                 const int *__restrict__ cost, int maxcost,
                                                                 int interleave, WorkItem *__restrict__ out)
{
    __shared__ int sH[NB_ORDER_BINS * NB_ORDER_COPIES];    // [bin][copy]
    __shared__ int sW[NB_ORDER_THREADS / 64];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6, cp = w % NB_ORDER_COPIES;
    for (int b = t; b < NB_ORDER_BINS * NB_ORDER_COPIES; b += NB_ORDER_THREADS) sH[b] = 0;
    __syncthreads();
    // The loads of a pass are issued together (indices clamped, so none is conditional).
    for (int k0 = t; k0 < nitems; k0 += NB_ORDER_UNROLL * NB_ORDER_THREADS) {
        int cc[NB_ORDER_UNROLL];
#pragma unroll
        for (int u = 0; u < NB_ORDER_UNROLL; u++) cc[u] = cost[min(k0 + u * NB_ORDER_THREADS, nitems - 1)];
#pragma unroll
        for (int u = 0; u < NB_ORDER_UNROLL; u++)
            if (k0 + u * NB_ORDER_THREADS < nitems) atomicAdd(&sH[(maxcost - cc[u]) * NB_ORDER_COPIES + cp], 1);
    }
    __syncthreads();
    // Exclusive scan over (bin, copy): eight consecutive counters per thread.
    constexpr int PER = NB_ORDER_BINS * NB_ORDER_COPIES / NB_ORDER_THREADS;
    int h[PER], sum = 0;
#pragma unroll
    for (int q = 0; q < PER; q++) {
        h[q] = sH[PER * t + q];
        sum += h[q];
    }
    int incl = sum;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int v = __shfl_up(incl, o, 64);
        if (lane >= o) incl += v;
    }
    if (lane == 63) sW[w] = incl;
    __syncthreads();
    int ex = incl - sum;
    for (int v = 0; v < w; v++) ex += sW[v];
#pragma unroll
    for (int q = 0; q < PER; q++) {
        sH[PER * t + q] = ex;
        ex += h[q];
    }
    __syncthreads();
    for (int k0 = t; k0 < nitems; k0 += NB_ORDER_UNROLL * NB_ORDER_THREADS) {
        int cc[NB_ORDER_UNROLL];
        int4 wi[NB_ORDER_UNROLL];
#pragma unroll
        for (int u = 0; u < NB_ORDER_UNROLL; u++) {
            const int k = min(k0 + u * NB_ORDER_THREADS, nitems - 1);
            cc[u] = cost[k];
            wi[u] = ((const int4 *)items)[k];
        }
        int r[NB_ORDER_UNROLL];
#pragma unroll
        for (int u = 0; u < NB_ORDER_UNROLL; u++)
            r[u] = k0 + u * NB_ORDER_THREADS < nitems ? atomicAdd(&sH[(maxcost - cc[u]) * NB_ORDER_COPIES + cp], 1) : -1;
#pragma unroll
        for (int u = 0; u < NB_ORDER_UNROLL; u++)
            if (r[u] >= 0) ((int4 *)out)[interleave ? queue_start(nitems, r[u] & 7) + (r[u] >> 3) : r[u]] = wi[u];
    }
}

__global__ __launch_bounds__(256) void k_item_scatter(const WorkItem *__restrict__ items, int nitems,
                                                      const int *__restrict__ cost, int maxcost, int interleave,
                                                      int *__restrict__ cursor, WorkItem *__restrict__ out)
{
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= nitems) return;
    const int r = atomicAdd(&cursor[maxcost - cost[k]], 1);
    out[interleave ? queue_start(nitems, r & 7) + (r >> 3) : r] = items[k];
}

// Work items of an already compacted list for another entries-per-item cut.
// images != 0: the blocks are (image, super-cluster) blocks of the image
// lists, and each item's sys field carries (image + 1) << 16.
__global__ __launch_bounds__(256) void k_cut_items(const int *__restrict__ counts, const int *__restrict__ eoff,
                                                   const int *__restrict__ ioff, int nblocks, int nsc, int per_item,
                                                   int sys, int images, WorkItem *__restrict__ items)
{
    const int s = blockIdx.x * 256 + threadIdx.x;
    if (s >= nblocks) return;
    const int cnt = counts[s], e0 = eoff[s], i0 = ioff[s];
    for (int k = 0; k * per_item < cnt; k++) {
        WorkItem wi;
        wi.sys = images ? (sys | ((s / nsc + 1) << 16)) : sys;
        wi.sc = s % nsc;
        wi.first = e0 + k * per_item;
        wi.count = min(per_item, cnt - k * per_item);
        items[i0 + k] = wi;
    }
}

// Entry and item offsets of the super-clusters from their entry counts, and
// the totals (entries, items, search capacity needed) into out[0..2], in one
// workgroup (each thread scans a contiguous chunk).
#define NB_OFFS_THREADS 1024
#define NB_OFFS_MAX (64 * NB_OFFS_THREADS)
__global__ __launch_bounds__(NB_OFFS_THREADS) void k_list_offsets(const int *__restrict__ counts, int m, int per_item,
                                                                  int *__restrict__ eoff, int *__restrict__ nits,
                                                                  int *__restrict__ ioff, const int *__restrict__ need,
                                                                  int *__restrict__ out)
{
    __shared__ int sE[NB_OFFS_THREADS / 64], sI[NB_OFFS_THREADS / 64];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int chunk = (m + NB_OFFS_THREADS - 1) / NB_OFFS_THREADS;
    const int lo = min(t * chunk, m), hi = min(lo + chunk, m);
    int se = 0, si = 0;
    for (int k = lo; k < hi; k++) {
        const int c = counts[k];
        se += c;
        si += (c + per_item - 1) / per_item;
    }
    int ie = se, ii = si;   // inclusive scans over the wave
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int v = __shfl_up(ie, o, 64), u = __shfl_up(ii, o, 64);
        if (lane >= o) {
            ie += v;
            ii += u;
        }
    }
    if (lane == 63) {
        sE[w] = ie;
        sI[w] = ii;
    }
    __syncthreads();
    int be = 0, bi = 0, te = 0, ti = 0;
    for (int v = 0; v < NB_OFFS_THREADS / 64; v++) {
        if (v < w) {
            be += sE[v];
            bi += sI[v];
        }
        te += sE[v];
        ti += sI[v];
    }
    int e = be + ie - se, i = bi + ii - si;
    for (int k = lo; k < hi; k++) {
        const int c = counts[k], ni = (c + per_item - 1) / per_item;
        eoff[k] = e;
        ioff[k] = i;
        nits[k] = ni;
        e += c;
        i += ni;
    }
    if (t == 0) {
        volatile int *o = out;
        o[0] = te;
        o[1] = ti;
        o[2] = *need;
    }
}

__global__ __launch_bounds__(256) void k_items_per_sc(const int *__restrict__ counts, int nsc, int per_item,
                                                      int *__restrict__ nitems)
{
    int s = blockIdx.x * 256 + threadIdx.x;
    if (s < nsc) nitems[s] = (counts[s] + per_item - 1) / per_item;
}

// Canonical pair keys max*n+min (for parity against the reference set).
__global__ __launch_bounds__(64) void k_export_pairs(const Entry *__restrict__ entries, int nent,
                                                     const int *__restrict__ slot_atom, long long n,
                                                     unsigned long long *__restrict__ cursor, long long cap,
                                                     long long *__restrict__ keys)
{
    int e = blockIdx.x;
    if (e >= nent) return;
    const Entry &en = entries[e];
    int lane = threadIdx.x, a = lane >> 3, b = lane & 7;
    int j = slot_atom[en.J * NB_CL + b];
    const int lb = en.bits[lane];
    for (int c = 0; c < NB_SCC; c++) {
        if (!((lb >> c) & 1)) continue;
        int i = slot_atom[(en.sc * NB_SCC + c) * NB_CL + a];
        long long hi = max(i, j), lo = min(i, j);
        unsigned long long pos = atomicAdd(cursor, 1ull);
        if ((long long)pos < cap) keys[pos] = hi * n + lo;
    }
}

// Candidate bound of each (image, super-cluster) block of the cross search:
// the image clusters whose box is within reach of the super-cluster's box
// (the first cull of k_search), zero when the image's box is out of reach.
__global__ __launch_bounds__(256) void k_image_bounds(SearchParams P, const float4 *__restrict__ blo,
                                                      const float4 *__restrict__ bhi, int *__restrict__ bound)
{
    __shared__ int sCnt[4];
    const int sc = blockIdx.x, img = blockIdx.y, I0 = sc * NB_SCC;
    float4 lo = blo[I0], hi = bhi[I0];
    for (int c = 1; c < NB_SCC; c++) {
        const float4 l = blo[I0 + c], h = bhi[I0 + c];
        lo.x = fminf(lo.x, l.x); lo.y = fminf(lo.y, l.y); lo.z = fminf(lo.z, l.z);
        hi.x = fmaxf(hi.x, h.x); hi.y = fmaxf(hi.y, h.y); hi.z = fmaxf(hi.z, h.z);
    }
    const float reach2 = P.reach * P.reach;
    const float4 ilo = make_float4(P.ibox[6 * img], P.ibox[6 * img + 1], P.ibox[6 * img + 2], 0.0f);
    const float4 ihi = make_float4(P.ibox[6 * img + 3], P.ibox[6 * img + 4], P.ibox[6 * img + 5], 0.0f);
    int cnt = 0;
    if (lo.x <= hi.x && box_gap2(lo, hi, ilo, ihi) < reach2) {
        const float4 *__restrict__ jl = P.iblo + (size_t)img * P.ncl;
        const float4 *__restrict__ jh = P.ibhi + (size_t)img * P.ncl;
        for (int J = threadIdx.x; J < P.ncl; J += 256)
            if (box_gap2(lo, hi, jl[J], jh[J]) < reach2) cnt++;
    }
    for (int o = 32; o >= 1; o >>= 1) cnt += __shfl_xor(cnt, o, 64);
    if ((threadIdx.x & 63) == 0) sCnt[threadIdx.x >> 6] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) bound[img * gridDim.x + sc] = sCnt[0] + sCnt[1] + sCnt[2] + sCnt[3];
}

__global__ __launch_bounds__(256) void k_add_int(int *__restrict__ p, int m, int v)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < m) p[i] += v;
}

// Pairs (i, j) of one image as keys i*n + j (i in the primary cell, j in the image).
__global__ __launch_bounds__(64) void k_export_image_pairs(const Entry *__restrict__ entries, int nent,
                                                           const int *__restrict__ slot_atom, long long n, int img,
                                                           unsigned long long *__restrict__ cursor, long long cap,
                                                           long long *__restrict__ keys)
{
    int e = blockIdx.x;
    if (e >= nent) return;
    const Entry &en = entries[e];
    if (en.img != img) return;
    int lane = threadIdx.x, a = lane >> 3, b = lane & 7;
    int j = slot_atom[en.J * NB_CL + b];
    const int lb = en.bits[lane];
    for (int c = 0; c < NB_SCC; c++) {
        if (!((lb >> c) & 1)) continue;
        int i = slot_atom[(en.sc * NB_SCC + c) * NB_CL + a];
        unsigned long long pos = atomicAdd(cursor, 1ull);
        if ((long long)pos < cap) keys[pos] = (long long)i * n + j;
    }
}

// ------------------------------------------------------------------------------------------------
// Image coordinates (GenerateImageLists, pMolecule-1.9.0/extensions/csource/
// NBModelABFS.c:841-1000), in the reference's operation order so that the
// image lists are searched with the reference's coordinates bit for bit:
// Coordinates3_Rotate + _Translate with the orthogonalized transformation
// (pCore-1.9.0/extensions/csource/Coordinates3.c:1473-1500,1760-1800), then per
// (a, b, c) of the transformation's sequence a translation by +d and one by -d.
// ------------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_image_transform(const double *__restrict__ xyz, int n,
                                                         const double *__restrict__ rt, double *__restrict__ out)
{
#pragma clang fp contract(off)
    const int i = blockIdx.x * 256 + threadIdx.x, t = blockIdx.y;
    if (i >= n) return;
    const double *R = rt + 12 * t, *T = R + 9;
    const double x0 = xyz[3 * (size_t)i], y0 = xyz[3 * (size_t)i + 1], z0 = xyz[3 * (size_t)i + 2];
    const double x1 = R[0] * x0 + R[1] * y0 + R[2] * z0;
    const double y1 = R[3] * x0 + R[4] * y0 + R[5] * z0;
    const double z1 = R[6] * x0 + R[7] * y0 + R[8] * z0;
    double *o = out + ((size_t)t * n + i) * 3;
    o[0] = x1 + T[0];
    o[1] = y1 + T[1];
    o[2] = z1 + T[2];
}

__global__ __launch_bounds__(256) void k_image_coords(const double *__restrict__ tcoords, int n,
                                                      const int *__restrict__ step_off,
                                                      const ImageStep *__restrict__ steps, double *__restrict__ out)
{
#pragma clang fp contract(off)
    const int i = blockIdx.x * 256 + threadIdx.x, t = blockIdx.y;
    if (i >= n) return;
    const double *c = tcoords + ((size_t)t * n + i) * 3;
    double x = c[0], y = c[1], z = c[2];
    for (int k = step_off[t]; k < step_off[t + 1]; k++) {
        const ImageStep q = steps[k];
        x += q.d[0];
        y += q.d[1];
        z += q.d[2];
        if (q.out >= 0) {
            double *o = out + ((size_t)q.out * n + i) * 3;
            o[0] = x;
            o[1] = y;
            o[2] = z;
        }
        x += -q.d[0];
        y += -q.d[1];
        z += -q.d[2];
    }
}

// Primary-grid cells of the i atoms of the image lists (the grid made from the
// current coordinates, NBModelABFSState_GridInitialize): cell indices with
// w = 1 on the grid, w = 0 off it.  Point/cell search: an atom off the grid
// searches around the cell indices decoded last in the reference's point loop
// (RegularGrid_CellIDToIndices, pCore-1.9.0/extensions/csource/RegularGrid.c:
// 69-85, writes nothing for a point with no cell), those of the latest earlier
// atom (index 0 up, in the MM selection) on the grid, else (-u, -u, -u).
__global__ __launch_bounds__(256) void k_image_cells(const double *__restrict__ xyz, int n, BuildGrid g, int pointcell,
                                                     int u, const unsigned char *__restrict__ free_flags,
                                                     int4 *__restrict__ icell)
{
#pragma clang fp contract(off)
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    int k[3], on = 1;
    for (int d = 0; d < 3; d++) {
        k[d] = (int)floor((xyz[3 * (size_t)i + d] - g.lower[d]) / g.h);
        if (k[d] < 0 || k[d] >= g.bins[d]) on = 0;
    }
    int4 c = make_int4(k[0], k[1], k[2], on);
    if (!on) {
        c = make_int4(-u, -u, -u, 0);
        if (pointcell)
            for (int q = i - 1; q >= 0; q--) {
                if (free_flags && (free_flags[q] & NB_SF_QC)) continue;
                int kq[3], onq = 1;
                for (int d = 0; d < 3; d++) {
                    kq[d] = (int)floor((xyz[3 * (size_t)q + d] - g.lower[d]) / g.h);
                    if (kq[d] < 0 || kq[d] >= g.bins[d]) onq = 0;
                }
                if (onq) {
                    c = make_int4(kq[0], kq[1], kq[2], 0);
                    break;
                }
            }
    }
    icell[i] = c;
}

// The grid of each image's points conforming to the primary grid
// (Coordinates3_MakeConformingGrid, pCore-1.9.0/extensions/csource/
// Coordinates3.c:998-1072): the box of the image points (Coordinates3_
// EnclosingOrthorhombicBox: minimum, and extent added back for the upper
// corner), the primary-grid cell indices of both corners (floor, not
// regularised), bins from the lower to the upper corner's cell, and the lower
// corner lower + lowN h.  One workgroup per image.
__global__ __launch_bounds__(256) void k_image_grid(const double *__restrict__ ixyz, int n, BuildGrid g,
                                                    ImageGrid *__restrict__ igrid)
{
#pragma clang fp contract(off)
    __shared__ double smn[3][256], smx[3][256];
    const double *x = ixyz + (size_t)blockIdx.x * 3 * n;
    double mn[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, mx[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    for (int i = threadIdx.x; i < n; i += 256)
        for (int d = 0; d < 3; d++) {
            const double v = x[3 * (size_t)i + d];
            mn[d] = fmin(mn[d], v);
            mx[d] = fmax(mx[d], v);
        }
    for (int d = 0; d < 3; d++) {
        smn[d][threadIdx.x] = mn[d];
        smx[d][threadIdx.x] = mx[d];
    }
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o)
            for (int d = 0; d < 3; d++) {
                smn[d][threadIdx.x] = fmin(smn[d][threadIdx.x], smn[d][threadIdx.x + o]);
                smx[d][threadIdx.x] = fmax(smx[d][threadIdx.x], smx[d][threadIdx.x + o]);
            }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        ImageGrid G;
        for (int d = 0; d < 3; d++) {
            const double lo = smn[d][0];
            const double up = (smx[d][0] - lo) + lo;
            const int lowN = (int)floor((lo - g.lower[d]) / g.h);
            const int upN = (int)floor((up - g.lower[d]) / g.h);
            G.lowN[d] = lowN;
            G.bins[d] = upN - lowN + 1;
            G.lower[d] = g.lower[d] + (double)lowN * g.h;
        }
        igrid[blockIdx.x] = G;
    }
}

// Full (both directions) CSR neighbour rows for the plain FP64 path.
__global__ __launch_bounds__(64) void k_csr_degree(const Entry *__restrict__ entries, int nent,
                                                   const int *__restrict__ slot_atom, int *__restrict__ deg)
{
    int e = blockIdx.x;
    if (e >= nent) return;
    const Entry &en = entries[e];
    int lane = threadIdx.x, a = lane >> 3, b = lane & 7;
    int j = slot_atom[en.J * NB_CL + b];
    const int lb = en.bits[lane];
    for (int c = 0; c < NB_SCC; c++) {
        if (!((lb >> c) & 1)) continue;
        int i = slot_atom[(en.sc * NB_SCC + c) * NB_CL + a];
        atomicAdd(&deg[i], 1);
        atomicAdd(&deg[j], 1);
    }
}

__global__ __launch_bounds__(64) void k_csr_fill(const Entry *__restrict__ entries, int nent,
                                                 const int *__restrict__ slot_atom, int *__restrict__ cursor,
                                                 int *__restrict__ cols)
{
    int e = blockIdx.x;
    if (e >= nent) return;
    const Entry &en = entries[e];
    int lane = threadIdx.x, a = lane >> 3, b = lane & 7;
    int j = slot_atom[en.J * NB_CL + b];
    const int lb = en.bits[lane];
    for (int c = 0; c < NB_SCC; c++) {
        if (!((lb >> c) & 1)) continue;
        int i = slot_atom[(en.sc * NB_SCC + c) * NB_CL + a];
        cols[atomicAdd(&cursor[i], 1)] = j;
        cols[atomicAdd(&cursor[j], 1)] = i;
    }
}

// Ascending sort of each CSR row, so that the plain path sums every row in the
// same order whatever order the fill's atomics took.  A row's entries are
// distinct atom indices.  Bitonic network in its "flip" form, which sorts any
// length: positions at or beyond the row end act as +infinity and never move.
#define NB_ROW_LDS 4096

template <class P>
__device__ void bitonic_row(P v, int len)
{
    int m = 1;
    while (m < len) m <<= 1;
    for (int k = 2; k <= m; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int t = threadIdx.x; t < m / 2; t += blockDim.x) {
                const int blk = t / j, off = t % j;
                int lo, hi;
                if (j == (k >> 1)) {        // flip comparator of the merge's first stage
                    const int base = (t / (k >> 1)) * k, o = t % (k >> 1);
                    lo = base + o;
                    hi = base + k - 1 - o;
                } else {
                    lo = blk * 2 * j + off;
                    hi = lo + j;
                }
                if (hi < len && v[lo] > v[hi]) {
                    const int x = v[lo];
                    v[lo] = v[hi];
                    v[hi] = x;
                }
            }
            __syncthreads();
        }
    }
}

__global__ __launch_bounds__(256) void k_csr_sort_rows(const int *__restrict__ rowptr, int n, int *__restrict__ cols)
{
    __shared__ int sv[NB_ROW_LDS];
    const int i = blockIdx.x;
    if (i >= n) return;
    const int r0 = rowptr[i], len = rowptr[i + 1] - r0;
    if (len <= 1) return;
    if (len <= NB_ROW_LDS) {
        for (int k = threadIdx.x; k < len; k += 256) sv[k] = cols[r0 + k];
        __syncthreads();
        bitonic_row(sv, len);
        for (int k = threadIdx.x; k < len; k += 256) cols[r0 + k] = sv[k];
    } else {
        bitonic_row((volatile int *)(cols + r0), len);
    }
}

// ------------------------------------------------------------------------------------------------
// Host-side launch wrappers.
// ------------------------------------------------------------------------------------------------
static inline int nblk(size_t m, int b) { return (int)((m + b - 1) / b); }

void launch_bbox(const double *xyz, int n, double *part, int nparts, double *out, hipStream_t s)
{
    hipLaunchKernelGGL(k_bbox_partial, dim3(nparts), dim3(256), 0, s, xyz, n, part);
    hipLaunchKernelGGL(k_bbox_final, dim3(1), dim3(256), 0, s, part, nparts, out);
}

void launch_bin(const double *xyz, int n, const BuildGrid &g, int4 *cellv, int *key, int *count, hipStream_t s)
{
    if (n > 0) hipLaunchKernelGGL(k_bin, dim3(nblk(n, 256)), dim3(256), 0, s, xyz, n, g, cellv, key, count);
}

// Exclusive scan of in[0..m) into out, using tmp (>= scan_tmp_size(m) ints).
size_t scan_tmp_size(int m)
{
    size_t t = 0;
    while (m > 1024) { int nb = nblk(m, 1024); t += 2 * (size_t)nb; m = nb; }
    return t + 2;
}

void launch_scan(const int *in, int *out, int m, int *tmp, hipStream_t s)
{
    if (m <= 0) return;
    int nb = nblk(m, 1024);
    if (nb == 1) {
        hipLaunchKernelGGL(k_scan_block, dim3(1), dim3(256), 0, s, in, out, m, (int *)nullptr);
        return;
    }
    int *bsum = tmp, *boff = tmp + nb;
    hipLaunchKernelGGL(k_scan_block, dim3(nb), dim3(256), 0, s, in, out, m, bsum);
    launch_scan(bsum, boff, nb, tmp + 2 * nb, s);
    hipLaunchKernelGGL(k_scan_add, dim3(nblk(m, 256)), dim3(256), 0, s, out, m, boff);
}

void launch_fill_int(int *p, size_t m, int v, hipStream_t s)
{
    if (m > 0) hipLaunchKernelGGL(k_fill_int, dim3(nblk(m, 256)), dim3(256), 0, s, p, m, v);
}

void launch_sort_slots(const int *key, int n, int nkeys, int ncol, int nzb, int *off, int *cursor, int *unsorted,
                       int *sorted, int *padded, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_scatter, dim3(nblk(n, 256)), dim3(256), 0, s, key, n, cursor, unsorted);
    hipLaunchKernelGGL(k_bucket_rank, dim3(nblk(n, 256)), dim3(256), 0, s, key, off, nkeys, n, unsorted, sorted);
    hipLaunchKernelGGL(k_colpad, dim3(nblk(ncol, 256)), dim3(256), 0, s, off, ncol, nzb, n, padded);
}

void launch_slots(const int *sorted, const int *key, int n, int nzb, const int *off, const int *colstart,
                  const int *padded, int ncol, const double *xyz, int nsc, int *slot_atom, int *atom_slot, int2 *colcl,
                  hipStream_t s)
{
    if (n > 0) {
        hipLaunchKernelGGL(k_slots, dim3(nblk(n, 256)), dim3(256), 0, s, sorted, key, n, nzb, off, colstart, slot_atom,
                           atom_slot);
        hipLaunchKernelGGL(k_sc_split, dim3(nsc), dim3(64), 0, s, xyz, slot_atom, atom_slot);
    }
    hipLaunchKernelGGL(k_colclusters, dim3(nblk(ncol, 256)), dim3(256), 0, s, colstart, padded, ncol, colcl);
}

void launch_cluster_geom(const double *xyz, const int *slot_atom, int ncl, const double *origin, float4 *blo,
                         float4 *bhi, float4 *centre, hipStream_t s)
{
    hipLaunchKernelGGL(k_cluster_geom, dim3(nblk(ncl, 256)), dim3(256), 0, s, xyz, slot_atom, ncl, origin[0],
                       origin[1], origin[2], blo, bhi, centre);
}

void launch_slot_records(const double *xyz, const int *slot_atom, int n_pad, const int4 *cellv,
                         const unsigned char *free_flags, int use_grid, int pointcell, int u, double4 *srec,
                         int4 *scell, hipStream_t s)
{
    if (n_pad > 0)
        hipLaunchKernelGGL(k_slot_records, dim3(nblk(n_pad, 256)), dim3(256), 0, s, xyz, slot_atom, n_pad, cellv,
                           free_flags, use_grid, pointcell, u, srec, scell);
}

void launch_search(const SearchParams &P, int nsc, const double *xyz, const int *slot_atom, const float4 *blo,
                   const float4 *bhi, const int2 *colcl, Entry *scratch, int *ncand, int *counts, int *overflow,
                   hipStream_t s)
{
    if (nsc <= 0) return;
    const int wst = 2 * P.u + 1;
    const int sten = !P.use_grid ? 0 : (wst * wst * wst <= NB_STENCIL_LDS ? 1 : 2);
    if (sten != 0 && P.pointcell) {
        if (sten == 1)
            hipLaunchKernelGGL((k_search<false, 1, true>), dim3(nsc), dim3(NB_SEARCH_THREADS), 0, s, P, xyz, slot_atom,
                               blo, bhi, colcl, scratch, ncand, counts, overflow);
        else
            hipLaunchKernelGGL((k_search<false, 2, true>), dim3(nsc), dim3(NB_SEARCH_THREADS), 0, s, P, xyz, slot_atom,
                               blo, bhi, colcl, scratch, ncand, counts, overflow);
        return;
    }
    if (sten == 0)
        hipLaunchKernelGGL((k_search<false, 0>), dim3(nsc), dim3(NB_SEARCH_THREADS), 0, s, P, xyz, slot_atom, blo, bhi,
                           colcl, scratch, ncand, counts, overflow);
    else if (sten == 1)
        hipLaunchKernelGGL((k_search<false, 1>), dim3(nsc), dim3(NB_SEARCH_THREADS), 0, s, P, xyz, slot_atom, blo, bhi,
                           colcl, scratch, ncand, counts, overflow);
    else
        hipLaunchKernelGGL((k_search<false, 2>), dim3(nsc), dim3(NB_SEARCH_THREADS), 0, s, P, xyz, slot_atom, blo, bhi,
                           colcl, scratch, ncand, counts, overflow);
}

void launch_search_images(const SearchParams &P, int nsc, int nimg, const double *xyz, const int *slot_atom,
                          const float4 *blo, const float4 *bhi, Entry *scratch, int *ncand, int *counts, int *overflow,
                          hipStream_t s)
{
    if (nsc <= 0 || nimg <= 0) return;
    if (P.use_grid && P.pointcell)
        hipLaunchKernelGGL((k_search<true, 0, true>), dim3(nsc, nimg), dim3(NB_SEARCH_THREADS), 0, s, P, xyz,
                           slot_atom, blo, bhi, (const int2 *)nullptr, scratch, ncand, counts, overflow);
    else
        hipLaunchKernelGGL((k_search<true, 0>), dim3(nsc, nimg), dim3(NB_SEARCH_THREADS), 0, s, P, xyz, slot_atom, blo,
                           bhi, (const int2 *)nullptr, scratch, ncand, counts, overflow);
}

void launch_cut_items(const int *counts, const int *eoff, const int *ioff, int nblocks, int nsc, int per_item,
                      int sys, WorkItem *items, hipStream_t s, bool images)
{
    if (nblocks > 0)
        hipLaunchKernelGGL(k_cut_items, dim3(nblk(nblocks, 256)), dim3(256), 0, s, counts, eoff, ioff, nblocks, nsc,
                           per_item, sys, images ? 1 : 0, items);
}

void launch_image_transform(const double *xyz, int n, const double *rt, int ntrans, double *out, hipStream_t s)
{
    if (n > 0 && ntrans > 0)
        hipLaunchKernelGGL(k_image_transform, dim3(nblk(n, 256), ntrans), dim3(256), 0, s, xyz, n, rt, out);
}

void launch_image_coords(const double *tcoords, int n, int ntrans, const int *step_off, const ImageStep *steps,
                         double *out, hipStream_t s)
{
    if (n > 0 && ntrans > 0)
        hipLaunchKernelGGL(k_image_coords, dim3(nblk(n, 256), ntrans), dim3(256), 0, s, tcoords, n, step_off, steps,
                           out);
}

void launch_image_cells(const double *xyz, int n, const BuildGrid &g, int pointcell, int u,
                        const unsigned char *free_flags, int4 *icell, hipStream_t s)
{
    if (n > 0)
        hipLaunchKernelGGL(k_image_cells, dim3(nblk(n, 256)), dim3(256), 0, s, xyz, n, g, pointcell, u, free_flags,
                           icell);
}

void launch_image_grid(const double *ixyz, int n, int nimg, const BuildGrid &g, ImageGrid *igrid, hipStream_t s)
{
    if (n > 0 && nimg > 0) hipLaunchKernelGGL(k_image_grid, dim3(nimg), dim3(256), 0, s, ixyz, n, g, igrid);
}

void launch_export_image_pairs(const Entry *entries, int nent, const int *slot_atom, long long n, int img,
                               unsigned long long *cursor, long long cap, long long *keys, hipStream_t s)
{
    if (nent > 0)
        hipLaunchKernelGGL(k_export_image_pairs, dim3(nent), dim3(64), 0, s, entries, nent, slot_atom, n, img, cursor,
                           cap, keys);
}

void launch_order_items(const WorkItem *items, int nitems, const Entry *entries, int maxcost, int interleave, int *cost,
                        int *hist, int *cursor, int *scantmp, WorkItem *out, hipStream_t s)
{
    if (nitems <= 0) return;
    if (maxcost + 1 <= NB_ORDER_BINS) {
        hipLaunchKernelGGL(k_item_cost, dim3(nblk(nitems, 4)), dim3(256), 0, s, items, nitems, entries, maxcost, cost,
                           (int *)nullptr);
        hipLaunchKernelGGL(k_item_order, dim3(1), dim3(NB_ORDER_THREADS), 0, s, items, nitems, cost, maxcost,
                           interleave, out);
        return;
    }
    launch_fill_int(hist, (size_t)maxcost + 1, 0, s);
    hipLaunchKernelGGL(k_item_cost, dim3(nblk(nitems, 4)), dim3(256), 0, s, items, nitems, entries, maxcost, cost,
                       hist);
    launch_scan(hist, cursor, maxcost + 1, scantmp, s);
    hipLaunchKernelGGL(k_item_scatter, dim3(nblk(nitems, 256)), dim3(256), 0, s, items, nitems, cost, maxcost,
                       interleave, cursor, out);
}

bool launch_list_offsets(const int *counts, int nsc, int per_item, int *eoff, int *nits, int *ioff, const int *need,
                         int *out, hipStream_t s)
{
    if (nsc <= 0 || nsc > NB_OFFS_MAX) return false;
    hipLaunchKernelGGL(k_list_offsets, dim3(1), dim3(NB_OFFS_THREADS), 0, s, counts, nsc, per_item, eoff, nits, ioff,
                       need, out);
    return true;
}

void launch_items_per_sc(const int *counts, int nsc, int per_item, int *nitems, hipStream_t s)
{
    hipLaunchKernelGGL(k_items_per_sc, dim3(nblk(nsc, 256)), dim3(256), 0, s, counts, nsc, per_item, nitems);
}

// Sum of the per-block pair counts into *out (one workgroup).
__global__ __launch_bounds__(256) void k_sum_u64(const unsigned long long *__restrict__ v, int m,
                                                 unsigned long long *__restrict__ out)
{
    __shared__ unsigned long long sP[4];
    unsigned long long t = 0;
    for (int k = threadIdx.x; k < m; k += 256) t += v[k];
    for (int o = 32; o >= 1; o >>= 1) t += __shfl_xor(t, o, 64);
    if ((threadIdx.x & 63) == 0) sP[threadIdx.x >> 6] = t;
    __syncthreads();
    if (threadIdx.x == 0) *out = (sP[0] + sP[1]) + (sP[2] + sP[3]);
}

void launch_compact(const Entry *scratch, int cap, const int *soff, const int *ncand, const int *counts,
                    const int *eoff, const int *ioff, int nblocks, int nsc, int per_item, int sys, Entry *entries,
                    WorkItem *items, unsigned long long *npairs, unsigned long long *npairs_block, hipStream_t s)
{
    if (nblocks <= 0) return;
    hipLaunchKernelGGL(k_compact, dim3(nblocks), dim3(NB_COMPACT_THREADS), 0, s, scratch, cap, soff, ncand, counts,
                       eoff, ioff, nsc, per_item, sys, entries, items, npairs_block);
    if (npairs) hipLaunchKernelGGL(k_sum_u64, dim3(1), dim3(256), 0, s, npairs_block, nblocks, npairs);
}

void launch_image_bounds(const SearchParams &P, int nsc, int nimg, const float4 *blo, const float4 *bhi, int *bound,
                         hipStream_t s)
{
    if (nsc > 0 && nimg > 0)
        hipLaunchKernelGGL(k_image_bounds, dim3(nsc, nimg), dim3(256), 0, s, P, blo, bhi, bound);
}

void launch_add_int(int *p, int m, int v, hipStream_t s)
{
    if (m > 0 && v != 0) hipLaunchKernelGGL(k_add_int, dim3(nblk(m, 256)), dim3(256), 0, s, p, m, v);
}

void launch_export_pairs(const Entry *entries, int nent, const int *slot_atom, long long n,
                         unsigned long long *cursor, long long cap, long long *keys, hipStream_t s)
{
    if (nent > 0)
        hipLaunchKernelGGL(k_export_pairs, dim3(nent), dim3(64), 0, s, entries, nent, slot_atom, n, cursor, cap, keys);
}

void launch_csr_degree(const Entry *entries, int nent, const int *slot_atom, int *deg, hipStream_t s)
{
    if (nent > 0) hipLaunchKernelGGL(k_csr_degree, dim3(nent), dim3(64), 0, s, entries, nent, slot_atom, deg);
}

void launch_csr_fill(const Entry *entries, int nent, const int *slot_atom, int *cursor, int *cols, hipStream_t s)
{
    if (nent > 0) hipLaunchKernelGGL(k_csr_fill, dim3(nent), dim3(64), 0, s, entries, nent, slot_atom, cursor, cols);
}

void launch_csr_sort_rows(const int *rowptr, int n, int *cols, hipStream_t s)
{
    if (n > 0) hipLaunchKernelGGL(k_csr_sort_rows, dim3(n), dim3(256), 0, s, rowptr, n, cols);
}

