// nbabfs_qcmm.hip -- the QC/MM electrostatics of the ABFS model on gfx950:
// the QC/MM pair lists, the electrostatic potentials at the QC atoms and the
// gradients for given QC charges.
//
// The QC side holds a few tens to a few hundred atoms, the MM side every
// atom: each kernel gives one workgroup to one QC atom (or pure QC atom) and
// lets its 256 lanes stride over the MM atoms or over the row's list.  All
// arithmetic is FP64 in the reference's expression order (no contraction);
// potentials and QC-atom gradient sums are block reductions in a fixed order,
// MM-side gradients go through the 64-bit fixed-point accumulators, so the
// results do not depend on scheduling.
//
// Reference: NBModelABFS_QCMMPotentials / _QCMMGradients (pMolecule-1.9.0/
// extensions/csource/NBModelABFS.c:384-498) over
// PairwiseInteractionABFS_QCMMPotentials / _QCMMGradients
// (PairwiseInteraction.c:536-620), the list generation of GenerateLists /
// PairListGenerator_CrossPairListFrom*Coordinates3 (NBModelABFS.c:1050-1108,
// pCore-1.9.0/extensions/csource/PairListGenerator.c:28-85,414-493), the link
// atom rules of QCAtomContainer.c:415-432,530-559,583-630 and the cubic
// spline of pCore-1.9.0/extensions/csource/CubicSpline.c:86-126.
#include <cfloat>

#include "nbabfs_kernels.h"

// QCAtomContainer_GetAtomCoordinates3: a link atom sits on the bond from its QC
// partner p to the boundary atom b, x = x_p + d (x_b - x_p) with d the link
// factor (ratio) or the link distance over |x_b - x_p|.
__device__ __forceinline__ void qc_coord(const QcmmGeom &G, int q, double x[3])
{
#pragma clang fp contract(off)
    const int i = G.qc_index[q], p = G.qc_partner[q];
    if (p < 0) {
        x[0] = G.xyz[3 * (size_t)i];
        x[1] = G.xyz[3 * (size_t)i + 1];
        x[2] = G.xyz[3 * (size_t)i + 2];
        return;
    }
    x[0] = G.xyz[3 * (size_t)p];
    x[1] = G.xyz[3 * (size_t)p + 1];
    x[2] = G.xyz[3 * (size_t)p + 2];
    const double xij = G.xyz[3 * (size_t)i] - x[0];
    const double yij = G.xyz[3 * (size_t)i + 1] - x[1];
    const double zij = G.xyz[3 * (size_t)i + 2] - x[2];
    double d = G.qc_lf[q];
    if (!G.linkRatio) d /= sqrt(xij * xij + yij * yij + zij * zij);
    x[0] += d * xij;
    x[1] += d * yij;
    x[2] += d * zij;
}

// An MM site: an atom, or (RC/RD coupling) the midpoint of a boundary atom and
// one of its MM partners (QCAtomContainer_GetMMCoordinates3).
__device__ __forceinline__ void site_coord(const QcmmGeom &G, int j, bool atoms, double x[3])
{
#pragma clang fp contract(off)
    if (atoms || !G.site_a) {
        x[0] = G.xyz[3 * (size_t)j];
        x[1] = G.xyz[3 * (size_t)j + 1];
        x[2] = G.xyz[3 * (size_t)j + 2];
        return;
    }
    const int a = G.site_a[j], p = G.site_p[j];
    x[0] = G.xyz[3 * (size_t)a];
    x[1] = G.xyz[3 * (size_t)a + 1];
    x[2] = G.xyz[3 * (size_t)a + 2];
    if (p >= 0) {
        x[0] = 0.5 * (x[0] + G.xyz[3 * (size_t)p]);
        x[1] = 0.5 * (x[1] + G.xyz[3 * (size_t)p + 1]);
        x[2] = 0.5 * (x[2] + G.xyz[3 * (size_t)p + 2]);
    }
}

// CubicSpline_Evaluate: bisection for the interval, then the symmetric cubic
// form.  Outside [x_0, x_n-1] both values are zero, as the reference leaves them.
__device__ __forceinline__ void spline_eval(const SplineDev &S, double x, double *f, double *g)
{
#pragma clang fp contract(off)
    if (x < S.x[0] || x > S.x[S.n - 1]) {
        if (f) *f = 0.0;
        if (g) *g = 0.0;
        return;
    }
    int l = 0, u = S.n - 1;
    while (u - l > 1) {
        const int i = (u + l) >> 1;
        if (S.x[i] > x) u = i;
        else l = i;
    }
    const double xl = S.x[l], xu = S.x[u];
    const double d = xu - xl;
    const double s = (x - xl) / d;
    const double t = (xu - x) / d;
    const double hl = S.h[l] * d / 6.0;
    const double hu = S.h[u] * d / 6.0;
    const double yl = S.y[l], yu = S.y[u];
    if (f) *f = t * yl + s * yu + d * (t * (t * t - 1.0) * hl + s * (s * s - 1.0) * hu);
    if (g) *g = (yu - yl) / d + (-(3.0 * t * t - 1.0) * hl + (3.0 * s * s - 1.0) * hu);
}

// The spline's arrays staged in LDS when they fit (every pair's bisection
// then reads LDS instead of L2); otherwise the global arrays are used.
#define NB_SPLINE_LDS 1024
__device__ __forceinline__ SplineDev stage_spline(const SplineDev &S, double *lds)
{
    if (S.n > NB_SPLINE_LDS) return S;
    for (int k = threadIdx.x; k < S.n; k += blockDim.x) {
        lds[k] = S.x[k];
        lds[NB_SPLINE_LDS + k] = S.y[k];
        lds[2 * NB_SPLINE_LDS + k] = S.h[k];
    }
    __syncthreads();
    return SplineDev{lds, lds + NB_SPLINE_LDS, lds + 2 * NB_SPLINE_LDS, S.n};
}

// Binary search in a sorted range [lo, hi).
__device__ __forceinline__ bool in_sorted(const int *__restrict__ a, int lo, int hi, int v)
{
    int n = hi - lo;
    while (n > 0) {
        const int half = n >> 1;
        if (a[lo + half] < v) {
            lo += half + 1;
            n -= half + 1;
        } else {
            n = half;
        }
    }
    return lo < hi && a[lo] == v;
}

// Block-wide FP64 sums of 256 threads (four wave sums, added in wave order).
__device__ __forceinline__ void block_sum_n(double *v, int k, double *sh)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int c = 0; c < k; c++) {
        const double s = wave_sum_f64(v[c]);
        if (lane == 0) sh[4 * c + w] = s;
    }
    __syncthreads();
    for (int c = 0; c < k; c++) v[c] = ((sh[4 * c] + sh[4 * c + 1]) + sh[4 * c + 2]) + sh[4 * c + 3];
    __syncthreads();
}

// ------------------------------------------------------------------------------------------------
// Cross lists.  Workgroup (row, chunk) tests the columns (atoms or MM sites)
// of its chunk of NB_XL_CHUNK in tiles of 256 in ascending order and compacts
// the accepted ones in that order; the chunks of a row are placed in chunk
// order, so a row's list is sorted.  Pass 0 counts per (row, chunk), pass 1
// fills at off[row * chunks + chunk] (the exclusive scan of the counts).
//
// Predicate (PairListGenerator.c:28-85): the column in the AND selection
// (atom flag bit 1 clear: not a pure QC atom), the OR test on the free flags
// (bit 0) of the row and column atoms when or_test is set, not in the row's
// sorted exclusion range, and (dx*dx + dy*dy + dz*dz) <= cutoff^2 with
// d = x_row - x_column in FP64.
//
// Grid method (P.grid; PairListGenerator.c:703-794,861-940 with RegularGrid.c:
// 341-405): the column must lie on the columns' grid, in cell kc, with the
// offset kc + lowN - I from the row's search cell I in the stencil.  Cell/cell:
// the row on the rows' grid, and the stencil flag 2 (the cell pair fully in
// range) or the distance test.  Point/cell: the smallest squared distance from
// the row point to cell kc below cutoff^2, then the largest at most cutoff^2
// or the distance test; a row off its grid searches around the cell of the
// latest earlier row on it, (-u, -u, -u) before any (RegularGrid_CellIDToIndices
// writes nothing for a point with no cell; the rows are the AND selection:
// the QC atoms in order, or the pure QC atoms by atom index).
// ------------------------------------------------------------------------------------------------
// Cell of point x on a grid (RegularGrid_FindCellIDOfPoint, RegularGrid.c:240-251): floor((x - lower)/h) per axis;
// returns whether the point is on the grid.
__device__ __forceinline__ bool grid_cell(const double x[3], const double *lower, double h, const int *bins, int k[3])
{
#pragma clang fp contract(off)
    bool on = true;
    for (int d = 0; d < 3; d++) {
        k[d] = (int)floor((x[d] - lower[d]) / h);
        on = on && k[d] >= 0 && k[d] < bins[d];
    }
    return on;
}

template <int PASS>
__global__ __launch_bounds__(256) void k_cross_list(CrossListParams P)
{
#pragma clang fp contract(off)
    const int r = blockIdx.x, chunk = blockIdx.y, nchunk = gridDim.y;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const QcmmGeom &G = P.G;
    double xr[3];
    int rowflag = 1, ex0 = 0, ex1 = 0, ra = -1;
    if (P.rows_are_qc) {
        qc_coord(G, r, xr);
        if (P.ex_ptr) {
            ex0 = P.ex_ptr[r];
            ex1 = P.ex_ptr[r + 1];
        }
    } else {
        const int a = P.row_atom[r];
        ra = a;
        xr[0] = G.xyz[3 * (size_t)a];
        xr[1] = G.xyz[3 * (size_t)a + 1];
        xr[2] = G.xyz[3 * (size_t)a + 2];
        if (P.flags) rowflag = P.flags[a];
        if (P.ex_ptr) {
            ex0 = P.ex_ptr[a];
            ex1 = P.ex_ptr[a + 1];
        }
    }
    // Grid method: the row's search cell and whether the row is on its grid (every thread, the same answer).
    const QcmmGrid *Gd = P.grid;
    int rc[3] = {0, 0, 0};
    bool ron = true;
    if (Gd) {
        ron = grid_cell(xr, Gd->lower1, P.h, Gd->bins1, rc);
        if (!ron && P.pointcell) {
            rc[0] = rc[1] = rc[2] = -P.u;
            if (P.rows_are_qc) {
                for (int q = r - 1; q >= 0; q--) {
                    double xq[3];
                    int kq[3];
                    qc_coord(G, q, xq);
                    if (grid_cell(xq, Gd->lower1, P.h, Gd->bins1, kq)) {
                        rc[0] = kq[0]; rc[1] = kq[1]; rc[2] = kq[2];
                        break;
                    }
                }
            } else {
                int best = -1;
                for (int q = 0; q < P.nrows; q++) {
                    const int a = P.row_atom[q];
                    if (a >= ra || a <= best) continue;
                    const double xq[3] = {G.xyz[3 * (size_t)a], G.xyz[3 * (size_t)a + 1], G.xyz[3 * (size_t)a + 2]};
                    int kq[3];
                    if (grid_cell(xq, Gd->lower1, P.h, Gd->bins1, kq)) {
                        best = a;
                        rc[0] = kq[0]; rc[1] = kq[1]; rc[2] = kq[2];
                    }
                }
            }
        }
    }
    const int wst = 2 * P.u + 1;
    const int ncol = P.col_atoms ? G.n : G.nsites;
    const int cbeg = chunk * NB_XL_CHUNK, cend = min(ncol, cbeg + NB_XL_CHUNK);
    __shared__ int sW[4];
    __shared__ int sBase;
    if (tid == 0) sBase = PASS ? P.off[(size_t)r * nchunk + chunk] : 0;
    __syncthreads();
    for (int base = cbeg; base < cend; base += 256) {
        const int j = base + tid;
        bool ok = j < cend;
        if (ok && P.col_atoms && P.flags) {
            const int fj = P.flags[j];
            ok = !(fj & 2);
            if (ok && P.or_test) ok = ((fj | rowflag) & 1) != 0;
        }
        if (ok) {
            double xc[3];
            site_coord(G, j, P.col_atoms != 0, xc);
            const double dx = xr[0] - xc[0], dy = xr[1] - xc[1], dz = xr[2] - xc[2];
            const bool near = (dx * dx + dy * dy + dz * dz) <= P.cutoff2;
            if (Gd) {
                const int4 cc = P.colcell[j];
                const int kc[3] = {cc.x, cc.y, cc.z};
                const bool con = cc.w != 0;
                const int ox = kc[0] + Gd->lowN[0] - rc[0] + P.u, oy = kc[1] + Gd->lowN[1] - rc[1] + P.u,
                          oz = kc[2] + Gd->lowN[2] - rc[2] + P.u;
                const bool inb = ox >= 0 && ox < wst && oy >= 0 && oy < wst && oz >= 0 && oz < wst;
                const int flag = inb ? (int)P.stencil[(ox * wst + oy) * wst + oz] : 0;
                if (P.pointcell) {
                    double minR2 = 0.0, maxR2 = 0.0;
                    if (con && flag != 0)
                        point_cell_r2(Gd->lower2, P.h, xr[0], xr[1], xr[2], kc[0], kc[1], kc[2], minR2, maxR2);
                    ok = con && flag != 0 && minR2 < P.cutoff2 && (maxR2 <= P.cutoff2 || near);
                } else {
                    ok = ron && con && flag != 0 && (flag == 2 || near);
                }
            } else {
                ok = near;
            }
        }
        if (ok && ex1 > ex0) ok = !in_sorted(P.ex_idx, ex0, ex1, j);
        const unsigned long long m = __ballot(ok);
        if (lane == 0) sW[w] = __popcll(m);
        __syncthreads();
        int before = 0;
        for (int k = 0; k < w; k++) before += sW[k];
        const int total = sW[0] + sW[1] + sW[2] + sW[3];
        if (PASS && ok) P.idx[sBase + before + __popcll(m & ((1ull << lane) - 1ull))] = j;
        __syncthreads();
        if (tid == 0) sBase += total;
        __syncthreads();
    }
    if (!PASS && tid == 0) P.count[(size_t)r * nchunk + chunk] = sBase;
}

__global__ __launch_bounds__(256) void k_row_pointers(const int *__restrict__ off, int nrows, int nchunk,
                                                      int *__restrict__ ptr)
{
    const int r = blockIdx.x * 256 + threadIdx.x;
    if (r <= nrows) ptr[r] = off[(size_t)r * nchunk];
}

// ------------------------------------------------------------------------------------------------
// Potentials (atomic units): pot[q] = sum_main (eScale q_m) f(r^2) + sum_1-4
// (eScale14 q_m) f(r^2) over r^2 <= outerCutoff^2.  Each QC atom's lists are
// shared by gridDim.y workgroups (chunk c takes entries c*256+t, stride
// gridDim.y*256); chunk sums go to part[2*(q*chunks + c) + list], added in
// chunk order by the host.
// ------------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_qcmm_potentials(QcmmGeom G, SplineDev S0, QcmmRows L, const double *site_q,
                                                        double escale, double escale14, double r2off, double *part)
{
#pragma clang fp contract(off)
    __shared__ double sh[8];
    __shared__ double sS[3 * NB_SPLINE_LDS];
    const SplineDev S = stage_spline(S0, sS);
    const int q = blockIdx.x, c = blockIdx.y, stride = gridDim.y * 256;
    double xq[3];
    qc_coord(G, q, xq);
    const bool atoms = G.site_a == nullptr;
    double v[2] = {0.0, 0.0};
    for (int list = 0; list < 2; list++) {
        const int *ptr = list ? L.ptr14 : L.ptr;
        const int *idx = list ? L.idx14 : L.idx;
        if (!ptr) continue;
        const double sc = list ? escale14 : escale;
        for (int k = ptr[q] + c * 256 + threadIdx.x; k < ptr[q + 1]; k += stride) {
            const int m = idx[k];
            double xm[3];
            site_coord(G, m, atoms, xm);
            const double xqm = xq[0] - xm[0], yqm = xq[1] - xm[1], zqm = xq[2] - xm[2];
            const double r2 = xqm * xqm + yqm * yqm + zqm * zqm;
            if (r2 > r2off) continue;
            double f;
            spline_eval(S, r2, &f, nullptr);
            v[list] += (sc * site_q[m]) * f;
        }
    }
    block_sum_n(v, 2, sh);
    if (threadIdx.x == 0) {
        part[2 * ((size_t)q * gridDim.y + c)] = v[0];
        part[2 * ((size_t)q * gridDim.y + c) + 1] = v[1];
    }
}

// ------------------------------------------------------------------------------------------------
// Gradients for the QC charges: per pair v = (x_q - x_m) * 2 f'(r^2) (scale
// qc_q) q_m; the MM site gets -v (onto its atom, or half each onto the
// boundary atom and partner), the QC atom the sum of v, passed through the
// link-atom rule (QCAtomContainer_SetAtomGradients3, linear in the sum, so each
// chunk of a QC atom's lists applies it to its own sum).  All sums are fixed
// point, so the chunking does not change the result.
// ------------------------------------------------------------------------------------------------
__device__ __forceinline__ void gacc_add(long long *gacc, int n, int a, const double v[3])
{
    for (int k = 0; k < 3; k++) fx_add(&gacc[(size_t)k * n + a], &gacc[(size_t)(k + 3) * n + a], v[k]);
}

__global__ __launch_bounds__(256) void k_qcmm_gradients(QcmmGeom G, SplineDev S0, QcmmRows L, const double *site_q,
                                                       const double *qcq, double scale, double scale14, double r2off,
                                                       long long *gacc)
{
#pragma clang fp contract(off)
    __shared__ double sh[12];
    __shared__ double sS[3 * NB_SPLINE_LDS];
    const SplineDev S = stage_spline(S0, sS);
    const int q = blockIdx.x, c = blockIdx.y, stride = gridDim.y * 256;
    double xq[3];
    qc_coord(G, q, xq);
    const bool atoms = G.site_a == nullptr;
    double g[3] = {0.0, 0.0, 0.0};
    for (int list = 0; list < 2; list++) {
        const int *ptr = list ? L.ptr14 : L.ptr;
        const int *idx = list ? L.idx14 : L.idx;
        if (!ptr) continue;
        const double qq = (list ? scale14 : scale) * qcq[q];
        for (int k = ptr[q] + c * 256 + threadIdx.x; k < ptr[q + 1]; k += stride) {
            const int m = idx[k];
            const double qqm = qq * site_q[m];
            double xm[3];
            site_coord(G, m, atoms, xm);
            double xqm = xq[0] - xm[0], yqm = xq[1] - xm[1], zqm = xq[2] - xm[2];
            const double r2 = xqm * xqm + yqm * yqm + zqm * zqm;
            if (r2 > r2off) continue;
            double dF;
            spline_eval(S, r2, nullptr, &dF);
            dF *= 2.0 * qqm;
            xqm *= dF;
            yqm *= dF;
            zqm *= dF;
            if (atoms || G.site_p[m] < 0) {
                const double mv[3] = {-xqm, -yqm, -zqm};
                gacc_add(gacc, G.n, atoms ? m : G.site_a[m], mv);
            } else {
                // QCAtomContainer_SetMMGradients3: half to the boundary atom, half to the partner.
                const double hv[3] = {-(0.5 * xqm), -(0.5 * yqm), -(0.5 * zqm)};
                gacc_add(gacc, G.n, G.site_a[m], hv);
                gacc_add(gacc, G.n, G.site_p[m], hv);
            }
            g[0] += xqm;
            g[1] += yqm;
            g[2] += zqm;
        }
    }
    block_sum_n(g, 3, sh);
    if (threadIdx.x != 0) return;
    const int i = G.qc_index[q], p = G.qc_partner[q];
    if (p < 0) {
        gacc_add(gacc, G.n, i, g);
        return;
    }
    gacc_add(gacc, G.n, p, g);
    double t[3];
    if (G.linkRatio) {
        for (int k = 0; k < 3; k++) t[k] = g[k] * G.qc_lf[q];
    } else {
        double u[3];
        for (int k = 0; k < 3; k++) u[k] = G.xyz[3 * (size_t)i + k] - G.xyz[3 * (size_t)p + k];
        const double d = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
        for (int k = 0; k < 3; k++) u[k] /= d;
        for (int k = 0; k < 3; k++) t[k] = g[k] * (G.qc_lf[q] / d);
        const double dg = u[0] * t[0] + u[1] * t[1] + u[2] * t[2];
        for (int k = 0; k < 3; k++) t[k] -= dg * u[k];
    }
    gacc_add(gacc, G.n, i, t);
    const double mt[3] = {-t[0], -t[1], -t[2]};
    gacc_add(gacc, G.n, p, mt);
}

// add = 0: grad = the accumulated gradient; 1: grad += it.
__global__ __launch_bounds__(256) void k_qcmm_grad_out(const long long *__restrict__ gacc, int n,
                                                      double *__restrict__ grad, int add)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    for (int k = 0; k < 3; k++) {
        const double g = (double)gacc[(size_t)(k + 3) * n + i] * NB_FX_HI_UNIT +
                         (double)gacc[(size_t)k * n + i] * (1.0 / NB_FX_LO_SCALE);
        grad[3 * (size_t)i + k] = add ? grad[3 * (size_t)i + k] + g : g;
    }
}

// The potentials from the per-(QC atom, chunk) partials, in the host call's
// order: the chunk sums of the main list, then of the 1-4 list, each added in
// turn to pot[q].
__global__ __launch_bounds__(256) void k_qcmm_pot_sum(const double *__restrict__ part, int nqc, int chunks,
                                                     double *__restrict__ pot)
{
#pragma clang fp contract(off)
    const int q = blockIdx.x * 256 + threadIdx.x;
    if (q >= nqc) return;
    double m = 0.0, m14 = 0.0;
    for (int c = 0; c < chunks; c++) {
        m += part[2 * ((size_t)q * chunks + c)];
        m14 += part[2 * ((size_t)q * chunks + c) + 1];
    }
    double p = pot[q];
    p += m;
    p += m14;
    pot[q] = p;
}

__global__ __launch_bounds__(256) void k_qcmm_coords(QcmmGeom G, double *qcxyz, double *sitexyz)
{
    const int k = blockIdx.x * 256 + threadIdx.x;
    double x[3];
    if (k < G.nqc) {
        qc_coord(G, k, x);
        for (int c = 0; c < 3; c++) qcxyz[3 * (size_t)k + c] = x[c];
    }
    if (k < G.nsites) {
        site_coord(G, k, false, x);
        for (int c = 0; c < 3; c++) sitexyz[3 * (size_t)k + c] = x[c];
    }
}

// ------------------------------------------------------------------------------------------------
// Launchers.
// ------------------------------------------------------------------------------------------------
// Min and max per axis over the workgroup's threads (every thread gets the result).
__device__ __forceinline__ void block_box(double *mn, double *mx, double (*smn)[256], double (*smx)[256])
{
    const int t = threadIdx.x;
    for (int d = 0; d < 3; d++) {
        smn[d][t] = mn[d];
        smx[d][t] = mx[d];
    }
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (t < o)
            for (int d = 0; d < 3; d++) {
                smn[d][t] = fmin(smn[d][t], smn[d][t + o]);
                smx[d][t] = fmax(smx[d][t], smx[d][t + o]);
            }
        __syncthreads();
    }
    for (int d = 0; d < 3; d++) {
        mn[d] = smn[d][0];
        mx[d] = smx[d][0];
    }
    __syncthreads();
}

// Per-workgroup boxes (min x, y, z, max x, y, z) of the QC coordinates (part[12 b + 0..5]) and of the electrostatic
// list's columns (part[12 b + 6..11]: the MM sites, or the MM selection's atoms with MM coupling).
__global__ __launch_bounds__(256) void k_qcmm_box_parts(QcmmGeom G, const unsigned char *__restrict__ flags,
                                                        double *__restrict__ part)
{
    __shared__ double smn[3][256], smx[3][256];
    const int t = threadIdx.x, b = blockIdx.x, stride = gridDim.x * 256;
    double mn[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, mx[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    for (int q = b * 256 + t; q < G.nqc; q += stride) {
        double x[3];
        qc_coord(G, q, x);
        for (int d = 0; d < 3; d++) {
            mn[d] = fmin(mn[d], x[d]);
            mx[d] = fmax(mx[d], x[d]);
        }
    }
    block_box(mn, mx, smn, smx);
    if (t == 0)
        for (int d = 0; d < 3; d++) {
            part[12 * b + d] = mn[d];
            part[12 * b + 3 + d] = mx[d];
        }
    const bool atoms = G.site_a == nullptr;
    const int ncol = atoms ? G.n : G.nsites;
    for (int d = 0; d < 3; d++) {
        mn[d] = DBL_MAX;
        mx[d] = -DBL_MAX;
    }
    for (int j = b * 256 + t; j < ncol; j += stride) {
        if (atoms && flags && (flags[j] & 2)) continue;
        double x[3];
        site_coord(G, j, atoms, x);
        for (int d = 0; d < 3; d++) {
            mn[d] = fmin(mn[d], x[d]);
            mx[d] = fmax(mx[d], x[d]);
        }
    }
    block_box(mn, mx, smn, smx);
    if (t == 0)
        for (int d = 0; d < 3; d++) {
            part[12 * b + 6 + d] = mn[d];
            part[12 * b + 9 + d] = mx[d];
        }
}

// The two grids of the QC/MM cross lists from the boxes (one workgroup; min and max are exact, so the boxes do not
// depend on the partition).  Electrostatic list: the QC coordinates' grid (Coordinates3_MakeGrid,
// pCore-1.9.0/extensions/csource/Coordinates3.c:1111-1158, no selection) for the rows and, for the columns, the grid
// conforming to it over the box of the MM selection's atoms (MM coupling: the atoms are the columns, the box leaves
// out the pure QC atoms) or of the MM sites (Coordinates3_MakeConformingGrid, Coordinates3.c:998-1072).  LJ list:
// the main list's grid for both.
__global__ __launch_bounds__(256) void k_qcmm_grids(const double *__restrict__ part, int nparts, double lx,
                                                    double ly, double lz, int bx, int by, int bz, double h,
                                                    QcmmGrid *__restrict__ out)
{
#pragma clang fp contract(off)
    __shared__ double smn[3][256], smx[3][256];
    const int t = threadIdx.x;
    double mn[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, mx[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    for (int b = t; b < nparts; b += 256)
        for (int d = 0; d < 3; d++) {
            mn[d] = fmin(mn[d], part[12 * b + d]);
            mx[d] = fmax(mx[d], part[12 * b + 3 + d]);
        }
    block_box(mn, mx, smn, smx);
    QcmmGrid E;
    E.bad = 0;
    for (int d = 0; d < 3; d++) {
        const double extent = mx[d] - mn[d];
        const double upper = mn[d] + extent;
        const double mid = 0.5 * (mn[d] + upper);
        const double cells = ceil(extent / h);
        E.bins1[d] = (int)cells;
        E.lower1[d] = mid - 0.5 * (cells * h);
        if (E.bins1[d] <= 0) E.bad = 1;
    }
    for (int d = 0; d < 3; d++) {
        mn[d] = DBL_MAX;
        mx[d] = -DBL_MAX;
    }
    for (int b = t; b < nparts; b += 256)
        for (int d = 0; d < 3; d++) {
            mn[d] = fmin(mn[d], part[12 * b + 6 + d]);
            mx[d] = fmax(mx[d], part[12 * b + 9 + d]);
        }
    block_box(mn, mx, smn, smx);
    for (int d = 0; d < 3; d++) {
        const double lo = mn[d];
        const double up = (mx[d] - lo) + lo;
        const int lowN = (int)floor((lo - E.lower1[d]) / h);
        const int upN = (int)floor((up - E.lower1[d]) / h);
        E.lowN[d] = lowN;
        E.bins2[d] = upN - lowN + 1;
        E.lower2[d] = E.lower1[d] + (double)lowN * h;
    }
    if (t == 0) {
        out[0] = E;
        QcmmGrid L;
        const double lw[3] = {lx, ly, lz};
        const int bn[3] = {bx, by, bz};
        for (int d = 0; d < 3; d++) {
            L.lower1[d] = L.lower2[d] = lw[d];
            L.bins1[d] = L.bins2[d] = bn[d];
            L.lowN[d] = 0;
        }
        L.bad = 0;
        out[1] = L;
    }
}

__global__ __launch_bounds__(256) void k_qcmm_colcells(QcmmGeom G, const QcmmGrid *__restrict__ grid, double h,
                                                       int4 *__restrict__ out)
{
    const bool atoms = G.site_a == nullptr;
    const int ncol = atoms ? G.n : G.nsites;
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= ncol) return;
    double x[3];
    site_coord(G, j, atoms, x);
    int k[3];
    const bool on = grid_cell(x, grid->lower2, h, grid->bins2, k);
    out[j] = make_int4(k[0], k[1], k[2], on ? 1 : 0);
}

void launch_qcmm_colcells(const QcmmGeom &G, const QcmmGrid *grid, double h, int4 *out, hipStream_t s)
{
    const int ncol = G.site_a == nullptr ? G.n : G.nsites;
    if (ncol > 0) hipLaunchKernelGGL(k_qcmm_colcells, dim3((ncol + 255) / 256), dim3(256), 0, s, G, grid, h, out);
}

void launch_qcmm_grids(const QcmmGeom &G, const unsigned char *flags, const double *lower, const int *bins, double h,
                       double *part, QcmmGrid *out, hipStream_t s)
{
    const int ncol = G.site_a == nullptr ? G.n : G.nsites;
    const int nparts = std::max(1, std::min(NB_QGRID_PARTS, (std::max(ncol, G.nqc) + 1023) / 1024));
    hipLaunchKernelGGL(k_qcmm_box_parts, dim3(nparts), dim3(256), 0, s, G, flags, part);
    hipLaunchKernelGGL(k_qcmm_grids, dim3(1), dim3(256), 0, s, part, nparts, lower[0], lower[1], lower[2], bins[0],
                       bins[1], bins[2], h, out);
}

void launch_cross_list(const CrossListParams &P, int pass, hipStream_t s)
{
    if (P.nrows <= 0) return;
    const dim3 grid(P.nrows, cross_list_chunks(P));
    if (pass == 0) hipLaunchKernelGGL(k_cross_list<0>, grid, dim3(256), 0, s, P);
    else hipLaunchKernelGGL(k_cross_list<1>, grid, dim3(256), 0, s, P);
}

void launch_row_pointers(const int *off, int nrows, int nchunk, int *ptr, hipStream_t s)
{
    hipLaunchKernelGGL(k_row_pointers, dim3((nrows + 1 + 255) / 256), dim3(256), 0, s, off, nrows, nchunk, ptr);
}

void launch_qcmm_potentials(const QcmmGeom &G, const SplineDev &S, const QcmmRows &L, const double *site_q,
                            double escale, double escale14, double r2off, int chunks, double *part, hipStream_t s)
{
    if (G.nqc > 0)
        hipLaunchKernelGGL(k_qcmm_potentials, dim3(G.nqc, chunks), dim3(256), 0, s, G, S, L, site_q, escale, escale14,
                           r2off, part);
}

void launch_qcmm_gradients(const QcmmGeom &G, const SplineDev &S, const QcmmRows &L, const double *site_q,
                           const double *qcq, double scale, double scale14, double r2off, int chunks, long long *gacc,
                           hipStream_t s)
{
    if (G.nqc > 0)
        hipLaunchKernelGGL(k_qcmm_gradients, dim3(G.nqc, chunks), dim3(256), 0, s, G, S, L, site_q, qcq, scale,
                           scale14, r2off, gacc);
}

void launch_qcmm_grad_out(const long long *gacc, int n, double *grad, bool add, hipStream_t s)
{
    if (n > 0) hipLaunchKernelGGL(k_qcmm_grad_out, dim3((n + 255) / 256), dim3(256), 0, s, gacc, n, grad, add ? 1 : 0);
}

void launch_qcmm_pot_sum(const double *part, int nqc, int chunks, double *pot, hipStream_t s)
{
    if (nqc > 0) hipLaunchKernelGGL(k_qcmm_pot_sum, dim3((nqc + 255) / 256), dim3(256), 0, s, part, nqc, chunks, pot);
}

void launch_qcmm_coords(const QcmmGeom &G, double *qcxyz, double *sitexyz, hipStream_t s)
{
    const int m = G.nqc > G.nsites ? G.nqc : G.nsites;
    if (m > 0) hipLaunchKernelGGL(k_qcmm_coords, dim3((m + 255) / 256), dim3(256), 0, s, G, qcxyz, sitexyz);
}
