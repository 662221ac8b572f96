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
// Cross lists.  One workgroup per row; the columns (atoms or MM sites) are
// tested in tiles of 256 in ascending order and the accepted ones are
// compacted in that order, so a row's list is sorted.  Pass 0 counts, pass 1
// fills at ptr[row].
//
// Predicate (PairListGenerator.c:28-85): the column in the AND selection
// (atom flag bit 1 clear: not a pure QC atom), the OR test on the free flags
// (bit 0) of the row and column atoms when or_test is set, not in the row's
// sorted exclusion range, and (dx*dx + dy*dy + dz*dz) <= cutoff^2 with
// d = x_row - x_column in FP64.
// ------------------------------------------------------------------------------------------------
template <int PASS>
__global__ __launch_bounds__(256) void k_cross_list(CrossListParams P)
{
#pragma clang fp contract(off)
    const int r = blockIdx.x;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const QcmmGeom &G = P.G;
    double xr[3];
    int rowflag = 1, ex0 = 0, ex1 = 0;
    if (P.rows_are_qc) {
        qc_coord(G, r, xr);
        if (P.ex_ptr) {
            ex0 = P.ex_ptr[r];
            ex1 = P.ex_ptr[r + 1];
        }
    } else {
        const int a = P.row_atom[r];
        xr[0] = G.xyz[3 * (size_t)a];
        xr[1] = G.xyz[3 * (size_t)a + 1];
        xr[2] = G.xyz[3 * (size_t)a + 2];
        if (P.flags) rowflag = P.flags[a];
        if (P.ex_ptr) {
            ex0 = P.ex_ptr[a];
            ex1 = P.ex_ptr[a + 1];
        }
    }
    const int ncol = P.col_atoms ? G.n : G.nsites;
    __shared__ int sW[4];
    __shared__ int sBase;
    if (tid == 0) sBase = PASS ? P.ptr[r] : 0;
    __syncthreads();
    for (int base = 0; base < ncol; base += 256) {
        const int j = base + tid;
        bool ok = j < ncol;
        if (ok && P.col_atoms && P.flags) {
            const int fj = P.flags[j];
            ok = !(fj & 2);
            if (ok && P.or_test) ok = ((fj | rowflag) & 1) != 0;
        }
        if (ok) {
            double xc[3];
            site_coord(G, j, P.col_atoms != 0, xc);
            const double dx = xr[0] - xc[0], dy = xr[1] - xc[1], dz = xr[2] - xc[2];
            ok = (dx * dx + dy * dy + dz * dz) <= P.cutoff2;
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
    if (!PASS && tid == 0) P.count[r] = sBase;
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
void launch_cross_list(const CrossListParams &P, int pass, hipStream_t s)
{
    if (P.nrows <= 0) return;
    if (pass == 0) hipLaunchKernelGGL(k_cross_list<0>, dim3(P.nrows), dim3(256), 0, s, P);
    else hipLaunchKernelGGL(k_cross_list<1>, dim3(P.nrows), dim3(256), 0, s, P);
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
