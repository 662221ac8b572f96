// nbabfs_build.hip -- pair-list construction on gfx950.
//
// Replaces PairListGenerator_SelfPairListFromCoordinates3 (CellCell and
// Direct, pCore-1.9.0/extensions/csource/PairListGenerator.c:946-1150) and
// RegularGridOccupancy_Fill (pCore-1.9.0/extensions/csource/
// RegularGridOccupancy.c:77-147).  The pair SET is the reference's exactly:
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

__global__ __launch_bounds__(64) void k_bbox_final(const double *__restrict__ part, int nblocks, double *__restrict__ out)
{
    if (threadIdx.x < 6) {
        int d = threadIdx.x;
        double v = part[d];
        for (int b = 1; b < nblocks; b++) v = (d < 3) ? fmin(v, part[6 * b + d]) : fmax(v, part[6 * b + d]);
        out[d] = v;
    }
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
                                                 int *__restrict__ sorted)
{
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    int pos = atomicAdd(&cursor[key[i]], 1);
    sorted[pos] = i;
}

__global__ __launch_bounds__(256) void k_bucket_sort(const int *__restrict__ off, int nkeys, int n, int *__restrict__ sorted)
{
    int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= nkeys) return;
    int b = off[k], e = (k + 1 < nkeys) ? off[k + 1] : n;
    for (int p = b + 1; p < e; p++) {
        int v = sorted[p], q = p - 1;
        while (q >= b && sorted[q] > v) { sorted[q + 1] = sorted[q]; q--; }
        sorted[q + 1] = v;
    }
}

// Column sizes padded to whole clusters.
__global__ __launch_bounds__(256) void k_colpad(const int *__restrict__ off, int ncol, int nzb, int n,
                                                int *__restrict__ padded)
{
    int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= ncol) return;
    int b = off[c * nzb];
    int e = (c + 1 < ncol) ? off[(c + 1) * nzb] : n;
    int cnt = e - b;
    padded[c] = (cnt + NB_CL - 1) / NB_CL * NB_CL;
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

// ------------------------------------------------------------------------------------------------
// The reference pair predicate for atoms i, j (PairListGenerator.c:103-143,
// 1020-1150): not excluded, OR selection (free atoms), and either the grid
// rule (both on grid; stencil flag 2 = whole cell pair within the cutoff,
// 1 = FP64 distance test, 0 = out of range) or, for the direct method, the
// FP64 distance test alone.
// ------------------------------------------------------------------------------------------------
__device__ __forceinline__ bool ref_pair_rule(int i, int j, const double *__restrict__ xyz, const int4 *__restrict__ cellv,
                                              const SearchParams &P)
{
#pragma clang fp contract(off)
    if (P.free_flags && !(P.free_flags[i] | P.free_flags[j])) return false;
    int flag = 1;
    if (P.use_grid) {
        int4 ci = cellv[i], cj = cellv[j];
        if (!(ci.w & cj.w)) return false;
        int dx = cj.x - ci.x, dy = cj.y - ci.y, dz = cj.z - ci.z;
        if (dx < -P.u || dx > P.u || dy < -P.u || dy > P.u || dz < -P.u || dz > P.u) return false;
        int w = 2 * P.u + 1;
        flag = P.stencil[((dx + P.u) * w + (dy + P.u)) * w + (dz + P.u)];
        if (flag == 0) return false;
    }
    if (flag == 1) {
        double xij = xyz[3 * (size_t)i] - xyz[3 * (size_t)j];
        double yij = xyz[3 * (size_t)i + 1] - xyz[3 * (size_t)j + 1];
        double zij = xyz[3 * (size_t)i + 2] - xyz[3 * (size_t)j + 2];
        double r2 = ((xij * xij) + (yij * yij)) + (zij * zij);
        if (!(r2 <= P.cutoff2)) return false;
    }
    for (int x = P.xptr[i]; x < P.xptr[i + 1]; x++)
        if (P.xidx[x] == j) return false;
    return true;
}

__device__ __forceinline__ float box_gap2(float4 alo, float4 ahi, float4 blo, float4 bhi)
{
    float gx = fmaxf(0.0f, fmaxf(alo.x - bhi.x, blo.x - ahi.x));
    float gy = fmaxf(0.0f, fmaxf(alo.y - bhi.y, blo.y - ahi.y));
    float gz = fmaxf(0.0f, fmaxf(alo.z - bhi.z, blo.z - ahi.z));
    return gx * gx + gy * gy + gz * gz;
}

// ------------------------------------------------------------------------------------------------
// Cluster-pair search: one wavefront per super-cluster.  Candidate j-clusters
// come from the x-y columns within reach, in a fixed order, so the entries of
// each super-cluster are produced deterministically.
// ------------------------------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_search(SearchParams P, const double *__restrict__ xyz,
                                               const int4 *__restrict__ cellv, const int *__restrict__ slot_atom,
                                               const float4 *__restrict__ blo, const float4 *__restrict__ bhi,
                                               const int2 *__restrict__ colcl, Entry *__restrict__ scratch,
                                               int *__restrict__ counts, int *__restrict__ overflow)
{
    const int sc = blockIdx.x;
    const int lane = threadIdx.x;
    const int a = lane >> 3, b = lane & 7;
    const int I0 = sc * NB_SCC;
    __shared__ float4 sIlo[NB_SCC], sIhi[NB_SCC];
    __shared__ int sIatom[NB_SC];
    if (lane < NB_SCC) {
        sIlo[lane] = blo[I0 + lane];
        sIhi[lane] = bhi[I0 + lane];
    }
    sIatom[lane] = slot_atom[I0 * NB_CL + lane];
    __syncthreads();
    float4 lo = sIlo[0], hi = sIhi[0];
    for (int c = 1; c < NB_SCC; c++) {
        lo.x = fminf(lo.x, sIlo[c].x); lo.y = fminf(lo.y, sIlo[c].y); lo.z = fminf(lo.z, sIlo[c].z);
        hi.x = fmaxf(hi.x, sIhi[c].x); hi.y = fmaxf(hi.y, sIhi[c].y); hi.z = fmaxf(hi.z, sIhi[c].z);
    }
    int cnt = 0;
    if (lo.x <= hi.x) {   // super-cluster holds at least one atom
        const float reach = P.reach;
        const float reach2 = reach * reach;
        int cx0 = (int)floorf((lo.x - reach) / P.cw), cx1 = (int)floorf((hi.x + reach) / P.cw);
        int cy0 = (int)floorf((lo.y - reach) / P.cw), cy1 = (int)floorf((hi.y + reach) / P.cw);
        cx0 = max(cx0, 0); cy0 = max(cy0, 0);
        cx1 = min(cx1, P.ncx - 1); cy1 = min(cy1, P.ncy - 1);
        const int Jmax = I0 + NB_SCC - 1;
        for (int cx = cx0; cx <= cx1; cx++) {
            for (int cy = cy0; cy <= cy1; cy++) {
                int2 r = colcl[cx * P.ncy + cy];
                int e = min(r.y, Jmax + 1);
                for (int base = r.x; base < e; base += 64) {
                    int J = base + lane;
                    int cm = 0;
                    if (J < e) {
                        float4 jl = blo[J], jh = bhi[J];
                        if (box_gap2(lo, hi, jl, jh) < reach2) {
                            for (int c = 0; c < NB_SCC; c++)
                                if (J <= I0 + c && box_gap2(sIlo[c], sIhi[c], jl, jh) < reach2) cm |= 1 << c;
                        }
                    }
                    unsigned long long cand = __ballot(cm != 0);
                    while (cand) {
                        int l = __ffsll((long long)cand) - 1;
                        cand &= cand - 1;
                        int Jc = __shfl(J, l, 64);
                        int cmc = __shfl(cm, l, 64);
                        int j = slot_atom[Jc * NB_CL + b];
                        int lbits = 0, cbits = 0;
                        for (int c = 0; c < NB_SCC; c++) {
                            bool ok = false;
                            if (cmc & (1 << c)) {
                                int i = sIatom[c * NB_CL + a];
                                ok = (i >= 0) && (j >= 0) && ((Jc < I0 + c) || (b < a)) && ref_pair_rule(i, j, xyz, cellv, P);
                            }
                            if (ok) lbits |= 1 << c;
                            if (__ballot(ok)) cbits |= 1 << c;
                        }
                        if (cbits) {
                            if (cnt < P.cap) {
                                Entry *en = &scratch[(size_t)sc * P.cap + cnt];
                                en->bits[lane] = (unsigned char)lbits;
                                if (lane == 0) {
                                    en->J = Jc;
                                    en->sc = sc;
                                    en->cbits = cbits;
                                    en->pad = 0;
                                }
                            }
                            cnt++;
                        }
                    }
                }
            }
        }
    }
    if (lane == 0) {
        counts[sc] = cnt;
        if (cnt > P.cap) atomicMax(overflow, cnt);
    }
}

// Compact per-super-cluster entries and cut them into work items.
__global__ __launch_bounds__(64) void k_compact(const Entry *__restrict__ scratch, int cap, const int *__restrict__ counts,
                                                const int *__restrict__ eoff, const int *__restrict__ ioff, int per_item,
                                                int sys, Entry *__restrict__ entries, WorkItem *__restrict__ items)
{
    int sc = blockIdx.x;
    int cnt = counts[sc], e0 = eoff[sc], i0 = ioff[sc];
    for (int k = threadIdx.x; k < cnt; k += 64) entries[e0 + k] = scratch[(size_t)sc * cap + k];
    int nit = (cnt + per_item - 1) / per_item;
    for (int k = threadIdx.x; k < nit; k += 64) {
        WorkItem w;
        w.sys = sys;
        w.sc = sc;
        w.first = e0 + k * per_item;
        w.count = min(per_item, cnt - k * per_item);
        items[i0 + k] = w;
    }
}

__global__ __launch_bounds__(256) void k_items_per_sc(const int *__restrict__ counts, int nsc, int per_item,
                                                      int *__restrict__ nitems)
{
    int s = blockIdx.x * 256 + threadIdx.x;
    if (s < nsc) nitems[s] = (counts[s] + per_item - 1) / per_item;
}

// Pair count of the list (popcount of every mask).
__global__ __launch_bounds__(256) void k_count_pairs(const Entry *__restrict__ entries, int nent,
                                                     unsigned long long *__restrict__ total)
{
    unsigned long long s = 0;
    for (int e = blockIdx.x * 256 + threadIdx.x; e < nent; e += gridDim.x * 256) {
        const unsigned long long *w = (const unsigned long long *)entries[e].bits;
        for (int k = 0; k < NB_SC / 8; k++) s += __popcll(w[k]);
    }
    for (int o = 32; o >= 1; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((threadIdx.x & 63) == 0 && s) atomicAdd(total, s);
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
    hipLaunchKernelGGL(k_bbox_final, dim3(1), dim3(64), 0, s, part, nparts, out);
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

void launch_sort_slots(const int *key, int n, int nkeys, int ncol, int nzb, int *off, int *cursor, int *sorted,
                       int *padded, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_scatter, dim3(nblk(n, 256)), dim3(256), 0, s, key, n, cursor, sorted);
    hipLaunchKernelGGL(k_bucket_sort, dim3(nblk(nkeys, 256)), dim3(256), 0, s, off, nkeys, n, sorted);
    hipLaunchKernelGGL(k_colpad, dim3(nblk(ncol, 256)), dim3(256), 0, s, off, ncol, nzb, n, padded);
}

void launch_slots(const int *sorted, const int *key, int n, int nzb, const int *off, const int *colstart,
                  const int *padded, int ncol, int *slot_atom, int *atom_slot, int2 *colcl, hipStream_t s)
{
    if (n > 0)
        hipLaunchKernelGGL(k_slots, dim3(nblk(n, 256)), dim3(256), 0, s, sorted, key, n, nzb, off, colstart, slot_atom,
                           atom_slot);
    hipLaunchKernelGGL(k_colclusters, dim3(nblk(ncol, 256)), dim3(256), 0, s, colstart, padded, ncol, colcl);
}

void launch_cluster_geom(const double *xyz, const int *slot_atom, int ncl, const double *origin, float4 *blo,
                         float4 *bhi, float4 *centre, hipStream_t s)
{
    hipLaunchKernelGGL(k_cluster_geom, dim3(nblk(ncl, 256)), dim3(256), 0, s, xyz, slot_atom, ncl, origin[0],
                       origin[1], origin[2], blo, bhi, centre);
}

void launch_search(const SearchParams &P, int nsc, const double *xyz, const int4 *cellv, const int *slot_atom,
                   const float4 *blo, const float4 *bhi, const int2 *colcl, Entry *scratch, int *counts,
                   int *overflow, hipStream_t s)
{
    if (nsc > 0)
        hipLaunchKernelGGL(k_search, dim3(nsc), dim3(64), 0, s, P, xyz, cellv, slot_atom, blo, bhi, colcl, scratch,
                           counts, overflow);
}

void launch_items_per_sc(const int *counts, int nsc, int per_item, int *nitems, hipStream_t s)
{
    hipLaunchKernelGGL(k_items_per_sc, dim3(nblk(nsc, 256)), dim3(256), 0, s, counts, nsc, per_item, nitems);
}

void launch_compact(const Entry *scratch, int cap, const int *counts, const int *eoff, const int *ioff, int nsc,
                    int per_item, int sys, Entry *entries, WorkItem *items, hipStream_t s)
{
    if (nsc > 0)
        hipLaunchKernelGGL(k_compact, dim3(nsc), dim3(64), 0, s, scratch, cap, counts, eoff, ioff, per_item, sys,
                           entries, items);
}

void launch_count_pairs(const Entry *entries, int nent, unsigned long long *total, hipStream_t s)
{
    if (nent > 0)
        hipLaunchKernelGGL(k_count_pairs, dim3(min(nblk(nent, 256), 1024)), dim3(256), 0, s, entries, nent, total);
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

