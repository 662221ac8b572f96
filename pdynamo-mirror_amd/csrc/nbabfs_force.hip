// nbabfs_force.hip -- per-call kernels of the ABFS energy/gradient on gfx950.
//
// Replaces PairwiseInteractionABFS_MMMMEnergy (analytic branch,
// pMolecule-1.9.0/extensions/csource/PairwiseInteraction.c:292-429) and the
// OpenMP gradient reduction (NBModelABFSState.c:19-38).
//
// Tiled (default, "mixed") path: one wavefront per work item (a super-cluster
// of 8 i-clusters against a run of j-clusters).  Lane l = 8a + b pairs i-atom
// a of cluster c with j-atom b.  1/r comes from the FP32 v_rsq seed refined by
// one FP64 Newton step; force vectors are summed in FP32 registers; i-forces
// stay in registers for the whole item, j-forces are reduced across the 8
// lanes that share b and added once per j-cluster.  All gradient sums land in
// 64-bit fixed point, so the result does not depend on the order in which
// items finish.
#include "nbabfs_internal.h"
#include "nbabfs_kernels.h"

#include <algorithm>

static inline int nblk(size_t m, int b) { return (int)((m + b - 1) / b); }

// ------------------------------------------------------------------------------------------------
// Per-pair terms.
// ------------------------------------------------------------------------------------------------
// FP64, reference expression order (pMolecule-1.9.0/extensions/cinclude/PairwiseInteraction.h:72-119).
__device__ __forceinline__ bool pair_fp64(const AbfsFactors &F, double r2, double qij, double A, double B,
                                          double &eel, double &elj, double &dF)
{
#pragma clang fp contract(off)
    if (r2 > F.r2Off) return false;
    double s, s2;
    if (r2 < F.r2Damp) { s2 = 0.0; s = 0.0; }
    else { s2 = 1.0 / r2; s = sqrt(s2); }
    double d = 0.0;
    if (r2 > F.r2On) {
        eel = qij * (s * (F.a - r2 * (F.b + r2 * (F.c + F.d * r2))) + F.qShift2);
        d += -qij * 0.5 * s * (F.a + r2 * (F.b + r2 * (3.0 * F.c + 5.0 * F.d * r2))) / r2;
    } else if (r2 > F.r2Damp) {
        eel = qij * (s + F.qShift1);
        d += -qij * 0.5 * s / r2;
    } else {
        eel = qij * (F.qF0 - F.qAlpha * r2);
        d += -qij * F.qAlpha;
    }
    double s6 = s2 * s2 * s2;
    if (r2 > F.r2On) {
        double l1 = s6 - F.aF6, l2 = (s / r2) - F.bF3;
        elj = A * F.aK12 * (l1 * l1) - B * F.bK6 * (l2 * l2);
        d += -3.0 * s6 * (2.0 * A * F.aK12 * l1 / r2 - B * F.bK6 * l2 / s);
    } else if (r2 > F.r2Damp) {
        elj = A * (s6 * s6 - F.aShift12) - B * (s6 - F.bShift6);
        d += -3.0 * s6 * (2.0 * A * s6 - B) / r2;
    } else {
        elj = A * (F.aF0 - F.aAlpha * r2) - B * (F.bF0 - F.bAlpha * r2);
        d += -A * F.aAlpha + B * F.bAlpha;
    }
    dF = d;
    return true;
}

// Tiled path: one pair with 1/r from the FP32 v_rsq seed refined by one
// Newton step in FP64 (relative error ~1e-14; no FP64 divide or square root),
// the three regions evaluated branch-free and selected per lane, the damped
// region (r < dampingCutoff) as a wave-uniform rare path.  On gfx950 FP64
// FMA issues at the FP32 rate, so the scalar core stays in FP64 and only the
// rsqrt seed and the force vectors are FP32.
__device__ __forceinline__ void pair_tiled(const AbfsFactors &F, double r2, double qq, double A, double B, bool act,
                                           double &fs, double &eel, double &elj)
{
    const bool sw = r2 > F.r2On;
    const double y = (double)__builtin_amdgcn_rsqf((float)r2);
    const double h = fma(-(r2 * y), y, 1.0);
    const double s = fma(0.5 * y, h, y);
    const double s2 = s * s;
    const double s3 = s2 * s;
    const double s6 = s3 * s3;
    const double P = sw ? fma(-r2, fma(r2, fma(F.d, r2, F.c), F.b), F.a) : 1.0;
    const double Q = sw ? fma(r2, fma(r2, fma(5.0 * F.d, r2, 3.0 * F.c), F.b), F.a) : 1.0;
    const double K = sw ? F.qShift2 : F.qShift1;
    double e1 = qq * fma(s, P, K);
    double d1 = -0.5 * qq * s3 * Q;
    const double l1 = s6 - F.aF6, l2 = s3 - F.bF3;
    const double Aa = A * F.aK12 * l1;
    const double Bb = B * F.bK6 * l2;
    const double e_sw = fma(Aa, l1, -Bb * l2);
    const double d_sw = -3.0 * s6 * fma(2.0 * Aa, s2, -Bb * (r2 * s));
    const double e_pl = fma(A, s6 * s6 - F.aShift12, -B * (s6 - F.bShift6));
    const double d_pl = -3.0 * s6 * s2 * fma(2.0 * A, s6, -B);
    double e2 = sw ? e_sw : e_pl;
    double d2 = sw ? d_sw : d_pl;
    const bool dmp = r2 < F.r2Damp;
    if (__any(act && dmp)) {
        if (dmp) {
            e1 = qq * fma(-F.qAlpha, r2, F.qF0);
            d1 = -qq * F.qAlpha;
            e2 = A * fma(-F.aAlpha, r2, F.aF0) - B * fma(-F.bAlpha, r2, F.bF0);
            d2 = -A * F.aAlpha + B * F.bAlpha;
        }
    }
    fs = act ? 2.0 * (d1 + d2) : 0.0;
    eel = act ? e1 : 0.0;
    elj = act ? e2 : 0.0;
}

// ------------------------------------------------------------------------------------------------
// Prepare: slot records (origin-relative FP64 position, scaled charge), slot
// types; zero the fixed-point accumulators; reset the work queues.  With
// chk.ref set it also applies CheckForUpdate as NBModelABFS_Update calls it
// (pMolecule-1.9.0/extensions/csource/NBModelABFS.c:538-545, 691-746: some
// free atom moved more than half the buffer since the list was built) and
// stores chk.epoch into *chk.flag when it holds.
// ------------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_prepare(const SysDesc *__restrict__ descs, int tiled, int *__restrict__ heads,
                                                 PrepareCheck chk)
{
    const SysDesc &D = descs[blockIdx.y];
    const int s = blockIdx.x * 256 + threadIdx.x;
    if (heads && blockIdx.y == 0 && blockIdx.x == 0 && threadIdx.x < 8) heads[32 * threadIdx.x] = 0;
    if (blockIdx.x == 0 && D.eacc)
        for (int k = threadIdx.x; k < NB_EACC_WORDS; k += 256) D.eacc[k] = 0;
    bool up = false;
    if (s < D.n_pad) {
        for (int k = 0; k < 6; k++) D.facc[(size_t)k * D.n_pad + s] = 0;
        const double *xyz = chk.xyz ? chk.xyz : D.xyz;
        const int a = D.slot_atom[s];
        double x = 0.0, y = 0.0, z = 0.0;
        if (a >= 0) {
            x = xyz[3 * (size_t)a];
            y = xyz[3 * (size_t)a + 1];
            z = xyz[3 * (size_t)a + 2];
            if (chk.ref && (!chk.free_flags || chk.free_flags[a])) {
#pragma clang fp contract(off)
                const double dx = x - chk.ref[3 * (size_t)a];
                const double dy = y - chk.ref[3 * (size_t)a + 1];
                const double dz = z - chk.ref[3 * (size_t)a + 2];
                up = dx * dx + dy * dy + dz * dz > chk.bufsq;
            }
        }
        if (tiled) {
            double4 v = make_double4(0.0, 0.0, 0.0, 0.0);
            int t = 0;
            if (a >= 0) {
                v = make_double4(x - D.origin[0], y - D.origin[1], z - D.origin[2], D.charge[a] * D.qscale);
                t = D.atype[a];
            }
            D.xyzq[s] = v;
            D.styp[s] = t;
        }
    }
    if (__any(up) && (threadIdx.x & 63) == 0) *(volatile int *)chk.flag = chk.epoch;
}

// ------------------------------------------------------------------------------------------------
// Tiled force kernel (persistent; one queue per XCD with stealing).  The LJ
// table of the item's system sits in LDS (ntypes <= NB_LDS_TYPES); the next
// entry's j-cluster record and masks are fetched while the current one is
// evaluated.
// ------------------------------------------------------------------------------------------------
#define NB_LDS_TYPES 16

__device__ __forceinline__ double4 load_slot(const GAS double *__restrict__ xr, int k)
{
    return make_double4(xr[4 * k], xr[4 * k + 1], xr[4 * k + 2], xr[4 * k + 3]);
}

// Entry fields through the vector memory path: J and cbits are read by every
// lane from one address (one request) and made uniform; each lane reads its
// own tile-mask byte (64 consecutive bytes per entry).
__device__ __forceinline__ int2 entry_jc(const GAS unsigned char *__restrict__ eb, int e)
{
    const GAS int *p = (const GAS int *)(eb + (size_t)e * sizeof(Entry));
    return make_int2(p[0], p[2]);
}

__device__ __forceinline__ int entry_bits(const GAS unsigned char *__restrict__ eb, int e, int lane)
{
    return eb[(size_t)e * sizeof(Entry) + 16 + lane];
}

// j-force partials (gx, gy, gz) of lane 8a+b are summed over a by a
// reduce-scatter: after three exchanges lanes a = 0, 2, 4 hold the x, y and z
// sums of j-atom b, so one fixed-point conversion serves all three components.
__device__ __forceinline__ float jforce_reduce_scatter(int a, float gx, float gy, float gz)
{
    const bool lo4 = a < 4;
    const float r1 = __shfl_xor(lo4 ? gz : gx, 32, 64);
    const float r2 = __shfl_xor(gy, 32, 64);
    const float k1 = (lo4 ? gx : gz) + r1;   // a < 4: x; a >= 4: z
    const float k2 = gy + r2;                // a < 4: y
    const bool hi2 = (a & 2) != 0;
    const float snd = lo4 ? (hi2 ? k1 : k2) : k1;
    float kp = lo4 ? (hi2 ? k2 : k1) : k1;
    kp += __shfl_xor(snd, 16, 64);
    kp += __shfl_xor(kp, 8, 64);
    return kp;
}

// LDS of the tiled kernel.  The i-force transpose at the end of an item
// reuses the space of the item's i-cluster records and LJ table.
template <bool LDS_LJ>
union TiledShared {
    struct {
        double4 sI[NB_SC];
        int sT[NB_SC];
        double2 sLJ[LDS_LJ ? NB_LDS_TYPES * NB_LDS_TYPES : 1];
    } in;
    float sF[NB_SCC * 3 * NB_SC];   // [c][component][lane]
};

template <bool LDS_LJ>
__global__ __launch_bounds__(64) void k_force_tiled(const SysDesc *__restrict__ descs,
                                                    const WorkItem *__restrict__ items, QueueParams qp, AbfsFactors F)
{
    const int lane = threadIdx.x;
    const int a = lane >> 3, b = lane & 7;
    __shared__ TiledShared<LDS_LJ> sh;
    double4 *sI = sh.in.sI;
    int *sT = sh.in.sT;
    double2 *sLJ = sh.in.sLJ;
    int q = blockIdx.x & 7;
    int tries = 0;
    // heads == null: one work item per workgroup (the hardware dispatcher
    // balances the load); otherwise a persistent loop over per-XCD queues whose
    // counters sit on separate 128-byte lines.
    bool done = false;
    while (!done && tries < 8) {
        int it;
        if (qp.heads == nullptr) {
            it = blockIdx.x;
            done = true;
        } else {
            int local = 0;
            if (lane == 0) local = atomicAdd(&qp.heads[32 * q], 1);
            local = __builtin_amdgcn_readfirstlane(local);
            const int qs = qp.starts[q], qe = qp.starts[q + 1];
            if (qs + local >= qe) {
                q = (q + 1) & 7;
                tries++;
                continue;
            }
            it = qs + local;
        }
        const WorkItem w = items[it];
        const SysDesc *D = &descs[w.sys];
        const GAS double *__restrict__ xq = (const GAS double *)D->xyzq;
        const GAS int *__restrict__ ty = as_global(D->styp);
        const GAS double *__restrict__ ljA = as_global(D->ljA);
        const GAS double *__restrict__ ljB = as_global(D->ljB);
        const GAS unsigned char *__restrict__ eb = (const GAS unsigned char *)D->entries;
        long long *facc = D->facc;
        const int n_pad = D->n_pad;
        const int nt = D->ntypes;
        const int I0 = w.sc * NB_SCC;

        __syncthreads();
        if (LDS_LJ)
            for (int k = lane; k < nt * nt; k += 64) sLJ[k] = make_double2(ljA[k], ljB[k]);
        sI[lane] = load_slot(xq, I0 * NB_CL + lane);
        sT[lane] = ty[I0 * NB_CL + lane] * nt;
        __syncthreads();

        float fx[NB_SCC], fy[NB_SCC], fz[NB_SCC];
#pragma unroll
        for (int c = 0; c < NB_SCC; c++) fx[c] = fy[c] = fz[c] = 0.0f;
        double eel = 0.0, elj = 0.0;
        const int e_end = w.first + w.count;
        // Two-deep pipeline: entry e+2's header and mask byte and entry e+1's
        // j-cluster record are in flight while entry e is evaluated.
        int2 jc0 = entry_jc(eb, w.first);
        int lb0 = entry_bits(eb, w.first, lane);
        int2 jc1 = jc0;
        int lb1 = 0;
        if (w.first + 1 < e_end) {
            jc1 = entry_jc(eb, w.first + 1);
            lb1 = entry_bits(eb, w.first + 1, lane);
        }
        int J = __builtin_amdgcn_readfirstlane(jc0.x);
        double4 xj = load_slot(xq, J * NB_CL + b);
        int tj = ty[J * NB_CL + b];
        for (int e = w.first; e < e_end; e++) {
            int2 jc2 = jc1;
            int lb2 = 0;
            if (e + 2 < e_end) {
                jc2 = entry_jc(eb, e + 2);
                lb2 = entry_bits(eb, e + 2, lane);
            }
            double4 xjn = xj;
            int tjn = tj;
            int Jn = J;
            if (e + 1 < e_end) {
                Jn = __builtin_amdgcn_readfirstlane(jc1.x);
                xjn = load_slot(xq, Jn * NB_CL + b);
                tjn = ty[Jn * NB_CL + b];
            }
            const int cb = __builtin_amdgcn_readfirstlane(jc0.y);
            const int lb = lb0;
            float gx = 0.0f, gy = 0.0f, gz = 0.0f;
#pragma unroll
            for (int c = 0; c < NB_SCC; c++) {
                if (cb & (1 << c)) {
                    const bool inl = (lb >> c) & 1;
                    const double4 xi = sI[c * NB_CL + a];
                    const int tij = sT[c * NB_CL + a] + tj;
                    double A, B;
                    if (LDS_LJ) {
                        const double2 ab = sLJ[tij];
                        A = ab.x;
                        B = ab.y;
                    } else {
                        A = ljA[tij];
                        B = ljB[tij];
                    }
                    const double dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
                    const double r2 = fma(dx, dx, fma(dy, dy, dz * dz));
                    const bool act = inl && (r2 <= F.r2Off);
                    if (__any(act)) {
                        double fs, x1, x2;
                        pair_tiled(F, r2, xi.w * xj.w, A, B, act, fs, x1, x2);
                        const float f = (float)fs;
                        const float px = f * (float)dx, py = f * (float)dy, pz = f * (float)dz;
                        fx[c] += px; fy[c] += py; fz[c] += pz;
                        gx -= px; gy -= py; gz -= pz;
                        eel += x1;
                        elj += x2;
                    }
                }
            }
            const float gj = jforce_reduce_scatter(a, gx, gy, gz);
            if (a == 0 || a == 2 || a == 4)
                fx_atomic_add(facc, n_pad, J * NB_CL + b, a >> 1, (double)gj);
            jc0 = jc1;
            lb0 = lb1;
            jc1 = jc2;
            lb1 = lb2;
            J = Jn;
            xj = xjn;
            tj = tjn;
        }
        // i-forces: transpose through LDS so that lane 8a+b sums atom a of
        // i-cluster c = b over the eight lanes that held its partials.
        __syncthreads();
#pragma unroll
        for (int c = 0; c < NB_SCC; c++) {
            sh.sF[(3 * c + 0) * NB_SC + lane] = fx[c];
            sh.sF[(3 * c + 1) * NB_SC + lane] = fy[c];
            sh.sF[(3 * c + 2) * NB_SC + lane] = fz[c];
        }
        __syncthreads();
        float ig[3];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const float4 *row = (const float4 *)&sh.sF[(3 * b + k) * NB_SC + NB_CL * a];
            const float4 u = row[0], v = row[1];
            ig[k] = ((u.x + u.y) + (u.z + u.w)) + ((v.x + v.y) + (v.z + v.w));
        }
        const int si = (I0 + b) * NB_CL + a;
#pragma unroll
        for (int k = 0; k < 3; k++) fx_atomic_add(facc, n_pad, si, k, (double)ig[k]);
        eel = wave_sum_f64(eel);
        elj = wave_sum_f64(elj);
        if (lane < 2) {
            const int slot = it & (NB_ESLOTS - 1);
            fx_add(eacc_word(D->eacc, lane, 0, slot), eacc_word(D->eacc, lane, 1, slot), lane == 0 ? eel : elj);
        }
    }
}

// ------------------------------------------------------------------------------------------------
// Plain FP64 path: one thread per atom over a full (both-direction) CSR list.
// ------------------------------------------------------------------------------------------------
template <int BS>
__device__ __forceinline__ void block_sum2(double &x, double &y)
{
    __shared__ double sx[BS], sy[BS];
    sx[threadIdx.x] = x;
    sy[threadIdx.x] = y;
    __syncthreads();
    for (int o = BS / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            sx[threadIdx.x] += sx[threadIdx.x + o];
            sy[threadIdx.x] += sy[threadIdx.x + o];
        }
        __syncthreads();
    }
    x = sx[0];
    y = sy[0];
}

__global__ __launch_bounds__(256) void k_force_plain(const SysDesc *__restrict__ descs, int sys, AbfsFactors F,
                                                     const int *__restrict__ rowptr, const int *__restrict__ cols)
{
#pragma clang fp contract(off)
    const SysDesc &D = descs[sys];
    const int i = blockIdx.x * 256 + threadIdx.x;
    double e1 = 0.0, e2 = 0.0;
    if (i < D.n) {
        double gx = 0.0, gy = 0.0, gz = 0.0;
        const double xi = D.xyz[3 * (size_t)i], yi = D.xyz[3 * (size_t)i + 1], zi = D.xyz[3 * (size_t)i + 2];
        const double qi = D.escale * D.charge[i];
        const int ti = D.ntypes * D.atype[i];
        for (int k = rowptr[i]; k < rowptr[i + 1]; k++) {
            const int j = cols[k];
            double xij = xi - D.xyz[3 * (size_t)j];
            double yij = yi - D.xyz[3 * (size_t)j + 1];
            double zij = zi - D.xyz[3 * (size_t)j + 2];
            const double r2 = (xij * xij + yij * yij + zij * zij);
            const double qij = qi * D.charge[j];
            const int tij = ti + D.atype[j];
            double x1, x2, dF;
            if (!pair_fp64(F, r2, qij, D.ljA[tij], D.ljB[tij], x1, x2, dF)) continue;
            e1 += x1;
            e2 += x2;
            gx += xij * (2.0 * dF);
            gy += yij * (2.0 * dF);
            gz += zij * (2.0 * dF);
        }
        D.gplain[3 * (size_t)i] = gx;
        D.gplain[3 * (size_t)i + 1] = gy;
        D.gplain[3 * (size_t)i + 2] = gz;
    }
    block_sum2<256>(e1, e2);
    if (threadIdx.x == 0) {
        D.eplain[2 * blockIdx.x] = e1;
        D.eplain[2 * blockIdx.x + 1] = e2;
    }
}

// ------------------------------------------------------------------------------------------------
// 1-4 list (NBModelABFS_MMMMEnergy's second pass, NBModelABFS.c:276-294): FP64
// with scale14 and the 1-4 LJ table; gradients into the fixed-point arrays.
// ------------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_pairs14(const SysDesc *__restrict__ descs, AbfsFactors F)
{
#pragma clang fp contract(off)
    const SysDesc &D = descs[blockIdx.y];
    if ((int)blockIdx.x >= D.n14_blocks) return;
    const int k = blockIdx.x * 256 + threadIdx.x;
    double e1 = 0.0, e2 = 0.0;
    if (k < D.n14) {
        const int i = D.p14[2 * k], j = D.p14[2 * k + 1];
        double xij = D.xyz[3 * (size_t)i] - D.xyz[3 * (size_t)j];
        double yij = D.xyz[3 * (size_t)i + 1] - D.xyz[3 * (size_t)j + 1];
        double zij = D.xyz[3 * (size_t)i + 2] - D.xyz[3 * (size_t)j + 2];
        const double r2 = (xij * xij + yij * yij + zij * zij);
        const double qij = (D.escale14 * D.charge[i]) * D.charge[j];
        const int tij = D.ntypes * D.atype[i] + D.atype[j];
        double x1, x2, dF;
        if (pair_fp64(F, r2, qij, D.ljA14[tij], D.ljB14[tij], x1, x2, dF)) {
            e1 = x1;
            e2 = x2;
            const int si = D.atom_slot[i], sj = D.atom_slot[j];
            const double px = xij * (2.0 * dF), py = yij * (2.0 * dF), pz = zij * (2.0 * dF);
            fx_atomic_add(D.facc, D.n_pad, si, 0, px);
            fx_atomic_add(D.facc, D.n_pad, si, 1, py);
            fx_atomic_add(D.facc, D.n_pad, si, 2, pz);
            fx_atomic_add(D.facc, D.n_pad, sj, 0, -px);
            fx_atomic_add(D.facc, D.n_pad, sj, 1, -py);
            fx_atomic_add(D.facc, D.n_pad, sj, 2, -pz);
        }
    }
    block_sum2<256>(e1, e2);
    if (threadIdx.x == 0) {
        D.epart14[2 * blockIdx.x] = e1;
        D.epart14[2 * blockIdx.x + 1] = e2;
    }
}

// ------------------------------------------------------------------------------------------------
// Finalize: gradients (fixed point -> FP64, plus the plain-path rows) added
// into the caller's array; block 0 of each system also sums the energies in a
// fixed order.
// ------------------------------------------------------------------------------------------------
__device__ void sum_range2(const double *__restrict__ p, int m, double &x, double &y)
{
    double sx = 0.0, sy = 0.0;
    for (int k = threadIdx.x; k < m; k += 256) {
        sx += p[2 * (size_t)k];
        sy += p[2 * (size_t)k + 1];
    }
    block_sum2<256>(sx, sy);
    x = sx;
    y = sy;
    __syncthreads();
}

__global__ __launch_bounds__(256) void k_finalize(const SysDesc *__restrict__ descs, int plain)
{
    const SysDesc &D = descs[blockIdx.y];
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < D.n && D.grad != nullptr) {
        const int s = D.atom_slot[i];
        for (int k = 0; k < 3; k++) {
            double g = fx_value(D.facc, D.n_pad, s, k);
            if (plain) g += D.gplain[3 * (size_t)i + k];
            D.grad[3 * (size_t)i + k] += g;
        }
    }
    if (blockIdx.x != 0) return;
    double el, lj, el14, lj14;
    if (plain) {
        sum_range2(D.eplain, D.n_plain_blocks, el, lj);
        el *= 0.5;
        lj *= 0.5;
    } else {
        double ex = 0.0, ey = 0.0;
        if (threadIdx.x < NB_ESLOTS) {
            const int t = threadIdx.x;
            ex = (double)*eacc_word(D.eacc, 0, 1, t) * NB_FX_HI_UNIT + (double)*eacc_word(D.eacc, 0, 0, t) / NB_FX_LO_SCALE;
            ey = (double)*eacc_word(D.eacc, 1, 1, t) * NB_FX_HI_UNIT + (double)*eacc_word(D.eacc, 1, 0, t) / NB_FX_LO_SCALE;
        }
        block_sum2<256>(ex, ey);
        el = ex;
        lj = ey;
        __syncthreads();
    }
    sum_range2(D.epart14, D.n14_blocks, el14, lj14);
    if (threadIdx.x == 0) {
        D.energies[0] = el;
        D.energies[1] = lj;
        D.energies[2] = el14;
        D.energies[3] = lj14;
    }
}

// ------------------------------------------------------------------------------------------------
// Launchers.
// ------------------------------------------------------------------------------------------------
void launch_prepare(const SysDesc *d_desc, int nsys, int max_npad, int tiled, int *heads, const PrepareCheck &chk,
                    hipStream_t s)
{
    if (nsys <= 0) return;
    hipLaunchKernelGGL(k_prepare, dim3(std::max(nblk(max_npad, 256), 1), nsys), dim3(256), 0, s, d_desc, tiled, heads,
                       chk);
}

void launch_force_tiled(const SysDesc *d_desc, const WorkItem *items, QueueParams q, const AbfsFactors &F,
                        int nblocks, int max_ntypes, hipStream_t s)
{
    if (nblocks <= 0) return;
    if (max_ntypes <= NB_LDS_TYPES)
        hipLaunchKernelGGL(k_force_tiled<true>, dim3(nblocks), dim3(64), 0, s, d_desc, items, q, F);
    else
        hipLaunchKernelGGL(k_force_tiled<false>, dim3(nblocks), dim3(64), 0, s, d_desc, items, q, F);
}

void launch_force_plain(const SysDesc *d_desc, int sys, int n, const AbfsFactors &F, const int *rowptr,
                        const int *cols, hipStream_t s)
{
    hipLaunchKernelGGL(k_force_plain, dim3(max(nblk(n, 256), 1)), dim3(256), 0, s, d_desc, sys, F, rowptr, cols);
}

void launch_pairs14(const SysDesc *d_desc, int nsys, int max_n14_blocks, const AbfsFactors &F, hipStream_t s)
{
    if (nsys > 0 && max_n14_blocks > 0)
        hipLaunchKernelGGL(k_pairs14, dim3(max_n14_blocks, nsys), dim3(256), 0, s, d_desc, F);
}

void launch_finalize(const SysDesc *d_desc, int nsys, int max_n, int plain, hipStream_t s)
{
    if (nsys <= 0) return;
    hipLaunchKernelGGL(k_finalize, dim3(std::max(nblk(max_n, 256), 1), nsys), dim3(256), 0, s, d_desc, plain);
}
