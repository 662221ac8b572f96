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
//
// Periodic images (MMMMImageEnergy, pMolecule-1.9.0/extensions/csource/
// NBModelABFS.c:1161-1290): an image item's j-clusters are transformed to the
// image on the fly (x' = R x + T, FP64), its energies and forces carry the
// image scale, its j-gradients are rotated back by R^T before they are added,
// and it sums the image-side gradients and their moments for dE/dM.
#include "nbabfs_internal.h"
#include "nbabfs_kernels.h"

#include <hip/hip_ext.h>

#include <algorithm>
#include <type_traits>

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
// Newton step in FP64 (relative error ~1e-14; no FP64 divide or square root).
// On gfx950 FP64 FMA issues at the FP32 rate, so the scalar core stays in FP64
// and only the rsqrt seed and the force vectors are FP32.
//
// Lennard-Jones in both undamped regions is one polynomial in s = 1/r,
//   E = C12 s^12 + C6 s^6 + C3 s^3 + C0,
// whose coefficients per type pair and region come from a row table (plain:
// A, -B, 0, B*bShift6 - A*aShift12; switched: the expanded squares
// A*aK12*(s^6-aF6)^2 - B*bK6*(s^3-bF3)^2), so the region choice is a row
// index instead of two evaluated branches.  The Coulomb switch keeps the
// reference polynomials, blended by a 0/1 region factor.  The damped region
// (r < dampingCutoff) is a wave-uniform rare path on the raw A, B.
__device__ __forceinline__ void pair_rows(const AbfsFactors &F, double r2, double qq, double4 C, bool sw,
                                          unsigned long long mact, const GAS double *__restrict__ ljA,
                                          const GAS double *__restrict__ ljB, int tij, double &fs, double &eel,
                                          double &elj)
{
    const double y = (double)__builtin_amdgcn_rsqf((float)r2);
    const double h = fma(-(r2 * y), y, 1.0);
    const double s = fma(0.5 * y, h, y);
    const double s2 = s * s;
    const double s3 = s2 * s;
    const double s6 = s3 * s3;
    // Coulomb in both undamped regions as one expression, m = 1 in the switched region and 0 in the plain
    // one (the reference's E = qq (s P + K) with P = 1, K = qShift1 below the inner cutoff):
    //   E / qq = s + qShift1 + m (s (P - 1) + qShift2 - qShift1),   P - 1 = (a - 1) - r2 (b + r2 (c + d r2)),
    //   qq Q   = qq + qq m (Q - 1),                                 Q - 1 = (a - 1) + r2 (b + r2 (3c + 5d r2)).
    // A product with m in {0, 1} is exact, so each region evaluates its own reference form.
    const double m = sw ? 1.0 : 0.0;
    const double pm1 = fma(-r2, fma(r2, fma(F.d, r2, F.c), F.b), F.a - 1.0);
    const double qm1 = fma(r2, fma(r2, fma(5.0 * F.d, r2, 3.0 * F.c), F.b), F.a - 1.0);
    double e1 = qq * fma(m, fma(s, pm1, F.qShift2 - F.qShift1), s + F.qShift1);
    const double qQ = fma(qq * m, qm1, qq);
    // Lennard-Jones row E = C12 s^12 + C6 s^6 + C3 s^3 + C0 (t1 = C12 s^6 + C6), and
    // inner = s^3 (4 C12 s^6 + 2 C6) + C3 = 2 s^3 (C12 s^6 + t1) + C3.
    const double t1 = fma(C.x, s6, C.y);
    double e2 = fma(t1, s6, fma(C.z, s3, C.w));
    const double inner = fma(2.0 * s3, fma(C.x, s6, t1), C.z);
    // Twice dE/d(r^2) of both terms (the gradient's factor 2 folded in; exact):
    //   Coulomb -qq s^3 Q, Lennard-Jones -3 s^5 inner, so fs = -s^3 (qq Q + 3 s^2 inner).
    double f = -s3 * fma(3.0 * s2, inner, qQ);
    if (__builtin_amdgcn_ballot_w64(r2 < F.r2Damp) & mact) {
        if (r2 < F.r2Damp) {
            const double A = ljA[tij], B = ljB[tij];
            e1 = qq * fma(-F.qAlpha, r2, F.qF0);
            e2 = A * fma(-F.aAlpha, r2, F.aF0) - B * fma(-F.bAlpha, r2, F.bF0);
            f = (-qq * F.qAlpha) * 2.0 + (-A * F.aAlpha + B * F.bAlpha) * 2.0;
        }
    }
    fs = f;
    eel = e1;
    elj = e2;
}

// ------------------------------------------------------------------------------------------------
// Prepare: slot records (origin-relative FP64 position, scaled charge), slot
// types; zero the fixed-point accumulators; reset the work queues.  With
// chk.ref set it also applies CheckForUpdate as NBModelABFS_Update calls it
// (pMolecule-1.9.0/extensions/csource/NBModelABFS.c:538-545, 691-746: some
// free atom moved more than half the buffer since the list was built) and
// stores chk.epoch into *chk.flag when it holds.
// ------------------------------------------------------------------------------------------------
// A guarded launch does nothing when its update test asked for a rebuild.
__device__ __forceinline__ bool guard_skip(const ForceGuard &g)
{
    if (!g.flag) return false;
    // Two relaxed loads that may be in flight together (volatile loads would be issued one after the other).
    const int epoch = g.epoch_ptr ? __hip_atomic_load(g.epoch_ptr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                  : g.epoch;
    return __hip_atomic_load(g.flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == epoch;
}

__global__ __launch_bounds__(256) void k_prepare(const SysDesc *__restrict__ descs, int tiled, int *__restrict__ heads,
                                                 PrepareCheck chk)
{
    const SysDesc &D = descs[blockIdx.y];
    const int s = blockIdx.x * 256 + threadIdx.x;
    if (heads && blockIdx.y == 0 && blockIdx.x == 0 && threadIdx.x < 8) heads[32 * threadIdx.x] = 0;
    if (blockIdx.x == 0 && D.eacc)
        for (int k = threadIdx.x; k < NB_EACC_WORDS; k += 256) D.eacc[k] = 0;
    if (D.gacc)
        for (int k = blockIdx.x * 256 + threadIdx.x; k < NB_GACC_WORDS_PER_IMAGE * D.nimg; k += gridDim.x * 256)
            D.gacc[k] = 0;
    bool up = false, nonfinite = false;
    if (s < D.n_pad) {
        for (int k = 0; k < 6; k++) D.facc[(size_t)k * D.n_pad + s] = 0;
        const double *xyz = chk.xyz ? chk.xyz : D.xyz;
        const int a = D.slot_atom[s];
        double x = 0.0, y = 0.0, z = 0.0;
        if (a >= 0) {
            x = xyz[3 * (size_t)a];
            y = xyz[3 * (size_t)a + 1];
            z = xyz[3 * (size_t)a + 2];
            nonfinite = !(isfinite(x) && isfinite(y) && isfinite(z));
            if (chk.ref && (!chk.free_flags || (chk.free_flags[a] & 1))) {
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
    // A NaN or infinite coordinate fails no distance test (its pairs drop out
    // silently), so it is reported here, where every atom is read.
    raise_error(D.err, nonfinite, NB_ERR_COORD);
    if (chk.counter) {
        // One 64-bit atomic per block carries both its completion (high word)
        // and whether any of its atoms moved (low word), so the last block knows
        // the answer without further fences; it alone publishes the flag and the
        // completion marker to the host (one system-scope fence).
        const int moved = __syncthreads_or(up);
        if (threadIdx.x == 0) {
            if (moved) *(volatile int *)chk.dflag = chk.epoch;
            const unsigned long long inc = (1ull << 32) + (moved ? 1ull : 0ull);
            const unsigned long long old = atomicAdd(chk.counter, inc);
            const unsigned long long total = (unsigned long long)gridDim.x * gridDim.y;
            if ((old >> 32) + 1 == total) {
                *(volatile unsigned long long *)chk.counter = 0ull;
                if (((old + inc) & 0xffffffffull) != 0) *(volatile int *)chk.flag = chk.epoch;
                __threadfence_system();
                *(volatile int *)chk.done = chk.epoch;
            }
        }
    } else if (chk.ref && __any(up) && (threadIdx.x & 63) == 0) {
        *(volatile int *)chk.flag = chk.epoch;
        *(volatile int *)chk.dflag = chk.epoch;
    }
}

// A wave-uniform FP64 value moved into scalar registers.
__device__ __forceinline__ double uniform_f64(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readfirstlane((int)(b & 0xffffffffll));
    const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// Cluster-frame image position x' = R x + T of a slot record (charge kept);
// X = R (9, row-major) then T (3), read from LDS at each use (broadcast reads)
// so that the transformation holds no registers across the entry loop.
__device__ __forceinline__ double4 image_slot(double4 v, const volatile double *X)
{
    return make_double4(fma(X[0], v.x, fma(X[1], v.y, fma(X[2], v.z, X[9]))),
                        fma(X[3], v.x, fma(X[4], v.y, fma(X[5], v.z, X[10]))),
                        fma(X[6], v.x, fma(X[7], v.y, fma(X[8], v.z, X[11]))), v.w);
}

// ------------------------------------------------------------------------------------------------
// The 1-4 list (the second pass of NBModelABFS_MMMMEnergy, pMolecule-1.9.0/
// extensions/csource/NBModelABFS.c:228-301): 64 pairs per block (one wave),
// FP64 in the reference's expression order, forces into the fixed-point
// accumulators, the block's two energies (wave sums) into epart14[blk].  Run
// by k_pairs14, or by extra workgroups of the main pair launch.
// ------------------------------------------------------------------------------------------------
__device__ __forceinline__ void pairs14_block(const SysDesc &D, const AbfsFactors &F, int blk, int lane)
{
#pragma clang fp contract(off)
    const int k = blk * 64 + lane;
    double e1 = 0.0, e2 = 0.0;
    bool bad = false;
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
            bad |= fx_atomic_add(D.facc, D.n_pad, si, 0, px);
            bad |= fx_atomic_add(D.facc, D.n_pad, si, 1, py);
            bad |= fx_atomic_add(D.facc, D.n_pad, si, 2, pz);
            bad |= fx_atomic_add(D.facc, D.n_pad, sj, 0, -px);
            bad |= fx_atomic_add(D.facc, D.n_pad, sj, 1, -py);
            bad |= fx_atomic_add(D.facc, D.n_pad, sj, 2, -pz);
        }
    }
    e1 = wave_sum_f64(e1);
    e2 = wave_sum_f64(e2);
    if (lane == 0) {
        D.epart14[2 * blk] = e1;
        D.epart14[2 * blk + 1] = e2;
    }
    raise_error(D.err, bad || !isfinite(e1) || !isfinite(e2), NB_ERR_RANGE);
}

// ------------------------------------------------------------------------------------------------
// Tiled force kernel: one work item per workgroup, or a persistent loop over
// per-XCD queues with stealing.  The LJ row table of the item's system sits
// in LDS when it fits; the next entry's j-cluster record and masks are
// fetched while the current one is evaluated.
// ------------------------------------------------------------------------------------------------

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

// Sum over the eight consecutive lanes 8b..8b+7 with DPP moves: quad
// exchanges (xor 1, xor 2) then the half-row mirror (lane i <-> 7-i).  Every
// lane of the group ends with the total, in the same order on every lane.
template <int CTRL>
__device__ __forceinline__ float dpp_f32(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}

__device__ __forceinline__ float sum8(float v)
{
    v += dpp_f32<0xB1>(v);    // quad_perm [1,0,3,2]
    v += dpp_f32<0x4E>(v);    // quad_perm [2,3,0,1]
    v += dpp_f32<0x141>(v);   // row_half_mirror
    return v;
}

// FP64 version (full-FP64 tiled kernel): a butterfly over the same 8 lanes;
// every lane adds the same operands in the same tree, so all end bit-equal.
__device__ __forceinline__ double sum8(double v)
{
    v += __shfl_xor(v, 1, 64);
    v += __shfl_xor(v, 2, 64);
    v += __shfl_xor(v, 4, 64);
    return v;
}

// LDS of the tiled kernel (dynamic).  A workgroup holds WPG wavefronts, each
// evaluating one work item.  At offset 0 the LJ row table of the group's
// system when it fits (ROWS_LDS: staged once per group), then one region per
// wave: the item's i-cluster records and slot types, then the per-entry
// j-force sums with their j-clusters, flushed to the fixed-point accumulators
// once at the end of the item; the i-force transpose at the end of the item
// reuses the start of the region once the j-forces are flushed.
#define NB_LDS_I_BYTES (NB_SC * 32 + NB_SC * 4)
#define NB_LDS_F_BYTES(acc) (NB_SCC * 3 * NB_SC * (acc))
#define NB_LDS_ROW_LIMIT 16384
#define NB_LDS_GROUP_LIMIT (160 * 1024 / 4)

static inline size_t tiled_lds_rows(int ntypes, bool rows_lds) { return rows_lds ? (size_t)2 * ntypes * ntypes * 32 : 0; }

// acc: bytes of a force accumulator (4: FP32, the mixed kernel; 8: FP64, the full-FP64 kernel).
static inline size_t tiled_lds_wave(int max_count, int acc)
{
    const size_t items = NB_LDS_I_BYTES + (size_t)max_count * (3 * NB_CL * acc + 4);
    return std::max(items, (size_t)NB_LDS_F_BYTES(acc));
}

// Wave-scope synchronisation of a wave's own LDS region (workgroups of several
// independent waves); a workgroup barrier when the workgroup is one wave.
template <int WPG>
__device__ __forceinline__ void wave_region_sync()
{
    if (WPG == 1) {
        __syncthreads();
    } else {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

// Four waves per SIMD: the register allocator fits the kernel in 128 VGPRs
// without spilling (137 unconstrained, three waves), measured 2-3% faster.
// IMG: the launch over the image lists' work items (the main list's items run
// in an instance without any image code); it runs at three waves per SIMD,
// whose register budget holds the transformation without spilling.
// F64: the full-FP64 instance (the NBABFS_TILED_FP64 comparison path): the same list, tiles, staging and
// fixed-point sums, with the reference's FP64 pair arithmetic (pair_fp64: divide and square root, no FP32
// seed, raw A and B per type pair) and FP64 force vectors; three waves per SIMD for its larger accumulators.
// WPG: wavefronts per workgroup.  With 4 (the one-item-per-wave launches of
// the main list), the row table is staged once for four items and a group's
// LDS stays small enough for four waves per SIMD; the four items of a group
// belong to one system (batches pad each system's items to a multiple of 4
// with empty ones).  Image items and the persistent queues run one wave per
// group.
template <bool ROWS_LDS, bool IMG, bool F64, int WPG>
__global__ __launch_bounds__(64 * WPG) __attribute__((amdgpu_waves_per_eu((IMG || F64) ? 3 : 4))) void k_force_tiled(
    const SysDesc *__restrict__ descs, const WorkItem *__restrict__ items, QueueParams qp, AbfsFactors F, int rows_bytes,
    int wave_bytes, ForceGuard guard, int nitems, int nforce_groups, int sys1)
{
    const int lane = threadIdx.x & 63;
    const int wave = WPG == 1 ? 0 : (int)__builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    // A group's start is a chain of dependent loads (its work item, the system's descriptor, the records the item
    // names).  The loads that depend on nothing go out together with the guard's flag: the work item of a wave
    // with one item, and (sys1 >= 0: every item belongs to system sys1) the descriptor and its row table.
    WorkItem w0 = {0, 0, 0, 0};
    const bool one_item = qp.heads == nullptr && (IMG || (int)blockIdx.x < nforce_groups);
    if (one_item && nitems > 0) w0 = items[min((int)blockIdx.x * WPG + wave, nitems - 1)];
    // guard_skip with both loads issued here unconditionally (a null pointer reads the descriptor instead) and
    // their values compared only where the answer is needed.
    const int *gfp = guard.flag ? guard.flag : (const int *)descs;
    const int *gep = guard.epoch_ptr ? guard.epoch_ptr : gfp;
    const int gflag = __hip_atomic_load(gfp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const int gepoch = __hip_atomic_load(gep, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    // Groups past the pair blocks evaluate the 1-4 list of system 0 (single-state launches), one block per wave.
    if (!IMG && (int)blockIdx.x >= nforce_groups) {
        const int blk = ((int)blockIdx.x - nforce_groups) * WPG + wave;
        if (!guard_skip(guard) && blk < descs[0].n14_blocks) pairs14_block(descs[0], F, blk, lane);
        return;
    }
    // Lane l = 8b + a: a is the i-atom within an i-cluster, b the j-atom of
    // the entry's j-cluster.  The entry masks are stored per l' = 8a + b.
    const int a = lane & 7, b = lane >> 3;
    const int bl = 8 * a + b;
    extern __shared__ __align__(32) unsigned char lds[];
    __shared__ __align__(16) double sX[12];     // image items: R then T of the item's image
    using acc_t = typename std::conditional<F64, double, float>::type;
    double4 *sR = (double4 *)lds;
    unsigned char *region = lds + rows_bytes + (size_t)wave * wave_bytes;
    double4 *sI = (double4 *)region;
    int *sT = (int *)(region + NB_SC * 32);
    acc_t *sF = (acc_t *)region;
    // The group's row table, staged once when the items are one per wave.
    int rows_sys = -1;
    if (ROWS_LDS && qp.heads == nullptr && WPG > 1) {
        rows_sys = sys1 >= 0 ? sys1 : items[(int)blockIdx.x * WPG].sys;
        const SysDesc *D0 = &descs[rows_sys];
        const GAS double *__restrict__ rows0 = as_global(D0->ljrow);
        const int nr = 2 * D0->ntypes * D0->ntypes;
        for (int k = threadIdx.x; k < nr; k += 64 * WPG) sR[k] = load_slot(rows0, k);
        __syncthreads();
    }
    // The guard's answer is needed only now (the loads above are reads; the flag is uniform over the launch).
    if (guard.flag != nullptr && gflag == (guard.epoch_ptr ? gepoch : guard.epoch)) return;
    int q = blockIdx.x & 7;
    int tries = 0;
    // heads == null: one work item per wave (the hardware dispatcher
    // balances the load); otherwise a persistent loop over per-XCD queues whose
    // counters sit on separate 128-byte lines (one wave per group).
    bool done = false;
    while (!done && tries < 8) {
        int it;
        if (qp.heads == nullptr) {
            it = (int)blockIdx.x * WPG + wave;
            done = true;
            if (it >= nitems) break;
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
        const WorkItem w = qp.heads == nullptr ? w0 : items[it];
        // Empty items pad a batch's systems to whole groups; an item of another system than the staged rows
        // would be a layout error (reported, never evaluated with the wrong rows).
        if (w.count <= 0) continue;
        if (rows_sys >= 0 && w.sys != rows_sys) {
            raise_error(descs[rows_sys].err, true, NB_ERR_RANGE);
            continue;
        }
        const SysDesc *D = &descs[sys1 >= 0 ? sys1 : (IMG ? (w.sys & 0xffff) : w.sys)];
        const GAS double *__restrict__ xq = (const GAS double *)D->xyzq;
        const GAS int *__restrict__ ty = as_global(D->styp);
        const GAS double *__restrict__ ljA = as_global(D->ljA);
        const GAS double *__restrict__ ljB = as_global(D->ljB);
        const GAS double *__restrict__ rows = as_global(D->ljrow);
        const GAS unsigned char *__restrict__ eb = (const GAS unsigned char *)D->entries;
        long long *facc = D->facc;
        const int n_pad = D->n_pad;
        const int nt = D->ntypes;
        const int I0 = w.sc * NB_SCC;
        acc_t *sJF = (acc_t *)(region + NB_LDS_I_BYTES);   // [entry][3 * NB_CL]
        int *sJ = (int *)(sJF + 3 * NB_CL * w.count);      // [entry]
        // The item's image (-1: main list); all of its entries share it.
        int img = -1;
        double isc = 1.0;
        const GAS double *X = nullptr;
        if (IMG) {
            img = __builtin_amdgcn_readfirstlane((w.sys >> 16) - 1);
            X = (const GAS double *)&D->imgxf[img];
            isc = uniform_f64(X[12]);
        }

        // The item's independent loads go out together: its super-cluster's records and types and its first two
        // entries' headers and mask bytes; then the first j-cluster's records, which need the first header.
        // Pipeline of the entry loop: while entry e is evaluated, entry e+1's j-cluster record and entry e+2's
        // header and mask byte are in flight.  The indices are clamped, so every iteration issues the same loads
        // and no atomic is pending inside the loop.
        const int e_last = w.first + w.count - 1;
        const double4 irec = load_slot(xq, I0 * NB_CL + lane);
        const int ityp = ty[I0 * NB_CL + lane];
        int2 jc0 = entry_jc(eb, w.first);
        int lb0 = entry_bits(eb, w.first, bl);
        int2 jc1 = entry_jc(eb, min(w.first + 1, e_last));
        int lb1 = entry_bits(eb, min(w.first + 1, e_last), bl);

        // A persistent wave's region still holds its previous item's data until every lane is past it.
        if (qp.heads != nullptr) wave_region_sync<WPG>();
        if (IMG && lane < 12) sX[lane] = X[lane];
        if (ROWS_LDS && rows_sys < 0)
            for (int k = lane; k < 2 * nt * nt; k += 64) sR[k] = load_slot(rows, k);
        sI[lane] = irec;
        sT[lane] = 64 * nt * ityp;   // byte offset of the type's first row pair
        int J = __builtin_amdgcn_readfirstlane(jc0.x);
        double4 xj = load_slot(xq, J * NB_CL + b);
        int tj = 64 * ty[J * NB_CL + b];
        wave_region_sync<WPG>();

        acc_t fx[NB_SCC], fy[NB_SCC], fz[NB_SCC];
#pragma unroll
        for (int c = 0; c < NB_SCC; c++) fx[c] = fy[c] = fz[c] = (acc_t)0;
        double eel = 0.0, elj = 0.0;
        for (int e = w.first; e <= e_last; e++) {
            const int Jn = __builtin_amdgcn_readfirstlane(jc1.x);
            const double4 xjn = load_slot(xq, Jn * NB_CL + b);
            const int tjn = 64 * ty[Jn * NB_CL + b];
            const int2 jc2 = entry_jc(eb, min(e + 2, e_last));
            const int lb2 = entry_bits(eb, min(e + 2, e_last), bl);
            const int cb = __builtin_amdgcn_readfirstlane(jc0.y);
            const int lb = lb0;
            // Image items: the j-atom's image position (its load was issued one entry earlier).
            const double4 xjt = IMG ? image_slot(xj, sX) : xj;
            acc_t gx = (acc_t)0, gy = (acc_t)0, gz = (acc_t)0;
#pragma unroll
            for (int c = 0; c < NB_SCC; c++) {
                if (cb & (1 << c)) {
                    const bool inl = __builtin_amdgcn_ubfe(lb, c, 1) != 0;
                    const double4 xi = sI[c * NB_CL + a];
                    const int roff = sT[c * NB_CL + a] + tj;   // byte offset of the pair's plain row
                    const double dx = xi.x - xjt.x, dy = xi.y - xjt.y, dz = xi.z - xjt.z;
                    const double r2 = fma(dx, dx, fma(dy, dy, dz * dz));
                    const bool inr = r2 <= F.r2Off;
                    // Ballots of the two compares (lane masks as they stand), combined in scalar registers; the
                    // active lanes are that mask again (an execution mask, no per-lane recomputation).
                    const unsigned long long mact = __builtin_amdgcn_ballot_w64(inl) & __builtin_amdgcn_ballot_w64(inr);
                    const bool act = __builtin_amdgcn_inverse_ballot_w64(mact);
                    if (mact != 0ull) {
                        if (F64) {
                            // PairwiseInteraction.h:72-119 in FP64 with the reference's operations.
                            if (act) {
                                const int tij = roff >> 6;
                                double x1, x2, dF;
                                pair_fp64(F, r2, xi.w * xjt.w, ljA[tij], ljB[tij], x1, x2, dF);
                                const double f = 2.0 * dF;
                                const double px = f * dx, py = f * dy, pz = f * dz;
                                fx[c] += px; fy[c] += py; fz[c] += pz;
                                gx -= px; gy -= py; gz -= pz;
                                eel += x1;
                                elj += x2;
                            }
                        } else {
                            const bool sw = r2 > F.r2On;
                            const int off = roff + (sw ? 32 : 0);
                            const double4 C = ROWS_LDS ? *(const double4 *)((const unsigned char *)sR + off)
                                                       : load_slot(rows, off >> 5);
                            double fs, x1, x2;
                            pair_rows(F, r2, xi.w * xjt.w, C, sw, mact, ljA, ljB, roff >> 6, fs, x1, x2);
                            // The sums take the active lanes only (the others may hold any value).
                            if (act) {
                                const float f = (float)fs;
                                const float px = f * (float)dx, py = f * (float)dy, pz = f * (float)dz;
                                fx[c] += px; fy[c] += py; fz[c] += pz;
                                gx -= px; gy -= py; gz -= pz;
                                eel += x1;
                                elj += x2;
                            }
                        }
                    }
                }
            }
            gx = sum8(gx);
            gy = sum8(gy);
            gz = sum8(gz);
            const int el = e - w.first;
            if (a < 3) sJF[el * 3 * NB_CL + 3 * b + a] = a == 0 ? gx : (a == 1 ? gy : gz);
            if (lane == 0) sJ[el] = J;
            jc0 = jc1;
            lb0 = lb1;
            jc1 = jc2;
            lb1 = lb2;
            J = Jn;
            xj = xjn;
            tj = tjn;
        }
        // The entries' j-force sums (written by other lanes) are complete.
        wave_region_sync<WPG>();
        bool bad = false;
        if (!IMG) {
            // j-forces of the whole item.
            for (int k = lane; k < 3 * NB_CL * w.count; k += 64) {
                const int el = k / (3 * NB_CL), r = k - el * 3 * NB_CL;
                bad |= fx_atomic_add(facc, n_pad, sJ[el] * NB_CL + r / 3, r % 3, (double)sJF[k]);
            }
        } else {
            // Image side: g'_j (scaled) is rotated back, g_j += R^T g'_j, and the
            // sums h = sum g'_j and G_ab = sum g'_j,a (x_j - origin)_b are kept
            // for the lattice derivatives (SymmetryParameterGradients_ImageDerivatives,
            // pMolecule-1.9.0/extensions/csource/SymmetryParameterGradients.c:142-216).
            // A pure translation needs only h (the moments' derivative term vanishes).
            const bool moments = __builtin_amdgcn_readfirstlane(X[13] == 0.0 ? 1 : 0) != 0;
            double G[12];
#pragma unroll
            for (int k = 0; k < 12; k++) G[k] = 0.0;
            for (int k = lane; k < NB_CL * w.count; k += 64) {
                const int el = k >> 3, jb = k & 7;
                const int sj = sJ[el] * NB_CL + jb;
                const acc_t *f = &sJF[el * 3 * NB_CL + 3 * jb];
                const double gx = isc * (double)f[0], gy = isc * (double)f[1], gz = isc * (double)f[2];
                bad |= fx_atomic_add(facc, n_pad, sj, 0, fma(sX[0], gx, fma(sX[3], gy, sX[6] * gz)));
                bad |= fx_atomic_add(facc, n_pad, sj, 1, fma(sX[1], gx, fma(sX[4], gy, sX[7] * gz)));
                bad |= fx_atomic_add(facc, n_pad, sj, 2, fma(sX[2], gx, fma(sX[5], gy, sX[8] * gz)));
                G[0] += gx; G[1] += gy; G[2] += gz;
                if (moments) {
                    const double4 xr = load_slot(xq, sj);
                    G[3] = fma(gx, xr.x, G[3]); G[4] = fma(gx, xr.y, G[4]); G[5] = fma(gx, xr.z, G[5]);
                    G[6] = fma(gy, xr.x, G[6]); G[7] = fma(gy, xr.y, G[7]); G[8] = fma(gy, xr.z, G[8]);
                    G[9] = fma(gz, xr.x, G[9]); G[10] = fma(gz, xr.y, G[10]); G[11] = fma(gz, xr.z, G[11]);
                }
            }
            long long *ga = D->gacc + (size_t)img * NB_GACC_WORDS_PER_IMAGE +
                            (size_t)(it & (NB_GSLOTS - 1)) * NB_GACC_WORDS_PER_SLOT;
            const int ncomp = moments ? 12 : 3;
#pragma unroll
            for (int k = 0; k < 12; k++) {
                if (k < ncomp) {
                    const double v = wave_sum_f64(G[k]);
                    if (lane == k) bad |= fx_add(&ga[2 * k], &ga[2 * k + 1], v);
                }
            }
        }
        // i-forces: transpose through LDS (over the records and the flushed
        // j-force sums) so that lane 8b+a sums atom a of i-cluster c = b over
        // the eight lanes that held its partials.
        wave_region_sync<WPG>();
#pragma unroll
        for (int c = 0; c < NB_SCC; c++) {
            sF[(3 * c + 0) * NB_SC + lane] = fx[c];
            sF[(3 * c + 1) * NB_SC + lane] = fy[c];
            sF[(3 * c + 2) * NB_SC + lane] = fz[c];
        }
        wave_region_sync<WPG>();
        // Partials of (c, a) sit at lanes 8b' + a, b' = 0..7 (stride 8).
        acc_t ig[3];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const acc_t *fr = &sF[(3 * b + k) * NB_SC + a];
            ig[k] = ((fr[0] + fr[8]) + (fr[16] + fr[24])) + ((fr[32] + fr[40]) + (fr[48] + fr[56]));
        }
        const int si = (I0 + b) * NB_CL + a;
#pragma unroll
        for (int k = 0; k < 3; k++) bad |= fx_atomic_add(facc, n_pad, si, k, IMG ? isc * (double)ig[k] : (double)ig[k]);
        eel = wave_sum_f64(eel);
        elj = wave_sum_f64(elj);
        if (lane < 2) {
            const int slot = it & (NB_ESLOTS - 1);
            const int kind = lane + (IMG ? 2 : 0);
            const double v = lane == 0 ? eel : elj;
            bad |= fx_add(eacc_word(D->eacc, kind, 0, slot), eacc_word(D->eacc, kind, 1, slot), IMG ? isc * v : v);
        }
        raise_error(D->err, bad, NB_ERR_RANGE);
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

__global__ __launch_bounds__(256) void k_force_plain(const SysDesc *__restrict__ descs, AbfsFactors F, ForceGuard guard)
{
#pragma clang fp contract(off)
    if (guard_skip(guard)) return;
    // blockIdx.y = system: one launch covers every system of a batch.
    const SysDesc &D = descs[blockIdx.y];
    if ((int)blockIdx.x >= D.n_plain_blocks) return;
    const int *__restrict__ rowptr = D.rowptr;
    const int *__restrict__ cols = D.cols;
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
// Plain FP64 image pass (MMMMImageEnergy, pMolecule-1.9.0/extensions/csource/NBModelABFS.c:1161-1290, with
// PairwiseInteractionABFS_MMMMEnergy's expression order, PairwiseInteraction.c:292-429): one wavefront per
// image work item, lane l = 8b + a pairing i-atom a of each i-cluster of the super-cluster with j-atom b of
// the entry's j-cluster.  Each item's energies are stored in FP64 and summed in item order by the finalize.  Absolute coordinates; the image coordinate of j is formed as Coordinates3_Rotate +
// _Translate form it (R' x, then + M (t + n)); qi = escale_image * q_i, qij = qi * q_j, A and B times the
// image scale; dF from pair_fp64; i += 2 dF r_ij, j (image side) -= 2 dF r_ij.  j-gradients are summed over
// the 8 lanes of the j-atom, rotated back (R'^T) and added per entry; i-gradients per item; all into the
// fixed-point accumulators, with the energies and the dE/dM sums as in the tiled image instance.
// ------------------------------------------------------------------------------------------------
__device__ __forceinline__ double sum_lanes_xor(double v, int m0, int m1, int m2)
{
    v += __shfl_xor(v, m0, 64);
    v += __shfl_xor(v, m1, 64);
    v += __shfl_xor(v, m2, 64);
    return v;
}

__global__ __launch_bounds__(64) void k_force_image_fp64(const SysDesc *__restrict__ descs,
                                                         const WorkItem *__restrict__ items, AbfsFactors F,
                                                         ForceGuard guard)
{
#pragma clang fp contract(off)
    if (guard_skip(guard)) return;
    const int it = blockIdx.x;
    const WorkItem w = items[it];
    const SysDesc &D = descs[w.sys & 0xffff];
    const int img = (w.sys >> 16) - 1;
    const ImageXf &X = D.imgxf[img];
    const int lane = threadIdx.x, a = lane & 7, b = lane >> 3;
    const int bl = 8 * a + b;
    const double isc = X.scale, esc = X.escale;
    const bool moments = X.translation == 0.0;
    const int nt = D.ntypes;
    double gi[NB_SCC][3];
    int ia[NB_SCC];
    double xi[NB_SCC][3], qi[NB_SCC];
    int ti[NB_SCC];
#pragma unroll
    for (int c = 0; c < NB_SCC; c++) {
        gi[c][0] = gi[c][1] = gi[c][2] = 0.0;
        ia[c] = D.slot_atom[(w.sc * NB_SCC + c) * NB_CL + a];
        const int i = ia[c] >= 0 ? ia[c] : 0;
        xi[c][0] = D.xyz[3 * (size_t)i];
        xi[c][1] = D.xyz[3 * (size_t)i + 1];
        xi[c][2] = D.xyz[3 * (size_t)i + 2];
        qi[c] = esc * D.charge[i];
        ti[c] = nt * D.atype[i];
    }
    double e1 = 0.0, e2 = 0.0, G[12];
    bool bad = false;
#pragma unroll
    for (int k = 0; k < 12; k++) G[k] = 0.0;
    for (int e = w.first; e < w.first + w.count; e++) {
        const Entry &en = D.entries[e];
        const int J = en.J, cb = en.cbits, lb = en.bits[bl];
        const int j = D.slot_atom[J * NB_CL + b];
        const int jj = j >= 0 ? j : 0;
        const double x0 = D.xyz[3 * (size_t)jj], y0 = D.xyz[3 * (size_t)jj + 1], z0 = D.xyz[3 * (size_t)jj + 2];
        const double xj = (X.R[0] * x0 + X.R[1] * y0 + X.R[2] * z0) + X.Tabs[0];
        const double yj = (X.R[3] * x0 + X.R[4] * y0 + X.R[5] * z0) + X.Tabs[1];
        const double zj = (X.R[6] * x0 + X.R[7] * y0 + X.R[8] * z0) + X.Tabs[2];
        const double qj = D.charge[jj];
        const int tj = D.atype[jj];
        double gx = 0.0, gy = 0.0, gz = 0.0;
#pragma unroll
        for (int c = 0; c < NB_SCC; c++) {
            if (!((cb >> c) & 1) || !((lb >> c) & 1) || ia[c] < 0 || j < 0) continue;
            double xij = xi[c][0] - xj, yij = xi[c][1] - yj, zij = xi[c][2] - zj;
            const double r2 = (xij * xij + yij * yij + zij * zij);
            const double qij = qi[c] * qj;
            const int tij = ti[c] + tj;
            double x1, x2, dF;
            if (!pair_fp64(F, r2, qij, D.ljA[tij] * isc, D.ljB[tij] * isc, x1, x2, dF)) continue;
            e1 += x1;
            e2 += x2;
            xij *= (2.0 * dF);
            yij *= (2.0 * dF);
            zij *= (2.0 * dF);
            gi[c][0] += xij;
            gi[c][1] += yij;
            gi[c][2] += zij;
            gx -= xij;
            gy -= yij;
            gz -= zij;
        }
        // The j-atom's image-side gradient: lanes 8b .. 8b+7.
        gx = sum_lanes_xor(gx, 1, 2, 4);
        gy = sum_lanes_xor(gy, 1, 2, 4);
        gz = sum_lanes_xor(gz, 1, 2, 4);
        if (a == 0 && j >= 0) {
            const int sj = J * NB_CL + b;
            // Coordinates3_Rotate by R'^T.
            bad |= fx_atomic_add(D.facc, D.n_pad, sj, 0, X.R[0] * gx + X.R[3] * gy + X.R[6] * gz);
            bad |= fx_atomic_add(D.facc, D.n_pad, sj, 1, X.R[1] * gx + X.R[4] * gy + X.R[7] * gz);
            bad |= fx_atomic_add(D.facc, D.n_pad, sj, 2, X.R[2] * gx + X.R[5] * gy + X.R[8] * gz);
            G[0] += gx;
            G[1] += gy;
            G[2] += gz;
            if (moments) {
                const double rx = x0 - D.origin[0], ry = y0 - D.origin[1], rz = z0 - D.origin[2];
                G[3] += gx * rx; G[4] += gx * ry; G[5] += gx * rz;
                G[6] += gy * rx; G[7] += gy * ry; G[8] += gy * rz;
                G[9] += gz * rx; G[10] += gz * ry; G[11] += gz * rz;
            }
        }
    }
    // i-gradients: lanes a, a+8, ..., a+56.
#pragma unroll
    for (int c = 0; c < NB_SCC; c++) {
        const double vx = sum_lanes_xor(gi[c][0], 8, 16, 32);
        const double vy = sum_lanes_xor(gi[c][1], 8, 16, 32);
        const double vz = sum_lanes_xor(gi[c][2], 8, 16, 32);
        if (b == 0 && ia[c] >= 0) {
            const int si = (w.sc * NB_SCC + c) * NB_CL + a;
            bad |= fx_atomic_add(D.facc, D.n_pad, si, 0, vx);
            bad |= fx_atomic_add(D.facc, D.n_pad, si, 1, vy);
            bad |= fx_atomic_add(D.facc, D.n_pad, si, 2, vz);
        }
    }
    long long *ga = D.gacc + (size_t)img * NB_GACC_WORDS_PER_IMAGE +
                    (size_t)(it & (NB_GSLOTS - 1)) * NB_GACC_WORDS_PER_SLOT;
#pragma unroll
    for (int k = 0; k < 12; k++) {
        const double v = wave_sum_f64(G[k]);
        if (lane == k && (moments || k < 3)) bad |= fx_add(&ga[2 * k], &ga[2 * k + 1], v);
    }
    e1 = wave_sum_f64(e1);
    e2 = wave_sum_f64(e2);
    if (lane == 0) {
        D.eimgpart[2 * (size_t)it] = e1;
        D.eimgpart[2 * (size_t)it + 1] = e2;
    }
    raise_error(D.err, bad || !isfinite(e1) || !isfinite(e2), NB_ERR_RANGE);
}

// ------------------------------------------------------------------------------------------------
// 1-4 list (NBModelABFS_MMMMEnergy's second pass, NBModelABFS.c:276-294): FP64
// with scale14 and the 1-4 LJ table; gradients into the fixed-point arrays.
// ------------------------------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_pairs14(const SysDesc *__restrict__ descs, AbfsFactors F, ForceGuard guard)
{
    const SysDesc &D = descs[blockIdx.y];
    if ((int)blockIdx.x >= D.n14_blocks || guard_skip(guard)) return;
    pairs14_block(D, F, blockIdx.x, threadIdx.x);
}

// ------------------------------------------------------------------------------------------------
// Finalize: gradients (fixed point -> FP64, plus the plain-path rows) added
// into the caller's array (or written over it in assign mode); block 0 of each system also sums the energies in a
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

__global__ __launch_bounds__(256) void k_finalize(const SysDesc *__restrict__ descs, int plain, ForceGuard guard)
{
    if (guard_skip(guard)) return;
    const SysDesc &D = descs[blockIdx.y];
    // The last block of each system sums its energies while the others add
    // the gradients (the two are independent, so neither waits for the other).
    // One thread per slot: the fixed-point words are read in slot order
    // (coalesced) and each atom's row is written through slot_atom.
    if (blockIdx.x + 1 != gridDim.x) {
        const int s = blockIdx.x * 256 + threadIdx.x;
        if (s < D.n_pad && D.grad != nullptr) {
            const int i = D.slot_atom[s];
            if (i >= 0) {
                for (int k = 0; k < 3; k++) {
                    double g = fx_value(D.facc, D.n_pad, s, k);
                    if (plain) g += D.gplain[3 * (size_t)i + k];
                    D.grad[3 * (size_t)i + k] = D.grad_assign ? g : D.grad[3 * (size_t)i + k] + g;
                }
            }
        }
        return;
    }
    double el, lj, el14, lj14, iel = 0.0, ilj = 0.0;
    {
        // Exact: each slot's low word is split at 2^48 (its upper part carried
        // into the high word), the words are summed as integers over the 64
        // slots, the total is brought to its unique form (low word in
        // [0, 2^48)) and converted once, so the result does not depend on which
        // items shared a slot.
        __shared__ double sE[NB_EKINDS];
        if (threadIdx.x < 64 * NB_EKINDS) {
            const int t = threadIdx.x & 63, kind = threadIdx.x >> 6;
            long long h = 0, l = 0;
            if (t < NB_ESLOTS) {
                const long long a0 = *eacc_word(D.eacc, kind, 0, t);
                h = *eacc_word(D.eacc, kind, 1, t) + (a0 >> 48);
                l = a0 & ((1ll << 48) - 1);
            }
#pragma unroll
            for (int o = 32; o >= 1; o >>= 1) {
                h += __shfl_xor(h, o, 64);
                l += __shfl_xor(l, o, 64);
            }
            if (t == 0) {
                // Canonical words of the total (0 <= low < 2^48), then one rounding.
                h += l >> 48;
                l &= (1ll << 48) - 1;
                sE[kind] = (double)h * NB_FX_HI_UNIT + (double)l / NB_FX_LO_SCALE;
            }
        }
        __syncthreads();
        el = sE[0];
        lj = sE[1];
        iel = sE[2];
        ilj = sE[3];
    }
    if (plain) {
        sum_range2(D.eplain, D.n_plain_blocks, el, lj);
        el *= 0.5;
        lj *= 0.5;
        if (D.eimgpart) sum_range2(D.eimgpart, D.n_img_items, iel, ilj);
    }
    sum_range2(D.epart14, D.n14_blocks, el14, lj14);
    double qlj = 0.0, qlj14 = 0.0;
    if (D.qljpart) sum_range2(D.qljpart, D.n_qlj, qlj, qlj14);
    if (threadIdx.x == 0) {
        D.energies[0] = el;
        D.energies[1] = lj;
        D.energies[2] = el14;
        D.energies[3] = lj14;
        D.energies[4] = iel;
        D.energies[5] = ilj;
        D.energies[6] = qlj;
        D.energies[7] = qlj14;
        // The call's error bits travel with its energies; the word starts the next call clear.
        int e = 0;
        if (D.err) {
            e = __hip_atomic_exchange(as_global(D.err), 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            // The plain path's energies are FP64 sums with no fixed-point stage.
            if (plain && !(isfinite(el) && isfinite(lj) && isfinite(iel) && isfinite(ilj))) e |= NB_ERR_RANGE;
        }
        D.energies[NB_ESTATUS] = (double)e;
    }
}

// ------------------------------------------------------------------------------------------------
// QC/MM LJ (NBModelABFS_QCMMEnergyLJ, pMolecule-1.9.0/extensions/csource/NBModelABFS.c:306-378):
// the analytic pair terms with no electrostatics over the QC/MM LJ list (the main LJ table) and
// its 1-4 list (the 1-4 table), NB_QCMM_CHUNKS workgroups per pure QC atom; gradients into the fixed-point
// accumulators, the two energies of each (row, chunk) into qljpart.
// ------------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_qclj(const SysDesc *__restrict__ descs, AbfsFactors F)
{
#pragma clang fp contract(off)
    const SysDesc &D = descs[0];
    const int r = blockIdx.x, c = blockIdx.y, stride = gridDim.y * 256;
    const int i = D.qlj_atom[r];
    const double xi = D.xyz[3 * (size_t)i], yi = D.xyz[3 * (size_t)i + 1], zi = D.xyz[3 * (size_t)i + 2];
    const int ti = D.ntypes * D.atype[i];
    const int si = D.atom_slot[i];
    double e[2] = {0.0, 0.0}, g[3] = {0.0, 0.0, 0.0};
    bool bad = false;
    for (int list = 0; list < 2; list++) {
        const int *ptr = list ? D.qlj14_ptr : D.qlj_ptr;
        const int *idx = list ? D.qlj14_idx : D.qlj_idx;
        const double *tA = list ? D.ljA14 : D.ljA;
        const double *tB = list ? D.ljB14 : D.ljB;
        for (int k = ptr[r] + c * 256 + threadIdx.x; k < ptr[r + 1]; k += stride) {
            const int j = idx[k];
            double xij = xi - D.xyz[3 * (size_t)j], yij = yi - D.xyz[3 * (size_t)j + 1],
                   zij = zi - D.xyz[3 * (size_t)j + 2];
            const double r2 = xij * xij + yij * yij + zij * zij;
            const int tij = ti + D.atype[j];
            double x1, x2, dF;
            if (!pair_fp64(F, r2, 0.0, tA[tij], tB[tij], x1, x2, dF)) continue;
            e[list] += x2;
            xij *= (2.0 * dF);
            yij *= (2.0 * dF);
            zij *= (2.0 * dF);
            g[0] += xij;
            g[1] += yij;
            g[2] += zij;
            const int sj = D.atom_slot[j];
            bad |= fx_atomic_add(D.facc, D.n_pad, sj, 0, -xij);
            bad |= fx_atomic_add(D.facc, D.n_pad, sj, 1, -yij);
            bad |= fx_atomic_add(D.facc, D.n_pad, sj, 2, -zij);
        }
    }
    raise_error(D.err, bad, NB_ERR_RANGE);
    // block_sum2's shared arrays are reused: a barrier after each sum.
    block_sum2<256>(e[0], e[1]);
    __syncthreads();
    double gz = g[2];
    block_sum2<256>(g[0], g[1]);
    __syncthreads();
    double zz = 0.0;
    block_sum2<256>(gz, zz);
    bool bad_i = false;
    if (threadIdx.x == 0) {
        D.qljpart[2 * ((size_t)r * gridDim.y + c)] = e[0];
        D.qljpart[2 * ((size_t)r * gridDim.y + c) + 1] = e[1];
        bad_i |= fx_atomic_add(D.facc, D.n_pad, si, 0, g[0]);
        bad_i |= fx_atomic_add(D.facc, D.n_pad, si, 1, g[1]);
        bad_i |= fx_atomic_add(D.facc, D.n_pad, si, 2, gz);
        bad_i |= !isfinite(e[0]) || !isfinite(e[1]);
    }
    raise_error(D.err, bad_i, NB_ERR_RANGE);
}

void launch_qclj(const SysDesc *d_desc, int nrows, const AbfsFactors &F, hipStream_t s)
{
    if (nrows > 0) hipLaunchKernelGGL(k_qclj, dim3(nrows, NB_QCMM_CHUNKS), dim3(256), 0, s, d_desc, F);
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

template <bool ROWS_LDS, bool IMG, bool F64, int WPG>
static void launch_tiled_instance(dim3 grid, size_t bytes, hipStream_t s, hipEvent_t ev_start, hipEvent_t ev_stop,
                                  const SysDesc *d_desc, const WorkItem *items, QueueParams q, const AbfsFactors &F,
                                  int rows_bytes, int wave_bytes, ForceGuard g, int nitems, int nforce_groups,
                                  int sys1)
{
    hipExtLaunchKernelGGL((k_force_tiled<ROWS_LDS, IMG, F64, WPG>), grid, dim3(64 * WPG), bytes, s, ev_start, ev_stop, 0,
                          d_desc, items, q, F, rows_bytes, wave_bytes, g, nitems, nforce_groups, sys1);
}

void launch_force_tiled(const SysDesc *d_desc, const WorkItem *items, QueueParams q, const AbfsFactors &F,
                        int nblocks, int max_ntypes, int max_count, ForceGuard g, hipStream_t s, bool images,
                        hipEvent_t ev_start, hipEvent_t ev_stop, int n14_blocks, bool full_fp64, int sys1)
{
    if (images) n14_blocks = 0;
    if (nblocks + n14_blocks <= 0) return;
    // One item per wave, four waves per group, unless the items come from the persistent queues or the images.
    const int wpg = (images || q.heads != nullptr) ? 1 : NB_TILED_WPG;
    const int ngroups = nblk(nblocks, wpg);
    const dim3 grid(ngroups + nblk(n14_blocks, wpg));
    const int acc = full_fp64 ? 8 : 4;
    const int wave_bytes = (int)tiled_lds_wave(max_count, acc);
    // The row table goes to LDS when the group's LDS still lets four groups of four waves (or ten single waves) share a CU.
    const size_t limit = wpg == 1 ? NB_LDS_ROW_LIMIT : NB_LDS_GROUP_LIMIT;
    const bool rows_lds = !full_fp64 && tiled_lds_rows(max_ntypes, true) + (size_t)wpg * wave_bytes <= limit;
    const int rows_bytes = (int)tiled_lds_rows(max_ntypes, rows_lds);
    const size_t bytes = rows_bytes + (size_t)wpg * wave_bytes;
#define NB_TILED_ARGS grid, bytes, s, ev_start, ev_stop, d_desc, items, q, F, rows_bytes, wave_bytes, g, nblocks, ngroups, \
                      sys1
    if (full_fp64) {
        // No row table: the FP64 pair terms read the raw A and B of the type pair.
        if (wpg == 1)
            launch_tiled_instance<false, false, true, 1>(NB_TILED_ARGS);
        else
            launch_tiled_instance<false, false, true, NB_TILED_WPG>(NB_TILED_ARGS);
    } else if (images) {
        if (rows_lds)
            launch_tiled_instance<true, true, false, 1>(NB_TILED_ARGS);
        else
            launch_tiled_instance<false, true, false, 1>(NB_TILED_ARGS);
    } else if (wpg == 1) {
        if (rows_lds)
            launch_tiled_instance<true, false, false, 1>(NB_TILED_ARGS);
        else
            launch_tiled_instance<false, false, false, 1>(NB_TILED_ARGS);
    } else {
        if (rows_lds)
            launch_tiled_instance<true, false, false, NB_TILED_WPG>(NB_TILED_ARGS);
        else
            launch_tiled_instance<false, false, false, NB_TILED_WPG>(NB_TILED_ARGS);
    }
#undef NB_TILED_ARGS
}

void launch_force_plain(const SysDesc *d_desc, int nsys, int max_n, const AbfsFactors &F, ForceGuard g, hipStream_t s,
                        hipEvent_t ev_start, hipEvent_t ev_stop)
{
    if (nsys <= 0) return;
    hipExtLaunchKernelGGL(k_force_plain, dim3(max(nblk(max_n, 256), 1), nsys), dim3(256), 0, s, ev_start, ev_stop, 0,
                          d_desc, F, g);
}

void launch_force_image_fp64(const SysDesc *d_desc, const WorkItem *items, int nitems, const AbfsFactors &F,
                             ForceGuard g, hipStream_t s, hipEvent_t ev_start, hipEvent_t ev_stop)
{
    if (nitems > 0)
        hipExtLaunchKernelGGL(k_force_image_fp64, dim3(nitems), dim3(64), 0, s, ev_start, ev_stop, 0, d_desc, items, F, g);
}

void launch_pairs14(const SysDesc *d_desc, int nsys, int max_n14_blocks, const AbfsFactors &F, hipStream_t s,
                    ForceGuard g)
{
    if (nsys > 0 && max_n14_blocks > 0)
        hipLaunchKernelGGL(k_pairs14, dim3(max_n14_blocks, nsys), dim3(64), 0, s, d_desc, F, g);
}

void launch_finalize(const SysDesc *d_desc, int nsys, int max_npad, int plain, hipStream_t s, ForceGuard g)
{
    if (nsys <= 0) return;
    hipLaunchKernelGGL(k_finalize, dim3(nblk(max_npad, 256) + 1, nsys), dim3(256), 0, s, d_desc, plain, g);
}

// ------------------------------------------------------------------------------------------------
// Virial of the last call's gradient (still in the fixed-point accumulators,
// plus the plain path's rows), W_ab = sum_i (x_i - o)_a g_i,b: wave sums in
// FP64 added into fixed point (slot = wave index modulo 64), so the total
// does not depend on which waves shared a slot.  vacc must be zero.
// ------------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_virial(const SysDesc *__restrict__ descs, int plain)
{
    const SysDesc &D = descs[0];
    const int i = blockIdx.x * 256 + threadIdx.x;
    double g[3] = {0.0, 0.0, 0.0}, x[3] = {0.0, 0.0, 0.0};
    if (i < D.n) {
        const int s = D.atom_slot[i];
        for (int k = 0; k < 3; k++) {
            g[k] = fx_value(D.facc, D.n_pad, s, k);
            if (plain) g[k] += D.gplain[3 * (size_t)i + k];
            x[k] = D.xyz[3 * (size_t)i + k] - D.origin[k];
        }
    }
    const int slot = ((blockIdx.x * 4 + (threadIdx.x >> 6)) & (NB_ESLOTS - 1));
#pragma unroll
    for (int ab = 0; ab < 9; ab++) {
        const double w = wave_sum_f64(x[ab / 3] * g[ab % 3]);
        if ((threadIdx.x & 63) == ab && w != 0.0)
            fx_add(&D.vacc[(2 * ab) * NB_ESLOTS + slot], &D.vacc[(2 * ab + 1) * NB_ESLOTS + slot], w);
    }
}

void launch_virial(const SysDesc *d_desc, int n, int plain, hipStream_t s)
{
    if (n > 0) hipLaunchKernelGGL(k_virial, dim3(nblk(n, 256)), dim3(256), 0, s, d_desc, plain);
}
