// nbabfs_internal.h -- shared definitions of the gfx950 NBModelABFS library.
//
// Data layout in HBM (per system):
//   xyz      FP64 n x 3, caller order (the reference Coordinates3 layout).
//   slots    atoms re-ordered into clusters of 8 (CL) along z inside x-y
//            columns; 8 clusters form a super-cluster (SC, 64 atoms = one
//            wavefront).  slot_atom[s] = atom or -1 (padding).
//   xyzq     double4 per slot: position relative to the system origin and the
//            charge times sqrt(Coulomb factor / dielectric); refreshed per call.
//   styp     int per slot: LJ type (0 for padding).
//   entries  one per (SC, j-cluster) pair: j-cluster index plus one byte per
//            lane l = 8a+b whose bit c says that i = atom a of i-cluster c and
//            j = atom b of the j-cluster are on the list.  The set bits are
//            exactly the reference pair list.  Entries of the periodic image
//            lists follow the main list's; their j-cluster is taken from an
//            image (a crystal transformation plus a lattice translation).
//   facc     six int64 arrays of n_pad (x,y,z low words, x,y,z high words):
//            deterministic fixed-point gradient accumulators.
//   eacc     fixed-point Coulomb and LJ energy accumulators of the main and
//            the image lists, 64 slots each.
//   gacc     per image, fixed-point sums of the image-side gradients g'_j
//            (3) and of g'_j (x) (x_j - origin) (9), for dE/dM.
//   vacc     fixed-point virial W_ab = sum_i (x_i - origin)_a g_i,b of the
//            call's gradient, 9 components x 64 slots.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define NB_CL 8           // atoms per cluster
#define NB_SCC 8          // clusters per super-cluster
#define NB_SC (NB_CL * NB_SCC)

// Fixed point: value = hi * 2^16 + lo * 2^-32.
#define NB_FX_HI_UNIT 65536.0
#define NB_FX_LO_SCALE 4294967296.0

// Switching factors, same field set as PairwiseInteractionABFSFactors
// (pMolecule-1.9.0/extensions/cinclude/PairwiseInteraction.h:40-66).
struct AbfsFactors {
    double r2Damp, r2Off, r2On;
    double a, b, c, d, qShift1, qShift2, qF0, qAlpha;
    double aF6, aK12, aShift12, aF0, aAlpha;
    double bF3, bK6, bShift6, bF0, bAlpha;
};

struct Entry {              // 80 bytes
    int J;                  // j-cluster (local index)
    int sc;                 // owning super-cluster
    int cbits;              // bit c set if i-cluster c has any pair with J
    int img;                // -1: main list; else the image whose copy of J is paired
    unsigned char bits[NB_SC];  // per lane l = 8a+b: bit c = pair (c,a)-(J,b) listed
};
static_assert(sizeof(Entry) == 80, "entries are copied as five 16-byte words");

// Real-space transformation of one image for the pair kernel, in the cluster
// frame: x' - origin = R (x - origin) + T.  scale multiplies the image's
// energies and forces (1/2 for an image that is its own inverse, or for every
// image when inverses are not used).  translation != 0 when the fractional
// rotation is the identity: dE/dM then needs only the gradient sum.
struct ImageXf {            // 160 bytes
    double R[9];
    double T[3];
    double scale;
    double translation;
    double Tabs[3];         // T of the plain path: x' = R x + Tabs on absolute coordinates
    double escale;          // plain path: ((1/dielectric) * scale) * Coulomb factor
    double pad[2];
};
// gacc words per image: NB_GSLOTS slots of 12 components (g' sum x3,
// g' (x) x x9) x (low, high); a work item adds into slot (item % NB_GSLOTS).
#define NB_GSLOTS 16
#define NB_GACC_WORDS_PER_SLOT 24
#define NB_GACC_WORDS_PER_IMAGE (NB_GSLOTS * NB_GACC_WORDS_PER_SLOT)

struct WorkItem {           // 16 bytes
    int sys;                // system index in the batch; image-list items: | (image + 1) << 16
    int sc;                 // super-cluster
    int first;              // first entry
    int count;              // entries
};
static_assert(sizeof(WorkItem) == 16, "work items are moved as one 16-byte word");

// Everything a per-call kernel needs to know about one system.
struct SysDesc {
    const double *xyz;          // n*3 current coordinates (device)
    double *grad;               // n*3 gradients to increment (device) or null
    const int *slot_atom;       // n_pad
    const int *atom_slot;       // n
    const float4 *centre;       // ncl
    double4 *xyzq;              // n_pad
    int *styp;                  // n_pad: LJ type (0 for padding)
    const int *atype;           // n: LJ types
    const double *charge;       // n
    const double *ljA, *ljB;    // ntypes*ntypes FP64 (main list)
    const double *ljA14, *ljB14;// ntypes*ntypes FP64 (1-4 list)
    const double *ljrow;        // 2*ntypes*ntypes rows of 4 FP64: [(ti*ntypes+tj)*2 + switched]
    const Entry *entries;
    const int *p14;             // 2*n14 (i, j)
    long long *facc;            // 6*n_pad
    double *gplain;             // n*3 plain-path gradients (or null)
    long long *eacc;            // NB_EACC_WORDS fixed-point energy accumulators (tiled path)
    const ImageXf *imgxf;       // per image transformation (null: no images)
    long long *gacc;            // NB_GACC_WORDS_PER_IMAGE per image
    long long *vacc;            // NB_VACC_WORDS fixed-point virial accumulators
    double *epart14;            // 2 per 1-4 block
    double *eplain;             // 2 per plain block
    double *energies;           // NB_NENERGIES outputs (device)
    double origin[3];           // FP64 frame origin of the cluster coordinates
    double escale;              // Coulomb factor / dielectric
    double escale14;            // times electrostaticScale14
    double qscale;              // sqrt(escale): charges are stored pre-scaled
    int n, n_pad, ncl, ntypes, n14;
    int item_first, item_count; // range of this system's items in the batch
    int n_plain_blocks, n14_blocks;
    int nimg;                   // images of the image table (gacc length / NB_GACC_WORDS_PER_IMAGE)
    int n_img_items;            // plain path: image work items (eimgpart holds 2 per item)
    double *eimgpart;           // plain path: FP64 image energies per work item, summed in order
    // QC/MM LJ: rows = pure QC atoms, CSR lists of MM atoms (main and 1-4).
    int n_qlj;
    const int *qlj_atom, *qlj_ptr, *qlj_idx, *qlj14_ptr, *qlj14_idx;
    double *qljpart;            // 2 per (row, chunk) (QC/MM LJ, QC/MM 1-4 LJ), summed in that order
    const int *rowptr, *cols;   // plain path: full CSR neighbour rows (n + 1, 2 npairs)
    int grad_assign;            // 1: the finalize pass overwrites grad instead of adding to it
    int *err;                   // error word (NB_ERR_* bits) of the current call; cleared by the finalize pass
};

// ------------------------------------------------------------------------------------------------
// Device helpers.
// ------------------------------------------------------------------------------------------------
// Pointers loaded from a descriptor are generic; accessing memory through
// address-space-1 pointers makes the compiler emit global_* operations (vmcnt
// only) instead of flat_* ones, which would also hold every LDS wait behind
// outstanding atomics.
#define GAS __attribute__((address_space(1)))
template <class T>
__device__ __forceinline__ GAS T *as_global(T *p)
{
    return (GAS T *)p;
}

// rint(x) as int64 for |x| < 2^51: x + 1.5*2^52 puts the rounded integer in
// the low mantissa bits.
__device__ __forceinline__ long long rint_i64(double x)
{
    const double magic = 6755399441055744.0;
    return __double_as_longlong(x + magic) - __double_as_longlong(magic);
}

// v = hi*2^16 + lo*2^-32 with hi = round(v/2^16) and the remainder |r| <= 2^15
// (exact) stored as round(r*2^32).  |v| is clamped to 2^66, far beyond any
// physical gradient or energy, and NaN is mapped to a finite value, so the
// integer sums stay defined; the return value says that v was outside
// (-2^66, 2^66) or NaN, and the callers raise the system's error word
// (NB_ERR_RANGE) so that the energy call reports it instead of a wrong value.
__device__ __forceinline__ bool fx_split(double v, long long &hi, long long &lo)
{
    const double magic = 6755399441055744.0;
    const bool bad = !(fabs(v) < 0x1p66);
    v = fmin(fmax(v, -0x1p66), 0x1p66);
    const double t = v * (1.0 / NB_FX_HI_UNIT) + magic;
    hi = __double_as_longlong(t) - __double_as_longlong(magic);
    const double r = fma(-(t - magic), NB_FX_HI_UNIT, v);
    lo = rint_i64(r * NB_FX_LO_SCALE);
    return bad;
}

__device__ __forceinline__ bool fx_add(long long *lo_word, long long *hi_word, double v)
{
    long long hi, lo;
    const bool bad = fx_split(v, hi, lo);
    if (lo != 0) __hip_atomic_fetch_add(as_global(lo_word), lo, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (hi != 0) __hip_atomic_fetch_add(as_global(hi_word), hi, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return bad;
}

__device__ __forceinline__ bool fx_atomic_add(long long *facc, int n_pad, int slot, int comp, double v)
{
    return fx_add(&facc[(size_t)comp * n_pad + slot], &facc[(size_t)(comp + 3) * n_pad + slot], v);
}

// Error word of a system (SysDesc::err), raised by the per-call kernels and
// reported by the energy calls (then cleared by the finalize pass):
// NB_ERR_COORD -- a coordinate is NaN or infinite; NB_ERR_RANGE -- a
// gradient or energy contribution was NaN or beyond the fixed-point range.
#define NB_ERR_COORD 1
#define NB_ERR_RANGE 2

// One wave-level report: the first lane of a wave with any flagged lane ORs
// the bits into *err (a vector atomic; every system owns one word).
__device__ __forceinline__ void raise_error(int *err, bool flagged, int bits)
{
    if (err && __any(flagged) && (threadIdx.x & 63) == 0)
        __hip_atomic_fetch_or(as_global(err), bits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ double fx_value(const long long *facc, int n_pad, int slot, int comp)
{
    return (double)facc[(size_t)(comp + 3) * n_pad + slot] * NB_FX_HI_UNIT +
           (double)facc[(size_t)comp * n_pad + slot] * (1.0 / NB_FX_LO_SCALE);
}

// Outputs of one system: MM/MM Elect., LJ, 1-4 Elect., 1-4 LJ, Image Elect., Image LJ, QC/MM LJ,
// QC/MM 1-4 LJ, then the call's error bits (NB_ERR_*, as a double) at NB_ESTATUS.
#define NB_NENERGIES 9
#define NB_ESTATUS 8

// Energy accumulators: [kind][lo word, hi word][slot], kinds 0 = Coulomb and
// 1 = LJ of the main list, 2 and 3 of the image lists; a work item adds into
// slot (item % NB_ESLOTS).
#define NB_ESLOTS 64
#define NB_EKINDS 4
#define NB_EACC_WORDS (2 * NB_EKINDS * NB_ESLOTS)
// Virial accumulators: [component ab = 3a+b][lo word, hi word][slot].
#define NB_VACC_WORDS (9 * 2 * NB_ESLOTS)

__device__ __forceinline__ long long *eacc_word(long long *eacc, int kind, int hi, int slot)
{
    return &eacc[(2 * kind + hi) * NB_ESLOTS + slot];
}

__device__ __forceinline__ double wave_sum_f64(double v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ float wave_sum_f32(float v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Squared smallest and largest distances between a point and grid cell n
// (RegularGrid_DistancesSquaredBetweenCellAndPoint, pCore-1.9.0/extensions/
// csource/RegularGrid.c:139-180, non-periodic branch, in its operation order:
// the cell spans [n h + lower, n h + lower + h] on each axis; the per-axis
// squares are summed in axis order).
__device__ __forceinline__ void point_cell_r2(const double *lower, double h, double x, double y, double z, int nx,
                                              int ny, int nz, double &minR2, double &maxR2)
{
#pragma clang fp contract(off)
    const double pt[3] = {x, y, z};
    const int nn[3] = {nx, ny, nz};
    double mn2[3], mx2[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const double lo = (double)nn[d] * h + lower[d];
        const double hi = lo + h;
        const double v = pt[d];
        const double dl = v - lo, du = hi - v;
        const double mx = v > hi ? dl : (v < lo ? du : (dl > du ? dl : du));
        const double mn = v > hi ? v - hi : (v < lo ? lo - v : 0.0);
        mn2[d] = mn * mn;
        mx2[d] = mx * mx;
    }
    minR2 = (mn2[0] + mn2[1]) + mn2[2];
    maxR2 = (mx2[0] + mx2[1]) + mx2[2];
}
