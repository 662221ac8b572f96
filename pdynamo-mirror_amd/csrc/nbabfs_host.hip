// nbabfs_host.hip -- host driver and C ABI of libnbabfs.so.
//
// The state object plays the role of NBModelABFSState (pMolecule-1.9.0/
// extensions/cinclude/NBModelABFSState.h:30-125): a per-configuration cache of
// device lists and reference coordinates, rebuilt on demand by the rule of
// NBModelABFS_Update (pMolecule-1.9.0/extensions/csource/NBModelABFS.c:508-623).
#include "nbabfs.h"
#include "nbabfs_kernels.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

// ------------------------------------------------------------------------------------------------
// Errors and device buffers.
// ------------------------------------------------------------------------------------------------
static thread_local std::string g_err;

static int fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIPCHK(x)                                                                                      \
    do {                                                                                               \
        hipError_t e_ = (x);                                                                           \
        if (e_ != hipSuccess) return fail(NBABFS_STATUS_LOGIC_ERROR, "%s: %s", #x, hipGetErrorString(e_)); \
    } while (0)

#define CHK(x)                                                                                         \
    do {                                                                                               \
        int s_ = (x);                                                                                  \
        if (s_ != NBABFS_STATUS_CONTINUE) return s_;                                                   \
    } while (0)

// The error bits a call's finalize pass stored beside its energies, as a status.
static int call_status(int bits, const char *what)
{
    if (bits == 0) return NBABFS_STATUS_CONTINUE;
    return fail(NBABFS_STATUS_VALUE_ERROR, "%s: %s%s%s", what, (bits & NB_ERR_COORD) ? "a coordinate is NaN or infinite" : "",
                (bits & NB_ERR_COORD) && (bits & NB_ERR_RANGE) ? "; " : "",
                (bits & NB_ERR_RANGE) ? "a gradient or energy contribution is NaN or beyond +-2^66" : "");
}

// A device buffer that grows on demand.  The device descriptors and batch
// layouts hold its address; they are compared field by field with what the
// current buffers give before every launch (ensure_desc, desc_current), so a
// reallocation is never launched on with a stale address.
template <class T>
struct DBuf {
    T *p = nullptr;
    size_t cap = 0;
    // Grow with 25% headroom so lists that fluctuate between rebuilds rarely reallocate.
    int ensure(size_t m)
    {
        if (p && m <= cap) return NBABFS_STATUS_CONTINUE;
        release();
        size_t mm = std::max<size_t>(m + m / 4, 16);
        if (hipMalloc((void **)&p, mm * sizeof(T)) != hipSuccess) {
            p = nullptr;
            (void)hipGetLastError();
            return fail(NBABFS_STATUS_OUT_OF_MEMORY, "hipMalloc of %zu bytes failed", mm * sizeof(T));
        }
        cap = mm;
        return NBABFS_STATUS_CONTINUE;
    }
    int upload(const T *h, size_t m, hipStream_t s)
    {
        CHK(ensure(m));
        if (m) HIPCHK(hipMemcpyAsync(p, h, m * sizeof(T), hipMemcpyHostToDevice, s));
        return NBABFS_STATUS_CONTINUE;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
    void swap(DBuf &o)
    {
        std::swap(p, o.p);
        std::swap(cap, o.cap);
    }
    ~DBuf() { release(); }
};

template <class T>
struct PinBuf {
    T *p = nullptr;
    size_t cap = 0;
    int ensure(size_t m)
    {
        if (p && m <= cap) return NBABFS_STATUS_CONTINUE;
        release();
        size_t mm = std::max<size_t>(m, 16);
        if (hipHostMalloc((void **)&p, mm * sizeof(T), hipHostMallocDefault) != hipSuccess) {
            p = nullptr;
            (void)hipGetLastError();
            return fail(NBABFS_STATUS_OUT_OF_MEMORY, "hipHostMalloc of %zu bytes failed", mm * sizeof(T));
        }
        cap = mm;
        return NBABFS_STATUS_CONTINUE;
    }
    void release()
    {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        cap = 0;
    }
    ~PinBuf() { release(); }
};

// Coherent pinned host memory that kernels write directly (rebuild flags):
// the host reads it after the stream has synchronised, with no copy.
struct MappedInts {
    int *h = nullptr;
    int *d = nullptr;
    size_t cap = 0;
    int ensure(size_t m)
    {
        if (h && m <= cap) return NBABFS_STATUS_CONTINUE;
        release();
        size_t mm = std::max<size_t>(m, 16);
        if (hipHostMalloc((void **)&h, mm * sizeof(int), hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess ||
            hipHostGetDevicePointer((void **)&d, h, 0) != hipSuccess) {
            (void)hipGetLastError();
            release();
            return fail(NBABFS_STATUS_OUT_OF_MEMORY, "mapped host allocation of %zu bytes failed", mm * sizeof(int));
        }
        std::memset(h, 0, mm * sizeof(int));
        cap = mm;
        return NBABFS_STATUS_CONTINUE;
    }
    void release()
    {
        if (h) (void)hipHostFree(h);
        h = d = nullptr;
        cap = 0;
    }
    ~MappedInts() { release(); }
};

// Epochs written by the update test; never 0, the initial flag value.
static inline int next_epoch(int e) { return e == 0x7fffffff ? 1 : e + 1; }

static inline int nblk(size_t m, int b) { return (int)((m + b - 1) / b); }

// ------------------------------------------------------------------------------------------------
// Reference constants and factors.
// ------------------------------------------------------------------------------------------------
// Coulomb factor 1e7 N_A e^2/(4 pi eps0) (pCore-1.9.0/extensions/cinclude/Units.h:18-30,48-49).
static double coulomb_factor()
{
    const double na = 6.0221415e+23, e = 1.60217653e-19, eps0 = 8.854187817e-12;
    return (1.0e+7 * na * e * e) / (4.0e+00 * M_PI * eps0);
}

// PairwiseInteractionABFS_MakeFactors (pMolecule-1.9.0/extensions/csource/PairwiseInteraction.c:89-141).
static void make_factors(double damp, double inner, double outer, AbfsFactors &F)
{
#pragma clang fp contract(off)
    const double r2d = damp * damp, r2off = outer * outer, r2on = inner * inner;
    F.r2Damp = r2d;
    F.r2Off = r2off;
    F.r2On = r2on;
    const double gamma = std::pow(r2off - r2on, 3);
    F.a = r2off * r2off * (r2off - 3.0 * r2on) / gamma;
    F.b = 6.0 * r2off * r2on / gamma;
    F.c = -(r2off + r2on) / gamma;
    F.d = 0.4 / gamma;
    F.qShift1 = 8.0 * (r2off * r2on * (outer - inner) - 0.2 * (outer * r2off * r2off - inner * r2on * r2on)) / gamma;
    F.qShift2 = -(F.a / outer) + F.b * outer + F.c * outer * r2off + F.d * outer * r2off * r2off;
    {   // Coulomb: value and slope of s + qShift1 matched at the damping radius.
        const double v = 1.0 / damp + F.qShift1, g = -1.0 / r2d;
        F.qF0 = v - 0.5 * damp * g;
        F.qAlpha = -0.5 * g / damp;
    }
    F.aF6 = 1.0 / (r2off * r2off * r2off);
    F.aK12 = std::pow(r2off, 3) / (std::pow(r2off, 3) - std::pow(r2on, 3));
    F.aShift12 = 1.0 / std::pow(inner * outer, 6);
    {
        const double s12 = 1.0 / std::pow(damp, 12), v = s12 - F.aShift12, g = -12.0 * s12 / damp;
        F.aF0 = v - 0.5 * damp * g;
        F.aAlpha = -0.5 * g / damp;
    }
    F.bF3 = 1.0 / (outer * r2off);
    F.bK6 = (outer * r2off) / (outer * r2off - inner * r2on);
    F.bShift6 = 1.0 / std::pow(inner * outer, 3);
    {
        const double s6 = 1.0 / std::pow(damp, 6), v = -s6 + F.bShift6, g = 6.0 * s6 / damp;
        F.bF0 = v - 0.5 * damp * g;
        F.bAlpha = -0.5 * g / damp;
    }
}

// ------------------------------------------------------------------------------------------------
// Device bookkeeping.
// ------------------------------------------------------------------------------------------------
static int g_cu_count = 0;
static int g_inited = 0;

extern "C" int nbabfs_init(void)
{
    if (g_inited) return NBABFS_STATUS_CONTINUE;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        (void)hipGetLastError();
        return fail(NBABFS_STATUS_LOGIC_ERROR, "no HIP device available (libnbabfs has no CPU path)");
    }
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, dev));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(NBABFS_STATUS_LOGIC_ERROR, "device %d is %s; libnbabfs is built for gfx950", dev, prop.gcnArchName);
    g_cu_count = prop.multiProcessorCount;
    g_inited = 1;
    return NBABFS_STATUS_CONTINUE;
}

extern "C" const char *nbabfs_last_error(void) { return g_err.c_str(); }
extern "C" int nbabfs_device_cu_count(void) { return g_cu_count; }

extern "C" void nbabfs_default_options(nbabfs_options *o)
{
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->dampingCutoff = 0.5;
    o->innerCutoff = 8.0;
    o->outerCutoff = 12.0;
    o->listCutoff = 13.5;
    o->dielectric = 1.0;
    o->electrostaticScale14 = 1.0;
    o->cutoffCellSizeFactor = 0.5;
    o->minimumCellExtent = 2;
    o->minimumPoints = 500;
    o->precision = NBABFS_MIXED;
    o->entriesPerItem = 16;
    o->wavesPerCU = 0;
    o->imageEntriesPerItem = 16;
    o->useGridByCell = 1;
}

// A system's items in a batch, followed by empty items up to `padded` (whole workgroups of the pair launch).
__global__ void k_remap_items(const WorkItem *__restrict__ src, int count, int padded, int sys,
                              WorkItem *__restrict__ dst)
{
    int k = blockIdx.x * 256 + threadIdx.x;
    if (k < count) {
        WorkItem w = src[k];
        w.sys = sys;
        dst[k] = w;
    } else if (k < padded) {
        dst[k] = WorkItem{sys, 0, 0, 0};
    }
}

// Totals of a list build: entries, work items, and the search's capacity
// need (0 when every super-cluster fitted), written to mapped host memory.
__global__ void k_list_totals(const int *__restrict__ eoff, const int *__restrict__ counts, const int *__restrict__ ioff,
                              const int *__restrict__ nits, int nsc, const int *__restrict__ need,
                              int *__restrict__ out)
{
    if (threadIdx.x == 0) {
        volatile int *o = out;
        o[0] = eoff[nsc - 1] + counts[nsc - 1];
        o[1] = ioff[nsc - 1] + nits[nsc - 1];
        o[2] = *need;
    }
}

struct CheckDesc {
    const double *xyz, *ref;
    const unsigned char *free_flags;
    int n;
    double bufsq;
};

// Update test of a batch, queued in front of its passes (and captured with
// them).  guard (device words): k_batch_epoch advances guard[1], the epoch of
// this test, and clears guard[0]; the test sets guard[0] to the epoch when
// some system must be rebuilt, and the passes behind it then do nothing.  The
// host reads the epoch (flags[nsys]) and the systems' flags (flags[k] == epoch:
// rebuild) from mapped memory after the call's one synchronisation.
__global__ void k_batch_epoch(int *__restrict__ guard, int *__restrict__ flags, int nsys)
{
    if (threadIdx.x == 0) {
        const int e = guard[1];
        const int n = (e <= 0 || e == 0x7fffffff) ? 1 : e + 1;
        guard[1] = n;
        guard[0] = 0;
        *(volatile int *)&flags[nsys] = n;
    }
}

__global__ __launch_bounds__(256) void k_check_batch(const CheckDesc *__restrict__ cd, int *__restrict__ flags,
                                                     int *__restrict__ guard)
{
#pragma clang fp contract(off)
    const CheckDesc &C = cd[blockIdx.y];
    const int epoch = guard[1];
    int i = blockIdx.x * 256 + threadIdx.x;
    bool up = false;
    if (C.ref && i < C.n && (!C.free_flags || (C.free_flags[i] & 1))) {
        double dx = C.xyz[3 * (size_t)i] - C.ref[3 * (size_t)i];
        double dy = C.xyz[3 * (size_t)i + 1] - C.ref[3 * (size_t)i + 1];
        double dz = C.xyz[3 * (size_t)i + 2] - C.ref[3 * (size_t)i + 2];
        up = dx * dx + dy * dy + dz * dz > C.bufsq;
    }
    if (__any(up) && threadIdx.x % 64 == 0) {
        *(volatile int *)&flags[blockIdx.y] = epoch;
        *(volatile int *)&guard[0] = epoch;
    }
}

// Largest displacement of a free atom since the last list build (CheckForUpdate's maximumDisplacement).
__global__ __launch_bounds__(256) void k_max_displacement(const double *__restrict__ xyz, const double *__restrict__ ref,
                                                          const unsigned char *__restrict__ free_flags, int n,
                                                          unsigned long long *__restrict__ out)
{
#pragma clang fp contract(off)
    const int i = blockIdx.x * 256 + threadIdx.x;
    double r2 = 0.0;
    if (i < n && (!free_flags || (free_flags[i] & 1))) {
        const double dx = xyz[3 * (size_t)i] - ref[3 * (size_t)i];
        const double dy = xyz[3 * (size_t)i + 1] - ref[3 * (size_t)i + 1];
        const double dz = xyz[3 * (size_t)i + 2] - ref[3 * (size_t)i + 2];
        r2 = dx * dx + dy * dy + dz * dz;
    }
    for (int o = 32; o >= 1; o >>= 1) r2 = fmax(r2, __shfl_xor(r2, o, 64));
    // Non-negative doubles order as their bit patterns.
    if ((threadIdx.x & 63) == 0 && r2 > 0.0) atomicMax(out, (unsigned long long)__double_as_longlong(r2));
}

// ------------------------------------------------------------------------------------------------
// The state.
// ------------------------------------------------------------------------------------------------
struct nbabfs_state {
    nbabfs_options opt;
    int n = 0, ntypes = 0, n14 = 0;
    AbfsFactors F;
    double escale = 0, escale14 = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    std::vector<long long> p14_keys;   // canonical 1-4 set (after the MM and free selections)
    std::vector<long long> p14_all;    // canonical 1-4 set as given
    std::vector<unsigned char> fixed_h;
    std::vector<int> xptr_h, xidx_h;   // exclusions CSR (host copy)
    std::vector<double> charges_h;
    const double *upd_xyz = nullptr;   // coordinates of the last update

    // Constant device data.
    DBuf<double> charge, ljA, ljB, ljA14, ljB14, ljrow;
    DBuf<int> atype, xptr, xidx, p14;
    // Atom flags: bit 0 = free (not fixed), bit 1 = pure QC atom (outside the MM selection).
    // has_flags: the list search reads them; has_fixed: some atom is fixed (displacement test).
    DBuf<unsigned char> freef;
    bool has_fixed = false, has_flags = false;

    // Coordinates.
    DBuf<double> xyz_stage, xyz_ref, grad_stage;
    bool have_ref = false;

    // Update state (NBModelABFSState.isNew / useGridSearch).
    bool isNew = true, useGrid = false, listValid = false, csrValid = false;
    long long calls = 0, updates = 0, npairs = 0;
    // The pair count of the last build reaches h_np without a synchronisation; it is read (and added to the
    // statistics when np_stat) at the next synchronisation that needs it.
    PinBuf<unsigned long long> h_np;
    bool np_pending = false, np_stat = false;
    std::vector<unsigned char> stencil_h;   // stencil flags last uploaded
    PinBuf<int> h_qs;                       // staging of the queue starts
    hipEvent_t ev_qs = nullptr, ev_desc = nullptr;   // completion of the last staging copies
    double sumPairs = 0.0;

    // List.
    BuildGrid grid;
    int u = 0;
    DBuf<unsigned char> stencil;
    int n_pad = 0, ncl = 0, nsc = 0, nent = 0, nitems = 0, cap = 512;
    DBuf<int> slot_atom, atom_slot;
    DBuf<int4> cellv;
    DBuf<float4> blo, bhi, centre;
    DBuf<unsigned long long> npblock;   // pairs per super-cluster of the last build
    DBuf<int> unsorted;                 // bucket contents in scatter order
    DBuf<double4> srec;      // slot-ordered search records (self-list search)
    DBuf<int4> scell;
    DBuf<Entry> entries, scratch;
    DBuf<WorkItem> items, items2;
    DBuf<int> icost, ihist, icur;
    DBuf<int> rowptr, cols;

    // Per call.
    DBuf<double4> xyzq;
    DBuf<int> styp;
    DBuf<long long> facc;
    DBuf<double> epart14, eplain, gplain, energies, eimgpart;
    DBuf<long long> eacc, vacc;
    DBuf<int> errw;                // error word of the per-call kernels (NB_ERR_* bits)
    bool err_dirty = false;        // a prepare pass ran since the last finalize pass read errw
    DBuf<SysDesc> d_desc;
    DBuf<int> heads, qstarts;
    PinBuf<SysDesc> h_desc;
    PinBuf<double> h_energies, h_grad;
    SysDesc desc_shadow;           // the descriptor last uploaded to d_desc (valid when desc_valid)
    bool desc_valid = false;
    long items_version = 0;        // bumped whenever the work items are cut or ordered anew
    int grad_assign = 0;           // NBABFS_GRADIENT_ASSIGN: energy calls overwrite the gradient
    const double *last_xyz = nullptr;
    double *last_grad = nullptr;
    // Coordinates for which the last update already queued the prepare pass
    // and the pair kernel (the energy call on the same array only finalizes).
    const double *forced_xyz = nullptr;
    MappedInts upflag, tail, prepdone;
    DBuf<int> dflag;
    DBuf<unsigned long long> prepcount;
    int epoch = 0;

    // Build temporaries.
    DBuf<int> key, count, off, cursor, sorted, padded, colstart, scantmp, counts, ncand, eoff, nits, ioff, flag;
    DBuf<int2> colcl;
    DBuf<double> bbpart, bbox;
    DBuf<unsigned long long> u64;
    PinBuf<double> h_bbox;
    PinBuf<int> h_int;

    // Crystal symmetry: the transformations given at set-up (fractional, 9 + 3
    // per item, with the identity and rotational inverses) and the crystal
    // parameters of the current call (NBModelABFSState_SetUp's transformations
    // and NBModelABFSState_Initialize's symmetryParameters).
    bool has_sym = false, sp_set = false;
    int ntrans = 0, identity = -1, checkInv = 1, expandFactor = 0;
    std::vector<double> trot, ttra;
    std::vector<int> tinv;
    double sp[6] = {0, 0, 0, 0, 0, 0}, M[9], Minv[9];
    double Mref[9];                 // M of the last image-list build (referenceSymmetryParameters)
    double Mxf[9];                  // M of the uploaded image transformations
    bool xf_valid = false;
    // Image slots: the (transformation, a, b, c) searched at the last build in
    // the reference's order; the image list proper is the slots with pairs.
    struct ImageSlot {
        int a, b, c, t;
        double scale;
        long long npairs;
    };
    std::vector<ImageSlot> slots;
    int nimages = 0, nent_main = 0, nitems_main = 0;
    long long img_npairs = 0;
    double sumImages = 0.0, sumImagePairs = 0.0;
    DBuf<double> itc, ixyz, rtbuf, ibb;
    DBuf<float4> iblo, ibhi, icen;
    DBuf<float> ibox;
    DBuf<QcmmGrid> qgrid;        // QC/MM cross lists: grids of the electrostatic and LJ lists
    DBuf<double> qgpart;         // their per-workgroup boxes
    DBuf<int> qoff;              // first list position of each (row, column chunk) of a cross list
    DBuf<int4> qcolcell;         // QC/MM electrostatic list: cells of its columns
    DBuf<int4> icell;            // image lists: primary-grid search cells of the i atoms
    DBuf<ImageGrid> igrid;       // image lists: conforming grid of each image
    DBuf<int> istep_off, ibound, isoff, icounts, incand, ieoff, inits, iioff;
    DBuf<ImageStep> isteps;
    DBuf<Entry> iscratch, entries_tmp;
    DBuf<unsigned long long> inpairs;
    DBuf<ImageXf> imgxf;
    DBuf<long long> gacc;
    PinBuf<double> h_ibb;
    PinBuf<ImageXf> h_imgxf;
    PinBuf<long long> h_gacc;
    PinBuf<unsigned long long> h_inpairs;

    // QC/MM (NBModelABFSState_SetUp with a QCAtomContainer): the QC atoms in
    // ascending atom order, link (boundary) atoms with their QC partner, link
    // factor and MM partners; the MM sites of the QC/MM electrostatics (the
    // atoms for MM coupling; the pure MM atoms plus the redistributed charges
    // for RC/RD); the lists of the last build and the electrostatic spline.
    bool has_qc = false;
    int nqc = 0, coupling = 0, linkRatio = 0, nsites = 0, nlj = 0;
    std::vector<unsigned char> qc_pure;          // per atom: pure QC atom
    std::vector<int> qc_index_h, qc_partner_h, lj_atom_h;
    DBuf<int> qc_index, qc_partner, site_a, site_p, qex_ptr, qex_idx, el14_ptr, el14_idx, lj_atom, lj14_ptr,
        lj14_idx, el_ptr, el_idx, lj_ptr, lj_idx, qcnt;
    DBuf<double> qc_lf, site_q, spx, spy, sph, qpot, qcq, qgrad, qljpart, qxyz_out;
    DBuf<long long> qgacc;
    PinBuf<double> h_qpot, h_qgrad;
    int nspline = 0;
    long long n_el = 0, n_lj = 0, n_el14 = 0, n_lj14 = 0;
    double sumQCMMElPairs = 0.0, sumQCMMLJPairs = 0.0;

    // Kernel timing.
    bool timing = false;
    std::vector<hipEvent_t> ev;
    size_t ev_used = 0;
    std::vector<int> ev_group;     // events per timed pass (2 or 4)
    double force_ms = 0.0;
    long long force_launches = 0;

    // Second stream for the image-list launch, which runs beside the main
    // list's launch (fork and join events on the state's stream).
    hipStream_t stream2 = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;

    ~nbabfs_state()
    {
        for (auto e : ev) (void)hipEventDestroy(e);
        if (ev_fork) (void)hipEventDestroy(ev_fork);
        if (ev_join) (void)hipEventDestroy(ev_join);
        if (ev_qs) (void)hipEventDestroy(ev_qs);
        if (ev_desc) (void)hipEventDestroy(ev_desc);
        if (stream2) (void)hipStreamDestroy(stream2);
        if (own_stream && stream) (void)hipStreamDestroy(stream);
    }
};

// Sum of timed force launches recorded so far (synchronises on the last event).
// Timed pair passes: each holds a group of 2 events (start, stop of one
// dispatch or of a same-stream sequence) or 4 (two dispatches on two streams,
// start/stop each); its time is the span from the first start to the last
// stop.  The events are the runtime's dispatch timestamps (hipExtLaunchKernel),
// so the time is that of the kernels themselves, not of the gaps before them.
static int collect_timing(nbabfs_state *S)
{
    if (S->ev_used == 0) return NBABFS_STATUS_CONTINUE;
    HIPCHK(hipDeviceSynchronize());
    size_t k = 0;
    for (int g : S->ev_group) {
        float t1 = 0.0f;
        HIPCHK(hipEventElapsedTime(&t1, S->ev[k], S->ev[k + 1]));
        float lo = 0.0f, hi = t1;
        if (g == 4) {
            float t2 = 0.0f, t3 = 0.0f;
            HIPCHK(hipEventElapsedTime(&t2, S->ev[k], S->ev[k + 2]));
            HIPCHK(hipEventElapsedTime(&t3, S->ev[k], S->ev[k + 3]));
            lo = std::min(lo, t2);
            hi = std::max(hi, t3);
        }
        S->force_ms += hi - lo;
        S->force_launches++;
        k += g;
    }
    S->ev_used = 0;
    S->ev_group.clear();
    return NBABFS_STATUS_CONTINUE;
}

// Reserve the events of one timed pass (2 or 4).
static int timing_events(nbabfs_state *S, int count, hipEvent_t *out)
{
    if (S->ev_used + count > S->ev.size()) {
        if (S->ev.size() >= 4096) {
            CHK(collect_timing(S));
        } else {
            for (int k = 0; k < 64; k++) {
                hipEvent_t e;
                HIPCHK(hipEventCreate(&e));
                S->ev.push_back(e);
            }
        }
    }
    for (int k = 0; k < count; k++) out[k] = S->ev[S->ev_used++];
    S->ev_group.push_back(count);
    return NBABFS_STATUS_CONTINUE;
}

// ------------------------------------------------------------------------------------------------
// Grid parameters (host FP64, reference operation order).
// ------------------------------------------------------------------------------------------------
// The box of a list build: k_bbox_partial turns a NaN or infinite coordinate
// into an infinite box.  A finite box may span at most 1e4 A per axis: the
// search culls with FP32 cluster boxes in the frame of the box minimum,
// padded by 1e-3 A, which covers FP32 rounding (half an ulp, <= 4.9e-4 A)
// only for coordinates below 2^14 A in that frame.
static int check_box(const double *bb)
{
    for (int d = 0; d < 3; d++) {
        if (!std::isfinite(bb[d]) || !std::isfinite(bb[3 + d]))
            return fail(NBABFS_STATUS_VALUE_ERROR, "a coordinate is NaN or infinite (list build)");
        if (bb[3 + d] - bb[d] > 1.0e4)
            return fail(NBABFS_STATUS_VALUE_ERROR, "coordinates span %.6g A in axis %d (at most 1e4 A)", bb[3 + d] - bb[d], d);
    }
    return NBABFS_STATUS_CONTINUE;
}

// PairListGenerator_DetermineMethod (pCore-1.9.0/extensions/csource/PairListGenerator.c:498-525).
static bool determine_method(const nbabfs_options &o, int n, const double *bb)
{
    double mext = 0.0;
    for (int d = 0; d < 3; d++) mext = std::max(mext, std::fabs(bb[3 + d] - bb[d]));
    const double cell = o.cutoffCellSizeFactor * o.listCutoff;
    const int mcells = (int)std::ceil(mext / cell);
    return (mext > o.cutoffCellSizeFactor * o.listCutoff) && (mcells > o.minimumCellExtent) && (n > o.minimumPoints);
}

// Coordinates3_MakeGrid (pCore-1.9.0/extensions/csource/Coordinates3.c:1111-1158).
static void reference_grid(const double *bb, double h, BuildGrid &g)
{
#pragma clang fp contract(off)
    for (int d = 0; d < 3; d++) {
        const double l0 = bb[d];
        const double extent = bb[3 + d] - bb[d];
        const double upper = l0 + extent;
        const double mid = 0.5 * (l0 + upper);
        const double cells = std::ceil(extent / h);
        g.bins[d] = (int)cells;
        g.lower[d] = mid - 0.5 * (cells * h);
    }
    g.h = h;
}

// RegularGrid_MakeSearchRange / DistancesSquaredBetweenCells (non-periodic),
// pCore-1.9.0/extensions/csource/RegularGrid.c:188-234,526-575: flag 2 when
// the whole cell pair is within the cutoff, 1 when it must be tested, 0 never.
static void make_stencil(double h, double cutoff, int &u, std::vector<unsigned char> &flags)
{
#pragma clang fp contract(off)
    u = (int)std::ceil(cutoff / h);
    const int w = 2 * u + 1;
    const double c2 = cutoff * cutoff;
    flags.assign((size_t)w * w * w, 0);
    for (int a = -u; a <= u; a++)
        for (int b = -u; b <= u; b++)
            for (int c = -u; c <= u; c++) {
                const int idx[3] = {a, b, c};
                double maxR2 = 0.0, minR2 = 0.0;
                for (int d = 0; d < 3; d++) {
                    const double delta = std::fabs((double)(0 - idx[d]));
                    const double mx = (delta + 1.0) * h;
                    const double mn = std::max((delta - 1.0) * h, 0.0);
                    maxR2 += mx * mx;
                    minR2 += mn * mn;
                }
                if (minR2 < c2) flags[((size_t)(a + u) * w + (b + u)) * w + (c + u)] = (maxR2 <= c2) ? 2 : 1;
            }
}

// ------------------------------------------------------------------------------------------------
// Crystal symmetry (host FP64 in the reference's operation order): 3x3
// matrices row-major.  pCore-1.9.0/extensions/csource/Matrix33.c,
// Transformation3.c, Transformation3Container.c; pMolecule-1.9.0/extensions/
// csource/SymmetryParameters.c.
// ------------------------------------------------------------------------------------------------
static double m33_det(const double *a)
{
#pragma clang fp contract(off)
    return a[0] * (a[4] * a[8] - a[7] * a[5]) - a[1] * (a[3] * a[8] - a[6] * a[5]) +
           a[2] * (a[3] * a[7] - a[6] * a[4]);
}

// Matrix33_Invert: cofactors, then a scale by 1/det.
static void m33_invert(const double *a, double *r)
{
#pragma clang fp contract(off)
    const double det = m33_det(a);
    r[0] = (a[4] * a[8] - a[5] * a[7]);
    r[1] = (a[2] * a[7] - a[8] * a[1]);
    r[2] = (a[1] * a[5] - a[4] * a[2]);
    r[3] = (a[5] * a[6] - a[3] * a[8]);
    r[4] = (a[0] * a[8] - a[2] * a[6]);
    r[5] = (a[2] * a[3] - a[0] * a[5]);
    r[6] = (a[3] * a[7] - a[4] * a[6]);
    r[7] = (a[1] * a[6] - a[0] * a[7]);
    r[8] = (a[0] * a[4] - a[3] * a[1]);
    const double sc = 1.0 / det;
    for (int k = 0; k < 9; k++) r[k] *= sc;
}

// Matrix33_PreMultiplyBy: self = o * self.
static void m33_premul(double *self, const double *o)
{
#pragma clang fp contract(off)
    double s[9], r[9];
    std::memcpy(s, self, sizeof(s));
    for (int i = 0; i < 3; i++) {
        const double o0 = o[3 * i], o1 = o[3 * i + 1], o2 = o[3 * i + 2];
        r[3 * i] = s[0] * o0 + s[3] * o1 + s[6] * o2;
        r[3 * i + 1] = s[1] * o0 + s[4] * o1 + s[7] * o2;
        r[3 * i + 2] = s[2] * o0 + s[5] * o1 + s[8] * o2;
    }
    std::memcpy(self, r, sizeof(r));
}

// Matrix33_PostMultiplyBy: self = self * o.
static void m33_postmul(double *self, const double *o)
{
#pragma clang fp contract(off)
    double s[9], r[9];
    std::memcpy(s, self, sizeof(s));
    for (int i = 0; i < 3; i++) {
        const double o0 = o[i], o1 = o[3 + i], o2 = o[6 + i];
        r[i] = s[0] * o0 + s[1] * o1 + s[2] * o2;
        r[3 + i] = s[3] * o0 + s[4] * o1 + s[5] * o2;
        r[6 + i] = s[6] * o0 + s[7] * o1 + s[8] * o2;
    }
    std::memcpy(self, r, sizeof(r));
}

// Matrix33_ApplyToVector3.
static void m33_apply(const double *m, double *v)
{
#pragma clang fp contract(off)
    const double x = v[0], y = v[1], z = v[2];
    v[0] = x * m[0] + y * m[1] + z * m[2];
    v[1] = x * m[3] + y * m[4] + z * m[5];
    v[2] = x * m[6] + y * m[7] + z * m[8];
}

static bool m33_equal_tol(const double *a, const double *b)
{
    for (int k = 0; k < 9; k++)
        if (std::fabs(a[k] - b[k]) > 1.0e-6) return false;
    return true;
}

// SymmetryParameters_MakeM: the columns of M are the lattice vectors.
static void symmetry_make_m(const double *sp, double *M, double *Minv)
{
#pragma clang fp contract(off)
    const double dr = M_PI / 180.0e+00;
    const double al = sp[3] * dr, be = sp[4] * dr, ga = sp[5] * dr;
    const double ca = std::cos(al), cb = std::cos(be), cg = std::cos(ga), sg = std::sin(ga);
    for (int k = 0; k < 9; k++) M[k] = 0.0;
    M[0] = sp[0];
    M[1] = sp[1] * cg;
    M[4] = sp[1] * sg;
    M[2] = sp[2] * cb;
    M[5] = sp[2] * (ca - cb * cg) / sg;
    M[8] = sp[2] * std::sqrt(1.0e+00 - ca * ca - cb * cb - cg * cg + 2.0e+00 * ca * cb * cg) / sg;
    m33_invert(M, Minv);
}

// Transformation3_Orthogonalize: R' = M R M^-1, T' = M t.
static void orthogonalize(const double *M, const double *Minv, const double *rot, const double *tr, double *Ro,
                          double *To)
{
    std::memcpy(Ro, rot, 9 * sizeof(double));
    m33_premul(Ro, M);
    m33_postmul(Ro, Minv);
    To[0] = tr[0];
    To[1] = tr[1];
    To[2] = tr[2];
    m33_apply(M, To);
}

// GetLimits of SymmetryParameters_FindBoxSearchLimits.
static bool get_limits(double bl, double bu, double il, double iu, double t, int *low, int *high)
{
#pragma clang fp contract(off)
    int n = 0;
    while (iu >= bl) { il -= t; iu -= t; n--; }
    while (iu < bl) { il += t; iu += t; n++; }
    if (il <= bu) {
        *low = n;
        while (il <= bu) { il += t; iu += t; n++; }
        *high = n - 1;
        return true;
    }
    return false;
}

// SymmetryParameters_FindBoxSearchLimits: lim = alow, ahigh, blow, bhigh, clow, chigh.
static void box_search_limits(const double *M, const double *lower, const double *upper, const double *ilower,
                              const double *iupper, int *lim)
{
#pragma clang fp contract(off)
    lim[0] = lim[2] = lim[4] = 0;
    lim[1] = lim[3] = lim[5] = -1;
    if (!(M[8] > 0.0 && M[4] > 0.0 && M[0] > 0.0)) return;
    if (!get_limits(lower[2], upper[2], ilower[2], iupper[2], M[8], &lim[4], &lim[5])) return;
    double d1 = lim[4] * M[5], d2 = lim[5] * M[5];
    double bl = lower[1] - std::max(d1, d2), bu = upper[1] - std::min(d1, d2);
    if (!get_limits(bl, bu, ilower[1], iupper[1], M[4], &lim[2], &lim[3])) return;
    d1 = lim[4] * M[2];
    d2 = lim[5] * M[2];
    bl = lower[0] - std::max(d1, d2);
    bu = upper[0] - std::min(d1, d2);
    d1 = lim[2] * M[1];
    d2 = lim[3] * M[1];
    bl -= std::max(d1, d2);
    bu -= std::min(d1, d2);
    get_limits(bl, bu, ilower[0], iupper[0], M[0], &lim[0], &lim[1]);
}

// Round (pCore-1.9.0/extensions/cinclude/Macros.h:34).
static int ref_round(double a) { return a >= 0 ? (int)(a + 0.5) : (int)(a - 0.5); }

// Transformation3Container_FindInverseIntegerTranslation.
static void inverse_integer_translation(const double *rot, const double *tr, int t, int tinv, int a, int b, int c,
                                        int *inv)
{
#pragma clang fp contract(off)
    double v[3] = {tr[3 * t] + (double)a, tr[3 * t + 1] + (double)b, tr[3 * t + 2] + (double)c};
    m33_apply(rot + 9 * tinv, v);
    for (int i = 0; i < 3; i++) {
        v[i] *= -1.0e+00;
        v[i] += -1.0e+00 * tr[3 * tinv + i];
        const int r = ref_round(v[i]);
        inv[i] = (std::fabs(v[i] - (double)r) < 1.0e-4) ? r : -999999;
    }
}

// Matrix33_InverseDerivative: d(M^-1)/dM_ij.
static void inverse_derivative(const double *m, int i, int j, double *o)
{
#pragma clang fp contract(off)
    const double m00 = m[0], m01 = m[1], m02 = m[2], m10 = m[3], m11 = m[4], m12 = m[5], m20 = m[6], m21 = m[7],
                 m22 = m[8];
    double ddet = 0.0;
    m33_invert(m, o);
    switch (3 * i + j) {
    case 0: ddet = -m12 * m21 + m11 * m22; break;
    case 1: ddet = m12 * m20 - m10 * m22; break;
    case 2: ddet = -m11 * m20 + m10 * m21; break;
    case 3: ddet = m02 * m21 - m01 * m22; break;
    case 4: ddet = -m02 * m20 + m00 * m22; break;
    case 5: ddet = m01 * m20 - m00 * m21; break;
    case 6: ddet = -m02 * m11 + m01 * m12; break;
    case 7: ddet = m02 * m10 - m00 * m12; break;
    default: ddet = -m01 * m10 + m00 * m11; break;
    }
    for (int k = 0; k < 9; k++) o[k] *= -ddet;
    switch (3 * i + j) {
    case 0: o[4] += m22; o[5] -= m12; o[7] -= m21; o[8] += m11; break;
    case 1: o[1] -= m22; o[2] += m12; o[7] += m20; o[8] -= m10; break;
    case 2: o[1] += m21; o[2] -= m11; o[4] -= m20; o[5] += m10; break;
    case 3: o[3] -= m22; o[5] += m02; o[6] += m21; o[8] -= m01; break;
    case 4: o[0] += m22; o[2] -= m02; o[6] -= m20; o[8] += m00; break;
    case 5: o[0] -= m21; o[2] += m01; o[3] += m20; o[5] -= m00; break;
    case 6: o[3] += m12; o[4] -= m02; o[6] -= m11; o[7] += m01; break;
    case 7: o[0] -= m12; o[1] += m02; o[6] += m10; o[7] -= m00; break;
    default: o[0] += m11; o[1] -= m01; o[3] -= m10; o[4] += m00; break;
    }
    const double sc = 1.0 / m33_det(m);
    for (int k = 0; k < 9; k++) o[k] *= sc;
}

// Work items longest-first (and, for the persistent queues, interleaved over
// the eight per-XCD queues), then the queue bounds.  The main list's items
// [0, nitems_main) and the image lists' items after them are ordered apart:
// they run in two launches.
// The pair count of the last build (synchronises the stream when it is still pending).
static int resolve_npairs(nbabfs_state *S)
{
    if (!S->np_pending) return NBABFS_STATUS_CONTINUE;
    HIPCHK(hipStreamSynchronize(S->stream));
    S->npairs = (long long)S->h_np.p[0];
    S->np_pending = false;
    if (S->np_stat) S->sumPairs += (double)S->npairs;
    S->np_stat = false;
    return NBABFS_STATUS_CONTINUE;
}

// Counts an update in the statistics (its pairs once the count is known).
static void count_update(nbabfs_state *S)
{
    S->updates++;
    if (S->np_pending) S->np_stat = true;
    else S->sumPairs += (double)S->npairs;
}

// Wait until the staging copy recorded by *ev has completed (an event, not the whole stream).
static int staging_free(hipEvent_t ev)
{
    if (ev) HIPCHK(hipEventSynchronize(ev));
    return NBABFS_STATUS_CONTINUE;
}

static int staging_done(hipEvent_t *ev, hipStream_t st)
{
    if (!*ev) HIPCHK(hipEventCreateWithFlags(ev, hipEventDisableTiming));
    HIPCHK(hipEventRecord(*ev, st));
    return NBABFS_STATUS_CONTINUE;
}

static int order_items(nbabfs_state *S)
{
    hipStream_t st = S->stream;
    const int n = S->nitems_main, ni = S->nitems - S->nitems_main;
    const int interleave = S->opt.wavesPerCU > 0 ? 1 : 0;
    std::vector<int> qs(9);
    if (n + ni > 0) {
        const int maxcost = NB_SCC * std::max(1, S->opt.entriesPerItem);
        CHK(S->icost.ensure(std::max(n, ni)));
        CHK(S->ihist.ensure(maxcost + 1));
        CHK(S->icur.ensure(maxcost + 1));
        CHK(S->items2.ensure(n + ni));
        CHK(S->scantmp.ensure(scan_tmp_size(maxcost + 1)));
        launch_order_items(S->items.p, n, S->entries.p, maxcost, interleave, S->icost.p, S->ihist.p, S->icur.p,
                           S->scantmp.p, S->items2.p, st);
        launch_order_items(S->items.p + n, ni, S->entries.p, maxcost, 0, S->icost.p, S->ihist.p, S->icur.p,
                           S->scantmp.p, S->items2.p + n, st);
        HIPCHK(hipGetLastError());
        S->items.swap(S->items2);
    }
    S->items_version++;
    qs[0] = 0;
    for (int q = 0; q < 8; q++) qs[q + 1] = interleave ? qs[q] + (n - q + 7) / 8 * (n > q) : (int)((long long)n * (q + 1) / 8);
    // Pinned staging: the copy is asynchronous; the staging is rewritten only after it completed.
    CHK(S->h_qs.ensure(9));
    CHK(S->qstarts.ensure(9));
    CHK(staging_free(S->ev_qs));
    std::memcpy(S->h_qs.p, qs.data(), 9 * sizeof(int));
    HIPCHK(hipMemcpyAsync(S->qstarts.p, S->h_qs.p, 9 * sizeof(int), hipMemcpyHostToDevice, st));
    CHK(staging_done(&S->ev_qs, st));
    return NBABFS_STATUS_CONTINUE;
}

static int build_qcmm_lists(nbabfs_state *S, const double *d_xyz);

// The error word collects the NB_ERR_* bits of the kernels since the last
// finalize pass, which reports and clears them.  Before a prepare pass, bits
// left by an abandoned one (an update with no energy call after it, the
// autotuner's passes) are cleared, so that they are not reported with the
// next call; on the usual update-then-energy sequence this costs nothing.
static void errw_before_prepare(nbabfs_state *S)
{
    if (S->err_dirty) launch_fill_int(S->errw.p, 1, 0, S->stream);
    S->err_dirty = true;
}

static void errw_clear(nbabfs_state *S)
{
    launch_fill_int(S->errw.p, 1, 0, S->stream);
    S->err_dirty = false;
}

// ------------------------------------------------------------------------------------------------
// List build (NBModelABFSState_GridInitialize + GenerateLists).
// ------------------------------------------------------------------------------------------------
static int build_list(nbabfs_state *S, const double *d_xyz)
{
    hipStream_t st = S->stream;
    const int n = S->n;
    const nbabfs_options &o = S->opt;
    CHK(resolve_npairs(S));
    S->listValid = false;
    S->csrValid = false;
    S->npairs = 0;
    S->nent = S->nent_main = 0;
    S->nitems = S->nitems_main = 0;
    S->slots.clear();
    S->nimages = 0;
    S->img_npairs = 0;
    errw_clear(S);      // a rebuild discards the passes queued on the old list
    if (n <= 0) {
        S->n_pad = S->ncl = S->nsc = 0;
        S->listValid = true;
        return NBABFS_STATUS_CONTINUE;
    }
    // Bounding box.
    const int nparts = std::min(nblk(n, 256), 1024);
    CHK(S->bbpart.ensure(6 * (size_t)nparts));
    CHK(S->bbox.ensure(6));
    CHK(S->h_bbox.ensure(6));
    launch_bbox(d_xyz, n, S->bbpart.p, nparts, S->bbox.p, st);
    HIPCHK(hipMemcpyAsync(S->h_bbox.p, S->bbox.p, 6 * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    double bb[6];
    std::memcpy(bb, S->h_bbox.p, sizeof(bb));
    CHK(check_box(bb));

    // Grid vs direct is decided once, on the first build (NBModelABFSState.c:212).
    if (S->isNew) S->useGrid = determine_method(o, n, bb);
    BuildGrid g;
    std::memset(&g, 0, sizeof(g));
    g.use_grid = S->useGrid ? 1 : 0;
    std::vector<unsigned char> flags;
    const double cell = o.cutoffCellSizeFactor * o.listCutoff;
    if (S->useGrid) {
        reference_grid(bb, cell, g);
        // A zero extent gives an axis with no cells, which RegularGridOccupancy_Allocate
        // rejects (pCore-1.9.0/extensions/csource/RegularGridOccupancy.c:29,54).
        if (g.bins[0] <= 0 || g.bins[1] <= 0 || g.bins[2] <= 0)
            return fail(NBABFS_STATUS_INVALID_ARGUMENT, "the cell grid has an axis with no cells (zero extent %s)",
                        g.bins[0] <= 0 ? "in x" : (g.bins[1] <= 0 ? "in y" : "in z"));
        make_stencil(cell, o.listCutoff, S->u, flags);
    } else {
        g.h = 1.0;
        S->u = 0;
        flags.assign(1, 1);
    }
    if (flags != S->stencil_h || !S->stencil.p) {
        CHK(S->stencil.upload(flags.data(), flags.size(), st));
        HIPCHK(hipStreamSynchronize(st));     // the copy reads pageable memory
        S->stencil_h = flags;
    }

    // Super-cluster columns: width so that 64 consecutive atoms of a column
    // (one super-cluster, split into 2 x 2 x 2 clusters) fill a rough cube.
    double ext[3], vol = 1.0;
    for (int d = 0; d < 3; d++) {
        ext[d] = bb[3 + d] - bb[d];
        vol *= std::max(ext[d], 1.0);
        g.corg[d] = bb[d];
    }
    double cw = std::cbrt((double)NB_SC * vol / (double)n);
    cw = std::max(cw, 3.0);
    long long maxkeys = 8LL * n + 4096;
    // A sparse system spread over a large box (atoms far apart) gets wider
    // columns, so that the column and key arrays stay O(n).
    while (std::ceil(ext[0] / cw) * std::ceil(ext[1] / cw) > (double)maxkeys) cw *= 2.0;
    g.cw = cw;
    g.ncx = std::max(1, (int)std::ceil(ext[0] / cw));
    g.ncy = std::max(1, (int)std::ceil(ext[1] / cw));
    double dz = 0.125 * cw;
    if ((long long)g.ncx * g.ncy * (long long)std::ceil(ext[2] / dz + 1.0) > maxkeys)
        dz = ext[2] / std::max(1.0, (double)(maxkeys / ((long long)g.ncx * g.ncy)) - 1.0);
    if (!(dz > 0.0)) dz = 1.0;
    g.dz = dz;
    g.nzb = std::max(1, (int)std::ceil(ext[2] / dz));
    S->grid = g;
    const int ncol = g.ncx * g.ncy;
    const int nkeys = ncol * g.nzb;

    // Bin, counting sort, deterministic bucket order, column padding.
    CHK(S->cellv.ensure(n));
    CHK(S->key.ensure(n));
    CHK(S->count.ensure(nkeys));
    CHK(S->off.ensure(nkeys));
    CHK(S->cursor.ensure(std::max(nkeys, n + 1)));
    CHK(S->sorted.ensure(n));
    CHK(S->padded.ensure(ncol));
    CHK(S->colstart.ensure(ncol));
    CHK(S->colcl.ensure(ncol));
    CHK(S->scantmp.ensure(scan_tmp_size(std::max(std::max(nkeys, n + 1), ncol)) + 1024));
    CHK(S->h_int.ensure(16));
    launch_fill_int(S->count.p, nkeys, 0, st);
    launch_bin(d_xyz, n, g, S->cellv.p, S->key.p, S->count.p, st);
    launch_scan(S->count.p, S->off.p, nkeys, S->scantmp.p, st);
    HIPCHK(hipMemcpyAsync(S->cursor.p, S->off.p, nkeys * sizeof(int), hipMemcpyDeviceToDevice, st));
    CHK(S->unsorted.ensure(n));
    launch_sort_slots(S->key.p, n, nkeys, ncol, g.nzb, S->off.p, S->cursor.p, S->unsorted.p, S->sorted.p, S->padded.p,
                      st);
    launch_scan(S->padded.p, S->colstart.p, ncol, S->scantmp.p, st);
    HIPCHK(hipMemcpyAsync(S->h_int.p, S->colstart.p + (ncol - 1), sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(S->h_int.p + 1, S->padded.p + (ncol - 1), sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    const int total_padded = S->h_int.p[0] + S->h_int.p[1];
    S->n_pad = (total_padded + NB_SC - 1) / NB_SC * NB_SC;
    S->ncl = S->n_pad / NB_CL;
    S->nsc = S->ncl / NB_SCC;
    CHK(S->slot_atom.ensure(S->n_pad));
    CHK(S->atom_slot.ensure(n));
    launch_fill_int(S->slot_atom.p, S->n_pad, -1, st);
    launch_slots(S->sorted.p, S->key.p, n, g.nzb, S->off.p, S->colstart.p, S->padded.p, ncol, d_xyz, S->nsc,
                 S->slot_atom.p, S->atom_slot.p, S->colcl.p, st);
    CHK(S->blo.ensure(S->ncl));
    CHK(S->bhi.ensure(S->ncl));
    CHK(S->centre.ensure(S->ncl));
    launch_cluster_geom(d_xyz, S->slot_atom.p, S->ncl, g.corg, S->blo.p, S->bhi.p, S->centre.p, st);
    CHK(S->srec.ensure(S->n_pad));
    CHK(S->scell.ensure(S->n_pad));
    const int pointcell = (g.use_grid && !o.useGridByCell) ? 1 : 0;
    launch_slot_records(d_xyz, S->slot_atom.p, S->n_pad, S->cellv.p, S->has_flags ? S->freef.p : nullptr, g.use_grid,
                        pointcell, S->u, S->srec.p, S->scell.p, st);

    // Cluster-pair search with the exact reference predicate.
    SearchParams P;
    std::memset(&P, 0, sizeof(P));
    P.cutoff2 = o.listCutoff * o.listCutoff;
    P.stencil = S->stencil.p;
    P.u = S->u;
    P.use_grid = g.use_grid;
    P.xptr = S->xptr.p;
    P.xidx = S->xidx.p;
    P.free_flags = S->has_flags ? S->freef.p : nullptr;
    P.reach = (float)(o.listCutoff + 0.01);
    P.cw = (float)cw;
    P.ncx = g.ncx;
    P.ncy = g.ncy;
    P.srec = S->srec.p;
    P.scell = S->scell.p;
    P.zoff = S->off.p;
    P.dz = (float)g.dz;
    P.nzb = g.nzb;
    P.ncol = ncol;
    P.n = n;
    P.pointcell = pointcell;
    for (int d = 0; d < 3; d++) P.glower[d] = g.lower[d];
    P.gh = g.h;
    CHK(S->counts.ensure(S->nsc));
    CHK(S->ncand.ensure(S->nsc));
    // Search, then the entry and item offsets; the totals and the capacity
    // check reach the host through mapped memory in one synchronisation.
    const int per_item = std::max(1, o.entriesPerItem > 0 ? o.entriesPerItem : 16);
    CHK(S->eoff.ensure(S->nsc + 1));
    CHK(S->nits.ensure(S->nsc));
    CHK(S->ioff.ensure(S->nsc + 1));
    CHK(S->tail.ensure(4));
    for (int attempt = 0; attempt < 4; attempt++) {
        P.cap = S->cap;
        CHK(S->scratch.ensure((size_t)S->nsc * S->cap));
        CHK(S->flag.ensure(1));
        launch_fill_int(S->flag.p, 1, 0, st);
        launch_search(P, S->nsc, d_xyz, S->slot_atom.p, S->blo.p, S->bhi.p, S->colcl.p, S->scratch.p, S->ncand.p,
                      S->counts.p, S->flag.p, st);
        if (!launch_list_offsets(S->counts.p, S->nsc, per_item, S->eoff.p, S->nits.p, S->ioff.p, S->flag.p, S->tail.d,
                                 st)) {
            launch_scan(S->counts.p, S->eoff.p, S->nsc, S->scantmp.p, st);
            launch_items_per_sc(S->counts.p, S->nsc, per_item, S->nits.p, st);
            launch_scan(S->nits.p, S->ioff.p, S->nsc, S->scantmp.p, st);
            hipLaunchKernelGGL(k_list_totals, dim3(1), dim3(64), 0, st, S->eoff.p, S->counts.p, S->ioff.p, S->nits.p,
                               S->nsc, S->flag.p, S->tail.d);
        }
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(st));
        const int need = ((volatile int *)S->tail.h)[2];
        if (need <= 0) break;
        if (attempt == 3) return fail(NBABFS_STATUS_LOGIC_ERROR, "pair search capacity did not converge");
        S->cap = need + need / 8 + 64;
    }
    S->nent = ((volatile int *)S->tail.h)[0];
    S->nitems = ((volatile int *)S->tail.h)[1];
    S->nent_main = S->nent;
    S->nitems_main = S->nitems;
    S->desc_valid = false;
    S->slots.clear();
    S->nimages = 0;
    S->img_npairs = 0;
    CHK(S->entries.ensure(S->nent));
    CHK(S->items.ensure(S->nitems));
    CHK(S->u64.ensure(2));
    CHK(S->npblock.ensure(S->nsc));
    launch_compact(S->scratch.p, S->cap, nullptr, S->ncand.p, S->counts.p, S->eoff.p, S->ioff.p, S->nsc, S->nsc,
                   per_item, 0, S->entries.p, S->items.p, S->u64.p, S->npblock.p, st);
    CHK(S->h_np.ensure(1));
    HIPCHK(hipMemcpyAsync(S->h_np.p, S->u64.p, sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
    S->np_pending = true;

    // Per-call buffers.
    CHK(S->xyzq.ensure(S->n_pad));
    CHK(S->styp.ensure(S->n_pad));
    CHK(S->facc.ensure(6 * (size_t)S->n_pad));
    CHK(S->eacc.ensure(NB_EACC_WORDS));
    CHK(S->vacc.ensure(NB_VACC_WORDS));
    CHK(S->energies.ensure(NB_NENERGIES));
    CHK(order_items(S));
    CHK(S->heads.ensure(8 * 32));

    // Plain path: full CSR rows (both directions).
    if (o.precision == NBABFS_PLAIN_FP64) {
        CHK(resolve_npairs(S));
        CHK(S->rowptr.ensure(n + 1));
        CHK(S->cols.ensure(std::max<size_t>(2 * (size_t)S->npairs, 1)));
        CHK(S->gplain.ensure(3 * (size_t)n));
        CHK(S->eplain.ensure(2 * (size_t)nblk(n, 256)));
        launch_fill_int(S->cursor.p, n + 1, 0, st);
        launch_csr_degree(S->entries.p, S->nent_main, S->slot_atom.p, S->cursor.p, st);
        launch_scan(S->cursor.p, S->rowptr.p, n + 1, S->scantmp.p, st);
        HIPCHK(hipMemcpyAsync(S->cursor.p, S->rowptr.p, (n + 1) * sizeof(int), hipMemcpyDeviceToDevice, st));
        launch_csr_fill(S->entries.p, S->nent_main, S->slot_atom.p, S->cursor.p, S->cols.p, st);
        launch_csr_sort_rows(S->rowptr.p, n, S->cols.p, st);
        S->csrValid = true;
    }
    HIPCHK(hipGetLastError());
    if (S->has_qc) CHK(build_qcmm_lists(S, d_xyz));
    S->listValid = true;
    return NBABFS_STATUS_CONTINUE;
}

static void fill_desc(nbabfs_state *S, SysDesc &D, const double *d_xyz, double *d_grad);

// Upload the state's descriptor when its content changed (a buffer was
// reallocated, the list's sizes or the coordinate/gradient arrays changed).
// The staging copy is reused only after the stream drained.
static int ensure_desc(nbabfs_state *S, const double *d_xyz, double *d_grad)
{
    CHK(S->d_desc.ensure(1));
    CHK(S->h_desc.ensure(1));
    SysDesc D;
    fill_desc(S, D, d_xyz, d_grad);
    S->last_xyz = d_xyz;
    S->last_grad = d_grad;
    if (S->desc_valid && std::memcmp(&D, &S->desc_shadow, sizeof(SysDesc)) == 0) return NBABFS_STATUS_CONTINUE;
    S->desc_valid = false;
    CHK(staging_free(S->ev_desc));
    S->h_desc.p[0] = D;
    HIPCHK(hipMemcpyAsync(S->d_desc.p, S->h_desc.p, sizeof(SysDesc), hipMemcpyHostToDevice, S->stream));
    CHK(staging_done(&S->ev_desc, S->stream));
    S->desc_shadow = D;
    S->desc_valid = true;
    return NBABFS_STATUS_CONTINUE;
}

// The uploaded descriptor describes the state's current buffers and sizes.
static bool desc_current(nbabfs_state *S)
{
    if (!S->desc_valid) return false;
    SysDesc D;
    fill_desc(S, D, S->last_xyz, S->last_grad);
    return std::memcmp(&D, &S->desc_shadow, sizeof(SysDesc)) == 0;
}

static int ensure_call_buffers(nbabfs_state *S)
{
    CHK(S->energies.ensure(NB_NENERGIES));
    if (S->has_qc) CHK(S->qljpart.ensure(2 * (size_t)std::max(S->nlj, 1) * NB_QCMM_CHUNKS));
    if (S->opt.precision == NBABFS_PLAIN_FP64) CHK(S->eimgpart.ensure(2 * (size_t)std::max(S->nitems - S->nitems_main, 1)));
    CHK(S->eacc.ensure(NB_EACC_WORDS));
    CHK(S->vacc.ensure(NB_VACC_WORDS));
    CHK(S->epart14.ensure(2 * (size_t)std::max(nblk(S->n14, NB_P14_BLOCK), 1)));
    return NBABFS_STATUS_CONTINUE;
}

static int launch_force(nbabfs_state *S, ForceGuard g);

// NBModelABFS_Update (NBModelABFS.c:508-623) for device coordinates.  When a
// list exists, the displacement test runs inside the prepare pass, and the
// pair kernel is queued right behind it, guarded by the test's device flag,
// before the host waits for the test: the GPU does not idle while the host
// decides.  If the list must be rebuilt, the guarded launch does nothing and
// the energy call prepares and evaluates again on the new list.
static int upload_image_xf(nbabfs_state *S);
static int build_images(nbabfs_state *S, const double *d_xyz);
static bool check_for_image_update(const nbabfs_state *S, double maxdisp);

static int update(nbabfs_state *S, const double *d_xyz, int *rebuilt)
{
    hipStream_t st = S->stream;
    S->upd_xyz = d_xyz;
    bool doUpdate = S->isNew || !S->have_ref || !S->listValid;
    bool doImages = false;
    S->forced_xyz = nullptr;
    if (S->has_sym && !S->sp_set)
        return fail(NBABFS_STATUS_INVALID_ARGUMENT, "crystal symmetry without symmetry parameters");
    // With images and a changed cell, CheckForImageUpdate needs the largest
    // displacement before anything is launched: no speculative pair kernel.
    const bool cellCheck = S->has_sym && S->nimages > 0 && std::memcmp(S->M, S->Mref, sizeof(S->M)) != 0;
    if (!doUpdate && S->has_sym && !S->xf_valid) CHK(upload_image_xf(S));
    if (!doUpdate && S->has_sym && std::memcmp(S->M, S->Mxf, sizeof(S->M)) != 0) CHK(upload_image_xf(S));
    if (!doUpdate) {
        const double buf = 0.5 * (S->opt.listCutoff - S->opt.outerCutoff);
        CHK(S->upflag.ensure(1));
        if (!S->dflag.p) {
            CHK(S->dflag.ensure(1));
            CHK(S->prepcount.ensure(1));
            CHK(S->prepdone.ensure(1));
            launch_fill_int(S->dflag.p, 1, 0, st);
            launch_fill_int((int *)S->prepcount.p, 2, 0, st);
        }
        CHK(ensure_call_buffers(S));
        CHK(ensure_desc(S, d_xyz, S->last_grad));
        S->epoch = next_epoch(S->epoch);
        const bool plain = S->opt.precision == NBABFS_PLAIN_FP64;
        PrepareCheck chk{d_xyz, S->xyz_ref.p, S->has_fixed ? S->freef.p : nullptr, buf * buf, S->upflag.d,
                         S->dflag.p, S->epoch, S->prepcount.p, S->prepdone.d};
        errw_before_prepare(S);
        launch_prepare(S->d_desc.p, 1, S->n_pad, plain ? 0 : 1, S->heads.p, chk, st);
        HIPCHK(hipGetLastError());
        // The descriptor's coordinates must be these for the speculative pass.
        if (S->last_xyz == d_xyz && S->n > 0 && !cellCheck) {
            CHK(launch_force(S, ForceGuard{S->dflag.p, S->epoch}));
            S->forced_xyz = d_xyz;
        }
        // Wait for the test through the mapped completion marker (no event in
        // the stream between the two kernels); the stream is checked as well,
        // so a failed launch cannot leave the host waiting.
        for (long spin = 0; *(volatile int *)S->prepdone.h != S->epoch; spin++) {
            if ((spin & 1023) == 1023) {
                const hipError_t q = hipStreamQuery(st);
                if (q != hipSuccess && q != hipErrorNotReady) return fail(NBABFS_STATUS_LOGIC_ERROR, "%s", hipGetErrorString(q));
                if (q == hipSuccess && *(volatile int *)S->prepdone.h != S->epoch)
                    return fail(NBABFS_STATUS_LOGIC_ERROR, "update test finished without its completion marker");
            }
        }
        doUpdate = *(volatile int *)S->upflag.h == S->epoch;
        if (doUpdate && S->forced_xyz) {
            // The guarded launch returns at once; keep it out of the timing.
            if (S->timing && !S->ev_group.empty()) {
                S->ev_used -= S->ev_group.back();
                S->ev_group.pop_back();
            }
            S->forced_xyz = nullptr;
        }
        if (!doUpdate && cellCheck) {
            CHK(S->u64.ensure(2));
            launch_fill_int((int *)S->u64.p, 2, 0, st);
            if (S->n > 0)
                hipLaunchKernelGGL(k_max_displacement, dim3(nblk(S->n, 256)), dim3(256), 0, st, d_xyz, S->xyz_ref.p,
                                   S->has_fixed ? S->freef.p : nullptr, S->n, S->u64.p);
            unsigned long long bits = 0;
            HIPCHK(hipMemcpyAsync(&bits, S->u64.p, sizeof(bits), hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
            double r2;
            std::memcpy(&r2, &bits, sizeof(r2));
            doImages = check_for_image_update(S, std::sqrt(r2));
        }
    }
    if (doUpdate) {
        CHK(build_list(S, d_xyz));
        CHK(S->xyz_ref.ensure(3 * (size_t)std::max(S->n, 1)));
        if (S->n) HIPCHK(hipMemcpyAsync(S->xyz_ref.p, d_xyz, 3 * (size_t)S->n * sizeof(double), hipMemcpyDeviceToDevice, st));
        S->have_ref = true;
    }
    if (S->has_sym && (doUpdate || doImages)) CHK(build_images(S, d_xyz));
    if (doUpdate || doImages) {
        count_update(S);
        S->sumImages += (double)S->nimages;
        S->sumImagePairs += (double)S->img_npairs;
    }
    S->isNew = false;
    if (rebuilt) *rebuilt = (doUpdate || doImages) ? 1 : 0;
    return NBABFS_STATUS_CONTINUE;
}

static void fill_desc(nbabfs_state *S, SysDesc &D, const double *d_xyz, double *d_grad)
{
    std::memset(&D, 0, sizeof(D));
    D.xyz = d_xyz;
    D.grad = d_grad;
    D.slot_atom = S->slot_atom.p;
    D.atom_slot = S->atom_slot.p;
    D.centre = S->centre.p;
    D.xyzq = S->xyzq.p;
    D.styp = S->styp.p;
    D.atype = S->atype.p;
    D.charge = S->charge.p;
    D.ljA = S->ljA.p;
    D.ljB = S->ljB.p;
    D.ljrow = S->ljrow.p;
    D.ljA14 = S->ljA14.p;
    D.ljB14 = S->ljB14.p;
    D.entries = S->entries.p;
    D.grad_assign = S->grad_assign;
    D.p14 = S->p14.p;
    D.facc = S->facc.p;
    D.gplain = S->gplain.p;
    D.eacc = S->eacc.p;
    D.vacc = S->vacc.p;
    D.epart14 = S->epart14.p;
    D.eplain = S->eplain.p;
    D.energies = S->energies.p;
    D.err = S->errw.p;
    D.rowptr = S->csrValid ? S->rowptr.p : nullptr;
    D.cols = S->csrValid ? S->cols.p : nullptr;
    for (int d = 0; d < 3; d++) D.origin[d] = S->grid.corg[d];
    D.escale = S->escale;
    D.escale14 = S->escale14;
    D.qscale = std::sqrt(S->escale);
    D.n = S->n;
    D.n_pad = S->n_pad;
    D.ncl = S->ncl;
    D.ntypes = S->ntypes;
    D.n14 = S->n14;
    D.n_plain_blocks = nblk(S->n, 256);
    D.n14_blocks = nblk(S->n14, NB_P14_BLOCK);
    D.nimg = (int)S->slots.size();
    D.imgxf = D.nimg ? S->imgxf.p : nullptr;
    D.gacc = D.nimg ? S->gacc.p : nullptr;
    D.n_qlj = S->has_qc ? S->nlj * NB_QCMM_CHUNKS : 0;   // (row, chunk) partials
    D.qlj_atom = S->lj_atom.p;
    D.qlj_ptr = S->lj_ptr.p;
    D.qlj_idx = S->lj_idx.p;
    D.qlj14_ptr = S->lj14_ptr.p;
    D.qlj14_idx = S->lj14_idx.p;
    D.qljpart = D.n_qlj > 0 ? S->qljpart.p : nullptr;
    D.n_img_items = S->nitems - S->nitems_main;
    D.eimgpart = (S->opt.precision == NBABFS_PLAIN_FP64 && D.n_img_items > 0) ? S->eimgpart.p : nullptr;
}

// Entries per work item of the image lists (they run in a launch of their own,
// one workgroup per item, so smaller items balance the launch).
static int image_per_item(const nbabfs_options &o)
{
    return std::max(1, std::min(o.imageEntriesPerItem > 0 ? o.imageEntriesPerItem : 16, NBABFS_MAX_ENTRIES_PER_ITEM));
}

// wavesPerCU > 0: persistent grid of that many waves per CU; otherwise one
// workgroup per work item.
static int n_persistent_blocks(const nbabfs_options &o, int nitems)
{
    if (o.wavesPerCU <= 0) return nitems;
    const int cu = g_cu_count > 0 ? g_cu_count : 256;
    return std::max(1, std::min(nitems, cu * o.wavesPerCU));
}

// NBModelABFS_MMMMEnergy (NBModelABFS.c:228-301) enqueued on the state's stream.
// The pair kernel of the state's precision (timed when timing is on).
static int launch_force(nbabfs_state *S, ForceGuard g)
{
    hipStream_t st = S->stream;
    // The kernels write through the descriptor's pointers: it must describe the current buffers.
    if (!desc_current(S)) return fail(NBABFS_STATUS_LOGIC_ERROR, "pair kernel launched with an out-of-date descriptor");
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    if (S->opt.precision != NBABFS_PLAIN_FP64) {
        const int nm = S->nitems_main, ni = S->nitems - S->nitems_main;
        // The image launch runs on the second stream, overlapping the main one.
        hipStream_t sti = st;
        if (ni > 0 && nm > 0) {
            if (!S->stream2) {
                HIPCHK(hipStreamCreateWithFlags(&S->stream2, hipStreamNonBlocking));
                HIPCHK(hipEventCreateWithFlags(&S->ev_fork, hipEventDisableTiming));
                HIPCHK(hipEventCreateWithFlags(&S->ev_join, hipEventDisableTiming));
            }
            HIPCHK(hipEventRecord(S->ev_fork, st));
            HIPCHK(hipStreamWaitEvent(S->stream2, S->ev_fork, 0));
            sti = S->stream2;
        }
        if (S->timing && nm + ni > 0) CHK(timing_events(S, (ni > 0 && nm > 0) ? 4 : 2, ev));
        if (ni > 0) {
            QueueParams q{nullptr, S->qstarts.p};
            const int o = nm > 0 ? 2 : 0;
            launch_force_tiled(S->d_desc.p, S->items.p + nm, q, S->F, ni, S->ntypes, image_per_item(S->opt), g, sti,
                               true, ev[o], ev[o + 1], 0, false, 0);
        }
        if (nm > 0) {
            QueueParams q{S->opt.wavesPerCU > 0 ? S->heads.p : nullptr, S->qstarts.p};
            launch_force_tiled(S->d_desc.p, S->items.p, q, S->F, n_persistent_blocks(S->opt, nm), S->ntypes,
                               S->opt.entriesPerItem, g, st, false, ev[0], ev[1], nblk(S->n14, NB_P14_BLOCK),
                               S->opt.precision == NBABFS_TILED_FP64, 0);
        } else if (S->n14 > 0) {
            QueueParams q{nullptr, S->qstarts.p};
            launch_force_tiled(S->d_desc.p, S->items.p, q, S->F, 0, S->ntypes, S->opt.entriesPerItem, g, st, false,
                               ni > 0 ? nullptr : ev[0], ni > 0 ? nullptr : ev[1], nblk(S->n14, NB_P14_BLOCK),
                               S->opt.precision == NBABFS_TILED_FP64, 0);
        }
        if (sti != st) {
            HIPCHK(hipEventRecord(S->ev_join, sti));
            HIPCHK(hipStreamWaitEvent(st, S->ev_join, 0));
        }
    } else {
        const int nm = S->nitems_main, ni = S->nitems - S->nitems_main;
        if (S->timing) CHK(timing_events(S, 2, ev));
        launch_force_plain(S->d_desc.p, 1, S->n, S->F, g, st, ev[0], ni > 0 ? nullptr : ev[1]);
        if (ni > 0) launch_force_image_fp64(S->d_desc.p, S->items.p + nm, ni, S->F, g, st, nullptr, ev[1]);
    }
    HIPCHK(hipGetLastError());
    return NBABFS_STATUS_CONTINUE;
}

static int mmmm_async(nbabfs_state *S, const double *d_xyz, double *d_grad)
{
    if (!S->listValid) return fail(NBABFS_STATUS_LOGIC_ERROR, "no pair list: call the update first");
    hipStream_t st = S->stream;
    const bool plain = S->opt.precision == NBABFS_PLAIN_FP64;
    CHK(ensure_call_buffers(S));
    // A gradient array other than the descriptor's changes only the
    // finalize pass, which reads the descriptor after the pair kernel.
    CHK(ensure_desc(S, d_xyz, d_grad));
    if (S->n == 0) {
        launch_fill_int((int *)S->energies.p, 2 * NB_NENERGIES, 0, st);
        HIPCHK(hipGetLastError());
        return NBABFS_STATUS_CONTINUE;
    }
    if (S->forced_xyz != d_xyz) {
        PrepareCheck none{nullptr, nullptr, nullptr, 0.0, nullptr, nullptr, 0, nullptr, nullptr};
        errw_before_prepare(S);
        launch_prepare(S->d_desc.p, 1, S->n_pad, plain ? 0 : 1, S->heads.p, none, st);
        CHK(launch_force(S, ForceGuard{nullptr, 0}));
    }
    S->forced_xyz = nullptr;
    // The tiled pair launch evaluates the 1-4 list in its extra workgroups.
    if (plain) launch_pairs14(S->d_desc.p, 1, nblk(S->n14, NB_P14_BLOCK), S->F, st);
    if (S->has_qc) launch_qclj(S->d_desc.p, S->nlj, S->F, st);
    launch_finalize(S->d_desc.p, 1, S->n_pad, plain ? 1 : 0, st);
    S->err_dirty = false;
    HIPCHK(hipGetLastError());
    return NBABFS_STATUS_CONTINUE;
}

// Cut the compacted entries of the current lists (main, then images) into
// work items of per_item entries (the entries and their order are unchanged)
// and order them longest-first.
static int recut_items(nbabfs_state *S, int per_item)
{
    hipStream_t st = S->stream;
    per_item = std::max(1, per_item > 0 ? per_item : 16);
    const int ipi = image_per_item(S->opt);
    if (S->nsc <= 0) return NBABFS_STATUS_CONTINUE;
    const int nb = (int)S->slots.size() * S->nsc;
    launch_items_per_sc(S->counts.p, S->nsc, per_item, S->nits.p, st);
    launch_scan(S->nits.p, S->ioff.p, S->nsc, S->scantmp.p, st);
    hipLaunchKernelGGL(k_list_totals, dim3(1), dim3(64), 0, st, S->eoff.p, S->counts.p, S->ioff.p, S->nits.p, S->nsc,
                       S->flag.p, S->tail.d);
    if (nb > 0) {
        launch_items_per_sc(S->icounts.p, nb, ipi, S->inits.p, st);
        launch_scan(S->inits.p, S->iioff.p, nb, S->scantmp.p, st);
        hipLaunchKernelGGL(k_list_totals, dim3(1), dim3(64), 0, st, S->ieoff.p, S->icounts.p, S->iioff.p, S->inits.p,
                           nb, S->flag.p, S->tail.d + 4);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(st));
    S->nitems_main = ((volatile int *)S->tail.h)[1];
    const int nimg_items = nb > 0 ? ((volatile int *)S->tail.h)[5] : 0;
    S->nitems = S->nitems_main + nimg_items;
    CHK(S->items.ensure(std::max(S->nitems, 1)));
    launch_cut_items(S->counts.p, S->eoff.p, S->ioff.p, S->nsc, S->nsc, per_item, 0, S->items.p, st);
    if (nb > 0) {
        launch_add_int(S->iioff.p, nb, S->nitems_main, st);
        launch_cut_items(S->icounts.p, S->ieoff.p, S->iioff.p, nb, S->nsc, ipi, 0, S->items.p, st, true);
    }
    HIPCHK(hipGetLastError());
    S->desc_valid = false;   // the descriptor carries the image item count
    return order_items(S);
}

// ------------------------------------------------------------------------------------------------
// Image lists (GenerateImageLists, pMolecule-1.9.0/extensions/csource/NBModelABFS.c:753-1045, the
// MM/MM list of a pure MM system): the host walks the transformations and
// (a, b, c) ranges as the reference does (inverse skipping and scales, box
// search limits, the box-overlap test, the +d/-d stepping of the image
// coordinates); the GPU generates the stepped image coordinates and runs one
// cross search per image slot.  Slots whose image has no pair are dropped from
// the reported image list (ImageList_CreateImage, ImageList.c:31-62).
// ------------------------------------------------------------------------------------------------
// Grow an entry buffer to hold `need` entries, keeping its first `keep`.
static int ensure_entries_keep(nbabfs_state *S, size_t need, size_t keep)
{
    if (S->entries.p && need <= S->entries.cap) return NBABFS_STATUS_CONTINUE;
    CHK(S->entries_tmp.ensure(need));
    if (keep) HIPCHK(hipMemcpyAsync(S->entries_tmp.p, S->entries.p, keep * sizeof(Entry), hipMemcpyDeviceToDevice, S->stream));
    HIPCHK(hipStreamSynchronize(S->stream));
    S->entries.swap(S->entries_tmp);
    S->entries_tmp.release();
    return NBABFS_STATUS_CONTINUE;
}

static int build_images(nbabfs_state *S, const double *d_xyz)
{
    hipStream_t st = S->stream;
    const int n = S->n;
    const nbabfs_options &o = S->opt;
    const double cutoff = o.listCutoff;
    S->slots.clear();
    S->nimages = 0;
    S->img_npairs = 0;
    S->nent = S->nent_main;
    if (S->has_sym && n > 0 && S->nsc > 0 && cutoff > 0.0) {
        // Box of the current coordinates (Coordinates3_EnclosingOrthorhombicBox; upper = extent + lower),
        // widened by the cutoff.
        const int nparts = std::min(nblk(n, 256), 1024);
        CHK(S->bbpart.ensure(6 * (size_t)nparts));
        CHK(S->bbox.ensure(6));
        CHK(S->h_bbox.ensure(6));
        launch_bbox(d_xyz, n, S->bbpart.p, nparts, S->bbox.p, st);
        HIPCHK(hipMemcpyAsync(S->h_bbox.p, S->bbox.p, 6 * sizeof(double), hipMemcpyDeviceToHost, st));
        // Transformations searched: with inverses, t is skipped when its inverse comes later.
        std::vector<int> tl;
        std::vector<double> rt;
        for (int t = 0; t < S->ntrans; t++) {
            const int ti = S->checkInv ? S->tinv[t] : -1;
            if (ti >= 0 && t < ti) continue;
            double Ro[9], To[3];
            orthogonalize(S->M, S->Minv, &S->trot[9 * t], &S->ttra[3 * t], Ro, To);
            tl.push_back(t);
            rt.insert(rt.end(), Ro, Ro + 9);
            rt.insert(rt.end(), To, To + 3);
        }
        const int nt = (int)tl.size();
        CHK(S->rtbuf.upload(rt.data(), rt.size(), st));
        CHK(S->itc.ensure((size_t)nt * 3 * n));
        launch_image_transform(d_xyz, n, S->rtbuf.p, nt, S->itc.p, st);
        CHK(S->ibb.ensure(6 * (size_t)nt));
        CHK(S->h_ibb.ensure(6 * (size_t)nt));
        for (int k = 0; k < nt; k++) launch_bbox(S->itc.p + (size_t)k * 3 * n, n, S->bbpart.p, nparts, S->ibb.p + 6 * k, st);
        HIPCHK(hipMemcpyAsync(S->h_ibb.p, S->ibb.p, 6 * (size_t)nt * sizeof(double), hipMemcpyDeviceToHost, st));
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(st));
        CHK(check_box(S->h_bbox.p));
        for (int k = 0; k < nt; k++) CHK(check_box(S->h_ibb.p + 6 * k));
        double lower[3], upper[3];
        {
#pragma clang fp contract(off)
            for (int d = 0; d < 3; d++) {
                lower[d] = S->h_bbox.p[d];
                upper[d] = (S->h_bbox.p[3 + d] - S->h_bbox.p[d]) + lower[d];
                lower[d] += -cutoff;
                upper[d] += cutoff;
            }
        }
        // Walk the (a, b, c) ranges.
        std::vector<ImageStep> steps;
        std::vector<int> step_off(nt + 1, 0);
        std::vector<float> boxes;
        for (int k = 0; k < nt; k++) {
#pragma clang fp contract(off)
            const int t = tl[k];
            const int ti = S->checkInv ? S->tinv[t] : -1;
            const double defscale = ti >= 0 ? 1.0 : 0.5;
            double ilower[3], iupper[3];
            for (int d = 0; d < 3; d++) {
                ilower[d] = S->h_ibb.p[6 * k + d];
                iupper[d] = (S->h_ibb.p[6 * k + 3 + d] - S->h_ibb.p[6 * k + d]) + ilower[d];
            }
            int lim[6];
            box_search_limits(S->M, lower, upper, ilower, iupper, lim);
            if (S->expandFactor > 0)
                for (int q = 0; q < 6; q++) lim[q] += (q & 1) ? S->expandFactor : -S->expandFactor;
            for (int a = lim[0]; a <= lim[1]; a++)
                for (int b = lim[2]; b <= lim[3]; b++)
                    for (int c = lim[4]; c <= lim[5]; c++) {
                        if (a == 0 && b == 0 && c == 0 && t == S->identity) continue;
                        double scale = defscale;
                        if (ti >= 0 && ti == t) {
                            int inv[3];
                            inverse_integer_translation(S->trot.data(), S->ttra.data(), t, ti, a, b, c, inv);
                            if (inv[0] >= lim[0] && inv[0] <= lim[1] && inv[1] >= lim[2] && inv[1] <= lim[3] &&
                                inv[2] >= lim[4] && inv[2] <= lim[5]) {
                                if (a == inv[0] && b == inv[1] && c == inv[2]) {
                                    scale = 0.5;
                                } else {
                                    scale = 1.0;
                                    if (a < inv[0] || (a == inv[0] && b < inv[1]) ||
                                        (a == inv[0] && b == inv[1] && c < inv[2]))
                                        continue;
                                }
                            } else {
                                scale = 1.0;
                            }
                        }
                        // SymmetryParameters_Displacement.
                        ImageStep q;
                        const double na = (double)a, nb = (double)b, nc = (double)c;
                        for (int i = 0; i < 3; i++) q.d[i] = na * S->M[3 * i] + nb * S->M[3 * i + 1] + nc * S->M[3 * i + 2];
                        q.out = -1;
                        q.pad = 0;
                        for (int d = 0; d < 3; d++) { ilower[d] += q.d[d]; iupper[d] += q.d[d]; }
                        if (ilower[0] <= upper[0] && ilower[1] <= upper[1] && ilower[2] <= upper[2] &&
                            iupper[0] >= lower[0] && iupper[1] >= lower[1] && iupper[2] >= lower[2]) {
                            q.out = (int)S->slots.size();
                            S->slots.push_back({a, b, c, t, scale, 0});
                            // Image box in the cluster frame, padded outwards for FP32.
                            for (int d = 0; d < 3; d++) boxes.push_back((float)(ilower[d] - S->grid.corg[d] - 1.0e-2));
                            for (int d = 0; d < 3; d++) boxes.push_back((float)(iupper[d] - S->grid.corg[d] + 1.0e-2));
                        }
                        steps.push_back(q);
                        for (int d = 0; d < 3; d++) {
                            const double md = -q.d[d];
                            ilower[d] += md;
                            iupper[d] += md;
                        }
                    }
            step_off[k + 1] = (int)steps.size();
        }
        const int ns = (int)S->slots.size();
        if (ns > 0) {
            const int nsc = S->nsc, ncl = S->ncl, nb = ns * nsc;
            CHK(S->isteps.upload(steps.data(), steps.size(), st));
            CHK(S->istep_off.upload(step_off.data(), step_off.size(), st));
            CHK(S->ixyz.ensure((size_t)ns * 3 * n));
            launch_image_coords(S->itc.p, n, nt, S->istep_off.p, S->isteps.p, S->ixyz.p, st);
            CHK(S->iblo.ensure((size_t)ns * ncl));
            CHK(S->ibhi.ensure((size_t)ns * ncl));
            CHK(S->icen.ensure(ncl));
            for (int k = 0; k < ns; k++)
                launch_cluster_geom(S->ixyz.p + (size_t)k * 3 * n, S->slot_atom.p, ncl, S->grid.corg,
                                    S->iblo.p + (size_t)k * ncl, S->ibhi.p + (size_t)k * ncl, S->icen.p, st);
            CHK(S->ibox.upload(boxes.data(), boxes.size(), st));
            // The primary clusters' boxes for the current coordinates: atoms may
            // have moved since the main list (and its boxes) was built.
            launch_cluster_geom(d_xyz, S->slot_atom.p, ncl, S->grid.corg, S->blo.p, S->bhi.p, S->centre.p, st);
            // Grid method: the i atoms' cells on the grid made from the current coordinates
            // (NBModelABFSState_GridInitialize) and each image's conforming grid.
            SearchParams P;
            std::memset(&P, 0, sizeof(P));
            if (S->useGrid) {
                BuildGrid g2;
                std::memset(&g2, 0, sizeof(g2));
                reference_grid(S->h_bbox.p, o.cutoffCellSizeFactor * o.listCutoff, g2);
                P.pointcell = o.useGridByCell ? 0 : 1;
                CHK(S->icell.ensure(n));
                launch_image_cells(d_xyz, n, g2, P.pointcell, S->u, S->has_flags ? S->freef.p : nullptr, S->icell.p,
                                   st);
                CHK(S->igrid.ensure(ns));
                launch_image_grid(S->ixyz.p, n, ns, g2, S->igrid.p, st);
                P.icell = S->icell.p;
                P.igrid = S->igrid.p;
                P.stencil = S->stencil.p;
                P.u = S->u;
                P.gh = g2.h;
            }
            P.cutoff2 = cutoff * cutoff;
            P.use_grid = S->useGrid ? 1 : 0;
            P.free_flags = S->has_flags ? S->freef.p : nullptr;
            P.reach = (float)(cutoff + 0.01);
            P.ixyz = S->ixyz.p;
            P.iblo = S->iblo.p;
            P.ibhi = S->ibhi.p;
            P.ibox = S->ibox.p;
            P.n = n;
            P.ncl = ncl;
            // Scratch: each (image, super-cluster) block gets its candidate bound.
            CHK(S->ibound.ensure(nb));
            CHK(S->isoff.ensure(nb + 1));
            CHK(S->scantmp.ensure(scan_tmp_size(nb + 1) + 1024));
            launch_image_bounds(P, nsc, ns, S->blo.p, S->bhi.p, S->ibound.p, st);
            launch_scan(S->ibound.p, S->isoff.p, nb, S->scantmp.p, st);
            CHK(S->tail.ensure(8));
            CHK(S->flag.ensure(1));
            launch_fill_int(S->flag.p, 1, 0, st);
            CHK(S->icounts.ensure(nb));
            CHK(S->incand.ensure(nb));
            CHK(S->ieoff.ensure(nb + 1));
            CHK(S->inits.ensure(nb));
            CHK(S->iioff.ensure(nb + 1));
            hipLaunchKernelGGL(k_list_totals, dim3(1), dim3(64), 0, st, S->isoff.p, S->ibound.p, S->isoff.p,
                               S->ibound.p, nb, S->flag.p, S->tail.d);
            HIPCHK(hipGetLastError());
            HIPCHK(hipStreamSynchronize(st));
            const int total_cand = ((volatile int *)S->tail.h)[0];
            CHK(S->iscratch.ensure(std::max(total_cand, 1)));
            P.soff = S->isoff.p;
            P.scap = S->ibound.p;
            launch_search_images(P, nsc, ns, d_xyz, S->slot_atom.p, S->blo.p, S->bhi.p, S->iscratch.p, S->incand.p,
                                 S->icounts.p, S->flag.p, st);
            launch_scan(S->icounts.p, S->ieoff.p, nb, S->scantmp.p, st);
            const int per_item = image_per_item(o);
            launch_items_per_sc(S->icounts.p, nb, per_item, S->inits.p, st);
            launch_scan(S->inits.p, S->iioff.p, nb, S->scantmp.p, st);
            hipLaunchKernelGGL(k_list_totals, dim3(1), dim3(64), 0, st, S->ieoff.p, S->icounts.p, S->iioff.p,
                               S->inits.p, nb, S->flag.p, S->tail.d);
            HIPCHK(hipGetLastError());
            HIPCHK(hipStreamSynchronize(st));
            if (((volatile int *)S->tail.h)[2] > 0)
                return fail(NBABFS_STATUS_LOGIC_ERROR, "image search exceeded its candidate bound");
            const int nent_img = ((volatile int *)S->tail.h)[0];
            const int nit_img = ((volatile int *)S->tail.h)[1];
            launch_add_int(S->ieoff.p, nb, S->nent_main, st);
            launch_add_int(S->iioff.p, nb, S->nitems_main, st);
            CHK(ensure_entries_keep(S, (size_t)S->nent_main + nent_img, (size_t)S->nent_main));
            CHK(S->items.ensure((size_t)S->nitems_main + nit_img + 1));
            CHK(S->inpairs.ensure(nb));
            CHK(S->h_inpairs.ensure(nb));
            launch_compact(S->iscratch.p, 0, S->isoff.p, S->incand.p, S->icounts.p, S->ieoff.p, S->iioff.p, nb, nsc,
                           per_item, 0, S->entries.p, S->items.p, nullptr, S->inpairs.p, st);
            HIPCHK(hipMemcpyAsync(S->h_inpairs.p, S->inpairs.p, nb * sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
            HIPCHK(hipGetLastError());
            HIPCHK(hipStreamSynchronize(st));
            for (int k = 0; k < ns; k++) {
                long long np = 0;
                for (int q = 0; q < nsc; q++) np += (long long)S->h_inpairs.p[(size_t)k * nsc + q];
                S->slots[k].npairs = np;
                if (np > 0) S->nimages++;
                S->img_npairs += np;
            }
            S->nent = S->nent_main + nent_img;
        }
    }
    std::memcpy(S->Mref, S->M, sizeof(S->Mref));
    S->xf_valid = false;
    CHK(recut_items(S, std::max(1, o.entriesPerItem > 0 ? o.entriesPerItem : 16)));
    CHK(upload_image_xf(S));
    return NBABFS_STATUS_CONTINUE;
}

// Energy-pass transformations of the image slots for the current M
// (MMMMImageEnergy: the fractional transformation plus (a, b, c),
// orthogonalized), expressed in the cluster frame, and the dE/dM accumulators.
static int upload_image_xf(nbabfs_state *S)
{
    const int ns = (int)S->slots.size();
    if (ns == 0) {
        S->xf_valid = true;
        std::memcpy(S->Mxf, S->M, sizeof(S->Mxf));
        return NBABFS_STATUS_CONTINUE;
    }
    CHK(S->h_imgxf.ensure(ns));
    CHK(S->imgxf.ensure(ns));
    CHK(S->gacc.ensure((size_t)NB_GACC_WORDS_PER_IMAGE * ns));
    HIPCHK(hipStreamSynchronize(S->stream));   // the staging copy may still be read by an earlier upload
    const double *og = S->grid.corg;
    for (int k = 0; k < ns; k++) {
        const auto &sl = S->slots[k];
        const double xtr[3] = {S->ttra[3 * sl.t] + (double)sl.a, S->ttra[3 * sl.t + 1] + (double)sl.b,
                               S->ttra[3 * sl.t + 2] + (double)sl.c};
        double Ro[9], To[3];
        orthogonalize(S->M, S->Minv, &S->trot[9 * sl.t], xtr, Ro, To);
        ImageXf &X = S->h_imgxf.p[k];
        std::memset(&X, 0, sizeof(X));
        for (int q = 0; q < 9; q++) X.R[q] = Ro[q];
        for (int r = 0; r < 3; r++) X.T[r] = (Ro[3 * r] * og[0] + Ro[3 * r + 1] * og[1] + Ro[3 * r + 2] * og[2]) + To[r] - og[r];
        X.scale = sl.scale;
        static const double I3[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
        X.translation = std::memcmp(&S->trot[9 * sl.t], I3, sizeof(I3)) == 0 ? 1.0 : 0.0;
        for (int r = 0; r < 3; r++) X.Tabs[r] = To[r];
        {
#pragma clang fp contract(off)
            // PairwiseInteractionABFS_MMMMEnergy's eScale for an image: (1/dielectric * scale) * factor.
            X.escale = ((1.0e+00 / S->opt.dielectric) * sl.scale) * coulomb_factor();
        }
    }
    HIPCHK(hipMemcpyAsync(S->imgxf.p, S->h_imgxf.p, ns * sizeof(ImageXf), hipMemcpyHostToDevice, S->stream));
    std::memcpy(S->Mxf, S->M, sizeof(S->Mxf));
    S->xf_valid = true;
    return NBABFS_STATUS_CONTINUE;
}

// CheckForImageUpdate (NBModelABFS.c:635-684): some image's lattice
// translation (transformation translation plus (a, b, c), through M - M_ref)
// moved by more than listCutoff - outerCutoff - the largest atom displacement.
static bool check_for_image_update(const nbabfs_state *S, double maxdisp)
{
#pragma clang fp contract(off)
    if (S->nimages <= 0) return false;
    const double buffac = S->opt.listCutoff - S->opt.outerCutoff - maxdisp;
    double dM[9];
    for (int k = 0; k < 9; k++) dM[k] = S->M[k] + -1.0e+00 * S->Mref[k];
    for (const auto &sl : S->slots) {
        if (sl.npairs <= 0) continue;
        double v[3] = {S->ttra[3 * sl.t] + (double)sl.a, S->ttra[3 * sl.t + 1] + (double)sl.b,
                       S->ttra[3 * sl.t + 2] + (double)sl.c};
        m33_apply(dM, v);
        const double di = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        if (di > buffac) return true;
    }
    return false;
}


// Time the pair kernel on the current list for each (entriesPerItem,
// wavesPerCU) candidate and keep the fastest.
static int autotune(nbabfs_state *S, const double *d_xyz, int *best_epi, int *best_wpc, double *best_ms)
{
    static const int cand[][2] = {{8, 0}, {12, 0}, {16, 0}, {24, 0}, {32, 0}, {48, 0}, {8, 12}, {16, 12}, {32, 12}};
    const int reps = 5;
    hipStream_t st = S->stream;
    CHK(update(S, d_xyz, nullptr));
    S->forced_xyz = nullptr;
    if (S->opt.precision == NBABFS_PLAIN_FP64 || S->n == 0) {
        if (best_epi) *best_epi = S->opt.entriesPerItem;
        if (best_wpc) *best_wpc = S->opt.wavesPerCU;
        if (best_ms) *best_ms = 0.0;
        return NBABFS_STATUS_CONTINUE;
    }
    CHK(ensure_call_buffers(S));
    CHK(ensure_desc(S, d_xyz, S->last_grad));
    PrepareCheck none{nullptr, nullptr, nullptr, 0.0, nullptr, nullptr, 0, nullptr, nullptr};
    errw_before_prepare(S);
    launch_prepare(S->d_desc.p, 1, S->n_pad, 1, S->heads.p, none, st);
    hipEvent_t ev[2 * reps];
    for (int k = 0; k < 2 * reps; k++) HIPCHK(hipEventCreate(&ev[k]));
    const nbabfs_options keep = S->opt;
    const bool was_timing = S->timing;
    S->timing = false;
    double best = 1e300;
    int bi = -1;
    int rc = NBABFS_STATUS_CONTINUE;
    const int nc = (int)(sizeof(cand) / sizeof(cand[0]));
    std::vector<double> tbest(nc, 1e300);
    // Two sweeps, the second in reverse order, best of the two per candidate: a drift of the device clocks over
    // the sweep (they ramp up under load) then favours no end of the candidate list.
    for (int sweep = 0; sweep < 2 * nc && rc == NBABFS_STATUS_CONTINUE; sweep++) {
        const int c = sweep < nc ? sweep : 2 * nc - 1 - sweep;
        S->opt.entriesPerItem = cand[c][0];
        S->opt.wavesPerCU = cand[c][1];
        rc = recut_items(S, cand[c][0]);
        // The launches below read the descriptor: refresh it after the new cut.
        if (rc == NBABFS_STATUS_CONTINUE) rc = ensure_desc(S, d_xyz, S->last_grad);
        // About 20 ms of untimed launches first (the re-cut left the GPU idle for a while and its clocks ramp up
        // again under load), then each launch timed alone; persistent queues are reset before every launch.
        const auto w0 = std::chrono::steady_clock::now();
        for (int w = 0; w < 4000 && rc == NBABFS_STATUS_CONTINUE; w++) {
            if (S->opt.wavesPerCU > 0) launch_fill_int(S->heads.p, 8 * 32, 0, st);
            rc = launch_force(S, ForceGuard{nullptr, 0});
            if ((w & 15) == 15) {
                if (hipStreamSynchronize(st) != hipSuccess) rc = fail(NBABFS_STATUS_LOGIC_ERROR, "synchronisation failed");
                if (std::chrono::steady_clock::now() - w0 > std::chrono::milliseconds(20)) break;
            }
        }
        for (int r = 0; r < reps && rc == NBABFS_STATUS_CONTINUE; r++) {
            if (S->opt.wavesPerCU > 0) launch_fill_int(S->heads.p, 8 * 32, 0, st);
            if (hipEventRecord(ev[2 * r], st) != hipSuccess) rc = fail(NBABFS_STATUS_LOGIC_ERROR, "event record failed");
            if (rc == NBABFS_STATUS_CONTINUE) rc = launch_force(S, ForceGuard{nullptr, 0});
            if (rc == NBABFS_STATUS_CONTINUE && hipEventRecord(ev[2 * r + 1], st) != hipSuccess)
                rc = fail(NBABFS_STATUS_LOGIC_ERROR, "event record failed");
        }
        if (rc != NBABFS_STATUS_CONTINUE) break;
        if (hipEventSynchronize(ev[2 * reps - 1]) != hipSuccess) {
            rc = fail(NBABFS_STATUS_LOGIC_ERROR, "event synchronisation failed");
            break;
        }
        double tot = 0.0;
        for (int r = 0; r < reps; r++) {
            float ms = 0.0f;
            if (hipEventElapsedTime(&ms, ev[2 * r], ev[2 * r + 1]) != hipSuccess) rc = fail(NBABFS_STATUS_LOGIC_ERROR, "event timing failed");
            tot += ms;
        }
        if (rc == NBABFS_STATUS_CONTINUE) tbest[c] = std::min(tbest[c], tot / reps);
    }
    for (int c = 0; c < nc; c++)
        if (tbest[c] < best) {
            best = tbest[c];
            bi = c;
        }
    for (int k = 0; k < 2 * reps; k++) (void)hipEventDestroy(ev[k]);
    S->timing = was_timing;
    S->opt = keep;
    if (rc == NBABFS_STATUS_CONTINUE && bi >= 0) {
        S->opt.entriesPerItem = cand[bi][0];
        S->opt.wavesPerCU = cand[bi][1];
    }
    // The accumulators hold the timing launches: the next energy call prepares again.
    CHK(recut_items(S, S->opt.entriesPerItem));
    if (rc != NBABFS_STATUS_CONTINUE) return rc;
    if (best_epi) *best_epi = S->opt.entriesPerItem;
    if (best_wpc) *best_wpc = S->opt.wavesPerCU;
    if (best_ms) *best_ms = best;
    return NBABFS_STATUS_CONTINUE;
}

// ------------------------------------------------------------------------------------------------
// C ABI.
// ------------------------------------------------------------------------------------------------
// The MM/MM 1-4 list of the state: SelfPairList_FromSelfPairList (pCore-1.9.0/extensions/csource/
// PairList.c:342-453) of the given 1-4 pairs with the MM selection (no pure QC atom) as the AND
// selection and the free atoms as the OR selection (NBModelABFS.c:1113-1128).
static int select_14(nbabfs_state *S)
{
    const int n = S->n;
    std::vector<long long> keep;
    keep.reserve(S->p14_all.size());
    for (long long key : S->p14_all) {
        const int i = (int)(key / n), j = (int)(key % n);
        if (!S->qc_pure.empty() && (S->qc_pure[i] || S->qc_pure[j])) continue;
        if (!S->fixed_h.empty() && S->fixed_h[i] && S->fixed_h[j]) continue;
        keep.push_back(key);
    }
    S->p14_keys = keep;
    S->n14 = (int)keep.size();
    std::vector<int> p14(2 * std::max<size_t>(keep.size(), 1), 0);
    for (size_t k = 0; k < keep.size(); k++) {
        p14[2 * k] = (int)(keep[k] / n);
        p14[2 * k + 1] = (int)(keep[k] % n);
    }
    return S->p14.upload(p14.data(), p14.size(), S->stream);
}

extern "C" int nbabfs_state_create(const nbabfs_options *opt, int natoms, const double *charges, const int *ljtypes,
                                   int ntypes, const double *tableA, const double *tableB, const double *tableA14,
                                   const double *tableB14, int nexclusions, const int *exclusions,
                                   int ninteractions14, const int *interactions14, const unsigned char *fixedAtoms,
                                   nbabfs_state **out)
{
    if (!out) return fail(NBABFS_STATUS_INVALID_ARGUMENT, "out is NULL");
    *out = nullptr;
    CHK(nbabfs_init());
    nbabfs_options o;
    nbabfs_default_options(&o);
    if (opt) o = *opt;
    if (o.entriesPerItem <= 0) o.entriesPerItem = 16;
    if (o.entriesPerItem > NBABFS_MAX_ENTRIES_PER_ITEM)
        return fail(NBABFS_STATUS_VALUE_ERROR, "entriesPerItem %d exceeds %d", o.entriesPerItem,
                    NBABFS_MAX_ENTRIES_PER_ITEM);
    if (o.precision != NBABFS_MIXED && o.precision != NBABFS_PLAIN_FP64 && o.precision != NBABFS_TILED_FP64)
        return fail(NBABFS_STATUS_VALUE_ERROR, "unknown precision %d", o.precision);
    if (o.wavesPerCU < 0) o.wavesPerCU = 0;
    if (o.imageEntriesPerItem <= 0) o.imageEntriesPerItem = 16;
    if (o.imageEntriesPerItem > NBABFS_MAX_ENTRIES_PER_ITEM)
        return fail(NBABFS_STATUS_VALUE_ERROR, "imageEntriesPerItem %d exceeds %d", o.imageEntriesPerItem,
                    NBABFS_MAX_ENTRIES_PER_ITEM);
    // NBModelABFS.SetOptions cutoff checks (pMolecule.NBModelABFS.pyx:169-176).
    if (o.dampingCutoff < 0.0 || o.innerCutoff < o.dampingCutoff || o.outerCutoff < o.innerCutoff ||
        o.listCutoff < o.outerCutoff)
        return fail(NBABFS_STATUS_VALUE_ERROR, "invalid cutoff values: damping %.3f inner %.3f outer %.3f list %.3f",
                    o.dampingCutoff, o.innerCutoff, o.outerCutoff, o.listCutoff);
    if (o.dampingCutoff <= 0.0)
        return fail(NBABFS_STATUS_VALUE_ERROR, "dampingCutoff must be positive for the analytic MM/MM form");
    if (o.cutoffCellSizeFactor <= 0.0 || o.dielectric == 0.0)
        return fail(NBABFS_STATUS_VALUE_ERROR, "invalid cell size factor or dielectric");
    if (natoms < 0 || ntypes <= 0 || (natoms > 0 && (!charges || !ljtypes)) || !tableA || !tableB)
        return fail(NBABFS_STATUS_INVALID_ARGUMENT, "missing atoms or LJ parameters");
    for (int i = 0; i < natoms; i++)
        if (ljtypes[i] < 0 || ljtypes[i] >= ntypes)
            return fail(NBABFS_STATUS_INVALID_ARGUMENT, "atom %d has LJ type %d outside [0,%d)", i, ljtypes[i], ntypes);
    if ((nexclusions > 0 && !exclusions) || (ninteractions14 > 0 && !interactions14) || nexclusions < 0 ||
        ninteractions14 < 0)
        return fail(NBABFS_STATUS_INVALID_ARGUMENT, "bad exclusion or 1-4 arrays");
    // Non-finite charges or LJ parameters would reach the energies unnoticed (the pair
    // kernels report non-finite coordinates and out-of-range sums, not parameters).
    for (int i = 0; i < natoms; i++)
        if (!std::isfinite(charges[i])) return fail(NBABFS_STATUS_VALUE_ERROR, "charge of atom %d is not finite", i);
    for (size_t k = 0; k < (size_t)ntypes * (ntypes + 1) / 2; k++)
        if (!std::isfinite(tableA[k]) || !std::isfinite(tableB[k]) || (tableA14 && !std::isfinite(tableA14[k])) ||
            (tableB14 && !std::isfinite(tableB14[k])))
            return fail(NBABFS_STATUS_VALUE_ERROR, "LJ table entry %zu is not finite", k);

    nbabfs_state *S = new nbabfs_state();
    S->opt = o;
    S->n = natoms;
    S->ntypes = ntypes;
    make_factors(o.dampingCutoff, o.innerCutoff, o.outerCutoff, S->F);
    S->escale = coulomb_factor() / o.dielectric;
    S->escale14 = S->escale * o.electrostaticScale14;
    if (hipStreamCreateWithFlags(&S->stream, hipStreamNonBlocking) != hipSuccess) {
        delete S;
        return fail(NBABFS_STATUS_LOGIC_ERROR, "hipStreamCreate failed");
    }
    S->own_stream = true;
    hipStream_t st = S->stream;
    int rc = NBABFS_STATUS_CONTINUE;
    auto bail = [&](int code) {
        delete S;
        return code;
    };

    // LJ tables expanded to ntypes x ntypes (tableindex of LJParameterContainer_Allocate,
    // pMolecule-1.9.0/extensions/csource/LJParameterContainer.c:39-60).
    const size_t nt2 = (size_t)ntypes * ntypes;
    std::vector<double> A(nt2), B(nt2), A14(nt2), B14(nt2);
    for (int i = 0; i < ntypes; i++)
        for (int j = 0; j < ntypes; j++) {
            const int hi = std::max(i, j), lo = std::min(i, j);
            const size_t t = (size_t)hi * (hi + 1) / 2 + lo;
            A[(size_t)i * ntypes + j] = tableA[t];
            B[(size_t)i * ntypes + j] = tableB[t];
            A14[(size_t)i * ntypes + j] = tableA14 ? tableA14[t] : tableA[t];
            B14[(size_t)i * ntypes + j] = tableB14 ? tableB14[t] : tableB[t];
        }
    // Polynomial rows of the tiled path: E = C12 s^12 + C6 s^6 + C3 s^3 + C0
    // for the plain region (r <= innerCutoff) and the force-switched region
    // (PairwiseInteraction.h:72-119 with the squares expanded).
    std::vector<double> R(8 * nt2);
    {
        const AbfsFactors &F = S->F;
        for (size_t k = 0; k < nt2; k++) {
            const double a = A[k], b = B[k], ak = a * F.aK12, bk = b * F.bK6;
            double *p = &R[8 * k];
            p[0] = a;
            p[1] = -b;
            p[2] = 0.0;
            p[3] = b * F.bShift6 - a * F.aShift12;
            p[4] = ak;
            p[5] = -(2.0 * ak * F.aF6 + bk);
            p[6] = 2.0 * bk * F.bF3;
            p[7] = ak * F.aF6 * F.aF6 - bk * F.bF3 * F.bF3;
        }
    }
    // Exclusions: symmetric CSR, ascending rows, no duplicates or self pairs
    // (SelfPairList_MakeConnections, pCore-1.9.0/extensions/csource/PairList.c:458-526).
    std::vector<long long> ek;
    ek.reserve((size_t)nexclusions);
    for (int k = 0; k < nexclusions; k++) {
        int i = exclusions[2 * k], j = exclusions[2 * k + 1];
        if (i < 0 || j < 0 || i >= natoms || j >= natoms) return bail(fail(NBABFS_STATUS_INVALID_ARGUMENT, "exclusion %d out of range", k));
        if (i == j) continue;
        ek.push_back((long long)std::max(i, j) * natoms + std::min(i, j));
    }
    std::sort(ek.begin(), ek.end());
    ek.erase(std::unique(ek.begin(), ek.end()), ek.end());
    std::vector<int> xptr((size_t)natoms + 1, 0), xidx(2 * ek.size());
    for (long long key : ek) {
        xptr[key / natoms + 1]++;
        xptr[key % natoms + 1]++;
    }
    for (int i = 0; i < natoms; i++) xptr[i + 1] += xptr[i];
    {
        std::vector<int> fillc(xptr.begin(), xptr.end() - 1);
        std::vector<std::pair<int, int>> rows;
        rows.reserve(2 * ek.size());
        for (long long key : ek) {
            rows.push_back({(int)(key / natoms), (int)(key % natoms)});
            rows.push_back({(int)(key % natoms), (int)(key / natoms)});
        }
        std::sort(rows.begin(), rows.end());
        for (auto &r : rows) xidx[fillc[r.first]++] = r.second;
    }
    // 1-4 list: SelfPairList_FromSelfPairList with no selections (PairList.c:342-453).
    std::vector<long long> pk;
    for (int k = 0; k < ninteractions14; k++) {
        int i = interactions14[2 * k], j = interactions14[2 * k + 1];
        if (i < 0 || j < 0 || i >= natoms || j >= natoms) return bail(fail(NBABFS_STATUS_INVALID_ARGUMENT, "1-4 pair %d out of range", k));
        if (i == j) continue;
        pk.push_back((long long)std::max(i, j) * natoms + std::min(i, j));
    }
    std::sort(pk.begin(), pk.end());
    pk.erase(std::unique(pk.begin(), pk.end()), pk.end());
    S->p14_all = pk;
    S->xptr_h = xptr;
    S->xidx_h = xidx;
    if (fixedAtoms) {
        std::vector<unsigned char> fr(natoms);
        bool any = false;
        for (int i = 0; i < natoms; i++) {
            fr[i] = fixedAtoms[i] ? 0 : 1;
            any = any || fixedAtoms[i];
        }
        if (any) {
            if ((rc = S->freef.upload(fr.data(), fr.size(), st)) != NBABFS_STATUS_CONTINUE) return bail(rc);
            S->has_fixed = S->has_flags = true;
            S->fixed_h.assign(fixedAtoms, fixedAtoms + natoms);
        }
    }
    // 1-4 list: SelfPairList_FromSelfPairList with the free atoms as the OR selection.
    if ((rc = select_14(S)) != NBABFS_STATUS_CONTINUE) return bail(rc);
    std::vector<double> q(charges, charges + natoms);
    S->charges_h = q;
    std::vector<int> t(ljtypes, ljtypes + natoms);
    if ((rc = S->charge.upload(q.data(), q.size(), st)) != NBABFS_STATUS_CONTINUE ||
        (rc = S->atype.upload(t.data(), t.size(), st)) != NBABFS_STATUS_CONTINUE ||
        (rc = S->ljA.upload(A.data(), nt2, st)) != NBABFS_STATUS_CONTINUE ||
        (rc = S->ljB.upload(B.data(), nt2, st)) != NBABFS_STATUS_CONTINUE ||
        (rc = S->ljrow.upload(R.data(), R.size(), st)) != NBABFS_STATUS_CONTINUE ||
        (rc = S->ljA14.upload(A14.data(), nt2, st)) != NBABFS_STATUS_CONTINUE ||
        (rc = S->ljB14.upload(B14.data(), nt2, st)) != NBABFS_STATUS_CONTINUE ||
        (rc = S->xptr.upload(xptr.data(), xptr.size(), st)) != NBABFS_STATUS_CONTINUE ||
        (rc = S->xidx.upload(xidx.data(), xidx.size(), st)) != NBABFS_STATUS_CONTINUE ||
        (rc = S->errw.ensure(1)) != NBABFS_STATUS_CONTINUE)
        return bail(rc);
    launch_fill_int(S->errw.p, 1, 0, st);
    if (hipStreamSynchronize(st) != hipSuccess) return bail(fail(NBABFS_STATUS_LOGIC_ERROR, "upload failed"));
    *out = S;
    return NBABFS_STATUS_CONTINUE;
}

extern "C" void nbabfs_state_destroy(nbabfs_state *S)
{
    if (!S) return;
    if (S->stream) (void)hipStreamSynchronize(S->stream);
    delete S;
}

extern "C" int nbabfs_set_stream(nbabfs_state *S, void *hip_stream)
{
    if (!S) return fail(NBABFS_STATUS_INVALID_ARGUMENT, "state is NULL");
    HIPCHK(hipStreamSynchronize(S->stream));
    if (S->stream2) HIPCHK(hipStreamSynchronize(S->stream2));
    if (S->own_stream) (void)hipStreamDestroy(S->stream);
    S->own_stream = false;
    if (hip_stream) {
        S->stream = (hipStream_t)hip_stream;
    } else {
        HIPCHK(hipStreamCreateWithFlags(&S->stream, hipStreamNonBlocking));
        S->own_stream = true;
    }
    return NBABFS_STATUS_CONTINUE;
}

extern "C" int nbabfs_update(nbabfs_state *S, const double *d_xyz, int *rebuilt)
{
    if (!S || (S->n > 0 && !d_xyz)) return fail(NBABFS_STATUS_INVALID_ARGUMENT, "null state or coordinates");
    S->calls++;
    return update(S, d_xyz, rebuilt);
}

extern "C" int nbabfs_mmmm_energy_async(nbabfs_state *S, const double *d_xyz, double *d_grad)
{
    if (!S || (S->n > 0 && !d_xyz)) return fail(NBABFS_STATUS_INVALID_ARGUMENT, "null state or coordinates");
    return mmmm_async(S, d_xyz, d_grad);
}

extern "C" int nbabfs_energies(nbabfs_state *S, double *energies, int *present)
{
    if (!S || !energies) return fail(NBABFS_STATUS_INVALID_ARGUMENT, "null argument");
    CHK(S->h_energies.ensure(NB_NENERGIES));
    CHK(S->energies.ensure(NB_NENERGIES));
    HIPCHK(hipMemcpyAsync(S->h_energies.p, S->energies.p, NB_NENERGIES * sizeof(double), hipMemcpyDeviceToHost,
                          S->stream));
    HIPCHK(hipStreamSynchronize(S->stream));
    std::memcpy(energies, S->h_energies.p, 4 * sizeof(double));
    CHK(resolve_npairs(S));
    if (present) {
        present[0] = S->npairs > 0;
        present[1] = S->n14 > 0;
    }
    return call_status((int)S->h_energies.p[NB_ESTATUS], "energy call");
}

extern "C" int nbabfs_energy_host(nbabfs_state *S, const double *xyz, double *grad_inout, double *energies,
                                  int *present, int *rebuilt)
{
    if (!S || !energies || (S->n > 0 && !xyz)) return fail(NBABFS_STATUS_INVALID_ARGUMENT, "null argument");
    hipStream_t st = S->stream;
    const size_t m = 3 * (size_t)S->n;
    CHK(S->xyz_stage.ensure(std::max<size_t>(m, 1)));
    if (m) HIPCHK(hipMemcpyAsync(S->xyz_stage.p, xyz, m * sizeof(double), hipMemcpyHostToDevice, st));
    S->calls++;
    CHK(update(S, S->xyz_stage.p, rebuilt));
    double *dg = nullptr;
    if (grad_inout && m) {
        CHK(S->grad_stage.ensure(m));
        if (!S->grad_assign) launch_fill_int((int *)S->grad_stage.p, 2 * m, 0, st);
        dg = S->grad_stage.p;
    }
    CHK(mmmm_async(S, S->xyz_stage.p, dg));
    if (dg) {
        CHK(S->h_grad.ensure(m));
        HIPCHK(hipMemcpyAsync(S->h_grad.p, dg, m * sizeof(double), hipMemcpyDeviceToHost, st));
    }
    CHK(nbabfs_energies(S, energies, present));
    if (dg) {
        if (S->grad_assign) std::memcpy(grad_inout, S->h_grad.p, m * sizeof(double));
        else
            for (size_t k = 0; k < m; k++) grad_inout[k] += S->h_grad.p[k];
    }
    return NBABFS_STATUS_CONTINUE;
}

extern "C" int nbabfs_set_gradient_mode(nbabfs_state *S, int mode)
{
    if (!S) return fail(NBABFS_STATUS_INVALID_ARGUMENT, "state is NULL");
    if (mode != NBABFS_GRADIENT_ADD && mode != NBABFS_GRADIENT_ASSIGN)
        return fail(NBABFS_STATUS_INVALID_ARGUMENT, "unknown gradient mode %d", mode);
    HIPCHK(hipStreamSynchronize(S->stream));
    S->grad_assign = mode == NBABFS_GRADIENT_ASSIGN ? 1 : 0;
    S->desc_valid = false;
    return NBABFS_STATUS_CONTINUE;
}

extern "C" int nbabfs_set_timing(nbabfs_state *S, int enable)
{
    if (!S) return fail(NBABFS_STATUS_INVALID_ARGUMENT, "state is NULL");
    CHK(collect_timing(S));
    S->timing = enable != 0;
    if (enable) {
        S->force_ms = 0.0;
        S->force_launches = 0;
    }
    return NBABFS_STATUS_CONTINUE;
}

extern "C" int nbabfs_stats_get(nbabfs_state *S, nbabfs_stats *out)
{
    if (!S || !out) return fail(NBABFS_STATUS_INVALID_ARGUMENT, "null argument");
    CHK(collect_timing(S));
    CHK(resolve_npairs(S));
    std::memset(out, 0, sizeof(*out));
    out->numberOfCalls = S->calls;
    out->numberOfUpdates = S->updates;
    out->numberOfMMMMPairs = S->sumPairs;
    out->npairs = S->npairs;
    out->npairs14 = S->n14;
    out->useGridSearch = S->useGrid ? 1 : 0;
    out->clusters = S->ncl;
    out->entries = S->nent;
    out->workItems = S->nitems;
    out->forceKernelMs = S->force_launches ? S->force_ms / (double)S->force_launches : 0.0;
    out->forceKernelLaunches = S->force_launches;
    out->images = S->nimages;
    out->imagePairs = S->img_npairs;
    out->numberOfMMMMImageImages = S->sumImages;
    out->numberOfMMMMImagePairs = S->sumImagePairs;
    out->qcmmElPairs = S->n_el;
    out->qcmmLJPairs = S->n_lj;
    out->numberOfQCMMElPairs = S->sumQCMMElPairs;
    out->numberOfQCMMLJPairs = S->sumQCMMLJPairs;
    return NBABFS_STATUS_CONTINUE;
}

// ------------------------------------------------------------------------------------------------
// Crystal symmetry.
// ------------------------------------------------------------------------------------------------
extern "C" int nbabfs_set_transformations(nbabfs_state *S, int ntransformations, const double *rotations,
                                          const double *translations, int checkForInverses, int imageExpandFactor)
{
    if (!S) return fail(NBABFS_STATUS_INVALID_ARGUMENT, "state is NULL");
    if (!S->isNew) return fail(NBABFS_STATUS_LOGIC_ERROR, "transformations must be set before the first update");
    if (S->has_qc && ntransformations > 0)
        return fail(NBABFS_STATUS_VALUE_ERROR, "QC/MM image lists are not supported: the state has QC atoms");
    if (S->opt.precision == NBABFS_TILED_FP64 && ntransformations > 0)
        return fail(NBABFS_STATUS_INVALID_ARGUMENT, "image lists are evaluated by the mixed and plain FP64 paths only");
    if (ntransformations < 0 || (ntransformations > 0 && (!rotations || !translations)) || imageExpandFactor < 0)
        return fail(NBABFS_STATUS_INVALID_ARGUMENT, "invalid transformations");
    S->has_sym = ntransformations > 0;
    S->ntrans = ntransformations;
    S->trot.assign(rotations, rotations + 9 * (size_t)ntransformations);
    S->ttra.assign(translations, translations + 3 * (size_t)ntransformations);
    S->checkInv = checkForInverses ? 1 : 0;
    S->expandFactor = imageExpandFactor;
    // Transformation3Container_FindIdentity / _FindInverses (rotations only, 1e-6 tolerance).
    static const double I3[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    S->identity = -1;
    for (int t = 0; t < S->ntrans; t++) {
        const double *tr = &S->ttra[3 * t];
        if (m33_equal_tol(&S->trot[9 * t], I3) && std::fabs(tr[0]) <= 1.0e-6 && std::fabs(tr[1]) <= 1.0e-6 &&
            std::fabs(tr[2]) <= 1.0e-6) {
            S->identity = t;
            break;
        }
    }
    S->tinv.assign(S->ntrans, -1);
    for (int i = 0; i < S->ntrans; i++) {
        if (S->tinv[i] >= 0) continue;
        double m[9];
        m33_invert(&S->trot[9 * i], m);
        for (int j = 0; j <= i; j++)
            if (S->tinv[j] < 0 && m33_equal_tol(m, &S->trot[9 * j])) {
                S->tinv[i] = j;
                S->tinv[j] = i;
                break;
            }
    }
    S->sp_set = false;
    return NBABFS_STATUS_CONTINUE;
}

extern "C" int nbabfs_set_symmetry_parameters(nbabfs_state *S, const double *crystal)
{
    if (!S || !crystal) return fail(NBABFS_STATUS_INVALID_ARGUMENT, "null argument");
    if (!S->has_sym) return fail(NBABFS_STATUS_LOGIC_ERROR, "the state has no transformations");
    if (!(crystal[0] > 0.0 && crystal[1] > 0.0 && crystal[2] > 0.0) || !(crystal[3] > 0.0 && crystal[3] < 180.0) ||
        !(crystal[4] > 0.0 && crystal[4] < 180.0) || !(crystal[5] > 0.0 && crystal[5] < 180.0))
        return fail(NBABFS_STATUS_INVALID_ARGUMENT, "invalid crystal parameters");
    for (int k = 0; k < 6; k++) S->sp[k] = crystal[k];
    symmetry_make_m(S->sp, S->M, S->Minv);
    if (!(std::isfinite(m33_det(S->M)) && m33_det(S->M) > 0.0))
        return fail(NBABFS_STATUS_INVALID_ARGUMENT, "crystal parameters give a singular cell");
    S->sp_set = true;
    return NBABFS_STATUS_CONTINUE;
}

// Images of the current image list (slots with pairs, reference order).
extern "C" int nbabfs_images(nbabfs_state *S, int cap, int *abct, double *scale, long long *npairs)
{
    if (!S) return -fail(NBABFS_STATUS_INVALID_ARGUMENT, "state is NULL");
    int k = 0;
    for (const auto &sl : S->slots) {
        if (sl.npairs <= 0) continue;
        if (k < cap) {
            if (abct) {
                abct[4 * k] = sl.a;
                abct[4 * k + 1] = sl.b;
                abct[4 * k + 2] = sl.c;
                abct[4 * k + 3] = sl.t;
            }
            if (scale) scale[k] = sl.scale;
            if (npairs) npairs[k] = sl.npairs;
        }
        k++;
    }
    return k;
}

// Image terms of the last energy call: the two energies and dE/dM (row-major,
// incremented) assembled from the per-slot sums h = sum g'_j and
// G_kl = sum g'_j,k x_j,l as SymmetryParameterGradients_ImageDerivatives does
// (pMolecule-1.9.0/extensions/csource/SymmetryParameterGradients.c:142-216):
// dE/dM_ab += sum_kl D_ab[k][l] G_kl + (t + n)_b h_a, with
// D_ab = M S d(M^-1)/dM_ab + E_a row_b(S M^-1).
extern "C" int nbabfs_image_terms(nbabfs_state *S, double *energies, double *dEdM, int *present)
{
    if (!S) return fail(NBABFS_STATUS_INVALID_ARGUMENT, "state is NULL");
    hipStream_t st = S->stream;
    const int ns = (int)S->slots.size();
    CHK(S->h_energies.ensure(NB_NENERGIES));
    CHK(S->energies.ensure(NB_NENERGIES));
    HIPCHK(hipMemcpyAsync(S->h_energies.p, S->energies.p, NB_NENERGIES * sizeof(double), hipMemcpyDeviceToHost, st));
    if (dEdM && ns > 0) {
        CHK(S->h_gacc.ensure((size_t)NB_GACC_WORDS_PER_IMAGE * ns));
        HIPCHK(hipMemcpyAsync(S->h_gacc.p, S->gacc.p, (size_t)NB_GACC_WORDS_PER_IMAGE * ns * sizeof(long long),
                              hipMemcpyDeviceToHost, st));
    }
    HIPCHK(hipStreamSynchronize(st));
    if (energies) {
        energies[0] = S->h_energies.p[4];
        energies[1] = S->h_energies.p[5];
    }
    if (present) *present = S->nimages > 0;
    CHK(call_status((int)S->h_energies.p[NB_ESTATUS], "energy call"));
    if (!dEdM || ns == 0) return NBABFS_STATUS_CONTINUE;
    const double *o = S->grid.corg;
    for (int k = 0; k < ns; k++) {
        const auto &sl = S->slots[k];
        if (sl.npairs <= 0) continue;
        double g[12];
        for (int q = 0; q < 12; q++) {
            long long lo = 0, hi = 0;
            for (int sl2 = 0; sl2 < NB_GSLOTS; sl2++) {
                const long long *w = S->h_gacc.p + (size_t)NB_GACC_WORDS_PER_IMAGE * k + (size_t)NB_GACC_WORDS_PER_SLOT * sl2;
                lo += w[2 * q];
                hi += w[2 * q + 1];
            }
            g[q] = (double)hi * NB_FX_HI_UNIT + (double)lo * (1.0 / NB_FX_LO_SCALE);
        }
        const double *h = g;
        double G[9];
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) G[3 * a + b] = g[3 + 3 * a + b] + h[a] * o[b];
        const double *rot = &S->trot[9 * sl.t];
        const double xtr[3] = {S->ttra[3 * sl.t] + (double)sl.a, S->ttra[3 * sl.t + 1] + (double)sl.b,
                               S->ttra[3 * sl.t + 2] + (double)sl.c};
        double ms[9], si[9];
        std::memcpy(ms, rot, sizeof(ms));
        m33_premul(ms, S->M);
        std::memcpy(si, rot, sizeof(si));
        m33_postmul(si, S->Minv);
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) {
                double di[9];
                inverse_derivative(S->M, a, b, di);
                m33_premul(di, ms);
                for (int q = 0; q < 3; q++) di[3 * a + q] += si[3 * b + q];
                double sum = 0.0;
                for (int q = 0; q < 9; q++) sum += di[q] * G[q];
                sum += xtr[b] * h[a];
                dEdM[3 * a + b] += sum;
            }
    }
    return NBABFS_STATUS_CONTINUE;
}

// Pairs of image k of the image list as sorted keys i*n + j.
extern "C" long long nbabfs_export_image_pairs(nbabfs_state *S, int image, long long *keys, long long cap)
{
    if (!S) return -fail(NBABFS_STATUS_INVALID_ARGUMENT, "state is NULL");
    int slot = -1, k = 0;
    for (int q = 0; q < (int)S->slots.size(); q++) {
        if (S->slots[q].npairs <= 0) continue;
        if (k == image) slot = q;
        k++;
    }
    if (slot < 0) return -fail(NBABFS_STATUS_INVALID_ARGUMENT, "no image %d", image);
    const long long m = S->slots[slot].npairs;
    if (!keys) return m;
    hipStream_t st = S->stream;
    DBuf<long long> dk;
    if (dk.ensure(std::max<long long>(m, 1)) != NBABFS_STATUS_CONTINUE) return -NBABFS_STATUS_OUT_OF_MEMORY;
    if (S->u64.ensure(2) != NBABFS_STATUS_CONTINUE) return -NBABFS_STATUS_OUT_OF_MEMORY;
    if (hipMemsetAsync(S->u64.p, 0, sizeof(unsigned long long), st) != hipSuccess) return -NBABFS_STATUS_LOGIC_ERROR;
    launch_export_image_pairs(S->entries.p + S->nent_main, S->nent - S->nent_main, S->slot_atom.p, S->n, slot, S->u64.p,
                              m, dk.p, st);
    std::vector<long long> hk((size_t)m);
    if (hipMemcpyAsync(hk.data(), dk.p, m * sizeof(long long), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return -fail(NBABFS_STATUS_LOGIC_ERROR, "image pair export failed");
    std::sort(hk.begin(), hk.end());
    for (long long q = 0; q < std::min(m, cap); q++) keys[q] = hk[q];
    return m;
}

extern "C" long long nbabfs_export_pairs(nbabfs_state *S, int which, long long *keys, long long cap)
{
    if (!S) return -fail(NBABFS_STATUS_INVALID_ARGUMENT, "state is NULL");
    if (which == 1) {
        long long m = (long long)S->p14_keys.size();
        if (keys)
            for (long long k = 0; k < std::min(m, cap); k++) keys[k] = S->p14_keys[k];
        return m;
    }
    if (!S->listValid) return -fail(NBABFS_STATUS_LOGIC_ERROR, "no pair list");
    if (resolve_npairs(S) != NBABFS_STATUS_CONTINUE) return -NBABFS_STATUS_LOGIC_ERROR;
    if (!keys || S->npairs == 0) return S->npairs;
    hipStream_t st = S->stream;
    DBuf<long long> dk;
    if (dk.ensure(std::max<long long>(S->npairs, 1)) != NBABFS_STATUS_CONTINUE) return -NBABFS_STATUS_OUT_OF_MEMORY;
    if (S->u64.ensure(2) != NBABFS_STATUS_CONTINUE) return -NBABFS_STATUS_OUT_OF_MEMORY;
    if (hipMemsetAsync(S->u64.p, 0, sizeof(unsigned long long), st) != hipSuccess) return -NBABFS_STATUS_LOGIC_ERROR;
    launch_export_pairs(S->entries.p, S->nent_main, S->slot_atom.p, S->n, S->u64.p, S->npairs, dk.p, st);
    std::vector<long long> hk((size_t)S->npairs);
    if (hipMemcpyAsync(hk.data(), dk.p, S->npairs * sizeof(long long), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return -fail(NBABFS_STATUS_LOGIC_ERROR, "pair export failed");
    std::sort(hk.begin(), hk.end());
    for (long long k = 0; k < std::min(S->npairs, cap); k++) keys[k] = hk[k];
    return S->npairs;
}

// ------------------------------------------------------------------------------------------------
// Device memory helpers.
// ------------------------------------------------------------------------------------------------
extern "C" int nbabfs_device_alloc(size_t bytes, void **out)
{
    if (!out) return fail(NBABFS_STATUS_INVALID_ARGUMENT, "out is NULL");
    *out = nullptr;
    CHK(nbabfs_init());
    if (hipMalloc(out, std::max<size_t>(bytes, 16)) != hipSuccess) {
        (void)hipGetLastError();
        *out = nullptr;
        return fail(NBABFS_STATUS_OUT_OF_MEMORY, "hipMalloc of %zu bytes failed", bytes);
    }
    return NBABFS_STATUS_CONTINUE;
}

extern "C" int nbabfs_device_free(void *ptr)
{
    if (ptr) HIPCHK(hipFree(ptr));
    return NBABFS_STATUS_CONTINUE;
}

extern "C" int nbabfs_memcpy_htod(nbabfs_state *S, void *dst, const void *src, size_t bytes)
{
    if (!bytes) return NBABFS_STATUS_CONTINUE;
    if (!dst || !src) return fail(NBABFS_STATUS_INVALID_ARGUMENT, "null pointer");
    if (S) {
        HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, S->stream));
        HIPCHK(hipStreamSynchronize(S->stream));
    } else {
        // The states' streams are non-blocking: a copy on the null stream does
        // not wait for their kernels, so the device is drained first.
        HIPCHK(hipDeviceSynchronize());
        HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    }
    return NBABFS_STATUS_CONTINUE;
}

extern "C" int nbabfs_memcpy_dtoh(nbabfs_state *S, void *dst, const void *src, size_t bytes)
{
    if (!bytes) return NBABFS_STATUS_CONTINUE;
    if (!dst || !src) return fail(NBABFS_STATUS_INVALID_ARGUMENT, "null pointer");
    if (S) {
        HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, S->stream));
        HIPCHK(hipStreamSynchronize(S->stream));
    } else {
        // As above: a stateless copy must not overtake work queued on the states' streams.
        HIPCHK(hipDeviceSynchronize());
        HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    }
    return NBABFS_STATUS_CONTINUE;
}

extern "C" int nbabfs_virial(nbabfs_state *S, double *virial)
{
    if (!S || !virial) return fail(NBABFS_STATUS_INVALID_ARGUMENT, "null argument");
    for (int k = 0; k < 9; k++) virial[k] = 0.0;
    // Nothing to report unless the accumulators and the descriptor are those of the last energy call.
    if (!S->vacc.p || S->n == 0 || !S->listValid || !desc_current(S)) return NBABFS_STATUS_CONTINUE;
    // From the last call's gradient, still in the fixed-point accumulators.
    launch_fill_int((int *)S->vacc.p, 2 * (size_t)NB_VACC_WORDS, 0, S->stream);
    launch_virial(S->d_desc.p, S->n, S->opt.precision == NBABFS_PLAIN_FP64 ? 1 : 0, S->stream);
    HIPCHK(hipGetLastError());
    std::vector<long long> w(NB_VACC_WORDS);
    HIPCHK(hipMemcpyAsync(w.data(), S->vacc.p, NB_VACC_WORDS * sizeof(long long), hipMemcpyDeviceToHost, S->stream));
    HIPCHK(hipStreamSynchronize(S->stream));
    // Integer sums over the slots in canonical form (low word in [0, 2^48)),
    // then one conversion: independent of which waves shared a slot.
    const long long mask = (1ll << 48) - 1;
    for (int ab = 0; ab < 9; ab++) {
        long long h = 0, l = 0;
        for (int q = 0; q < NB_ESLOTS; q++) {
            const long long lo = w[(2 * ab) * NB_ESLOTS + q], hi = w[(2 * ab + 1) * NB_ESLOTS + q];
            h += hi + (lo >> 48);
            l += lo & mask;
        }
        h += l >> 48;
        l &= mask;
        virial[ab] = (double)h * NB_FX_HI_UNIT + (double)l / NB_FX_LO_SCALE;
    }
    return NBABFS_STATUS_CONTINUE;
}

extern "C" int nbabfs_autotune(nbabfs_state *S, const double *d_xyz, int *best_entries_per_item, int *best_waves_per_cu,
                               double *best_kernel_ms)
{
    if (!S || !d_xyz) return fail(NBABFS_STATUS_INVALID_ARGUMENT, "null argument");
    return autotune(S, d_xyz, best_entries_per_item, best_waves_per_cu, best_kernel_ms);
}

extern "C" int nbabfs_zero_async(nbabfs_state *S, void *dst, size_t bytes)
{
    if (!bytes) return NBABFS_STATUS_CONTINUE;
    if (!dst || bytes % 4) return fail(NBABFS_STATUS_INVALID_ARGUMENT, "zero fill needs a 4-byte multiple");
    CHK(nbabfs_init());
    if (!S) HIPCHK(hipDeviceSynchronize());
    launch_fill_int((int *)dst, bytes / 4, 0, S ? S->stream : nullptr);
    HIPCHK(hipGetLastError());
    if (!S) HIPCHK(hipDeviceSynchronize());
    return NBABFS_STATUS_CONTINUE;
}

extern "C" int nbabfs_synchronize(nbabfs_state *S)
{
    if (S) HIPCHK(hipStreamSynchronize(S->stream));
    else HIPCHK(hipDeviceSynchronize());
    return NBABFS_STATUS_CONTINUE;
}

// ------------------------------------------------------------------------------------------------
// Batches.
// ------------------------------------------------------------------------------------------------
struct nbabfs_batch {
    std::vector<nbabfs_state *> st;
    int use_graph = 0;
    hipStream_t stream = nullptr;
    DBuf<SysDesc> d_desc;
    PinBuf<SysDesc> h_desc;
    DBuf<WorkItem> items;
    DBuf<int> heads, qstarts;
    MappedInts flags;               // [k]: epoch when system k must be rebuilt; [nsys]: epoch of the last test
    // Guard words of the batch's passes: [0] the epoch when some system must be rebuilt, [1] the test's epoch.
    DBuf<int> gguard;
    std::vector<CheckDesc> check_shadow;   // the test's descriptors on the device
    DBuf<CheckDesc> d_check;
    PinBuf<CheckDesc> h_check;
    PinBuf<double> h_energies;
    DBuf<double> energies;
    int nitems = 0, max_npad = 0, max_n = 0, max_n14b = 0;
    std::vector<SysDesc> shadow;   // the descriptors of the current layout (graph_valid implies they are current)
    bool layout_valid = false;
    std::vector<long> last_items;
    hipGraphExec_t gexec = nullptr;
    bool graph_valid = false;
    bool mixed = true;             // the tiled kernels run the pair passes (precision mixed or tiled FP64)
    int precision = NBABFS_MIXED;
    long long n_captures = 0, n_launches = 0, n_layouts = 0;
    // Timing of the batched pair launch: with timing on, the calls enqueue the
    // passes directly (no graph) with the runtime's dispatch events around the
    // pair kernel.
    bool timing = false, timed_pass = false;
    hipEvent_t ev[2] = {nullptr, nullptr};
    double force_ms = 0.0;
    long long force_launches = 0;
    ~nbabfs_batch()
    {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        for (auto e : ev)
            if (e) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

extern "C" int nbabfs_batch_create(nbabfs_state *const *states, int nsystems, int use_graph, nbabfs_batch **out)
{
    if (!out || !states || nsystems <= 0) return fail(NBABFS_STATUS_INVALID_ARGUMENT, "bad batch arguments");
    *out = nullptr;
    CHK(nbabfs_init());
    nbabfs_batch *B = new nbabfs_batch();
    for (int k = 0; k < nsystems; k++) {
        if (!states[k]) {
            delete B;
            return fail(NBABFS_STATUS_INVALID_ARGUMENT, "state %d is NULL", k);
        }
        if (states[k]->has_sym) {
            delete B;
            return fail(NBABFS_STATUS_INVALID_ARGUMENT, "state %d has crystal symmetry; batches take non-periodic states", k);
        }
        if (states[k]->has_qc) {
            delete B;
            return fail(NBABFS_STATUS_INVALID_ARGUMENT, "state %d has QC atoms; batches take MM states", k);
        }
        B->st.push_back(states[k]);
    }
    B->precision = states[0]->opt.precision;
    B->mixed = B->precision != NBABFS_PLAIN_FP64;     // the tiled kernels (mixed or full FP64)
    for (auto *s : B->st)
        if (s->opt.precision != B->precision) {
            delete B;
            return fail(NBABFS_STATUS_INVALID_ARGUMENT, "all states of a batch must share one precision");
        }
    B->use_graph = use_graph;
    if (hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking) != hipSuccess) {
        delete B;
        return fail(NBABFS_STATUS_LOGIC_ERROR, "hipStreamCreate failed");
    }
    B->last_items.assign(nsystems, -1);
    *out = B;
    return NBABFS_STATUS_CONTINUE;
}

extern "C" void nbabfs_batch_destroy(nbabfs_batch *B)
{
    if (!B) return;
    if (B->stream) (void)hipStreamSynchronize(B->stream);
    delete B;
}

// Concatenate the systems' work items and descriptors.
static int batch_layout(nbabfs_batch *B, const double *const *d_xyz, double *const *d_grad)
{
    const int ns = (int)B->st.size();
    hipStream_t st = B->stream;
    int total = 0;
    B->max_npad = B->max_n = B->max_n14b = 0;
    for (auto *s : B->st) {
        total += nblk(s->nitems, NB_TILED_WPG) * NB_TILED_WPG;
        B->max_npad = std::max(B->max_npad, s->n_pad);
        B->max_n = std::max(B->max_n, s->n);
        B->max_n14b = std::max(B->max_n14b, nblk(s->n14, NB_P14_BLOCK));
    }
    B->nitems = total;
    CHK(B->items.ensure(std::max(total, 1)));
    CHK(B->d_desc.ensure(ns));
    CHK(B->h_desc.ensure(ns));
    CHK(B->energies.ensure(NB_NENERGIES * (size_t)ns));
    CHK(B->heads.ensure(8 * 32));
    int first = 0;
    for (int k = 0; k < ns; k++) {
        nbabfs_state *s = B->st[k];
        CHK(ensure_call_buffers(s));
        SysDesc &D = B->h_desc.p[k];
        fill_desc(s, D, d_xyz[k], d_grad ? d_grad[k] : nullptr);
        D.energies = B->energies.p + NB_NENERGIES * (size_t)k;
        const int padded = nblk(s->nitems, NB_TILED_WPG) * NB_TILED_WPG;
        if (padded > 0)
            hipLaunchKernelGGL(k_remap_items, dim3(nblk(padded, 256)), dim3(256), 0, st, s->items.p, s->nitems, padded,
                               k, B->items.p + first);
        first += padded;
        B->last_items[k] = s->items_version;
    }
    HIPCHK(hipMemcpyAsync(B->d_desc.p, B->h_desc.p, ns * sizeof(SysDesc), hipMemcpyHostToDevice, st));
    B->shadow.assign(B->h_desc.p, B->h_desc.p + ns);
    std::vector<int> qs(9);
    for (int q = 0; q <= 8; q++) qs[q] = (int)((long long)total * q / 8);
    CHK(B->qstarts.upload(qs.data(), 9, st));
    HIPCHK(hipStreamSynchronize(st));
    B->layout_valid = true;
    B->n_layouts++;
    if (B->gexec) {
        (void)hipGraphExecDestroy(B->gexec);
        B->gexec = nullptr;
    }
    B->graph_valid = false;
    return NBABFS_STATUS_CONTINUE;
}

static int batch_enqueue(nbabfs_batch *B, hipEvent_t ev_start = nullptr, hipEvent_t ev_stop = nullptr)
{
    hipStream_t st = B->stream;
    const int ns = (int)B->st.size();
    nbabfs_state *s0 = B->st[0];
    const ForceGuard g{B->gguard.p, 0, B->gguard.p + 1};
    hipLaunchKernelGGL(k_batch_epoch, dim3(1), dim3(64), 0, st, B->gguard.p, B->flags.d, ns);
    if (B->max_n > 0)
        hipLaunchKernelGGL(k_check_batch, dim3(nblk(B->max_n, 256), ns), dim3(256), 0, st, B->d_check.p, B->flags.d,
                           B->gguard.p);
    PrepareCheck none{nullptr, nullptr, nullptr, 0.0, nullptr, nullptr, 0, nullptr, nullptr};
    launch_prepare(B->d_desc.p, ns, B->max_npad, B->mixed ? 1 : 0, B->heads.p, none, st);
    if (B->mixed) {
        if (B->nitems > 0) {
            QueueParams q{s0->opt.wavesPerCU > 0 ? B->heads.p : nullptr, B->qstarts.p};
            int maxnt = 0, maxcount = 0;
            for (auto *s : B->st) {
                maxnt = std::max(maxnt, s->ntypes);
                maxcount = std::max(maxcount, s->opt.entriesPerItem);
            }
            launch_force_tiled(B->d_desc.p, B->items.p, q, s0->F, n_persistent_blocks(s0->opt, B->nitems), maxnt,
                               maxcount, g, st, false, ev_start, ev_stop, 0, B->precision == NBABFS_TILED_FP64);
        }
    } else {
        launch_force_plain(B->d_desc.p, ns, B->max_n, s0->F, g, st, ev_start, ev_stop);
    }
    launch_pairs14(B->d_desc.p, ns, B->max_n14b, s0->F, st, g);
    launch_finalize(B->d_desc.p, ns, B->max_npad, B->mixed ? 0 : 1, st, g);
    HIPCHK(hipMemcpyAsync(B->h_energies.p, B->energies.p, NB_NENERGIES * (size_t)ns * sizeof(double),
                          hipMemcpyDeviceToHost, st));
    return NBABFS_STATUS_CONTINUE;
}

// A new layout (and graph) when some list was rebuilt or re-cut, or when a
// system's descriptor (its buffers, sizes or arrays) is not the one laid out.
static int batch_relayout_if_needed(nbabfs_batch *B, const double *const *d_xyz, double *const *d_grad)
{
    const int ns = (int)B->st.size();
    bool relayout = !B->layout_valid || (int)B->shadow.size() != ns;
    for (int k = 0; k < ns && !relayout; k++) {
        SysDesc D;
        fill_desc(B->st[k], D, d_xyz[k], d_grad ? d_grad[k] : nullptr);
        D.energies = B->energies.p + NB_NENERGIES * (size_t)k;
        relayout = B->last_items[k] != B->st[k]->items_version || std::memcmp(&D, &B->shadow[k], sizeof(SysDesc)) != 0;
    }
    if (relayout) CHK(batch_layout(B, d_xyz, d_grad));
    // The batch's prepare and finalize passes report into each state's error
    // word: clear what an abandoned single-state pass left there.
    for (auto *s : B->st)
        if (s->err_dirty) {
            launch_fill_int(s->errw.p, 1, 0, B->stream);
            s->err_dirty = false;
        }
    return NBABFS_STATUS_CONTINUE;
}

// The batch's passes (graph replay, or enqueued directly when timed or without graphs).
static int batch_launch(nbabfs_batch *B)
{
    hipStream_t st = B->stream;
    const bool timed = B->timing && (B->mixed ? B->nitems > 0 : true);
    B->timed_pass = timed;
    if (timed) {
        for (auto &e : B->ev)
            if (!e) HIPCHK(hipEventCreate(&e));
        CHK(batch_enqueue(B, B->ev[0], B->ev[1]));
    } else if (B->use_graph) {
        if (!B->graph_valid) {
            hipGraph_t graph = nullptr;
            HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            int rc = batch_enqueue(B);
            hipError_t ce = hipStreamEndCapture(st, &graph);
            if (rc != NBABFS_STATUS_CONTINUE) return rc;
            if (ce != hipSuccess) return fail(NBABFS_STATUS_LOGIC_ERROR, "graph capture failed: %s", hipGetErrorString(ce));
            hipError_t ie = hipGraphInstantiate(&B->gexec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (ie != hipSuccess) return fail(NBABFS_STATUS_LOGIC_ERROR, "graph instantiate failed: %s", hipGetErrorString(ie));
            B->graph_valid = true;
            B->n_captures++;
        }
        HIPCHK(hipGraphLaunch(B->gexec, st));
    } else {
        CHK(batch_enqueue(B));
    }
    return NBABFS_STATUS_CONTINUE;
}

// The timed pair launch of the last completed pass (the stream has synchronised).
static int batch_collect_timing(nbabfs_batch *B)
{
    if (!B->timed_pass) return NBABFS_STATUS_CONTINUE;
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, B->ev[0], B->ev[1]));
    B->force_ms += ms;
    B->force_launches++;
    return NBABFS_STATUS_CONTINUE;
}

// Rebuilds system k's list (NBModelABFS_Update's rebuild branch).
static int batch_rebuild(nbabfs_batch *B, int k, const double *d_xyz)
{
    nbabfs_state *s = B->st[k];
    HIPCHK(hipStreamSynchronize(s->stream));
    CHK(build_list(s, d_xyz));
    CHK(s->xyz_ref.ensure(3 * (size_t)std::max(s->n, 1)));
    if (s->n)
        HIPCHK(hipMemcpyAsync(s->xyz_ref.p, d_xyz, 3 * (size_t)s->n * sizeof(double), hipMemcpyDeviceToDevice, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    s->have_ref = true;
    count_update(s);
    return NBABFS_STATUS_CONTINUE;
}

// The update test's descriptors (uploaded when they changed: new coordinate arrays or reference copies).
static int batch_check_descs(nbabfs_batch *B, const double *const *d_xyz)
{
    const int ns = (int)B->st.size();
    std::vector<CheckDesc> cd(ns);
    for (int k = 0; k < ns; k++) {
        nbabfs_state *s = B->st[k];
        const double buf = 0.5 * (s->opt.listCutoff - s->opt.outerCutoff);
        cd[k] = CheckDesc{d_xyz[k], s->have_ref ? s->xyz_ref.p : nullptr, s->has_fixed ? s->freef.p : nullptr, s->n,
                          buf * buf};
    }
    if ((int)B->check_shadow.size() == ns && std::memcmp(cd.data(), B->check_shadow.data(), ns * sizeof(CheckDesc)) == 0)
        return NBABFS_STATUS_CONTINUE;
    // The staging copy of the last upload has completed: every call ends with a synchronisation.
    std::memcpy(B->h_check.p, cd.data(), ns * sizeof(CheckDesc));
    HIPCHK(hipMemcpyAsync(B->d_check.p, B->h_check.p, ns * sizeof(CheckDesc), hipMemcpyHostToDevice, B->stream));
    B->check_shadow = cd;
    return NBABFS_STATUS_CONTINUE;
}

extern "C" int nbabfs_batch_energy(nbabfs_batch *B, const double *const *d_xyz, double *const *d_grad,
                                   double *energies, int *rebuilt)
{
    if (!B || !d_xyz || !energies) return fail(NBABFS_STATUS_INVALID_ARGUMENT, "null argument");
    const int ns = (int)B->st.size();
    hipStream_t st = B->stream;
    for (auto *s : B->st)
        if (s->F.r2Off != B->st[0]->F.r2Off || s->F.r2On != B->st[0]->F.r2On || s->F.r2Damp != B->st[0]->F.r2Damp)
            return fail(NBABFS_STATUS_INVALID_ARGUMENT, "all states of a batch must share the cutoffs");
    CHK(B->d_check.ensure(ns));
    CHK(B->h_check.ensure(ns));
    CHK(B->flags.ensure(ns + 1));
    CHK(B->h_energies.ensure(NB_NENERGIES * (size_t)ns));
    if (!B->gguard.p) {
        CHK(B->gguard.ensure(2));
        launch_fill_int(B->gguard.p, 2, 0, st);
    }
    // Rebuilds known before any test (first call, no list) come first.
    for (int k = 0; k < ns; k++) {
        nbabfs_state *s = B->st[k];
        s->forced_xyz = nullptr;
        s->calls++;
        const bool need = s->isNew || !s->have_ref || !s->listValid;
        if (need) CHK(batch_rebuild(B, k, d_xyz[k]));
        s->isNew = false;
        if (rebuilt) rebuilt[k] = need ? 1 : 0;
    }
    // The update test of every system and the batched passes behind it, guarded by it: when no system moved
    // too far (the usual case) the call waits for the GPU once; otherwise the passes do nothing, the lists that
    // must be rebuilt are rebuilt, and the test and passes run again.
    for (int pass = 0; pass < 2; pass++) {
        CHK(batch_check_descs(B, d_xyz));
        CHK(batch_relayout_if_needed(B, d_xyz, d_grad));
        CHK(batch_launch(B));
        HIPCHK(hipStreamSynchronize(st));
        const int epoch = ((volatile int *)B->flags.h)[ns];
        bool any = false;
        for (int k = 0; k < ns; k++) {
            if (((volatile int *)B->flags.h)[k] != epoch) continue;
            if (pass == 1) return fail(NBABFS_STATUS_LOGIC_ERROR, "batch system %d moved beyond its new list", k);
            CHK(batch_rebuild(B, k, d_xyz[k]));
            if (rebuilt) rebuilt[k] = 1;
            any = true;
        }
        if (!any) break;
    }
    B->n_launches++;
    CHK(batch_collect_timing(B));
    int bad = -1;
    for (int k = 0; k < ns; k++) {
        std::memcpy(energies + 4 * (size_t)k, B->h_energies.p + NB_NENERGIES * (size_t)k, 4 * sizeof(double));
        if (bad < 0 && B->h_energies.p[NB_NENERGIES * (size_t)k + NB_ESTATUS] != 0.0) bad = k;
    }
    if (bad >= 0) {
        char what[64];
        std::snprintf(what, sizeof(what), "batch system %d", bad);
        return call_status((int)B->h_energies.p[NB_NENERGIES * (size_t)bad + NB_ESTATUS], what);
    }
    return NBABFS_STATUS_CONTINUE;
}

// Batch work-item autotuner: times the batched pair launch (dispatch events, as with timing on) on the systems'
// lists for each entriesPerItem candidate, every state re-cut alike (one work item per wave), and keeps the fastest
// in every state.  The calls pass no gradient arrays (nothing of the caller's is written) and leave the call counts
// (the states' numberOfCalls, the batch's launches) as they were.
static int batch_autotune(nbabfs_batch *B, const double *const *d_xyz, int *best_epi, double *best_ms)
{
    static const int cand[] = {8, 12, 16, 20, 24, 32, 40, 48};
    const int reps = 8;
    const int ns = (int)B->st.size();
    std::vector<double> e(4 * (size_t)ns);
    std::vector<long long> calls(ns);
    for (int k = 0; k < ns; k++) calls[k] = B->st[k]->calls;
    const bool was_timing = B->timing;
    const long long launches = B->n_launches;
    const int keep = B->st[0]->opt.entriesPerItem;
    int rc = nbabfs_batch_energy(B, d_xyz, nullptr, e.data(), nullptr);
    const int nc = (int)(sizeof(cand) / sizeof(cand[0]));
    std::vector<double> tbest(nc, 1e300);
    // Two sweeps, the second in reverse order, best of the two per candidate: a drift of the device clocks over
    // the sweep (they ramp up under load) then favours no end of the candidate list.
    for (int sweep = 0; sweep < 2 * nc && rc == NBABFS_STATUS_CONTINUE && B->mixed; sweep++) {
        const int c = sweep < nc ? sweep : 2 * nc - 1 - sweep;
        for (auto *s : B->st) {
            s->opt.entriesPerItem = cand[c];
            s->opt.wavesPerCU = 0;
            rc = recut_items(s, cand[c]);
            if (rc != NBABFS_STATUS_CONTINUE) break;
        }
        B->timing = true;
        // Untimed calls for about 20 ms first (the first re-lays the batch out): the re-cuts left the GPU idle
        // for a while and its clocks ramp up again under load.
        const auto w0 = std::chrono::steady_clock::now();
        for (int w = 0; w < 1000 && rc == NBABFS_STATUS_CONTINUE; w++) {
            rc = nbabfs_batch_energy(B, d_xyz, nullptr, e.data(), nullptr);
            if (std::chrono::steady_clock::now() - w0 > std::chrono::milliseconds(20)) break;
        }
        B->force_ms = 0.0;
        B->force_launches = 0;
        for (int r = 0; r < reps && rc == NBABFS_STATUS_CONTINUE; r++)
            rc = nbabfs_batch_energy(B, d_xyz, nullptr, e.data(), nullptr);
        if (rc == NBABFS_STATUS_CONTINUE && B->force_launches > 0)
            tbest[c] = std::min(tbest[c], B->force_ms / B->force_launches);
    }
    double best = 1e300;
    int bi = -1;
    for (int c = 0; c < nc; c++)
        if (tbest[c] < best) {
            best = tbest[c];
            bi = c;
        }
    const int pick = bi >= 0 ? cand[bi] : keep;
    for (auto *s : B->st) {
        s->opt.entriesPerItem = pick;
        const int r2 = recut_items(s, pick);
        if (rc == NBABFS_STATUS_CONTINUE) rc = r2;
    }
    B->timing = was_timing;
    B->force_ms = 0.0;
    B->force_launches = 0;
    for (int k = 0; k < ns; k++) B->st[k]->calls = calls[k];
    B->n_launches = launches;
    if (rc != NBABFS_STATUS_CONTINUE) return rc;
    if (best_epi) *best_epi = pick;
    if (best_ms) *best_ms = bi >= 0 ? best : 0.0;
    return NBABFS_STATUS_CONTINUE;
}

extern "C" int nbabfs_batch_autotune(nbabfs_batch *B, const double *const *d_xyz, int *best_entries_per_item,
                                     double *best_kernel_ms)
{
    if (!B || !d_xyz) return fail(NBABFS_STATUS_INVALID_ARGUMENT, "null argument");
    return batch_autotune(B, d_xyz, best_entries_per_item, best_kernel_ms);
}

extern "C" int nbabfs_batch_info(nbabfs_batch *B, long long *graph_captures, long long *launches, long long *layouts,
                                 double *force_kernel_ms, long long *timed_launches)
{
    if (!B) return fail(NBABFS_STATUS_INVALID_ARGUMENT, "batch is NULL");
    if (graph_captures) *graph_captures = B->n_captures;
    if (launches) *launches = B->n_launches;
    if (layouts) *layouts = B->n_layouts;
    if (force_kernel_ms) *force_kernel_ms = B->force_launches ? B->force_ms / (double)B->force_launches : 0.0;
    if (timed_launches) *timed_launches = B->force_launches;
    return NBABFS_STATUS_CONTINUE;
}

extern "C" int nbabfs_batch_set_timing(nbabfs_batch *B, int enable)
{
    if (!B) return fail(NBABFS_STATUS_INVALID_ARGUMENT, "batch is NULL");
    HIPCHK(hipStreamSynchronize(B->stream));
    B->timing = enable != 0;
    if (B->timing) {
        B->force_ms = 0.0;
        B->force_launches = 0;
    }
    return NBABFS_STATUS_CONTINUE;
}

// ------------------------------------------------------------------------------------------------
// QC/MM (NBModelABFSState_SetUp with QC atoms, pMolecule-1.9.0/extensions/csource/
// NBModelABFSState.c:323-361; GenerateLists / GenerateLists14, NBModelABFS.c:1050-1155;
// NBModelABFS_QCMMEnergyLJ / _QCMMPotentials / _QCMMGradients, NBModelABFS.c:306-498).
// ------------------------------------------------------------------------------------------------
static const double kHartreesToKJ = 2625.5;   // UNITS_ENERGY_HARTREES_TO_KILOJOULES_PER_MOLE (Units.h:52)

static QcmmGeom qcmm_geom(nbabfs_state *S, const double *d_xyz)
{
    QcmmGeom G;
    G.xyz = d_xyz;
    G.n = S->n;
    G.qc_index = S->qc_index.p;
    G.qc_partner = S->qc_partner.p;
    G.qc_lf = S->qc_lf.p;
    G.nqc = S->nqc;
    G.linkRatio = S->linkRatio;
    G.site_a = S->coupling == NBABFS_QCMM_MM ? nullptr : S->site_a.p;
    G.site_p = S->coupling == NBABFS_QCMM_MM ? nullptr : S->site_p.p;
    G.nsites = S->nsites;
    return G;
}

static int upload_csr(DBuf<int> &ptr, DBuf<int> &idx, const std::vector<std::vector<int>> &rows, hipStream_t st)
{
    std::vector<int> p(rows.size() + 1, 0), x;
    for (size_t r = 0; r < rows.size(); r++) {
        p[r + 1] = p[r] + (int)rows[r].size();
        x.insert(x.end(), rows[r].begin(), rows[r].end());
    }
    if (x.empty()) x.push_back(0);
    CHK(ptr.upload(p.data(), p.size(), st));
    CHK(idx.upload(x.data(), x.size(), st));
    return NBABFS_STATUS_CONTINUE;
}

extern "C" int nbabfs_set_qc_atoms(nbabfs_state *S, int nqc, const int *index, const int *boundary,
                                   const int *qcpartner, const double *linkfactor, int linkRatio,
                                   const int *mmpartnerOffsets, const int *mmpartners, int coupling)
{
    if (!S) return fail(NBABFS_STATUS_INVALID_ARGUMENT, "state is NULL");
    if (!S->isNew || S->listValid) return fail(NBABFS_STATUS_LOGIC_ERROR, "QC atoms must be set before the first update");
    if (S->has_qc) return fail(NBABFS_STATUS_LOGIC_ERROR, "QC atoms are already set");
    if (S->has_sym) return fail(NBABFS_STATUS_VALUE_ERROR, "QC/MM image lists are not supported: the state has symmetry");
    if (nqc <= 0 || !index) return fail(NBABFS_STATUS_INVALID_ARGUMENT, "no QC atoms");
    if (coupling < NBABFS_QCMM_MM || coupling > NBABFS_QCMM_RD)
        return fail(NBABFS_STATUS_VALUE_ERROR, "unknown QC/MM coupling %d", coupling);
    const int n = S->n;
    std::vector<unsigned char> full(n, 0), pure(n, 0), bnd(n, 0);
    for (int q = 0; q < nqc; q++) {
        if (index[q] < 0 || index[q] >= n || (q > 0 && index[q] <= index[q - 1]))
            return fail(NBABFS_STATUS_INVALID_ARGUMENT, "QC atom indices must be ascending atoms in [0,%d)", n);
        full[index[q]] = 1;
    }
    std::vector<int> part(nqc, -1);
    std::vector<double> lf(nqc, 0.0);
    std::vector<std::vector<int>> mmp(nqc);
    for (int q = 0; q < nqc; q++) {
        const bool b = boundary && boundary[q];
        if (!b) {
            pure[index[q]] = 1;
            continue;
        }
        if (!qcpartner || !linkfactor || !mmpartnerOffsets)
            return fail(NBABFS_STATUS_INVALID_ARGUMENT, "boundary atom %d without partner data", index[q]);
        bnd[index[q]] = 1;
        part[q] = qcpartner[q];
        lf[q] = linkfactor[q];
        const int o0 = mmpartnerOffsets[q], o1 = mmpartnerOffsets[q + 1];
        if (o1 < o0 || (o1 > o0 && !mmpartners))
            return fail(NBABFS_STATUS_INVALID_ARGUMENT, "bad MM partner range of boundary atom %d", index[q]);
        if (o1 == o0 && coupling != NBABFS_QCMM_MM)
            return fail(NBABFS_STATUS_VALUE_ERROR, "boundary atom %d has no MM partners (RC/RD coupling)", index[q]);
        for (int k = o0; k < o1; k++) mmp[q].push_back(mmpartners[k]);
    }
    for (int q = 0; q < nqc; q++) {
        if (part[q] < 0 && !(boundary && boundary[q])) continue;
        if (part[q] < 0 || part[q] >= n || !full[part[q]] || bnd[part[q]])
            return fail(NBABFS_STATUS_INVALID_ARGUMENT, "boundary atom %d: QC partner %d is not a pure QC atom",
                        index[q], part[q]);
        for (int p : mmp[q])
            if (p < 0 || p >= n || pure[p])
                return fail(NBABFS_STATUS_INVALID_ARGUMENT, "boundary atom %d: MM partner %d is not an MM atom",
                            index[q], p);
    }
    hipStream_t st = S->stream;
    S->nqc = nqc;
    S->coupling = coupling;
    S->linkRatio = linkRatio ? 1 : 0;
    S->qc_pure = pure;
    S->qc_index_h.assign(index, index + nqc);
    S->qc_partner_h = part;
    // Atom flags: bit 0 free, bit 1 pure QC.
    std::vector<unsigned char> fl(n);
    for (int i = 0; i < n; i++)
        fl[i] = (unsigned char)(((S->fixed_h.empty() || !S->fixed_h[i]) ? 1 : 0) | (pure[i] ? 2 : 0));
    CHK(S->freef.upload(fl.data(), fl.size(), st));
    S->has_flags = true;
    CHK(select_14(S));
    CHK(S->qc_index.upload(index, nqc, st));
    CHK(S->qc_partner.upload(part.data(), nqc, st));
    CHK(S->qc_lf.upload(lf.data(), nqc, st));
    // 1-4 neighbours and exclusion neighbours per atom.
    std::vector<std::vector<int>> n14(n);
    for (long long key : S->p14_all) {
        const int i = (int)(key / n), j = (int)(key % n);
        n14[i].push_back(j);
        n14[j].push_back(i);
    }
    auto freeat = [&](int a) { return S->fixed_h.empty() || !S->fixed_h[a]; };
    // MM sites and charges: MMAtomContainer_GetCharges of the active atoms (MM coupling) or
    // QCAtomContainer_GetMMCharges (RC/RD, QCAtomContainer.c:465-525).
    std::vector<double> sq;
    std::vector<int> sa, sp;
    if (coupling == NBABFS_QCMM_MM) {
        sq = S->charges_h;
        for (int i = 0; i < n; i++)
            if (pure[i]) sq[i] = 0.0;
        S->nsites = n;
    } else {
        std::vector<int> pos(n, -1);
        for (int i = 0; i < n; i++)
            if (!full[i]) {
                pos[i] = (int)sa.size();
                sa.push_back(i);
                sp.push_back(-1);
                sq.push_back(S->charges_h[i]);
            }
        for (int q = 0; q < nqc; q++) {
            if (part[q] < 0) continue;
            const int b = index[q];
            const double q0 = S->charges_h[b] / (double)mmp[q].size();
            for (int p : mmp[q]) {
                sa.push_back(b);
                sp.push_back(p);
                if (bnd[p] || coupling == NBABFS_QCMM_RC) {
                    sq.push_back(q0);
                } else {
                    sq.push_back(2.0 * q0);
                    sq[pos[p]] -= q0;
                }
            }
        }
        S->nsites = (int)sa.size();
        CHK(S->site_a.upload(sa.data(), sa.size(), st));
        CHK(S->site_p.upload(sp.data(), sp.size(), st));
    }
    CHK(S->site_q.upload(sq.data(), sq.size(), st));
    // qcmmExclusions (SelfPairList_ToCrossPairList of the exclusions, QC atoms x MM selection,
    // the boundary atom itself included) and the QC/MM electrostatic 1-4 list (MM coupling only).
    std::vector<std::vector<int>> qex(nqc), e14(nqc);
    for (int q = 0; q < nqc; q++) {
        const int i = index[q];
        for (int k = S->xptr_h[i]; k < S->xptr_h[i + 1]; k++)
            if (!pure[S->xidx_h[k]]) qex[q].push_back(S->xidx_h[k]);
        if (!pure[i]) qex[q].push_back(i);
        std::sort(qex[q].begin(), qex[q].end());
        if (coupling == NBABFS_QCMM_MM)
            for (int j : n14[i])
                if (!pure[j] && j != i) e14[q].push_back(j);
        std::sort(e14[q].begin(), e14[q].end());
    }
    if (coupling == NBABFS_QCMM_MM) CHK(upload_csr(S->qex_ptr, S->qex_idx, qex, st));
    CHK(upload_csr(S->el14_ptr, S->el14_idx, e14, st));
    S->n_el14 = 0;
    for (auto &r : e14) S->n_el14 += (long long)r.size();
    // QC/MM LJ rows (pure QC atoms) and their 1-4 lists (SelfPairList_ToCrossPairList with the
    // free atoms as the OR selection).
    std::vector<int> ljat;
    std::vector<std::vector<int>> l14;
    for (int q = 0; q < nqc; q++) {
        const int i = index[q];
        if (!pure[i]) continue;
        ljat.push_back(i);
        std::vector<int> r;
        for (int j : n14[i])
            if (!pure[j] && (freeat(i) || freeat(j))) r.push_back(j);
        std::sort(r.begin(), r.end());
        l14.push_back(r);
    }
    S->nlj = (int)ljat.size();
    S->lj_atom_h = ljat;
    if (ljat.empty()) ljat.push_back(0);
    CHK(S->lj_atom.upload(ljat.data(), ljat.size(), st));
    CHK(upload_csr(S->lj14_ptr, S->lj14_idx, l14, st));
    S->n_lj14 = 0;
    for (auto &r : l14) S->n_lj14 += (long long)r.size();
    HIPCHK(hipStreamSynchronize(st));
    S->has_qc = true;
    S->desc_valid = false;
    return NBABFS_STATUS_CONTINUE;
}

// Tridiagonal solve by Gaussian elimination with partial pivoting (LAPACK dgtsv's
// published algorithm, the solver CubicSpline_MakeFromReal1DArrays calls): sub-diagonal dl
// (n-1), diagonal d (n), super-diagonal du (n-1), right-hand side b (n, overwritten).
static bool tridiagonal_solve(std::vector<double> dl, std::vector<double> d, std::vector<double> du,
                              std::vector<double> &b)
{
#pragma clang fp contract(off)
    const int n = (int)d.size();
    for (int i = 0; i + 1 < n; i++) {
        const bool last = i == n - 2;
        if (std::fabs(d[i]) >= std::fabs(dl[i])) {
            if (d[i] == 0.0) return false;
            const double fact = dl[i] / d[i];
            d[i + 1] = d[i + 1] - fact * du[i];
            b[i + 1] = b[i + 1] - fact * b[i];
            if (!last) dl[i] = 0.0;
        } else {
            const double fact = d[i] / dl[i];
            d[i] = dl[i];
            const double temp = d[i + 1];
            d[i + 1] = du[i] - fact * temp;
            if (!last) {
                dl[i] = du[i + 1];
                du[i + 1] = -fact * dl[i];
            }
            du[i] = temp;
            const double tb = b[i];
            b[i] = b[i + 1];
            b[i + 1] = tb - fact * b[i + 1];
        }
    }
    if (d[n - 1] == 0.0) return false;
    b[n - 1] = b[n - 1] / d[n - 1];
    if (n > 1) b[n - 2] = (b[n - 2] - du[n - 2] * b[n - 1]) / d[n - 2];
    for (int i = n - 3; i >= 0; i--) b[i] = (b[i] - du[i] * b[i + 1] - dl[i] * b[i + 2]) / d[i];
    return true;
}

extern "C" int nbabfs_set_qcmm_spline(nbabfs_state *S, int npoints, const double *x, const double *y)
{
#pragma clang fp contract(off)
    if (!S || !x || !y) return fail(NBABFS_STATUS_INVALID_ARGUMENT, "null argument");
    if (npoints < 2) return fail(NBABFS_STATUS_VALUE_ERROR, "a spline needs at least two points");
    std::vector<double> X(x, x + npoints), Y(y, y + npoints);
    // Decreasing abscissae are reversed, as the reference does.
    if (X[0] > X[npoints - 1]) {
        std::reverse(X.begin(), X.end());
        std::reverse(Y.begin(), Y.end());
    }
    for (int i = 1; i < npoints; i++)
        if (!(X[i] > X[i - 1])) return fail(NBABFS_STATUS_VALUE_ERROR, "spline abscissae must be strictly monotonic");
    const int n = npoints;
    std::vector<double> dl(n - 1), d(n), du(n - 1), b(n);
    // Clamped ends (first derivatives 0) and the interior continuity conditions.
    double h0 = X[1] - X[0];
    d[0] = h0 / 3.0;
    du[0] = h0 / 6.0;
    b[0] = (Y[1] - Y[0]) / h0 - 0.0;
    for (int i = 1; i < n - 1; i++) {
        const double hl = X[i] - X[i - 1], hu = X[i + 1] - X[i];
        dl[i - 1] = hl / 6.0;
        d[i] = (hl + hu) / 3.0;
        du[i] = hu / 6.0;
        b[i] = (Y[i + 1] - Y[i]) / hu + (Y[i - 1] - Y[i]) / hl;
    }
    const double hn = X[n - 1] - X[n - 2];
    dl[n - 2] = hn / 6.0;
    d[n - 1] = hn / 3.0;
    b[n - 1] = 0.0 - (Y[n - 1] - Y[n - 2]) / hn;
    if (!tridiagonal_solve(dl, d, du, b)) return fail(NBABFS_STATUS_VALUE_ERROR, "singular spline system");
    hipStream_t st = S->stream;
    CHK(S->spx.upload(X.data(), n, st));
    CHK(S->spy.upload(Y.data(), n, st));
    CHK(S->sph.upload(b.data(), n, st));
    HIPCHK(hipStreamSynchronize(st));
    S->nspline = n;
    return NBABFS_STATUS_CONTINUE;
}

// The QC/MM electrostatic list (QC atoms x MM sites, with the qcmmExclusions for MM coupling)
// and the QC/MM LJ list (pure QC atoms x MM atoms, exclusions and the free-atom OR test), each
// by a count pass, a scan and a fill pass.
static int build_qcmm_lists(nbabfs_state *S, const double *d_xyz)
{
    hipStream_t st = S->stream;
    const QcmmGeom G = qcmm_geom(S, d_xyz);
    const double c2 = S->opt.listCutoff * S->opt.listCutoff;
    CHK(S->h_int.ensure(2));
    // Grid method: the cross lists search the grids of NBModelABFSState_GridInitialize (pMolecule-1.9.0/
    // extensions/csource/NBModelABFSState.c:207-226): the electrostatic list the QC coordinates' grid when there is
    // more than one QC atom, the LJ list the grid of all atoms (the main list's).
    if (S->useGrid) {
        CHK(S->qgrid.ensure(2));
        CHK(S->qgpart.ensure(12 * NB_QGRID_PARTS));
        launch_qcmm_grids(G, S->freef.p, S->grid.lower, S->grid.bins, S->grid.h, S->qgpart.p, S->qgrid.p, st);
        CHK(S->qcolcell.ensure(std::max(std::max(S->n, S->nsites), 1)));
        launch_qcmm_colcells(G, S->qgrid.p, S->grid.h, S->qcolcell.p, st);
    }
    for (int list = 0; list < 2; list++) {
        CrossListParams P{};
        P.G = G;
        P.cutoff2 = c2;
        P.flags = S->freef.p;
        if (S->useGrid && (list == 1 || S->nqc > 1)) {
            P.grid = S->qgrid.p + list;
            // The LJ list's columns are the atoms on the main list's grid: their cells from the binning pass.
            P.colcell = list == 1 ? S->cellv.p : S->qcolcell.p;
            P.stencil = S->stencil.p;
            P.u = S->u;
            P.h = S->grid.h;
            P.pointcell = S->opt.useGridByCell ? 0 : 1;
        }
        DBuf<int> &ptr = list ? S->lj_ptr : S->el_ptr;
        DBuf<int> &idx = list ? S->lj_idx : S->el_idx;
        if (list == 0) {
            P.rows_are_qc = 1;
            P.nrows = S->nqc;
            P.col_atoms = S->coupling == NBABFS_QCMM_MM;
            if (P.col_atoms) {
                P.ex_ptr = S->qex_ptr.p;
                P.ex_idx = S->qex_idx.p;
            }
        } else {
            P.rows_are_qc = 0;
            P.row_atom = S->lj_atom.p;
            P.nrows = S->nlj;
            P.col_atoms = 1;
            P.or_test = 1;
            P.ex_ptr = S->xptr.p;
            P.ex_idx = S->xidx.p;
        }
        CHK(ptr.ensure(P.nrows + 1));
        long long total = 0;
        if (P.nrows > 0) {
            // Counts per (row, column chunk) plus a zero, their exclusive scan, the row pointers from it.
            const int nchunk = cross_list_chunks(P), m = P.nrows * nchunk + 1;
            CHK(S->qcnt.ensure(m));
            CHK(S->qoff.ensure(m));
            CHK(S->scantmp.ensure(scan_tmp_size(m)));
            launch_fill_int(S->qcnt.p, m, 0, st);
            P.count = S->qcnt.p;
            launch_cross_list(P, 0, st);
            launch_scan(S->qcnt.p, S->qoff.p, m, S->scantmp.p, st);
            launch_row_pointers(S->qoff.p, P.nrows, nchunk, ptr.p, st);
            HIPCHK(hipMemcpyAsync(S->h_int.p, ptr.p + P.nrows, sizeof(int), hipMemcpyDeviceToHost, st));
            S->h_int.p[1] = 0;
            if (P.grid)
                HIPCHK(hipMemcpyAsync(S->h_int.p + 1, &P.grid->bad, sizeof(int), hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
            // A grid axis with no cells has no occupancy (RegularGridOccupancy_Allocate rejects zero cells,
            // pCore-1.9.0/extensions/csource/RegularGridOccupancy.c:29,54).
            if (S->h_int.p[1])
                return fail(NBABFS_STATUS_INVALID_ARGUMENT,
                            "the QC atoms' cell grid has an axis with no cells (zero extent)");
            total = S->h_int.p[0];
        } else {
            launch_fill_int(ptr.p, 1, 0, st);
        }
        CHK(idx.ensure(std::max<long long>(total, 1)));
        if (total > 0) {
            P.off = S->qoff.p;
            P.idx = idx.p;
            launch_cross_list(P, 1, st);
        }
        HIPCHK(hipGetLastError());
        if (list == 0) S->n_el = total;
        else S->n_lj = total;
    }
    S->sumQCMMElPairs += (double)S->n_el;
    S->sumQCMMLJPairs += (double)S->n_lj;
    S->desc_valid = false;
    return NBABFS_STATUS_CONTINUE;
}

static int qcmm_ready(nbabfs_state *S, bool spline)
{
    if (!S) return fail(NBABFS_STATUS_INVALID_ARGUMENT, "state is NULL");
    if (!S->has_qc) return fail(NBABFS_STATUS_LOGIC_ERROR, "the state has no QC atoms");
    if (!S->listValid || !S->upd_xyz) return fail(NBABFS_STATUS_LOGIC_ERROR, "no lists: call an update first");
    if (spline && S->nspline < 2) return fail(NBABFS_STATUS_LOGIC_ERROR, "no QC/MM spline: call nbabfs_set_qcmm_spline");
    return NBABFS_STATUS_CONTINUE;
}

extern "C" int nbabfs_qcmm_terms(nbabfs_state *S, double *energies, int *present)
{
    if (!S || !energies) return fail(NBABFS_STATUS_INVALID_ARGUMENT, "null argument");
    if (!S->has_qc) return fail(NBABFS_STATUS_LOGIC_ERROR, "the state has no QC atoms");
    CHK(S->h_energies.ensure(NB_NENERGIES));
    CHK(S->energies.ensure(NB_NENERGIES));
    HIPCHK(hipMemcpyAsync(S->h_energies.p, S->energies.p, NB_NENERGIES * sizeof(double), hipMemcpyDeviceToHost,
                          S->stream));
    HIPCHK(hipStreamSynchronize(S->stream));
    energies[0] = S->h_energies.p[6];
    energies[1] = S->h_energies.p[7];
    if (present) {
        present[0] = S->n_lj > 0;
        present[1] = S->n_lj14 > 0;
    }
    return NBABFS_STATUS_CONTINUE;
}

// QC/MM potentials (NBModelABFS_QCMMPotentials) into the per-(QC atom, chunk)
// partials on the state's stream.
static int qcmm_potentials_enqueue(nbabfs_state *S)
{
    hipStream_t st = S->stream;
    const QcmmGeom G = qcmm_geom(S, S->upd_xyz);
    const SplineDev sp{S->spx.p, S->spy.p, S->sph.p, S->nspline};
    const bool mm = S->coupling == NBABFS_QCMM_MM;
    const QcmmRows L{S->el_ptr.p, S->el_idx.p, mm ? S->el14_ptr.p : nullptr, mm ? S->el14_idx.p : nullptr};
    double escale, escale14;
    {
#pragma clang fp contract(off)
        escale = 1.0 / S->opt.dielectric;
        escale14 = escale * S->opt.electrostaticScale14;
    }
    CHK(S->qpot.ensure(2 * (size_t)S->nqc * NB_QCMM_CHUNKS));
    launch_qcmm_potentials(G, sp, L, S->site_q.p, escale, escale14, S->opt.outerCutoff * S->opt.outerCutoff,
                           NB_QCMM_CHUNKS, S->qpot.p, st);
    HIPCHK(hipGetLastError());
    return NBABFS_STATUS_CONTINUE;
}

extern "C" int nbabfs_qcmm_potentials(nbabfs_state *S, double *potentials)
{
    CHK(qcmm_ready(S, true));
    if (!potentials) return fail(NBABFS_STATUS_INVALID_ARGUMENT, "potentials is NULL");
    hipStream_t st = S->stream;
    const int ch = NB_QCMM_CHUNKS;
    const size_t np = 2 * (size_t)S->nqc * ch;
    CHK(qcmm_potentials_enqueue(S));
    CHK(S->h_qpot.ensure(np));
    HIPCHK(hipMemcpyAsync(S->h_qpot.p, S->qpot.p, np * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    // Chunk sums in chunk order; the main list's total, then the 1-4 list's, as the reference adds them.
    for (int q = 0; q < S->nqc; q++) {
        double m = 0.0, m14 = 0.0;
        for (int c = 0; c < ch; c++) {
            m += S->h_qpot.p[2 * ((size_t)q * ch + c)];
            m14 += S->h_qpot.p[2 * ((size_t)q * ch + c) + 1];
        }
        potentials[q] += m;
        potentials[q] += m14;
    }
    return NBABFS_STATUS_CONTINUE;
}

extern "C" int nbabfs_qcmm_potentials_async(nbabfs_state *S, double *d_potentials)
{
    CHK(qcmm_ready(S, true));
    if (!d_potentials) return fail(NBABFS_STATUS_INVALID_ARGUMENT, "potentials is NULL");
    CHK(qcmm_potentials_enqueue(S));
    launch_qcmm_pot_sum(S->qpot.p, S->nqc, NB_QCMM_CHUNKS, d_potentials, S->stream);
    HIPCHK(hipGetLastError());
    return NBABFS_STATUS_CONTINUE;
}

// QC/MM electrostatic gradients (NBModelABFS_QCMMGradients) for the QC charges
// at d_q (device) into the state's fixed-point accumulators, then out: written
// to out (add = false) or added into it.
static int qcmm_gradients_enqueue(nbabfs_state *S, const double *d_q, double *out, bool add)
{
    hipStream_t st = S->stream;
    const int n = S->n;
    const QcmmGeom G = qcmm_geom(S, S->upd_xyz);
    const SplineDev sp{S->spx.p, S->spy.p, S->sph.p, S->nspline};
    const bool mm = S->coupling == NBABFS_QCMM_MM;
    const QcmmRows L{S->el_ptr.p, S->el_idx.p, mm ? S->el14_ptr.p : nullptr, mm ? S->el14_idx.p : nullptr};
    double scale, scale14;
    {
#pragma clang fp contract(off)
        const double e = 1.0 / S->opt.dielectric;
        scale = kHartreesToKJ * e;
        scale14 = kHartreesToKJ * (e * S->opt.electrostaticScale14);
    }
    CHK(S->qgacc.ensure(6 * (size_t)std::max(n, 1)));
    launch_fill_int((int *)S->qgacc.p, 12 * (size_t)n, 0, st);
    launch_qcmm_gradients(G, sp, L, S->site_q.p, d_q, scale, scale14, S->opt.outerCutoff * S->opt.outerCutoff,
                          NB_QCMM_CHUNKS, S->qgacc.p, st);
    launch_qcmm_grad_out(S->qgacc.p, n, out, add, st);
    HIPCHK(hipGetLastError());
    return NBABFS_STATUS_CONTINUE;
}

extern "C" int nbabfs_qcmm_gradients_host(nbabfs_state *S, const double *qcCharges, double *grad_inout)
{
    CHK(qcmm_ready(S, true));
    if (!qcCharges || !grad_inout) return fail(NBABFS_STATUS_INVALID_ARGUMENT, "null argument");
    hipStream_t st = S->stream;
    const int n = S->n;
    CHK(S->qcq.upload(qcCharges, S->nqc, st));
    CHK(S->qgrad.ensure(3 * (size_t)std::max(n, 1)));
    CHK(S->h_qgrad.ensure(3 * (size_t)std::max(n, 1)));
    CHK(qcmm_gradients_enqueue(S, S->qcq.p, S->qgrad.p, false));
    HIPCHK(hipMemcpyAsync(S->h_qgrad.p, S->qgrad.p, 3 * (size_t)n * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    for (size_t k = 0; k < 3 * (size_t)n; k++) grad_inout[k] += S->h_qgrad.p[k];
    return NBABFS_STATUS_CONTINUE;
}

extern "C" int nbabfs_qcmm_gradients_async(nbabfs_state *S, const double *d_qcCharges, double *d_grad)
{
    CHK(qcmm_ready(S, true));
    if (!d_qcCharges || !d_grad) return fail(NBABFS_STATUS_INVALID_ARGUMENT, "null argument");
    return qcmm_gradients_enqueue(S, d_qcCharges, d_grad, true);
}

extern "C" long long nbabfs_export_qcmm_pairs(nbabfs_state *S, int which, long long *keys, long long cap)
{
    if (qcmm_ready(S, false) != NBABFS_STATUS_CONTINUE) return -1;
    if (which < 0 || which > 3) return fail(NBABFS_STATUS_INVALID_ARGUMENT, "which must be 0..3"), -1;
    hipStream_t st = S->stream;
    const bool rows_qc = which == 0 || which == 2;
    const int nrows = rows_qc ? S->nqc : S->nlj;
    DBuf<int> &ptr = which == 0 ? S->el_ptr : which == 1 ? S->lj_ptr : which == 2 ? S->el14_ptr : S->lj14_ptr;
    DBuf<int> &idx = which == 0 ? S->el_idx : which == 1 ? S->lj_idx : which == 2 ? S->el14_idx : S->lj14_idx;
    std::vector<int> p(nrows + 1, 0);
    if (nrows > 0) HIPCHK(hipMemcpy(p.data(), ptr.p, (nrows + 1) * sizeof(int), hipMemcpyDeviceToHost));
    const long long total = p[nrows];
    std::vector<int> x(std::max<long long>(total, 1));
    if (total > 0) HIPCHK(hipMemcpy(x.data(), idx.p, total * sizeof(int), hipMemcpyDeviceToHost));
    if (!keys) return total;
    std::vector<long long> k;
    k.reserve(total);
    for (int r = 0; r < nrows; r++) {
        const long long row = rows_qc ? r : S->lj_atom_h[r];
        for (int e = p[r]; e < p[r + 1]; e++) k.push_back((row << 32) | (long long)x[e]);
    }
    std::sort(k.begin(), k.end());
    for (long long e = 0; e < total && e < cap; e++) keys[e] = k[e];
    (void)st;
    return total;
}

extern "C" long long nbabfs_qcmm_sites(nbabfs_state *S, double *qcxyz, double *sitexyz, double *siteq, long long cap)
{
    if (qcmm_ready(S, false) != NBABFS_STATUS_CONTINUE) return -1;
    hipStream_t st = S->stream;
    const QcmmGeom G = qcmm_geom(S, S->upd_xyz);
    const int m = S->nsites;
    CHK(S->qxyz_out.ensure(3 * ((size_t)S->nqc + m)));
    launch_qcmm_coords(G, S->qxyz_out.p, S->qxyz_out.p + 3 * (size_t)S->nqc, st);
    HIPCHK(hipGetLastError());
    std::vector<double> h(3 * ((size_t)S->nqc + m)), q(m);
    HIPCHK(hipMemcpyAsync(h.data(), S->qxyz_out.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(q.data(), S->site_q.p, m * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (qcxyz) std::memcpy(qcxyz, h.data(), 3 * (size_t)S->nqc * sizeof(double));
    const long long w = std::min<long long>(cap, m);
    if (sitexyz) std::memcpy(sitexyz, h.data() + 3 * (size_t)S->nqc, 3 * (size_t)w * sizeof(double));
    if (siteq) std::memcpy(siteq, q.data(), (size_t)w * sizeof(double));
    return m;
}
