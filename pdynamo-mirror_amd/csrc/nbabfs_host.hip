// nbabfs_host.hip -- host driver and C ABI of libnbabfs.so.
//
// The state object plays the role of NBModelABFSState (pMolecule-1.9.0/
// extensions/cinclude/NBModelABFSState.h:30-125): a per-configuration cache of
// device lists and reference coordinates, rebuilt on demand by the rule of
// NBModelABFS_Update (pMolecule-1.9.0/extensions/csource/NBModelABFS.c:508-623).
#include "nbabfs.h"
#include "nbabfs_kernels.h"

#include <algorithm>
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

static long g_alloc_version = 0;

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
        g_alloc_version++;
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
        g_alloc_version++;
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
}

__global__ void k_remap_items(const WorkItem *__restrict__ src, int count, int sys, WorkItem *__restrict__ dst)
{
    int k = blockIdx.x * 256 + threadIdx.x;
    if (k < count) {
        WorkItem w = src[k];
        w.sys = sys;
        dst[k] = w;
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

__global__ __launch_bounds__(256) void k_check_batch(const CheckDesc *__restrict__ cd, int *__restrict__ flags, int epoch)
{
#pragma clang fp contract(off)
    const CheckDesc &C = cd[blockIdx.y];
    int i = blockIdx.x * 256 + threadIdx.x;
    bool up = false;
    if (C.ref && i < C.n && (!C.free_flags || C.free_flags[i])) {
        double dx = C.xyz[3 * (size_t)i] - C.ref[3 * (size_t)i];
        double dy = C.xyz[3 * (size_t)i + 1] - C.ref[3 * (size_t)i + 1];
        double dz = C.xyz[3 * (size_t)i + 2] - C.ref[3 * (size_t)i + 2];
        up = dx * dx + dy * dy + dz * dz > C.bufsq;
    }
    if (__any(up) && threadIdx.x % 64 == 0) *(volatile int *)&flags[blockIdx.y] = epoch;
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
    std::vector<long long> p14_keys;   // canonical 1-4 set

    // Constant device data.
    DBuf<double> charge, ljA, ljB, ljA14, ljB14, ljrow;
    DBuf<int> atype, xptr, xidx, p14;
    DBuf<unsigned char> freef;
    bool has_fixed = false;

    // Coordinates.
    DBuf<double> xyz_stage, xyz_ref, grad_stage;
    bool have_ref = false;

    // Update state (NBModelABFSState.isNew / useGridSearch).
    bool isNew = true, useGrid = false, listValid = false, csrValid = false;
    long long calls = 0, updates = 0, npairs = 0;
    double sumPairs = 0.0;

    // List.
    BuildGrid grid;
    int u = 0;
    DBuf<unsigned char> stencil;
    int n_pad = 0, ncl = 0, nsc = 0, nent = 0, nitems = 0, cap = 512;
    DBuf<int> slot_atom, atom_slot;
    DBuf<int4> cellv;
    DBuf<float4> blo, bhi, centre;
    DBuf<Entry> entries, scratch;
    DBuf<WorkItem> items, items2;
    DBuf<int> icost, ihist, icur;
    DBuf<int> rowptr, cols;

    // Per call.
    DBuf<double4> xyzq;
    DBuf<int> styp;
    DBuf<long long> facc;
    DBuf<double> epart14, eplain, gplain, energies;
    DBuf<long long> eacc, vacc;
    DBuf<SysDesc> d_desc;
    DBuf<int> heads, qstarts;
    PinBuf<SysDesc> h_desc;
    PinBuf<double> h_energies, h_grad;
    long desc_version = -1;
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

    // Kernel timing.
    bool timing = false;
    std::vector<hipEvent_t> ev;
    size_t ev_used = 0;
    double force_ms = 0.0;
    long long force_launches = 0;

    ~nbabfs_state()
    {
        for (auto e : ev) (void)hipEventDestroy(e);
        if (own_stream && stream) (void)hipStreamDestroy(stream);
    }
};

// Sum of timed force launches recorded so far (synchronises on the last event).
static int collect_timing(nbabfs_state *S)
{
    if (S->ev_used == 0) return NBABFS_STATUS_CONTINUE;
    HIPCHK(hipEventSynchronize(S->ev[S->ev_used - 1]));
    for (size_t k = 0; k + 1 < S->ev_used; k += 2) {
        float ms = 0.0f;
        HIPCHK(hipEventElapsedTime(&ms, S->ev[k], S->ev[k + 1]));
        S->force_ms += ms;
        S->force_launches++;
    }
    S->ev_used = 0;
    return NBABFS_STATUS_CONTINUE;
}

static int record_event(nbabfs_state *S)
{
    if (S->ev_used == S->ev.size()) {
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
    HIPCHK(hipEventRecord(S->ev[S->ev_used++], S->stream));
    return NBABFS_STATUS_CONTINUE;
}

// ------------------------------------------------------------------------------------------------
// Grid parameters (host FP64, reference operation order).
// ------------------------------------------------------------------------------------------------
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

// Work items longest-first (and, for the persistent queues, interleaved over
// the eight per-XCD queues), then the queue bounds.
static int order_items(nbabfs_state *S)
{
    hipStream_t st = S->stream;
    const int n = S->nitems;
    const int interleave = S->opt.wavesPerCU > 0 ? 1 : 0;
    std::vector<int> qs(9);
    if (n > 0) {
        const int maxcost = NB_SCC * std::max(1, S->opt.entriesPerItem);
        CHK(S->icost.ensure(n));
        CHK(S->ihist.ensure(maxcost + 1));
        CHK(S->icur.ensure(maxcost + 1));
        CHK(S->items2.ensure(n));
        CHK(S->scantmp.ensure(scan_tmp_size(maxcost + 1)));
        launch_order_items(S->items.p, n, S->entries.p, maxcost, interleave, S->icost.p, S->ihist.p, S->icur.p,
                           S->scantmp.p, S->items2.p, st);
        HIPCHK(hipGetLastError());
        S->items.swap(S->items2);
    }
    qs[0] = 0;
    for (int q = 0; q < 8; q++) qs[q + 1] = interleave ? qs[q] + (n - q + 7) / 8 * (n > q) : (int)((long long)n * (q + 1) / 8);
    CHK(S->qstarts.upload(qs.data(), 9, st));
    return NBABFS_STATUS_CONTINUE;
}

// ------------------------------------------------------------------------------------------------
// List build (NBModelABFSState_GridInitialize + GenerateLists).
// ------------------------------------------------------------------------------------------------
static int build_list(nbabfs_state *S, const double *d_xyz)
{
    hipStream_t st = S->stream;
    const int n = S->n;
    const nbabfs_options &o = S->opt;
    S->listValid = false;
    S->csrValid = false;
    S->npairs = 0;
    S->nent = 0;
    S->nitems = 0;
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

    // Grid vs direct is decided once, on the first build (NBModelABFSState.c:212).
    if (S->isNew) S->useGrid = determine_method(o, n, bb);
    BuildGrid g;
    std::memset(&g, 0, sizeof(g));
    g.use_grid = S->useGrid ? 1 : 0;
    std::vector<unsigned char> flags;
    const double cell = o.cutoffCellSizeFactor * o.listCutoff;
    if (S->useGrid) {
        reference_grid(bb, cell, g);
        make_stencil(cell, o.listCutoff, S->u, flags);
    } else {
        g.h = 1.0;
        S->u = 0;
        flags.assign(1, 1);
    }
    CHK(S->stencil.upload(flags.data(), flags.size(), st));

    // Cluster columns: width so that an 8-atom cluster is roughly cubic.
    double ext[3], vol = 1.0;
    for (int d = 0; d < 3; d++) {
        ext[d] = bb[3 + d] - bb[d];
        vol *= std::max(ext[d], 1.0);
        g.corg[d] = bb[d];
    }
    double cw = std::cbrt((double)NB_CL * vol / (double)n);
    cw = std::min(std::max(cw, 1.5), 1.0e4);
    g.cw = cw;
    g.ncx = std::max(1, (int)std::ceil(ext[0] / cw));
    g.ncy = std::max(1, (int)std::ceil(ext[1] / cw));
    double dz = 0.25 * cw;
    long long maxkeys = 8LL * n + 4096;
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
    launch_sort_slots(S->key.p, n, nkeys, ncol, g.nzb, S->off.p, S->cursor.p, S->sorted.p, S->padded.p, st);
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
    launch_slots(S->sorted.p, S->key.p, n, g.nzb, S->off.p, S->colstart.p, S->padded.p, ncol, S->slot_atom.p,
                 S->atom_slot.p, S->colcl.p, st);
    CHK(S->blo.ensure(S->ncl));
    CHK(S->bhi.ensure(S->ncl));
    CHK(S->centre.ensure(S->ncl));
    launch_cluster_geom(d_xyz, S->slot_atom.p, S->ncl, g.corg, S->blo.p, S->bhi.p, S->centre.p, st);

    // Cluster-pair search with the exact reference predicate.
    SearchParams P;
    P.cutoff2 = o.listCutoff * o.listCutoff;
    P.stencil = S->stencil.p;
    P.u = S->u;
    P.use_grid = g.use_grid;
    P.xptr = S->xptr.p;
    P.xidx = S->xidx.p;
    P.free_flags = S->has_fixed ? S->freef.p : nullptr;
    P.reach = (float)(o.listCutoff + 0.01);
    P.cw = (float)cw;
    P.ncx = g.ncx;
    P.ncy = g.ncy;
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
        launch_search(P, S->nsc, d_xyz, S->cellv.p, S->slot_atom.p, S->blo.p, S->bhi.p, S->colcl.p, S->scratch.p,
                      S->ncand.p, S->counts.p, S->flag.p, st);
        launch_scan(S->counts.p, S->eoff.p, S->nsc, S->scantmp.p, st);
        launch_items_per_sc(S->counts.p, S->nsc, per_item, S->nits.p, st);
        launch_scan(S->nits.p, S->ioff.p, S->nsc, S->scantmp.p, st);
        hipLaunchKernelGGL(k_list_totals, dim3(1), dim3(64), 0, st, S->eoff.p, S->counts.p, S->ioff.p, S->nits.p,
                           S->nsc, S->flag.p, S->tail.d);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(st));
        const int need = ((volatile int *)S->tail.h)[2];
        if (need <= 0) break;
        if (attempt == 3) return fail(NBABFS_STATUS_LOGIC_ERROR, "pair search capacity did not converge");
        S->cap = need + need / 8 + 64;
    }
    S->nent = ((volatile int *)S->tail.h)[0];
    S->nitems = ((volatile int *)S->tail.h)[1];
    CHK(S->entries.ensure(S->nent));
    CHK(S->items.ensure(S->nitems));
    CHK(S->u64.ensure(2));
    HIPCHK(hipMemsetAsync(S->u64.p, 0, 2 * sizeof(unsigned long long), st));
    launch_compact(S->scratch.p, S->cap, S->ncand.p, S->counts.p, S->eoff.p, S->ioff.p, S->nsc, per_item, 0,
                   S->entries.p, S->items.p, S->u64.p, st);
    unsigned long long np = 0;
    HIPCHK(hipMemcpyAsync(&np, S->u64.p, sizeof(np), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    S->npairs = (long long)np;

    // Per-call buffers.
    CHK(S->xyzq.ensure(S->n_pad));
    CHK(S->styp.ensure(S->n_pad));
    CHK(S->facc.ensure(6 * (size_t)S->n_pad));
    CHK(S->eacc.ensure(NB_EACC_WORDS));
    CHK(S->vacc.ensure(NB_VACC_WORDS));
    CHK(S->energies.ensure(4));
    CHK(order_items(S));
    CHK(S->heads.ensure(8 * 32));

    // Plain path: full CSR rows (both directions).
    if (o.precision == NBABFS_PLAIN_FP64) {
        CHK(S->rowptr.ensure(n + 1));
        CHK(S->cols.ensure(std::max<size_t>(2 * (size_t)S->npairs, 1)));
        CHK(S->gplain.ensure(3 * (size_t)n));
        CHK(S->eplain.ensure(2 * (size_t)nblk(n, 256)));
        launch_fill_int(S->cursor.p, n + 1, 0, st);
        launch_csr_degree(S->entries.p, S->nent, S->slot_atom.p, S->cursor.p, st);
        launch_scan(S->cursor.p, S->rowptr.p, n + 1, S->scantmp.p, st);
        HIPCHK(hipMemcpyAsync(S->cursor.p, S->rowptr.p, (n + 1) * sizeof(int), hipMemcpyDeviceToDevice, st));
        launch_csr_fill(S->entries.p, S->nent, S->slot_atom.p, S->cursor.p, S->cols.p, st);
        launch_csr_sort_rows(S->rowptr.p, n, S->cols.p, st);
        S->csrValid = true;
    }
    HIPCHK(hipGetLastError());
    S->listValid = true;
    return NBABFS_STATUS_CONTINUE;
}

static void fill_desc(nbabfs_state *S, SysDesc &D, const double *d_xyz, double *d_grad);

// Upload the state's descriptor when allocations or the coordinate/gradient
// arrays changed.  The staging copy is reused only after the stream drained.
static int ensure_desc(nbabfs_state *S, const double *d_xyz, double *d_grad)
{
    CHK(S->d_desc.ensure(1));
    CHK(S->h_desc.ensure(1));
    if (S->desc_version == g_alloc_version && S->last_xyz == d_xyz && S->last_grad == d_grad)
        return NBABFS_STATUS_CONTINUE;
    HIPCHK(hipStreamSynchronize(S->stream));
    fill_desc(S, S->h_desc.p[0], d_xyz, d_grad);
    HIPCHK(hipMemcpyAsync(S->d_desc.p, S->h_desc.p, sizeof(SysDesc), hipMemcpyHostToDevice, S->stream));
    S->desc_version = g_alloc_version;
    S->last_xyz = d_xyz;
    S->last_grad = d_grad;
    return NBABFS_STATUS_CONTINUE;
}

static int ensure_call_buffers(nbabfs_state *S)
{
    CHK(S->energies.ensure(4));
    CHK(S->eacc.ensure(NB_EACC_WORDS));
    CHK(S->vacc.ensure(NB_VACC_WORDS));
    CHK(S->epart14.ensure(2 * (size_t)std::max(nblk(S->n14, 256), 1)));
    return NBABFS_STATUS_CONTINUE;
}

static int launch_force(nbabfs_state *S, ForceGuard g);

// NBModelABFS_Update (NBModelABFS.c:508-623) for device coordinates.  When a
// list exists, the displacement test runs inside the prepare pass, and the
// pair kernel is queued right behind it, guarded by the test's device flag,
// before the host waits for the test: the GPU does not idle while the host
// decides.  If the list must be rebuilt, the guarded launch does nothing and
// the energy call prepares and evaluates again on the new list.
static int update(nbabfs_state *S, const double *d_xyz, int *rebuilt)
{
    hipStream_t st = S->stream;
    bool doUpdate = S->isNew || !S->have_ref || !S->listValid;
    S->forced_xyz = nullptr;
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
        launch_prepare(S->d_desc.p, 1, S->n_pad, plain ? 0 : 1, S->heads.p, chk, st);
        HIPCHK(hipGetLastError());
        // The descriptor's coordinates must be these for the speculative pass.
        if (S->last_xyz == d_xyz && S->n > 0) {
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
            if (S->timing && S->ev_used >= 2) S->ev_used -= 2;
            S->forced_xyz = nullptr;
        }
    }
    if (doUpdate) {
        CHK(build_list(S, d_xyz));
        CHK(S->xyz_ref.ensure(3 * (size_t)std::max(S->n, 1)));
        if (S->n) HIPCHK(hipMemcpyAsync(S->xyz_ref.p, d_xyz, 3 * (size_t)S->n * sizeof(double), hipMemcpyDeviceToDevice, st));
        S->have_ref = true;
        S->updates++;
        S->sumPairs += (double)S->npairs;
    }
    S->isNew = false;
    if (rebuilt) *rebuilt = doUpdate ? 1 : 0;
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
    D.p14 = S->p14.p;
    D.facc = S->facc.p;
    D.gplain = S->gplain.p;
    D.eacc = S->eacc.p;
    D.vacc = S->vacc.p;
    D.epart14 = S->epart14.p;
    D.eplain = S->eplain.p;
    D.energies = S->energies.p;
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
    D.n14_blocks = nblk(S->n14, 256);
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
    if (S->timing) CHK(record_event(S));
    if (S->opt.precision != NBABFS_PLAIN_FP64) {
        if (S->nitems > 0) {
            QueueParams q{S->opt.wavesPerCU > 0 ? S->heads.p : nullptr, S->qstarts.p};
            launch_force_tiled(S->d_desc.p, S->items.p, q, S->F, n_persistent_blocks(S->opt, S->nitems), S->ntypes,
                               S->opt.entriesPerItem, g, st);
        }
    } else {
        launch_force_plain(S->d_desc.p, 0, S->n, S->F, S->rowptr.p, S->cols.p, g, st);
    }
    if (S->timing) CHK(record_event(S));
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
        launch_fill_int((int *)S->energies.p, 8, 0, st);
        HIPCHK(hipGetLastError());
        return NBABFS_STATUS_CONTINUE;
    }
    if (S->forced_xyz != d_xyz) {
        PrepareCheck none{nullptr, nullptr, nullptr, 0.0, nullptr, nullptr, 0, nullptr, nullptr};
        launch_prepare(S->d_desc.p, 1, S->n_pad, plain ? 0 : 1, S->heads.p, none, st);
        CHK(launch_force(S, ForceGuard{nullptr, 0}));
    }
    S->forced_xyz = nullptr;
    launch_pairs14(S->d_desc.p, 1, nblk(S->n14, 256), S->F, st);
    launch_finalize(S->d_desc.p, 1, S->n, plain ? 1 : 0, st);
    HIPCHK(hipGetLastError());
    return NBABFS_STATUS_CONTINUE;
}

// Re-cut the compacted entries of the current list into work items of
// per_item entries (the entries and their order are unchanged).
static int recut_items(nbabfs_state *S, int per_item)
{
    hipStream_t st = S->stream;
    if (S->nsc <= 0) return NBABFS_STATUS_CONTINUE;
    launch_items_per_sc(S->counts.p, S->nsc, per_item, S->nits.p, st);
    launch_scan(S->nits.p, S->ioff.p, S->nsc, S->scantmp.p, st);
    hipLaunchKernelGGL(k_list_totals, dim3(1), dim3(64), 0, st, S->eoff.p, S->counts.p, S->ioff.p, S->nits.p, S->nsc,
                       S->flag.p, S->tail.d);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(st));
    S->nitems = ((volatile int *)S->tail.h)[1];
    CHK(S->items.ensure(std::max(S->nitems, 1)));
    launch_cut_items(S->counts.p, S->eoff.p, S->ioff.p, S->nsc, per_item, 0, S->items.p, st);
    HIPCHK(hipGetLastError());
    return order_items(S);
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
    launch_prepare(S->d_desc.p, 1, S->n_pad, 1, S->heads.p, none, st);
    hipEvent_t ev[2 * reps];
    for (int k = 0; k < 2 * reps; k++) HIPCHK(hipEventCreate(&ev[k]));
    const nbabfs_options keep = S->opt;
    const bool was_timing = S->timing;
    S->timing = false;
    double best = 1e300;
    int bi = -1;
    int rc = NBABFS_STATUS_CONTINUE;
    for (size_t c = 0; c < sizeof(cand) / sizeof(cand[0]) && rc == NBABFS_STATUS_CONTINUE; c++) {
        S->opt.entriesPerItem = cand[c][0];
        S->opt.wavesPerCU = cand[c][1];
        rc = recut_items(S, cand[c][0]);
        // Each launch is timed alone; persistent queues are reset before it.
        for (int r = -1; r < reps && rc == NBABFS_STATUS_CONTINUE; r++) {
            if (S->opt.wavesPerCU > 0) launch_fill_int(S->heads.p, 8 * 32, 0, st);
            if (r >= 0 && hipEventRecord(ev[2 * r], st) != hipSuccess) rc = fail(NBABFS_STATUS_LOGIC_ERROR, "event record failed");
            if (rc == NBABFS_STATUS_CONTINUE) rc = launch_force(S, ForceGuard{nullptr, 0});
            if (rc == NBABFS_STATUS_CONTINUE && r >= 0 && hipEventRecord(ev[2 * r + 1], st) != hipSuccess)
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
        if (rc == NBABFS_STATUS_CONTINUE && tot / reps < best) {
            best = tot / reps;
            bi = (int)c;
        }
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
    if (o.wavesPerCU < 0) o.wavesPerCU = 0;
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
    S->p14_keys = pk;
    S->n14 = (int)pk.size();
    std::vector<int> p14(2 * pk.size());
    for (size_t k = 0; k < pk.size(); k++) {
        p14[2 * k] = (int)(pk[k] / natoms);
        p14[2 * k + 1] = (int)(pk[k] % natoms);
    }
    if (fixedAtoms) {
        std::vector<unsigned char> fr(natoms);
        bool any = false;
        for (int i = 0; i < natoms; i++) {
            fr[i] = fixedAtoms[i] ? 0 : 1;
            any = any || fixedAtoms[i];
        }
        if (any) {
            if ((rc = S->freef.upload(fr.data(), fr.size(), st)) != NBABFS_STATUS_CONTINUE) return bail(rc);
            S->has_fixed = true;
        }
    }
    std::vector<double> q(charges, charges + natoms);
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
        (rc = S->p14.upload(p14.data(), p14.size(), st)) != NBABFS_STATUS_CONTINUE)
        return bail(rc);
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
    CHK(S->h_energies.ensure(4));
    CHK(S->energies.ensure(4));
    HIPCHK(hipMemcpyAsync(S->h_energies.p, S->energies.p, 4 * sizeof(double), hipMemcpyDeviceToHost, S->stream));
    HIPCHK(hipStreamSynchronize(S->stream));
    std::memcpy(energies, S->h_energies.p, 4 * sizeof(double));
    if (present) {
        present[0] = S->npairs > 0;
        present[1] = S->n14 > 0;
    }
    return NBABFS_STATUS_CONTINUE;
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
        launch_fill_int((int *)S->grad_stage.p, 2 * m, 0, st);
        dg = S->grad_stage.p;
    }
    CHK(mmmm_async(S, S->xyz_stage.p, dg));
    if (dg) {
        CHK(S->h_grad.ensure(m));
        HIPCHK(hipMemcpyAsync(S->h_grad.p, dg, m * sizeof(double), hipMemcpyDeviceToHost, st));
    }
    CHK(nbabfs_energies(S, energies, present));
    if (dg)
        for (size_t k = 0; k < m; k++) grad_inout[k] += S->h_grad.p[k];
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
    return NBABFS_STATUS_CONTINUE;
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
    if (!keys || S->npairs == 0) return S->npairs;
    hipStream_t st = S->stream;
    DBuf<long long> dk;
    if (dk.ensure(std::max<long long>(S->npairs, 1)) != NBABFS_STATUS_CONTINUE) return -NBABFS_STATUS_OUT_OF_MEMORY;
    if (S->u64.ensure(2) != NBABFS_STATUS_CONTINUE) return -NBABFS_STATUS_OUT_OF_MEMORY;
    if (hipMemsetAsync(S->u64.p, 0, sizeof(unsigned long long), st) != hipSuccess) return -NBABFS_STATUS_LOGIC_ERROR;
    launch_export_pairs(S->entries.p, S->nent, S->slot_atom.p, S->n, S->u64.p, S->npairs, dk.p, st);
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
        HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    }
    return NBABFS_STATUS_CONTINUE;
}

extern "C" int nbabfs_virial(nbabfs_state *S, double *virial)
{
    if (!S || !virial) return fail(NBABFS_STATUS_INVALID_ARGUMENT, "null argument");
    for (int k = 0; k < 9; k++) virial[k] = 0.0;
    if (!S->vacc.p || S->n == 0) return NBABFS_STATUS_CONTINUE;
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
    MappedInts flags;
    int epoch = 0;
    DBuf<CheckDesc> d_check;
    PinBuf<CheckDesc> h_check;
    PinBuf<double> h_energies;
    DBuf<double> energies;
    int nitems = 0, max_npad = 0, max_n = 0, max_n14b = 0;
    long layout_version = -1;
    std::vector<const double *> last_xyz;
    std::vector<double *> last_grad;
    hipGraphExec_t gexec = nullptr;
    bool graph_valid = false;
    bool mixed = true;
    ~nbabfs_batch()
    {
        if (gexec) (void)hipGraphExecDestroy(gexec);
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
        B->st.push_back(states[k]);
    }
    B->mixed = states[0]->opt.precision == NBABFS_MIXED;
    for (auto *s : B->st)
        if ((s->opt.precision == NBABFS_MIXED) != B->mixed) {
            delete B;
            return fail(NBABFS_STATUS_INVALID_ARGUMENT, "all states of a batch must share one precision");
        }
    B->use_graph = use_graph;
    if (hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking) != hipSuccess) {
        delete B;
        return fail(NBABFS_STATUS_LOGIC_ERROR, "hipStreamCreate failed");
    }
    B->last_xyz.assign(nsystems, nullptr);
    B->last_grad.assign(nsystems, nullptr);
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
        total += s->nitems;
        B->max_npad = std::max(B->max_npad, s->n_pad);
        B->max_n = std::max(B->max_n, s->n);
        B->max_n14b = std::max(B->max_n14b, nblk(s->n14, 256));
    }
    B->nitems = total;
    CHK(B->items.ensure(std::max(total, 1)));
    CHK(B->d_desc.ensure(ns));
    CHK(B->h_desc.ensure(ns));
    CHK(B->energies.ensure(4 * (size_t)ns));
    CHK(B->heads.ensure(8 * 32));
    int first = 0;
    for (int k = 0; k < ns; k++) {
        nbabfs_state *s = B->st[k];
        CHK(ensure_call_buffers(s));
        SysDesc &D = B->h_desc.p[k];
        fill_desc(s, D, d_xyz[k], d_grad ? d_grad[k] : nullptr);
        D.energies = B->energies.p + 4 * (size_t)k;
        if (s->nitems > 0)
            hipLaunchKernelGGL(k_remap_items, dim3(nblk(s->nitems, 256)), dim3(256), 0, st, s->items.p, s->nitems, k,
                               B->items.p + first);
        first += s->nitems;
        B->last_xyz[k] = d_xyz[k];
        B->last_grad[k] = d_grad ? d_grad[k] : nullptr;
    }
    HIPCHK(hipMemcpyAsync(B->d_desc.p, B->h_desc.p, ns * sizeof(SysDesc), hipMemcpyHostToDevice, st));
    std::vector<int> qs(9);
    for (int q = 0; q <= 8; q++) qs[q] = (int)((long long)total * q / 8);
    CHK(B->qstarts.upload(qs.data(), 9, st));
    HIPCHK(hipStreamSynchronize(st));
    B->layout_version = g_alloc_version;
    if (B->gexec) {
        (void)hipGraphExecDestroy(B->gexec);
        B->gexec = nullptr;
    }
    B->graph_valid = false;
    return NBABFS_STATUS_CONTINUE;
}

static int batch_enqueue(nbabfs_batch *B)
{
    hipStream_t st = B->stream;
    const int ns = (int)B->st.size();
    nbabfs_state *s0 = B->st[0];
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
                               maxcount, ForceGuard{nullptr, 0}, st);
        }
    } else {
        for (int k = 0; k < ns; k++)
            launch_force_plain(B->d_desc.p, k, B->st[k]->n, B->st[k]->F, B->st[k]->rowptr.p, B->st[k]->cols.p,
                               ForceGuard{nullptr, 0}, st);
    }
    launch_pairs14(B->d_desc.p, ns, B->max_n14b, s0->F, st);
    launch_finalize(B->d_desc.p, ns, B->max_n, B->mixed ? 0 : 1, st);
    HIPCHK(hipMemcpyAsync(B->h_energies.p, B->energies.p, 4 * (size_t)ns * sizeof(double), hipMemcpyDeviceToHost, st));
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
    // Rebuild decisions for all systems with one launch and one synchronisation.
    CHK(B->d_check.ensure(ns));
    CHK(B->h_check.ensure(ns));
    CHK(B->flags.ensure(ns));
    CHK(B->h_energies.ensure(4 * (size_t)ns));
    int maxn = 0;
    for (int k = 0; k < ns; k++) {
        nbabfs_state *s = B->st[k];
        const double buf = 0.5 * (s->opt.listCutoff - s->opt.outerCutoff);
        B->h_check.p[k] = CheckDesc{d_xyz[k], (s->isNew || !s->have_ref) ? nullptr : s->xyz_ref.p,
                                    s->has_fixed ? s->freef.p : nullptr, s->n, buf * buf};
        maxn = std::max(maxn, s->n);
    }
    HIPCHK(hipMemcpyAsync(B->d_check.p, B->h_check.p, ns * sizeof(CheckDesc), hipMemcpyHostToDevice, st));
    B->epoch = next_epoch(B->epoch);
    if (maxn > 0)
        hipLaunchKernelGGL(k_check_batch, dim3(nblk(maxn, 256), ns), dim3(256), 0, st, B->d_check.p, B->flags.d, B->epoch);
    HIPCHK(hipStreamSynchronize(st));
    bool any = false;
    for (int k = 0; k < ns; k++) {
        nbabfs_state *s = B->st[k];
        s->forced_xyz = nullptr;
        const bool need = s->isNew || !s->have_ref || !s->listValid || ((volatile int *)B->flags.h)[k] == B->epoch;
        s->calls++;
        if (need) {
            HIPCHK(hipStreamSynchronize(s->stream));
            CHK(build_list(s, d_xyz[k]));
            CHK(s->xyz_ref.ensure(3 * (size_t)std::max(s->n, 1)));
            if (s->n)
                HIPCHK(hipMemcpyAsync(s->xyz_ref.p, d_xyz[k], 3 * (size_t)s->n * sizeof(double), hipMemcpyDeviceToDevice,
                                      s->stream));
            HIPCHK(hipStreamSynchronize(s->stream));
            s->have_ref = true;
            s->updates++;
            s->sumPairs += (double)s->npairs;
            any = true;
        }
        s->isNew = false;
        if (rebuilt) rebuilt[k] = need ? 1 : 0;
    }
    bool relayout = any || B->layout_version != g_alloc_version;
    for (int k = 0; k < ns && !relayout; k++)
        relayout = B->last_xyz[k] != d_xyz[k] || B->last_grad[k] != (d_grad ? d_grad[k] : nullptr);
    if (relayout) CHK(batch_layout(B, d_xyz, d_grad));

    if (B->use_graph) {
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
        }
        HIPCHK(hipGraphLaunch(B->gexec, st));
    } else {
        CHK(batch_enqueue(B));
    }
    HIPCHK(hipStreamSynchronize(st));
    std::memcpy(energies, B->h_energies.p, 4 * (size_t)ns * sizeof(double));
    return NBABFS_STATUS_CONTINUE;
}
