// Host dispatch of point synthesis and analysis onto the GPU.
//
// charm_shs_point / charm_sha_point keep CHarm's host-pointer contract: the
// inputs are copied to HBM, the transform runs there, and the results are
// copied back.  The charm_amd_*_dev entry points take device pointers and
// leave everything resident.  Each host thread owns one non-blocking HIP
// stream and its own FFT plans, so concurrent calls from several threads do
// not share mutable state; the recurrence-coefficient table is shared and
// only ever grows.
#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <mutex>
#include <tuple>
#include <vector>

#include "internal.h"

namespace charm_amd {

namespace {

const double PI = 3.14159265358979323846;

// ---------------------------------------------------------------------------
// Errors
// ---------------------------------------------------------------------------
struct HipFail {
    hipError_t code;
    const char *what;
    const char *file;
    int line;
};

#define HIPCHK(expr)                                                  \
    do {                                                              \
        hipError_t e_ = (expr);                                       \
        if (e_ != hipSuccess) throw HipFail{e_, #expr, __FILE__, __LINE__}; \
    } while (0)

#define FFTCHK(expr)                                                        \
    do {                                                                    \
        hipfftResult r_ = (expr);                                           \
        if (r_ != HIPFFT_SUCCESS)                                           \
            throw HipFail{(hipError_t)(1000 + (int)r_), #expr, __FILE__, __LINE__}; \
    } while (0)

void report(charm_err *err, const HipFail &f, const char *func)
{
    char msg[CHARM_ERR_MAX_MSG];
    int code = CHARM_EFUNCARG;
    if ((int)f.code >= 1000) {
        std::snprintf(msg, sizeof msg, "hipFFT failure %d in \"%s\".", (int)f.code - 1000, f.what);
        code = CHARM_EFFTWINIT;
    } else {
        std::snprintf(msg, sizeof msg, "HIP error \"%s\" in \"%s\".", hipGetErrorString(f.code),
                      f.what);
        if (f.code == hipErrorOutOfMemory || f.code == hipErrorMemoryAllocation) code = CHARM_EMEM;
    }
    err_set(err, f.file, (size_t)f.line, func, code, msg);
}

// ---------------------------------------------------------------------------
// Per-thread stream, device memory and FFT plans
// ---------------------------------------------------------------------------
std::once_flag g_pool_once;

void init_pool()
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return;
    hipMemPool_t pool;
    if (hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess) {
        uint64_t thr = UINT64_MAX;
        (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &thr);
    }
}

struct ThreadState {
    hipStream_t stream = nullptr;
    std::map<std::tuple<int, size_t, size_t, size_t, hipStream_t>, hipfftHandle> plans;
    void *pinned = nullptr;
    size_t pinned_size = 0;
    hipEvent_t stage_ev[2] = {nullptr, nullptr};  // last copy out of each staging half
    ~ThreadState()
    {
        for (auto &kv : plans) hipfftDestroy(kv.second);
        for (hipEvent_t e : stage_ev)
            if (e) (void)hipEventDestroy(e);
        if (pinned) (void)hipHostFree(pinned);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

ThreadState &tls()
{
    thread_local ThreadState st;
    return st;
}

hipStream_t thread_stream()
{
    ThreadState &st = tls();
    if (!st.stream) {
        std::call_once(g_pool_once, init_pool);
        HIPCHK(hipStreamCreateWithFlags(&st.stream, hipStreamNonBlocking));
    }
    return st.stream;
}

void *pinned_buffer(size_t bytes)
{
    ThreadState &st = tls();
    if (st.pinned_size < bytes) {
        if (st.pinned) HIPCHK(hipHostFree(st.pinned));
        st.pinned = nullptr;
        st.pinned_size = 0;
        HIPCHK(hipHostMalloc(&st.pinned, bytes, hipHostMallocDefault));
        st.pinned_size = bytes;
    }
    return st.pinned;
}

// stream-ordered device buffer
// Per-call scratch comes from the device's default stream-ordered pool.  Its
// release threshold is raised once per device so that freed scratch stays
// in the pool across calls: with the default threshold of 0 every
// synchronisation (hipFFT's, a caller's) hands the memory back to the driver
// and the next call maps it again, which at nmax = 10800 (several GB of
// spectra per call) costs more than the transforms.
void keep_pool_memory()
{
    static std::mutex mu;
    static std::vector<char> done;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return;
    std::lock_guard<std::mutex> lk(mu);
    if ((size_t)dev >= done.size()) done.resize(dev + 1, 0);
    if (done[dev]) return;
    done[dev] = 1;
    hipMemPool_t pool;
    if (hipDeviceGetDefaultMemPool(&pool, dev) != hipSuccess) return;
    uint64_t keep = UINT64_MAX;
    (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
}

// Scratch released by a call stays cached for later calls, by stream and
// size (the stream-ordered allocator slows down badly with thousands of
// allocations queued behind a deep stream).  A cached buffer carries an event
// recorded after its last use; the next user waits for it on its own stream,
// so reuse is ordered even across streams (or a recycled stream handle).
// Cached bytes are capped; the oldest buffers go first.  No buffer is made
// under stream capture: it would become a graph memory-allocation node, and
// the library records no such nodes (the device-resident entry points refuse
// a capturing stream before they enqueue anything; graphs come from the
// batch API, whose buffers are allocated when the batch is created).
struct CachedBuf {
    void *p;
    size_t bytes;
    hipEvent_t done;
    uint64_t stamp;
};
std::mutex g_sc_mu;
std::map<hipStream_t, std::multimap<size_t, CachedBuf>> g_sc;
size_t g_sc_bytes = 0;
uint64_t g_sc_clock = 0;
constexpr size_t SCRATCH_CACHE_CAP = size_t(32) << 30;

bool capturing(hipStream_t st)
{
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    return hipStreamIsCapturing(st, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone;
}

// Entry points whose work needs per-call scratch: false (and err set) when
// the stream is capturing, before anything is enqueued.
bool capture_refused(hipStream_t st, charm_err *err, const char *func)
{
    if (!capturing(st)) return false;
    err_set(err, __FILE__, (size_t)__LINE__, func, CHARM_EFUNCARG,
            "This device entry point allocates per-call scratch and cannot be recorded under "
            "stream capture: use charm_amd_batch_* (graphs recorded by the library) or "
            "charm_amd_shs_sctr_dev_noalloc.");
    return true;
}

// with g_sc_mu held
void scratch_evict_oldest()
{
    auto best_s = g_sc.end();
    std::multimap<size_t, CachedBuf>::iterator best;
    for (auto si = g_sc.begin(); si != g_sc.end(); ++si)
        for (auto it = si->second.begin(); it != si->second.end(); ++it)
            if (best_s == g_sc.end() || it->second.stamp < best->second.stamp) {
                best_s = si;
                best = it;
            }
    if (best_s == g_sc.end()) return;
    CachedBuf b = best->second;
    best_s->second.erase(best);
    g_sc_bytes -= b.bytes;
    (void)hipEventSynchronize(b.done);
    (void)hipEventDestroy(b.done);
    (void)hipFree(b.p);
}

struct DevBuf {
    void *p = nullptr;
    hipStream_t s = nullptr;
    size_t cap = 0;
    hipEvent_t done = nullptr;  // non-null: the buffer goes back to the cache
    DevBuf() = default;
    DevBuf(size_t bytes, hipStream_t st) { alloc(bytes, st); }
    void alloc(size_t bytes, hipStream_t st)
    {
        s = st;
        if (!bytes) return;
        if (capturing(st))
            throw HipFail{hipErrorStreamCaptureUnsupported, "scratch allocation under stream capture",
                          __FILE__, __LINE__};
        {
            std::lock_guard<std::mutex> lk(g_sc_mu);
            auto sit = g_sc.find(st);
            if (sit != g_sc.end()) {
                auto it = sit->second.lower_bound(bytes);
                if (it != sit->second.end() && it->first <= bytes + bytes / 4) {
                    const CachedBuf b = it->second;
                    sit->second.erase(it);
                    g_sc_bytes -= b.bytes;
                    p = b.p;
                    cap = b.bytes;
                    done = b.done;
                }
            }
        }
        if (p) {
            HIPCHK(hipStreamWaitEvent(st, done, 0));
            return;
        }
        keep_pool_memory();
        cap = bytes;
        HIPCHK(hipMallocAsync(&p, bytes, st));
        if (hipEventCreateWithFlags(&done, hipEventDisableTiming) != hipSuccess) done = nullptr;
    }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf()
    {
        if (!p) return;
        if (!done || hipEventRecord(done, s) != hipSuccess) {
            if (done) (void)hipEventDestroy(done);
            (void)hipFreeAsync(p, s);
            return;
        }
        std::lock_guard<std::mutex> lk(g_sc_mu);
        g_sc[s].emplace(cap, CachedBuf{p, cap, done, ++g_sc_clock});
        g_sc_bytes += cap;
        while (g_sc_bytes > SCRATCH_CACHE_CAP) scratch_evict_oldest();
    }
    template <class T>
    T *as() const
    {
        return (T *)p;
    }
};

// ---------------------------------------------------------------------------
// Optional per-kernel timing: HIP events recorded on the launching stream
// around each launch; resolved when read.
// ---------------------------------------------------------------------------
struct ProfRec {
    const char *name;
    hipEvent_t a, b;
};
std::mutex g_prof_mu;
int g_prof_on = 0;  // 0 off, 1 every named kernel, 2 the recurrence kernels only
std::vector<ProfRec> g_prof_pending;
std::map<std::string, std::pair<double, long>> g_prof_acc;

struct ProfScope {
    ProfRec rec{nullptr, nullptr, nullptr};
    hipStream_t s;
    ProfScope(const char *name, hipStream_t st) : s(st)
    {
        int on;
        {
            std::lock_guard<std::mutex> lk(g_prof_mu);
            on = g_prof_on;
        }
        if (on == 0) return;
        if (on == 2 && std::strcmp(name, "shs_grid_kernel") != 0 && std::strcmp(name, "sha_kernel") != 0 &&
            std::strcmp(name, "sctr_kernel") != 0)
            return;
        rec.name = name;
        HIPCHK(hipEventCreate(&rec.a));
        HIPCHK(hipEventCreate(&rec.b));
        HIPCHK(hipEventRecord(rec.a, s));
    }
    ~ProfScope()
    {
        if (!rec.name) return;
        if (hipEventRecord(rec.b, s) != hipSuccess) return;
        std::lock_guard<std::mutex> lk(g_prof_mu);
        g_prof_pending.push_back(rec);
    }
};

// kind 0: c2r (Z2D), 1: r2c (D2Z); n = transform length; batch rows.  Plans are
// kept per host thread and stream: a plan's work area must not be shared by
// transforms running concurrently on different streams.
hipfftHandle fft_plan(int kind, size_t n, size_t batch, size_t ldc, hipStream_t stream)
{
    ThreadState &st = tls();
    auto key = std::make_tuple(kind, n, batch, ldc, stream);
    auto it = st.plans.find(key);
    hipfftHandle h;
    if (it == st.plans.end()) {
        FFTCHK(hipfftCreate(&h));
        int nn = (int)n;
        int cemb = (int)ldc, remb = (int)n;
        if (kind == 0)
            FFTCHK(hipfftPlanMany(&h, 1, &nn, &cemb, 1, (int)ldc, &remb, 1, (int)n, HIPFFT_Z2D,
                                  (int)batch));
        else
            FFTCHK(hipfftPlanMany(&h, 1, &nn, &remb, 1, (int)n, &cemb, 1, (int)ldc, HIPFFT_D2Z,
                                  (int)batch));
        st.plans[key] = h;
    } else {
        h = it->second;
    }
    FFTCHK(hipfftSetStream(h, stream));
    return h;
}

// Row transforms of nrow rows of n reals <-> n/2+1 complex values: the LDS
// kernels (one pass up to 2 x 9400 points, two passes for longer rows with a
// short factor), hipFFT otherwise.
void rows_c2r(size_t n, size_t nrow, const void *spec, size_t ldc, double *f, hipStream_t s)
{
    if (row_fft_supported(n)) {
        DevBuf scratch(row_fft_scratch(n, nrow), s);
        HIPCHK((hipError_t)launch_row_fft(1, n, nrow, f, n, const_cast<void *>(spec), ldc,
                                          scratch.p, s));
    } else {
        hipfftHandle h = fft_plan(0, n, nrow, ldc, s);
        FFTCHK(hipfftExecZ2D(h, (hipfftDoubleComplex *)spec, f));
    }
}

void rows_r2c(size_t n, size_t nrow, const double *f, void *spec, size_t ldc, hipStream_t s)
{
    if (row_fft_supported(n)) {
        DevBuf scratch(row_fft_scratch(n, nrow), s);
        HIPCHK((hipError_t)launch_row_fft(0, n, nrow, const_cast<double *>(f), n, spec, ldc,
                                          scratch.p, s));
    } else {
        hipfftHandle h = fft_plan(1, n, nrow, ldc, s);
        FFTCHK(hipfftExecD2Z(h, (hipfftDoubleReal *)f, (hipfftDoubleComplex *)spec));
    }
}

// plans (and their tables) made before a stream capture
void rows_prepare(int kind, size_t n, size_t nrow, size_t ldc, hipStream_t s)
{
    if (row_fft_supported(n))
        HIPCHK((hipError_t)prepare_row_fft(n, s));
    else
        (void)fft_plan(kind, n, nrow, ldc, s);
}

// ---------------------------------------------------------------------------
// Shared tables: (a_nm, b_nm) for all m <= n <= nt, and dm[n], n <= nt
// (src/leg/leg_func_r_ri.c, leg_func_dm.c, leg_func_anm_bnm.c)
// ---------------------------------------------------------------------------
struct Tables {
    double *al = nullptr;  // alpha_nm of the scaled recurrence, m-major, degree nt
    double *sn = nullptr;  // s_nm, same layout
    double *dm = nullptr;  // dm[n] of the sectorial seeds
    int nt = -1;
};
std::mutex g_tab_mu;
Tables g_tab;
std::vector<void *> g_retired;  // earlier (smaller) tables stay valid for in-flight work

Tables tables(int nmax, hipStream_t stream)
{
    std::lock_guard<std::mutex> lk(g_tab_mu);
    if (g_tab.nt >= nmax) return g_tab;
    int nt = std::max(nmax, std::max(64, g_tab.nt + g_tab.nt / 4));
    // r[k] = sqrt(k), ri[k] = 1/sqrt(k) (src/leg/leg_func_r_ri.c:15-30) and
    // dm[n] (src/leg/leg_func_dm.c:14-23)
    std::vector<double> r, ri, dm;
    leg_host_tables(nt, r, ri, dm);
    void *ab = nullptr;
    double *al = nullptr, *sn = nullptr, *ddm = nullptr, *dr = nullptr, *dri = nullptr;
    HIPCHK(hipMalloc(&ab, tri_len(nt) * 2 * sizeof(double)));
    HIPCHK(hipMalloc((void **)&al, tri_len(nt) * sizeof(double)));
    HIPCHK(hipMalloc((void **)&sn, tri_len(nt) * sizeof(double)));
    HIPCHK(hipMalloc((void **)&ddm, dm.size() * sizeof(double)));
    HIPCHK(hipMalloc((void **)&dr, r.size() * sizeof(double)));
    HIPCHK(hipMalloc((void **)&dri, ri.size() * sizeof(double)));
    HIPCHK(hipMemcpyAsync(ddm, dm.data(), dm.size() * sizeof(double), hipMemcpyHostToDevice, stream));
    HIPCHK(hipMemcpyAsync(dr, r.data(), r.size() * sizeof(double), hipMemcpyHostToDevice, stream));
    HIPCHK(hipMemcpyAsync(dri, ri.data(), ri.size() * sizeof(double), hipMemcpyHostToDevice, stream));
    HIPCHK((hipError_t)launch_leg_tables(nt, dr, dri, ab, al, sn, stream));
    HIPCHK(hipStreamSynchronize(stream));
    HIPCHK(hipFree(dr));
    HIPCHK(hipFree(dri));
    HIPCHK(hipFree(ab));
    if (g_tab.al) {
        g_retired.push_back(g_tab.al);
        g_retired.push_back(g_tab.sn);
        g_retired.push_back(g_tab.dm);
    }
    g_tab.al = al;
    g_tab.sn = sn;
    g_tab.dm = ddm;
    g_tab.nt = nt;
    return g_tab;
}

template <class T>
void h2d(T *dst, const T *src, size_t n, hipStream_t s)
{
    if (n) HIPCHK(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyHostToDevice, s));
}

// Grid geometry shared by the synthesis and analysis paths
struct GridGeom {
    size_t nlat = 0, nlon = 0, nring = 0;
    // latitude count the mirror rows are taken from: ring i mirrors to row
    // mir_nlat - 1 - i (nlat, or nlat + 1 on Driscoll--Healy grids, which
    // lack the south pole)
    size_t mir_nlat = 0;
    bool symm = false;
    std::vector<double> t, u;             // [nring]
    std::vector<unsigned char> mirror;    // [nring]
};

// Symmetry of a grid's latitudes (src/shs/shs_point_grd.c:133-197,
// src/crd/crd_grd_check_symm.c:21-72, crd_point_quad_equator.c:22-53,
// src/misc/misc_arr_chck_symm.c).  Driscoll--Healy grids count one more
// latitude, the missing south pole (src/shs/shs_point_grd.c:116-130,
// src/sha/sha_point.c:190-192): rings 0 (north pole, no mirror) .. L
// (equator), ring 0 < i < L mirrored to row 2L - i.
GridGeom grid_geometry(const charm_point *pnt)
{
    GridGeom g;
    g.nlat = pnt->nlat;
    g.nlon = pnt->nlon;
    g.mir_nlat = g.nlat;
    const size_t nlat = g.nlat;
    if (pnt->type == CHARM_CRD_POINT_GRID_DH1 || pnt->type == CHARM_CRD_POINT_GRID_DH2) {
        const size_t L = nlat / 2;
        g.symm = true;
        g.mir_nlat = nlat + 1;
        g.nring = L + 1;
        g.t.resize(g.nring);
        g.u.resize(g.nring);
        g.mirror.resize(g.nring);
        for (size_t i = 0; i < g.nring; i++) {
            g.t[i] = std::sin(pnt->lat[i]);
            g.u[i] = std::cos(pnt->lat[i]);
            g.mirror[i] = (i > 0 && i < L) ? 1 : 0;
        }
        return g;
    }
    const bool even = !(nlat % 2);
    const bool gl = pnt->type == CHARM_CRD_POINT_GRID_GL;
    if (nlat == 1) {
        g.symm = false;
    } else if (gl) {
        g.symm = true;
    } else {
        bool ok = true;
        for (size_t i = 0; i < nlat / 2 && ok; i++)
            ok = nearly_equal(pnt->lat[i] - 0.0, 0.0 - pnt->lat[nlat - 1 - i], charm_glob_threshold2);
        if (ok && (nlat % 2)) ok = nearly_equal(0.0, pnt->lat[nlat / 2], charm_glob_threshold2);
        g.symm = ok;
    }
    g.nring = g.symm ? (nlat + 1 - (even ? 1 : 0)) / 2 : nlat;
    const size_t nmax_grd = nlat - 1;
    const size_t equator = (gl && (nmax_grd % 2)) ? SIZE_MAX : nlat / 2;
    g.t.resize(g.nring);
    g.u.resize(g.nring);
    g.mirror.resize(g.nring);
    for (size_t i = 0; i < g.nring; i++) {
        g.t[i] = std::sin(pnt->lat[i]);
        g.u[i] = std::cos(pnt->lat[i]);
        unsigned char mr = 0;
        if (g.symm) {
            if (gl)
                mr = (i < equator) ? 1 : 0;
            else
                mr = (!even && i == g.nring - 1) ? 0 : 1;
        }
        g.mirror[i] = mr;
    }
    return g;
}

// FFT along parallels possible?  (src/shs/shs_grd_point_fft_check.c:14-40)
bool fft_applicable(const charm_point *pnt, unsigned long nmax)
{
    const size_t nlon = pnt->nlon;
    if (nlon == 0) return false;
    const double dl = (nlon > 1) ? pnt->lon[1] - pnt->lon[0] : 0.0;
    return ((nlon - 1) / 2 >= nmax) && nearly_equal(pnt->lon[0], 0.0, charm_glob_threshold) &&
           nearly_equal(pnt->lon[nlon - 1] + dl, 2.0 * PI, charm_glob_threshold);
}

int ring_tiles(size_t nring, int K)
{
    return (int)((nring + 64 * (size_t)K - 1) / (64 * (size_t)K));
}

// Waves whose column is all zero.  A wave (order m, ring tile) whose rings
// stay below 2^-1100 in every term of the column contributes exactly zero to
// synthesis and analysis; which waves those are depends only on the rings,
// the degree and the rings per lane K.  For a device-resident grid the first
// transform at (K, nmax) records one flag per wave (kernels with dead_mode 1)
// and later transforms skip the flagged waves (dead_mode 2), once an event
// shows that the recording launch has completed.  Tables are only used with
// the polar optimisation off and, for synthesis, r = R.
struct DeadFlags {
    struct Entry {
        unsigned char *p = nullptr;
        hipEvent_t done = nullptr;
        bool recorded = false, ready = false;
    };
    std::mutex mu;
    std::map<std::pair<int, unsigned long>, Entry> tab;

    ~DeadFlags()
    {
        for (auto &kv : tab) {
            if (kv.second.p) (void)hipFree(kv.second.p);
            if (kv.second.done) (void)hipEventDestroy(kv.second.done);
        }
    }

    // Flags for a launch with K rings per lane and ntile ring tiles; mode is
    // the kernels' dead_mode (0 when no table applies).
    unsigned char *acquire(int K, unsigned long nmax, int ntile, hipStream_t s, int &mode)
    {
        mode = 0;
        hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
        const bool capturing =
            hipStreamIsCapturing(s, &cap) != hipSuccess || cap != hipStreamCaptureStatusNone;
        std::lock_guard<std::mutex> lk(mu);
        auto it = tab.find({K, nmax});
        if (it == tab.end()) {
            if (capturing) return nullptr;  // no allocation or recording in a graph
            Entry e;
            const size_t n = (size_t)(nmax + 1) * (size_t)ntile;
            if (hipMalloc((void **)&e.p, n) != hipSuccess) {
                (void)hipGetLastError();
                return nullptr;
            }
            if (hipMemsetAsync(e.p, 0, n, s) != hipSuccess ||
                hipEventCreateWithFlags(&e.done, hipEventDisableTiming) != hipSuccess) {
                (void)hipGetLastError();
                (void)hipStreamSynchronize(s);
                (void)hipFree(e.p);
                if (e.done) (void)hipEventDestroy(e.done);
                return nullptr;
            }
            tab[{K, nmax}] = e;
            mode = 1;
            return e.p;
        }
        Entry &e = it->second;
        if (!e.ready && e.recorded && !capturing && hipEventQuery(e.done) == hipSuccess)
            e.ready = true;
        if (!e.ready) return nullptr;
        mode = 2;
        return e.p;
    }

    // after the recording launch, on its stream
    void recorded(int K, unsigned long nmax, hipStream_t s)
    {
        std::lock_guard<std::mutex> lk(mu);
        auto it = tab.find({K, nmax});
        if (it != tab.end() && !it->second.recorded) {
            (void)hipEventRecord(it->second.done, s);
            it->second.recorded = true;
        }
    }
};

// Lumped sums + longitude transform for the rings of a grid whose geometry
// is already on the device.
struct DevGrid {
    const double *t, *u;
    const unsigned char *mirror;
    const double *sx;
    const int *six;
    size_t sld;
    const double *ratio;   // nullptr when all radii equal R
    const double *rp;
    size_t rld;
    DeadFlags *dead = nullptr;  // device-resident grids only
};

// ---------------------------------------------------------------------------
// Ring pairs per lane K of the grid kernels.
//
// By default a rule fitted on MI355X to the kernel times of every K on
// Gauss-Legendre grids of degree 64 to 10800 (`python -m charm_amd.tune`,
// which reruns that sweep; analysis timed with its slot reduction): few
// rings leave the chip short of waves, so small grids want few rings per
// lane; large grids amortise the per-degree work of a wave (column staging,
// analysis's cross-lane reduction) over more rings.  In ring pairs n:
// synthesis K = 1 up to 450, 2 up to 900, 3 beyond (K = 4, 3 waves per
// SIMD, is 0-2.5% faster from nmax 3500 on and a tuner candidate only);
// analysis K = 2 up to 351, 3 up to 551, 5 up to 951, 6 beyond (K <= 4 run
// 4 waves per SIMD, K = 5 and 6 three, K = 8 two).  Above 1000 ring pairs
// K = 5, 6 and 8 are within
// a few per cent of each other, the order set by how the rings fill the last
// tile; K = 6 was fastest or within 4% at 13 of 16 sampled degrees (at nmax
// 2160: 0.540 ms against 0.573 (K = 5) and 0.580 (K = 8)), 8.5% behind K = 5
// at nmax 2500.
//
// With CHARM_AMD_TUNE=1 an autotuner refines the rule at run time: the first
// large transform of a kind in a ring-count bucket (a quarter octave) times
// every candidate with HIP events on the caller's stream (one warm-up launch
// each, then the fastest of three launches in interleaved rounds) and keeps
// the fastest for the bucket.  It never runs under stream capture or on a
// thread that has switched tuning off (the serving worker), and
// CHARM_AMD_SHS_K / CHARM_AMD_SHA_K force a value.
// ---------------------------------------------------------------------------
std::mutex g_tune_mu;
std::map<std::pair<int, int>, int> g_tuned;
std::atomic<int> g_last_k[2];
thread_local bool t_no_tune = false;

int fitted_k(int kind, size_t nring)
{
    if (kind == 0) return nring <= 450 ? 1 : (nring <= 900 ? 2 : 3);
    if (nring <= 351) return 2;
    if (nring <= 551) return 3;
    return nring <= 951 ? 5 : 6;
}

bool tuning_on()
{
    const char *v = std::getenv("CHARM_AMD_TUNE");
    return v && v[0] == '1' && !t_no_tune;
}

// Ring pairs per lane that each grid kernel is built for
constexpr int SHS_KS[] = {1, 2, 3, 4};
constexpr int SHA_KS[] = {2, 3, 4, 5, 6, 8};

bool valid_k(int kind, int k)
{
    for (int v : (kind == 0 ? std::vector<int>(std::begin(SHS_KS), std::end(SHS_KS))
                            : std::vector<int>(std::begin(SHA_KS), std::end(SHA_KS))))
        if (v == k) return true;
    return false;
}

template <class Launch>
int tuned_k(int kind, int nring, unsigned long nmax, hipStream_t s, Launch launch)
{
    const char *env = kind == 0 ? "CHARM_AMD_SHS_K" : "CHARM_AMD_SHA_K";
    if (const char *v = std::getenv(env)) {
        const int k = std::atoi(v);
        if (valid_k(kind, k)) return k;
    }
    const int rule = fitted_k(kind, (size_t)nring);
    const double work = (double)nring * (double)(nmax + 1) * (double)(nmax + 1);
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    if (!tuning_on() || work < 2e8 || hipStreamIsCapturing(s, &cap) != hipSuccess ||
        cap != hipStreamCaptureStatusNone)
        return rule;
    const int bucket = (int)std::lround(4.0 * std::log2((double)std::max(nring, 1)));
    const auto key = std::make_pair(kind, bucket);
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tuned.find(key);
        if (it != g_tuned.end()) return it->second;
    }
    const std::vector<int> ks = kind == 0 ? std::vector<int>(std::begin(SHS_KS), std::end(SHS_KS))
                                          : std::vector<int>(std::begin(SHA_KS), std::end(SHA_KS));
    int best = rule;
    float best_ms = 0.0f;
    hipEvent_t a, b;
    HIPCHK(hipEventCreate(&a));
    HIPCHK(hipEventCreate(&b));
    std::vector<float> kms(ks.size(), 0.0f);
    for (int k : ks) launch(k);  // warm-up (code load, caches)
    for (int round = 0; round < 3; round++) {
        for (size_t i = 0; i < ks.size(); i++) {
            HIPCHK(hipEventRecord(a, s));
            launch(ks[i]);
            HIPCHK(hipEventRecord(b, s));
            HIPCHK(hipEventSynchronize(b));
            float ms = 0.0f;
            HIPCHK(hipEventElapsedTime(&ms, a, b));
            if (round == 0 || ms < kms[i]) kms[i] = ms;
        }
    }
    for (size_t i = 0; i < ks.size(); i++) {
        const float ms = kms[i];
        if (std::getenv("CHARM_AMD_TUNE_LOG"))
            std::fprintf(stderr, "charm_amd tune: kind %d nring %d nmax %lu K=%d %.3f ms\n", kind,
                         nring, nmax, ks[i], (double)ms);
        if (i == 0 || ms < best_ms) {
            best_ms = ms;
            best = ks[i];
        }
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (std::getenv("CHARM_AMD_TUNE_LOG"))
        std::fprintf(stderr, "charm_amd tune: kind %d ring bucket %d -> K=%d (%.3f ms; rule %d)\n",
                     kind, bucket, best, (double)best_ms, rule);
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tuned[key] = best;
    return best;
}

// Polar optimisation on a grid: the first order skipped per ring, for CHarm's
// ring blocks of synthesis (analysis = 0) or analysis (analysis = 1), on the
// device; p stays null when the optimisation is off.  The rings are CHarm's
// "do" rings in its loop order (src/shs/shs_point_grd.c:594-600,860-862,
// src/sha/sha_point.c:716-718).
// The per-ring cos(lat) u_dev are the grid's (glibc cos on the host), and the
// device applies CHarm's arithmetic exactly, so the result equals the host's
// polar_ring_m0.
struct PolarRings {
    DevBuf buf, umax;
    const int *p = nullptr;
    PolarRings(const GridGeom &g, const double *u_dev, unsigned long nmax, int analysis,
               hipStream_t s)
    {
        const double thr = polar_threshold(nmax);
        if (thr < 0.0 || g.nring == 0) return;
        const size_t bs = polar_block(analysis);
        buf.alloc(g.nring * sizeof(int), s);
        umax.alloc(((g.nring + bs - 1) / bs) * sizeof(double), s);
        HIPCHK((hipError_t)launch_polar_rings_ws(g.nring, u_dev, bs, (int)nmax, thr,
                                                 umax.as<double>(), buf.as<int>(), s));
        p = buf.as<int>();
    }
};

// Lumped sums of a grid synthesis: the Hermitian half-spectra of the rows
// (fft = true, ld = nlon/2 + 1 values per row, ready for c2r) or the raw
// (a_m, b_m) (fft = false, ld = nmax + 1 per row)
void shs_grid_lumped(const GridGeom &g, const DevGrid &d, bool fft, unsigned long nmax, double mur,
                     const void *cs_dev, const Tables &tb, void *out, size_t ld,
                     hipStream_t stream)
{
    ShsGridArgs a{};
    a.nmax = (int)nmax;
    a.nring = (int)g.nring;
    a.nlat = (int)g.mir_nlat;
    a.unit = d.ratio == nullptr;
    a.t = d.t;
    a.sx = d.sx;
    a.six = d.six;
    a.sld = d.sld;
    a.rp = d.rp;
    a.ratio = d.ratio;
    a.rld = d.rld;
    a.mirror = d.mirror;
    a.al = tb.al;
    a.al_nt = tb.nt;
    PolarRings pol(g, d.u, nmax, 0, stream);
    a.pm0 = pol.p;
    // coefficients times s_nm for the scaled recurrence, as the kernel stages them
    a.cs = cs_dev;
    a.sn = tb.sn;
    a.mur = mur;
    a.out = out;
    a.ldo = ld;
    a.mcount = (int)ld;
    a.fft_layout = fft ? 1 : 0;
    a.k = tuned_k(0, (int)g.nring, nmax, stream, [&](int k) {
        ShsGridArgs y = a;
        y.k = k;
        y.ntile = ring_tiles(g.nring, k);
        HIPCHK((hipError_t)launch_shs_grid(y, stream));
    });
    a.ntile = ring_tiles(g.nring, a.k);
    g_last_k[0].store(a.k);
    if (d.dead && a.unit && a.pm0 == nullptr)
        a.dead = d.dead->acquire(a.k, nmax, a.ntile, stream, a.dead_mode);
    {
        ProfScope ps("shs_grid_kernel", stream);
        HIPCHK((hipError_t)launch_shs_grid(a, stream));
    }
    if (a.dead_mode == 1) d.dead->recorded(a.k, nmax, stream);
}

void run_shs_grid(const GridGeom &g, const DevGrid &d, const double *lon_dev, bool use_fft,
                  unsigned long nmax, double mur, const void *cs_dev, const Tables &tb,
                  double *f_dev, hipStream_t stream)
{
    const size_t nlat = g.nlat, nlon = g.nlon;
    if (use_fft) {
        const size_t nfc = nlon / 2 + 1;
        DevBuf spec(nlat * nfc * 2 * sizeof(double), stream);
        shs_grid_lumped(g, d, true, nmax, mur, cs_dev, tb, spec.p, nfc, stream);
        ProfScope ps("fft_c2r", stream);
        rows_c2r(nlon, nlat, spec.p, nfc, f_dev, stream);
    } else {
        const size_t ldl = nmax + 1;
        DevBuf lc(nlat * ldl * 2 * sizeof(double), stream);
        shs_grid_lumped(g, d, false, nmax, mur, cs_dev, tb, lc.p, ldl, stream);
        HIPCHK((hipError_t)launch_lon_direct((int)nlat, (int)nmax, nlon, lon_dev, lc.p, ldl, mur,
                                             f_dev, stream));
    }
}

// ---------------------------------------------------------------------------
// Plain per-call synthesis on a grid (host arrays)
// ---------------------------------------------------------------------------
// Custom grids: longitudes must be linearly increasing
// (src/shs/shs_point_grd.c:205-236)
bool check_grid_lons(const charm_point *pnt, charm_err *err)
{
    const size_t nlon = pnt->nlon;
    if (pnt->type != CHARM_CRD_POINT_GRID || nlon < 2) return true;
    const double d0 = pnt->lon[1] - pnt->lon[0];
    for (size_t j = 1; j < nlon; j++)
        if (!nearly_equal(pnt->lon[j] - pnt->lon[j - 1], d0, charm_glob_threshold2)) {
            CHARM_AMD_ERR(err, CHARM_EFUNCARG,
                          "\"pnt->lon\" is not a linearly increasing array within the "
                          "\"threshold2\".");
            return false;
        }
    return true;
}

// Device copies of a grid's rings, sectorial seeds, radius powers and, for
// the direct longitude sums, longitudes, for syntheses up to degree nmax.
struct GridSetup {
    GridGeom g;
    bool use_fft = false, unit = true;
    DevBuf dt, du, dmir, sx, six, ratio, rp, lon;
    DevGrid d{};
};

void grid_setup(const charm_point *pnt, const charm_shc *shcs, unsigned long nmax,
                const Tables &tb, hipStream_t s, GridSetup &G)
{
    const size_t nlat = pnt->nlat, nlon = pnt->nlon;
    G.g = grid_geometry(pnt);
    G.use_fft = fft_applicable(pnt, nmax);
    G.unit = true;
    for (size_t i = 0; i < nlat && G.unit; i++)
        G.unit = nearly_equal(pnt->r[i], shcs->r, charm_glob_threshold);
    const size_t nr = G.g.nring;
    G.dt.alloc(nr * sizeof(double), s);
    G.du.alloc(nr * sizeof(double), s);
    G.dmir.alloc(nr, s);
    h2d(G.dt.as<double>(), G.g.t.data(), nr, s);
    h2d(G.du.as<double>(), G.g.u.data(), nr, s);
    h2d(G.dmir.as<unsigned char>(), G.g.mirror.data(), nr, s);
    const size_t sld = nr;
    G.sx.alloc(std::max<size_t>(nmax, 1) * sld * sizeof(double), s);
    G.six.alloc(std::max<size_t>(nmax, 1) * sld * sizeof(int), s);
    HIPCHK((hipError_t)launch_seeds((int)nr, G.du.as<double>(), tb.dm, (int)nmax, sld,
                                    G.sx.as<double>(), G.six.as<int>(), s));
    G.d = DevGrid{G.dt.as<double>(), G.du.as<double>(), G.dmir.as<unsigned char>(),
                  G.sx.as<double>(), G.six.as<int>(), sld, nullptr, nullptr, 0};
    if (!G.unit) {
        std::vector<double> q(nlat);
        for (size_t i = 0; i < nlat; i++) q[i] = shcs->r / pnt->r[i];
        G.ratio.alloc(nlat * sizeof(double), s);
        G.rp.alloc((nmax + 2) * nlat * sizeof(double), s);
        h2d(G.ratio.as<double>(), q.data(), nlat, s);
        HIPCHK((hipError_t)launch_rpows((int)nlat, G.ratio.as<double>(), (int)nmax + 1, nlat,
                                        G.rp.as<double>(), s));
        G.d.ratio = G.ratio.as<double>();
        G.d.rp = G.rp.as<double>();
        G.d.rld = nlat;
    }
    if (!G.use_fft) {
        G.lon.alloc(nlon * sizeof(double), s);
        h2d(G.lon.as<double>(), pnt->lon, nlon, s);
    }
}

// Coefficients of degree nmax as (C, S) pairs on the device, packed into two
// pinned staging halves of at most 2^20 pairs (16 MiB) each and copied half
// by half; a half is refilled once its previous copy has completed.  A large
// model so pins at most 32 MiB of host memory (pinning the whole triangle,
// 933 MB at nmax 10800, made the first call of a process at that degree take
// twice as long as the next ones).
void upload_coeffs(const charm_shc *shcs, unsigned long nmax, hipStream_t s, DevBuf &cs)
{
    constexpr size_t STAGE_PAIRS = size_t(1) << 20;
    const size_t P = tri_len(nmax);
    cs.alloc(P * 2 * sizeof(double), s);
    const size_t half = std::min(P, STAGE_PAIRS);
    double *h = (double *)pinned_buffer(2 * half * 2 * sizeof(double));
    ThreadState &st = tls();
    for (hipEvent_t &e : st.stage_ev)
        if (!e) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    unsigned long m = 0, n = 0;  // the next pair to pack
    bool used[2] = {false, false};
    int which = 0;
    for (size_t q0 = 0; q0 < P; q0 += half, which ^= 1) {
        const size_t cnt = std::min(half, P - q0);
        double *buf = h + (size_t)which * half * 2;
        if (used[which]) HIPCHK(hipEventSynchronize(st.stage_ev[which]));
        for (size_t i = 0; i < cnt; i++) {
            buf[2 * i] = shcs->c[m][n - m];
            buf[2 * i + 1] = (m == 0) ? 0.0 : shcs->s[m][n - m];
            if (++n > nmax) n = ++m;
        }
        h2d(cs.as<double>() + 2 * q0, buf, 2 * cnt, s);
        HIPCHK(hipEventRecord(st.stage_ev[which], s));
        used[which] = true;
    }
}

void shs_grid_host(const charm_point *pnt, const charm_shc *shcs, unsigned long nmax, double *f,
                   charm_err *err)
{
    if (!check_grid_lons(pnt, err)) return;
    const size_t nlat = pnt->nlat, nlon = pnt->nlon;
    hipStream_t s = thread_stream();
    Tables tb = tables((int)std::max<unsigned long>(nmax, 1), s);
    DevBuf cs;
    upload_coeffs(shcs, nmax, s, cs);
    GridSetup G;
    grid_setup(pnt, shcs, nmax, tb, s, G);
    DevBuf df(nlat * nlon * sizeof(double), s);
    run_shs_grid(G.g, G.d, G.lon.as<double>(), G.use_fft, nmax, shcs->mu / shcs->r, cs.p, tb,
                 df.as<double>(), s);
    HIPCHK(hipMemcpyAsync(f, df.p, nlat * nlon * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
}

// ---------------------------------------------------------------------------
// Scattered points
// ---------------------------------------------------------------------------
// Points per lane of a scattered synthesis: 4 for large point sets (1e6
// points at nmax 2160: 0.436 s against 0.459 s with 2), 2 otherwise, where
// 4 would leave too few waves to fill the device.
int sctr_k(size_t npoint)
{
    if (const char *v = std::getenv("CHARM_AMD_SCTR_K")) {
        const int k = std::atoi(v);
        if (k == 1 || k == 2 || k == 4) return k;
    }
    return npoint >= (size_t(1) << 18) ? 4 : 2;
}

// Order chunks of equal work (boundaries m, nchunk + 1 entries): more chunks
// when there are few points, so that the grid still fills the device
// The chunks depend on the point count and degree only (not on the points
// per lane), so batched and standalone calls sum in the same order.  Aim:
// about 16384 waves of 128 points (two chunks at 1e6 points: 1.5% faster
// than one, the waves of a long column balance better).
std::vector<int> sctr_chunks(size_t npoint, unsigned long nmax)
{
    int buf[SCTR_MAX_CHUNKS + 1];
    const int nchunk = sctr_order_chunks(npoint, nmax, buf);
    return std::vector<int>(buf, buf + nchunk + 1);
}

// Scattered synthesis on device arrays: lat/lon/r of npoint points, unscaled
// coefficients cs (interleaved pairs, degree nmax).  Points are processed in
// the order of |sin lat| (results go back to the caller's order).
// unit_dev non-null: the radii test was made on the device (*unit_dev) and
// overrides unit
void run_sctr(size_t npoint, const double *lat_dev, const double *lon_dev, const double *r_dev,
              bool unit, unsigned long nmax, double mu, double R, const void *cs_dev,
              const Tables &tb, double *f_dev, hipStream_t s, const int *unit_dev = nullptr)
{
    const int K = sctr_k(npoint);
    DevBuf dt(npoint * sizeof(double), s), du(npoint * sizeof(double), s);
    DevBuf dl, dr, perm;
    const double *lon_use = lon_dev, *r_use = r_dev;
    if (npoint < (size_t)INT32_MAX) {
        perm.alloc(npoint * sizeof(int), s);
        dl.alloc(npoint * sizeof(double), s);
        dr.alloc(npoint * sizeof(double), s);
        HIPCHK((hipError_t)launch_sort_points(npoint, lat_dev, lon_dev, r_dev, perm.as<int>(),
                                              dt.as<double>(), du.as<double>(), dl.as<double>(),
                                              dr.as<double>(), s));
        lon_use = dl.as<double>();
        r_use = dr.as<double>();
    } else {
        HIPCHK((hipError_t)launch_sincos(npoint, lat_dev, dt.as<double>(), du.as<double>(), s));
    }
    // the order chunks go to the device as kernel arguments (no copy from
    // host memory, which a stream under capture would record by address)
    IntVals mcv{};
    const int nchunk = sctr_order_chunks(npoint, nmax, mcv.v);
    mcv.n = nchunk + 1;
    DevBuf dmc((nchunk + 1) * sizeof(int), s);
    HIPCHK((hipError_t)launch_set_ints(dmc.as<int>(), mcv, s));
    DevBuf part((size_t)nchunk * npoint * sizeof(double), s);
    SctrArgs a{};
    a.nmax = (int)nmax;
    a.npoint = npoint;
    a.nchunk = nchunk;
    a.mchunk = dmc.as<int>();
    a.unit = unit;
    a.unit_dev = unit_dev;
    a.R = R;
    a.lat_t = dt.as<double>();
    a.lat_u = du.as<double>();
    a.lon = lon_use;
    a.r = r_use;
    a.dm = tb.dm;
    a.al = tb.al;
    a.al_nt = tb.nt;
    a.cs = cs_dev;
    a.sn = tb.sn;
    a.part = part.as<double>();
    a.k = K;
    const double thr = polar_threshold(nmax);
    DevBuf pm0;
    if (thr >= 0.0) {
        pm0.alloc(npoint * sizeof(int), s);
        HIPCHK((hipError_t)launch_polar_points(npoint, lat_dev, perm.p ? perm.as<int>() : nullptr,
                                               polar_block(0), (int)nmax, thr, pm0.as<int>(), s));
        a.pm0 = pm0.as<int>();
    }
    {
        ProfScope ps("sctr_kernel", s);
        HIPCHK((hipError_t)launch_sctr(a, s));
    }
    HIPCHK((hipError_t)launch_sctr_reduce(npoint, nchunk, part.as<double>(), perm.as<int>(),
                                          mu / R, f_dev, s));
}

void shs_sctr_host(const charm_point *pnt, const charm_shc *shcs, unsigned long nmax, double *f,
                   charm_err *err)
{
    if (pnt->nlat != pnt->nlon) {
        char msg[CHARM_ERR_MAX_MSG];
        std::snprintf(msg, sizeof msg,
                      "The number of latitudes \"%zu\" do not match the number of longitudes "
                      "\"%zu\".",
                      pnt->nlat, pnt->nlon);
        CHARM_AMD_ERR(err, CHARM_EFUNCARG, msg);
        return;
    }
    const size_t npoint = pnt->nlat;
    hipStream_t s = thread_stream();
    bool unit = true;
    for (size_t i = 0; i < npoint && unit; i++)
        unit = nearly_equal(pnt->r[i], shcs->r, charm_glob_threshold);
    Tables tb = tables((int)std::max<unsigned long>(nmax, 1), s);
    DevBuf cs;
    upload_coeffs(shcs, nmax, s, cs);
    DevBuf dlat(npoint * sizeof(double), s), dl(npoint * sizeof(double), s);
    DevBuf dr(npoint * sizeof(double), s);
    h2d(dlat.as<double>(), pnt->lat, npoint, s);
    h2d(dl.as<double>(), pnt->lon, npoint, s);
    h2d(dr.as<double>(), pnt->r, npoint, s);
    DevBuf df(npoint * sizeof(double), s);
    run_sctr(npoint, dlat.as<double>(), dl.as<double>(), dr.as<double>(), unit, nmax, shcs->mu,
             shcs->r, cs.p, tb, df.as<double>(), s);
    HIPCHK(hipMemcpyAsync(f, df.p, npoint * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
}

// ---------------------------------------------------------------------------
// Gradients and single derivatives (src/shs/shs.h.in:176-436)
//
// With g_ijk = 1/(r^(j+k) cos^k lat) d^(i+j+k) f / dr^i dlat^j dlon^k, the
// radial derivatives multiply degree n by rho_1 = -(n+1), rho_2 = (n+1)(n+2)
// and add 1/r^i; the longitude derivatives act on (C, S) per order; the
// latitude derivatives follow from
//   u dP_nm/dlat       = -n t P_nm + e_nm P_{n-1,m}
//   d^2 P_nm / dlat^2  = (t/u) dP_nm/dlat + (m^2/u^2 - n(n+1)) P_nm
// (src/leg/leg_func_xnum.h:35-44), where the P_{n-1,m} term is a synthesis
// over the table e_{n+1,m} X_{n+1,m} carrying one more power of q = R/r.
// Terms of one output with the same weight share one table.
// ---------------------------------------------------------------------------
struct GradPlan {
    std::vector<GradTable> tabs;
    std::vector<std::vector<GradTerm>> outs;
    GradPlan() = default;
    explicit GradPlan(int nout) : outs(nout) {}
    void term(int o, double coef, GradTable t, int pt, int pu, int pr, int pq)
    {
        int k = 0;
        for (; k < (int)tabs.size(); k++) {
            const GradTable &x = tabs[k];
            if (x.c0 == t.c0 && x.c1 == t.c1 && x.c2 == t.c2 && x.shift == t.shift &&
                x.lon == t.lon)
                break;
        }
        if (k == (int)tabs.size()) tabs.push_back(t);
        outs[o].push_back(GradTerm{coef, k, pt, pu, pr, pq, 0});
    }
};

//                                c0  c1  c2 shift lon
constexpr GradTable TAB_ONE    = {1, 0, 0, 0, 0};
constexpr GradTable TAB_RAD1   = {-1, -1, 0, 0, 0};   // -(n+1)
constexpr GradTable TAB_RAD2   = {2, 3, 1, 0, 0};     // (n+1)(n+2)
constexpr GradTable TAB_LON1   = {1, 0, 0, 0, 1};
constexpr GradTable TAB_LON2   = {1, 0, 0, 0, 2};
constexpr GradTable TAB_N      = {0, 1, 0, 0, 0};     // n
constexpr GradTable TAB_NN1    = {0, 1, 1, 0, 0};     // n(n+1)
constexpr GradTable TAB_SH     = {1, 0, 0, 1, 0};
constexpr GradTable TAB_NLON1  = {0, 1, 0, 0, 1};
constexpr GradTable TAB_SHLON1 = {1, 0, 0, 1, 1};
constexpr GradTable TAB_NRAD1  = {0, -1, -1, 0, 0};   // -n(n+1)
constexpr GradTable TAB_SHRAD1 = {-1, -1, 0, 1, 0};
constexpr GradTable TAB_RAD1LON1 = {-1, -1, 0, 0, 1};

// g_{dr dlat dlon} as one output
GradPlan guru_plan(unsigned dr, unsigned dlat, unsigned dlon)
{
    GradPlan p(1);
    const int key = 100 * (int)dr + 10 * (int)dlat + (int)dlon;
    switch (key) {
    case 0: p.term(0, 1, TAB_ONE, 0, 0, 0, 0); break;
    case 100: p.term(0, 1, TAB_RAD1, 0, 0, 1, 0); break;
    case 200: p.term(0, 1, TAB_RAD2, 0, 0, 2, 0); break;
    case 1: p.term(0, 1, TAB_LON1, 0, 1, 1, 0); break;
    case 2: p.term(0, 1, TAB_LON2, 0, 2, 2, 0); break;
    case 10:
        p.term(0, -1, TAB_N, 1, 1, 1, 0);
        p.term(0, 1, TAB_SH, 0, 1, 1, 1);
        break;
    case 20:
        p.term(0, -1, TAB_N, 2, 2, 2, 0);
        p.term(0, 1, TAB_SH, 1, 2, 2, 1);
        p.term(0, -1, TAB_NN1, 0, 0, 2, 0);
        p.term(0, -1, TAB_LON2, 0, 2, 2, 0);
        break;
    case 110:
        p.term(0, -1, TAB_NRAD1, 1, 1, 2, 0);
        p.term(0, 1, TAB_SHRAD1, 0, 1, 2, 1);
        break;
    case 101: p.term(0, 1, TAB_RAD1LON1, 0, 1, 2, 0); break;
    case 11:
        p.term(0, -1, TAB_NLON1, 1, 2, 2, 0);
        p.term(0, 1, TAB_SHLON1, 0, 2, 2, 1);
        break;
    default: break;
    }
    return p;
}

// f_x = g_010, f_y = -g_001, f_z = g_100 (src/shs/shs.h.in:176-197)
GradPlan grad1_plan()
{
    GradPlan p(3);
    p.term(0, -1, TAB_N, 1, 1, 1, 0);
    p.term(0, 1, TAB_SH, 0, 1, 1, 1);
    p.term(1, -1, TAB_LON1, 0, 1, 1, 0);
    p.term(2, 1, TAB_RAD1, 0, 0, 1, 0);
    return p;
}

// Reed (1973) after src/shs/shs.h.in:278-312:
//   f_xx = g_020 + g_100/r          f_xy = -g_011 - (t/u) g_001/r
//   f_xz = g_110 - g_010/r          f_yy = g_002 + g_100/r - (t/u) g_010/r
//   f_yz = -g_101 + g_001/r         f_zz = g_200
GradPlan grad2_plan()
{
    GradPlan p(6);
    // xx: the r^-2 terms -n(n+1) - (n+1) = -(n+1)^2 share a table
    p.term(0, -1, TAB_N, 2, 2, 2, 0);
    p.term(0, 1, TAB_SH, 1, 2, 2, 1);
    p.term(0, 1, GradTable{-1, -2, -1, 0, 0}, 0, 0, 2, 0);
    p.term(0, -1, TAB_LON2, 0, 2, 2, 0);
    // xy: n - 1 times the longitude table
    p.term(1, 1, GradTable{-1, 1, 0, 0, 1}, 1, 2, 2, 0);
    p.term(1, -1, TAB_SHLON1, 0, 2, 2, 1);
    // xz: n(n+1) + n and -(n+1) - 1 on the shifted table
    p.term(2, 1, GradTable{0, 2, 1, 0, 0}, 1, 1, 2, 0);
    p.term(2, 1, GradTable{-2, -1, 0, 1, 0}, 0, 1, 2, 1);
    // yy
    p.term(3, 1, TAB_LON2, 0, 2, 2, 0);
    p.term(3, 1, TAB_RAD1, 0, 0, 2, 0);
    p.term(3, 1, TAB_N, 2, 2, 2, 0);
    p.term(3, -1, TAB_SH, 1, 2, 2, 1);
    // yz: 1 + (n+1) times the longitude table
    p.term(4, 1, GradTable{2, 1, 0, 0, 1}, 0, 1, 2, 0);
    // zz
    p.term(5, 1, TAB_RAD2, 0, 0, 2, 0);
    return p;
}

// Per-row sin and cos of a grid's latitudes, as the synthesis uses them:
// mirrored rows take -t of their ring
void grid_rows_tu(const GridGeom &g, std::vector<double> &t, std::vector<double> &u)
{
    t.assign(g.nlat, 0.0);
    u.assign(g.nlat, 1.0);
    for (size_t i = 0; i < g.nring; i++) {
        if (i < g.nlat) {
            t[i] = g.t[i];
            u[i] = g.u[i];
        }
        if (g.mirror[i]) {
            const size_t j = g.mir_nlat - 1 - i;
            if (j < g.nlat) {
                t[j] = -g.t[i];
                u[j] = g.u[i];
            }
        }
    }
}

// One synthesis per table of the plan on a grid: tabs[k * nlat * nlon + p]
void grad_tables_grid(const GridGeom &g, const DevGrid &d, const double *lon_dev, bool use_fft,
                      unsigned long nmax, double mur, const void *cs_dev, const Tables &tb,
                      const GradPlan &plan, double *tabs, hipStream_t s)
{
    const size_t npoint = g.nlat * g.nlon;
    DevBuf ctab(tri_len(nmax) * 2 * sizeof(double), s);
    for (size_t k = 0; k < plan.tabs.size(); k++) {
        HIPCHK((hipError_t)launch_grad_table((int)nmax, cs_dev, plan.tabs[k], ctab.p, s));
        run_shs_grid(g, d, lon_dev, use_fft, nmax, mur, ctab.p, tb, tabs + k * npoint, s);
    }
}

// Grids through the FFT: every table is c0 + c1 n + c2 n^2 (shifted or not)
// times a longitude map, and the longitude maps are (i m)^lon on the
// spectrum, so the outputs combine the half-spectra of at most five basis
// syntheses (n^j X, j = 0..2, and the shifted e X, e n X) row by row, then
// take one c2r each: grad2 costs five syntheses and six row transforms.
bool grad_grid_fft(const GradPlan &plan, const GridGeom &g, const DevGrid &d, unsigned long nmax,
                   double mur, const void *cs_dev, const Tables &tb, const double *trow,
                   const double *urow, const double *rrow, bool unit, double R, double *out,
                   hipStream_t s)
{
    const size_t nlat = g.nlat, nlon = g.nlon, nfc = nlon / 2 + 1;
    if (nlat > 65535) return false;
    // basis index of (power j, shift)
    int bidx[3][2];
    for (auto &v : bidx) v[0] = v[1] = -1;
    std::vector<GradTable> basis;
    SpecCombine c{};
    c.nout = (int)plan.outs.size();
    int q = 0;
    for (int o = 0; o < c.nout; o++) {
        c.beg[o] = q;
        for (const GradTerm &x : plan.outs[o]) {
            const GradTable &tab = plan.tabs[x.tab];
            const double cj[3] = {tab.c0, tab.c1, tab.c2};
            for (int j = 0; j < 3; j++) {
                if (cj[j] == 0.0) continue;
                int &b = bidx[j][tab.shift];
                if (b < 0) {
                    b = (int)basis.size();
                    basis.push_back(GradTable{j == 0 ? 1.0 : 0.0, j == 1 ? 1.0 : 0.0,
                                              j == 2 ? 1.0 : 0.0, tab.shift, 0});
                }
                if (q >= SPEC_MAXTERM) return false;
                c.term[q++] = SpecTerm{x.coef * cj[j], b, tab.lon, x.pt, x.pu, x.pr, x.pq};
            }
        }
    }
    c.beg[c.nout] = q;
    const size_t per = nlat * nfc * 2 * sizeof(double);
    DevBuf ctab(tri_len(nmax) * 2 * sizeof(double), s), bspec(basis.size() * per, s);
    for (size_t b = 0; b < basis.size(); b++) {
        HIPCHK((hipError_t)launch_grad_table((int)nmax, cs_dev, basis[b], ctab.p, s));
        shs_grid_lumped(g, d, true, nmax, mur, ctab.p, tb, (char *)bspec.p + b * per, nfc, s);
    }
    DevBuf ospec(c.nout * per, s);
    for (int o = 0; o < c.nout; o++) c.out[o] = (char *)ospec.p + o * per;
    {
        ProfScope ps("grad_combine_kernel", s);
        HIPCHK((hipError_t)launch_grad_spec_combine(nlat, nfc, trow, urow, rrow, unit ? 1 : 0, R,
                                                    bspec.p, c, s));
    }
    ProfScope ps("fft_c2r", s);
    for (int o = 0; o < c.nout; o++)
        rows_c2r(nlon, nlat, c.out[o], nfc, out + (size_t)o * nlat * nlon, s);
    return true;
}

// One synthesis per table of the plan at scattered points
void grad_tables_sctr(size_t npoint, const double *lat_dev, const double *lon_dev,
                      const double *r_dev, bool unit, unsigned long nmax, double mu, double R,
                      const void *cs_dev, const Tables &tb, const GradPlan &plan, double *tabs,
                      hipStream_t s)
{
    DevBuf ctab(tri_len(nmax) * 2 * sizeof(double), s);
    for (size_t k = 0; k < plan.tabs.size(); k++) {
        HIPCHK((hipError_t)launch_grad_table((int)nmax, cs_dev, plan.tabs[k], ctab.p, s));
        run_sctr(npoint, lat_dev, lon_dev, r_dev, unit, nmax, mu, R, ctab.p, tb,
                 tabs + k * npoint, s);
    }
}

// Outputs o = 0..nout-1 into out + o * nrow * ncol
void grad_combine(const GradPlan &plan, size_t nrow, size_t ncol, const double *t_dev,
                  const double *u_dev, const double *r_dev, bool unit, double R,
                  const double *tabs, double *out, hipStream_t s)
{
    const int nout = (int)plan.outs.size();
    size_t nterm = 0;
    for (const auto &o : plan.outs) nterm += o.size();
    if (nout > GRAD_MAXOUT || nterm > (size_t)GRAD_MAXTERM)
        throw HipFail{hipErrorInvalidValue, "derivative plan too large", __FILE__, __LINE__};
    const size_t npoint = nrow * ncol;
    GradCombine c{};
    c.nout = nout;
    int q = 0;
    for (int o = 0; o < nout; o++) {
        c.beg[o] = q;
        for (const GradTerm &x : plan.outs[o]) c.term[q++] = x;
        c.out[o] = out + (size_t)o * npoint;
    }
    c.beg[nout] = q;
    ProfScope ps("grad_combine_kernel", s);
    HIPCHK((hipError_t)launch_grad_combine(nrow, ncol, t_dev, u_dev, r_dev, unit ? 1 : 0, R, tabs,
                                           npoint, c, s));
}

void shs_grad_host(const charm_point *pnt, const charm_shc *shcs, unsigned long nmax,
                   const GradPlan &plan, double **f, charm_err *err)
{
    const bool sctr = pnt->type == CHARM_CRD_POINT_SCATTERED;
    if (sctr && pnt->nlat != pnt->nlon) {
        CHARM_AMD_ERR(err, CHARM_EFUNCARG,
                      "The number of latitudes and longitudes in the \"pnt\" structure must be "
                      "the same to perform point-wise spherical harmonic synthesis.");
        return;
    }
    if (!sctr && !check_grid_lons(pnt, err)) return;
    const int nout = (int)plan.outs.size();
    hipStream_t s = thread_stream();
    Tables tb = tables((int)std::max<unsigned long>(nmax, 1), s);
    DevBuf cs;
    upload_coeffs(shcs, nmax, s, cs);
    const size_t ntab = plan.tabs.size();
    const size_t nrow = pnt->nlat, ncol = sctr ? 1 : pnt->nlon, npoint = nrow * ncol;
    // the per-table grids are needed only off the spectral path
    DevBuf tabs, outs(nout * npoint * sizeof(double), s);
    DevBuf dt(nrow * sizeof(double), s), du(nrow * sizeof(double), s);
    DevBuf dr(nrow * sizeof(double), s);
    h2d(dr.as<double>(), pnt->r, nrow, s);
    bool unit = true;
    for (size_t i = 0; i < nrow && unit; i++)
        unit = nearly_equal(pnt->r[i], shcs->r, charm_glob_threshold);
    if (sctr) {
        tabs.alloc(ntab * npoint * sizeof(double), s);
        DevBuf dlat(npoint * sizeof(double), s), dlon(npoint * sizeof(double), s);
        h2d(dlat.as<double>(), pnt->lat, npoint, s);
        h2d(dlon.as<double>(), pnt->lon, npoint, s);
        HIPCHK((hipError_t)launch_sincos(npoint, dlat.as<double>(), dt.as<double>(),
                                         du.as<double>(), s));
        grad_tables_sctr(npoint, dlat.as<double>(), dlon.as<double>(), dr.as<double>(), unit,
                         nmax, shcs->mu, shcs->r, cs.p, tb, plan, tabs.as<double>(), s);
    } else {
        GridSetup G;
        grid_setup(pnt, shcs, nmax, tb, s, G);
        std::vector<double> ht, hu;
        grid_rows_tu(G.g, ht, hu);
        h2d(dt.as<double>(), ht.data(), nrow, s);
        h2d(du.as<double>(), hu.data(), nrow, s);
        if (!G.use_fft ||
            !grad_grid_fft(plan, G.g, G.d, nmax, shcs->mu / shcs->r, cs.p, tb, dt.as<double>(),
                           du.as<double>(), dr.as<double>(), unit, shcs->r, outs.as<double>(),
                           s)) {
            tabs.alloc(ntab * npoint * sizeof(double), s);
            grad_tables_grid(G.g, G.d, G.lon.as<double>(), G.use_fft, nmax, shcs->mu / shcs->r,
                             cs.p, tb, plan, tabs.as<double>(), s);
            grad_combine(plan, nrow, ncol, dt.as<double>(), du.as<double>(), dr.as<double>(),
                         unit, shcs->r, tabs.as<double>(), outs.as<double>(), s);
        }
    }
    if (sctr)
        grad_combine(plan, nrow, ncol, dt.as<double>(), du.as<double>(), dr.as<double>(), unit,
                     shcs->r, tabs.as<double>(), outs.as<double>(), s);
    for (int o = 0; o < nout; o++)
        HIPCHK(hipMemcpyAsync(f[o], outs.as<double>() + (size_t)o * npoint,
                              npoint * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
}

// Plan of a derivative request: CHARM_AMD_DERIV_GRAD1, CHARM_AMD_DERIV_GRAD2
// or 100 dr + 10 dlat + dlon with dr + dlat + dlon <= 2; false if invalid
bool deriv_plan(int what, GradPlan &p)
{
    if (what == CHARM_AMD_DERIV_GRAD1) {
        p = grad1_plan();
        return true;
    }
    if (what == CHARM_AMD_DERIV_GRAD2) {
        p = grad2_plan();
        return true;
    }
    if (what < 0 || what > 200) return false;
    const int dr = what / 100, dlat = (what / 10) % 10, dlon = what % 10;
    if (dlat > 2 || dlon > 2 || dr + dlat + dlon > 2) return false;
    p = guru_plan((unsigned)dr, (unsigned)dlat, (unsigned)dlon);
    return true;
}

// ---------------------------------------------------------------------------
// Analysis
// ---------------------------------------------------------------------------
// c = pi/(nmax_grd+1) * 1/(4 pi) * r0/mu (src/sha/sha_point.c:313-322)
// c = pi / (nmax_grd + 1) (Gauss--Legendre, DH1) or pi / (2 nmax_grd + 2)
// (DH2), times 1/(4 pi) r0/mu (src/sha/sha_point.c:309-322)
double sha_const(int type, size_t nlat, double r0, double mu)
{
    const bool dh = type == CHARM_CRD_POINT_GRID_DH1 || type == CHARM_CRD_POINT_GRID_DH2;
    const double nmax_grd = dh ? (double)(nlat / 2 - 1) : (double)(nlat - 1);
    double c = (type == CHARM_CRD_POINT_GRID_DH2) ? PI / (2.0 * nmax_grd + 2.0) : PI / (nmax_grd + 1.0);
    c *= 1.0 / (4.0 * PI) * (r0 / mu);
    return c;
}

// (r0/R)^n, n = 0..nmax_out, when r0 differs from R (src/shc/shc_rescale.c:30-46)
bool rescale_table(double r0, double R, unsigned long nmax_out, std::vector<double> &sc)
{
    if (nearly_equal(R, r0, charm_glob_threshold)) return false;
    const double mu_ratio = 1.0, r_ratio = r0 / R;
    sc.resize(nmax_out + 1);
    double tmp2 = mu_ratio * r_ratio;
    sc[0] = mu_ratio;
    for (unsigned long n = 1; n <= nmax_out; n++) {
        sc[n] = tmp2;
        tmp2 *= r_ratio;
    }
    return true;
}

// blocks an analysis launch aims at (a few waves of resident blocks)
constexpr int SHA_BLOCKS = 4096;

// Device table of (r0/R)^n, n = 0..nmax_out (the rescale of src/shc/shc_rescale.c:24-90 with
// mu unchanged), or nullptr when r0 equals R; built on the device, so a call under
// stream capture copies nothing from host memory.
struct RescaleTable {
    DevBuf buf;
    const double *p = nullptr;
    RescaleTable(double r0, double R, unsigned long nmax_out, hipStream_t s)
    {
        if (nearly_equal(R, r0, charm_glob_threshold)) return;
        buf.alloc((nmax_out + 1) * sizeof(double), s);
        HIPCHK((hipError_t)launch_powers((int)nmax_out, r0 / R, buf.as<double>(), s));
        p = buf.as<double>();
    }
};

// Partial slots of an analysis: ring tiles are grouped so that the launch has
// about SHA_BLOCKS blocks, each block adding its group's tiles, in order, into
// one slot.  Few slots keep the partial sums small at high degree (one per
// tile would be 20-40 GB at nmax = 10800).  From SHA_BLOCKS / 2 order quads
// on, one slot (every block runs all tiles of its quad): at nmax = 10800 that
// is 5% faster than two slots (half the partial-sum writes, one slot to
// reduce), while at nmax = 2160 fewer, longer blocks are slower (5 slots:
// 0.54 ms, 3: 0.60 ms, 1: 1.14 ms).
int sha_slots(size_t nring, unsigned long nmax)
{
    const int nquad = (int)(nmax + 1 + 3) / 4;
    return nquad >= SHA_BLOCKS / 2
               ? 1
               : std::max(1, std::min(ring_tiles(nring, 2), (SHA_BLOCKS + nquad - 1) / nquad));
}

void run_sha(const GridGeom &g, const DevGrid &d, const double *w_dev, double cws,
             const double *f_dev, unsigned long nmax, const Tables &tb, const double *scale_dev,
             unsigned long nmax_out, double *c_dev, double *s_dev, size_t stride, hipStream_t s)
{
    const size_t nlat = g.nlat, nlon = g.nlon, nfc = nlon / 2 + 1;
    DevBuf spec(nlat * nfc * 2 * sizeof(double), s);
    {
        ProfScope ps("fft_r2c", s);
        rows_r2c(nlon, nlat, f_dev, spec.p, nfc, s);
    }
    ShaArgs a{};
    a.nmax = (int)nmax;
    a.nring = (int)g.nring;
    a.nlat = (int)g.mir_nlat;
    a.nrow = (int)g.nlat;
    a.t = d.t;
    a.sx = d.sx;
    a.six = d.six;
    a.sld = d.sld;
    a.mirror = d.mirror;
    a.cw = w_dev;
    a.cws = cws;
    a.spec = spec.p;
    a.lds = nfc;
    a.al = tb.al;
    a.al_nt = tb.nt;
    PolarRings pol(g, d.u, nmax, 1, s);
    a.pm0 = pol.p;
    const int nslot = sha_slots(g.nring, nmax);
    auto group = [&](ShaArgs &y) {
        y.ntile = ring_tiles(g.nring, y.k);
        y.tpg = (y.ntile + nslot - 1) / nslot;
    };
    // With one slot and the output at the transform's degree the analysis
    // kernel finishes each order quad itself; otherwise sha_reduce_kernel adds
    // the slots (and zero-fills degrees above nmax).
    const bool fuse = nslot == 1 && nmax_out == nmax;
    DevBuf part((size_t)nslot * tri_len(nmax) * 2 * sizeof(double), s);
    a.part = part.p;
    a.k = tuned_k(1, (int)g.nring, nmax, s, [&](int k) {
        ShaArgs y = a;
        y.k = k;
        group(y);
        HIPCHK((hipError_t)launch_sha(y, s));
    });
    group(a);
    g_last_k[1].store(a.k);
    if (fuse) {
        a.sn = tb.sn;
        a.scale = scale_dev;
        a.fc = c_dev;
        a.fs = s_dev;
        a.fstride = stride;
    }
    if (d.dead && a.pm0 == nullptr) a.dead = d.dead->acquire(a.k, nmax, a.ntile, s, a.dead_mode);
    {
        ProfScope ps("sha_kernel", s);
        HIPCHK((hipError_t)launch_sha(a, s));
    }
    if (a.dead_mode == 1) d.dead->recorded(a.k, nmax, s);
    if (!fuse) {
        ProfScope ps("sha_reduce_kernel", s);
        const int nused = (a.ntile + a.tpg - 1) / a.tpg;
        HIPCHK((hipError_t)launch_sha_reduce((int)nmax, nused, part.p, tb.sn, tb.nt,
                                             (int)nmax_out, scale_dev, c_dev, s_dev, stride, s));
    }
}

void sha_host(const charm_point *pnt, const double *f, unsigned long nmax, charm_shc *shcs,
              charm_err *err)
{
    const int type = pnt->type;
    if (type != CHARM_CRD_POINT_GRID_GL && type != CHARM_CRD_POINT_GRID_DH1 &&
        type != CHARM_CRD_POINT_GRID_DH2) {
        CHARM_AMD_ERR(err, CHARM_EFUNCARG,
                      "Unsupported \"pnt->type\" for spherical harmonic analysis of point data "
                      "values.");
        return;
    }
    const size_t nlat = pnt->nlat, nlon = pnt->nlon;
    const double r0 = pnt->r[0];
    for (size_t i = 1; i < nlat; i++)
        if (!nearly_equal(pnt->r[i], r0, charm_glob_threshold)) {
            CHARM_AMD_ERR(err, CHARM_EFUNCARG, "All spherical radii in \"pnt->r\" must be equal.");
            return;
        }
    const unsigned long nmax_grd =
        (type == CHARM_CRD_POINT_GRID_GL) ? (unsigned long)nlat - 1 : ((unsigned long)nlat - 2) / 2;
    if (nmax > nmax_grd) {
        CHARM_AMD_ERR(err, CHARM_EFUNCARG,
                      "The input data grid \"pnt\" was created for a maximum degree that is not "
                      "high enough to recover harmonics up to the specified maximum degree of "
                      "the analysis.");
        return;
    }
    hipStream_t s = thread_stream();
    GridGeom g = grid_geometry(pnt);
    Tables tb = tables((int)std::max<unsigned long>(nmax, 1), s);
    const size_t nr = g.nring;
    const double c = sha_const(type, nlat, r0, shcs->mu);
    DevBuf dt(nr * sizeof(double), s), du(nr * sizeof(double), s), dmir(nr, s);
    DevBuf dw(nlat * sizeof(double), s);
    h2d(dt.as<double>(), g.t.data(), nr, s);
    h2d(du.as<double>(), g.u.data(), nr, s);
    h2d(dmir.as<unsigned char>(), g.mirror.data(), nr, s);
    h2d(dw.as<double>(), pnt->w, nlat, s);
    const size_t sld = nr;
    DevBuf sx(std::max<size_t>(nmax, 1) * sld * sizeof(double), s);
    DevBuf six(std::max<size_t>(nmax, 1) * sld * sizeof(int), s);
    HIPCHK((hipError_t)launch_seeds((int)nr, du.as<double>(), tb.dm, (int)nmax, sld,
                                    sx.as<double>(), six.as<int>(), s));
    DevGrid d{dt.as<double>(), du.as<double>(), dmir.as<unsigned char>(), sx.as<double>(),
              six.as<int>(), sld, nullptr, nullptr, 0};
    DevBuf df(nlat * nlon * sizeof(double), s);
    h2d(df.as<double>(), f, nlat * nlon, s);
    RescaleTable sc(r0, shcs->r, shcs->nmax, s);
    const size_t nc = tri_len(shcs->nmax);
    DevBuf dc(nc * sizeof(double), s), ds(nc * sizeof(double), s);
    run_sha(g, d, dw.as<double>(), c, df.as<double>(), nmax, tb, sc.p, shcs->nmax,
            dc.as<double>(), ds.as<double>(), 1, s);
    HIPCHK(hipMemcpyAsync(shcs->c[0], dc.p, nc * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(shcs->s[0], ds.p, nc * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
}

}  // namespace

void set_thread_tuning(bool on)
{
    t_no_tune = !on;
}

size_t sctr_call_scratch(size_t npoint, unsigned long nmax)
{
    int mc[SCTR_MAX_CHUNKS + 1];
    const int nchunk = sctr_order_chunks(npoint, nmax, mc);
    size_t b = npoint * (4 * sizeof(double) + sizeof(int));   // sorted t, u, lon, r; permutation
    b += sort_points_ws_bytes(npoint);
    b += (size_t)nchunk * npoint * sizeof(double);           // partial sums per order chunk
    b += npoint * (sizeof(int) + sizeof(double));            // polar optimisation
    return b + 4096;                                         // order chunks, radii flag
}

size_t grid_call_scratch(int analysis, size_t nring, size_t nlat, size_t nlon, unsigned long nmax)
{
    size_t b = nlat * (nlon / 2 + 1) * 2 * sizeof(double);   // half-spectra of the rows
    b += row_fft_scratch(nlon, nlat);
    b += nring * (sizeof(int) + sizeof(double));             // polar optimisation
    b += (nmax + 2 + nlat) * sizeof(double);                 // radius powers or rescale table
    if (analysis) b += (size_t)sha_slots(nring, nmax) * tri_len(nmax) * 2 * sizeof(double);
    return b + 4096;
}

size_t scratch_cached_bytes()
{
    std::lock_guard<std::mutex> lk(g_sc_mu);
    return g_sc_bytes;
}

void scratch_trim(size_t keep_bytes)
{
    std::lock_guard<std::mutex> lk(g_sc_mu);
    while (g_sc_bytes > keep_bytes) scratch_evict_oldest();
}

int run_sctr_jobs(const SctrJob *jobs, int njob, void *stream)
{
    hipStream_t s = (hipStream_t)stream;
    try {
        unsigned long nt = 1;
        for (int j = 0; j < njob; j++) nt = std::max(nt, jobs[j].nmax);
        Tables tb = tables((int)nt, s);
        const int K = batch_ring_k(2);
        std::vector<std::unique_ptr<DevBuf>> keep;
        auto buf = [&](size_t bytes) {
            keep.emplace_back(new DevBuf(std::max<size_t>(bytes, 16), s));
            return keep.back()->p;
        };
        std::vector<SctrArgs> args(njob);
        struct Item {
            long cost;
            int2 w;
        };
        std::vector<Item> items;
        std::vector<int> nchunk(njob, 0);
        std::vector<const int *> perm(njob, nullptr);
        for (int j = 0; j < njob; j++) {
            const SctrJob &q = jobs[j];
            const size_t np = q.npoint;
            if (np == 0) continue;
            if (np >= (size_t)INT32_MAX) return (int)hipErrorInvalidValue;
            // as run_sctr does for one call: points by latitude, radii test on
            // the device, the same order chunks (so the sums are the same)
            int *pm = (int *)buf(np * sizeof(int));
            double *t = (double *)buf(np * sizeof(double)), *u = (double *)buf(np * sizeof(double));
            double *l = (double *)buf(np * sizeof(double)), *r = (double *)buf(np * sizeof(double));
            HIPCHK((hipError_t)launch_sort_points(np, q.lat, q.lon, q.r, pm, t, u, l, r, s));
            int *flag = (int *)buf(sizeof(int));
            HIPCHK((hipError_t)launch_all_near_dev(np, q.r, q.R, charm_glob_threshold, flag, s));
            IntVals mcv{};
            nchunk[j] = sctr_order_chunks(np, q.nmax, mcv.v);
            mcv.n = nchunk[j] + 1;
            int *dmc = (int *)buf(mcv.n * sizeof(int));
            HIPCHK((hipError_t)launch_set_ints(dmc, mcv, s));
            SctrArgs &a = args[j];
            a.nmax = (int)q.nmax;
            a.npoint = np;
            a.nchunk = nchunk[j];
            a.mchunk = dmc;
            a.unit_dev = flag;
            a.R = q.R;
            a.lat_t = t;
            a.lat_u = u;
            a.lon = l;
            a.r = r;
            a.dm = tb.dm;
            a.al = tb.al;
            a.al_nt = tb.nt;
            a.cs = q.cs;
            a.sn = tb.sn;
            a.part = (double *)buf((size_t)nchunk[j] * np * sizeof(double));
            a.k = K;
            const double thr = polar_threshold(q.nmax);
            if (thr >= 0.0) {
                int *m0 = (int *)buf(np * sizeof(int));
                double *um = (double *)buf(np * sizeof(double));
                HIPCHK((hipError_t)launch_polar_points_ws(np, q.lat, pm, polar_block(0),
                                                          (int)q.nmax, thr, um, m0, s));
                a.pm0 = m0;
            }
            perm[j] = pm;
            const int nblk = (int)((np + 256 * (size_t)K - 1) / (256 * (size_t)K));
            for (int c = 0; c < nchunk[j]; c++) {
                long work = 0;
                for (int m = mcv.v[c]; m < mcv.v[c + 1]; m++) work += (long)q.nmax + 1 - m;
                for (int bk = 0; bk < nblk; bk++) items.push_back({work, make_int2(j, (bk << 6) | c)});
            }
        }
        if (items.empty()) return 0;
        std::stable_sort(items.begin(), items.end(),
                         [](const Item &x, const Item &y) { return x.cost > y.cost; });
        std::vector<int2> w(items.size());
        for (size_t i = 0; i < items.size(); i++) w[i] = items[i].w;
        void *dargs = buf(args.size() * sizeof(SctrArgs));
        void *ditems = buf(w.size() * sizeof(int2));
        int *counter = (int *)buf(sizeof(int));
        h2d((SctrArgs *)dargs, args.data(), args.size(), s);
        h2d((int2 *)ditems, w.data(), w.size(), s);
        HIPCHK((hipError_t)launch_zero_ints(counter, 1, s));
        {
            ProfScope ps("sctr_kernel", s);
            HIPCHK((hipError_t)launch_sctr_batch(dargs, ditems, (int)w.size(), counter,
                                                 batch_grid_size(2), s));
        }
        for (int j = 0; j < njob; j++)
            if (jobs[j].npoint)
                HIPCHK((hipError_t)launch_sctr_reduce(jobs[j].npoint, nchunk[j], args[j].part,
                                                      perm[j], jobs[j].mu / jobs[j].R, jobs[j].f, s));
        return 0;
    } catch (const HipFail &e) {
        return (int)e.code;
    }
}

// Plan of a scattered synthesis shared with the allocation-free entry point
int sctr_points_per_lane(size_t npoint)
{
    return sctr_k(npoint);
}

int sctr_order_chunks(size_t npoint, unsigned long nmax, int *mc)
{
    const size_t nwave = (npoint + 127) / 128;
    int nchunk = (int)std::min<size_t>(SCTR_MAX_CHUNKS,
                                       std::max<size_t>(1, 16384 / std::max<size_t>(nwave, 1)));
    nchunk = (int)std::min<unsigned long>((unsigned long)nchunk, nmax + 1);
    const double total = 0.5 * (double)(nmax + 1) * (double)(nmax + 2);
    double acc = 0.0;
    int c = 1;
    mc[0] = 0;
    for (unsigned long m = 0; m <= nmax && c < nchunk; m++) {
        acc += (double)(nmax + 1 - m);
        if (acc >= total * c / nchunk) mc[c++] = (int)m + 1;
    }
    while (c < nchunk) mc[c++] = (int)nmax + 1;
    mc[nchunk] = (int)nmax + 1;
    return nchunk;
}

// r[k] = sqrt(k), ri[k] = 1/sqrt(k), k <= 2 nt + 3, and dm[n], n <= nt, as
// tables() computes them
void leg_host_tables(int nt, std::vector<double> &r, std::vector<double> &ri,
                     std::vector<double> &dm)
{
    r.assign(2 * (size_t)nt + 4, 0.0);
    ri.assign(2 * (size_t)nt + 4, 0.0);
    dm.assign((size_t)nt + 1, 0.0);
    for (size_t k = 1; k <= 2 * (size_t)nt + 3; k++) {
        double w = std::sqrt((double)k);
        r[k] = w;
        ri[k] = 1.0 / w;
    }
    for (size_t n = 1; n <= (size_t)nt; n++) dm[n] = r[2 * n + 3] * ri[2 * n + 2];
}

}  // namespace charm_amd

// ---------------------------------------------------------------------------
// Device-resident GL grid
// ---------------------------------------------------------------------------
struct charm_amd_grid {
    charm_amd::GridGeom g;
    int type = 0;
    double r0 = 0.0;
    unsigned long nmax_grd = 0;
    double *t = nullptr, *u = nullptr, *w = nullptr, *lon = nullptr, *sx = nullptr;
    double *trow = nullptr, *urow = nullptr, *rrow = nullptr;  // per row, for derivatives
    int *six = nullptr;
    unsigned char *mirror = nullptr;
    size_t sld = 0;
    mutable charm_amd::DeadFlags dead;  // a cache: logically const
};

using namespace charm_amd;

extern "C" {

void charm_amd_prof_enable(int on)
{
    std::lock_guard<std::mutex> lk(g_prof_mu);
    g_prof_on = (on == 2) ? 2 : (on != 0 ? 1 : 0);
}

void charm_amd_prof_reset(void)
{
    std::lock_guard<std::mutex> lk(g_prof_mu);
    for (auto &r : g_prof_pending) {
        (void)hipEventSynchronize(r.b);
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    g_prof_pending.clear();
    g_prof_acc.clear();
}

int charm_amd_last_ring_k(int kind)
{
    return (kind == 0 || kind == 1) ? g_last_k[kind].load() : 0;
}

int charm_amd_prof_read(const char *name, double *total_ms, long *count)
{
    std::lock_guard<std::mutex> lk(g_prof_mu);
    for (auto &r : g_prof_pending) {
        float ms = 0.0f;
        if (hipEventSynchronize(r.b) == hipSuccess && hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
            auto &acc = g_prof_acc[r.name];
            acc.first += ms;
            acc.second += 1;
        }
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    g_prof_pending.clear();
    auto it = g_prof_acc.find(name);
    if (it == g_prof_acc.end()) {
        *total_ms = 0.0;
        *count = 0;
        return 0;
    }
    *total_ms = it->second.first;
    *count = it->second.second;
    return 1;
}

int charm_amd_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

// src/shs/shs_point.c:26-129
void charm_shs_point(const charm_point *pnt, const charm_shc *shcs, unsigned long nmax,
                     double *f, charm_err *err)
{
    if (nmax > shcs->nmax) {
        char msg[CHARM_ERR_MAX_MSG];
        std::snprintf(msg, sizeof msg,
                      "Maximum harmonic degree of the synthesis \"nmax = %lu\" cannot be larger "
                      "than maximum harmonic degree of spherical harmonic coefficients "
                      "\"shcs->nmax = %lu\".",
                      nmax, shcs->nmax);
        CHARM_AMD_ERR(err, CHARM_EFUNCARG, msg);
        return;
    }
    const int type = pnt->type;
    const bool sctr = type == CHARM_CRD_POINT_SCATTERED;
    const bool grid = type == CHARM_CRD_POINT_GRID || type == CHARM_CRD_POINT_GRID_GL ||
                      type == CHARM_CRD_POINT_GRID_DH1 || type == CHARM_CRD_POINT_GRID_DH2;
    if (!sctr && !grid) {
        CHARM_AMD_ERR(err, CHARM_EFUNCARG,
                      "Unsupported \"pnt->type\" for spherical harmonic synthesis of point "
                      "values.");
        return;
    }
    if (pnt->npoint == 0) return;
    try {
        if (sctr)
            shs_sctr_host(pnt, shcs, nmax, f, err);
        else
            shs_grid_host(pnt, shcs, nmax, f, err);
    } catch (const HipFail &e) {
        report(err, e, __func__);
    }
    if (!charm_err_isempty(err)) err_propagate(err, __FILE__, __LINE__, __func__);
}

// Shared front end of the derivative syntheses: the argument checks of
// src/shs/shs_point_gradn.c:49-83 and src/shs/shs_point_guru.c:30-66
static void shs_deriv_entry(const charm_point *pnt, const charm_shc *shcs, unsigned long nmax,
                            const GradPlan &plan, double **f, charm_err *err)
{
    if (nmax > shcs->nmax) {
        CHARM_AMD_ERR(err, CHARM_EFUNCARG,
                      "Maximum harmonic degree of the synthesis (\"nmax\") cannot be larger than "
                      "maximum harmonic degree of spherical harmonic coefficients "
                      "(\"shcs->nmax\").");
        return;
    }
    const int type = pnt->type;
    const bool sctr = type == CHARM_CRD_POINT_SCATTERED;
    const bool grid = type == CHARM_CRD_POINT_GRID || type == CHARM_CRD_POINT_GRID_GL ||
                      type == CHARM_CRD_POINT_GRID_DH1 || type == CHARM_CRD_POINT_GRID_DH2;
    if (!sctr && !grid) {
        CHARM_AMD_ERR(err, CHARM_EFUNCARG,
                      "Unsupported \"pnt->type\" for spherical harmonic synthesis of point "
                      "values.");
        return;
    }
    if (pnt->npoint == 0) return;
    try {
        shs_grad_host(pnt, shcs, nmax, plan, f, err);
    } catch (const HipFail &e) {
        report(err, e, "shs_deriv_entry");
    }
}

// src/shs/shs_point_gradn.c:31-173 (GRADN = GRAD_1); f[0..2] = x, y, z
void charm_shs_point_grad1(const charm_point *pnt, const charm_shc *shcs, unsigned long nmax,
                           double **f, charm_err *err)
{
    shs_deriv_entry(pnt, shcs, nmax, grad1_plan(), f, err);
    if (!charm_err_isempty(err)) err_propagate(err, __FILE__, __LINE__, __func__);
}

// src/shs/shs_point_gradn.c:31-173 (GRADN = GRAD_2); f[0..5] = xx, xy, xz, yy, yz, zz
void charm_shs_point_grad2(const charm_point *pnt, const charm_shc *shcs, unsigned long nmax,
                           double **f, charm_err *err)
{
    shs_deriv_entry(pnt, shcs, nmax, grad2_plan(), f, err);
    if (!charm_err_isempty(err)) err_propagate(err, __FILE__, __LINE__, __func__);
}

// src/shs/shs_point_guru.c:21-119, derivative checks of
// src/shs/shs_check_single_derivative.c:17-64
void charm_shs_point_guru(const charm_point *pnt, const charm_shc *shcs, unsigned long nmax,
                          unsigned dr, unsigned dlat, unsigned dlon, double *f, charm_err *err)
{
    if (nmax > shcs->nmax) {
        CHARM_AMD_ERR(err, CHARM_EFUNCARG,
                      "Maximum harmonic degree of the synthesis (\"nmax\") cannot be larger than "
                      "maximum harmonic degree of spherical harmonic coefficients "
                      "(\"shcs->nmax\").");
    } else if ((int)dr < 0 || (int)dlat < 0 || (int)dlon < 0) {
        char msg[CHARM_ERR_MAX_MSG];
        const char *which = (int)dr < 0 ? "dr" : ((int)dlat < 0 ? "dlat" : "dlon");
        const int v = (int)dr < 0 ? (int)dr : ((int)dlat < 0 ? (int)dlat : (int)dlon);
        std::snprintf(msg, sizeof msg, "\"%s\" is \"%d\", but it must be non-negative.", which, v);
        CHARM_AMD_ERR(err, CHARM_EFUNCARG, msg);
    } else if ((int)dr + (int)dlat + (int)dlon > 2) {
        char msg[CHARM_ERR_MAX_MSG];
        std::snprintf(msg, sizeof msg,
                      "The sum \"dr + dlat + dlon\" is \"%d\", but it cannot be larger than "
                      "\"%d\".",
                      (int)dr + (int)dlat + (int)dlon, 2);
        CHARM_AMD_ERR(err, CHARM_EFUNCARG, msg);
    } else {
        double *fp[1] = {f};
        shs_deriv_entry(pnt, shcs, nmax, guru_plan(dr, dlat, dlon), fp, err);
    }
    if (!charm_err_isempty(err)) err_propagate(err, __FILE__, __LINE__, __func__);
}

// src/sha/sha_point.c:124-1125
void charm_sha_point(const charm_point *pnt, const double *f, unsigned long nmax,
                     charm_shc *shcs, charm_err *err)
{
    if (nmax > shcs->nmax) {
        CHARM_AMD_ERR(err, CHARM_EFUNCARG,
                      "Maximum harmonic degree of the analysis (\"nmax\") cannot be larger than "
                      "maximum harmonic degree of spherical harmonic coefficients "
                      "(\"shcs->nmax\").");
        return;
    }
    try {
        sha_host(pnt, f, nmax, shcs, err);
    } catch (const HipFail &e) {
        report(err, e, __func__);
    }
    if (!charm_err_isempty(err)) err_propagate(err, __FILE__, __LINE__, __func__);
}

charm_amd_grid *charm_amd_grid_create(const charm_point *pnt, charm_err *err)
{
    if (pnt == nullptr || pnt->type != CHARM_CRD_POINT_GRID_GL || pnt->nlat == 0) {
        CHARM_AMD_ERR(err, CHARM_EFUNCARG, "A Gauss--Legendre grid is required.");
        return nullptr;
    }
    const double r0 = pnt->r[0];
    for (size_t i = 1; i < pnt->nlat; i++)
        if (!nearly_equal(pnt->r[i], r0, charm_glob_threshold)) {
            CHARM_AMD_ERR(err, CHARM_EFUNCARG, "All spherical radii in \"pnt->r\" must be equal.");
            return nullptr;
        }
    charm_amd_grid *grd = new charm_amd_grid();
    try {
        grd->g = grid_geometry(pnt);
        grd->type = pnt->type;
        grd->r0 = r0;
        grd->nmax_grd = pnt->nlat - 1;
        hipStream_t s = thread_stream();
        const size_t nr = grd->g.nring, nlat = pnt->nlat;
        const unsigned long nm = std::max<unsigned long>(grd->nmax_grd, 1);
        Tables tb = tables((int)nm, s);
        HIPCHK(hipMalloc((void **)&grd->t, nr * sizeof(double)));
        HIPCHK(hipMalloc((void **)&grd->u, nr * sizeof(double)));
        HIPCHK(hipMalloc((void **)&grd->mirror, nr));
        HIPCHK(hipMalloc((void **)&grd->w, nlat * sizeof(double)));
        HIPCHK(hipMalloc((void **)&grd->lon, std::max<size_t>(pnt->nlon, 1) * sizeof(double)));
        HIPCHK(hipMalloc((void **)&grd->trow, nlat * sizeof(double)));
        HIPCHK(hipMalloc((void **)&grd->urow, nlat * sizeof(double)));
        HIPCHK(hipMalloc((void **)&grd->rrow, nlat * sizeof(double)));
        std::vector<double> ht, hu;
        grid_rows_tu(grd->g, ht, hu);
        const std::vector<double> hr(nlat, r0);
        HIPCHK(hipMemcpy(grd->trow, ht.data(), nlat * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(grd->urow, hu.data(), nlat * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(grd->rrow, hr.data(), nlat * sizeof(double), hipMemcpyHostToDevice));
        grd->sld = nr;
        HIPCHK(hipMalloc((void **)&grd->sx, nm * nr * sizeof(double)));
        HIPCHK(hipMalloc((void **)&grd->six, nm * nr * sizeof(int)));
        h2d(grd->t, grd->g.t.data(), nr, s);
        h2d(grd->u, grd->g.u.data(), nr, s);
        h2d(grd->mirror, grd->g.mirror.data(), nr, s);
        h2d(grd->w, pnt->w, nlat, s);
        h2d(grd->lon, pnt->lon, pnt->nlon, s);
        HIPCHK((hipError_t)launch_seeds((int)nr, grd->u, tb.dm, (int)grd->nmax_grd, grd->sld,
                                        grd->sx, grd->six, s));
        HIPCHK(hipStreamSynchronize(s));
    } catch (const HipFail &e) {
        report(err, e, __func__);
        charm_amd_grid_free(grd);
        return nullptr;
    }
    return grd;
}

void charm_amd_grid_free(charm_amd_grid *grd)
{
    if (!grd) return;
    (void)hipDeviceSynchronize();
    (void)hipFree(grd->t);
    (void)hipFree(grd->u);
    (void)hipFree(grd->mirror);
    (void)hipFree(grd->w);
    (void)hipFree(grd->lon);
    (void)hipFree(grd->sx);
    (void)hipFree(grd->six);
    (void)hipFree(grd->trow);
    (void)hipFree(grd->urow);
    (void)hipFree(grd->rrow);
    delete grd;
}

void charm_amd_shs_grid_dev(const charm_amd_grid *grd, unsigned long nmax, double mu, double R,
                            const double *cs_dev, double *f_dev, void *stream, charm_err *err)
{
    if (!grd || nmax > grd->nmax_grd) {
        CHARM_AMD_ERR(err, CHARM_EFUNCARG, "\"nmax\" exceeds the degree of the grid.");
        return;
    }
    if (capture_refused((hipStream_t)stream, err, __func__)) return;
    try {
        hipStream_t s = (hipStream_t)stream;
        Tables tb = tables((int)std::max<unsigned long>(nmax, 1), s);
        DevGrid d{grd->t, grd->u, grd->mirror, grd->sx, grd->six, grd->sld, nullptr, nullptr, 0,
                  &grd->dead};
        DevBuf ratio, rp;
        if (!nearly_equal(grd->r0, R, charm_glob_threshold)) {
            // one radius for all rows: a single column of powers (row stride
            // 0), made on the device (no host copy under stream capture)
            const double q = R / grd->r0;
            ratio.alloc(grd->g.nlat * sizeof(double), s);
            rp.alloc((nmax + 2) * sizeof(double), s);
            HIPCHK((hipError_t)launch_fill(grd->g.nlat, q, ratio.as<double>(), s));
            HIPCHK((hipError_t)launch_powers((int)nmax + 1, q, rp.as<double>(), s));
            d.ratio = ratio.as<double>();
            d.rp = rp.as<double>();
            d.rld = 0;
        }
        // grids made by charm_crd_point_gl always admit the FFT
        run_shs_grid(grd->g, d, grd->lon, true, nmax, mu / R, cs_dev, tb, f_dev, s);
        HIPCHK(hipGetLastError());
    } catch (const HipFail &e) {
        report(err, e, __func__);
    }
}

void charm_amd_sha_grid_dev(const charm_amd_grid *grd, unsigned long nmax, double mu, double R,
                            const double *f_dev, double *cs_dev, void *stream, charm_err *err)
{
    if (!grd || nmax > grd->nmax_grd) {
        CHARM_AMD_ERR(err, CHARM_EFUNCARG, "\"nmax\" exceeds the degree of the grid.");
        return;
    }
    if (capture_refused((hipStream_t)stream, err, __func__)) return;
    try {
        hipStream_t s = (hipStream_t)stream;
        Tables tb = tables((int)std::max<unsigned long>(nmax, 1), s);
        const size_t nlat = grd->g.nlat;
        const double c = sha_const(grd->type, nlat, grd->r0, mu);
        RescaleTable sc(grd->r0, R, nmax, s);
        DevGrid d{grd->t, grd->u, grd->mirror, grd->sx, grd->six, grd->sld, nullptr, nullptr, 0,
                  &grd->dead};
        run_sha(grd->g, d, grd->w, c, f_dev, nmax, tb, sc.p, nmax, cs_dev, cs_dev + 1, 2, s);
        HIPCHK(hipGetLastError());
    } catch (const HipFail &e) {
        report(err, e, __func__);
    }
}

// nset coefficient sets (set-major, each an m-major triangle of interleaved
// (C, S) pairs of degree nmax) -> nset grids [set][row][lon]; the sums over
// degrees run as one FP64 matrix-core contraction per order.
void charm_amd_shs_grid_multi_dev(const charm_amd_grid *grd, unsigned long nmax, size_t nset,
                                  double mu, double R, const double *cs_dev, double *f_dev,
                                  void *stream, charm_err *err)
{
    if (!grd || nmax > grd->nmax_grd) {
        CHARM_AMD_ERR(err, CHARM_EFUNCARG, "\"nmax\" exceeds the degree of the grid.");
        return;
    }
    if (!nearly_equal(grd->r0, R, charm_glob_threshold)) {
        CHARM_AMD_ERR(err, CHARM_EFUNCARG,
                      "Multi-set synthesis needs the grid radius to equal the reference radius.");
        return;
    }
    if (nset == 0) return;
    if (capture_refused((hipStream_t)stream, err, __func__)) return;
    try {
        hipStream_t s = (hipStream_t)stream;
        Tables tb = tables((int)std::max<unsigned long>(nmax, 1), s);
        const size_t P = tri_len(nmax), nlat = grd->g.nlat, nlon = grd->g.nlon, nfc = nlon / 2 + 1;
        if (nset > (size_t)(INT32_MAX / 16) * 16 - 16) {
            CHARM_AMD_ERR(err, CHARM_EFUNCARG, "Too many coefficient sets.");
            return;
        }
        const int nsetp = (int)((nset + 15) / 16 * 16);
        DevBuf cs2((size_t)nsetp * P * 2 * sizeof(double), s);
        HIPCHK((hipError_t)launch_scale_cs_multi((int)nmax, nset, cs_dev, tb.sn, tb.nt, cs2.p, nsetp,
                                                 s));
        DevBuf spec(nset * nlat * nfc * 2 * sizeof(double), s);
        MultiShsArgs a{};
        a.nmax = (int)nmax;
        a.nring = (int)grd->g.nring;
        a.nlat = (int)grd->g.mir_nlat;
        a.nset = (int)nset;
        a.mcount = (int)nfc;
        a.nrt = (int)((grd->g.nring + 63) / 64);
        a.nst = nsetp / 16;
        a.nunit = (int)((nfc + 7) / 8) * a.nst;
        a.nsetp = nsetp;
        a.t = grd->t;
        a.sx = grd->sx;
        a.six = grd->six;
        a.sld = grd->sld;
        a.mirror = grd->mirror;
        a.al = tb.al;
        a.al_nt = tb.nt;
        a.cs = cs2.p;
        a.spec = spec.p;
        a.lds = nfc;
        a.mur = mu / R;
        PolarRings pol(grd->g, grd->u, nmax, 0, s);
        a.pm0 = pol.p;
        {
            ProfScope ps("shs_multi_kernel", s);
            HIPCHK((hipError_t)launch_shs_multi(a, s));
        }
        {
            ProfScope ps("fft_c2r_multi", s);
            rows_c2r(nlon, nset * nlat, spec.p, nfc, f_dev, s);
        }
        HIPCHK(hipGetLastError());
    } catch (const HipFail &e) {
        report(err, e, __func__);
    }
}

void charm_amd_cs_truncate(const double *src_dev, unsigned long nmax_src, double *dst_dev,
                           unsigned long nmax, void *stream, charm_err *err)
{
    if (nmax > nmax_src) {
        CHARM_AMD_ERR(err, CHARM_EFUNCARG, "\"nmax\" exceeds the degree of the source model.");
        return;
    }
    try {
        HIPCHK((hipError_t)launch_cs_truncate(src_dev, nmax_src, dst_dev, nmax, stream));
    } catch (const HipFail &e) {
        report(err, e, __func__);
    }
}

void charm_amd_shs_sctr_dev(size_t npoint, const double *lat_dev, const double *lon_dev,
                            const double *r_dev, unsigned long nmax, double mu, double R,
                            const double *cs_dev, double *f_dev, void *stream, charm_err *err)
{
    if (npoint == 0) return;
    if (capture_refused((hipStream_t)stream, err, __func__)) return;
    try {
        hipStream_t s = (hipStream_t)stream;
        Tables tb = tables((int)std::max<unsigned long>(nmax, 1), s);
        // radii equal to R within glob_threshold (src/shs/shs_r_eq_rref.c:17-30),
        // decided on the device: no host round trip, nothing a stream under
        // capture could not record
        DevBuf flag(sizeof(int), s);
        HIPCHK((hipError_t)launch_all_near_dev(npoint, r_dev, R, charm_glob_threshold,
                                               flag.as<int>(), s));
        run_sctr(npoint, lat_dev, lon_dev, r_dev, false, nmax, mu, R, cs_dev, tb, f_dev, s,
                 flag.as<int>());
        HIPCHK(hipGetLastError());
    } catch (const HipFail &e) {
        report(err, e, __func__);
    }
}

// Derivatives on a device-resident grid: outputs one after another in f_dev
void charm_amd_shs_grid_deriv_dev(const charm_amd_grid *grd, unsigned long nmax, double mu,
                                  double R, const double *cs_dev, int what, double *f_dev,
                                  void *stream, charm_err *err)
{
    GradPlan plan;
    if (!deriv_plan(what, plan)) {
        CHARM_AMD_ERR(err, CHARM_EFUNCARG, "Unsupported derivative request \"what\".");
        return;
    }
    if (!grd || nmax > grd->nmax_grd) {
        CHARM_AMD_ERR(err, CHARM_EFUNCARG, "\"nmax\" exceeds the degree of the grid.");
        return;
    }
    if (capture_refused((hipStream_t)stream, err, __func__)) return;
    try {
        hipStream_t s = (hipStream_t)stream;
        Tables tb = tables((int)std::max<unsigned long>(nmax, 1), s);
        DevGrid d{grd->t, grd->u, grd->mirror, grd->sx, grd->six, grd->sld, nullptr, nullptr, 0,
                  &grd->dead};
        DevBuf ratio, rp;
        const bool unit = nearly_equal(grd->r0, R, charm_glob_threshold);
        if (!unit) {
            const double q = R / grd->r0;
            ratio.alloc(grd->g.nlat * sizeof(double), s);
            rp.alloc((nmax + 2) * sizeof(double), s);
            HIPCHK((hipError_t)launch_fill(grd->g.nlat, q, ratio.as<double>(), s));
            HIPCHK((hipError_t)launch_powers((int)nmax + 1, q, rp.as<double>(), s));
            d.ratio = ratio.as<double>();
            d.rp = rp.as<double>();
            d.rld = 0;
        }
        const size_t nrow = grd->g.nlat, ncol = grd->g.nlon;
        if (!grad_grid_fft(plan, grd->g, d, nmax, mu / R, cs_dev, tb, grd->trow, grd->urow,
                           grd->rrow, unit, R, f_dev, s)) {
            DevBuf tabs(plan.tabs.size() * nrow * ncol * sizeof(double), s);
            grad_tables_grid(grd->g, d, grd->lon, true, nmax, mu / R, cs_dev, tb, plan,
                             tabs.as<double>(), s);
            grad_combine(plan, nrow, ncol, grd->trow, grd->urow, grd->rrow, unit, R,
                         tabs.as<double>(), f_dev, s);
        }
        HIPCHK(hipGetLastError());
    } catch (const HipFail &e) {
        report(err, e, __func__);
    }
}

// Derivatives at scattered points given by device arrays
void charm_amd_shs_sctr_deriv_dev(size_t npoint, const double *lat_dev, const double *lon_dev,
                                  const double *r_dev, unsigned long nmax, double mu, double R,
                                  const double *cs_dev, int what, double *f_dev, void *stream,
                                  charm_err *err)
{
    GradPlan plan;
    if (!deriv_plan(what, plan)) {
        CHARM_AMD_ERR(err, CHARM_EFUNCARG, "Unsupported derivative request \"what\".");
        return;
    }
    if (npoint == 0) return;
    if (capture_refused((hipStream_t)stream, err, __func__)) return;
    try {
        hipStream_t s = (hipStream_t)stream;
        Tables tb = tables((int)std::max<unsigned long>(nmax, 1), s);
        int unit = 0;
        HIPCHK((hipError_t)launch_all_near(npoint, r_dev, R, charm_glob_threshold, &unit, s));
        DevBuf tabs(plan.tabs.size() * npoint * sizeof(double), s);
        DevBuf dt(npoint * sizeof(double), s), du(npoint * sizeof(double), s);
        HIPCHK((hipError_t)launch_sincos(npoint, lat_dev, dt.as<double>(), du.as<double>(), s));
        grad_tables_sctr(npoint, lat_dev, lon_dev, r_dev, unit != 0, nmax, mu, R, cs_dev, tb, plan,
                         tabs.as<double>(), s);
        grad_combine(plan, npoint, 1, dt.as<double>(), du.as<double>(), r_dev, unit != 0, R,
                     tabs.as<double>(), f_dev, s);
        HIPCHK(hipGetLastError());
    } catch (const HipFail &e) {
        report(err, e, __func__);
    }
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Batched execution
// ---------------------------------------------------------------------------
struct charm_amd_batch {
    std::vector<charm_amd_request> req;
    std::vector<void *> owned;
    ShsGridArgs *d_shs = nullptr;
    ShaArgs *d_sha = nullptr;
    int2 *d_shs_items = nullptr, *d_sha_items = nullptr;
    int n_shs_items = 0, n_sha_items = 0;
    int *d_counter = nullptr;
    struct Job {
        size_t r;
        void *spec;
        void *part;   // SHA: tile partials; SHS: coefficients times s_nm
        int ntile;
        const double *scale;
        size_t nlat, nlon, nfc;
    };
    std::vector<Job> shs, sha;
    const double *sn = nullptr;  // s_nm table (degree nt) the batch was built with
    int nt = 0;
    struct SJob {
        size_t r;
        void *cs2;     // coefficients times s_nm
        double *part;  // [chunk][point]
        int nchunk;
        const int *perm;
    };
    std::vector<SJob> sct;
    SctrArgs *d_sct = nullptr;
    int2 *d_sct_items = nullptr;
    int n_sct_items = 0;
    int grid_sct = 0;
    int grid_shs = 0, grid_sha = 0;
    // the per-request steps around the grouped kernels, each in one launch:
    // row transforms of the grids with single-row plans (the others run per
    // request), coefficient scaling and the reductions
    std::vector<charm_amd::RowGroupPart> r2c_group, c2r_group;
    std::vector<size_t> r2c_alone, c2r_alone;  // indices into sha / shs
    charm_amd::ScaleJob *d_scale_shs = nullptr, *d_scale_sct = nullptr;
    charm_amd::ShaReduceJob *d_red = nullptr;
    charm_amd::SctrReduceJob *d_sred = nullptr;
    int nmax_shs = 0, nmax_sha = 0, nmax_sct = 0;
    size_t npoint_max = 0;
    bool use_graph = false;
    hipGraphExec_t exec = nullptr;
    hipStream_t cap = nullptr;
};

namespace charm_amd {
namespace {

void *dmalloc(charm_amd_batch *b, size_t bytes)
{
    void *p = nullptr;
    HIPCHK(hipMalloc(&p, std::max<size_t>(bytes, 16)));
    b->owned.push_back(p);
    return p;
}

template <class T>
T *dupload(charm_amd_batch *b, const std::vector<T> &v, hipStream_t s)
{
    T *p = (T *)dmalloc(b, v.size() * sizeof(T));
    h2d(p, v.data(), v.size(), s);
    return p;
}

// First skipped order per ring for the polar optimisation (PolarRings), in
// memory the batch owns; nullptr when the optimisation is off
const int *polar_rings_owned(charm_amd_batch *b, const GridGeom &g, unsigned long nmax,
                             int analysis, hipStream_t s)
{
    const double thr = polar_threshold(nmax);
    if (thr < 0.0 || g.nring == 0) return nullptr;
    std::vector<int> m0(g.nring);
    polar_ring_m0(g.u.data(), g.nring, nmax, thr, polar_block(analysis), m0.data());
    return dupload(b, m0, s);
}

void batch_enqueue(charm_amd_batch *b, hipStream_t s)
{
    HIPCHK((hipError_t)launch_zero_ints(b->d_counter, 3, s));
    if (!b->sha.empty()) {
        HIPCHK((hipError_t)row_group_launch(0, b->r2c_group, s));
        for (size_t i : b->r2c_alone) {
            const auto &j = b->sha[i];
            rows_r2c(j.nlon, j.nlat, b->req[j.r].f, j.spec, j.nfc, s);
        }
        HIPCHK((hipError_t)launch_sha_batch(b->d_sha, b->d_sha_items, b->n_sha_items,
                                            b->d_counter + 1, b->grid_sha, s));
        HIPCHK((hipError_t)launch_sha_reduce_group(b->d_red, (int)b->sha.size(), b->nmax_sha, b->sn,
                                                   b->nt, s));
    }
    if (!b->shs.empty()) {
        HIPCHK((hipError_t)launch_scale_cs_group(b->d_scale_shs, (int)b->shs.size(), b->nmax_shs,
                                                 b->sn, b->nt, s));
        HIPCHK((hipError_t)launch_shs_batch(b->d_shs, b->d_shs_items, b->n_shs_items, b->d_counter,
                                            b->grid_shs, s));
        HIPCHK((hipError_t)row_group_launch(1, b->c2r_group, s));
        for (size_t i : b->c2r_alone) {
            const auto &j = b->shs[i];
            rows_c2r(j.nlon, j.nlat, j.spec, j.nfc, b->req[j.r].f, s);
        }
    }
    if (!b->sct.empty()) {
        HIPCHK((hipError_t)launch_scale_cs_group(b->d_scale_sct, (int)b->sct.size(), b->nmax_sct,
                                                 b->sn, b->nt, s));
        HIPCHK((hipError_t)launch_sctr_batch(b->d_sct, b->d_sct_items, b->n_sct_items,
                                             b->d_counter + 2, b->grid_sct, s));
        HIPCHK((hipError_t)launch_sctr_reduce_group(b->d_sred, (int)b->sct.size(), b->npoint_max,
                                                    s));
    }
}

}  // namespace
}  // namespace charm_amd

extern "C" {

charm_amd_batch *charm_amd_batch_create(const charm_amd_request *reqs, size_t nreq, int use_graph,
                                        charm_err *err)
{
    for (size_t i = 0; i < nreq; i++) {
        const charm_amd_request &q = reqs[i];
        const bool sct = q.kind == CHARM_AMD_SHS_SCTR;
        const bool bad =
            sct ? (!q.lat || !q.lon || !q.r || q.npoint == 0 || q.npoint >= (size_t)INT32_MAX ||
                   (q.npoint + 511) / 512 >= (1u << 24))
                : (!q.grid || (q.kind != CHARM_AMD_SHS_GRID && q.kind != CHARM_AMD_SHA_GRID) ||
                   q.nmax > q.grid->nmax_grd);
        if (bad || !q.cs || !q.f) {
            char msg[CHARM_ERR_MAX_MSG];
            std::snprintf(msg, sizeof msg, "Invalid batch request %zu.", i);
            CHARM_AMD_ERR(err, CHARM_EFUNCARG, msg);
            return nullptr;
        }
    }
    charm_amd_batch *b = new charm_amd_batch();
    b->req.assign(reqs, reqs + nreq);
    b->use_graph = use_graph != 0;
    try {
        hipStream_t s = thread_stream();
        unsigned long nmax_all = 1;
        for (auto &q : b->req) nmax_all = std::max(nmax_all, q.nmax);
        Tables tb = tables((int)nmax_all, s);
        b->sn = tb.sn;
        b->nt = tb.nt;
        std::vector<ShsGridArgs> shs_args;
        std::vector<ShaArgs> sha_args;
        struct Item {
            long cost;
            int2 w;
        };
        std::vector<Item> shs_items, sha_items;
        std::vector<SctrArgs> sct_args;
        std::vector<Item> sct_items;
        for (size_t i = 0; i < nreq; i++) {
            const charm_amd_request &q = b->req[i];
            if (q.kind == CHARM_AMD_SHS_SCTR) {
                // points ordered by latitude once, as charm_amd_shs_sctr_dev does per call
                const size_t np = q.npoint;
                const int K = batch_ring_k(2);
                int *perm = (int *)dmalloc(b, np * sizeof(int));
                double *pt = (double *)dmalloc(b, np * sizeof(double));
                double *pu = (double *)dmalloc(b, np * sizeof(double));
                double *pl = (double *)dmalloc(b, np * sizeof(double));
                double *pr = (double *)dmalloc(b, np * sizeof(double));
                HIPCHK((hipError_t)launch_sort_points(np, q.lat, q.lon, q.r, perm, pt, pu, pl, pr, s));
                int unit = 0;
                HIPCHK((hipError_t)launch_all_near(np, q.r, q.R, charm_glob_threshold, &unit, s));
                const std::vector<int> mc = sctr_chunks(np, q.nmax);
                const int nchunk = (int)mc.size() - 1;
                SctrArgs a{};
                a.nmax = (int)q.nmax;
                a.npoint = np;
                a.nchunk = nchunk;
                a.mchunk = dupload(b, mc, s);
                a.unit = unit != 0;
                a.R = q.R;
                a.lat_t = pt;
                a.lat_u = pu;
                a.lon = pl;
                a.r = pr;
                a.dm = tb.dm;
                a.al = tb.al;
                a.al_nt = tb.nt;
                void *cs2 = dmalloc(b, tri_len(q.nmax) * 2 * sizeof(double));
                a.cs = cs2;
                double *part = (double *)dmalloc(b, (size_t)nchunk * np * sizeof(double));
                a.part = part;
                a.k = K;
                const double thr = polar_threshold(q.nmax);
                if (thr >= 0.0) {
                    int *m0 = (int *)dmalloc(b, np * sizeof(int));
                    HIPCHK((hipError_t)launch_polar_points(np, q.lat, perm, polar_block(0),
                                                           (int)q.nmax, thr, m0, s));
                    a.pm0 = m0;
                }
                const int idx = (int)sct_args.size();
                sct_args.push_back(a);
                b->sct.push_back({i, cs2, part, nchunk, perm});
                const int nblk = (int)((np + 256 * K - 1) / (256 * K));
                for (int c = 0; c < nchunk; c++) {
                    long work = 0;
                    for (int m = mc[c]; m < mc[c + 1]; m++) work += (long)q.nmax + 1 - m;
                    for (int bk = 0; bk < nblk; bk++)
                        sct_items.push_back({work, make_int2(idx, (bk << 6) | c)});
                }
                continue;
            }
            const charm_amd_grid *g = q.grid;
            const size_t nlat = g->g.nlat, nlon = g->g.nlon, nfc = nlon / 2 + 1;
            const int nmax = (int)q.nmax;
            if (q.kind == CHARM_AMD_SHS_GRID) {
                const int K = batch_ring_k(0);
                ShsGridArgs a{};
                a.nmax = nmax;
                a.nring = (int)g->g.nring;
                a.nlat = (int)g->g.mir_nlat;
                a.k = K;
                a.ntile = ring_tiles(g->g.nring, K);
                a.t = g->t;
                a.sx = g->sx;
                a.six = g->six;
                a.sld = g->sld;
                a.mirror = g->mirror;
                a.al = tb.al;
                a.al_nt = tb.nt;
                a.pm0 = polar_rings_owned(b, g->g, q.nmax, 0, s);
                void *cs2 = dmalloc(b, tri_len(q.nmax) * 2 * sizeof(double));
                a.cs = cs2;
                a.mur = q.mu / q.R;
                a.fft_layout = 1;
                a.mcount = (int)nfc;
                a.ldo = nfc;
                a.unit = 1;
                if (!nearly_equal(g->r0, q.R, charm_glob_threshold)) {
                    const double qq = q.R / g->r0;
                    std::vector<double> pw(q.nmax + 2), rat(nlat, qq);
                    pw[0] = 1.0;
                    for (unsigned long k = 1; k <= q.nmax + 1; k++) pw[k] = pw[k - 1] * qq;
                    a.ratio = dupload(b, rat, s);
                    a.rp = dupload(b, pw, s);
                    a.rld = 0;
                    a.unit = 0;
                }
                void *spec = dmalloc(b, nlat * nfc * 2 * sizeof(double));
                a.out = spec;
                const int idx = (int)shs_args.size();
                shs_args.push_back(a);
                b->shs.push_back({i, spec, cs2, a.ntile, nullptr, nlat, nlon, nfc});
                const int nquad = (a.mcount + 3) / 4;
                for (int qd = 0; qd < nquad; qd++)
                    for (int tl = 0; tl < a.ntile; tl++) {
                        const long m0 = 4L * qd;
                        const long cost = (m0 <= nmax ? (nmax + 1 - m0) : 0) + 1;
                        shs_items.push_back({cost, make_int2(idx, (tl << 16) | qd)});
                    }
            } else {
                const int K = batch_ring_k(1);
                ShaArgs a{};
                a.nmax = nmax;
                a.nring = (int)g->g.nring;
                a.nlat = (int)g->g.mir_nlat;
                a.nrow = (int)g->g.nlat;
                a.k = K;
                a.ntile = ring_tiles(g->g.nring, K);
                a.tpg = 1;
                a.t = g->t;
                a.sx = g->sx;
                a.six = g->six;
                a.sld = g->sld;
                a.mirror = g->mirror;
                a.cw = g->w;
                a.cws = sha_const(g->type, nlat, g->r0, q.mu);
                void *spec = dmalloc(b, nlat * nfc * 2 * sizeof(double));
                a.spec = spec;
                a.lds = nfc;
                a.al = tb.al;
                a.al_nt = tb.nt;
                a.pm0 = polar_rings_owned(b, g->g, q.nmax, 1, s);
                void *part = dmalloc(b, (size_t)a.ntile * tri_len(q.nmax) * 2 * sizeof(double));
                a.part = part;
                std::vector<double> sc;
                const double *scale = nullptr;
                if (rescale_table(g->r0, q.R, q.nmax, sc)) scale = dupload(b, sc, s);
                const int idx = (int)sha_args.size();
                sha_args.push_back(a);
                b->sha.push_back({i, spec, part, a.ntile, scale, nlat, nlon, nfc});
                const int nquad = (nmax + 1 + 3) / 4;
                for (int qd = 0; qd < nquad; qd++)
                    for (int tl = 0; tl < a.ntile; tl++)
                        sha_items.push_back({(long)(nmax + 1 - 4 * qd), make_int2(idx, (tl << 16) | qd)});
            }
        }
        // descriptors of the grouped per-request steps
        {
            std::vector<RowJob> rj;
            for (size_t i = 0; i < b->sha.size(); i++) {
                const auto &j = b->sha[i];
                if (row_group_supported(j.nlon))
                    rj.push_back({j.nlon, j.nlat, b->req[j.r].f, j.nlon, j.spec, j.nfc});
                else
                    b->r2c_alone.push_back(i);
            }
            if (!rj.empty())
                HIPCHK((hipError_t)row_group_build(rj.data(), (int)rj.size(), s, &b->r2c_group,
                                                   &b->owned));
            rj.clear();
            for (size_t i = 0; i < b->shs.size(); i++) {
                const auto &j = b->shs[i];
                if (row_group_supported(j.nlon))
                    rj.push_back({j.nlon, j.nlat, b->req[j.r].f, j.nlon, j.spec, j.nfc});
                else
                    b->c2r_alone.push_back(i);
            }
            if (!rj.empty())
                HIPCHK((hipError_t)row_group_build(rj.data(), (int)rj.size(), s, &b->c2r_group,
                                                   &b->owned));
            std::vector<ScaleJob> sj;
            for (const auto &j : b->shs) {
                const charm_amd_request &q = b->req[j.r];
                sj.push_back({(int)q.nmax, q.cs, j.part});
                b->nmax_shs = std::max(b->nmax_shs, (int)q.nmax);
            }
            if (!sj.empty()) b->d_scale_shs = dupload(b, sj, s);
            sj.clear();
            std::vector<SctrReduceJob> sr;
            for (const auto &j : b->sct) {
                const charm_amd_request &q = b->req[j.r];
                sj.push_back({(int)q.nmax, q.cs, j.cs2});
                sr.push_back({q.npoint, j.nchunk, j.part, j.perm, q.mu / q.R, q.f});
                b->nmax_sct = std::max(b->nmax_sct, (int)q.nmax);
                b->npoint_max = std::max(b->npoint_max, q.npoint);
            }
            if (!sj.empty()) b->d_scale_sct = dupload(b, sj, s);
            if (!sr.empty()) b->d_sred = dupload(b, sr, s);
            std::vector<ShaReduceJob> rd;
            for (const auto &j : b->sha) {
                const charm_amd_request &q = b->req[j.r];
                rd.push_back({(int)q.nmax, j.ntile, j.part, j.scale, q.cs});
                b->nmax_sha = std::max(b->nmax_sha, (int)q.nmax);
            }
            if (!rd.empty()) b->d_red = dupload(b, rd, s);
        }
        auto sort_items = [](std::vector<Item> &v) {
            std::stable_sort(v.begin(), v.end(),
                             [](const Item &x, const Item &y) { return x.cost > y.cost; });
            std::vector<int2> w(v.size());
            for (size_t i = 0; i < v.size(); i++) w[i] = v[i].w;
            return w;
        };
        std::vector<int2> wi = sort_items(shs_items), wa = sort_items(sha_items);
        std::vector<int2> wsc = sort_items(sct_items);
        b->n_sct_items = (int)wsc.size();
        if (!sct_args.empty()) b->d_sct = dupload(b, sct_args, s);
        if (!wsc.empty()) b->d_sct_items = dupload(b, wsc, s);
        b->grid_sct = batch_grid_size(2);
        b->n_shs_items = (int)wi.size();
        b->n_sha_items = (int)wa.size();
        if (!shs_args.empty()) b->d_shs = dupload(b, shs_args, s);
        if (!sha_args.empty()) b->d_sha = dupload(b, sha_args, s);
        if (!wi.empty()) b->d_shs_items = dupload(b, wi, s);
        if (!wa.empty()) b->d_sha_items = dupload(b, wa, s);
        b->d_counter = (int *)dmalloc(b, 3 * sizeof(int));
        b->grid_shs = batch_grid_size(0);
        b->grid_sha = batch_grid_size(1);
        HIPCHK(hipStreamSynchronize(s));
    } catch (const HipFail &e) {
        report(err, e, __func__);
        charm_amd_batch_free(b);
        return nullptr;
    }
    return b;
}

void charm_amd_batch_run(charm_amd_batch *b, void *stream, charm_err *err)
{
    if (!b) {
        CHARM_AMD_ERR(err, CHARM_EFUNCARG, "NULL batch.");
        return;
    }
    hipStream_t s = (hipStream_t)stream;
    try {
        if (b->exec) {
            HIPCHK(hipGraphLaunch(b->exec, s));
            return;
        }
        batch_enqueue(b, s);
        if (b->use_graph) {
            // record the same work once for later replays
            HIPCHK(hipStreamSynchronize(s));
            if (!b->cap) HIPCHK(hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking));
            // FFT plans for the capture stream are made (and warmed) before the
            // capture: plan creation allocates and synchronises
            for (const auto &j : b->sha) rows_prepare(1, j.nlon, j.nlat, j.nfc, b->cap);
            for (const auto &j : b->shs) rows_prepare(0, j.nlon, j.nlat, j.nfc, b->cap);
            batch_enqueue(b, b->cap);
            HIPCHK(hipStreamSynchronize(b->cap));
            hipGraph_t graph = nullptr;
            HIPCHK(hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal));
            try {
                batch_enqueue(b, b->cap);
            } catch (...) {
                (void)hipStreamEndCapture(b->cap, &graph);
                if (graph) (void)hipGraphDestroy(graph);
                throw;
            }
            HIPCHK(hipStreamEndCapture(b->cap, &graph));
            hipError_t e = hipGraphInstantiate(&b->exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (e != hipSuccess) {
                b->exec = nullptr;
                b->use_graph = false;
            }
        }
    } catch (const HipFail &e) {
        report(err, e, __func__);
    }
}

int charm_amd_batch_uses_graph(const charm_amd_batch *b)
{
    return (b && b->exec) ? 1 : 0;
}

void charm_amd_batch_free(charm_amd_batch *b)
{
    if (!b) return;
    (void)hipDeviceSynchronize();
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->cap) (void)hipStreamDestroy(b->cap);
    for (void *p : b->owned) (void)hipFree(p);
    delete b;
}

}  // extern "C"
