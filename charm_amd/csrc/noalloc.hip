// Scattered-point synthesis without device allocation
// (charm_amd_shs_sctr_dev_noalloc), for callers that cannot allocate: code
// under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc.  Every buffer the transform needs beyond its inputs and output
// lives in one file-scope device array, `scratch`, of NOALLOC_BYTES per GPU:
//
//   [tables]   alpha_nm, s_nm (m-major, degree NOALLOC_NMAX) and dm[n], built
//              once per device by the first call on it
//   [flags]    the radii flag (all r equal to R) and the order chunks
//   [points]   permutation, sorted sin/cos lat, lon, r, polar first orders
//   [partial]  per-chunk partial sums
//   [sort]     key/index ping-pong buffers and the sort's own temporary space
//              (after the sort: the per-block polar cosines)
//
// The kernels are those of charm_amd_shs_sctr_dev and take the scratch
// address as an argument, so both entry points give the same results.  Every
// call on a GPU shares that GPU's array and overwrites the per-call regions:
// calls that overlap on one GPU race, and must be serialised by the caller.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <mutex>

#include "internal.h"

namespace {

constexpr unsigned long NOALLOC_NMAX = 2190;
constexpr size_t NOALLOC_NPOINT = size_t(1) << 20;
constexpr size_t NOALLOC_BYTES = size_t(128) << 20;
constexpr int NOALLOC_MAXDEV = 64;

static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

inline size_t al256(size_t x)
{
    return (x + 255) & ~(size_t)255;
}

// Offsets into the scratch array for a call with np points and nchunk order
// chunks; end is the bytes used.
struct Layout {
    size_t al, sn, dm, flag, mc, perm, t, u, lon, r, m0, part, sort, sort_bytes, end;
    Layout(size_t np, int nchunk, size_t sort_ws)
    {
        const size_t P = charm_amd::tri_len(NOALLOC_NMAX);
        al = 0;
        sn = al + al256(P * sizeof(double));
        dm = sn + al256(P * sizeof(double));
        flag = dm + al256((NOALLOC_NMAX + 1) * sizeof(double));
        mc = flag + 256;
        perm = mc + al256((charm_amd::SCTR_MAX_CHUNKS + 1) * sizeof(int));
        t = perm + al256(np * sizeof(int));
        u = t + al256(np * sizeof(double));
        lon = u + al256(np * sizeof(double));
        r = lon + al256(np * sizeof(double));
        m0 = r + al256(np * sizeof(double));
        part = m0 + al256(np * sizeof(int));
        sort = part + al256((size_t)nchunk * np * sizeof(double));
        sort_bytes = sort_ws;
        end = sort + al256(sort_ws);
    }
};

struct DevScratch {
    char *ptr = nullptr;  // the device address of this device's scratch array
    bool ready = false;
};
std::mutex g_mu;
DevScratch g_dev[NOALLOC_MAXDEV];

struct Fail {
    hipError_t code;
    const char *what;
    int line;
};

#define NACHK(expr)                                               \
    do {                                                          \
        hipError_t e_ = (expr);                                   \
        if (e_ != hipSuccess) throw Fail{e_, #expr, __LINE__};    \
    } while (0)

// Set-up, once per device with that device current: the device address of
// its scratch array, and the recurrence tables of degree NOALLOC_NMAX in it.
char *device_scratch(hipStream_t s)
{
    int dev = 0;
    NACHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= NOALLOC_MAXDEV) throw Fail{hipErrorInvalidDevice, "device index", __LINE__};
    std::lock_guard<std::mutex> lk(g_mu);
    DevScratch &d = g_dev[dev];
    if (d.ready) return d.ptr;
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    NACHK(hipStreamIsCapturing(s, &cap));
    if (cap != hipStreamCaptureStatusNone)
        throw Fail{hipErrorStreamCaptureUnsupported, "first call on a device under stream capture",
                   __LINE__};
    void *p = nullptr;
    NACHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    char *base = (char *)p;
    // the tables, built through the per-point region (free at this point)
    const int nt = (int)NOALLOC_NMAX;
    std::vector<double> r, ri, dm;
    charm_amd::leg_host_tables(nt, r, ri, dm);
    const Layout L(0, 0, 0);
    double *dr = (double *)(base + L.perm);
    double *dri = dr + r.size();
    void *ab = (void *)(base + al256(L.perm + 2 * r.size() * sizeof(double)));
    if ((char *)ab - base + charm_amd::tri_len(nt) * 2 * sizeof(double) > NOALLOC_BYTES)
        throw Fail{hipErrorInvalidValue, "scratch size", __LINE__};
    NACHK(hipMemcpyAsync(dr, r.data(), r.size() * sizeof(double), hipMemcpyHostToDevice, s));
    NACHK(hipMemcpyAsync(dri, ri.data(), ri.size() * sizeof(double), hipMemcpyHostToDevice, s));
    NACHK(hipMemcpyAsync(base + L.dm, dm.data(), dm.size() * sizeof(double), hipMemcpyHostToDevice,
                         s));
    NACHK((hipError_t)charm_amd::launch_leg_tables(nt, dr, dri, ab, (double *)(base + L.al),
                                                   (double *)(base + L.sn), s));
    NACHK(hipStreamSynchronize(s));
    d.ptr = base;
    d.ready = true;
    return base;
}

}  // namespace

extern "C" {

void charm_amd_shs_sctr_dev_noalloc(size_t npoint, const double *lat_dev, const double *lon_dev,
                                    const double *r_dev, unsigned long nmax, double mu, double R,
                                    const double *cs_dev, double *f_dev, void *stream,
                                    charm_err *err)
{
    using namespace charm_amd;
    if (npoint > NOALLOC_NPOINT || nmax > NOALLOC_NMAX) {
        char msg[CHARM_ERR_MAX_MSG];
        std::snprintf(msg, sizeof msg,
                      "charm_amd_shs_sctr_dev_noalloc takes at most %zu points and degree %lu "
                      "(got %zu points, degree %lu).",
                      NOALLOC_NPOINT, NOALLOC_NMAX, npoint, nmax);
        CHARM_AMD_ERR(err, CHARM_EFUNCARG, msg);
        return;
    }
    if (npoint == 0) return;
    hipStream_t s = (hipStream_t)stream;
    try {
        char *base = device_scratch(s);
        int mc[SCTR_MAX_CHUNKS + 1];
        const int nchunk = sctr_order_chunks(npoint, nmax, mc);
        const size_t ws = sort_points_ws_bytes(npoint);
        if (!ws) throw Fail{hipErrorInvalidValue, "sort workspace query", __LINE__};
        const Layout L(npoint, nchunk, ws);
        if (L.end > NOALLOC_BYTES) throw Fail{hipErrorInvalidValue, "scratch size", __LINE__};
        int *flag = (int *)(base + L.flag);
        int *dmc = (int *)(base + L.mc);
        // radii equal to R within glob_threshold (src/shs/shs_r_eq_rref.c:17-30),
        // decided on the device
        NACHK((hipError_t)launch_all_near_dev(npoint, r_dev, R, charm_glob_threshold, flag, s));
        IntVals v{};
        v.n = nchunk + 1;
        for (int c = 0; c <= nchunk; c++) v.v[c] = mc[c];
        NACHK((hipError_t)launch_set_ints(dmc, v, s));
        double *t = (double *)(base + L.t), *u = (double *)(base + L.u);
        double *lon = (double *)(base + L.lon), *r = (double *)(base + L.r);
        const int *perm = nullptr;
        NACHK((hipError_t)launch_sort_points_ws(npoint, lat_dev, lon_dev, r_dev,
                                                (int *)(base + L.perm), base + L.sort, L.sort_bytes,
                                                t, u, lon, r, &perm, s));
        SctrArgs a{};
        a.nmax = (int)nmax;
        a.npoint = npoint;
        a.nchunk = nchunk;
        a.mchunk = dmc;
        a.unit_dev = flag;
        a.R = R;
        a.lat_t = t;
        a.lat_u = u;
        a.lon = lon;
        a.r = r;
        a.dm = (const double *)(base + L.dm);
        a.al = (const double *)(base + L.al);
        a.al_nt = (int)NOALLOC_NMAX;
        a.cs = cs_dev;
        a.sn = (const double *)(base + L.sn);
        a.part = (double *)(base + L.part);
        a.k = sctr_points_per_lane(npoint);
        const double thr = polar_threshold(nmax);
        if (thr >= 0.0) {
            // the sort's buffers are free again: per-block cosines there
            NACHK((hipError_t)launch_polar_points_ws(npoint, lat_dev, perm, polar_block(0),
                                                     (int)nmax, thr, (double *)(base + L.sort),
                                                     (int *)(base + L.m0), s));
            a.pm0 = (const int *)(base + L.m0);
        }
        NACHK((hipError_t)launch_sctr(a, s));
        NACHK((hipError_t)launch_sctr_reduce(npoint, nchunk, a.part, perm, mu / R, f_dev, s));
    } catch (const Fail &f) {
        char msg[CHARM_ERR_MAX_MSG];
        std::snprintf(msg, sizeof msg, "HIP error \"%s\" in \"%s\".", hipGetErrorString(f.code),
                      f.what);
        err_set(err, __FILE__, (size_t)f.line, __func__, CHARM_EFUNCARG, msg);
    }
}

}  // extern "C"
