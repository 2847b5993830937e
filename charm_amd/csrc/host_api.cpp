// Host-side data types of the C ABI: charm_err, charm_shc, charm_point,
// Gauss-Legendre grids and the global tunables.  Written for this library;
// behaviour (argument checks, ownership, error codes) follows CHarm 0.4.9.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

#include "charm_amd.h"
#include "internal.h"

// ---------------------------------------------------------------------------
// Global tunables (src/glob/glob.c:16-23)
// ---------------------------------------------------------------------------
extern "C" {
double charm_glob_threshold = 100.0 * 2.220446049250313080847263336181640625e-16;
double charm_glob_threshold2 = 100000.0 * 2.220446049250313080847263336181640625e-16;
unsigned long charm_glob_polar_optimization_a1 = 100;
double charm_glob_polar_optimization_a2 = -1.0;
// Ring blocks of the polar optimisation: CHarm tests it jointly over
// SIMD_SIZE * BLOCK_S rings (synthesis) or SIMD_SIZE * BLOCK_A rings
// (analysis), with BLOCK_S/BLOCK_A = glob_{shs,sha}_block_lat_multiplier
// (src/glob/glob.c:26-29, src/simd/simd.h:298-304,609-611).  The defaults are
// the AVX-512 double build: 8 lanes, multipliers 4 and 8, blocks of 32 and 64.
size_t charm_glob_sha_block_lat_multiplier = 8;
size_t charm_glob_shs_block_lat_multiplier = 4;
int charm_amd_polar_simd_size = 8;
}

namespace charm_amd {

double polar_threshold(unsigned long nmax)
{
    if (charm_glob_polar_optimization_a2 < 0.0) return -1.0;
    return std::max((double)charm_glob_polar_optimization_a1,
                    charm_glob_polar_optimization_a2 * (double)nmax);
}

size_t polar_block(int analysis)
{
    const size_t v = charm_amd_polar_simd_size > 0 ? (size_t)charm_amd_polar_simd_size : 1;
    const size_t b = analysis ? charm_glob_sha_block_lat_multiplier
                              : charm_glob_shs_block_lat_multiplier;
    return v * (b > 0 ? b : 1);
}

// First order skipped by the polar optimisation at a block whose largest
// cos(lat) is umax: misc_polar_optimization_apply skips m when
// thr < m - nmax * umax (src/misc/misc_polar_optimization_apply.c:30-37),
// with the product and the difference rounded separately as CHarm's SIMD
// code does.  Both roundings are monotone, so the skipped orders of a block
// are m0..nmax; returns nmax + 1 when none is skipped or thr < 0.
int polar_first_skip(unsigned long nmax, double thr, double umax)
{
    if (thr < 0.0) return (int)nmax + 1;
    const double c = (double)nmax * umax;
    long m = (long)std::floor(thr + c);
    if (m < 0) m = 0;
    if (m > (long)nmax + 1) m = (long)nmax + 1;
    while (m > 0 && thr < (double)(m - 1) - c) m--;
    while (m <= (long)nmax && !(thr < (double)m - c)) m++;
    return (int)m;
}

void polar_ring_m0(const double *u, size_t nring, unsigned long nmax, double thr, size_t block,
                   int *m0)
{
    for (size_t b0 = 0; b0 < nring; b0 += block) {
        const size_t b1 = std::min(nring, b0 + block);
        double umax = u[b0];
        for (size_t i = b0 + 1; i < b1; i++) umax = std::max(umax, u[i]);
        const int v = polar_first_skip(nmax, thr, umax);
        for (size_t i = b0; i < b1; i++) m0[i] = v;
    }
}

// misc_is_nearly_equal (src/misc/misc_is_nearly_equal.c:16-32)
bool nearly_equal(double a, double b, double eps)
{
    double aa = std::fabs(a), ba = std::fabs(b), d = std::fabs(a - b), s = aa + ba;
    if (a == b) return true;
    if (a == 0.0 || b == 0.0) return d <= eps;
    if (s == aa || s == ba) return d <= eps;
    return d / (aa > ba ? aa : ba) <= eps;
}

// err_inc_level (src/err/err_inc_level.c:12-25): one more level; the
// structure saturates at CHARM_ERR_MAX_LEVEL levels
static void inc_level(charm_err *err)
{
    if (err == nullptr || err->saturated) return;
    err->level += 1;
    if (err->level >= CHARM_ERR_MAX_LEVEL) err->saturated = 1;
}

// err_set (src/err/err_set.c:16-57)
void err_set(charm_err *err, const char *file, size_t line, const char *func, int code,
             const char *msg)
{
    if (err == nullptr || err->saturated) return;
    size_t lvl = err->level;
    std::snprintf(err->file[lvl], CHARM_ERR_MAX_FILE, "%s", file);
    err->line[lvl] = line;
    std::snprintf(err->func[lvl], CHARM_ERR_MAX_FUNC, "%s", func);
    err->code = code;
    std::snprintf(err->msg, CHARM_ERR_MAX_MSG, "%s", msg);
    inc_level(err);
}

// err_propagate (src/err/err_propagate.c): records one more traceback level
void err_propagate(charm_err *err, const char *file, size_t line, const char *func)
{
    if (err == nullptr || err->saturated) return;
    size_t lvl = err->level;
    std::snprintf(err->file[lvl], CHARM_ERR_MAX_FILE, "%s", file);
    err->line[lvl] = line;
    std::snprintf(err->func[lvl], CHARM_ERR_MAX_FUNC, "%s", func);
    inc_level(err);
}

}  // namespace charm_amd

using charm_amd::err_set;

extern "C" {

// ---------------------------------------------------------------------------
// charm_err (src/err/err_init.c:17-130, err_free.c, err_reset.c,
// err_handler.c:15-56, err_isempty.c)
// ---------------------------------------------------------------------------
charm_err *charm_err_init(void)
{
    charm_err *err = (charm_err *)std::malloc(sizeof(charm_err));
    if (!err) return nullptr;
    err->file = (char **)std::calloc(CHARM_ERR_MAX_LEVEL, sizeof(char *));
    err->line = (size_t *)std::calloc(CHARM_ERR_MAX_LEVEL, sizeof(size_t));
    err->func = (char **)std::calloc(CHARM_ERR_MAX_LEVEL, sizeof(char *));
    err->msg = (char *)std::malloc(CHARM_ERR_MAX_MSG);
    bool ok = err->file && err->line && err->func && err->msg;
    for (size_t e = 0; ok && e < CHARM_ERR_MAX_LEVEL; e++) {
        err->file[e] = (char *)std::malloc(CHARM_ERR_MAX_FILE);
        err->func[e] = (char *)std::malloc(CHARM_ERR_MAX_FUNC);
        ok = err->file[e] && err->func[e];
    }
    if (!ok) {
        charm_err_free(err);
        return nullptr;
    }
    err->distributed = 0;
    charm_err_reset(err);
    return err;
}

void charm_err_free(charm_err *err)
{
    if (!err) return;
    if (err->file)
        for (size_t e = 0; e < CHARM_ERR_MAX_LEVEL; e++) std::free(err->file[e]);
    if (err->func)
        for (size_t e = 0; e < CHARM_ERR_MAX_LEVEL; e++) std::free(err->func[e]);
    std::free(err->file);
    std::free(err->line);
    std::free(err->func);
    std::free(err->msg);
    std::free(err);
}

void charm_err_reset(charm_err *err)
{
    if (!err) return;
    err->level = 0;
    for (size_t e = 0; e < CHARM_ERR_MAX_LEVEL; e++) {
        std::memset(err->file[e], 0, CHARM_ERR_MAX_FILE);
        std::memset(err->func[e], 0, CHARM_ERR_MAX_FUNC);
    }
    std::memset(err->line, 0, CHARM_ERR_MAX_LEVEL * sizeof(size_t));
    err->code = CHARM_SUCCESS;
    std::memset(err->msg, 0, CHARM_ERR_MAX_MSG);
    err->saturated = 0;
}

CHARM_AMD_BOOL charm_err_isempty(const charm_err *err)
{
    if (err == nullptr) return 0;
    return err->code == CHARM_SUCCESS;
}

void charm_err_handler(charm_err *err, CHARM_AMD_BOOL terminate)
{
    if (err == nullptr || charm_err_isempty(err)) return;
    std::fprintf(stderr,
                 "\n-----------\nCHarm ERROR\n-----------\n"
                 "Error code: %u                   Traceback (most recent call last)\n\n",
                 (unsigned)err->code);
    for (size_t e = err->level; e-- > 0;)
        std::fprintf(stderr, "    File \"%s\", line: %zu, function: \"%s\"\n\n", err->file[e],
                     err->line[e], err->func[e]);
    if (err->saturated)
        std::fprintf(stderr, "%s",
                     "    Warning: The error structure is saturated.  Most recent function "
                     "calls may therefore not be reported.\n\n");
    std::fprintf(stderr, "Error message: %s\n-----------\n", err->msg);
    if (terminate) std::exit(CHARM_FAILURE);
    charm_err_reset(err);
}

// ---------------------------------------------------------------------------
// charm_shc (src/shc/shc_alloc.c:15-69, shc_init_chunk.c:23-199,
// shc_free.c:14-47)
// ---------------------------------------------------------------------------
charm_shc *charm_shc_init(unsigned long nmax, double mu, double r, double *c, double *s)
{
    if (r <= 0.0) return nullptr;  // misc_check_radius
    charm_shc *shcs = (charm_shc *)std::malloc(sizeof(charm_shc));
    if (!shcs) return nullptr;
    const size_t nmax1 = (size_t)nmax + 1;
    shcs->nmax = nmax;
    shcs->mu = mu;
    shcs->r = r;
    shcs->nc = shcs->ns = ((nmax1 + 1) * nmax1) / 2;
    shcs->owner = 0;
    shcs->distributed = 0;
    shcs->c = (double **)std::malloc(nmax1 * sizeof(double *));
    shcs->s = (double **)std::malloc(nmax1 * sizeof(double *));
    if (!shcs->c || !shcs->s) {
        std::free(shcs->c);
        std::free(shcs->s);
        std::free(shcs);
        return nullptr;
    }
    size_t idx = 0;
    for (unsigned long m = 0; m <= nmax; m++) {
        shcs->c[m] = c + idx;
        shcs->s[m] = s + idx;
        idx += nmax1 - m;
    }
    return shcs;
}

static charm_shc *shc_alloc(unsigned long nmax, double mu, double r, bool zero)
{
    const size_t nmax1 = (size_t)nmax + 1, len = ((nmax1 + 1) * nmax1) / 2;
    double *c = (double *)(zero ? std::calloc(len, sizeof(double)) : std::malloc(len * sizeof(double)));
    double *s = (double *)(zero ? std::calloc(len, sizeof(double)) : std::malloc(len * sizeof(double)));
    if (!c || !s) {
        std::free(c);
        std::free(s);
        return nullptr;
    }
    charm_shc *shcs = charm_shc_init(nmax, mu, r, c, s);
    if (!shcs) {
        std::free(c);
        std::free(s);
        return nullptr;
    }
    shcs->owner = 1;
    return shcs;
}

charm_shc *charm_shc_malloc(unsigned long nmax, double mu, double r)
{
    return shc_alloc(nmax, mu, r, false);
}

charm_shc *charm_shc_calloc(unsigned long nmax, double mu, double r)
{
    return shc_alloc(nmax, mu, r, true);
}

void charm_shc_free(charm_shc *shcs)
{
    if (!shcs) return;
    if (shcs->owner && shcs->c && shcs->s) {
        std::free(shcs->c[0]);
        std::free(shcs->s[0]);
    }
    std::free(shcs->c);
    std::free(shcs->s);
    std::free(shcs);
}

// ---------------------------------------------------------------------------
// charm_point (src/crd/crd_point_alloc.c:18-99, crd_point_init.c:22-81,
// crd_point_free.c:14-31, crd_point_npoint.c:16-26,
// crd_point_check_inputs.c:14-32)
// ---------------------------------------------------------------------------
static bool is_point_type(int type)
{
    return type == CHARM_CRD_POINT_SCATTERED || type == CHARM_CRD_POINT_GRID ||
           type == CHARM_CRD_POINT_GRID_GL || type == CHARM_CRD_POINT_GRID_DH1 ||
           type == CHARM_CRD_POINT_GRID_DH2;
}

static int point_check_inputs(int type, size_t nlat, size_t nlon)
{
    if (!is_point_type(type)) return 1;
    if (type == CHARM_CRD_POINT_SCATTERED && nlat != nlon) return 2;
    return 0;
}

charm_point *charm_crd_point_init(int type, size_t nlat, size_t nlon, double *lat, double *lon,
                                  double *r)
{
    if (point_check_inputs(type, nlat, nlon)) return nullptr;
    charm_point *pnt = (charm_point *)std::malloc(sizeof(charm_point));
    if (!pnt) return nullptr;
    pnt->lat = pnt->lon = pnt->r = pnt->w = nullptr;
    if (nlat > 0) {
        pnt->lat = lat;
        pnt->r = r;
    }
    if (nlon > 0) pnt->lon = lon;
    pnt->nlat = nlat;
    pnt->nlon = nlon;
    pnt->npoint = (type == CHARM_CRD_POINT_SCATTERED) ? nlat : nlat * nlon;
    pnt->type = type;
    pnt->owner = 0;
    pnt->distributed = 0;
    return pnt;
}

static charm_point *point_alloc(int type, size_t nlat, size_t nlon, bool zero)
{
    if (point_check_inputs(type, nlat, nlon)) return nullptr;
    auto alloc = [zero](size_t n) -> double * {
        return (double *)(zero ? std::calloc(n, sizeof(double)) : std::malloc(n * sizeof(double)));
    };
    double *lat = nullptr, *lon = nullptr, *r = nullptr, *w = nullptr;
    charm_point *pnt = nullptr;
    if (nlat > 0) {
        lat = alloc(nlat);
        r = alloc(nlat);
        if (!lat || !r) goto fail;
    }
    if (nlon > 0) {
        lon = alloc(nlon);
        if (!lon) goto fail;
    }
    pnt = charm_crd_point_init(type, nlat, nlon, lat, lon, r);
    if (!pnt) goto fail;
    if (type == CHARM_CRD_POINT_GRID_GL || type == CHARM_CRD_POINT_GRID_DH1 ||
        type == CHARM_CRD_POINT_GRID_DH2) {
        w = alloc(nlat);
        if (!w && nlat > 0) goto fail;
        pnt->w = w;
    }
    pnt->owner = 1;
    return pnt;
fail:
    std::free(lat);
    std::free(lon);
    std::free(r);
    std::free(w);
    std::free(pnt);
    return nullptr;
}

charm_point *charm_crd_point_malloc(int type, size_t nlat, size_t nlon)
{
    return point_alloc(type, nlat, nlon, false);
}

charm_point *charm_crd_point_calloc(int type, size_t nlat, size_t nlon)
{
    return point_alloc(type, nlat, nlon, true);
}

void charm_crd_point_free(charm_point *pnt)
{
    if (!pnt) return;
    if (pnt->owner) {
        std::free(pnt->lat);
        std::free(pnt->lon);
        std::free(pnt->r);
        std::free(pnt->w);
    }
    std::free(pnt);
}

// ---------------------------------------------------------------------------
// Gauss-Legendre grids (src/crd/crd_point_gl.c:15-20,
// crd_point_gl_chunk.c:34-234, crd_point_gl_shape.c:14-24,
// crd_point_quad_l.c:17-20)
// ---------------------------------------------------------------------------
void charm_crd_point_gl_shape(unsigned long nmax, size_t *nlat, size_t *nlon)
{
    const unsigned long L = nmax + 1;
    *nlat = L;
    *nlon = 2 * (size_t)L;
}

charm_point *charm_crd_point_gl(unsigned long nmax, double r)
{
    if (r <= 0.0) return nullptr;
    size_t nlat, nlon;
    charm_crd_point_gl_shape(nmax, &nlat, &nlon);
    charm_point *glg = charm_crd_point_calloc(CHARM_CRD_POINT_GRID_GL, nlat, nlon);
    if (!glg) return nullptr;
    const unsigned long L = nmax + 1;
    const double Lf = (double)L, c1 = Lf + 0.5;
    const double pi = 3.14159265358979323846;
    const size_t nnorth = (nlat + 1) / 2;
    const double thold = charm_glob_threshold;
    const double eps = 2.220446049250313080847263336181640625e-16;
    // Newton iterations on P_L for each northern node; nodes are independent,
    // so they are split over host threads.
    std::vector<int> failed(nnorth, 0);
    auto work = [&](size_t i0, size_t i1) {
        for (size_t i = i0; i < i1; i++) {
            double z = std::cos(pi * ((double)(i + 1) - 0.25) / c1), z1, p1, p2, p3, pp;
            unsigned long it = 0;
            do {
                p1 = 1.0;
                p2 = 0.0;
                for (unsigned long j = 1; j <= L; j++) {
                    p3 = p2;
                    p2 = p1;
                    p1 = ((double)(2 * j - 1) * z * p2 - (double)(j - 1) * p3) / (double)j;
                }
                pp = Lf * (z * p1 - p2) / (z * z - 1.0);
                z1 = z;
                z -= p1 / pp;
                it++;
            } while (std::fabs(z - z1) > eps && it < 1000UL);
            if (it >= 1000UL || !std::isfinite(z)) failed[i] = 1;
            const size_t north = i, south = nlat - 1 - i;
            glg->lat[south] = -std::asin(z);
            glg->w[south] = 2.0 / ((1.0 - z * z) * pp * pp);
            if (charm_amd::nearly_equal(glg->lat[south], 0.0, thold)) continue;
            glg->lat[north] = -glg->lat[south];
            glg->w[north] = glg->w[south];
        }
    };
    size_t nthr = std::thread::hardware_concurrency();
    if (nthr == 0) nthr = 1;
    if (nthr > 16) nthr = 16;
    if (nnorth < 64 || nthr == 1) {
        work(0, nnorth);
    } else {
        std::vector<std::thread> th;
        // interleaved chunks: node cost is uniform, so contiguous ranges suffice
        size_t chunk = (nnorth + nthr - 1) / nthr;
        for (size_t t = 0; t < nthr; t++) {
            size_t i0 = t * chunk, i1 = std::min(nnorth, i0 + chunk);
            if (i0 < i1) th.emplace_back(work, i0, i1);
        }
        for (auto &x : th) x.join();
    }
    for (size_t i = 0; i < nnorth; i++)
        if (failed[i]) {
            charm_crd_point_free(glg);
            return nullptr;
        }
    for (size_t i = 0; i < nlat; i++) glg->r[i] = r;
    const double c = pi / (double)L;
    for (size_t j = 0; j < nlon; j++) glg->lon[j] = c * (double)j;
    return glg;
}

// ---------------------------------------------------------------------------
// Driscoll--Healy grids (src/crd/crd_point_dh1.c, crd_point_dh2.c,
// crd_point_dh_chunk.c:19-150, crd_point_dh_lat_w_chunk.c:16-100,
// crd_point_dh{1,2}_shape.c): 2L latitudes from the north pole (co-latitude
// pi i / 2L, no south pole), 2L (DH1) or 4L (DH2) longitudes, L = nmax + 1
// ---------------------------------------------------------------------------
void charm_crd_point_dh1_shape(unsigned long nmax, size_t *nlat, size_t *nlon)
{
    const size_t L = (size_t)nmax + 1;
    *nlat = 2 * L;
    *nlon = 2 * L;
}

void charm_crd_point_dh2_shape(unsigned long nmax, size_t *nlat, size_t *nlon)
{
    const size_t L = (size_t)nmax + 1;
    *nlat = 2 * L;
    *nlon = 4 * L;
}

static charm_point *dh_grid(unsigned long nmax, double r, int type)
{
    if (r <= 0.0) return nullptr;
    size_t nlat, nlon;
    if (type == CHARM_CRD_POINT_GRID_DH1)
        charm_crd_point_dh1_shape(nmax, &nlat, &nlon);
    else
        charm_crd_point_dh2_shape(nmax, &nlat, &nlon);
    charm_point *dhg = charm_crd_point_calloc(type, nlat, nlon);
    if (!dhg) return nullptr;
    const unsigned long L = nmax + 1;
    const double pi = 3.14159265358979323846;
    const double c = 2.0 / (double)L, pi_2L = pi / (2.0 * (double)L);
    const double thold = charm_glob_threshold;
    // rings 0 (north pole) .. L (equator); ring i > 0 and its mirror 2L - i
    // share the weight w = c sin(th) sum_{k<L} sin((2k+1) th)/(2k+1), the
    // sines by the Chebyshev recurrence
    auto work = [&](size_t i0, size_t i1) {
        for (size_t i = i0; i < i1; i++) {
            const double clti = pi_2L * (double)i;
            dhg->lat[i] = pi / 2.0 - clti;
            const double sclti = std::sin(clti);
            double w;
            if (nmax == 0) {
                w = c * sclti;
            } else {
                double s0 = 0.0, s1 = sclti, s2, acc = s1;
                const double cc2 = 2.0 * std::cos(clti);
                for (unsigned long k = 1; k < L; k++) {
                    s2 = cc2 * s1 - s0;
                    s0 = s1;
                    s1 = s2;
                    s2 = cc2 * s1 - s0;
                    acc += s2 / (double)(2 * k + 1);
                    s0 = s1;
                    s1 = s2;
                }
                w = c * sclti * acc;
            }
            dhg->w[i] = w;
            if (i == 0 || charm_amd::nearly_equal(dhg->lat[i], 0.0, thold)) continue;
            dhg->w[nlat - i] = w;
            dhg->lat[nlat - i] = -dhg->lat[i];
        }
    };
    const size_t nnorth = L + 1;
    size_t nthr = std::thread::hardware_concurrency();
    if (nthr == 0) nthr = 1;
    if (nthr > 16) nthr = 16;
    if (nnorth < 64 || nthr == 1) {
        work(0, nnorth);
    } else {
        // later rings cost the same (L terms each): contiguous ranges
        std::vector<std::thread> th;
        const size_t chunk = (nnorth + nthr - 1) / nthr;
        for (size_t t = 0; t < nthr; t++) {
            const size_t i0 = t * chunk, i1 = std::min(nnorth, i0 + chunk);
            if (i0 < i1) th.emplace_back(work, i0, i1);
        }
        for (auto &x : th) x.join();
    }
    for (size_t i = 0; i < nlat; i++) dhg->r[i] = r;
    const double cl = (type == CHARM_CRD_POINT_GRID_DH1) ? pi / (double)L : pi / (2.0 * (double)L);
    for (size_t j = 0; j < nlon; j++) dhg->lon[j] = cl * (double)j;
    return dhg;
}

charm_point *charm_crd_point_dh1(unsigned long nmax, double r)
{
    return dh_grid(nmax, r, CHARM_CRD_POINT_GRID_DH1);
}

charm_point *charm_crd_point_dh2(unsigned long nmax, double r)
{
    return dh_grid(nmax, r, CHARM_CRD_POINT_GRID_DH2);
}

}  // extern "C"
