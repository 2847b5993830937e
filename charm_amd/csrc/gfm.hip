// Global spectral gravity forward modelling: the potential of the masses
// between a sphere of radius R and a shape r(lat, lon) with a density that is
// constant, lateral or a polynomial in r (src/gfm/gfm.h.in:222-320,
// src/gfm/gfm_sgfm.c with SGFM_GLOBAL).
//
// For each topography power p = pmin..pmax the shape is synthesised on the
// Gauss--Legendre grid of degree p * shape_nmax + max density degree (so the
// p-th power has no aliasing), turned into ((r - R)/R)^p (times R^i and the
// i-th density term), analysed on the same grid, scaled per degree by
//   c binom(n + i + 3, p) 2 / ((2n + 1)(n + i + 3)),   c = 2 pi R^3 / M [rho]
// and added to the potential.  Grid, fields and coefficients stay in HBM for
// the whole loop; only the output (and any exported powers) come back.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <memory>
#include <mutex>
#include <cstring>
#include <string>
#include <vector>

#include "charm_amd.h"
#include "internal.h"

namespace charm_amd {
namespace {

constexpr double PI_G = 3.14159265358979323846264338327950288;

__global__ void gfm_power_kernel(size_t n, const double *__restrict__ shape, double R,
                                 unsigned p, const double *__restrict__ rho, double rpow,
                                 double *__restrict__ out)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double h = (shape[i] - R) / R;
    double v = h;
    for (unsigned k = 2; k <= p; k++) v *= h;
    out[i] = rho ? rho[i] * (v * rpow) : v;
}

// pot (degree npot) += vnm[n] * cs (degree ncs <= npot), pairs (C, S)
__global__ void gfm_add_kernel(int ncs, const double2 *__restrict__ cs, const double *__restrict__ vnm,
                               int npot, double2 *__restrict__ pot)
{
    const int m = blockIdx.y;
    const double2 *src = cs + (size_t)m * (ncs + 1) - ((size_t)m * (m - 1)) / 2 - m;
    double2 *dst = pot + (size_t)m * (npot + 1) - ((size_t)m * (m - 1)) / 2 - m;
    for (int n = m + blockIdx.x * blockDim.x + threadIdx.x; n <= ncs; n += gridDim.x * blockDim.x) {
        const double2 x = src[n];
        const double v = vnm[n];
        double2 y = dst[n];
        y.x += x.x * v;
        y.y += x.y * v;
        dst[n] = y;
    }
}

struct Fail {
    int code;
    std::string msg;
};

void hchk(hipError_t e, const char *what)
{
    if (e != hipSuccess)
        throw Fail{e == hipErrorOutOfMemory ? CHARM_EMEM : CHARM_EFUNCARG,
                   std::string("HIP error \"") + hipGetErrorString(e) + "\" in " + what + "."};
}

struct Dev {
    void *p = nullptr;
    explicit Dev(size_t bytes)
    {
        if (bytes) hchk(hipMalloc(&p, bytes), "hipMalloc");
    }
    ~Dev()
    {
        if (p) (void)hipFree(p);
    }
    Dev(const Dev &) = delete;
    Dev &operator=(const Dev &) = delete;
    double *d() const { return (double *)p; }
};

void err_check(charm_err *e, const char *what)
{
    if (!charm_err_isempty(e)) throw Fail{e->code, std::string(what) + ": " + e->msg};
}

// (C, S) pairs of degree nmax from a charm_shc
std::vector<double> pack(const charm_shc *s, unsigned long nmax)
{
    std::vector<double> v(2 * tri_len(nmax));
    size_t q = 0;
    for (unsigned long m = 0; m <= nmax; m++)
        for (unsigned long n = m; n <= nmax; n++, q++) {
            v[2 * q] = s->c[m][n - m];
            v[2 * q + 1] = (m == 0) ? 0.0 : s->s[m][n - m];
        }
    return v;
}

// src/gfm/gfm_sgfm.c:208-250 (global_vnmc)
void global_vnm(unsigned long nmax, unsigned p, unsigned i, double c, std::vector<double> &v)
{
    v.assign(nmax + 1, 0.0);
    const double pr = (double)p;
    for (unsigned long n = 0; n <= nmax; n++) {
        const double np3 = (double)(n + i + 3);
        double s = np3 / pr;
        for (unsigned k = 1; k < p; k++) s *= (np3 - (double)k) / (pr - (double)k);
        s *= 2.0 / ((2.0 * (double)n + 1.0) * np3);
        v[n] = c * s;
    }
}

int ndigits(unsigned x)
{
    return x > 0 ? (int)std::ceil(std::log10((double)x)) : 1;
}

// src/gfm/gfm_sgfm.c:129-200: <path>-p<p>-i<i>.shcs (bin) or .txt
void export_powers(const std::vector<double> &cs, unsigned long nmax, double mu, double r,
                   const char *path, const char *fmt, unsigned p, unsigned pmax, unsigned i,
                   unsigned imax)
{
    char name[CHARM_ERR_MAX_FILE];
    const bool bin = !std::strcmp(fmt, "bin");
    const int ret = std::snprintf(name, sizeof name, bin ? "%s-p%0*u-i%0*u.shcs" : "%s-p%0*u-i%0*u.txt",
                                  path, ndigits(pmax), p, ndigits(imax), i);
    if (ret <= 0 || (size_t)ret >= sizeof name)
        throw Fail{CHARM_EFUNCARG, "Failed to prepare the output file name string."};
    charm_shc *s = charm_shc_calloc(nmax, mu, r);
    if (!s) throw Fail{CHARM_EMEM, "Failed to allocate the coefficients to export."};
    size_t q = 0;
    for (unsigned long m = 0; m <= nmax; m++)
        for (unsigned long n = m; n <= nmax; n++, q++) {
            s->c[m][n - m] = cs[2 * q];
            s->s[m][n - m] = cs[2 * q + 1];
        }
    charm_err *e = charm_err_init();
    if (bin)
        charm_shc_write_bin(s, nmax, name, e);
    else if (!std::strcmp(fmt, "mtx"))
        charm_shc_write_mtx(s, nmax, "%0.16e", name, e);
    else if (!std::strcmp(fmt, "tbl"))
        charm_shc_write_tbl(s, nmax, "%0.16e", CHARM_SHC_WRITE_N, name, e);
    else
        charm_shc_write_dov(s, nmax, "%0.16e", CHARM_SHC_WRITE_N, name, e);
    charm_shc_free(s);
    const bool ok = charm_err_isempty(e);
    const std::string msg = ok ? "" : std::string(e->msg);
    charm_err_free(e);
    if (!ok) throw Fail{CHARM_EFILEIO, msg};
}

// Gauss--Legendre grids of radius 1 kept on the device across calls (a
// modelling run visits one grid per topography power; repeated runs reuse
// them).  At most GRID_CACHE grids, the least recently used is dropped.
constexpr size_t GRID_CACHE = 8;
struct CachedGrid {
    unsigned long nmax;
    charm_amd_grid *grd;
    size_t nlat, nlon;
    unsigned long use;
};
std::mutex g_grid_mu;
std::vector<CachedGrid> g_grids;
unsigned long g_grid_clock = 0;

// The cached grid of degree nmax (created if needed).  Callers hold
// g_grid_mu for the whole modelling run, so no grid in use is dropped.
CachedGrid unit_gl_grid(unsigned long nmax, charm_err *e)
{
    for (auto &g : g_grids)
        if (g.nmax == nmax) {
            g.use = ++g_grid_clock;
            return g;
        }
    charm_point *gl = charm_crd_point_gl(nmax, 1.0);
    if (!gl) throw Fail{CHARM_EFUNCARG, "Failed to compute the Gauss--Legendre grid."};
    charm_amd_grid *grd = charm_amd_grid_create(gl, e);
    const size_t nlat = gl->nlat, nlon = gl->nlon;
    charm_crd_point_free(gl);
    err_check(e, "grid");
    if (g_grids.size() >= GRID_CACHE) {
        auto old = std::min_element(g_grids.begin(), g_grids.end(),
                                    [](const CachedGrid &a, const CachedGrid &b) {
                                        return a.use < b.use;
                                    });
        charm_amd_grid_free(old->grd);
        g_grids.erase(old);
    }
    g_grids.push_back(CachedGrid{nmax, grd, nlat, nlon, ++g_grid_clock});
    return g_grids.back();
}

bool is_one(double x)
{
    return nearly_equal(x, 1.0, charm_glob_threshold);
}

// The shared driver: density terms i = 0..imax from density_shcs (rho_const
// when density_shcs is null)
void sgfm_global(const charm_shc *shape_shcs, unsigned long shape_nmax, double R,
                 double rho_const, charm_shc **density_shcs, const unsigned long *density_nmax,
                 unsigned density_order, double G, double M, unsigned pmin, unsigned pmax,
                 unsigned long nmax_pot, const char *hro_path, const char *pot_path,
                 const char *fmt, charm_shc *pot, charm_err *err)
{
    char msg[CHARM_ERR_MAX_MSG];
    const bool rho_is_const = density_shcs == nullptr;
    if (!rho_is_const)
        for (unsigned i = 0; i <= density_order; i++) {
            if (density_nmax[i] > density_shcs[i]->nmax) {
                std::snprintf(msg, sizeof msg,
                              "\"density_nmax[%u]\" is \"%lu\", but it cannot be larger than "
                              "\"density_shcs[%u]->nmax = %lu\".",
                              i, density_nmax[i], i, density_shcs[i]->nmax);
                CHARM_AMD_ERR(err, CHARM_EFUNCARG, msg);
                return;
            }
            if (!is_one(density_shcs[i]->mu) || !is_one(density_shcs[i]->r)) {
                std::snprintf(msg, sizeof msg,
                              "\"density_shcs[%u]->mu\" and \"density_shcs[%u]->r\" have to be "
                              "\"1.0\".",
                              i, i);
                CHARM_AMD_ERR(err, CHARM_EFUNCARG, msg);
                return;
            }
        }
    if (shape_nmax > shape_shcs->nmax) {
        std::snprintf(msg, sizeof msg,
                      "\"shape_nmax\" is \"%lu\", but it cannot be larger than "
                      "\"shape_shcs->nmax = %lu\".",
                      shape_nmax, shape_shcs->nmax);
        CHARM_AMD_ERR(err, CHARM_EFUNCARG, msg);
        return;
    }
    if (!is_one(shape_shcs->mu) || !is_one(shape_shcs->r)) {
        CHARM_AMD_ERR(err, CHARM_EFUNCARG,
                      "\"shape_shcs->mu\" and \"shape_shcs->r\" have to be \"1.0\".");
        return;
    }
    if (R <= 0.0) {
        CHARM_AMD_ERR(err, CHARM_EFUNCARG, "\"shape_ref_radius\" has to be positive.");
        return;
    }
    if (nearly_equal(M, 0.0, charm_glob_threshold)) {
        CHARM_AMD_ERR(err, CHARM_EFUNCARG, "\"mass\" cannot be zero.");
        return;
    }
    if (pmin < 1 || pmax < 1) {
        CHARM_AMD_ERR(err, CHARM_EFUNCARG, "Topography power must be positive.");
        return;
    }
    if (pmin > pmax) {
        std::snprintf(msg, sizeof msg,
                      "\"shape_power_min\" is \"%u\", but it cannot be larger than "
                      "\"shape_power_max = %u\".",
                      pmin, pmax);
        CHARM_AMD_ERR(err, CHARM_EFUNCARG, msg);
        return;
    }
    if (pot && nmax_pot > pot->nmax) {
        std::snprintf(msg, sizeof msg,
                      "\"potential_shcs_nmax\" is \"%lu\", but it cannot be larger than "
                      "\"potential_shcs->nmax = %lu\".",
                      nmax_pot, pot->nmax);
        CHARM_AMD_ERR(err, CHARM_EFUNCARG, msg);
        return;
    }
    const bool exp_hro = hro_path != nullptr, exp_pot = pot_path != nullptr;
    if ((exp_hro || exp_pot) &&
        (!fmt || (std::strcmp(fmt, "bin") && std::strcmp(fmt, "mtx") && std::strcmp(fmt, "tbl") &&
                  std::strcmp(fmt, "dov")))) {
        CHARM_AMD_ERR(err, CHARM_EFUNCARG, "Unsupported value of \"shcs_file_format\".");
        return;
    }
    if (pot) {
        std::fill(pot->c[0], pot->c[0] + pot->nc, 0.0);
        std::fill(pot->s[0], pot->s[0] + pot->ns, 0.0);
    }
    const unsigned imax = rho_is_const ? 0 : density_order;
    unsigned long dens_nmax_max = 0;
    if (!rho_is_const)
        for (unsigned i = 0; i <= imax; i++) dens_nmax_max = std::max(dens_nmax_max, density_nmax[i]);
    double c = (2.0 * PI_G * std::pow(R, 3)) / M;
    if (rho_is_const) c *= rho_const;
    const double gm = G * M;

    // one modelling run at a time per process: the runs share the grid cache
    std::lock_guard<std::mutex> run_lock(g_grid_mu);
    hipStream_t st = nullptr;
    charm_err *e = charm_err_init();
    try {
        hchk(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "hipStreamCreateWithFlags");
        // shape and density coefficients, and the potential, resident in HBM
        const std::vector<double> hs = pack(shape_shcs, shape_nmax);
        Dev dshape(hs.size() * sizeof(double));
        hchk(hipMemcpyAsync(dshape.p, hs.data(), hs.size() * sizeof(double),
                            hipMemcpyHostToDevice, st), "hipMemcpyAsync");
        std::vector<std::unique_ptr<Dev>> ddens;
        for (unsigned i = 0; !rho_is_const && i <= imax; i++) {
            const std::vector<double> hd = pack(density_shcs[i], density_nmax[i]);
            ddens.emplace_back(new Dev(hd.size() * sizeof(double)));
            hchk(hipMemcpy(ddens.back()->p, hd.data(), hd.size() * sizeof(double),
                           hipMemcpyHostToDevice), "hipMemcpy");
        }
        const size_t npot_pairs = tri_len(nmax_pot);
        Dev dpot(2 * npot_pairs * sizeof(double));
        hchk(hipMemsetAsync(dpot.p, 0, 2 * npot_pairs * sizeof(double), st), "hipMemsetAsync");
        std::vector<double> vnm, hcs;
        // buffers for the largest grid (the last power), reused by every power
        const unsigned long nmaxp_top = (unsigned long)pmax * shape_nmax + dens_nmax_max;
        const unsigned long thr_top = std::min(nmaxp_top, nmax_pot);
        const size_t npt_top = (size_t)(nmaxp_top + 1) * (2 * nmaxp_top + 2);
        Dev dsh(npt_top * sizeof(double)), dfield(npt_top * sizeof(double));
        Dev dcs(2 * tri_len(thr_top) * sizeof(double)), dv((thr_top + 1) * sizeof(double));
        for (unsigned p = pmin; p <= pmax; p++) {
            const unsigned long nmaxp = (unsigned long)p * shape_nmax + dens_nmax_max;
            const unsigned long thr = std::min(nmaxp, nmax_pot);
            const CachedGrid cg = unit_gl_grid(nmaxp, e);
            charm_amd_grid *grd = cg.grd;
            const size_t npt = cg.nlat * cg.nlon;
            charm_amd_shs_grid_dev(grd, shape_nmax, 1.0, 1.0, dshape.d(), dsh.d(), st, e);
            err_check(e, "shape synthesis");
            double rpow = 1.0;
            for (unsigned i = 0; i <= imax; i++) {
                const double *rho = nullptr;
                if (!rho_is_const) {
                    charm_amd_shs_grid_dev(grd, density_nmax[i], 1.0, 1.0, ddens[i]->d(),
                                           dfield.d(), st, e);
                    err_check(e, "density synthesis");
                    rho = dfield.d();
                }
                hipLaunchKernelGGL(gfm_power_kernel, dim3((unsigned)((npt + 255) / 256)), dim3(256),
                                   0, st, npt, (const double *)dsh.d(), R, p, rho, rpow,
                                   dfield.d());
                hchk(hipGetLastError(), "gfm_power_kernel");
                rpow *= R;
                charm_amd_sha_grid_dev(grd, thr, 1.0, 1.0, dfield.d(), dcs.d(), st, e);
                err_check(e, "analysis");
                if (exp_hro || exp_pot) {
                    hcs.resize(2 * tri_len(thr));
                    hchk(hipMemcpyAsync(hcs.data(), dcs.p, hcs.size() * sizeof(double),
                                        hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
                    hchk(hipStreamSynchronize(st), "hipStreamSynchronize");
                }
                if (exp_hro) export_powers(hcs, thr, 1.0, 1.0, hro_path, fmt, p, pmax, i, imax);
                global_vnm(thr, p, i, c, vnm);
                if (exp_pot) {
                    std::vector<double> w(hcs.size());
                    size_t q = 0;
                    for (unsigned long m = 0; m <= thr; m++)
                        for (unsigned long n = m; n <= thr; n++, q++) {
                            w[2 * q] = hcs[2 * q] * vnm[n];
                            w[2 * q + 1] = hcs[2 * q + 1] * vnm[n];
                        }
                    export_powers(w, thr, gm, R, pot_path, fmt, p, pmax, i, imax);
                }
                if (pot) {
                    hchk(hipMemcpyAsync(dv.p, vnm.data(), (thr + 1) * sizeof(double),
                                        hipMemcpyHostToDevice, st), "hipMemcpyAsync");
                    hipLaunchKernelGGL(gfm_add_kernel, dim3((unsigned)((thr + 256) / 256),
                                                            (unsigned)thr + 1),
                                       dim3(256), 0, st, (int)thr, (const double2 *)dcs.p,
                                       (const double *)dv.d(), (int)nmax_pot, (double2 *)dpot.p);
                    hchk(hipGetLastError(), "gfm_add_kernel");
                    // vnm is reused by the next term: the copy must be done
                    hchk(hipStreamSynchronize(st), "hipStreamSynchronize");
                }
            }
        }
        if (pot) {
            std::vector<double> hp(2 * npot_pairs);
            hchk(hipMemcpyAsync(hp.data(), dpot.p, hp.size() * sizeof(double),
                                hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
            hchk(hipStreamSynchronize(st), "hipStreamSynchronize");
            size_t q = 0;
            for (unsigned long m = 0; m <= nmax_pot; m++)
                for (unsigned long n = m; n <= nmax_pot; n++, q++) {
                    pot->c[m][n - m] = hp[2 * q];
                    pot->s[m][n - m] = hp[2 * q + 1];
                }
            pot->mu = gm;
            pot->r = R;
        }
    } catch (const Fail &f) {
        CHARM_AMD_ERR(err, f.code, f.msg.c_str());
    }
    if (st) {
        (void)hipStreamSynchronize(st);
        (void)hipStreamDestroy(st);
    }
    charm_err_free(e);
}

}  // namespace
}  // namespace charm_amd

using namespace charm_amd;

extern "C" {

// src/gfm/gfm.h.in:222-239 (src/gfm/gfm_global_density_3d.c)
void charm_gfm_global_density_3d(const charm_shc *shape_shcs, unsigned long shape_nmax,
                                 double shape_radius_ref, charm_shc **density_shcs,
                                 unsigned long *density_nmax, unsigned density_order,
                                 double grav_const, double mass, unsigned shape_power_min,
                                 unsigned shape_power_max, unsigned long potential_shcs_nmax,
                                 const char *shape_density_shcs_path,
                                 const char *potential_shcs_path, const char *shcs_file_format,
                                 charm_shc *potential_shcs, charm_err *err)
{
    sgfm_global(shape_shcs, shape_nmax, shape_radius_ref, 0.0, density_shcs, density_nmax,
                density_order, grav_const, mass, shape_power_min, shape_power_max,
                potential_shcs_nmax, shape_density_shcs_path, potential_shcs_path,
                shcs_file_format, potential_shcs, err);
    if (!charm_err_isempty(err)) err_propagate(err, __FILE__, __LINE__, __func__);
}

// src/gfm/gfm.h.in:255-270 (src/gfm/gfm_global_density_lateral.c): the 3D
// case with one density term
void charm_gfm_global_density_lateral(const charm_shc *shape_shcs, unsigned long shape_nmax,
                                      double shape_radius_ref, charm_shc *density_shcs,
                                      unsigned long density_nmax, double grav_const, double mass,
                                      unsigned shape_power_min, unsigned shape_power_max,
                                      unsigned long potential_shcs_nmax,
                                      const char *shape_density_shcs_path,
                                      const char *potential_shcs_path,
                                      const char *shcs_file_format, charm_shc *potential_shcs,
                                      charm_err *err)
{
    unsigned long dn = density_nmax;
    charm_gfm_global_density_3d(shape_shcs, shape_nmax, shape_radius_ref, &density_shcs, &dn, 0,
                                grav_const, mass, shape_power_min, shape_power_max,
                                potential_shcs_nmax, shape_density_shcs_path, potential_shcs_path,
                                shcs_file_format, potential_shcs, err);
}

// src/gfm/gfm.h.in:305-319 (src/gfm/gfm_global_density_const.c)
void charm_gfm_global_density_const(const charm_shc *shape_shcs, unsigned long shape_nmax,
                                    double shape_radius_ref, double density, double grav_const,
                                    double mass, unsigned shape_power_min,
                                    unsigned shape_power_max, unsigned long potential_shcs_nmax,
                                    const char *shape_shcs_path, const char *potential_shcs_path,
                                    const char *shcs_file_format, charm_shc *potential_shcs,
                                    charm_err *err)
{
    sgfm_global(shape_shcs, shape_nmax, shape_radius_ref, density, nullptr, nullptr, 0,
                grav_const, mass, shape_power_min, shape_power_max, potential_shcs_nmax,
                shape_shcs_path, potential_shcs_path, shcs_file_format, potential_shcs, err);
    if (!charm_err_isempty(err)) err_propagate(err, __FILE__, __LINE__, __func__);
}

}  // extern "C"
