/*
 * CPU oracle for CHarm's point spherical-harmonic synthesis and analysis.
 *
 * TEST INFRASTRUCTURE ONLY.  Only the tests, the smoke check and the CPU
 * baseline leg of the benchmark may load this code, and only as the checker.
 * The product (libcharm_amd.so) never links, loads or calls it.
 *
 * This file is a plain, scalar restatement of the reference algorithm, written
 * from scratch.  One source builds two libraries:
 *   - liborc_d.so : REAL = double      (restates CHarm built in double)
 *   - liborc_q.so : REAL = __float128  (restates CHarm built in quad, which
 *                                       is the higher-precision oracle)
 * It is compiled with -ffp-contract=off, so every multiply and add rounds
 * separately, in the operation order of the reference.
 *
 * Array arguments.  Every floating-point array of n values is passed as
 * NCOMP*n doubles: NCOMP = 1 for the double build; NCOMP = 3 for the quad
 * build, where value i is hi[i] + mid[i] + lo[i] with hi = x[i], mid = x[n+i],
 * lo = x[2n+i] (a non-overlapping triple-double, exact for any __float128).
 * Outputs use the same layout.  orc_ncomp() returns NCOMP.
 *
 * Reference files followed (paths relative to the CHarm 0.4.9 tree):
 *   src/leg/leg_func_r_ri.c:15-30, leg_func_dm.c:14-23,
 *   leg_func_anm_bnm.c:14-65, leg_func_prepare.c:18-69,
 *   leg_func_xnum.h:29-31,299-443, src/prec.h:248-326 (quad: :160-240)
 *   src/shs/shs_point_kernel.c:745-1623 (DR = DLAT = DLON = 0 path)
 *   src/shs/shs_point_grd.c:97-1076, shs_grd_fft_lc.c:54-171,
 *   shs_grd_fft.c:23-136, shs_grd_point_fft_check.c:14-40, shs_rpows.c:16-37,
 *   shs_r_eq_rref.c:17-30, shs_point_sctr.c:86-728, shs_sctr_mulc.c:17-61
 *   src/sha/sha_point.c:124-1125, src/shc/shc_rescale.c:24-90
 *   src/crd/crd_point_gl_chunk.c:34-234, crd_grd_check_symm.c:21-72,
 *   crd_point_quad_equator.c:22-53
 *   src/misc/misc_is_nearly_equal.c:16-32, src/glob/glob.c:16-23
 *   src/misc/misc_polar_optimization_threshold.c:16-20,
 *   misc_polar_optimization_apply.c:16-41 (polar optimisation, off by
 *   default; CHarm tests the criterion jointly for a block of SIMD_SIZE x
 *   BLOCK_S rings or points in synthesis, SIMD_SIZE x BLOCK_A rings in
 *   analysis, and skips the order for the whole block when every member
 *   passes: shs_point_grd.c:594-600,860-862, shs_point_sctr.c:291-320,518-520,
 *   sha_point.c:716-718.  The block sizes are set with orc_set_polar; the
 *   AVX-512 double build has 32 and 64.)
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>

#ifdef ORACLE_QUAD
#include <quadmath.h>
typedef __float128 real;
#define NCOMP 3
#define RSIN sinq
#define RCOS cosq
#define RASIN asinq
#define RFABS fabsq
#define RSQRT sqrtq
#define RPI M_PIq
#define REPS FLT128_EPSILON
/* X-number radix 2^16000 in quad (src/prec.h:160-240) */
#define XEXP 16000
#else
typedef double real;
#define NCOMP 1
#define RSIN sin
#define RCOS cos
#define RASIN asin
#define RFABS fabs
#define RSQRT sqrt
#define RPI 3.14159265358979323846
#define REPS DBL_EPSILON
/* X-number radix 2^960 in double (src/prec.h:283-320) */
#define XEXP 960
#endif

static real BIG, BIGI, BIGS, BIGSI;
static int consts_ready = 0;

static real pow2i(int e)
{
    real x = 1;
    if (e >= 0)
        for (int i = 0; i < e; i++) x *= 2;
    else
        for (int i = 0; i < -e; i++) x /= 2;
    return x;
}

static void init_consts(void)
{
    if (consts_ready) return;
    BIG = pow2i(XEXP);
    BIGI = pow2i(-XEXP);
    BIGS = pow2i(XEXP / 2);
    BIGSI = pow2i(-XEXP / 2);
    consts_ready = 1;
}

/* ------------------------------------------------------------------------ */
/* Array I/O in the NCOMP layout                                            */
/* ------------------------------------------------------------------------ */
static inline real ld(const double *x, size_t n, size_t i)
{
#ifdef ORACLE_QUAD
    return ((real)x[i] + (real)x[n + i]) + (real)x[2 * n + i];
#else
    (void)n;
    return x[i];
#endif
}

static inline void st(double *y, size_t n, size_t i, real v)
{
#ifdef ORACLE_QUAD
    double hi = (double)v;
    real r1 = v - (real)hi;
    double mid = (double)r1;
    double lo = (double)(r1 - (real)mid);
    y[i] = hi;
    y[n + i] = mid;
    y[2 * n + i] = lo;
#else
    (void)n;
    y[i] = v;
#endif
}

static real *ld_all(const double *x, size_t n)
{
    real *v = (real *)malloc((n ? n : 1) * sizeof(real));
    for (size_t i = 0; i < n; i++) v[i] = ld(x, n, i);
    return v;
}

/* misc_is_nearly_equal (src/misc/misc_is_nearly_equal.c:16-32) */
static int nearly_equal(real a, real b, real eps)
{
    real aa = RFABS(a), ba = RFABS(b), d = RFABS(a - b), s = aa + ba;
    if (a == b) return 1;
    if (a == 0 || b == 0) return d <= eps;
    if (s == aa || s == ba) return d <= eps;
    return d / (aa > ba ? aa : ba) <= eps;
}

#define GLOB_THRESHOLD (((real)100) * REPS) /* src/glob/glob.c:16 */
#define GLOB_THRESHOLD2 (((real)100000) * REPS) /* src/glob/glob.c:17 */

/* ------------------------------------------------------------------------ */
/* Legendre recurrence coefficients                                         */
/* ------------------------------------------------------------------------ */
/* r[k] = sqrt(k), ri[k] = 1/sqrt(k), k <= 2*nmax+3 (leg_func_r_ri.c:15-30);
 * dm[n] = sqrt(2n+3)/sqrt(2n+2) (leg_func_dm.c:14-23) */
static void leg_tables(unsigned long nmax, real *r, real *ri, real *dm)
{
    r[0] = ri[0] = 0;
    for (unsigned long k = 1; k <= 2 * nmax + 3; k++) {
        real w = RSQRT((real)k);
        r[k] = w;
        ri[k] = ((real)1) / w;
    }
    dm[0] = 0;
    for (unsigned long n = 1; n <= nmax; n++) dm[n] = r[2 * n + 3] * ri[2 * n + 2];
}

/* a_nm, b_nm for one order m (leg_func_anm_bnm.c:14-65) */
static void leg_anm_bnm(unsigned long nmax, unsigned long m, const real *r,
                        const real *ri, real *anm, real *bnm)
{
    anm[m] = bnm[m] = 0;
    if (m == nmax) return;
    anm[m + 1] = r[2 * m + 3];
    bnm[m + 1] = 0;
    for (unsigned long n = m + 2; n <= nmax; n++) {
        real w = r[2 * n + 1] * ri[n - m] * ri[n + m];
        anm[n] = r[2 * n - 1] * w;
        bnm[n] = r[n - m - 1] * r[n + m - 1] * ri[2 * n - 3] * w;
    }
}

/* Renormalise an X-number into [BIGSI, BIGS) (leg_func_xnum.h:318-330) */
#define XNORM(x, ix)                                  \
    do {                                              \
        real w_ = RFABS(x);                           \
        if (w_ >= BIGS) { (x) *= BIGI; (ix) += 1; }   \
        else if (w_ < BIGSI) { (x) *= BIG; (ix) -= 1; } \
    } while (0)

/* X-number to an ordinary floating-point number (leg_func_xnum.h:304-314) */
static inline real x2f(real x, int64_t ix)
{
    if (ix == 0) return x;
    if (ix < -1) return 0;
    if (ix < 0) return x * BIGI;
    return x * BIG;
}

/* Sectorial seeds P_mm, m = 1..nmax, as X-numbers (leg_func_prepare.c:18-69).
 * ps[m-1], ips[m-1] hold P_mm. */
static void leg_seed(real u, const real *dm, unsigned long nmax, real *ps, int64_t *ips)
{
    if (nmax == 0) return;
    real x = RSQRT((real)3) * u;
    int64_t ix = 0;
    ps[0] = x;
    ips[0] = ix;
    for (unsigned long n = 1; n < nmax; n++) {
        x = (dm[n] * u) * x;
        XNORM(x, ix);
        ps[n] = x;
        ips[n] = ix;
    }
}

/* ------------------------------------------------------------------------ */
/* Lumped coefficients for one ring and one order                           */
/* ------------------------------------------------------------------------ */
/* Follows shs_point_kernel.c:1014-1623 for DR = DLAT = DLON = 0, one lane.
 * Cm, Sm point at C_{m,m}, S_{m,m} (column m of the triangle).  rp[k] and,
 * for the mirror ring, rp2[k] hold (R/r)^k; the n-th term is weighted by
 * rp[n+1].  a2/b2 receive the mirror-ring sums when rp2 != NULL, using the
 * parity P_nm(-t) = (-1)^(n+m) P_nm(t).
 *
 * Per ring, X-numbers are used from the sectorial term until both the
 * current and the previous tesseral terms have a zero exponent; from then on
 * the ordinary recurrence is used.  This is the per-lane behaviour of the
 * reference, which in its blocked form only delays the switch of the whole
 * block, not the arithmetic of any one lane. */
static void lumped_ring(unsigned long nmax, unsigned long m, const real *Cm,
                        const real *Sm, const real *anm, const real *bnm,
                        real t, const real *ps, const int64_t *ips,
                        const real *rp, const real *rp2,
                        real *pa, real *pb, real *pa2, real *pb2)
{
    real a, b, a2 = 0, b2 = 0;
    const int symm = (rp2 != NULL);
    size_t idx = 0;

    if (m == 0) {
        real p0 = 1, p1, p2;
        a = rp[1] * (p0 * Cm[idx]);
        b = 0;
        if (symm) { a2 = a2 + rp2[1] * (p0 * Cm[idx]); b2 = 0; }
        idx++;
        if (nmax >= 1) {
            p1 = RSQRT((real)3) * t;
            a = a + rp[2] * (p1 * Cm[idx]);
            if (symm) a2 = a2 - rp2[2] * (p1 * Cm[idx]);
            idx++;
            int npm_even = 1;
            for (unsigned long n = 2; n <= nmax; n++, npm_even = !npm_even) {
                p2 = (anm[n] * t) * p1 - bnm[n] * p0;
                real lc = p2 * Cm[idx];
                a = a + rp[n + 1] * lc;
                p0 = p1;
                p1 = p2;
                if (symm) {
                    if (npm_even) a2 = a2 + rp2[n + 1] * lc;
                    else          a2 = a2 - rp2[n + 1] * lc;
                }
                idx++;
            }
        }
    } else {
        real x = ps[m - 1], y = 0, z;
        int64_t ix = ips[m - 1], iy = 0, iz;
        real p = x2f(x, ix);
        real lc = p * Cm[idx], ls = p * Sm[idx];
        a = rp[m + 1] * lc;
        b = rp[m + 1] * ls;
        if (symm) { a2 = a2 + rp2[m + 1] * lc; b2 = b2 + rp2[m + 1] * ls; }
        idx++;
        if (m < nmax) {
            /* semi-sectorial (leg_func_xnum.h:336-363) */
            y = x; iy = ix;
            x = (anm[m + 1] * t) * y;
            ix = iy;
            XNORM(x, ix);
            p = x2f(x, ix);
            lc = p * Cm[idx]; ls = p * Sm[idx];
            a = a + rp[m + 2] * lc;
            b = b + rp[m + 2] * ls;
            if (symm) { a2 = a2 - rp2[m + 2] * lc; b2 = b2 - rp2[m + 2] * ls; }
            idx++;
            int ds = 0, npm_even = 1;
            for (unsigned long n = m + 2; n <= nmax; n++, npm_even = !npm_even) {
                if (ds) {
                    p = (anm[n] * t) * x - bnm[n] * y;
                    y = x;
                    x = p;
                } else {
                    /* tesseral X-step (leg_func_xnum.h:366-443) */
                    int64_t ixy = ix - iy;
                    if (ixy == 0) { z = (anm[n] * t) * x - bnm[n] * y; iz = ix; }
                    else if (ixy == 1) { z = (anm[n] * t) * x - bnm[n] * (y * BIGI); iz = ix; }
                    else if (ixy == -1) { z = (anm[n] * t) * (x * BIGI) - bnm[n] * y; iz = iy; }
                    else if (ixy > 1) { z = (anm[n] * t) * x; iz = ix; }
                    else { z = -bnm[n] * y; iz = iy; }
                    XNORM(z, iz);
                    y = x; iy = ix;
                    x = z; ix = iz;
                    p = x2f(z, iz);
                    if (ix == 0 && iy == 0) ds = 1;
                }
                lc = p * Cm[idx]; ls = p * Sm[idx];
                a = a + rp[n + 1] * lc;
                b = b + rp[n + 1] * ls;
                if (symm) {
                    if (npm_even) { a2 = a2 + rp2[n + 1] * lc; b2 = b2 + rp2[n + 1] * ls; }
                    else          { a2 = a2 - rp2[n + 1] * lc; b2 = b2 - rp2[n + 1] * ls; }
                }
                idx++;
            }
        }
    }
    *pa = a; *pb = b;
    if (symm) { *pa2 = a2; *pb2 = b2; }
}

/* (R/r)^k, k = 0..kmax, by repeated products (shs_rpows.c:16-37).  With
 * unit = 1 all entries are 1 (shs_point_grd.c:542-556). */
static void rpows(real R, real r, int unit, unsigned long kmax, real *rp)
{
    rp[0] = 1;
    if (unit) { for (unsigned long k = 1; k <= kmax; k++) rp[k] = 1; return; }
    real ratio = R / r;
    for (unsigned long k = 1; k <= kmax; k++) rp[k] = rp[k - 1] * ratio;
}

static size_t tri_len(unsigned long nmax)
{
    return ((size_t)(nmax + 1) * (size_t)(nmax + 2)) / 2;
}

/* offset of column m in the m-major triangle of a model of degree nmax_shc
 * (shc_init_chunk.c:139-150) */
static size_t col_off(unsigned long nmax_shc, unsigned long m)
{
    return (size_t)m * (nmax_shc + 1) - ((size_t)m * (m - 1)) / 2;
}

/* ------------------------------------------------------------------------ */
/* Discrete Fourier sums along a parallel                                   */
/* ------------------------------------------------------------------------ */
/* Unnormalised c2r of X_0 = a_0, X_m = (a_m - i b_m)/2 (m = 1..nmax) on nlon
 * equispaced longitudes, as FFTW computes it (shs_grd_fft_lc.c:24-32,
 * shs_grd_fft.c:110-112): out_j = a_0 + sum_m (a_m cos + b_m sin)(2 pi m j /
 * nlon).  Evaluated directly with exactly reduced angles. */
static void dft_c2r_direct(unsigned long nmax, const real *a, const real *b,
                           size_t nlon, const real *ct, const real *stb, real *out)
{
    for (size_t j = 0; j < nlon; j++) {
        real s = a[0];
        size_t k = 0;
        for (unsigned long m = 1; m <= nmax; m++) {
            k += j;
            if (k >= nlon) k %= nlon;
            s = s + (a[m] * ct[k] + b[m] * stb[k]);
        }
        out[j] = s;
    }
}

/* Forward r2c, returning Re and -Im of F_k = sum_j f_j e^{-2 pi i jk/n} for
 * k = 0..kmax (sha_point.c:520-529 before the weighting). */
static void dft_r2c_direct(size_t nlon, const real *f, unsigned long kmax,
                           const real *ct, const real *stb, real *re, real *mim)
{
    for (unsigned long k = 0; k <= kmax; k++) {
        real sr = 0, si = 0;
        size_t idx = 0;
        for (size_t j = 0; j < nlon; j++) {
            sr = sr + f[j] * ct[idx];
            si = si + f[j] * stb[idx];
            idx += k;
            if (idx >= nlon) idx %= nlon;
        }
        re[k] = sr;
        mim[k] = si;
    }
}

static void twiddles(size_t nlon, real **ct, real **stb)
{
    *ct = (real *)malloc(nlon * sizeof(real));
    *stb = (real *)malloc(nlon * sizeof(real));
    for (size_t k = 0; k < nlon; k++) {
        real ang = ((real)2 * RPI) * (real)k / (real)nlon;
        (*ct)[k] = RCOS(ang);
        (*stb)[k] = RSIN(ang);
    }
}

/* FFT applicability (shs_grd_point_fft_check.c:14-40) */
static int fft_ok(size_t nlon, const real *lon, unsigned long nmax)
{
    if (nlon == 0) return 0;
    real dl = (nlon > 1) ? lon[1] - lon[0] : 0;
    return ((nlon - 1) / 2 >= nmax) && nearly_equal(lon[0], 0, GLOB_THRESHOLD) &&
           nearly_equal(lon[nlon - 1] + dl, (real)2 * RPI, GLOB_THRESHOLD);
}

/* ------------------------------------------------------------------------ */
/* Public oracle API                                                        */
/* ------------------------------------------------------------------------ */
/* Polar optimisation (glob_polar_optimization_a1/a2): order m is skipped for
 * a block of rings or points when every member, cos(lat) = u, passes
 * m - nmax u > max(a1, a2 nmax); a2 < 0 turns it off
 * (misc_polar_optimization_threshold.c:16-20,
 * misc_polar_optimization_apply.c:16-41).  Blocks hold bs_s consecutive rings
 * or points in synthesis and bs_a rings in analysis (SIMD_SIZE times
 * BLOCK_S / BLOCK_A of the CHarm build, src/simd/simd.h:298-304,609-611). */
static unsigned long polar_a1 = 100;
static double polar_a2 = -1.0;
static size_t polar_bs_s = 32, polar_bs_a = 64;

void orc_set_polar(unsigned long a1, double a2, size_t bs_s, size_t bs_a)
{
    polar_a1 = a1;
    polar_a2 = a2;
    polar_bs_s = bs_s ? bs_s : 1;
    polar_bs_a = bs_a ? bs_a : 1;
}

/* the test of one member, cos(lat) = u, for order m */
static int polar_pass(unsigned long m, unsigned long nmax, real u)
{
    if (polar_a2 < 0.0) return 0;
    real thr = (real)polar_a1;
    real t2 = (real)polar_a2 * (real)nmax;
    if (t2 > thr) thr = t2;
    return thr < (real)m - (real)nmax * u;
}

/* order m skipped for the block of cos(lat) values u[0..n) */
static int polar_skip_block(unsigned long m, unsigned long nmax, const real *u, size_t n)
{
    for (size_t i = 0; i < n; i++)
        if (!polar_pass(m, nmax, u[i])) return 0;
    return n > 0;
}

/* misc_arr_chck_symm (src/misc/misc_arr_chck_symm.c:16-52) about 0 */
static int arr_symm(const real *a, size_t n, real eps);

/* The polar block of each row of a grid: rb[row] = first ring of the block,
 * rn[row] = its length, over the rings CHarm evaluates directly.  On a
 * symmetric grid these are the first nlatdo = (n + 1 - even) / 2 rings, with
 * n = nlat + 1 on Driscoll-Healy grids (the missing south pole) and n = nlat
 * otherwise; a mirror row n - 1 - i takes ring i's block.  Otherwise every
 * ring is evaluated directly (shs_point_grd.c:116-197,594-600,
 * crd_grd_check_symm.c:21-72).  type: 0 custom grid (symmetric when
 * misc_arr_chck_symm passes at glob_threshold2), 1 or 3 Gauss-Legendre,
 * 4/5 Driscoll-Healy. */
static void polar_rows(int type, size_t nlat, const real *lat, size_t bs, size_t *rb, size_t *rn);

static int arr_symm(const real *a, size_t n, real eps)
{
    for (size_t i = 0; i < n / 2; i++)
        if (!nearly_equal(a[i] - 0, 0 - a[n - 1 - i], eps)) return 0;
    if (n % 2 && !nearly_equal(0, a[n / 2], eps)) return 0;
    return 1;
}

static void polar_rows(int type, size_t nlat, const real *lat, size_t bs, size_t *rb, size_t *rn)
{
    const int dh = type == 4 || type == 5;
    const size_t n = dh ? nlat + 1 : nlat;
    int symm;
    if (dh) symm = 1;
    else if (nlat == 1) symm = 0;
    else if (type == 1 || type == 3) symm = 1;
    else symm = arr_symm(lat, nlat, GLOB_THRESHOLD2);
    const int even = !(n % 2);
    const size_t nlatdo = symm ? (n + 1 - even) / 2 : n;
    for (size_t i = 0; i < nlatdo; i++) {
        const size_t b0 = (i / bs) * bs;
        const size_t len = (b0 + bs < nlatdo ? b0 + bs : nlatdo) - b0;
        rb[i] = b0;
        rn[i] = len;
        if (symm) {
            const size_t j = n - 1 - i;
            if (j < nlat && j != i) { rb[j] = b0; rn[j] = len; }
        }
    }
}

int orc_ncomp(void) { return NCOMP; }

/* Gauss-Legendre grid (crd_point_gl_chunk.c:105-229).  lat, w: nlat = nmax+1
 * values; lon: nlon = 2(nmax+1) values.  Returns 0 on success. */
int orc_gl_grid(unsigned long nmax, double *lat_o, double *w_o, double *lon_o)
{
    init_consts();
    const unsigned long L = nmax + 1;
    const size_t nlat = L, nlon = 2 * L;
    const size_t nnorth = (nlat + 1) / 2;
    const real Lf = (real)L, c1 = Lf + (real)0.5;
    real *lat = (real *)calloc(nlat, sizeof(real));
    real *w = (real *)calloc(nlat, sizeof(real));
    int err = 0;
    for (size_t i = 0; i < nnorth; i++) {
        real z = RCOS(RPI * ((real)(i + 1) - (real)0.25) / c1), z1, p1, p2, p3, pp;
        unsigned long it = 0;
        do {
            p1 = 1;
            p2 = 0;
            for (unsigned long j = 1; j <= L; j++) {
                p3 = p2;
                p2 = p1;
                p1 = ((real)(2 * j - 1) * z * p2 - (real)(j - 1) * p3) / (real)j;
            }
            pp = Lf * (z * p1 - p2) / (z * z - (real)1);
            z1 = z;
            z -= p1 / pp;
            it++;
        } while (RFABS(z - z1) > REPS && it < 1000UL);
        if (it >= 1000UL) err = 1;
        size_t north = i, south = nlat - 1 - i;
        lat[south] = -RASIN(z);
        w[south] = (real)2 / (((real)1 - z * z) * pp * pp);
        if (nearly_equal(lat[south], 0, GLOB_THRESHOLD)) continue;
        lat[north] = -lat[south];
        w[north] = w[south];
    }
    for (size_t i = 0; i < nlat; i++) { st(lat_o, nlat, i, lat[i]); st(w_o, nlat, i, w[i]); }
    const real c = RPI / (real)L;
    for (size_t j = 0; j < nlon; j++) st(lon_o, nlon, j, c * (real)j);
    free(lat);
    free(w);
    return err;
}

/* Driscoll--Healy grids (crd_point_dh_chunk.c:19-150,
 * crd_point_dh_lat_w_chunk.c:16-100, crd_point_dh{1,2}_shape.c): type 4 = DH1
 * (nlon = 2L), 5 = DH2 (nlon = 4L); nlat = 2L, L = nmax+1, co-latitudes
 * pi i/(2L) from the north pole (no south pole).  Weights
 * w = (2/L) sin(th) sum_{k<L} sin((2k+1) th)/(2k+1), the sines by the
 * Chebyshev recurrence in CHarm's order; rows i and 2L-i share them.
 * Returns 0 on success, 1 for an unknown type. */
int orc_dh_grid(unsigned long nmax, int type, double *lat_o, double *w_o, double *lon_o)
{
    init_consts();
    if (type != 4 && type != 5) return 1;
    const unsigned long L = nmax + 1;
    const size_t nlat = 2 * L, nlon = (type == 4) ? 2 * L : 4 * L;
    const real Lf = (real)L, c = (real)2 / Lf, pi_2L = RPI / ((real)2 * Lf);
    real *lat = (real *)calloc(nlat, sizeof(real));
    real *w = (real *)calloc(nlat, sizeof(real));
    for (size_t i = 0; i <= L; i++) {
        const real clti = pi_2L * (real)i;
        lat[i] = RPI / 2 - clti;
        const real sclti = RSIN(clti);
        if (nmax == 0) {
            w[i] = c * sclti;
        } else {
            real s0 = 0, s1 = sclti, s2, cc2 = (real)2 * RCOS(clti), acc = s1;
            for (unsigned long k = 1; k < L; k++) {
                s2 = cc2 * s1 - s0;
                s0 = s1;
                s1 = s2;
                s2 = cc2 * s1 - s0;
                acc += s2 / (real)(2 * k + 1);
                s0 = s1;
                s1 = s2;
            }
            w[i] = c * sclti * acc;
        }
        if (i == 0) continue;                                /* no south pole */
        if (nearly_equal(lat[i], 0, GLOB_THRESHOLD)) continue; /* equator */
        w[nlat - i] = w[i];
        lat[nlat - i] = -lat[i];
    }
    for (size_t i = 0; i < nlat; i++) { st(lat_o, nlat, i, lat[i]); st(w_o, nlat, i, w[i]); }
    const real cl = (type == 4) ? RPI / Lf : RPI / ((real)2 * Lf);
    for (size_t j = 0; j < nlon; j++) st(lon_o, nlon, j, cl * (real)j);
    free(lat);
    free(w);
    return 0;
}

/* Point synthesis on a grid (shs_point_grd.c:97-1076).
 *   gl      : 1 if the grid is a Gauss-Legendre grid (then it is symmetric
 *             by definition and the southern rings reuse the northern
 *             latitudes, as the reference does), 0 for a custom grid (each
 *             ring evaluated at its own latitude, which is bitwise the same
 *             as the reference's fold for an exactly symmetric grid), 4 or 5
 *             for a Driscoll-Healy grid (evaluated as a custom grid; the type
 *             sets the polar-optimisation ring blocks, polar_rows).
 *   C, S    : m-major triangles of a model of degree nmax_shc >= nmax.
 *   mu, R   : one value each.  lat, r: nlat values; lon: nlon values.
 *   rings   : if non-NULL, only these nring ring indices are synthesised and
 *             f holds nring*nlon values (row k = ring rings[k]).
 * Returns 0 on success. */
int orc_shs_grid(unsigned long nmax, unsigned long nmax_shc, const double *C_i,
                 const double *S_i, const double *mu_i, const double *R_i,
                 int gl, size_t nlat, const double *lat_i, const double *r_i,
                 size_t nlon, const double *lon_i, const size_t *rings,
                 size_t nring, double *f_o)
{
    init_consts();
    const size_t ncs = tri_len(nmax_shc);
    real *C = ld_all(C_i, ncs), *S = ld_all(S_i, ncs);
    real mu = ld(mu_i, 1, 0), R = ld(R_i, 1, 0);
    real *lat = ld_all(lat_i, nlat), *rr = ld_all(r_i, nlat), *lon = ld_all(lon_i, nlon);
    const real mur = mu / R;

    /* all radii equal to R? (shs_r_eq_rref.c:17-30) */
    int unit = 1;
    for (size_t i = 0; i < nlat; i++)
        if (!nearly_equal(rr[i], R, GLOB_THRESHOLD)) { unit = 0; break; }

    real *rt = (real *)malloc((2 * nmax + 4) * sizeof(real));
    real *rit = (real *)malloc((2 * nmax + 4) * sizeof(real));
    real *dm = (real *)malloc((nmax + 1) * sizeof(real));
    leg_tables(nmax, rt, rit, dm);
    const int use_fft = fft_ok(nlon, lon, nmax);
    real *ct = NULL, *stb = NULL;
    if (use_fft) twiddles(nlon, &ct, &stb);

    const size_t nout = rings ? nring : nlat;
    const int type = gl;
    gl = gl == 1;
    const size_t equator = (gl && !(nlat - 1) % 2) ? nlat / 2 : (size_t)-1;
    /* polar optimisation blocks (polar_rows) */
    real *ucos = (real *)malloc(nlat * sizeof(real));
    size_t *rb = (size_t *)malloc(nlat * sizeof(size_t)), *rn = (size_t *)malloc(nlat * sizeof(size_t));
    for (size_t i = 0; i < nlat; i++) ucos[i] = RCOS(lat[i]);
    polar_rows(type, nlat, lat, polar_bs_s, rb, rn);

#pragma omp parallel
    {
    real *anm = (real *)malloc((nmax + 2) * sizeof(real));
    real *bnm = (real *)malloc((nmax + 2) * sizeof(real));
    real *ps = (real *)malloc((nmax + 1) * sizeof(real));
    int64_t *ips = (int64_t *)malloc((nmax + 1) * sizeof(int64_t));
    real *rp = (real *)malloc((nmax + 3) * sizeof(real));
    real *a = (real *)calloc(nmax + 1, sizeof(real)), *b = (real *)calloc(nmax + 1, sizeof(real));
    real *out = (real *)malloc((nlon ? nlon : 1) * sizeof(real));
#pragma omp for schedule(dynamic)
    for (size_t k = 0; k < nout; k++) {
        size_t i = rings ? rings[k] : k;
        /* latitude actually used for ring i */
        real t, u;
        if (gl && nlat > 1 && i >= nlat - nlat / 2 && !(i == equator)) {
            /* southern ring: mirror of northern ring nlat-1-i
             * (shs_grd_fft.c:117-130; crd_grd_check_symm.c:44-57) */
            size_t in = nlat - 1 - i;
            t = -RSIN(lat[in]);
            u = RCOS(lat[in]);
        } else {
            t = RSIN(lat[i]);
            u = RCOS(lat[i]);
        }
        leg_seed(u, dm, nmax, ps, ips);
        rpows(R, rr[i], unit, nmax + 1, rp);
        for (unsigned long m = 0; m <= nmax; m++) {
            if (polar_skip_block(m, nmax, ucos + rb[i], rn[i])) { a[m] = 0; b[m] = 0; continue; }
            leg_anm_bnm(nmax, m, rt, rit, anm, bnm);
            size_t off = col_off(nmax_shc, m);
            lumped_ring(nmax, m, C + off, S + off, anm, bnm, t, ps, ips, rp, NULL,
                        &a[m], &b[m], NULL, NULL);
        }
        if (use_fft) {
            dft_c2r_direct(nmax, a, b, nlon, ct, stb, out);
        } else {
            for (size_t j = 0; j < nlon; j++) {
                real s = 0;
                for (unsigned long m = 0; m <= nmax; m++) {
                    real ml = (real)m * lon[j];
                    s = s + (a[m] * RCOS(ml) + b[m] * RSIN(ml));
                }
                out[j] = s;
            }
        }
        for (size_t j = 0; j < nlon; j++) st(f_o, nout * nlon, k * nlon + j, mur * out[j]);
    }
    free(anm); free(bnm); free(ps); free(ips); free(rp); free(a); free(b); free(out);
    }
    free(C); free(S); free(lat); free(rr); free(lon); free(rt); free(rit); free(dm);
    free(ct); free(stb); free(ucos); free(rb); free(rn);
    return 0;
}

/* Lumped coefficients a_m, b_m (m = 0..nmax) at one latitude and radius,
 * before the longitude sum and without mu/R (shs_point_kernel.c).  Outputs
 * nmax+1 values each. */
int orc_lumped(unsigned long nmax, unsigned long nmax_shc, const double *C_i,
               const double *S_i, const double *R_i, const double *lat_i,
               const double *r_i, double *a_o, double *b_o)
{
    init_consts();
    const size_t ncs = tri_len(nmax_shc);
    real *C = ld_all(C_i, ncs), *S = ld_all(S_i, ncs);
    real R = ld(R_i, 1, 0), lat = ld(lat_i, 1, 0), r = ld(r_i, 1, 0);
    real *rt = (real *)malloc((2 * nmax + 4) * sizeof(real));
    real *rit = (real *)malloc((2 * nmax + 4) * sizeof(real));
    real *dm = (real *)malloc((nmax + 1) * sizeof(real));
    leg_tables(nmax, rt, rit, dm);
    real *anm = (real *)malloc((nmax + 2) * sizeof(real));
    real *bnm = (real *)malloc((nmax + 2) * sizeof(real));
    real *ps = (real *)malloc((nmax + 1) * sizeof(real));
    int64_t *ips = (int64_t *)malloc((nmax + 1) * sizeof(int64_t));
    real *rp = (real *)malloc((nmax + 3) * sizeof(real));
    leg_seed(RCOS(lat), dm, nmax, ps, ips);
    rpows(R, r, nearly_equal(r, R, GLOB_THRESHOLD), nmax + 1, rp);
    real t = RSIN(lat);
    for (unsigned long m = 0; m <= nmax; m++) {
        real a = 0, b = 0;
        const real ul = RCOS(lat);
        if (!polar_skip_block(m, nmax, &ul, 1)) {
            leg_anm_bnm(nmax, m, rt, rit, anm, bnm);
            size_t off = col_off(nmax_shc, m);
            lumped_ring(nmax, m, C + off, S + off, anm, bnm, t, ps, ips, rp, NULL, &a, &b, NULL, NULL);
        }
        st(a_o, nmax + 1, m, a);
        st(b_o, nmax + 1, m, b);
    }
    free(C); free(S); free(rt); free(rit); free(dm); free(anm); free(bnm); free(ps);
    free(ips); free(rp);
    return 0;
}

/* Point synthesis at scattered points (shs_point_sctr.c:86-728,
 * shs_sctr_mulc.c:17-61): f_i = mu/R * sum_m (a_m cos m lon_i + b_m sin m
 * lon_i), lumped coefficients without the symmetry fold. */
int orc_shs_sctr(unsigned long nmax, unsigned long nmax_shc, const double *C_i,
                 const double *S_i, const double *mu_i, const double *R_i,
                 size_t npoint, const double *lat_i, const double *lon_i,
                 const double *r_i, double *f_o)
{
    init_consts();
    const size_t ncs = tri_len(nmax_shc);
    real *C = ld_all(C_i, ncs), *S = ld_all(S_i, ncs);
    real mu = ld(mu_i, 1, 0), R = ld(R_i, 1, 0);
    real *lat = ld_all(lat_i, npoint), *lon = ld_all(lon_i, npoint), *rr = ld_all(r_i, npoint);
    int unit = 1;
    for (size_t i = 0; i < npoint; i++)
        if (!nearly_equal(rr[i], R, GLOB_THRESHOLD)) { unit = 0; break; }
    real *rt = (real *)malloc((2 * nmax + 4) * sizeof(real));
    real *rit = (real *)malloc((2 * nmax + 4) * sizeof(real));
    real *dm = (real *)malloc((nmax + 1) * sizeof(real));
    leg_tables(nmax, rt, rit, dm);
    const real mur = mu / R;
    /* polar optimisation blocks: polar_bs_s consecutive points */
    real *ucos = (real *)malloc((npoint ? npoint : 1) * sizeof(real));
    for (size_t i = 0; i < npoint; i++) ucos[i] = RCOS(lat[i]);
    const size_t bs = polar_bs_s;
#pragma omp parallel
    {
    real *anm = (real *)malloc((nmax + 2) * sizeof(real));
    real *bnm = (real *)malloc((nmax + 2) * sizeof(real));
    real *ps = (real *)malloc((nmax + 1) * sizeof(real));
    int64_t *ips = (int64_t *)malloc((nmax + 1) * sizeof(int64_t));
    real *rp = (real *)malloc((nmax + 3) * sizeof(real));
#pragma omp for schedule(dynamic)
    for (size_t i = 0; i < npoint; i++) {
        real t = RSIN(lat[i]), u = RCOS(lat[i]);
        const size_t b0 = (i / bs) * bs, bn = (b0 + bs < npoint ? b0 + bs : npoint) - b0;
        leg_seed(u, dm, nmax, ps, ips);
        rpows(R, rr[i], unit, nmax + 1, rp);
        real fi = 0;
        for (unsigned long m = 0; m <= nmax; m++) {
            real a, b;
            if (polar_skip_block(m, nmax, ucos + b0, bn)) continue;
            leg_anm_bnm(nmax, m, rt, rit, anm, bnm);
            size_t off = col_off(nmax_shc, m);
            lumped_ring(nmax, m, C + off, S + off, anm, bnm, t, ps, ips, rp, NULL, &a, &b, NULL, NULL);
            real ml = (real)m * lon[i];
            fi = fi + (a * RCOS(ml) + b * RSIN(ml));
        }
        st(f_o, npoint, i, mur * fi);
    }
    free(anm); free(bnm); free(ps); free(ips); free(rp);
    }
    free(C); free(S); free(lat); free(lon); free(rr); free(rt); free(rit); free(dm); free(ucos);
    return 0;
}

/* Point analysis on a Gauss-Legendre grid (sha_point.c:124-1125).
 *   f        : nlat*nlon values (row-major, ring i = row i).
 *   lat, w   : nlat values; r0 = radius of the grid (one value).
 *   mu, R    : the output model's constants; coefficients are rescaled from
 *              r0 to R when they differ (sha_point.c:1088-1108).
 *   C, S     : outputs, m-major triangles of degree nmax_shc >= nmax; every
 *              entry is written (zeros above nmax, sha_point.c:381).
 * Returns 0 on success, 1 on invalid input. */
static int orc_sha_quad(int type, unsigned long nmax, unsigned long nmax_shc, size_t nlat,
                        size_t nlon, const double *f_i, const double *lat_i, const double *w_i,
                        const double *r0_i, const double *mu_i, const double *R_i,
                        double *C_o, double *S_o);

int orc_sha_gl(unsigned long nmax, unsigned long nmax_shc, size_t nlat, size_t nlon,
               const double *f_i, const double *lat_i, const double *w_i,
               const double *r0_i, const double *mu_i, const double *R_i,
               double *C_o, double *S_o)
{
    return orc_sha_quad(3, nmax, nmax_shc, nlat, nlon, f_i, lat_i, w_i, r0_i, mu_i, R_i, C_o, S_o);
}

/* The same on a Driscoll--Healy grid (type 4 = DH1, 5 = DH2): nmax_grd =
 * nlat/2 - 1, c = pi/(nmax_grd+1) (DH1) or pi/(2 nmax_grd+2) (DH2)
 * (sha_point.c:313-317); the grid counts one more latitude (the missing
 * south pole, sha_point.c:190-192), the north pole has no mirror, ring i
 * < nlat/2 mirrors to row nlat - i (crd_grd_check_symm.c:40-52). */
int orc_sha_dh(int type, unsigned long nmax, unsigned long nmax_shc, size_t nlat, size_t nlon,
               const double *f_i, const double *lat_i, const double *w_i,
               const double *r0_i, const double *mu_i, const double *R_i,
               double *C_o, double *S_o)
{
    if (type != 4 && type != 5) return 1;
    return orc_sha_quad(type, nmax, nmax_shc, nlat, nlon, f_i, lat_i, w_i, r0_i, mu_i, R_i, C_o,
                        S_o);
}

static int orc_sha_quad(int type, unsigned long nmax, unsigned long nmax_shc, size_t nlat,
                        size_t nlon, const double *f_i, const double *lat_i, const double *w_i,
                        const double *r0_i, const double *mu_i, const double *R_i,
                        double *C_o, double *S_o)
{
    init_consts();
    const int dh = type != 3;
    if (nlat == 0 || (dh && nlat % 2)) return 1;
    const unsigned long nmax_grd = dh ? nlat / 2 - 1 : nlat - 1;
    if (nmax > nmax_grd || nmax > nmax_shc) return 1;
    real *f = ld_all(f_i, nlat * nlon), *lat = ld_all(lat_i, nlat), *w = ld_all(w_i, nlat);
    real r0 = ld(r0_i, 1, 0), mu = ld(mu_i, 1, 0), R = ld(R_i, 1, 0);
    const size_t ncs = tri_len(nmax_shc);
    real *C = (real *)calloc(ncs, sizeof(real)), *S = (real *)calloc(ncs, sizeof(real));

    /* c = pi/(nmax_grd+1) (GL, DH1) or pi/(2 nmax_grd + 2) (DH2), times
     * 1/(4 pi) * r0/mu (sha_point.c:313-322) */
    real c = (type == 5) ? RPI / (real)(2 * nmax_grd + 2) : RPI / (real)(nmax_grd + 1);
    c *= ((real)1) / ((real)4 * RPI) * (r0 / mu);

    real *ct, *stb;
    twiddles(nlon, &ct, &stb);
    /* DH grids count the missing south pole (sha_point.c:190-192) */
    const size_t nlat_s = dh ? nlat + 1 : nlat;
    const int even = !(nlat_s % 2);
    const size_t equator = dh ? nlat / 2 : ((!(nmax_grd % 2)) ? nlat / 2 : (size_t)-1);
    const size_t nlatdo = (nlat_s > 1) ? (nlat_s + 1 - even) / 2 : nlat_s;
    /* lumped sums per ring pair: A, B for the ring and A2, B2 for its mirror */
    real *re = (real *)malloc((nmax + 1) * sizeof(real)), *mim = (real *)malloc((nmax + 1) * sizeof(real));
    real *amp = (real *)malloc((nmax + 1) * nlatdo * sizeof(real));
    real *amm = (real *)malloc((nmax + 1) * nlatdo * sizeof(real));
    real *bmp = (real *)malloc((nmax + 1) * nlatdo * sizeof(real));
    real *bmm = (real *)malloc((nmax + 1) * nlatdo * sizeof(real));
    real *tt = (real *)malloc(nlatdo * sizeof(real)), *uu = (real *)malloc(nlatdo * sizeof(real));
    real *A = (real *)malloc((nmax + 1) * sizeof(real)), *B = (real *)malloc((nmax + 1) * sizeof(real));
    for (size_t i = 0; i < nlatdo; i++) {
        tt[i] = RSIN(lat[i]);
        uu[i] = RCOS(lat[i]);
        int symm = (nlat > 1) && (i < equator) && !(dh && i == 0);
        dft_r2c_direct(nlon, f + i * nlon, nmax, ct, stb, re, mim);
        real cw = c * w[i];
        for (unsigned long k = 0; k <= nmax; k++) { A[k] = cw * re[k]; B[k] = cw * mim[k]; }
        real *A2 = (real *)calloc(nmax + 1, sizeof(real)), *B2 = (real *)calloc(nmax + 1, sizeof(real));
        if (symm) {
            size_t is = nlat_s - 1 - i;
            dft_r2c_direct(nlon, f + is * nlon, nmax, ct, stb, re, mim);
            real cw2 = c * w[is];
            for (unsigned long k = 0; k <= nmax; k++) { A2[k] = cw2 * re[k]; B2[k] = cw2 * mim[k]; }
        }
        real sy = symm ? 1 : 0;
        for (unsigned long m = 0; m <= nmax; m++) {
            size_t q = m * nlatdo + i;
            amp[q] = (A[m] + sy * A2[m]);
            amm[q] = (A[m] - sy * A2[m]);
            bmp[q] = (B[m] + sy * B2[m]);
            bmm[q] = (B[m] - sy * B2[m]);
        }
        free(A2); free(B2);
    }

    real *rt = (real *)malloc((2 * nmax + 4) * sizeof(real));
    real *rit = (real *)malloc((2 * nmax + 4) * sizeof(real));
    real *dm = (real *)malloc((nmax + 1) * sizeof(real));
    leg_tables(nmax, rt, rit, dm);
    real *anm = (real *)malloc((nmax + 2) * sizeof(real));
    real *bnm = (real *)malloc((nmax + 2) * sizeof(real));
    real *ps = (real *)malloc(((size_t)nmax + 1) * nlatdo * sizeof(real));
    int64_t *ips = (int64_t *)malloc(((size_t)nmax + 1) * nlatdo * sizeof(int64_t));
    for (size_t i = 0; i < nlatdo; i++) leg_seed(uu[i], dm, nmax, ps + i * (nmax + 1), ips + i * (nmax + 1));

    for (unsigned long m = 0; m <= nmax; m++) {
        leg_anm_bnm(nmax, m, rt, rit, anm, bnm);
        real *Cm = C + col_off(nmax_shc, m), *Sm = S + col_off(nmax_shc, m);
        for (size_t i = 0; i < nlatdo; i++) {
            const size_t b0 = (i / polar_bs_a) * polar_bs_a;
            const size_t bn = (b0 + polar_bs_a < nlatdo ? b0 + polar_bs_a : nlatdo) - b0;
            if (polar_skip_block(m, nmax, uu + b0, bn)) continue;
            real t = tt[i];
            size_t q = m * nlatdo + i;
            real ap = amp[q], am = amm[q], bp = bmp[q], bm = bmm[q];
            const real *psi = ps + i * (nmax + 1);
            const int64_t *ipsi = ips + i * (nmax + 1);
            /* per ring: P_nm(t) in the same sequence as the synthesis, then
             * C_nm += P (n+m even ? amp : amm) (sha_point.c:756-969) */
            if (m == 0) {
                real p0 = 1, p1, p2;
                Cm[0] = Cm[0] + p0 * ap;
                if (nmax >= 1) {
                    p1 = RSQRT((real)3) * t;
                    Cm[1] = Cm[1] + p1 * am;
                    int npm_even = 1;
                    for (unsigned long n = 2; n <= nmax; n++, npm_even = !npm_even) {
                        p2 = (anm[n] * t) * p1 - bnm[n] * p0;
                        p0 = p1;
                        p1 = p2;
                        Cm[n] = Cm[n] + p2 * (npm_even ? ap : am);
                    }
                }
            } else {
                real x = psi[m - 1], y = 0, z, p;
                int64_t ix = ipsi[m - 1], iy = 0, iz;
                p = x2f(x, ix);
                Cm[0] = Cm[0] + p * ap;
                Sm[0] = Sm[0] + p * bp;
                if (m < nmax) {
                    y = x; iy = ix;
                    x = (anm[m + 1] * t) * y;
                    ix = iy;
                    XNORM(x, ix);
                    p = x2f(x, ix);
                    Cm[1] = Cm[1] + p * am;
                    Sm[1] = Sm[1] + p * bm;
                    int ds = 0, npm_even = 1;
                    for (unsigned long n = m + 2; n <= nmax; n++, npm_even = !npm_even) {
                        if (ds) {
                            p = (anm[n] * t) * x - bnm[n] * y;
                            y = x;
                            x = p;
                        } else {
                            int64_t ixy = ix - iy;
                            if (ixy == 0) { z = (anm[n] * t) * x - bnm[n] * y; iz = ix; }
                            else if (ixy == 1) { z = (anm[n] * t) * x - bnm[n] * (y * BIGI); iz = ix; }
                            else if (ixy == -1) { z = (anm[n] * t) * (x * BIGI) - bnm[n] * y; iz = iy; }
                            else if (ixy > 1) { z = (anm[n] * t) * x; iz = ix; }
                            else { z = -bnm[n] * y; iz = iy; }
                            XNORM(z, iz);
                            y = x; iy = ix;
                            x = z; ix = iz;
                            p = x2f(z, iz);
                            if (ix == 0 && iy == 0) ds = 1;
                        }
                        Cm[n - m] = Cm[n - m] + p * (npm_even ? ap : am);
                        Sm[n - m] = Sm[n - m] + p * (npm_even ? bp : bm);
                    }
                }
            }
        }
    }

    /* rescale from r0 to R (sha_point.c:1088-1108, shc_rescale.c:24-90) */
    if (!nearly_equal(R, r0, GLOB_THRESHOLD)) {
        real r_ratio = r0 / R;
        real tmp2 = r_ratio;  /* mu_ratio = 1 */
        real *sc = (real *)malloc((nmax_shc + 1) * sizeof(real));
        sc[0] = 1;
        for (unsigned long n = 1; n <= nmax_shc; n++) { sc[n] = tmp2; tmp2 *= r_ratio; }
        for (unsigned long m = 0; m <= nmax_shc; m++) {
            real *Cm = C + col_off(nmax_shc, m), *Sm = S + col_off(nmax_shc, m);
            for (unsigned long n = m; n <= nmax_shc; n++) { Cm[n - m] *= sc[n]; Sm[n - m] *= sc[n]; }
        }
        free(sc);
    }
    for (size_t q = 0; q < ncs; q++) { st(C_o, ncs, q, C[q]); st(S_o, ncs, q, S[q]); }

    free(f); free(lat); free(w); free(C); free(S); free(ct); free(stb); free(re); free(mim);
    free(amp); free(amm); free(bmp); free(bmm); free(tt); free(uu); free(A); free(B);
    free(rt); free(rit); free(dm); free(anm); free(bnm); free(ps); free(ips);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Derivatives of the synthesis                                             */
/* ------------------------------------------------------------------------ */
/* P_nm(t), n = m..nmax, of one order into p[n - m], by the recurrences of
 * lumped_ring (the same X-number switch per lane). */
static void leg_column(unsigned long nmax, unsigned long m, const real *anm,
                       const real *bnm, real t, const real *ps, const int64_t *ips,
                       real *p)
{
    if (m == 0) {
        real p0 = 1, p1, p2;
        p[0] = p0;
        if (nmax >= 1) {
            p1 = RSQRT((real)3) * t;
            p[1] = p1;
            for (unsigned long n = 2; n <= nmax; n++) {
                p2 = (anm[n] * t) * p1 - bnm[n] * p0;
                p[n] = p2;
                p0 = p1;
                p1 = p2;
            }
        }
        return;
    }
    real x = ps[m - 1], y = 0, z;
    int64_t ix = ips[m - 1], iy = 0, iz;
    p[0] = x2f(x, ix);
    if (m == nmax) return;
    y = x; iy = ix;
    x = (anm[m + 1] * t) * y;
    ix = iy;
    XNORM(x, ix);
    p[1] = x2f(x, ix);
    int ds = 0;
    for (unsigned long n = m + 2; n <= nmax; n++) {
        real v;
        if (ds) {
            v = (anm[n] * t) * x - bnm[n] * y;
            y = x;
            x = v;
        } else {
            int64_t ixy = ix - iy;
            if (ixy == 0) { z = (anm[n] * t) * x - bnm[n] * y; iz = ix; }
            else if (ixy == 1) { z = (anm[n] * t) * x - bnm[n] * (y * BIGI); iz = ix; }
            else if (ixy == -1) { z = (anm[n] * t) * (x * BIGI) - bnm[n] * y; iz = iy; }
            else if (ixy > 1) { z = (anm[n] * t) * x; iz = ix; }
            else { z = -bnm[n] * y; iz = iy; }
            XNORM(z, iz);
            y = x; iy = ix;
            x = z; ix = iz;
            v = x2f(z, iz);
            if (ix == 0 && iy == 0) ds = 1;
        }
        p[n - m] = v;
    }
}

/* The ten g_ijk = 1/(r^(j+k) cos^k lat) d^(i+j+k) f / dr^i dlat^j dlon^k,
 * i + j + k <= 2, at one point, in the order
 *   000, 100, 010, 001, 200, 020, 002, 110, 101, 011.
 * Per term, as CHarm's kernels evaluate them (shs_point_kernel.c:176-487,
 * leg_func_xnum.h:35-44, leg_func_enm.c:40-41, shs_get_mur_dorder_npar.c):
 *   dP_nm  = e_nm / u P_{n-1,m} - n t/u P_nm
 *   ddP_nm = t/u dP_nm + (m^2/u^2 - n(n+1)) P_nm
 * radial factors -(n+1) and (n+1)(n+2) times (R/r)^(n+1), longitude
 * derivatives of a cos + b sin, and 1/r^(i+j+k) with r = R when every radius
 * equals R. */
static void guru_point(unsigned long nmax, unsigned long nmax_shc, const real *C,
                       const real *S, real mu, real R, real lat, real lon, real r,
                       int unit, const real *rt, const real *rit, const real *dm,
                       real *anm, real *bnm, real *enm, real *ps, int64_t *ips, real *rp,
                       real *p, const real *ub, size_t nb, real *g)
{
    const real t = RSIN(lat), u = RCOS(lat);
    const real tu = t / u, ui = ((real)1) / u, ui2 = ui * ui;
    leg_seed(u, dm, nmax, ps, ips);
    rpows(R, r, unit, nmax + 1, rp);
    real F[3][3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            for (int k = 0; k < 3; k++) F[i][j][k] = 0;
    for (unsigned long m = 0; m <= nmax; m++) {
        if (polar_skip_block(m, nmax, ub, nb)) continue;
        leg_anm_bnm(nmax, m, rt, rit, anm, bnm);
        enm[m] = 0;
        for (unsigned long n = m + 1; n <= nmax; n++)
            enm[n] = rt[2 * n + 1] * rt[n - m] * rt[n + m] * rit[2 * n - 1];
        leg_column(nmax, m, anm, bnm, t, ps, ips, p);
        const size_t off = col_off(nmax_shc, m);
        real A[3][3], B[3][3];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) A[i][j] = B[i][j] = 0;
        const real m2 = (real)m * (real)m;
        for (unsigned long n = m; n <= nmax; n++) {
            const real pn = p[n - m], pp = (n > m) ? p[n - m - 1] : 0;
            const real dn = (real)n;
            const real dp = (enm[n] * ui) * pp - (dn * tu) * pn;
            const real ddp = tu * dp + (m2 * ui2 - dn * (dn + 1)) * pn;
            const real L[3] = {pn, dp, ddp};
            const real W[3] = {rp[n + 1], (-(dn + 1)) * rp[n + 1],
                               ((dn + 1) * (dn + 2)) * rp[n + 1]};
            const real c = C[off + (n - m)], sv = (m == 0) ? 0 : S[off + (n - m)];
            for (int i = 0; i < 3; i++)
                for (int j = 0; i + j < 3; j++) {
                    const real wl = W[i] * L[j];
                    A[i][j] = A[i][j] + wl * c;
                    B[i][j] = B[i][j] + wl * sv;
                }
        }
        const real ml = (real)m * lon, cm = RCOS(ml), sm = RSIN(ml);
        for (int i = 0; i < 3; i++)
            for (int j = 0; i + j < 3; j++) {
                F[i][j][0] = F[i][j][0] + (A[i][j] * cm + B[i][j] * sm);
                if (i + j < 2)
                    F[i][j][1] = F[i][j][1] + (real)m * (B[i][j] * cm - A[i][j] * sm);
                if (i + j == 0)
                    F[i][j][2] = F[i][j][2] - m2 * (A[i][j] * cm + B[i][j] * sm);
            }
    }
    const real rr = unit ? R : r, ri = ((real)1) / rr, mur = mu / R;
    static const int IJK[10][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {2, 0, 0},
                                   {0, 2, 0}, {0, 0, 2}, {1, 1, 0}, {1, 0, 1}, {0, 1, 1}};
    for (int q = 0; q < 10; q++) {
        const int i = IJK[q][0], j = IJK[q][1], k = IJK[q][2];
        real v = mur * F[i][j][k];
        for (int e = 0; e < i + j + k; e++) v = v * ri;
        for (int e = 0; e < k; e++) v = v * ui;
        g[q] = v;
    }
}

/* Derivatives at scattered points (src/shs/shs_point_sctr.c with
 * dr, dlat, dlon as in shs_point_guru.c and shs_point_gradn.c).
 *   kind 0..9 : one g_ijk in the order of guru_point, 1 output;
 *   kind 10   : grad1 (fx, fy, fz) = (g010, -g001, g100), 3 outputs;
 *   kind 11   : grad2 after Reed (1973) (src/shs/shs.h.in:278-312),
 *               (fxx, fxy, fxz, fyy, fyz, fzz), 6 outputs.
 * f receives the outputs one after another, npoint values each. */
int orc_shs_deriv_sctr(int kind, unsigned long nmax, unsigned long nmax_shc,
                       const double *C_i, const double *S_i, const double *mu_i,
                       const double *R_i, size_t npoint, const double *lat_i,
                       const double *lon_i, const double *r_i, double *f_o)
{
    if (kind < 0 || kind > 11) return 1;
    init_consts();
    const size_t ncs = tri_len(nmax_shc);
    real *C = ld_all(C_i, ncs), *S = ld_all(S_i, ncs);
    real mu = ld(mu_i, 1, 0), R = ld(R_i, 1, 0);
    real *lat = ld_all(lat_i, npoint), *lon = ld_all(lon_i, npoint), *rr = ld_all(r_i, npoint);
    int unit = 1;
    for (size_t i = 0; i < npoint; i++)
        if (!nearly_equal(rr[i], R, GLOB_THRESHOLD)) { unit = 0; break; }
    real *rt = (real *)malloc((2 * nmax + 4) * sizeof(real));
    real *rit = (real *)malloc((2 * nmax + 4) * sizeof(real));
    real *dm = (real *)malloc((nmax + 1) * sizeof(real));
    leg_tables(nmax, rt, rit, dm);
    const int nout = kind < 10 ? 1 : (kind == 10 ? 3 : 6);
    /* polar optimisation blocks: polar_bs_s consecutive points */
    real *ucos = (real *)malloc((npoint ? npoint : 1) * sizeof(real));
    for (size_t i = 0; i < npoint; i++) ucos[i] = RCOS(lat[i]);
    const size_t bs = polar_bs_s;
#pragma omp parallel
    {
    real *anm = (real *)malloc((nmax + 2) * sizeof(real));
    real *bnm = (real *)malloc((nmax + 2) * sizeof(real));
    real *enm = (real *)malloc((nmax + 2) * sizeof(real));
    real *ps = (real *)malloc((nmax + 1) * sizeof(real));
    int64_t *ips = (int64_t *)malloc((nmax + 1) * sizeof(int64_t));
    real *rp = (real *)malloc((nmax + 3) * sizeof(real));
    real *p = (real *)malloc((nmax + 1) * sizeof(real));
#pragma omp for schedule(dynamic)
    for (size_t i = 0; i < npoint; i++) {
        real g[10], o[6];
        const size_t b0 = (i / bs) * bs, bn = (b0 + bs < npoint ? b0 + bs : npoint) - b0;
        guru_point(nmax, nmax_shc, C, S, mu, R, lat[i], lon[i], rr[i], unit, rt, rit, dm,
                   anm, bnm, enm, ps, ips, rp, p, ucos + b0, bn, g);
        const real r = unit ? R : rr[i], ri = ((real)1) / r;
        const real tu = RSIN(lat[i]) / RCOS(lat[i]);
        if (kind < 10) {
            o[0] = g[kind];
        } else if (kind == 10) {
            o[0] = g[2];
            o[1] = -g[3];
            o[2] = g[1];
        } else {
            o[0] = g[5] + g[1] * ri;
            o[1] = -g[9] - tu * g[3] * ri;
            o[2] = g[7] - g[2] * ri;
            o[3] = g[6] + g[1] * ri - tu * g[2] * ri;
            o[4] = -g[8] + g[3] * ri;
            o[5] = g[4];
        }
        for (int q = 0; q < nout; q++) st(f_o, nout * npoint, q * npoint + i, o[q]);
    }
    free(anm); free(bnm); free(enm); free(ps); free(ips); free(rp); free(p);
    }
    free(C); free(S); free(lat); free(lon); free(rr); free(rt); free(rit); free(dm); free(ucos);
    return 0;
}

/* Fully-normalised P_nm(sin lat) for all 0 <= m <= n <= nmax at one latitude,
 * through the same recurrences (m-major triangle output, ncs values). */
int orc_leg_pnm(unsigned long nmax, const double *lat_i, double *p_o)
{
    init_consts();
    real lat = ld(lat_i, 1, 0), t = RSIN(lat), u = RCOS(lat);
    const size_t ncs = tri_len(nmax);
    real *C = (real *)calloc(ncs, sizeof(real)), *S = (real *)calloc(ncs, sizeof(real));
    real *rt = (real *)malloc((2 * nmax + 4) * sizeof(real));
    real *rit = (real *)malloc((2 * nmax + 4) * sizeof(real));
    real *dm = (real *)malloc((nmax + 1) * sizeof(real));
    leg_tables(nmax, rt, rit, dm);
    real *anm = (real *)malloc((nmax + 2) * sizeof(real));
    real *bnm = (real *)malloc((nmax + 2) * sizeof(real));
    real *ps = (real *)malloc((nmax + 1) * sizeof(real));
    int64_t *ips = (int64_t *)malloc((nmax + 1) * sizeof(int64_t));
    real *rp = (real *)malloc((nmax + 3) * sizeof(real));
    leg_seed(u, dm, nmax, ps, ips);
    rpows(1, 1, 1, nmax + 1, rp);
    /* P_nm = lumped sum with a single unit coefficient C_nm = 1 */
    for (unsigned long m = 0; m <= nmax; m++) {
        leg_anm_bnm(nmax, m, rt, rit, anm, bnm);
        size_t off = col_off(nmax, m);
        for (unsigned long n = m; n <= nmax; n++) {
            real a, b;
            C[off + (n - m)] = 1;
            lumped_ring(nmax, m, C + off, S + off, anm, bnm, t, ps, ips, rp, NULL, &a, &b, NULL, NULL);
            C[off + (n - m)] = 0;
            st(p_o, ncs, off + (n - m), a);
        }
    }
    free(C); free(S); free(rt); free(rit); free(dm); free(anm); free(bnm); free(ps);
    free(ips); free(rp);
    return 0;
}
