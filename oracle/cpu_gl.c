/*
 * CPU restatement of CHarm's Gauss--Legendre synthesis and analysis, for the
 * benchmark's CPU baseline.
 *
 * TEST INFRASTRUCTURE ONLY.  Only the tests and the CPU baseline leg of the
 * benchmark may load this code.  The product (libcharm_amd.so) never links,
 * loads or calls it.
 *
 * The arithmetic is the oracle's (charm_oracle.c: the same Legendre tables,
 * sectorial seeds as X-numbers, the same X-number rules per ring until two
 * consecutive terms are ordinary numbers), organised the way CHarm runs it
 * on a CPU, so that the baseline is a fair CPU implementation rather than a
 * checker:
 *   - ring pairs folded by the parity P_nm(-t) = (-1)^(n+m) P_nm(t)
 *     (shs_point_grd.c:97-1076, sha_point.c:124-1125), one recurrence for
 *     both rings of a pair;
 *   - blocks of LANES northern rings advanced together in loops the compiler
 *     vectorises (CHarm's SIMD_SIZE x BLOCK_S blocks, src/simd/simd.h:298-304,
 *     32 rings in its AVX-512 double build);
 *   - OpenMP threads over (order, ring block) tasks in synthesis and over
 *     orders in analysis;
 *   - the longitude sums by an O(n log n) FFT, done by the caller
 *     (scipy.fft over the rows) on the spectra this file produces or takes.
 * The polar optimisation is off (CHarm's default, src/glob/glob.c:20-23).
 * The recurrence regions run with subnormals flushed to zero (MXCSR FTZ and
 * DAZ, restored on exit), as CHarm's benchmark build with -ffast-math does
 * (docs/src/benchmarks.rst:9-14): the masked X-number steps compute every
 * branch, and subnormal results would otherwise cost a microcode assist each;
 * only terms below 2^-1022 change, to zero.
 * Sums are formed in a different order from the oracle's, so results agree
 * with it to rounding, not bit for bit.
 */
#include <math.h>
#include <omp.h>
#include <xmmintrin.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define LANES 32
#define GROUP 16
#define XEXP 960

static double BIG, BIGI, BIGS, BIGSI;

static void init_consts(void)
{
    BIG = ldexp(1.0, XEXP);
    BIGI = ldexp(1.0, -XEXP);
    BIGS = ldexp(1.0, XEXP / 2);
    BIGSI = ldexp(1.0, -XEXP / 2);
}

static inline void xnorm(double *x, int64_t *ix)
{
    double w = fabs(*x);
    if (w >= BIGS) { *x *= BIGI; *ix += 1; }
    else if (w < BIGSI) { *x *= BIG; *ix -= 1; }
}

static inline double x2f(double x, int64_t ix)
{
    if (ix == 0) return x;
    if (ix < -1) return 0;
    if (ix < 0) return x * BIGI;
    return x * BIG;
}

/* sqrt(k), 1/sqrt(k) and dm[n] = sqrt(2n+3)/sqrt(2n+2) (leg_func_r_ri.c,
 * leg_func_dm.c) */
static void tables(unsigned long nmax, double *r, double *ri, double *dm)
{
    r[0] = ri[0] = 0;
    for (unsigned long k = 1; k <= 2 * nmax + 3; k++) {
        r[k] = sqrt((double)k);
        ri[k] = 1.0 / r[k];
    }
    dm[0] = 0;
    for (unsigned long n = 1; n <= nmax; n++) dm[n] = r[2 * n + 3] * ri[2 * n + 2];
}

/* a_nm, b_nm of order m (leg_func_anm_bnm.c:14-65) */
static void anm_bnm(unsigned long nmax, unsigned long m, const double *r, const double *ri,
                    double *anm, double *bnm)
{
    anm[m] = bnm[m] = 0;
    if (m == nmax) return;
    anm[m + 1] = r[2 * m + 3];
    bnm[m + 1] = 0;
    for (unsigned long n = m + 2; n <= nmax; n++) {
        double w = r[2 * n + 1] * ri[n - m] * ri[n + m];
        anm[n] = r[2 * n - 1] * w;
        bnm[n] = r[n - m - 1] * r[n + m - 1] * ri[2 * n - 3] * w;
    }
}

/* Sectorial seeds P_mm (m = 1..nmax) of one ring as X-numbers
 * (leg_func_prepare.c:18-69) */
static void seeds(double u, const double *dm, unsigned long nmax, double *ps, int64_t *ips)
{
    if (nmax == 0) return;
    double x = sqrt(3.0) * u;
    int64_t ix = 0;
    ps[0] = x;
    ips[0] = ix;
    for (unsigned long n = 1; n < nmax; n++) {
        x = (dm[n] * u) * x;
        xnorm(&x, &ix);
        ps[n] = x;
        ips[n] = ix;
    }
}

/* Order m's Legendre functions at the LANES rings of a block, from the
 * sectorial seeds (xs, ixs per lane), with X-numbers until every lane has
 * had two consecutive ordinary terms (leg_func_xnum.h:336-443; a lane that
 * got there continues with the ordinary recurrence, masked, as CHarm's SIMD
 * blocks do): p[(n - m) * LANES + l] receives P_nm for n below the returned
 * degree nd, and x, y receive P_{nd-1}, P_{nd-2} per lane, from which the
 * ordinary recurrence continues. */
static unsigned long xblock(unsigned long nmax, unsigned long m, const double *anm,
                            const double *bnm, const double *t, const double *xs,
                            const int64_t *ixs, double *p, double *x, double *y)
{
    if (m == 0) {
        for (int l = 0; l < LANES; l++) {
            p[l] = 1.0;
            x[l] = 1.0;
            y[l] = 0.0;
        }
        if (nmax == 0) return 1;
        const double s3 = sqrt(3.0);
        for (int l = 0; l < LANES; l++) {
            p[LANES + l] = s3 * t[l];
            y[l] = 1.0;
            x[l] = p[LANES + l];
        }
        return 2;
    }
    int64_t ix[LANES], iy[LANES], ds[LANES];
    for (int l = 0; l < LANES; l++) {
        x[l] = xs[l];
        ix[l] = ixs[l];
        p[l] = x2f(x[l], ix[l]);
    }
    if (m == nmax) {
        for (int l = 0; l < LANES; l++) {
            x[l] = p[l];
            y[l] = 0.0;
        }
        return m + 1;
    }
    const double a1 = anm[m + 1];
    for (int l = 0; l < LANES; l++) {
        y[l] = x[l];
        iy[l] = ix[l];
        double z = (a1 * t[l]) * y[l];
        int64_t iz = iy[l];
        const double w = fabs(z);
        const int64_t up = w >= BIGS, dn = w < BIGSI;
        z = up ? z * BIGI : (dn ? z * BIG : z);
        iz += up - dn;
        x[l] = z;
        ix[l] = iz;
        p[LANES + l] = x2f(z, iz);
        ds[l] = 0;
    }
    /* the masked X-number phase runs per group of GROUP lanes (one vector),
     * each group until all its lanes are ordinary */
    unsigned long ndg[LANES / GROUP], nd = m + 2;
    for (int g = 0; g < LANES / GROUP; g++) {
        const int l0 = g * GROUP;
        unsigned long n = m + 2;
        for (; n <= nmax; n++) {
            const double an = anm[n], bn = bnm[n];
            double *pn = p + (n - m) * LANES;
            int64_t all = 1;
            for (int l = l0; l < l0 + GROUP; l++) {
                const double xl = x[l], yl = y[l];
                const int64_t ixl = ix[l], iyl = iy[l], d = ds[l];
                const int64_t ixy = ixl - iyl;
                const double at = an * t[l];
                const double xa = (ixy == -1) ? xl * BIGI : xl;
                const double yb = (ixy == 1) ? yl * BIGI : yl;
                double z = (ixy > 1) ? at * xl : ((ixy < -1) ? -bn * yl : at * xa - bn * yb);
                int64_t iz = (ixy >= 0) ? ixl : iyl;
                const double w = fabs(z);
                const int64_t up = !d && w >= BIGS, dn = !d && w < BIGSI;
                z = up ? z * BIGI : (dn ? z * BIG : z);
                iz += up - dn;
                y[l] = xl;
                iy[l] = ixl;
                x[l] = z;
                ix[l] = iz;
                pn[l] = (iz == 0) ? z : ((iz < -1) ? 0.0 : ((iz < 0) ? z * BIGI : z * BIG));
                const int64_t dl = d | (iz == 0 && ixl == 0);
                ds[l] = dl;
                all &= dl;
            }
            if (all) break;
        }
        ndg[g] = (n <= nmax) ? n + 1 : nmax + 1;
        if (ndg[g] > nd) nd = ndg[g];
    }
    /* groups that became ordinary early continue plainly up to nd */
    for (int g = 0; g < LANES / GROUP; g++) {
        const int l0 = g * GROUP;
        for (unsigned long n = ndg[g]; n < nd; n++) {
            const double an = anm[n], bn = bnm[n];
            double *pn = p + (n - m) * LANES;
            for (int l = l0; l < l0 + GROUP; l++) {
                const double z = (an * t[l]) * x[l] - bn * y[l];
                y[l] = x[l];
                x[l] = z;
                pn[l] = z;
            }
        }
    }
    return nd;
}

void cgl_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }

static size_t col_off(unsigned long nmax, unsigned long m)
{
    return (size_t)m * (nmax + 1) - ((size_t)m * (m - 1)) / 2;
}

/* Synthesis on the Gauss--Legendre grid of nlat = nmax_grd + 1 rings at one
 * radius r: the spectra of every ring, spec[(i * nspec + k) * 2 + {0, 1}] =
 * X_k with X_0 = s a_0, X_k = s (a_k - i b_k)/2 (k = 1..nmax), zero above,
 * s = mu/R, so that an unnormalised complex-to-real transform of length
 * nlon gives f_j = s (a_0 + sum_k a_k cos + b_k sin)(2 pi k j / nlon).
 * lat holds the nlat latitudes, north first; coefficients are m-major
 * triangles of degree nmax.  Returns 0, or 1 on invalid input. */
int cgl_shs_spectra(unsigned long nmax, const double *C, const double *S, double mu, double R,
                    double r, size_t nlat, const double *lat, size_t nspec, double *spec)
{
    if (nlat < nmax + 1 || nspec < nmax + 1) return 1;
    init_consts();
    const size_t nring = (nlat + 1) / 2;
    const size_t equator = (nlat % 2) ? nlat / 2 : (size_t)-1;
    const size_t nblk = (nring + LANES - 1) / LANES;
    const size_t ncs = (size_t)(nmax + 1) * (nmax + 2) / 2;
    double *rt = malloc((2 * nmax + 4) * sizeof(double)), *rit = malloc((2 * nmax + 4) * sizeof(double));
    double *dm = malloc((nmax + 1) * sizeof(double));
    tables(nmax, rt, rit, dm);
    /* coefficients times (R/r)^(n+1) (shs_rpows.c:16-37) */
    double *Cs = malloc(ncs * sizeof(double)), *Ss = malloc(ncs * sizeof(double));
    double *rp = malloc((nmax + 2) * sizeof(double));
    rp[0] = 1;
    for (unsigned long k = 1; k <= nmax + 1; k++) rp[k] = (r == R) ? 1.0 : rp[k - 1] * (R / r);
    for (unsigned long m = 0; m <= nmax; m++)
        for (unsigned long n = m; n <= nmax; n++) {
            size_t q = col_off(nmax, m) + n - m;
            Cs[q] = C[q] * rp[n + 1];
            Ss[q] = S[q] * rp[n + 1];
        }
    double *tt = malloc(nblk * LANES * sizeof(double));
    double *ps = malloc(nring * (nmax + 1) * sizeof(double));
    int64_t *ips = malloc(nring * (nmax + 1) * sizeof(int64_t));
    /* lumped sums of the ring pair, even and odd n + m: [m][ring][4] */
    double *lump = malloc((nmax + 1) * nblk * LANES * 4 * sizeof(double));
    for (size_t i = 0; i < nblk * LANES; i++) tt[i] = (i < nring) ? sin(lat[i]) : 0.0;
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < nring; i++)
        seeds(cos(lat[i]), dm, nmax, ps + i * (nmax + 1), ips + i * (nmax + 1));

#pragma omp parallel
    {
        const unsigned int csr = _mm_getcsr();
        _mm_setcsr(csr | 0x8040);  /* FTZ | DAZ */
        double *anm = malloc((nmax + 2) * sizeof(double)), *bnm = malloc((nmax + 2) * sizeof(double));
        double *pb = malloc((nmax + 2) * LANES * sizeof(double));
        unsigned long mlast = (unsigned long)-1;
#pragma omp for collapse(2) schedule(dynamic, 1)
        for (unsigned long m = 0; m <= nmax; m++)
            for (size_t b = 0; b < nblk; b++) {
                if (m != mlast) {
                    anm_bnm(nmax, m, rt, rit, anm, bnm);
                    mlast = m;
                }
                const double *Cm = Cs + col_off(nmax, m), *Sm = Ss + col_off(nmax, m);
                double x[LANES], y[LANES], t[LANES], ae[LANES], ao[LANES], be[LANES], bo[LANES];
                double xs[LANES];
                int64_t ixs[LANES];
                for (int l = 0; l < LANES; l++) {
                    const size_t i = b * LANES + l;
                    ae[l] = ao[l] = be[l] = bo[l] = 0;
                    t[l] = tt[i];
                    xs[l] = (m && i < nring) ? ps[i * (nmax + 1) + m - 1] : 0.0;
                    ixs[l] = (m && i < nring) ? ips[i * (nmax + 1) + m - 1] : 0;
                }
                const unsigned long nd = xblock(nmax, m, anm, bnm, t, xs, ixs, pb, x, y);
                for (unsigned long n = m; n < nd; n++) {
                    const double c = Cm[n - m], s = Sm[n - m], *pn = pb + (n - m) * LANES;
                    if ((n + m) % 2 == 0)
                        for (int l = 0; l < LANES; l++) { ae[l] += pn[l] * c; be[l] += pn[l] * s; }
                    else
                        for (int l = 0; l < LANES; l++) { ao[l] += pn[l] * c; bo[l] += pn[l] * s; }
                }
                unsigned long n = nd;
                /* two degrees per pass: one of each parity */
                if (n <= nmax && (n + m) % 2) {
                    const double an = anm[n], bn = bnm[n], c = Cm[n - m], s = Sm[n - m];
                    for (int l = 0; l < LANES; l++) {
                        const double p = (an * t[l]) * x[l] - bn * y[l];
                        y[l] = x[l];
                        x[l] = p;
                        ao[l] += p * c;
                        bo[l] += p * s;
                    }
                    n++;
                }
                for (; n + 1 <= nmax; n += 2) {
                    const double an0 = anm[n], bn0 = bnm[n], c0 = Cm[n - m], s0 = Sm[n - m];
                    const double an1 = anm[n + 1], bn1 = bnm[n + 1], c1 = Cm[n + 1 - m],
                                 s1 = Sm[n + 1 - m];
                    for (int l = 0; l < LANES; l++) {
                        const double p0 = (an0 * t[l]) * x[l] - bn0 * y[l];
                        const double p1 = (an1 * t[l]) * p0 - bn1 * x[l];
                        y[l] = p0;
                        x[l] = p1;
                        ae[l] += p0 * c0;
                        be[l] += p0 * s0;
                        ao[l] += p1 * c1;
                        bo[l] += p1 * s1;
                    }
                }
                if (n <= nmax) {
                    const double an = anm[n], bn = bnm[n], c = Cm[n - m], s = Sm[n - m];
                    for (int l = 0; l < LANES; l++) {
                        const double p = (an * t[l]) * x[l] - bn * y[l];
                        ae[l] += p * c;
                        be[l] += p * s;
                    }
                }
                double *o = lump + ((size_t)m * nblk + b) * LANES * 4;
                for (int l = 0; l < LANES; l++) {
                    o[4 * l] = ae[l];
                    o[4 * l + 1] = ao[l];
                    o[4 * l + 2] = be[l];
                    o[4 * l + 3] = bo[l];
                }
            }
        free(anm); free(bnm); free(pb);
        _mm_setcsr(csr);
    }
    const double sc = mu / R;
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < nring; i++) {
        double *zn = spec + i * nspec * 2, *zs = spec + (nlat - 1 - i) * nspec * 2;
        const int mirror = i != equator;
        for (size_t k = 0; k < nspec; k++) {
            double an = 0, bn = 0, as = 0, bs = 0;
            if (k <= nmax) {
                const double *o = lump + ((size_t)k * nblk * LANES + i) * 4;
                an = o[0] + o[1];
                bn = o[2] + o[3];
                as = o[0] - o[1];
                bs = o[2] - o[3];
            }
            const double h = k ? 0.5 * sc : sc;
            zn[2 * k] = h * an;
            zn[2 * k + 1] = -h * bn;
            if (mirror) {
                zs[2 * k] = h * as;
                zs[2 * k + 1] = -h * bs;
            }
        }
    }
    free(rt); free(rit); free(dm); free(Cs); free(Ss); free(rp); free(tt); free(ps); free(ips);
    free(lump);
    return 0;
}

/* Analysis on the Gauss--Legendre grid from the ring spectra
 * spec[(i * nspec + k) * 2 + {0, 1}] = F_k = sum_j f_j e^{-2 pi i j k / nlon}
 * (an unnormalised real-to-complex transform of each ring), with the
 * quadrature weights w, the grid radius r0 and the model's mu and R
 * (sha_point.c:313-322, 520-529, 716-969, 1088-1108).  C and S receive the
 * m-major triangles of degree nmax.  Returns 0, or 1 on invalid input. */
int cgl_sha_spectra(unsigned long nmax, size_t nlat, size_t nspec, const double *spec,
                    const double *lat, const double *w, double r0, double mu, double R,
                    double *C, double *S)
{
    if (nlat == 0 || nmax > nlat - 1 || nspec < nmax + 1) return 1;
    init_consts();
    const unsigned long nmax_grd = nlat - 1;
    const size_t nring = (nlat + 1) / 2;
    const size_t equator = (nmax_grd % 2 == 0) ? nlat / 2 : (size_t)-1;
    const size_t nblk = (nring + LANES - 1) / LANES, nr = nblk * LANES;
    double c = M_PI / (double)(nmax_grd + 1);
    c *= 1.0 / (4.0 * M_PI) * (r0 / mu);
    /* folded spectra [m][ring]: A + A2, A - A2, B + B2, B - B2 */
    double *amp = calloc((nmax + 1) * nr, sizeof(double)), *amm = calloc((nmax + 1) * nr, sizeof(double));
    double *bmp = calloc((nmax + 1) * nr, sizeof(double)), *bmm = calloc((nmax + 1) * nr, sizeof(double));
    double *tt = calloc(nr, sizeof(double));
    double *ps = malloc(nring * (nmax + 1) * sizeof(double));
    int64_t *ips = malloc(nring * (nmax + 1) * sizeof(int64_t));
    double *rt = malloc((2 * nmax + 4) * sizeof(double)), *rit = malloc((2 * nmax + 4) * sizeof(double));
    double *dm = malloc((nmax + 1) * sizeof(double));
    tables(nmax, rt, rit, dm);
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < nring; i++) {
        tt[i] = sin(lat[i]);
        seeds(cos(lat[i]), dm, nmax, ps + i * (nmax + 1), ips + i * (nmax + 1));
        const int symm = nlat > 1 && i != equator;
        const size_t is = nlat - 1 - i;
        const double cw = c * w[i], cw2 = symm ? c * w[is] : 0.0;
        const double *zn = spec + i * nspec * 2, *zs = spec + is * nspec * 2;
        for (unsigned long k = 0; k <= nmax; k++) {
            const double A = cw * zn[2 * k], B = -cw * zn[2 * k + 1];
            const double A2 = symm ? cw2 * zs[2 * k] : 0.0, B2 = symm ? -cw2 * zs[2 * k + 1] : 0.0;
            amp[k * nr + i] = A + A2;
            amm[k * nr + i] = A - A2;
            bmp[k * nr + i] = B + B2;
            bmm[k * nr + i] = B - B2;
        }
    }
#pragma omp parallel
    {
        const unsigned int csr = _mm_getcsr();
        _mm_setcsr(csr | 0x8040);  /* FTZ | DAZ */
        double *anm = malloc((nmax + 2) * sizeof(double)), *bnm = malloc((nmax + 2) * sizeof(double));
        double *pb = malloc((nmax + 2) * LANES * sizeof(double));
        double *accc = malloc((nmax + 1) * LANES * sizeof(double));
        double *accs = malloc((nmax + 1) * LANES * sizeof(double));
#pragma omp for schedule(dynamic, 1)
        for (unsigned long m = 0; m <= nmax; m++) {
            anm_bnm(nmax, m, rt, rit, anm, bnm);
            const size_t len = nmax + 1 - m;
            double *Cm = C + col_off(nmax, m), *Sm = S + col_off(nmax, m);
            memset(Cm, 0, len * sizeof(double));
            memset(Sm, 0, len * sizeof(double));
            memset(accc, 0, len * LANES * sizeof(double));
            memset(accs, 0, len * LANES * sizeof(double));
            for (size_t b = 0; b < nblk; b++) {
                double x[LANES], y[LANES], t[LANES], ap[LANES], am[LANES], bp[LANES], bm[LANES];
                double xs[LANES];
                int64_t ixs[LANES];
                for (int l = 0; l < LANES; l++) {
                    const size_t i = b * LANES + l;
                    t[l] = tt[i];
                    ap[l] = amp[m * nr + i];
                    am[l] = amm[m * nr + i];
                    bp[l] = bmp[m * nr + i];
                    bm[l] = bmm[m * nr + i];
                    xs[l] = (m && i < nring) ? ps[i * (nmax + 1) + m - 1] : 0.0;
                    ixs[l] = (m && i < nring) ? ips[i * (nmax + 1) + m - 1] : 0;
                }
                const unsigned long nd = xblock(nmax, m, anm, bnm, t, xs, ixs, pb, x, y);
                for (unsigned long n = m; n < nd; n++) {
                    const double *pn = pb + (n - m) * LANES;
                    double *ac = accc + (n - m) * LANES, *as = accs + (n - m) * LANES;
                    const double *fa = ((n + m) % 2 == 0) ? ap : am;
                    const double *fb = ((n + m) % 2 == 0) ? bp : bm;
                    for (int l = 0; l < LANES; l++) {
                        ac[l] += pn[l] * fa[l];
                        as[l] += pn[l] * fb[l];
                    }
                }
                unsigned long n = nd;
                for (; n <= nmax; n++) {
                    const double an = anm[n], bn = bnm[n];
                    double *ac = accc + (n - m) * LANES, *as = accs + (n - m) * LANES;
                    const double *fa = ((n + m) % 2 == 0) ? ap : am;
                    const double *fb = ((n + m) % 2 == 0) ? bp : bm;
                    for (int l = 0; l < LANES; l++) {
                        const double p = (an * t[l]) * x[l] - bn * y[l];
                        y[l] = x[l];
                        x[l] = p;
                        ac[l] += p * fa[l];
                        as[l] += p * fb[l];
                    }
                }
            }
            for (size_t q = 0; q < len; q++) {
                double sc = 0, ss = 0;
                for (int l = 0; l < LANES; l++) {
                    sc += accc[q * LANES + l];
                    ss += accs[q * LANES + l];
                }
                Cm[q] += sc;
                Sm[q] += ss;
            }
            if (m == 0)
                for (size_t q = 0; q < len; q++) Sm[q] = 0;
        }
        free(anm); free(bnm); free(pb); free(accc); free(accs);
        _mm_setcsr(csr);
    }
    /* rescale from r0 to R (shc_rescale.c:24-90) */
    if (r0 != R) {
        const double q = r0 / R;
        for (unsigned long m = 0; m <= nmax; m++) {
            double *Cm = C + col_off(nmax, m), *Sm = S + col_off(nmax, m);
            double f = 1.0;  /* (r0/R)^n by repeated products */
            for (unsigned long n = 0; n < m; n++) f *= q;
            for (unsigned long n = m; n <= nmax; n++) {
                Cm[n - m] *= f;
                Sm[n - m] *= f;
                f *= q;
            }
        }
    }
    free(amp); free(amm); free(bmp); free(bmm); free(tt); free(ps); free(ips);
    free(rt); free(rit); free(dm);
    return 0;
}
