// HIP kernels for point synthesis (SHS) and analysis (SHA) on gfx950.
//
// Mapping.  Lanes run over latitude rings (or scattered points), work-groups
// over orders m; the degree recurrence runs sequentially in each lane, with
// the recurrence coefficients (a_nm, b_nm) and the model coefficients
// (C_nm, S_nm) as wave-uniform scalar loads.  Equatorially symmetric rings are
// folded: one Legendre evaluation per ring pair, with even and odd (n + m)
// sums kept apart so that P_nm(-t) = (-1)^(n+m) P_nm(t) gives the mirror ring.
//
// X-numbers (extended exponent, radix 2^960) carry the sectorial seeds and the
// first tesseral terms of each column exactly as in CHarm
// (src/leg/leg_func_xnum.h:299-443); once both terms of a lane have a zero
// exponent that lane continues in ordinary doubles.  A wave switches to the
// tight all-double loop when every lane has switched.
#include <hip/hip_runtime.h>

#include "internal.h"

namespace charm_amd {

namespace {

constexpr double BIG = 0x1p960;
constexpr double BIGI = 0x1p-960;
constexpr double BIGS = 0x1p480;
constexpr double BIGSI = 0x1p-480;
// src/prec.h:301 (ROOT3 in the double build)
constexpr double ROOT3 = 1.73205080756887729353;

__device__ __forceinline__ size_t dcol_off(long nmax, long m)
{
    return (size_t)m * (size_t)(nmax + 1) - ((size_t)m * (size_t)(m - 1)) / 2;
}

__device__ __forceinline__ size_t tri_len_dev(long nmax)
{
    return ((size_t)(nmax + 1) * (size_t)(nmax + 2)) / 2;
}

// Keep |x| in [BIGSI, BIGS) (src/leg/leg_func_xnum.h:318-330)
__device__ __forceinline__ void xnorm(double &x, int &ix)
{
    double w = fabs(x);
    if (w >= BIGS) {
        x *= BIGI;
        ix += 1;
    } else if (w < BIGSI) {
        x *= BIG;
        ix -= 1;
    }
}

__device__ __forceinline__ bool wave_any(bool p)
{
    return __ballot(p) != 0;
}

__device__ __forceinline__ bool wave_all(bool p)
{
    return __ballot(!p) == 0;
}

// Column state of one lane: P_{n-1} = x 2^e and P_{n-2} = y 2^e.
//
// CHarm carries the first terms of a column as X-numbers (mantissa in
// [2^-480, 2^480) and exponent in units of 2^960, src/leg/leg_func_xnum.h)
// and switches to plain doubles once both terms have a zero exponent, i.e.
// once both true values lie in [2^-480, 2^480).  Scaling by a power of two is
// exact, so the same true values come out of the plain recurrence on scaled
// mantissas with a separate binary exponent e that is refreshed only every
// 8 degrees; a lane drops e (e = 0) at the same point CHarm leaves
// X-numbers.  P_n itself is ldexp(z, e), which rounds (also into the
// subnormal range, and to 0 far below it) exactly as CHarm's x * BIG^ix.
__device__ __forceinline__ void lane_renorm(double &x, double &y, int &e)
{
    if (e != 0) {
        int ex, ey;
        (void)frexp(x, &ex);
        (void)frexp(y, &ey);
        if (e + ex > -480 && e + ey > -480) {
            x = ldexp(x, e);
            y = ldexp(y, e);
            e = 0;
        } else if (ex > 256 || ex < -256) {
            x = ldexp(x, -ex);
            y = ldexp(y, -ex);
            e += ex;
        }
    }
}

// True |P| stays below 2^-1100 for the next 8 degrees (|a_nm t| < 2^7), so
// these terms are exactly zero in double precision.
__device__ __forceinline__ bool lane_deep(double x, double y, int e)
{
    int ex, ey;
    (void)frexp(x, &ex);
    (void)frexp(y, &ey);
    return e != 0 && e + max(ex, ey) < -1160;
}

__device__ __forceinline__ int uniform(int v)
{
    return __builtin_amdgcn_readfirstlane(v);
}

}  // namespace

// ---------------------------------------------------------------------------
// (a_nm, b_nm) table, m-major, degree nt (src/leg/leg_func_anm_bnm.c:14-65)
// ---------------------------------------------------------------------------
__global__ void ab_table_kernel(int nt, const double *__restrict__ r, const double *__restrict__ ri,
                                double2 *__restrict__ ab)
{
    const int m = blockIdx.x;
    double2 *col = ab + dcol_off(nt, m) - m;
    for (int n = m + threadIdx.x; n <= nt; n += blockDim.x) {
        double a = 0.0, b = 0.0;
        if (n == m + 1) {
            a = r[2 * m + 3];
        } else if (n >= m + 2) {
            double w = r[2 * n + 1] * ri[n - m] * ri[n + m];
            a = r[2 * n - 1] * w;
            b = r[n - m - 1] * r[n + m - 1] * ri[2 * n - 3] * w;
        }
        col[n] = make_double2(a, b);
    }
}

// ---------------------------------------------------------------------------
// Sectorial seeds P_mm, m = 1..nmax, as X-numbers, one lane per ring
// (src/leg/leg_func_prepare.c:18-69).  Row m-1 of the table holds P_mm.
// ---------------------------------------------------------------------------
__global__ void seed_kernel(int nring, const double *__restrict__ u, const double *__restrict__ dm,
                            int nmax, size_t sld, double *__restrict__ sx, int *__restrict__ six)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nring) return;
    const double uu = u[i];
    double x = ROOT3 * uu;
    int ix = 0;
    sx[i] = x;
    six[i] = 0;
    for (int n = 1; n < nmax; n++) {
        x = (dm[n] * uu) * x;
        xnorm(x, ix);
        sx[(size_t)n * sld + i] = x;
        six[(size_t)n * sld + i] = ix;
    }
}

// (R/r)^k, k = 0..kmax, by repeated products (src/shs/shs_rpows.c:16-37)
__global__ void rpows_kernel(int nrow, const double *__restrict__ ratio, int kmax, size_t rld,
                             double *__restrict__ rp)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nrow) return;
    const double q = ratio[i];
    double v = 1.0;
    rp[i] = v;
    for (int k = 1; k <= kmax; k++) {
        v = v * q;
        rp[(size_t)k * rld + i] = v;
    }
}

// ---------------------------------------------------------------------------
// SHS: lumped coefficients on grid rings
// (src/shs/shs_point_kernel.c:1014-1623 for DR = DLAT = DLON = 0)
//
// Block = 4 waves = 4 consecutive orders m0..m0+3 for one tile of 64*K ring
// pairs.  Results are staged in LDS and written as 64-byte runs per row.
// Blocks are numbered so that one XCD (blockIdx mod 8) owns every 8th order
// quad: all tiles of a column share that XCD's L2.
// ---------------------------------------------------------------------------
template <int K, bool UNIT>
__device__ __forceinline__ void shs_grid_item(const ShsGridArgs &A, int tile, int quad,
                                              double2 (*stage)[64 * K][4])
{
    const int lane = threadIdx.x & 63;
    const int wave = uniform(threadIdx.x >> 6);
    const int m0 = quad * 4;
    if (m0 >= A.mcount) return;
    const int m = m0 + wave;
    const int nmax = A.nmax;

    int ring[K];
    bool valid[K];
    double t[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
        ring[k] = tile * 64 * K + k * 64 + lane;
        valid[k] = ring[k] < A.nring;
        t[k] = valid[k] ? A.t[ring[k]] : 0.0;
    }

    double aN[K], bN[K], aS[K], bS[K];
#pragma unroll
    for (int k = 0; k < K; k++) aN[k] = bN[k] = aS[k] = bS[k] = 0.0;

    if (m <= nmax) {
        const double2 *__restrict__ abm = (const double2 *)A.ab + dcol_off(A.ab_nt, m) - m;
        const double2 *__restrict__ csm = (const double2 *)A.cs + dcol_off(nmax, m) - m;
        double x[K], y[K];
        int e[K];
        // even/odd (n + m) partial sums (UNIT) or north/south sums (!UNIT)
        double aE[K], aO[K], bE[K], bO[K];
        double rpN[K], rpS[K], qN[K], qS[K];
#pragma unroll
        for (int k = 0; k < K; k++) {
            aE[k] = aO[k] = bE[k] = bO[k] = 0.0;
            x[k] = y[k] = 0.0;
            e[k] = 0;
            rpN[k] = rpS[k] = qN[k] = qS[k] = 1.0;
            if (!UNIT && valid[k]) {
                const int rs = A.mirror[ring[k]] ? A.nlat - 1 - ring[k] : ring[k];
                qN[k] = A.ratio[ring[k]];
                qS[k] = A.ratio[rs];
                rpN[k] = A.rp[(size_t)(m + 1) * A.rld + ring[k]];
                rpS[k] = A.rp[(size_t)(m + 1) * A.rld + rs];
            }
        }
        // accumulate term n with Legendre value p; par = (n + m) & 1
#define SHS_ACC(k, p, c, par)                                              \
    do {                                                                   \
        if (UNIT) {                                                        \
            if ((par) == 0) {                                              \
                aE[k] = fma((p), (c).x, aE[k]);                            \
                bE[k] = fma((p), (c).y, bE[k]);                            \
            } else {                                                       \
                aO[k] = fma((p), (c).x, aO[k]);                            \
                bO[k] = fma((p), (c).y, bO[k]);                            \
            }                                                              \
        } else {                                                           \
            const double pn_ = rpN[k] * (p);                               \
            const double ps_ = ((par) == 0) ? rpS[k] * (p) : -(rpS[k] * (p)); \
            aE[k] = fma(pn_, (c).x, aE[k]);                                \
            bE[k] = fma(pn_, (c).y, bE[k]);                                \
            aO[k] = fma(ps_, (c).x, aO[k]);                                \
            bO[k] = fma(ps_, (c).y, bO[k]);                                \
            rpN[k] *= qN[k];                                               \
            rpS[k] *= qS[k];                                               \
        }                                                                  \
    } while (0)
        // advance the radius powers over terms that are exactly zero
#define SHS_SKIP(k)                                                        \
    do {                                                                   \
        if (!UNIT) {                                                       \
            rpN[k] *= qN[k];                                               \
            rpS[k] *= qS[k];                                               \
        }                                                                  \
    } while (0)

        int n;
        if (m == 0) {
            // zonal: P00 = 1, P10 = sqrt(3) t, then the ordinary recurrence
            const double2 c0 = csm[0];
#pragma unroll
            for (int k = 0; k < K; k++) SHS_ACC(k, 1.0, c0, 0);
            if (nmax >= 1) {
                const double2 c1 = csm[1];
#pragma unroll
                for (int k = 0; k < K; k++) {
                    const double p1 = ROOT3 * t[k];
                    SHS_ACC(k, p1, c1, 1);
                    x[k] = p1;
                    y[k] = 1.0;
                }
            }
            n = 2;
        } else {
            // sectorial seed P_mm (X-number from the seed table) and
            // semi-sectorial P_{m+1,m} = (a t) P_mm
            const double2 cmm = csm[m];
            const double2 ab1 = (m < nmax) ? abm[m + 1] : make_double2(0.0, 0.0);
            const double2 c1 = (m < nmax) ? csm[m + 1] : make_double2(0.0, 0.0);
#pragma unroll
            for (int k = 0; k < K; k++) {
                const double xs = valid[k] ? A.sx[(size_t)(m - 1) * A.sld + ring[k]] : 0.0;
                e[k] = valid[k] ? 960 * A.six[(size_t)(m - 1) * A.sld + ring[k]] : 0;
                SHS_ACC(k, ldexp(xs, e[k]), cmm, 0);
                if (m < nmax) {
                    y[k] = xs;
                    x[k] = (ab1.x * t[k]) * xs;
                    SHS_ACC(k, ldexp(x[k], e[k]), c1, 1);
                    lane_renorm(x[k], y[k], e[k]);
                }
            }
            n = m + 2;
        }

        // Tiles of 8 degrees; (n + m) is even at every tile start.  While a
        // lane of the wave still carries an exponent, tiles convert each term
        // with ldexp (or, when every lane is far below the double range, run
        // the bare recurrence); afterwards the wave stays in the tight loop.
        for (;;) {
            if (n + 7 > nmax) break;
            bool anyx = false, deep = true;
#pragma unroll
            for (int k = 0; k < K; k++) {
                anyx = anyx || (e[k] != 0);
                deep = deep && (!valid[k] || lane_deep(x[k], y[k], e[k]));
            }
            if (!wave_any(anyx)) break;
            if (wave_all(deep)) {
#pragma unroll
                for (int j = 0; j < 8; j++) {
                    const double2 abj = abm[n + j];
#pragma unroll
                    for (int k = 0; k < K; k++) {
                        const double z = fma(abj.x * t[k], x[k], -(abj.y * y[k]));
                        y[k] = x[k];
                        x[k] = z;
                        SHS_SKIP(k);
                    }
                }
            } else {
#pragma unroll
                for (int j = 0; j < 8; j++) {
                    const double2 abj = abm[n + j];
                    const double2 cj = csm[n + j];
#pragma unroll
                    for (int k = 0; k < K; k++) {
                        const double z = fma(abj.x * t[k], x[k], -(abj.y * y[k]));
                        y[k] = x[k];
                        x[k] = z;
                        SHS_ACC(k, ldexp(z, e[k]), cj, j & 1);
                    }
                }
            }
#pragma unroll
            for (int k = 0; k < K; k++) lane_renorm(x[k], y[k], e[k]);
            n += 8;
        }
        for (; n + 7 <= nmax; n += 8) {
#pragma unroll
            for (int j = 0; j < 8; j += 2) {
                const double2 ab0 = abm[n + j], ab1 = abm[n + j + 1];
                const double2 c0 = csm[n + j], c1 = csm[n + j + 1];
#pragma unroll
                for (int k = 0; k < K; k++) {
                    const double z = fma(ab0.x * t[k], x[k], -(ab0.y * y[k]));
                    SHS_ACC(k, z, c0, 0);
                    const double w = fma(ab1.x * t[k], z, -(ab1.y * x[k]));
                    SHS_ACC(k, w, c1, 1);
                    y[k] = z;
                    x[k] = w;
                }
            }
        }
        for (; n <= nmax; n++) {
            const double2 abj = abm[n];
            const double2 cj = csm[n];
            const int par = (n + m) & 1;
#pragma unroll
            for (int k = 0; k < K; k++) {
                const double z = fma(abj.x * t[k], x[k], -(abj.y * y[k]));
                y[k] = x[k];
                x[k] = z;
                SHS_ACC(k, ldexp(z, e[k]), cj, par);
            }
        }
#undef SHS_ACC
#undef SHS_SKIP
        if (m == 0) {
            // S_n0 does not exist; the zonal sums have no sine part
#pragma unroll
            for (int k = 0; k < K; k++) bE[k] = bO[k] = 0.0;
        }
#pragma unroll
        for (int k = 0; k < K; k++) {
            if (UNIT) {
                aN[k] = aE[k] + aO[k];
                bN[k] = bE[k] + bO[k];
                aS[k] = aE[k] - aO[k];
                bS[k] = bE[k] - bO[k];
            } else {
                aN[k] = aE[k];
                bN[k] = bE[k];
                aS[k] = aO[k];
                bS[k] = bO[k];
            }
        }
    }

    // stage: X_m = c (a_m - i b_m) mu/R with c = 1 (m = 0) or 1/2
    // (src/shs/shs_grd_fft_lc.c:24-44,72; shs_grd_fft.c:110-112)
    const double sc = A.fft_layout ? ((m == 0) ? A.mur : 0.5 * A.mur) : 1.0;
    const double sb = A.fft_layout ? -sc : sc;
#pragma unroll
    for (int k = 0; k < K; k++) {
        stage[0][k * 64 + lane][wave] = make_double2(sc * aN[k], sb * bN[k]);
        stage[1][k * 64 + lane][wave] = make_double2(sc * aS[k], sb * bS[k]);
    }
    __syncthreads();
    double2 *out = (double2 *)A.out;
    for (int e = threadIdx.x; e < 2 * 64 * K * 4; e += 256) {
        const int col = e & 3, rowsel = e >> 2;
        const int ns = rowsel / (64 * K), rl = rowsel % (64 * K);
        const int i = tile * 64 * K + rl;
        const int mm = m0 + col;
        if (i >= A.nring || mm >= A.mcount) continue;
        int row = i;
        if (ns) {
            if (!A.mirror[i]) continue;
            row = A.nlat - 1 - i;
        }
        out[(size_t)row * A.ldo + mm] = stage[ns][rl][col];
    }
}

template <int K, bool UNIT>
__global__ __launch_bounds__(256) void shs_grid_kernel(ShsGridArgs A)
{
    __shared__ double2 stage[2][64 * K][4];
    const int b = blockIdx.x;
    shs_grid_item<K, UNIT>(A, (b >> 3) % A.ntile, ((b >> 3) / A.ntile) * 8 + (b & 7), stage);
}

// Persistent form: blocks pull (request, ring tile, order quad) items from a
// list sorted by decreasing cost until the list is exhausted.
template <int K>
__global__ __launch_bounds__(256) void shs_batch_kernel(const ShsGridArgs *__restrict__ reqs,
                                                        const int2 *__restrict__ items, int nitems,
                                                        int *counter)
{
    __shared__ double2 stage[2][64 * K][4];
    __shared__ int cur;
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) cur = atomicAdd(counter, 1);
        __syncthreads();
        const int it = cur;
        if (it >= nitems) break;
        const int2 w = items[it];
        const ShsGridArgs A = reqs[w.x];
        if (A.unit)
            shs_grid_item<K, true>(A, w.y >> 16, w.y & 0xFFFF, stage);
        else
            shs_grid_item<K, false>(A, w.y >> 16, w.y & 0xFFFF, stage);
    }
}

// ---------------------------------------------------------------------------
// Longitude sums for grids that do not admit an FFT (custom longitudes):
// f_ij = mu/R sum_m (a_im cos m lon_j + b_im sin m lon_j)
// (src/shs/shs_grd_lr.c; evaluated directly)
// ---------------------------------------------------------------------------
__global__ void lon_direct_kernel(int nlat, int nmax, size_t nlon, const double *__restrict__ lon,
                                  const double2 *__restrict__ lc, size_t ldl, double mur,
                                  double *__restrict__ f)
{
    const size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y;
    if (j >= nlon || i >= nlat) return;
    const double l = lon[j];
    double s = 0.0;
    for (int m = 0; m <= nmax; m++) {
        double sn, cn;
        sincos((double)m * l, &sn, &cn);
        const double2 ab = lc[(size_t)i * ldl + m];
        s = s + (ab.x * cn + ab.y * sn);
    }
    f[(size_t)i * nlon + j] = mur * s;
}

// ---------------------------------------------------------------------------
// SHA: C_nm = sum over ring pairs of P_nm(t_i) (A_i +- A_mirror)
// (src/sha/sha_point.c:490-980)
//
// Block = 4 waves = orders m0..m0+3 for one tile of 64*K ring pairs.  Each lane
// sums its K ring pairs in registers; every 8 degrees the 64 lanes' partial
// sums are transposed through LDS and reduced in a fixed order.  One partial
// per (tile, n, m) goes to global memory; sha_reduce_kernel adds the tiles.
// ---------------------------------------------------------------------------
namespace {
constexpr int RED_LD = 66;  // padded row (16-byte aligned) of the per-wave transpose buffer

__device__ __forceinline__ double dpp_quad_xor1(double v)
{
    int2 w = __builtin_bit_cast(int2, v);
    w.x = __builtin_amdgcn_mov_dpp(w.x, 0xB1, 0xF, 0xF, false);  // quad_perm [1,0,3,2]
    w.y = __builtin_amdgcn_mov_dpp(w.y, 0xB1, 0xF, 0xF, false);
    return __builtin_bit_cast(double, w);
}

__device__ __forceinline__ double dpp_quad_xor2(double v)
{
    int2 w = __builtin_bit_cast(int2, v);
    w.x = __builtin_amdgcn_mov_dpp(w.x, 0x4E, 0xF, 0xF, false);  // quad_perm [2,3,0,1]
    w.y = __builtin_amdgcn_mov_dpp(w.y, 0x4E, 0xF, 0xF, false);
    return __builtin_bit_cast(double, w);
}

// Cross-lane reduction of 16 values per lane (8 degrees x (C, S)) over the 64
// lanes of a wave, software-pipelined across tiles: a tile's values are
// written to LDS after the tile is computed, read back at the start of the
// next tile and summed after that tile's arithmetic, so the LDS latency hides
// behind the recurrence.  Lane l sums value l/4 over 16 lanes; the 4 lanes of
// a quad then combine through DPP.  The order of additions is fixed.
struct WaveReducer {
    double *red;      // [16][RED_LD] doubles, this wave's slice of LDS
    double *dst;      // destination of the pending tile (nullptr: none)
    int cnt;          // degrees in the pending tile
    double r[16];     // pending values read back from LDS
    int lane;

    __device__ __forceinline__ void begin()
    {
        if (dst) {
            const int j = lane >> 2, q = lane & 3;
            const double2 *row = (const double2 *)(red + j * RED_LD + q * 16);
#pragma unroll
            for (int i = 0; i < 8; i++) {
                const double2 w = row[i];
                r[2 * i] = w.x;
                r[2 * i + 1] = w.y;
            }
        }
    }
    __device__ __forceinline__ void finish()
    {
        if (dst) {
            double s = r[0];
#pragma unroll
            for (int i = 1; i < 16; i++) s += r[i];
            s += dpp_quad_xor1(s);
            s += dpp_quad_xor2(s);
            const int j = lane >> 2;
            if ((lane & 3) == 0 && j < 2 * cnt) dst[j] = s;
            dst = nullptr;
        }
    }
    __device__ __forceinline__ void post(const double v[16], int c, double *d)
    {
#pragma unroll
        for (int j = 0; j < 16; j++) red[j * RED_LD + lane] = v[j];
        dst = d;
        cnt = c;
    }
};
}  // namespace

template <int K>
struct ShaLds {
    static constexpr int STAGE = 2 * 64 * K * 4 * 2;  // doubles
    static constexpr int RED = 4 * 16 * RED_LD;       // doubles
    static constexpr int SIZE = STAGE > RED ? STAGE : RED;
};

template <int K>
__device__ __forceinline__ void sha_item(const ShaArgs &A, int tile, int quad, double *lds)
{
    const int lane = threadIdx.x & 63;
    const int wave = uniform(threadIdx.x >> 6);
    const int m0 = quad * 4;
    if (m0 > A.nmax) return;
    const int m = m0 + wave;
    const int nmax = A.nmax;

    // stage the weighted spectra A = cw Re F, B = -cw Im F of the tile's rows
    // for orders m0..m0+3 (64-byte runs per row)
    double2 *stage = (double2 *)lds;
    const double2 *spec = (const double2 *)A.spec;
    for (int e = threadIdx.x; e < 2 * 64 * K * 4; e += 256) {
        const int col = e & 3, rowsel = e >> 2;
        const int ns = rowsel / (64 * K), rl = rowsel % (64 * K);
        const int i = tile * 64 * K + rl;
        const int mm = m0 + col;
        double2 v = make_double2(0.0, 0.0);
        if (i < A.nring && mm <= nmax) {
            int row = i;
            if (ns) row = A.mirror[i] ? A.nlat - 1 - i : -1;
            if (row >= 0) {
                const double2 F = spec[(size_t)row * A.lds + mm];
                const double cw = A.cw[row];
                v = make_double2(cw * F.x, -cw * F.y);
            }
        }
        stage[e] = v;
    }
    __syncthreads();

    int ring[K];
    double t[K], ampC[K], ammC[K], ampS[K], ammS[K];
    bool valid[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
        ring[k] = tile * 64 * K + k * 64 + lane;
        valid[k] = ring[k] < A.nring;
        t[k] = valid[k] ? A.t[ring[k]] : 0.0;
        const double2 N = stage[((0 * 64 * K) + k * 64 + lane) * 4 + wave];
        const double2 S = stage[((1 * 64 * K) + k * 64 + lane) * 4 + wave];
        // fold (src/sha/sha_point.c:728-748); rows without a mirror hold S = 0
        ampC[k] = N.x + S.x;
        ammC[k] = N.x - S.x;
        ampS[k] = N.y + S.y;
        ammS[k] = N.y - S.y;
    }
    __syncthreads();
    if (m > nmax) return;

    const double2 *__restrict__ abm = (const double2 *)A.ab + dcol_off(A.ab_nt, m) - m;
    double *dst = (double *)((double2 *)A.part + (size_t)tile * tri_len_dev(nmax) + dcol_off(nmax, m) - m);
    WaveReducer rd;
    rd.red = lds + wave * 16 * RED_LD;
    rd.dst = nullptr;
    rd.cnt = 0;
    rd.lane = lane;

    double x[K], y[K];
    int e[K];
    double v[16];
#pragma unroll
    for (int j = 0; j < 16; j++) v[j] = 0.0;
#pragma unroll
    for (int k = 0; k < K; k++) {
        x[k] = y[k] = 0.0;
        e[k] = 0;
    }

    // degrees m and m + 1
    if (m == 0) {
#pragma unroll
        for (int k = 0; k < K; k++) {
            v[0] += ampC[k];
            if (nmax >= 1) {
                const double p1 = ROOT3 * t[k];
                v[2] = fma(p1, ammC[k], v[2]);
                x[k] = p1;
                y[k] = 1.0;
            }
        }
    } else {
        const double2 ab1 = (m < nmax) ? abm[m + 1] : make_double2(0.0, 0.0);
#pragma unroll
        for (int k = 0; k < K; k++) {
            const double xs = valid[k] ? A.sx[(size_t)(m - 1) * A.sld + ring[k]] : 0.0;
            e[k] = valid[k] ? 960 * A.six[(size_t)(m - 1) * A.sld + ring[k]] : 0;
            const double p0 = ldexp(xs, e[k]);
            v[0] = fma(p0, ampC[k], v[0]);
            v[1] = fma(p0, ampS[k], v[1]);
            if (m < nmax) {
                y[k] = xs;
                x[k] = (ab1.x * t[k]) * xs;
                const double p1 = ldexp(x[k], e[k]);
                v[2] = fma(p1, ammC[k], v[2]);
                v[3] = fma(p1, ammS[k], v[3]);
                lane_renorm(x[k], y[k], e[k]);
            }
        }
    }
    rd.post(v, (m < nmax) ? 2 : 1, dst + 2 * m);

    int n = m + 2;
    for (;;) {
        if (n + 7 > nmax) break;
        bool anyx = false, deep = true;
#pragma unroll
        for (int k = 0; k < K; k++) {
            anyx = anyx || (e[k] != 0);
            deep = deep && (!valid[k] || lane_deep(x[k], y[k], e[k]));
        }
        if (!wave_any(anyx)) break;
        rd.begin();
        if (wave_all(deep)) {
            // every term of the tile is exactly zero
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const double2 abj = abm[n + j];
#pragma unroll
                for (int k = 0; k < K; k++) {
                    const double z = fma(abj.x * t[k], x[k], -(abj.y * y[k]));
                    y[k] = x[k];
                    x[k] = z;
                }
            }
            rd.finish();
            if (lane < 16) dst[2 * n + lane] = 0.0;
        } else {
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const double2 abj = abm[n + j];
                double c = 0.0, sv = 0.0;
#pragma unroll
                for (int k = 0; k < K; k++) {
                    const double z = fma(abj.x * t[k], x[k], -(abj.y * y[k]));
                    y[k] = x[k];
                    x[k] = z;
                    const double p = ldexp(z, e[k]);
                    c = fma(p, (j & 1) ? ammC[k] : ampC[k], c);
                    sv = fma(p, (j & 1) ? ammS[k] : ampS[k], sv);
                }
                v[2 * j] = c;
                v[2 * j + 1] = sv;
            }
            rd.finish();
            rd.post(v, 8, dst + 2 * n);
        }
#pragma unroll
        for (int k = 0; k < K; k++) lane_renorm(x[k], y[k], e[k]);
        n += 8;
    }
    for (; n + 7 <= nmax; n += 8) {
        rd.begin();
#pragma unroll
        for (int j = 0; j < 8; j += 2) {
            const double2 ab0 = abm[n + j], ab1 = abm[n + j + 1];
            double c0 = 0.0, s0 = 0.0, c1 = 0.0, s1 = 0.0;
#pragma unroll
            for (int k = 0; k < K; k++) {
                const double z = fma(ab0.x * t[k], x[k], -(ab0.y * y[k]));
                c0 = fma(z, ampC[k], c0);
                s0 = fma(z, ampS[k], s0);
                const double w = fma(ab1.x * t[k], z, -(ab1.y * x[k]));
                c1 = fma(w, ammC[k], c1);
                s1 = fma(w, ammS[k], s1);
                y[k] = z;
                x[k] = w;
            }
            v[2 * j] = c0;
            v[2 * j + 1] = s0;
            v[2 * j + 2] = c1;
            v[2 * j + 3] = s1;
        }
        rd.finish();
        rd.post(v, 8, dst + 2 * n);
    }
    if (n <= nmax) {
        const int cnt = nmax - n + 1;
        rd.begin();
#pragma unroll
        for (int j = 0; j < 8; j++) {
            double c = 0.0, sv = 0.0;
            if (j < cnt) {
                const double2 abj = abm[n + j];
#pragma unroll
                for (int k = 0; k < K; k++) {
                    const double z = fma(abj.x * t[k], x[k], -(abj.y * y[k]));
                    y[k] = x[k];
                    x[k] = z;
                    const double p = ldexp(z, e[k]);
                    c = fma(p, (j & 1) ? ammC[k] : ampC[k], c);
                    sv = fma(p, (j & 1) ? ammS[k] : ampS[k], sv);
                }
            }
            v[2 * j] = c;
            v[2 * j + 1] = sv;
        }
        rd.finish();
        rd.post(v, cnt, dst + 2 * n);
    }
    rd.begin();
    rd.finish();
}

template <int K>
__global__ __launch_bounds__(256) void sha_kernel(ShaArgs A)
{
    __shared__ __align__(16) double lds[ShaLds<K>::SIZE];
    const int b = blockIdx.x;
    sha_item<K>(A, (b >> 3) % A.ntile, ((b >> 3) / A.ntile) * 8 + (b & 7), lds);
}

template <int K>
__global__ __launch_bounds__(256) void sha_batch_kernel(const ShaArgs *__restrict__ reqs,
                                                        const int2 *__restrict__ items, int nitems,
                                                        int *counter)
{
    __shared__ __align__(16) double lds[ShaLds<K>::SIZE];
    __shared__ int cur;
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) cur = atomicAdd(counter, 1);
        __syncthreads();
        const int it = cur;
        if (it >= nitems) break;
        const int2 w = items[it];
        const ShaArgs A = reqs[w.x];
        sha_item<K>(A, w.y >> 16, w.y & 0xFFFF, lds);
    }
}

// Sum the tile partials in tile order, rescale by scale[n] (if given) and
// write the coefficients of a model of degree nmax_out (zeros above nmax)
// (src/sha/sha_point.c:381,1088-1108; src/shc/shc_rescale.c:24-90).
__global__ void sha_reduce_kernel(int nmax, int ntile, const double2 *__restrict__ part,
                                  int nmax_out, const double *__restrict__ scale,
                                  double *__restrict__ c, double *__restrict__ s, size_t stride)
{
    const int m = blockIdx.y;
    const size_t P = tri_len_dev(nmax);
    const size_t out0 = dcol_off(nmax_out, m) - m;
    for (int n = m + blockIdx.x * blockDim.x + threadIdx.x; n <= nmax_out;
         n += gridDim.x * blockDim.x) {
        double vc = 0.0, vs = 0.0;
        if (m <= nmax && n <= nmax) {
            const size_t q = dcol_off(nmax, m) + (n - m);
            for (int tl = 0; tl < ntile; tl++) {
                const double2 p = part[(size_t)tl * P + q];
                vc += p.x;
                vs += p.y;
            }
            if (scale) {
                vc *= scale[n];
                vs *= scale[n];
            }
            if (m == 0) vs = 0.0;
        }
        c[(out0 + n) * stride] = vc;
        s[(out0 + n) * stride] = vs;
    }
}

// ---------------------------------------------------------------------------
// SHS at scattered points (src/shs/shs_point_sctr.c:294-687,
// shs_sctr_mulc.c:36-57).  One lane per point, sequential over the orders of
// one chunk; the sectorial seed and (R/r)^(m+1) advance with m in the lane.
// ---------------------------------------------------------------------------
template <bool UNIT>
__global__ __launch_bounds__(256) void sctr_kernel(SctrArgs A)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int chunk = blockIdx.y;
    const int mb = A.mchunk[chunk], me = A.mchunk[chunk + 1];
    const bool valid = i < A.npoint;
    const double t = valid ? A.lat_t[i] : 0.0;
    const double u = valid ? A.lat_u[i] : 0.0;
    const double lon = valid ? A.lon[i] : 0.0;
    const double q = (!UNIT && valid) ? A.R / A.r[i] : 1.0;
    const int nmax = A.nmax;

    // P_{mb,mb} (for mb >= 1) as an X-number, and (R/r)^(mb+1)
    double xs = ROOT3 * u;
    int ixs = 0;
    for (int mm = 1; mm < mb; mm++) {
        xs = (A.dm[mm] * u) * xs;
        xnorm(xs, ixs);
    }
    double rpc = 1.0;
    if (!UNIT)
        for (int k = 1; k <= mb + 1; k++) rpc = rpc * q;

    double f = 0.0;
    for (int m = mb; m < me; m++) {
        const double2 *__restrict__ abm = (const double2 *)A.ab + dcol_off(A.ab_nt, m) - m;
        const double2 *__restrict__ csm = (const double2 *)A.cs + dcol_off(nmax, m) - m;
        double a = 0.0, b = 0.0, rp = rpc;
        double x = 0.0, y = 0.0;
        int e = 0;
#define SC_ACC(p, c)                              \
    do {                                          \
        const double pq_ = UNIT ? (p) : rp * (p); \
        a = fma(pq_, (c).x, a);                   \
        b = fma(pq_, (c).y, b);                   \
        if (!UNIT) rp *= q;                       \
    } while (0)
        int n;
        if (m == 0) {
            SC_ACC(1.0, csm[0]);
            if (nmax >= 1) {
                x = ROOT3 * t;
                y = 1.0;
                SC_ACC(x, csm[1]);
            }
            n = 2;
        } else {
            e = valid ? 960 * ixs : 0;
            SC_ACC(ldexp(xs, e), csm[m]);
            if (m < nmax) {
                y = xs;
                x = (abm[m + 1].x * t) * xs;
                SC_ACC(ldexp(x, e), csm[m + 1]);
                lane_renorm(x, y, e);
            }
            n = m + 2;
        }
        for (;;) {
            if (n + 7 > nmax || !wave_any(e != 0)) break;
            if (wave_all(!valid || lane_deep(x, y, e))) {
#pragma unroll
                for (int j = 0; j < 8; j++) {
                    const double2 ab = abm[n + j];
                    const double z = fma(ab.x * t, x, -(ab.y * y));
                    y = x;
                    x = z;
                    if (!UNIT) rp *= q;
                }
            } else {
#pragma unroll
                for (int j = 0; j < 8; j++) {
                    const double2 ab = abm[n + j];
                    const double z = fma(ab.x * t, x, -(ab.y * y));
                    y = x;
                    x = z;
                    SC_ACC(ldexp(z, e), csm[n + j]);
                }
            }
            lane_renorm(x, y, e);
            n += 8;
        }
        for (; n + 7 <= nmax; n += 8) {
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const double2 ab = abm[n + j];
                const double z = fma(ab.x * t, x, -(ab.y * y));
                SC_ACC(z, csm[n + j]);
                y = x;
                x = z;
            }
        }
        for (; n <= nmax; n++) {
            const double2 ab = abm[n];
            const double z = fma(ab.x * t, x, -(ab.y * y));
            y = x;
            x = z;
            SC_ACC(ldexp(z, e), csm[n]);
        }
#undef SC_ACC
        if (m == 0) b = 0.0;
        double sn, cn;
        sincos((double)m * lon, &sn, &cn);
        f = f + (a * cn + b * sn);
        // advance the seed to P_{m+1,m+1} and the radius power
        if (m >= 1) {
            xs = (A.dm[m] * u) * xs;
            xnorm(xs, ixs);
        }
        if (!UNIT) rpc = rpc * q;
    }
    if (valid) A.part[(size_t)chunk * A.npoint + i] = f;
}

__global__ void sctr_reduce_kernel(size_t npoint, int nchunk, const double *__restrict__ part,
                                   double mur, double *__restrict__ f)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npoint) return;
    double s = part[i];
    for (int c = 1; c < nchunk; c++) s += part[(size_t)c * npoint + i];
    f[i] = mur * s;
}

__global__ void sincos_kernel(size_t n, const double *__restrict__ lat, double *__restrict__ t,
                              double *__restrict__ u)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) sincos(lat[i], &t[i], &u[i]);
}

// misc_is_nearly_equal (src/misc/misc_is_nearly_equal.c:16-32) for every
// x[i] against ref; clears *flag if any value differs
__global__ void all_near_kernel(size_t n, const double *__restrict__ x, double ref, double eps,
                                int *flag)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double a = x[i], aa = fabs(a), ba = fabs(ref), d = fabs(a - ref), s = aa + ba;
    bool ok;
    if (a == ref)
        ok = true;
    else if (a == 0.0 || ref == 0.0 || s == aa || s == ba)
        ok = d <= eps;
    else
        ok = d / fmax(aa, ba) <= eps;
    if (!ok) *flag = 0;
}

__global__ void pack_cs_kernel(const double *__restrict__ c, const double *__restrict__ s,
                               size_t n, double2 *__restrict__ cs)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) cs[i] = make_double2(c[i], s[i]);
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
static inline hipStream_t S(void *s)
{
    return (hipStream_t)s;
}

int launch_ab_table(int nt, const double *r, const double *ri, void *ab, void *stream)
{
    hipLaunchKernelGGL(ab_table_kernel, dim3(nt + 1), dim3(256), 0, S(stream), nt, r, ri,
                       (double2 *)ab);
    return (int)hipGetLastError();
}

int launch_seeds(int nring, const double *u, const double *dm, int nmax, size_t sld, double *sx,
                 int *six, void *stream)
{
    if (nring <= 0 || nmax <= 0) return 0;
    hipLaunchKernelGGL(seed_kernel, dim3((nring + 63) / 64), dim3(64), 0, S(stream), nring, u, dm,
                       nmax, sld, sx, six);
    return (int)hipGetLastError();
}

int launch_rpows(int nrow, const double *ratio, int kmax, size_t rld, double *rp, void *stream)
{
    if (nrow <= 0) return 0;
    hipLaunchKernelGGL(rpows_kernel, dim3((nrow + 63) / 64), dim3(64), 0, S(stream), nrow, ratio,
                       kmax, rld, rp);
    return (int)hipGetLastError();
}

int ring_k(int nring, int kmin, int kmax)
{
    // rings per lane: the fewest idle lanes in the last tile, then the most
    // rings per lane
    int best = kmax;
    long best_cap = -1;
    for (int k = kmax; k >= kmin; k--) {
        const long tiles = (nring + 64L * k - 1) / (64L * k);
        const long cap = tiles * 64L * k;
        if (best_cap < 0 || cap < best_cap) {
            best_cap = cap;
            best = k;
        }
    }
    return best;
}

template <int K>
static void shs_launch_k(const ShsGridArgs &a, unsigned nblk, hipStream_t s)
{
    if (a.unit)
        hipLaunchKernelGGL((shs_grid_kernel<K, true>), dim3(nblk), dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL((shs_grid_kernel<K, false>), dim3(nblk), dim3(256), 0, s, a);
}

int launch_shs_grid(const ShsGridArgs &a, void *stream)
{
    const int nquad = (a.mcount + 3) / 4;
    const int nqg = (nquad + 7) / 8;
    const unsigned nblk = 8u * (unsigned)nqg * (unsigned)a.ntile;
    switch (a.k) {
    case 1: shs_launch_k<1>(a, nblk, S(stream)); break;
    case 3: shs_launch_k<3>(a, nblk, S(stream)); break;
    default: shs_launch_k<2>(a, nblk, S(stream)); break;
    }
    return (int)hipGetLastError();
}

int launch_lon_direct(int nlat, int nmax, size_t nlon, const double *lon, const void *lc,
                      size_t ldl, double mur, double *f, void *stream)
{
    dim3 grid((unsigned)((nlon + 127) / 128), (unsigned)nlat);
    hipLaunchKernelGGL(lon_direct_kernel, grid, dim3(128), 0, S(stream), nlat, nmax, nlon, lon,
                       (const double2 *)lc, ldl, mur, f);
    return (int)hipGetLastError();
}

int launch_sha(const ShaArgs &a, void *stream)
{
    const int nquad = (a.nmax + 1 + 3) / 4;
    const int nqg = (nquad + 7) / 8;
    const unsigned nblk = 8u * (unsigned)nqg * (unsigned)a.ntile;
    switch (a.k) {
    case 2: hipLaunchKernelGGL((sha_kernel<2>), dim3(nblk), dim3(256), 0, S(stream), a); break;
    case 3: hipLaunchKernelGGL((sha_kernel<3>), dim3(nblk), dim3(256), 0, S(stream), a); break;
    default: hipLaunchKernelGGL((sha_kernel<4>), dim3(nblk), dim3(256), 0, S(stream), a); break;
    }
    return (int)hipGetLastError();
}

int launch_sha_reduce(int nmax, int ntile, const void *part, int nmax_out, const double *scale,
                      double *c, double *s, size_t stride, void *stream)
{
    const unsigned gx = (unsigned)((nmax_out + 1 + 255) / 256);
    hipLaunchKernelGGL(sha_reduce_kernel, dim3(gx, nmax_out + 1), dim3(256), 0, S(stream), nmax,
                       ntile, (const double2 *)part, nmax_out, scale, c, s, stride);
    return (int)hipGetLastError();
}

int launch_sctr(const SctrArgs &a, void *stream)
{
    dim3 grid((unsigned)((a.npoint + 255) / 256), (unsigned)a.nchunk);
    if (a.unit)
        hipLaunchKernelGGL((sctr_kernel<true>), grid, dim3(256), 0, S(stream), a);
    else
        hipLaunchKernelGGL((sctr_kernel<false>), grid, dim3(256), 0, S(stream), a);
    return (int)hipGetLastError();
}

int launch_sctr_reduce(size_t npoint, int nchunk, const double *part, double mur, double *f,
                       void *stream)
{
    hipLaunchKernelGGL(sctr_reduce_kernel, dim3((unsigned)((npoint + 255) / 256)), dim3(256), 0,
                       S(stream), npoint, nchunk, part, mur, f);
    return (int)hipGetLastError();
}

int launch_sincos(size_t n, const double *lat, double *t, double *u, void *stream)
{
    hipLaunchKernelGGL(sincos_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, S(stream), n,
                       lat, t, u);
    return (int)hipGetLastError();
}

int launch_all_near(size_t n, const double *x, double ref, double eps, int *result, void *stream)
{
    int *flag = nullptr;
    hipError_t e = hipMallocAsync((void **)&flag, sizeof(int), S(stream));
    if (e != hipSuccess) return (int)e;
    const int one = 1;
    e = hipMemcpyAsync(flag, &one, sizeof(int), hipMemcpyHostToDevice, S(stream));
    if (e == hipSuccess) {
        hipLaunchKernelGGL(all_near_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                           S(stream), n, x, ref, eps, flag);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(result, flag, sizeof(int), hipMemcpyDeviceToHost, S(stream));
    if (e == hipSuccess) e = hipStreamSynchronize(S(stream));
    hipFreeAsync(flag, S(stream));
    return (int)e;
}

// (C, S) pairs of a model of degree ns, cut to degree nmax (m-major order)
__global__ void cs_truncate_kernel(const double2 *__restrict__ src, int ns, double2 *__restrict__ dst,
                                   int nmax)
{
    const int m = blockIdx.y;
    const double2 *sc = src + dcol_off(ns, m) - m;
    double2 *dc = dst + dcol_off(nmax, m) - m;
    for (int n = m + blockIdx.x * blockDim.x + threadIdx.x; n <= nmax; n += gridDim.x * blockDim.x)
        dc[n] = sc[n];
}

int launch_cs_truncate(const double *src, unsigned long ns, double *dst, unsigned long nmax,
                       void *stream)
{
    const unsigned gx = (unsigned)((nmax + 1 + 255) / 256);
    hipLaunchKernelGGL(cs_truncate_kernel, dim3(gx, (unsigned)nmax + 1), dim3(256), 0, S(stream),
                       (const double2 *)src, (int)ns, (double2 *)dst, (int)nmax);
    return (int)hipGetLastError();
}

__global__ void zero_ints_kernel(int *p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

int launch_zero_ints(int *p, int n, void *stream)
{
    hipLaunchKernelGGL(zero_ints_kernel, dim3((n + 255) / 256), dim3(256), 0, S(stream), p, n);
    return (int)hipGetLastError();
}

constexpr int BATCH_SHS_K = 2;
constexpr int BATCH_SHA_K = 3;

int batch_ring_k(int kind)
{
    return kind == 0 ? BATCH_SHS_K : BATCH_SHA_K;
}

int batch_grid_size(int kind)
{
    int dev = 0, ncu = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    hipError_t e = (kind == 0)
        ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, shs_batch_kernel<BATCH_SHS_K>, 256, 0)
        : hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, sha_batch_kernel<BATCH_SHA_K>, 256, 0);
    if (e != hipSuccess || per < 1) per = 1;
    return ncu * per;
}

int launch_shs_batch(const void *reqs, const void *items, int nitems, int *counter, int grid,
                     void *stream)
{
    if (nitems <= 0) return 0;
    hipLaunchKernelGGL((shs_batch_kernel<BATCH_SHS_K>), dim3(grid), dim3(256), 0, S(stream),
                       (const ShsGridArgs *)reqs, (const int2 *)items, nitems, counter);
    return (int)hipGetLastError();
}

int launch_sha_batch(const void *reqs, const void *items, int nitems, int *counter, int grid,
                     void *stream)
{
    if (nitems <= 0) return 0;
    hipLaunchKernelGGL((sha_batch_kernel<BATCH_SHA_K>), dim3(grid), dim3(256), 0, S(stream),
                       (const ShaArgs *)reqs, (const int2 *)items, nitems, counter);
    return (int)hipGetLastError();
}

int launch_pack_cs(const double *c, const double *s, size_t n, void *cs, void *stream)
{
    hipLaunchKernelGGL(pack_cs_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, S(stream),
                       c, s, n, (double2 *)cs);
    return (int)hipGetLastError();
}

}  // namespace charm_amd
