// HIP kernels for point synthesis (SHS) and analysis (SHA) on gfx950.
//
// Mapping.  Lanes run over latitude rings (or scattered points), work-groups
// over orders m; the degree recurrence runs sequentially in each lane.  The
// per-degree column data of an order (the recurrence factor alpha_nm and, for
// synthesis, the scaled coefficients) are wave-uniform: each wave streams its
// column through a 64-entry LDS window that is refilled from registers loaded
// one window ahead, so the global loads never sit on the critical path.
// Equatorially symmetric rings are folded: one Legendre evaluation per ring
// pair, with even and odd (n + m) sums kept apart so that
// P_nm(-t) = (-1)^(n+m) P_nm(t) gives the mirror ring.
//
// Recurrence.  CHarm runs P_n = (a_n t) P_{n-1} - b_n P_{n-2}
// (src/leg/leg_func_xnum.h:29-31).  Here the same recurrence runs on
// Q_n = P_n / s_n with s_m = s_{m+1} = 1 and s_n = b_n s_{n-2}:
//   Q_n = (alpha_n Q_{n-1}) t - Q_{n-2},   alpha_n = a_n s_{n-1} / s_n,
// two FP64 instructions per term instead of three, built from CHarm's own
// (double) a_n, b_n.  Synthesis uses coefficients pre-multiplied by s_n,
// analysis multiplies its sums by s_n; s_n stays within [0.1, 1.2].
//
// X-numbers.  CHarm carries the sectorial seeds and the first terms of each
// column as X-numbers (mantissa in [2^-480, 2^480), exponent in units of
// 2^960, src/leg/leg_func_xnum.h:299-443) and continues in doubles once both
// terms of a lane are representable.  Here a lane carries scaled mantissas
// with its own binary exponent, checked every 16 degrees and renormalised
// only when a lane of the wave needs it; synthesis keeps its sums in the
// same scaled units, so an X-number term costs what a plain one does.  A wave
// runs the plain loop once no lane has an exponent left.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>

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

// Lane state in the 16-degree tiles of the X-number phase.  The true column
// values are x 2^e and y 2^e; ex is the binary exponent of |x| + |y| (that of
// the larger value, within one), so e + ex locates the true magnitude.
//  * A lane leaves the scaled form (e = 0) once e + ex > XS_PLAIN: from there
//    on every value and product stays a normal double, so the plain
//    recurrence gives the same roundings as the scaled one; a smaller
//    previous value y only loses bits below 2^-1074, against a column value
//    of at least 2^-800.  CHarm leaves X-numbers at 2^-480
//    (src/leg/leg_func_xnum.h:436-442); the terms in between are the same
//    doubles in both forms.
//  * Mantissas are rescaled (e += ex) once |x| + |y| leaves [2^-256, 2^256]:
//    16 degrees grow them by less than 2^144 (|alpha t| + 1 < 2^9 per
//    degree), far from overflow.
//  * A lane is deep when e + ex < XS_DEEP: its true values stay below 2^-1100
//    for the next 16 degrees, so every term of the tile is exactly zero.
constexpr int XS_PLAIN = -800;
constexpr int XS_DEEP = -1250;
// The analysis makes a returning ring's amplitudes again from the spectrum
// (a global round trip for its wave): once one ring of a wave returns, every
// scaled ring of the wave above 2^XS_BACK returns with it.  Its values and
// their previous ones (at least 2^-1009: a column grows by less than 2^9 per
// degree) are normal doubles, so the recurrence rounds as in the scaled form.
constexpr int XS_BACK = -1000;

__device__ __forceinline__ int xs_exp(double x, double y)
{
    return __builtin_amdgcn_frexp_exp(fabs(x) + fabs(y));
}

__device__ __forceinline__ bool xs_need(int e, int ex)
{
    return e != 0 && (e + ex > XS_PLAIN || (unsigned)(ex + 256) > 512u);
}

// power-of-two shift applied to a lane's mantissas (and scaled sums) in a
// renormalisation: back to plain doubles, or mantissas back to [0.5, 1)
__device__ __forceinline__ int xs_shift(int e, int ex)
{
    return (e == 0 || e + ex > XS_PLAIN) ? e : -ex;
}

// The 8 alpha values of a tile from the LDS window, read before any other
// LDS access of the tile: LDS requests of a wave complete in order, so a
// wait for a value read later would also wait for everything issued before.
__device__ __forceinline__ void load_alpha8(const double *al, double a[8])
{
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const double2 w = *(const double2 *)(al + 2 * i);
        a[2 * i] = w.x;
        a[2 * i + 1] = w.y;
    }
    __builtin_amdgcn_sched_barrier(0);
}

// A column's alpha values are the same for every lane of a wave: read from a
// wave-uniform address through the constant address space, they come in by
// scalar loads (s_load_dwordx16) into SGPRs, which the FP64 multiplies take
// directly, instead of through an LDS window (measured: analysis at nmax 2160
// 0.541 -> 0.534 ms, with 14 fewer VGPRs).
typedef const __attribute__((address_space(4))) double cdouble;
__device__ __forceinline__ void load_alpha8_const(cdouble *al, double a[8])
{
#pragma unroll
    for (int i = 0; i < 8; i++) a[i] = al[i];
}

__device__ __forceinline__ int uniform(int v)
{
    return __builtin_amdgcn_readfirstlane(v);
}

// Orders the LDS accesses of one wave around a refill of its window (LDS
// instructions of a wave execute in issue order).
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Per-degree column data of one order for one wave: alpha_n and (CS) the
// scaled (C_n, S_n), in 64-degree windows.  Lane l loads degree n0 + 64 + l of
// the next window while the wave works on the current one.
template <bool CS>
struct ColStage {
    double *al;          // LDS window [64]
    double2 *cs;         // LDS window [64] (CS)
    const double *gal;   // entry n of the column
    const double2 *gcs;  // entry n of the column (CS)
    const double *gsn;   // entry n of the column of s_nm (CS; nullptr: gcs is scaled)
    int nlast;           // last degree of the column
    int n0;              // first degree of the window in LDS
    int lane;
    double pal;
    double2 pcs;

    __device__ __forceinline__ void fetch(int nb)
    {
        const int n = nb + lane;
        const bool ok = n <= nlast;
        pal = ok ? gal[n] : 0.0;
        if (CS) {
            pcs = ok ? gcs[n] : make_double2(0.0, 0.0);
            if (gsn) {
                const double f = ok ? gsn[n] : 0.0;
                pcs.x *= f;
                pcs.y *= f;
            }
        }
    }
    __device__ __forceinline__ void store()
    {
        al[lane] = pal;
        if (CS) cs[lane] = pcs;
        wave_lds_sync();
    }
    __device__ __forceinline__ void start(int nb)
    {
        n0 = nb;
        fetch(nb);
        store();
        fetch(nb + 64);
    }
    __device__ __forceinline__ void next()
    {
        wave_lds_sync();
        n0 += 64;
        store();
        fetch(n0 + 64);
    }
};

}  // namespace

// ---------------------------------------------------------------------------
// Recurrence tables of degree nt, m-major (entry n of column m at
// dcol_off(nt, m) + n - m).
// ---------------------------------------------------------------------------
// (a_nm, b_nm) as CHarm computes them (src/leg/leg_func_anm_bnm.c:14-65)
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

// alpha_nm and s_nm of the scaled recurrence, one thread per column
__global__ void alpha_table_kernel(int nt, const double2 *__restrict__ ab, double *__restrict__ al,
                                   double *__restrict__ sn)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m > nt) return;
    const size_t o = dcol_off(nt, m) - m;
    const double2 *c = ab + o;
    double *A = al + o;
    double *S = sn + o;
    A[m] = 0.0;
    S[m] = 1.0;
    if (m + 1 > nt) return;
    A[m + 1] = c[m + 1].x;
    S[m + 1] = 1.0;
    double s2 = 1.0, s1 = 1.0;
    for (int n = m + 2; n <= nt; n++) {
        const double2 v = c[n];
        const double s = v.y * s2;
        A[n] = (v.x * s1) / s;
        S[n] = s;
        s2 = s1;
        s1 = s;
    }
}

// (C_nm s_nm, S_nm s_nm) for a model of degree nmax
__global__ void scale_cs_kernel(int nmax, const double2 *__restrict__ cs, const double *__restrict__ sn,
                                int nt, double2 *__restrict__ out)
{
    const int m = blockIdx.y;
    const double2 *src = cs + dcol_off(nmax, m) - m;
    double2 *dst = out + dcol_off(nmax, m) - m;
    const double *s = sn + dcol_off(nt, m) - m;
    for (int n = m + blockIdx.x * blockDim.x + threadIdx.x; n <= nmax; n += gridDim.x * blockDim.x) {
        const double2 v = src[n];
        const double f = s[n];
        dst[n] = make_double2(v.x * f, v.y * f);
    }
}

// scale_cs_kernel for many coefficient sets in one launch: set blockIdx.z
__global__ void scale_cs_group_kernel(const ScaleJob *__restrict__ jobs,
                                      const double *__restrict__ sn, int nt)
{
    const ScaleJob q = jobs[blockIdx.z];
    const int m = blockIdx.y;
    if (m > q.nmax) return;
    const double2 *src = (const double2 *)q.cs + dcol_off(q.nmax, m) - m;
    double2 *dst = (double2 *)q.out + dcol_off(q.nmax, m) - m;
    const double *s = sn + dcol_off(nt, m) - m;
    for (int n = m + blockIdx.x * blockDim.x + threadIdx.x; n <= q.nmax;
         n += gridDim.x * blockDim.x) {
        const double2 v = src[n];
        const double f = s[n];
        dst[n] = make_double2(v.x * f, v.y * f);
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
template <int K>
struct ShsSmem {
    double2 stage[2][64 * K][4];
    double al[4][64];
    double2 cs[4][64];
};

template <int K, bool UNIT>
__device__ __forceinline__ void shs_grid_item(const ShsGridArgs &A, int tile, int rbase, int quad,
                                              ShsSmem<K> &sm)
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
        ring[k] = rbase + k * 64 + lane;
        valid[k] = ring[k] < A.nring;
        t[k] = valid[k] ? A.t[ring[k]] : 0.0;
    }

    double aN[K], bN[K], aS[K], bS[K];
#pragma unroll
    for (int k = 0; k < K; k++) aN[k] = bN[k] = aS[k] = bS[k] = 0.0;

    // polar optimisation: order m is dropped at every ring of a CHarm ring
    // block once m - nmax u > threshold holds for all rings of the block
    // (src/misc/misc_polar_optimization_apply.c:16-41); pm0 holds the first
    // dropped order per ring
    bool skip[K], allskip = true;
#pragma unroll
    for (int k = 0; k < K; k++) {
        skip[k] = !valid[k] || (A.pm0 != nullptr && m >= A.pm0[ring[k]]);
        allskip = allskip && skip[k];
    }
    allskip = wave_all(allskip);
    // a wave whose column was found all zero by an earlier transform on this
    // grid (same degree and tiling) contributes nothing
    const bool dead_skip = A.dead_mode == 2 && m <= nmax && A.dead[(size_t)m * A.ntile + tile] != 0;

    if (m <= nmax && !allskip && !dead_skip) {
        const double2 *__restrict__ csm = (const double2 *)A.cs + dcol_off(nmax, m) - m;
        double x[K], y[K];
        int e[K];
        // even/odd (n + m) partial sums (UNIT) or north/south sums (!UNIT)
        double aE[K], aO[K], bE[K], bO[K];
        double rpN[K], rpS[K], qN[K], qS[K];
#pragma unroll
        for (int k = 0; k < K; k++) {
            aE[k] = aO[k] = bE[k] = bO[k] = 0.0;
            rpN[k] = rpS[k] = qN[k] = qS[k] = 1.0;
            if (!UNIT && valid[k]) {
                const int rs = A.mirror[ring[k]] ? A.nlat - 1 - ring[k] : ring[k];
                qN[k] = A.ratio[ring[k]];
                qS[k] = A.ratio[rs];
                // rld = 0: one radius for every row, a single column of powers
                const size_t col = A.rld ? (size_t)(m + 1) * A.rld : (size_t)(m + 1);
                rpN[k] = A.rp[col + (A.rld ? ring[k] : 0)];
                rpS[k] = A.rp[col + (A.rld ? rs : 0)];
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

        // degree m: P_00 = 1 or the sectorial seed P_mm (X-number).  The sums
        // are kept in the lane's scaled units (true sum = sum 2^e) and are
        // rescaled with the mantissas, so terms need no per-term ldexp.
        {
            const double2 cm = csm[m];
#pragma unroll
            for (int k = 0; k < K; k++) {
                if (m == 0) {
                    x[k] = skip[k] ? 0.0 : 1.0;
                    e[k] = 0;
                } else {
                    x[k] = skip[k] ? 0.0 : A.sx[(size_t)(m - 1) * A.sld + ring[k]];
                    e[k] = skip[k] ? 0 : 960 * A.six[(size_t)(m - 1) * A.sld + ring[k]];
                }
                y[k] = 0.0;
                SHS_ACC(k, x[k], cm, 0);
            }
        }

        bool dk[K], dead = true;
#pragma unroll
        for (int k = 0; k < K; k++) {
            const int ex = xs_exp(x[k], y[k]);
            dk[k] = !valid[k] || (e[k] != 0 && e[k] + ex < XS_DEEP);
            dead = dead && dk[k];
        }
        // every term of the column is exactly zero while the wave stays deep
        // from the seed to the last degree
        dead = wave_all(dead);

        // degrees m+1..nmax; (n + m) is odd at every tile start
        if (m < nmax) {
            ColStage<true> st;
            st.al = sm.al[wave];
            st.cs = sm.cs[wave];
            st.gal = A.al + dcol_off(A.al_nt, m) - m;
            st.gcs = csm;
            // s_m = 1: the seed term at degree m needs no scaling
            st.gsn = A.sn ? A.sn + dcol_off(A.al_nt, m) - m : nullptr;
            st.nlast = nmax;
            st.lane = lane;
            st.start(m + 1);
            int n = m + 1;
            // phase 1: 16-degree tiles while a lane of the wave still carries
            // an exponent
            for (; n + 15 <= nmax; n += 16) {
                if (n - st.n0 == 64) st.next();
                const int o = n - st.n0;
                bool anyx = false, deep = true;
#pragma unroll
                for (int k = 0; k < K; k++) {
                    anyx = anyx || (e[k] != 0);
                    deep = deep && dk[k];
                }
                if (!wave_any(anyx)) {
                    dead = false;
                    break;
                }
                const bool alldeep = wave_all(deep);
                dead = dead && alldeep;
                if (alldeep) {
#pragma unroll
                    for (int j = 0; j < 16; j++) {
                        const double al = st.al[o + j];
#pragma unroll
                        for (int k = 0; k < K; k++) {
                            const double z = fma(al * x[k], t[k], -y[k]);
                            y[k] = x[k];
                            x[k] = z;
                            SHS_SKIP(k);
                        }
                    }
                } else {
#pragma unroll
                    for (int j = 0; j < 16; j += 2) {
                        const double2 a2 = *(const double2 *)(st.al + o + j);
                        const double2 c0 = st.cs[o + j], c1 = st.cs[o + j + 1];
#pragma unroll
                        for (int k = 0; k < K; k++) {
                            const double z = fma(a2.x * x[k], t[k], -y[k]);
                            SHS_ACC(k, z, c0, 1);
                            const double w = fma(a2.y * z, t[k], -x[k]);
                            SHS_ACC(k, w, c1, 0);
                            y[k] = z;
                            x[k] = w;
                        }
                    }
                }
                int ex[K], d[K];
                bool need = false;
#pragma unroll
                for (int k = 0; k < K; k++) {
                    ex[k] = xs_exp(x[k], y[k]);
                    d[k] = e[k] + ex[k];
                    need = need || xs_need(e[k], ex[k]);
                }
                if (wave_any(need)) {
#pragma unroll
                    for (int k = 0; k < K; k++) {
                        const int sh = xs_shift(e[k], ex[k]);
                        x[k] = ldexp(x[k], sh);
                        y[k] = ldexp(y[k], sh);
                        aE[k] = ldexp(aE[k], sh);
                        aO[k] = ldexp(aO[k], sh);
                        bE[k] = ldexp(bE[k], sh);
                        bO[k] = ldexp(bO[k], sh);
                        e[k] -= sh;
                    }
                }
#pragma unroll
                for (int k = 0; k < K; k++) dk[k] = !valid[k] || (e[k] != 0 && d[k] < XS_DEEP);
            }
            // the last (fewer than 16) degrees: zero terms if still deep
            {
                bool deep = true;
#pragma unroll
                for (int k = 0; k < K; k++) deep = deep && dk[k];
                dead = dead && wave_all(deep);
            }
            // phase 2: 8-degree tiles; lanes still scaled here have fewer
            // than 16 degrees left
            for (; n + 7 <= nmax; n += 8) {
                if (n - st.n0 == 64) st.next();
                const int o = n - st.n0;
#pragma unroll
                for (int j = 0; j < 8; j += 2) {
                    const double2 a2 = *(const double2 *)(st.al + o + j);
                    const double2 c0 = st.cs[o + j], c1 = st.cs[o + j + 1];
#pragma unroll
                    for (int k = 0; k < K; k++) {
                        const double z = fma(a2.x * x[k], t[k], -y[k]);
                        SHS_ACC(k, z, c0, 1);
                        const double w = fma(a2.y * z, t[k], -x[k]);
                        SHS_ACC(k, w, c1, 0);
                        y[k] = z;
                        x[k] = w;
                    }
                }
            }
            // last, partial tile
            if (n <= nmax) {
                if (n - st.n0 == 64) st.next();
                const int o = n - st.n0;
                const int cnt = nmax - n + 1;
#pragma unroll
                for (int j = 0; j < 8; j++) {
                    if (j < cnt) {
                        const double al = st.al[o + j];
                        const double2 c = st.cs[o + j];
#pragma unroll
                        for (int k = 0; k < K; k++) {
                            const double z = fma(al * x[k], t[k], -y[k]);
                            y[k] = x[k];
                            x[k] = z;
                            SHS_ACC(k, z, c, (j & 1) ^ 1);
                        }
                    }
                }
            }
        }
#undef SHS_ACC
#undef SHS_SKIP
        // sums of lanes that never left the scaled form back to true units
        // (tiny, or exactly zero)
#pragma unroll
        for (int k = 0; k < K; k++) {
            aE[k] = ldexp(aE[k], e[k]);
            aO[k] = ldexp(aO[k], e[k]);
            bE[k] = ldexp(bE[k], e[k]);
            bO[k] = ldexp(bO[k], e[k]);
        }
        if (A.dead_mode == 1 && lane == 0) A.dead[(size_t)m * A.ntile + tile] = dead ? 1 : 0;
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
        sm.stage[0][k * 64 + lane][wave] = make_double2(sc * aN[k], sb * bN[k]);
        sm.stage[1][k * 64 + lane][wave] = make_double2(sc * aS[k], sb * bS[k]);
    }
    __syncthreads();
    double2 *out = (double2 *)A.out;
    for (int ei = threadIdx.x; ei < 2 * 64 * K * 4; ei += 256) {
        const int col = ei & 3, rowsel = ei >> 2;
        const int ns = rowsel / (64 * K), rl = rowsel % (64 * K);
        const int i = rbase + rl;
        const int mm = m0 + col;
        if (i >= A.nring || mm >= A.mcount) continue;
        int row = i;
        if (ns) {
            if (!A.mirror[i]) continue;
            row = A.nlat - 1 - i;
        }
        out[(size_t)row * A.ldo + mm] = sm.stage[ns][rl][col];
    }
}

// Rings per lane for ring tile `tile` of a launch with K: the last tile runs
// with the fewest that cover its rings (its other lanes would idle).
template <int K>
__device__ __forceinline__ int tile_k(int nring, int tile)
{
    const int rem = nring - tile * 64 * K;
    return uniform(rem >= 64 * K ? K : max(1, (rem + 63) / 64));
}

template <int K, bool UNIT>
__device__ __forceinline__ void shs_tile(const ShsGridArgs &A, int tile, int quad, ShsSmem<K> &sm)
{
    const int rbase = tile * 64 * K;
    const int kt = tile_k<K>(A.nring, tile);
    if constexpr (K >= 2) {
        if (kt == 1) {
            shs_grid_item<1, UNIT>(A, tile, rbase, quad, reinterpret_cast<ShsSmem<1> &>(sm));
            return;
        }
    }
    if constexpr (K >= 3) {
        if (kt == 2) {
            shs_grid_item<2, UNIT>(A, tile, rbase, quad, reinterpret_cast<ShsSmem<2> &>(sm));
            return;
        }
    }
    if constexpr (K >= 4) {
        if (kt == 3) {
            shs_grid_item<3, UNIT>(A, tile, rbase, quad, reinterpret_cast<ShsSmem<3> &>(sm));
            return;
        }
    }
    shs_grid_item<K, UNIT>(A, tile, rbase, quad, sm);
}

template <int K, bool UNIT>
__global__ __launch_bounds__(256) void shs_grid_kernel(ShsGridArgs A)
{
    __shared__ __align__(16) ShsSmem<K> sm;
    const int b = blockIdx.x;
    shs_tile<K, UNIT>(A, (b >> 3) % A.ntile, ((b >> 3) / A.ntile) * 8 + (b & 7), sm);
}

// Persistent form: blocks pull (request, ring tile, order quad) items from a
// list sorted by decreasing cost until the list is exhausted.
template <int K>
__global__ __launch_bounds__(256) void shs_batch_kernel(const ShsGridArgs *__restrict__ reqs,
                                                        const int2 *__restrict__ items, int nitems,
                                                        int *counter)
{
    __shared__ __align__(16) ShsSmem<K> sm;
    __shared__ int cur;
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) cur = atomicAdd(counter, 1);
        __syncthreads();
        const int it = uniform(cur);
        if (it >= nitems) break;
        const int2 w = items[it];
        const ShsGridArgs A = reqs[uniform(w.x)];
        const int wy = uniform(w.y);
        if (A.unit)
            shs_tile<K, true>(A, wy >> 16, wy & 0xFFFF, sm);
        else
            shs_tile<K, false>(A, wy >> 16, wy & 0xFFFF, sm);
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
// per (tile, n, m) goes to global memory; sha_reduce_kernel adds the tiles and
// applies s_nm.
// ---------------------------------------------------------------------------
namespace {
// Row pitch (doubles) of the per-wave transpose buffer: a multiple of 8 but
// not of 16, so that the 16-byte reads of 16 lanes cover all 64 banks.
constexpr int RED_LD = 72;

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
// behind the recurrence.  Lane l sums value l/4 over the 16 columns
// 2q + 8i, 2q + 8i + 1 (q = l mod 4, i = 0..7); the 4 lanes of a quad then
// combine through DPP.  The order of additions is fixed.
struct WaveReducer {
    double *red;      // [16][RED_LD] doubles, this wave's slice of LDS
    double *dst;      // destination of the pending tile (nullptr: none)
    int cnt;          // degrees in the pending tile
    double r[16];     // pending values read back from LDS
    int lane;
    bool acc;         // add to dst (a later ring tile of the same slot)

    // The reads are issued whether or not a tile is pending (the buffer is
    // always this wave's own LDS; unused values are dropped by finish), so
    // that no branch makes the compiler wait for them early.
    __device__ __forceinline__ void begin()
    {
        const int j = lane >> 2, q = lane & 3;
        const double *row = red + j * RED_LD + 2 * q;
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const double2 w = *(const double2 *)(row + 8 * i);
            r[2 * i] = w.x;
            r[2 * i + 1] = w.y;
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
            // the same lane owns output j for every tile of the slot, so the
            // earlier tile's value is its own store
            if ((lane & 3) == 0 && j < 2 * cnt) dst[j] = acc ? dst[j] + s : s;
            dst = nullptr;
        }
    }
    // values of degree j of the tile being computed (written as they are
    // formed; the pending tile's reads were issued before and execute first)
    __device__ __forceinline__ void put(int j, double c, double s)
    {
        red[(2 * j) * RED_LD + lane] = c;
        red[(2 * j + 1) * RED_LD + lane] = s;
    }
    // the tile just written becomes the pending one
    __device__ __forceinline__ void commit(int c, double *d)
    {
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
struct ShaSmem {
    double buf[ShaLds<K>::SIZE];
};

// The folded amplitudes of ring ic at order m, as the staging of sha_item
// forms them (same products and sums, so the same doubles): weighted spectrum
// of the northern row (cw Re F, -cw Im F) plus and minus that of its mirror
// (zero without one).
__device__ __forceinline__ void sha_amps(const ShaArgs &A, int ic, int m, double &pC, double &mC,
                                         double &pS, double &mS)
{
    const double2 *spec = (const double2 *)A.spec;
    const int rn = ic, rs = min(A.nlat - 1 - ic, A.nrow - 1);
    const double2 FN = spec[(size_t)rn * A.lds + m];
    const double2 FS = spec[(size_t)rs * A.lds + m];
    const double cn = A.cws * A.cw[rn], cs = A.cws * A.cw[rs];
    const bool mir = A.mirror[ic] != 0;
    const double nx = cn * FN.x, ny = -(cn * FN.y);
    const double sx = mir ? cs * FS.x : 0.0, sy = mir ? -(cs * FS.y) : 0.0;
    pC = nx + sx;
    mC = nx - sx;
    pS = ny + sy;
    mS = ny - sy;
}

// The ring tile starting at ring `rbase`, order quad `quad`; its sums go to
// partial slot `slot`,
// written (acc false: the slot's first tile) or added to (acc true).
template <int K>
__device__ __forceinline__ void sha_item(const ShaArgs &A, int tile, int rbase, int quad, int slot, bool acc,
                                         ShaSmem<K> &sm)
{
    const int lane = threadIdx.x & 63;
    const int wave = uniform(threadIdx.x >> 6);
    const int m0 = quad * 4;
    if (m0 > A.nmax) return;
    const int m = m0 + wave;
    const int nmax = A.nmax;
    double *lds = sm.buf;
    // K = 5 and K = 8: each tile's transpose reads are issued before the next
    // tile's arithmetic and summed after it, which hides the LDS latency at
    // the 3 (K = 5) or 2 (K = 8) waves per SIMD these kernels run at.  Other
    // K: an immediate sum, whose 32 fewer VGPRs let K <= 4 run at 4 waves per
    // SIMD and K = 6 at 3 (nmax 2160, K = 4: 0.55 ms against 0.58 ms
    // pipelined; K = 6 with its slot reduction: 0.534 ms against 0.542 ms
    // pipelined at 2 waves).
    constexpr bool PIPE = K == 5 || K >= 8;

    // the wave's per-ring data (sin of the latitude, the sectorial seed of
    // order m, the polar optimisation's first dropped order) and its
    // all-zero flag, requested first and from indices clamped into the
    // arrays, so that these loads and the staging's overlap
    int ring[K];
    double t[K], sxs[K];
    int sxe[K], pm[K];
    bool valid[K];
    int rc[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
        ring[k] = rbase + k * 64 + lane;
        valid[k] = ring[k] < A.nring;
        rc[k] = min(ring[k], A.nring - 1);
        t[k] = A.t[rc[k]];
        pm[k] = A.pm0 != nullptr ? A.pm0[rc[k]] : INT_MAX;
    }
    {
        // seed row m - 1, clamped (the tables hold at least one row); the
        // values are used for 0 < m <= nmax only
        const size_t srow = (size_t)max(0, min(m - 1, nmax - 1)) * A.sld;
#pragma unroll
        for (int k = 0; k < K; k++) {
            sxs[k] = A.sx[srow + rc[k]];
            sxe[k] = A.six[srow + rc[k]];
        }
    }
    // (read from the mirror flags when there is no table, so that the load
    // needs no branch)
    const int dead_v = (A.dead_mode == 2 ? A.dead + (size_t)min(m, nmax) * A.ntile + tile : A.mirror)[0];

    // stage the weighted spectra A = cw Re F, B = -cw Im F of the tile's rows
    // for orders m0..m0+3 (64-byte runs per row): thread (r, c) = (tid / 4,
    // tid mod 4) takes order m0 + c of the rings rbase + r + 64k and of their
    // mirrors.  The loads are unconditional, from indices clamped into the
    // arrays, so that a thread issues all of them before it waits for any;
    // what lies outside the tile, above nmax or has no mirror is dropped
    // afterwards.  The southern row nlat-1-i is clamped to the last row (on
    // Driscoll--Healy grids the north pole's would be row nrow, the missing
    // south pole).
    double2 *stage = (double2 *)lds;
    {
        const double2 *spec = (const double2 *)A.spec;
        const int c = threadIdx.x & 3, r = threadIdx.x >> 2;
        const int mm = m0 + c, mc = min(mm, nmax);
        double2 F[2 * K];
        double w[2 * K];
        int mir[K];
#pragma unroll
        for (int k = 0; k < K; k++) {
            const int ic = min(rbase + r + 64 * k, A.nring - 1);
            const int rows[2] = {ic, min(A.nlat - 1 - ic, A.nrow - 1)};
            mir[k] = A.mirror[ic];
#pragma unroll
            for (int h = 0; h < 2; h++) {
                F[h * K + k] = spec[(size_t)rows[h] * A.lds + mc];
                w[h * K + k] = A.cw[rows[h]];
            }
        }
#pragma unroll
        for (int j = 0; j < 2 * K; j++) {
            const int k = j % K;
            const bool in = rbase + r + 64 * k < A.nring && mm <= nmax && (j < K || mir[k] != 0);
            const double cw = A.cws * w[j];
            stage[(j * 64 + r) * 4 + c] =
                in ? make_double2(cw * F[j].x, -cw * F[j].y) : make_double2(0.0, 0.0);
        }
    }
    __syncthreads();

    double ampC[K], ammC[K], ampS[K], ammS[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
        if (!valid[k]) t[k] = 0.0;
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

    double *dst = (double *)((double2 *)A.part + (size_t)slot * tri_len_dev(nmax) + dcol_off(nmax, m) - m);
    // polar optimisation: order m is dropped at every ring of a CHarm ring
    // block once m - nmax u > threshold holds for all rings of the block
    // (src/misc/misc_polar_optimization_apply.c:16-41); pm0 holds the first
    // dropped order per ring
    bool skip[K], allskip = true;
#pragma unroll
    for (int k = 0; k < K; k++) {
        skip[k] = !valid[k] || m >= pm[k];
        allskip = allskip && skip[k];
    }
    // a wave whose column was found all zero by an earlier transform on this
    // grid (same degree and tiling) contributes nothing
    const bool dead_skip = A.dead_mode == 2 && dead_v != 0;
    if (wave_all(allskip) || dead_skip) {
        if (!acc)
            for (int n = m + lane; n <= nmax; n += 64) {
                dst[2 * n] = 0.0;
                dst[2 * n + 1] = 0.0;
            }
        return;
    }
    WaveReducer rd;
    rd.red = lds + wave * 16 * RED_LD;
    rd.dst = nullptr;
    rd.cnt = 0;
    rd.lane = lane;
    rd.acc = acc;

    double x[K], y[K];
    int e[K];
    double v0 = 0.0, v1 = 0.0;

    // degree m: P_00 = 1 or the sectorial seed P_mm (X-number).  A ring whose
    // column is in the scaled form (e != 0) has values below 2^XS_PLAIN =
    // 2^-800 at the last check: its terms are left out (its amplitudes are
    // zero until it leaves the scaled form, when they are made again from the
    // spectrum), so that scaled and plain rings run the same arithmetic and
    // no term needs an ldexp.  What is left out is below 2^-500 times the
    // amplitudes (checks are 32 degrees apart, which grow a column by less
    // than 2^288), far below the rounding of any coefficient.
#pragma unroll
    for (int k = 0; k < K; k++) {
        // P_00 = 1 at m = 0
        x[k] = skip[k] ? 0.0 : (m == 0 ? 1.0 : sxs[k]);
        e[k] = (skip[k] || m == 0) ? 0 : 960 * sxe[k];
        y[k] = 0.0;
        if (e[k] != 0) ampC[k] = ammC[k] = ampS[k] = ammS[k] = 0.0;
        v0 = fma(x[k], ampC[k], v0);
        v1 = fma(x[k], ampS[k], v1);
    }
    // The wave's state, made at the seed and every 32 degrees while a ring is
    // scaled: xph (some ring scaled), allx (every ring scaled: every term is
    // left out, so the tiles run the recurrence only and write zero sums),
    // dead (every ring scaled from the seed on: the whole column is zero).
    bool xph, allx, dead;
    {
        bool anyx = false, all = true;
#pragma unroll
        for (int k = 0; k < K; k++) {
            anyx = anyx || e[k] != 0;
            all = all && (!valid[k] || e[k] != 0);
        }
        xph = wave_any(anyx);
        allx = wave_all(all);
        dead = allx;
    }
    rd.put(0, v0, v1);
    rd.commit(1, dst + 2 * m);
    if (!PIPE) {
        rd.begin();
        rd.finish();
    }
    if (m == nmax) {
        if (A.dead_mode == 1 && lane == 0) A.dead[(size_t)m * A.ntile + tile] = dead ? 1 : 0;
        rd.begin();
        rd.finish();
        return;
    }

    // degrees m+1..nmax in tiles of 8; (n + m) is odd at every tile start
    cdouble *calc = (cdouble *)(A.al + dcol_off(A.al_nt, m) - m);
    int n = m + 1;
    // a[] holds the alphas of the 8-degree tile at n.  Each tile requests the
    // next tile's alphas before its arithmetic (the last whole tile requests
    // its own again, which stays in bounds): a scalar load's wait is a wait
    // for every LDS access issued before it as well, so a load made at the
    // top of the tile would also make the reducer's reads complete there.
    double a[8];
    if (n + 7 <= nmax) load_alpha8_const(calc + n, a);
    int win = 0;  // tiles since the last check
    // While a ring of the wave is scaled: one degree at a time (the sums of
    // two degrees at once, as below, need registers that the rings'
    // exponents and the checks take here), checks every 32 degrees.
    for (; xph && n + 7 <= nmax; n += 8) {
        double an[8];
        load_alpha8_const(calc + (n + 15 <= nmax ? n + 8 : n), an);
        if (PIPE) rd.begin();
        if (allx) {
            // every term of the tile is left out
#pragma unroll
            for (int j = 0; j < 8; j++) {
#pragma unroll
                for (int k = 0; k < K; k++) {
                    const double z = fma(a[j] * x[k], t[k], -y[k]);
                    y[k] = x[k];
                    x[k] = z;
                }
            }
            if (PIPE) rd.finish();
            if (lane < 16 && !acc) dst[2 * n + lane] = 0.0;
        } else {
#pragma unroll
            for (int j = 0; j < 8; j++) {
                double c = 0.0, sv = 0.0;
#pragma unroll
                for (int k = 0; k < K; k++) {
                    const double z = fma(a[j] * x[k], t[k], -y[k]);
                    y[k] = x[k];
                    x[k] = z;
                    c = fma(z, (j & 1) ? ampC[k] : ammC[k], c);
                    sv = fma(z, (j & 1) ? ampS[k] : ammS[k], sv);
                }
                rd.put(j, c, sv);
            }
            if (PIPE) rd.finish();
            rd.commit(8, dst + 2 * n);
            if (!PIPE) {
                rd.begin();
                rd.finish();
            }
        }
#pragma unroll
        for (int j = 0; j < 8; j++) a[j] = an[j];
        if (++win == 4) {
            // every 32 degrees while a ring is scaled: rescale mantissas that
            // left [2^-256, 2^256], return rings above 2^XS_PLAIN to plain
            // doubles (and their amplitudes to them)
            win = 0;
            int d[K];
            bool need = false, must = false;
#pragma unroll
            for (int k = 0; k < K; k++) {
                const int ex = xs_exp(x[k], y[k]);
                d[k] = e[k] + ex;
                need = need || xs_need(e[k], ex);
                must = must || (e[k] != 0 && d[k] > XS_PLAIN);
            }
            bool back[K], anyback = false;
#pragma unroll
            for (int k = 0; k < K; k++) back[k] = false;
            if (wave_any(need)) {
                const int plain = wave_any(must) ? XS_BACK : XS_PLAIN;
#pragma unroll
                for (int k = 0; k < K; k++) {
                    const int sh = (e[k] == 0 || d[k] > plain) ? e[k] : -(d[k] - e[k]);
                    x[k] = ldexp(x[k], sh);
                    y[k] = ldexp(y[k], sh);
                    back[k] = e[k] != 0 && sh == e[k];
                    anyback = anyback || back[k];
                    e[k] -= sh;
                }
            }
            bool anyx = false, all = true;
#pragma unroll
            for (int k = 0; k < K; k++) {
                anyx = anyx || e[k] != 0;
                all = all && (!valid[k] || e[k] != 0);
            }
            xph = wave_any(anyx);
            allx = wave_all(all);
            dead = dead && allx;
            if (wave_any(anyback)) {
                // (the order and ring base through readfirstlane, so that the
                // addresses are formed in this rare branch)
                const int mr = uniform(m), rb = uniform(rbase);
#pragma unroll
                for (int k = 0; k < K; k++)
                    if (back[k])
                        sha_amps(A, min(rb + k * 64 + lane, A.nring - 1), mr, ampC[k], ammC[k],
                                 ampS[k], ammS[k]);
            }
        }
    }
    // Every ring plain: two degrees at a time
    for (; n + 7 <= nmax; n += 8) {
        double an[8];
        load_alpha8_const(calc + (n + 15 <= nmax ? n + 8 : n), an);
        if (PIPE) rd.begin();
#pragma unroll
        for (int j = 0; j < 8; j += 2) {
            double c0 = 0.0, s0 = 0.0, c1 = 0.0, s1 = 0.0;
#pragma unroll
            for (int k = 0; k < K; k++) {
                const double z = fma(a[j] * x[k], t[k], -y[k]);
                c0 = fma(z, ammC[k], c0);
                s0 = fma(z, ammS[k], s0);
                const double w = fma(a[j + 1] * z, t[k], -x[k]);
                c1 = fma(w, ampC[k], c1);
                s1 = fma(w, ampS[k], s1);
                y[k] = z;
                x[k] = w;
            }
            rd.put(j, c0, s0);
            rd.put(j + 1, c1, s1);
        }
        if (PIPE) rd.finish();
        rd.commit(8, dst + 2 * n);
        if (!PIPE) {
            rd.begin();
            rd.finish();
        }
#pragma unroll
        for (int j = 0; j < 8; j++) a[j] = an[j];
    }
    if (A.dead_mode == 1 && lane == 0) A.dead[(size_t)m * A.ntile + tile] = dead ? 1 : 0;
    // last, partial tile
    if (n <= nmax) {
        const int cnt = nmax - n + 1;
        if (PIPE) rd.begin();
#pragma unroll
        for (int j = 0; j < 8; j++) {
            double c = 0.0, sv = 0.0;
            if (j < cnt) {
                const double al = calc[n + j];
#pragma unroll
                for (int k = 0; k < K; k++) {
                    const double z = fma(al * x[k], t[k], -y[k]);
                    y[k] = x[k];
                    x[k] = z;
                    c = fma(z, (j & 1) ? ampC[k] : ammC[k], c);
                    sv = fma(z, (j & 1) ? ampS[k] : ammS[k], sv);
                }
            }
            rd.put(j, c, sv);
        }
        if (PIPE) rd.finish();
        rd.commit(cnt, dst + 2 * n);
    }
    rd.begin();
    rd.finish();
}

// Ring tile `tile` (the last one with the fewest rings per lane that cover
// its rings, as for synthesis)
template <int K>
__device__ __forceinline__ void sha_tile(const ShaArgs &A, int tile, int quad, int slot, bool acc,
                                         ShaSmem<K> &sm)
{
    const int rbase = tile * 64 * K;
    const int kt = tile_k<K>(A.nring, tile);
    if constexpr (K >= 2) {
        if (kt == 1) {
            sha_item<1>(A, tile, rbase, quad, slot, acc, reinterpret_cast<ShaSmem<1> &>(sm));
            return;
        }
    }
    if constexpr (K >= 3) {
        if (kt == 2) {
            sha_item<2>(A, tile, rbase, quad, slot, acc, reinterpret_cast<ShaSmem<2> &>(sm));
            return;
        }
    }
    if constexpr (K >= 4) {
        if (kt == 3) {
            sha_item<3>(A, tile, rbase, quad, slot, acc, reinterpret_cast<ShaSmem<3> &>(sm));
            return;
        }
    }
    if constexpr (K >= 5) {
        if (kt == 4) {
            sha_item<4>(A, tile, rbase, quad, slot, acc, reinterpret_cast<ShaSmem<4> &>(sm));
            return;
        }
    }
    if constexpr (K == 6) {
        if (kt == 5) {
            sha_item<5>(A, tile, rbase, quad, slot, acc, reinterpret_cast<ShaSmem<5> &>(sm));
            return;
        }
    }
    if constexpr (K >= 7) {
        if (kt == 5 || kt == 6) {
            sha_item<6>(A, tile, rbase, quad, slot, acc, reinterpret_cast<ShaSmem<6> &>(sm));
            return;
        }
    }
    sha_item<K>(A, tile, rbase, quad, slot, acc, sm);
}

// Block b runs on XCD b mod 8 (round-robin dispatch).  XCD x takes the order
// quad pairs p = 8g + x; its consecutive blocks alternate between the two
// quads of a pair for each ring tile, so the two 64-byte halves of every
// 128-byte spectrum line are read through the same L2.
// Coefficients of order quad `quad` from the ngrp partial slots, added in
// slot order (as sha_reduce_kernel does for a whole triangle).  Degrees above
// nmax are not written (the fused finish requires nmax_out == nmax).
__device__ void sha_finish_quad(const ShaArgs &A, int quad, int ngrp)
{
    const int nmax = A.nmax;
    const size_t P = tri_len_dev(nmax);
    const double2 *part = (const double2 *)A.part;
    for (int w = 0; w < 4; w++) {
        const int m = quad * 4 + w;
        if (m > nmax) break;
        const size_t col = dcol_off(nmax, m) - m;
        const double *snm = A.sn + dcol_off(A.al_nt, m) - m;
        for (int n = m + (int)threadIdx.x; n <= nmax; n += blockDim.x) {
            double vc = 0.0, vs = 0.0;
            for (int sl = 0; sl < ngrp; sl++) {
                const double2 p = part[(size_t)sl * P + col + n];
                vc += p.x;
                vs += p.y;
            }
            const double f = snm[n];
            vc *= f;
            vs *= f;
            if (A.scale) {
                vc *= A.scale[n];
                vs *= A.scale[n];
            }
            if (m == 0) vs = 0.0;
            A.fc[(col + n) * A.fstride] = vc;
            A.fs[(col + n) * A.fstride] = vs;
        }
    }
}

template <int K>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(K <= 4 ? 4 : (K <= 6 ? 3 : 2))))
void sha_kernel(ShaArgs A)
{
    __shared__ __align__(16) ShaSmem<K> sm;
    const int b = blockIdx.x;
    const int ngrp = (A.ntile + A.tpg - 1) / A.tpg;
    const int s = b >> 3, h = s & 1, grp = (s >> 1) % ngrp, g = (s >> 1) / ngrp;
    const int quad = 2 * (8 * g + (b & 7)) + h;
    const int t1 = min(A.ntile, (grp + 1) * A.tpg);
    for (int tile = grp * A.tpg; tile < t1; tile++) {
        if (tile > grp * A.tpg) __syncthreads();  // the block's LDS is reused
        sha_tile<K>(A, tile, quad, grp, tile > grp * A.tpg, sm);
    }
    // One slot: this block alone summed the quad's columns; its threads read
    // back what the block wrote (workgroup scope).  With more slots a
    // last-block reduction would need a device-scope fence per block, whose L2
    // write-back and invalidation cost more than sha_reduce_kernel (0.61 ->
    // 0.96 ms at nmax 2160).
    if (!A.fc || ngrp != 1 || quad * 4 > A.nmax) return;
    __syncthreads();
    sha_finish_quad(A, quad, 1);
}

template <int K>
__global__ __launch_bounds__(256) void sha_batch_kernel(const ShaArgs *__restrict__ reqs,
                                                        const int2 *__restrict__ items, int nitems,
                                                        int *counter)
{
    __shared__ __align__(16) ShaSmem<K> sm;
    __shared__ int cur;
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) cur = atomicAdd(counter, 1);
        __syncthreads();
        const int it = uniform(cur);
        if (it >= nitems) break;
        const int2 w = items[it];
        const ShaArgs A = reqs[uniform(w.x)];
        const int wy = uniform(w.y);
        sha_tile<K>(A, wy >> 16, wy & 0xFFFF, wy >> 16, false, sm);
    }
}

// Sum the tile partials in tile order, apply s_nm and scale[n] (if given),
// and write the coefficients of a model of degree nmax_out (zeros above nmax)
// (src/sha/sha_point.c:381,1088-1108; src/shc/shc_rescale.c:24-90).
// Sum of the partial slots of one (n, m) in slot order: slots are read four
// at a time from clamped indices, all four loads issued before the adds.
__device__ __forceinline__ double2 slot_sum(const double2 *__restrict__ part, size_t P, size_t q,
                                            int nslot)
{
    double vc = 0.0, vs = 0.0;
    for (int t0 = 0; t0 < nslot; t0 += 4) {
        double2 p[4];
#pragma unroll
        for (int j = 0; j < 4; j++) p[j] = part[(size_t)min(t0 + j, nslot - 1) * P + q];
#pragma unroll
        for (int j = 0; j < 4; j++)
            if (t0 + j < nslot) {
                vc += p[j].x;
                vs += p[j].y;
            }
    }
    return make_double2(vc, vs);
}

__global__ void sha_reduce_kernel(int nmax, int ntile, const double2 *__restrict__ part,
                                  const double *__restrict__ sn, int nt, int nmax_out,
                                  const double *__restrict__ scale, double *__restrict__ c,
                                  double *__restrict__ s, size_t stride)
{
    const int m = blockIdx.y;
    const size_t P = tri_len_dev(nmax);
    const size_t out0 = dcol_off(nmax_out, m) - m;
    for (int n = m + blockIdx.x * blockDim.x + threadIdx.x; n <= nmax_out;
         n += gridDim.x * blockDim.x) {
        double vc = 0.0, vs = 0.0;
        if (m <= nmax && n <= nmax) {
            const size_t q = dcol_off(nmax, m) + (n - m);
            const double f = sn[dcol_off(nt, m) + (n - m)];
            const double g = scale ? scale[n] : 1.0;
            const double2 v = slot_sum(part, P, q, ntile);
            vc = v.x * f;
            vs = v.y * f;
            if (scale) {
                vc *= g;
                vs *= g;
            }
            if (m == 0) vs = 0.0;
        }
        c[(out0 + n) * stride] = vc;
        s[(out0 + n) * stride] = vs;
    }
}

// sha_reduce_kernel for many analyses in one launch (analysis blockIdx.z;
// output at its own degree, interleaved C, S)
__global__ void sha_reduce_group_kernel(const ShaReduceJob *__restrict__ jobs,
                                        const double *__restrict__ sn, int nt)
{
    const ShaReduceJob q = jobs[blockIdx.z];
    const int m = blockIdx.y, nmax = q.nmax;
    if (m > nmax) return;
    const size_t P = tri_len_dev(nmax);
    const size_t col = dcol_off(nmax, m) - m;
    const double2 *part = (const double2 *)q.part;
    for (int n = m + blockIdx.x * blockDim.x + threadIdx.x; n <= nmax;
         n += gridDim.x * blockDim.x) {
        const double f = sn[dcol_off(nt, m) + (n - m)];
        const double g = q.scale ? q.scale[n] : 1.0;
        const double2 v = slot_sum(part, P, col + n, q.ntile);
        double vc = v.x * f, vs = v.y * f;
        if (q.scale) {
            vc *= g;
            vs *= g;
        }
        if (m == 0) vs = 0.0;
        q.cs[2 * (col + n)] = vc;
        q.cs[2 * (col + n) + 1] = vs;
    }
}

// ---------------------------------------------------------------------------
// SHS at scattered points (src/shs/shs_point_sctr.c:294-687,
// shs_sctr_mulc.c:36-57).  K points per lane (points sorted by latitude, so
// that a wave's points leave the X-number range together), sequential over
// the orders of one chunk; the sectorial seed and (R/r)^(m+1) advance with m
// in the lane.  With r != R the factor (R/r)^(n+1) is carried by the
// recurrence itself: Q~_n = alpha_n (t q) Q~_{n-1} - q^2 Q~_{n-2}, q = R/r.
// ---------------------------------------------------------------------------
struct SctrSmem {
    double al[4][64];
    double2 cs[4][64];
};

// block blk of 256 K points, order chunk `chunk`
template <int K, bool UNIT>
__device__ __forceinline__ void sctr_item(const SctrArgs &A, int blk, int chunk, SctrSmem &sm)
{
    const int lane = threadIdx.x & 63;
    const int wave = uniform(threadIdx.x >> 6);
    const size_t base = ((size_t)blk * 4 + wave) * 64 * K;
    const int mb = A.mchunk[chunk], me = A.mchunk[chunk + 1];
    const int nmax = A.nmax;

    bool valid[K];
    double t[K], tq[K], q2[K], u[K], lon[K], q[K];
    double xs[K], rpc[K], f[K];
    int ixs[K], m0[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
        const size_t i = base + k * 64 + lane;
        valid[k] = i < A.npoint;
        m0[k] = (A.pm0 != nullptr && valid[k]) ? A.pm0[i] : nmax + 1;
        t[k] = valid[k] ? A.lat_t[i] : 0.0;
        u[k] = valid[k] ? A.lat_u[i] : 0.0;
        lon[k] = valid[k] ? A.lon[i] : 0.0;
        q[k] = (!UNIT && valid[k]) ? A.R / A.r[i] : 1.0;
        tq[k] = UNIT ? t[k] : t[k] * q[k];
        q2[k] = q[k] * q[k];
        // P_{mb,mb} (mb >= 1) as an X-number, and (R/r)^(mb+1)
        xs[k] = ROOT3 * u[k];
        ixs[k] = 0;
        for (int mm = 1; mm < mb; mm++) {
            xs[k] = (A.dm[mm] * u[k]) * xs[k];
            xnorm(xs[k], ixs[k]);
        }
        rpc[k] = q[k];
        if (!UNIT)
            for (int kk = 2; kk <= mb + 1; kk++) rpc[k] = rpc[k] * q[k];
        f[k] = 0.0;
    }

    ColStage<true> st;
    st.al = sm.al[wave];
    st.cs = sm.cs[wave];
    st.nlast = nmax;
    st.lane = lane;
    for (int m = mb; m < me; m++) {
        const double2 *__restrict__ csm = (const double2 *)A.cs + dcol_off(nmax, m) - m;
        double x[K], y[K], a[K], b[K];
        int e[K];
        const double2 cm = csm[m];
        // polar optimisation: order m is dropped at every point of a CHarm
        // point block once m - nmax u > threshold holds for all its points
        // (src/misc/misc_polar_optimization_apply.c:16-41); m0 holds the
        // first dropped order per point
        bool skip[K], allskip = true;
#pragma unroll
        for (int k = 0; k < K; k++) {
            skip[k] = !valid[k] || m >= m0[k];
            allskip = allskip && skip[k];
        }
        allskip = wave_all(allskip);
#pragma unroll
        for (int k = 0; k < K; k++) {
            // Q~_m = P_mm (R/r)^(m+1) as a mantissa and a binary exponent
            double xm_ = (m == 0) ? 1.0 : xs[k];
            int em = (m == 0) ? 0 : 960 * ixs[k];
            if (!UNIT) {
                int er;
                const double rm = frexp(rpc[k], &er);
                xm_ *= rm;
                em += er;
            }
            x[k] = skip[k] ? 0.0 : xm_;
            e[k] = skip[k] ? 0 : em;
            y[k] = 0.0;
            // sums in the lane's scaled units (true sum = sum 2^e)
            a[k] = x[k] * cm.x;
            b[k] = x[k] * cm.y;
        }
        if (m < nmax && !allskip) {
            st.gal = A.al + dcol_off(A.al_nt, m) - m;
            st.gcs = csm;
            // s_m = 1: the seed term at degree m needs no scaling
            st.gsn = A.sn ? A.sn + dcol_off(A.al_nt, m) - m : nullptr;
            st.start(m + 1);
#define SC_REC(k, al)                                                      \
    (UNIT ? fma((al) * x[k], tq[k], -y[k]) : fma((al) * x[k], tq[k], -(q2[k] * y[k])))
            int n = m + 1;
            bool dk[K];
#pragma unroll
            for (int k = 0; k < K; k++) {
                const int ex = xs_exp(x[k], y[k]);
                dk[k] = !valid[k] || (e[k] != 0 && e[k] + ex < XS_DEEP);
            }
            // phase 1: 16-degree tiles while a lane of the wave still carries
            // an exponent
            for (; n + 15 <= nmax; n += 16) {
                if (n - st.n0 == 64) st.next();
                const int o = n - st.n0;
                bool anyx = false, deep = true;
#pragma unroll
                for (int k = 0; k < K; k++) {
                    anyx = anyx || (e[k] != 0);
                    deep = deep && dk[k];
                }
                if (!wave_any(anyx)) break;
                if (wave_all(deep)) {
#pragma unroll
                    for (int j = 0; j < 16; j++) {
                        const double al = st.al[o + j];
#pragma unroll
                        for (int k = 0; k < K; k++) {
                            const double z = SC_REC(k, al);
                            y[k] = x[k];
                            x[k] = z;
                        }
                    }
                } else {
#pragma unroll
                    for (int j = 0; j < 16; j += 2) {
                        const double2 a2 = *(const double2 *)(st.al + o + j);
                        const double2 c0 = st.cs[o + j], c1 = st.cs[o + j + 1];
#pragma unroll
                        for (int k = 0; k < K; k++) {
                            const double z = SC_REC(k, a2.x);
                            a[k] = fma(z, c0.x, a[k]);
                            b[k] = fma(z, c0.y, b[k]);
                            y[k] = x[k];
                            x[k] = z;
                            const double w = SC_REC(k, a2.y);
                            a[k] = fma(w, c1.x, a[k]);
                            b[k] = fma(w, c1.y, b[k]);
                            y[k] = x[k];
                            x[k] = w;
                        }
                    }
                }
                int ex[K], d[K];
                bool need = false;
#pragma unroll
                for (int k = 0; k < K; k++) {
                    ex[k] = xs_exp(x[k], y[k]);
                    d[k] = e[k] + ex[k];
                    need = need || xs_need(e[k], ex[k]);
                }
                if (wave_any(need)) {
#pragma unroll
                    for (int k = 0; k < K; k++) {
                        const int sh = xs_shift(e[k], ex[k]);
                        x[k] = ldexp(x[k], sh);
                        y[k] = ldexp(y[k], sh);
                        a[k] = ldexp(a[k], sh);
                        b[k] = ldexp(b[k], sh);
                        e[k] -= sh;
                    }
                }
#pragma unroll
                for (int k = 0; k < K; k++) dk[k] = !valid[k] || (e[k] != 0 && d[k] < XS_DEEP);
            }
            // phase 2: 8-degree tiles; lanes still scaled here have fewer
            // than 16 degrees left
            for (; n + 7 <= nmax; n += 8) {
                if (n - st.n0 == 64) st.next();
                const int o = n - st.n0;
#pragma unroll
                for (int j = 0; j < 8; j += 2) {
                    const double2 a2 = *(const double2 *)(st.al + o + j);
                    const double2 c0 = st.cs[o + j], c1 = st.cs[o + j + 1];
#pragma unroll
                    for (int k = 0; k < K; k++) {
                        const double z = SC_REC(k, a2.x);
                        a[k] = fma(z, c0.x, a[k]);
                        b[k] = fma(z, c0.y, b[k]);
                        y[k] = x[k];
                        x[k] = z;
                        const double w = SC_REC(k, a2.y);
                        a[k] = fma(w, c1.x, a[k]);
                        b[k] = fma(w, c1.y, b[k]);
                        y[k] = x[k];
                        x[k] = w;
                    }
                }
            }
            // last, partial tile
            if (n <= nmax) {
                if (n - st.n0 == 64) st.next();
                const int o = n - st.n0;
                const int cnt = nmax - n + 1;
#pragma unroll
                for (int j = 0; j < 8; j++) {
                    if (j < cnt) {
                        const double al = st.al[o + j];
                        const double2 c = st.cs[o + j];
#pragma unroll
                        for (int k = 0; k < K; k++) {
                            const double z = SC_REC(k, al);
                            y[k] = x[k];
                            x[k] = z;
                            a[k] = fma(z, c.x, a[k]);
                            b[k] = fma(z, c.y, b[k]);
                        }
                    }
                }
            }
#undef SC_REC
        }
#pragma unroll
        for (int k = 0; k < K; k++) {
            // sums of lanes still in the scaled form back to true units
            a[k] = ldexp(a[k], e[k]);
            b[k] = ldexp(b[k], e[k]);
            if (m == 0) b[k] = 0.0;
            double sn, cn;
            sincos((double)m * lon[k], &sn, &cn);
            f[k] = f[k] + (a[k] * cn + b[k] * sn);
            // advance the seed to P_{m+1,m+1} and the radius power
            if (m >= 1) {
                xs[k] = (A.dm[m] * u[k]) * xs[k];
                xnorm(xs[k], ixs[k]);
            }
            if (!UNIT) rpc[k] = rpc[k] * q[k];
        }
    }
#pragma unroll
    for (int k = 0; k < K; k++) {
        const size_t i = base + k * 64 + lane;
        if (valid[k]) A.part[(size_t)chunk * A.npoint + i] = f[k];
    }
}

template <int K, bool UNIT>
__global__ __launch_bounds__(256) void sctr_kernel(SctrArgs A)
{
    __shared__ __align__(16) SctrSmem sm;
    if (A.unit_dev != nullptr && ((*A.unit_dev != 0) != UNIT)) return;
    sctr_item<K, UNIT>(A, blockIdx.x, blockIdx.y, sm);
}

// Persistent form: blocks pull (request, point block, order chunk) items,
// packed as (request, block << 6 | chunk), from a list sorted by decreasing
// cost until the list is exhausted.
template <int K>
__global__ __launch_bounds__(256) void sctr_batch_kernel(const SctrArgs *__restrict__ reqs,
                                                         const int2 *__restrict__ items, int nitems,
                                                         int *counter)
{
    __shared__ __align__(16) SctrSmem sm;
    __shared__ int cur;
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) cur = atomicAdd(counter, 1);
        __syncthreads();
        const int it = uniform(cur);
        if (it >= nitems) break;
        const int2 w = items[it];
        const SctrArgs A = reqs[uniform(w.x)];
        const int wy = uniform(w.y);
        if (A.unit_dev ? (*A.unit_dev != 0) : (A.unit != 0))
            sctr_item<K, true>(A, wy >> 6, wy & 63, sm);
        else
            sctr_item<K, false>(A, wy >> 6, wy & 63, sm);
    }
}

// f[perm[i]] = mu/R * sum over chunks (perm == nullptr: identity)
__global__ void sctr_reduce_kernel(size_t npoint, int nchunk, const double *__restrict__ part,
                                   const int *__restrict__ perm, double mur, double *__restrict__ f)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npoint) return;
    double s = part[i];
    for (int c = 1; c < nchunk; c++) s += part[(size_t)c * npoint + i];
    f[perm ? (size_t)perm[i] : i] = mur * s;
}

// sctr_reduce_kernel for many point sets in one launch (set blockIdx.y)
__global__ void sctr_reduce_group_kernel(const SctrReduceJob *__restrict__ jobs)
{
    const SctrReduceJob q = jobs[blockIdx.y];
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= q.npoint) return;
    double s = q.part[i];
    for (int c = 1; c < q.nchunk; c++) s += q.part[(size_t)c * q.npoint + i];
    q.f[q.perm ? (size_t)q.perm[i] : i] = q.mur * s;
}

// points in the order of perm: t, u from lat; lon and r gathered
__global__ void gather_points_kernel(size_t n, const int *__restrict__ perm,
                                     const double *__restrict__ lat, const double *__restrict__ lon,
                                     const double *__restrict__ r, double *__restrict__ t_out,
                                     double *__restrict__ u_out, double *__restrict__ lon_out,
                                     double *__restrict__ r_out)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const size_t p = (size_t)perm[i];
    sincos(lat[p], &t_out[i], &u_out[i]);
    lon_out[i] = lon[p];
    r_out[i] = r[p];
}

// Polar optimisation at scattered points: CHarm tests it jointly over blocks
// of bs consecutive input points (src/shs/shs_point_sctr.c:291-320,518-520).
// The largest cos(lat) of each block ...
__global__ void polar_block_umax_kernel(size_t n, size_t bs, const double *__restrict__ lat,
                                        double *__restrict__ umax)
{
    const size_t b = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t i0 = b * bs;
    if (i0 >= n) return;
    const size_t i1 = i0 + bs < n ? i0 + bs : n;
    double u = cos(lat[i0]);
    for (size_t i = i0 + 1; i < i1; i++) u = fmax(u, cos(lat[i]));
    umax[b] = u;
}

// ... and the first order skipped at each point (nmax + 1: none), for the
// points in the order of perm (nullptr: input order).  The test
// thr < m - nmax umax (src/misc/misc_polar_optimization_apply.c:30-37) rounds
// the product and the difference separately, as CHarm's SIMD code does.
__device__ int polar_first_skip_dev(int nmax, double thr, double umax)
{
#pragma clang fp contract(off)
    const double c = (double)nmax * umax;
    long m = (long)floor(thr + c);
    if (m < 0) m = 0;
    if (m > (long)nmax + 1) m = (long)nmax + 1;
    while (m > 0 && thr < (double)(m - 1) - c) m--;
    while (m <= (long)nmax && !(thr < (double)m - c)) m++;
    return (int)m;
}

// The same for rings whose cos(lat) is given: blocks of bs consecutive rings
__global__ void polar_block_umax_u_kernel(size_t n, size_t bs, const double *__restrict__ u,
                                          double *__restrict__ umax)
{
    const size_t b = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t i0 = b * bs;
    if (i0 >= n) return;
    const size_t i1 = i0 + bs < n ? i0 + bs : n;
    double v = u[i0];
    for (size_t i = i0 + 1; i < i1; i++) v = fmax(v, u[i]);
    umax[b] = v;
}

__global__ void polar_point_m0_kernel(size_t n, size_t bs, const int *__restrict__ perm,
                                      const double *__restrict__ umax, int nmax, double thr,
                                      int *__restrict__ m0)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const size_t p = perm ? (size_t)perm[i] : i;
    m0[i] = polar_first_skip_dev(nmax, thr, umax[p / bs]);
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

int launch_leg_tables(int nt, const double *r, const double *ri, void *ab_tmp, double *al,
                      double *sn, void *stream)
{
    hipLaunchKernelGGL(ab_table_kernel, dim3(nt + 1), dim3(256), 0, S(stream), nt, r, ri,
                       (double2 *)ab_tmp);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return (int)e;
    hipLaunchKernelGGL(alpha_table_kernel, dim3((nt + 1 + 63) / 64), dim3(64), 0, S(stream), nt,
                       (const double2 *)ab_tmp, al, sn);
    return (int)hipGetLastError();
}

int launch_scale_cs(int nmax, const void *cs, const double *sn, int nt, void *out, void *stream)
{
    const unsigned gx = (unsigned)((nmax + 1 + 255) / 256);
    hipLaunchKernelGGL(scale_cs_kernel, dim3(gx, (unsigned)nmax + 1), dim3(256), 0, S(stream), nmax,
                       (const double2 *)cs, sn, nt, (double2 *)out);
    return (int)hipGetLastError();
}

int launch_scale_cs_group(const ScaleJob *jobs, int njob, int nmax_max, const double *sn, int nt,
                          void *stream)
{
    if (njob <= 0) return 0;
    const dim3 grid((unsigned)((nmax_max + 1 + 255) / 256), (unsigned)nmax_max + 1, (unsigned)njob);
    hipLaunchKernelGGL(scale_cs_group_kernel, grid, dim3(256), 0, S(stream), jobs, sn, nt);
    return (int)hipGetLastError();
}

int launch_sha_reduce_group(const ShaReduceJob *jobs, int njob, int nmax_max, const double *sn,
                            int nt, void *stream)
{
    if (njob <= 0) return 0;
    const dim3 grid((unsigned)((nmax_max + 1 + 255) / 256), (unsigned)nmax_max + 1, (unsigned)njob);
    hipLaunchKernelGGL(sha_reduce_group_kernel, grid, dim3(256), 0, S(stream), jobs, sn, nt);
    return (int)hipGetLastError();
}

int launch_sctr_reduce_group(const SctrReduceJob *jobs, int njob, size_t npoint_max, void *stream)
{
    if (njob <= 0) return 0;
    const dim3 grid((unsigned)((npoint_max + 255) / 256), (unsigned)njob);
    hipLaunchKernelGGL(sctr_reduce_group_kernel, grid, dim3(256), 0, S(stream), jobs);
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
    case 4: shs_launch_k<4>(a, nblk, S(stream)); break;
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
    const int nqg = (nquad + 15) / 16;  // groups of 8 quad pairs
    const int ngrp = (a.ntile + a.tpg - 1) / a.tpg;
    const unsigned nblk = 16u * (unsigned)nqg * (unsigned)ngrp;
    switch (a.k) {
    case 2: hipLaunchKernelGGL((sha_kernel<2>), dim3(nblk), dim3(256), 0, S(stream), a); break;
    case 3: hipLaunchKernelGGL((sha_kernel<3>), dim3(nblk), dim3(256), 0, S(stream), a); break;
    case 5: hipLaunchKernelGGL((sha_kernel<5>), dim3(nblk), dim3(256), 0, S(stream), a); break;
    case 6: hipLaunchKernelGGL((sha_kernel<6>), dim3(nblk), dim3(256), 0, S(stream), a); break;
    case 8: hipLaunchKernelGGL((sha_kernel<8>), dim3(nblk), dim3(256), 0, S(stream), a); break;
    default: hipLaunchKernelGGL((sha_kernel<4>), dim3(nblk), dim3(256), 0, S(stream), a); break;
    }
    return (int)hipGetLastError();
}

int launch_sha_reduce(int nmax, int ntile, const void *part, const double *sn, int nt, int nmax_out,
                      const double *scale, double *c, double *s, size_t stride, void *stream)
{
    const unsigned gx = (unsigned)((nmax_out + 1 + 255) / 256);
    hipLaunchKernelGGL(sha_reduce_kernel, dim3(gx, nmax_out + 1), dim3(256), 0, S(stream), nmax,
                       ntile, (const double2 *)part, sn, nt, nmax_out, scale, c, s, stride);
    return (int)hipGetLastError();
}

int sctr_points_per_block(int k)
{
    return 256 * k;
}

template <int K>
static void sctr_launch_k(const SctrArgs &a, hipStream_t s)
{
    dim3 grid((unsigned)((a.npoint + 256 * K - 1) / (256 * K)), (unsigned)a.nchunk);
    if (a.unit_dev != nullptr || a.unit)
        hipLaunchKernelGGL((sctr_kernel<K, true>), grid, dim3(256), 0, s, a);
    if (a.unit_dev != nullptr || !a.unit)
        hipLaunchKernelGGL((sctr_kernel<K, false>), grid, dim3(256), 0, s, a);
}

int launch_sctr(const SctrArgs &a, void *stream)
{
    switch (a.k) {
    case 1: sctr_launch_k<1>(a, S(stream)); break;
    case 4: sctr_launch_k<4>(a, S(stream)); break;
    default: sctr_launch_k<2>(a, S(stream)); break;
    }
    return (int)hipGetLastError();
}

int launch_sctr_reduce(size_t npoint, int nchunk, const double *part, const int *perm, double mur,
                       double *f, void *stream)
{
    hipLaunchKernelGGL(sctr_reduce_kernel, dim3((unsigned)((npoint + 255) / 256)), dim3(256), 0,
                       S(stream), npoint, nchunk, part, perm, mur, f);
    return (int)hipGetLastError();
}

// Points ordered by |sin lat|: a stable LSD radix sort on a 16-bit key,
// floor(|sin lat| 2^16) (16 bits group the points of a wave closely enough in
// latitude; the order within a key is the input order), in two 8-bit passes of
// three kernels each: per-tile digit counts, one exclusive scan over (digit,
// tile), and a stable scatter.  Written for this library instead of a library
// sort: it clears nothing with memsets, keeps its whole state in a caller
// workspace and issues no host copies or synchronisations, so it can be
// captured into a HIP graph.
namespace {
constexpr int SORT_THREADS = 256;

struct SortPlan {
    size_t n, tile, nblk, bytes, o_idx, o_hist;
    explicit SortPlan(size_t n_) : n(n_)
    {
        auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
        // at most 8192 tiles, multiples of the block size, at least 1024
        const size_t need = (n + 8191) / 8192;
        tile = std::max<size_t>(1024, (need + SORT_THREADS - 1) / SORT_THREADS * SORT_THREADS);
        nblk = (n + tile - 1) / tile;
        o_idx = al(n * sizeof(unsigned));                 // keys of pass 1 at 0
        o_hist = o_idx + al(n * sizeof(int));
        bytes = o_hist + al(256 * nblk * sizeof(int));
    }
};

__device__ __forceinline__ unsigned lat_key(double lat)
{
    const double v = fabs(sin(lat)) * 65536.0;
    return v >= 65535.0 ? 65535u : (unsigned)v;
}

// digit counts of one tile: hist[d * nblk + b]
__global__ __launch_bounds__(SORT_THREADS) void sort_hist_kernel(size_t n, size_t tile, size_t nblk,
                                                                 const double *__restrict__ lat,
                                                                 const unsigned *__restrict__ key,
                                                                 int shift, int *__restrict__ hist)
{
    __shared__ int h[256];
    const size_t b = blockIdx.x;
    h[threadIdx.x] = 0;
    __syncthreads();
    const size_t i0 = b * tile, i1 = i0 + tile < n ? i0 + tile : n;
    for (size_t i = i0 + threadIdx.x; i < i1; i += SORT_THREADS) {
        const unsigned k = lat ? lat_key(lat[i]) : key[i];
        atomicAdd(&h[(k >> shift) & 255], 1);
    }
    __syncthreads();
    hist[(size_t)threadIdx.x * nblk + b] = h[threadIdx.x];
}

// exclusive scan of cnt entries in place, one block of 1024 threads
__global__ __launch_bounds__(1024) void sort_scan_kernel(size_t cnt, int *__restrict__ v)
{
    __shared__ int part[1024];
    const size_t per = (cnt + 1023) / 1024;
    const size_t a = threadIdx.x * per, e = a + per < cnt ? a + per : cnt;
    int sum = 0;
    for (size_t i = a; i < e; i++) sum += v[i];
    part[threadIdx.x] = sum;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const int x = threadIdx.x >= (unsigned)off ? part[threadIdx.x - off] : 0;
        __syncthreads();
        part[threadIdx.x] += x;
        __syncthreads();
    }
    int run = part[threadIdx.x] - sum;
    for (size_t i = a; i < e; i++) {
        const int c = v[i];
        v[i] = run;
        run += c;
    }
}

// stable scatter of one tile: the tile is taken in chunks of 256 elements in
// order; within a chunk an element's place among equal digits comes from
// lane ballots (waves in order), so equal digits keep their input order
__global__ __launch_bounds__(SORT_THREADS) void sort_scatter_kernel(
    size_t n, size_t tile, size_t nblk, const double *__restrict__ lat,
    const unsigned *__restrict__ key, const int *__restrict__ idx, int shift,
    const int *__restrict__ offs, unsigned *__restrict__ key_out, int *__restrict__ idx_out)
{
    __shared__ int run[256];
    __shared__ int cnt[SORT_THREADS / 64][256];
    const size_t b = blockIdx.x;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    run[t] = offs[(size_t)t * nblk + b];
    const size_t i0 = b * tile, i1 = i0 + tile < n ? i0 + tile : n;
    const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (size_t c0 = i0; c0 < i1; c0 += SORT_THREADS) {
        for (int w = 0; w < SORT_THREADS / 64; w++) cnt[w][t] = 0;
        __syncthreads();
        const size_t i = c0 + t;
        const bool ok = i < i1;
        unsigned k = 0;
        int id = 0;
        if (ok) {
            k = lat ? lat_key(lat[i]) : key[i];
            id = lat ? (int)i : idx[i];
        }
        const int d = (k >> shift) & 255;
        unsigned long long peers = __ballot(ok);
#pragma unroll
        for (int bit = 0; bit < 8; bit++) {
            const unsigned long long on = __ballot(((d >> bit) & 1) != 0);
            peers &= ((d >> bit) & 1) ? on : ~on;
        }
        const int below = __popcll(peers & lt);
        if (ok && below == 0) cnt[wave][d] = __popcll(peers);
        __syncthreads();
        if (ok) {
            int pos = run[d] + below;
            for (int w = 0; w < wave; w++) pos += cnt[w][d];
            if (key_out) key_out[pos] = k;
            idx_out[pos] = id;
        }
        __syncthreads();
        int add = 0;
        for (int w = 0; w < SORT_THREADS / 64; w++) add += cnt[w][t];
        run[t] += add;
        __syncthreads();
    }
}
}  // namespace

size_t sort_points_ws_bytes(size_t n)
{
    return SortPlan(std::max<size_t>(n, 1)).bytes;
}

int launch_sort_points_ws(size_t n, const double *lat, const double *lon, const double *r,
                          int *perm, void *ws, size_t ws_bytes, double *t_out, double *u_out,
                          double *lon_out, double *r_out, const int **perm_out, void *stream)
{
    hipStream_t s = S(stream);
    *perm_out = perm;
    if (n == 0) return 0;
    if (n >= (size_t)INT32_MAX) return (int)hipErrorInvalidValue;
    const SortPlan P(n);
    if (P.bytes > ws_bytes) return (int)hipErrorInvalidValue;
    char *w = (char *)ws;
    unsigned *k1 = (unsigned *)w;
    int *i1 = (int *)(w + P.o_idx), *hist = (int *)(w + P.o_hist);
    const dim3 grid((unsigned)P.nblk);
    const size_t cnt = 256 * P.nblk;
    // pass 1: low byte, keys computed from the latitudes
    hipLaunchKernelGGL(sort_hist_kernel, grid, dim3(SORT_THREADS), 0, s, n, P.tile, P.nblk, lat,
                       (const unsigned *)nullptr, 0, hist);
    hipLaunchKernelGGL(sort_scan_kernel, dim3(1), dim3(1024), 0, s, cnt, hist);
    hipLaunchKernelGGL(sort_scatter_kernel, grid, dim3(SORT_THREADS), 0, s, n, P.tile, P.nblk, lat,
                       (const unsigned *)nullptr, (const int *)nullptr, 0, (const int *)hist, k1,
                       i1);
    // pass 2: high byte; only the permutation is kept
    hipLaunchKernelGGL(sort_hist_kernel, grid, dim3(SORT_THREADS), 0, s, n, P.tile, P.nblk,
                       (const double *)nullptr, (const unsigned *)k1, 8, hist);
    hipLaunchKernelGGL(sort_scan_kernel, dim3(1), dim3(1024), 0, s, cnt, hist);
    hipLaunchKernelGGL(sort_scatter_kernel, grid, dim3(SORT_THREADS), 0, s, n, P.tile, P.nblk,
                       (const double *)nullptr, (const unsigned *)k1, (const int *)i1, 8,
                       (const int *)hist, (unsigned *)nullptr, perm);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) {
        hipLaunchKernelGGL(gather_points_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                           n, (const int *)perm, lat, lon, r, t_out, u_out, lon_out, r_out);
        e = hipGetLastError();
    }
    return (int)e;
}

// The same with the workspace from the stream-ordered pool (a graph
// allocation node under capture)
int launch_sort_points(size_t n, const double *lat, const double *lon, const double *r, int *perm,
                       double *t_out, double *u_out, double *lon_out, double *r_out, void *stream)
{
    hipStream_t s = S(stream);
    const size_t bytes = sort_points_ws_bytes(n);
    void *ws = nullptr;
    hipError_t e = hipMallocAsync(&ws, bytes, s);
    if (e != hipSuccess) return (int)e;
    const int *cur = perm;
    e = (hipError_t)launch_sort_points_ws(n, lat, lon, r, perm, ws, bytes, t_out, u_out, lon_out,
                                          r_out, &cur, stream);
    (void)hipFreeAsync(ws, s);
    return (int)e;
}

int launch_polar_points_ws(size_t n, const double *lat, const int *perm, size_t bs, int nmax,
                           double thr, double *umax_ws, int *m0, void *stream)
{
    hipStream_t s = S(stream);
    const size_t nb = (n + bs - 1) / bs;
    hipLaunchKernelGGL(polar_block_umax_kernel, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, s,
                       n, bs, lat, umax_ws);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) {
        hipLaunchKernelGGL(polar_point_m0_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                           s, n, bs, perm, umax_ws, nmax, thr, m0);
        e = hipGetLastError();
    }
    return (int)e;
}

int launch_polar_rings_ws(size_t n, const double *u, size_t bs, int nmax, double thr,
                          double *umax_ws, int *m0, void *stream)
{
    hipStream_t s = S(stream);
    const size_t nb = (n + bs - 1) / bs;
    hipLaunchKernelGGL(polar_block_umax_u_kernel, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0,
                       s, n, bs, u, umax_ws);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) {
        hipLaunchKernelGGL(polar_point_m0_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                           s, n, bs, (const int *)nullptr, umax_ws, nmax, thr, m0);
        e = hipGetLastError();
    }
    return (int)e;
}

// out[k] = q^k, k = 0..n, by repeated products in order (one thread: the
// same sequence of roundings as the host loops of shc_rescale and shs_rpows)
__global__ void powers_kernel(int n, double q, double *__restrict__ out)
{
    if (threadIdx.x != 0) return;
    double v = 1.0;
    out[0] = v;
    for (int k = 1; k <= n; k++) {
        v = v * q;
        out[k] = v;
    }
}

__global__ void fill_kernel(size_t n, double v, double *__restrict__ out)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = v;
}

int launch_powers(int n, double q, double *out, void *stream)
{
    hipLaunchKernelGGL(powers_kernel, dim3(1), dim3(64), 0, S(stream), n, q, out);
    return (int)hipGetLastError();
}

int launch_fill(size_t n, double v, double *out, void *stream)
{
    if (!n) return 0;
    hipLaunchKernelGGL(fill_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, S(stream), n, v,
                       out);
    return (int)hipGetLastError();
}

int launch_polar_points(size_t n, const double *lat, const int *perm, size_t bs, int nmax,
                        double thr, int *m0, void *stream)
{
    hipStream_t s = S(stream);
    const size_t nb = (n + bs - 1) / bs;
    double *umax = nullptr;
    hipError_t e = hipMallocAsync((void **)&umax, std::max<size_t>(nb, 1) * sizeof(double), s);
    if (e != hipSuccess) return (int)e;
    e = (hipError_t)launch_polar_points_ws(n, lat, perm, bs, nmax, thr, umax, m0, stream);
    (void)hipFreeAsync(umax, s);
    return (int)e;
}

int launch_sincos(size_t n, const double *lat, double *t, double *u, void *stream)
{
    hipLaunchKernelGGL(sincos_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, S(stream), n,
                       lat, t, u);
    return (int)hipGetLastError();
}

__global__ void set_ints_kernel(int *p, IntVals v)
{
    const int i = (int)threadIdx.x;
    if (i < v.n) p[i] = v.v[i];
}

int launch_set_ints(int *p, const IntVals &v, void *stream)
{
    if (v.n < 0 || v.n > 72) return (int)hipErrorInvalidValue;
    hipLaunchKernelGGL(set_ints_kernel, dim3(1), dim3(128), 0, S(stream), p, v);
    return (int)hipGetLastError();
}

int launch_all_near_dev(size_t n, const double *x, double ref, double eps, int *flag, void *stream)
{
    IntVals one{};
    one.n = 1;
    one.v[0] = 1;
    hipError_t e = (hipError_t)launch_set_ints(flag, one, stream);
    if (e == hipSuccess && n) {
        hipLaunchKernelGGL(all_near_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                           S(stream), n, x, ref, eps, flag);
        e = hipGetLastError();
    }
    return (int)e;
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
    (void)hipFreeAsync(flag, S(stream));
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

// ---------------------------------------------------------------------------
// Derivatives: transformed coefficient tables and their per-point combination
// ---------------------------------------------------------------------------
__global__ void grad_table_kernel(int nmax, const double2 *__restrict__ cs, GradTable g,
                                  double2 *__restrict__ out)
{
    const int m = blockIdx.y;
    const double2 *src = cs + dcol_off(nmax, m) - m;
    double2 *dst = out + dcol_off(nmax, m) - m;
    const double dm = (double)m;
    for (int nd = m + blockIdx.x * blockDim.x + threadIdx.x; nd <= nmax;
         nd += gridDim.x * blockDim.x) {
        const int n = nd + g.shift;
        double2 v = make_double2(0.0, 0.0);
        if (n <= nmax) {
            const double2 x = src[n];
            const double dn = (double)n;
            double f = g.c0 + dn * (g.c1 + dn * g.c2);
            // e_nm = sqrt(2n+1) sqrt(n-m) sqrt(n+m) / sqrt(2n-1), n > m here
            if (g.shift)
                f *= sqrt(2.0 * dn + 1.0) * sqrt(dn - dm) * sqrt(dn + dm) *
                     (1.0 / sqrt(2.0 * dn - 1.0));
            if (g.lon == 1)
                v = make_double2(f * (dm * x.y), f * (-dm * x.x));
            else if (g.lon == 2)
                v = make_double2(f * (-(dm * dm) * x.x), f * (-(dm * dm) * x.y));
            else
                v = make_double2(f * x.x, f * x.y);
        }
        dst[nd] = v;
    }
}

__global__ void grad_combine_kernel(size_t nrow, size_t ncol, const double *__restrict__ t,
                                    const double *__restrict__ u, const double *__restrict__ r,
                                    int unit, double R, const double *__restrict__ tabs, size_t tld,
                                    GradCombine c)
{
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nrow * ncol) return;
    const size_t row = (ncol == 1) ? p : p / ncol;
    const double tt = t[row], ui = 1.0 / u[row];
    const double rr = unit ? R : r[row];
    const double ri = 1.0 / rr, q = unit ? 1.0 : R / rr;
    const double tp[3] = {1.0, tt, tt * tt}, up[3] = {1.0, ui, ui * ui};
    const double rp[3] = {1.0, ri, ri * ri}, qp[2] = {1.0, q};
    for (int o = 0; o < c.nout; o++) {
        double acc = 0.0;
        for (int k = c.beg[o]; k < c.beg[o + 1]; k++) {
            const GradTerm &g = c.term[k];
            const double w = g.coef * ((tp[g.pt] * up[g.pu]) * (rp[g.pr] * qp[g.pq]));
            acc += w * tabs[(size_t)g.tab * tld + p];
        }
        c.out[o][p] = acc;
    }
}

__global__ void grad_spec_combine_kernel(size_t nrow, size_t ld, const double *__restrict__ t,
                                         const double *__restrict__ u,
                                         const double *__restrict__ r, int unit, double R,
                                         const double2 *__restrict__ basis, SpecCombine c)
{
    const size_t m = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t row = blockIdx.y;
    if (m >= ld) return;
    const double tt = t[row], ui = 1.0 / u[row];
    const double rr = unit ? R : r[row];
    const double ri = 1.0 / rr, q = unit ? 1.0 : R / rr;
    const double tp[3] = {1.0, tt, tt * tt}, up[3] = {1.0, ui, ui * ui};
    const double rp[3] = {1.0, ri, ri * ri}, qp[2] = {1.0, q};
    const double dm = (double)m;
    for (int o = 0; o < c.nout; o++) {
        double ax = 0.0, ay = 0.0;
        for (int k = c.beg[o]; k < c.beg[o + 1]; k++) {
            const SpecTerm &g = c.term[k];
            const double w = g.coef * ((tp[g.pt] * up[g.pu]) * (rp[g.pr] * qp[g.pq]));
            const double2 b = basis[((size_t)g.basis * nrow + row) * ld + m];
            // (i m)^lon
            double bx = b.x, by = b.y;
            if (g.lon == 1) {
                bx = -dm * b.y;
                by = dm * b.x;
            } else if (g.lon == 2) {
                bx = -(dm * dm) * b.x;
                by = -(dm * dm) * b.y;
            }
            ax += w * bx;
            ay += w * by;
        }
        ((double2 *)c.out[o])[row * ld + m] = make_double2(ax, ay);
    }
}

int launch_grad_spec_combine(size_t nrow, size_t ld, const double *t, const double *u,
                             const double *r, int unit, double R, const void *basis,
                             const SpecCombine &c, void *stream)
{
    if (nrow == 0 || ld == 0) return 0;
    hipLaunchKernelGGL(grad_spec_combine_kernel, dim3((unsigned)((ld + 255) / 256), (unsigned)nrow),
                       dim3(256), 0, S(stream), nrow, ld, t, u, r, unit, R,
                       (const double2 *)basis, c);
    return (int)hipGetLastError();
}

int launch_grad_table(int nmax, const void *cs, const GradTable &g, void *out, void *stream)
{
    const unsigned gx = (unsigned)((nmax + 1 + 255) / 256);
    hipLaunchKernelGGL(grad_table_kernel, dim3(gx, (unsigned)nmax + 1), dim3(256), 0, S(stream),
                       nmax, (const double2 *)cs, g, (double2 *)out);
    return (int)hipGetLastError();
}

int launch_grad_combine(size_t nrow, size_t ncol, const double *t, const double *u,
                        const double *r, int unit, double R, const double *tabs, size_t tld,
                        const GradCombine &c, void *stream)
{
    const size_t n = nrow * ncol;
    if (n == 0) return 0;
    hipLaunchKernelGGL(grad_combine_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       S(stream), nrow, ncol, t, u, r, unit, R, tabs, tld, c);
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

constexpr int BATCH_SHS_K = 3;
constexpr int BATCH_SHA_K = 4;
constexpr int BATCH_SCTR_K = 2;

int batch_ring_k(int kind)
{
    return kind == 0 ? BATCH_SHS_K : (kind == 1 ? BATCH_SHA_K : BATCH_SCTR_K);
}

// Resident blocks per CU of a 256-thread kernel on gfx950: 8 wave slots per
// SIMD, a 512-entry VGPR file per SIMD lane (allocated in granules of 8) and
// 160 KiB of LDS per CU.  (hipOccupancyMaxActiveBlocksPerMultiprocessor
// budgets 256 VGPRs per lane and under-fills the persistent grids.)
template <class F>
int resident_blocks(F kernel)
{
    hipFuncAttributes fa{};
    if (hipFuncGetAttributes(&fa, (const void *)kernel) != hipSuccess) return 1;
    const int vg = std::max(8, (fa.numRegs + 7) / 8 * 8);
    int per = std::min(8, 512 / vg);
    const size_t lds = fa.sharedSizeBytes;
    if (lds > 0) per = std::min<int>(per, (int)((160 * 1024) / lds));
    return std::max(per, 1);
}

int batch_grid_size(int kind)
{
    int dev = 0, ncu = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    const int per = (kind == 0)   ? resident_blocks(shs_batch_kernel<BATCH_SHS_K>)
                    : (kind == 1) ? resident_blocks(sha_batch_kernel<BATCH_SHA_K>)
                                  : resident_blocks(sctr_batch_kernel<BATCH_SCTR_K>);
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

int launch_sctr_batch(const void *reqs, const void *items, int nitems, int *counter, int grid,
                      void *stream)
{
    if (nitems <= 0) return 0;
    hipLaunchKernelGGL((sctr_batch_kernel<BATCH_SCTR_K>), dim3(grid), dim3(256), 0, S(stream),
                       (const SctrArgs *)reqs, (const int2 *)items, nitems, counter);
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
