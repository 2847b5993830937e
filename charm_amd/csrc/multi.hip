// Many coefficient sets on one Gauss-Legendre grid: the sums over degrees
// (synthesis) are dense contractions of a latitude panel of Legendre values
// with a block of coefficient sets, and run on the FP64 matrix cores
// (v_mfma_f64_16x16x4f64).
//
// Synthesis, per order m: a[ring][set] = sum_n Q_n(ring) (C_n s_n)[set], with
// even and odd (n + m) kept apart for the equatorial fold
// (src/shs/shs_point_kernel.c:1014-1623 evaluates the same sums one set at a
// time).  One wave handles (m, 16 sets, 64 ring pairs): each lane runs the
// scaled recurrence for its ring pair, eight degrees at a time; the eight
// values go through LDS into MFMA A-operands (16 rings x 4 degrees), the
// coefficient block through LDS into B-operands (4 degrees x 16 sets), and 16
// accumulator tiles (4 ring tiles x {C, S} x {even, odd}) collect the sums.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "internal.h"

namespace charm_amd {

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ size_t mcol_off(long nmax, long m)
{
    return (size_t)m * (size_t)(nmax + 1) - ((size_t)m * (size_t)(m - 1)) / 2;
}

__device__ __forceinline__ int uni(int v)
{
    return __builtin_amdgcn_readfirstlane(v);
}

__device__ __forceinline__ bool any64(bool p)
{
    return __ballot(p) != 0;
}

__device__ __forceinline__ bool all64(bool p)
{
    return __ballot(!p) == 0;
}

__device__ __forceinline__ void wave_sync_lds()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// exponent handling of the scaled mantissas (as in the single-set kernels)
__device__ __forceinline__ void renorm(double &x, double &y, int &e)
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

}  // namespace

// Block b (one wave) -> (m, set tile, ring tile).  Blocks run on XCD b mod 8;
// each XCD owns a fixed share of the (set tile, ring tile) pairs and runs the
// 8 orders of an order group back to back for each pair, so the 16-byte
// outputs of consecutive orders meet in its L2 as whole lines.
__global__ __launch_bounds__(64) void shs_multi_kernel(MultiShsArgs A)
{
    __shared__ __align__(16) double pb[8][64];     // Q of 8 degrees x 64 ring pairs
    __shared__ __align__(16) double cbc[16][16];   // C s of 16 degrees x 16 sets
    __shared__ __align__(16) double cbs[16][16];   // S s
    __shared__ __align__(16) double alw[64];       // alpha window

    const int lane = threadIdx.x;
    const int b = blockIdx.x;
    const int x = b & 7, s = b >> 3;
    const int j8 = s & 7;
    const int ppx = (A.npair + 7) / 8;
    const int pp = (s >> 3) % ppx, g = (s >> 3) / ppx;
    const int pair = pp * 8 + x;
    if (pair >= A.npair) return;
    const int st = pair / A.nrt, rt = pair % A.nrt;
    const int m = uni(8 * g + j8);
    if (m >= A.mcount) return;
    const int nmax = A.nmax;
    const int ring = rt * 64 + lane;
    const bool valid = ring < A.nring;
    const int set0 = st * 16;

    d4 acc[4][2][2];
#pragma unroll
    for (int q = 0; q < 4; q++)
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
            for (int p = 0; p < 2; p++) acc[q][c][p] = (d4){0.0, 0.0, 0.0, 0.0};

    // polar optimisation: order m is dropped at every ring of a CHarm ring
    // block (src/misc/misc_polar_optimization_apply.c:16-41); pm0 holds the
    // first dropped order per ring
    const bool skip = !valid || (A.pm0 != nullptr && m >= A.pm0[ring]);
    if (m <= nmax && !all64(skip)) {
        const double t = valid ? A.t[ring] : 0.0;
        double xq, yq = 0.0;
        int e = 0;
        if (m == 0) {
            xq = skip ? 0.0 : 1.0;
        } else {
            xq = skip ? 0.0 : A.sx[(size_t)(m - 1) * A.sld + ring];
            e = skip ? 0 : 960 * A.six[(size_t)(m - 1) * A.sld + ring];
        }
        const double *alm = A.al + mcol_off(A.al_nt, m) - m;
        // coefficient block loader: lane -> set set0 + lane/4, degrees 4(lane%4)..+3
        const int cset = set0 + (lane >> 2), cq = (lane & 3) * 4;
        const bool cvalid = cset < A.nset;
        const double2 *ccol = (const double2 *)A.cs + (size_t)(cvalid ? cset : 0) * A.cs_stride + mcol_off(nmax, m) - m;
        double2 pre[4];
        auto cfetch = [&](int n0) {
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int n = n0 + cq + i;
                pre[i] = (cvalid && n <= nmax) ? ccol[n] : make_double2(0.0, 0.0);
            }
        };
        auto cstore = [&]() {
#pragma unroll
            for (int i = 0; i < 4; i++) {
                cbc[cq + i][lane >> 2] = pre[i].x;
                cbs[cq + i][lane >> 2] = pre[i].y;
            }
        };
        // alpha window of 64 degrees, refilled from a register prefetch
        double apre = (m + lane <= A.al_nt) ? alm[m + lane] : 0.0;
        int a0 = m;
        alw[lane] = apre;
        apre = (m + 64 + lane <= A.al_nt) ? alm[m + 64 + lane] : 0.0;
        cfetch(m);
        cstore();
        cfetch(m + 16);
        wave_sync_lds();

        bool xmode = true;
        // degrees n0..n0+7, n0 = m + 8i; (n + m) even for j even
        for (int n0 = m; n0 <= nmax; n0 += 8) {
            if (n0 - a0 == 64) {
                wave_sync_lds();
                a0 += 64;
                alw[lane] = apre;
                apre = (a0 + 64 + lane <= A.al_nt) ? alm[a0 + 64 + lane] : 0.0;
            }
            if (((n0 - m) & 15) == 0 && n0 > m) {
                wave_sync_lds();
                cstore();
                cfetch(n0 + 16);
            }
            const int ao = n0 - a0;
            bool deep = false;
            if (xmode) {
                int ex, ey;
                (void)frexp(xq, &ex);
                (void)frexp(yq, &ey);
                deep = all64(!valid || (e != 0 && e + max(ex, ey) < -1180));
            }
            // Q values of the 8 degrees
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const int n = n0 + j;
                double z;
                if (n == m) {
                    z = xq;
                } else {
                    z = fma(alw[ao + j] * xq, t, -yq);
                    yq = xq;
                    xq = z;
                }
                pb[j][lane] = (n <= nmax && !deep) ? (xmode ? ldexp(z, e) : z) : 0.0;
            }
            if (xmode) {
                renorm(xq, yq, e);
                xmode = any64(e != 0);
            }
            wave_sync_lds();
            if (!deep) {
                const int co = (n0 - m) & 15;  // 0 or 8 within the coefficient block
                const int kk = lane >> 4, col = lane & 15;
#pragma unroll
                for (int p = 0; p < 2; p++) {
                    const double bc = cbc[co + 2 * kk + p][col];
                    const double bs = cbs[co + 2 * kk + p][col];
#pragma unroll
                    for (int q = 0; q < 4; q++) {
                        const double a = pb[2 * kk + p][16 * q + col];
                        acc[q][0][p] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bc, acc[q][0][p], 0, 0, 0);
                        acc[q][1][p] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bs, acc[q][1][p], 0, 0, 0);
                    }
                }
            }
            wave_sync_lds();
        }
    }

    // accumulator element r of lane l: ring pair 16q + l/16 + 4r, set l%16;
    // X_m = c (a_m - i b_m) mu/R, c = 1 (m = 0) or 1/2 (src/shs/shs_grd_fft_lc.c:24-44)
    const int set = set0 + (lane & 15);
    if (set >= A.nset) return;
    const double sc = (m == 0) ? A.mur : 0.5 * A.mur;
    double2 *out = (double2 *)A.spec + (size_t)set * A.nlat * A.lds;
#pragma unroll
    for (int q = 0; q < 4; q++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int rp = rt * 64 + 16 * q + (lane >> 4) + 4 * r;
            if (rp >= A.nring) continue;
            const double ec = acc[q][0][0][r], oc = acc[q][0][1][r];
            const double es = (m == 0) ? 0.0 : acc[q][1][0][r];
            const double os = (m == 0) ? 0.0 : acc[q][1][1][r];
            out[(size_t)rp * A.lds + m] = make_double2(sc * (ec + oc), -sc * (es + os));
            if (A.mirror[rp])
                out[(size_t)(A.nlat - 1 - rp) * A.lds + m] = make_double2(sc * (ec - oc), -sc * (es - os));
        }
}

// (C_nm s_nm, S_nm s_nm) for nset models of degree nmax, set-major
__global__ void scale_cs_multi_kernel(int nmax, const double2 *__restrict__ cs, size_t stride,
                                      const double *__restrict__ sn, int nt, double2 *__restrict__ out)
{
    const int m = blockIdx.y;
    const size_t set = blockIdx.z;
    const double2 *src = cs + set * stride + mcol_off(nmax, m) - m;
    double2 *dst = out + set * stride + mcol_off(nmax, m) - m;
    const double *f = sn + mcol_off(nt, m) - m;
    for (int n = m + blockIdx.x * blockDim.x + threadIdx.x; n <= nmax; n += gridDim.x * blockDim.x) {
        const double2 v = src[n];
        const double w = f[n];
        dst[n] = make_double2(v.x * w, v.y * w);
    }
}

int launch_shs_multi(const MultiShsArgs &a, void *stream)
{
    const int ngroups = (a.mcount + 7) / 8;
    const int ppx = (a.npair + 7) / 8;
    const unsigned nblk = 8u * 8u * (unsigned)ppx * (unsigned)ngroups;
    hipLaunchKernelGGL(shs_multi_kernel, dim3(nblk), dim3(64), 0, (hipStream_t)stream, a);
    return (int)hipGetLastError();
}

int launch_scale_cs_multi(int nmax, size_t nset, const void *cs, const double *sn, int nt, void *out,
                          void *stream)
{
    const unsigned gx = (unsigned)((nmax + 1 + 255) / 256);
    hipLaunchKernelGGL(scale_cs_multi_kernel, dim3(gx, (unsigned)nmax + 1, (unsigned)nset), dim3(256), 0,
                       (hipStream_t)stream, nmax, (const double2 *)cs, tri_len((unsigned long)nmax),
                       sn, nt, (double2 *)out);
    return (int)hipGetLastError();
}

}  // namespace charm_amd
