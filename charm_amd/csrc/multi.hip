// Many coefficient sets on one Gauss-Legendre grid: the sums over degrees
// (synthesis) are dense contractions of a latitude panel of Legendre values
// with a block of coefficient sets, and run on the FP64 matrix cores
// (v_mfma_f64_16x16x4f64).
//
// Synthesis, per order m: a[ring][set] = sum_n Q_n(ring) (C_n s_n)[set], with
// even and odd (n + m) kept apart for the equatorial fold
// (src/shs/shs_point_kernel.c:1014-1623 evaluates the same sums one set at a
// time).  One wave handles (m, 16 sets, 64 ring pairs), one ring pair per
// lane, eight degrees at a time:
//  * each lane runs the scaled recurrence of its ring pair (the alphas of the
//    column are wave-uniform: scalar loads);
//  * the eight values of a lane go into the MFMA B-operands (4 degrees of one
//    parity x 16 ring pairs) by a 4x4 transpose across the wave's four 16-lane
//    rows with the gfx950 lane swaps (v_permlane32_swap, v_permlane16_swap),
//    without LDS: row q of lanes holds ring tile q, register d degree d, and
//    after the transpose register q of row k holds degree 2k + parity of
//    ring tile q;
//  * the A-operands (16 sets x 4 degrees) come straight from a copy of the
//    coefficients whose triangle entries hold the sets side by side (one
//    coalesced 16-byte load per lane and parity), fetched a tile ahead;
//  * 16 accumulator tiles (4 ring tiles x {C, S} x {even, odd}) collect
//    D[set][ring pair].
#include <hip/hip_runtime.h>

#include <algorithm>

#include "internal.h"

namespace charm_amd {

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef const __attribute__((address_space(4))) double cdouble;

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

// x in rows 2, 3 <-> y in rows 0, 1 (rows of 16 lanes)
__device__ __forceinline__ void swap32(double &x, double &y)
{
    const int2 a = __builtin_bit_cast(int2, x), b = __builtin_bit_cast(int2, y);
    const auto lo = __builtin_amdgcn_permlane32_swap(a.x, b.x, false, false);
    const auto hi = __builtin_amdgcn_permlane32_swap(a.y, b.y, false, false);
    x = __builtin_bit_cast(double, make_int2(lo[0], hi[0]));
    y = __builtin_bit_cast(double, make_int2(lo[1], hi[1]));
}

// x in rows 1, 3 <-> y in rows 0, 2
__device__ __forceinline__ void swap16(double &x, double &y)
{
    const int2 a = __builtin_bit_cast(int2, x), b = __builtin_bit_cast(int2, y);
    const auto lo = __builtin_amdgcn_permlane16_swap(a.x, b.x, false, false);
    const auto hi = __builtin_amdgcn_permlane16_swap(a.y, b.y, false, false);
    x = __builtin_bit_cast(double, make_int2(lo[0], hi[0]));
    y = __builtin_bit_cast(double, make_int2(lo[1], hi[1]));
}

// v[c] of row r -> v[r] of row c (4x4 over the wave's rows and the values)
__device__ __forceinline__ void transpose4(double v[4])
{
    swap32(v[0], v[2]);
    swap32(v[1], v[3]);
    swap16(v[0], v[1]);
    swap16(v[2], v[3]);
}

}  // namespace

// Block b (one wave) -> (m, set tile, ring tile).  Blocks run on XCD b mod 8.
// A unit is (order group of 8 orders, set tile); XCD x takes the units
// u = x, x + 8, ... and runs, per unit, its ring tiles one after another with
// the 8 orders of the group back to back: the ring tiles of one (order, set
// tile) read the same coefficient column through that XCD's L2, and the
// 16-byte outputs of consecutive orders meet in it as whole lines.
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) void shs_multi_kernel(MultiShsArgs A)
{
    const int lane = threadIdx.x;
    const int b = blockIdx.x;
    const int x = b & 7, s = b >> 3;
    const int per = A.nrt * 8;
    const int u = (s / per) * 8 + x;
    if (u >= A.nunit) return;
    const int rem = s % per;
    const int rt = rem >> 3;
    const int g = u / A.nst, st = u % A.nst;
    const int m = uni(8 * g + (rem & 7));
    if (m >= A.mcount) return;
    const int nmax = A.nmax;
    const int ring = rt * 64 + lane;
    const bool valid = ring < A.nring;
    const int rc = min(ring, A.nring - 1);

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
    const bool skip = !valid || (A.pm0 != nullptr && m >= A.pm0[rc]);
    if (m <= nmax && !all64(skip)) {
        const double t = valid ? A.t[rc] : 0.0;
        double xq, yq = 0.0;
        int e = 0;
        if (m == 0) {
            xq = skip ? 0.0 : 1.0;
        } else {
            const size_t srow = (size_t)(m - 1) * A.sld + rc;
            xq = skip ? 0.0 : A.sx[srow];
            e = skip ? 0 : 960 * A.six[srow];
        }
        cdouble *calc = (cdouble *)(A.al + mcol_off(A.al_nt, m) - m);
        // A-operand of lane l: set st*16 + l%16, degree n0 + 2(l/16) + parity
        // (the padding sets of the copy hold zeros)
        const int k = lane >> 4;
        const int set = st * 16 + (lane & 15);
        const double2 *ccol = (const double2 *)A.cs + (mcol_off(nmax, m) - m) * (size_t)A.nsetp + set;
        const double2 *pk = ccol + (size_t)(2 * k) * A.nsetp;
        // blocks whose 8 degrees are all <= nmax
        auto aload = [&](int n0, double2 *a2) {
            const double2 *q = pk + (size_t)n0 * A.nsetp;
            a2[0] = q[0];
            a2[1] = q[A.nsetp];
        };
        // any block (degrees above nmax read degree nmax; their Q values are 0)
        auto aload_clamped = [&](int n0, double2 *a2) {
#pragma unroll
            for (int p = 0; p < 2; p++) a2[p] = ccol[(size_t)min(n0 + 2 * k + p, nmax) * A.nsetp];
        };
        // 16 MFMAs of one block: per parity, the lanes' values of the four
        // degrees of that parity transposed into the B-operands of the four
        // ring tiles
        auto mma = [&](const double v[8], const double2 *a2) {
#pragma unroll
            for (int p = 0; p < 2; p++) {
                double bq[4] = {v[p], v[2 + p], v[4 + p], v[6 + p]};
                transpose4(bq);
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    acc[q][0][p] = __builtin_amdgcn_mfma_f64_16x16x4f64(a2[p].x, bq[q], acc[q][0][p], 0, 0, 0);
                    acc[q][1][p] = __builtin_amdgcn_mfma_f64_16x16x4f64(a2[p].y, bq[q], acc[q][1][p], 0, 0, 0);
                }
            }
        };
        bool xmode = any64(e != 0);
        // whole wave below 2^-1100 for the next 8 degrees (its Q values are 0)
        auto deep_now = [&]() {
            if (!xmode) return false;
            int ex, ey;
            (void)frexp(xq, &ex);
            (void)frexp(yq, &ey);
            return all64(!valid || (e != 0 && e + max(ex, ey) < -1180));
        };
        // first or last block: degrees n0..n0+7, those above nmax zero, the
        // seed at degree m (alphas read one by one, clamped into the column)
        auto edge_block = [&](int n0, double2 *a2) {
            aload_clamped(n0, a2);
            const bool deep = deep_now();
            double v[8];
#pragma unroll
            for (int j = 0; j < 8; j++) {
                double z = xq;
                if (n0 + j != m) {
                    z = fma(calc[min(n0 + j, A.al_nt)] * xq, t, -yq);
                    yq = xq;
                    xq = z;
                }
                v[j] = (n0 + j <= nmax) ? (xmode ? ldexp(z, e) : z) : 0.0;
            }
            if (xmode) {
                renorm(xq, yq, e);
                xmode = any64(e != 0);
            }
            if (!deep) mma(v, a2);
        };
        // a whole block: degrees n0..n0+7 <= nmax, alphas and A-operands
        // requested earlier; requests the next whole block's alphas
        double alc[8];
        auto block = [&](int n0, const double2 *a2) {
            double al[8];
#pragma unroll
            for (int j = 0; j < 8; j++) al[j] = alc[j];
            if (n0 + 15 <= nmax) {
#pragma unroll
                for (int j = 0; j < 8; j++) alc[j] = calc[n0 + 8 + j];
            }
            const bool deep = deep_now();
            double v[8];
            if (xmode) {
#pragma unroll
                for (int j = 0; j < 8; j++) {
                    const double z = fma(al[j] * xq, t, -yq);
                    yq = xq;
                    xq = z;
                    v[j] = ldexp(z, e);
                }
                renorm(xq, yq, e);
                xmode = any64(e != 0);
            } else {
#pragma unroll
                for (int j = 0; j < 8; j++) {
                    const double z = fma(al[j] * xq, t, -yq);
                    yq = xq;
                    xq = z;
                    v[j] = z;
                }
            }
            if (!deep) mma(v, a2);
        };

        double2 aA[2], aB[2];
        edge_block(m, aA);
        int n0 = m + 8;
        // whole blocks in pairs, alternating A-operand buffers; a buffer is
        // refilled right after its block, two blocks before it is used
        if (n0 + 7 <= nmax) {
            aload(n0, aA);
#pragma unroll
            for (int j = 0; j < 8; j++) alc[j] = calc[n0 + j];
        }
        if (n0 + 15 <= nmax) aload(n0 + 8, aB);
        for (; n0 + 15 <= nmax; n0 += 16) {
            block(n0, aA);
            if (n0 + 23 <= nmax) aload(n0 + 16, aA);
            block(n0 + 8, aB);
            if (n0 + 31 <= nmax) aload(n0 + 24, aB);
        }
        if (n0 + 7 <= nmax) {
            block(n0, aA);
            n0 += 8;
        }
        if (n0 <= nmax) edge_block(n0, aB);
    }

    // accumulator element r of lane l: set st*16 + l/16 + 4r, ring pair
    // 64 rt + 16q + l%16; X_m = c (a_m - i b_m) mu/R, c = 1 (m = 0) or 1/2
    // (src/shs/shs_grd_fft_lc.c:24-44)
    const double sc = (m == 0) ? A.mur : 0.5 * A.mur;
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const int os = st * 16 + (lane >> 4) + 4 * r;
        if (os >= A.nset) continue;
        double2 *out = (double2 *)A.spec + (size_t)os * A.nlat * A.lds;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int rp = rt * 64 + 16 * q + (lane & 15);
            if (rp >= A.nring) continue;
            const double ec = acc[q][0][0][r], oc = acc[q][0][1][r];
            const double es = (m == 0) ? 0.0 : acc[q][1][0][r];
            const double os_ = (m == 0) ? 0.0 : acc[q][1][1][r];
            out[(size_t)rp * A.lds + m] = make_double2(sc * (ec + oc), -sc * (es + os_));
            if (A.mirror[rp])
                out[(size_t)(A.nlat - 1 - rp) * A.lds + m] =
                    make_double2(sc * (ec - oc), -sc * (es - os_));
        }
    }
}

// (C_nm s_nm, S_nm s_nm) for nset models of degree nmax (set-major input,
// stride P pairs per set) into entry (m, n) x nsetp sets of the output, the
// padding sets zeroed.  A block transposes a tile of 32 degrees x 64 sets of
// one order through LDS: read along the degrees of each set, written along
// the sets of each degree.
constexpr int SCM_N = 32, SCM_S = 64;

__global__ __launch_bounds__(256) void scale_cs_multi_kernel(int nmax, const double2 *__restrict__ cs,
                                                             size_t stride, int nset, int nsetp,
                                                             const double *__restrict__ sn, int nt,
                                                             double2 *__restrict__ out)
{
    __shared__ double2 tile[SCM_N][SCM_S + 1];
    const int m = blockIdx.y;
    const int n0 = m + (int)blockIdx.x * SCM_N;
    if (n0 > nmax) return;
    const int set0 = (int)blockIdx.z * SCM_S;
    const size_t col = mcol_off(nmax, m) - m;
    const double *f = sn + mcol_off(nt, m) - m;
    for (int i = threadIdx.x; i < SCM_N * SCM_S; i += blockDim.x) {
        const int sl = i / SCM_N, nl = i % SCM_N;
        const int set = set0 + sl, n = n0 + nl;
        double2 v = make_double2(0.0, 0.0);
        if (set < nset && n <= nmax) {
            v = cs[(size_t)set * stride + col + n];
            const double w = f[n];
            v = make_double2(v.x * w, v.y * w);
        }
        tile[nl][sl] = v;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < SCM_N * SCM_S; i += blockDim.x) {
        const int nl = i / SCM_S, sl = i % SCM_S;
        const int set = set0 + sl, n = n0 + nl;
        if (set < nsetp && n <= nmax) out[(col + n) * (size_t)nsetp + set] = tile[nl][sl];
    }
}

int launch_shs_multi(const MultiShsArgs &a, void *stream)
{
    const size_t nblk = 8 * (size_t)((a.nunit + 7) / 8) * (size_t)a.nrt * 8;
    if (nblk > (size_t)INT32_MAX) return (int)hipErrorInvalidValue;
    hipLaunchKernelGGL(shs_multi_kernel, dim3((unsigned)nblk), dim3(64), 0, (hipStream_t)stream, a);
    return (int)hipGetLastError();
}

int launch_scale_cs_multi(int nmax, size_t nset, const void *cs, const double *sn, int nt, void *out,
                          int nsetp, void *stream)
{
    if (nset == 0 || nsetp < (int)nset || nsetp % 16) return (int)hipErrorInvalidValue;
    const unsigned gx = (unsigned)((nmax + SCM_N) / SCM_N);
    hipLaunchKernelGGL(scale_cs_multi_kernel,
                       dim3(gx, (unsigned)nmax + 1, (unsigned)((nsetp + SCM_S - 1) / SCM_S)), dim3(256),
                       0, (hipStream_t)stream, nmax, (const double2 *)cs, tri_len((unsigned long)nmax),
                       (int)nset, nsetp, sn, nt, (double2 *)out);
    return (int)hipGetLastError();
}

}  // namespace charm_amd
