// Building blocks of the longitude transforms (host and device).
//
// The real transforms of length 2L along a parallel reduce to one complex DFT
// of length L (even samples in the real part, odd samples in the imaginary
// part).  A complex DFT of any length L runs as Bluestein's chirp convolution
// over a power-of-two length M >= 2L - 1: forward FFT in decimation-in-
// frequency passes (natural order in, digit-reversed order out), pointwise
// product with the chirp spectrum held in the same digit-reversed order, and
// the inverse FFT in decimation-in-time passes (digit-reversed in, natural
// out), so no reordering pass is needed.  All passes work in place on one
// array; element e is stored at slot(e) = e + e/16, which spreads the short
// strides of the last passes over the LDS banks.
#ifndef CHARM_AMD_FFT_CORE_H
#define CHARM_AMD_FFT_CORE_H

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define FFT_HD __host__ __device__ __forceinline__
#else
#include <cmath>
#define FFT_HD inline
#endif

namespace charm_amd {
namespace fft {

struct alignas(16) cplx {
    double x, y;
};

FFT_HD cplx mk(double x, double y)
{
    cplx c;
    c.x = x;
    c.y = y;
    return c;
}
FFT_HD cplx add(cplx a, cplx b)
{
    return mk(a.x + b.x, a.y + b.y);
}
FFT_HD cplx sub(cplx a, cplx b)
{
    return mk(a.x - b.x, a.y - b.y);
}
FFT_HD cplx mul(cplx a, cplx b)
{
    return mk(fma(a.x, b.x, -(a.y * b.y)), fma(a.x, b.y, a.y * b.x));
}
FFT_HD cplx conj(cplx a)
{
    return mk(a.x, -a.y);
}
FFT_HD cplx scale(cplx a, double s)
{
    return mk(a.x * s, a.y * s);
}
// a * (SIGN i): SIGN = -1 multiplies by -i
template <int SIGN>
FFT_HD cplx rot90(cplx a)
{
    return SIGN < 0 ? mk(a.y, -a.x) : mk(-a.y, a.x);
}

FFT_HD int slot(int e)
{
    return e + (e >> 4);
}
// slots of an M-point array
FFT_HD int slots(int M)
{
    return M + (M >> 4);
}

// In-place DFT of R points with kernel exp(SIGN 2 pi i jk / R), natural order.
template <int SIGN>
FFT_HD void dft2(cplx *v)
{
    const cplx a = v[0], b = v[1];
    v[0] = add(a, b);
    v[1] = sub(a, b);
}

template <int SIGN>
FFT_HD void dft4(cplx *v)
{
    const cplx c0 = add(v[0], v[2]), c1 = sub(v[0], v[2]);
    const cplx c2 = add(v[1], v[3]), c3 = rot90<SIGN>(sub(v[1], v[3]));
    v[0] = add(c0, c2);
    v[2] = sub(c0, c2);
    v[1] = add(c1, c3);
    v[3] = sub(c1, c3);
}

template <int SIGN>
FFT_HD void dft8(cplx *v)
{
    constexpr double H = 0.70710678118654752440;  // sqrt(1/2)
    cplx a[4], b[4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
        a[r] = add(v[r], v[r + 4]);
        b[r] = sub(v[r], v[r + 4]);
    }
    // b_r *= exp(SIGN 2 pi i r / 8)
    b[1] = scale(add(b[1], rot90<SIGN>(b[1])), H);
    b[2] = rot90<SIGN>(b[2]);
    b[3] = scale(sub(rot90<SIGN>(b[3]), b[3]), H);
    dft4<SIGN>(a);
    dft4<SIGN>(b);
#pragma unroll
    for (int r = 0; r < 4; r++) {
        v[2 * r] = a[r];
        v[2 * r + 1] = b[r];
    }
}

template <int R, int SIGN>
FFT_HD void dft(cplx *v)
{
    if (R == 2) dft2<SIGN>(v);
    if (R == 4) dft4<SIGN>(v);
    if (R == 8) dft8<SIGN>(v);
}

// Twiddles w^r, r = 1..R-1, with w = tw[q] = exp(-2 pi i q / M) (tw has M
// entries); three table reads at most, two roundings at most.
template <int R>
FFT_HD void twiddles(const cplx *tw, int q, int M, cplx *w)
{
    w[1] = tw[q];
    if (R >= 4) {
        w[2] = tw[(2 * q) & (M - 1)];
        w[3] = mul(w[1], w[2]);
    }
    if (R >= 8) {
        w[4] = tw[(4 * q) & (M - 1)];
        w[5] = mul(w[4], w[1]);
        w[6] = mul(w[4], w[2]);
        w[7] = mul(w[4], w[3]);
    }
}

// One butterfly of the decimation-in-frequency pass of radix R on blocks of
// S points: butterfly t of M/R.
template <int R>
FFT_HD void dif_butterfly(cplx *x, const cplx *tw, int M, int S, int t)
{
    const int st = S / R;
    const int b = t / st, j = t - b * st;
    const int base = b * S + j;
    cplx v[R];
#pragma unroll
    for (int r = 0; r < R; r++) v[r] = x[slot(base + r * st)];
    dft<R, -1>(v);
    if (j != 0) {
        cplx w[R];
        twiddles<R>(tw, j * (M / S), M, w);
#pragma unroll
        for (int r = 1; r < R; r++) v[r] = mul(v[r], w[r]);
    }
#pragma unroll
    for (int r = 0; r < R; r++) x[slot(base + r * st)] = v[r];
}

// One butterfly of the decimation-in-time pass (inverse kernel) of radix R on
// blocks of S points.
template <int R>
FFT_HD void dit_butterfly(cplx *x, const cplx *tw, int M, int S, int t)
{
    const int st = S / R;
    const int b = t / st, j = t - b * st;
    const int base = b * S + j;
    cplx v[R];
#pragma unroll
    for (int r = 0; r < R; r++) v[r] = x[slot(base + r * st)];
    if (j != 0) {
        cplx w[R];
        twiddles<R>(tw, j * (M / S), M, w);
#pragma unroll
        for (int r = 1; r < R; r++) v[r] = mul(v[r], conj(w[r]));
    }
    dft<R, +1>(v);
#pragma unroll
    for (int r = 0; r < R; r++) x[slot(base + r * st)] = v[r];
}

// Radix of pass i (0 = first decimation-in-frequency pass, S = M) for
// M = 2^p: radix 8 throughout, the remainder (2 or 4) in the last pass.
FFT_HD int pass_count(int p)
{
    return (p + 2) / 3;
}
FFT_HD int pass_radix(int p, int i)
{
    const int n8 = p / 3, rem = p - 3 * n8;
    if (i < n8) return 8;
    return rem == 1 ? 2 : 4;
}

// exp(SIGN i pi (n^2 mod 2L) / L) with the exponent reduced exactly
FFT_HD cplx chirp(long long n, long long L, int sign)
{
    const long long q = (n * n) % (2 * L);
    const double ang = 3.14159265358979323846 * (double)q / (double)L;
    double s, c;
#if defined(__HIP_DEVICE_COMPILE__)
    sincos(ang, &s, &c);
#else
    s = std::sin(ang);
    c = std::cos(ang);
#endif
    return mk(c, sign * s);
}

}  // namespace fft
}  // namespace charm_amd

#endif
