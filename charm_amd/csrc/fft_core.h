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

// 16 points as 4 x 4: DFTs of the 4 interleaved columns, twiddles
// exp(SIGN 2 pi i n2 k1 / 16), DFTs of the rows
template <int SIGN>
FFT_HD void dft16(cplx *v)
{
    constexpr double C1 = 0.92387953251128675613;  // cos(pi/8)
    constexpr double S1 = 0.38268343236508977173;  // sin(pi/8)
    constexpr double H = 0.70710678118654752440;
    // cos and sin of 2 pi m / 16 for m = 0..9
    constexpr double CS[10][2] = {{1.0, 0.0}, {C1, S1}, {H, H},   {S1, C1},  {0.0, 1.0},
                                  {-S1, C1},  {-H, H},  {-C1, S1}, {-1.0, 0.0}, {-C1, -S1}};
    cplx y[4][4];
#pragma unroll
    for (int n2 = 0; n2 < 4; n2++) {
        cplx c[4] = {v[n2], v[4 + n2], v[8 + n2], v[12 + n2]};
        dft4<SIGN>(c);
#pragma unroll
        for (int k1 = 0; k1 < 4; k1++) {
            const int m = n2 * k1;
            if (m == 0)
                y[n2][k1] = c[k1];
            else
                y[n2][k1] = mul(c[k1], mk(CS[m][0], SIGN * CS[m][1]));
        }
    }
#pragma unroll
    for (int k1 = 0; k1 < 4; k1++) {
        cplx r[4] = {y[0][k1], y[1][k1], y[2][k1], y[3][k1]};
        dft4<SIGN>(r);
#pragma unroll
        for (int k2 = 0; k2 < 4; k2++) v[k1 + 4 * k2] = r[k2];
    }
}

template <int R, int SIGN>
FFT_HD void dft(cplx *v)
{
    if (R == 2) dft2<SIGN>(v);
    if (R == 4) dft4<SIGN>(v);
    if (R == 8) dft8<SIGN>(v);
    if (R == 16) dft16<SIGN>(v);
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
    if (R >= 16) {
        w[8] = tw[(8 * q) & (M - 1)];
        for (int r = 9; r < 16; r++) w[r] = mul(w[8], w[r - 8]);
    }
}

// The three table values a radix-8 pass needs: w^1, w^2, w^4
struct Tw8 {
    cplx w1, w2, w4;
};
FFT_HD Tw8 tw8_load(const cplx *tw, int q, int M)
{
    Tw8 t;
    t.w1 = tw[q];
    t.w2 = tw[(2 * q) & (M - 1)];
    t.w4 = tw[(4 * q) & (M - 1)];
    return t;
}
FFT_HD void tw8_expand(const Tw8 &t, cplx *w)
{
    w[1] = t.w1;
    w[2] = t.w2;
    w[3] = mul(t.w1, t.w2);
    w[4] = t.w4;
    w[5] = mul(t.w4, t.w1);
    w[6] = mul(t.w4, t.w2);
    w[7] = mul(t.w4, w[3]);
}

// Radix-8 butterflies with the twiddles already loaded (j = 0 gives w = 1)
FFT_HD void dif_butterfly8(cplx *x, const Tw8 &t, int S, int b, int j)
{
    const int st = S / 8;
    const int base = b * S + j;
    cplx v[8];
#pragma unroll
    for (int r = 0; r < 8; r++) v[r] = x[slot(base + r * st)];
    dft8<-1>(v);
    cplx w[8];
    tw8_expand(t, w);
#pragma unroll
    for (int r = 1; r < 8; r++) v[r] = mul(v[r], w[r]);
#pragma unroll
    for (int r = 0; r < 8; r++) x[slot(base + r * st)] = v[r];
}
FFT_HD void dit_butterfly8(cplx *x, const Tw8 &t, int S, int b, int j)
{
    const int st = S / 8;
    const int base = b * S + j;
    cplx v[8];
#pragma unroll
    for (int r = 0; r < 8; r++) v[r] = x[slot(base + r * st)];
    cplx w[8];
    tw8_expand(t, w);
#pragma unroll
    for (int r = 1; r < 8; r++) v[r] = mul(v[r], conj(w[r]));
    dft8<+1>(v);
#pragma unroll
    for (int r = 0; r < 8; r++) x[slot(base + r * st)] = v[r];
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

// Pass plan for M = 2^p (pass 0 = first decimation-in-frequency pass, on
// blocks of S = M): an optional radix-2 or -4 pass, radix-8 passes, and a
// final radix-16 pass on contiguous 16-point blocks (p = 3: one radix-8
// pass).  Every pass but the last then has a stride S/R of at least 16, so
// 16 consecutive lanes touch 16 consecutive elements, and in the last pass
// lane b touches 16b..16b+15, i.e. slots 17b..17b+15: distinct banks either way.
constexpr FFT_HD int plan_first(int p)
{
    if (p <= 4) return 0;
    const int rem = (p - 4) % 3;
    return rem == 0 ? 0 : (1 << rem);
}
constexpr FFT_HD int plan_n8(int p)
{
    return p == 3 ? 1 : (p <= 4 ? 0 : (p - 4) / 3);
}
constexpr FFT_HD int pass_count(int p)
{
    return (plan_first(p) ? 1 : 0) + plan_n8(p) + (p >= 4 ? 1 : 0);
}
constexpr FFT_HD int pass_radix(int p, int i)
{
    if (plan_first(p)) {
        if (i == 0) return plan_first(p);
        i--;
    }
    return i < plan_n8(p) ? 8 : 16;
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
