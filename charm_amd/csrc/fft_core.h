// Building blocks of the longitude transforms (host and device).
//
// The real transforms of length 2L along a parallel reduce to one complex DFT
// of length L (even samples in the real part, odd samples in the imaginary
// part).  That DFT runs in LDS as mixed-radix passes (radices 2, 3, 4, 5, 7,
// 8, 9 with twiddles, 16 as a last pass):
//
// * directly, when L = 2^a 3^b 5^c 7^d: decimation-in-frequency passes
//   (natural order in), the output read through a mixed-radix digit-reversal
//   table;
// * otherwise as Bluestein's chirp convolution over such a smooth M >= 2L - 1:
//   the forward passes in decimation in frequency, the pointwise product with
//   the chirp spectrum held in the same digit-reversed order, the inverse
//   passes in decimation in time (digit-reversed in, natural out), so no
//   reordering pass is needed.
//
// All passes work in place on one array; element e is stored at
// slot(e) = e + e/16, which spreads the short strides of the last passes over
// the LDS banks.  The plan (radices and their order) is chosen on the host by
// an operation-count model; the same plan drives the kernels and the host
// check of these blocks.
#ifndef CHARM_AMD_FFT_CORE_H
#define CHARM_AMD_FFT_CORE_H

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define FFT_HD __host__ __device__ __forceinline__
#else
#define FFT_HD inline
#endif

#include <cmath>
#include <vector>

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
// a + s * b
FFT_HD cplx axpy(double s, cplx b, cplx a)
{
    return mk(fma(s, b.x, a.x), fma(s, b.y, a.y));
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
FFT_HD void dft3(cplx *v)
{
    constexpr double S = 0.86602540378443864676;  // sin(2 pi / 3)
    const cplx s = add(v[1], v[2]);
    const cplx d = rot90<SIGN>(scale(sub(v[1], v[2]), S));
    const cplx m = axpy(-0.5, s, v[0]);
    v[0] = add(v[0], s);
    v[1] = add(m, d);
    v[2] = sub(m, d);
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

// Pairs k, R - k share cos(2 pi jk / R) on the sums and sin on the
// differences (R odd).
template <int SIGN>
FFT_HD void dft5(cplx *v)
{
    constexpr double C1 = 0.30901699437494742410, C2 = -0.80901699437494742410;
    constexpr double S1 = 0.95105651629515357212, S2 = 0.58778525229247312917;
    const cplx t1 = add(v[1], v[4]), t2 = add(v[2], v[3]);
    const cplx d1 = sub(v[1], v[4]), d2 = sub(v[2], v[3]);
    const cplx m1 = axpy(C2, t2, axpy(C1, t1, v[0]));
    const cplx m2 = axpy(C1, t2, axpy(C2, t1, v[0]));
    const cplx n1 = rot90<SIGN>(axpy(S2, d2, scale(d1, S1)));
    const cplx n2 = rot90<SIGN>(axpy(-S1, d2, scale(d1, S2)));
    v[0] = add(v[0], add(t1, t2));
    v[1] = add(m1, n1);
    v[4] = sub(m1, n1);
    v[2] = add(m2, n2);
    v[3] = sub(m2, n2);
}

template <int SIGN>
FFT_HD void dft7(cplx *v)
{
    constexpr double C1 = 0.62348980185873353053, C2 = -0.22252093395631440429,
                     C3 = -0.90096886790241912624;
    constexpr double S1 = 0.78183148246802980871, S2 = 0.97492791218182360702,
                     S3 = 0.43388373911755812048;
    const cplx t1 = add(v[1], v[6]), t2 = add(v[2], v[5]), t3 = add(v[3], v[4]);
    const cplx d1 = sub(v[1], v[6]), d2 = sub(v[2], v[5]), d3 = sub(v[3], v[4]);
    const cplx m1 = axpy(C3, t3, axpy(C2, t2, axpy(C1, t1, v[0])));
    const cplx m2 = axpy(C1, t3, axpy(C3, t2, axpy(C2, t1, v[0])));
    const cplx m3 = axpy(C2, t3, axpy(C1, t2, axpy(C3, t1, v[0])));
    const cplx n1 = rot90<SIGN>(axpy(S3, d3, axpy(S2, d2, scale(d1, S1))));
    const cplx n2 = rot90<SIGN>(axpy(-S1, d3, axpy(-S3, d2, scale(d1, S2))));
    const cplx n3 = rot90<SIGN>(axpy(S2, d3, axpy(-S1, d2, scale(d1, S3))));
    v[0] = add(v[0], add(t1, add(t2, t3)));
    v[1] = add(m1, n1);
    v[6] = sub(m1, n1);
    v[2] = add(m2, n2);
    v[5] = sub(m2, n2);
    v[3] = add(m3, n3);
    v[4] = sub(m3, n3);
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

// 9 points as 3 x 3: DFTs of the 3 interleaved columns, twiddles
// exp(SIGN 2 pi i n2 k1 / 9), DFTs of the rows
template <int SIGN>
FFT_HD void dft9(cplx *v)
{
    // cos and sin of 2 pi m / 9 for m = 1, 2, 4
    constexpr double CS1[2] = {0.76604444311897803520, 0.64278760968653932632};
    constexpr double CS2[2] = {0.17364817766693034885, 0.98480775301220805936};
    constexpr double CS4[2] = {-0.93969262078590838405, 0.34202014332566873304};
    cplx y[3][3];
#pragma unroll
    for (int n2 = 0; n2 < 3; n2++) {
        cplx c[3] = {v[n2], v[3 + n2], v[6 + n2]};
        dft3<SIGN>(c);
        y[n2][0] = c[0];
        y[n2][1] = c[1];
        y[n2][2] = c[2];
    }
    y[1][1] = mul(y[1][1], mk(CS1[0], SIGN * CS1[1]));
    y[1][2] = mul(y[1][2], mk(CS2[0], SIGN * CS2[1]));
    y[2][1] = mul(y[2][1], mk(CS2[0], SIGN * CS2[1]));
    y[2][2] = mul(y[2][2], mk(CS4[0], SIGN * CS4[1]));
#pragma unroll
    for (int k1 = 0; k1 < 3; k1++) {
        cplx r[3] = {y[0][k1], y[1][k1], y[2][k1]};
        dft3<SIGN>(r);
#pragma unroll
        for (int k2 = 0; k2 < 3; k2++) v[k1 + 3 * k2] = r[k2];
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
    if (R == 3) dft3<SIGN>(v);
    if (R == 4) dft4<SIGN>(v);
    if (R == 5) dft5<SIGN>(v);
    if (R == 7) dft7<SIGN>(v);
    if (R == 8) dft8<SIGN>(v);
    if (R == 9) dft9<SIGN>(v);
    if (R == 16) dft16<SIGN>(v);
}

// R-point DFT from a table w[q] = exp(SIGN 2 pi i q / R) (primes without a
// codelet); pairs k, R - k share the cosine sums of the pair sums and the sine
// sums of the pair differences, as in dft5 and dft7 (R odd)
template <int R>
FFT_HD void dft_table(cplx *v, const cplx *w)
{
    static_assert(R % 2 == 1, "odd R");
    cplx t[R / 2 + 1], d[R / 2 + 1];
#pragma unroll
    for (int j = 1; j <= R / 2; j++) {
        t[j] = add(v[j], v[R - j]);
        d[j] = sub(v[j], v[R - j]);
    }
    cplx out[R];
    cplx s0 = v[0];
#pragma unroll
    for (int j = 1; j <= R / 2; j++) s0 = add(s0, t[j]);
    out[0] = s0;
#pragma unroll
    for (int k = 1; k <= R / 2; k++) {
        // m = v0 + sum_j t_j cos(2 pi jk/R); n = i sign sum_j d_j sin(2 pi jk/R)
        cplx m = v[0], n = mk(0.0, 0.0);
#pragma unroll
        for (int j = 1; j <= R / 2; j++) {
            const cplx wk = w[(j * k) % R];
            m = axpy(wk.x, t[j], m);
            n = add(n, mk(-wk.y * d[j].y, wk.y * d[j].x));
        }
        out[k] = add(m, n);
        out[R - k] = sub(m, n);
    }
#pragma unroll
    for (int k = 0; k < R; k++) v[k] = out[k];
}

// exp(-2 pi i q / M) for 0 <= q < M from two short tables, hi[q / 64] =
// exp(-2 pi i 64 (q / 64) / M) and lo[q mod 64] = exp(-2 pi i (q mod 64) / M):
// ceil(M / 64) + 64 entries (a few KB, kept in LDS beside the row), one
// rounding beyond the tables' own (exact for q < 64, hi[0] = 1).
struct TwSplit {
    const cplx *hi, *lo;
    FFT_HD cplx operator()(int q) const
    {
        return mul(hi[q >> 6], lo[q & 63]);
    }
};
constexpr int TW_LO = 64;
FFT_HD int tw_hi_count(int M)
{
    return (M + TW_LO - 1) / TW_LO;
}

FFT_HD cplx sqr(cplx a)
{
    return mk(fma(a.x, a.x, -(a.y * a.y)), 2.0 * a.x * a.y);
}

// Twiddles w^r, r = 1..R-1, with w = exp(-2 pi i q / M): w from the tables,
// w^2, w^4, w^8 by squaring, the others as products of those (a few ulp at
// most; the table reads of the higher powers would collide in LDS banks,
// their indices r q stepping by even multiples across the lanes).
template <int R>
FFT_HD void twiddles(const TwSplit &tw, int q, cplx *w)
{
#pragma unroll
    for (int r = 1; r < R; r++) {
        const int hb = (r >= 8) ? 8 : (r >= 4) ? 4 : (r >= 2) ? 2 : 1;
        if (r == 1)
            w[1] = tw(q);
        else if (r == hb)
            w[r] = sqr(w[r / 2]);
        else
            w[r] = mul(w[hb], w[r - hb]);
    }
}

// Geometry of a twiddled pass: radix R on blocks of S points, stride
// st = S / R.  When st is a power of two of at least 16 (every pass of the
// plans above but the last, unless the plan is odd radices only) butterfly t
// splits by shifts and its R points sit at constant slot offsets
// r (st + st/16) from the first one.
struct PassGeom {
    int S, st, sh;   // sh = log2(st) if st = 2^sh >= 16, else -1
    int tws;         // twiddle index step M / S
};
FFT_HD PassGeom pass_geom(int M, int S, int R)
{
    PassGeom g;
    g.S = S;
    g.st = S / R;
    g.tws = M / S;
    g.sh = -1;
    if (g.st >= 16 && (g.st & (g.st - 1)) == 0) {
        int k = 0;
        while ((1 << k) < g.st) k++;
        g.sh = k;
    }
    return g;
}

// Slots of the R points of butterfly t, and its index j within the block
template <int R, bool POW2>
FFT_HD int butterfly_slots(const PassGeom &g, int t, int *sl)
{
    int b, j;
    if (POW2) {
        b = t >> g.sh;
        j = t & (g.st - 1);
        const int s0 = slot(b * g.S + j), ds = g.st + (g.st >> 4);
#pragma unroll
        for (int r = 0; r < R; r++) sl[r] = s0 + r * ds;
    } else {
        b = t / g.st;
        j = t - b * g.st;
        const int base = b * g.S + j;
#pragma unroll
        for (int r = 0; r < R; r++) sl[r] = slot(base + r * g.st);
    }
    return j;
}

// One butterfly of the decimation-in-frequency pass of radix R: butterfly t
// of M/R (DFT first, twiddles after).
template <int R, bool POW2>
FFT_HD void dif_butterfly(cplx *x, const TwSplit &tw, const PassGeom &g, int t)
{
    int sl[R];
    const int j = butterfly_slots<R, POW2>(g, t, sl);
    cplx v[R];
#pragma unroll
    for (int r = 0; r < R; r++) v[r] = x[sl[r]];
    dft<R, -1>(v);
    if (j != 0) {
        cplx w[R];
        twiddles<R>(tw, j * g.tws, w);
#pragma unroll
        for (int r = 1; r < R; r++) v[r] = mul(v[r], w[r]);
    }
#pragma unroll
    for (int r = 0; r < R; r++) x[sl[r]] = v[r];
}

// One butterfly of the decimation-in-time pass (inverse kernel) of radix R
// (conjugate twiddles first, DFT after).
template <int R, bool POW2>
FFT_HD void dit_butterfly(cplx *x, const TwSplit &tw, const PassGeom &g, int t)
{
    int sl[R];
    const int j = butterfly_slots<R, POW2>(g, t, sl);
    cplx v[R];
#pragma unroll
    for (int r = 0; r < R; r++) v[r] = x[sl[r]];
    if (j != 0) {
        cplx w[R];
        twiddles<R>(tw, j * g.tws, w);
#pragma unroll
        for (int r = 1; r < R; r++) v[r] = mul(v[r], conj(w[r]));
    }
    dft<R, +1>(v);
#pragma unroll
    for (int r = 0; r < R; r++) x[sl[r]] = v[r];
}

// Slot of point r of the contiguous block u of R points; for R dividing 16
// the block never straddles a padding step, so the offsets are constant
template <int R>
FFT_HD int block_slot(int u, int r)
{
    return (16 % R == 0) ? slot(u * R) + r : slot(u * R + r);
}

// The last decimation-in-frequency pass: blocks of R contiguous points, no
// twiddles (block u).
template <int R>
FFT_HD void last_block(cplx *x, int u)
{
    cplx v[R];
#pragma unroll
    for (int r = 0; r < R; r++) v[r] = x[block_slot<R>(u, r)];
    dft<R, -1>(v);
#pragma unroll
    for (int r = 0; r < R; r++) x[block_slot<R>(u, r)] = v[r];
}

// Last forward pass, product with the chirp spectrum, first inverse pass, in
// registers (block u of R contiguous points).  The first half of the block's
// spectrum values is requested before the LDS reads and the forward DFT, the
// second half right after the first half's products: two overlapped global
// round trips, where loads placed at their use (the register pressure of the
// 16-point DFT sinks them there) cost four in a row.
template <int R>
FFT_HD void middle_block(cplx *x, const cplx *B, int u)
{
    constexpr int H = (R + 1) / 2;
    const cplx *b = B + u * R;
    cplx w[H];
#pragma unroll
    for (int r = 0; r < H; r++) w[r] = b[r];
    cplx v[R];
#pragma unroll
    for (int r = 0; r < R; r++) v[r] = x[block_slot<R>(u, r)];
    dft<R, -1>(v);
#pragma unroll
    for (int r = 0; r < H; r++) v[r] = mul(v[r], w[r]);
#pragma unroll
    for (int r = H; r < R; r++) w[r - H] = b[r];
#pragma unroll
    for (int r = H; r < R; r++) v[r] = mul(v[r], w[r - H]);
    dft<R, +1>(v);
#pragma unroll
    for (int r = 0; r < R; r++) x[block_slot<R>(u, r)] = v[r];
}

// ---------------------------------------------------------------------------
// Plans (host)
// ---------------------------------------------------------------------------
constexpr int MAXPASS = 16;
// LDS holds slots(M) complex values and the twiddle tables: M <= MMAX keeps
// one row within 160 KB
constexpr int MMAX = 9400;

// How the length-L DFT runs: passes on L itself, Bluestein's chirp
// convolution over M >= 2L - 1, or (L prime) Rader's cyclic convolution of
// length M = L - 1 with the points reordered by powers of a primitive root g
enum { DFT_DIRECT = 0, DFT_BLUESTEIN = 1, DFT_RADER = 2 };

struct MixPlan {
    int M = 1;           // transform length (direct) or convolution length
    int np = 0;          // passes, decimation-in-frequency order
    int R[MAXPASS] = {};
    int mode = DFT_DIRECT;
    int g = 0;           // primitive root (Rader)
};

// Radices of M = 2^a 3^b 5^c 7^d (false if M has another prime factor).  The
// last pass (contiguous blocks, no twiddles) takes 16 when 16 divides M; the
// twiddled passes come odd radices first, then 2 or 4, then 8s, so that with
// a last pass of 16 every earlier pass has a stride of at least 16.
inline bool factor_plan(int M, MixPlan &p)
{
    if (M < 1) return false;
    int a = 0, b = 0, c = 0, d = 0, v = M;
    while (v % 2 == 0) v /= 2, a++;
    while (v % 3 == 0) v /= 3, b++;
    while (v % 5 == 0) v /= 5, c++;
    while (v % 7 == 0) v /= 7, d++;
    if (v != 1) return false;
    p = MixPlan();
    p.M = M;
    std::vector<int> tw;
    int last = 0;
    if (a >= 4) {
        last = 16;
        a -= 4;
    }
    for (; d > 0; d--) tw.push_back(7);
    for (; c > 0; c--) tw.push_back(5);
    for (; b >= 2; b -= 2) tw.push_back(9);
    if (b == 1) tw.push_back(3);
    if (a % 3 == 1) tw.push_back(2);
    if (a % 3 == 2) tw.push_back(4);
    for (int k = 0; k < a / 3; k++) tw.push_back(8);
    if (!last && !tw.empty()) {
        last = tw.back();
        tw.pop_back();
    }
    if ((int)tw.size() + (last ? 1 : 0) > MAXPASS) return false;
    for (int r : tw) p.R[p.np++] = r;
    if (last) p.R[p.np++] = last;
    return true;
}

// Relative cost of one point through one pass of radix R (LDS round trip,
// butterfly arithmetic, twiddle products; the last pass has no twiddles)
inline double pass_cost(int R, bool last)
{
    double c = 8.0;
    switch (R) {
    case 2: c += 2.0; break;
    case 3: c += 5.3; break;
    case 4: c += 4.0; break;
    case 5: c += 8.8; break;
    case 7: c += 15.0; break;
    case 8: c += 7.5; break;
    case 9: c += 13.0; break;
    default: c += 11.4; break;
    }
    if (!last) c += 6.0 * (R - 1) / R;
    return c;
}

inline double plan_cost(const MixPlan &p)
{
    double s = 0.0;
    for (int i = 0; i < p.np; i++) s += pass_cost(p.R[i], i == p.np - 1);
    return s * p.M;
}

inline long long pow_mod(long long b, long long e, long long m)
{
    long long r = 1 % m;
    b %= m;
    while (e > 0) {
        if (e & 1) r = r * b % m;
        b = b * b % m;
        e >>= 1;
    }
    return r;
}

inline bool is_prime(int n)
{
    if (n < 2) return false;
    for (int d = 2; (long)d * d <= n; d++)
        if (n % d == 0) return false;
    return true;
}

// Smallest primitive root of the prime L: g^((L-1)/f) != 1 for every prime f
// dividing L - 1
inline int primitive_root(int L)
{
    std::vector<int> f;
    int v = L - 1;
    for (int d = 2; (long)d * d <= v; d++)
        if (v % d == 0) {
            f.push_back(d);
            while (v % d == 0) v /= d;
        }
    if (v > 1) f.push_back(v);
    for (int g = 2; g < L; g++) {
        bool ok = true;
        for (int q : f)
            if (pow_mod(g, (L - 1) / q, L) == 1) {
                ok = false;
                break;
            }
        if (ok) return g;
    }
    return 1;
}

// Plan of a complex DFT of length L: direct if L is smooth (and within LDS);
// else the cheaper of Bluestein over the cheapest smooth M in
// [2L - 1, 2 (2L - 1)] and, for L prime with L - 1 smooth, Rader over L - 1.
// Returns false if no plan fits (then the caller uses a library FFT).
inline bool choose_plan(int L, MixPlan &p)
{
    MixPlan q;
    if (L <= MMAX && factor_plan(L, q)) {
        p = q;
        p.mode = DFT_DIRECT;
        return true;
    }
    const long lo = 2L * L - 1;
    double best = -1.0;
    for (long M = lo; M <= 2 * lo && M <= MMAX; M++) {
        if (!factor_plan((int)M, q)) continue;
        const double c = 2.0 * plan_cost(q);
        if (best < 0.0 || c < best) {
            best = c;
            p = q;
            p.mode = DFT_BLUESTEIN;
        }
    }
    if (L > 2 && L - 1 <= MMAX && is_prime(L) && factor_plan(L - 1, q) && q.np > 0) {
        // the reordering gather and scatter cost about one more pass
        const double c = 2.0 * plan_cost(q) + 10.0 * L;
        if (best < 0.0 || c < best) {
            best = c;
            p = q;
            p.mode = DFT_RADER;
            p.g = primitive_root(L);
        }
    }
    return best >= 0.0;
}

// Rader tables of the prime L with primitive root g: lg[n] = log_g n for
// n = 1..L-1 (lg[0] = 0), and the filter b_m = exp(-2 pi i (g^-m mod L) / L),
// m = 0..L-2, with the exponent reduced exactly.
inline std::vector<int> rader_logs(int L, int g)
{
    std::vector<int> lg(L, 0);
    long long v = 1;
    for (int q = 0; q < L - 1; q++) {
        lg[(int)v] = q;
        v = v * g % L;
    }
    return lg;
}

inline std::vector<cplx> rader_filter(int L, int g)
{
    std::vector<cplx> b(L - 1);
    const long long gi = pow_mod(g, L - 2, L);  // g^-1
    long long v = 1;
    for (int m = 0; m < L - 1; m++) {
        const double a = 2.0 * 3.14159265358979323846 * (double)v / (double)L;
        b[m] = mk(std::cos(a), -std::sin(a));
        v = v * gi % L;
    }
    return b;
}

// Where the decimation-in-frequency passes leave output k: digits of k in the
// mixed radix R[0], R[1], ... (least significant first) weighted by the
// block strides M / (R[0] ... R[i]).
inline std::vector<int> digit_reversal(const MixPlan &p)
{
    std::vector<int> rev(p.M);
    for (int k = 0; k < p.M; k++) {
        int rem = k, pos = 0, span = p.M;
        for (int i = 0; i < p.np; i++) {
            span /= p.R[i];
            pos += (rem % p.R[i]) * span;
            rem /= p.R[i];
        }
        rev[k] = pos;
    }
    return rev;
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
