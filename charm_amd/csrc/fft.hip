// Longitude transforms of Gauss-Legendre rows (nlon = 2L points per row):
// r2c (analysis) and c2r (synthesis) with hipFFT's unnormalised conventions,
// as one complex DFT of length L per row evaluated by Bluestein's chirp
// convolution of length M = 2^P <= 8192 entirely in LDS, one row per
// work-group (src/shs/shs_grd_fft.c:99-130 and src/sha/sha_point.c:490-553
// use FFTW's r2c/c2r for the same transforms).
#include <hip/hip_runtime.h>

#include <map>
#include <algorithm>
#include <mutex>
#include <vector>

#include "fft_core.h"
#include "internal.h"

namespace charm_amd {

using fft::cplx;

namespace {

struct RowArgs {
    int L;
    const cplx *w;    // [L]   exp(-i pi n^2 / L)
    const cplx *pt;   // [L+1] exp(-i pi k / L)
    const cplx *B;    // [M]   DIF_M of the chirp filter (digit-reversed order)
    const cplx *tw;   // [M]   exp(-2 pi i q / M)
    double *f;        // rows of 2L reals
    size_t ldf;       // doubles between rows of f
    cplx *spec;       // rows of L+1 complex values
    size_t lds;       // complex values between rows of spec
};

template <int P>
struct Cfg {
    static constexpr int M = 1 << P;
    // threads: one radix-8 butterfly each
    static constexpr int T = (M / 8 < 64) ? 64 : (M / 8 > 1024 ? 1024 : M / 8);
    static constexpr int NB8 = (M / 8 + T - 1) / T;  // radix-8 butterflies per thread
    static constexpr int SLOTS = M + (M >> 4);
};

// Twiddle indices of radix-8 pass on blocks of S points for butterfly t
__device__ __forceinline__ fft::Tw8 tw8_for(const cplx *tw, int M, int S, int t)
{
    const int st = S / 8;
    const int j = t % st;
    return fft::tw8_load(tw, j * (M / S), M);
}

// conv = M * (a * b~) in natural order from a in LDS (natural order):
// decimation-in-frequency passes down to 16-point blocks; per block the last
// forward pass, the product with B and the first inverse pass in registers;
// decimation-in-time passes back up.  The twiddles of each radix-8 pass are
// requested before the barrier that precedes it.
template <int P>
__device__ __forceinline__ void convolve(cplx *x, const cplx *B, const cplx *tw)
{
    constexpr int M = Cfg<P>::M, T = Cfg<P>::T, NB = Cfg<P>::NB8;
    constexpr int F = fft::plan_first(P), N8 = (P == 3) ? 0 : fft::plan_n8(P);
    constexpr int LAST = (P == 3) ? 8 : 16;  // radix of the fused middle pass
    const int t = threadIdx.x;
    constexpr int S8 = (F ? M / F : M);      // block size of the first radix-8 pass

    // forward
    fft::Tw8 nxt[NB];
#pragma unroll
    for (int q = 0; q < NB; q++)
        if (N8 > 0 && t + q * T < M / 8) nxt[q] = tw8_for(tw, M, S8, t + q * T);
    if (F) {
        for (int u = t; u < M / F; u += T) {
            if (F == 2)
                fft::dif_butterfly<2>(x, tw, M, M, u);
            else
                fft::dif_butterfly<4>(x, tw, M, M, u);
        }
        __syncthreads();
    }
    int S = S8;
#pragma unroll
    for (int i = 0; i < N8; i++) {
        fft::Tw8 cur[NB];
#pragma unroll
        for (int q = 0; q < NB; q++) {
            const int u = t + q * T;
            cur[q] = nxt[q];
            if (i + 1 < N8 && u < M / 8) nxt[q] = tw8_for(tw, M, S / 8, u);
        }
#pragma unroll
        for (int q = 0; q < NB; q++) {
            const int u = t + q * T;
            if (u < M / 8) {
                const int st = S / 8, b = u / st, j = u - b * st;
                fft::dif_butterfly8(x, cur[q], S, b, j);
            }
        }
        S /= 8;
        __syncthreads();
    }
    // middle: contiguous blocks of LAST points (the twiddles of the first
    // inverse pass are requested after it: nothing else stays live across it)
    for (int u = t; u < M / LAST; u += T) {
        cplx v[LAST];
#pragma unroll
        for (int r = 0; r < LAST; r++) v[r] = x[fft::slot(u * LAST + r)];
        fft::dft<LAST, -1>(v);
#pragma unroll
        for (int r = 0; r < LAST; r++) v[r] = fft::mul(v[r], B[u * LAST + r]);
        fft::dft<LAST, +1>(v);
#pragma unroll
        for (int r = 0; r < LAST; r++) x[fft::slot(u * LAST + r)] = v[r];
    }
#pragma unroll
    for (int q = 0; q < NB; q++)
        if (N8 > 0 && t + q * T < M / 8) nxt[q] = tw8_for(tw, M, LAST * 8, t + q * T);
    __syncthreads();
    // inverse
    S = LAST * 8;
#pragma unroll
    for (int i = N8 - 1; i >= 0; i--) {
        fft::Tw8 cur[NB];
#pragma unroll
        for (int q = 0; q < NB; q++) {
            const int u = t + q * T;
            cur[q] = nxt[q];
            if (i > 0 && u < M / 8) nxt[q] = tw8_for(tw, M, S * 8, u);
        }
#pragma unroll
        for (int q = 0; q < NB; q++) {
            const int u = t + q * T;
            if (u < M / 8) {
                const int st = S / 8, b = u / st, j = u - b * st;
                fft::dit_butterfly8(x, cur[q], S, b, j);
            }
        }
        S *= 8;
        __syncthreads();
    }
    if (F) {
        for (int u = t; u < M / F; u += T) {
            if (F == 2)
                fft::dit_butterfly<2>(x, tw, M, M, u);
            else
                fft::dit_butterfly<4>(x, tw, M, M, u);
        }
        __syncthreads();
    }
}

// f (2L reals per row) -> spec (L+1 complex per row), F_k = sum_j f_j e^{-2 pi i jk/2L}
template <int P>
__global__ __launch_bounds__(Cfg<P>::T) void r2c_kernel(RowArgs a)
{
    constexpr int M = Cfg<P>::M, T = Cfg<P>::T;
    __shared__ cplx x[Cfg<P>::SLOTS];
    const int L = a.L;
    const size_t row = blockIdx.x;
    const cplx *z = (const cplx *)(a.f + row * a.ldf);
    for (int n = threadIdx.x; n < M; n += T)
        x[fft::slot(n)] = (n < L) ? fft::mul(z[n], a.w[n]) : fft::mk(0.0, 0.0);
    __syncthreads();
    convolve<P>(x, a.B, a.tw);
    // Z_k = w_k conv_k / M; F_k = E_k + e^{-i pi k/L} O_k with
    // E_k = (Z_k + conj Z_{L-k}) / 2, O_k = (Z_k - conj Z_{L-k}) / (2i)
    const double h = 0.5 / (double)M;
    cplx *out = a.spec + row * a.lds;
    for (int k = threadIdx.x; k <= L; k += T) {
        const int kk = (k == L) ? 0 : k, kr = (k == 0) ? 0 : L - k;
        const cplx zk = fft::mul(a.w[kk], x[fft::slot(kk)]);
        const cplx zr = fft::conj(fft::mul(a.w[kr], x[fft::slot(kr)]));
        const cplx e = fft::scale(fft::add(zk, zr), h);
        const cplx o = fft::rot90<-1>(fft::scale(fft::sub(zk, zr), h));
        out[k] = fft::add(e, fft::mul(a.pt[k], o));
    }
}

// spec (L+1 complex per row) -> f (2L reals per row),
// f_j = sum_{m=0}^{2L-1} X_m e^{+2 pi i jm/2L} over the Hermitian extension
// (imaginary parts of X_0 and X_L ignored)
template <int P>
__global__ __launch_bounds__(Cfg<P>::T) void c2r_kernel(RowArgs a)
{
    constexpr int M = Cfg<P>::M, T = Cfg<P>::T;
    __shared__ cplx x[Cfg<P>::SLOTS];
    const int L = a.L;
    const size_t row = blockIdx.x;
    const cplx *X = a.spec + row * a.lds;
    for (int n = threadIdx.x; n < M; n += T) {
        cplx v = fft::mk(0.0, 0.0);
        if (n < L) {
            cplx xk = X[n], xr = X[L - n];
            if (n == 0) {
                xk.y = 0.0;
                xr.y = 0.0;
            }
            // E = (X_n + conj X_{L-n}) / 2, O = (X_n - conj X_{L-n}) e^{i pi n/L} / 2,
            // Z = E + i O; the inverse DFT of Z is conj(DFT(conj Z))
            const cplx e = fft::scale(fft::add(xk, fft::conj(xr)), 0.5);
            const cplx o = fft::mul(fft::scale(fft::sub(xk, fft::conj(xr)), 0.5), fft::conj(a.pt[n]));
            const cplx zz = fft::add(e, fft::rot90<+1>(o));
            v = fft::mul(fft::conj(zz), a.w[n]);
        }
        x[fft::slot(n)] = v;
    }
    __syncthreads();
    convolve<P>(x, a.B, a.tw);
    // z_n = conj(w_n conv_n / M); f_2n = 2 Re z_n, f_2n+1 = 2 Im z_n
    const double s2 = 2.0 / (double)M;
    cplx *f = (cplx *)(a.f + row * a.ldf);
    for (int n = threadIdx.x; n < L; n += T) {
        const cplx y = fft::mul(a.w[n], x[fft::slot(n)]);
        f[n] = fft::mk(s2 * y.x, -s2 * y.y);
    }
}

// Spectrum of the chirp filter b~ (b~_n = conj w_n for |n| < L, cyclic) in the
// order the forward passes of convolve() leave data in
template <int P>
__global__ __launch_bounds__(Cfg<P>::T) void filter_kernel(int L, const cplx *w, const cplx *tw,
                                                           cplx *B)
{
    constexpr int M = Cfg<P>::M, T = Cfg<P>::T;
    constexpr int F = fft::plan_first(P), N8 = (P == 3) ? 0 : fft::plan_n8(P);
    constexpr int LAST = (P == 3) ? 8 : 16;
    __shared__ cplx x[Cfg<P>::SLOTS];
    for (int n = threadIdx.x; n < M; n += T) {
        cplx v = fft::mk(0.0, 0.0);
        if (n < L)
            v = fft::conj(w[n]);
        else if (M - n < L)
            v = fft::conj(w[M - n]);
        x[fft::slot(n)] = v;
    }
    __syncthreads();
    const int t = threadIdx.x;
    if (F) {
        for (int u = t; u < M / F; u += T) {
            if (F == 2)
                fft::dif_butterfly<2>(x, tw, M, M, u);
            else
                fft::dif_butterfly<4>(x, tw, M, M, u);
        }
        __syncthreads();
    }
    int S = F ? M / F : M;
    for (int i = 0; i < N8; i++) {
        for (int u = t; u < M / 8; u += T) fft::dif_butterfly<8>(x, tw, M, S, u);
        S /= 8;
        __syncthreads();
    }
    for (int u = t; u < M / LAST; u += T) {
        cplx v[LAST];
#pragma unroll
        for (int r = 0; r < LAST; r++) v[r] = x[fft::slot(u * LAST + r)];
        fft::dft<LAST, -1>(v);
#pragma unroll
        for (int r = 0; r < LAST; r++) B[u * LAST + r] = v[r];
    }
}

__global__ void chirp_table_kernel(int L, cplx *w, cplx *pt)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n < L) w[n] = fft::chirp(n, L, -1);
    if (n <= L) {
        double s, c;
        sincos(3.14159265358979323846 * (double)n / (double)L, &s, &c);
        // exact at the quarter points
        if (2 * n == L) {
            c = 0.0;
            s = 1.0;
        }
        pt[n] = fft::mk(c, -s);
    }
}

__global__ void twiddle_table_kernel(int M, cplx *tw)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= M) return;
    double s, c;
    sincos(2.0 * 3.14159265358979323846 * (double)q / (double)M, &s, &c);
    if (4 * q == M) {
        c = 0.0;
        s = 1.0;
    } else if (2 * q == M) {
        c = -1.0;
        s = 0.0;
    } else if (4 * q == 3 * M) {
        c = 0.0;
        s = -1.0;
    }
    tw[q] = fft::mk(c, -s);
}

struct Plan {
    int L = 0, P = 0;
    cplx *w = nullptr, *pt = nullptr, *B = nullptr;
    const cplx *tw = nullptr;
};

std::mutex g_fft_mu;
std::map<int, cplx *> g_tw;   // M -> twiddles
std::map<int, Plan> g_plans;  // L -> plan

int log2_ceil(long v)
{
    int p = 0;
    while ((1L << p) < v) p++;
    return p;
}

template <int P>
void launch_rows(bool c2r, const RowArgs &a, int nrow, hipStream_t s)
{
    if (c2r)
        hipLaunchKernelGGL(c2r_kernel<P>, dim3(nrow), dim3(Cfg<P>::T), 0, s, a);
    else
        hipLaunchKernelGGL(r2c_kernel<P>, dim3(nrow), dim3(Cfg<P>::T), 0, s, a);
}

template <int P>
void launch_filter(int L, const cplx *w, const cplx *tw, cplx *B, hipStream_t s)
{
    hipLaunchKernelGGL(filter_kernel<P>, dim3(1), dim3(Cfg<P>::T), 0, s, L, w, tw, B);
}

#define FFT_DISPATCH(P, CALL)                                                                   \
    switch (P) {                                                                                \
    case 3: CALL(3); break;                                                                     \
    case 4: CALL(4); break;                                                                     \
    case 5: CALL(5); break;                                                                     \
    case 6: CALL(6); break;                                                                     \
    case 7: CALL(7); break;                                                                     \
    case 8: CALL(8); break;                                                                     \
    case 9: CALL(9); break;                                                                     \
    case 10: CALL(10); break;                                                                   \
    case 11: CALL(11); break;                                                                   \
    case 12: CALL(12); break;                                                                   \
    default: CALL(13); break;                                                                   \
    }

// Plan of length L (created once per process; the tables stay resident)
int get_plan(int L, hipStream_t s, Plan *out)
{
    std::lock_guard<std::mutex> lk(g_fft_mu);
    auto it = g_plans.find(L);
    if (it != g_plans.end()) {
        *out = it->second;
        return 0;
    }
    Plan p;
    p.L = L;
    p.P = std::max(3, log2_ceil(2L * L - 1));
    const int M = 1 << p.P;
    hipError_t e = hipSuccess;
    cplx *tw = nullptr;
    auto tit = g_tw.find(M);
    if (tit == g_tw.end()) {
        e = hipMalloc((void **)&tw, M * sizeof(cplx));
        if (e != hipSuccess) return (int)e;
        hipLaunchKernelGGL(twiddle_table_kernel, dim3((M + 255) / 256), dim3(256), 0, s, M, tw);
        g_tw[M] = tw;
    } else {
        tw = tit->second;
    }
    p.tw = tw;
    if ((e = hipMalloc((void **)&p.w, L * sizeof(cplx))) != hipSuccess) return (int)e;
    if ((e = hipMalloc((void **)&p.pt, (L + 1) * sizeof(cplx))) != hipSuccess) return (int)e;
    if ((e = hipMalloc((void **)&p.B, M * sizeof(cplx))) != hipSuccess) return (int)e;
    hipLaunchKernelGGL(chirp_table_kernel, dim3((L + 1 + 255) / 256), dim3(256), 0, s, L, p.w, p.pt);
#define FFT_FILTER(PP) launch_filter<PP>(L, p.w, p.tw, p.B, s)
    FFT_DISPATCH(p.P, FFT_FILTER)
#undef FFT_FILTER
    if ((e = hipGetLastError()) != hipSuccess) return (int)e;
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return (int)e;
    g_plans[L] = p;
    *out = p;
    return 0;
}

}  // namespace

bool row_fft_supported(size_t nlon)
{
    return nlon >= 2 && nlon % 2 == 0 && nlon / 2 <= 4096;
}

int launch_row_fft(int c2r, size_t nlon, size_t nrow, double *f, size_t ldf, void *spec, size_t lds,
                   void *stream)
{
    if (!row_fft_supported(nlon)) return (int)hipErrorInvalidValue;
    if (nrow == 0) return 0;
    hipStream_t s = (hipStream_t)stream;
    Plan p;
    const int L = (int)(nlon / 2);
    if (int e = get_plan(L, s, &p)) return e;
    RowArgs a;
    a.L = L;
    a.w = p.w;
    a.pt = p.pt;
    a.B = p.B;
    a.tw = p.tw;
    a.f = f;
    a.ldf = ldf;
    a.spec = (cplx *)spec;
    a.lds = lds;
#define FFT_ROWS(PP) launch_rows<PP>(c2r != 0, a, (int)nrow, s)
    FFT_DISPATCH(p.P, FFT_ROWS)
#undef FFT_ROWS
    return (int)hipGetLastError();
}

// Plans made ahead of a graph capture (plan creation synchronises)
int prepare_row_fft(size_t nlon, void *stream)
{
    if (!row_fft_supported(nlon)) return 0;
    Plan p;
    return get_plan((int)(nlon / 2), (hipStream_t)stream, &p);
}

}  // namespace charm_amd
