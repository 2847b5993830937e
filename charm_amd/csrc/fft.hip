// Longitude transforms of Gauss-Legendre rows (nlon = 2L points per row):
// r2c (analysis) and c2r (synthesis) with hipFFT's unnormalised conventions,
// as one complex DFT of length L per row, one row per work-group, entirely
// in LDS: mixed-radix passes directly when L = 2^a 3^b 5^c 7^d, else Rader's
// cyclic convolution of length L - 1 (L prime, L - 1 smooth) or Bluestein's
// chirp convolution over a smooth length M >= 2L - 1, whichever is cheaper
// (src/shs/shs_grd_fft.c:99-130 and src/sha/sha_point.c:490-553 use FFTW's
// r2c/c2r for the same transforms).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <vector>

#include "fft_core.h"
#include "internal.h"

namespace charm_amd {

using fft::cplx;

namespace {

struct RowArgs {
    fft::MixPlan plan;
    int L;
    const cplx *w;    // [L]   exp(-i pi n^2 / L) (Bluestein)
    const cplx *pt;   // [L+1] exp(-i pi k / L)
    const cplx *B;    // [M]   chirp filter spectrum, digit-reversed (Bluestein)
    const int *rev;   // [L]   position of output k after the passes (direct)
    const int *lg;    // [L]   log_g n (Rader)
    const cplx *tw;   // [M]   exp(-2 pi i q / M)
    double *f;        // rows of 2L reals
    size_t ldf;       // doubles between rows of f
    cplx *spec;       // rows of L+1 complex values
    size_t lds;       // complex values between rows of spec
};

template <int R, bool INV, bool POW2>
__device__ __forceinline__ void run_pass_g(cplx *x, const fft::TwSplit &tw, int M,
                                           const fft::PassGeom &g)
{
    for (int u = threadIdx.x; u < M / R; u += blockDim.x) {
        if (INV)
            fft::dit_butterfly<R, POW2>(x, tw, g, u);
        else
            fft::dif_butterfly<R, POW2>(x, tw, g, u);
    }
}

template <int R, bool INV>
__device__ __forceinline__ void run_pass(cplx *x, const fft::TwSplit &tw, int M, int S)
{
    const fft::PassGeom g = fft::pass_geom(M, S, R);
    if (g.sh >= 0)
        run_pass_g<R, INV, true>(x, tw, M, g);
    else
        run_pass_g<R, INV, false>(x, tw, M, g);
}

template <bool INV>
__device__ __forceinline__ void pass(cplx *x, const fft::TwSplit &tw, int M, int S, int R)
{
    switch (R) {
    case 2: run_pass<2, INV>(x, tw, M, S); break;
    case 3: run_pass<3, INV>(x, tw, M, S); break;
    case 4: run_pass<4, INV>(x, tw, M, S); break;
    case 5: run_pass<5, INV>(x, tw, M, S); break;
    case 7: run_pass<7, INV>(x, tw, M, S); break;
    case 8: run_pass<8, INV>(x, tw, M, S); break;
    default: run_pass<9, INV>(x, tw, M, S); break;
    }
}

template <int R>
__device__ __forceinline__ void run_last(cplx *x, int M)
{
    for (int u = threadIdx.x; u < M / R; u += blockDim.x) fft::last_block<R>(x, u);
}

template <int R>
__device__ __forceinline__ void run_middle(cplx *x, const cplx *B, int M)
{
    for (int u = threadIdx.x; u < M / R; u += blockDim.x) fft::middle_block<R>(x, B, u);
}

// Forward passes 0..np-2 (twiddled), leaving blocks of R[np-1] points
__device__ __forceinline__ void forward_head(cplx *x, const fft::TwSplit &tw, const fft::MixPlan &p)
{
    int S = p.M;
    for (int i = 0; i + 1 < p.np; i++) {
        pass<false>(x, tw, p.M, S, p.R[i]);
        S /= p.R[i];
        __syncthreads();
    }
}

// The full forward DFT (direct plans): output k at position rev[k]
__device__ __forceinline__ void forward_dft(cplx *x, const fft::TwSplit &tw, const fft::MixPlan &p)
{
    if (p.np == 0) return;
    forward_head(x, tw, p);
    switch (p.R[p.np - 1]) {
    case 2: run_last<2>(x, p.M); break;
    case 3: run_last<3>(x, p.M); break;
    case 4: run_last<4>(x, p.M); break;
    case 5: run_last<5>(x, p.M); break;
    case 7: run_last<7>(x, p.M); break;
    case 8: run_last<8>(x, p.M); break;
    case 9: run_last<9>(x, p.M); break;
    default: run_last<16>(x, p.M); break;
    }
    __syncthreads();
}

// conv = M (a * b~) in natural order from a in natural order (Bluestein)
__device__ __forceinline__ void convolve(cplx *x, const cplx *B, const fft::TwSplit &tw,
                                         const fft::MixPlan &p)
{
    forward_head(x, tw, p);
    switch (p.R[p.np - 1]) {
    case 2: run_middle<2>(x, B, p.M); break;
    case 3: run_middle<3>(x, B, p.M); break;
    case 4: run_middle<4>(x, B, p.M); break;
    case 5: run_middle<5>(x, B, p.M); break;
    case 7: run_middle<7>(x, B, p.M); break;
    case 8: run_middle<8>(x, B, p.M); break;
    case 9: run_middle<9>(x, B, p.M); break;
    default: run_middle<16>(x, B, p.M); break;
    }
    __syncthreads();
    // inverse passes np-2..0 on blocks of S_i = M / (R[0] ... R[i-1])
    int S = p.R[p.np - 1];
    for (int i = p.np - 2; i >= 0; i--) {
        S *= p.R[i];
        pass<true>(x, tw, p.M, S, p.R[i]);
        __syncthreads();
    }
}

extern __shared__ cplx lds_rows[];

constexpr int NWSUM = 16;  // per-wave partial sums (Rader), after the tables

// The row's slots, then the twiddle tables, in LDS; the tables are filled
// from the global table (visible after the caller's first barrier)
__device__ __forceinline__ fft::TwSplit lds_twiddles(const cplx *tw, int M)
{
    cplx *hi = lds_rows + fft::slots(M);
    const int nh = fft::tw_hi_count(M);
    cplx *lo = hi + nh;
    for (int i = threadIdx.x; i < nh + fft::TW_LO; i += blockDim.x) {
        if (i < nh)
            hi[i] = tw[i * fft::TW_LO];
        else if (i - nh < M)
            lo[i - nh] = tw[i - nh];
    }
    fft::TwSplit t;
    t.hi = hi;
    t.lo = lo;
    return t;
}

__device__ __forceinline__ cplx* wave_sums(int M)
{
    return lds_rows + fft::slots(M) + fft::tw_hi_count(M) + fft::TW_LO;
}

__device__ __forceinline__ cplx wave_reduce(cplx v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        v.x += __shfl_xor(v.x, o);
        v.y += __shfl_xor(v.y, o);
    }
    return v;
}

// Output k of the row's DFT after the transform, in two steps: row_out_at reads
// what output k needs from the global tables (its LDS slot; Bluestein's
// chirp factor), row_out the LDS (v0 = input 0, used by Rader).
struct OutAt {
    int s;
    cplx w;
};

template <int MODE>
__device__ __forceinline__ OutAt row_out_at(const RowArgs &a, int k)
{
    const int M = a.plan.M;
    OutAt o;
    o.w = fft::mk(0.0, 0.0);
    if (MODE == fft::DFT_DIRECT) {
        o.s = fft::slot(a.rev[k]);
    } else if (MODE == fft::DFT_BLUESTEIN) {
        o.s = fft::slot(k);
        o.w = a.w[k];
    } else {
        const int q = a.lg[k];
        o.s = fft::slot(q == 0 ? 0 : M - q);
    }
    return o;
}

template <int MODE>
__device__ __forceinline__ cplx row_out(const RowArgs &a, const cplx *x, int k, const OutAt &o,
                                        cplx v0)
{
    const int M = a.plan.M;
    if (MODE == fft::DFT_DIRECT) return x[o.s];
    if (MODE == fft::DFT_BLUESTEIN) return fft::scale(fft::mul(o.w, x[o.s]), 1.0 / (double)M);
    // Rader: Z_0 = sum of the inputs; Z_{g^-p} = v0 + conv_p / M
    if (k == 0) {
        const cplx *ws = wave_sums(M);
        cplx s = fft::mk(0.0, 0.0);
        const int nw = (blockDim.x + 63) >> 6;
        for (int i = 0; i < nw; i++) s = fft::add(s, ws[i]);
        return s;
    }
    return fft::axpy(1.0 / (double)M, x[o.s], v0);
}

// The whole-row kernels below read and write a row in batches that cover it
// in one step at the common sizes (L = 2161 over 256 threads): a thread
// issues all global loads of a batch (values, index and twiddle tables)
// before it waits for any, and the LDS twiddle tables are fetched together
// with the row's first batch.  Outputs k and L - k (r2c) and inputs n and
// L - n (c2r) come from the same two values of the row and are made by one
// thread; pt_{L-n} = -conj pt_n holds exactly (chirp_table_kernel).
constexpr int POINT_BATCH = 9;

// Pairs per batch (Bluestein pairs also load two chirp factors)
template <int MODE>
constexpr int pair_batch()
{
    return MODE == fft::DFT_BLUESTEIN ? 3 : 5;
}

__device__ __forceinline__ int tw_entries(int M)
{
    return fft::tw_hi_count(M) + min(M, fft::TW_LO);
}

// Entry i of the LDS twiddle tables (hi: every TW_LO-th power, then lo: the
// first TW_LO powers), which sit one after the other behind the row's slots
__device__ __forceinline__ cplx tw_entry(const cplx *tw, int nh, int i)
{
    return tw[i < nh ? i * fft::TW_LO : i - nh];
}

// This thread's first table entry, loaded before the row's first batch
__device__ __forceinline__ cplx tw_fetch(const cplx *tw, int M)
{
    const int i = threadIdx.x;
    return i < tw_entries(M) ? tw_entry(tw, fft::tw_hi_count(M), i) : fft::mk(0.0, 0.0);
}

// Writes the first entry and loads and writes the rest (threads beyond the
// block size: short rows); visible after the caller's next barrier
__device__ __forceinline__ fft::TwSplit tw_store(const cplx *tw, int M, cplx first)
{
    const int nh = fft::tw_hi_count(M), ne = tw_entries(M);
    cplx *t = lds_rows + fft::slots(M);
    if ((int)threadIdx.x < ne) t[threadIdx.x] = first;
    for (int i = threadIdx.x + blockDim.x; i < ne; i += blockDim.x) t[i] = tw_entry(tw, nh, i);
    fft::TwSplit r;
    r.hi = t;
    r.lo = t + nh;
    return r;
}

// Where input n of the row's DFT goes (Rader: n >= 1 to slot(log_g n)) and
// its factor (Bluestein: the chirp w_n)
struct InAt {
    int s;
    cplx w;
};

template <int MODE>
__device__ __forceinline__ InAt row_in_at(const RowArgs &a, int n)
{
    InAt o;
    o.w = fft::mk(0.0, 0.0);
    if (MODE == fft::DFT_RADER) {
        o.s = fft::slot(a.lg[n]);
    } else {
        o.s = fft::slot(n);
        if (MODE == fft::DFT_BLUESTEIN) o.w = a.w[n];
    }
    return o;
}

template <int MODE>
__device__ __forceinline__ cplx row_in(const InAt &o, cplx v)
{
    return MODE == fft::DFT_BLUESTEIN ? fft::mul(v, o.w) : v;
}

// Zero padding of the convolution (Bluestein) and the sum of the inputs
// (Rader, per wave beside the tables); then the transform
template <int MODE>
__device__ __forceinline__ void row_transform(const RowArgs &a, cplx *x, const fft::TwSplit &tw,
                                              cplx acc)
{
    const int L = a.L, M = a.plan.M;
    if (MODE == fft::DFT_BLUESTEIN)
        for (int n = L + threadIdx.x; n < M; n += blockDim.x) x[fft::slot(n)] = fft::mk(0.0, 0.0);
    if (MODE == fft::DFT_RADER) {
        acc = wave_reduce(acc);
        if ((threadIdx.x & 63) == 0) wave_sums(M)[threadIdx.x >> 6] = acc;
    }
    __syncthreads();
    if (MODE == fft::DFT_DIRECT)
        forward_dft(x, tw, a.plan);
    else
        convolve(x, a.B, tw, a.plan);
}

// f (2L reals per row) -> spec (L+1 complex per row), F_k = sum_j f_j e^{-2 pi i jk/2L}
template <int MODE>
__device__ __forceinline__ void r2c_row(const RowArgs &a, size_t row)
{
    cplx *x = lds_rows;
    const int L = a.L, M = a.plan.M, bd = blockDim.x;
    const cplx *z = (const cplx *)(a.f + row * a.ldf);
    const cplx tw0 = tw_fetch(a.tw, M);
    cplx acc = fft::mk(0.0, 0.0);
    // the first batch outside the loop, so that its loads are issued while the
    // twiddle load is still in flight (a loop header waits for all of them)
    auto gather = [&](int n0) {
        cplx v[POINT_BATCH];
        InAt o[POINT_BATCH];
#pragma unroll
        for (int j = 0; j < POINT_BATCH; j++) {
            const int n = min(n0 + j * bd, L - 1);
            v[j] = z[n];
            o[j] = row_in_at<MODE>(a, n);
        }
#pragma unroll
        for (int j = 0; j < POINT_BATCH; j++) {
            const int n = n0 + j * bd;
            if (n >= L) continue;
            if (MODE == fft::DFT_RADER) {
                acc = fft::add(acc, v[j]);
                if (n >= 1) x[o[j].s] = v[j];
            } else {
                x[o[j].s] = row_in<MODE>(o[j], v[j]);
            }
        }
    };
    gather(threadIdx.x);
    for (int n0 = threadIdx.x + POINT_BATCH * bd; n0 < L; n0 += POINT_BATCH * bd) gather(n0);
    // input 0 (Rader outputs), read once the row's loads are on their way: a
    // scalar load before them makes every thread wait for it first
    const cplx z0 = (MODE == fft::DFT_RADER) ? z[0] : fft::mk(0.0, 0.0);
    row_transform<MODE>(a, x, tw_store(a.tw, M, tw0), acc);
    // Z_k; F_k = E_k + e^{-i pi k/L} O_k with E_k = (Z_k + conj Z_{L-k}) / 2,
    // O_k = (Z_k - conj Z_{L-k}) / (2i); F_{L-k} = conj(E_k - e^{-i pi k/L} O_k)
    // (Z_L = Z_0)
    cplx *out = a.spec + row * a.lds;
    const int nk = (L + 1) / 2;  // pairs (k, L - k), 0 <= k < nk
    constexpr int PAIR_BATCH = pair_batch<MODE>();
    for (int k0 = threadIdx.x; k0 < nk; k0 += PAIR_BATCH * bd) {
        OutAt ok[PAIR_BATCH], orr[PAIR_BATCH];
        cplx pw[PAIR_BATCH];
#pragma unroll
        for (int j = 0; j < PAIR_BATCH; j++) {
            const int k = min(k0 + j * bd, nk - 1), kr = (k == 0) ? 0 : L - k;
            ok[j] = row_out_at<MODE>(a, k);
            orr[j] = row_out_at<MODE>(a, kr);
            pw[j] = a.pt[k];
        }
#pragma unroll
        for (int j = 0; j < PAIR_BATCH; j++) {
            const int k = k0 + j * bd;
            if (k >= nk) continue;
            const int kr = (k == 0) ? 0 : L - k;
            const cplx zk = row_out<MODE>(a, x, k, ok[j], z0);
            const cplx zr = fft::conj(row_out<MODE>(a, x, kr, orr[j], z0));
            const cplx e = fft::scale(fft::add(zk, zr), 0.5);
            const cplx o = fft::rot90<-1>(fft::scale(fft::sub(zk, zr), 0.5));
            const cplx m = fft::mul(pw[j], o);
            out[k] = fft::add(e, m);
            out[L - k] = fft::conj(fft::sub(e, m));
        }
    }
    if (!(L & 1) && threadIdx.x == 0) {
        const int k = L / 2;
        const cplx zk = row_out<MODE>(a, x, k, row_out_at<MODE>(a, k), z0);
        const cplx zr = fft::conj(zk);
        const cplx e = fft::scale(fft::add(zk, zr), 0.5);
        const cplx o = fft::rot90<-1>(fft::scale(fft::sub(zk, zr), 0.5));
        out[k] = fft::add(e, fft::mul(a.pt[k], o));
    }
}

template <int MODE>
__global__ __launch_bounds__(1024) void r2c_kernel(RowArgs a)
{
    r2c_row<MODE>(a, blockIdx.x);
}

// Input n of the inverse transform of a spectrum row: conj Z_n with
// E = (X_n + conj X_{L-n}) / 2, O = (X_n - conj X_{L-n}) e^{i pi n/L} / 2,
// Z = E + i O (imaginary parts of X_0 and X_L ignored); the inverse DFT of
// Z is conj(DFT(conj Z))
__device__ __forceinline__ cplx c2r_input(const RowArgs &a, const cplx *X, int n)
{
    const int L = a.L;
    cplx xk = X[n], xr = X[L - n];
    if (n == 0) {
        xk.y = 0.0;
        xr.y = 0.0;
    }
    const cplx e = fft::scale(fft::add(xk, fft::conj(xr)), 0.5);
    const cplx o = fft::mul(fft::scale(fft::sub(xk, fft::conj(xr)), 0.5), fft::conj(a.pt[n]));
    return fft::conj(fft::add(e, fft::rot90<+1>(o)));
}

// c2r_input(n) and c2r_input(L - n) for 1 <= n, 2n != L, from X_n, X_{L-n}
// and pt_n: with E and O of input n, Z_{L-n} = conj E + i conj O
__device__ __forceinline__ void c2r_pair(cplx xk, cplx xr, cplx p, cplx &in_n, cplx &in_r)
{
    const cplx e = fft::scale(fft::add(xk, fft::conj(xr)), 0.5);
    const cplx o = fft::mul(fft::scale(fft::sub(xk, fft::conj(xr)), 0.5), fft::conj(p));
    in_n = fft::conj(fft::add(e, fft::rot90<+1>(o)));
    in_r = fft::add(e, fft::rot90<-1>(o));
}

// spec (L+1 complex per row) -> f (2L reals per row),
// f_j = sum_{m=0}^{2L-1} X_m e^{+2 pi i jm/2L} over the Hermitian extension
template <int MODE>
__device__ __forceinline__ void c2r_row(const RowArgs &a, size_t row)
{
    cplx *x = lds_rows;
    const int L = a.L, M = a.plan.M, bd = blockDim.x;
    const cplx *X = a.spec + row * a.lds;
    const cplx tw0 = tw_fetch(a.tw, M);
    const int np = (L - 1) / 2;  // pairs (n, L - n), 1 <= n <= np
    cplx acc = fft::mk(0.0, 0.0);
    constexpr int PAIR_BATCH = pair_batch<MODE>();
    // the first batch outside the loop (as in r2c_row)
    auto gather = [&](int j0) {
        cplx xk[PAIR_BATCH], xr[PAIR_BATCH], pw[PAIR_BATCH];
        InAt ia[PAIR_BATCH], ib[PAIR_BATCH];
#pragma unroll
        for (int j = 0; j < PAIR_BATCH; j++) {
            const int n = 1 + min(j0 + j * bd, np - 1);
            xk[j] = X[n];
            xr[j] = X[L - n];
            pw[j] = a.pt[n];
            ia[j] = row_in_at<MODE>(a, n);
            ib[j] = row_in_at<MODE>(a, L - n);
        }
#pragma unroll
        for (int j = 0; j < PAIR_BATCH; j++) {
            if (j0 + j * bd >= np) continue;
            cplx u, v;
            c2r_pair(xk[j], xr[j], pw[j], u, v);
            if (MODE == fft::DFT_RADER) acc = fft::add(acc, fft::add(u, v));
            x[ia[j].s] = row_in<MODE>(ia[j], u);
            x[ib[j].s] = row_in<MODE>(ib[j], v);
        }
    };
    if (np > 0) {
        gather(threadIdx.x);
        for (int j0 = threadIdx.x + PAIR_BATCH * bd; j0 < np; j0 += PAIR_BATCH * bd) gather(j0);
    }
    // input 0, read once the row's loads are on their way (as in r2c_row)
    const cplx v0 = c2r_input(a, X, 0);
    if (threadIdx.x == 0) {
        // input 0 (Rader: only in the sum) and, for even L, input L / 2
        if (MODE == fft::DFT_RADER) {
            acc = fft::add(acc, v0);
        } else {
            const InAt o = row_in_at<MODE>(a, 0);
            x[o.s] = row_in<MODE>(o, v0);
        }
        if (!(L & 1)) {
            const InAt o = row_in_at<MODE>(a, L / 2);
            const cplx vh = c2r_input(a, X, L / 2);
            if (MODE == fft::DFT_RADER) acc = fft::add(acc, vh);
            x[o.s] = row_in<MODE>(o, vh);
        }
    }
    row_transform<MODE>(a, x, tw_store(a.tw, M, tw0), acc);
    // z_n = conj(DFT_n); f_2n = 2 Re z_n, f_2n+1 = 2 Im z_n
    cplx *f = (cplx *)(a.f + row * a.ldf);
    for (int n0 = threadIdx.x; n0 < L; n0 += POINT_BATCH * bd) {
        OutAt o[POINT_BATCH];
#pragma unroll
        for (int j = 0; j < POINT_BATCH; j++) o[j] = row_out_at<MODE>(a, min(n0 + j * bd, L - 1));
#pragma unroll
        for (int j = 0; j < POINT_BATCH; j++) {
            const int n = n0 + j * bd;
            if (n >= L) continue;
            const cplx y = row_out<MODE>(a, x, n, o[j], v0);
            f[n] = fft::mk(2.0 * y.x, -2.0 * y.y);
        }
    }
}

template <int MODE>
__global__ __launch_bounds__(1024) void c2r_kernel(RowArgs a)
{
    c2r_row<MODE>(a, blockIdx.x);
}

// Rows of many transforms of one mode in one launch: block b runs row
// b - row0[j] of job j, the last job with row0[j] <= b (row0 ascending,
// row0[0] = 0).  For batches of small grids, whose rows would otherwise take
// one launch per grid.
template <int MODE, bool C2R>
__global__ __launch_bounds__(1024) void rows_group_kernel(const RowArgs *__restrict__ jobs,
                                                          const unsigned *__restrict__ row0,
                                                          int njob)
{
    const unsigned b = blockIdx.x;
    int lo = 0, hi = njob - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (row0[mid] <= b)
            lo = mid;
        else
            hi = mid - 1;
    }
    const RowArgs &a = jobs[lo];
    if (C2R)
        c2r_row<MODE>(a, b - row0[lo]);
    else
        r2c_row<MODE>(a, b - row0[lo]);
}

// Spectrum of the convolution filter in the order the forward passes of
// convolve() leave data in: src (Rader's b), or the chirp b~ (b~_n = conj w_n
// for |n| < L, cyclic) when src is null
__global__ __launch_bounds__(1024) void filter_kernel(RowArgs a, const cplx *src, cplx *B)
{
    cplx *x = lds_rows;
    const fft::MixPlan &p = a.plan;
    const int L = a.L, M = p.M;
    const fft::TwSplit tw = lds_twiddles(a.tw, M);
    for (int n = threadIdx.x; n < M; n += blockDim.x) {
        cplx v = fft::mk(0.0, 0.0);
        if (src)
            v = src[n];
        else if (n < L)
            v = fft::conj(a.w[n]);
        else if (M - n < L)
            v = fft::conj(a.w[M - n]);
        x[fft::slot(n)] = v;
    }
    __syncthreads();
    forward_dft(x, tw, p);
    for (int e = threadIdx.x; e < M; e += blockDim.x) B[e] = x[fft::slot(e)];
}

__global__ void chirp_table_kernel(int L, cplx *w, cplx *pt)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n < L) w[n] = fft::chirp(n, L, -1);
    if (n <= L) {
        // pt_n from the angle pi min(n, L - n) / L, pt_{L-n} = -conj pt_n
        // exactly (the row kernels make outputs and inputs n, L - n in pairs)
        const bool upper = 2 * n > L;
        double s, c;
        sincos(3.14159265358979323846 * (double)(upper ? L - n : n) / (double)L, &s, &c);
        // exact at the quarter points
        if (2 * n == L) {
            c = 0.0;
            s = 1.0;
        }
        pt[n] = upper ? fft::mk(-c, -s) : fft::mk(c, -s);
    }
}

// tw[q] = exp(-2 pi i q / M), exact at the quarter points
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

// ---------------------------------------------------------------------------
// Rows longer than one LDS transform: L = L1 L2 with a short factor L1 (one
// of the codelet sizes) and L2 <= MMAX with an LDS plan, as the four-step
// algorithm: with n = n1 + L1 n2 and k = k2 + L2 k1,
//   Z_k = sum_n1 W_L1^(n1 k1) [ W_L^(n1 k2) sum_n2 z_(n1 + L1 n2) W_L2^(n2 k2) ].
// Pass A (one work-group per row and n1) gathers the stride-L1 subsequence,
// runs the length-L2 LDS transform and applies the twiddle; pass B (one
// thread per row and k2) runs the L1-point DFTs and the real/complex packing.
// ---------------------------------------------------------------------------
constexpr int SPLIT_R[] = {2, 3, 4, 5, 7, 8, 9, 16, 11, 13};

struct SplitArgs {
    RowArgs sub;       // the length-L2 plan
    int L, L1, L2;
    const cplx *ptf;   // [L+1] exp(-i pi k / L)
    const cplx *twf;   // [L]   exp(-2 pi i j / L)
    const cplx *w1;    // [L1]  exp(-2 pi i q / L1) (short factors without a codelet)
    cplx *Y;           // [row][n1][k2]
    double *f;
    size_t ldf;
    cplx *spec;
    size_t lds;
    size_t nrow;
};

// Input n of the inverse transform (c2r_input for the full row length)
__device__ __forceinline__ cplx c2r_input_full(int L, const cplx *pt, const cplx *X, int n)
{
    cplx xk = X[n], xr = X[L - n];
    if (n == 0) {
        xk.y = 0.0;
        xr.y = 0.0;
    }
    const cplx e = fft::scale(fft::add(xk, fft::conj(xr)), 0.5);
    const cplx o = fft::mul(fft::scale(fft::sub(xk, fft::conj(xr)), 0.5), fft::conj(pt[n]));
    return fft::conj(fft::add(e, fft::rot90<+1>(o)));
}

// Points per thread and batch of the column blocks (a column of L2 = 1543
// points over 320 threads in one batch; 4 where a point of the inverse
// transform also carries the chirp factor, which would spill at 5)
constexpr int COL_BATCH = 5;

template <int MODE, bool C2R>
constexpr int col_batch()
{
    return (C2R && MODE == fft::DFT_BLUESTEIN) ? 4 : COL_BATCH;
}

// Loads the L inputs get(n) of a column's DFT into LDS in the order MODE
// needs, the first batch's loads issued while the twiddle load (tw0) is in
// flight, and runs the transform
template <int MODE, int COL_BATCH, class Get>
__device__ __forceinline__ void column_dft(const RowArgs &a, cplx *x, cplx tw0, Get get)
{
    const int L = a.L, M = a.plan.M, bd = blockDim.x;
    cplx acc = fft::mk(0.0, 0.0);
    auto gather = [&](int n0) {
        cplx v[COL_BATCH];
        InAt o[COL_BATCH];
#pragma unroll
        for (int j = 0; j < COL_BATCH; j++) {
            const int n = min(n0 + j * bd, L - 1);
            v[j] = get(n);
            o[j] = row_in_at<MODE>(a, n);
        }
#pragma unroll
        for (int j = 0; j < COL_BATCH; j++) {
            const int n = n0 + j * bd;
            if (n >= L) continue;
            if (MODE == fft::DFT_RADER) {
                acc = fft::add(acc, v[j]);
                if (n >= 1) x[o[j].s] = v[j];
            } else {
                x[o[j].s] = row_in<MODE>(o[j], v[j]);
            }
        }
    };
    gather(threadIdx.x);
    for (int n0 = threadIdx.x + COL_BATCH * bd; n0 < L; n0 += COL_BATCH * bd) gather(n0);
    row_transform<MODE>(a, x, tw_store(a.tw, M, tw0), acc);
}

template <int MODE, bool C2R>
__global__ __launch_bounds__(1024) void split_a_kernel(SplitArgs a)
{
    cplx *x = lds_rows;
    const int L = a.L, L1 = a.L1, L2 = a.L2;
    // Block b runs on XCD b mod 8.  A group of 8 L1 consecutive blocks holds
    // rows 8g..8g+7, and the L1 column blocks of one row share b mod 8: they
    // gather the same cache lines of their row (stride L1) through one L2,
    // at about the same time.
    const unsigned b = blockIdx.x, span = 8u * (unsigned)L1;
    const size_t row = (size_t)(b / span) * 8 + (b % span) % 8;
    const int n1 = (int)((b % span) / 8);
    if (row >= a.nrow) return;
    const cplx tw0 = tw_fetch(a.sub.tw, a.sub.plan.M);
    const cplx *X = a.spec + row * a.lds;
    const cplx *z = (const cplx *)(a.f + row * a.ldf);
    auto get = [&](int n2) {
        const int n = n1 + L1 * n2;
        return C2R ? c2r_input_full(L, a.ptf, X, n) : z[n];
    };
    column_dft<MODE, col_batch<MODE, C2R>()>(a.sub, x, tw0, get);
    // input 0 (Rader outputs; read after the column's loads)
    const cplx v0 = (MODE == fft::DFT_RADER) ? get(0) : fft::mk(0.0, 0.0);
    cplx *y = a.Y + (row * L1 + n1) * (size_t)L2;
    const int bd = blockDim.x;
    for (int k0 = threadIdx.x; k0 < L2; k0 += COL_BATCH * bd) {
        OutAt o[COL_BATCH];
        cplx w[COL_BATCH];
#pragma unroll
        for (int i = 0; i < COL_BATCH; i++) {
            const int k2 = min(k0 + i * bd, L2 - 1);
            o[i] = row_out_at<MODE>(a.sub, k2);
            w[i] = a.twf[(int)(((long)n1 * k2) % L)];
        }
#pragma unroll
        for (int i = 0; i < COL_BATCH; i++) {
            const int k2 = k0 + i * bd;
            if (k2 < L2) y[k2] = fft::mul(row_out<MODE>(a.sub, x, k2, o[i], v0), w[i]);
        }
    }
}

// The L1-point DFT of column k2 of a row of Y
template <int R1>
__device__ __forceinline__ void split_column(const SplitArgs &a, size_t row, int k2, cplx *v)
{
    const cplx *y = a.Y + row * (size_t)R1 * a.L2 + k2;
#pragma unroll
    for (int n1 = 0; n1 < R1; n1++) v[n1] = y[(size_t)n1 * a.L2];
    if constexpr (R1 == 11 || R1 == 13)
        fft::dft_table<R1>(v, a.w1);
    else
        fft::dft<R1, -1>(v);
}

template <int R1, bool C2R>
__global__ __launch_bounds__(256) void split_b_kernel(SplitArgs a)
{
    const int k2 = blockIdx.y * blockDim.x + threadIdx.x;
    const size_t row = blockIdx.x;
    const int L = a.L, L2 = a.L2;
    if (k2 >= L2) return;
    cplx v[R1];
    split_column<R1>(a, row, k2, v);
    if (C2R) {
        cplx *f = (cplx *)(a.f + row * a.ldf);
#pragma unroll
        for (int k1 = 0; k1 < R1; k1++) f[k2 + L2 * k1] = fft::mk(2.0 * v[k1].x, -2.0 * v[k1].y);
        return;
    }
    // Z_(L-k) for k = k2 + L2 k1 lies in column (L2 - k2) mod L2, at row
    // L1 - 1 - k1 (k2 > 0) or (L1 - k1) mod L1 (k2 = 0)
    cplx w[R1];
    const int k2p = k2 ? L2 - k2 : 0;
    if (k2p == k2) {
#pragma unroll
        for (int i = 0; i < R1; i++) w[i] = v[i];
    } else {
        split_column<R1>(a, row, k2p, w);
    }
    cplx *out = a.spec + row * a.lds;
#pragma unroll
    for (int k1 = 0; k1 < R1; k1++) {
        const int k = k2 + L2 * k1;
        const int k1r = k2 ? R1 - 1 - k1 : (R1 - k1) % R1;
        const cplx zk = v[k1], zr = fft::conj(w[k1r]);
        const cplx e = fft::scale(fft::add(zk, zr), 0.5);
        const cplx o = fft::rot90<-1>(fft::scale(fft::sub(zk, zr), 0.5));
        out[k] = fft::add(e, fft::mul(a.ptf[k], o));
        if (k == 0) out[L] = fft::add(e, fft::mul(a.ptf[L], o));
    }
}

__global__ void full_tables_kernel(int L, cplx *pt, cplx *tw)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n <= L) {
        double s, c;
        sincos(3.14159265358979323846 * (double)n / (double)L, &s, &c);
        if (2 * n == L) {
            c = 0.0;
            s = 1.0;
        }
        pt[n] = fft::mk(c, -s);
    }
    if (n < L) {
        double s, c;
        sincos(2.0 * 3.14159265358979323846 * (double)n / (double)L, &s, &c);
        if (4 * n == L) {
            c = 0.0;
            s = 1.0;
        } else if (2 * n == L) {
            c = -1.0;
            s = 0.0;
        } else if (4 * n == 3 * L) {
            c = 0.0;
            s = -1.0;
        }
        tw[n] = fft::mk(c, -s);
    }
}

struct Plan {
    fft::MixPlan mp;
    int L = 0;
    int threads = 64;
    cplx *w = nullptr, *pt = nullptr, *B = nullptr;
    int *rev = nullptr, *lg = nullptr;
    const cplx *tw = nullptr;
};

std::mutex g_fft_mu;
std::map<int, cplx *> g_tw;   // M -> twiddles
std::map<int, Plan> g_plans;  // L -> plan
bool g_attr_set = false;

size_t lds_bytes(int M)
{
    return (size_t)(fft::slots(M) + fft::tw_hi_count(M) + fft::TW_LO + NWSUM) * sizeof(cplx);
}

// Up to one radix-8 butterfly per thread, in whole waves, and no more waves
// than let every block that fits in LDS stay resident at 16 waves per CU
// (4 per SIMD: the row kernels use up to 128 VGPRs)
int threads_for(const fft::MixPlan &p)
{
    int most = 1;
    for (int i = 0; i < p.np; i++) most = std::max(most, p.M / p.R[i]);
    most = std::min(most, std::max(1, p.M / 8));
    int t = std::min(1024, std::max(64, (most + 63) / 64 * 64));
    const int per_cu = std::max(1, (int)(163840 / lds_bytes(p.M)));
    const int cap = std::max(64, (16 / std::min(per_cu, 16)) * 64);
    return std::min(t, cap);
}

hipError_t set_lds_limits()
{
    if (g_attr_set) return hipSuccess;
    const int bytes = (int)lds_bytes(fft::MMAX);
    const void *fns[] = {(const void *)r2c_kernel<fft::DFT_DIRECT>,
                         (const void *)r2c_kernel<fft::DFT_BLUESTEIN>,
                         (const void *)r2c_kernel<fft::DFT_RADER>,
                         (const void *)c2r_kernel<fft::DFT_DIRECT>,
                         (const void *)c2r_kernel<fft::DFT_BLUESTEIN>,
                         (const void *)c2r_kernel<fft::DFT_RADER>,
                         (const void *)filter_kernel,
                         (const void *)split_a_kernel<fft::DFT_DIRECT, false>,
                         (const void *)split_a_kernel<fft::DFT_BLUESTEIN, false>,
                         (const void *)split_a_kernel<fft::DFT_RADER, false>,
                         (const void *)split_a_kernel<fft::DFT_DIRECT, true>,
                         (const void *)split_a_kernel<fft::DFT_BLUESTEIN, true>,
                         (const void *)split_a_kernel<fft::DFT_RADER, true>,
                         (const void *)rows_group_kernel<fft::DFT_DIRECT, false>,
                         (const void *)rows_group_kernel<fft::DFT_BLUESTEIN, false>,
                         (const void *)rows_group_kernel<fft::DFT_RADER, false>,
                         (const void *)rows_group_kernel<fft::DFT_DIRECT, true>,
                         (const void *)rows_group_kernel<fft::DFT_BLUESTEIN, true>,
                         (const void *)rows_group_kernel<fft::DFT_RADER, true>};
    for (const void *fn : fns) {
        hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
        if (e != hipSuccess) return e;
    }
    g_attr_set = true;
    return hipSuccess;
}

RowArgs row_args(const Plan &p)
{
    RowArgs a{};
    a.plan = p.mp;
    a.L = p.L;
    a.w = p.w;
    a.pt = p.pt;
    a.B = p.B;
    a.rev = p.rev;
    a.lg = p.lg;
    a.tw = p.tw;
    return a;
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
    if (!fft::choose_plan(L, p.mp)) return (int)hipErrorInvalidValue;
    p.threads = threads_for(p.mp);
    const int M = p.mp.M;
    hipError_t e = set_lds_limits();
    if (e != hipSuccess) return (int)e;
    cplx *tw = nullptr;
    auto tit = g_tw.find(M);
    if (tit == g_tw.end()) {
        if ((e = hipMalloc((void **)&tw, M * sizeof(cplx))) != hipSuccess) return (int)e;
        hipLaunchKernelGGL(twiddle_table_kernel, dim3((M + 255) / 256), dim3(256), 0, s, M, tw);
        g_tw[M] = tw;
    } else {
        tw = tit->second;
    }
    p.tw = tw;
    if ((e = hipMalloc((void **)&p.pt, (L + 1) * sizeof(cplx))) != hipSuccess) return (int)e;
    if ((e = hipMalloc((void **)&p.w, L * sizeof(cplx))) != hipSuccess) return (int)e;
    // host tables go up through pageable copies, waited for below
    std::vector<int> itab;
    std::vector<cplx> btab;
    cplx *bsrc = nullptr;
    if (p.mp.mode != fft::DFT_DIRECT) {
        if ((e = hipMalloc((void **)&p.B, M * sizeof(cplx))) != hipSuccess) return (int)e;
    }
    if (p.mp.mode == fft::DFT_DIRECT) {
        itab = fft::digit_reversal(p.mp);
        if ((e = hipMalloc((void **)&p.rev, L * sizeof(int))) != hipSuccess) return (int)e;
        if ((e = hipMemcpyAsync(p.rev, itab.data(), L * sizeof(int), hipMemcpyHostToDevice, s)) !=
            hipSuccess)
            return (int)e;
    } else if (p.mp.mode == fft::DFT_RADER) {
        itab = fft::rader_logs(L, p.mp.g);
        btab = fft::rader_filter(L, p.mp.g);
        if ((e = hipMalloc((void **)&p.lg, L * sizeof(int))) != hipSuccess) return (int)e;
        if ((e = hipMalloc((void **)&bsrc, M * sizeof(cplx))) != hipSuccess) return (int)e;
        if ((e = hipMemcpyAsync(p.lg, itab.data(), L * sizeof(int), hipMemcpyHostToDevice, s)) !=
            hipSuccess)
            return (int)e;
        if ((e = hipMemcpyAsync(bsrc, btab.data(), M * sizeof(cplx), hipMemcpyHostToDevice, s)) !=
            hipSuccess)
            return (int)e;
    }
    hipLaunchKernelGGL(chirp_table_kernel, dim3((L + 1 + 255) / 256), dim3(256), 0, s, L, p.w, p.pt);
    if (p.mp.mode != fft::DFT_DIRECT) {
        RowArgs a = row_args(p);
        hipLaunchKernelGGL(filter_kernel, dim3(1), dim3(p.threads), lds_bytes(M), s, a, bsrc, p.B);
    }
    if ((e = hipGetLastError()) != hipSuccess) return (int)e;
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return (int)e;
    if (bsrc) (void)hipFree(bsrc);
    g_plans[L] = p;
    *out = p;
    return 0;
}

// Whether an LDS plan exists for L; the search over Bluestein lengths costs
// O(L) factorisations, so the answer is kept per L
bool plan_exists(int L)
{
    static std::mutex mu;
    static std::map<int, bool> known;
    std::lock_guard<std::mutex> lk(mu);
    auto it = known.find(L);
    if (it != known.end()) return it->second;
    fft::MixPlan mp;
    const bool ok = fft::choose_plan(L, mp);
    known[L] = ok;
    return ok;
}

// The short factor of a split row length (0: none)
int split_factor(int L)
{
    static std::mutex mu;
    static std::map<int, int> known;
    std::lock_guard<std::mutex> lk(mu);
    auto it = known.find(L);
    if (it != known.end()) return it->second;
    int best = 0;
    for (int r : SPLIT_R) {
        if (L % r || L / r > fft::MMAX) continue;
        fft::MixPlan mp;
        if (fft::choose_plan(L / r, mp)) {
            best = r;
            break;
        }
    }
    known[L] = best;
    return best;
}

struct SplitPlan {
    int L1 = 0;
    cplx *pt = nullptr, *tw = nullptr, *w1 = nullptr;
};
std::map<int, SplitPlan> g_split;

int get_split(int L, hipStream_t s, SplitPlan *out)
{
    {
        std::lock_guard<std::mutex> lk(g_fft_mu);
        auto it = g_split.find(L);
        if (it != g_split.end()) {
            *out = it->second;
            return 0;
        }
    }
    SplitPlan sp;
    sp.L1 = split_factor(L);
    if (!sp.L1) return (int)hipErrorInvalidValue;
    hipError_t e;
    if ((e = hipMalloc((void **)&sp.pt, (L + 1) * sizeof(cplx))) != hipSuccess) return (int)e;
    if ((e = hipMalloc((void **)&sp.tw, L * sizeof(cplx))) != hipSuccess) return (int)e;
    hipLaunchKernelGGL(full_tables_kernel, dim3((L + 1 + 255) / 256), dim3(256), 0, s, L, sp.pt,
                       sp.tw);
    if ((e = hipMalloc((void **)&sp.w1, sp.L1 * sizeof(cplx))) != hipSuccess) return (int)e;
    hipLaunchKernelGGL(twiddle_table_kernel, dim3(1), dim3(64), 0, s, sp.L1, sp.w1);
    if ((e = hipGetLastError()) != hipSuccess) return (int)e;
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return (int)e;
    std::lock_guard<std::mutex> lk(g_fft_mu);
    g_split[L] = sp;
    *out = sp;
    return 0;
}

bool single_row_plan(size_t nlon)
{
    return nlon / 2 <= (size_t)fft::MMAX && plan_exists((int)(nlon / 2));
}

}  // namespace

bool row_fft_supported(size_t nlon)
{
    if (nlon < 2 || nlon % 2 != 0 || nlon / 2 > (size_t)INT32_MAX / 16) return false;
    return single_row_plan(nlon) || split_factor((int)(nlon / 2)) != 0;
}

size_t row_fft_scratch(size_t nlon, size_t nrow)
{
    if (!row_fft_supported(nlon) || single_row_plan(nlon)) return 0;
    return nrow * (nlon / 2) * sizeof(cplx);
}

int launch_row_fft(int c2r, size_t nlon, size_t nrow, double *f, size_t ldf, void *spec, size_t lds,
                   void *scratch, void *stream)
{
    if (!row_fft_supported(nlon)) return (int)hipErrorInvalidValue;
    if (nrow == 0) return 0;
    hipStream_t s = (hipStream_t)stream;
    Plan p;
    const int L = (int)(nlon / 2);
    if (!single_row_plan(nlon)) {
        if (!scratch) return (int)hipErrorInvalidValue;
        SplitPlan sp;
        if (int e = get_split(L, s, &sp)) return e;
        const int L2 = L / sp.L1;
        if (int e = get_plan(L2, s, &p)) return e;
        SplitArgs a{};
        a.sub = row_args(p);
        a.L = L;
        a.L1 = sp.L1;
        a.L2 = L2;
        a.ptf = sp.pt;
        a.twf = sp.tw;
        a.w1 = sp.w1;
        a.Y = (cplx *)scratch;
        a.f = f;
        a.ldf = ldf;
        a.spec = (cplx *)spec;
        a.lds = lds;
        const size_t sh = lds_bytes(p.mp.M);
        a.nrow = nrow;
        // one block per (row, column n1), rows padded to a multiple of 8
        const dim3 ga((unsigned)(((nrow + 7) / 8) * 8 * (size_t)sp.L1)), ba((unsigned)p.threads);
#define SPLIT_A(C2R)                                                                               \
    switch (p.mp.mode) {                                                                           \
    case fft::DFT_DIRECT:                                                                          \
        hipLaunchKernelGGL((split_a_kernel<fft::DFT_DIRECT, C2R>), ga, ba, sh, s, a);              \
        break;                                                                                     \
    case fft::DFT_BLUESTEIN:                                                                       \
        hipLaunchKernelGGL((split_a_kernel<fft::DFT_BLUESTEIN, C2R>), ga, ba, sh, s, a);           \
        break;                                                                                     \
    default: hipLaunchKernelGGL((split_a_kernel<fft::DFT_RADER, C2R>), ga, ba, sh, s, a); break;   \
    }
        if (c2r) {
            SPLIT_A(true)
        } else {
            SPLIT_A(false)
        }
#undef SPLIT_A
        const dim3 gb((unsigned)nrow, (unsigned)((L2 + 255) / 256)), bb(256);
#define SPLIT_B(R1)                                                                                \
    case R1:                                                                                       \
        if (c2r)                                                                                   \
            hipLaunchKernelGGL((split_b_kernel<R1, true>), gb, bb, 0, s, a);                       \
        else                                                                                       \
            hipLaunchKernelGGL((split_b_kernel<R1, false>), gb, bb, 0, s, a);                      \
        break;
        switch (sp.L1) {
            SPLIT_B(2) SPLIT_B(3) SPLIT_B(4) SPLIT_B(5) SPLIT_B(7) SPLIT_B(8) SPLIT_B(9) SPLIT_B(16)
            SPLIT_B(11) SPLIT_B(13)
        default: return (int)hipErrorInvalidValue;
        }
#undef SPLIT_B
        return (int)hipGetLastError();
    }
    if (int e = get_plan(L, s, &p)) return e;
    RowArgs a = row_args(p);
    a.f = f;
    a.ldf = ldf;
    a.spec = (cplx *)spec;
    a.lds = lds;
    const size_t sh = lds_bytes(p.mp.M);
    const dim3 grid((unsigned)nrow), block((unsigned)p.threads);
#define ROW_LAUNCH(KERNEL)                                                                         \
    switch (p.mp.mode) {                                                                           \
    case fft::DFT_DIRECT:                                                                          \
        hipLaunchKernelGGL(KERNEL<fft::DFT_DIRECT>, grid, block, sh, s, a);                        \
        break;                                                                                     \
    case fft::DFT_BLUESTEIN:                                                                       \
        hipLaunchKernelGGL(KERNEL<fft::DFT_BLUESTEIN>, grid, block, sh, s, a);                     \
        break;                                                                                     \
    default: hipLaunchKernelGGL(KERNEL<fft::DFT_RADER>, grid, block, sh, s, a); break;             \
    }
    if (c2r) {
        ROW_LAUNCH(c2r_kernel)
    } else {
        ROW_LAUNCH(r2c_kernel)
    }
#undef ROW_LAUNCH
    return (int)hipGetLastError();
}

bool row_group_supported(size_t nlon)
{
    return nlon >= 2 && nlon % 2 == 0 && nlon / 2 <= (size_t)INT32_MAX / 16 &&
           single_row_plan(nlon);
}

// Device descriptors of grouped row transforms: one part per transform mode
// (one launch each).  Plans are made here (this synchronises the stream).
int row_group_build(const RowJob *jobs, int njob, void *stream, std::vector<RowGroupPart> *parts,
                    std::vector<void *> *owned)
{
    hipStream_t s = (hipStream_t)stream;
    std::vector<RowArgs> args[3];
    std::vector<unsigned> row0[3];
    unsigned nrow[3] = {0, 0, 0};
    int threads[3] = {64, 64, 64};
    size_t lds[3] = {0, 0, 0};
    for (int j = 0; j < njob; j++) {
        const RowJob &q = jobs[j];
        if (!row_group_supported(q.nlon) || q.nrow == 0 || q.nrow > (size_t)INT32_MAX)
            return (int)hipErrorInvalidValue;
        Plan p;
        if (int e = get_plan((int)(q.nlon / 2), s, &p)) return e;
        RowArgs a = row_args(p);
        a.f = q.f;
        a.ldf = q.ldf;
        a.spec = (cplx *)q.spec;
        a.lds = q.lds;
        const int md = p.mp.mode;
        if ((size_t)nrow[md] + q.nrow > (size_t)INT32_MAX) return (int)hipErrorInvalidValue;
        args[md].push_back(a);
        row0[md].push_back(nrow[md]);
        nrow[md] += (unsigned)q.nrow;
        threads[md] = std::max(threads[md], p.threads);
        lds[md] = std::max(lds[md], lds_bytes(p.mp.M));
    }
    if (hipError_t e = set_lds_limits()) return (int)e;
    for (int md = 0; md < 3; md++) {
        if (args[md].empty()) continue;
        RowGroupPart part;
        part.mode = md;
        part.njob = (int)args[md].size();
        part.nrow = nrow[md];
        part.threads = threads[md];
        part.lds = lds[md];
        void *da = nullptr, *dr = nullptr;
        hipError_t e = hipMalloc(&da, args[md].size() * sizeof(RowArgs));
        if (e != hipSuccess) return (int)e;
        owned->push_back(da);
        if ((e = hipMalloc(&dr, row0[md].size() * sizeof(unsigned))) != hipSuccess) return (int)e;
        owned->push_back(dr);
        if ((e = hipMemcpyAsync(da, args[md].data(), args[md].size() * sizeof(RowArgs),
                                hipMemcpyHostToDevice, s)) != hipSuccess ||
            (e = hipMemcpyAsync(dr, row0[md].data(), row0[md].size() * sizeof(unsigned),
                                hipMemcpyHostToDevice, s)) != hipSuccess ||
            (e = hipStreamSynchronize(s)) != hipSuccess)
            return (int)e;
        part.args = da;
        part.row0 = (const unsigned *)dr;
        parts->push_back(part);
    }
    return 0;
}

int row_group_launch(int c2r, const std::vector<RowGroupPart> &parts, void *stream)
{
    hipStream_t s = (hipStream_t)stream;
    for (const RowGroupPart &p : parts) {
        const RowArgs *a = (const RowArgs *)p.args;
        const dim3 grid(p.nrow), block((unsigned)p.threads);
#define GROUP_LAUNCH(MODE)                                                                             if (c2r)                                                                                               hipLaunchKernelGGL((rows_group_kernel<MODE, true>), grid, block, p.lds, s, a, p.row0,                                 p.njob);                                                                    else                                                                                                   hipLaunchKernelGGL((rows_group_kernel<MODE, false>), grid, block, p.lds, s, a, p.row0,                                p.njob);
        switch (p.mode) {
        case fft::DFT_DIRECT: GROUP_LAUNCH(fft::DFT_DIRECT) break;
        case fft::DFT_BLUESTEIN: GROUP_LAUNCH(fft::DFT_BLUESTEIN) break;
        default: GROUP_LAUNCH(fft::DFT_RADER) break;
        }
#undef GROUP_LAUNCH
    }
    return (int)hipGetLastError();
}

// Plans made ahead of a graph capture (plan creation synchronises)
int prepare_row_fft(size_t nlon, void *stream)
{
    if (!row_fft_supported(nlon)) return 0;
    Plan p;
    const int L = (int)(nlon / 2);
    if (single_row_plan(nlon)) return get_plan(L, (hipStream_t)stream, &p);
    SplitPlan sp;
    if (int e = get_split(L, (hipStream_t)stream, &sp)) return e;
    return get_plan(L / sp.L1, (hipStream_t)stream, &p);
}

}  // namespace charm_amd
