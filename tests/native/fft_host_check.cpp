// Host build of the longitude-transform building blocks (charm_amd/csrc/
// fft_core.h), driven serially with the same plans and in the same pass order
// as the HIP kernels, so that the CPU test suite checks the butterflies,
// twiddles, plan choice, digit reversal and the real/complex packing against
// numpy.fft.
#include <cstddef>
#include <vector>

#include "fft_core.h"

using namespace charm_amd::fft;

namespace {

struct HostPlan {
    MixPlan mp;
    bool ok = false;
    int L = 0;
    std::vector<cplx> w, pt, B, hi, lo;
    std::vector<int> rev, lg;
    TwSplit tw() const
    {
        TwSplit t;
        t.hi = hi.data();
        t.lo = lo.data();
        return t;
    }
};

void pass(std::vector<cplx> &x, const HostPlan &p, int S, int R, bool inv)
{
    const int M = p.mp.M;
    const TwSplit tw = p.tw();
    const PassGeom g = pass_geom(M, S, R);
    for (int u = 0; u < M / R; u++) {
        cplx *d = x.data();
#define HOST_PASS(RR)                                                                              \
    if (R == RR) {                                                                                 \
        if (inv) {                                                                                 \
            if (g.sh >= 0)                                                                         \
                dit_butterfly<RR, true>(d, tw, g, u);                                              \
            else                                                                                   \
                dit_butterfly<RR, false>(d, tw, g, u);                                             \
        } else {                                                                                   \
            if (g.sh >= 0)                                                                         \
                dif_butterfly<RR, true>(d, tw, g, u);                                              \
            else                                                                                   \
                dif_butterfly<RR, false>(d, tw, g, u);                                             \
        }                                                                                          \
    }
        HOST_PASS(2) HOST_PASS(3) HOST_PASS(4) HOST_PASS(5) HOST_PASS(7) HOST_PASS(8) HOST_PASS(9)
#undef HOST_PASS
    }
}

void forward_head(std::vector<cplx> &x, const HostPlan &p)
{
    int S = p.mp.M;
    for (int i = 0; i + 1 < p.mp.np; i++) {
        pass(x, p, S, p.mp.R[i], false);
        S /= p.mp.R[i];
    }
}

void forward_dft(std::vector<cplx> &x, const HostPlan &p)
{
    if (p.mp.np == 0) return;
    forward_head(x, p);
    const int R = p.mp.R[p.mp.np - 1];
    for (int u = 0; u < p.mp.M / R; u++) {
#define HOST_LAST(RR)                                                                              \
    if (R == RR) last_block<RR>(x.data(), u);
        HOST_LAST(2) HOST_LAST(3) HOST_LAST(4) HOST_LAST(5) HOST_LAST(7) HOST_LAST(8) HOST_LAST(9)
        HOST_LAST(16)
#undef HOST_LAST
    }
}

void convolve(std::vector<cplx> &x, const HostPlan &p)
{
    forward_head(x, p);
    const int R = p.mp.R[p.mp.np - 1];
    for (int u = 0; u < p.mp.M / R; u++) {
#define HOST_MID(RR)                                                                               \
    if (R == RR) middle_block<RR>(x.data(), p.B.data(), u);
        HOST_MID(2) HOST_MID(3) HOST_MID(4) HOST_MID(5) HOST_MID(7) HOST_MID(8) HOST_MID(9)
        HOST_MID(16)
#undef HOST_MID
    }
    int S = R;
    for (int i = p.mp.np - 2; i >= 0; i--) {
        S *= p.mp.R[i];
        pass(x, p, S, p.mp.R[i], true);
    }
}

HostPlan make_plan(int L)
{
    HostPlan p;
    p.L = L;
    if (!choose_plan(L, p.mp)) return p;
    p.ok = true;
    const int M = p.mp.M;
    std::vector<cplx> tw(M);
    for (int q = 0; q < M; q++) {
        const double a = 2.0 * 3.14159265358979323846 * q / M;
        tw[q] = mk(std::cos(a), -std::sin(a));
    }
    for (int i = 0; i < tw_hi_count(M); i++) p.hi.push_back(tw[i * TW_LO]);
    for (int i = 0; i < TW_LO; i++) p.lo.push_back(i < M ? tw[i] : mk(0, 0));
    p.w.resize(L);
    p.pt.resize(L + 1);
    for (int n = 0; n < L; n++) p.w[n] = chirp(n, L, -1);
    for (int n = 0; n <= L; n++) {
        const double a = 3.14159265358979323846 * n / L;
        p.pt[n] = mk(std::cos(a), -std::sin(a));
    }
    if (p.mp.mode == DFT_DIRECT) {
        p.rev = digit_reversal(p.mp);
        return p;
    }
    std::vector<cplx> x(slots(M), mk(0, 0));
    if (p.mp.mode == DFT_RADER) {
        p.lg = rader_logs(L, p.mp.g);
        const std::vector<cplx> b = rader_filter(L, p.mp.g);
        for (int n = 0; n < M; n++) x[slot(n)] = b[n];
    } else {
        for (int n = 0; n < M; n++) {
            cplx v = mk(0, 0);
            if (n < L)
                v = conj(p.w[n]);
            else if (M - n < L)
                v = conj(p.w[M - n]);
            x[slot(n)] = v;
        }
    }
    forward_dft(x, p);
    p.B.resize(M);
    for (int e = 0; e < M; e++) p.B[e] = x[slot(e)];
    return p;
}

// Z_k = sum_n z_n exp(-2 pi i nk / L), k = 0..L-1, through the plan, with the
// kernels' input placement and output reads
std::vector<cplx> row_dft(const HostPlan &p, const std::vector<cplx> &z)
{
    const int L = p.L, M = p.mp.M;
    std::vector<cplx> x(slots(M), mk(0, 0)), Z(L);
    cplx sum = mk(0, 0);
    if (p.mp.mode == DFT_RADER) {
        for (int n = 0; n < L; n++) {
            sum = add(sum, z[n]);
            if (n >= 1) x[slot(p.lg[n])] = z[n];
        }
        convolve(x, p);
        Z[0] = sum;
        for (int k = 1; k < L; k++) {
            const int q = p.lg[k];
            Z[k] = axpy(1.0 / M, x[slot(q == 0 ? 0 : M - q)], z[0]);
        }
    } else if (p.mp.mode == DFT_BLUESTEIN) {
        for (int n = 0; n < L; n++) x[slot(n)] = mul(z[n], p.w[n]);
        convolve(x, p);
        for (int k = 0; k < L; k++) Z[k] = scale(mul(p.w[k], x[slot(k)]), 1.0 / M);
    } else {
        for (int n = 0; n < L; n++) x[slot(n)] = z[n];
        forward_dft(x, p);
        for (int k = 0; k < L; k++) Z[k] = x[slot(p.rev[k])];
    }
    return Z;
}

// Short factor of a row too long for one LDS transform, in the kernels'
// order of preference (0: none)
int split_factor(int L)
{
    static const int rs[] = {2, 3, 4, 5, 7, 8, 9, 16};
    for (int r : rs) {
        if (L % r || L / r > MMAX) continue;
        MixPlan mp;
        if (choose_plan(L / r, mp)) return r;
    }
    return 0;
}

cplx w_full(long j, int L)
{
    const double a = 2.0 * 3.14159265358979323846 * (double)j / (double)L;
    return mk(std::cos(a), -std::sin(a));
}

// Z = DFT_L(z) by the two passes of the split kernels: pass A per n1 (the
// stride-L1 subsequence through the length-L2 plan, twiddle W_L^(n1 k2)),
// pass B per k2 (the L1-point DFT, output k = k2 + L2 k1)
std::vector<cplx> split_dft(int L, int L1, const std::vector<cplx> &z)
{
    const int L2 = L / L1;
    const HostPlan p = make_plan(L2);
    std::vector<cplx> Y((size_t)L1 * L2), Z(L);
    for (int n1 = 0; n1 < L1; n1++) {
        std::vector<cplx> sub(L2);
        for (int n2 = 0; n2 < L2; n2++) sub[n2] = z[n1 + L1 * n2];
        const std::vector<cplx> S = row_dft(p, sub);
        for (int k2 = 0; k2 < L2; k2++) Y[(size_t)n1 * L2 + k2] = mul(S[k2], w_full((long)n1 * k2 % L, L));
    }
    for (int k2 = 0; k2 < L2; k2++)
        for (int k1 = 0; k1 < L1; k1++) {
            cplx acc = mk(0, 0);
            for (int n1 = 0; n1 < L1; n1++)
                acc = add(acc, mul(Y[(size_t)n1 * L2 + k2], w_full((long)n1 * k1 * L2 % L, L)));
            Z[k2 + L2 * k1] = acc;
        }
    return Z;
}

}  // namespace

extern "C" {

// the short factor the split kernels use for a row of L complex points
int host_split_factor(int L)
{
    return split_factor(L);
}

// r2c / c2r of rows of 2L reals through split_dft (L with a short factor)
void host_r2c_split(int L, int nrow, const double *f, double *spec)
{
    const int L1 = split_factor(L);
    if (!L1) return;
    const double PI = 3.14159265358979323846;
    for (int row = 0; row < nrow; row++) {
        const double *zr0 = f + (size_t)row * 2 * L;
        std::vector<cplx> z(L);
        for (int n = 0; n < L; n++) z[n] = mk(zr0[2 * n], zr0[2 * n + 1]);
        const std::vector<cplx> Z = split_dft(L, L1, z);
        for (int k = 0; k <= L; k++) {
            const int kk = (k == L) ? 0 : k, kr = (k == 0) ? 0 : L - k;
            const cplx zk = Z[kk], zr = conj(Z[kr]);
            const cplx e = scale(add(zk, zr), 0.5);
            const cplx o = rot90<-1>(scale(sub(zk, zr), 0.5));
            const cplx pt = mk(std::cos(PI * k / L), -std::sin(PI * k / L));
            const cplx F = add(e, mul(pt, o));
            spec[((size_t)row * (L + 1) + k) * 2] = F.x;
            spec[((size_t)row * (L + 1) + k) * 2 + 1] = F.y;
        }
    }
}

void host_c2r_split(int L, int nrow, const double *spec, double *f)
{
    const int L1 = split_factor(L);
    if (!L1) return;
    const double PI = 3.14159265358979323846;
    for (int row = 0; row < nrow; row++) {
        const double *X = spec + (size_t)row * (L + 1) * 2;
        std::vector<cplx> v(L);
        for (int n = 0; n < L; n++) {
            cplx xk = mk(X[2 * n], X[2 * n + 1]), xr = mk(X[2 * (L - n)], X[2 * (L - n) + 1]);
            if (n == 0) {
                xk.y = 0.0;
                xr.y = 0.0;
            }
            const cplx pt = mk(std::cos(PI * n / L), -std::sin(PI * n / L));
            const cplx e = scale(add(xk, conj(xr)), 0.5);
            const cplx o = mul(scale(sub(xk, conj(xr)), 0.5), conj(pt));
            v[n] = conj(add(e, rot90<+1>(o)));
        }
        const std::vector<cplx> Y = split_dft(L, L1, v);
        for (int n = 0; n < L; n++) {
            f[(size_t)row * 2 * L + 2 * n] = 2.0 * Y[n].x;
            f[(size_t)row * 2 * L + 2 * n + 1] = -2.0 * Y[n].y;
        }
    }
}

// plan of length L: M, pass count, radices, mode; returns 0 if none
int host_plan(int L, int *M, int *np, int *R, int *mode)
{
    MixPlan mp;
    if (!choose_plan(L, mp)) return 0;
    *M = mp.M;
    *np = mp.np;
    for (int i = 0; i < mp.np; i++) R[i] = mp.R[i];
    *mode = mp.mode;
    return 1;
}

// f: nrow x 2L reals -> spec: nrow x (L+1) complex (interleaved)
void host_r2c(int L, int nrow, const double *f, double *spec)
{
    HostPlan p = make_plan(L);
    if (!p.ok) return;
    for (int row = 0; row < nrow; row++) {
        const double *zr0 = f + (size_t)row * 2 * L;
        std::vector<cplx> z(L);
        for (int n = 0; n < L; n++) z[n] = mk(zr0[2 * n], zr0[2 * n + 1]);
        const std::vector<cplx> Z = row_dft(p, z);
        for (int k = 0; k <= L; k++) {
            const int kk = (k == L) ? 0 : k, kr = (k == 0) ? 0 : L - k;
            const cplx zk = Z[kk], zr = conj(Z[kr]);
            const cplx e = scale(add(zk, zr), 0.5);
            const cplx o = rot90<-1>(scale(sub(zk, zr), 0.5));
            const cplx F = add(e, mul(p.pt[k], o));
            spec[((size_t)row * (L + 1) + k) * 2] = F.x;
            spec[((size_t)row * (L + 1) + k) * 2 + 1] = F.y;
        }
    }
}

// spec: nrow x (L+1) complex -> f: nrow x 2L reals
void host_c2r(int L, int nrow, const double *spec, double *f)
{
    HostPlan p = make_plan(L);
    if (!p.ok) return;
    for (int row = 0; row < nrow; row++) {
        const double *X = spec + (size_t)row * (L + 1) * 2;
        std::vector<cplx> v(L);
        for (int n = 0; n < L; n++) {
            cplx xk = mk(X[2 * n], X[2 * n + 1]), xr = mk(X[2 * (L - n)], X[2 * (L - n) + 1]);
            if (n == 0) {
                xk.y = 0.0;
                xr.y = 0.0;
            }
            const cplx e = scale(add(xk, conj(xr)), 0.5);
            const cplx o = mul(scale(sub(xk, conj(xr)), 0.5), conj(p.pt[n]));
            v[n] = conj(add(e, rot90<+1>(o)));
        }
        const std::vector<cplx> Y = row_dft(p, v);
        for (int n = 0; n < L; n++) {
            f[(size_t)row * 2 * L + 2 * n] = 2.0 * Y[n].x;
            f[(size_t)row * 2 * L + 2 * n + 1] = -2.0 * Y[n].y;
        }
    }
}

}
