// Host build of the longitude-transform building blocks (charm_amd/csrc/
// fft_core.h), driven serially in the same pass order as the HIP kernels, so
// that the CPU test suite checks the butterflies, twiddles, pass plan and
// the real/complex packing against numpy.fft.
#include <cstddef>
#include <vector>

#include "fft_core.h"

using namespace charm_amd::fft;

namespace {

int log2_ceil(long v)
{
    int p = 0;
    while ((1L << p) < v) p++;
    return p;
}

struct HostPlan {
    int L, P, M;
    std::vector<cplx> w, pt, B, tw;
};

void dif(std::vector<cplx> &x, const HostPlan &p)
{
    int S = p.M;
    for (int i = 0; i < pass_count(p.P); i++) {
        const int R = pass_radix(p.P, i);
        for (int t = 0; t < p.M / R; t++) {
            if (R == 16) dif_butterfly<16>(x.data(), p.tw.data(), p.M, S, t);
            if (R == 8) dif_butterfly<8>(x.data(), p.tw.data(), p.M, S, t);
            if (R == 4) dif_butterfly<4>(x.data(), p.tw.data(), p.M, S, t);
            if (R == 2) dif_butterfly<2>(x.data(), p.tw.data(), p.M, S, t);
        }
        S /= R;
    }
}

void dit(std::vector<cplx> &x, const HostPlan &p)
{
    for (int i = pass_count(p.P) - 1; i >= 0; i--) {
        int S = p.M;
        for (int q = 0; q < i; q++) S /= pass_radix(p.P, q);
        const int R = pass_radix(p.P, i);
        for (int t = 0; t < p.M / R; t++) {
            if (R == 16) dit_butterfly<16>(x.data(), p.tw.data(), p.M, S, t);
            if (R == 8) dit_butterfly<8>(x.data(), p.tw.data(), p.M, S, t);
            if (R == 4) dit_butterfly<4>(x.data(), p.tw.data(), p.M, S, t);
            if (R == 2) dit_butterfly<2>(x.data(), p.tw.data(), p.M, S, t);
        }
    }
}

HostPlan make_plan(int L)
{
    HostPlan p;
    p.L = L;
    p.P = log2_ceil(2L * L - 1);
    if (p.P < 3) p.P = 3;
    p.M = 1 << p.P;
    p.tw.resize(p.M);
    for (int q = 0; q < p.M; q++) {
        const double a = 2.0 * 3.14159265358979323846 * q / p.M;
        p.tw[q] = mk(std::cos(a), -std::sin(a));
    }
    p.w.resize(L);
    p.pt.resize(L + 1);
    for (int n = 0; n < L; n++) p.w[n] = chirp(n, L, -1);
    for (int n = 0; n <= L; n++) {
        const double a = 3.14159265358979323846 * n / L;
        p.pt[n] = mk(std::cos(a), -std::sin(a));
    }
    std::vector<cplx> x(slots(p.M), mk(0, 0));
    for (int n = 0; n < p.M; n++) {
        cplx v = mk(0, 0);
        if (n < L)
            v = conj(p.w[n]);
        else if (p.M - n < L)
            v = conj(p.w[p.M - n]);
        x[slot(n)] = v;
    }
    dif(x, p);
    p.B.resize(p.M);
    for (int e = 0; e < p.M; e++) p.B[e] = x[slot(e)];
    return p;
}

void convolve(std::vector<cplx> &x, const HostPlan &p)
{
    dif(x, p);
    for (int e = 0; e < p.M; e++) x[slot(e)] = mul(x[slot(e)], p.B[e]);
    dit(x, p);
}

}  // namespace

extern "C" {

// f: nrow x 2L reals -> spec: nrow x (L+1) complex (interleaved)
void host_r2c(int L, int nrow, const double *f, double *spec)
{
    HostPlan p = make_plan(L);
    for (int row = 0; row < nrow; row++) {
        std::vector<cplx> x(slots(p.M), mk(0, 0));
        const double *z = f + (size_t)row * 2 * L;
        for (int n = 0; n < L; n++) x[slot(n)] = mul(mk(z[2 * n], z[2 * n + 1]), p.w[n]);
        convolve(x, p);
        const double h = 0.5 / p.M;
        for (int k = 0; k <= L; k++) {
            const int kk = (k == L) ? 0 : k, kr = (k == 0) ? 0 : L - k;
            const cplx zk = mul(p.w[kk], x[slot(kk)]);
            const cplx zr = conj(mul(p.w[kr], x[slot(kr)]));
            const cplx e = scale(add(zk, zr), h);
            const cplx o = rot90<-1>(scale(sub(zk, zr), h));
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
    for (int row = 0; row < nrow; row++) {
        const double *X = spec + (size_t)row * (L + 1) * 2;
        std::vector<cplx> x(slots(p.M), mk(0, 0));
        for (int n = 0; n < L; n++) {
            cplx xk = mk(X[2 * n], X[2 * n + 1]), xr = mk(X[2 * (L - n)], X[2 * (L - n) + 1]);
            if (n == 0) {
                xk.y = 0.0;
                xr.y = 0.0;
            }
            const cplx e = scale(add(xk, conj(xr)), 0.5);
            const cplx o = mul(scale(sub(xk, conj(xr)), 0.5), conj(p.pt[n]));
            const cplx zz = add(e, rot90<+1>(o));
            x[slot(n)] = mul(conj(zz), p.w[n]);
        }
        convolve(x, p);
        const double s2 = 2.0 / p.M;
        for (int n = 0; n < L; n++) {
            const cplx y = mul(p.w[n], x[slot(n)]);
            f[(size_t)row * 2 * L + 2 * n] = s2 * y.x;
            f[(size_t)row * 2 * L + 2 * n + 1] = -s2 * y.y;
        }
    }
}

}
