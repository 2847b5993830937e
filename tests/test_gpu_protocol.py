"""CHarm's own accuracy protocol and its per-degree metric.

* Protocol (bench/bench.c:192-300, docs/src/benchmarks.rst:36-76): C_nm, S_nm
  from U[-1, 1), mu = R = r = 1, synthesis on the Gauss-Legendre grid of
  degree N, analysis back to N; eps_max = max |C - C_ref| and
  eps_rms = sqrt(2 sum (C - C_ref)^2 / ((N + 1)(N + 2))) over C and S.  The
  published values (double, read off CHarm's plot, about +-15%) are
  8.5e-13 / 5.3e-14 at N = 1000, 5.3e-12 / 1.65e-13 at N = 2500 and
  4.2e-11 / 8.5e-13 at N = 10000, for coefficients drawn by rand().
* Per-degree differences (src/shc/shc_dda.c:16-36, shc_ddv):
  dda_n = sqrt(sum_m (dC_nm^2 + dS_nm^2)), compared degree by degree with the
  oracle's (CHarm's algorithm in double) on the same input.
"""
import os

import numpy as np
import pytest

from charm_amd.device import DevArray, DeviceGrid, Stream
from charm_amd.workloads import GM, R_EARTH, degree_order, kaula_coeffs, uniform_coeffs
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

# CHarm's published values, interpolated log-log in N between the plotted
# degrees (N = 2160: between 1000 and 2500; N = 10800: extrapolated from
# 10000 with the 2500 -> 10000 slope)
PUBLISHED = {1000: (8.5e-13, 5.3e-14), 2160: (4.0e-12, 1.38e-13), 10800: (4.9e-11, 9.1e-13)}


def dda(C1, S1, C2, S2, nmax):
    n, m = degree_order(nmax)
    d = (C1 - C2) ** 2 + (S1 - S2) ** 2
    return np.sqrt(np.bincount(n, weights=d, minlength=nmax + 1))


def roundtrip_dev(N, C, S, mu, R):
    grid = DeviceGrid(N, R)
    st = Stream()
    cs = DevArray.from_numpy(np.stack([C, S], axis=1))
    f = DevArray((grid.nlat, grid.nlon))
    grid.shs(cs, N, mu, R, f, st)
    out = DevArray(cs.shape)
    grid.sha(f, N, mu, R, out, st)
    st.synchronize()
    o = out.numpy()
    return o[:, 0], o[:, 1], grid


def eps(Co, So, C, S, N):
    d = np.concatenate([Co - C, So - S])
    return float(np.abs(d).max()), float(np.sqrt(2.0 * np.sum(d * d) / ((N + 1) * (N + 2))))


@pytest.mark.parametrize("N", [1000, 2160, 10800])
def test_charm_protocol_roundtrip(N):
    C, S = uniform_coeffs(N, N)
    Co, So, _ = roundtrip_dev(N, C, S, 1.0, 1.0)
    e_max, e_rms = eps(Co, So, C, S, N)
    p_max, p_rms = PUBLISHED[N]
    print("protocol N=%d: eps_max %.3e eps_rms %.3e (CHarm published %.2e / %.2e)"
          % (N, e_max, e_rms, p_max, p_rms))
    assert e_max <= 2.0 * p_max and e_rms <= 2.0 * p_rms


@pytest.mark.parametrize("N", [1000, 2160])
def test_protocol_equals_charm_algorithm_in_double(N):
    """The protocol's errors are those of CHarm's algorithm in double: the
    CPU restatement of CHarm's GL transforms (oracle/cpu_gl.c: folded ring
    pairs, CHarm's recurrences and X-numbers, FFT rows) on the same
    coefficients gives eps_max and eps_rms within 1.3x of the GPU's (measured:
    3.39e-12 / 1.324e-13 against 3.45e-12 / 1.318e-13 at N = 2160; at N =
    10800, 7.35e-11 / 8.60e-13 against 7.60e-11 / 8.65e-13)."""
    from oracle.cpu_gl import CpuGL

    C, S = uniform_coeffs(N, N)
    Co, So, grid = roundtrip_dev(N, C, S, 1.0, 1.0)
    g_max, g_rms = eps(Co, So, C, S, N)
    p = grid.pnt
    cg = CpuGL(max(1, min(16, len(os.sched_getaffinity(0)))))
    f = cg.shs(N, C, S, 1.0, 1.0, 1.0, p.lat, len(p.lon))
    Cc, Sc = cg.sha(N, f, p.lat, p.w, 1.0, 1.0, 1.0)
    c_max, c_rms = eps(Cc, Sc, C, S, N)
    print("N=%d GPU eps_max %.3e eps_rms %.3e; CHarm's algorithm on the CPU %.3e / %.3e"
          % (N, g_max, g_rms, c_max, c_rms))
    assert g_max <= 1.3 * c_max and g_rms <= 1.3 * c_rms


@pytest.mark.parametrize("N,kind", [(1000, "uniform"), (720, "kaula")])
def test_per_degree_error_equals_oracles(N, kind):
    """Round trip: the per-degree differences of the HIP path are those of
    CHarm's algorithm in double, degree by degree (within a factor that
    allows for rounding noise), and the analysis of the oracle's grid agrees
    with the oracle's analysis."""
    if kind == "uniform":
        C, S = uniform_coeffs(N, 7 * N)
        mu = R = 1.0
    else:
        C, S = kaula_coeffs(N, 7 * N)
        mu, R = GM, R_EARTH
    Co, So, grid = roundtrip_dev(N, C, S, mu, R)
    orc = Oracle("d")
    p = grid.pnt
    f = orc.shs_grid(N, C, S, mu, R, p.lat, p.r, p.lon, gl=True)
    Cr, Sr = orc.sha_gl(N, f, p.lat, p.w, R, mu, R)
    d_gpu = dda(Co, So, C, S, N)
    d_orc = dda(Cr, Sr, C, S, N)
    da = dda(C, S, 0 * C, 0 * S, N)                    # degree amplitudes
    ratio = d_gpu[1:] / d_orc[1:]
    print("%s N=%d per degree: dda_gpu/dda_oracle median %.2f, 95%% %.2f, max %.2f; "
          "max relative dda_gpu/da %.2e (oracle %.2e)"
          % (kind, N, np.median(ratio), np.quantile(ratio, 0.95), ratio.max(),
             (d_gpu[2:] / da[2:]).max(), (d_orc[2:] / da[2:]).max()))
    assert np.median(ratio) <= 1.5 and np.quantile(ratio, 0.95) <= 3.0
    # the same grid analysed by both: agreement well below the round-trip error
    g2 = DevArray.from_numpy(f)
    out = DevArray((C.size, 2))
    st = Stream()
    grid.sha(g2, N, mu, R, out, st)
    st.synchronize()
    o = out.numpy()
    d_sha = dda(o[:, 0], o[:, 1], Cr, Sr, N)
    print("%s N=%d: analysis of the oracle's grid vs the oracle's analysis, max dda %.2e"
          % (kind, N, d_sha.max()))
    scale = max(np.abs(Cr).max(), np.abs(Sr).max())
    assert max(np.abs(o[:, 0] - Cr).max(), np.abs(o[:, 1] - Sr).max()) <= 1e-12 * scale
