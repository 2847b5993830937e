"""The CPU-baseline restatement (oracle/cpu_gl.c: folded ring pairs, ring
blocks in vector loops with a masked X-number phase, OpenMP, FFT rows by
scipy.fft) agrees with the scalar oracle to rounding on GL grids of odd and
even ring counts (with and without an equator ring), at the reference radius
and above it, for degrees where the X-number phase matters (polar rings at
high orders); and both ISA builds give the same results to rounding."""
import numpy as np
import pytest

from charm_amd.workloads import GM, R_EARTH, kaula_coeffs, uniform_coeffs
from oracle.cpu_gl import CpuGL, host_isa
from oracle.oracle import Oracle


@pytest.fixture(scope="module")
def orc():
    return Oracle("d")


def rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.mark.parametrize("N", [0, 1, 2, 5, 30, 63, 64, 200])
@pytest.mark.parametrize("above", [False, True])
def test_restatement_matches_oracle(orc, N, above):
    if N >= 2:
        C, S = kaula_coeffs(N, 11 + N)
    else:
        C, S = uniform_coeffs(N, 11 + N)
    lat, w, lon = orc.gl_grid(N)
    r = R_EARTH + (7000.0 if above else 0.0)
    f0 = orc.shs_grid(N, C, S, GM, R_EARTH, lat, np.full(len(lat), r), lon, gl=True)
    cg = CpuGL(4)
    f1 = cg.shs(N, C, S, GM, R_EARTH, r, lat, len(lon))
    assert f1.shape == f0.shape
    assert rel(f1, f0) <= 1e-13
    C0, S0 = orc.sha_gl(N, f0, lat, w, r, GM, R_EARTH)
    C1, S1 = cg.sha(N, f0, lat, w, r, GM, R_EARTH)
    scale = np.abs(C0).max()
    assert np.abs(C1 - C0).max() <= 1e-13 * scale
    assert np.abs(S1 - S0).max() <= 1e-13 * scale


def test_restatement_high_degree_roundtrip(orc):
    """N = 700: rings near the poles underflow at high orders, so blocks run
    their masked X-number phase; the round trip recovers the coefficients and
    sampled rings equal the oracle's."""
    N = 700
    C, S = uniform_coeffs(N, 700)
    lat, w, lon = orc.gl_grid(N)
    cg = CpuGL(8)
    f = cg.shs(N, C, S, 1.0, 1.0, 1.0, lat, len(lon))
    rings = [0, 1, 2, 17, 175, 349, 350, 351, 699, 700]
    f0 = orc.shs_grid(N, C, S, 1.0, 1.0, lat, np.ones(len(lat)), lon, gl=True, rings=rings)
    # 245 thousand unit-size terms per value, summed in another order
    assert rel(f[rings], f0) <= 1e-12
    Co, So = cg.sha(N, f, lat, w, 1.0, 1.0, 1.0)
    assert max(np.abs(Co - C).max(), np.abs(So - S).max()) <= 1e-12


def test_both_isa_builds_agree(orc):
    if host_isa() != "avx512":
        pytest.skip("host has no AVX-512")
    N = 150
    C, S = kaula_coeffs(N, 5)
    lat, w, lon = orc.gl_grid(N)
    fa = CpuGL(2, "avx2").shs(N, C, S, GM, R_EARTH, R_EARTH, lat, len(lon))
    fb = CpuGL(2, "avx512").shs(N, C, S, GM, R_EARTH, R_EARTH, lat, len(lon))
    assert rel(fa, fb) <= 1e-14
