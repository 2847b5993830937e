"""Host build of the longitude-transform building blocks (the butterflies,
twiddles, pass plan and real/complex packing of charm_amd/csrc/fft_core.h,
driven in the same pass order as the HIP kernels) against numpy.fft, which
matches the unnormalised r2c/c2r conventions of FFTW used by CHarm
(src/shs/shs_grd_fft.c:99-130, src/sha/sha_point.c:490-553)."""
import ctypes as ct
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


@pytest.fixture(scope="module")
def fftlib(tmp_path_factory):
    out = tmp_path_factory.mktemp("fft") / "libfftcheck.so"
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared",
                    "-I" + os.path.join(ROOT, "charm_amd", "csrc"),
                    os.path.join(HERE, "native", "fft_host_check.cpp"), "-o", str(out)],
                   check=True)
    lib = ct.CDLL(str(out))
    dp = ct.POINTER(ct.c_double)
    lib.host_r2c.argtypes = [ct.c_int, ct.c_int, dp, dp]
    lib.host_c2r.argtypes = [ct.c_int, ct.c_int, dp, dp]
    return lib


def _p(a):
    return a.ctypes.data_as(ct.POINTER(ct.c_double))


# L = nlon / 2: tiny, prime (2161 = nmax 2160 + 1), composite, powers of two,
# and the largest length the LDS kernels take
@pytest.mark.parametrize("L", [1, 2, 3, 5, 8, 91, 361, 1000, 2161, 4096])
def test_row_r2c_matches_numpy(fftlib, L):
    rng = np.random.default_rng(L)
    f = rng.standard_normal((3, 2 * L))
    spec = np.zeros((3, L + 1, 2))
    fftlib.host_r2c(L, 3, _p(f), _p(spec))
    ref = np.fft.rfft(f, axis=1)
    got = spec[..., 0] + 1j * spec[..., 1]
    assert np.abs(got - ref).max() <= 1e-14 * np.abs(ref).max()


@pytest.mark.parametrize("L", [1, 2, 3, 5, 8, 91, 361, 1000, 2161, 4096])
def test_row_c2r_matches_numpy(fftlib, L):
    rng = np.random.default_rng(10 + L)
    X = rng.standard_normal((3, L + 1)) + 1j * rng.standard_normal((3, L + 1))
    Xi = np.ascontiguousarray(np.stack([X.real, X.imag], -1))
    out = np.zeros((3, 2 * L))
    fftlib.host_c2r(L, 3, _p(Xi), _p(out))
    # unnormalised inverse; the imaginary parts of X_0 and X_L are ignored
    ref = np.fft.irfft(X, n=2 * L, axis=1) * (2 * L)
    assert np.abs(out - ref).max() <= 1e-14 * np.abs(ref).max()


def test_row_roundtrip_single_harmonic(fftlib):
    # a single cosine of order m: r2c puts L at index m only (closed form)
    L, m = 2161, 37
    lon = np.arange(2 * L) * np.pi / L
    f = np.cos(m * lon)[None, :].copy()
    spec = np.zeros((1, L + 1, 2))
    fftlib.host_r2c(L, 1, _p(f), _p(spec))
    F = spec[0, :, 0] + 1j * spec[0, :, 1]
    expect = np.zeros(L + 1, complex)
    expect[m] = L
    assert np.abs(F - expect).max() <= 1e-12 * L
