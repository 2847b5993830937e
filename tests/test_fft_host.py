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
    ip = ct.POINTER(ct.c_int)
    lib.host_plan.argtypes = [ct.c_int, ip, ip, ip, ip]
    lib.host_r2c_split.argtypes = [ct.c_int, ct.c_int, dp, dp]
    lib.host_c2r_split.argtypes = [ct.c_int, ct.c_int, dp, dp]
    lib.host_split_factor.argtypes = [ct.c_int]
    return lib


DIRECT, BLUESTEIN, RADER = 0, 1, 2


def plan(lib, L):
    M, np_, mode = ct.c_int(), ct.c_int(), ct.c_int()
    R = (ct.c_int * 16)()
    if not lib.host_plan(L, ct.byref(M), ct.byref(np_), R, ct.byref(mode)):
        return None
    return M.value, [R[i] for i in range(np_.value)], mode.value


def is_prime(n):
    return n > 1 and all(n % d for d in range(2, int(n ** 0.5) + 1))


def test_plans(fftlib):
    """Plans multiply out to M, put 16 last when 16 | M, run smooth lengths
    directly, primes with a smooth L - 1 by Rader, and the rest by Bluestein
    over a smooth length >= 2L - 1."""
    assert plan(fftlib, 1) == (1, [], DIRECT)
    assert plan(fftlib, 360) == (360, [5, 9, 8], DIRECT)
    assert plan(fftlib, 8192) == (8192, [8, 8, 8, 16], DIRECT)
    assert plan(fftlib, 2161) == (2160, [5, 9, 3, 16], RADER)   # nmax 2160
    for L in range(1, 600):
        M, R, mode = plan(fftlib, L)
        assert int(np.prod(R)) == M
        assert all(r in (2, 3, 4, 5, 7, 8, 9, 16) for r in R)
        assert all(r != 16 for r in R[:-1])
        if mode == DIRECT:
            assert M == L
        elif mode == RADER:
            assert is_prime(L) and M == L - 1
        else:
            assert 2 * L - 1 <= M
    assert plan(fftlib, 9601) is None    # prime beyond the largest LDS row
    assert plan(fftlib, 4603) == (9216, [9, 8, 8, 16], BLUESTEIN)   # 4602 = 2 3 13 59


def _p(a):
    return a.ctypes.data_as(ct.POINTER(ct.c_double))


# L = nlon / 2: tiny; every radix as a direct plan (3, 5, 7, 9, 16, 360 =
# 5 9 8, 1000, 4096, 2^4 3^2 5 7 = 5040, 9216 near the largest row);
# Bluestein over a smooth M (91, 361, 2161 = nmax 2160 + 1, 4603 with M = 9216)
# Rader on primes with smooth L - 1 (11, 97, 1801, 2161, 4801)
LENGTHS = [1, 2, 3, 5, 7, 8, 9, 11, 16, 91, 97, 360, 361, 1000, 1801, 2161, 4096, 4603, 4801,
           5040, 9216]


@pytest.mark.parametrize("L", LENGTHS)
def test_row_r2c_matches_numpy(fftlib, L):
    rng = np.random.default_rng(L)
    f = rng.standard_normal((3, 2 * L))
    spec = np.zeros((3, L + 1, 2))
    fftlib.host_r2c(L, 3, _p(f), _p(spec))
    ref = np.fft.rfft(f, axis=1)
    got = spec[..., 0] + 1j * spec[..., 1]
    assert np.abs(got - ref).max() <= 1e-14 * np.abs(ref).max()


@pytest.mark.parametrize("L", LENGTHS)
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


# Rows longer than one LDS transform: L = L1 L2 through two passes
# (10801 = 7 x 1543 is the nmax = 10800 GL row; 10000 = 2 x 5000; 12285 =
# 3 x 4095 with a Bluestein L2; 18000 = 2 x 9000)
SPLIT = [(10801, 7), (10000, 2), (12285, 3), (18000, 2)]


@pytest.mark.parametrize("L,L1", SPLIT)
def test_split_r2c_c2r_match_numpy(fftlib, L, L1):
    assert fftlib.host_split_factor(L) == L1
    rng = np.random.default_rng(L)
    f = rng.standard_normal((2, 2 * L))
    spec = np.zeros((2, L + 1, 2))
    fftlib.host_r2c_split(L, 2, _p(f), _p(spec))
    ref = np.fft.rfft(f, axis=1)
    got = spec[..., 0] + 1j * spec[..., 1]
    assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max()
    X = rng.standard_normal((2, L + 1)) + 1j * rng.standard_normal((2, L + 1))
    Xi = np.ascontiguousarray(np.stack([X.real, X.imag], -1))
    out = np.zeros((2, 2 * L))
    fftlib.host_c2r_split(L, 2, _p(Xi), _p(out))
    ref = np.fft.irfft(X, n=2 * L, axis=1) * (2 * L)
    assert np.abs(out - ref).max() <= 1e-13 * np.abs(ref).max()


def test_split_factor_choice(fftlib):
    """No split below the LDS row limit or for a prime; the smallest codelet
    size whose cofactor has an LDS plan otherwise."""
    assert fftlib.host_split_factor(10007) == 0          # prime
    assert fftlib.host_split_factor(2 * 9403) == 0       # 9403 prime, > 9400
    assert fftlib.host_split_factor(9 * 9397) == 0       # 9397 prime, no LDS plan
    assert fftlib.host_split_factor(2 ** 17) == 16       # 8192 per pass
    assert fftlib.host_split_factor(27000) == 3          # 13500 > 9400, 9000 direct
