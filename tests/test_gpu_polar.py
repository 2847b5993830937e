"""Polar optimisation (charm_glob_polar_optimization_a1/a2, off by default):
order m is dropped at points where m - nmax cos(lat) > max(a1, a2 nmax)
(src/misc/misc_polar_optimization_threshold.c:16-20,
misc_polar_optimization_apply.c:16-41).  The HIP path and the oracle apply it
per ring / per point; CHarm tests it jointly for blocks of SIMD_SIZE x BLOCK_S
points, a build-dependent size, and no CHarm test or fixture covers it, so the
parity here is against the oracle's restatement (parity unpinned against
CHarm itself)."""
import ctypes as ct

import numpy as np
import pytest

import charm_amd
from charm_amd import crd, sha, shc, shs
from charm_amd.device import DevArray, DeviceGrid, Stream
from charm_amd.workloads import GM, R_EARTH, kaula_coeffs, scattered_points
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

TOL = 1e-12
A1, A2 = 20, 0.0   # threshold max(20, 0 * nmax) = 20


@pytest.fixture
def polar_on():
    lib = charm_amd.lib()
    a1 = ct.c_ulong.in_dll(lib, "charm_glob_polar_optimization_a1")
    a2 = ct.c_double.in_dll(lib, "charm_glob_polar_optimization_a2")
    old = (a1.value, a2.value)
    orc = Oracle("d")
    a1.value, a2.value = A1, A2
    orc.set_polar(A1, A2)
    try:
        yield orc
    finally:
        a1.value, a2.value = old
        orc.set_polar(100, -1.0)


def relmax(x, ref):
    return float(np.max(np.abs(np.asarray(x) - ref)) / np.max(np.abs(ref)))


def model(nmax, C, S):
    return shc.Shc.from_arrays(nmax, C, S, GM, R_EARTH)


def test_polar_grid_synthesis(polar_on):
    N = 180
    C, S = kaula_coeffs(N, 31)
    g = crd.PointGridGL(N, R_EARTH)
    f = shs.point(g, model(N, C, S), N)
    ref = polar_on.shs_grid(N, C, S, GM, R_EARTH, g.lat, g.r, g.lon, gl=True)
    assert relmax(f, ref) <= TOL
    # the optimisation is applied: terms were dropped somewhere
    polar_on.set_polar(100, -1.0)
    full = polar_on.shs_grid(N, C, S, GM, R_EARTH, g.lat, g.r, g.lon, gl=True)
    polar_on.set_polar(A1, A2)
    assert np.any(ref != full)


def test_polar_scattered_synthesis(polar_on):
    N = 180
    C, S = kaula_coeffs(N, 32)
    lat, lon, r = scattered_points(2000, 33)
    f = shs.point(crd.PointSctr(lat, lon, r), model(N, C, S), N)
    ref = polar_on.shs_sctr(N, C, S, GM, R_EARTH, lat, lon, r)
    assert relmax(f, ref) <= TOL


def test_polar_analysis(polar_on):
    N = 120
    C, S = kaula_coeffs(N, 34)
    g = crd.PointGridGL(N, R_EARTH)
    rng = np.random.default_rng(35)
    f = rng.standard_normal(g.shape(N))
    out = sha.point(g, f, N, GM, R_EARTH)
    Co, So = polar_on.sha_gl(N, f, g.lat, g.w, R_EARTH, GM, R_EARTH)
    scale = max(np.abs(Co).max(), np.abs(So).max())
    assert np.abs(out.c - Co).max() <= TOL * scale
    assert np.abs(out.s - So).max() <= TOL * scale


def test_polar_multi_set_matches_single(polar_on):
    N, nset = 160, 16
    grid = DeviceGrid(N, R_EARTH)
    st = Stream()
    cs = np.stack([np.stack(kaula_coeffs(N, 40 + i), axis=1) for i in range(nset)])
    out = DevArray((nset, grid.nlat, grid.nlon))
    grid.shs_multi(DevArray.from_numpy(cs), N, nset, GM, R_EARTH, out, st)
    st.synchronize()
    got = out.numpy()
    for i in (0, nset - 1):
        f = DevArray((grid.nlat, grid.nlon))
        grid.shs(DevArray.from_numpy(cs[i]), N, GM, R_EARTH, f, st)
        st.synchronize()
        ref = f.numpy()
        assert relmax(got[i], ref) <= TOL
