"""Accuracy of the HIP path against the oracle built in double (CHarm's
algorithm in double) and in __float128 (CHarm's quad build, the
higher-precision reference) at the benchmark configurations:

* C1: nmax = 360, GL grid 361 x 722, synthesis and analysis;
* C2: nmax = 2160, GL grid 2161 x 4322 (sampled rings) and scattered points
  with r = R + U(0, 9000 m) (sampled points);
* C3: nmax = 10800, GL grid 10801 x 21602 (sampled rings; the quad check at a
  few longitudes of two rings).

Bound: max|f - f_ref| <= 1e-12 max|f_ref| (synthesis), max|C - C_ref| <=
1e-12 max|C_ref| (analysis).  Each test prints its errors.
"""
import numpy as np
import pytest

from charm_amd import crd, sha, shc, shs
from charm_amd.device import DevArray, DeviceGrid, Stream, shs_sctr
from charm_amd.workloads import GM, R_EARTH, kaula_coeffs, scattered_points
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

TOL = 1e-12


def relmax(x, ref):
    return float(np.max(np.abs(np.asarray(x) - ref)) / np.max(np.abs(ref)))


@pytest.fixture(scope="module")
def orcd():
    return Oracle("d")


@pytest.fixture(scope="module")
def orcq():
    return Oracle("q")


def test_c1_nmax360_double_and_quad(orcd, orcq):
    N = 360
    C, S = kaula_coeffs(N, N)
    g = crd.PointGridGL(N, R_EARTH)
    m = shc.Shc.from_arrays(N, C, S, GM, R_EARTH)
    f = shs.point(g, m, N)
    fd = orcd.shs_grid(N, C, S, GM, R_EARTH, g.lat, g.r, g.lon, gl=True)
    fq = orcq.shs_grid(N, C, S, GM, R_EARTH, g.lat, g.r, g.lon, gl=True)
    e_d, e_q = relmax(f, fd), relmax(f, fq)
    out = sha.point(g, f, N, GM, R_EARTH)
    Cq, Sq = orcq.sha_gl(N, f, g.lat, g.w, R_EARTH, GM, R_EARTH)
    scale = max(np.abs(Cq).max(), np.abs(Sq).max())
    e_sha_q = max(np.abs(out.c - Cq).max(), np.abs(out.s - Sq).max()) / scale
    e_rt = max(np.abs(out.c - C).max(), np.abs(out.s - S).max()) / scale
    print("C1 N=360: SHS vs double %.2e, vs quad %.2e; SHA vs quad %.2e; round trip %.2e"
          % (e_d, e_q, e_sha_q, e_rt))
    assert max(e_d, e_q, e_sha_q, e_rt) <= TOL


def test_c2_nmax2160_sampled_rings_quad(orcq):
    N = 2160
    C, S = kaula_coeffs(N, N)
    grid = DeviceGrid(N, R_EARTH)
    st = Stream()
    cs = DevArray.from_numpy(np.stack([C, S], axis=1))
    f = DevArray((grid.nlat, grid.nlon))
    grid.shs(cs, N, GM, R_EARTH, f, st)
    st.synchronize()
    fn = f.numpy()
    rings = np.array([0, 3, 540, 1080, 2160], dtype=np.int64)
    p = grid.pnt
    fq = orcq.shs_grid(N, C, S, GM, R_EARTH, p.lat, p.r, p.lon, gl=True, rings=rings)
    e_q = relmax(fn[rings], fq)
    print("C2 N=2160 GL, rings %s: vs quad %.2e" % (rings.tolist(), e_q))
    assert e_q <= TOL


def test_c2_scattered_sampled_points_quad(orcd, orcq):
    N = 2160
    C, S = kaula_coeffs(N, N)
    lat, lon, r = scattered_points(100000, 12345)
    st = Stream()
    cs = DevArray.from_numpy(np.stack([C, S], axis=1))
    d = [DevArray.from_numpy(a) for a in (lat, lon, r)]
    out = DevArray(lat.size)
    shs_sctr(*d, cs, N, GM, R_EARTH, out, st)
    st.synchronize()
    f = out.numpy()
    idx = np.random.default_rng(7).choice(lat.size, 48, replace=False)
    fd = orcd.shs_sctr(N, C, S, GM, R_EARTH, lat[idx], lon[idx], r[idx])
    fq = orcq.shs_sctr(N, C, S, GM, R_EARTH, lat[idx], lon[idx], r[idx])
    e_d, e_q = relmax(f[idx], fd), relmax(f[idx], fq)
    print("C2 N=2160 scattered (48 of 1e5 points, r != R): vs double %.2e, vs quad %.2e"
          % (e_d, e_q))
    assert max(e_d, e_q) <= TOL


def test_c3_nmax10800_sampled_rings(orcd, orcq):
    N = 10800
    C, S = kaula_coeffs(N, N)
    grid = DeviceGrid(N, R_EARTH)
    st = Stream()
    cs = DevArray.from_numpy(np.stack([C, S], axis=1))
    f = DevArray((grid.nlat, grid.nlon))
    grid.shs(cs, N, GM, R_EARTH, f, st)
    st.synchronize()
    fn = f.numpy()
    # round trip through the analysis (rows of 21602 points: two-pass r2c)
    out = DevArray(cs.shape)
    grid.sha(f, N, GM, R_EARTH, out, st)
    st.synchronize()
    o = out.numpy()
    rt = float(max(np.abs(o[:, 0] - C).max(), np.abs(o[:, 1] - S).max()) / np.abs(C).max())
    del f, out
    p = grid.pnt
    rings = np.array([1, 1543, 5400, 10800], dtype=np.int64)
    fd = orcd.shs_grid(N, C, S, GM, R_EARTH, p.lat, p.r, p.lon, gl=True, rings=rings)
    e_d = relmax(fn[rings], fd)
    # quad: two rings at 8 longitudes (direct sums over orders)
    qr = np.array([1543, 5400], dtype=np.int64)
    jl = np.linspace(0, grid.nlon - 1, 8).round().astype(np.int64)
    fq = orcq.shs_grid(N, C, S, GM, R_EARTH, p.lat, p.r, p.lon[jl], gl=True, rings=qr)
    e_q = relmax(fn[np.ix_(qr, jl)], fq)
    print("C3 N=10800 GL: rings %s vs double %.2e; rings %s x 8 longitudes vs quad %.2e; "
          "round trip %.2e" % (rings.tolist(), e_d, qr.tolist(), e_q, rt))
    assert max(e_d, e_q, rt) <= TOL
