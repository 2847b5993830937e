"""Parity of the HIP path (through the C ABI) with the oracle and with CHarm's
own golden fixtures.

Tolerances:
* CHarm's fixtures: the reference's own test bound, misc_is_nearly_equal with
  10*glob_threshold (2.2e-13) per value (10*glob_threshold2 for the dsun
  files), plus the north-star bound max|f - f_ref| <= 1e-12 max|f_ref|.
* Oracle comparisons at larger degrees: max|f - f_oracle| <= 1e-12 max|f|
  (synthesis) and max|C - C_oracle| <= 1e-12 max|C| (analysis).
"""
import numpy as np
import pytest

import charm_amd
from charm_amd import crd, sha, shc, shs
from charm_amd.workloads import GM, R_EARTH, kaula_coeffs, scattered_points
from charm_fixtures import (BREAK_SYMM, CUSTOM_GRIDS, DATA, DELTAR, THRESH, THRESH2, custom_grid,
                            egm96_test_model, nearly_equal, read_fixture)
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

TOL = 1e-12


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if charm_amd.lib().charm_amd_device_count() < 1:
        pytest.fail("no HIP device visible: the GPU tests need an MI355X")


@pytest.fixture(scope="module")
def orc():
    return Oracle("d")


def relmax(x, ref):
    x, ref = np.asarray(x), np.asarray(ref)
    return float(np.max(np.abs(x - ref)) / max(np.max(np.abs(ref)), 1e-300))


def model(nmax, C, S, mu, R):
    return shc.Shc.from_arrays(nmax, np.array(C, dtype=float), np.array(S, dtype=float), mu, R)


# --------------------------------------------------------------------------
# CHarm golden fixtures
# --------------------------------------------------------------------------
@pytest.mark.parametrize("nmax", range(0, 5))
@pytest.mark.parametrize("deltar", (0, 1))
def test_shs_gl_fixtures(nmax, deltar):
    mu, R, C, S = egm96_test_model()
    m = model(10, C, S, mu, R)
    g = crd.PointGridGL(nmax, R + DELTAR * deltar)
    f = shs.point(g, m, nmax).ravel()
    ref = read_fixture("double", "shs_p_nx%d_dr%d_fft1_gl.txt" % (nmax, deltar))
    assert nearly_equal(f, ref, 10 * THRESH["double"]).all(), relmax(f, ref)
    assert relmax(f, ref) <= TOL
    # error against CHarm built in quad precision
    refq = read_fixture("quad", "shs_p_nx%d_dr%d_fft1_gl.txt" % (nmax, deltar))
    assert relmax(f, refq) <= TOL


def _custom_cases():
    import os
    out = []
    for nmax in range(0, 5):
        for i in range(len(CUSTOM_GRIDS)):
            for fft in (0, 1):
                for s in (0, 1):
                    name = "shs_p_nx%d_n%d_dr%%d_fft%d_s%d.txt" % (nmax, i, fft, s)
                    if os.path.exists(os.path.join(DATA, "double", name % 0)):
                        out.append((nmax, i, fft, s, name))
    return out


@pytest.mark.parametrize("nmax,i,fft,s,name", _custom_cases())
def test_shs_custom_grid_fixtures(nmax, i, fft, s, name):
    """Custom grids of CHarm's point-synthesis test (tests/check_shs_point_all.c:
    537-620): with the FFT along parallels (fft1) and with direct longitude
    sums (fft0, half the circle), symmetric and not, at R and R + 1 km."""
    mu, R, C, S = egm96_test_model()
    m = model(10, C, S, mu, R)
    lat, lon = custom_grid(i, fft, s == 1)
    for deltar in (0, 1):
        g = crd.PointGrid(lat, lon, np.full(lat.size, R + DELTAR * deltar))
        f = shs.point(g, m, nmax).ravel()
        ref = read_fixture("double", name % deltar)
        assert nearly_equal(f, ref, 100 * THRESH["double"]).all(), relmax(f, ref)
        assert relmax(f, read_fixture("quad", name % deltar)) <= TOL


@pytest.mark.parametrize("nmax", range(0, 5))
def test_shs_sctr_fixtures(nmax):
    from charm_fixtures import scattered_points as fixture_points

    mu, R, C, S = egm96_test_model()
    m = model(10, C, S, mu, R)
    for i, npt in enumerate([1, 2, 3, 4, 5, 6, 7, 8, 9, 20, 30, 31]):
        lat, lon = fixture_points(npt, "double")
        for deltar in (0, 1):
            p = crd.PointSctr(lat, lon, np.full(npt, R + DELTAR * deltar))
            f = shs.point(p, m, nmax)
            ref = read_fixture("double", "shs_p_nx%d_n%d_dr%d_sctr.txt" % (nmax, i, deltar))
            assert nearly_equal(f, ref, 10 * THRESH["double"]).all(), (i, deltar, relmax(f, ref))


@pytest.mark.parametrize("nmax", range(150, 155))
@pytest.mark.parametrize("s", (0, 1))
def test_shs_dsun_fixtures(nmax, s):
    """GL latitudes, one longitude: the lumped sums themselves (non-FFT path)."""
    ncs = (nmax + 1) * (nmax + 2) // 2
    from charm_amd.workloads import degree_order

    n_of, m_of = degree_order(nmax)
    C = 1.0 / (n_of + 1.0)
    S = np.where(m_of > 0, C, 0.0)
    gl = crd.PointGridGL(nmax, 1.0)
    lat = gl.lat.copy()
    if s == 0:
        lat[0] -= BREAK_SYMM
    for deltar in (0, 1):
        # a GL-typed point set with the shifted latitude and lon = [0]
        pnt = crd.PointGridGL(nmax, 1.0 + DELTAR * deltar)
        pnt.lat[:] = lat
        pnt._Point.contents.nlon = 1
        pnt._Point.contents.npoint = len(lat)
        m = model(nmax, C, S, 1.0, 1.0)
        f = shs.point(pnt, m, nmax).ravel()[: len(lat)]
        ref = read_fixture("double", "shs_p_nx%d_n0_dr%d_fft0_s%d_dsun.txt" % (nmax, deltar, s))
        assert nearly_equal(f, ref, 10 * THRESH2["double"]).all(), relmax(f, ref)
        assert relmax(f, ref) <= TOL
    assert ncs == C.size


# --------------------------------------------------------------------------
# Oracle comparisons (Kaula-like models)
# --------------------------------------------------------------------------
def test_shs_gl_360_all_rings(orc):
    N = 360
    C, S = kaula_coeffs(N, 1)
    g = crd.PointGridGL(N, R_EARTH)
    f = shs.point(g, model(N, C, S, GM, R_EARTH), N)
    ref = orc.shs_grid(N, C, S, GM, R_EARTH, g.lat, g.r, g.lon, gl=True)
    assert relmax(f, ref) <= TOL, relmax(f, ref)


def test_shs_gl_2160_sampled_rings(orc):
    """Full-size grid; the oracle checks polar rings (X-number region),
    mid-latitude rings, the equator and southern rings."""
    N = 2160
    C, S = kaula_coeffs(N, 2)
    g = crd.PointGridGL(N, R_EARTH)
    f = shs.point(g, model(N, C, S, GM, R_EARTH), N)
    rings = [0, 1, 2, 37, 300, 700, 1080, 1500, 2100, 2160]
    ref = orc.shs_grid(N, C, S, GM, R_EARTH, g.lat, g.r, g.lon, gl=True, rings=rings)
    assert relmax(f[rings], ref) <= TOL, relmax(f[rings], ref)


def test_shs_gl_r_differs_from_R(orc):
    N = 120
    C, S = kaula_coeffs(N, 3)
    g = crd.PointGridGL(N, R_EARTH + 5000.0)
    f = shs.point(g, model(N, C, S, GM, R_EARTH), N)
    ref = orc.shs_grid(N, C, S, GM, R_EARTH, g.lat, g.r, g.lon, gl=True)
    assert relmax(f, ref) <= TOL


@pytest.mark.parametrize("symmetric", (True, False))
@pytest.mark.parametrize("nlon", (7, 400))
def test_shs_custom_grid(orc, symmetric, nlon):
    """Custom grids: folded when the latitudes are symmetric; FFT when the
    longitudes allow it (nlon = 400), direct longitude sums otherwise."""
    N = 90
    C, S = kaula_coeffs(N, 4)
    if symmetric:
        h = np.linspace(1.5, 0.1, 20)
        lat = np.concatenate([h, [0.0], -h[::-1]])
    else:
        lat = np.linspace(1.4, -0.9, 31)
    lon = np.arange(nlon) * (2 * np.pi / nlon) if nlon == 400 else np.linspace(0.2, 1.0, nlon)
    r = np.full(lat.size, R_EARTH + 100.0)
    p = crd.PointGrid(lat, lon, r)
    f = shs.point(p, model(N, C, S, GM, R_EARTH), N)
    ref = orc.shs_grid(N, C, S, GM, R_EARTH, lat, r, lon, gl=False)
    assert relmax(f, ref) <= TOL, relmax(f, ref)


def test_shs_sctr_vs_oracle(orc):
    N = 360
    C, S = kaula_coeffs(N, 5)
    lat, lon, r = scattered_points(3000, 6)
    f = shs.point(crd.PointSctr(lat, lon, r), model(N, C, S, GM, R_EARTH), N)
    ref = orc.shs_sctr(N, C, S, GM, R_EARTH, lat, lon, r)
    assert relmax(f, ref) <= TOL, relmax(f, ref)


def test_sha_gl_vs_oracle(orc):
    N = 200
    C, S = kaula_coeffs(N, 7)
    g = crd.PointGridGL(N, R_EARTH)
    f = shs.point(g, model(N, C, S, GM, R_EARTH), N)
    out = sha.point(g, f, N, GM, R_EARTH)
    Co, So = orc.sha_gl(N, f, g.lat, g.w, R_EARTH, GM, R_EARTH)
    scale = max(np.abs(Co).max(), np.abs(So).max())
    assert np.abs(out.c - Co).max() <= TOL * scale
    assert np.abs(out.s - So).max() <= TOL * scale
    # the analysis inverts the synthesis (exact quadrature)
    assert np.abs(out.c - C).max() <= TOL * scale
    assert np.abs(out.s - S).max() <= TOL * scale


def test_sha_roundtrip_rescale_and_zero_fill(orc):
    """Grid at R + 1000 m, coefficients returned for radius R; the output
    model is of higher degree than the analysis, so the rest is zero."""
    N = 150
    C, S = kaula_coeffs(N, 8)
    R1 = R_EARTH + 1000.0
    g = crd.PointGridGL(N, R1)
    f = shs.point(g, model(N, C, S, GM, R_EARTH), N)
    import ctypes as ct

    out = shc.Shc.from_garbage(N + 10, GM, R_EARTH)
    out.c[:] = np.nan
    fn = charm_amd.lib().charm_sha_point
    fn.restype = None
    fn.argtypes = [ct.POINTER(crd._Point), ct.POINTER(ct.c_double), ct.c_ulong,
                   ct.POINTER(shc._Shc), ct.POINTER(charm_amd._err._Err)]
    err = charm_amd._err.init()
    fc = np.ascontiguousarray(f)
    fn(g._Point, fc.ctypes.data_as(ct.POINTER(ct.c_double)), N, out._Shc, err)
    charm_amd._err.handler(err)
    charm_amd._err.free(err)
    Co, So = orc.sha_gl(N, f, g.lat, g.w, R1, GM, R_EARTH, nmax_shc=N + 10)
    scale = np.abs(Co).max()
    assert np.isfinite(out.c).all()
    assert np.abs(out.c - Co).max() <= TOL * scale
    assert np.abs(out.s - So).max() <= TOL * scale
    from charm_amd.shc import index

    assert out.c[index(N + 10, N + 5, 3)] == 0.0
    assert out.mu == GM and out.r == R_EARTH


def test_sha_roundtrip_2160():
    N = 2160
    C, S = kaula_coeffs(N, 9)
    g = crd.PointGridGL(N, R_EARTH)
    f = shs.point(g, model(N, C, S, GM, R_EARTH), N)
    out = sha.point(g, f, N, GM, R_EARTH)
    assert np.abs(out.c - C).max() <= TOL * np.abs(C).max()
    assert np.abs(out.s - S).max() <= TOL * np.abs(C).max()


# --------------------------------------------------------------------------
# Device-resident entry points
# --------------------------------------------------------------------------
def test_device_grid_matches_host_api():
    from charm_amd.device import DevArray, DeviceGrid, Stream

    N = 300
    C, S = kaula_coeffs(N, 10)
    g = crd.PointGridGL(N, R_EARTH)
    f_host = shs.point(g, model(N, C, S, GM, R_EARTH), N)
    dg = DeviceGrid(N, R_EARTH)
    st = Stream()
    cs = DevArray.from_numpy(np.stack([C, S], axis=1))
    f = DevArray((dg.nlat, dg.nlon))
    dg.shs(cs, N, GM, R_EARTH, f, st)
    out = DevArray(cs.shape)
    dg.sha(f, N, GM, R_EARTH, out, st)
    st.synchronize()
    assert np.array_equal(f.numpy(), f_host)
    cs_np = out.numpy()
    assert np.abs(cs_np[:, 0] - C).max() <= TOL * np.abs(C).max()
    assert np.abs(cs_np[:, 1] - S).max() <= TOL * np.abs(C).max()


def test_device_grid_radius_differs_from_R():
    """A device grid whose radius is not R: one column of (R/r)^k for every
    row, values equal to the host entry point's."""
    from charm_amd.device import DevArray, DeviceGrid, Stream

    N, Ng = 150, 300
    C, S = kaula_coeffs(N, 21)
    r0 = R_EARTH + 7000.0
    g = crd.PointGridGL(Ng, r0)
    f_host = shs.point(g, model(N, C, S, GM, R_EARTH), N)
    dg = DeviceGrid(Ng, r0)
    st = Stream()
    cs = DevArray.from_numpy(np.stack([C, S], axis=1))
    f = DevArray((dg.nlat, dg.nlon))
    dg.shs(cs, N, GM, R_EARTH, f, st)
    st.synchronize()
    assert np.array_equal(f.numpy(), f_host)


@pytest.mark.parametrize("K", (2, 3, 4, 5, 6, 8))
def test_device_grid_skips_all_zero_waves(K, monkeypatch):
    """Repeated transforms on one device grid skip the waves (order, ring
    tile) whose every term the first transform found to be exactly zero.  At
    nmax = 1600 several hundred waves near the poles are such.  Synthesis:
    every call equals the host entry point (which never skips) bit for bit.
    Analysis on a fresh grid: the first call records the flags without
    skipping, later calls skip and must equal it bit for bit.  K = 5, 6 and 8
    (analysis only; synthesis keeps its rule): 801 ring pairs are two full
    tiles and one of 161 rings (K = 5), two full tiles and one of 33 rings
    (K = 6), or one full tile and one of 289 rings, run at 6 pairs per lane
    (K = 8); K = 4: three full tiles and one of 33 rings."""
    from charm_amd.device import DevArray, DeviceGrid, Stream

    monkeypatch.setenv("CHARM_AMD_SHS_K", str(K))
    monkeypatch.setenv("CHARM_AMD_SHA_K", str(K))
    N = 1600
    C, S = kaula_coeffs(N, 33)
    g = crd.PointGridGL(N, R_EARTH)
    f_host = shs.point(g, model(N, C, S, GM, R_EARTH), N)
    st = Stream()
    cs = DevArray.from_numpy(np.stack([C, S], axis=1))
    dg = DeviceGrid(N, R_EARTH)
    f = DevArray((dg.nlat, dg.nlon))
    for call in range(3):
        dg.shs(cs, N, GM, R_EARTH, f, st)
        st.synchronize()
        assert np.array_equal(f.numpy(), f_host), "synthesis call %d" % call
    dg2 = DeviceGrid(N, R_EARTH)
    fd = DevArray.from_numpy(f_host)
    out = DevArray(cs.shape)
    first = None
    for call in range(3):
        dg2.sha(fd, N, GM, R_EARTH, out, st)
        st.synchronize()
        o = out.numpy()
        if first is None:
            first = o
            assert np.abs(o[:, 0] - C).max() <= TOL * np.abs(C).max()
            assert np.abs(o[:, 1] - S).max() <= TOL * np.abs(C).max()
        else:
            assert np.array_equal(o, first), "analysis call %d" % call


def test_every_ring_pairs_per_lane(monkeypatch):
    """Synthesis at every built K gives the same grid bit for bit (a ring's
    sums do not depend on how rings are tiled), and analysis at every K
    recovers the coefficients.  nmax = 810: 406 ring pairs, at K = 4 one full
    tile and one of 150 rings, run at 3 pairs per lane; at K = 6 one tile of
    384 and one of 22; at K = 8 one of 406, run at 8."""
    from charm_amd.device import DevArray, DeviceGrid, Stream

    N = 810
    C, S = kaula_coeffs(N, 41)
    st = Stream()
    cs = DevArray.from_numpy(np.stack([C, S], axis=1))
    dg = DeviceGrid(N, R_EARTH)
    f = DevArray((dg.nlat, dg.nlon))
    ref = None
    for K in (1, 2, 3, 4):
        monkeypatch.setenv("CHARM_AMD_SHS_K", str(K))
        dg.shs(cs, N, GM, R_EARTH, f, st)
        st.synchronize()
        assert charm_amd.lib().charm_amd_last_ring_k(0) == K
        if ref is None:
            ref = f.numpy()
        else:
            assert np.array_equal(f.numpy(), ref), "synthesis K=%d" % K
    out = DevArray(cs.shape)
    first = None
    for K in (2, 3, 4, 5, 6, 8):
        monkeypatch.setenv("CHARM_AMD_SHA_K", str(K))
        dg.sha(f, N, GM, R_EARTH, out, st)
        st.synchronize()
        assert charm_amd.lib().charm_amd_last_ring_k(1) == K
        o = out.numpy()
        assert np.abs(o[:, 0] - C).max() <= TOL * np.abs(C).max(), "analysis K=%d" % K
        assert np.abs(o[:, 1] - S).max() <= TOL * np.abs(C).max(), "analysis K=%d" % K
        if first is None:
            first = o
        else:
            # the same sums, added across lanes and tiles in another order
            assert np.abs(o - first).max() <= 1e-14 * np.abs(C).max(), "analysis K=%d" % K


def test_device_grid_one_slot_rescale_roundtrip():
    """From 2048 order quads on (nmax >= 8188) analysis keeps one partial slot
    and its blocks apply s_nm and the rescale (r0/R)^n themselves.  Round trip
    at nmax = 8200 on a GL grid of radius R + 500 m: synthesis, then analysis
    with the rescale, recovers the coefficients.  (R/r0)^n stays above 0.52 up
    to n = 8200, so the round trip is well conditioned, while a rescale left
    out would be off by 4e-10 of max|C| already at degree 2.  (At R + 5 km the
    factor reaches 1.6e-3 and both analysis paths, one slot or several, lose
    1e-11 of max|C| in the high zonal degrees.)  Then the per-degree relative
    error of a second round trip, with U[-1, 1) coefficients."""
    from charm_amd.device import DevArray, DeviceGrid, Stream

    N = 8200
    C, S = kaula_coeffs(N, 8200)
    r0 = R_EARTH + 500.0
    dg = DeviceGrid(N, r0)
    st = Stream()
    cs = DevArray.from_numpy(np.stack([C, S], axis=1))
    f = DevArray((dg.nlat, dg.nlon))
    dg.shs(cs, N, GM, R_EARTH, f, st)
    out = DevArray(cs.shape)
    dg.sha(f, N, GM, R_EARTH, out, st)
    st.synchronize()
    o = out.numpy()
    err = max(np.abs(o[:, 0] - C).max(), np.abs(o[:, 1] - S).max()) / np.abs(C).max()
    print("nmax 8200, r0 = R + 500 m: round trip %.2e" % err)
    assert err <= TOL
    # Per degree (as CHarm's dda, src/shc/shc_dda.c:16-36), on coefficients
    # drawn from U[-1, 1) as CHarm's accuracy protocol draws them, so that
    # every degree carries signal (Kaula's fall as 1e-5 / n^2, below the
    # rounding of C00 = 1 at high degree): the error over a degree's orders
    # against that degree's coefficients.  A wrong s_nm offset, a missing
    # rescale or a missed write at any degree would show as an O(1e-4 n) or
    # O(1) relative error there.
    rng = np.random.default_rng(8201)
    U = rng.uniform(-1.0, 1.0, (C.size, 2))
    m_of = np.repeat(np.arange(N + 1), np.arange(N + 1, 0, -1))
    U[m_of == 0, 1] = 0.0
    cu = DevArray.from_numpy(U)
    dg.shs(cu, N, GM, R_EARTH, f, st)
    dg.sha(f, N, GM, R_EARTH, out, st)
    st.synchronize()
    o = out.numpy()
    first = np.concatenate(([0], np.cumsum(np.arange(N + 1, 0, -1))[:-1]))
    n_of = m_of + (np.arange(C.size) - np.repeat(first, np.arange(N + 1, 0, -1)))
    d2 = np.bincount(n_of, ((o - U) ** 2).sum(axis=1), N + 1)
    c2 = np.bincount(n_of, (U ** 2).sum(axis=1), N + 1)
    rel = np.sqrt(d2[1:] / c2[1:])
    print("U[-1,1): per-degree relative error max %.2e (degree %d), degree 0 %.2e"
          % (rel.max(), 1 + int(rel.argmax()), abs(o[0, 0] - U[0, 0])))
    assert rel.max() <= 1e-10 and abs(o[0, 0] - U[0, 0]) <= 1e-10


@pytest.mark.parametrize("nlon", (20000, 21602, 24570, 10802, 19682))
def test_split_row_lengths_synthesis(orc, nlon):
    """Rows longer than one LDS transform (L = 10000 = 2 x 5000, 10801 =
    7 x 1543, 12285 = 3 x 4095), or with no single LDS plan (L = 5401 =
    11 x 491 and 9841 = 13 x 757, split with the table DFTs of 11 and 13
    points): the two-pass c2r against the oracle's direct longitude sums."""
    N = 60
    C, S = kaula_coeffs(N, 17)
    lat = np.array([1.2, 0.4, -0.05, -0.9])
    lon = np.arange(nlon) * (2.0 * np.pi / nlon)
    g = crd.PointGrid(lat, lon, np.full(lat.size, R_EARTH))
    f = shs.point(g, model(N, C, S, GM, R_EARTH), N)
    ref = orc.shs_grid(N, C, S, GM, R_EARTH, lat, np.full(lat.size, R_EARTH), lon, gl=False)
    assert relmax(f, ref) <= TOL, relmax(f, ref)


@pytest.mark.parametrize("N", [4999, 5400])
def test_split_row_lengths_roundtrip(N):
    """DH2 grid of nmax 4999 (10000 x 20000) and GL grid of nmax 5400
    (5401 x 10802, rows split 11 x 491): two-pass c2r and r2c."""
    from charm_amd.device import DevArray, Stream

    C, S = kaula_coeffs(N, 5)
    g = crd.PointGridDH2(N, R_EARTH) if N == 4999 else crd.PointGridGL(N, R_EARTH)
    m = model(N, C, S, GM, R_EARTH)
    f = shs.point(g, m, N)
    out = sha.point(g, f, N, GM, R_EARTH)
    err = max(np.abs(out.c - C).max(), np.abs(out.s - S).max()) / np.abs(C).max()
    assert err <= TOL, err


def test_device_sctr_matches_host_api():
    from charm_amd.device import DevArray, shs_sctr, synchronize

    N = 200
    C, S = kaula_coeffs(N, 11)
    lat, lon, r = scattered_points(1000, 12)
    f_host = shs.point(crd.PointSctr(lat, lon, r), model(N, C, S, GM, R_EARTH), N)
    cs = DevArray.from_numpy(np.stack([C, S], axis=1))
    out = DevArray(lat.size)
    shs_sctr(DevArray.from_numpy(lat), DevArray.from_numpy(lon), DevArray.from_numpy(r), cs, N,
             GM, R_EARTH, out)
    synchronize()
    assert relmax(out.numpy(), f_host) <= 1e-14
