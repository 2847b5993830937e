"""Polar optimisation (charm_glob_polar_optimization_a1/a2, off by default).

CHarm skips order m for a whole block of rings or points once every member
passes m - nmax cos(lat) > max(a1, a2 nmax)
(src/misc/misc_polar_optimization_threshold.c:16-20,
misc_polar_optimization_apply.c:16-41), the blocks being SIMD_SIZE x BLOCK_S
rings/points in synthesis and SIMD_SIZE x BLOCK_A rings in analysis
(src/shs/shs_point_grd.c:860-862, shs_point_sctr.c:518-520,
src/sha/sha_point.c:716-718).  The library reproduces the AVX-512 double
build by default (blocks of 32 and 64) and other builds through
charm_amd_polar_simd_size and charm_glob_{shs,sha}_block_lat_multiplier.
The oracle's block rule is pinned in tests/test_oracle_polar.py."""
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


class Polar:
    """Sets the library's and the oracle's polar optimisation together."""

    def __init__(self):
        lib = charm_amd.lib()
        self.a1 = ct.c_ulong.in_dll(lib, "charm_glob_polar_optimization_a1")
        self.a2 = ct.c_double.in_dll(lib, "charm_glob_polar_optimization_a2")
        self.simd = ct.c_int.in_dll(lib, "charm_amd_polar_simd_size")
        self.ms = ct.c_size_t.in_dll(lib, "charm_glob_shs_block_lat_multiplier")
        self.ma = ct.c_size_t.in_dll(lib, "charm_glob_sha_block_lat_multiplier")
        self.old = (self.a1.value, self.a2.value, self.simd.value, self.ms.value, self.ma.value)
        self.orc = Oracle("d")

    def set(self, a1, a2, simd=8, ms=4, ma=8):
        self.a1.value, self.a2.value = a1, a2
        self.simd.value, self.ms.value, self.ma.value = simd, ms, ma
        self.orc.set_polar(a1, a2, simd * ms, simd * ma)

    def restore(self):
        self.a1.value, self.a2.value, self.simd.value, self.ms.value, self.ma.value = self.old
        self.orc.set_polar(100, -1.0)


@pytest.fixture
def polar():
    p = Polar()
    assert p.old == (100, -1.0, 8, 4, 8)   # CHarm's defaults, AVX-512 blocks
    p.set(A1, A2)
    try:
        yield p
    finally:
        p.restore()


def relmax(x, ref):
    return float(np.max(np.abs(np.asarray(x) - ref)) / np.max(np.abs(ref)))


def model(nmax, C, S):
    return shc.Shc.from_arrays(nmax, C, S, GM, R_EARTH)


@pytest.mark.parametrize("blocks", [(8, 4, 8), (1, 8, 8), (4, 4, 8)])
def test_polar_grid_synthesis(polar, blocks):
    N = 180
    polar.set(A1, A2, *blocks)
    C, S = kaula_coeffs(N, 31)
    g = crd.PointGridGL(N, R_EARTH)
    f = shs.point(g, model(N, C, S), N)
    ref = polar.orc.shs_grid(N, C, S, GM, R_EARTH, g.lat, g.r, g.lon, gl=True)
    assert relmax(f, ref) <= TOL
    # the optimisation is applied: terms were dropped somewhere
    polar.orc.set_polar(100, -1.0)
    full = polar.orc.shs_grid(N, C, S, GM, R_EARTH, g.lat, g.r, g.lon, gl=True)
    assert np.any(ref != full)


def test_polar_grid_synthesis_dh_and_custom(polar):
    N = 120
    C, S = kaula_coeffs(N, 36)
    g = crd.PointGridDH1(N, R_EARTH + 300.0)
    f = shs.point(g, model(N, C, S), N)
    ref = polar.orc.shs_grid(N, C, S, GM, R_EARTH, g.lat, g.r, g.lon, gl=False, dh=1)
    assert relmax(f, ref) <= TOL
    north = np.linspace(np.pi / 2 - 0.001, 0.02, 70)
    lat = np.concatenate([north, -north[::-1]])
    lon = np.linspace(0, 2 * np.pi, 2 * N + 2, endpoint=False)
    r = np.full(lat.size, R_EARTH)
    g = crd.PointGrid(lat, lon, r)
    f = shs.point(g, model(N, C, S), N)
    ref = polar.orc.shs_grid(N, C, S, GM, R_EARTH, lat, r, lon, gl=False)
    assert relmax(f, ref) <= TOL


@pytest.mark.parametrize("npoint", [2000, 300000])
def test_polar_scattered_synthesis(polar, npoint):
    """Point blocks follow the input order, which the latitude sort of the
    HIP path reorders; the first points sit in the polar caps so that whole
    blocks pass the test."""
    N = 180
    C, S = kaula_coeffs(N, 32)
    lat, lon, r = scattered_points(npoint, 33)
    rng = np.random.default_rng(34)
    ncap = npoint // 2
    lat[:ncap] = np.arcsin(rng.uniform(0.85, 1.0, ncap)) * rng.choice([-1.0, 1.0], ncap)
    f = shs.point(crd.PointSctr(lat, lon, r), model(N, C, S), N)
    idx = np.arange(npoint) if npoint <= 4000 else \
        np.concatenate([np.arange(512), rng.choice(npoint, 512, replace=False)])
    # blocks of 32 input points: sample whole blocks
    blk = np.unique(idx // 32)
    sel = (blk[:, None] * 32 + np.arange(32)[None, :]).ravel()
    sel = sel[sel < npoint]
    ref = polar.orc.shs_sctr(N, C, S, GM, R_EARTH, lat[sel], lon[sel], r[sel])
    assert relmax(f[sel], ref) <= TOL
    polar.orc.set_polar(100, -1.0)
    full = polar.orc.shs_sctr(N, C, S, GM, R_EARTH, lat[sel[:4096]], lon[sel[:4096]],
                              r[sel[:4096]])
    assert np.any(full != ref[:4096])


@pytest.mark.parametrize("blocks,K", [((8, 4, 8), None), ((1, 8, 8), None), ((8, 4, 8), 8),
                                      ((8, 4, 8), 6)])
def test_polar_analysis(polar, blocks, K, monkeypatch):
    """Analysis with the block rule, also with 6 and 8 ring pairs per lane
    (at nmax 1100, 551 ring pairs: a full tile of 512 and one of 39 rings for
    K = 8, a full tile of 384 and one of 167 rings for K = 6)."""
    N = 200 if K is None else 1100
    if K is not None:
        monkeypatch.setenv("CHARM_AMD_SHA_K", str(K))
    polar.set(A1, A2, *blocks)
    g = crd.PointGridGL(N, R_EARTH)
    rng = np.random.default_rng(35)
    f = rng.standard_normal(g.shape(N))
    out = sha.point(g, f, N, GM, R_EARTH)
    Co, So = polar.orc.sha_gl(N, f, g.lat, g.w, R_EARTH, GM, R_EARTH)
    scale = max(np.abs(Co).max(), np.abs(So).max())
    assert np.abs(out.c - Co).max() <= TOL * scale
    assert np.abs(out.s - So).max() <= TOL * scale


def test_polar_device_grid_and_multi_set(polar):
    """Device-resident grid synthesis/analysis and the multi-set kernel use
    the same blocks (and never the all-zero-wave flags, which assume the
    optimisation off)."""
    N, nset = 160, 16
    grid = DeviceGrid(N, R_EARTH)
    st = Stream()
    cs = np.stack([np.stack(kaula_coeffs(N, 40 + i), axis=1) for i in range(nset)])
    ref0 = polar.orc.shs_grid(N, cs[0, :, 0], cs[0, :, 1], GM, R_EARTH, grid.pnt.lat,
                              grid.pnt.r, grid.pnt.lon, gl=True)
    out = DevArray((nset, grid.nlat, grid.nlon))
    grid.shs_multi(DevArray.from_numpy(cs), N, nset, GM, R_EARTH, out, st)
    st.synchronize()
    got = out.numpy()
    assert relmax(got[0], ref0) <= TOL
    for i in (0, nset - 1):
        f = DevArray((grid.nlat, grid.nlon))
        grid.shs(DevArray.from_numpy(cs[i]), N, GM, R_EARTH, f, st)
        st.synchronize()
        assert relmax(got[i], f.numpy()) <= TOL
    fr = f.numpy()
    co = DevArray(cs[0].shape)
    grid.sha(DevArray.from_numpy(fr), N, GM, R_EARTH, co, st)
    st.synchronize()
    Co, So = polar.orc.sha_gl(N, fr, grid.pnt.lat, grid.pnt.w, R_EARTH, GM, R_EARTH)
    got = co.numpy()
    scale = max(np.abs(Co).max(), np.abs(So).max())
    assert np.abs(got[:, 0] - Co).max() <= TOL * scale
    assert np.abs(got[:, 1] - So).max() <= TOL * scale


def test_polar_documented_accuracy(polar):
    """CHarm's documented property (docs/src/benchmarks.rst:66-76): with
    a1 = 100, a2 = 0.01 the GL round trip of U[-1, 1) coefficients
    (mu = R = r = 1, bench/bench.c:192-300) is as accurate as without, and
    as accurate as CHarm's algorithm in double (the oracle) on the same
    coefficients."""
    N = 1000
    rng = np.random.default_rng(1000)
    P = (N + 1) * (N + 2) // 2
    C = rng.uniform(-1, 1, P)
    S = rng.uniform(-1, 1, P)
    S[: N + 1] = 0.0
    g = crd.PointGridGL(N, 1.0)
    m = shc.Shc.from_arrays(N, C, S, 1.0, 1.0)

    def stats(Co, So):
        d = np.concatenate([Co - C, So - S])
        return np.abs(d).max(), np.sqrt(2.0 * np.sum(d * d) / ((N + 1) * (N + 2)))

    def eps():
        f = shs.point(g, m, N)
        out = sha.point(g, f, N, 1.0, 1.0)
        return stats(out.c, out.s)

    polar.set(100, -1.0)
    e_off = eps()
    polar.set(100, 0.01)
    e_on = eps()
    polar.set(100, -1.0)
    orc = polar.orc
    f = orc.shs_grid(N, C, S, 1.0, 1.0, g.lat, g.r, g.lon, gl=True)
    e_orc = stats(*orc.sha_gl(N, f, g.lat, g.w, 1.0, 1.0, 1.0))
    print("N=1000 round trip eps_max/eps_rms: polar off %.3e/%.3e, on %.3e/%.3e, "
          "oracle %.3e/%.3e" % (e_off + e_on + e_orc))
    assert e_on[0] <= 1.5 * e_off[0] and e_on[1] <= 1.5 * e_off[1]
    assert e_off[0] <= 1.5 * e_orc[0] and e_off[1] <= 1.5 * e_orc[1]
