"""The oracle's polar optimisation, pinned without CHarm's output (no CHarm
test or fixture covers it).

CHarm skips order m for a whole block of rings (points) when every member,
cos(lat) = u, passes thr < m - nmax u, thr = max(a1, a2 nmax)
(src/misc/misc_polar_optimization_threshold.c:16-20,
misc_polar_optimization_apply.c:16-41).  Blocks are consecutive rings of the
loop over the rings evaluated directly -- the northern half with the equator
on symmetric grids, all rings otherwise (src/shs/shs_point_grd.c:116-197,
594-600, 860-862; src/sha/sha_point.c:716-718) -- or consecutive input points
(src/shs/shs_point_sctr.c:291-320, 518-520).

Two pins:
* identities: synthesis with the optimisation equals synthesis without it
  after zeroing, per row, the lumped coefficients of the orders its block
  skips (the block rule written out here in numpy); analysis with it equals
  the sum over ring blocks of the analyses of each block's rows alone, each
  block contributing only the orders it keeps;
* CHarm's documented property: with a1 = 100, a2 = 0.01 the Gauss-Legendre
  round trip of U[-1, 1) coefficients is as accurate as without the
  optimisation (docs/src/benchmarks.rst:66-76, protocol of
  bench/bench.c:192-300).
"""
import numpy as np
import pytest

from oracle.oracle import Oracle


def first_skip(nmax, thr, umax):
    """Smallest m with thr < m - nmax * umax (product and difference rounded
    separately), nmax + 1 if none."""
    m = np.arange(nmax + 1, dtype=np.float64)
    ok = thr < (m - np.float64(nmax) * np.float64(umax))
    return int(np.argmax(ok)) if ok.any() else nmax + 1


def row_first_skip(kind, lat, bs, nmax, thr):
    """First skipped order of every row of a grid: kind 'gl', 'dh' or
    'custom' (symmetric if lat[i] = -lat[n-1-i])."""
    nlat = len(lat)
    n = nlat + 1 if kind == "dh" else nlat
    if kind == "dh":
        symm = True
    elif nlat == 1:
        symm = False
    elif kind == "gl":
        symm = True
    else:
        symm = bool(np.allclose(lat, -lat[::-1], rtol=0, atol=1e-14))
    ndo = (n + 1 - (n % 2 == 0)) // 2 if symm else n
    m0 = np.full(nlat, -1)
    u = np.cos(np.asarray(lat, dtype=np.float64))
    for b0 in range(0, ndo, bs):
        v = first_skip(nmax, thr, u[b0:min(ndo, b0 + bs)].max())
        for i in range(b0, min(ndo, b0 + bs)):
            m0[i] = v
            if symm and n - 1 - i < nlat:
                m0[n - 1 - i] = v
    assert (m0 >= 0).all()
    return m0


def synth_from_lumped(orc, N, C, S, mu, R, lat, r, lon, m0):
    """Rows of mu/R sum_m (a_m cos m lon + b_m sin m lon) with the lumped
    coefficients of the unoptimised oracle, orders m >= m0[row] dropped."""
    orc.set_polar(100, -1.0)
    f = np.zeros((len(lat), len(lon)))
    m = np.arange(N + 1)
    for i in range(len(lat)):
        a, b = orc.lumped(N, C, S, R, lat[i], r[i])
        keep = m < m0[i]
        a, b = a * keep, b * keep
        ml = np.outer(m, lon)
        f[i] = mu / R * (a @ np.cos(ml) + b @ np.sin(ml))
    return f


def coeffs(N, seed):
    rng = np.random.default_rng(seed)
    ncs = (N + 1) * (N + 2) // 2
    C = rng.uniform(-1, 1, ncs)
    S = rng.uniform(-1, 1, ncs)
    S[: N + 1] = 0.0   # S_n0
    return C, S


@pytest.fixture
def orc():
    o = Oracle("d")
    yield o
    o.set_polar(100, -1.0)


THR = 10.0   # a1 = 10, a2 = 0


@pytest.mark.parametrize("N,bs", [(60, 8), (61, 8), (120, 32), (47, 5)])
def test_grid_synthesis_gl(orc, N, bs):
    C, S = coeffs(N, N + bs)
    lat, w, lon = orc.gl_grid(N)
    r = np.ones_like(lat)
    orc.set_polar(10, 0.0, bs, 64)
    f = orc.shs_grid(N, C, S, 1.0, 1.0, lat, r, lon, gl=True)
    m0 = row_first_skip("gl", lat, bs, N, THR)
    assert (m0 <= N).any() and (m0 > N).any()   # skipped somewhere, not everywhere
    ref = synth_from_lumped(orc, N, C, S, 1.0, 1.0, lat, r, lon, m0)
    assert np.abs(f - ref).max() <= 1e-13 * np.abs(ref).max()
    # the block rule differs from a per-ring test on these grids
    m0_ring = row_first_skip("gl", lat, 1, N, THR)
    if bs > 1:
        assert (m0_ring != m0).any()


@pytest.mark.parametrize("dh", [1, 2])
def test_grid_synthesis_dh(orc, dh):
    N, bs = 40, 8
    C, S = coeffs(N, 7 + dh)
    lat, w, lon = orc.dh_grid(N, dh)
    r = np.ones_like(lat)
    orc.set_polar(10, 0.0, bs, 64)
    f = orc.shs_grid(N, C, S, 1.0, 1.0, lat, r, lon, gl=False, dh=dh)
    m0 = row_first_skip("dh", lat, bs, N, THR)
    assert (m0 <= N).any()
    ref = synth_from_lumped(orc, N, C, S, 1.0, 1.0, lat, r, lon, m0)
    assert np.abs(f - ref).max() <= 1e-13 * np.abs(ref).max()


@pytest.mark.parametrize("symmetric", [True, False])
def test_grid_synthesis_custom(orc, symmetric):
    N, bs = 50, 4
    C, S = coeffs(N, 11)
    if symmetric:
        north = np.linspace(np.pi / 2, 0.05, 23)
        lat = np.concatenate([north, [0.0], -north[::-1]])
    else:
        lat = np.linspace(np.pi / 2 - 0.01, -np.pi / 2 + 0.3, 37)
    lon = np.linspace(0, 2 * np.pi, 2 * N + 2, endpoint=False)
    r = np.ones_like(lat)
    orc.set_polar(10, 0.0, bs, 64)
    f = orc.shs_grid(N, C, S, 1.0, 1.0, lat, r, lon, gl=False)
    m0 = row_first_skip("custom", lat, bs, N, THR)
    assert (m0 <= N).any()
    ref = synth_from_lumped(orc, N, C, S, 1.0, 1.0, lat, r, lon, m0)
    assert np.abs(f - ref).max() <= 1e-13 * np.abs(ref).max()


@pytest.mark.parametrize("bs", [1, 8, 32])
def test_scattered_synthesis(orc, bs):
    N = 60
    C, S = coeffs(N, 21)
    rng = np.random.default_rng(22)
    n = 150
    # the first 96 points in the polar caps, so that whole blocks pass
    cap = np.arcsin(rng.uniform(0.8, 1.0, 96)) * rng.choice([-1.0, 1.0], 96)
    lat = np.concatenate([cap, np.arcsin(rng.uniform(-1, 1, n - 96))])
    lon = rng.uniform(0, 2 * np.pi, n)
    r = 1.0 + rng.uniform(0, 0.01, n)
    orc.set_polar(10, 0.0, bs, 64)
    f = orc.shs_sctr(N, C, S, 1.0, 1.0, lat, lon, r)
    u = np.cos(lat)
    orc.set_polar(100, -1.0)
    ref = np.zeros(n)
    m = np.arange(N + 1)
    nskip = 0
    for i in range(n):
        b0 = (i // bs) * bs
        m0 = first_skip(N, THR, u[b0:b0 + bs].max())
        nskip += N + 1 - m0
        a, b = orc.lumped(N, C, S, 1.0, lat[i], r[i])
        keep = m < m0
        ref[i] = (a * keep) @ np.cos(m * lon[i]) + (b * keep) @ np.sin(m * lon[i])
    assert nskip > 0
    assert np.abs(f - ref).max() <= 1e-13 * np.abs(ref).max()


@pytest.mark.parametrize("kind,bs,N", [("gl", 8, 48), ("gl", 64, 200), ("dh", 8, 48)])
def test_analysis_block_sum(orc, kind, bs, N):
    """SHA with the optimisation = sum over ring blocks of the SHA of each
    block's rows alone (others zero), each block keeping orders m < m0."""
    if kind == "gl":
        lat, w, lon = orc.gl_grid(N)
        n = len(lat)
        sha = lambda f: orc.sha_gl(N, f, lat, w, 1.0, 1.0, 1.0)
    else:
        lat, w, lon = orc.dh_grid(N, 1)
        n = len(lat) + 1
        sha = lambda f: orc.sha_dh(1, N, f, lat, w, 1.0, 1.0, 1.0)
    rng = np.random.default_rng(5)
    f = rng.standard_normal((len(lat), len(lon)))
    orc.set_polar(10, 0.0, 32, bs)
    Con, Son = sha(f)
    orc.set_polar(100, -1.0)
    m0 = row_first_skip(kind, lat, bs, N, THR)
    ndo = (n + 1 - (n % 2 == 0)) // 2
    mm = np.concatenate([np.full(N + 1 - m, m) for m in range(N + 1)])   # order of each entry
    Cref = np.zeros_like(Con)
    Sref = np.zeros_like(Son)
    for b0 in range(0, ndo, bs):
        rows = [i for i in range(len(lat)) if min(i, n - 1 - i) // bs == b0 // bs
                and min(i, n - 1 - i) < ndo]
        fb = np.zeros_like(f)
        fb[rows] = f[rows]
        Cb, Sb = sha(fb)
        keep = mm < m0[b0]
        Cref += Cb * keep
        Sref += Sb * keep
    assert (m0 <= N).any()
    scale = max(np.abs(Cref).max(), np.abs(Sref).max())
    assert np.abs(Con - Cref).max() <= 1e-13 * scale
    assert np.abs(Son - Sref).max() <= 1e-13 * scale


def test_documented_accuracy_property(orc):
    """a1 = 100, a2 = 0.01: the round-trip errors equal those without the
    optimisation (docs/src/benchmarks.rst:66-76), while orders are skipped."""
    N = 300
    C, S = coeffs(N, 300)
    lat, w, lon = orc.gl_grid(N)
    r = np.ones_like(lat)

    def roundtrip():
        f = orc.shs_grid(N, C, S, 1.0, 1.0, lat, r, lon, gl=True)
        Co, So = orc.sha_gl(N, f, lat, w, 1.0, 1.0, 1.0)
        d = np.concatenate([Co - C, So - S])
        return np.abs(d).max(), np.sqrt(2.0 * np.sum(d * d) / ((N + 1) * (N + 2)))

    e_off = roundtrip()
    orc.set_polar(100, 0.01)
    e_on = roundtrip()
    m0 = row_first_skip("gl", lat, 32, N, 100.0)
    assert (m0 <= N).sum() > len(lat) // 4
    assert e_on[0] <= 1.5 * e_off[0] and e_on[1] <= 1.5 * e_off[1], (e_on, e_off)
    assert e_off[0] < 1e-12
    # an aggressive threshold is not harmless (the test can fail)
    orc.set_polar(0, 0.0)
    e_bad = roundtrip()
    assert e_bad[0] > 1e3 * e_off[0]
