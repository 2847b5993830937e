"""Pins the CPU oracle against CHarm's own golden fixtures (double and quad).

Tolerances are the ones CHarm's own test-suite applies to the same files
(tests/check_crd_point_quad.c:59-76, tests/check_shs_point_all.c:511-517,
730-736, 890-900): misc_is_nearly_equal with 10*glob_threshold, and
10*glob_threshold2 for the dsun files.
"""
import numpy as np
import pytest

from charm_fixtures import (BREAK_SYMM, DELTAR, THRESH, THRESH2, custom_grid,
                            egm96_test_model, nearly_equal, nearly_equal_exact, read_fixture,
                            scattered_points, truncate)
from oracle.oracle import Oracle

PRECS = [("double", "d"), ("quad", "q")]


@pytest.fixture(scope="module")
def orc():
    return {"d": Oracle("d"), "q": Oracle("q")}


def _check(prec, got, ref, eps):
    if prec == "double":
        ok = nearly_equal(got, ref, eps)
    else:
        ok = nearly_equal_exact(got, ref, eps)
    assert ok.all(), (np.flatnonzero(~ok)[:5], got[:3], ref[:3])


def _radius(prec, base, deltar):
    if prec == "double":
        return base + DELTAR * deltar
    import mpmath

    with mpmath.workprec(113):
        return +(mpmath.mpf(base) + DELTAR * deltar)


@pytest.mark.parametrize("prec,p", PRECS)
@pytest.mark.parametrize("nmax", range(0, 11))
def test_gl_grid_fixtures(orc, prec, p, nmax):
    exact = p == "q"
    lat, w, lon = orc[p].gl_grid(nmax, exact=exact)
    eps = 10 * THRESH[prec]
    for d in (0, 1):
        _check(prec, lat, read_fixture(prec, "crd_nx%d_dr%d_gl_lat.txt" % (nmax, d), exact), eps)
        _check(prec, w, read_fixture(prec, "crd_nx%d_dr%d_gl_w.txt" % (nmax, d), exact), eps)
        _check(prec, lon, read_fixture(prec, "crd_nx%d_dr%d_gl_lon.txt" % (nmax, d), exact), eps)
        r = read_fixture(prec, "crd_nx%d_dr%d_gl_r.txt" % (nmax, d), exact)
        assert all(float(v) == 1.0 + DELTAR * d for v in r)


@pytest.mark.parametrize("prec,p", PRECS)
@pytest.mark.parametrize("nmax", range(0, 11))
@pytest.mark.parametrize("dh", (1, 2))
def test_dh_grid_fixtures(orc, prec, p, nmax, dh):
    exact = p == "q"
    lat, w, lon = orc[p].dh_grid(nmax, dh, exact=exact)
    eps = 10 * THRESH[prec]
    for d in (0, 1):
        name = "crd_nx%d_dr%d_dh%d_%%s.txt" % (nmax, d, dh)
        _check(prec, lat, read_fixture(prec, name % "lat", exact), eps)
        _check(prec, w, read_fixture(prec, name % "w", exact), eps)
        _check(prec, lon, read_fixture(prec, name % "lon", exact), eps)
        assert all(float(v) == 1.0 + DELTAR * d for v in read_fixture(prec, name % "r", exact))


@pytest.mark.parametrize("prec,p", PRECS)
@pytest.mark.parametrize("nmax", range(0, 5))
@pytest.mark.parametrize("deltar", (0, 1))
@pytest.mark.parametrize("dh", (1, 2))
def test_shs_dh_fixtures(orc, prec, p, nmax, deltar, dh):
    """Synthesis on Driscoll-Healy grids (CHarm's shs_p_*_fft1_dh{1,2} files);
    the oracle evaluates every ring directly (no south pole, so no mirror)."""
    exact = p == "q"
    mu, R, C, S = egm96_test_model(exact)
    lat, w, lon = orc[p].dh_grid(nmax, dh, exact=exact)
    rr = _radius(prec, R, deltar)
    r = np.array([rr] * len(lat), dtype=object if exact else np.float64)
    f = orc[p].shs_grid(nmax, C, S, mu, R, lat, r, lon, gl=False, exact=exact)
    ref = read_fixture(prec, "shs_p_nx%d_dr%d_fft1_dh%d.txt" % (nmax, deltar, dh), exact)
    _check(prec, f.ravel(), ref, 10 * THRESH[prec])


def _custom_cases():
    out = []
    for nmax in range(0, 5):
        for i in range(4):
            for fft in (0, 1):
                for s in (0, 1):
                    for deltar in (0, 1):
                        name = "shs_p_nx%d_n%d_dr%d_fft%d_s%d.txt" % (nmax, i, deltar, fft, s)
                        import os
                        from charm_fixtures import DATA
                        if os.path.exists(os.path.join(DATA, "double", name)):
                            out.append((nmax, i, fft, s, deltar, name))
    return out


@pytest.mark.parametrize("prec,p", PRECS)
@pytest.mark.parametrize("nmax,i,fft,s,deltar,name", _custom_cases())
def test_shs_custom_grid_fixtures(orc, prec, p, nmax, i, fft, s, deltar, name):
    """Custom grids of CHarm's point-synthesis test, with (fft1) and without
    (fft0, longitudes over half the circle) the FFT along parallels,
    symmetric (s1) and not (s0)."""
    exact = p == "q"
    mu, R, C, S = egm96_test_model(exact)
    lat, lon = custom_grid(i, fft, s == 1, prec)
    rr = _radius(prec, R, deltar)
    r = np.array([rr] * len(lat), dtype=object if exact else np.float64)
    f = orc[p].shs_grid(nmax, C, S, mu, R, lat, r, lon, gl=False, exact=exact)
    ref = read_fixture(prec, name, exact)
    _check(prec, f.ravel(), ref, 100 * THRESH[prec])


@pytest.mark.parametrize("nmax", (0, 1, 6, 25))
@pytest.mark.parametrize("dh", (1, 2))
def test_sha_dh_roundtrip(orc, nmax, dh):
    """Driscoll-Healy quadrature is exact to degree nmax: the analysis of a
    synthesis returns the coefficients (as CHarm's check_sha_point.c checks)."""
    rng = np.random.default_rng(100 + nmax)
    ncs = (nmax + 1) * (nmax + 2) // 2
    C, S = rng.standard_normal(ncs), rng.standard_normal(ncs)
    S[:nmax + 1] = 0.0   # S_n0
    mu, R = 3.986004415e14, 6378136.3
    lat, w, lon = orc["d"].dh_grid(nmax, dh)
    r = np.full(len(lat), R)
    f = orc["d"].shs_grid(nmax, C, S, mu, R, lat, r, lon, gl=False)
    Co, So = orc["d"].sha_dh(dh, nmax, f, lat, w, R, mu, R)
    assert np.abs(Co - C).max() <= 1e-12 * np.abs(C).max()
    assert np.abs(So - S).max() <= 1e-12 * np.abs(C).max()


@pytest.mark.parametrize("prec,p", PRECS)
@pytest.mark.parametrize("nmax", range(0, 5))
@pytest.mark.parametrize("deltar", (0, 1))
def test_shs_gl_fixtures(orc, prec, p, nmax, deltar):
    exact = p == "q"
    mu, R, C, S = egm96_test_model(exact)
    lat, w, lon = orc[p].gl_grid(nmax, exact=exact)
    rr = _radius(prec, R, deltar)
    r = np.array([rr] * len(lat), dtype=object if exact else np.float64)
    f = orc[p].shs_grid(nmax, C, S, mu, R, lat, r, lon, gl=True, exact=exact)
    ref = read_fixture(prec, "shs_p_nx%d_dr%d_fft1_gl.txt" % (nmax, deltar), exact)
    _check(prec, f.ravel(), ref, 10 * THRESH[prec])


@pytest.mark.parametrize("prec,p", PRECS)
@pytest.mark.parametrize("nmax", range(0, 5))
@pytest.mark.parametrize("i,npt", list(enumerate([1, 2, 3, 4, 5, 6, 7, 8, 9, 20, 30, 31])))
def test_shs_sctr_fixtures(orc, prec, p, nmax, i, npt):
    exact = p == "q"
    mu, R, C, S = egm96_test_model(exact)
    lat, lon = scattered_points(npt, prec)
    for deltar in (0, 1):
        rr = _radius(prec, R, deltar)
        r = np.array([rr] * npt, dtype=object if exact else np.float64)
        f = orc[p].shs_sctr(nmax, C, S, mu, R, lat, lon, r, exact=exact)
        ref = read_fixture(prec, "shs_p_nx%d_n%d_dr%d_sctr.txt" % (nmax, i, deltar), exact)
        _check(prec, f, ref, 10 * THRESH[prec])


def dsun_inputs(nmax, exact):
    """tests/check_shs_point_all.c:768-810: C_nm = S_nm = 1/(n+1) (S_n0 = 0),
    mu = R = 1."""
    ncs = (nmax + 1) * (nmax + 2) // 2
    C = np.zeros(ncs, dtype=object if exact else np.float64)
    S = np.zeros(ncs, dtype=object if exact else np.float64)
    k = 0
    if exact:
        import mpmath
        ctx = mpmath.workprec(113)
        ctx.__enter__()
    for m in range(nmax + 1):
        for n in range(m, nmax + 1):
            if exact:
                v = mpmath.mpf(1) / (n + 1)
            else:
                v = 1.0 / float(n + 1)
            C[k] = v
            S[k] = v if m > 0 else 0
            k += 1
    if exact:
        ctx.__exit__(None, None, None)
    return C, S


@pytest.mark.parametrize("prec,p", PRECS)
@pytest.mark.parametrize("nmax", range(150, 155))
@pytest.mark.parametrize("s", (0, 1))
def test_shs_dsun_fixtures(orc, prec, p, nmax, s):
    """GL latitudes with one longitude (lon = 0): f = mu/R sum_m a_m, the
    lumped sums themselves.  s = 0 shifts lat[0] by BREAK_SYMM while the grid
    stays a GL grid, so the mirror ring reuses the shifted latitude."""
    exact = p == "q"
    C, S = dsun_inputs(nmax, exact)
    lat, w, lon = orc[p].gl_grid(nmax, exact=exact)
    if s == 0:
        if exact:
            import mpmath

            with mpmath.workprec(113):
                lat[0] = +(lat[0] - mpmath.mpf(BREAK_SYMM))
        else:
            lat[0] -= BREAK_SYMM
    lon1 = np.zeros(1, dtype=object if exact else np.float64)
    for deltar in (0, 1):
        r = np.array([1.0 + DELTAR * deltar] * len(lat), dtype=object if exact else np.float64)
        f = orc[p].shs_grid(nmax, C, S, 1.0, 1.0, lat, r, lon1, gl=True, exact=exact)
        ref = read_fixture(prec, "shs_p_nx%d_n0_dr%d_fft0_s%d_dsun.txt" % (nmax, deltar, s), exact)
        _check(prec, f.ravel(), ref, 10 * THRESH2[prec])
        # the oracle meets a much tighter bound than the reference's own test
        _check(prec, f.ravel(), ref, 1e3 * THRESH[prec] if prec == "double" else 1e-28)
