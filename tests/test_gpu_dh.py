"""Driscoll--Healy grids (charm_crd_point_dh1/dh2) through the HIP path:
synthesis against CHarm's own shs_p_*_fft1_dh{1,2} fixtures (double, and the
error against the quad build) and against the oracle at larger degrees;
analysis against the oracle and as the inverse of synthesis (Driscoll--Healy
quadrature is exact to the grid's degree).  The north pole has no southern
counterpart (src/shs/shs_point_grd.c:116-130, src/sha/sha_point.c:190-192)."""
import numpy as np
import pytest

import charm_amd
from charm_amd import crd, sha, shc, shs
from charm_amd.workloads import GM, R_EARTH, kaula_coeffs
from charm_fixtures import DELTAR, THRESH, egm96_test_model, nearly_equal, read_fixture
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

TOL = 1e-12
GRIDS = {1: crd.PointGridDH1, 2: crd.PointGridDH2}


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


@pytest.mark.parametrize("dh", (1, 2))
@pytest.mark.parametrize("nmax", range(0, 5))
@pytest.mark.parametrize("deltar", (0, 1))
def test_shs_dh_fixtures(dh, nmax, deltar):
    mu, R, C, S = egm96_test_model()
    m = model(10, C, S, mu, R)
    g = GRIDS[dh](nmax, R + DELTAR * deltar)
    f = shs.point(g, m, nmax).ravel()
    name = "shs_p_nx%d_dr%d_fft1_dh%d.txt" % (nmax, deltar, dh)
    ref = read_fixture("double", name)
    assert nearly_equal(f, ref, 10 * THRESH["double"]).all(), relmax(f, ref)
    assert relmax(f, read_fixture("quad", name)) <= TOL


@pytest.mark.parametrize("dh", (1, 2))
def test_shs_dh_vs_oracle(orc, dh):
    N = 181
    C, S = kaula_coeffs(N, 50 + dh)
    g = GRIDS[dh](N, R_EARTH + 300.0)
    f = shs.point(g, model(N, C, S, GM, R_EARTH), N)
    ref = orc.shs_grid(N, C, S, GM, R_EARTH, g.lat, g.r, g.lon, gl=False)
    assert relmax(f, ref) <= TOL, relmax(f, ref)


@pytest.mark.parametrize("dh", (1, 2))
@pytest.mark.parametrize("N", (1, 64, 255))
def test_sha_dh_vs_oracle_and_roundtrip(orc, dh, N):
    C, S = kaula_coeffs(N, 60 + N)
    g = GRIDS[dh](N, R_EARTH)
    f = shs.point(g, model(N, C, S, GM, R_EARTH), N)
    out = sha.point(g, f, N, GM, R_EARTH)
    Co, So = orc.sha_dh(dh, N, f, g.lat, g.w, R_EARTH, GM, R_EARTH)
    scale = max(np.abs(Co).max(), np.abs(So).max())
    assert np.abs(out.c - Co).max() <= TOL * scale
    assert np.abs(out.s - So).max() <= TOL * scale
    assert np.abs(out.c - C).max() <= TOL * scale
    assert np.abs(out.s - S).max() <= TOL * scale


def test_sha_dh_lower_degree_and_rescale(orc):
    """Analysis to a degree below the grid's, with the grid at R + 2 km and
    the coefficients returned for R."""
    Ng, N = 120, 90
    C, S = kaula_coeffs(Ng, 70)
    R1 = R_EARTH + 2000.0
    g = crd.PointGridDH2(Ng, R1)
    f = shs.point(g, model(Ng, C, S, GM, R_EARTH), Ng)
    out = sha.point(g, f, N, GM, R_EARTH)
    Co, So = orc.sha_dh(2, N, f, g.lat, g.w, R1, GM, R_EARTH)
    scale = max(np.abs(Co).max(), np.abs(So).max())
    assert np.abs(out.c - Co).max() <= TOL * scale
    assert np.abs(out.s - So).max() <= TOL * scale
