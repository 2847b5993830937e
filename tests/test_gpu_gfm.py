"""GPU parity of the global spectral gravity forward modelling
(charm_gfm_global_density_{3d,lateral,const}): CHarm's lunar fixtures
gfm_sgfm_g_d{0,1,2}_nx10 (degree amplitudes, CHarm's tolerance
glob_threshold2, and against CHarm in quad precision), the oracle's
coefficients (1e-12 of the largest), a spherical shell in closed form, and
the exported per-power coefficient files."""
import math
import os

import numpy as np
import pytest

import charm_amd
from charm_amd import gfm, shc
from charm_fixtures import (GOLDEN, GRAV_CONST, MOON_DENSITY_CONST, MOON_MASS, MOON_NMAX_POT,
                            MOON_PMAX, MOON_PMIN, MOON_RREF, THRESH2, degree_amplitudes,
                            nearly_equal, read_fixture)
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

TOL = 1e-12


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if charm_amd.lib().charm_amd_device_count() < 1:
        pytest.fail("no HIP device visible: the GPU tests need an MI355X")


def moon():
    shape = shc.Shc.from_file("tbl", os.path.join(GOLDEN, "MoonTopo2600p_to10-tbl.txt"))
    d0 = shc.Shc.from_file("tbl", os.path.join(GOLDEN, "moon-rho0_to10.tbl"))
    d1 = shc.Shc.from_file("tbl", os.path.join(GOLDEN, "moon-rho1_to10.tbl"))
    return shape, d0, d1


def run(dens, shape, d0, d1, **kw):
    common = (GRAV_CONST, MOON_MASS, MOON_PMIN, MOON_PMAX, MOON_NMAX_POT)
    if dens == 0:
        return gfm.global_density_3d(shape, 10, MOON_RREF, [d0, d1], [10, 10], 1, *common, **kw)
    if dens == 1:
        return gfm.global_density_lateral(shape, 10, MOON_RREF, d0, 10, *common, **kw)
    return gfm.global_density_const(shape, 10, MOON_RREF, MOON_DENSITY_CONST, *common, **kw)


@pytest.mark.parametrize("dens", (0, 1, 2))
def test_moon_fixtures_and_oracle(dens):
    shape, d0, d1 = moon()
    pot = run(dens, shape, d0, d1)
    assert pot.mu == GRAV_CONST * MOON_MASS and pot.r == MOON_RREF
    da = degree_amplitudes(pot.c, pot.s, MOON_NMAX_POT)
    ref = read_fixture("double", "gfm_sgfm_g_d%d_nx10.txt" % dens)
    assert nearly_equal(da, ref, THRESH2["double"]).all(), np.max(np.abs(da / ref - 1))
    refq = read_fixture("quad", "gfm_sgfm_g_d%d_nx10.txt" % dens)
    assert np.max(np.abs(da - refq)) <= 1e-13 * np.max(refq)
    density = {0: [(d0.c, d0.s, 10), (d1.c, d1.s, 10)], 1: [(d0.c, d0.s, 10)],
               2: MOON_DENSITY_CONST}[dens]
    C, S = Oracle("d").gfm_global((shape.c, shape.s), 10, MOON_RREF, density, GRAV_CONST,
                                  MOON_MASS, MOON_PMIN, MOON_PMAX, MOON_NMAX_POT)
    scale = np.max(np.abs(C))
    assert np.max(np.abs(pot.c - C)) <= TOL * scale and np.max(np.abs(pot.s - S)) <= TOL * scale


def test_spherical_shell_closed_form():
    R, h, rho, M = 1.0e6, 2.5e4, 2670.0, 1.0e22
    shape = shc.Shc.from_zeros(6)
    shape.c[0] = R + h
    want = 4.0 / 3.0 * math.pi * ((R + h) ** 3 - R ** 3) * rho / M
    for pmax in (3, 6):
        pot = gfm.global_density_const(shape, 6, R, rho, 6.674e-11, M, 1, pmax, 12)
        assert abs(pot.c[0] - want) <= 1e-14 * want
        assert np.max(np.abs(pot.c[1:])) <= 1e-15 * want and np.max(np.abs(pot.s)) <= 1e-15 * want


def test_degree_90_against_oracle():
    """A degree-90 random shape, powers 1..4: GL grids up to degree 360."""
    from charm_amd.workloads import kaula_coeffs

    C, S = kaula_coeffs(90, 3)
    C = C * 2.0e5
    S = S * 2.0e5
    C[0] = 1.74e6
    shape = shc.Shc.from_arrays(90, C, S)
    pot = gfm.global_density_const(shape, 90, 1.73e6, 2500.0, GRAV_CONST, MOON_MASS, 1, 4, 200)
    Co, So = Oracle("d").gfm_global((C, S), 90, 1.73e6, 2500.0, GRAV_CONST, MOON_MASS, 1, 4, 200)
    scale = np.max(np.abs(Co))
    assert np.max(np.abs(pot.c - Co)) <= TOL * scale and np.max(np.abs(pot.s - So)) <= TOL * scale


def test_exported_powers_sum_to_the_potential(tmp_path):
    shape, d0, d1 = moon()
    hro = str(tmp_path / "hro")
    pp = str(tmp_path / "pot")
    pot = run(0, shape, d0, d1, shape_density_shcs_path=hro, potential_shcs_path=pp,
              shcs_file_format="tbl")
    total_c = np.zeros_like(pot.c)
    total_s = np.zeros_like(pot.s)
    for p in range(MOON_PMIN, MOON_PMAX + 1):
        for i in (0, 1):
            name = "%s-p%d-i%d.txt" % (pp, p, i)
            assert os.path.exists("%s-p%d-i%d.txt" % (hro, p, i))
            x = shc.Shc.from_file("tbl", name)
            assert x.mu == GRAV_CONST * MOON_MASS and x.r == MOON_RREF
            n = min(x.nmax, pot.nmax)
            for m in range(n + 1):
                for d in range(m, n + 1):
                    a = shc.index(pot.nmax, d, m)
                    b = shc.index(x.nmax, d, m)
                    total_c[a] += x.c[b]
                    total_s[a] += x.s[b]
    scale = np.max(np.abs(pot.c))
    assert np.max(np.abs(total_c - pot.c)) <= 1e-14 * scale
    assert np.max(np.abs(total_s - pot.s)) <= 1e-14 * scale
