"""Pins the oracle's global gravity forward modelling (composed from the
oracle's SHS/SHA on GL grids, src/gfm/gfm_sgfm.c) against CHarm's
gfm_sgfm_g_d{0,1,2}_nx10 fixtures -- degree amplitudes of the lunar
potential for 3D, lateral and constant density, validated in CHarm at
glob_threshold2 (tests/check_gfm_sgfm_all.c.in:165-220) -- and against the
closed form of a spherical shell."""
import math
import os

import numpy as np
import pytest

from charm_fixtures import (GOLDEN, GRAV_CONST, MOON_DENSITY_CONST, MOON_MASS, MOON_NMAX_POT,
                            MOON_PMAX, MOON_PMIN, MOON_RREF, THRESH2, degree_amplitudes,
                            nearly_equal, read_fixture, read_tbl)
from oracle.oracle import Oracle


def moon_inputs(exact=False):
    sh = read_tbl(os.path.join(GOLDEN, "MoonTopo2600p_to10-tbl.txt"), exact)
    d0 = read_tbl(os.path.join(GOLDEN, "moon-rho0_to10.tbl"), exact)
    d1 = read_tbl(os.path.join(GOLDEN, "moon-rho1_to10.tbl"), exact)
    return sh, d0, d1


@pytest.mark.parametrize("dens", (0, 1, 2))
def test_moon_fixtures(dens):
    o = Oracle("d")
    sh, d0, d1 = moon_inputs()
    density = {0: [(d0[3], d0[4], 10), (d1[3], d1[4], 10)], 1: [(d0[3], d0[4], 10)],
               2: MOON_DENSITY_CONST}[dens]
    C, S = o.gfm_global((sh[3], sh[4]), 10, MOON_RREF, density, GRAV_CONST, MOON_MASS,
                        MOON_PMIN, MOON_PMAX, MOON_NMAX_POT)
    da = degree_amplitudes(C, S, MOON_NMAX_POT)
    ref = read_fixture("double", "gfm_sgfm_g_d%d_nx10.txt" % dens)
    assert nearly_equal(da, ref, THRESH2["double"]).all(), np.max(np.abs(da / ref - 1))
    refq = read_fixture("quad", "gfm_sgfm_g_d%d_nx10.txt" % dens)
    assert np.max(np.abs(da - refq)) <= 1e-13 * np.max(refq)


def test_spherical_shell_closed_form():
    """Shape r = R + h everywhere, constant density: only degree 0, equal to
    the shell mass 4/3 pi ((R+h)^3 - R^3) rho / M; the power series ends at
    p = 3, so pmax = 3 and pmax = 5 give the same."""
    o = Oracle("d")
    R, h, rho, M = 1.0e6, 2.5e4, 2670.0, 1.0e22
    nmax = 4
    ncs = (nmax + 1) * (nmax + 2) // 2
    C = np.zeros(ncs)
    C[0] = R + h
    S = np.zeros(ncs)
    want = 4.0 / 3.0 * math.pi * ((R + h) ** 3 - R ** 3) * rho / M
    for pmax in (3, 5):
        Cp, Sp = o.gfm_global((C, S), nmax, R, rho, 6.674e-11, M, 1, pmax, 6)
        assert abs(Cp[0] - want) <= 1e-14 * want
        assert np.max(np.abs(Cp[1:])) <= 1e-15 * want and np.max(np.abs(Sp)) <= 1e-15 * want
