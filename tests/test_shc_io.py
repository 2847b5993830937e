"""Coefficient files: charm_shc_read_{gfc,tbl,dov,mtx,bin} and
charm_shc_write_{tbl,dov,mtx,bin}.

Follows tests/check_shc_read_gfc.c (static and time variable gfc against
the shc_nx10_*_gfc_* fixtures), tests/check_shc_read_type.c (tbl, dov and
mtx read the same model as the gfc file; reading the degree only) and
tests/check_shc_write_type.c (write, read back, compare).  The input files
are CHarm's data/input files EGM96-degree10{.gfc,-tbl.txt,-dov.txt,-mtx.txt}
and ITSG-Grace2018s.gfc.  The ICGEM 2.0 records, the "dot" keyword and the
epoch rules, which no CHarm fixture covers, are checked against the ICGEM
formulas evaluated in Python on files written by the test.
"""
import math
import os

import numpy as np
import pytest

import charm_amd
from charm_amd import shc
from charm_fixtures import GOLDEN, THRESH, nearly_equal, read_fixture

GFC = os.path.join(GOLDEN, "EGM96-degree10.gfc")
GFC_TVG = os.path.join(GOLDEN, "ITSG-Grace2018s.gfc")
TXT = {t: os.path.join(GOLDEN, "EGM96-degree10-%s.txt" % t) for t in ("tbl", "dov", "mtx")}
EPOCH = "20020305.1245"    # tests/parameters.h:285-286


def test_gfc_static_fixtures():
    m = shc.Shc.from_file("gfc", GFC, 10)
    assert m.nmax == 10 and m.mu == 3.986004415e14 and m.r == 6378136.3
    for cs, arr in (("c", m.c), ("s", m.s)):
        ref = read_fixture("double", "shc_nx10_%s_gfc_static.txt" % cs)
        assert nearly_equal(arr, ref, 10 * THRESH["double"]).all()
        assert np.array_equal(arr, ref)


def test_gfc_time_variable_fixtures():
    """ITSG-Grace2018s (ICGEM 1.0: gfct, trnd, acos, asin) at 2002-03-05
    12:45."""
    m = shc.Shc.from_file("gfc", GFC_TVG, 10, epoch=EPOCH)
    for cs, arr in (("c", m.c), ("s", m.s)):
        ref = read_fixture("double", "shc_nx10_%s_gfc_tvg.txt" % cs)
        assert nearly_equal(arr, ref, 10 * THRESH["double"]).all()
        refq = read_fixture("quad", "shc_nx10_%s_gfc_tvg.txt" % cs)
        assert np.max(np.abs(arr - refq)) <= 1e-15 * np.max(np.abs(refq))


@pytest.mark.parametrize("ftype", ("tbl", "dov", "mtx"))
def test_text_formats_read_the_same_model(ftype):
    ref = shc.Shc.from_file("gfc", GFC, 10)
    m = shc.Shc.from_file(ftype, TXT[ftype], 10)
    assert (m.mu, m.r) == (ref.mu, ref.r)
    assert nearly_equal(m.c, ref.c, 10 * THRESH["double"]).all()
    assert nearly_equal(m.s, ref.s, 10 * THRESH["double"]).all()


@pytest.mark.parametrize("ftype", ("gfc", "tbl", "dov", "mtx"))
def test_degree_only(ftype):
    path = GFC if ftype == "gfc" else TXT[ftype]
    assert shc.Shc.nmax_from_file(ftype, path) == 10


@pytest.mark.parametrize("ftype,ordering", [("bin", 0), ("mtx", 0), ("tbl", shc.WRITE_N),
                                            ("tbl", shc.WRITE_M), ("dov", shc.WRITE_N),
                                            ("dov", shc.WRITE_M)])
def test_write_read_round_trip(tmp_path, ftype, ordering):
    ref = shc.Shc.from_file("gfc", GFC, 10)
    path = str(tmp_path / ("model." + ftype))
    ref.to_file(ftype, path, formatting="%0.16e", ordering=ordering)
    assert shc.Shc.nmax_from_file(ftype, path) == 10
    m = shc.Shc.from_file(ftype, path)
    assert (m.nmax, m.mu, m.r) == (10, ref.mu, ref.r)
    assert np.array_equal(m.c, ref.c) and np.array_equal(m.s, ref.s)


def test_written_text_layouts(tmp_path):
    """The layouts of src/shc/shc_write_{mtdt,tbl,dov,mtx}.c."""
    c = np.arange(1.0, 7.0)
    s = np.array([0.0, 0.0, 0.0, 7.0, 8.0, 9.0])
    m = shc.Shc.from_arrays(2, c, s, 2.0, 3.0)
    p = str(tmp_path / "x")
    m.to_file("tbl", p, formatting="%g", ordering=shc.WRITE_M)
    assert open(p).read() == ("2 2 3\n0 0 1 0\n1 0 2 0\n1 1 4 7\n2 0 3 0\n2 1 5 8\n"
                              "2 2 6 9\n")
    m.to_file("dov", p, formatting="%g", ordering=shc.WRITE_N)
    assert open(p).read() == ("2 2 3\n0  0 1\n1  0 2\n2  0 3\n1  1 4\n1 -1 7\n2  1 5\n"
                              "2 -1 8\n2  2 6\n2 -2 9\n")
    m.to_file("mtx", p, formatting="%g")
    assert open(p).read() == "2 2 3\n1 7 8\n2 4 9\n3 5 6\n"


def test_partial_read_zero_fills(tmp_path):
    ref = shc.Shc.from_file("gfc", GFC, 10)
    for ftype in ("gfc", "tbl", "dov", "mtx", "bin"):
        path = GFC if ftype == "gfc" else str(tmp_path / ("m." + ftype))
        if ftype != "gfc":
            ref.to_file(ftype, path)
        big = shc.Shc.from_zeros(12)
        big.c[:] = 5.0
        big.s[:] = 5.0
        shc._read(ftype, path, 6, big._Shc, None)
        for n in range(13):
            for mm in range(n + 1):
                cv, sv = big.get_coeffs(n, mm)
                want = ref.get_coeffs(n, mm) if n <= 6 else (0.0, 0.0)
                assert (cv, sv) == want, (ftype, n, mm)


def _err_code(ftype, path, nmax, target, epoch=None):
    with pytest.raises(charm_amd.CHarmError) as e:
        shc._read(ftype, path, nmax, target._Shc if target is not None else None, epoch)
    return str(e.value)


def test_read_errors(tmp_path):
    small = shc.Shc.from_zeros(5)
    assert "only" in _err_code("tbl", TXT["tbl"], 11, shc.Shc.from_zeros(12))
    assert "cannot be smaller" in _err_code("mtx", TXT["mtx"], 8, small)
    assert "Couldn't open" in _err_code("dov", str(tmp_path / "none"), 3, small)
    bad = tmp_path / "bad.tbl"
    bad.write_text("3 1 1\n0 0 1\n1 1 2\n")   # order 1 without S
    assert "sine" in _err_code("tbl", str(bad), 3, small)
    bad.write_text("3 1 -1\n")
    with pytest.raises(charm_amd.CHarmError):
        shc.Shc.nmax_from_file("tbl", str(bad)) and shc.Shc.from_file("tbl", str(bad))


# --------------------------------------------------------------------------
# ICGEM 2.0, "dot" and epochs: the formulas of the ICGEM format evaluated here
# --------------------------------------------------------------------------
def year_fraction(date):
    """yyyyMMdd[.hhmm] -> year + (day of year - 1 + hours/24) / days in year."""
    y, mo, d = int(date[:4]), int(date[4:6]), int(date[6:8])
    h = int(date[9:11]) if len(date) > 8 else 0
    mi = int(date[11:13]) if len(date) > 8 else 0
    leap = (y % 4 == 0 and y % 100 != 0) or y % 400 == 0
    days = [31, 29 if leap else 28, 31, 30, 31, 30, 31, 31, 30, 31, 30, 31]
    doy = sum(days[:mo - 1]) + d - 1
    return y + (doy + (h + mi / 60.0) / 24.0) / (366.0 if leap else 365.0)


HEAD = """model with a time variable part
product_type gravity_field
earth_gravity_constant 3.986004415E+14
radius 6378136.3
max_degree 3
norm fully_normalized
errors %s
%s
key L M C S
end_of_head ==========
"""


def test_icgem2_intervals(tmp_path):
    """gfct/trnd/acos/asin records valid on [t0, t1) (ICGEM 2.0)."""
    body = [
        "gfc 0 0 1.0 0.0",
        "gfct 2 0 -4.8E-4 0.0 20000101 20100101",
        "trnd 2 0 1.0E-11 0.0 20000101 20100101",
        "acos 2 0 2.0E-11 0.0 20000101 20100101 1.0",
        "asin 2 0 3.0E-11 0.0 20000101 20100101 0.5",
        "gfct 2 0 -4.9E-4 0.0 20100101 20200101",
        "trnd 2 0 4.0E-11 0.0 20100101 20200101",
        "acos 2 0 5.0E-11 0.0 20100101 20200101 1.0",
        "asin 2 0 6.0E-11 0.0 20100101 20200101 1.0",
        "gfct 3 1 1.0E-7 2.0E-7 20000101 20200101",
        "dot 3 1 1.0E-10 -1.0E-10 20000101 20200101",
    ]
    path = tmp_path / "m2.gfc"
    path.write_text(HEAD % ("no", "format icgem2.0") + "\n".join(body) + "\n")
    for epoch in ("20050617.0630", "20120229", "20191231.2300", "20100101"):
        m = shc.Shc.from_file("gfc", str(path), epoch=epoch)
        t = year_fraction(epoch)
        if t < year_fraction("20100101"):
            t0 = year_fraction("20000101")
            c20 = (-4.8e-4 + 1e-11 * (t - t0) + 2e-11 * math.cos(2 * math.pi * (t - t0))
                   + 3e-11 * math.sin(2 * math.pi / 0.5 * (t - t0)))
        else:
            t0 = year_fraction("20100101")
            c20 = (-4.9e-4 + 4e-11 * (t - t0) + 5e-11 * math.cos(2 * math.pi * (t - t0))
                   + 6e-11 * math.sin(2 * math.pi * (t - t0)))
        dt31 = t - year_fraction("20000101")
        c31, s31 = 1e-7 + 1e-10 * dt31, 2e-7 - 1e-10 * dt31
        got = m.get_coeffs(2, 0), m.get_coeffs(3, 1)
        assert abs(got[0][0] - c20) <= 1e-16 * abs(c20), epoch
        assert abs(got[1][0] - c31) <= 1e-16 * abs(c31) and abs(got[1][1] - s31) <= 1e-16 * abs(s31)
    # 2019-12-31 24:00 is 2020-01-01, outside [t0, t1) of every record
    m = shc.Shc.from_file("gfc", str(path), epoch="20191231.2400")
    assert m.get_coeffs(2, 0) == (0.0, 0.0) and m.get_coeffs(0, 0) == (1.0, 0.0)
    # ICGEM 2.0 needs an epoch
    with pytest.raises(charm_amd.CHarmError, match="epoch"):
        shc.Shc.from_file("gfc", str(path))


def test_icgem1_default_epoch_and_errors_columns(tmp_path):
    """ICGEM 1.0 with formal errors; no epoch: each record's own t0, so the
    trend and the periodic terms add dt = 0 (cos 0 = 1, sin 0 = 0)."""
    body = ["gfc 0 0 1.0 0.0 0.0 0.0",
            "gfct 2 1 1.0E-9 2.0E-9 1E-12 1E-12 20040101",
            "trnd 2 1 1.0E-11 2.0E-11 1E-12 1E-12",
            "acos 2 1 3.0E-11 4.0E-11 1E-12 1E-12 1.0",
            "asin 2 1 5.0E-11 6.0E-11 1E-12 1E-12 1.0"]
    path = tmp_path / "m1.gfc"
    path.write_text(HEAD % ("formal", "") + "\n".join(body) + "\n")
    m = shc.Shc.from_file("gfc", str(path))
    c, s = m.get_coeffs(2, 1)
    assert c == 1.0e-9 + 3.0e-11 and s == 2.0e-9 + 4.0e-11
    m = shc.Shc.from_file("gfc", str(path), epoch="20060701")
    dt = year_fraction("20060701") - year_fraction("20040101")
    want = 1e-9 + 1e-11 * dt + 3e-11 * math.cos(2 * math.pi * dt) + 5e-11 * math.sin(2 * math.pi * dt)
    assert abs(m.get_coeffs(2, 1)[0] - want) <= 2e-16 * abs(want)


@pytest.mark.parametrize("epoch", ["2002030", "20021305", "20020230", "20030229",
                                   "20020305.2460", "20020305.1261", "20020305-1200",
                                   "-20020305"])
def test_invalid_epochs(epoch):
    with pytest.raises(charm_amd.CHarmError):
        shc.Shc.from_file("gfc", GFC_TVG, 10, epoch=epoch)


@pytest.mark.parametrize("epoch", ["20000229", "20020305.2400", "20020306.0000"])
def test_valid_epochs(epoch):
    m = shc.Shc.from_file("gfc", GFC_TVG, 4, epoch=epoch)
    assert m.nmax == 4


def test_large_model_round_trip(tmp_path):
    """A degree-600 model (180k lines) through every text format: the
    parallel parse keeps file order and exact values."""
    from charm_amd.workloads import kaula_coeffs

    C, S = kaula_coeffs(600, 9)
    m = shc.Shc.from_arrays(600, C, S, 3.986004415e14, 6378136.3)
    for ftype in ("tbl", "dov", "mtx", "bin"):
        path = str(tmp_path / ("big." + ftype))
        m.to_file(ftype, path, ordering=shc.WRITE_M)
        x = shc.Shc.from_file(ftype, path)
        assert np.array_equal(x.c, C) and np.array_equal(x.s, S), ftype


def test_bin_reads_columns_and_survives_corrupt_headers(tmp_path):
    """charm_shc_read_bin reads order by order (shc_read_bin.c:203-256): a
    low-degree read of a file whose header claims a huge degree touches only
    what it needs, a truncated file reports CHARM_EFILEIO with CHarm's
    message, and nothing escapes the C ABI as an exception."""
    import struct

    from charm_amd.workloads import kaula_coeffs

    C, S = kaula_coeffs(40, 3)
    m = shc.Shc.from_arrays(40, C, S, 2.0, 3.0)
    path = str(tmp_path / "m.bin")
    m.to_file("bin", path)
    low = shc.Shc.from_file("bin", path, 7)
    ref = shc.Shc.from_file("bin", path)
    for n in range(8):
        for mm in range(n + 1):
            assert low.get_coeffs(n, mm) == ref.get_coeffs(n, mm)
    # header claiming degree 2^40 with 41 degrees of data behind it
    raw = open(path, "rb").read()
    huge = tmp_path / "huge.bin"
    huge.write_bytes(struct.pack("<Q", 2 ** 40) + raw[8:])
    assert shc.Shc.nmax_from_file("bin", str(huge)) == 2 ** 40
    msg = _err_code("bin", str(huge), 5, shc.Shc.from_zeros(5))
    assert "Failed to read the \"C\" coefficients" in msg
    # truncated inside the S table
    cut = tmp_path / "cut.bin"
    cut.write_bytes(raw[:-100])
    assert "Failed to read the \"S\" coefficients" in _err_code("bin", str(cut), 40,
                                                                shc.Shc.from_zeros(40))
    short = tmp_path / "short.bin"
    short.write_bytes(raw[:12])
    assert "scaling parameter" in _err_code("bin", str(short), 3, shc.Shc.from_zeros(3))
