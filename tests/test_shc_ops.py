"""Coefficient utilities of the C ABI (charm_shc_copy, add/sub/mul/div, the
degree- and order-wise products, dv/da/ddv/dda, rescale) against CHarm's own
tests and golden outputs:

* tests/check_shc_arithmetics.c:22-306 (files shc_{add,sub,mul,div}_*),
* tests/check_shc_arithmetics_wise.c:22-221 (shc_{mul,div}_{degree,order}_wise_*),
* tests/check_shc_dav.c / check_shc_ddav.c (shc_nx4_d{v,a}, zeros for ddv),
* tests/check_shc_rescale.c:20-79 (shc_rescale_{c,s}),
* tests/check_shc_copy.c:20-234.

Beyond the fixtures, random models of higher degree are checked against
numpy expressions of the documented definitions (src/shc/shc.h.in:1320-1722).
Host code only: no GPU is used.
"""
import ctypes as ct
import os

import numpy as np
import pytest

from charm_amd import CHarmError, shc
from charm_amd._err import CHARM_EFUNCARG
from charm_fixtures import GOLDEN, THRESH, nearly_equal, read_fixture

EPS10 = 10 * THRESH["double"]


def fixture_ok(x, name, eps=EPS10):
    ref = read_fixture("double", name)
    assert x.shape == ref.shape
    ok = nearly_equal(x, ref, eps)
    assert ok.all(), (name, x[~ok], ref[~ok])


def filled(nmax, cval, sval, mu=1.0, r=1.0):
    """calloc'ed model with every C = cval and every S_{n,m>0} = sval."""
    out = shc.Shc.from_zeros(nmax, mu, r)
    n, m = out.get_degrees_orders()
    out.c[:] = cval
    out.s[:] = np.where(m > 0, sval, 0.0)
    return out


def random_model(nmax, seed, mu=3.986004415e14, r=6378136.3):
    rng = np.random.default_rng(seed)
    out = shc.Shc.from_zeros(nmax, mu, r)
    n, m = out.get_degrees_orders()
    out.c[:] = rng.standard_normal(out.c.size)
    out.s[:] = np.where(m > 0, rng.standard_normal(out.s.size), 0.0)
    return out


# ---------------------------------------------------------------------------
# add, sub, mul, div: tests/check_shc_arithmetics.c
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("op", ["add", "sub", "mul", "div"])
def test_arithmetics_golden(op):
    # rop of degree 4, op1 of degree 4 (C = 1, S = 2), op2 of degree 3 (5 for
    # div) with C = 10, S = 20; all with mu = r = 1
    rop = shc.Shc.from_zeros(4)
    op1 = filled(4, 1.0, 2.0)
    op2 = filled(5 if op == "div" else 3, 10.0, 20.0)
    for nmin, nmax in ((0, 4), (1, 2)):
        rop.c[:] = 0.0
        rop.s[:] = 0.0
        shc._call(_binary_fn(op), rop._Shc, op1._Shc, op2._Shc, nmin, nmax)
        fixture_ok(rop.c, "shc_%s_c_nmin%d_nmax%d.txt" % (op, nmin, nmax))
        fixture_ok(rop.s, "shc_%s_s_nmin%d_nmax%d.txt" % (op, nmin, nmax))
    # the same through the Python mirror (out of place, degree 4)
    out = getattr(op1, op)(op2, 0, 4, inplace=False)
    fixture_ok(out.c, "shc_%s_c_nmin0_nmax4.txt" % op)


def _binary_fn(op):
    from charm_amd import _err

    f = getattr(shc.lib(), "charm_shc_" + op)
    f.restype = None
    f.argtypes = [ct.POINTER(shc._Shc)] * 3 + [ct.c_ulong, ct.c_ulong, ct.POINTER(_err._Err)]
    return f


@pytest.mark.parametrize("op", ["add", "sub", "mul", "div"])
def test_arithmetics_errors(op):
    """The argument errors tests/check_shc_arithmetics.c:178-300 expects."""
    rop = shc.Shc.from_zeros(4)
    op1 = filled(4, 1.0, 2.0)
    op2 = filled(5 if op == "div" else 3, 10.0, 20.0)
    f = _binary_fn(op)
    for nmin, nmax in ((0, 5), (5, 6), (4, 3)):
        with pytest.raises(CHarmError) as e:
            shc._call(f, rop._Shc, op1._Shc, op2._Shc, nmin, nmax)
        assert e.value.code == CHARM_EFUNCARG
    for obj in (rop, op1, op2):
        for attr in ("mu", "r"):
            setattr(obj, attr, getattr(obj, attr) + 0.1)
            with pytest.raises(CHarmError):
                shc._call(f, rop._Shc, op1._Shc, op2._Shc, 0, 4)
            setattr(obj, attr, getattr(obj, attr) - 0.1)
    # mu and r equal within charm_glob_threshold pass
    op1.mu = 1.0 + 10 * np.finfo(float).eps
    shc._call(f, rop._Shc, op1._Shc, op2._Shc, 0, 3)


def test_div_zero_and_padding_errors():
    """charm_shc_div refuses a padded divisor and zero divisors, except the
    S_{n,0} (src/shc/shc_div.c:33-79), and writes nothing when it does."""
    op1 = filled(4, 1.0, 2.0)
    op2 = filled(3, 10.0, 20.0)
    rop = shc.Shc.from_zeros(4)
    f = _binary_fn("div")
    with pytest.raises(CHarmError):
        shc._call(f, rop._Shc, op1._Shc, op2._Shc, 0, 4)
    op2 = filled(4, 10.0, 20.0)
    op2.set_coeffs(3, 2, c=0.0)
    rop.c[:] = 7.0
    with pytest.raises(CHarmError):
        shc._call(f, rop._Shc, op1._Shc, op2._Shc, 0, 4)
    assert (rop.c == 7.0).all()
    # the zero lies outside 0..2: fine
    shc._call(f, rop._Shc, op1._Shc, op2._Shc, 0, 2)
    # S_{n,0} are zero in op2 and not divided; rop's S_{n,0} stay as they were
    op2.set_coeffs(3, 2, c=10.0)
    rop.s[:] = 5.0
    shc._call(f, rop._Shc, op1._Shc, op2._Shc, 0, 4)
    n, m = rop.get_degrees_orders()
    assert (rop.s[m == 0] == 5.0).all()
    assert np.allclose(rop.s[m > 0], 0.1)
    # op1 shorter than the range: zeros there
    short = filled(2, 1.0, 2.0)
    shc._call(f, rop._Shc, short._Shc, op2._Shc, 0, 4)
    assert (rop.c[n > 2] == 0.0).all() and (rop.s[n > 2] == 0.0).all()


@pytest.mark.parametrize("N1,N2,nmin,nmax", [(30, 20, 0, 30), (20, 30, 5, 25), (25, 25, 3, 25),
                                             (10, 12, 0, 40)])
def test_arithmetics_random_zero_padding(N1, N2, nmin, nmax):
    """Out-of-place results against numpy on padded copies of the operands;
    degrees outside nmin..nmax of rop are untouched."""
    a, b = random_model(N1, 1), random_model(N2, 2)
    for op, np_op in (("add", np.add), ("sub", np.subtract), ("mul", np.multiply)):
        rop = random_model(nmax, 3)
        before_c, before_s = rop.c.copy(), rop.s.copy()
        shc._call(_binary_fn(op), rop._Shc, a._Shc, b._Shc, nmin, nmax)
        n, m = rop.get_degrees_orders()
        pa = shc.Shc.from_zeros(nmax, a.mu, a.r)
        pb = shc.Shc.from_zeros(nmax, a.mu, a.r)
        for src, dst in ((a, pa), (b, pb)):
            top = min(src.nmax, nmax)
            for mm in range(top + 1):
                c, s = src.get_coeffs(m=mm)
                k = top + 1 - mm
                i0 = shc.index(nmax, mm, mm)
                dst.c[i0:i0 + k] = c[:k]
                dst.s[i0:i0 + k] = s[:k]
        inr = (n >= nmin) & (n <= nmax)
        assert np.array_equal(rop.c[inr], np_op(pa.c, pb.c)[inr])
        assert np.array_equal(rop.s[inr], np_op(pa.s, pb.s)[inr])
        assert np.array_equal(rop.c[~inr], before_c[~inr])
        assert np.array_equal(rop.s[~inr], before_s[~inr])


def test_arithmetics_in_place_and_operators():
    a, b = random_model(15, 4), random_model(12, 5)
    ref = a.c.copy()
    c = a + b
    assert c.nmax == 15
    a += b
    assert np.array_equal(a.c, c.c) and np.array_equal(a.s, c.s)
    a -= b
    assert np.allclose(a.c, ref, rtol=0, atol=1e-14)
    # op1 == op2 == rop: doubling
    d = shc.Shc.from_copy(b)
    d.add(d)
    assert np.array_equal(d.c, 2 * b.c)
    q = b / b
    qm = q.get_degrees_orders()[1]
    assert np.array_equal(q.c, np.ones(q.c.size))
    assert np.array_equal(q.s[qm > 0], np.ones(int((qm > 0).sum())))
    assert np.array_equal(q.s[qm == 0], np.zeros(int((qm == 0).sum())))
    p = b * b
    assert np.array_equal(p.c, b.c * b.c)


# ---------------------------------------------------------------------------
# Degree- and order-wise products: tests/check_shc_arithmetics_wise.c
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("op", ["mul_degree_wise", "mul_order_wise", "div_degree_wise",
                                "div_order_wise"])
def test_wise_golden(op):
    a = np.arange(2, 7, dtype=np.float64)       # a[i] = i + 2, i = 0..4
    s = filled(4, 1.0, 2.0)
    getattr(s, op)(a, 0, 4)
    fixture_ok(s.c, "shc_%s_c_nmin0_nmax4.txt" % op)
    fixture_ok(s.s, "shc_%s_s_nmin0_nmax4.txt" % op)
    s = filled(4, 1.0, 2.0)
    getattr(s, op)(a[1:] if "degree" in op else a, 1, 2)
    fixture_ok(s.c, "shc_%s_c_nmin1_nmax2.txt" % op)
    fixture_ok(s.s, "shc_%s_s_nmin1_nmax2.txt" % op)
    for nmin, nmax in ((0, 5), (5, 6), (4, 3)):
        with pytest.raises((CHarmError, ValueError)):
            getattr(s, op)(np.ones(8), nmin, nmax)
    if op.startswith("div"):
        with pytest.raises(CHarmError) as e:
            getattr(s, op)(np.zeros(5), 0, 4)
        assert e.value.code == CHARM_EFUNCARG


def test_wise_errors_through_c_abi():
    """Range errors reported by the C functions themselves (the Python
    mirror lets them through)."""
    from charm_amd import _err

    s = filled(4, 1.0, 2.0)
    a = np.ones(16)
    for name in ("mul_degree_wise", "mul_order_wise", "div_degree_wise", "div_order_wise"):
        f = getattr(shc.lib(), "charm_shc_" + name)
        f.restype = None
        f.argtypes = [ct.POINTER(shc._Shc), ct.POINTER(ct.c_double), ct.c_ulong, ct.c_ulong,
                      ct.POINTER(_err._Err)]
        for nmin, nmax in ((0, 5), (5, 6), (4, 3)):
            with pytest.raises(CHarmError) as e:
                shc._call(f, s._Shc, a.ctypes.data_as(ct.POINTER(ct.c_double)), nmin, nmax)
            assert e.value.code == CHARM_EFUNCARG


def test_wise_random():
    b = random_model(40, 6)
    n, m = b.get_degrees_orders()
    n, m = n.astype(int), m.astype(int)
    a = np.linspace(0.5, 3.0, 41)
    for nmin, nmax in ((0, 40), (7, 31)):
        inr = (n >= nmin) & (n <= nmax)
        x = shc.Shc.from_copy(b)
        x.mul_degree_wise(a[nmin:nmax + 1], nmin, nmax)
        assert np.array_equal(x.c[inr], b.c[inr] * a[n[inr]])
        assert np.array_equal(x.c[~inr], b.c[~inr])
        x.div_degree_wise(a[nmin:nmax + 1], nmin, nmax)
        assert np.allclose(x.c, b.c, rtol=1e-15, atol=0)
        y = shc.Shc.from_copy(b)
        y.mul_order_wise(a, nmin, nmax)
        assert np.array_equal(y.s[inr], b.s[inr] * a[m[inr]])
        y.div_order_wise(a, nmin, nmax)
        assert np.allclose(y.s, b.s, rtol=1e-15, atol=0)


# ---------------------------------------------------------------------------
# Spectra: tests/check_shc_dav.c, check_shc_ddav.c
# ---------------------------------------------------------------------------
def earth2014_deg4():
    return shc.Shc.from_file("mtx", os.path.join(GOLDEN, "Earth2014-RET2014-degree4-mtx.txt"))


@pytest.mark.parametrize("what", ["dv", "da"])
def test_dv_da_golden(what):
    s = earth2014_deg4()
    ref = read_fixture("double", "shc_nx4_d%s.txt" % what[1])
    for nmax in range(5):
        x = getattr(s, what)(nmax)
        ok = nearly_equal(x, ref[:nmax + 1], EPS10)
        assert ok.all(), (nmax, x, ref)


@pytest.mark.parametrize("what", ["ddv", "dda"])
def test_ddv_dda_of_itself_is_zero(what):
    s = earth2014_deg4()
    for nmax in range(5):
        assert np.array_equal(getattr(s, what)(s, nmax), np.zeros(nmax + 1))


def test_spectra_random_and_errors():
    a, b = random_model(60, 7), random_model(50, 8)
    n, m = a.get_degrees_orders()
    dv = np.bincount(n.astype(int), weights=a.c ** 2 + a.s ** 2)
    assert np.allclose(a.dv(), dv, rtol=1e-14, atol=0)
    assert np.allclose(a.da(), np.sqrt(dv), rtol=1e-14, atol=0)
    # the difference spectrum is the spectrum of the difference
    d = a.sub(b, nmax=50, inplace=False)
    assert np.allclose(a.ddv(b), d.dv(), rtol=1e-14, atol=0)
    assert np.array_equal(a.dda(b, 20), np.sqrt(a.ddv(b, 20)))
    assert len(a.dda(b)) == 51
    # errors from the C functions: nmax beyond a model, mu or r differ
    from charm_amd import _err

    f = shc.lib().charm_shc_ddv
    f.restype = None
    f.argtypes = [ct.POINTER(shc._Shc)] * 2 + [ct.c_ulong, ct.POINTER(ct.c_double),
                                               ct.POINTER(_err._Err)]
    out = np.zeros(70)
    ptr = out.ctypes.data_as(ct.POINTER(ct.c_double))
    with pytest.raises(CHarmError):
        shc._call(f, a._Shc, b._Shc, 55, ptr)
    b.mu = b.mu * 1.01
    with pytest.raises(CHarmError):
        shc._call(f, a._Shc, b._Shc, 10, ptr)
    b.mu = a.mu
    b.r = b.r * 1.01
    with pytest.raises(CHarmError):
        shc._call(f, a._Shc, b._Shc, 10, ptr)
    g = shc.lib().charm_shc_dv
    g.restype = None
    g.argtypes = [ct.POINTER(shc._Shc), ct.c_ulong, ct.POINTER(ct.c_double),
                  ct.POINTER(_err._Err)]
    with pytest.raises(CHarmError) as e:
        shc._call(g, a._Shc, 61, ptr)
    assert e.value.code == CHARM_EFUNCARG


# ---------------------------------------------------------------------------
# rescale: tests/check_shc_rescale.c
# ---------------------------------------------------------------------------
def test_rescale_golden():
    s = shc.Shc.from_file("mtx", os.path.join(GOLDEN, "EGM96-degree10-mtx.txt"))
    s.rescale(s.mu * 1.1, s.r * 1.5)
    fixture_ok(s.c, "shc_rescale_c.txt", THRESH["double"])
    fixture_ok(s.s, "shc_rescale_s.txt", THRESH["double"])


def test_rescale_random_round_trip():
    a = random_model(80, 9)
    b = shc.Shc.from_copy(a)
    mu, r = a.mu, a.r
    b.rescale(mu * 0.97, r * 1.2)
    assert b.mu == mu * 0.97 and b.r == r * 1.2
    n, _ = a.get_degrees_orders()
    f = (mu / (mu * 0.97)) * (r / (r * 1.2)) ** n.astype(float)
    assert np.allclose(b.c, a.c * f, rtol=1e-13, atol=0)
    b.rescale(mu=mu)
    b.rescale(r=r)
    assert np.allclose(b.c, a.c, rtol=1e-13, atol=0)
    assert np.allclose(b.s, a.s, rtol=1e-13, atol=0)


# ---------------------------------------------------------------------------
# copy: tests/check_shc_copy.c
# ---------------------------------------------------------------------------
def test_copy():
    ref = shc.Shc.from_zeros(10, 1.1, 2.2)
    n, m = ref.get_degrees_orders()
    ref.c[:] = n
    ref.s[:] = n - m
    for nmin, nmax in ((0, 10), (2, 2), (5, 7)):
        x = shc.Shc.from_copy(ref, nmin, nmax)
        inr = (n >= nmin) & (n <= nmax)
        assert np.array_equal(x.c, np.where(inr, ref.c, 0.0))
        assert np.array_equal(x.s, np.where(inr, ref.s, 0.0))
        assert x.nmax == 10 and x.mu == 1.1 and x.r == 2.2 and x.owner
    big = shc.Shc.from_copy(ref, 0, 10, 14)
    assert big.nmax == 14
    assert big.get_coeffs(7, 3) == ref.get_coeffs(7, 3)
    assert big.get_coeffs(12, 3) == (0.0, 0.0)
    # invalid ranges: the C function returns NULL
    f = shc.lib().charm_shc_copy
    f.restype = ct.POINTER(shc._Shc)
    f.argtypes = [ct.POINTER(shc._Shc), ct.c_ulong, ct.c_ulong, ct.c_ulong]
    for args in ((11, 10, 10), (10, 11, 10), (10, 9, 10), (0, 10, 9)):
        assert not f(ref._Shc, *args)
    ref._Shc.contents.distributed = True
    assert not f(ref._Shc, 0, 10, 10)
    ref._Shc.contents.distributed = False


def test_get_set_coeffs():
    s = shc.Shc.from_zeros(6)
    s.set_coeffs(4, 2, c=1.5, s=-2.5)
    assert s.get_coeffs(4, 2) == (1.5, -2.5)
    s.set_coeffs(n=3, c=np.arange(4.0))
    assert np.array_equal(s.get_coeffs(n=3)[0], np.arange(4.0))
    s.set_coeffs(m=5, s=np.array([7.0, 8.0]))
    assert np.array_equal(s.get_coeffs(m=5)[1], [7.0, 8.0])
    s.set_coeffs([1, 6], [0, 6], c=np.array([9.0, 10.0]))
    assert np.array_equal(s.get_coeffs([1, 6], [0, 6])[0], [9.0, 10.0])
    n, m = s.get_degrees_orders()
    assert np.array_equal(shc.index(6, n.astype(int), m.astype(int)), np.arange(n.size))
    with pytest.raises(ValueError):
        s.get_coeffs(3, 4)
    with pytest.raises(ValueError):
        s.set_coeffs(n=3, c=np.arange(3.0))
    with pytest.raises(ValueError):
        s.get_coeffs([1], 0)
