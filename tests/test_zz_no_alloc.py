"""charm_amd_shs_sctr_dev_noalloc (no counterpart in CHarm): scattered
synthesis with every buffer in a static device array, for callers that
cannot allocate.  Checked through the C ABI against charm_amd_shs_sctr_dev on
the same device inputs, on consecutive calls of decreasing size (the array
keeps the previous call's contents), with and without r = R, with the polar
optimisation, and at the largest size it takes."""
import ctypes as ct

import numpy as np
import pytest

import charm_amd
from charm_amd import _err
from charm_amd.device import DevArray, Stream, shs_sctr
from charm_amd.workloads import GM, R_EARTH, kaula_coeffs, scattered_points

pytestmark = pytest.mark.gpu

NMAX_MAX = 2190
NPOINT_MAX = 1 << 20


def noalloc(lat, lon, r, cs, nmax, mu, R, out, stream):
    f = charm_amd.lib().charm_amd_shs_sctr_dev_noalloc
    f.restype = None
    f.argtypes = [ct.c_size_t, ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_ulong, ct.c_double,
                  ct.c_double, ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.POINTER(_err._Err)]
    err = _err.init()
    try:
        f(lat.size, lat.ptr, lon.ptr, r.ptr, int(nmax), float(mu), float(R), cs.ptr, out.ptr,
          stream.handle, err)
        _err.handler(err, True)
    finally:
        _err.free(err)


def both(npoint, nmax, seed, unit, st):
    lat, lon, r = scattered_points(npoint, seed)
    if unit:
        r = np.full(npoint, R_EARTH)
    C, S = kaula_coeffs(nmax, seed + 1)
    d = [DevArray.from_numpy(x) for x in (lat, lon, r)]
    cs = DevArray.from_numpy(np.stack([C, S], axis=1))
    f0, f1 = DevArray(npoint), DevArray(npoint)
    shs_sctr(*d, cs, nmax, GM, R_EARTH, f0, st)
    noalloc(*d, cs, nmax, GM, R_EARTH, f1, st)
    st.synchronize()
    a, b = f0.numpy(), f1.numpy()
    assert np.abs(a - b).max() <= 1e-12 * np.abs(a).max()
    assert np.array_equal(a, b), (npoint, nmax, unit)
    return b


def test_noalloc_equals_allocating_entry_point():
    st = Stream()
    print("noalloc: 300000 points nmax 2190 r != R, then 5000 points nmax 700 r = R", flush=True)
    both(300000, 2190, 11, False, st)
    both(5000, 700, 12, True, st)
    both(100, 3, 13, False, st)


def test_noalloc_largest_size_and_polar():
    st = Stream()
    print("noalloc: 1048576 points nmax 2190", flush=True)
    both(NPOINT_MAX, NMAX_MAX, 21, False, st)
    lib = charm_amd.lib()
    a1 = ct.c_ulong.in_dll(lib, "charm_glob_polar_optimization_a1")
    a2 = ct.c_double.in_dll(lib, "charm_glob_polar_optimization_a2")
    old = (a1.value, a2.value)
    a1.value, a2.value = 20, 0.0
    try:
        print("noalloc: polar optimisation, 20000 points nmax 600", flush=True)
        both(20000, 600, 22, False, st)
    finally:
        a1.value, a2.value = old


def test_noalloc_rejects_beyond_limits():
    st = Stream()
    lat, lon, r = (DevArray(NPOINT_MAX + 1) for _ in range(3))
    cs = DevArray((2, 2))
    f = DevArray(NPOINT_MAX + 1)
    with pytest.raises(Exception, match="at most 1048576 points and degree 2190"):
        noalloc(lat, lon, r, cs, 10, GM, R_EARTH, f, st)
    with pytest.raises(Exception, match="at most 1048576 points and degree 2190"):
        noalloc(lat, lon, r, cs, NMAX_MAX + 1, GM, R_EARTH, f, st)
