"""Spherical harmonic synthesis (wrap/pyharm/shs.py:27-73)."""
import ctypes as _ct

import numpy as _np

from . import _err
from . import crd as _crd
from . import shc as _shc
from ._lib import lib


def _check(pnt, shcs, nmax):
    if not isinstance(pnt, _crd._PointBase):
        raise TypeError("pnt must be a point class instance")
    if not isinstance(shcs, _shc.Shc):
        raise TypeError("shcs must be an Shc instance")
    nmax = int(nmax)
    if nmax < 0:
        raise ValueError("nmax must be non-negative")
    if nmax > shcs.nmax:
        raise ValueError("nmax cannot exceed shcs.nmax")
    return nmax


def _shape(pnt):
    if isinstance(pnt, _crd.PointSctr):
        return (pnt.npoint,)
    return (pnt.lat.size, pnt.lon.size)


def point(pnt, shcs, nmax):
    """Synthesis of ``shcs`` up to degree ``nmax`` at the points ``pnt``.

    Returns a (nlat, nlon) array for grids and an (npoint,) array for
    scattered points."""
    nmax = _check(pnt, shcs, nmax)
    f = lib().charm_shs_point
    f.restype = None
    f.argtypes = [_ct.POINTER(_crd._Point), _ct.POINTER(_shc._Shc), _ct.c_ulong,
                  _ct.POINTER(_ct.c_double), _ct.POINTER(_err._Err)]
    out = _np.zeros(_shape(pnt), dtype=_np.float64, order="C")
    err = _err.init()
    try:
        f(pnt._Point, shcs._Shc, _ct.c_ulong(nmax), out.ctypes.data_as(_ct.POINTER(_ct.c_double)),
          err)
        _err.handler(err, True)
    finally:
        _err.free(err)
    return out


def _point_gradn(pnt, shcs, nmax, gradn):
    nmax = _check(pnt, shcs, nmax)
    f = lib().charm_shs_point_grad1 if gradn == 1 else lib().charm_shs_point_grad2
    f.restype = None
    f.argtypes = [_ct.POINTER(_crd._Point), _ct.POINTER(_shc._Shc), _ct.c_ulong,
                  _ct.POINTER(_ct.POINTER(_ct.c_double)), _ct.POINTER(_err._Err)]
    nout = 3 if gradn == 1 else 6
    outs = [_np.zeros(_shape(pnt), dtype=_np.float64, order="C") for _ in range(nout)]
    ptrs = (_ct.POINTER(_ct.c_double) * nout)(
        *[o.ctypes.data_as(_ct.POINTER(_ct.c_double)) for o in outs])
    err = _err.init()
    try:
        f(pnt._Point, shcs._Shc, _ct.c_ulong(nmax), ptrs, err)
        _err.handler(err, True)
    finally:
        _err.free(err)
    return tuple(outs)


def point_grad1(pnt, shcs, nmax):
    """First-order gradient (fx, fy, fz) in the local north-oriented frame
    (wrap/pyharm/shs.py:76-102)."""
    return _point_gradn(pnt, shcs, nmax, 1)


def point_grad2(pnt, shcs, nmax):
    """Second-order gradient (fxx, fxy, fxz, fyy, fyz, fzz)
    (wrap/pyharm/shs.py:105-137)."""
    return _point_gradn(pnt, shcs, nmax, 2)


def point_guru(pnt, shcs, nmax, dr, dlat, dlon):
    """1/(r^(dlat+dlon) cos^dlon lat) times the derivative of orders dr,
    dlat, dlon in r, lat, lon, dr + dlat + dlon <= 2
    (wrap/pyharm/shs.py:191-281)."""
    nmax = _check(pnt, shcs, nmax)
    for name, d in (("dr", dr), ("dlat", dlat), ("dlon", dlon)):
        if int(d) != d:
            raise TypeError("'%s' must be an integer" % name)
        if d < 0:
            raise ValueError("The derivative order '%s' is '%d', but it cannot be negative."
                             % (name, d))
    f = lib().charm_shs_point_guru
    f.restype = None
    f.argtypes = [_ct.POINTER(_crd._Point), _ct.POINTER(_shc._Shc), _ct.c_ulong, _ct.c_uint,
                  _ct.c_uint, _ct.c_uint, _ct.POINTER(_ct.c_double), _ct.POINTER(_err._Err)]
    out = _np.zeros(_shape(pnt), dtype=_np.float64, order="C")
    err = _err.init()
    try:
        f(pnt._Point, shcs._Shc, _ct.c_ulong(nmax), _ct.c_uint(int(dr)), _ct.c_uint(int(dlat)),
          _ct.c_uint(int(dlon)), out.ctypes.data_as(_ct.POINTER(_ct.c_double)), err)
        _err.handler(err, True)
    finally:
        _err.free(err)
    return out
