"""Spherical harmonic synthesis (wrap/pyharm/shs.py:27-73)."""
import ctypes as _ct

import numpy as _np

from . import _err
from . import crd as _crd
from . import shc as _shc
from ._lib import lib


def point(pnt, shcs, nmax):
    """Synthesis of ``shcs`` up to degree ``nmax`` at the points ``pnt``.

    Returns a (nlat, nlon) array for grids and an (npoint,) array for
    scattered points."""
    if not isinstance(pnt, _crd._PointBase):
        raise TypeError("pnt must be a point class instance")
    if not isinstance(shcs, _shc.Shc):
        raise TypeError("shcs must be an Shc instance")
    nmax = int(nmax)
    if nmax < 0:
        raise ValueError("nmax must be non-negative")
    if nmax > shcs.nmax:
        raise ValueError("nmax cannot exceed shcs.nmax")
    f = lib().charm_shs_point
    f.restype = None
    f.argtypes = [_ct.POINTER(_crd._Point), _ct.POINTER(_shc._Shc), _ct.c_ulong,
                  _ct.POINTER(_ct.c_double), _ct.POINTER(_err._Err)]
    if isinstance(pnt, _crd.PointSctr):
        shape = (pnt.npoint,)
    else:
        shape = (pnt.lat.size, pnt.lon.size)
    out = _np.zeros(shape, dtype=_np.float64, order="C")
    err = _err.init()
    try:
        f(pnt._Point, shcs._Shc, _ct.c_ulong(nmax), out.ctypes.data_as(_ct.POINTER(_ct.c_double)),
          err)
        _err.handler(err, True)
    finally:
        _err.free(err)
    return out
