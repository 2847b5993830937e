"""Spherical harmonic analysis (wrap/pyharm/sha.py:36-103)."""
import ctypes as _ct

import numpy as _np

from . import _err
from . import crd as _crd
from . import shc as _shc
from ._lib import lib


def point(pnt, f, nmax, mu=1.0, r=1.0):
    """Analysis of ``f`` sampled on the quadrature grid ``pnt`` up to degree
    ``nmax``; returns an :class:`~charm_amd.shc.Shc` with constants mu, r."""
    if not isinstance(pnt, (_crd.PointGridGL,)) and not (
            isinstance(pnt, _crd._PointBase) and pnt._type in (_crd.POINT_GRID_DH1,
                                                                 _crd.POINT_GRID_DH2)):
        raise TypeError("pnt must be a quadrature grid")
    f = _np.ascontiguousarray(f, dtype=_np.float64)
    if f.ndim != 2:
        raise ValueError("f must be a 2-D array")
    if f.shape[0] != pnt.lat.size:
        raise ValueError("The number of elements in 'f' along the axis '0' does not match "
                         "the number of latitudes in 'pnt' (%d vs. %d)."
                         % (f.shape[0], pnt.lat.size))
    if f.shape[1] != pnt.lon.size:
        raise ValueError("The number of elements in 'f' along the axis '1' does not match "
                         "the number of longitudes in 'pnt' (%d vs. %d)."
                         % (f.shape[1], pnt.lon.size))
    nmax = int(nmax)
    if nmax < 0:
        raise ValueError("nmax must be non-negative")
    fn = lib().charm_sha_point
    fn.restype = None
    fn.argtypes = [_ct.POINTER(_crd._Point), _ct.POINTER(_ct.c_double), _ct.c_ulong,
                   _ct.POINTER(_shc._Shc), _ct.POINTER(_err._Err)]
    shcs = _shc.Shc.from_garbage(nmax, mu, r)
    err = _err.init()
    try:
        fn(pnt._Point, f.ctypes.data_as(_ct.POINTER(_ct.c_double)), _ct.c_ulong(nmax),
           shcs._Shc, err)
        _err.handler(err, True)
    finally:
        _err.free(err)
    return shcs
