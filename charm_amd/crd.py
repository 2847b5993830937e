"""Evaluation points (mirrors PyHarm's ``crd`` classes,
wrap/pyharm/crd.py:43-716)."""
import ctypes as _ct

import numpy as _np

from ._lib import lib

POINT_SCATTERED = 1
POINT_GRID = 2
POINT_GRID_GL = 3
POINT_GRID_DH1 = 4
POINT_GRID_DH2 = 5


class _Point(_ct.Structure):
    # src/crd/crd.h.in:125-332 (non-MPI build)
    _fields_ = [("type", _ct.c_int),
                ("nlat", _ct.c_size_t),
                ("nlon", _ct.c_size_t),
                ("npoint", _ct.c_size_t),
                ("lat", _ct.POINTER(_ct.c_double)),
                ("lon", _ct.POINTER(_ct.c_double)),
                ("r", _ct.POINTER(_ct.c_double)),
                ("w", _ct.POINTER(_ct.c_double)),
                ("owner", _ct.c_bool),
                ("distributed", _ct.c_bool)]


def _dp(a):
    return a.ctypes.data_as(_ct.POINTER(_ct.c_double))


class _PointBase:
    _type = None

    def _from_arrays(self, lat, lon, r):
        self._arrays = [_np.ascontiguousarray(lat, dtype=_np.float64),
                        _np.ascontiguousarray(lon, dtype=_np.float64),
                        _np.ascontiguousarray(r, dtype=_np.float64)]
        lat, lon, r = self._arrays
        if lat.ndim != 1 or lon.ndim != 1 or r.ndim != 1 or r.size != lat.size:
            raise ValueError("lat, lon and r must be 1-D; r must match lat")
        f = lib().charm_crd_point_init
        f.restype = _ct.POINTER(_Point)
        f.argtypes = [_ct.c_int, _ct.c_size_t, _ct.c_size_t, _ct.POINTER(_ct.c_double),
                      _ct.POINTER(_ct.c_double), _ct.POINTER(_ct.c_double)]
        p = f(self._type, lat.size, lon.size, _dp(lat), _dp(lon), _dp(r))
        if not p:
            raise ValueError("charm_crd_point_init rejected the input")
        self._Point = p

    def _arr(self, name, n):
        ptr = getattr(self._Point.contents, name)
        if not ptr or n == 0:
            return _np.zeros(0)
        return _np.ctypeslib.as_array(ptr, shape=(n,))

    @property
    def lat(self):
        return self._arr("lat", int(self._Point.contents.nlat))

    @property
    def lon(self):
        return self._arr("lon", int(self._Point.contents.nlon))

    @property
    def r(self):
        return self._arr("r", int(self._Point.contents.nlat))

    @property
    def w(self):
        return self._arr("w", int(self._Point.contents.nlat))

    @property
    def npoint(self):
        return int(self._Point.contents.npoint)

    def __del__(self):
        if getattr(self, "_Point", None) is not None:
            f = lib().charm_crd_point_free
            f.restype = None
            f.argtypes = [_ct.POINTER(_Point)]
            f(self._Point)
            self._Point = None


class PointSctr(_PointBase):
    """Scattered points (lat, lon, r of equal length)."""
    _type = POINT_SCATTERED

    def __init__(self, lat, lon, r):
        self._from_arrays(lat, lon, r)


class PointGrid(_PointBase):
    """Custom grid: the tensor product of lat (with radii r) and lon."""
    _type = POINT_GRID

    def __init__(self, lat, lon, r):
        self._from_arrays(lat, lon, r)


class _PointGridDH(_PointBase):
    _fn = None

    def __init__(self, nmax, r=1.0):
        f = getattr(lib(), "charm_crd_point_%s" % self._fn)
        f.restype = _ct.POINTER(_Point)
        f.argtypes = [_ct.c_ulong, _ct.c_double]
        p = f(int(nmax), float(r))
        if not p:
            raise ValueError("charm_crd_point_%s failed" % self._fn)
        self._Point = p
        self._arrays = None

    @staticmethod
    def get_nmax(nlat):
        return nlat // 2 - 1

    @classmethod
    def shape(cls, nmax):
        f = getattr(lib(), "charm_crd_point_%s_shape" % cls._fn)
        f.restype = None
        f.argtypes = [_ct.c_ulong, _ct.POINTER(_ct.c_size_t), _ct.POINTER(_ct.c_size_t)]
        a, b = _ct.c_size_t(), _ct.c_size_t()
        f(int(nmax), _ct.byref(a), _ct.byref(b))
        return a.value, b.value


class PointGridDH1(_PointGridDH):
    """Non-equiangular Driscoll-Healy grid for degree nmax on a sphere of
    radius r (charm_crd_point_dh1): 2(nmax+1) x 2(nmax+1) points."""
    _type = POINT_GRID_DH1
    _fn = "dh1"


class PointGridDH2(_PointGridDH):
    """Equiangular Driscoll-Healy grid (charm_crd_point_dh2): 2(nmax+1) x
    4(nmax+1) points."""
    _type = POINT_GRID_DH2
    _fn = "dh2"


class PointGridGL(_PointBase):
    """Gauss-Legendre grid for degree nmax on a sphere of radius r
    (charm_crd_point_gl)."""
    _type = POINT_GRID_GL

    def __init__(self, nmax, r=1.0):
        f = lib().charm_crd_point_gl
        f.restype = _ct.POINTER(_Point)
        f.argtypes = [_ct.c_ulong, _ct.c_double]
        p = f(int(nmax), float(r))
        if not p:
            raise ValueError("charm_crd_point_gl failed")
        self._Point = p
        self._arrays = None

    @staticmethod
    def get_nmax(nlat):
        return nlat - 1

    @staticmethod
    def shape(nmax):
        f = lib().charm_crd_point_gl_shape
        f.restype = None
        f.argtypes = [_ct.c_ulong, _ct.POINTER(_ct.c_size_t), _ct.POINTER(_ct.c_size_t)]
        a, b = _ct.c_size_t(), _ct.c_size_t()
        f(int(nmax), _ct.byref(a), _ct.byref(b))
        return a.value, b.value
