"""Spherical harmonic coefficients (mirrors PyHarm's ``shc.Shc``,
wrap/pyharm/shc.py:65-528).

Coefficients are stored as in CHarm: m-major triangles C_{m,m..nmax} for
m = 0..nmax, contiguous from ``c[0]``; ``S_{n,0}`` are zeros.
"""
import ctypes as _ct

import numpy as _np

from ._lib import lib

_MU = 1.0
_R = 1.0

# src/shc/shc.h.in:60-108
NMAX_MODEL = 2 ** 64 - 2     # CHARM_SHC_NMAX_MODEL (ULONG_MAX - 1, LP64)
NMAX_ERROR = 2 ** 64 - 1     # CHARM_SHC_NMAX_ERROR
WRITE_N = 0                  # CHARM_SHC_WRITE_N: degree varies fastest
WRITE_M = 1                  # CHARM_SHC_WRITE_M: order varies fastest
_FILE_TYPES = ("gfc", "tbl", "dov", "mtx", "bin")
_DEFAULT_FORMAT = "%0.16e"


class _Shc(_ct.Structure):
    # src/shc/shc.h.in:139-381 (non-MPI build)
    _fields_ = [("nmax", _ct.c_ulong),
                ("mu", _ct.c_double),
                ("r", _ct.c_double),
                ("nc", _ct.c_size_t),
                ("ns", _ct.c_size_t),
                ("c", _ct.POINTER(_ct.POINTER(_ct.c_double))),
                ("s", _ct.POINTER(_ct.POINTER(_ct.c_double))),
                ("owner", _ct.c_bool),
                ("distributed", _ct.c_bool)]


def nc(nmax):
    return (nmax + 1) * (nmax + 2) // 2


def index(nmax, n, m):
    """Position of (n, m) in the coefficient arrays of a model of degree nmax."""
    return m * (nmax + 1) - (m * (m - 1)) // 2 + (n - m)


class Shc:
    """Spherical harmonic coefficients owned by CHarm's C allocator (or views
    of caller arrays with :meth:`from_arrays`)."""

    def __init__(self, nmax, mu, r, method, c=None, s=None):
        nmax, mu, r = int(nmax), float(mu), float(r)
        if nmax < 0:
            raise ValueError("nmax must be non-negative")
        if r <= 0.0:
            raise ValueError("The radius must be positive.")
        self._arrays = None
        L = lib()
        if method == "arrays":
            c = _np.ascontiguousarray(c, dtype=_np.float64)
            s = _np.ascontiguousarray(s, dtype=_np.float64)
            if c.size != nc(nmax) or s.size != nc(nmax):
                raise ValueError("c and s must hold (nmax+1)(nmax+2)/2 values")
            self._arrays = (c, s)
            f = L.charm_shc_init
            f.restype = _ct.POINTER(_Shc)
            f.argtypes = [_ct.c_ulong, _ct.c_double, _ct.c_double,
                          _ct.POINTER(_ct.c_double), _ct.POINTER(_ct.c_double)]
            p = f(nmax, mu, r, c.ctypes.data_as(_ct.POINTER(_ct.c_double)),
                  s.ctypes.data_as(_ct.POINTER(_ct.c_double)))
        else:
            f = L.charm_shc_calloc if method == "zeros" else L.charm_shc_malloc
            f.restype = _ct.POINTER(_Shc)
            f.argtypes = [_ct.c_ulong, _ct.c_double, _ct.c_double]
            p = f(nmax, mu, r)
        if not p:
            raise MemoryError("Failed to initialize the charm_shc structure.")
        self._Shc = p

    @classmethod
    def from_garbage(cls, nmax, mu=_MU, r=_R):
        return cls(nmax, mu, r, "garbage")

    @classmethod
    def from_zeros(cls, nmax, mu=_MU, r=_R):
        return cls(nmax, mu, r, "zeros")

    @classmethod
    def from_arrays(cls, nmax, c, s, mu=_MU, r=_R):
        return cls(nmax, mu, r, "arrays", c, s)

    @classmethod
    def from_file(cls, file_type, pathname, nmax=-1, epoch=None):
        """Coefficients up to degree ``nmax`` (all of them when -1) from a
        ``gfc``, ``tbl``, ``dov``, ``mtx`` or ``bin`` file; ``epoch``
        ("yyyyMMdd" or "yyyyMMdd.hhmm") evaluates time variable gfc models
        (wrap/pyharm/shc.py:575-641)."""
        nmax_file = cls.nmax_from_file(file_type, pathname)
        nmax = nmax_file if nmax < 0 else int(nmax)
        if nmax > nmax_file:
            raise ValueError("nmax = %d exceeds the maximum degree %d of the file"
                             % (nmax, nmax_file))
        out = cls.from_zeros(nmax)
        _read(file_type, pathname, nmax, out._Shc, epoch)
        return out

    @staticmethod
    def nmax_from_file(file_type, pathname):
        """Maximum harmonic degree stored in a coefficient file
        (wrap/pyharm/shc.py:962-997)."""
        return _read(file_type, pathname, NMAX_MODEL, None, None)

    def to_file(self, file_type, pathname, nmax=-1, formatting=_DEFAULT_FORMAT,
                ordering=WRITE_N):
        """Writes the coefficients up to degree ``nmax`` (all when -1)
        (wrap/pyharm/shc.py:644-703)."""
        from . import _err

        if file_type not in ("tbl", "dov", "mtx", "bin"):
            raise ValueError("Unsupported file type \"%s\"" % file_type)
        nmax = self.nmax if nmax < 0 else int(nmax)
        L = lib()
        err = _err.init()
        try:
            if file_type == "bin":
                f = L.charm_shc_write_bin
                f.argtypes = [_ct.POINTER(_Shc), _ct.c_ulong, _ct.c_char_p, _ct.POINTER(_err._Err)]
                f.restype = None
                f(self._Shc, nmax, pathname.encode(), err)
            elif file_type == "mtx":
                f = L.charm_shc_write_mtx
                f.argtypes = [_ct.POINTER(_Shc), _ct.c_ulong, _ct.c_char_p, _ct.c_char_p,
                              _ct.POINTER(_err._Err)]
                f.restype = None
                f(self._Shc, nmax, formatting.encode(), pathname.encode(), err)
            else:
                f = L.charm_shc_write_tbl if file_type == "tbl" else L.charm_shc_write_dov
                f.argtypes = [_ct.POINTER(_Shc), _ct.c_ulong, _ct.c_char_p, _ct.c_int,
                              _ct.c_char_p, _ct.POINTER(_err._Err)]
                f.restype = None
                f(self._Shc, nmax, formatting.encode(), int(ordering), pathname.encode(), err)
            _err.handler(err, True)
        finally:
            _err.free(err)

    @property
    def nmax(self):
        return int(self._Shc.contents.nmax)

    @property
    def mu(self):
        return float(self._Shc.contents.mu)

    @mu.setter
    def mu(self, v):
        self._Shc.contents.mu = float(v)

    @property
    def r(self):
        return float(self._Shc.contents.r)

    @r.setter
    def r(self, v):
        if v <= 0.0:
            raise ValueError("The radius must be positive.")
        self._Shc.contents.r = float(v)

    @property
    def c(self):
        return _np.ctypeslib.as_array(self._Shc.contents.c[0], shape=(int(self._Shc.contents.nc),))

    @property
    def s(self):
        return _np.ctypeslib.as_array(self._Shc.contents.s[0], shape=(int(self._Shc.contents.ns),))

    @property
    def owner(self):
        return bool(self._Shc.contents.owner)

    def get_coeffs(self, n, m):
        i = index(self.nmax, n, m)
        return self.c[i], self.s[i]

    def _free(self):
        if getattr(self, "_Shc", None) is not None:
            f = lib().charm_shc_free
            f.restype = None
            f.argtypes = [_ct.POINTER(_Shc)]
            f(self._Shc)
            self._Shc = None

    def __del__(self):
        self._free()


def _read(file_type, pathname, nmax, shc_ptr, epoch):
    """charm_shc_read_<file_type>; returns the file's maximum degree."""
    from . import _err

    if file_type not in _FILE_TYPES:
        raise ValueError("Unsupported file type \"%s\"; supported: %s"
                         % (file_type, ", ".join(_FILE_TYPES)))
    f = getattr(lib(), "charm_shc_read_" + file_type)
    f.restype = _ct.c_ulong
    ptr = shc_ptr if shc_ptr is not None else _ct.POINTER(_Shc)()
    err = _err.init()
    try:
        if file_type == "gfc":
            f.argtypes = [_ct.c_char_p, _ct.c_ulong, _ct.c_char_p, _ct.POINTER(_Shc),
                          _ct.POINTER(_err._Err)]
            v = f(pathname.encode(), nmax, epoch.encode() if epoch else None, ptr, err)
        else:
            f.argtypes = [_ct.c_char_p, _ct.c_ulong, _ct.POINTER(_Shc), _ct.POINTER(_err._Err)]
            v = f(pathname.encode(), nmax, ptr, err)
        _err.handler(err, True)
    finally:
        _err.free(err)
    return int(v)
