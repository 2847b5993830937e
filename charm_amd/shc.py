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

    # ------------------------------------------------------------------
    # Coefficient access (wrap/pyharm/shc.py:1000-1358)
    # ------------------------------------------------------------------
    def _deg(self, v, what):
        if isinstance(v, bool) or not isinstance(v, (int, _np.integer)):
            raise TypeError("The %s must be an integer." % what)
        v = int(v)
        if v < 0 or v > self.nmax:
            raise ValueError("The %s %d must be within 0..%d." % (what, v, self.nmax))
        return v

    def _nm_index(self, n, m):
        n, m = self._deg(n, "degree"), self._deg(m, "order")
        if m > n:
            raise ValueError("The order m = %d cannot exceed the degree n = %d." % (m, n))
        return index(self.nmax, n, m)

    def _select(self, n, m):
        """Indices into c/s of the (n, m) selection, and whether it is a
        scalar: (n, m) integers, a degree n (orders 0..n), an order m
        (degrees m..nmax), equal-length lists of degrees and orders, or
        everything when both are None."""
        lists = (list, tuple, _np.ndarray)
        if n is None and m is None:
            return slice(None), False
        if isinstance(n, lists) and isinstance(m, lists):
            if len(n) != len(m):
                raise ValueError("The lengths of 'n' and 'm' do not match (%d vs. %d)."
                                 % (len(n), len(m)))
            return _np.array([self._nm_index(a, b) for a, b in zip(n, m)], dtype=_np.int64), False
        if isinstance(n, lists) or isinstance(m, lists):
            raise ValueError("Either both or none of 'n' and 'm' can be lists.")
        if n is not None and m is not None:
            return self._nm_index(n, m), True
        if n is not None:
            n = self._deg(n, "degree")
            return _np.array([index(self.nmax, n, k) for k in range(n + 1)], dtype=_np.int64), False
        m = self._deg(m, "order")
        i0 = index(self.nmax, m, m)
        return slice(i0, i0 + self.nmax + 1 - m), False

    def get_coeffs(self, n=None, m=None):
        """C_nm and S_nm of the selection (see ``_select``); arrays are
        copies."""
        idx, scalar = self._select(n, m)
        if scalar:
            return float(self.c[idx]), float(self.s[idx])
        return self.c[idx].copy(), self.s[idx].copy()

    def set_coeffs(self, n=None, m=None, c=None, s=None):
        """Writes ``c`` and/or ``s`` into the selection of ``get_coeffs``;
        arrays must have the selection's length."""
        idx, scalar = self._select(n, m)
        size = 1 if scalar else len(self.c[idx])
        for name, v, dst in (("c", c, self.c), ("s", s, self.s)):
            if v is None:
                continue
            if scalar:
                if not isinstance(v, (float, int, _np.floating, _np.integer)):
                    raise TypeError("'%s' must be a floating point scalar." % name)
                dst[idx] = float(v)
                continue
            v = _np.asarray(v)
            if v.ndim != 1 or not _np.issubdtype(v.dtype, _np.floating):
                raise TypeError("'%s' must be a 1D floating point array." % name)
            if v.shape != (size,):
                raise ValueError("'%s' must have %d element(s), not %d." % (name, size, v.size))
            dst[idx] = v

    def get_degrees_orders(self):
        """Degrees and orders matching the entries of ``c`` and ``s``."""
        N = self.nmax
        m = _np.repeat(_np.arange(N + 1, dtype=_np.uint), _np.arange(N + 1, 0, -1))
        n = _np.concatenate([_np.arange(k, N + 1, dtype=_np.uint) for k in range(N + 1)])
        return n, m

    # ------------------------------------------------------------------
    # Copies, arithmetic, spectra, rescaling (wrap/pyharm/shc.py:398-445,
    # 528-573, 707-948, 1360-1485, 1859-2100)
    # ------------------------------------------------------------------
    @classmethod
    def from_copy(cls, shcs, nmin=0, nmax=None, nmax_shcs_out=None):
        """Copy of degrees ``nmin``..``nmax`` (default: all) of ``shcs`` in a
        new instance of degree ``nmax_shcs_out`` (default ``shcs.nmax``),
        zeros elsewhere (charm_shc_copy)."""
        _check_shc(shcs, "shcs")
        nmax = shcs.nmax if nmax is None else int(nmax)
        nout = shcs.nmax if nmax_shcs_out is None else int(nmax_shcs_out)
        nmin = int(nmin)
        if nmin > nmax or nmax > shcs.nmax or nmax > nout or nmin < 0:
            raise ValueError("Invalid range: nmin = %d, nmax = %d, shcs.nmax = %d, "
                             "nmax_shcs_out = %d." % (nmin, nmax, shcs.nmax, nout))
        f = lib().charm_shc_copy
        f.restype = _ct.POINTER(_Shc)
        f.argtypes = [_ct.POINTER(_Shc), _ct.c_ulong, _ct.c_ulong, _ct.c_ulong]
        p = f(shcs._Shc, nmin, nmax, nout)
        if not p:
            raise MemoryError("charm_shc_copy failed.")
        out = cls.__new__(cls)
        out._arrays = None
        out._Shc = p
        return out

    def _binary(self, name, op2, nmin, nmax, inplace):
        from . import _err

        _check_shc(op2, "op2")
        if not isinstance(inplace, bool):
            raise TypeError("'inplace' must be boolean.")
        if nmax is None:
            nmax = self.nmax if inplace else max(self.nmax, op2.nmax)
        nmin, nmax = int(nmin), int(nmax)
        if nmin < 0 or nmax < 0:
            raise ValueError("nmin and nmax must be non-negative.")
        rop = self if inplace else Shc.from_zeros(nmax, self.mu, self.r)
        f = getattr(lib(), "charm_shc_" + name)
        f.restype = None
        f.argtypes = [_ct.POINTER(_Shc)] * 3 + [_ct.c_ulong, _ct.c_ulong, _ct.POINTER(_err._Err)]
        _call(f, rop._Shc, self._Shc, op2._Shc, nmin, nmax)
        return None if inplace else rop

    def add(self, op2, nmin=0, nmax=None, inplace=True):
        """Adds ``op2`` on degrees ``nmin``..``nmax``, zero padding either
        operand beyond its degree; in place (returns None) or into a new
        instance of degree ``nmax`` (default max(self.nmax, op2.nmax))."""
        return self._binary("add", op2, nmin, nmax, inplace)

    def sub(self, op2, nmin=0, nmax=None, inplace=True):
        return self._binary("sub", op2, nmin, nmax, inplace)

    def mul(self, op2, nmin=0, nmax=None, inplace=True):
        return self._binary("mul", op2, nmin, nmax, inplace)

    def div(self, op2, nmin=0, nmax=None, inplace=True):
        return self._binary("div", op2, nmin, nmax, inplace)

    def __add__(self, other):
        return self.add(other, inplace=False)

    def __iadd__(self, other):
        self.add(other, inplace=True)
        return self

    def __sub__(self, other):
        return self.sub(other, inplace=False)

    def __isub__(self, other):
        self.sub(other, inplace=True)
        return self

    def __mul__(self, other):
        return self.mul(other, inplace=False)

    def __imul__(self, other):
        self.mul(other, inplace=True)
        return self

    def __truediv__(self, other):
        return self.div(other, inplace=False)

    def __itruediv__(self, other):
        self.div(other, inplace=True)
        return self

    def _wise(self, name, a, nmin, nmax):
        from . import _err

        a = _np.asarray(a)
        if a.ndim != 1 or not _np.issubdtype(a.dtype, _np.floating):
            raise TypeError("'a' must be a 1D floating point array.")
        a = _np.ascontiguousarray(a, dtype=_np.float64)
        nmin = int(nmin)
        nmax = self.nmax if nmax is None else int(nmax)
        if nmin < 0 or nmax < 0:
            raise ValueError("nmin and nmax must be non-negative.")
        need = nmax - nmin + 1 if "degree" in name else nmax + 1
        if a.size < need:
            raise ValueError("'a' must hold %d factors, not %d." % (need, a.size))
        f = getattr(lib(), "charm_shc_" + name)
        f.restype = None
        f.argtypes = [_ct.POINTER(_Shc), _ct.POINTER(_ct.c_double), _ct.c_ulong, _ct.c_ulong,
                      _ct.POINTER(_err._Err)]
        _call(f, self._Shc, a.ctypes.data_as(_ct.POINTER(_ct.c_double)), nmin, nmax)

    def mul_degree_wise(self, a, nmin=0, nmax=None):
        """Degree n (nmin..nmax) times a[n - nmin]."""
        self._wise("mul_degree_wise", a, nmin, nmax)

    def mul_order_wise(self, a, nmin=0, nmax=None):
        """Order m of degrees nmin..nmax times a[m]."""
        self._wise("mul_order_wise", a, nmin, nmax)

    def div_degree_wise(self, a, nmin=0, nmax=None):
        self._wise("div_degree_wise", a, nmin, nmax)

    def div_order_wise(self, a, nmin=0, nmax=None):
        self._wise("div_order_wise", a, nmin, nmax)

    def rescale(self, mu=None, r=None):
        """Rescales to a new ``mu`` and/or ``r`` (unchanged when None)."""
        from . import _err

        mu = self.mu if mu is None else float(mu)
        r = self.r if r is None else float(r)
        if r <= 0.0:
            raise ValueError("The radius must be positive.")
        f = lib().charm_shc_rescale
        f.restype = None
        f.argtypes = [_ct.POINTER(_Shc), _ct.c_double, _ct.c_double, _ct.POINTER(_err._Err)]
        _call(f, self._Shc, mu, r)

    def _spectrum(self, name, nmax, other=None):
        from . import _err

        out = _np.zeros(nmax + 1, dtype=_np.float64)
        f = getattr(lib(), "charm_shc_" + name)
        f.restype = None
        shcs = [_ct.POINTER(_Shc)] * (1 if other is None else 2)
        f.argtypes = shcs + [_ct.c_ulong, _ct.POINTER(_ct.c_double), _ct.POINTER(_err._Err)]
        args = [self._Shc] + ([] if other is None else [other._Shc])
        _call(f, *args, nmax, out.ctypes.data_as(_ct.POINTER(_ct.c_double)))
        return out

    def _spectrum_nmax(self, nmax, other=None):
        top = self.nmax if other is None else min(self.nmax, other.nmax)
        nmax = top if nmax is None else int(nmax)
        if nmax < 0 or nmax > top:
            raise ValueError("nmax = %d must be within 0..%d." % (nmax, top))
        return nmax

    def dv(self, nmax=None):
        """Degree variances 0..nmax (default self.nmax)."""
        return self._spectrum("dv", self._spectrum_nmax(nmax))

    def da(self, nmax=None):
        """Degree amplitudes 0..nmax."""
        return self._spectrum("da", self._spectrum_nmax(nmax))

    def ddv(self, shcs, nmax=None):
        """Difference degree variances with respect to ``shcs``, 0..nmax
        (default min(self.nmax, shcs.nmax))."""
        _check_shc(shcs, "shcs")
        return self._spectrum("ddv", self._spectrum_nmax(nmax, shcs), shcs)

    def dda(self, shcs, nmax=None):
        """Difference degree amplitudes with respect to ``shcs``."""
        _check_shc(shcs, "shcs")
        return self._spectrum("dda", self._spectrum_nmax(nmax, shcs), shcs)

    def _free(self):
        if getattr(self, "_Shc", None) is not None:
            f = lib().charm_shc_free
            f.restype = None
            f.argtypes = [_ct.POINTER(_Shc)]
            f(self._Shc)
            self._Shc = None

    def __del__(self):
        self._free()


def _check_shc(v, name):
    if not isinstance(v, Shc):
        raise TypeError("'%s' must be an Shc instance." % name)


def _call(f, *args):
    """Calls a charm_shc_* routine whose last argument is a charm_err and
    raises CHarmError on a reported error."""
    from . import _err

    err = _err.init()
    try:
        f(*args, err)
        _err.handler(err, True)
    finally:
        _err.free(err)


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
