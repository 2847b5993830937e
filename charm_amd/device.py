"""Device-resident entry points (no counterpart in CHarm).

Device memory comes from the same HIP runtime the library links
(libamdhip64.so.7 from ROCm), through :class:`DevArray`; the library only sees
raw pointers and a stream handle.  Coefficients are interleaved
(C_nm, S_nm) pairs in CHarm's m-major order, i.e. float64 arrays of shape
(P, 2), P = (nmax+1)(nmax+2)/2.
"""
import ctypes as _ct

import numpy as _np

from . import _err
from . import crd as _crd
from ._lib import lib

_vp = _ct.c_void_p
_hip = None


def hip():
    """The HIP runtime the library is linked against."""
    global _hip
    if _hip is None:
        lib()  # loads libamdhip64.so.7 as a dependency
        h = _ct.CDLL("libamdhip64.so.7")
        h.hipMalloc.argtypes = [_ct.POINTER(_vp), _ct.c_size_t]
        h.hipFree.argtypes = [_vp]
        h.hipMemcpy.argtypes = [_vp, _vp, _ct.c_size_t, _ct.c_int]
        h.hipMemsetAsync.argtypes = [_vp, _ct.c_int, _ct.c_size_t, _vp]
        h.hipStreamCreate.argtypes = [_ct.POINTER(_vp)]
        h.hipStreamSynchronize.argtypes = [_vp]
        h.hipStreamDestroy.argtypes = [_vp]
        h.hipGetErrorString.restype = _ct.c_char_p
        _hip = h
    return _hip


def _check(rc, what):
    if rc != 0:
        raise RuntimeError("%s failed: %s" % (what, hip().hipGetErrorString(rc).decode()))


def synchronize():
    _check(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")


class Stream:
    """A HIP stream of the library's runtime (default priority)."""

    def __init__(self):
        self.handle = _vp()
        _check(hip().hipStreamCreate(_ct.byref(self.handle)), "hipStreamCreate")

    def synchronize(self):
        _check(hip().hipStreamSynchronize(self.handle), "hipStreamSynchronize")

    def __del__(self):
        if getattr(self, "handle", None) and self.handle.value:
            hip().hipStreamDestroy(self.handle)
            self.handle = None


class DevArray:
    """A float64 array in device memory."""

    def __init__(self, shape, dtype=_np.float64):
        self.shape = tuple(int(s) for s in (shape if hasattr(shape, "__len__") else (shape,)))
        self.dtype = _np.dtype(dtype)
        self.nbytes = int(_np.prod(self.shape)) * self.dtype.itemsize
        self.ptr = _vp()
        _check(hip().hipMalloc(_ct.byref(self.ptr), max(self.nbytes, 1)), "hipMalloc")

    @classmethod
    def from_numpy(cls, a):
        a = _np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        _check(hip().hipMemcpy(d.ptr, a.ctypes.data_as(_vp), d.nbytes, 1), "hipMemcpy H2D")
        return d

    def copy_from_numpy(self, a):
        """Overwrite the contents with the host array ``a`` (same size)."""
        a = _np.ascontiguousarray(a, dtype=self.dtype)
        if a.nbytes != self.nbytes:
            raise ValueError("size mismatch: %d bytes into %d" % (a.nbytes, self.nbytes))
        _check(hip().hipMemcpy(self.ptr, a.ctypes.data_as(_vp), self.nbytes, 1), "hipMemcpy H2D")

    def numpy(self):
        a = _np.empty(self.shape, dtype=self.dtype)
        _check(hip().hipMemcpy(a.ctypes.data_as(_vp), self.ptr, self.nbytes, 2), "hipMemcpy D2H")
        return a

    @property
    def size(self):
        return int(_np.prod(self.shape))

    def __del__(self):
        if getattr(self, "ptr", None) is not None and self.ptr.value:
            hip().hipFree(self.ptr)
            self.ptr = None


def _s(stream):
    return stream.handle if isinstance(stream, Stream) else _vp(stream or 0)


def _call(name, argtypes, *args):
    f = getattr(lib(), name)
    f.restype = None
    f.argtypes = argtypes + [_ct.POINTER(_err._Err)]
    err = _err.init()
    try:
        f(*args, err)
        _err.handler(err, True)
    finally:
        _err.free(err)


class DeviceGrid:
    """A Gauss-Legendre grid with its latitude tables resident in HBM."""

    def __init__(self, nmax, r=1.0):
        self.pnt = _crd.PointGridGL(nmax, r)
        self.nmax = int(nmax)
        self.nlat, self.nlon = _crd.PointGridGL.shape(nmax)
        f = lib().charm_amd_grid_create
        f.restype = _vp
        f.argtypes = [_ct.POINTER(_crd._Point), _ct.POINTER(_err._Err)]
        err = _err.init()
        try:
            h = f(self.pnt._Point, err)
            _err.handler(err, True)
        finally:
            _err.free(err)
        self._h = _vp(h)

    def shs(self, cs, nmax, mu, R, out, stream=None):
        """Synthesis into ``out`` (nlat, nlon) from ``cs`` (P, 2)."""
        _call("charm_amd_shs_grid_dev",
              [_vp, _ct.c_ulong, _ct.c_double, _ct.c_double, _vp, _vp, _vp],
              self._h, int(nmax), float(mu), float(R), cs.ptr, out.ptr, _s(stream))
        return out

    def sha(self, f, nmax, mu, R, out, stream=None):
        """Analysis of ``f`` (nlat, nlon) into ``out`` (P, 2)."""
        _call("charm_amd_sha_grid_dev",
              [_vp, _ct.c_ulong, _ct.c_double, _ct.c_double, _vp, _vp, _vp],
              self._h, int(nmax), float(mu), float(R), f.ptr, out.ptr, _s(stream))
        return out

    def shs_multi(self, cs, nmax, nset, mu, R, out, stream=None):
        """Synthesis of ``nset`` coefficient sets ``cs`` (nset, P, 2) into
        ``out`` (nset, nlat, nlon), the sums over degrees on the FP64 matrix
        cores; the grid radius must equal ``R``."""
        _call("charm_amd_shs_grid_multi_dev",
              [_vp, _ct.c_ulong, _ct.c_size_t, _ct.c_double, _ct.c_double, _vp, _vp, _vp],
              self._h, int(nmax), int(nset), float(mu), float(R), cs.ptr, out.ptr, _s(stream))
        return out

    def shs_deriv(self, cs, nmax, mu, R, what, out, stream=None):
        """Gradient or guru derivative into ``out`` (nout, nlat, nlon):
        ``what`` is DERIV_GRAD1, DERIV_GRAD2 or 100 dr + 10 dlat + dlon."""
        _call("charm_amd_shs_grid_deriv_dev",
              [_vp, _ct.c_ulong, _ct.c_double, _ct.c_double, _vp, _ct.c_int, _vp, _vp],
              self._h, int(nmax), float(mu), float(R), cs.ptr, int(what), out.ptr, _s(stream))
        return out

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            f = lib().charm_amd_grid_free
            f.restype = None
            f.argtypes = [_vp]
            f(h)
            self._h = None


def shs_sctr(lat, lon, r, cs, nmax, mu, R, out, stream=None):
    """Synthesis at scattered points given as device arrays."""
    _call("charm_amd_shs_sctr_dev",
          [_ct.c_size_t, _vp, _vp, _vp, _ct.c_ulong, _ct.c_double, _ct.c_double, _vp, _vp, _vp],
          lat.size, lat.ptr, lon.ptr, r.ptr, int(nmax), float(mu), float(R), cs.ptr, out.ptr,
          _s(stream))
    return out


def shs_sctr_deriv(lat, lon, r, cs, nmax, mu, R, what, out, stream=None):
    """Gradient or guru derivative at scattered points into ``out``
    (nout, npoint)."""
    _call("charm_amd_shs_sctr_deriv_dev",
          [_ct.c_size_t, _vp, _vp, _vp, _ct.c_ulong, _ct.c_double, _ct.c_double, _vp, _ct.c_int,
           _vp, _vp],
          lat.size, lat.ptr, lon.ptr, r.ptr, int(nmax), float(mu), float(R), cs.ptr, int(what),
          out.ptr, _s(stream))
    return out


DERIV_GRAD1 = -1
DERIV_GRAD2 = -2
DERIV_NOUT = {DERIV_GRAD1: 3, DERIV_GRAD2: 6}

SHS_GRID = 0
SHA_GRID = 1
SHS_SCTR = 2


class _Request(_ct.Structure):
    _fields_ = [("kind", _ct.c_int),
                ("nmax", _ct.c_ulong),
                ("mu", _ct.c_double),
                ("R", _ct.c_double),
                ("grid", _vp),
                ("cs", _vp),
                ("f", _vp),
                ("npoint", _ct.c_size_t),
                ("lat", _vp),
                ("lon", _vp),
                ("r", _vp)]


class Batch:
    """Many independent grid requests executed together (one persistent
    launch per transform kind), optionally replayed as a HIP graph.

    ``requests``: sequence of (kind, nmax, mu, R, where, cs DevArray,
    f DevArray) with kind SHS_GRID (cs -> f on the DeviceGrid ``where``),
    SHA_GRID (f -> cs on the DeviceGrid ``where``) or SHS_SCTR (cs -> f at the
    points ``where`` = (lat, lon, r) DevArrays, read once at creation)."""

    def __init__(self, requests, use_graph=True):
        self._keep = list(requests)
        arr = (_Request * len(self._keep))()
        for i, (kind, nmax, mu, R, where, cs, f) in enumerate(self._keep):
            if int(kind) == SHS_SCTR:
                lat, lon, r = where
                arr[i] = _Request(int(kind), int(nmax), float(mu), float(R), None, cs.ptr, f.ptr,
                                  lat.size, lat.ptr, lon.ptr, r.ptr)
            else:
                arr[i] = _Request(int(kind), int(nmax), float(mu), float(R), where._h, cs.ptr,
                                  f.ptr, 0, None, None, None)
        fn = lib().charm_amd_batch_create
        fn.restype = _vp
        fn.argtypes = [_ct.POINTER(_Request), _ct.c_size_t, _ct.c_int, _ct.POINTER(_err._Err)]
        err = _err.init()
        try:
            h = fn(arr, len(self._keep), 1 if use_graph else 0, err)
            _err.handler(err, True)
        finally:
            _err.free(err)
        self._h = _vp(h)

    def run(self, stream=None):
        _call("charm_amd_batch_run", [_vp, _vp], self._h, _s(stream))

    @property
    def uses_graph(self):
        f = lib().charm_amd_batch_uses_graph
        f.restype = _ct.c_int
        f.argtypes = [_vp]
        return bool(f(self._h))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            f = lib().charm_amd_batch_free
            f.restype = None
            f.argtypes = [_vp]
            f(h)
            self._h = None
