"""Device-resident entry points (no counterpart in CHarm).

Inputs and outputs are device tensors (``torch`` on a ROCm build supplies the
HBM allocations and the stream); the library only sees raw pointers.
Coefficients are interleaved (C_nm, S_nm) pairs in CHarm's m-major order,
i.e. a float64 tensor of shape (P, 2), P = (nmax+1)(nmax+2)/2.
"""
import ctypes as _ct

from . import _err
from . import crd as _crd
from ._lib import lib

_vp = _ct.c_void_p


def _stream_ptr(stream):
    if stream is None:
        import torch

        stream = torch.cuda.current_stream()
    return _vp(stream.cuda_stream if hasattr(stream, "cuda_stream") else int(stream))


def _ptr(t):
    if not t.is_contiguous():
        raise ValueError("device tensors must be contiguous")
    return _vp(t.data_ptr())


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
              self._h, int(nmax), float(mu), float(R), _ptr(cs), _ptr(out), _stream_ptr(stream))
        return out

    def sha(self, f, nmax, mu, R, out, stream=None):
        """Analysis of ``f`` (nlat, nlon) into ``out`` (P, 2)."""
        _call("charm_amd_sha_grid_dev",
              [_vp, _ct.c_ulong, _ct.c_double, _ct.c_double, _vp, _vp, _vp],
              self._h, int(nmax), float(mu), float(R), _ptr(f), _ptr(out), _stream_ptr(stream))
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
    """Synthesis at scattered points given as device tensors."""
    _call("charm_amd_shs_sctr_dev",
          [_ct.c_size_t, _vp, _vp, _vp, _ct.c_ulong, _ct.c_double, _ct.c_double, _vp, _vp, _vp],
          lat.numel(), _ptr(lat), _ptr(lon), _ptr(r), int(nmax), float(mu), float(R), _ptr(cs),
          _ptr(out), _stream_ptr(stream))
    return out
