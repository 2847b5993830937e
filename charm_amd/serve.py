"""Request serving over the C ABI's server (charm_amd_server_*).

A worker thread in the library admits requests under an HBM budget (in
arrival order; a request that does not fit yet lets at most 8 later ones
pass), overlaps them on the device (one stream per launch in flight), runs
small scattered-point requests admitted together as one persistent launch and
completes them asynchronously; models and Gauss-Legendre grids stay resident.
Submitting copies a request's inputs, so the caller may reuse its arrays at
once; the output array is filled by :meth:`Ticket.wait`.
"""
import ctypes as _ct
import threading as _th

import numpy as _np

from . import _err
from . import shc as _shc
from ._lib import lib

_vp = _ct.c_void_p
_dp = _ct.POINTER(_ct.c_double)

SHS_SCTR = 0
SHS_GRID = 1
SHA_GRID = 2


class _SrvStats(_ct.Structure):
    _fields_ = [(n, _ct.c_ulong) for n in ("submitted", "completed", "failed", "cancelled",
                                           "launches", "grouped", "max_group")]


class _SrvRequest(_ct.Structure):
    _fields_ = [("kind", _ct.c_int),
                ("model", _ct.c_int),
                ("nmax", _ct.c_ulong),
                ("npoint", _ct.c_size_t),
                ("lat", _dp), ("lon", _dp), ("r", _dp),
                ("grid_nmax", _ct.c_ulong),
                ("r_grid", _ct.c_double),
                ("f_in", _dp),
                ("mu", _ct.c_double), ("R", _ct.c_double),
                ("out", _dp)]


def _p(a):
    return a.ctypes.data_as(_dp) if a is not None else None


class Ticket:
    def __init__(self, server, tid, out):
        self.server, self.id, self.out = server, tid, out
        self.latency_ms = None
        self._done = False

    def wait(self):
        """Block until done; returns the output array.  A failed or cancelled
        request raises :class:`charm_amd.CHarmError` with the library's code."""
        if not self._done:
            self._done = True
            f = lib().charm_amd_server_wait
            f.restype = _ct.c_int
            f.argtypes = [_vp, _ct.c_long, _ct.POINTER(_ct.c_double), _ct.POINTER(_err._Err)]
            lat = _ct.c_double()
            err = _err.init()
            try:
                f(self.server._h, self.id, _ct.byref(lat), err)
                self.latency_ms = lat.value
                _err.handler(err, True)
            finally:
                _err.free(err)
        return self.out

    def cancel(self):
        """Cancel: a queued request is dropped, an in-flight one has its
        result discarded.  Returns False if the request had already completed."""
        f = lib().charm_amd_server_cancel
        f.restype = _ct.c_int
        f.argtypes = [_vp, _ct.c_long, _ct.POINTER(_err._Err)]
        err = _err.init()
        try:
            rc = f(self.server._h, self.id, err)
            _err.handler(err, True)
        finally:
            _err.free(err)
        return rc == 0


class Server:
    def __init__(self, hbm_budget_bytes=200 * 2**30, max_inflight=4):
        f = lib().charm_amd_server_create
        f.restype = _vp
        f.argtypes = [_ct.c_size_t, _ct.c_int, _ct.POINTER(_err._Err)]
        err = _err.init()
        try:
            h = f(int(hbm_budget_bytes), int(max_inflight), err)
            _err.handler(err, True)
        finally:
            _err.free(err)
        self._h = _vp(h)
        self._lock = _th.Lock()

    def add_model(self, shcs, nmax=None):
        """Make ``shcs`` (an :class:`~charm_amd.shc.Shc`) resident; returns its id."""
        nmax = shcs.nmax if nmax is None else int(nmax)
        f = lib().charm_amd_server_add_model
        f.restype = _ct.c_int
        f.argtypes = [_vp, _ct.POINTER(_shc._Shc), _ct.c_ulong, _ct.POINTER(_err._Err)]
        err = _err.init()
        try:
            mid = f(self._h, shcs._Shc, nmax, err)
            _err.handler(err, True)
        finally:
            _err.free(err)
        return mid

    def _submit(self, req, out):
        f = lib().charm_amd_server_submit
        f.restype = _ct.c_long
        f.argtypes = [_vp, _ct.POINTER(_SrvRequest), _ct.POINTER(_err._Err)]
        err = _err.init()
        try:
            tid = f(self._h, _ct.byref(req), err)
            _err.handler(err, True)
        finally:
            _err.free(err)
        return Ticket(self, tid, out)

    def shs_sctr(self, model, nmax, lat, lon, r):
        lat, lon, r = (_np.ascontiguousarray(a, dtype=_np.float64) for a in (lat, lon, r))
        out = _np.empty(lat.size)
        req = _SrvRequest(SHS_SCTR, int(model), int(nmax), lat.size, _p(lat), _p(lon), _p(r),
                          0, 0.0, None, 0.0, 0.0, _p(out))
        return self._submit(req, out)

    def shs_grid(self, model, nmax, grid_nmax, r_grid):
        out = _np.empty((grid_nmax + 1, 2 * (grid_nmax + 1)))
        req = _SrvRequest(SHS_GRID, int(model), int(nmax), 0, None, None, None, int(grid_nmax),
                          float(r_grid), None, 0.0, 0.0, _p(out))
        return self._submit(req, out)

    def sha_grid(self, f_grid, nmax, r_grid, mu, R):
        f_grid = _np.ascontiguousarray(f_grid, dtype=_np.float64)
        grid_nmax = f_grid.shape[0] - 1
        P = (nmax + 1) * (nmax + 2) // 2
        out = _np.empty((P, 2))
        req = _SrvRequest(SHA_GRID, -1, int(nmax), 0, None, None, None, int(grid_nmax),
                          float(r_grid), _p(f_grid), float(mu), float(R), _p(out))
        return self._submit(req, out)

    def inject_error(self, count, code):
        """Make the next ``count`` launches fail with CHARM error ``code``, as a
        device error would (fault injection from the host)."""
        f = lib().charm_amd_server_inject_error
        f.restype = None
        f.argtypes = [_vp, _ct.c_int, _ct.c_int]
        f(self._h, int(count), int(code))

    def stats(self):
        """Counters: submitted, completed, failed, cancelled, launches,
        grouped (requests that ran in a group of several), max_group."""
        f = lib().charm_amd_server_get_stats
        f.restype = None
        f.argtypes = [_vp, _ct.POINTER(_SrvStats)]
        st = _SrvStats()
        f(self._h, _ct.byref(st))
        return {n: getattr(st, n) for n, _ in _SrvStats._fields_}

    def in_use(self):
        f = lib().charm_amd_server_in_use
        f.restype = _ct.c_size_t
        f.argtypes = [_vp]
        return f(self._h)

    def close(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            f = lib().charm_amd_server_free
            f.restype = None
            f.argtypes = [_vp]
            f(h)
            self._h = None

    def __del__(self):
        self.close()
