"""ctypes front end of the CPU oracle (``charm_oracle.c``).

TEST INFRASTRUCTURE ONLY: imported by the tests, the smoke check and the CPU
baseline leg of the benchmark, as the checker.  The product never imports it.

Two builds of the same C source are exposed:

* ``Oracle('d')`` -- the algorithm in IEEE double, as CHarm built in double;
* ``Oracle('q')`` -- the algorithm in ``__float128``, as CHarm built in quad.

Floating-point arrays cross the boundary in the oracle's NCOMP layout (see the
header of ``charm_oracle.c``).  For the quad build, inputs may be given as
float64 arrays (exact doubles) or as sequences of ``mpmath.mpf`` values (any
quad value); outputs are returned as ``mpmath.mpf`` arrays (object dtype) when
``exact=True`` and as float64 otherwise.
"""
import ctypes as ct
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_BUILD = os.path.join(_HERE, "_build")

_dp = ct.POINTER(ct.c_double)
_ul = ct.c_ulong
_sz = ct.c_size_t


def build():
    """Compile both oracle builds (gcc; a few seconds)."""
    import subprocess

    subprocess.run(["make", "-s", "-C", _HERE, "all"], check=True)


class Oracle:
    def __init__(self, prec="d"):
        if prec not in ("d", "q"):
            raise ValueError("prec must be 'd' or 'q'")
        path = os.path.join(_BUILD, "liborc_%s.so" % prec)
        if not os.path.exists(path):
            build()
        self.prec = prec
        self.lib = ct.CDLL(path)
        self.ncomp = self.lib.orc_ncomp()
        L = self.lib
        L.orc_gl_grid.argtypes = [_ul, _dp, _dp, _dp]
        L.orc_shs_grid.argtypes = [_ul, _ul, _dp, _dp, _dp, _dp, ct.c_int, _sz, _dp,
                                   _dp, _sz, _dp, ct.POINTER(_sz), _sz, _dp]
        L.orc_lumped.argtypes = [_ul, _ul, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
        L.orc_shs_sctr.argtypes = [_ul, _ul, _dp, _dp, _dp, _dp, _sz, _dp, _dp, _dp, _dp]
        L.orc_sha_gl.argtypes = [_ul, _ul, _sz, _sz, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
        L.orc_dh_grid.argtypes = [_ul, ct.c_int, _dp, _dp, _dp]
        L.orc_sha_dh.argtypes = [ct.c_int, _ul, _ul, _sz, _sz, _dp, _dp, _dp, _dp, _dp, _dp,
                                 _dp, _dp]
        L.orc_leg_pnm.argtypes = [_ul, _dp, _dp]
        L.orc_shs_deriv_sctr.argtypes = [ct.c_int, _ul, _ul, _dp, _dp, _dp, _dp, _sz, _dp, _dp,
                                         _dp, _dp]
        L.orc_set_polar.argtypes = [_ul, ct.c_double, _sz, _sz]

    # -- layout helpers ------------------------------------------------------
    def pack(self, x):
        """Values -> contiguous NCOMP*n float64 buffer."""
        if self.ncomp == 1:
            return np.ascontiguousarray(np.asarray(x, dtype=np.float64).ravel())
        arr = np.asarray(x, dtype=object).ravel() if not isinstance(x, np.ndarray) or x.dtype == object \
            else np.asarray(x).ravel()
        n = arr.size
        out = np.zeros(3 * n, dtype=np.float64)
        if arr.dtype != object:
            out[:n] = arr.astype(np.float64)
            return out
        import mpmath

        with mpmath.workprec(400):
            for i, v in enumerate(arr):
                v = mpmath.mpf(v)
                hi = float(v)
                r1 = v - hi
                mid = float(r1)
                lo = float(r1 - mid)
                out[i], out[n + i], out[2 * n + i] = hi, mid, lo
        return out

    def empty(self, n):
        return np.zeros(self.ncomp * n, dtype=np.float64)

    def unpack(self, y, n, exact=False):
        if self.ncomp == 1:
            return y[:n].copy()
        if not exact:
            return y[:n] + y[n:2 * n] + y[2 * n:3 * n]
        import mpmath

        out = np.empty(n, dtype=object)
        with mpmath.workprec(400):
            for i in range(n):
                out[i] = mpmath.mpf(y[i]) + mpmath.mpf(y[n + i]) + mpmath.mpf(y[2 * n + i])
        return out

    @staticmethod
    def _p(a):
        return a.ctypes.data_as(_dp)

    # -- API -------------------------------------------------------------------
    def set_polar(self, a1, a2, block_shs=32, block_sha=64):
        """Polar optimisation constants (a2 < 0: off, the default) and the
        ring/point blocks over which CHarm tests it jointly: SIMD_SIZE x
        BLOCK_S in synthesis, SIMD_SIZE x BLOCK_A in analysis (32 and 64 for
        the AVX-512 double build)."""
        self.lib.orc_set_polar(int(a1), float(a2), int(block_shs), int(block_sha))

    def gl_grid(self, nmax, exact=False):
        L = nmax + 1
        lat, w, lon = self.empty(L), self.empty(L), self.empty(2 * L)
        rc = self.lib.orc_gl_grid(nmax, self._p(lat), self._p(w), self._p(lon))
        if rc:
            raise RuntimeError("GL iteration failed")
        return (self.unpack(lat, L, exact), self.unpack(w, L, exact),
                self.unpack(lon, 2 * L, exact))

    def dh_grid(self, nmax, dh, exact=False):
        """Driscoll-Healy grid dh = 1 or 2: (lat, w) of 2L rings, lon."""
        L = nmax + 1
        nlat, nlon = 2 * L, (2 if dh == 1 else 4) * L
        lat, w, lon = self.empty(nlat), self.empty(nlat), self.empty(nlon)
        rc = self.lib.orc_dh_grid(nmax, 3 + dh, self._p(lat), self._p(w), self._p(lon))
        if rc:
            raise ValueError("dh must be 1 or 2")
        return (self.unpack(lat, nlat, exact), self.unpack(w, nlat, exact),
                self.unpack(lon, nlon, exact))

    def shs_grid(self, nmax, C, S, mu, R, lat, r, lon, gl=True, rings=None, exact=False,
                 dh=0):
        """Grid synthesis; gl selects a Gauss-Legendre grid, dh = 1 or 2 a
        Driscoll-Healy grid (evaluated ring by ring; dh sets the polar
        optimisation's ring blocks), otherwise a custom grid."""
        nmax_shc = _nmax_from_len(len(C))
        nlat, nlon = len(lat), len(lon)
        Cb, Sb = self.pack(C), self.pack(S)
        mub, Rb = self.pack([mu]), self.pack([R])
        latb, rb, lonb = self.pack(lat), self.pack(r), self.pack(lon)
        if rings is None:
            rp, nring, nout = None, 0, nlat
        else:
            rarr = np.ascontiguousarray(np.asarray(rings, dtype=np.uintp))
            rp, nring, nout = rarr.ctypes.data_as(ct.POINTER(_sz)), rarr.size, rarr.size
        f = self.empty(nout * nlon)
        self.lib.orc_shs_grid(nmax, nmax_shc, self._p(Cb), self._p(Sb), self._p(mub),
                              self._p(Rb), 3 + dh if dh else (1 if gl else 0), nlat,
                              self._p(latb), self._p(rb),
                              nlon, self._p(lonb), rp, nring, self._p(f))
        out = self.unpack(f, nout * nlon, exact)
        return out.reshape(nout, nlon)

    def lumped(self, nmax, C, S, R, lat, r, exact=False):
        nmax_shc = _nmax_from_len(len(C))
        Cb, Sb = self.pack(C), self.pack(S)
        a, b = self.empty(nmax + 1), self.empty(nmax + 1)
        self.lib.orc_lumped(nmax, nmax_shc, self._p(Cb), self._p(Sb), self._p(self.pack([R])),
                            self._p(self.pack([lat])), self._p(self.pack([r])),
                            self._p(a), self._p(b))
        return self.unpack(a, nmax + 1, exact), self.unpack(b, nmax + 1, exact)

    def shs_sctr(self, nmax, C, S, mu, R, lat, lon, r, exact=False):
        nmax_shc = _nmax_from_len(len(C))
        n = len(lat)
        f = self.empty(n)
        args = [self.pack(v) for v in (C, S, [mu], [R], lat, lon, r)]
        self.lib.orc_shs_sctr(nmax, nmax_shc, *(self._p(a) for a in args[:4]), n,
                              *(self._p(a) for a in args[4:]), self._p(f))
        return self.unpack(f, n, exact)

    def sha_gl(self, nmax, f, lat, w, r0, mu, R, nmax_shc=None, exact=False):
        f = np.asarray(f)
        nlat, nlon = f.shape
        nmax_shc = nmax if nmax_shc is None else nmax_shc
        ncs = (nmax_shc + 1) * (nmax_shc + 2) // 2
        C, S = self.empty(ncs), self.empty(ncs)
        args = [self.pack(v) for v in (f, lat, w, [r0], [mu], [R])]
        rc = self.lib.orc_sha_gl(nmax, nmax_shc, nlat, nlon, *(self._p(a) for a in args),
                                 self._p(C), self._p(S))
        if rc:
            raise ValueError("invalid SHA input")
        return self.unpack(C, ncs, exact), self.unpack(S, ncs, exact)

    def sha_dh(self, dh, nmax, f, lat, w, r0, mu, R, nmax_shc=None, exact=False):
        """Analysis on a Driscoll-Healy grid (dh = 1 or 2)."""
        f = np.asarray(f)
        nlat, nlon = f.shape
        nmax_shc = nmax if nmax_shc is None else nmax_shc
        ncs = (nmax_shc + 1) * (nmax_shc + 2) // 2
        C, S = self.empty(ncs), self.empty(ncs)
        args = [self.pack(v) for v in (f, lat, w, [r0], [mu], [R])]
        rc = self.lib.orc_sha_dh(3 + dh, nmax, nmax_shc, nlat, nlon,
                                 *(self._p(a) for a in args), self._p(C), self._p(S))
        if rc:
            raise ValueError("invalid SHA input")
        return self.unpack(C, ncs, exact), self.unpack(S, ncs, exact)

    # g_ijk order of orc_shs_deriv_sctr
    GURU = ((0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (2, 0, 0), (0, 2, 0), (0, 0, 2),
            (1, 1, 0), (1, 0, 1), (0, 1, 1))

    def shs_deriv_sctr(self, what, nmax, C, S, mu, R, lat, lon, r, exact=False):
        """Derivatives at scattered points: what = (dr, dlat, dlon) for one
        g_ijk, or "grad1" (fx, fy, fz), or "grad2" (fxx, fxy, fxz, fyy, fyz,
        fzz).  Returns an array of shape (nout, npoint)."""
        if what == "grad1":
            kind, nout = 10, 3
        elif what == "grad2":
            kind, nout = 11, 6
        else:
            kind, nout = self.GURU.index(tuple(what)), 1
        nmax_shc = _nmax_from_len(len(C))
        n = len(lat)
        f = self.empty(nout * n)
        args = [self.pack(v) for v in (C, S, [mu], [R], lat, lon, r)]
        rc = self.lib.orc_shs_deriv_sctr(kind, nmax, nmax_shc, *(self._p(a) for a in args[:4]),
                                         n, *(self._p(a) for a in args[4:]), self._p(f))
        if rc:
            raise ValueError("invalid derivative kind")
        return self.unpack(f, nout * n, exact).reshape(nout, n)

    def shs_deriv_grid(self, what, nmax, C, S, mu, R, lat, lon, r, exact=False):
        """shs_deriv_sctr on every point of a grid (rows lat with radii r,
        columns lon); returns (nout, nlat, nlon)."""
        nlat, nlon = len(lat), len(lon)
        la = np.repeat(np.asarray(lat, dtype=object if exact else None), nlon)
        lo = np.tile(np.asarray(lon, dtype=object if exact else None), nlat)
        ra = np.repeat(np.asarray(r, dtype=object if exact else None), nlon)
        out = self.shs_deriv_sctr(what, nmax, C, S, mu, R, la, lo, ra, exact)
        return out.reshape(out.shape[0], nlat, nlon)

    def gfm_global(self, shape, shape_nmax, R, density, G, M, pmin, pmax, nmax_pot,
                   exact=False):
        """Global spectral gravity forward modelling (src/gfm/gfm_sgfm.c with
        SGFM_GLOBAL, :777-1060).  shape = (C, S) of the shape (mu = r = 1);
        density = a number (constant density) or a list of (C, S, nmax), the
        terms rho_i of rho = sum_i rho_i r^i.  Returns (C, S) of the potential
        to degree nmax_pot (mu = G M, r = R)."""
        import math

        const = not isinstance(density, (list, tuple))
        terms = [None] if const else list(density)
        dmax = 0 if const else max(t[2] for t in terms)
        ncs = (nmax_pot + 1) * (nmax_pot + 2) // 2
        if exact:
            import mpmath

            mp = mpmath.mp
            mp.prec = 113
            num = mpmath.mpf
            PI = mpmath.pi
        else:
            num = float
            PI = math.pi
        Cp = np.zeros(ncs, dtype=object if exact else np.float64)
        Sp = np.zeros(ncs, dtype=object if exact else np.float64)
        R = num(R)
        c = (2 * PI * R ** 3) / num(M)
        if const:
            c = c * num(density)
        for p in range(pmin, pmax + 1):
            nmaxp = p * shape_nmax + dmax
            thr = min(nmaxp, nmax_pot)
            lat, w, lon = self.gl_grid(nmaxp, exact=exact)
            ones = np.array([num(1)] * len(lat), dtype=object if exact else np.float64)
            H = self.shs_grid(shape_nmax, shape[0], shape[1], 1.0, 1.0, lat, ones, lon,
                              exact=exact)
            for i, term in enumerate(terms):
                h = (H - R) / R
                v = h.copy()
                for _ in range(2, p + 1):
                    v = v * h
                if term is None:
                    f = v
                else:
                    rho = self.shs_grid(term[2], term[0], term[1], 1.0, 1.0, lat, ones, lon,
                                        exact=exact)
                    f = rho * (v * R ** i)
                C, S = self.sha_gl(thr, f, lat, w, 1.0, 1.0, 1.0, exact=exact)
                pr = num(p)
                for n in range(thr + 1):
                    np3 = num(n + i + 3)
                    snp = np3 / pr
                    for k in range(1, p):
                        snp = snp * ((np3 - k) / (pr - k))
                    snp = snp * (num(2) / ((2 * num(n) + 1) * np3))
                    vn = c * snp
                    for m in range(n + 1):
                        a = m * (thr + 1) - (m * (m - 1)) // 2 + (n - m)
                        b = m * (nmax_pot + 1) - (m * (m - 1)) // 2 + (n - m)
                        Cp[b] = Cp[b] + C[a] * vn
                        Sp[b] = Sp[b] + S[a] * vn
        return Cp, Sp

    def leg_pnm(self, nmax, lat, exact=False):
        ncs = (nmax + 1) * (nmax + 2) // 2
        p = self.empty(ncs)
        self.lib.orc_leg_pnm(nmax, self._p(self.pack([lat])), self._p(p))
        return self.unpack(p, ncs, exact)


def _nmax_from_len(ncs):
    n = int(round(((8 * ncs + 1) ** 0.5 - 3) / 2))
    if (n + 1) * (n + 2) // 2 != ncs:
        raise ValueError("coefficient array length %d is not triangular" % ncs)
    return n


def tri_index(nmax, n, m):
    """Index of (n, m) in the m-major triangle of degree nmax."""
    return m * (nmax + 1) - (m * (m - 1)) // 2 + (n - m)
