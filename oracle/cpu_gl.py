"""ctypes front end of ``cpu_gl.c``, the CPU restatement of CHarm's
Gauss-Legendre synthesis and analysis that the benchmark times as its CPU
baseline (symmetry fold, ring blocks advanced in vectorised loops, OpenMP
threads, FFTs along the parallels by ``scipy.fft``).

TEST INFRASTRUCTURE ONLY: imported by the tests and the CPU baseline leg of the
benchmark.  The product never imports it.
"""
import ctypes as ct
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_BUILD = os.path.join(_HERE, "_build")
_dp = ct.POINTER(ct.c_double)
_ul, _sz = ct.c_ulong, ct.c_size_t


def host_isa():
    """'avx512' when the host CPU has AVX-512F/DQ/VL, else 'avx2'."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("flags"):
                    fl = set(line.split(":", 1)[1].split())
                    return "avx512" if {"avx512f", "avx512dq", "avx512vl"} <= fl else "avx2"
    except OSError:
        pass
    return "avx2"


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def physical_cores():
    """Physical cores among the CPUs this process may run on (one per
    (package, core) pair in /sys), or the CPU count if that is unreadable."""
    cpus = sorted(os.sched_getaffinity(0))
    seen = set()
    for c in cpus:
        base = "/sys/devices/system/cpu/cpu%d/topology/" % c
        try:
            with open(base + "physical_package_id") as f:
                pkg = f.read().strip()
            with open(base + "core_id") as f:
                core = f.read().strip()
        except OSError:
            return len(cpus)
        seen.add((pkg, core))
    return len(seen) or len(cpus)


class CpuGL:
    def __init__(self, threads, isa=None):
        isa = isa or host_isa()
        path = os.path.join(_BUILD, "libcpu_gl_%s.so" % isa)
        if not os.path.exists(path):
            import subprocess

            subprocess.run(["make", "-s", "-C", _HERE, "all"], check=True)
        self.isa, self.threads = isa, int(threads)
        L = self.lib = ct.CDLL(path)
        L.cgl_set_threads.argtypes = [ct.c_int]
        L.cgl_shs_spectra.argtypes = [_ul, _dp, _dp, ct.c_double, ct.c_double, ct.c_double, _sz,
                                      _dp, _sz, _dp]
        L.cgl_sha_spectra.argtypes = [_ul, _sz, _sz, _dp, _dp, _dp, ct.c_double, ct.c_double,
                                      ct.c_double, _dp, _dp]
        L.cgl_set_threads(self.threads)

    @staticmethod
    def _p(a):
        return a.ctypes.data_as(_dp)

    def shs(self, nmax, C, S, mu, R, r, lat, nlon):
        """Synthesis on the GL grid with latitudes ``lat`` (north first) and
        ``nlon`` longitudes 2 pi j / nlon at radius r: an (nlat, nlon) array."""
        import scipy.fft

        C, S, lat = (np.ascontiguousarray(a, dtype=np.float64) for a in (C, S, lat))
        nspec = nlon // 2 + 1
        spec = np.empty((len(lat), nspec), dtype=np.complex128)
        rc = self.lib.cgl_shs_spectra(nmax, self._p(C), self._p(S), mu, R, r, len(lat),
                                      self._p(lat), nspec, spec.ctypes.data_as(_dp))
        if rc:
            raise ValueError("invalid synthesis input")
        return scipy.fft.irfft(spec, n=nlon, axis=1, norm="forward", workers=self.threads)

    def sha(self, nmax, f, lat, w, r0, mu, R):
        """Analysis of the (nlat, nlon) GL grid values f: (C, S) m-major
        triangles of degree nmax."""
        import scipy.fft

        f, lat, w = (np.ascontiguousarray(a, dtype=np.float64) for a in (f, lat, w))
        spec = np.ascontiguousarray(scipy.fft.rfft(f, axis=1, workers=self.threads))
        P = (nmax + 1) * (nmax + 2) // 2
        C, S = np.empty(P), np.empty(P)
        rc = self.lib.cgl_sha_spectra(nmax, f.shape[0], spec.shape[1], spec.ctypes.data_as(_dp),
                                      self._p(lat), self._p(w), r0, mu, R, self._p(C), self._p(S))
        if rc:
            raise ValueError("invalid analysis input")
        return C, S
