"""The plain API is reentrant: charm_shs_point / charm_sha_point called from
several host threads at once (ctypes releases the GIL during the calls), at
different degrees, on grids and scattered points, give exactly what each
call gives alone.  CHarm itself is not reentrant (it creates and cleans up
FFTW state inside every call, src/shs/shs_point_grd.c:390-433,1037-1045);
the library keeps one stream, FFT plans and staging buffers per thread."""
import threading

import numpy as np
import pytest

from charm_amd import crd, sha, shc, shs
from charm_amd.workloads import GM, R_EARTH, kaula_coeffs, scattered_points

pytestmark = pytest.mark.gpu


def jobs():
    out = []
    for i, N in enumerate((180, 517, 1024, 2160)):
        C, S = kaula_coeffs(N, 900 + i)
        m = shc.Shc.from_arrays(N, C, S, GM, R_EARTH)
        g = crd.PointGridGL(N, R_EARTH + 100.0 * i)
        out.append(("grid", N, m, g))
        lat, lon, r = scattered_points(20000 + 5000 * i, 950 + i)
        out.append(("sctr", N, m, crd.PointSctr(lat, lon, r)))
    return out


def run(job, f_in=None):
    kind, N, m, p = job
    f = shs.point(p, m, N)
    if kind == "grid":
        o = sha.point(p, f, N, GM, R_EARTH)
        return f, np.stack([o.c, o.s])
    return f, None


def test_concurrent_plain_calls_equal_solo():
    js = jobs()
    solo = [run(j) for j in js]
    results = [None] * len(js)
    errors = []

    def worker(idx):
        try:
            for rep in range(3):
                for i in idx:
                    results[i] = run(js[i])
                    f, c = results[i]
                    assert np.array_equal(f, solo[i][0]), (i, rep)
                    if c is not None:
                        assert np.array_equal(c, solo[i][1]), (i, rep)
        except Exception as e:  # reported in the main thread
            errors.append(e)

    threads = [threading.Thread(target=worker, args=([k, (k + 4) % 8],)) for k in range(4)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors, errors
    assert all(r is not None for r in results)
