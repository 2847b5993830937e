"""Shared test helpers: fixture readers and the inputs of CHarm's own tests.

The fixture files under ``tests/golden/charm_data/{double,quad}`` are data
files copied from CHarm 0.4.9's ``data/tests`` folder (outputs of CHarm built
in double and in quad precision).  The input models are CHarm's
``data/input/EGM96-degree10-mtx.txt`` and, for the degree variance tests,
``data/input/Earth2014-RET2014-degree4-mtx.txt``.
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
DATA = os.path.join(GOLDEN, "charm_data")

DBL_EPS = np.finfo(np.float64).eps
QUAD_EPS = 2.0 ** -112
# glob_threshold = 100 eps, glob_threshold2 = 1e5 eps (src/glob/glob.c:16-17)
THRESH = {"double": 100 * DBL_EPS, "quad": 100 * QUAD_EPS}
THRESH2 = {"double": 1e5 * DBL_EPS, "quad": 1e5 * QUAD_EPS}
# tests/parameters.h:167-174, 222-229, 239-240
DELTAR = 1000.0
C00, C10, C11, S11 = 0.0000001, 0.0000002, 0.0000003, 0.0000004
BREAK_SYMM = 0.001


def tri_index(nmax, n, m):
    return m * (nmax + 1) - (m * (m - 1)) // 2 + (n - m)


def read_fixture(prec, name, exact=False):
    """Read a golden text file: float64 array, or mpf values if exact."""
    path = os.path.join(DATA, prec, name)
    with open(path) as fh:
        toks = fh.read().split()
    if not exact:
        return np.array([float(t) for t in toks])
    import mpmath

    with mpmath.workprec(113):
        return np.array([mpmath.mpf(t) for t in toks], dtype=object)


def read_mtx(path, nmax, exact=False):
    """Coefficients of an ``mtx`` file (src/shc/shc_read_mtx.c:150-250).

    Entry (row, col) is C_{row,col} when row >= col, else S_{col,row+1}.
    Returns (mu, R, C, S) with C, S as m-major triangles of degree nmax.
    Quad values are rounded to 113 bits like strtoflt128.
    """
    with open(path) as fh:
        head = fh.readline().split()
        rows = [ln.split() for ln in fh if ln.strip()]
    conv = float
    if exact:
        import mpmath

        def conv(s):
            with mpmath.workprec(113):
                return +mpmath.mpf(s)
    mu, R = conv(head[1]), conv(head[2])
    ncs = (nmax + 1) * (nmax + 2) // 2
    C = np.zeros(ncs, dtype=object if exact else np.float64)
    S = np.zeros(ncs, dtype=object if exact else np.float64)
    for row in range(nmax + 1):
        for col in range(nmax + 1):
            v = conv(rows[row][col])
            if row >= col:
                C[tri_index(nmax, row, col)] = v
            else:
                S[tri_index(nmax, col, row + 1)] = v
    return mu, R, C, S


def egm96_test_model(exact=False):
    """EGM96 to degree 10 with CHarm's modified low degrees
    (tests/modify_low_degree_coefficients.c:20-34).  In the quad build the
    replaced values are the double literals converted to quad."""
    mu, R, C, S = read_mtx(os.path.join(GOLDEN, "EGM96-degree10-mtx.txt"), 10, exact)
    nm = 10
    C[tri_index(nm, 0, 0)] = C00
    C[tri_index(nm, 1, 0)] = C10
    C[tri_index(nm, 1, 1)] = C11
    S[tri_index(nm, 1, 1)] = S11
    if exact:
        import mpmath

        with mpmath.workprec(113):
            for arr in (C, S):
                for i in range(arr.size):
                    arr[i] = +mpmath.mpf(arr[i])
    return mu, R, C, S


def truncate(C, nmax_from, nmax_to):
    """m-major triangle of degree nmax_from cut to degree nmax_to."""
    out = []
    for m in range(nmax_to + 1):
        for n in range(m, nmax_to + 1):
            out.append(C[tri_index(nmax_from, n, m)])
    return np.array(out, dtype=C.dtype)


def scattered_points(nlat, prec):
    """tests/generate_point.c:16-60 for scattered points with lat_rng = pi,
    lon_rng = 2 pi.  Arithmetic in the precision of the build."""
    if prec == "double":
        PI = np.pi
        lat = np.array([np.pi / 2 - (float(l) / float(nlat)) * PI for l in range(nlat)])
        lon = np.array([(float(l) / float(nlat)) * (2.0 * PI) for l in range(nlat)])
        return lat, lon
    import mpmath

    with mpmath.workprec(113):
        PI = +mpmath.pi
        PI_2 = PI / 2
        two_pi = 2 * PI
        lat = np.array([PI_2 - (mpmath.mpf(l) / nlat) * PI for l in range(nlat)], dtype=object)
        lon = np.array([(mpmath.mpf(l) / nlat) * two_pi for l in range(nlat)], dtype=object)
    return lat, lon


def nearly_equal(x, ref, eps):
    """Elementwise misc_is_nearly_equal (src/misc/misc_is_nearly_equal.c)."""
    x = np.asarray(x, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    ax, ar = np.abs(x), np.abs(ref)
    d = np.abs(x - ref)
    s = ax + ar
    big = np.maximum(ax, ar)
    with np.errstate(divide="ignore", invalid="ignore"):
        rel = np.where(big > 0, d / big, 0.0)
    absolute = (x == 0) | (ref == 0) | (s == ax) | (s == ar)
    ok = (x == ref) | np.where(absolute, d <= eps, rel <= eps)
    return ok


def nearly_equal_exact(x, ref, eps):
    """misc_is_nearly_equal on mpf values (quad comparisons)."""
    import mpmath

    ok = []
    with mpmath.workprec(300):
        for a, b in zip(x, ref):
            a, b = mpmath.mpf(a), mpmath.mpf(b)
            d = abs(a - b)
            if a == b:
                ok.append(True)
            elif a == 0 or b == 0:
                ok.append(d <= eps)
            else:
                ok.append(d / max(abs(a), abs(b)) <= eps)
    return np.array(ok)


# tests/check_shs_point_all.c:537-600: custom grids (nlat, nlon) of CHarm's
# point-synthesis test, fft = 0 (longitudes over pi, no FFT) or 1 (over 2 pi)
CUSTOM_GRIDS = [(1, 1), (1, 2), (3, 8), (10, 22)]


def custom_grid(i, fft, symm, prec="double"):
    """lat, lon of custom grid i as tests/generate_point.c:19-53 makes it
    (lat from pi/2 down over pi, lon = l/nlon * range), the first latitude
    moved by BREAK_SYMM (tests/check_shs_point_all.c:589-595) when symm is
    false; arithmetic in the precision of the build.  None for the
    combinations the reference test skips."""
    nlat, nlon = CUSTOM_GRIDS[i]
    if nlat == 1 and symm:
        return None
    nlat2 = nlat - 1 if nlat > 1 else nlat
    if prec == "double":
        lat = np.array([np.pi / 2 - (float(l) / float(nlat2)) * np.pi for l in range(nlat)])
        lon_rng = 2.0 * np.pi if fft else np.pi
        lon = np.array([(float(l) / float(nlon)) * lon_rng for l in range(nlon)])
        if not symm:
            lat[0] -= BREAK_SYMM
        return lat, lon
    import mpmath

    with mpmath.workprec(113):
        PI = +mpmath.pi
        lat = np.array([PI / 2 - (mpmath.mpf(l) / nlat2) * PI for l in range(nlat)], dtype=object)
        lon_rng = 2 * PI if fft else PI
        lon = np.array([(mpmath.mpf(l) / nlon) * lon_rng for l in range(nlon)], dtype=object)
        if not symm:
            lat[0] = lat[0] - mpmath.mpf(BREAK_SYMM)
    return lat, lon



# --------------------------------------------------------------------------
# Gradient and guru fixtures (tests/check_shs_point_all.c)
# --------------------------------------------------------------------------
GRAD_NSCTR = [1, 2, 3, 4, 5, 6, 7, 8, 9, 20, 30, 31]
GURU_ORDER = ((0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (2, 0, 0), (0, 2, 0), (0, 0, 2),
              (1, 1, 0), (1, 0, 1), (0, 1, 1))


def elements(what):
    """File suffixes of the outputs of "grad1", "grad2" or a guru (i, j, k)."""
    if what == "grad1":
        return ["x", "y", "z"]
    if what == "grad2":
        return ["xx", "xy", "xz", "yy", "yz", "zz"]
    return ["guru%d%d%d" % tuple(what)]


def singular(what, j):
    """tests/check_shs_point_all.c:256-300: which outputs are zeroed near
    the poles."""
    if what == "grad1":
        return j != 2
    if what == "grad2":
        return j != 5
    _, dlat, dlon = what
    return dlat > 0 or dlon > 0


def zero_polar(prec, lat_rows, f, ncol):
    """tests/check_shs_point_all.c:229-252 on one output (rows of ncol)."""
    f = np.array(f, dtype=object if f.dtype == object else np.float64).reshape(-1, ncol)
    if prec == "double":
        lim = np.pi / 2 - 100 * THRESH[prec]
        polar = [abs(float(la)) > lim for la in lat_rows]
    else:
        import mpmath

        with mpmath.workprec(113):
            lim = +(mpmath.pi / 2 - 100 * mpmath.mpf(THRESH[prec]))
            polar = [abs(mpmath.mpf(la)) > lim for la in lat_rows]
    for i, pl in enumerate(polar):
        if pl:
            f[i, :] = 0
    return f.ravel()


def dsun_rows(lat):
    """Rows of a GL-typed grid whose first latitude was moved: the mirror of
    ring i is the row nlat-1-i at -lat[i] (src/shs/shs_grd_fft.c:117-130)."""
    rows = np.array(lat, copy=True)
    n = len(rows)
    for i in range(n // 2):
        rows[n - 1 - i] = -rows[i]
    return rows


def read_tbl(path, exact=False):
    """(nmax, mu, r, C, S) of a tbl file (src/shc/shc_read_tbl.c): first line
    nmax mu r, then "n m C [S]" lines."""
    with open(path) as fh:
        head = fh.readline().split()
        rows = [ln.split() for ln in fh if ln.strip()]
    conv = float
    if exact:
        import mpmath

        def conv(s):
            with mpmath.workprec(113):
                return +mpmath.mpf(s)
    nmax = int(head[0])
    ncs = (nmax + 1) * (nmax + 2) // 2
    C = np.zeros(ncs, dtype=object if exact else np.float64)
    S = np.zeros(ncs, dtype=object if exact else np.float64)
    for r in rows:
        n, m = int(r[0]), int(r[1])
        C[tri_index(nmax, n, m)] = conv(r[2])
        S[tri_index(nmax, n, m)] = conv(r[3]) if len(r) > 3 else conv("0")
    return nmax, conv(head[1]), conv(head[2]), C, S


# tests/parameters.h:298-356: the lunar gravity forward modelling case
MOON_RREF = 1728000.0
MOON_MASS = 7.346e22
GRAV_CONST = 6.67430e-11
MOON_PMIN, MOON_PMAX = 1, 4
MOON_NMAX_POT = 10
MOON_DENSITY_CONST = 2550.0


def degree_amplitudes(C, S, nmax):
    """src/shc/shc_da.c: sqrt(sum_m C_nm^2 + S_nm^2)."""
    out = []
    for n in range(nmax + 1):
        v = 0.0
        for m in range(n + 1):
            i = tri_index(nmax, n, m)
            v += float(C[i]) ** 2 + float(S[i]) ** 2
        out.append(np.sqrt(v))
    return np.array(out)
