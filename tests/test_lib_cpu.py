"""CPU-side checks of the C ABI library: it loads, exports every symbol the
header declares, has CHarm's struct layouts, and its host-only functions
(errors, coefficients, points, Gauss-Legendre grids) behave like CHarm's."""
import ctypes as ct
import os
import re

import numpy as np
import pytest

import charm_amd
from charm_amd import crd, shc
from charm_fixtures import THRESH, nearly_equal, read_fixture

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "charm_amd.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    funcs = re.findall(r"\b(charm_[a-z0-9_]+)\s*\(", txt)
    data = re.findall(r"extern\s+[a-z_ ]+\s+(charm_[a-z0-9_]+)\s*;", txt)
    return sorted(set(funcs)), sorted(set(data))


def test_exports_every_declared_symbol():
    lib = charm_amd.lib()
    funcs, data = header_symbols()
    assert "charm_shs_point" in funcs and "charm_sha_point" in funcs
    missing = [s for s in funcs + data if not hasattr(lib, s)]
    assert not missing, missing


def test_struct_layouts_match_charm():
    # non-MPI x86-64 layouts of src/{shc,crd,err}/*.h.in
    assert ct.sizeof(shc._Shc) == 64
    assert shc._Shc.c.offset == 40 and shc._Shc.owner.offset == 56
    assert ct.sizeof(crd._Point) == 72
    assert crd._Point.lat.offset == 32 and crd._Point.owner.offset == 64
    assert ct.sizeof(charm_amd._err._Err) == 56
    assert charm_amd._err._Err.msg.offset == 40


# CHarm's non-MPI x86-64 layouts (src/shc/shc.h.in:139-381,
# src/crd/crd.h.in:125-332, src/err/err.h.in:159-222): size, then offsets
CHARM_LAYOUTS = {
    "charm_shc": (64, [("nmax", 0), ("mu", 8), ("r", 16), ("nc", 24), ("ns", 32), ("c", 40),
                       ("s", 48), ("owner", 56), ("distributed", 57)]),
    "charm_point": (72, [("type", 0), ("nlat", 8), ("nlon", 16), ("npoint", 24), ("lat", 32),
                         ("lon", 40), ("r", 48), ("w", 56), ("owner", 64),
                         ("distributed", 65)]),
    "charm_err": (56, [("level", 0), ("file", 8), ("line", 16), ("func", 24), ("code", 32),
                       ("msg", 40), ("saturated", 48), ("distributed", 49)]),
}


def test_header_struct_layouts_compiled(tmp_path):
    """The C header itself, compiled with gcc, has CHarm's struct layouts."""
    import subprocess
    body = ""
    for name, (_, fields) in CHARM_LAYOUTS.items():
        body += '    printf("%%zu", sizeof(%s));\n' % name
        body += "".join('    printf(" %%zu", offsetof(%s, %s));\n' % (name, f) for f, _ in fields)
        body += '    printf("\\n");\n'
    src = tmp_path / "charm_layout.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "charm_amd.h"\n'
                   "int main(void) {\n" + body + "    return 0;\n}\n")
    exe = tmp_path / "charm_layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o",
                    str(exe)], check=True)
    lines = subprocess.run([str(exe)], check=True, capture_output=True,
                           text=True).stdout.split("\n")
    for line, (name, (size, fields)) in zip(lines, CHARM_LAYOUTS.items()):
        got = [int(v) for v in line.split()]
        assert got == [size] + [o for _, o in fields], name


def test_request_layout_matches_header(tmp_path):
    """The ctypes mirror of charm_amd_request has the header's offsets."""
    from charm_amd import device
    fields = [f for f, _ in device._Request._fields_]
    src = tmp_path / "layout.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "charm_amd.h"\n'
                   "int main(void) {\n"
                   '    printf("%zu", sizeof(charm_amd_request));\n'
                   + "".join('    printf(" %%zu", offsetof(charm_amd_request, %s));\n' % f
                             for f in fields)
                   + "    return 0;\n}\n")
    exe = tmp_path / "layout"
    import subprocess
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    got = [int(v) for v in subprocess.run([str(exe)], check=True, capture_output=True,
                                          text=True).stdout.split()]
    want = [ct.sizeof(device._Request)] + [getattr(device._Request, f).offset for f in fields]
    assert got == want


def test_globals_default_values():
    lib = charm_amd.lib()
    eps = np.finfo(float).eps
    assert ct.c_double.in_dll(lib, "charm_glob_threshold").value == 100 * eps
    assert ct.c_double.in_dll(lib, "charm_glob_threshold2").value == 1e5 * eps
    assert ct.c_ulong.in_dll(lib, "charm_glob_polar_optimization_a1").value == 100
    assert ct.c_double.in_dll(lib, "charm_glob_polar_optimization_a2").value == -1.0
    # polar-optimisation blocks of CHarm's AVX-512 double build (src/simd/simd.h:303-304)
    assert ct.c_size_t.in_dll(lib, "charm_glob_shs_block_lat_multiplier").value == 4
    assert ct.c_size_t.in_dll(lib, "charm_glob_sha_block_lat_multiplier").value == 8
    assert ct.c_int.in_dll(lib, "charm_amd_polar_simd_size").value == 8


def test_noalloc_rejects_sizes_beyond_its_limits():
    """The allocation-free entry point checks its limits before any device
    work (null device pointers are never touched)."""
    from charm_amd import _err
    f = charm_amd.lib().charm_amd_shs_sctr_dev_noalloc
    f.restype = None
    f.argtypes = [ct.c_size_t, ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_ulong, ct.c_double,
                  ct.c_double, ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.POINTER(_err._Err)]
    for npoint, nmax in ((1 << 20) + 1, 10), (100, 2191):
        err = _err.init()
        try:
            f(npoint, None, None, None, nmax, 1.0, 1.0, None, None, None, err)
            assert err.contents.code == 2   # CHARM_EFUNCARG
            assert b"at most 1048576 points and degree 2190" in ct.string_at(err.contents.msg)
        finally:
            _err.free(err)


@pytest.mark.parametrize("nmax", range(0, 11))
def test_gl_grid_matches_fixtures(nmax):
    eps = 10 * THRESH["double"]
    for d in (0, 1):
        g = crd.PointGridGL(nmax, 1.0 + 1000.0 * d)
        assert nearly_equal(g.lat, read_fixture("double", "crd_nx%d_dr%d_gl_lat.txt" % (nmax, d)), eps).all()
        assert nearly_equal(g.w, read_fixture("double", "crd_nx%d_dr%d_gl_w.txt" % (nmax, d)), eps).all()
        assert nearly_equal(g.lon, read_fixture("double", "crd_nx%d_dr%d_gl_lon.txt" % (nmax, d)), eps).all()
        assert (g.r == 1.0 + 1000.0 * d).all()


@pytest.mark.parametrize("nmax", range(0, 11))
@pytest.mark.parametrize("dh", (1, 2))
def test_dh_grid_matches_fixtures(nmax, dh):
    """charm_crd_point_dh1/dh2 against CHarm's crd_nx*_dr*_dh{1,2} files at
    CHarm's tolerance (tests/check_crd_point_quad.c:59-76)."""
    for d in (0, 1):
        g = (crd.PointGridDH1 if dh == 1 else crd.PointGridDH2)(nmax, 1.0 + 1000.0 * d)
        assert (g.lat.size, g.lon.size) == (crd.PointGridDH1 if dh == 1 else
                                            crd.PointGridDH2).shape(nmax)
        name = "crd_nx%d_dr%d_dh%d_%%s.txt" % (nmax, d, dh)
        for key, got in (("lat", g.lat), ("lon", g.lon), ("w", g.w), ("r", g.r)):
            ref = read_fixture("double", name % key)
            assert nearly_equal(got, ref, 10 * THRESH["double"]).all(), key


def test_gl_grid_matches_oracle_large():
    from oracle.oracle import Oracle

    lat, w, lon = Oracle("d").gl_grid(400)
    g = crd.PointGridGL(400, 2.0)
    assert np.array_equal(g.lat, lat) and np.array_equal(g.w, w) and np.array_equal(g.lon, lon)
    assert abs(g.w.sum() - 2.0) < 1e-13


def test_shc_layout_and_views():
    c = np.arange(21, dtype=float)
    s = -np.arange(21, dtype=float)
    x = shc.Shc.from_arrays(5, c, s, 2.0, 3.0)
    assert x.nmax == 5 and x.mu == 2.0 and x.r == 3.0 and not x.owner
    assert x.get_coeffs(3, 2) == (c[shc.index(5, 3, 2)], s[shc.index(5, 3, 2)])
    z = shc.Shc.from_zeros(7)
    assert z.owner and z.c.size == 36 and not z.c.any()
    with pytest.raises(ValueError):
        shc.Shc.from_zeros(3, 1.0, -1.0)


def test_point_checks():
    with pytest.raises(ValueError):
        crd.PointSctr(np.zeros(3), np.zeros(4), np.ones(3))  # nlat != nlon
    p = crd.PointSctr(np.zeros(3), np.zeros(3), np.ones(3))
    assert p.npoint == 3
    g = crd.PointGrid(np.zeros(3), np.zeros(5), np.ones(3))
    assert g.npoint == 15


def test_error_structure_roundtrip():
    err = charm_amd._err.init()
    assert err.contents.code == 0 and err.contents.level == 0
    charm_amd._err.free(err)


def test_nmax_check_reports_efuncarg():
    # argument checks run before any device work (src/shs/shs_point.c:37-48)
    s = shc.Shc.from_zeros(3)
    g = crd.PointGridGL(5, 1.0)
    with pytest.raises(ValueError):
        charm_amd.shs.point(g, s, 4)
    f = charm_amd.lib().charm_shs_point
    f.restype = None
    f.argtypes = [ct.POINTER(crd._Point), ct.POINTER(shc._Shc), ct.c_ulong,
                  ct.POINTER(ct.c_double), ct.POINTER(charm_amd._err._Err)]
    err = charm_amd._err.init()
    out = np.zeros(6 * 12)
    f(g._Point, s._Shc, 4, out.ctypes.data_as(ct.POINTER(ct.c_double)), err)
    assert err.contents.code == charm_amd._err.CHARM_EFUNCARG
    charm_amd._err.free(err)


def test_derivative_argument_errors():
    """Checks made before any device work (src/shs/shs_point_guru.c:30-62,
    src/shs/shs_check_single_derivative.c:17-64)."""
    s = shc.Shc.from_zeros(3, 1.0, 1.0)
    g = crd.PointGridGL(3, 1.0)
    with pytest.raises(charm_amd.CHarmError, match="cannot be larger than \"2\""):
        charm_amd.shs.point_guru(g, s, 3, 1, 1, 1)
    with pytest.raises(ValueError):
        charm_amd.shs.point_guru(g, s, 3, -1, 0, 0)
    with pytest.raises(ValueError):
        charm_amd.shs.point_grad1(g, s, 4)
    # zero points: nothing to do
    p = crd.PointSctr(np.zeros(0), np.zeros(0), np.zeros(0))
    assert charm_amd.shs.point_guru(p, s, 3, 1, 0, 0).size == 0
    assert all(x.size == 0 for x in charm_amd.shs.point_grad2(p, s, 3))


def test_gfm_argument_errors():
    """src/gfm/gfm_sgfm.c:400-690: checked before any device work."""
    s = shc.Shc.from_zeros(4)
    s.c[0] = 1.1
    g = charm_amd.gfm.global_density_const
    for args in ((s, 5, 1.0, 1.0, 1.0, 1.0, 1, 2, 4),      # shape_nmax > shcs nmax
                 (s, 4, -1.0, 1.0, 1.0, 1.0, 1, 2, 4),     # negative radius
                 (s, 4, 1.0, 1.0, 1.0, 0.0, 1, 2, 4),      # zero mass
                 (s, 4, 1.0, 1.0, 1.0, 1.0, 0, 2, 4),      # power 0
                 (s, 4, 1.0, 1.0, 1.0, 1.0, 3, 2, 4)):     # pmin > pmax
        with pytest.raises(charm_amd.CHarmError):
            g(*args)
    with pytest.raises(charm_amd.CHarmError, match="shcs_file_format"):
        g(s, 4, 1.0, 1.0, 1.0, 1.0, 1, 2, 4, shape_shcs_path="/tmp/x", shcs_file_format="gfc")
    with pytest.raises(charm_amd.CHarmError, match="potential_shcs_nmax"):
        charm_amd.lib()  # potential degree above the output's degree
        out = shc.Shc.from_zeros(2)
        import ctypes as ct
        from charm_amd import _err
        f = charm_amd.lib().charm_gfm_global_density_const
        f.restype = None
        err = _err.init()
        f.argtypes = [ct.POINTER(shc._Shc), ct.c_ulong, ct.c_double, ct.c_double, ct.c_double,
                      ct.c_double, ct.c_uint, ct.c_uint, ct.c_ulong, ct.c_char_p, ct.c_char_p,
                      ct.c_char_p, ct.POINTER(shc._Shc), ct.POINTER(_err._Err)]
        f(s._Shc, 4, 1.0, 1.0, 1.0, 1.0, 1, 2, 4, None, None, None, out._Shc, err)
        _err.handler(err, True)
    bad = shc.Shc.from_zeros(4, 2.0, 1.0)
    with pytest.raises(charm_amd.CHarmError, match="have to be"):
        g(bad, 4, 1.0, 1.0, 1.0, 1.0, 1, 2, 4)


def test_sha_rejects_scattered_points():
    s = shc.Shc.from_zeros(2)
    p = crd.PointSctr(np.zeros(3), np.zeros(3), np.ones(3))
    with pytest.raises(TypeError):
        charm_amd.sha.point(p, np.zeros((3, 3)), 2)


def test_no_device_fails_loudly():
    """Without a HIP device the transforms raise; there is no CPU fallback."""
    if charm_amd.lib().charm_amd_device_count() > 0:
        pytest.skip("a GPU is visible")
    s = shc.Shc.from_zeros(4, 1.0, 1.0)
    g = crd.PointGridGL(4, 1.0)
    with pytest.raises(charm_amd.CHarmError):
        charm_amd.shs.point(g, s, 4)
