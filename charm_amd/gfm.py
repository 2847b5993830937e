"""Global spectral gravity forward modelling (mirrors PyHarm's ``gfm``
global functions, wrap/pyharm/gfm.py:36-406)."""
import ctypes as _ct

from . import _err
from . import shc as _shc
from ._lib import lib

_P = _ct.POINTER(_shc._Shc)


def _s(x):
    return x.encode() if x is not None else None


def _call(name, argtypes, args):
    f = getattr(lib(), name)
    f.restype = None
    f.argtypes = argtypes + [_ct.POINTER(_err._Err)]
    err = _err.init()
    try:
        f(*args, err)
        _err.handler(err, True)
    finally:
        _err.free(err)


_TAIL = [_ct.c_double, _ct.c_double, _ct.c_uint, _ct.c_uint, _ct.c_ulong, _ct.c_char_p,
         _ct.c_char_p, _ct.c_char_p, _P]


def global_density_3d(shape_shcs, shape_nmax, shape_radius_ref, density_shcs, density_nmax,
                      density_order, grav_const, mass, shape_power_min, shape_power_max,
                      potential_shcs_nmax, shape_density_shcs_path=None, potential_shcs_path=None,
                      shcs_file_format="tbl", return_output=True):
    """Density sum_i rho_i(lat, lon) r^i, i = 0..density_order."""
    k = int(density_order) + 1
    arr = (_P * k)(*[d._Shc for d in density_shcs[:k]])
    nm = (_ct.c_ulong * k)(*[int(v) for v in density_nmax[:k]])
    out = _shc.Shc.from_zeros(potential_shcs_nmax) if return_output else None
    _call("charm_gfm_global_density_3d",
          [_P, _ct.c_ulong, _ct.c_double, _ct.POINTER(_P), _ct.POINTER(_ct.c_ulong), _ct.c_uint]
          + _TAIL,
          [shape_shcs._Shc, int(shape_nmax), float(shape_radius_ref), arr, nm, int(density_order),
           float(grav_const), float(mass), int(shape_power_min), int(shape_power_max),
           int(potential_shcs_nmax), _s(shape_density_shcs_path), _s(potential_shcs_path),
           _s(shcs_file_format), out._Shc if out is not None else None])
    return out


def global_density_lateral(shape_shcs, shape_nmax, shape_radius_ref, density_shcs, density_nmax,
                           grav_const, mass, shape_power_min, shape_power_max,
                           potential_shcs_nmax, shape_density_shcs_path=None,
                           potential_shcs_path=None, shcs_file_format="tbl", return_output=True):
    """Laterally varying density rho(lat, lon)."""
    out = _shc.Shc.from_zeros(potential_shcs_nmax) if return_output else None
    _call("charm_gfm_global_density_lateral",
          [_P, _ct.c_ulong, _ct.c_double, _P, _ct.c_ulong] + _TAIL,
          [shape_shcs._Shc, int(shape_nmax), float(shape_radius_ref), density_shcs._Shc,
           int(density_nmax), float(grav_const), float(mass), int(shape_power_min),
           int(shape_power_max), int(potential_shcs_nmax), _s(shape_density_shcs_path),
           _s(potential_shcs_path), _s(shcs_file_format), out._Shc if out is not None else None])
    return out


def global_density_const(shape_shcs, shape_nmax, shape_radius_ref, density, grav_const, mass,
                         shape_power_min, shape_power_max, potential_shcs_nmax,
                         shape_shcs_path=None, potential_shcs_path=None, shcs_file_format="tbl",
                         return_output=True):
    """Constant density."""
    out = _shc.Shc.from_zeros(potential_shcs_nmax) if return_output else None
    _call("charm_gfm_global_density_const",
          [_P, _ct.c_ulong, _ct.c_double, _ct.c_double] + _TAIL,
          [shape_shcs._Shc, int(shape_nmax), float(shape_radius_ref), float(density),
           float(grav_const), float(mass), int(shape_power_min), int(shape_power_max),
           int(potential_shcs_nmax), _s(shape_shcs_path), _s(potential_shcs_path),
           _s(shcs_file_format), out._Shc if out is not None else None])
    return out
