"""charm_amd: CHarm's point spherical-harmonic synthesis and analysis on
AMD MI355X (gfx950), behind CHarm's C ABI and a PyHarm-style Python API.

Modules mirror PyHarm: :mod:`shc` (coefficients), :mod:`crd` (points),
:mod:`shs` (synthesis), :mod:`sha` (analysis), :mod:`gfm` (global gravity
forward modelling); :mod:`device` holds the
device-resident entry points.
"""
from . import _err, crd, device, gfm, serve, sha, shc, shs, workloads  # noqa: F401
from ._err import CHarmError  # noqa: F401
from ._lib import LIB_PATH, build, lib  # noqa: F401
