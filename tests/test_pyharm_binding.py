"""The PyHarm binding of INTEGRATION.md, run against a stand-in for
libcharm.so: PyHarm reads and writes CHarm's globals with
ctypes.c_double.in_dll(_libcharm, 'charm_glob_threshold') and the like
(wrap/pyharm/glob.py:35-120), which resolves through _libcharm._handle.
With libcharm_amd.so loaded globally before CHarm and the dispatch object
handing in_dll its handle, those writes reach libcharm_amd.so, and CHarm's
own code (here the stand-in's) reads the same variables."""
import os
import subprocess
import sys
import textwrap

import charm_amd

STANDIN = r"""
double charm_glob_threshold = 1.0;
double charm_glob_threshold2 = 2.0;
unsigned long charm_glob_polar_optimization_a1 = 7;
double charm_glob_polar_optimization_a2 = 3.0;
double standin_threshold(void) { return charm_glob_threshold; }
double standin_threshold2(void) { return charm_glob_threshold2; }
unsigned long standin_a1(void) { return charm_glob_polar_optimization_a1; }
double standin_a2(void) { return charm_glob_polar_optimization_a2; }
void charm_shs_point(void) {}
void charm_crd_point_free(void) {}
"""

SCRIPT = r"""
import ctypes as _ct, sys
amd_path, standin_path = sys.argv[1], sys.argv[2]

# --- the binding of INTEGRATION.md (PyHarm's _load_lib is a ctypes.CDLL) ---
_libamd = _ct.CDLL(amd_path, mode=_ct.RTLD_GLOBAL)   # before CHarm
_libcharm = _ct.CDLL(standin_path)
_AMD_NAMES = {'charm_shs_point', 'charm_sha_point'}

class _Dispatch:
    def __init__(self, main, amd):
        self._main, self._amd = main, amd
        self._handle = amd._handle
    def __getitem__(self, name):
        if name in _AMD_NAMES:
            return self._amd[name]
        return self._main[name]
    def __getattr__(self, name):
        return getattr(self._main, name)

_libcharm = _Dispatch(_libcharm, _libamd)
# --- end of the binding ---

main = _libcharm._main
for f in ("standin_threshold", "standin_threshold2", "standin_a2"):
    getattr(main, f).restype = _ct.c_double
main.standin_a1.restype = _ct.c_ulong
eps = 2.220446049250313e-16
thr = _ct.c_double.in_dll(_libcharm, "charm_glob_threshold")
a1 = _ct.c_ulong.in_dll(_libcharm, "charm_glob_polar_optimization_a1")
a2 = _ct.c_double.in_dll(_libcharm, "charm_glob_polar_optimization_a2")
# PyHarm sees libcharm_amd's variables (CHarm's defaults), and so does CHarm's code
assert thr.value == 100 * eps and a1.value == 100 and a2.value == -1.0
assert main.standin_threshold() == 100 * eps and main.standin_a1() == 100
# a PyHarm write reaches both libraries
thr.value = 1e-12
a1.value = 55
a2.value = 0.25
_ct.c_double.in_dll(_libcharm, "charm_glob_threshold2").value = 3e-9
amd = _libcharm._amd
assert _ct.c_double.in_dll(amd, "charm_glob_threshold").value == 1e-12
assert _ct.c_ulong.in_dll(amd, "charm_glob_polar_optimization_a1").value == 55
assert _ct.c_double.in_dll(amd, "charm_glob_polar_optimization_a2").value == 0.25
assert _ct.c_double.in_dll(amd, "charm_glob_threshold2").value == 3e-9
assert main.standin_threshold() == 1e-12 and main.standin_a1() == 55
assert main.standin_a2() == 0.25 and main.standin_threshold2() == 3e-9
# without the handle, in_dll would land on the stand-in's own, unused copy
assert _ct.c_double.in_dll(main, "charm_glob_threshold").value == 1.0
# hot calls come from libcharm_amd.so, the rest from CHarm
addr = lambda f: _ct.cast(f, _ct.c_void_p).value
assert addr(_libcharm["charm_shs_point"]) == addr(amd["charm_shs_point"])
assert addr(_libcharm["charm_crd_point_free"]) == addr(main["charm_crd_point_free"])
print("binding ok")
"""


def test_pyharm_binding_shares_globals(tmp_path):
    src = tmp_path / "standin.c"
    src.write_text(STANDIN)
    so = tmp_path / "libcharm_standin.so"
    subprocess.run(["gcc", "-shared", "-fPIC", "-O2", str(src), "-o", str(so)], check=True)
    script = tmp_path / "binding.py"
    script.write_text(textwrap.dedent(SCRIPT))
    amd = os.path.join(os.path.dirname(charm_amd.__file__), "libcharm_amd.so")
    out = subprocess.run([sys.executable, str(script), amd, str(so)], capture_output=True,
                         text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    assert "binding ok" in out.stdout
