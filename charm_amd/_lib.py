"""Loads libcharm_amd.so (built in-tree next to this file).

There is no fallback: without the library every call raises.
"""
import ctypes as _ct
import os as _os

_HERE = _os.path.dirname(_os.path.abspath(__file__))
LIB_PATH = _os.path.join(_HERE, "libcharm_amd.so")

_lib = None


def lib():
    """The loaded C ABI library (raises ``OSError`` if it is not built)."""
    global _lib
    if _lib is None:
        if not _os.path.exists(LIB_PATH):
            raise OSError("%s is missing: build it with charm_amd/csrc/build.sh "
                          "(or __graft_entry__.build())" % LIB_PATH)
        _lib = _ct.CDLL(LIB_PATH)
    return _lib


def build():
    """Compile libcharm_amd.so for gfx950 with hipcc."""
    import subprocess

    subprocess.run(["sh", _os.path.join(_HERE, "csrc", "build.sh"), LIB_PATH], check=True)
