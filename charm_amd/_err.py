"""charm_err mirror and error handling (wrap/pyharm/_err.py:24-121)."""
import ctypes as _ct

from ._lib import lib

CHARM_SUCCESS = 0
CHARM_EMEM = 1
CHARM_EFUNCARG = 2
CHARM_EFILEIO = 3
CHARM_EFFTWINIT = 4
CHARM_EMPIPROCESS = 5


class CHarmError(Exception):
    """Raised when a charm_* call reports an error."""

    def __init__(self, msg, code=None):
        super().__init__(msg)
        self.code = code


class _Err(_ct.Structure):
    # src/err/err.h.in:159-222 (non-MPI build)
    _fields_ = [("level", _ct.c_size_t),
                ("file", _ct.POINTER(_ct.c_char_p)),
                ("line", _ct.POINTER(_ct.c_size_t)),
                ("func", _ct.POINTER(_ct.c_char_p)),
                ("code", _ct.c_int),
                ("msg", _ct.c_char_p),
                ("saturated", _ct.c_bool),
                ("distributed", _ct.c_bool)]


def init():
    f = lib().charm_err_init
    f.restype = _ct.POINTER(_Err)
    f.argtypes = []
    err = f()
    if not err:
        raise MemoryError("charm_err_init failed")
    return err


def free(err):
    f = lib().charm_err_free
    f.restype = None
    f.argtypes = [_ct.POINTER(_Err)]
    f(err)


def handler(err, raise_error=True):
    """Raise CHarmError if err holds an error (the Python counterpart of
    charm_err_handler(err, 1))."""
    e = err.contents
    if e.code == CHARM_SUCCESS:
        return
    msg = e.msg.decode() if e.msg else ""
    code = e.code
    tb = []
    for i in range(e.level - 1, -1, -1):
        tb.append('File "%s", line %d, function "%s"' % (
            e.file[i].decode(), e.line[i], e.func[i].decode()))
    f = lib().charm_err_reset
    f.restype = None
    f.argtypes = [_ct.POINTER(_Err)]
    f(err)
    if raise_error:
        raise CHarmError("%s\nTraceback (C): %s" % (msg, "; ".join(tb)), code)
