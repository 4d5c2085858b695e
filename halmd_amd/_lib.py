"""Loader of the HIP backend library (libhalmd_hip.so) and its C-ABI.

The ctypes signatures are derived from include/halmd_hip.h, so the Python
front-end and the C header cannot drift apart.  There is no fallback: if the
library is missing or a call fails, a HalmdError is raised.
"""
import ctypes
import os
import re

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libhalmd_hip.so")
HEADER_PATH = os.path.join(os.path.dirname(_PKG), "include", "halmd_hip.h")

_lib = None
_signatures = None


class HalmdError(RuntimeError):
    """Error reported by the HIP backend (status code and message)."""

    def __init__(self, code, message):
        super().__init__(f"[{code}] {message}")
        self.code = code


_SCALARS = {
    "double": ctypes.c_double,
    "unsigned": ctypes.c_uint,
    "int": ctypes.c_int,
    "size_t": ctypes.c_size_t,
    "uint64_t": ctypes.c_uint64,
    "int64_t": ctypes.c_int64,
}
_POINTERS = {
    "double": ctypes.POINTER(ctypes.c_double),
    "float": ctypes.POINTER(ctypes.c_float),
    "unsigned": ctypes.POINTER(ctypes.c_uint),
    "int": ctypes.POINTER(ctypes.c_int),
    "size_t": ctypes.POINTER(ctypes.c_size_t),
    "uint64_t": ctypes.POINTER(ctypes.c_uint64),
    "int64_t": ctypes.POINTER(ctypes.c_int64),
    "uint32_t": ctypes.c_void_p,   # device pointers in the algorithm entries
    "char": ctypes.c_char_p,
    "void": ctypes.c_void_p,
}


def _ctype(decl):
    decl = decl.strip()
    if decl == "void":
        return None
    decl = re.sub(r"\bconst\b", "", decl)
    stars = decl.count("*")
    base = decl.replace("*", " ").split()[0]
    if stars == 0:
        return _SCALARS[base]
    if base.startswith("halmd_"):
        return ctypes.c_void_p if stars == 1 else ctypes.POINTER(ctypes.c_void_p)
    if stars == 1:
        return _POINTERS[base]
    if stars == 2 and base == "void":
        return ctypes.POINTER(ctypes.c_void_p)
    if stars == 2 and base == "char":
        return ctypes.POINTER(ctypes.c_char_p)
    raise ValueError(f"unsupported C type: {decl}")


def parse_header(path=HEADER_PATH, prefix=r"HALMD_API\s+"):
    """Return {name: (restype, [argtypes])} for every function of the header
    (declarations that start with `prefix`)."""
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    sigs = {}
    for m in re.finditer(prefix + r"([\w\s\*]+?)\s*\b(halmd_\w+)\s*\(([^)]*)\)\s*;", text):
        ret, name, params = m.group(1).strip(), m.group(2), m.group(3)
        restype = ctypes.c_char_p if "char" in ret else ctypes.c_int
        args = []
        if params.strip() not in ("", "void"):
            for p in params.split(","):
                p = p.strip()
                # every parameter is named: drop the trailing identifier
                args.append(_ctype(re.sub(r"\w+\s*$", "", p)))
        sigs[name] = (restype, args)
    return sigs


def lib():
    """The loaded library with argtypes set; raises if it is not built."""
    global _lib, _signatures
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is not built: run `python -c 'import __graft_entry__ as g; g.build()'` "
                "or `make -C halmd_amd/csrc`")
        L = ctypes.CDLL(LIB_PATH)
        _signatures = parse_header()
        for name, (restype, args) in _signatures.items():
            fn = getattr(L, name)
            fn.restype = restype
            fn.argtypes = args
        _lib = L
    return _lib


def signatures():
    lib()
    return _signatures


def call(name, *args):
    """Call a C-ABI function and raise HalmdError on a non-zero status."""
    L = lib()
    rc = getattr(L, name)(*args)
    if rc != 0:
        msg = L.halmd_last_error()
        raise HalmdError(rc, msg.decode() if msg else "unknown error")
    return rc


# ---- numpy helpers ------------------------------------------------------------
def f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None:
        a = a.reshape(shape)
    return a


def ptr(a, ctype):
    return a.ctypes.data_as(ctypes.POINTER(ctype)) if a is not None else None


def dptr(a):
    return ptr(a, ctypes.c_double)


def uptr(a):
    return ptr(a, ctypes.c_uint)


def iptr(a):
    return ptr(a, ctypes.c_int)


def u64ptr(a):
    return ptr(a, ctypes.c_uint64)


def out_handle():
    return ctypes.c_void_p()


class Handle:
    """Owner of one C-ABI object; destroyed with the given function."""

    _destroy = None

    def __init__(self):
        self._h = None

    @property
    def handle(self):
        if self._h is None or not self._h.value:
            raise HalmdError(1, f"{type(self).__name__} is not initialised")
        return self._h

    def close(self):
        """Destroy the C-ABI object now (a force module then leaves its particle)."""
        h = getattr(self, "_h", None)
        if h is not None and h.value and self._destroy and _lib is not None:
            getattr(_lib, self._destroy)(h)
            self._h = None

    def __del__(self):
        self.close()


def device_info():
    name = ctypes.create_string_buffer(64)
    cu = ctypes.c_int()
    mem = ctypes.c_size_t()
    call("halmd_device_info", name, 64, ctypes.byref(cu), ctypes.byref(mem))
    return name.value.decode(), cu.value, mem.value
