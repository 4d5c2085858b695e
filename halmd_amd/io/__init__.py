"""halmd.io: H5MD writers and readers over include/halmd_h5md.h.

Mirrors lua/halmd/io/writers/h5md.lua.in (file(path, email, overwrite) and
file:writer{location, mode}) and lua/halmd/io/readers/h5md.lua.in
(file(path) and file:reader{location, mode}); the group writers and readers
mirror halmd/io/writers/h5md/{append,truncate}.cpp and
halmd/io/readers/h5md/{append,truncate}.cpp: slots registered with
on_write / on_read, then write() or read_at_step() / read_at_time().

The library (libhalmd_h5md.so) is host code linked against HDF5; it is
separate from the HIP library, and loading it does not touch the GPU.
"""
import ctypes
import os

import numpy as np

from .. import _lib

LIB_PATH = os.path.join(os.path.dirname(_lib.LIB_PATH), "libhalmd_h5md.so")
HEADER_PATH = os.path.join(os.path.dirname(_lib.HEADER_PATH), "halmd_h5md.h")

F64, F32, I32, U32, I64, U64 = range(6)
_CODES = {np.dtype(np.float64): F64, np.dtype(np.float32): F32, np.dtype(np.int32): I32,
          np.dtype(np.uint32): U32, np.dtype(np.int64): I64, np.dtype(np.uint64): U64}

_h5 = None


class H5mdError(RuntimeError):
    def __init__(self, code, message):
        super().__init__(f"[{code}] {message}")
        self.code = code


def lib():
    global _h5
    if _h5 is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run `make -C halmd_amd/h5md`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (restype, args) in _lib.parse_header(HEADER_PATH, prefix=r"(?m)^").items():
            fn = getattr(L, name)
            fn.restype = restype
            fn.argtypes = args
        _h5 = L
    return _h5


def call(name, *args):
    L = lib()
    rc = getattr(L, name)(*args)
    if rc != 0:
        msg = L.halmd_h5md_last_error()
        raise H5mdError(rc, msg.decode() if msg else "unknown error")
    return rc


def _path(location):
    if isinstance(location, str):
        return location
    return "/".join(str(x) for x in location if x is not None)


def _array(value):
    a = np.asarray(value)   # (a scalar stays 0-d)
    if not a.flags.c_contiguous:
        a = a.copy(order="C")
    if a.dtype == np.bool_:
        a = a.astype(np.int32)
    if a.dtype not in _CODES:
        a = a.astype(np.float64 if a.dtype.kind == "f" else np.int64)
    return a


def _dims(a):
    d = np.asarray(a.shape, dtype=np.uint64)
    return a.ndim, (d.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)) if a.ndim else None), d


class _file:
    def __init__(self, handle, path):
        self._h = handle
        self.path = path

    @property
    def handle(self):
        if self._h is None or not self._h.value:
            raise H5mdError(1, "file is closed")
        return self._h

    def flush(self):
        call("halmd_h5md_file_flush", self.handle)

    def close(self):
        if self._h is not None and self._h.value and _h5 is not None:
            _h5.halmd_h5md_file_close(self._h)
        self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def version(self):
        v = (ctypes.c_int * 2)()
        call("halmd_h5md_file_version", self.handle, v)
        return [v[0], v[1]]

    def exists(self, location):
        v = ctypes.c_int()
        call("halmd_h5md_exists", self.handle, _path(location).encode(), ctypes.byref(v))
        return bool(v.value)

    def attribute_string(self, location, name):
        buf = ctypes.create_string_buffer(4096)
        call("halmd_h5md_read_attribute_string", self.handle, _path(location).encode(), name.encode(), buf, 4096)
        return buf.value.decode()

    def attribute_int(self, location, name, n=1):
        v = (ctypes.c_int * n)()
        call("halmd_h5md_read_attribute_int", self.handle, _path(location).encode(), name.encode(), v, n)
        return list(v) if n > 1 else v[0]

    def write_attribute_int(self, location, name, value):
        v = np.atleast_1d(np.asarray(value, dtype=np.int32))
        call("halmd_h5md_write_attribute_int", self.handle, _path(location).encode(), name.encode(),
             v.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), v.size)

    def write_attribute_strings(self, location, name, values):
        arr = (ctypes.c_char_p * len(values))(*[s.encode() for s in values])
        call("halmd_h5md_write_attribute_strings", self.handle, _path(location).encode(), name.encode(), arr,
             len(values))

    def shape(self, path):
        rank = ctypes.c_int()
        dims = (ctypes.c_uint64 * 8)()
        call("halmd_h5md_dataset_shape", self.handle, _path(path).encode(), ctypes.byref(rank), dims)
        return tuple(dims[k] for k in range(rank.value))

    def read_dataset(self, path, dtype=np.float64):
        """A whole dataset (truncate mode), e.g. particles/all/box/edges."""
        p = _path(path)
        loc, _, name = p.rpartition("/")
        out = np.empty(self.shape(p), dtype=dtype)
        rank, dptr, _ = _dims(out)
        call("halmd_h5md_truncate_read", self.handle, loc.encode(), name.encode(), _CODES[out.dtype],
             out.ctypes.data_as(ctypes.c_void_p), rank, dptr)
        return out

    def read_sample(self, location, index, dtype=np.float64):
        """Sample `index` of the append-mode field group at `location`."""
        p = _path(location)
        shape = self.shape(p + "/value")[1:]
        out = np.empty(shape, dtype=dtype)
        rank, dptr, _ = _dims(out)
        call("halmd_h5md_append_read", self.handle, p.encode(), int(index), _CODES[out.dtype],
             out.ctypes.data_as(ctypes.c_void_p), rank, dptr)
        return out

    def samples(self, location):
        v = ctypes.c_uint64()
        call("halmd_h5md_append_samples", self.handle, _path(location).encode(), ctypes.byref(v))
        return v.value

    def step_time(self, location, index):
        s, t = ctypes.c_int64(), ctypes.c_double()
        call("halmd_h5md_append_step_time", self.handle, _path(location).encode(), int(index), ctypes.byref(s),
             ctypes.byref(t))
        return s.value, t.value


class _append_writer:
    """halmd/io/writers/h5md/append.cpp:39-118: shared step and time, one
    value dataset per field; write() appends one sample of every field."""

    def __init__(self, file, location, clock):
        self.file = file
        self.location = list(location) if not isinstance(location, str) else location.split("/")
        self.clock = clock
        self._h = ctypes.c_void_p()
        call("halmd_h5md_append_create", file.handle, _path(location).encode(), ctypes.byref(self._h))
        self._slots = []
        self._prepend, self._append = [], []

    @property
    def group(self):
        return _path(self.location)

    def on_write(self, slot, location):
        self._slots.append((_path(location), slot))

    def on_prepend_write(self, slot):
        self._prepend.append(slot)

    def on_append_write(self, slot):
        self._append.append(slot)

    def write(self):
        for f in self._prepend:
            f()
        call("halmd_h5md_append_begin", self._h, int(self.clock.step), float(self.clock.time))
        for name, slot in self._slots:
            a = _array(slot())
            rank, dptr, _ = _dims(a)
            call("halmd_h5md_append_field", self._h, name.encode(), _CODES[a.dtype],
                 a.ctypes.data_as(ctypes.c_void_p), rank, dptr)
        for f in self._append:
            f()

    def __del__(self):
        if getattr(self, "_h", None) is not None and self._h.value and _h5 is not None:
            _h5.halmd_h5md_append_destroy(self._h)
            self._h = None


class _truncate_writer:
    """halmd/io/writers/h5md/truncate.cpp:36-103: each field a dataset,
    rewritten whole by write()."""

    def __init__(self, file, location):
        self.file = file
        self.location = list(location) if not isinstance(location, str) else location.split("/")
        self._slots = []

    @property
    def group(self):
        return _path(self.location)

    def on_write(self, slot, location):
        self._slots.append((_path(location), slot))

    def write(self):
        for name, slot in self._slots:
            a = _array(slot())
            sub, _, leaf = name.rpartition("/")
            loc = _path(self.location + ([sub] if sub else []))
            rank, dptr, _ = _dims(a)
            call("halmd_h5md_truncate_write", self.file.handle, loc.encode(), leaf.encode(), _CODES[a.dtype],
                 a.ctypes.data_as(ctypes.c_void_p), rank, dptr)


class _append_reader:
    """halmd/io/readers/h5md/append.cpp:38-160: slots receive the sample of
    their field at the step (or time) asked for."""

    def __init__(self, file, location):
        self.file = file
        self.location = list(location) if not isinstance(location, str) else location.split("/")
        self._slots = []

    def on_read(self, slot, location, dtype=np.float64):
        self._slots.append((_path(self.location + [_path(location)]), slot, dtype))

    def _read(self, index_of):
        for loc, slot, dtype in self._slots:
            slot(self.file.read_sample(loc, index_of(loc), dtype))

    def read_at_step(self, offset):
        def index_of(loc):
            v = ctypes.c_uint64()
            call("halmd_h5md_append_index_at_step", self.file.handle, loc.encode(), int(offset), ctypes.byref(v))
            return v.value
        self._read(index_of)

    def read_at_time(self, offset):
        def index_of(loc):
            v = ctypes.c_uint64()
            call("halmd_h5md_append_index_at_time", self.file.handle, loc.encode(), float(offset),
                 ctypes.byref(v))
            return v.value
        self._read(index_of)


class _truncate_reader:
    def __init__(self, file, location):
        self.file = file
        self.location = list(location) if not isinstance(location, str) else location.split("/")
        self._slots = []

    def on_read(self, slot, location, dtype=np.float64):
        self._slots.append((_path(self.location + [_path(location)]), slot, dtype))

    def read(self):
        for path, slot, dtype in self._slots:
            slot(self.file.read_dataset(path, dtype))


class _clock:
    """Step and time of an MD core (halmd/mdsim/clock.hpp): time = step * dt."""

    def __init__(self, core):
        self.core = core

    @property
    def step(self):
        return self.core.step

    @property
    def time(self):
        return self.core.time


class writers:
    class h5md(_file):
        """lua/halmd/io/writers/h5md.lua.in: create the file with the H5MD
        metadata (an existing file is kept unless overwrite is set)."""

        def __init__(self, path, email="", overwrite=False, author=""):
            self._h = None
            h = ctypes.c_void_p()
            call("halmd_h5md_file_create", path.encode(), author.encode(), email.encode(), int(bool(overwrite)),
                 ctypes.byref(h))
            super().__init__(h, path)

        def writer(self, location, mode, clock=None):
            if mode == "append":
                if clock is None:
                    raise ValueError("append writer: a clock (MD core) is required")
                return _append_writer(self, location, clock if hasattr(clock, "time") else _clock(clock))
            if mode == "truncate":
                return _truncate_writer(self, location)
            raise ValueError(f"invalid mode: {mode}")


class readers:
    class h5md(_file):
        """lua/halmd/io/readers/h5md.lua.in: open read-only and check the
        H5MD metadata."""

        def __init__(self, path):
            self._h = None
            h = ctypes.c_void_p()
            call("halmd_h5md_file_open", path.encode(), ctypes.byref(h))
            super().__init__(h, path)

        def reader(self, location, mode):
            if mode == "append":
                return _append_reader(self, location)
            if mode == "truncate":
                return _truncate_reader(self, location)
            raise ValueError(f"invalid mode: {mode}")


def version():
    """H5MD version written by this library (writers/h5md/file.cpp: {1, 0})."""
    return [1, 0]
