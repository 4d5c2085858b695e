"""H5MD writers and readers (host library libhalmd_h5md.so) on the CPU.

The layout follows the reference's writers and readers
(halmd/io/writers/h5md/{file,append,truncate}.cpp, readers/h5md/*.cpp):
/h5md metadata, append groups with shared step and time, truncate datasets.
The index rules of the append reader are those of readers/h5md/append.cpp:
read_step_index and read_time_index.
"""
import os

import numpy as np
import pytest

from halmd_amd import io


class _clock:
    def __init__(self):
        self.step, self.time = 0, 0.0


def test_file_metadata_and_overwrite(tmp_path):
    path = str(tmp_path / "a.h5")
    f = io.writers.h5md(path, email="md@example.org", author="Jane Doe")
    f.close()
    r = io.readers.h5md(path)
    assert r.version == [1, 0] == io.version()
    assert r.attribute_string("h5md/author", "name") == "Jane Doe"
    assert r.attribute_string("h5md/author", "email") == "md@example.org"
    assert r.attribute_string("h5md/creator", "name")
    r.close()
    # an existing file is kept unless overwrite is set (writers/h5md/file.cpp:45-52)
    with pytest.raises(io.H5mdError):
        io.writers.h5md(path)
    io.writers.h5md(path, overwrite=True).close()
    # not an H5MD file
    bad = tmp_path / "b.h5"
    bad.write_bytes(b"not hdf5")
    with pytest.raises(io.H5mdError):
        io.readers.h5md(str(bad))


def test_append_and_truncate_round_trip(tmp_path):
    path = str(tmp_path / "t.h5")
    clock = _clock()
    f = io.writers.h5md(path)
    rng = np.random.default_rng(1)
    pos = [rng.normal(size=(7, 3)) for _ in range(4)]
    spec = rng.integers(0, 3, 7).astype(np.uint32)
    w = f.writer(location=["particles", "all"], mode="append", clock=clock)
    k = {"i": 0}
    w.on_write(lambda: pos[k["i"]], ["position"])
    w.on_write(lambda: spec, ["species"])
    w.on_write(lambda: float(k["i"]) * 0.5, ["scalar"])
    for i, step in enumerate((0, 10, 20, 30)):
        k["i"] = i
        clock.step, clock.time = step, step * 0.005
        w.write()
    t = f.writer(location=["observables"], mode="truncate")
    vol = {"v": 12.5}
    t.on_write(lambda: vol["v"], ["volume"])
    t.write()
    vol["v"] = 13.5
    t.write()   # rewritten in place
    f.write_attribute_int(["particles", "all", "box"], "dimension", 3)
    f.write_attribute_strings(["particles", "all", "box"], "boundary", ["periodic"] * 3)
    del w
    f.close()

    r = io.readers.h5md(path)
    assert r.samples(["particles", "all", "position"]) == 4
    assert r.shape("particles/all/position/value") == (4, 7, 3)
    assert r.shape("particles/all/scalar/value") == (4,)
    for i in range(4):
        assert np.array_equal(r.read_sample(["particles", "all", "position"], i), pos[i])
        assert r.step_time("particles/all/position", i) == (10 * i, 10 * i * 0.005)
    assert np.array_equal(r.read_sample("particles/all/species", 2, np.uint32), spec)
    assert r.read_dataset("observables/volume") == 13.5
    assert r.attribute_int("particles/all/box", "dimension") == 3
    # step and time are shared: the same objects linked into every field group
    for name in ("position", "species", "scalar"):
        assert r.step_time(f"particles/all/{name}", 3) == (30, 30 * 0.005)

    # append reader (readers/h5md/append.cpp:65-160)
    rd = r.reader(location=["particles", "all"], mode="append")
    got = {}
    rd.on_read(lambda a: got.__setitem__("p", a), ["position"])
    rd.read_at_step(-1)          # the last step
    assert np.array_equal(got["p"], pos[3])
    rd.read_at_step(20)
    assert np.array_equal(got["p"], pos[2])
    rd.read_at_step(-11)         # offset + last step + 1 = 20
    assert np.array_equal(got["p"], pos[2])
    rd.read_at_time(0.05)
    assert np.array_equal(got["p"], pos[1])
    rd.read_at_time(-0.1)        # offset + last time = 0.05 (relative tolerance)
    assert np.array_equal(got["p"], pos[1])
    with pytest.raises(io.H5mdError, match="nonexistent step"):
        rd.read_at_step(15)
    with pytest.raises(io.H5mdError, match="nonexistent time"):
        rd.read_at_time(0.07)
    r.close()


def test_append_rules(tmp_path):
    path = str(tmp_path / "r.h5")
    clock = _clock()
    f = io.writers.h5md(path)
    w = f.writer(location="observables", mode="append", clock=clock)
    w.on_write(lambda: 1.0, ["a"])
    w.write()
    # a field cannot join after the first sample
    w.on_write(lambda: 2.0, ["b"])
    with pytest.raises(io.H5mdError):
        w.write()
    # a sample with a changed shape is refused
    f2 = io.writers.h5md(str(tmp_path / "s.h5"))
    w2 = f2.writer(location="observables", mode="append", clock=clock)
    v = {"x": np.zeros(3)}
    w2.on_write(lambda: v["x"], ["x"])
    w2.write()
    v["x"] = np.zeros(4)
    with pytest.raises(io.H5mdError):
        w2.write()
    # duplicated steps are ambiguous for the reader
    f3 = io.writers.h5md(str(tmp_path / "d.h5"))
    w3 = f3.writer(location="observables", mode="append", clock=clock)
    w3.on_write(lambda: 1.0, ["x"])
    w3.write()
    w3.write()
    del w3
    f3.close()
    r = io.readers.h5md(str(tmp_path / "d.h5"))
    rd = r.reader(location="observables", mode="append")
    rd.on_read(lambda a: None, ["x"])
    with pytest.raises(io.H5mdError, match="ambiguous step"):
        rd.read_at_step(0)


def test_libraries_are_separate():
    # the device library does not link HDF5; the H5MD library does not link HIP
    from halmd_amd import _lib
    hip = open(_lib.LIB_PATH, "rb").read()
    assert b"H5Fcreate" not in hip
    h5 = open(io.LIB_PATH, "rb").read()
    assert b"hipMalloc" not in h5 and os.path.getsize(io.LIB_PATH) < 4 << 20
