"""CPU checks of the C-ABI library: it loads, exports every declared symbol,
and its host-side logic (parameter checks, FP64 potential evaluation,
wavevector shells, input generators) agrees with the oracle.  No call here
touches a GPU."""
import ctypes
import json
import os
import re
import subprocess

import numpy as np
import pytest

from halmd_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def _declared():
    text = re.sub(r"/\*.*?\*/", "", open(_lib.HEADER_PATH).read(), flags=re.S)
    return set(re.findall(r"HALMD_API[^;(]*\b(halmd_\w+)\s*\(", text))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    declared = _declared()
    assert len(declared) > 100
    for name in declared:
        assert hasattr(L, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T halmd_" in ln}
    assert exported == declared


def test_signatures_parse():
    sigs = _lib.signatures()
    assert sigs["halmd_mdsim_particle_create"][1][:3] == [ctypes.c_uint] * 3
    assert sigs["halmd_last_error"][0] is ctypes.c_char_p


def test_errors_are_status_codes():
    h = ctypes.c_void_p()
    L = _lib.lib()
    bad = np.array([10.0, -1.0, 10.0])
    rc = L.halmd_mdsim_box_create(_lib.dptr(bad), ctypes.byref(h))
    assert rc == 1
    assert b"positive" in L.halmd_last_error()
    with pytest.raises(_lib.HalmdError):
        _lib.call("halmd_mdsim_box_create", None, ctypes.byref(h))
    from halmd_amd.mdsim.potentials import pair
    lj = pair.lennard_jones(epsilon=1.0, sigma=1.0)
    with pytest.raises(ValueError):
        lj.truncate("no_such_truncation", cutoff=2.5)
    with pytest.raises(_lib.HalmdError):
        _lib.call("halmd_mdsim_potentials_pair_truncate", lj.handle, b"bogus", _lib.dptr(np.ones(1)), 0.0)


@pytest.mark.parametrize("truncation", ["sharp", "shifted", "force_shifted", "smooth_r4"])
def test_host_potential_matches_golden_tables(truncation):
    from halmd_amd.mdsim.potentials import pair
    g = json.load(open(os.path.join(GOLDEN, "potential_tables.json")))["lennard_jones"]
    lj = pair.lennard_jones(epsilon=g["epsilon"], sigma=g["sigma"]).truncate(truncation, cutoff=g["r_cut_sigma"],
                                                                             smoothing=g["h"])
    eps = np.finfo(float).eps * g["tolerance_eps"]
    for key, (a, b) in (("aa", (0, 0)), ("ab", (0, 1)), ("bb", (1, 1))):
        for r, fref, uref in g["tables"][truncation][key]:
            f, u = lj(r * r, a, b)
            for x, ref in ((f[0], fref), (u[0], uref)):
                assert (x == ref) if ref == 0 else abs(x - ref) <= eps * abs(ref), (key, r, x, ref)


def test_host_morse_matches_golden_tables():
    from halmd_amd.mdsim.potentials import pair
    g = json.load(open(os.path.join(GOLDEN, "potential_tables.json")))["morse"]
    m = pair.morse(epsilon=g["epsilon"], sigma=g["sigma"], r_min=g["r_min"], distortion=g["distortion"])
    m.truncate("shifted", cutoff=g["r_cut_sigma"])
    for key, (a, b) in (("aa", (0, 0)), ("ab", (0, 1)), ("bb", (1, 1))):
        tol = g["tolerance_eps"][key] * np.finfo(float).eps
        for r, fref, uref in g["tables"][key]:
            f, u = m(r * r, a, b)
            assert (f[0] == fref) if fref == 0 else abs(f[0] - fref) <= 2 * tol * abs(fref)
            assert (u[0] == uref) if uref == 0 else abs(u[0] - uref) <= tol * abs(uref)


def test_host_power_law_and_mie_match_golden_tables():
    from halmd_amd.mdsim.potentials import pair
    g = json.load(open(os.path.join(GOLDEN, "potential_tables.json")))
    pl = g["power_law"]
    p = pair.power_law(epsilon=pl["epsilon"], sigma=pl["sigma"], index=pl["index"])
    p.truncate("shifted", cutoff=pl["r_cut_sigma"])
    mi = g["mie"]
    m = pair.mie(epsilon=mi["epsilon"], sigma=mi["sigma"], index_repulsion=mi["index_m"],
                 index_attraction=mi["index_n"])
    m.truncate("shifted", cutoff=mi["r_cut_sigma"])
    for pot, tables, tol_of in ((p, pl["tables"], lambda a, b: (1 + pl["index"][a][b] // 2)),
                                (m, mi["tables"], lambda a, b: mi["tolerance_eps"])):
        for key, (a, b) in (("aa", (0, 0)), ("ab", (0, 1)), ("bb", (1, 1))):
            tol = tol_of(a, b) * np.finfo(float).eps
            for r, fref, uref in tables[key]:
                f, u = pot(r * r, a, b)
                for x, ref in ((f[0], fref), (u[0], uref)):
                    assert (x == ref) if ref == 0 else abs(x - ref) <= tol * abs(ref), (key, r, x, ref)


def test_host_hard_core_matches_golden_table():
    from halmd_amd.mdsim.potentials import pair
    g = json.load(open(os.path.join(GOLDEN, "potential_tables.json")))["power_law_hard_core"]
    p = pair.power_law(epsilon=g["epsilon"], sigma=g["sigma"], index=g["index"])
    p.modify("hard_core", radius=g["r_core_sigma"]).truncate("shifted", cutoff=g["r_cut_sigma"])
    for key, (a, b) in (("aa", (0, 0)), ("ab", (0, 1)), ("bb", (1, 1))):
        for r, fref, uref in g["tables"][key]:
            rs = r / g["sigma"][a][b]
            tol = np.finfo(float).eps * g["index"][a][b] * (1 + rs / (rs - g["r_core_sigma"][a][b]))
            f, u = p(r * r, a, b)
            for x, ref in ((f[0], fref), (u[0], uref)):
                assert (x == ref) if ref == 0 else abs(x - ref) <= tol * abs(ref), (key, r, x, ref)
    # modify comes before truncate, and Morse has no hard-core form
    with pytest.raises(_lib.HalmdError):
        pair.lennard_jones().truncate("shifted", cutoff=2.5).modify("hard_core", radius=0.3)
    with pytest.raises(_lib.HalmdError):
        pair.morse().modify("hard_core", radius=0.3)


def test_mie_rejects_odd_or_inverted_indices():
    from halmd_amd.mdsim.potentials import pair
    with pytest.raises(_lib.HalmdError):
        pair.mie(index_repulsion=12, index_attraction=5)
    with pytest.raises(_lib.HalmdError):
        pair.mie(index_repulsion=6, index_attraction=8)
    with pytest.raises(_lib.HalmdError):
        pair.power_law(index=0)
    # modified_lennard_jones is the reference's deprecated name of mie
    assert pair.modified_lennard_jones is pair.mie


def test_wavevector_shells_match_oracle(oracle):
    from halmd_amd import mdsim
    from halmd_amd.observables.utility import wavevector
    for L, k in ((37.35, [2.0, 4.0, 7.25, 10.0]), (107.49, np.linspace(0.0585 * 2, 10, 25))):
        box = mdsim.box(length=[L] * 3)
        w = wavevector(box, k, tolerance=0.01, max_count=8)
        q, shell = oracle.wavevector_shell(k, [L] * 3, 0.01, 8)
        keep = shell[:, 1] > shell[:, 0]
        assert np.allclose(w.wavenumber, np.asarray(k)[keep])
        assert np.array_equal(w.value, q)


def test_dense_wavevector_grid():
    from halmd_amd import mdsim
    from halmd_amd.observables.utility import wavevector
    box = mdsim.box(length=[10.0, 12.0, 14.0])
    w = wavevector(box, [1.0, 2.0, 3.0], dense=True)
    qn = np.linalg.norm(w.value, axis=1)
    assert np.all(qn < 3.0)
    lo = 0.0
    for (a, b), k in zip(w.shell, w.wavenumber):
        assert np.all((qn[a:b] >= lo) & (qn[a:b] < k))
        lo = k
    # every lattice vector below 3 is present exactly once
    u = 2 * np.pi / box.length
    n = np.floor(3.0 / u).astype(int)
    cnt = sum(1 for x in range(-n[0], n[0] + 1) for y in range(-n[1], n[1] + 1) for z in range(-n[2], n[2] + 1)
              if (x * u[0]) ** 2 + (y * u[1]) ** 2 + (z * u[2]) ** 2 < 9.0)
    assert len(w.value) == cnt


def test_input_generators_match_oracle(oracle):
    from halmd_amd.mdsim import positions, velocities
    for n, L in ((4000, [16.796] * 3), (3999, [16.796] * 3), (62500, [37.35] * 3), (1000, [9.0, 11.0, 13.0])):
        r, a = positions.fcc(n, L)
        ro, ao = oracle.lattice_fcc(n, L)
        assert a == ao
        assert np.array_equal(r, ro)
    v = velocities.maxwell_boltzmann(10000, 1.5, 1.0, seed=4)
    assert np.abs(v.mean(0)).max() < 1e-15
    assert abs(np.mean(v ** 2) - 1.5) < 0.05


def test_serving_stream_statistics():
    from halmd_amd.workloads import serving_stream
    jobs = serving_stream(rate=50.0, duration=60.0, seed=7)
    # Poisson arrivals: about rate * duration jobs, increasing times in [0, duration)
    assert 2700 < len(jobs) < 3300
    t = np.array([j["arrival"] for j in jobs])
    assert np.all(np.diff(t) > 0) and t[0] >= 0 and t[-1] < 60.0
    n = np.array([j["n"] for j in jobs])
    assert n.min() >= 1000 and n.max() <= 200000
    assert 15000 < np.median(n) < 25000
    steps = np.array([j["steps"] for j in jobs])
    assert steps.min() >= 1000 and steps.max() <= 100000
    assert {j["potential"] for j in jobs} == {"lj", "ka", "wca"}
    scaled = serving_stream(rate=50.0, duration=5.0, seed=7, step_scale=0.01)
    assert all(1 <= j["steps"] <= 1000 for j in scaled)


def test_skin_per_potential():
    # a skin for all systems, or one per potential name (the ensemble
    # autotuner's result); the specs of the batched mix draw all three names
    from halmd_amd.workloads import batched_specs, spec_skin
    specs = batched_specs(count=32, seed=5)
    assert {sp["potential"] for sp in specs} == {"lj", "wca", "morse"}
    skins = {"lj": 0.35, "wca": 0.5, "morse": 0.3}
    for sp in specs:
        assert spec_skin(0.3, sp) == 0.3
        assert spec_skin(skins, sp) == skins[sp["potential"]]
