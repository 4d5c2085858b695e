"""Density modes, blocking-scheme MSD/ISF on the GPU against the oracle."""
import numpy as np
import pytest

from halmd_amd import mdsim, observables, workloads
from halmd_amd.mdsim import positions
from halmd_amd.observables.dynamics import blocking_scheme
from halmd_amd.observables.utility import wavevector

pytestmark = pytest.mark.gpu


def test_density_modes_perfect_lattice():
    # test/unit/observables/ssf.cpp:237-242: rho_k = N at Bragg vectors of the
    # fcc crystal (|rho| within 50 N eps_float), ~0 at other box vectors
    n = 4000
    L = np.full(3, 15.0)
    box = mdsim.box(length=L)
    p = mdsim.particle(particles=n)
    r, a = positions.fcc(n, L)
    p.data["position"] = r
    b = 2 * np.pi / a
    w = wavevector(box, [2 * b, np.sqrt(3) * b, 2 * np.pi / 15.0, b], tolerance=1e-9, max_count=20)
    rho = observables.density_mode(p, w).acquire()
    # Bragg vectors of the fcc crystal: q a / 2 pi integer and all of one parity
    hkl = w.value * a / (2 * np.pi)
    integer = np.all(np.abs(hkl - np.rint(hkl)) < 1e-9, axis=1)
    par = np.rint(hkl).astype(int) % 2
    bragg = integer & (np.all(par == par[:, :1], axis=1))
    assert bragg.sum() >= 4 and (~bragg).sum() >= 4
    assert np.all(np.abs(np.abs(rho[bragg]) - n) < 50 * n * np.finfo(np.float32).eps)
    assert np.all(np.abs(rho[~bragg]) < 50 * n * np.finfo(np.float32).eps)


@pytest.mark.parametrize("n", [4000, 64 ** 3 * 4])
def test_density_modes_match_fp64(oracle, n):
    w = workloads.lennard_jones(n, density=0.8442, jitter=0.3, seed=61)
    k = np.linspace(2 * np.pi / w.length[0], 10.0, 40)
    wv = wavevector(w.box, k, tolerance=0.02, max_count=5)
    dm = observables.density_mode(w.particle, wv)
    rho = dm.acquire()
    pos = w.particle.data["position"]
    sub = slice(None) if n < 10000 else slice(0, len(wv.value), 7)
    ref = oracle.density_mode(pos, wv.value[sub])
    err = np.abs(rho[sub] - ref)
    print(f"density modes: nq={len(wv.value)}, max |d rho| = {err.max():.3e} (N={n})")
    assert np.all(err < 2e-7 * n)


def test_msd_isf_blocking_scheme_matches_oracle(oracle):
    w = workloads.lennard_jones(2048, density=0.8442, temperature=1.5, jitter=0.1, seed=71, timestep=0.005,
                                sort_interval=25)
    wv = wavevector(w.box, [3.0, 7.25], tolerance=0.02, max_count=8)
    dm = observables.density_mode(w.particle, wv)
    bs = blocking_scheme(w.box, w.particle, timestep=0.005, max_lag=200 * 0.005, resolution=2 * 0.005,
                         block_size=5, shift=2, separation=5)
    bs.correlate_msd()
    bs.correlate_isf(dm)
    oc = oracle.Correlator(200, 2, 5, 2, 5, 0, 2048, nq=len(wv.value), shell=wv.shell)
    assert list(oc.intervals()) == list(bs.interval)
    L = w.length
    for step in range(0, 301):
        if step:
            w.core.mdstep()
        # feed both correlators the same samples (positions unwrapped, ID order)
        due = any(step % iv == 0 for iv in bs.interval)
        bs.sample(step)
        if due:
            pos = w.particle.ordered("position") + w.particle.ordered("image") * L
            oc.sample(step, pos, dm.acquire())
    bs.finalise()
    oc.finalise()
    c, m, e = bs.result("msd")
    oc_c, oc_m, oc_e = oc.result("msd")
    assert np.array_equal(c, oc_c)
    nz = c > 0
    assert np.allclose(m[nz], oc_m[nz], rtol=1e-9, atol=1e-14)
    assert np.allclose(e[nz], oc_e[nz], rtol=1e-5, atol=1e-12)
    ic, im_, ie = bs.result("isf")
    oic, oim, oie = oc.result("isf")
    assert np.array_equal(ic, oic)
    assert np.allclose(im_[ic > 0], oim[ic > 0], rtol=1e-9, atol=1e-12)
    # physics: MSD grows with lag; ISF at lag 0 is S(k) > 0 and decays
    lvl0 = m[0, 1:][c[0, 1:] > 0]
    assert np.all(np.diff(lvl0) > 0)


def test_mqd_vacf_blocking_scheme_matches_oracle(oracle):
    w = workloads.lennard_jones(2048, density=0.8442, temperature=1.5, jitter=0.1, seed=73, timestep=0.005,
                                sort_interval=25)
    bs = blocking_scheme(w.box, w.particle, timestep=0.005, max_lag=200 * 0.005, resolution=2 * 0.005,
                         block_size=5, shift=2, separation=5)
    bs.correlate_msd()
    bs.correlate_mqd()
    bs.correlate_vacf()
    oc = oracle.Correlator(200, 2, 5, 2, 5, 0, 2048)
    L = w.length
    for step in range(0, 301):
        if step:
            w.core.mdstep()
        due = any(step % iv == 0 for iv in bs.interval)
        bs.sample(step)
        if due:
            pos = w.particle.ordered("position") + w.particle.ordered("image") * L
            oc.sample(step, pos, vel=w.particle.ordered("velocity"))
    bs.finalise()
    oc.finalise()
    for name in ("msd", "mqd", "vacf"):
        c, m, e = bs.result(name)
        oc_c, oc_m, oc_e = oc.result(name)
        assert np.array_equal(c, oc_c), name
        nz = c > 0
        print(f"{name}: max rel err {np.max(np.abs(m[nz] - oc_m[nz]) / np.maximum(np.abs(oc_m[nz]), 1e-300)):.2e}")
        assert np.allclose(m[nz], oc_m[nz], rtol=1e-9, atol=1e-14), name
        assert np.allclose(e[nz], oc_e[nz], rtol=1e-5, atol=1e-12), name
    # physics: VACF(0) = <v^2> = 2 e_kin (unit mass) and the VACF decays;
    # MQD(0) = MSD(0) = 0; MQD >= MSD^2 (Jensen)
    vc, vm, _ = bs.result("vacf")
    assert np.isclose(vm[0, 0], 2 * w.thermodynamics.en_kin(), rtol=0.1)
    assert np.all(np.abs(vm[vc > 0]) <= 1.05 * vm[0, 0])
    qc, qm, _ = bs.result("mqd")
    sc, sm, _ = bs.result("msd")
    assert qm[0, 0] == 0 and sm[0, 0] == 0
    nz = (qc > 0) & (sm > 0)
    assert np.all(qm[nz] >= sm[nz] ** 2 * (1 - 1e-12))


def test_sampler_density_mode_slot():
    w = workloads.lennard_jones(4000, density=0.8442, seed=72)
    wv = wavevector(w.box, np.linspace(0.5, 10, 20), tolerance=0.05, max_count=10)
    dm = observables.density_mode(w.particle, wv)
    k = w.sampler.on_sample_density_mode(dm, 10)
    w.sampler.run(30)
    steps, rho = w.sampler.density_mode_records(k)
    assert list(steps) == [0, 10, 20, 30]
    assert np.allclose(rho[-1], dm.acquire())


def test_phase_space_acquire_and_set():
    w = workloads.lennard_jones(4000, density=0.8442, temperature=1.5, seed=74, sort_interval=25)
    w.sampler.run(120)   # particles sorted (memory order != ID order) and crossing the box
    ps = observables.phase_space(w.box, w.particle)
    L = w.length
    img0 = w.particle.ordered("image")
    assert np.any(img0 != 0)
    assert not np.array_equal(w.particle.data["id"], np.arange(4000))
    r = ps.acquire_position()
    assert np.array_equal(r, w.particle.ordered("position") + img0 * L)
    v = ps.acquire_velocity()
    assert np.array_equal(v, w.particle.ordered("velocity"))
    assert np.array_equal(ps.acquire_species(), w.particle.ordered("species"))
    assert np.array_equal(ps.acquire_mass(), w.particle.ordered("mass"))
    en_pot, en_kin = w.thermodynamics.en_pot(), w.thermodynamics.en_kin()
    # set: the same physical positions shifted by whole boxes come back as
    # positions in the box with the shifts in the images; doubled velocities
    shift = np.random.default_rng(3).integers(-3, 4, r.shape)
    ps.set({"position": r + shift * L, "velocity": 2 * v})
    assert np.array_equal(w.particle.ordered("image"), img0 + shift)
    assert np.allclose(ps.acquire_position(), r + shift * L, rtol=0, atol=1e-11)
    inbox = w.particle.data["position"]
    assert np.all(np.abs(inbox) <= L / 2 * (1 + 1e-12))
    assert np.isclose(w.thermodynamics.en_pot(), en_pot, rtol=1e-6)
    assert np.isclose(w.thermodynamics.en_kin(), 4 * en_kin, rtol=1e-12)
    # species and mass round trip in ID order
    sp = np.zeros(4000, np.uint32)
    ps.set({"mass": np.full(4000, 2.0), "species": sp})
    assert np.all(ps.acquire_mass() == 2.0) and np.all(ps.acquire_species() == 0)
    with pytest.raises(Exception):
        ps.acquire("bogus")
