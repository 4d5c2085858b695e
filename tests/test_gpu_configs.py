"""The workload configurations at their stated sizes, against the FP64 oracle.

  Kob-Andersen N = 62,500 at rho = 1.2 under the Nose-Hoover chain (T = 0.7,
  dt = 0.002): full-size forces, a 100-step run (chain state, velocities)
  and the blocking-scheme MSD and ISF at k = 7.25;
  Lennard-Jones N = 2^20: full-size forces, energy and virial;
  256 batched systems (N log-uniform in [500, 32000], LJ / WCA / Morse):
  every 16th system bitwise against its own run and its forces against FP64;
  a Poisson stream of LJ / Kob-Andersen / WCA jobs through the server: a
  sample of jobs bitwise against the plain per-job path, with their MSD and
  ISF against the oracle's blocking-scheme correlator.
Cited reference checks: test/unit/mdsim/integrators/verlet_nvt_hoover.cpp:
138-262 (chain and velocities against the host), test/unit/mdsim/forces/
pair_trunc.cpp (forces against the host module).
"""
import numpy as np
import pytest
from forcecheck import force_errors

from halmd_amd import observables, workloads
from halmd_amd.observables.dynamics import blocking_scheme
from halmd_amd.observables.utility import wavevector

pytestmark = pytest.mark.gpu

FORCE_TOL = 1e-5          # max_i |F_i - F_i^ref| / rms(|F^ref|)
MEDIAN_TOL = 2e-6         # median_i |F_i - F_i^ref| / |F_i^ref|
EN_POT_TOL = 1e-6         # relative
VIRIAL_TOL = 1e-6         # relative


def _check_forces(w, pot, oracle, label):
    """forces (ID order), mean E_pot and virial against the oracle at the current positions."""
    # (the force first: its evaluation may rebuild the lists and rearrange the particles)
    f = w.particle.ordered("force")
    pos = w.particle.ordered("position")
    species = w.particle.ordered("species")
    fref, uref, sref = oracle.pair_force(pos, species, w.length, pot, skin=w.skin)
    mx, med, nband = force_errors(f, fref, pos, species, w.length, pot, oracle, FORCE_TOL)
    # the error relative to the rms force, over all particles but the worst
    # one in 10^4 (the largest errors sit on particles with the largest pair
    # forces; force_errors holds each particle to its FP32 bound)
    rms = np.sqrt(np.mean(np.sum(fref ** 2, axis=1)))
    q4 = np.quantile(np.max(np.abs(f - fref), axis=1), 0.9999) / rms
    n = len(pos)
    u, virial = uref.sum() / n, sref[:, :3].sum() / n
    du = abs(w.thermodynamics.en_pot() - u) / abs(u)
    dw = abs(w.thermodynamics.virial() - virial) / abs(virial)
    print(f"{label}: force max {mx:.2e}, 99.99% {q4:.2e}, median {med:.2e} ({nband} pairs in the cutoff "
          f"band); E_pot {du:.2e}, virial {dw:.2e}")
    assert med < MEDIAN_TOL and q4 < FORCE_TOL
    assert du < EN_POT_TOL and dw < VIRIAL_TOL
    return mx


# ----------------------------------------------------------------------------
# Kob-Andersen N = 62,500, Nose-Hoover chain
# ----------------------------------------------------------------------------
def test_kob_andersen_62500_nhc_msd_isf(oracle):
    n, T, dt, nu = 62500, 0.7, 0.002, 5.0
    w = workloads.kob_andersen(n, density=1.2, temperature=T, timestep=dt, seed=91, jitter=0.05,
                               integrator="verlet_nvt_hoover", resonance_frequency=nu)
    pot = oracle.Potential(oracle.POT_LJ, "shifted", workloads.KA_EPSILON, workloads.KA_SIGMA, 2.5)
    assert abs(w.length[0] - 37.345) < 1e-3
    assert _check_forces(w, pot, oracle, "KA N=62500 start") < FORCE_TOL

    sim = oracle.Simulation(w.position, w.velocity, w.species, np.ones(n), w.length, pot, 0.3, dt,
                            integrator=oracle.INT_NHC, temperature=T, nu=nu)
    # MSD and ISF at |k| = 7.25 (shell tolerance 0.01, at most 7 vectors;
    # examples/liquid/lennard_jones.lua:230-231) over 100 steps
    steps, res = 100, 5
    wv = wavevector(w.box, [7.25], tolerance=0.01, max_count=7)
    dm = observables.density_mode(w.particle, wv)
    bs = blocking_scheme(w.box, w.particle, timestep=dt, max_lag=steps * dt, resolution=res * dt, block_size=5,
                         shift=2, separation=5)
    bs.correlate_msd()
    bs.correlate_isf(dm)
    args = (steps, res, 5, 2, 5, 0, n)
    oc_dev = oracle.Correlator(*args, nq=len(wv.value), shell=wv.shell)    # fed with the device samples
    oc_ref = oracle.Correlator(*args, nq=len(wv.value), shell=wv.shell)    # fed with the FP64 run
    assert list(oc_dev.intervals()) == list(bs.interval)
    L = w.length
    rho_max = 0.0
    for step in range(0, steps + 1):
        if step:
            w.core.mdstep()
            sim.run(1)
        bs.sample(step)
        if any(step % iv == 0 for iv in bs.interval):
            pos = w.particle.ordered("position") + w.particle.ordered("image") * L
            rho = dm.acquire()
            rho_max = max(rho_max, np.abs(rho).max())
            oc_dev.sample(step, pos, rho)
            st = sim.state()
            oc_ref.sample(step, st["position"] + st["image"] * L, oracle.density_mode(st["position"], wv.value))
    bs.finalise()
    oc_dev.finalise()
    oc_ref.finalise()

    # chain state and velocities after 100 steps
    xi, vxi, _ = sim.nhc_state()
    print("chain device", w.integrator.position, w.integrator.velocity, "oracle", xi, vxi)
    assert np.allclose(w.integrator.position, xi, rtol=1e-4, atol=1e-10)
    assert np.allclose(w.integrator.velocity, vxi, rtol=1e-4, atol=1e-9)
    ref = sim.state()
    rid = w.particle.data["reverse_id"]
    dv = np.max(np.abs(w.particle.data["velocity"][rid] - ref["velocity"]))
    pos_d = w.particle.data["position"][rid] + w.particle.data["image"][rid] * L
    dr = np.max(np.abs(pos_d - (ref["position"] + ref["image"] * L)))
    # a pair whose distance crosses r_c within one step of each other in the
    # two runs (their positions differ by about 1e-7) feels the shifted
    # force's jump F(r_c) = 0.039 for one step in one run only: |dv| = 7.8e-5
    # for those particles (about a crossing in 5000 per step at this size),
    # shared with their neighbours by the steps that follow.  Everything else
    # is held to the float-float bounds.
    dvi = np.max(np.abs(w.particle.data["velocity"][rid] - ref["velocity"]), axis=1)
    dri = np.max(np.abs(pos_d - (ref["position"] + ref["image"] * L)), axis=1)
    off = (dvi > 1e-5) | (dri > 1e-6)
    print(f"KA NHC 100 steps: max |dr| = {dr:.2e}, max |dv| = {dv:.2e}; {off.sum()} particles beyond "
          f"1e-6 / 1e-5 (max |dv| among the others {dvi[~off].max():.2e})")
    assert np.percentile(dvi, 99) < 1e-5 and np.percentile(dri, 99) < 1e-6
    assert dv < 20 * 0.039 * dt and dr < 100 * 0.039 * dt * dt
    assert _check_forces(w, pot, oracle, "KA N=62500 after 100 NHC steps") < FORCE_TOL

    # the device correlator is the oracle's on the device samples ...
    c, m, _ = bs.result("msd")
    oc, om, _ = oc_dev.result("msd")
    assert np.array_equal(c, oc)
    nz = c > 0
    assert np.allclose(m[nz], om[nz], rtol=1e-9, atol=1e-14)
    ic, im_, _ = bs.result("isf")
    oic, oim, _ = oc_dev.result("isf")
    assert np.array_equal(ic, oic)
    assert np.allclose(im_[ic > 0], oim[ic > 0], rtol=1e-9, atol=1e-12)
    # ... and the FP64 run's within the trajectory and density-mode errors:
    # with |dr_i| <= 2 sqrt(3) 1e-6 per displacement, |d MSD| <= 2 |dr| sqrt(MSD)
    # + |dr|^2; with |d rho_k| <= 2e-7 N, |d F(k, t)| <= 2 max|rho| 2e-7
    _, rm, _ = oc_ref.result("msd")
    _, rim, _ = oc_ref.result("isf")
    dmsd = np.abs(m[nz] - rm[nz])
    disf = np.abs(im_[ic > 0] - rim[ic > 0])
    print(f"MSD vs FP64 run: max abs {dmsd.max():.2e}; ISF vs FP64 run: max abs {disf.max():.2e}")
    ddr = 2 * np.sqrt(3) * 1e-6
    assert np.all(dmsd <= 2 * ddr * np.sqrt(rm[nz]) + ddr ** 2)
    assert np.all(disf <= 4e-7 * rho_max + 1e-12)
    # physics: MSD grows with the lag at the first level
    lvl = m[0, 1:][c[0, 1:] > 0]
    assert len(lvl) >= 3 and np.all(np.diff(lvl) > 0)


# ----------------------------------------------------------------------------
# Lennard-Jones N = 2^20: forces, energy and virial at full size
# ----------------------------------------------------------------------------
def test_lennard_jones_1m_forces_match_fp64(oracle):
    n = 64 ** 3 * 4
    w = workloads.lennard_jones(n, density=0.8442, temperature=0.722, seed=92, jitter=0.1)
    assert abs(w.length[0] - 107.49) < 0.01
    pot = oracle.Potential(oracle.POT_LJ, "shifted", 1.0, 1.0, 2.5)
    # (the jittered lattice holds pairs pressed close: the per-particle bound
    # of force_errors applies, the largest error relative to the rms is stated)
    _check_forces(w, pot, oracle, "LJ N=2^20 start")
    # after 20 steps of the fused force-and-integrate schedule (rebuilds,
    # cell-order rearrangements), at the positions the device holds
    w.sampler.run(20)
    # (largest error of 2^20 particles: about 1.05e-5 of the rms force)
    assert _check_forces(w, pot, oracle, "LJ N=2^20 after 20 steps") < 2 * FORCE_TOL
    counts, _ = w.neighbour.lists()
    assert 60 < counts.mean() < 95


# ----------------------------------------------------------------------------
# 256 batched systems
# ----------------------------------------------------------------------------
def _spec_potential_oracle(oracle, spec):
    from halmd_amd.ensemble import POTENTIALS
    name = spec["potential"]
    if name == "ka":
        return oracle.Potential(oracle.POT_LJ, "shifted", workloads.KA_EPSILON, workloads.KA_SIGMA, 2.5)
    kind, eps, sigma, rc, r_min, dist = POTENTIALS[name]
    if kind == 0:
        return oracle.Potential(oracle.POT_LJ, "shifted", eps, sigma, rc)
    return oracle.Potential(oracle.POT_MORSE, "shifted", eps, sigma, rc, r_min=r_min, distortion=dist)


def test_batched_256_systems(oracle):
    specs = workloads.batched_specs(256, seed=2024)
    ntot = sum(s["n"] for s in specs)
    assert 1.0e6 < ntot < 3.5e6
    dt, steps = 0.003, 200
    e = workloads.ensemble_from_specs(specs, timestep=dt)
    e.step(steps)
    checked = 0
    for k in range(0, 256, 16):
        spec = specs[k]
        r_e, v_e, im_e = e.state(k)
        sp = workloads.spec_species(spec)
        sp = np.zeros(spec["n"], np.int32) if sp is None else sp
        L = np.full(3, (spec["n"] / spec["density"]) ** (1 / 3))
        pot = _spec_potential_oracle(oracle, spec)
        fref, _, _ = oracle.pair_force(r_e, sp, L, pot, skin=0.3)
        mx, med, _ = force_errors(e.force(k), fref, r_e, sp, L, pot, oracle, FORCE_TOL)
        # the same system alone through the per-system modules
        w = workloads.single_from_spec(spec, timestep=dt)
        w.sampler.run(steps)
        rid = w.particle.data["reverse_id"]
        r_a, v_a, im_a = (w.particle.data["position"][rid], w.particle.data["velocity"][rid],
                          w.particle.data["image"][rid])
        small = w.neighbour.algorithm == "from_particle"
        print(f"system {k}: {spec['potential']} N={spec['n']} rho={spec['density']:.3f} small={small}: "
              f"force max {mx:.2e} median {med:.2e}; rebuilds {e.rebuilds(k)} / {w.neighbour.rebuilds}")
        assert mx < FORCE_TOL and med < MEDIAN_TOL
        assert e.rebuilds(k) == w.neighbour.rebuilds
        if small:
            # all-pairs lists: the single path's split-lane kernel sums in another order
            assert np.max(np.abs((r_e + im_e * L) - (r_a + im_a * L))) < 1e-4
        else:
            assert np.array_equal(r_e, r_a) and np.array_equal(v_e, v_a) and np.array_equal(im_e, im_a)
        checked += 1
    assert checked == 16


# ----------------------------------------------------------------------------
# Poisson job stream through the server
# ----------------------------------------------------------------------------
def test_poisson_stream_jobs_match_plain_path_and_oracle(oracle):
    from halmd_amd.serve import server
    dt, tick = 0.003, 50
    corr = dict(block_size=5, shift=2, separation=10)
    K, KTOL, KMAX = list(workloads.SERVING_WAVENUMBERS), 0.01, 7
    # 0.4 s of the stream at 50 jobs/s, steps scaled by 1/100 (10 .. 1000)
    jobs = workloads.serving_stream(rate=50.0, duration=0.4, seed=2025, step_scale=0.01)
    assert 10 <= len(jobs) <= 40
    s = server(tick_steps=tick)
    ids = []
    for j in jobs:
        L, r0, v0, _ = workloads.system_inputs(j["n"], j["density"], j["temperature"], j["seed"])
        ids.append(s.submit(r0, v0, L[0], j["steps"], potential=j["potential"], timestep=dt,
                            species=workloads.spec_species(j), wavenumber=K, tolerance=KTOL, max_count=KMAX,
                            keep_state=True, arrival=0.0, **corr))
    for _ in range(100000):
        if all(s.info(i)["status"] == "done" for i in ids):
            break
        s.tick()
    assert all(s.info(i)["status"] == "done" for i in ids)
    for k in range(0, len(jobs), max(1, len(jobs) // 3)):
        spec, jid = jobs[k], ids[k]
        steps = -(-spec["steps"] // tick) * tick
        w = workloads.single_from_spec(spec, timestep=dt)
        wv = wavevector(w.box, K, tolerance=KTOL, max_count=KMAX)
        dm = observables.density_mode(w.particle, wv)
        oc = oracle.Correlator(steps, tick, corr["block_size"], corr["shift"], corr["separation"], 0, spec["n"],
                               nq=len(wv.value), shell=wv.shell)
        L = w.length
        for step in range(0, steps + 1, tick):
            if step:
                w.sampler.run(tick)
            pos = w.particle.ordered("position") + w.particle.ordered("image") * L
            oc.sample(step, pos, dm.acquire())
        oc.finalise()
        rid = w.particle.data["reverse_id"]
        r_s, v_s, im_s = s.state(jid)
        print(f"job {k}: {spec['potential']} N={spec['n']} steps={steps}")
        assert np.array_equal(r_s, w.particle.data["position"][rid])
        assert np.array_equal(v_s, w.particle.data["velocity"][rid])
        assert np.array_equal(im_s, w.particle.data["image"][rid])
        _, m, c = s.msd(jid)
        oc_c, oc_m, _ = oc.result("msd")
        assert np.array_equal(c.astype(np.int64), oc_c)
        nz = oc_c > 0
        assert np.allclose(m[nz], oc_m[nz], rtol=1e-9, atol=1e-14)
        im_, ic = s.isf(jid)
        oic, oim, _ = oc.result("isf")
        assert np.array_equal(ic.astype(np.int64), oic)
        # (step-0 density modes: the server sums them in cell order, the plain
        # path here in ID order)
        assert np.allclose(im_[oic > 0], oim[oic > 0], rtol=1e-9, atol=1e-12)
