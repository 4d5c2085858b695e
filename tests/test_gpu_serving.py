"""Serving: jobs stream through one running ensemble; each result is the same
alone, in the mix, and on the plain per-job path."""
import numpy as np
import pytest

from halmd_amd import observables, workloads
from halmd_amd.observables.dynamics import blocking_scheme
from halmd_amd.observables.utility import wavevector
from halmd_amd.serve import server

pytestmark = pytest.mark.gpu

DT = 0.003
TICK = 50
CORR = dict(block_size=5, shift=2, separation=10)
K = [2.5, 7.25]
KTOL, KMAX = 0.02, 6


def _submit(s, spec, keep_state=True, arrival=None):
    L, r0, v0, _ = workloads.system_inputs(spec["n"], spec["density"], spec["temperature"], spec["seed"])
    return s.submit(r0, v0, L[0], spec["steps"], potential=spec["potential"], timestep=DT,
                    species=workloads.spec_species(spec), wavenumber=K, tolerance=KTOL, max_count=KMAX,
                    keep_state=keep_state, arrival=arrival, **CORR)


def _run_until_done(s, ids, max_ticks=10000):
    for _ in range(max_ticks):
        if all(s.info(j)["status"] == "done" for j in ids):
            return
        s.tick()
    raise AssertionError("jobs did not finish")


def _plain(spec):
    """The job on the per-system modules: trajectory, blocking scheme, energies."""
    w = workloads.single_from_spec(spec, timestep=DT)
    steps = -(-spec["steps"] // TICK) * TICK
    bs = blocking_scheme(w.box, w.particle, timestep=DT, max_lag=steps * DT, resolution=TICK * DT, **CORR)
    bs.correlate_msd()
    dm = observables.density_mode(w.particle, wavevector(w.box, K, tolerance=KTOL, max_count=KMAX))
    bs.correlate_isf(dm)
    w.sampler.on_sample_blocking_scheme(bs)
    w.sampler.run(steps)
    bs.finalise()
    rid = w.particle.data["reverse_id"]
    state = (w.particle.data["position"][rid], w.particle.data["velocity"][rid], w.particle.data["image"][rid])
    return w, bs, state


SPEC_KA = dict(n=4000, density=1.2, temperature=0.8, potential="ka", steps=600, seed=501)
SPEC_LJ = dict(n=6000, density=0.85, temperature=1.0, potential="lj", steps=420, seed=502)


def test_served_job_matches_plain_path():
    s = server(tick_steps=TICK)
    jid = _submit(s, SPEC_KA)
    _run_until_done(s, [jid])
    w, bs, (r_p, v_p, im_p) = _plain(SPEC_KA)
    r_s, v_s, im_s = s.state(jid)
    print(f"KA job: max |dr| = {np.max(np.abs(r_s - r_p)):.2e}")
    assert np.array_equal(r_s, r_p) and np.array_equal(v_s, v_p) and np.array_equal(im_s, im_p)
    t, m, c = s.msd(jid)
    pc, pm, _ = bs.result("msd")
    assert np.array_equal(c, pc)
    assert np.array_equal(m[c > 0], pm[pc > 0])
    assert np.allclose(t, bs.time)
    im_, ic = s.isf(jid)
    pic, pim, _ = bs.result("isf")
    assert np.array_equal(ic, pic)
    # the plain path takes its step-0 density modes before the first list
    # build (particles in ID order), the server after placing the job (cell
    # order): the same positions summed in another order
    assert np.allclose(im_[ic > 0], pim[pic > 0], rtol=1e-12, atol=1e-15)
    info = s.info(jid)
    th = w.thermodynamics
    assert abs(info["en_pot"] - th.en_pot()) < 1e-6 * abs(th.en_pot())
    assert abs(info["en_kin"] - th.en_kin()) < 1e-6 * th.en_kin()
    # physics: the MSD grows with the lag on the first level
    lvl = m[0, 1:][c[0, 1:] > 0]
    assert len(lvl) > 2 and np.all(np.diff(lvl) > 0)


def test_job_same_alone_and_in_the_mix():
    alone = server(tick_steps=TICK)
    jid = _submit(alone, SPEC_LJ)
    _run_until_done(alone, [jid])
    ref = alone.state(jid), alone.msd(jid), alone.isf(jid)

    mix = server(tick_steps=TICK)
    early = dict(n=3000, density=1.0, temperature=0.9, potential="wca", steps=150, seed=503)
    long_ = dict(n=5000, density=1.15, temperature=0.7, potential="ka", steps=900, seed=504)
    later = dict(n=2500, density=0.9, temperature=1.3, potential="lj", steps=200, seed=505)
    ids = [_submit(mix, early, keep_state=False), _submit(mix, long_, keep_state=False)]
    for _ in range(2):
        mix.tick()
    j = _submit(mix, SPEC_LJ)   # joins running jobs; one leaves before it, one after
    mix.tick()
    ids.append(_submit(mix, later, keep_state=False))
    _run_until_done(mix, ids + [j])
    got = mix.state(j), mix.msd(j), mix.isf(j)
    for a, b in zip(ref[0], got[0]):
        assert np.array_equal(a, b)
    for a, b in zip(ref[1] + ref[2], got[1] + got[2]):
        assert np.array_equal(a, b)


def test_admission_waits_for_memory():
    probe = server(tick_steps=TICK)
    _submit(probe, SPEC_LJ, keep_state=False)
    probe.tick()
    fp = probe.stats()["used_bytes"]
    assert fp > 0
    s = server(budget_bytes=int(1.5 * fp), tick_steps=TICK)
    a = _submit(s, SPEC_LJ, keep_state=False)
    b = _submit(s, dict(SPEC_LJ, seed=506), keep_state=False)
    s.tick()
    assert s.info(a)["status"] == "running" and s.info(b)["status"] == "queued"
    assert s.stats()["used_bytes"] == fp
    _run_until_done(s, [a, b])
    ia, ib = s.info(a), s.info(b)
    assert ib["admitted"] >= ia["done"]
    assert s.stats()["used_bytes"] == 0


def test_invalid_job_is_refused_and_the_queue_goes_on():
    # a Kob-Andersen job of a few hundred particles at rho = 1.2 fits fewer than
    # 3 cells of rc + skin, which the batched layout holds for one species only:
    # submit() refuses it, and the valid jobs before and after it run to the end
    s = server(tick_steps=TICK)
    a = _submit(s, dict(SPEC_LJ, steps=150), keep_state=False)
    bad = dict(n=300, density=1.2, temperature=0.8, potential="ka", steps=100, seed=507)
    with pytest.raises(Exception, match="3 cells"):
        _submit(s, bad, keep_state=False)
    b = _submit(s, dict(SPEC_KA, steps=150), keep_state=False)
    _run_until_done(s, [a, b])
    assert s.info(a)["status"] == s.info(b)["status"] == "done"
    assert s.error(a) == ""
    assert s.stats()["used_bytes"] == 0
