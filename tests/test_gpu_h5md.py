"""Trajectory and observable files, and continuation, on the GPU.

  - the reference's continuation check (test/integration/continuation/
    lennard_jones.cmake.in with compare_trajectory.cpp:43-48): a run writes its
    trajectory; a second run reads the last sample, restores the system and
    writes its first sample; both samples are equal bit for bit;
  - a restart from the file continues bit for bit: the float-float words,
    images and the Nose-Hoover chain go through the file, and the first run
    resets its particle order at the checkpoint, as the restarted run starts;
  - the thermodynamics writer records what the observables report.
"""
import numpy as np
import pytest

from halmd_amd import io, mdsim, observables, workloads

pytestmark = pytest.mark.gpu

FIELDS = ("position", "image", "velocity", "species", "mass")


def _system(n=2000, density=0.8442, integrator="verlet", seed=61, **kw):
    return workloads.lennard_jones(n, density=density, temperature=1.2, seed=seed, timestep=0.002, jitter=0.1,
                                   integrator=integrator, **kw)


def test_continuation_sample_round_trip(tmp_path):
    path1, path2 = str(tmp_path / "run.h5"), str(tmp_path / "continued.h5")
    a = _system()
    f1 = io.writers.h5md(path1)
    ps_a = observables.phase_space(a.box, a.particle)
    ps_a.writer(f1, FIELDS, every=50, sampler=a.sampler)
    a.thermodynamics.writer(f1, every=10, sampler=a.sampler)
    a.sampler.sample()
    a.sampler.run(150)
    a.sampler.finish()
    f1.close()

    # restore from the last sample (lennard_jones.lua --input)
    r = io.readers.h5md(path1)
    assert r.samples("particles/all/position") == 4
    edges = mdsim.box.reader(r, ["particles", "all"])
    reader, sample = observables.phase_space.reader(r, FIELDS)
    reader.read_at_step(-1)
    step, _ = r.step_time("particles/all/position", 3)
    assert step == 150
    thermo_samples = r.samples("observables/potential_energy")
    r.close()
    assert thermo_samples == 16   # steps 0, 10, ..., 150

    box = mdsim.box(edges=edges)
    n = len(sample["species"])
    p = mdsim.particle(particles=n, species=1)
    p.data["position"] = sample["position"]
    p.data["image"] = sample["image"]
    p.data["velocity"] = sample["velocity"]
    p.data["species"] = sample["species"]
    p.data["mass"] = sample["mass"]
    pot = mdsim.potentials.pair.lennard_jones().truncate("shifted", cutoff=2.5)
    nb = mdsim.neighbour(particle=p, box=box, r_cut=pot.r_cut, skin=0.3, disable_sorting=True)
    fm = mdsim.forces.pair_trunc(box=box, particle=p, potential=pot, neighbour=nb)
    integ = mdsim.integrators.verlet(box=box, particle=p, timestep=0.001)
    core = mdsim.core(p)
    core.set_integrator(integ)
    core.step = step
    smp = observables.sampler(core)
    f2 = io.writers.h5md(path2)
    observables.phase_space(box, p).writer(f2, FIELDS, every=1000, sampler=smp)
    smp.sample()
    smp.run(20)
    f2.close()

    r1, r2 = io.readers.h5md(path1), io.readers.h5md(path2)
    for name in FIELDS:
        dt = {"image": np.int32, "species": np.uint32}.get(name, np.float64)
        last = r1.read_sample(f"particles/all/{name}", r1.samples(f"particles/all/{name}") - 1, dt)
        first = r2.read_sample(f"particles/all/{name}", 0, dt)
        assert last.dtype == first.dtype and np.array_equal(last, first), name
    assert np.array_equal(r1.read_dataset("particles/all/box/edges"), r2.read_dataset("particles/all/box/edges"))
    assert r2.step_time("particles/all/position", 0)[0] == 150
    r1.close()
    r2.close()
    del fm, nb


def _checkpoint(w, path):
    f = io.writers.h5md(path, overwrite=True)
    ps = observables.phase_space(w.box, w.particle)
    wr = f.writer(["particles", "all"], mode="append", clock=w.core)
    for name in ("position_parts", "velocity_parts", "image", "species", "mass"):
        wr.on_write(ps.data(name), [name])
    wr.write()
    if hasattr(w.integrator, "set_chain"):
        t = f.writer(["integrator"], mode="truncate")
        t.on_write(lambda: w.integrator.position, ["chain_position"])
        t.on_write(lambda: w.integrator.velocity, ["chain_velocity"])
        t.write()
    del wr
    f.close()


def _restore(w, path):
    r = io.readers.h5md(path)
    reader, sample = observables.phase_space.reader(r, ("position_parts", "velocity_parts", "image", "species",
                                                         "mass"))
    reader.read_at_step(-1)
    step, _ = r.step_time("particles/all/position_parts", 0)
    w.particle.reset_order()   # memory index = particle ID, as the file's order
    for name, a in sample.items():
        w.particle.data[name] = a
    if hasattr(w.integrator, "set_chain"):
        w.integrator.set_chain(r.read_dataset("integrator/chain_position"),
                               r.read_dataset("integrator/chain_velocity"))
    w.core.step = step
    r.close()


@pytest.mark.parametrize("integrator", ["verlet", "verlet_nvt_hoover"])
def test_restart_continues_bit_for_bit(tmp_path, integrator):
    path = str(tmp_path / "checkpoint.h5")
    kw = dict(resonance_frequency=5.0) if integrator == "verlet_nvt_hoover" else {}
    a = _system(4000, integrator=integrator, seed=62, **kw)
    a.sampler.run(73)
    _checkpoint(a, path)
    a.particle.reset_order()
    a.sampler.run(120)

    b = _system(4000, integrator=integrator, seed=999, **kw)   # other initial state, overwritten
    _restore(b, path)
    assert b.core.step == 73
    b.sampler.run(120)
    for name in ("position_parts", "velocity_parts", "image"):
        assert np.array_equal(a.particle.ordered(name), b.particle.ordered(name)), name
    if integrator == "verlet_nvt_hoover":
        assert np.array_equal(a.integrator.position, b.integrator.position)
        assert np.array_equal(a.integrator.velocity, b.integrator.velocity)
    assert a.neighbour.rebuilds > 2


def test_thermodynamics_writer_records_the_observables(tmp_path):
    path = str(tmp_path / "thermo.h5")
    w = _system(3000, seed=63)
    f = io.writers.h5md(path)
    fields = ("potential_energy", "kinetic_energy", "internal_energy", "pressure", "temperature",
              "center_of_mass_velocity", "virial")
    w.thermodynamics.writer(f, every=25, sampler=w.sampler, fields=fields)
    ref = {}
    w.sampler.on_sample(lambda: ref.setdefault(w.core.step, (w.thermodynamics.en_pot(), w.thermodynamics.pressure())),
                        25, 0)
    w.sampler.sample()
    w.sampler.run(100)
    f.close()
    r = io.readers.h5md(path)
    assert r.attribute_int("h5md/modules/thermodynamics", "version", 2) == [1, 0]
    assert r.attribute_int("observables", "dimension") == 3
    assert r.read_dataset("observables/particle_number") == 3000
    assert np.isclose(r.read_dataset("observables/volume"), w.box.volume, rtol=1e-15)
    assert r.samples("observables/potential_energy") == 5
    for k in range(5):
        step, time = r.step_time("observables/potential_energy", k)
        assert step == 25 * k and np.isclose(time, step * 0.002, rtol=1e-15)
        u = r.read_sample("observables/potential_energy", k)
        pr = r.read_sample("observables/pressure", k)
        assert u == ref[step][0] and pr == ref[step][1]
    assert r.read_sample("observables/center_of_mass_velocity", 4).shape == (3,)
    r.close()
