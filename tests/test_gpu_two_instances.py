"""Neighbour lists and truncated forces between two particle instances.

The Kob-Andersen mixture split into one instance per species, with the four
force modules AA, AB, BA, BB (test/integration/modularity/two_particles.lua:
66-130): forces, energies and a short trajectory equal those of the FP64
oracle on the one-instance mixture.
"""
import numpy as np
import pytest

from halmd_amd import mdsim, observables, workloads

pytestmark = pytest.mark.gpu

FORCE_TOL = 1e-5          # max_i |F_i - F_i^ref| / rms(|F^ref|)


def _force_error(f, fref):
    rms = np.sqrt(np.mean(np.sum(fref ** 2, axis=1)))
    return np.max(np.abs(f - fref)) / rms


def _split_system(n, truncation, seed, temperature=1.0, cross_weight=0.5, sorting=True):
    L, r0, v0, _ = workloads.system_inputs(n, 1.2, temperature, seed, jitter=0.1)
    species = workloads.ka_species(n, seed)
    ia, ib = np.flatnonzero(species == 0), np.flatnonzero(species == 1)
    box = mdsim.box(length=L)
    pot = mdsim.potentials.pair.lennard_jones(epsilon=workloads.KA_EPSILON, sigma=workloads.KA_SIGMA).truncate(
        truncation, cutoff=2.5)
    parts = {}
    for label, idx, s in (("A", ia, 0), ("B", ib, 1)):
        p = mdsim.particle(particles=len(idx), species=2)
        p.data["position"] = r0[idx]
        p.data["velocity"] = v0[idx]
        p.data["species"] = np.full(len(idx), s, np.uint32)
        parts[label] = p
    binning = {k: mdsim.binning(p, box, pot.r_cut, skin=0.3) for k, p in parts.items()}
    nb, force = {}, {}
    for l1 in "AB":
        for l2 in "AB":
            p1, p2 = parts[l1], parts[l2]
            if l1 == l2:
                nb[l1 + l2] = mdsim.neighbour(particle=p1, box=box, r_cut=pot.r_cut, skin=0.3, binning=binning[l1],
                                              disable_sorting=not sorting)
                force[l1 + l2] = mdsim.forces.pair_trunc(box=box, particle=p1, potential=pot,
                                                         neighbour=nb[l1 + l2])
            else:
                nb[l1 + l2] = mdsim.neighbour(particle=(p1, p2), box=box, r_cut=pot.r_cut, skin=0.3,
                                              binning=(binning[l1], binning[l2]))
                force[l1 + l2] = mdsim.forces.pair_trunc(box=box, particle=(p1, p2), potential=pot,
                                                         neighbour=nb[l1 + l2], weight=cross_weight)
    return dict(L=L, r0=r0, v0=v0, species=species, ia=ia, ib=ib, box=box, pot=pot, parts=parts,
                binning=binning, nb=nb, force=force)


def _oracle_pot(oracle, truncation):
    return oracle.Potential(oracle.POT_LJ, truncation, workloads.KA_EPSILON, workloads.KA_SIGMA, 2.5, h=0.005)


def test_two_instance_forces_match_fp64(oracle):
    s = _split_system(4000, "shifted", seed=95)
    A, B = s["parts"]["A"], s["parts"]["B"]
    fref, uref, sref = oracle.pair_force(s["r0"], s["species"], s["L"], _oracle_pot(oracle, "shifted"), skin=0.3)
    fa, fb = A.ordered("force"), B.ordered("force")
    ea, eb = _force_error(fa, fref[s["ia"]]), _force_error(fb, fref[s["ib"]])
    print(f"two instances: force error A {ea:.2e}, B {eb:.2e}")
    assert ea < FORCE_TOL and eb < FORCE_TOL
    # energies: the cross modules carry weight 1/2, so the two instances'
    # sums add up to the mixture's
    ta = observables.thermodynamics(box=s["box"], particle=A)
    tb = observables.thermodynamics(box=s["box"], particle=B)
    na, nb_ = len(s["ia"]), len(s["ib"])
    u = ta.en_pot() * na + tb.en_pot() * nb_
    w = ta.virial() * na + tb.virial() * nb_
    assert abs(u - uref.sum()) < 1e-6 * abs(uref.sum())
    assert abs(w - sref[:, :3].sum()) < 1e-6 * abs(sref[:, :3].sum())
    # the AB lists: every B particle within r_cut(A, B) of an A particle is
    # listed, and nothing beyond r_cut + skin
    counts, entries = s["nb"]["AB"].lists()
    ra, rb = A.data["position"], B.data["position"]
    L = s["L"]
    rc = 2.5 * 0.8
    for i in range(0, len(ra), 37):
        d = rb - ra[i]
        d -= np.rint(d / L) * L
        r = np.sqrt(np.sum(d ** 2, axis=1))
        listed = set(entries[i, :counts[i]].tolist())
        assert set(np.flatnonzero(r < rc).tolist()) <= listed, i
        assert all(r[j] <= rc + 0.3 + 1e-5 for j in listed), i


def test_two_instance_trajectory_matches_fp64(oracle):
    # both instances stepped by one core (every instance integrates, then
    # every instance finalizes), across list rebuilds, Hilbert sorts and
    # cell-order rearrangements of each instance
    s = _split_system(4000, "smooth_r4", seed=96, temperature=1.2)
    A, B = s["parts"]["A"], s["parts"]["B"]
    box = s["box"]
    ia, ib, L = s["ia"], s["ib"], s["L"]
    int_a = mdsim.integrators.verlet(box=box, particle=A, timestep=0.002)
    int_b = mdsim.integrators.verlet(box=box, particle=B, timestep=0.002)
    core = mdsim.core(A)
    core.set_integrator(int_a)
    core.add_integrator(int_b)
    sim = oracle.Simulation(s["r0"], s["v0"], s["species"], np.ones(4000), L, _oracle_pot(oracle, "smooth_r4"), 0.3,
                            0.002)
    steps = 50   # as the one-instance trajectory check (test_nve_trajectory_matches_fp64)
    for _ in range(steps):
        core.mdstep()
    sim.run(steps)
    ref = sim.state()
    rref = ref["position"] + ref["image"] * L
    dr = max(np.max(np.abs(A.ordered("position") + A.ordered("image") * L - rref[ia])),
             np.max(np.abs(B.ordered("position") + B.ordered("image") * L - rref[ib])))
    dv = max(np.max(np.abs(A.ordered("velocity") - ref["velocity"][ia])),
             np.max(np.abs(B.ordered("velocity") - ref["velocity"][ib])))
    rebuilds = {k: v.rebuilds for k, v in s["nb"].items()}
    # control: the same mixture as one instance on the device, from the same
    # state, against the same FP64 run (the split adds FP32 roundings of its
    # own: the cross force is summed apart and added to the own one)
    one = mdsim.particle(particles=4000, species=2)
    one.data["position"] = s["r0"]
    one.data["velocity"] = s["v0"]
    one.data["species"] = s["species"]
    nb1 = mdsim.neighbour(particle=one, box=box, r_cut=s["pot"].r_cut, skin=0.3)
    f1 = mdsim.forces.pair_trunc(box=box, particle=one, potential=s["pot"], neighbour=nb1)
    core1 = mdsim.core(one)
    core1.set_integrator(mdsim.integrators.verlet(box=box, particle=one, timestep=0.002))
    for _ in range(steps):
        core1.mdstep()
    dr1 = np.max(np.abs(one.ordered("position") + one.ordered("image") * L - rref))
    dv1 = np.max(np.abs(one.ordered("velocity") - ref["velocity"]))
    print(f"two instances, {steps} steps: max |dr| = {dr:.2e}, max |dv| = {dv:.2e}; rebuilds {rebuilds}; "
          f"one instance: max |dr| = {dr1:.2e}, max |dv| = {dv1:.2e}")
    assert dr < 1e-6 and dv < max(1e-5, 2 * dv1)
    del f1
    assert min(rebuilds.values()) >= 3
    fa = A.ordered("force")
    assert _force_error(fa, ref["force"][ia]) < FORCE_TOL
