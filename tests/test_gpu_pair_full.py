"""Untruncated forces over all pairs (forces.pair_full) through the C-ABI,
against the oracle's FP64 brute-force sum with a sharp cutoff beyond every
minimum-image distance (which leaves the potential untruncated).
"""
import numpy as np
import pytest

from halmd_amd import mdsim
from halmd_amd.workloads import system_inputs

pytestmark = pytest.mark.gpu

FORCE_TOL = 1e-5      # max_i |F_i - F_i^ref| / rms(|F^ref|)
AUX_TOL = 1e-5        # relative, summed potential energy and virial
NO_CUTOFF = 1e4       # in sigma: beyond every minimum-image distance of these boxes


def _force_error(f, fref):
    rms = np.sqrt(np.mean(np.sum(fref ** 2, axis=1)))
    return np.max(np.abs(f - fref)) / rms


def _device_potential(kind, species):
    pair = mdsim.potentials.pair
    if kind == "lj":
        return pair.lennard_jones(epsilon=1.0, sigma=1.0)
    if kind == "ka":
        return pair.lennard_jones(epsilon=[[1.0, 1.5], [1.5, 0.5]], sigma=[[1.0, 0.8], [0.8, 0.88]], species=2)
    if kind == "morse":
        return pair.morse(epsilon=1.0, sigma=1.0, r_min=1.2, distortion=1.3)
    return pair.mie(index_repulsion=10, index_attraction=4)


def _oracle_potential(oracle, kind):
    if kind == "lj":
        return oracle.Potential(oracle.POT_LJ, "sharp", 1.0, 1.0, NO_CUTOFF)
    if kind == "ka":
        return oracle.Potential(oracle.POT_LJ, "sharp", [[1.0, 1.5], [1.5, 0.5]], [[1.0, 0.8], [0.8, 0.88]],
                                NO_CUTOFF)
    if kind == "morse":
        return oracle.Potential(oracle.POT_MORSE, "sharp", 1.0, 1.0, NO_CUTOFF, r_min=1.2, distortion=1.3)
    return oracle.Potential(oracle.POT_MIE, "sharp", 1.0, 1.0, NO_CUTOFF, index_m=10, index_n=4)


def _system(n, density, kind, seed):
    L, r0, _, _ = system_inputs(n, density, 1.0, seed=seed, jitter=0.08)
    species = np.zeros(n, np.uint32)
    if kind == "ka":
        species[np.random.default_rng(seed).permutation(n)[: n // 5]] = 1
    return L, r0, species


@pytest.mark.parametrize("kind,n", [("lj", 2048), ("ka", 4000), ("morse", 1372), ("mie", 500)])
def test_pair_full_matches_fp64(oracle, kind, n):
    L, r0, species = _system(n, 0.8, kind, seed=n)
    ns = 2 if kind == "ka" else 1
    box = mdsim.box(length=L)
    p = mdsim.particle(particles=n, species=ns)
    p.data["position"] = r0
    p.data["species"] = species
    pot = _device_potential(kind, ns)
    force = mdsim.forces.pair(box=box, particle=p, potential=pot)
    assert isinstance(force, mdsim.forces.pair_full)
    f = p.data["force"]
    u = p.data["potential_energy"]
    s = p.data["stress_pot"]
    fref, uref, sref = oracle.pair_force_brute(r0, species, L, _oracle_potential(oracle, kind))
    err = _force_error(f, fref)
    print(f"pair_full {kind} N={n}: max|dF|/rms = {err:.3e}")
    assert err < FORCE_TOL
    assert abs(u.sum() - uref.sum()) / abs(uref.sum()) < AUX_TOL
    vir, vref = s[:, :3].sum(), sref[:, :3].sum()
    assert abs(vir - vref) / abs(vref) < AUX_TOL
    # off-diagonal stress: compare against the scale of the diagonal
    assert np.max(np.abs(s[:, 3:].sum(axis=0) - sref[:, 3:].sum(axis=0))) < AUX_TOL * abs(vref)
    # Newton's third law over all pairs
    rms = np.sqrt(np.mean(np.sum(fref ** 2, axis=1)))
    assert np.all(np.abs(f.sum(axis=0)) < 1e-5 * rms * np.sqrt(n))


def test_pair_full_two_instances(oracle):
    # force on instance 1 from every particle of instance 2 = (force on 1 in
    # the union) - (force on 1 from itself); moving instance 2 invalidates it
    n1, n2 = 300, 1100
    L, r0, species = _system(n1 + n2, 0.8, "lj", seed=5)
    box = mdsim.box(length=L)
    p1 = mdsim.particle(particles=n1)
    p2 = mdsim.particle(particles=n2)
    p1.data["position"] = r0[:n1]
    p2.data["position"] = r0[n1:]
    pot = _device_potential("lj", 1)
    force = mdsim.forces.pair_full(box=box, particle=(p1, p2), potential=pot)  # noqa: F841 (kept alive)
    opot = _oracle_potential(oracle, "lj")

    def reference(r):
        fu, _, _ = oracle.pair_force_brute(r, np.zeros(n1 + n2, np.uint32), L, opot)
        fs, _, _ = oracle.pair_force_brute(r[:n1], np.zeros(n1, np.uint32), L, opot)
        return fu[:n1] - fs

    f = p1.data["force"]
    fref = reference(r0)
    assert _force_error(f, fref) < FORCE_TOL
    r1 = r0.copy()
    r1[n1:] += np.random.default_rng(6).uniform(-0.05, 0.05, (n2, 3))
    p2.data["position"] = r1[n1:]
    f1 = p1.data["force"]
    fref1 = reference(r1)
    assert not np.allclose(f1, f)
    assert _force_error(f1, fref1) < FORCE_TOL


def test_pair_full_nve_conserves_energy():
    # untruncated LJ liquid, velocity Verlet: total energy stays within a small
    # band over 400 steps (dt = 0.002)
    n = 864
    L, r0, v0, _ = system_inputs(n, 0.8, 1.0, seed=3, jitter=0.02)
    box = mdsim.box(length=L)
    p = mdsim.particle(particles=n)
    p.data["position"] = r0
    p.data["velocity"] = v0
    force = mdsim.forces.pair(box=box, particle=p, potential=_device_potential("lj", 1))  # noqa: F841
    integ = mdsim.integrators.verlet(box=box, particle=p, timestep=0.002)
    core = mdsim.core(p)
    core.set_integrator(integ)

    def energy():
        v = p.data["velocity"]
        return (0.5 * np.sum(v * v) + p.data["potential_energy"].sum()) / n

    for _ in range(100):
        core.mdstep()
    e0 = energy()
    es = []
    for _ in range(6):
        for _ in range(50):
            core.mdstep()
        es.append(energy())
    print(f"pair_full NVE: E0 = {e0:.6f}, max |E - E0| = {np.max(np.abs(np.array(es) - e0)):.2e}")
    assert np.max(np.abs(np.array(es) - e0)) < 2e-4 * abs(e0) + 1e-4


def test_pair_full_rejects_truncated_potential():
    box = mdsim.box(length=[10.0, 10.0, 10.0])
    p = mdsim.particle(particles=64)
    pot = mdsim.potentials.pair.lennard_jones().truncate("shifted", cutoff=2.5)
    with pytest.raises(ValueError):
        mdsim.forces.pair_full(box=box, particle=p, potential=pot)
