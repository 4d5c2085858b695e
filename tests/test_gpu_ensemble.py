"""Batched ensembles on the GPU: each system's result is the same alone and in the mix."""
import numpy as np
import pytest
from forcecheck import force_errors

from halmd_amd import workloads

pytestmark = pytest.mark.gpu

# a small mix: the three potentials, a box under 3 cells (all-pairs lists),
# uneven sizes and densities
SPECS = [
    dict(n=4000, density=0.8442, temperature=1.2, potential="lj", seed=11),
    dict(n=3000, density=1.0, temperature=0.8, potential="wca", seed=12),
    dict(n=2500, density=0.9, temperature=1.0, potential="morse", seed=13),
    dict(n=500, density=1.2, temperature=1.5, potential="lj", seed=14),
    dict(n=13500, density=0.75, temperature=0.6, potential="lj", seed=15),
    dict(n=6000, density=1.15, temperature=1.8, potential="wca", seed=16),
]
DT = 0.003


def _alone(spec, steps):
    w = workloads.single_from_spec(spec, timestep=DT)
    w.sampler.run(steps)
    rid = w.particle.data["reverse_id"]
    pos = w.particle.data["position"][rid]
    vel = w.particle.data["velocity"][rid]
    img = w.particle.data["image"][rid]
    return pos, vel, img, w


FORCE_TOL = 1e-5          # max_i |F_i - F_i^ref| / rms(|F^ref|) against the FP64 oracle


def _oracle_potential(oracle, spec):
    from halmd_amd.ensemble import POTENTIALS
    name = spec["potential"]
    tr = spec.get("truncation", "shifted")
    if name == "ka":
        return oracle.Potential(oracle.POT_LJ, tr, workloads.KA_EPSILON, workloads.KA_SIGMA, 2.5, h=0.005)
    kind, eps, sigma, rc, r_min, dist = POTENTIALS[name]
    if kind == 0:
        return oracle.Potential(oracle.POT_LJ, tr, eps, sigma, rc, h=0.005)
    return oracle.Potential(oracle.POT_MORSE, tr, eps, sigma, rc, r_min=r_min, distortion=dist, h=0.005)


def _force_error(f, fref):
    rms = np.sqrt(np.mean(np.sum(fref ** 2, axis=1)))
    return np.max(np.abs(f - fref)) / rms


def test_each_system_same_alone_and_in_the_mix(oracle):
    steps = 150
    e = workloads.ensemble_from_specs(SPECS, timestep=DT)
    e.step(steps)
    for k, spec in enumerate(SPECS):
        r_e, v_e, im_e = e.state(k)
        r_a, v_a, im_a, w = _alone(spec, steps)
        L = w.length
        small = int(L[0] / (w.potential.r_cut[0, 0] + 0.3)) < 3
        print(f"system {k} ({spec['potential']}, N={spec['n']}, small={small}): rebuilds {e.rebuilds(k)} vs "
              f"{w.neighbour.rebuilds}; max |dr| {np.max(np.abs((r_e + im_e * L) - (r_a + im_a * L))):.2e}")
        # every system's forces at its current positions against the FP64 oracle
        # (the all-pairs path of boxes under 3 cells included)
        sp = workloads.spec_species(spec)
        sp = np.zeros(spec["n"], np.int32) if sp is None else sp
        pot = _oracle_potential(oracle, spec)
        fref, _, _ = oracle.pair_force(r_e, sp, L, pot, skin=0.3)
        err, _, _ = force_errors(e.force(k), fref, r_e, sp, L, pot, oracle, FORCE_TOL)
        print(f"  force vs FP64: {err:.2e}")
        assert err < FORCE_TOL
        if small:
            # all-pairs lists: the single path's force kernel differs in rounding
            assert np.max(np.abs((r_e + im_e * L) - (r_a + im_a * L))) < 1e-4
            assert np.max(np.abs(v_e - v_a)) < 1e-3
            assert e.rebuilds(k) == w.neighbour.rebuilds
        else:
            # the same device code and the same rebuild rule: bitwise
            assert np.array_equal(r_e, r_a)
            assert np.array_equal(v_e, v_a)
            assert np.array_equal(im_e, im_a)
            assert e.rebuilds(k) == w.neighbour.rebuilds


def test_truncations_same_alone_and_in_the_mix():
    # every truncation in batched mode, smooth_r4 included, with the systems of
    # different truncations in one ensemble: bitwise the single path
    specs = [dict(n=4000, density=0.8442, temperature=1.2, potential="lj", seed=21, truncation="smooth_r4"),
             dict(n=3000, density=1.0, temperature=0.8, potential="lj", seed=22, truncation="force_shifted"),
             dict(n=2500, density=0.9, temperature=1.0, potential="morse", seed=23, truncation="smooth_r4"),
             dict(n=5000, density=0.9, temperature=1.0, potential="lj", seed=24, truncation="sharp"),
             dict(n=4000, density=1.2, temperature=0.8, potential="ka", seed=25, truncation="smooth_r4")]
    steps = 120
    e = workloads.ensemble_from_specs(specs, timestep=DT)
    e.step(steps)
    en_pot, _ = e.energies()
    for k, spec in enumerate(specs):
        r_e, v_e, im_e = e.state(k)
        r_a, v_a, im_a, w = _alone(spec, steps)
        assert np.array_equal(r_e, r_a), (k, spec["truncation"])
        assert np.array_equal(v_e, v_a) and np.array_equal(im_e, im_a)
        u = w.thermodynamics.en_pot()
        assert abs(en_pot[k] - u) < 1e-5 * max(1.0, abs(u))


def test_ensemble_energies_match_single_path():
    e = workloads.ensemble_from_specs(SPECS, timestep=DT)
    e.step(40)
    en_pot, en_kin = e.energies()
    for k, spec in enumerate(SPECS):
        _, _, _, w = _alone(spec, 40)
        t = w.thermodynamics
        assert abs(en_pot[k] - t.en_pot()) < 1e-5 * max(1.0, abs(t.en_pot()))
        assert abs(en_kin[k] - t.en_kin()) < 1e-6 * t.en_kin()


def test_ensemble_conserves_energy():
    e = workloads.ensemble_from_specs(SPECS, timestep=DT)
    p0, k0 = e.energies()
    e.step(1500)
    p1, k1 = e.energies()
    drift = np.abs((p1 + k1) - (p0 + k0))
    print("energy change per particle:", drift)
    assert np.all(drift < 5e-3)
    # and the state reads back in the box, with finite values
    for k in range(len(SPECS)):
        r, v, _ = e.state(k)
        assert np.all(np.isfinite(r)) and np.all(np.isfinite(v))


def test_ensemble_overflow_regrows_lists():
    # a cluster at three times the mean density overflows the list capacity
    # and the LDS staging sized from the mean; the ensemble regrows both (new
    # list layout for every system) and rebuilds
    from halmd_amd.ensemble import ensemble
    rng = np.random.default_rng(3)
    e = ensemble()
    for sp in SPECS[:2]:
        L, r0, v0, _ = workloads.system_inputs(sp["n"], sp["density"], sp["temperature"], sp["seed"])
        e.add(r0, v0, L[0], potential=sp["potential"], timestep=DT)
    n, L = 6000, 24.0
    m = int(np.ceil(n ** (1 / 3)))
    g = np.stack(np.meshgrid(*[np.arange(m)] * 3, indexing="ij"), -1).reshape(-1, 3)[:n]
    r = -L / 2 + (g + 0.5) * 0.69 + rng.normal(0, 0.01, (n, 3))   # rho ~ 3 in a corner
    e.add(r, rng.normal(0, 0.5, (n, 3)), L, potential="wca", timestep=DT)
    e.commit()
    e.step(5)
    pot, kin = e.energies()
    assert np.all(np.isfinite(pot)) and np.all(np.isfinite(kin))
    # the other systems are unaffected by the relayout: same as alone
    r_e, v_e, _ = e.state(0)
    r_a, v_a, _, _ = _alone(SPECS[0], 5)
    assert np.array_equal(r_e, r_a) and np.array_equal(v_e, v_a)


# ----------------------------------------------------------------------------
# rebuilds decided on the device: the same trajectories as the host decision
# ----------------------------------------------------------------------------
def _device_and_host(specs, steps, prepare=None):
    out = []
    for device in (True, False):
        e = workloads.ensemble_from_specs(specs, timestep=DT)
        e.set_device_steps(device)
        if prepare:
            prepare(e)
        e.step(steps)
        out.append(e)
    return out


def _same_trajectories(ed, eh, count):
    for k in range(count):
        for a, b in zip(ed.state(k), eh.state(k)):
            assert np.all(np.isfinite(a)) and np.array_equal(a, b), k
        assert ed.rebuilds(k) == eh.rebuilds(k), k


def test_device_decided_steps_match_host_decided_steps():
    # stretches of 64 steps and a remainder
    ed, eh = _device_and_host(SPECS, 150)
    print("rebuilds:", [ed.rebuilds(k) for k in range(len(SPECS))], "recoveries:", ed.recoveries)
    _same_trajectories(ed, eh, len(SPECS))


def test_stretches_replayed_from_graphs_match_host_decided_steps(monkeypatch):
    # HALMD_ENSEMBLE_GRAPHS=1: each 64-step stretch captured once per key
    # (snapshot parity, pending kick) and then replayed; 5 stretches replay
    # two of the captured graphs, with the program's readbacks between them
    monkeypatch.setenv("HALMD_ENSEMBLE_GRAPHS", "1")
    ed = workloads.ensemble_from_specs(SPECS, timestep=DT)
    monkeypatch.delenv("HALMD_ENSEMBLE_GRAPHS")
    eh = workloads.ensemble_from_specs(SPECS, timestep=DT)
    eh.set_device_steps(False)
    for chunk in (128, 192):
        ed.step(chunk)
        eh.step(chunk)
        _same_trajectories(ed, eh, len(SPECS))


def test_tight_list_capacity_recovers_on_the_host():
    # capacities at the lattice's list length: the melting fluid outgrows them
    # in device-decided stretches, which stop those systems, rebuild them on
    # the host with more room and let them catch up
    def tight(e):
        for k in range(len(SPECS)):
            e.set_list_capacity(k, 0)
    ed, eh = _device_and_host(SPECS, 200, tight)
    print("recoveries:", ed.recoveries)
    assert ed.recoveries > 0
    _same_trajectories(ed, eh, len(SPECS))


def test_particle_past_the_neighbouring_cells_recovers():
    # one particle of a dilute gas crosses several cells per step: the
    # device's merge of the old cell order cannot place it, the system stops
    # and is rebuilt (full sort) on the host; the other system runs on
    from halmd_amd.ensemble import ensemble
    runs = []
    for device in (True, False):
        e = ensemble()
        L, r0, v0, _ = workloads.system_inputs(500, 0.0002, 1.0, seed=5)
        v0[0] = [1100.0, 530.0, 270.0]
        e.add(r0, v0, L[0], potential="wca", timestep=DT)
        sp = SPECS[0]
        L, r0, v0, _ = workloads.system_inputs(sp["n"], sp["density"], sp["temperature"], sp["seed"])
        e.add(r0, v0, L[0], potential=sp["potential"], timestep=DT)
        e.commit()
        e.set_device_steps(device)
        e.step(30)
        runs.append(e)
    ed, eh = runs
    print("recoveries:", ed.recoveries, "rebuilds:", ed.rebuilds(0), ed.rebuilds(1))
    assert ed.recoveries > 0 and ed.rebuilds(0) >= 25
    _same_trajectories(ed, eh, 2)


def test_segment_shapes_change_only_the_rounding():
    # the workgroup size does not enter the arithmetic: bitwise the same;
    # the particles per segment set how the last wave of a segment splits its
    # lists (the order of the FP32 sums), so another segment size gives the
    # same lists and forces that agree to rounding
    ref = workloads.ensemble_from_specs(SPECS, timestep=DT)
    e = workloads.ensemble_from_specs(SPECS, timestep=DT, segments=(200, 128))
    other = workloads.ensemble_from_specs(SPECS, timestep=DT, segments=(120, 256))
    # (from the lattice the forces cancel; after 20 steps they do not, and the
    # two runs have drifted apart only by the rounding of those steps)
    ref.step(20)
    other.step(20)
    for k in range(len(SPECS)):
        f0, f1 = ref.force(k), other.force(k)
        rms = np.sqrt(np.mean(np.sum(f0 ** 2, axis=1)))
        assert np.max(np.abs(f1 - f0)) < 1e-5 * rms, k
        assert np.max(np.abs(ref.state(k)[0] - other.state(k)[0])) < 1e-5, k
    e.step(20)
    _same_trajectories(e, ref, len(SPECS))
    ref.step(130)
    e.step(130)
    _same_trajectories(e, ref, len(SPECS))


def test_ensemble_autotuner_returns_a_measured_choice():
    best, table = workloads.tune_ensemble(SPECS[:3], timestep=DT, targets=(150, 200), threads=(256,),
                                          skins=(0.3,), class_skins=(), warmup=4, steps=16)
    assert len(table) == 2 and all(t["ms_per_step"] > 0 for t in table)
    assert best["ms_per_step"] == min(t["ms_per_step"] for t in table)


def test_ensemble_autotuner_skins_per_potential():
    # the per-potential pass: every trial is measured, the choice is one of
    # them, and its skins are those of the potentials present
    specs = SPECS[:3]
    best, table = workloads.tune_ensemble(specs, timestep=DT, targets=(200,), threads=(256,), skins=(0.3,),
                                          class_skins=(0.3, 0.5), warmup=4, steps=16)
    names = {sp["potential"] for sp in specs}
    assert isinstance(best["skin"], dict) and set(best["skin"]) == names
    assert all(t["ms_per_step"] > 0 for t in table)
    assert len(table) == 1 + len(names)   # the grid's one point, then one alternative per potential
    # kept: the grid's skin for every potential, or a measured alternative
    assert best["skin"] == {name: 0.3 for name in names} or \
        any(t["skin"] == best["skin"] and t["ms_per_step"] == best["ms_per_step"] for t in table[1:])
