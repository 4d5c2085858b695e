"""Parity of the HIP MD path with the oracle, through the C-ABI.

Integer results (cell lists, neighbour lists with their order, Hilbert
permutations) are compared bit-exactly with the FP32 restatements of the
device rules; floating-point results (forces, energies, virial, trajectories)
with the FP64 host-backend restatement, within the tolerances stated below.
"""
import numpy as np
import pytest

from halmd_amd import mdsim, workloads
from halmd_amd.mdsim import positions

pytestmark = pytest.mark.gpu

# per-result tolerances against the FP64 oracle
FORCE_TOL = 1e-5          # max_i |F_i - F_i^ref| / rms(|F^ref|)
MEDIAN_TOL = 2e-6         # median_i |F_i - F_i^ref| / |F_i^ref| (FP32 pair terms, ~80 per particle)
EN_POT_TOL = 1e-6         # relative, mean potential energy per particle
VIRIAL_TOL = 1e-6         # relative, mean virial per particle


def _pos4(pos, species):
    p4 = np.zeros((len(pos), 4), dtype=np.float32)
    p4[:, :3] = np.asarray(pos, dtype=np.float64).astype(np.float32)
    p4[:, 3] = np.asarray(species, dtype=np.uint32).view(np.float32)
    return p4


def _force_error(f, fref):
    rms = np.sqrt(np.mean(np.sum(fref ** 2, axis=1)))
    return np.max(np.abs(f - fref)) / rms


# ----------------------------------------------------------------------------
# binning and neighbour lists: bit-exact against the FP32 rules
# ----------------------------------------------------------------------------
def test_fixed_point_snapshot_bit_exact(oracle):
    # the integer positions every binning / neighbour / force kernel reads are
    # exactly the FP64 rule applied to the stored float-float positions
    w = workloads.kob_andersen(8788, density=1.2, jitter=0.2, seed=2)
    assert np.array_equal(w.particle.ordered("position_fixed"), oracle.fixed_point(w.position, w.length))
    w.sampler.run(20)   # written by the integrate kernel from then on
    assert np.array_equal(w.particle.data["position_fixed"],
                          oracle.fixed_point(w.particle.data["position"], w.length))


@pytest.mark.parametrize("n,rho", [(4000, 0.8442), (13500, 1.2)])
def test_cell_lists_bit_exact(oracle, n, rho):
    w = workloads.lennard_jones(n, density=rho, jitter=0.1, seed=n)
    b = w.neighbour.binning
    ncell = b.ncell
    assert list(ncell) == list(oracle.binning_ncell(w.length, 2.5, 0.3))
    begin, end, index = b.cell_lists()
    # memory order (the neighbour module keeps the particles in cell order)
    cell = oracle.fixed_cell_index(w.particle.data["position_fixed"], ncell)
    expect = np.argsort(cell, kind="stable")
    assert np.array_equal(index, expect)
    counts = np.bincount(cell, minlength=int(np.prod(ncell)))
    assert np.array_equal(end - begin, counts)
    assert np.array_equal(begin, np.concatenate([[0], np.cumsum(counts)[:-1]]))


@pytest.mark.parametrize("kind", ["lj", "ka", "lj_3cells", "lj_4cells"])
def test_neighbour_lists_bit_exact(oracle, kind):
    if kind == "lj":
        w = workloads.lennard_jones(6912, density=0.8442, jitter=0.15, seed=11)
    elif kind == "ka":
        w = workloads.kob_andersen(8788, density=1.2, jitter=0.15, seed=12)
    else:
        # the smallest boxes on cell segments: 3 cells a side (every cell is
        # in every neighbourhood, and the 16-bit minimum image picks the
        # nearest image across the whole box) and 4
        n = 864 if kind == "lj_3cells" else 1372
        w = workloads.lennard_jones(n, density=0.8442, jitter=0.15, seed=13)
    w.sampler.run(37)    # positions written by the integrate kernel
    ncell = w.neighbour.binning.ncell
    if kind.endswith("cells"):
        assert tuple(ncell) == (int(kind[3]),) * 3
    rcs = w.potential.r_cut + w.skin
    # rebuild from the current positions (which reorders the particles into
    # cell order), then compare with the rule on them in that memory order
    w.neighbour.update()
    u = w.particle.data["position_fixed"]
    species = w.particle.data["species"]
    counts, entries = w.neighbour.lists()
    ocounts, olists, m = oracle.staged_neighbour_list(u, species, ncell, w.length, rcs, entries.shape[1])
    assert m <= entries.shape[1]
    assert np.array_equal(counts, ocounts)
    for i in range(len(counts)):
        assert np.array_equal(entries[i, :counts[i]], olists[i, :counts[i]]), i


def test_neighbour_list_overflow_grows(oracle):
    # start with a list capacity far below the neighbour count: the module
    # must detect the overflow on the device and rebuild with a larger size
    # (cell-order module, first on its particle)
    n = 4000
    L = np.full(3, (n / 0.8442) ** (1 / 3))
    box = mdsim.box(length=L)
    p = mdsim.particle(particles=n)
    r, _ = positions.fcc(n, L)
    p.data["position"] = r + np.random.default_rng(5).normal(0, 0.1, r.shape)
    nb = mdsim.neighbour(particle=p, box=box, r_cut=np.array([[2.5]]), skin=0.3, occupancy=1.0,
                         disable_sorting=True)
    small = nb.size
    assert nb.entry_bytes == 2
    nb.update()
    counts, entries = nb.lists()
    assert nb.size >= counts.max() > small * 0.5
    u = p.data["position_fixed"]
    ocounts, olists, _ = oracle.staged_neighbour_list(u, p.data["species"], nb.binning.ncell, L, [[2.8]],
                                                      entries.shape[1])
    assert np.array_equal(counts, ocounts)
    assert np.array_equal(entries, olists)


def test_second_neighbour_module_global_lists(oracle):
    # a second neighbour module on the same particles does not own their order:
    # it builds global-index lists (and grows them from a small capacity);
    # forces through either list agree
    w = workloads.lennard_jones(4000, density=0.8442, seed=6, jitter=0.1)
    f_owner = w.particle.ordered("force")
    nb2 = mdsim.neighbour(particle=w.particle, box=w.box, r_cut=w.potential.r_cut, skin=0.3, occupancy=1.0,
                          disable_sorting=True)
    assert nb2.entry_bytes == 4
    nb2.update()
    counts, entries = nb2.lists()
    assert nb2.size >= counts.max()
    u = w.particle.data["position_fixed"]
    ocounts, olists, _ = oracle.fixed_neighbour_list(u, w.particle.data["species"], nb2.binning.ncell, w.length,
                                                     np.array([[2.8 ** 2]], np.float32), entries.shape[1])
    assert np.array_equal(counts, ocounts)
    assert np.array_equal(entries, olists)
    # the same potential through the second list, as the only force module;
    # both evaluations agree with FP64 (they differ in FP32 rounding only)
    w.force.close()
    f2 = mdsim.forces.pair_trunc(box=w.box, particle=w.particle, potential=w.potential, neighbour=nb2)
    f_global = w.particle.ordered("force")
    pot = oracle.Potential(oracle.POT_LJ, "shifted", 1.0, 1.0, 2.5)
    rid = w.particle.data["reverse_id"]
    fref, _, _ = oracle.pair_force(w.particle.data["position"][rid], np.zeros(4000, np.uint32), w.length, pot,
                                   skin=0.3)
    assert _force_error(f_owner, fref) < FORCE_TOL
    assert _force_error(f_global, fref) < FORCE_TOL
    assert _force_error(f_global, f_owner) < 2 * FORCE_TOL
    del f2


def test_autotune_keeps_parity(oracle):
    # the autotuner rebuilds the lists for several segment shapes; the shape it
    # keeps must give lists and forces that still follow the rules
    w = workloads.lennard_jones(13500, density=0.8442, temperature=1.0, jitter=0.1, seed=36)
    cells = w.force.tune()
    S, R, T = w.neighbour.segments
    assert cells == S >= 1 and R >= 1 and T in (128, 192, 256)
    f = w.particle.ordered("force")
    pot = oracle.Potential(oracle.POT_LJ, "shifted", 1.0, 1.0, 2.5)
    fref, _, _ = oracle.pair_force(w.position, w.species, w.length, pot, skin=0.3)
    assert _force_error(f, fref) < FORCE_TOL
    w.sampler.run(40)
    w.neighbour.update()
    counts, entries = w.neighbour.lists()
    u = w.particle.data["position_fixed"]
    ocounts, olists, _ = oracle.staged_neighbour_list(u, w.particle.data["species"], w.neighbour.binning.ncell,
                                                      w.length, [[2.8]], entries.shape[1])
    assert np.array_equal(counts, ocounts)
    assert np.array_equal(entries, olists)


def test_step_tuner_restores_the_state(oracle):
    # core.tune times MD steps of each candidate shape from a snapshot: the
    # state, step counter and physics afterwards are those before tuning
    w = workloads.lennard_jones(13500, density=0.8442, temperature=1.0, jitter=0.1, seed=37)
    w.sampler.run(15)
    step = w.core.step
    r0, v0, im0 = (w.particle.ordered(k) for k in ("position", "velocity", "image"))
    e0 = w.thermodynamics.en_tot()
    w.core.tune(steps=8)
    S, R, T = w.neighbour.segments
    assert S >= 1 and R >= 1 and T in (128, 192, 256)
    assert w.core.step == step
    for k, a in (("position", r0), ("velocity", v0), ("image", im0)):
        assert np.array_equal(w.particle.ordered(k), a), k
    # (the energy sums follow the lists of the new shape: FP32 summation order)
    assert w.thermodynamics.en_tot() == pytest.approx(e0, rel=1e-7)
    # the run continues: forces against FP64 and the rules of the lists
    f = w.particle.ordered("force")
    pot = oracle.Potential(oracle.POT_LJ, "shifted", 1.0, 1.0, 2.5)
    fref, _, _ = oracle.pair_force(r0, w.particle.ordered("species"), w.length, pot, skin=0.3)
    assert _force_error(f, fref) < FORCE_TOL
    w.sampler.run(30)
    assert abs(w.thermodynamics.en_tot() - e0) < 2e-3


@pytest.mark.parametrize("n,rho", [(500, 1.2), (256, 0.8)])
def test_from_particle_small_box(oracle, n, rho):
    # boxes under 3 cells of rc + skin: all-pairs lists (ascending j, the
    # fixed-point rule), forces and a short NVE run against FP64
    L = np.full(3, (n / rho) ** (1 / 3))
    rng = np.random.default_rng(n)
    m = int(np.ceil(n ** (1 / 3)))
    g = np.stack(np.meshgrid(*[np.arange(m)] * 3, indexing="ij"), -1).reshape(-1, 3)[:n]
    pos = -L / 2 + (g + 0.5) * (L / m) + rng.normal(0, 0.05, (n, 3))
    pot = mdsim.potentials.pair.lennard_jones(epsilon=1.0, sigma=1.0).truncate("shifted", cutoff=2.5)
    box = mdsim.box(length=L)
    p = mdsim.particle(particles=n)
    p.data["position"] = pos
    nb = mdsim.neighbour(particle=p, box=box, r_cut=pot.r_cut, skin=0.3, disable_sorting=True)
    assert nb.algorithm == "from_particle" and nb.entry_bytes == 4
    force = mdsim.forces.pair_trunc(box=box, particle=p, potential=pot, neighbour=nb)  # noqa: F841 (kept alive)
    counts, entries = nb.lists()
    u = p.data["position_fixed"]
    ocounts, olists, _ = oracle.all_pairs_neighbour_list(u, p.data["species"], L, np.array([[2.8 ** 2]], np.float32),
                                                         entries.shape[1])
    assert np.array_equal(counts, ocounts)
    assert np.array_equal(entries, olists)
    f = p.ordered("force")
    opot = oracle.Potential(oracle.POT_LJ, "shifted", 1.0, 1.0, 2.5)
    fref, _, _ = oracle.pair_force_brute(pos, np.zeros(n, np.uint32), L, opot)
    assert _force_error(f, fref) < FORCE_TOL


@pytest.mark.parametrize("fraction", [0.5, 0.15])
def test_clustered_configuration_restages(oracle, fraction):
    # particles packed into a slab of the box: cell occupancies far above the
    # mean overflow the initial LDS staging (and, for the thin slab, force
    # narrower segments); lists and forces must still follow the rules
    n = 6000
    L = np.full(3, 24.0)
    rng = np.random.default_rng(int(fraction * 100))
    # simple cubic lattice inside the slab x in [-L/2, -L/2 + fraction L)
    a = (fraction * L[0] * L[1] * L[2] / n) ** (1 / 3)
    nxs, nys = int(fraction * L[0] / a), int(L[1] / a)
    nzs = -(-n // (nxs * nys))
    g = np.stack(np.meshgrid(np.arange(nxs), np.arange(nys), np.arange(nzs), indexing="ij"), -1).reshape(-1, 3)[:n]
    step = np.array([fraction * L[0] / nxs, L[1] / nys, L[2] / nzs])
    pos = -L / 2 + (g + 0.5) * step + rng.normal(0, 0.02, (n, 3))
    pot = mdsim.potentials.pair.lennard_jones(epsilon=1.0, sigma=1.0).truncate("shifted", cutoff=2.5)
    box = mdsim.box(length=L)
    p = mdsim.particle(particles=n)
    p.data["position"] = pos
    nb = mdsim.neighbour(particle=p, box=box, r_cut=pot.r_cut, skin=0.3, disable_sorting=True)
    force = mdsim.forces.pair_trunc(box=box, particle=p, potential=pot, neighbour=nb)  # noqa: F841 (kept alive)
    f = p.ordered("force")
    counts, entries = nb.lists()
    u = p.data["position_fixed"]
    ocounts, olists, _ = oracle.staged_neighbour_list(u, p.data["species"], nb.binning.ncell, L, [[2.8]],
                                                      entries.shape[1])
    assert np.array_equal(counts, ocounts)
    assert np.array_equal(entries, olists)
    opot = oracle.Potential(oracle.POT_LJ, "shifted", 1.0, 1.0, 2.5)
    fref, _, _ = oracle.pair_force(pos, np.zeros(n, np.uint32), L, opot, skin=0.3)
    assert _force_error(f, fref) < FORCE_TOL


# ----------------------------------------------------------------------------
# forces against the FP64 host backend
# ----------------------------------------------------------------------------
def _thermalised(w, steps=300):
    """Run the device system for a while and return its state (memory order)
    as FP64 arrays: the oracle then evaluates exactly the same positions."""
    w.sampler.run(steps)
    return w.particle.data["position"], w.particle.data["species"]


def _median_rel(f, fref):
    return np.median(np.linalg.norm(f - fref, axis=1) / np.linalg.norm(fref, axis=1))


@pytest.mark.parametrize("truncation", ["shifted", "sharp", "force_shifted", "smooth_r4"])
def test_lj_forces_match_fp64(oracle, truncation):
    w = workloads.lennard_jones(6912, density=0.8442, temperature=1.2, jitter=0.05, seed=21, truncation=truncation)
    pos, species = _thermalised(w)
    pot = oracle.Potential(oracle.POT_LJ, truncation, 1.0, 1.0, 2.5, h=0.005)
    fref, uref, sref = oracle.pair_force(pos, species, w.length, pot, skin=0.3)
    f = w.particle.data["force"]
    print(f"{truncation}: max|dF|/rms = {_force_error(f, fref):.3e}, median rel = {_median_rel(f, fref):.3e}")
    assert _force_error(f, fref) < FORCE_TOL
    assert _median_rel(f, fref) < MEDIAN_TOL
    u = w.particle.data["potential_energy"]
    s = w.particle.data["stress_pot"]
    # force_shifted subtracts nearly equal terms near r_c in FP32 (the
    # reference's GPU test widens its tolerance 500-fold for it)
    # (and the virial likewise: its sum of force_cut r over all pairs is ten
    # times the virial itself and carries the FP32 rounding of force_cut)
    widen = 10 if truncation == "force_shifted" else 1
    tr, trref = s[:, :3].sum(), sref[:, :3].sum()
    print(f"{truncation}: en_pot rel = {abs(u.sum() - uref.sum()) / abs(uref.sum()):.3e}, "
          f"virial rel = {abs(tr - trref) / abs(trref):.3e}")
    assert abs(u.sum() - uref.sum()) / abs(uref.sum()) < EN_POT_TOL * widen
    assert np.max(np.abs(u - uref)) < 1e-5 * np.max(np.abs(uref))
    assert abs(tr - trref) / abs(trref) < VIRIAL_TOL * widen
    assert np.max(np.abs(s - sref)) < 1e-5 * np.max(np.abs(sref))


@pytest.mark.parametrize("n,ncell", [(864, 3), (1372, 4)])
def test_lj_forces_smallest_cell_boxes_match_fp64(oracle, n, ncell):
    # 3 and 4 cells of rc + skin a side on cell segments: the minimum image
    # of the fixed-point differences across the whole box, in the list build
    # and in the force kernel
    w = workloads.lennard_jones(n, density=0.8442, temperature=1.2, jitter=0.05, seed=22)
    assert tuple(w.neighbour.binning.ncell) == (ncell,) * 3
    pos, species = _thermalised(w)
    pot = oracle.Potential(oracle.POT_LJ, "shifted", 1.0, 1.0, 2.5, h=0.005)
    fref, uref, sref = oracle.pair_force(pos, species, w.length, pot, skin=0.3)
    f = w.particle.data["force"]
    print(f"N={n}: max|dF|/rms = {_force_error(f, fref):.3e}, median rel = {_median_rel(f, fref):.3e}")
    assert _force_error(f, fref) < FORCE_TOL
    assert _median_rel(f, fref) < MEDIAN_TOL
    u = w.particle.data["potential_energy"]
    assert abs(u.sum() - uref.sum()) / abs(uref.sum()) < EN_POT_TOL


def test_ka_forces_match_fp64(oracle):
    w = workloads.kob_andersen(8788, density=1.2, temperature=1.0, jitter=0.05, seed=22, timestep=0.002)
    pos, species = _thermalised(w)
    pot = oracle.Potential(oracle.POT_LJ, "shifted", workloads.KA_EPSILON, workloads.KA_SIGMA, 2.5)
    fref, uref, _ = oracle.pair_force(pos, species, w.length, pot, skin=0.3)
    f = w.particle.data["force"]
    print(f"KA: max|dF|/rms = {_force_error(f, fref):.3e}, median rel = {_median_rel(f, fref):.3e}")
    assert _force_error(f, fref) < FORCE_TOL
    u = w.particle.data["potential_energy"]
    assert abs(u.sum() - uref.sum()) / abs(uref.sum()) < EN_POT_TOL


def test_morse_forces_match_fp64(oracle):
    # Morse D=1, a=6/sigma, r0=sigma, r_c=2.5 sigma: sigma_M = 1/6, r_cut_sigma = 15
    pot_d = mdsim.potentials.pair.morse(epsilon=1.0, sigma=1.0 / 6, r_min=1.0, distortion=1.0).truncate(
        "shifted", cutoff=15.0)
    w = workloads.fluid(4000, 0.9, 1.0, pot_d, jitter=0.05, seed=23, timestep=0.002)
    pos, species = _thermalised(w)
    pot = oracle.Potential(oracle.POT_MORSE, "shifted", 1.0, 1.0 / 6, 15.0, r_min=1.0, distortion=1.0)
    fref, uref, _ = oracle.pair_force(pos, species, w.length, pot, skin=0.3)
    f = w.particle.data["force"]
    print(f"Morse: max|dF|/rms = {_force_error(f, fref):.3e}, median rel = {_median_rel(f, fref):.3e}")
    assert _force_error(f, fref) < FORCE_TOL
    u = w.particle.data["potential_energy"]
    assert abs(u.sum() - uref.sum()) / abs(uref.sum()) < 1e-5


@pytest.mark.parametrize("kind", ["mie", "power_law"])
def test_mie_and_power_law_forces_match_fp64(oracle, kind):
    # a thermalised fluid: Mie 10-4 (even indices), soft-sphere power law 9
    # (odd index, the sqrt branch), both shifted at 2.5 sigma
    if kind == "mie":
        pot_d = mdsim.potentials.pair.mie(index_repulsion=10, index_attraction=4).truncate("shifted", cutoff=2.5)
        pot = oracle.Potential(oracle.POT_MIE, "shifted", 1.0, 1.0, 2.5, index_m=10, index_n=4)
    else:
        pot_d = mdsim.potentials.pair.power_law(index=9).truncate("shifted", cutoff=2.5)
        pot = oracle.Potential(oracle.POT_POWER_LAW, "shifted", 1.0, 1.0, 2.5, index=9)
    w = workloads.fluid(4000, 0.85, 1.0, pot_d, jitter=0.05, seed=29, timestep=0.002)
    pos, species = _thermalised(w)
    fref, uref, _ = oracle.pair_force(pos, species, w.length, pot, skin=0.3)
    f = w.particle.data["force"]
    print(f"{kind}: max|dF|/rms = {_force_error(f, fref):.3e}, median rel = {_median_rel(f, fref):.3e}")
    assert _force_error(f, fref) < FORCE_TOL
    assert _median_rel(f, fref) < MEDIAN_TOL
    u = w.particle.data["potential_energy"]
    assert abs(u.sum() - uref.sum()) / abs(uref.sum()) < 1e-5


@pytest.mark.parametrize("kind,truncation", [("lj", "shifted"), ("lj", "force_shifted"), ("lj", "smooth_r4"),
                                             ("ka", "shifted"), ("morse", "shifted"), ("power_law", "shifted"),
                                             ("power_law_odd", "force_shifted"), ("mie", "shifted"),
                                             ("mie", "smooth_r4"), ("power_law_hard_core", "shifted"),
                                             ("lj_hard_core", "force_shifted")])
def test_pair_table(oracle, kind, truncation):
    # isolated pairs at known separations (the reference's neighbour_chain
    # idea, test/unit/mdsim/potentials/pair/gpu/neighbour_chain.hpp): the force
    # on each particle is a single pair term, compared with FP64 per pair
    rng = np.random.default_rng(7)
    g = 6
    spacing = 8.0
    L = np.full(3, g * spacing)
    centres = (np.stack(np.meshgrid(*[np.arange(g)] * 3, indexing="ij"), -1).reshape(-1, 3) + 0.5) * spacing - L / 2
    m = len(centres)
    if kind == "morse":
        pot_d = mdsim.potentials.pair.morse(epsilon=1.0, sigma=1.0 / 6, r_min=1.0, distortion=1.0)
        pot = oracle.Potential(oracle.POT_MORSE, truncation, 1.0, 1.0 / 6, 15.0, r_min=1.0, distortion=1.0)
        cut, r_lo, nsp = 15.0, 0.75, 1
    elif kind in ("power_law", "power_law_odd"):
        # binary mixture with species-dependent indices (the odd index takes
        # the sqrt branch of the power)
        idx = [[12, 9], [9, 6]] if kind == "power_law" else [[7, 9], [9, 5]]
        eps, sig = [[1.0, 0.7], [0.7, 0.5]], [[1.0, 1.1], [1.1, 1.2]]
        pot_d = mdsim.potentials.pair.power_law(epsilon=eps, sigma=sig, index=idx)
        pot = oracle.Potential(oracle.POT_POWER_LAW, truncation, eps, sig, 2.5, index=idx)
        cut, r_lo, nsp = 2.5, 0.8, 2
    elif kind == "power_law_hard_core":
        idx, core = [[6, 12], [12, 7]], [[0.375, 0.3], [0.3, 0.25]]
        eps, sig = [[1.0, 0.7], [0.7, 0.5]], [[1.0, 1.1], [1.1, 1.2]]
        pot_d = mdsim.potentials.pair.power_law(epsilon=eps, sigma=sig, index=idx).modify("hard_core", radius=core)
        pot = oracle.Potential(oracle.POT_POWER_LAW, truncation, eps, sig, 2.5, index=idx, r_core_sigma=core)
        cut, r_lo, nsp = 2.5, 0.95, 2
    elif kind == "lj_hard_core":
        pot_d = mdsim.potentials.pair.lennard_jones(epsilon=1.0, sigma=1.0).modify("hard_core", radius=0.3)
        pot = oracle.Potential(oracle.POT_LJ, truncation, 1.0, 1.0, 2.5, r_core_sigma=0.3)
        cut, r_lo, nsp = 2.5, 1.05, 1
    elif kind == "mie":
        im, in_ = [[12, 10], [10, 8]], [[6, 4], [4, 2]]
        eps, sig = [[1.0, 0.7], [0.7, 0.5]], [[1.0, 1.1], [1.1, 1.2]]
        pot_d = mdsim.potentials.pair.mie(epsilon=eps, sigma=sig, index_repulsion=im, index_attraction=in_)
        pot = oracle.Potential(oracle.POT_MIE, truncation, eps, sig, 2.5, index_m=im, index_n=in_, h=0.005)
        cut, r_lo, nsp = 2.5, 0.85, 2
    elif kind == "ka":
        pot_d = mdsim.potentials.pair.lennard_jones(epsilon=workloads.KA_EPSILON, sigma=workloads.KA_SIGMA)
        pot = oracle.Potential(oracle.POT_LJ, truncation, workloads.KA_EPSILON, workloads.KA_SIGMA, 2.5)
        cut, r_lo, nsp = 2.5, 0.75, 2
    else:
        pot_d = mdsim.potentials.pair.lennard_jones(epsilon=1.0, sigma=1.0)
        pot = oracle.Potential(oracle.POT_LJ, truncation, 1.0, 1.0, 2.5, h=0.005)
        cut, r_lo, nsp = 2.5, 0.85, 1
    pot_d.truncate(truncation, cutoff=cut, smoothing=0.005)
    sp = rng.integers(0, nsp, (m, 2))
    rc_pair = pot.r_cut[sp[:, 0], sp[:, 1]]
    r = r_lo * (rc_pair / rc_pair.max()) + rng.uniform(0, 1, m) * (rc_pair * 0.999 - r_lo * rc_pair / rc_pair.max())
    u = rng.normal(size=(m, 3))
    u /= np.linalg.norm(u, axis=1)[:, None]
    pos = np.concatenate([centres, centres + r[:, None] * u])
    species = np.concatenate([sp[:, 0], sp[:, 1]]).astype(np.uint32)
    box = mdsim.box(length=L)
    p = mdsim.particle(particles=2 * m, species=nsp)
    p.data["position"] = pos
    p.data["species"] = species
    nb = mdsim.neighbour(particle=p, box=box, r_cut=pot_d.r_cut, skin=0.3, disable_sorting=True)
    force = mdsim.forces.pair_trunc(box=box, particle=p, potential=pot_d, neighbour=nb)  # noqa: F841 (kept alive)
    f = p.ordered("force")[:m]
    rr = np.sum((pos[:m] - pos[m:]) ** 2, axis=1)
    fref = np.empty_like(f)
    for k in range(m):
        fv, _ = pot(rr[k], int(sp[k, 0]), int(sp[k, 1]))
        fref[k] = fv[0] * (pos[k] - pos[m + k])
    # relative error with the denominator floored at one force unit (epsilon /
    # sigma): near the zero crossing of the LJ force (r = 2^(1/6) sigma) the
    # pair force itself vanishes and a per-pair ratio measures nothing
    rel = np.linalg.norm(f - fref, axis=1) / np.maximum(np.linalg.norm(fref, axis=1), 1.0)
    print(f"pair table {kind}/{truncation}: max rel = {rel.max():.3e}, median rel = {np.median(rel):.3e}")
    assert np.median(rel) < 1e-6
    assert rel.max() < 1e-5


def test_forces_large_box_newton3():
    # size-independent property at the benchmark size: total force vanishes
    w = workloads.lennard_jones(64 ** 3 * 4, density=0.8442, jitter=0.1, seed=24)
    f = w.particle.data["force"]
    rms = np.sqrt(np.mean(np.sum(f ** 2, axis=1)))
    assert np.all(np.abs(f.sum(axis=0)) < 1e-3 * rms * np.sqrt(len(f)) * 1e-2)
    counts, _ = w.neighbour.lists()
    assert 60 < counts.mean() < 95


# ----------------------------------------------------------------------------
# integrators
# ----------------------------------------------------------------------------
def test_nve_trajectory_matches_fp64(oracle):
    # short horizon, before chaos amplifies rounding: positions agree closely.
    # smooth_r4 keeps the force continuous at r_c, so a pair that FP32 and FP64
    # place on different sides of the cutoff cannot kick the trajectories apart
    w = workloads.lennard_jones(4000, density=0.8442, temperature=1.2, jitter=0.05, seed=31, timestep=0.005,
                                truncation="smooth_r4")
    pot = oracle.Potential(oracle.POT_LJ, "smooth_r4", 1.0, 1.0, 2.5, h=0.005)
    sim = oracle.Simulation(w.position, w.velocity, w.species, np.ones(4000), w.length, pot, 0.3, 0.005)
    steps = 50
    w.sampler.run(steps)
    sim.run(steps)
    ref = sim.state()
    rid = w.particle.data["reverse_id"]
    pos = w.particle.data["position"][rid]
    img = w.particle.data["image"][rid]
    vel = w.particle.data["velocity"][rid]
    dr = (pos + img * w.length) - (ref["position"] + ref["image"] * w.length)
    print(f"max |dr| = {np.max(np.abs(dr)):.3e}, max |dv| = {np.max(np.abs(vel - ref['velocity'])):.3e}")
    assert np.max(np.abs(dr)) < 1e-6
    assert np.max(np.abs(vel - ref["velocity"])) < 1e-5


def test_rebuild_rule_matches_fp64(oracle):
    # max_displacement (max_displacement.cpp:60-106) and the rebuild rule
    # (from_binning.cpp:130-131): the displacement the lists are checked
    # against is max_i |r_i - r0_i| (minimum image) of the stored positions
    # since the last rebuild, within FP32 rounding (it is taken on the hi
    # words), and the lists are rebuilt exactly at the steps where it exceeds
    # skin / 2.  Over the first 100 steps the rebuild steps are also those of
    # the FP64 oracle run from the same input.
    w = workloads.lennard_jones(4000, density=0.8442, temperature=1.2, seed=36, timestep=0.005, jitter=0.05,
                                truncation="smooth_r4")
    pot = oracle.Potential(oracle.POT_LJ, "smooth_r4", 1.0, 1.0, 2.5, h=0.005)
    sim = oracle.Simulation(w.position, w.velocity, w.species, np.ones(4000), w.length, pot, 0.3, 0.005)
    L, half = w.length, w.skin / 2
    band = 5e-6            # FP32 rounding of the hi-word displacement
    disp = w.neighbour.displacement
    w.particle.data["force"]                 # lists built at the input positions
    sim.state()
    r0 = w.particle.ordered("position")
    prev, sched, osched = w.neighbour.rebuilds, [], []
    oprev = sim.rebuilds()
    assert prev == oprev == 1
    for step in range(1, 301):
        w.sampler.run(1)
        r = w.particle.ordered("position")   # (reading evaluates the owed force: the list check)
        dr = r - r0
        dr -= np.rint(dr / L) * L
        m = np.sqrt(np.max(np.sum(dr ** 2, axis=1)))
        now = w.neighbour.rebuilds
        if now > prev:
            assert now == prev + 1
            assert m > half or abs(m - half) < band, (step, m)
            assert disp.compute() == 0.0
            r0 = r
            sched.append(step)
        else:
            assert m <= half or abs(m - half) < band, (step, m)
            assert abs(disp.compute() - m) < band, (step, disp.compute(), m)
        prev = now
        if step <= 100:
            sim.run(1)
            if sim.rebuilds() > oprev:
                osched.append(step)
            oprev = sim.rebuilds()
    print(f"rebuild steps: {sched}")
    assert len(sched) >= 10
    assert [s for s in sched if s <= 100] == osched


def test_device_rebuild_leaves_displacement_current():
    # all-pairs lists rebuilt on the device reset r0 there: the displacement
    # module reports the displacement from the reference positions now in
    # memory (0 right after a rebuild, below skin / 2 otherwise)
    w = workloads.kob_andersen(2048, density=1.2, temperature=1.0, timestep=0.002, seed=18, binning=False)
    assert w.neighbour.algorithm == "from_particle"
    disp = w.neighbour.displacement
    w.particle.data["force"]
    prev, hits = w.neighbour.rebuilds, 0
    for _ in range(200):
        w.sampler.run(1)
        w.particle.data["force"]            # evaluates the owed force: the device decides
        now = w.neighbour.rebuilds
        d = disp.compute()
        if now > prev:
            assert d == 0.0
            hits += 1
        else:
            assert d <= w.skin / 2
        prev = now
    assert hits >= 3


@pytest.mark.parametrize("n, steps", [(6912, 300), (256000, 150)])
def test_speculative_force_is_bitwise_identical(n, steps):
    # the force launched before the displacement check comes back (and
    # skipped on the device when the lists are stale) gives the trajectory of
    # the waiting schedule bit for bit, across many list rebuilds; the
    # speculative launches store the displacement maximum into host-mapped
    # memory (workgroup 0 of about a thousand at N = 256,000), the waiting
    # schedule copies the words after the stream
    runs = []
    for spec in (True, False):
        w = workloads.lennard_jones(n, density=0.8442, temperature=1.5, seed=81, timestep=0.005)
        w.particle.set_speculative_force(spec)
        w.sampler.run(steps)
        runs.append((w.particle.ordered("position"), w.particle.ordered("velocity"),
                     w.thermodynamics.en_pot(), w.neighbour.rebuilds))
    assert runs[0][3] == runs[1][3] and runs[0][3] >= 10
    assert np.array_equal(runs[0][0], runs[1][0]) and np.array_equal(runs[0][1], runs[1][1])
    assert runs[0][2] == runs[1][2]


def test_device_decided_rebuild_is_bitwise_identical(oracle):
    # all-pairs lists that rebuild on the device (no host displacement check)
    # give the trajectory, the rebuild count and the lists of the host-checked
    # schedule bit for bit; the lists stay the oracle's
    runs = []
    for device in (True, False):
        w = workloads.kob_andersen(2048, density=1.2, temperature=1.0, timestep=0.002, seed=17, binning=False)
        assert w.neighbour.algorithm == "from_particle"
        w.particle.set_speculative_force(device)
        w.sampler.run(400)
        counts, entries = w.neighbour.lists()
        runs.append((w.particle.ordered("position"), w.particle.ordered("velocity"), w.neighbour.rebuilds,
                     counts, entries, w.particle.data["position_fixed"], w.particle.data["species"]))
    assert runs[0][2] == runs[1][2] and runs[0][2] >= 10
    assert np.array_equal(runs[0][0], runs[1][0]) and np.array_equal(runs[0][1], runs[1][1])
    assert np.array_equal(runs[0][3], runs[1][3]) and np.array_equal(runs[0][4], runs[1][4])
    # the lists were built at the positions of the last rebuild; at the current
    # positions every pair within its cutoff must be listed
    u, sp = runs[0][5], runs[0][6]
    L = w.length
    rc = 2.5 * np.array([[1.0, 0.8], [0.8, 0.88]])   # the KA cutoffs 2.5 sigma_ab
    ocounts, olists, _ = oracle.all_pairs_neighbour_list(u, sp, L, (rc ** 2).astype(np.float32),
                                                         runs[0][4].shape[1])
    for i in range(0, len(u), 97):
        assert set(olists[i, :ocounts[i]]) <= set(runs[0][4][i, :runs[0][3][i]])


def test_device_decided_rebuild_after_host_position_change(oracle):
    # positions set from the host between device-decided steps: the host check
    # takes over (against the reference positions the device last set), and the
    # forces before and after further steps are those of the FP64 oracle
    def run(steps):
        w = workloads.kob_andersen(2048, density=1.2, temperature=1.0, timestep=0.002, seed=19, binning=False)
        w.sampler.run(steps)
        return w
    # positions of the same trajectory 60 steps on (a liquid configuration,
    # displaced past half the skin), set into the run at step 120
    pos = run(180).particle.data["position"]
    w = run(120)
    L = w.length
    w.particle.data["position"] = pos
    species = w.particle.data["species"]
    pot = oracle.Potential(oracle.POT_LJ, "shifted", workloads.KA_EPSILON, workloads.KA_SIGMA, 2.5)
    fref, _, _ = oracle.pair_force(pos, species, L, pot, skin=0.3)
    assert _force_error(w.particle.data["force"], fref) < FORCE_TOL
    w.sampler.run(60)
    pos = w.particle.data["position"]
    fref, _, _ = oracle.pair_force(pos, species, L, pot, skin=0.3)
    assert _force_error(w.particle.data["force"], fref) < FORCE_TOL


@pytest.mark.parametrize("kind", ["lj", "ka", "ka_all_pairs", "lj_256k"])
def test_fused_step_is_bitwise_identical(kind):
    # the force kernel that integrates each particle as it finishes it gives
    # the trajectory of separate force and integrate kernels bit for bit,
    # across rebuilds, Hilbert sorts and samples taken in between; the fused
    # launches publish their displacement maximum to the host (the last of
    # about a thousand workgroups at N = 256,000), the separate ones are read
    # back by a copy, and both must ask for the same rebuilds
    runs = []
    for fused in (True, False):
        if kind == "lj":
            w = workloads.lennard_jones(6912, density=0.8442, temperature=1.5, seed=82, sort_interval=40)
        elif kind == "lj_256k":
            w = workloads.lennard_jones(256000, density=0.8442, temperature=1.5, seed=85, sort_interval=100)
        elif kind == "ka_all_pairs":
            # all-pairs lists (split-lane force kernel, rebuilds decided on the device)
            w = workloads.kob_andersen(3000, density=1.2, temperature=0.9, seed=84, binning=False)
        else:
            w = workloads.kob_andersen(8788, density=1.2, temperature=0.9, seed=83, sort_interval=40)
        w.particle.set_fused_step(fused)
        s = w.sampler.on_sample_thermodynamics(w.thermodynamics, 25)
        w.sampler.run(260)
        runs.append((w.particle.ordered("position"), w.particle.ordered("velocity"), w.particle.ordered("image"),
                     w.particle.ordered("force"), w.sampler.thermodynamics_records(s), w.neighbour.rebuilds))
    a, b = runs
    for k in range(5):
        assert np.array_equal(a[k], b[k]), k
    assert a[5] == b[5] > 5


def test_deferred_finalize_is_bitwise_identical():
    # the finalize kick folded into the next integrate pass computes exactly
    # what the separate pass computes, also across rebuilds, sorts and samples
    a = workloads.lennard_jones(4000, density=0.8442, temperature=1.2, seed=35, sort_interval=20)
    b = workloads.lennard_jones(4000, density=0.8442, temperature=1.2, seed=35, sort_interval=20)
    b.integrator.defer_finalize(False)
    sa = a.sampler.on_sample_thermodynamics(a.thermodynamics, 7)
    sb = b.sampler.on_sample_thermodynamics(b.thermodynamics, 7)
    a.sampler.run(90)
    b.sampler.run(90)
    for name in ("position", "velocity", "image", "id", "force"):
        assert np.array_equal(a.particle.data[name], b.particle.data[name]), name
    assert np.array_equal(a.sampler.thermodynamics_records(sa), b.sampler.thermodynamics_records(sb))
    assert a.neighbour.rebuilds == b.neighbour.rebuilds > 1


def test_nve_energy_conservation_and_drift(oracle):
    w = workloads.lennard_jones(4000, density=0.8442, temperature=1.2, seed=32, timestep=0.005)
    slot = w.sampler.on_sample_thermodynamics(w.thermodynamics, 100)
    w.sampler.run(4000)
    rec = w.sampler.thermodynamics_records(slot)
    etot = rec[:, 1] + rec[:, 2]
    t = rec[:, 0] * 0.005
    late = t >= 5.0
    slope = np.polyfit(t[late], etot[late], 1)[0]
    # shifted (force-discontinuous) cutoff: compare with the FP64 oracle's own drift
    pot = oracle.Potential(oracle.POT_LJ, "shifted", 1.0, 1.0, 2.5)
    sim = oracle.Simulation(w.position, w.velocity, w.species, np.ones(4000), w.length, pot, 0.3, 0.005)
    orec = sim.run(4000, every=100, record0=True)
    oslope = np.polyfit(orec[late, 0] * 0.005, (orec[:, 1] + orec[:, 2])[late], 1)[0]
    print(f"NVE drift per particle per tau: device {slope:.3e}, FP64 oracle {oslope:.3e}")
    assert abs(slope) < 1e-4
    assert abs(slope - oslope) < 1e-5
    assert np.max(np.abs(etot[late] - etot[late].mean())) < 2e-3


def test_nhc_matches_fp64(oracle):
    w = workloads.lennard_jones(4000, density=0.8442, temperature=1.0, seed=33, timestep=0.002,
                                integrator="verlet_nvt_hoover", resonance_frequency=5.0, jitter=0.05,
                                truncation="smooth_r4")
    pot = oracle.Potential(oracle.POT_LJ, "smooth_r4", 1.0, 1.0, 2.5, h=0.005)
    sim = oracle.Simulation(w.position, w.velocity, w.species, np.ones(4000), w.length, pot, 0.3, 0.002,
                            integrator=oracle.INT_NHC, temperature=1.0, nu=5.0)
    steps = 40
    w.sampler.run(steps)
    sim.run(steps)
    xi, vxi, _ = sim.nhc_state()
    print("chain device", w.integrator.position, w.integrator.velocity, "oracle", xi, vxi)
    assert np.allclose(w.integrator.position, xi, rtol=1e-4, atol=1e-10)
    assert np.allclose(w.integrator.velocity, vxi, rtol=1e-4, atol=1e-9)
    rid = w.particle.data["reverse_id"]
    vel = w.particle.data["velocity"][rid]
    assert np.max(np.abs(vel - sim.state()["velocity"])) < 1e-5
    # conserved extended energy over a longer run
    slot = w.sampler.on_sample_thermodynamics(w.thermodynamics, 50)
    w.sampler.run(3000)
    rec = w.sampler.thermodynamics_records(slot)
    econs = rec[:, 1] + rec[:, 2] + rec[:, 4]
    assert np.max(np.abs(econs - econs[0])) < 5e-3


def test_andersen_matches_fp64(oracle):
    # the heat bath draws from a counter generator of (seed, step, particle id),
    # which the oracle restates: the same particles collide at the same steps,
    # and the trajectory (resampled velocities included) is the FP64 one: a
    # collision set that differed anywhere would leave O(1) velocity differences
    w = workloads.lennard_jones(4000, density=0.8442, temperature=1.2, seed=34, timestep=0.005, jitter=0.05,
                                truncation="smooth_r4", integrator="verlet_nvt_andersen", collision_rate=5.0)
    pot = oracle.Potential(oracle.POT_LJ, "smooth_r4", 1.0, 1.0, 2.5, h=0.005)
    sim = oracle.Simulation(w.position, w.velocity, w.species, np.ones(4000), w.length, pot, 0.3, 0.005,
                            integrator=oracle.INT_ANDERSEN, temperature=1.2, rate=5.0, seed=34)
    steps = 40
    w.sampler.run(steps)
    sim.run(steps)
    ref = sim.state()
    rid = w.particle.data["reverse_id"]
    vel = w.particle.data["velocity"][rid]
    pos = w.particle.data["position"][rid] + w.particle.data["image"][rid] * w.length
    dv = np.max(np.abs(vel - ref["velocity"]))
    dr = np.max(np.abs(pos - (ref["position"] + ref["image"] * w.length)))
    # ~ N steps rate dt = 4000 collisions, all at the oracle's particles and steps
    print(f"Andersen vs FP64: {sim.collisions()} collisions, max |dv| = {dv:.3e}, max |dr| = {dr:.3e}")
    assert 3000 < sim.collisions() < 5000
    assert dv < 1e-5 and dr < 1e-6


def test_andersen_ideal_gas_statistics():
    # test/unit/mdsim/integrators/verlet_nvt_andersen.cpp:101-200 (GPU set-up:
    # 5000 particles without interactions, box ratios 1 : 2 : 1.01, rho = 0.3,
    # dt = 0.01, T = 1, collision rate 10, 500 tau sampled every 3 collision
    # times): centre-of-mass velocity, temperature and specific heat within 4.5
    # sigma, and the collision count within 4.5 sigma of N steps rate dt
    import andersen_checks as ac
    from halmd_amd import observables
    from halmd_amd.mdsim import velocities
    n = 5000
    L = ac.box_edges(n)
    box = mdsim.box(length=L)
    p = mdsim.particle(particles=n)
    r, _ = positions.fcc(n, L)
    p.data["position"] = r
    p.data["velocity"] = velocities.maxwell_boltzmann(n, ac.TEMPERATURE, seed=71)
    integ = mdsim.integrators.verlet_nvt_andersen(box=box, particle=p, timestep=ac.TIMESTEP,
                                                  temperature=ac.TEMPERATURE, rate=ac.RATE, seed=72)
    core = mdsim.core(p)
    core.set_integrator(integ)
    thermo = observables.thermodynamics(box=box, particle=p, integrator=integ)
    sampler = observables.sampler(core)
    steps = int(np.ceil(500 / ac.TIMESTEP))
    period = ac.sample_period()
    temps, vcms = [], []
    changed = 0
    for k in range(0, steps, period):
        if k == 0:
            # without forces a velocity changes only in a collision: count them
            # over one step
            v_before = p.data["velocity"]
            sampler.run(1)
            changed = int(np.sum(np.any(p.data["velocity"] != v_before, axis=1)))
        else:
            sampler.run(period)
        temps.append(thermo.temperature())
        vcms.append(thermo.center_of_mass_velocity())
    done = 1 + (len(temps) - 1) * period
    sampler.run(steps - done)
    ac.check(temps, vcms, thermo.temperature(), thermo.center_of_mass_velocity(), n)
    # one step's collisions: binomial(N, rate dt)
    pc = float(np.float32(ac.RATE * ac.TIMESTEP))
    print(f"collisions in one step: {changed} (expected {n * pc:.0f})")
    assert abs(changed - n * pc) < 4.5 * np.sqrt(n * pc * (1 - pc))


# ----------------------------------------------------------------------------
# Hilbert sort
# ----------------------------------------------------------------------------
def test_hilbert_order_bit_exact(oracle):
    w = workloads.lennard_jones(13500, density=0.8442, jitter=0.2, seed=41)
    s = mdsim.sorts.hilbert(particle=w.particle, box=w.box)
    assert s.depth == oracle.hilbert_depth(w.length)
    f_before = w.particle.ordered("force")
    # memory order before the sort (the force evaluation left it in cell order)
    p_before = w.particle.data["position"]
    id_before = w.particle.data["id"]
    code = oracle.fp32_hilbert_code(_pos4(p_before, w.particle.data["species"]), w.length, s.depth)
    s.order()
    perm = np.argsort(code, kind="stable")
    assert np.array_equal(w.particle.data["id"], id_before[perm])
    rid = w.particle.data["reverse_id"]
    assert np.array_equal(rid[id_before[perm]], np.arange(len(perm)))
    assert np.array_equal(w.particle.data["position"], p_before[perm])
    # forces in ID order are unchanged by the reordering up to the FP32
    # summation order of the rebuilt lists
    f_after = w.particle.ordered("force")
    assert _force_error(f_after, f_before) < FORCE_TOL


def test_sorted_run_matches_unsorted():
    a = workloads.lennard_jones(4000, density=0.8442, seed=42, jitter=0.05, sort_interval=10)
    b = workloads.lennard_jones(4000, density=0.8442, seed=42, jitter=0.05)
    a.sampler.run(30)
    b.sampler.run(30)
    pa = a.particle.ordered("position") + a.particle.ordered("image") * a.length
    pb = b.particle.ordered("position") + b.particle.ordered("image") * b.length
    # identical physics; only the FP32 summation order of the forces differs
    assert np.max(np.abs(pa - pb)) < 1e-6


# ----------------------------------------------------------------------------
# thermodynamics
# ----------------------------------------------------------------------------
def test_thermodynamics_match_fp64(oracle):
    w = workloads.kob_andersen(8788, density=1.2, temperature=0.7, jitter=0.1, seed=51)
    pot = oracle.Potential(oracle.POT_LJ, "shifted", workloads.KA_EPSILON, workloads.KA_SIGMA, 2.5)
    sim = oracle.Simulation(w.position, w.velocity, w.species, np.ones(8788), w.length, pot, 0.3, 0.005)
    rec = sim.run(0, every=1, record0=True)[0]
    t = w.thermodynamics
    assert abs(t.en_pot() - rec[1]) < EN_POT_TOL * abs(rec[1])
    assert abs(t.en_kin() - rec[2]) < 1e-12 * abs(rec[2])
    assert abs(t.virial() - rec[3]) < VIRIAL_TOL * abs(rec[3])
    assert np.allclose(t.center_of_mass_velocity(), 0, atol=1e-14)
    rho = 8788 / np.prod(w.length)
    assert abs(t.pressure() - rho * (2 * rec[2] / 3 + rec[3] / 3)) < 1e-5 * abs(t.pressure())
