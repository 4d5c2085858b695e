"""The binning's merged cell order is the sorted one, bit for bit.

Between two rebuilds the displacement check bounds every particle's move by
about half the skin, far less than a cell, so the particles reach at most the
26 cells around their old one and the new stable cell order is a merge of the
old one (binning.hip: ens_rank_kernel, ens_cell_count_kernel,
ens_cell_carry_kernel, ens_order_kernel).  A rebuild after another
rearrangement (here the Hilbert sort every 100 steps) sorts.
"""
import numpy as np
import pytest

from halmd_amd import workloads

pytestmark = pytest.mark.gpu


def _slab(sort_interval):
    # a non-cubic cell grid: 5 x 5 x 10 cells of r_c + skin
    from types import SimpleNamespace
    from halmd_amd import mdsim, observables
    rng = np.random.default_rng(8)
    L = np.array([14.5, 14.5, 29.0])
    m = np.array([13, 13, 26])
    n = int(m.prod())
    g = np.stack(np.meshgrid(*[np.arange(k) for k in m], indexing="ij"), -1).reshape(-1, 3)
    r = -L / 2 + (g + 0.5) * (L / m) + rng.normal(0, 0.02, (n, 3))
    pot = mdsim.potentials.pair.lennard_jones().truncate("shifted", cutoff=2.5)
    box = mdsim.box(length=L)
    particle = mdsim.particle(dimension=3, particles=n, species=1)
    particle.data["position"] = r
    particle.data["velocity"] = rng.normal(0, 1.0, (n, 3))
    nb = mdsim.neighbour(particle=particle, box=box, r_cut=pot.r_cut, skin=0.3, disable_sorting=True)
    force = mdsim.forces.pair_trunc(box=box, particle=particle, potential=pot, neighbour=nb)
    integ = mdsim.integrators.verlet(box=box, particle=particle, timestep=0.004)
    core = mdsim.core(particle)
    core.set_integrator(integ)
    sort = mdsim.sorts.hilbert(particle=particle, box=box)
    core.set_sort(sort, sort_interval)
    return SimpleNamespace(particle=particle, neighbour=nb, force=force, core=core, sort=sort,
                           sampler=observables.sampler(core))


@pytest.mark.parametrize("box", ["cubic", "slab"])
def test_merged_cell_order_is_the_sorted_one(monkeypatch, box):
    runs = []
    for merge in ("1", "0"):
        monkeypatch.setenv("HALMD_BINNING_MERGE", merge)
        if box == "cubic":
            w = workloads.lennard_jones(4000, density=0.8442, temperature=1.2, seed=7, timestep=0.005,
                                        sort_interval=100, jitter=0.1)
        else:
            w = _slab(sort_interval=100)
        w.sampler.run(450)
        runs.append(w)
    a, b = runs
    for name in ("position", "velocity", "image"):
        assert np.array_equal(a.particle.ordered(name), b.particle.ordered(name)), name
    assert a.neighbour.rebuilds == b.neighbour.rebuilds
    merged, sorted_ = a.neighbour.reorders
    print("rebuilds", a.neighbour.rebuilds, "merged", merged, "sorted", sorted_)
    assert merged > 0 and sorted_ >= 1
    assert b.neighbour.reorders[0] == 0
