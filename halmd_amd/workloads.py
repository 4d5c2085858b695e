"""Named workloads built with the module API (synthetic, seeded inputs).

Each builder returns a namespace holding the modules of one simulation and the
FP64 initial state (positions, velocities, species) so that any checker can be
started from exactly the same input.
"""
from types import SimpleNamespace

import numpy as np

from . import mdsim, observables
from .mdsim import positions, velocities

__all__ = ["fluid", "lennard_jones", "kob_andersen", "KA_EPSILON", "KA_SIGMA", "serving_stream"]

KA_EPSILON = np.array([[1.0, 1.5], [1.5, 0.5]])
KA_SIGMA = np.array([[1.0, 0.8], [0.8, 0.88]])


def system_inputs(n, density, temperature, seed=1, jitter=0.0):
    """(box edges, positions, velocities, lattice constant) of an fcc fluid
    start: lattice at `density`, optional uniform jitter (in lattice
    constants), Maxwell-Boltzmann velocities at `temperature` (seed + 1)."""
    L = np.full(3, (n / density) ** (1.0 / 3))
    r0, a = positions.fcc(n, L)
    rng = np.random.default_rng(seed)
    if jitter:
        r0 = r0 + rng.uniform(-jitter, jitter, r0.shape) * a
        r0 = (r0 + L / 2) % L - L / 2
    v0 = velocities.maxwell_boltzmann(n, temperature, 1.0, seed=seed + 1)
    return L, r0, v0, a


def fluid(nparticle, density, temperature, potential, skin=0.3, timestep=0.005, seed=1, species=None,
          integrator="verlet", resonance_frequency=5.0, collision_rate=1.0, sort_interval=0, hilbert_on_rebuild=False,
          jitter=0.0, binning=None):
    """fcc lattice at `density`, Maxwell-Boltzmann velocities at `temperature`;
    binning=False: neighbour lists from all particle pairs (from_particle)."""
    n = int(nparticle)
    L, r0, v0, a = system_inputs(n, density, temperature, seed, jitter)
    nspecies = 1 if species is None else int(np.max(species)) + 1
    nspecies = max(nspecies, potential.species)
    box = mdsim.box(length=L)
    particle = mdsim.particle(dimension=3, particles=n, species=nspecies)
    particle.data["position"] = r0
    particle.data["velocity"] = v0
    sp = np.zeros(n, dtype=np.uint32) if species is None else np.asarray(species, dtype=np.uint32)
    if species is not None:
        particle.data["species"] = sp
    neighbour = mdsim.neighbour(particle=particle, box=box, r_cut=potential.r_cut, skin=skin, binning=binning,
                                disable_sorting=not hilbert_on_rebuild)
    force = mdsim.forces.pair_trunc(box=box, particle=particle, potential=potential, neighbour=neighbour)
    if integrator == "verlet":
        integ = mdsim.integrators.verlet(box=box, particle=particle, timestep=timestep)
    elif integrator == "verlet_nvt_hoover":
        integ = mdsim.integrators.verlet_nvt_hoover(box=box, particle=particle, timestep=timestep,
                                                    temperature=temperature,
                                                    resonance_frequency=resonance_frequency)
    elif integrator == "verlet_nvt_andersen":
        integ = mdsim.integrators.verlet_nvt_andersen(box=box, particle=particle, timestep=timestep,
                                                      temperature=temperature, rate=collision_rate, seed=seed)
    else:
        raise ValueError(integrator)
    core = mdsim.core(particle)
    core.set_integrator(integ)
    sort = None
    if sort_interval:
        sort = mdsim.sorts.hilbert(particle=particle, box=box)
        core.set_sort(sort, sort_interval)
    sampler = observables.sampler(core)
    thermo = observables.thermodynamics(box=box, particle=particle, integrator=integ)
    return SimpleNamespace(box=box, particle=particle, potential=potential, neighbour=neighbour, force=force,
                           integrator=integ, core=core, sort=sort, sampler=sampler, thermodynamics=thermo,
                           length=L, position=r0, velocity=v0, species=sp, temperature=temperature,
                           timestep=timestep, skin=skin, lattice_constant=a)


def lennard_jones(nparticle, density=0.8442, temperature=1.2, r_cut=2.5, truncation="shifted", **kw):
    pot = mdsim.potentials.pair.lennard_jones(epsilon=1.0, sigma=1.0).truncate(truncation, cutoff=r_cut)
    return fluid(nparticle, density, temperature, pot, **kw)


def kob_andersen(nparticle, density=1.2, temperature=0.7, fraction_b=0.2, seed=1, **kw):
    """80:20 Kob-Andersen mixture, species assigned by a seeded shuffle."""
    n = int(nparticle)
    species = ka_species(n, seed, fraction_b)
    pot = mdsim.potentials.pair.lennard_jones(epsilon=KA_EPSILON, sigma=KA_SIGMA).truncate("shifted", cutoff=2.5)
    return fluid(n, density, temperature, pot, species=species, seed=seed, **kw)


# ---------------------------------------------------------------------------
# batched ensembles of independent systems
# ---------------------------------------------------------------------------
def batched_specs(count=256, seed=2024, n_range=(500, 32000), density_range=(0.7, 1.2),
                  temperature_range=(0.5, 2.0), potentials=("lj", "wca", "morse")):
    """Seeded system mix: N log-uniform in n_range, density and temperature
    uniform, potential drawn from `potentials` (names of ensemble.POTENTIALS)."""
    rng = np.random.default_rng(seed)
    lo, hi = np.log(n_range[0]), np.log(n_range[1])
    out = []
    for k in range(count):
        out.append(dict(n=int(np.exp(rng.uniform(lo, hi))), density=float(rng.uniform(*density_range)),
                        temperature=float(rng.uniform(*temperature_range)),
                        potential=str(potentials[rng.integers(len(potentials))]), seed=int(seed * 1000 + k)))
    return out


def spec_potential(name, truncation="shifted", smoothing=0.005):
    """The single-system potential object of an ensemble potential name."""
    from .ensemble import POTENTIALS
    kind, eps, sigma, rc, r_min, dist = POTENTIALS[name]
    if kind == 0:
        return mdsim.potentials.pair.lennard_jones(epsilon=eps, sigma=sigma).truncate(truncation, cutoff=rc,
                                                                                      smoothing=smoothing)
    return mdsim.potentials.pair.morse(epsilon=eps, sigma=sigma, r_min=r_min, distortion=dist).truncate(
        truncation, cutoff=rc, smoothing=smoothing)


def ka_species(n, seed, fraction_b=0.2):
    """Species of the Kob-Andersen mixture: a seeded shuffle, as kob_andersen() assigns them."""
    rng = np.random.default_rng(seed + 7)
    species = np.zeros(n, dtype=np.uint32)
    species[rng.permutation(n)[:int(round(fraction_b * n))]] = 1
    return species


def spec_species(spec):
    return ka_species(spec["n"], spec["seed"]) if spec["potential"] == "ka" else None


def spec_skin(skin, spec):
    """The skin of `spec`'s system: `skin` itself, or skin[potential name] for a
    mapping of potential names to skins."""
    return skin[spec["potential"]] if isinstance(skin, dict) else skin


def single_from_spec(spec, timestep=0.003, skin=0.3):
    """The system of `spec` alone, through the per-system modules (skin: a
    number, or a mapping of potential names to skins)."""
    pot = spec_potential(spec["potential"], spec.get("truncation", "shifted"), spec.get("smoothing", 0.005))
    return fluid(spec["n"], spec["density"], spec["temperature"], pot, skin=spec_skin(skin, spec),
                 timestep=timestep, seed=spec["seed"], species=spec_species(spec))


def ensemble_from_specs(specs, timestep=0.003, skin=0.3, segments=None):
    """All systems of `specs` in one committed ensemble (inputs as fluid() makes
    them); skin: a number, or a mapping of potential names to skins;
    segments: (target particles per segment, workgroup size)."""
    from .ensemble import ensemble
    e = ensemble()
    if segments is not None:
        e.set_segments(*segments)
    for sp in specs:
        L, r0, v0, _ = system_inputs(sp["n"], sp["density"], sp["temperature"], sp["seed"])
        e.add(r0, v0, L[0], potential=sp["potential"], truncation=sp.get("truncation", "shifted"),
              smoothing=sp.get("smoothing", 0.005), timestep=timestep, skin=spec_skin(skin, sp),
              species=spec_species(sp))
    e.commit()
    return e


def tune_ensemble(specs, timestep=0.003, targets=(150, 200, 260), threads=(128, 256), skins=(0.25, 0.3, 0.35),
                  class_skins=(0.2, 0.3, 0.5, 0.7), warmup=32, steps=128, margin=0.02):
    """Autotuner of the batched path: cell size (through the skin: cells are
    rc + skin wide), segment shape (particles per segment) and workgroup size,
    each candidate timed over `steps` device-decided steps of the systems of
    `specs` from their start.  After the grid over (skin, target, threads),
    each potential's systems in turn try the skins of `class_skins` with the
    others fixed (a potential with a short cutoff rebuilds less often with a
    wider skin at little pair cost); a change is kept when the whole ensemble
    is faster by `margin`.  Returns (best, table): best = dict(skin (a mapping
    of potential names to skins when class_skins is not empty), segments =
    (target, threads), ms_per_step); table = every candidate's ms per step
    (None where a staging does not fit)."""
    from ._lib import HalmdError

    def trial(skin, segments):
        try:
            e = ensemble_from_specs(specs, timestep=timestep, skin=skin, segments=segments)
            e.step(warmup)
            return e.step_timed(steps) / steps
        except HalmdError:
            return None   # e.g. a cell neighbourhood beyond the staging limit

    table = []
    for skin in skins:
        for tg in targets:
            for th in threads:
                table.append(dict(skin=skin, segments=(tg, th), ms_per_step=trial(skin, (tg, th))))
    best = dict(min((t for t in table if t["ms_per_step"] is not None), key=lambda t: t["ms_per_step"]))
    if class_skins:
        names = sorted({sp["potential"] for sp in specs})
        best["skin"] = {name: best["skin"] for name in names}
        for name in names:
            for alt in class_skins:
                if alt == best["skin"][name]:
                    continue
                skin = dict(best["skin"], **{name: alt})
                ms = trial(skin, best["segments"])
                table.append(dict(skin=skin, segments=best["segments"], ms_per_step=ms))
                if ms is not None and ms < best["ms_per_step"] * (1 - margin):
                    best = dict(skin=skin, segments=best["segments"], ms_per_step=ms)
    return best, table


# ---------------------------------------------------------------------------
# served job streams
# ---------------------------------------------------------------------------
SERVING_WAVENUMBERS = (1.5, 2.5, 3.5, 4.5, 5.5, 6.5, 7.25, 8.5)


def serving_stream(rate=50.0, duration=60.0, seed=2025, step_scale=1.0, n_median=2.0e4, n_sigma=0.75,
                   n_clip=(1000, 200000), density_range=(0.8, 1.2), temperature_range=(0.45, 1.5),
                   steps_range=(1.0e3, 1.0e5), potentials=("lj", "ka", "wca")):
    """Poisson arrivals of MD jobs: arrival times (s) at `rate` over `duration`,
    N log-normal (median n_median, log-sigma n_sigma, clipped), density and
    temperature uniform, steps log-uniform in steps_range times step_scale,
    potential drawn from `potentials`."""
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    lo, hi = np.log(steps_range[0]), np.log(steps_range[1])
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration:
            break
        n = int(np.clip(np.exp(rng.normal(np.log(n_median), n_sigma)), *n_clip))
        steps = max(1, int(round(np.exp(rng.uniform(lo, hi)) * step_scale)))
        out.append(dict(arrival=t, n=n, density=float(rng.uniform(*density_range)),
                        temperature=float(rng.uniform(*temperature_range)),
                        potential=str(potentials[rng.integers(len(potentials))]), steps=steps,
                        seed=int(seed * 100000 + len(out))))
    return out
