"""Benchmark: Lennard-Jones fluid, N = 1,048,576, on one MI355X.

Workload (one "step" = one velocity-Verlet MD step of the whole system):
fcc 64^3 lattice at rho = 0.8442, Maxwell-Boltzmann velocities at T = 0.722
(seeded), LJ truncated and shifted at r_c = 2.5 sigma, neighbour skin 0.3 sigma
with rebuilds on the half-skin criterion, NVE dt = 0.005 tau, Hilbert-curve
sort every 100 steps, density modes every 1000 steps for 201 wavevectors with
|k| in [2 pi / L, 10].  All state is resident on the device; the timed region
is bracketed by device synchronisations.

Prints one JSON line with the throughput (particle-steps/s), the HBM roofline
of the dominant kernel (the pair-force kernel) measured with HIP events on its
stream, and the FP64 CPU oracle timed on a bounded sample on this host.  The
same line carries, as extra keys, the metric's second half and context:
"batched" (256 seeded systems stepped together for 5000 steps against the
plain per-job path), "serving" (a Poisson job stream through the
continuous-batching server against the plain per-job path run first-in
first-out on the same arrivals, both measured), and "reference_benchmarks"
(the reference's published LJ N=64,000 and Kob-Andersen N=256,000 setups).
"""
import argparse
import gc
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N = 64 ** 3 * 4
DENSITY = 0.8442
TEMPERATURE = 0.722
R_CUT = 2.5
SKIN = 0.3
TIMESTEP = 0.005
SORT_INTERVAL = 100
MODES_INTERVAL = 1000
HBM_PEAK_GBS = 8000.0
# hardware-counter summary of the pair-force kernel on this workload (rocprofv3
# --pmc FETCH_SIZE / WRITE_SIZE, calibrated against a copy kernel); committed
# beside the kernel-time summary it was measured with
TRAFFIC_PROFILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                               "pair_force_lj_n1m_counters.md")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def build_workload(seed):
    from halmd_amd import observables, workloads
    from halmd_amd.observables.utility import wavevector
    w = workloads.lennard_jones(N, density=DENSITY, temperature=TEMPERATURE, r_cut=R_CUT, skin=SKIN,
                                timestep=TIMESTEP, seed=seed, sort_interval=SORT_INTERVAL)
    k = np.linspace(2 * np.pi / w.length[0], 10.0, 34)
    wv = wavevector(w.box, k, tolerance=0.01, max_count=6)
    dm = observables.density_mode(w.particle, wv)
    w.sampler.on_sample_density_mode(dm, MODES_INTERVAL)
    w.wavevector, w.density_mode = wv, dm
    return w


def kernel_times(w, steps):
    """Average duration per launch (ms) of each kernel class over `steps` steps.

    The pass waits for each step's neighbour check before the force launch, so
    that every timed force launch does the full work (in the timed region a
    force launch may start before the check is back and return at once when the
    lists turn out stale; the module then rebuilds and launches again)."""
    from halmd_amd import _lib
    mods = {"force": w.force, "neighbour": w.neighbour, "binning": w.neighbour.binning,
            "integrate": w.integrator, "finalize": w.integrator, "sort": w.sort}
    for name, m in mods.items():
        _lib.call("halmd_profile_enable", m.handle, name.encode(), 4 * steps + 16)
    w.particle.set_speculative_force(False)
    w.sampler.run(steps)
    w.particle.set_speculative_force(True)
    out = {}
    for name, m in mods.items():
        import ctypes
        ms, n = ctypes.c_double(), ctypes.c_uint64()
        _lib.call("halmd_profile_collect", m.handle, name.encode(), ctypes.byref(ms), ctypes.byref(n))
        _lib.call("halmd_profile_enable", m.handle, name.encode(), 0)
        out[name] = (ms.value, n.value)
    return out


def calibration_valu(blocks=2048, iterations=8192):
    """The VALU rate probe: a known number of FP32 FMA wave instructions with
    no memory traffic; its hardware-counter reading and duration calibrate
    the VALU issue fraction of the other kernels of the same profile."""
    import ctypes
    from halmd_amd import _lib
    n, ms = ctypes.c_uint64(), ctypes.c_double()
    _lib.call("halmd_device_valu_probe", blocks, iterations, ctypes.byref(n), ctypes.byref(ms))
    log(f"calibration VALU probe: {n.value} FMA wave instructions in {ms.value:.3f} ms "
        f"({n.value / (ms.value * 1e-3):.3e} per second)")


def calibration_copy(nbytes):
    import ctypes
    from halmd_amd import _lib
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    _lib.call("halmd_device_alloc", nbytes, ctypes.byref(a))
    _lib.call("halmd_device_alloc", nbytes, ctypes.byref(b))
    _lib.call("halmd_device_copy", b, a, nbytes)
    _lib.call("halmd_device_free", a)
    _lib.call("halmd_device_free", b)
    log(f"calibration copy: {nbytes} bytes read + {nbytes} bytes written")


def counter_summary(kernel):
    """Figures recorded for `kernel` in the committed counter summary:
    lines 'key: value' inside the section headed by the kernel's name."""
    try:
        text = open(TRAFFIC_PROFILE).read()
    except OSError:
        return {}
    out, inside = {}, False
    for line in text.splitlines():
        if line.startswith("## "):
            inside = line[3:].strip().strip("`") == kernel
        elif inside and ":" in line and line.startswith("- "):
            k, v = line[2:].split(":", 1)
            try:
                out[k.strip()] = float(v.split()[0])
            except ValueError:
                out[k.strip()] = v.strip()
            except IndexError:
                pass
    return out


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(w, steps):
    """FP64 oracle (host backend restatement), one thread, same initial state,
    `steps` NVE steps after the setup (the reference host backend is
    single-threaded: doc/benchmarks.rst.in:17-27)."""
    from oracle import oracle as orc
    pot = orc.Potential(orc.POT_LJ, "shifted", 1.0, 1.0, R_CUT)
    sim = orc.Simulation(w.position, w.velocity, w.species, np.ones(N), w.length, pot, SKIN, TIMESTEP)
    sim.time(0)                     # initial binning, lists and forces
    builds0 = sim.rebuilds()
    secs = sim.time(steps)
    return {"value": N * steps / secs, "unit": "particle-steps/s", "cores": 1, "kind": "port",
            "sample": f"FP64 oracle (host-backend restatement, single thread) on {cpu_model()}, same LJ N={N} "
                      f"state, {steps} NVE steps after setup ({sim.rebuilds() - builds0} neighbour-list builds), "
                      f"{secs:.1f} s",
            "cpu_model": cpu_model()}


def skin_text(skin):
    """'skin 0.3', or 'skins lj 0.3 / morse 0.3 / wca 0.5' for a skin per potential."""
    if isinstance(skin, dict):
        return "skins " + " / ".join(f"{k} {v}" for k, v in sorted(skin.items()))
    return f"skin {skin}"


def batched_run(systems, steps, warmup, seed, plain_steps=1000, pinned=None):
    """Batched ensembles: `systems` seeded systems (N log-uniform in [500,
    32000], rho in [0.7, 1.2], T in [0.5, 2.0], LJ / WCA / Morse), NVE dt =
    0.003, stepped together for `steps` steps; against the plain per-job path
    (each system alone through the per-system modules, one after another) on
    every 16th system.  Returns the figures as a dict."""
    from halmd_amd import workloads
    specs = workloads.batched_specs(systems, seed=seed)
    ntot = sum(s["n"] for s in specs)
    # the autotuner picks cell size (skin), segment shape and workgroup size
    # by timing the same systems from their start
    t0 = time.perf_counter()
    if pinned:   # (skin, particles per segment, workgroup size), e.g. for profiling
        skin, segments, table = pinned[0], (pinned[1], int(pinned[2])), []
        log(f"ensemble shape pinned: skin {skin}, {segments[0]} particles per segment, {segments[1]} threads")
    else:
        best, table = workloads.tune_ensemble(specs, timestep=0.003)
        skin, segments = best["skin"], best["segments"]
        log(f"ensemble autotuner ({len(table)} candidates, {time.perf_counter() - t0:.1f} s): "
            f"{skin_text(skin)}, {segments[0]} particles per segment, {segments[1]} threads per workgroup "
            f"({best['ms_per_step']:.3f} ms/step over 128 steps)")
    t0 = time.perf_counter()
    e = workloads.ensemble_from_specs(specs, timestep=0.003, skin=skin, segments=segments)
    log(f"ensemble of {len(specs)} systems, {ntot} particles, setup {time.perf_counter() - t0:.1f} s")
    e.step(warmup)
    pot0, kin0 = e.energies()
    ms = e.step_timed(steps)
    pot1, kin1 = e.energies()
    value = ntot * steps / (ms / 1e3)
    log(f"batched: {ms / steps:.3f} ms/step, {value:.4e} particle-steps/s")
    sample = specs[::16] if plain_steps else []
    psum, tsum = 0.0, 1e-300
    for sp in sample:
        w = workloads.single_from_spec(sp, timestep=0.003, skin=skin)
        w.sampler.run(5)
        t = w.sampler.run_timed(plain_steps)
        psum += sp["n"] * plain_steps
        tsum += t / 1e3
    plain = psum / tsum if sample else None
    if sample:
        log(f"plain per-job path on {len(sample)} systems: {plain:.4e} particle-steps/s ({value / plain:.1f}x)")
    drift = np.abs((pot1 + kin1) - (pot0 + kin0)) / (steps * 0.003)
    return {
        "workload": f"{len(specs)} independent NVE systems, N log-uniform in [500, 32000] (total {ntot}), "
                    f"rho U[0.7, 1.2], T U[0.5, 2.0], LJ rc 2.5 / WCA / Morse rc 2.5, dt 0.003, {skin_text(skin)} "
                    + ("(pinned)" if pinned else "(autotuned)"),
        "autotuned": {"skin": skin, "segment_particles": segments[0], "workgroup": segments[1],
                      "candidates": [[t["skin"], t["segments"][0], t["segments"][1],
                                      None if t["ms_per_step"] is None else round(t["ms_per_step"], 4)]
                                     for t in table]},
        "systems": len(specs),
        "particles": ntot,
        "steps": steps,
        "ms_per_step": ms / steps,
        "particle_steps_per_s": value,
        "plain_per_job_particle_steps_per_s": plain,
        "plain_sample": f"every 16th system ({len(sample)}), {plain_steps} timed steps each, one after another",
        "speedup_vs_plain": value / plain if plain else None,
        "energy_drift_per_particle_per_tau_max": float(drift.max()),
    }


def batched(args):
    """--workload batched: the batched figures as the JSON line."""
    pinned = [float(x) for x in args.pin_ensemble.split(",")] if args.pin_ensemble else None
    b = batched_run(args.systems, args.steps, args.warmup, args.seed, plain_steps=args.plain_steps, pinned=pinned)
    out = {
        "metric": "aggregate particle-steps/s, batched ensemble of 256 systems vs plain per-job GPU path",
        "value": b["particle_steps_per_s"],
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": b["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": b,
    }
    print(json.dumps(out), flush=True)


TINY_H100 = 4.23e7   # particle-steps/s, KA "tiny" N=4096 on H100, unrolled force loop (23.6 ns)


def tiny(args):
    """The reference's "tiny" variant (doc/benchmarks.rst.in:109-127): Kob-Andersen
    N=4096, rho=1.2, r_c=2.5 sigma shifted, skin 0.3, NVE dt=0.001, neighbour
    lists built from all particle pairs (no binning), no sorting."""
    from halmd_amd import workloads
    n = 4096
    w = workloads.kob_andersen(n, density=1.2, temperature=0.7, timestep=0.001, seed=args.seed, binning=False)
    assert w.neighbour.algorithm == "from_particle"
    w.sampler.run(args.warmup)
    ms = w.sampler.run_timed(args.steps)
    value = n * args.steps / (ms / 1e3)
    log(f"tiny: {ms / args.steps * 1e3:.2f} us/step, {ms / args.steps * 1e6 / n:.2f} ns per step and particle, "
        f"{value:.4e} particle-steps/s; {w.neighbour.rebuilds} list builds")
    out = {
        "metric": "particle-steps/s, Kob-Andersen N=4096 (tiny)",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": value / TINY_H100,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "Kob-Andersen 80:20 N=4096, rho=1.2, T=0.7, r_c=2.5 sigma shifted, skin 0.3, NVE "
                        "dt=0.001, neighbour lists from all pairs (from_particle), no sorting",
            "nparticle": n,
            "ns_per_step_and_particle": ms / args.steps * 1e6 / n,
            "neighbour_rebuilds": int(w.neighbour.rebuilds),
            "baseline": "H100, unrolled force loop, 23.6 ns per step and particle (4.23e7 particle-steps/s)",
        },
    }
    print(json.dumps(out), flush=True)


# the skin of each served job by potential, the same on the plain path: the
# ensemble autotuner's per-potential choice on the batched mix (WCA 0.7, LJ
# 0.35); Kob-Andersen keeps 0.3
SERVING_SKINS = {"lj": 0.35, "ka": 0.3, "wca": 0.7}


def parse_skins(text):
    """'lj:0.35,ka:0.3,wca:0.7' -> the mapping (None: the defaults)."""
    if not text:
        return None
    out = dict(SERVING_SKINS)
    for part in text.split(","):
        name, value = part.split(":")
        out[name.strip()] = float(value)
    return out


def serving_run(rate, duration, seed, step_scale, tick=50, max_jobs=256, budget_gib=128, skins=None):
    """Served job stream: Poisson arrivals of LJ / Kob-Andersen / WCA jobs (N
    log-normal, median 2e4, clipped to [1e3, 2e5]; rho in [0.8, 1.2]; T in
    [0.45, 1.5]; steps log-uniform in [1e3, 1e5] times step_scale), each with
    blocking-scheme MSD and ISF at 8 wavenumbers.  First through one server in
    real time, then through the plain per-job path: the same jobs with the same
    correlators, one at a time in arrival order, each built, run and read out
    through the per-system modules before the next starts, arrivals replayed in
    real time.  Both latencies are measured.  Returns the figures as a dict."""
    from collections import deque
    from halmd_amd import observables, workloads
    from halmd_amd.observables.dynamics import blocking_scheme
    from halmd_amd.observables.utility import wavevector
    from halmd_amd.serve import server
    stream = workloads.serving_stream(rate=rate, duration=duration, seed=seed, step_scale=step_scale)
    skins = dict(SERVING_SKINS if skins is None else skins)   # per potential, the same on both paths
    dt = 0.003
    k = list(workloads.SERVING_WAVENUMBERS)
    corr = dict(block_size=10, shift=2, separation=10, wavenumber=k, tolerance=0.01, max_count=7)
    t0 = time.perf_counter()
    inputs = [workloads.system_inputs(j["n"], j["density"], j["temperature"], j["seed"]) for j in stream]
    log(f"{len(stream)} jobs over {duration:.0f} s, inputs ready in {time.perf_counter() - t0:.1f} s")
    s = server(budget_bytes=budget_gib << 30, max_jobs=max_jobs, tick_steps=tick)
    work = sum(j["n"] * (-(-j["steps"] // tick) * tick) for j in stream)
    offered = work / duration
    pending = deque(range(len(stream)))
    ids = {}
    start = s.now()
    ticks = 0
    while pending or s.stats()["active"] or s.stats()["queued"]:
        now = s.now() - start
        while pending and stream[pending[0]]["arrival"] <= now:
            q = pending.popleft()
            j = stream[q]
            L, r0, v0, _ = inputs[q]
            ids[q] = s.submit(r0, v0, L[0], j["steps"], potential=j["potential"], timestep=dt,
                              skin=skins[j["potential"]], species=workloads.spec_species(j),
                              arrival=start + j["arrival"], **corr)
        st = s.stats()
        if not st["active"] and not st["queued"]:
            time.sleep(max(0.0, stream[pending[0]]["arrival"] - (s.now() - start)))
            continue
        s.tick()
        ticks += 1
        if ticks % 100 == 0:
            log(f"served t = {s.now() - start:6.1f} s: {st['active']} running, {st['queued']} queued, "
                f"{len(ids)} submitted of {len(stream)}")
    end = s.now()
    lat = np.array([s.info(ids[q])["done"] - (start + stream[q]["arrival"]) for q in range(len(stream))])
    wait = np.array([s.info(ids[q])["admitted"] - (start + stream[q]["arrival"]) for q in range(len(stream))])
    psteps = s.stats()["particle_steps"]
    phases = s.phases()
    for q in range(len(stream)):
        s.release(ids[q])
    served = psteps / (end - start)
    log(f"served: {served:.4e} particle-steps/s over {end - start:.1f} s; latency p50 {np.percentile(lat, 50):.2f} s, "
        f"p99 {np.percentile(lat, 99):.2f} s; offered {offered:.3e} particle-steps/s")
    del s

    # plain per-job path, first in first out on the same arrivals
    plat = np.zeros(len(stream))
    pstart = time.perf_counter()
    for q, j in enumerate(stream):
        now = time.perf_counter() - pstart
        if now < j["arrival"]:
            time.sleep(j["arrival"] - now)
        steps = -(-j["steps"] // tick) * tick   # whole ticks, as the server runs them
        w = workloads.single_from_spec(j, timestep=dt, skin=skins)
        bs = blocking_scheme(w.box, w.particle, timestep=dt, max_lag=steps * dt, resolution=tick * dt,
                             block_size=10, shift=2, separation=10)
        bs.correlate_msd()
        bs.correlate_isf(observables.density_mode(w.particle, wavevector(w.box, k, tolerance=0.01, max_count=7)))
        w.sampler.on_sample_blocking_scheme(bs)
        w.sampler.run(steps)
        bs.finalise()
        bs.result("msd")
        bs.result("isf")
        plat[q] = time.perf_counter() - pstart - j["arrival"]
        if q % 100 == 99:
            log(f"plain path: {q + 1} of {len(stream)} jobs done at {time.perf_counter() - pstart:.1f} s")
    pend = time.perf_counter() - pstart
    plain = work / pend
    log(f"plain per-job path (first in first out, measured): {plain:.4e} particle-steps/s over {pend:.1f} s; "
        f"latency p50 {np.percentile(plat, 50):.2f} s, p99 {np.percentile(plat, 99):.2f} s")
    return {
        "workload": f"Poisson {rate:g} jobs/s for {duration:g} s ({len(stream)} jobs): LJ / KA / WCA, "
                    "N log-normal (median 2e4, log-sigma 0.75, clipped [1e3, 2e5]), rho U[0.8, 1.2], "
                    f"T U[0.45, 1.5], steps log-uniform [1e3, 1e5] x {step_scale:g} (whole ticks of {tick}), "
                    f"dt {dt}, {skin_text(skins)}, MSD + ISF at 8 wavenumbers (blocking scheme, block 10)",
        "jobs": len(stream),
        "offered_particle_steps_per_s": offered,
        "particle_steps_per_s": served,
        "latency_p50_s": float(np.percentile(lat, 50)),
        "latency_p99_s": float(np.percentile(lat, 99)),
        "queue_wait_p50_s": float(np.percentile(wait, 50)),
        "queue_wait_p99_s": float(np.percentile(wait, 99)),
        "makespan_s": end - start,
        "ticks": ticks,
        "tick_steps": tick,
        "tick_phase_seconds": phases,
        "plain_per_job_particle_steps_per_s": plain,
        "plain_latency_p50_s": float(np.percentile(plat, 50)),
        "plain_latency_p99_s": float(np.percentile(plat, 99)),
        "plain_makespan_s": pend,
        "plain_method": "the same jobs and correlators one at a time in arrival order through the per-system "
                        "modules (build, run, read out), arrivals replayed in real time; measured",
        "speedup_vs_plain": served / plain,
    }


def serving(args):
    """--workload serving: the serving figures as the JSON line."""
    r = serving_run(args.rate, args.duration, args.seed, args.step_scale, args.tick, args.max_jobs,
                    args.budget_gib, parse_skins(args.serving_skins))
    out = {
        "metric": "served job stream: aggregate particle-steps/s and p50/p99 job latency vs plain per-job GPU path",
        "value": r["particle_steps_per_s"],
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": r["ticks"],
        "warmup": 0,
        "ms_per_step": r["makespan_s"] * 1e3 / max(r["ticks"], 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": r,
    }
    print(json.dumps(out), flush=True)


# the reference's published setups (doc/benchmarks.rst.in): H100 double-single
REFERENCE_SETUPS = {
    "lj64k": dict(n=64000, density=0.4, r_cut=3.0, skin=0.7, timestep=0.002, h100_ns=3.4,
                  source="doc/benchmarks.rst.in:17-29"),
    "ka256k": dict(n=256000, density=1.2, r_cut=2.5, skin=0.3, timestep=0.001, h100_ns=1.79, occupancy=0.7,
                   source="doc/benchmarks.rst.in:62-79"),
}


def reference_setup_run(name, steps, warmup, seed):
    """One of the reference's published benchmark setups, NVE, from an fcc
    start at T = 1.2 (LJ) or 0.7 (Kob-Andersen): ns per step and particle
    beside the reference's H100 figure (context; the setups differ from the
    workloads of the metric)."""
    from halmd_amd import workloads
    c = REFERENCE_SETUPS[name]
    if name == "lj64k":
        w = workloads.lennard_jones(c["n"], density=c["density"], temperature=1.2, r_cut=c["r_cut"], skin=c["skin"],
                                    timestep=c["timestep"], seed=seed)
    else:
        w = workloads.kob_andersen(c["n"], density=c["density"], temperature=0.7, skin=c["skin"],
                                   timestep=c["timestep"], seed=seed)
    w.core.tune()
    w.sampler.run(warmup)
    ms = w.sampler.run_timed(steps)
    ns = ms / steps * 1e6 / c["n"]
    log(f"{name}: {ns:.3f} ns per step and particle (H100 double-single: {c['h100_ns']} ns)")
    return {"particles": c["n"], "steps": steps, "ns_per_step_and_particle": ns,
            "particle_steps_per_s": c["n"] * steps / (ms / 1e3), "h100_ns_per_step_and_particle": c["h100_ns"],
            "ratio_to_h100": c["h100_ns"] / ns, "source": c["source"],
            "setup": {k: v for k, v in c.items() if k not in ("h100_ns", "source")}}


def reference_setup(args):
    r = reference_setup_run(args.workload, args.steps, args.warmup, args.seed)
    out = {
        "metric": f"particle-steps/s, reference benchmark setup {args.workload}",
        "value": r["particle_steps_per_s"],
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ns_per_step_and_particle"] * r["particles"] * 1e-6,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": r["ratio_to_h100"],
        "dtype": "f32",
        "data": "synthetic",
        "config": r,
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--warmup", type=int, default=400)
    ap.add_argument("--profile-steps", type=int, default=500)
    ap.add_argument("--cpu-steps", type=int, default=200,
                    help="NVE steps of the FP64 CPU baseline (one thread) after its setup")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="LJ workload only: leave out the batched, serving and reference-setup figures")
    ap.add_argument("--batched-steps", type=int, default=5000, help="steps of the batched ensemble figure")
    ap.add_argument("--no-tune", action="store_true", help="keep the default cell-segment shape")
    ap.add_argument("--segments", default=None,
                    help="cell-segment shape 'cells,rows,threads' instead of autotuning (profiling runs)")
    ap.add_argument("--workload", default="lj", choices=["lj", "batched", "serving", "tiny", "lj64k", "ka256k"],
                    help="lj: LJ fluid N = 2^20 (default); batched: ensemble of independent systems; "
                         "serving: Poisson job stream through the continuous-batching server; "
                         "tiny: Kob-Andersen N = 4096 with all-pairs neighbour lists; lj64k, ka256k: the "
                         "reference's published benchmark setups")
    ap.add_argument("--rate", type=float, default=50.0, help="serving: job arrivals per second")
    ap.add_argument("--duration", type=float, default=4.0, help="serving: seconds of arrivals")
    ap.add_argument("--step-scale", type=float, default=0.2,
                    help="serving: factor on the job step counts (0.2 keeps the offered load, about 5e9 "
                         "particle-steps/s, above what one GPU serves, as the full-length stream's is)")
    ap.add_argument("--tick", type=int, default=50, help="serving: steps per tick")
    ap.add_argument("--serving-skins", default=None,
                    help="serving: skin per potential, e.g. 'lj:0.35,ka:0.3,wca:0.7' (both paths)")
    ap.add_argument("--max-jobs", type=int, default=256, help="serving: running-job cap")
    ap.add_argument("--budget-gib", type=int, default=128, help="serving: device memory budget for jobs")
    ap.add_argument("--systems", type=int, default=256, help="systems of the batched workload")
    ap.add_argument("--pin-ensemble", default=None,
                    help="batched: SKIN,PARTICLES_PER_SEGMENT,THREADS instead of the autotuner's choice")
    ap.add_argument("--plain-steps", type=int, default=1000,
                    help="batched: steps of each sampled plain per-job run (0: none)")
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--calibration-copy-mib", type=int, default=0,
                    help="run one copy kernel of this size first (hardware-counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs one GPU per process; --gpus {args.gpus} is not supported "
              "(independent systems shard across GPUs without exchange)", file=sys.stderr)
        sys.exit(2)

    import halmd_amd
    if args.workload == "batched":
        return batched(args)
    if args.workload == "serving":
        return serving(args)
    if args.workload == "tiny":
        return tiny(args)
    if args.workload in REFERENCE_SETUPS:
        return reference_setup(args)
    arch, cus, mem = halmd_amd.device_info()
    if args.calibration_copy_mib:
        calibration_copy(args.calibration_copy_mib << 20)
        calibration_valu()
    log(f"device {arch}, {cus} CUs, {mem / 2**30:.0f} GiB; building LJ N={N}")
    t0 = time.perf_counter()
    w = build_workload(args.seed)
    if args.segments:
        w.neighbour.set_segments(*[int(x) for x in args.segments.split(",")])
        log(f"cell segments set to {w.neighbour.segments} (cells, rows, workgroup size)")
    elif not args.no_tune:
        w.core.tune()
        log(f"autotuned cell segments: {w.neighbour.segments} (cells, rows, workgroup size)")
    log(f"setup {time.perf_counter() - t0:.1f} s; warmup {args.warmup} steps")
    w.sampler.run(args.warmup)
    log(f"timed {args.steps} steps")
    ms = w.sampler.run_timed(args.steps)
    value = N * args.steps / (ms / 1e3)
    log(f"{ms / args.steps:.4f} ms/step, {value:.4e} particle-steps/s; per-kernel timing pass")

    kt = kernel_times(w, args.profile_steps)
    counts = w.neighbour.counts().astype(np.float64)
    n_nb = counts.mean()
    f_ms, f_n = kt["force"]
    f_avg = f_ms / max(f_n, 1)
    # algorithmic bytes per particle for the force kernel (fixed definition of
    # unique compulsory traffic with float-float state)
    staged = w.neighbour.entry_bytes == 2
    # the Verlet update ran inside the force kernel (all but the steps after a
    # sample, whose force was already evaluated)
    fused = kt["integrate"][1] * 10 < kt["force"][1]
    if fused:
        # force kernel that integrates each particle as it finishes it: r_i
        # (hi+lo, 24 B) and (n_nb + 1) list words of 4 B for the force, r and v
        # read and written (96 B), r0 for the displacement (12 B); the force
        # itself never goes to memory
        force_bytes = N * (24 + 4 * (n_nb + 1) + 96 + 12)
    else:
        # (unique compulsory traffic with float-float state): r_i (hi+lo, 24 B),
        # (n_nb + 1) list words of 4 B (entries + count), force (12 B)
        force_bytes = N * (24 + 4 * (n_nb + 1) + 12)
    achieved = force_bytes / (f_avg * 1e-3) / 1e9
    kname = ("pair_force_staged_kernel<0, 1, false, true, false, %s>" % ("true" if fused else "false") if staged
             else "pair_force_kernel<0, 1, false>")
    prof = counter_summary(kname)
    traffic = prof.get("traffic_bytes_per_launch")
    kernels = {k: {"launches": int(n), "avg_ms": (ms_ / n if n else None)} for k, (ms_, n) in kt.items()}
    log(f"force kernel {f_avg:.4f} ms/launch, {achieved:.0f} GB/s algorithmic; n_nb = {n_nb:.2f}")
    nb_ms, nb_n = kt["neighbour"]
    nb_avg = nb_ms / nb_n if nb_n else None
    nb_bytes = N * (24 + 13 + 4 * (n_nb + 1))
    vf = prof.get("valu_issue_frac")
    limiter = ("valu" if vf is not None and vf > achieved / HBM_PEAK_GBS else "hbm")
    cell_segments = list(w.neighbour.segments)
    rebuilds = int(w.neighbour.rebuilds)
    nwave = int(len(w.wavevector.value))
    thermo = w.thermodynamics
    en_tot = thermo.en_tot()
    # the Hilbert sort (every 100 steps) and the density modes (every 1000
    # steps) amortised per step, when the timed window holds fewer steps
    import halmd_amd._lib as _hl
    _hl.call("halmd_device_synchronize")
    t0 = time.perf_counter()
    for _ in range(3):
        w.density_mode.acquire()
    modes_ms = (time.perf_counter() - t0) / 3 * 1e3
    sort_ms = kt["sort"][0] / max(kt["sort"][1], 1)
    amortised = ms / args.steps
    if args.steps < SORT_INTERVAL:
        amortised += sort_ms / SORT_INTERVAL
    if args.steps < MODES_INTERVAL:
        amortised += modes_ms / MODES_INTERVAL

    cpu = None
    if not args.no_cpu_baseline:
        log(f"CPU baseline (FP64 oracle, 1 thread, {args.cpu_steps} steps)")
        cpu = cpu_baseline(w, args.cpu_steps)
        cpu["host_cpus"] = os.cpu_count()

    extras = {}
    if not args.no_extras:
        # the LJ modules go first (their device memory and streams)
        del w, thermo, kt
        gc.collect()
        log("batched ensemble figure")
        extras["batched"] = batched_run(args.systems, args.batched_steps, 100, args.seed)
        log("served stream figure")
        extras["serving"] = serving_run(args.rate, args.duration, args.seed, args.step_scale, args.tick,
                                        args.max_jobs, args.budget_gib, parse_skins(args.serving_skins))
        log("reference benchmark setups")
        extras["reference_benchmarks"] = {name: reference_setup_run(name, 2000, 200, args.seed)
                                          for name in REFERENCE_SETUPS}

    out = {
        "metric": "particle-steps/s, LJ N=1M",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "Lennard-Jones fluid N=1048576 (fcc 64^3, rho=0.8442, T=0.722), r_c=2.5 shifted, "
                        "skin=0.3, NVE dt=0.005, Hilbert sort every 100 steps, density modes every 1000 steps",
            "nparticle": N,
            "precision": "FP32 pair forces, float-float positions/velocities, FP64 reductions",
            "wavevectors": nwave,
            "neighbours_mean": n_nb,
            "neighbour_rebuilds": rebuilds,
            "cell_segments": cell_segments,
            "en_tot_per_particle_end": en_tot,
            "device": arch,
            "ms_per_step_amortised": amortised,
            "amortised": f"Hilbert sort {sort_ms:.3f} ms / {SORT_INTERVAL} steps, density modes {modes_ms:.3f} ms "
                         f"/ {MODES_INTERVAL} steps, added where the timed window holds fewer steps",
        },
        "roofline": {
            # the roofline is HBM (no MFMA on this path); what limits the kernel
            # is read from its counters: VALU issue against the measured rate
            # of the VALU probe, or HBM
            "bound": "hbm",
            "limiter": limiter,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "kernel": kname,
            "bytes_per_launch": force_bytes,
            "avg_ms": f_avg,
            # from the counter summary: VALU wave instructions per second of
            # the kernel over those of the VALU probe in the same profile
            "valu_issue_frac": prof.get("valu_issue_frac"),
            "counters": os.path.relpath(TRAFFIC_PROFILE, ROOT) if prof else None,
            "counter_shape": prof.get("segments"),
        },
        "neighbour": {
            # the list build against its fixed algorithmic bytes: r 24 B, cell
            # slots ~13 B, list words (n_nb + 1) 4 B per particle
            "avg_ms": nb_avg,
            "bytes_per_launch": nb_bytes,
            "achieved": nb_bytes / (nb_avg * 1e-3) / 1e9 if nb_avg else None,
            "frac": nb_bytes / (nb_avg * 1e-3) / 1e9 / HBM_PEAK_GBS if nb_avg else None,
            "unit": "GB/s",
        },
        "kernels": kernels,
        "cpu_baseline": cpu,
        **extras,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
