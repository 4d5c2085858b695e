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
stream, and the FP64 CPU oracle timed on a bounded sample on this host.
"""
import argparse
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
            except (ValueError, IndexError):
                pass
    return out


def cpu_baseline(w, budget_s):
    """FP64 oracle (host backend restatement), one thread, same initial state."""
    from oracle import oracle as orc
    pot = orc.Potential(orc.POT_LJ, "shifted", 1.0, 1.0, R_CUT)
    t0 = time.perf_counter()
    sim = orc.Simulation(w.position, w.velocity, w.species, np.ones(N), w.length, pot, SKIN, TIMESTEP)
    sim.time(0)                     # initial binning, lists and forces
    setup = time.perf_counter() - t0
    dt1 = sim.time(2)
    steps = max(2, int((budget_s - setup) / max(dt1 / 2, 1e-6)))
    steps = min(steps, 200)
    dt = sim.time(steps)
    total_steps = steps + 2
    secs = dt + dt1
    return {"value": N * total_steps / secs, "unit": "particle-steps/s", "cores": 1, "kind": "port",
            "sample": f"FP64 oracle (host-backend restatement, single thread), same LJ N={N} state, "
                      f"{total_steps} NVE steps after setup ({sim.rebuilds()} neighbour-list builds incl. setup), "
                      f"{secs:.1f} s"}


def batched(args):
    """Batched ensembles: 256 seeded systems (N log-uniform in [500, 32000], rho
    in [0.7, 1.2], T in [0.5, 2.0], LJ / WCA / Morse), NVE dt = 0.003, stepped
    together; against the plain per-job path (each system alone through the
    per-system modules, one after another) on every 16th system."""
    from halmd_amd import workloads
    specs = workloads.batched_specs(args.systems, seed=args.seed)
    ntot = sum(s["n"] for s in specs)
    t0 = time.perf_counter()
    e = workloads.ensemble_from_specs(specs, timestep=0.003)
    log(f"ensemble of {len(specs)} systems, {ntot} particles, setup {time.perf_counter() - t0:.1f} s")
    e.step(args.warmup)
    ms = e.step_timed(args.steps)
    value = ntot * args.steps / (ms / 1e3)
    log(f"batched: {ms / args.steps:.3f} ms/step, {value:.4e} particle-steps/s")
    pot0, kin0 = e.energies()
    sample = specs[::16]
    plain_steps = max(10, args.steps // 5)
    psum, tsum = 0.0, 0.0
    for sp in sample:
        w = workloads.single_from_spec(sp, timestep=0.003)
        w.sampler.run(5)
        t = w.sampler.run_timed(plain_steps)
        psum += sp["n"] * plain_steps
        tsum += t / 1e3
    plain = psum / tsum
    log(f"plain per-job path on {len(sample)} systems: {plain:.4e} particle-steps/s")
    out = {
        "metric": "aggregate particle-steps/s, batched ensemble of 256 systems vs plain per-job GPU path",
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
            "workload": f"{len(specs)} independent NVE systems, N log-uniform in [500, 32000] (total {ntot}), "
                        "rho U[0.7, 1.2], T U[0.5, 2.0], LJ rc 2.5 / WCA / Morse rc 2.5, dt 0.003, skin 0.3",
            "systems": len(specs),
            "particles": ntot,
            "plain_per_job_particle_steps_per_s": plain,
            "plain_sample": f"every 16th system ({len(sample)}), {plain_steps} timed steps each, sequential",
            "speedup_vs_plain": value / plain,
            "en_tot_per_particle_range": [float(np.min(pot0 + kin0)), float(np.max(pot0 + kin0))],
        },
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


def serving(args):
    """Served job stream: Poisson arrivals of LJ / Kob-Andersen / WCA jobs (N
    log-normal, median 2e4, clipped to [1e3, 2e5]; rho in [0.8, 1.2]; T in
    [0.45, 1.5]; steps log-uniform in [1e3, 1e5] times --step-scale), each with
    blocking-scheme MSD and ISF at 8 wavenumbers, through one server in real
    time; against the plain per-job path replayed first-in first-out with
    service times measured on a sample of the stream's jobs."""
    from collections import deque
    from halmd_amd import workloads
    from halmd_amd.serve import server
    stream = workloads.serving_stream(rate=args.rate, duration=args.duration, seed=args.seed,
                                      step_scale=args.step_scale)
    dt = 0.003
    k = list(workloads.SERVING_WAVENUMBERS)
    corr = dict(block_size=10, shift=2, separation=10, wavenumber=k, tolerance=0.01, max_count=7)
    t0 = time.perf_counter()
    inputs = [workloads.system_inputs(j["n"], j["density"], j["temperature"], j["seed"]) for j in stream]
    log(f"{len(stream)} jobs over {args.duration:.0f} s, inputs ready in {time.perf_counter() - t0:.1f} s")
    s = server(budget_bytes=args.budget_gib << 30, max_jobs=args.max_jobs, tick_steps=args.tick)
    offered = sum(j["n"] * j["steps"] for j in stream) / args.duration
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
                              species=workloads.spec_species(j), arrival=start + j["arrival"], **corr)
            inputs[q] = None
        st = s.stats()
        if not st["active"] and not st["queued"]:
            time.sleep(max(0.0, stream[pending[0]]["arrival"] - (s.now() - start)))
            continue
        s.tick()
        ticks += 1
        if ticks % 50 == 0:
            log(f"t = {s.now() - start:6.1f} s: {st['active']} running, {st['queued']} queued, "
                f"{len(ids)} submitted of {len(stream)}")
    end = s.now()
    lat = np.array([s.info(ids[q])["done"] - (start + stream[q]["arrival"]) for q in range(len(stream))])
    wait = np.array([s.info(ids[q])["admitted"] - (start + stream[q]["arrival"]) for q in range(len(stream))])
    psteps = s.stats()["particle_steps"]
    phases = s.phases()
    log("tick phases (s): " + ", ".join(f"{k} {v:.2f}" for k, v in phases.items()))
    value = psteps / (end - start)
    log(f"served: {value:.4e} particle-steps/s over {end - start:.1f} s; latency p50 {np.percentile(lat, 50):.2f} s, "
        f"p99 {np.percentile(lat, 99):.2f} s; offered {offered:.3e} particle-steps/s")

    # plain per-job path: one job at a time through the per-system modules
    # (with the same correlators); time per step fitted as a + b N on a sample
    from halmd_amd import observables
    from halmd_amd.observables.dynamics import blocking_scheme
    from halmd_amd.observables.utility import wavevector
    rng = np.random.default_rng(args.seed)
    sample = sorted(rng.choice(len(stream), size=min(8, len(stream)), replace=False),
                    key=lambda q: stream[q]["n"])
    ns, ts, setups = [], [], []
    for q in sample:
        j = stream[q]
        c0 = time.perf_counter()
        w = workloads.single_from_spec(j, timestep=dt)
        steps = -(-j["steps"] // args.tick) * args.tick   # whole ticks, as the server runs them
        nsteps = min(steps, 300)
        bs = blocking_scheme(w.box, w.particle, timestep=dt, max_lag=steps * dt, resolution=args.tick * dt,
                             block_size=10, shift=2, separation=10)
        bs.correlate_msd()
        bs.correlate_isf(observables.density_mode(w.particle, wavevector(w.box, k, tolerance=0.01, max_count=7)))
        w.sampler.on_sample_blocking_scheme(bs)
        w.sampler.run(1)
        setups.append(time.perf_counter() - c0)
        ms = w.sampler.run_timed(nsteps)
        ns.append(j["n"])
        ts.append(ms / 1e3 / nsteps)
    b, a = np.polyfit(ns, ts, 1)
    setup = float(np.median(setups))
    service = [setup + -(-j["steps"] // args.tick) * args.tick * max(a + b * j["n"], 0.0) for j in stream]
    fin, plat = 0.0, []
    for j, sv in zip(stream, service):
        fin = max(fin, j["arrival"]) + sv
        plat.append(fin - j["arrival"])
    plat = np.array(plat)
    plain = sum(j["n"] * j["steps"] for j in stream) / fin
    log(f"plain per-job path (FIFO replay): {plain:.4e} particle-steps/s; latency p50 "
        f"{np.percentile(plat, 50):.2f} s, p99 {np.percentile(plat, 99):.2f} s")
    out = {
        "metric": "served job stream: aggregate particle-steps/s and p50/p99 job latency vs plain per-job GPU path",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": ticks,
        "warmup": 0,
        "ms_per_step": (end - start) * 1e3 / max(ticks, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"Poisson {args.rate:g} jobs/s for {args.duration:g} s ({len(stream)} jobs): LJ / KA / WCA, "
                        "N log-normal (median 2e4, log-sigma 0.75, clipped [1e3, 2e5]), rho U[0.8, 1.2], "
                        f"T U[0.45, 1.5], steps log-uniform [1e3, 1e5] x {args.step_scale:g}, dt {dt}, "
                        "MSD + ISF at 8 wavenumbers (blocking scheme, block 10)",
            "tick_steps": args.tick,
            "max_jobs": args.max_jobs,
            "budget_gib": args.budget_gib,
            "offered_particle_steps_per_s": offered,
            "latency_p50_s": float(np.percentile(lat, 50)),
            "latency_p99_s": float(np.percentile(lat, 99)),
            "queue_wait_p50_s": float(np.percentile(wait, 50)),
            "queue_wait_p99_s": float(np.percentile(wait, 99)),
            "makespan_s": end - start,
            "tick_phase_seconds": phases,
            "plain_per_job_particle_steps_per_s": plain,
            "plain_latency_p50_s": float(np.percentile(plat, 50)),
            "plain_latency_p99_s": float(np.percentile(plat, 99)),
            "plain_method": f"FIFO replay of the same arrivals; time per step a + b N fitted on {len(sample)} "
                            f"sampled jobs ({min(ns)}..{max(ns)} particles, up to 300 timed steps each, same "
                            f"correlators), median setup {setup:.3f} s per job",
            "speedup_vs_plain": value / plain,
        },
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--warmup", type=int, default=400)
    ap.add_argument("--profile-steps", type=int, default=500)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="keep the default cell-segment shape")
    ap.add_argument("--segments", default=None,
                    help="cell-segment shape 'cells,rows,threads' instead of autotuning (profiling runs)")
    ap.add_argument("--workload", default="lj", choices=["lj", "batched", "serving", "tiny"],
                    help="lj: LJ fluid N = 2^20 (default); batched: ensemble of independent systems; "
                         "serving: Poisson job stream through the continuous-batching server; "
                         "tiny: Kob-Andersen N = 4096 with all-pairs neighbour lists")
    ap.add_argument("--rate", type=float, default=50.0, help="serving: job arrivals per second")
    ap.add_argument("--duration", type=float, default=10.0, help="serving: seconds of arrivals")
    ap.add_argument("--step-scale", type=float, default=1.0 / 30, help="serving: factor on the job step counts")
    ap.add_argument("--tick", type=int, default=50, help="serving: steps per tick")
    ap.add_argument("--max-jobs", type=int, default=256, help="serving: running-job cap")
    ap.add_argument("--budget-gib", type=int, default=128, help="serving: device memory budget for jobs")
    ap.add_argument("--systems", type=int, default=256, help="systems of the batched workload")
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
    arch, cus, mem = halmd_amd.device_info()
    if args.calibration_copy_mib:
        calibration_copy(args.calibration_copy_mib << 20)
    log(f"device {arch}, {cus} CUs, {mem / 2**30:.0f} GiB; building LJ N={N}")
    t0 = time.perf_counter()
    w = build_workload(args.seed)
    if args.segments:
        w.neighbour.set_segments(*[int(x) for x in args.segments.split(",")])
        log(f"cell segments set to {w.neighbour.segments} (cells, rows, workgroup size)")
    elif not args.no_tune:
        cells = w.force.tune()
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
    thermo = w.thermodynamics
    en_tot = thermo.en_tot()

    cpu = None
    if not args.no_cpu_baseline:
        log("CPU baseline (FP64 oracle, 1 thread)")
        cpu = cpu_baseline(w, args.cpu_seconds)
        cpu["host_cpus"] = os.cpu_count()

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
            "wavevectors": int(len(w.wavevector.value)),
            "neighbours_mean": n_nb,
            "neighbour_rebuilds": int(w.neighbour.rebuilds),
            "cell_segments": list(w.neighbour.segments),
            "en_tot_per_particle_end": en_tot,
            "device": arch,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "kernel": kname,
            "bytes_per_launch": force_bytes,
            "avg_ms": f_avg,
            # from the counter summary: fraction of the VALU issue rate
            # (wave64 instruction = 4 cycles on 1024 SIMDs) the kernel used
            "valu_issue_frac": prof.get("valu_issue_frac"),
            "counters": os.path.relpath(TRAFFIC_PROFILE, ROOT) if prof else None,
        },
        "kernels": kernels,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
