"""halmd.observables: sampler, thermodynamics, density_mode, phase_space, utility, dynamics.

Mirrors lua/halmd/observables/*.lua.in over include/halmd_hip.h.
"""
import ctypes

import numpy as np

from .._lib import Handle, call, dptr, out_handle

__all__ = ["sampler", "thermodynamics", "density_mode", "phase_space"]


class phase_space(Handle):
    """Particle samples in particle-ID order (lua/halmd/observables/phase_space.lua.in:165-260,
    observables/gpu/phase_space.cpp).  acquire(name) for name in position
    (extended by the periodic image, r + image L), velocity, species, mass;
    set({name: array}) writes them back (positions reduced into the box, their
    images kept)."""

    _destroy = "halmd_observables_phase_space_destroy"
    _SHAPES = {"position": (3, np.float64), "velocity": (3, np.float64), "species": (1, np.uint32),
               "mass": (1, np.float64)}

    def __init__(self, box, particle):
        super().__init__()
        self._deps = (box, particle)
        self.box, self.particle = box, particle
        self._h = out_handle()
        call("halmd_observables_phase_space_create", particle.handle, box.handle, ctypes.byref(self._h))

    def _shape(self, name):
        if name not in self._SHAPES:
            raise ValueError(f"phase_space: unknown sample {name!r}")
        d, t = self._SHAPES[name]
        n = self.particle.nparticle
        return ((n, d) if d > 1 else (n,)), t

    def acquire(self, name):
        shape, t = self._shape(name)
        out = np.empty(shape, dtype=t)
        call("halmd_observables_phase_space_acquire", self.handle, name.encode(), out.ctypes.data_as(ctypes.c_void_p))
        return out

    def acquire_position(self):
        return self.acquire("position")

    def acquire_velocity(self):
        return self.acquire("velocity")

    def acquire_species(self):
        return self.acquire("species")

    def acquire_mass(self):
        return self.acquire("mass")

    # fields a writer can take, in particle-ID order: the phase-space data of
    # the reference (position reduced into the box, image, velocity, species,
    # mass, force) and the float-float words of position and velocity, which
    # carry the state exactly for a restart
    _FIELDS = ("position", "image", "velocity", "species", "mass", "force", "potential_energy",
               "position_parts", "velocity_parts")

    def data(self, name):
        """Slot returning the sample `name` in particle-ID order (phase_space:data)."""
        if name not in self._FIELDS:
            raise ValueError(f"phase_space: unknown data field {name!r}")
        return lambda: self.particle.ordered(name)

    def writer(self, file, fields, location=("particles", "all"), every=None, sampler=None):
        """lua/halmd/observables/phase_space.lua.in:236-275: an append writer at
        `location` with one field per name (a dict maps file names to data
        names), the box written once beside it, and the writer connected to
        the sampler every `every` steps (else at the start and the finish)."""
        if sampler is None:
            raise ValueError("phase_space writer: the sampler to connect to is required")
        pairs = fields.items() if isinstance(fields, dict) else [(f, f) for f in fields]
        w = file.writer(location=list(location), mode="append", clock=sampler.core)
        for name, field in pairs:
            w.on_write(self.data(field), [name])
        self.box.writer(file=file, location=list(location))
        if every:
            sampler.on_sample(w.write, int(every), sampler.core.step)
        else:
            sampler.on_start(w.write)
            sampler.on_finish(w.write)
        return w

    @staticmethod
    def reader(file, fields, location=("particles", "all")):
        """lua/halmd/observables/phase_space.lua.in:306-370: (reader, sample);
        reader.read_at_step(offset) / read_at_time(offset) fill sample[name]
        (particle-ID order) from the file."""
        pairs = fields.items() if isinstance(fields, dict) else [(f, f) for f in fields]
        r = file.reader(location=list(location), mode="append")
        sample = {}
        dtypes = {"image": np.int32, "species": np.uint32, "position_parts": np.float32,
                  "velocity_parts": np.float32}
        for name, field in pairs:
            def slot(a, field=field):
                sample[field] = a
            r.on_read(slot, [name], dtypes.get(field, np.float64))
        return r, sample

    def set(self, samples):
        """samples: {name: array in particle-ID order} (phase_space.lua.in:116-130)."""
        for name, value in samples.items():
            shape, t = self._shape(name)
            a = np.ascontiguousarray(np.broadcast_to(np.asarray(value, dtype=t), shape))
            call("halmd_observables_phase_space_set", self.handle, name.encode(), a.ctypes.data_as(ctypes.c_void_p))


class thermodynamics(Handle):
    """Mean per-particle thermodynamic quantities, FP64 device reductions
    (lua/halmd/observables/thermodynamics.lua.in; gpu/thermodynamics.cpp:67-191)."""

    _destroy = "halmd_observables_thermodynamics_destroy"

    def __init__(self, box, particle, integrator=None):
        super().__init__()
        self._deps = (box, particle, integrator)
        self.box, self.particle = box, particle
        self._h = out_handle()
        call("halmd_observables_thermodynamics_create", particle.handle, box.handle,
             integrator.handle if integrator is not None else None, ctypes.byref(self._h))

    def _get(self, name):
        v = ctypes.c_double()
        call(f"halmd_observables_thermodynamics_{name}", self.handle, ctypes.byref(v))
        return v.value

    def en_kin(self):
        return self._get("en_kin")

    def en_pot(self):
        return self._get("en_pot")

    def virial(self):
        return self._get("virial")

    def pressure(self):
        return self._get("pressure")

    def temperature(self):
        return self._get("temperature")

    def en_tot(self):
        return self._get("en_tot")

    def center_of_mass_velocity(self):
        v = np.zeros(3)
        call("halmd_observables_thermodynamics_center_of_mass_velocity", self.handle, dptr(v))
        return v

    def stress_tensor_pot(self):
        v = np.zeros(6)
        call("halmd_observables_thermodynamics_stress_tensor_pot", self.handle, dptr(v))
        return v

    def particle_number(self):
        return self.particle.nparticle

    def density(self):
        return self.particle.nparticle / self.box.volume

    def volume(self):
        return self.box.volume

    # H5MD names of the observables (lua/halmd/observables/thermodynamics.lua.in:58-150)
    _OBSERVABLES = {"potential_energy": "en_pot", "kinetic_energy": "en_kin", "internal_energy": "en_tot",
                    "pressure": "pressure", "temperature": "temperature", "virial": "virial",
                    "center_of_mass_velocity": "center_of_mass_velocity", "stress_tensor": "stress_tensor_pot",
                    "particle_number": "particle_number", "density": "density", "volume": "volume"}

    def writer(self, file, every, sampler, fields=("potential_energy", "pressure", "temperature",
                                                    "center_of_mass_velocity"), location=("observables",)):
        """lua/halmd/observables/thermodynamics.lua.in:213-284: the H5MD module
        thermodynamics {1, 0}; dimension, volume, particle number and density
        written once (truncate); then one append sample of each field every
        `every` steps."""
        pairs = fields.items() if isinstance(fields, dict) else [(f, f) for f in fields]
        file.write_attribute_int("h5md/modules/thermodynamics", "version", [1, 0])
        t = file.writer(location=list(location), mode="truncate")
        file.write_attribute_int(list(location), "dimension", 3)
        for v in ("volume", "particle_number", "density"):
            t.on_write(getattr(self, v), [v])
        t.write()
        w = file.writer(location=list(location), mode="append", clock=sampler.core)
        for name, field in pairs:
            if field in ("particle_number", "density", "volume"):
                continue   # written once above (a group of fixed size)
            w.on_write(getattr(self, self._OBSERVABLES[field]), [name])
        sampler.on_sample(w.write, int(every), sampler.core.step)
        return w


class density_mode(Handle):
    """rho_k = sum_j exp(-i k.r_j) (lua/halmd/observables/density_mode.lua.in;
    gpu/density_mode.cpp:140-167)."""

    _destroy = "halmd_observables_density_mode_destroy"

    def __init__(self, particle, wavevector):
        super().__init__()
        self._deps = (particle, wavevector)
        self.wavevector = wavevector
        self._h = out_handle()
        call("halmd_observables_density_mode_create", particle.handle, wavevector.handle, ctypes.byref(self._h))

    def acquire(self):
        out = np.zeros(2 * len(self.wavevector.value))
        call("halmd_observables_density_mode_acquire", self.handle, dptr(out))
        return out[0::2] + 1j * out[1::2]


class sampler(Handle):
    """Step loop with sample slots (lua/halmd/observables/sampler.lua.in; sampler.cpp:42-130)."""

    _destroy = "halmd_observables_sampler_destroy"

    def __init__(self, core):
        super().__init__()
        self.core = core
        self._slots = []
        self._h = out_handle()
        call("halmd_observables_sampler_create", core.handle, ctypes.byref(self._h))

    def on_sample_thermodynamics(self, thermodynamics, every):
        call("halmd_observables_sampler_add_thermodynamics", self.handle, thermodynamics.handle, int(every))
        self._slots.append(thermodynamics)
        return len([s for s in self._slots if isinstance(s, globals()["thermodynamics"])]) - 1

    def on_sample_blocking_scheme(self, blocking_scheme):
        call("halmd_observables_sampler_add_blocking_scheme", self.handle, blocking_scheme.handle)
        self._slots.append(blocking_scheme)

    def on_sample_density_mode(self, density_mode, every):
        call("halmd_observables_sampler_add_density_mode", self.handle, density_mode.handle, int(every))
        self._slots.append(density_mode)
        self._modes = getattr(self, "_modes", []) + [density_mode]
        return len(self._modes) - 1

    def density_mode_records(self, k=0):
        """(steps, rho[nrec, nq]) of density-mode slot k."""
        nq = len(self._modes[k].wavevector.value)
        n = ctypes.c_size_t()
        call("halmd_observables_sampler_density_mode_records", self.handle, int(k), None, 0, ctypes.byref(n))
        out = np.zeros((n.value, 1 + 2 * nq))
        call("halmd_observables_sampler_density_mode_records", self.handle, int(k), dptr(out), n.value,
             ctypes.byref(n))
        return out[:, 0].astype(np.int64), out[:, 1::2] + 1j * out[:, 2::2]

    # ---- slots of the host (observables/sampler.cpp:42-100): on_start,
    # on_sample every `interval` steps counted from `start`, on_finish; the
    # device slots above run inside the native loop
    def on_start(self, slot):
        self._py_start = getattr(self, "_py_start", []) + [slot]

    def on_sample(self, slot, interval, start=0):
        if interval < 1:
            raise ValueError("sampler: interval must be positive")
        self._py_sample = getattr(self, "_py_sample", []) + [(slot, int(interval), int(start))]

    def on_finish(self, slot):
        self._py_finish = getattr(self, "_py_finish", []) + [slot]

    def sample(self):
        """Sample the current state: every on_sample slot due at this step
        (sampler.cpp:42-47)."""
        step = self.core.step
        for slot, interval, start in getattr(self, "_py_sample", []):
            if (step - start) % interval == 0:
                slot()

    def start(self):
        for slot in getattr(self, "_py_start", []):
            slot()

    def finish(self):
        for slot in getattr(self, "_py_finish", []):
            slot()

    def run(self, steps):
        """Advance `steps` steps (sampler.cpp:49-83): the on_start slots at the
        first run, the on_sample slots after each step at which they are due."""
        if not getattr(self, "_started", False):
            self._started = True
            self.start()
        slots = getattr(self, "_py_sample", [])
        if not slots:
            call("halmd_observables_sampler_run", self.handle, int(steps))
            return
        step, end = self.core.step, self.core.step + int(steps)
        while step < end:
            nxt = min([step + interval - (step - start) % interval for _, interval, start in slots] + [end])
            call("halmd_observables_sampler_run", self.handle, int(nxt - step))
            step = nxt
            for slot, interval, start in slots:
                if (step - start) % interval == 0:
                    slot()

    def run_timed(self, steps):
        ms = ctypes.c_double()
        call("halmd_observables_sampler_run_timed", self.handle, int(steps), ctypes.byref(ms))
        return ms.value

    def thermodynamics_records(self, k=0):
        """(step, en_pot, en_kin, virial, en_nhc) per record."""
        n = ctypes.c_size_t()
        call("halmd_observables_sampler_thermodynamics_records", self.handle, int(k), None, 0, ctypes.byref(n))
        out = np.zeros((n.value, 5))
        call("halmd_observables_sampler_thermodynamics_records", self.handle, int(k), dptr(out), n.value,
             ctypes.byref(n))
        return out


from . import utility, dynamics  # noqa: E402,F401
