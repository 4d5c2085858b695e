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

    def run(self, steps):
        call("halmd_observables_sampler_run", self.handle, int(steps))

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
