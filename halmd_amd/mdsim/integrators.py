"""halmd.mdsim.integrators: verlet, verlet_nvt_hoover, verlet_nvt_andersen.

Mirrors lua/halmd/mdsim/integrators/*.lua.in over include/halmd_hip.h.
"""
import ctypes

import numpy as np

from .._lib import Handle, call, dptr, out_handle

__all__ = ["verlet", "verlet_nvt_hoover", "verlet_nvt_andersen"]


class _integrator(Handle):
    _destroy = "halmd_mdsim_integrators_destroy"

    def integrate(self):
        call("halmd_mdsim_integrators_integrate", self.handle)

    def finalize(self):
        call("halmd_mdsim_integrators_finalize", self.handle)

    def set_timestep(self, timestep):
        call("halmd_mdsim_integrators_set_timestep", self.handle, float(timestep))

    @property
    def timestep(self):
        v = ctypes.c_double()
        call("halmd_mdsim_integrators_timestep", self.handle, ctypes.byref(v))
        return v.value


class verlet(_integrator):
    """Velocity-Verlet (integrators/verlet.lua.in:90-126; gpu/integrators/verlet.cpp:124-155)."""

    def __init__(self, box, particle, timestep):
        super().__init__()
        self._deps = (box, particle)
        self._h = out_handle()
        call("halmd_mdsim_integrators_verlet_create", particle.handle, box.handle, float(timestep),
             ctypes.byref(self._h))

    def defer_finalize(self, enable=True):
        """Fold the finalize half kick into the next integrate pass (default on)."""
        call("halmd_mdsim_integrators_verlet_defer_finalize", self.handle, int(bool(enable)))


class verlet_nvt_hoover(_integrator):
    """Nose-Hoover chain (M=2) thermostat (integrators/verlet_nvt_hoover.lua.in:165-255;
    gpu/integrators/verlet_nvt_hoover.cpp:270-313).  The chain lives on the device."""

    def __init__(self, box, particle, timestep, temperature, resonance_frequency):
        super().__init__()
        self._deps = (box, particle)
        self._h = out_handle()
        call("halmd_mdsim_integrators_verlet_nvt_hoover_create", particle.handle, box.handle, float(timestep),
             float(temperature), float(resonance_frequency), ctypes.byref(self._h))

    def set_temperature(self, temperature):
        call("halmd_mdsim_integrators_verlet_nvt_hoover_set_temperature", self.handle, float(temperature))

    @property
    def position(self):
        xi = np.zeros(2); v = np.zeros(2)
        call("halmd_mdsim_integrators_verlet_nvt_hoover_chain", self.handle, dptr(xi), dptr(v))
        return xi

    @property
    def velocity(self):
        xi = np.zeros(2); v = np.zeros(2)
        call("halmd_mdsim_integrators_verlet_nvt_hoover_chain", self.handle, dptr(xi), dptr(v))
        return v

    def set_chain(self, position, velocity):
        xi = np.ascontiguousarray(position, dtype=np.float64)
        v = np.ascontiguousarray(velocity, dtype=np.float64)
        call("halmd_mdsim_integrators_verlet_nvt_hoover_set_chain", self.handle, dptr(xi), dptr(v))

    @property
    def mass(self):
        m = np.zeros(2)
        call("halmd_mdsim_integrators_verlet_nvt_hoover_mass", self.handle, dptr(m))
        return m

    def set_mass(self, mass):
        m = np.ascontiguousarray(mass, dtype=np.float64)
        call("halmd_mdsim_integrators_verlet_nvt_hoover_set_mass", self.handle, dptr(m))

    @property
    def internal_energy(self):
        v = ctypes.c_double()
        call("halmd_mdsim_integrators_internal_energy", self.handle, ctypes.byref(v))
        return v.value


class verlet_nvt_andersen(_integrator):
    """Andersen thermostat (integrators/verlet_nvt_andersen.lua.in;
    gpu/integrators/verlet_nvt_andersen.cpp:146-183); counter-based RNG."""

    def __init__(self, box, particle, timestep, temperature, rate, seed=5489):
        super().__init__()
        self._deps = (box, particle)
        self._h = out_handle()
        call("halmd_mdsim_integrators_verlet_nvt_andersen_create", particle.handle, box.handle, int(seed),
             float(timestep), float(temperature), float(rate), ctypes.byref(self._h))
