"""halmd.observables.dynamics: blocking_scheme with MSD and ISF correlations.

Mirrors lua/halmd/observables/dynamics/{blocking_scheme,correlation,
mean_square_displacement,intermediate_scattering_function}.lua.in.  Position
samples stay on the device (FP64, unwrapped, ID order); results are Welford
accumulators per (level, lag[, shell]) as in correlation.hpp:207-224.
"""
import ctypes

import numpy as np

from .._lib import Handle, call, dptr, out_handle, u64ptr

__all__ = ["blocking_scheme"]


class blocking_scheme(Handle):
    """blocking_scheme{max_lag, every (resolution in steps*timestep), size, shift, separation}."""

    _destroy = "halmd_observables_dynamics_blocking_scheme_destroy"

    def __init__(self, box, particle, timestep, max_lag, resolution, block_size=10, shift=None, separation=None):
        super().__init__()
        if shift is None:
            shift = int(np.floor(np.sqrt(block_size)))
        if separation is None:
            separation = block_size
        self._deps = (box, particle)
        self.timestep = float(timestep)
        self._h = out_handle()
        call("halmd_observables_dynamics_blocking_scheme_create", particle.handle, box.handle, float(timestep),
             float(max_lag), float(resolution), int(block_size), int(shift), int(separation), ctypes.byref(self._h))
        lv, bs = ctypes.c_uint(), ctypes.c_uint()
        call("halmd_observables_dynamics_blocking_scheme_levels", self.handle, ctypes.byref(lv), ctypes.byref(bs))
        self.levels, self.block_size = lv.value, bs.value
        iv = np.zeros(self.levels, dtype=np.uint64)
        call("halmd_observables_dynamics_blocking_scheme_intervals", self.handle, u64ptr(iv))
        self.interval = iv.astype(np.int64)
        self.density_mode = None

    @property
    def time(self):
        """Lag times [levels][block_size] (blocking_scheme.cpp time_ grid)."""
        return self.interval[:, None] * np.arange(self.block_size)[None, :] * self.timestep

    def correlate_msd(self):
        call("halmd_observables_dynamics_blocking_scheme_add_msd", self.handle)

    def correlate_mqd(self):
        call("halmd_observables_dynamics_blocking_scheme_add_mqd", self.handle)

    def correlate_vacf(self):
        call("halmd_observables_dynamics_blocking_scheme_add_vacf", self.handle)

    def correlate_isf(self, density_mode):
        call("halmd_observables_dynamics_blocking_scheme_add_isf", self.handle, density_mode.handle)
        self.density_mode = density_mode

    def sample(self, step):
        call("halmd_observables_dynamics_blocking_scheme_sample", self.handle, int(step))

    def finalise(self):
        call("halmd_observables_dynamics_blocking_scheme_finalise", self.handle)

    def result(self, which):
        """(count, mean, error) arrays; which = 'msd', 'mqd', 'vacf' or 'isf'."""
        if which in ("msd", "mqd", "vacf"):
            shape = (self.levels, self.block_size)
        else:
            shape = (self.levels, self.block_size, len(self.density_mode.wavevector.shell))
        size = int(np.prod(shape))
        c = np.zeros(size, dtype=np.uint64)
        m = np.zeros(size)
        e = np.zeros(size)
        call("halmd_observables_dynamics_blocking_scheme_result", self.handle, which.encode(), u64ptr(c), dptr(m),
             dptr(e))
        return c.reshape(shape).astype(np.int64), m.reshape(shape), e.reshape(shape)
