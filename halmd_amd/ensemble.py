"""Batched ensembles: many independent NVE systems stepped in shared launches.

Each system is a cubic box of one species (unit mass) with a truncated pair
potential and velocity-Verlet integration; all systems advance together (one
integrate launch, one batched rebuild of the systems that need it, one force
launch per potential class per step).  Over include/halmd_hip.h
(halmd_ensemble_*).
"""
import ctypes

import numpy as np

from ._lib import Handle, call, dptr, f64, iptr, out_handle

__all__ = ["ensemble", "POTENTIALS"]

# potential name -> (kind, epsilon, sigma, cutoff in sigma, r_min, distortion):
# Lennard-Jones r_c = 2.5 sigma, WCA (LJ cut at its minimum 2^(1/6) sigma) and
# Morse with D = 1, a = 6 / sigma_LJ, r0 = sigma_LJ, r_c = 2.5 sigma_LJ
POTENTIALS = {
    "lj": (0, 1.0, 1.0, 2.5, 1.0, 1.0),
    "wca": (0, 1.0, 1.0, 2.0 ** (1.0 / 6), 1.0, 1.0),
    "morse": (1, 1.0, 1.0 / 6, 15.0, 1.0, 1.0),
}


class ensemble(Handle):
    _destroy = "halmd_ensemble_destroy"

    def __init__(self):
        super().__init__()
        self._h = out_handle()
        call("halmd_ensemble_create", ctypes.byref(self._h))
        self.n = []

    def add(self, position, velocity, length, potential="lj", truncation="shifted", timestep=0.005, skin=0.3,
            smoothing=0.005):
        """Add one system (position, velocity: n x 3); returns its index."""
        r = f64(position)
        v = f64(velocity)
        kind, eps, sigma, rc, r_min, dist = POTENTIALS[potential] if isinstance(potential, str) else potential
        idx = ctypes.c_uint()
        call("halmd_ensemble_add", self.handle, len(r), float(length), int(kind), truncation.encode(), float(eps),
             float(sigma), float(rc), float(r_min), float(dist), float(smoothing), float(timestep), float(skin),
             dptr(r), dptr(v), ctypes.byref(idx))
        self.n.append(len(r))
        return idx.value

    def commit(self):
        call("halmd_ensemble_commit", self.handle)

    def step(self, steps=1):
        call("halmd_ensemble_step", self.handle, int(steps))

    def step_timed(self, steps):
        """Steps with HIP events around them; elapsed milliseconds."""
        ms = ctypes.c_double()
        call("halmd_ensemble_step_timed", self.handle, int(steps), ctypes.byref(ms))
        return ms.value

    def thermo(self):
        """Per system: (sum m v^2, sum en_pot, sum of stress traces, n)."""
        out = np.zeros((len(self.n), 4))
        call("halmd_ensemble_thermo", self.handle, dptr(out))
        return out

    def energies(self):
        """Per system: (en_pot, en_kin) per particle."""
        t = self.thermo()
        return t[:, 1] / t[:, 3], 0.5 * t[:, 0] / t[:, 3]

    def state(self, k):
        """(positions, velocities, images) of system k in particle-ID order."""
        n = self.n[k]
        r = np.zeros((n, 3))
        v = np.zeros((n, 3))
        im = np.zeros((n, 3), dtype=np.int32)
        call("halmd_ensemble_state", self.handle, int(k), dptr(r), dptr(v), iptr(im))
        return r, v, im

    def rebuilds(self, k):
        c = ctypes.c_uint64()
        call("halmd_ensemble_rebuilds", self.handle, int(k), ctypes.byref(c))
        return c.value
