"""Batched ensembles: many independent NVE systems stepped in shared launches.

Each system is a cubic box of unit-mass particles with a truncated pair
potential and velocity-Verlet integration; all systems advance together (one
integrate launch, one batched rebuild of the systems that need it, one force
launch per potential class per step).  Systems may be added and retired
between steps; the next step (or apply()) places them without touching the
others' state.  Over include/halmd_hip.h (halmd_ensemble_*).
"""
import ctypes

import numpy as np

from ._lib import Handle, call, dptr, f64, iptr, out_handle

__all__ = ["ensemble", "POTENTIALS", "potential_matrices", "KA_EPSILON", "KA_SIGMA"]

KA_EPSILON = np.array([[1.0, 1.5], [1.5, 0.5]])
KA_SIGMA = np.array([[1.0, 0.8], [0.8, 0.88]])

# potential name -> (kind, epsilon, sigma, cutoff in sigma, r_min, distortion):
# Lennard-Jones r_c = 2.5 sigma, WCA (LJ cut at its minimum 2^(1/6) sigma),
# Morse with D = 1, a = 6 / sigma_LJ, r0 = sigma_LJ, r_c = 2.5 sigma_LJ, and
# the Kob-Andersen binary Lennard-Jones mixture (r_c = 2.5 sigma_ab)
POTENTIALS = {
    "lj": (0, 1.0, 1.0, 2.5, 1.0, 1.0),
    "wca": (0, 1.0, 1.0, 2.0 ** (1.0 / 6), 1.0, 1.0),
    "morse": (1, 1.0, 1.0 / 6, 15.0, 1.0, 1.0),
    "ka": (0, KA_EPSILON, KA_SIGMA, 2.5, 1.0, 1.0),
}


def potential_matrices(potential):
    """(kind, ntype, epsilon, sigma, r_cut_sigma, r_min, distortion) as ntype x ntype arrays."""
    kind, eps, sigma, rc, r_min, dist = POTENTIALS[potential] if isinstance(potential, str) else potential
    eps = np.atleast_2d(np.asarray(eps, dtype=np.float64))
    ntype = eps.shape[0]
    full = lambda x: np.ascontiguousarray(np.broadcast_to(np.asarray(x, dtype=np.float64), (ntype, ntype)))
    return kind, ntype, full(eps), full(sigma), full(rc), full(r_min), full(dist)


class ensemble(Handle):
    _destroy = "halmd_ensemble_destroy"

    def __init__(self):
        super().__init__()
        self._h = out_handle()
        call("halmd_ensemble_create", ctypes.byref(self._h))
        self.n = []

    def add(self, position, velocity, length, potential="lj", truncation="shifted", timestep=0.005, skin=0.3,
            smoothing=0.005, species=None):
        """Add one system (position, velocity: n x 3); returns its index."""
        r = f64(position)
        v = f64(velocity)
        kind, ntype, eps, sigma, rc, r_min, dist = potential_matrices(potential)
        idx = ctypes.c_uint()
        if ntype == 1 and species is None:
            call("halmd_ensemble_add", self.handle, len(r), float(length), int(kind), truncation.encode(),
                 float(eps[0, 0]), float(sigma[0, 0]), float(rc[0, 0]), float(r_min[0, 0]), float(dist[0, 0]),
                 float(smoothing), float(timestep), float(skin), dptr(r), dptr(v), ctypes.byref(idx))
        else:
            sp = np.ascontiguousarray(np.zeros(len(r)) if species is None else species, dtype=np.uint32)
            call("halmd_ensemble_add_mixture", self.handle, len(r), float(length), int(kind), truncation.encode(),
                 int(ntype), dptr(eps), dptr(sigma), dptr(rc), dptr(r_min), dptr(dist), float(smoothing),
                 float(timestep), float(skin), dptr(r), dptr(v), sp.ctypes.data_as(ctypes.POINTER(ctypes.c_uint)),
                 ctypes.byref(idx))
        self.n.append(len(r))
        return idx.value

    def commit(self):
        call("halmd_ensemble_commit", self.handle)

    def retire(self, k):
        """Drop system k at the next step or apply() (later indices shift down)."""
        call("halmd_ensemble_retire", self.handle, int(k))
        self._retired = getattr(self, "_retired", []) + [k]

    def apply(self):
        call("halmd_ensemble_apply", self.handle)
        for k in sorted(getattr(self, "_retired", []), reverse=True):
            del self.n[k]
        self._retired = []

    def step(self, steps=1):
        if getattr(self, "_retired", None):
            self.apply()
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

    def force(self, k):
        """Pair force on each particle of system k at its current positions (particle-ID order)."""
        f = np.zeros((self.n[k], 3))
        call("halmd_ensemble_force", self.handle, int(k), dptr(f))
        return f

    def rebuilds(self, k):
        c = ctypes.c_uint64()
        call("halmd_ensemble_rebuilds", self.handle, int(k), ctypes.byref(c))
        return c.value

    def set_device_steps(self, on=True):
        """Rebuild decisions on the device (default) or on the host after each step."""
        call("halmd_ensemble_set_device_steps", self.handle, int(bool(on)))

    @property
    def recoveries(self):
        """Systems rebuilt on the host after a device-decided stretch stopped them."""
        c = ctypes.c_uint64()
        call("halmd_ensemble_recoveries", self.handle, ctypes.byref(c))
        return c.value

    def set_segments(self, target=200.0, threads=256):
        """About `target` particles per cell segment, `threads` per workgroup (before commit)."""
        call("halmd_ensemble_set_segments", self.handle, float(target), int(threads))

    def set_list_capacity(self, k, capacity=0):
        """List capacity of system k (0: its largest present list); lists rebuilt at once."""
        call("halmd_ensemble_set_list_capacity", self.handle, int(k), int(capacity))
