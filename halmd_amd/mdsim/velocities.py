"""halmd.mdsim.velocities.boltzmann: Maxwell-Boltzmann velocities on the host.

Follows halmd/mdsim/host/velocities/boltzmann.cpp:53-97: v ~ N(0, sqrt(T/m))
per component, then the centre-of-mass velocity is removed.  The random stream
is numpy's PCG64 seeded with `seed`, so the same seed gives the same FP64
velocities to the device and to any checker.
"""
import numpy as np

__all__ = ["maxwell_boltzmann", "boltzmann"]


def maxwell_boltzmann(n, temperature, mass=1.0, seed=0):
    rng = np.random.default_rng(seed)
    m = np.broadcast_to(np.asarray(mass, dtype=np.float64), (n,))
    v = rng.normal(0.0, np.sqrt(temperature), (n, 3)) / np.sqrt(m)[:, None]
    v_cm = (m[:, None] * v).sum(axis=0) / m.sum()
    return v - v_cm


class boltzmann:
    """velocities.boltzmann{particle, temperature}: set() assigns the velocities."""

    def __init__(self, particle, temperature, seed=0):
        self.particle, self.temperature, self.seed = particle, float(temperature), seed

    def set(self):
        v = maxwell_boltzmann(self.particle.nparticle, self.temperature, self.particle.data["mass"], self.seed)
        self.particle.data["velocity"] = v
        return v
