"""halmd.mdsim.forces: pair_trunc, pair_full and the pair convenience constructor.

Mirrors lua/halmd/mdsim/forces/{pair,pair_trunc,pair_full}.lua.in over include/halmd_hip.h.
"""
import ctypes

from .._lib import Handle, call, out_handle
from . import neighbour as _neighbour

__all__ = ["pair_trunc", "pair_full", "pair"]


class pair_trunc(Handle):
    """Truncated pair force (lua/halmd/mdsim/forces/pair_trunc.lua.in:109-167;
    gpu/forces/pair_trunc.hpp:340-374).  Registers itself with the particle
    instance, whose force is then evaluated lazily when read."""

    _destroy = "halmd_mdsim_forces_pair_trunc_destroy"

    def __init__(self, box, particle, potential, neighbour=None, weight=1.0):
        super().__init__()
        pair = None
        if isinstance(particle, (tuple, list)):
            if len(particle) != 2:
                raise ValueError("bad argument 'particle'")
            if particle[0] is not particle[1]:
                pair = tuple(particle)
            particle = particle[0]
        if getattr(potential, "r_cut", None) is None:
            raise ValueError("pair_trunc: the potential must be truncated first")
        if neighbour is None or isinstance(neighbour, dict):
            kw = dict(neighbour or {})
            neighbour = _neighbour(particle=pair or particle, box=box, r_cut=potential.r_cut, **kw)
        self.neighbour = neighbour
        self.potential = potential
        self.particle = particle
        self._deps = (box, pair)
        self._h = out_handle()
        if pair:
            # the force on particle[0] from the particles of particle[1]
            call("halmd_mdsim_forces_pair_trunc_create2", potential.handle, pair[0].handle, pair[1].handle,
                 box.handle, neighbour.handle, float(weight), ctypes.byref(self._h))
        else:
            call("halmd_mdsim_forces_pair_trunc_create", potential.handle, particle.handle, box.handle,
                 neighbour.handle, float(weight), ctypes.byref(self._h))

    def apply(self):
        call("halmd_mdsim_forces_pair_trunc_apply", self.handle)

    def tune(self):
        """Autotune the cell-segment width of the neighbour lists (timed on the
        device); returns the cells per segment kept, 0 if not applicable."""
        v = ctypes.c_uint()
        call("halmd_mdsim_forces_pair_trunc_tune", self.handle, ctypes.byref(v))
        return v.value


class pair_full(Handle):
    """Untruncated pair force over all pairs (lua/halmd/mdsim/forces/pair_full.lua.in;
    gpu/forces/pair_full.hpp:116-313).  particle is one instance or a pair
    (particle1, particle2): the force on particle1 from every particle of
    particle2.  The potential must not be truncated."""

    _destroy = "halmd_mdsim_forces_pair_full_destroy"

    def __init__(self, box, particle, potential, weight=1.0):
        super().__init__()
        if not isinstance(particle, (tuple, list)):
            particle = (particle, particle)
        if len(particle) != 2:
            raise ValueError("bad argument 'particle'")
        if getattr(potential, "r_cut", None) is not None:
            raise ValueError("pair_full: the potential is truncated (use pair_trunc)")
        self.potential = potential
        self.particle = particle[0]
        self._deps = (box, particle[1])
        self._h = out_handle()
        call("halmd_mdsim_forces_pair_full_create", potential.handle, particle[0].handle, particle[1].handle,
             box.handle, float(weight), ctypes.byref(self._h))

    def apply(self):
        call("halmd_mdsim_forces_pair_full_apply", self.handle)


def pair(box, particle, potential, neighbour=None, weight=1.0):
    """forces.pair{box, particle, potential, neighbour} (lua/halmd/mdsim/forces/pair.lua.in):
    truncated potentials use pair_trunc, untruncated ones pair_full."""
    if getattr(potential, "r_cut", None) is None:
        return pair_full(box=box, particle=particle, potential=potential, weight=weight)
    return pair_trunc(box=box, particle=particle, potential=potential, neighbour=neighbour, weight=weight)
