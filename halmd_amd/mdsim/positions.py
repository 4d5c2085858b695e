"""halmd.mdsim.positions.lattice: fcc placement on the host.

Follows halmd/mdsim/host/positions/lattice.cpp:65-170 (smallest lattice
constant whose unit cells hold all particles, vacancies spread uniformly) and
halmd/mdsim/positions/lattice_primitive.hpp:121-129 (close-packed unit cell,
x-fastest unit-cell index).  Positions are FP64 and shared unchanged with any
checker run on the same seed.
"""
import numpy as np

__all__ = ["fcc", "lattice"]


def fcc(npart, length, slab=(1.0, 1.0, 1.0)):
    """Return (positions [npart, 3], lattice constant a)."""
    length = np.asarray(length, dtype=np.float64) * np.asarray(slab, dtype=np.float64)
    offset = -length / 2
    u = 4.0
    V = np.prod(length) / np.ceil(npart / u)
    a = V ** (1.0 / 3)
    n = (length / a).astype(np.int64)
    while npart > u * np.prod(n):
        t = length / (n + 1)
        a = t.max()
        m = n.copy()
        n = np.maximum(m, (length / a).astype(np.int64))
        if np.array_equal(m, n):
            n = n + 1
    N = int(u * np.prod(n))
    skip = int(np.ceil(N / (N - npart))) if N > npart else 0
    idx = np.arange(N, dtype=np.int64)
    if skip:
        idx = idx[(idx % skip) != skip - 1][:npart]
    else:
        idx = idx[:npart]
    cell = idx >> 2
    cx = cell % n[0]
    cy = (cell // n[0]) % n[1]
    cz = cell // n[0] // n[1]
    r = np.empty((npart, 3))
    r[:, 0] = cx + (((idx ^ (idx >> 1)) & 1) * 2 + 1) / 4.0
    r[:, 1] = cy + ((idx & 1) * 2 + 1) / 4.0
    r[:, 2] = cz + ((idx & 2) + 1) / 4.0
    return r * a + offset, a


class lattice:
    """positions.lattice{box, particle, slab}: set() places an fcc lattice."""

    def __init__(self, box, particle, slab=(1.0, 1.0, 1.0)):
        self.box, self.particle, self.slab = box, particle, slab

    def set(self):
        r, _ = fcc(self.particle.nparticle, self.box.length, self.slab)
        self.particle.data["position"] = r
        self.particle.data["image"] = np.zeros((self.particle.nparticle, 3), dtype=np.int32)
        return r
