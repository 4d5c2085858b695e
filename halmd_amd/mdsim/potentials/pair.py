"""halmd.mdsim.potentials.pair: lennard_jones, morse, power_law and mie with
truncations.

Mirrors lua/halmd/mdsim/potentials/pair/{lennard_jones,morse,power_law,mie}.lua.in and
adapters/truncate.lua.in over include/halmd_hip.h.  The cutoff of a
truncation is given in units of sigma (r_c = cutoff * sigma), as in the
reference (potentials/pair/truncations/shifted.hpp constructor).
"""
import ctypes

import numpy as np

from ..._lib import Handle, call, dptr, f64, out_handle

__all__ = ["lennard_jones", "morse", "power_law", "mie", "modified_lennard_jones", "TRUNCATIONS"]

TRUNCATIONS = ("sharp", "shifted", "force_shifted", "smooth_r4")


def _species_matrix(value, species=None):
    a = np.atleast_2d(np.asarray(value, dtype=np.float64))
    if species is not None and a.shape != (species, species):
        a = np.broadcast_to(a, (species, species))
    return f64(a)


class _pair_potential(Handle):
    _destroy = "halmd_mdsim_potentials_pair_destroy"
    description = "pair potential"

    def modify(self, modification, radius=None):
        """potential:modify{"hard_core", radius = r_core / sigma}: U(r) -> U(r - r_core)
        (lua/halmd/mdsim/potentials/pair/adapters/modify.lua.in:38-64); before truncate."""
        if isinstance(modification, (tuple, list)):
            modification, radius = modification[0], modification[1] if len(modification) > 1 else radius
        if modification != "hard_core":
            raise ValueError(f"unknown modification {modification!r}")
        if radius is None:
            raise ValueError("hard_core: radius is required")
        rc = _species_matrix(radius, self.species)
        call("halmd_mdsim_potentials_pair_modify_hard_core", self.handle, dptr(rc))
        self.r_core_sigma = rc.copy()
        return self

    def truncate(self, truncation="shifted", cutoff=None, smoothing=0.005):
        """potential:truncate{truncation, cutoff = r_c / sigma [, smoothing = h]}."""
        if isinstance(truncation, (tuple, list)):
            truncation, cutoff = truncation[0], truncation[1] if len(truncation) > 1 else cutoff
        if truncation not in TRUNCATIONS:
            raise ValueError(f"unknown truncation {truncation!r}")
        if cutoff is None:
            raise ValueError("truncate: cutoff is required")
        rc = _species_matrix(cutoff, self.species)
        call("halmd_mdsim_potentials_pair_truncate", self.handle, truncation.encode(), dptr(rc), float(smoothing))
        self.truncation = truncation
        self.r_cut_sigma = rc.copy()
        out = np.zeros(self.species * self.species)
        call("halmd_mdsim_potentials_pair_r_cut", self.handle, dptr(out))
        self.r_cut = out.reshape(self.species, self.species)
        return self

    def __call__(self, rr, a=0, b=0):
        """(fval, U) at squared distances rr, FP64 host evaluation."""
        rr = f64(np.atleast_1d(rr))
        f = np.zeros_like(rr)
        u = np.zeros_like(rr)
        call("halmd_mdsim_potentials_pair_evaluate", self.handle, int(a), int(b), len(rr), dptr(rr), dptr(f), dptr(u))
        return f, u


class lennard_jones(_pair_potential):
    """U(r) = 4 eps ((sigma/r)^12 - (sigma/r)^6) (lennard_jones.cpp:45-71)."""

    description = "Lennard-Jones potential"

    def __init__(self, epsilon=1.0, sigma=1.0, species=None):
        super().__init__()
        eps = _species_matrix(epsilon)
        n = species or eps.shape[0]
        eps = _species_matrix(epsilon, n)
        sig = _species_matrix(sigma, n)
        self.species = n
        self.epsilon, self.sigma = eps.copy(), sig.copy()
        self._h = out_handle()
        call("halmd_mdsim_potentials_pair_lennard_jones_create", n, n, dptr(eps), dptr(sig), ctypes.byref(self._h))


class morse(_pair_potential):
    """U(r) = eps/(2B^2-1) [exp(-2(r-r_min)B/sigma) - 2B^2 exp(-(r-r_min)/(B sigma))]
    (morse.cpp:41-61, lua/halmd/mdsim/potentials/pair/morse.lua.in:38-52)."""

    description = "Morse potential"

    def __init__(self, epsilon=1.0, sigma=1.0, r_min=1.0, distortion=1.0, species=None):
        super().__init__()
        n = species or _species_matrix(epsilon).shape[0]
        eps = _species_matrix(epsilon, n)
        sig = _species_matrix(sigma, n)
        rmin = _species_matrix(r_min, n)
        B = _species_matrix(distortion, n)
        self.species = n
        self.epsilon, self.sigma, self.r_min, self.distortion = eps.copy(), sig.copy(), rmin.copy(), B.copy()
        self._h = out_handle()
        call("halmd_mdsim_potentials_pair_morse_create", n, n, dptr(eps), dptr(sig), dptr(rmin), dptr(B),
             ctypes.byref(self._h))


def _index_matrix(value, n):
    a = np.atleast_2d(np.asarray(value))
    if a.shape != (n, n):
        a = np.broadcast_to(a, (n, n))
    if np.any(a != np.round(a)) or np.any(a < 0):
        raise ValueError("power law indices must be non-negative integers")
    return np.ascontiguousarray(a, dtype=np.uint32)


class power_law(_pair_potential):
    """U(r) = eps (sigma/r)^n with integer index n per species pair
    (lua/halmd/mdsim/potentials/pair/power_law.lua.in:130-180; default index 12)."""

    description = "power law potential"

    def __init__(self, epsilon=1.0, sigma=1.0, index=12, species=None):
        super().__init__()
        n = species or max(_species_matrix(epsilon).shape[0], np.atleast_2d(index).shape[0])
        eps = _species_matrix(epsilon, n)
        sig = _species_matrix(sigma, n)
        idx = _index_matrix(index, n)
        self.species = n
        self.epsilon, self.sigma, self.index = eps.copy(), sig.copy(), idx.copy()
        self._h = out_handle()
        call("halmd_mdsim_potentials_pair_power_law_create", n, n, dptr(eps), dptr(sig),
             idx.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), ctypes.byref(self._h))


class mie(_pair_potential):
    """U(r) = C(m, n) eps [(sigma/r)^m - (sigma/r)^n], C = m/(m-n) (m/n)^(n/(m-n)),
    even indices m > n (lua/halmd/mdsim/potentials/pair/mie.lua.in:140-200:
    index_repulsion m, index_attraction n; mie(m=12, n=6) is Lennard-Jones)."""

    description = "Mie potential"

    def __init__(self, epsilon=1.0, sigma=1.0, index_repulsion=12, index_attraction=6, species=None):
        super().__init__()
        n = species or max(_species_matrix(epsilon).shape[0], np.atleast_2d(index_repulsion).shape[0])
        eps = _species_matrix(epsilon, n)
        sig = _species_matrix(sigma, n)
        m_ = _index_matrix(index_repulsion, n)
        n_ = _index_matrix(index_attraction, n)
        self.species = n
        self.epsilon, self.sigma = eps.copy(), sig.copy()
        self.index_repulsion, self.index_attraction = m_.copy(), n_.copy()
        self._h = out_handle()
        call("halmd_mdsim_potentials_pair_mie_create", n, n, dptr(eps), dptr(sig),
             m_.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), n_.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
             ctypes.byref(self._h))


# deprecated alias of the reference (lua/halmd/mdsim/potentials/pair/modified_lennard_jones.lua.in:35-39)
modified_lennard_jones = mie
