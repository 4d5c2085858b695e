"""halmd.observables.utility.wavevector (utility/wavevector.cpp:39-219)."""
import ctypes

import numpy as np

from .._lib import Handle, call, dptr, f64, out_handle, u64ptr

__all__ = ["wavevector"]


class wavevector(Handle):
    """Wavevector shells |q| ~ wavenumber on the reciprocal box lattice.

    With tolerance/max_count: lattice points picked along coprime Miller
    directions (pick_lattice_points.hpp); without: the dense grid |q| < max k,
    partitioned into shells.
    """

    _destroy = "halmd_observables_utility_wavevector_destroy"

    def __init__(self, box, wavenumber, tolerance=None, max_count=None, dense=False):
        super().__init__()
        k = f64(np.atleast_1d(wavenumber))
        L = f64(box.length)
        self._h = out_handle()
        if dense:
            call("halmd_observables_utility_wavevector_create_dense", dptr(k), len(k), dptr(L), ctypes.byref(self._h))
        else:
            call("halmd_observables_utility_wavevector_create", dptr(k), len(k), dptr(L),
                 float(tolerance if tolerance is not None else 0.01),
                 int(max_count if max_count is not None else 7), ctypes.byref(self._h))
        nq, ns = ctypes.c_size_t(), ctypes.c_size_t()
        call("halmd_observables_utility_wavevector_count", self.handle, ctypes.byref(nq), ctypes.byref(ns))
        q = np.zeros((nq.value, 3))
        shell = np.zeros((ns.value, 2), dtype=np.uint64)
        wn = np.zeros(ns.value)
        call("halmd_observables_utility_wavevector_get", self.handle, dptr(q), u64ptr(shell), dptr(wn))
        self.value = q
        self.shell = shell.astype(np.int64)
        self.wavenumber = wn
