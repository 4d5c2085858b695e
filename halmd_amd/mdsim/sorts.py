"""halmd.mdsim.sorts.hilbert (lua/halmd/mdsim/sorts/hilbert.lua.in; gpu/sorts/hilbert.cpp:129-156)."""
import ctypes

from .._lib import Handle, call, out_handle

__all__ = ["hilbert"]


class hilbert(Handle):
    """Reorder particles along a 3D Hilbert curve (stable radix sort of the codes)."""

    _destroy = "halmd_mdsim_sorts_hilbert_destroy"

    def __init__(self, particle, box):
        super().__init__()
        self._deps = (particle, box)
        self._h = out_handle()
        call("halmd_mdsim_sorts_hilbert_create", particle.handle, box.handle, ctypes.byref(self._h))

    def order(self):
        call("halmd_mdsim_sorts_hilbert_order", self.handle)

    @property
    def depth(self):
        v = ctypes.c_uint()
        call("halmd_mdsim_sorts_hilbert_depth", self.handle, ctypes.byref(v))
        return v.value
