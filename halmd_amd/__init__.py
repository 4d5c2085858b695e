"""MI355X (gfx950) molecular-dynamics backend with the module API of HAL's MD package.

The hot path (binning, neighbour lists, truncated pair forces, Verlet
integrators, Hilbert sort, reductions, density modes and correlators) runs in
hand-written HIP kernels behind the C-ABI of include/halmd_hip.h.  This
package mirrors the Lua module tree of the reference (halmd.mdsim.*,
halmd.observables.*) on top of that C-ABI.
"""
from ._lib import HalmdError, device_info  # noqa: F401
from . import mdsim, observables  # noqa: F401
