"""halmd.mdsim.potentials (pair potentials)."""
from . import pair  # noqa: F401
