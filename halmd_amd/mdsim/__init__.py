"""halmd.mdsim: box, particle, binning, neighbour lists, core.

Mirrors lua/halmd/mdsim/{box,particle,binning,neighbour,max_displacement,core}.lua.in
over the C-ABI of include/halmd_hip.h.
"""
import ctypes

import numpy as np

from .._lib import Handle, call, dptr, f64, iptr, out_handle, u64ptr, uptr

__all__ = ["box", "particle", "binning", "max_displacement", "neighbour", "core"]


def _matrix(value, n):
    """Species-pair matrix as row-major float64 (scalar or n x n)."""
    return f64(np.broadcast_to(np.asarray(value, dtype=np.float64), (n, n)))


class box(Handle):
    """Cuboid box centred at the origin (lua/halmd/mdsim/box.lua.in; box.cpp:37-62)."""

    _destroy = "halmd_mdsim_box_destroy"

    def __init__(self, length=None, edges=None):
        super().__init__()
        if length is None:
            e = np.asarray(edges, dtype=np.float64)
            if np.any(e != np.diag(np.diag(e))):
                raise ValueError("non-cuboid geometries are not implemented")
            length = np.diag(e)
        L = f64(length)
        if L.shape != (3,):
            raise ValueError("only three-dimensional boxes are supported")
        self._h = out_handle()
        call("halmd_mdsim_box_create", dptr(L), ctypes.byref(self._h))
        self.length = L.copy()

    @property
    def volume(self):
        return float(np.prod(self.length))

    @property
    def lowest_corner(self):
        return -self.length / 2

    @property
    def edges(self):
        return np.diag(self.length)

    def writer(self, file, location):
        """lua/halmd/mdsim/box.lua.in:138-170 (box fixed in time): group
        location/box with attributes dimension and boundary, dataset edges."""
        loc = list(location) + ["box"]
        file.write_attribute_int(loc, "dimension", 3)
        file.write_attribute_strings(loc, "boundary", ["periodic"] * 3)
        w = file.writer(location=list(location), mode="truncate")
        w.on_write(lambda: self.edges, ["box", "edges"])
        w.write()
        return w

    @staticmethod
    def reader(file, location):
        """lua/halmd/mdsim/box.lua.in:190-215: the edge vectors stored at
        location/box/edges."""
        return file.read_dataset(list(location) + ["box", "edges"])


class _data:
    """particle.data[name] get/set in memory order (particle.lua.in data table)."""

    def __init__(self, p):
        self._p = p

    def __getitem__(self, name):
        p = self._p
        n = p.nparticle
        if name in ("position", "velocity", "force"):
            out = np.zeros((n, 3))
            call(f"halmd_mdsim_particle_get_{name}", p.handle, dptr(out))
            return out
        if name == "image":
            out = np.zeros((n, 3), dtype=np.int32)
            call("halmd_mdsim_particle_get_image", p.handle, iptr(out))
            return out
        if name in ("species", "id", "reverse_id"):
            out = np.zeros(n, dtype=np.uint32)
            call(f"halmd_mdsim_particle_get_{name}", p.handle, uptr(out))
            return out
        if name == "mass":
            out = np.zeros(n)
            call("halmd_mdsim_particle_get_mass", p.handle, dptr(out))
            return out
        if name == "potential_energy":
            out = np.zeros(n)
            call("halmd_mdsim_particle_get_potential_energy", p.handle, dptr(out))
            return out
        if name in ("position_parts", "velocity_parts"):
            # the float-float words (hi, lo) per component, exact
            out = np.zeros((n, 3, 2), dtype=np.float32)
            call(f"halmd_mdsim_particle_get_{name}", p.handle, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
            return out
        if name == "position_fixed":
            out = np.zeros((n, 3), dtype=np.uint32)
            call("halmd_mdsim_particle_get_position_fixed", p.handle, out.ctypes.data_as(ctypes.c_void_p))
            return out
        if name == "stress_pot":
            out = np.zeros((n, 6))
            call("halmd_mdsim_particle_get_stress_pot", p.handle, dptr(out))
            return out
        raise KeyError(name)

    def __setitem__(self, name, value):
        p = self._p
        n = p.nparticle
        if name in ("position", "velocity"):
            a = f64(value, (n, 3))
            call(f"halmd_mdsim_particle_set_{name}", p.handle, dptr(a))
        elif name in ("position_parts", "velocity_parts"):
            a = np.ascontiguousarray(value, dtype=np.float32).reshape(n, 3, 2)
            call(f"halmd_mdsim_particle_set_{name}", p.handle, a.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
        elif name == "image":
            a = np.ascontiguousarray(value, dtype=np.int32).reshape(n, 3)
            call("halmd_mdsim_particle_set_image", p.handle, iptr(a))
        elif name == "species":
            a = np.ascontiguousarray(value, dtype=np.uint32).reshape(n)
            call("halmd_mdsim_particle_set_species", p.handle, uptr(a))
        elif name == "mass":
            a = f64(np.broadcast_to(np.asarray(value, dtype=np.float64), (n,)))
            call("halmd_mdsim_particle_set_mass", p.handle, dptr(a))
        else:
            raise KeyError(name)


class particle(Handle):
    """System state of N particles (lua/halmd/mdsim/particle.lua.in:229-290;
    halmd/mdsim/gpu/particle.cpp:55-160)."""

    _destroy = "halmd_mdsim_particle_destroy"

    def __init__(self, dimension=3, particles=None, species=1, memory="gpu", precision="double-single"):
        super().__init__()
        if memory != "gpu":
            raise ValueError("this backend runs on the GPU only (memory='gpu')")
        self._h = out_handle()
        call("halmd_mdsim_particle_create", dimension, int(particles), int(species), ctypes.byref(self._h))
        self.nparticle = int(particles)
        self.nspecies = int(species)
        self.dimension = dimension
        self.data = _data(self)

    def aux_enable(self):
        call("halmd_mdsim_particle_aux_enable", self.handle)

    def set_speculative_force(self, enable=True):
        """Launch the force before the neighbour check is back (default on;
        results are the same either way)."""
        call("halmd_mdsim_particle_set_speculative_force", self.handle, int(bool(enable)))

    def set_fused_step(self, enable=True):
        """Plain NVE Verlet with one pair force: integrate each particle in the
        force kernel as soon as its force is summed (default on; results are
        the same either way)."""
        call("halmd_mdsim_particle_set_fused_step", self.handle, int(bool(enable)))

    def rearrange(self, index):
        idx = np.ascontiguousarray(index, dtype=np.uint32)
        call("halmd_mdsim_particle_rearrange", self.handle, uptr(idx))

    def ordered(self, name):
        """data[name] in particle-ID order (reverse_id maps ID -> memory index)."""
        return self.data[name][self.data["reverse_id"]]

    def reset_order(self):
        """Rearrange the particles into particle-ID order (memory index =
        ID), as a freshly created instance has them: neighbour lists and cell
        order are rebuilt at the next force evaluation.  With the float-float
        state restored exactly, a run continued from here and a run restarted
        from that state then proceed bit for bit alike."""
        self.rearrange(self.data["reverse_id"])


class binning(Handle):
    """Cell lists (lua/halmd/mdsim/binning.lua.in:67-100; gpu/binning.cpp:240-247)."""

    _destroy = "halmd_mdsim_binning_destroy"

    def __init__(self, particle, box, r_cut, skin=0.5, occupancy=0.5):
        super().__init__()
        self._deps = (particle, box)
        rc = _matrix(r_cut, particle.nspecies)
        self._h = out_handle()
        call("halmd_mdsim_binning_create", particle.handle, box.handle, dptr(rc), float(skin), float(occupancy),
             ctypes.byref(self._h))
        self.r_skin = float(skin)
        self.particle = particle

    @property
    def ncell(self):
        out = np.zeros(3, dtype=np.uint32)
        call("halmd_mdsim_binning_ncell", self.handle, uptr(out))
        return out

    def update(self):
        call("halmd_mdsim_binning_update", self.handle)

    def cell_lists(self):
        """(begin, end, index): cell c holds index[begin[c]:end[c]]."""
        nc = int(np.prod(self.ncell))
        b = np.zeros(nc, dtype=np.uint32)
        e = np.zeros(nc, dtype=np.uint32)
        idx = np.zeros(self.particle.nparticle, dtype=np.uint32)
        call("halmd_mdsim_binning_cell_lists", self.handle, uptr(b), uptr(e), uptr(idx))
        return b, e, idx


class max_displacement(Handle):
    """lua/halmd/mdsim/max_displacement.lua.in; gpu/max_displacement.cpp:60-106."""

    _destroy = "halmd_mdsim_max_displacement_destroy"

    def __init__(self, particle, box):
        super().__init__()
        self._deps = (particle, box)
        self._h = out_handle()
        call("halmd_mdsim_max_displacement_create", particle.handle, box.handle, ctypes.byref(self._h))

    def compute(self):
        v = ctypes.c_double()
        call("halmd_mdsim_max_displacement_compute", self.handle, ctypes.byref(v))
        return v.value

    def zero(self):
        call("halmd_mdsim_max_displacement_zero", self.handle)


class neighbour(Handle):
    """Verlet neighbour lists from cell lists
    (lua/halmd/mdsim/neighbour.lua.in:151-306; neighbours/from_binning.cpp:328-339)."""

    _destroy = "halmd_mdsim_neighbours_from_binning_destroy"

    def __init__(self, particle, box, r_cut, skin=0.5, occupancy=0.4, binning=None, displacement=None,
                 algorithm="shared_mem", disable_sorting=False):
        super().__init__()
        if isinstance(particle, (tuple, list)):
            if len(particle) != 2:
                raise ValueError("bad argument 'particle'")
            if particle[0] is not particle[1]:
                self._init_two(particle, box, r_cut, skin, occupancy, binning, displacement)
                return
            particle = particle[0]
        if isinstance(binning, (tuple, list)):
            binning = binning[0]
        if isinstance(displacement, (tuple, list)):
            displacement = displacement[0]
        rc = _matrix(r_cut, particle.nspecies)
        if binning is None:
            binning = globals()["binning"](particle, box, rc, skin=skin)
        # the cell algorithm needs 3 cells of rc + skin per dimension; fall
        # back to all pairs otherwise (neighbour.lua.in:216-226), or when
        # binning=False is passed
        if binning is not False and min(binning.ncell) < 3:
            binning = False
        if displacement is None:
            displacement = max_displacement(particle, box)
        self.binning = binning or None
        self.algorithm = "from_binning" if binning else "from_particle"
        self.displacement = displacement
        self.particle = particle
        self._deps = (particle, box)
        self._h = out_handle()
        if binning:
            call("halmd_mdsim_neighbours_from_binning_create", particle.handle, binning.handle, displacement.handle,
                 box.handle, dptr(rc), float(skin), float(occupancy), ctypes.byref(self._h))
        else:
            call("halmd_mdsim_neighbours_from_particle_create", particle.handle, displacement.handle, box.handle,
                 dptr(rc), float(skin), float(occupancy), ctypes.byref(self._h))
        self.r_skin = float(skin)
        self.sort = None
        if not disable_sorting:
            from .sorts import hilbert
            self.sort = hilbert(particle=particle, box=box)
            self.on_prepend_update(self.sort)

    def _init_two(self, particle, box, r_cut, skin, occupancy, binning, displacement):
        """Lists of particle[0]'s particles holding indices of particle[1]
        (neighbour.lua.in:165-245 with two instances): a binning and a
        displacement module of each (created when not given).  No Hilbert sort
        is attached: each instance's order belongs to its own cell-order
        neighbour module, which rearranges it at every rebuild."""
        p1, p2 = particle
        rc = np.ascontiguousarray(np.broadcast_to(np.atleast_2d(np.asarray(r_cut, dtype=np.float64)),
                                                  (p1.nspecies, p2.nspecies)))
        if binning is None:
            binning = (globals()["binning"](p1, box, rc, skin=skin), globals()["binning"](p2, box, rc, skin=skin))
        if binning is False or min(binning[1].ncell) < 3:
            raise ValueError("neighbour: lists between two particle instances need at least 3 cells of "
                             "rc + skin per dimension")
        if displacement is None:
            displacement = (max_displacement(p1, box), max_displacement(p2, box))
        if binning[0].particle is not p1 or binning[1].particle is not p2:
            raise ValueError("'particle' instances of binning modules do not match with 'particle' argument")
        self.binning = tuple(binning)
        self.displacement = tuple(displacement)
        self.algorithm = "from_binning"
        self.particle = p1
        self.particles = (p1, p2)
        self._deps = (p1, p2, box, self.binning, self.displacement)
        self._h = out_handle()
        call("halmd_mdsim_neighbours_from_binning_create2", p1.handle, p2.handle, binning[0].handle,
             binning[1].handle, displacement[0].handle, displacement[1].handle, box.handle, dptr(rc), float(skin),
             float(occupancy), ctypes.byref(self._h))
        self.r_skin = float(skin)
        self.sort = None

    def on_prepend_update(self, sort):
        call("halmd_mdsim_neighbours_from_binning_on_prepend_update_sort", self.handle, sort.handle)
        self._sort_ref = sort

    def update(self):
        call("halmd_mdsim_neighbours_from_binning_update", self.handle)

    @property
    def size(self):
        v = ctypes.c_uint()
        call("halmd_mdsim_neighbours_from_binning_size", self.handle, ctypes.byref(v))
        return v.value

    @property
    def segments(self):
        """(cells along x, rows along y, workgroup size) of cell-segment lists, zeros otherwise."""
        c, r, t = ctypes.c_uint(), ctypes.c_uint(), ctypes.c_uint()
        call("halmd_mdsim_neighbours_from_binning_segments", self.handle, ctypes.byref(c), ctypes.byref(r),
             ctypes.byref(t))
        return c.value, r.value, t.value

    def set_segments(self, cells, rows, threads):
        """Sets the cell-segment shape; the lists are rebuilt at their next use."""
        call("halmd_mdsim_neighbours_from_binning_set_segments", self.handle, int(cells), int(rows), int(threads))

    @property
    def entry_bytes(self):
        """Bytes per list entry: 2 (16-bit slots, cell-ordered particles) or 4 (indices)."""
        v = ctypes.c_uint()
        call("halmd_mdsim_neighbours_from_binning_entry_bytes", self.handle, ctypes.byref(v))
        return v.value

    @property
    def rebuilds(self):
        v = ctypes.c_uint64()
        call("halmd_mdsim_neighbours_from_binning_rebuilds", self.handle, ctypes.byref(v))
        return v.value

    @property
    def reorders(self):
        """(merged, sorted): how the rebuilds put the particles into cell order."""
        m, s = ctypes.c_uint64(), ctypes.c_uint64()
        call("halmd_mdsim_neighbours_from_binning_reorders", self.handle, ctypes.byref(m), ctypes.byref(s))
        return m.value, s.value

    def counts(self):
        out = np.zeros(self.particle.nparticle, dtype=np.uint32)
        call("halmd_mdsim_neighbours_from_binning_counts", self.handle, uptr(out))
        return out

    def lists(self):
        """(counts, entries): entries[i, :counts[i]] are the neighbours of particle i (memory order)."""
        n = self.particle.nparticle
        stride = self.size
        counts = np.zeros(n, dtype=np.uint32)
        entries = np.zeros((n, stride), dtype=np.uint32)
        call("halmd_mdsim_neighbours_from_binning_lists", self.handle, uptr(counts), uptr(entries), stride)
        return counts, entries


class core(Handle):
    """MD step sequencing (lua/halmd/mdsim/core.lua.in; core.cpp:30-63)."""

    _destroy = "halmd_mdsim_core_destroy"

    def __init__(self, particle):
        super().__init__()
        self.particle = particle
        self._h = out_handle()
        call("halmd_mdsim_core_create", particle.handle, ctypes.byref(self._h))
        self.integrator = None
        self.sort = None

    def set_integrator(self, integrator):
        call("halmd_mdsim_core_set_integrator", self.handle, integrator.handle)
        self.integrator = integrator
        self.integrators = [integrator]

    def add_integrator(self, integrator):
        """Integrator of a further particle instance: a step integrates every
        instance, then finalizes every instance (core.cpp:30-63)."""
        call("halmd_mdsim_core_add_integrator", self.handle, integrator.handle)
        self.integrators.append(integrator)

    def set_sort(self, sort, interval):
        call("halmd_mdsim_core_set_sort", self.handle, sort.handle if sort else None, int(interval))
        self.sort = sort

    def mdstep(self):
        call("halmd_mdsim_core_mdstep", self.handle)

    @property
    def step(self):
        v = ctypes.c_uint64()
        call("halmd_mdsim_core_step", self.handle, ctypes.byref(v))
        return v.value

    @step.setter
    def step(self, value):
        call("halmd_mdsim_core_set_step", self.handle, int(value))

    def tune(self, steps=40):
        """Choose the cell-segment shape of the pair force by timing `steps`
        MD steps per candidate shape from one snapshot of the state (restored
        afterwards): the project's autotuner of the step the run takes."""
        call("halmd_mdsim_core_tune", self.handle, int(steps))

    @property
    def time(self):
        """Simulation time of the current step (halmd/mdsim/clock.hpp): the
        step count times the integrator's timestep."""
        return self.step * (self.integrator.timestep if self.integrator is not None else 0.0)


from . import potentials, forces, integrators, sorts, positions, velocities  # noqa: E402,F401
