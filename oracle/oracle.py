"""ctypes front-end of the CPU oracle (oracle/oracle.cpp) — TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this
module, as a checker or as the timed CPU baseline.  Nothing in the product
package (halmd_amd/) imports it.

Functions mirror the reference host backend; see the header of oracle.cpp for
the reference file:line each restatement follows and for how each is pinned.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

_d = ctypes.POINTER(ctypes.c_double)
_f = ctypes.POINTER(ctypes.c_float)
_u = ctypes.POINTER(ctypes.c_uint)
_i = ctypes.POINTER(ctypes.c_int)
_l = ctypes.POINTER(ctypes.c_long)
_vp = ctypes.c_void_p

POT_LJ, POT_MORSE, POT_POWER_LAW, POT_MIE = 0, 1, 2, 3
TRUNCATIONS = {"sharp": 0, "shifted": 1, "force_shifted": 2, "smooth_r4": 3}
INT_VERLET, INT_NHC, INT_ANDERSEN = 0, 1, 2


def build():
    """Compile the oracle with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        sig = {
            "oracle_potential_create": (_vp, [ctypes.c_int, ctypes.c_int, ctypes.c_int, _d, _d, _d, _d, _u, _u, _vp,
                                              _d, ctypes.c_double]),
            "oracle_potential_destroy": (None, [_vp]),
            "oracle_potential_eval": (None, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _d, _d, _d]),
            "oracle_potential_cutoffs": (None, [_vp, _d, _d, _d]),
            "oracle_binning_ncell": (None, [_d, ctypes.c_double, ctypes.c_double, _i]),
            "oracle_binning_cell": (None, [ctypes.c_int, _d, _d, ctypes.c_double, ctypes.c_double, _u]),
            "oracle_neighbour_half": (ctypes.c_long, [ctypes.c_int, _d, _i, _d, _vp, ctypes.c_double, ctypes.c_long, _l, _u]),
            "oracle_pair_force": (None, [ctypes.c_int, _d, _i, _d, _vp, ctypes.c_double, ctypes.c_double, _d, _d, _d]),
            "oracle_pair_force_brute": (None, [ctypes.c_int, _d, _i, _d, _vp, _d, _d, _d]),
            "oracle_cutoff_band": (ctypes.c_long, [ctypes.c_int, _d, _i, _d, _vp, ctypes.c_double, _d]),
            "oracle_pair_force_scale": (None, [ctypes.c_int, _d, _i, _d, _vp, _d]),
            "oracle_fixed_point": (None, [ctypes.c_int, _d, _d, _u]),
            "oracle_fixed_cell_index": (None, [ctypes.c_int, _u, _u, _u]),
            "oracle_fixed_neighbour_list": (ctypes.c_int, [ctypes.c_int, _u, _u, _u, _f, ctypes.c_int, _f, ctypes.c_int, _u, _u]),
            "oracle_fixed_neighbour_list16": (ctypes.c_int, [ctypes.c_int, _u, _u, _u, ctypes.c_double, ctypes.c_int, _d, ctypes.c_int, _u, _u]),
            "oracle_all_pairs_neighbour_list": (ctypes.c_int, [ctypes.c_int, _u, _u, _f, ctypes.c_int, _f, ctypes.c_int, _u, _u]),
            "oracle_fp32_hilbert_code": (None, [ctypes.c_int, _f, _f, ctypes.c_uint, _u]),
            "oracle_hilbert_depth": (ctypes.c_uint, [_d]),
            "oracle_hilbert_map_f64": (ctypes.c_uint, [ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_uint]),
            "oracle_lattice_fcc": (ctypes.c_double, [ctypes.c_int, _d, _d, _d]),
            "oracle_sim_create": (_vp, [ctypes.c_int, _d, _vp, ctypes.c_double, ctypes.c_double, ctypes.c_int, _d, _d, _d, _i]),
            "oracle_sim_destroy": (None, [_vp]),
            "oracle_sim_set_nhc": (None, [_vp, ctypes.c_double, ctypes.c_double]),
            "oracle_sim_set_andersen": (None, [_vp, ctypes.c_double, ctypes.c_double, ctypes.c_uint64]),
            "oracle_sim_collisions": (ctypes.c_long, [_vp]),
            "oracle_sim_max_displacement": (ctypes.c_double, [_vp]),
            "oracle_sim_run": (ctypes.c_long, [_vp, ctypes.c_long, ctypes.c_long, ctypes.c_int, _d]),
            "oracle_sim_get": (None, [_vp, _d, _d, _l, _d, _d, _d]),
            "oracle_sim_nhc_state": (None, [_vp, _d, _d, _d]),
            "oracle_sim_rebuilds": (ctypes.c_long, [_vp]),
            "oracle_sim_time": (ctypes.c_double, [_vp, ctypes.c_long]),
            "oracle_density_mode": (None, [ctypes.c_int, _d, ctypes.c_int, _d, _d, _d]),
            "oracle_wavevector_shell": (ctypes.c_long, [ctypes.c_int, _d, _d, ctypes.c_double, ctypes.c_uint, ctypes.c_long, _d, _l]),
            "oracle_correlator_create": (_vp, [ctypes.c_long, ctypes.c_long, ctypes.c_uint, ctypes.c_uint, ctypes.c_uint, ctypes.c_long, ctypes.c_int, ctypes.c_int, ctypes.c_int, _l]),
            "oracle_correlator_destroy": (None, [_vp]),
            "oracle_correlator_levels": (ctypes.c_int, [_vp]),
            "oracle_correlator_intervals": (None, [_vp, _l]),
            "oracle_correlator_sample": (None, [_vp, ctypes.c_long, _d, _d, _d]),
            "oracle_correlator_finalise": (None, [_vp]),
            "oracle_correlator_result": (None, [_vp, ctypes.c_int, _l, _d, _d]),
            "oracle_accumulate": (None, [ctypes.c_int, _d, _l, _d, _d]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t) if a is not None else None


def _dd(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Potential:
    """Pair potential with truncation (host semantics, FP64)."""

    def __init__(self, kind, truncation, epsilon, sigma, r_cut_sigma, r_min=None, distortion=None, h=0.005,
                 index=None, index_m=None, index_n=None, r_core_sigma=None):
        """index: power law index n (POT_POWER_LAW); index_m, index_n: Mie indices (POT_MIE);
        r_core_sigma: hard-core radius / sigma (adapters/hard_core)."""
        eps = np.atleast_2d(_dd(epsilon))
        ntype = eps.shape[0]
        sig = _dd(np.broadcast_to(np.atleast_2d(sigma), (ntype, ntype)))
        rcs = _dd(np.broadcast_to(np.atleast_2d(r_cut_sigma), (ntype, ntype)))
        rmin = _dd(np.broadcast_to(np.atleast_2d(r_min if r_min is not None else 1.0), (ntype, ntype)))
        B = _dd(np.broadcast_to(np.atleast_2d(distortion if distortion is not None else 1.0), (ntype, ntype)))
        im = index if index is not None else index_m
        m = np.ascontiguousarray(np.broadcast_to(np.atleast_2d(im if im is not None else 0), (ntype, ntype)),
                                 dtype=np.uint32)
        n = np.ascontiguousarray(np.broadcast_to(np.atleast_2d(index_n if index_n is not None else 0),
                                                 (ntype, ntype)), dtype=np.uint32)
        core = None if r_core_sigma is None else _dd(np.broadcast_to(np.atleast_2d(r_core_sigma), (ntype, ntype)))
        self.ntype = ntype
        self.kind = kind
        self._keep = (eps, sig, rcs, rmin, B, m, n, core)
        self.h = lib().oracle_potential_create(kind, TRUNCATIONS[truncation], ntype, _p(eps, _d), _p(sig, _d),
                                               _p(rmin, _d), _p(B, _d), _p(m, _u), _p(n, _u),
                                               None if core is None else core.ctypes.data_as(ctypes.c_void_p),
                                               _p(rcs, _d), h)
        rc = np.zeros(ntype * ntype)
        ec = np.zeros(ntype * ntype)
        fc = np.zeros(ntype * ntype)
        lib().oracle_potential_cutoffs(self.h, _p(rc, _d), _p(ec, _d), _p(fc, _d))
        self.r_cut = rc.reshape(ntype, ntype)
        self.en_cut = ec.reshape(ntype, ntype)
        self.force_cut = fc.reshape(ntype, ntype)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.oracle_potential_destroy(self.h)

    def __call__(self, rr, a, b):
        rr = _dd(np.atleast_1d(rr))
        f = np.empty_like(rr)
        u = np.empty_like(rr)
        lib().oracle_potential_eval(self.h, a, b, len(rr), _p(rr, _d), _p(f, _d), _p(u, _d))
        return f, u


def lattice_fcc(n, L, slab=(1.0, 1.0, 1.0)):
    L = _dd(L)
    s = _dd(slab)
    out = np.empty((n, 3))
    a = lib().oracle_lattice_fcc(n, _p(L, _d), _p(s, _d), _p(out, _d))
    return out, a


def binning_ncell(L, r_cut_max, skin):
    L = _dd(L)
    nc = np.zeros(3, dtype=np.int32)
    lib().oracle_binning_ncell(_p(L, _d), r_cut_max, skin, _p(nc, _i))
    return nc


def binning_cell(pos, L, r_cut_max, skin):
    pos = _dd(pos)
    L = _dd(L)
    c = np.zeros(len(pos), dtype=np.uint32)
    lib().oracle_binning_cell(len(pos), _p(pos, _d), _p(L, _d), r_cut_max, skin, _p(c, _u))
    return c


def neighbour_half(pos, species, L, pot, skin):
    pos = _dd(pos)
    sp = np.ascontiguousarray(species, dtype=np.int32)
    L = _dd(L)
    n = len(pos)
    cap = n * 200
    off = np.zeros(n + 1, dtype=np.int64)
    lst = np.zeros(cap, dtype=np.uint32)
    tot = lib().oracle_neighbour_half(n, _p(pos, _d), _p(sp, _i), _p(L, _d), pot.h, skin, cap, _p(off, _l), _p(lst, _u))
    assert tot >= 0
    return off, lst[:tot]


def pair_force(pos, species, L, pot, skin=0.3, aux_weight=1.0):
    pos = _dd(pos)
    sp = np.ascontiguousarray(species, dtype=np.int32)
    L = _dd(L)
    n = len(pos)
    f = np.zeros((n, 3))
    u = np.zeros(n)
    s = np.zeros((n, 6))
    lib().oracle_pair_force(n, _p(pos, _d), _p(sp, _i), _p(L, _d), pot.h, skin, aux_weight, _p(f, _d), _p(u, _d), _p(s, _d))
    return f, u, s


def pair_force_brute(pos, species, L, pot):
    pos = _dd(pos)
    sp = np.ascontiguousarray(species, dtype=np.int32)
    L = _dd(L)
    n = len(pos)
    f = np.zeros((n, 3))
    u = np.zeros(n)
    s = np.zeros((n, 6))
    lib().oracle_pair_force_brute(n, _p(pos, _d), _p(sp, _i), _p(L, _d), pot.h, _p(f, _d), _p(u, _d), _p(s, _d))
    return f, u, s


def cutoff_band(pos, species, L, pot, band=1e-6):
    """(number of pairs, per-particle sum of |F_pair|) over the pairs whose
    squared distance is within a relative `band` of the squared cutoff."""
    pos = _dd(pos)
    sp = np.ascontiguousarray(species, dtype=np.int32)
    L = _dd(L)
    jump = np.zeros(len(pos))
    cnt = lib().oracle_cutoff_band(len(pos), _p(pos, _d), _p(sp, _i), _p(L, _d), pot.h, float(band), _p(jump, _d))
    return cnt, jump


def pair_force_scale(pos, species, L, pot):
    """Per-particle sum of the pair-force magnitudes within the cutoff."""
    pos = _dd(pos)
    sp = np.ascontiguousarray(species, dtype=np.int32)
    L = _dd(L)
    out = np.zeros(len(pos))
    lib().oracle_pair_force_scale(len(pos), _p(pos, _d), _p(sp, _i), _p(L, _d), pot.h, _p(out, _d))
    return out


def fixed_point(pos, L):
    """uint32 fixed-point coordinates [n, 3] of FP64 positions (device rule)."""
    pos = _dd(pos)
    L = _dd(L)
    out = np.zeros(pos.shape, dtype=np.uint32)
    lib().oracle_fixed_point(len(pos), _p(pos, _d), _p(L, _d), _p(out, _u))
    return out


def fixed_cell_index(u, ncell):
    u = np.ascontiguousarray(u, dtype=np.uint32)
    nc = np.ascontiguousarray(ncell, dtype=np.uint32)
    out = np.zeros(len(u), dtype=np.uint32)
    lib().oracle_fixed_cell_index(len(u), _p(u, _u), _p(nc, _u), _p(out, _u))
    return out


def fixed_unit(L):
    """float(L) * 2^-32 per dimension (FP32 length of one fixed-point unit)."""
    return np.ldexp(np.asarray(L, dtype=np.float64).astype(np.float32), -32).astype(np.float32)


def all_pairs_neighbour_list(u, species, L, rr_cut_skin, cap):
    """Device from_particle rule: (counts, lists[n, cap], max count), ascending j."""
    u = np.ascontiguousarray(u, dtype=np.uint32)
    sp = np.ascontiguousarray(species, dtype=np.uint32)
    q = np.ascontiguousarray(fixed_unit(L), dtype=np.float32)
    rr = np.ascontiguousarray(rr_cut_skin, dtype=np.float32)
    ntype = rr.shape[0]
    n = len(u)
    counts = np.zeros(n, dtype=np.uint32)
    lst = np.zeros(n * cap, dtype=np.uint32)
    m = lib().oracle_all_pairs_neighbour_list(n, _p(u, _u), _p(sp, _u), _p(q, _f), ntype, _p(rr, _f), cap,
                                              _p(counts, _u), _p(lst, _u))
    return counts, lst.reshape(n, cap), m


def fixed_neighbour_list(u, species, ncell, L, rr_cut_skin, cap):
    u = np.ascontiguousarray(u, dtype=np.uint32)
    sp = np.ascontiguousarray(species, dtype=np.uint32)
    nc = np.ascontiguousarray(ncell, dtype=np.uint32)
    q = np.ascontiguousarray(fixed_unit(L), dtype=np.float32)
    rr = np.ascontiguousarray(rr_cut_skin, dtype=np.float32)
    ntype = rr.shape[0]
    n = len(u)
    counts = np.zeros(n, dtype=np.uint32)
    lst = np.zeros(n * cap, dtype=np.uint32)
    m = lib().oracle_fixed_neighbour_list(n, _p(u, _u), _p(sp, _u), _p(nc, _u), _p(q, _f), ntype, _p(rr, _f), cap,
                                          _p(counts, _u), _p(lst, _u))
    return counts, lst.reshape(n, cap), m


def staged_neighbour_list(u, species, ncell, L, r_cut_skin, cap):
    """Lists of the cell-segment build: the 16-bit integer rule in a cubic box,
    the FP32 rule of fixed_neighbour_list otherwise.  r_cut_skin: ntype x ntype
    distances rc + skin (FP64).  (counts, lists[n, cap], max count)."""
    L = np.broadcast_to(np.asarray(L, dtype=np.float64), (3,))
    rcs = np.ascontiguousarray(np.atleast_2d(r_cut_skin), dtype=np.float64)
    if not (L[0] == L[1] == L[2]):
        return fixed_neighbour_list(u, species, ncell, L, (rcs ** 2).astype(np.float32), cap)
    u = np.ascontiguousarray(u, dtype=np.uint32)
    sp = np.ascontiguousarray(species, dtype=np.uint32)
    nc = np.ascontiguousarray(ncell, dtype=np.uint32)
    n = len(u)
    counts = np.zeros(n, dtype=np.uint32)
    lst = np.zeros(n * cap, dtype=np.uint32)
    m = lib().oracle_fixed_neighbour_list16(n, _p(u, _u), _p(sp, _u), _p(nc, _u), float(L[0]), rcs.shape[0],
                                            _p(rcs, _d), cap, _p(counts, _u), _p(lst, _u))
    return counts, lst.reshape(n, cap), m


def fp32_hilbert_code(pos4, L, depth):
    pos4 = np.ascontiguousarray(pos4, dtype=np.float32)
    Lf = np.ascontiguousarray(L, dtype=np.float32)
    out = np.zeros(len(pos4), dtype=np.uint32)
    lib().oracle_fp32_hilbert_code(len(pos4), _p(pos4, _f), _p(Lf, _f), depth, _p(out, _u))
    return out


def hilbert_depth(L):
    L = _dd(L)
    return int(lib().oracle_hilbert_depth(_p(L, _d)))


def hilbert_map_f64(x, y, z, depth):
    return int(lib().oracle_hilbert_map_f64(x, y, z, depth))


class Simulation:
    """FP64 host MD (binning + half lists + pair_trunc + verlet / NHC)."""

    def __init__(self, pos, vel, species, mass, L, pot, skin, dt, integrator=INT_VERLET, temperature=None, nu=None,
                 rate=None, seed=0):
        self.pot = pot
        pos = _dd(pos)
        vel = _dd(vel)
        m = _dd(mass)
        sp = np.ascontiguousarray(species, dtype=np.int32)
        L = _dd(L)
        self.n = len(pos)
        self.h = lib().oracle_sim_create(self.n, _p(L, _d), pot.h, skin, dt, integrator, _p(pos, _d), _p(vel, _d),
                                         _p(m, _d), _p(sp, _i))
        if integrator == INT_NHC:
            lib().oracle_sim_set_nhc(self.h, temperature, nu)
        elif integrator == INT_ANDERSEN:
            lib().oracle_sim_set_andersen(self.h, temperature, rate, seed)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.oracle_sim_destroy(self.h)

    def run(self, steps, every=0, record0=False):
        nrec = (steps // every + 1) if every else 0
        out = np.zeros((max(nrec, 1), 5))
        k = lib().oracle_sim_run(self.h, steps, every, int(record0), _p(out, _d))
        return out[:k]

    def time(self, steps):
        return lib().oracle_sim_time(self.h, steps)

    def state(self):
        n = self.n
        pos = np.zeros((n, 3)); vel = np.zeros((n, 3)); img = np.zeros((n, 3), dtype=np.int64)
        f = np.zeros((n, 3)); u = np.zeros(n); s = np.zeros((n, 6))
        lib().oracle_sim_get(self.h, _p(pos, _d), _p(vel, _d), _p(img, _l), _p(f, _d), _p(u, _d), _p(s, _d))
        return dict(position=pos, velocity=vel, image=img, force=f, en_pot=u, stress=s)

    def nhc_state(self):
        xi = np.zeros(2); vxi = np.zeros(2); e = np.zeros(1)
        lib().oracle_sim_nhc_state(self.h, _p(xi, _d), _p(vxi, _d), _p(e, _d))
        return xi, vxi, e[0]

    def rebuilds(self):
        return int(lib().oracle_sim_rebuilds(self.h))

    def collisions(self):
        """Heat-bath collisions so far (Andersen)."""
        return int(lib().oracle_sim_collisions(self.h))

    def max_displacement(self):
        """max_i |r_i - r0_i| since the last neighbour-list rebuild (FP64)."""
        return float(lib().oracle_sim_max_displacement(self.h))


def density_mode(pos, q):
    pos = _dd(pos)
    q = _dd(q)
    re = np.zeros(len(q))
    im = np.zeros(len(q))
    lib().oracle_density_mode(len(pos), _p(pos, _d), len(q), _p(q, _d), _p(re, _d), _p(im, _d))
    return re + 1j * im


def wavevector_shell(wavenumber, L, tolerance, max_count):
    k = _dd(wavenumber)
    L = _dd(L)
    cap = 200000
    out = np.zeros((cap, 3))
    shell = np.zeros((len(k), 2), dtype=np.int64)
    w = lib().oracle_wavevector_shell(len(k), _p(k, _d), _p(L, _d), tolerance, max_count, cap, _p(out, _d), _p(shell, _l))
    assert w >= 0
    return out[:w].copy(), shell


class Correlator:
    """Blocking-scheme MSD / MQD / VACF / ISF correlator driven by explicit samples."""

    def __init__(self, max_interval, resolution, block_size, shift, separation, step0, n, nq=0, shell=None):
        shell = np.zeros((0, 2), dtype=np.int64) if shell is None else np.ascontiguousarray(shell, dtype=np.int64)
        self.nshell = len(shell)
        self.block_size = block_size
        self._shell = shell
        self.h = lib().oracle_correlator_create(max_interval, resolution, block_size, shift, separation, step0, n, nq,
                                                self.nshell, _p(shell, _l))
        self.levels = lib().oracle_correlator_levels(self.h)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.oracle_correlator_destroy(self.h)

    def intervals(self):
        out = np.zeros(self.levels, dtype=np.int64)
        lib().oracle_correlator_intervals(self.h, _p(out, _l))
        return out

    def sample(self, step, pos=None, rho=None, vel=None):
        p = _dd(pos) if pos is not None else None
        v = _dd(vel) if vel is not None else None
        r = None
        if rho is not None:
            r = np.ascontiguousarray(np.stack([np.real(rho), np.imag(rho)], axis=-1), dtype=np.float64)
        lib().oracle_correlator_sample(self.h, step, _p(p, _d), _p(r, _d), _p(v, _d))

    def finalise(self):
        lib().oracle_correlator_finalise(self.h)

    def result(self, which):
        """(count, mean, error of mean); which = 'msd', 'mqd', 'vacf' or 'isf'."""
        code = {"msd": 0, "isf": 1, "mqd": 2, "vacf": 3}[which]
        shape = (self.levels, self.block_size) if code != 1 else (self.levels, self.block_size, self.nshell)
        size = int(np.prod(shape))
        c = np.zeros(size, dtype=np.int64); m = np.zeros(size); e = np.zeros(size)
        lib().oracle_correlator_result(self.h, code, _p(c, _l), _p(m, _d), _p(e, _d))
        return c.reshape(shape), m.reshape(shape), e.reshape(shape)


def accumulate(x):
    x = _dd(x)
    c = np.zeros(1, dtype=np.int64); m = np.zeros(1); v = np.zeros(1)
    lib().oracle_accumulate(len(x), _p(x, _d), _p(c, _l), _p(m, _d), _p(v, _d))
    return int(c[0]), m[0], v[0]
