"""Pins the CPU oracle (oracle/oracle.cpp) before it is trusted as a checker.

Each check is chosen so that a plausible mistake in the restatement (a dropped
term, a wrong sign or index, a transposed operand) fails: golden tables from
the reference's own unit tests, brute force on small inputs, conservation
laws and closed forms.
"""
import json
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
EPS = np.finfo(np.float64).eps


def _close_fraction(x, ref, tol):
    # BOOST_CHECK_CLOSE_FRACTION (strong): relative to both operands
    if ref == 0.0 or x == 0.0:
        return x == ref
    d = abs(x - ref)
    return d <= tol * abs(ref) and d <= tol * abs(x)


# ----------------------------------------------------------------------------
# potentials: reference golden tables
# ----------------------------------------------------------------------------
@pytest.mark.parametrize("truncation", ["sharp", "shifted", "force_shifted", "smooth_r4"])
def test_lennard_jones_tables(oracle, truncation):
    g = json.load(open(os.path.join(GOLDEN, "potential_tables.json")))["lennard_jones"]
    pot = oracle.Potential(oracle.POT_LJ, truncation, g["epsilon"], g["sigma"], g["r_cut_sigma"], h=g["h"])
    tol = g["tolerance_eps"] * EPS
    for pair, (a, b) in (("aa", (0, 0)), ("ab", (0, 1)), ("bb", (1, 1))):
        for r, fref, uref in g["tables"][truncation][pair]:
            f, u = pot(r * r, a, b)
            assert _close_fraction(f[0], fref, tol), (truncation, pair, r, f[0], fref)
            assert _close_fraction(u[0], uref, tol), (truncation, pair, r, u[0], uref)


def test_morse_tables(oracle):
    g = json.load(open(os.path.join(GOLDEN, "potential_tables.json")))["morse"]
    pot = oracle.Potential(oracle.POT_MORSE, g["truncation"], g["epsilon"], g["sigma"], g["r_cut_sigma"],
                           r_min=g["r_min"], distortion=g["distortion"])
    for pair, (a, b) in (("aa", (0, 0)), ("ab", (0, 1)), ("bb", (1, 1))):
        tol = g["tolerance_eps"][pair] * EPS
        for r, fref, uref in g["tables"][pair]:
            f, u = pot(r * r, a, b)
            assert _close_fraction(f[0], fref, 2 * tol), (pair, r, f[0], fref)
            assert _close_fraction(u[0], uref, tol), (pair, r, u[0], uref)


def test_power_law_tables(oracle):
    g = json.load(open(os.path.join(GOLDEN, "potential_tables.json")))["power_law"]
    pot = oracle.Potential(oracle.POT_POWER_LAW, g["truncation"], g["epsilon"], g["sigma"], g["r_cut_sigma"],
                           index=g["index"])
    for pair, (a, b) in (("aa", (0, 0)), ("ab", (0, 1)), ("bb", (1, 1))):
        tol = (1 + g["index"][a][b] // 2) * EPS
        for r, fref, uref in g["tables"][pair]:
            f, u = pot(r * r, a, b)
            assert _close_fraction(f[0], fref, tol), (pair, r, f[0], fref)
            assert _close_fraction(u[0], uref, tol), (pair, r, u[0], uref)


def test_mie_tables(oracle):
    g = json.load(open(os.path.join(GOLDEN, "potential_tables.json")))["mie"]
    pot = oracle.Potential(oracle.POT_MIE, g["truncation"], g["epsilon"], g["sigma"], g["r_cut_sigma"],
                           index_m=g["index_m"], index_n=g["index_n"])
    tol = g["tolerance_eps"] * EPS
    for pair, (a, b) in (("aa", (0, 0)), ("ab", (0, 1)), ("bb", (1, 1))):
        for r, fref, uref in g["tables"][pair]:
            f, u = pot(r * r, a, b)
            assert _close_fraction(f[0], fref, tol), (pair, r, f[0], fref)
            assert _close_fraction(u[0], uref, tol), (pair, r, u[0], uref)


def test_power_law_hard_core_tables(oracle):
    g = json.load(open(os.path.join(GOLDEN, "potential_tables.json")))["power_law_hard_core"]
    pot = oracle.Potential(oracle.POT_POWER_LAW, g["truncation"], g["epsilon"], g["sigma"], g["r_cut_sigma"],
                           index=g["index"], r_core_sigma=g["r_core_sigma"])
    for pair, (a, b) in (("aa", (0, 0)), ("ab", (0, 1)), ("bb", (1, 1))):
        for r, fref, uref in g["tables"][pair]:
            rs = r / g["sigma"][a][b]
            tol = EPS * g["index"][a][b] * (1 + rs / (rs - g["r_core_sigma"][a][b]))
            f, u = pot(r * r, a, b)
            assert _close_fraction(f[0], fref, tol), (pair, r, f[0], fref)
            assert _close_fraction(u[0], uref, tol), (pair, r, u[0], uref)


def test_hard_core_lennard_jones_is_shifted_argument(oracle):
    # hard core over Lennard-Jones: U(r) = U_LJ(r - r_core), -U'(r)/r by
    # central differences
    hc = oracle.Potential(oracle.POT_LJ, "sharp", 1.0, 1.0, 3.0, r_core_sigma=0.4)
    lj = oracle.Potential(oracle.POT_LJ, "sharp", 1.0, 1.0, 3.0)
    for r in (1.1, 1.3, 1.6, 2.2, 2.9):
        _, u = hc(r * r, 0, 0)
        _, ul = lj((r - 0.4) ** 2, 0, 0)
        assert u[0] == ul[0]
        h = 1e-6
        _, up = hc((r + h) ** 2, 0, 0)
        _, um = hc((r - h) ** 2, 0, 0)
        f, _ = hc(r * r, 0, 0)
        assert abs(f[0] - (-(up[0] - um[0]) / (2 * h) / r)) < 1e-6 * max(1.0, abs(f[0]))


@pytest.mark.parametrize("kind,args", [("power_law", dict(index=7)), ("mie", dict(index_m=10, index_n=4))])
def test_power_potentials_derivative_and_mie_limit(oracle, kind, args):
    # fval = -U'(r)/r by central differences (odd power law index: the sqrt
    # branch); Mie with m = 12, n = 6 is Lennard-Jones (C(12, 6) = 4)
    k = oracle.POT_POWER_LAW if kind == "power_law" else oracle.POT_MIE
    pot = oracle.Potential(k, "sharp", 0.7, 1.3, 3.0, **args)
    for r in (0.9, 1.1, 1.5, 2.5, 3.5):
        h = 1e-6
        _, up = pot((r + h) ** 2, 0, 0)
        _, um = pot((r - h) ** 2, 0, 0)
        f, _ = pot(r * r, 0, 0)
        assert abs(f[0] - (-(up[0] - um[0]) / (2 * h) / r)) < 1e-6 * max(1.0, abs(f[0]))
    mie = oracle.Potential(oracle.POT_MIE, "shifted", 0.7, 1.3, 2.5, index_m=12, index_n=6)
    lj = oracle.Potential(oracle.POT_LJ, "shifted", 0.7, 1.3, 2.5)
    rr = np.linspace(1.0, 3.2, 23) ** 2
    fm, um = mie(rr, 0, 0)
    fl, ul = lj(rr, 0, 0)
    assert np.allclose(fm, fl, rtol=1e-13) and np.allclose(um, ul, rtol=1e-12, atol=1e-15)


def test_morse_derivative(oracle):
    # fval = -U'(r)/r by central differences (the check the reference's
    # compute_potential_reference.py makes), for the Morse parameters of the
    # batched workload: D=1, a=6/sigma (sigma_M = 1/6), r0 = 1, B = 1
    pot = oracle.Potential(oracle.POT_MORSE, "shifted", 1.0, 1.0 / 6, 15.0, r_min=1.0, distortion=1.0)
    for r in (0.8, 0.95, 1.0, 1.1, 1.5, 2.2):
        h = 1e-6
        _, up = pot((r + h) ** 2, 0, 0)
        _, um = pot((r - h) ** 2, 0, 0)
        f, _ = pot(r * r, 0, 0)
        assert abs(f[0] - (-(up[0] - um[0]) / (2 * h) / r)) < 1e-6 * max(1.0, abs(f[0]))


# ----------------------------------------------------------------------------
# helper configurations
# ----------------------------------------------------------------------------
def _liquid(oracle, n, rho, seed, jitter=0.2, species_frac=None):
    L = np.full(3, (n / rho) ** (1 / 3))
    pos, a = oracle.lattice_fcc(n, L)
    rng = np.random.default_rng(seed)
    pos = pos + rng.uniform(-jitter, jitter, pos.shape) * a / 4
    pos = (pos + L / 2) % L - L / 2
    species = np.zeros(n, dtype=np.int32)
    if species_frac is not None:
        nb = int(round(species_frac * n))
        species[rng.permutation(n)[:nb]] = 1
    return pos, L, species, rng


def _ka_potential(oracle):
    eps = [[1.0, 1.5], [1.5, 0.5]]
    sig = [[1.0, 0.8], [0.8, 0.88]]
    return oracle.Potential(oracle.POT_LJ, "shifted", eps, sig, 2.5)


# ----------------------------------------------------------------------------
# forces
# ----------------------------------------------------------------------------
@pytest.mark.parametrize("truncation", ["shifted", "force_shifted", "smooth_r4", "sharp"])
def test_force_half_list_equals_brute_force(oracle, truncation):
    pos, L, species, _ = _liquid(oracle, 864, 0.8442, 1)
    pot = oracle.Potential(oracle.POT_LJ, truncation, 1.0, 1.0, 2.5)
    f, u, s = oracle.pair_force(pos, species, L, pot, skin=0.3)
    fb, ub, sb = oracle.pair_force_brute(pos, species, L, pot)
    scale = np.sqrt(np.mean(np.sum(fb ** 2, axis=1)))
    assert np.max(np.abs(f - fb)) < 1e-12 * scale
    assert np.max(np.abs(u - ub)) < 1e-12
    assert np.max(np.abs(s - sb)) < 1e-11
    # Newton's third law
    assert np.max(np.abs(f.sum(axis=0))) < 1e-10 * scale


def test_force_binary_mixture_brute(oracle):
    pos, L, species, _ = _liquid(oracle, 1372, 1.2, 2, species_frac=0.2)
    pot = _ka_potential(oracle)
    f, u, s = oracle.pair_force(pos, species, L, pot, skin=0.3)
    fb, ub, sb = oracle.pair_force_brute(pos, species, L, pot)
    scale = np.sqrt(np.mean(np.sum(fb ** 2, axis=1)))
    assert np.max(np.abs(f - fb)) < 1e-12 * scale
    assert np.max(np.abs(u - ub)) < 1e-12


def test_force_is_minus_energy_gradient(oracle):
    # smooth_r4 makes U C^2, so central differences converge cleanly
    pos, L, species, rng = _liquid(oracle, 500, 0.8442, 3)
    pot = oracle.Potential(oracle.POT_LJ, "smooth_r4", 1.0, 1.0, 2.5, h=0.005)
    f, u, _ = oracle.pair_force_brute(pos, species, L, pot)
    h = 1e-6
    for i in rng.choice(len(pos), 5, replace=False):
        for k in range(3):
            p1 = pos.copy(); p1[i, k] += h
            p2 = pos.copy(); p2[i, k] -= h
            _, u1, _ = oracle.pair_force_brute(p1, species, L, pot)
            _, u2, _ = oracle.pair_force_brute(p2, species, L, pot)
            fd = -(u1.sum() - u2.sum()) / (2 * h)
            assert abs(fd - f[i, k]) < 1e-5 * max(1.0, abs(f[i, k]))


def test_force_two_particles_closed_form(oracle):
    # two particles at distance r along x: F = fval * r, U_i = U/2 each
    L = np.full(3, 20.0)
    pot = oracle.Potential(oracle.POT_LJ, "shifted", 1.0, 1.0, 2.5)
    for r in (0.95, 1.12, 1.7, 2.4):
        pos = np.array([[0.0, 0.0, 0.0], [r, 0.0, 0.0]])
        f, u, s = oracle.pair_force_brute(pos, [0, 0], L, pot)
        fval, uu = pot(r * r, 0, 0)
        assert np.allclose(f[0], [fval[0] * (-r), 0, 0], rtol=1e-14)
        assert np.allclose(f[1], [fval[0] * r, 0, 0], rtol=1e-14)
        assert np.allclose(u, [uu[0] / 2, uu[0] / 2], rtol=1e-14)
        assert np.isclose(s[0, 0], 0.5 * fval[0] * r * r, rtol=1e-14)


def test_force_small_box_is_brute_force(oracle):
    # a box under 3 cells of rc + skin: the cell lists would visit a cell
    # twice through the wrap; the oracle sums over all pairs instead
    rng = np.random.default_rng(7)
    n = 400
    L = np.full(3, (n / 1.1) ** (1 / 3))
    assert L[0] < 3 * 2.8
    pos = rng.uniform(-L / 2, L / 2, (n, 3))
    species = (rng.random(n) < 0.2).astype(np.int32)
    pot = _ka_potential(oracle)
    f, u, s = oracle.pair_force(pos, species, L, pot, skin=0.3, aux_weight=0.5)
    fb, ub, sb = oracle.pair_force_brute(pos, species, L, pot)
    assert np.array_equal(f, fb)
    assert np.allclose(u, 0.5 * ub, rtol=1e-15) and np.allclose(s, 0.5 * sb, rtol=1e-15)


def test_pair_force_scale(oracle):
    # closed form on three particles, then a brute-force sweep on a mixture
    L = np.full(3, 20.0)
    pot = oracle.Potential(oracle.POT_LJ, "shifted", 1.0, 1.0, 2.5)
    pos = np.array([[0.0, 0, 0], [1.1, 0, 0], [0, 2.0, 0]])

    def mag(r):
        return abs(48 * r ** -13 - 24 * r ** -7)
    d12 = np.hypot(1.1, 2.0)
    s = oracle.pair_force_scale(pos, np.zeros(3, np.int32), L, pot)
    assert np.allclose(s, [mag(1.1) + mag(2.0), mag(1.1) + mag(d12), mag(2.0) + mag(d12)], rtol=1e-13)
    pos, L, species, _ = _liquid(oracle, 1372, 1.2, 2, species_frac=0.2)
    pot = _ka_potential(oracle)
    s = oracle.pair_force_scale(pos, species, L, pot)
    rc = np.array([[2.5, 2.0], [2.0, 2.2]])
    eps = np.array([[1.0, 1.5], [1.5, 0.5]])
    sig = np.array([[1.0, 0.8], [0.8, 0.88]])
    d = pos[:, None, :] - pos[None, :, :]
    d -= L * np.round(d / L)
    r = np.sqrt(np.sum(d * d, axis=-1))
    np.fill_diagonal(r, np.inf)
    e, sg = eps[species][:, species], sig[species][:, species]
    x = sg / r
    f = np.abs(24 * e / r * (2 * x ** 12 - x ** 6))
    f[r >= rc[species][:, species]] = 0
    assert np.allclose(s, f.sum(axis=1), rtol=1e-12)


def test_cutoff_band_pairs(oracle):
    # closed form: two particles at r_c (1 +- 4e-7) are in a 1e-6 band with
    # |F(r_c)| = |48 r^-13 - 24 r^-7| each; at r_c (1 - 4e-6) they are not
    L = np.full(3, 20.0)
    pot = oracle.Potential(oracle.POT_LJ, "shifted", 1.0, 1.0, 2.5)
    for r, inside in ((2.5 * (1 - 4e-7), True), (2.5 * (1 + 4e-7), True), (2.5 * (1 - 4e-6), False)):
        cnt, jump = oracle.cutoff_band(np.array([[0.0, 0, 0], [0, 0, r]]), np.zeros(2, np.int32), L, pot, 1e-6)
        assert cnt == int(inside)
        f = abs(48 * r ** -13 - 24 * r ** -7) if inside else 0.0
        assert np.allclose(jump, [f, f], rtol=1e-12, atol=0)
    # a liquid: the band pairs are those a brute-force sweep finds, with the
    # per-species cutoffs of a mixture, across the periodic boundary
    pos, L, species, _ = _liquid(oracle, 1372, 1.2, 2, species_frac=0.2)
    pot = _ka_potential(oracle)
    band = 2e-3
    cnt, jump = oracle.cutoff_band(pos, species, L, pot, band)
    rc = np.array([[2.5, 2.0], [2.0, 2.2]])
    d = pos[:, None, :] - pos[None, :, :]
    d -= L * np.round(d / L)
    rr = np.sum(d * d, axis=-1)
    rc2 = rc[species][:, species] ** 2
    hit = np.triu(np.abs(rr - rc2) <= band * rc2, 1)
    assert cnt == hit.sum() and cnt > 10
    deg = hit.sum(axis=0) + hit.sum(axis=1)
    assert np.array_equal(jump > 0, deg > 0)


# ----------------------------------------------------------------------------
# binning and neighbour lists
# ----------------------------------------------------------------------------
@pytest.mark.parametrize("scale", [0.0, 0.5, 0.9])
def test_binning_invariant(oracle, scale):
    # sine-compressed fcc lattice as in test/unit/mdsim/binning.cpp:74-98; every
    # particle is binned exactly once, into floor((r + L/2) * ncell / L) up to the
    # periodic offset of the host rule (index = floor(r/len) mod ncell)
    n = 4000
    L = np.array([17.0, 19.0, 23.0])
    pos, _ = oracle.lattice_fcc(n, L)
    k = 2 * np.pi / L
    pos = pos + (scale / k) * np.sin(k * pos)
    ncell = oracle.binning_ncell(L, 2.5, 0.3)
    assert list(ncell) == [int(L[0] // 2.8), int(L[1] // 2.8), int(L[2] // 2.8)]
    cell = oracle.binning_cell(pos, L, 2.5, 0.3)
    idx = np.floor(pos * ncell / L).astype(np.int64) % ncell
    expect = (idx[:, 2] * ncell[1] + idx[:, 1]) * ncell[0] + idx[:, 0]
    assert np.array_equal(cell, expect)
    assert np.bincount(cell, minlength=int(np.prod(ncell))).sum() == n


def _brute_pairs(pos, L, species, rr_cut_skin, strict=True):
    d = pos[:, None, :] - pos[None, :, :]
    d -= L * np.where(d > L / 2, 1, np.where(d < -L / 2, -1, 0))
    rr = np.sum(d * d, axis=-1)
    cut = rr_cut_skin[species[:, None], species[None, :]]
    m = (rr < cut) if strict else (rr <= cut)
    np.fill_diagonal(m, False)
    return m


def test_neighbour_half_list_is_brute_force_set(oracle):
    pos, L, species, _ = _liquid(oracle, 1372, 1.2, 4, species_frac=0.2)
    pot = _ka_potential(oracle)
    off, lst = oracle.neighbour_half(pos, species, L, pot, 0.3)
    rcs = (pot.r_cut + 0.3) ** 2
    m = _brute_pairs(pos, L, species, rcs)
    got = np.zeros_like(m)
    for i in range(len(pos)):
        js = lst[off[i]:off[i + 1]]
        assert len(np.unique(js)) == len(js)
        got[i, js] = True
    assert not np.any(got & got.T)            # each pair once
    assert np.array_equal(got | got.T, m)     # and every pair


def test_fixed_point_rule(oracle):
    # u = floor(frac(x / L + 1/2) 2^32): closed form on exactly representable
    # inputs, periodicity, and one unit per L / 2^32
    L = np.array([16.0, 32.0, 8.0])
    x = np.array([[-8.0, -16.0, -4.0], [0.0, 0.0, 0.0], [4.0, -8.0, 1.0], [8.0, 16.0, 4.0 - 2.0 ** -30]])
    u = oracle.fixed_point(x, L)
    assert list(u[0]) == [0, 0, 0]
    assert list(u[1]) == [2 ** 31] * 3
    assert list(u[2]) == [3 * 2 ** 30, 2 ** 30, 2 ** 31 + 2 ** 29]
    assert u[3, 0] == 0 and u[3, 1] == 0          # +L/2 wraps to -L/2
    assert u[3, 2] == 2 ** 32 - 1                 # 4 - 2^-30 in a box of 8: half a unit below +L/2
    rng = np.random.default_rng(3)
    L = np.array([107.49, 37.35, 17.0])
    p = rng.uniform(-L / 2, L / 2, (10000, 3))
    u = oracle.fixed_point(p, L).astype(np.float64)
    assert np.all(np.abs(u / 2 ** 32 - (p / L + 0.5)) < 2e-9)
    # differences are minimum-image separations
    i, j = rng.integers(0, 10000, (2, 1000))
    d = (oracle.fixed_point(p, L)[i] - oracle.fixed_point(p, L)[j]).astype(np.int32) * (L / 2 ** 32)
    dr = p[i] - p[j]
    dr -= L * np.rint(dr / L)
    assert np.max(np.abs(d - dr)) < 1e-7


def test_fixed_cell_index_is_floor_rule(oracle):
    # the reference binning invariant cell == floor(r ncell / L), for positions
    # measured from the lower box corner (test/unit/mdsim/binning.cpp:146)
    L = np.array([17.0, 19.0, 23.0])
    pos, _ = oracle.lattice_fcc(4000, L)
    k = 2 * np.pi / L
    pos = pos + (0.9 / k) * np.sin(k * pos)
    ncell = np.floor(L / 2.8).astype(np.uint32)
    u = oracle.fixed_point(pos, L)
    c = oracle.fixed_cell_index(u, ncell)
    idx = np.floor((pos + L / 2) * ncell / L).astype(np.int64) % ncell
    expect = (idx[:, 2] * ncell[1] + idx[:, 1]) * ncell[0] + idx[:, 0]
    assert np.array_equal(c, expect)


def test_cubic_rule_cell_and_all_pairs_agree(oracle):
    # cubic box: the fixed-point-units predicate.  The cell algorithm and the
    # all-pairs algorithm are separate code paths over the same predicate:
    # their sets agree, all-pairs lists ascend, both are symmetric, and both
    # equal the FP64 sets outside a band of rounding width around rc + skin
    pos, L, species, _ = _liquid(oracle, 1500, 0.8442, 7, species_frac=0.3)
    pot = _ka_potential(oracle)
    skin = 0.3
    ncell = np.floor(L / (np.max(pot.r_cut) + skin)).astype(np.uint32)
    rcs = ((pot.r_cut + skin) ** 2).astype(np.float32)
    u = oracle.fixed_point(pos, L)
    c1, l1, _ = oracle.fixed_neighbour_list(u, species, ncell, L, rcs, 256)
    c2, l2, _ = oracle.all_pairs_neighbour_list(u, species, L, rcs, 256)
    assert np.array_equal(c1, c2)
    got = np.zeros((len(u), len(u)), bool)
    for i in range(len(u)):
        assert np.array_equal(np.sort(l1[i, :c1[i]]), l2[i, :c2[i]])
        assert np.all(np.diff(l2[i, :c2[i]].astype(np.int64)) > 0)
        got[i, l2[i, :c2[i]]] = True
    assert np.array_equal(got, got.T)
    dr = pos[:, None, :] - pos[None, :, :]
    dr -= L * np.rint(dr / L)
    r64 = np.sqrt(np.sum(dr * dr, -1))
    cut = np.sqrt(rcs[species[:, None], species[None, :]].astype(np.float64))
    clear = np.abs(r64 - cut) > 1e-5
    np.fill_diagonal(clear, False)
    assert np.array_equal(got[clear], (r64 < cut)[clear])


def test_fixed_neighbour_list_rules(oracle):
    # device rule: full lists; the set equals the predicate evaluated on all
    # pairs, the lists are symmetric, and entries follow the documented order.
    # A slightly non-cubic box exercises the FP32 physical-units predicate,
    # which numpy evaluates exactly as written (no fused products)
    pos, L, species, _ = _liquid(oracle, 2048, 0.8442, 5, species_frac=0.3)
    L = L * np.array([1.0, 1.0 + 1e-3, 1.0 - 1e-3])
    pos = pos * (L / L[0])
    pot = _ka_potential(oracle)
    skin = 0.3
    ncell = np.floor(L / (np.max(pot.r_cut) + skin)).astype(np.uint32)
    rcs = ((pot.r_cut + skin) ** 2).astype(np.float32)
    u = oracle.fixed_point(pos, L)
    counts, lst, m = oracle.fixed_neighbour_list(u, species, ncell, L, rcs, 256)
    assert m <= 256
    q = oracle.fixed_unit(L)
    assert not (q[0] == q[1] == q[2])
    D = (u[:, None, :] - u[None, :, :]).astype(np.int32)
    d = D.astype(np.float32) * q
    rr = (d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2]
    expect = rr <= rcs[species[:, None], species[None, :]]
    np.fill_diagonal(expect, False)
    got = np.zeros_like(expect)
    for i in range(len(u)):
        got[i, lst[i, :counts[i]]] = True
        assert len(np.unique(lst[i, :counts[i]])) == counts[i]
    assert np.array_equal(got, expect)
    assert np.array_equal(got, got.T)
    # and equals the FP64 set except for pairs within rounding of rc + skin
    dr = pos[:, None, :] - pos[None, :, :]
    dr -= L * np.rint(dr / L)
    r64 = np.sqrt(np.sum(dr * dr, -1))
    cut = np.sqrt(rcs[species[:, None], species[None, :]].astype(np.float64))
    clear = np.abs(r64 - cut) > 1e-5
    np.fill_diagonal(clear, False)
    assert np.array_equal(got[clear], (r64 < cut)[clear])
    # order: rows (dz, dy) lexicographic, cells x-1, x, x+1 in each row,
    # ascending particle index in each cell
    cell = oracle.fixed_cell_index(u, ncell)
    cxyz = np.stack([cell % ncell[0], (cell // ncell[0]) % ncell[1], cell // ncell[0] // ncell[1]], 1).astype(int)
    for i in range(0, len(u), 97):
        js = lst[i, :counts[i]]
        off = (cxyz[js] - cxyz[i] + 1) % ncell.astype(int) - 1
        keys = [((o[2], o[1], o[0]), j) for o, j in zip(off, js)]
        assert keys == sorted(keys)
        assert np.all(np.abs(off) <= 1)


# ----------------------------------------------------------------------------
# Hilbert curve
# ----------------------------------------------------------------------------
@pytest.mark.parametrize("depth", [1, 2, 3, 4])
def test_hilbert_curve_is_a_continuous_bijection(oracle, depth):
    # cells of the 2^depth grid (in fractional coordinates frac(r/L), which the
    # map takes first) map one-to-one onto codes 0 .. 8^depth-1, and consecutive
    # codes are face-adjacent cells: the defining property of a Hilbert curve,
    # which any wrong vertex-table entry breaks
    n = 1 << depth
    codes = {}
    for x in range(n):
        for y in range(n):
            for z in range(n):
                c = oracle.hilbert_map_f64((x + 0.5) / n, (y + 0.5) / n, (z + 0.5) / n, depth)
                codes[c] = (x, y, z)
    assert sorted(codes) == list(range(n ** 3))
    for c in range(n ** 3 - 1):
        a, b = np.array(codes[c]), np.array(codes[c + 1])
        assert np.sum(np.abs(a - b)) == 1


def test_hilbert_fp32_code_snaps_to_cell(oracle):
    L = np.array([107.49, 107.49, 107.49])
    depth = oracle.hilbert_depth(L)
    assert depth == 7
    rng = np.random.default_rng(0)
    pos4 = np.zeros((5000, 4), np.float32)
    pos4[:, :3] = rng.uniform(-L / 2, L / 2, (5000, 3))
    code = oracle.fp32_hilbert_code(pos4, L, depth)
    n = 1 << depth
    u = (pos4[:, :3] / L.astype(np.float32)).astype(np.float64)
    cellf = np.floor((u - np.floor(u)) * n)
    for i in range(0, 5000, 50):
        x, y, z = cellf[i]
        ref = oracle.hilbert_map_f64((x + 0.5) / n, (y + 0.5) / n, (z + 0.5) / n, depth)
        assert code[i] == ref


# ----------------------------------------------------------------------------
# lattice
# ----------------------------------------------------------------------------
def test_fcc_lattice(oracle):
    n = 4000
    L = np.full(3, (n / 0.8442) ** (1 / 3))
    pos, a = oracle.lattice_fcc(n, L)
    assert np.isclose(a, L[0] / 10)
    assert np.all(pos >= -L / 2) and np.all(pos < L / 2)
    # fcc: 12 nearest neighbours at a/sqrt(2)
    d = pos[:50, None, :] - pos[None, :, :]
    d -= L * np.rint(d / L)
    r = np.sqrt(np.sum(d * d, -1))
    r[np.arange(50), np.arange(50)] = np.inf
    assert np.allclose(np.sort(r, axis=1)[:, :12], a / np.sqrt(2))
    assert np.all(np.sort(r, axis=1)[:, 12] > a / np.sqrt(2) * 1.2)
    # first site of each unit cell and its x-fastest cell order
    assert np.allclose(pos[0], -L / 2 + a * np.array([0.25, 0.25, 0.25]))
    assert np.allclose(pos[4], -L / 2 + a * np.array([1.25, 0.25, 0.25]))


# ----------------------------------------------------------------------------
# integrators
# ----------------------------------------------------------------------------
def test_verlet_ideal_gas(oracle):
    # test/unit/mdsim/integrators/verlet.cpp:106-140: no interactions, E_kin and
    # momentum exactly conserved, r advances linearly (mod L)
    n = 1000
    L = np.full(3, 40.0)
    rng = np.random.default_rng(7)
    pos = rng.uniform(-L / 2, L / 2, (n, 3))
    vel = rng.normal(0, 1, (n, 3))
    vel -= vel.mean(0)
    pot = oracle.Potential(oracle.POT_LJ, "shifted", 1e-30, 1.0, 2.5)   # negligible forces
    sim = oracle.Simulation(pos, vel, np.zeros(n, np.int32), np.ones(n), L, pot, 0.3, 0.001)
    sim.run(1000)
    st = sim.state()
    expect = pos + vel * 1.0
    unwrapped = st["position"] + st["image"] * L
    assert np.allclose(unwrapped, expect, atol=1e-9)
    assert np.allclose(st["velocity"], vel, atol=1e-12)


def test_nve_energy_conservation(oracle):
    n = 864
    pos, L, species, rng = _liquid(oracle, n, 0.8442, 8, jitter=0.0)
    vel = rng.normal(0, np.sqrt(1.2), (n, 3))
    vel -= vel.mean(0)
    pot = oracle.Potential(oracle.POT_LJ, "shifted", 1.0, 1.0, 2.5)
    sim = oracle.Simulation(pos, vel, species, np.ones(n), L, pot, 0.3, 0.002)
    rec = sim.run(2000, every=50, record0=True)
    etot = rec[:, 1] + rec[:, 2]
    assert np.max(np.abs(etot[5:] - etot[5])) < 2e-3
    assert sim.rebuilds() > 5
    st = sim.state()
    assert np.max(np.abs(st["velocity"].sum(0))) < 1e-9


def test_nhc_conserved_energy(oracle):
    # test/unit/mdsim/integrators/verlet_nvt_hoover.cpp:203-206: the extended
    # energy E_pot + E_kin + E_nhc is conserved; the temperature relaxes to T
    n = 864
    pos, L, species, rng = _liquid(oracle, n, 0.8442, 9, jitter=0.0)
    vel = rng.normal(0, np.sqrt(2.0), (n, 3))
    vel -= vel.mean(0)
    pot = oracle.Potential(oracle.POT_LJ, "shifted", 1.0, 1.0, 2.5)
    sim = oracle.Simulation(pos, vel, species, np.ones(n), L, pot, 0.3, 0.002, integrator=oracle.INT_NHC,
                            temperature=1.0, nu=5.0)
    rec = sim.run(4000, every=50, record0=False)
    econs = rec[:, 1] + rec[:, 2] + rec[:, 4]
    assert np.max(np.abs(econs - econs[0])) < 5e-3
    temp = 2 * rec[-20:, 2] / 3
    assert abs(temp.mean() - 1.0) < 0.1


@pytest.mark.parametrize("temp,duration", [(1.0, 500), (2.5, 100)])
def test_andersen_ideal_gas_statistics(oracle, temp, duration):
    # test/unit/mdsim/integrators/verlet_nvt_andersen.cpp:101-200 (host set-up:
    # 1500 particles, box ratios 1 : 2 : 1.01, rho = 0.3, dt = 0.01, T = 1,
    # collision rate 10, a run of 500 tau sampled every 3 collision times);
    # a second, shorter run at T = 2.5 pins the sqrt(T) scale of the draws
    import andersen_checks as ac
    n = 1500
    L = ac.box_edges(n)
    rng = np.random.default_rng(61)
    pos = rng.uniform(-L / 2, L / 2, (n, 3))
    vel = rng.normal(0, np.sqrt(temp), (n, 3))
    vel -= vel.mean(0)
    pot = oracle.Potential(oracle.POT_LJ, "sharp", 0.0, 1.0, 1.0)   # no interactions
    sim = oracle.Simulation(pos, vel, np.zeros(n, np.int32), np.ones(n), L, pot, 1.0, ac.TIMESTEP,
                            integrator=oracle.INT_ANDERSEN, temperature=temp, rate=ac.RATE, seed=5)
    steps = int(np.ceil(duration / ac.TIMESTEP))
    period = ac.sample_period()
    temps, vcms = [], []
    for k in range(0, steps, period):
        # samples at steps 0, period, 2 period, ... taken after the step (the
        # reference samples after finalize when i % period == 0)
        sim.run(1 if k == 0 else period)
        v = sim.state()["velocity"]
        temps.append(np.mean(np.sum(v ** 2, axis=1)) / 3)
        vcms.append(v.mean(0))
    done = 1 + (len(temps) - 1) * period
    sim.run(steps - done)
    v = sim.state()["velocity"]
    ac.check(temps, vcms, np.mean(np.sum(v ** 2, axis=1)) / 3, v.mean(0), n, collisions=sim.collisions(),
             particle_steps=n * steps, temp=temp)


def test_andersen_without_collisions_is_verlet(oracle):
    # rate 0: no particle ever collides, and the trajectory is velocity Verlet's
    n = 500
    pos, L, species, rng = _liquid(oracle, n, 0.8442, 10, jitter=0.05)
    vel = rng.normal(0, 1.0, (n, 3))
    vel -= vel.mean(0)
    pot = oracle.Potential(oracle.POT_LJ, "shifted", 1.0, 1.0, 2.5)
    a = oracle.Simulation(pos, vel, species, np.ones(n), L, pot, 0.3, 0.002, integrator=oracle.INT_ANDERSEN,
                          temperature=1.0, rate=0.0, seed=3)
    b = oracle.Simulation(pos, vel, species, np.ones(n), L, pot, 0.3, 0.002)
    a.run(50)
    b.run(50)
    assert a.collisions() == 0
    assert np.array_equal(a.state()["velocity"], b.state()["velocity"])
    assert np.array_equal(a.state()["position"], b.state()["position"])


# ----------------------------------------------------------------------------
# observables
# ----------------------------------------------------------------------------
def test_density_modes_perfect_lattice(oracle):
    # test/unit/observables/ssf.cpp:237-242: on a perfect fcc lattice rho_k = N
    # (up to a phase) at reciprocal-lattice vectors, ~0 elsewhere
    n = 4000
    L = np.full(3, 10.0 * 1.5)
    pos, a = oracle.lattice_fcc(n, L)
    b = 2 * np.pi / a
    q_bragg = np.array([[2 * b, 0, 0], [b, b, b], [2 * b, 2 * b, 0]])
    q_off = np.array([[2 * np.pi / L[0], 0, 0], [b, 0, 0], [b, b, 0]])   # box vectors, not fcc Bragg
    rho = oracle.density_mode(pos, q_bragg)
    assert np.allclose(np.abs(rho), n, rtol=0, atol=50 * n * EPS * 100)
    rho0 = oracle.density_mode(pos, q_off)
    assert np.all(np.abs(rho0) < 1e-8 * n)


def test_wavevector_shells(oracle):
    L = np.full(3, 37.35)
    k = np.array([2.0, 4.0, 7.25, 10.0])
    q, shell = oracle.wavevector_shell(k, L, 0.01, 7)
    u = 2 * np.pi / L[0]
    for s, kk in enumerate(k):
        qs = q[shell[s, 0]:shell[s, 1]]
        assert 0 < len(qs) <= 7
        assert np.all(np.abs(np.linalg.norm(qs, axis=1) - kk) <= kk * 0.01 + 1e-12)
        assert np.all(qs >= 0)
        # integer multiples of the reciprocal box vector
        m = qs / u
        assert np.allclose(m, np.rint(m), atol=1e-9)


def test_accumulator(oracle):
    x = np.random.default_rng(1).normal(3.0, 2.0, 10001)
    c, m, v = oracle.accumulate(x)
    assert c == len(x)
    assert np.isclose(m, x.mean(), rtol=1e-14)
    assert np.isclose(v, x.var(), rtol=1e-12)


def test_blocking_scheme_intervals(oracle):
    # resolution 1 step, max lag 1000, block size 10, shift 3: 1, 3, 10, 30, 100, 300, 1000
    c = oracle.Correlator(1000, 1, 10, 3, 10, 0, 4)
    assert list(c.intervals()) == [1, 3, 10, 30, 100, 300, 1000]
    c = oracle.Correlator(1000, 2, 4, 1, 4, 0, 4)
    assert list(c.intervals()) == [2, 8, 32, 128, 512]


def test_msd_ballistic_and_isf_lag0(oracle):
    # ballistic particles: R(t) = R0 + v t, so MSD(lag) = <v^2> lag^2 exactly at
    # every level; ISF at lag 0 is <|rho_k|^2>/N over the shell
    n = 64
    rng = np.random.default_rng(2)
    r0 = rng.uniform(-5, 5, (n, 3))
    v = rng.normal(0, 1, (n, 3))
    dt = 0.01
    q = np.array([[1.0, 0, 0], [0, 1.0, 0]])
    shell = np.array([[0, 2]])
    c = oracle.Correlator(200, 1, 5, 2, 5, 0, n, nq=2, shell=shell)
    rhos = {}
    for step in range(0, 401):
        pos = r0 + v * step * dt
        rho = oracle.density_mode(pos, q)
        rhos[step] = rho
        c.sample(step, pos, rho)
    c.finalise()
    cnt, mean, err = c.result("msd")
    iv = c.intervals()
    v2 = np.mean(np.sum(v * v, 1))
    for lvl in range(len(iv)):
        for lag in range(5):
            if cnt[lvl, lag]:
                assert np.isclose(mean[lvl, lag], v2 * (lag * iv[lvl] * dt) ** 2, rtol=1e-9, atol=1e-15)
    icnt, imean, _ = c.result("isf")
    assert icnt[0, 0, 0] > 0
    # level 0, lag 0 averages |rho_k|^2/N over all processed origins; check bounds
    vals = [np.mean(np.abs(rhos[s]) ** 2) / n for s in rhos]
    assert min(vals) - 1e-12 <= imean[0, 0, 0] <= max(vals) + 1e-12


def test_mqd_ballistic_and_vacf_alternating(oracle):
    # ballistic R(t) = R0 + v t: MQD(lag) = <|v|^4> lag^4; velocities that flip
    # sign every step: VACF(lag) = (-1)^(lag interval) <|v|^2> (pins the sample
    # pairing and index)
    n = 48
    rng = np.random.default_rng(5)
    r0 = rng.uniform(-5, 5, (n, 3))
    v = rng.normal(0, 1, (n, 3))
    dt = 0.01
    c = oracle.Correlator(300, 1, 6, 2, 6, 0, n)
    for step in range(0, 601):
        c.sample(step, pos=r0 + v * step * dt, vel=v * (-1.0) ** step)
    c.finalise()
    iv = c.intervals()
    v2 = np.mean(np.sum(v * v, 1))
    v4 = np.mean(np.sum(v * v, 1) ** 2)
    mc, mm, _ = c.result("mqd")
    vc, vm, _ = c.result("vacf")
    checked = 0
    for lvl in range(len(iv)):
        for lag in range(6):
            if mc[lvl, lag]:
                assert np.isclose(mm[lvl, lag], v4 * (lag * iv[lvl] * dt) ** 4, rtol=1e-9, atol=1e-15)
            if vc[lvl, lag]:
                assert np.isclose(vm[lvl, lag], (-1.0) ** (lag * iv[lvl]) * v2, rtol=1e-12)
                checked += 1
    assert checked > 10
    # MSD from the same position samples is unchanged by the extra correlations
    sc, sm, _ = c.result("msd")
    assert np.array_equal(sc, mc)


@pytest.mark.parametrize("n,rho,L_override", [(2048, 0.8442, None), (1500, 1.2, None), (400, 0.9, 8.5)])
def test_staged_list_rule16_contains_the_sphere(oracle, n, rho, L_override):
    # cubic-box rule of the cell-segment build (16-bit integer distance):
    # checked against exact integer geometry, not a restatement of the formula.
    # (1) every pair whose exact fixed-point distance is within rc + skin is in
    # the list; (2) no listed pair is farther than rc + skin + 2 (sqrt 3 +
    # 0.001) units of L / 2^16; (3) the set equals the predicate evaluated on all
    # pairs by brute force (the cell traversal misses no cell); (4) symmetric.
    # The small box (3 cells per side) exercises the wrapped 16-bit differences.
    pos, L, species, _ = _liquid(oracle, n, rho, 9, species_frac=0.3)
    if L_override is not None:
        pos = pos * (L_override / L[0])
        L = np.full(3, L_override)
    pot = _ka_potential(oracle)
    skin = 0.3
    rcs = pot.r_cut + skin
    ncell = np.floor(L / np.max(rcs)).astype(np.uint32)
    assert ncell.min() >= 3
    u = oracle.fixed_point(pos, L)
    counts, lst, m = oracle.staged_neighbour_list(u, species, ncell, L, rcs, 512)
    assert m <= 512
    got = np.zeros((n, n), bool)
    for i in range(n):
        got[i, lst[i, :counts[i]]] = True
        assert len(np.unique(lst[i, :counts[i]])) == counts[i]
    assert np.array_equal(got, got.T)
    # exact distances of the fixed-point positions (minimum image), FP64
    D = (u[:, None, :] - u[None, :, :]).astype(np.int32).astype(np.float64) * (L[0] / 2 ** 32)
    r = np.sqrt(np.sum(D * D, -1))
    cut = rcs[species[:, None], species[None, :]]
    np.fill_diagonal(r, np.inf)
    assert np.all(got[r <= cut])
    unit = L[0] / 65536
    assert np.all(r[got] <= cut[got] + 2 * (np.sqrt(3) + 0.001) * unit)
    # brute force over all pairs with numpy's integer arithmetic
    h = (u >> 16).astype(np.int64)
    d = (h[:, None, :] - h[None, :, :]) % 65536
    d = np.where(d >= 32768, d - 65536, d)
    rq = np.sum(d * d, -1)
    thr = np.floor((rcs * 65536 / L[0] + np.sqrt(3) + 0.001) ** 2)
    expect = rq <= thr[species[:, None], species[None, :]]
    np.fill_diagonal(expect, False)
    assert np.array_equal(got, expect)
    # the margin is thin: the list exceeds the exact sphere by well under 1 %
    assert got.sum() < 1.01 * np.sum(r <= cut)


def test_untruncated_brute_force_closed_forms(oracle):
    # the reference for pair_full: a sharp cutoff beyond every minimum-image
    # distance leaves the potential untruncated.  Two particles: the
    # Lennard-Jones closed form F = 24 eps (2 (s/r)^12 - (s/r)^6) / r, U/2 each,
    # the virial trace F r / 2 each; three particles: the sum of the pair terms
    L = np.array([40.0, 40.0, 40.0])
    pot = oracle.Potential(oracle.POT_LJ, "sharp", 1.0, 1.0, 1e4)
    for r in (0.95, 1.12, 1.7, 3.9, 12.0):
        pos = np.array([[0.0, 0.0, 0.0], [r, 0.0, 0.0]])
        f, u, s = oracle.pair_force_brute(pos, np.zeros(2, np.uint32), L, pot)
        sr6 = r ** -6
        fmag = 24 * (2 * sr6 * sr6 - sr6) / r
        uu = 4 * (sr6 * sr6 - sr6)
        assert np.allclose(f[0], [-fmag, 0, 0], rtol=1e-13, atol=1e-15)
        assert np.allclose(f[1], [fmag, 0, 0], rtol=1e-13, atol=1e-15)
        assert np.allclose(u, [uu / 2, uu / 2], rtol=1e-13, atol=1e-16)
        assert np.allclose(s[:, 0], [fmag * r / 2, fmag * r / 2], rtol=1e-13, atol=1e-16)
    # minimum image across the boundary: 39.0 apart in a box of 40 is 1.0 apart
    pos = np.array([[-19.5, 0.0, 0.0], [19.5, 0.0, 0.0]])
    f, u, _ = oracle.pair_force_brute(pos, np.zeros(2, np.uint32), L, pot)
    assert np.allclose(f[0], [24.0, 0.0, 0.0], rtol=1e-13) and np.allclose(u, [0.0, 0.0], atol=1e-14)
    rng = np.random.default_rng(8)
    pos = rng.uniform(-3, 3, (3, 3))
    f3, u3, _ = oracle.pair_force_brute(pos, np.zeros(3, np.uint32), L, pot)
    fsum = np.zeros((3, 3))
    usum = np.zeros(3)
    for a, b in ((0, 1), (0, 2), (1, 2)):
        fp, up, _ = oracle.pair_force_brute(pos[[a, b]], np.zeros(2, np.uint32), L, pot)
        fsum[[a, b]] += fp
        usum[[a, b]] += up
    assert np.allclose(f3, fsum, rtol=1e-12, atol=1e-12) and np.allclose(u3, usum, rtol=1e-12, atol=1e-14)
