"""The ensemble's device rebuild orders particles by cell without a sort.

Between rebuilds a particle moves into one of the 27 cells around its old
cell, so the stable order by new cell is a merge of the old cell order: new
cell c takes, from each source cell c + e in ascending index, the particles
that moved by -e, in memory order.  This restates the kernels' arithmetic
(ens_rank_kernel, ens_cell_count_kernel, ens_scatter_kernel in binning.hip:
move codes, ranks per (old cell, code), running sums in ascending source
order) on the host and checks that it gives the permutation and the cell
offsets of a stable sort, on random moves with periodic wrap (3 to 6 cells
per side, so that the wrap at both ends and the smallest ring are covered).
"""
import numpy as np


def _ring_step(a, b, nc):
    d = a - b
    if d > 1:
        d -= nc
    if d < -1:
        d += nc
    return d if -1 <= d <= 1 else 2


def _ascending(a, nc):
    if a == 0:
        return (0, 1, -1)
    if a == nc - 1:
        return (1, -1, 0)
    return (-1, 0, 1)


def _code(ex, ey, ez):
    return (ez + 1) * 9 + (ey + 1) * 3 + (ex + 1)


def _coords(c, nc):
    return c % nc, (c // nc) % nc, c // (nc * nc)


def _lin(x, y, z, nc):
    return (z * nc + y) * nc + x


def _merge_order(old_cells, new_cells, nc):
    """Slots of the particles (memory order = sorted old cells) and new offsets."""
    C = nc ** 3
    n = len(old_cells)
    off = np.searchsorted(old_cells, np.arange(C + 1))
    rank = np.zeros(n, int)
    code = np.zeros(n, int)
    cnt = np.zeros((C, 27), int)
    for cp in range(C):                       # ens_rank_kernel: one old cell at a time
        xo, yo, zo = _coords(cp, nc)
        for i in range(off[cp], off[cp + 1]):
            x, y, z = _coords(new_cells[i], nc)
            e = (_ring_step(xo, x, nc), _ring_step(yo, y, nc), _ring_step(zo, z, nc))
            assert 2 not in e
            code[i] = _code(*e)
            rank[i] = cnt[cp, code[i]]
            cnt[cp, code[i]] += 1
    base27 = np.zeros((C, 27), int)           # ens_cell_count_kernel
    count = np.zeros(C, int)
    for c in range(C):
        x, y, z = _coords(c, nc)
        v = 0
        for ez in _ascending(z, nc):
            for ey in _ascending(y, nc):
                for ex in _ascending(x, nc):
                    cd = _code(ex, ey, ez)
                    base27[c, cd] = v
                    v += cnt[_lin((x + ex) % nc, (y + ey) % nc, (z + ez) % nc, nc), cd]
        count[c] = v
    offsets = np.concatenate([[0], np.cumsum(count)])   # with ens_cell_carry_kernel
    slot = offsets[new_cells] + base27[new_cells, code] + rank   # ens_scatter_kernel
    return slot, offsets


def test_merge_is_the_stable_sort_by_cell():
    rng = np.random.default_rng(7)
    for trial in range(150):
        nc = int(rng.integers(3, 7))
        n = int(rng.integers(1, 300))
        old = np.sort(rng.integers(0, nc ** 3, n))
        new = np.empty(n, int)
        for i in range(n):
            x, y, z = _coords(old[i], nc)
            e = rng.integers(-1, 2, 3) if rng.random() < 0.5 else np.zeros(3, int)
            new[i] = _lin((x + e[0]) % nc, (y + e[1]) % nc, (z + e[2]) % nc, nc)
        slot, offsets = _merge_order(old, new, nc)
        perm = np.argsort(new, kind="stable")
        assert np.array_equal(slot[perm], np.arange(n)), trial
        assert np.array_equal(offsets, np.searchsorted(np.sort(new), np.arange(nc ** 3 + 1))), trial


def test_ascending_source_order_covers_every_wrap():
    # the three source coordinates a + e, in the order the kernels visit them, ascend
    for nc in (3, 4, 7):
        for a in range(nc):
            src = [(a + e) % nc for e in _ascending(a, nc)]
            assert src == sorted(src) and len(set(src)) == 3
