// Device side of the cell-segment decomposition: staging of a segment's
// neighbourhood union in LDS (shared by the neighbour-list build and the pair
// force kernels, which must agree slot for slot).
//
// A segment is S cells along x by R rows along y in one z-plane.  Its union
// holds 3 (R + 2) rows (dz outer, then y), each the cells x0 - 1 .. x0 + S of
// one (y, z) row; a particle's 9 candidate rows (dz, dy) are 9 of them, in
// the same (dz, dy) order.
#pragma once

#include "kernels.hpp"

namespace halmd {

struct seg_stage {
    seg_place q;
    unsigned nr;                            // union rows, 3 (h + 2)
    unsigned base[SEG_NR + 1];              // LDS start of each union row; base[nr] = U
    unsigned g[SEG_NR][2], n[SEG_NR][2];    // memory ranges of each union row
    unsigned own_begin[SEG_RMAX];           // memory start of the segment's particles in row j
    unsigned own_pre[SEG_RMAX + 1];         // particles of rows before j; own_pre[h] = all
};

constexpr unsigned SEG_BOUND = SEG_MAX + 3;   // bound words per union row

__device__ __forceinline__ unsigned seg_row_start(seg_params const& sp, seg_place const& q, unsigned ur) {
    unsigned const nx = sp.grid.ncell[0], ny = sp.grid.ncell[1], nz = sp.grid.ncell[2];
    unsigned const hh = q.h + 2;
    unsigned const kz = (q.z + nz + ur / hh - 1) % nz, ky = (q.y0 + ny + ur % hh - 1) % ny;
    return (kz * ny + ky) * nx;
}

// own particle t (0 .. own_pre[h] - 1) of the segment: its row j and memory index
__device__ __forceinline__ void seg_own(seg_stage const& st, unsigned t, unsigned& j, unsigned& i) {
    unsigned r = 0;
    for (unsigned k = 1; k < st.q.h; ++k) r += t >= st.own_pre[k] ? 1u : 0u;
    j = r;
    i = st.own_begin[r] + (t - st.own_pre[r]);
}

// Called by the whole workgroup (at least 64 threads).  Fills st (and, if
// bound is given, the LDS start of each of the w + 2 cells of every union row:
// bound[ur * SEG_BOUND + k], k = 0 .. w + 2), then, if U + 1 <= sp.lds_cap,
// copies the union into s_rec[0 .. U) and a sentinel record into s_rec[U]:
// fixed-point y and z half a box from the centre of the segment's rows and
// plane, hence at least one cell width (>= rc + skin) from each of its
// particles in both (h <= ny - 2, nz >= 3).  Returns the union size U; every
// thread sees the same value.
__device__ __forceinline__ unsigned stage_union(seg_params const& sp, unsigned seg, uint32_t const* __restrict__ off,
                                                uint4 const* __restrict__ posq, uint4* s_rec, seg_stage& st,
                                                unsigned* bound) {
    unsigned const tid = threadIdx.x;
    seg_place const q = seg_locate(sp, seg);
    unsigned const nr = 3 * (q.h + 2);
    if (tid < nr) seg_row(sp, q, tid, off, st.g[tid], st.n[tid]);
    if (tid >= 32 && tid < 32 + q.h) {
        unsigned b, c;
        seg_own_row(sp, q, tid - 32, off, b, c);
        st.own_begin[tid - 32] = b;
        st.own_pre[tid - 32 + 1] = c;
    }
    if (tid == 0) {
        st.q = q;
        st.nr = nr;
    }
    __syncthreads();
    if (tid == 0) {
        unsigned b = 0;
        for (unsigned r = 0; r < nr; ++r) {
            st.base[r] = b;
            b += st.n[r][0] + st.n[r][1];
        }
        st.base[nr] = b;
        unsigned p = 0;
        st.own_pre[0] = 0;
        for (unsigned j = 0; j < q.h; ++j) {
            p += st.own_pre[j + 1];
            st.own_pre[j + 1] = p;
        }
    }
    __syncthreads();
    unsigned const U = st.base[nr];
    if (bound && tid < nr) {
        // cells x0 - 1 .. x0 + w (periodic) of union row tid, in LDS order
        unsigned const nx = sp.grid.ncell[0];
        unsigned const row = seg_row_start(sp, q, tid);
        unsigned* const bd = bound + tid * SEG_BOUND;
        unsigned b = st.base[tid];
        bd[0] = b;
        for (unsigned k = 0; k < q.w + 2; ++k) {
            unsigned const x = (q.x0 + nx - 1 + k) % nx;
            b += off[row + x + 1] - off[row + x];
            bd[k + 1] = b;
        }
    }
    if (U + 1 <= sp.lds_cap) {
        if (tid == 0) {
            unsigned const ny = sp.grid.ncell[1], nz = sp.grid.ncell[2];
            uint32_t const cy = uint32_t(((2ull * q.y0 + q.h) << 31) / ny);   // (y0 + h/2) 2^32 / ny
            uint32_t const cz = uint32_t(((2ull * q.z + 1) << 31) / nz);
            s_rec[U] = make_uint4(0u, cy + 0x80000000u, cz + 0x80000000u, 0u);
        }
        // flattened copy: every thread issues eight independent loads before
        // it stores any of them, so a workgroup waits for about
        // U / (8 * blockDim) memory latencies instead of one per row.  Slots
        // past U load record 0 (U >= 1 inside the loop) and store nothing.
        auto fetch = [&](unsigned u) -> uint4 {
            u = u < U ? u : 0u;
            unsigned lo = 0, hi = nr;   // union row: last r with base[r] <= u
            while (hi - lo > 1) {
                unsigned const m = (lo + hi) / 2;
                if (st.base[m] <= u) lo = m; else hi = m;
            }
            unsigned const o = u - st.base[lo];
            unsigned const src = o < st.n[lo][0] ? st.g[lo][0] + o : st.g[lo][1] + (o - st.n[lo][0]);
            return posq[src];
        };
        unsigned const B = blockDim.x;
        for (unsigned u0 = tid; u0 < U; u0 += 8 * B) {
            uint4 const a0 = fetch(u0), a1 = fetch(u0 + B), a2 = fetch(u0 + 2 * B), a3 = fetch(u0 + 3 * B);
            uint4 const a4 = fetch(u0 + 4 * B), a5 = fetch(u0 + 5 * B), a6 = fetch(u0 + 6 * B),
                        a7 = fetch(u0 + 7 * B);
            s_rec[u0] = a0;
            if (u0 + B < U) s_rec[u0 + B] = a1;
            if (u0 + 2 * B < U) s_rec[u0 + 2 * B] = a2;
            if (u0 + 3 * B < U) s_rec[u0 + 3 * B] = a3;
            if (u0 + 4 * B < U) s_rec[u0 + 4 * B] = a4;
            if (u0 + 5 * B < U) s_rec[u0 + 5 * B] = a5;
            if (u0 + 6 * B < U) s_rec[u0 + 6 * B] = a6;
            if (u0 + 7 * B < U) s_rec[u0 + 7 * B] = a7;
        }
    }
    __syncthreads();
    return U;
}

} // namespace halmd
