// Device side of the cell-segment decomposition: staging of a segment's
// neighbourhood union in LDS (shared by the neighbour-list build and the pair
// force kernels, which must agree slot for slot).
#pragma once

#include "kernels.hpp"

namespace halmd {

struct seg_stage {
    unsigned base[10];             // LDS start of each row; base[9] = union size
    unsigned g[9][2], n[9][2];     // memory ranges of each row
    unsigned x0, w, pb, pe;        // first cell, cells, particle range of the segment
};

__device__ __forceinline__ unsigned seg_row_start(seg_params const& sp, unsigned seg, unsigned r) {
    unsigned const nx = sp.grid.ncell[0], ny = sp.grid.ncell[1], nz = sp.grid.ncell[2];
    unsigned const t = seg / sp.nsx;
    unsigned const y = t % ny, z = t / ny;
    unsigned const kz = (z + nz + r / 3 - 1) % nz, ky = (y + ny + r % 3 - 1) % ny;
    return (kz * ny + ky) * nx;
}

// Called by the whole workgroup.  Fills st (and, if bound is given, the LDS
// start of each of the w + 2 cells of every row: bound[r * (SEG_MAX + 3) + k],
// k = 0 .. w + 2), then, if U + 1 <= sp.lds_cap, copies the union into
// s_rec[0 .. U) and a sentinel record into s_rec[U]: fixed-point y and z half
// a box from the centre of the segment's row, hence at least L / 3 >= rc +
// skin from each of its particles in both (ny, nz >= 3).  Returns the union
// size U; every thread sees the same value.
__device__ __forceinline__ unsigned stage_union(seg_params const& sp, unsigned seg, uint32_t const* __restrict__ off,
                                                uint4 const* __restrict__ posq, uint4* s_rec, seg_stage& st,
                                                unsigned* bound) {
    unsigned const tid = threadIdx.x;
    if (tid < 9) seg_row(sp, seg, tid, off, st.g[tid], st.n[tid]);
    if (tid == 9) seg_particles(sp, seg, off, st.x0, st.w, st.pb, st.pe);
    __syncthreads();
    if (tid == 0) {
        unsigned b = 0;
        for (unsigned r = 0; r < 9; ++r) {
            st.base[r] = b;
            b += st.n[r][0] + st.n[r][1];
        }
        st.base[9] = b;
    }
    __syncthreads();
    unsigned const U = st.base[9];
    if (bound && tid < 9) {
        // cells x0 - 1 .. x0 + w (periodic) of row tid, in LDS order
        unsigned const nx = sp.grid.ncell[0];
        unsigned const row = seg_row_start(sp, seg, tid);
        unsigned* const bd = bound + tid * (SEG_MAX + 3);
        unsigned b = st.base[tid];
        bd[0] = b;
        for (unsigned k = 0; k < st.w + 2; ++k) {
            unsigned const x = (st.x0 + nx - 1 + k) % nx;
            b += off[row + x + 1] - off[row + x];
            bd[k + 1] = b;
        }
    }
    if (U + 1 <= sp.lds_cap) {
        if (tid == 0) {
            unsigned const ny = sp.grid.ncell[1], nz = sp.grid.ncell[2];
            unsigned const t = seg / sp.nsx;
            unsigned const y = t % ny, z = t / ny;
            uint32_t const cy = uint32_t(((2ull * y + 1) << 31) / ny);   // (y + 1/2) 2^32 / ny
            uint32_t const cz = uint32_t(((2ull * z + 1) << 31) / nz);
            s_rec[U] = make_uint4(0u, cy + 0x80000000u, cz + 0x80000000u, 0u);
        }
        // flattened copy: every thread issues eight independent loads before
        // it stores any of them, so a workgroup waits for about
        // U / (8 * blockDim) memory latencies instead of one per row.  Slots
        // past U load record 0 (U >= 1 inside the loop) and store nothing.
        auto fetch = [&](unsigned u) -> uint4 {
            u = u < U ? u : 0u;
            unsigned r = 0;
#pragma unroll
            for (unsigned q = 1; q < 9; ++q) r += u >= st.base[q] ? 1u : 0u;
            unsigned const o = u - st.base[r];
            unsigned const src = o < st.n[r][0] ? st.g[r][0] + o : st.g[r][1] + (o - st.n[r][0]);
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
