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

// the largest of the DISP_SLOTS displacement words at slots exceeds bits
// (false without slots).  Each wave reads the 64 words itself, so the
// decision is uniform without a barrier.  publish (optional): lane 0 of
// workgroup 0 stores (seq << 32) | the largest word into publish[seq %
// PUBLISH_RING], host-mapped memory, by a vector store at system scope.
__device__ __forceinline__ bool displacement_aborts(uint32_t const* __restrict__ slots, uint32_t bits,
                                                    unsigned long long* publish = nullptr, uint32_t seq = 0) {
    if (!slots) return false;
    static_assert(DISP_SLOTS == WAVE, "one displacement slot per lane");
    unsigned m = slots[(threadIdx.x % WAVE) * DISP_STRIDE];
    for (int o = WAVE / 2; o > 0; o >>= 1) m = max(m, (unsigned)__shfl_xor((int)m, o, WAVE));
    if (publish && blockIdx.x == 0 && threadIdx.x == 0)
        __hip_atomic_store(publish + seq % PUBLISH_RING, (static_cast<unsigned long long>(seq) << 32) | m,
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    return m > bits;
}

// inclusive sum over the lanes of a wave
__device__ __forceinline__ unsigned wave_inclusive_sum(unsigned x) {
    unsigned const lane = threadIdx.x % WAVE;
#pragma unroll
    for (unsigned d = 1; d < WAVE; d <<= 1) {
        unsigned const t = __shfl_up(x, d, WAVE);
        if (lane >= d) x += t;
    }
    return x;
}

// Staged record formats: the full fixed-point record (pair forces; non-cubic
// list build), or for the cubic-box list build the top 16 bits of the x and y
// coordinates in one word (x low) and of z with the species above them in the
// other.
__device__ __forceinline__ uint2 pack_rec16(uint4 const r) {
    return make_uint2((r.x >> 16) | (r.y & 0xffff0000u), (r.z >> 16) | ((r.w & 3u) << 16));
}
__device__ __forceinline__ void put_rec(uint4* s, unsigned k, uint4 const r) { s[k] = r; }
__device__ __forceinline__ void put_rec(uint2* s, unsigned k, uint4 const r) { s[k] = pack_rec16(r); }

// Called by the whole workgroup (at least 64 threads).  Fills st (and, if
// bound is given, the LDS start of each of the w + 2 cells of every union row:
// bound[ur * bstride + k], k = 0 .. w + 2; bstride >= w + 3), then, if U + 1 <= sp.lds_cap,
// copies the union into s_rec[0 .. U) (in the format of Rec) and a sentinel
// record into s_rec[U]:
// fixed-point y and z half a box from the centre of the segment's rows and
// plane, hence at least one cell width (>= rc + skin) from each of its
// particles in both (h <= ny - 2, nz >= 3).  Returns the union size U; every
// thread sees the same value.
//
// Wave 0 locates the segment and its rows (lanes 0 .. nr - 1 the union rows,
// lanes 32 .. 32 + h - 1 the own rows) and forms both prefix sums by a wave
// scan; then wave v copies union rows v, v + W, ... (W waves), each in chunks
// of four 64-record loads per lane issued before their stores.
template <class Rec>
__device__ __forceinline__ unsigned stage_union(seg_params const& sp, unsigned seg, uint32_t const* __restrict__ off,
                                                uint4 const* __restrict__ posq, Rec* s_rec, seg_stage& st,
                                                unsigned* bound, unsigned bstride = SEG_BOUND) {
    unsigned const tid = threadIdx.x, lane = tid % WAVE;
    if (tid < WAVE) {
        seg_place const q = seg_locate(sp, seg);
        unsigned const nr = 3 * (q.h + 2);   // <= SEG_NR = 30 < 32
        unsigned rows = 0, own = 0;
        if (lane < nr) {
            seg_row(sp, q, lane, off, st.g[lane], st.n[lane]);
            rows = st.n[lane][0] + st.n[lane][1];
        }
        if (lane >= 32 && lane < 32 + q.h) {
            unsigned b;
            seg_own_row(sp, q, lane - 32, off, b, own);
            st.own_begin[lane - 32] = b;
        }
        unsigned const rsum = wave_inclusive_sum(rows);   // lanes >= nr carry the total U
        unsigned const osum = wave_inclusive_sum(own);    // (own is zero below lane 32)
        if (lane < nr) st.base[lane] = rsum - rows;
        if (lane == nr) st.base[nr] = rsum;
        if (lane >= 32 && lane < 32 + q.h) st.own_pre[lane - 32 + 1] = osum;
        if (lane == 0) {
            st.q = q;
            st.nr = nr;
            st.own_pre[0] = 0;
        }
    }
    __syncthreads();
    seg_place const q = st.q;
    unsigned const nr = st.nr, U = st.base[nr];
    if (bound && tid < nr) {
        // cells x0 - 1 .. x0 + w (periodic) of union row tid, in LDS order
        unsigned const nx = sp.grid.ncell[0];
        unsigned const row = seg_row_start(sp, q, tid);
        unsigned* const bd = bound + tid * bstride;
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
            put_rec(s_rec, U, make_uint4(0u, cy + 0x80000000u, cz + 0x80000000u, 0u));
        }
        unsigned const W = blockDim.x / WAVE;
        for (unsigned r = tid / WAVE; r < nr; r += W) {   // wave-uniform
            unsigned const b = st.base[r], g0 = st.g[r][0], n0 = st.n[r][0], g1 = st.g[r][1];
            unsigned const n = n0 + st.n[r][1];
            for (unsigned u0 = 0; u0 < n; u0 += 4 * WAVE) {
                // slots at or past n read record 0 (one cached line) and
                // store nothing; the four loads are unconditional, so that
                // they issue back to back
                auto src = [&](unsigned u) { return u < n ? (u < n0 ? g0 + u : g1 + (u - n0)) : 0u; };
                unsigned const u = u0 + lane;
                uint4 const a0 = posq[src(u)];
                uint4 const a1 = posq[src(u + WAVE)];
                uint4 const a2 = posq[src(u + 2 * WAVE)];
                uint4 const a3 = posq[src(u + 3 * WAVE)];
                if (u < n) put_rec(s_rec, b + u, a0);
                if (u + WAVE < n) put_rec(s_rec, b + u + WAVE, a1);
                if (u + 2 * WAVE < n) put_rec(s_rec, b + u + 2 * WAVE, a2);
                if (u + 3 * WAVE < n) put_rec(s_rec, b + u + 3 * WAVE, a3);
            }
        }
    }
    __syncthreads();
    return U;
}

} // namespace halmd
