// Cell binning and neighbour-list construction for gfx950.
//
// Both work on the fixed-point snapshot of the positions (posq: 32 bits per
// coordinate, one unit = L / 2^32, written by the integrator from the
// double-word state).  Integer coordinates make the rules exact and cheap:
//  - cell index per component: (u * ncell) >> 32 (the reference device rule
//    uint(r / len + ncell) mod ncell of halmd/mdsim/gpu/binning_kernel.cu:34-54,
//    with the grid anchored at the lower box corner), linearised x fastest;
//  - minimum-image separation: int32(u_i - u_j) (wrap-around subtraction),
//    converted to FP32 and scaled by q = float(L) 2^-32.
// Instead of fixed-size cell slots with an overflow/retry loop
// (binning.cpp:166-221) the cells are compressed ranges: a stable radix sort
// of (cell, particle) puts the particles of each cell in ascending index
// order, and offset[c] .. offset[c + 1] is the range of cell c (an exclusive
// scan over cells, so the cells x-1, x, x+1 of one row are one contiguous
// range).  A compact copy of the positions in that order carries the particle
// index and species in its fourth word: one 16-byte record per candidate.
//
// Neighbour lists are full lists (i -> j and j -> i) with the predicate
// rr <= (rc_ab + skin)^2 in FP32 (from_binning_kernel.cu:144-163), rr formed
// as (dx*dx + dy*dy) + dz*dz without contraction.  Candidate order: the 9 rows
// (dz, dy) in lexicographic order, in each row the cells x-1, x, x+1, in each
// cell ascending particle index.  The list is stored per wavefront of 64
// particles, column-major inside the wavefront: entry k of particle i sits at
// (i / 64) * 64 * cap + k * 64 + i % 64, so one wave reading entry k touches
// 256 contiguous bytes.  A per-particle count replaces the placeholder
// terminator, and no memset of the list is needed.
#include "kernels.hpp"

namespace halmd {
namespace {

__device__ __forceinline__ unsigned cell_component(uint32_t u, unsigned nc) {
    return unsigned((uint64_t(u) * nc) >> 32);
}

__global__ void compute_cell_kernel(uint4 const* __restrict__ posq, unsigned n, cell_grid g,
                                    uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint4 const r = posq[i];
    unsigned const x = cell_component(r.x, g.ncell[0]);
    unsigned const y = cell_component(r.y, g.ncell[1]);
    unsigned const z = cell_component(r.z, g.ncell[2]);
    keys[i] = (z * g.ncell[1] + y) * g.ncell[0] + x;
    vals[i] = i;
}

// offset[c] = first k with keys[k] >= c, for c = 0 .. ncells (k = 0 .. n)
__global__ void cell_offsets_kernel(uint32_t const* __restrict__ keys, unsigned n, unsigned ncells,
                                    uint32_t* __restrict__ offset) {
    unsigned const k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k > n) return;
    int const prev = k == 0 ? -1 : int(keys[k - 1]);
    int const cur = k == n ? int(ncells) : int(keys[k]);
    for (int c = prev + 1; c <= cur; ++c) offset[c] = k;
}

// compact record: fixed-point position, w = (particle index << 2) | species
__global__ void gather_cell_pos_kernel(uint4 const* __restrict__ posq, uint32_t const* __restrict__ vals, unsigned n,
                                       uint4* __restrict__ cell_pos) {
    unsigned k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    uint32_t const j = vals[k];
    uint4 r = posq[j];
    r.w = (j << 2) | (r.w & 3u);
    cell_pos[k] = r;
}

struct nb_state {
    uint4 ri;
    unsigned i, ti, c, cap;
    uint32_t* out;
};

__device__ __forceinline__ void test_candidate(nb_state& s, uint4 const rj, neighbour_params const& p,
                                               float const* s_rr) {
#pragma clang fp contract(off)
    uint32_t const j = rj.w >> 2;
    float const dx = float(int32_t(s.ri.x - rj.x)) * p.q[0];
    float const dy = float(int32_t(s.ri.y - rj.y)) * p.q[1];
    float const dz = float(int32_t(s.ri.z - rj.z)) * p.q[2];
    float rr = dx * dx;
    rr = rr + dy * dy;
    rr = rr + dz * dz;
    if (j != s.i && rr <= s_rr[s.ti + (rj.w & 3u)]) {
        if (s.c < s.cap) s.out[size_t(s.c) * WAVE] = j;
        ++s.c;
    }
}

__device__ __forceinline__ void scan_range(nb_state& s, unsigned b, unsigned e, uint4 const* __restrict__ cell_pos,
                                           neighbour_params const& p, float const* s_rr) {
    unsigned k = b;
    for (; k + 4 <= e; k += 4) {
        uint4 const r0 = cell_pos[k], r1 = cell_pos[k + 1], r2 = cell_pos[k + 2], r3 = cell_pos[k + 3];
        test_candidate(s, r0, p, s_rr);
        test_candidate(s, r1, p, s_rr);
        test_candidate(s, r2, p, s_rr);
        test_candidate(s, r3, p, s_rr);
    }
    for (; k < e; ++k) test_candidate(s, cell_pos[k], p, s_rr);
}

__global__ __launch_bounds__(256) void build_neighbours_kernel(
    uint4 const* __restrict__ posq, unsigned n, uint32_t const* __restrict__ offset,
    uint4 const* __restrict__ cell_pos, neighbour_params p, unsigned cap, uint32_t* __restrict__ list,
    uint32_t* __restrict__ count, uint32_t* __restrict__ max_count)
{
    __shared__ float s_rr[MAX_TYPES * MAX_TYPES];
    if (threadIdx.x < MAX_TYPES * MAX_TYPES) s_rr[threadIdx.x] = p.rr_cut_skin[threadIdx.x];
    __syncthreads();

    unsigned const i = blockIdx.x * blockDim.x + threadIdx.x;
    nb_state s;
    s.c = 0;
    if (i < n) {
        s.ri = posq[i];
        s.i = i;
        s.ti = (s.ri.w & 3u) * p.ntype;
        s.cap = cap;
        s.out = list + size_t(i & ~(WAVE - 1)) * cap + (i & (WAVE - 1));
        unsigned const nx = p.grid.ncell[0], ny = p.grid.ncell[1], nz = p.grid.ncell[2];
        unsigned const cx = cell_component(s.ri.x, nx);
        unsigned const cy = cell_component(s.ri.y, ny);
        unsigned const cz = cell_component(s.ri.z, nz);
        unsigned const xm = (cx + nx - 1) % nx, xp = (cx + 1) % nx;
        for (int dz = -1; dz <= 1; ++dz) {
            unsigned const kz = (cz + nz + dz) % nz;
            for (int dy = -1; dy <= 1; ++dy) {
                unsigned const ky = (cy + ny + dy) % ny;
                unsigned const row = (kz * ny + ky) * nx;
                if (cx > 0 && cx + 1 < nx) {
                    // cells x-1, x, x+1 are one contiguous range
                    scan_range(s, offset[row + xm], offset[row + xp + 1], cell_pos, p, s_rr);
                } else if (cx == 0) {
                    scan_range(s, offset[row + xm], offset[row + xm + 1], cell_pos, p, s_rr);
                    scan_range(s, offset[row], offset[row + 2], cell_pos, p, s_rr);
                } else {
                    scan_range(s, offset[row + xm], offset[row + nx], cell_pos, p, s_rr);
                    scan_range(s, offset[row], offset[row + 1], cell_pos, p, s_rr);
                }
            }
        }
        count[i] = s.c;
    }
    // largest count of the wave -> one atomic per wave
    unsigned m = s.c;
    for (int off = WAVE / 2; off > 0; off >>= 1) m = max(m, (unsigned)__shfl_xor((int)m, off, WAVE));
    if ((threadIdx.x % WAVE) == 0 && m > 0) atomicMax(max_count, m);
}

} // namespace

void launch_compute_cell(uint4 const* posq, unsigned n, cell_grid g, uint32_t* keys, uint32_t* vals, hipStream_t s) {
    if (!n) return;
    compute_cell_kernel<<<div_up(n, 256), 256, 0, s>>>(posq, n, g, keys, vals);
    HIP_CHECK(hipGetLastError());
}

void launch_cell_offsets(uint32_t const* keys, unsigned n, unsigned ncells, uint32_t* offset, hipStream_t s) {
    cell_offsets_kernel<<<div_up(size_t(n) + 1, 256), 256, 0, s>>>(keys, n, ncells, offset);
    HIP_CHECK(hipGetLastError());
}

void launch_gather_cell_pos(uint4 const* posq, uint32_t const* vals, unsigned n, uint4* cell_pos, hipStream_t s) {
    if (!n) return;
    gather_cell_pos_kernel<<<div_up(n, 256), 256, 0, s>>>(posq, vals, n, cell_pos);
    HIP_CHECK(hipGetLastError());
}

void launch_build_neighbours(uint4 const* posq, unsigned n, uint32_t const* offset, uint4 const* cell_pos,
                             neighbour_params const& p, unsigned cap, uint32_t* list, uint32_t* count,
                             uint32_t* max_count, hipStream_t s) {
    if (!n) return;
    build_neighbours_kernel<<<div_up(n, 256), 256, 0, s>>>(posq, n, offset, cell_pos, p, cap, list, count,
                                                           max_count);
    HIP_CHECK(hipGetLastError());
}

} // namespace halmd
