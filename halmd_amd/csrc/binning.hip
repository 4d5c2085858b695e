// Cell binning and neighbour-list construction for gfx950.
//
// Binning follows the reference device rule for the cell index
// (halmd/mdsim/gpu/binning_kernel.cu:34-75): idx = uint(r / len + ncell) mod
// ncell per component, x fastest.  Instead of fixed-size cell slots with an
// overflow/retry loop (binning.cpp:166-221) the cells are compressed ranges:
// a stable radix sort of (cell, particle) puts the particles of each cell in
// ascending index order, and offset[c] .. offset[c + 1] is the range of cell c
// (an exclusive scan over cells, so the cells x-1, x, x+1 of one row are one
// contiguous range).  A compact copy of the positions in that order carries
// the particle index and species in its fourth word, so the neighbour search
// reads one 16-byte record per candidate.
//
// Neighbour lists are full lists (i -> j and j -> i) with the reference
// device predicate rr <= (rc_ab + skin)^2 in FP32 on the high words of the
// positions (from_binning_kernel.cu:144-163).  Candidate order: the 9 rows
// (dz, dy) in lexicographic order, in each row the cells x-1, x, x+1, in each
// cell ascending particle index.  The list is stored per wavefront of 64
// particles, column-major inside the wavefront: entry k of particle i sits at
// (i / 64) * 64 * cap + k * 64 + i % 64, so one wave reading entry k touches
// 256 contiguous bytes.  A per-particle count replaces the placeholder
// terminator, and no memset of the list is needed.
#include "kernels.hpp"

namespace halmd {
namespace {

__device__ __forceinline__ unsigned cell_component(float r, float len, unsigned nc) {
#pragma clang fp contract(off)
    float q = r / len;
    float s = q + float(nc);
    return unsigned(s) % nc;
}

__device__ __forceinline__ void cell_of(float4 const& r, cell_grid const& g, unsigned& x, unsigned& y, unsigned& z) {
    x = cell_component(r.x, g.cell_length[0], g.ncell[0]);
    y = cell_component(r.y, g.cell_length[1], g.ncell[1]);
    z = cell_component(r.z, g.cell_length[2], g.ncell[2]);
}

__global__ void compute_cell_kernel(float4 const* __restrict__ pos, unsigned n, cell_grid g,
                                    uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    unsigned x, y, z;
    cell_of(pos[2 * size_t(i)], g, x, y, z);
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

// compact record: high-word position, w = (particle index << 2) | species
__global__ void gather_cell_pos_kernel(float4 const* __restrict__ pos, uint32_t const* __restrict__ vals, unsigned n,
                                       float4* __restrict__ cell_pos) {
    unsigned k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    uint32_t const j = vals[k];
    float4 r = pos[2 * size_t(j)];
    r.w = __uint_as_float((j << 2) | (__float_as_uint(r.w) & 3u));
    cell_pos[k] = r;
}

struct nb_state {
    float4 ri;
    unsigned i, ti, c, cap;
    uint32_t* out;
};

__device__ __forceinline__ void test_candidate(nb_state& s, float4 const rj, neighbour_params const& p,
                                               float const* s_rr) {
#pragma clang fp contract(off)
    uint32_t const w = __float_as_uint(rj.w);
    uint32_t const j = w >> 2;
    float dx = s.ri.x - rj.x, dy = s.ri.y - rj.y, dz = s.ri.z - rj.z;
    dx = dx - rintf(dx * p.invL[0]) * p.L[0];
    dy = dy - rintf(dy * p.invL[1]) * p.L[1];
    dz = dz - rintf(dz * p.invL[2]) * p.L[2];
    float rr = dx * dx;
    rr = rr + dy * dy;
    rr = rr + dz * dz;
    if (j != s.i && rr <= s_rr[s.ti + (w & 3u)]) {
        if (s.c < s.cap) s.out[size_t(s.c) * WAVE] = j;
        ++s.c;
    }
}

__device__ __forceinline__ void scan_range(nb_state& s, unsigned b, unsigned e, float4 const* __restrict__ cell_pos,
                                           neighbour_params const& p, float const* s_rr) {
    unsigned k = b;
    for (; k + 4 <= e; k += 4) {
        float4 const r0 = cell_pos[k], r1 = cell_pos[k + 1], r2 = cell_pos[k + 2], r3 = cell_pos[k + 3];
        test_candidate(s, r0, p, s_rr);
        test_candidate(s, r1, p, s_rr);
        test_candidate(s, r2, p, s_rr);
        test_candidate(s, r3, p, s_rr);
    }
    for (; k < e; ++k) test_candidate(s, cell_pos[k], p, s_rr);
}

__global__ __launch_bounds__(256) void build_neighbours_kernel(
    float4 const* __restrict__ pos, unsigned n, uint32_t const* __restrict__ offset,
    float4 const* __restrict__ cell_pos, neighbour_params p, unsigned cap, uint32_t* __restrict__ list,
    uint32_t* __restrict__ count, uint32_t* __restrict__ max_count)
{
    __shared__ float s_rr[MAX_TYPES * MAX_TYPES];
    if (threadIdx.x < MAX_TYPES * MAX_TYPES) s_rr[threadIdx.x] = p.rr_cut_skin[threadIdx.x];
    __syncthreads();

    unsigned const i = blockIdx.x * blockDim.x + threadIdx.x;
    nb_state s;
    s.c = 0;
    if (i < n) {
        s.ri = pos[2 * size_t(i)];
        s.i = i;
        s.ti = __float_as_uint(s.ri.w) * p.ntype;
        s.cap = cap;
        s.out = list + size_t(i & ~(WAVE - 1)) * cap + (i & (WAVE - 1));
        unsigned cx, cy, cz;
        cell_of(s.ri, p.grid, cx, cy, cz);
        unsigned const nx = p.grid.ncell[0], ny = p.grid.ncell[1], nz = p.grid.ncell[2];
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

void launch_compute_cell(float4 const* pos, unsigned n, cell_grid g, uint32_t* keys, uint32_t* vals, hipStream_t s) {
    if (!n) return;
    compute_cell_kernel<<<div_up(n, 256), 256, 0, s>>>(pos, n, g, keys, vals);
    HIP_CHECK(hipGetLastError());
}

void launch_cell_offsets(uint32_t const* keys, unsigned n, unsigned ncells, uint32_t* offset, hipStream_t s) {
    cell_offsets_kernel<<<div_up(size_t(n) + 1, 256), 256, 0, s>>>(keys, n, ncells, offset);
    HIP_CHECK(hipGetLastError());
}

void launch_gather_cell_pos(float4 const* pos, uint32_t const* vals, unsigned n, float4* cell_pos, hipStream_t s) {
    if (!n) return;
    gather_cell_pos_kernel<<<div_up(n, 256), 256, 0, s>>>(pos, vals, n, cell_pos);
    HIP_CHECK(hipGetLastError());
}

void launch_build_neighbours(float4 const* pos, unsigned n, uint32_t const* offset, float4 const* cell_pos,
                             neighbour_params const& p, unsigned cap, uint32_t* list, uint32_t* count,
                             uint32_t* max_count, hipStream_t s) {
    if (!n) return;
    build_neighbours_kernel<<<div_up(n, 256), 256, 0, s>>>(pos, n, offset, cell_pos, p, cap, list, count,
                                                           max_count);
    HIP_CHECK(hipGetLastError());
}

} // namespace halmd
