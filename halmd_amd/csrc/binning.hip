// Cell binning and neighbour-list construction for gfx950.
//
// Binning follows the reference device rule for the cell index
// (halmd/mdsim/gpu/binning_kernel.cu:34-75): idx = uint(r / len + ncell) mod
// ncell per component, x fastest.  Instead of fixed-size cell slots with an
// overflow/retry loop (binning.cpp:166-221) the cells are compressed ranges:
// a stable radix sort of (cell, particle) gives, per cell, the contiguous range
// [begin, end) of particles in ascending index order, and a compact copy of the
// positions in that order (cell_pos) so that the neighbour search reads each
// cell's particles as one contiguous run.
//
// Neighbour lists are full lists (i -> j and j -> i) with the reference
// device predicate rr <= (rc_ab + skin)^2 in FP32 on the high words of the
// positions (from_binning_kernel.cu:144-163) and its cell order: 13 pairs of
// opposite neighbour cells, then the own cell (from_binning_kernel.cu:231-259).
// The list is stored per wavefront of 64 particles, column-major inside the
// wavefront: entry k of particle i sits at (i / 64) * 64 * cap + k * 64 + i % 64,
// so one wave reading entry k touches 256 contiguous bytes.  A per-particle
// count replaces the placeholder terminator, and no memset of the list is
// needed.
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

__global__ void cell_bounds_kernel(uint32_t const* __restrict__ keys, unsigned n, uint32_t* __restrict__ begin,
                                   uint32_t* __restrict__ end) {
    unsigned k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    uint32_t c = keys[k];
    if (k == 0 || keys[k - 1] != c) begin[c] = k;
    if (k + 1 == n || keys[k + 1] != c) end[c] = k + 1;
}

__global__ void gather_cell_pos_kernel(float4 const* __restrict__ pos, uint32_t const* __restrict__ vals, unsigned n,
                                       float4* __restrict__ cell_pos) {
    unsigned k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    cell_pos[k] = pos[2 * size_t(vals[k])];
}

// neighbour cell offsets: 13 pairs (j, -j) for the offsets preceding zero in
// lexicographic (x, y, z) order, then the own cell
__constant__ signed char c_stencil[27][3] = {
    {-1, -1, -1}, {1, 1, 1},   {-1, -1, 0}, {1, 1, 0},   {-1, -1, 1}, {1, 1, -1},  {-1, 0, -1},
    {1, 0, 1},    {-1, 0, 0},  {1, 0, 0},   {-1, 0, 1},  {1, 0, -1},  {-1, 1, -1}, {1, -1, 1},
    {-1, 1, 0},   {1, -1, 0},  {-1, 1, 1},  {1, -1, -1}, {0, -1, -1}, {0, 1, 1},   {0, -1, 0},
    {0, 1, 0},    {0, -1, 1},  {0, 1, -1},  {0, 0, -1},  {0, 0, 1},   {0, 0, 0},
};

__global__ __launch_bounds__(256) void build_neighbours_kernel(
    float4 const* __restrict__ pos, unsigned n, uint32_t const* __restrict__ cell_begin,
    uint32_t const* __restrict__ cell_end, float4 const* __restrict__ cell_pos, uint32_t const* __restrict__ cell_val,
    neighbour_params p, unsigned cap, uint32_t* __restrict__ list, uint32_t* __restrict__ count,
    uint32_t* __restrict__ max_count)
{
#pragma clang fp contract(off)
    __shared__ float s_rr[MAX_TYPES * MAX_TYPES];
    if (threadIdx.x < MAX_TYPES * MAX_TYPES) s_rr[threadIdx.x] = p.rr_cut_skin[threadIdx.x];
    __syncthreads();

    unsigned const i = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned c = 0;
    if (i < n) {
        float4 const ri = pos[2 * size_t(i)];
        unsigned const ti = __float_as_uint(ri.w);
        unsigned cx, cy, cz;
        cell_of(ri, p.grid, cx, cy, cz);
        unsigned const nx = p.grid.ncell[0], ny = p.grid.ncell[1], nz = p.grid.ncell[2];
        uint32_t* const out = list + size_t(i & ~(WAVE - 1)) * cap + (i & (WAVE - 1));
        for (unsigned s = 0; s < 27; ++s) {
            unsigned const kx = (cx + nx + c_stencil[s][0]) % nx;
            unsigned const ky = (cy + ny + c_stencil[s][1]) % ny;
            unsigned const kz = (cz + nz + c_stencil[s][2]) % nz;
            unsigned const cell = (kz * ny + ky) * nx + kx;
            unsigned const b = cell_begin[cell], e = cell_end[cell];
            bool const self = (s == 26);
            for (unsigned k = b; k < e; ++k) {
                unsigned const j = cell_val[k];
                if (self && j == i) continue;
                float4 const rj = cell_pos[k];
                float dx = ri.x - rj.x, dy = ri.y - rj.y, dz = ri.z - rj.z;
                dx = dx - rintf(dx * p.invL[0]) * p.L[0];
                dy = dy - rintf(dy * p.invL[1]) * p.L[1];
                dz = dz - rintf(dz * p.invL[2]) * p.L[2];
                float rr = dx * dx;
                rr = rr + dy * dy;
                rr = rr + dz * dz;
                if (rr <= s_rr[ti * p.ntype + __float_as_uint(rj.w)]) {
                    if (c < cap) out[size_t(c) * WAVE] = j;
                    ++c;
                }
            }
        }
        count[i] = c;
    }
    // largest count of the wave -> one atomic per wave
    unsigned m = c;
    for (int off = WAVE / 2; off > 0; off >>= 1) m = max(m, (unsigned)__shfl_xor((int)m, off, WAVE));
    if ((threadIdx.x % WAVE) == 0 && m > 0) atomicMax(max_count, m);
}

} // namespace

void launch_compute_cell(float4 const* pos, unsigned n, cell_grid g, uint32_t* keys, uint32_t* vals, hipStream_t s) {
    if (!n) return;
    compute_cell_kernel<<<div_up(n, 256), 256, 0, s>>>(pos, n, g, keys, vals);
    HIP_CHECK(hipGetLastError());
}

void launch_cell_bounds(uint32_t const* keys, unsigned n, unsigned ncells, uint32_t* begin, uint32_t* end,
                        hipStream_t s) {
    launch_fill_u32(begin, 0, ncells, s);
    launch_fill_u32(end, 0, ncells, s);
    if (!n) return;
    cell_bounds_kernel<<<div_up(n, 256), 256, 0, s>>>(keys, n, begin, end);
    HIP_CHECK(hipGetLastError());
}

void launch_gather_cell_pos(float4 const* pos, uint32_t const* vals, unsigned n, float4* cell_pos, hipStream_t s) {
    if (!n) return;
    gather_cell_pos_kernel<<<div_up(n, 256), 256, 0, s>>>(pos, vals, n, cell_pos);
    HIP_CHECK(hipGetLastError());
}

void launch_build_neighbours(float4 const* pos, unsigned n, uint32_t const* cell_begin, uint32_t const* cell_end,
                             float4 const* cell_pos, uint32_t const* cell_val, neighbour_params const& p,
                             unsigned cap, uint32_t* list, uint32_t* count, uint32_t* max_count, hipStream_t s) {
    if (!n) return;
    build_neighbours_kernel<<<div_up(n, 256), 256, 0, s>>>(pos, n, cell_begin, cell_end, cell_pos, cell_val, p, cap,
                                                           list, count, max_count);
    HIP_CHECK(hipGetLastError());
}

} // namespace halmd
