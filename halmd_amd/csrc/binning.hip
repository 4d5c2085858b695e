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
// cell ascending particle index.
//
// List layout (blocked by wavefront, 16 entries per lane and block): entry k
// of particle i sits at (i / 64) * 64 * cap + (k / 16) * 1024 + (i % 64) * 16 +
// k % 16.  Each lane stages its next 16 entries in LDS and stores them as one
// 64-byte block (one whole L2 line), so the HBM write is whole lines rather
// than 4-byte stores scattered over many partially written lines; a wave
// reading one block reads 4 KiB contiguously.  A per-particle count replaces
// the placeholder terminator, and no memset of the list is needed.
#include "segment.hpp"

#include <cstring>
#include <vector>

namespace halmd {
namespace {

__device__ __forceinline__ unsigned cell_component(uint32_t u, unsigned nc) {
    return unsigned((uint64_t(u) * nc) >> 32);
}

// neighbour predicate on fixed-point records (cubic box: fixed-point units)
template <bool ISO>
__device__ __forceinline__ bool pair_inside(uint4 const ri, uint4 const rj, neighbour_params const& p, float cut) {
#pragma clang fp contract(off)
    if (ISO) {
        float const dx = float(int32_t(ri.x - rj.x));
        float const dy = float(int32_t(ri.y - rj.y));
        float const dz = float(int32_t(ri.z - rj.z));
        float rq = dx * dx;
        rq = __builtin_fmaf(dy, dy, rq);
        rq = __builtin_fmaf(dz, dz, rq);
        return rq <= cut;
    } else {
        float const dx = float(int32_t(ri.x - rj.x)) * p.q[0];
        float const dy = float(int32_t(ri.y - rj.y)) * p.q[1];
        float const dz = float(int32_t(ri.z - rj.z)) * p.q[2];
        float rr = dx * dx;
        rr = rr + dy * dy;
        rr = rr + dz * dz;
        return rr <= cut;
    }
}

__global__ void compute_cell_kernel(uint4 const* __restrict__ posq, unsigned n, cell_grid g,
                                    uint32_t* __restrict__ keys, uint32_t* __restrict__ vals, unsigned sub_bits) {
    unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint4 const r = posq[i];
    // (x sub-bin and cell: floor(u nx 2^b / 2^32) = cell 2^b + sub-bin)
    unsigned const xs = unsigned((uint64_t(r.x) * (g.ncell[0] << sub_bits)) >> 32);
    unsigned const y = cell_component(r.y, g.ncell[1]);
    unsigned const z = cell_component(r.z, g.ncell[2]);
    keys[i] = (((z * g.ncell[1] + y) * g.ncell[0]) << sub_bits) + xs;
    vals[i] = i;
}

// offset[c] = first k with (keys[k] >> shift) >= c, for c = 0 .. nbins (k = 0 .. n)
__global__ void cell_offsets_kernel(uint32_t const* __restrict__ keys, unsigned n, unsigned nbins,
                                    uint32_t* __restrict__ offset, unsigned shift) {
    unsigned const k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k > n) return;
    int const prev = k == 0 ? -1 : int(keys[k - 1] >> shift);
    int const cur = k == n ? int(nbins) : int(keys[k] >> shift);
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
    uint32_t* out;      // first block of this particle: wave base + lane * NB_BLOCK
    uint32_t* buf;      // this thread's NB_BLOCK staging words in LDS
};

// LDS staging stride per thread: odd, so that the 64 lanes of a wave hit
// distinct banks whatever their entry counts
constexpr unsigned NB_STAGE = NB_BLOCK + 1;

__device__ __forceinline__ void flush_block(nb_state& s, unsigned block) {
    if ((block + 1) * NB_BLOCK <= s.cap) {
        uint4* const dst = reinterpret_cast<uint4*>(s.out + size_t(block) * NB_BLOCK * WAVE);
#pragma unroll
        for (unsigned v = 0; v < NB_BLOCK / 4; ++v)
            dst[v] = make_uint4(s.buf[4 * v], s.buf[4 * v + 1], s.buf[4 * v + 2], s.buf[4 * v + 3]);
    }
}

template <bool ISO>
__device__ __forceinline__ void test_candidate(nb_state& s, uint4 const rj, neighbour_params const& p,
                                               float const* s_rr) {
    uint32_t const j = rj.w >> 2;
    if (j != s.i && pair_inside<ISO>(s.ri, rj, p, s_rr[s.ti + (rj.w & 3u)])) {
        s.buf[s.c % NB_BLOCK] = j;
        if (s.c % NB_BLOCK == NB_BLOCK - 1) flush_block(s, s.c / NB_BLOCK);
        ++s.c;
    }
}

template <bool ISO>
__device__ __forceinline__ void scan_range(nb_state& s, unsigned b, unsigned e, uint4 const* __restrict__ cell_pos,
                                           neighbour_params const& p, float const* s_rr) {
    unsigned k = b;
    for (; k + 4 <= e; k += 4) {
        uint4 const r0 = cell_pos[k], r1 = cell_pos[k + 1], r2 = cell_pos[k + 2], r3 = cell_pos[k + 3];
        test_candidate<ISO>(s, r0, p, s_rr);
        test_candidate<ISO>(s, r1, p, s_rr);
        test_candidate<ISO>(s, r2, p, s_rr);
        test_candidate<ISO>(s, r3, p, s_rr);
    }
    for (; k < e; ++k) test_candidate<ISO>(s, cell_pos[k], p, s_rr);
}

template <bool ISO>
__global__ __launch_bounds__(256) void build_neighbours_kernel(
    uint4 const* __restrict__ posq, unsigned n, uint32_t const* __restrict__ offset,
    uint4 const* __restrict__ cell_pos, neighbour_params p, unsigned cap, uint32_t* __restrict__ list,
    uint32_t* __restrict__ count, uint32_t* __restrict__ max_count, bool self)
{
    __shared__ float s_rr[MAX_TYPES * MAX_TYPES];
    __shared__ uint32_t s_buf[256 * NB_STAGE];
    if (threadIdx.x < MAX_TYPES * MAX_TYPES)
        s_rr[threadIdx.x] = ISO ? p.rr_cut_skin_q[threadIdx.x] : p.rr_cut_skin[threadIdx.x];
    __syncthreads();

    unsigned const i = blockIdx.x * blockDim.x + threadIdx.x;
    nb_state s;
    s.c = 0;
    if (i < n) {
        s.ri = posq[i];
        // the records of a second particle instance (cell_pos) are never
        // particle i itself: no index matches
        s.i = self ? i : 0xffffffffu;
        s.ti = (s.ri.w & 3u) * p.ntype;
        s.cap = cap;
        s.out = list + size_t(i & ~(WAVE - 1)) * cap + (i & (WAVE - 1)) * NB_BLOCK;
        s.buf = s_buf + threadIdx.x * NB_STAGE;
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
                    scan_range<ISO>(s, offset[row + xm], offset[row + xp + 1], cell_pos, p, s_rr);
                } else if (cx == 0) {
                    scan_range<ISO>(s, offset[row + xm], offset[row + xm + 1], cell_pos, p, s_rr);
                    scan_range<ISO>(s, offset[row], offset[row + 2], cell_pos, p, s_rr);
                } else {
                    scan_range<ISO>(s, offset[row + xm], offset[row + nx], cell_pos, p, s_rr);
                    scan_range<ISO>(s, offset[row], offset[row + 1], cell_pos, p, s_rr);
                }
            }
        }
        if (s.c % NB_BLOCK) flush_block(s, s.c / NB_BLOCK);   // last, partial block
        count[i] = s.c;
    }
    // largest count of the wave -> one atomic per wave
    unsigned m = s.c;
    for (int off = WAVE / 2; off > 0; off >>= 1) m = max(m, (unsigned)__shfl_xor((int)m, off, WAVE));
    if ((threadIdx.x % WAVE) == 0 && m > 0) atomicMax(max_count, m);
}

// --- all pairs (from_particle): waves over particles, ascending j -----------
// The workgroup stages the records in tiles of 256 through LDS; each of its
// waves takes AP_PPW particles, and for each lane l tests j = base + l against
// them.  A ballot of the hits gives each its rank, so every list holds the j
// in ascending order (the order of a sequential scan) while every step of a
// wave covers 64 candidates, and each tile read from L2 serves 4 AP_PPW
// particles.  Entry k goes to the blocked layout of the other builds,
// (i / 64) 64 cap + (k / 16) 1024 + (i % 64) 16 + k % 16; entries past the
// capacity are dropped and the count still says how many there were.
//
// With cond (a device-side rebuild decision): the kernel does nothing unless
// the displacement words cond.slots exceed cond.keep_bits; when it rebuilds,
// it also sets the reference positions r0[i] = hi(pos[i]) of the displacement
// check and workgroup 0 counts the rebuild.
constexpr unsigned AP_PPW = 2;                        // particles per wave
constexpr unsigned AP_PER_GROUP = 256 / WAVE * AP_PPW;   // particles per 256-thread workgroup

template <bool ISO>
__global__ __launch_bounds__(256) void build_all_pairs_kernel(uint4 const* __restrict__ posq, unsigned n,
                                                              neighbour_params p, unsigned cap,
                                                              uint32_t* __restrict__ list,
                                                              uint32_t* __restrict__ count,
                                                              uint32_t* __restrict__ max_count,
                                                              rebuild_condition cond) {
    __shared__ float s_rr[MAX_TYPES * MAX_TYPES];
    __shared__ uint4 s_tile[256];
    __shared__ unsigned s_max[256 / WAVE];
    unsigned const lane = threadIdx.x % WAVE, w = threadIdx.x / WAVE;
    unsigned const i0 = blockIdx.x * AP_PER_GROUP;
    if (cond.slots) {
        if (!displacement_aborts(cond.slots, cond.keep_bits)) return;   // the whole workgroup
        if (threadIdx.x < AP_PER_GROUP && i0 + threadIdx.x < n)
            cond.r0[i0 + threadIdx.x] = cond.pos[2 * size_t(i0 + threadIdx.x)];
        if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(cond.rebuilt, 1u);
    }
    if (threadIdx.x < MAX_TYPES * MAX_TYPES)
        s_rr[threadIdx.x] = ISO ? p.rr_cut_skin_q[threadIdx.x] : p.rr_cut_skin[threadIdx.x];
    uint64_t const below = (uint64_t(1) << lane) - 1u;
    unsigned ii[AP_PPW], ti[AP_PPW], c[AP_PPW];
    uint4 ri[AP_PPW];
#pragma unroll
    for (unsigned q = 0; q < AP_PPW; ++q) {
        ii[q] = i0 + w * AP_PPW + q;   // wave-uniform
        ri[q] = posq[ii[q] < n ? ii[q] : 0u];
        ti[q] = (ri[q].w & 3u) * p.ntype;
        c[q] = 0;
    }
    for (unsigned base = 0; base < n; base += 256) {
        __syncthreads();   // the previous tile is consumed (and s_rr written, first time)
        if (base + threadIdx.x < n) s_tile[threadIdx.x] = posq[base + threadIdx.x];
        __syncthreads();
        unsigned const m = min(256u, n - base);
        for (unsigned k = 0; k < m; k += WAVE) {
            bool const inb = k + lane < m;
            unsigned const j = base + k + lane;
            uint4 const rj = s_tile[inb ? k + lane : 0u];
#pragma unroll
            for (unsigned q = 0; q < AP_PPW; ++q) {
                bool const hit = inb && ii[q] < n && j != ii[q] &&
                                 pair_inside<ISO>(ri[q], rj, p, s_rr[ti[q] + (rj.w & 3u)]);
                uint64_t const mm = __ballot(hit);
                if (hit) {
                    unsigned const cc = c[q] + unsigned(__popcll(mm & below));
                    if (cc < cap)
                        list[size_t(ii[q] & ~(WAVE - 1)) * cap + (ii[q] & (WAVE - 1)) * NB_BLOCK +
                             size_t(cc / NB_BLOCK) * (NB_BLOCK * WAVE) + cc % NB_BLOCK] = j;
                }
                c[q] += unsigned(__popcll(mm));
            }
        }
    }
    unsigned cm = 0;
#pragma unroll
    for (unsigned q = 0; q < AP_PPW; ++q) {
        if (ii[q] < n) {
            if (lane == 0) count[ii[q]] = c[q];
            cm = max(cm, c[q]);
        }
    }
    // largest count of the workgroup -> one atomic
    if (lane == 0) s_max[w] = cm;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned mx = 0;
        for (unsigned v = 0; v < blockDim.x / WAVE; ++v) mx = max(mx, s_max[v]);
        if (mx > 0) atomicMax(max_count, mx);
    }
}

// --- cell segments: 16-bit slots into the LDS-staged neighbourhood union ----

// Byte offset of entry c of a lane from the workgroup's list base: H + 2048
// (c / 16) + 32 lane + 2 (c % 16), with H (a multiple of 2048) the lane's
// wave block row 0.  The counter q keeps it with the lane field (bits 5..10)
// set to ones, q = H + 2048 (c / 16) + 0x7e0 + 2 (c % 16): adding 2 then
// carries from the entry field straight into the block field, and one bit
// field insert of the lane bits gives the address.
constexpr uint32_t NB16_LANE_BITS = 0x7e0u;

struct nb16_state {
    uint4 ri;
    unsigned own, ti;
    unsigned q, q0, qmax;   // counter; its value at c = 0 and at c = cap - 1
    unsigned lane32;        // 32 (lane % 64)
    char* list;             // the workgroup's list base (uniform)
};

__device__ __forceinline__ unsigned nb16_count(nb16_state const& s) {
    return (((s.q >> 11) - (s.q0 >> 11)) << 4) + ((s.q & 31u) >> 1);
}

// 2-byte stores straight to the blocked list: a lane fills one 32-byte block
// at a time and a wave's blocks share 2 KiB, so the lines are complete in L2
// before they are written back (no staging buffer, hence more workgroups per
// CU; gathering entries in LDS and storing whole blocks measured 6% slower).
// Entries past the capacity overwrite the particle's last slot instead of
// branching: the count still exceeds the capacity, and the caller rebuilds
// larger lists.
__device__ __forceinline__ void append16(nb16_state& s, unsigned entry) {
    unsigned const qa = min(s.q, s.qmax);
    unsigned const o = (s.lane32 & NB16_LANE_BITS) | (qa & ~NB16_LANE_BITS);
    *reinterpret_cast<uint16_t*>(s.list + o) = uint16_t(entry);
    s.q = (s.q + 2u) | NB16_LANE_BITS;
}

// the list entry is the record's byte offset in the full-record union (16
// bytes per record, the layout the force kernel stages)

// non-cubic box: FP32 squared distance of the scaled fixed-point differences
template <bool MULTI, bool SELF>
__device__ __forceinline__ void test16(nb16_state& s, uint4 const rj, unsigned slot, neighbour_params const& p,
                                       uint32_t const* s_rr, uint32_t rr0) {
#pragma clang fp contract(off)
    float const dx = float(int32_t(s.ri.x - rj.x)) * p.q[0];
    float const dy = float(int32_t(s.ri.y - rj.y)) * p.q[1];
    float const dz = float(int32_t(s.ri.z - rj.z)) * p.q[2];
    float rr = dx * dx;
    rr = rr + dy * dy;
    rr = rr + dz * dz;
    bool const inside = rr <= __uint_as_float(MULTI ? s_rr[s.ti + (rj.w & 3u)] : rr0);
    if ((!SELF || slot != s.own) && inside) append16(s, slot * 16u);
}

// cubic box, packed records: exact integer squared distance of the 16-bit
// differences of the top halves (wrapped, hence the minimum image), at most
// 3 * 2^30 and compared unsigned; s.ri.x and s.ri.y hold the particle's own
// packed record
typedef short short2v __attribute__((ext_vector_type(2)));
typedef unsigned short ushort2v __attribute__((ext_vector_type(2)));
template <bool MULTI, bool SELF>
__device__ __forceinline__ void test16p(nb16_state& s, uint2 const rj, unsigned slot, uint32_t const* s_rr,
                                        uint32_t rr0) {
    ushort2v const dxy = __builtin_bit_cast(ushort2v, s.ri.x) - __builtin_bit_cast(ushort2v, rj.x);
    int const dz = short(s.ri.y - rj.y);
    short2v const d = __builtin_bit_cast(short2v, dxy);
    uint32_t const rq = uint32_t(__builtin_amdgcn_sdot2(d, d, dz * dz, false));
    bool const inside = rq <= (MULTI ? s_rr[s.ti + (rj.y >> 16)] : rr0);
    if ((!SELF || slot != s.own) && inside) append16(s, slot * 16u);
}

// the 16-bit predicate of test16p alone
template <bool MULTI>
__device__ __forceinline__ bool pass16(nb16_state const& s, uint2 const rj, uint32_t const* s_rr, uint32_t rr0) {
    ushort2v const dxy = __builtin_bit_cast(ushort2v, s.ri.x) - __builtin_bit_cast(ushort2v, rj.x);
    int const dz = short(s.ri.y - rj.y);
    short2v const d = __builtin_bit_cast(short2v, dxy);
    uint32_t const rq = uint32_t(__builtin_amdgcn_sdot2(d, d, dz * dz, false));
    return rq <= (MULTI ? s_rr[s.ti + (rj.y >> 16)] : rr0);
}

// candidates [t, e) of one row, four records in flight
template <bool MULTI, bool SELF>
__device__ __forceinline__ void scan16(nb16_state& s, unsigned t, unsigned const e, uint4 const* s_rec,
                                       neighbour_params const& p, uint32_t const* s_rr, uint32_t rr0) {
    for (; t + 4 <= e; t += 4) {
        uint4 const r0 = s_rec[t], r1 = s_rec[t + 1], r2 = s_rec[t + 2], r3 = s_rec[t + 3];
        test16<MULTI, SELF>(s, r0, t, p, s_rr, rr0);
        test16<MULTI, SELF>(s, r1, t + 1, p, s_rr, rr0);
        test16<MULTI, SELF>(s, r2, t + 2, p, s_rr, rr0);
        test16<MULTI, SELF>(s, r3, t + 3, p, s_rr, rr0);
    }
    for (; t < e; ++t) test16<MULTI, SELF>(s, s_rec[t], t, p, s_rr, rr0);
}
// cubic box, packed records: the hits of up to 32 candidates gathered in a
// bit mask with no branch, then appended in ascending candidate order: one
// store instruction per round of the wave's remaining hits instead of one per
// candidate.  Each candidate costs its test and one v_addc_co_u32 (m = 2 m +
// the lane's bit of the compare's lane mask, the mask as carry-in).
// Candidates go in groups of 4; a group may read up to 3 records past the
// row's end (the launches give every workgroup at least 4 records of LDS past
// its staging, the sentinel's slot and the cell-start table after it
// included), whose bits one mask per chunk clears, as it clears the
// particle's own bit.

// m = 2 m + (this lane's bit of lanes)
__device__ __forceinline__ uint32_t shift_in(uint32_t m, uint64_t lanes) {
    uint32_t r;
    uint64_t carry;
    asm("v_addc_co_u32_e64 %0, %1, %2, %2, %3" : "=v"(r), "=s"(carry) : "v"(m), "s"(lanes));
    return r;
}

// the lanes whose candidate rj passes the test of pass16 (compare lane mask)
template <bool MULTI>
__device__ __forceinline__ uint64_t pass16_lanes(nb16_state const& s, uint2 const rj, uint32_t const* s_rr,
                                                 uint32_t rr0) {
    ushort2v const dxy = __builtin_bit_cast(ushort2v, s.ri.x) - __builtin_bit_cast(ushort2v, rj.x);
    int const dz = short(s.ri.y - rj.y);
    short2v const d = __builtin_bit_cast(short2v, dxy);
    uint32_t const rq = uint32_t(__builtin_amdgcn_sdot2(d, d, dz * dz, false));
    uint32_t const thr = MULTI ? s_rr[s.ti + (rj.y >> 16)] : rr0;
    return __builtin_amdgcn_uicmp(rq, thr, 37);   // 37: unsigned less or equal
}

template <bool MULTI, bool SELF>
__device__ __forceinline__ void scan16(nb16_state& s, unsigned t, unsigned const e, uint2 const* s_rec,
                                       neighbour_params const&, uint32_t const* s_rr, uint32_t rr0) {
    while (t < e) {
        unsigned const n = min(e - t, 32u);
        uint32_t m = 0;
        unsigned k = 0;
        for (; k < n; k += 4) {
            uint2 const r0 = s_rec[t + k], r1 = s_rec[t + k + 1], r2 = s_rec[t + k + 2], r3 = s_rec[t + k + 3];
            m = shift_in(m, pass16_lanes<MULTI>(s, r0, s_rr, rr0));
            m = shift_in(m, pass16_lanes<MULTI>(s, r1, s_rr, rr0));
            m = shift_in(m, pass16_lanes<MULTI>(s, r2, s_rr, rr0));
            m = shift_in(m, pass16_lanes<MULTI>(s, r3, s_rr, rr0));
        }
        // k (n rounded up to whole groups) bits shifted in: candidate t + c
        // at bit k - 1 - c, reversed at bit 32 - k + c = c - base (mod 2^32)
        m = __builtin_bitreverse32(m);
        unsigned const base = t + k - 32u;
        m &= ~0u >> (k - n);   // candidates t + n .. t + k - 1 (k - n <= 3)
        if (SELF) {
            unsigned const ob = s.own - base;   // below 32 only inside the chunk or its empty low bits
            m &= ob < 32u ? ~(1u << ob) : ~0u;
        }
        while (m) {
            append16(s, (base + unsigned(__builtin_ctz(m))) * 16u);
            m &= m - 1u;
        }
        t += n;
    }
}

// One cell segment of the staged list build (the whole workgroup calls it);
// shared by the single-system kernel and the ensemble kernel.  The union is
// staged as packed 16-bit records in a cubic box (8 bytes per record of the
// dynamic LDS), else as full records (16 bytes).  stats[0] receives the
// largest count, stats[1] the staging size U + 1.
template <bool MULTI, bool ISO>
__device__ __forceinline__ void staged_build_segment(
    uint4 const* __restrict__ posq, uint32_t const* __restrict__ off, seg_params const& sp, unsigned seg,
    neighbour_params const& p, unsigned cap, uint16_t* __restrict__ list, uint32_t* __restrict__ count,
    uint32_t* __restrict__ stats, uint4* s_lds, seg_stage& st, unsigned* s_bound, uint32_t* s_rr,
    unsigned bstride = SEG_BOUND, uint32_t* overflow = nullptr)
{
    using rec_t = typename std::conditional<ISO, uint2, uint4>::type;
    rec_t* const s_rec = reinterpret_cast<rec_t*>(s_lds);
    if (threadIdx.x < MAX_TYPES * MAX_TYPES)
        s_rr[threadIdx.x] = ISO ? p.rr16[threadIdx.x] : __float_as_uint(p.rr_cut_skin[threadIdx.x]);
    unsigned const U = stage_union(sp, seg, off, posq, s_rec, st, s_bound, bstride);
    if (threadIdx.x == 0) atomicMax(&stats[1], U + 1);
    if (U + 1 > sp.lds_cap) {   // uniform: the caller enlarges the staging and rebuilds
        if (overflow && threadIdx.x == 0) atomicOr(overflow, ENS_SKIP_OVERFLOW);
        return;
    }

    unsigned const nx = p.grid.ncell[0];
    uint32_t const rr0 = ISO ? p.rr16[0] : __float_as_uint(p.rr_cut_skin[0]);
    unsigned const hh = st.q.h + 2, T = st.own_pre[st.q.h];
    // the segment's particles follow the first own row in memory (its rows
    // y0 .. y0 + h - 1 do not wrap), so the list offsets from the lane block
    // of its first particle fit 32 bits (particles of one z-plane of cells)
    unsigned const i0 = st.own_begin[0] & ~(WAVE - 1);
    char* const wlist = reinterpret_cast<char*>(list + size_t(i0) * cap);
    unsigned cmax = 0;
    for (unsigned t = threadIdx.x; t < T; t += blockDim.x) {
        unsigned j, i;
        seg_own(st, t, j, i);
        unsigned const self = hh + j + 1;   // union row of the particle's own (y, z) row
        nb16_state s;
        s.own = s_bound[self * bstride + 1] + (t - st.own_pre[j]);
        uint4 const full = posq[i];          // = the staged record s.own, in full
        if (ISO) {
            uint2 const pk = pack_rec16(full);
            s.ri = make_uint4(pk.x, pk.y, 0u, 0u);
        } else {
            s.ri = full;
        }
        s.ti = (full.w & 3u) * p.ntype;
        unsigned const H = 2 * ((i & ~(WAVE - 1)) - i0) * cap;   // multiple of 2048
        s.q0 = H | NB16_LANE_BITS;
        s.q = s.q0;
        s.qmax = s.q0 + (((cap - 1) >> 4) << 11) + (((cap - 1) & 15u) << 1);
        s.lane32 = (i & (WAVE - 1)) * (2 * NB16_BLOCK);
        s.list = wlist;
        unsigned const k = cell_component(full.x, nx) - st.q.x0;   // 0 .. w - 1
        // the 9 rows (dz, dy) in lexicographic order; only the own row holds
        // the particle itself
        for (unsigned dz = 0; dz < 3; ++dz)
            for (unsigned dy = 0; dy < 3; ++dy) {
                unsigned const ur = dz * hh + j + dy;
                unsigned const b = s_bound[ur * bstride + k];
                unsigned const e = s_bound[ur * bstride + k + 3];
                if (ur == self)
                    scan16<MULTI, true>(s, b, e, s_rec, p, s_rr, rr0);
                else
                    scan16<MULTI, false>(s, b, e, s_rec, p, s_rr, rr0);
            }
        unsigned const c = nb16_count(s);
        // pad with the sentinel to a multiple of 8 (cap is a multiple of 16)
        if (c <= cap)
            while (s.q & 15u) append16(s, U * 16u);
        count[i] = c;
        cmax = max(cmax, c);
    }
    for (int o = WAVE / 2; o > 0; o >>= 1) cmax = max(cmax, (unsigned)__shfl_xor((int)cmax, o, WAVE));
    if ((threadIdx.x % WAVE) == 0 && cmax > 0) atomicMax(&stats[0], cmax);
    if (overflow && (threadIdx.x % WAVE) == 0 && cmax > cap) atomicOr(overflow, ENS_SKIP_OVERFLOW);
}

template <bool MULTI, bool ISO>
__global__ __launch_bounds__(SEG_THREADS_MAX) void build_staged_kernel(
    uint4 const* __restrict__ posq, uint32_t const* __restrict__ off, seg_params sp, neighbour_params p,
    unsigned cap, uint16_t* __restrict__ list, uint32_t* __restrict__ count, uint32_t* __restrict__ stats)
{
    extern __shared__ uint4 s_rec[];
    __shared__ seg_stage st;
    __shared__ uint32_t s_rr[MAX_TYPES * MAX_TYPES];

    unsigned const seg = seg_of_block(sp, blockIdx.x);
    if (seg >= sp.nseg) return;   // the whole workgroup
    // the cell starts of the union rows follow the records in the dynamic LDS
    // (3 (R + 2) rows of S + 3 words), sized to the segment shape
    unsigned* const s_bound = reinterpret_cast<unsigned*>(reinterpret_cast<char*>(s_rec) +
                                                          size_t(sp.lds_cap) * (ISO ? sizeof(uint2) : sizeof(uint4)));
    staged_build_segment<MULTI, ISO>(posq, off, sp, seg, p, cap, list, count, stats, s_rec, st, s_bound, s_rr,
                                     sp.S + 3);
}

// --- cell segments with x sub-bins: per-particle candidate windows ---------
//
// The particle arrays are in (cell, sub-bin) order: each cell is cut along x
// into SUB_BINS slabs, sub-bin = (u_x * nx * SUB_BINS) >> 32, and a stable
// sort puts a cell's particles slab by slab.  A union row is then sorted by
// sub-bin, and the records of any x interval form one contiguous range of it.
// For each of its 9 rows (dz, dy) a lane takes the smallest |dy|, |dz| (in
// 16-bit units) that a particle of that row can have, the largest |dx| that
// can still pass the 16-bit predicate, and the sub-bins that interval covers:
// its window, a range of the row's 3-cell scan that holds every candidate that
// can pass.  The lists are those of staged_build_segment entry for entry (same
// predicate, same ascending slots per row), over about half the candidates.
// The own row's window is split at the particle itself.  The windows (at most
// WIN_MAX per lane) go to LDS, and the lane walks them as one sequence, four
// candidates per step, so that every lane of a wave is busy until the wave's
// longest sequence ends (rows are no longer the unit of divergence).
//
// Dynamic LDS: lds_cap + 3 packed records (reads run up to 3 records past a
// window's end), then the sub-bin starts of the union rows (3 (R + 2) rows of
// (S + 2) SUB_BINS + 1 words), then WIN_MAX words per thread.
__host__ __device__ inline unsigned windowed_sub_stride(unsigned S) { return (S + 2) * SUB_BINS + 1; }

template <bool MULTI>
__device__ __forceinline__ void windowed_build_segment(
    uint4 const* __restrict__ posq, uint32_t const* __restrict__ off, uint32_t const* __restrict__ soff,
    seg_params const& sp, unsigned seg, neighbour_params const& p, unsigned cap, uint16_t* __restrict__ list,
    uint32_t* __restrict__ count, uint32_t* __restrict__ stats, uint4* s_lds, seg_stage& st, unsigned* s_bound,
    uint32_t* s_rr, unsigned* s_yz)
{
    uint2* const s_rec = reinterpret_cast<uint2*>(s_lds);
    if (threadIdx.x < MAX_TYPES * MAX_TYPES) s_rr[threadIdx.x] = p.rr16[threadIdx.x];
    unsigned const U = stage_union(sp, seg, off, posq, s_rec, st, s_bound);
    if (threadIdx.x == 0) atomicMax(&stats[1], U + 1);
    if (U + 1 > sp.lds_cap) return;   // uniform: the caller enlarges the staging and rebuilds

    seg_place const q = st.q;
    unsigned const nx = sp.grid.ncell[0], ny = sp.grid.ncell[1], nz = sp.grid.ncell[2];
    unsigned const hh = q.h + 2, nr = st.nr;
    unsigned const sstride = windowed_sub_stride(sp.S);
    unsigned* const s_sub = reinterpret_cast<unsigned*>(s_rec + sp.lds_cap + 3);
    uint32_t* const s_win = s_sub + 3 * (sp.R + 2) * sstride;
    // LDS start of sub-bin b of relative cell c (cells x0 - 1 .. x0 + w) of
    // each union row: the cell's start plus the records of its sub-bins below b
    unsigned const wsub = (q.w + 2) * SUB_BINS;
    for (unsigned e = threadIdx.x; e < nr * (wsub + 1); e += blockDim.x) {
        unsigned const ur = e / (wsub + 1), rs = e % (wsub + 1), c = rs / SUB_BINS;
        unsigned v = s_bound[ur * SEG_BOUND + c];
        if (c < q.w + 2) {
            unsigned const sb = (seg_row_start(sp, q, ur) + (q.x0 + nx - 1 + c) % nx) * SUB_BINS;
            v += soff[sb + rs % SUB_BINS] - soff[sb];
        }
        s_sub[ur * sstride + rs] = v;
    }
    // 16-bit lower boundaries floor(c 2^16 / n) of the segment's rows y0 .. y0
    // + h (h + 1 values) and of its plane z and the next; a particle of cell c
    // has u >> 16 in [floor(c 2^16 / n), floor((c + 1) 2^16 / n)]
    if (threadIdx.x <= q.h) s_yz[threadIdx.x] = ((q.y0 + threadIdx.x) << 16) / ny;
    if (threadIdx.x == q.h + 1) s_yz[q.h + 1] = (q.z << 16) / nz;
    if (threadIdx.x == q.h + 2) s_yz[q.h + 2] = ((q.z + 1) << 16) / nz;
    __syncthreads();

    uint32_t const rr0 = p.rr16[0];
    unsigned const T = st.own_pre[q.h];
    unsigned const nxs = nx * SUB_BINS;
    unsigned const i0 = st.own_begin[0] & ~(WAVE - 1);
    char* const wlist = reinterpret_cast<char*>(list + size_t(i0) * cap);
    unsigned const zlo = s_yz[q.h + 1], zhi = s_yz[q.h + 2];
    unsigned cmax = 0;
    for (unsigned t = threadIdx.x; t < T; t += blockDim.x) {
        unsigned j, i;
        seg_own(st, t, j, i);
        unsigned const self = hh + j + 1;
        nb16_state s;
        s.own = s_bound[self * SEG_BOUND + 1] + (t - st.own_pre[j]);
        uint4 const full = posq[i];
        uint2 const pk = pack_rec16(full);
        s.ri = make_uint4(pk.x, pk.y, 0u, 0u);
        s.ti = (full.w & 3u) * p.ntype;
        unsigned const H = 2 * ((i & ~(WAVE - 1)) - i0) * cap;
        s.q0 = H | NB16_LANE_BITS;
        s.q = s.q0;
        s.qmax = s.q0 + (((cap - 1) >> 4) << 11) + (((cap - 1) & 15u) << 1);
        s.lane32 = (i & (WAVE - 1)) * (2 * NB16_BLOCK);
        s.list = wlist;

        uint32_t rrmax = rr0;
        if (MULTI)
            for (unsigned b = 0; b < p.ntype; ++b) rrmax = max(rrmax, s_rr[s.ti + b]);
        unsigned const x16 = full.x >> 16, y16 = full.y >> 16, z16 = full.z >> 16;
        // smallest |d| to the rows below / above and the planes below / above
        int const dyl = int(y16 - s_yz[j]), dyh = int(s_yz[j + 1] - y16);
        int const dzl = int(z16 - zlo), dzh = int(zhi - z16);
        unsigned const sbi = __umulhi(full.x, nxs);                  // own sub-bin (absolute)
        int const k = int(sbi / SUB_BINS) - int(q.x0);                // own cell - 1, relative to the row
        int const rsi = (k + 1) * int(SUB_BINS) + int(sbi % SUB_BINS);
        unsigned nw = 0;
        for (unsigned dz = 0; dz < 3; ++dz) {
            int const mz = dz == 0 ? dzl : (dz == 2 ? dzh : 0);
            for (unsigned dy = 0; dy < 3; ++dy) {
                int const my = dy == 0 ? dyl : (dy == 2 ? dyh : 0);
                int const rem = int(rrmax) - my * my - mz * mz;
                if (rem < 0) continue;   // no particle of this row can pass
                unsigned const rx = unsigned(__builtin_sqrtf(float(rem))) + 1u;
                uint32_t const ulo = ((x16 - rx) & 0xffffu) << 16, uhi = (((x16 + rx) & 0xffffu) << 16) | 0xffffu;
                int dlo = int(sbi) - int(__umulhi(ulo, nxs));
                int dhi = int(__umulhi(uhi, nxs)) - int(sbi);
                dlo += dlo < 0 ? int(nxs) : 0;   // the interval wraps across the box edge
                dhi += dhi < 0 ? int(nxs) : 0;
                int const lo = max(rsi - dlo, k * int(SUB_BINS));
                int const hi = min(rsi + dhi, (k + 3) * int(SUB_BINS) - 1);
                unsigned const ur = dz * hh + j + dy;
                unsigned const b = s_sub[ur * sstride + lo], e = s_sub[ur * sstride + hi + 1];
                if (ur == self) {
                    if (b < s.own) s_win[nw++ * blockDim.x + threadIdx.x] = b | (s.own << 16);
                    if (s.own + 1 < e) s_win[nw++ * blockDim.x + threadIdx.x] = (s.own + 1) | (e << 16);
                } else if (b < e) {
                    s_win[nw++ * blockDim.x + threadIdx.x] = b | (e << 16);
                }
            }
        }
        // the windows as one sequence, four candidates per step
        unsigned r = 0, c0 = 0, ce = 0;
        if (nw) {
            uint32_t const wv = s_win[threadIdx.x];
            c0 = wv & 0xffffu;
            ce = wv >> 16;
        }
        while (r < nw) {
            uint2 const a0 = s_rec[c0], a1 = s_rec[c0 + 1], a2 = s_rec[c0 + 2], a3 = s_rec[c0 + 3];
            // the four predicates first (no branches between them), then the appends
            bool const h0 = pass16<MULTI>(s, a0, s_rr, rr0);
            bool const h1 = pass16<MULTI>(s, a1, s_rr, rr0) && c0 + 1 < ce;
            bool const h2 = pass16<MULTI>(s, a2, s_rr, rr0) && c0 + 2 < ce;
            bool const h3 = pass16<MULTI>(s, a3, s_rr, rr0) && c0 + 3 < ce;
            if (h0) append16(s, c0 * 16u);
            if (h1) append16(s, c0 * 16u + 16u);
            if (h2) append16(s, c0 * 16u + 32u);
            if (h3) append16(s, c0 * 16u + 48u);
            c0 += 4;
            if (c0 >= ce) {
                ++r;
                uint32_t const wv = s_win[min(r, WIN_MAX - 1) * blockDim.x + threadIdx.x];
                c0 = wv & 0xffffu;
                ce = wv >> 16;
            }
        }
        unsigned const c = nb16_count(s);
        if (c <= cap)
            while (s.q & 15u) append16(s, U * 16u);
        count[i] = c;
        cmax = max(cmax, c);
    }
    for (int o = WAVE / 2; o > 0; o >>= 1) cmax = max(cmax, (unsigned)__shfl_xor((int)cmax, o, WAVE));
    if ((threadIdx.x % WAVE) == 0 && cmax > 0) atomicMax(&stats[0], cmax);
}

template <bool MULTI>
__global__ __launch_bounds__(SEG_THREADS_MAX) void build_windowed_kernel(
    uint4 const* __restrict__ posq, uint32_t const* __restrict__ off, uint32_t const* __restrict__ soff,
    seg_params sp, neighbour_params p, unsigned cap, uint16_t* __restrict__ list, uint32_t* __restrict__ count,
    uint32_t* __restrict__ stats)
{
    extern __shared__ uint4 s_rec[];
    __shared__ seg_stage st;
    __shared__ unsigned s_bound[SEG_NR * SEG_BOUND];
    __shared__ uint32_t s_rr[MAX_TYPES * MAX_TYPES];
    __shared__ unsigned s_yz[SEG_RMAX + 3];

    unsigned const seg = seg_of_block(sp, blockIdx.x);
    if (seg >= sp.nseg) return;   // the whole workgroup
    windowed_build_segment<MULTI>(posq, off, soff, sp, seg, p, cap, list, count, stats, s_rec, st, s_bound, s_rr,
                                  s_yz);
}

// --- ensembles: rebuild of a listed subset of systems --------------------

__device__ __forceinline__ unsigned cell_of(uint4 const r, cell_grid const& g) {
    unsigned const x = cell_component(r.x, g.ncell[0]);
    unsigned const y = cell_component(r.y, g.ncell[1]);
    unsigned const z = cell_component(r.z, g.ncell[2]);
    return (z * g.ncell[1] + y) * g.ncell[0] + x;
}

__global__ __launch_bounds__(256) void ens_cells_kernel(uint4 const* __restrict__ posq,
                                                        ens_desc const* __restrict__ descs,
                                                        uint32_t const* __restrict__ rsys,
                                                        uint32_t const* __restrict__ rpos,
                                                        uint32_t const* __restrict__ rblk, unsigned cell_bits,
                                                        uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    uint32_t const e = rblk[blockIdx.x];
    unsigned const rank = e >> 16, lb = e & 0xffffu;
    ens_desc const& d = descs[rsys[rank]];
    unsigned const li = lb * 256 + threadIdx.x;
    if (li >= d.n) return;
    unsigned const c = rpos[rank] + li;
    keys[c] = (rank << cell_bits) | cell_of(posq[d.base + li], d.np.grid);
    vals[c] = li;
}

// offset[c] = first slot of cell c in each listed system's sorted keys: the
// particle at slot k writes the cells after its predecessor's up to its own
// (the last particle also those after it, with n), as cell_offsets_kernel does
__global__ __launch_bounds__(256) void ens_offsets_kernel(ens_desc const* __restrict__ descs,
                                                          uint32_t const* __restrict__ rsys,
                                                          uint32_t const* __restrict__ rpos,
                                                          uint32_t const* __restrict__ rblk, unsigned cell_bits,
                                                          uint32_t const* __restrict__ keys,
                                                          uint32_t* __restrict__ offsets) {
    uint32_t const e = rblk[blockIdx.x];
    unsigned const rank = e >> 16, lb = e & 0xffffu;
    ens_desc const& d = descs[rsys[rank]];
    unsigned const k = lb * 256 + threadIdx.x;
    if (k >= d.n) return;
    uint32_t const mask = (1u << cell_bits) - 1u;
    unsigned const lo = rpos[rank];
    int const prev = k == 0 ? -1 : int(keys[lo + k - 1] & mask);
    int const cur = int(keys[lo + k] & mask);
    uint32_t* const off = offsets + d.cell_base;
    for (int c = prev + 1; c <= cur; ++c) off[c] = k;
    if (k + 1 == d.n)
        for (int c = cur + 1; c <= int(d.ncells); ++c) off[c] = d.n;
}

__global__ __launch_bounds__(256) void ens_gather_kernel(
    float4 const* __restrict__ pos, float4 const* __restrict__ vel, int4 const* __restrict__ image,
    uint32_t const* __restrict__ id, uint4 const* __restrict__ posq, float4* __restrict__ pos_t,
    float4* __restrict__ vel_t, int4* __restrict__ image_t, uint32_t* __restrict__ id_t, uint4* __restrict__ posq_t,
    ens_desc const* __restrict__ descs, uint32_t const* __restrict__ rsys, uint32_t const* __restrict__ rpos,
    uint32_t const* __restrict__ rblk, uint32_t const* __restrict__ vals) {
    uint32_t const e = rblk[blockIdx.x];
    unsigned const rank = e >> 16, lb = e & 0xffffu;
    ens_desc const& d = descs[rsys[rank]];
    unsigned const li = lb * 256 + threadIdx.x;
    if (li >= d.n) return;
    size_t const dst = d.base + li, src = d.base + vals[rpos[rank] + li];
    pos_t[2 * dst] = pos[2 * src];
    pos_t[2 * dst + 1] = pos[2 * src + 1];
    vel_t[2 * dst] = vel[2 * src];
    vel_t[2 * dst + 1] = vel[2 * src + 1];
    image_t[dst] = image[src];
    id_t[dst] = id[src];
    posq_t[dst] = posq[src];
}

__global__ __launch_bounds__(256) void ens_copyback_kernel(
    float4* __restrict__ pos, float4* __restrict__ vel, int4* __restrict__ image, uint32_t* __restrict__ id,
    uint4* __restrict__ posq, float4* __restrict__ r0, float4 const* __restrict__ pos_t,
    float4 const* __restrict__ vel_t, int4 const* __restrict__ image_t, uint32_t const* __restrict__ id_t,
    uint4 const* __restrict__ posq_t, ens_desc const* __restrict__ descs, uint32_t const* __restrict__ rsys,
    uint32_t const* __restrict__ rblk) {
    uint32_t const e = rblk[blockIdx.x];
    unsigned const rank = e >> 16, lb = e & 0xffffu;
    ens_desc const& d = descs[rsys[rank]];
    unsigned const li = lb * 256 + threadIdx.x;
    if (li >= d.n) return;
    size_t const i = d.base + li;
    float4 const h = pos_t[2 * i];
    pos[2 * i] = h;
    pos[2 * i + 1] = pos_t[2 * i + 1];
    vel[2 * i] = vel_t[2 * i];
    vel[2 * i + 1] = vel_t[2 * i + 1];
    image[i] = image_t[i];
    id[i] = id_t[i];
    posq[i] = posq_t[i];
    r0[i] = h;
}

// entries (s << 16) | block
__global__ __launch_bounds__(256) void ens_r0_kernel(float4 const* __restrict__ pos, float4* __restrict__ r0,
                                                     ens_desc const* __restrict__ descs,
                                                     uint32_t const* __restrict__ rblk) {
    uint32_t const e = rblk[blockIdx.x];
    ens_desc const& d = descs[e >> 16];
    unsigned const li = (e & 0xffffu) * 256 + threadIdx.x;
    if (li < d.n) r0[d.base + li] = pos[2 * size_t(d.base + li)];
}

template <bool MULTI>
__global__ __launch_bounds__(SEG_THREADS_MAX) void ens_build_kernel(
    uint4 const* __restrict__ posq, uint32_t const* __restrict__ offsets, ens_desc const* __restrict__ descs,
    uint32_t const* __restrict__ rseg, uint16_t* __restrict__ list, uint32_t* __restrict__ count,
    uint32_t* __restrict__ stats)
{
    extern __shared__ uint4 s_rec[];
    __shared__ seg_stage st;
    __shared__ unsigned s_bound[SEG_NR * SEG_BOUND];
    __shared__ uint32_t s_rr[MAX_TYPES * MAX_TYPES];
    uint32_t const e = rseg[blockIdx.x];
    unsigned const sys = e >> 16, seg = e & 0xffffu;
    ens_desc const& d = descs[sys];
    staged_build_segment<MULTI, true>(posq + d.base, offsets + d.cell_base, d.sp, seg, d.np, d.cap,
                                      list + d.list_off, count + d.base, stats + 2 * sys, s_rec, st, s_bound, s_rr);
}

// all-pairs lists of small systems: the workgroup builds the lists of
// particles y ENS_SMALL_P .. of system sys, with the whole system staged in
// LDS; one wave per particle, lane l testing j = base + l, a ballot giving each
// hit its rank (entries are particle indices of the system, ascending)
__device__ __forceinline__ void ens_build_small_group(uint4 const* __restrict__ posq, ens_desc const& d, unsigned sys,
                                                      unsigned y, uint16_t* __restrict__ list,
                                                      uint32_t* __restrict__ count, uint32_t* __restrict__ stats,
                                                      uint4* s_rec, uint32_t* overflow) {
    unsigned const i0 = y * ENS_SMALL_P;
    if (i0 >= d.n) return;   // the whole workgroup
    unsigned const n = d.n, cap = d.cap;
    for (unsigned t = threadIdx.x; t < n; t += blockDim.x) s_rec[t] = posq[d.base + t];
    __syncthreads();
    float const cut = d.np.rr_cut_skin_q[0];
    uint16_t* const L = list + d.list_off;
    unsigned const lane = threadIdx.x % WAVE, W = blockDim.x / WAVE;
    uint64_t const below = (uint64_t(1) << lane) - 1u;
    unsigned cmax = 0;
    unsigned const i1 = min(n, i0 + ENS_SMALL_P);
    for (unsigned i = i0 + threadIdx.x / WAVE; i < i1; i += W) {   // wave-uniform
        uint4 const ri = s_rec[i];
        uint16_t* const out = L + size_t(i & ~(WAVE - 1)) * cap + (i & (WAVE - 1)) * NB16_BLOCK;
        unsigned c = 0;
        for (unsigned base = 0; base < n; base += WAVE) {
            unsigned const j = base + lane;
            bool const hit = j < n && j != i && pair_inside<true>(ri, s_rec[j < n ? j : i], d.np, cut);
            uint64_t const m = __ballot(hit);
            if (hit) {
                unsigned const cc = c + unsigned(__popcll(m & below));
                if (cc < cap) out[size_t(cc / NB16_BLOCK) * (WAVE * NB16_BLOCK) + cc % NB16_BLOCK] = uint16_t(j);
            }
            c += unsigned(__popcll(m));
        }
        if (lane == 0) count[d.base + i] = c;
        cmax = max(cmax, c);
    }
    if (lane == 0 && cmax > 0) atomicMax(&stats[2 * sys], cmax);
    if (overflow && lane == 0 && cmax > cap) atomicOr(overflow, ENS_SKIP_OVERFLOW);
    if (y == 0 && threadIdx.x == 0) atomicMax(&stats[2 * sys + 1], n + 1);
}

// workgroup (x, y): system rsmall[x], group y
__global__ __launch_bounds__(256) void ens_build_small_kernel(uint4 const* __restrict__ posq,
                                                              ens_desc const* __restrict__ descs,
                                                              uint32_t const* __restrict__ rsmall,
                                                              uint16_t* __restrict__ list,
                                                              uint32_t* __restrict__ count,
                                                              uint32_t* __restrict__ stats) {
    extern __shared__ uint4 s_rec[];
    unsigned const sys = rsmall[blockIdx.x];
    ens_build_small_group(posq, descs[sys], sys, blockIdx.y, list, count, stats, s_rec, nullptr);
}

// --- ensembles: rebuilds decided on the device ---------------------------

// k with pre[k] <= g < pre[k + 1], for g < pre[m] (pre ascending, pre[0] = 0)
__device__ __forceinline__ unsigned work_find(uint32_t const* __restrict__ pre, unsigned m, uint32_t g) {
    unsigned lo = 0, hi = m;   // pre[lo] <= g < pre[hi]
    while (hi - lo > 1) {
        unsigned const mid = (lo + hi) / 2;
        if (pre[mid] <= g) lo = mid; else hi = mid;
    }
    return lo;
}

// exclusive scans of V values over the workgroup (blockDim.x a multiple of
// 64, at most 1024): e[k] the prefix of v[k], t[k] its total.  s_tmp: 17 V
// words of LDS.  Two barriers (and one more before s_tmp is reused).
template <int V>
__device__ __forceinline__ void block_exclusive_scan_n(uint32_t const (&v)[V], uint32_t (&e)[V], uint32_t (&t)[V],
                                                       uint32_t* s_tmp) {
    unsigned const lane = threadIdx.x % WAVE, wid = threadIdx.x / WAVE, nw = blockDim.x / WAVE;
    uint32_t x[V];
#pragma unroll
    for (int k = 0; k < V; ++k) {
        x[k] = v[k];
        for (unsigned o = 1; o < WAVE; o <<= 1) {
            uint32_t const y = __shfl_up(x[k], o, WAVE);
            if (lane >= o) x[k] += y;
        }
    }
    if (lane == WAVE - 1)
#pragma unroll
        for (int k = 0; k < V; ++k) s_tmp[wid * V + k] = x[k];
    __syncthreads();
    if (threadIdx.x < unsigned(V)) {   // value k: exclusive scan over the waves
        uint32_t acc = 0;
        for (unsigned w = 0; w < nw; ++w) {
            uint32_t const y = s_tmp[w * V + threadIdx.x];
            s_tmp[w * V + threadIdx.x] = acc;
            acc += y;
        }
        s_tmp[16 * V + threadIdx.x] = acc;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < V; ++k) {
        e[k] = s_tmp[wid * V + k] + x[k] - v[k];
        t[k] = s_tmp[16 * V + k];
    }
    __syncthreads();
}

__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* s_tmp, uint32_t& total) {
    uint32_t const vv[1] = {v};
    uint32_t e[1], t[1];
    block_exclusive_scan_n<1>(vv, e, t, s_tmp);
    total = t[0];
    return e[0];
}

// one workgroup of 1024 threads
__global__ __launch_bounds__(1024) void ens_decide_kernel(ens_desc const* __restrict__ descs, unsigned nsys,
                                                          uint32_t const* __restrict__ skip,
                                                          uint32_t* __restrict__ disp, uint32_t* __restrict__ steps,
                                                          uint32_t* __restrict__ rebuilds,
                                                          uint32_t* __restrict__ stats, ens_work w,
                                                          uint32_t* __restrict__ flags) {
    constexpr int V = 8;   // cell systems: flag, cells, blocks, segments, cell blocks; all-pairs: flag, blocks, groups
    __shared__ uint32_t s_tmp[17 * V];
    uint32_t carry[V] = {};
    for (unsigned base = 0; base < nsys; base += blockDim.x) {
        unsigned const s = base + threadIdx.x;
        uint32_t v[V] = {};
        if (s < nsys) {
            uint32_t* const slots = disp + size_t(s) * ENS_DISP_WORDS;
            uint32_t m = 0;
#pragma unroll
            for (unsigned k = 0; k < ENS_DISP_SLOTS; ++k) m = max(m, slots[k * ENS_DISP_STRIDE]);
#pragma unroll
            for (unsigned k = 0; k < ENS_DISP_SLOTS; ++k) slots[k * ENS_DISP_STRIDE] = 0u;
            ens_desc const& d = descs[s];
            uint32_t flag = 0;
            if (!skip[s]) {
                atomicAdd(&steps[s], 1u);   // (no result: nothing waits for it)
                if (m > d.keep_bits) {      // bits of non-negative floats order as the floats
                    flag = 1;
                    atomicAdd(&rebuilds[s], 1u);
                    stats[2 * s] = 0u;
                    stats[2 * s + 1] = 0u;
                    unsigned const nblk = (d.n + ENS_BLOCK - 1) / ENS_BLOCK;
                    if (d.small) {
                        v[5] = 1; v[6] = nblk; v[7] = (d.n + ENS_SMALL_P - 1) / ENS_SMALL_P;
                    } else {
                        v[0] = 1; v[1] = d.ncells; v[2] = nblk; v[3] = d.sp.nseg;
                        v[4] = (d.ncells + ENS_CELL_BLOCK - 1) / ENS_CELL_BLOCK;
                    }
                }
            }
            flags[s] = flag;
        }
        uint32_t e[V], t[V];
        block_exclusive_scan_n<V>(v, e, t, s_tmp);
        if (v[0]) {
            unsigned const r = carry[0] + e[0];
            w.lsys[r] = s;
            w.lcell[r] = carry[1] + e[1];
            w.lblk[r] = carry[2] + e[2];
            w.lseg[r] = carry[3] + e[3];
            w.lcb[r] = carry[4] + e[4];
        }
        if (v[5]) {
            unsigned const r = carry[5] + e[5];
            w.ssys[r] = s;
            w.sblk[r] = carry[6] + e[6];
            w.sgrp[r] = carry[7] + e[7];
        }
#pragma unroll
        for (int k = 0; k < V; ++k) carry[k] += t[k];
    }
    if (threadIdx.x == 0) {
        w.head[0] = carry[0];
        w.head[1] = carry[5];
        w.lcell[carry[0]] = carry[1];
        w.lblk[carry[0]] = carry[2];
        w.lseg[carry[0]] = carry[3];
        w.lcb[carry[0]] = carry[4];
        w.sblk[carry[5]] = carry[6];
        w.sgrp[carry[5]] = carry[7];
    }
}

__device__ __forceinline__ void cell_coords(unsigned c, cell_grid const& g, unsigned& x, unsigned& y, unsigned& z) {
    x = c % g.ncell[0];
    unsigned const r = c / g.ncell[0];
    y = r % g.ncell[1];
    z = r / g.ncell[1];
}

__device__ __forceinline__ unsigned cell_index(unsigned x, unsigned y, unsigned z, cell_grid const& g) {
    return (z * g.ncell[1] + y) * g.ncell[0] + x;
}

// a - b on a ring of nc >= 3 cells, as -1, 0, +1 (or 2 when farther)
__device__ __forceinline__ int ring_step(unsigned a, unsigned b, unsigned nc) {
    int d = int(a) - int(b);
    if (d > 1) d -= int(nc);
    if (d < -1) d += int(nc);
    return (d >= -1 && d <= 1) ? d : 2;
}

__device__ __forceinline__ unsigned ring_add(unsigned a, int e, unsigned nc) {
    int const r = int(a) + e;
    return unsigned(r < 0 ? r + int(nc) : (r >= int(nc) ? r - int(nc) : r));
}

// the offsets -1, 0, +1 along one ring in the ascending order of a + e
__device__ __forceinline__ void ascending_steps(unsigned a, unsigned nc, int o[3]) {
    if (a == 0) { o[0] = 0; o[1] = 1; o[2] = -1; }
    else if (a == nc - 1) { o[0] = 1; o[1] = -1; o[2] = 0; }
    else { o[0] = -1; o[1] = 0; o[2] = 1; }
}

__device__ __forceinline__ unsigned step_code(int ex, int ey, int ez) { return unsigned((ez + 1) * 9 + (ey + 1) * 3 + (ex + 1)); }

// waves over runs of consecutive old cells of the flagged cell systems, a
// cell per half-wave (32 lanes: most cells hold fewer particles; a cell with
// more takes more chunks); workgroups of 256 threads: s_cnt holds one row per
// half-wave
__global__ __launch_bounds__(256) void ens_rank_kernel(uint4 const* __restrict__ posq,
                                                       uint32_t const* __restrict__ offsets,
                                                       ens_desc const* __restrict__ descs, uint32_t* skip,
                                                       ens_work w, uint32_t* __restrict__ keys,
                                                       uint32_t* __restrict__ vals, uint32_t* __restrict__ cnt27) {
    // per half-wave: particles so far of each of the 27 codes in its cell
    __shared__ uint32_t s_cnt[256 / WAVE][2][32];
    unsigned const nl = w.head[0];
    uint32_t const total = w.lcell[nl];
    unsigned const lane = threadIdx.x % WAVE, half = lane / 32, hl = lane % 32;
    uint32_t* const wc = s_cnt[threadIdx.x / WAVE][half];
    unsigned const waves = gridDim.x * (blockDim.x / WAVE);
    unsigned const wave = (blockIdx.x * blockDim.x + threadIdx.x) / WAVE;
    uint64_t const below = (uint64_t(1) << lane) - 1u;
    uint64_t const mine = half ? ~uint64_t(0) << 32 : ~uint64_t(0) >> 32;
    // an even number of cells per wave: a pair at a time
    uint32_t const run = ((total + waves - 1) / waves + 1u) & ~1u;
    uint32_t const g0 = min(total, wave * run), g1 = min(total, g0 + run);
    if (g0 >= g1) return;   // the whole wave
    unsigned k = work_find(w.lcell, nl, g0);   // (per lane from here: the halves may be in two systems)
    for (uint32_t gp = g0; gp < g1; gp += 2) {   // wave-uniform
        uint32_t const g = gp + half;
        bool const has = g < g1;
        if (has)
            while (g >= w.lcell[k + 1]) ++k;
        unsigned const s = w.lsys[k];
        ens_desc const& d = descs[s];
        cell_grid const& cg = d.sp.grid;
        unsigned const cp = has ? g - w.lcell[k] : 0u;
        unsigned xo, yo, zo;
        cell_coords(cp, cg, xo, yo, zo);
        uint32_t const* const off = offsets + d.cell_base;
        unsigned const b = has ? off[cp] : 0u, len = has ? off[cp + 1] - b : 0u;
        unsigned const lmax = max(len, unsigned(__shfl_xor(int(len), 32, WAVE)));   // the longer cell of the pair
        wc[hl] = 0u;
        __builtin_amdgcn_wave_barrier();
        bool far = false;
        for (unsigned i0 = 0; i0 < lmax; i0 += 32) {
            unsigned const i = b + i0 + hl;
            bool const act = i0 + hl < len;
            unsigned code = 31u, c = 0;
            if (act) {
                uint4 const q = posq[d.base + i];
                unsigned const x = cell_component(q.x, cg.ncell[0]), y = cell_component(q.y, cg.ncell[1]),
                               z = cell_component(q.z, cg.ncell[2]);
                c = cell_index(x, y, z, cg);
                int const ex = ring_step(xo, x, cg.ncell[0]), ey = ring_step(yo, y, cg.ncell[1]),
                          ez = ring_step(zo, z, cg.ncell[2]);
                if (ex == 2 || ey == 2 || ez == 2)
                    far = true;
                else
                    code = step_code(ex, ey, ez);
            }
            uint64_t peers = mine;   // lanes of this half with the same code
#pragma unroll
            for (unsigned bit = 0; bit < 5; ++bit) {
                bool const on = (code >> bit) & 1u;
                uint64_t const m = __ballot(on);
                peers &= on ? m : ~m;
            }
            unsigned const rank = unsigned(__popcll(peers & below));
            if (act && code < 27u) {
                // the wave's LDS instructions run in order: every lane reads
                // its code's count before the lowest lane of each code adds
                // the chunk's particles of that code
                unsigned const base = wc[code];
                keys[d.base + i] = c | (code << ENS_CELL_KEY_BITS);
                vals[d.base + i] = base + rank;
                if (rank == 0) wc[code] = base + unsigned(__popcll(peers));
            }
            __builtin_amdgcn_wave_barrier();
        }
        if ((__ballot(far) & mine) != 0 && hl == 0) atomicOr(&skip[s], ENS_SKIP_FAR);
        if (has && hl < 27) cnt27[(size_t(d.cell_base) + cp) * 27 + hl] = wc[hl];
        __builtin_amdgcn_wave_barrier();
    }
}

// one thread per new cell of the flagged cell systems (workgroups of
// ENS_CELL_BLOCK cells of one system): the particles that reach cell c from
// each of its 27 source cells, in ascending source index, as running sums
// (base27), their total scanned within the block (offsets_t) and the block
// sums (bsum, one per block of the work table)
__global__ __launch_bounds__(ENS_CELL_BLOCK) void ens_cell_count_kernel(
    ens_desc const* __restrict__ descs, uint32_t const* skip, ens_work w, uint32_t const* __restrict__ cnt27,
    uint32_t* __restrict__ base27, uint32_t* __restrict__ offsets_t, uint32_t* __restrict__ bsum) {
    __shared__ uint32_t s_tmp[17];
    unsigned const nl = w.head[0];
    uint32_t const total = w.lcb[nl];
    for (uint32_t g = blockIdx.x; g < total; g += gridDim.x) {   // workgroup-uniform
        unsigned const k = work_find(w.lcb, nl, g);
        unsigned const s = w.lsys[k];
        if (skip[s]) continue;   // (no writer of skip runs during this kernel: uniform)
        ens_desc const& d = descs[s];
        cell_grid const& cg = d.sp.grid;
        unsigned const C = d.ncells;
        unsigned const c = (g - w.lcb[k]) * ENS_CELL_BLOCK + threadIdx.x;
        uint32_t v = 0;
        if (c < C) {
            uint32_t const* const cnt = cnt27 + size_t(d.cell_base) * 27;
            uint32_t* const out = base27 + (size_t(d.cell_base) + c) * 27;
            unsigned x, y, z;
            cell_coords(c, cg, x, y, z);
            int ox[3], oy[3], oz[3];
            ascending_steps(x, cg.ncell[0], ox);
            ascending_steps(y, cg.ncell[1], oy);
            ascending_steps(z, cg.ncell[2], oz);
            uint32_t n27[27];
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
                for (int b = 0; b < 3; ++b)
#pragma unroll
                    for (int e = 0; e < 3; ++e) {
                        unsigned const src = cell_index(ring_add(x, ox[e], cg.ncell[0]), ring_add(y, oy[b], cg.ncell[1]),
                                                        ring_add(z, oz[a], cg.ncell[2]), cg);
                        n27[a * 9 + b * 3 + e] = cnt[size_t(src) * 27 + step_code(ox[e], oy[b], oz[a])];
                    }
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
                for (int b = 0; b < 3; ++b)
#pragma unroll
                    for (int e = 0; e < 3; ++e) {
                        out[step_code(ox[e], oy[b], oz[a])] = v;
                        v += n27[a * 9 + b * 3 + e];
                    }
        }
        uint32_t tot;
        uint32_t const ex = block_exclusive_scan(v, s_tmp, tot);
        if (c < C) offsets_t[d.cell_base + c] = ex;
        if (threadIdx.x == 0) bsum[g] = tot;
    }
}

// the block sums of the earlier blocks of the same system added; the last
// block also writes offsets_t[n_cells] = n
__global__ __launch_bounds__(ENS_CELL_BLOCK) void ens_cell_carry_kernel(ens_desc const* __restrict__ descs,
                                                                        uint32_t const* skip, ens_work w,
                                                                        uint32_t const* __restrict__ bsum,
                                                                        uint32_t* __restrict__ offsets_t) {
    __shared__ uint32_t s_tmp[17];
    unsigned const nl = w.head[0];
    uint32_t const total = w.lcb[nl];
    for (uint32_t g = blockIdx.x; g < total; g += gridDim.x) {   // workgroup-uniform
        unsigned const k = work_find(w.lcb, nl, g);
        unsigned const s = w.lsys[k];
        if (skip[s]) continue;
        ens_desc const& d = descs[s];
        unsigned const cb = g - w.lcb[k], C = d.ncells;
        uint32_t part = 0;
        for (unsigned j = threadIdx.x; j < cb; j += blockDim.x) part += bsum[w.lcb[k] + j];
        uint32_t carry;
        (void)block_exclusive_scan(part, s_tmp, carry);   // carry: the sum over the workgroup
        unsigned const c = cb * ENS_CELL_BLOCK + threadIdx.x;
        if (c < C) offsets_t[d.cell_base + c] += carry;
        if (threadIdx.x == 0 && (cb + 1) * ENS_CELL_BLOCK >= C) offsets_t[d.cell_base + C] = carry + bsum[g];
    }
}

// each particle of the flagged cell systems to its slot in the new order
__global__ __launch_bounds__(256) void ens_scatter_kernel(
    float4 const* __restrict__ pos, float4 const* __restrict__ vel, int4 const* __restrict__ image,
    uint32_t const* __restrict__ id, uint4 const* __restrict__ posq, float4* __restrict__ pos_t,
    float4* __restrict__ vel_t, int4* __restrict__ image_t, uint32_t* __restrict__ id_t, uint4* __restrict__ posq_t,
    ens_desc const* __restrict__ descs, uint32_t const* skip, ens_work w, uint32_t const* __restrict__ keys,
    uint32_t const* __restrict__ vals, uint32_t const* __restrict__ base27, uint32_t const* __restrict__ offsets_t) {
    unsigned const nl = w.head[0];
    uint32_t const total = w.lblk[nl];
    for (uint32_t g = blockIdx.x; g < total; g += gridDim.x) {   // workgroup-uniform
        unsigned const k = work_find(w.lblk, nl, g);
        unsigned const s = w.lsys[k];
        if (skip[s]) continue;
        ens_desc const& d = descs[s];
        unsigned const li = (g - w.lblk[k]) * ENS_BLOCK + threadIdx.x;
        if (li >= d.n) continue;
        uint32_t const key = keys[d.base + li];
        unsigned const c = key & ((1u << ENS_CELL_KEY_BITS) - 1u), code = key >> ENS_CELL_KEY_BITS;
        size_t const cc = size_t(d.cell_base) + c;
        size_t const dst = d.base + offsets_t[cc] + base27[cc * 27 + code] + vals[d.base + li];
        size_t const src = d.base + li;
        pos_t[2 * dst] = pos[2 * src];
        pos_t[2 * dst + 1] = pos[2 * src + 1];
        vel_t[2 * dst] = vel[2 * src];
        vel_t[2 * dst + 1] = vel[2 * src + 1];
        image_t[dst] = image[src];
        id_t[dst] = id[src];
        posq_t[dst] = posq[src];
    }
}

// the same slots as a permutation: perm[slot] = li (a gather list, for a
// particle instance that rearranges itself)
__global__ __launch_bounds__(256) void ens_order_kernel(ens_desc const* __restrict__ descs, uint32_t const* skip,
                                                        ens_work w, uint32_t const* __restrict__ keys,
                                                        uint32_t const* __restrict__ vals,
                                                        uint32_t const* __restrict__ base27,
                                                        uint32_t const* __restrict__ offsets_t,
                                                        uint32_t* __restrict__ perm) {
    unsigned const nl = w.head[0];
    uint32_t const total = w.lblk[nl];
    for (uint32_t g = blockIdx.x; g < total; g += gridDim.x) {   // workgroup-uniform
        unsigned const k = work_find(w.lblk, nl, g);
        unsigned const s = w.lsys[k];
        if (skip[s]) continue;
        ens_desc const& d = descs[s];
        unsigned const li = (g - w.lblk[k]) * ENS_BLOCK + threadIdx.x;
        if (li >= d.n) continue;
        uint32_t const key = keys[d.base + li];
        unsigned const c = key & ((1u << ENS_CELL_KEY_BITS) - 1u), code = key >> ENS_CELL_KEY_BITS;
        size_t const cc = size_t(d.cell_base) + c;
        perm[d.base + offsets_t[cc] + base27[cc * 27 + code] + vals[d.base + li]] = li;
    }
}

// twins back into place, r0 := new positions (hi words), new cell offsets
__global__ __launch_bounds__(256) void ens_copyback_flagged_kernel(
    float4* __restrict__ pos, float4* __restrict__ vel, int4* __restrict__ image, uint32_t* __restrict__ id,
    uint4* __restrict__ posq, float4* __restrict__ r0, float4 const* __restrict__ pos_t,
    float4 const* __restrict__ vel_t, int4 const* __restrict__ image_t, uint32_t const* __restrict__ id_t,
    uint4 const* __restrict__ posq_t, uint32_t* __restrict__ offsets, uint32_t const* __restrict__ offsets_t,
    ens_desc const* __restrict__ descs, uint32_t const* skip, ens_work w) {
    unsigned const nl = w.head[0];
    uint32_t const total = w.lblk[nl];
    for (uint32_t g = blockIdx.x; g < total; g += gridDim.x) {   // workgroup-uniform
        unsigned const k = work_find(w.lblk, nl, g);
        unsigned const s = w.lsys[k];
        if (skip[s]) continue;
        ens_desc const& d = descs[s];
        unsigned const lb = g - w.lblk[k], nblk = (d.n + ENS_BLOCK - 1) / ENS_BLOCK;
        // this block's share of the n_cells + 1 offsets
        size_t const C1 = size_t(d.ncells) + 1;
        size_t const c_lo = C1 * lb / nblk, c_hi = C1 * (lb + 1) / nblk;
        for (size_t c = c_lo + threadIdx.x; c < c_hi; c += blockDim.x)
            offsets[d.cell_base + c] = offsets_t[d.cell_base + c];
        unsigned const li = lb * ENS_BLOCK + threadIdx.x;
        if (li >= d.n) continue;
        size_t const i = d.base + li;
        float4 const h = pos_t[2 * i];
        pos[2 * i] = h;
        pos[2 * i + 1] = pos_t[2 * i + 1];
        vel[2 * i] = vel_t[2 * i];
        vel[2 * i + 1] = vel_t[2 * i + 1];
        image[i] = image_t[i];
        id[i] = id_t[i];
        posq[i] = posq_t[i];
        r0[i] = h;
    }
}

// the staged list build over the flagged segments
template <bool MULTI>
__global__ __launch_bounds__(SEG_THREADS_MAX) void ens_build_flagged_kernel(
    uint4 const* __restrict__ posq, uint32_t const* __restrict__ offsets, ens_desc const* __restrict__ descs,
    uint32_t* skip, ens_work w, uint16_t* __restrict__ list, uint32_t* __restrict__ count,
    uint32_t* __restrict__ stats)
{
    extern __shared__ uint4 s_rec[];
    __shared__ seg_stage st;
    __shared__ unsigned s_bound[SEG_NR * SEG_BOUND];
    __shared__ uint32_t s_rr[MAX_TYPES * MAX_TYPES];
    __shared__ uint32_t s_skip;
    unsigned const nl = w.head[0];
    uint32_t const total = w.lseg[nl];
    for (uint32_t g = blockIdx.x; g < total; g += gridDim.x) {   // workgroup-uniform
        unsigned const k = work_find(w.lseg, nl, g);
        unsigned const s = w.lsys[k];
        // (other workgroups of this launch may set the system's overflow bit:
        // one read, shared, keeps the branch uniform)
        if (threadIdx.x == 0) s_skip = skip[s];
        __syncthreads();
        bool const skipped = s_skip != 0;
        if (!skipped) {
            ens_desc const& d = descs[s];
            staged_build_segment<MULTI, true>(posq + d.base, offsets + d.cell_base, d.sp, g - w.lseg[k], d.np, d.cap,
                                              list + d.list_off, count + d.base, stats + 2 * s, s_rec, st, s_bound,
                                              s_rr, SEG_BOUND, &skip[s]);
        }
        __syncthreads();   // the next segment reuses the LDS
    }
}

// r0 := positions of the flagged all-pairs systems
__global__ __launch_bounds__(256) void ens_r0_flagged_kernel(float4 const* __restrict__ pos, float4* __restrict__ r0,
                                                             ens_desc const* __restrict__ descs, ens_work w) {
    unsigned const ns = w.head[1];
    uint32_t const total = w.sblk[ns];
    for (uint32_t g = blockIdx.x; g < total; g += gridDim.x) {
        unsigned const k = work_find(w.sblk, ns, g);
        ens_desc const& d = descs[w.ssys[k]];
        unsigned const li = (g - w.sblk[k]) * ENS_BLOCK + threadIdx.x;
        if (li < d.n) r0[d.base + li] = pos[2 * size_t(d.base + li)];
    }
}

__global__ __launch_bounds__(256) void ens_build_small_flagged_kernel(uint4 const* __restrict__ posq,
                                                                      ens_desc const* __restrict__ descs,
                                                                      uint32_t* skip, ens_work w,
                                                                      uint16_t* __restrict__ list,
                                                                      uint32_t* __restrict__ count,
                                                                      uint32_t* __restrict__ stats) {
    extern __shared__ uint4 s_rec[];
    unsigned const ns = w.head[1];
    uint32_t const total = w.sgrp[ns];
    for (uint32_t g = blockIdx.x; g < total; g += gridDim.x) {   // workgroup-uniform
        unsigned const k = work_find(w.sgrp, ns, g);
        unsigned const s = w.ssys[k];
        ens_build_small_group(posq, descs[s], s, g - w.sgrp[k], list, count, stats, s_rec, &skip[s]);
        __syncthreads();   // the next group restages the LDS
    }
}

} // namespace

void launch_compute_cell(uint4 const* posq, unsigned n, cell_grid g, uint32_t* keys, uint32_t* vals, hipStream_t s,
                         unsigned sub_bits) {
    if (!n) return;
    compute_cell_kernel<<<div_up(n, 256), 256, 0, s>>>(posq, n, g, keys, vals, sub_bits);
    HIP_CHECK(hipGetLastError());
}

void launch_cell_offsets(uint32_t const* keys, unsigned n, unsigned nbins, uint32_t* offset, hipStream_t s,
                         unsigned shift) {
    cell_offsets_kernel<<<div_up(size_t(n) + 1, 256), 256, 0, s>>>(keys, n, nbins, offset, shift);
    HIP_CHECK(hipGetLastError());
}

void launch_gather_cell_pos(uint4 const* posq, uint32_t const* vals, unsigned n, uint4* cell_pos, hipStream_t s) {
    if (!n) return;
    gather_cell_pos_kernel<<<div_up(n, 256), 256, 0, s>>>(posq, vals, n, cell_pos);
    HIP_CHECK(hipGetLastError());
}

void launch_build_neighbours(uint4 const* posq, unsigned n, uint32_t const* offset, uint4 const* cell_pos,
                             neighbour_params const& p, unsigned cap, uint32_t* list, uint32_t* count,
                             uint32_t* max_count, hipStream_t s, bool self) {
    if (!n) return;
    if (p.iso)
        build_neighbours_kernel<true><<<div_up(n, 256), 256, 0, s>>>(posq, n, offset, cell_pos, p, cap, list, count,
                                                                     max_count, self);
    else
        build_neighbours_kernel<false><<<div_up(n, 256), 256, 0, s>>>(posq, n, offset, cell_pos, p, cap, list, count,
                                                                      max_count, self);
    HIP_CHECK(hipGetLastError());
}

} // namespace halmd

namespace halmd {

void launch_build_neighbours_staged(uint4 const* posq, uint32_t const* offset, seg_params const& sp,
                                    neighbour_params const& p, unsigned cap, uint16_t* list, uint32_t* count,
                                    uint32_t* stats, hipStream_t s, uint32_t const* sub_offset) {
    if (!sp.nseg) return;
    unsigned const grid = SEG_XCD * sp.chunk;
    if (sub_offset) {
        if (!p.iso || sp.grid.ncell[0] < 4)
            throw error(HALMD_ERR_INVALID_ARGUMENT, "windowed list build: cubic box with 4 or more cells along x");
        size_t const wl = windowed_build_lds(sp);
        if (p.ntype > 1)
            build_windowed_kernel<true><<<grid, sp.threads, wl, s>>>(posq, offset, sub_offset, sp, p, cap, list,
                                                                    count, stats);
        else
            build_windowed_kernel<false><<<grid, sp.threads, wl, s>>>(posq, offset, sub_offset, sp, p, cap, list,
                                                                     count, stats);
        HIP_CHECK(hipGetLastError());
        return;
    }
    size_t const lds = size_t(sp.lds_cap) * (p.iso ? sizeof(uint2) : sizeof(uint4)) +
                       size_t(3) * (sp.R + 2) * (sp.S + 3) * sizeof(unsigned);
    if (p.ntype > 1) {
        if (p.iso)
            build_staged_kernel<true, true><<<grid, sp.threads, lds, s>>>(posq, offset, sp, p, cap, list, count, stats);
        else
            build_staged_kernel<true, false><<<grid, sp.threads, lds, s>>>(posq, offset, sp, p, cap, list, count, stats);
    } else {
        if (p.iso)
            build_staged_kernel<false, true><<<grid, sp.threads, lds, s>>>(posq, offset, sp, p, cap, list, count, stats);
        else
            build_staged_kernel<false, false><<<grid, sp.threads, lds, s>>>(posq, offset, sp, p, cap, list, count,
                                                                          stats);
    }
    HIP_CHECK(hipGetLastError());
}

} // namespace halmd

namespace halmd {

void launch_build_neighbours_all_pairs(uint4 const* posq, unsigned n, neighbour_params const& p, unsigned cap,
                                       uint32_t* list, uint32_t* count, uint32_t* max_count, hipStream_t s,
                                       rebuild_condition const& cond) {
    if (!n) return;
    if (cond.slots && !(cond.pos && cond.r0 && cond.rebuilt))
        throw error(HALMD_ERR_INVALID_ARGUMENT, "all-pairs build: incomplete rebuild condition");
    unsigned const grid = div_up(n, AP_PER_GROUP);
    if (p.iso)
        build_all_pairs_kernel<true><<<grid, 256, 0, s>>>(posq, n, p, cap, list, count, max_count, cond);
    else
        build_all_pairs_kernel<false><<<grid, 256, 0, s>>>(posq, n, p, cap, list, count, max_count, cond);
    HIP_CHECK(hipGetLastError());
}

} // namespace halmd

namespace halmd {

void launch_ens_cells(uint4 const* posq, ens_desc const* descs, uint32_t const* rsys, uint32_t const* rpos,
                      uint32_t const* rblk, unsigned nrblk, unsigned cell_bits, uint32_t* keys, uint32_t* vals,
                      hipStream_t s) {
    if (!nrblk) return;
    ens_cells_kernel<<<nrblk, 256, 0, s>>>(posq, descs, rsys, rpos, rblk, cell_bits, keys, vals);
    HIP_CHECK(hipGetLastError());
}

void launch_ens_offsets(ens_desc const* descs, uint32_t const* rsys, uint32_t const* rpos, uint32_t const* rblk,
                        unsigned nrblk, unsigned cell_bits, uint32_t const* sorted_keys, uint32_t* offsets,
                        hipStream_t s) {
    if (!nrblk) return;
    ens_offsets_kernel<<<nrblk, 256, 0, s>>>(descs, rsys, rpos, rblk, cell_bits, sorted_keys, offsets);
    HIP_CHECK(hipGetLastError());
}

void launch_ens_reorder(float4* pos, float4* vel, int4* image, uint32_t* id, uint4* posq, float4* r0,
                        float4* pos_t, float4* vel_t, int4* image_t, uint32_t* id_t, uint4* posq_t,
                        ens_desc const* descs, uint32_t const* rsys, uint32_t const* rpos, uint32_t const* rblk,
                        unsigned nrblk, uint32_t const* sorted_vals, hipStream_t s) {
    if (!nrblk) return;
    ens_gather_kernel<<<nrblk, 256, 0, s>>>(pos, vel, image, id, posq, pos_t, vel_t, image_t, id_t, posq_t, descs,
                                            rsys, rpos, rblk, sorted_vals);
    HIP_CHECK(hipGetLastError());
    ens_copyback_kernel<<<nrblk, 256, 0, s>>>(pos, vel, image, id, posq, r0, pos_t, vel_t, image_t, id_t, posq_t,
                                              descs, rsys, rblk);
    HIP_CHECK(hipGetLastError());
}

void launch_ens_r0(float4 const* pos, float4* r0, ens_desc const* descs, uint32_t const* rblk_sys, unsigned nrblk,
                   hipStream_t s) {
    if (!nrblk) return;
    ens_r0_kernel<<<nrblk, 256, 0, s>>>(pos, r0, descs, rblk_sys);
    HIP_CHECK(hipGetLastError());
}

void launch_ens_build(uint4 const* posq, uint32_t const* offsets, ens_desc const* descs, uint32_t const* rseg,
                      unsigned nrseg, unsigned threads, unsigned lds_records, bool multi, uint16_t* list,
                      uint32_t* count, uint32_t* stats, hipStream_t s) {
    if (!nrseg) return;
    // cubic boxes: packed records, and 4 records of slack for the scan's last group
    size_t const lds = (size_t(lds_records) + 4) * sizeof(uint2);
    if (multi)
        ens_build_kernel<true><<<nrseg, threads, lds, s>>>(posq, offsets, descs, rseg, list, count, stats);
    else
        ens_build_kernel<false><<<nrseg, threads, lds, s>>>(posq, offsets, descs, rseg, list, count, stats);
    HIP_CHECK(hipGetLastError());
}

ens_work ens_work_at(uint32_t* words, size_t nsys) {
    ens_work w;
    w.head = words;
    uint32_t* p = words + 16;
    size_t const m = nsys + 1;
    w.lsys = p; p += m;
    w.lcell = p; p += m;
    w.lblk = p; p += m;
    w.lseg = p; p += m;
    w.lcb = p; p += m;
    w.ssys = p; p += m;
    w.sblk = p; p += m;
    w.sgrp = p;
    return w;
}

void launch_ens_decide(ens_desc const* descs, unsigned nsys, uint32_t const* skip, uint32_t* disp, uint32_t* steps,
                       uint32_t* rebuilds, uint32_t* stats, ens_work w, uint32_t* flags, hipStream_t s) {
    ens_decide_kernel<<<1, 1024, 0, s>>>(descs, nsys, skip, disp, steps, rebuilds, stats, w, flags);
    HIP_CHECK(hipGetLastError());
}

// fixed grids (the work is on the device): 4 workgroups per CU
constexpr unsigned ENS_WORK_GRID = 1024;

void launch_ens_rebuild_device(ens_rebuild_args const& a, hipStream_t s) {
    ens_rank_kernel<<<ENS_WORK_GRID, 256, 0, s>>>(a.posq, a.offsets, a.descs, a.skip, a.w, a.keys, a.vals, a.cnt27);
    HIP_CHECK(hipGetLastError());
    ens_cell_count_kernel<<<ENS_WORK_GRID, ENS_CELL_BLOCK, 0, s>>>(a.descs, a.skip, a.w, a.cnt27, a.base27,
                                                                  a.offsets_t, a.bsum);
    HIP_CHECK(hipGetLastError());
    ens_cell_carry_kernel<<<ENS_WORK_GRID, ENS_CELL_BLOCK, 0, s>>>(a.descs, a.skip, a.w, a.bsum, a.offsets_t);
    HIP_CHECK(hipGetLastError());
    ens_scatter_kernel<<<ENS_WORK_GRID, 256, 0, s>>>(a.pos, a.vel, a.image, a.id, a.posq, a.pos_t, a.vel_t,
                                                    a.image_t, a.id_t, a.posq_t, a.descs, a.skip, a.w, a.keys,
                                                    a.vals, a.base27, a.offsets_t);
    HIP_CHECK(hipGetLastError());
    ens_copyback_flagged_kernel<<<ENS_WORK_GRID, 256, 0, s>>>(a.pos, a.vel, a.image, a.id, a.posq, a.r0, a.pos_t,
                                                             a.vel_t, a.image_t, a.id_t, a.posq_t, a.offsets,
                                                             a.offsets_t, a.descs, a.skip, a.w);
    HIP_CHECK(hipGetLastError());
    if (a.lds_records) {
        // cubic boxes: packed records, and 4 records of slack for the scan's last group
        size_t const lds = (size_t(a.lds_records) + 4) * sizeof(uint2);
        if (a.multi)
            ens_build_flagged_kernel<true><<<ENS_WORK_GRID, a.threads, lds, s>>>(a.posq, a.offsets, a.descs, a.skip,
                                                                                a.w, a.list, a.count, a.stats);
        else
            ens_build_flagged_kernel<false><<<ENS_WORK_GRID, a.threads, lds, s>>>(a.posq, a.offsets, a.descs, a.skip,
                                                                                 a.w, a.list, a.count, a.stats);
        HIP_CHECK(hipGetLastError());
    }
    if (a.small_max_n) {
        ens_r0_flagged_kernel<<<ENS_WORK_GRID, 256, 0, s>>>(a.pos, a.r0, a.descs, a.w);
        HIP_CHECK(hipGetLastError());
        ens_build_small_flagged_kernel<<<ENS_WORK_GRID, 256, size_t(a.small_max_n) * sizeof(uint4), s>>>(
            a.posq, a.descs, a.skip, a.w, a.list, a.count, a.stats);
        HIP_CHECK(hipGetLastError());
    }
}

void merge_tables_upload(cell_grid const& g, unsigned n, merge_scratch& m, hipStream_t s) {
    m.n = n;
    m.ncells = g.ncell[0] * g.ncell[1] * g.ncell[2];
    ens_desc d;
    std::memset(&d, 0, sizeof(d));
    d.n = n;
    d.ncells = g.ncell[0] * g.ncell[1] * g.ncell[2];
    d.sp.grid = g;
    std::vector<uint32_t> words(ens_work_words(1), 0u);
    ens_work const w = ens_work_at(words.data(), 1);
    w.head[0] = 1;                    // one cell system, flagged
    w.lsys[0] = 0;
    w.lcell[1] = d.ncells;
    w.lblk[1] = div_up(n, ENS_BLOCK);
    w.lcb[1] = div_up(d.ncells, ENS_CELL_BLOCK);
    HIP_CHECK(hipMemcpyAsync(m.desc, &d, sizeof(d), hipMemcpyHostToDevice, s));
    HIP_CHECK(hipMemcpyAsync(m.work, words.data(), words.size() * sizeof(uint32_t), hipMemcpyHostToDevice, s));
    HIP_CHECK(hipStreamSynchronize(s));   // the host copies are released
}

void launch_merge_cell_order(uint4 const* posq, uint32_t const* old_offsets, uint32_t* new_offsets, uint32_t* perm,
                             merge_scratch const& m, hipStream_t s) {
    ens_work const w = ens_work_at(m.work, 1);
    // grids sized to this one system (the kernels stride over any grid): a
    // wave per pair of old cells, a workgroup per block of cells or particles
    unsigned const rank_grid = std::max(1u, std::min(div_up(m.ncells, 2 * (256 / WAVE)), 65535u));
    unsigned const cell_grid_n = std::max(1u, std::min(div_up(m.ncells, ENS_CELL_BLOCK), 65535u));
    unsigned const blk_grid = std::max(1u, std::min(div_up(m.n, ENS_BLOCK), 65535u));
    launch_fill_u32(m.flag, 0u, 1, s);
    ens_rank_kernel<<<rank_grid, 256, 0, s>>>(posq, old_offsets, m.desc, m.flag, w, m.keys, m.rank, m.cnt27);
    HIP_CHECK(hipGetLastError());
    ens_cell_count_kernel<<<cell_grid_n, ENS_CELL_BLOCK, 0, s>>>(m.desc, m.flag, w, m.cnt27, m.base27, new_offsets,
                                                                m.bsum);
    HIP_CHECK(hipGetLastError());
    ens_cell_carry_kernel<<<cell_grid_n, ENS_CELL_BLOCK, 0, s>>>(m.desc, m.flag, w, m.bsum, new_offsets);
    HIP_CHECK(hipGetLastError());
    ens_order_kernel<<<blk_grid, 256, 0, s>>>(m.desc, m.flag, w, m.keys, m.rank, m.base27, new_offsets, perm);
    HIP_CHECK(hipGetLastError());
}

void launch_ens_build_small(uint4 const* posq, ens_desc const* descs, uint32_t const* rsmall, unsigned nsmall,
                            unsigned max_n, uint16_t* list, uint32_t* count, uint32_t* stats, hipStream_t s) {
    if (!nsmall) return;
    dim3 const grid(nsmall, div_up(max_n, ENS_SMALL_P));
    ens_build_small_kernel<<<grid, 256, size_t(max_n) * sizeof(uint4), s>>>(posq, descs, rsmall, list, count,
                                                                            stats);
    HIP_CHECK(hipGetLastError());
}

} // namespace halmd
