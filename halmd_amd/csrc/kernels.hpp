// Host-side launchers of the gfx950 kernels.  Every launcher is stream-ordered
// and synchronises nothing; device pointers and POD parameters only.
#pragma once

#include "common.hpp"

namespace halmd {

// maximum number of particle species (pair tables are MAX_TYPES^2)
constexpr unsigned MAX_TYPES = 4;

// ---------------------------------------------------------------------------
// scan / radix sort (sort.hip)
// ---------------------------------------------------------------------------
struct scan_workspace {
    dbuf<uint32_t> partial;       // per-level block sums
    std::vector<size_t> level_offset;
    void reserve(size_t n);
};
// exclusive prefix sum of m uint32 values, in place
void scan_exclusive(uint32_t* data, size_t m, scan_workspace& ws, hipStream_t s);

struct radix_sort_workspace {
    dbuf<uint32_t> keys_alt, vals_alt, counts;
    scan_workspace scan;
    size_t n_max = 0;
    void reserve(size_t n);
};
// stable LSD radix sort of (key, value) pairs by the low key_bits bits of key,
// in place; 8-bit digits, ceil(key_bits / 8) passes
void radix_sort(uint32_t* keys, uint32_t* vals, size_t n, unsigned key_bits, radix_sort_workspace& ws, hipStream_t s);
// the same sort on a file-scope device scratch array (no allocation): at most
// NOALLOC_SORT_MAX elements; calls that overlap on one device race
constexpr size_t NOALLOC_SORT_MAX = size_t(1) << 21;
void radix_sort_noalloc(uint32_t* keys, uint32_t* vals, size_t n, unsigned key_bits, hipStream_t s);

void launch_iota(uint32_t* out, size_t n, hipStream_t s);
// out[i] := i for i < n and dst[k] := src[k] for k < m, in one launch
void launch_iota_copy_u32(uint32_t* out, size_t n, uint32_t const* src, uint32_t* dst, size_t m, hipStream_t s);
void launch_fill_u32(uint32_t* out, uint32_t value, size_t n, hipStream_t s);
// copy of bytes / 16 16-byte words (a multiple of 16 bytes)
void launch_copy(void const* src, void* dst, size_t bytes, hipStream_t s);
// n words (a kernel rather than a DMA copy: shorter latency for small arrays)
void launch_copy_u32(uint32_t const* src, uint32_t* dst, size_t n, hipStream_t s);
// dst[k] := (seq << 32) | *src.word[k] for k < n (MAILBOX_WORDS at most; a
// null source reads as 0), stored at system scope into host-mapped memory
constexpr unsigned MAILBOX_WORDS = 4;
struct mailbox_sources {
    uint32_t const* word[MAILBOX_WORDS];
};
void launch_post_words(mailbox_sources const& src, unsigned n, unsigned long long* dst, uint32_t seq, hipStream_t s);
// VALU rate probe (blocks workgroups of 256 threads); returns the number of
// FP32 FMA wave instructions it issues
uint64_t launch_valu_probe(float* out, unsigned blocks, unsigned iterations, hipStream_t s);

// ---------------------------------------------------------------------------
// fixed-point position snapshot (integrate.hip)
// ---------------------------------------------------------------------------
// posq[i] = {ux, uy, uz, species} with u = floor(frac(x / L + 1/2) * 2^32),
// evaluated in FP64 from the double-word position x = hi + lo.  Differences
// int32(ui - uj) are minimum-image separations in units of L / 2^32.
struct quant_params {
    double invL[3];     // 1 / L in FP64
};
void launch_quantize(float4 const* pos, unsigned n, quant_params const& q, uint4* posq, hipStream_t s);

// ---------------------------------------------------------------------------
// binning and neighbour lists (binning.hip)
// ---------------------------------------------------------------------------
struct cell_grid {
    unsigned ncell[3];
};
// cell index per component: (ux * ncell) >> 32, linearised x fastest; with
// sub_bits > 0 the key is (cell << sub_bits) | the x sub-bin of the cell,
// (ux * ncell << sub_bits) >> 32 mod 2^sub_bits
void launch_compute_cell(uint4 const* posq, unsigned n, cell_grid g, uint32_t* keys, uint32_t* vals, hipStream_t s,
                         unsigned sub_bits = 0);
// offset[c] = first sorted index with (key >> shift) >= c, c = 0 .. nbins
// (nbins + 1 entries)
void launch_cell_offsets(uint32_t const* sorted_keys, unsigned n, unsigned nbins, uint32_t* offset, hipStream_t s,
                         unsigned shift = 0);
// x sub-bins per cell of the particle order that cell-segment lists use
constexpr unsigned SUB_BITS = 3;
constexpr unsigned SUB_BINS = 1u << SUB_BITS;
// cell_pos[k] = fixed-point position of particle vals[k], w = (vals[k] << 2) | species
void launch_gather_cell_pos(uint4 const* posq, uint32_t const* sorted_vals, unsigned n, uint4* cell_pos, hipStream_t s);

struct neighbour_params {
    cell_grid grid;
    float q[3];         // float(L) * 2^-32: length of one fixed-point unit
    unsigned ntype;
    float rr_cut_skin[MAX_TYPES * MAX_TYPES];
    // cubic box (q equal in all dimensions): predicate in fixed-point units,
    // rq = fma(dz, dz, fma(dy, dy, dx * dx)) <= float(rr_cut_skin / q^2)
    unsigned iso;
    float rr_cut_skin_q[MAX_TYPES * MAX_TYPES];
    // cubic box, cell-segment build: integer predicate on the top 16 bits of
    // each fixed-point coordinate, dx^2 + dy^2 + dz^2 <= rr16 with d the
    // wrapped 16-bit difference (exact); see list_threshold16
    uint32_t rr16[MAX_TYPES * MAX_TYPES];
};
// rr16 for the distance r_cut + r_skin in a cubic box of edge L: the 16-bit
// differences are within 1 of the exact differences in units of L / 2^16, so
// the 16-bit distance is within sqrt(3) of the exact one, and every pair at an
// exact distance <= r_cut + r_skin passes (pairs up to 2 sqrt(3) units beyond
// may pass as well); the margin over sqrt(3) absorbs the FP64 rounding here
inline uint32_t list_threshold16(double r_cut_skin, double L) {
    double const r = r_cut_skin * 65536.0 / L + 1.7330508075688772;
    return uint32_t(r * r);
}
// full neighbour lists, blocked layout: entry k of particle i at
// (i / 64) * 64 * cap + (k / NB_BLOCK) * 64 * NB_BLOCK + (i % 64) * NB_BLOCK + k % NB_BLOCK
// (cap a multiple of NB_BLOCK); *max_count receives the largest per-particle
// count (may exceed cap)
constexpr unsigned NB_BLOCK = 16;
__host__ __device__ inline size_t list_index(size_t i, size_t k, size_t cap) {
    return (i & ~size_t(WAVE - 1)) * cap + (k / NB_BLOCK) * WAVE * NB_BLOCK + (i & (WAVE - 1)) * NB_BLOCK +
           k % NB_BLOCK;
}
// self = false: cell_pos and offset are the binning of a second particle
// instance (no entry is skipped as particle i itself)
void launch_build_neighbours(uint4 const* posq, unsigned n, uint32_t const* offset, uint4 const* cell_pos,
                             neighbour_params const& p, unsigned cap, uint32_t* list, uint32_t* count,
                             uint32_t* max_count, hipStream_t s, bool self = true);

// A rebuild decided on the device: rebuild only if the largest of the
// displacement words at slots (DISP_SLOTS words, DISP_STRIDE apart) exceeds
// keep_bits, and then also set r0[i] = hi(pos[i]) and add 1 to *rebuilt.
// slots == nullptr: rebuild unconditionally (the other fields unused).
struct rebuild_condition {
    uint32_t const* slots = nullptr;
    uint32_t keep_bits = 0;
    float4 const* pos = nullptr;
    float4* r0 = nullptr;
    uint32_t* rebuilt = nullptr;
};

// lists from all particle pairs (small boxes): ascending j, the same predicate
// and list layout as launch_build_neighbours
void launch_build_neighbours_all_pairs(uint4 const* posq, unsigned n, neighbour_params const& p, unsigned cap,
                                       uint32_t* list, uint32_t* count, uint32_t* max_count, hipStream_t s,
                                       rebuild_condition const& cond = rebuild_condition());

// ---------------------------------------------------------------------------
// cell segments (binning.hip, force.hip): particles stored in cell order, one
// workgroup per segment of up to S consecutive cells along x in one (y, z)
// row.  The union of the 27-cell neighbourhoods of a segment's cells is 9 rows
// (dz, dy) of S' + 2 cells (x0 - 1 .. x0 + S', periodic), each one or two
// contiguous ranges of particle memory; it is staged in LDS, rows in (dz, dy)
// lexicographic order, cells in that x order.  Neighbour lists then hold
// 16-bit LDS slots of that union.
// ---------------------------------------------------------------------------
constexpr unsigned SEG_MAX = 32;          // largest S
constexpr unsigned SEG_THREADS_MAX = 256; // largest workgroup of the segment kernels
constexpr unsigned SEG_XCD = 8;           // workgroups are dealt round-robin to 8 XCDs
constexpr unsigned SEG_RMAX = 8;           // largest R
constexpr unsigned SEG_NR = 3 * (SEG_RMAX + 2);   // most union rows
struct seg_params {
    cell_grid grid;
    unsigned S;           // cells per segment along x, 1 <= S <= min(SEG_MAX, nx - 2)
    unsigned R;           // rows per segment along y, 1 <= R <= min(SEG_RMAX, ny - 2)
    unsigned nsx;         // segments along x, ceil(nx / S)
    unsigned nsy;         // segments along y, ceil(ny / R)
    unsigned nseg;        // nsx * nsy * nz
    unsigned chunk;       // ceil(nseg / SEG_XCD): consecutive segments per XCD
    unsigned lds_cap;     // LDS records per workgroup (>= every union size)
    unsigned threads;     // workgroup size (a multiple of 64, <= SEG_THREADS_MAX)
};
// workgroup b serves segment (b % 8) * chunk + b / 8, so that each XCD walks a
// contiguous run of segments and re-reads shared halo rows from its own L2
__host__ __device__ inline unsigned seg_of_block(seg_params const& sp, unsigned b) {
    return (b % SEG_XCD) * sp.chunk + b / SEG_XCD;
}
// position of segment `seg`: cells x0 .. x0 + w - 1 of rows y0 .. y0 + h - 1 in plane z
struct seg_place {
    unsigned x0, w, y0, h, z;
};
__host__ __device__ inline seg_place seg_locate(seg_params const& sp, unsigned seg) {
    unsigned const nx = sp.grid.ncell[0], ny = sp.grid.ncell[1];
    seg_place q;
    unsigned const sx = seg % sp.nsx, t = seg / sp.nsx;
    unsigned const sy = t % sp.nsy;
    q.z = t / sp.nsy;
    q.x0 = sx * sp.S;
    q.w = q.x0 + sp.S <= nx ? sp.S : nx - q.x0;
    q.y0 = sy * sp.R;
    q.h = q.y0 + sp.R <= ny ? sp.R : ny - q.y0;
    return q;
}
// Union row ur (0 .. 3 (h + 2) - 1; plane dz = ur / (h + 2) - 1, row y0 - 1 +
// ur % (h + 2), periodic) of the segment's 27-cell neighbourhoods: cells
// x0 - 1 .. x0 + w as up to two memory ranges (begin, length), from the cell
// offsets (device or host memory)
__host__ __device__ inline void seg_row(seg_params const& sp, seg_place const& q, unsigned ur, uint32_t const* off,
                                        unsigned* g, unsigned* len) {
    unsigned const nx = sp.grid.ncell[0], ny = sp.grid.ncell[1], nz = sp.grid.ncell[2];
    unsigned const hh = q.h + 2;
    unsigned const kz = (q.z + nz + ur / hh - 1) % nz, ky = (q.y0 + ny + ur % hh - 1) % ny;
    unsigned const row = (kz * ny + ky) * nx;
    unsigned const x0 = q.x0, w = q.w;
    if (x0 > 0 && x0 + w < nx) {
        g[0] = off[row + x0 - 1]; len[0] = off[row + x0 + w + 1] - g[0];
        g[1] = 0; len[1] = 0;
    } else if (x0 == 0) {
        g[0] = off[row + nx - 1]; len[0] = off[row + nx] - g[0];
        g[1] = off[row]; len[1] = off[row + w + 1] - g[1];
    } else {
        g[0] = off[row + x0 - 1]; len[0] = off[row + nx] - g[0];
        g[1] = off[row]; len[1] = off[row + 1] - g[1];
    }
}
// the segment's own particles of row j (0 .. h - 1): memory range [begin, begin + count)
__host__ __device__ inline void seg_own_row(seg_params const& sp, seg_place const& q, unsigned j, uint32_t const* off,
                                            unsigned& begin, unsigned& count) {
    unsigned const nx = sp.grid.ncell[0], ny = sp.grid.ncell[1];
    unsigned const row = (q.z * ny + q.y0 + j) * nx;
    begin = off[row + q.x0];
    count = off[row + q.x0 + q.w] - begin;
}
// 16-bit slot lists (an entry is the byte offset 16 * slot of the record in
// the staged union, so a union holds at most 4096 records): entry k of
// particle i (memory order) at
// (i / 64) * 64 * cap + (k / 16) * 1024 + (i % 64) * 16 + k % 16 (uint16 units;
// one lane block = 16 entries = 32 bytes, the write granule of the memory
// system; cap a multiple of 16)
constexpr unsigned NB16_BLOCK = 16;
constexpr unsigned SEG_MAX_RECORDS = 4096;   // staged union plus sentinel
__host__ __device__ inline size_t list16_index(size_t i, size_t k, size_t cap) {
    return (i & ~size_t(WAVE - 1)) * cap + (k / NB16_BLOCK) * WAVE * NB16_BLOCK + (i & (WAVE - 1)) * NB16_BLOCK +
           k % NB16_BLOCK;
}
// Lists are padded with the slot U of a sentinel record (staged after the
// union, at least a third of the box away from every particle of the segment
// in y and z) to a multiple of 8 entries; count[i] excludes the padding.
// stats[0]: largest count (may exceed cap); stats[1]: largest staging size
// U + 1 (a segment whose staging exceeds sp.lds_cap builds nothing; the caller
// retries)
//
// With sub_offset (the offsets of the (cell, x sub-bin) order the particles
// are in, ncells SUB_BINS + 1 entries; cubic box, nx >= 4) each lane scans
// only the candidate windows its position allows (same lists); the dynamic
// LDS is then windowed_build_lds() bytes.
void launch_build_neighbours_staged(uint4 const* posq, uint32_t const* offset, seg_params const& sp,
                                    neighbour_params const& p, unsigned cap, uint16_t* list, uint32_t* count,
                                    uint32_t* stats, hipStream_t s, uint32_t const* sub_offset = nullptr);
constexpr unsigned WIN_MAX = 10;   // candidate windows per lane: 9 rows, the own row split
inline size_t windowed_build_lds(seg_params const& sp) {
    return (size_t(sp.lds_cap) + 3) * 8 + size_t(3) * (sp.R + 2) * ((sp.S + 2) * SUB_BINS + 1) * 4 +
           size_t(WIN_MAX) * sp.threads * 4;
}

// ---------------------------------------------------------------------------
// pair forces (force.hip)
// ---------------------------------------------------------------------------
// POTENTIAL_POWER_LAW and POTENTIAL_MIE share one device form (POTENTIAL_POWER_LAW):
// U = c_m (sigma/r)^m - c_n (sigma/r)^n with integer indices
enum potential_kind : int { POTENTIAL_LENNARD_JONES = 0, POTENTIAL_MORSE = 1, POTENTIAL_POWER_LAW = 2, POTENTIAL_MIE = 3 };
enum truncation_kind : int { TRUNC_SHARP = 0, TRUNC_SHIFTED = 1, TRUNC_FORCE_SHIFTED = 2, TRUNC_SMOOTH_R4 = 3 };

// per species pair, FP32
struct pair_param {
    float p[4];        // LJ: {48 eps / sigma^2, 4 eps, sigma^2, -}; Morse: {eps/A, 1/(sigma B), r_min_sigma/B, A};
                       // power law / Mie: {c_m, c_n, sigma^2, -}
    float p_lo[4];     // LJ, cubic boxes: sigma^2 / q^2 as a double word in p_lo[0], p_lo[1] (the
                       // kernels take the nearest float, p_lo[0]);
                       // power law / Mie: the indices {m, n} (exact small integers)
    float rr_cut, r_cut, en_cut, force_cut;
    float rr_cut_q;    // rr_cut in fixed-point units squared, rr_cut / q^2 (cubic boxes)
};
struct force_params {
    int kind, trunc;
    unsigned ntype;
    float rri_smooth;
    float q[3], q_lo[3];   // L * 2^-32 as a double word
    unsigned iso;          // cubic box: q the same in all dimensions
    float q2[2];           // q^2 as a double word (cubic boxes)
    float q_p0[2], q2_p0[2];   // q p[0] and q^2 p[0] of pair 0 as double words (cubic boxes, one species)
    pair_param pair[MAX_TYPES * MAX_TYPES];
};
// force_zero: overwrite force/en_pot/stress, else add to the stored values.
// fuse: the force kernel makes the Verlet step of each particle with its force
// (no aux; only for pair_force_fusable(n) particles)
struct fused_integrate;
bool pair_force_fusable(unsigned n);
// posq2: the positions the list indices refer to (a second particle instance;
// null: posq)
void launch_pair_force(uint4 const* posq, unsigned n, uint32_t const* list, uint32_t const* count, unsigned cap,
                       force_params const& p, bool aux, bool force_zero, float aux_weight, float4* force,
                       float* stress, unsigned stride, hipStream_t s, fused_integrate const* fuse = nullptr,
                       uint4 const* posq2 = nullptr);

// untruncated forces over all pairs of two particle instances (same: one
// instance, j != i); p as from potential::device_params for pair_full (sharp,
// infinite cutoff).  partial: pair_full_partial_floats(n1, n2) floats of
// scratch, at most PAIR_FULL_MAX_PARTS * PAIR_FULL_COMPONENTS * n1.
constexpr unsigned PAIR_FULL_MAX_PARTS = 64;
constexpr unsigned PAIR_FULL_COMPONENTS = 10;
size_t pair_full_partial_floats(unsigned n1, unsigned n2);
void launch_pair_full(uint4 const* posq1, unsigned n1, uint4 const* posq2, unsigned n2, bool same,
                      force_params const& p, bool aux, bool force_zero, float aux_weight, float4* force,
                      float* stress, unsigned stride, float* partial, hipStream_t s);

// velocity-Verlet step parameters (integrate.hip, and the fused force kernel)
struct integrate_params {
    float dt_hi, dt_lo, half_dt;
    float L_hi[3], L_lo[3], invL[3];
    quant_params qp;
};

// The velocity-Verlet step of each particle, done by the force kernel as soon
// as the particle's force is summed (instead of writing the force): the
// deferred finalize kick pre_half_dt and the half kick with that force, the
// drift, wrap and image, the fixed-point snapshot into posq_out (not the posq
// the kernel stages from), and the displacement from r0 folded into the slot
// set disp_write; disp_clear is zeroed for the update after.
struct fused_integrate {
    float4* pos;
    float4* vel;
    int4* image;
    uint4* posq_out;
    float4 const* r0;
    uint32_t* disp_write;
    uint32_t* disp_clear;
    integrate_params ip;
    float pre_half_dt;
};

// A speculative launch's check: it does nothing when the maximum of the
// DISP_SLOTS displacement words at slots exceeds bits (its lists are stale).
// Given publish, the launch also stores (seq << 32) | that maximum into
// publish[seq % PUBLISH_RING] (host-mapped memory) as it starts, so the host
// learns the displacement without a copy behind the stream.
struct displacement_check {
    uint32_t const* slots;
    uint32_t bits;
    unsigned long long* publish;
    uint32_t seq;
};
constexpr unsigned PUBLISH_RING = 4;

// the same over 16-bit slot lists of cell segments (particles in cell order);
// check (optional): a speculative launch; fuse (optional, no aux): integrate
// each particle instead of writing its force
void launch_pair_force_staged(uint4 const* posq, uint32_t const* offset, seg_params const& sp, uint16_t const* list,
                              uint32_t const* count, unsigned cap, force_params const& p, bool aux, bool force_zero,
                              float aux_weight, float4* force, float* stress, unsigned stride, hipStream_t s,
                              displacement_check const* check = nullptr, fused_integrate const* fuse = nullptr);

// ---------------------------------------------------------------------------
// integrators (integrate.hip)
// ---------------------------------------------------------------------------
// displacement maxima: each wave folds its maximum into slot (global wave
// index) % DISP_SLOTS of a slot set, the slots DISP_STRIDE words (256 bytes)
// apart so that their atomics proceed on different cache lines; the reader
// takes the largest slot
constexpr unsigned DISP_SLOTS = 64;
constexpr unsigned DISP_STRIDE = 64;
constexpr unsigned DISP_SET = DISP_SLOTS * DISP_STRIDE;   // words per slot set
// kick-drift-wrap and fixed-point snapshot; optional deferred finalize kick
// of the previous step (pre_half_dt != 0: v += F dt/2m first, the same
// arithmetic as the finalize kernel), optional velocity pre-scale (*scale,
// device) and optional max-displacement tracking against r0 (max |r - r0|^2 as
// float bits folded into the slot set at max_disp2; the slots of the set at
// clear_slot, that of the next step, are set to zero)
void launch_verlet_integrate(float4* pos, float4* vel, float4 const* force, int4* image, uint4* posq, unsigned n,
                             integrate_params const& p, float pre_half_dt, double const* scale, float4 const* r0,
                             uint32_t* max_disp2, uint32_t* clear_slot, hipStream_t s);
void launch_verlet_finalize(float4* vel, float4 const* force, unsigned n, float half_dt, hipStream_t s);
void launch_rescale_velocity(float4* vel, unsigned n, double const* scale, hipStream_t s);
void launch_copy_hi(float4 const* pos, float4* r0, unsigned n, hipStream_t s);
void launch_max_displacement(float4 const* pos, float4 const* r0, unsigned n, float const* L, float const* invL,
                             uint32_t* max_disp2, hipStream_t s);
// Andersen collisions: with probability p per particle resample v ~ N(0, sqrt(T))
// (counter-based RNG keyed by seed, step and particle id), else kick
void launch_andersen_finalize(float4* vel, float4 const* force, uint32_t const* id, unsigned n, float half_dt,
                              float coll_prob, float sqrt_temperature, uint64_t seed, uint64_t step, hipStream_t s);

// ---------------------------------------------------------------------------
// ensembles: many independent systems in shared arrays (ensemble.cpp drives
// these).  System s owns particles [base, base + n) (base a multiple of 256),
// cell offsets [cell_base, cell_base + ncells] (particle indices relative to
// base), list entries from list_off (uint16 units, per-system capacity cap).
// Each kernel runs all systems (or a listed subset) in one launch.
// ---------------------------------------------------------------------------
constexpr unsigned ENS_BLOCK = 256;          // particles per integrate block
struct ens_desc {
    unsigned base, n, blk_base;              // first particle, particles, first integrate block
    unsigned cell_base, ncells;
    unsigned small;                          // all-pairs lists (fewer than 3 cells per dimension)
    unsigned cap;                            // list entries per particle
    uint32_t keep_bits;                      // rebuild when the displacement maximum's float bits exceed it
    unsigned long long list_off;
    float skin_half_sq;                      // rebuild when |r - r0|^2 exceeds it
    seg_params sp;
    neighbour_params np;
    integrate_params ip;
    force_params P;
};
// displacement maxima: ENS_DISP_SLOTS words per system, ENS_DISP_STRIDE words
// apart (a wave folds its maximum into slot (global wave index) mod
// ENS_DISP_SLOTS, so that the waves of a large system spread their atomics);
// the maximum over a system's slots is its displacement maximum
constexpr unsigned ENS_DISP_SLOTS = 16;
constexpr unsigned ENS_DISP_STRIDE = 16;
constexpr unsigned ENS_DISP_WORDS = ENS_DISP_SLOTS * ENS_DISP_STRIDE;   // per system
// per-system state words of the device-decided step (skip[s]): a system with
// any bit set is left as it is by every kernel of the step
constexpr uint32_t ENS_SKIP_OVERFLOW = 1;   // its rebuild overflowed the list or staging capacity
constexpr uint32_t ENS_SKIP_FAR = 2;        // a particle left the 27 cells around its old cell
constexpr uint32_t ENS_SKIP_HOLD = 4;       // held by the host (catching up other systems)
// step part 1: deferred kick (if pending), kick, drift, snapshot; per-system
// displacement maximum (float bits) into the slots at disp + s ENS_DISP_WORDS;
// systems with skip[s] set (skip may be null) are left as they are
void launch_ens_integrate(float4* pos, float4* vel, float4 const* force, int4* image, uint4* posq,
                          float4 const* r0, ens_desc const* descs, uint16_t const* block_sys, unsigned nblocks,
                          bool pending, uint32_t* disp, uint32_t const* skip, hipStream_t s);
// the deferred finalize kick of every system, applied now
void launch_ens_settle(float4* vel, float4 const* force, ens_desc const* descs, uint16_t const* block_sys,
                       unsigned nblocks, hipStream_t s);
// rebuild of the listed systems.  rblk: per 256-particle block of a listed
// cell system, (rank << 16) | block; rpos[rank]: its first compacted slot
void launch_ens_cells(uint4 const* posq, ens_desc const* descs, uint32_t const* rsys, uint32_t const* rpos,
                      uint32_t const* rblk, unsigned nrblk, unsigned cell_bits, uint32_t* keys, uint32_t* vals,
                      hipStream_t s);
void launch_ens_offsets(ens_desc const* descs, uint32_t const* rsys, uint32_t const* rpos, uint32_t const* rblk,
                        unsigned nrblk, unsigned cell_bits, uint32_t const* sorted_keys, uint32_t* offsets,
                        hipStream_t s);
// reorder each listed system into its cell order (sorted vals are local
// indices in the old order); r0 := new positions (hi words)
void launch_ens_reorder(float4* pos, float4* vel, int4* image, uint32_t* id, uint4* posq, float4* r0,
                        float4* pos_t, float4* vel_t, int4* image_t, uint32_t* id_t, uint4* posq_t,
                        ens_desc const* descs, uint32_t const* rsys, uint32_t const* rpos, uint32_t const* rblk,
                        unsigned nrblk, uint32_t const* sorted_vals, hipStream_t s);
// r0 := positions of the listed systems (all-pairs systems, no reorder)
// moves of whole ranges between two arrays of 16-byte units: range k copies
// units [src_k * num / den, + n_k * num / den) of src to dst from dst_k * num / den
// (num / den: 16-byte units per element, every product an integer)
struct ens_move {
    unsigned long long src, dst, n;
};
void launch_ens_move(uint4* dst, uint4 const* src, ens_move const* moves, unsigned nmoves, unsigned num, unsigned den,
                     hipStream_t s);
void launch_ens_r0(float4 const* pos, float4* r0, ens_desc const* descs, uint32_t const* rblk_sys, unsigned nrblk,
                   hipStream_t s);
// particles per workgroup of the small-system (all-pairs) build and force
constexpr unsigned ENS_SMALL_P = 16;

// list builds: rseg (s << 16) | segment for cell systems; rsmall the
// all-pairs systems (a grid of workgroups each, ENS_SMALL_P particles per
// workgroup); stats[2 s .. 2 s + 1].  multi: cutoffs
// by species pair (the same lists as the one-species rule when ntype is 1)
void launch_ens_build(uint4 const* posq, uint32_t const* offsets, ens_desc const* descs, uint32_t const* rseg,
                      unsigned nrseg, unsigned threads, unsigned lds_records, bool multi, uint16_t* list,
                      uint32_t* count, uint32_t* stats, hipStream_t s);
void launch_ens_build_small(uint4 const* posq, ens_desc const* descs, uint32_t const* rsmall, unsigned nsmall,
                            unsigned max_n, uint16_t* list, uint32_t* count, uint32_t* stats, hipStream_t s);
// forces of one (kind, truncation, species) class: fseg (s << 16) | segment,
// fsmall the all-pairs systems; aux: energy in force.w and stress (stride =
// total); multi: pair parameters by species (Lennard-Jones, cell systems)
// fuse (optional, no aux): instead of writing the force, each particle makes
// its Verlet step with it (the pending finalize kick first when pending), the
// new snapshot goes to posq_out and the displacement maximum to the system's
// slots at disp; skipped systems have their snapshot copied to posq_out
struct ens_fuse {
    float4 *pos, *vel;
    int4* image;
    uint4* posq_out;
    float4 const* r0;
    uint32_t* disp;
    bool pending;
    // pass 1: only the systems with flags[s] == 0, pass 2: those with
    // flags[s] == 1 (0: all); the two passes of a step together take each
    // system once
    int pass;
    uint32_t const* flags;
};
void launch_ens_force(int kind, int trunc, bool aux, bool multi, uint4 const* posq, uint32_t const* offsets,
                      ens_desc const* descs, uint32_t const* fseg, unsigned nfseg, unsigned threads,
                      unsigned lds_records, uint32_t const* fsmall, unsigned nsmall, unsigned small_max_n,
                      uint16_t const* list, uint32_t const* count, float4* force, float* stress, unsigned stride,
                      uint32_t const* skip, ens_fuse const* fuse, hipStream_t s);

// Device-decided rebuilds (a step with no host round trip).  After the
// integrate launch, ens_decide flags each system whose displacement maximum
// exceeds its keep_bits and writes the work tables of the step's rebuild: the
// flagged cell systems in ascending index with exclusive prefix sums of their
// cells, 256-particle blocks and segments, and the flagged all-pairs systems
// with prefix sums of their blocks and ENS_SMALL_P-particle groups.  The
// rebuild kernels stride over those tables with fixed grids.
struct ens_work {
    uint32_t* head;          // [0] flagged cell systems, [1] flagged all-pairs systems
    uint32_t *lsys, *lcell, *lblk, *lseg, *lcb;   // cell systems (prefixes have one more entry)
    uint32_t *ssys, *sblk, *sgrp;                 // all-pairs systems
};
constexpr size_t ens_work_words(size_t nsys) { return 16 + 8 * (nsys + 1); }
constexpr unsigned ENS_CELL_BLOCK = 256;   // cells per workgroup of the new-offset scan
ens_work ens_work_at(uint32_t* words, size_t nsys);
// work tables, per-system step and rebuild counts (+1 each), flags[s] = 1
// for a system to rebuild (else 0), the displacement slots cleared, stats of
// flagged systems cleared
void launch_ens_decide(ens_desc const* descs, unsigned nsys, uint32_t const* skip, uint32_t* disp,
                       uint32_t* steps, uint32_t* rebuilds, uint32_t* stats, ens_work w, uint32_t* flags,
                       hipStream_t s);
// Rebuild of the flagged cell systems.  The particles of a system are in the
// order of its previous rebuild (cells ascending, offsets[] describe it), and
// none moves farther than into one of the 26 cells around its old one (one
// that does sets ENS_SKIP_FAR, and the host rebuilds that system).  The stable
// order by new cell is then a merge: cell c takes, from each of the 27 cells c
// + e in ascending index, the particles that moved by -e, in memory order.
//  - rank: waves over runs of old cells; keys[i] = new cell | code(e) << 27,
//    vals[i] = rank of i among the particles of its old cell with that code,
//    cnt27[27 c' + code] = their number
//  - cell_count: one thread per new cell; base27[27 c + code] = particles
//    that reach c from sources below c + e, offsets_t = the exclusive scan of
//    the counts within each block of ENS_CELL_BLOCK cells, bsum = block sums
//  - cell_carry: the sums of the earlier blocks added (offsets_t final)
//  - scatter: each particle to its new slot (into the twin arrays)
//  - copyback: twins back, r0 := new positions, offsets_t back
//  - build: the staged list build over the flagged segments
// Afterwards the all-pairs systems: r0 and their lists.
struct ens_rebuild_args {
    float4 *pos, *vel;
    int4* image;
    uint32_t* id;
    uint4* posq;
    float4* r0;
    float4 *pos_t, *vel_t;
    int4* image_t;
    uint32_t* id_t;
    uint4* posq_t;
    uint32_t *offsets, *offsets_t, *keys, *vals, *cnt27, *base27, *bsum;
    ens_desc const* descs;
    ens_work w;
    uint32_t* skip;
    unsigned threads, lds_records;   // list build (lds_records 0: no cell systems)
    bool multi;
    unsigned small_max_n;            // 0: no all-pairs systems
    uint16_t* list;
    uint32_t *count, *stats;
};
void launch_ens_rebuild_device(ens_rebuild_args const& a, hipStream_t s);
constexpr unsigned ENS_CELL_KEY_BITS = 27;   // cells per system below 2^27

// The same merge for one particle instance: the stable cell order of its
// present positions from the cell order of its previous binning (old_offsets
// describe the memory order; every particle has moved at most into one of the
// 26 cells around its old one).  Writes new_offsets (n_cells + 1 entries) and
// perm (the gather list of the new order); flag receives ENS_SKIP_FAR if a
// particle moved farther (then new_offsets and perm are not written).  The
// tables come from merge_tables_upload for this grid and n.
struct merge_scratch {
    ens_desc* desc;         // 1 descriptor
    uint32_t* work;         // ens_work_words(1) words
    uint32_t *keys, *rank;  // n words each
    uint32_t *cnt27, *base27, *bsum;   // 27 n_cells, 27 n_cells, n_cells / ENS_CELL_BLOCK + 2 words
    uint32_t* flag;         // 1 word
    unsigned n = 0, ncells = 0;   // set by merge_tables_upload
};
void merge_tables_upload(cell_grid const& g, unsigned n, merge_scratch& m, hipStream_t s);
void launch_merge_cell_order(uint4 const* posq, uint32_t const* old_offsets, uint32_t* new_offsets, uint32_t* perm,
                             merge_scratch const& m, hipStream_t s);
// per-system FP64 sums, one workgroup per listed system: out[4 s] = sum m v^2
// (velocities after the pending kick, when pending), [1] = sum en_pot, [2] =
// sum of the stress traces (both from aux_force / stress when given), [3] = n
void launch_ens_thermo(float4 const* vel, float4 const* force, float4 const* aux_force, float const* stress,
                       unsigned stride, ens_desc const* descs, uint32_t const* sys, unsigned nsys, bool pending,
                       double* out, hipStream_t s);

// ---------------------------------------------------------------------------
// reductions (reduce.hip): deterministic two-pass FP64 block sums
// ---------------------------------------------------------------------------
constexpr unsigned REDUCE_BLOCKS = 240;
constexpr unsigned REDUCE_THREADS = 256;
// per-block partial sums of: [0] sum m v^2, [1..3] sum m v, [4] sum m,
// [5] sum en_pot, [6] sum tr(stress), [7..12] sum stress components
constexpr unsigned THERMO_NQ = 13;
void launch_thermo_partials(float4 const* vel, float4 const* force, float const* stress, unsigned stride,
                            unsigned n, bool aux, double* partials, hipStream_t s);
// sum partials[REDUCE_BLOCKS][nq] in a fixed order into out[nq]
void launch_reduce_final(double const* partials, unsigned nq, double* out, hipStream_t s);

// NHC chain state on the device
struct nhc_state {
    double xi[2], v_xi[2], mass_xi[2];
    double temperature, en_kin_target_2, timestep;
    double scale;           // last velocity scale factor
    double en_nhc_raw;      // T (dN xi0 + xi1) + sum m_xi v_xi^2 / 2 (not divided by N)
};
// sum m v^2 partials -> chain half-step; writes state->scale
void launch_nhc_propagate(double const* partials, nhc_state* state, hipStream_t s);
void launch_mv2_partials(float4 const* vel, unsigned n, double* partials, hipStream_t s);

// ---------------------------------------------------------------------------
// Hilbert sort and particle rearrangement (hilbert.hip)
// ---------------------------------------------------------------------------
void launch_hilbert_map(float4 const* pos, unsigned n, float const* L, unsigned depth, uint32_t* code,
                        uint32_t* index, hipStream_t s);
// rid (optional): rid[id] := the particle's new index (the reverse map);
// r0 (optional): r0[i] := the hi word of the new position i
void launch_rearrange(float4 const* pos_in, float4 const* vel_in, int4 const* image_in, uint32_t const* id_in,
                      uint4 const* posq_in, uint32_t const* perm, unsigned n, float4* pos_out, float4* vel_out,
                      int4* image_out, uint32_t* id_out, uint4* posq_out, hipStream_t s, uint32_t* rid = nullptr,
                      float4* r0 = nullptr);
void launch_reverse_id(uint32_t const* id, unsigned n, uint32_t* rid, hipStream_t s);

// ---------------------------------------------------------------------------
// observables (observables.hip)
// ---------------------------------------------------------------------------
// rho_k = sum_j exp(-i k.r_j); q as double3 [nq]; out [nq] (re, im) pairs
constexpr unsigned DM_KCHUNK = 8;
constexpr unsigned DM_BLOCKS = 240;
void launch_density_mode(float4 const* pos, unsigned n, double const* q, unsigned nq, double* partials, double* out,
                         hipStream_t s);
// unwrapped positions r + image L in ID order, FP64
void launch_sample_position(float4 const* pos, int4 const* image, uint32_t const* id, unsigned n, double const* L,
                            double* out, hipStream_t s);
// The same two samples of many systems in one launch each.  Task t covers the
// virtual blocks [vb0, vb0 + its block count) of its launch (ascending vb0,
// the first at 0) and the density modes' final waves [vk0, vk0 + nq); each
// virtual block computes exactly what the block of the one-system launch with
// those indices computes, so the results are the same bit for bit.
struct position_task {
    float4 const* pos;
    int4 const* image;
    uint32_t const* id;
    double* out;
    double L[3];
    unsigned n, vb0;   // blocks: ceil(n / 256)
};
struct density_task {
    float4 const* pos;
    double const* q;
    double* partials;   // launch_density_mode's partials (its nblk blocks)
    double* out;        // 2 nq
    unsigned n, nq;
    unsigned nblk;      // particle blocks, as launch_density_mode chooses them for n
    unsigned vb0, vk0;  // blocks: nblk * ceil(nq / DM_KCHUNK); final waves: nq
};
unsigned density_mode_blocks(unsigned n);   // nblk of launch_density_mode
// tasks: device tables; vblocks / vwaves: the totals of the counts above
void launch_sample_tasks(position_task const* pos_tasks, unsigned npos, unsigned pos_vblocks,
                         density_task const* dm_tasks, unsigned ndm, unsigned dm_vblocks, unsigned dm_vwaves,
                         hipStream_t s);
// per-block partials of sum |r2 - r1|^2 and sum |r2 - r1|^4 over n particles,
// written to partials[block * nq + qoff + {0, 1}]
void launch_msd_partials(double const* r1, double const* r2, unsigned n, double* partials, unsigned nq, unsigned qoff,
                         hipStream_t s);
// the pairs (r1, r2[l]), l < nl, into partials[block][qoff + 2 l + {0, 1}] with
// one launch per MSD_LAGS pairs; each pair sums exactly as launch_msd_partials
constexpr unsigned MSD_LAGS = 32;
struct msd_samples {
    double const* p[MSD_LAGS];
};
void launch_msd_partials_lags(double const* r1, double const* const* r2, unsigned nl, unsigned n, double* partials,
                              unsigned nq, unsigned qoff, hipStream_t s);
// the same for the time correlations (halmd/observables/dynamics/*.hpp) of
// position samples, mean square (TCF_MSD) and quartic (TCF_MQD) displacement,
// and of velocity samples, the velocity autocorrelation (TCF_VACF)
enum tcf_kind : int { TCF_MSD = 0, TCF_MQD = 1, TCF_VACF = 2 };
void launch_tcf_partials_lags(int kind, double const* r1, double const* const* r2, unsigned nl, unsigned n,
                              double* partials, unsigned nq, unsigned qoff, hipStream_t s);
// velocity sample in particle-ID order: out[3 id + k] = v_hi + v_lo (FP64)
void launch_sample_velocity(float4 const* vel, uint32_t const* id, unsigned n, double* out, hipStream_t s);

} // namespace halmd
