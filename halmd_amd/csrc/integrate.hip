// Velocity-Verlet integrators on double-word (float-float) positions and
// velocities for gfx950.
//
// integrate: v += f dt/2 / m; r += v dt; wrap r into the box, image += shift
//            (halmd/mdsim/gpu/integrators/verlet_kernel.cu:37-72)
// finalize:  v += f dt/2 / m (verlet_kernel.cu:77-98)
// Nose-Hoover: the velocity scale factor lives on the device (written by the
// chain kernel in reduce.hip), so neither half-step needs a host round trip
// (the reference propagates the chain on the host, verlet_nvt_hoover.cpp:188-222).
//
// The time step is carried as a double word (dt_hi + dt_lo) in the drift, and
// the box length as L_hi + L_lo in the wrap, so that neither the FP32 rounding
// of dt nor of L enters the trajectory.  The integrate kernel also tracks the
// largest squared displacement since the last neighbour-list build (high
// words, minimum image), reduced per wave and folded with atomicMax into one
// of DISP_SLOTS device words on separate cache lines (by wave index, so that
// no line sees many atomics; the reader takes the largest): the rebuild criterion of
// halmd/mdsim/gpu/max_displacement.cpp without a separate pass over the
// positions.
#include "integrate_device.hpp"

namespace halmd {
namespace {

__global__ void quantize_kernel(float4 const* __restrict__ pos, unsigned n, quant_params q, uint4* __restrict__ posq) {
    unsigned const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) posq[i] = quantize4(pos[2 * size_t(i)], pos[2 * size_t(i) + 1], q);
}

__global__ __launch_bounds__(256) void verlet_integrate_kernel(
    float4* __restrict__ pos, float4* __restrict__ vel, float4 const* __restrict__ force, int4* __restrict__ image,
    uint4* __restrict__ posq, unsigned n, integrate_params p, float pre_half_dt, double const* __restrict__ scale,
    float4 const* __restrict__ r0, uint32_t* __restrict__ max_disp2, uint32_t* __restrict__ clear_slot)
{
    unsigned const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (clear_slot && i < DISP_SLOTS) clear_slot[i * DISP_STRIDE] = 0u;
    float d2 = 0.0f;
    if (i < n) d2 = integrate_particle(pos, vel, force, image, posq, i, p, pre_half_dt, scale, r0);
    if (max_disp2) wave_max_to(d2, disp_slot(max_disp2));
}

__global__ __launch_bounds__(256) void verlet_finalize_kernel(float4* __restrict__ vel, float4 const* __restrict__ force,
                                                              unsigned n, float half_dt) {
    unsigned const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 const vh = vel[2 * size_t(i)], vl = vel[2 * size_t(i) + 1];
    float4 const f = force[i];
    float const m = vh.w;
    float const k = half_dt / m;   // the same kick as the deferred one in integrate_particle
    ff const vx = ff_add_f(ff{vh.x, vl.x}, f.x * k);
    ff const vy = ff_add_f(ff{vh.y, vl.y}, f.y * k);
    ff const vz = ff_add_f(ff{vh.z, vl.z}, f.z * k);
    vel[2 * size_t(i)] = make_float4(vx.hi, vy.hi, vz.hi, m);
    vel[2 * size_t(i) + 1] = make_float4(vx.lo, vy.lo, vz.lo, 0.0f);
}

__global__ __launch_bounds__(256) void rescale_velocity_kernel(float4* __restrict__ vel, unsigned n,
                                                               double const* __restrict__ scale) {
    unsigned const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    ff const s = ff_from_double(*scale);
    float4 const vh = vel[2 * size_t(i)], vl = vel[2 * size_t(i) + 1];
    ff const vx = ff_mul(ff{vh.x, vl.x}, s);
    ff const vy = ff_mul(ff{vh.y, vl.y}, s);
    ff const vz = ff_mul(ff{vh.z, vl.z}, s);
    vel[2 * size_t(i)] = make_float4(vx.hi, vy.hi, vz.hi, vh.w);
    vel[2 * size_t(i) + 1] = make_float4(vx.lo, vy.lo, vz.lo, 0.0f);
}

__global__ void copy_hi_kernel(float4 const* __restrict__ pos, float4* __restrict__ r0, unsigned n) {
    unsigned const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) r0[i] = pos[2 * size_t(i)];
}

__global__ __launch_bounds__(256) void max_displacement_kernel(float4 const* __restrict__ pos,
                                                               float4 const* __restrict__ r0, unsigned n, float Lx,
                                                               float Ly, float Lz, float ix_, float iy_, float iz_,
                                                               uint32_t* __restrict__ max_disp2) {
    unsigned const i = blockIdx.x * blockDim.x + threadIdx.x;
    float d2 = 0.0f;
    if (i < n) {
        float4 const r = pos[2 * size_t(i)], q = r0[i];
        float dx = r.x - q.x, dy = r.y - q.y, dz = r.z - q.z;
        dx -= rintf(dx * ix_) * Lx;
        dy -= rintf(dy * iy_) * Ly;
        dz -= rintf(dz * iz_) * Lz;
        d2 = dx * dx + dy * dy + dz * dz;
    }
    wave_max_to(d2, disp_slot(max_disp2));
}

// SplitMix64 finaliser (Steele, Lea & Flood, "Fast splittable pseudorandom
// number generators", OOPSLA 2014) as a counter-based generator
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// uniform in (0, 1]
__device__ __forceinline__ float u01(uint64_t x) { return float((x >> 40) + 1) * (1.0f / 16777216.0f); }

__global__ __launch_bounds__(256) void andersen_finalize_kernel(float4* __restrict__ vel,
                                                                float4 const* __restrict__ force,
                                                                uint32_t const* __restrict__ id, unsigned n,
                                                                float half_dt, float coll_prob, float sqrt_t,
                                                                uint64_t seed, uint64_t step) {
    unsigned const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 const vh = vel[2 * size_t(i)], vl = vel[2 * size_t(i) + 1];
    float const m = vh.w;
    // counter: (seed, step, particle id) -> independent of the memory order
    uint64_t const key = mix64(seed ^ mix64(step * 0x632be59bd9b4e019ull + id[i]));
    if (u01(mix64(key)) > coll_prob) {
        float4 const f = force[i];
        float const k = half_dt / m;
        ff const vx = ff_add_f(ff{vh.x, vl.x}, f.x * k);
        ff const vy = ff_add_f(ff{vh.y, vl.y}, f.y * k);
        ff const vz = ff_add_f(ff{vh.z, vl.z}, f.z * k);
        vel[2 * size_t(i)] = make_float4(vx.hi, vy.hi, vz.hi, m);
        vel[2 * size_t(i) + 1] = make_float4(vx.lo, vy.lo, vz.lo, 0.0f);
    } else {
        // Box-Muller from four uniforms; the heat bath draws N(0, sqrt(T)) per
        // component (verlet_nvt_andersen.cpp finalize)
        float const u1 = u01(mix64(key + 1)), u2 = u01(mix64(key + 2));
        float const u3 = u01(mix64(key + 3)), u4 = u01(mix64(key + 4));
        float const r1 = sqrt_t * sqrtf(-2.0f * logf(u1)), r2 = sqrt_t * sqrtf(-2.0f * logf(u3));
        float s1, c1, s2, c2;
        sincospif(2.0f * u2, &s1, &c1);
        sincospif(2.0f * u4, &s2, &c2);
        vel[2 * size_t(i)] = make_float4(r1 * c1, r1 * s1, r2 * c2, m);
        vel[2 * size_t(i) + 1] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
}

// ensembles: block b serves system block_sys[b] (systems start on block
// boundaries); one displacement maximum per system, spread over its slots
__global__ __launch_bounds__(ENS_BLOCK) void ens_integrate_kernel(
    float4* __restrict__ pos, float4* __restrict__ vel, float4 const* __restrict__ force, int4* __restrict__ image,
    uint4* __restrict__ posq, float4 const* __restrict__ r0, ens_desc const* __restrict__ descs,
    uint16_t const* __restrict__ block_sys, bool pending, uint32_t* __restrict__ disp, uint32_t const* skip)
{
    unsigned const s = block_sys[blockIdx.x];
    if (skip && skip[s]) return;   // the whole workgroup (one system per block)
    ens_desc const& d = descs[s];
    unsigned const li = (blockIdx.x - d.blk_base) * ENS_BLOCK + threadIdx.x;
    float d2 = 0.0f;
    if (li < d.n)
        d2 = integrate_particle(pos, vel, force, image, posq, d.base + li, d.ip, pending ? d.ip.half_dt : 0.0f,
                                nullptr, r0);
    unsigned const slot = ((blockIdx.x * ENS_BLOCK + threadIdx.x) / WAVE) % ENS_DISP_SLOTS;
    wave_max_to(d2, disp + size_t(s) * ENS_DISP_WORDS + slot * ENS_DISP_STRIDE);
}

// apply the deferred finalize kick of every system (before velocities are read)
__global__ __launch_bounds__(ENS_BLOCK) void ens_settle_kernel(float4* __restrict__ vel, float4 const* __restrict__ force,
                                                               ens_desc const* __restrict__ descs,
                                                               uint16_t const* __restrict__ block_sys) {
    ens_desc const& d = descs[block_sys[blockIdx.x]];
    unsigned const li = (blockIdx.x - d.blk_base) * ENS_BLOCK + threadIdx.x;
    if (li >= d.n) return;
    size_t const i = d.base + li;
    float4 const vh = vel[2 * i], vl = vel[2 * i + 1];
    float4 const f = force[i];
    float const m = vh.w;
    float const k = d.ip.half_dt / m;
    ff const vx = ff_add_f(ff{vh.x, vl.x}, f.x * k);
    ff const vy = ff_add_f(ff{vh.y, vl.y}, f.y * k);
    ff const vz = ff_add_f(ff{vh.z, vl.z}, f.z * k);
    vel[2 * i] = make_float4(vx.hi, vy.hi, vz.hi, m);
    vel[2 * i + 1] = make_float4(vx.lo, vy.lo, vz.lo, 0.0f);
}

} // namespace

void launch_verlet_integrate(float4* pos, float4* vel, float4 const* force, int4* image, uint4* posq, unsigned n,
                             integrate_params const& p, float pre_half_dt, double const* scale, float4 const* r0,
                             uint32_t* max_disp2, uint32_t* clear_slot, hipStream_t s) {
    if (!n) return;
    verlet_integrate_kernel<<<div_up(n, 256), 256, 0, s>>>(pos, vel, force, image, posq, n, p, pre_half_dt, scale, r0,
                                                           max_disp2, clear_slot);
    HIP_CHECK(hipGetLastError());
}

void launch_quantize(float4 const* pos, unsigned n, quant_params const& q, uint4* posq, hipStream_t s) {
    if (!n) return;
    quantize_kernel<<<div_up(n, 256), 256, 0, s>>>(pos, n, q, posq);
    HIP_CHECK(hipGetLastError());
}

void launch_verlet_finalize(float4* vel, float4 const* force, unsigned n, float half_dt, hipStream_t s) {
    if (!n) return;
    verlet_finalize_kernel<<<div_up(n, 256), 256, 0, s>>>(vel, force, n, half_dt);
    HIP_CHECK(hipGetLastError());
}

void launch_rescale_velocity(float4* vel, unsigned n, double const* scale, hipStream_t s) {
    if (!n) return;
    rescale_velocity_kernel<<<div_up(n, 256), 256, 0, s>>>(vel, n, scale);
    HIP_CHECK(hipGetLastError());
}

void launch_copy_hi(float4 const* pos, float4* r0, unsigned n, hipStream_t s) {
    if (!n) return;
    copy_hi_kernel<<<div_up(n, 256), 256, 0, s>>>(pos, r0, n);
    HIP_CHECK(hipGetLastError());
}

void launch_max_displacement(float4 const* pos, float4 const* r0, unsigned n, float const* L, float const* invL,
                             uint32_t* max_disp2, hipStream_t s) {
    if (!n) return;
    max_displacement_kernel<<<div_up(n, 256), 256, 0, s>>>(pos, r0, n, L[0], L[1], L[2], invL[0], invL[1], invL[2],
                                                           max_disp2);
    HIP_CHECK(hipGetLastError());
}

void launch_andersen_finalize(float4* vel, float4 const* force, uint32_t const* id, unsigned n, float half_dt,
                              float coll_prob, float sqrt_temperature, uint64_t seed, uint64_t step, hipStream_t s) {
    if (!n) return;
    andersen_finalize_kernel<<<div_up(n, 256), 256, 0, s>>>(vel, force, id, n, half_dt, coll_prob, sqrt_temperature,
                                                            seed, step);
    HIP_CHECK(hipGetLastError());
}

void launch_ens_integrate(float4* pos, float4* vel, float4 const* force, int4* image, uint4* posq,
                          float4 const* r0, ens_desc const* descs, uint16_t const* block_sys, unsigned nblocks,
                          bool pending, uint32_t* disp, uint32_t const* skip, hipStream_t s) {
    if (!nblocks) return;
    ens_integrate_kernel<<<nblocks, ENS_BLOCK, 0, s>>>(pos, vel, force, image, posq, r0, descs, block_sys, pending,
                                                       disp, skip);
    HIP_CHECK(hipGetLastError());
}

// blockIdx.y selects the range, blockIdx.x strides through it
__global__ __launch_bounds__(256) void ens_move_kernel(uint4* __restrict__ dst, uint4 const* __restrict__ src,
                                                       ens_move const* __restrict__ moves, unsigned num,
                                                       unsigned den) {
    ens_move const m = moves[blockIdx.y];
    size_t const s0 = size_t(m.src) * num / den, d0 = size_t(m.dst) * num / den, len = size_t(m.n) * num / den;
    for (size_t k = size_t(blockIdx.x) * blockDim.x + threadIdx.x; k < len; k += size_t(gridDim.x) * blockDim.x)
        dst[d0 + k] = src[s0 + k];
}

void launch_ens_move(uint4* dst, uint4 const* src, ens_move const* moves, unsigned nmoves, unsigned num, unsigned den,
                     hipStream_t s) {
    if (!nmoves) return;
    ens_move_kernel<<<dim3(32, nmoves), 256, 0, s>>>(dst, src, moves, num, den);
    HIP_CHECK(hipGetLastError());
}

void launch_ens_settle(float4* vel, float4 const* force, ens_desc const* descs, uint16_t const* block_sys,
                       unsigned nblocks, hipStream_t s) {
    if (!nblocks) return;
    ens_settle_kernel<<<nblocks, ENS_BLOCK, 0, s>>>(vel, force, descs, block_sys);
    HIP_CHECK(hipGetLastError());
}

} // namespace halmd
