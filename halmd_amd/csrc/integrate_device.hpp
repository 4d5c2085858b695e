// Device side of the velocity-Verlet step on double-word (float-float)
// positions and velocities, shared by the integrate kernels (integrate.hip)
// and the pair-force kernel that integrates its particles as it finishes them
// (force.hip).
#pragma once

#include "kernels.hpp"

namespace halmd {

// fixed-point coordinate u = floor(frac(x / L + 1/2) * 2^32) of the double-word
// x = hi + lo, in FP64 without contraction (so that the CPU restatement of the
// rule gives the same integers)
__device__ __forceinline__ uint32_t quantize(float hi, float lo, double invL) {
#pragma clang fp contract(off)
    double const x = double(hi) + double(lo);
    double t = x * invL;
    t = t + 0.5;
    t = t - floor(t);
    return uint32_t(uint64_t(t * 4294967296.0));
}

__device__ __forceinline__ uint4 quantize4(float4 const h, float4 const l, quant_params const& q) {
    return make_uint4(quantize(h.x, l.x, q.invL[0]), quantize(h.y, l.y, q.invL[1]), quantize(h.z, l.z, q.invL[2]),
                      __float_as_uint(h.w));
}

// fold the wave's maximum of d2 (>= 0) into *word (float bits) with an atomic
// whose result is not used: no barrier, and no wave waits for it
__device__ __forceinline__ void wave_max_to(float d2, uint32_t* word) {
    d2 = wave_max(d2);
    if (threadIdx.x % WAVE == 0 && d2 > 0.0f) atomicMax(word, __float_as_uint(d2));
}

// the slot of this wave in a slot set
__device__ __forceinline__ uint32_t* disp_slot(uint32_t* slots) {
    return slots + ((blockIdx.x * blockDim.x + threadIdx.x) / WAVE) % DISP_SLOTS * DISP_STRIDE;
}

__device__ __forceinline__ ff ff_from_double(double x) {
    float const hi = float(x);
    return ff{hi, float(x - double(hi))};
}

// Kick with the force f (optionally after the previous step's deferred
// finalize kick and a velocity scale), drift, wrap, image update and
// fixed-point snapshot (into posq) of particle i; returns |r - r0|^2 (0
// without r0).  Shared by the single-system and ensemble integrate kernels and
// the pair-force kernel that integrates its particles as it finishes them.
__device__ __forceinline__ float integrate_particle_f(float4* __restrict__ pos, float4* __restrict__ vel,
                                                     float4 const f, int4* __restrict__ image,
                                                     uint4* __restrict__ posq, unsigned i,
                                                     integrate_params const& p, float pre_half_dt,
                                                     double const* __restrict__ scale,
                                                     float4 const* __restrict__ r0) {
    float d2 = 0.0f;
    // every load up front (r0 too): a load issued after the stores would
    // wait for them and add a second memory round trip to each wave
    float4 rh = pos[2 * size_t(i)], rl = pos[2 * size_t(i) + 1];
    float4 vh = vel[2 * size_t(i)], vl = vel[2 * size_t(i) + 1];
    float4 const q0 = r0 ? r0[i] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    float const m = vh.w;
    ff vx{vh.x, vl.x}, vy{vh.y, vl.y}, vz{vh.z, vl.z};
    if (pre_half_dt != 0.0f) {
        // finalize of the previous step, as verlet_finalize_kernel computes it
        float const k = pre_half_dt / m;
        vx = ff_add_f(vx, f.x * k);
        vy = ff_add_f(vy, f.y * k);
        vz = ff_add_f(vz, f.z * k);
    }
    if (scale) {
        ff const s = ff_from_double(*scale);
        vx = ff_mul(vx, s); vy = ff_mul(vy, s); vz = ff_mul(vz, s);
    }
    float const k = p.half_dt / m;
    vx = ff_add_f(vx, f.x * k);
    vy = ff_add_f(vy, f.y * k);
    vz = ff_add_f(vz, f.z * k);
    ff const dt{p.dt_hi, p.dt_lo};
    ff rx = ff_add(ff{rh.x, rl.x}, ff_mul(vx, dt));
    ff ry = ff_add(ff{rh.y, rl.y}, ff_mul(vy, dt));
    ff rz = ff_add(ff{rh.z, rl.z}, ff_mul(vz, dt));
    float const ix = rintf(rx.hi * p.invL[0]);
    float const iy = rintf(ry.hi * p.invL[1]);
    float const iz = rintf(rz.hi * p.invL[2]);
    if (ix != 0.0f) { rx = ff_add_f(rx, -ix * p.L_hi[0]); rx = ff_add_f(rx, -ix * p.L_lo[0]); }
    if (iy != 0.0f) { ry = ff_add_f(ry, -iy * p.L_hi[1]); ry = ff_add_f(ry, -iy * p.L_lo[1]); }
    if (iz != 0.0f) { rz = ff_add_f(rz, -iz * p.L_hi[2]); rz = ff_add_f(rz, -iz * p.L_lo[2]); }
    float4 const nh = make_float4(rx.hi, ry.hi, rz.hi, rh.w), nl = make_float4(rx.lo, ry.lo, rz.lo, 0.0f);
    pos[2 * size_t(i)] = nh;
    pos[2 * size_t(i) + 1] = nl;
    posq[i] = quantize4(nh, nl, p.qp);
    vel[2 * size_t(i)] = make_float4(vx.hi, vy.hi, vz.hi, m);
    vel[2 * size_t(i) + 1] = make_float4(vx.lo, vy.lo, vz.lo, 0.0f);
    if (ix != 0.0f || iy != 0.0f || iz != 0.0f) {
        int4 im = image[i];
        im.x += int(ix); im.y += int(iy); im.z += int(iz);
        image[i] = im;
    }
    if (r0) {
        float dx = rx.hi - q0.x, dy = ry.hi - q0.y, dz = rz.hi - q0.z;
        dx -= rintf(dx * p.invL[0]) * p.L_hi[0];
        dy -= rintf(dy * p.invL[1]) * p.L_hi[1];
        dz -= rintf(dz * p.invL[2]) * p.L_hi[2];
        d2 = dx * dx + dy * dy + dz * dz;
    }
    return d2;
}

// the same with the force read from memory
__device__ __forceinline__ float integrate_particle(float4* __restrict__ pos, float4* __restrict__ vel,
                                                   float4 const* __restrict__ force, int4* __restrict__ image,
                                                   uint4* __restrict__ posq, unsigned i, integrate_params const& p,
                                                   float pre_half_dt, double const* __restrict__ scale,
                                                   float4 const* __restrict__ r0) {
    return integrate_particle_f(pos, vel, force[i], image, posq, i, p, pre_half_dt, scale, r0);
}

} // namespace halmd
