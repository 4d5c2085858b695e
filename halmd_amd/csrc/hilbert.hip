// Hilbert-curve ordering of particles for gfx950.
//
// Device rule of halmd/mdsim/gpu/sorts/hilbert_kernel.cu:38-69 and
// halmd/mdsim/sorts/hilbert_kernel.hpp:53-230: the position is divided by the
// box length (FP32), snapped to the centre of its cell on the 2^depth grid of
// fractional coordinates, and mapped to a 3*depth-bit code by the vertex-table
// recursion of Wang & Shan ("Space-filling curve based point clouds index",
// GeoComputation 2005).  The stable radix sort of (code, index) gives the
// permutation; the rearrange kernel gathers the 32-byte position and velocity
// records, the image and the particle id, and reverse_id is rebuilt by a
// scatter (the reference radix-sorts the ids instead, particle.cpp:214-217;
// both give reverse_id[id[i]] = i).
#include "kernels.hpp"

namespace halmd {
namespace {

__device__ __forceinline__ void vswap(unsigned& v, unsigned& a, unsigned& b) {
    unsigned const va = (v >> a) & 7u;
    unsigned const vb = (v >> b) & 7u;
    v = v ^ (va << a) ^ (vb << b) ^ (va << b) ^ (vb << a);
    unsigned const t = a;
    a = b;
    b = t;
}

__device__ unsigned hilbert_code(float x, float y, float z, unsigned depth) {
#pragma clang fp contract(off)
    float const n = float(1u << depth);
    float r[3] = {x, y, z};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        r[k] = (r[k] - floorf(r[k])) * n;
        r[k] = (floorf(r[k]) + 0.5f) / n;
        r[k] -= 0.5f;
    }
    // code-to-vertex bit positions and vertex-to-code table
    unsigned a = 21, b = 18, c = 12, d = 15, e = 3, f = 0, g = 6, h = 9;
    unsigned vc = 1u << b ^ 2u << c ^ 3u << d ^ 4u << e ^ 5u << f ^ 6u << g ^ 7u << h;
    unsigned code = 0;
    for (unsigned i = 0; i < depth; ++i) {
        unsigned const bx = signbit(r[0]) ? 1u : 0u;
        unsigned const by = signbit(r[1]) ? 1u : 0u;
        unsigned const bz = signbit(r[2]) ? 1u : 0u;
        unsigned const v = (vc >> (3 * (bx + (by << 1) + (bz << 2)))) & 7u;
        r[0] = r[0] * 2.0f; r[0] = r[0] + float(bx); r[0] = r[0] - 0.5f;
        r[1] = r[1] * 2.0f; r[1] = r[1] + float(by); r[1] = r[1] - 0.5f;
        r[2] = r[2] * 2.0f; r[2] = r[2] + float(bz); r[2] = r[2] - 0.5f;
        if (v == 0) { vswap(vc, b, h); vswap(vc, c, e); }
        else if (v == 1 || v == 2) { vswap(vc, c, g); vswap(vc, d, h); }
        else if (v == 3 || v == 4) { vswap(vc, a, c); vswap(vc, f, h); }
        else if (v == 5 || v == 6) { vswap(vc, a, e); vswap(vc, b, f); }
        else { vswap(vc, a, g); vswap(vc, d, f); }
        code = (code << 3) + v;
    }
    return code;
}

__global__ void hilbert_map_kernel(float4 const* __restrict__ pos, unsigned n, float Lx, float Ly, float Lz,
                                   unsigned depth, uint32_t* __restrict__ code, uint32_t* __restrict__ index) {
#pragma clang fp contract(off)
    unsigned const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 const r = pos[2 * size_t(i)];
    code[i] = hilbert_code(r.x / Lx, r.y / Ly, r.z / Lz, depth);
    index[i] = i;
}

__global__ void rearrange_kernel(float4 const* __restrict__ pos_in, float4 const* __restrict__ vel_in,
                                 int4 const* __restrict__ image_in, uint32_t const* __restrict__ id_in,
                                 uint4 const* __restrict__ posq_in, uint32_t const* __restrict__ perm, unsigned n,
                                 float4* __restrict__ pos_out, float4* __restrict__ vel_out,
                                 int4* __restrict__ image_out, uint32_t* __restrict__ id_out,
                                 uint4* __restrict__ posq_out, uint32_t* __restrict__ rid,
                                 float4* __restrict__ r0) {
    unsigned const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    size_t const j = perm[i];
    if (rid) rid[id_in[j]] = i;   // the reverse map of the new order
    float4 const hi = pos_in[2 * j];
    if (r0) r0[i] = hi;
    pos_out[2 * size_t(i)] = hi;
    pos_out[2 * size_t(i) + 1] = pos_in[2 * j + 1];
    vel_out[2 * size_t(i)] = vel_in[2 * j];
    vel_out[2 * size_t(i) + 1] = vel_in[2 * j + 1];
    image_out[i] = image_in[j];
    id_out[i] = id_in[j];
    posq_out[i] = posq_in[j];
}

__global__ void reverse_id_kernel(uint32_t const* __restrict__ id, unsigned n, uint32_t* __restrict__ rid) {
    unsigned const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) rid[id[i]] = i;
}

} // namespace

void launch_hilbert_map(float4 const* pos, unsigned n, float const* L, unsigned depth, uint32_t* code, uint32_t* index,
                        hipStream_t s) {
    if (!n) return;
    hilbert_map_kernel<<<div_up(n, 256), 256, 0, s>>>(pos, n, L[0], L[1], L[2], depth, code, index);
    HIP_CHECK(hipGetLastError());
}

void launch_rearrange(float4 const* pos_in, float4 const* vel_in, int4 const* image_in, uint32_t const* id_in,
                      uint4 const* posq_in, uint32_t const* perm, unsigned n, float4* pos_out, float4* vel_out,
                      int4* image_out, uint32_t* id_out, uint4* posq_out, hipStream_t s, uint32_t* rid,
                      float4* r0) {
    if (!n) return;
    rearrange_kernel<<<div_up(n, 256), 256, 0, s>>>(pos_in, vel_in, image_in, id_in, posq_in, perm, n, pos_out,
                                                    vel_out, image_out, id_out, posq_out, rid, r0);
    HIP_CHECK(hipGetLastError());
}

void launch_reverse_id(uint32_t const* id, unsigned n, uint32_t* rid, hipStream_t s) {
    if (!n) return;
    reverse_id_kernel<<<div_up(n, 256), 256, 0, s>>>(id, n, rid);
    HIP_CHECK(hipGetLastError());
}

} // namespace halmd
